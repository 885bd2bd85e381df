"""The C-ABI library loads and exports every entry point include/vse.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from taichi_sim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "vse.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vse_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert set(header_symbols()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_library_runs_on_the_rocm_runtime_it_was_built_with():
    # tests/conftest.py loads libvse before any test module imports torch,
    # whose wheel carries another libamdhip64 under the same library name
    rt = _lib.runtime_path()
    assert rt is not None and rt.startswith(_lib.ROCM_PATH + os.sep), rt


def test_abi_version_and_arg_validation_without_gpu():
    lib = _lib.load()
    assert lib.vse_abi_version() == _lib.ABI_VERSION == 3
    # argument validation happens before any device call
    d = _lib.VseDesc()
    d.ndim = 4
    h = ctypes.c_void_p()
    assert lib.vse_create(ctypes.byref(d), ctypes.byref(h)) == _lib.VSE_ERR_ARG
    assert not h.value
    # grids beyond the kernels' 32-bit buffer addressing are refused, not run wrong:
    # 2D with 2^30 cells or more, 3D with y-z planes of 2^30 cells or more
    d = _lib.VseDesc(ndim=2, nx=32768, ny=32768, nz=1, dx=0.1, dy=0.1, dz=1.0, dt=1e-3, ord=2, variant=0, visco=0,
                     n_sls=1, n_rows=1)
    assert lib.vse_create(ctypes.byref(d), ctypes.byref(h)) == _lib.VSE_ERR_ARG and not h.value
    d.ndim, d.nx, d.ny, d.nz, d.dz = 3, 8, 32768, 32768, 0.1
    assert lib.vse_create(ctypes.byref(d), ctypes.byref(h)) == _lib.VSE_ERR_ARG and not h.value
    assert lib.vse_last_error(None) == b"null handle"
    # batches: member validation precedes any device call
    bh = ctypes.c_void_p()
    assert lib.vse_batch_create(None, 1, ctypes.byref(bh)) == _lib.VSE_ERR_ARG
    arr = (ctypes.c_void_p * 1)(None)
    assert lib.vse_batch_create(arr, 1, ctypes.byref(bh)) == _lib.VSE_ERR_ARG
    assert lib.vse_batch_create(arr, 0, ctypes.byref(bh)) == _lib.VSE_ERR_ARG
    assert lib.vse_batch_create(arr, _lib.BATCH_MAX + 1, ctypes.byref(bh)) == _lib.VSE_ERR_ARG
    assert not bh.value
    assert lib.vse_batch_size(None) == 0
    assert lib.vse_batch_last_error(None) == b"null handle"
    assert lib.vse_batch_run(None, 1, 1, None) == _lib.VSE_ERR_ARG
    # bandwidth probes: the limits are checked before anything is allocated
    g = ctypes.c_float(0.0)
    assert lib.vse_stream_bandwidth(1 << 20, 30, 3, 1, ctypes.byref(g)) == _lib.VSE_ERR_ARG   # > 32 arrays
    assert lib.vse_stream_bandwidth(1 << 20, 2, 0, 1, ctypes.byref(g)) == _lib.VSE_ERR_ARG    # nothing written
    assert lib.vse_stream_bandwidth(100, 2, 2, 1, ctypes.byref(g)) == _lib.VSE_ERR_ARG        # too small
    assert lib.vse_copy_bandwidth(100, 1, ctypes.byref(g)) == _lib.VSE_ERR_ARG
