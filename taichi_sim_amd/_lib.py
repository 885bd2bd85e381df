"""ctypes binding of libvse (include/vse.h).

The library is built in-tree (``taichi_sim_amd/libvse.so``).  There is no
fallback: if it is missing or cannot be loaded, every entry point raises.

HIP runtime: libvse.so needs libamdhip64.so.7 and libhsa-runtime64.so.1.  The
dynamic linker resolves those names to whichever copy the process loaded
first, and the torch wheel ships its own copies (an older ROCm release) under
the same names.  Load this library before importing torch to run it on the
ROCm runtime it was compiled with; ``runtime_path()`` names the copy in use
and ``load()`` raises when it lies outside that ROCm installation.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvse.so")
ROCM_PATH = os.path.realpath(os.environ.get("ROCM_PATH", "/opt/rocm"))


class _DlInfo(ctypes.Structure):
    _fields_ = [("dli_fname", ctypes.c_char_p), ("dli_fbase", ctypes.c_void_p),
                ("dli_sname", ctypes.c_char_p), ("dli_saddr", ctypes.c_void_p)]


def runtime_path(lib=None):
    """File of the HIP runtime libvse's own calls bind to (dladdr of hipMalloc
    as resolved from the library's handle), or None if it cannot be told."""
    lib = lib if lib is not None else load()
    try:
        addr = ctypes.cast(lib.hipMalloc, ctypes.c_void_p).value
        info = _DlInfo()
        dladdr = ctypes.CDLL(None).dladdr
        dladdr.argtypes = [ctypes.c_void_p, ctypes.POINTER(_DlInfo)]
        if not addr or dladdr(addr, ctypes.byref(info)) == 0 or not info.dli_fname:
            return None
        return os.path.realpath(info.dli_fname.decode())
    except (AttributeError, OSError):
        return None


VSE_OK = 0
VSE_ERR_ARG = -1
VSE_ERR_HIP = -2
VSE_ERR_UNSTABLE = -3
VSE_ERR_STATE = -4
VARIANT_WGSL = 0
VARIANT_TAICHI = 1
FIELD = {"vx": 0, "vy": 1, "vz": 2, "sxx": 3, "syy": 4, "szz": 5, "sxy": 6, "sxz": 7, "syz": 8}
MEMORY = {"xx": 0, "yy": 1, "zz": 2, "xy": 3, "xz": 4, "yz": 5}

# every symbol include/vse.h declares
EXPORTS = ["vse_abi_version", "vse_create", "vse_set_material", "vse_set_cpml", "vse_set_sources",
           "vse_set_receivers", "vse_reset", "vse_run", "vse_run_noalloc", "vse_profile", "vse_set_tuning", "vse_set_math", "vse_set_schedule", "vse_read_field", "vse_read_traces",
           "vse_read_memory", "vse_read_v2max", "vse_get_step", "vse_field_device_ptr",
           "vse_device_bytes", "vse_device_name", "vse_last_error", "vse_destroy", "vse_copy_bandwidth",
           "vse_stream_bandwidth",
           "vse_batch_create", "vse_batch_run", "vse_batch_profile", "vse_batch_set_rows", "vse_batch_set_group_bytes", "vse_batch_set_schedule", "vse_batch_groups",
           "vse_batch_size", "vse_grad_create", "vse_grad_set_tau_scale", "vse_grad_forward", "vse_grad_misfit",
           "vse_grad_backward", "vse_grad_set_flush", "vse_grad_set_overlap", "vse_grad_read", "vse_grad_read_state", "vse_grad_segment", "vse_grad_device_bytes",
           "vse_grad_last_error", "vse_grad_destroy", "vse_batch_last_error", "vse_batch_destroy"]
BATCH_MAX = 256
ABI_VERSION = 3
NOALLOC_SLOTS = 65536


class VseDesc(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float),
                ("ord", ctypes.c_int), ("coefs", ctypes.c_float * 8), ("variant", ctypes.c_int),
                ("visco", ctypes.c_int), ("n_sls", ctypes.c_int), ("tau", ctypes.c_float * 32),
                ("sum_alpha", ctypes.c_float * 2), ("n_rows", ctypes.c_int)]


class VSEError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libvse error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libvse.so; raises if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with taichi_sim_amd._build.build() "
                          "(hipcc --offload-arch=gfx950); the product has no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "vse_abi_version": (I, []),
        "vse_create": (I, [ctypes.POINTER(VseDesc), ctypes.POINTER(P)]),
        "vse_set_material": (I, [P, fp, fp, fp]),
        "vse_set_cpml": (I, [P, I, fp, fp, fp, fp, fp, fp]),
        "vse_set_sources": (I, [P, I, ip, ip, I, fp]),
        "vse_set_receivers": (I, [P, I, ip, ip, ip, I, I]),
        "vse_reset": (I, [P]),
        "vse_run": (I, [P, I, I, fp]),
        "vse_run_noalloc": (I, [P, I, P]),
        "vse_profile": (I, [P, I, fp]),
        "vse_set_tuning": (I, [P, I, I, I, I, I]),
        "vse_set_math": (I, [P, I]),
        "vse_set_schedule": (I, [P, I]),
        "vse_read_field": (I, [P, I, fp]),
        "vse_read_traces": (I, [P, fp]),
        "vse_read_memory": (I, [P, I, fp]),
        "vse_read_v2max": (I, [P, fp]),
        "vse_get_step": (I, [P, ip]),
        "vse_field_device_ptr": (P, [P, I]),
        "vse_device_bytes": (ctypes.c_size_t, [P]),
        "vse_device_name": (ctypes.c_char_p, [P]),
        "vse_last_error": (ctypes.c_char_p, [P]),
        "vse_destroy": (None, [P]),
        "vse_copy_bandwidth": (I, [ctypes.c_size_t, I, fp]),
        "vse_stream_bandwidth": (I, [ctypes.c_size_t, I, I, I, fp]),
        "vse_batch_create": (I, [ctypes.POINTER(P), I, ctypes.POINTER(P)]),
        "vse_batch_run": (I, [P, I, I, fp]),
        "vse_batch_profile": (I, [P, I, fp]),
        "vse_batch_set_rows": (I, [P, I]),
        "vse_batch_set_group_bytes": (I, [P, ctypes.c_size_t]),
        "vse_batch_set_schedule": (I, [P, I]),
        "vse_batch_groups": (I, [P, ip, I]),
        "vse_batch_size": (I, [P]),
        "vse_grad_create": (I, [P, I, ctypes.POINTER(P)]),
        "vse_grad_set_tau_scale": (I, [P, fp]),
        "vse_grad_forward": (I, [P, I, I, fp, fp]),
        "vse_grad_misfit": (I, [P, fp, ctypes.POINTER(ctypes.c_double)]),
        "vse_grad_backward": (I, [P, fp, fp]),
        "vse_grad_set_flush": (I, [P, I]),
        "vse_grad_set_overlap": (I, [P, I]),
        "vse_grad_read": (I, [P, I, ctypes.POINTER(ctypes.c_double)]),
        "vse_grad_read_state": (I, [P, I, fp]),
        "vse_grad_segment": (I, [P]),
        "vse_grad_device_bytes": (ctypes.c_size_t, [P]),
        "vse_grad_last_error": (ctypes.c_char_p, [P]),
        "vse_grad_destroy": (None, [P]),
        "vse_batch_last_error": (ctypes.c_char_p, [P]),
        "vse_batch_destroy": (None, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if lib.vse_abi_version() != ABI_VERSION:
        raise ImportError(f"libvse ABI version {lib.vse_abi_version()}, this binding expects {ABI_VERSION}: rebuild it")
    rt = runtime_path(lib)
    if rt is not None and not rt.startswith(ROCM_PATH + os.sep):
        raise ImportError(f"libvse is bound to the HIP runtime {rt}, loaded earlier by another package, not the one "
                          f"under {ROCM_PATH} it was built with; import taichi_sim_amd (or call "
                          f"taichi_sim_amd._lib.load()) before that package, e.g. before torch")
    _lib = lib
    return lib


def mapped_runtimes():
    """Files of every HIP and HSA runtime mapped into this process (Linux).
    A second copy (another package's) stays idle as long as only libvse's is called."""
    try:
        with open("/proc/self/maps") as f:
            names = {ln.split()[-1] for ln in f if "libamdhip64" in ln or "libhsa-runtime64" in ln}
    except OSError:
        return []
    return sorted(n for n in names if n.startswith("/"))
