import ast
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")

# libvse, and with it the ROCm runtime it was compiled with, before any test
# module imports torch (whose wheel carries another HIP runtime under the
# same library names)
try:
    from taichi_sim_amd import _lib as _vse_lib
    _vse_lib.load()
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built libvse.so")


def pytest_report_header(config):
    # which HIP runtime libvse is bound to, and every runtime mapped in this process
    try:
        from taichi_sim_amd import _lib
        lib = _lib.load()
        return [f"libvse HIP runtime: {_lib.runtime_path(lib)}", f"mapped HIP/HSA runtimes: {_lib.mapped_runtimes()}"]
    except ImportError as e:
        return [f"libvse not loaded: {e}"]


def pytest_terminal_summary(terminalreporter):
    # after collection and every test: the runtimes the whole session mapped
    try:
        from taichi_sim_amd import _lib
        terminalreporter.write_line(f"mapped HIP/HSA runtimes at the end: {_lib.mapped_runtimes()}")
    except ImportError:
        pass


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work, so a crash can be attributed
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield


def pedro_config(n_steps=None):
    """The reference's teste_pedro configuration with its maps loaded from the fixture."""
    d = np.load(os.path.join(GOLDEN, "pedro_teste_aco.npz"), allow_pickle=False)
    cfg = ast.literal_eval(str(d["config"]))
    for k in ("rho", "cp", "cs"):
        cfg["specimen_params"][k + "_map"] = d[k]
    cfg["simul_configs"]["results_dir"] = ""
    if n_steps is not None:
        cfg["simul_params"]["time_steps"] = n_steps
    return cfg, d["bscan"]
