"""vse_run_noalloc (no device allocation, no host synchronisation) against
vse_run through the C-ABI: identical fields, memory variables, traces, step
counter and running |v|^2 max, on two calls in a row (the second on a smaller
grid, after the first left its values in the shared scratch), in 2D and 3D,
and split into several calls; plus the refusals (grid beyond the scratch,
handle not ready)."""
import numpy as np
import pytest

from taichi_sim_amd import _lib, workloads
from taichi_sim_amd.backend import attenuation_from_config
from taichi_sim_amd.engine import VSESim
from taichi_sim_amd.sim_support.simulator import Simulator
from taichi_sim_amd.sim_support.simulator3d import Simulator3D

pytestmark = pytest.mark.gpu


def _state(g):
    names = ("vx", "vy", "sxx", "syy", "sxy") if g.ndim == 2 else ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy",
                                                                   "sxz", "syz")
    out = {k: g.field(k) for k in names}
    if g.visco:
        for c in (("xx", "yy", "xy") if g.ndim == 2 else ("xx", "yy", "zz", "xy", "xz", "yz")):
            out["r" + c] = g.memory(c)
    out["traces"] = g.traces()
    out["step"] = np.array([g.step()])
    out["v2max"] = np.array([g.v2max()], dtype=np.float32)
    return out


def _pair(sim, att, channel=None):
    mk = lambda: VSESim.from_simulator(sim, variant="wgsl", att=att, channel=channel)
    return mk(), mk()


def _same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), (what, k)


def test_noalloc_matches_run_two_calls_second_smaller():
    print("noalloc: 2D 160x150 then 2D 96x90", flush=True)
    cfg1 = workloads.visco_2d_array(n=(160, 150), steps=220, n_elem=8, n_voids=2)
    cfg2 = workloads.visco_2d_array(n=(96, 90), steps=130, n_elem=4, n_voids=1)
    for cfg, steps in ((cfg1, 220), (cfg2, 130)):
        sim = Simulator(cfg)
        ref, got = _pair(sim, attenuation_from_config(cfg))
        ref.run(steps, check_every=50)
        got.run_noalloc(steps)
        a, b = _state(ref), _state(got)
        assert b["step"][0] == steps and b["v2max"][0] > 0
        _same(a, b, sim._nx)
        ref.close()
        got.close()


def test_noalloc_split_calls_and_velocity_channel():
    print("noalloc: split calls, vy channel", flush=True)
    cfg = workloads.visco_2d_array(n=(120, 104), steps=180, n_elem=6, n_voids=2)
    sim = Simulator(cfg)
    ref, got = _pair(sim, attenuation_from_config(cfg), channel="vy")
    ref.run(180)
    got.run_noalloc(50)
    got.run(30)                 # a plain run continues from the counter the scratch handed back
    got.run_noalloc(100)
    _same(_state(ref), _state(got), "split")
    ref.close()
    got.close()


def test_noalloc_3d_matches_run():
    print("noalloc: 3D 58x56x54, then 3D 50x52x48", flush=True)
    for n, steps in (((58, 56, 54), 50), ((50, 52, 48), 30)):
        cfg = workloads.visco_3d_block(n=n, steps=steps, n_incl=2, n_rec_side=3, t0=0.3)
        sim = Simulator3D(cfg)
        ref, got = _pair(sim, attenuation_from_config(cfg))
        ref.run(steps)
        got.run_noalloc(steps)
        _same(_state(ref), _state(got), n)
        ref.close()
        got.close()


def test_noalloc_refusals():
    print("noalloc: refusals", flush=True)
    # a handle whose state is not allocated yet
    g = VSESim(2, (64, 64), (0.1, 0.1), 1e-3, [1.0, 0.0], n_rows=4)
    with pytest.raises(_lib.VSEError, match="vse_reset"):
        g.run_noalloc(2)
    g.close()
    # a 2D grid with more velocity workgroups than the scratch holds
    sim = Simulator(workloads.elastic_2d_point(n=(4200, 4096), steps=4, n_rec=2))
    big = VSESim.from_simulator(sim, variant="wgsl")
    assert (4200 // 4) * (4096 // 64) > _lib.NOALLOC_SLOTS
    with pytest.raises(_lib.VSEError, match="VSE_NOALLOC_SLOTS"):
        big.run_noalloc(1)
    big.close()
