"""HIP path (libvse through its C-ABI) against the CPU oracle.

Same setup arrays feed both; after N steps every wavefield, memory variable
and receiver trace must match the FP32 oracle within 1e-5 relative L2 (the
library keeps the reference's FP32 operation order, so the observed gap is
usually zero).  The FP64 oracle gap is asserted alongside.
"""
import numpy as np
import pytest

import oracle
from taichi_sim_amd import workloads
from taichi_sim_amd._lib import VSEError
from taichi_sim_amd.backend import attenuation_from_config
from taichi_sim_amd.engine import VSESim
from taichi_sim_amd.sim_support import StabilityError
from taichi_sim_amd.sim_support.attenuation import AttenuationCoefficients
from taichi_sim_amd.sim_support.simulator import Simulator
from taichi_sim_amd.sim_support.simulator3d import Simulator3D
from tests.conftest import pedro_config

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _compare(sim, variant, steps, att=None, visco=None, check_f64=True, channel=None):
    o = oracle.Oracle(sim, variant=variant, att=att, visco=visco, channel=channel).run(steps)
    g = VSESim.from_simulator(sim, variant=variant, att=att, visco=visco, channel=channel)
    g.run(steps, check_every=64)
    assert g.step() == steps
    names = list(o.f)
    worst = 0.0
    for k in names:
        ref = o.f[k]
        got = g.field(k)
        if np.abs(ref).max() == 0:
            assert np.abs(got).max() == 0, k
            continue
        rel = oracle.rel_l2(got, ref)
        worst = max(worst, rel)
        assert rel <= TOL, (k, rel, np.abs(got - ref).max())
    if g.visco:
        for ci, comp in ([(0, "xx"), (1, "yy"), (3, "xy")] if g.ndim == 2 else
                         list(enumerate(("xx", "yy", "zz", "xy", "xz", "yz")))):
            ref = o.r[ci]
            got = g.memory(comp)
            if np.abs(ref).max() > 0:
                assert oracle.rel_l2(got, ref) <= TOL, comp
    tr = g.traces()
    assert tr.shape == o.traces[:, :tr.shape[1]].shape
    if np.abs(o.traces).max() > 0:
        rel = oracle.rel_l2(tr, o.traces)
        worst = max(worst, rel)
        assert rel <= TOL, ("traces", rel)
    # running max of |v|^2 over every cell and step (the stability check's input)
    assert o.v2max[0] > 0 and abs(g.v2max() - o.v2max[0]) <= TOL * o.v2max[0], (g.v2max(), o.v2max[0])
    gap64 = None
    if check_f64:
        o64 = oracle.Oracle(sim, variant=variant, att=att, visco=visco, precision="f64", channel=channel).run(steps)
        key = "syy" if g.ndim == 2 else "szz"
        gap64 = oracle.rel_l2(g.field(key), o64.f[key])
        assert gap64 < 1e-4
        if np.abs(o64.traces).max() > 0:
            gap64 = (gap64, oracle.rel_l2(tr, o64.traces))
    print(f"worst rel-L2 vs FP32 oracle: {worst:.3e}; vs FP64 oracle (field, traces): {gap64}")
    g.close()
    return worst


@pytest.mark.parametrize("variant", ["wgsl", "taichi"])
def test_2d_elastic_point(variant):
    sim = Simulator(workloads.elastic_2d_point(n=96, steps=300, n_rec=16))
    _compare(sim, variant, 300)


@pytest.mark.parametrize("channel", ["syy", "vy"])
@pytest.mark.parametrize("variant", ["wgsl", "taichi"])
def test_2d_elastic_point_full_size(variant, channel):
    """The CPU-runnable configuration at its full size: 2D elastic aluminium
    256^2 incl. CPML, centred 5 MHz point force, 1000 steps, one line of 128
    receivers recording sigma_yy (the reference's channel) or vy -- the vertical
    particle velocity, "vz" in the ROI's axes."""
    sim = Simulator(workloads.elastic_2d_point(n=256, steps=1000, n_rec=128))
    assert sim._nx == sim._ny == 256 and sim._n_rec == 128
    _compare(sim, variant, 1000, channel=channel)


@pytest.mark.parametrize("channel", ["vz", "vx", "sxz"])
def test_3d_receiver_channels(channel):
    cfg = workloads.visco_3d_block(n=(58, 56, 54), steps=60, n_incl=2, n_rec_side=4, t0=0.3)
    _compare(Simulator3D(cfg), "wgsl", 60, att=attenuation_from_config(cfg), check_f64=False, channel=channel)


def test_2d_visco_velocity_channel():
    cfg = workloads.visco_2d_array(n=(112, 100), steps=200, n_elem=8, n_voids=2)
    _compare(Simulator(cfg), "wgsl", 200, att=attenuation_from_config(cfg), channel="vy")


def test_2d_visco_array_voids():
    cfg = workloads.visco_2d_array(n=128, steps=300, n_elem=8, n_voids=3)
    _compare(Simulator(cfg), "wgsl", 300, att=attenuation_from_config(cfg))


def test_2d_visco_reference_tau_short():
    # the reference's own tau (seconds vs dt in us: r alternates sign each step)
    cfg = workloads.visco_2d_array(n=80, steps=60, n_elem=4, n_voids=1)
    _compare(Simulator(cfg), "wgsl", 60, att=AttenuationCoefficients(), check_f64=False)


@pytest.mark.parametrize("acc", [3, 4, 5, 6])
def test_2d_visco_ragged_high_order(acc):
    # nx != ny, neither a multiple of the block shape
    cfg = workloads.visco_2d_array(n=(108, 101), steps=150, n_elem=6, n_voids=2)
    cfg["simul_params"]["acc"] = acc
    _compare(Simulator(cfg), "wgsl", 150, att=attenuation_from_config(cfg))


def test_2d_partial_emitters_wgsl_stride():
    # only some elements emit: WGSL reads the source with stride n_src (wgsl2d :67-71)
    cfg = workloads.visco_2d_array(n=96, steps=120, n_elem=8, n_voids=0,
                                   emitters=["True", "False", "True", "False", "False", "True", "False", "False"])
    _compare(Simulator(cfg), "wgsl", 120, att=attenuation_from_config(cfg))


def test_source_cell_listed_twice_is_rejected():
    # the reference's source map holds one column per cell
    from taichi_sim_amd._lib import VSEError
    sim = Simulator(workloads.elastic_2d_point(n=64, steps=20, n_rec=4))
    g = VSESim.from_simulator(sim, variant="wgsl")
    cells = np.array([100, 200, 100], dtype=np.int32)
    with pytest.raises(VSEError, match="listed twice"):
        g.set_sources(cells, np.zeros(3, np.int32), np.ones((g.n_rows, 1), np.float32))
    g.close()


@pytest.mark.parametrize("visco", [True, False])
def test_3d_block(visco):
    cfg = workloads.visco_3d_block(n=56, steps=80, n_incl=4, n_rec_side=4, t0=0.3)
    _compare(Simulator3D(cfg), "wgsl", 80, att=attenuation_from_config(cfg), visco=visco)


@pytest.mark.parametrize("acc", [4, 5, 6])
def test_3d_high_order(acc):
    # FD orders 4-6 on a ragged grid, x-march in tile order 2 for both passes
    cfg = workloads.visco_3d_block(n=(50, 48, 52), steps=40, n_incl=2, n_rec_side=3, t0=0.3)
    cfg["simul_params"]["acc"] = acc
    sim = Simulator3D(cfg)
    att = attenuation_from_config(cfg)
    o = oracle.Oracle(sim, variant="wgsl", att=att).run(40)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    g.set_tuning(0, 4, 0, 4, 2)
    g.set_tuning(1, 1, 0, 8, 2)
    g.run(40)
    for k in o.f:
        if np.abs(o.f[k]).max() > 0:
            assert oracle.rel_l2(g.field(k), o.f[k]) <= TOL, (acc, k)
    assert oracle.rel_l2(g.traces(), o.traces) <= TOL
    g.close()


def test_3d_launch_shapes_bit_identical():
    """Every 3D launch shape the autotuner may pick gives the same bits."""
    cfg = workloads.visco_3d_block(n=(70, 64, 130), steps=40, n_incl=3, n_rec_side=3, t0=0.3)     # ny/ty % 8 == 0
    sim = Simulator3D(cfg)
    att = attenuation_from_config(cfg)
    o = oracle.Oracle(sim, variant="wgsl", att=att).run(40)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    ref = None
    for ty, nt, xc, rm in [(4, 0, 0, 0), (4, 0, 1, 0), (1, 0, 7, 1), (1, 1, 16, 0), (4, 1, 3, 1), (4, 1, 64, 1),
                           (1, 0, 1000, 0), (8, 0, 8, 1), (8, 1, 5, 0), (4, 0, 8, 2), (1, 1, 3, 2), (8, 0, 9, 2)]:
        g.set_tuning(0, ty, nt, xc, rm)
        g.set_tuning(1, ty, 1 - nt, xc // 2, (rm + 1) % 3)
        g.reset()
        g.run(40)
        out = [g.field(k) for k in ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")] + [g.traces()]
        if ref is None:
            ref = out
            assert oracle.rel_l2(out[5], o.f["szz"]) <= TOL
        else:
            for a, b in zip(ref, out):
                np.testing.assert_array_equal(a, b)
    g.close()


@pytest.mark.parametrize("variant,channel", [("wgsl", None), ("wgsl", "vy"), ("taichi", None)])
def test_2d_launch_shapes_bit_identical(variant, channel):
    """Every 2D launch shape the autotuner may pick (row groups per workgroup,
    tile order) gives the same bits: fields, memory variables, traces, |v|^2
    max; an nx that no row count divides exercises the ragged last tile."""
    cfg = workloads.visco_2d_array(n=(203, 160), steps=60, n_elem=8, n_voids=2)
    sim = Simulator(cfg)
    att = attenuation_from_config(cfg)
    g = VSESim.from_simulator(sim, variant=variant, att=att, channel=channel)
    ref = None
    for rows, rm, march in [(1, 1, 0), (1, 0, 0), (2, 1, 0), (3, 0, 0), (4, 1, 0), (8, 0, 0), (16, 1, 0),
                            (1, 1, 1), (2, 0, 1), (5, 1, 1), (8, 1, 1), (13, 0, 1), (16, 1, 1)]:
        g.set_tuning(0, 4, 0, rows, rm)
        g.set_tuning(1, 4, march, 17 - rows, 1 - rm)
        g.reset()
        g.run(60, check_every=20)
        out = [g.field(k) for k in ("vx", "vy", "sxx", "syy", "sxy")] + [g.traces(), np.array([g.v2max()])]
        if variant == "wgsl":
            out += [g.memory(c) for c in ("xx", "yy", "xy")]
        if ref is None:
            ref = out
            assert np.abs(out[3]).max() > 0 and np.abs(out[5]).max() > 0
        else:
            for a, b in zip(ref, out):
                np.testing.assert_array_equal(a, b)
    with pytest.raises(VSEError):
        g.set_tuning(0, 4, 0, 17, 1)
    with pytest.raises(VSEError):
        g.set_tuning(0, 4, 1, 4, 1)          # the x-row march is a velocity-pass shape
    g.close()


@pytest.mark.parametrize("acc", [2, 4])
def test_3d_lds_velocity_shapes_bit_identical(acc):
    """The velocity pass with LDS-staged y-neighbours (nt bit 1) gives the same
    bits as the register x-march, on a grid whose ny and nz no tile divides."""
    cfg = workloads.visco_3d_block(n=(50, 61, 130), steps=30, n_incl=3, n_rec_side=3, t0=0.3)
    cfg["simul_params"]["acc"] = acc
    sim = Simulator3D(cfg)
    att = attenuation_from_config(cfg)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    assert g.ord == acc
    g.set_tuning(0, 1, 0, 4, 2)
    shapes = [(1, 0, 8, 2)] + [(ty, nt, xc, rm) for ty, nt, xc, rm in
                               [(4, 2, 8, 2), (8, 2, 1, 0), (8, 3, 3, 1), (4, 3, 64, 0), (8, 2, 16, 2), (4, 2, 5, 1)]
                               if ty >= 2 * acc]
    ref = None
    for shape in shapes:
        g.set_tuning(1, *shape)
        g.reset()
        g.run(30, check_every=10)
        out = _state3d(g)
        if ref is None:
            ref = out
            assert np.abs(out["vz"]).max() > 0
        else:
            for k in ref:
                np.testing.assert_array_equal(out[k], ref[k], err_msg=f"{shape} {k}")
    with pytest.raises(VSEError):
        g.set_tuning(0, 8, 2, 4, 2)          # LDS staging is for the velocity pass only
    with pytest.raises(VSEError):
        g.set_tuning(1, 8, 2, 0, 2)          # not with the plain kernel
    if acc == 4:
        with pytest.raises(VSEError):
            g.set_tuning(1, 4, 2, 8, 2)      # ty < 2 * FD order
    g.close()


def _state3d(g):
    out = {k: g.field(k) for k in ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")}
    for c in ("xx", "yy", "zz", "xy", "xz", "yz"):
        out["r" + c] = g.memory(c)
    out["traces"] = g.traces()
    out["step"] = np.array([g.step()])
    out["v2max"] = np.array([g.v2max()], dtype=np.float32)
    return out


@pytest.mark.parametrize("math", ["fast", "strict"])
@pytest.mark.parametrize("channel", [None, "vz"])
def test_3d_fused_step_bit_identical(math, channel):
    """The fused schedule (velocity of step n and stress of step n+1 in one
    kernel) ends bit for bit where the two-pass schedule ends, for every launch
    shape, for runs split into calls of 1, 2 and odd numbers of steps, with
    CPML bands on every face and a piston source on the top face."""
    steps = 91
    cfg = workloads.visco_3d_block(n=(70, 66, 130), steps=steps, n_incl=3, n_rec_side=3, t0=0.3)
    sim = Simulator3D(cfg)
    att = attenuation_from_config(cfg)
    ref = VSESim.from_simulator(sim, variant="wgsl", att=att, channel=channel)
    ref.set_math(math)
    ref.run(steps, check_every=10)
    want = _state3d(ref)
    ref.close()
    assert np.abs(want["szz"]).max() > 0 and np.abs(want["traces"]).max() > 0
    g = VSESim.from_simulator(sim, variant="wgsl", att=att, channel=channel)
    g.set_math(math)
    g.set_schedule("fused")
    print(f"fused {math} channel {channel}", flush=True)
    for ty, xc, rm, split in [(4, 16, 2, (91,)), (8, 16, 2, (1, 2, 88)), (4, 1, 0, (90, 1)), (8, 3, 1, (5, 86)),
                              (4, 1000, 0, (91,)), (8, 7, 2, (50, 41))]:
        assert sum(split) == steps
        g.set_tuning(2, ty, 0, xc, rm)
        g.reset()
        for n in split:
            g.run(n, check_every=8)
        got = _state3d(g)
        for k in want:
            assert np.array_equal(got[k], want[k]), ((ty, xc, rm, split), k, float(np.abs(got[k] - want[k]).max()))
    ms = g.profile(3)
    assert ms[0] > 0 and ms[1] == 0
    g.close()


def _state2d(g):
    out = {k: g.field(k) for k in ("vx", "vy", "sxx", "syy", "sxy")}
    if g.visco and g.variant == "wgsl":
        for c in ("xx", "yy", "xy"):
            out["r" + c] = g.memory(c)
    out["traces"] = g.traces()
    out["step"] = np.array([g.step()])
    out["v2max"] = np.array([g.v2max()], dtype=np.float32)
    return out


@pytest.mark.parametrize("variant,visco,acc,math,channel,n_sls", [
    ("wgsl", True, 2, "strict", None, 3), ("wgsl", True, 2, "fast", "vy", 3), ("wgsl", True, 4, "strict", None, 3),
    ("wgsl", True, 2, "strict", None, 2), ("wgsl", False, 3, "strict", None, 3), ("taichi", False, 2, "strict", None, 3)])
def test_2d_fused_step_bit_identical(variant, visco, acc, math, channel, n_sls):
    """The fused 2D schedule (one x-marching launch per step: velocity of step
    n, then stress of step n+1, y-neighbour velocities by cross-lane shuffles,
    sigma / v / velocity-pass psi double-buffered) ends bit for bit where the
    two-pass schedule ends: fields, memory variables, traces (sigma or a
    velocity channel), |v|^2 max and step counter, for several x-chunks, for
    runs split into calls of 1, 2 and odd numbers of steps, on a grid whose
    extents no strip or chunk divides, with voids, CPML on every side and an
    array of sources inside the marched rows."""
    steps = 75
    cfg = workloads.visco_2d_array(n=(203, 171), steps=steps, n_elem=8, n_voids=2)
    cfg["simul_params"]["acc"] = acc
    cfg["attenuation"]["n_sls"] = n_sls          # 2: the kernels' runtime mechanism loop
    sim = Simulator(cfg)
    att = attenuation_from_config(cfg)
    ref = VSESim.from_simulator(sim, variant=variant, visco=visco, att=att, channel=channel)
    ref.set_math(math)
    ref.run(steps, check_every=10)
    want = _state2d(ref)
    ref.close()
    assert np.abs(want["syy"]).max() > 0 and np.abs(want["traces"]).max() > 0
    g = VSESim.from_simulator(sim, variant=variant, visco=visco, att=att, channel=channel)
    g.set_math(math)
    g.set_schedule("fused")
    print(f"fused 2D {variant} visco {visco} acc {acc} {math} channel {channel} L {n_sls}", flush=True)
    for xc, rm, split in [(16, 1, (75,)), (8, 0, (1, 2, 72)), (7, 1, (74, 1)), (37, 0, (6, 69)), (1000, 1, (75,)),
                          (11, 1, (40, 35))]:
        assert sum(split) == steps
        g.set_tuning(2, 4, 0, xc, rm)
        g.reset()
        for n in split:
            g.run(n, check_every=8)
        got = _state2d(g)
        for k in want:
            assert np.array_equal(got[k], want[k]), ((xc, rm, split), k, float(np.abs(got[k] - want[k]).max()))
    ms = g.profile(3)
    assert ms[0] > 0 and ms[1] == 0
    with pytest.raises(VSEError):
        g.set_tuning(2, 4, 0, 1, 1)         # more workgroups than |v|^2 slots
    g.close()


def test_3d_fused_schedule_refusals():
    cfg = workloads.visco_3d_block(n=56, steps=4, n_incl=1, n_rec_side=2)
    cfg["simul_params"]["acc"] = 4
    sim = Simulator3D(cfg)
    g = VSESim.from_simulator(sim, variant="wgsl", att=attenuation_from_config(cfg))
    from taichi_sim_amd._lib import VSEError
    with pytest.raises(VSEError, match="FD order 2"):
        g.set_schedule("fused")
    g.close()
    g = VSESim.from_simulator(Simulator(workloads.elastic_2d_point(n=64, steps=4, n_rec=2)), variant="wgsl")
    g.set_tuning(2, 4, 0, 16, 1)
    g.set_schedule("fused")                  # 2D: any variant and order
    with pytest.raises(VSEError):
        g.set_tuning(2, 4, 0, 1, 1)          # one |v|^2 slot per workgroup: x-chunks of ~4 rows at least
    g.close()


@pytest.mark.parametrize("visco", [True, False])
def test_3d_fast_math_within_tolerance(visco):
    """FAST arithmetic (precomputed coef/dx, 1/k) stays within 1e-5 of the oracle."""
    cfg = workloads.visco_3d_block(n=(64, 60, 70), steps=120, n_incl=4, n_rec_side=4, t0=0.3)
    sim = Simulator3D(cfg)
    att = attenuation_from_config(cfg)
    o = oracle.Oracle(sim, variant="wgsl", att=att, visco=visco).run(120)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att, visco=visco)
    g.set_math("fast")
    g.run(120)
    worst = 0.0
    for k in o.f:
        rel = oracle.rel_l2(g.field(k), o.f[k])
        worst = max(worst, rel)
        assert rel <= TOL, (k, rel)
    rel = oracle.rel_l2(g.traces(), o.traces)
    assert rel <= TOL, rel
    print(f"fast math worst field rel-L2 {worst:.2e}, traces {rel:.2e}")
    g.close()


@pytest.mark.parametrize("acc", [2, 4])
def test_2d_fast_math_within_tolerance(acc):
    """FAST arithmetic in 2D (precomputed coef/d, 1/k) stays within 1e-5 of the oracle."""
    cfg = workloads.visco_2d_array(n=(120, 104), steps=300, n_elem=8, n_voids=3)
    cfg["simul_params"]["acc"] = acc
    sim = Simulator(cfg)
    att = attenuation_from_config(cfg)
    o = oracle.Oracle(sim, variant="wgsl", att=att).run(300)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    g.set_math("fast")
    g.run(300, check_every=64)
    worst = 0.0
    for k in o.f:
        rel = oracle.rel_l2(g.field(k), o.f[k])
        worst = max(worst, rel)
        assert rel <= TOL, (k, rel)
    for ci, comp in ((0, "xx"), (1, "yy"), (3, "xy")):
        assert oracle.rel_l2(g.memory(comp), o.r[ci]) <= TOL, comp
    rel = oracle.rel_l2(g.traces(), o.traces)
    assert rel <= TOL, rel
    print(f"2D fast math worst field rel-L2 {worst:.2e}, traces {rel:.2e}")
    g.close()


def test_3d_ragged_order3():
    cfg = workloads.visco_3d_block(n=(55, 53, 50), steps=50, n_incl=2, n_rec_side=3, t0=0.3)
    cfg["simul_params"]["acc"] = 3
    _compare(Simulator3D(cfg), "wgsl", 50, att=attenuation_from_config(cfg))


@pytest.mark.parametrize("variant", ["taichi", "wgsl"])
def test_golden_bscan_teste_pedro(variant):
    """The reference's B-scan (Taichi FP32 on a GPU), all 10000 steps."""
    cfg, bscan = pedro_config()
    sim = Simulator(cfg)
    g = VSESim.from_simulator(sim, variant=variant)
    g.run(sim._n_steps, check_every=1000)
    tr = g.traces()
    g.close()
    assert np.nonzero(np.any(tr[1:] != 0, axis=1))[0][0] + 1 == 335
    rel = oracle.rel_l2(tr[1:], bscan[1:])
    print(f"{variant}: rel-L2 vs golden {rel:.3e}")
    assert rel < 2e-5


def test_stability_error_on_overflow():
    cfg = workloads.elastic_2d_point(n=64, steps=200, n_rec=2)
    sim = Simulator(cfg)
    sim._source_term = (sim._source_term * np.float32(3e37)).astype(np.float32)
    g = VSESim.from_simulator(sim, variant="wgsl")
    with pytest.raises(StabilityError):
        g.run(200, check_every=50)
    g.close()


def test_stability_error_on_nan():
    # the reference's `vsn2 > 1e38` test misses NaN; the device max catches it
    cfg = workloads.elastic_2d_point(n=64, steps=60, n_rec=2)
    sim = Simulator(cfg)
    st = sim._source_term.copy()
    st[10] = np.nan
    sim._source_term = st
    g = VSESim.from_simulator(sim, variant="wgsl")
    with pytest.raises(StabilityError):
        g.run(60, check_every=20)
    g.close()


def test_reset_and_rerun_is_deterministic():
    cfg = workloads.visco_2d_array(n=96, steps=100, n_elem=4, n_voids=1)
    sim = Simulator(cfg)
    g = VSESim.from_simulator(sim, variant="wgsl", att=attenuation_from_config(cfg))
    g.run(100)
    a = g.field("syy"), g.traces()
    g.reset()
    g.run(37)
    g.run(63, check_every=7)      # different graph chunking, same result
    b = g.field("syy"), g.traces()
    g.close()
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_bandwidth_probes_run():
    # the roofline's reference points: plain copy and the stencil passes' read/write mixes
    import ctypes
    from taichi_sim_amd import _lib
    lib = _lib.load()
    g = ctypes.c_float(0.0)
    assert lib.vse_copy_bandwidth(64 << 20, 3, ctypes.byref(g)) == 0 and 100.0 < g.value < 20000.0
    for r, w in ((6, 24), (7, 3), (0, 1), (0, 32)):
        assert lib.vse_stream_bandwidth(1 << 22, r, w, 3, ctypes.byref(g)) == 0, (r, w)
        assert 100.0 < g.value < 20000.0, (r, w, g.value)


@pytest.mark.parametrize("n_sls", [1, 2, 5])
def test_2d_visco_mechanism_counts(n_sls):
    # L != 3 runs the kernels' runtime mechanism loop (the L = 3 kernels issue
    # every memory-variable load up front)
    print(f"2D viscoelastic, {n_sls} mechanisms", flush=True)
    cfg = workloads.visco_2d_array(n=80, steps=200, n_elem=8, n_voids=2)
    cfg["attenuation"]["n_sls"] = n_sls
    _compare(Simulator(cfg), "wgsl", 200, att=attenuation_from_config(cfg))


@pytest.mark.parametrize("n_sls", [2, 5])
def test_3d_visco_mechanism_counts(n_sls):
    print(f"3D viscoelastic, {n_sls} mechanisms", flush=True)
    cfg = workloads.visco_3d_block(n=48, steps=50, n_incl=2, n_rec_side=3, t0=0.3)
    cfg["attenuation"]["n_sls"] = n_sls
    _compare(Simulator3D(cfg), "wgsl", 50, att=attenuation_from_config(cfg), check_f64=False)
