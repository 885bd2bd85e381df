"""The drop-in hook and the multi-shot configurations on the GPU.

* ``implementation()`` of each backend class (the reference's plugin API,
  sim_support/simulator.py:337,608-611; Simulator3D.py:356): both result-key
  families, the field and traces against the CPU oracle, a positive loop time;
  and its show_debug / show_anim output at the reference's display steps.
* Full-matrix capture batches (the batched multi-shot configuration):
  reduced sizes against solo runs bit for bit and one member against the
  oracle; the full sizes (512^2 .. 2048^2, 16 shots each) against solo runs
  over a short window.
* A batch whose members record a velocity channel.
* The torch.autograd.Function with the default tau scale (None).
"""
import numpy as np
import pytest

import oracle
from taichi_sim_amd import workloads
from taichi_sim_amd.backend import (SimulatorHIP, SimulatorHIP3D, SimulatorHIPTaichiElastic,
                                    attenuation_from_config)
from taichi_sim_amd.engine import VSEBatch, VSESim
from taichi_sim_amd.sim_support.simulator import Simulator

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _check_result(res, o, field_key):
    assert set(res) >= {"stress", "sens_stress", "pressure", "sens_pressure", "gpu_str", "sim_time"}
    assert res["stress"] is res["pressure"] and res["sens_stress"] is res["sens_pressure"]
    assert isinstance(res["gpu_str"], str) and res["gpu_str"]
    assert res["sim_time"] > 0.0
    assert res["stress"].dtype == np.float32 and res["stress"].shape == o.f[field_key].shape
    assert oracle.rel_l2(res["stress"], o.f[field_key]) <= TOL
    assert res["sens_stress"].shape == o.traces.shape
    assert np.abs(o.traces).max() > 0
    assert oracle.rel_l2(res["sens_stress"], o.traces) <= TOL


def test_implementation_2d_visco_wgsl():
    cfg = workloads.visco_2d_array(n=(112, 100), steps=240, n_elem=8, n_voids=2)
    sim = SimulatorHIP(cfg)
    res = sim.implementation()
    o = oracle.Oracle(sim, variant="wgsl", att=attenuation_from_config(cfg)).run(240)
    _check_result(res, o, "syy")


def test_implementation_2d_elastic_taichi():
    cfg = workloads.elastic_2d_point(n=128, steps=400, n_rec=24)
    sim = SimulatorHIPTaichiElastic(cfg)
    res = sim.implementation()
    o = oracle.Oracle(sim, variant="taichi").run(400)
    _check_result(res, o, "syy")


@pytest.mark.parametrize("cls,dims", [(SimulatorHIP, 2), (SimulatorHIPTaichiElastic, 2), (SimulatorHIP3D, 3)])
def test_implementation_display_steps(cls, dims, capsys):
    # show_debug / show_anim: after step it with it % it_display == 0 or it == 5
    # the reference prints two lines and shows sigma_yy (2D) or sigma_zz at y
    # index gcd(ny, 16) // 2 (3D) over the ROI (sim_webgpu_viscoelastic.py
    # :519-531, sim_taichi_elastic.py:320-328, sim_webgpu_3D_viscoelastic.py
    # :680-689).  The run is split at those steps; its results must not change.
    if dims == 2:
        cfg = (workloads.elastic_2d_point(n=96, steps=230, n_rec=12) if cls is SimulatorHIPTaichiElastic
               else workloads.visco_2d_array(n=96, steps=230, n_elem=8, n_voids=2))
        field = "syy"
    else:
        cfg = workloads.visco_3d_block(n=(52, 48, 56), steps=46, n_incl=2, n_rec_side=3, t0=0.3)
        field = "szz"
    n = cfg["simul_params"]["time_steps"]
    cfg["simul_params"]["it_display"] = 20 if dims == 2 else 9
    plain = cls(cfg).implementation()
    cfg["simul_configs"]["show_debug"] = True
    cfg["simul_configs"]["show_anim"] = True
    sim = cls(cfg)
    frames = []
    sim.frame_hook = lambda it, frame: frames.append((it, frame.copy()))
    capsys.readouterr()
    res = sim.implementation()
    out = capsys.readouterr().out
    k = cfg["simul_params"]["it_display"]
    steps = sorted(set(range(k, n + 1, k)) | {5})
    assert [it for it, _ in frames] == steps
    assert out.count("Time step # ") == len(steps) and f"Time step # {steps[-1]} out of {n}" in out
    assert ("Max norm velocity vector (m/s) = " if cls is SimulatorHIPTaichiElastic
            else "Max norm velocity vector V (m/s) = ") in out
    for key in ("stress", "sens_stress"):
        assert np.array_equal(res[key], plain[key]), key
    # a frame is the displayed slice of the field after its step
    roi = sim._roi
    it, last = frames[-1]
    g = VSESim.from_simulator(sim, variant=sim._variant, att=attenuation_from_config(cfg))
    g.run(it)
    full = g.field(field)
    g.close()
    if dims == 2:
        want = full[roi.get_ix_min():roi.get_ix_max(), roi.get_iz_min():roi.get_iz_max()]
    else:
        want = full[roi.get_ix_min():roi.get_ix_max(), int(np.gcd(int(roi.get_ny()), 16)) // 2,
                    roi.get_iz_min():roi.get_iz_max()]
    assert last.shape == want.shape and np.array_equal(last, want)
    assert np.abs(last).max() > 0


def test_implementation_3d_visco():
    cfg = workloads.visco_3d_block(n=(52, 48, 56), steps=70, n_incl=3, n_rec_side=4, t0=0.3)
    sim = SimulatorHIP3D(cfg)
    res = sim.implementation()
    o = oracle.Oracle(sim, variant="wgsl", att=attenuation_from_config(cfg)).run(70)
    _check_result(res, o, "szz")


def _fmc_members(sizes, n_elem, steps, n_voids, channel=None):
    out = []
    for cfg, laws in workloads.fmc_2d(sizes=sizes, n_elem=n_elem, steps=steps, n_voids=n_voids):
        sim = Simulator(cfg)
        sim._emission_laws = laws
        att = attenuation_from_config(cfg)
        for law in range(laws.shape[0]):
            sim.select_law(law)
            out.append((sim, law, att, VSESim.from_simulator(sim, variant="wgsl", att=att, channel=channel)))
    return out


def _solo(sim, law, att, steps, channel=None):
    sim.select_law(law)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att, channel=channel)
    g.run(steps, check_every=100)
    return g


def test_fmc_batch_reduced_sizes_bit_identical_and_oracle():
    """FMC (law k fires element k) on four plates of uneven sizes, 4 shots
    each, all 16 shots in one batch: each member ends bit-identical to its solo
    run, and one member matches the oracle."""
    steps = 260
    members = _fmc_members((64, 96, 128, 160), 4, steps, n_voids=2)
    batch = VSEBatch([m[3] for m in members])
    batch.run(steps, check_every=100)
    batch.close()
    for sim, law, att, g in members:
        solo = _solo(sim, law, att, steps)
        for k in ("vx", "vy", "sxx", "syy", "sxy"):
            assert np.array_equal(g.field(k), solo.field(k)), (sim._nx, law, k)
        for c in ("xx", "yy", "xy"):
            assert np.array_equal(g.memory(c), solo.memory(c)), (sim._nx, law, c)
        assert np.array_equal(g.traces(), solo.traces()), (sim._nx, law)
        solo.close()
    # the shots differ: each fires its own element
    tr = [m[3].traces() for m in members[-4:]]
    assert all(not np.array_equal(tr[0], t) for t in tr[1:])
    sim, law, att, g = members[-1]
    sim.select_law(law)
    o = oracle.Oracle(sim, variant="wgsl", att=att).run(steps)
    assert np.abs(o.traces).max() > 0
    assert oracle.rel_l2(g.traces(), o.traces) <= TOL
    assert oracle.rel_l2(g.field("syy"), o.f["syy"]) <= TOL
    for m in members:
        m[3].close()


def test_fmc_batch_full_sizes_match_solo_runs():
    """The full configuration -- 512^2, 1024^2, 1536^2, 2048^2 plates with 16
    voids, 16 FMC shots each, 64 simulations in one batch -- over a short
    window: every member bit-identical to its solo run."""
    steps = 24
    members = _fmc_members((512, 1024, 1536, 2048), 16, steps, n_voids=16)
    assert len(members) == 64
    batch = VSEBatch([m[3] for m in members])
    print(f"full FMC batch: {len(batch.groups())} groups", flush=True)
    batch.run(steps, check_every=12)
    batch.close()
    for sim, law, att, g in members:
        solo = _solo(sim, law, att, steps)
        for k in ("vx", "vy", "syy"):
            assert np.array_equal(g.field(k), solo.field(k)), (sim._nx, law, k)
        assert np.array_equal(g.traces(), solo.traces()), (sim._nx, law)
        assert g.v2max() == solo.v2max()
        solo.close()
        g.close()


def test_batch_velocity_channel_matches_solo():
    steps = 200
    members = _fmc_members((80, 112), 3, steps, n_voids=1, channel="vy")
    for grouping in (0, None):
        for m in members:
            m[3].reset()
        batch = VSEBatch([m[3] for m in members])
        if grouping == 0:
            batch.set_group_bytes(0)
        batch.run(steps, check_every=64)
        batch.close()
        for sim, law, att, g in members:
            solo = _solo(sim, law, att, steps, channel="vy")
            assert np.array_equal(g.traces(), solo.traces()), (grouping, sim._nx, law)
            assert np.array_equal(g.field("vy"), solo.field("vy"))
            solo.close()
    sim, law, att, g = members[-1]
    sim.select_law(law)
    o = oracle.Oracle(sim, variant="wgsl", att=att, channel="vy").run(steps)
    assert np.abs(o.traces).max() > 0
    assert oracle.rel_l2(g.traces(), o.traces) <= TOL
    for m in members:
        m[3].close()


def test_autograd_function_default_tau_scale():
    """ViscoTraces with tscale None (every relaxation time at scale 1) gives
    gradients for rho, cp, cs and none for the scale; a node whose engine ran
    another forward refuses to differentiate."""
    import torch
    from taichi_sim_amd.adjoint import VSEGrad, ViscoTraces
    steps = 120
    cfg = workloads.visco_2d_array(n=64, steps=steps, n_elem=4, n_voids=0)
    sim = Simulator(cfg)
    att = attenuation_from_config(cfg)
    eng = VSEGrad(VSESim.from_simulator(sim, variant="wgsl", att=att))
    maps = [torch.tensor(np.asarray(a, dtype=np.float64), requires_grad=True)
            for a in (sim._rho_grid_vx, sim._cp_grid_vx, sim._cs_grid_vx)]
    tr = ViscoTraces.apply(eng, *maps, None)
    (0.5 * (tr ** 2).sum()).backward()
    for p in maps:
        assert p.grad is not None and torch.isfinite(p.grad).all() and p.grad.abs().max() > 0
    # two nodes on one engine: the first one's checkpoints are gone
    a = ViscoTraces.apply(eng, *maps, None)
    b = ViscoTraces.apply(eng, *maps, None)
    (b ** 2).sum().backward(retain_graph=True)
    with pytest.raises(RuntimeError, match="another forward"):
        (a ** 2).sum().backward()
    eng.sim.close()
