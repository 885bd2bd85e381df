"""Checkpointed forward and hand-written adjoint (vse_grad_*) on the GPU.

* The out-of-place forward at tau scale 1 is bit-identical to vse_run.
* With a per-cell tau scale it matches the FP64 torch restatement.
* The gradient with respect to rho, cp, cs and the tau scale matches FP64
  torch autograd through that restatement on a 64x64 grid (ROI 22x22).
* Fourth-order central differences of the GPU misfit along 10 seeded
  directions (all four maps perturbed together) agree with the adjoint's
  directional derivative to 1e-3.
* The checkpoint spacing changes nothing (bit-identical gradients).
* The FP32 partials flushed into FP64 every 32 steps stay within 1e-6 of a
  per-step FP64 time accumulation.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle.visco2d_torch import Visco2D
from taichi_sim_amd import workloads
from taichi_sim_amd._lib import VSEError
from taichi_sim_amd.adjoint import GRAD_NAMES, VSEGrad, ViscoTraces, chain_to_cells, misfit_and_gradient
from taichi_sim_amd.backend import attenuation_from_config
from taichi_sim_amd.engine import VSESim
from taichi_sim_amd.sim_support.simulator import Simulator

pytestmark = pytest.mark.gpu


def _setup(n, steps, voids=0, n_elem=4):
    cfg = workloads.visco_2d_array(n=n, steps=steps, n_elem=n_elem, n_voids=voids)
    sim = Simulator(cfg)
    return sim, attenuation_from_config(cfg)


def _maps(sim):
    return [np.asarray(a, dtype=np.float32).copy() for a in (sim._rho_grid_vx, sim._cp_grid_vx, sim._cs_grid_vx)]


def _smooth_field(shape, seed, amp):
    """Seeded smooth perturbation (sum of a few Gaussians), zero near the edges."""
    rng = np.random.default_rng(seed)
    nx, ny = shape
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    f = np.zeros(shape)
    for _ in range(4):
        cx, cy = rng.uniform(0.3, 0.7) * nx, rng.uniform(0.3, 0.7) * ny
        s = rng.uniform(0.05, 0.12) * min(nx, ny)
        f += rng.uniform(-1, 1) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s))
    return amp * f / max(np.abs(f).max(), 1e-30)


def test_closing_the_simulation_closes_its_gradient_engine():
    from taichi_sim_amd._lib import VSEError
    sim, att = _setup(48, 20)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    eng = VSEGrad(g)
    eng.forward(20)
    g.close()
    assert eng._h is None
    with pytest.raises(VSEError):
        eng.forward(20)
    with pytest.raises(ValueError):
        VSEGrad(g)


def test_checkpointed_forward_is_bit_identical_to_run():
    sim, att = _setup(96, 200, voids=2, n_elem=8)
    print("checkpointed forward vs vse_run", flush=True)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    g.run(200, check_every=64)
    ref_tr, ref_f = g.traces(), {k: g.field(k) for k in ("vx", "vy", "sxx", "syy", "sxy")}
    g.close()
    for seg in (0, 7):
        eng = VSEGrad(VSESim.from_simulator(sim, variant="wgsl", att=att), seg_len=seg)
        tr = eng.forward(200)
        assert np.array_equal(tr, ref_tr)
        for k, v in ref_f.items():
            assert np.array_equal(eng.state(k), v), (seg, k)
        eng.close()


def test_tau_scale_forward_vs_fp64_restatement():
    sim, att = _setup(64, 200)
    ts = (1.0 + _smooth_field((sim._nx, sim._ny), 3, 0.4)).astype(np.float32)
    eng = VSEGrad(VSESim.from_simulator(sim, variant="wgsl", att=att))
    rho, cp, cs = _maps(sim)
    eng.set_model(rho, cp, cs, ts)
    tr = eng.forward(200)
    m = Visco2D(sim, att, n_steps=200)
    t64 = lambda a: torch.tensor(np.asarray(a, dtype=np.float64))
    ref, st = m.forward(t64(rho), t64(cp), t64(cs), t64(ts), return_state=True)
    rel = oracle.rel_l2(tr, ref.numpy())
    print(f"traces rel-L2 vs FP64: {rel:.2e}")
    assert rel < 1e-4
    assert oracle.rel_l2(eng.state("syy"), st["syy"].numpy()) < 1e-4
    # and the scale matters (well above the FP32 gap just measured)
    eng.set_model(rho, cp, cs, None)
    assert oracle.rel_l2(eng.forward(200), tr) > 100 * rel
    eng.close()


def _near_array_field(shape, seed, y0, y1):
    """Seeded smooth field of Gaussians centred at depths y0..y1 (cells),
    normalised to max |f| = 1."""
    rng = np.random.default_rng(seed)
    nx, ny = shape
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    f = np.zeros(shape)
    for _ in range(4):
        cx, cy = rng.uniform(0.3, 0.7) * nx, rng.uniform(y0, y1)
        s = rng.uniform(3.0, 8.0)
        f += rng.uniform(-1, 1) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s))
    return f / max(np.abs(f).max(), 1e-30)


def _perturbed(sim, amp=1.0):
    rho, cp, cs = _maps(sim)
    shape = rho.shape
    true = (rho, cp, cs, np.ones(shape, np.float32))
    start = (rho * (1 + _smooth_field(shape, 5, 0.03 * amp)).astype(np.float32),
             cp * (1 - _smooth_field(shape, 6, 0.04 * amp)).astype(np.float32),
             cs * (1 + _smooth_field(shape, 7, 0.04 * amp)).astype(np.float32),
             (1 + _smooth_field(shape, 8, 0.2 * amp)).astype(np.float32))
    return true, start


def test_gradient_vs_fp64_autograd_64():
    steps = 250
    sim, att = _setup(64, steps)
    true, start = _perturbed(sim)
    eng = VSEGrad(VSESim.from_simulator(sim, variant="wgsl", att=att), seg_len=16)
    eng.set_model(*true)
    obs = eng.forward(steps)
    eng.set_model(*start)
    eng.forward(steps)
    J = eng.misfit(obs)
    coef = eng.backward()
    g_rho, g_cp, g_cs = (t.numpy() for t in chain_to_cells(coef, *start[:3]))
    eng.close()

    m = Visco2D(sim, att, n_steps=steps)
    params = [torch.tensor(a.astype(np.float64), requires_grad=True) for a in start]
    tr = m.forward(*params)
    J64 = 0.5 * ((tr - torch.tensor(obs.astype(np.float64))) ** 2).sum()
    J64.backward()
    print(f"misfit GPU {J:.6e} FP64 {float(J64):.6e}")
    assert abs(J - float(J64)) <= 1e-4 * abs(float(J64))
    for name, got, p in (("rho", g_rho, params[0]), ("cp", g_cp, params[1]), ("cs", g_cs, params[2]),
                         ("tau", coef["tau"], params[3])):
        ref = p.grad.numpy()
        rel = oracle.rel_l2(got, ref)
        print(f"d/d{name}: rel-L2 vs FP64 autograd {rel:.2e}")
        assert np.abs(ref).max() > 0
        assert rel < 1e-3, name


def test_fp32_partials_against_per_step_fp64_accumulation():
    # The gradient maps take each step's FP32 terms into FP32 partials and add
    # those into FP64 every 32 steps; with a flush every step the time
    # accumulation is FP64 throughout.  The default grouping must stay within
    # 1e-6 rel-L2 of it on every map (measured 6e-8: the FP32 rounding of
    # 32-term partial sums), and both as close to FP64 autograd as the FP32
    # wavefields allow (the same 7e-6 .. 1.3e-5 for both).
    steps = 250
    sim, att = _setup(64, steps)
    true, start = _perturbed(sim)
    eng = VSEGrad(VSESim.from_simulator(sim, variant="wgsl", att=att), seg_len=16)
    eng.set_model(*true)
    obs = eng.forward(steps)
    res = {}
    for every in (32, 1, 7):
        eng.set_flush(every)
        eng.set_model(*start)
        eng.forward(steps)
        eng.misfit(obs)
        res[every] = eng.backward()
    for bad in (0, 4097):
        with pytest.raises(VSEError):
            eng.set_flush(bad)
    eng.close()

    m = Visco2D(sim, att, n_steps=steps)
    params = [torch.tensor(a.astype(np.float64), requires_grad=True) for a in start]
    tr = m.forward(*params)
    (0.5 * ((tr - torch.tensor(obs.astype(np.float64))) ** 2).sum()).backward()
    cells = {every: [t.numpy() for t in chain_to_cells(r, *start[:3])] + [r["tau"]] for every, r in res.items()}
    for name in GRAD_NAMES:
        d32, d7 = oracle.rel_l2(res[32][name], res[1][name]), oracle.rel_l2(res[7][name], res[1][name])
        print(f"{name}: flush 32 vs every step {d32:.2e}, flush 7 vs every step {d7:.2e}", flush=True)
        assert np.abs(res[1][name]).max() > 0
        assert d32 < 1e-6 and d7 < 1e-6, name
    for i, (name, p) in enumerate(zip(("rho", "cp", "cs", "tau"), params)):
        ref = p.grad.numpy()
        e32, e1 = oracle.rel_l2(cells[32][i], ref), oracle.rel_l2(cells[1][i], ref)
        print(f"d/d{name} vs FP64 autograd: flush 32 {e32:.2e}, every step {e1:.2e}", flush=True)
        assert e32 < 1e-3 and e1 < 1e-3, name


def test_gradient_central_differences_10_directions():
    # Misfit against zero data, J = 0.5 * ||traces||^2.  The array both emits
    # and records, so against a true model's data the residual is ~1e-4 of the
    # direct wave and the FP32 rounding of the traces alone is ~1e-3 of J's
    # differences; against zero data the FP32 forward resolves J to ~1e-7.
    # (The gradient of a true-model misfit is checked against FP64 autograd.)
    steps = 300
    sim, att = _setup(128, steps)
    _, start = _perturbed(sim)
    eng = VSEGrad(VSESim.from_simulator(sim, variant="wgsl", att=att))
    obs = np.zeros((steps, sim._n_rec), np.float32)

    def misfit(model):
        eng.set_model(*model)
        eng.forward(steps)
        return eng.misfit(obs)

    J0 = misfit(start)
    print(f"misfit {J0:.6e}")
    coef = eng.backward()
    grads = list(t.numpy() for t in chain_to_cells(coef, *start[:3])) + [coef["tau"]]
    # each direction moves all four parameter maps at once (smooth seeded
    # fields under the array, where the recorded waves are sensitive to the
    # medium, scaled to each map's magnitude): no direction is close to
    # orthogonal to the gradient, where an FP32 forward could not resolve the
    # difference quotient
    rng = np.random.default_rng(21)
    y_top = int(np.min(np.nonzero(np.asarray(sim._pos_sources) != -1)[1]))
    errs = []
    for k in range(10):
        dirs = [_near_array_field(start[0].shape, 100 + 4 * k + p, y_top + 2, y_top + 25) * rng.choice([-1.0, 1.0])
                * float(np.abs(start[p]).max()) for p in range(4)]
        # relative step 2e-2: the error of the fourth-order quotient falls as h
        # grows up to here (FP32 rounding of J dominates below) and shows no
        # truncation term yet (measured: 1e-5..2e-4 at 1e-2 and 2e-2)
        h = 2e-2

        models = {t: [(start[p] + t * h * dirs[p]).astype(np.float32) for p in range(4)] for t in (-2, -1, 1, 2)}
        J = {t: misfit(m) for t, m in models.items()}
        # fourth-order central difference of J along the direction ...
        fd = (8.0 * (J[1] - J[-1]) - (J[2] - J[-2])) / 12.0
        # ... and the adjoint's prediction for the same combination of the
        # steps actually taken (the FP32 models round start + t*h*dir)
        step = lambda p, t: models[t][p].astype(np.float64) - start[p].astype(np.float64)
        ad = sum(float((grads[p] * (8.0 * (step(p, 1) - step(p, -1)) - (step(p, 2) - step(p, -2))) / 12.0).sum())
                 for p in range(4))
        err = abs(fd - ad) / max(abs(fd), 1e-30)
        errs.append(err)
        print(f"direction {k}: FD {fd:.6e} adjoint {ad:.6e} rel {err:.1e}", flush=True)
    print(f"worst FD disagreement {max(errs):.1e}")
    eng.close()
    assert max(errs) <= 1e-3


def test_segment_length_does_not_change_gradient():
    steps = 120
    sim, att = _setup(80, steps)
    true, start = _perturbed(sim)
    res = []
    for seg in (2, 11, 0, 120):
        eng = VSEGrad(VSESim.from_simulator(sim, variant="wgsl", att=att), seg_len=seg)
        eng.set_model(*true)
        obs = eng.forward(steps)
        eng.set_model(*start)
        eng.forward(steps)
        eng.misfit(obs)
        res.append(eng.backward())
        eng.close()
    for r in res[1:]:
        for k in GRAD_NAMES:
            assert np.array_equal(r[k], res[0][k]), k


def test_autograd_function_and_fused_path_agree():
    steps = 150
    sim, att = _setup(72, steps)
    true, start = _perturbed(sim)
    eng = VSEGrad(VSESim.from_simulator(sim, variant="wgsl", att=att))
    eng.set_model(*true)
    obs = eng.forward(steps)
    # torch.autograd.Function: misfit in FP64 torch on the host
    params = [torch.tensor(a.astype(np.float64), requires_grad=True) for a in start]
    tr = ViscoTraces.apply(eng, *params)
    J = 0.5 * ((tr - torch.tensor(obs.astype(np.float64))) ** 2).sum()
    J.backward()
    # fused path: (rho, lambda, mu) parameterisation, FP64 misfit on the GPU
    rho, cp, cs, ts = (a.astype(np.float64) for a in start)
    lam, mu = rho * (cp * cp - 2 * cs * cs), rho * cs * cs
    J2, g = misfit_and_gradient(eng, obs, rho, lam, mu, ts.astype(np.float32))
    assert abs(J2 - float(J)) <= 1e-5 * abs(float(J))
    assert oracle.rel_l2(g["tau"], params[3].grad.numpy()) < 1e-4
    # chain (rho, cp, cs) -> (rho, lambda, mu): d/dlambda = d/dcp * dcp/dlambda
    dcp_dlam = 1.0 / (2.0 * rho * cp)
    inner = (slice(4, -4), slice(4, -4))
    assert oracle.rel_l2(g["lambda"][inner], (params[1].grad.numpy() * dcp_dlam)[inner]) < 1e-4
    eng.close()


def test_overlapped_reverse_sweep_matches_serial():
    # With sqrt(N) checkpoints the default sweep recomputes segment s-1 on a
    # second stream while the adjoint of segment s runs; serial recomputes
    # each segment before its adjoint on one stream.  Same kernels on the same
    # inputs: the gradient must have the same bits, with the last segment
    # stored by the forward or recomputed (a second backward), for several
    # spacings (a partial last segment included).
    steps = 130
    sim, att = _setup(80, steps)
    true, start = _perturbed(sim)
    for seg in (2, 11, 0):
        print(f"segment length {seg}", flush=True)
        eng = VSEGrad(VSESim.from_simulator(sim, variant="wgsl", att=att), seg_len=seg)
        eng.set_model(*true)
        obs = eng.forward(steps)
        eng.set_model(*start)
        eng.forward(steps)
        eng.misfit(obs)
        eng.set_overlap(False)
        ref = eng.backward()
        eng.set_overlap(True)
        runs = [eng.backward()]                 # last segment recomputed on the second stream
        eng.forward(steps)
        eng.misfit(obs)
        runs.append(eng.backward())             # last segment as the forward stored it
        for got in runs:
            for k in GRAD_NAMES:
                assert np.array_equal(got[k], ref[k]), (seg, k)
        with pytest.raises(VSEError):
            eng._check(eng._lib.vse_grad_set_overlap(eng._h, 2))
        eng.close()


def test_gradient_two_mechanisms_vs_fp64_autograd():
    # L = 2: the adjoint's runtime mechanism loop (L = 3 has its own kernels)
    steps = 250
    cfg = workloads.visco_2d_array(n=64, steps=steps, n_elem=4, n_voids=0)
    cfg["attenuation"]["n_sls"] = 2
    sim = Simulator(cfg)
    att = attenuation_from_config(cfg)
    true, start = _perturbed(sim)
    eng = VSEGrad(VSESim.from_simulator(sim, variant="wgsl", att=att), seg_len=16)
    eng.set_model(*true)
    obs = eng.forward(steps)
    eng.set_model(*start)
    eng.forward(steps)
    J = eng.misfit(obs)
    coef = eng.backward()
    g_rho, g_cp, g_cs = (t.numpy() for t in chain_to_cells(coef, *start[:3]))
    eng.close()
    assert J > 0
    m = Visco2D(sim, att, n_steps=steps)
    params = [torch.tensor(a.astype(np.float64), requires_grad=True) for a in start]
    tr = m.forward(*params)
    J64 = 0.5 * ((tr - torch.tensor(obs.astype(np.float64))) ** 2).sum()
    J64.backward()
    assert abs(J - float(J64.detach())) <= 1e-4 * abs(float(J64.detach()))
    for name, got, p in (("rho", g_rho, params[0]), ("cp", g_cp, params[1]), ("cs", g_cs, params[2]),
                         ("tau", coef["tau"], params[3])):
        ref = p.grad.numpy()
        rel = oracle.rel_l2(got, ref)
        print(f"2 mechanisms, d/d{name}: rel-L2 vs FP64 autograd {rel:.2e}", flush=True)
        assert np.abs(ref).max() > 0
        assert rel < 1e-3, name
