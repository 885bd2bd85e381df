"""The host setup mirror (taichi_sim_amd.sim_support) against the reference's
own setup modules, through the data they produced
(tests/golden/setup_fixtures.npz, written by make_setup_fixtures.py).

Every comparison is exact: ROI sizes, steps and decimals, the CPML (a, b, k)
profiles on the full and half grids of every axis, nearest-grid indices,
probe points and receiver indices, receiver delays, float32 source terms
before and after an emission law, .law text and parsing, the FD coefficient
table and the default Zener constants.
"""
import json
import os
import tempfile

import numpy as np
import pytest

from taichi_sim_amd.sim_support import coefs_Lui
from taichi_sim_amd.sim_support.attenuation import AttenuationCoefficients
from taichi_sim_amd.sim_support.emission_law import EmissionLaw
from taichi_sim_amd.sim_support.simul_classes import (SimulationROI, SimulationProbeLinearArray,
                                                      SimulationProbePoint, SimulationProbeMonoCirc)

from tests.golden.setup_cases import setup_cases

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "setup_fixtures.npz"), allow_pickle=False)
CASES = sorted({k.split("/")[0] for k in FIX.files if "/" in k})


CFG = setup_cases()


def _roi(name):
    cfg, st, maps = CFG[name]
    assert json.loads(str(FIX[name + "/roi_cfg"])) == json.loads(json.dumps(cfg["roi"])), "config drifted"
    acc = int(FIX[name + "/acc"])
    if maps is not None:
        return SimulationROI(**cfg["roi"], pad=acc - 1, rho_map=maps["rho"]), cfg, st
    return SimulationROI(**cfg["roi"], pad=acc - 1), cfg, st


def test_fixture_has_all_cases():
    assert CASES == sorted(CFG)


@pytest.mark.parametrize("name", sorted(CFG))
def test_roi_and_cpml_profiles(name):
    roi, cfg, st = _roi(name)
    p = name + "/"
    assert [roi.get_nx(), roi.get_ny(), roi.get_nz()] == FIX[p + "shape"].tolist()
    assert np.array_equal(np.array([roi.get_dx(), roi.get_dy(), roi.get_dz()], dtype=np.float64), FIX[p + "steps"])
    assert tuple(np.asarray(roi.get_dec()).tolist()) == tuple(FIX[p + "dec"].tolist())
    dt, cp, alpha_max = FIX[p + "pml_args"]
    for axis in (("x", "z") if st == "2d" else ("x", "y", "z")):
        for grid in ("f", "h"):
            got = np.stack(roi.calc_pml_array(axis=axis, grid=grid, dt=float(dt), cp=float(cp),
                                              alpha_max=float(alpha_max)))
            ref = FIX[p + f"pml_{axis}{grid}"]
            assert got.dtype == ref.dtype and got.shape == ref.shape, (axis, grid)
            assert np.array_equal(got, ref), (axis, grid, np.abs(got - ref).max())


@pytest.mark.parametrize("name", sorted(CFG))
def test_nearest_grid_index(name):
    roi, _, _ = _roi(name)
    pts, idx = FIX[name + "/nearest_pts"], FIX[name + "/nearest_idx"]
    assert (idx[:, 0] != -999).sum() >= 20 and (idx[:, 0] == -999).sum() >= 1
    for q, want in zip(pts, idx):
        try:
            got = list(np.asarray(roi.get_nearest_grid_idx(q)).reshape(-1))
        except IndexError:
            got = [-999, -999, -999]
        assert got == want.tolist(), (q, got, want)


@pytest.mark.parametrize("name", sorted(CFG))
def test_probe_points_sources_and_delays(name):
    roi, cfg, st = _roi(name)
    dt = float(FIX[name + "/pml_args"][0])
    cls = {"linear": SimulationProbeLinearArray, "point": SimulationProbePoint, "mono": SimulationProbeMonoCirc}
    for i, pr_cfg in enumerate(cfg["probes"][:6]):
        q = name + f"/probe{i}_"
        kind, kw = next(iter(pr_cfg.items()))
        assert str(FIX[q + "kind"]) == kind
        pr = cls[kind](**kw, dec=roi.get_dec())
        for d_ in ("e", "r"):
            pts, idx = pr.get_points_roi(roi, simul_type=st, dir=d_)
            pts = np.asarray(pts, dtype=np.int64).reshape(-1, 3) if len(pts) else np.zeros((0, 3), np.int64)
            assert np.array_equal(pts, FIX[q + f"points_{d_}"]), (i, d_)
            assert np.array_equal(np.asarray(idx, dtype=np.int64).reshape(-1), FIX[q + f"index_{d_}"]), (i, d_)
        assert np.array_equal(np.asarray(pr.get_delay_rx(), dtype=np.float64).reshape(-1), FIX[q + "delay_rx"])
        samples = int(FIX[q + "samples"])
        if q + "src" in FIX.files:
            got = np.asarray(pr.get_source_term(samples=samples, dt=dt), dtype=np.float32)
            assert np.array_equal(got, FIX[q + "src"]), (i, np.abs(got - FIX[q + "src"]).max())
        if kind == "linear":
            pr.set_t0(FIX[q + "law"])
            got = np.asarray(pr.get_source_term(samples=samples, dt=dt), dtype=np.float32)
            assert np.array_equal(got, FIX[q + "src_law"]), i


def test_law_files():
    with tempfile.TemporaryDirectory() as tmp:
        root = os.path.join(tmp, "law")
        EmissionLaw().write_law(root, FIX["law_in"], emitter_amp=FIX["law_amp_in"])
        with open(root + ".law") as f:
            assert f.read() == str(FIX["law_text"])
        with open(root + ".law", "w") as f:       # parse the reference-written text
            f.write(str(FIX["law_text"]))
        delay, amp = EmissionLaw().read_law(root + ".law")
    assert np.array_equal(delay, FIX["law_delay"]) and np.array_equal(amp, FIX["law_amp"])


def test_coefficients_and_default_zener():
    assert json.loads(str(FIX["coefs_lui"])) == [list(map(float, c)) for c in coefs_Lui]
    att = AttenuationCoefficients()
    for k in ("_tau_epsilon_p", "_tau_sigma_p", "_tau_epsilon_s", "_tau_sigma_s"):
        assert np.array_equal(getattr(att, k), FIX["att" + k])
    assert np.array_equal(np.array([att._sum_alpha_p, att._sum_alpha_s], dtype=np.float64), FIX["att_sum_alpha"])
