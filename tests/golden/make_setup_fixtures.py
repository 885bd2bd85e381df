"""Write tests/golden/setup_fixtures.npz from the reference's own setup modules.

The host setup layer -- SimulationROI step rounding and calc_pml_array's
np.round chain, get_nearest_grid_idx tie-breaking, probe point rasterisation,
float32 source-term arithmetic, receiver indices and delays, .law files, the
FD coefficient table and the default Zener constants -- has no reference
test and no closed form that pins its rounding.  This script imports the
reference's numpy/scipy-only modules (sim_support/__init__.py,
simul_classes.py, gauss_wavelets.py, emission_law.py, attenuation.py) in
this container, evaluates them on the configurations the tests use, and
stores inputs and outputs as data.  Nothing of the reference is copied: the
.npz holds arrays and strings only.

Usage: python tests/golden/make_setup_fixtures.py [/path/to/reference]
"""
import json
import os
import sys
import tempfile

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "setup_fixtures.npz")

sys.path.insert(0, REF)
import sim_support as ref_init                                          # noqa: E402
from sim_support.attenuation import AttenuationCoefficients            # noqa: E402
from sim_support.emission_law import EmissionLaw                        # noqa: E402
from sim_support.simul_classes import (SimulationROI, SimulationProbeLinearArray,  # noqa: E402
                                       SimulationProbePoint, SimulationProbeMonoCirc)
sys.path.remove(REF)
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)
from setup_cases import setup_cases, probe_points                       # noqa: E402  (configurations only)


def cases():
    for name, (cfg, st, maps) in setup_cases().items():
        yield name, cfg, st, maps


out = {"coefs_lui": np.array(json.dumps([list(map(float, c)) for c in ref_init.coefs_Lui]))}
att = AttenuationCoefficients()
for k in ("_tau_epsilon_p", "_tau_sigma_p", "_tau_epsilon_s", "_tau_sigma_s"):
    out["att" + k] = np.asarray(getattr(att, k))
out["att_sum_alpha"] = np.array([att._sum_alpha_p, att._sum_alpha_s], dtype=np.float64)

rng = np.random.default_rng(7)
for name, cfg, st, maps in cases():
    acc = int(cfg["simul_params"].get("acc", 2))
    roi_kw = dict(cfg["roi"])
    if maps is not None:
        roi = SimulationROI(**roi_kw, pad=acc - 1, rho_map=maps["rho"])
    else:
        roi = SimulationROI(**roi_kw, pad=acc - 1)
    p = name + "/"
    out[p + "roi_cfg"] = np.array(json.dumps(roi_kw))
    out[p + "acc"] = np.array(acc)
    out[p + "shape"] = np.array([roi.get_nx(), roi.get_ny(), roi.get_nz()])
    out[p + "steps"] = np.array([roi.get_dx(), roi.get_dy(), roi.get_dz()], dtype=np.float64)
    out[p + "dec"] = np.array(roi.get_dec())
    dt = float(np.float32(cfg["simul_params"]["dt"]))
    cp = float(cfg["specimen_params"]["cp"])
    alpha_max = float(np.pi * 5.0)
    out[p + "pml_args"] = np.array([dt, cp, alpha_max])
    for axis in (("x", "z") if st == "2d" else ("x", "y", "z")):
        for grid in ("f", "h"):
            a, b, k = roi.calc_pml_array(axis=axis, grid=grid, dt=dt, cp=cp, alpha_max=alpha_max)
            out[p + f"pml_{axis}{grid}"] = np.stack([a, b, k])
    # nearest grid index: points inside, on the grid, half-way and outside
    pts = probe_points(cfg, st, seed=len(name))
    nearest = []
    for q in pts:
        try:
            nearest.append(list(np.asarray(roi.get_nearest_grid_idx(q)).reshape(-1)))
        except IndexError:
            nearest.append([-999, -999, -999])
    out[p + "nearest_pts"] = pts
    out[p + "nearest_idx"] = np.array(nearest, dtype=np.int64)
    # probes
    n_steps = int(cfg["simul_params"]["time_steps"])
    samples = min(n_steps, 1200)
    probes = []
    for pr in cfg["probes"]:
        kind, kw = next(iter(pr.items()))
        cls = {"linear": SimulationProbeLinearArray, "point": SimulationProbePoint,
               "mono": SimulationProbeMonoCirc}[kind]
        probes.append((kind, cls(**kw, dec=roi.get_dec())))
    for i, (kind, pr) in enumerate(probes[:6]):
        q = p + f"probe{i}_"
        out[q + "kind"] = np.array(kind)
        for d_ in ("e", "r"):
            pts_, idx_ = pr.get_points_roi(roi, simul_type=st, dir=d_)
            out[q + f"points_{d_}"] = np.asarray(pts_, dtype=np.int64).reshape(-1, 3) if len(pts_) else \
                np.zeros((0, 3), np.int64)
            out[q + f"index_{d_}"] = np.asarray(idx_, dtype=np.int64).reshape(-1)
        out[q + "delay_rx"] = np.asarray(pr.get_delay_rx(), dtype=np.float64).reshape(-1)
        if any(pr.emitters):
            if kind == "point":
                out[q + "src"] = np.asarray(pr.get_source_term(samples=samples, dt=dt), dtype=np.float32)
            else:
                out[q + "src"] = np.asarray(pr.get_source_term(samples=samples, dt=dt), dtype=np.float32)
        if kind == "linear":
            law = rng.uniform(0.5, 2.0, size=pr.num_elem)
            pr.set_t0(law)
            out[q + "law"] = law
            out[q + "src_law"] = np.asarray(pr.get_source_term(samples=samples, dt=dt), dtype=np.float32)
        out[q + "samples"] = np.array(samples)

# .law files: the reference's writer and reader
law = np.round(rng.uniform(0.0, 3.0, size=(3, 8)), 4)
amp = np.round(rng.uniform(0.5, 1.0, size=(3, 8)), 3)
with tempfile.TemporaryDirectory() as tmp:
    root = os.path.join(tmp, "law")
    EmissionLaw().write_law(root, law, emitter_amp=amp)
    with open(root + ".law") as f:
        out["law_text"] = np.array(f.read())
    delay, amp_read = EmissionLaw().read_law(root + ".law")
out["law_in"] = law
out["law_amp_in"] = amp
out["law_delay"] = np.asarray(delay)
out["law_amp"] = np.asarray(amp_read)

np.savez_compressed(OUT, **out)
print(OUT, os.path.getsize(OUT), len(out), "arrays")
