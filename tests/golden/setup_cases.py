"""Configurations whose host setup is compared with the reference's own setup
modules (make_setup_fixtures.py writes the fixture, test_setup_fixtures.py
checks the mirror against it): (name, config, simulation type, maps)."""
import ast
import os

import numpy as np

from taichi_sim_amd import workloads

HERE = os.path.dirname(os.path.abspath(__file__))


def setup_cases():
    d = np.load(os.path.join(HERE, "pedro_teste_aco.npz"), allow_pickle=False)
    out = {"pedro": (ast.literal_eval(str(d["config"])), "2d", {"rho": d["rho"]})}
    cfg = workloads.elastic_2d_point(n=256, steps=400, n_rec=16)
    for i, p in enumerate(cfg["probes"][1:]):            # receivers with delays
        p["point"]["t0_reception"] = float(np.round(0.137 * (i + 1), 3))
    out["el256"] = (cfg, "2d", None)
    cfg = workloads.visco_2d_array(n=200, steps=300, n_elem=12, n_voids=0)
    cfg["probes"][0]["linear"]["t0_reception"] = [float(v) for v in np.round(np.linspace(0.0, 1.9, 12), 3)]
    out["va200"] = (cfg, "2d", None)
    out["vb64"] = (workloads.visco_3d_block(n=64, steps=100, n_incl=2, n_rec_side=4), "3d", None)
    return out


def probe_points(cfg, st, seed):
    """40 seeded points over (and a little beyond) the ROI's physical extent,
    some on and half-way between grid lines."""
    r = cfg["roi"]
    ref0 = np.asarray(r.get("coord_ref", [[0, 0, 0]]), dtype=np.float64).reshape(3)
    ext = np.array([r["width"], r.get("depth", 0.0), r["height"]], dtype=np.float64)
    rng = np.random.default_rng(seed)
    pts = ref0 + rng.uniform(-0.05, 1.05, size=(40, 3)) * ext
    pts[:12] = np.round(pts[:12], 1)
    pts[12:18] = np.round(pts[12:18], 1) + 0.05
    if st == "2d":
        pts[:, 1] = ref0[1]
    return pts.astype(np.float32)
