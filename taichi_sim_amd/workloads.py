"""Synthetic workloads in the reference's configuration format.

Each builder returns a configuration dict accepted by ``Simulator`` /
``Simulator3D`` (material maps as arrays instead of .npy paths).  Grid spacing
is 0.1 mm, every side has a 20-cell CPML, and dt sits at CFL 0.5:
dt = 0.5 dx / (cp_max sqrt(ndim)).  Random media are seeded.

* ``elastic_2d_point``: 2D isotropic elastic, aluminium, point force at the
  centre, a line of receivers (the CPU-runnable case).
* ``visco_2d_array``: 2D Zener (L=3, constant Q fitted over 2-8 MHz) steel
  plate with random circular voids and a focusing linear array.
* ``focal_laws`` / ``multi_shot_2d``: emission laws (one focal point per
  shot) for ``visco_2d_array`` specimens of several sizes -- the multi-shot
  workload of a full focal-law scan.
* ``fmc_2d``: full-matrix capture (law k fires element k only) on
  ``visco_2d_array`` specimens of several sizes.
* ``adjoint_2d``: the inversion test problem -- a true model with seeded
  Gaussian velocity anomalies, a homogeneous starting model, FMC shots.
* ``visco_3d_block``: 3D Zener two-layer block (aluminium over steel) with
  random ellipsoidal density inclusions, a circular piston on the top face and
  a square grid of surface receivers.
"""
import numpy as np

AL = (6.32, 3.13, 2700.0)
STEEL = (5.9, 3.2, 7850.0)
PML = 20
DX = 0.1


def cfl_dt(cp_max, ndim, dx=DX):
    return float(np.float32(0.5 * dx / (cp_max * np.sqrt(ndim))))


def _roi(n_full, ndim, acc=2):
    """ROI giving a full grid of n_full cells per axis (an int, or a tuple
    (nx, nz) in 2D / (nx, ny, nz) in 3D): n = len + 2*PML + 2*pad."""
    pad = acc - 1
    if np.isscalar(n_full):
        n_full = (n_full,) * ndim
    lens = [int(n) - 2 * PML - 2 * pad for n in n_full]
    lx, lz = lens[0], lens[-1]
    roi = {"coord_ref": [[0.0, 0.0, 0.0]], "width": lx * DX, "w_len": lx,
           "height": lz * DX, "h_len": lz, "depth": 0.0, "d_len": 1,
           "len_pml_xmin": PML, "len_pml_xmax": PML, "len_pml_ymin": PML if ndim == 3 else 10,
           "len_pml_ymax": PML if ndim == 3 else 10, "len_pml_zmin": PML, "len_pml_zmax": PML,
           "rcoef_pml": 0.0001, "npower_pml": 2.0, "k_max_pml": 1.0}
    if ndim == 3:
        roi["depth"] = lens[1] * DX
        roi["d_len"] = lens[1]
    return roi, tuple(lens)


def _configs(visco, **extra):
    c = {"n_iter": 1, "show_anim": 0, "show_debug": 0, "save_results": 0, "save_bscan": 0,
         "viscoelastic_attn": int(visco), "emission_laws": "", "results_dir": ""}
    c.update(extra)
    return c


def elastic_2d_point(n=256, steps=1000, n_rec=128, material=AL):
    roi, (L, Lz) = _roi(n, 2)
    cp, cs, rho = material
    w, h = L * DX, Lz * DX
    probes = [{"point": {"coord_center": [w / 2, 0.0, h / 2], "freq": 5.0, "bw": 0.5, "gain": 1.0,
                         "emitter": "True", "receiver": "False", "t0_emission": 1.0}}]
    xs = np.linspace(0.1 * w, 0.8 * w, n_rec)
    for x in xs:
        probes.append({"point": {"coord_center": [float(np.floor(x * 100) / 100), 0.0, float(np.floor(0.25 * h * 100) / 100)],
                                 "freq": 5.0, "bw": 0.5, "gain": 1.0, "emitter": "False",
                                 "receiver": "True", "t0_reception": 0.0}})
    return {"roi": roi, "probes": probes,
            "simul_params": {"time_steps": steps, "dt": cfl_dt(cp, 2), "acc": 2, "it_display": 100},
            "simul_configs": _configs(False),
            "specimen_params": {"cp": cp, "cs": cs, "rho": rho}}


def _voids_2d(shape, n_voids, seed, margin_cells):
    rng = np.random.default_rng(seed)
    nx, nz = shape
    mask = np.zeros(shape, dtype=bool)
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    for _ in range(n_voids):
        r = rng.uniform(0.3, 1.5) / DX
        cx = rng.uniform(margin_cells, nx - margin_cells)
        cz = rng.uniform(min(margin_cells + 40, nz // 3), nz - margin_cells)
        mask |= (X - cx) ** 2 + (Z - cz) ** 2 <= r * r
    return mask


def visco_2d_array(n=2048, steps=4000, n_elem=64, n_voids=16, seed=2, q=50.0, emitters="all",
                   t0_emission=None):
    roi, (L, Lz) = _roi(n, 2)
    cp, cs, rho = STEEL
    mask = _voids_2d((L, Lz), n_voids, seed, margin_cells=min(20, min(L, Lz) // 8))
    cp_m = np.where(mask, 0.0, cp).astype(np.float32)
    cs_m = np.where(mask, 0.0, cs).astype(np.float32)
    rho_m = np.where(mask, 0.0, rho).astype(np.float32)
    w = L * DX
    pitch = 0.5
    if t0_emission is None:
        rng = np.random.default_rng(seed + 100)
        fx, fz = rng.uniform(0.3 * w, 0.7 * w), rng.uniform(0.3 * Lz * DX, 0.7 * Lz * DX)
        ex = w / 2 + (np.arange(n_elem) - (n_elem - 1) / 2) * pitch
        tof = np.sqrt((ex - fx) ** 2 + fz ** 2) / cp
        t0_emission = [float(v) for v in (tof.max() - tof + 1.0)]
    probes = [{"linear": {"id": "array", "num_elem": n_elem, "dim_a": 0.42, "dim_p": 10.0, "inter_elem": 0.08,
                          "freq": 5.0, "bw": 0.5, "gain": 1e2, "pulse_type": "gaussian",
                          "coord_center": [w / 2, 0.0, 0.0], "emitters": emitters,
                          "t0_emission": t0_emission, "receivers": "all", "t0_reception": 0.0}}]
    return {"roi": roi, "probes": probes,
            "simul_params": {"time_steps": steps, "dt": cfl_dt(cp, 2), "acc": 2, "it_display": 100},
            "simul_configs": _configs(True),
            "specimen_params": {"cp": cp, "cs": cs, "rho": rho, "cp_map": cp_m, "cs_map": cs_m, "rho_map": rho_m},
            "attenuation": {"q_p": q, "q_s": q, "f_min": 2.0, "f_max": 8.0, "n_sls": 3}}


def focal_laws(n_elem, n_shots, width, height, cp, pitch=0.5, seed=7):
    """(n_shots, n_elem) emission delays in microseconds: shot i focuses the
    array (centred on the top face, element pitch ``pitch`` mm) at a seeded
    point in the middle of the specimen (the retE column of a .law file)."""
    rng = np.random.default_rng(seed)
    ex = width / 2 + (np.arange(n_elem) - (n_elem - 1) / 2) * pitch
    laws = np.empty((n_shots, n_elem))
    for i in range(n_shots):
        fx = rng.uniform(0.25 * width, 0.75 * width)
        fz = rng.uniform(0.3 * height, 0.7 * height)
        tof = np.sqrt((ex - fx) ** 2 + fz ** 2) / cp
        laws[i] = tof.max() - tof + 1.0
    return laws


def multi_shot_2d(sizes=(512, 1024, 1536, 2048), n_shots=64, steps=3000, n_elem=32, seed=2):
    """[(config, laws)] per specimen: a ``visco_2d_array`` steel plate of each
    size (full grid incl. CPML) and ``n_shots`` focal laws for its array."""
    out = []
    for k, n in enumerate(sizes):
        cfg = visco_2d_array(n=n, steps=steps, n_elem=n_elem, n_voids=max(2, n // 128), seed=seed + k)
        pr = cfg["probes"][0]["linear"]
        w = cfg["roi"]["width"]
        h = cfg["roi"]["height"]
        pitch = pr["dim_a"] + pr["inter_elem"]
        out.append((cfg, focal_laws(n_elem, n_shots, w, h, STEEL[0], pitch=pitch, seed=seed + 100 + k)))
    return out


def fmc_2d(sizes=(512, 1024, 1536, 2048), n_elem=16, steps=3000, seeds=(40, 41, 42, 43), n_voids=16,
           silent_t0=1.0e6):
    """[(config, laws)] per specimen for a full-matrix capture: a
    ``visco_2d_array`` steel plate of each size whose ``n_elem`` elements all
    emit and receive, and ``n_elem`` emission laws where law k fires element k
    at t0 = 1 us and delays every other element by ``silent_t0`` us, past any
    simulated time, so its pulse is exactly zero.  (A per-shot emitter mask
    would instead leave a single emitting element, which the WGSL source
    indexing reads as column 0 only: shots k > 0 would be silent.)"""
    out = []
    for n, seed in zip(sizes, seeds):
        cfg = visco_2d_array(n=n, steps=steps, n_elem=n_elem, n_voids=n_voids, seed=seed,
                             t0_emission=[1.0] * n_elem)
        laws = np.full((n_elem, n_elem), float(silent_t0))
        np.fill_diagonal(laws, 1.0)
        out.append((cfg, laws))
    return out


def adjoint_2d(n=1024, steps=3000, n_shots=8, n_anom=8, seed=5, amp=0.05, depth=(4.0, 30.0)):
    """(config, laws, true_maps): a steel plate of n^2 cells (incl. CPML), an
    ``n_shots``-element array on the top face firing one element per shot
    (law k fires element k), and the true model: the background with
    ``n_anom`` seeded Gaussian anomalies of +-amp in cp and cs (sigma 2-5 mm)
    centred ``depth`` mm below the array -- within reach of a reflection in
    3000 steps (18 us of two-way time in steel is ~53 mm) -- as (rho, cp, cs)
    maps over the full grid.  The configuration itself holds the homogeneous
    background -- the starting model of an inversion."""
    cfg, laws = fmc_2d(sizes=(n,), n_elem=n_shots, steps=steps, seeds=(seed,), n_voids=0)[0]
    L, Lz = cfg["roi"]["w_len"], cfg["roi"]["h_len"]
    nx, nz = L + 2 * PML + 2, Lz + 2 * PML + 2
    rng = np.random.default_rng(seed)
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    pert_p = np.zeros((nx, nz))
    pert_s = np.zeros((nx, nz))
    z_top = PML + 1                                    # grid row of the top ROI face (the array)
    for _ in range(n_anom):
        cx = rng.uniform(0.2, 0.8) * nx
        cz = z_top + rng.uniform(depth[0], depth[1]) / DX
        sig = rng.uniform(2.0, 5.0) / DX
        g = np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (2 * sig * sig))
        pert_p += rng.choice([-1.0, 1.0]) * amp * g
        pert_s += rng.choice([-1.0, 1.0]) * amp * g
    cp, cs, rho = STEEL
    true = (np.full((nx, nz), rho, np.float32), (cp * (1 + pert_p)).astype(np.float32),
            (cs * (1 + pert_s)).astype(np.float32))
    return cfg, laws, true


def visco_3d_block(n=512, steps=2000, n_incl=24, seed=3, q=50.0, n_rec_side=32, rec_pitch=1.0, t0=1.0):
    roi, (L, Ly, Lz) = _roi(n, 3)
    rng = np.random.default_rng(seed)
    al_cp, al_cs, al_rho = AL
    st_cp, st_cs, st_rho = STEEL
    # layer interface at grid z = nz/2 (ROI index nz/2 - PML - pad)
    zi = (Lz + 2 * PML + 2) // 2 - PML - 1
    z = np.arange(Lz)
    top = (z < zi)[None, None, :]
    shape = (L, Ly, Lz)
    cp_m = np.broadcast_to(np.where(top, al_cp, st_cp), shape).astype(np.float32)
    cs_m = np.broadcast_to(np.where(top, al_cs, st_cs), shape).astype(np.float32)
    rho_m = np.broadcast_to(np.where(top, al_rho, st_rho), shape).astype(np.float32).copy()
    ext = np.array(shape)
    for _ in range(n_incl):
        ax = rng.uniform(0.5, 2.0, size=3) / DX
        c = rng.uniform(0.15, 0.85, size=3) * ext
        s = np.float32(rng.uniform(0.5, 2.0))
        lo = np.maximum(np.floor(c - ax).astype(int), 0)
        hi = np.minimum(np.ceil(c + ax).astype(int) + 1, ext)
        gx, gy, gz = np.ogrid[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        inside = ((gx - c[0]) / ax[0]) ** 2 + ((gy - c[1]) / ax[1]) ** 2 + ((gz - c[2]) / ax[2]) ** 2 <= 1.0
        blk = rho_m[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        blk[inside] *= s
    w, d = L * DX, Ly * DX
    radius = min(0.125, 0.3 * min(w, d) / 25.4)            # inches (simul_classes.py:548)
    rec_pitch = min(rec_pitch, 0.8 * min(w, d) / n_rec_side)
    probes = [{"mono": {"coord_center": [w / 2, d / 2, 0.0], "radius": radius, "freq": 5.0, "bw": 0.5,
                        "gain": 1.0, "emitters": "True", "receivers": "False", "t0_emission": float(t0),
                        "t0_reception": 0.0}}]
    off = (n_rec_side - 1) / 2 * rec_pitch
    for i in range(n_rec_side):
        for j in range(n_rec_side):
            probes.append({"point": {"coord_center": [float(np.round(w / 2 - off + i * rec_pitch, 2)),
                                                      float(np.round(d / 2 - off + j * rec_pitch, 2)), 0.0],
                                     "freq": 5.0, "bw": 0.5, "gain": 1.0, "emitter": "False",
                                     "receiver": "True", "t0_reception": 0.0}})
    return {"roi": roi, "probes": probes,
            "simul_params": {"time_steps": steps, "dt": cfl_dt(max(al_cp, st_cp), 3), "acc": 2, "it_display": 100},
            "simul_configs": _configs(True),
            "specimen_params": {"cp": al_cp, "cs": al_cs, "rho": al_rho, "cp_map": cp_m, "cs_map": cs_m,
                                "rho_map": rho_m},
            "attenuation": {"q_p": q, "q_s": q, "f_min": 2.0, "f_max": 8.0, "n_sls": 3}}
