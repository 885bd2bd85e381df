"""Headline benchmark: 3D viscoelastic (Zener L=3) FP32 time loop, 512^3 cells.

Workload "3D viscoelastic Zener FD 512^3": 512^3 full grid
incl. 20-cell CPML on every face, 0.1 mm cells, aluminium over steel with
seeded ellipsoidal density inclusions, L=3 mechanisms fitted to Q=50 over
2-8 MHz, 5 MHz circular piston on the top face, 32x32 surface receivers,
dt at CFL 0.5.  A step = stress pass + velocity pass + source/receiver pass
over the whole grid; inputs are resident in HBM before the timed region.

``--workload shots`` measures the multi-shot path instead: a full-matrix
capture on 4 steel plates with voids (512^2, 1024^2, 1536^2, 2048^2, Zener
L=3, 16-element array, shot k fires element k), all 64 simulations advanced
as one batch (one launch per pass per step), against the same 64 simulations
run one after another as the reference runs its emission laws.
value = 3000-step shots per second.

``--workload 2d`` measures the 2D path on its headline configuration: 2D
viscoelastic (Zener L=3) 2048^2 incl. 20-cell CPML, steel plate with 16
seeded voids, 64-element array with a seeded focusing law, 64 receivers.
value = Gcell-updates/s; roofline of the 2D stress pass; the CPU oracle on a
bounded sample beside it.

``--workload adjoint`` measures the gradient: 2D viscoelastic 1024^2, L=3,
3000 steps, 8 FMC shots, observed traces from a true model with 8 seeded
Gaussian velocity anomalies (+-5%, 4-30 mm under the array), gradient at the homogeneous starting
model with sqrt(N) checkpointing.  value = time(checkpointed forward +
recompute + adjoint) / time(plain forward), per shot, device time.

Prints one JSON line.  value = Gcell-updates/s of the whole job (one GPU).
The default line also carries the rest of the metric and the 2D line as
extra keys: shots_per_s (batched FMC against the per-shot loop),
adjoint_over_forward (one shot, sqrt(N) checkpoints) and visco2d.
The headline runs the FAST arithmetic (FP32 with the FD coefficients
pre-divided by the spacing, within 1e-5 rel-L2 of the reference's operation
order); STRICT, the reference's order term for term, is reported beside it.
roofline: the dominant kernel (stress pass) -- algorithmic bytes per launch /
its average duration measured here with HIP events on the library's stream.
cpu_baseline: the FP32 CPU oracle (oracle/, the reference's update order
restated in C, OpenMP) on a bounded sample of the same grid.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0

# The result line goes to the real stdout; everything else the run prints
# (including the simulator setup's own messages) goes to stderr, so stdout
# holds exactly one JSON line.
_STDOUT = sys.stdout


def emit(obj):
    print(json.dumps(obj), file=_STDOUT, flush=True)

# Launch shapes (ty, non-temporal, x-chunk, tile order) the 3D autotuner
# chooses from in this benchmark: the fastest ones on MI355X, each with its
# stress-kernel HBM traffic measured below.
SHAPES_3D = [(4, 0, 4, 2), (4, 0, 8, 2), (4, 0, 16, 2), (1, 0, 4, 2), (1, 0, 8, 2), (1, 0, 16, 2),
             (4, 0, 4, 1), (4, 0, 8, 1), (1, 0, 4, 1), (1, 0, 8, 1)]
# the velocity pass also tries its LDS-staged y-neighbour shapes (nt bit 1)
VELOCITY_SHAPES_3D = SHAPES_3D + [(4, 2, 8, 2), (4, 2, 16, 2)]
# HBM bytes per launch of the 512^3 viscoelastic FAST stress kernel by launch
# shape, measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE over
# `bench.py --shape-sweep` (scaled by 2.0 and 1.0 from the copy-kernel
# calibration); the table and commands are in profiles/visco3d_512cube_kernels.md.
STRESS_TRAFFIC_512 = {(4, 0, 4, 2): 31.86e9, (4, 0, 8, 2): 32.02e9, (4, 0, 16, 2): 32.79e9,
                      (1, 0, 4, 2): 31.78e9, (1, 0, 8, 2): 31.64e9, (1, 0, 16, 2): 32.25e9,
                      (4, 0, 4, 1): 32.10e9, (4, 0, 8, 1): 32.13e9, (1, 0, 4, 1): 32.05e9,
                      (1, 0, 8, 1): 31.75e9}
# The same for the 2048^2 2D viscoelastic stress kernel at its tuned shape
# (profiles/visco2d_2048_kernels.md).
STRESS2D_TRAFFIC_2048 = 489.2e6


def bytes_per_cell(g, L):
    """Algorithmic bytes per cell and step: every state field read once and
    written once, materials read once, psi only inside its axis's CPML band."""
    n = g.shape
    band = sum(48.0 * g.band_cells()[a] / n[a] for a in range(3))      # 6 psi x (r+w) x 4 B per axis
    step = 4.0 * (2 * (3 + 6 + 6 * L) + 3) + band
    stress = 4.0 * (3 + 6 + 6 * L + 3 + 6 + 6 * L) + band / 2          # v, sigma, r, rho/cp/cs in; sigma, r out
    velocity = 4.0 * (6 + 3 + 1 + 3) + band / 2                         # sigma, v, rho in; v out
    return step, stress, velocity


def bytes_per_cell_2d(g, L):
    """2D: every state field read and written once, materials read once, psi
    (4 slots per axis, split between the passes) only inside its band."""
    n = g.shape
    band = sum(32.0 * g.band_cells()[a] / n[a] for a in range(2))      # 4 psi x (r+w) x 4 B per axis
    step = 4.0 * (2 * (2 + 3 + 3 * L) + 3) + band
    stress = 4.0 * (2 + 3 + 3 * L + 3 + 3 + 3 * L) + band / 2          # v, sigma, r, rho/cp/cs in; sigma, r out
    velocity = 4.0 * (3 + 2 + 1 + 2) + band / 2                         # sigma, v, rho in; v out
    return step, stress, velocity


def bench_2d(args):
    """2D viscoelastic time loop at 2048^2 (see module doc)."""
    emit(measure_2d(args.n2d, args.steps, args.warmup, args.math2d, args.profile_steps, args.cpu_steps,
                    args.fused_sweep))


def measure_2d(n, steps, warmup, math, profile_steps, cpu_steps, fused_sweep=False):
    """The 2D line: Gcell-updates/s of the n^2 viscoelastic loop, its stress
    pass against the roofline and the CPU oracle on a bounded sample."""
    from taichi_sim_amd import autotune, workloads
    from taichi_sim_amd.backend import attenuation_from_config
    from taichi_sim_amd.engine import VSESim
    from taichi_sim_amd.sim_support.simulator import Simulator

    cfg = workloads.visco_2d_array(n=n, steps=warmup + steps)
    sim = Simulator(cfg)
    att = attenuation_from_config(cfg)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    g.set_math(math)
    ncell = int(np.prod(g.shape))
    print(f"# setup done: grid {g.shape}, {g.device_bytes() / 2**30:.2f} GiB on {g.device_name()}",
          file=sys.stderr, flush=True)
    best = autotune.tune(g, steps=3, cache=None)           # before the timed region; it advances the state
    shapes = {k: list(best[k]) for k in ("stress", "velocity")}
    print(f"# launch shapes (ty, nt, rows, remap, step ms): {shapes}", file=sys.stderr, flush=True)
    if warmup:
        g.reset()
        g.run(warmup, check_every=warmup)
    g.reset()
    g.run(min(steps, 256), check_every=256)           # capture the graph outside the timed region
    g.reset()
    t0 = time.perf_counter()
    dev_ms = g.run(steps, check_every=500, timed=True)
    wall = time.perf_counter() - t0
    gcell = ncell * steps / wall / 1e9
    k_ms = g.profile(profile_steps)
    b_step, b_stress, b_vel = bytes_per_cell_2d(g, g.n_sls)
    achieved = b_stress * ncell / (k_ms[0] * 1e-3) / 1e9
    fused = None
    if fused_sweep:
        # the opt-in fused schedule (velocity of step n + stress of step n+1 per
        # launch) at several x-chunks, timed like the line above, then the
        # two-pass step again in the same process
        fused = {}
        for xc in (8, 16, 32, 64):
            g.set_tuning(2, 4, 0, xc, 1)
            g.set_schedule("fused")
            g.reset()
            g.run(min(steps, 256), check_every=256)
            g.reset()
            t = time.perf_counter()
            g.run(steps, check_every=500)
            fused[f"xchunk {xc}"] = round(ncell * steps / (time.perf_counter() - t) / 1e9, 3)
            print(f"# fused schedule, x-chunk {xc}: {fused[f'xchunk {xc}']} Gcell/s", file=sys.stderr, flush=True)
        g.set_schedule("two-pass")
        g.reset()
        t = time.perf_counter()
        g.run(steps, check_every=500)
        fused["two-pass again"] = round(ncell * steps / (time.perf_counter() - t) / 1e9, 3)
    g.close()
    print(f"# timed {steps} steps: wall {wall:.3f} s, device {dev_ms / 1e3:.3f} s", file=sys.stderr, flush=True)

    cpu = None
    if cpu_steps > 0:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        steps_cpu = 10 * cpu_steps
        o = oracle.Oracle(sim, variant="wgsl", att=att, n_threads=threads)
        t = time.perf_counter()
        o.run(steps_cpu)
        cpu_s = time.perf_counter() - t
        cpu = {"value": round(ncell * steps_cpu / cpu_s / 1e9, 6), "unit": "Gcell-upd/s", "cores": threads,
               "kind": "port", "sample": f"{steps_cpu} steps of the same {n}^2 grid, FP32 C oracle "
                                         f"(reference update order, OpenMP), {cpu_s:.1f} s"}
        del o

    line = {
        "metric": "Gcell-upd/s (2D viscoelastic Zener L=3 FP32 forward)",
        "value": round(gcell, 4), "unit": "Gcell-upd/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(1e3 * wall / steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"2D viscoelastic Zener FD {n}^2 (incl. 20-cell CPML), L=3 Q=50, steel plate with "
                               f"16 seeded voids, 64-element array with a focusing law, 64 receivers",
                   "grid": [n, n], "n_sls": 3, "fd_order": 2, "parallelism": "single GPU",
                   "math": math},
        "launch_shapes": shapes,
        "hbm_frac_step": round(gcell * 1e9 * b_step / (HBM_PEAK_GBS * 1e9), 4),
        "bytes_per_cell_step": round(b_step, 2),
        "kernel_ms": {"stress": round(k_ms[0], 4), "velocity": round(k_ms[1], 4),
                      "step_counter": round(k_ms[2], 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": STRESS2D_TRAFFIC_2048 if n == 2048 else None,
                     "traffic_source": None if n != 2048 else
                     "rocprofv3 FETCH_SIZE + WRITE_SIZE of the tuned shape, profiles/visco2d_2048_kernels.md",
                     "kernel": "stress2d", "bytes_per_cell": round(b_stress, 2)},
        "cpu_baseline": cpu,
    }
    if fused is not None:
        line["fused_schedule_gcell_per_s"] = fused
    return line


def shape_sweep(g, k):
    """k stress and velocity launches (HIP-event timed) of the FAST stress pass
    at each shape of SHAPES_3D in turn, the velocity pass fixed: the run that
    rocprofv3 --pmc counts per dispatch (after 2 warm-up dispatches at the
    first shape, dispatch 2 + i of the stress kernel has shape SHAPES_3D[i // k]);
    then the same for the velocity pass over VELOCITY_SHAPES_3D (its dispatch
    2 + k * len(SHAPES_3D) + i has shape VELOCITY_SHAPES_3D[i // k])."""
    g.set_math("fast")
    g.set_tuning(0, *SHAPES_3D[0])
    g.set_tuning(1, *SHAPES_3D[0])
    g.profile(2)                                  # warm-up, not counted per shape
    out = []
    for shape in SHAPES_3D:
        g.set_tuning(0, *shape)
        ms = g.profile(k)
        out.append({"shape": list(shape), "stress_ms": round(ms[0], 4)})
        print(f"# stress shape {shape}: {ms[0]:.4f} ms", file=sys.stderr, flush=True)
    # then the velocity pass at each shape, the stress pass fixed at the first
    g.set_tuning(0, *SHAPES_3D[0])
    vel = []
    for shape in VELOCITY_SHAPES_3D:
        g.set_tuning(1, *shape)
        ms = g.profile(k)
        vel.append({"shape": list(shape), "velocity_ms": round(ms[1], 4)})
        print(f"# velocity shape {shape}: {ms[1]:.4f} ms", file=sys.stderr, flush=True)
    g.close()
    emit({"shape_sweep": out, "velocity_sweep": vel, "dispatches_per_shape": k})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=10, help="steps of the CPU-oracle sample (0: skip)")
    ap.add_argument("--workload", choices=("3d", "2d", "shots", "adjoint"), default="3d")
    ap.add_argument("--tuning", default="",
                    help="3D: pin the launch shapes instead of autotuning, "
                         "ty,nt,xchunk,remap/ty,nt,xchunk,remap (stress/velocity)")
    ap.add_argument("--shape-sweep", type=int, default=0,
                    help="3D: only time k launches of the stress pass, then of the velocity pass, "
                         "at each shape of SHAPES_3D (the run the traffic table is counted on)")
    ap.add_argument("--n2d", type=int, default=2048, help="grid size (--workload 2d)")
    ap.add_argument("--math2d", choices=("strict", "fast"), default="strict",
                    help="arithmetic mode (--workload 2d): strict = the reference's operation order")
    ap.add_argument("--fd-dirs", type=int, default=2, help="finite-difference directions (--workload adjoint)")
    ap.add_argument("--store-all", type=int, default=1,
                    help="also time the gradient with every state stored, no recompute (--workload adjoint)")
    ap.add_argument("--adj-shots", type=int, default=8, help="shots (--workload adjoint)")
    ap.add_argument("--adj-schedules", type=int, default=0,
                    help="--workload adjoint: also time shot 0's reverse sweep with and without the "
                         "second-stream recompute, with sqrt(N) checkpoints and with every state stored, "
                         "and check that the gradients have the same bits")
    ap.add_argument("--sizes", default="512,1024,1536,2048",
                    help="specimen grid sizes (--workload shots), comma separated")
    ap.add_argument("--shots", type=int, default=16, help="array elements = FMC shots per specimen (--workload shots)")
    ap.add_argument("--fused-sweep", action="store_true",
                    help="--workload 2d: also time the opt-in fused schedule at x-chunks 8..64")
    ap.add_argument("--group-sweep", action="store_true",
                    help="also time the batch at cache-group budgets 0..320 MB (--workload shots)")
    ap.add_argument("--extras", type=int, default=1,
                    help="3D line: also measure batched FMC shots/s, the adjoint/forward ratio and the 2D "
                         "2048^2 line (key visco2d, reference operation order) (0: skip)")
    ap.add_argument("--shots-window", type=int, default=3000,
                    help="steps of each shot timed for the 3D line's shots/s (batch and per-shot loop)")
    ap.add_argument("--copy-only", action="store_true",
                    help="only run the plain copy kernel (1 GiB, 5 launches): rocprof counter calibration")
    ap.add_argument("--bandwidth-probe", action="store_true",
                    help="only measure the in-place read/write mixes of the stencil passes (vse_stream_bandwidth)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the shot/slab sharding "
              "across GPUs is a design, not part of this build.", file=sys.stderr)
        sys.exit(2)

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    if args.copy_only:
        import ctypes
        from taichi_sim_amd import _lib
        gbs = ctypes.c_float(0.0)
        rc = _lib.load().vse_copy_bandwidth(1 << 30, 5, ctypes.byref(gbs))
        emit({"copy_bytes_per_launch": 2 << 30, "launches": 6, "gbs": round(gbs.value, 1), "rc": rc})
        sys.exit(0 if rc == 0 else 1)
    if args.bandwidth_probe:
        return bandwidth_probe()
    if args.workload == "shots":
        return bench_shots(args)
    if args.workload == "2d":
        return bench_2d(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    from taichi_sim_amd import workloads, autotune
    from taichi_sim_amd.backend import attenuation_from_config
    from taichi_sim_amd.engine import VSESim
    from taichi_sim_amd.sim_support.simulator3d import Simulator3D

    total = args.warmup + args.steps
    cfg = workloads.visco_3d_block(n=args.n, steps=total)
    sim = Simulator3D(cfg)
    att = attenuation_from_config(cfg)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    ncell = int(np.prod(g.shape))
    print(f"# setup done: grid {g.shape}, {g.device_bytes() / 2**30:.1f} GiB on {g.device_name()}",
          file=sys.stderr, flush=True)

    # launch shapes from the autotuner (before the timed region; it advances
    # the state, so the simulation restarts from zero afterwards)
    if args.shape_sweep:
        return shape_sweep(g, args.shape_sweep)
    tuned = {}
    if args.tuning:
        pinned = [tuple(int(v) for v in part.split(",")) for part in args.tuning.split("/")]
        if len(pinned) != 2 or any(len(p) != 4 for p in pinned):
            raise SystemExit("--tuning takes ty,nt,xchunk,remap/ty,nt,xchunk,remap (stress/velocity)")
        for math in ("fast", "strict"):
            tuned[math] = {"stress": pinned[0] + (None,), "velocity": pinned[1] + (None,)}
    else:
        for math in ("fast", "strict"):
            g.set_math(math)
            best = autotune.tune(g, steps=2, candidates={"stress": SHAPES_3D, "velocity": VELOCITY_SHAPES_3D},
                                 cache=None)
            tuned[math] = {k: best[k] for k in ("stress", "velocity")}
    print(f"# launch shapes (ty, nt, xchunk, remap, step ms): {tuned}", file=sys.stderr, flush=True)

    def timed_run(math, steps):
        g.set_math(math)
        for k, name in enumerate(("stress", "velocity")):
            g.set_tuning(k, *tuned[math][name][:4])
        g.reset()
        if args.warmup:
            g.run(args.warmup, check_every=args.warmup)
        t0 = time.perf_counter()
        dev_ms = g.run(steps, check_every=500, timed=True)      # synchronises before returning
        wall = time.perf_counter() - t0
        return wall, dev_ms

    # headline: FAST arithmetic (FP32, FD coefficients pre-divided by the
    # spacing), which stays within 1e-5 rel-L2 of the reference operation order
    # (tests/test_gpu_z_fullsize.py at this size); STRICT, term for term the
    # reference's order, is timed beside it
    wall_strict, _ = timed_run("strict", max(1, args.steps // 4))
    gcell_strict = ncell * max(1, args.steps // 4) / wall_strict / 1e9
    wall, dev_ms = timed_run("fast", args.steps)
    ms_per_step = 1e3 * wall / args.steps
    gcell = ncell * args.steps / wall / 1e9
    print(f"# timed {args.steps} steps: wall {wall:.3f} s, device {dev_ms / 1e3:.3f} s",
          file=sys.stderr, flush=True)

    # per-kernel durations (HIP events around each launch), after the timed loop
    k_ms = g.profile(args.profile_steps)
    traffic = STRESS_TRAFFIC_512.get(tuple(tuned["fast"]["stress"][:4])) if args.n == 512 else None
    b_step, b_stress, b_vel = bytes_per_cell(g, g.n_sls)
    achieved = b_stress * ncell / (k_ms[0] * 1e-3) / 1e9
    step_frac = gcell * 1e9 * b_step / (HBM_PEAK_GBS * 1e9)
    g.close()

    # achievable-HBM reference: a plain float4 copy of 2 x 2 GiB
    import ctypes
    from taichi_sim_amd import _lib
    copy_gbs = ctypes.c_float(0.0)
    if _lib.load().vse_copy_bandwidth(2 << 30, 20, ctypes.byref(copy_gbs)) != 0:
        copy_gbs.value = float("nan")

    cpu = None
    if args.cpu_steps > 0:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        o = oracle.Oracle(sim, variant="wgsl", att=att, n_threads=threads)
        t = time.perf_counter()
        o.run(args.cpu_steps)
        cpu_s = time.perf_counter() - t
        cpu = {"value": round(ncell * args.cpu_steps / cpu_s / 1e9, 6), "unit": "Gcell-upd/s",
               "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same {args.n}^3 grid, FP32 C oracle "
                         f"(reference update order, OpenMP), {cpu_s:.1f} s"}
        del o

    line = {
        "metric": "Gcell-upd/s (3D viscoelastic Zener L=3 FP32 forward)",
        "value": round(gcell, 4), "unit": "Gcell-upd/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3D viscoelastic Zener FD {args.n}^3 (incl. 20-cell CPML), L=3 Q=50, "
                               f"Al over steel with 24 seeded inclusions, piston source, 32x32 receivers",
                   "grid": [args.n] * 3, "n_sls": 3, "fd_order": 2, "parallelism": "single GPU"},
        "math": "fast: FP32, FD coefficients pre-divided by the spacing, 1/k and 1/3 multiplied; "
                "<=1e-5 rel-L2 from the reference operation order (tested at 512^3)",
        "value_strict_math": round(gcell_strict, 4),
        "launch_shapes": tuned,
        "hbm_frac_step": round(step_frac, 4),
        "copy_kernel_gbs": round(copy_gbs.value, 1),
        "hbm_frac_step_vs_copy": round(gcell * b_step / copy_gbs.value, 4),
        "bytes_per_cell_step": round(b_step, 2),
        "kernel_ms": {"stress": round(k_ms[0], 4), "velocity": round(k_ms[1], 4), "sparse": round(k_ms[2], 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "traffic_source": None if traffic is None else
                     "rocprofv3 FETCH_SIZE + WRITE_SIZE of this launch shape, profiles/visco3d_512cube_kernels.md",
                     "kernel": "stress3d", "bytes_per_cell": round(b_stress, 2)},
        "cpu_baseline": cpu,
    }
    if args.extras:
        # the rest of the metric: batched shots/s and the gradient's cost, each on its
        # own configuration, measured after the 3D handle is released
        s = measure_shots((512, 1024, 1536, 2048), 16, args.shots_window, profile_steps=0, persistent=False)
        line["shots_per_s"] = {"batched": s["value"], "per_shot_loop": s["per_shot_loop_shots_per_s"],
                               "speedup": s["batch_speedup"], "shots": s["config"]["shots"],
                               "window_steps": s["config"]["window_steps"], "shot_steps": 3000,
                               "cache_groups": s["config"]["cache_groups"], "tuning": s["tuning"],
                               "workload": s["config"]["workload"]}
        c2 = measure_2d(2048, 2000, 10, "strict", 10, 0)
        line["visco2d"] = {"value": c2["value"], "unit": c2["unit"], "ms_per_step": c2["ms_per_step"],
                           "steps": c2["steps"], "math": "strict", "kernel_ms": c2["kernel_ms"],
                           "roofline": c2["roofline"], "hbm_frac_step": c2["hbm_frac_step"],
                           "launch_shapes": c2["launch_shapes"], "workload": c2["config"]["workload"]}
        a = measure_adjoint(1, store_all=True)
        line["adjoint_over_forward"] = {"value": a["value"], "forward_ms": a["forward_ms_per_shot"],
                                        "checkpointed_forward_ms": a["checkpointed_forward_ms_per_shot"],
                                        "recompute_adjoint_ms": a["recompute_adjoint_ms_per_shot"],
                                        "segment": a["config"]["segment"], "shots": 1, "steps": a["steps"],
                                        "every_state_stored": a["every_state_stored"],
                                        "gradient_accumulation": a["gradient_accumulation"],
                                        "workload": a["config"]["workload"]}
    emit(line)


def bandwidth_probe():
    """GB/s the device sustains at each stencil pass's read/write mix, without the
    arithmetic (vse_stream_bandwidth: n_read read-only arrays, n_rw arrays read and
    rewritten in place, 512^3 floats each), beside the plain copy kernel."""
    import ctypes
    from taichi_sim_amd import _lib
    lib = _lib.load()
    n = 512 ** 3
    mixes = {"3D stress pass (6 read-only, 24 read and written)": (6, 24),
             "3D velocity pass (7 read-only, 3 read and written)": (7, 3),
             "6 read-only, 12 read and written": (6, 12),
             "2D stress pass (5 read-only, 12 read and written)": (5, 12),
             "2D velocity pass (4 read-only, 2 read and written)": (4, 2),
             "24 read and written only": (0, 24)}
    out = {}
    gbs = ctypes.c_float(0.0)
    rc = lib.vse_copy_bandwidth(1 << 30, 10, ctypes.byref(gbs))
    out["float4 copy, 2 x 1 GiB"] = round(gbs.value, 1) if rc == 0 else None
    for name, (r, w) in mixes.items():
        rc = lib.vse_stream_bandwidth(n, r, w, 5, ctypes.byref(gbs))
        out[name] = round(gbs.value, 1) if rc == 0 else None
        print(f"# {name}: {out[name]} GB/s", file=sys.stderr, flush=True)
    emit({"metric": "sustained GB/s by read/write mix", "floats_per_array": n, "gbs": out})


def bench_shots(args):
    """Batched multi-shot 2D throughput vs the per-shot loop (see module doc)."""
    sizes = tuple(int(v) for v in args.sizes.split(","))
    emit(measure_shots(sizes, args.shots, args.steps, args.warmup, args.profile_steps, args.group_sweep))


def measure_shots(sizes, n_shots, steps, warmup=0, profile_steps=10, group_sweep=False, persistent=True):
    """FMC on plates of the given sizes x n_shots: every shot as one batch
    against the same shots one after another, over a window of `steps` steps of
    each shot; shots/s is for 3000-step shots (window time x 3000 / steps)."""
    from taichi_sim_amd import autotune, workloads
    from taichi_sim_amd.backend import attenuation_from_config
    from taichi_sim_amd.engine import VSEBatch, VSESim
    from taichi_sim_amd.sim_support.simulator import Simulator

    shot_len = 3000                                   # steps of one shot
    t_setup = time.perf_counter()
    sims, cells = [], 0
    for cfg, laws in workloads.fmc_2d(sizes=sizes, n_elem=n_shots, steps=shot_len,
                                      seeds=tuple(40 + i for i in range(len(sizes)))):
        sim = Simulator(cfg)
        sim._emission_laws = laws
        att = attenuation_from_config(cfg)
        for law in range(n_shots):
            sim.select_law(law)
            sims.append(VSESim.from_simulator(sim, variant="wgsl", att=att))
            cells += int(np.prod(sims[-1].shape))
    gib = sum(g.device_bytes() for g in sims) / 2**30
    print(f"# setup: {len(sims)} simulations, {cells / 1e6:.0f} Mcell, {gib:.1f} GiB, "
          f"{time.perf_counter() - t_setup:.0f} s", file=sys.stderr, flush=True)

    # launch shapes: the 2D autotuner once per member shape (the per-shot loop
    # and the batch's groups of one run with them), then the batch's row groups
    # and cache-group budget
    t_tune = time.perf_counter()
    tuned2d = {}
    for g in sims:
        if g.shape not in tuned2d:
            best = autotune.tune(g, steps=3, cache=None)
            tuned2d[g.shape] = {k: best[k] for k in ("stress", "velocity")}
        for k, name in enumerate(("stress", "velocity")):
            g.set_tuning(k, *tuned2d[g.shape][name][:4])
    batch = VSEBatch(sims)
    bt = autotune.tune_batch(batch, cache=None)
    tuning = {"solo": {"x".join(map(str, shp)): {k: list(v[:4]) for k, v in t.items()} for shp, t in tuned2d.items()},
              "batch": {"rows": bt["rows"], "group_mb": bt["group_mb"], "groups": bt["groups"]}}
    print(f"# tuning ({time.perf_counter() - t_tune:.0f} s): {tuning}", file=sys.stderr, flush=True)
    groups = batch.groups()
    k_batch = k_solo = None
    if profile_steps > 0:
        k_batch = batch.profile(profile_steps)        # all members in one launch per pass
        k_solo = [0.0, 0.0, 0.0]
        for g in sims:
            t = g.profile(profile_steps)
            k_solo = [a + b for a, b in zip(k_solo, t)]

    def timed(fn):
        # both paths run hipGraphs of min(steps, 256) steps; an untimed call of
        # that length captures them first
        for g in sims:
            g.reset()
        fn(min(steps, 256))
        for g in sims:
            g.reset()
        t0 = time.perf_counter()
        dev = fn(steps)
        return time.perf_counter() - t0, dev

    check = min(steps, 500)
    wall_b, dev_b = timed(lambda n: batch.run(n, check_every=check, timed=True))
    print(f"# batch ({len(groups)} groups): {steps} steps in {wall_b:.3f} s (device {dev_b / 1e3:.3f} s)",
          file=sys.stderr, flush=True)
    wall_l, _ = timed(lambda n: [g.run(n, check_every=check) for g in sims])
    print(f"# per-shot loop: {wall_l:.3f} s", file=sys.stderr, flush=True)
    # the same batch under the persistent schedule: each group's window as
    # cooperative launches of up to `check` steps (grid-wide barriers between passes)
    wall_p = None
    if persistent:
        batch.set_schedule("persistent")
        wall_p, dev_p = timed(lambda n: batch.run(n, check_every=check, timed=True))
        batch.set_schedule("graph")
        print(f"# batch, persistent schedule: {steps} steps in {wall_p:.3f} s (device {dev_p / 1e3:.3f} s)",
              file=sys.stderr, flush=True)
    sweep = None
    if group_sweep:
        # batch Gcell/s at each cache-group budget (MB of member device memory per group)
        sweep, n_sw = {}, min(steps, 200)
        for mb in (0, 32, 64, 96, 128, 160, 192, 224, 256, 320):
            batch.set_group_bytes(mb << 20)
            key = f"{mb}MB/{len(batch.groups())}g"
            for rep in range(2):                      # the first run captures the groups' graphs
                for g in sims:
                    g.reset()
                t0 = time.perf_counter()
                batch.run(n_sw, check_every=n_sw)
                w = time.perf_counter() - t0
            sweep[key] = round(cells * n_sw / w / 1e9, 2)
            print(f"# group budget {key}: {sweep[key]} Gcell/s", file=sys.stderr, flush=True)
    batch.close()
    name = sims[0].device_name()
    for g in sims:
        g.close()

    shots = len(sims) * steps / shot_len
    ms = lambda k: None if k is None else {n: round(v, 4) for n, v in zip(("stress", "velocity", "sparse"), k)}
    return {
        "metric": "shots/s (2D viscoelastic Zener L=3 FP32 FMC, 3000-step shots, batched)",
        "value": round(shots / wall_b, 4), "unit": "shots/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(1e3 * wall_b / steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"FMC: {len(sizes)} steel plates with voids {'/'.join(f'{n}^2' for n in sizes)} "
                               f"(incl. CPML) x {n_shots} shots ({n_shots}-element array, all receiving)",
                   "shots": len(sims), "cells_per_step": cells, "shot_steps": shot_len, "window_steps": steps,
                   "parallelism": "single GPU, one batch", "cache_groups": len(groups)},
        "per_shot_loop_shots_per_s": round(shots / wall_l, 4),
        "batched_persistent_shots_per_s": None if wall_p is None else round(shots / wall_p, 4),
        "group_sweep_gcell_per_s": sweep,
        "batch_speedup": round(wall_l / wall_b, 3),
        "gcell_per_s_batched": round(cells * steps / wall_b / 1e9, 3),
        "gcell_per_s_loop": round(cells * steps / wall_l / 1e9, 3),
        "kernel_ms_all_in_one_launch": ms(k_batch),
        "kernel_ms_loop_sum": ms(k_solo),
        "tuning": tuning,
        "device": name,
    }


def bench_adjoint(args):
    """Adjoint/forward time ratio of the C-ABI gradient (see module doc)."""
    emit(measure_adjoint(args.adj_shots, args.fd_dirs, args.store_all, args.adj_schedules))


def measure_adjoint(n_shots, fd_dirs=0, store_all=False, schedules=False):
    """Gradient of the trace misfit on the 1024^2 inversion problem: device
    time of the checkpointed forward + recompute + adjoint over the plain
    forward, per shot; optional finite-difference directions and the
    every-level-stored variant."""
    from taichi_sim_amd import autotune, workloads
    from taichi_sim_amd.adjoint import VSEGrad, chain_to_cells
    from taichi_sim_amd.backend import attenuation_from_config
    from taichi_sim_amd.engine import VSESim
    from taichi_sim_amd.sim_support.simulator import Simulator

    n, steps = 1024, 3000
    cfg, laws, true = workloads.adjoint_2d(n=n, steps=steps, n_shots=n_shots)
    sim = Simulator(cfg)
    sim._emission_laws = laws
    att = attenuation_from_config(cfg)
    g = VSESim.from_simulator(sim, variant="wgsl", visco=True, att=att)
    # the plain forward, the ratio's denominator, runs at its tuned launch shapes
    fwd_shapes = {k: list(v[:4]) for k, v in autotune.tune(g, steps=3, cache=None).items() if k != "table"}
    start = [np.asarray(a, np.float32).copy() for a in (sim._rho_grid_vx, sim._cp_grid_vx, sim._cs_grid_vx)]
    eng = VSEGrad(g)
    seg = eng.seg_len
    print(f"# setup: grid {g.shape}, segment {seg} steps, {eng.device_bytes() / 2**30:.1f} GiB "
          f"of checkpoints/adjoint state", file=sys.stderr, flush=True)
    ncell = int(np.prod(g.shape))

    # observed data: the true model through the product forward (vse_run)
    obs = []
    g.set_material(*true)
    for k in range(n_shots):
        sim.select_law(k)
        g.load_sources(sim)
        g.reset()
        g.run(steps, check_every=500)
        obs.append(g.traces().copy())

    t_fwd, t_gfwd, t_bwd, J = [], [], [], 0.0
    grad = None
    for k in range(n_shots):
        sim.select_law(k)
        g.load_sources(sim)
        # the plain forward of the same shot at the starting model: the median of
        # three runs (one run alone varies by a few percent between calls)
        g.set_material(*start)
        runs = []
        for _ in range(3):
            g.reset()
            runs.append(g.run(steps, check_every=500, timed=True))
        t_fwd.append(float(np.median(runs)))
        # checkpointed forward + FP64 misfit + recompute/adjoint
        eng.set_model(*start)
        eng.forward(steps, check_every=500)
        t_gfwd.append(eng.ms["forward"])
        J += eng.misfit(obs[k])
        coef = eng.backward()
        t_bwd.append(eng.ms["backward"])
        grad = coef if grad is None else {key: grad[key] + coef[key] for key in grad}
        if k == 0:
            coef0 = {key: v.copy() for key, v in coef.items()}
        print(f"# shot {k}: forward {t_fwd[-1]:.1f} ms, checkpointed forward {t_gfwd[-1]:.1f} ms, "
              f"recompute+adjoint {t_bwd[-1]:.1f} ms", file=sys.stderr, flush=True)
    g_rho, g_cp, g_cs = (t.numpy() for t in chain_to_cells(grad, *start))

    # shot 0 again with the FP32 partials flushed into FP64 every step: how far
    # the default 32-step grouping is from a per-step FP64 time accumulation
    sim.select_law(0)
    g.load_sources(sim)
    eng.set_flush(1)
    eng.set_model(*start)
    eng.forward(steps, check_every=500)
    eng.misfit(obs[0])
    c_every = eng.backward()
    eng.set_flush(32)
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    accumulation = {"flush_steps": 32, "every_step_recompute_adjoint_ms": round(eng.ms["backward"], 2),
                    "rel_l2_vs_every_step_fp64": {key: float(f"{rel(coef0[key], c_every[key]):.2e}")
                                                  for key in coef0}}
    print(f"# FP32 partials vs per-step FP64 accumulation: {accumulation}", file=sys.stderr, flush=True)

    # finite-difference spot check at full size on shot 0.  The misfit of the
    # inversion is tiny at the starting model (smooth anomalies reflect
    # weakly), where a central difference measures its cubic term; the check
    # therefore differentiates J = 0.5*||traces||^2 (zero data), which the
    # FP32 forward resolves to ~1e-7 of J, along directions under the array
    # moving rho, cp and cs, with a fourth-order central quotient.  Direction
    # k is the adjoint gradient of each map weighted by W_k, a seeded sum of
    # positive Gaussian bumps under the array: d.g = sum W_k g^2 cannot
    # cancel, whereas a smooth direction of random sign can fall near the
    # oscillating sensitivity's null space, where J changes by less than the
    # FP32 forward resolves.  The difference quotient uses forward runs only;
    # random directions at 128^2 are checked in the test suite.
    zero = np.zeros_like(obs[0])
    energy = None
    if fd_dirs > 0:
        sim.select_law(0)
        g.load_sources(sim)
        eng.set_model(*start)
        eng.forward(steps, check_every=500)
        energy = eng.misfit(zero)
        e_coef = eng.backward()
        e_rho, e_cp, e_cs = (t.numpy() for t in chain_to_cells(e_coef, *start))

    def misfit_all(model):
        eng.set_model(*model)
        eng.forward(steps, check_every=500)
        return eng.misfit(zero)

    fd_rel = []
    if fd_dirs > 0:
        rng = np.random.default_rng(5)
        y_top = int(np.min(np.nonzero(np.asarray(sim._pos_sources) != -1)[1]))
        X, Y = np.meshgrid(np.arange(g.shape[0]), np.arange(g.shape[1]), indexing="ij")
        for d in range(fd_dirs):
            w = np.zeros(g.shape)
            for _ in range(4):
                cx, cy = rng.uniform(0.2, 0.8) * g.shape[0], rng.uniform(y_top + 2, y_top + 40)
                sg = rng.uniform(10.0, 40.0)
                w += rng.uniform(0.5, 1.0) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sg * sg))
            h = 2e-2
            dirs = [w * e / np.abs(w * e).max() * float(m.max()) for e, m in zip((e_rho, e_cp, e_cs), start)]
            models = {t: [(start[p] + t * h * dirs[p]).astype(np.float32) for p in range(3)] for t in (-2, -1, 1, 2)}
            Jt = {t: misfit_all(m) for t, m in models.items()}
            fd = (8.0 * (Jt[1] - Jt[-1]) - (Jt[2] - Jt[-2])) / 12.0
            step = lambda p, t: models[t][p].astype(np.float64) - start[p].astype(np.float64)
            ad = sum(float((gp * (8.0 * (step(p, 1) - step(p, -1)) - (step(p, 2) - step(p, -2))) / 12.0).sum())
                     for p, gp in enumerate((e_rho, e_cp, e_cs)))
            fd_rel.append(abs(fd - ad) / max(abs(fd), 1e-30))
            print(f"# FD direction {d}: FD {fd:.6e} adjoint {ad:.6e} rel {fd_rel[-1]:.1e}", file=sys.stderr, flush=True)
    eng.close()

    # the same gradient with every state kept (segment = N levels, no
    # recompute): 3000 levels of 1024^2 take ~180 GB, which one MI355X holds
    stored = None
    if store_all:
        sim.select_law(0)
        g.load_sources(sim)
        ea = VSEGrad(g, seg_len=steps)
        ea.set_model(*start)
        ta = []
        for _ in range(2):
            ea.forward(steps, check_every=500)
            ea.misfit(obs[0])
            c_all = ea.backward()
            ta.append((ea.ms["forward"], ea.ms["backward"]))
        stored = {"segment": ea.seg_len, "device_gib": round(ea.device_bytes() / 2**30, 1),
                  "forward_ms": round(ta[-1][0], 2), "adjoint_ms": round(ta[-1][1], 2),
                  "same_gradient_bits": all(np.array_equal(c_all[key], coef0[key]) for key in coef0)}
        ea.close()
        print(f"# every state stored: {stored}", file=sys.stderr, flush=True)
    # the reverse sweep of shot 0 with the next segment recomputed on a second
    # stream (the default) and with one stream; bit-identical gradients
    sched = None
    if schedules:
        sim.select_law(0)
        g.load_sources(sim)
        sched = {}
        for seg in (0, steps):
            es = VSEGrad(g, seg_len=seg)
            es.set_model(*start)
            ref = None
            for ov in (False, True):
                es.set_overlap(ov)
                es.forward(steps, check_every=500)
                es.misfit(obs[0])
                c = es.backward()
                ref = c if ref is None else ref
                key = f"{'sqrtN' if seg == 0 else 'every_state'}_{'overlapped' if ov else 'one_stream'}"
                sched[key] = {"backward_ms": round(es.ms["backward"], 2),
                              "same_gradient_bits": all(np.array_equal(c[k], ref[k]) for k in ref)}
                print(f"# reverse sweep {key}: {sched[key]}", file=sys.stderr, flush=True)
            es.close()

    g.close()

    fwd, gf, bw = (float(np.mean(t)) for t in (t_fwd, t_gfwd, t_bwd))
    ratio = (gf + bw) / fwd
    if stored is not None:
        stored["ratio"] = round((stored["forward_ms"] + stored["adjoint_ms"]) / fwd, 3)
    return {
        "metric": "adjoint/forward time (2D viscoelastic 1024^2 L=3, 3000 steps, sqrt(N) checkpoints)",
        "value": round(ratio, 3), "unit": "x", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round((gf + bw) / steps, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (FP64 misfit and gradient accumulators)", "data": "synthetic",
        "config": {"workload": f"gradient of the L2 trace misfit, {n_shots} FMC shots, true model with 8 Gaussian "
                               f"velocity anomalies (+-5%), homogeneous start", "grid": [n, n], "n_sls": 3,
                   "segment": seg, "parallelism": "single GPU"},
        "forward_ms_per_shot": round(fwd, 2), "checkpointed_forward_ms_per_shot": round(gf, 2),
        "recompute_adjoint_ms_per_shot": round(bw, 2), "misfit": J, "trace_energy_shot0": energy,
        "gradient_norms": {"rho": float(np.linalg.norm(g_rho)), "cp": float(np.linalg.norm(g_cp)),
                           "cs": float(np.linalg.norm(g_cs)), "tau": float(np.linalg.norm(grad["tau"]))},
        "gcell_per_s_forward": round(ncell * steps / fwd / 1e6, 3),
        "fd_rel_err_trace_energy": [float(f"{e:.2e}") for e in fd_rel],
        "gradient_accumulation": accumulation,
        "every_state_stored": stored,
        "reverse_schedules": sched,
        "forward_launch_shapes": fwd_shapes,
    }


if __name__ == "__main__":
    import contextlib
    with contextlib.redirect_stdout(sys.stderr):
        main()
