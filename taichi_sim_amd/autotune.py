"""Autotuner for the launch shapes of the 2D, 3D and batched kernels.

``tune(g)`` times, for one simulation handle, every candidate shape of the
stress and velocity passes with HIP events (``VSESim.profile``), re-times the
fastest few over more steps, and keeps the one with the fastest *step*
(stress + velocity back to back: one pass's speed depends on what the other
leaves in L2 / MALL), tuning stress first, then velocity.  3D shapes are (tile
height ty, non-temporal streaming nt, x-chunk length, tile order); 2D shapes
are (row groups per workgroup, tile order).

``tune_batch(b)`` picks a batch's row groups per workgroup and its cache-group
budget (how many members' state is stepped together before the next group)
by the wall time of short batched runs.

Both cache the choice per problem key (device, arithmetic mode, grid(s), FD
order, viscoelastic flag, mechanisms) in a JSON file.  All shapes compute
identical results, so tuning only changes speed.  The handles' state advances
while tuning; call ``reset()`` before a measured run.
"""
import json
import os
import time

DEFAULT_CACHE = os.environ.get("VSE_TUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                            "tuning_cache.json"))
TY = (1, 4, 8)
NT = (0, 1)
XCHUNK = (1, 2, 4, 8, 16, 32)
PLAIN = (4, 0, 0, 0)      # xchunk 0: one cell per thread
REMAP = (0, 1, 2)
ROWS_2D = (1, 2, 4, 8)
MARCH_ROWS_2D = (2, 4, 8, 16)
REMAP_2D = (0, 1)
BATCH_ROWS = (1, 2, 4, 8)
BATCH_GROUP_MB = (0, 64, 128, 224, 320)


def problem_key(g):
    """Cache key: the device (name, architecture, CU count) and arithmetic mode
    as well as the problem, so a shape tuned on one part is not reused on another."""
    dev = g.device_name() if hasattr(g, "device_name") else "any"
    math = getattr(g, "math", "strict")
    return f"{dev}|{g.ndim}d:{'x'.join(map(str, g.shape))}:ord{g.ord}:visco{int(g.visco)}:L{g.n_sls}:{math}"


def batch_key(b):
    """Cache key of a batch: the device, the members' shapes in order, and what
    the batched kernels are compiled for."""
    g = b.sims[0]
    dev = g.device_name() if hasattr(g, "device_name") else "any"
    shapes = ",".join("x".join(map(str, s.shape)) for s in b.sims)
    return (f"{dev}|batch:{g.variant}:ord{g.ord}:visco{int(g.visco)}:L{g.n_sls}:"
            f"{getattr(g, 'math', 'strict')}:{shapes}")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, key, value):
    if not path:
        return
    db = _load(path)
    db[key] = value
    try:
        with open(path, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    except OSError:
        pass


def candidates_for(g, name="stress"):
    """Every launch shape (ty, nt, xchunk, remap) ``tune`` tries on pass ``name``
    of ``g``; the 3D velocity pass adds the LDS-staged shapes (nt bit 1, ty >=
    2 * FD order)."""
    if g.ndim == 2:
        out = [(4, 0, rows, rm) for rows in ROWS_2D for rm in REMAP_2D]
        if name == "velocity":      # the x-row march with register windows (nt = 1)
            out += [(4, 1, rows, rm) for rows in MARCH_ROWS_2D for rm in REMAP_2D]
        return out
    out = [PLAIN] + [(ty, nt, xc, rm) for ty in TY for nt in NT for xc in XCHUNK for rm in REMAP]
    if name == "velocity":
        out += [(ty, 2 | nt, xc, rm) for ty in TY if ty >= 2 * g.ord for nt in NT for xc in XCHUNK for rm in REMAP]
    return out


def tune(g, steps=3, candidates=None, cache=DEFAULT_CACHE, verbose=False, use_cache=True, finalists=3,
         final_steps=8):
    """Pick and apply the fastest launch shape of each pass of ``g`` (2D or 3D).

    Every candidate is timed over ``steps`` steps; the ``finalists`` fastest are
    timed again, in turn, over ``final_steps`` steps (shapes within a few percent
    of each other are not told apart reliably by a short timing), and the fastest
    of that second timing pass is kept.

    ``candidates`` is one list for both passes or {"stress": [...], "velocity": [...]}.
    Returns {"stress": (ty, nt, xchunk, remap, step_ms), "velocity": (...), "table": [...]}.
    """
    key = problem_key(g)
    db = _load(cache) if cache else {}
    if use_cache and key in db:
        best = db[key]
        for k, name in enumerate(("stress", "velocity")):
            ty, nt, xc, rm, _ = best[name]
            g.set_tuning(k, ty, nt, xc, rm)
        return best
    table = []
    best = {}
    for k, name in enumerate(("stress", "velocity")):
        if isinstance(candidates, dict):
            cands = candidates[name]
        else:
            cands = candidates or candidates_for(g, name)
        results = []
        for ty, nt, xc, rm in cands:
            g.set_tuning(k, ty, nt, xc, rm)
            t = g.profile(steps)
            ms = t[0] + t[1]
            results.append((ms, ty, nt, xc, rm))
            table.append((name, ty, nt, xc, rm, ms))
            if verbose:
                print(f"  {name:8s} ty={ty} nt={nt} xchunk={xc:3d} remap={rm}: {ms:.3f} ms", flush=True)
        final = []
        for _, ty, nt, xc, rm in sorted(results)[:max(1, finalists)]:
            g.set_tuning(k, ty, nt, xc, rm)
            t = g.profile(final_steps)
            final.append((t[0] + t[1], ty, nt, xc, rm))
            if verbose:
                print(f"  {name:8s} finalist ty={ty} nt={nt} xchunk={xc:3d} remap={rm}: {final[-1][0]:.3f} ms",
                      flush=True)
        ms, ty, nt, xc, rm = min(final)
        g.set_tuning(k, ty, nt, xc, rm)
        best[name] = (ty, nt, xc, rm, ms)
    _store(cache, key, {k: best[k] for k in ("stress", "velocity")})
    best["table"] = table
    return best


def tune_batch(b, steps=24, rows=BATCH_ROWS, group_mb=BATCH_GROUP_MB, cache=DEFAULT_CACHE, verbose=False,
               use_cache=True, finalists=3):
    """Pick and apply a batch's row groups per workgroup and cache-group budget.

    Each (rows, budget) pair is timed as the wall time of ``b.run(steps)``
    after an untimed run of the same length (which captures the groups'
    graphs); the ``finalists`` fastest are timed again and the fastest kept.
    Returns {"rows": r, "group_mb": mb, "groups": n, "ms_per_step": t, "table": [...]}.
    """
    key = batch_key(b)
    db = _load(cache) if cache else {}
    if use_cache and key in db:
        best = db[key]
        b.set_rows(best["rows"])
        b.set_group_bytes(best["group_mb"] << 20)
        return best

    def run(r, mb, reps=1):
        b.set_rows(r)
        b.set_group_bytes(mb << 20)
        b.run(steps, check_every=steps)                 # graph capture
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            b.run(steps, check_every=steps)
            t.append((time.perf_counter() - t0) * 1e3 / steps)
        return min(t)

    table = []
    for r in rows:
        for mb in group_mb:
            ms = run(r, mb)
            table.append((ms, r, mb, len(b.groups())))
            if verbose:
                print(f"  batch rows={r} group={mb} MB ({table[-1][3]} groups): {ms:.4f} ms/step", flush=True)
    final = []
    for _, r, mb, _ in sorted(table)[:max(1, finalists)]:
        final.append((run(r, mb, reps=2), r, mb))
    ms, r, mb = min(final)
    b.set_rows(r)
    b.set_group_bytes(mb << 20)
    best = {"rows": r, "group_mb": mb, "groups": len(b.groups()), "ms_per_step": ms}
    _store(cache, key, best)
    best["table"] = [(rr, m, ng, round(t, 4)) for t, rr, m, ng in table]
    return best
