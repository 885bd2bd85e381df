"""Autotuner for the 3D kernels' launch shapes.

For one simulation handle it times every candidate (tile height ty, non-
temporal streaming nt, x-chunk length, tile order) of the stress and velocity
passes with HIP events (``VSESim.profile``), re-times the fastest few over
more steps, and keeps the one with the fastest *step* (stress + velocity back to back: one pass's speed depends on what the
other leaves in L2 / MALL), tuning stress first, then velocity, and caches the
choice per problem key (ndim, grid, FD order, visco, L) in a JSON file.  All
shapes compute identical results, so tuning only changes speed.  The handle's
state advances while tuning; call ``reset()`` before a measured run.
"""
import json
import os

DEFAULT_CACHE = os.environ.get("VSE_TUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                            "tuning_cache.json"))
TY = (1, 4, 8)
NT = (0, 1)
XCHUNK = (1, 2, 4, 8, 16, 32)
PLAIN = (4, 0, 0, 0)      # xchunk 0: one cell per thread
REMAP = (0, 1, 2)


def problem_key(g):
    """Cache key: the device (name, architecture, CU count) and arithmetic mode
    as well as the problem, so a shape tuned on one part is not reused on another."""
    dev = g.device_name() if hasattr(g, "device_name") else "any"
    math = getattr(g, "math", "strict")
    return f"{dev}|{g.ndim}d:{'x'.join(map(str, g.shape))}:ord{g.ord}:visco{int(g.visco)}:L{g.n_sls}:{math}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(g, steps=3, candidates=None, cache=DEFAULT_CACHE, verbose=False, use_cache=True, finalists=3,
         final_steps=8):
    """Pick and apply the fastest launch shape of each 3D pass of ``g``.

    Every candidate is timed over ``steps`` steps; the ``finalists`` fastest are
    timed again, in turn, over ``final_steps`` steps (shapes within a few percent
    of each other are not told apart reliably by a short timing), and the fastest
    of that second timing pass is kept.

    Returns {"stress": (ty, nt, xchunk, remap, step_ms), "velocity": (...), "table": [...]}.
    """
    if g.ndim != 3:
        return {}
    key = problem_key(g)
    db = _load(cache) if cache else {}
    if use_cache and key in db:
        best = db[key]
        for k, name in enumerate(("stress", "velocity")):
            ty, nt, xc, rm, _ = best[name]
            g.set_tuning(k, ty, nt, xc, rm)
        return best
    cands = candidates or ([PLAIN] + [(ty, nt, xc, rm) for ty in TY for nt in NT for xc in XCHUNK for rm in REMAP])
    table = []
    best = {}
    for k, name in enumerate(("stress", "velocity")):
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
    if cache:
        db[key] = best
        try:
            with open(cache, "w") as f:
                json.dump(db, f, indent=1, sort_keys=True)
        except OSError:
            pass
    best["table"] = table
    return best
