"""Autotuner selection logic on the CPU, with a stand-in for the simulation
handle whose timings are scripted."""
import json

from taichi_sim_amd import autotune


class FakeSim:
    """Times each launch shape from a table; the first timing of a shape is
    scripted separately, to model a noisy short run."""
    ndim, shape, ord, visco, n_sls = 3, (64, 64, 64), 2, True, 3

    def __init__(self, first, steady):
        self.first, self.steady = first, steady
        self.tuning = {0: None, 1: None}
        self.seen = set()
        self.calls = []

    def set_tuning(self, k, ty, nt, xc, rm):
        self.tuning[k] = (ty, nt, xc, rm)

    def profile(self, steps):
        key = (self.tuning[0], self.tuning[1])
        self.calls.append((key, steps))
        table = self.first if key not in self.seen else self.steady
        self.seen.add(key)
        ms = table.get(self.tuning[0], 9.0) + (table.get(("v",) + self.tuning[1], 2.0) if self.tuning[1] else 2.0)
        return (ms * 0.75, ms * 0.25, 0.0)


def test_finalists_are_retimed_and_the_steady_best_kept(tmp_path):
    a, b, c, d = (4, 0, 4, 2), (4, 0, 8, 2), (1, 0, 4, 2), (1, 0, 8, 1)
    # short first timings rank b first, but over more steps c is the fastest
    first = {a: 8.6, b: 8.4, c: 8.5, d: 8.7}
    steady = {a: 8.6, b: 8.55, c: 8.45, d: 8.7}
    g = FakeSim(first, steady)
    cache = tmp_path / "tune.json"
    best = autotune.tune(g, steps=2, candidates=[a, b, c, d], cache=str(cache), finalists=3, final_steps=8)
    assert best["stress"][:4] == c
    assert g.tuning[0] == c
    # every candidate timed once per pass, then three finalists per pass over 8 steps
    assert sum(1 for _, s in g.calls if s == 8) == 6
    saved = json.loads(cache.read_text())
    assert saved[autotune.problem_key(g)]["stress"][:4] == list(c)
    # a second call reads the cache and times nothing
    g2 = FakeSim(first, steady)
    again = autotune.tune(g2, steps=2, candidates=[a, b, c, d], cache=str(cache))
    assert tuple(again["stress"][:4]) == c and g2.calls == [] and g2.tuning[0] == c


def test_two_dimensional_handles_tune_rows_and_tile_order(tmp_path):
    cands = autotune.candidates_for(type("G", (), {"ndim": 2})())
    assert sorted({c[2] for c in cands}) == list(autotune.ROWS_2D) and {c[3] for c in cands} == {0, 1}
    best_shape = (4, 0, 4, 0)
    steady = {c: 1.0 + 0.01 * i for i, c in enumerate(cands)}
    steady[best_shape] = 0.5
    g = FakeSim(steady, steady)
    g.ndim, g.shape = 2, (2048, 2048)
    best = autotune.tune(g, steps=2, cache=str(tmp_path / "t.json"))
    assert tuple(best["stress"][:4]) == best_shape and g.tuning[0] == best_shape
    vel = autotune.candidates_for(g, "velocity")
    assert {e[1:5] for e in best["table"] if e[0] == "stress"} == set(cands)
    assert {e[1:5] for e in best["table"] if e[0] == "velocity"} == set(vel)
    # the velocity pass also tries the x-row march (nt = 1); the stress pass does not
    assert {c[1] for c in vel} == {0, 1} and {c[1] for c in cands} == {0}


class FakeBatch:
    """A batch whose step time is scripted per (rows, group budget); run()
    advances a fake clock by that much per step."""

    def __init__(self, times, clock):
        self.times, self.clock = times, clock
        self.rows, self.mb = 1, 224
        self.runs = []
        member = type("M", (), {"ndim": 2, "shape": (512, 512), "ord": 2, "visco": True, "n_sls": 3,
                                "variant": "wgsl", "device_name": lambda self: "fake"})
        self.sims = [member(), member()]

    def set_rows(self, r):
        self.rows = r

    def set_group_bytes(self, nbytes):
        self.mb = nbytes >> 20

    def groups(self):
        return [0] if self.mb == 0 else [0, 1]

    def run(self, steps, check_every=100):
        self.runs.append((self.rows, self.mb, steps))
        self.clock[0] += steps * self.times.get((self.rows, self.mb), 1e-3)


def test_batch_tuner_keeps_the_fastest_rows_and_budget(tmp_path, monkeypatch):
    clock = [0.0]
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: clock[0])
    times = {(r, mb): 1e-3 for r in autotune.BATCH_ROWS for mb in autotune.BATCH_GROUP_MB}
    times[(2, 128)] = 0.6e-3
    times[(4, 0)] = 0.7e-3
    b = FakeBatch(times, clock)
    cache = tmp_path / "t.json"
    best = autotune.tune_batch(b, steps=10, cache=str(cache))
    assert (best["rows"], best["group_mb"]) == (2, 128) and (b.rows, b.mb) == (2, 128)
    assert abs(best["ms_per_step"] - 0.6) < 1e-9 and best["groups"] == 2
    n_cand = len(autotune.BATCH_ROWS) * len(autotune.BATCH_GROUP_MB)
    # each candidate: one capture run and one timed run; three finalists: capture and two timed runs
    assert len(b.runs) == 2 * n_cand + 3 * 3
    saved = json.loads(cache.read_text())
    assert saved[autotune.batch_key(b)]["rows"] == 2
    b2 = FakeBatch(times, clock)
    again = autotune.tune_batch(b2, steps=10, cache=str(cache))
    assert b2.runs == [] and (b2.rows, b2.mb) == (2, 128) and again["group_mb"] == 128
