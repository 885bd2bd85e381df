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


def test_two_dimensional_handles_are_not_tuned():
    g = FakeSim({}, {})
    g.ndim = 2
    assert autotune.tune(g, cache=None) == {}
    assert g.calls == []
