"""Batched 2D simulations (vse_batch_*): many shots and specimens per launch.

Every member of a batch must end exactly where the same simulation run
alone ends (same kernels, same per-cell arithmetic, only the workgroup ->
tile mapping differs), so the comparisons here are bit-exact; one member is
also checked against the CPU oracle.
"""
import numpy as np
import pytest

import oracle
from taichi_sim_amd import workloads
from taichi_sim_amd.backend import SimulatorHIP, attenuation_from_config
from taichi_sim_amd.engine import VSEBatch, VSESim
from taichi_sim_amd.sim_support import StabilityError
from taichi_sim_amd.sim_support.simulator import Simulator

pytestmark = pytest.mark.gpu


def _shots(sizes, n_shots, steps, variant="wgsl", channel=None):
    """One VSESim per (specimen, focal law), plus the Simulator and law of each."""
    out = []
    for cfg, laws in workloads.multi_shot_2d(sizes=sizes, n_shots=n_shots, steps=steps, n_elem=8):
        sim = Simulator(cfg)
        sim._emission_laws = laws
        att = attenuation_from_config(cfg)
        for law in range(n_shots):
            sim.select_law(law)
            out.append((sim, law, att, VSESim.from_simulator(sim, variant=variant, att=att, channel=channel)))
    return out


def _state(g):
    f = {k: g.field(k) for k in ("vx", "vy", "sxx", "syy", "sxy")}
    if g.visco:
        for c in ("xx", "yy", "xy"):
            f["r" + c] = g.memory(c)
    f["traces"] = g.traces()
    f["v2max"] = np.array([g.v2max()], dtype=np.float32)     # running max of |v|^2
    return f


def _assert_same(a, b, what):
    for k in a:
        assert a[k].shape == b[k].shape, (what, k)
        assert np.array_equal(a[k], b[k]), (what, k, float(np.abs(a[k] - b[k]).max()))


@pytest.mark.parametrize("variant,grouping", [("wgsl", "lockstep"), ("wgsl", "mixed"), ("taichi", "default")])
def test_batch_uneven_specimens_match_solo_runs(variant, grouping):
    steps = 240
    print(f"batch uneven specimens {variant} {grouping}", flush=True)
    members = _shots((96, 160, 131), 2, steps, variant)       # ragged sizes, 6 members
    batch = VSEBatch([m[3] for m in members])
    if grouping == "lockstep":
        batch.set_group_bytes(0)
        assert batch.groups() == [0]
    elif grouping == "mixed":
        # a budget that pairs the two 96^2 shots and leaves the larger ones alone
        sizes = [m[3].device_bytes() for m in members]
        batch.set_group_bytes(sizes[0] + sizes[1])
        assert batch.groups() == [0, 2, 3, 4, 5]
    batch.run(steps, check_every=100)                          # two graphs of 100 + 40 single steps
    batch.close()
    for sim, law, att, g in members:
        assert g.step() == steps
        got = _state(g)
        sim.select_law(law)
        solo = VSESim.from_simulator(sim, variant=variant, att=att)
        solo.run(steps, check_every=100)
        _assert_same(got, _state(solo), (sim._nx, law))
        solo.close()
    # one member against the CPU oracle (its reference-order restatement)
    sim, law, att, g = members[-1]
    sim.select_law(law)
    o = oracle.Oracle(sim, variant=variant, att=att, visco=variant == "wgsl").run(steps)
    for k in ("vx", "vy", "syy"):
        assert oracle.rel_l2(g.field(k), o.f[k]) <= 1e-5, k
    assert o.v2max[0] > 0 and abs(g.v2max() - o.v2max[0]) <= 1e-5 * o.v2max[0]
    if np.abs(o.traces).max() > 0:
        assert oracle.rel_l2(g.traces(), o.traces) <= 1e-5
    for m in members:
        m[3].close()


@pytest.mark.parametrize("variant,grouping,channel",
                         [("wgsl", "mixed", None), ("wgsl", "lockstep", "vy"), ("taichi", "default", None)])
def test_batch_persistent_schedule_bit_identical(variant, grouping, channel):
    """The persistent schedule (each group's steps in cooperative launches,
    grid-wide barriers between the passes; groups of one member included)
    leaves every member exactly where the graph schedule does: fields,
    memory variables, traces (sigma or a velocity channel), |v|^2 max and the
    step counter, over launches of 64, 36 and 140 steps."""
    steps = 240
    print(f"persistent batch {variant} {grouping} {channel}", flush=True)
    runs = []
    for schedule in ("graph", "persistent"):
        members = _shots((96, 160, 131), 2, steps, variant, channel)
        batch = VSEBatch([m[3] for m in members])
        if grouping == "lockstep":
            batch.set_group_bytes(0)
        elif grouping == "mixed":
            sizes = [m[3].device_bytes() for m in members]
            batch.set_group_bytes(sizes[0] + sizes[1])
            assert batch.groups() == [0, 2, 3, 4, 5]
        batch.set_schedule(schedule)
        batch.run(100, check_every=64)
        batch.run(steps - 100, check_every=0)
        batch.close()
        assert [m[3].step() for m in members] == [steps] * len(members)
        runs.append([_state(m[3]) for m in members])
        for m in members:
            m[3].close()
    assert any(np.abs(f["traces"]).max() > 0 for f in runs[0])
    for i, (a, b) in enumerate(zip(*runs)):
        _assert_same(b, a, ("member", i))


def test_batch_resumes_across_calls_and_reset():
    steps = 150
    members = _shots((112,), 3, steps)
    gs = [m[3] for m in members]
    batch = VSEBatch(gs)
    batch.run(70, check_every=32)
    batch.run(steps - 70, check_every=0)
    first = [_state(g) for g in gs]
    for g in gs:
        g.reset()
    batch.run(steps, check_every=50)
    for g, f in zip(gs, first):
        assert g.step() == steps
        _assert_same(_state(g), f, "rerun")
    # a member advanced alone between batch calls keeps its own step counter
    gs[0].reset()
    gs[0].run(10)
    for g in gs[1:]:
        g.reset()
    batch.run(5)
    assert [g.step() for g in gs] == [15, 5, 5]
    batch.close()
    for g in gs:
        g.close()


@pytest.mark.parametrize("schedule", ["graph", "persistent"])
def test_batch_instability_names_member(schedule):
    print(f"batch instability {schedule}", flush=True)
    members = _shots((96,), 2, 80)
    gs = [m[3] for m in members]
    sim, law = members[1][0], members[1][1]
    sim.select_law(law)
    sim._source_term = (sim._source_term * np.float32(3e37)).astype(np.float32)   # |v|^2 overflows
    gs[1].load_sources(sim)
    gs[1].reset()
    batch = VSEBatch(gs)
    batch.set_schedule(schedule)
    with pytest.raises(StabilityError) as e:
        batch.run(80, check_every=20)
    assert "member 1" in str(e.value)
    batch.close()
    for g in gs:
        g.close()


def test_batch_rejects_mixed_members():
    a = _shots((96,), 1, 20)[0][3]
    b = _shots((96,), 1, 20, variant="taichi")[0][3]
    with pytest.raises(Exception):
        VSEBatch([a, b])
    with pytest.raises(Exception):
        VSEBatch([a, a])
    a.close()
    b.close()


def test_closing_a_member_closes_its_batch():
    # the batch keeps raw member handles: closing a member must not leave it dangling
    gs = [m[3] for m in _shots((96,), 2, 20)]
    batch = VSEBatch(gs)
    batch.run(5)
    gs[0].close()
    with pytest.raises(ValueError, match="closed"):
        batch.run(5)
    with pytest.raises(ValueError):
        VSEBatch(gs)
    assert gs[1].step() == 5
    gs[1].close()


def test_run_laws_batched_equals_sequential(tmp_path):
    """Simulator.run() with laws handed over in groups gives the same
    per-law results, in the same order, as one law per implementation()."""
    cfg, laws = workloads.multi_shot_2d(sizes=(128,), n_shots=5, steps=200, n_elem=8)[0]
    res = {}
    for k in (1, 3):
        sim = SimulatorHIP(cfg)
        sim._emission_laws = laws
        res[k] = sim.run(laws_per_call=k)
    assert len(res[1]) == len(res[3]) == 5
    for a, b in zip(res[1], res[3]):
        assert np.array_equal(a["stress"], b["stress"])
        assert np.array_equal(a["sens_stress"], b["sens_stress"])
    # different laws give different traces
    assert not np.array_equal(res[1][0]["sens_stress"], res[1][1]["sens_stress"])


def test_batch_row_groups_bit_identical():
    steps = 120
    ref = None
    for rows in (1, 3, 16):
        members = _shots((96, 131), 2, steps)
        gs = [m[3] for m in members]
        b = VSEBatch(gs)
        b.set_rows(rows)
        b.run(steps, check_every=50)
        b.close()
        st = [_state(g) for g in gs]
        if ref is None:
            ref = st
        else:
            for a, c in zip(st, ref):
                _assert_same(a, c, f"rows {rows}")
        for g in gs:
            g.close()


def test_batch_fast_math_matches_solo_fast_runs():
    """A batch takes its members' arithmetic mode: FAST batch members end exactly
    where FAST solo runs end, and members whose modes differ are refused."""
    steps = 150
    print("batch fast math", flush=True)
    members = _shots((96, 131), 2, steps)
    gs = [m[3] for m in members]
    for g in gs:
        g.set_math("fast")
    batch = VSEBatch(gs)
    batch.set_group_bytes(0)
    batch.run(steps, check_every=64)
    for sim, law, att, g in members:
        sim.select_law(law)
        solo = VSESim.from_simulator(sim, variant="wgsl", att=att)
        solo.set_math("fast")
        solo.run(steps, check_every=64)
        _assert_same(_state(g), _state(solo), ("fast", sim._nx, law))
        strict = VSESim.from_simulator(sim, variant="wgsl", att=att)
        strict.run(steps, check_every=64)
        # one extra FP32 rounding per term: close to, but not the same bits as, the strict run
        rel = oracle.rel_l2(_state(g)["syy"], _state(strict)["syy"])
        assert rel <= 1e-5, rel
        solo.close()
        strict.close()
    gs[0].set_math("strict")
    with pytest.raises(Exception, match="arithmetic mode"):
        batch.run(5)
    batch.close()
    for g in gs:
        g.close()


@pytest.mark.parametrize("schedule", ["graph", "persistent"])
def test_batch_two_mechanisms_matches_solo_runs(schedule):
    # L = 2 mechanisms: the batched and solo kernels' runtime mechanism loop,
    # under both batch schedules (the persistent kernel's runtime-L instance)
    from taichi_sim_amd.sim_support.attenuation import fit_constant_q
    steps = 150
    att = fit_constant_q(50.0, 50.0, 2.0, 8.0, n_sls=2)
    members = []
    for cfg, laws in workloads.multi_shot_2d(sizes=(96, 131), n_shots=2, steps=steps, n_elem=8):
        sim = Simulator(cfg)
        sim._emission_laws = laws
        for law in range(2):
            sim.select_law(law)
            members.append((sim, law, VSESim.from_simulator(sim, variant="wgsl", att=att)))
    print(f"batch of 4 members with 2 mechanisms, lockstep, {schedule} schedule", flush=True)
    batch = VSEBatch([m[2] for m in members])
    batch.set_group_bytes(0)
    batch.set_schedule(schedule)
    batch.run(steps, check_every=100)
    batch.close()
    for sim, law, g in members:
        got = _state(g)
        sim.select_law(law)
        solo = VSESim.from_simulator(sim, variant="wgsl", att=att)
        solo.run(steps, check_every=100)
        _assert_same(got, _state(solo), (sim._nx, law))
        solo.close()
    sim, law, g = members[-1]
    sim.select_law(law)
    o = oracle.Oracle(sim, variant="wgsl", att=att).run(steps)
    for k in ("vx", "vy", "syy"):
        assert oracle.rel_l2(g.field(k), o.f[k]) <= 1e-5, k
    for m in members:
        m[2].close()
