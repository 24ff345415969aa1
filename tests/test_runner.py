"""The campaign runner of the library (Runner / Stop / Grid / Record of crates/com-sim)
through the C ABI, with scenarios given as callbacks. CPU only: no kernel runs here.

Mirrors the reference's tests: crates/com-sim/src/runner.rs:487-580,
crates/comrs/tests/simulation.rs:24-61, stopping.rs and stats.rs tests.
"""
import math

import pytest

import comfec as F
import oracle as O


def coin_scenario(seed, name="coin", p=0.3, with_mean=True):
    """runner.rs:487-493: a coin of probability p and a uniform value per trial, each trial
    drawing from stream(seed, name, condition, trial)."""
    def run(index, cond, t0, t1, rec):
        heads = 0
        for t in range(t0, t1):
            r = O.Rng.stream(seed, name, index, t)
            heads += r.uniform() < cond.get("p", p)
            if with_mean:
                rec.value("u", r.uniform())
        rec.errors("heads", heads, t1 - t0)
    return run


def test_max_trials_is_met_exactly_whatever_the_batch():
    g = F.Grid().axis("x", [1])
    ref = None
    for batch in (1000, 32, 7, 1, 5000):
        r = F.Runner().seed(5).batch(batch).stop(F.Stop.max_trials(1000)).run(
            "coin", coin_scenario(5, with_mean=False), g, metrics=("heads",))
        assert r[0].trials == 1000 and r[0].status == "completed"
        m = r[0].metric("heads")
        if ref is None:
            ref = (m.count, m.total)
        assert (m.count, m.total) == ref
    # an error target stops at the end of a batch, the cap exactly; at least one trial
    r = F.Runner().batch(32).stop(F.Stop.min_errors("heads", 10**6) | F.Stop.max_trials(70)).run(
        "coin", coin_scenario(0), g, metrics=("heads",))
    assert r[0].trials == 70
    r = F.Runner().batch(32).stop(F.Stop.max_trials(0)).run("coin", coin_scenario(0), g, metrics=("heads",))
    assert r[0].trials == 1


def test_min_errors_stops_at_a_batch_boundary():
    g = F.Grid().axis("p", [0.5])
    r = F.Runner().seed(1).batch(10).stop(F.Stop.min_errors("heads", 25) | F.Stop.max_trials(10**6)).run(
        "coin", coin_scenario(1), g, metrics=("heads",))
    m = r[0].metric("heads")
    assert m.count >= 25 and r[0].trials % 10 == 0 and m.total == r[0].trials
    # the count at the previous batch was below the target
    prev = sum(O.Rng.stream(1, "coin", 0, t).uniform() < 0.5 for t in range(r[0].trials - 10))
    assert prev < 25


def test_grid_order_and_conditions():
    seen = []

    def scen(index, cond, t0, t1, rec):
        if t0 == 0:
            seen.append((index, dict(cond)))
        rec.errors("e", 0, t1 - t0)
    g = F.Grid().axis("a", [1, 2]).axis("b", [10, 20, 30]).fixed("c", 7)
    r = F.Runner().stop(F.Stop.max_trials(3)).run("grid", scen, g, metrics=("e",))
    assert len(r) == 6
    # the last axis fastest (grid.rs:321-361)
    assert [c for _, c in seen] == [{"a": a, "b": b, "c": 7.0} for a in (1.0, 2.0) for b in (10.0, 20.0, 30.0)]
    with pytest.raises(F.ComfecError) as e:
        F.Runner().stop(F.Stop.max_trials(3)).run("grid", scen, F.Grid().axis("a", [1]).axis("a", [2]))
    assert e.value.kind == "Config" and '"a" twice' in str(e.value)


def test_a_stopping_target_that_is_not_recorded_is_an_error():
    g = F.Grid().axis("x", [1, 2])
    rule = F.Stop.min_errors("ber", 100) | F.Stop.max_trials(10_000)
    with pytest.raises(F.ComfecError) as e:
        F.Runner().batch(4).stop(rule).run("coin", coin_scenario(0), g, metrics=("heads",))
    msg = str(e.value)
    assert "`ber`" in msg and "recorded: heads, u" in msg and "x=1" in msg
    rule = F.Stop.min_errors("u", 100) | F.Stop.max_trials(10_000)
    with pytest.raises(F.ComfecError) as e:
        F.Runner().batch(4).stop(rule).run("coin", coin_scenario(0), g, metrics=("heads",))
    assert "is a mean" in str(e.value)
    # a dev run checks it too
    rule = F.Stop.precision("evm", 0.1) | F.Stop.max_trials(10)
    with pytest.raises(F.ComfecError):
        F.Runner().fast_dev_run(True).stop(rule).run("coin", coin_scenario(0), g)


def test_confidence_outside_zero_one_and_unbounded_rules_are_errors():
    g = F.Grid()
    for c in (95.0, 0.0, 1.0, float("nan")):
        with pytest.raises(F.ComfecError) as e:
            F.Runner().confidence(c).stop(F.Stop.max_trials(1)).run("coin", coin_scenario(0), g)
        assert "between 0 and 1" in str(e.value)
    with pytest.raises(F.ComfecError):
        F.Runner().stop(F.Stop.precision("u", 0.1, 95.0) | F.Stop.max_trials(4)).run("coin", coin_scenario(0), g)
    r = F.Runner().batch(2).stop(F.Stop.precision("u", 1e-9, 0.5) | F.Stop.max_trials(4)).run(
        "coin", coin_scenario(0), g, metrics=("u",))
    assert r[0].trials == 4
    with pytest.raises(F.ComfecError) as e:
        F.Runner().stop(F.Stop.min_errors("heads", 10)).run("coin", coin_scenario(0), g)
    assert "can run forever" in str(e.value)


def test_fast_dev_run_runs_one_batch():
    r = F.Runner().batch(8).fast_dev_run(True).stop(F.Stop.max_trials(100)).run(
        "coin", coin_scenario(0), F.Grid().axis("x", [1, 2]), metrics=("heads",))
    assert [c.trials for c in r] == [8, 8] and r[0].status == "dev_run"


def test_wilson_intervals_and_means_match_the_oracle():
    r = F.Runner().seed(9).batch(64).stop(F.Stop.max_trials(1000)).run(
        "coin", coin_scenario(9), F.Grid().axis("x", [1]), metrics=("heads", "u"))
    m = r[0].metric("heads")
    lo, hi = O.wilson(m.count, m.total, 0.95)
    assert (m.lo, m.hi) == pytest.approx((lo, hi), rel=1e-15)
    u = r[0].metric("u")
    vals = [O.Rng.stream(9, "coin", 0, t) for t in range(1000)]
    xs = []
    for rr in vals:
        rr.uniform()
        xs.append(rr.uniform())
    mean = sum(xs) / len(xs)
    assert u.is_rate == 0 and u.count == 1000 and u.value == pytest.approx(mean, rel=1e-12)
    sd = math.sqrt(sum((x - mean) ** 2 for x in xs) / 999)
    z = O.lib.or_z_score(0.95)
    assert u.hi - u.value == pytest.approx(z * sd / math.sqrt(1000), rel=1e-9)


def test_dropped_values_reach_the_result():
    # runner.rs:569-579: non-finite values are dropped and counted
    def scen(index, cond, t0, t1, rec):
        for t in range(t0, t1):
            rec.value("snr_db", float("inf") if O.Rng.stream(0, "inf", index, t).uniform() < 0.5 else 10.0)
    r = F.Runner().batch(8).stop(F.Stop.max_trials(64)).run("inf", scen, F.Grid(), metrics=("snr_db",))
    m = r[0].metric("snr_db")
    assert m.count > 0 and m.dropped > 0 and m.count + m.dropped == 64


def test_a_failing_trial_stops_the_run_with_context():
    # crates/comrs/tests/simulation.rs:115-128
    def scen(index, cond, t0, t1, rec):
        raise F.ComfecError(F.CF_NUMERICAL, "numerical failure: no convergence")
    with pytest.raises(F.ComfecError) as e:
        F.Runner().batch(4).stop(F.Stop.max_trials(8)).run("bad", scen, F.Grid().axis("x", [3]))
    assert e.value.kind == "Numerical" and "x=3" in str(e.value) and "no convergence" in str(e.value)


def test_wilson_of_the_library_equals_the_oracle():
    for k, n in [(0, 1000), (37, 1000), (1, 1), (200, 1_000_000), (999, 1000)]:
        for c in (0.95, 0.99):
            assert F.wilson(k, n, c) == pytest.approx(O.wilson(k, n, c), rel=1e-15, abs=0)
    with pytest.raises(F.ComfecError):
        F.wilson(5, 3)
