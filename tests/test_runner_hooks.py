"""The runner's callbacks and loggers (crates/com-sim/src/callbacks.rs, loggers.rs, the hooks
of runner.rs:309-415) through the C ABI, with scenarios given as callbacks drawing from the
oracle's restatement of the reference's trial streams. CPU only: no kernel runs here.

Mirrors the reference's own tests and doc examples: runner.rs:550-561 (a callback stops a
condition), callbacks.rs:19-41 (SkipAfter), loggers.rs:20-41 (CSV header and rows, JSON
scenario), loggers.rs:323-339 (the dropped values in the JSON report).
"""
import json
import math
import os

import pytest

import comfec as F
import oracle as O


def coin(seed, name="coin", p=0.3):
    """runner.rs:487-493: heads with probability p (or the condition's "p") and a uniform value."""
    def run(index, cond, t0, t1, rec):
        heads = 0
        for t in range(t0, t1):
            r = O.Rng.stream(seed, name, index, t)
            heads += r.uniform() < cond.get("p", p)
            rec.value("u", r.uniform())
        rec.errors("heads", heads, t1 - t0)
    return run


def toss(seed=0, name="toss"):
    """loggers.rs:30: one fair bit per trial as the error rate "p"."""
    def run(index, cond, t0, t1, rec):
        rec.errors("p", sum(O.Rng.stream(seed, name, index, t).bit() for t in range(t0, t1)), t1 - t0)
    return run


class Once:
    def on_batch_end(self, progress):
        return F.Action.STOP_CONDITION


def test_a_callback_stops_a_condition_after_a_batch():
    # runner.rs:550-561
    r = F.Runner().batch(8).stop(F.Stop.max_trials(100)).callback(Once()).run("coin", coin(0), F.Grid())
    assert (r[0].status, r[0].trials) == ("stopped", 8)
    # the rule wins when it is met at the same batch
    r = F.Runner().batch(8).stop(F.Stop.max_trials(8)).callback(Once()).run("coin", coin(0), F.Grid())
    assert (r[0].status, r[0].trials) == ("completed", 8)


def test_skip_after_a_floor_along_an_axis():
    # callbacks.rs:19-41: p = 10^(-snr/5); the point at 30 dB is skipped once the estimate at
    # 20 dB is below 1e-3
    def decay(index, cond, t0, t1, rec):
        p = 10.0 ** (-cond["snr_db"] / 5.0)
        rec.errors("ber", sum(O.Rng.stream(0, "decay", index, t).uniform() < p for t in range(t0, t1)), t1 - t0)
    g = F.Grid().axis("snr_db", [0.0, 10.0, 20.0, 30.0])
    r = F.Runner().stop(F.Stop.max_trials(2000)).callback(F.SkipAfter.below("ber", 1e-3, "snr_db")).run(
        "decay", decay, g)
    assert [c.status for c in r] == ["completed", "completed", "completed", "skipped"]
    assert r[3].trials == 0 and r[3].metrics == {}
    # along the last axis only: each value of the other key has its own sweep
    g = F.Grid().axis("x", [1, 2]).axis("snr_db", [20.0, 30.0])
    r = F.Runner().stop(F.Stop.max_trials(2000)).callback(F.SkipAfter.below("ber", 1e-3, "snr_db")).run(
        "decay", decay, g)
    assert [c.status for c in r] == ["completed", "skipped", "completed", "skipped"]


class Recorder:
    """Every hook, in the order the runner calls them."""

    def __init__(self, events, tag, stop_run_after=None):
        self.events, self.tag, self.stop_run_after = events, tag, stop_run_after

    def on_run_start(self, info):
        self.events.append((self.tag, "run_start", info.scenario, info.n_conditions))

    def on_condition_start(self, index, condition):
        self.events.append((self.tag, "condition_start", index, condition))

    def on_batch_end(self, p):
        m = p.metric("heads")
        self.events.append((self.tag, "batch_end", p.condition, p.trials, m.total))

    def on_condition_end(self, r):
        self.events.append((self.tag, "condition_end", r.index, r.status, r.trials))
        return F.Action.STOP_RUN if r.index == self.stop_run_after else F.Action.CONTINUE

    def on_run_end(self, results):
        self.events.append((self.tag, "run_end", [c.status for c in results]))


class ListLogger:
    def __init__(self, events):
        self.events = events

    def start(self, info):
        self.events.append(("log", "start", info.seed))

    def log(self, r, confidence):
        self.events.append(("log", "log", r.index, confidence))

    def finish(self, info, results, wall_s):
        self.events.append(("log", "finish", len(results), wall_s >= 0.0))


def test_the_hooks_in_order_and_a_run_stopped_by_a_callback():
    ev = []
    g = F.Grid().axis("p", [0.1, 0.5, 0.9])
    r = (F.Runner().seed(4).batch(10).stop(F.Stop.max_trials(25)).confidence(0.9)
         .callback(Recorder(ev, "a", stop_run_after=1)).callback(Recorder(ev, "b")).logger(ListLogger(ev))
         .run("coin", coin(4), g))
    assert [c.status for c in r] == ["completed", "completed", "skipped"]
    assert [c.trials for c in r] == [25, 25, 0]
    want = [("a", "run_start", "coin", 3), ("b", "run_start", "coin", 3), ("log", "start", 4)]
    for i in (0, 1):
        want += [("a", "condition_start", i, {"p": [0.1, 0.5][i]}), ("b", "condition_start", i, {"p": [0.1, 0.5][i]})]
        for t in (10, 20, 25):
            want += [("a", "batch_end", i, t, t), ("b", "batch_end", i, t, t)]
        want += [("a", "condition_end", i, "completed", 25), ("b", "condition_end", i, "completed", 25),
                 ("log", "log", i, 0.9)]
    # after StopRun the next condition is skipped without asking the callbacks
    want += [("a", "condition_end", 2, "skipped", 0), ("b", "condition_end", 2, "skipped", 0), ("log", "log", 2, 0.9)]
    want += [("a", "run_end", ["completed", "completed", "skipped"]),
             ("b", "run_end", ["completed", "completed", "skipped"]), ("log", "finish", 3, True)]
    assert ev == want


def test_a_skip_from_any_callback_skips_and_every_callback_is_asked():
    ev = []

    class SkipFirst:
        def on_condition_start(self, index, condition):
            return F.Action.SKIP if index == 0 else F.Action.CONTINUE

    r = (F.Runner().stop(F.Stop.max_trials(5)).callback(SkipFirst()).callback(Recorder(ev, "r"))
         .run("coin", coin(0), F.Grid().axis("x", [1, 2])))
    assert [c.status for c in r] == ["skipped", "completed"]
    assert ("r", "condition_start", 0, {"x": 1.0}) in ev


def test_csv_report_header_rows_and_json_report(tmp_path):
    # loggers.rs:20-41
    csv_path, json_path = tmp_path / "out" / "toss.csv", tmp_path / "toss.json"
    r = (F.Runner().stop(F.Stop.max_trials(64)).logger(F.CsvLogger(csv_path)).logger(F.JsonLogger(json_path))
         .run("toss", toss(), F.Grid().axis("n", [1, 2])))
    text = csv_path.read_text()
    assert text.startswith("index,n,status,trials,wall_s,p,p_lo,p_hi,p_count,p_total\n")
    lines = text.splitlines()
    assert len(lines) == 3
    for i, line in enumerate(lines[1:]):
        cells = line.split(",")
        heads = sum(O.Rng.stream(0, "toss", i, t).bit() for t in range(64))
        lo, hi = O.wilson(heads, 64, 0.95)
        assert cells[:4] == [str(i), str(i + 1), "completed", "64"]
        assert len(cells[4].split(".")[1]) == 6  # the wall time to the microsecond
        assert float(cells[5]) == heads / 64 and float(cells[6]) == pytest.approx(lo, rel=1e-15)
        assert float(cells[7]) == pytest.approx(hi, rel=1e-15)
        assert cells[8:] == [str(heads), "64"]
        assert "e" in cells[5] and "e+" not in cells[5] and "e-0" not in cells[5]
    doc = json.loads(json_path.read_text())
    assert '"scenario": "toss"' in json_path.read_text()
    assert doc["seed"] == 0 and doc["batch"] == 32 and doc["stop"] == "max_trials(64)"
    assert doc["confidence"] == 0.95 and doc["config"] is None and doc["version"]
    c = doc["commit"]
    h = c[:-len("-dirty")] if c.endswith("-dirty") else c
    assert h == "unknown" or (len(h) >= 40 and all(ch in "0123456789abcdef" for ch in h))
    assert [x["condition"] for x in doc["conditions"]] == [{"n": 1}, {"n": 2}]
    p = doc["conditions"][0]["metrics"]["p"]
    assert p["kind"] == "rate" and p["total"] == 64 and p["value"] == r[0].metric("p").value
    assert doc["conditions"][0]["dropped"] == {} and doc["conditions"][0]["status"] == "completed"


def test_the_json_report_lists_the_dropped_values(tmp_path):
    # loggers.rs:323-339
    path = tmp_path / "r.json"

    def drops(index, cond, t0, t1, rec):
        for _ in range(t0, t1):
            rec.value("snr_db", float("inf"))
            rec.value("evm", 0.1)
    F.Runner().batch(4).stop(F.Stop.max_trials(6)).logger(F.JsonLogger(path)).run("drops", drops, F.Grid())
    text = path.read_text()
    assert '"dropped": {"snr_db": 6}' in text
    evm = json.loads(text)["conditions"][0]["metrics"]["evm"]
    assert evm["kind"] == "mean" and evm["count"] == 6 and evm["value"] == pytest.approx(0.1, rel=1e-15)


def test_csv_rows_of_skipped_conditions_and_empty_cells(tmp_path):
    path = tmp_path / "s.csv"

    class SkipFirst:
        def on_condition_start(self, index, condition):
            return F.Action.SKIP if index == 0 else F.Action.CONTINUE

    # a mean of one value has an infinite interval: its bounds are empty cells
    def one(index, cond, t0, t1, rec):
        rec.value("m", 2.5)
    (F.Runner().stop(F.Stop.max_trials(1)).callback(SkipFirst()).logger(F.CsvLogger(path))
     .run("one", one, F.Grid().axis("x", [7, 8])))
    lines = path.read_text().splitlines()
    assert lines[0] == "index,x,status,trials,wall_s,m,m_lo,m_hi,m_count,m_total"
    # the skipped condition waits for the header and keeps its place
    assert lines[1] == "0,7,skipped,0,0.000000,,,,,"
    assert lines[2].startswith("1,8,completed,1,") and lines[2].endswith(",2.5e0,,,1,1")
    # every condition skipped: the header of the first, no metric columns
    path2 = tmp_path / "all.csv"

    class SkipAll:
        def on_condition_start(self, index, condition):
            return F.Action.SKIP

    F.Runner().stop(F.Stop.max_trials(1)).callback(SkipAll()).logger(F.CsvLogger(path2)).run(
        "one", one, F.Grid().axis("x", [7, 8]))
    assert path2.read_text().splitlines() == ["index,x,status,trials,wall_s", "0,7,skipped,0,0.000000",
                                              "1,8,skipped,0,0.000000"]


def test_progress_bar_lines(capfd):
    # callbacks.rs:107-166
    g = F.Grid().axis("p", [0.25])
    r = (F.Runner().seed(3).batch(50).threads(2).stop(F.Stop.max_trials(100)).callback(F.ProgressBar(0.0))
         .run("coin", coin(3), g))
    err = capfd.readouterr().err
    assert err.startswith("[coin] 1 conditions, seed 3, batches of 50, 2 threads, stop: max_trials(100)\n")
    assert "\r  condition 1/1: 50 trials, 0 s" in err and "\r  condition 1/1: 100 trials, 0 s" in err
    h, u = r[0].metric("heads"), r[0].metric("u")
    last = err.splitlines()[-1].split("\r")[-1]

    def e(x, d):
        m, ex = f"{x:.{d}e}".split("e")
        return f"{m}e{int(ex)}"
    want = (f"[1/1] (p=0.25) completed: 100 trials, {r[0].wall_s:.2f} s  heads = {e(h.value, 4)} "
            f"[{e(h.lo, 2)}, {e(h.hi, 2)}], u = {e(u.value, 4)} [{e(u.lo, 2)}, {e(u.hi, 2)}]")
    assert last.split(" s  ")[1] == want.split(" s  ")[1]
    assert last.startswith("[1/1] (p=0.25) completed: 100 trials, ")


def test_run_information_threads_config_and_dev_run(tmp_path):
    infos = []

    class Info:
        def on_run_start(self, info):
            infos.append(info)

    path = tmp_path / "c.json"
    F.Runner().seed(11).batch(7).threads(3).config('{"k": [1, 2]}').stop(F.Stop.max_trials(7)).callback(
        Info()).logger(F.JsonLogger(path)).run("cfg", coin(11), F.Grid())
    i = infos[0]
    assert (i.scenario, i.seed, i.batch, i.threads, i.stop, i.n_conditions) == ("cfg", 11, 7, 3, "max_trials(7)", 1)
    assert i.config == '{"k": [1, 2]}' and abs(i.started_unix_s - __import__("time").time()) < 600
    assert json.loads(path.read_text())["config"] == {"k": [1, 2]}
    F.Runner().fast_dev_run(True).callback(Info()).run("cfg", coin(11), F.Grid())
    assert infos[1].stop == "fast_dev_run" and infos[1].threads >= 1


def test_logger_and_callback_failures_end_the_run(tmp_path):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    with pytest.raises(F.ComfecError) as e:
        F.Runner().stop(F.Stop.max_trials(2)).logger(F.CsvLogger(blocker / "sub" / "r.csv")).run(
            "coin", coin(0), F.Grid())
    assert e.value.kind == "Io"

    class Bad:
        def log(self, r, confidence):
            raise ValueError("disk full")
    with pytest.raises(ValueError, match="disk full"):
        F.Runner().stop(F.Stop.max_trials(2)).logger(Bad()).run("coin", coin(0), F.Grid().axis("x", [1, 2]))

    seen = []

    class Boom:
        def on_batch_end(self, p):
            seen.append(p.trials)
            raise RuntimeError("callback failed")
    with pytest.raises(RuntimeError, match="callback failed"):
        F.Runner().batch(4).stop(F.Stop.max_trials(100)).callback(Boom()).run("coin", coin(0), F.Grid())
    assert seen == [4]  # the condition stopped at the first batch


def test_the_record_of_a_progress_report_is_read_only():
    class Writer:
        def on_batch_end(self, p):
            F._check(F.lib().cf_record_errors(p._rec, b"x", 1, 1))
    with pytest.raises(F.ComfecError) as e:
        F.Runner().stop(F.Stop.max_trials(2)).callback(Writer()).run("coin", coin(0), F.Grid())
    assert "read-only" in str(e.value)


def test_csv_fields_are_quoted(tmp_path):
    # loggers.rs:84-90 on a metric name that needs quotes
    path = tmp_path / "q.csv"

    def q(index, cond, t0, t1, rec):
        rec.errors('a, "b"', 0, t1 - t0)
    F.Runner().stop(F.Stop.max_trials(3)).logger(F.CsvLogger(path)).run("q", q, F.Grid())
    header = path.read_text().splitlines()[0]
    assert header == 'index,status,trials,wall_s,"a, ""b""","a, ""b""_lo","a, ""b""_hi","a, ""b""_count","a, ""b""_total"'
    row = path.read_text().splitlines()[1].split(",")
    assert row[4] == "0e0" and float(row[6]) == pytest.approx(O.wilson(0, 3, 0.95)[1], rel=1e-15)
