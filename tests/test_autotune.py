"""The autotuner's search (coordinate descent over the device options) against synthetic scorers with
known optima: no device is needed, the scorer stands in for the timed calibration workload."""
import math

from minirender_amd import autotune

KNOBS = {"path_capacity": [1 << 22, 1 << 24, 1 << 26, 1 << 28], "persistent_blocks_per_cu": [0, 4, 6],
         "traversal": ["split", "fused", "refill", "auto"], "use_graph": [False, True], "max_jobs_in_flight": [16, 64]}


def separable(opts):
    """A separable score whose optimum is (2^26, 6, refill, graph, 16 jobs)."""
    v = -abs(math.log2(opts["path_capacity"]) - 26)
    v += {0: 0.0, 4: 0.5, 6: 1.0}[opts["persistent_blocks_per_cu"]]
    v += {"split": 0.0, "fused": 0.3, "refill": 0.9, "auto": 0.6}[opts["traversal"]]
    v += 0.2 if opts["use_graph"] else 0.0
    v += 0.1 if opts.get("max_jobs_in_flight") == 16 else 0.0
    return 100.0 + v, [0.01]


def test_coordinate_descent_finds_a_separable_optimum():
    calls = []

    def scorer(o):
        calls.append(dict(o))
        return separable(o)

    r = autotune.tune(None, None, schedule="batched", knobs=KNOBS, scorer=scorer)
    assert r["options"] == {"path_capacity": 1 << 26, "persistent_blocks_per_cu": 6, "traversal": "refill",
                            "use_graph": True, "max_jobs_in_flight": 16}
    assert r["samples_per_s"] == separable(r["options"])[0]
    # every distinct option set is measured once (the defaults are a candidate of every sweep)
    keys = {tuple(sorted(c.items())) for c in calls}
    assert len(keys) == len(calls) == len(r["trials"])
    assert len(calls) <= 1 + sum(len(v) for v in KNOBS.values())


def test_plain_schedules_do_not_tune_jobs_in_flight():
    r = autotune.tune(None, None, schedule="plain", knobs=KNOBS, scorer=separable)
    assert "max_jobs_in_flight" not in r["options"]
    assert all("max_jobs_in_flight" not in t["options"] for t in r["trials"])


def test_descent_keeps_the_defaults_when_nothing_beats_them():
    flat = lambda o: (1.0, [1.0])  # noqa: E731
    r = autotune.tune(None, None, schedule="batched", knobs=KNOBS, scorer=flat)
    assert r["options"] == {k: autotune.DEFAULTS[k] for k in KNOBS}


def test_second_round_follows_a_coupled_optimum():
    """The best pool size depends on the traversal kernel: one pass finds the kernel, a second pass
    re-sweeps the pool size under it."""
    def coupled(o):
        if o["traversal"] != "refill":
            return 0.0, [0.0]
        return {1 << 26: 0.5, 1 << 28: 1.0}.get(o["path_capacity"], 0.0), [0.0]

    knobs = {k: KNOBS[k] for k in ("path_capacity", "traversal")}
    one = autotune.tune(None, None, schedule="plain", knobs=knobs, rounds=1, scorer=coupled)
    two = autotune.tune(None, None, schedule="plain", knobs=knobs, rounds=2, scorer=coupled)
    assert one["options"] == {"path_capacity": 1 << 26, "traversal": "refill"} and one["samples_per_s"] == 0.5
    assert two["options"] == {"path_capacity": 1 << 28, "traversal": "refill"} and two["samples_per_s"] == 1.0


def test_save_and_load_round_trip(tmp_path):
    r = autotune.tune(None, None, schedule="batched", knobs=KNOBS, scorer=separable)
    path = tmp_path / "tuned.json"
    autotune.save(r, path)
    assert autotune.load(path) == r["options"]


def test_refused_candidates_score_zero_and_are_recorded():
    """A candidate whose pools leave no room for a job in the HBM budget makes the device refuse every
    job (NotAdmitted); the search records it as infeasible and keeps going."""
    from minirender_amd import _native as N

    def scorer(o):
        if o["path_capacity"] == 1 << 28 and o.get("batches_in_flight") == 2:
            raise N.NotAdmitted(3, "job exceeds the device HBM budget")
        v = math.log2(o["path_capacity"]) + (0.5 if o.get("batches_in_flight") == 2 else 0.0)
        return v, [0.01]

    knobs = {"path_capacity": [1 << 26, 1 << 28], "batches_in_flight": [1, 2]}
    r = autotune.tune(None, None, schedule="plain", knobs=knobs, scorer=scorer)
    assert r["options"] == {"path_capacity": 1 << 28, "batches_in_flight": 1}
    refused = [t for t in r["trials"] if "infeasible" in t]
    assert len(refused) == 1 and refused[0]["samples_per_s"] == 0.0
    assert refused[0]["options"] == {"path_capacity": 1 << 28, "batches_in_flight": 2}
