"""Schedule tuner (seisfat2d_amd/tune.py) on the CPU: the workload key, the
coordinate-descent search over seisfat_sched_opts and the on-disk cache,
with a deterministic cost model in place of GPU timing; the option struct's
validation.  The GPU side (every option bit-exact, a real tuning run) is in
test_gpu_tune.py."""
import json
from types import SimpleNamespace

import pytest

from seisfat2d_amd import capi, tune

EIK = SimpleNamespace(nz=1001, nx=4001, dz=5.0, dx=5.0)
SHOTS = [(100.0 * k, 0.0, 10 * k, 1200 + 64 * (k % 5), 0, 0) for k in range(20)]


def cost_model(best):
    """Fake sweep time: minimal at `best`, each differing option costing a
    distinct amount; records every call."""
    calls = []

    def measure(eik, shots, rx, rz, type, opts, runs=5, warmup=2):
        calls.append(dict(opts))
        ms = 10.0 + sum((i + 1) * 0.3 for i, n in enumerate(tune.ORDER) if opts[n] != best[n])
        return ms, dict(opts)
    return measure, calls


def test_key_separates_geometry_width_mix_and_type():
    k = tune.workload_key(EIK, SHOTS, device="dev")
    assert k == tune.workload_key(EIK, list(reversed(SHOTS)), device="dev")   # order-free histogram
    assert k != tune.workload_key(SimpleNamespace(nz=1002, nx=4001, dz=5.0, dx=5.0), SHOTS, device="dev")
    assert k != tune.workload_key(EIK, SHOTS[:-1], device="dev")
    assert k != tune.workload_key(EIK, SHOTS, type=capi.ANI, device="dev")
    wider = [s[:3] + (s[3] + 64,) + s[4:] for s in SHOTS]
    assert k != tune.workload_key(EIK, wider, device="dev")
    assert k != tune.workload_key(EIK, SHOTS, device="other")
    assert tune.width_histogram([(0, 0, 0, 63), (0, 0, 0, 64), (0, 0, 0, 127)]) == "0:1,1:2"


@pytest.mark.parametrize("best", [tune.DEFAULT, {"pairs": 0, "wpcu": 8, "skip": 0, "priority": 1},
                                  {"pairs": 1, "wpcu": 6, "skip": 1, "priority": 0}])
def test_descent_finds_the_optimum_and_caches_it(tmp_path, monkeypatch, best):
    monkeypatch.setenv("SEISFAT_TUNE_CACHE", str(tmp_path / "c.json"))
    measure, calls = cost_model(best)
    r = tune.tune(EIK, SHOTS, measure=measure, device="dev")
    assert r["opts"] == best and not r["cached"]
    assert len(calls) == len({tuple(c[n] for n in tune.ORDER) for c in calls})   # no candidate twice
    assert len(calls) <= 1 + sum(len(v) - 1 for v in tune.SPACE.values())
    assert all(c[n] in tune.SPACE[n] for c in calls for n in tune.ORDER)
    d = json.loads((tmp_path / "c.json").read_text())
    assert d[r["key"]]["opts"] == best
    # cached: no measurement; force: measured again
    r2 = tune.tune(EIK, SHOTS, measure=measure, device="dev")
    assert r2["cached"] and r2["opts"] == best and len(calls) == len(r["trials"])
    r3 = tune.tune(EIK, SHOTS, measure=measure, device="dev", force=True)
    assert not r3["cached"] and r3["opts"] == best and len(calls) == 2 * len(r["trials"])


def test_small_gains_keep_the_defaults(tmp_path, monkeypatch):
    monkeypatch.setenv("SEISFAT_TUNE_CACHE", str(tmp_path / "c.json"))

    def measure(eik, shots, rx, rz, type, opts, runs=5, warmup=2):
        return (10.0 if opts == tune.DEFAULT else 10.0 * (1 - tune.MIN_GAIN / 2)), dict(opts)
    assert tune.tune(EIK, SHOTS, measure=measure, device="dev")["opts"] == tune.DEFAULT


def test_unreadable_cache_is_remeasured_and_rewritten(tmp_path, monkeypatch):
    p = tmp_path / "sub" / "c.json"
    monkeypatch.setenv("SEISFAT_TUNE_CACHE", str(p))
    p.parent.mkdir()
    p.write_text("{not json")
    measure, calls = cost_model(tune.DEFAULT)
    r = tune.tune(EIK, SHOTS, measure=measure, device="dev")
    assert calls and not r["cached"]
    assert json.loads(p.read_text())[r["key"]]["opts"] == tune.DEFAULT


def test_sched_opts_struct():
    o = capi.SchedOpts.of({"wpcu": 7, "skip": 0})
    assert o.as_dict() == {"pairs": 1, "wpcu": 7, "skip": 0, "priority": 1}
    assert capi.SchedOpts.of(None).as_dict() == tune.DEFAULT
    with pytest.raises(ValueError):
        capi.SchedOpts.of({"tile": 32})
