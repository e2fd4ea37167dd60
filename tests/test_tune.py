"""The autotuner's host side (mhd_simulator_amd/tune.py) and the process-wide launch shapes of
libmhdp.so (mhdp_get/set_launch_tuning: host state only, no GPU call): candidate enumeration,
validation, the search's choice and bookkeeping with a scripted clock, persistence."""
import os

import pytest

from mhd_simulator_amd import tune

DEFAULT = {"grid_cap": 1024, "brick_cap": 1024, "march_cap": 1024, "march_z": 32, "brick_occ": 2}


@pytest.fixture
def restore():
    before = tune.current()
    yield before
    tune.set_launch(before)


def test_defaults_and_every_candidate_is_a_valid_shape(restore, monkeypatch):
    for k in ("MHDP_GRID_CAP", "MHDP_BRICK_CAP", "MHDP_MARCH_CAP", "MHDP_MARCH_Z", "MHDP_BRICK_OCC"):
        assert k not in os.environ  # the defaults below are the built-in ones
    assert restore == DEFAULT
    cands = tune.candidates()
    assert len(cands) == 3 * 2 * 2 * 3
    assert len({str(c) for c in cands}) == len(cands)
    for c in cands:
        tune.set_launch(c["launch"])
        assert {k: tune.current()[k] for k in c["launch"]} == c["launch"]
        assert c["opts"]["coarse_visits"] in (0, 1, 2)


def test_inner_tolerance_axis_enters_the_schedule_only_when_set():
    cands = tune.candidates(march_z=(32,), march_cap=(1024,), brick_occ=(2,), coarse_visits=(0,),
                            inner_tol=(0.0, 3e-5))
    assert [c["opts"] for c in cands] == [{"coarse_visits": 0}, {"coarse_visits": 0, "inner_tol": 3e-5}]


@pytest.mark.parametrize("bad", [{"march_z": 3}, {"march_z": 0}, {"march_z": 514}, {"march_cap": 4},
                                 {"grid_cap": 5000}, {"brick_occ": 1}, {"brick_occ": 5}])
def test_invalid_shapes_are_rejected_and_change_nothing(restore, bad):
    with pytest.raises(ValueError):
        tune.set_launch(bad)
    assert tune.current() == restore


def test_unknown_shape_key_is_an_error(restore):
    with pytest.raises(KeyError):
        tune.set_launch({"warp_size": 32})


def test_search_picks_the_fastest_converging_candidate_and_restores_the_shapes(restore):
    cands = tune.candidates(march_z=(16, 32, 64), march_cap=(1024,), brick_occ=(2,), coarse_visits=(1,))
    cost = {16: 5.0, 32: 2.0, 64: 1.0}  # ms per call; z 64 is fastest but fails to converge
    seen = []
    clock_t = [0.0]

    def clock():
        return clock_t[0]

    def run(c):
        z = c["launch"]["march_z"]
        assert tune.current()["march_z"] == z  # the candidate's shapes are in force while it runs
        seen.append(z)
        clock_t[0] += cost[z] * 1e-3
        if z == 64:
            raise RuntimeError("did not converge")
        return True

    best, table = tune.search(run, cands, repeats=3, clock=clock)
    assert best["launch"]["march_z"] == 32 and best["ms"] == pytest.approx(2.0)
    assert [r["ok"] for r in table] == [True, True, False] and "did not converge" in table[2]["error"]
    assert seen == [16] * 4 + [32] * 4 + [64]  # one untimed call + 3 timed per candidate
    assert tune.current() == restore


def test_search_without_a_converging_candidate_raises(restore):
    cands = tune.candidates(march_z=(16,), march_cap=(1024,), brick_occ=(2,), coarse_visits=(1, 2))
    with pytest.raises(RuntimeError, match="no candidate converged"):
        tune.search(lambda c: False, cands, repeats=1)
    assert tune.current() == restore


def test_results_persist_by_key(tmp_path, restore):
    path = str(tmp_path / "tuning.json")
    a = {"key": "64x64x64|mixed|1e-08", "launch": {"march_z": 16, "march_cap": 2048, "brick_occ": 3},
         "opts": {"coarse_visits": 2}, "ms": 4.5, "table": []}
    b = dict(a, key="32x32x32|mixed|1e-08", ms=1.0)
    tune.save(a, path)
    tune.save(b, path)
    got = tune.load(path, a["key"])
    assert got["launch"] == a["launch"] and got["opts"] == a["opts"] and "table" not in got
    assert tune.apply(got) == {"coarse_visits": 2}
    assert {k: tune.current()[k] for k in a["launch"]} == a["launch"]
    with pytest.raises(KeyError):
        tune.load(path, "missing")


def test_schedule_stage_covers_smoothers_weights_and_scales():
    cands = tune.schedules()
    assert len(cands) == 2 * 3 * 4 * 3
    assert all(c["launch"] == tune.BUILTIN_LAUNCH for c in cands)
    assert {c["opts"]["smoother"] for c in cands} == {"jacobi", "red-black"}
    # the library's defaults are left out of the options
    assert {c["opts"].get("smooth_weight", 0.0) for c in cands} == {0.0, 0.8, 0.95}
    assert {c["opts"].get("coarse_scale", 0.0) for c in cands} == {0.0, 1.4, 1.8, 1.95}
    assert not any("smooth_weight" in c["opts"] and c["opts"]["smooth_weight"] == 0.0 for c in cands)


def test_two_stage_search_tunes_the_launch_at_the_winning_schedule(restore):
    """Stage 1 picks the fastest schedule at the built-in shapes, stage 2 times the launch shapes
    with that schedule; a schedule whose cycle breaks down (the run raises) never wins."""
    clock_t = [0.0]
    seen = []

    def cost(c):
        o = c["opts"]
        ms = 10.0 - (3.0 if o["smoother"] == "red-black" else 0.0) - (1.0 if o.get("coarse_scale") == 1.4 else 0.0)
        return ms - (0.5 if c["launch"]["march_z"] == 64 else 0.0)

    def run(c):
        o = c["opts"]
        if o["smoother"] == "red-black" and o.get("coarse_scale") == 1.95:
            raise RuntimeError("the cycle broke down")
        seen.append(c)
        clock_t[0] += cost(c) * 1e-3
        return True

    best, table = tune.two_stage(run, 1, tune.schedules(smooth_weight=(0.0,), coarse_visits=(1,)),
                                 {"march_z": (32, 64), "march_cap": (1024,), "brick_occ": (2,)},
                                 clock=lambda: clock_t[0])
    assert best["opts"] == {"coarse_visits": 1, "smoother": "red-black", "coarse_scale": 1.4}
    assert best["launch"]["march_z"] == 64 and best["ms"] == pytest.approx(5.5)
    stage2 = table[-2:]
    assert all(r["opts"] == best["opts"] for r in stage2)
    assert any(not r["ok"] and "broke down" in r.get("error", "") for r in table)
    assert tune.current() == restore
