"""Host logic of the launch autotuner: the lane tables it derives from the
single-system timings (LDS budget and fewer-lanes tolerance) and the choice
tune_batch makes among whole-batch timings, with the device timing replaced
by a table of made-up costs."""
import pytest

import mole_amd as M
from mole_amd import autotune as A

# N=2: L=2 fastest, L=4 12.5 % and L=1 30 % slower; N=5: L=4 25 % slower than L=8,
# L=16 over the device; LDS shrinks with L
TIMINGS = {"2:1": 1.04, "2:2": 0.80, "2:4": 0.90,
           "5:4": 2.0, "5:8": 1.6, "5:16": None,
           "_lds": {"2:1": 8000, "2:2": 4000, "2:4": 2000, "5:4": 30000, "5:8": 15000, "5:16": 9000}}


def test_table_for_budget():
    assert A.table_for_budget(TIMINGS, None) == {2: 2, 5: 8}
    assert A.table_for_budget(TIMINGS, 3000) == {2: 4, 5: 8}  # 5 has nothing under 3000: smallest timed footprint
    assert A.table_for_budget(TIMINGS, 20000) == {2: 2, 5: 8}


def test_table_within():
    assert A.table_within(TIMINGS, 0.0) == {2: 2, 5: 8}
    assert A.table_within(TIMINGS, 0.15) == {2: 2, 5: 8}
    assert A.table_within(TIMINGS, 0.31) == {2: 1, 5: 4}
    assert A.table_within(TIMINGS, 0.26) == {2: 2, 5: 4}
    assert A.table_within(TIMINGS, 0.24) == {2: 2, 5: 8}


class _Ctx:
    def __init__(self):
        self.schedule = None

    def set_tile_schedule(self, s):
        self.schedule = s


@pytest.mark.parametrize("fewer_lanes_wins", [False, True])
def test_tune_batch_picks_fastest(monkeypatch, fewer_lanes_wins):
    applied = []

    def cost(ctx, systems, table, schedule, chunk, reps):
        base = 1.0 if schedule == M.SCHEDULE_CLASSES else 1.3
        if table == {2: 1, 5: 4}:
            base *= 0.8 if fewer_lanes_wins else 1.2
        return base * (1.0 - 0.01 * chunk)

    monkeypatch.setattr(A, "_time_config", cost)
    monkeypatch.setattr(A, "apply", lambda ctx, systems, table: applied.append(dict(table)))
    ctx = _Ctx()
    best, trials = A.tune_batch(ctx, [], TIMINGS, budgets=(3000, None), tolerances=(0.15, 0.31))
    # tolerance 0.15 repeats the budget-None table and is not timed again
    tol_trials = [t for t in trials if t.get("tolerance") is not None and t["chunk"] == 5]
    assert [t["tolerance"] for t in tol_trials] == [0.31]
    assert tol_trials[0]["schedule"] == M.SCHEDULE_CLASSES
    assert best["chunk"] == 20 and best["schedule"] == M.SCHEDULE_CLASSES
    assert ctx.schedule == M.SCHEDULE_CLASSES
    if fewer_lanes_wins:
        assert best["table"] == {2: 1, 5: 4} and best["tolerance"] == 0.31
    else:
        assert best["table"] in ({2: 2, 5: 8}, {2: 4, 5: 8}) and "tolerance" not in best
    assert applied[-1] == best["table"]
