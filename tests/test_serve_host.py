"""CPU checks of the serving form's scheduling policy (no GPU)."""
import numpy as np

from moraine_amd.serve import PatchSpec, admit_order, footprint_bytes, plan_tick, stream_specs


def test_admission_is_fifo_and_bounded_by_budget():
    q = [(0, 40), (1, 30), (2, 10), (3, 5)]
    assert admit_order(q, used=0, budget=100, max_active=8, n_active=0) == [0, 1, 2, 3]
    # stops at the first request that does not fit, even if a later one would
    assert admit_order(q, used=50, budget=100, max_active=8, n_active=0) == [0]
    assert admit_order(q, used=70, budget=100, max_active=8, n_active=0) == []
    assert admit_order(q, used=0, budget=100, max_active=3, n_active=1) == [0, 1]


def test_tick_plan_respects_pixel_and_coherence_budgets_oldest_first():
    P = lambda n: 8 * (n * (n - 1) // 2)
    active = [(7, 60, 100), (8, 30, 1000), (9, 200, 50)]
    plan = plan_tick(active, px_budget=600, coh_budget=10 ** 12)
    assert plan == [(7, 100), (8, 500)]
    plan = plan_tick(active, px_budget=10 ** 9, coh_budget=100 * P(60) + 10 * P(30))
    assert plan == [(7, 100), (8, 10)]
    # finished requests are skipped; an exhausted budget ends the tick
    assert plan_tick([(1, 17, 0), (2, 17, 5)], 3, 10 ** 9) == [(2, 3)]
    assert plan_tick([(1, 200, 5)], 10, P(200) - 1) == []


def test_footprint_counts_stack_mask_and_outputs():
    s = PatchSpec(0, 60, 100, 200, 7, 10)
    px = 100 * 200
    assert footprint_bytes(s) >= px * 60 * 8 * 2 + px * 315
    assert s.win == 315 and s.npx == px


def test_stream_specs_follow_the_request_model():
    specs = stream_specs(200, rate=50.0, seed=1)
    arr = np.array([s.arrival for s in specs])
    assert np.all(np.diff(arr) > 0)
    assert 2.0 < arr[-1] < 6.0  # ~200 / 50 s
    for s in specs:
        assert s.n in (30, 60, 120, 200) and s.h == s.w and 128 <= s.h <= 1024 and s.h % 32 == 0
        assert 11 <= 2 * s.az + 1 <= 21 and 15 <= 2 * s.r + 1 <= 41
    assert len({s.seed for s in specs}) == 200
