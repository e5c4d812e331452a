"""The sweep's precision choice (aphi_amd.sweep.PrecisionPolicy), host logic only."""
import math

import aphi_amd as A


def rep(steps, used, factor_ms, solve_ms):
    return {"refinement_steps": steps, "precision_used": used, "factorize_ms": factor_ms, "solve_ms": solve_ms}


def test_starts_with_fp32_and_learns_costs():
    p = A.PrecisionPolicy()
    assert p.choose(10.0) == A.APHI_FP32
    p.record(10.0, A.APHI_FP32, rep(3, A.APHI_FP32, 800.0, 4 * 25.0))
    assert p.factor32 == 800.0 and p.solve32 == 25.0
    # a neighbour that needed 3 solves: FP32 (800 + 4*25) beats the FP64 prior (1600 + 25)
    assert p.choose(20.0) == A.APHI_FP32


def test_fallback_marks_neighbourhood_fp64_and_bounds_it():
    p = A.PrecisionPolicy()
    p.record(1e3, A.APHI_FP32, rep(3, A.APHI_FP32, 800.0, 100.0))
    p.record(1e7, A.APHI_FP32, rep(30, A.APHI_FP64, 2500.0, 900.0))
    assert math.isinf(p.points[7.0])
    assert p.choose(5e7) == A.APHI_FP64       # beyond the stalled point
    assert p.choose(3e6) == A.APHI_FP64       # between an easy and a stalled point: the worse neighbour
    assert p.choose(100.0) == A.APHI_FP32     # next to easy points only
    # an FP64 point records timings, not an FP32 solve count
    p.record(5e7, A.APHI_FP64, rep(0, A.APHI_FP64, 1700.0, 36.0))
    assert p.factor64 == 1700.0 and 7.69 not in p.points


def test_cost_model_threshold():
    p = A.PrecisionPolicy()
    p.record(1e3, A.APHI_FP32, rep(3, A.APHI_FP32, 800.0, 100.0))        # 25 ms per solve
    p.record(1e5, A.APHI_FP64, rep(0, A.APHI_FP64, 1650.0, 35.0))
    p.record(1e4, A.APHI_FP32, rep(12, A.APHI_FP32, 800.0, 13 * 25.0))
    # between them: about 5.7 solves, far from the budget: the plain cost decides,
    # 800 + 6.7*25 = 968 <= 1650 + 35
    assert p.choose(3e3) == A.APHI_FP32
    p.record(1e6, A.APHI_FP32, rep(30, A.APHI_FP32, 800.0, 31 * 25.0))   # slow but converged
    # just above it the rising trend predicts more than the budget: 800 + 32.5*25 =
    # 1612 would beat 1685 if it converged, but failure is then certain and costs
    # 800 + 15*25 + 1685 (half the budget in solves), so FP64
    assert p.choose(1.2e6) == A.APHI_FP64
    p.record(2e6, A.APHI_FP32, rep(36, A.APHI_FP32, 800.0, 37 * 25.0))
    assert p.choose(3e6) == A.APHI_FP64


def test_failure_risk_near_the_budget():
    """The FP32 cost is weighted by the likelihood of a failed attempt,
    ((expected solves - budget/2) / (budget/2))^2: none up to half the budget,
    certain at the budget; a failure costs half the budget in solves."""
    p = A.PrecisionPolicy(budget=48)
    assert p.stall_solves == 24.0
    p.record(1e2, A.APHI_FP32, rep(2, A.APHI_FP32, 560.0, 3 * 16.6))
    p.record(1e5, A.APHI_FP64, rep(0, A.APHI_FP64, 1330.0, 30.0))
    fp64 = 1330.0 + 30.0
    failed = 560.0 + 24.0 * 16.6 + fp64
    ok = lambda k: 560.0 + (1.0 + k) * 16.6  # noqa: E731
    assert abs(p.fp32_cost(0.0, fp64) - ok(0.0)) < 1e-9                    # nothing to refine: no risk
    assert abs(p.fp32_cost(24.0, fp64) - ok(24.0)) < 1e-9                  # half the budget: still no risk
    r = (12.0 / 24.0) ** 2
    assert abs(p.fp32_cost(36.0, fp64) - ((1 - r) * ok(36.0) + r * failed)) < 1e-9
    assert abs(p.fp32_cost(48.0, fp64) - failed) < 1e-9                    # certain failure
    assert abs(p.fp32_cost(60.0, fp64) - failed) < 1e-9
    # the measured counts around the breakdown of the 32^3 sweep: 26 and 33 solves
    # converged (12.9 and 17.3 MHz), the next point needed about the budget
    assert p.fp32_cost(26.0, fp64) < fp64
    assert p.fp32_cost(31.9, fp64) < fp64                                   # predicted from 21 and 26
    assert p.fp32_cost(41.8, fp64) > fp64                                   # predicted from 26 and 33
    assert ok(41.8) < fp64                                                  # cheaper only if it converged


def test_interpolates_between_an_easy_and_a_failed_neighbour():
    p = A.PrecisionPolicy(budget=30)
    p.record(10 ** 5.97, A.APHI_FP32, rep(7, A.APHI_FP32, 670.0, 8 * 23.0))     # 23 ms per solve
    p.record(10 ** 7.24, A.APHI_FP32, rep(6, A.APHI_FP64, 2100.0, 600.0))      # stalled, fell back
    p.record(10 ** 7.62, A.APHI_FP64, rep(0, A.APHI_FP64, 1415.0, 30.0))
    # log(1 + k) interpolated: 7 solves at 5.97, 60 (twice the budget) at 7.24
    k = p.expected_solves(10 ** 6.35)
    lo, hi = math.log1p(7), math.log1p(60)
    t = (6.35 - 5.97) / (7.24 - 5.97)
    assert abs(k - math.expm1(lo + t * (hi - lo))) < 1e-9
    assert 7 < k < 60
    assert p.choose(10 ** 6.35) == A.APHI_FP32   # 670 + (1 + k) * 23 < 1415 + 30
    assert p.choose(10 ** 7.1) == A.APHI_FP64    # close to the failed point
    assert math.isinf(p.expected_solves(10 ** 7.5))  # only failed (or no FP32) neighbours
    assert p.choose(10 ** 7.5) == A.APHI_FP64


def test_extrapolates_a_rising_trend_in_an_ascending_sweep():
    p = A.PrecisionPolicy(budget=30)
    p.record(10 ** 5.0, A.APHI_FP32, rep(3, A.APHI_FP32, 610.0, 4 * 21.0))
    p.record(10 ** 6.0, A.APHI_FP32, rep(9, A.APHI_FP32, 610.0, 10 * 21.0))
    lo, hi = math.log1p(3), math.log1p(9)
    k = p.expected_solves(10 ** 6.5)
    assert abs(k - math.expm1(hi + 0.5 * (hi - lo))) < 1e-9 and k > 9
    assert p.choose(10 ** 6.5) == A.APHI_FP32              # 610 + (1 + k) * 21 < 2 * 610 + 21
    assert p.choose(10 ** 7.5) == A.APHI_FP64              # the trend passes the break-even point
    assert p.expected_solves(10 ** 9) == math.expm1(math.log1p(60))  # capped at a failure
    # a falling trend is not carried on: the nearest point's count stands
    q = A.PrecisionPolicy(budget=30)
    q.record(10 ** 5.0, A.APHI_FP32, rep(9, A.APHI_FP32, 610.0, 10 * 21.0))
    q.record(10 ** 6.0, A.APHI_FP32, rep(3, A.APHI_FP32, 610.0, 4 * 21.0))
    assert q.expected_solves(10 ** 6.5) == 3.0
