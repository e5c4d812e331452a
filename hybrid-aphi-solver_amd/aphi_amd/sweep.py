"""Precision choice across the points of a frequency sweep.

The FP32 factor plus GMRES refinement reaches the backward-error target in a few
solves at most frequencies, but where the equilibrated matrix is badly
conditioned (here: high frequencies, skin depth far below the mesh size) the
FP32 factor is too inaccurate a preconditioner and refinement needs tens of
solves, or stalls and falls back to an FP64 factor. The reference's sweep
(tools/solve_mesh.cpp:152-259) factors every point in complex128; this policy
keeps that as the choice wherever the points already solved nearby show the
mixed-precision path would be slower, and uses FP32 everywhere else.
"""
import math

from ._capi import APHI_FP32, APHI_FP64


class PrecisionPolicy:
    """Chooses APHI_FP32 or APHI_FP64 for a frequency from the points solved so far.

    Cost model per point: FP32 = factor32 + (1 + solves) * solve32, FP64 =
    factor64 + solve64, with the times taken from the solve reports as they come
    in, each the smallest seen so far (a report taken while other work shared the
    device, or one that includes a first call's graph capture, overstates it;
    factor64 starts as `fp64_factor_ratio` times factor32, the measured ratio of
    the two MFMA paths on the 32^3 cube). The refinement solves expected at a new frequency are interpolated
    between its nearest solved neighbours below and above, linearly in
    log-frequency on log(1 + solves); with solved points on one side only (a
    sweep in frequency order) the line through the two nearest ones is carried
    on when it rises towards the new frequency. A point where FP32 refinement did
    not converge counts as twice the refinement budget, and a frequency whose
    nearest neighbours all failed is left to FP64.

    An FP32 attempt whose expected refinement runs towards the budget may fail:
    refinement runs out of budget or stalls, and the point is refactored in FP64.
    The FP32 cost is therefore an expectation: with the failure likelihood taken
    as ((expected solves - budget/2) / (budget/2))^2, zero up to half the budget
    and rising to 1 at the budget, a failure costing the FP32 factor,
    ``stall_solves`` solves (by default half the budget) and the FP64 point on top.
    On the 32^3 cube, FP32 attempts without a stall rule converged after 21, 26
    and 33 solves at 9.6, 12.9 and 17.3 MHz, reached 3e-12 after 40 at 23.2 MHz and
    did not converge at 41.6 MHz; the refinement's stall rule judges the rate
    late, where GMRES has sped up.
    """

    def __init__(self, fp64_factor_ratio=2.3, budget=30, stall_solves=None):
        self.points = {}          # log10(f) -> refinement solves (inf: FP32 not viable)
        self.factor32 = None
        self.factor64 = None
        self.solve32 = None
        self.solve64 = None
        self.ratio = fp64_factor_ratio
        self.budget = budget
        self.stall_solves = 0.5 * budget if stall_solves is None else stall_solves

    def expected_solves(self, freq):
        if not self.points:
            return 0.0
        x = math.log10(freq)
        below = [k for k in self.points if k <= x]
        above = [k for k in self.points if k >= x]
        near = []
        if below:
            near.append(max(below))
        if above:
            near.append(min(above))
        vals = [self.points[k] for k in near]
        if all(math.isinf(v) for v in vals):
            return math.inf

        def level(v):
            return math.log1p(2.0 * self.budget if math.isinf(v) else v)

        if len(near) == 1 or near[0] == near[-1]:
            # one side only (a sweep in frequency order): carry the trend of the
            # two nearest points on that side, when it rises towards x
            side = sorted(below if below else above, key=lambda k: abs(k - x))
            v0 = self.points[side[0]]
            if len(side) >= 2 and not math.isinf(v0) and not math.isinf(self.points[side[1]]):
                x0, x1 = side[0], side[1]
                slope = (level(v0) - level(self.points[x1])) / (x0 - x1)
                if slope * (x - x0) > 0:
                    return math.expm1(min(level(v0) + slope * (x - x0), level(math.inf)))
            return math.expm1(level(v0))
        t = (x - near[0]) / (near[1] - near[0])
        return math.expm1((1.0 - t) * level(vals[0]) + t * level(vals[1]))

    def choose(self, freq):
        k = self.expected_solves(freq)
        if self.factor32 is None or self.solve32 is None:
            return APHI_FP32
        if math.isinf(k):
            return APHI_FP64
        f64 = self.factor64 if self.factor64 is not None else self.ratio * self.factor32
        s64 = self.solve64 if self.solve64 is not None else self.solve32
        return APHI_FP32 if self.fp32_cost(k, f64 + s64) <= f64 + s64 else APHI_FP64

    def fp32_cost(self, k, fp64_point):
        """Expected cost of an FP32 attempt expecting k refinement solves."""
        ok = self.factor32 + (1.0 + k) * self.solve32
        half = 0.5 * self.budget
        risk = min(1.0, max(0.0, (k - half) / half)) ** 2
        failed = self.factor32 + self.stall_solves * self.solve32 + fp64_point
        return (1.0 - risk) * ok + risk * failed

    def record(self, freq, precision, report):
        """Feeds back one solve_symmetric report (precision = the one requested)."""
        x = math.log10(freq)
        steps = report["refinement_steps"]
        solve_each = report["solve_ms"] / max(1, steps + 1)
        low = lambda old, new: new if old is None else min(old, new)  # noqa: E731
        if precision == APHI_FP32:
            if report["precision_used"] == APHI_FP32:
                self.points[x] = float(steps)
                self.factor32 = low(self.factor32, report["factorize_ms"])
                self.solve32 = low(self.solve32, solve_each)
            else:
                self.points[x] = math.inf
        else:
            # the FP64 timings only: what FP32 would have needed here stays unknown
            self.factor64 = low(self.factor64, report["factorize_ms"])
            self.solve64 = low(self.solve64, solve_each)
