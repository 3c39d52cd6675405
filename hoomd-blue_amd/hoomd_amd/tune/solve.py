"""Solvers that move a tunable x toward an optimum of a measured y, the roles of
hoomd/tune/solve.py and hoomd/tune/attr_tuner.py: the tunable definition, the grid optimizer and
the gradient-descent optimizer that the neighbour-list buffer tuner uses.

Each solver keeps one small search-state object per tunable (``_GridSearch``,
``_DescentState``); ``solve`` advances every tunable by one measurement and reports whether all
of them have settled.
"""

import math
from collections import deque

from ..variant import Constant, Variant


class ManualTuneDefinition:
    """y = f(x) for a solver (attr_tuner.py:150-230): get_y, get_x and set_x callables and an
    optional (lo, hi) domain, either end None for unbounded. Setting x clamps it into the
    domain. A y of None means that no measurement is available yet."""

    def __init__(self, get_y, target, get_x, set_x, domain=None):
        if domain is not None and len(domain) != 2:
            raise ValueError("domain must be a sequence of length two.")
        self._get_y = get_y
        self._get_x = get_x
        self._set_x = set_x
        self.target = target
        self.domain = None if domain is None else tuple(domain)

    def clamp_into_domain(self, value):
        if self.domain is None:
            return value
        lo, hi = self.domain
        if lo is not None:
            value = max(lo, value)
        if hi is not None:
            value = min(hi, value)
        return value

    @property
    def y(self):
        return self._get_y()

    @property
    def x(self):
        return self._get_x()

    @x.setter
    def x(self, value):
        self._set_x(self.clamp_into_domain(value))


class SolverStep:
    """A solver advanced one measurement per ``solve`` call. Tunables without a measurement
    (y is None) are left alone and count as unsettled."""

    def solve_one(self, tunable):
        raise NotImplementedError

    def reset(self):
        raise NotImplementedError

    def solve(self, tunables):
        settled = True
        for t in tunables:
            if t.y is None:
                settled = False
            elif not self.solve_one(t):
                settled = False
        return settled


class Optimizer(SolverStep):
    """A solver that seeks an extremum of y rather than a target value."""


class _GridSearch:
    """Search state of one tunable: the interval being binned, the round, and the y measured
    at the centre of each bin visited so far in this round."""

    __slots__ = ("lo", "hi", "round", "ys", "done")

    def __init__(self, lo, hi):
        self.lo, self.hi = float(lo), float(hi)
        self.round = 1
        self.ys = []
        self.done = False

    def centre(self, n_bins, k):
        return self.lo + (k + 0.5) * (self.hi - self.lo) / n_bins

    def edges(self, n_bins, k):
        w = (self.hi - self.lo) / n_bins
        return self.lo + k * w, self.lo + (k + 1) * w


class GridOptimizer(Optimizer):
    """GridOptimizer(n_bins=5, n_rounds=1, maximize=True), the search of solve.py:511-614: the
    domain is cut into n_bins equal bins and the centre of each is measured in turn; the bin
    with the best y (the first of equals) becomes the next round's domain, and after n_rounds
    x stays at the centre of the last best bin."""

    def __init__(self, n_bins=5, n_rounds=1, maximize=True):
        self._n_bins = int(n_bins)
        self._n_rounds = int(n_rounds)
        self._maximize = bool(maximize)
        self.reset()

    def reset(self):
        self._search = {}

    def is_searching(self, tunable):
        return tunable in self._search

    def final_bin(self, tunable):
        """(lo, hi) of the bin the search settled in, or None while it runs."""
        s = self._search.get(tunable)
        return (s.lo, s.hi) if s is not None and s.done else None

    def _best(self, ys):
        best = 0
        for k, y in enumerate(ys):
            if (y > ys[best]) if self._maximize else (y < ys[best]):
                best = k
        return best

    def solve_one(self, tunable):
        s = self._search.get(tunable)
        if s is None:
            lo, hi = tunable.domain if tunable.domain is not None else (None, None)
            if lo is None or hi is None:
                raise ValueError("GridOptimizer requires max and min x value to tune.")
            s = self._search[tunable] = _GridSearch(lo, hi)
            tunable.x = s.centre(self._n_bins, 0)
            return False
        if s.done:
            return True
        s.ys.append(tunable.y)
        if len(s.ys) < self._n_bins:
            tunable.x = s.centre(self._n_bins, len(s.ys))
            return False
        s.lo, s.hi = s.edges(self._n_bins, self._best(s.ys))
        s.ys = []
        if s.round >= self._n_rounds:
            s.done = True
            tunable.x = (s.lo + s.hi) / 2
            return True
        s.round += 1
        tunable.x = s.centre(self._n_bins, 0)
        return False


class _DescentState:
    """Search state of one tunable: the previous (x, y), the recent gradients (newest first)
    and how many measurements in a row left y unchanged."""

    __slots__ = ("x", "y", "grads", "still", "done")

    def __init__(self, x, y, depth):
        self.x, self.y = x, y
        self.grads = deque([0.0] * depth, maxlen=depth) if depth else None
        self.still = 0
        self.done = False


class GradientDescent(Optimizer):
    """GradientDescent(alpha=0.1, kappa=None, tol=1e-5, maximize=True, max_delta=None), the
    optimizer of solve.py:320-480: a forward-difference gradient from the last two
    measurements and the step x <- x - alpha * ((1 - sum(kappa)) g + sum_i kappa_i g_i), where
    g_i are the previous gradients (newest first), negated when maximizing; |step| is capped at
    max_delta. alpha may be a variant of the number of solve calls. The first measurement
    perturbs x by 10 % to get a difference; a step below tol settles the tunable. When y stays
    unchanged for more than three measurements in a row, x is nudged on by half its last
    change."""

    _STILL_LIMIT = 3

    def __init__(self, alpha=0.1, kappa=None, tol=1e-5, maximize=True, max_delta=None):
        self.alpha = alpha
        self.kappa = None if kappa is None else [float(k) for k in kappa]
        self.tol = float(tol)
        self.maximize = bool(maximize)
        self.max_delta = max_delta
        self.reset()

    @property
    def alpha(self):
        return self._alpha(self._calls)

    @alpha.setter
    def alpha(self, value):
        if isinstance(value, Variant):
            self._alpha = value
        elif isinstance(value, (int, float)):
            self._alpha = Constant(float(value))
        else:
            raise TypeError("Expected either a hoomd.variant.variant_like object.")

    def reset(self):
        self._calls = 0
        self._state = {}

    def solve(self, tunables):
        self._calls += 1
        return super().solve(tunables)

    def _step(self, s, grad):
        if self.kappa is None:
            d = grad
        else:
            d = (1.0 - sum(self.kappa)) * grad + sum(k * g for k, g in zip(self.kappa, s.grads))
        d *= self.alpha
        if self.max_delta is not None and abs(d) > self.max_delta:
            d = math.copysign(self.max_delta, d)
        return d

    def solve_one(self, tunable):
        x, y = tunable.x, tunable.y
        s = self._state.get(tunable)
        if s is None:
            self._state[tunable] = _DescentState(x, y, len(self.kappa or ()))
            for factor in (1.1, 0.9):
                nx = tunable.clamp_into_domain(x * factor)
                if nx != x:
                    tunable.x = nx
                    return False
            raise RuntimeError("Unable to perturb x for secant solver.")
        if s.done:
            return True
        if y == s.y:
            s.still += 1
            if s.still > self._STILL_LIMIT:
                s.still = 0
                tunable.x = x + 0.5 * (x - s.x)
            return False
        s.still = 0
        grad = (y - s.y) / (x - s.x)
        if self.maximize:
            grad = -grad
        d = self._step(s, grad)
        if abs(d) < self.tol:
            s.done = True
            return True
        nx = tunable.clamp_into_domain(x - d)
        if not math.isclose(nx, x):
            tunable.x = nx
            s.x, s.y = x, y
            if s.grads is not None:
                s.grads.appendleft(grad)
        return False
