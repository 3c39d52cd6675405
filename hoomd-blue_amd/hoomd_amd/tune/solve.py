"""Solvers for tuning x toward an optimum of a measured y (hoomd/tune/solve.py,
hoomd/tune/attr_tuner.py): the tunable definition and the grid optimizer the neighbour-list
buffer tuner uses."""

from collections import defaultdict

import numpy as np


class ManualTuneDefinition:
    """y = f(x) for a solver (attr_tuner.py:150-230): get_y, get_x and set_x callables; x is
    clamped into the domain when set. A y of None means no measurement is available."""

    def __init__(self, get_y, target, get_x, set_x, domain=None):
        if domain is not None and len(domain) != 2:
            raise ValueError("domain must be a sequence of length two.")
        self._get_y = get_y
        self._get_x = get_x
        self._set_x = set_x
        self.target = target
        self.domain = domain

    @property
    def y(self):
        return self._get_y()

    @property
    def x(self):
        return self._get_x()

    @x.setter
    def x(self, value):
        if self.domain is not None:
            lo, hi = self.domain
            if lo is not None and value < lo:
                value = lo
            if hi is not None and value > hi:
                value = hi
        self._set_x(value)


class SolverStep:
    """One iteration toward a solution per call (solve.py:20-80); tunables whose y is None are
    skipped and count as not solved."""

    def solve_one(self, tunable):
        raise NotImplementedError

    def reset(self):
        raise NotImplementedError

    def _solve_one_internal(self, tunable):
        if tunable.y is None:
            return False
        return self.solve_one(tunable)

    def solve(self, tunables):
        return all([self._solve_one_internal(t) for t in tunables])


class Optimizer(SolverStep):
    """A solver that seeks an extremum of y rather than a target value."""


class GridOptimizer(Optimizer):
    """GridOptimizer(n_bins=5, n_rounds=1, maximize=True) (solve.py:511-614): the domain is cut
    into n_bins bins whose centres are measured in turn; the best bin becomes the next round's
    domain, and after n_rounds the best bin's centre is the answer."""

    def __init__(self, n_bins=5, n_rounds=1, maximize=True):
        self._n_bins = int(n_bins)
        self._n_rounds = int(n_rounds)
        self._opt = max if maximize else min
        self.reset()

    def solve_one(self, tunable):
        if self._solved[tunable]:
            return True
        if tunable not in self._bins:
            lo, hi = tunable.domain if tunable.domain is not None else (None, None)
            if lo is None or hi is None:
                raise ValueError("GridOptimizer requires max and min x value to tune.")
            self._bins[tunable] = np.linspace(lo, hi, self._n_bins + 1)
            tunable.x = self._centre(tunable, 0)
            return False
        bin_y = self._bin_y[tunable]
        bin_y.append(tunable.y)
        if len(bin_y) == self._n_bins:
            index = bin_y.index(self._opt(bin_y))
            bounds = self._bins[tunable][index:index + 2]
            if self._round[tunable] == self._n_rounds:
                tunable.x = float(sum(bounds) / 2)
                self._final_bins[tunable] = tuple(float(b) for b in bounds)
                self._solved[tunable] = True
                return True
            self._bins[tunable] = np.linspace(*bounds, self._n_bins + 1)
            bin_y.clear()
            tunable.x = self._centre(tunable, 0)
            self._round[tunable] += 1
            return False
        tunable.x = self._centre(tunable, len(bin_y))
        return False

    def _centre(self, tunable, index):
        return float(sum(self._bins[tunable][index:index + 2]) / 2)

    def reset(self):
        self._bins = {}
        self._round = defaultdict(lambda: 1)
        self._bin_y = defaultdict(list)
        self._final_bins = {}
        self._solved = defaultdict(lambda: False)
