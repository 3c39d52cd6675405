"""hoomd.md.tune.NeighborListBuffer (hoomd/md/tune/nlist_buffer.py): tune the neighbour-list
buffer for the largest measured timesteps per second."""

import time

from ..trigger import as_trigger
from ..tune.solve import GradientDescent, GridOptimizer, ManualTuneDefinition, Optimizer
from ..variant import Ramp


class NeighborListBuffer:
    _log_namespace = ("md", "tune", "NeighborListBuffer")
    _loggables = {"max_tps": ("scalar", True), "best_buffer_size": ("scalar", True),
                  "last_tps": ("scalar", True)}

    """Acts on the steps its trigger selects, before the step runs (as tuners do in
    System::run, System.cc:131-137): the TPS since the previous act in the same run is the
    measurement for the current buffer (nlist_buffer.py:20-57, the first act of a run only
    starts the clock), the solver proposes the next buffer, and the engine's neighbour list
    takes it at once. Tuned once the solver reports solved twice in a row."""

    def __init__(self, trigger, nlist, solver, maximum_buffer, minimum_buffer=0.0):
        if not isinstance(solver, Optimizer):
            raise TypeError("solver must be a hoomd_amd.tune Optimizer")
        self.trigger = as_trigger(trigger)
        self.nlist = nlist
        self.solver = solver
        self.maximum_buffer = float(maximum_buffer)
        self.minimum_buffer = float(minimum_buffer)
        self._tuned = 0
        self._last_tps = 0.0
        self._max_tps = 0.0
        self._best_buffer_size = nlist.buffer
        self._clock = None  # (run id, timestep, wall time) of the previous act
        self._sim = None
        self._tunable = ManualTuneDefinition(
            get_y=lambda: self._y, target=0.0, get_x=lambda: self.nlist.buffer,
            set_x=self._set_buffer, domain=(self.minimum_buffer, self.maximum_buffer))
        self._y = None

    @classmethod
    def with_grid(cls, trigger, nlist, maximum_buffer, minimum_buffer=0.0, n_bins=5,
                  n_rounds=1):
        return cls(trigger, nlist, GridOptimizer(n_bins, n_rounds, True),
                   maximum_buffer=maximum_buffer, minimum_buffer=minimum_buffer)

    @classmethod
    def with_gradient_descent(cls, trigger, nlist, maximum_buffer, minimum_buffer=0.0,
                              alpha=Ramp(1e-5, 1e-6, 0, 30), kappa=(0.33, 0.165), tol=1e-5,
                              max_delta=None):
        """The tuner with a GradientDescent solver maximising TPS (nlist_buffer.py:242-305).
        The domain is (minimum_buffer, maximum_buffer); alpha may vary with the number of
        solver calls."""
        return cls(trigger, nlist, GradientDescent(alpha, kappa, tol, True, max_delta),
                   maximum_buffer=maximum_buffer, minimum_buffer=minimum_buffer)

    def _signature(self):
        return ("NeighborListBuffer", id(self))

    def _set_buffer(self, value):
        self.nlist.buffer = float(value)
        if self._sim is not None:
            self._sim._set_nlist_buffer(self.nlist)

    def _act(self, sim, timestep):
        self._sim = sim
        now = time.perf_counter()
        run_id = sim._run_id
        y = None
        if self._clock is not None and self._clock[0] == run_id and timestep > self._clock[1]:
            y = (timestep - self._clock[1]) / (now - self._clock[2])
        self._y = y
        if y is not None:
            self._last_tps = y
            if y > self._max_tps:
                self._max_tps = y
                self._best_buffer_size = self.nlist.buffer
        if not self.tuned:
            if self.solver.solve([self._tunable]):
                self._tuned += 1
            else:
                self._tuned = 0
        # the clock restarts after the buffer change, so the next interval measures it
        self._clock = (run_id, timestep, time.perf_counter())

    @property
    def tuned(self):
        return self._tuned > 1

    @property
    def max_tps(self):
        return self._max_tps

    @property
    def best_buffer_size(self):
        return self._best_buffer_size

    @property
    def last_tps(self):
        return self._last_tps

    @property
    def final_bin(self):
        if not isinstance(self.solver, GridOptimizer):
            return None
        return self.solver.final_bin(self._tunable)

    def reset(self):
        self._tuned = 0
        self.solver.reset()
