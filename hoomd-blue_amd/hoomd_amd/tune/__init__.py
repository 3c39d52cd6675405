"""hoomd.tune subset: ParticleSorter (hoomd/tune/sorter.py), GridOptimizer, GradientDescent and
ManualTuneDefinition (hoomd/tune/solve.py, attr_tuner.py)."""

import math

from ..trigger import Periodic  # noqa: F401
from .solve import (GradientDescent, GridOptimizer, ManualTuneDefinition, Optimizer,  # noqa: F401
                    SolverStep)


class ParticleSorter:
    """Order particles in memory along a Hilbert curve every ``trigger`` steps
    (SFCPackTuner, hoomd/SFCPackTuner.cc). grid defaults to 256 in 3D and rounds up to a power
    of two (sorter.py:44-56)."""

    def __init__(self, trigger=200, grid=None):
        self.trigger = trigger
        self.grid = grid

    @property
    def grid(self):
        return self._grid_v

    @grid.setter
    def grid(self, value):
        if value is None:
            self._grid_v = None
            return
        value = int(value)
        if value <= 0:
            raise ValueError("grid must be a natural number")
        self._grid_v = int(2.0 ** math.ceil(math.log2(value)))

    def _period(self):
        t = self.trigger
        period = getattr(t, "period", t)
        if getattr(t, "phase", 0):
            raise NotImplementedError("ParticleSorter: only Periodic triggers with phase 0")
        return int(period)

    def _grid(self):
        return 256 if self._grid_v is None else self._grid_v

    def _signature(self):
        return (self._period(), self._grid())

    def __call__(self, timestep):
        return False  # the engine sorts natively on its own period
