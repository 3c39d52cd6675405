"""hoomd_amd -- the short-range MD timestep of HOOMD-blue, MI355X-native.

A drop-in for the ``hoomd`` / ``hoomd.md`` surface used by short-range simulations on
``hoomd.device.GPU`` (hoomd/__init__.py, hoomd/md/__init__.py): the same class names, argument
meanings and per-particle outputs, backed by hand-written gfx950 kernels in libhoomd_amd.so.

    import hoomd_amd as hoomd
    sim = hoomd.Simulation(device=hoomd.device.GPU(), seed=1)
"""

from . import device, filter, logging, md, trigger, tune, variant, write  # noqa: F401
from .simulation import Simulation  # noqa: F401
from .snapshot import Snapshot  # noqa: F401
from .state import State  # noqa: F401

__version__ = "0.1.0"
