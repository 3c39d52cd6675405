"""hoomd.md subset: Integrator, nlist.Cell, pair.LJ, bond.FENEWCA, methods, compute."""

from . import bond, compute, methods, nlist, pair, tune  # noqa: F401
from .integrate import Integrator  # noqa: F401
