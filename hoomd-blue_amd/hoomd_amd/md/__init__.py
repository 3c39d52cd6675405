"""hoomd.md subset: Integrator, nlist.Cell, pair.LJ/Table, bond.FENEWCA/Harmonic, methods,
compute, and the learned coarse-grained site energy (mlcg)."""

from . import bond, compute, methods, mlcg, nlist, pair, tune  # noqa: F401
from .integrate import Integrator  # noqa: F401
