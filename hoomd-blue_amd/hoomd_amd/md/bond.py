"""Bond forces (hoomd/md/bond.py). ``FENEWCA`` drives the EvaluatorBondFENE kernel and
``Harmonic`` the EvaluatorBondHarmonic kernel."""

import numpy as np

from .. import _lib
from .force import Force
from .typeparam import TypeParameter


class FENEWCA(Force):
    """hoomd.md.bond.FENEWCA(): params[bond_type] = dict(k, r0, epsilon, sigma, delta)
    (EvaluatorBondFENE.h:30-156, bond.py:108-163)."""

    _log_namespace = ("md", "bond", "FENEWCA")

    def __init__(self):
        super().__init__()
        self.params = TypeParameter()

    def _signature(self):
        return ("FENEWCA", tuple(sorted((k, tuple(sorted(v.items()))) for k, v in self.params.items())))

    def _attach(self, sim, h):
        snap = sim.state._snapshot
        b = snap.bonds
        types = list(b.types) if b.types else ["A-A"]
        P = (_lib.HmFeneParams * len(types))()
        for q, t in enumerate(types):
            p = self.params[t]
            P[q].k, P[q].r0 = float(p["k"]), float(p["r0"])
            P[q].epsilon, P[q].sigma = float(p["epsilon"]), float(p["sigma"])
            P[q].delta = float(p.get("delta", 0.0))
        g = np.ascontiguousarray(b.group, np.uint32).reshape(-1)
        tid = np.ascontiguousarray(b.typeid, np.uint32)
        fid = _lib.lib().hm_sys_add_bond_fene(h, b.N, g.ctypes.data, tid.ctypes.data, len(types), P)
        _lib.check(fid, "bond.FENEWCA")
        self._force_id = fid
        self._simulation = sim


class Harmonic(Force):
    """hoomd.md.bond.Harmonic(): params[bond_type] = dict(k, r0), U = k (r - r0)^2 / 2
    (EvaluatorBondHarmonic.h:37-140, bond.py:57-106)."""

    _log_namespace = ("md", "bond", "Harmonic")

    def __init__(self):
        super().__init__()
        self.params = TypeParameter()

    def _signature(self):
        return ("Harmonic",
                tuple(sorted((k, tuple(sorted(v.items()))) for k, v in self.params.items())))

    def _attach(self, sim, h):
        b = sim.state._snapshot.bonds
        types = list(b.types) if b.types else ["A-A"]
        P = (_lib.HmHarmonicParams * len(types))()
        for q, t in enumerate(types):
            p = self.params[t]
            P[q].k, P[q].r0 = float(p["k"]), float(p["r0"])
        g = np.ascontiguousarray(b.group, np.uint32).reshape(-1)
        tid = np.ascontiguousarray(b.typeid, np.uint32)
        fid = _lib.lib().hm_sys_add_bond_harmonic(h, b.N, g.ctypes.data, tid.ctypes.data,
                                                  len(types), P)
        _lib.check(fid, "bond.Harmonic")
        self._force_id = fid
        self._simulation = sim
