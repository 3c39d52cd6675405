"""Pair forces (hoomd/md/pair/pair.py). ``LJ`` and ``Table`` drive the pair kernels with
the EvaluatorPairLJ and EvaluatorPairTable evaluators."""

import numpy as np

from .. import _lib
from .force import Force
from .typeparam import TypePairParameter


class Pair(Force):
    """Base of pair potentials: nlist, r_cut, r_on per type pair, energy shift mode
    (pair.py:65-146). The evaluator runs in FP32 on the GPU with a full neighbour list."""

    _modes = {"none": 0, "shift": 1, "xplor": 2}

    def __init__(self, nlist, default_r_cut=None, default_r_on=0.0, mode="none"):
        super().__init__()
        if mode not in self._modes:
            raise ValueError(f"mode must be one of {list(self._modes)}")
        self.nlist = nlist
        self.mode = mode
        self.r_cut = TypePairParameter(default_r_cut)
        self.r_on = TypePairParameter(default_r_on)

    def _matrices(self, types):
        nt = len(types)
        rc = np.zeros((nt, nt))
        ro = np.zeros((nt, nt))
        for i, a in enumerate(types):
            for j, b in enumerate(types):
                r = self.r_cut.get(a, b)
                if r is None:
                    raise RuntimeError(f"r_cut for ({a}, {b}) is not set")
                rc[i, j] = r
                ro[i, j] = self.r_on.get(a, b, 0.0)
        return rc, ro


class LJ(Pair):
    """hoomd.md.pair.LJ(nlist, default_r_cut=None, default_r_on=0., mode='none'):
    params[(a, b)] = dict(epsilon=..., sigma=...) (EvaluatorPairLJ.h:99-155)."""

    _log_namespace = ("md", "pair", "LJ")

    def __init__(self, nlist, default_r_cut=None, default_r_on=0.0, mode="none",
                 tail_correction=False):
        super().__init__(nlist, default_r_cut, default_r_on, mode)
        if tail_correction:
            raise NotImplementedError("tail_correction is not supported")
        self.params = TypePairParameter()

    def _signature(self):
        return ("LJ", self.mode, self.params.signature(), self.r_cut.signature(),
                self.r_on.signature(), self.nlist._signature())

    def _attach(self, sim, h):
        types = sim.state.particle_types
        nt = len(types)
        rc, ro = self._matrices(types)
        P = (_lib.HmLJParams * (nt * nt))()
        for i, a in enumerate(types):
            for j, b in enumerate(types):
                p = self.params.get(a, b)
                if p is None:
                    raise RuntimeError(f"params for ({a}, {b}) are not set")
                P[i * nt + j].sigma = float(p["sigma"])
                P[i * nt + j].epsilon = float(p["epsilon"])
        rcf = np.ascontiguousarray(rc.ravel())
        rof = np.ascontiguousarray(ro.ravel())
        fid = _lib.lib().hm_sys_add_pair_lj(h, rcf.ctypes.data, rof.ctypes.data, P,
                                            self._modes[self.mode])
        _lib.check(fid, "pair.LJ")
        self._force_id = fid
        self._simulation = sim
        self.nlist._simulation = sim


class Table(Pair):
    """hoomd.md.pair.Table(nlist, default_r_cut=None) (pair.py:520-618): params[(a, b)] =
    dict(r_min=..., U=array, F=array), U and F (= -dU/dr) on the points
    numpy.linspace(r_min, r_cut, len(U), endpoint=False), interpolated linearly
    (EvaluatorPairTable.h:152-196). Mode 'none' only."""

    _modes = {"none": 0}
    _log_namespace = ("md", "pair", "Table")

    def __init__(self, nlist, default_r_cut=None):
        super().__init__(nlist, default_r_cut, 0.0, "none")
        self.params = TypePairParameter()

    @staticmethod
    def _table(p):
        missing = {"r_min", "U", "F"} - set(p)
        if missing:
            raise ValueError(f"Table params need {sorted(missing)}")
        U = np.asarray(p["U"], np.float64).ravel()
        F = np.asarray(p["F"], np.float64).ravel()
        if U.size != F.size:
            raise RuntimeError("The length of V and F arrays must be equal.")
        if U.size == 0:
            raise RuntimeError("The length of V and F must not be zero.")
        return float(p["r_min"]), U, F

    def _signature(self):
        # the layout (r_min and widths) only: new U and F values of an attached table are
        # uploaded in place by _sync, so a ramped potential needs no new system
        tabs = []
        for k, v in sorted(self.params._d.items()):
            r_min, U, F = self._table(v)
            tabs.append((k, r_min, U.size))
        return ("Table", tuple(tabs), self.r_cut.signature(), self.nlist._signature())

    def _values(self, types):
        Us, Fs = [], []
        for a in types:
            for b in types:
                _, U, F = self._table(self.params.get(a, b))
                Us.append(U)
                Fs.append(F)
        return (np.ascontiguousarray(np.concatenate(Us)),
                np.ascontiguousarray(np.concatenate(Fs)))

    def _sync(self, sim):
        U, F = self._values(sim.state.particle_types)
        if getattr(self, "_uploaded", None) is not None and \
                np.array_equal(self._uploaded[0], U) and np.array_equal(self._uploaded[1], F):
            return
        _lib.check(_lib.lib().hm_sys_update_pair_table(sim._system, self._force_id,
                                                       U.ctypes.data, F.ctypes.data),
                   "pair.Table values")
        self._uploaded = (U, F)

    def _attach(self, sim, h):
        types = sim.state.particle_types
        nt = len(types)
        rc, _ = self._matrices(types)
        r_min = np.zeros(nt * nt)
        width = np.zeros(nt * nt, np.uint32)
        Us, Fs = [], []
        for i, a in enumerate(types):
            for j, b in enumerate(types):
                p = self.params.get(a, b)
                if p is None:
                    raise RuntimeError(f"params for ({a}, {b}) are not set")
                rm, U, F = self._table(p)
                if rc[i, j] > 0 and not rm < rc[i, j]:
                    raise ValueError(f"r_min must be below r_cut for ({a}, {b})")
                r_min[i * nt + j] = rm
                width[i * nt + j] = U.size
                Us.append(U)
                Fs.append(F)
        U = np.ascontiguousarray(np.concatenate(Us))
        F = np.ascontiguousarray(np.concatenate(Fs))
        rcf = np.ascontiguousarray(rc.ravel())
        fid = _lib.lib().hm_sys_add_pair_table(h, rcf.ctypes.data, r_min.ctypes.data,
                                               width.ctypes.data, U.ctypes.data, F.ctypes.data)
        _lib.check(fid, "pair.Table")
        self._uploaded = (U, F)
        self._force_id = fid
        self._simulation = sim
        self.nlist._simulation = sim
