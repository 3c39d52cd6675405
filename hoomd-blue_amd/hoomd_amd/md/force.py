"""Force base (hoomd/md/force.py): per-particle outputs of one ForceCompute, by tag.

The arrays come from the engine's force buffers: fx, fy, fz and energy (FP32 accumulation) and
the six virial components (ForceCompute.h get*Python; energy sum in double, ForceCompute.cc:
213-233)."""

import numpy as np

from .. import _lib


class Force:
    # md/force.py Force loggables: the total energy by default, per-particle arrays on request
    _loggables = {"energy": ("scalar", True), "energies": ("particle", False),
                  "forces": ("particle", False), "virials": ("particle", False)}

    def __init__(self):
        self._simulation = None
        self._force_id = None

    def _fetch(self, virial=False):
        sim = self._simulation
        if sim is None or sim._system is None or self._force_id is None:
            return None, None
        N = sim.state.N_particles
        f4 = np.zeros((N, 4))
        v6 = np.zeros((6, N)) if virial else None
        _lib.check(_lib.lib().hm_sys_get_forces(sim._system, self._force_id, f4.ctypes.data,
                                                v6.ctypes.data if virial else None), "forces")
        return f4, v6

    @property
    def forces(self):
        f4, _ = self._fetch()
        return None if f4 is None else f4[:, :3].copy()

    @property
    def energies(self):
        f4, _ = self._fetch()
        return None if f4 is None else f4[:, 3].copy()

    @property
    def energy(self):
        f4, _ = self._fetch()
        return None if f4 is None else float(np.sum(f4[:, 3], dtype=np.float64))

    @property
    def virials(self):
        _, v6 = self._fetch(virial=True)
        return None if v6 is None else v6.T.copy()
