"""State (hoomd/state.py): the system's particles, held on the GPU once a simulation runs."""

import copy

import numpy as np

from . import _lib
from .snapshot import Snapshot


class State:
    def __init__(self, simulation, snapshot):
        self._simulation = simulation
        box = snapshot.configuration.box
        if any(abs(t) > 0 for t in box[3:6]):
            raise ValueError("hoomd_amd supports orthorhombic boxes only (tilt factors must be 0)")
        self._snapshot = copy.deepcopy(snapshot)
        p = self._snapshot.particles
        L = np.asarray(box[:3], float)
        if p.N and ((p.position < -L / 2).any() or (p.position >= L / 2).any()):
            raise RuntimeError("Not all particles were found inside the given box")

    # -------------------------------------------------------------- properties
    @property
    def N_particles(self):
        return self._snapshot.particles.N

    @property
    def box(self):
        return list(self._snapshot.configuration.box)

    @property
    def particle_types(self):
        return list(self._snapshot.particles.types)

    @property
    def bond_types(self):
        return list(self._snapshot.bonds.types)

    # -------------------------------------------------------------- snapshots
    def _sync_from_device(self):
        sysh = self._simulation._system
        if sysh is None:
            return
        p = self._snapshot.particles
        N = p.N
        pos = np.zeros((N, 3))
        vel = np.zeros((N, 3))
        img = np.zeros((N, 3), np.int32)
        acc = np.zeros((N, 3))
        _lib.check(_lib.lib().hm_sys_get_state(sysh, pos.ctypes.data, vel.ctypes.data,
                                               img.ctypes.data, acc.ctypes.data), "get_state")
        p.position[:] = pos
        p.velocity[:] = vel
        p.image[:] = img
        p.acceleration[:] = acc

    def get_snapshot(self):
        self._sync_from_device()
        return copy.deepcopy(self._snapshot)

    def set_snapshot(self, snapshot):
        """Replace the system's particles (hoomd/state.py set_snapshot). Keeping N, the box and
        the types writes positions, velocities and images into the running system; any other
        change rebuilds the device system on next use, as initializeFromSnapshot does."""
        box = snapshot.configuration.box
        if any(abs(t) > 0 for t in box[3:6]):
            raise ValueError("hoomd_amd supports orthorhombic boxes only (tilt factors must be 0)")
        sim = self._simulation
        same = (snapshot.particles.N == self.N_particles
                and list(box[:3]) == list(self._snapshot.configuration.box[:3])
                and list(snapshot.particles.types) == self.particle_types
                and snapshot.bonds.N == self._snapshot.bonds.N
                and np.array_equal(snapshot.particles.typeid, self._snapshot.particles.typeid)
                and np.array_equal(snapshot.particles.mass, self._snapshot.particles.mass)
                and np.array_equal(snapshot.bonds.group, self._snapshot.bonds.group))
        if not same:
            sim._destroy_system()
            sim._invalidate()
        self._snapshot = copy.deepcopy(snapshot)
        sysh = sim._system
        if sysh is not None:
            p = self._snapshot.particles
            pos = np.ascontiguousarray(p.position, np.float64)
            vel = np.ascontiguousarray(p.velocity, np.float64)
            img = np.ascontiguousarray(p.image, np.int32)
            _lib.check(_lib.lib().hm_sys_set_state(sysh, pos.ctypes.data, vel.ctypes.data,
                                                   img.ctypes.data), "set_state")

    def replicate(self, nx, ny, nz=1):
        """State.replicate (hoomd/state.py:580-609): the snapshot replicated nx x ny x nz."""
        snap = self.get_snapshot()
        snap.replicate(nx, ny, nz)
        self.set_snapshot(snap)

    def thermalize_particle_momenta(self, filter, kT):
        """Draw Maxwell-Boltzmann velocities (ParticleGroup::thermalizeParticleMomenta) on the
        GPU with the simulation seed and current timestep, then zero the COM momentum."""
        self._simulation._ensure_system()
        _lib.check(_lib.lib().hm_sys_thermalize(self._simulation._system, float(kT)),
                   "thermalize")
        self._sync_from_device()


def empty_snapshot():
    return Snapshot()
