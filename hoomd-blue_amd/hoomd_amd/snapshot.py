"""Snapshot (hoomd/snapshot.py): host-side system description, arrays by tag."""

import numpy as np


class _Configuration:
    def __init__(self):
        self.box = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
        self.dimensions = 3


class _Particles:
    def __init__(self):
        self._N = 0
        self.types = ["A"]
        self._alloc(0)

    def _alloc(self, N):
        self.position = np.zeros((N, 3))
        self.velocity = np.zeros((N, 3))
        self.acceleration = np.zeros((N, 3))
        self.typeid = np.zeros(N, np.uint32)
        self.mass = np.ones(N)
        self.image = np.zeros((N, 3), np.int32)

    @property
    def N(self):
        return self._N

    @N.setter
    def N(self, n):
        self._N = int(n)
        self._alloc(self._N)


class _Bonds:
    def __init__(self):
        self._N = 0
        self.types = []
        self.typeid = np.zeros(0, np.uint32)
        self.group = np.zeros((0, 2), np.uint32)

    @property
    def N(self):
        return self._N

    @N.setter
    def N(self, n):
        self._N = int(n)
        self.typeid = np.zeros(self._N, np.uint32)
        self.group = np.zeros((self._N, 2), np.uint32)


class _Communicator:
    rank = 0
    num_ranks = 1


class Snapshot:
    """hoomd.Snapshot: configuration (box), particles and bonds, all by tag."""

    def __init__(self, communicator=None):
        self.configuration = _Configuration()
        self.particles = _Particles()
        self.bonds = _Bonds()
        self.communicator = _Communicator()

    def replicate(self, nx, ny, nz):
        """Snapshot.replicate: tile the box nx x ny x nz (positions wrapped into the new box)."""
        Lx, Ly, Lz = self.configuration.box[:3]
        p = self.particles
        N = p.N
        shifts = np.array([[i, j, k] for i in range(nx) for j in range(ny) for k in range(nz)],
                          float) * [Lx, Ly, Lz]
        pos = (p.position[None, :, :] + shifts[:, None, :]).reshape(-1, 3)
        newL = np.array([nx * Lx, ny * Ly, nz * Lz])
        pos = pos - (newL - [Lx, Ly, Lz]) / 2
        pos = (pos + newL / 2) % newL - newL / 2
        reps = len(shifts)
        vel = np.tile(p.velocity, (reps, 1))
        typeid = np.tile(p.typeid, reps)
        mass = np.tile(p.mass, reps)
        groups = [self.bonds.group + r * N for r in range(reps)]
        btype = np.tile(self.bonds.typeid, reps)
        types = list(p.types)
        p.N = N * reps
        p.types = types
        p.position[:] = pos
        p.velocity[:] = vel
        p.typeid[:] = typeid
        p.mass[:] = mass
        nb = self.bonds.N
        if nb:
            bt = list(self.bonds.types)
            self.bonds.N = nb * reps
            self.bonds.types = bt
            self.bonds.group[:] = np.concatenate(groups)
            self.bonds.typeid[:] = btype
        self.configuration.box = [float(newL[0]), float(newL[1]), float(newL[2]), 0, 0, 0]
        return self
