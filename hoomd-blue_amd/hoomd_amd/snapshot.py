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

    def replicate(self, nx, ny, nz=1):
        """Snapshot.replicate (SnapshotParticleData::replicate, ParticleData.cc:3742-3811): each
        particle is unwrapped with its image, and copy j (loops x, y, z, z fastest) gets tag
        j*N + i at fraction f/n + l/n of the new box, with a fresh image flag. Bond tags shift by
        j*N (BondedGroupData.cc:1366-1400)."""
        nx, ny, nz = int(nx), int(ny), int(nz)
        if min(nx, ny, nz) < 1:
            raise ValueError("replicate counts must be positive")
        L = np.asarray(self.configuration.box[:3], float)
        n = np.array([nx, ny, nz], float)
        newL = L * n
        p = self.particles
        N = p.N
        frac = (p.position + p.image * L + L / 2) / L
        lmn = np.array([[l, m, k] for l in range(nx) for m in range(ny) for k in range(nz)], float)
        f_new = (frac[None, :, :] / n + lmn[:, None, :] / n).reshape(-1, 3)
        q = -newL / 2 + f_new * newL
        g = (q + newL / 2) / newL - 0.5
        img = np.where(g >= 0, g + 0.5, g - 0.5).astype(np.int64)
        q = q - img * newL
        hi = q >= newL / 2
        q = np.where(hi, q - newL, q)
        img = img + hi
        lo = q < -newL / 2
        q = np.where(lo, q + newL, q)
        img = img - lo
        reps = len(lmn)
        vel = np.tile(p.velocity, (reps, 1))
        acc = np.tile(p.acceleration, (reps, 1))
        typeid = np.tile(p.typeid, reps)
        mass = np.tile(p.mass, reps)
        groups = [self.bonds.group + r * N for r in range(reps)]
        btype = np.tile(self.bonds.typeid, reps)
        types = list(p.types)
        p.N = N * reps
        p.types = types
        p.position[:] = q
        p.image[:] = img
        p.velocity[:] = vel
        p.acceleration[:] = acc
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
