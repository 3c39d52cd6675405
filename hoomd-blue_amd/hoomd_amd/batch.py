"""Batched independent replicas: many LJ systems of uneven size advanced together on one GPU.

The reference runs one system per ``Simulation``; a sweep over state points or seeds runs the
systems one after another, each too small to fill the device. ``ReplicaBatch`` keeps every
replica in one set of device arrays and runs each phase of the timestep (step one with the Bussi
factor, cell and list rebuilds where a replica's distance check asks for them, LJ forces, step
two) as one launch over all replicas, chunks of steps captured as hipGraphs.

Every replica is the system that ``Simulation`` + ``md.nlist.Cell(buffer)`` +
``md.pair.LJ(mode="none")`` + ``md.methods.ConstantVolume(thermostat=Bussi(kT, tau))`` would run
(hoomd/md/TwoStepConstantVolume.cc:7-260, hoomd/md/Thermostat.h:357-446), in a cubic periodic
box with its own kT and seed.
"""

import ctypes as C

import numpy as np

from . import _lib


def fcc_fill(N, L, rng):
    """N sites of the smallest FCC lattice (n^3 cells, 4 sites each) that holds N, chosen at
    random without replacement (kept in lattice order), in the box [-L/2, L/2)^3."""
    n = int(np.ceil((N / 4.0) ** (1.0 / 3.0) - 1e-9))
    while 4 * n ** 3 < N:
        n += 1
    a = L / n
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    g = np.arange(n)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    sites = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) + 0.25) * a - L / 2
    keep = np.sort(rng.choice(len(sites), size=N, replace=False))
    return sites[keep]


def seeded_replicas(nrep=256, seed=1, n_range=(500, 20000), rho_range=(0.70, 0.90),
                    kT_range=(0.8, 1.5)):
    """The batched-replica workload drawn from one seed (numpy default_rng(seed)):

        N_r   = round(exp(U[ln n_min, ln n_max]))   (log-uniform), all nrep drawn first
        rho_r = U[rho_min, rho_max]                 then all densities
        kT_r  = U[kT_min, kT_max]                   then all temperatures
        L_r   = (N_r / rho_r)^(1/3)
        positions: N_r random sites of the smallest FCC lattice holding N_r (fcc_fill)
        seed_r = (seed + r) mod 2^16 (the replica's thermostat / thermalize seed)

    Returns (n, L, kT, seeds, positions list)."""
    rng = np.random.default_rng(seed)
    n = np.rint(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), nrep))).astype(int)
    rho = rng.uniform(rho_range[0], rho_range[1], nrep)
    kT = rng.uniform(kT_range[0], kT_range[1], nrep)
    L = (n / rho) ** (1.0 / 3.0)
    seeds = (seed + np.arange(nrep)) % 65536
    pos = [fcc_fill(int(n[r]), float(L[r]), rng) for r in range(nrep)]
    return n, L, kT, seeds, pos


class ReplicaBatch:
    """ReplicaBatch(positions, L, kT, seeds, velocities=None, epsilon=1, sigma=1, r_cut=2.5,
    buffer=0.4, dt=0.005, tau=0.0)

    positions: list of [N_r, 3] arrays; L, kT, seeds: one per replica. velocities: list of
    [N_r, 3] arrays or None (then call ``thermalize``). Every sort_period steps (0: never) each
    replica's particles are re-laid in cell order in device memory; ``state`` and ``forces``
    always return them in the order given here."""

    def __init__(self, positions, L, kT, seeds, velocities=None, epsilon=1.0, sigma=1.0,
                 r_cut=2.5, buffer=0.4, dt=0.005, tau=0.0, sort_period=100):
        self.nrep = len(positions)
        self.n = np.array([len(p) for p in positions], np.uint32)
        self.L = np.ascontiguousarray(L, np.float64).reshape(self.nrep)
        self.kT = np.ascontiguousarray(kT, np.float64).reshape(self.nrep)
        self.seeds = np.ascontiguousarray(seeds, np.uint16).reshape(self.nrep)
        self.offsets = np.concatenate([[0], np.cumsum(self.n.astype(np.int64))])
        self.N = int(self.offsets[-1])
        pos = np.ascontiguousarray(np.concatenate(positions), np.float64)
        vel = None
        if velocities is not None:
            vel = np.ascontiguousarray(np.concatenate(velocities), np.float64)
        self._h = _lib.lib().hm_batch_create(
            self.nrep, self.n.ctypes.data, self.L.ctypes.data, self.kT.ctypes.data,
            self.seeds.ctypes.data, pos.ctypes.data, vel.ctypes.data if vel is not None else None,
            None, float(epsilon), float(sigma), float(r_cut), float(buffer), float(dt),
            float(tau))
        if not self._h:
            raise _lib.HoomdAmdError(-1, "hm_batch_create: " + _lib.lib().hm_last_error().decode())
        self.timestep = 0
        _lib.check(_lib.lib().hm_batch_set_sort_period(self._h, int(sort_period)), "sort period")

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.lib().hm_batch_destroy(self._h)
            self._h = None

    @classmethod
    def seeded(cls, nrep=256, seed=1, **kw):
        """The seeded workload of ``seeded_replicas``, velocities thermalized at each kT_r."""
        n, L, kT, seeds, pos = seeded_replicas(nrep, seed)
        b = cls(pos, L, kT, seeds, **kw)
        b.thermalize()
        return b

    def thermalize(self, timestep=None):
        t = self.timestep if timestep is None else int(timestep)
        _lib.check(_lib.lib().hm_batch_thermalize(self._h, t), "thermalize")

    def set_timestep(self, t):
        _lib.check(_lib.lib().hm_batch_set_timestep(self._h, int(t)), "timestep")
        self.timestep = int(t)

    def run(self, steps):
        _lib.check(_lib.lib().hm_batch_run(self._h, int(steps)), "run")
        self.timestep += int(steps)

    def prime(self, steps):
        return _lib.check(_lib.lib().hm_batch_prime(self._h, int(steps)), "prime")

    def tune(self):
        """Autotune the pair launch; returns the sweep rows (tpp, block, ms)."""
        log = np.zeros((64, 3), np.float32)
        rows = _lib.check(_lib.lib().hm_batch_tune(self._h, log.ctypes.data, 64), "tune")
        return log[:rows]

    def set_pair_launch(self, tpp, block):
        _lib.check(_lib.lib().hm_batch_set_pair_launch(self._h, int(tpp), int(block)), "launch")

    @property
    def config(self):
        out = np.zeros(6, np.uint32)
        _lib.check(_lib.lib().hm_batch_config(self._h, out.ctypes.data), "config")
        return dict(tpp=int(out[0]), block=int(out[1]), autotuned=bool(out[2]), cmax=int(out[3]),
                    nmax=int(out[4]), cells=int(out[5]))

    def _split(self, a):
        return [a[self.offsets[r]:self.offsets[r + 1]] for r in range(self.nrep)]

    def state(self):
        """Lists of positions, velocities and images per replica."""
        p = np.zeros((self.N, 3))
        v = np.zeros((self.N, 3))
        im = np.zeros((self.N, 3), np.int32)
        _lib.check(_lib.lib().hm_batch_get_state(self._h, p.ctypes.data, v.ctypes.data,
                                                 im.ctypes.data), "state")
        return self._split(p), self._split(v), self._split(im)

    def forces(self):
        f = np.zeros((self.N, 4))
        _lib.check(_lib.lib().hm_batch_get_forces(self._h, f.ctypes.data), "forces")
        return self._split(f)

    def replica_stats(self):
        """[nrep, 4]: kinetic energy, pair energy, list builds, neighbour-list entries."""
        out = np.zeros((self.nrep, 4))
        _lib.check(_lib.lib().hm_batch_replica_stats(self._h, out.ctypes.data), "stats")
        return out

    def stats(self):
        out = np.zeros(4, np.uint64)
        _lib.check(_lib.lib().hm_batch_stats(self._h, out.ctypes.data), "stats")
        return dict(particles=int(out[0]), steps=int(out[1]), replays=int(out[2]),
                    builds=int(out[3]))

    def time_pair(self, iters=20):
        ms = C.c_double()
        _lib.check(_lib.lib().hm_batch_time_pair(self._h, int(iters), C.byref(ms)), "time")
        return ms.value
