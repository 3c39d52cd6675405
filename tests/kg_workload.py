"""The Kremer-Grest bead-spring melt workload, built through the hoomd_amd surface (shared by
the GPU tests and the benchmark).

Chains are random walks with bond length 0.97 at bead density 0.85. The overlaps of the ideal
chains are pushed apart by a soft cosine pair potential U = A (1 + cos(pi r / r_c)),
r_c = 2^(1/6), whose amplitude is ramped linearly (a variant.Ramp) from 0 to a_max over the
push-off steps, with FENE bonds (k = 30, R0 = 1.5, WCA eps = sigma = 1) and Langevin dynamics
(kT = 1, gamma = 1, dt = 0.01). Then the pair potential becomes WCA (LJ eps = sigma = 1,
r_cut = 2^(1/6), mode shift) with a buffer of 0.4 and bond exclusions.
"""

import numpy as np

from lattices import random_walk_chains

WCA_RC = 2.0 ** (1.0 / 6.0)


def melt_snapshot(hoomd, n_chains, chain_len, rho=0.85, bond=0.97, seed=1):
    pos, image, bonds, L = random_walk_chains(n_chains, chain_len, rho, bond, seed)
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = len(pos)
    snap.particles.types = ["A"]
    snap.particles.position[:] = pos
    snap.particles.image[:] = image
    snap.bonds.N = len(bonds)
    snap.bonds.types = ["backbone"]
    snap.bonds.group[:] = bonds
    return snap, L


def fene():
    import hoomd_amd as hoomd
    f = hoomd.md.bond.FENEWCA()
    f.params["backbone"] = dict(k=30.0, r0=1.5, epsilon=1.0, sigma=1.0, delta=0.0)
    return f


def pushed_off_melt(hoomd, n_chains, chain_len, pushoff_steps=10000, ramp_updates=100,
                    a_max=100.0, rho=0.85, seed=1, log=None):
    """A Simulation of the melt after the push-off, running WCA + FENE + Langevin."""
    from hoomd_amd.md.mlcg import soft_cosine_table
    from hoomd_amd.variant import Ramp
    snap, L = melt_snapshot(hoomd, n_chains, chain_len, rho, seed=seed)
    sim = hoomd.Simulation(device=hoomd.device.GPU(), seed=seed)
    sim.create_state_from_snapshot(snap)
    sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.0)
    nl = hoomd.md.nlist.Cell(buffer=0.4, exclusions=("bond",))
    soft = hoomd.md.pair.Table(nlist=nl, default_r_cut=WCA_RC)
    ramp = Ramp(0.0, a_max, 0, pushoff_steps)
    soft.params[("A", "A")] = soft_cosine_table(ramp(0), WCA_RC)
    langevin = hoomd.md.methods.Langevin(filter=hoomd.filter.All(), kT=1.0)
    langevin.gamma["A"] = 1.0
    sim.operations.integrator = hoomd.md.Integrator(dt=0.01, forces=[soft, fene()],
                                                    methods=[langevin])
    chunk = max(1, pushoff_steps // ramp_updates)
    done = 0
    while done < pushoff_steps:
        k = min(chunk, pushoff_steps - done)
        sim.run(k)
        done += k
        soft.params[("A", "A")] = soft_cosine_table(ramp(done), WCA_RC)
    if log:
        log(f"push-off: {pushoff_steps} steps, A ramped to {a_max}")
    wca = hoomd.md.pair.LJ(nlist=nl, default_r_cut=WCA_RC, mode="shift")
    wca.params[("A", "A")] = dict(epsilon=1.0, sigma=1.0)
    sim.operations.integrator = hoomd.md.Integrator(dt=0.01, forces=[wca, fene()],
                                                    methods=[langevin])
    return sim, L
