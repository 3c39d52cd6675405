"""GPU parity of the batched replicas (hm_batch_*) against the oracle's CPU step loop, one oracle
System per replica on identical states.

Bars: first forces and energies 1e-5 relative RMS per replica against FP64; thermalized
velocities bit-exact against the oracle's thermalize (same Philox stream per replica seed);
10-step Bussi trajectories at the single-system engine's bar (velocity change 1e-5 relative
RMS, largest velocity deviation 1e-5).
"""

import numpy as np
import pytest

import oracle as O
from gpu_helpers import rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    from hoomd_amd import _lib, batch
    if _lib.lib().hm_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return batch


def small_replicas(B, nrep=5, seed=11, n_range=(500, 3000)):
    return B.seeded_replicas(nrep, seed, n_range=n_range)


def test_first_forces_match_oracle(B):
    n, L, kT, seeds, pos = small_replicas(B)
    print(f"replicas N={list(n)} L={np.round(L, 2).tolist()}", flush=True)
    b = B.ReplicaBatch(pos, L, kT, seeds)
    forces = b.forces()
    for r in range(len(n)):
        nl, nn = O.nlist_binned(O.Box(L[r]), pos[r], 2.5, 0.4, full=True)
        rf, _ = O.pair_lj(O.Box(L[r]), pos[r], nl, nn, third_law=False)
        ef = rel_rms(forces[r][:, :3], rf[:, :3])
        ee = rel_rms(forces[r][:, 3], rf[:, 3])
        print(f"replica {r}: N={n[r]} force rel RMS {ef:.2e}, energy rel RMS {ee:.2e}",
              flush=True)
        assert ef < 1e-5 and ee < 1e-5
    st = b.replica_stats()
    for r in range(len(n)):
        nl, nn = O.nlist_binned(O.Box(L[r]), pos[r], 2.5, 0.4, full=True)
        assert st[r, 3] == nn.sum()  # list entries equal the oracle's binned full list


def test_thermalize_matches_oracle(B):
    n, L, kT, seeds, pos = small_replicas(B, nrep=4, seed=3)
    b = B.ReplicaBatch(pos, L, kT, seeds)
    b.thermalize(0)
    _, vel, _ = b.state()
    for r in range(len(n)):
        ref = O.thermalize(np.ones(n[r]), np.arange(n[r]), kT[r], 0, int(seeds[r]))
        assert np.abs(vel[r] - ref).max() < 1e-12


def test_bussi_trajectories_match_oracle(B):
    n, L, kT, seeds, pos = small_replicas(B, nrep=4, seed=5)
    vel = [O.thermalize(np.ones(n[r]), np.arange(n[r]), kT[r], 0, int(seeds[r]))
           for r in range(len(n))]
    b = B.ReplicaBatch(pos, L, kT, seeds, velocities=vel)
    b.run(10)
    _, v, _ = b.state()
    for r in range(len(n)):
        ref = O.System(pos[r], vel[r], L[r], dt=0.005, seed=int(seeds[r]))
        ref.set_bussi(float(kT[r]), 0.0)
        ref.run(10)
        _, rv, _ = ref.state()
        e = rel_rms(v[r] - vel[r], rv - vel[r])
        print(f"replica {r}: N={n[r]} velocity change rel RMS {e:.2e}, max dev "
              f"{np.abs(v[r] - rv).max():.2e}", flush=True)
        assert e < 1e-5
        assert np.abs(v[r] - rv).max() < 1e-5


def test_replica_independent_of_batch(B):
    """A replica advanced inside a batch follows the same trajectory as the same replica alone,
    bit for bit: nothing of one replica enters another's step, and cells are ordered by particle
    index and sums taken in a fixed order, so scheduling changes nothing (cell-order sorts at
    steps 0 and 100 included)."""
    n, L, kT, seeds, pos = small_replicas(B, nrep=3, seed=9)
    vel = [O.thermalize(np.ones(n[r]), np.arange(n[r]), kT[r], 0, int(seeds[r]))
           for r in range(len(n))]
    both = B.ReplicaBatch(pos, L, kT, seeds, velocities=vel)
    alone = B.ReplicaBatch(pos[1:2], L[1:2], kT[1:2], seeds[1:2], velocities=vel[1:2])
    both.run(200)
    alone.run(200)
    p2, v2, i2 = both.state()
    p1, v1, i1 = alone.state()
    assert np.array_equal(p2[1], p1[0]) and np.array_equal(v2[1], v1[0])
    assert np.array_equal(i2[1], i1[0])


def test_overflow_replay_matches_oracle(B):
    """Capacities far below the need: the next step's rebuild overflows, the chunk is replayed
    from its start with grown lists, and the trajectory still matches the oracle."""
    n, L, kT, seeds, pos = small_replicas(B, nrep=3, seed=13)
    vel = [O.thermalize(np.ones(n[r]), np.arange(n[r]), kT[r], 0, int(seeds[r]))
           for r in range(len(n))]
    b = B.ReplicaBatch(pos, L, kT, seeds, velocities=vel)
    b.forces()  # prepares with the default capacities
    from hoomd_amd import _lib
    _lib.check(_lib.lib().hm_batch_set_list_sizes(b._h, 4, 8), "list sizes")
    b.run(10)
    st = b.stats()
    cfg = b.config
    print(f"replays {st['replays']}, cmax {cfg['cmax']}, nmax {cfg['nmax']}", flush=True)
    assert st["replays"] >= 1 and cfg["cmax"] > 4 and cfg["nmax"] > 8
    _, v, _ = b.state()
    for r in range(len(n)):
        ref = O.System(pos[r], vel[r], L[r], dt=0.005, seed=int(seeds[r]))
        ref.set_bussi(float(kT[r]), 0.0)
        ref.run(10)
        _, rv, _ = ref.state()
        assert rel_rms(v[r] - vel[r], rv - vel[r]) < 1e-5


def test_tuner_and_long_run(B):
    """Autotuned launch, 1000 steps over 12 replicas: every replica stays finite, rebuilds its
    lists on its own schedule and holds its own temperature (Bussi at kT_r)."""
    n, L, kT, seeds, pos = B.seeded_replicas(12, 21, n_range=(500, 4000))
    b = B.ReplicaBatch(pos, L, kT, seeds)
    b.thermalize(0)
    rows = b.tune()
    assert len(rows) > 0 and b.config["autotuned"]
    b.run(1000)
    st = b.replica_stats()
    T = 2.0 * st[:, 0] / (3.0 * n - 3.0)
    print("kT targets", np.round(kT, 3).tolist(), "measured", np.round(T, 3).tolist(), flush=True)
    assert np.all(np.isfinite(st))
    assert np.all(st[:, 2] > 1)  # every replica rebuilt its lists after the first build
    # instantaneous temperature of a canonical system fluctuates by sqrt(2 / (3N)) relative
    assert np.all(np.abs(T / kT - 1.0) < 6 * np.sqrt(2.0 / (3.0 * n)) + 0.02)
    # a second batch of the same shape takes the cached tuning
    b2 = B.ReplicaBatch(pos, L, kT, seeds)
    b2.thermalize(0)
    assert len(b2.tune()) == 0 and b2.config["tpp"] == b.config["tpp"]
