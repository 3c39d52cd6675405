"""Engine limits and recovery: the box-size error (NeighborList::checkBoxSize,
hoomd/md/NeighborList.cc:538-561) and the overflow replay of a chunk whose cell or row capacity
was too small (the reference's regrow-and-repeat, NeighborList.cc:411-427, CellList.cc:183-196),
checked against the oracle's trajectory."""

import numpy as np
import pytest

import oracle as O
from test_gpu_engine import lj_state, make_sim

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hoomd():
    import hoomd_amd
    from hoomd_amd import _lib
    if _lib.lib().hm_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return hoomd_amd


def test_box_too_small_for_the_list(hoomd):
    from hoomd_amd import _lib
    rng = np.random.default_rng(0)
    L = 5.7  # 2 (r_cut + buffer) = 5.8
    pos = rng.uniform(-L / 2, L / 2, (20, 3))
    sim, _ = make_sim(hoomd, pos, np.zeros((20, 3)), L)
    with pytest.raises(_lib.HoomdAmdError) as e:
        sim.run(0)
    assert e.value.code == -9 and "too small" in str(e.value)


@pytest.mark.parametrize("cmax,nmax", [(4, 256), (256, 16), (3, 8)])
def test_overflow_replay_matches_oracle(hoomd, cmax, nmax):
    """Capacities far below the need in the middle of a run: the next build overflows, the chunk
    restarts from its first state with grown capacities, and the trajectory is the oracle's."""
    from hoomd_amd import _lib
    pos, vel, L = lj_state(8)
    sim, lj = make_sim(hoomd, pos, vel, L)
    sim.run(10)
    st = np.zeros(9, np.uint64)
    _lib.check(_lib.lib().hm_sys_stats(sim._system, st.ctypes.data))
    rb0 = int(st[5])
    _lib.check(_lib.lib().hm_sys_set_list_capacity(sim._system, cmax, nmax), "capacity")
    sim.run(40)
    _lib.check(_lib.lib().hm_sys_stats(sim._system, st.ctypes.data))
    print(f"replays {int(st[5]) - rb0}, nmax {int(st[1])}, cmax {int(st[2])}", flush=True)
    assert int(st[5]) > rb0
    if cmax < 16:  # a cell holds about 21 particles here
        assert int(st[2]) > cmax
    if nmax < 64:  # a row holds about 90 neighbours here
        assert int(st[1]) > 64
    ref = O.System(pos, vel, L, dt=0.005)
    ref.run(50)
    rp, rv, ri = ref.state()
    snap = sim.state.get_snapshot()
    gp = snap.particles.position + snap.particles.image * L
    assert np.abs(gp - (rp + ri * L)).max() < 1e-6
    assert np.abs(snap.particles.velocity - rv).max() < 1e-5
    assert sim.timestep == 50
