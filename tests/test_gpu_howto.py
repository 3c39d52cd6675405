"""GPU parity of the reference's how-to benchmark workload: spheres.gsd replicated, WCA
(LJ r_cut = 2^(1/6)), Cell(buffer = 0.2), Bussi(kT = 1.2), dt = 0.001; and the GSD writer
driving split runs."""

import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN
from gpu_helpers import rel_rms

pytestmark = pytest.mark.gpu

SPHERES = os.path.join(GOLDEN, "spheres.gsd")
WCA = 2 ** (1 / 6)


@pytest.fixture(scope="module")
def hoomd():
    import hoomd_amd
    from hoomd_amd import _lib
    if _lib.lib().hm_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return hoomd_amd


def howto_sim(hoomd, rep, seed=5, writer=None):
    sim = hoomd.Simulation(device=hoomd.device.GPU(), seed=seed)
    sim.timestep = 0
    sim.create_state_from_gsd(SPHERES)
    sim.state.replicate(*rep)
    sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.2)
    cell = hoomd.md.nlist.Cell(buffer=0.2)
    lj = hoomd.md.pair.LJ(nlist=cell)
    lj.params[("A", "A")] = dict(sigma=1, epsilon=1)
    lj.r_cut[("A", "A")] = WCA
    bussi = hoomd.md.methods.thermostats.Bussi(kT=1.2)
    method = hoomd.md.methods.ConstantVolume(filter=hoomd.filter.All(), thermostat=bussi)
    sim.operations.integrator = hoomd.md.Integrator(dt=0.001, methods=[method], forces=[lj])
    sim.operations.tuners.clear()
    if writer is not None:
        sim.operations.writers.append(writer)
    return sim, lj


def test_howto_state_forces_and_list(hoomd):
    sim, lj = howto_sim(hoomd, (2, 2, 1))
    snap = sim.state.get_snapshot()
    assert snap.particles.N == 4 * 2048
    L = np.array(snap.configuration.box[:3])
    v = snap.particles.velocity
    # thermalized: zero COM momentum, temperature near 1.2
    assert np.abs(v.sum(axis=0)).max() < 1e-9
    assert abs((v**2).sum() / (3 * len(v) - 3) / 1.2 - 1) < 0.05
    sim.run(0)
    pos = snap.particles.position
    box = O.Box(*L)
    nl, nn = O.nlist_binned(box, pos, WCA, 0.2, full=True)
    rf, rv = O.pair_lj(box, pos, nl, nn, third_law=False, r_cut=WCA)
    assert rel_rms(lj.forces, rf[:, :3]) < 1e-5
    assert rel_rms(lj.energies, rf[:, 3]) < 1e-5
    got = lj.nlist.pair_list.astype(np.int64)
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    onl, onn = O.nlist_binned(box, pos, WCA, 0.2, full=False)
    assert np.array_equal(got, O.pair_array(onl, onn))


def test_howto_bussi_steps_match_oracle(hoomd):
    sim, lj = howto_sim(hoomd, (2, 1, 1), seed=11)
    snap = sim.state.get_snapshot()
    L = np.array(snap.configuration.box[:3])
    pos0 = snap.particles.position
    vel0 = snap.particles.velocity
    sim.run(40)
    ref = O.System(pos0, vel0, L, r_cut=WCA, mode="none", buffer=0.2, dt=0.001, seed=11)
    ref.set_bussi(1.2, 0.0)
    ref.run(40)
    rp, rv, ri = ref.state()
    out = sim.state.get_snapshot()
    gp = out.particles.position + (out.particles.image - snap.particles.image) * L
    assert np.abs(gp - (rp + ri * L)).max() < 1e-6
    assert rel_rms(out.particles.velocity - vel0, rv - vel0) < 1e-5
    assert lj.nlist.num_builds == ref.n_builds


def test_gsd_writer_split_run(hoomd, tmp_path):
    from hoomd_amd import gsd
    fn = str(tmp_path / "traj.gsd")
    w = hoomd.write.GSD(trigger=hoomd.trigger.Periodic(25), filename=fn, mode="wb",
                        dynamic=["property", "momentum"])
    a, la = howto_sim(hoomd, (2, 1, 1), seed=3, writer=w)
    b, lb = howto_sim(hoomd, (2, 1, 1), seed=3)
    a.run(110, write_at_start=True)
    b.run(110)
    w.close()
    g = gsd.GSDFile(fn)
    steps = [int(g.read_chunk(k, "configuration/step")[0]) for k in range(len(g))]
    assert steps == [0, 25, 50, 75, 100]
    sa, sb = a.state.get_snapshot(), b.state.get_snapshot()
    assert a.timestep == b.timestep == 110
    assert la.nlist.num_builds == lb.nlist.num_builds
    np.testing.assert_allclose(sa.particles.position, sb.particles.position, atol=1e-6)
    # the frame at step 100 holds the state of a 100-step run
    c, _ = howto_sim(hoomd, (2, 1, 1), seed=3)
    c.run(100)
    sc = c.state.get_snapshot()
    last = gsd.read_snapshot(fn, -1)
    np.testing.assert_allclose(last.particles.position, sc.particles.position.astype(np.float32),
                               atol=2e-6)
    np.testing.assert_array_equal(last.particles.image, sc.particles.image)


def test_nlist_buffer_tuner_grid(hoomd):
    """NeighborListBuffer.with_grid (md/tune/nlist_buffer.py): acts every 100 steps; the first
    act of the run starts the clock, the second bins the domain, five measured bins, then the
    best bin's centre, solved twice -> tuned. Forces stay exact through the buffer changes."""
    sim, lj = howto_sim(hoomd, (2, 2, 2), seed=2)
    tuner = hoomd.md.tune.NeighborListBuffer.with_grid(
        trigger=100, nlist=lj.nlist, maximum_buffer=0.45, minimum_buffer=0.05, n_bins=5)
    sim.operations.add(tuner)
    sim.run(900)
    assert tuner.tuned
    lo, hi = tuner.final_bin
    centres = [0.09, 0.17, 0.25, 0.33, 0.41]
    assert any(abs((lo + hi) / 2 - c) < 1e-12 for c in centres)
    assert abs(lj.nlist.buffer - (lo + hi) / 2) < 1e-12
    assert tuner.max_tps > 0 and 0.05 <= tuner.best_buffer_size <= 0.45
    assert sim.timestep == 900
    snap = sim.state.get_snapshot()
    L = np.array(snap.configuration.box[:3])
    pos = snap.particles.position
    box = O.Box(*L)
    nl, nn = O.nlist_binned(box, pos, WCA, 0.0, full=True)
    rf, _ = O.pair_lj(box, pos, nl, nn, third_law=False, r_cut=WCA)
    assert rel_rms(lj.forces, rf[:, :3]) < 1e-5
    # a later run keeps the tuned buffer without rebuilding the system
    h = sim._system
    sim.run(50)
    assert sim._system == h and abs(lj.nlist.buffer - (lo + hi) / 2) < 1e-12


def test_logger_table_during_run(hoomd):
    """A Logger of the simulation, thermodynamic quantities and the pair force, written by
    write.Table every 10 steps of a run: a header and one row per trigger step, the last row
    holding the values the objects report after the run."""
    import io
    sim, lj = howto_sim(hoomd, (1, 1, 1), seed=4)
    thermo = hoomd.md.compute.ThermodynamicQuantities(hoomd.filter.All())
    sim.operations.add(thermo)
    log = hoomd.logging.Logger(categories=["scalar", "string"])
    log.add(sim, quantities=["timestep"])
    log.add(thermo, quantities=["kinetic_temperature", "potential_energy"])
    log.add(lj, quantities=["energy"])
    out = io.StringIO()
    sim.operations.add(hoomd.write.Table(trigger=10, logger=log, output=out))
    sim.run(30)
    lines = out.getvalue().splitlines()
    assert lines[0].split() == ["Simulation.timestep",
                                "md.compute.ThermodynamicQuantities.kinetic_temperature",
                                "md.compute.ThermodynamicQuantities.potential_energy",
                                "md.pair.LJ.energy"]
    rows = [r.split() for r in lines[1:]]
    assert [int(r[0]) for r in rows] == [10, 20, 30]
    last = rows[-1]
    assert abs(float(last[1]) - thermo.kinetic_temperature) < 1e-4 * thermo.kinetic_temperature
    assert abs(float(last[3]) - lj.energy) < 1e-3 * abs(lj.energy) + 1e-6
