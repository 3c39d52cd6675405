"""GPU parity of the Table pair evaluator (md.pair.Table; EvaluatorPairTable.h) through the
driver (hm_pair_table) and the engine's pair kernels, against the FP64 oracle."""

import math

import numpy as np
import pytest

import oracle as O
from gpu_helpers import nlist, pair_table, rel_rms
from lattices import maxwell_boltzmann, perturbed_fcc

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5  # relative RMS vs FP64 per quantity


@pytest.fixture(scope="module")
def hoomd():
    import hoomd_amd
    from hoomd_amd import _lib
    if _lib.lib().hm_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return hoomd_amd


def lj_table(r_min, r_cut, width):
    r = np.linspace(r_min, r_cut, width, endpoint=False)
    return 4 * (r**-12 - r**-6), 24 * (2 * r**-13 - r**-7)


def test_table_golden_two_particles(hoomd, golden):
    """forces_and_energies.json 'Table' at r = 0.75 and 1.5 (rtol 5e-6, test_potential.py)."""
    fx = golden("table_forces_and_energies.json")
    t = fx["Table"]
    for p, forces, energies in zip(t["params"], t["forces"], t["energies"]):
        for i, d in enumerate(fx["distances"]):
            pos = np.array([[0.0, 0.0, 0.1], [0.0, 0.0, d + 0.1]])
            nl = np.array([[1] + [0] * 7, [0] * 8], np.uint32)
            nn = np.array([1, 1], np.uint32)
            f, _ = pair_table(pos, 20.0, nl, nn, fx["r_cut"], p["r_min"], p["U"], p["F"])
            assert math.isclose(f[:, 3].sum(), energies[i], rel_tol=5e-6, abs_tol=1e-6)
            assert math.isclose(f[0, 2], forces[i], rel_tol=5e-6, abs_tol=1e-6)


@pytest.mark.parametrize("tpp", [1, 4, 16])
def test_table_driver_matches_fp64(hoomd, tpp):
    pos, L = perturbed_fcc(10, 0.8442, r=0.05, seed=9)
    nl, nn, _ = nlist(pos, L, 2.5, 0.4)
    U, F = lj_table(0.8, 2.5, 301)
    f, v = pair_table(pos, L, nl, nn, 2.5, 0.8, U, F, tpp=tpp)
    rf, rv = O.pair_table(O.Box(L), pos, nl, nn, False, 2.5, 0.8, U, F)
    assert rel_rms(f[:, :3], rf[:, :3]) < FORCE_TOL
    assert rel_rms(f[:, 3], rf[:, 3]) < FORCE_TOL
    assert rel_rms(v, rv) < FORCE_TOL


def test_table_driver_rejects_shift(hoomd):
    from hoomd_amd import _lib
    import ctypes as C
    a = _lib.HmPairArgs()
    a.N, a.ntypes, a.nmax, a.shift_mode = 1, 1, 8, 1
    a.pos = a.nlist = a.n_neigh = a.force = 16
    P = (_lib.HmTableParams * 1)()
    assert _lib.lib().hm_pair_table(C.byref(a), P, None) != 0


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_table_engine_forces_match_oracle(hoomd, kind):
    """The engine's Table force on an LJ liquid (the cell-tiled request falls back to the
    relative gather) against the oracle's FP64 table evaluation on the oracle's list."""
    from hoomd_amd import _lib
    pos, L = perturbed_fcc(8, 0.8442, r=0.05, seed=2)
    vel = maxwell_boltzmann(len(pos), 1.0, seed=2)
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = len(pos)
    snap.particles.position[:] = pos
    snap.particles.velocity[:] = vel
    sim = hoomd.Simulation(device=hoomd.device.GPU(), seed=1)
    sim.create_state_from_snapshot(snap)
    nl = hoomd.md.nlist.Cell(buffer=0.4)
    table = hoomd.md.pair.Table(nlist=nl, default_r_cut=2.5)
    U, F = lj_table(0.85, 2.5, 500)
    table.params[("A", "A")] = dict(r_min=0.85, U=U, F=F)
    sim.operations.integrator = hoomd.md.Integrator(
        dt=0.005, forces=[table], methods=[hoomd.md.methods.ConstantVolume(hoomd.filter.All())])
    sim.operations.tuners.clear()
    sim.run(0)
    _lib.check(_lib.lib().hm_sys_set_pair_kernel(sim._system, kind, 0, 0))
    sim.run(0)
    cfg = np.zeros(5, np.uint32)
    _lib.check(_lib.lib().hm_sys_pair_config(sim._system, cfg.ctypes.data))
    assert cfg[0] == (2 if kind == 3 else kind)
    onl, onn = O.nlist_binned(O.Box(L), pos, 2.5, 0.4, full=True)
    rf, rv = O.pair_table(O.Box(L), pos, onl, onn, False, 2.5, 0.85, U, F)
    assert rel_rms(table.forces, rf[:, :3]) < FORCE_TOL
    assert rel_rms(table.energies, rf[:, 3]) < FORCE_TOL
    assert rel_rms(table.virials.T, rv) < FORCE_TOL
    # a few steps run and stay finite, and the fine table tracks the LJ force closely
    sim.run(20)
    assert np.isfinite(table.forces).all()
    lf, _ = O.pair_lj(O.Box(L), pos, onl, onn, third_law=False)
    assert rel_rms(rf[:, :3], lf[:, :3]) < 1e-3


def test_table_rejects_bad_params(hoomd):
    nl = hoomd.md.nlist.Cell(buffer=0.4)
    t = hoomd.md.pair.Table(nlist=nl, default_r_cut=2.5)
    with pytest.raises(RuntimeError):
        t._table(dict(r_min=0.5, U=[1.0, 2.0], F=[1.0]))
    with pytest.raises(RuntimeError):
        t._table(dict(r_min=0.5, U=[], F=[]))
    with pytest.raises(ValueError):
        t._table(dict(U=[1.0], F=[1.0]))
