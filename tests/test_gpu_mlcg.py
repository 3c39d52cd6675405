"""The learned CG site energy through the engine and the hoomd_amd surface (md.mlcg): forces of
a running system against the FP64 oracle, force-matching training, and the whole
fine-grained -> coarse-grained workflow on a small Kremer-Grest melt, where the trained network
must lower the held-out force residual and bring the CG liquid's site RDF far closer to the
mapped one than the same CG model without it.

Tolerances: engine forces and site energies against the bf16-rounding oracle 5e-4 relative RMS
(the kernel-level tests state why); loss 2e-3 relative and gradient 3e-3 relative RMS against
the same oracle after training moved the parameters."""

import numpy as np
import pytest

import oracle as O
from oracle import mlp_cg as M
from test_gpu_mlp import cg_config, rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hoomd():
    from hoomd_amd import _lib
    if _lib.lib().hm_device_count() < 1:
        pytest.skip("no GPU")
    import hoomd_amd
    return hoomd_amd


def cg_sim(hoomd, pos, L, bonds=None, params=None, gmean=None, gscale=None, exclusions=()):
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = len(pos)
    snap.particles.types = ["S"]
    snap.particles.position[:] = pos
    forces = []
    nl = hoomd.md.nlist.Cell(buffer=0.4, exclusions=exclusions)
    mlp = hoomd.md.mlcg.SiteEnergyMLP(nlist=nl, params=params)
    if gmean is not None:
        mlp.normalization = (gmean, gscale)
    forces.append(mlp)
    harm = None
    if bonds is not None:
        snap.bonds.N = len(bonds)
        snap.bonds.types = ["cg"]
        snap.bonds.group[:] = bonds
        harm = hoomd.md.bond.Harmonic()
        harm.params["cg"] = dict(k=5.0, r0=1.6)
        forces.append(harm)
    sim = hoomd.Simulation(device=hoomd.device.GPU(), seed=3)
    sim.create_state_from_snapshot(snap)
    lang = hoomd.md.methods.Langevin(filter=hoomd.filter.All(), kT=1.0)
    sim.operations.integrator = hoomd.md.Integrator(dt=0.02, forces=forces, methods=[lang])
    sim.operations.tuners.clear()
    return sim, mlp, harm


def stats_model(pos, L, nl, nn):
    G = M.descriptor(M.Model(), L, pos, nl, nn)
    gmean, gscale = G.mean(axis=0), 1.0 / (G.std(axis=0) + 1e-3)
    # the engine's production force pass gathers the bf16 copy of dE/dG (oracle bf16 mode 2);
    # its training reads the FP32 dE/dG (mode 1)
    return M.Model(gmean=gmean, gscale=gscale, bf16=2), gmean, gscale


def test_engine_forces_match_oracle(hoomd):
    print("engine mlp forces: N=2500", flush=True)
    pos, L = cg_config(2500, seed=11)
    nl, nn = O.nlist_binned(O.Box(L), pos, 3.0, 0.0, full=True)
    om, gmean, gscale = stats_model(pos, L, nl, nn)
    p = M.init_params(21)
    sim, mlp, _ = cg_sim(hoomd, pos, L, params=p.astype(np.float32), gmean=gmean, gscale=gscale)
    sim.run(0)
    _, oF, oE, _, _ = M.forces(om, p.astype(np.float32).astype(np.float64), L, pos, nl, nn)
    eF, eE = rel_rms(mlp.forces, oF), rel_rms(mlp.energies, oE)
    print(f"engine forces rel RMS {eF:.2e}, site energies {eE:.2e}", flush=True)
    assert eF < 5e-4 and eE < 5e-4
    # one step later (the list with its buffer, new positions) still matches
    sim.run(5)
    s = sim.state.get_snapshot()
    p2 = s.particles.position
    nl2, nn2 = O.nlist_binned(O.Box(L), p2, 3.0, 0.0, full=True)
    _, oF2, _, _, _ = M.forces(om, p.astype(np.float32).astype(np.float64), L, p2, nl2, nn2)
    assert rel_rms(mlp.forces, oF2) < 5e-4


def test_engine_bonds_and_exclusions(hoomd):
    print("engine mlp + harmonic bonds", flush=True)
    pos, L = cg_config(1200, seed=12)
    # bond each even site to its nearest odd partner within 2.6
    bonds = []
    used = set()
    for i in range(0, len(pos), 2):
        d = pos[1::2] - pos[i]
        d -= L * np.rint(d / L)
        r = np.sqrt((d * d).sum(axis=1))
        j = int(np.argmin(r)) * 2 + 1
        if r.min() < 2.6 and j not in used:
            bonds.append((i, j))
            used.add(j)
    bonds = np.asarray(bonds, np.uint32)
    excl = [[] for _ in range(len(pos))]
    for a, b in bonds:
        excl[a].append(int(b))
        excl[b].append(int(a))
    nl, nn = O.nlist_binned(O.Box(L), pos, 3.0, 0.0, full=True, excl=excl)
    om, gmean, gscale = stats_model(pos, L, nl, nn)
    p = M.init_params(22).astype(np.float32)
    sim, mlp, harm = cg_sim(hoomd, pos, L, bonds=bonds, params=p, gmean=gmean, gscale=gscale,
                            exclusions=("bond",))
    sim.run(0)
    _, oF, _, _, _ = M.forces(om, p.astype(np.float64), L, pos, nl, nn)
    assert rel_rms(mlp.forces, oF) < 5e-4
    from hoomd_amd.md.mlcg import harmonic_bond_forces
    hb = harmonic_bond_forces(pos, L, bonds, 5.0, 1.6)
    assert rel_rms(harm.forces, hb) < 1e-5


def test_training_teacher_student(hoomd):
    """Targets from a teacher network (FP64 oracle forces); a student from other random
    parameters learns them: the loss falls by more than 20x, and the engine's loss and gradient
    at the trained parameters match the oracle."""
    from hoomd_amd.md.mlcg import FrameSet
    N, nfr = 800, 6
    confs = [cg_config(N, seed=30 + k) for k in range(nfr)]
    L = confs[0][1]
    nl0, nn0 = O.nlist_binned(O.Box(L), confs[0][0], 3.0, 0.0, full=True)
    om, gmean, gscale = stats_model(confs[0][0], L, nl0, nn0)
    om64 = M.Model(gmean=gmean, gscale=gscale)
    teacher = M.init_params(40)
    fs = FrameSet(nfr, N)
    for k, (p, _) in enumerate(confs):
        nl, nn = O.nlist_binned(O.Box(L), p, 3.0, 0.0, full=True)
        _, F, _, _, _ = M.forces(om64, teacher, L, p, nl, nn)
        fs.set(k, positions=p, forces=F)
    sim, mlp, _ = cg_sim(hoomd, confs[0][0], L, params=M.init_params(41).astype(np.float32),
                         gmean=gmean, gscale=gscale)
    loss = mlp.train(fs, 600, lr=3e-3)
    first, last = loss[:nfr].mean(), loss[-nfr:].mean()
    print(f"teacher-student loss {first:.4e} -> {last:.4e}", flush=True)
    assert np.isfinite(loss).all() and last < first / 20
    # engine loss and gradient at the trained parameters against the oracle (frame 2)
    pt = mlp.params.astype(np.float64)
    l_gpu, g_gpu = mlp.loss_grad(fs, 2)
    p2 = confs[2][0]
    nl, nn = O.nlist_binned(O.Box(L), p2, 3.0, 0.0, full=True)
    T = fs.forces(2).astype(np.float64)
    l_or, g_or, _ = M.loss_grad(om.with_bf16(1), pt, L, fs.positions(2).astype(np.float64), nl,
                                nn, T)
    print(f"frame 2: loss {l_gpu:.6e} vs oracle {l_or:.6e}, grad rel RMS "
          f"{rel_rms(g_gpu, g_or):.2e}", flush=True)
    # after training the residual F - T is ~4 % of F, so the kernels' force differences from
    # the bf16-rounding oracle (~1e-4 of F) weigh ~2e-3 of the loss at most
    assert abs(l_gpu - l_or) <= 2e-3 * l_or
    assert rel_rms(g_gpu, g_or) < 3e-3
    # a minibatch step: its loss is the mean of its frames' losses, and its Adam step (the
    # first after a parameter reset: -lr g / (|g| + eps) per parameter) uses their mean gradient
    p0 = mlp.params
    ls, gs = zip(*[mlp.loss_grad(fs, k) for k in (0, 3, 5)])
    mlp.params = p0  # resets the Adam moments
    lb = mlp.train(fs, 1, lr=1e-3, eps=1e-8, order=[0, 3, 5], frames_per_step=3)
    assert abs(lb[0] - np.mean(ls)) <= 1e-6 * np.mean(ls)
    g = np.mean(gs, axis=0)
    step = mlp.params - p0
    big = np.abs(g) > 1e-4 * np.abs(g).max()
    expect = -1e-3 * g / (np.abs(g) + 1e-8)
    assert np.abs(step[big] - expect[big]).max() < 2e-6


def _soft_table(r_min=0.5, r_cut=3.0, width=300):
    r = r_min + np.arange(width) * (r_cut - r_min) / width
    U = 2.0 * (1.0 - r / r_cut) ** 2 - 0.3 * np.exp(-((r - 1.6) / 0.3) ** 2)
    F = 4.0 / r_cut * (1.0 - r / r_cut) - 0.3 * 2.0 * (r - 1.6) / 0.09 * np.exp(
        -((r - 1.6) / 0.3) ** 2)
    return dict(r_min=r_min, U=U, F=F)


def test_pair_table_in_force_pass(hoomd):
    """A tabulated pair term inside the site energy's force pass (SiteEnergyMLP.pair_table) gives
    the forces, energies and virials of the same site energy plus a separate md.pair.Table with
    the same table (1e-5 relative RMS: the same FP32 evaluation, summed in another order). An
    in-place update of the table (same width) takes effect without new graphs, and the training
    loss includes the term: with a zero network and targets equal to the table's forces, the
    loss vanishes."""
    from hoomd_amd.md import mlcg
    print("pair table in the force pass: N=2500", flush=True)
    pos, L = cg_config(2500, seed=17)
    p = M.init_params(5).astype(np.float32)
    tab = _soft_table()
    sim_a, mlp_a, _ = cg_sim(hoomd, pos, L, params=p)
    mlp_a.pair_table = tab
    sim_b, mlp_b, _ = cg_sim(hoomd, pos, L, params=p)
    table_b = hoomd.md.pair.Table(nlist=mlp_b.nlist, default_r_cut=3.0)
    table_b.params[("S", "S")] = tab
    lang = hoomd.md.methods.Langevin(filter=hoomd.filter.All(), kT=1.0)
    sim_b.operations.integrator = hoomd.md.Integrator(dt=0.02, forces=[mlp_b, table_b],
                                                      methods=[lang])
    sim_a.run(0)
    sim_b.run(0)
    fb = mlp_b.forces + table_b.forces
    eb = mlp_b.energies + table_b.energies
    vb = mlp_b.virials + table_b.virials
    ef, ee = rel_rms(mlp_a.forces, fb), rel_rms(mlp_a.energies, eb)
    ev = rel_rms(mlp_a.virials, vb)
    print(f"fused vs separate: force {ef:.2e}, energy {ee:.2e}, virial {ev:.2e}", flush=True)
    assert ef < 1e-5 and ee < 1e-5 and ev < 1e-5
    # in place, after step graphs were captured: doubling U and F adds the table's forces
    # once more (the forces without the table taken at the same positions)
    sim_a.run(10)
    f1 = mlp_a.forces.copy()
    mlp_a.pair_table = dict(r_min=tab["r_min"], U=2 * tab["U"], F=2 * tab["F"])
    sim_a.run(0)
    f2 = mlp_a.forces.copy()
    mlp_a.pair_table = None
    sim_a.run(0)
    f0 = mlp_a.forces.copy()
    assert rel_rms(f2 - f1, f1 - f0) < 1e-5
    sim_a.run(10)
    assert np.all(np.isfinite(mlp_a.forces))
    # training loss includes the term
    mlp_a.params = np.zeros(mlcg.NPARAM, np.float32)
    mlp_a.pair_table = tab
    sim_a.run(0)
    x = sim_a.state.get_snapshot().particles.position
    ft = mlp_a.forces  # the table alone: the zero network has no force
    fs = mlcg.FrameSet(1, len(pos))
    fs.set(0, positions=x, forces=ft)
    loss, _ = mlp_a.loss_grad(fs, 0)
    mlp_a.pair_table = None
    loss_none, _ = mlp_a.loss_grad(fs, 0)
    print(f"loss with the term {loss:.3e}, without {loss_none:.3e}", flush=True)
    # (the frame holds FP32 positions: the target differs by that rounding only)
    assert loss < 1e-6 * loss_none and loss_none > 0


def test_trained_params_survive_system_rebuild(hoomd):
    """A parameter change of another force (here the CG bonds) makes the simulation rebuild its
    device system; the site energy keeps the parameters training left on the device."""
    from scipy.spatial import cKDTree
    from hoomd_amd.md import mlcg
    pos, L = cg_config(1500, seed=23)
    x = np.mod(pos + L / 2, L)
    _, nn = cKDTree(x, boxsize=L).query(x, k=2)
    pairs = np.unique(np.sort(np.stack([np.arange(len(pos)), nn[:, 1]], 1), axis=1), axis=0)
    sim, mlp, harm = cg_sim(hoomd, pos, L, bonds=pairs.astype(np.uint32),
                            params=M.init_params(3).astype(np.float32), exclusions=("bond",))
    sim.run(0)
    fs = mlcg.FrameSet(1, len(pos))
    fs.set(0, positions=pos, forces=np.zeros((len(pos), 3)))
    from hoomd_amd import _lib
    before = mlp.params
    mlp.train(fs, 3, lr=1e-2)
    # read the device copy directly: the property would refresh the host copy itself
    trained = np.zeros(mlcg.NPARAM, np.float32)
    _lib.check(_lib.lib().hm_sys_mlp_get_params(sim._system, mlp._force_id, trained.ctypes.data))
    assert not np.array_equal(before, trained)
    harm.params["cg"] = dict(k=6.0, r0=1.6)
    sim.run(1)
    assert np.array_equal(mlp.params, trained)


def test_kg_to_cg_network_gains(hoomd):
    """Fine-grained Kremer-Grest melt (200 chains x 50 beads, random walks pushed off, WCA +
    FENE, Langevin) mapped 5:1 on the GPU. The CG model without the network: harmonic CG bonds
    from Boltzmann inversion and a repulsive pair prior from the short-range Boltzmann inversion
    of the mapped site RDF. The network is trained on what they leave of the mapped force. With
    it, the held-out force residual must be lower and the CG liquid's site RDF (bonded pairs out)
    closer to the mapped one than without it."""
    from kg_workload import pushed_off_melt
    from hoomd_amd.md import mlcg
    n_chains, chain_len, per = 200, 50, 5
    print("kg -> cg: push-off and equilibration", flush=True)
    sim, L = pushed_off_melt(hoomd, n_chains, chain_len, pushoff_steps=3000, ramp_updates=30,
                             seed=5)
    sim.run(5000)
    smap = mlcg.SiteMap.linear_chains(n_chains, chain_len, per)
    smap.attach(sim)
    nfr, every = 150, 40
    frames = mlcg.FrameSet(nfr, smap.nsites)
    for k in range(nfr):
        sim.run(every)
        frames.record(sim, k)
    cgb = smap.bonds()
    r = np.concatenate([mlcg.bond_lengths(frames.positions(k).astype(np.float64), L, cgb)
                        for k in range(nfr)])
    kb, r0 = mlcg.boltzmann_harmonic(r, kT=1.0, nbins=60, min_count=20)
    print(f"CG bonds by Boltzmann inversion: k = {kb:.3f}, r0 = {r0:.3f}", flush=True)
    assert 0.5 < r0 < 4.0 and kb > 0
    rr, g_map = None, []
    for k in range(nfr):
        rr, g = mlcg.rdf(frames.positions(k), L, 4.0, 80, exclude_bonded=cgb)
        g_map.append(g)
    g_map = np.mean(g_map, axis=0)
    prior_params, rc_prior = mlcg.rdf_prior_table(rr, g_map)
    print(f"RDF prior: cut-off {rc_prior:.3f}, U(0) = {prior_params['U'][0]:.2f}", flush=True)
    target = mlcg.FrameSet(nfr, smap.nsites)
    for k in range(nfr):
        target.set(k, positions=frames.positions(k), forces=frames.forces(k))
    # targets: mapped force minus the CG bond force, on the device; checked against the host
    # restatement (FP64 from the same FP32 frames) on the first and last frames
    target.subtract_harmonic(cgb, kb, r0, L)
    for k in (0, nfr - 1):
        p = frames.positions(k).astype(np.float64)
        T = frames.forces(k).astype(np.float64) - mlcg.harmonic_bond_forces(p, L, cgb, kb, r0)
        assert rel_rms(target.forces(k), T) < 1e-6
    # the CG system: sites, CG bonds, the prior and the network (bonded pairs excluded)
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = smap.nsites
    snap.particles.types = ["S"]
    snap.particles.position[:] = frames.positions(0)
    snap.particles.mass[:] = float(per)
    snap.bonds.N = len(cgb)
    snap.bonds.types = ["cg"]
    snap.bonds.group[:] = cgb
    cg = hoomd.Simulation(device=hoomd.device.GPU(), seed=7)
    cg.create_state_from_snapshot(snap)
    nl = hoomd.md.nlist.Cell(buffer=0.4, exclusions=("bond",))
    mlp = hoomd.md.mlcg.SiteEnergyMLP(nlist=nl, seed=3)
    harm = hoomd.md.bond.Harmonic()
    harm.params["cg"] = dict(k=kb, r0=r0)
    prior = hoomd.md.pair.Table(nlist=nl, default_r_cut=rc_prior)
    prior.params[("S", "S")] = prior_params
    lang = hoomd.md.methods.Langevin(filter=hoomd.filter.All(), kT=1.0)
    lang.gamma["S"] = float(per)
    cg.operations.integrator = hoomd.md.Integrator(dt=0.02, forces=[mlp, harm, prior],
                                                   methods=[lang])
    cg.operations.tuners.clear()
    target.subtract_pair(cg, prior)
    G = np.concatenate([mlp.descriptors(target, k) for k in range(0, nfr, 10)])
    mlp.fit_normalization(G)
    train, held = np.arange(0, nfr - 20), np.arange(nfr - 20, nfr)
    zero = np.zeros(mlcg.NPARAM, np.float32)
    mlp.params = zero
    base = np.mean([mlp.loss_grad(target, int(k))[0] for k in held])
    mlp.params = mlcg.init_params(3)
    order = np.random.default_rng(0).choice(train, size=3000)
    loss = mlp.train(target, 3000, lr=1e-3, order=order)
    trained = np.mean([mlp.loss_grad(target, int(k))[0] for k in held])
    print(f"held-out force residual: without the network {base:.4f}, with it {trained:.4f} "
          f"({trained / base:.4f}); training loss {loss[:100].mean():.4f} -> "
          f"{loss[-100:].mean():.4f}", flush=True)
    assert np.isfinite(loss).all()
    # most of a mapped force is thermal noise that no function of the site positions explains
    # (bond vibrations inside the blobs), so the learnable part is a small share of the residual:
    # the network must lower it on frames it was not trained on
    assert trained < base
    trained_params = mlp.params

    def cg_rdf(params):
        mlp.params = params
        cg.run(2000)
        gs = []
        for _ in range(40):
            cg.run(100)
            gs.append(mlcg.rdf(cg.state.get_snapshot().particles.position, L, 4.0, 80,
                               exclude_bonded=cgb)[1])
        return np.mean(gs, axis=0)

    e_net = np.abs(cg_rdf(trained_params) - g_map).sum() / g_map.sum()
    e_none = np.abs(cg_rdf(zero) - g_map).sum() / g_map.sum()
    print(f"site RDF relative L1 error: with the network {e_net:.4f}, without {e_none:.4f}",
          flush=True)
    assert e_net < e_none
