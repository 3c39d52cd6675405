"""Benchmark: the 1M-particle LJ liquid timestep on one MI355X.

    python bench.py [--gpus 1] [--steps 10000] [--warmup 1000] [--n 1048576] [--no-cpu]

Workload (the benchmark's LJ-at-scale configuration): N = 1,048,576 particles (FCC 64^3 x 4) in a
cubic periodic box at number density 0.8442, LJ eps = sigma = 1, r_cut = 2.5 (unshifted), cell
neighbour list with buffer 0.4, Maxwell-Boltzmann velocities at kT = 1.2 (Philox thermalize,
COM momentum zeroed), ConstantVolume + Bussi(kT = 1.2, tau = 0), dt = 0.005, Hilbert sort every
500 steps (grid 256). Synthetic input; timed steps run with all state resident in HBM.

Prints ONE JSON line. "value" is timesteps/s of the whole job (one GPU); particle-steps/s, the
pair kernel's roofline and the CPU baseline (the oracle's C restatement of the reference CPU path,
timed on a bounded sample of the same workload on this host) are reported beside it.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hoomd-blue_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = ("timesteps/s + particle-steps/s (1M LJ; MLP-CG vs plain) at 1 GPU; % HBM peak")
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fcc(n, rho):
    from lattices import fcc_lattice
    return fcc_lattice(n, rho)


def build_sim(hoomd, n_cells, rho, seed, sort_period=500):
    pos, L = fcc(n_cells, rho)
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = len(pos)
    snap.particles.types = ["A"]
    snap.particles.position[:] = pos
    sim = hoomd.Simulation(device=hoomd.device.GPU(), seed=seed)
    sim.create_state_from_snapshot(snap)
    sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.2)
    nl = hoomd.md.nlist.Cell(buffer=0.4)
    lj = hoomd.md.pair.LJ(nlist=nl, default_r_cut=2.5, mode="none")
    lj.params[("A", "A")] = dict(epsilon=1.0, sigma=1.0)
    bussi = hoomd.md.methods.thermostats.Bussi(kT=1.2, tau=0.0)
    method = hoomd.md.methods.ConstantVolume(filter=hoomd.filter.All(), thermostat=bussi)
    sim.operations.integrator = hoomd.md.Integrator(dt=0.005, forces=[lj], methods=[method])
    sim.operations.tuners.clear()
    sim.operations.add(hoomd.tune.ParticleSorter(trigger=sort_period, grid=256))
    return sim, lj, pos, L


PAIR_KERNELS = {
    # kind: (kernel, bytes per neighbour-list entry, bytes per particle)
    # 16-bit local rows; n_neigh + FP64 position + force, plus the cell's slot-index table
    0: ("pair_cell_kernel<EvaluatorPairLJ>", 2, 4 + 32 + 16 + 4),
    # 32-bit rows; n_neigh + FP64 position + force
    1: ("pair_kernel<EvaluatorPairLJ>", 4, 4 + 32 + 16),
    # 32-bit rows; n_neigh + FP32 cell-relative position + force
    2: ("pair_rel_kernel<EvaluatorPairLJ>", 4, 4 + 16 + 16),
    # 16-bit local rows; n_neigh + FP32 cell-relative position + cell slot table + force
    3: ("pair_tile_kernel<EvaluatorPairLJ>", 2, 4 + 16 + 4 + 16),
}


def pair_roofline(sim, lj, N, kind, period=500, samples=10, launches=20):
    """Average pair-kernel launch duration (HIP events on the engine stream around back-to-back
    launches) and the algorithmic bytes of one launch: each neighbour-list entry once plus, per
    particle, its n_neigh, its own position and its force (PAIR_KERNELS); every neighbour
    position gathered is some particle's own position, so it is counted once.

    The launch time depends on how far particles have drifted from their sorted order, so it is
    sampled at `samples` points spread over one sort period of steps (as the timed steps are),
    `launches` launches each; the entry count is averaged the same way."""
    import ctypes as C

    from hoomd_amd import _lib
    ms = C.c_double()
    ent = C.c_uint64()
    tot_ms, tot_ent = 0.0, 0
    for _ in range(samples):
        sim.run(period // samples)
        _lib.check(_lib.lib().hm_sys_time_pair(sim._system, lj._force_id, launches, C.byref(ms),
                                               C.byref(ent)), "time pair")
        tot_ms += ms.value
        tot_ent += ent.value
    ms_avg = tot_ms / samples
    ent_avg = tot_ent / samples
    _, per_entry, per_particle = PAIR_KERNELS[kind]
    B = float(per_entry) * ent_avg + N * per_particle
    achieved = B / (ms_avg * 1e-3) / 1e9
    return ms_avg, ent_avg, B, achieved


def nlist_pair_fraction(sim, N, entries, pair_ms, steps, builds):
    """The nlist + pair roofline fraction with the list cost amortised over k = steps per build
    measured in the timed steps:
        (B_pair + (B_nl + B_cl) / k) / (t_pair + (t_nl + t_cl) / k) / peak
    with the algorithmic bytes of FP64 position records (P = 32 bytes):
        B_pair = 4 sum n + N (4 + P + 16)          (row entries, n_neigh, own position, force)
        B_nl   = N (P + 16 + 4 + 4 + P) + 4 sum n  (own position, cell record, cell size,
                                                    n_neigh, last position; row entries)
        B_cl   = N (P + 16 + 4) + 4 cells          (position, cell record, slot; cell sizes)
    t_cl and t_nl are complete rebuilds at the current state (HIP events, 5 rebuilds)."""
    import ctypes as C

    from hoomd_amd import _lib
    t = (C.c_double * 2)()
    _lib.check(_lib.lib().hm_sys_time_build(sim._system, 5, t), "time build")
    P = 32
    k = steps / max(builds, 1)
    b_pair = 4.0 * entries + N * (4 + P + 16)
    b_nl = N * (P + 16 + 4 + 4 + P) + 4.0 * entries
    cells = int(np.prod(np.floor(np.asarray(sim.state.box[:3]) / (2.5 + 0.4))))
    b_cl = N * (P + 16 + 4) + 4.0 * cells
    t_s = pair_ms * 1e-3 + (t[0] + t[1]) * 1e-3 / k
    achieved = (b_pair + (b_nl + b_cl) / k) / t_s / 1e9
    return {"k": k, "t_pair_us": pair_ms * 1e3, "t_cell_us": t[0] * 1e3, "t_nlist_us": t[1] * 1e3,
            "bytes_per_step": b_pair + (b_nl + b_cl) / k, "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "target_frac": 0.60}


def load_traffic(N):
    """HBM bytes per pair launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled
    by a copy-kernel calibration), committed under profiles/ for this workload."""
    path = os.path.join(ROOT, "profiles", "pair_lj_1M_traffic.json")
    if N != 1048576 or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def host_cpu():
    """(CPU model, threads used for the CPU baseline): the host's cores, at most 16 (the share
    of one GPU on these machines)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return model, max(1, min(16, n))


def parity_block(sim, lj, L):
    """Per-quantity error of the engine's current forces, energies and virials against the
    oracle's FP64 re-evaluation on the same state (full binned list, no third law), the
    liquid the timed steps left. Test infrastructure: the oracle is the checker here."""
    import oracle as O
    from gpu_helpers import rel_rms
    t0 = time.perf_counter()
    O.set_threads(host_cpu()[1])
    pos = np.asarray(sim.state.get_snapshot().particles.position, np.float64)
    f, e, v = lj.forces, lj.energies, lj.virials
    nl, nn = O.nlist_binned(O.Box(L), pos, 2.5, 0.4, full=True)
    rf, rv = O.pair_lj(O.Box(L), pos, nl, nn, third_law=False)
    out = {"state": "after the timed steps", "reference": "oracle FP64 re-evaluation",
           "force_rel_rms": rel_rms(f, rf[:, :3]), "energy_rel_rms": rel_rms(e, rf[:, 3]),
           "virial_rel_rms": rel_rms(np.asarray(v).T, rv), "bar": 1e-5,
           "pairs_within_r_list": int(nn.sum() // 2)}
    log(f"[bench] parity vs FP64: force {out['force_rel_rms']:.3e}, energy "
        f"{out['energy_rel_rms']:.3e}, virial {out['virial_rel_rms']:.3e} "
        f"({time.perf_counter() - t0:.1f} s)")
    return out


def cpu_baseline(n_cells, rho, seed, steps):
    """Oracle (FP64 C restatement of the reference CPU path: Bussi, binned lists) on the host's
    cores (full lists on P threads; the serial form keeps the half list with the third law), on
    the same 1M-particle state for a bounded number of steps."""
    import oracle as O
    from lattices import fcc_lattice
    pos, L = fcc_lattice(n_cells, rho)
    N = len(pos)
    O.set_threads(host_cpu()[1])
    vel = O.thermalize(np.ones(N), np.arange(N), 1.2, 0, seed)
    s = O.System(pos, vel, L, r_cut=2.5, mode="none", buffer=0.4, dt=0.005, seed=seed)
    s.set_bussi(1.2, 0.0)
    s.set_sorter(500, 256)
    s.run(1)  # initial build, first step and sort outside the timed sample
    # a short probe sizes the sample to about 15 s of host work (at least `steps` steps)
    t0 = time.perf_counter()
    s.run(3)
    per = (time.perf_counter() - t0) / 3
    steps = max(steps, min(2000, int(15.0 / max(per, 1e-6))))
    b0 = s.n_builds
    t0 = time.perf_counter()
    s.run(steps)
    dt = time.perf_counter() - t0
    O.set_threads(1)
    return steps / dt, dt, int(s.n_builds - b0), steps


# sphinx-doc/howto/determine-the-most-efficient-device.rst: one A100; one EPYC 7742 MPI rank
HOWTO_A100_TPS = {1: 15955.0, 4: 7276.5}
HOWTO_EPYC_1RANK_TPS = {1: 2699.0, 4: 36.072}


def howto(replicate, steps, warmup, seed, cpu_seconds):
    """The reference's how-to device benchmark (sphinx-doc/howto/determine-the-most-efficient-
    device.py): spheres.gsd (2048 particles at density 1) replicated replicate^3 times, momenta
    thermalized at kT = 1.2, WCA (LJ sigma = epsilon = 1, r_cut = 2^(1/6)), Cell(buffer = 0.2),
    ConstantVolume + Bussi(kT = 1.2), dt = 0.001; run(100) until tuning completes, warm up, then
    time. Prints one JSON line, the reference's published A100 figure beside it."""
    import hoomd_amd as hoomd
    from hoomd_amd import _lib
    sim = hoomd.Simulation(device=hoomd.device.GPU(), seed=seed)
    sim.create_state_from_gsd(os.path.join(ROOT, "tests", "golden", "spheres.gsd"))
    sim.state.replicate(replicate, replicate, replicate)
    sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.2)
    cell = hoomd.md.nlist.Cell(buffer=0.2)
    lj = hoomd.md.pair.LJ(nlist=cell)
    lj.params[("A", "A")] = dict(sigma=1, epsilon=1)
    lj.r_cut[("A", "A")] = 2 ** (1 / 6)
    bussi = hoomd.md.methods.thermostats.Bussi(kT=1.2)
    method = hoomd.md.methods.ConstantVolume(filter=hoomd.filter.All(), thermostat=bussi)
    sim.operations.integrator = hoomd.md.Integrator(dt=0.001, methods=[method], forces=[lj])
    N = sim.state.N_particles
    snap0 = sim.state.get_snapshot()
    while not sim.operations.is_tuning_complete:
        sim.run(100)
    log(f"[bench] how-to N={N}: warmup {warmup} steps")
    sim.run(warmup)
    _lib.check(_lib.lib().hm_synchronize())
    _lib.check(_lib.lib().hm_sys_prime(sim._system, steps), "prime graphs")
    t0 = time.perf_counter()
    sim.run(steps)
    _lib.check(_lib.lib().hm_synchronize())
    elapsed = time.perf_counter() - t0
    tps = steps / elapsed
    ref = HOWTO_A100_TPS.get(replicate)
    cpu = None
    if cpu_seconds > 0:
        cps, cdt, csteps = howto_cpu(snap0, seed, cpu_seconds)
        ref_cpu = HOWTO_EPYC_1RANK_TPS.get(replicate)
        cpu = {"value": cps, "unit": "timesteps/s", "cores": 1, "kind": "port",
               "sample": f"N={N} how-to WCA state after thermalize, {csteps} steps in {cdt:.1f} s "
                         "on one host core",
               "reference_epyc7742_1rank_tps": ref_cpu,
               "ratio_to_reference_1rank": cps / ref_cpu if ref_cpu else None}
        log(f"[bench] CPU oracle {cps:.1f} steps/s")
    print(json.dumps({
        "metric": "timesteps/s (reference how-to device benchmark, WCA)", "value": tps,
        "unit": "timesteps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True,
        "reference_a100_tps": ref, "ratio_to_reference_a100": tps / ref if ref else None,
        "dtype": "f32 pair arithmetic, f64 positions/integration",
        "data": "spheres.gsd (the reference's how-to input) replicated, thermalized",
        "config": {"workload": f"how-to WCA N={N} (spheres.gsd x {replicate}^3), buffer 0.2, "
                               "Bussi kT=1.2, dt=0.001", "N": N,
                   "nlist_builds_in_timed_steps": cell.num_builds},
        "cpu_baseline": cpu}), flush=True)
    return 0


def howto_cpu(snap, seed, seconds):
    """The oracle (FP64 C restatement of the reference CPU path) on the how-to state, one host
    core: a 20-step probe sizes a sample of about ``seconds``."""
    import oracle as O
    L = np.array(snap.configuration.box[:3])
    s = O.System(snap.particles.position, snap.particles.velocity, L, r_cut=2 ** (1 / 6),
                 mode="none", buffer=0.2, dt=0.001, seed=seed)
    s.set_bussi(1.2, 0.0)
    s.set_sorter(200, 256)
    t0 = time.perf_counter()
    s.run(20)
    probe = time.perf_counter() - t0
    steps = max(20, int(seconds * 20 / max(probe, 1e-9)))
    t0 = time.perf_counter()
    s.run(steps)
    dt = time.perf_counter() - t0
    return steps / dt, dt, steps


def kg(steps, warmup, seed, cpu_steps, pushoff=10000):
    """Kremer-Grest melt production step: 20 000 chains x 50 beads at rho 0.85, random-walk
    chains (bond 0.97) pushed apart by a soft cosine ramped over `pushoff` steps, then FENE
    (k 30, r0 1.5) + WCA bonds, WCA pairs (r_cut 2^(1/6), mode shift, bond exclusions, buffer
    0.4), Langevin kT 1, gamma 1, dt 0.01. Prints one JSON line."""
    import hoomd_amd as hoomd
    from hoomd_amd import _lib
    from kg_workload import pushed_off_melt
    t0 = time.perf_counter()
    sim, L = pushed_off_melt(hoomd, 20000, 50, pushoff_steps=pushoff, seed=seed, log=log)
    _lib.check(_lib.lib().hm_synchronize())
    log(f"[bench] push-off {pushoff} steps in {time.perf_counter() - t0:.1f} s")
    snap0 = sim.state.get_snapshot()
    N = snap0.particles.N
    bonds = np.array(snap0.bonds.group, np.uint32)
    nl = sim.operations.integrator.forces[0].nlist
    log(f"[bench] KG melt N={N}, {len(bonds)} bonds: warmup {warmup} steps")
    sim.run(warmup)
    _lib.check(_lib.lib().hm_synchronize())
    _lib.check(_lib.lib().hm_sys_prime(sim._system, steps), "prime graphs")
    t0 = time.perf_counter()
    sim.run(steps)
    _lib.check(_lib.lib().hm_synchronize())
    elapsed = time.perf_counter() - t0
    builds = nl.num_builds  # builds of the last run
    tps = steps / elapsed
    cpu = None
    if cpu_steps:
        import oracle as O
        model, P = host_cpu()
        O.set_threads(P)
        snap = sim.state.get_snapshot()
        s = O.System(snap.particles.position, snap.particles.velocity, L, r_cut=2 ** (1 / 6),
                     mode="shift", buffer=0.4, dt=0.01, seed=seed)
        s.set_bonds(bonds, 30.0, 1.5, 1.0, 1.0, 0.0, exclude=True)
        s.set_langevin(1.0, 1.0)
        s.set_sorter(200, 256)
        s.run(1)
        # a bounded sample of about 15 s: probe two steps, then at least cpu_steps
        t1 = time.perf_counter()
        s.run(2)
        per = (time.perf_counter() - t1) / 2
        cpu_steps = max(cpu_steps, int(15.0 / max(per, 1e-6)))
        t1 = time.perf_counter()
        s.run(cpu_steps)
        cdt = time.perf_counter() - t1
        cpu = {"value": cpu_steps / cdt, "unit": "timesteps/s", "cores": P, "kind": "port",
               "cpu_model": model,
               "sample": f"N={N} KG melt state after the timed steps, {cpu_steps} steps, "
                         f"{cdt:.1f} s on {P} host threads"}
    print(json.dumps({
        "metric": "timesteps/s (Kremer-Grest melt, Langevin)", "value": tps,
        "unit": "timesteps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True,
        "particle_steps_per_s": tps * N,
        "dtype": "f32 pair/bond arithmetic, f64 positions/integration",
        "data": f"synthetic (random-walk chains, bond 0.97, soft-cosine push-off {pushoff} "
                "steps, thermalized)",
        "config": {"workload": f"KG melt {N // 50} x 50 beads, rho 0.85, FENE k 30 r0 1.5 + WCA, "
                               "WCA pairs shift, bond exclusions, buffer 0.4, Langevin kT 1 "
                               "gamma 1, dt 0.01", "N": N, "bonds": int(len(bonds)),
                   "nlist_builds_in_timed_steps": builds},
        "cpu_baseline": cpu}), flush=True)
    return 0


def replicas(nrep, steps, warmup, seed, cpu_seconds):
    """Batched uneven LJ replicas: nrep independent periodic LJ systems drawn from the seed
    (hoomd_amd.batch.seeded_replicas: N log-uniform in [500, 20000], density U[0.70, 0.90], kT
    U[0.8, 1.5]), eps = sigma = 1, r_cut 2.5, buffer 0.4, ConstantVolume + Bussi(kT_r, tau 0),
    dt 0.005, each phase one launch over all replicas, chunks captured as hipGraphs, pair launch
    from the project autotuner. Prints one JSON line; value = aggregate particle-steps/s."""
    from hoomd_amd import batch
    from hoomd_amd import _lib
    n, L, kT, seeds, pos = batch.seeded_replicas(nrep, seed)
    N = int(n.sum())
    log(f"[bench] {nrep} replicas, N total {N} (min {n.min()}, max {n.max()})")
    b = batch.ReplicaBatch(pos, L, kT, seeds)
    b.thermalize(0)
    rows = b.tune()
    for r in rows:
        log(f"[bench] autotuner tpp={int(r[0])} block={int(r[1])} {r[2] * 1e3:.1f} us")
    log(f"[bench] launch {b.config}; warmup {warmup} steps")
    b.run(warmup)
    primed = b.prime(steps)
    st0 = b.stats()
    _lib.check(_lib.lib().hm_synchronize())
    t0 = time.perf_counter()
    b.run(steps)
    _lib.check(_lib.lib().hm_synchronize())
    elapsed = time.perf_counter() - t0
    st1 = b.stats()
    tps = steps / elapsed
    rs = b.replica_stats()
    # the pair launch slows as particles drift from their cell order between sorts (every 100
    # steps): sample it at 10 points over one sort period, entries likewise
    ms, entries = 0.0, 0.0
    for _ in range(10):
        b.run(10)
        ms += b.time_pair(10) / 10
        entries += float(b.replica_stats()[:, 3].sum()) / 10
    Bp = 4.0 * entries + N * (4 + 32 + 16)
    achieved = Bp / (ms * 1e-3) / 1e9
    T = 2.0 * rs[:, 0] / (3.0 * n - 3.0)
    log(f"[bench] {tps:.1f} steps/s, {tps * N:.3e} particle-steps/s; pair {ms * 1e3:.1f} us")
    cpu = None
    if cpu_seconds > 0:
        import oracle as O
        from concurrent.futures import ThreadPoolExecutor
        # a bounded sample: the first 3 P replicas in seed order, one replica per host thread at
        # a time (independent replicas spread over the cores, as separate CPU runs would be),
        # each run 20 steps after its first; particle-steps/s over the wall time of the sample
        model, P = host_cpu()
        O.set_threads(1)
        sample = np.arange(min(nrep, 3 * P))
        systems = []
        for r in sample[np.argsort(-n[sample], kind="stable")]:  # largest first: less idle tail
            v = O.thermalize(np.ones(n[r]), np.arange(n[r]), kT[r], 0, int(seeds[r]))
            s = O.System(pos[r], v, L[r], dt=0.005, seed=int(seeds[r]))
            s.set_bussi(float(kT[r]), 0.0)
            s.run(1)
            systems.append(s)
        k = 20
        t1 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=P) as ex:
            list(ex.map(lambda s: s.run(k), systems))
        spent = time.perf_counter() - t1
        done_ps = float(k * n[sample].sum())
        cpu = {"value": done_ps / spent, "unit": "particle-steps/s", "cores": P, "kind": "port",
               "cpu_model": model,
               "sample": f"the oracle's FP64 step loop on the first {len(sample)} replicas "
                         f"(N_total {int(n[sample].sum())}), 20 steps each after the first, one "
                         f"replica per thread on {P} host threads, {spent:.1f} s"}
    print(json.dumps({
        "metric": "particle-steps/s (batched uneven LJ replicas)", "value": tps * N,
        "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True, "timesteps_per_s": tps,
        "dtype": "f32 pair arithmetic, f64 positions/integration",
        "data": "synthetic (seeded replicas on random FCC sites, Philox thermalized)",
        "config": {"workload": f"{nrep} LJ replicas, N log-uniform [500, 20000], rho U[0.70, "
                               "0.90], kT U[0.8, 1.5], r_cut 2.5, buffer 0.4, Bussi tau 0, "
                               "dt 0.005", "replicas": nrep, "N_total": N,
                   "launch": b.config, "graphs_primed": primed,
                   "list_builds_in_timed_steps": st1["builds"] - st0["builds"],
                   "overflow_replays": st1["replays"]},
        "replica_kT_rel_dev_max": float(np.abs(T / kT - 1).max()),
        "roofline": {"bound": "hbm", "kernel": "bpair_kernel", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "avg_launch_us": ms * 1e3, "algorithmic_bytes": Bp,
                     "nlist_entries": int(entries)},
        "cpu_baseline": cpu}), flush=True)
    return 0


MLP_SITE_FLOP = 2 * (16 * 64 + 64 * 64) + 2 * 64 + 2 * (64 * 64 + 64 * 16)  # forward + dE/dG


def mlcg(n_chains, steps, warmup, seed, nframes, frame_every, train_steps, pushoff, lr=2e-3,
         frames_per_step=1, ibi_iters=12, cg_dt=0.02, pressure_match=False, cg_buffer=0.4):
    """Learned CG of the Kremer-Grest melt, end to end on the GPU, against the plain run:

    1. fine-grained melt: n_chains x 50 random-walk chains (bond 0.97, rho 0.85) pushed off with
       the ramped soft cosine (pushoff steps), then WCA + FENE + Langevin (kT 1, gamma 1,
       dt 0.01); its production rate is timed (the plain run);
    2. nframes frames, one every frame_every steps, mapped 5:1 on the device (site centres of
       mass and summed forces);
    3. harmonic CG bonds by Boltzmann inversion of the mapped bond lengths; targets = mapped
       force - CG bond force (device);
    4. the MLP site energy (16 Gaussian densities, 16 -> 64 -> 64 -> 1 SiLU, bf16 MFMA) trained
       by force matching with the project's Adam kernels, train_steps steps over the frames;
    5. the CG system (sites of frame 0, CG bonds, the MLP; Langevin kT 1, dt cg_dt, site mass 5,
       gamma 5) timed, then sampled for RDF, bond-length distribution and pressure against the
       mapped frames / the plain run;
    6. structure refinement: ibi_iters iterative Boltzmann inversion steps on the pair prior
       (a Table out to the MLP cut-off) with the network in place, then the same measurements;
       the comparison model without the network gets the same refinement from the same prior.

    Prints one JSON line: value = fine-grained-equivalent simulated time per wall-second of the
    CG run over that of the plain run (tau_FG advanced per second)."""
    import ctypes as Cc

    import hoomd_amd as hoomd
    from hoomd_amd import _lib
    from hoomd_amd.md import mlcg as MC
    from kg_workload import pushed_off_melt
    t_all = time.perf_counter()
    chain_len, per = 50, 5
    log(f"[mlcg] melt {n_chains} x {chain_len}: push-off {pushoff} steps")
    sim, L = pushed_off_melt(hoomd, n_chains, chain_len, pushoff_steps=pushoff,
                             ramp_updates=100, seed=seed)
    N = sim.state.N_particles
    sim.run(warmup)
    thermo_fg = hoomd.md.compute.ThermodynamicQuantities(hoomd.filter.All())
    sim.operations.add(thermo_fg)
    # plain run rate
    _lib.check(_lib.lib().hm_sys_prime(sim._system, steps), "prime")
    _lib.check(_lib.lib().hm_synchronize())
    t0 = time.perf_counter()
    sim.run(steps)
    _lib.check(_lib.lib().hm_synchronize())
    tps_fg = steps / (time.perf_counter() - t0)
    log(f"[mlcg] plain KG: {tps_fg:.1f} steps/s at N={N}")
    # frames
    smap = MC.SiteMap.linear_chains(n_chains, chain_len, per)
    smap.attach(sim)
    fs = MC.FrameSet(nframes, smap.nsites)
    p_fg = []
    for k in range(nframes):
        sim.run(frame_every)
        fs.record(sim, k)
        if k % max(1, nframes // 20) == 0:
            p_fg.append(thermo_fg.pressure)
    log(f"[mlcg] {nframes} frames of {smap.nsites} sites recorded")
    cgb = smap.bonds()
    sample = np.linspace(0, nframes - 1, min(nframes, 20)).astype(int)
    r_map = np.concatenate([MC.bond_lengths(fs.positions(k).astype(np.float64), L, cgb)
                            for k in sample])
    kb, r0 = MC.boltzmann_harmonic(r_map, kT=1.0)
    log(f"[mlcg] CG bonds by Boltzmann inversion: k = {kb:.3f}, r0 = {r0:.3f}")
    rdf_edges_max, nb = 4.0, 80
    g_map = np.mean([MC.rdf(fs.positions(k), L, rdf_edges_max, nb, exclude_bonded=cgb)[1]
                     for k in sample], axis=0)
    r_bins = (np.arange(nb) + 0.5) * rdf_edges_max / nb
    prior_params, rc_prior = MC.rdf_prior_table(r_bins, g_map)
    log(f"[mlcg] RDF prior: cut-off {rc_prior:.3f}")
    fs.subtract_harmonic(cgb, kb, r0, L)
    # CG system
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = smap.nsites
    snap.particles.types = ["S"]
    snap.particles.position[:] = fs.positions(0)
    snap.particles.mass[:] = float(per)
    snap.bonds.N = len(cgb)
    snap.bonds.types = ["cg"]
    snap.bonds.group[:] = cgb
    cg = hoomd.Simulation(device=hoomd.device.GPU(), seed=seed + 1)
    cg.create_state_from_snapshot(snap)
    nl = hoomd.md.nlist.Cell(buffer=cg_buffer, exclusions=("bond",))
    mlp = hoomd.md.mlcg.SiteEnergyMLP(nlist=nl, seed=seed)
    harm = hoomd.md.bond.Harmonic()
    harm.params["cg"] = dict(k=kb, r0=r0)
    # the pair prior, out to the MLP cut-off, evaluated inside the site energy's force pass
    prior_params = MC.regrid_table(prior_params, rc_prior, mlp.r_cut, width=600)
    mlp.pair_table = prior_params
    lang = hoomd.md.methods.Langevin(filter=hoomd.filter.All(), kT=1.0)
    lang.gamma["S"] = float(per)
    cg.operations.integrator = hoomd.md.Integrator(dt=cg_dt, forces=[mlp, harm],
                                                   methods=[lang])
    cg.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.0)
    G = np.concatenate([mlp.descriptors(fs, int(k)) for k in sample[:5]])
    mlp.fit_normalization(G)
    ntrain = max(1, int(nframes * 0.9))
    held = np.arange(ntrain, nframes)
    zero = np.zeros(MC.NPARAM, np.float32)
    mlp.params = zero
    base = float(np.mean([mlp.loss_grad(fs, int(k))[0] for k in held[:10]]))
    mlp.params = MC.init_params(seed)
    order = np.random.default_rng(seed).integers(0, ntrain, size=train_steps * frames_per_step)
    t1 = time.perf_counter()
    loss = mlp.train(fs, train_steps, lr=lr, order=order, frames_per_step=frames_per_step)
    _lib.check(_lib.lib().hm_synchronize())
    t_train = time.perf_counter() - t1
    trained = float(np.mean([mlp.loss_grad(fs, int(k))[0] for k in held[:10]]))
    log(f"[mlcg] trained {train_steps} steps in {t_train:.1f} s: held-out force residual "
        f"{trained:.4f} vs bonds-only {base:.4f}")
    # training leaves the CG system at the sites of its last frame
    cg.run(warmup)
    thermo = hoomd.md.compute.ThermodynamicQuantities(hoomd.filter.All())
    cg.operations.add(thermo)
    _lib.check(_lib.lib().hm_sys_prime(cg._system, steps), "prime")
    _lib.check(_lib.lib().hm_synchronize())
    t0 = time.perf_counter()
    cg.run(steps)
    _lib.check(_lib.lib().hm_synchronize())
    tps_cg = steps / (time.perf_counter() - t0)
    log(f"[mlcg] CG run: {tps_cg:.1f} steps/s at {smap.nsites} sites")
    # structure of the CG run, with the network and without it (bonds + prior)
    p_fg_mean = float(np.mean(p_fg))

    def ibi(params, iters, relax=1000, frames=2, every=200):
        """Iterative Boltzmann inversion of the pair prior against the mapped RDF (with
        pressure_match, plus the linear pressure correction toward the plain run's pressure),
        and of the harmonic CG bond
        against the mapped bond-length mean and variance (r0 moves by the difference of the
        means, k scales by the square root of the variance ratio)."""
        errs = []
        k_b, r_b = kb, r0
        m_map, v_map = float(r_map.mean()), float(r_map.var())
        for _ in range(iters):
            cg.run(relax)
            gs, rs, ps = [], [], []
            for _ in range(frames):
                cg.run(every)
                pc = cg.state.get_snapshot().particles.position
                gs.append(MC.rdf(pc, L, rdf_edges_max, nb, exclude_bonded=cgb)[1])
                rs.append(MC.bond_lengths(pc, L, cgb))
                ps.append(thermo.pressure)
            g_it = np.mean(gs, axis=0)
            rb = np.concatenate(rs)
            errs.append(float(np.abs(g_it - g_map).sum() / g_map.sum()))
            params = MC.ibi_update(params, mlp.r_cut, r_bins, g_it, g_map, kT=1.0)
            if pressure_match:
                params = MC.pressure_ramp(params, mlp.r_cut, float(np.mean(ps)) - p_fg_mean)
            mlp.pair_table = params
            r_b += m_map - float(rb.mean())
            k_b *= float(np.sqrt(rb.var() / v_map))
            harm.params["cg"] = dict(k=k_b, r0=r_b)
        return params, errs, (k_b, r_b)

    kT_cg = []

    def structure(nsamp=10, every=250):
        gs, ps, rs = [], [], []
        for k in range(nsamp):
            cg.run(every)
            pc = cg.state.get_snapshot().particles.position
            gs.append(MC.rdf(pc, L, rdf_edges_max, nb, exclude_bonded=cgb)[1])
            rs.append(MC.bond_lengths(pc, L, cgb))
            ps.append(thermo.pressure)
            kT_cg.append(thermo.kinetic_temperature)
        return np.mean(gs, axis=0), np.mean(ps), np.concatenate(rs)

    g_cg, p_cg_mean, r_cg = structure()
    t_ibi = time.perf_counter()
    ibi_table, ibi_errs, ibi_bond = ibi(prior_params, ibi_iters)
    g_ibi, p_ibi, r_ibi = structure()
    log(f"[mlcg] IBI {ibi_iters} steps with the network ({time.perf_counter() - t_ibi:.1f} s): "
        f"RDF L1 {' '.join(f'{e:.3f}' for e in ibi_errs)}; bond k {ibi_bond[0]:.3f} r0 "
        f"{ibi_bond[1]:.3f}")
    # the refined model's rate (the same kernels; the table is updated in place, the bond
    # refinement rebuilt the system before the structure sampling above); every graph the timed
    # run launches is captured first
    cg.run(0)
    _lib.check(_lib.lib().hm_sys_prime(cg._system, steps), "prime")
    _lib.check(_lib.lib().hm_synchronize())
    t0 = time.perf_counter()
    cg.run(steps)
    _lib.check(_lib.lib().hm_synchronize())
    tps_cg_ibi = steps / (time.perf_counter() - t0)
    builds_cg = nl.num_builds  # builds of the timed run
    log(f"[mlcg] final model timed: {tps_cg_ibi:.1f} steps/s, {builds_cg} list builds in {steps} "
        f"steps (buffer {cg_buffer})")
    nl_out = np.zeros(3, np.int32)
    nl_log = np.zeros((32, 3), np.float32)
    rows = _lib.lib().hm_sys_nlist_launch(cg._system, nl_out.ctypes.data, nl_log.ctypes.data, 32)
    for r in nl_log[:max(rows, 0)]:
        log(f"[mlcg] CG list-build tuner shape={int(r[0])} grid={int(r[1])} {r[2] * 1e3:.1f} us")
    cg_nl_launch = {"shape": int(nl_out[0]), "grid": int(nl_out[1])}
    trained_params = mlp.params
    mlp.params = np.zeros(MC.NPARAM, np.float32)
    mlp.pair_table = prior_params
    harm.params["cg"] = dict(k=kb, r0=r0)
    cg.run(2000)
    g_none, p_none, r_none = structure()
    _, ibi_errs_none, ibi_bond_none = ibi(prior_params, ibi_iters)
    g_none_ibi, p_none_ibi, r_none_ibi = structure()
    mlp.params = trained_params
    dr = rdf_edges_max / nb
    rdf_l1 = float(np.abs(g_cg - g_map).sum() / g_map.sum())
    rdf_l2 = float(np.sqrt(((g_cg - g_map) ** 2).sum() / (g_map ** 2).sum()))
    rdf_l1_none = float(np.abs(g_none - g_map).sum() / g_map.sum())
    hb_e = np.linspace(min(r_map.min(), r_cg.min()), max(r_map.max(), r_cg.max()), 61)
    hm_, _ = np.histogram(r_map, hb_e)
    hc_, _ = np.histogram(r_cg, hb_e)
    bond_tv = float(0.5 * np.abs(hm_ / hm_.sum() - hc_ / hc_.sum()).sum())
    def tv(rs):
        h, _ = np.histogram(np.clip(rs, hb_e[0], hb_e[-1]), hb_e)
        return float(0.5 * np.abs(hm_ / hm_.sum() - h / h.sum()).sum())

    def l1(g):
        return float(np.abs(g - g_map).sum() / g_map.sum())

    bond_tv_none = tv(r_none)
    ms3 = (Cc.c_double * 3)()
    _lib.check(_lib.lib().hm_sys_time_mlp(cg._system, mlp._force_id, 20, ms3), "time mlp")
    site_flops = MLP_SITE_FLOP * smap.nsites
    mfma = site_flops / (ms3[1] * 1e-3) / 2.5e15
    speedup = (tps_cg_ibi * cg_dt) / (tps_fg * 0.01)
    log(f"[mlcg] final model: speed-up {speedup:.2f}x, RDF L1 error {l1(g_ibi):.4f}, bond TV "
        f"{tv(r_ibi):.4f}, pressure {p_ibi:.3f} (plain run {p_fg_mean:.3f}); without the "
        f"network: RDF {l1(g_none_ibi):.4f}, bond TV {tv(r_none_ibi):.4f}")
    print(json.dumps({
        "metric": "fine-grained-equivalent simulated time per wall-second, MLP-CG / plain KG",
        "value": speedup, "unit": "x", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "higher_is_better": True,
        "plain": {"timesteps_per_s": tps_fg, "particles": N, "dt": 0.01,
                  "tau_per_s": tps_fg * 0.01, "pressure": float(np.mean(p_fg))},
        "cg": {"timesteps_per_s": tps_cg_ibi, "sites": smap.nsites, "dt": cg_dt,
               "nlist_buffer": cg_buffer, "list_builds_in_timed_steps": builds_cg,
               "nlist_launch": cg_nl_launch,
               "tau_per_s": tps_cg_ibi * cg_dt, "pressure": float(p_ibi),
               "model": "trained site energy + refined pair prior (in the force pass) + "
                        "refined harmonic bonds",
               "kinetic_temperature": float(np.mean(kT_cg)),
               "kernel_ms": {"densities": ms3[0], "site_mlp": ms3[1], "forces": ms3[2]},
               "mfma": {"site_flop": MLP_SITE_FLOP, "achieved_tflops": site_flops /
                        (ms3[1] * 1e-3) / 1e12, "peak_tflops_bf16": 2500.0,
                        "utilisation": mfma}},
        "training": {"frames": nframes, "frame_every": frame_every, "steps": train_steps,
                     "lr": lr, "frames_per_step": frames_per_step,
                     "seconds": t_train, "loss_first50": float(loss[:50].mean()),
                     "loss_last50": float(loss[-50:].mean()),
                     "heldout_residual_bonds_only": base, "heldout_residual_trained": trained},
        "cg_bonds": {"k": kb, "r0": r0},
        "errors": {"rdf_rel_l1": l1(g_ibi),
                   "rdf_rel_l2": float(np.sqrt(((g_ibi - g_map) ** 2).sum()
                                               / (g_map ** 2).sum())),
                   "bond_length_tv": tv(r_ibi),
                   "pressure_rel": float(abs(p_ibi - p_fg_mean) / abs(p_fg_mean))},
        "without_network": {"rdf_rel_l1": l1(g_none_ibi), "bond_length_tv": tv(r_none_ibi),
                            "pressure_rel": float(abs(p_none_ibi - p_fg_mean)
                                                  / abs(p_fg_mean)),
                            "bond_k_r0": list(ibi_bond_none),
                            "refinement_rdf_l1_per_iteration": ibi_errs_none},
        "refinement": {"iterations": ibi_iters, "relax_steps": 1000, "frames_per_iteration": 2,
                       "pressure_match": bool(pressure_match),
                       "bond_k_r0": list(ibi_bond), "rdf_l1_per_iteration": ibi_errs},
        "force_matched_only": {
            "timesteps_per_s": tps_cg, "speedup": (tps_cg * cg_dt) / (tps_fg * 0.01),
            "with_network": {"rdf_rel_l1": rdf_l1, "rdf_rel_l2": rdf_l2,
                             "bond_length_tv": bond_tv,
                             "pressure_rel": float(abs(p_cg_mean - p_fg_mean) / abs(p_fg_mean))},
            "without_network": {"rdf_rel_l1": rdf_l1_none, "bond_length_tv": bond_tv_none,
                                "pressure_rel": float(abs(p_none - p_fg_mean)
                                                      / abs(p_fg_mean))}},
        "prior": {"r_cut": rc_prior},
        "rdf": {"r": [float((i + 0.5) * dr) for i in range(nb)],
                "mapped": g_map.tolist(), "cg": g_cg.tolist(), "cg_without_network":
                g_none.tolist(), "cg_ibi": g_ibi.tolist(),
                "cg_without_network_ibi": g_none_ibi.tolist()},
        "dtype": "bf16 MFMA (fp32 accumulate) MLP, f32 densities/forces, f64 integration",
        "data": "synthetic (seeded random-walk melt, pushed off)",
        "config": {"workload": f"KG melt {n_chains} x {chain_len} -> 5:1 CG "
                               f"({smap.nsites} sites), {nframes} frames every {frame_every} "
                               "steps, MLP 16-64-64-1", "wall_s": time.perf_counter() - t_all},
    }), flush=True)
    return 0


def cg_step(n_chains, steps, warmup, seed, mlp_only=0):
    """Profiling mode of the learned-CG step alone: n_chains x 10 CG sites as random walks of
    step 1.7 at site density 0.17 (the 5:1-mapped melt's), the site energy with random
    parameters (init_params(seed) x 0.05) and a repulsive pair table in its force pass, harmonic
    bonds k 2.3 r0 1.74, Langevin kT 1 gamma 5, site mass 5, dt 0.02; relaxed for warmup steps,
    then steps timed. Prints one JSON line (timesteps/s and the three site-energy kernels).
    With mlp_only > 0 it launches only the three site-energy kernels that many times each after
    the relaxation (outside any graph, for counter passes) and exits."""
    import ctypes as Cc

    import hoomd_amd as hoomd
    from hoomd_amd import _lib
    from hoomd_amd.md import mlcg as MC
    per = 10
    N = n_chains * per
    L = (N / 0.17) ** (1 / 3)
    rng = np.random.default_rng(seed)
    start = rng.uniform(-L / 2, L / 2, (n_chains, 1, 3))
    stp = rng.normal(size=(n_chains, per, 3))
    stp /= np.linalg.norm(stp, axis=2, keepdims=True)
    pos = ((start + 1.7 * np.cumsum(stp, axis=1)).reshape(-1, 3) + L / 2) % L - L / 2
    idx = np.arange(N).reshape(n_chains, per)
    bonds = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1).astype(np.uint32)
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = N
    snap.particles.types = ["S"]
    snap.particles.position[:] = pos
    snap.particles.mass[:] = 5.0
    snap.bonds.N = len(bonds)
    snap.bonds.types = ["cg"]
    snap.bonds.group[:] = bonds
    cg = hoomd.Simulation(device=hoomd.device.GPU(), seed=seed)
    cg.create_state_from_snapshot(snap)
    nl = hoomd.md.nlist.Cell(buffer=0.4, exclusions=("bond",))
    mlp = hoomd.md.mlcg.SiteEnergyMLP(nlist=nl, seed=seed)
    r = (np.arange(80) + 0.5) * 0.05
    prior, rc = MC.rdf_prior_table(r, np.clip((r - 0.3) / 1.5, 0.0, 1.0))
    mlp.pair_table = MC.regrid_table(prior, rc, mlp.r_cut, width=600)
    harm = hoomd.md.bond.Harmonic()
    harm.params["cg"] = dict(k=2.3, r0=1.74)
    lang = hoomd.md.methods.Langevin(filter=hoomd.filter.All(), kT=1.0)
    lang.gamma["S"] = 5.0
    cg.operations.integrator = hoomd.md.Integrator(dt=0.02, forces=[mlp, harm], methods=[lang])
    cg.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.0)
    mlp.params = MC.init_params(seed) * 0.05
    log(f"[cg-step] {N} sites, L {L:.3f}: relax {warmup} steps")
    cg.run(warmup)
    if mlp_only:
        ms3 = (Cc.c_double * 3)()
        _lib.check(_lib.lib().hm_sys_time_mlp(cg._system, mlp._force_id, mlp_only, ms3),
                   "time mlp")
        log(f"[cg-step] {mlp_only} launches each: densities {ms3[0] * 1e3:.1f} us, site MLP "
            f"{ms3[1] * 1e3:.1f} us, forces {ms3[2] * 1e3:.1f} us")
        return 0
    _lib.check(_lib.lib().hm_sys_prime(cg._system, steps), "prime")
    _lib.check(_lib.lib().hm_synchronize())
    t0 = time.perf_counter()
    cg.run(steps)
    _lib.check(_lib.lib().hm_synchronize())
    el = time.perf_counter() - t0
    builds = nl.num_builds  # builds of the last run
    ms3 = (Cc.c_double * 3)()
    _lib.check(_lib.lib().hm_sys_time_mlp(cg._system, mlp._force_id, 20, ms3), "time mlp")
    site_flops = MLP_SITE_FLOP * N
    log(f"[cg-step] {steps / el:.1f} steps/s; densities {ms3[0] * 1e3:.1f} us, site MLP "
        f"{ms3[1] * 1e3:.1f} us, forces {ms3[2] * 1e3:.1f} us")
    print(json.dumps({
        "metric": "timesteps/s (learned-CG step, synthetic state)", "value": steps / el,
        "unit": "timesteps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * el / steps, "higher_is_better": True,
        "kernel_ms": {"densities": ms3[0], "site_mlp": ms3[1], "forces": ms3[2]},
        "mfma": {"site_flop": MLP_SITE_FLOP, "achieved_tflops": site_flops / (ms3[1] * 1e-3)
                 / 1e12, "peak_tflops_bf16": 2500.0,
                 "utilisation": site_flops / (ms3[1] * 1e-3) / 2.5e15},
        "dtype": "bf16 MFMA (fp32 accumulate) MLP, f32 densities/forces, f64 integration",
        "data": "synthetic (random-walk CG chains, random network parameters)",
        "config": {"workload": f"learned-CG step, {N} sites ({n_chains} x {per}), MLP "
                               "16-64-64-1 + pair table in the force pass, harmonic bonds, "
                               "Langevin dt 0.02", "list_builds_in_timed_steps": builds}}),
          flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--n", type=int, default=1048576, help="particles (4 * n_cells^3)")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the FP64 re-evaluation of the final state (LJ workload)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pair-only", type=int, default=0,
                    help="profiling mode: after warmup launch only the pair kernel this many "
                         "times (for rocprofv3 counter passes) and exit")
    ap.add_argument("--build-only", type=int, default=0,
                    help="profiling mode: after warmup run this many full cell + list rebuilds "
                         "outside any graph (for rocprofv3 counter passes) and exit")
    ap.add_argument("--pair-kind", type=int, default=2, choices=(0, 1, 2, 3),
                    help="pair kernel: 0 FP64 cell tiles, 1 FP64 gather, 2 FP32 cell-relative "
                         "gather, 3 FP32 cell tiles")
    ap.add_argument("--pair-tpp", type=int, default=0, help="pair kernel lanes per particle")
    ap.add_argument("--pair-block", type=int, default=0, help="pair kernel block size")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="profiling mode: device-to-device copy of this many MiB (the known-size "
                         "kernel that scales FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--howto", type=int, default=0, metavar="R",
                    help="run the reference's how-to device benchmark instead: spheres.gsd "
                         "replicated R x R x R (1: N = 2048, 4: N = 131072)")
    ap.add_argument("--kg", action="store_true",
                    help="run the Kremer-Grest melt (10^6 beads, Langevin) instead")
    ap.add_argument("--mlcg", type=int, default=0, metavar="CHAINS",
                    help="run the learned-CG workflow on a melt of CHAINS x 50 beads instead "
                         "(20000 for the full 10^6-bead melt)")
    ap.add_argument("--frames", type=int, default=2000)
    ap.add_argument("--frame-every", type=int, default=20)
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--pushoff", type=int, default=10000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--frames-per-step", type=int, default=1)
    ap.add_argument("--ibi-iters", type=int, default=12)
    ap.add_argument("--cg-dt", type=float, default=0.02,
                    help="timestep of the CG run (the learned-CG configuration's 0.02; 0.04 "
                         "gives the same structure and temperature)")
    ap.add_argument("--cg-buffer", type=float, default=0.4,
                    help="neighbour-list buffer of the CG run (site cut-off 3.0)")
    ap.add_argument("--pressure-match", action="store_true",
                    help="add the linear pressure correction to each refinement step")
    ap.add_argument("--cg-step", type=int, default=0, metavar="CHAINS",
                    help="profiling mode: the learned-CG step alone on CHAINS x 10 synthetic "
                         "sites (20000 for the 200 000-site system)")
    ap.add_argument("--mlp-only", type=int, default=0,
                    help="with --cg-step: only launch the site-energy kernels this many times "
                         "each after the relaxation (for rocprofv3 counter passes)")
    ap.add_argument("--replicas", type=int, default=0, metavar="R",
                    help="run R batched uneven LJ replicas instead (256 for the batched workload)")
    a = ap.parse_args()
    if a.mlcg:
        if a.gpus != 1:
            print("bench.py runs on exactly one GPU", file=sys.stderr)
            return 2
        return mlcg(a.mlcg, a.steps, a.warmup, a.seed, a.frames, a.frame_every, a.train_steps,
                    a.pushoff, a.lr, a.frames_per_step, a.ibi_iters, a.cg_dt,
                    a.pressure_match, a.cg_buffer)
    if a.cg_step:
        if a.gpus != 1:
            print("bench.py runs on exactly one GPU", file=sys.stderr)
            return 2
        return cg_step(a.cg_step, a.steps, a.warmup, a.seed, a.mlp_only)
    if a.replicas:
        if a.gpus != 1:
            print("bench.py runs on exactly one GPU", file=sys.stderr)
            return 2
        return replicas(a.replicas, a.steps, a.warmup, a.seed, 0 if a.no_cpu else 10.0)
    if a.kg:
        if a.gpus != 1:
            print("bench.py runs on exactly one GPU", file=sys.stderr)
            return 2
        return kg(a.steps, a.warmup, a.seed, 0 if a.no_cpu else a.cpu_steps, a.pushoff)
    if a.howto:
        if a.gpus != 1:
            print("bench.py runs on exactly one GPU", file=sys.stderr)
            return 2
        return howto(a.howto, a.steps, a.warmup, a.seed, 0 if a.no_cpu else 8.0)
    if a.copy_calibration:
        from hoomd_amd import _lib
        nb = a.copy_calibration << 20
        src = _lib.DeviceArray((nb,), "u1", fill=1)
        dst = _lib.DeviceArray((nb,), "u1", fill=0)
        for _ in range(3):
            _lib.check(_lib.lib().hm_memcpy_dtod(dst.ptr, src.ptr, nb))
        _lib.check(_lib.lib().hm_synchronize())
        log(f"[bench] copied {a.copy_calibration} MiB x3")
        return 0
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(no domain decomposition); run with --gpus 1", file=sys.stderr)
        return 2
    n_cells = int(round((a.n / 4) ** (1 / 3)))
    if 4 * n_cells**3 != a.n:
        print("--n must be 4 * k^3 (FCC)", file=sys.stderr)
        return 2
    rho = 0.8442

    import hoomd_amd as hoomd
    from hoomd_amd import _lib

    log(f"[bench] building N={a.n} LJ system")
    sim, lj, pos, L = build_sim(hoomd, n_cells, rho, a.seed)
    N = len(pos)
    sim.run(0)
    _lib.check(_lib.lib().hm_sys_set_pair_kernel(sim._system, a.pair_kind, a.pair_tpp,
                                                 a.pair_block))
    log(f"[bench] warmup {a.warmup} steps")
    sim.run(a.warmup)
    _lib.check(_lib.lib().hm_synchronize())
    if a.pair_only:
        import ctypes as C
        ms = C.c_double()
        _lib.check(_lib.lib().hm_sys_time_pair(sim._system, lj._force_id, a.pair_only,
                                               C.byref(ms), None))
        log(f"[bench] {a.pair_only} pair launches, {ms.value * 1e3:.1f} us each")
        return 0
    if a.build_only:
        import ctypes as C
        t2 = (C.c_double * 2)()
        _lib.check(_lib.lib().hm_sys_time_build(sim._system, a.build_only, t2), "time build")
        log(f"[bench] {a.build_only} full rebuilds: cells {t2[0] * 1e3:.1f} us, list "
            f"{t2[1] * 1e3:.1f} us each")
        return 0
    pcfg = np.zeros(5, np.uint32)
    _lib.check(_lib.lib().hm_sys_pair_config(sim._system, pcfg.ctypes.data))
    launch = {"kind": int(pcfg[0]), "tpp": int(pcfg[1]), "block": int(pcfg[2]),
              "batch": int(pcfg[4]), "autotuned": bool(pcfg[3])}
    log(f"[bench] pair launch {launch}")
    if launch["autotuned"]:
        tl = np.zeros((64, 4), np.float32)
        rows = _lib.lib().hm_sys_tune_log(sim._system, tl.ctypes.data, 64)
        for r in tl[:max(rows, 0)]:
            log(f"[bench] autotuner tpp={int(r[0])} block={int(r[1])} batch={int(r[2])} "
                f"{r[3] * 1e3:.1f} us")
    st0 = np.zeros(9, np.uint64)
    _lib.check(_lib.lib().hm_sys_stats(sim._system, st0.ctypes.data))
    # every hipGraph the timed run launches is captured and instantiated here, so the timed
    # region holds no capture
    primed = _lib.lib().hm_sys_prime(sim._system, a.steps)
    _lib.check(primed, "prime graphs")
    log(f"[bench] {primed} step graphs primed; timing {a.steps} steps")
    _lib.check(_lib.lib().hm_synchronize())
    t0 = time.perf_counter()
    sim.run(a.steps)
    _lib.check(_lib.lib().hm_synchronize())
    elapsed = time.perf_counter() - t0
    st1 = np.zeros(9, np.uint64)
    _lib.check(_lib.lib().hm_sys_stats(sim._system, st1.ctypes.data))
    tps = a.steps / elapsed
    builds = int(st1[0] - st0[0])
    log(f"[bench] {tps:.1f} steps/s, {builds} neighbour-list builds in the timed steps")

    ms, entries, B, achieved = pair_roofline(sim, lj, N, a.pair_kind)
    log(f"[bench] pair kernel {ms * 1e3:.1f} us, {B / 1e9:.3f} GB algorithmic, {achieved:.0f} GB/s")
    traffic = load_traffic(N)
    nlp = nlist_pair_fraction(sim, N, entries, ms, a.steps, builds)
    log(f"[bench] nlist+pair: cell {nlp['t_cell_us']:.1f} us + list {nlp['t_nlist_us']:.1f} us "
        f"per build, k = {nlp['k']:.2f} steps per build -> {nlp['achieved']:.0f} GB/s "
        f"({nlp['frac']:.3f} of peak)")
    nl_out = np.zeros(3, np.int32)
    nl_log = np.zeros((32, 3), np.float32)
    rows = _lib.lib().hm_sys_nlist_launch(sim._system, nl_out.ctypes.data, nl_log.ctypes.data, 32)
    for r in nl_log[:max(rows, 0)]:
        log(f"[bench] list-build tuner shape={int(r[0])} grid={int(r[1])} {r[2] * 1e3:.1f} us")
    launch["nlist_shape"], launch["nlist_grid"] = int(nl_out[0]), int(nl_out[1])

    parity = None if a.no_parity else parity_block(sim, lj, L)
    cpu = None
    if not a.no_cpu:
        model, P = host_cpu()
        log(f"[bench] CPU baseline: oracle, {a.cpu_steps} steps of the same workload, {P} threads")
        cps, cdt, cbuilds, csteps = cpu_baseline(n_cells, rho, a.seed, a.cpu_steps)
        cpu = {"value": cps, "unit": "timesteps/s", "cores": P, "kind": "port",
               "cpu_model": model,
               "sample": f"N={N} LJ Bussi workload, {csteps} steps after the first four "
                         f"(includes {cbuilds} list rebuilds), {cdt:.1f} s on {P} host threads "
                         f"({model}); {cps * N:.3e} particle-steps/s",
               "particle_steps_per_s": cps * N}
        log(f"[bench] CPU {cps:.3f} steps/s")

    out = {
        "metric": METRIC,
        "value": tps,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 positions/integration",
        "data": "synthetic (FCC lattice, Philox Maxwell-Boltzmann velocities)",
        "particle_steps_per_s": tps * N,
        "config": {"workload": f"LJ liquid N={N} (FCC {n_cells}^3x4), rho=0.8442, r_cut=2.5, "
                               "buffer=0.4, Bussi kT=1.2 tau=0, dt=0.005, Hilbert sort every 500",
                   "N": N, "box_L": L, "nlist_builds_in_timed_steps": builds,
                   "pair_launch": launch, "timed_region_captures": 0,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": PAIR_KERNELS[a.pair_kind][0],
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_us": ms * 1e3, "algorithmic_bytes": B,
                     "nlist_entries": int(entries)},
        "roofline_nlist_pair": nlp,
        "parity": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
