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


def load_traffic(N):
    """HBM bytes per pair launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled
    by a copy-kernel calibration), committed under profiles/ for this workload."""
    path = os.path.join(ROOT, "profiles", "pair_lj_1M_traffic.json")
    if N != 1048576 or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(n_cells, rho, seed, steps):
    """Oracle (FP64 C restatement of the reference CPU path: half list, 3rd law, Bussi) on one
    host core, on the same 1M-particle state for a bounded number of steps."""
    import oracle as O
    from lattices import fcc_lattice
    pos, L = fcc_lattice(n_cells, rho)
    N = len(pos)
    vel = O.thermalize(np.ones(N), np.arange(N), 1.2, 0, seed)
    s = O.System(pos, vel, L, r_cut=2.5, mode="none", buffer=0.4, dt=0.005, seed=seed)
    s.set_bussi(1.2, 0.0)
    s.set_sorter(500, 256)
    s.run(1)  # initial build, first step and sort outside the timed sample
    b0 = s.n_builds
    t0 = time.perf_counter()
    s.run(steps)
    dt = time.perf_counter() - t0
    return steps / dt, dt, int(s.n_builds - b0)


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


def kg_melt(nchain=20000, length=50, rho=0.85):
    """Kremer-Grest melt without overlaps: beads on a simple-cubic lattice at density rho, chains
    of `length` consecutive sites along z rows (rows of n sites hold n / length chains), so
    every bond starts at the lattice spacing (1.056 at rho 0.85) and no two beads overlap."""
    N = nchain * length
    n = int(round(N ** (1 / 3)))
    if n ** 3 != N or n % length:
        raise ValueError("nchain * length must be a cube whose edge is a multiple of length")
    a = rho ** (-1 / 3)
    L = n * a
    g = np.arange(n)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")  # x varies slowest: index = (x n + y) n + z
    # chains run along z (consecutive indices), length beads each
    pos = (np.stack([x, y, z], -1).reshape(-1, 3) + 0.5) * a - L / 2
    idx = np.arange(N).reshape(-1, length)
    bonds = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1).astype(np.uint32)
    return pos, bonds, L


def kg(steps, warmup, seed, cpu_steps):
    """Kremer-Grest melt production step: 20 000 chains x 50 beads at rho 0.85, FENE (k 30,
    r0 1.5) + WCA bonds, WCA pairs (r_cut 2^(1/6), mode shift, bond exclusions, buffer 0.4),
    Langevin kT 1, gamma 1, dt 0.01. Prints one JSON line."""
    import hoomd_amd as hoomd
    from hoomd_amd import _lib
    pos, bonds, L = kg_melt()
    N = len(pos)
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = N
    snap.particles.types = ["A"]
    snap.particles.position[:] = pos
    snap.bonds.N = len(bonds)
    snap.bonds.types = ["A-A"]
    snap.bonds.group[:] = bonds
    sim = hoomd.Simulation(device=hoomd.device.GPU(), seed=seed)
    sim.create_state_from_snapshot(snap)
    sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.0)
    nl = hoomd.md.nlist.Cell(buffer=0.4, exclusions=("bond",))
    wca = hoomd.md.pair.LJ(nlist=nl, default_r_cut=2 ** (1 / 6), mode="shift")
    wca.params[("A", "A")] = dict(epsilon=1.0, sigma=1.0)
    fene = hoomd.md.bond.FENEWCA()
    fene.params["A-A"] = dict(k=30.0, r0=1.5, epsilon=1.0, sigma=1.0, delta=0.0)
    lang = hoomd.md.methods.Langevin(hoomd.filter.All(), kT=1.0, default_gamma=1.0)
    sim.operations.integrator = hoomd.md.Integrator(dt=0.01, forces=[wca, fene], methods=[lang])
    log(f"[bench] KG melt N={N}, {len(bonds)} bonds: warmup {warmup} steps")
    sim.run(warmup)
    _lib.check(_lib.lib().hm_synchronize())
    t0 = time.perf_counter()
    sim.run(steps)
    _lib.check(_lib.lib().hm_synchronize())
    elapsed = time.perf_counter() - t0
    tps = steps / elapsed
    cpu = None
    if cpu_steps:
        import oracle as O
        snap = sim.state.get_snapshot()
        s = O.System(snap.particles.position, snap.particles.velocity, L, r_cut=2 ** (1 / 6),
                     mode="shift", buffer=0.4, dt=0.01, seed=seed)
        s.set_bonds(bonds, 30.0, 1.5, 1.0, 1.0, 0.0, exclude=True)
        s.set_langevin(1.0, 1.0)
        s.set_sorter(200, 256)
        s.run(1)
        t1 = time.perf_counter()
        s.run(cpu_steps)
        cdt = time.perf_counter() - t1
        cpu = {"value": cpu_steps / cdt, "unit": "timesteps/s", "cores": 1, "kind": "port",
               "sample": f"N={N} KG melt state after the timed steps, {cpu_steps} steps, "
                         f"{cdt:.1f} s on one host core"}
    print(json.dumps({
        "metric": "timesteps/s (Kremer-Grest melt, Langevin)", "value": tps,
        "unit": "timesteps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True,
        "particle_steps_per_s": tps * N,
        "dtype": "f32 pair/bond arithmetic, f64 positions/integration",
        "data": "synthetic (lattice chains, thermalized)",
        "config": {"workload": f"KG melt {N // 50} x 50 beads, rho 0.85, FENE k 30 r0 1.5 + WCA, "
                               "WCA pairs shift, bond exclusions, buffer 0.4, Langevin kT 1 "
                               "gamma 1, dt 0.01", "N": N, "bonds": int(len(bonds)),
                   "nlist_builds_in_timed_steps": nl.num_builds},
        "cpu_baseline": cpu}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--n", type=int, default=1048576, help="particles (4 * n_cells^3)")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pair-only", type=int, default=0,
                    help="profiling mode: after warmup launch only the pair kernel this many "
                         "times (for rocprofv3 counter passes) and exit")
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
    a = ap.parse_args()
    if a.kg:
        if a.gpus != 1:
            print("bench.py runs on exactly one GPU", file=sys.stderr)
            return 2
        return kg(a.steps, a.warmup, a.seed, 0 if a.no_cpu else a.cpu_steps)
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
    log(f"[bench] timing {a.steps} steps")
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

    cpu = None
    if not a.no_cpu:
        log(f"[bench] CPU baseline: oracle, {a.cpu_steps} steps of the same workload, 1 core")
        cps, cdt, cbuilds = cpu_baseline(n_cells, rho, a.seed, a.cpu_steps)
        cpu = {"value": cps, "unit": "timesteps/s", "cores": 1, "kind": "port",
               "sample": f"N={N} LJ Bussi workload, {a.cpu_steps} steps after the first "
                         f"(includes {cbuilds} list rebuilds), {cdt:.1f} s on one host core; "
                         f"{cps * N:.3e} particle-steps/s",
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
                   "pair_launch": launch,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": PAIR_KERNELS[a.pair_kind][0],
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_us": ms * 1e3, "algorithmic_bytes": B,
                     "nlist_entries": int(entries)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
