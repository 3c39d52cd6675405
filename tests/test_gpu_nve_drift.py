"""NVE energy drift of the plain LJ fluid on the HIP engine over 100,000 steps, against the
oracle's precision study from the same start (tests/golden/nve_drift_lj16384.json).

System: N = 16,384 (FCC 16^3 x 4), number density 0.8442, eps = sigma = 1, r_cut 2.5, buffer
0.4, velocities from the thermalize stream at kT 1.2, seed 1 (bit-equal to the oracle's), dt
0.005, ConstantVolume without a thermostat, the default particle sorter. The energy is kinetic
+ pair energy with the pair potential shifted to zero at r_cut (mode "shift": the same forces as
the unshifted potential, hence the same trajectory; the shifted energy is the one the
truncated force conserves).

Over 1e5 steps at dt 0.005 the drift of every variant is a random walk driven by the force
discontinuity at the unshifted cut-off, not by rounding: FP64, FP64 state with FP32 pairs and
all-FP32 give slopes of 2.5e-7, 2.3e-7 and 0.03e-7 per particle per unit time on the same start,
with fluctuations of 1.7e-5 per particle about the fit. Trajectories diverge within a few
thousand steps, so the engine is held to that spread rather than to one variant's value.

Bar: the engine's drift (least-squares slope of E/N against time, and the end-minus-start
change of E/N) is at most three times the largest of the three variants' in magnitude.
"""

import json
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "nve_drift_lj16384.json")


def test_nve_drift_lj16384():
    import hoomd_amd as hoomd
    from hoomd_amd import _lib
    from lattices import fcc_lattice
    if _lib.lib().hm_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    with open(FIXTURE) as f:
        fx = json.load(f)
    pos, L = fcc_lattice(16, fx["rho"])
    assert len(pos) == fx["N"] and abs(L - fx["L"]) < 1e-12
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = len(pos)
    snap.particles.types = ["A"]
    snap.particles.position[:] = pos
    sim = hoomd.Simulation(device=hoomd.device.GPU(), seed=fx["seed"])
    sim.create_state_from_snapshot(snap)
    sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=fx["kT"])
    nl = hoomd.md.nlist.Cell(buffer=fx["r_buff"])
    lj = hoomd.md.pair.LJ(nlist=nl, default_r_cut=fx["r_cut"], mode="shift")
    lj.params[("A", "A")] = dict(epsilon=1.0, sigma=1.0)
    sim.operations.integrator = hoomd.md.Integrator(
        dt=fx["dt"], forces=[lj], methods=[hoomd.md.methods.ConstantVolume(hoomd.filter.All())])
    thermo = hoomd.md.compute.ThermodynamicQuantities(hoomd.filter.All())
    sim.operations.add(thermo)
    sim.run(0)
    ke = [thermo.kinetic_energy]
    pe = [thermo.potential_energy]
    ref0 = fx["variants"]["fp64"]
    assert abs(ke[0] - ref0["kinetic"][0]) < 1e-9 * ke[0]
    assert abs(pe[0] - ref0["potential"][0]) < 1e-6 * abs(pe[0])
    t0 = time.perf_counter()
    for k in range(fx["steps"] // fx["every"]):
        sim.run(fx["every"])
        ke.append(thermo.kinetic_energy)
        pe.append(thermo.potential_energy)
        if (k + 1) % 20 == 0:
            print(f"step {(k + 1) * fx['every']}: E/N {(ke[-1] + pe[-1]) / fx['N']:.6f} "
                  f"({time.perf_counter() - t0:.1f} s)", flush=True)
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_nve_drift", os.path.join(HERE, "golden", "make_nve_drift.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    drift_stats = gen.drift_stats
    e = np.array(ke) + np.array(pe)
    assert np.all(np.isfinite(e))
    gpu = drift_stats(e, fx["N"], fx["every"], fx["dt"])
    slopes = {v: r["slope_per_tau"] for v, r in fx["variants"].items()}
    for v, s in slopes.items():
        print(f"{v}: slope {s:.3e}, end-start {fx['variants'][v]['end_minus_start']:.3e}")
    print(f"engine: slope {gpu['slope_per_tau']:.3e}, end-start {gpu['end_minus_start']:.3e}, "
          f"list builds {nl.num_builds}", flush=True)
    worst_slope = max(abs(s) for s in slopes.values())
    worst_change = max(abs(r["end_minus_start"]) for r in fx["variants"].values())
    assert abs(gpu["slope_per_tau"]) <= 3.0 * worst_slope
    assert abs(gpu["end_minus_start"]) <= 3.0 * worst_change
