"""Write tests/golden/nve_drift_lj16384.json: the NVE energy drift of the plain LJ fluid in the
three precision variants of the oracle's precision study (oracle/nve_precision.c).

    python tests/golden/make_nve_drift.py [THREADS]

System: N = 16,384 (FCC 16^3 x 4) at number density 0.8442 in a cubic periodic box, eps =
sigma = 1, r_cut 2.5 (unshifted forces), list buffer 0.4 with distance check, velocities
Maxwell-Boltzmann at kT 1.2 from seed 1 (the oracle's thermalize, COM momentum zeroed), dt
0.005, 100,000 velocity-Verlet NVE steps. Energies (kinetic + pair energy shifted to zero at
r_cut, summed in FP64) every 1,000 steps. Nothing of the reference is run: the variants are
the oracle's own code at three precisions. About 10 minutes per variant on 8 host threads.
"""

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from lattices import fcc_lattice  # noqa: E402

STEPS = 100_000
EVERY = 1_000
DT = 0.005


def drift_stats(e, n, every=EVERY, dt=DT):
    """Per-particle drift figures of one energy series: least-squares slope over the whole run
    and over its second half (per unit time), and end minus start."""
    t = np.arange(len(e)) * every * dt
    epp = np.asarray(e) / n
    half = len(e) // 2
    return {"slope_per_tau": float(np.polyfit(t, epp, 1)[0]),
            "slope_second_half_per_tau": float(np.polyfit(t[half:], epp[half:], 1)[0]),
            "end_minus_start": float(epp[-1] - epp[0]),
            "rms_about_fit": float(np.std(epp - np.polyval(np.polyfit(t, epp, 1), t)))}


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    pos, L = fcc_lattice(16, 0.8442)
    n = len(pos)
    vel = O.thermalize(np.ones(n), np.arange(n), 1.2, 0, 1)
    out = {"N": n, "L": L, "rho": 0.8442, "kT": 1.2, "seed": 1, "r_cut": 2.5, "r_buff": 0.4,
           "dt": DT, "steps": STEPS, "every": EVERY, "variants": {}}
    for v in O.NVE_VARIANTS:
        t0 = time.time()
        ke, pe, _, _, builds = O.nve_precision(v, pos, vel, L, STEPS, EVERY, threads=threads)
        e = ke + pe
        out["variants"][v] = {"kinetic": ke.tolist(), "potential": pe.tolist(),
                              "list_builds": builds, **drift_stats(e, n)}
        print(f"{v}: {time.time() - t0:.0f} s, builds {builds}, "
              f"slope {out['variants'][v]['slope_per_tau']:.3e} per particle per tau", flush=True)
    with open(os.path.join(HERE, "nve_drift_lj16384.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
