"""The oracle's NVE precision study (oracle/nve_precision.c) pinned to the oracle's System, and
the committed drift fixture checked for consistency.

* FP64 variant: the same trajectory as the oracle's System in NVE (mode "shift", whose energy
  the study measures) step by step: positions within 1e-12, total energy within 1e-11 relative
  after 200 steps.
* The FP32 variants really compute in FP32: their 20-step energies leave the FP64 one by more
  than FP64 rounding (1e-11 relative) and stay within FP32 rounding growth (1e-5).
* The fixture's drift figures follow from its own samples.
"""

import json
import os

import numpy as np
import pytest

import oracle as O
from lattices import fcc_lattice

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "nve_drift_lj16384.json")


@pytest.fixture(scope="module")
def start():
    pos, L = fcc_lattice(8, 0.8442)
    n = len(pos)
    vel = O.thermalize(np.ones(n), np.arange(n), 1.2, 0, 1)
    return pos, vel, L


def test_fp64_variant_is_the_oracle_system(start):
    pos, vel, L = start
    n = len(pos)
    s = O.System(pos, vel, L, r_cut=2.5, mode="shift", buffer=0.4, dt=0.005, seed=1)
    s.run(0)
    e_sys = [s.kinetic_energy() + s.potential_energy()]
    for _ in range(20):
        s.run(10)
        e_sys.append(s.kinetic_energy() + s.potential_energy())
    ke, pe, p, v, builds = O.nve_precision("fp64", pos, vel, L, 200, 10, threads=2)
    rp, rv, _ = s.state()
    d = p - rp
    d -= L * np.rint(d / L)
    assert np.abs(d).max() < 1e-12 and np.abs(v - rv).max() < 1e-11
    e = ke + pe
    assert np.abs(e - np.array(e_sys)).max() / np.abs(e_sys).max() < 1e-11
    assert builds == s.n_builds
    # the energy is the one the run conserves: the velocity-Verlet error over 200 steps at
    # dt 0.005 is far below the energy scale (kinetic ~ 1.8 per particle)
    assert abs(e[-1] - e[0]) / n < 1e-3


def test_fp32_variants_round_in_fp32(start):
    pos, vel, L = start
    e = {}
    for v in O.NVE_VARIANTS:
        ke, pe, _, _, _ = O.nve_precision(v, pos, vel, L, 20, 20, threads=2)
        e[v] = ke + pe
    for v in ("fp64_state_fp32_pairs", "fp32"):
        dev = np.abs(e[v] - e["fp64"]).max() / np.abs(e["fp64"]).max()
        assert 1e-11 < dev < 1e-5, (v, dev)


def test_threads_do_not_change_the_trajectory(start):
    pos, vel, L = start
    a = O.nve_precision("fp32", pos, vel, L, 50, 50, threads=1)
    b = O.nve_precision("fp32", pos, vel, L, 50, 50, threads=3)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


def test_bad_arguments_raise(start):
    pos, vel, L = start
    with pytest.raises(ValueError):
        O.nve_precision("fp64", pos, vel, L, 25, 10)  # steps not a multiple of every
    with pytest.raises(ValueError):
        O.nve_precision("fp64", pos[:8], vel[:8], 5.0, 10, 10)  # fewer than 3 cells per side


def test_fixture_figures_follow_from_samples():
    if not os.path.exists(FIXTURE):
        pytest.skip("drift fixture not generated (tests/golden/make_nve_drift.py)")
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_nve_drift", os.path.join(HERE, "golden", "make_nve_drift.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    with open(FIXTURE) as f:
        fx = json.load(f)
    assert fx["N"] == 16384 and fx["steps"] == 100_000 and set(fx["variants"]) == set(
        O.NVE_VARIANTS)
    for v, rec in fx["variants"].items():
        e = np.array(rec["kinetic"]) + np.array(rec["potential"])
        assert len(e) == fx["steps"] // fx["every"] + 1
        st = gen.drift_stats(e, fx["N"], fx["every"], fx["dt"])
        for k, val in st.items():
            assert np.isclose(rec[k], val, rtol=1e-12, atol=1e-15), (v, k)
    # same start for every variant: the step-0 energies agree to FP32 rounding of the state
    e0 = [fx["variants"][v]["kinetic"][0] + fx["variants"][v]["potential"][0]
          for v in fx["variants"]]
    assert np.ptp(e0) / abs(e0[0]) < 1e-6
