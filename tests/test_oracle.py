"""CPU checks that pin the oracle (FP64 restatement of the reference CPU path).

Each test names what fixes the expected answer: a golden fixture from the reference's own
tests, the reference's own headers compiled in oracle/_ref, a closed form, an invariant, or a
brute-force recomputation. None of them retypes the oracle's formula.
"""

import math

import numpy as np
import pytest

import oracle as O
from lattices import fcc_lattice, maxwell_boltzmann, perturbed_fcc, random_fluid, sc_lattice

H = lambda s: int(s, 16) if isinstance(s, str) else int(s)  # noqa: E731


# ------------------------------------------------------------------ RNG
def test_philox_known_answers(golden):
    kat = golden("rng_known_answers.json")
    for v in kat["philox4x32_10"]:
        out = O.philox([H(x) for x in v["ctr"]], [H(x) for x in v["key"]])
        assert [int(x) for x in out] == [H(x) for x in v["out"]]


def test_seed_counter_u01_known_answers(golden):
    kat = golden("rng_known_answers.json")
    s = kat["seed"]
    key = O.seed_key(s["id"], H(s["timestep"]), H(s["seed"]))
    assert [int(x) for x in key] == [H(x) for x in s["key"]]
    for c in kat["counters"]:
        assert [int(x) for x in O.counter(*[H(a) for a in c["args"]])] == [H(x) for x in c["ctr"]]
    assert O.u01(0) == kat["u01_double"]["smallest"]
    assert O.u01(2**64 - 1) == kat["u01_double"]["largest"]
    # uneg11 endpoints: (int64) min * 2^-63 + 2^-64 = -1 + 2^-64 -> -1.0; max -> 1.0
    assert O.uneg11(1 << 63) == -1.0
    assert O.uneg11((1 << 63) - 1) == 1.0


def test_normal_and_gamma_moments():
    """Statistical pin of the Box-Muller normal and the Marsaglia-Tsang gamma (the stream
    itself is unpinned; random_numbers_test.cc checks moments the same way)."""
    # thermalize draws N(0, sqrt(kT/m)) per axis from Philox; 60k particles x 3 axes
    N = 60000
    v = O.thermalize(np.ones(N), np.arange(N), kT=2.0, timestep=5, seed=11)
    x = v.ravel()
    assert abs(x.mean()) < 4 * math.sqrt(2.0 / x.size)
    assert abs(x.var() / 2.0 - 1.0) < 0.02
    kurt = ((x - x.mean()) ** 4).mean() / x.var() ** 2
    assert abs(kurt - 3.0) < 0.1
    # Bussi with tau = 0 draws K' = alpha^2 K = (kT/2)(R^2 + 2 Gamma((g-1)/2)); mean g kT/2
    g, kT, K = 30.0, 1.3, 7.0
    ks = np.array([O.bussi_alpha(K, g, 0.005, 0.0, kT, t, 3) ** 2 * K for t in range(6000)])
    mean, var = ks.mean(), ks.var()
    assert abs(mean - g * kT / 2) < 4 * math.sqrt(g / 2 * kT**2 / len(ks))
    assert abs(var / (g / 2 * kT**2) - 1.0) < 0.1  # Gamma(g/2, kT) variance


def test_stream_distributions_ks():
    """Kolmogorov-Smirnov checks of the streams whose exact values cannot be pinned here (the
    reference's RandomNumbers.h needs the Random123 headers, an un-vendored submodule, and its
    tests hold no normal or gamma vectors). A distribution test cannot catch a change that
    keeps the distribution (sin for cos in Box-Muller, a swapped word pair); it does catch a
    wrong scale, shape, a lost draw or a biased branch.

    * thermalize: each axis N(0, sqrt(kT/m)) over 3 x 40,000 particles, m = 1 and m = 2.5;
      axes uncorrelated;
    * Bussi at tau = 0: K' = alpha^2 K is Gamma(g/2, scale kT) for g = 2 (the shape-1/2 gamma
      path with its boost), 5, 30 and 2,997 degrees of freedom, 20,000 timesteps each;
    * Langevin noise: uniform on [-1, 1] per axis."""
    from scipy import stats
    N = 40000
    for m, kT in ((1.0, 1.2), (2.5, 0.7)):
        v = O.thermalize(np.full(N, m), np.arange(N), kT, 3, 21)
        sd = math.sqrt(kT / m)
        for d in range(3):
            assert stats.kstest(v[:, d], "norm", args=(0.0, sd)).pvalue > 1e-3
        c = np.corrcoef(v.T)
        assert np.abs(c[np.triu_indices(3, 1)]).max() < 4 / math.sqrt(N)
    for g, kT in ((2.0, 1.0), (5.0, 1.3), (30.0, 0.8), (2997.0, 1.2)):
        K = 0.5 * g * kT * 1.7  # any current kinetic energy: K' does not depend on it at tau 0
        ks = np.array([O.bussi_alpha(K, g, 0.005, 0.0, kT, t, 7) ** 2 * K for t in range(20000)])
        assert stats.kstest(ks, "gamma", args=(g / 2, 0.0, kT)).pvalue > 1e-3, g
    vel = np.zeros((N, 3))
    f4 = np.zeros((N, 4))
    gamma, dt = 1.0, 0.01
    _, a = O.langevin_step_two(vel, np.ones(N), np.arange(N), f4, gamma, 1.0, dt, 4, 9)
    r = a / math.sqrt(6 * gamma * 1.0 / dt)  # the uniform draws, F = r sqrt(6 gamma kT / dt)
    for d in range(3):
        assert stats.kstest(r[:, d], "uniform", args=(-1.0, 2.0)).pvalue > 1e-3


# ------------------------------------------------------------------ evaluators
def test_lj_golden(golden):
    """forces_and_energies.json 'LJ' (Mathematica), rtol 5e-6 as test_potential.py:994."""
    fx = golden("lj_forces_and_energies.json")
    lj = fx["LJ"]
    for p, forces, energies in zip(lj["params"], lj["forces"], lj["energies"]):
        for i, d in enumerate(fx["distances"]):
            ok, fdivr, e = O.lj(d * d, fx["r_cut"] ** 2, p["sigma"], p["epsilon"])
            assert ok
            # particle 0 at z=0.1, particle 1 at z=d+0.1: F_0z = f/r * (z0 - z1) = -f/r * d
            assert math.isclose(-fdivr * d, forces[i], rel_tol=5e-6, abs_tol=1e-6)
            assert math.isclose(e, energies[i], rel_tol=5e-6, abs_tol=1e-6)


def test_lj_matches_reference_header():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import ctypes as C
    f, e = C.c_double(), C.c_double()
    rng = np.random.default_rng(0)
    for _ in range(3000):
        r = rng.uniform(0.7, 3.0)
        s, eps = rng.uniform(0.5, 1.5), rng.uniform(0.1, 2.0)
        for shift in (0, 1):
            ok = R.ref_lj(r * r, 6.25, s, eps, shift, C.byref(f), C.byref(e))
            ok2, f2, e2 = O.lj(r * r, 6.25, s, eps, shift)
            assert bool(ok) == ok2 and f.value == f2 and e.value == e2


def test_lj_shift_closed_form():
    """mode='shift' makes V(r_cut) = 0 (test_potential.py:780-855)."""
    rc = 2.5
    ok, f, e = O.lj((rc - 1e-12) ** 2, rc * rc, 1.0, 1.0, shift=True)
    assert ok and abs(e) < 1e-10
    ok, f, e = O.lj(rc * rc, rc * rc, 1.0, 1.0, shift=True)  # r_cut strict
    assert not ok


def test_table_golden(golden):
    """forces_and_energies.json 'Table' (three tables, r = 0.75 and 1.5, r_cut 2.5) through the
    pair accumulation of a two-particle system: the energies of the two particles sum to the
    pair energy, particle 0 (below particle 1 on z) feels the reference's force along z."""
    fx = golden("table_forces_and_energies.json")
    t = fx["Table"]
    box = O.Box(20.0)
    for p, forces, energies in zip(t["params"], t["forces"], t["energies"]):
        for i, d in enumerate(fx["distances"]):
            pos = np.array([[0, 0, 0.1], [0, 0, d + 0.1]])
            nl = np.array([[1], [0]], np.uint32)
            nn = np.array([1, 1], np.uint32)
            f, v = O.pair_table(box, pos, nl, nn, False, fx["r_cut"], p["r_min"], p["U"], p["F"])
            assert math.isclose(f[:, 3].sum(), energies[i], rel_tol=5e-6, abs_tol=1e-6)
            assert math.isclose(f[0, 2], forces[i], rel_tol=5e-6, abs_tol=1e-6)
            assert abs(f[0, 0]) + abs(f[0, 1]) == 0 and f[1, 2] == -f[0, 2]
            # the half list with the third law gives the same forces
            f3, _ = O.pair_table(box, pos, np.array([[1], [0]], np.uint32),
                                 np.array([1, 0], np.uint32), True, fx["r_cut"], p["r_min"],
                                 p["U"], p["F"])
            np.testing.assert_array_equal(f3, f)


def test_table_interpolation_closed_forms():
    """Linear interpolation on r_min + k dr, dr = (r_cut - r_min) / width; zero below r_min and
    from r_cut on; the interval after the last point runs linearly to zero."""
    rmin, rc, w = 0.8, 2.0, 6
    dr = (rc - rmin) / w
    U = np.array([5.0, 3.0, 2.0, 1.5, 1.0, 0.5])
    F = np.array([9.0, 4.0, 2.0, 1.0, 0.5, 0.25])
    for k in range(w):
        r = rmin + k * dr
        ok, fdr, e = O.table(r * r, rc * rc, rmin, U, F)
        assert ok and math.isclose(e, U[k], rel_tol=1e-12) and math.isclose(fdr * r, F[k],
                                                                              rel_tol=1e-12)
        rm = r + 0.5 * dr
        nxt = (U[k + 1], F[k + 1]) if k + 1 < w else (0.0, 0.0)
        ok, fdr, e = O.table(rm * rm, rc * rc, rmin, U, F)
        assert math.isclose(e, 0.5 * (U[k] + nxt[0]), rel_tol=1e-9)
        assert math.isclose(fdr * rm, 0.5 * (F[k] + nxt[1]), rel_tol=1e-9)
    assert not O.table((rmin - 1e-9) ** 2, rc * rc, rmin, U, F)[0]
    assert not O.table(rc * rc, rc * rc, rmin, U, F)[0]
    ok, fdr, e = O.table((rc - 1e-12) ** 2, rc * rc, rmin, U, F)
    assert ok and abs(e) < 1e-9
    # a fine table of LJ reproduces LJ to second order in dr
    rs = np.linspace(0.9, 2.5, 4000, endpoint=False)
    Ul = 4 * (rs**-12 - rs**-6)
    Fl = 24 * (2 * rs**-13 - rs**-7)
    for r in (0.95, 1.12, 1.5, 2.2):
        ok, fdr, e = O.table(r * r, 6.25, 0.9, Ul, Fl)
        ok2, fl, el = O.lj(r * r, 6.25, 1.0, 1.0)
        assert abs(e - el) < 1e-4 * max(1.0, abs(el)) and abs(fdr - fl) < 1e-4 * max(1.0, abs(fl))


def test_fene_golden(golden):
    fx = golden("fene_bond.json")
    d = fx["d"]
    for c in fx["cases"]:
        p = c["params"]
        ok, fdivr, e = O.fene(d * d, p["k"], p["r0"], p["sigma"], p["epsilon"], p["delta"])
        assert ok
        # particle 0 at -d/2, 1 at +d/2, bond (0, 1): F_0x = -f/r * (x1 - x0) = -f/r * d
        assert math.isclose(-fdivr * d, c["force"], rel_tol=fx["rtol"])
        assert math.isclose(e, c["energy"], rel_tol=fx["rtol"])


def test_fene_matches_reference_header():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import ctypes as C
    f, e = C.c_double(), C.c_double()
    rng = np.random.default_rng(1)
    for _ in range(3000):
        r = rng.uniform(0.6, 1.6)
        args = (30.0, 1.5, 1.0, 1.0, rng.choice([0.0, -0.3, 0.2]))
        ok = R.ref_fene(r * r, *args, C.byref(f), C.byref(e))
        ok2, f2, e2 = O.fene(r * r, *args)
        assert bool(ok) == ok2 and f.value == f2 and e.value == e2


def test_fene_finite_difference():
    args = (30.0, 1.5, 1.0, 1.0, 0.0)
    for r in (0.9, 1.0, 1.12, 1.3):
        h = 1e-6
        _, _, ep = O.fene((r + h) ** 2, *args)
        _, _, em = O.fene((r - h) ** 2, *args)
        _, fdivr, _ = O.fene(r * r, *args)
        assert math.isclose(fdivr * r, -(ep - em) / (2 * h), rel_tol=1e-6, abs_tol=1e-6)


# ------------------------------------------------------------------ box
def test_box_matches_reference_header():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import ctypes as C
    rng = np.random.default_rng(2)
    for L in (10.0, 26.8735, 107.494):
        box = O.Box(L)
        for _ in range(500):
            w = rng.uniform(-0.99 * L, 0.99 * L, 3)
            a = w.copy()
            R.ref_min_image(L, L, L, a.ctypes.data_as(C.c_void_p))
            assert (a == box.min_image(w)).all()
            f = np.zeros(3)
            R.ref_make_fraction(L, L, L, w.ctypes.data_as(C.c_void_p), f.ctypes.data_as(C.c_void_p))
            assert (f == box.make_fraction(w)).all()
            img = np.array([1, -2, 3], np.int32)
            w2 = w.copy()
            R.ref_wrap(L, L, L, w2.ctypes.data_as(C.c_void_p), img.ctypes.data_as(C.c_void_p))
            w3, img3 = box.wrap(w, [1, -2, 3])
            assert (w2 == w3).all() and (img == img3).all()


def test_min_image_closed_form():
    box = O.Box(10.0)
    assert np.allclose(box.min_image([6.0, -7.0, 4.9]), [-4.0, 3.0, 4.9])


# ------------------------------------------------------------------ cell list
def test_cell_dims_known_answers(golden):
    for c in golden("cell_dims.json")["cases"]:
        dim = O.cell_dims(O.Box(*c["L"]), c["width"])
        assert list(dim) == c["dim"]


@pytest.mark.parametrize("dim", [(5, 5, 5), (3, 4, 6), (2, 5, 7), (1, 1, 4)])
def test_cell_adjacency_invariants(dim):
    adj = O.cell_adj(dim)
    nx, ny, nz = dim
    assert adj.shape[1] == min(nx, 3) * min(ny, 3) * min(nz, 3)
    for c in range(nx * ny * nz):
        row = adj[c]
        assert (np.diff(row.astype(np.int64)) > 0).all()  # sorted, unique
        assert c in row
        i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
        expect = {(((k + dk) % nz) * ny + (j + dj) % ny) * nx + (i + di) % nx
                  for di in (-1, 0, 1) for dj in (-1, 0, 1) for dk in (-1, 0, 1)}
        assert set(int(x) for x in row) == expect


def test_cell_list_invariants():
    pos, L = random_fluid(3000, 0.8, seed=4)
    box = O.Box(L)
    dim = O.cell_dims(box, 2.9)
    err, size, idx, cell_of, ovf = O.cell_list(box, dim, pos, 64)
    assert err == 0 and ovf == 0
    assert size.sum() == len(pos)
    # each particle in exactly one cell, the cell given by floor(f * dim)
    seen = np.concatenate([idx[c, :size[c]] for c in range(len(size))])
    assert sorted(seen.tolist()) == list(range(len(pos)))
    f = (pos + L / 2) / L
    ijk = np.floor(f * dim).astype(int) % dim
    expect = (ijk[:, 2] * dim[1] + ijk[:, 1]) * dim[0] + ijk[:, 0]
    assert (cell_of == expect).all()
    # overflow reports the needed capacity
    err, size2, idx2, _, ovf = O.cell_list(box, dim, pos, 2)
    assert ovf == size.max()


def test_cell_list_errors():
    box = O.Box(10.0)
    dim = O.cell_dims(box, 1.0)
    pos = np.zeros((3, 3))
    pos[1] = np.nan
    assert O.cell_list(box, dim, pos, 8)[0] == -2
    pos = np.zeros((3, 3))
    pos[2, 0] = 7.0  # outside [-5, 5)
    assert O.cell_list(box, dim, pos, 8)[0] == -3


# ------------------------------------------------------------------ neighbour list
def test_true_pair_list(golden):
    """true_pair_list.json: SC 7^3, a=1, r_cut 1.1, buffer 0 -> the 6 nearest neighbours."""
    fx = golden("true_pair_list.json")
    pos, L = sc_lattice(fx["lattice"]["n"], fx["lattice"]["a"])
    box = O.Box(L)
    for full in (True, False):
        nl, nn = O.nlist_binned(box, pos, fx["r_cut"], fx["buffer"], full=full)
        got = O.pair_set(nl, nn)
        expect = {tuple(sorted(p)) for p in fx["pairs"]}
        assert got == expect
        assert nn.sum() == (2 if full else 1) * len(expect)


@pytest.mark.parametrize("N,rho,rcut,buff", [(1500, 0.84, 2.5, 0.4), (800, 0.3, 1.12246, 0.4),
                                            (300, 0.84, 1.5, 0.0)])
def test_binned_equals_brute_force(N, rho, rcut, buff):
    pos, L = random_fluid(N, rho, seed=N)
    box = O.Box(L)
    nl, nn = O.nlist_binned(box, pos, rcut, buff, full=True)
    bl, bn = O.nlist_brute(box, pos, rcut + buff, full=True)
    assert (nn == bn).all()
    assert O.pair_set(nl, nn) == O.pair_set(bl, bn)
    hl, hn = O.nlist_binned(box, pos, rcut, buff, full=False)
    assert hn.sum() * 2 == nn.sum()
    assert O.pair_set(hl, hn) == O.pair_set(nl, nn)


def test_nlist_exclusions():
    pos, L = random_fluid(400, 0.8, seed=9)
    box = O.Box(L)
    nl, nn = O.nlist_binned(box, pos, 2.5, 0.4, full=True)
    # exclude each particle's first neighbour (symmetrically)
    excl = [[] for _ in range(len(pos))]
    for i in range(len(pos)):
        if nn[i]:
            j = int(nl[i, 0])
            if j not in excl[i]:
                excl[i].append(j)
                excl[j].append(i)
    el, en = O.nlist_binned(box, pos, 2.5, 0.4, full=True, excl=excl)
    full = O.pair_set(nl, nn)
    ex = {tuple(sorted((i, j))) for i in range(len(excl)) for j in excl[i]}
    assert O.pair_set(el, en) == full - ex


def test_distance_check():
    pos, L = random_fluid(200, 0.5, seed=5)
    box = O.Box(L)
    last = pos.copy()
    assert not O.distance_check(box, pos, last, 2.5, 0.4)
    moved = pos.copy()
    moved[17, 0] += 0.1999
    assert not O.distance_check(box, moved, last, 2.5, 0.4)
    moved[17, 0] += 0.0002
    assert O.distance_check(box, moved, last, 2.5, 0.4)


# ------------------------------------------------------------------ pair / bond forces
def test_pair_third_law_equals_full_list():
    pos, L = perturbed_fcc(5, 0.8442, r=0.1)
    box = O.Box(L)
    fl, fn = O.nlist_binned(box, pos, 2.5, 0.4, full=True)
    hl, hn = O.nlist_binned(box, pos, 2.5, 0.4, full=False)
    for mode in ("none", "shift", "xplor"):
        ff, fv = O.pair_lj(box, pos, fl, fn, third_law=False, mode=mode, r_on=2.0)
        hf, hv = O.pair_lj(box, pos, hl, hn, third_law=True, mode=mode, r_on=2.0)
        assert np.allclose(ff, hf, rtol=1e-10, atol=1e-10)
        assert np.allclose(fv, hv, rtol=1e-10, atol=1e-10)
    assert np.allclose(ff[:, :3].sum(axis=0), 0.0, atol=1e-9)  # Newton's third law


def test_pair_force_is_minus_energy_gradient():
    """F = -dE/dx by central differences on the total energy (test_potential.py:858-949)."""
    pos, L = perturbed_fcc(3, 0.8442, r=0.1, seed=8)
    box = O.Box(L)
    nl, nn = O.nlist_binned(box, pos, 2.5, 1.0, full=False)
    f, _ = O.pair_lj(box, pos, nl, nn, third_law=True)
    h = 1e-6
    for i in (0, 7, 33):
        for d in range(3):
            p = pos.copy()
            p[i, d] += h
            ep = O.pair_lj(box, p, nl, nn, third_law=True)[0][:, 3].sum()
            p[i, d] -= 2 * h
            em = O.pair_lj(box, p, nl, nn, third_law=True)[0][:, 3].sum()
            assert math.isclose(f[i, d], -(ep - em) / (2 * h), rel_tol=1e-5, abs_tol=1e-5)


def test_pair_virial_trace_is_pair_sum():
    """Sum of per-particle virial xx+yy+zz equals 1/2 sum_pairs r.F for a two-particle system."""
    box = O.Box(20.0)
    pos = np.array([[-0.55, 0, 0.1], [0.55, 0, 0.1]])
    nl, nn = O.nlist_binned(box, pos, 2.5, 0.4, full=False)
    f, v = O.pair_lj(box, pos, nl, nn, third_law=True)
    _, fdivr, e = O.lj(1.1**2, 6.25, 1.0, 1.0)
    assert math.isclose(f[:, 3].sum(), e)
    assert math.isclose(v[0].sum(), fdivr * 1.1**2)  # xx of both halves
    assert math.isclose(f[0, 0], -fdivr * 1.1)


def test_bond_fene_system(golden):
    fx = golden("fene_bond.json")
    d = fx["d"]
    box = O.Box(fx["L"])
    pos = np.array([[-d / 2, 0, 0.1], [d / 2, 0, 0.1]])
    for c in fx["cases"]:
        p = c["params"]
        r, f, v = O.bond_fene(box, pos, [[0, 1]], p["k"], p["r0"], p["sigma"], p["epsilon"],
                              p["delta"])
        assert r == 0
        assert math.isclose(f[:, 3].sum(), c["energy"], rel_tol=fx["rtol"])
        assert np.allclose(f[0, :3], [c["force"], 0, 0], rtol=fx["rtol"], atol=1e-5)
        assert np.allclose(f[1, :3], [-c["force"], 0, 0], rtol=fx["rtol"], atol=1e-5)
    # stretched beyond r0 -> error code for the bond
    far = np.array([[-1.0, 0, 0], [1.0, 0, 0]])
    r, _, _ = O.bond_fene(box, far, [[0, 1]], 30.0, 1.5, 1.0, 1.0, 0.0)
    assert r == -1


# ------------------------------------------------------------------ integrators
def test_nvt_step_closed_form():
    box = O.Box(10.0)
    pos = np.array([[4.999, 0.0, -1.0]])
    vel = np.array([[1.0, -2.0, 0.5]])
    acc = np.array([[0.5, 0.25, -1.0]])
    dt, alpha = 0.01, 0.9
    p, v, im = O.nvt_step_one(box, pos, vel, acc, np.zeros((1, 3), np.int32), dt, alpha)
    v_exp = (vel + 0.5 * acc * dt) * alpha
    p_exp = pos + dt * v_exp
    assert np.allclose(v, v_exp, rtol=1e-15)
    assert np.allclose(p, p_exp - [[10.0, 0, 0]], rtol=1e-13)  # wrapped once in +x
    assert im.tolist() == [[1, 0, 0]]
    f4 = np.array([[2.0, -4.0, 1.0, 0.0]])
    v2, a2 = O.nvt_step_two(v, [2.0], f4, dt, 1.0)
    assert np.allclose(a2, f4[:, :3] / 2.0)
    assert np.allclose(v2, v + 0.5 * dt * a2)


def test_langevin_zero_gamma_is_nve():
    vel = np.array([[1.0, 2.0, 3.0]])
    f4 = np.array([[1.0, 1.0, 1.0, 0.0]])
    v1, a1 = O.langevin_step_two(vel, [1.0], [0], f4, 0.0, 1.0, 0.01, 7, 1)
    v2, a2 = O.nvt_step_two(vel, [1.0], f4, 0.01)
    assert np.allclose(v1, v2) and np.allclose(a1, a2)


def test_langevin_noise_statistics():
    """Uniform noise in [-1, 1] per axis: F_R = r sqrt(6 gamma kT / dt) has variance 2 gamma kT/dt."""
    N = 20000
    vel = np.zeros((N, 3))
    f4 = np.zeros((N, 4))
    gamma, kT, dt = 1.0, 1.0, 0.01
    v, a = O.langevin_step_two(vel, np.ones(N), np.arange(N), f4, gamma, kT, dt, 3, 2)
    var = a.var()
    assert abs(var / (2 * gamma * kT / dt) - 1.0) < 0.03
    assert np.abs(a).max() <= math.sqrt(6 * gamma * kT / dt) + 1e-9


def _lj_system(n=5, kT=1.2, dt=0.005, seed=3):
    pos, L = perturbed_fcc(n, 0.8442, r=0.02, seed=seed)
    vel = maxwell_boltzmann(len(pos), kT, seed=seed)
    return pos, vel, L


@pytest.mark.slow
def test_nve_energy_conservation():
    pos, vel, L = _lj_system(5)
    s = O.System(pos, vel, L, r_cut=2.5, mode="shift", buffer=0.4, dt=0.002)
    s.run(1)
    e0 = s.kinetic_energy() + s.potential_energy()
    es = []
    for _ in range(20):
        s.run(50)
        es.append(s.kinetic_energy() + s.potential_energy())
    drift = np.abs(np.array(es) - e0).max() / len(pos)
    assert drift < 2e-4
    assert s.n_builds > 1


def test_time_reversibility():
    pos, vel, L = _lj_system(4)
    s = O.System(pos, vel, L, r_cut=2.5, mode="shift", buffer=0.4, dt=0.002)
    s.run(100)
    p1, v1, im1 = s.state()
    s2 = O.System(p1, -v1, L, r_cut=2.5, mode="shift", buffer=0.4, dt=0.002)
    s2.run(100)
    p2, v2, im2 = s2.state()
    unwrapped = p2 + (im2 + im1) * L
    assert np.allclose(unwrapped, pos, atol=1e-8)
    assert np.allclose(-v2, vel, atol=1e-8)


def test_bussi_reaches_target_temperature():
    pos, vel, L = _lj_system(4, kT=3.0)
    s = O.System(pos, vel, L, r_cut=2.5, buffer=0.4, dt=0.005, seed=5)
    s.set_bussi(1.0, tau=0.0)
    s.run(50)
    ks = []
    for _ in range(40):
        s.run(5)
        ks.append(s.kinetic_energy())
    N = len(pos)
    assert abs(np.mean(ks) / ((3 * N - 3) / 2 * 1.0) - 1.0) < 0.05


def test_thermalize_momentum_and_temperature():
    N = 5000
    m = np.full(N, 2.0)
    v = O.thermalize(m, np.arange(N), 1.5, 0, 42)
    assert np.allclose((m[:, None] * v).sum(axis=0), 0.0, atol=1e-9)
    ke = 0.5 * (m[:, None] * v * v).sum()
    assert abs(ke / (1.5 * N * 1.5) - 1.0) < 0.05
    # the stream depends only on (seed, timestep, tag): reordering particles permutes rows
    perm = np.random.default_rng(0).permutation(N)
    v2 = O.thermalize(m[perm], perm, 1.5, 0, 42)
    assert np.allclose(v2, v[perm])


# ------------------------------------------------------------------ Hilbert sort
@pytest.mark.parametrize("grid", [2, 4, 16, 64])
def test_hilbert_traversal_is_a_hilbert_curve(grid):
    t = O.hilbert_traversal(grid)
    assert sorted(t.tolist()) == list(range(grid**3))
    order = np.argsort(t)  # bins in curve order; bin = i*g*g + j*g + k
    i, j, k = order // (grid * grid), (order // grid) % grid, order % grid
    step = np.abs(np.diff(i)) + np.abs(np.diff(j)) + np.abs(np.diff(k))
    assert (step == 1).all()  # consecutive bins are face neighbours
    assert order[0] == 0  # starts at the origin


def test_sort_order_properties():
    pos, L = random_fluid(5000, 0.8, seed=12)
    box = O.Box(L)
    grid = 16
    t = O.hilbert_traversal(grid)
    order = O.sort_order(box, pos, grid, t)
    assert sorted(order.tolist()) == list(range(len(pos)))
    f = (pos + L / 2) / L
    b = (np.floor(f * grid).astype(int) % grid)
    key = t[b[:, 0] * grid * grid + b[:, 1] * grid + b[:, 2]]
    ks = key[order]
    assert (np.diff(ks.astype(np.int64)) >= 0).all()
    same = np.diff(ks.astype(np.int64)) == 0
    assert (np.diff(order.astype(np.int64))[same] > 0).all()  # stable ties


def test_sorted_system_matches_unsorted():
    """Sorting permutes memory only: the trajectory by tag is unchanged (FP64, same sums up to
    order -> tiny differences)."""
    pos, vel, L = _lj_system(4)
    a = O.System(pos, vel, L, dt=0.005)
    b = O.System(pos, vel, L, dt=0.005)
    b.set_sorter(10, 16)
    a.run(30)
    b.run(30)
    pa, va, _ = a.state()
    pb, vb, _ = b.state()
    assert np.allclose(pa, pb, atol=1e-9) and np.allclose(va, vb, atol=1e-9)


def test_fcc_lattice_inputs():
    pos, L = fcc_lattice(16)
    assert len(pos) == 16384 and math.isclose(L, 26.8735, rel_tol=1e-5)
    assert (pos >= -L / 2).all() and (pos < L / 2).all()


def test_rebuild_policy_counts():
    """NeighborList::needsUpdating (NeighborList.cc:1325-1418): without distance checks the list
    is rebuilt exactly every rebuild_check_delay steps (every step for 0); with them, no check
    happens before the delay has passed, a build on the first check after a build is counted as
    dangerous, and the shortest period is the delay or longer."""
    from lattices import maxwell_boltzmann, perturbed_fcc
    pos, L = perturbed_fcc(4, 0.8442, r=0.05, seed=2)
    vel = maxwell_boltzmann(len(pos), 1.2, seed=2)
    for delay, dist, steps, expect in ((5, False, 50, 1 + 10), (0, False, 20, 1 + 20),
                                       (3, False, 31, 1 + 10)):
        s = O.System(pos, vel, L, dt=0.005)
        s.set_nlist_policy(delay, dist)
        s.run(steps)
        assert s.n_builds == expect
        assert s.n_dangerous == 0
        assert s.shortest_rebuild == max(delay, 1) if delay else s.shortest_rebuild == 1
    # hot system, distance checks from the 4th step after a build: every build is dangerous
    hot = maxwell_boltzmann(len(pos), 40.0, seed=3)
    s = O.System(pos, hot, L, dt=0.005)
    s.set_nlist_policy(4, True)
    s.run(40)
    assert s.n_builds > 3
    assert s.n_dangerous == s.n_builds - 1
    assert s.shortest_rebuild == 4
    # default policy: the first check after a build is never dangerous
    s = O.System(pos, hot, L, dt=0.005)
    s.run(40)
    assert s.n_dangerous == 0
    assert 1 <= s.shortest_rebuild < 100
