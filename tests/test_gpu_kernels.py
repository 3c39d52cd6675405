"""GPU parity of each kernel driver against the oracle, through the C ABI.

Bars: bit-exact for cell assignments, neighbour lists (row contents and order), the Hilbert
order and the FP64 integrator; pair/bond forces, energies and virials within 1e-5 relative RMS
of an FP64 re-evaluation on the same neighbour list (FP32 pair arithmetic).
"""

import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from gpu_helpers import cell_list, dev, entries_idx, nlist, pair_lj, pos4, rel_rms
from lattices import perturbed_fcc, random_fluid, sc_lattice

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5  # relative RMS vs FP64 per quantity


@pytest.fixture(scope="module")
def H():
    from hoomd_amd import _lib
    L = _lib.lib()
    if L.hm_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return _lib


# ------------------------------------------------------------------ cell list
@pytest.mark.parametrize("N,rho,width", [(20000, 0.8442, 2.9), (3000, 0.5, 1.5225), (50, 0.3, 2.9)])
def test_cell_list_bit_exact(H, N, rho, width):
    pos, L = random_fluid(N, rho, seed=N)
    cl = cell_list(pos, L, width, cmax=96)
    size, cells = entries_idx(cl)
    assert (cl["d_cond"].numpy() == 0).all()
    box = O.Box(L)
    err, osize, oidx, cell_of, ovf = O.cell_list(box, cl["dim"], pos, 96)
    assert err == 0 and ovf == 0
    assert (size == osize).all()
    for c in range(cl["ncell"]):
        assert cells[c] == oidx[c, :osize[c]].tolist()
    assert np.array_equal(cl["d_last"].numpy().reshape(-1, 3), pos)


def test_cell_list_conditions(H):
    pos, L = random_fluid(2000, 0.8, seed=2)
    cl = cell_list(pos, L, 2.9, cmax=4)  # overflow
    size, _ = entries_idx(cl)
    assert cl["d_cond"].numpy()[0] == size.max()
    bad = pos.copy()
    bad[17, 1] = np.nan
    bad[99, 0] = L  # outside [-L/2, L/2)
    cl = cell_list(bad, L, 2.9, cmax=96)
    c = cl["d_cond"].numpy()
    assert c[1] == 18 and c[2] == 100


# ------------------------------------------------------------------ neighbour list
@pytest.mark.parametrize("N,rho,rcut,buff", [(16384, 0.8442, 2.5, 0.4), (4000, 0.85, 2**(1 / 6), 0.4),
                                            (2000, 0.84, 2.5, 0.0), (300, 0.3, 2.5, 0.4)])
def test_nlist_rows_equal_oracle(H, N, rho, rcut, buff):
    pos, L = random_fluid(N, rho, seed=N + 1)
    nl, nn, cond = nlist(pos, L, rcut, buff)
    assert cond[3] == 0
    onl, onn = O.nlist_binned(O.Box(L), pos, rcut, buff, full=True)
    assert (nn == onn).all()
    for i in range(N):
        assert (nl[i, :nn[i]] == onl[i, :onn[i]]).all()


def test_true_pair_list(H, golden):
    fx = golden("true_pair_list.json")
    pos, L = sc_lattice(fx["lattice"]["n"], fx["lattice"]["a"])
    nl, nn, cond = nlist(pos, L, fx["r_cut"], fx["buffer"])
    assert O.pair_set(nl, nn) == {tuple(sorted(p)) for p in fx["pairs"]}
    assert nn.sum() == 2 * len(fx["pairs"])


def test_nlist_fcc_lattice_bit_exact(H):
    """Perfect lattice: many exactly equal distances; the FP64 test must agree entry for entry."""
    from lattices import fcc_lattice
    pos, L = fcc_lattice(12)
    nl, nn, _ = nlist(pos, L, 2.5, 0.4)
    onl, onn = O.nlist_binned(O.Box(L), pos, 2.5, 0.4, full=True)
    assert (nn == onn).all()
    assert np.array_equal(O.pair_array(nl, nn), O.pair_array(onl, onn))


def test_nlist_exclusions_and_types(H):
    pos, L = random_fluid(3000, 0.8, seed=5)
    N = len(pos)
    excl = [[(i + 1) % N] if i % 2 == 0 else [] for i in range(N)]
    for i in range(0, N, 2):
        excl[(i + 1) % N].append(i)
    nl, nn, _ = nlist(pos, L, 2.5, 0.4, excl=excl)
    onl, onn = O.nlist_binned(O.Box(L), pos, 2.5, 0.4, full=True, excl=excl)
    assert (nn == onn).all() and O.pair_set(nl, nn) == O.pair_set(onl, onn)
    types = (np.arange(N) % 2).astype(np.uint32)
    rc = np.array([[2.5, 1.2], [1.2, 0.0]])
    nl, nn, _ = nlist(pos, L, rc, 0.3, types=types, ntypes=2)
    onl, onn = O.nlist_binned(O.Box(L), pos, rc, 0.3, full=True, types=types, ntypes=2)
    assert (nn == onn).all()
    for i in range(N):
        assert (nl[i, :nn[i]] == onl[i, :onn[i]]).all()


def test_nlist_overflow_reports_needed_size(H):
    pos, L = random_fluid(2000, 0.84, seed=8)
    _, nn, cond = nlist(pos, L, 2.5, 0.4, nmax=8)
    assert cond[3] == nn.max() and nn.max() > 8


def test_distance_check(H):
    pos, L = random_fluid(5000, 0.8, seed=3)
    last = pos.copy()
    moved = pos.copy()
    moved[1234, 2] += 0.2001
    for p, expect in ((pos, 0), (moved, 1)):
        d_pos, d_last = dev(pos4(p)), dev(last)
        flag = H.DeviceArray((1,), np.uint32, fill=0)
        H.check(H.lib().hm_nlist_check(len(p), d_pos.ptr, d_last.ptr, H.box(L), 2.5, 0.4,
                                       flag.ptr, None))
        H.check(H.lib().hm_synchronize())
        assert flag.numpy()[0] == expect
        assert O.distance_check(O.Box(L), p, last, 2.5, 0.4) == bool(expect)


# ------------------------------------------------------------------ pair forces
def _fp64_forces(pos, L, nl, nn, **kw):
    return O.pair_lj(O.Box(L), pos, nl, nn, third_law=False, **kw)


@pytest.mark.parametrize("mode", ["none", "shift", "xplor"])
@pytest.mark.parametrize("tpp", [1, 4, 16])
def test_pair_lj_within_tolerance_of_fp64(H, mode, tpp):
    pos, L = perturbed_fcc(16, 0.8442, r=0.08, seed=tpp)
    nl, nn, _ = nlist(pos, L, 2.5, 0.4)
    m = {"none": 0, "shift": 1, "xplor": 2}[mode]
    f, v = pair_lj(pos, L, nl, nn, mode=m, r_on=2.0, tpp=tpp, block=256)
    rf, rv = _fp64_forces(pos, L, nl, nn, mode=mode, r_on=2.0)
    errs = {"force": rel_rms(f[:, :3], rf[:, :3]), "energy": rel_rms(f[:, 3], rf[:, 3]),
            "virial": rel_rms(v, rv)}
    print(f"pair lj mode={mode} tpp={tpp} rel RMS: {errs}")
    for k, e in errs.items():
        assert e < FORCE_TOL, (k, e)


@pytest.mark.parametrize("tpp,block", [(2, 64), (8, 512), (4, 1024)])
def test_pair_lj_launch_shapes_agree(H, tpp, block):
    pos, L = perturbed_fcc(10, 0.8442, r=0.05, seed=4)
    nl, nn, _ = nlist(pos, L, 2.5, 0.4)
    f0, v0 = pair_lj(pos, L, nl, nn, tpp=1, block=256)
    f1, v1 = pair_lj(pos, L, nl, nn, tpp=tpp, block=block)
    assert rel_rms(f1, f0) < 1e-6 and rel_rms(v1, v0) < 1e-6


def test_pair_lj_golden_two_particles(H, golden):
    fx = golden("lj_forces_and_energies.json")
    lj = fx["LJ"]
    for p, forces, energies in zip(lj["params"], lj["forces"], lj["energies"]):
        for i, d in enumerate(fx["distances"]):
            pos = np.array([[0.0, 0.0, 0.1], [0.0, 0.0, d + 0.1]])
            nl = np.array([[1] + [0] * 7, [0] * 8], np.uint32)
            nn = np.array([1, 1], np.uint32)
            f, _ = pair_lj(pos, 20.0, nl, nn, r_cut=fx["r_cut"], sigma=p["sigma"],
                           epsilon=p["epsilon"])
            assert math.isclose(f[:, 3].sum(), energies[i], rel_tol=5e-6, abs_tol=1e-6)
            assert math.isclose(f[0, 2], forces[i], rel_tol=5e-6, abs_tol=1e-6)


def test_pair_lj_multitype(H):
    pos, L = random_fluid(4000, 0.8, seed=6)
    pos = perturbed_fcc(10, 0.8442, r=0.05, seed=6)[0]
    L = perturbed_fcc(10, 0.8442, r=0.05, seed=6)[1]
    N = len(pos)
    types = (np.arange(N) % 2).astype(np.uint32)
    rc = np.array([[2.5, 2.0], [2.0, 1.5]])
    params = [(1.0, 1.0), (0.9, 0.5), (0.9, 0.5), (0.8, 1.5)]
    nl, nn, _ = nlist(pos, L, rc, 0.4, types=types, ntypes=2)
    f, v = pair_lj(pos, L, nl, nn, r_cut=rc, types=types, ntypes=2, params=params)
    # FP64 re-evaluation per type pair with the oracle evaluator
    ref = np.zeros((N, 4))
    for i in range(N):
        for j in nl[i, :nn[i]]:
            t = types[i] * 2 + types[j]
            dx = O.Box(L).min_image(pos[i] - pos[j])
            ok, fd, e = O.lj(dx @ dx, rc.ravel()[t] ** 2, *params[t])
            if ok:
                ref[i, :3] += fd * dx
                ref[i, 3] += 0.5 * e
    assert rel_rms(f[:, :3], ref[:, :3]) < FORCE_TOL
    assert rel_rms(f[:, 3], ref[:, 3]) < FORCE_TOL


# ------------------------------------------------------------------ bonds
def _bond_tables(N, bonds):
    deg = np.zeros(N + 1, np.uint32)
    for a, b in bonds:
        deg[a] += 1
        deg[b] += 1
    row = np.zeros(N + 1, np.uint32)
    row[1:] = np.cumsum(deg[:N])
    other = np.zeros(max(1, row[-1]), np.uint32)
    fill = row[:-1].copy()
    for a, b in bonds:
        other[fill[a]] = b
        fill[a] += 1
        other[fill[b]] = a
        fill[b] += 1
    return row, other


def _gpu_fene(H, pos, L, bonds, p, perm=None):
    N = len(pos)
    tag = np.arange(N, dtype=np.uint32) if perm is None else perm.astype(np.uint32)
    rtag = np.argsort(tag).astype(np.uint32)
    row, other = _bond_tables(N, bonds)
    bufs = [dev(pos4(pos)), dev(tag), dev(rtag), dev(row), dev(other)]
    d_f = H.DeviceArray((N * 4,), np.float32, fill=0)
    d_v = H.DeviceArray((N * 6,), np.float32, fill=0)
    flag = H.DeviceArray((1,), np.uint32, fill=0)
    a = H.HmBondArgs()
    a.N, a.pos, a.box = N, bufs[0].ptr, H.box(L)
    a.tag, a.rtag, a.bond_row, a.bond_other, a.bond_type = bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, bufs[4].ptr, None
    a.nbond_types, a.compute_virial, a.force, a.virial, a.flag = 1, 1, d_f.ptr, d_v.ptr, flag.ptr
    P = H.HmFeneParams(p["k"], p["r0"], p["epsilon"], p["sigma"], p["delta"])
    H.check(H.lib().hm_bond_fene(C.byref(a), C.byref(P), None))
    H.check(H.lib().hm_synchronize())
    return d_f.numpy().reshape(N, 4).astype(float), d_v.numpy().reshape(6, N).astype(float), flag.numpy()[0]


def test_bond_fene_golden(H, golden):
    fx = golden("fene_bond.json")
    d = fx["d"]
    pos = np.array([[-d / 2, 0, 0.1], [d / 2, 0, 0.1]])
    for c in fx["cases"]:
        f, _, flag = _gpu_fene(H, pos, fx["L"], [(0, 1)], c["params"])
        assert flag == 0
        assert math.isclose(f[:, 3].sum(), c["energy"], rel_tol=fx["rtol"])
        assert np.allclose(f[0, :3], [c["force"], 0, 0], rtol=fx["rtol"], atol=1e-5)
        assert np.allclose(f[1, :3], [-c["force"], 0, 0], rtol=fx["rtol"], atol=1e-5)


def test_bond_fene_chains_vs_fp64(H):
    rng = np.random.default_rng(0)
    nch, M = 200, 50
    N = nch * M
    L = (N / 0.85) ** (1 / 3)
    pos = np.zeros((N, 3))
    bonds = []
    for c in range(nch):
        x = rng.uniform(-L / 2, L / 2, 3)
        for k in range(M):
            i = c * M + k
            if k:
                step = rng.normal(size=3)
                x = x + 0.97 * step / np.linalg.norm(step)
                bonds.append((i - 1, i))
            pos[i] = x
    pos = (pos + L / 2) % L - L / 2
    p = dict(k=30.0, r0=1.5, epsilon=1.0, sigma=1.0, delta=0.0)
    perm = rng.permutation(N)  # arbitrary tag order: bonds are by tag
    pos_idx = pos[perm]
    f, v, flag = _gpu_fene(H, pos_idx, L, bonds, p, perm=perm)
    assert flag == 0
    r, rf, rv = O.bond_fene(O.Box(L), pos, bonds, **p)
    assert r == 0
    f_tag = np.zeros_like(f)
    f_tag[perm] = f
    v_tag = np.zeros_like(v)
    v_tag[:, perm] = v
    assert rel_rms(f_tag[:, :3], rf[:, :3]) < FORCE_TOL
    assert rel_rms(f_tag[:, 3], rf[:, 3]) < FORCE_TOL
    assert rel_rms(v_tag, rv) < FORCE_TOL


def test_bond_fene_stretched_flag(H):
    pos = np.array([[-1.0, 0, 0], [1.0, 0, 0]])
    _, _, flag = _gpu_fene(H, pos, 10.0, [(0, 1)], dict(k=30.0, r0=1.5, epsilon=1.0, sigma=1.0,
                                                        delta=0.0))
    assert flag == 1


# ------------------------------------------------------------------ integrators (FP64, bit-exact)
def test_nvt_steps_bit_exact(H):
    rng = np.random.default_rng(1)
    N = 10000
    pos, L = random_fluid(N, 0.8, seed=9)
    vel = rng.normal(size=(N, 3))
    acc = rng.normal(size=(N, 3)) * 10
    img = rng.integers(-3, 3, size=(N, 3)).astype(np.int32)
    mass = rng.uniform(0.5, 2.0, N)
    dt, alpha = 0.005, 0.987
    v4 = np.zeros((N, 4))
    v4[:, :3], v4[:, 3] = vel, mass
    d_pos, d_vel, d_acc, d_img = dev(pos4(pos)), dev(v4), dev(acc), dev(img)
    H.check(H.lib().hm_nvt_step_one(N, d_pos.ptr, d_vel.ptr, d_acc.ptr, d_img.ptr, H.box(L), dt,
                                    alpha, None))
    H.check(H.lib().hm_synchronize())
    p1, v1, i1 = O.nvt_step_one(O.Box(L), pos, vel, acc, img, dt, alpha)
    assert np.array_equal(d_pos.numpy()[:, :3], p1)
    assert np.array_equal(d_vel.numpy()[:, :3], v1)
    assert np.array_equal(d_img.numpy(), i1)
    f = (rng.normal(size=(N, 4)) * 5).astype(np.float32)
    d_f = dev(f)
    ptrs = H.pointer_array([d_f.ptr])
    H.check(H.lib().hm_nvt_step_two(N, d_vel.ptr, d_acc.ptr, ptrs, 1, dt, 1.0, None))
    H.check(H.lib().hm_synchronize())
    v2, a2 = O.nvt_step_two(v1, mass, f.astype(np.float64), dt, 1.0)
    assert np.array_equal(d_vel.numpy()[:, :3], v2)
    assert np.array_equal(d_acc.numpy(), a2)


def test_langevin_step_two_bit_exact(H):
    rng = np.random.default_rng(2)
    N = 5000
    pos, L = random_fluid(N, 0.8, seed=10)
    vel = rng.normal(size=(N, 3))
    mass = np.ones(N)
    tag = rng.permutation(N).astype(np.uint32)
    f = (rng.normal(size=(N, 4)) * 5).astype(np.float32)
    v4 = np.zeros((N, 4))
    v4[:, :3], v4[:, 3] = vel, mass
    d_pos, d_vel, d_acc, d_tag, d_f = dev(pos4(pos)), dev(v4), dev(np.zeros((N, 3))), dev(tag), dev(f)
    gamma = np.array([1.0])
    d_g = dev(gamma)
    H.check(H.lib().hm_langevin_step_two(N, d_vel.ptr, d_acc.ptr, d_pos.ptr, d_tag.ptr,
                                         H.pointer_array([d_f.ptr]), 1, d_g.ptr, 1, 1.0, 0.01,
                                         12345, 7, None))
    H.check(H.lib().hm_synchronize())
    v2, a2 = O.langevin_step_two(vel, mass, tag, f.astype(np.float64), gamma, 1.0, 0.01, 12345, 7)
    assert np.array_equal(d_vel.numpy()[:, :3], v2)
    assert np.array_equal(d_acc.numpy(), a2)


def test_bussi_host_matches_oracle(H):
    for t in (0, 1, 500, 10**6):
        for K, g in ((12.0, 30.0), (15000.0, 3 * 16384 - 3.0)):
            a = H.lib().hm_bussi_alpha_host(K, g, 0.005, 0.0, 1.2, t, 5, 0)
            b = O.bussi_alpha(K, g, 0.005, 0.0, 1.2, t, 5, 0)
            assert a == b


def test_thermalize_matches_oracle(H):
    N = 20000
    mass = np.ones(N)
    tag = np.random.default_rng(3).permutation(N).astype(np.uint32)
    v4 = np.zeros((N, 4))
    v4[:, 3] = mass
    d_vel, d_tag = dev(v4), dev(tag)
    H.check(H.lib().hm_thermalize(N, d_vel.ptr, d_tag.ptr, 1.2, 0, 42, None))
    H.check(H.lib().hm_synchronize())
    got = d_vel.numpy()[:, :3]
    ref = O.thermalize(mass, tag, 1.2, 0, 42)
    # device sin/log vs the host libm may differ in the last bit; the COM sum order differs
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
    assert np.allclose(got.sum(axis=0), 0.0, atol=1e-9)


def test_thermo_sums(H):
    rng = np.random.default_rng(4)
    N = 30000
    v4 = np.zeros((N, 4))
    v4[:, :3] = rng.normal(size=(N, 3))
    v4[:, 3] = rng.uniform(0.5, 2, N)
    f = rng.normal(size=(N, 4)).astype(np.float32)
    vir = rng.normal(size=(6, N)).astype(np.float32)
    d_v, d_f, d_vir = dev(v4), dev(f), dev(vir)
    out = H.DeviceArray((15,), np.float64, fill=0)
    H.check(H.lib().hm_thermo(N, d_v.ptr, H.pointer_array([d_f.ptr]), H.pointer_array([d_vir.ptr]),
                              1, out.ptr, None))
    H.check(H.lib().hm_synchronize())
    o = out.numpy()
    ke = 0.5 * (v4[:, 3] * (v4[:, :3] ** 2).sum(axis=1)).sum()
    assert math.isclose(o[0], ke, rel_tol=1e-12)
    assert math.isclose(o[1], f[:, 3].astype(float).sum(), rel_tol=1e-9, abs_tol=1e-9)
    vs = vir.astype(float).sum(axis=1)
    assert np.allclose(o[3:9], vs, rtol=1e-9, atol=1e-9)
    assert math.isclose(o[2], (vs[0] + vs[3] + vs[5]) / 3, rel_tol=1e-9, abs_tol=1e-9)
    m, v = v4[:, 3], v4[:, :3]
    kin = [np.sum(m * v[:, a] * v[:, b]) for a, b in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))]
    assert np.allclose(o[9:15], kin, rtol=1e-12, atol=1e-9)
    assert math.isclose(o[9] + o[12] + o[14], 2 * o[0], rel_tol=1e-12)


# ------------------------------------------------------------------ Hilbert sort
@pytest.mark.parametrize("grid", [16, 64])
def test_hilbert_traversal_matches_oracle(H, grid):
    out = H.DeviceArray((grid**3,), np.uint32)
    H.check(H.lib().hm_sfc_traversal(grid, out.ptr, None))
    H.check(H.lib().hm_synchronize())
    assert np.array_equal(out.numpy(), O.hilbert_traversal(grid))


@pytest.mark.parametrize("grid", [8, 256])
def test_sfc_order_matches_oracle(H, grid):
    pos, L = random_fluid(50000, 0.84, seed=grid)
    d_pos = dev(pos4(pos))
    order = H.DeviceArray((len(pos),), np.uint32)
    H.check(H.lib().hm_sfc_order(len(pos), d_pos.ptr, H.box(L), grid, order.ptr, None))
    H.check(H.lib().hm_synchronize())
    assert np.array_equal(order.numpy(), O.sort_order(O.Box(L), pos, grid))


# ------------------------------------------------------------------ cell-tiled pair kernel
from gpu_helpers import CellPipeline  # noqa: E402


@pytest.mark.parametrize("mode", ["none", "shift", "xplor"])
@pytest.mark.parametrize("tpp,wg", [(4, 0), (8, 0), (16, 0), (8, 512)])
def test_pair_cells_within_tolerance_of_fp64(H, mode, tpp, wg):
    pos, L = perturbed_fcc(16, 0.8442, r=0.08, seed=tpp + wg)
    cp = CellPipeline(pos, L, 2.5, 0.4)
    nl, nn = cp.lists()
    m = {"none": 0, "shift": 1, "xplor": 2}[mode]
    f, v = cp.forces(pos, mode=m, r_on=2.0, tpp=tpp, workgroups=wg)
    rf, rv = _fp64_forces(pos, L, nl, nn, mode=mode, r_on=2.0)
    errs = {"force": rel_rms(f[:, :3], rf[:, :3]), "energy": rel_rms(f[:, 3], rf[:, 3]),
            "virial": rel_rms(v, rv)}
    for k, e in errs.items():
        assert e < FORCE_TOL, (k, e)


def test_pair_cells_after_motion_and_wrap(H):
    """Lists built at one configuration, forces at a later one (every displacement below half the
    buffer, particles wrapped across the box edges): the cell-tiled kernel must use the current
    positions in the right periodic frame."""
    pos, L = perturbed_fcc(12, 0.8442, r=0.05, seed=21)
    cp = CellPipeline(pos, L, 2.5, 0.4)
    nl, nn = cp.lists()
    rng = np.random.default_rng(21)
    step = rng.normal(size=pos.shape)
    step *= 0.19 / np.linalg.norm(step, axis=1)[:, None] * rng.uniform(0, 1, (len(pos), 1))
    moved = pos + step
    moved = (moved + L / 2) % L - L / 2
    assert not O.distance_check(O.Box(L), moved, pos, 2.5, 0.4)
    f, v = cp.forces(moved)
    rf, rv = _fp64_forces(moved, L, nl, nn)
    assert rel_rms(f[:, :3], rf[:, :3]) < FORCE_TOL
    assert rel_rms(f[:, 3], rf[:, 3]) < FORCE_TOL
    assert rel_rms(v, rv) < FORCE_TOL


def test_pair_cells_fallback_path(H):
    """More than 1024 candidates per home cell (long cutoff at liquid density) takes the
    global-gather fallback inside the same kernel."""
    pos, L = perturbed_fcc(10, 0.8442, r=0.05, seed=5)
    cp = CellPipeline(pos, L, 3.3, 0.4, cmax=96)
    nl, nn = cp.lists()
    assert cp.cond[3] == 0
    f, v = cp.forces(pos)
    rf, rv = O.pair_lj(O.Box(L), pos, nl, nn, third_law=False, r_cut=3.3)
    assert rel_rms(f[:, :3], rf[:, :3]) < FORCE_TOL
    assert rel_rms(f[:, 3], rf[:, 3]) < FORCE_TOL
