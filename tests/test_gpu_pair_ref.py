"""The reference-layout pair driver (include/hoomd_amd_pair.h) on buffers laid out exactly as
PotentialPairGPU hands them to gpu_compute_pair_forces (hoomd/md/PotentialPairGPU.cuh:44-98):
Scalar4 positions with the type as int bits, a head list with per-type row capacities, a padded
virial pitch, device r_cut^2 / r_on^2 matrices for 3 types (one type pair switched off).

References: FP64 LJ per type pair (EvaluatorPairLJ's closed form, pinned by the reference's
forces_and_energies fixtures in the oracle tests) over the oracle's binned full list; the
oracle's XPLOR pair loop; in a tilted box, the nearest of the 27 images found by brute force.
Bar: 1e-5 relative RMS per quantity (force, energy, virial). The example plugin, built as its
own library from one evaluator and one instantiation line, is checked against the closed form
and parameter grid of the reference's example_plugins/pair_plugin/pytest/test_example_pair.py.
"""

import ctypes as C
import itertools
import os

import numpy as np
import pytest

import oracle as O
from gpu_helpers import rel_rms
from lattices import perturbed_fcc

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def H():
    from hoomd_amd import _lib
    if _lib.lib().hm_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return _lib


class BoxRef(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("L", C.c_double * 3), ("xy", C.c_double),
                ("xz", C.c_double), ("yz", C.c_double), ("periodic", C.c_int * 3)]


class PairRefArgs(C.Structure):
    _fields_ = [("d_force", C.c_void_p), ("d_virial", C.c_void_p), ("virial_pitch", C.c_size_t),
                ("N", C.c_uint), ("n_max", C.c_uint), ("d_pos", C.c_void_p),
                ("d_charge", C.c_void_p), ("box", BoxRef), ("d_n_neigh", C.c_void_p),
                ("d_nlist", C.c_void_p), ("d_head_list", C.c_void_p), ("d_rcutsq", C.c_void_p),
                ("d_ronsq", C.c_void_p), ("size_neigh_list", C.c_size_t), ("ntypes", C.c_uint),
                ("block_size", C.c_uint), ("shift_mode", C.c_uint),
                ("compute_virial", C.c_uint), ("threads_per_particle", C.c_uint)]


def scalar4_positions(pos, types):
    """Scalar4 (x, y, z, __int_as_scalar(type)): the int in the low 32 bits of w."""
    p4 = np.zeros((len(pos), 4))
    p4[:, :3] = pos
    p4[:, 3] = np.asarray(types, np.int64).view(np.float64)
    return p4


def head_list_layout(nl, nn, types, ntypes):
    """NeighborList::buildHeadList: each particle's row holds Nmax[type] entries, Nmax the
    largest count of that type rounded up to a multiple of 4 (plus a spare block of 4)."""
    nmax_t = np.zeros(ntypes, np.int64)
    for t in range(ntypes):
        sel = types == t
        nmax_t[t] = (int(nn[sel].max()) + 3) // 4 * 4 + 4 if sel.any() else 4
    caps = nmax_t[types]
    head = np.concatenate([[0], np.cumsum(caps)[:-1]]).astype(np.uint64)
    flat = np.full(int(caps.sum()), 0xFFFFFFFF, np.uint32)  # unused slots hold garbage
    for i in range(len(nn)):
        flat[head[i]:head[i] + nn[i]] = nl[i, :nn[i]]
    return flat, head


def run_driver(H, fn, pos, types, L, nl, nn, rcut, ron, params, shift_mode, tpp, block,
               ntypes, tilt=(0.0, 0.0, 0.0), virial=True):
    N = len(pos)
    flat, head = head_list_layout(nl, nn, np.asarray(types), ntypes)
    pitch = (N + 63) // 64 * 64 + 64
    d_pos = H.DeviceArray.from_numpy(scalar4_positions(pos, types))
    d_nl = H.DeviceArray.from_numpy(flat)
    d_head = H.DeviceArray.from_numpy(head)
    d_nn = H.DeviceArray.from_numpy(np.asarray(nn, np.uint32))
    d_rc = H.DeviceArray.from_numpy((np.asarray(rcut, np.float64) ** 2).ravel())
    d_ron = H.DeviceArray.from_numpy((np.asarray(ron, np.float64) ** 2).ravel())
    d_par = H.DeviceArray.from_numpy(np.asarray(params, np.float64))
    d_f = H.DeviceArray.from_numpy(np.full((N, 4), np.nan))
    d_v = H.DeviceArray.from_numpy(np.full(6 * pitch, np.nan))
    a = PairRefArgs()
    a.d_force, a.d_virial, a.virial_pitch = d_f.ptr, d_v.ptr, pitch
    a.N, a.n_max, a.d_pos, a.d_charge = N, N, d_pos.ptr, None
    Ls = np.broadcast_to(np.asarray(L, np.float64), (3,))
    for k in range(3):
        a.box.L[k] = Ls[k]
        a.box.lo[k] = -Ls[k] / 2
        a.box.periodic[k] = 1
    a.box.xy, a.box.xz, a.box.yz = tilt
    a.d_n_neigh, a.d_nlist, a.d_head_list = d_nn.ptr, d_nl.ptr, d_head.ptr
    a.d_rcutsq, a.d_ronsq, a.size_neigh_list = d_rc.ptr, d_ron.ptr, len(flat)
    a.ntypes, a.block_size, a.shift_mode = ntypes, block, shift_mode
    a.compute_virial, a.threads_per_particle = int(virial), tpp
    H.check(fn(C.byref(a), C.c_void_p(d_par.ptr), None), "pair ref driver")
    H.check(H.lib().hm_synchronize())
    return d_f.numpy(), d_v.numpy().reshape(6, pitch)[:, :N]


def lj_pairs_fp64(pos, types, L, nl, nn, rcut, sig_eps, shift):
    """FP64 force, energy and virial of LJ per type pair over the list (min image, cubic box)."""
    N = len(pos)
    i = np.repeat(np.arange(N), nn)
    j = np.concatenate([nl[k, :nn[k]] for k in range(N)]).astype(np.int64)
    dx = pos[i] - pos[j]
    dx -= L * np.rint(dx / L)
    return pair_sums(N, i, j, dx, types, rcut, sig_eps, shift)


def pair_sums(N, i, j, dx, types, rcut, sig_eps, shift):
    rsq = (dx * dx).sum(axis=1)
    tp = types[i] * len(rcut) + types[j]
    rcsq = np.asarray(rcut, np.float64).ravel()[tp] ** 2
    sig, eps = np.asarray(sig_eps)[tp, 0], np.asarray(sig_eps)[tp, 1]
    lj1 = 4 * eps * sig ** 12
    lj2 = 4 * eps * sig ** 6
    inr = (rsq < rcsq) & (lj1 != 0)
    r2i = np.where(inr, 1.0 / np.where(rsq > 0, rsq, 1.0), 0.0)
    r6i = r2i ** 3
    fdr = np.where(inr, r2i * r6i * (12 * lj1 * r6i - 6 * lj2), 0.0)
    e = np.where(inr, r6i * (lj1 * r6i - lj2), 0.0)
    if shift:
        rc2i = 1.0 / np.where(rcsq > 0, rcsq, 1.0)
        rc6i = rc2i ** 3
        e -= np.where(inr, rc6i * (lj1 * rc6i - lj2), 0.0)
    f = np.zeros((N, 4))
    np.add.at(f[:, :3], i, dx * fdr[:, None])
    np.add.at(f[:, 3], i, 0.5 * e)
    v = np.zeros((6, N))
    for m, (a, b) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))):
        np.add.at(v[m], i, 0.5 * fdr * dx[:, a] * dx[:, b])
    return f, v


def sig_eps_table(ntypes):
    rng = np.random.default_rng(3)
    t = np.zeros((ntypes, ntypes, 2))
    for a in range(ntypes):
        for b in range(a, ntypes):
            t[a, b] = t[b, a] = (rng.uniform(0.8, 1.1), rng.uniform(0.5, 1.5))
    return t.reshape(-1, 2)


def lj_param_ref(sig_eps):
    """EvaluatorPairLJ::param_type records (sigma^6, 4 epsilon)."""
    return np.stack([sig_eps[:, 0] ** 6, 4 * sig_eps[:, 1]], axis=1)


@pytest.mark.parametrize("tpp,block,shift", [(1, 256, 0), (4, 128, 1), (32, 1024, 0),
                                             (8, 64, 1)])
def test_lj_reference_layout_three_types(H, tpp, block, shift):
    pos, L = perturbed_fcc(9, 0.8442, r=0.05, seed=17)  # 2916 particles
    N = len(pos)
    types = (np.arange(N) * 7 % 3).astype(np.int64)
    rcut = np.array([[2.5, 2.0, 1.8], [2.0, 1.5, 0.0], [1.8, 0.0, 2.2]])
    ron = np.zeros_like(rcut)
    se = sig_eps_table(3)
    nl, nn = O.nlist_binned(O.Box(L), pos, rcut, 0.4, full=True,
                            types=types.astype(np.uint32), ntypes=3)
    f, v = run_driver(H, H.lib().hm_compute_pair_forces_lj, pos, types, L, nl, nn, rcut, ron,
                      lj_param_ref(se), shift, tpp, block, 3)
    rf, rv = lj_pairs_fp64(pos, types, L, nl, nn, rcut, se, shift == 1)
    ef, ee, ev = (rel_rms(f[:, :3], rf[:, :3]), rel_rms(f[:, 3], rf[:, 3]), rel_rms(v, rv))
    print(f"tpp {tpp} block {block} shift {shift}: force {ef:.2e} energy {ee:.2e} "
          f"virial {ev:.2e}", flush=True)
    assert np.isfinite(f).all() and np.isfinite(v).all()
    assert ef < TOL and ee < TOL and ev < TOL


def test_lj_reference_layout_xplor_against_oracle(H):
    pos, L = perturbed_fcc(8, 0.8442, r=0.05, seed=4)
    N = len(pos)
    types = np.zeros(N, np.int64)
    nl, nn = O.nlist_binned(O.Box(L), pos, 2.5, 0.4, full=True)
    f, v = run_driver(H, H.lib().hm_compute_pair_forces_lj, pos, types, L, nl, nn,
                      [[2.5]], [[2.0]], lj_param_ref(np.array([[1.0, 1.0]])), 2, 4, 256, 1)
    rf, rv = O.pair_lj(O.Box(L), pos, nl, nn, third_law=False, r_cut=2.5, r_on=2.0,
                       mode="xplor")
    assert rel_rms(f[:, :3], rf[:, :3]) < TOL
    assert rel_rms(f[:, 3], rf[:, 3]) < TOL
    assert rel_rms(v, rv) < TOL


def test_lj_reference_layout_tilted_box(H):
    """Tilt factors xy, xz, yz: the kernel's minimum image must pick the nearest image, found
    here by brute force over the 27 images of each pair."""
    rng = np.random.default_rng(8)
    Lb = np.array([12.0, 12.5, 13.0])
    xy, xz, yz = 0.2, -0.1, 0.15
    Hm = np.array([[Lb[0], xy * Lb[1], xz * Lb[2]], [0, Lb[1], yz * Lb[2]], [0, 0, Lb[2]]])
    n = np.array([10, 10, 12])
    g = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3)
    frac = (g + 0.5) / n - 0.5 + rng.uniform(-0.02, 0.02, g.shape)
    N = len(frac)
    pos = frac @ Hm.T
    shifts = np.array(list(itertools.product((-1, 0, 1), repeat=3))) @ Hm.T
    i, j = np.triu_indices(N, 1)
    d0 = pos[i] - pos[j]
    best = np.full(len(i), np.inf)
    dx = np.zeros_like(d0)
    for s in shifts:
        d = d0 + s
        r = (d * d).sum(axis=1)
        upd = r < best
        best[upd] = r[upd]
        dx[upd] = d[upd]
    keep = best < (2.5 + 0.4) ** 2
    ii = np.concatenate([i[keep], j[keep]])
    jj = np.concatenate([j[keep], i[keep]])
    dd = np.concatenate([dx[keep], -dx[keep]])
    order = np.argsort(ii, kind="stable")
    ii, jj, dd = ii[order], jj[order], dd[order]
    nn = np.bincount(ii, minlength=N).astype(np.uint32)
    nl = np.zeros((N, int(nn.max())), np.uint32)
    start = np.concatenate([[0], np.cumsum(nn.astype(np.int64))[:-1]])
    for k in range(N):
        nl[k, :nn[k]] = jj[start[k]:start[k] + nn[k]]
    types = np.zeros(N, np.int64)
    se = np.array([[0.6, 1.0]])
    f, v = run_driver(H, H.lib().hm_compute_pair_forces_lj, pos, types, Lb, nl, nn, [[2.5]],
                      [[0.0]], lj_param_ref(se), 0, 4, 256, 1, tilt=(xy, xz, yz))
    rf, rv = pair_sums(N, ii, jj, dd, types, [[2.5]], se, False)
    assert rel_rms(f[:, :3], rf[:, :3]) < TOL
    assert rel_rms(f[:, 3], rf[:, 3]) < TOL
    assert rel_rms(v, rv) < TOL


def test_driver_rejects_bad_launch(H):
    a = PairRefArgs()
    a.N, a.ntypes, a.block_size, a.threads_per_particle = 10, 1, 100, 4
    assert H.lib().hm_compute_pair_forces_lj(C.byref(a), C.c_void_p(16), None) == -2


# ------------------------------------------------------------------ plugin
def harm_force_and_energy(dx, k, sigma, r_cut, shift=False):
    """example_plugins/pair_plugin/pytest/test_example_pair.py:16-28"""
    dr = np.linalg.norm(dx)
    if dr >= r_cut:
        return np.zeros(3), 0.0
    f = k * (sigma - dr) * np.asarray(dx, np.float64) / dr
    e = 0.5 * k * (sigma - dr) ** 2
    if shift:
        e -= 0.5 * k * (r_cut - sigma) ** 2
    return f, e


@pytest.fixture(scope="module")
def plugin(H):
    from conftest import ROOT
    path = os.path.join(ROOT, "hoomd-blue_amd", "plugins", "pair_example", "libpair_example.so")
    lib = C.CDLL(path)
    fn = lib.hm_compute_pair_forces_example
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    return fn


@pytest.mark.parametrize("distance, k, sigma, mode", list(itertools.product(
    np.linspace(0.1, 2.0, 3), [0.5, 2.0, 5.0], [0.5, 1.0, 1.5], ["none", "shift"])))
def test_example_plugin_two_particles(H, plugin, distance, k, sigma, mode):
    pos = np.array([[-distance / 2, 0.0, 0.0], [distance / 2, 0.0, 0.0]])
    nl = np.array([[1], [0]], np.uint32)
    nn = np.array([1, 1], np.uint32)
    f, _ = run_driver(H, plugin, pos, np.zeros(2, np.int64), 20.0, nl, nn, [[sigma]], [[0.0]],
                      np.array([[k, sigma]]), 1 if mode == "shift" else 0, 1, 64, 1)
    ref_f, e = harm_force_and_energy(pos[1] - pos[0], k, sigma, sigma, mode == "shift")
    np.testing.assert_array_almost_equal(f[:, :3], [-ref_f, ref_f], decimal=6)
    np.testing.assert_array_almost_equal(f[:, 3], [e / 2, e / 2], decimal=6)
