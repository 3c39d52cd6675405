"""Helpers that drive the HIP kernels through the C ABI for the parity tests."""

import ctypes as C

import numpy as np

from hoomd_amd import _lib as H


def dev(a, dtype=None):
    return H.DeviceArray.from_numpy(a, dtype)


def pos4(pos, types=None):
    p = np.zeros((len(pos), 4))
    p[:, :3] = pos
    if types is not None:
        p[:, 3] = types
    return p


def cell_list(pos, L, width, cmax=64):
    """Run hm_cell_list_build; returns (dim, cell_size, entries structured, cond, device bufs)."""
    L3 = np.broadcast_to(np.asarray(L, np.float64), (3,))
    b = H.box(L3)
    dim = (C.c_uint * 3)()
    H.check(H.lib().hm_cell_dims(C.byref(b), width, dim))
    ncell = dim[0] * dim[1] * dim[2]
    d_pos = dev(pos4(pos))
    d_size = H.DeviceArray((ncell,), np.uint32)
    d_ent = H.DeviceArray((ncell * cmax * 4,), np.float64)  # 32-byte entries
    d_last = H.DeviceArray((len(pos) * 3,), np.float64)
    d_cond = H.DeviceArray((3,), np.uint32, fill=0)
    a = H.HmCellArgs()
    a.N = len(pos)
    a.pos = d_pos.ptr
    a.box = b
    for k in range(3):
        a.dim[k] = dim[k]
    a.cmax = cmax
    a.cell_size = d_size.ptr
    a.entries = d_ent.ptr
    a.last_pos = d_last.ptr
    a.conditions = d_cond.ptr
    H.check(H.lib().hm_cell_list_build(C.byref(a), None), "cell")
    H.check(H.lib().hm_synchronize())
    return dict(dim=tuple(dim), ncell=ncell, cmax=cmax, b=b, d_pos=d_pos, d_size=d_size,
                d_ent=d_ent, d_last=d_last, d_cond=d_cond, N=len(pos))


def entries_idx(cl):
    """cell_size and per-cell sorted index lists from a cell_list() result."""
    size = cl["d_size"].numpy()
    raw = cl["d_ent"].numpy().view(np.uint32).reshape(cl["ncell"], cl["cmax"], 8)
    idx = raw[:, :, 6]
    cells = [sorted(idx[c, :min(size[c], cl["cmax"])].tolist()) for c in range(cl["ncell"])]
    return size, cells


def nlist(pos, L, r_cut, r_buff, nmax=None, cmax=64, excl=None, types=None, ntypes=1):
    """Cell list + full neighbour list on the GPU. Returns (nlist[N, nmax], n_neigh, cond)."""
    N = len(pos)
    rc = np.broadcast_to(np.asarray(r_cut, np.float64), (ntypes, ntypes))
    width = rc.max() + r_buff
    L3 = np.broadcast_to(np.asarray(L, np.float64), (3,))
    b = H.box(L3)
    dim = (C.c_uint * 3)()
    H.check(H.lib().hm_cell_dims(C.byref(b), width, dim))
    ncell = dim[0] * dim[1] * dim[2]
    n_adj = H.lib().hm_cell_n_adj(dim)
    d_pos = dev(pos4(pos, types))
    d_size = H.DeviceArray((ncell,), np.uint32)
    d_ent = H.DeviceArray((ncell * cmax * 4,), np.float64)
    d_adj = H.DeviceArray((ncell * n_adj,), np.uint32)
    d_cond = H.DeviceArray((4,), np.uint32, fill=0)
    H.check(H.lib().hm_cell_adj(dim, d_adj.ptr, None))
    a = H.HmCellArgs()
    a.N, a.pos, a.box = N, d_pos.ptr, b
    for k in range(3):
        a.dim[k] = dim[k]
    a.cmax, a.cell_size, a.entries, a.last_pos, a.conditions = cmax, d_size.ptr, d_ent.ptr, None, d_cond.ptr
    H.check(H.lib().hm_cell_list_build(C.byref(a), None), "cell")
    if nmax is None:
        rho = N / float(np.prod(L3))
        nmax = int(np.ceil(rho * 4 / 3 * np.pi * width**3 * 1.5 + 16))
        nmax = (nmax + 7) & ~7
    d_nl = H.DeviceArray((N * nmax,), np.uint32, fill=0xff)
    d_nn = H.DeviceArray((N,), np.uint32, fill=0)
    keep = [d_pos, d_size, d_ent, d_adj, d_nl, d_nn, d_cond]
    na = H.HmNlistArgs()
    na.N, na.pos, na.box = N, d_pos.ptr, b
    for k in range(3):
        na.dim[k] = dim[k]
    na.cmax, na.cell_size, na.entries, na.cell_adj, na.n_adj = cmax, d_size.ptr, d_ent.ptr, d_adj.ptr, n_adj
    na.ntypes = ntypes
    for q, v in enumerate(rc.ravel()):
        na.r_cut[q] = v
    na.r_buff = r_buff
    if excl is not None:
        mx = max(1, max(len(e) for e in excl))
        ne = np.array([len(e) for e in excl], np.uint32)
        ex = np.full((N, mx), 0xffffffff, np.uint32)
        for i, e in enumerate(excl):
            ex[i, :len(e)] = e
        d_ne, d_ex = dev(ne), dev(ex)
        keep += [d_ne, d_ex]
        na.n_excl, na.excl, na.max_excl = d_ne.ptr, d_ex.ptr, mx
    na.nmax, na.nlist, na.n_neigh = nmax, d_nl.ptr, d_nn.ptr
    na.conditions = d_cond.ptr + 12  # 4th word: needed nmax
    H.check(H.lib().hm_nlist_build(C.byref(na), None), "nlist")
    H.check(H.lib().hm_synchronize())
    cond = d_cond.numpy()
    return d_nl.numpy().reshape(N, nmax), d_nn.numpy(), cond


def pair_lj(pos, L, nl, nn, r_cut=2.5, sigma=1.0, epsilon=1.0, mode=0, r_on=0.0, virial=True,
            tpp=4, block=256, types=None, ntypes=1, params=None):
    N = len(pos)
    nmax = nl.shape[1]
    d_pos = dev(pos4(pos, types))
    d_nl = dev(nl, np.uint32)
    d_nn = dev(nn, np.uint32)
    d_f = H.DeviceArray((N * 4,), np.float32, fill=0)
    d_v = H.DeviceArray((N * 6,), np.float32, fill=0)
    a = H.HmPairArgs()
    a.N, a.pos, a.box = N, d_pos.ptr, H.box(L)
    a.nlist, a.nmax, a.n_neigh, a.ntypes = d_nl.ptr, nmax, d_nn.ptr, ntypes
    rc = np.broadcast_to(np.asarray(r_cut, np.float64), (ntypes, ntypes)).ravel()
    ro = np.broadcast_to(np.asarray(r_on, np.float64), (ntypes, ntypes)).ravel()
    for q in range(ntypes * ntypes):
        a.r_cut[q] = rc[q]
        a.r_on[q] = ro[q]
    a.shift_mode, a.compute_virial = mode, int(virial)
    a.force, a.virial, a.tpp, a.block_size = d_f.ptr, d_v.ptr, tpp, block
    P = (H.HmLJParams * (ntypes * ntypes))()
    for q in range(ntypes * ntypes):
        s, e = (sigma, epsilon) if params is None else params[q]
        P[q].sigma, P[q].epsilon = s, e
    H.check(H.lib().hm_pair_lj(C.byref(a), P, None), "pair")
    H.check(H.lib().hm_synchronize())
    return d_f.numpy().reshape(N, 4).astype(np.float64), d_v.numpy().reshape(6, N).astype(np.float64)


def rel_rms(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.sqrt(np.mean(ref**2))
    return np.sqrt(np.mean((got - ref) ** 2)) / den if den > 0 else np.sqrt(np.mean(got**2))


class CellPipeline:
    """Cell list (with slot index table) + neighbour list (with 16-bit local rows) built on the
    GPU from build_pos; forces can then be evaluated at later positions with the same lists."""

    def __init__(self, build_pos, L, r_cut, r_buff, cmax=64, nmax=None):
        N = len(build_pos)
        self.N, self.L, self.r_cut = N, L, r_cut
        width = r_cut + r_buff
        b = H.box(L)
        dim = (C.c_uint * 3)()
        H.check(H.lib().hm_cell_dims(C.byref(b), width, dim))
        self.dim = tuple(dim)
        ncell = dim[0] * dim[1] * dim[2]
        self.n_adj = H.lib().hm_cell_n_adj(dim)
        self.cmax = cmax
        self.d_pos = dev(pos4(build_pos))
        self.d_size = H.DeviceArray((ncell,), np.uint32)
        self.d_idx = H.DeviceArray((ncell * cmax,), np.uint32, fill=0)
        self.d_ent = H.DeviceArray((ncell * cmax * 4,), np.float64)
        self.d_adj = H.DeviceArray((ncell * self.n_adj,), np.uint32)
        self.d_cond = H.DeviceArray((4,), np.uint32, fill=0)
        H.check(H.lib().hm_cell_adj(dim, self.d_adj.ptr, None))
        a = H.HmCellArgs()
        a.N, a.pos, a.box = N, self.d_pos.ptr, b
        for k in range(3):
            a.dim[k] = dim[k]
        a.cmax, a.cell_size, a.entries, a.last_pos = cmax, self.d_size.ptr, self.d_ent.ptr, None
        a.conditions, a.cell_idx = self.d_cond.ptr, self.d_idx.ptr
        H.check(H.lib().hm_cell_list_build(C.byref(a), None), "cell")
        if nmax is None:
            rho = N / float(L) ** 3
            nmax = int(np.ceil(rho * 4 / 3 * np.pi * width**3 * 1.5 + 16))
            nmax = (nmax + 7) & ~7
        self.nmax = nmax
        self.d_nl = H.DeviceArray((N * nmax,), np.uint32, fill=0)
        self.d_lr = H.DeviceArray((N * nmax,), np.uint16, fill=0)
        self.d_nn = H.DeviceArray((N,), np.uint32, fill=0)
        na = H.HmNlistArgs()
        na.N, na.pos, na.box = N, self.d_pos.ptr, b
        for k in range(3):
            na.dim[k] = dim[k]
        na.cmax, na.cell_size, na.entries = cmax, self.d_size.ptr, self.d_ent.ptr
        na.cell_adj, na.n_adj, na.ntypes = self.d_adj.ptr, self.n_adj, 1
        na.r_cut[0], na.r_buff = r_cut, r_buff
        na.nmax, na.nlist, na.n_neigh = nmax, self.d_nl.ptr, self.d_nn.ptr
        na.conditions, na.local_rows = self.d_cond.ptr + 12, self.d_lr.ptr
        H.check(H.lib().hm_nlist_build(C.byref(na), None), "nlist")
        H.check(H.lib().hm_synchronize())
        self.cond = self.d_cond.numpy()

    def lists(self):
        return (self.d_nl.numpy().reshape(self.N, self.nmax), self.d_nn.numpy())

    def forces(self, pos, mode=0, r_on=0.0, virial=True, tpp=8, workgroups=0):
        N = self.N
        d_pos = dev(pos4(pos))
        d_f = H.DeviceArray((N * 4,), np.float32, fill=0)
        d_v = H.DeviceArray((N * 6,), np.float32, fill=0)
        a = H.HmPairCellArgs()
        a.N, a.pos, a.box = N, d_pos.ptr, H.box(self.L)
        for k in range(3):
            a.dim[k] = self.dim[k]
        a.cmax, a.cell_size, a.cell_idx = self.cmax, self.d_size.ptr, self.d_idx.ptr
        a.cell_adj, a.n_adj, a.local_rows = self.d_adj.ptr, self.n_adj, self.d_lr.ptr
        a.nmax, a.n_neigh, a.ntypes = self.nmax, self.d_nn.ptr, 1
        a.r_cut[0], a.r_on[0], a.shift_mode, a.compute_virial = self.r_cut, r_on, mode, int(virial)
        a.force, a.virial, a.tpp, a.workgroups = d_f.ptr, d_v.ptr, tpp, workgroups
        P = (H.HmLJParams * 1)()
        P[0].sigma, P[0].epsilon = 1.0, 1.0
        H.check(H.lib().hm_pair_lj_cells(C.byref(a), P, None), "pair cells")
        H.check(H.lib().hm_synchronize())
        return (d_f.numpy().reshape(N, 4).astype(np.float64),
                d_v.numpy().reshape(6, N).astype(np.float64))


def pair_table(pos, L, nl, nn, r_cut, r_min, U, F, virial=True, tpp=4, block=256):
    """hm_pair_table for one particle type; U and F uploaded as FP32 device tables."""
    N = len(pos)
    d_pos = dev(pos4(pos))
    d_nl = dev(nl, np.uint32)
    d_nn = dev(nn, np.uint32)
    d_U = dev(np.asarray(U, np.float32), np.float32)
    d_F = dev(np.asarray(F, np.float32), np.float32)
    d_f = H.DeviceArray((N * 4,), np.float32, fill=0)
    d_v = H.DeviceArray((N * 6,), np.float32, fill=0)
    a = H.HmPairArgs()
    a.N, a.pos, a.box = N, d_pos.ptr, H.box(L)
    a.nlist, a.nmax, a.n_neigh, a.ntypes = d_nl.ptr, nl.shape[1], d_nn.ptr, 1
    a.r_cut[0] = r_cut
    a.shift_mode, a.compute_virial = 0, int(virial)
    a.force, a.virial, a.tpp, a.block_size = d_f.ptr, d_v.ptr, tpp, block
    P = (H.HmTableParams * 1)()
    P[0].r_min, P[0].width, P[0].U, P[0].F = r_min, len(U), d_U.ptr, d_F.ptr
    H.check(H.lib().hm_pair_table(C.byref(a), P, None), "pair table")
    H.check(H.lib().hm_synchronize())
    return d_f.numpy().reshape(N, 4).astype(np.float64), d_v.numpy().reshape(6, N).astype(np.float64)
