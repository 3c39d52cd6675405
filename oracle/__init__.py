"""ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import this module. It loads ``oracle/liboracle.so`` (the FP64 C restatement of the
HOOMD-blue CPU path, see ``hoomd_oracle.h``) and, when built, ``oracle/_ref/libhoomd_ref.so``
(the reference's own evaluator and box headers).
"""

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_d = C.c_double
_u = C.c_uint
_i = C.c_int
_sz = C.c_size_t
_p = C.c_void_p


def build():
    """Compile liboracle.so (and the reference driver when the reference sources exist)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if os.path.isdir("/root/reference/hoomd"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        sigs = {
            "ob_box_init": (None, [_p, _d, _d, _d]),
            "ob_make_fraction": (None, [_p, _p, _p]),
            "ob_min_image": (None, [_p, _p]),
            "ob_wrap": (None, [_p, _p, _p]),
            "orng_philox4x32_10": (None, [_p, _p, _p]),
            "orng_seed": (None, [C.c_uint8, C.c_uint64, C.c_uint16, _p]),
            "orng_counter": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint16, _p]),
            "orng_u01_u64": (_d, [C.c_uint64]),
            "orng_uneg11_u64": (_d, [C.c_uint64]),
            "oe_lj": (_i, [_d, _d, _d, _d, _i, _p, _p]),
            "oe_fene": (_i, [_d, _d, _d, _d, _d, _d, _p, _p]),
            "oe_table": (_i, [_d, _d, _d, _u, _p, _p, _p, _p]),
            "oc_compute_dims": (None, [_p, _d, _p]),
            "oc_n_adj": (_u, [_p]),
            "oc_cell_adj": (None, [_p, _p]),
            "oc_compute_cell_list": (_i, [_p, _p, _sz, _p, _u, _p, _p, _p, _p]),
            "on_build_binned": (_u, [_p, _sz, _p, _p, _u, _p, _d, _i, _p, _p, _u, _u, _p, _p]),
            "on_build_brute": (_u, [_p, _sz, _p, _d, _i, _u, _p, _p]),
            "on_distance_check": (_i, [_p, _sz, _p, _p, _d, _d]),
            "op_pair_lj": (None, [_p, _sz, _p, _p, _p, _u, _i, _d, _d, _d, _d, _i, _i, _p, _p]),
            "op_pair_table": (None, [_p, _sz, _p, _p, _p, _u, _i, _d, _d, _u, _p, _p, _i, _p,
                                     _p]),
            "op_bond_fene": (_i, [_p, _sz, _p, _sz, _p, _d, _d, _d, _d, _d, _i, _p, _p]),
            "oi_nvt_step_one": (None, [_p, _sz, _p, _p, _p, _p, _d, _d]),
            "oi_nvt_step_two": (None, [_sz, _p, _p, _p, _p, _d, _d]),
            "oi_langevin_step_two": (None, [_sz, _p, _p, _p, _p, _p, _p, _p, _d, _d,
                                            C.c_uint64, C.c_uint16]),
            "oi_kinetic_energy": (_d, [_sz, _p, _p]),
            "oi_bussi_alpha": (_d, [_d, _d, _d, _d, _d, C.c_uint64, C.c_uint16, C.c_uint32]),
            "oi_thermalize": (None, [_sz, _p, _p, _p, _d, C.c_uint64, C.c_uint16]),
            "os_hilbert_traversal": (None, [_u, _p]),
            "os_sort_order": (None, [_p, _sz, _p, _u, _p, _p]),
            "os_create": (_p, [_sz, _d, _d, _d, _p, _p, _p, _d, _d, _d, _i, _d, _d, C.c_uint16]),
            "os_destroy": (None, [_p]),
            "oracle_set_threads": (None, [_i]),
            "oracle_threads": (_i, []),
            "os_set_thermostat": (None, [_p, _i, _d, _d]),
            "os_set_sorter": (None, [_p, _u, _u]),
            "os_set_bonds": (None, [_p, _sz, _p, _d, _d, _d, _d, _d, _i]),
            "os_run": (_i, [_p, C.c_uint64]),
            "os_get_state": (None, [_p, _p, _p, _p]),
            "os_get_forces": (None, [_p, _p, _p, _p]),
            "os_timestep": (C.c_uint64, [_p]),
            "os_n_builds": (C.c_uint64, [_p]),
            "os_set_nlist_policy": (None, [_p, C.c_uint, C.c_int]),
            "os_n_dangerous": (C.c_uint64, [_p]),
            "os_shortest_rebuild": (C.c_uint, [_p]),
            "os_kinetic_energy": (_d, [_p]),
            "os_potential_energy": (_d, [_p]),
            "nvp_run": (C.c_longlong, [_i, _sz, _d, _p, _p, _d, _d, _d, C.c_uint64, C.c_uint64,
                                       _i, _p, _p, _p, _p]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def ref():
    """The reference-header library, or None when it was not built (no reference sources)."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libhoomd_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_lj.argtypes = [_d, _d, _d, _d, _i, _p, _p]
        R.ref_lj.restype = _i
        R.ref_fene.argtypes = [_d, _d, _d, _d, _d, _d, _p, _p]
        R.ref_fene.restype = _i
        R.ref_make_fraction.argtypes = [_d, _d, _d, _p, _p]
        R.ref_min_image.argtypes = [_d, _d, _d, _p]
        R.ref_wrap.argtypes = [_d, _d, _d, _p, _p]
        _REF = R
    return _REF


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class Box:
    """ob_box (orthorhombic, periodic)."""

    SIZE = 8 * 12 + 4 * 3 + 4  # lo, hi, L, Linv doubles + 3 ints (+pad)

    def __init__(self, Lx, Ly=None, Lz=None):
        Ly = Lx if Ly is None else Ly
        Lz = Lx if Lz is None else Lz
        self.L = (float(Lx), float(Ly), float(Lz))
        self.buf = C.create_string_buffer(self.SIZE)
        lib().ob_box_init(self.buf, *self.L)

    @property
    def ptr(self):
        return C.cast(self.buf, C.c_void_p)

    def make_fraction(self, v):
        v = _f64(v)
        f = np.zeros(3)
        lib().ob_make_fraction(self.ptr, _ptr(v), _ptr(f))
        return f

    def min_image(self, w):
        w = _f64(w).copy()
        lib().ob_min_image(self.ptr, _ptr(w))
        return w

    def wrap(self, w, img):
        w = _f64(w).copy()
        img = np.ascontiguousarray(img, dtype=np.int32).copy()
        lib().ob_wrap(self.ptr, _ptr(w), _ptr(img))
        return w, img


# ------------------------------------------------------------------ RNG
def philox(ctr, key):
    c = _u32(ctr)
    k = _u32(key)
    out = np.zeros(4, np.uint32)
    lib().orng_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def seed_key(id_, timestep, seed):
    k = np.zeros(2, np.uint32)
    lib().orng_seed(id_, timestep, seed, _ptr(k))
    return k


def counter(a=0, b=0, c=0, d=0):
    out = np.zeros(4, np.uint32)
    lib().orng_counter(a, b, c, d, _ptr(out))
    return out


def u01(x):
    return lib().orng_u01_u64(x)


def uneg11(x):
    return lib().orng_uneg11_u64(x)


# ------------------------------------------------------------------ evaluators
def lj(rsq, rcutsq, sigma, epsilon, shift=False):
    f = C.c_double()
    e = C.c_double()
    ok = lib().oe_lj(rsq, rcutsq, sigma, epsilon, int(shift), C.byref(f), C.byref(e))
    return bool(ok), f.value, e.value


def table(rsq, rcutsq, r_min, U, F):
    """EvaluatorPairTable on one pair: (evaluated, force / r, energy)."""
    U = _f64(U)
    F = _f64(F)
    f = C.c_double()
    e = C.c_double()
    ok = lib().oe_table(rsq, rcutsq, r_min, len(U), _ptr(U), _ptr(F), C.byref(f), C.byref(e))
    return bool(ok), f.value, e.value


def fene(rsq, k, r0, sigma, epsilon, delta):
    f = C.c_double()
    e = C.c_double()
    ok = lib().oe_fene(rsq, k, r0, sigma, epsilon, delta, C.byref(f), C.byref(e))
    return bool(ok), f.value, e.value


# ------------------------------------------------------------------ cell list / nlist
def cell_dims(box, width):
    dim = np.zeros(3, np.uint32)
    lib().oc_compute_dims(box.ptr, width, _ptr(dim))
    return dim


def cell_adj(dim):
    dim = _u32(dim)
    n_adj = lib().oc_n_adj(_ptr(dim))
    adj = np.zeros(int(np.prod(dim.astype(np.int64))) * n_adj, np.uint32)
    lib().oc_cell_adj(_ptr(dim), _ptr(adj))
    return adj.reshape(-1, n_adj)


def cell_list(box, dim, pos, cmax):
    """Returns (err, cell_size, cell_idx[ncell, cmax], cell_of[N], overflow)."""
    pos = _f64(pos)
    dim = _u32(dim)
    ncell = int(np.prod(dim.astype(np.int64)))
    size = np.zeros(ncell, np.uint32)
    idx = np.zeros(ncell * cmax, np.uint32)
    cell_of = np.zeros(len(pos), np.uint32)
    ovf = C.c_uint()
    err = lib().oc_compute_cell_list(box.ptr, _ptr(dim), len(pos), _ptr(pos), cmax, _ptr(size),
                                     _ptr(idx), _ptr(cell_of), C.byref(ovf))
    return err, size, idx.reshape(ncell, cmax), cell_of, ovf.value


def nlist_binned(box, pos, r_cut, r_buff, full=True, types=None, ntypes=1, excl=None):
    """Binned build; r_cut is a scalar or an ntypes x ntypes matrix. excl: list of lists of
    excluded indices per particle. Returns (nlist [N, nmax], n_neigh)."""
    pos = _f64(pos)
    N = len(pos)
    rc = _f64(np.broadcast_to(np.asarray(r_cut, np.float64), (ntypes, ntypes)).ravel())
    t = _u32(types) if types is not None else None
    n_ex = ex = None
    max_ex = 0
    if excl is not None:
        max_ex = max(1, max(len(e) for e in excl))
        n_ex = np.array([len(e) for e in excl], np.uint32)
        ex = np.zeros((N, max_ex), np.uint32)
        for i, e in enumerate(excl):
            ex[i, :len(e)] = e
        ex = ex.ravel()
    nmax = 4
    while True:
        nl = np.zeros(N * nmax, np.uint32)
        nn = np.zeros(N, np.uint32)
        need = lib().on_build_binned(box.ptr, N, _ptr(pos), _ptr(t), ntypes, _ptr(rc), r_buff,
                                     int(full), _ptr(n_ex), _ptr(ex), max_ex, nmax, _ptr(nl),
                                     _ptr(nn))
        if need <= nmax:
            return nl.reshape(N, nmax), nn
        nmax = (need + 3) & ~3


def nlist_brute(box, pos, r_list, full=True):
    pos = _f64(pos)
    N = len(pos)
    nmax = max(4, N)
    nl = np.zeros(N * nmax, np.uint32)
    nn = np.zeros(N, np.uint32)
    lib().on_build_brute(box.ptr, N, _ptr(pos), r_list, int(full), nmax, _ptr(nl), _ptr(nn))
    return nl.reshape(N, nmax), nn


def pair_set(nlist, n_neigh, tags=None):
    """Set of frozenset({i, j}) pairs, optionally mapped through tags."""
    out = set()
    for i in range(len(n_neigh)):
        for j in nlist[i, :n_neigh[i]]:
            a, b = (i, int(j)) if tags is None else (int(tags[i]), int(tags[j]))
            out.add((a, b) if a < b else (b, a))
    return out


def pair_array(nlist, n_neigh, tags=None):
    """Sorted unique (min, max) pairs as an int64 array [P, 2] (fast for large lists)."""
    N = len(n_neigh)
    nn = n_neigh.astype(np.int64)
    i = np.repeat(np.arange(N, dtype=np.int64), nn)
    mask = np.arange(nlist.shape[1])[None, :] < nn[:, None]
    j = nlist[mask].astype(np.int64)
    if tags is not None:
        tg = np.asarray(tags, np.int64)
        i, j = tg[i], tg[j]
    a = np.minimum(i, j)
    b = np.maximum(i, j)
    key = np.unique(a * (1 << 32) + b)
    return np.stack([key >> 32, key & 0xffffffff], axis=1)


def distance_check(box, pos, last, rcut_max, r_buff):
    pos = _f64(pos)
    last = _f64(last)
    return bool(lib().on_distance_check(box.ptr, len(pos), _ptr(pos), _ptr(last), rcut_max, r_buff))


# ------------------------------------------------------------------ forces
def pair_lj(box, pos, nlist, n_neigh, third_law, sigma=1.0, epsilon=1.0, r_cut=2.5, r_on=0.0,
            mode="none", virial=True):
    pos = _f64(pos)
    N = len(pos)
    nl = _u32(nlist)
    nn = _u32(n_neigh)
    f = np.zeros((N, 4))
    v = np.zeros((6, N))
    m = {"none": 0, "shift": 1, "xplor": 2}[mode]
    lib().op_pair_lj(box.ptr, N, _ptr(pos), _ptr(nl), _ptr(nn), nl.shape[1], int(third_law), sigma,
                     epsilon, r_cut, r_on, m, int(virial), _ptr(f), _ptr(v))
    return f, v


def pair_table(box, pos, nlist, n_neigh, third_law, r_cut, r_min, U, F, virial=True):
    pos = _f64(pos)
    N = len(pos)
    nl = _u32(nlist)
    nn = _u32(n_neigh)
    U = _f64(U)
    Fv = _f64(F)
    f = np.zeros((N, 4))
    v = np.zeros((6, N))
    lib().op_pair_table(box.ptr, N, _ptr(pos), _ptr(nl), _ptr(nn), nl.shape[1], int(third_law),
                        r_cut, r_min, len(U), _ptr(U), _ptr(Fv), int(virial), _ptr(f), _ptr(v))
    return f, v


def bond_fene(box, pos, bonds, k, r0, sigma, epsilon, delta, virial=True):
    pos = _f64(pos)
    N = len(pos)
    b = _u32(bonds).ravel()
    f = np.zeros((N, 4))
    v = np.zeros((6, N))
    r = lib().op_bond_fene(box.ptr, N, _ptr(pos), len(b) // 2, _ptr(b), k, r0, sigma, epsilon,
                           delta, int(virial), _ptr(f), _ptr(v))
    return r, f, v


# ------------------------------------------------------------------ integrators
def nvt_step_one(box, pos, vel, accel, image, dt, alpha=1.0):
    pos = _f64(pos).copy()
    vel = _f64(vel).copy()
    accel = _f64(accel)
    image = np.ascontiguousarray(image, np.int32).copy()
    lib().oi_nvt_step_one(box.ptr, len(pos), _ptr(pos), _ptr(vel), _ptr(accel), _ptr(image), dt,
                          alpha)
    return pos, vel, image


def nvt_step_two(vel, mass, net_force4, dt, alpha=1.0):
    vel = _f64(vel).copy()
    acc = np.zeros_like(vel)
    mass = _f64(mass)
    f = _f64(net_force4)
    lib().oi_nvt_step_two(len(vel), _ptr(vel), _ptr(acc), _ptr(mass), _ptr(f), dt, alpha)
    return vel, acc


def langevin_step_two(vel, mass, tag, net_force4, gamma, kT, dt, timestep, seed, types=None):
    vel = _f64(vel).copy()
    acc = np.zeros_like(vel)
    mass = _f64(mass)
    tag = _u32(tag)
    f = _f64(net_force4)
    g = _f64(np.atleast_1d(gamma))
    t = _u32(types) if types is not None else None
    lib().oi_langevin_step_two(len(vel), _ptr(vel), _ptr(acc), _ptr(mass), _ptr(tag), _ptr(f),
                               _ptr(g), _ptr(t), kT, dt, timestep, seed)
    return vel, acc


def kinetic_energy(vel, mass):
    vel = _f64(vel)
    mass = _f64(mass)
    return lib().oi_kinetic_energy(len(vel), _ptr(vel), _ptr(mass))


def bussi_alpha(K, ndof, dt, tau, kT, timestep, seed, instance_tag=0):
    return lib().oi_bussi_alpha(K, ndof, dt, tau, kT, timestep, seed, instance_tag)


def thermalize(mass, tag, kT, timestep, seed):
    mass = _f64(mass)
    tag = _u32(tag)
    vel = np.zeros((len(mass), 3))
    lib().oi_thermalize(len(mass), _ptr(vel), _ptr(mass), _ptr(tag), kT, timestep, seed)
    return vel


# ------------------------------------------------------------------ Hilbert sort
def hilbert_traversal(grid):
    t = np.zeros(grid ** 3, np.uint32)
    lib().os_hilbert_traversal(grid, _ptr(t))
    return t


def sort_order(box, pos, grid, traversal=None):
    pos = _f64(pos)
    if traversal is None:
        traversal = hilbert_traversal(grid)
    order = np.zeros(len(pos), np.uint32)
    lib().os_sort_order(box.ptr, len(pos), _ptr(pos), grid, _ptr(_u32(traversal)), _ptr(order))
    return order


# ------------------------------------------------------------------ full system
def set_threads(n):
    """Host threads of the step loop (1: the serial order the tests compare against)."""
    lib().oracle_set_threads(int(n))


def threads():
    return lib().oracle_threads()


class System:
    """The restated CPU step loop (half list + 3rd law), by tag."""

    def __init__(self, pos, vel, L, mass=None, sigma=1.0, epsilon=1.0, r_cut=2.5, mode="none",
                 buffer=0.4, dt=0.005, seed=0):
        pos = _f64(pos)
        vel = _f64(vel)
        self.N = len(pos)
        Ls = np.broadcast_to(np.asarray(L, np.float64), (3,))
        m = _f64(mass) if mass is not None else None
        self.h = lib().os_create(self.N, Ls[0], Ls[1], Ls[2], _ptr(pos), _ptr(vel), _ptr(m), sigma,
                                 epsilon, r_cut, {"none": 0, "shift": 1}[mode], buffer, dt, seed)

    def __del__(self):
        if getattr(self, "h", None):
            lib().os_destroy(self.h)
            self.h = None

    def set_bussi(self, kT, tau=0.0):
        lib().os_set_thermostat(self.h, 1, kT, tau)

    def set_langevin(self, kT, gamma=1.0):
        lib().os_set_thermostat(self.h, 2, kT, gamma)

    def set_sorter(self, period, grid=256):
        lib().os_set_sorter(self.h, period, grid)

    def set_bonds(self, bonds, k, r0, sigma, epsilon, delta, exclude=True):
        b = _u32(bonds).ravel()
        lib().os_set_bonds(self.h, len(b) // 2, _ptr(b), k, r0, sigma, epsilon, delta,
                           int(exclude))

    def run(self, steps):
        r = lib().os_run(self.h, steps)
        if r != 0:
            raise RuntimeError(f"oracle run failed with code {r}")

    def state(self):
        p = np.zeros((self.N, 3))
        v = np.zeros((self.N, 3))
        im = np.zeros((self.N, 3), np.int32)
        lib().os_get_state(self.h, _ptr(p), _ptr(v), _ptr(im))
        return p, v, im

    def forces(self):
        f = np.zeros((self.N, 4))
        vir = np.zeros((6, self.N))
        bf = np.zeros((self.N, 4))
        lib().os_get_forces(self.h, _ptr(f), _ptr(vir), _ptr(bf))
        return f, vir, bf

    @property
    def timestep(self):
        return lib().os_timestep(self.h)

    @property
    def n_builds(self):
        return lib().os_n_builds(self.h)

    def set_nlist_policy(self, rebuild_check_delay=1, check_dist=True):
        lib().os_set_nlist_policy(self.h, int(rebuild_check_delay), int(bool(check_dist)))

    @property
    def n_dangerous(self):
        return lib().os_n_dangerous(self.h)

    @property
    def shortest_rebuild(self):
        return lib().os_shortest_rebuild(self.h)

    def kinetic_energy(self):
        return lib().os_kinetic_energy(self.h)

    def potential_energy(self):
        return lib().os_potential_energy(self.h)


NVE_VARIANTS = {"fp64": 0, "fp64_state_fp32_pairs": 1, "fp32": 2}


def nve_precision(variant, pos, vel, L, steps, every, r_cut=2.5, r_buff=0.4, dt=0.005,
                  threads=1):
    """NVE LJ run of the precision study (nve_precision.c) in one variant of NVE_VARIANTS.
    Returns (kinetic, shifted potential) energy samples at steps 0, every, ..., the final
    positions and velocities, and the number of list builds."""
    pos = _f64(pos)
    vel = _f64(vel)
    n = len(pos)
    ns = steps // every + 1
    ke = np.zeros(ns)
    pe = np.zeros(ns)
    po = np.zeros((n, 3))
    vo = np.zeros((n, 3))
    builds = lib().nvp_run(NVE_VARIANTS[variant], n, float(L), _ptr(pos), _ptr(vel), r_cut,
                           r_buff, dt, steps, every, threads, _ptr(ke), _ptr(pe), _ptr(po),
                           _ptr(vo))
    if builds < 0:
        raise ValueError("nvp_run: bad arguments or allocation failure")
    return ke, pe, po, vo, int(builds)
