"""ctypes bindings of the MLP coarse-grained oracle (oracle/mlp_cg.c; TEST INFRASTRUCTURE
ONLY: tests/, smoke() and bench.py's checks import it, the product never does)."""

import ctypes as C

import numpy as np

from . import _f64, _ptr, _u32, lib

NK = 16
H = 64
NPARAM = H * NK + H + H * H + H + H + 1

_d = C.c_double
_p = C.c_void_p
_sz = C.c_size_t
_u = C.c_uint

# offsets of the flat parameter vector (W1[64][16], b1, W2[64][64], b2, w3, b3)
OFF = {"W1": 0, "b1": H * NK, "W2": H * NK + H, "b2": H * NK + H + H * H,
       "w3": H * NK + 2 * H + H * H, "b3": H * NK + 3 * H + H * H}


class _Model(C.Structure):
    _fields_ = [("rcut", _d), ("width", _d), ("mu0", _d), ("dmu", _d), ("gmean", _d * NK),
                ("gscale", _d * NK), ("bf16", C.c_int)]


def _sigs():
    L = lib()
    if getattr(L, "_ocg_ready", False):
        return L
    sigs = {
        "ocg_basis": (None, [_p, _d, _p, _p]),
        "ocg_site": (None, [_p, _p, _p, _p, _p]),
        "ocg_descriptor": (None, [_p, _p, _p, _sz, _p, _p, _u, _p]),
        "ocg_forces": (_d, [_p, _p, _p, _p, _sz, _p, _p, _u, _p, _p, _p, _p]),
        "ocg_loss_grad": (_d, [_p, _p, _p, _p, _sz, _p, _p, _u, _p, _p, _p]),
        "ocg_adam": (None, [_p, _p, _p, _p, _sz, _d, _d, _d, _d, _u]),
        "ocg_bf16_round": (_d, [_d]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L._ocg_ready = True
    return L


class Model:
    """The descriptor constants and normalisation: centres mu0 + k dmu, Gaussian width, cosine
    cutoff r_cut; x = (G - gmean) * gscale. bf16 = 1 (or True) rounds matrix-product operands as
    the GPU; bf16 = 2 also rounds dE/dG where the force pass reads it (the GPU's bf16 copy)."""

    def __init__(self, rcut=3.0, width=0.2, mu0=0.0, dmu=0.2, gmean=None, gscale=None,
                 bf16=False):
        self.s = _Model()
        self.s.rcut, self.s.width, self.s.mu0, self.s.dmu = rcut, width, mu0, dmu
        gm = np.zeros(NK) if gmean is None else np.asarray(gmean, np.float64)
        gs = np.ones(NK) if gscale is None else np.asarray(gscale, np.float64)
        for k in range(NK):
            self.s.gmean[k] = gm[k]
            self.s.gscale[k] = gs[k]
        self.s.bf16 = int(bf16)

    @property
    def ptr(self):
        return C.cast(C.pointer(self.s), C.c_void_p)

    def with_bf16(self, on=True):
        """The same model with bf16 mode on (1 or True, or 2: see the class) or off."""
        m = Model(self.s.rcut, self.s.width, self.s.mu0, self.s.dmu, list(self.s.gmean),
                  list(self.s.gscale), int(on))
        return m


def basis(model, r):
    phi = np.zeros(NK)
    dphi = np.zeros(NK)
    _sigs().ocg_basis(model.ptr, float(r), _ptr(phi), _ptr(dphi))
    return phi, dphi


def site(model, params, x):
    params = _f64(params)
    x = _f64(x)
    E = C.c_double()
    g = np.zeros(NK)
    _sigs().ocg_site(model.ptr, _ptr(params), _ptr(x), C.byref(E), _ptr(g))
    return E.value, g


def _L3(L):
    return _f64(np.broadcast_to(np.asarray(L, np.float64), (3,)))


def descriptor(model, L, pos, nlist, n_neigh):
    pos = _f64(pos)
    nl = _u32(nlist)
    nn = _u32(n_neigh)
    G = np.zeros((len(pos), NK))
    _sigs().ocg_descriptor(model.ptr, _ptr(_L3(L)), _ptr(pos), len(pos), _ptr(nl), _ptr(nn),
                           nl.shape[1], _ptr(G))
    return G


def forces(model, params, L, pos, nlist, n_neigh):
    """(total energy, F[N,3], site energies[N], virial[6,N], dE/dG[N,16])"""
    pos = _f64(pos)
    nl = _u32(nlist)
    nn = _u32(n_neigh)
    N = len(pos)
    F = np.zeros((N, 3))
    E = np.zeros(N)
    V = np.zeros((6, N))
    A = np.zeros((N, NK))
    Et = _sigs().ocg_forces(model.ptr, _ptr(_f64(params)), _ptr(_L3(L)), _ptr(pos), N, _ptr(nl),
                            _ptr(nn), nl.shape[1], _ptr(F), _ptr(E), _ptr(V), _ptr(A))
    return Et, F, E, V, A


def loss_grad(model, params, L, pos, nlist, n_neigh, target):
    """(loss, grad[NPARAM], F[N,3]) of the force-matching loss sum |F - T|^2 / (3N)."""
    pos = _f64(pos)
    nl = _u32(nlist)
    nn = _u32(n_neigh)
    N = len(pos)
    g = np.zeros(NPARAM)
    F = np.zeros((N, 3))
    loss = _sigs().ocg_loss_grad(model.ptr, _ptr(_f64(params)), _ptr(_L3(L)), _ptr(pos), N,
                                 _ptr(nl), _ptr(nn), nl.shape[1], _ptr(_f64(target)), _ptr(g),
                                 _ptr(F))
    return loss, g, F


def adam(params, m1, m2, grad, lr, beta1, beta2, eps, t):
    """One Adam step in place on float64 arrays (t counts from 1)."""
    for a in (params, m1, m2):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    _sigs().ocg_adam(_ptr(params), _ptr(m1), _ptr(m2), _ptr(_f64(grad)), len(params), lr, beta1,
                     beta2, eps, t)


def bf16_round(v):
    return _sigs().ocg_bf16_round(float(v))


def init_params(seed=0, scale=1.0):
    """Random parameters: Glorot-uniform weights, zero biases (a test input, not a trained
    model)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    a1 = np.sqrt(6.0 / (NK + H)) * scale
    a2 = np.sqrt(6.0 / (H + H)) * scale
    a3 = np.sqrt(6.0 / (H + 1)) * scale
    p[OFF["W1"]:OFF["b1"]] = rng.uniform(-a1, a1, H * NK)
    p[OFF["W2"]:OFF["b2"]] = rng.uniform(-a2, a2, H * H)
    p[OFF["w3"]:OFF["b3"]] = rng.uniform(-a3, a3, H)
    p[OFF["b1"]:OFF["W2"]] = rng.normal(0, 0.1, H) * scale
    p[OFF["b2"]:OFF["w3"]] = rng.normal(0, 0.1, H) * scale
    return p
