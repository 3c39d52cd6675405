"""GPU parity of the learned coarse-grained site energy (hm_mlp_cg_* and the engine) against
the FP64 oracle (oracle/mlp_cg.c).

Two oracle modes are used: plain FP64, and FP64 with the matrix-product operands rounded to
bfloat16 at the points where the kernels round them (the kernels then differ from it only by
FP32 arithmetic and accumulation order, and by the rare operand whose FP32 value sits on a
bf16 rounding boundary). Tolerances, relative RMS over all entries:
  densities G against FP64 ............................. 1e-5 (FP32 sums of FP32 exp)
  site energies, dE/dG, forces, virials against bf16 .... 5e-4 (about 1e-4 measured: an
  operand whose FP32 and FP64 values round to different bf16 values moves its site by ~1e-3)
  the same against plain FP64 ........................... 3e-2 (the bf16 rounding itself)
  loss against bf16 ..................................... 1e-4
  parameter gradient against bf16 ....................... 3e-3; against FP64 5e-2
"""

import ctypes as C

import numpy as np
import pytest

import oracle as O
from gpu_helpers import dev, pos4
from oracle import mlp_cg as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    from hoomd_amd import _lib
    if _lib.lib().hm_device_count() < 1:
        pytest.skip("no GPU")
    return _lib


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).mean() / max((b ** 2).mean(), 1e-300)))


def cg_config(N, rho=0.17, seed=1, rmin=0.8):
    """Random sites at CG density with no pair closer than rmin (grid-accelerated)."""
    rng = np.random.default_rng(seed)
    L = (N / rho) ** (1 / 3)
    nc = max(1, int(L // rmin))
    grid = {}
    pos = []
    while len(pos) < N:
        p = rng.uniform(-L / 2, L / 2, 3)
        c = tuple(((p + L / 2) / L * nc).astype(int) % nc)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                        d = pos[q] - p
                        d -= L * np.rint(d / L)
                        if (d * d).sum() < rmin * rmin:
                            ok = False
        if ok:
            grid.setdefault(c, []).append(len(pos))
            pos.append(p)
    return np.asarray(pos), L


class Setup:
    """Buffers of hm_mlp_cg_args for one configuration, its oracle list and a model."""

    def __init__(self, H, N, seed, params_seed, bf16_a=False, pos32=False):
        self.H = H
        pos, L = cg_config(N, seed=seed)
        self.pos, self.L, self.N = pos, L, N
        nl, nn = O.nlist_binned(O.Box(L), pos, 3.0, 0.4, full=True)
        self.nl, self.nn = nl, nn
        m0 = M.Model()
        G0 = M.descriptor(m0, L, pos, nl, nn)
        self.gmean = G0.mean(axis=0)
        self.gscale = 1.0 / (G0.std(axis=0) + 1e-3)
        self.om = M.Model(gmean=self.gmean, gscale=self.gscale)
        self.omb = self.om.with_bf16()
        self.params = M.init_params(params_seed).astype(np.float32)
        self.p64 = self.params.astype(np.float64)
        self.model = H.HmMlpModel()
        self.model.r_cut, self.model.width, self.model.mu0, self.model.dmu = 3.0, 0.2, 0.0, 0.2
        for k in range(16):
            self.model.gmean[k] = self.gmean[k]
            self.model.gscale[k] = self.gscale[k]
        nmax = nl.shape[1]
        self.d_pos = dev(pos4(pos))
        self.d_nl = dev(np.ascontiguousarray(nl, np.uint32))
        self.d_nn = dev(np.ascontiguousarray(nn, np.uint32))
        self.d_par = dev(self.params)
        self.d_bf = H.DeviceArray((H.MLP_NBF16,), np.uint16, fill=0)
        self.d_G = H.DeviceArray((16 * N,), np.float32, fill=0)
        self.d_A = H.DeviceArray((16 * N,), np.float32, fill=0)
        self.d_E = H.DeviceArray((N,), np.float32, fill=0)
        self.d_F = H.DeviceArray((4 * N,), np.float32, fill=0)
        self.d_V = H.DeviceArray((6 * N,), np.float32, fill=0)
        self.d_T = H.DeviceArray((4 * N,), np.float32, fill=0)
        self.d_err = H.DeviceArray((4 * N,), np.float32, fill=0)
        self.d_C = H.DeviceArray((16 * N,), np.float32, fill=0)
        self.d_lp = H.DeviceArray((H.lib().hm_mlp_cg_loss_blocks(N),), np.float32, fill=0)
        self.d_gp = H.DeviceArray((H.MLP_GRAD_BLOCKS * H.MLP_NPARAM,), np.float32, fill=0)
        a = H.HmMlpArgs()
        a.N, a.pos, a.box = N, self.d_pos.ptr, H.box(L)
        a.nlist, a.nmax, a.n_neigh = self.d_nl.ptr, nmax, self.d_nn.ptr
        a.params, a.params_bf16 = self.d_par.ptr, self.d_bf.ptr
        a.G, a.A, a.Esite, a.force, a.virial = (self.d_G.ptr, self.d_A.ptr, self.d_E.ptr,
                                                self.d_F.ptr, self.d_V.ptr)
        a.target, a.err, a.C = self.d_T.ptr, self.d_err.ptr, self.d_C.ptr
        if bf16_a:
            self.d_Ab = H.DeviceArray((16 * N,), np.uint16, fill=0)
            a.A_bf16 = self.d_Ab.ptr
        if pos32:
            p32 = np.zeros((N, 4), np.float32)
            p32[:, :3] = pos.astype(np.float32)
            self.d_p32 = dev(p32)
            a.pos32 = self.d_p32.ptr
        a.loss_partial, a.grad_partial = self.d_lp.ptr, self.d_gp.ptr
        self.args = a
        H.check(H.lib().hm_mlp_cg_pack(self.d_par.ptr, self.d_bf.ptr, None), "pack")

    def forward(self):
        L = self.H.lib()
        for fn in (L.hm_mlp_cg_descriptor, L.hm_mlp_cg_forward):
            self.H.check(fn(C.byref(self.args), C.byref(self.model), None))
        self.H.check(L.hm_mlp_cg_forces(C.byref(self.args), C.byref(self.model), 0, None))
        self.H.check(L.hm_synchronize())


@pytest.mark.parametrize("N,seed", [(3000, 1), (517, 2)])
def test_mlp_forward_matches_oracle(H, N, seed):
    print(f"mlp forward: N={N} seed={seed}", flush=True)
    S = Setup(H, N, seed, params_seed=seed + 10)
    S.forward()
    G = S.d_G.numpy().reshape(N, 16)
    A = S.d_A.numpy().reshape(N, 16)
    E = S.d_E.numpy()
    F = S.d_F.numpy().reshape(N, 4)
    V = S.d_V.numpy().reshape(6, N)
    oG = M.descriptor(S.om, S.L, S.pos, S.nl, S.nn)
    Et, oF, oE, oV, oA = M.forces(S.omb, S.p64, S.L, S.pos, S.nl, S.nn)
    _, fF, fE, fV, fA = M.forces(S.om, S.p64, S.L, S.pos, S.nl, S.nn)
    errs = dict(G=rel_rms(G, oG), E=rel_rms(E, oE), A=rel_rms(A, oA), F=rel_rms(F[:, :3], oF),
                V=rel_rms(V, oV), E64=rel_rms(E, fE), F64=rel_rms(F[:, :3], fF),
                V64=rel_rms(V, fV))
    print("rel RMS:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert errs["G"] < 1e-5
    for k in ("E", "A", "F", "V"):
        assert errs[k] < 5e-4, k
    for k in ("E64", "F64", "V64"):
        assert errs[k] < 3e-2, k
    np.testing.assert_array_equal(F[:, 3], E)
    # third law: the forces sum to zero
    assert np.abs(F[:, :3].astype(np.float64).sum(axis=0)).max() < 1e-3 * np.abs(F[:, :3]).max()


def bf16_rne(x):
    """FP32 -> bf16 bits, round to nearest even (finite values)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


@pytest.mark.parametrize("N,seed,pos32", [(3000, 4, False), (333, 5, False), (3000, 6, True),
                                          (517, 7, True)])
def test_mlp_bf16_dEdG_forces_match_oracle(H, N, seed, pos32):
    """The force pass on the bf16 copy of dE/dG: the copy is A rounded to nearest even, bit for
    bit, and the forces follow the oracle that rounds a(i) in pass 2 (bf16 mode 2). With the
    FP32 position copy both pair passes match the oracle evaluated at the FP32-rounded
    positions (the densities to the same 1e-5 as from FP64 positions)."""
    print(f"mlp forward, bf16 dE/dG: N={N} seed={seed} pos32={pos32}", flush=True)
    S = Setup(H, N, seed, params_seed=seed + 10, bf16_a=True, pos32=pos32)
    if pos32:
        S.pos = S.pos.astype(np.float32).astype(np.float64)
    S.forward()
    G = S.d_G.numpy().reshape(N, 16)
    assert rel_rms(G, M.descriptor(S.om, S.L, S.pos, S.nl, S.nn)) < 1e-5
    A = S.d_A.numpy()
    np.testing.assert_array_equal(S.d_Ab.numpy(), bf16_rne(A))
    F = S.d_F.numpy().reshape(N, 4)
    E = S.d_E.numpy()
    V = S.d_V.numpy().reshape(6, N)
    om2 = S.om.with_bf16(2)
    _, oF, oE, oV, oA = M.forces(om2, S.p64, S.L, S.pos, S.nl, S.nn)
    _, fF, _, _, _ = M.forces(S.om, S.p64, S.L, S.pos, S.nl, S.nn)
    _, bF, _, _, _ = M.forces(S.omb, S.p64, S.L, S.pos, S.nl, S.nn)
    errs = dict(F=rel_rms(F[:, :3], oF), V=rel_rms(V, oV), E=rel_rms(E, oE),
                A=rel_rms(A.reshape(N, 16), oA), F64=rel_rms(F[:, :3], fF),
                F_vs_mode1=rel_rms(F[:, :3], bF))
    print("rel RMS:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k in ("F", "V", "E", "A"):
        assert errs[k] < 5e-4, k
    assert errs["F64"] < 3e-2
    # the oracle's mode 2 is what the kernel computes: closer than mode 1 (FP32 dE/dG)
    assert errs["F"] < errs["F_vs_mode1"]
    # pairs use the same rounded values on both sides: the forces still sum to zero
    assert np.abs(F[:, :3].astype(np.float64).sum(axis=0)).max() < 1e-3 * np.abs(F[:, :3]).max()
    # without A the site kernel writes only the copy, and the forces are unchanged
    Ab = S.d_Ab.numpy()
    H.check(H.lib().hm_memset(S.d_Ab.ptr, 0, Ab.nbytes))
    H.check(H.lib().hm_memset(S.d_F.ptr, 0, 16 * N))
    S.args.A = None
    S.forward()
    np.testing.assert_array_equal(S.d_Ab.numpy(), Ab)
    np.testing.assert_array_equal(S.d_F.numpy().reshape(N, 4), F)


def test_mlp_gradient_and_adam_match_oracle(H):
    N = 1500
    print(f"mlp gradient: N={N}", flush=True)
    S = Setup(H, N, 3, params_seed=7)
    T = np.random.default_rng(5).normal(size=(N, 3))
    T4 = np.zeros((N, 4), np.float32)
    T4[:, :3] = T
    H.check(H.lib().hm_memcpy_htod(S.d_T.ptr, T4.ctypes.data, T4.nbytes))
    L = H.lib()
    for fn in (L.hm_mlp_cg_descriptor, L.hm_mlp_cg_forward):
        H.check(fn(C.byref(S.args), C.byref(S.model), None))
    H.check(L.hm_mlp_cg_forces(C.byref(S.args), C.byref(S.model), 1, None))
    H.check(L.hm_mlp_cg_backward(C.byref(S.args), C.byref(S.model), None))
    d_loss = H.DeviceArray((1,), np.float64, fill=0)
    d_grad = H.DeviceArray((H.MLP_NPARAM,), np.float32, fill=0)
    H.check(L.hm_mlp_cg_loss(S.d_lp.ptr, N, d_loss.ptr, None))
    H.check(L.hm_mlp_cg_grad(S.d_gp.ptr, d_grad.ptr, None))
    H.check(L.hm_synchronize())
    loss = d_loss.numpy()[0] / (3 * N)
    grad = d_grad.numpy()
    ol, og, _ = M.loss_grad(S.omb, S.p64, S.L, S.pos, S.nl, S.nn, T.astype(np.float32))
    fl, fg, _ = M.loss_grad(S.om, S.p64, S.L, S.pos, S.nl, S.nn, T.astype(np.float32))
    e_b, e_64 = rel_rms(grad, og), rel_rms(grad, fg)
    print(f"loss {loss:.6e} oracle(bf16) {ol:.6e} oracle(fp64) {fl:.6e}; "
          f"grad rel RMS vs bf16 {e_b:.2e}, vs fp64 {e_64:.2e}", flush=True)
    assert abs(loss - ol) <= 1e-4 * ol
    assert e_b < 3e-3
    assert e_64 < 5e-2
    # gradient blocks: each parameter group matches on its own
    for name, lo, hi in (("W1", 0, 1024), ("b1", 1024, 1088), ("W2", 1088, 5184),
                         ("b2", 5184, 5248), ("w3", 5248, 5312)):
        assert rel_rms(grad[lo:hi], og[lo:hi]) < 1e-2, name
    assert grad[5312] == 0.0
    # one Adam step from zero moments, against the oracle's step from the GPU gradient
    d_m1 = H.DeviceArray((H.MLP_NPARAM,), np.float32, fill=0)
    d_m2 = H.DeviceArray((H.MLP_NPARAM,), np.float32, fill=0)
    H.check(L.hm_mlp_cg_adam(S.d_par.ptr, d_m1.ptr, d_m2.ptr, S.d_gp.ptr, S.d_bf.ptr, 1e-3,
                             0.9, 0.999, 1e-8, 1, None))
    H.check(L.hm_synchronize())
    p = S.p64.copy()
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    M.adam(p, m1, m2, grad.astype(np.float64), 1e-3, 0.9, 0.999, 1e-8, 1)
    np.testing.assert_allclose(S.d_par.numpy(), p, rtol=0, atol=2e-7)
    # the bf16 copies follow the new parameters (round to nearest even)
    bf = S.d_bf.numpy()
    newp = S.d_par.numpy()
    w1 = np.array([M.bf16_round(v) for v in newp[:1024]], np.float32)
    got = (bf[:1024].astype(np.uint32) << 16).view(np.float32)
    np.testing.assert_array_equal(got, w1)
