"""CPU pins of the MLP coarse-grained oracle (oracle/mlp_cg.c).

The model is this project's own (the reference has no learned potential), so the oracle is
pinned by mathematics rather than by reference fixtures:

* forces against the central finite-difference gradient of the total energy, and against
  torch autograd of the same energy (an independent derivative of the model definition);
* parameter gradients of the force-matching loss against central finite differences of the
  loss, and against torch double-backward autograd;
* a two-site system against its closed form, the descriptor against a brute-force sum, the
  Adam step against its first-step closed form and the bf16 rounding against known answers.
"""

import numpy as np
import pytest

import oracle as O
from oracle import mlp_cg as M

torch = pytest.importorskip("torch")


def cg_config(N=140, rho=0.17, seed=1, rmin=0.9):
    """Random sites at CG density with no pair closer than rmin (a hard-core walk)."""
    rng = np.random.default_rng(seed)
    L = (N / rho) ** (1 / 3)
    pos = []
    while len(pos) < N:
        p = rng.uniform(-L / 2, L / 2, 3)
        if pos:
            d = np.asarray(pos) - p
            d -= L * np.rint(d / L)
            if (d ** 2).sum(axis=1).min() < rmin ** 2:
                continue
        pos.append(p)
    return np.asarray(pos), L


def model_for(pos, L, bf16=False):
    """A model whose normalisation comes from the configuration's own descriptor statistics."""
    m0 = M.Model()
    nl, nn = O.nlist_binned(O.Box(L), pos, 3.0, 0.0, full=True)
    G = M.descriptor(m0, L, pos, nl, nn)
    return M.Model(gmean=G.mean(axis=0), gscale=1.0 / (G.std(axis=0) + 1e-3), bf16=bf16), nl, nn


def torch_energy(model, params, L, pos_t, pairs):
    """Total energy of the model written with torch operations (FP64) over explicit pairs."""
    s = model.s
    i, j = pairs[:, 0], pairs[:, 1]
    d = pos_t[i] - pos_t[j]
    d = d - L * torch.round(d / L)
    r = torch.sqrt((d * d).sum(dim=1))
    mu = torch.tensor([s.mu0 + k * s.dmu for k in range(M.NK)], dtype=torch.float64)
    fc = 0.5 * (torch.cos(np.pi * r / s.rcut) + 1.0)
    phi = torch.exp(-(r[:, None] - mu[None, :]) ** 2 / (2 * s.width ** 2)) * fc[:, None]
    G = torch.zeros((pos_t.shape[0], M.NK), dtype=torch.float64).index_add(0, i, phi)
    x = (G - torch.tensor(list(s.gmean), dtype=torch.float64)) * \
        torch.tensor(list(s.gscale), dtype=torch.float64)
    W1 = params[M.OFF["W1"]:M.OFF["b1"]].reshape(M.H, M.NK)
    b1 = params[M.OFF["b1"]:M.OFF["W2"]]
    W2 = params[M.OFF["W2"]:M.OFF["b2"]].reshape(M.H, M.H)
    b2 = params[M.OFF["b2"]:M.OFF["w3"]]
    w3 = params[M.OFF["w3"]:M.OFF["b3"]]
    b3 = params[M.OFF["b3"]]
    h1 = x @ W1.T + b1
    h2 = torch.nn.functional.silu(h1) @ W2.T + b2
    E = torch.nn.functional.silu(h2) @ w3 + b3
    return E.sum()


def pair_list(nl, nn, pos, L, rcut=3.0):
    out = []
    for i in range(len(pos)):
        for q in range(nn[i]):
            j = int(nl[i, q])
            d = pos[i] - pos[j]
            d -= L * np.rint(d / L)
            if (d * d).sum() < rcut * rcut:
                out.append((i, j))
    return np.asarray(out, np.int64)


def test_bf16_rounding_known_answers():
    assert M.bf16_round(1.0) == 1.0
    assert M.bf16_round(1.0 + 2 ** -8) == 1.0                     # tie -> even
    assert M.bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6       # tie -> even (up)
    assert M.bf16_round(1.0 + 2 ** -8 + 2 ** -20) == 1.0 + 2 ** -7
    assert M.bf16_round(-3.0) == -3.0
    assert M.bf16_round(1e-3) == np.float32(np.frombuffer(
        (np.frombuffer(np.float32(1e-3).tobytes(), np.uint32) + 0x7fff + 0 & 0xffff0000)
        .astype(np.uint32).tobytes(), np.float32)[0])


def test_basis_closed_form():
    m = M.Model()
    for r in (0.3, 1.37, 2.999):
        phi, dphi = M.basis(m, r)
        mu = 0.2 * np.arange(16)
        fc = 0.5 * (np.cos(np.pi * r / 3.0) + 1)
        np.testing.assert_allclose(phi, np.exp(-(r - mu) ** 2 / 0.08) * fc, rtol=1e-13)
        h = 1e-6
        fd = (M.basis(m, r + h)[0] - M.basis(m, r - h)[0]) / (2 * h)
        np.testing.assert_allclose(dphi, fd, rtol=1e-7, atol=1e-12)
    phi, dphi = M.basis(m, 3.0)
    assert not phi.any() and not dphi.any()


def test_descriptor_is_brute_force_sum():
    pos, L = cg_config(80, seed=2)
    m = M.Model()
    nl, nn = O.nlist_binned(O.Box(L), pos, 3.0, 0.4, full=True)
    G = M.descriptor(m, L, pos, nl, nn)
    ref = np.zeros_like(G)
    for i in range(len(pos)):
        for j in range(len(pos)):
            if i == j:
                continue
            d = pos[i] - pos[j]
            d -= L * np.rint(d / L)
            r = np.sqrt((d * d).sum())
            if r < 3.0:
                ref[i] += M.basis(m, r)[0]
    np.testing.assert_allclose(G, ref, rtol=1e-12, atol=1e-15)


def test_two_site_closed_form():
    L = 20.0
    r = 1.7
    pos = np.array([[0.0, 0.0, 0.0], [r * 0.6, r * 0.8, 0.0]])
    m = M.Model(gmean=np.full(16, 0.1), gscale=np.full(16, 2.0))
    p = M.init_params(5)
    nl = np.array([[1], [0]], np.uint32)
    nn = np.array([1, 1], np.uint32)
    Et, F, E, V, A = M.forces(m, p, L, pos, nl, nn)
    phi, dphi = M.basis(m, r)
    x = (phi - 0.1) * 2.0
    Es, g = M.site(m, p, x)
    a = 2.0 * g
    assert Et == pytest.approx(2 * Es, rel=1e-14)
    rhat = (pos[0] - pos[1]) / r
    np.testing.assert_allclose(F[0], -2.0 * (a @ dphi) * rhat, rtol=1e-12)
    np.testing.assert_allclose(F[1], -F[0], rtol=1e-14)
    np.testing.assert_allclose(V[:, 0], 0.5 * (-2.0 * (a @ dphi) / r) * np.array(
        [rhat[0] ** 2, rhat[0] * rhat[1], 0, rhat[1] ** 2, 0, 0]) * r * r, atol=1e-14)


def test_bf16_mode2_rounds_dEdG_in_the_force_pass():
    """bf16 mode 2 (the GPU force pass on the bf16 copy of dE/dG): on the two-site closed form
    the pair force uses bf16(a) of both sites; site energies and the returned dE/dG are those
    of mode 1; the forces move away from mode 1 by the rounding and stay close to FP64."""
    L = 20.0
    r = 1.7
    pos = np.array([[0.0, 0.0, 0.0], [r * 0.6, r * 0.8, 0.0]])
    m = M.Model(gmean=np.full(16, 0.1), gscale=np.full(16, 2.0))
    p = M.init_params(5)
    nl = np.array([[1], [0]], np.uint32)
    nn = np.array([1, 1], np.uint32)
    _, F1, E1, _, A1 = M.forces(m.with_bf16(1), p, L, pos, nl, nn)
    Et, F2, E2, _, A2 = M.forces(m.with_bf16(2), p, L, pos, nl, nn)
    np.testing.assert_array_equal(E2, E1)
    np.testing.assert_array_equal(A2, A1)
    _, dphi = M.basis(m, r)
    ab = np.array([M.bf16_round(v) for v in A1[0]])
    assert not np.array_equal(ab, A1[0])  # the rounding is visible at this size
    rhat = (pos[0] - pos[1]) / r
    np.testing.assert_allclose(F2[0], -2.0 * (ab @ dphi) * rhat, rtol=1e-12)
    np.testing.assert_allclose(F2[1], -F2[0], rtol=1e-14)
    # a liquid-like configuration: within bf16 rounding of mode 1 and within 3e-2 of FP64
    pos, L = cg_config(140, seed=8)
    m, nl, nn = model_for(pos, L)
    pr = M.init_params(6)
    _, F, _, _, _ = M.forces(m, pr, L, pos, nl, nn)
    _, Fb, _, _, _ = M.forces(m.with_bf16(1), pr, L, pos, nl, nn)
    _, Fb2, _, _, _ = M.forces(m.with_bf16(2), pr, L, pos, nl, nn)
    d12 = np.sqrt(((Fb2 - Fb) ** 2).mean() / (Fb ** 2).mean())
    d64 = np.sqrt(((Fb2 - F) ** 2).mean() / (F ** 2).mean())
    assert 1e-5 < d12 < 1e-2 and d64 < 3e-2


def test_forces_are_minus_energy_gradient():
    pos, L = cg_config(60, seed=3)
    m, nl, nn = model_for(pos, L)
    p = M.init_params(1)
    _, F, _, _, _ = M.forces(m, p, L, pos, nl, nn)
    h = 1e-6
    rng = np.random.default_rng(0)
    for i in rng.choice(len(pos), 8, replace=False):
        for c in range(3):
            pp = pos.copy()
            pp[i, c] += h
            Ep = M.forces(m, p, L, pp, nl, nn)[0]
            pp[i, c] -= 2 * h
            Em = M.forces(m, p, L, pp, nl, nn)[0]
            assert F[i, c] == pytest.approx(-(Ep - Em) / (2 * h), rel=1e-6, abs=1e-7)


def test_forces_match_torch_autograd():
    pos, L = cg_config(140, seed=4)
    m, nl, nn = model_for(pos, L)
    p = M.init_params(2)
    Et, F, _, V, _ = M.forces(m, p, L, pos, nl, nn)
    pairs = torch.from_numpy(pair_list(nl, nn, pos, L))
    pt = torch.tensor(pos, dtype=torch.float64, requires_grad=True)
    E = torch_energy(m, torch.tensor(p), L, pt, pairs)
    (g,) = torch.autograd.grad(E, pt)
    assert Et == pytest.approx(E.item(), rel=1e-12)
    np.testing.assert_allclose(F, -g.numpy(), rtol=1e-9, atol=1e-11)
    # sum of forces vanishes (third law); virial trace equals the pair sum r . f / 2 per site
    np.testing.assert_allclose(F.sum(axis=0), 0, atol=1e-10)


def test_parameter_gradient_matches_torch_double_backward():
    pos, L = cg_config(100, seed=5)
    m, nl, nn = model_for(pos, L)
    p = M.init_params(3)
    T = np.random.default_rng(1).normal(size=pos.shape)
    loss, grad, F = M.loss_grad(m, p, L, pos, nl, nn, T)
    pairs = torch.from_numpy(pair_list(nl, nn, pos, L))
    pt = torch.tensor(pos, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(p, requires_grad=True)
    E = torch_energy(m, th, L, pt, pairs)
    (g,) = torch.autograd.grad(E, pt, create_graph=True)
    Ft = -g
    lt = ((Ft - torch.tensor(T)) ** 2).sum() / (3 * len(pos))
    (gt,) = torch.autograd.grad(lt, th)
    assert loss == pytest.approx(lt.item(), rel=1e-12)
    np.testing.assert_allclose(F, Ft.detach().numpy(), rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(grad, gt.numpy(), rtol=1e-8, atol=1e-12 * np.abs(gt.numpy()).max())
    assert grad[M.OFF["b3"]] == 0.0  # the energy offset does not move the forces


def test_parameter_gradient_matches_finite_differences():
    pos, L = cg_config(50, seed=6)
    m, nl, nn = model_for(pos, L)
    p = M.init_params(4)
    T = np.random.default_rng(2).normal(size=pos.shape)
    _, grad, _ = M.loss_grad(m, p, L, pos, nl, nn, T)
    rng = np.random.default_rng(3)
    idx = list(rng.choice(M.NPARAM - 1, 24, replace=False)) + [M.OFF["b1"], M.OFF["w3"] + 3]
    h = 1e-6
    for k in idx:
        q = p.copy()
        q[k] += h
        lp = M.loss_grad(m, q, L, pos, nl, nn, T)[0]
        q[k] -= 2 * h
        lm = M.loss_grad(m, q, L, pos, nl, nn, T)[0]
        assert grad[k] == pytest.approx((lp - lm) / (2 * h), rel=2e-5, abs=1e-9)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(7)
    p = rng.normal(size=100)
    g = rng.normal(size=100)
    p0 = p.copy()
    m1 = np.zeros(100)
    m2 = np.zeros(100)
    M.adam(p, m1, m2, g, 1e-3, 0.9, 0.999, 1e-8, 1)
    np.testing.assert_allclose(p, p0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    # second step follows the running moments
    g2 = rng.normal(size=100)
    M.adam(p, m1, m2, g2, 1e-3, 0.9, 0.999, 1e-8, 2)
    mm = 0.9 * 0.1 * g + 0.1 * g2
    vv = 0.999 * 0.001 * g * g + 0.001 * g2 * g2
    np.testing.assert_allclose(m1, mm, rtol=1e-12)
    np.testing.assert_allclose(
        p, p0 - 1e-3 * g / (np.abs(g) + 1e-8)
        - 1e-3 * (mm / (1 - 0.81)) / (np.sqrt(vv / (1 - 0.999 ** 2)) + 1e-8), rtol=1e-10)


def test_bf16_mode_stays_close_to_fp64():
    pos, L = cg_config(140, seed=8)
    m, nl, nn = model_for(pos, L)
    p = M.init_params(6)
    _, F, E, _, _ = M.forces(m, p, L, pos, nl, nn)
    _, Fb, Eb, _, _ = M.forces(m.with_bf16(), p, L, pos, nl, nn)
    rel = np.sqrt(((Fb - F) ** 2).mean() / (F ** 2).mean())
    assert 1e-5 < rel < 3e-2
    T = np.random.default_rng(9).normal(size=pos.shape)
    _, g, _ = M.loss_grad(m, p, L, pos, nl, nn, T)
    _, gb, _ = M.loss_grad(m.with_bf16(), p, L, pos, nl, nn, T)
    assert np.sqrt(((gb - g) ** 2).sum() / (g ** 2).sum()) < 5e-2
