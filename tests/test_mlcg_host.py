"""CPU checks of the learned-CG host tools: the RDF prior, the Boltzmann-inverted bonds and the
RDF estimator, against closed forms."""

import numpy as np


def test_rdf_prior_table_inverts_a_known_rdf():
    from hoomd_amd.md.mlcg import rdf_prior_table
    # g(r) = exp(-U(r)/kT) for a soft repulsion U = 3 (1 - r/1.5)^2 below 1.5, zero beyond
    r = (np.arange(80) + 0.5) * 0.05
    U = np.where(r < 1.5, 3.0 * (1 - r / 1.5) ** 2, 0.0)
    g = np.exp(-U)
    params, rc = rdf_prior_table(r, g, kT=1.0, g_join=0.9)
    # the cut-off is the first bin where g reaches 0.9: U = -ln 0.9
    assert abs(rc - r[np.nonzero(g >= 0.9)[0][0]]) < 1e-12
    rr = np.linspace(0.0, rc, len(params["U"]), endpoint=False)
    inside = rr >= r[0]
    # U recovered up to the shift that makes it vanish at the cut-off
    ref = np.interp(rr, r, U) - np.interp(rc, r, U)
    assert np.abs(params["U"][inside] - ref[inside]).max() < 1e-9
    # F = -dU/dr of the tabulated U, positive (repulsive) everywhere
    assert np.all(params["F"] > 0)
    dU = np.gradient(params["U"], rr)
    assert np.allclose(params["F"], -dU)


def test_rdf_prior_wall_below_sampled_range():
    from hoomd_amd.md.mlcg import rdf_prior_table
    r = (np.arange(80) + 0.5) * 0.05
    g = np.clip((r - 0.6) / 0.8, 0.0, 1.0)  # no samples below 0.6
    params, rc = rdf_prior_table(r, g, kT=1.0, g_join=0.9, min_count_g=0.02)
    U = params["U"]
    assert np.all(np.isfinite(U)) and np.all(np.diff(U) <= 1e-12)  # a wall, falling with r
    assert U[0] > U[len(U) // 2] > 0


def test_boltzmann_harmonic_recovers_spring():
    from hoomd_amd.md.mlcg import boltzmann_harmonic
    rng = np.random.default_rng(0)
    k, r0 = 4.0, 1.7
    # bond vectors with U = k (r - r0)^2 / 2 in 3D: sample r with density r^2 exp(-U)
    r = rng.normal(r0, 1 / np.sqrt(k), 2_000_000)
    w = r ** 2
    keep = rng.uniform(0, w.max(), len(r)) < w
    kk, rr = boltzmann_harmonic(r[keep], kT=1.0)
    assert abs(kk / k - 1) < 0.03 and abs(rr - r0) < 0.01


def test_rdf_of_ideal_gas_is_one():
    from hoomd_amd.md.mlcg import rdf
    rng = np.random.default_rng(1)
    L = 20.0
    pos = rng.uniform(-L / 2, L / 2, (20000, 3))
    r, g = rdf(pos, L, 4.0, 20)
    assert np.abs(g[4:] - 1).max() < 0.05
