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


def test_rdf_counts_pairs_of_a_known_lattice():
    """Simple-cubic lattice, spacing 1 (rho = 1): the first shell holds 6 neighbours at r = 1
    and the second 12 at sqrt(2), so g times the shell volumes sums to those counts."""
    from hoomd_amd.md.mlcg import rdf
    n = 12
    g1 = np.arange(n) - n / 2 + 0.5
    pos = np.stack(np.meshgrid(g1, g1, g1, indexing="ij"), -1).reshape(-1, 3)
    r, g = rdf(pos, float(n), 1.6, 32)
    dr = 1.6 / 32
    vol = 4.0 / 3.0 * np.pi * ((r + dr / 2) ** 3 - (r - dr / 2) ** 3)
    counts = g * vol
    assert abs(counts[(r > 0.95) & (r < 1.05)].sum() - 6) < 1e-9
    assert abs(counts[(r > 1.39) & (r < 1.44)].sum() - 12) < 1e-9
    assert abs(counts.sum() - 18) < 1e-9


def test_regrid_table_keeps_forces():
    from hoomd_amd.md.mlcg import regrid_table, rdf_prior_table
    r = (np.arange(80) + 0.5) * 0.05
    U = np.where(r < 1.5, 3.0 * (1 - r / 1.5) ** 2, 0.0)
    params, rc = rdf_prior_table(r, np.exp(-U), kT=1.0, g_join=0.9)
    ext = regrid_table(params, rc, 3.0, width=600)
    rr = np.linspace(0.0, 3.0, 600, endpoint=False)
    ro = np.linspace(0.0, rc, len(params["U"]), endpoint=False)
    assert np.all(ext["U"][rr >= rc] == 0.0)
    inside = rr < ro[-1]
    assert np.abs(ext["U"][inside] - np.interp(rr[inside], ro, params["U"])).max() < 1e-12


def test_ibi_update_moves_toward_target():
    """Where the model over-populates (g > g_target) the potential rises by damping kT ln(g /
    g_target); where the RDFs agree it is unchanged; U(r_cut) stays 0; F = -dU/dr."""
    from hoomd_amd.md.mlcg import ibi_update
    rc = 3.0
    rt = np.linspace(0.0, rc, 600, endpoint=False)
    params = dict(r_min=0.0, U=np.zeros(600), F=np.zeros(600))
    r = (np.arange(80) + 0.5) * 0.05
    g_t = np.clip((r - 0.6) / 0.6, 0.0, 1.0)
    g_c = g_t.copy()
    bump = (r > 1.3) & (r < 1.7)
    g_c[bump] *= 1.2
    out = ibi_update(params, rc, r, g_c, g_t, kT=1.0, damping=0.5)
    U = out["U"]
    assert np.all(np.isfinite(U)) and np.allclose(out["F"], -np.gradient(U, rt))
    mid = (rt > 1.4) & (rt < 1.6)
    assert np.allclose(U[mid], 0.5 * np.log(1.2), rtol=0.05)
    far = (rt > 2.0) & (rt < 2.6)
    assert np.abs(U[far]).max() < 1e-12
    assert abs(U[-1]) < 0.5 * np.log(1.2) * 1e-2
    # identical RDFs: no change
    same = ibi_update(params, rc, r, g_t, g_t)
    assert np.abs(same["U"]).max() < 1e-15


def test_pressure_ramp_sign_and_shape():
    """A model below its target pressure (dp < 0) gets U += A (1 - r / r_cut) with A > 0: a
    repulsive force A / r_cut everywhere, zero energy change at r_cut; A is capped at scale kT."""
    from hoomd_amd.md.mlcg import pressure_ramp
    rc = 3.0
    rt = np.linspace(0.0, rc, 300, endpoint=False)
    params = dict(r_min=0.0, U=np.zeros(300), F=np.zeros(300))
    out = pressure_ramp(params, rc, dp=-4.0, kT=1.0, scale=0.1)
    assert np.allclose(out["U"], 0.1 * (1 - rt / rc))
    assert np.allclose(out["F"], 0.1 / rc)
    small = pressure_ramp(params, rc, dp=0.5, kT=1.0, scale=0.1, dp_scale=1.0)
    assert np.allclose(small["U"], -0.05 * (1 - rt / rc))
