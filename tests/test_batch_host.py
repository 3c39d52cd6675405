"""CPU checks of the batched-replica workload recipe and the batch C-ABI's argument checks (no
compute calls: the checks below fail before anything touches a GPU)."""

import numpy as np


def test_seeded_recipe_ranges_and_determinism():
    from hoomd_amd.batch import seeded_replicas
    n, L, kT, seeds, pos = seeded_replicas(256, 1)
    n2, L2, kT2, seeds2, pos2 = seeded_replicas(256, 1)
    assert np.array_equal(n, n2) and np.array_equal(L, L2) and np.array_equal(kT, kT2)
    assert all(np.array_equal(a, b) for a, b in zip(pos, pos2))
    assert n.min() >= 500 and n.max() <= 20000
    rho = n / L ** 3
    assert rho.min() >= 0.70 - 1e-12 and rho.max() <= 0.90 + 1e-12
    assert kT.min() >= 0.8 and kT.max() <= 1.5
    assert len(set(seeds.tolist())) == 256
    # log-uniform N: the mean of ln N is the middle of [ln 500, ln 20000] within 4 sigma
    mid = 0.5 * (np.log(500) + np.log(20000))
    sd = (np.log(20000) - np.log(500)) / np.sqrt(12) / np.sqrt(256)
    assert abs(np.log(n).mean() - mid) < 4 * sd
    # total particles: mean of the log-uniform law times 256 within 20 %
    mean_n = (20000 - 500) / np.log(20000 / 500)
    assert abs(n.sum() / (256 * mean_n) - 1) < 0.2


def test_fcc_fill_sites_inside_box_and_apart():
    from scipy.spatial import cKDTree

    from hoomd_amd.batch import fcc_fill
    rng = np.random.default_rng(0)
    for N, rho in ((500, 0.9), (777, 0.7), (4001, 0.85)):
        L = (N / rho) ** (1 / 3)
        p = fcc_fill(N, L, rng)
        assert p.shape == (N, 3)
        assert p.min() >= -L / 2 and p.max() < L / 2
        n = int(np.ceil((N / 4) ** (1 / 3) - 1e-9))
        d, _ = cKDTree(np.mod(p + L / 2, L), boxsize=L).query(np.mod(p + L / 2, L), k=2)
        # nearest FCC neighbours sit a / sqrt(2) apart
        assert d[:, 1].min() > 0.99 * (L / n) / np.sqrt(2)


def test_create_rejects_small_box_and_bad_arguments():
    from hoomd_amd import _lib
    L = _lib.lib()
    n = np.array([10], np.uint32)
    box = np.array([5.0])  # 2 (r_cut + r_buff) = 5.8 > 5
    kT = np.array([1.0])
    seeds = np.array([1], np.uint16)
    pos = np.zeros((10, 3))
    h = L.hm_batch_create(1, n.ctypes.data, box.ctypes.data, kT.ctypes.data, seeds.ctypes.data,
                          pos.ctypes.data, None, None, 1.0, 1.0, 2.5, 0.4, 0.005, 0.0)
    assert not h
    assert b"too small" in L.hm_last_error()
    h = L.hm_batch_create(1, n.ctypes.data, box.ctypes.data, kT.ctypes.data, seeds.ctypes.data,
                          pos.ctypes.data, None, None, 1.0, 1.0, -1.0, 0.4, 0.005, 0.0)
    assert not h
    assert b"invalid" in L.hm_last_error()
