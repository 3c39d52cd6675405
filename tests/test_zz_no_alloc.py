"""hm_thermo_noalloc (partial sums in the library's per-device array, no allocation) against
hm_thermo (partial sums in stream-allocated scratch) through the C ABI: two calls in a row, the
second on a smaller input, each bit-identical to hm_thermo on the same input (same kernels, same
fixed summation order)."""

import numpy as np
import pytest

from gpu_helpers import dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    from hoomd_amd import _lib
    if _lib.lib().hm_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return _lib


def thermo_both(H, N, nforces, seed):
    rng = np.random.default_rng(seed)
    v4 = np.zeros((N, 4))
    v4[:, :3] = rng.normal(size=(N, 3))
    v4[:, 3] = rng.uniform(0.5, 2, N)
    fs = [dev(rng.normal(size=(N, 4)).astype(np.float32)) for _ in range(nforces)]
    vs = [dev(rng.normal(size=(6, N)).astype(np.float32)) for _ in range(nforces)]
    d_v = dev(v4)
    fp = H.pointer_array([f.ptr for f in fs])
    vp = H.pointer_array([v.ptr for v in vs])
    a = H.DeviceArray((15,), np.float64, fill=0)
    b = H.DeviceArray((15,), np.float64, fill=0)
    H.check(H.lib().hm_thermo(N, d_v.ptr, fp, vp, nforces, a.ptr, None), "hm_thermo")
    H.check(H.lib().hm_thermo_noalloc(N, d_v.ptr, fp, vp, nforces, b.ptr, None),
            "hm_thermo_noalloc")
    H.check(H.lib().hm_synchronize())
    ke = 0.5 * (v4[:, 3] * (v4[:, :3] ** 2).sum(axis=1)).sum()
    return a.numpy(), b.numpy(), ke


def test_noalloc_equals_allocating_entry_point(H):
    for N, nf, seed in ((200000, 2, 1), (3000, 1, 2)):  # second call smaller than the first
        ref, got, ke = thermo_both(H, N, nf, seed)
        print(f"N={N} nforces={nf}: KE {got[0]:.12e} (host {ke:.12e})", flush=True)
        assert np.array_equal(ref, got)
        assert abs(got[0] - ke) <= 1e-12 * ke


def test_noalloc_rejects_bad_arguments(H):
    out = H.DeviceArray((15,), np.float64, fill=0)
    assert H.lib().hm_thermo_noalloc(10, None, None, None, 0, out.ptr, None) == -2
    v = H.DeviceArray((40,), np.float64, fill=0)
    assert H.lib().hm_thermo_noalloc(10, v.ptr, None, None, 5, out.ptr, None) == -2
