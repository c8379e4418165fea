"""pxrhip_particle_bin_sorting_noalloc (the sort with a fixed device scratch
array) against pxrhip_particle_bin_sorting (scratch it allocates), through
the C ABI: bit-identical results on two calls in a row, the second smaller,
and the oracle's stable counting sort (sorting.F90:247-361) on the same
input; limits and key errors rejected before anything moves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP_MAX = 1 << 20
NBINS_MAX = 1 << 20


@pytest.fixture(scope="module")
def lib():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)
    return _lib


def particles(rng, n, lo, hi):
    P = [rng.uniform(lo, hi, n) for _ in range(3)] + [rng.standard_normal(n) for _ in range(4)]
    P.append(np.arange(n, dtype=np.float64))  # pid(:, 1): the input order
    return P


def upload(P):
    from picsar_amd.devmem import DeviceArray
    return [DeviceArray.from_numpy(p) for p in P]


def sort_both(lib, orc, P, box):
    from picsar_amd.devmem import synchronize
    n = len(P[0])
    A, B = upload(P), upload(P)
    lib.call("pxrhip_particle_bin_sorting", n, *[a.ptr for a in A], 1, *box)
    lib.call("pxrhip_particle_bin_sorting_noalloc", n, *[b.ptr for b in B], 1, *box)
    synchronize()
    Q = [p.copy() for p in P]
    assert orc.particle_bin_sorting(*Q, *box) == 0
    for a, b, q in zip(A, B, Q):
        ha, hb = a.numpy(), b.numpy()
        np.testing.assert_array_equal(hb, ha)
        np.testing.assert_array_equal(hb, q)


def test_noalloc_sort_matches_the_allocating_sort(lib, orc):
    rng = np.random.default_rng(2024)
    # call 1: a 24^3-node tile with half-cell margins, 400k particles
    box1 = (-0.5, -0.5, -0.5, 24.5, 24.5, 24.5, 1.0, 1.0, 1.0)
    sort_both(lib, orc, particles(rng, 400000, -0.5, 24.5), box1)
    # call 2, smaller: a 10^3-node tile with 2-cell sorting cells, 30k particles;
    # the fixed array still holds call 1's larger contents
    box2 = (0.0, 0.0, 0.0, 10.0, 10.0, 10.0, 0.5, 0.5, 0.5)
    sort_both(lib, orc, particles(rng, 30000, 0.0, 10.0), box2)


def test_noalloc_sort_limits_and_errors(lib):
    from picsar_amd import PxrHipError
    from picsar_amd._lib import ERR_ARG, ERR_BOUNDS
    from picsar_amd.devmem import DeviceArray
    d = [DeviceArray(4) for _ in range(8)]
    ptrs = [a.ptr for a in d]
    # more particles than the fixed array holds: rejected before any device work
    with pytest.raises(PxrHipError) as ei:
        lib.call("pxrhip_particle_bin_sorting_noalloc", NP_MAX + 1, *ptrs, 1,
                 0, 0, 0, 4, 4, 4, 1, 1, 1)
    assert ei.value.code == ERR_ARG and "exceeds" in str(ei.value)
    # more bins than it holds
    with pytest.raises(PxrHipError) as ei:
        lib.call("pxrhip_particle_bin_sorting_noalloc", 4, *ptrs, 1,
                 0, 0, 0, 1025, 1024, 1, 1, 1, 1)
    assert ei.value.code == ERR_ARG and "bins" in str(ei.value)
    # a key outside the sorting grid: reported, nothing moved
    P = [DeviceArray.from_numpy(np.array([0.5, 0.5, 9.0])) for _ in range(3)] + \
        [DeviceArray.from_numpy(np.array([1.0, 2.0, 3.0])) for _ in range(5)]
    with pytest.raises(PxrHipError) as ei:
        lib.call("pxrhip_particle_bin_sorting_noalloc", 3, *[p.ptr for p in P], 1,
                 0, 0, 0, 2, 2, 2, 1, 1, 1)
    assert ei.value.code == ERR_BOUNDS
    np.testing.assert_array_equal(P[7].numpy(), [1.0, 2.0, 3.0])
    # the error left the counter clear: the next call sorts
    P = [DeviceArray.from_numpy(np.array([1.5, 0.5, 1.2])) for _ in range(3)] + \
        [DeviceArray.from_numpy(np.array([1.0, 2.0, 3.0])) for _ in range(5)]
    lib.call("pxrhip_particle_bin_sorting_noalloc", 3, *[p.ptr for p in P], 1,
             0, 0, 0, 2, 2, 2, 1, 1, 1)
    np.testing.assert_array_equal(P[7].numpy(), [2.0, 1.0, 3.0])
    # np2 = 0 is a no-op
    lib.call("pxrhip_particle_bin_sorting_noalloc", 0, *ptrs, 1, 0, 0, 0, 4, 4, 4, 1, 1, 1)
