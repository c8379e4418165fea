"""The CPU baseline's OpenMP tile loop (oracle/ref_tile_driver.c) over the
reference's production-built kernels: it runs, counts whole tiles of both
species, and leaves the caller's pool untouched (each tile starts from a
copy).  Skipped where the reference kernels were not built."""
import numpy as np
import pytest


def test_tile_driver_counts_whole_tiles():
    from oracle import load_ref_driver
    drv = load_ref_driver()
    if drv is None:
        pytest.skip("oracle/_ref/libpxrref_driver.so not built (no reference tree)")
    rng = np.random.default_rng(0)
    nc, ng, ntile, nthreads = 7, 3, 4096, 2
    shape = (nc + 1 + 2 * ng,) * 3
    fields = [np.asfortranarray(rng.standard_normal(shape) * 1e-3) for _ in range(6)]
    pool = np.empty((nthreads, 2, 8, ntile))
    for t in range(nthreads):
        for s, uth in enumerate((0.01, 0.001)):
            pool[t, s, :3] = rng.uniform(-0.5, nc + 0.5, (3, ntile))
            pool[t, s, 3:6] = rng.standard_normal((3, ntile)) * uth
            pool[t, s, 6] = 1.0 / np.sqrt(1.0 + np.sum(pool[t, s, 3:6] ** 2, axis=0))
            pool[t, s, 7] = 1.0
    before = pool.copy()
    steps, wall = drv.time_tiles(nthreads, 0.2, nc, ng, fields, pool, [-1.0, 1.0], [1.0, 1836.0],
                                 0.5)
    assert wall > 0.0 and steps > 0
    assert steps % (2 * ntile) == 0  # whole tiles, both species
    np.testing.assert_array_equal(pool, before)
