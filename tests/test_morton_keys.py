"""The Morton order of the reference's recursive radix partition
(src/applications/particles.c:1090-1221, restated by the oracle's
oracle_zsort) expressed as a plain sort, the form the GPU sort uses:

  * u_d = (uint32)((x_d - min_d) / (max_d - min_d) * UINT32_MAX) per point;
  * key = sum over bits b and dimensions d of bit b of u_d at position
    b * ndim + d (the partition visits bit 31 of dimension ndim-1 first,
    then bit 31 of ndim-2, ..., bit 0 of dimension 0);
  * the partition moves 1-bit items to the back in reverse order, so two
    points with equal keys swap once per 1 bit of the key: among equal keys
    the original order is kept when the key has an even number of 1 bits
    and reversed when odd.
Sorting by (key, index or reversed index by that parity) must give the
oracle's permutation exactly, duplicates included.
"""
import numpy as np

from oracle import oracle as O


def _keyed_order(points):
    ndim, n = points.shape
    mn = points.min(axis=1)
    mx = points.max(axis=1)
    u = ((points - mn[:, None]) / (mx - mn)[:, None] *
         4294967295.0).astype(np.uint32)
    # key as a Python int per point (ndim * 32 bits)
    keys = []
    for j in range(n):
        k = 0
        for b in range(31, -1, -1):
            for d in range(ndim - 1, -1, -1):
                k = (k << 1) | ((int(u[d, j]) >> b) & 1)
        keys.append(k)
    tb = [(n - 1 - j) if bin(keys[j]).count("1") & 1 else j
          for j in range(n)]
    return sorted(range(n), key=lambda j: (keys[j], tb[j]))


def _check(points):
    _, order = O.zsort(points)
    assert list(order) == _keyed_order(points)


def test_keyed_order_random_points():
    rng = np.random.default_rng(1)
    for ndim in (1, 2, 3, 4):
        _check(rng.random((ndim, 700)))


def test_keyed_order_with_many_duplicates():
    rng = np.random.default_rng(2)
    for ndim in (2, 3):
        # a coarse grid: every key occurs many times, with both parities
        pts = rng.integers(0, 5, (ndim, 900)).astype(np.float64) / 4.0
        _check(pts)


def test_keyed_order_clustered_points():
    rng = np.random.default_rng(3)
    centres = rng.random((3, 10))
    pts = (centres[:, rng.integers(0, 10, 800)] +
           1e-9 * rng.standard_normal((3, 800)))
    _check(pts)
