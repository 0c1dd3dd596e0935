"""The oracle's exact engines pinned against independent computations:

  oracle_dlrsdd (sequential/dense/dsdd.c:16-66) against numpy's SVD of the
  same tile with the rank rule written out from dsvfr.c's definition;
  oracle_dlrqp3 (sequential/dense/dqp3.c:16-93) against scipy's pivoted QR
  (scipy.linalg.qr(pivoting=True)) followed by numpy's SVD of the leading
  rows of R.

Tiles have a prescribed spectrum (random orthogonal factors, geometric
singular values) so that ranks and pivots have no near ties; the cases
cover low rank, rank above maxrank, and rank >= min(m, n)/2 (a false
far-field tile for both engines, unlike the randomized engine), and for
RRQR a tile whose range exceeds maxrank + oversample.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O

pytestmark = pytest.mark.skipif(not O.lapack_load(),
                                reason="no host LAPACK for the oracle")


def _tile(m, n, decay, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    mn = min(m, n)
    U, _ = np.linalg.qr(rng.standard_normal((m, mn)))
    V, _ = np.linalg.qr(rng.standard_normal((n, mn)))
    s = scale * decay ** np.arange(mn)
    return (U * s) @ V.T, s


def _rank(s, tol):
    """Smallest r with sum_{i >= r} s_i^2 <= tol^2 sum s_i^2 (at least 1
    for a nonzero s), the rule of dsvfr.c:16-47."""
    s2 = np.asarray(s, dtype=np.float64) ** 2
    total = s2.sum()
    if total == 0:
        return 0
    tail = np.concatenate([np.cumsum(s2[::-1])[::-1], [0.0]])
    for r in range(1, len(s2) + 1):
        if tail[r] <= tol * tol * total:
            return r
    return len(s2)


@pytest.mark.parametrize("m,n,decay,maxrank,tol,want", [
    (120, 100, 0.7, 60, 1e-6, "low"),      # rank 39 < 50, <= 60
    (120, 100, 0.7, 20, 1e-6, "dense"),    # rank > maxrank
    (100, 100, 0.95, 90, 1e-6, "dense"),   # rank >= mn/2
    (80, 150, 0.5, 40, 1e-9, "low"),       # wide tile
])
def test_dlrsdd_against_numpy_svd(m, n, decay, maxrank, tol, want):
    D, _ = _tile(m, n, decay, seed=m + n)
    r, U, V, S = O.dlrsdd(D.copy(), maxrank, tol)
    s = np.linalg.svd(D, compute_uv=False)
    np.testing.assert_allclose(S, s, rtol=1e-12, atol=1e-14 * s[0])
    rr = _rank(s, tol)
    mn = min(m, n)
    expect = rr if (rr < mn // 2 and rr <= maxrank) else -1
    assert r == expect
    assert (r >= 0) == (want == "low")
    if r > 0:
        err = np.linalg.norm(D - U @ V.T) / np.linalg.norm(D)
        assert err <= tol
        # U orthonormal, V = (S Vs^T)^T
        np.testing.assert_allclose(U.T @ U, np.eye(r), atol=1e-12)
        np.testing.assert_allclose(np.linalg.norm(V, axis=0), s[:r],
                                   rtol=1e-10, atol=1e-14 * s[0])


@pytest.mark.parametrize("m,n,decay,maxrank,os_,tol", [
    (120, 100, 0.7, 60, 10, 1e-6),     # range below mn2: the SVD result
    (150, 140, 0.9, 30, 10, 1e-5),     # range beyond mn2 = 40
    (100, 130, 0.8, 25, 5, 1e-4),      # wide, range beyond mn2
])
def test_dlrqp3_against_scipy_pivoted_qr(m, n, decay, maxrank, os_, tol):
    D, s_true = _tile(m, n, decay, seed=3 * m + n)
    # distinct column norms (no pivot ties): scale columns
    D = D * np.linspace(1.0, 2.0, n)[None, :]
    r, U, V, S, piv = O.dlrqp3(D.copy(), maxrank, os_, tol)
    Qs, Rs, Ps = scipy.linalg.qr(D, pivoting=True, mode="economic")
    mn = min(m, n)
    mn2 = min(maxrank + os_, mn)
    np.testing.assert_array_equal(piv, Ps)
    s = np.linalg.svd(Rs[:mn2, :], compute_uv=False)
    np.testing.assert_allclose(S, s, rtol=1e-10, atol=1e-13 * s[0])
    rr = _rank(s, tol)
    expect = rr if (rr < mn // 2 and rr <= maxrank) else -1
    assert r == expect
    if r > 0:
        # U V^T is the best rank-r approximation of Q1 Q1^T D
        Q1 = Qs[:, :mn2]
        P = Q1 @ (Q1.T @ D)
        err = np.linalg.norm(P - U @ V.T) / np.linalg.norm(P)
        assert err <= tol


def test_dlrqp3_equals_dlrsdd_when_range_fits():
    """When the tile's range beyond maxrank + oversample dimensions is
    negligible (here below 1e-11 |D|), R[:mn2, :] holds the rest and both
    engines agree to that level."""
    D, _ = _tile(90, 90, 0.6, seed=5)
    r1, U1, V1, S1 = O.dlrsdd(D.copy(), 40, 1e-8)
    r2, U2, V2, S2, _ = O.dlrqp3(D.copy(), 40, 10, 1e-8)
    assert r1 == r2 > 0
    np.testing.assert_allclose(S2[:r1], S1[:r1], rtol=0, atol=1e-12 * S1[0])
    np.testing.assert_allclose(U1 @ V1.T, U2 @ V2.T, atol=1e-10)
