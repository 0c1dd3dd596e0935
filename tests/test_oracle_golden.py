"""Pins the oracle (CPU restatement of the reference) before it is trusted.

Each function of oracle/oracle.c is checked against something that fixes
its answer independently: the host LAPACK's dlarnv output (golden
fixtures), closed forms and scipy's Bessel functions, brute force on tiny
inputs, library SVD/QR routines, and mathematical invariants.
"""
import ctypes as C
import hashlib
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- dlarnv test matrix ------------------------------------------------

def test_omega_matches_lapack_dlarnv_prefix():
    g = np.load(os.path.join(GOLD, "omega_golden.npz"))
    x = O.omega(1024, 1).ravel()
    assert np.array_equal(x, g["normals"])


def test_omega_matches_lapack_dlarnv_full_digests():
    g = np.load(os.path.join(GOLD, "omega_golden.npz"))
    for key, val in zip(g["digest_keys"], g["digest_vals"]):
        ncols, k = map(int, str(key).split("x"))
        x = O.omega(ncols, k).T.ravel()     # stream order = column-major
        assert hashlib.sha256(x.tobytes()).hexdigest() == str(val), key


def test_omega_uniform_stream_closed_form():
    # the uniforms of dlaruv are the 48-bit MCG states times 2^-48
    g = np.load(os.path.join(GOLD, "omega_golden.npz"))
    a, s, us = 33952834046453, 1, []
    for _ in range(300):
        s = (a * s) % (1 << 48)
        us.append(s / float(1 << 48))
    assert np.array_equal(np.array(us), g["uniforms"])
    # 2048 uniforms consumed by 1024 normals: seed advanced by a^2048
    s = pow(a, 2048, 1 << 48)
    limbs = [(s >> 36) & 4095, (s >> 24) & 4095, (s >> 12) & 4095, s & 4095]
    assert limbs == list(g["seed_after"])


def test_omega_is_prefix_reshaped_for_ragged_tiles():
    # a narrower tile uses the first ncols*k values with ld = ncols
    full = O.omega(1600, 416).T.ravel()
    small = O.omega(160, 400).T.ravel()
    assert np.array_equal(full[:small.size], small)


# ---- kernels ------------------------------------------------------------

def _kv(family, nu=None):
    g = json.load(open(os.path.join(GOLD, "kernel_golden.json")))
    return g


@pytest.mark.parametrize("family,key,nu", [
    ("exp", "exp", 0.5), ("sqrexp", "sqrexp", 0.5),
    ("matern", "matern_0.5", 0.5), ("matern", "matern_1.5", 1.5),
    ("matern", "matern_2.5", 2.5), ("matern2", "matern2_0.5", 0.5),
    ("matern2", "matern2_1.5", 1.5), ("matern2", "matern2_2.5", 2.5),
    ("coulomb", "coulomb", 0.5), ("sin", "sin", 0.5), ("cos", "cos", 0.5)]
    + [(fam, f"{fam}_{nu}", nu) for fam in ("matern", "matern2")
       for nu in (0.3, 0.8, 1.2, 2.2, 3.7, 6.0)])
def test_kernel_values_match_closed_forms(family, key, nu):
    g = json.load(open(os.path.join(GOLD, "kernel_golden.json")))
    kd = O.kernel(family, 3, beta=g["beta"], nu=nu, noise=g["noise"],
                  sigma=g["sigma"], wave_k=g["wave_k"], diag=g["diag"])
    for rec in g["values"]:
        x1 = np.zeros((3, 1))
        x2 = np.array([[rec["r"]], [0.0], [0.0]])
        v = O.kernel_block(kd, x1, x2, [0], [0])[0, 0]
        want = rec[key]
        # general nu: scipy's kv (AMOS) is itself accurate to a few 1e-14
        tol = 4e-14 if nu in (0.5, 1.5, 2.5) else 2e-13
        assert abs(v - want) <= tol * max(abs(want), 1e-300), \
            (key, rec["r"], v, want)


def test_kernel_tile_is_column_major_with_leading_dimension():
    rng = np.random.default_rng(0)
    x = rng.random((2, 30))
    kd = O.kernel("exp", 2, beta=0.3)
    A = O.kernel_block(kd, x, x, np.arange(5, 17), np.arange(2, 9))
    for i in range(12):
        for j in range(7):
            r = math.dist(x[:, 5 + i], x[:, 2 + j])
            want = 1.0 if r == 0 else math.exp(-r / 0.3)
            assert abs(A[i, j] - want) <= 1e-15


# ---- dsvfr --------------------------------------------------------------

def _rank_exact(S, tol):
    """Smallest r >= 1 with sum_{i>=r} S_i^2 <= tol^2 sum S_i^2, exactly."""
    Sf = [Fraction(s) ** 2 for s in S]
    tot = sum(Sf)
    if tot == 0:
        return 0
    t2 = Fraction(tol) ** 2
    for r in range(1, len(S) + 1):
        if sum(Sf[r:]) <= t2 * tot:
            return r
    return len(S)


def test_dsvfr_table():
    assert O.dsvfr([0.0, 0.0, 0.0], 1e-3) == 0
    assert O.dsvfr([5.0], 1e-3) == 1
    assert O.dsvfr([1.0, 0.0, 0.0], 0.5) == 1      # never 0 unless S == 0
    assert O.dsvfr([1.0, 1.0, 1.0], 0.0) == 3
    # tie exactly at the threshold is accepted (>= comparison):
    # S = [1, 1, 1, 1], tol^2 = 1/4: limit 1, tail from 3 is 1 -> rank 3
    assert O.dsvfr([1.0, 1.0, 1.0, 1.0], 0.5) == 3
    assert O.dsvfr([2.0, 1.0], 0.5) == 1
    assert O.dsvfr([2.0, 1.0], 0.4) == 2


def test_dsvfr_matches_exact_rule_away_from_ties():
    rng = np.random.default_rng(1)
    for trial in range(300):
        k = int(rng.integers(1, 40))
        S = np.sort(rng.random(k) ** rng.integers(1, 20))[::-1]
        tol = 10.0 ** rng.uniform(-9, -1)
        r = O.dsvfr(S, tol)
        rx = _rank_exact(S, tol)
        if r != rx:
            tot = np.sum(S ** 2)
            tail = np.sum(S[rx:] ** 2)
            assert abs(tail - tol * tol * tot) <= 1e-12 * tol * tol * tot


# ---- dense RSDD ----------------------------------------------------------

def _spectrum(m, n, s, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, len(s))))
    V, _ = np.linalg.qr(rng.standard_normal((n, len(s))))
    return (U * s) @ V.T


@pytest.mark.parametrize("m,n,decay,tol", [
    (120, 120, 0.5, 1e-9), (150, 100, 0.7, 1e-6), (90, 130, 0.3, 1e-12)])
def test_dlrrsdd_rank_on_prescribed_spectrum(m, n, decay, tol):
    s = decay ** np.arange(min(m, n) // 2)
    A = _spectrum(m, n, s, 3)
    r, U, V, S = O.dlrrsdd(A, 60, 10, tol)
    assert r == _rank_exact(s, tol) or abs(
        np.sum(s[r:] ** 2) - tol ** 2 * np.sum(s ** 2)) < 1e-6 * tol ** 2
    # singular values of B equal the prescribed ones to rounding
    kk = min(len(S), len(s))
    assert np.max(np.abs(S[:kk] - s[:kk])) <= 1e-13 * s[0]
    err = np.linalg.norm(A - U @ V.T)
    assert abs(err - np.sqrt(np.sum(s[r:] ** 2))) <= 1e-12 * s[0]


def test_dlrrsdd_matches_library_qr_svd_on_same_sketch():
    """Same Omega and steps with scipy's QR and SVD give the same S, rank
    and approximation error."""
    rng = np.random.default_rng(5)
    x = rng.random((2, 400))
    kd = O.kernel("sqrexp", 2, beta=0.2)
    A = O.kernel_block(kd, x, x, np.arange(0, 200), np.arange(200, 400))
    maxrank, os_, tol = 80, 10, 1e-5
    r, U, V, S = O.dlrrsdd(A, maxrank, os_, tol)
    k = min(maxrank + os_, 200)
    Om = O.omega(200, k)
    Q, _ = sl.qr(A @ Om, mode="economic")
    B = Q.T @ A
    Ss = sl.svd(B, compute_uv=False)
    assert np.max(np.abs(Ss - S)) <= 1e-13 * Ss[0]
    rs = O.dsvfr(Ss, tol)
    assert 0 < rs <= maxrank and r == rs
    err_o = np.linalg.norm(A - U @ V.T)
    Ub, sb, Vbt = sl.svd(B, full_matrices=False)
    err_s = np.linalg.norm(A - (Q @ Ub[:, :r]) @ (sb[:r, None] * Vbt[:r]))
    assert abs(err_o - err_s) <= 1e-10 * np.linalg.norm(A)


def test_dlrrsdd_false_far_when_rank_exceeds_maxrank():
    A = _spectrum(100, 100, np.ones(50), 9)
    r, U, V, S = O.dlrrsdd(A, 20, 5, 1e-9)
    assert r == -1


def test_dlrrsdd_zero_tile_has_rank_zero():
    r, U, V, S = O.dlrrsdd(np.zeros((40, 40)), 10, 5, 1e-9)
    assert r == 0


# ---- particles ------------------------------------------------------------

def test_quasiuniform1_matches_libc_rand_reproduction():
    libc = C.CDLL(None)
    libc.srand(C.c_uint(3))
    count, ndim = 50, 2
    side = math.ceil(count ** (1.0 / ndim))
    mult, add = 0.8 / side / 2147483647, 0.1 / side
    want = np.zeros((ndim, count))
    idx = [0] * ndim
    for i in range(count):
        for d in range(ndim):
            want[d, i] = idx[d] / side + add + mult * libc.rand()
        d = ndim - 1
        idx[d] += 1
        while idx[d] == side:
            idx[d] = 0
            if d > 0:
                d -= 1
                idx[d] += 1
    got = O.gen_quasiuniform1(count, ndim, 3)
    assert np.array_equal(got, want)


def _morton_keys(pts):
    ndim, n = pts.shape
    mn, mx = pts.min(1), pts.max(1)
    u = ((pts - mn[:, None]) / (mx - mn)[:, None] * 4294967295.0).astype(
        np.uint64)
    keys = []
    for j in range(n):
        key = 0
        for bit in range(31, -1, -1):
            for d in range(ndim - 1, -1, -1):
                key = (key << 1) | int((u[d, j] >> np.uint64(bit)) & 1)
        keys.append(key)
    return keys


@pytest.mark.parametrize("ndim,n", [(2, 300), (3, 200)])
def test_zsort_is_morton_order(ndim, n):
    pts = O.gen_rand(n, ndim, 11)
    srt, order = O.zsort(pts)
    keys = _morton_keys(pts)
    ks = [keys[o] for o in order]
    assert all(ks[i] <= ks[i + 1] for i in range(n - 1))
    assert sorted(order.tolist()) == list(range(n))
    assert np.array_equal(srt, pts[:, order])


def test_zsort_tie_order_reverses_set_bit_groups():
    # identical keys: the partition writes 1-bit items from the back, so a
    # group of equal points with all bits set ends reversed an even/odd
    # number of times; check against a direct restatement of the partition.
    pts = np.array([[0.0, 1.0, 1.0, 1.0, 0.5], [0.0, 1.0, 1.0, 1.0, 0.5]])
    srt, order = O.zsort(pts)

    def part(order, keys, ndim, sdim, sbit, lo, hi):
        tmp = list(order)
        a, b = lo, hi
        for i in range(lo, hi + 1):
            if (keys[sdim][order[i]] >> sbit) & 1 == 0:
                tmp[a] = order[i]
                a += 1
            else:
                tmp[b] = order[i]
                b -= 1
        order[lo:hi + 1] = tmp[lo:hi + 1]
        if sdim > 0:
            nd, nb = sdim - 1, sbit
        elif sbit > 0:
            nd, nb = ndim - 1, sbit - 1
        else:
            return
        if a - lo > 1:
            part(order, keys, ndim, nd, nb, lo, a - 1)
        if hi - b > 1:
            part(order, keys, ndim, nd, nb, b + 1, hi)

    u = ((pts - pts.min(1)[:, None]) / np.ptp(pts, 1)[:, None] *
         4294967295.0).astype(np.uint64).astype(object)
    want = list(range(5))
    part(want, [list(u[0]), list(u[1])], 2, 1, 31, 0, 4)
    assert order.tolist() == want


# ---- error measure ---------------------------------------------------------

def test_dfe_matches_dense_assembly():
    rng = np.random.default_rng(2)
    n, nb = 90, 30
    pts = O.zsort(rng.random((2, n)))[0]
    kd = O.kernel("exp", 2, beta=0.2)
    starts, sizes = [0, 30, 60], [30, 30, 30]
    tiles = [(i, j) for i in range(3) for j in range(i + 1)]
    res = O.compress(kd, pts, starts, sizes, tiles, 12, 4, 1e-4)
    far = [t for t, r in zip(tiles, res["rank"]) if r >= 0]
    near = [t for t, r in zip(tiles, res["rank"]) if r < 0]
    Us = [u for u, r in zip(res["U"], res["rank"]) if r >= 0]
    Vs = [v for v, r in zip(res["V"], res["rank"]) if r >= 0]
    ranks = [r for r in res["rank"] if r >= 0]
    g, _, _ = O.dfe(kd, pts, starts, sizes, "S", far, ranks, Us, Vs, near)
    A = O.kernel_block(kd, pts, pts, np.arange(n), np.arange(n))
    Ah = A.copy()
    for (i, j), u, v in zip(far, Us, Vs):
        blk = u @ v.T
        Ah[starts[i]:starts[i] + nb, starts[j]:starts[j] + nb] = blk
        if i != j:
            Ah[starts[j]:starts[j] + nb, starts[i]:starts[i] + nb] = blk.T
    want = np.linalg.norm(A - Ah) / np.linalg.norm(A)
    assert abs(g - want) <= 1e-10 * want + 1e-15
