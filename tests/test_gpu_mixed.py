"""GPU checks of the batched mixed-precision form (STARSH_BACKEND=HIP_MIXED).

The mixed form does not promise the reference's exact ranks (its sketch
width adapts per tile); it promises the requested accuracy.  Tests check the
int8 sketch against an exact product, per-tile and global relative
Frobenius errors against FP64 regenerated tiles (starsh_blrm__dfe_hip and
the oracle), and ranks against the plain FP64 form.
"""
import ctypes as C
import os

import numpy as np
import pytest

import starsh_hip as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_int8_sketch_matches_exact_product(tmp_path):
    """Y from the int8 digit planes equals A @ round(32*Omega) / 32 up to
    the representation error of each row of A: entries are rounded to
    multiples of 2^(e-30) with 2^e <= 2 * (row maximum)."""
    _say("case: int8 sketch of 2D sqexp N=3200 nb=800")
    os.environ["STARSH_HIP_DEBUG_DUMP"] = str(tmp_path)
    L = S.lib()
    # one bucket only (the widest), so the dump holds tile (0, 0)
    assert L.starsh_hip_set_option(b"mixed_first_bucket", C.c_double(4)) == 0
    try:
        fn = S.init("HIP_MIXED", oversample=16)
        p = S.TLRProblem("spatial", 3200, 2, 800, symm="S", seed=4,
                         kernel="sqrexp", beta=0.1)
        assert p.approximate(200, 1e-7, 0, fn) == 0
    finally:
        del os.environ["STARSH_HIP_DEBUG_DUMP"]
        L.starsh_hip_set_option(b"mixed_first_bucket", C.c_double(-1))
    Y = np.fromfile(tmp_path / "Y.bin")
    m = 800
    k = Y.size // m
    Y = Y.reshape(k, m).T
    pts = p.points()
    st, sz = p.cluster_ranges()
    i, j = 0, 0
    A = O.kernel_block(O.kernel("sqrexp", 2, beta=0.1), pts, pts,
                       np.arange(st[i], st[i] + sz[i]),
                       np.arange(st[j], st[j] + sz[j]))
    om = np.clip(np.rint(32.0 * O.omega(800, k)), -127, 127)
    want = A @ om / 32.0
    rowmax = np.max(np.abs(A), axis=1)
    bound = 2.0 ** -30 * rowmax * np.sum(np.abs(om), axis=0).max() / 32.0
    assert np.all(np.max(np.abs(Y - want), axis=1) <= bound + 1e-300)
    p.free_matrix()


def test_int8_projection_matches_exact_product(tmp_path):
    """Bt from the int8 digit planes of A^T and Q equals A^T Q up to the
    rounding of both operands: A per (column, chunk of 64 rows) and Q per
    column to multiples of 2^-30 of a power of two at most twice the
    largest entry, so each term is off by at most 2^-29 of the product of
    those maxima."""
    _say("case: int8 projection of 2D sqexp N=3200 nb=800")
    os.environ["STARSH_HIP_DEBUG_DUMP"] = str(tmp_path)
    L = S.lib()
    assert L.starsh_hip_set_option(b"mixed_first_bucket", C.c_double(4)) == 0
    try:
        fn = S.init("HIP_MIXED", oversample=16)
        p = S.TLRProblem("spatial", 3200, 2, 800, symm="S", seed=4,
                         kernel="sqrexp", beta=0.1)
        assert p.approximate(200, 1e-7, 0, fn) == 0
    finally:
        del os.environ["STARSH_HIP_DEBUG_DUMP"]
        L.starsh_hip_set_option(b"mixed_first_bucket", C.c_double(-1))
    m = n = 800
    Q = np.fromfile(tmp_path / "Q.bin")
    k = Q.size // m
    Q = Q.reshape(k, m).T
    Bt = np.fromfile(tmp_path / "Bt.bin").reshape(k, n).T
    pts = p.points()
    st, sz = p.cluster_ranges()
    A = O.kernel_block(O.kernel("sqrexp", 2, beta=0.1), pts, pts,
                       np.arange(st[0], st[0] + sz[0]),
                       np.arange(st[0], st[0] + sz[0]))
    # the first columns of Q are kept ones (the tile's rank is above 4)
    c = 4
    q = Q[:, :c]
    np.testing.assert_allclose(q.T @ q, np.eye(c), atol=1e-12)
    want = A.T @ q
    amax = np.array([np.abs(A[i:i + 64]).max(axis=0)
                     for i in range(0, m, 64)])               # chunk x j
    qmax = np.abs(q).max(axis=0)
    # sum over chunks of 64 * 2^-29 * amax * qmax, per (j, c)
    bound = 2.0 ** -29 * 64 * np.outer(amax.sum(axis=0), qmax)
    err = np.abs(Bt[:, :c] - want)
    assert np.all(err <= bound)
    assert err.max() > 0       # the int8 path ran (FP64 would be ~1e-16)
    p.free_matrix()


def _mixed_vs_plain(kind, N, ndim, nb, maxrank, os_, tol, **kw):
    res = {}
    for backend in ("HIP", "HIP_MIXED"):
        fn = S.init(backend, oversample=os_)
        p = S.TLRProblem(kind, N, ndim, nb, symm="S", **kw)
        assert p.approximate(maxrank, tol, 0, fn) == 0
        g, terr, tnorm = p.dfe_tiles()
        res[backend] = (p.ranks(), g, terr, tnorm, p.tiles(), p)
    return res


def test_mixed_sqexp2d_accuracy_and_ranks():
    _say("case: mixed vs plain, 2D sqexp N=12800 nb=1600 tol 1e-7")
    res = _mixed_vs_plain("spatial", 12800, 2, 1600, 400, 16, 1e-7, seed=1,
                          kernel="sqrexp", beta=0.1)
    rp, gp, _, _, tp, pp = res["HIP"]
    rm, gm, terr, tnorm, tm, pm = res["HIP_MIXED"]
    _say(f"  plain err {gp:.3e} mixed err {gm:.3e}; ranks plain "
         f"{rp.min()}..{rp.max()} mixed {rm.min()}..{rm.max()}")
    assert tp == tm
    assert gm <= 1e-7
    big = tnorm > 1e-12 * tnorm.max()
    assert np.all(terr[big] <= 2e-7)
    assert np.max(np.abs(rm - rp)) <= 3
    # oracle check of the mixed factors' error
    pts = pm.points()
    st, sz = pm.cluster_ranges()
    Us = [pm.factors(b)[0] for b in range(len(rm))]
    Vs = [pm.factors(b)[1] for b in range(len(rm))]
    go, _, _ = O.dfe(O.kernel("sqrexp", 2, beta=0.1), pts, st, sz, "S", tm,
                     rm, Us, Vs, pm.near_tiles())
    assert abs(go - gm) <= 1e-6 * go
    pp.free_matrix()
    pm.free_matrix()


def test_mixed_exp2d_tight_tolerance_uses_fp64_sketch():
    _say("case: mixed vs plain, 2D exp N=2500 nb=250 tol 1e-9")
    res = _mixed_vs_plain("spatial", 2500, 2, 250, 100, 10, 1e-9, seed=0,
                          kernel="exp", beta=0.1)
    rp, gp, _, _, tp, pp = res["HIP"]
    rm, gm, _, _, tm, pm = res["HIP_MIXED"]
    _say(f"  plain err {gp:.3e} mixed err {gm:.3e}")
    assert gm <= max(2 * gp, 1e-9)
    assert pm.near_tiles() == pp.near_tiles()
    pp.free_matrix()
    pm.free_matrix()


@pytest.mark.parametrize("kw", [
    dict(kernel="exp", beta=0.1, sigma=2.0, noise=0.05),
    dict(kernel="matern", beta=0.05, nu=0.5, sigma=1.5),
    dict(kernel="matern2", beta=0.1, nu=2.5),
])
def test_mixed_sigma_noise_and_matern_variants(kw):
    """The int8 paths fold sigma into the exp table, scale the points per
    family and select sigma + noise at distance 0 only for batches with a
    noise: the mixed form keeps the accuracy and the false far-field tiles
    (the diagonal tiles with a nugget) of the plain form."""
    _say(f"case: mixed vs plain, 2D {kw} N=6400 nb=800 tol 1e-7")
    res = _mixed_vs_plain("spatial", 6400, 2, 800, 200, 16, 1e-7, seed=2,
                          **kw)
    rp, gp, _, _, tp, pp = res["HIP"]
    rm, gm, _, _, tm, pm = res["HIP_MIXED"]
    _say(f"  plain err {gp:.3e} mixed err {gm:.3e}; false far plain "
         f"{len(pp.near_tiles())} mixed {len(pm.near_tiles())}")
    assert tp == tm
    assert pm.near_tiles() == pp.near_tiles()
    assert gm <= 1e-7
    assert abs(gm - gp) <= 0.05 * gp
    assert np.max(np.abs(rm - rp)) <= 3
    pp.free_matrix()
    pm.free_matrix()


def test_mixed_ladder1_3d_matern(tmp_path):
    """Bucket ladder 1 (a 384-wide bucket; an unfinished tile moves to the
    width its computed rank needs when its sketch was not saturated) on 3D
    Matern nu=1.5 from first width 256: tiles that need more than 256
    columns finish at width 384, with the accuracy of the plain form and
    its ranks to within 3."""
    _say("case: mixed ladder 1, 3D Matern nu=1.5 N=32768 nb=2048 tol 1e-7")
    L = S.lib()
    trace = tmp_path / "rounds.txt"
    kw = dict(seed=2, place=S.STARSH_PARTICLES_RAND, kernel="matern",
              beta=0.1, nu=1.5)
    fn = S.init("HIP", oversample=16)
    pp = S.TLRProblem("spatial", 32768, 3, 2048, symm="S", **kw)
    assert pp.approximate(512, 1e-7, 0, fn) == 0
    rp = pp.ranks()
    gp, _, _ = pp.dfe_tiles()
    assert L.starsh_hip_set_option(b"mixed_first_bucket", C.c_double(3)) == 0
    assert L.starsh_hip_set_option(b"mixed_ladder", C.c_double(1)) == 0
    os.environ["STARSH_HIP_MIXED_TRACE"] = str(trace)
    try:
        fn = S.init("HIP_MIXED", oversample=16)
        pm = S.TLRProblem("spatial", 32768, 3, 2048, symm="S", **kw)
        assert pm.approximate(512, 1e-7, 0, fn) == 0
    finally:
        del os.environ["STARSH_HIP_MIXED_TRACE"]
        L.starsh_hip_set_option(b"mixed_first_bucket", C.c_double(-1))
        L.starsh_hip_set_option(b"mixed_ladder", C.c_double(-1))
    rm = pm.ranks()
    gm, terr, tnorm = pm.dfe_tiles()
    # round, problem, tile, width, rank, accepted, ...
    rows = [line.split() for line in trace.read_text().splitlines()]
    w384 = [r for r in rows if int(r[3]) == 384]
    _say(f"  plain err {gp:.3e} mixed err {gm:.3e}; {len(rows)} tile "
         f"rounds, {len(w384)} at width 384; ranks plain "
         f"{rp.min()}..{rp.max()} mixed {rm.min()}..{rm.max()}")
    assert len(w384) > 0 and all(int(r[5]) == 1 for r in w384)
    assert all(int(r[0]) == 0 for r in rows if int(r[3]) == 256)
    assert pm.tiles() == pp.tiles()
    assert gm <= 1e-7
    big = tnorm > 1e-12 * tnorm.max()
    assert np.all(terr[big] <= 2e-7)
    assert np.max(np.abs(rm - rp)) <= 3
    pp.free_matrix()
    pm.free_matrix()
