"""GPU parity of the plain FP64 HIP backend against the oracle.

Every test calls the product through its C ABI (libstarsh_hip.so) and
compares with the CPU restatement of the reference (oracle/).  Tolerances:
kernel entries and the test matrix within a few ulps (device exp/log/cos/sin
may differ from libm in the last bit); ranks equal except where the
reference's rank decision is a near tie; errors within 1e-6 relative of the
oracle's (both compute |A - U V^T|_F of the same approximation up to
rounding of order 1e-16 |A|).
"""
import ctypes as C

import numpy as np
import pytest

import starsh_hip as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _dev(arr):
    L = S.lib()
    ptr = C.c_void_p()
    assert L.starsh_hip_malloc(C.byref(ptr), C.c_size_t(arr.nbytes)) == 0
    assert L.starsh_hip_memcpy_h2d(ptr, arr.ctypes.data_as(C.c_void_p),
                                   C.c_size_t(arr.nbytes)) == 0
    return ptr


def _free(ptr):
    S.lib().starsh_hip_free(ptr)


def _say(msg):
    print(msg, flush=True)


def test_omega_matches_oracle():
    _say("case: omega 1600 x 416 and 160 x 400")
    L = S.lib()
    for ncols, k in [(1600, 416), (160, 400), (7, 3)]:
        out = np.zeros(ncols * k)
        d = C.c_void_p()
        assert L.starsh_hip_malloc(C.byref(d), C.c_size_t(out.nbytes)) == 0
        assert L.starsh_hip_omega(d, C.c_int(ncols), C.c_int(k)) == 0
        assert L.starsh_hip_memcpy_d2h(out.ctypes.data_as(C.c_void_p), d,
                                       C.c_size_t(out.nbytes)) == 0
        _free(d)
        ref = O.omega(ncols, k).T.ravel()
        rel = np.max(np.abs(out - ref) / np.maximum(np.abs(ref), 1e-300))
        assert np.max(np.abs(out - ref)) <= 1e-14 * np.max(np.abs(ref)), rel


def _gen_case(family, ndim, tol=8e-16, **kw):
    rng = np.random.default_rng(7)
    pts = rng.random((ndim, 700))
    pts[:, 5] = pts[:, 300]      # a coincident pair: r == 0 off the diagonal
    kd = O.kernel(family, ndim, **kw)
    L = S.lib()
    d_pts = _dev(pts)
    r0 = np.array([0, 300, 650], dtype=np.int64)
    c0 = np.array([0, 290, 0], dtype=np.int64)
    m = np.array([300, 64, 50], dtype=np.int32)
    n = np.array([300, 100, 37], dtype=np.int32)
    off = np.array([0, 90000, 90000 + 6400], dtype=np.int64)
    out = np.zeros(int(off[-1] + 50 * 37))
    d_out = C.c_void_p()
    assert L.starsh_hip_malloc(C.byref(d_out), C.c_size_t(out.nbytes)) == 0
    kdesc = S.KernelDesc(kd.family, ndim, kd.beta, kd.nu, kd.noise, kd.sigma,
                         kd.wave_k, kd.diag)
    info = L.starsh_hip_generate_tiles(
        C.byref(kdesc), d_pts, C.c_int64(700), d_pts, C.c_int64(700),
        C.c_int(3), r0.ctypes.data_as(C.c_void_p),
        c0.ctypes.data_as(C.c_void_p), m.ctypes.data_as(C.c_void_p),
        n.ctypes.data_as(C.c_void_p), off.ctypes.data_as(C.c_void_p), d_out)
    assert info == 0
    assert L.starsh_hip_memcpy_d2h(out.ctypes.data_as(C.c_void_p), d_out,
                                   C.c_size_t(out.nbytes)) == 0
    _free(d_out)
    _free(d_pts)
    for t in range(3):
        ref = O.kernel_block(kd, pts, pts, np.arange(r0[t], r0[t] + m[t]),
                             np.arange(c0[t], c0[t] + n[t]))
        got = out[off[t]:off[t] + m[t] * n[t]].reshape(n[t], m[t]).T
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(got - ref)) <= tol * scale, (family, t)


@pytest.mark.parametrize("family,ndim,kw", [
    ("exp", 2, dict(beta=0.1, noise=0.5)),
    ("sqrexp", 2, dict(beta=0.1, sigma=2.0)),
    ("matern", 3, dict(beta=0.1, nu=1.5)),
    ("matern2", 3, dict(beta=0.2, nu=2.5)),
    ("coulomb", 3, {}),
    ("sin", 3, dict(wave_k=50.0, diag=1.0)),
    ("cos", 3, dict(wave_k=6.0, diag=2.0)),
])
def test_generated_tiles_match_oracle(family, ndim, kw):
    _say(f"case: generate {family} {ndim}d")
    _gen_case(family, ndim, **kw)


@pytest.mark.parametrize("family,ndim,kw", [
    ("matern", 3, dict(beta=0.1, nu=0.8, noise=0.3)),
    ("matern", 2, dict(beta=0.05, nu=2.2)),
    ("matern2", 3, dict(beta=0.2, nu=3.7, sigma=1.7)),
])
def test_generated_matern_general_nu_match_oracle(family, ndim, kw):
    """Matern with nu outside {1/2, 3/2, 5/2}: the device evaluates the
    reference's sigma 2^(1-nu)/Gamma(nu) x^nu K_nu(x) with its own K_nu
    (continued fraction / trapezoid), the oracle with an independent finer
    quadrature; both agree with scipy's kv to a few 1e-14, so the tiles
    agree to 1e-13 of their largest entry."""
    _say(f"case: generate {family} {ndim}d nu {kw['nu']}")
    _gen_case(family, ndim, tol=1e-13, **kw)


def _spectrum_matrix(m, n, sigmas, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, len(sigmas))))
    V, _ = np.linalg.qr(rng.standard_normal((n, len(sigmas))))
    return (U * sigmas) @ V.T


def test_dense_rsdd_matches_oracle():
    _say("case: dense rsdd on explicit tiles")
    L = S.lib()
    tiles = []
    # known spectra: geometric decay, and an exactly rank-deficient tile
    tiles.append(_spectrum_matrix(200, 200, 0.5 ** np.arange(60), 1))
    tiles.append(_spectrum_matrix(150, 170, np.r_[np.ones(5), np.zeros(0)], 2))
    tiles.append(_spectrum_matrix(300, 300, 0.8 ** np.arange(150), 3))
    tiles.append(np.zeros((64, 64)))
    # kernel tiles
    rng = np.random.default_rng(3)
    pts = rng.random((2, 800))
    kd = O.kernel("exp", 2, beta=0.1)
    tiles.append(O.kernel_block(kd, pts, pts, np.arange(0, 250),
                                np.arange(0, 250)))
    tiles.append(O.kernel_block(kd, pts, pts, np.arange(250, 500),
                                np.arange(500, 800)))
    maxrank, os_, tol = 100, 10, 1e-9
    m = np.array([t.shape[0] for t in tiles], dtype=np.int32)
    n = np.array([t.shape[1] for t in tiles], dtype=np.int32)
    off = np.cumsum([0] + [t.size for t in tiles])[:-1].astype(np.int64)
    flat = np.concatenate([t.T.ravel() for t in tiles])
    d_a = _dev(flat)
    uoff = np.cumsum([0] + [int(mm) * maxrank for mm in m])[:-1].astype(np.int64)
    voff = np.cumsum([0] + [int(nn) * maxrank for nn in n])[:-1].astype(np.int64)
    k = np.minimum(maxrank + os_, np.minimum(m, n))
    soff = np.cumsum([0] + list(k))[:-1].astype(np.int64)
    U = np.zeros(int(m.sum()) * maxrank)
    V = np.zeros(int(n.sum()) * maxrank)
    Sv = np.zeros(int(k.sum()))
    rank = np.zeros(len(tiles), dtype=np.int32)
    info = L.starsh_hip_dense_rsdd(
        C.c_int(len(tiles)), m.ctypes.data_as(C.c_void_p),
        n.ctypes.data_as(C.c_void_p), d_a, off.ctypes.data_as(C.c_void_p),
        C.c_int(maxrank), C.c_int(os_), C.c_double(tol),
        rank.ctypes.data_as(C.c_void_p), U.ctypes.data_as(C.c_void_p),
        uoff.ctypes.data_as(C.c_void_p), V.ctypes.data_as(C.c_void_p),
        voff.ctypes.data_as(C.c_void_p), Sv.ctypes.data_as(C.c_void_p),
        soff.ctypes.data_as(C.c_void_p))
    _free(d_a)
    assert info == 0
    for t, A in enumerate(tiles):
        r_ref, Ur, Vr, Sr = O.dlrrsdd(A, maxrank, os_, tol)
        r = int(rank[t])
        _say(f"  tile {t}: {A.shape} rank gpu {r} oracle {r_ref}")
        assert r == r_ref, (t, r, r_ref)
        s = Sv[soff[t]:soff[t] + k[t]]
        # singular values that matter for the rank agree to 1e-13 * s_max
        assert np.max(np.abs(s[:max(r, 1)] - Sr[:max(r, 1)])) <= \
            1e-12 * max(Sr[0], 1e-300) + 1e-300
        if r > 0:
            Ug = U[uoff[t]:uoff[t] + m[t] * r].reshape(r, m[t]).T
            Vg = V[voff[t]:voff[t] + n[t] * r].reshape(r, n[t]).T
            nA = np.linalg.norm(A)
            e_g = np.linalg.norm(A - Ug @ Vg.T) / nA
            e_r = np.linalg.norm(A - Ur @ Vr.T) / nA
            assert abs(e_g - e_r) <= 1e-6 * e_r + 1e-14, (t, e_g, e_r)


def _compress_vs_oracle(kind, N, ndim, nb, maxrank, os_, tol, symm="S",
                        **kw):
    fn = S.init("HIP", oversample=os_)
    p = S.TLRProblem(kind, N, ndim, nb, symm=symm, **kw)
    tiles0 = p.tiles()
    pts = p.points()
    st, sz = p.cluster_ranges()
    fam = kw.get("kernel", "sqrexp") if kind == "spatial" else kind
    kd = O.kernel(fam, ndim, **{a: kw[a] for a in
                                ("beta", "nu", "noise", "sigma", "wave_k",
                                 "diag") if a in kw})
    ref = O.compress(kd, pts, st, sz, tiles0, maxrank, os_, tol,
                     want_factors=True, want_S=True)
    info = p.approximate(maxrank, tol, 0, fn)
    assert info == 0
    nfalse_ref = int(np.sum(ref["rank"] < 0))
    # the format was rewritten: far tiles keep the original relative order
    far_ref = [t for t, r in zip(tiles0, ref["rank"]) if r >= 0]
    assert p.tiles() == far_ref, "false far-field sets differ"
    assert len(p.near_tiles()) == nfalse_ref
    rg = p.ranks()
    rr = np.array([r for r in ref["rank"] if r >= 0])
    mism = np.nonzero(rg != rr)[0]
    for b in mism:
        # allowed only at a near tie of the reference's rule
        i = [q for q, t in enumerate(tiles0) if t == far_ref[b]][0]
        s = ref["S"][i]
        tot = np.sum(s * s)
        tail = np.sum(s[rr[b]:] ** 2)
        assert abs(tail - tol * tol * tot) <= 1e-6 * tol * tol * tot, \
            (far_ref[b], rg[b], rr[b])
    assert len(mism) <= max(1, len(rr) // 100)
    err = p.dfe()
    Us = [p.factors(b)[0] for b in range(len(rg))]
    Vs = [p.factors(b)[1] for b in range(len(rg))]
    g_mine, _, _ = O.dfe(kd, pts, st, sz, symm, p.tiles(), rg, Us, Vs,
                         p.near_tiles())
    g_ref, _, _ = O.dfe(kd, pts, st, sz, symm, far_ref, rr, ref["U"] and
                        [u for u, r in zip(ref["U"], ref["rank"]) if r >= 0],
                        [v for v, r in zip(ref["V"], ref["rank"]) if r >= 0],
                        p.near_tiles())
    _say(f"  dfe gpu {err:.6e} oracle-of-gpu {g_mine:.6e} "
         f"reference {g_ref:.6e} ranks {rg.min()}..{rg.max()} "
         f"false-far {nfalse_ref}")
    assert abs(err - g_mine) <= 1e-6 * g_mine
    assert abs(g_mine - g_ref) <= 0.05 * g_ref
    assert err / tol <= 10
    if near := p.near_tiles():
        for b, (i, j) in enumerate(near):
            D = p.near_block(b)
            R = O.kernel_block(kd, pts, pts, np.arange(st[i], st[i] + sz[i]),
                               np.arange(st[j], st[j] + sz[j]))
            assert np.max(np.abs(D - R)) <= 8e-16 * np.max(np.abs(R))
    p.free_matrix()


def test_compress_c1_exp2d_matches_oracle():
    _say("case: compress 2D exp N=2500 nb=250 maxrank=100 tol=1e-9")
    _compress_vs_oracle("spatial", 2500, 2, 250, 100, 10, 1e-9, seed=0,
                        kernel="exp", beta=0.1)


def test_compress_sqexp2d_ragged_nonsymmetric():
    _say("case: compress 2D sqexp N=3000 nb=700 symm N (ragged)")
    _compress_vs_oracle("spatial", 3000, 2, 700, 150, 16, 1e-7, symm="N",
                        seed=1, kernel="sqrexp", beta=0.1)


@pytest.mark.parametrize("backend", ["HIP", "HIP_MIXED"])
def test_compress_matern_general_nu(backend):
    """3D Matern with nu = 0.8 (no closed form), N=4000 uniform points,
    tiles 500, maxrank 200, tol 1e-7: the reference's acceptance bound and,
    for the plain form, ranks equal to the oracle's on sampled tiles (up to
    near-ties of the rank rule)."""
    _say(f"case: {backend} compress 3D matern nu 0.8 N=4000 nb=500")
    fn = S.init(backend, oversample=10)
    p = S.TLRProblem("spatial", 4000, 3, 500, symm="S", seed=5,
                     place=S.STARSH_PARTICLES_RAND, kernel="matern",
                     beta=0.1, nu=0.8)
    assert p.approximate(200, 1e-7, 0, fn) == 0
    err = p.dfe()
    _say(f"  error {err:.3e}, ranks {p.ranks().tolist()[:12]}")
    assert 0 < err <= 1e-6
    if backend == "HIP":
        tiles, ranks = p.tiles(), p.ranks()
        st, sz = p.cluster_ranges()
        kd = O.kernel("matern", 3, beta=0.1, nu=0.8)
        pick = [0, 7, 20, len(tiles) - 1]
        ref = O.compress(kd, p.points(), st, sz, [tiles[b] for b in pick],
                         200, 10, 1e-7, want_factors=False, want_S=True)
        for q, b in enumerate(pick):
            rr, rg = int(ref["rank"][q]), int(ranks[b])
            if rr != rg:
                s = ref["S"][q]
                tot = np.sum(s * s)
                assert abs(np.sum(s[rr:] ** 2) - 1e-14 * tot) <= \
                    1e-6 * 1e-14 * tot, (tiles[b], rg, rr)
    p.free_matrix()
