"""ORACLE -- test infrastructure only.

ctypes wrapper of liboracle.so, the CPU restatement of the reference
algorithm (see oracle.c for the reference file:line each function follows).
Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
use it; the product library never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

FAMILY = {"exp": 1, "sqrexp": 2, "matern": 3, "matern2": 4, "coulomb": 5,
          "sin": 6, "cos": 7}


class Kernel(C.Structure):
    _fields_ = [("family", C.c_int), ("ndim", C.c_int), ("beta", C.c_double),
                ("nu", C.c_double), ("noise", C.c_double),
                ("sigma", C.c_double), ("wave_k", C.c_double),
                ("diag", C.c_double)]


def kernel(family, ndim, beta=0.1, nu=0.5, noise=0.0, sigma=1.0, wave_k=1.0,
           diag=1.0):
    return Kernel(FAMILY[family] if isinstance(family, str) else family,
                  ndim, beta, nu, noise, sigma, wave_k, diag)


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_int64)


def _d(a):
    return a.ctypes.data_as(_dp)


def _l(a):
    return a.ctypes.data_as(_lp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", HERE])
        L = C.CDLL(LIB)
        L.oracle_dsvfr.argtypes = [C.c_int, _dp, C.c_double]
        L.oracle_dfe.restype = C.c_double
        L.oracle_lapack_name.restype = C.c_char_p
        _lib = L
    return _lib


def lapack_load():
    return lib().oracle_lapack_load(None) == 0


def omega(ncols, k):
    out = np.empty(ncols * k)
    lib().oracle_omega(C.c_int(ncols), C.c_int(k), _d(out))
    return out.reshape(k, ncols).T          # column-major ncols x k


def dsvfr(S, tol):
    S = np.ascontiguousarray(S, dtype=np.float64)
    return lib().oracle_dsvfr(len(S), _d(S), tol)


def kernel_block(kd, xr, xc, irow, icol):
    """Tile A[i, j] = f(xr[:, irow[i]], xc[:, icol[j]]) (xr: ndim x nr)."""
    xr = np.ascontiguousarray(xr, dtype=np.float64)
    xc = np.ascontiguousarray(xc, dtype=np.float64)
    irow = np.ascontiguousarray(irow, dtype=np.int64)
    icol = np.ascontiguousarray(icol, dtype=np.int64)
    m, n = len(irow), len(icol)
    out = np.empty(m * n)
    lib().oracle_kernel_block(C.byref(kd), _d(xr), C.c_int64(xr.shape[1]),
                              _d(xc), C.c_int64(xc.shape[1]), C.c_int(m),
                              C.c_int(n), _l(irow), _l(icol), _d(out),
                              C.c_int(m))
    return out.reshape(n, m).T


def dlrrsdd(D, maxrank, oversample, tol):
    """Returns (rank, U, V, S) of the reference dense RSDD of D."""
    assert lapack_load(), "no host LAPACK found"
    D = np.asfortranarray(D, dtype=np.float64)
    m, n = D.shape
    k = min(maxrank + oversample, m, n)
    U = np.zeros((m, maxrank), order="F")
    V = np.zeros((n, maxrank), order="F")
    S = np.zeros(k)
    r = lib().oracle_dlrrsdd(C.c_int(m), C.c_int(n), _d(D), C.c_int(m),
                             _d(U), C.c_int(m), _d(V), C.c_int(n),
                             C.c_int(maxrank), C.c_int(oversample),
                             C.c_double(tol), _d(S))
    rr = max(r, 0)
    return r, U[:, :rr].copy(), V[:, :rr].copy(), S


def dlrsdd(D, maxrank, tol):
    """Returns (rank, U, V, S) of the reference dense SVD engine
    (sequential/dense/dsdd.c:16-66) on D."""
    assert lapack_load(), "no host LAPACK found"
    D = np.array(D, dtype=np.float64, order="F")
    m, n = D.shape
    mn = min(m, n)
    U = np.zeros((m, maxrank), order="F")
    V = np.zeros((n, maxrank), order="F")
    S = np.zeros(mn)
    r = lib().oracle_dlrsdd(C.c_int(m), C.c_int(n), _d(D), C.c_int(m),
                            _d(U), C.c_int(m), _d(V), C.c_int(n),
                            C.c_int(maxrank), C.c_double(tol), _d(S))
    rr = max(r, 0)
    return r, U[:, :rr].copy(), V[:, :rr].copy(), S


def dlrqp3(D, maxrank, oversample, tol):
    """Returns (rank, U, V, S, pivots) of the reference dense RRQR engine
    (sequential/dense/dqp3.c:16-93) on D; pivots are 0-based."""
    assert lapack_load(), "no host LAPACK found"
    D = np.array(D, dtype=np.float64, order="F")
    m, n = D.shape
    k = min(maxrank + oversample, m, n)
    U = np.zeros((m, maxrank), order="F")
    V = np.zeros((n, maxrank), order="F")
    S = np.zeros(k)
    piv = np.zeros(n, dtype=np.int32)
    r = lib().oracle_dlrqp3(C.c_int(m), C.c_int(n), _d(D), C.c_int(m),
                            _d(U), C.c_int(m), _d(V), C.c_int(n),
                            C.c_int(maxrank), C.c_int(oversample),
                            C.c_double(tol), _d(S), piv.ctypes.data_as(_ip))
    rr = max(r, 0)
    return r, U[:, :rr].copy(), V[:, :rr].copy(), S, piv.astype(np.int64)


ENGINES = {"rsvd": 0, "sdd": 1, "rrqr": 2}


def compress(kd, pts, starts, sizes, tiles, maxrank, oversample, tol,
             nthreads=0, want_factors=True, want_S=False, engine="rsvd"):
    """Reference tile loop on plain tiling (engine "rsvd": openmp/blrm/
    drsdd.c, "sdd": dsdd.c, "rrqr": dqp3.c).  Returns dict with ranks and
    optionally U, V (lists) and S (ntiles x k)."""
    assert lapack_load(), "no host LAPACK found"
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    ndata = pts.shape[1]
    starts = np.ascontiguousarray(starts, dtype=np.int64)
    sizes = np.ascontiguousarray(sizes, dtype=np.int64)
    brow = np.array([t[0] for t in tiles], dtype=np.int64)
    bcol = np.array([t[1] for t in tiles], dtype=np.int64)
    nt = len(tiles)
    rank = np.zeros(nt, dtype=np.int32)
    uoff = np.zeros(nt, dtype=np.int64)
    voff = np.zeros(nt, dtype=np.int64)
    su = sv = 0
    for b in range(nt):
        uoff[b] = su
        voff[b] = sv
        su += sizes[brow[b]] * maxrank
        sv += sizes[bcol[b]] * maxrank
    U = np.zeros(max(su, 1)) if want_factors else None
    V = np.zeros(max(sv, 1)) if want_factors else None
    kmax = int(min(maxrank + oversample, sizes.max()))
    if engine == "sdd":
        kmax = int(sizes.max())
    S = np.zeros(nt * kmax) if want_S else None
    info = lib().oracle_compress_engine(
        C.c_int(ENGINES[engine]),
        C.byref(kd), _d(pts), C.c_int64(ndata), _d(pts), C.c_int64(ndata),
        C.c_int64(nt), _l(brow), _l(bcol), _l(starts), _l(sizes), _l(starts),
        _l(sizes), C.c_int(maxrank), C.c_int(oversample), C.c_double(tol),
        C.c_int(nthreads), rank.ctypes.data_as(_ip),
        _d(U) if want_factors else None, _l(uoff),
        _d(V) if want_factors else None, _l(voff),
        _d(S) if want_S else None, C.c_int(kmax))
    assert info == 0
    out = {"rank": rank.astype(np.int64)}
    if want_factors:
        Us, Vs = [], []
        for b in range(nt):
            m, n, r = sizes[brow[b]], sizes[bcol[b]], max(rank[b], 0)
            Us.append(U[uoff[b]:uoff[b] + m * r].reshape(r, m).T.copy())
            Vs.append(V[voff[b]:voff[b] + n * r].reshape(r, n).T.copy())
        out["U"], out["V"] = Us, Vs
    if want_S:
        out["S"] = S.reshape(nt, kmax)
    return out


def dfe(kd, pts, starts, sizes, symm, far_tiles, ranks, Us, Vs, near_tiles,
        nthreads=0):
    """Global relative error (openmp/blrm/dfe.c); also per-tile norms."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    ndata = pts.shape[1]
    starts = np.ascontiguousarray(starts, dtype=np.int64)
    sizes = np.ascontiguousarray(sizes, dtype=np.int64)
    nf, nn = len(far_tiles), len(near_tiles)
    fbrow = np.array([t[0] for t in far_tiles] + [0], dtype=np.int64)
    fbcol = np.array([t[1] for t in far_tiles] + [0], dtype=np.int64)
    nbrow = np.array([t[0] for t in near_tiles] + [0], dtype=np.int64)
    nbcol = np.array([t[1] for t in near_tiles] + [0], dtype=np.int64)
    rank = np.array(list(ranks) + [0], dtype=np.int32)
    uoff = np.zeros(nf + 1, dtype=np.int64)
    voff = np.zeros(nf + 1, dtype=np.int64)
    su = sv = 0
    for b in range(nf):
        uoff[b], voff[b] = su, sv
        su += Us[b].size
        sv += Vs[b].size
    U = np.concatenate([u.T.ravel() for u in Us] + [np.zeros(1)])
    V = np.concatenate([v.T.ravel() for v in Vs] + [np.zeros(1)])
    fn = np.zeros(nf + 1)
    fd = np.zeros(nf + 1)
    g = lib().oracle_dfe(C.byref(kd), _d(pts), C.c_int64(ndata), _d(pts),
                         C.c_int64(ndata), C.c_char(symm.encode()),
                         C.c_int64(nf), _l(fbrow), _l(fbcol),
                         rank.ctypes.data_as(_ip), _d(U), _l(uoff), _d(V),
                         _l(voff), C.c_int64(nn), _l(nbrow), _l(nbcol),
                         _l(starts), _l(sizes), _l(starts), _l(sizes),
                         C.c_int(nthreads), _d(fn), _d(fd))
    return g, fn[:nf], fd[:nf]


def gen_quasiuniform1(count, ndim, seed):
    out = np.empty(count * ndim)
    assert lib().oracle_gen_quasiuniform1(C.c_int64(count), C.c_int(ndim),
                                          C.c_uint(seed), _d(out)) == 0
    return out.reshape(ndim, count)


def gen_rand(count, ndim, seed):
    out = np.empty(count * ndim)
    lib().oracle_gen_rand(C.c_int64(count), C.c_int(ndim), C.c_uint(seed),
                          _d(out))
    return out.reshape(ndim, count)


def zsort(points):
    pts = np.ascontiguousarray(points, dtype=np.float64).copy()
    ndim, count = pts.shape
    order = np.empty(count, dtype=np.int64)
    lib().oracle_zsort(C.c_int64(count), C.c_int(ndim), _d(pts), _l(order))
    return pts, order


def dmml(starts, sizes, symm, far_tiles, Us, Vs, near_tiles, Ds, A, alpha,
         beta, B, col_starts=None):
    """B = alpha * M * A + beta * B for a compressed matrix M, restating
    starsh_blrm__dmml (src/backends/sequential/blrm/dmml.c:16-145): B is
    scaled first (set to zero when beta == 0, :52-59); far tile (i, j) adds
    alpha U (V^T A_j) to B_i and, when symm == 'S' and i != j, alpha V (U^T
    A_i) to B_j (:61-91); near tile (i, j) adds alpha D A_j and, symmetric
    with i != j, alpha D^T A_i (:92-145).  starts/sizes describe the row
    blocks, col_starts (default: the same) the column blocks; A and B are
    (rows x nrhs) arrays in cluster order.  Returns the new B."""
    cs = starts if col_starts is None else col_starts
    B = np.zeros_like(B) if beta == 0.0 else beta * B
    A = np.asarray(A, dtype=np.float64)
    for (i, j), U, V in zip(far_tiles, Us, Vs):
        ri = slice(starts[i], starts[i] + U.shape[0])
        cj = slice(cs[j], cs[j] + V.shape[0])
        if U.shape[1] > 0:
            B[ri] += alpha * (U @ (V.T @ A[cj]))
            if symm == "S" and i != j:
                B[cj] += alpha * (V @ (U.T @ A[ri]))
    for (i, j), D in zip(near_tiles, Ds):
        ri = slice(starts[i], starts[i] + D.shape[0])
        cj = slice(cs[j], cs[j] + D.shape[1])
        B[ri] += alpha * (D @ A[cj])
        if symm == "S" and i != j:
            B[cj] += alpha * (D.T @ A[ri])
    return B


def cg(matvec, B, X, tol):
    """Conjugate gradients restating starsh_itersolvers__dcg_omp
    (src/itersolvers/cg.c:24-95): R = B - M X, P = R; column q stops once
    |r_q| < tol |r_q(0)|; returns (iteration at which every column stopped,
    or -1 after n iterations, X).  matvec(P) returns M P."""
    X = np.array(X, dtype=np.float64)
    n, nrhs = B.shape
    R = B - matvec(X)
    P = R.copy()
    rsold = np.linalg.norm(R, axis=0)
    rscheck = rsold * tol
    rsold = rsold ** 2
    finished = 0
    for it in range(n):
        Q = matvec(P)
        for q in range(nrhs):
            if rscheck[q] < 0:
                continue
            a = rsold[q] / (P[:, q] @ Q[:, q])
            X[:, q] += a * P[:, q]
            R[:, q] -= a * Q[:, q]
            rsnew = np.linalg.norm(R[:, q])
            if rsnew < rscheck[q]:
                finished += 1
                rscheck[q] = -1.0
                continue
            rsnew *= rsnew
            P[:, q] = R[:, q] + (rsnew / rsold[q]) * P[:, q]
            rsold[q] = rsnew
        if finished == nrhs:
            return it, X
    return -1, X
