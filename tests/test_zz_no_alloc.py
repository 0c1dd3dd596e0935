"""The allocation-free product starsh_hip_tlr_dmml_noalloc against
starsh_hip_tlr_dmml, through the C ABI.

Both run the same kernels in the same order on the same operator; the
only difference is where the coefficient blocks live (a static device
array instead of the operator's workspace), so the results must be
bit-identical.  Two calls in a row, the second on a smaller operator with
fewer right-hand sides (the array keeps the first call's contents), then a
call whose coefficients exceed the array, which must be refused before any
GPU work.
"""
import ctypes as C

import numpy as np
import pytest

import starsh_hip as S

pytestmark = pytest.mark.gpu

dp = C.POINTER(C.c_double)


def _say(msg):
    print(msg, flush=True)


class _Dev:
    """Device buffer through the library's own helpers."""

    def __init__(self, L, host):
        self.L = L
        self.n = host.size
        self.p = C.c_void_p()
        assert L.starsh_hip_malloc(C.byref(self.p),
                                   C.c_size_t(8 * max(self.n, 1))) == 0
        assert L.starsh_hip_memcpy_h2d(self.p, host.ctypes.data_as(dp),
                                       C.c_size_t(8 * self.n)) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.L.starsh_hip_memcpy_d2h(out.ctypes.data_as(dp), self.p,
                                            C.c_size_t(8 * self.n)) == 0
        return out

    def free(self):
        self.L.starsh_hip_free(self.p)


def _operator(L, N, nb, seed, kernel, symm):
    fn = S.init("HIP", oversample=10)
    p = S.TLRProblem("spatial", N, 2, nb, symm=symm, seed=seed,
                     kernel=kernel, beta=0.1)
    assert p.approximate(100, 1e-9, 0, fn) == 0
    T = C.c_void_p()
    assert L.starsh_hip_tlr_new(C.byref(T), C.cast(p.M, C.c_void_p)) == 0
    return p, T


def _both(L, T, N, nrhs, alpha, beta, rng):
    X = np.asfortranarray(rng.standard_normal((N, nrhs)))
    Y0 = np.asfortranarray(rng.standard_normal((N, nrhs)))
    dx = _Dev(L, X.ravel(order="F"))
    dy1 = _Dev(L, Y0.ravel(order="F"))
    dy2 = _Dev(L, Y0.ravel(order="F"))
    try:
        assert L.starsh_hip_tlr_dmml(T, nrhs, C.c_double(alpha), dx.p, N,
                                     C.c_double(beta), dy1.p, N) == 0
        assert L.starsh_hip_tlr_dmml_noalloc(T, nrhs, C.c_double(alpha),
                                             dx.p, N, C.c_double(beta),
                                             dy2.p, N) == 0
        assert L.starsh_hip_synchronize() == 0
        return dy1.get(), dy2.get()
    finally:
        dx.free()
        dy1.free()
        dy2.free()


def test_noalloc_product_matches_workspace_product():
    L = S.lib()
    rng = np.random.default_rng(5)
    # 1: 2D exp N=2500, tiles 250, symmetric, false far-field tiles
    # (near D), 3 right-hand sides
    _say("case: noalloc product, 2D exp N=2500 nb=250 'S', nrhs 3")
    p1, T1 = _operator(L, 2500, 250, 0, "exp", "S")
    try:
        a, b = _both(L, T1, 2500, 3, 0.7, -0.4, rng)
        assert np.array_equal(a, b)
        assert np.abs(a).max() > 0
        # coefficients beyond the static array (2^22 doubles): 2000
        # right-hand sides need 2 x (sum of ranks, about 1800 here) x 2000;
        # refused before any GPU work
        ranks = p1.ranks()
        assert 2 * int(ranks.sum()) * 2000 > (1 << 22)
        X = np.zeros(2500 * 2000)
        dx, dy = _Dev(L, X), _Dev(L, X)
        try:
            assert L.starsh_hip_tlr_dmml_noalloc(
                T1, 2000, C.c_double(1.0), dx.p, 2500, C.c_double(0.0),
                dy.p, 2500) == S.STARSH_WRONG_PARAMETER
        finally:
            dx.free()
            dy.free()
    finally:
        L.starsh_hip_tlr_free(T1)
        p1.free_matrix()
    # 2: smaller operator, one right-hand side, nonsymmetric
    _say("case: noalloc product, 2D sqexp N=1200 nb=300 'N', nrhs 1")
    p2, T2 = _operator(L, 1200, 300, 3, "sqrexp", "N")
    try:
        a, b = _both(L, T2, 1200, 1, 1.0, 0.0, rng)
        assert np.array_equal(a, b)
        assert np.abs(a).max() > 0
    finally:
        L.starsh_hip_tlr_free(T2)
        p2.free_matrix()
