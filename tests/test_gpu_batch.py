"""Batch API (starsh_blrm_approximate_batch) on the GPU.

Several independent problems go through one call; each problem's result
must be what a call of its own gives.  The acceptance bound on the error is
the reference tests' rel_err / tol <= 10 (testing/spatial.c:105-109), and
per-tile errors are checked against the exact tiles by starsh_blrm__dfe_hip.
"""
import numpy as np
import pytest

import starsh_hip as S

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _problems(specs):
    out = []
    for (kind, kernel, N, ndim, nb, beta, seed) in specs:
        out.append(S.TLRProblem(kind, N, ndim, nb, symm="S", seed=seed,
                                kernel=kernel, beta=beta,
                                nu=1.5 if kernel == "matern2" else 0.5))
    return out


def _free(ps):
    for p in ps:
        p.free_matrix()


@pytest.mark.parametrize("backend,engine", [
    ("HIP", "RSVD"), ("HIP_MIXED", "RSVD"), ("HIP", "SVD"), ("HIP", "RRQR"),
])
def test_batch_of_identical_problems_equals_single_call(backend, engine):
    """Two copies of one problem in a batch give exactly the single-call
    ranks and factors (same tile shapes, so the same launches per tile),
    for the randomized engine and the exact ones."""
    _say(f"case: {backend} {engine} batch of 2 identical 2D exp N=3000 "
         "nb=300")
    try:
        _identical_pair(backend, engine)
    finally:
        S.init("HIP", lrengine="RSVD")


def _identical_pair(backend, engine):
    fn = S.init(backend, oversample=10, lrengine=engine)
    spec = ("spatial", "exp", 3000, 2, 300, 0.1, 7)
    single = _problems([spec])[0]
    assert single.approximate(150, 1e-7, 0, fn) == 0
    pair = _problems([spec, spec])
    assert S.approximate_batch(pair, 150, 1e-7) == 0
    r0 = single.ranks()
    for p in pair:
        assert p.tiles() == single.tiles()
        np.testing.assert_array_equal(p.ranks(), r0)
        for b in range(len(r0)):
            if r0[b] > 0:
                u0, v0 = single.factors(b)
                u1, v1 = p.factors(b)
                np.testing.assert_array_equal(u1, u0)
                np.testing.assert_array_equal(v1, v0)
    _free(pair + [single])


@pytest.mark.parametrize("backend", ["HIP", "HIP_MIXED"])
def test_batch_mixed_sizes_and_kernels(backend):
    """Problems of different sizes, tile sizes, kernels and dimensions in
    one call: every problem meets the accuracy bound, and its ranks agree
    with a call of its own up to near-ties of the rank rule."""
    _say(f"case: {backend} batch of 5 problems (exp 2D, sqexp 2D, "
         "matern2 3D)")
    fn = S.init(backend, oversample=10)
    specs = [("spatial", "exp", 2000, 2, 256, 0.05, 1),
             ("spatial", "exp", 4500, 2, 448, 0.2, 2),
             ("spatial", "sqrexp", 3600, 2, 384, 0.1, 3),
             ("spatial", "exp", 2900, 2, 320, 0.3, 4),
             ("spatial", "matern2", 4096, 3, 512, 0.1, 5)]
    tol, maxrank = 1e-7, 200
    batch = _problems(specs)
    assert S.approximate_batch(batch, maxrank, tol) == 0
    singles = _problems(specs)
    for pb, ps, spec in zip(batch, singles, specs):
        assert ps.approximate(maxrank, tol, 0, fn) == 0
        g, terr, tnorm = pb.dfe_tiles()
        _say(f"  {spec[1]} N={spec[2]}: error {g:.3e}")
        assert g <= 10 * tol
        rb, rs = pb.ranks(), ps.ranks()
        if pb.tiles() == ps.tiles():
            assert np.mean(np.abs(rb - rs) <= 1) >= 0.95
    _free(batch + singles)


def test_batch_empty_and_bad_arguments():
    S.init("HIP", oversample=10)
    assert S.approximate_batch([], 10, 1e-7) == 0
    p = _problems([("spatial", "exp", 1000, 2, 250, 0.1, 0)])
    assert S.approximate_batch(p, 0, 1e-7) != 0
    assert p[0].M is None
