"""The headline workload at full size on the GPU: 2D squared exponential,
beta 0.1, N=64000 (253 x 253 jittered grid, srand 1, Morton order), tile
1600, maxrank 400, oversample 16, eps 1e-7, symmetric format (820 tiles).

Plain form (promise: the reference's ranks and errors):
  * global relative Frobenius error <= eps, every tile regenerated
    (starsh_blrm__dfe_hip_tiles), and every per-tile error <= 1.01 eps
    (the rank rule bounds the truncation relative to |B|_F, and the
    sketch of these smooth tiles captures the range almost exactly);
  * 10 sampled tiles -- four diagonal tiles, the first sub-diagonal, and
    far tiles across the matrix -- compressed by the oracle (the CPU
    restatement of dlrrsdd with the host LAPACK): equal ranks (or a
    near-tie of the rank rule), and per-tile errors within 2% of the
    oracle's error for the same tile.
Mixed form: global error <= eps, per-tile errors <= 2 eps, ranks within 3
of the plain form's on the same tiles.
Product with the mixed form's factors (one and three right-hand sides):
within 1e-13 relative of the oracle's dmml on the same factors.
"""
import numpy as np
import pytest

import starsh_hip as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL, MAXRANK, OS = 1e-7, 400, 16


def _say(msg):
    print(msg, flush=True)


def _c2():
    return S.TLRProblem("spatial", 64000, 2, 1600, symm="S", seed=1,
                        place=S.STARSH_PARTICLES_QUASIUNIFORM1,
                        kernel="sqrexp", beta=0.1)


def _sample(tiles):
    want = [(0, 0), (13, 13), (27, 27), (39, 39), (1, 0), (20, 19),
            (39, 38), (10, 0), (30, 5), (39, 0)]
    return [tiles.index(t) for t in want if t in tiles]


@pytest.fixture(scope="module")
def plain():
    _say("case: C2 full size, plain form")
    fn = S.init("HIP", oversample=OS)
    p = _c2()
    assert p.N == 64000
    assert len(p.tiles()) == 820
    assert p.approximate(MAXRANK, TOL, 0, fn) == 0
    st = S.stats()
    g, terr, tnorm = p.dfe_tiles()
    ranks = p.ranks()
    _say(f"  plain: device {st.device_seconds:.4f} s, wall "
         f"{st.wall_seconds:.4f} s, ranks {ranks.min()}..{ranks.max()}, "
         f"error {g:.4e}, max tile error {terr.max():.4e}")
    yield p, g, terr, ranks
    p.free_matrix()


def test_c2_plain_errors(plain):
    p, g, terr, ranks = plain
    assert len(p.near_tiles()) == 0
    assert g <= TOL
    assert terr.max() <= 1.01 * TOL


def test_c2_plain_sampled_tiles_match_oracle(plain):
    p, g, terr, ranks = plain
    tiles = p.tiles()
    pick = _sample(tiles)
    assert len(pick) >= 8
    st, sz = p.cluster_ranges()
    pts = p.points()
    kd = O.kernel("sqrexp", 2, beta=0.1)
    sub = [tiles[b] for b in pick]
    ref = O.compress(kd, pts, st, sz, sub, MAXRANK, OS, TOL, want_S=True)
    # the oracle's per-tile errors on the same tiles (its own factors)
    _, tile_norm, tile_diff = O.dfe(kd, pts, st, sz, "N", sub, ref["rank"],
                                    ref["U"], ref["V"], [])
    for q, b in enumerate(pick):
        rr, rg = int(ref["rank"][q]), int(ranks[b])
        e_ref = tile_diff[q] / tile_norm[q]
        _say(f"  tile {tiles[b]}: rank gpu {rg} oracle {rr}, error gpu "
             f"{terr[b]:.4e} oracle {e_ref:.4e}")
        if rr != rg:
            # only at a near-tie of the rank rule
            s = ref["S"][q]
            tot = np.sum(s * s)
            tail = np.sum(s[rr:] ** 2)
            assert abs(tail - TOL * TOL * tot) <= 1e-6 * TOL * TOL * tot, \
                (tiles[b], rg, rr)
            continue
        assert abs(terr[b] - e_ref) <= 0.02 * e_ref + 1e-15, \
            (tiles[b], terr[b], e_ref)


def test_c2_mixed_errors_and_ranks(plain):
    p0, g0, terr0, ranks0 = plain
    plain_rank = dict(zip(p0.tiles(), ranks0.tolist()))
    _say("case: C2 full size, mixed form")
    fn = S.init("HIP_MIXED", oversample=OS)
    p = _c2()
    assert p.approximate(MAXRANK, TOL, 0, fn) == 0
    st = S.stats()
    g, terr, _ = p.dfe_tiles()
    ranks = p.ranks()
    _say(f"  mixed: device {st.device_seconds:.4f} s, wall "
         f"{st.wall_seconds:.4f} s, ranks {ranks.min()}..{ranks.max()}, "
         f"error {g:.4e}, max tile error {terr.max():.4e}")
    assert g <= TOL
    assert terr.max() <= 2 * TOL
    assert len(p.near_tiles()) == 0
    for t, r in zip(p.tiles(), ranks.tolist()):
        assert abs(r - plain_rank[t]) <= 3, (t, r, plain_rank[t])
    p.free_matrix()


def test_c2_product_matches_oracle():
    """The product with the compressed matrix at the headline shape (the
    mixed form's factors, ranks 7-31, 780 tiles with a transposed twin),
    one and three right-hand sides, against the oracle's restatement of
    dmml.c on the same factors."""
    _say("case: C2 full size, product with the mixed form's factors")
    fn = S.init("HIP_MIXED", oversample=OS)
    p = _c2()
    assert p.approximate(MAXRANK, TOL, 0, fn) == 0
    st, sz = p.cluster_ranges()
    far = p.tiles()
    UV = [p.factors(b) for b in range(len(far))]
    near = p.near_tiles()
    Ds = [p.near_block(b) for b in range(len(near))]
    rng = np.random.default_rng(11)
    for nrhs in (1, 3):
        A = rng.standard_normal((64000, nrhs))
        got = p.dmml(A)
        want = O.dmml(st, sz, "S", far, [u for u, _ in UV],
                      [v for _, v in UV], near, Ds, A, 1.0, 0.0,
                      np.zeros((64000, nrhs)))
        rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        _say(f"  nrhs {nrhs}: rel diff {rel:.2e}")
        assert rel <= 1e-13
    p.free_matrix()
