"""The oracle's restatements of the TLR product (dmml.c) and of conjugate
gradients (cg.c), pinned against dense linear algebra on the CPU.

dmml is checked against the product with the dense matrix assembled block
by block from the same factors (U V^T at (i, j) and, symmetric, its
transpose at (j, i); near tiles D and D^T), for symmetric and
nonsymmetric formats, several right-hand sides, alpha and beta (beta = 0
must ignore NaN in B, as the reference zeroes B first).  cg is checked
against numpy.linalg.solve on a well-conditioned SPD matrix, and the
stopping rule against the residual it reports.
"""
import numpy as np

from oracle import oracle as O


def _random_tlr(rng, sizes, symm, rank=3):
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(int)
    nb = len(sizes)
    far, Us, Vs, near, Ds = [], [], [], [], []
    for i in range(nb):
        for j in range(nb if symm == "N" else i + 1):
            if (i + j) % 3 == 0:
                near.append((i, j))
                Ds.append(rng.standard_normal((sizes[i], sizes[j])))
            else:
                r = int(rng.integers(0, rank + 1))
                far.append((i, j))
                Us.append(rng.standard_normal((sizes[i], r)))
                Vs.append(rng.standard_normal((sizes[j], r)))
    return starts, far, Us, Vs, near, Ds


def _dense(starts, sizes, symm, far, Us, Vs, near, Ds):
    n = int(sum(sizes))
    A = np.zeros((n, n))
    blocks = [((i, j), U @ V.T) for (i, j), U, V in zip(far, Us, Vs)]
    blocks += list(zip(near, Ds))
    for (i, j), Bk in blocks:
        A[starts[i]:starts[i] + sizes[i], starts[j]:starts[j] + sizes[j]] = Bk
        if symm == "S" and i != j:
            A[starts[j]:starts[j] + sizes[j],
              starts[i]:starts[i] + sizes[i]] = Bk.T
    return A


def test_dmml_matches_dense_assembly():
    rng = np.random.default_rng(0)
    sizes = [7, 5, 9, 4]
    for symm in ("S", "N"):
        starts, far, Us, Vs, near, Ds = _random_tlr(rng, sizes, symm)
        A = _dense(starts, sizes, symm, far, Us, Vs, near, Ds)
        n = A.shape[0]
        X = rng.standard_normal((n, 3))
        Y = rng.standard_normal((n, 3))
        got = O.dmml(starts, sizes, symm, far, Us, Vs, near, Ds, X, 0.7,
                     0.3, Y)
        np.testing.assert_allclose(got, 0.7 * A @ X + 0.3 * Y, rtol=1e-13,
                                   atol=1e-12)
        Yn = np.full((n, 3), np.nan)
        got = O.dmml(starts, sizes, symm, far, Us, Vs, near, Ds, X, -1.0,
                     0.0, Yn)
        np.testing.assert_allclose(got, -A @ X, rtol=1e-13, atol=1e-12)


def test_cg_solves_spd_system_and_stops_by_the_rule():
    rng = np.random.default_rng(1)
    n = 60
    G = rng.standard_normal((n, n))
    A = G @ G.T / n + np.eye(n)
    B = rng.standard_normal((n, 2))
    X0 = np.zeros((n, 2))
    tol = 1e-10
    it, X = O.cg(lambda P: A @ P, B, X0, tol)
    assert 0 <= it < n
    np.testing.assert_allclose(X, np.linalg.solve(A, B), rtol=1e-7,
                               atol=1e-8)
    # every column's true residual is at the tolerance of its start
    r0 = np.linalg.norm(B - A @ X0, axis=0)
    r = np.linalg.norm(B - A @ X, axis=0)
    assert np.all(r <= 10 * tol * r0)
