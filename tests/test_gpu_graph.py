"""Gram-Schmidt panels as one captured hipGraph (starsh_hip_set_option
"qr_graph" 1) against the default host-launched panels: the same ranks and
the same error, bit for bit, in both forms, on a first call (capture) and a
second one (replay of the kept graph)."""
import ctypes as C

import pytest

import starsh_hip as S

pytestmark = pytest.mark.gpu


def _opt(name, value):
    assert S.lib().starsh_hip_set_option(name.encode(),
                                         C.c_double(value)) == 0, name


@pytest.mark.parametrize("backend", ["HIP", "HIP_MIXED"])
def test_graph_panels_match_host_launches(backend):
    print(f"case: qr_graph {backend} 2D sqexp N=12800 nb=1600", flush=True)
    res = {}
    try:
        for g in (0, 1):
            _opt("qr_graph", g)
            fn = S.init(backend, oversample=16)
            out = []
            for _ in range(2):
                p = S.TLRProblem("spatial", 12800, 2, 1600, symm="S", seed=1,
                                 kernel="sqrexp", beta=0.1)
                assert p.approximate(400, 1e-7, 0, fn) == 0
                out.append((p.ranks().tolist(), p.dfe()))
                p.free_matrix()
            res[g] = out
    finally:
        _opt("qr_graph", 0)
    for (r0, e0), (r1, e1) in zip(res[0], res[1]):
        assert r0 == r1
        assert e0 == e1
