"""A C program written against the reference API only
(tests/c_api/spatial_hip.c: the call sequence of the reference's
testing/spatial.c:55-109, including starsh_blrm__dfe_omp and
starsh_blrm__dmml) is compiled with gcc against include/, linked with
-lstarsh_hip, and run with STARSH_BACKEND=HIP and HIP_MIXED and with the
engines STARSH_LRENGINE=RSVD, SVD and RRQR.  It must pass the reference's
acceptance rule rel_err / tol <= 10, and its error and ranks must agree
with the same problem driven through the Python mirror.
"""
import os
import re
import subprocess

import pytest

import starsh_hip as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "stars-h_amd")


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("capi") / "spatial_hip")
    subprocess.check_call(["gcc", "-std=gnu11", "-O2", "-Wall",
                           "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "c_api",
                                        "spatial_hip.c"),
                           "-L", PKG, "-lstarsh_hip",
                           "-Wl,-rpath," + PKG, "-o", out])
    return out


EXP = ["2", "6", "1", "0.1", "0.5", "2500", "250", "100", "1e-9"]
SQREXP = ["2", "6", "2", "0.1", "0.5", "6400", "800", "200", "1e-7"]


@pytest.mark.parametrize("backend,engine,args,tol", [
    # 2D exp, jittered grid, N=2500, tiles 250, maxrank 100, eps 1e-9
    ("HIP", "RSVD", EXP, 1e-9),
    ("HIP_MIXED", "RSVD", EXP, 1e-9),
    ("HIP", "SVD", EXP, 1e-9),
    ("HIP", "RRQR", EXP, 1e-9),
    # 2D sqexp, N=6400, tiles 800, maxrank 200, eps 1e-7
    ("HIP", "RSVD", SQREXP, 1e-7),
    ("HIP_MIXED", "RSVD", SQREXP, 1e-7),
    ("HIP_MIXED", "SVD", SQREXP, 1e-7),
])
def test_reference_style_c_caller(prog, backend, engine, args, tol):
    print(f"case: C caller {backend} {engine} {' '.join(args)}", flush=True)
    env = dict(os.environ, STARSH_BACKEND=backend, STARSH_LRENGINE=engine,
               STARSH_OVERSAMPLE="10")
    res = subprocess.run([prog] + args, capture_output=True, text=True,
                         env=env, timeout=300)
    print(res.stdout[-2000:], res.stderr[-2000:], flush=True)
    assert res.returncode == 0
    err = float(re.search(r"RELATIVE ERROR: (\S+)", res.stdout).group(1))
    assert 0 < err <= 10 * tol
    assert re.search(r"DMML INFO 0 ", res.stdout)
    # the same problem through the Python mirror ('N' format, srand(0))
    fn = S.init(backend, oversample=10, lrengine=engine)
    kern = {"1": "exp", "2": "sqrexp"}[args[2]]
    p = S.TLRProblem("spatial", int(args[5]), int(args[0]), int(args[6]),
                     symm="N", seed=0, kernel=kern, beta=float(args[3]))
    assert p.approximate(int(args[7]), tol, 0, fn) == 0
    e2 = p.dfe()
    r = p.ranks()
    p.free_matrix()
    assert abs(e2 - err) <= 1e-6 * err
    m = re.search(r"RANKS (\d+) (\S+) (\d+)", res.stdout)
    assert (int(m.group(1)), int(m.group(3))) == (int(r.min()), int(r.max()))
    S.init("HIP", lrengine="RSVD")
