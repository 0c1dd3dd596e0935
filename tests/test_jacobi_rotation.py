"""Host test of the plane rotation the Jacobi core kernels use
(stars-h_amd/csrc/hip/jacobi_rotation.h, compiled here with g++).

A pair of columns with squared norms al, be and inner product ga is
rotated by (c, s).  The rotated pair's inner product
    cs (al - be) + (c^2 - s^2) ga
is evaluated exactly (rational arithmetic on the doubles c, s, al, be, ga)
and must vanish to working precision relative to the rotated norms.  The
cases cover ordinary magnitudes and the ends of the FP64 range, where the
squares of d = be - al and h = 2 ga inside the formula underflow or
overflow (al, be near 1e-160, 1e-200, 1e-300, 1e200, 1e300; d == 0 with
h*h below the smallest double), and no value may become NaN or Inf.
Columns of two rows check the whole step (sums, test, rotation, update) at
magnitudes whose squares are normal numbers.
"""
import math
import os
import subprocess
from fractions import Fraction as Fr

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "stars-h_amd", "csrc", "hip")


@pytest.fixture(scope="module")
def rot_bin(tmp_path_factory):
    out = tmp_path_factory.mktemp("jr") / "jr"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                           "-I", HDR,
                           os.path.join(ROOT, "tests",
                                        "jacobi_rotation_host.cpp"),
                           "-o", str(out)])
    return str(out)


def _run(binary, lines):
    res = subprocess.run([binary], input="\n".join(lines) + "\n",
                         capture_output=True, text=True, check=True)
    return [[float(x) for x in line.split()]
            for line in res.stdout.strip().splitlines()]


def _gram_cases():
    rng = np.random.default_rng(7)
    out = []
    for scale in (1.0, 1e-100, 1e-160, 1e-200, 1e-300, 1e100, 1e200, 1e300):
        for _ in range(30):
            a, b = rng.uniform(0.1, 1.0, 2)
            cosang = rng.uniform(-0.999, 0.999)
            out.append((a * scale, b * scale,
                        cosang * math.sqrt(a * b) * scale))
        # equal norms (d == 0) and a tiny inner product: h*h underflows
        out.append((scale, scale, 3e-3 * scale))
        out.append((scale, scale, 1e-12 * scale))
    # the case of a far tile of a fast-decaying kernel
    out.append((1e-160, 1e-160, 3e-161))
    out.append((1e-160, 1.0000001e-160, 3e-161))
    return out


def test_rotation_orthogonalises_across_the_fp64_range(rot_bin):
    cases = _gram_cases()
    res = _run(rot_bin, [f"g {float(al)!r} {float(be)!r} {float(ga)!r}"
                         for al, be, ga in cases])
    assert len(res) == len(cases)
    for (al, be, ga), (rot, c, s) in zip(cases, res):
        assert rot == 1, (al, be, ga)
        assert math.isfinite(c) and math.isfinite(s), (al, be, ga, c, s)
        assert abs(c * c + s * s - 1.0) <= 1e-15, (al, be, ga)
        C, Sn = Fr(c), Fr(s)
        A, B, G = Fr(al), Fr(be), Fr(ga)
        off = C * Sn * (A - B) + (C * C - Sn * Sn) * G
        na = C * C * A - 2 * C * Sn * G + Sn * Sn * B
        nb = Sn * Sn * A + 2 * C * Sn * G + C * C * B
        assert na > 0 and nb > 0
        rel = abs(off) / Fr(math.sqrt(float(na / A))) / \
            Fr(math.sqrt(float(nb / A))) / A
        assert float(rel) <= 2e-15, (al, be, ga, float(rel))


def test_rotation_test_does_not_underflow(rot_bin):
    """|ga| > ctol |a||b| with squared norms near 1e-160: al*be underflows
    to 0 and would accept every pair; the test takes the norms one at a
    time."""
    res = _run(rot_bin, ["g 1e-160 1e-160 0",          # orthogonal
                         "g 1e-160 1e-160 1e-190",     # orthogonal to 1e-30
                         "g 1e-160 1e-160 3e-161"])    # far from it
    assert [r[0] for r in res] == [0, 0, 1]


def test_column_pairs(rot_bin):
    rng = np.random.default_rng(11)
    cols = []
    for scale in (1.0, 1e-75, 1e-60, 1e60, 1e75, 1e100, 1e150):
        for _ in range(20):
            cols.append(tuple(rng.standard_normal(4) * scale))
        cols.append((scale, scale * 1e-3, scale * 1e-3, scale))
    res = _run(rot_bin, ["c " + " ".join(repr(float(v)) for v in c)
                         for c in cols])
    for (a0, a1, b0, b1), (rot, c, s, x0, x1, y0, y1) in zip(cols, res):
        for v in (c, s, x0, x1, y0, y1):
            assert math.isfinite(v), (a0, a1, b0, b1)
        if not rot:
            continue
        sc = max(abs(x0), abs(x1), abs(y0), abs(y1))
        u = np.array([x0, x1]) / sc
        v = np.array([y0, y1]) / sc
        rel = abs(u @ v) / (np.linalg.norm(u) * np.linalg.norm(v))
        assert rel <= 5e-14, (a0, a1, b0, b1, rel)
