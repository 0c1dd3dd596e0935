"""The byte-digit split of the int8 sketch, restated in numpy.

The mixed sketch (stars-h_amd/csrc/hip/sketch_i8.hip) rounds each entry to
an integer X with |X| <= 2^30 and reads four signed int8 digits off the
bytes of Z = (X + 0x808080) ^ 0x808080, gathering byte s of four entries
into one word with v_perm_b32.  These tests check, on the CPU, that the
digits recombine to X, that they stay in int8 range, that a chunk's int32
sums cannot overflow, and that the v_perm_b32 selectors (emulated from the
instruction's definition) gather the right bytes.
"""
import numpy as np


def _digits(x):
    z = ((x.astype(np.int64) + 0x808080) & 0xFFFFFFFF) ^ 0x808080
    z = z.astype(np.uint32)
    d = np.stack([(z >> (8 * s)) & 0xFF for s in range(4)]).astype(np.int64)
    return np.where(d >= 128, d - 256, d)       # int8 view of each byte


def test_digits_recombine_and_stay_in_range():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.integers(-2**30, 2**30 + 1, 200000),
                        np.array([0, 1, -1, 127, 128, -128, -129, 2**30,
                                  -2**30, 0x7F7F7F, -0x808080, 0x808080])])
    d = _digits(x)
    back = d[3] * 2**24 + d[2] * 2**16 + d[1] * 2**8 + d[0]
    np.testing.assert_array_equal(back, x)
    assert d[:3].min() >= -128 and d[:3].max() <= 127
    assert np.abs(d[3]).max() <= 64


def test_chunk_sums_fit_int32_and_recombine_exactly():
    rng = np.random.default_rng(1)
    for _ in range(20):
        x = rng.integers(-2**30, 2**30 + 1, 64)
        x[rng.integers(0, 64, 4)] = rng.choice([2**30, -2**30], 4)
        om = rng.integers(-127, 128, 64)
        om[:8] = rng.choice([127, -127], 8)
        d = _digits(x)
        s = d @ om
        assert np.abs(s).max() < 2**21
        hi = s[3] * 256 + s[2]
        lo = s[1] * 256 + s[0]
        assert abs(hi) < 2**31 and abs(lo) < 2**31
        assert float(hi) * 65536.0 + float(lo) == float(int(x @ om))


def _perm(a, b, sel):
    """v_perm_b32 D = perm(S0=a, S1=b, S2=sel): result byte i is byte
    sel.byte[i] of the 64-bit value {a, b} (b in the low half) for
    selectors 0-7."""
    data = (int(a) << 32) | int(b)
    out = 0
    for i in range(4):
        k = (sel >> (8 * i)) & 0xFF
        assert k < 8
        out |= ((data >> (8 * k)) & 0xFF) << (8 * i)
    return out


def test_byte_plane_selectors():
    rng = np.random.default_rng(2)
    for _ in range(200):
        z = [int(v) for v in rng.integers(0, 2**32, 4, dtype=np.uint64)]
        t01l = _perm(z[1], z[0], 0x05010400)
        t01h = _perm(z[1], z[0], 0x07030602)
        t23l = _perm(z[3], z[2], 0x05010400)
        t23h = _perm(z[3], z[2], 0x07030602)
        p = [_perm(t23l, t01l, 0x05040100), _perm(t23l, t01l, 0x07060302),
             _perm(t23h, t01h, 0x05040100), _perm(t23h, t01h, 0x07060302)]
        for s in range(4):
            want = sum(((z[e] >> (8 * s)) & 0xFF) << (8 * e) for e in range(4))
            assert p[s] == want


def _fma(a, b, c):
    L = np.longdouble
    return (L(a) * L(b) + L(c)).astype(np.float64)


def _sexp_nonpos(x, s=1.0):
    """Numpy restatement of sexp_nonpos (stars-h_amd/csrc/hip/hip_common.h):
    Tang's table-driven exp with 1024 table entries s*2^(j/1024), a
    degree-3 Taylor polynomial and a one-constant reduction; fma emulated
    in extended precision."""
    x = np.maximum(x, -1000.0)
    shift = 1.5 * 2.0**52
    sh = _fma(x, 1477.3197218702985, shift)
    n = sh.view(np.int64).astype(np.int32)         # low 32 bits
    nd = sh - shift
    r = _fma(-nd, 0.0006769015435155716, x)
    p = _fma(r, 1.0 / 6.0, 0.5)
    p = _fma(p, r, 1.0)
    p = p * r
    tab = s * np.exp2(np.arange(1024) / 1024.0)
    t = tab[n & 1023]
    return np.ldexp(_fma(t, p, t), (n >> 10).astype(np.int32))


def test_table_exp_accuracy():
    rng = np.random.default_rng(3)
    x = np.concatenate([-rng.uniform(0, 1000, 200000),
                        -rng.uniform(0, 1, 20000), np.array([0.0, -1e-300])])
    got = _sexp_nonpos(x)
    want = np.exp(x.astype(np.longdouble))
    ok = want > 1e-300          # normal range
    rel = np.abs((got[ok] - want[ok]) / want[ok]).astype(np.float64)
    assert rel.max() <= 2.0**-44
    assert _sexp_nonpos(np.array([-1e6, -1000.0]))[0] == 0.0


def test_table_exp_folded_factor():
    """The factor in the table gives s*exp(x), exactly s at x = 0 (the
    value at distance 0 without a select)."""
    s = 1.7
    assert _sexp_nonpos(np.array([0.0, -0.0]), s).tolist() == [s, s]
    x = -np.linspace(0, 50, 1001)
    want = s * np.exp(x.astype(np.longdouble))
    rel = np.abs((_sexp_nonpos(x, s) - want) / want).astype(np.float64)
    assert rel.max() <= 2.0**-44


def test_scaled_byte_digits_rounding():
    """scaled_byte_digits: the low word of 1.5*2^52 + 0x808080 + a*scale
    (a*scale exact for a power-of-two scale, so one rounding) equals
    byte_digits(rint(a*scale)), ties to even, over [-2^30, 2^30]."""
    rng = np.random.default_rng(4)
    ex = rng.integers(-60, 60, 50000)
    scale = np.ldexp(1.0, 30 - ex)
    f = rng.uniform(-1.0, 1.0, 50000)
    a = np.ldexp(f, ex)
    a[:8] = np.ldexp(np.array([0.5, -0.5, 1.5, -1.5, 2.5, 0.0, 1.0, -1.0]),
                     -30 + ex[:8])           # exact ties and extremes
    sh = a * scale + (1.5 * 2.0**52 + 8421504.0)
    low = (sh.view(np.int64) & 0xFFFFFFFF).astype(np.uint64)
    z = (low ^ 0x808080).astype(np.int64)
    want = np.rint(a * scale).astype(np.int64)
    assert np.abs(want).max() <= 2**30
    got = np.stack([(z >> (8 * s)) & 0xFF for s in range(4)])
    got = np.where(got >= 128, got - 256, got)
    back = got[3] * 2**24 + got[2] * 2**16 + got[1] * 2**8 + got[0]
    np.testing.assert_array_equal(back, want)
