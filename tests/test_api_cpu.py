"""Host side of the product library (no GPU needed): the C ABI loads and
exports every declared symbol, backend selection follows the reference's
init.c, and the host data model (points, Morton order, clusters, formats,
kernels, dsvfr, file I/O) matches the oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import starsh_hip as S
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _declared_functions():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|double|const char \*)\s+"
                             r"(?:\*\s*)?([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = S.lib()
    names = _declared_functions()
    assert len(names) > 120
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # global data of the API
    C.c_void_p.in_dll(L, "starsh_blrm_approximate")
    S.Stats()


def test_backend_selection_follows_init_table():
    L = S.lib()
    assert L.starsh_set_backend(b"HIP") == 0
    assert L.starsh_set_lrengine(b"RSVD") == 0
    f1 = S.approximate_fn()
    assert f1 is not None
    assert L.starsh_set_backend(b"HIP_MIXED") == 0
    f2 = S.approximate_fn()
    assert C.cast(f1, C.c_void_p).value != C.cast(f2, C.c_void_p).value
    # reference backends exist in the table but are not built here
    assert L.starsh_set_backend(b"OPENMP") == S.STARSH_WRONG_PARAMETER
    assert L.starsh_set_backend(b"NO_SUCH") == S.STARSH_WRONG_PARAMETER
    assert L.starsh_set_lrengine(b"NO_SUCH") == S.STARSH_WRONG_PARAMETER
    assert L.starsh_set_oversample(b"0") == S.STARSH_WRONG_PARAMETER
    assert L.starsh_set_oversample(b"16") == 0
    assert C.c_int.in_dll(L, "starsh_params").value == 10  # backend enum


def test_init_reads_environment(monkeypatch):
    L = S.lib()
    monkeypatch.setenv("STARSH_BACKEND", "HIP_MIXED")
    monkeypatch.setenv("STARSH_LRENGINE", "RSVD")
    monkeypatch.setenv("STARSH_OVERSAMPLE", "16")
    assert L.starsh_init() == 0
    params = (C.c_int * 3).in_dll(L, "starsh_params")
    assert list(params) == [10, 3, 16]
    monkeypatch.setenv("STARSH_BACKEND", "BOGUS")
    monkeypatch.setenv("STARSH_OVERSAMPLE", "-3")
    assert L.starsh_init() == 0       # falls back to the defaults
    assert list(params) == [9, 3, 10]


@pytest.mark.parametrize("backend", [b"HIP", b"HIP_MIXED"])
def test_engine_table(backend):
    """STARSH_LRENGINE selects the engine as the reference's tables do
    (include/control/init.h:111-190): SVD and DCSVD the SVD engine, RRQR
    the pivoted-QR engine, RSVD and CROSS the randomized SVD (the mixed
    form under HIP_MIXED)."""
    L = S.lib()
    addr = {name: C.cast(getattr(L, name), C.c_void_p).value
            for name in ("starsh_blrm__dsdd_hip", "starsh_blrm__dqp3_hip",
                         "starsh_blrm__drsdd_hip",
                         "starsh_blrm__drsdd_hip_mixed")}
    rsvd = ("starsh_blrm__drsdd_hip_mixed" if backend == b"HIP_MIXED" else
            "starsh_blrm__drsdd_hip")
    want = {b"SVD": "starsh_blrm__dsdd_hip", b"DCSVD": "starsh_blrm__dsdd_hip",
            b"RRQR": "starsh_blrm__dqp3_hip", b"RSVD": rsvd, b"CROSS": rsvd}
    assert L.starsh_set_backend(backend) == 0
    try:
        for eng, name in want.items():
            assert L.starsh_set_lrengine(eng) == 0
            fp = C.c_void_p.in_dll(L, "starsh_blrm_approximate").value
            assert fp == addr[name], eng
    finally:
        assert L.starsh_set_backend(b"HIP") == 0
        assert L.starsh_set_lrengine(b"RSVD") == 0


@pytest.mark.parametrize("place,ndim,seed", [(6, 2, 0), (6, 3, 1), (1, 2, 5),
                                              (1, 3, 2)])
def test_generated_points_match_oracle(place, ndim, seed):
    N = 2000
    p = S.TLRProblem("spatial", N, ndim, 500, seed=seed, place=place,
                     kernel="exp")
    raw = (O.gen_quasiuniform1(N, ndim, seed) if place == 6 else
           O.gen_rand(N, ndim, seed))
    ref, _ = O.zsort(raw)
    assert np.array_equal(p.points(), ref)


def test_c2_points_are_the_jittered_grid():
    p = S.TLRProblem("spatial", 64000, 2, 1600, seed=1, kernel="sqrexp")
    pts = p.points()
    h = 1.0 / 253
    cell = np.floor(pts / h)
    off = pts - cell * h
    assert np.all(off >= 0.1 * h - 1e-12) and np.all(off <= 0.9 * h + 1e-12)
    assert len(p.tiles()) == 820


def test_spheres_generator():
    L = S.lib()
    S.srand(3)
    part = C.POINTER(S.Particles)()
    assert L.starsh_particles_generate_spheres(C.byref(part), S.STARSH_int(20),
                                               S.STARSH_int(50),
                                               C.c_double(0.02)) == 0
    pp = part.contents
    pts = np.ctypeslib.as_array(pp.point, shape=(3 * 1000,)).reshape(3, 1000)
    # points lie on spheres of radius 0.02: every point has 49 others within
    # 0.04 (its own sphere)
    d = np.sqrt(((pts[:, :, None] - pts[:, None, :]) ** 2).sum(0))
    assert np.all(np.sum(d <= 0.04 + 1e-12, axis=1) >= 50)
    L.starsh_particles_free(part)


def test_tlr_format_structure():
    p = S.TLRProblem("spatial", 1030, 2, 100, symm="S", seed=0, kernel="exp")
    st, sz = p.cluster_ranges()
    assert len(st) == 11 and sz[-1] == 30 and all(s == 100 for s in sz[:-1])
    assert st == [100 * i for i in range(11)]
    tiles = p.tiles()
    assert tiles == [(i, j) for i in range(11) for j in range(i + 1)]
    q = S.TLRProblem("spatial", 1030, 2, 100, symm="N", seed=0, kernel="exp")
    assert q.tiles() == [(i, j) for i in range(11) for j in range(11)]
    F = q.F.contents
    # CSR/CSC views of the COO list
    for i in range(11):
        row = [F.brow_far[x] for x in range(F.brow_far_start[i],
                                            F.brow_far_start[i + 1])]
        assert [tiles_q for tiles_q in row] == [i * 11 + j for j in range(11)]
        col = [F.bcol_far[x] for x in range(F.bcol_far_start[i],
                                            F.bcol_far_start[i + 1])]
        assert col == [r * 11 + i for r in range(11)]


@pytest.mark.parametrize("kernel,kw", [
    ("exp", dict(beta=0.1)), ("sqrexp", dict(beta=0.2)),
    ("matern", dict(beta=0.1, nu=1.5)), ("matern2", dict(beta=0.1, nu=2.5)),
    ("matern", dict(beta=0.1, nu=0.8)), ("matern2", dict(beta=0.3, nu=4.3))])
def test_host_spatial_kernels_match_oracle(kernel, kw):
    p = S.TLRProblem("spatial", 900, 3, 300, seed=2, place=1, kernel=kernel,
                     **kw)
    pts = p.points()
    kd = O.kernel(kernel, 3, **kw)
    for i, j in [(0, 0), (2, 1)]:
        A = p.host_block(i, j)
        R = O.kernel_block(kd, pts, pts, np.arange(300 * i, 300 * i + 300),
                           np.arange(300 * j, 300 * j + 300))
        if kw.get("nu", 0.5) in (0.5, 1.5, 2.5):
            assert np.array_equal(A, R)
        else:
            # general nu: two K_nu methods, each within a few 1e-14
            assert np.max(np.abs(A - R)) <= 1e-13 * np.max(np.abs(R))


@pytest.mark.parametrize("kind", ["coulomb", "sin", "cos"])
def test_host_particle_kernels_match_oracle(kind):
    p = S.TLRProblem(kind, 800, 3, 200, seed=2, place=1, wave_k=6.0,
                     diag=1.5)
    pts = p.points()
    kd = O.kernel(kind, 3, wave_k=6.0, diag=1.5)
    A = p.host_block(1, 1)
    R = O.kernel_block(kd, pts, pts, np.arange(200, 400),
                       np.arange(200, 400))
    assert np.array_equal(A, R)
    assert np.all(np.diag(A) == (0.0 if kind == "coulomb" else 1.5))


def test_matern_accepts_general_nu_and_rejects_zero():
    """Any nu > 0 is a Matern kernel (general nu through K_nu); nu = 0
    makes 2^(1-nu)/Gamma(nu) x^nu K_nu(x) degenerate and is refused."""
    L = S.lib()
    for nu, want in ((0.7, 0), (3.3, 0), (0.0, S.STARSH_WRONG_PARAMETER)):
        data, kern = C.c_void_p(), C.c_void_p()
        info = L.starsh_application(
            C.byref(data), C.byref(kern), S.STARSH_int(100), C.c_char(b"d"),
            C.c_int(S.STARSH_SPATIAL), C.c_int(S.STARSH_SPATIAL_MATERN),
            C.c_int(S.STARSH_SPATIAL_NU), C.c_double(nu), C.c_int(0))
        assert info == want, nu
        if info == 0:
            L.starsh_ssdata_free(data)


def test_product_dsvfr_equals_oracle():
    L = S.lib()
    rng = np.random.default_rng(4)
    for _ in range(200):
        k = int(rng.integers(1, 60))
        s = np.sort(rng.random(k) ** rng.integers(1, 30))[::-1].copy()
        tol = 10.0 ** rng.uniform(-12, -1)
        got = L.starsh_dense_dsvfr(k, s.ctypes.data_as(C.POINTER(C.c_double)),
                                   tol)
        assert got == O.dsvfr(s, tol)


def test_particles_io_roundtrip_reference_fixture(tmp_path):
    """The reference's particles_io test on its own fixture
    (testing/particles.c:20-89): ASCII read, ASCII and binary write/read
    give identical points."""
    L = S.lib()
    a = C.POINTER(S.Particles)()
    b = C.POINTER(S.Particles)()
    assert L.starsh_particles_read_from_file(
        C.byref(a), os.path.join(GOLD, "particles.txt").encode(), 1) == 0
    assert a.contents.count == 36 and a.contents.ndim == 3
    pts = np.ctypeslib.as_array(a.contents.point, shape=(108,)).copy()
    assert pts[0] == 1.0 and pts[36] == 2.0 and pts[72] == 3.0
    for ftype, name in ((1, "p.txt"), (2, "p.bin")):
        path = str(tmp_path / name).encode()
        assert L.starsh_particles_write_to_file(a, path, ftype) == 0
        assert L.starsh_particles_read_from_file(C.byref(b), path, ftype) == 0
        got = np.ctypeslib.as_array(b.contents.point, shape=(108,))
        assert np.array_equal(got, pts)
        L.starsh_particles_free(b)
    L.starsh_particles_free(a)
    assert L.starsh_particles_read_from_file(
        C.byref(b), str(tmp_path / "missing").encode(), 1) == 2


def test_hip_driver_rejects_foreign_kernel_before_touching_gpu():
    """A kernel pointer the library does not own has no device generator:
    the driver returns STARSH_WRONG_PARAMETER (the reference's GPU path
    silently assumed a 2D exponential kernel)."""
    L = S.lib()
    fn = S.init("HIP")
    p = S.TLRProblem("spatial", 400, 2, 100, seed=0, kernel="exp")
    KT = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_int)
    foreign = KT(lambda *a: None)
    p.P.contents.kernel = C.cast(foreign, C.c_void_p)
    assert p.approximate(20, 1e-6, 0, fn) == S.STARSH_WRONG_PARAMETER


def test_gram_schmidt_panel_option():
    """qr_panel takes 0 (by batch size, the default), 4 or 8 and rejects
    other widths; setting it touches no GPU."""
    L = S.lib()
    for w in (4, 8, 0):
        assert L.starsh_hip_set_option(b"qr_panel", C.c_double(w)) == 0
    assert L.starsh_hip_set_option(b"qr_panel", C.c_double(5)) != 0
    assert L.starsh_hip_set_option(b"qr_panel", C.c_double(0)) == 0


@pytest.mark.parametrize("backend", ["HIP", "HIP_MIXED"])
def test_oversized_tiles_rejected_before_gpu_work(backend):
    """Tiles over 4096 rows/columns and sketch widths min(maxrank +
    oversample, m, n) over 1024 are refused by argument checking, before the
    driver touches the GPU (this container has none)."""
    fn = S.init(backend, oversample=16)
    # one 5000 x 5000 tile
    p = S.TLRProblem("spatial", 5000, 2, 5000, symm="S", seed=0,
                     kernel="exp", beta=0.1)
    assert p.approximate(100, 1e-7, 0, fn) == S.STARSH_WRONG_PARAMETER
    assert not p.M
    # tiles of 2000 with maxrank 1100: sketch width 1116
    p = S.TLRProblem("spatial", 4000, 2, 2000, symm="S", seed=0,
                     kernel="exp", beta=0.1)
    assert p.approximate(1100, 1e-7, 0, fn) == S.STARSH_WRONG_PARAMETER
    assert not p.M
    assert len(p.tiles()) == 3


def test_jacobi_sweep_option_is_checked():
    L = S.lib()
    for bad in (-1, 61, 1000):
        assert L.starsh_hip_set_option(b"jacobi_max_sweeps",
                                       C.c_double(bad)) == \
            S.STARSH_WRONG_PARAMETER
    assert L.starsh_hip_set_option(b"jacobi_max_sweeps", C.c_double(5)) == 0
    assert L.starsh_hip_set_option(b"jacobi_max_sweeps", C.c_double(0)) == 0


def test_mixed_schedule_options_are_checked():
    """The mixed form's schedule options take their documented values and
    refuse others, without touching a GPU: mixed_ladder -1 (the autotuner's
    table), 0 or 1; mixed_first_bucket -1 or an index 0..4;
    tile_schedule 0 or 1."""
    L = S.lib()
    W = S.STARSH_WRONG_PARAMETER
    for name, good, bad in ((b"mixed_ladder", (0, 1, -1), (2, -2, 5)),
                            (b"mixed_first_bucket", (0, 4, -1), (5, -2)),
                            (b"tile_schedule", (0, 1), (2, -1))):
        for v in bad:
            assert L.starsh_hip_set_option(name, C.c_double(v)) == W, \
                (name, v)
        for v in good:
            assert L.starsh_hip_set_option(name, C.c_double(v)) == 0, \
                (name, v)
