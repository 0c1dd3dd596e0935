"""GPU Morton sort (starsh_hip_zsort, starsh_particles_zsort_inplace_hip)
against the oracle's restatement of the reference's recursive partition
(src/applications/particles.c:1090-1221): the permuted coordinates and the
permutation must be identical, ties included.

Point sets: the headline workload's points (253 x 253 jittered grid
truncated to 64,000, srand 1) and the 3D Matern workload's (131,072
uniform, srand 2), each in a shuffled order; coarse grids where every key
occurs many times (both tie parities); 1D and 4D points; and the binary
particle file format (starsh_particles_write_to_file / read_from_file,
testing/particles.c) feeding the GPU sort.
"""
import ctypes as C

import numpy as np
import pytest

import starsh_hip as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

dp = C.POINTER(C.c_double)


def _say(msg):
    print(msg, flush=True)


def _gpu_zsort(points):
    """(sorted points, order) from starsh_hip_zsort."""
    L = S.lib()
    pts = np.ascontiguousarray(points, dtype=np.float64)
    ndim, n = pts.shape
    dx, dy, do = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert L.starsh_hip_malloc(C.byref(dx), C.c_size_t(8 * pts.size)) == 0
    assert L.starsh_hip_malloc(C.byref(dy), C.c_size_t(8 * pts.size)) == 0
    assert L.starsh_hip_malloc(C.byref(do), C.c_size_t(8 * n)) == 0
    try:
        assert L.starsh_hip_memcpy_h2d(dx, pts.ctypes.data_as(dp),
                                       C.c_size_t(8 * pts.size)) == 0
        assert L.starsh_hip_zsort(dx, C.c_int64(n), C.c_int(ndim), dy,
                                  do) == 0
        out = np.empty_like(pts)
        order = np.empty(n, dtype=np.int64)
        assert L.starsh_hip_memcpy_d2h(out.ctypes.data_as(dp), dy,
                                       C.c_size_t(8 * pts.size)) == 0
        assert L.starsh_hip_memcpy_d2h(order.ctypes.data_as(C.c_void_p), do,
                                       C.c_size_t(8 * n)) == 0
    finally:
        for p in (dx, dy, do):
            L.starsh_hip_free(p)
    return out, order


def _check(points, name):
    _say(f"case: zsort {name}, ndim {points.shape[0]}, n {points.shape[1]}")
    want, worder = O.zsort(points)
    got, gorder = _gpu_zsort(points)
    assert np.array_equal(gorder, worder), name
    assert np.array_equal(got, want), name


def _shuffled(pts, seed):
    perm = np.random.default_rng(seed).permutation(pts.shape[1])
    return np.ascontiguousarray(pts[:, perm])


def test_zsort_headline_points():
    p = S.TLRProblem("spatial", 64000, 2, 1600, symm="S", seed=1,
                     place=S.STARSH_PARTICLES_QUASIUNIFORM1,
                     kernel="sqrexp", beta=0.1)
    pts = p.points()
    _check(pts, "C2 points as generated (already in Morton order)")
    _check(_shuffled(pts, 11), "C2 points shuffled")


def test_zsort_matern3d_points():
    p = S.TLRProblem("spatial", 131072, 3, 2048, symm="S", seed=2,
                     place=S.STARSH_PARTICLES_RAND, kernel="matern",
                     beta=0.1, nu=1.5)
    _check(_shuffled(p.points(), 12), "C3 points shuffled")


@pytest.mark.parametrize("ndim", [1, 2, 3, 4])
def test_zsort_duplicates_and_dimensions(ndim):
    rng = np.random.default_rng(20 + ndim)
    grid = rng.integers(0, 6, (ndim, 5000)).astype(np.float64) / 5.0
    _check(grid, f"{ndim}D coarse grid with duplicates")
    _check(rng.random((ndim, 3001)), f"{ndim}D uniform, ragged tile")


def test_zsort_from_binary_particle_file(tmp_path):
    """Binary particle file -> starsh_particles_read_from_file ->
    starsh_particles_zsort_inplace_hip, against the oracle's sort of the
    same points."""
    L = S.lib()
    rng = np.random.default_rng(31)
    pts = rng.random((3, 20000))
    pts[:, 5000:6000] = pts[:, :1000]          # exact duplicates
    libc = C.CDLL(None)
    libc.malloc.restype = C.c_void_p
    buf = libc.malloc(C.c_size_t(8 * pts.size))
    C.memmove(buf, np.ascontiguousarray(pts).ctypes.data, 8 * pts.size)
    a = C.POINTER(S.Particles)()
    assert L.starsh_particles_init(C.byref(a), S.STARSH_int(20000), 3,
                                   C.c_void_p(buf)) == 0
    path = str(tmp_path / "pts.bin").encode()
    assert L.starsh_particles_write_to_file(a, path, 2) == 0
    L.starsh_particles_free(a)
    b = C.POINTER(S.Particles)()
    assert L.starsh_particles_read_from_file(C.byref(b), path, 2) == 0
    assert b.contents.count == 20000 and b.contents.ndim == 3
    _say("case: zsort from a binary particle file, 3D n 20000")
    assert L.starsh_particles_zsort_inplace_hip(b) == 0
    got = np.ctypeslib.as_array(b.contents.point, shape=(60000,)).reshape(
        3, 20000).copy()
    L.starsh_particles_free(b)
    want, _ = O.zsort(pts)
    assert np.array_equal(got, want)
