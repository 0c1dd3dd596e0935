"""Python mirror of the STARS-H C API served by libstarsh_hip.so.

The library is the product: a C ABI (include/starsh*.h) whose
starsh_blrm_approximate is the HIP tile low-rank compressor.  This module
only loads it with ctypes and mirrors the reference's call sequence
(testing/spatial.c:55-109 in ecrc/stars-h 0.3.1): starsh_init,
starsh_application, starsh_problem_new, starsh_cluster_new_plain,
starsh_blrf_new_tlr, starsh_blrm_approximate, starsh_blrm__dfe.

Loading fails loudly when the shared library has not been built: there is
no fallback implementation.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstarsh_hip.so")

# enum values (include/starsh-constants.h and application headers)
STARSH_SUCCESS = 0
STARSH_MALLOC_ERROR = 1
STARSH_WRONG_PARAMETER = 4
STARSH_SPATIAL = 3
STARSH_ELECTROSTATICS = 5
STARSH_ELECTRODYNAMICS = 6
STARSH_SPATIAL_EXP, STARSH_SPATIAL_SQREXP = 1, 2
STARSH_SPATIAL_MATERN, STARSH_SPATIAL_MATERN2 = 3, 4
STARSH_SPATIAL_EXP_SIMD, STARSH_SPATIAL_SQREXP_SIMD = 11, 12
STARSH_SPATIAL_NDIM, STARSH_SPATIAL_BETA, STARSH_SPATIAL_NU = 1, 2, 3
STARSH_SPATIAL_NOISE, STARSH_SPATIAL_PLACE, STARSH_SPATIAL_SIGMA = 4, 5, 6
STARSH_ELECTROSTATICS_COULOMB_POTENTIAL = 1
STARSH_ELECTRODYNAMICS_SIN, STARSH_ELECTRODYNAMICS_COS = 1, 2
STARSH_PARTICLES_RAND, STARSH_PARTICLES_UNIFORM = 1, 3
STARSH_PARTICLES_QUASIUNIFORM1 = 6
FAMILY = {"exp": 1, "sqrexp": 2, "matern": 3, "matern2": 4, "coulomb": 5,
          "sin": 6, "cos": 7}

STARSH_int = C.c_ssize_t


class Array(C.Structure):
    _fields_ = [("ndim", C.c_int), ("shape", C.POINTER(C.c_int)),
                ("stride", C.POINTER(C.c_ssize_t)), ("order", C.c_char),
                ("size", C.c_size_t), ("dtype", C.c_char),
                ("dtype_size", C.c_size_t), ("nbytes", C.c_size_t),
                ("data_nbytes", C.c_size_t), ("data", C.c_void_p)]


class Particles(C.Structure):
    _fields_ = [("count", STARSH_int), ("ndim", C.c_int),
                ("point", C.POINTER(C.c_double))]


class SSData(C.Structure):
    _fields_ = [("particles", Particles), ("beta", C.c_double),
                ("nu", C.c_double), ("noise", C.c_double),
                ("sigma", C.c_double), ("nu2", C.c_double),
                ("sigma2", C.c_double), ("corr", C.c_double)]


class EDData(C.Structure):
    _fields_ = [("particles", Particles), ("k", C.c_double),
                ("diag", C.c_double)]


class Cluster(C.Structure):
    _fields_ = [("data", C.c_void_p), ("ndata", STARSH_int),
                ("pivot", C.POINTER(STARSH_int)), ("nblocks", STARSH_int),
                ("nlevels", STARSH_int), ("level", C.POINTER(STARSH_int)),
                ("start", C.POINTER(STARSH_int)),
                ("size", C.POINTER(STARSH_int)),
                ("parent", C.POINTER(STARSH_int)),
                ("child_start", C.POINTER(STARSH_int)),
                ("child", C.POINTER(STARSH_int)), ("type", C.c_int)]


class Problem(C.Structure):
    _fields_ = [("ndim", C.c_int), ("shape", C.POINTER(STARSH_int)),
                ("symm", C.c_char), ("dtype", C.c_char),
                ("dtype_size", C.c_size_t), ("entry_size", C.c_size_t),
                ("row_data", C.c_void_p), ("col_data", C.c_void_p),
                ("kernel", C.c_void_p), ("name", C.c_char_p)]


_P = C.POINTER(STARSH_int)


class BLRF(C.Structure):
    _fields_ = [("problem", C.POINTER(Problem)), ("symm", C.c_char),
                ("row_cluster", C.POINTER(Cluster)),
                ("col_cluster", C.POINTER(Cluster)),
                ("nbrows", STARSH_int), ("nbcols", STARSH_int),
                ("nblocks_far", STARSH_int), ("block_far", _P),
                ("brow_far_start", _P), ("brow_far", _P),
                ("bcol_far_start", _P), ("bcol_far", _P),
                ("nblocks_far_local", STARSH_int), ("block_far_local", _P),
                ("nblocks_near", STARSH_int), ("block_near", _P),
                ("brow_near_start", _P), ("brow_near", _P),
                ("bcol_near_start", _P), ("bcol_near", _P),
                ("nblocks_near_local", STARSH_int), ("block_near_local", _P),
                ("type", C.c_int)]


class BLRM(C.Structure):
    _fields_ = [("format", C.POINTER(BLRF)), ("far_rank", C.POINTER(C.c_int)),
                ("far_U", C.POINTER(C.POINTER(Array))),
                ("far_V", C.POINTER(C.POINTER(Array))), ("onfly", C.c_int),
                ("near_D", C.POINTER(C.POINTER(Array))),
                ("alloc_U", C.c_void_p), ("alloc_V", C.c_void_p),
                ("alloc_D", C.c_void_p), ("alloc_type", C.c_char),
                ("nbytes", C.c_size_t), ("data_nbytes", C.c_size_t)]


class Stats(C.Structure):
    _fields_ = [("device_seconds", C.c_double), ("wall_seconds", C.c_double),
                ("copy_seconds", C.c_double),
                ("dominant_kernel_seconds", C.c_double),
                ("dominant_kernel_launches", C.c_longlong),
                ("ntiles", C.c_longlong), ("nfalse_far", C.c_longlong),
                ("rank_min", C.c_int), ("rank_max", C.c_int),
                ("rank_avg", C.c_double), ("flops_algorithmic", C.c_double),
                ("sketch_bytes_or_flops", C.c_double),
                ("prepare_seconds", C.c_double),
                ("batch_seconds", C.c_double),
                ("d2h_seconds", C.c_double),
                ("scatter_seconds", C.c_double),
                ("finish_seconds", C.c_double),
                ("proj_kernel_seconds", C.c_double),
                ("early_dense", C.c_longlong),
                ("tile_rounds", C.c_longlong)]


class KernelDesc(C.Structure):
    _fields_ = [("family", C.c_int), ("ndim", C.c_int), ("beta", C.c_double),
                ("nu", C.c_double), ("noise", C.c_double),
                ("sigma", C.c_double), ("wave_k", C.c_double),
                ("diag", C.c_double)]


_lib = None


def lib():
    """The loaded product library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libstarsh_hip.so is not built; run "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        L.starsh_blrm__dfe_hip.restype = C.c_double
        L.starsh_blrm__dfe_hip_tiles.restype = C.c_double
        L.starsh_dense_dsvfr.argtypes = [C.c_int, C.POINTER(C.c_double),
                                         C.c_double]
        dp = C.POINTER(C.c_double)
        L.starsh_blrm__dmml_hip.argtypes = [C.c_void_p, C.c_int, C.c_double,
                                            dp, C.c_int, C.c_double, dp,
                                            C.c_int]
        L.starsh_itersolvers__dcg_hip.argtypes = [C.c_void_p, C.c_int, dp,
                                                  C.c_int, dp, C.c_int,
                                                  C.c_double, dp]
        L.starsh_hip_tlr_new.argtypes = [C.POINTER(C.c_void_p), C.c_void_p]
        L.starsh_hip_tlr_bench.argtypes = [C.c_void_p, C.c_int, C.c_int, dp,
                                           dp]
        L.starsh_hip_tlr_free.argtypes = [C.c_void_p]
        L.starsh_hip_tlr_free.restype = None
        _lib = L
    return _lib


def approximate_fn():
    """Current value of the starsh_blrm_approximate function pointer."""
    L = lib()
    fp = C.c_void_p.in_dll(L, "starsh_blrm_approximate").value
    if not fp:
        return None
    proto = C.CFUNCTYPE(C.c_int, C.POINTER(C.POINTER(BLRM)), C.POINTER(BLRF),
                        C.c_int, C.c_double, C.c_int)
    return proto(fp)


def init(backend="HIP", oversample=None, lrengine="RSVD"):
    L = lib()
    assert L.starsh_set_backend(backend.encode()) == 0, backend
    assert L.starsh_set_lrengine(lrengine.encode()) == 0
    if oversample is not None:
        assert L.starsh_set_oversample(str(int(oversample)).encode()) == 0
    return approximate_fn()


def stats():
    s = Stats()
    lib().starsh_hip_get_stats(C.byref(s))
    return s


def srand(seed):
    C.CDLL(None).srand(C.c_uint(seed))


def spheres(nspheres, npoints, radius, seed):
    """Points of starsh_particles_generate_spheres after srand(seed), as a
    3 x N array (SoA)."""
    L = lib()
    srand(seed)
    data = C.POINTER(Particles)()
    assert L.starsh_particles_generate_spheres(
        C.byref(data), STARSH_int(nspheres), STARSH_int(npoints),
        C.c_double(radius)) == 0
    part = data.contents
    n = part.count * part.ndim
    pts = np.ctypeslib.as_array(part.point, shape=(n,)).reshape(
        part.ndim, part.count).copy()
    L.starsh_particles_free(data)
    return pts


def approximate_batch(problems, maxrank, tol, onfly=0):
    """starsh_blrm_approximate_batch over TLRProblem objects; each
    problem's matrix is stored in its .M.  maxrank / tol: scalars or one
    value per problem."""
    n = len(problems)
    mr = list(maxrank) if hasattr(maxrank, "__len__") else [maxrank] * n
    tl = list(tol) if hasattr(tol, "__len__") else [tol] * n
    Ms = (C.POINTER(BLRM) * n)()
    Fs = (C.POINTER(BLRF) * n)(*[p.F for p in problems])
    info = lib().starsh_blrm_approximate_batch(
        Ms, Fs, C.c_int(n), (C.c_int * n)(*mr), (C.c_double * n)(*tl),
        C.c_int(onfly))
    for i, p in enumerate(problems):
        p.M = Ms[i] if info == 0 else None
    return info


class Tuning(C.Structure):
    """starsh_hip_tuning."""
    _fields_ = [("form", C.c_int), ("first_bucket", C.c_int),
                ("tile_schedule", C.c_int), ("qr_panel", C.c_int),
                ("sketch_bn", C.c_int), ("sketch_rw", C.c_int),
                ("proj_bn", C.c_int), ("proj_rw", C.c_int),
                ("ntrials", C.c_int),
                ("first_seconds", C.c_double * 5),
                ("schedule_seconds", C.c_double),
                ("untuned_seconds", C.c_double),
                ("best_seconds", C.c_double),
                ("ladder", C.c_int),
                ("ladder_seconds", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["first_seconds"] = [x for x in self.first_seconds]
        return d


def autotune_batch(problems, maxrank, tol, form=1, onfly=0):
    """starsh_hip_autotune_batch over TLRProblem objects (form 0 plain,
    1 mixed); returns the Tuning record as a dict."""
    n = len(problems)
    mr = list(maxrank) if hasattr(maxrank, "__len__") else [maxrank] * n
    tl = list(tol) if hasattr(tol, "__len__") else [tol] * n
    Fs = (C.POINTER(BLRF) * n)(*[p.F for p in problems])
    out = Tuning()
    info = lib().starsh_hip_autotune_batch(
        Fs, C.c_int(n), (C.c_int * n)(*mr), (C.c_double * n)(*tl),
        C.c_int(onfly), C.c_int(form), C.byref(out))
    assert info == 0, info
    return out.as_dict()


class TLRProblem:
    """A problem built through the C API, ready for compression.

    kind: "spatial" (kernel in exp/sqrexp/matern/matern2), "coulomb",
    "sin" or "cos".  Points come from starsh_application's placement
    generator (rand() after srand(seed)) unless `points` (ndim x N, SoA)
    is given, in which case they are Morton-sorted by
    starsh_particles_zsort_inplace like the reference's generators do.
    """

    def __init__(self, kind, N, ndim, block_size, symm="S", seed=0,
                 place=STARSH_PARTICLES_QUASIUNIFORM1, beta=0.1, nu=0.5,
                 noise=0.0, sigma=1.0, wave_k=1.0, diag=1.0, points=None,
                 kernel="sqrexp"):
        L = lib()
        self.L = L
        self.kind = kind
        self.ndim = ndim
        self.data = C.c_void_p()
        self.kernel = C.c_void_p()
        if points is not None:
            pts = np.ascontiguousarray(points, dtype=np.float64)
            assert pts.shape == (ndim, N)
            libc = C.CDLL(None)
            libc.malloc.restype = C.c_void_p
            ptr = libc.malloc(C.c_size_t(pts.nbytes))
            C.memmove(ptr, pts.ctypes.data, pts.nbytes)
            if kind == "spatial":
                info = L.starsh_ssdata_init(C.byref(self.data), STARSH_int(N),
                                            C.c_int(ndim), C.c_void_p(ptr),
                                            C.c_double(beta), C.c_double(nu),
                                            C.c_double(noise),
                                            C.c_double(sigma))
                part = C.cast(self.data, C.POINTER(Particles))
            elif kind == "coulomb":
                info = L.starsh_esdata_init(C.byref(self.data), STARSH_int(N),
                                            C.c_int(ndim), C.c_void_p(ptr))
                part = C.cast(self.data, C.POINTER(Particles))
            else:
                info = L.starsh_eddata_init(C.byref(self.data), STARSH_int(N),
                                            C.c_int(ndim), C.c_void_p(ptr),
                                            C.c_double(wave_k),
                                            C.c_double(diag))
                part = C.cast(self.data, C.POINTER(Particles))
            assert info == 0, info
            assert L.starsh_particles_zsort_inplace(part) == 0
            self._kernel_from_type(kind, kernel)
        else:
            srand(seed)
            if kind == "spatial":
                ktype = {"exp": STARSH_SPATIAL_EXP,
                         "sqrexp": STARSH_SPATIAL_SQREXP,
                         "matern": STARSH_SPATIAL_MATERN,
                         "matern2": STARSH_SPATIAL_MATERN2}[kernel]
                info = L.starsh_application(
                    C.byref(self.data), C.byref(self.kernel), STARSH_int(N),
                    C.c_char(b"d"), C.c_int(STARSH_SPATIAL), C.c_int(ktype),
                    C.c_int(STARSH_SPATIAL_NDIM), C.c_int(ndim),
                    C.c_int(STARSH_SPATIAL_BETA), C.c_double(beta),
                    C.c_int(STARSH_SPATIAL_NU), C.c_double(nu),
                    C.c_int(STARSH_SPATIAL_NOISE), C.c_double(noise),
                    C.c_int(STARSH_SPATIAL_PLACE), C.c_int(place),
                    C.c_int(STARSH_SPATIAL_SIGMA), C.c_double(sigma),
                    C.c_int(0))
            elif kind == "coulomb":
                info = L.starsh_application(
                    C.byref(self.data), C.byref(self.kernel), STARSH_int(N),
                    C.c_char(b"d"), C.c_int(STARSH_ELECTROSTATICS),
                    C.c_int(STARSH_ELECTROSTATICS_COULOMB_POTENTIAL),
                    C.c_int(1), C.c_int(ndim), C.c_int(2), C.c_int(place),
                    C.c_int(0))
            else:
                info = L.starsh_application(
                    C.byref(self.data), C.byref(self.kernel), STARSH_int(N),
                    C.c_char(b"d"), C.c_int(STARSH_ELECTRODYNAMICS),
                    C.c_int(STARSH_ELECTRODYNAMICS_SIN if kind == "sin" else
                            STARSH_ELECTRODYNAMICS_COS),
                    C.c_int(1), C.c_int(ndim), C.c_int(2), C.c_double(wave_k),
                    C.c_int(3), C.c_double(diag), C.c_int(4), C.c_int(place),
                    C.c_int(0))
            assert info == 0, info
        self.N = N
        self.block_size = block_size
        self.symm = symm
        shape = (STARSH_int * 2)(N, N)
        self.P = C.POINTER(Problem)()
        info = L.starsh_problem_new(C.byref(self.P), C.c_int(2), shape,
                                    C.c_char(symm.encode()), C.c_char(b"d"),
                                    self.data, self.data, self.kernel,
                                    C.c_char_p(b"problem"))
        assert info == 0, info
        self.C = C.POINTER(Cluster)()
        info = L.starsh_cluster_new_plain(C.byref(self.C), self.data,
                                          STARSH_int(N),
                                          STARSH_int(block_size))
        assert info == 0, info
        self.F = C.POINTER(BLRF)()
        info = L.starsh_blrf_new_tlr(C.byref(self.F), self.P,
                                     C.c_char(symm.encode()), self.C, self.C)
        assert info == 0, info
        self.M = None

    def _kernel_from_type(self, kind, kernel):
        L = self.L
        if kind == "spatial":
            ktype = {"exp": STARSH_SPATIAL_EXP, "sqrexp": STARSH_SPATIAL_SQREXP,
                     "matern": STARSH_SPATIAL_MATERN,
                     "matern2": STARSH_SPATIAL_MATERN2}[kernel]
            info = L.starsh_ssdata_get_kernel(C.byref(self.kernel), self.data,
                                              C.c_int(ktype))
        elif kind == "coulomb":
            info = L.starsh_esdata_get_kernel(
                C.byref(self.kernel), self.data,
                C.c_int(STARSH_ELECTROSTATICS_COULOMB_POTENTIAL))
        else:
            info = L.starsh_eddata_get_kernel(
                C.byref(self.kernel), self.data,
                C.c_int(STARSH_ELECTRODYNAMICS_SIN if kind == "sin" else
                        STARSH_ELECTRODYNAMICS_COS))
        assert info == 0, info

    # -- views of host data
    def points(self):
        part = C.cast(self.data, C.POINTER(Particles)).contents
        n = part.count * part.ndim
        return np.ctypeslib.as_array(part.point, shape=(n,)).reshape(
            part.ndim, part.count).copy()

    def tiles(self):
        """List of far-field tiles (i, j) of the current format."""
        F = self.F.contents
        return [(F.block_far[2 * b], F.block_far[2 * b + 1])
                for b in range(F.nblocks_far)]

    def near_tiles(self):
        F = self.F.contents
        return [(F.block_near[2 * b], F.block_near[2 * b + 1])
                for b in range(F.nblocks_near)]

    def cluster_ranges(self):
        Cl = self.C.contents
        return ([Cl.start[i] for i in range(Cl.nblocks)],
                [Cl.size[i] for i in range(Cl.nblocks)])

    def approximate(self, maxrank, tol, onfly=0, fn=None):
        fn = fn or approximate_fn()
        self.M = C.POINTER(BLRM)()
        info = fn(C.byref(self.M), self.F, maxrank, tol, onfly)
        return info

    def autotune(self, maxrank, tol, onfly=0):
        """Runs starsh_hip_autotune; returns (first bucket, seconds per
        candidate tried)."""
        first = C.c_int(-1)
        secs = (C.c_double * 5)(*([-1.0] * 5))
        info = self.L.starsh_hip_autotune(self.F, C.c_int(maxrank),
                                          C.c_double(tol), C.c_int(onfly),
                                          C.byref(first), secs)
        assert info == 0, info
        return first.value, [x for x in secs if x >= 0]

    def ranks(self):
        M = self.M.contents
        nfar = M.format.contents.nblocks_far
        return np.array([M.far_rank[b] for b in range(nfar)], dtype=np.int64)

    def factors(self, b):
        """(U, V) of far-field tile b as numpy arrays (copies)."""
        M = self.M.contents
        r = M.far_rank[b]
        U = M.far_U[b].contents
        V = M.far_V[b].contents
        m, n = U.shape[0], V.shape[0]
        u = np.ctypeslib.as_array(C.cast(U.data, C.POINTER(C.c_double)),
                                  shape=(m * max(r, 1),))[:m * r]
        v = np.ctypeslib.as_array(C.cast(V.data, C.POINTER(C.c_double)),
                                  shape=(n * max(r, 1),))[:n * r]
        return (u.reshape(r, m).T.copy(), v.reshape(r, n).T.copy())

    def near_block(self, b):
        M = self.M.contents
        D = M.near_D[b].contents
        m, n = D.shape[0], D.shape[1]
        d = np.ctypeslib.as_array(C.cast(D.data, C.POINTER(C.c_double)),
                                  shape=(m * n,))
        return d.reshape(n, m).T.copy()

    def dfe(self):
        return self.L.starsh_blrm__dfe_hip(self.M)

    def dfe_tiles(self):
        nfar = self.M.contents.format.contents.nblocks_far
        err = np.zeros(max(nfar, 1))
        nrm = np.zeros(max(nfar, 1))
        g = self.L.starsh_blrm__dfe_hip_tiles(
            self.M, err.ctypes.data_as(C.POINTER(C.c_double)),
            nrm.ctypes.data_as(C.POINTER(C.c_double)))
        return g, err[:nfar], nrm[:nfar]

    def host_block(self, i, j):
        """Dense tile (i, j) from the host STARSH_kernel (API path)."""
        shape = (C.c_int * 2)()
        D = C.c_void_p()
        assert self.L.starsh_blrf_get_block(self.F, STARSH_int(i),
                                            STARSH_int(j), shape,
                                            C.byref(D)) == 0
        m, n = shape[0], shape[1]
        arr = np.ctypeslib.as_array(C.cast(D, C.POINTER(C.c_double)),
                                    shape=(m * n,)).reshape(n, m).T.copy()
        C.CDLL(None).free(D)
        return arr

    def dmml(self, A, alpha=1.0, beta=0.0, B=None):
        """starsh_blrm__dmml_hip: returns alpha * M * A + beta * B (A, B:
        rows x nrhs arrays in cluster order)."""
        A = np.asfortranarray(A, dtype=np.float64)
        n, nrhs = A.shape
        B = (np.zeros((n, nrhs), order="F") if B is None else
             np.array(B, dtype=np.float64, order="F"))
        dp = C.POINTER(C.c_double)
        info = self.L.starsh_blrm__dmml_hip(
            C.cast(self.M, C.c_void_p), nrhs, alpha, A.ctypes.data_as(dp), n,
            beta, B.ctypes.data_as(dp), n)
        assert info == 0, info
        return B

    def cg(self, B, X0, tol):
        """starsh_itersolvers__dcg_hip: returns (iteration, X)."""
        B = np.asfortranarray(B, dtype=np.float64)
        X = np.array(X0, dtype=np.float64, order="F")
        n, nrhs = B.shape
        work = np.zeros(3 * n * nrhs + 3 * nrhs)
        dp = C.POINTER(C.c_double)
        it = self.L.starsh_itersolvers__dcg_hip(
            C.cast(self.M, C.c_void_p), nrhs, B.ctypes.data_as(dp), n,
            X.ctypes.data_as(dp), n, tol, work.ctypes.data_as(dp))
        return it, X

    def tlr_bench(self, nrhs=1, reps=10):
        """Seconds per device-resident product with nrhs columns and the
        bytes it must move at least (starsh_hip_tlr_bench)."""
        T = C.c_void_p()
        assert self.L.starsh_hip_tlr_new(C.byref(T),
                                         C.cast(self.M, C.c_void_p)) == 0
        sec, nbytes = C.c_double(0.0), C.c_double(0.0)
        try:
            info = self.L.starsh_hip_tlr_bench(T, nrhs, reps, C.byref(sec),
                                               C.byref(nbytes))
            assert info == 0, info
        finally:
            self.L.starsh_hip_tlr_free(T)
        return sec.value, nbytes.value

    def free_matrix(self):
        if self.M:
            self.L.starsh_blrm_free(self.M)
            self.M = None
