"""Writes the golden fixtures of the oracle tests (run in a container that
has a LAPACK; the fixtures are committed, so tests do not need it).

omega_golden.npz  LAPACK dlarnv(idist, iseed={0,0,0,1}, n) called in the
                  host LAPACK (MKL libmkl_rt or scipy's OpenBLAS): the first
                  1024 normals (idist=3), 300 uniforms (idist=1), the final
                  seed after 1024 normals, and SHA-256 digests of the full
                  test matrices Omega (ncols x k) of the benchmark shapes.
kernel_golden.json  kernel values at chosen distances from closed forms and,
                  for Matern, from scipy.special.kv / gamma (the general
                  2^(1-nu)/Gamma(nu) x^nu K_nu(x) formula the reference
                  evaluates with GSL), for nu = 1/2, 3/2, 5/2 and the
                  general values GENERAL_NU; distances reach scaled
                  arguments x from 1e-5 to about 100.
particles.txt     copy of the reference's own fixture testing/particles.txt
                  (36 3-D points) used by its particles_io test.
"""
import ctypes
import hashlib
import json
import math
import os

import numpy as np
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))
# Matern smoothness values without a closed form (general nu through
# scipy.special.kv, the reference's GSL formula)
GENERAL_NU = (0.3, 0.8, 1.2, 2.2, 3.7, 6.0)


def lapack():
    for path, pre in (("/opt/conda/lib/libmkl_rt.so", ""),):
        if os.path.exists(path):
            L = ctypes.CDLL(path)
            return getattr(L, pre + "dlarnv_"), path
    import scipy
    d = os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs")
    for f in os.listdir(d):
        if f.startswith("libscipy_openblas"):
            L = ctypes.CDLL(os.path.join(d, f))
            return L.scipy_dlarnv_, f
    raise RuntimeError("no LAPACK")


def dlarnv(fn, idist, n):
    iseed = (ctypes.c_int * 4)(0, 0, 0, 1)
    x = np.zeros(n)
    fn(ctypes.byref(ctypes.c_int(idist)), iseed,
       ctypes.byref(ctypes.c_int(n)), x.ctypes.data_as(ctypes.c_void_p))
    return x, list(iseed)


def main():
    fn, name = lapack()
    normals, seed_after = dlarnv(fn, 3, 1024)
    uniforms, _ = dlarnv(fn, 1, 300)
    digests = {}
    for ncols, k in [(1600, 416), (2048, 528), (1536, 400), (160, 400),
                     (250, 110)]:
        x, _ = dlarnv(fn, 3, ncols * k)
        digests[f"{ncols}x{k}"] = hashlib.sha256(x.tobytes()).hexdigest()
    np.savez(os.path.join(HERE, "omega_golden.npz"), normals=normals,
             uniforms=uniforms, seed_after=np.array(seed_after),
             digest_keys=np.array(list(digests.keys())),
             digest_vals=np.array(list(digests.values())),
             lapack=np.array(os.path.basename(name)))
    # kernel values: distance r between (0,..,0) and (r,0,..)
    rs = [0.0, 1e-6, 1e-3, 0.01, 0.05, 0.1, 0.15, 0.2, 0.3, 0.5, 1.0,
          1.4142135623730951]
    out = []
    for r in rs:
        rec = {"r": r}
        beta, sigma, noise = 0.1, 1.3, 0.25
        rec["exp"] = sigma + noise if r == 0 else sigma * math.exp(-r / beta)
        rec["sqrexp"] = sigma + noise if r == 0 else \
            sigma * math.exp(-r * r / (2 * beta * beta))
        for nu in (0.5, 1.5, 2.5) + GENERAL_NU:
            for var, x in (("matern", math.sqrt(2 * nu) * r / beta),
                           ("matern2", r / beta)):
                rec[f"{var}_{nu}"] = sigma + noise if r == 0 else \
                    sigma * 2 ** (1 - nu) / sp.gamma(nu) * x ** nu * \
                    sp.kv(nu, x)
        rec["coulomb"] = 0.0 if r == 0 else 1.0 / r
        kw, diag = 50.0, 0.75
        rec["sin"] = diag if r == 0 else math.sin(kw * r) / r
        rec["cos"] = diag if r == 0 else math.cos(kw * r) / r
        out.append(rec)
    with open(os.path.join(HERE, "kernel_golden.json"), "w") as f:
        json.dump({"beta": 0.1, "sigma": 1.3, "noise": 0.25, "wave_k": 50.0,
                   "diag": 0.75, "values": out}, f, indent=1)


if __name__ == "__main__":
    main()
