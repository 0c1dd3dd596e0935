"""Benchmark: seconds to TLR-compress N=64,000 2D squared-exponential
(beta=0.1, jittered 253x253 grid truncated to 64,000 points, seed 1,
Morton-sorted), tile 1600, maxrank 400, oversample 16, eps=1e-7, symmetric
format (820 tiles) -- plain FP64 HIP form and batched mixed-precision form
on one MI355X, through starsh_blrm_approximate.

A step is one full compression of that matrix.  Inputs are resident on the
GPU during the timed device region; `value` is the device time of the
headline form (mixed when available, otherwise plain): HIP events from the
first kernel to the last.  The JSON line also carries, per form, the wall
time of the starsh_blrm_approximate call (`api_wall_seconds`: argument
checks, point upload, device batches and the PCIe copy of U and V into the
caller's malloc'd buffers) with its host-side timeline, the roofline of the
dominant kernel (the headline form's at the top level), and a CPU baseline:
the oracle (CPU restatement of the reference's OpenMP RSDD) on every tile
of the workload (C2) or a spread sample (others), best of 3, one OpenMP
thread per host core.

--config selects one of the other measured workloads (same JSON layout):
  C3          3D Matern nu=1.5 beta 0.1, N=131072 uniform (srand 2), tile
              2048, maxrank 512, oversample 16, eps 1e-7 (2080 tiles);
  C4-coulomb  3D 1/r on 200 spheres r=0.02 x 500 points (srand 3),
  C4-sin      3D sin(50 r)/r (diag 1.0) on the same points; N=100000,
              tile 1536 (ragged 160), maxrank 384, oversample 16, eps 1e-6;
  C5          48 2D exp problems (beta U[0.05,0.3], N log-uniform in
              [2000, 40000], tile round64(sqrt(400 N)), numpy seed 4)
              through starsh_blrm_approximate_batch, maxrank 384,
              oversample 16, eps 1e-7.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "stars-h_amd"), ROOT]

PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 (vector and matrix)

# Per-launch counter readings of the dominant kernels at C2 (rocprofv3
# --pmc, one step, one run per counter group; commands and raw values in
# profiles/c2_counters.md).  HBM bytes are FETCH_SIZE and WRITE_SIZE (KiB)
# scaled by the same counters on a 1 GiB device copy (fetch x2.00, write
# x1.00).
C2_PLAIN_SKETCH_TRAFFIC = 7.332e9   # gen_gemm_kernel<2,2,256,1,false>
C2_MIXED_PROJ_TRAFFIC = 0.660e9     # proj_i8_kernel<2,2,64,false>
C2_MIXED_PROJ_VALU_INSTS = 1.030e9  # SQ_INSTS_VALU (wave64 instructions)
C2_MIXED_PROJ_VALU_BUSY = 0.69      # VALUBusy at the kernel's clock
C2_MIXED_PROJ_MFMA_UTIL = 0.18      # MfmaUtil (int8 MFMA)
# VALU issue peak of MI355X: 256 CUs x one wave64 instruction per cycle x
# 64 lanes x 2.4 GHz
PEAK_VALU_LANE_INSTS = 256 * 64 * 2.4e9 / 1e12   # T lane-instructions/s

WORKLOADS = {
    "C2": dict(kind="spatial", kernel="sqrexp", N=64000, ndim=2, nb=1600,
               maxrank=400, os=16, tol=1e-7, beta=0.1, nu=0.5, seed=1,
               place="quasiuniform1",
               desc="C2: 2D sqexp beta=0.1, N=64000, tile 1600, maxrank 400, "
                    "oversample 16, eps 1e-7, symm S (820 tiles)",
               metric="sec to TLR-compress N=64k sqexp at eps=1e-7: FP64 "
                      "plain vs batched mixed-prec"),
    "C3": dict(kind="spatial", kernel="matern", N=131072, ndim=3, nb=2048,
               maxrank=512, os=16, tol=1e-7, beta=0.1, nu=1.5, seed=2,
               place="rand",
               desc="C3: 3D Matern nu=1.5 beta=0.1, N=131072 uniform, tile "
                    "2048, maxrank 512, oversample 16, eps 1e-7, symm S "
                    "(2080 tiles)",
               metric="sec to TLR-compress N=131072 3D Matern at eps=1e-7"),
    "C4-coulomb": dict(kind="coulomb", kernel="coulomb", N=100000, ndim=3,
                       nb=1536, maxrank=384, os=16, tol=1e-6, seed=3,
                       place="spheres",
                       desc="C4: 3D Coulomb 1/r, 200 spheres x 500 points, "
                            "tile 1536 (ragged 160), maxrank 384, oversample"
                            " 16, eps 1e-6, symm S (2211 tiles)",
                       metric="sec to TLR-compress N=100000 Coulomb at "
                              "eps=1e-6"),
    "C4-sin": dict(kind="sin", kernel="sin", N=100000, ndim=3, nb=1536,
                   maxrank=384, os=16, tol=1e-6, seed=3, place="spheres",
                   wave_k=50.0, diag=1.0,
                   desc="C4: 3D sin(50r)/r, 200 spheres x 500 points, tile "
                        "1536 (ragged 160), maxrank 384, oversample 16, eps "
                        "1e-6, symm S (2211 tiles)",
                   metric="sec to TLR-compress N=100000 sin(kr)/r at "
                          "eps=1e-6"),
    "C5": dict(kind="batch", kernel="exp", ndim=2, maxrank=384, os=16,
               tol=1e-7, seed=4,
               desc="C5: 48 2D exp problems, beta U[0.05,0.3], N log-uniform"
                    " in [2000, 40000], tile round64(sqrt(400N)), maxrank "
                    "384, oversample 16, eps 1e-7, one batch call",
               metric="sec to TLR-compress 48 2D exp problems in one batch "
                      "call at eps=1e-7"),
}


def c5_specs(seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(48):
        beta = float(rng.uniform(0.05, 0.3))
        N = int(round(math.exp(rng.uniform(math.log(2000),
                                           math.log(40000)))))
        out.append((N, int(64 * round(math.sqrt(400 * N) / 64)), beta))
    return out


def make_problems(S, w):
    """TLRProblem objects of a workload (a list; one except for C5)."""
    if w["kind"] == "batch":
        return [S.TLRProblem("spatial", N, 2, nb, symm="S", seed=100 + q,
                             kernel="exp", beta=beta)
                for q, (N, nb, beta) in enumerate(c5_specs(w["seed"]))]
    if w["place"] == "spheres":
        pts = S.spheres(200, 500, 0.02, w["seed"])
        return [S.TLRProblem(w["kind"], pts.shape[1], 3, w["nb"], symm="S",
                             points=pts, wave_k=w.get("wave_k", 1.0),
                             diag=w.get("diag", 1.0))]
    place = {"quasiuniform1": S.STARSH_PARTICLES_QUASIUNIFORM1,
             "rand": S.STARSH_PARTICLES_RAND}[w["place"]]
    return [S.TLRProblem("spatial", w["N"], w["ndim"], w["nb"], symm="S",
                         seed=w["seed"], place=place, kernel=w["kernel"],
                         beta=w["beta"], nu=w["nu"])]


def compress(S, ps, w, fn):
    if w["kind"] != "batch":
        return ps[0].approximate(w["maxrank"], w["tol"], 0, fn)
    return S.approximate_batch(ps, w["maxrank"], w["tol"])


def _set(S, name, value):
    assert S.lib().starsh_hip_set_option(name.encode(),
                                         ctypes.c_double(value)) == 0, name


def _timed(S, w, fn, steps):
    """Device seconds of `steps` compressions of fresh problems."""
    out = []
    for _ in range(steps):
        ps = make_problems(S, w)
        assert compress(S, ps, w, fn) == 0
        out.append(S.stats().device_seconds)
        for p in ps:
            p.free_matrix()
    return out


def run_form(S, w, backend, steps, warmup, autotune, lrengine="RSVD"):
    fn = S.init(backend, oversample=w["os"], lrengine=lrengine)
    mixed = backend == "HIP_MIXED" and lrengine == "RSVD"
    schedule = None
    if autotune and lrengine == "RSVD":
        # plan step (outside the timed steps): knobs by trial
        # (starsh_hip_autotune_batch; all problems of C5 as one batch)
        t = S.autotune_batch(make_problems(S, w), w["maxrank"], w["tol"],
                             form=1 if mixed else 0)
        schedule = {k: (round(v, 6) if isinstance(v, float) else
                        [round(x, 6) for x in v] if isinstance(v, list)
                        else v) for k, v in t.items()}
        if mixed:
            schedule["first_bucket_width"] = (
                [32, 64, 128, 256, 384, 512] if t["ladder"] else
                [32, 64, 128, 256, 512])[t["first_bucket"]]
    if mixed:
        # the timed steps do not replay the per-tile outcomes recorded on
        # this matrix by the plan step (reported apart below)
        _set(S, "tile_schedule", 0)
    ps = make_problems(S, w)
    dev, wall, dom, proj = [], [], [], []
    parts = ("prepare", "batch", "d2h", "scatter", "finish")
    tl = {k: [] for k in parts}
    for it in range(warmup + steps):
        if it > 0:
            # a compression rewrites the format when it finds false
            # far-field tiles; every step starts from a fresh format
            ps = make_problems(S, w)
        info = compress(S, ps, w, fn)
        assert info == 0, info
        st = S.stats()
        if it >= warmup:
            dev.append(st.device_seconds)
            wall.append(st.wall_seconds)
            dom.append(st.dominant_kernel_seconds)
            proj.append(st.proj_kernel_seconds)
            for k in parts:
                tl[k].append(getattr(st, k + "_seconds"))
        if it < warmup + steps - 1:
            for p in ps:
                p.free_matrix()
    errs, tmax, ranks = [], 0.0, []
    for p in ps:
        e, terr, _ = p.dfe_tiles()
        errs.append(e)
        if len(terr):
            tmax = max(tmax, float(terr.max()))
        ranks.extend(p.ranks().tolist())
    t = sum(dev) / len(dev)
    res = {
        "device_seconds": t,
        "step_device_seconds": [round(x, 6) for x in dev],
        "api_wall_seconds": sum(wall) / len(wall),
        "step_api_wall_seconds": [round(x, 6) for x in wall],
        # host-side timeline of the API call, mean over the timed steps:
        # argument checks + point upload + U/V malloc, device batches (wall),
        # pack + D2H of U/V, copy into the caller's buffers, format rewrite
        # + near tiles + blrm
        "api_timeline_seconds": {k: round(sum(v) / len(v), 6)
                                 for k, v in tl.items()},
        # summed over the launches of one step
        "dominant_kernel_seconds": sum(dom) / len(dom),
        "dominant_kernel_launches": int(st.dominant_kernel_launches),
        # summed over the launches of one step, HIP events
        "sketch_kernel_seconds": sum(dom) / len(dom),
        "proj_kernel_seconds": sum(proj) / len(proj),
        "sketch_flops": st.sketch_bytes_or_flops,
        "algorithmic_flops": st.flops_algorithmic,
        "tiles": int(st.ntiles),
        "tiles_per_second": st.ntiles / t if t > 0 else 0.0,
        "false_far": int(st.nfalse_far),
        "early_dense": int(st.early_dense),
        "tile_rounds": int(st.tile_rounds),
        "rank_min": int(min(ranks)) if ranks else 0,
        "rank_max": int(max(ranks)) if ranks else 0,
        "rank_avg": float(sum(ranks) / len(ranks)) if ranks else 0.0,
        "global_rel_error": max(errs),
        "max_tile_rel_error": tmax,
        "problems": ps,
    }
    if schedule is not None:
        res["schedule"] = schedule
    if mixed:
        _set(S, "tile_schedule", 1)
        if schedule is not None and schedule["tile_schedule"]:
            # recompression of the same matrix replaying the per-tile
            # schedule of the plan step (each tile starts in the bucket it
            # finished in; tiles found false far-field are stored densely
            # at once): the case of a repeated compression of one problem
            _timed(S, w, fn, 1)
            rs = _timed(S, w, fn, steps)
            res["tile_schedule_replay_device_seconds"] = sum(rs) / len(rs)
            res["step_tile_schedule_replay_device_seconds"] = [
                round(x, 6) for x in rs]
    return res


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, p, nsample, repeats):
    """The oracle (OpenMP restatement of the reference tile loop:
    openmp/blrm/drsdd.c + sequential/dense/drsdd.c with the host LAPACK in
    sequential mode) on every tile of the workload (nsample == 0) or on a
    spread sample scaled to all tiles; one warm-up on a few tiles, then the
    best of `repeats` timed runs, one OpenMP thread per host core."""
    from oracle import oracle as O
    nthreads = os.cpu_count() or 1
    kd = O.kernel(w["kernel"], w["ndim"], beta=w.get("beta", 0.1),
                  nu=w.get("nu", 0.5), wave_k=w.get("wave_k", 1.0),
                  diag=w.get("diag", 1.0))
    tiles = p.tiles()
    if nsample <= 0 or nsample >= len(tiles):
        sample, step = tiles, 1
    else:
        step = max(1, len(tiles) // nsample)
        sample = tiles[::step][:nsample]
    st, sz = p.cluster_ranges()
    pts = p.points()
    O.compress(kd, pts, st, sz, sample[:nthreads], w["maxrank"], w["os"],
               w["tol"], nthreads=nthreads, want_factors=False)
    times = []
    for _ in range(max(1, repeats)):
        t0 = time.perf_counter()
        O.compress(kd, pts, st, sz, sample, w["maxrank"], w["os"],
                   w["tol"], nthreads=nthreads, want_factors=False)
        times.append(time.perf_counter() - t0)
    dt = min(times)
    scaled = len(sample) < len(tiles)
    what = (f"all {len(tiles)} tiles" if not scaled else
            f"{len(sample)} of {len(tiles)} tiles (every {step}th), scaled "
            f"to all tiles")
    return {"value": round(dt * len(tiles) / len(sample), 4), "unit": "s",
            "cores": nthreads, "kind": "port",
            "sample": f"{what}; best of {len(times)} runs "
                      f"({', '.join(f'{x:.2f}' for x in times)} s) after a "
                      f"warm-up, {nthreads} OpenMP threads (one per host "
                      f"core), LAPACK sequential",
            "cpu_model": _cpu_model(),
            "lapack": O.lib().oracle_lapack_name().decode()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="C2", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="tiles of the CPU baseline (0: all; default all "
                         "for C2, 256 otherwise)")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="plain,mixed",
                    help="comma-separated: plain, mixed (randomized SVD), "
                         "sdd, rrqr (exact engines, STARSH_LRENGINE=SVD / "
                         "RRQR)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the FP64 rate probe and the product timing "
                         "(short runs under a profiler)")
    ap.add_argument("--first-bucket", type=int, default=None,
                    help="force the mixed form's first rank bucket (index)")
    ap.add_argument("--set", action="append", default=[],
                    metavar="NAME=VALUE",
                    help="starsh_hip_set_option before the run")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB "
                         "(reference kernel for HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: the compression path runs on a "
              "single GPU in this build (multi-GPU sharding is not built); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n}))
        return
    import starsh_hip as S
    w = WORKLOADS[args.config]
    for kv in args.set:
        name, val = kv.split("=")
        assert S.lib().starsh_hip_set_option(
            name.encode(), ctypes.c_double(float(val))) == 0, kv
    if args.first_bucket is not None:
        assert S.lib().starsh_hip_set_option(
            b"mixed_first_bucket", ctypes.c_double(args.first_bucket)) == 0
    forms = {}
    for form in args.forms.split(","):
        backend, lre = {"plain": ("HIP", "RSVD"),
                        "mixed": ("HIP_MIXED", "RSVD"),
                        "sdd": ("HIP", "SVD"),
                        "rrqr": ("HIP", "RRQR")}[form]
        print(f"bench: {args.config} {form} form, {args.warmup} warmup + "
              f"{args.steps} steps", file=sys.stderr, flush=True)
        forms[form] = run_form(S, w, backend, args.steps, args.warmup,
                               not args.no_autotune, lre)
    S.init("HIP", lrengine="RSVD")
    plain = forms.get("plain")
    head_name = ("mixed" if "mixed" in forms else
                 "plain" if "plain" in forms else next(iter(forms)))
    head = forms[head_name]
    cpu = None
    if not args.no_cpu and plain is not None and w["kind"] != "batch":
        print("bench: CPU baseline sample", file=sys.stderr, flush=True)
        ns = args.cpu_sample
        if ns is None:
            ns = 0 if args.config == "C2" else 256
        cpu = cpu_baseline(w, plain["problems"][0], ns, args.cpu_repeats)
    # attainable FP64 rates of this GPU (MFMA chains, vector FMA)
    mf, vf = ctypes.c_double(0.0), ctypes.c_double(0.0)
    probe = None
    if not args.no_extras and S.lib().starsh_hip_fp64_probe(
            ctypes.byref(mf), ctypes.byref(vf)) == 0:
        probe = {"mfma": round(mf.value, 2), "valu": round(vf.value, 2)}
    # product with the compressed matrix (starsh_hip_tlr_*, factors
    # resident in HBM): one right-hand side, 20 products
    matvec = None
    if w["kind"] != "batch" and not args.no_extras:
        sec, nbytes = head["problems"][0].tlr_bench(1, 20)
        if sec > 0:
            matvec = {"form": head_name, "nrhs": 1,
                      "seconds": round(sec, 7),
                      "bytes": nbytes,
                      "GB_per_s": round(nbytes / sec / 1e9, 1),
                      "frac_of_hbm_peak": round(nbytes / sec / 8e12, 4)}
    c2 = args.config == "C2"
    rooflines = {}
    if plain is not None:
        # plain form: the fused sketch GEMM on FP64 MFMA; algorithmic
        # flops 2 m n k per tile, time from HIP events around the launch
        t = plain["sketch_kernel_seconds"]
        a = plain["sketch_flops"] / t / 1e12 if t > 0 else 0.0
        rooflines["plain"] = {
            "kernel": "gen_gemm_kernel (sketch Y = A Omega, FP64 MFMA, "
                      "tile generated in registers)",
            "bound": "mfma", "achieved": round(a, 3), "peak": PEAK_FP64,
            "unit": "TFLOP/s", "frac": round(a / PEAK_FP64, 4),
            "frac_of_attainable": (round(a / probe["mfma"], 4)
                                   if probe else None),
            "traffic": C2_PLAIN_SKETCH_TRAFFIC if c2 else None}
    if "mixed" in forms:
        # mixed form: the int8 projection B^T = A^T Q dominates; it is
        # bound by VALU issue (FP64 tile generation and digit splitting),
        # not by MFMA or HBM: achieved = VALU lane-instructions per second
        # (counter reading per launch / launch time from HIP events)
        mx = forms["mixed"]
        t = mx["proj_kernel_seconds"] / max(1, mx["dominant_kernel_launches"])
        a = (C2_MIXED_PROJ_VALU_INSTS * 64 / t / 1e12 if t > 0 and c2
             else None)
        rooflines["mixed"] = {
            "kernel": "proj_i8_kernel (projection B^T = A^T Q: FP64 tile "
                      "generation, int8 digit products on MFMA)",
            "bound": "valu", "achieved": round(a, 3) if a else None,
            "peak": round(PEAK_VALU_LANE_INSTS, 2),
            "unit": "T VALU lane-instructions/s",
            "frac": round(a / PEAK_VALU_LANE_INSTS, 4) if a else None,
            "valu_busy_counter": C2_MIXED_PROJ_VALU_BUSY if c2 else None,
            "mfma_util_counter": C2_MIXED_PROJ_MFMA_UTIL if c2 else None,
            "launch_seconds": round(t, 7),
            "traffic": C2_MIXED_PROJ_TRAFFIC if c2 else None}
    summary = {}
    for name, r in forms.items():
        summary[name] = {k: (float(f"{v:.6g}") if isinstance(v, float)
                             else v)
                         for k, v in r.items() if k != "problems"}
        if name in rooflines:
            summary[name]["roofline"] = rooflines[name]
    line = {
        "metric": w["metric"],
        "value": round(head["device_seconds"], 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["api_wall_seconds"] * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "i8+f64" if head_name == "mixed" else "f64",
        "data": "synthetic (seeded point sets, glibc rand / numpy)",
        "config": {"workload": w["desc"], "form": head_name},
        "forms": summary,
        "speedup_mixed_vs_plain": (round(plain["device_seconds"] /
                                         forms["mixed"]["device_seconds"], 3)
                                   if "mixed" in forms and plain
                                   else None),
        # the headline form's dominant kernel
        "roofline": rooflines.get(head_name),
        "api_wall_speedup_mixed_vs_plain": (
            round(plain["api_wall_seconds"] /
                  forms["mixed"]["api_wall_seconds"], 3)
            if "mixed" in forms and plain else None),
        "cpu_baseline": cpu,
        "fp64_attainable_tflops": probe,
        "tlr_matvec": matvec,
    }
    for r in forms.values():
        for p in r["problems"]:
            p.free_matrix()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
