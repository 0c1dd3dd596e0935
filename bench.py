"""Benchmark: seconds to TLR-compress N=64,000 2D squared-exponential
(beta=0.1, jittered 253x253 grid truncated to 64,000 points, seed 1,
Morton-sorted), tile 1600, maxrank 400, oversample 16, eps=1e-7, symmetric
format (820 tiles) -- plain FP64 HIP form and batched mixed-precision form
on one MI355X, through starsh_blrm_approximate.

A step is one full compression of that matrix.  Inputs are resident on the
GPU during the timed device region; `value` is the device time of the
headline form (mixed when available, otherwise plain).  The JSON line also
carries the API wall time (incl. factor copy-back), the dominant kernel's
roofline, and a CPU baseline: the oracle (CPU restatement of the
reference's OpenMP RSDD) timed on a bounded sample of tiles on the host
cores and scaled to the 820 tiles.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "stars-h_amd"), ROOT]

N, NB, MAXRANK, OVERSAMPLE, TOL, BETA, SEED = 64000, 1600, 400, 16, 1e-7, \
    0.1, 1
PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 (vector and matrix)


def run_form(S, backend, steps, warmup, autotune):
    fn = S.init(backend, oversample=OVERSAMPLE)
    p = S.TLRProblem("spatial", N, 2, NB, symm="S", seed=SEED,
                     kernel="sqrexp", beta=BETA)
    schedule = None
    if backend == "HIP_MIXED" and autotune:
        # plan step (outside the timed steps): first rank bucket by trial
        first, secs = p.autotune(MAXRANK, TOL)
        schedule = {"first_bucket_width": [32, 64, 128, 256, 512][first],
                    "candidate_device_seconds": [round(x, 6) for x in secs]}
    dev, wall, dom, launches = [], [], 0.0, 0
    for it in range(warmup + steps):
        info = p.approximate(MAXRANK, TOL, 0, fn)
        assert info == 0, info
        st = S.stats()
        if it >= warmup:
            dev.append(st.device_seconds)
            wall.append(st.wall_seconds)
            dom += st.dominant_kernel_seconds
            launches += st.dominant_kernel_launches
        if it < warmup + steps - 1:
            p.free_matrix()
    err, terr, tnorm = p.dfe_tiles()
    ranks = p.ranks()
    res = {
        "device_seconds": sum(dev) / len(dev),
        "api_wall_seconds": sum(wall) / len(wall),
        "dominant_kernel_seconds": dom / max(launches, 1),
        "sketch_flops": st.sketch_bytes_or_flops,
        "algorithmic_flops": st.flops_algorithmic,
        "tiles": int(st.ntiles),
        "false_far": int(st.nfalse_far),
        "rank_min": int(ranks.min()), "rank_max": int(ranks.max()),
        "rank_avg": float(ranks.mean()),
        "global_rel_error": err,
        "max_tile_rel_error": float(terr.max()),
        "points": p.points(),
        "cluster": p.cluster_ranges(),
        "tiles_list": p.tiles(),
    }
    if schedule is not None:
        res["schedule"] = schedule
    p.free_matrix()
    return res


def cpu_baseline(points, cluster, tiles, nsample):
    from oracle import oracle as O
    nthreads = min(16, os.cpu_count() or 1)
    kd = O.kernel("sqrexp", 2, beta=BETA)
    step = max(1, len(tiles) // nsample)
    sample = tiles[::step][:nsample]
    t0 = time.perf_counter()
    O.compress(kd, points, cluster[0], cluster[1], sample, MAXRANK,
               OVERSAMPLE, TOL, nthreads=nthreads, want_factors=False)
    dt = time.perf_counter() - t0
    return {"value": dt * len(tiles) / len(sample), "unit": "s",
            "cores": nthreads, "kind": "port",
            "sample": f"{len(sample)} of {len(tiles)} tiles (every {step}th)"
                      f" compressed in {dt:.2f} s with {nthreads} OpenMP "
                      f"threads, MKL sequential; scaled to all tiles",
            "lapack": O.lib().oracle_lapack_name().decode()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="plain,mixed")
    ap.add_argument("--no-autotune", action="store_true")
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
    for kv in args.set:
        name, val = kv.split("=")
        assert S.lib().starsh_hip_set_option(
            name.encode(), ctypes.c_double(float(val))) == 0, kv
    if args.first_bucket is not None:
        assert S.lib().starsh_hip_set_option(
            b"mixed_first_bucket", ctypes.c_double(args.first_bucket)) == 0
    forms = {}
    for form in args.forms.split(","):
        backend = {"plain": "HIP", "mixed": "HIP_MIXED"}[form]
        try:
            print(f"bench: {form} form, {args.warmup} warmup + {args.steps} "
                  f"steps", file=sys.stderr, flush=True)
            forms[form] = run_form(S, backend, args.steps, args.warmup,
                                   not args.no_autotune)
        except AssertionError as e:
            print(f"bench: {form} form unavailable ({e})", file=sys.stderr)
    plain = forms.get("plain")
    head = forms.get("mixed") or plain
    head_name = "mixed" if "mixed" in forms else "plain"
    cpu = None
    if not args.no_cpu and plain is not None:
        print("bench: CPU baseline sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(plain["points"], plain["cluster"],
                           plain["tiles_list"], args.cpu_sample)
    dom = plain or head
    dom_t = dom["dominant_kernel_seconds"]
    achieved = dom["sketch_flops"] / dom_t / 1e12 if dom_t > 0 else 0.0
    summary = {}
    for name, r in forms.items():
        summary[name] = {k: (float(f"{v:.6g}") if isinstance(v, float)
                             else v)
                         for k, v in r.items()
                         if k not in ("points", "cluster", "tiles_list")}
    line = {
        "metric": "sec to TLR-compress N=64k sqexp at eps=1e-7: FP64 plain "
                  "vs batched mixed-prec",
        "value": round(head["device_seconds"], 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["api_wall_seconds"] * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if head_name == "plain" else "i8+f64",
        "data": "synthetic (seeded jittered grid, glibc rand)",
        "config": {"workload": "C2: 2D sqexp beta=0.1, N=64000, tile 1600, "
                               "maxrank 400, oversample 16, eps 1e-7, "
                               "symm S (820 tiles)",
                   "form": head_name},
        "forms": summary,
        "speedup_mixed_vs_plain": (round(plain["device_seconds"] /
                                         forms["mixed"]["device_seconds"], 3)
                                   if "mixed" in forms and plain
                                   else None),
        "roofline": {"kernel": "sketch GEMM Y=A*Omega (FP64 MFMA, fused "
                               "tile generation), plain form",
                     "bound": "mfma", "achieved": round(achieved, 3),
                     "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64, 4),
                     "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
