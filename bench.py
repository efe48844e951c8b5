"""Benchmark of the 1D k-mer association scan on one MI355X.

Workload (default): n=2048 phenotyped accessions, p=32 U columns, m=1e8 k-mers
of seeded synthetic bits resident in HBM (25.6 GB packed).  One step = one
full scan of the m k-mers (bulk int8-limb MFMA pass + FP64 refinement of the
candidates + host decisions), exactly what kgwas does per batch.

Printed JSON (one line): value = mixed-precision k-mer tests/s; beside it the
plain FP32 GPU kernel, the CPU restatement of the reference's cols path timed
on this host, the bulk kernel's roofline, and every path's statistic error vs
the FP64 oracle on a sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "k-mer tests/sec @n=2048,p=32 mixed vs FP32 GPU vs CPU; stat rel err vs FP64"
HBM_PEAK_GBS = 8000.0
I8_PEAK_TOPS = 5000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--m", type=int, default=100_000_000, help="k-mers per step")
    ap.add_argument("--fp32-m", type=int, default=10_000_000, help="k-mers per FP32 step")
    ap.add_argument("--fp32-steps", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=1_000_000, help="k-mers for the CPU baseline")
    ap.add_argument("--err-sample", type=int, default=50_000, help="k-mers for error vs FP64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--limbs", type=int, default=3)
    ap.add_argument("--profile-only", action="store_true",
                    help="only the mixed scans (for rocprofv3 counter runs)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="copy this many bytes device-to-device and exit (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(k-mer shards are independent, see DESIGN.md for sharding)"}))
        sys.exit(2)

    import oracle
    from supergwas_amd import synth
    from supergwas_amd.api import KgwasHip, ScanParams
    from tests.helpers import T_from, rel_err_T

    if args.copy_calibration:
        k = KgwasHip()
        a, b = k.alloc(args.copy_calibration), k.alloc(args.copy_calibration)
        from supergwas_amd.api import lib
        for _ in range(3):
            k._check(lib().kg_memcpy(k._h, b.ptr, a.ptr, args.copy_calibration, 2))
        log(f"[bench] copied {args.copy_calibration} bytes device-to-device x3")
        return

    cfg = dict(synth.CONFIGS["mixed"])
    cfg["m"] = args.m
    t0 = time.time()
    ds = synth.make_dataset(seed=0, **cfg)
    n_acc, n, p = ds.n_acc, ds.n, ds.p
    rw = synth.row_words_for(n_acc)
    log(f"[bench] dataset n={n} p={p} rw={rw} ({time.time() - t0:.1f}s)")
    prm = ScanParams(0.2, 6.0, 15.0)

    mixed = KgwasHip(precision="mixed", limbs=args.limbs)
    dev = mixed.alloc(args.m * rw * 8)
    mixed.set_model(n_acc, ds.idxs, ds.U, ds.y)
    t1 = time.time()
    mixed.synth_bits(dev.ptr, ds, 0, args.m, rw)
    log(f"[bench] synth {args.m} rows on device ({time.time() - t1:.1f}s)")

    # ---- mixed precision: warmup + timed steps ----
    for _ in range(args.warmup):
        mixed.scan_device(dev.ptr, args.m, rw, prm)
    bulk_ms, ref_ms, cands = [], [], []
    t2 = time.perf_counter()
    for _ in range(args.steps):
        hits, npass = mixed.scan_device(dev.ptr, args.m, rw, prm)
        tm = mixed.timing()
        bulk_ms.append(tm["bulk_ms"])
        ref_ms.append(tm["refine_ms"])
        cands.append(tm["candidates"])
    el = time.perf_counter() - t2
    ms_step = 1e3 * el / args.steps
    value = args.m * args.steps / el
    log(f"[bench] mixed: {ms_step:.2f} ms/step, {value:.3e} k-mers/s, bulk {np.mean(bulk_ms):.3f} ms,"
        f" refine {np.mean(ref_ms):.3f} ms, candidates {int(np.mean(cands))}, hits {len(hits)},"
        f" passed {int(npass[0])}")

    if args.profile_only:
        return

    # ---- plain FP32 kernel ----
    fp32 = KgwasHip(precision="fp32")
    fp32.set_model(n_acc, ds.idxs, ds.U, ds.y)
    mf = min(args.fp32_m, args.m)
    fp32.scan_device(dev.ptr, mf, rw, prm)
    t3 = time.perf_counter()
    f_bulk = []
    for _ in range(args.fp32_steps):
        fp32.scan_device(dev.ptr, mf, rw, prm)
        f_bulk.append(fp32.timing()["bulk_ms"])
    fel = time.perf_counter() - t3
    fp32_value = mf * args.fp32_steps / fel
    log(f"[bench] fp32: {1e3 * fel / args.fp32_steps:.2f} ms/step over {mf} k-mers, "
        f"{fp32_value:.3e} k-mers/s, bulk {np.mean(f_bulk):.3f} ms")

    # ---- statistic error vs FP64 oracle on a sample ----
    es = min(args.err_sample, args.m)
    sample = dev.to_host(np.uint64, (es, rw), nbytes=es * rw * 8)
    ref = oracle.scan_f64(sample, ds.idxs, ds.y, ds.U, 0.2)
    yyM = oracle.yyM(ds.y, ds.U)
    stm = mixed.stats_device(dev.ptr, es, rw)
    stf = fp32.stats_device(dev.ptr, es, rw)
    err_m = rel_err_T(T_from(stm["xxM"], stm["xyM"][:, 0], yyM, n), ref["T"])
    err_f = rel_err_T(T_from(stf["xxM"], stf["xyM"][:, 0], yyM, n), ref["T"])
    counts_exact = bool(np.array_equal(stm["c"], ref["c"]))
    # reported k-mers vs oracle on the sample
    hs, nps = mixed.scan_device(dev.ptr, es, rw, prm)
    exp = np.nonzero(ref["passed"] & (ref["p"] <= 1e-6))[0]
    flips = int(len(np.setxor1d(exp, hs["idx"]))) + abs(int(nps[0]) - ref["n_passed"])
    hf, npf = fp32.scan_device(dev.ptr, es, rw, prm)
    flips_f = int(len(np.setxor1d(exp, hf["idx"]))) + abs(int(npf[0]) - ref["n_passed"])

    # ---- CPU baseline: restatement of the reference's FP32 cols path ----
    cpu = None
    if not args.no_cpu:
        cs = min(args.cpu_sample, args.m)
        X = oracle.unpack_rows(dev.to_host(np.uint64, (cs, rw), nbytes=cs * rw * 8), ds.idxs)
        threads = len(os.sched_getaffinity(0))
        t4 = time.perf_counter()
        res = oracle.cols_fp32(X, ds.y, ds.U, 0.2, 6.0, 15.0, threads)
        cel = time.perf_counter() - t4
        Tc_k = T_from(res["xxM"].astype(np.float64), res["xyM"].astype(np.float64), yyM, n)
        inside = res["keep"] < es   # error on the part of the sample the oracle covers
        err_c = rel_err_T(Tc_k[inside], ref["T"][res["keep"][inside]])
        cpu = {"value": cs / cel, "unit": "k-mer tests/s", "cores": threads, "kind": "port",
               "sample": f"{cs} k-mers of the same workload (n=2048, p=32), in-memory float "
                         f"matrix; pre-filter {res['t_stage'][0]:.2f}s + LA "
                         f"{res['t_stage'][1]:.2f}s + emit {res['t_stage'][2]:.2f}s",
               "stat_max_rel_err_vs_fp64": float(np.nanmax(err_c))}
        log(f"[bench] cpu: {cs / cel:.3e} k-mers/s on {threads} threads ({cel:.1f}s)")

    # ---- roofline of the bulk kernel ----
    # HBM bytes per k-mer measured with rocprofv3 FETCH_SIZE/WRITE_SIZE (copy-calibrated)
    traffic = None
    tp = os.path.join(ROOT, "profiles", "traffic_scan_i8_n2048_p32.md")
    if os.path.exists(tp):
        for ln in open(tp):
            if ln.startswith("traffic_bytes_per_kmer:"):
                traffic = float(ln.split(":")[1]) * args.m
    bulk_s = np.mean(bulk_ms) / 1e3
    alg_bytes = args.m * (n_acc / 8.0)
    achieved = alg_bytes / bulk_s / 1e9
    alg_ops = 2.0 * n * (p + 1) * args.m
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "k-mer tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic (seeded; bits generated on device)",
        "config": {"workload": f"mixed: n=2048, p=32, m={args.m:.0e} k-mers, 1 trait, "
                               f"{args.limbs} int8 limbs", "n": n, "p": p, "m": args.m,
                   "corr_threshold": 0.2, "log10_thresh": 6.0},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "mfma": {"algorithmic_TOPs": alg_ops / bulk_s / 1e12,
                 "executed_TOPs": 2.0 * n_acc * (16 * int(np.ceil(args.limbs * (p) / 16.0)))
                 * args.m / bulk_s / 1e12, "peak_int8_TOPs": I8_PEAK_TOPS},
        "bulk_kernel_ms": float(np.mean(bulk_ms)),
        "refine_kernel_ms": float(np.mean(ref_ms)),
        "candidates_per_step": int(np.mean(cands)),
        "fp32_gpu": {"value": fp32_value, "unit": "k-mer tests/s", "m": mf,
                     "bulk_kernel_ms": float(np.mean(f_bulk)),
                     "stat_max_rel_err_vs_fp64": float(err_f.max()), "flips": flips_f},
        "speedup_vs_fp32_gpu": value / fp32_value,
        "error_vs_fp64": {"sample": es, "stat_max_rel_err": float(err_m.max()),
                          "flips": flips, "counts_bit_exact": counts_exact},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
