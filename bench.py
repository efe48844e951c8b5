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
FP6_PEAK_TOPS = 10100.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """CPU model and online logical CPUs of this host (for the CPU baseline records)."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}


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
    ap.add_argument("--profile-c5", action="store_true",
                    help="only HBM-resident scans of the n=4096, p=64, 8-trait workload (--m rows, "
                         "--steps timed scans; for rocprofv3 runs)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="copy this many bytes device-to-device and exit (counter calibration)")
    ap.add_argument("--legs", default="c2,c4,c5,pairs,ingest,dedup",
                    help="further workloads beside the headline: c2 (FP32 kernel, n=1024 p=20, "
                         "50 batches of 1e6), c4 (400 uneven shards), c5 (n=4096 p=64 8 traits, "
                         "streamed from page-locked host memory), pairs (all-pairs epistasis "
                         "scan of 20,000 rows), ingest (ASCII columns -> packed rows on the "
                         "host), dedup (sorted batches with runs of equal rows); '' for none")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--c5-stream-kmers", type=int, default=1_000_000_000,
                    help="k-mers the c5 leg streams from page-locked host memory")
    ap.add_argument("--screen", default="auto", choices=("auto", "on", "off"),
                    help="pre-filter screen of the headline scans (auto: screen while it pays)")
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

    if args.profile_c5:
        profile_c5(args, synth)
        return

    cfg = dict(synth.CONFIGS["mixed"])
    cfg["m"] = args.m
    t0 = time.time()
    ds = synth.make_dataset(seed=0, **cfg)
    n_acc, n, p = ds.n_acc, ds.n, ds.p
    rw = synth.row_words_for(n_acc)
    log(f"[bench] dataset n={n} p={p} rw={rw} ({time.time() - t0:.1f}s)")
    prm = ScanParams(0.2, 6.0, 15.0)

    mixed = KgwasHip(precision="mixed", limbs=args.limbs, screen=args.screen)
    dev = mixed.alloc(args.m * rw * 8)
    mixed.set_model(n_acc, ds.idxs, ds.U, ds.y)
    t1 = time.time()
    mixed.synth_bits(dev.ptr, ds, 0, args.m, rw)
    log(f"[bench] synth {args.m} rows on device ({time.time() - t1:.1f}s)")

    tune = None
    if not args.no_autotune:
        t_at = time.time()
        tune = mixed.autotune(2_000_000)
        log(f"[bench] autotune: {tune} ({time.time() - t_at:.1f}s)")

    # ---- mixed precision: warmup + timed steps ----
    for _ in range(args.warmup):
        mixed.scan_device(dev.ptr, args.m, rw, prm)
    bulk_ms, ref_ms, cands, scr_ms, surv, tot_ms, rchk = [], [], [], [], [], [], []
    t2 = time.perf_counter()
    for _ in range(args.steps):
        hits, npass = mixed.scan_device(dev.ptr, args.m, rw, prm)
        tm = mixed.timing()
        bulk_ms.append(tm["bulk_ms"])
        ref_ms.append(tm["refine_ms"])
        cands.append(tm["candidates"])
        scr_ms.append(tm["screen_ms"])
        surv.append(tm["survivors"])
        tot_ms.append(tm["total_ms"])
        rchk.append(tm["rechecked"])
    el = time.perf_counter() - t2
    ms_step = 1e3 * el / args.steps
    value = args.m * args.steps / el
    log(f"[bench] mixed: {ms_step:.2f} ms/step, {value:.3e} k-mers/s, bulk {np.mean(bulk_ms):.3f} ms,"
        f" refine {np.mean(ref_ms):.3f} ms, candidates {int(np.mean(cands))}, hits {len(hits)},"
        f" passed {int(npass[0])}, screen {np.mean(scr_ms):.3f} ms, survivors {int(np.mean(surv))},"
        f" rechecked {int(np.mean(rchk))}, device total {np.mean(tot_ms):.3f} ms")

    if args.profile_only:
        return

    # ---- the same scan with the survivors' kernel in its other forms: the
    # resident kernel on two int8 limbs or three FP6 digits, and the streaming
    # kernel (3 int8 limbs, weights staged chunk by chunk) ----
    cur = mixed.get_tune()
    shape = (tune["tiles_per_wave"] if tune else 0, tune["workgroups_per_cu"] if tune else 1,
             tune["screen_wgs_per_cu"] if tune else 1)
    variants = {}
    for vname, kw in (("resident_int8_2limbs", dict(resident_form=1)),
                      ("resident_fp6_3digits", dict(resident_form=3)),
                      ("streamed_weights_3limbs", dict(stream_weights=1))):
        mixed.set_tune(*shape, **kw)
        mixed.scan_device(dev.ptr, args.m, rw, prm)
        t2b = time.perf_counter()
        b2, c2 = [], []
        for _ in range(max(args.steps // 2, 1)):
            h2, np2 = mixed.scan_device(dev.ptr, args.m, rw, prm)
            b2.append(mixed.timing()["bulk_ms"])
            c2.append(mixed.timing()["candidates"])
        v2 = args.m * max(args.steps // 2, 1) / (time.perf_counter() - t2b)
        same2 = bool(np.array_equal(h2["idx"], hits["idx"]) and np.array_equal(np2, npass))
        variants[vname] = {"value": v2, "unit": "k-mer tests/s", "bulk_kernel_ms": float(np.mean(b2)),
                           "candidates_per_step": int(np.mean(c2)), "same_hits_and_counts": same2}
        log(f"[bench] mixed, {vname}: {v2:.3e} k-mers/s, bulk {np.mean(b2):.3f} ms, "
            f"candidates {int(np.mean(c2))}, same hits and counts: {same2}")
    mixed.set_tune(*shape, stream_weights=cur["stream_weights"], resident_form=cur["resident_form"])

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
    rep_err = float(np.max(rel_err_T(hs["T"], ref["T"][hs["idx"]]))) if len(hs) else 0.0

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
               **host_cpu(),
               "sample": f"{cs} k-mers of the same workload (n=2048, p=32), in-memory float "
                         f"matrix; pre-filter {res['t_stage'][0]:.2f}s + LA "
                         f"{res['t_stage'][1]:.2f}s + emit {res['t_stage'][2]:.2f}s",
               "stat_max_rel_err_vs_fp64": float(np.nanmax(err_c))}
        log(f"[bench] cpu: {cs / cel:.3e} k-mers/s on {threads} threads ({cel:.1f}s)")

    # ---- roofline of the dominant kernel ----
    # Each bulk kernel streams its k-mers' packed rows once: n_acc / 8 bytes per
    # k-mer (the screen over all m rows, the full kernel over the survivors).
    # HBM bytes per k-mer measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    # (copy-calibrated) are read from the committed traffic summary.
    traffic_per = {}
    tp = os.path.join(ROOT, "profiles", "traffic_scan_n2048_p32.md")
    if os.path.exists(tp):
        for ln in open(tp):
            if ln.startswith("traffic_bytes_per_kmer "):
                kname, val = ln[len("traffic_bytes_per_kmer "):].split(":")
                traffic_per[kname.strip()] = float(val)
    row_bytes = n_acc / 8.0
    n_surv = float(np.mean(surv))
    # the one-trait screen streams its rows through an LDS ring (screen shapes 3
    # and 4) or loads them into registers (screen_i8)
    scr_name = "screen_ring" if mixed.get_tune()["screen_wgs_per_cu"] in (3, 4) else "screen_i8"
    kern = {scr_name: {"ms": float(np.mean(scr_ms)), "kmers": args.m},
            "scan_i8": {"ms": float(np.mean(bulk_ms)), "kmers": n_surv}}
    for kname, kv in kern.items():
        kv["achieved_GBs"] = (kv["kmers"] * row_bytes / (kv["ms"] / 1e3) / 1e9
                              if kv["ms"] > 0 else 0.0)
    dom = max(kern, key=lambda x: kern[x]["ms"])
    achieved = kern[dom]["achieved_GBs"]
    traffic = (traffic_per[dom] * kern[dom]["kmers"]) if dom in traffic_per else None
    bulk_s = np.mean(bulk_ms) / 1e3
    # weight rows the survivors' kernel executes: 2 or 3 digits / limbs per
    # column with resident weights (FP6 x FP4 or int8 MFMA), else args.limbs
    tnow = mixed.get_tune()
    form = tnow["resident_form"] if tnow["resident_active"] else 0
    limbs_exec = {0: args.limbs, 1: 2, 2: 2, 3: 3}[form]
    ncol_exec = 16 * int(np.ceil(limbs_exec * p / 16.0))
    surv_ops = {0: f"{args.limbs} int8 limbs (streamed weights)",
                1: "2 int8 limbs with weights resident in LDS",
                2: "2 FP6 digits x FP4 bits with weights resident in LDS (3 int8 limbs for the "
                   "k-mers its bounds leave open)",
                3: "3 FP6 digits x FP4 bits with weights resident in LDS"}[form]
    surv_peak = FP6_PEAK_TOPS if form in (2, 3) else I8_PEAK_TOPS
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
        "dtype": ("fp6 x fp4 with exact fp32 sums (screen and survivors)" if form in (2, 3)
                  else "int8 (survivors), fp6 x fp4 with exact fp32 sums (screen)"),
        "data": "synthetic (seeded; bits generated on device)",
        "config": {"workload": f"mixed: n=2048, p=32, m={args.m:.0e} k-mers, 1 trait; screen "
                               f"5 FP6 digits x FP4 bits, survivors {surv_ops}, statistic "
                               f"{args.limbs} int8 limbs, FP64 refinement",
                   "n": n, "p": p, "m": args.m,
                   "corr_threshold": 0.2, "log10_thresh": 6.0},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "kernels": kern,
        "mfma": {"scan_i8_executed_TOPs": 2.0 * n_acc * ncol_exec * n_surv / bulk_s / 1e12
                 if bulk_s > 0 else 0.0,
                 "operands": "fp6 x fp4" if form in (2, 3) else "int8",
                 "peak_TOPs": surv_peak},
        "screen_kernel_ms": float(np.mean(scr_ms)),
        "survivors_per_step": int(n_surv),
        "bulk_kernel_ms": float(np.mean(bulk_ms)),
        "refine_kernel_ms": float(np.mean(ref_ms)),
        "candidates_per_step": int(np.mean(cands)),
        "fp32_gpu": {"value": fp32_value, "unit": "k-mer tests/s", "m": mf,
                     "bulk_kernel_ms": float(np.mean(f_bulk)),
                     "stat_max_rel_err_vs_fp64": float(err_f.max()), "flips": flips_f},
        "rechecked_per_step": int(np.mean(rchk)),
        "survivor_kernel_forms": variants,
        "speedup_vs_fp32_gpu": value / fp32_value,
        "error_vs_fp64": {"sample": es, "stat_max_rel_err": float(err_m.max()),
                          "reported_stat_max_rel_err": rep_err,
                          "flips": flips, "counts_bit_exact": counts_exact},
        "cpu_baseline": cpu,
    }
    if tune is not None:
        line["config"]["autotune"] = tune
    dev.free()
    legs = [x for x in args.legs.split(",") if x]
    if legs:
        line["legs"] = {}
    for name in legs:
        t_leg = time.time()
        line["legs"][name] = LEGS[name](args, synth, ds)
        log(f"[bench] leg {name} done ({time.time() - t_leg:.1f}s)")
    print(json.dumps(line), flush=True)


def profile_c5(args, synth):
    """HBM-resident scans of the n=4096, p=64, 8-trait workload only (the c5 leg's
    kernel-only part), for profiler runs: --warmup untimed and --steps timed scans of --m rows."""
    from supergwas_amd.api import KgwasHip, ScanParams
    ds = synth.make_dataset(seed=5, n_struct=1024, **dict(synth.CONFIGS["multitrait"], m=args.m))
    rw = synth.row_words_for(ds.n_acc)
    k = KgwasHip(precision="mixed", screen=args.screen)
    k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
    dev = k.alloc(args.m * rw * 8)
    k.synth_bits(dev.ptr, ds, 0, args.m, rw)
    prm = ScanParams(0.2, 6.0, 0.0)
    for _ in range(args.warmup):
        k.scan_device(dev.ptr, args.m, rw, prm)
    for _ in range(args.steps):
        h, npass = k.scan_device(dev.ptr, args.m, rw, prm)
        tm = k.timing()
        log(f"[bench] c5 profile: m={args.m} screen {tm['screen_ms']:.3f} ms, full {tm['bulk_ms']:.3f} "
            f"ms, refine {tm['refine_ms']:.3f} ms, survivors {tm['survivors']}, hits {len(h)}")
    dev.free()
    k.close()


def leg_c2(args, synth, ds_c3):
    """Plain FP32 kernel at scale: n=1024, p=20, m=5e7 in 50 batches of 1e6, hits at p < 1e-5
    and a full-statistic checksum (sum of T in FP64, sum of allele counts exactly)."""
    from supergwas_amd.api import KgwasHip, ScanParams
    from tests.helpers import T_from
    import oracle
    cfg = dict(synth.CONFIGS["fp32_scale"])
    ds = synth.make_dataset(seed=2, **cfg)
    rw = synth.row_words_for(ds.n_acc)
    nb, bm = 50, 1_000_000
    prm = ScanParams(0.2, 5.0, 0.0)
    out = {"workload": "FP32 GPU kernel: n=1024, p=20, m=5e7 as 50 batches of 1e6"}
    yy = oracle.yyM(ds.y, ds.U)
    for prec in ("fp32", "mixed"):
        k = KgwasHip(precision=prec)
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.alloc(nb * bm * rw * 8)
        k.synth_bits(dev.ptr, ds, 0, nb * bm, rw)
        k.scan_device(dev.ptr, bm, rw, prm)  # warm-up
        t0 = time.perf_counter()
        bulk, hits = 0.0, 0
        for b in range(nb):
            h, _ = k.scan_device(dev.ptr + b * bm * rw * 8, bm, rw, prm)
            tm = k.timing()
            bulk += tm["bulk_ms"] + tm["screen_ms"] + tm["refine_ms"]
            hits += len(h)
        el = time.perf_counter() - t0
        sT, sc = 0.0, 0
        for b in range(nb):
            st = k.stats_device(dev.ptr + b * bm * rw * 8, bm, rw)
            sT += float(T_from(st["xxM"], st["xyM"][:, 0], yy, ds.n).sum())
            sc += int(st["c"].astype(np.int64).sum())
        out[prec] = {"kmers_per_s": nb * bm / el, "kernel_kmers_per_s": nb * bm / (bulk / 1e3),
                     "hits_p_lt_1e-5": hits, "checksum_sum_T": sT, "checksum_sum_c": sc}
        log(f"[bench] c2 {prec}: {nb * bm / el:.3e} k-mers/s ({nb * bm / (bulk / 1e3):.3e} kernel),"
            f" hits {hits}, sum T {sT:.6e}, sum c {sc}")
        dev.free()
        k.close()
    out["checksum_rel_diff_T"] = abs(out["fp32"]["checksum_sum_T"] - out["mixed"]["checksum_sum_T"]) / \
        max(abs(out["mixed"]["checksum_sum_T"]), 1.0)
    # both paths' statistics against the FP64 oracle on the first batch of 1e6
    bits = ds.fill_bits(0, bm)
    ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, 0.2)
    sT64 = float(ref["T"].sum())
    out["fp64_first_batch"] = {"kmers": bm, "sum_T": sT64, "sum_c": int(ref["c"].astype(np.int64).sum()),
                               "hits_p_lt_1e-5": int(np.sum(ref["passed"] & (ref["p"] <= 1e-5)))}
    for prec in ("fp32", "mixed"):
        k = KgwasHip(precision=prec)
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dbits = k.to_device(bits)
        st = k.stats_device(dbits.ptr, bm, rw)
        dbits.free()
        T = T_from(st["xxM"], st["xyM"][:, 0], yy, ds.n)
        h, _ = k.scan(bits, prm)
        exp = np.nonzero(ref["passed"] & (ref["p"] <= 1e-5))[0]
        out[prec]["first_batch_vs_fp64"] = {
            "sum_T_rel_diff": abs(float(T.sum()) - sT64) / sT64,
            "stat_max_rel_err": float(np.max(np.abs(T - ref["T"]) / np.maximum(ref["T"], 1.0))),
            "counts_bit_exact": bool(np.array_equal(st["c"], ref["c"])),
            "hit_flips": int(len(np.setxor1d(exp, h["idx"])))}
        k.close()
    log(f"[bench] c2 vs FP64 (first batch): fp32 {out['fp32']['first_batch_vs_fp64']}, "
        f"mixed {out['mixed']['first_batch_vs_fp64']}")
    out["checksum_c_equal"] = out["fp32"]["checksum_sum_c"] == out["mixed"]["checksum_sum_c"]
    return out


def leg_c4(args, synth, ds):
    """400 shards, sizes log-uniform in [1e3, 2e6], each with its own 0-15% missing-phenotype
    mask (own null model), n=2048, p=32: one persistent launch (and its HIP-graph replay)
    against one launch per shard."""
    from supergwas_amd.api import SHARDS_GRAPH, SHARDS_PER_LAUNCH, KgwasHip, ScanParams
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(2e6), 400)).astype(np.int64)
    total = int(sizes.sum())
    rw = synth.row_words_for(ds.n_acc)
    k = KgwasHip(precision="mixed")
    mids, models = [], []
    t0 = time.time()
    for s in range(len(sizes)):
        keep = np.ones(ds.n, bool)
        drop = rng.choice(ds.n, size=int(rng.uniform(0, 0.15) * ds.n), replace=False)
        keep[drop] = False
        idxs = ds.idxs[keep]
        U, _, _ = np.linalg.svd(ds.U[keep], full_matrices=False)
        y = ds.y[keep]
        models.append((idxs, U, y))
        if s == 0:
            k.set_model(ds.n_acc, idxs, U, y)
            mids.append(0)
        else:
            mids.append(k.add_model(ds.n_acc, idxs, U, y))
    log(f"[bench] c4: {len(sizes)} shard models in {time.time() - t0:.1f}s, {total} k-mers")
    dev = k.alloc(total * rw * 8)
    k.synth_bits(dev.ptr, ds, 0, total, rw)
    tune = k.autotune(2_000_000)
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    shards = [(dev.ptr + int(o) * rw * 8, int(m), rw, mid) for o, m, mid in zip(off, sizes, mids)]
    prm = ScanParams(0.2, 6.0, 0.0)
    out = {"workload": f"{len(sizes)} shards log-uniform in [1e3, 2e6], total {total} k-mers, "
                       "n=2048, p=32, per-shard masks 0-15% missing",
           "autotune": tune}
    ref = None
    runs = (("persistent", 0, "auto"), ("graph", SHARDS_GRAPH, "auto"),
            ("persistent_unscreened", 0, "off"), ("per_shard", SHARDS_PER_LAUNCH, "off"))
    for name, flags, scr in runs:
        k.set_screen(scr)
        k.scan_shards(shards, prm, flags)  # warm-up (and graph capture)
        ts, ks = [], []
        for _ in range(3):
            t1 = time.perf_counter()
            hits, npass = k.scan_shards(shards, prm, flags)
            ts.append(time.perf_counter() - t1)
            ks.append(k.timing()["total_ms"])
        same = True
        if ref is None:
            ref = (hits.copy(), npass.copy())
        else:
            same = bool(np.array_equal(hits["idx"], ref[0]["idx"]) and
                        np.array_equal(hits["shard"], ref[0]["shard"]) and
                        np.array_equal(npass, ref[1]))
        out[name] = {"kmers_per_s": total / min(ts), "device_ms": min(ks), "wall_ms": 1e3 * min(ts),
                     "hits": int(len(hits)), "matches_persistent": same}
        log(f"[bench] c4 {name}: {total / min(ts):.3e} k-mers/s (device {min(ks):.1f} ms,"
            f" wall {1e3 * min(ts):.1f} ms), hits {len(hits)}, same {same}")
    # the persistent launch's per-shard counts and hits against the FP64 oracle on the
    # eight smallest shards (each with its own mask, U and phenotype)
    import oracle
    h0, n0 = ref
    checked, count_eq, hit_eq = 0, 0, 0
    for s in np.argsort(sizes)[:8]:
        m, o = int(sizes[s]), int(off[s])
        bits = dev.to_host(np.uint64, (m, rw), nbytes=m * rw * 8, offset=o * rw * 8)
        idxs, U, y = models[s]
        r = oracle.scan_f64(bits, idxs, y, U, 0.2)
        exp = np.nonzero(r["passed"] & (r["p"] <= 1e-6))[0]
        checked += 1
        count_eq += int(n0[s, 0] == r["n_passed"])
        hit_eq += int(np.array_equal(h0[h0["shard"] == s]["idx"], exp))
    out["oracle_check"] = {"shards": checked, "counts_equal": count_eq, "hits_equal": hit_eq}
    log(f"[bench] c4 oracle check on {checked} shards: counts equal {count_eq}, hits equal {hit_eq}")
    out["speedup_persistent_vs_per_shard"] = (out["per_shard"]["wall_ms"] /
                                              out["persistent_unscreened"]["wall_ms"])
    out["speedup_screen"] = out["persistent_unscreened"]["wall_ms"] / out["persistent"]["wall_ms"]
    dev.free()
    k.close()
    return out


def leg_c5(args, synth, ds_c3):
    """n=4096, p=64, 8 traits (correlation 0.2, 100 causal each) in one pass: kernel-only rate on
    HBM-resident rows, every trait's statistic against the FP64 oracle on a sample, and the
    end-to-end rate of a 1e9-k-mer stream from page-locked host memory (64 MiB chunks on a copy
    stream): kg_scan over a 4 GB page-locked ring of 8e6 rows, 125 passes, every pass copying
    its rows over the host link again."""
    import oracle
    from supergwas_amd.api import KgwasHip, ScanParams
    from tests.helpers import T_from, rel_err_T
    cfg = dict(synth.CONFIGS["multitrait"])
    cfg["m"] = 10_000_000
    ds = synth.make_dataset(seed=5, n_struct=1024, **cfg)
    rw = synth.row_words_for(ds.n_acc)
    t = ds.Y.shape[1]
    k = KgwasHip(precision="mixed")
    k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
    prm = ScanParams(0.2, 6.0, 0.0)
    mk, mh = 10_000_000, 8_000_000
    dev = k.alloc(mk * rw * 8)
    k.synth_bits(dev.ptr, ds, 0, mk, rw)
    tune = k.autotune(1_000_000)
    k.scan_device(dev.ptr, mk, rw, prm)
    bulk = []
    for _ in range(3):
        k.scan_device(dev.ptr, mk, rw, prm)
        tm = k.timing()
        bulk.append(tm["bulk_ms"] + tm["screen_ms"] + tm["refine_ms"])
    surv = tm["survivors"]
    kb = min(bulk) / 1e3
    # statistic of every trait against FP64 on a sample of the same rows
    es = min(args.err_sample // 2, mk)
    sample = dev.to_host(np.uint64, (es, rw), nbytes=es * rw * 8)
    st = k.stats_device(dev.ptr, es, rw)
    refined = k.timing()["stat_refined"]
    hs, nps = k.scan_device(dev.ptr, es, rw, prm)
    err, flips, counts_exact = 0.0, 0, True
    for tr in range(t):
        ref = oracle.scan_f64(sample, ds.idxs, ds.Y[:, tr], ds.U, 0.2)
        counts_exact &= bool(np.array_equal(st["c"], ref["c"]))
        Tb = T_from(st["xxM"], st["xyM"][:, tr], oracle.yyM(ds.Y[:, tr], ds.U), ds.n)
        err = max(err, float(rel_err_T(Tb, ref["T"]).max()))
        exp = np.nonzero(ref["passed"] & (ref["p"] <= 1e-6))[0]
        got = hs[hs["trait"] == tr]["idx"]
        flips += int(len(np.setxor1d(exp, got))) + abs(int(nps[tr]) - ref["n_passed"])
    host = k.host_alloc((mh, rw))
    from supergwas_amd.api import lib
    k._check(lib().kg_memcpy(k._h, host.ctypes.data, dev.ptr, mh * rw * 8, 1))
    dev.free()
    k.scan(host, prm)  # warm-up of the chunk buffers
    passes = max(1, int(round(args.c5_stream_kmers / mh)))
    t0 = time.perf_counter()
    copies, kern, nhits, npass = 0.0, 0.0, 0, np.zeros(t, np.int64)
    for _ in range(passes):
        hits, np_ = k.scan(host, prm)
        tm = k.timing()
        copies += tm["copy_ms"]
        kern += tm["bulk_ms"] + tm["screen_ms"]
        nhits += len(hits)
        npass += np_
    w = time.perf_counter() - t0
    k.host_free(host)
    total = passes * mh
    out = {"workload": "n=4096, p=64, t=8 (corr 0.2, 100 causal each), 3 int8 limbs",
           "autotune": tune,
           "kernel_only": {"m": mk, "kmers_per_s": mk / kb, "GBps": mk * rw * 8 / kb / 1e9,
                           "trait_tests_per_s": t * mk / kb, "survivors": int(surv)},
           "error_vs_fp64": {"sample": es, "traits": t, "stat_max_rel_err": err,
                             "flips": flips, "counts_bit_exact": counts_exact,
                             "stat_refined_in_fp64": int(refined)},
           "streamed": {"kmers": total, "passes": passes, "kmers_per_pass": mh,
                        "kmers_per_s": total / w, "GBps": total * rw * 8 / w / 1e9,
                        "copy_ms": copies, "kernel_ms": kern, "wall_ms": 1e3 * w,
                        "chunk_MiB": 64, "host_memory": "page-locked ring of 8e6 rows "
                        "(kg_host_alloc), scanned by kg_scan once per pass",
                        "hits": int(nhits), "passed_per_trait": [int(x) for x in npass]}}
    log(f"[bench] c5 kernel {mk / kb:.3e} k-mers/s ({mk * rw * 8 / kb / 1e9:.0f} GB/s), stat err "
        f"{err:.2e} (refined {refined}), flips {flips}, counts exact {counts_exact}; streamed "
        f"{total:.2e} k-mers at {total / w:.3e} k-mers/s ({total * rw * 8 / w / 1e9:.1f} GB/s, "
        f"copy {copies:.0f} ms, kernel {kern:.0f} ms, wall {1e3 * w:.0f} ms)")
    k.close()
    return out


def leg_pairs(args, synth, ds):
    """All-pairs epistasis scan (kgwas_2d, src/kgwas_2d.cpp:258-401) of 20,000 rows at n=2048,
    p=32: every pair i < j tested as x_i AND x_j, p <= 1e-6 (the 2D default threshold).  Rows
    are HBM-resident; the CPU baseline is the FP64 oracle on the pairs of the first rows, which
    the GPU must reproduce exactly."""
    import oracle
    from supergwas_amd.api import KgwasHip, ScanParams
    N, l10 = 20_000, 6.0
    rw = synth.row_words_for(ds.n_acc)
    k = KgwasHip(precision="mixed")
    k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
    dev = k.alloc(N * rw * 8)
    k.synth_bits(dev.ptr, ds, 0, N, rw)
    prm = ScanParams(0.0, l10, 0.0)
    k.scan_pairs(dev.ptr, N, rw, prm)
    walls, kern = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        hits = k.scan_pairs(dev.ptr, N, rw, prm)
        walls.append(time.perf_counter() - t0)
        tm = k.timing()
        kern.append(tm["bulk_ms"] + tm["refine_ms"])
    npairs = N * (N - 1) // 2
    w, kb = min(walls), min(kern) / 1e3
    # MFMA work the kernel executes per pair, counted as for the headline: 3 limbs x p weight
    # columns (the intercept direction comes from the allele count) in 16-row tiles
    rows16 = 16 * int(np.ceil(3 * ds.p / 16.0))
    exec_ops = 2.0 * ds.n_acc * rows16 * npairs
    out = {"workload": f"all pairs of {N} rows, n={ds.n}, p={ds.p}, p <= 1e-{l10:g}",
           "pairs": npairs, "pairs_per_s": npairs / w, "kernel_pairs_per_s": npairs / kb,
           "kernel_ms": 1e3 * kb, "wall_ms": 1e3 * w, "hits": int(len(hits)),
           "candidates": int(tm["candidates"]),
           "executed_int8_TOPs": exec_ops / kb / 1e12,
           "algorithmic_TFLOPs": 2.0 * ds.n * (ds.p + 1) * npairs / kb / 1e12}
    log(f"[bench] pairs: {npairs / w:.3e} pairs/s end to end, {npairs / kb:.3e} pairs/s kernel "
        f"({1e3 * kb:.1f} ms, {exec_ops / kb / 1e12:.0f} executed int8 TOP/s), {len(hits)} hits")
    if not args.no_cpu:
        R = 1500
        bits = dev.to_host(np.uint64, (R, rw), nbytes=R * rw * 8)
        t0 = time.perf_counter()
        ref = oracle.pairs_f64(bits, ds.idxs, ds.y, ds.U, l10)
        tc = time.perf_counter() - t0
        sub = k.scan_pairs(dev.ptr, R, rw, prm)
        same = (np.array_equal(sub["shard"].astype(np.int64), ref["i"]) and
                np.array_equal(sub["idx"], ref["j"]) and np.array_equal(sub["c"], ref["c"]))
        terr = float(np.max(np.abs(sub["T"] - ref["T"]) / np.maximum(ref["T"], 1.0),
                            initial=0.0)) if same else None
        rp = R * (R - 1) // 2
        out["cpu_baseline"] = {"value": rp / tc, "unit": "pairs/s", "cores": os.cpu_count(),
                               "kind": "port", **host_cpu(), "sample": f"all {rp} pairs of the first {R} rows "
                               "(FP64 oracle, OpenMP over i)"}
        out["parity_first_rows"] = {"rows": R, "hits": int(len(sub)), "identical_pairs": same,
                                    "T_max_rel_err_vs_fp64": terr}
        log(f"[bench] pairs CPU oracle {rp / tc:.3e} pairs/s on {R} rows; GPU identical: {same}")
    dev.free()
    k.close()
    return out


def leg_ingest(args, synth, ds):
    """Cols-mode ingest: the drop-in kgwas turns per-sample ASCII columns
    (bits/col_%06d.txt, one byte per k-mer) into packed k-mer-major rows with SSE2 tile
    transposes (KG_PACK_ONLY stops it there), beside the reference's loader restated
    (load_matrix_rows, src/kgwas_utils.hpp:225-282: a float n x N matrix).  250,000
    k-mers x 2048 accessions (512 MB of ASCII) from the page cache, 16 threads each."""
    import re
    import shutil
    import subprocess
    import tempfile
    import oracle
    from supergwas_amd import batchdir
    N, threads = 250_000, 16
    bits = ds.fill_bits(0, N)
    root = tempfile.mkdtemp(prefix="kg_ingest_")
    try:
        paths = batchdir.write_inputs(root, ds, bits, rows_file=False)
        exe = os.path.join(ROOT, "supergwas_amd", "bin", "kgwas")
        out = os.path.join(root, "rows.bin")
        env = dict(os.environ, KG_PACK_ONLY=out, LOAD_THREADS=str(threads))
        cmd = [exe, paths["samples"], paths["pheno"], paths["batch"], paths["ubin"],
               os.path.join(root, "o.tsv")]
        walls, packs = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=300)
            walls.append(time.perf_counter() - t0)
            if r.returncode != 0:
                raise RuntimeError("kgwas pack failed: " + r.stderr[-500:])
            m = re.search(r"\[LOAD\] \d+/\d+  \(([0-9.e+-]+) s", r.stderr)
            packs.append(float(m.group(1)))
        rows = np.fromfile(out, np.uint64).reshape(N, -1)
        mask = np.zeros(bits.shape[1], np.uint64)
        for a in ds.idxs:
            mask[a >> 6] |= np.uint64(1) << np.uint64(a & 63)
        exact = bool(np.array_equal(rows, bits & mask))
        del rows
        tref = []
        for _ in range(2):
            t0 = time.perf_counter()
            X = oracle.load_columns(os.path.join(paths["batch"], "bits"), ds.idxs, N, threads)
            tref.append(time.perf_counter() - t0)
            del X
    finally:
        shutil.rmtree(root, ignore_errors=True)
    ascii_gb = ds.n * N / 1e9
    out = {"workload": f"{N} k-mers x {ds.n} phenotyped accessions, plain .txt columns "
                       f"({ascii_gb * 1e3:.0f} MB ASCII, page cache)",
           "threads": threads, "pack_s": min(packs), "pack_GBps": ascii_gb / min(packs),
           "process_wall_s": min(walls), "rows_bit_exact": exact,
           "cpu_baseline": {"value": ascii_gb / min(tref), "unit": "GB/s of ASCII",
                            "cores": threads, "kind": "port", **host_cpu(),
                            "sample": "load_matrix_rows restated (float n x N), same columns",
                            "seconds": min(tref)}}
    log(f"[bench] ingest: pack {min(packs):.3f} s ({ascii_gb / min(packs):.2f} GB/s), process "
        f"{min(walls):.2f} s, bit-exact {exact}; reference-style float load {min(tref):.3f} s "
        f"({ascii_gb / min(tref):.2f} GB/s)")
    return out


def leg_dedup(args, synth, ds):
    """Duplicate-row scans on sorted batches (src/convert_matrix.cpp:366-368 sorts k-mers by
    bit pattern, so equal rows are adjacent): 5e6 distinct rows of the n=2048, p=32 workload,
    each repeated a geometric number of times (mean 2.9), about 1.4e7 rows in HBM; the screened
    scan with and without the duplicate check, same hits and counts required."""
    from supergwas_amd.api import KgwasHip, ScanParams
    rng = np.random.default_rng(11)
    mb = 5_000_000
    lens = rng.geometric(0.35, mb)
    rows = np.repeat(ds.fill_bits(0, mb), lens, axis=0)
    m, rw = rows.shape
    prm = ScanParams(0.2, 6.0, 15.0)
    out = {"workload": f"{m} rows = {mb} distinct rows in runs (mean {m / mb:.2f}), n=2048, p=32"}
    res = {}
    for dd in (False, True):
        k = KgwasHip(precision="mixed", screen="on", dedup=dd)
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.to_device(rows)
        k.scan_device(dev.ptr, m, rw, prm)
        walls, dev_ms = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            hits, npass = k.scan_device(dev.ptr, m, rw, prm)
            walls.append(time.perf_counter() - t0)
            tm = k.timing()
            dev_ms.append(tm["total_ms"])
        res[dd] = (hits, npass)
        out["dedup" if dd else "plain"] = {
            "kmers_per_s": m / min(walls), "device_ms": min(dev_ms),
            "screen_ms": tm["screen_ms"], "bulk_ms": tm["bulk_ms"],
            "survivors": int(tm["survivors"]), "hits": int(len(hits)), "passed": int(npass[0])}
        dev.free()
        k.close()
    (h0, p0), (h1, p1) = res[False], res[True]
    out["identical"] = bool(np.array_equal(p0, p1) and len(h0) == len(h1) and
                            all(np.array_equal(h0[f], h1[f]) for f in h0.dtype.names))
    out["speedup"] = out["dedup"]["kmers_per_s"] / out["plain"]["kmers_per_s"]
    log(f"[bench] dedup: {m} rows, plain {out['plain']['kmers_per_s']:.3e} k-mers/s "
        f"({out['plain']['survivors']} survivors), dedup {out['dedup']['kmers_per_s']:.3e} "
        f"({out['dedup']['survivors']} survivors), identical {out['identical']}")
    return out


LEGS = {"c2": leg_c2, "c4": leg_c4, "c5": leg_c5, "pairs": leg_pairs, "ingest": leg_ingest,
        "dedup": leg_dedup}


if __name__ == "__main__":
    main()
