"""GPU parity of the HIP scan against the FP64 oracle (and the FP32 CPU restatement).

Bar (per the workload definition): allele counts bit-exact, reported k-mers
identical to the FP64 oracle's outside a +-1e-6 relative band around the
thresholds, reported statistics equal to FP64 (they are FP64-refined), and the
bulk statistic of every k-mer within 1e-5 relative error (|T - T64| / max(T64, 1)).
"""
import numpy as np
import pytest

import oracle
from supergwas_amd import synth
from tests.helpers import T_from, expected_hits, pearson, rel_err_T

pytestmark = pytest.mark.gpu

STAT_TOL = 1e-5      # bulk-kernel statistic, relative (|T - T64| / max(T64, 1))
REPORT_TOL = 1e-9    # FP64-refined reported values
BAND = 1e-6          # threshold band inside which a flip is not counted

_DS = {}


def dataset(key, **kw):
    if key not in _DS:
        _DS[key] = synth.make_dataset(**kw)
    return _DS[key]


CASES = {
    # name: (dataset kwargs, m, corr, log10, log10_2d)
    "small_n190_p5": (dict(seed=0, n_acc=200, n_missing=10, npop=3, p=5, m=30000, n_causal=5,
                           effect=0.5), 30000, 0.2, 2.0, 3.0),
    "n1024_p20": (dict(seed=1, n_acc=1024, n_missing=50, npop=8, p=20, m=8269, n_causal=20,
                       effect_sd=0.3), 8269, 0.2, 3.0, 5.0),
    "n2048_p32": (dict(seed=2, n_acc=2048, n_missing=0, npop=8, p=32, m=12000, n_causal=50,
                       effect_sd=0.3), 12000, 0.2, 4.0, 8.0),
    "n333_nointercept": (dict(seed=3, n_acc=333, n_missing=7, npop=4, p=6, m=5001, n_causal=5,
                              effect=0.5, intercept=False), 5001, 0.1, 2.0, 4.0),
    "n2048_p32_missing": (dict(seed=4, n_acc=2048, n_missing=60, npop=8, p=32, m=12000,
                               n_causal=50, effect_sd=0.3), 12000, 0.2, 4.0, 8.0),
}


def _run_case(name, precision="mixed", limbs=3):
    from supergwas_amd.api import KgwasHip, ScanParams
    kw, m, corr, l10, l2d = CASES[name]
    ds = dataset(name, **kw)
    bits = ds.fill_bits(0, m)
    ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, corr)
    with KgwasHip(precision=precision, limbs=limbs) as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        hits, npass = k.scan(bits, ScanParams(corr, l10, l2d))
        tim = k.timing()
    return ds, bits, ref, hits, npass, (corr, l10, l2d), tim


def _band_sets(ds, bits, ref, corr, l10):
    cor = pearson(bits, ds.idxs, ds.y)
    with np.errstate(divide="ignore", invalid="ignore"):
        near_c = np.abs(np.abs(cor) / corr - 1.0) <= BAND
        near_p = np.abs(ref["p"] / 10.0 ** (-l10) - 1.0) <= BAND
    return near_c, near_p


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("limbs", [3, 2])
def test_mixed_scan_matches_oracle(built, name, limbs):
    ds, bits, ref, hits, npass, (corr, l10, l2d), tim = _run_case(name, "mixed", limbs)
    print(f"[case] {name} limbs={limbs} n={ds.n} p={ds.p} m={len(bits)} "
          f"hits={len(hits)} passed={npass[0]} oracle_passed={ref['n_passed']} "
          f"candidates={tim['candidates']}", flush=True)
    near_c, near_p = _band_sets(ds, bits, ref, corr, l10)
    # pre-filter count: identical up to k-mers inside the band
    assert abs(int(npass[0]) - ref["n_passed"]) <= int(near_c.sum())
    exp_idx, exp_2d = expected_hits(ref, l10, l2d)
    got_idx = hits["idx"]
    assert np.all(np.diff(got_idx) > 0), "hits must be sorted by k-mer index"
    sym = np.setxor1d(exp_idx, got_idx)
    assert np.all(near_c[sym] | near_p[sym]), f"flips outside band: {sym[:10]}"
    common, ie, ig = np.intersect1d(exp_idx, got_idx, return_indices=True)
    h = hits[ig]
    assert np.array_equal(h["c"], ref["c"][common])
    for key, tol in (("T", REPORT_TOL), ("effect", REPORT_TOL)):
        err = np.abs(h[key] - ref[key][common]) / np.maximum(np.abs(ref[key][common]), 1e-300)
        assert err.max(initial=0) <= tol, (key, err.max())
    perr = np.abs(h["pval"] - ref["p"][common]) / np.maximum(ref["p"][common], 1e-300)
    assert perr.max(initial=0) <= 1e-7
    assert np.array_equal((h["flags"] & 1).astype(bool), exp_2d[ie])


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_bulk_statistic_error(built, name):
    """Full-statistic mode: bulk (int8-limb MFMA + FP64 epilogue) values of every k-mer."""
    from supergwas_amd.api import KgwasHip
    kw, m, corr, l10, l2d = CASES[name]
    ds = dataset(name, **kw)
    bits = ds.fill_bits(0, m)
    ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, corr)
    with KgwasHip(precision="mixed") as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.to_device(bits)
        st = k.stats_device(dev.ptr, m, bits.shape[1])
        refined = k.timing()["stat_refined"]
        dev.free()
    assert np.array_equal(st["c"], ref["c"]), "allele counts must be bit-exact"
    yyM = oracle.yyM(ds.y, ds.U)
    T = T_from(st["xxM"], st["xyM"][:, 0], yyM, ds.n)
    err = rel_err_T(T, ref["T"])
    print(f"[stat] {name}: max rel err T = {err.max():.3e}, refined in FP64 "
          f"{refined} of {m}", flush=True)
    assert err.max() <= STAT_TOL


def test_fp32_matches_cpu_restatement(built):
    """Plain FP32 kernel vs the FP32 CPU restatement of run_cols: same semantics,
    summation order differs, so compare within FP32 tolerance."""
    from supergwas_amd.api import KgwasHip, ScanParams
    name = "small_n190_p5"
    kw, m, corr, l10, l2d = CASES[name]
    ds = dataset(name, **kw)
    bits = ds.fill_bits(0, m)
    X = oracle.unpack_rows(bits, ds.idxs)
    cpu = oracle.cols_fp32(X, ds.y, ds.U, corr, l10, l2d, 4)
    with KgwasHip(precision="fp32") as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        hits, npass = k.scan(bits, ScanParams(corr, l10, l2d))
        dev = k.to_device(bits)
        st = k.stats_device(dev.ptr, m, bits.shape[1])
        dev.free()
    assert abs(int(npass[0]) - cpu["n_passed"]) <= max(2, cpu["n_passed"] // 1000)
    kept = cpu["keep"]
    scale = np.maximum(st["c"][kept].astype(np.float64), 1.0)
    assert np.max(np.abs(st["xxM"][kept] - cpu["xxM"]) / scale) < 1e-4
    assert np.max(np.abs(st["xyM"][kept, 0] - cpu["xyM"]) / scale) < 1e-4
    assert abs(len(hits) - cpu["hits_1d"]) <= max(2, cpu["hits_1d"] // 100)


def test_synth_device_matches_host(built):
    from supergwas_amd.api import KgwasHip
    ds = dataset("n333", seed=5, n_acc=333, n_missing=3, npop=4, p=3, m=1000)
    rw = synth.row_words_for(ds.n_acc)
    m = 777
    with KgwasHip() as k:
        dev = k.alloc(m * rw * 8)
        k.synth_bits(dev.ptr, ds, 12345, m, rw)
        got = dev.to_host(np.uint64, (m, rw))
        dev.free()
    assert np.array_equal(got, ds.fill_bits(12345, m, rw))


def test_edge_cases(built):
    from supergwas_amd.api import KgwasHip, ScanParams
    ds = dataset("small_n190_p5", **CASES["small_n190_p5"][0])
    rw = synth.row_words_for(ds.n_acc)
    bits = ds.fill_bits(0, 300)
    # monomorphic rows over the phenotyped samples: all zero / all one
    bits[0, :] = 0
    allone = np.zeros(rw, np.uint64)
    for a in ds.idxs:
        allone[a >> 6] |= np.uint64(1) << np.uint64(a & 63)
    bits[1, :] = allone
    ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, 0.0)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        # empty scan
        h, npass = k.scan(bits[:0], ScanParams(0.2, 6, 15))
        assert len(h) == 0 and npass[0] == 0
        # tau = 0: every polymorphic k-mer passes; L = 0: every pass is reported
        h, npass = k.scan(bits, ScanParams(0.0, 0.0, 0.0))
        poly = (ref["c"] > 0) & (ref["c"] < ds.n)
        assert npass[0] == poly.sum() == ref["n_passed"]
        assert np.array_equal(h["idx"], np.nonzero(poly)[0])
        # one k-mer
        h, npass = k.scan(bits[5:6], ScanParams(0.0, 0.0, 0.0))
        assert len(h) == 1 and h["idx"][0] == 0 and h["c"][0] == ref["c"][5]
    # constant phenotype: |y - mean| = 0 -> nothing passes (src/kgwas.cpp:465)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, np.full(ds.n, 2.5))
        h, npass = k.scan(bits, ScanParams(0.2, 0.0, 0.0))
        assert npass[0] == 0 and len(h) == 0


def test_many_candidates_read_back(built):
    """tau = 0 and p <= 1 make every polymorphic k-mer a reported candidate.  90,000 of them are
    more than the 4,096 returned in place through page-locked memory (the rest are copied after
    the synchronisation), and more than the initial candidate capacity of 65,536 (the scan then
    reruns with a larger one).  Every k-mer comes back with the oracle's count and FP64
    statistic, and a later, smaller scan on the grown buffers reports the same values."""
    from supergwas_amd.api import KgwasHip, ScanParams
    ds = dataset("small_n190_p5", **CASES["small_n190_p5"][0])
    m = 90000
    bits = ds.fill_bits(0, m)
    ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, 0.0)
    poly = (ref["c"] > 0) & (ref["c"] < ds.n)
    print(f"[case] many candidates m={m} polymorphic={int(poly.sum())}", flush=True)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        h, npass = k.scan(bits, ScanParams(0.0, 0.0, 0.0))
        tim = k.timing()
        h2, _ = k.scan(bits[:5000], ScanParams(0.0, 0.0, 0.0))
    print(f"[case] many candidates: candidates={tim['candidates']} hits={len(h)}", flush=True)
    assert tim["candidates"] == poly.sum() > 65536
    assert npass[0] == poly.sum()
    assert np.array_equal(h["idx"], np.nonzero(poly)[0])
    assert np.array_equal(h["c"], ref["c"][poly])
    dT = np.abs(h["T"] - ref["T"][poly])
    assert np.all(dT <= REPORT_TOL * ref["T"][poly] + 1e-10), dT.max()
    first = h["idx"] < 5000
    assert np.array_equal(h2["idx"], h["idx"][first])
    assert np.array_equal(h2["T"], h["T"][first])


def test_multi_trait_equals_single_trait_runs(built):
    from supergwas_amd.api import KgwasHip, ScanParams
    ds = synth.make_dataset(seed=7, n_acc=512, n_missing=20, npop=4, p=8, m=4000, n_causal=10,
                            effect_sd=0.5, traits=3, trait_corr=0.2)
    bits = ds.fill_bits(0, 4000)
    prm = ScanParams(0.15, 2.0, 4.0)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
        hits, npass = k.scan(bits, prm)
    for tr in range(3):
        ref = oracle.scan_f64(bits, ds.idxs, ds.Y[:, tr], ds.U, 0.15)
        exp_idx, _ = expected_hits(ref, 2.0, 4.0)
        got = hits[hits["trait"] == tr]
        assert npass[tr] == ref["n_passed"]
        assert np.array_equal(got["idx"], exp_idx)
        assert np.allclose(got["T"], ref["T"][exp_idx], rtol=1e-9, atol=0)


@pytest.mark.parametrize("p,t,n_acc", [(40, 1, 1100), (56, 1, 777), (64, 8, 1100), (80, 4, 1300)])
def test_wide_tiles_match_oracle(built, p, t, n_acc):
    """Weight tiles past 128 columns (half- and quarter-chunk LDS stages) and the
    multi-trait epilogue: p + t up to 84 base columns at 3 limbs."""
    from supergwas_amd.api import KgwasHip, ScanParams
    m = 3000
    ds = synth.make_dataset(seed=40 + p, n_acc=n_acc, n_missing=37, npop=5, p=p, m=m,
                            n_causal=10, effect_sd=0.5, traits=t, trait_corr=0.2)
    bits = ds.fill_bits(0, m)
    prm = ScanParams(0.1, 2.0, 4.0)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
        hits, npass = k.scan(bits, prm)
        dev = k.to_device(bits)
        st = k.stats_device(dev.ptr, m, bits.shape[1])
        dev.free()
    print(f"[wide] p={p} t={t} n={ds.n} hits={len(hits)} passed={list(npass)}", flush=True)
    for tr in range(t):
        ref = oracle.scan_f64(bits, ds.idxs, ds.Y[:, tr], ds.U, 0.1)
        exp_idx, _ = expected_hits(ref, 2.0, 4.0)
        got = hits[hits["trait"] == tr]
        assert npass[tr] == ref["n_passed"], tr
        assert np.array_equal(got["idx"], exp_idx), tr
        assert np.allclose(got["T"], ref["T"][exp_idx], rtol=1e-9, atol=0)
        assert np.array_equal(st["c"], ref["c"])
        Tb = T_from(st["xxM"], st["xyM"][:, tr], oracle.yyM(ds.Y[:, tr], ds.U), ds.n)
        err = rel_err_T(Tb, ref["T"]).max()
        print(f"[wide] trait {tr}: max rel err T {err:.3e}", flush=True)
        assert err <= STAT_TOL, tr


def test_streamed_scan_matches_device_scan(built):
    """kg_scan streams host rows in 64 MiB chunks (copy of chunk i+1 overlapping
    the scan of chunk i), from pageable memory through bounce buffers and from
    page-locked memory directly; both equal the one-shot device scan."""
    from supergwas_amd.api import KgwasHip, ScanParams
    m = 600_000  # 2048 accessions: 256 B rows, 3 chunks
    ds = synth.make_dataset(seed=11, n_acc=2048, n_missing=30, npop=4, p=16, m=m, n_causal=20,
                            effect_sd=0.4)
    bits = ds.fill_bits(0, m)
    prm = ScanParams(0.2, 5.0, 9.0)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.to_device(bits)
        h_dev, n_dev = k.scan_device(dev.ptr, m, bits.shape[1], prm)
        dev.free()
        h_pg, n_pg = k.scan(bits, prm)
        tm_pg = k.timing()
        pin = k.host_alloc(bits.shape)
        pin[:] = bits
        h_pin, n_pin = k.scan(pin, prm)
        tm_pin = k.timing()
        k.host_free(pin)
    print(f"[stream] hits={len(h_dev)} passed={int(n_dev[0])} pageable copy {tm_pg['copy_ms']:.1f} ms"
          f" wall {tm_pg['wall_ms']:.1f} ms; pinned copy {tm_pin['copy_ms']:.1f} ms"
          f" wall {tm_pin['wall_ms']:.1f} ms", flush=True)
    assert len(h_dev) > 0
    for h, n in ((h_pg, n_pg), (h_pin, n_pin)):
        assert np.array_equal(n, n_dev)
        assert np.array_equal(h["idx"], h_dev["idx"])
        assert np.array_equal(h["T"], h_dev["T"])
    assert tm_pin["kmers"] == m and tm_pin["copy_ms"] > 0 and tm_pin["wall_ms"] > 0


def test_autotune_keeps_results(built):
    """kg_autotune only changes the launch shape: results are identical before and after."""
    from supergwas_amd.api import KgwasHip, ScanParams
    kw, m, corr, l10, l2d = CASES["n2048_p32"]
    ds = dataset("n2048_p32", **kw)
    bits = ds.fill_bits(0, m)
    prm = ScanParams(corr, l10, l2d)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        h0, n0 = k.scan(bits, prm)
        tune = k.autotune(300_000)
        h1, n1 = k.scan(bits, prm)
    print(f"[autotune] {tune}", flush=True)
    assert tune["tiles_per_wave"] in (2, 4) and tune["workgroups_per_cu"] in (1, 2, 4)
    assert tune["screen_wgs_per_cu"] in (1, 2, 3, 4)
    assert np.array_equal(n0, n1)
    assert np.array_equal(h0["idx"], h1["idx"]) and np.array_equal(h0["T"], h1["T"])


@pytest.mark.parametrize("t", [1, 3])
def test_every_launch_shape_gives_identical_results(built, t):
    """Every shape the autotuner can pick (tiles per wave x full-kernel workgroups per CU x
    screen workgroups per CU), screened and unscreened, with the resident-weight kernel and
    with streamed weights, reports the same hits, values and counts as the default shape."""
    from supergwas_amd.api import KgwasHip, ScanParams
    m = 20_000
    ds = synth.make_dataset(seed=80 + t, n_acc=2048, n_missing=40, npop=8, p=32, m=m,
                            n_causal=20, effect_sd=0.4, traits=t, trait_corr=0.2)
    bits = ds.fill_bits(0, m)
    prm = ScanParams(0.2, 3.0, 6.0)
    out = {}
    with KgwasHip(screen="on") as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
        dev = k.to_device(bits)
        for tiles in (0, 2):
            for gm in (1, 2, 4):
                for sw in (1, 2, 3):
                    for stw in (0, 1):
                        k.set_tune(tiles, gm, sw, stw)
                        for mode in ("on", "off"):
                            k.set_screen(mode)
                            out[(tiles, gm, sw, stw, mode)] = (
                                k.scan_device(dev.ptr, m, bits.shape[1], prm), k.get_tune())
        dev.free()
    (h0, n0), tune0 = out[(0, 1, 1, 0, "on")]
    assert len(h0) > 0
    # one trait at n=2048, p=32: the two-limb weights (128 KiB) sit in LDS
    assert tune0["resident_active"] == (1 if t == 1 else 0)
    assert out[(0, 1, 1, 1, "on")][1]["resident_active"] == 0
    kpw = set()
    for key, ((h, n), tune) in out.items():
        kpw.add((tune["kmers_per_workgroup"], tune["screen_kmers_per_workgroup"]))
        assert np.array_equal(n, n0), key
        for f in ("idx", "trait", "c", "T", "effect", "pval", "flags"):
            assert np.array_equal(h[f], h0[f]), (key, f)
    print(f"[shapes] t={t}: {len(out)} shapes, (k-mers per workgroup, screen) {sorted(kpw)}",
          flush=True)
    if t == 1:  # one trait: 4 or 2 tiles per wave, 16- or 8-wave screen workgroups, the ring
        assert len(kpw) >= 2


@pytest.mark.parametrize("name", ["n2048_p32", "n2048_p32_missing", "n1024_p20",
                                  "n333_nointercept"])
def test_resident_forms_identical(built, name):
    """The resident kernel's operand forms -- two int8 limbs, two FP6 digits against FP4 bits
    (with the three-limb recheck pass of the k-mers its bounds leave open), three FP6 digits --
    and the streaming kernel report the same hits, values and counts, equal to the FP64
    oracle outside the +-1e-6 band.  The FP6 kernel's allele count comes three ways: a bit count
    of the unmasked bits (n2048_p32: every accession phenotyped, 32 columns fill 4 tiles), of
    the masked bits (n2048_p32_missing: 60 accessions without phenotype), or the count column
    in a spare column slot (n1024_p20, n333_nointercept).  The p threshold sits on a passing k-mer's own p-value near
    T = 2, in the bulk of the statistic's distribution, so that many k-mers fall inside the
    two-digit bounds and the recheck pass has work."""
    from supergwas_amd.api import KgwasHip, ScanParams
    kw, m, corr, _, _ = CASES[name]
    ds = dataset(name, **kw)
    bits = ds.fill_bits(0, m)
    ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, corr)
    pas = np.nonzero(ref["passed"])[0]
    j = pas[np.argmin(np.abs(ref["T"][pas] - 2.0))]
    l10 = float(-np.log10(ref["p"][j]))
    l2d = l10 + 1.0
    prm = ScanParams(corr, l10, l2d)
    out = {}
    with KgwasHip(screen="on") as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.to_device(bits)
        for form in (2, 1, 3):
            for stw in (0, 1):
                k.set_tune(0, 1, 1, stw, form)
                out[(form, stw)] = (k.scan_device(dev.ptr, m, bits.shape[1], prm), k.timing(),
                                    k.get_tune())
        dev.free()
    (h0, n0), t0, tune0 = out[(2, 0)]
    print(f"[resident forms] {name} n={ds.n} p={ds.p} m={m} log10={l10:.6f} hits={len(h0)} "
          f"passed={int(n0[0])} survivors={t0['survivors']} rechecked={t0['rechecked']} "
          f"candidates={t0['candidates']}", flush=True)
    assert tune0["resident_form"] == 2 and tune0["resident_active"] == 1
    assert out[(1, 0)][2]["resident_form"] == 1 and out[(3, 0)][2]["resident_form"] == 3
    assert t0["rechecked"] > 0
    for (form, stw), ((h, n), tm, _) in out.items():
        if form != 2 or stw:
            assert tm["rechecked"] == 0, (form, stw)
        assert np.array_equal(n, n0), (form, stw)
        for f in ("idx", "trait", "c", "T", "effect", "pval", "flags"):
            assert np.array_equal(h[f], h0[f]), (form, stw, f)
    cor = np.abs(pearson(bits, ds.idxs, ds.y))
    with np.errstate(divide="ignore", invalid="ignore"):
        near_c = np.abs(cor / corr - 1.0) <= BAND
        near_p = np.abs(ref["p"] / 10.0 ** (-l10) - 1.0) <= BAND
    assert abs(int(n0[0]) - ref["n_passed"]) <= int(near_c.sum())
    exp_idx, _ = expected_hits(ref, l10, l2d)
    sym = np.setxor1d(exp_idx, h0["idx"])
    assert np.all((near_c | near_p)[sym]), sym[:10]
    assert np.array_equal(h0["c"], ref["c"][h0["idx"]])
    assert np.allclose(h0["T"], ref["T"][h0["idx"]], rtol=REPORT_TOL, atol=0)


def _shard_setup(k, nsh=5, seed=30):
    """Shards of uneven size, each with its own phenotyped-sample mask (own model)."""
    sizes = [1, 300, 0, 5003, 2500][:nsh]
    rng = np.random.default_rng(seed)
    base = synth.make_dataset(seed=seed, n_acc=640, n_missing=0, npop=4, p=8, m=10000,
                              n_causal=10, effect_sd=0.5)
    rw = synth.row_words_for(base.n_acc)
    shards, refs, devs = [], [], []
    k0 = 0
    for s, m in enumerate(sizes):
        miss = rng.choice(base.n_acc, size=int(rng.uniform(0, 0.15) * base.n_acc), replace=False)
        mask = np.ones(base.n_acc, bool)
        mask[miss] = False
        idxs = np.nonzero(mask)[0].astype(np.int32)
        # per-shard U: orthonormal basis of [1, PCs restricted to the shard's samples]
        Upcs = base.U[np.isin(base.idxs, idxs)]
        U, _, _ = np.linalg.svd(Upcs, full_matrices=False)
        y = rng.standard_normal(len(idxs)) + 0.2
        mid = k.add_model(base.n_acc, idxs, U, y) if s else None
        if s == 0:
            k.set_model(base.n_acc, idxs, U, y)
            mid = 0
        bits = base.fill_bits(k0, max(m, 1))[:m]
        k0 += m
        dev = k.to_device(bits) if m else k.alloc(rw * 8)
        devs.append(dev)
        shards.append((dev.ptr, m, rw, mid))
        refs.append((bits, idxs, y, U))
    return shards, refs, devs


def test_batched_shards_match_oracle_and_modes(built):
    from supergwas_amd.api import SHARDS_GRAPH, SHARDS_PER_LAUNCH, KgwasHip, ScanParams
    prm = ScanParams(0.1, 1.5, 3.0)
    with KgwasHip() as k:
        shards, refs, devs = _shard_setup(k)
        results = {}
        for name, flags in (("one", 0), ("per", SHARDS_PER_LAUNCH), ("graph", SHARDS_GRAPH),
                            ("graph_per", SHARDS_GRAPH | SHARDS_PER_LAUNCH)):
            reps = 3 if flags & SHARDS_GRAPH else 1
            for r in range(reps):   # graph replays, each checked
                hits, npass = k.scan_shards(shards, prm, flags)
                results[(name, r)] = (hits.copy(), npass.copy())
                print(f"[shards] mode={name} replay={r} hits={len(hits)}", flush=True)
        for d in devs:
            d.free()
    h0, n0 = results[("one", 0)]
    for key, (h, n) in results.items():
        assert np.array_equal(n, n0), key
        assert np.array_equal(h["idx"], h0["idx"]) and np.array_equal(h["shard"], h0["shard"]), key
        assert np.array_equal(h["T"], h0["T"]), key
    for s, (bits, idxs, y, U) in enumerate(refs):
        if len(bits) == 0:
            assert n0[s, 0] == 0 and not np.any(h0["shard"] == s)
            continue
        ref = oracle.scan_f64(bits, idxs, y, U, 0.1)
        exp, _ = expected_hits(ref, 1.5, 3.0)
        got = h0[h0["shard"] == s]
        assert n0[s, 0] == ref["n_passed"], s
        assert np.array_equal(got["idx"], exp), s
        assert np.allclose(got["T"], ref["T"][exp], rtol=1e-9, atol=0)


SCREEN_CASES = {
    # name: (dataset kwargs, m, corr, log10, log10_2d)
    "n2048_p32_t1": (dict(seed=2, n_acc=2048, n_missing=0, npop=8, p=32, m=40000, n_causal=50,
                          effect_sd=0.3), 40000, 0.2, 4.0, 8.0),
    "n700_p6_t2_missing": (dict(seed=51, n_acc=700, n_missing=41, npop=4, p=6, m=9001,
                                n_causal=10, effect_sd=0.5, traits=2, trait_corr=0.3),
                           9001, 0.12, 2.0, 4.0),
    "n1500_p12_t5": (dict(seed=52, n_acc=1500, n_missing=13, npop=6, p=12, m=7000, n_causal=10,
                          effect_sd=0.5, traits=5, trait_corr=0.2), 7000, 0.15, 2.0, 4.0),
    "n4096_p64_t8": (dict(seed=53, n_acc=4096, n_missing=0, npop=8, p=64, m=6000, n_causal=20,
                          effect_sd=0.3, traits=8, trait_corr=0.2, n_struct=512),
                     6000, 0.2, 3.0, 6.0),
    "n8192_p8_t1_high_tau": (dict(seed=54, n_acc=8192, n_missing=100, npop=8, p=8, m=3000,
                                  n_causal=5, effect_sd=0.3, n_struct=256), 3000, 0.6, 2.0, 4.0),
}


def test_bulk_statistic_n4096_p64_t8(built):
    """Full-statistic mode at the multi-trait shape (n=4096, p=64, 8 traits): every k-mer's
    statistic within 1e-5 of FP64 for every trait, allele counts bit-exact."""
    from supergwas_amd.api import KgwasHip
    kw, m, corr, l10, l2d = SCREEN_CASES["n4096_p64_t8"]
    ds = dataset("n4096_p64_t8", **kw)
    bits = ds.fill_bits(0, m)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
        dev = k.to_device(bits)
        st = k.stats_device(dev.ptr, m, bits.shape[1])
        refined = k.timing()["stat_refined"]
        dev.free()
    worst = 0.0
    for tr in range(ds.Y.shape[1]):
        ref = oracle.scan_f64(bits, ds.idxs, ds.Y[:, tr], ds.U, corr)
        assert np.array_equal(st["c"], ref["c"])
        Tb = T_from(st["xxM"], st["xyM"][:, tr], oracle.yyM(ds.Y[:, tr], ds.U), ds.n)
        worst = max(worst, rel_err_T(Tb, ref["T"]).max())
    print(f"[stat] n4096_p64_t8: max rel err T over 8 traits {worst:.3e}, "
          f"refined in FP64 {refined} of {8 * m}", flush=True)
    assert worst <= STAT_TOL


@pytest.mark.parametrize("name", list(SCREEN_CASES))
def test_screen_on_off_identical(built, name):
    """The pre-filter screen drops only k-mers whose pre-filter certainly fails,
    so screened and unscreened scans report the same hits, statistics and
    counts; the screen keeps at least every k-mer that passes for some trait."""
    from supergwas_amd.api import KgwasHip, ScanParams
    kw, m, corr, l10, l2d = SCREEN_CASES[name]
    ds = dataset(name, **kw)
    bits = ds.fill_bits(0, m)
    prm = ScanParams(corr, l10, l2d)
    out = {}
    for mode in ("off", "on"):
        with KgwasHip(screen=mode) as k:
            k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
            dev = k.to_device(bits)
            h, n = k.scan_device(dev.ptr, m, bits.shape[1], prm)
            out[mode] = (h, n, k.timing())
            dev.free()
    (h0, n0, t0), (h1, n1, t1) = out["off"], out["on"]
    print(f"[screen] {name} t={ds.Y.shape[1]} n={ds.n} m={m} hits={len(h1)} passed={list(n1)} "
          f"survivors={t1['survivors']} screen_ms={t1['screen_ms']:.3f}", flush=True)
    assert t0["survivors"] == m and t0["screen_ms"] == 0.0
    assert np.array_equal(n0, n1)
    assert np.array_equal(h0["idx"], h1["idx"]) and np.array_equal(h0["trait"], h1["trait"])
    assert np.array_equal(h0["c"], h1["c"]) and np.array_equal(h0["T"], h1["T"])
    # survivors cover the union of passes: at least the largest per-trait count
    assert int(n1.max()) <= t1["survivors"] <= m
    # against the oracle, trait by trait
    for tr in range(ds.Y.shape[1]):
        ref = oracle.scan_f64(bits, ds.idxs, ds.Y[:, tr], ds.U, corr)
        exp_idx, _ = expected_hits(ref, l10, l2d)
        assert n1[tr] == ref["n_passed"], tr
        assert np.array_equal(h1[h1["trait"] == tr]["idx"], exp_idx), tr


def test_screen_shards_with_empty_survivor_lists(built):
    """Batched shards where the screen keeps nothing of some shards (high
    threshold): the device-built pass-2 table skips them; one launch, HIP-graph
    replays and the unscreened per-shard launches agree."""
    from supergwas_amd.api import SHARDS_GRAPH, SHARDS_PER_LAUNCH, KgwasHip, ScanParams
    for corr in (0.08, 0.9):
        prm = ScanParams(corr, 1.0, 3.0)
        with KgwasHip(screen="on") as k:
            shards, refs, devs = _shard_setup(k, seed=31)
            res = []
            for flags in (0, SHARDS_GRAPH, SHARDS_GRAPH, SHARDS_PER_LAUNCH):
                h, n = k.scan_shards(shards, prm, flags)
                res.append((h.copy(), n.copy(), k.timing()["survivors"]))
            for d in devs:
                d.free()
        print(f"[screen shards] corr={corr} hits={len(res[0][0])} passed={res[0][1][:, 0].tolist()} "
              f"survivors={res[0][2]}", flush=True)
        h0, n0, _ = res[0]
        for h, n, _ in res[1:]:
            assert np.array_equal(n, n0)
            assert np.array_equal(h["idx"], h0["idx"]) and np.array_equal(h["shard"], h0["shard"])
            assert np.array_equal(h["T"], h0["T"])
        for s, (bits, idxs, y, U) in enumerate(refs):
            if len(bits) == 0:
                continue
            ref = oracle.scan_f64(bits, idxs, y, U, corr)
            assert n0[s, 0] == ref["n_passed"], (corr, s)
            exp, _ = expected_hits(ref, 1.0, 3.0)
            assert np.array_equal(h0[h0["shard"] == s]["idx"], exp), (corr, s)


def test_screen_at_the_threshold(built):
    """Thresholds placed exactly on k-mers' own FP64 |cor| (the k-mers the
    screen's margin exists for): screened, unscreened and auto scans agree
    bit for bit, and both match the FP64 oracle outside the +-1e-6 band."""
    from supergwas_amd.api import KgwasHip, ScanParams
    kw, m, _, l10, l2d = CASES["n2048_p32"]
    ds = dataset("n2048_p32", **kw)
    bits = ds.fill_bits(0, m)
    cor = np.abs(pearson(bits, ds.idxs, ds.y))
    order = np.argsort(-np.nan_to_num(cor))
    # taus at the |cor| of the 10th, 300th and 3000th strongest k-mers
    taus = [float(cor[order[j]]) for j in (9, 299, 2999)]
    with KgwasHip(screen="on") as kon, KgwasHip(screen="off") as koff, KgwasHip() as kauto:
        for k in (kon, koff, kauto):
            k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = kon.to_device(bits)
        for tau in taus:
            prm = ScanParams(tau, l10, l2d)
            h_on, n_on = kon.scan_device(dev.ptr, m, bits.shape[1], prm)
            surv = kon.timing()["survivors"]
            h_off, n_off = koff.scan(bits, prm)
            h_au, n_au = kauto.scan(bits, prm)
            ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, tau)
            print(f"[screen tau] tau={tau:.9f} passed={int(n_on[0])} oracle={ref['n_passed']} "
                  f"survivors={surv} hits={len(h_on)}", flush=True)
            for h, n in ((h_off, n_off), (h_au, n_au)):
                assert np.array_equal(n, n_on)
                assert np.array_equal(h["idx"], h_on["idx"]) and np.array_equal(h["T"], h_on["T"])
            with np.errstate(divide="ignore", invalid="ignore"):
                near_c = np.abs(cor / tau - 1.0) <= BAND
            assert abs(int(n_on[0]) - ref["n_passed"]) <= int(near_c.sum())
            assert surv >= int(n_on[0])
            exp_idx, _ = expected_hits(ref, l10, l2d)
            sym = np.setxor1d(exp_idx, h_on["idx"])
            assert np.all(near_c[sym]), sym[:10]
        dev.free()


def test_screen_auto_small_scans(built):
    """Auto mode on scans of up to 2^22 k-mers: the first scan at a new size runs unscreened to
    measure the full kernel, the next one is screened, and later ones follow whichever was
    faster (with a re-measurement every 32 scans).  Every scan reports what the forced modes
    report."""
    from supergwas_amd.api import KgwasHip, ScanParams
    kw, m, corr, l10, l2d = CASES["n2048_p32"]
    ds = dataset("n2048_p32", **kw)
    bits = ds.fill_bits(0, m)
    prm = ScanParams(corr, l10, l2d)
    with KgwasHip(screen="on") as kon:
        kon.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        h_on, n_on = kon.scan(bits, prm)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.to_device(bits)
        surv = []
        for _ in range(12):
            h, n = k.scan_device(dev.ptr, m, bits.shape[1], prm)
            surv.append(k.timing()["survivors"])
            assert np.array_equal(n, n_on)
            assert np.array_equal(h["idx"], h_on["idx"]) and np.array_equal(h["T"], h_on["T"])
        dev.free()
    print(f"[screen auto] m={m} survivors per scan {surv}", flush=True)
    assert surv[0] == m and surv[1] < m


def test_screen_models_of_different_width(built):
    """One shard table whose models have different accession counts (1100, 600
    and 2600: 3, 2 and 6 chunks of 512), interleaved, so the screen re-stages a
    differently sized LDS table between shards; screened, unscreened and the
    oracle agree."""
    from supergwas_amd.api import KgwasHip, ScanParams
    specs = [(1100, 2100, 61), (600, 1700, 62), (2600, 2400, 63), (1100, 900, 64)]
    prm = ScanParams(0.12, 2.0, 4.0)
    data = []
    for n_acc, m, seed in specs:
        ds = synth.make_dataset(seed=seed, n_acc=n_acc, n_missing=n_acc // 20, npop=4, p=7,
                                m=m, n_causal=8, effect_sd=0.5, n_struct=512)
        data.append((ds, ds.fill_bits(0, m)))
    res = {}
    for mode in ("on", "off"):
        with KgwasHip(screen=mode) as k:
            shards, devs = [], []
            for s, (ds, bits) in enumerate(data):
                mid = 0
                if s == 0:
                    k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
                else:
                    mid = k.add_model(ds.n_acc, ds.idxs, ds.U, ds.y)
                dev = k.to_device(bits)
                devs.append(dev)
                shards.append((dev.ptr, len(bits), bits.shape[1], mid))
            h, n = k.scan_shards(shards, prm, 0)
            res[mode] = (h.copy(), n.copy(), k.timing()["survivors"])
            for d in devs:
                d.free()
    (h1, n1, sv1), (h0, n0, sv0) = res["on"], res["off"]
    print(f"[screen widths] hits={len(h1)} passed={n1[:, 0].tolist()} survivors={sv1}", flush=True)
    assert np.array_equal(n1, n0)
    assert np.array_equal(h1["idx"], h0["idx"]) and np.array_equal(h1["shard"], h0["shard"])
    assert np.array_equal(h1["T"], h0["T"])
    assert sv1 < sum(len(b) for _, b in data)
    for s, (ds, bits) in enumerate(data):
        ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, 0.12)
        exp, _ = expected_hits(ref, 2.0, 4.0)
        assert n1[s, 0] == ref["n_passed"], s
        assert np.array_equal(h1[h1["shard"] == s]["idx"], exp), s


@pytest.mark.parametrize("n_acc,m,t", [(65, 1, 1), (65, 31, 1), (511, 33, 2), (513, 1000, 3),
                                       (1024, 4097, 1), (3000, 1537, 8), (2048, 70000, 1)])
def test_screen_odd_shapes(built, n_acc, m, t):
    """Bitmap words, compaction chunks and survivor blocks at awkward sizes:
    k-mer counts that are not multiples of 32 or of a screen block, accession
    counts just off a 512-sample chunk; screened == unscreened, and == oracle."""
    from supergwas_amd.api import KgwasHip, ScanParams
    ds = synth.make_dataset(seed=70 + n_acc + t, n_acc=n_acc, n_missing=n_acc // 13, npop=3,
                            p=4, m=m, n_causal=min(4, m), effect_sd=0.6, traits=t,
                            trait_corr=0.2, n_struct=256)
    bits = ds.fill_bits(0, m)
    prm = ScanParams(0.1, 1.5, 3.0)
    out = {}
    for mode in ("on", "off"):
        with KgwasHip(screen=mode) as k:
            k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
            h, n = k.scan(bits, prm)
            out[mode] = (h, n, k.timing()["survivors"])
    (h1, n1, sv), (h0, n0, _) = out["on"], out["off"]
    print(f"[odd] n_acc={n_acc} m={m} t={t} hits={len(h1)} passed={list(n1)} survivors={sv}",
          flush=True)
    assert np.array_equal(n1, n0)
    assert np.array_equal(h1["idx"], h0["idx"]) and np.array_equal(h1["trait"], h0["trait"])
    assert np.array_equal(h1["T"], h0["T"])
    for tr in range(t):
        ref = oracle.scan_f64(bits, ds.idxs, ds.Y[:, tr], ds.U, 0.1)
        exp, _ = expected_hits(ref, 1.5, 3.0)
        assert n1[tr] == ref["n_passed"], tr
        assert np.array_equal(h1[h1["trait"] == tr]["idx"], exp), tr


@pytest.mark.parametrize("n_acc", [2048, 4000])
def test_ring_screen_identical(built, n_acc):
    """The one-trait screen that streams its rows through an LDS ring (kg_tune screen shapes 3
    and 4) keeps exactly the k-mers the register-loaded screen (shape 1) keeps: same survivor
    counts, hits and pass counts, on uneven shards of two models with missing samples (4 and
    8 chunks, partial last blocks, an empty shard), with and without the duplicate-row check,
    and equal to the FP64 oracle per shard."""
    from supergwas_amd.api import KgwasHip, ScanParams
    from tests.test_gpu_dedup import planted_rows
    base = synth.make_dataset(seed=90 + n_acc % 7, n_acc=n_acc, n_missing=0, npop=6, p=12,
                              m=30000, n_causal=20, effect_sd=0.4)
    rng = np.random.default_rng(n_acc)
    rw = synth.row_words_for(n_acc)
    models = []
    for mi in range(2):
        miss = rng.choice(n_acc, size=int(0.02 * (mi + 1) * n_acc), replace=False)
        idxs = np.setdiff1d(np.arange(n_acc), miss).astype(np.int32)
        U, _, _ = np.linalg.svd(base.U[np.isin(base.idxs, idxs)], full_matrices=False)
        y = base.y[np.isin(base.idxs, idxs)] + 0.3 * rng.standard_normal(len(idxs))
        models.append((idxs, U, y))
    sizes = [7001, 1, 0, 12345, 4096, 577]
    rows_all = planted_rows(base, 12000, seed=n_acc)[: sum(sizes)]
    prm = ScanParams(0.15, 2.0, 3.0)
    out = {}
    with KgwasHip(screen="on") as k:
        k.set_model(n_acc, *models[0])
        k.add_model(n_acc, *models[1])
        shards, devs, refs, k0 = [], [], [], 0
        for si, m in enumerate(sizes):
            bits = rows_all[k0:k0 + m]
            k0 += m
            dev = k.to_device(bits) if m else k.alloc(rw * 8)
            devs.append(dev)
            shards.append((dev.ptr, m, rw, si % 2))
            refs.append(bits)
        shapes = (1, 3, 4) if n_acc <= 2048 else (1, 3)
        for dd in (False, True):
            k.set_dedup(dd)
            for sw in shapes:
                k.set_tune(0, 1, sw, 0)
                h, n = k.scan_shards(shards, prm)
                out[(dd, sw)] = (h.copy(), n.copy(), k.timing()["survivors"], k.get_tune())
        for d in devs:
            d.free()
    h0, n0, s0, _ = out[(False, 1)]
    print(f"[ring screen] n_acc={n_acc} hits={len(h0)} passed={n0[:, 0].tolist()} survivors={s0}",
          flush=True)
    assert out[(False, 3)][3]["screen_wgs_per_cu"] == 3
    for (dd, sw), (h, n, sv, tune) in out.items():
        assert np.array_equal(n, n0), (dd, sw)
        if not dd:
            assert sv == s0, (dd, sw, sv, s0)
        else:
            assert sv == out[(True, 1)][2], (dd, sw)
        for f in ("idx", "shard", "c", "T", "flags"):
            assert np.array_equal(h[f], h0[f]), (dd, sw, f)
    for si, bits in enumerate(refs):
        if len(bits) == 0:
            continue
        idxs, U, y = models[si % 2]
        ref = oracle.scan_f64(bits, idxs, y, U, 0.15)
        assert n0[si, 0] == ref["n_passed"], si
        exp, _ = expected_hits(ref, 2.0, 3.0)
        assert np.array_equal(h0[h0["shard"] == si]["idx"], exp), si
