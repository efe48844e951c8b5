"""GPU parity at the shapes of the benchmark configurations and at the model limits.

  * the plain FP32 kernel at n=1024, p=20 (the reference's cols arithmetic,
    src/kgwas.cpp:146-158) against the FP64 oracle: exact allele counts, hits
    equal outside the band the FP32 error bounds define, the statistic's
    checksum against FP64;
  * batched uneven shards at n=2048, p=32, each with its own missing-phenotype
    mask and U (the reference runs one kgwas process per batch,
    scripts/run_pipeline.sh:243-253, each with its own overlap,
    src/kgwas_utils.hpp:291-314), in the persistent, HIP-graph and per-shard
    launch modes;
  * models past the screen's 8192 accessions (the unscreened fallback over more
    than 16 chunks of 512 samples) and a U.bin with r = 0
    (src/kgwas_utils.hpp:135-157 accepts it);
  * the headline workload's shape (n=2048, p=32) at 1.2e6 k-mers, and three
    traits at 4e5 k-mers, where every workgroup of the persistent kernels walks
    several blocks, against the oracle.
"""
import numpy as np
import pytest

import oracle
from supergwas_amd import synth
from tests.helpers import (T_from, T_interval, expected_hits, fp32_error_bounds, pearson,
                           rel_err_T)

pytestmark = pytest.mark.gpu

STAT_TOL = 1e-5
BAND = 1e-6


def test_fp32_kernel_at_n1024_p20(built):
    """Plain FP32 kernel at the n=1024, p=20 workload (2e4 k-mers, hits at p < 1e-5)."""
    from supergwas_amd.api import KgwasHip, ScanParams
    m, corr, l10 = 20_000, 0.2, 5.0
    ds = synth.make_dataset(seed=2, **dict(synth.CONFIGS["fp32_scale"], m=m))
    bits = ds.fill_bits(0, m)
    ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, corr)
    with KgwasHip(precision="fp32") as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        hits, npass = k.scan(bits, ScanParams(corr, l10, 0.0))
        dev = k.to_device(bits)
        st = k.stats_device(dev.ptr, m, bits.shape[1])
        dev.free()
    assert np.array_equal(st["c"], ref["c"]), "allele counts must be bit-exact"
    Bxx, Bxy, Bcor = fp32_error_bounds(bits, ds.idxs, ds.y, ds.U, ds.n_acc)
    yyM = oracle.yyM(ds.y, ds.U)
    # the kernel's FP32 values lie inside their bounds
    assert np.all(np.abs(st["xxM"] - ref["xxM"]) <= Bxx)
    assert np.all(np.abs(st["xyM"][:, 0] - ref["xyM"]) <= Bxy)
    # decisions: a k-mer may differ from the FP64 oracle only inside the band
    cor = np.abs(np.nan_to_num(pearson(bits, ds.idxs, ds.y)))
    tau32 = float(np.float32(corr))
    amb_c = (cor + Bcor >= min(corr, tau32) * (1 - BAND)) & \
            (cor - Bcor <= max(corr, tau32) * (1 + BAND))
    T_lo, T_hi = T_interval(ref["xxM"], ref["xyM"], yyM, ds.n, Bxx, Bxy)
    Ts = oracle.T_threshold(10.0 ** -l10)
    amb_p = (T_hi >= Ts * (1 - BAND)) & (T_lo <= Ts * (1 + BAND))
    exp_idx, _ = expected_hits(ref, l10, 0.0)
    sym = np.setxor1d(exp_idx, hits["idx"])
    print(f"[c2 fp32] n={ds.n} p={ds.p} m={m} hits={len(hits)} oracle={len(exp_idx)} "
          f"flips={len(sym)} in-band pre-filter={int(amb_c.sum())} in-band p={int(amb_p.sum())} "
          f"passed={int(npass[0])} oracle_passed={ref['n_passed']}", flush=True)
    assert len(exp_idx) > 0
    assert np.all(amb_c[sym] | amb_p[sym]), f"flips outside the FP32 band: {sym[:10]}"
    assert abs(int(npass[0]) - ref["n_passed"]) <= int(amb_c.sum())
    common, ie, ig = np.intersect1d(exp_idx, hits["idx"], return_indices=True)
    assert np.array_equal(hits["c"][ig], ref["c"][common])
    # checksum of the statistic over every k-mer against FP64, inside the summed bounds
    T32 = T_from(st["xxM"], st["xyM"][:, 0], yyM, ds.n)
    sT32, sT64 = float(T32.sum()), float(ref["T"].sum())
    slack = float(np.sum(np.maximum(T_hi - ref["T"], ref["T"] - T_lo)))
    print(f"[c2 fp32] sum T fp32 {sT32:.9e} fp64 {sT64:.9e} rel diff "
          f"{abs(sT32 - sT64) / sT64:.3e}; max rel err T {rel_err_T(T32, ref['T']).max():.3e}",
          flush=True)
    assert abs(sT32 - sT64) <= slack


def _uneven_shards(k, base, nsh, lo, hi, seed):
    """nsh shards of log-uniform size in [lo, hi] k-mers of `base`'s workload, each with its
    own 0-15% missing-phenotype mask, its own U (orthonormal basis of the base U restricted
    to the shard's samples) and phenotype; returns (shards, refs, devs)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), nsh)).astype(np.int64)
    rw = synth.row_words_for(base.n_acc)
    shards, refs, devs = [], [], []
    k0 = 0
    for s, m in enumerate(sizes):
        keep = np.ones(base.n, bool)
        keep[rng.choice(base.n, size=int(rng.uniform(0, 0.15) * base.n), replace=False)] = False
        idxs = base.idxs[keep]
        U, _, _ = np.linalg.svd(base.U[keep], full_matrices=False)
        y = base.y[keep]
        if s == 0:
            k.set_model(base.n_acc, idxs, U, y)
            mid = 0
        else:
            mid = k.add_model(base.n_acc, idxs, U, y)
        bits = base.fill_bits(k0, int(m))
        k0 += int(m)
        dev = k.to_device(bits)
        devs.append(dev)
        shards.append((dev.ptr, int(m), rw, mid))
        refs.append((bits, idxs, y, U))
    return shards, refs, devs


def test_uneven_shards_at_n2048_p32(built):
    """24 shards log-uniform in [1e3, 3e4] k-mers at n=2048, p=32 with per-shard masks and
    models: per-shard pre-filter counts and hits equal to the FP64 oracle outside the band,
    in one persistent launch (screened), its HIP-graph replays and one launch per shard."""
    from supergwas_amd.api import SHARDS_GRAPH, SHARDS_PER_LAUNCH, KgwasHip, ScanParams
    corr, l10 = 0.2, 3.0
    base = synth.make_dataset(seed=0, **dict(synth.CONFIGS["mixed"], m=1_000_000))
    prm = ScanParams(corr, l10, 6.0)
    results = {}
    # screen on: auto mode would measure the first scan of a new size unscreened
    with KgwasHip(screen="on") as k:
        shards, refs, devs = _uneven_shards(k, base, 24, 1e3, 3e4, seed=4)
        total = sum(s[1] for s in shards)
        for name, flags in (("persistent", 0), ("graph", SHARDS_GRAPH),
                            ("per_shard", SHARDS_PER_LAUNCH)):
            for r in range(3 if flags & SHARDS_GRAPH else 1):
                h, n = k.scan_shards(shards, prm, flags)
                results[(name, r)] = (h.copy(), n.copy(), k.timing()["survivors"])
                print(f"[c4] mode={name} replay={r} shards={len(shards)} k-mers={total} "
                      f"hits={len(h)} survivors={results[(name, r)][2]}", flush=True)
        for d in devs:
            d.free()
    h0, n0, _ = results[("persistent", 0)]
    for key, (h, n, _) in results.items():
        assert np.array_equal(n, n0), key
        assert np.array_equal(h["shard"], h0["shard"]) and np.array_equal(h["idx"], h0["idx"]), key
        assert np.array_equal(h["T"], h0["T"]), key
    assert results[("persistent", 0)][2] < total      # the screen ran
    nhits = 0
    for s, (bits, idxs, y, U) in enumerate(refs):
        ref = oracle.scan_f64(bits, idxs, y, U, corr)
        cor = np.abs(np.nan_to_num(pearson(bits, idxs, y)))
        near_c = np.abs(cor / corr - 1.0) <= BAND
        with np.errstate(divide="ignore", invalid="ignore"):
            near_p = np.abs(ref["p"] / 10.0 ** -l10 - 1.0) <= BAND
        assert abs(int(n0[s, 0]) - ref["n_passed"]) <= int(near_c.sum()), s
        exp, exp2d = expected_hits(ref, l10, 6.0)
        got = h0[h0["shard"] == s]
        sym = np.setxor1d(exp, got["idx"])
        assert np.all(near_c[sym] | near_p[sym]), (s, sym[:10])
        common, ie, ig = np.intersect1d(exp, got["idx"], return_indices=True)
        assert np.array_equal(got["c"][ig], ref["c"][common]), s
        assert np.allclose(got["T"][ig], ref["T"][common], rtol=1e-9, atol=0), s
        assert np.array_equal((got["flags"][ig] & 1).astype(bool), exp2d[ie]), s
        nhits += len(exp)
    assert nhits > 0


@pytest.mark.parametrize("t", [1, 3])
def test_wide_model_unscreened_fallback(built, t):
    """n_acc = 9500 (19 chunks of 512): the screen's LDS table covers at most 8192
    accessions, so the scan runs the full kernel on every k-mer; counts, hits and the bulk
    statistic against the oracle."""
    from supergwas_amd.api import KgwasHip, ScanParams
    m, corr, l10 = 4000, 0.2, 3.0
    ds = synth.make_dataset(seed=90 + t, n_acc=9500, n_missing=150, npop=6, p=8, m=m,
                            n_causal=10, effect_sd=0.3, traits=t, trait_corr=0.2, n_struct=512)
    bits = ds.fill_bits(0, m)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
        hits, npass = k.scan(bits, ScanParams(corr, l10, 6.0))
        tim = k.timing()
        dev = k.to_device(bits)
        st = k.stats_device(dev.ptr, m, bits.shape[1])
        dev.free()
    print(f"[wide n] n_acc={ds.n_acc} n={ds.n} t={t} hits={len(hits)} passed={list(npass)} "
          f"survivors={tim['survivors']}", flush=True)
    assert tim["survivors"] == m and tim["screen_ms"] == 0.0
    for tr in range(t):
        ref = oracle.scan_f64(bits, ds.idxs, ds.Y[:, tr], ds.U, corr)
        exp, _ = expected_hits(ref, l10, 6.0)
        got = hits[hits["trait"] == tr]
        assert npass[tr] == ref["n_passed"], tr
        assert np.array_equal(got["idx"], exp), tr
        assert np.allclose(got["T"], ref["T"][exp], rtol=1e-9, atol=0), tr
        assert np.array_equal(st["c"], ref["c"])
        Tb = T_from(st["xxM"], st["xyM"][:, tr], oracle.yyM(ds.Y[:, tr], ds.U), ds.n)
        err = rel_err_T(Tb, ref["T"]).max()
        print(f"[wide n] trait {tr}: bulk max rel err T {err:.3e}", flush=True)
        assert err <= STAT_TOL, tr


def test_model_without_structure_columns(built):
    """U.bin with r = 0 (src/kgwas_utils.hpp:135-157 reads it): xxM = c, xyM = x'y,
    yyM = y'y; screened and unscreened scans against the oracle."""
    from supergwas_amd.api import KgwasHip, ScanParams
    m, corr, l10 = 6000, 0.1, 2.0
    ds = synth.make_dataset(seed=95, n_acc=700, n_missing=20, npop=4, p=3, m=m, n_causal=10,
                            effect_sd=0.5)
    bits = ds.fill_bits(0, m)
    U0 = np.zeros((ds.n, 0))
    ref = oracle.scan_f64(bits, ds.idxs, ds.y, U0, corr)
    out = {}
    for mode in ("on", "off"):
        with KgwasHip(screen=mode) as k:
            k.set_model(ds.n_acc, ds.idxs, U0, ds.y)
            out[mode] = k.scan(bits, ScanParams(corr, l10, 4.0))
            dev = k.to_device(bits)
            st = k.stats_device(dev.ptr, m, bits.shape[1])
            dev.free()
    exp, _ = expected_hits(ref, l10, 4.0)
    print(f"[p=0] n={ds.n} hits={len(out['on'][0])} oracle={len(exp)} passed={int(out['on'][1][0])}",
          flush=True)
    for mode, (h, n) in out.items():
        assert n[0] == ref["n_passed"], mode
        assert np.array_equal(h["idx"], exp), mode
        assert np.allclose(h["T"], ref["T"][exp], rtol=1e-9, atol=0), mode
    assert np.array_equal(st["c"], ref["c"])
    Tb = T_from(st["xxM"], st["xyM"][:, 0], oracle.yyM(ds.y, U0), ds.n)
    assert rel_err_T(Tb, ref["T"]).max() <= STAT_TOL


def test_headline_shape_over_many_blocks(built):
    """n=2048, p=32, 1.2e6 k-mers in one screened scan: every workgroup of the persistent
    kernels walks several blocks (about 9 screen blocks and 1-2 survivor blocks each), so the
    ring's refills across blocks, the survivor kernel's next-block row loads and the recheck
    pass all run as at the benchmark's size.  Pre-filter count and hits against the FP64
    oracle outside the +-1e-6 band (the oracle in pieces of 1e5 k-mers; the |cor| band from
    oracle runs at tau (1 -+ 1e-6))."""
    from supergwas_amd.api import KgwasHip, ScanParams
    m, corr, l10, l2d = 1_200_000, 0.2, 4.0, 8.0
    ds = synth.make_dataset(seed=11, **dict(synth.CONFIGS["mixed"], m=m))
    bits = ds.fill_bits(0, m)
    prm = ScanParams(corr, l10, l2d)
    with KgwasHip(screen="on") as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.to_device(bits)
        h, n = k.scan_device(dev.ptr, m, bits.shape[1], prm)
        tm, tune = k.timing(), k.get_tune()
        dev.free()
    print(f"[headline shape] m={m} hits={len(h)} passed={int(n[0])} survivors={tm['survivors']} "
          f"rechecked={tm['rechecked']} candidates={tm['candidates']}", flush=True)
    assert tune["resident_active"] == 1 and tune["screen_wgs_per_cu"] == 3
    assert tm["survivors"] < m // 4
    exp_all, near_all, c_all, T_all = [], np.zeros(m, bool), np.zeros(m, np.int64), np.zeros(m)
    npass, nnear_c = 0, 0
    for a in range(0, m, 100_000):
        b = bits[a:a + 100_000]
        ref = oracle.scan_f64(b, ds.idxs, ds.y, ds.U, corr)
        lo = oracle.scan_f64(b, ds.idxs, ds.y, ds.U, corr * (1.0 - BAND))["passed"]
        hi = oracle.scan_f64(b, ds.idxs, ds.y, ds.U, corr * (1.0 + BAND))["passed"]
        near_c = lo & ~hi
        with np.errstate(divide="ignore", invalid="ignore"):
            near_p = np.abs(ref["p"] / 10.0 ** -l10 - 1.0) <= BAND
        exp, _ = expected_hits(ref, l10, l2d)
        exp_all.append(exp + a)
        near_all[a:a + len(b)] = near_c | near_p
        c_all[a:a + len(b)] = ref["c"]
        T_all[a:a + len(b)] = ref["T"]
        npass += int(ref["n_passed"])
        nnear_c += int(near_c.sum())
    exp_all = np.concatenate(exp_all)
    assert abs(int(n[0]) - npass) <= nnear_c
    sym = np.setxor1d(exp_all, h["idx"])
    assert np.all(near_all[sym]), sym[:10]
    assert np.array_equal(h["c"], c_all[h["idx"]])
    assert np.allclose(h["T"], T_all[h["idx"]], rtol=1e-9, atol=0)
    assert len(exp_all) > 10


def test_three_traits_over_many_blocks(built):
    """Three traits at n=2048, p=32, 4e5 k-mers in one screened scan at tau = 0.1 (most
    k-mers survive for some trait): the multi-trait screen and the streaming full kernel
    (weight stages through LDS) walk several blocks per workgroup, so the next block's first
    bits chunk and weight stage load during a block's last stage.  Per-trait pre-filter
    counts and hits against the FP64 oracle outside the +-1e-6 band."""
    from supergwas_amd.api import KgwasHip, ScanParams
    m, corr, l10, l2d = 400_000, 0.1, 4.0, 8.0
    ds = synth.make_dataset(seed=12, n_acc=2048, n_missing=0, npop=8, p=32, m=m, n_causal=50,
                            effect_sd=0.3, traits=3, trait_corr=0.2)
    bits = ds.fill_bits(0, m)
    prm = ScanParams(corr, l10, l2d)
    with KgwasHip(screen="on") as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
        dev = k.to_device(bits)
        h, n = k.scan_device(dev.ptr, m, bits.shape[1], prm)
        tm = k.timing()
        dev.free()
    print(f"[three traits] m={m} hits={len(h)} passed={list(map(int, n))} "
          f"survivors={tm['survivors']} candidates={tm['candidates']}", flush=True)
    assert 256 * 256 * 2 < tm["survivors"] < m   # more than two 256-k-mer blocks per workgroup
    for tr in range(3):
        y = ds.Y[:, tr]
        ref = oracle.scan_f64(bits, ds.idxs, y, ds.U, corr)
        lo = oracle.scan_f64(bits, ds.idxs, y, ds.U, corr * (1.0 - BAND))["passed"]
        hi = oracle.scan_f64(bits, ds.idxs, y, ds.U, corr * (1.0 + BAND))["passed"]
        near_c = lo & ~hi
        with np.errstate(divide="ignore", invalid="ignore"):
            near_p = np.abs(ref["p"] / 10.0 ** -l10 - 1.0) <= BAND
        assert abs(int(n[tr]) - ref["n_passed"]) <= int(near_c.sum()), tr
        exp, _ = expected_hits(ref, l10, l2d)
        got = h[h["trait"] == tr]
        sym = np.setxor1d(exp, got["idx"])
        assert np.all((near_c | near_p)[sym]), (tr, sym[:10])
        assert np.array_equal(got["c"], ref["c"][got["idx"]]), tr
        assert np.allclose(got["T"], ref["T"][got["idx"]], rtol=1e-9, atol=0), tr
        assert len(exp) > 0, tr
