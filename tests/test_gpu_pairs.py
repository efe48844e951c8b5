"""GPU parity of the all-pairs scan (kg_scan_pairs) against the FP64 oracle.

The scan replaces emit_pairs_self (src/kgwas_2d.cpp:258-401): every row pair
i < j is tested as x_i AND x_j with the 1D statistic, no pre-filter.  Bar:
reported pairs identical to the oracle's outside a +-1e-6 relative band
around the p threshold, allele counts bit-exact, reported T and effect equal
to FP64 (they are FP64-refined).
"""
import numpy as np
import pytest

import oracle
from supergwas_amd import synth

pytestmark = pytest.mark.gpu

REPORT_TOL = 1e-9
BAND = 1e-6

PAIR_CASES = {
    # name: (dataset kwargs, rows N, log10_thresh)
    "n190_p5": (dict(seed=40, n_acc=200, n_missing=10, npop=3, p=5, m=300, n_causal=5,
                     effect=0.5), 300, 2.0),
    "n2048_p32": (dict(seed=41, n_acc=2048, n_missing=0, npop=8, p=32, m=160, n_causal=8,
                       effect_sd=0.3), 160, 2.0),
    "n333_nointercept": (dict(seed=42, n_acc=333, n_missing=7, npop=4, p=6, m=257, n_causal=5,
                              effect=0.5, intercept=False), 257, 1.5),
    "n600_p40": (dict(seed=43, n_acc=600, n_missing=20, npop=5, p=40, m=130, n_causal=5,
                      effect=0.5), 130, 2.0),
}

_DS = {}


def _case(name):
    kw, N, l10 = PAIR_CASES[name]
    if name not in _DS:
        ds = synth.make_dataset(**kw)
        _DS[name] = (ds, ds.fill_bits(0, N))
    ds, bits = _DS[name]
    return ds, bits, l10


def _scan_pairs(ds, bits, l10, ranges=None, **kw):
    from supergwas_amd.api import KgwasHip, ScanParams
    with KgwasHip(**kw) as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.to_device(bits)
        prm = ScanParams(0.2, l10, 0.0)
        if ranges is None:
            out = k.scan_pairs(dev.ptr, bits.shape[0], bits.shape[1], prm)
        else:
            out = [k.scan_pairs(dev.ptr, bits.shape[0], bits.shape[1], prm, i0, i1)
                   for (i0, i1) in ranges]
        tim = k.timing()
        dev.free()
    return out, tim


def _compare(hits, ref, l10):
    got = hits["shard"].astype(np.int64) * 1_000_000 + hits["idx"]
    exp = ref["i"] * 1_000_000 + ref["j"]
    assert np.all(np.diff(got) > 0), "pairs must be sorted by (i, j)"
    sym = np.setxor1d(exp, got)
    if len(sym):
        # a flip is allowed only for pairs whose FP64 p sits inside the band
        pv = {int(e): p for e, p in zip(exp, ref["p"])}
        for e in sym:
            p = pv.get(int(e))
            assert p is not None and abs(p / 10.0 ** (-l10) - 1.0) <= BAND, f"flip {divmod(e, 1_000_000)}"
    common, ie, ig = np.intersect1d(exp, got, return_indices=True)
    h = hits[ig]
    assert np.array_equal(h["c"], ref["c"][ie])
    for key, rk in (("T", "T"), ("effect", "effect")):
        err = np.abs(h[key] - ref[rk][ie]) / np.maximum(np.abs(ref[rk][ie]), 1e-300)
        assert err.max(initial=0) <= REPORT_TOL, (key, err.max())
    perr = np.abs(h["pval"] - ref["p"][ie]) / np.maximum(ref["p"][ie], 1e-300)
    assert perr.max(initial=0) <= 1e-7
    return len(common)


@pytest.mark.parametrize("name", list(PAIR_CASES))
def test_pairs_match_oracle(built, name):
    ds, bits, l10 = _case(name)
    print(f"[pairs] {name} n={ds.n} p={ds.p} N={len(bits)}", flush=True)
    ref = oracle.pairs_f64(bits, ds.idxs, ds.y, ds.U, l10)
    hits, tim = _scan_pairs(ds, bits, l10)
    k = _compare(hits, ref, l10)
    print(f"[pairs] {name} hits={len(hits)} oracle={len(ref['i'])} common={k} "
          f"candidates={tim['candidates']} pairs={tim['kmers']}", flush=True)
    assert len(ref["i"]) > 10
    N = len(bits)
    assert tim["kmers"] == N * (N - 1) // 2


def test_pairs_row_ranges(built):
    ds, bits, l10 = _case("n190_p5")
    N = len(bits)
    cuts = [0, 1, 37, 150, 298, N]
    ranges = list(zip(cuts[:-1], cuts[1:])) + [(5, 5), (N - 1, N)]
    print(f"[pairs] ranges {ranges}", flush=True)
    (whole,), _ = _scan_pairs(ds, bits, l10, ranges=[(0, N)])
    parts, _ = _scan_pairs(ds, bits, l10, ranges=ranges)
    joined = np.concatenate(parts[:len(cuts) - 1])
    assert np.array_equal(joined["shard"], whole["shard"])
    assert np.array_equal(joined["idx"], whole["idx"])
    assert np.array_equal(joined["T"], whole["T"])
    assert len(parts[-2]) == 0 and len(parts[-1]) == 0
    for (i0, i1), part in zip(ranges, parts):
        assert np.all((part["shard"] >= i0) & (part["shard"] < i1))


def test_pairs_every_testable_pair_and_full_rows(built):
    """p_thresh = 1 reports every pair with xxM, yyM > 1e-12; a pair carried by
    every phenotyped sample (c = n) is tested when 1 is outside span(U)."""
    ds, bits, _ = _case("n333_nointercept")
    rows = bits[:40].copy()
    for r in (3, 7):                       # carried by every phenotyped sample
        for a in ds.idxs:
            rows[r, a >> 6] |= np.uint64(1) << np.uint64(a & 63)
    ref = oracle.pairs_f64(rows, ds.idxs, ds.y, ds.U, 0.0)
    hits, _ = _scan_pairs(ds, rows, 0.0)
    print(f"[pairs] p<=1: hits={len(hits)} oracle={len(ref['i'])}", flush=True)
    assert len(ref["i"]) > 700
    assert _compare(hits, ref, 0.0) == len(ref["i"]) == len(hits)
    full = (hits["shard"] == 3) & (hits["idx"] == 7)
    assert full.sum() == 1 and hits["c"][full][0] == ds.n


def test_pairs_equal_scan_of_and_rows(built):
    """The pair (i, j) is scanned as the row x_i AND x_j: kg_scan_device over
    materialised AND rows (no pre-filter) reports the same values."""
    from supergwas_amd.api import KgwasHip, ScanParams
    ds, bits, l10 = _case("n2048_p32")
    hits, _ = _scan_pairs(ds, bits, l10)
    ii, jj = hits["shard"].astype(np.int64), hits["idx"]
    and_rows = np.ascontiguousarray(bits[ii] & bits[jj])
    with KgwasHip(screen="off") as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        one, _ = k.scan(and_rows, ScanParams(0.0, l10, 0.0))
    assert np.array_equal(one["idx"], np.arange(len(hits)))
    assert np.array_equal(one["c"], hits["c"])
    assert np.array_equal(one["T"], hits["T"])
    assert np.array_equal(one["effect"], hits["effect"])


def test_pairs_argument_errors(built):
    from supergwas_amd.api import KgError, KgwasHip, ScanParams
    ds, bits, l10 = _case("n190_p5")
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.to_device(bits)
        N, rw = bits.shape
        for (i0, i1) in ((-1, 5), (6, 5), (0, N + 1)):
            with pytest.raises(KgError):
                k.scan_pairs(dev.ptr, N, rw, ScanParams(0.2, l10, 0.0), i0, i1)
        assert len(k.scan_pairs(dev.ptr, 1, rw, ScanParams(0.2, l10, 0.0))) == 0
        assert len(k.scan_pairs(dev.ptr, 0, rw, ScanParams(0.2, l10, 0.0))) == 0
        dev.free()
    with KgwasHip(precision="fp32") as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.to_device(bits)
        with pytest.raises(KgError):
            k.scan_pairs(dev.ptr, 4, bits.shape[1], ScanParams(0.2, l10, 0.0))
        dev.free()


def test_pairs_across_batches(built):
    """25,000 rows = 3.1e8 pairs: the scan runs as two batches of rows i (at most ~2.7e8
    pairs each), each with its own shard table and pair masks.  The result equals row-range
    scans that each fit one batch, and the pairs of sampled rows i equal the oracle's."""
    ds = synth.make_dataset(seed=44, n_acc=200, n_missing=10, npop=3, p=5, m=25000, n_causal=5,
                            effect=0.5)
    N = 25000
    bits = ds.fill_bits(0, N)
    l10 = 4.0
    print(f"[pairs] batches N={N} pairs={N * (N - 1) // 2}", flush=True)
    (whole,), tim = _scan_pairs(ds, bits, l10, ranges=[(0, N)])
    parts, _ = _scan_pairs(ds, bits, l10, ranges=[(0, 3000), (3000, N)])
    joined = np.concatenate(parts)
    print(f"[pairs] batches hits={len(whole)} candidates={tim['candidates']}", flush=True)
    assert tim["kmers"] == N * (N - 1) // 2
    assert len(whole) > 100
    for key in ("shard", "idx", "c", "T", "effect", "pval"):
        assert np.array_equal(joined[key], whole[key]), key
    for i in (0, 1, 2999, 3000, 12345, 24990, 24998):
        ref = oracle.pairs_f64(bits, ds.idxs, ds.y, ds.U, l10, i, i + 1)
        got = whole[whole["shard"] == i]
        assert np.array_equal(got["idx"], ref["j"]), i
        assert np.array_equal(got["c"], ref["c"]), i
        if len(got):
            assert np.max(np.abs(got["T"] - ref["T"]) / ref["T"]) <= REPORT_TOL


def test_pairs_multi_trait(built):
    """A model with 3 traits: each pair is tested per trait (hits carry the trait), and
    every trait's pairs equal the oracle's for that phenotype column."""
    ds = synth.make_dataset(seed=45, n_acc=300, n_missing=12, npop=4, p=6, m=140, n_causal=5,
                            effect=0.6, traits=3, trait_corr=0.3)
    bits = ds.fill_bits(0, 140)
    l10 = 2.0
    from supergwas_amd.api import KgwasHip, ScanParams
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.Y)
        dev = k.to_device(bits)
        hits = k.scan_pairs(dev.ptr, bits.shape[0], bits.shape[1], ScanParams(0.2, l10, 0.0))
        dev.free()
    print(f"[pairs] 3 traits: hits={len(hits)}", flush=True)
    key = (hits["shard"].astype(np.int64) * 1000 + hits["idx"]) * 8 + hits["trait"]
    assert np.all(np.diff(key) > 0), "sorted by (i, j, trait)"
    for tr in range(3):
        ref = oracle.pairs_f64(bits, ds.idxs, ds.Y[:, tr], ds.U, l10)
        assert len(ref["i"]) > 10
        _compare(hits[hits["trait"] == tr], ref, l10)
