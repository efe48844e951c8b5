"""Duplicate-row scans (kg_set_dedup): equal adjacent rows are scanned once.

The batches of src/convert_matrix.cpp:366-368 are sorted by bit pattern, so
k-mers with equal rows are adjacent.  With the duplicate check on, the screen
compares each row with the previous one under the accession mask, the full
kernel runs on the first row of each run, pre-filter passes count the whole
run and every member of a reported run is reported.  The bar: hits (every
field) and pass counts identical to the scan without the check, on rows with
planted runs (across wave tiles and blocks, long runs, rows that differ only
at accessions without a phenotype).
"""
import numpy as np
import pytest

from supergwas_amd import synth

pytestmark = pytest.mark.gpu


def planted_rows(ds, m_base, seed):
    """Rows built from m_base distinct rows repeated in runs (geometric lengths, and runs of
    300, 129 and 64 rows)."""
    rng = np.random.default_rng(seed)
    base = ds.fill_bits(0, m_base)
    lens = rng.geometric(0.35, m_base)
    lens[rng.choice(m_base, 3, replace=False)] = [300, 129, 64]
    rows = np.repeat(base, lens, axis=0)
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
    # inside some runs, flip a bit at an accession without a phenotype: still equal
    # under the mask; flip a phenotyped bit in others: a new run starts there
    unph = np.setdiff1d(np.arange(ds.n_acc), ds.idxs)
    for s, L in zip(starts[::7], lens[::7]):
        if L > 2 and len(unph):
            a = int(rng.choice(unph))
            rows[s + 1, a >> 6] ^= np.uint64(1) << np.uint64(a & 63)
    for s, L in zip(starts[3::11], lens[3::11]):
        if L > 3:
            a = int(ds.idxs[rng.integers(len(ds.idxs))])
            rows[s + 2, a >> 6] ^= np.uint64(1) << np.uint64(a & 63)
    return np.ascontiguousarray(rows)


def _scan(ds, rows, prm, dedup, Y=None):
    from supergwas_amd.api import KgwasHip
    with KgwasHip(screen="on", dedup=dedup) as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y if Y is None else Y)
        hits, npass = k.scan(rows, prm)
        tim = k.timing()
    return hits, npass, tim


def _same(a, b):
    assert len(a) == len(b), (len(a), len(b))
    for key in a.dtype.names:
        assert np.array_equal(a[key], b[key]), key


CASES = {
    # name: (dataset kwargs, distinct rows, corr, log10, traits)
    "t1_n2048": (dict(seed=50, n_acc=2048, n_missing=40, npop=6, p=12, m=4000, n_causal=8,
                      effect=0.6), 4000, 0.1, 2.0, 1),
    "t1_n190": (dict(seed=51, n_acc=200, n_missing=10, npop=3, p=5, m=3000, n_causal=5,
                     effect=0.5), 3000, 0.2, 2.0, 1),
    "t3_n600": (dict(seed=52, n_acc=600, n_missing=25, npop=4, p=8, m=2500, n_causal=5,
                     effect=0.6, traits=3, trait_corr=0.3), 2500, 0.15, 2.0, 3),
    "t8_n333": (dict(seed=53, n_acc=333, n_missing=7, npop=4, p=6, m=2000, n_causal=5,
                     effect=0.6, traits=8, trait_corr=0.2), 2000, 0.15, 2.0, 8),
}


@pytest.mark.parametrize("name", list(CASES))
def test_dedup_identical_to_plain_scan(built, name):
    from supergwas_amd.api import ScanParams
    kw, mb, corr, l10, t = CASES[name]
    ds = synth.make_dataset(**kw)
    rows = planted_rows(ds, mb, seed=kw["seed"])
    prm = ScanParams(corr, l10, l10 + 1.0)
    Y = ds.Y if t > 1 else None
    print(f"[dedup] {name} rows={len(rows)} distinct<={mb} t={t}", flush=True)
    h0, p0, t0 = _scan(ds, rows, prm, False, Y)
    h1, p1, t1 = _scan(ds, rows, prm, True, Y)
    print(f"[dedup] {name} hits={len(h0)} passed={p0.tolist()} survivors {t0['survivors']} -> "
          f"{t1['survivors']}", flush=True)
    assert np.array_equal(p0, p1)
    _same(h0, h1)
    assert len(h0) > 0 and int(p0.sum()) > 0
    assert t1["survivors"] < t0["survivors"] / 2


def test_dedup_shards(built):
    """A shard table: runs stop at shard ends (each shard's first row is its own)."""
    from supergwas_amd.api import KgwasHip, ScanParams
    kw, mb, corr, l10, t = CASES["t1_n190"]
    ds = synth.make_dataset(**kw)
    rows = planted_rows(ds, mb, seed=7)
    rw = rows.shape[1]
    cuts = [0, 1000, 1001, 2500, 5000, len(rows)]
    prm = ScanParams(corr, l10, 0.0)
    out = []
    for dd in (False, True):
        with KgwasHip(screen="on", dedup=dd) as k:
            k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
            dev = k.to_device(rows)
            shards = [(dev.ptr + a * rw * 8, b - a, rw, 0) for a, b in zip(cuts[:-1], cuts[1:])]
            out.append(k.scan_shards(shards, prm))
            dev.free()
    (h0, p0), (h1, p1) = out
    print(f"[dedup] shards hits={len(h0)} passed={p0.ravel().tolist()}", flush=True)
    assert np.array_equal(p0, p1)
    _same(h0, h1)


def test_dedup_set_and_options(built):
    from supergwas_amd.api import KgError, KgwasHip, ScanParams
    kw, mb, corr, l10, t = CASES["t1_n190"]
    ds = synth.make_dataset(**kw)
    rows = planted_rows(ds, 400, seed=9)
    with KgwasHip(screen="on") as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        h0, p0 = k.scan(rows, ScanParams(corr, l10, 0.0))
        s0 = k.timing()["survivors"]
        k.set_dedup(True)
        h1, p1 = k.scan(rows, ScanParams(corr, l10, 0.0))
        s1 = k.timing()["survivors"]
        k.set_dedup(False)
        h2, p2 = k.scan(rows, ScanParams(corr, l10, 0.0))
        with pytest.raises(KgError):
            from supergwas_amd.api import lib
            k._check(lib().kg_set_dedup(k._h, 2))
    assert s1 < s0
    assert np.array_equal(p0, p1) and np.array_equal(p0, p2)
    _same(h0, h1)
    _same(h0, h2)


def test_dedup_streamed_chunks(built):
    """Host rows streamed in 64 MiB chunks (262,144 rows at 2048 accessions): runs that
    straddle a chunk boundary are split there; results equal the plain scan's."""
    from supergwas_amd.api import ScanParams
    ds = synth.make_dataset(seed=54, n_acc=2048, n_missing=30, npop=6, p=12, m=120000,
                            n_causal=8, effect=0.6)
    rows = planted_rows(ds, 120000, seed=54)
    assert len(rows) > 262144 + 1000
    prm = ScanParams(0.1, 3.0, 4.0)
    print(f"[dedup] streamed rows={len(rows)}", flush=True)
    h0, p0, t0 = _scan(ds, rows, prm, False)
    h1, p1, t1 = _scan(ds, rows, prm, True)
    print(f"[dedup] streamed hits={len(h0)} passed={int(p0[0])} survivors {t0['survivors']} -> "
          f"{t1['survivors']}", flush=True)
    assert np.array_equal(p0, p1)
    _same(h0, h1)
