"""The `kgwas` drop-in binary: CLI, formats, error behaviour, end-to-end output.

CPU tests exercise every input error the reference turns into "[ERROR] ..." +
exit 1 (src/kgwas_utils.hpp, src/kgwas.cpp:67-126); those are raised before any
GPU call.  GPU tests compare the written .tsv / .2d.tsv / .count with the text
the reference's emit rule gives for the FP64 oracle's values.
"""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from supergwas_amd import batchdir, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KGWAS = os.path.join(ROOT, "supergwas_amd", "bin", "kgwas")


@pytest.fixture(scope="module")
def small(tmp_path_factory, built):
    if not os.path.exists(KGWAS):
        from supergwas_amd import build
        build.build_kgwas_bin()
    ds = synth.make_dataset(seed=21, n_acc=150, n_missing=9, npop=3, p=4, m=3000, n_causal=5,
                            effect=1.0)
    bits = ds.fill_bits(0, 3000)
    root = str(tmp_path_factory.mktemp("kg"))
    paths = batchdir.write_inputs(root, ds, bits)
    return ds, bits, paths, root


def run(args, env=None, cwd=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([KGWAS, *args], capture_output=True, text=True, env=e, cwd=cwd,
                          timeout=300)


def base_args(p, out, inp=None):
    return [p["samples"], p["pheno"], inp or p["batch"], p["ubin"], out]


def test_usage_exits_1(small):
    r = run([])
    assert r.returncode == 1 and "Usage" in r.stderr


@pytest.mark.parametrize("bad,msg", [("-1", "threshold must be >= 0"),
                                     ("abc", "invalid argument: abc")])
def test_bad_threshold(small, tmp_path, bad, msg):
    ds, bits, p, root = small
    r = run(base_args(p, str(tmp_path / "o.tsv")) + [bad])
    assert r.returncode == 1 and f"[ERROR] {msg}" in r.stderr


def test_bad_mode(small, tmp_path):
    ds, bits, p, root = small
    r = run(base_args(p, str(tmp_path / "o.tsv")) + ["--mode", "diag"])
    assert r.returncode == 1 and "--mode must be 'cols' or 'rows'" in r.stderr


def test_ubin_n_mismatch(small, tmp_path):
    ds, bits, p, root = small
    bad = tmp_path / "U.bin"
    with open(bad, "wb") as f:
        f.write(np.array([ds.n + 1, 2], np.int32).tobytes())
        f.write(np.zeros((ds.n + 1) * 2).tobytes())
    r = run([p["samples"], p["pheno"], p["batch"], str(bad), str(tmp_path / "o.tsv")])
    assert r.returncode == 1 and "U.bin n mismatch" in r.stderr


def test_short_ubin_body(small, tmp_path):
    ds, bits, p, root = small
    bad = tmp_path / "U.bin"
    with open(bad, "wb") as f:
        f.write(np.array([ds.n, 3], np.int32).tobytes())
        f.write(np.zeros(ds.n).tobytes())
    r = run([p["samples"], p["pheno"], p["batch"], str(bad), str(tmp_path / "o.tsv")])
    assert r.returncode == 1 and "short U.bin body" in r.stderr


def test_no_overlap(small, tmp_path):
    ds, bits, p, root = small
    ph = tmp_path / "ph.txt"
    ph.write_text("nobody 1.0\n")
    r = run([p["samples"], str(ph), p["batch"], p["ubin"], str(tmp_path / "o.tsv")])
    assert r.returncode == 1 and "no phenotype overlap" in r.stderr


def test_missing_and_short_columns(small, tmp_path):
    ds, bits, p, root = small
    b2 = tmp_path / "batch"
    shutil.copytree(p["batch"], b2)
    a = int(ds.idxs[3])
    col = b2 / "bits" / f"col_{a:06d}.txt"
    col.write_bytes(b"0101\n")
    r = run(base_args(p, str(tmp_path / "o.tsv"), str(b2)))
    assert r.returncode == 1 and "short column" in r.stderr
    col.unlink()
    r = run(base_args(p, str(tmp_path / "o.tsv"), str(b2)))
    assert r.returncode == 1 and "missing column" in r.stderr


def test_gz_column_too_long(small, tmp_path):
    ds, bits, p, root = small
    b2 = tmp_path / "batch"
    shutil.copytree(p["batch"], b2)
    a = int(ds.idxs[0])
    col = b2 / "bits" / f"col_{a:06d}.txt"
    data = col.read_bytes()
    col.unlink()
    with gzip.open(str(col) + ".gz", "wb") as f:
        f.write(data[:-1] + b"1\n")
    r = run(base_args(p, str(tmp_path / "o.tsv"), str(b2)))
    # the first phenotyped column defines N, so N grows by one and kmers.txt.gz is short
    assert r.returncode == 1 and "kmers lines != N" in r.stderr
    a2 = int(ds.idxs[1])
    col2 = b2 / "bits" / f"col_{a2:06d}.txt"
    d2 = col2.read_bytes()
    col2.unlink()
    with gzip.open(str(col2) + ".gz", "wb") as f:
        f.write(d2[:-1] + b"11\n")
    (b2 / "bits" / f"col_{a:06d}.txt.gz").unlink()
    col.write_bytes(data)
    r = run(base_args(p, str(tmp_path / "o.tsv"), str(b2)))
    assert r.returncode == 1 and "gz column longer than expected" in r.stderr


def test_kmer_count_mismatch(small, tmp_path):
    ds, bits, p, root = small
    b2 = tmp_path / "batch"
    shutil.copytree(p["batch"], b2)
    with gzip.open(b2 / "kmers.txt.gz", "wt") as f:
        f.write("A\nB\n")
    r = run(base_args(p, str(tmp_path / "o.tsv"), str(b2)))
    assert r.returncode == 1 and "kmers lines != N (got 2, expected 3000)" in r.stderr


def test_no_gpu_fails_cleanly(small, tmp_path):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    ds, bits, p, root = small
    r = run(base_args(p, str(tmp_path / "o.tsv")))
    assert r.returncode == 1 and "[ERROR] GPU init failed" in r.stderr


def _expected(ds, bits, names, corr, l10, l2d):
    ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, corr)
    t1, t2 = batchdir.expected_text(ref, names, ds.idxs, bits, l10, l2d)
    return ref, t1, t2


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["cols", "cols_gz", "rows", "rows_gz"])
def test_end_to_end_outputs(small, tmp_path, mode):
    ds, bits, p, root = small
    if mode == "cols_gz":
        p = batchdir.write_inputs(str(tmp_path / "in"), ds, bits, gz_cols=True)
    if mode == "rows_gz":
        p = batchdir.write_inputs(str(tmp_path / "in"), ds, bits, rows_gz=True)
    inp = p["rows"] if mode.startswith("rows") else p["batch"]
    out = str(tmp_path / "res.batch_000.tsv")
    l10, l2d = 2.0, 3.0
    r = run(base_args(p, out, inp) + [str(l10), "--log10-thresh-2d", str(l2d)],
            env={"CORR_THRESHOLD": "0.15", "KG_BATCH_KMERS": "1000"})
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    ref, t1, t2 = _expected(ds, bits, p["names"], 0.15, l10, l2d)
    assert open(out).read() == t1
    assert open(str(tmp_path / "res.batch_000.2d.tsv")).read() == t2
    assert open(out + ".count").read() == f"res.batch_000\t{ref['n_passed']}\n"
    if mode.startswith("cols"):  # the reference's stage lines (src/kgwas.cpp:164,199,217)
        for line in ("passed correlation filter: ", "[INFO] Correlation filter (",
                     "[INFO] LA (", "[INFO] Count written to "):
            assert line in r.stderr, line


@pytest.mark.gpu
def test_2d_disabled_and_env_override(small, tmp_path):
    ds, bits, p, root = small
    out = str(tmp_path / "o.tsv")
    r = run(base_args(p, out) + ["1"], env={"LOG10_THRESH_2D_INPUT": "0"})
    assert r.returncode == 0, r.stderr
    assert not os.path.exists(str(tmp_path / "o.2d.tsv"))
    ref, t1, _ = _expected(ds, bits, p["names"], 0.2, 1.0, 0.0)
    assert open(out).read() == t1


def _overlap_n(p, tmp_path, samples_text, pheno_text):
    """Overlap size the binary computes, read from its U.bin mismatch message."""
    s = tmp_path / "s.txt"
    s.write_bytes(samples_text.encode())
    ph = tmp_path / "ph.txt"
    ph.write_text(pheno_text)
    ub = tmp_path / "U0.bin"
    ub.write_bytes(np.array([99999, 0], np.int32).tobytes())
    r = run([str(s), str(ph), p["batch"], str(ub), str(tmp_path / "o.tsv")])
    assert r.returncode == 1, r.stderr
    msg = [l for l in r.stderr.splitlines() if "U.bin n mismatch" in l][0]
    return int(msg.split("overlap n=")[1])


def test_parse_quirks(small, tmp_path):
    """load_lines / load_pheno quirks (src/kgwas_utils.hpp:51-70)."""
    ds, bits, p, root = small
    # CRLF and empty lines in ACCESSIONS are tolerated
    assert _overlap_n(p, tmp_path, "a\r\n\r\nb\r\nc\n", "a 1\nb 2\nc 3\n") == 3
    # reading stops at the first value that does not parse (NA)
    assert _overlap_n(p, tmp_path, "a\nb\nc\n", "a 1\nb NA\nc 3\n") == 1
    # duplicate ids: one entry (last value wins)
    assert _overlap_n(p, tmp_path, "a\nb\n", "a 1\na 2\nb 3\n") == 2
    # accession order defines the overlap, not phenotype order
    assert _overlap_n(p, tmp_path, "x\nb\na\n", "a 1\nb 2\n") == 2


@pytest.mark.parametrize("gz", [False, True])
def test_packed_rows_bit_exact(small, tmp_path, gz):
    """The SSE2 tile packer (16 accessions x 16 k-mers per transpose) against the
    expected rows: the phenotyped accessions' bits, zero elsewhere, including the
    k-mers past the last full tile of 16 (3000 = 187 x 16 + 8)."""
    ds, bits, p, root = small
    batch = p["batch"]
    if gz:
        batch = str(tmp_path / "batch")
        shutil.copytree(p["batch"], batch)
        for a in range(0, ds.n_acc, 3):
            col = os.path.join(batch, "bits", f"col_{a:06d}.txt")
            data = open(col, "rb").read()
            os.remove(col)
            with gzip.open(col + ".gz", "wb") as f:
                f.write(data)
    out = tmp_path / "rows.bin"
    r = run(base_args(p, str(tmp_path / "o.tsv"), batch),
            env={"KG_PACK_ONLY": str(out), "LOAD_THREADS": "3"})
    assert r.returncode == 0, r.stderr
    rows = np.fromfile(out, np.uint64).reshape(len(bits), -1)
    mask = np.zeros(bits.shape[1], np.uint64)
    for a in ds.idxs:
        mask[a >> 6] |= np.uint64(1) << np.uint64(a & 63)
    assert np.array_equal(rows, bits & mask)


@pytest.mark.gpu
def test_dedup_outputs_identical(small, tmp_path):
    """KG_DEDUP=1 on a batch with runs of equal rows (as sorted batches have): the same
    .tsv / .2d.tsv / .count text as the per-k-mer oracle."""
    ds, bits, p, root = small
    rng = np.random.default_rng(3)
    lens = rng.geometric(0.4, 1200)
    rows = np.ascontiguousarray(np.repeat(bits[:1200], lens, axis=0))
    q = batchdir.write_inputs(str(tmp_path / "in"), ds, rows, rows_file=False)
    out = str(tmp_path / "res.batch_000.tsv")
    l10, l2d = 2.0, 3.0
    r = run(base_args(q, out) + [str(l10), "--log10-thresh-2d", str(l2d)],
            env={"CORR_THRESHOLD": "0.15", "KG_DEDUP": "1"})
    assert r.returncode == 0, r.stderr
    ref, t1, t2 = _expected(ds, rows, q["names"], 0.15, l10, l2d)
    print(f"[cli] dedup rows={len(rows)} hits={t1.count(chr(10))} passed={ref['n_passed']}",
          flush=True)
    assert open(out).read() == t1
    assert open(str(tmp_path / "res.batch_000.2d.tsv")).read() == t2
    assert open(out + ".count").read() == f"res.batch_000\t{ref['n_passed']}\n"
    r = run(base_args(q, out), env={"KG_DEDUP": "yes"})
    assert r.returncode == 1 and "KG_DEDUP must be 0 or 1" in r.stderr


def _write_ubin(path, U):
    with open(path, "wb") as f:
        f.write(np.array(U.shape, np.int32).tobytes())
        f.write(np.asfortranarray(U, np.float64).tobytes(order="F"))


@pytest.mark.gpu
def test_too_many_structure_columns(small, tmp_path):
    """p + t above the int8 kernel's 84 base columns (at 3 limbs) ends with [ERROR] + exit 1."""
    ds, bits, p, root = small
    rng = np.random.default_rng(5)
    U, _, _ = np.linalg.svd(rng.standard_normal((ds.n, 85)), full_matrices=False)
    ub = tmp_path / "U85.bin"
    _write_ubin(ub, U)
    r = run([p["samples"], p["pheno"], p["batch"], str(ub), str(tmp_path / "o.tsv")])
    print(r.stderr[-600:])
    assert r.returncode == 1
    assert "[ERROR] model set-up failed: p + t exceeds the int8 kernel's weight columns" in r.stderr


@pytest.mark.gpu
def test_too_many_samples(tmp_path, built):
    """More than 60000 phenotyped samples (the int8 kernel's exact-sum limit) ends with
    [ERROR] + exit 1; the inputs are 60001 accessions with 2 k-mers."""
    n = 60001
    ids = [f"s{j}" for j in range(n)]
    (tmp_path / "acc.txt").write_text("\n".join(ids) + "\n")
    rng = np.random.default_rng(6)
    y = rng.standard_normal(n)
    (tmp_path / "ph.txt").write_text("".join(f"{a} {v!r}\n" for a, v in zip(ids, y.tolist())))
    _write_ubin(tmp_path / "U.bin", np.full((n, 1), 1.0 / np.sqrt(n)))
    bd = tmp_path / "batch_000" / "bits"
    bd.mkdir(parents=True)
    for j in range(n):
        (bd / f"col_{j:06d}.txt").write_bytes(b"01\n" if j % 2 else b"10\n")
    with gzip.open(tmp_path / "batch_000" / "kmers.txt.gz", "wt") as f:
        f.write("K0\nK1\n")
    r = run([str(tmp_path / "acc.txt"), str(tmp_path / "ph.txt"), str(tmp_path / "batch_000"),
             str(tmp_path / "U.bin"), str(tmp_path / "o.tsv")])
    print(r.stderr[-600:])
    assert r.returncode == 1
    assert "[ERROR] model set-up failed: more than 60000 phenotyped samples" in r.stderr
