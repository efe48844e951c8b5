"""The `kgwas_2d` drop-in binary: d4 index input, CLI, outputs, error behaviour.

CPU tests exercise the input errors that end a run with "[ERROR] ..." + exit 1
before any GPU call.  GPU tests compare the written pairs with the FP64
oracle's all-pairs scan (src/kgwas_2d.cpp:258-401) and the .kmers file with the
index's k-mers in encoded order (src/kgwas_2d.cpp:141-150).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from supergwas_amd import batchdir, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KGWAS_2D = os.path.join(ROOT, "supergwas_amd", "bin", "kgwas_2d")
K = 31
BASES = "ACGT"


def _encode(seq):
    e = 0
    for ch in seq:
        e = (e << 2) | BASES.index(ch)
    return e


def write_index(root, shards, k=K):
    """shards: list of lists of (kmer string, bitstring str); writes the d4 index
    (index.meta, shard_%03u.umap, shard_%03u.bits; src/map_utils.hpp:319-400)."""
    os.makedirs(root, exist_ok=True)
    with open(os.path.join(root, "index.meta"), "w") as f:
        f.write(f"K={k}\nM={len(shards)}\nN={sum(len(s) for s in shards)}\n")
    for sid, recs in enumerate(shards):
        with open(os.path.join(root, f"shard_{sid:03d}.umap"), "wb") as f:
            f.write(b"UMAPd4\0\0" + np.uint32(k).tobytes() + np.uint64(len(recs)).tobytes())
            for seq, _ in recs:
                f.write(int(_encode(seq)).to_bytes(16, "little"))
                f.write(np.float64(0.5).tobytes() + np.float64(1e-3).tobytes())
        with open(os.path.join(root, f"shard_{sid:03d}.bits"), "wb") as f:
            f.write(np.uint64(len(recs)).tobytes())
            for _, b in recs:
                f.write(np.uint32(len(b)).tobytes() + b.encode())


@pytest.fixture(scope="module")
def index(tmp_path_factory, built):
    if not os.path.exists(KGWAS_2D):
        from supergwas_amd import build
        build.build_kgwas_bin()
    ds = synth.make_dataset(seed=22, n_acc=150, n_missing=9, npop=3, p=4, m=90, n_causal=5,
                            effect=1.0)
    bits = ds.fill_bits(0, 90)
    root = str(tmp_path_factory.mktemp("kg2d"))
    paths = batchdir.write_inputs(root, ds, bits[:1])          # samples, phenotype, U.bin
    X = batchdir.unpack_acc(bits, ds.n_acc).astype(np.uint8)    # 90 x n_acc
    rng = np.random.default_rng(5)
    seqs = ["".join(rng.choice(list(BASES), K)) for _ in range(90)]
    recs = []
    for r in range(90):
        if r % 2 == 0:   # accession-length bitstring, read at the overlap positions
            b = "".join("1" if v else "0" for v in X[r])
        else:            # phenotyped-length bitstring, read directly
            b = "".join("1" if v else "0" for v in X[r, ds.idxs])
        recs.append((seqs[r], b))
    extra = [(seqs[0][:-1] + "A" if seqs[0][-1] != "A" else seqs[0][:-1] + "C", "0101"),  # dropped
             ("".join(rng.choice(list(BASES), K)), "")]                              # no bits
    # a key repeated inside shard 0: the last non-empty bitstring wins
    dup_seq = seqs[1]
    shard0 = [(dup_seq, "1" * ds.n)] + recs[:45] + extra
    shard1 = recs[45:]
    idx = os.path.join(root, "index")
    write_index(idx, [shard0, shard1])
    # expected k-mer order and rows (phenotyped order)
    keep = sorted(range(90), key=lambda r: _encode(seqs[r]))
    Xp = X[keep][:, ds.idxs]
    kmers = [seqs[r] for r in keep]
    return ds, Xp, kmers, paths, idx, root


def run(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([KGWAS_2D, *args], capture_output=True, text=True, env=e, timeout=300)


def _args(paths, idx, out):
    return [paths["samples"], paths["pheno"], paths["ubin"], idx, out]


def _expected(ds, Xp, l10, i0=0, i1=None):
    n = ds.n
    rw = ((n + 63) // 64 + 7) // 8 * 8
    bits = np.zeros((Xp.shape[0], rw), np.uint64)
    for j in range(n):
        bits[:, j >> 6] |= Xp[:, j].astype(np.uint64) << np.uint64(j & 63)
    return oracle.pairs_f64(bits, np.arange(n, dtype=np.int32), ds.y, ds.U, l10, i0, i1)


def _read_pairs(path):
    rows = [l.split("\t") for l in open(path).read().splitlines()]
    return rows


def _check(rows, ref):
    assert [(int(r[0]), int(r[1])) for r in rows] == list(zip(ref["i"].tolist(),
                                                              ref["j"].tolist()))
    for r, e, p in zip(rows, ref["effect"], ref["p"]):
        assert abs(float(r[2]) - e) <= 1e-7 * abs(e) + 1e-300
        assert abs(float(r[3]) - p) <= 1e-7 * p + 1e-300
        assert "e" in r[3] and len(r[3].split("e")[0].split(".")[1]) == 9   # %.9e


def test_usage_exits_1(index):
    r = run([])
    assert r.returncode == 1 and "Usage" in r.stderr


@pytest.mark.parametrize("case", ["not_a_dir", "no_meta", "bad_magic", "k_mismatch",
                                  "bits_count", "short_bits", "bad_threshold"])
def test_index_errors(index, tmp_path, case):
    ds, Xp, kmers, paths, idx, root = index
    bad = str(tmp_path / "idx")
    recs = [("A" * K, "1" * ds.n), ("C" * K, "0" * ds.n)]
    write_index(bad, [recs])
    um, bi = os.path.join(bad, "shard_000.umap"), os.path.join(bad, "shard_000.bits")
    args = _args(paths, bad, str(tmp_path / "o"))
    if case == "not_a_dir":
        args[3] = paths["samples"]
        msg = "Index path is not a directory"
    elif case == "no_meta":
        os.remove(os.path.join(bad, "index.meta"))
        msg = "cannot read index.meta"
    elif case == "bad_magic":
        data = bytearray(open(um, "rb").read())
        data[5] = ord("3")
        open(um, "wb").write(bytes(data))
        msg = "bad magic"
    elif case == "k_mismatch":
        with open(os.path.join(bad, "index.meta"), "w") as f:
            f.write("K=25\nM=1\nN=2\n")
        msg = "K mismatch"
    elif case == "bits_count":
        data = bytearray(open(bi, "rb").read())
        data[0] = 3
        open(bi, "wb").write(bytes(data))
        msg = ".bits N=3 != .umap N=2"
    elif case == "short_bits":
        data = open(bi, "rb").read()
        open(bi, "wb").write(data[:-5])
        msg = "short read (bitstring@1)"
    else:
        args.append("x1")
        msg = "invalid log10_thresh: x1"
    r = run(args)
    assert r.returncode == 1 and f"[ERROR] {msg}" in r.stderr, r.stderr


def test_no_valid_kmers(index, tmp_path):
    ds, Xp, kmers, paths, idx, root = index
    bad = str(tmp_path / "idx")
    write_index(bad, [[("A" * K, "01"), ("C" * K, "")]])
    r = run(_args(paths, bad, str(tmp_path / "o")))
    assert r.returncode == 1 and "[ERROR] No valid k-mers after bitstring parsing" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("l10", [2.0, 6.0])
def test_end_to_end_pairs(index, tmp_path, l10):
    ds, Xp, kmers, paths, idx, root = index
    out = str(tmp_path / "res")
    r = run(_args(paths, idx, out) + [f"{l10:g}"])
    assert r.returncode == 0, r.stderr
    ref = _expected(ds, Xp, l10)
    rows = _read_pairs(out + ".tsv")
    print(f"[2d] log10={l10} pairs={len(rows)} oracle={len(ref['i'])}", flush=True)
    _check(rows, ref)
    assert open(out + ".kmers").read().splitlines() == kmers
    if l10 == 2.0:
        assert len(rows) > 20


@pytest.mark.gpu
def test_task_split(index, tmp_path):
    ds, Xp, kmers, paths, idx, root = index
    out = str(tmp_path / "res")
    parts = []
    for t in range(3):
        r = run(_args(paths, idx, out) + ["2", "--task-id", str(t), "--num-tasks", "3",
                                          "--emit-threads", "4"])
        assert r.returncode == 0, r.stderr
        parts += _read_pairs(f"{out}.part_{t}.tsv")
        assert ("k-mers to" in r.stderr) == (t == 0)      # .kmers from task 0 only
    assert open(out + ".kmers").read().splitlines() == kmers
    ref = _expected(ds, Xp, 2.0)
    _check(parts, ref)
    # 60 tasks of ceil(90 / 60) = 2 rows: task 50 starts at row 100 >= N (src/kgwas_2d.cpp:287-290)
    r = run(_args(paths, idx, str(tmp_path / "late")) + ["2", "--task-id", "50", "--num-tasks", "60"])
    assert r.returncode == 0 and "no rows to process" in r.stderr
    assert not os.path.exists(str(tmp_path / "late") + ".part_50.tsv")
