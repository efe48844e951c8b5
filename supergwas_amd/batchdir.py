"""Write kgwas inputs on disk from a synthetic dataset.

Produces the files the `kgwas` binary reads, in the reference's formats:
ACCESSIONS (one id per line), PHENOTYPE ("id value" lines), U.bin
(int32 n, int32 r, float64 column-major; src/compute_structure_all.cpp:80-88),
a cols-mode batch directory (bits/col_%06d.txt[.gz] holding one '0'/'1' byte
per k-mer and a newline, kmers.txt.gz; src/convert_matrix.cpp:209-258) and a
rows-mode table ("kmer\\tbitstring" lines over all accessions).
"""
import gzip
import os

import numpy as np


def kmer_names(m, k0=0):
    return [f"K{i:09d}" for i in range(k0, k0 + m)]


def unpack_acc(bits, n_acc):
    a = np.arange(n_acc)
    return ((bits[:, a >> 6] >> (a & 63).astype(np.uint64)) & 1).astype(np.uint8)  # m x n_acc


def write_inputs(root, ds, bits, gz_cols=False, rows_gz=False, missing_pheno_token=None,
                 rows_file=True):
    """Writes everything under `root` (the rows-mode file only if rows_file); returns a
    dict of paths."""
    os.makedirs(root, exist_ok=True)
    ids = [f"acc{j:05d}" for j in range(ds.n_acc)]
    acc = os.path.join(root, "samples.txt")
    with open(acc, "w") as f:
        f.write("\n".join(ids) + "\n")
    ph = os.path.join(root, "phenotype.txt")
    with open(ph, "w") as f:
        for j, a in enumerate(ds.idxs):
            f.write(f"{ids[a]} {float(ds.y[j])!r}\n")
        if missing_pheno_token:
            f.write(missing_pheno_token)
    ub = os.path.join(root, "U.bin")
    with open(ub, "wb") as f:
        f.write(np.array([ds.n, ds.p], np.int32).tobytes())
        f.write(np.asfortranarray(ds.U, np.float64).tobytes(order="F"))
    m = bits.shape[0]
    X = unpack_acc(bits, ds.n_acc)  # m x n_acc
    names = kmer_names(m)
    batch = os.path.join(root, "batch_000")
    os.makedirs(os.path.join(batch, "bits"), exist_ok=True)
    for a in range(ds.n_acc):
        data = (X[:, a] + ord("0")).astype(np.uint8).tobytes() + b"\n"
        p = os.path.join(batch, "bits", f"col_{a:06d}.txt")
        if gz_cols:
            with gzip.open(p + ".gz", "wb") as f:
                f.write(data)
        else:
            with open(p, "wb") as f:
                f.write(data)
    with gzip.open(os.path.join(batch, "kmers.txt.gz"), "wt") as f:
        f.write("\n".join(names) + "\n")
    rows = os.path.join(root, "rows.tsv.gz" if rows_gz else "rows.tsv")
    if rows_file:
        opener = gzip.open if rows_gz else open
        with opener(rows, "wt") as f:
            chars = (X + ord("0")).astype(np.uint8)
            for i in range(m):
                f.write(names[i] + "\t" + chars[i].tobytes().decode() + "\n")
    return dict(samples=acc, pheno=ph, ubin=ub, batch=batch, rows=rows, names=names)


def expected_text(ref, names, idxs, bits, log10_thresh, log10_thresh_2d):
    """(1D text, 2D text) the reference's emit rule prints for FP64 values."""
    p_thr = 10.0 ** (-log10_thresh)
    rep = np.nonzero(ref["passed"] & (ref["p"] <= p_thr))[0]
    o1, o2 = [], []
    idxs = np.asarray(idxs, np.int64)
    for i in rep:
        e, p = ref["effect"][i], ref["p"][i]
        o1.append(f"{names[i]}\t{e:g}\t{p:g}\n")
        if log10_thresh_2d > 0 and p <= 10.0 ** (-log10_thresh_2d):
            b = ((bits[i, idxs >> 6] >> (idxs & 63).astype(np.uint64)) & 1).astype(np.uint8)
            o2.append(f"{names[i]}\t{(b + ord('0')).tobytes().decode()}\t{e:g}\t{p:g}\n")
    return "".join(o1), "".join(o2)
