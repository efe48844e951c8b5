"""Seeded synthetic k-mer GWAS inputs (accessions, populations, bits, U, phenotype).

The bit matrix is defined by csrc/synth_core.h (integer-only hash), so the host
fill (libkgsynth.so) and the device fill (kg_synth_bits in libkgwas_hip.so)
produce identical rows.  Population frequency profiles, the structure matrix U
and the phenotype are drawn here with numpy's PCG64 from the seed.

Model choices (the workload descriptions leave these open):
  * per-population allele frequency ~ Beta(f_anc, 1 - f_anc) around an
    ancestral frequency f_anc ~ U(0.05, 0.95) (Balding-Nichols with F = 0.5,
    which is exactly Beta(0.5, 0.5) at f_anc = 0.5);
  * MAF >= 0.05 is applied to each profile's expected frequency over the
    accession population mix (profiles failing it are redrawn);
  * p is the column count of U (U.bin's r, the width the scan contracts
    with).  PCs = top left singular vectors of the column-standardised matrix
    of `n_struct` extra k-mers; U = orthonormal basis of [1, PC_1..PC_{p-1}]
    (intercept direction included), the shape the reference's U.bin producer
    writes (src/compute_structure_all.cpp:176-186).  With intercept=False U is
    the bare top-p PCs (orthogonal to the all-ones vector);
  * phenotype = N(0, 1) + sum of causal effects x + 0.3 * PC1 (PC1 = U[:, 0]
    z-scored); missing-phenotype accessions are a seeded random subset.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "lib", "libkgsynth.so")
_lib = None

STRUCT_KMER_BASE = 1 << 40   # k-mer ids used for U estimation, disjoint from tested ids
CAUSAL_KMER_BASE = 0         # causal k-mers are tested k-mers (ids spread over [0, m))


def _synth_lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"{_LIB} not built; run __graft_entry__.build()")
        L = ctypes.CDLL(_LIB)
        L.ks_fill_bits_host.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.c_uint32, ctypes.c_int64, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_void_p]
        L.ks_fill_bits_host.restype = None
        _lib = L
    return _lib


def row_words_for(n_acc, align_words=8):
    """uint64 words per packed k-mer row: ceil(n_acc/64) rounded up to `align_words`."""
    w = (n_acc + 63) // 64
    return ((w + align_words - 1) // align_words) * align_words


@dataclass
class Dataset:
    seed: int
    n_acc: int
    npop: int
    pop: np.ndarray          # uint8 [n_acc]
    thr: np.ndarray          # uint32 [n_profiles, npop]
    idxs: np.ndarray         # int32 [n] phenotyped accessions, accession order
    U: np.ndarray            # float64 [n, p]
    Y: np.ndarray            # float64 [n, t]
    causal: list = field(default_factory=list)   # per trait: (ids, effects)
    m: int = 0

    @property
    def n(self):
        return len(self.idxs)

    @property
    def p(self):
        return self.U.shape[1]

    @property
    def y(self):
        return self.Y[:, 0]

    @property
    def n_profiles(self):
        return self.thr.shape[0]

    def fill_bits(self, kmer0, m, row_words=None):
        """Packed rows for k-mer ids kmer0 .. kmer0+m-1 (uint64 [m, row_words])."""
        rw = row_words or row_words_for(self.n_acc)
        out = np.empty((m, rw), np.uint64)
        pop = np.ascontiguousarray(self.pop, np.uint8)
        thr = np.ascontiguousarray(self.thr, np.uint32)
        _synth_lib().ks_fill_bits_host(self.seed, kmer0, m, self.n_acc, rw,
                                       pop.ctypes.data, thr.ctypes.data, self.npop,
                                       self.n_profiles, out.ctypes.data)
        return out


def _profiles(rng, npop, weights, n_profiles):
    thr = np.empty((n_profiles, npop), np.uint32)
    k = 0
    while k < n_profiles:
        f_anc = rng.uniform(0.05, 0.95)
        fr = rng.beta(max(f_anc, 1e-3), max(1.0 - f_anc, 1e-3), npop)
        fbar = float(weights @ fr)
        if min(fbar, 1.0 - fbar) < 0.05:
            continue
        thr[k] = np.minimum(np.floor(fr * 65536.0), 65536).astype(np.uint32)
        k += 1
    return thr


def make_dataset(seed=0, n_acc=200, n_missing=10, npop=3, p=5, m=100_000, n_causal=5,
                 effect=0.5, effect_sd=None, pc_effect=0.3, traits=1, trait_corr=0.0,
                 n_struct=4096, n_profiles=4096, intercept=True):
    """Build a seeded dataset; effect_sd switches causal effects to N(0, effect_sd^2)."""
    rng = np.random.default_rng(seed)
    pop = rng.integers(0, npop, n_acc).astype(np.uint8)
    weights = np.bincount(pop, minlength=npop) / n_acc
    thr = _profiles(rng, npop, weights, n_profiles)
    missing = rng.choice(n_acc, size=n_missing, replace=False) if n_missing else []
    mask = np.ones(n_acc, bool)
    mask[missing] = False
    idxs = np.nonzero(mask)[0].astype(np.int32)
    ds = Dataset(seed=seed, n_acc=n_acc, npop=npop, pop=pop, thr=thr, idxs=idxs,
                 U=np.zeros((len(idxs), 0)), Y=np.zeros((len(idxs), traits)), m=m)
    # structure: top-p left singular vectors of the standardised sample matrix
    S = ds.fill_bits(STRUCT_KMER_BASE, n_struct)
    Z = _unpack(S, idxs)                       # n x n_struct
    Z -= Z.mean(axis=0)
    sd = Z.std(axis=0, ddof=1)
    sd[sd == 0] = 1.0
    Z /= sd
    Uf, _, _ = np.linalg.svd(Z, full_matrices=False)
    n = len(idxs)
    if intercept:
        pcs = Uf[:, :max(p - 1, 1)]
        U, _, _ = np.linalg.svd(np.column_stack([np.ones(n), pcs[:, :p - 1]]),
                                full_matrices=False)
    else:
        pcs = Uf[:, :p]
        U = pcs
    ds.U = np.ascontiguousarray(U)
    pc1 = pcs[:, 0] - pcs[:, 0].mean()
    pc1 /= pc1.std(ddof=1)
    # correlated noise across traits
    C = np.full((traits, traits), trait_corr) + (1 - trait_corr) * np.eye(traits)
    noise = rng.standard_normal((n, traits)) @ np.linalg.cholesky(C).T
    Y = noise + pc_effect * pc1[:, None]
    ds.causal = []
    for t in range(traits):
        ids = np.sort(rng.choice(m, size=min(n_causal, m), replace=False)).astype(np.int64)
        eff = (rng.normal(0.0, effect_sd, len(ids)) if effect_sd is not None
               else np.full(len(ids), effect))
        for cid, e in zip(ids, eff):
            Y[:, t] += e * _unpack(ds.fill_bits(int(cid), 1), idxs)[:, 0]
        ds.causal.append((ids, eff))
    ds.Y = Y
    return ds


def _unpack(bits, idxs):
    idxs = np.asarray(idxs, np.int64)
    w = bits[:, idxs >> 6]
    return ((w >> (idxs & 63).astype(np.uint64)) & 1).T.astype(np.float64)


# Named workloads (sizes as the benchmark descriptions give them).
CONFIGS = {
    # plain method, small: n_acc=200 (10 without phenotype), p=5, m=1e5, 3 populations
    "small": dict(n_acc=200, n_missing=10, npop=3, p=5, m=100_000, n_causal=5, effect=0.5),
    # plain FP32 at scale: n=1024, p=20, 50 batches of 1e6, 8 populations, 50 causal N(0,.3^2)
    "fp32_scale": dict(n_acc=1024, n_missing=0, npop=8, p=20, m=50_000_000, n_causal=50,
                       effect_sd=0.3),
    # mixed precision headline: n=2048, p=32, m=1e8
    "mixed": dict(n_acc=2048, n_missing=0, npop=8, p=32, m=100_000_000, n_causal=50,
                  effect_sd=0.3),
    # multi-trait: n=4096, p=64, 8 traits (corr 0.2), 100 causal each
    "multitrait": dict(n_acc=4096, n_missing=0, npop=8, p=64, m=1_000_000_000, n_causal=100,
                       effect_sd=0.3, traits=8, trait_corr=0.2),
}
