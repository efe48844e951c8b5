"""Writes tests/golden/golden_small.npz: golden vectors for the 1D statistic.

The reference ships no tests or fixtures, so the expected values here come
from an independent derivation, not from the oracle's formula:
  * T = n ln(RSS0 / RSS1), with RSS0 / RSS1 the residual sums of squares of
    least-squares fits (numpy lstsq) of y on [U] and on [U, x]; for an
    orthonormal U this is the reference's n ln(1 / (1 - r^2))
    (src/kgwas.cpp:44-56);
  * effect = beta_x / RSS0, beta_x the lstsq coefficient of x in the [U, x]
    fit (Frisch-Waugh: beta_x = xyM / xxM, so xyM / (xxM yyM) = beta_x / yyM);
  * p = scipy.stats.chi2.sf(T, 1);
  * the pre-filter correlation is numpy.corrcoef(x, y);
  * c = number of phenotyped samples carrying the k-mer.
Case A enumerates all 2^12 bit patterns over 12 phenotyped samples (brute force).
Run: python tests/golden/make_golden.py
"""
import os

import numpy as np
from scipy import stats

HERE = os.path.dirname(os.path.abspath(__file__))


def orthonormal(rng, n, p, intercept):
    A = rng.standard_normal((n, p))
    if intercept:
        A[:, 0] = 1.0
    Q, _ = np.linalg.qr(A)
    return Q


def pack(xacc):
    """xacc: m x n_acc 0/1 -> packed uint64 rows (row_words = multiple of 8)."""
    m, n_acc = xacc.shape
    rw = ((n_acc + 63) // 64 + 7) // 8 * 8
    out = np.zeros((m, rw), np.uint64)
    for a in range(n_acc):
        out[:, a >> 6] |= (xacc[:, a].astype(np.uint64) << np.uint64(a & 63))
    return out


def expected(X, y, U):
    """X: m x n (phenotyped samples), y: n, U: n x p orthonormal."""
    m, n = X.shape
    res0 = y - U @ np.linalg.lstsq(U, y, rcond=None)[0]
    rss0 = res0 @ res0
    T = np.zeros(m)
    eff = np.zeros(m)
    for i in range(m):
        A = np.column_stack([U, X[i]])
        coef, _, rank, _ = np.linalg.lstsq(A, y, rcond=None)
        if rank <= U.shape[1]:       # x in span(U): xxM = 0 -> (0, 0, 1)
            continue
        r1 = y - A @ coef
        rss1 = r1 @ r1
        T[i] = n * np.log(rss0 / rss1)
        eff[i] = coef[-1] / rss0
    p = stats.chi2.sf(T, 1)
    p[T == 0] = 1.0
    with np.errstate(invalid="ignore", divide="ignore"):
        cor = np.array([np.corrcoef(x, y)[0, 1] if 0 < x.sum() < n else np.nan for x in X])
    return T, p, eff, cor, X.sum(axis=1).astype(np.int32)


def case(rng, n_acc, missing, p, intercept, X=None, m=0):
    mask = np.ones(n_acc, bool)
    mask[list(missing)] = False
    idxs = np.nonzero(mask)[0].astype(np.int32)
    n = len(idxs)
    U = orthonormal(rng, n, p, intercept)
    y = rng.standard_normal(n) + 0.7
    if X is None:
        X = (rng.random((m, n)) < rng.uniform(0.1, 0.9, (m, 1))).astype(np.float64)
    xacc = (rng.random((len(X), n_acc)) < 0.5).astype(np.float64)  # noise at unphenotyped
    xacc[:, idxs] = X
    T, pv, eff, cor, c = expected(X, y, U)
    return dict(bits=pack(xacc), idxs=idxs, y=y, U=U, T=T, p=pv, effect=eff, cor=cor, c=c)


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    allx = ((np.arange(1 << 12)[:, None] >> np.arange(12)[None, :]) & 1).astype(np.float64)
    cases = {
        "A": case(rng, 14, [3, 9], 3, True, X=allx),
        "B": case(rng, 16, [], 2, False, m=500),
        "C": case(rng, 70, [0, 65, 69, 33, 12, 40], 3, True, m=300),
    }
    for name, d in cases.items():
        for k, v in d.items():
            out[f"{name}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
