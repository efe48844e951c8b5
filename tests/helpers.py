"""Shared test helpers: expected reports from the FP64 oracle and comparisons."""
import numpy as np

import oracle


def expected_hits(ref, log10_thresh, log10_thresh_2d):
    """Reported k-mers per the reference's emit rule (src/kgwas.cpp:249-254) on oracle values."""
    p_thr = 10.0 ** (-log10_thresh)
    rep = ref["passed"] & (ref["p"] <= p_thr)
    idx = np.nonzero(rep)[0]
    flag2d = (ref["p"][idx] <= 10.0 ** (-log10_thresh_2d)) if log10_thresh_2d > 0 else \
        np.zeros(len(idx), bool)
    return idx, flag2d


def near_threshold(ref, corr_thr, log10_thresh, band=1e-6):
    """K-mers whose oracle pre-filter |cor| or p lies within a relative band of a threshold."""
    n = None
    p_thr = 10.0 ** (-log10_thresh)
    with np.errstate(divide="ignore", invalid="ignore"):
        near_p = np.abs(ref["p"] / p_thr - 1.0) <= band
    near_c = np.zeros(len(ref["p"]), bool)
    if "cor" in ref:
        near_c = np.abs(np.abs(ref["cor"]) / corr_thr - 1.0) <= band
    return near_p | near_c


def pearson(bits, idxs, y):
    X = oracle.unpack_rows(bits, idxs).astype(np.float64)  # n x m
    yc = y - y.mean()
    Xc = X - X.mean(axis=0)
    num = Xc.T @ yc
    den = np.sqrt((Xc * Xc).sum(axis=0)) * np.sqrt(yc @ yc)
    with np.errstate(divide="ignore", invalid="ignore"):
        return num / den


def rel_err_T(T, T64):
    return np.abs(T - T64) / np.maximum(np.abs(T64), 1.0)


def T_from(xxM, xyM, yyM, n):
    xxM = np.asarray(xxM, np.float64)
    xyM = np.asarray(xyM, np.float64)
    ok = (xxM > 1e-12) & (yyM > 1e-12)
    with np.errstate(divide="ignore", invalid="ignore"):
        r2 = np.where(ok, xyM * xyM / (xxM * yyM), 0.0)
    r2 = np.clip(r2, 0.0, None)
    r2 = np.where(r2 >= 1, 1 - 1e-15, r2)
    return np.where(ok, n * np.log(1.0 / (1.0 - r2)), 0.0)
