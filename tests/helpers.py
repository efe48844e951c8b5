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


# ---------------------------------------------------------------------------
# Error bounds of the plain FP32 kernel (the reference's cols-mode arithmetic,
# src/kgwas.cpp:146-158 and the FP32 pre-filter of src/kgwas.cpp:183-200).
# Every FP32 sum is bounded as a sum of K terms with one rounding per addition
# (|error| <= gamma_K * sum |terms|, gamma_K = K u / (1 - K u), u = 2^-24),
# whatever order the kernel adds in; every FP32 input carries its own rounding
# (u relative).  The bounds are worst-case, so the true FP32 values lie inside
# them and a k-mer whose FP64 decision differs from the FP32 one must have its
# FP64 statistic inside the band they define.

U32 = 2.0 ** -24


def _gamma(k):
    k = np.asarray(k, np.float64)
    return k * U32 / (1.0 - k * U32)


def fp32_error_bounds(bits, idxs, y, U, n_acc):
    """Per-k-mer bounds (Bxx, Bxy, Bcor) on |FP32 value - exact value| of xxM, xyM and the
    pre-filter correlation as the FP32 kernel forms them."""
    X = oracle.unpack_rows(bits, idxs).astype(np.float64)   # n x m
    n = X.shape[0]
    p = U.shape[1]
    y = np.asarray(y, np.float64)
    c = X.sum(axis=0)
    gK = float(_gamma(n_acc))          # device sums run over all accession positions
    gn = float(_gamma(n + 1))
    gp = float(_gamma(p + 2))
    # U'x (device MFMA sums of FP32 U entries over the set bits)
    UX = X.T @ U                                             # m x p exact
    A = X.T @ np.abs(U)
    B = (U32 + gK * (1 + U32)) * A
    dsux = (2 * np.abs(UX) * B + B * B).sum(axis=1) + gp * ((np.abs(UX) + B) ** 2).sum(axis=1)
    xx = c - (UX * UX).sum(axis=1)
    Bxx = dsux + U32 * (np.abs(xx) + dsux)
    # U'y on the host in FP32 (sequential, FP32 inputs)
    Uy = U.T @ y
    dUy = (3 * U32 + gn) * (1 + 3 * U32) * (np.abs(U).T @ np.abs(y))
    # x'y on the device, then x'y - (U'x)'(U'y) in FP32
    dXy = (U32 + gK * (1 + U32)) * (X.T @ np.abs(y))
    duyux = ((np.abs(Uy) * B + np.abs(UX) * dUy + B * dUy).sum(axis=1) +
             gp * ((np.abs(UX) + B) * (np.abs(Uy) + dUy)).sum(axis=1))
    xy = X.T @ y - UX @ Uy
    Bxy = dXy + duyux + U32 * (np.abs(xy) + dXy + duyux)
    # pre-filter: FP32 mean of y, centred y, its norm, S = x'(y - mean) with the
    # reference's mean-of-x correction, and the closed-form |x - mean x|
    ybar = y.mean()
    sabs = np.abs(y).sum()
    dm = ((U32 + gn * (1 + U32)) * sabs) / n + 2 * U32 * (abs(ybar) + sabs / n)
    d = U32 * np.abs(y) + dm + U32 * (np.abs(y - ybar) + U32 * np.abs(y) + dm)
    yc = y - ybar
    ycn = np.sqrt(yc @ yc)
    ycabs = np.abs(yc) + d
    dS = X.T @ d + gK * (X.T @ ycabs)
    ysum = d.sum() + gn * ycabs.sum()
    S = X.T @ yc
    dS = dS + (c / n) * (1 + 4 * U32) * ysum + U32 * (np.abs(S) + dS + (c / n) * ysum)
    dny = np.sqrt(d @ d) + (gn / 2 + 2 * U32) * (ycn + np.sqrt(d @ d))
    xcn = np.sqrt(np.maximum(c * (n - c) / n, 0.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        cor = np.where(xcn > 0, S / (xcn * ycn), 0.0)
        Bcor = np.where(xcn > 0, dS / (xcn * ycn) + np.abs(cor) * (dny / ycn + 16 * U32), np.inf)
    return Bxx, Bxy, Bcor


def T_interval(xxM, xyM, yyM, n, Bxx, Bxy):
    """[T_lo, T_hi] of compute_hit's T over xxM +- Bxx, xyM +- Bxy (FP64 yyM)."""
    lo_xy = np.maximum(np.abs(xyM) - Bxy, 0.0)
    hi_xy = np.abs(xyM) + Bxy
    with np.errstate(divide="ignore", invalid="ignore"):
        r2_lo = np.where(xxM + Bxx > 1e-12, lo_xy ** 2 / ((xxM + Bxx) * yyM), 0.0)
        r2_hi = np.where(xxM - Bxx > 1e-12, hi_xy ** 2 / ((xxM - Bxx) * yyM), 1.0)
    r2_lo = np.clip(r2_lo, 0.0, 1 - 1e-15)
    r2_hi = np.clip(r2_hi, 0.0, 1 - 1e-15)
    T_lo = np.where(xxM - Bxx > 1e-12, n * -np.log1p(-r2_lo), 0.0)  # xxM may reach eps: T = 0
    return T_lo, n * -np.log1p(-r2_hi)
