"""FP64 oracle and FP32 CPU restatement of the kgwas cols-mode scan.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The product package never
imports this module.  See kgwas_oracle.c / cols_fp32.cpp for the reference
lines each function follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libkgoracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_compute_hit.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, _dp, _dp, _dp]
        L.or_compute_hit.restype = None
        L.or_chi2_sf_k1.argtypes = [ctypes.c_double]
        L.or_chi2_sf_k1.restype = ctypes.c_double
        L.or_yyM.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int]
        L.or_yyM.restype = ctypes.c_double
        L.or_T_threshold.argtypes = [ctypes.c_double]
        L.or_T_threshold.restype = ctypes.c_double
        L.or_scan_f64.argtypes = [_u64p, ctypes.c_int64, ctypes.c_int64, _i32p, ctypes.c_int,
                                  _dp, _dp, ctypes.c_int, ctypes.c_double,
                                  _i32p, _u8p, _dp, _dp, _dp, _dp, _dp]
        L.or_scan_f64.restype = ctypes.c_int64
        L.or_pairs_f64.argtypes = [_u64p, ctypes.c_int64, ctypes.c_int64, _i32p, ctypes.c_int,
                                   _dp, _dp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_double, ctypes.c_int64, _i64p, _i32p, _dp, _dp, _dp,
                                   _dp, _dp]
        L.or_pairs_f64.restype = ctypes.c_int64
        L.or_unpack_rows.argtypes = [_u64p, ctypes.c_int64, ctypes.c_int64, _i32p, ctypes.c_int,
                                     _fp]
        L.or_unpack_rows.restype = None
        L.cpu_cols_fp32.argtypes = [_fp, ctypes.c_int, ctypes.c_int64, _dp, _dp, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, _i64p, _i64p, _dp, _fp, _fp, _i64p, _i64p]
        L.cpu_cols_fp32.restype = ctypes.c_int64
        L.cpu_load_columns.argtypes = [ctypes.c_char_p, _i32p, ctypes.c_int, ctypes.c_int64, _fp,
                                       ctypes.c_int]
        L.cpu_load_columns.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def compute_hit(xxM, xyM, yyM, n):
    """(effect, T, pval) of src/kgwas.cpp:44-56."""
    e, t, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().or_compute_hit(float(xxM), float(xyM), float(yyM), int(n),
                         ctypes.byref(e), ctypes.byref(t), ctypes.byref(p))
    return e.value, t.value, p.value


def chi2_sf_k1(T):
    return lib().or_chi2_sf_k1(float(T))


def T_threshold(p_thresh):
    return lib().or_T_threshold(float(p_thresh))


def yyM(y, U):
    y = np.ascontiguousarray(y, dtype=np.float64)
    Uf = np.asfortranarray(U, dtype=np.float64)
    return lib().or_yyM(_ptr(y, _dp), len(y), _ptr(Uf, _dp), Uf.shape[1])


def scan_f64(bits, idxs, y, U, corr_thr):
    """FP64 statistic for every row of `bits` (uint64 [m, row_words]).

    Returns dict(c, passed, xxM, xyM, T, p, effect, n_passed).
    """
    bits = np.ascontiguousarray(bits, dtype=np.uint64)
    m, rw = bits.shape
    idxs = np.ascontiguousarray(idxs, dtype=np.int32)
    y = np.ascontiguousarray(y, dtype=np.float64)
    Uf = np.asfortranarray(U, dtype=np.float64)
    n, p = Uf.shape
    assert len(idxs) == n and len(y) == n
    c = np.empty(m, np.int32)
    ok = np.empty(m, np.uint8)
    xx, xy, T, pv, ef = (np.empty(m) for _ in range(5))
    npass = lib().or_scan_f64(_ptr(bits, _u64p), m, rw, _ptr(idxs, _i32p), n, _ptr(y, _dp),
                              _ptr(Uf, _dp), p, float(corr_thr), _ptr(c, _i32p), _ptr(ok, _u8p),
                              _ptr(xx, _dp), _ptr(xy, _dp), _ptr(T, _dp), _ptr(pv, _dp),
                              _ptr(ef, _dp))
    return dict(c=c, passed=ok.astype(bool), xxM=xx, xyM=xy, T=T, p=pv, effect=ef,
                n_passed=int(npass))


def pairs_f64(bits, idxs, y, U, log10_thresh, i0=0, i1=None):
    """FP64 all-pairs scan (src/kgwas_2d.cpp:258-401) of rows i < j, i in [i0, i1).

    Returns dict(i, j, c, xxM, xyM, effect, T, p) of the pairs with
    p <= 10^-log10_thresh, sorted by (i, j).
    """
    bits = np.ascontiguousarray(bits, dtype=np.uint64)
    N, rw = bits.shape
    idxs = np.ascontiguousarray(idxs, dtype=np.int32)
    y = np.ascontiguousarray(y, dtype=np.float64)
    Uf = np.asfortranarray(U, dtype=np.float64)
    n, p = Uf.shape
    assert len(idxs) == n and len(y) == n
    i1 = N if i1 is None else i1
    cap = 1024
    while True:
        ij = np.empty((cap, 2), np.int64)
        c = np.empty(cap, np.int32)
        xx, xy, ef, T, pv = (np.empty(cap) for _ in range(5))
        k = lib().or_pairs_f64(_ptr(bits, _u64p), N, rw, _ptr(idxs, _i32p), n, _ptr(y, _dp),
                               _ptr(Uf, _dp), p, int(i0), int(i1), 10.0 ** (-log10_thresh), cap,
                               _ptr(ij, _i64p), _ptr(c, _i32p), _ptr(xx, _dp), _ptr(xy, _dp),
                               _ptr(ef, _dp), _ptr(T, _dp), _ptr(pv, _dp))
        if k <= cap:
            break
        cap = int(k)
    o = np.lexsort((ij[:k, 1], ij[:k, 0]))
    return dict(i=ij[:k, 0][o], j=ij[:k, 1][o], c=c[:k][o], xxM=xx[:k][o], xyM=xy[:k][o],
                effect=ef[:k][o], T=T[:k][o], p=pv[:k][o])


def unpack_rows(bits, idxs):
    """Row-major float32 matrix X[n, m] (phenotyped samples x k-mers) from packed bits,
    the layout src/kgwas_utils.hpp:225-282 builds."""
    bits = np.ascontiguousarray(bits, dtype=np.uint64)
    m, rw = bits.shape
    idxs = np.ascontiguousarray(idxs, dtype=np.int32)
    X = np.empty((len(idxs), m), np.float32)
    lib().or_unpack_rows(_ptr(bits, _u64p), m, rw, _ptr(idxs, _i32p), len(idxs), _ptr(X, _fp))
    return X


def cols_fp32(X, y, U, corr_thr, log10_thresh, log10_thresh_2d, threads):
    """FP32 restatement of run_cols (pre-filter + LA + emit) on X[n, N] row-major."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, N = X.shape
    y = np.ascontiguousarray(y, dtype=np.float64)
    Uf = np.asfortranarray(U, dtype=np.float64)
    h1, h2, tb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    ts = np.zeros(3)
    xx = np.empty(N, np.float32)
    xy = np.empty(N, np.float32)
    keep = np.empty(N, np.int64)
    npass = lib().cpu_cols_fp32(_ptr(X, _fp), n, N, _ptr(y, _dp), _ptr(Uf, _dp), Uf.shape[1],
                                float(corr_thr), float(log10_thresh), float(log10_thresh_2d),
                                int(threads), ctypes.byref(h1), ctypes.byref(h2), _ptr(ts, _dp),
                                _ptr(xx, _fp), _ptr(xy, _fp), _ptr(keep, _i64p), ctypes.byref(tb))
    return dict(n_passed=int(npass), hits_1d=h1.value, hits_2d=h2.value, t_stage=ts,
                xxM=xx[:npass], xyM=xy[:npass], keep=keep[:npass], text_bytes=tb.value)


def load_columns(bits_dir, idxs, N, threads):
    """load_matrix_rows restated (src/kgwas_utils.hpp:225-282): X[n, N] float32 from
    plain .txt columns."""
    idxs = np.ascontiguousarray(idxs, dtype=np.int32)
    X = np.empty((len(idxs), N), np.float32)
    rc = lib().cpu_load_columns(bits_dir.encode(), _ptr(idxs, _i32p), len(idxs), int(N),
                                _ptr(X, _fp), int(threads))
    if rc != 0:
        raise RuntimeError(f"column of sample {-1 - rc} missing or short")
    return X
