"""Python host binding of libkgwas_hip's C ABI (include/kgwas_hip.h) via ctypes.

This mirrors the reference's in-process seam of `kgwas --mode cols`
(src/kgwas.cpp:183-289): set the null model once (U.bin + aligned phenotype),
then scan packed k-mer rows and receive the reported k-mers with the values
compute_hit prints (src/kgwas.cpp:44-56) plus the pre-filter pass count
(`.count`, src/kgwas.cpp:292-304).  Errors raise KgError carrying the
library's message, as die() ends the reference with "[ERROR] msg".

There is no CPU fallback: without the built HIP library every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libkgwas_hip.so")

KG_OK, KG_ERR_ARG, KG_ERR_DEVICE, KG_ERR_STATE, KG_ERR_CAPACITY, KG_ERR_SCRATCH = 0, 1, 2, 3, 4, 5
PRECISIONS = {"mixed": 0, "fp32": 1}
# pre-filter screen of mixed scans (KG_SCREEN_*)
SCREEN = {"auto": 0, "off": 1, "on": 2}
HIT_2D = 1

HIT_DTYPE = np.dtype([("idx", np.int64), ("trait", np.int32), ("c", np.int32),
                      ("flags", np.uint32), ("shard", np.int32), ("xxM", np.float64),
                      ("xyM", np.float64), ("effect", np.float64), ("T", np.float64),
                      ("pval", np.float64)])


class KgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[kgwas_hip error {code}] {msg}")
        self.code = code


class _Opts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("limbs", ctypes.c_int32), ("screen", ctypes.c_int32),
                ("dedup", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


class ScanParams(ctypes.Structure):
    """CORR_THRESHOLD, log10 threshold, 2D log10 threshold (src/kgwas.cpp:28-30)."""
    _fields_ = [("corr_threshold", ctypes.c_double), ("log10_thresh", ctypes.c_double),
                ("log10_thresh_2d", ctypes.c_double)]

    def __init__(self, corr_threshold=0.2, log10_thresh=6.0, log10_thresh_2d=15.0):
        super().__init__(corr_threshold, log10_thresh, log10_thresh_2d)


class Shard(ctypes.Structure):
    """One shard of a batched scan (kg_shard): device rows + model id."""
    _fields_ = [("d_bits", ctypes.c_void_p), ("m", ctypes.c_int64), ("row_words", ctypes.c_int64),
                ("model", ctypes.c_int32), ("pad", ctypes.c_int32)]


SHARDS_PER_LAUNCH = 1
SHARDS_GRAPH = 2


class Tune(ctypes.Structure):
    """kg_tune: launch shape of the int8 scan."""
    _fields_ = [("tiles_per_wave", ctypes.c_int32), ("workgroups_per_cu", ctypes.c_int32),
                ("screen_wgs_per_cu", ctypes.c_int32), ("kmers_per_workgroup", ctypes.c_int32),
                ("screen_kmers_per_workgroup", ctypes.c_int32),
                ("stream_weights", ctypes.c_int32), ("resident_active", ctypes.c_int32),
                ("resident_form", ctypes.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class _Timing(ctypes.Structure):
    _fields_ = [("bulk_ms", ctypes.c_double), ("refine_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("candidates", ctypes.c_int64),
                ("kmers", ctypes.c_int64), ("copy_ms", ctypes.c_double),
                ("wall_ms", ctypes.c_double), ("screen_ms", ctypes.c_double),
                ("survivors", ctypes.c_int64), ("stat_refined", ctypes.c_int64),
                ("rechecked", ctypes.c_int64)]


_lib = None
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32

# Every entry point the header declares, with its ctypes signature.
SIGNATURES = {
    "kg_create": ([ctypes.POINTER(_vp), ctypes.POINTER(_Opts)], ctypes.c_int),
    "kg_destroy": ([_vp], None),
    "kg_last_error": ([_vp], ctypes.c_char_p),
    "kg_set_model": ([_vp, _i32, _vp, _i32, _vp, _i32, _vp, _i32], ctypes.c_int),
    "kg_scan": ([_vp, _vp, _i64, _i64, ctypes.POINTER(ScanParams), _vp, _i64,
                 ctypes.POINTER(_i64), _vp], ctypes.c_int),
    "kg_scan_device": ([_vp, _vp, _i64, _i64, ctypes.POINTER(ScanParams), _vp, _i64,
                        ctypes.POINTER(_i64), _vp], ctypes.c_int),
    "kg_scan_device_noalloc": ([_vp, _vp, _i64, _i64, ctypes.POINTER(ScanParams), _vp, _i64,
                                ctypes.POINTER(_i64), _vp], ctypes.c_int),
    "kg_add_model": ([_vp, _i32, _vp, _i32, _vp, _i32, _vp, _i32, ctypes.POINTER(_i32)],
                     ctypes.c_int),
    "kg_scan_shards": ([_vp, _vp, _i32, ctypes.POINTER(ScanParams), ctypes.c_uint32, _vp, _i64,
                        ctypes.POINTER(_i64), _vp], ctypes.c_int),
    "kg_scan_pairs": ([_vp, _vp, _i64, _i64, _i64, _i64, ctypes.POINTER(ScanParams), _vp, _i64,
                       ctypes.POINTER(_i64)], ctypes.c_int),
    "kg_stats_device": ([_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp], ctypes.c_int),
    "kg_last_timing": ([_vp, ctypes.POINTER(_Timing)], ctypes.c_int),
    "kg_set_screen": ([_vp, _i32], ctypes.c_int),
    "kg_set_dedup": ([_vp, _i32], ctypes.c_int),
    "kg_dev_alloc": ([_vp, ctypes.c_size_t, ctypes.POINTER(_vp)], ctypes.c_int),
    "kg_dev_free": ([_vp, _vp], ctypes.c_int),
    "kg_host_alloc": ([_vp, ctypes.c_size_t, ctypes.POINTER(_vp)], ctypes.c_int),
    "kg_autotune": ([_vp, _i64, ctypes.POINTER(Tune)], ctypes.c_int),
    "kg_set_tune": ([_vp, ctypes.POINTER(Tune)], ctypes.c_int),
    "kg_get_tune": ([_vp, ctypes.POINTER(Tune)], ctypes.c_int),
    "kg_host_free": ([_vp, _vp], ctypes.c_int),
    "kg_memcpy": ([_vp, _vp, _vp, ctypes.c_size_t, _i32], ctypes.c_int),
    "kg_synth_bits": ([_vp, _vp, ctypes.c_uint64, _i64, _i64, _i32, _i64, _vp, _vp, _i32, _i32],
                      ctypes.c_int),
}


def lib():
    """Load libkgwas_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KgError(KG_ERR_STATE, f"{LIB_PATH} missing: build the HIP library first "
                          "(python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


class DeviceArray:
    """Device allocation owned by a KgwasHip context (freed on close/garbage)."""

    def __init__(self, ctx, ptr, nbytes):
        self._ctx = ctx
        self.ptr = ptr
        self.nbytes = nbytes
        self.shape = None
        self.dtype = None

    def to_host(self, dtype=None, shape=None, nbytes=None, offset=0):
        dtype = np.dtype(dtype or self.dtype)
        nb = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(nb // dtype.itemsize, dtype)
        self._ctx._check(lib().kg_memcpy(self._ctx._h, _p(out), self.ptr + offset, out.nbytes, 1))
        return out.reshape(shape) if shape is not None else out

    def free(self):
        if self.ptr and self._ctx._h and self._ctx._h.value:
            lib().kg_dev_free(self._ctx._h, self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class KgwasHip:
    """One device context (kg_create) with one null model (kg_set_model)."""

    def __init__(self, device=0, precision="mixed", limbs=3, screen="auto", dedup=False):
        L = lib()
        opts = _Opts(device, PRECISIONS[precision], limbs, SCREEN[screen], 1 if dedup else 0)
        h = _vp()
        rc = L.kg_create(ctypes.byref(h), ctypes.byref(opts))
        self._h = h
        if rc != KG_OK:
            msg = L.kg_last_error(h).decode() if h.value else "kg_create failed"
            if h.value:
                L.kg_destroy(h)
            self._h = _vp()
            raise KgError(rc, msg)
        self.precision = precision
        self.t = 0
        self.n = 0
        self._pinned = {}

    def close(self):
        if self._h and self._h.value:
            for ptr in list(getattr(self, "_pinned", {}).values()):
                lib().kg_host_free(self._h, ptr)
            self._pinned = {}
            lib().kg_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != KG_OK:
            raise KgError(rc, lib().kg_last_error(self._h).decode())

    def set_model(self, n_acc, idxs, U, Y):
        """idxs: phenotyped accession positions (increasing); U: n x p; Y: n or n x t."""
        idxs = np.ascontiguousarray(idxs, np.int32)
        U = np.asfortranarray(np.asarray(U, np.float64).reshape(len(idxs), -1))
        Y = np.asarray(Y, np.float64)
        if Y.ndim == 1:
            Y = Y[:, None]
        Y = np.asfortranarray(Y)
        self.t = Y.shape[1]
        self.n = len(idxs)
        self._check(lib().kg_set_model(self._h, int(n_acc), _p(idxs), len(idxs),
                                       _p(U) if U.shape[1] else None, U.shape[1], _p(Y),
                                       Y.shape[1]))

    def add_model(self, n_acc, idxs, U, Y):
        """A further null model (own sample mask / U / phenotypes); returns its id."""
        idxs = np.ascontiguousarray(idxs, np.int32)
        U = np.asfortranarray(np.asarray(U, np.float64).reshape(len(idxs), -1))
        Y = np.asarray(Y, np.float64)
        if Y.ndim == 1:
            Y = Y[:, None]
        Y = np.asfortranarray(Y)
        mid = _i32(0)
        self._check(lib().kg_add_model(self._h, int(n_acc), _p(idxs), len(idxs),
                                       _p(U) if U.shape[1] else None, U.shape[1], _p(Y),
                                       Y.shape[1], ctypes.byref(mid)))
        return mid.value

    def scan_shards(self, shards, params=None, flags=0, cap=1 << 16):
        """shards: list of (dev_ptr, m, row_words, model_id).  Returns (hits, n_passed[s, t])."""
        params = params or ScanParams()
        arr = (Shard * len(shards))(*[Shard(p, m, rw, mid, 0) for (p, m, rw, mid) in shards])
        np_ = np.zeros(len(shards) * max(self.t, 1), np.int64)
        while True:
            out = np.zeros(max(cap, 1), HIT_DTYPE)
            n_out = _i64(0)
            rc = lib().kg_scan_shards(self._h, ctypes.cast(arr, _vp), len(shards),
                                      ctypes.byref(params), flags, _p(out), cap,
                                      ctypes.byref(n_out), _p(np_))
            if rc == KG_ERR_CAPACITY:
                cap = int(n_out.value)
                continue
            self._check(rc)
            return out[:n_out.value], np_.reshape(len(shards), -1).copy()

    def _hit_buffer(self, cap):
        # reused across calls; callers get a copy of the filled part
        buf = getattr(self, "_hits", None)
        if buf is None or len(buf) < cap:
            buf = np.empty(max(cap, 1), HIT_DTYPE)
            self._hits = buf
        return buf

    def _scan(self, fn, ptr, m, row_words, params, cap):
        params = params or ScanParams()
        np_ = np.zeros(max(self.t, 1), np.int64)
        while True:
            out = self._hit_buffer(cap)
            n_out = _i64(0)
            rc = fn(self._h, ptr, int(m), int(row_words), ctypes.byref(params), _p(out), cap,
                    ctypes.byref(n_out), _p(np_))
            if rc == KG_ERR_CAPACITY:
                cap = int(n_out.value)
                continue
            self._check(rc)
            return out[:n_out.value].copy(), np_.copy()

    def scan(self, bits, params=None, cap=1 << 16):
        """Host rows (uint64 [m, row_words]) -> (hits, n_passed per trait)."""
        bits = np.ascontiguousarray(bits, np.uint64)
        return self._scan(lib().kg_scan, _p(bits), bits.shape[0], bits.shape[1], params, cap)

    def scan_device(self, dev_ptr, m, row_words, params=None, cap=1 << 16):
        """Rows already in device memory at dev_ptr (e.g. a torch tensor's data_ptr())."""
        return self._scan(lib().kg_scan_device, dev_ptr, m, row_words, params, cap)

    def scan_device_noalloc(self, dev_ptr, m, row_words, params=None, cap=1 << 16):
        """kg_scan_device without any allocation (fixed per-GPU scratch; mixed precision)."""
        return self._scan(lib().kg_scan_device_noalloc, dev_ptr, m, row_words, params, cap)

    def scan_pairs(self, dev_ptr, n_rows, row_words, params=None, i0=0, i1=None, cap=1 << 16):
        """All-pairs scan of device rows (src/kgwas_2d.cpp:258-401): hits with
        shard = i and idx = j (i < j, i in [i0, i1)), sorted by (i, j, trait)."""
        params = params or ScanParams()
        i1 = n_rows if i1 is None else i1
        while True:
            out = self._hit_buffer(cap)
            n_out = _i64(0)
            rc = lib().kg_scan_pairs(self._h, dev_ptr, int(n_rows), int(row_words), int(i0),
                                     int(i1), ctypes.byref(params), _p(out), cap,
                                     ctypes.byref(n_out))
            if rc == KG_ERR_CAPACITY:
                cap = int(n_out.value)
                continue
            self._check(rc)
            return out[:n_out.value].copy()

    def stats_device(self, dev_ptr, m, row_words):
        """Bulk-kernel values for every k-mer: c, xxM, xyM[m, t], S[m, t]."""
        t = max(self.t, 1)
        c = np.zeros(m, np.int32)
        xx = np.zeros(m)
        xy = np.zeros((m, t))
        S = np.zeros((m, t))
        self._check(lib().kg_stats_device(self._h, dev_ptr, int(m), int(row_words), _p(c), _p(xx),
                                          _p(xy), _p(S)))
        return dict(c=c, xxM=xx, xyM=xy, S=S)

    def synth_bits(self, dev_ptr, ds, kmer0, m, row_words):
        pop = np.ascontiguousarray(ds.pop, np.uint8)
        thr = np.ascontiguousarray(ds.thr, np.uint32)
        self._check(lib().kg_synth_bits(self._h, dev_ptr, ds.seed, int(kmer0), int(m),
                                        ds.n_acc, int(row_words), _p(pop), _p(thr), ds.npop,
                                        ds.n_profiles))

    # ---- device memory (no PyTorch needed) ----
    def alloc(self, nbytes):
        p = _vp()
        self._check(lib().kg_dev_alloc(self._h, int(nbytes), ctypes.byref(p)))
        return DeviceArray(self, p.value, int(nbytes))

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        d = self.alloc(arr.nbytes)
        self._check(lib().kg_memcpy(self._h, d.ptr, _p(arr), arr.nbytes, 0))
        d.shape, d.dtype = arr.shape, arr.dtype
        return d

    def set_dedup(self, on):
        """Duplicate-row scans on (True) or off for later scans (kg_set_dedup)."""
        self._check(lib().kg_set_dedup(self._h, 1 if on else 0))

    def set_screen(self, mode):
        """Pre-filter screen mode for later scans: "auto", "off" or "on"."""
        self._check(lib().kg_set_screen(self._h, SCREEN[mode]))

    def timing(self):
        t = _Timing()
        self._check(lib().kg_last_timing(self._h, ctypes.byref(t)))
        return dict(bulk_ms=t.bulk_ms, refine_ms=t.refine_ms, total_ms=t.total_ms,
                    candidates=t.candidates, kmers=t.kmers, copy_ms=t.copy_ms, wall_ms=t.wall_ms,
                    screen_ms=t.screen_ms, survivors=t.survivors, stat_refined=t.stat_refined,
                    rechecked=t.rechecked)

    def autotune(self, sample=2_000_000):
        """Pick the launch shape (k-mer tiles per wave, full-kernel and screen workgroups
        per CU) by timing `sample` random rows with model 0; returns it as a dict."""
        t = Tune()
        self._check(lib().kg_autotune(self._h, int(sample), ctypes.byref(t)))
        return t.as_dict()

    def set_tune(self, tiles_per_wave=0, workgroups_per_cu=1, screen_wgs_per_cu=1,
                 stream_weights=0, resident_form=0):
        """Set the launch shape explicitly (kg_set_tune); stream_weights=1 keeps
        single-trait screened scans off the resident-weight kernel; resident_form
        picks its operands (1 int8 limbs, 2 or 3 FP6 digits, 0 the default)."""
        t = Tune(tiles_per_wave, workgroups_per_cu, screen_wgs_per_cu)
        t.stream_weights = int(stream_weights)
        t.resident_form = int(resident_form)
        self._check(lib().kg_set_tune(self._h, ctypes.byref(t)))

    def get_tune(self):
        t = Tune()
        self._check(lib().kg_get_tune(self._h, ctypes.byref(t)))
        return t.as_dict()

    def host_alloc(self, shape, dtype=np.uint64):
        """Page-locked host array (numpy view); kg_scan streams such rows without a
        bounce copy.  Free with host_free(arr) before the context closes."""
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = _vp()
        self._check(lib().kg_host_alloc(self._h, nbytes, ctypes.byref(p)))
        buf = (ctypes.c_uint8 * max(nbytes, 1)).from_address(p.value)
        arr = np.frombuffer(buf, dtype=np.uint8, count=nbytes).view(dtype).reshape(shape)
        self._pinned[arr.ctypes.data] = p.value
        return arr

    def host_free(self, arr):
        ptr = self._pinned.pop(arr.ctypes.data, None)
        if ptr:
            self._check(lib().kg_host_free(self._h, ptr))
