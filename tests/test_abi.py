"""CPU checks of the C ABI: the library loads, exports exactly what the header
declares, and fails cleanly (no crash, a message) when no GPU is present."""
import ctypes
import os
import re

import pytest

from supergwas_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kgwas_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(kg_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("kg_create", "kg_destroy", "kg_last_error", "kg_set_model", "kg_scan",
              "kg_scan_device", "kg_stats_device", "kg_last_timing", "kg_synth_bits",
              "kg_add_model", "kg_scan_shards", "kg_dev_alloc", "kg_dev_free", "kg_memcpy"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    L = api.lib()
    for s in declared_symbols():
        assert hasattr(L, s), s
    assert set(declared_symbols()) == set(api.SIGNATURES)


def test_hit_struct_layout_matches_header():
    # kg_hit: int64, int32, int32, uint32, uint32, 5 doubles
    assert api.HIT_DTYPE.itemsize == 64


def test_create_without_gpu_reports_error():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(api.KgError) as ei:
        api.KgwasHip()
    assert ei.value.code == api.KG_ERR_DEVICE
    assert "device" in str(ei.value)


def test_bad_options_rejected():
    L = api.lib()
    h = ctypes.c_void_p()
    opts = api._Opts(0, 7, 3)
    assert L.kg_create(ctypes.byref(h), ctypes.byref(opts)) == api.KG_ERR_ARG
    opts = api._Opts(0, 0, 9)
    assert L.kg_create(ctypes.byref(h), ctypes.byref(opts)) == api.KG_ERR_ARG
    opts = api._Opts(0, 0, 3, 5)   # pre-filter screen mode out of range
    assert L.kg_create(ctypes.byref(h), ctypes.byref(opts)) == api.KG_ERR_ARG
    opts = api._Opts(0, 0, 3, 0, 2)   # duplicate-row flag is 0 or 1
    assert L.kg_create(ctypes.byref(h), ctypes.byref(opts)) == api.KG_ERR_ARG
    assert L.kg_set_screen(None, 0) == api.KG_ERR_ARG
    assert L.kg_set_dedup(None, 1) == api.KG_ERR_ARG


def test_option_and_timing_layouts():
    # kg_opts: 8 int32; kg_timing: 3 doubles, 2 int64, 3 doubles, 3 int64; kg_tune: 8 int32
    assert ctypes.sizeof(api._Opts) == 32
    assert ctypes.sizeof(api._Timing) == 88
    assert api._Opts.screen.offset == 12 and api._Timing.survivors.offset == 64
    assert api._Timing.stat_refined.offset == 72 and api._Timing.rechecked.offset == 80
    assert api._Opts.dedup.offset == 16
    assert ctypes.sizeof(api.Tune) == 32 and api.Tune.resident_form.offset == 28


def test_null_context_calls():
    L = api.lib()
    assert L.kg_last_error(None) == b"null context"
    assert L.kg_set_model(None, 1, None, 1, None, 0, None, 1) == api.KG_ERR_ARG
    L.kg_destroy(None)
