"""kg_scan_device_noalloc: the allocation-free scan equals kg_scan_device.

Two calls in a row on one context, the second with a smaller input (the
scratch array keeps the first call's contents), each compared with
kg_scan_device on the same rows: same hits (index, trait, count, FP64 values,
2D flag) and the same pre-filter counts; and a scan with more candidates than
the scratch holds fails with KG_ERR_SCRATCH and a message.
"""
import numpy as np
import pytest

from supergwas_amd import synth

pytestmark = pytest.mark.gpu


def _same(a, b):
    (h0, n0), (h1, n1) = a, b
    assert np.array_equal(n0, n1)
    assert len(h0) == len(h1)
    for f in h0.dtype.names:
        assert np.array_equal(h0[f], h1[f]), f


def test_noalloc_equals_scan_device(built):
    from supergwas_amd.api import KG_ERR_SCRATCH, KgError, KgwasHip, ScanParams
    m1, m2 = 200_000, 37_001
    ds = synth.make_dataset(seed=2, **dict(synth.CONFIGS["mixed"], m=m1))
    bits = ds.fill_bits(0, m1)
    prm = ScanParams(0.2, 3.0, 6.0)
    with KgwasHip() as k:
        k.set_model(ds.n_acc, ds.idxs, ds.U, ds.y)
        dev = k.to_device(bits)
        rw = bits.shape[1]
        for m in (m1, m2):
            ref = k.scan_device(dev.ptr, m, rw, prm)
            got = k.scan_device_noalloc(dev.ptr, m, rw, prm)
            print(f"[noalloc] m={m} hits={len(got[0])} passed={int(got[1][0])} "
                  f"candidates={k.timing()['candidates']}", flush=True)
            assert len(ref[0]) > 0
            _same(ref, got)
        # tau = 0 and p <= 1: every polymorphic k-mer is a candidate, more than the scratch holds
        with pytest.raises(KgError) as ei:
            k.scan_device_noalloc(dev.ptr, m1, rw, ScanParams(0.0, 0.0, 0.0))
        assert ei.value.code == KG_ERR_SCRATCH and "candidates exceed" in str(ei.value)
        # the context and the scratch still work after the refusal
        _same(k.scan_device(dev.ptr, m2, rw, prm), k.scan_device_noalloc(dev.ptr, m2, rw, prm))
        dev.free()
