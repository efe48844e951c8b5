"""Host-side synthetic data: determinism and model properties."""
import numpy as np

from supergwas_amd import synth


def test_fill_is_deterministic_and_offset_consistent(built):
    ds = synth.make_dataset(seed=11, n_acc=130, n_missing=5, npop=3, p=4, m=500)
    a = ds.fill_bits(0, 200)
    b = ds.fill_bits(50, 100)
    assert np.array_equal(a[50:150], b)
    # bits beyond n_acc are zero
    assert np.all(a[:, 2] >> np.uint64(130 - 128) == 0)
    assert np.all(a[:, 3:] == 0)


def test_U_orthonormal_with_intercept(built):
    ds = synth.make_dataset(seed=1, n_acc=300, n_missing=10, npop=4, p=6, m=100)
    U = ds.U
    assert U.shape == (290, 6)
    assert np.allclose(U.T @ U, np.eye(6), atol=1e-12)
    one = np.ones(290) / np.sqrt(290)
    assert np.allclose(U @ (U.T @ one), one, atol=1e-12)   # 1 in span(U)


def test_allele_frequencies_respect_maf(built):
    ds = synth.make_dataset(seed=2, n_acc=1000, n_missing=0, npop=5, p=3, m=100)
    bits = ds.fill_bits(0, 3000)
    import oracle
    X = oracle.unpack_rows(bits, ds.idxs)
    f = X.mean(axis=0)
    # expected MAF >= 0.05 per profile; realised frequencies stay near that range
    assert np.mean((f < 0.02) | (f > 0.98)) < 0.01
