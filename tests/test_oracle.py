"""Pins the FP64 oracle (oracle/kgwas_oracle.c) before it is trusted as the checker.

Checks: golden vectors from an independent least-squares derivation and brute
force over all bit patterns (tests/golden/make_golden.py), scipy's chi2.sf,
numpy's corrcoef for the pre-filter, the special cases of compute_hit, and
invariances (sample permutation, U -> UQ).  Planted-mistake checks show the
golden vectors reject plausible slips (effect = xyM/xxM, a flipped sign, a
dropped projection term).
"""
import os

import numpy as np
import pytest
from scipy import stats

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "golden_small.npz"))


def golden(name):
    return {k.split("_", 1)[1]: G[k] for k in G.files if k.startswith(name + "_")}


@pytest.mark.parametrize("name", ["A", "B", "C"])
def test_oracle_matches_golden_vectors(built, name):
    g = golden(name)
    r = oracle.scan_f64(g["bits"], g["idxs"], g["y"], g["U"], 0.0)
    assert np.array_equal(r["c"], g["c"])
    ok = g["T"] > 1e-9
    assert np.allclose(r["T"][ok], g["T"][ok], rtol=1e-9, atol=1e-12)
    assert np.allclose(r["T"][~ok], g["T"][~ok], atol=1e-9)
    assert np.allclose(r["effect"], g["effect"], rtol=1e-9, atol=1e-13)
    assert np.allclose(r["p"], g["p"], rtol=1e-9, atol=1e-300)


@pytest.mark.parametrize("tau", [0.0, 0.2, 0.5])
@pytest.mark.parametrize("name", ["A", "B", "C"])
def test_oracle_prefilter_matches_corrcoef(built, name, tau):
    g = golden(name)
    r = oracle.scan_f64(g["bits"], g["idxs"], g["y"], g["U"], tau)
    cor = g["cor"]
    poly = ~np.isnan(cor)
    exp = poly & (np.abs(np.nan_to_num(cor)) >= tau)
    band = poly & (np.abs(np.abs(np.nan_to_num(cor)) - tau) <= 1e-12)
    assert np.array_equal(r["passed"][~band], exp[~band])
    assert r["n_passed"] == int(r["passed"].sum())
    # monomorphic k-mers are dropped even at tau = 0 (src/kgwas.cpp:465)
    assert not r["passed"][~poly].any()


def test_chi2_sf_matches_scipy(built):
    for T in np.concatenate([np.linspace(0.01, 50, 200), [100.0, 500.0, 1000.0, 1400.0]]):
        assert oracle.chi2_sf_k1(T) == pytest.approx(stats.chi2.sf(T, 1), rel=1e-12, abs=1e-300)
    assert oracle.chi2_sf_k1(0.0) == 1.0
    assert oracle.chi2_sf_k1(-1.0) == 1.0


def test_compute_hit_special_cases(built):
    n = 100
    assert oracle.compute_hit(1e-13, 5.0, 10.0, n) == (0.0, 0.0, 1.0)   # xxM <= eps
    assert oracle.compute_hit(2.0, 5.0, 1e-13, n) == (0.0, 0.0, 1.0)    # yyM <= eps
    e, T, p = oracle.compute_hit(1.0, 10.0, 1.0, n)                     # r2 >= 1 clamp
    assert T == pytest.approx(n * np.log(1e15), rel=1e-3)
    assert e == 10.0
    e, T, p = oracle.compute_hit(4.0, 2.0, 9.0, n)                      # r2 = 1/9
    assert T == pytest.approx(n * np.log(9.0 / 8.0), rel=1e-14)
    assert e == pytest.approx(2.0 / 36.0, rel=1e-15)
    assert p == pytest.approx(stats.chi2.sf(T, 1), rel=1e-12)


def test_T_threshold(built):
    # T*(p) for the 1D and 2D defaults (10^-6, 10^-15)
    for L, T_exp in ((6, 23.928127), (15, 64.430464)):
        Ts = oracle.T_threshold(10.0 ** -L)
        assert Ts == pytest.approx(T_exp, abs=2e-6)
        assert oracle.chi2_sf_k1(Ts) <= 10.0 ** -L
        assert oracle.chi2_sf_k1(np.nextafter(Ts, 0)) > 10.0 ** -L


def test_invariance_permutation_and_rotation(built):
    g = golden("C")
    r0 = oracle.scan_f64(g["bits"], g["idxs"], g["y"], g["U"], 0.1)
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    r1 = oracle.scan_f64(g["bits"], g["idxs"], g["y"], g["U"] @ Q, 0.1)
    assert np.allclose(r0["T"], r1["T"], rtol=1e-10, atol=1e-12)
    assert np.array_equal(r0["passed"], r1["passed"])
    # permuting samples: reorder the accession bits, idxs stay increasing
    n = len(g["idxs"])
    perm = rng.permutation(n)
    X = oracle.unpack_rows(g["bits"], g["idxs"]).T            # m x n
    from tests.golden.make_golden import pack
    xacc = np.zeros((X.shape[0], n))
    xacc[:, :] = X[:, perm]
    bits_p = pack(xacc)
    r2 = oracle.scan_f64(bits_p, np.arange(n, dtype=np.int32), g["y"][perm], g["U"][perm], 0.1)
    assert np.allclose(r0["T"], r2["T"], rtol=1e-10, atol=1e-12)
    assert np.array_equal(r0["passed"], r2["passed"])


def test_planted_mistakes_are_caught(built):
    """Plausible slips in the statistic would fail the golden check."""
    g = golden("A")
    r = oracle.scan_f64(g["bits"], g["idxs"], g["y"], g["U"], 0.0)
    ok = (g["T"] > 1e-6) & (np.abs(r["effect"]) > 1e-6)
    yyM = oracle.yyM(g["y"], g["U"])
    beta = r["xyM"] / np.where(r["xxM"] > 1e-12, r["xxM"], np.nan)
    assert not np.allclose(beta[ok], g["effect"][ok], rtol=1e-6)        # effect = OLS beta
    assert np.allclose(beta[ok] / yyM, g["effect"][ok], rtol=1e-9)      # the actual formula
    assert not np.allclose(-r["effect"][ok], g["effect"][ok], rtol=1e-6)  # sign slip
    # dropping the projection term from xxM (xxM = c) changes T
    n = len(g["idxs"])
    T_bad = n * np.log(1.0 / (1.0 - r["xyM"] ** 2 / (r["c"] * yyM)))
    assert not np.allclose(T_bad[ok], g["T"][ok], rtol=1e-6)


def test_fp32_restatement_close_to_fp64(built):
    from supergwas_amd import synth
    ds = synth.make_dataset(seed=4, n_acc=200, n_missing=10, npop=3, p=5, m=5000)
    bits = ds.fill_bits(0, 5000)
    ref = oracle.scan_f64(bits, ds.idxs, ds.y, ds.U, 0.2)
    X = oracle.unpack_rows(bits, ds.idxs)
    cpu = oracle.cols_fp32(X, ds.y, ds.U, 0.2, 2.0, 4.0, 2)
    assert abs(cpu["n_passed"] - ref["n_passed"]) <= 3
    keep = cpu["keep"]
    both = ref["passed"][keep]
    c = ref["c"][keep][both].astype(np.float64)
    assert np.max(np.abs(cpu["xxM"][both] - ref["xxM"][keep][both]) / c) < 1e-4
    assert np.max(np.abs(cpu["xyM"][both] - ref["xyM"][keep][both]) / c) < 1e-4


# ---- all-pairs scan (src/kgwas_2d.cpp:258-401) ----

def _pair_setup(seed, intercept, n_acc=40, N=26, p=3):
    from tests.golden.make_golden import orthonormal, pack
    rng = np.random.default_rng(seed)
    xacc = (rng.random((N, n_acc)) < rng.uniform(0.3, 0.9, (N, 1))).astype(np.uint8)
    idxs = np.sort(rng.choice(n_acc, size=n_acc - 6, replace=False)).astype(np.int32)
    # two rows carried by every phenotyped sample: their pair has c = n
    xacc[3, idxs] = 1
    xacc[7, idxs] = 1
    n = len(idxs)
    U = orthonormal(rng, n, p, intercept)
    y = rng.standard_normal(n) + 0.8 * xacc[1, idxs] * xacc[2, idxs]
    return pack(xacc), xacc[:, idxs].astype(np.float64), idxs, y, U


def _pairs_by_lstsq(X, y, U, log10, op=np.logical_and):
    """Expected pairs from independent least-squares fits of y on [U] and [U, x_i AND x_j]."""
    from tests.golden.make_golden import expected
    N = X.shape[0]
    ii, jj = np.triu_indices(N, 1)
    X12 = op(X[ii] > 0, X[jj] > 0).astype(np.float64)
    c = X12.sum(axis=1)
    T, p, eff = expected(X12, y, U)[:3]
    keep = (c > 0) & (T > 0) & (p <= 10.0 ** (-log10))
    return ii[keep], jj[keep], c[keep].astype(np.int32), T[keep], eff[keep], p[keep]


@pytest.mark.parametrize("intercept", [True, False])
def test_pairs_oracle_matches_lstsq(built, intercept):
    bits, X, idxs, y, U = _pair_setup(11 if intercept else 12, intercept)
    log10 = 0.4
    r = oracle.pairs_f64(bits, idxs, y, U, log10)
    ei, ej, ec, eT, eeff, ep = _pairs_by_lstsq(X, y, U, log10)
    assert len(ei) > 20
    assert np.array_equal(r["i"], ei) and np.array_equal(r["j"], ej)
    assert np.array_equal(r["c"], ec)
    assert np.allclose(r["T"], eT, rtol=1e-9, atol=1e-12)
    assert np.allclose(r["effect"], eeff, rtol=1e-9, atol=1e-13)
    assert np.allclose(r["p"], ep, rtol=1e-9, atol=1e-300)
    # the pair (3, 7) has c = n: tested only when 1 is outside span(U)
    full = (r["i"] == 3) & (r["j"] == 7)
    assert r["c"][full].tolist() in ([], [len(idxs)])
    T37 = _pairs_by_lstsq(X, y, U, 0.0)
    k37 = (T37[0] == 3) & (T37[1] == 7)
    assert k37.sum() == (0 if intercept else 1)


def test_pairs_oracle_equals_1d_on_and_rows(built):
    """Each pair's statistic is the 1D statistic of the AND row (src/kgwas_2d.cpp:372-390)."""
    bits, X, idxs, y, U = _pair_setup(13, True)
    r = oracle.pairs_f64(bits, idxs, y, U, 0.0)
    and_rows = bits[r["i"]] & bits[r["j"]]
    one = oracle.scan_f64(and_rows, idxs, y, U, 0.0)
    assert np.array_equal(one["c"], r["c"])
    assert np.allclose(one["T"], r["T"], rtol=1e-12, atol=1e-12)
    assert np.allclose(one["effect"], r["effect"], rtol=1e-12, atol=1e-14)


def test_pairs_oracle_row_ranges_and_planted_or(built):
    bits, X, idxs, y, U = _pair_setup(14, True)
    whole = oracle.pairs_f64(bits, idxs, y, U, 0.4)
    a = oracle.pairs_f64(bits, idxs, y, U, 0.4, 0, 9)
    b = oracle.pairs_f64(bits, idxs, y, U, 0.4, 9, None)
    assert np.array_equal(np.concatenate([a["i"], b["i"]]), whole["i"])
    assert np.array_equal(np.concatenate([a["j"], b["j"]]), whole["j"])
    assert np.all(a["i"] < 9) and np.all(b["i"] >= 9)
    # x_i OR x_j in place of AND would be caught
    oi, oj = _pairs_by_lstsq(X, y, U, 0.4, op=np.logical_or)[:2]
    assert not (np.array_equal(oi, whole["i"]) and np.array_equal(oj, whole["j"]))
