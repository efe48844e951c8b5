// FP64 ORACLE for the 1D k-mer association scan -- TEST INFRASTRUCTURE ONLY.
//
// This file is a CPU restatement of SuperGWAS's per-k-mer statistic in double
// precision.  It is the checker the HIP kernels are compared against and is
// never linked into, loaded by, or called from the product library or the
// kgwas drop-in binary.  Only tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg use it.
//
// Followed, line for line in FP64:
//   src/kgwas.cpp:44-56    compute_hit (r^2 clamp, T = n ln(1/(1-r^2)),
//                          p = erfc(sqrt(T/2)), effect = xyM/(xxM yyM))
//   src/kgwas.cpp:387-393  centred phenotype, Uy = U'y, yyM = y'y - |Uy|^2
//   src/kgwas.cpp:454-459  x(j) = bit at accession idxs[j]
//   src/kgwas.cpp:462-468  Pearson pre-filter (|x_c|<1e-12 or |y_c|<1e-12
//                          -> drop; |cor| < tau -> drop)
//   src/kgwas.cpp:472-474  Ux = U'x, xxM = x'x - |Ux|^2, xyM = x'y - Ux'Uy
//   src/chisq.hpp:6-9      chi2_sf_k1 (identical to the erfc above)
//
//   src/kgwas_2d.cpp:219-251, 258-401  all-pairs scan (or_pairs_f64)
//
// Pinned by tests/test_oracle.py: golden vectors written by
// tests/golden/make_golden.py from an independent least-squares fit
// (T = n ln(RSS0/RSS1), effect = beta_x / yyM), scipy's chi2.sf, numpy's
// corrcoef, brute force over all bit patterns of tiny n, and invariances
// (sample permutation, U -> UQ).
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

// src/kgwas.cpp:44-56
void or_compute_hit(double xxM, double xyM, double yyM, int n, double* effect, double* T,
                    double* pval) {
  const double eps = 1e-12;
  double e = 0.0, t = 0.0, p = 1.0;
  if (xxM > eps && yyM > eps) {
    double r2 = (xyM * xyM) / (xxM * yyM);
    if (r2 < 0) r2 = 0;
    if (r2 >= 1) r2 = 1 - 1e-15;
    t = (double)n * log(1.0 / (1.0 - r2));
    p = erfc(sqrt(0.5 * t));
    e = xyM / (xxM * yyM);
  }
  *effect = e;
  *T = t;
  *pval = p;
}

// src/chisq.hpp:6-9
double or_chi2_sf_k1(double T) {
  if (T <= 0.0) return 1.0;
  return erfc(sqrt(0.5 * T));
}

// Row-major copy of the column-major n x p U: the per-k-mer loops add U's
// row j for every set bit j, and a contiguous row keeps those loads in cache
// (the sums and their order are unchanged).
static double* rows_of(const double* U, int n, int p) {
  double* Ur = (double*)malloc(sizeof(double) * (size_t)n * (size_t)(p > 0 ? p : 1));
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < p; ++k) Ur[(size_t)j * p + k] = U[(size_t)k * n + j];
  return Ur;
}

// yyM = y'y - |U'y|^2 (src/kgwas.cpp:391-392, src/kgwas_utils.hpp:151-152).
double or_yyM(const double* y, int n, const double* U, int p) {
  double yy = 0.0, uu = 0.0;
  for (int j = 0; j < n; ++j) yy += y[j] * y[j];
  for (int k = 0; k < p; ++k) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += U[(size_t)k * n + j] * y[j];
    uu += s * s;
  }
  return yy - uu;
}

// Full FP64 scan over m packed k-mer rows (row_words uint64 words each,
// accession a = bit a%64 of word a/64).  idxs[n] are the phenotyped accession
// positions in accession order, y[n] the aligned phenotype, U[n*p]
// column-major.  For every k-mer the statistic is evaluated (whether or not it
// passes the pre-filter) and pass[i] records the pre-filter decision.
// Returns the number of k-mers passing the pre-filter (the .count value).
int64_t or_scan_f64(const uint64_t* bits, int64_t m, int64_t row_words, const int32_t* idxs,
                    int n, const double* y, const double* U, int p, double corr_thr,
                    int32_t* c_out, uint8_t* pass_out, double* xxM_out, double* xyM_out,
                    double* T_out, double* p_out, double* eff_out) {
  // kgwas.cpp:387-393
  double ymean = 0.0;
  for (int j = 0; j < n; ++j) ymean += y[j];
  ymean /= n;
  double* yc = (double*)malloc(sizeof(double) * (size_t)n);
  double ycn = 0.0;
  for (int j = 0; j < n; ++j) {
    yc[j] = y[j] - ymean;
    ycn += yc[j] * yc[j];
  }
  ycn = sqrt(ycn);
  double* Uy = (double*)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1));
  for (int k = 0; k < p; ++k) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += U[(size_t)k * n + j] * y[j];
    Uy[k] = s;
  }
  const double yyM = or_yyM(y, n, U, p);
  const double eps = 1e-12;
  int64_t passed = 0;
  double* Ur = rows_of(U, n, p);

#pragma omp parallel
  {
    double* x = (double*)malloc(sizeof(double) * (size_t)n);
    double* Ux = (double*)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1));
#pragma omp for schedule(dynamic, 256) reduction(+ : passed)
    for (int64_t i = 0; i < m; ++i) {
      const uint64_t* row = bits + (size_t)i * (size_t)row_words;
      // kgwas.cpp:455-459
      double xs = 0.0;
      int32_t c = 0;
      for (int j = 0; j < n; ++j) {
        const int a = idxs[j];
        x[j] = (double)((row[a >> 6] >> (a & 63)) & 1ULL);
        xs += x[j];
        c += (int32_t)x[j];
      }
      // kgwas.cpp:462-468
      const double xmean = xs / n;
      double xcn = 0.0, dot = 0.0;
      for (int j = 0; j < n; ++j) {
        const double d = x[j] - xmean;
        xcn += d * d;
        dot += d * yc[j];
      }
      xcn = sqrt(xcn);
      int ok = 1;
      if (xcn < eps || ycn < eps) ok = 0;
      else if (fabs(dot / (xcn * ycn)) < corr_thr) ok = 0;
      // kgwas.cpp:472-474
      for (int k = 0; k < p; ++k) Ux[k] = 0.0;
      double xx = 0.0, xy = 0.0;
      for (int j = 0; j < n; ++j) {
        if (x[j] != 0.0) {
          xx += x[j] * x[j];
          xy += x[j] * y[j];
          const double* Uj = Ur + (size_t)j * p;
          for (int k = 0; k < p; ++k) Ux[k] += Uj[k];
        }
      }
      double uu = 0.0, uy = 0.0;
      for (int k = 0; k < p; ++k) {
        uu += Ux[k] * Ux[k];
        uy += Ux[k] * Uy[k];
      }
      const double xxM = xx - uu, xyM = xy - uy;
      double e, t, pv;
      or_compute_hit(xxM, xyM, yyM, n, &e, &t, &pv);
      c_out[i] = c;
      pass_out[i] = (uint8_t)ok;
      xxM_out[i] = xxM;
      xyM_out[i] = xyM;
      T_out[i] = t;
      p_out[i] = pv;
      eff_out[i] = e;
      passed += ok;
    }
    free(x);
    free(Ux);
  }
  free(yc);
  free(Uy);
  free(Ur);
  return passed;
}

// All-pairs epistasis scan in FP64 (src/kgwas_2d.cpp:258-401, emit_pairs_self
// with the caches of build_stats2d_self, src/kgwas_2d.cpp:219-251): for rows
// i < j, i in [i0, i1), x12 = x_i * x_j over the phenotyped accessions
// (src/kgwas_2d.cpp:372); skip unless x_i, x_j and x12 are nonzero
// (:364, :369, :374); t12 = U'x12, xxM = |x12|^2 - |t12|^2, xyM = x12'y -
// t12'U'y (:373-382); skip if xxM or yyM <= 1e-12 (:380); compute_hit and keep
// p <= p_thresh (:383-388).  Rows i are spread over OpenMP threads, so kept
// pairs land in no particular order (the Python wrapper sorts them by (i, j),
// the reference's output order, :415-418); at most `cap` are written and the
// return value is how many pass (a caller can retry with a larger buffer).
int64_t or_pairs_f64(const uint64_t* bits, int64_t N, int64_t row_words, const int32_t* idxs,
                     int n, const double* y, const double* U, int p, int64_t i0, int64_t i1,
                     double p_thresh, int64_t cap, int64_t* ij_out, int32_t* c_out,
                     double* xxM_out, double* xyM_out, double* eff_out, double* T_out,
                     double* p_out) {
  const double eps = 1e-12;
  double* Uy = (double*)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1));
  for (int k = 0; k < p; ++k) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += U[(size_t)k * n + j] * y[j];
    Uy[k] = s;
  }
  const double yyM = or_yyM(y, n, U, p);
  double* Ur = rows_of(U, n, p);
  // x_i over the phenotyped accessions, one byte per sample
  uint8_t* X = (uint8_t*)malloc((size_t)(N > 0 ? N : 1) * (size_t)n);
  for (int64_t r = 0; r < N; ++r)
    for (int j = 0; j < n; ++j) {
      const int a = idxs[j];
      X[(size_t)r * n + j] = (uint8_t)((bits[(size_t)r * row_words + (a >> 6)] >> (a & 63)) & 1ULL);
    }
  int64_t found = 0;
  const int64_t iend = i1 < N ? i1 : N;
#pragma omp parallel
  {
  double* t12 = (double*)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1));
#pragma omp for schedule(dynamic, 1)
  for (int64_t i = i0; i < iend; ++i) {
    const uint8_t* xi = X + (size_t)i * n;
    int ci = 0;
    for (int j = 0; j < n; ++j) ci += xi[j];
    if (ci == 0) continue;
    for (int64_t jr = i + 1; jr < N; ++jr) {
      const uint8_t* xj = X + (size_t)jr * n;
      int cj = 0, c12 = 0;
      double b3 = 0.0;
      for (int k = 0; k < p; ++k) t12[k] = 0.0;
      for (int j = 0; j < n; ++j) {
        cj += xj[j];
        if (xi[j] & xj[j]) {
          ++c12;
          b3 += y[j];
          const double* Uj = Ur + (size_t)j * p;
          for (int k = 0; k < p; ++k) t12[k] += Uj[k];
        }
      }
      if (cj == 0 || c12 == 0) continue;
      double uu = 0.0, uy = 0.0;
      for (int k = 0; k < p; ++k) {
        uu += t12[k] * t12[k];
        uy += t12[k] * Uy[k];
      }
      const double xxM = (double)c12 - uu;
      if (xxM <= eps || yyM <= eps) continue;
      const double xyM = b3 - uy;
      double e, T, pv;
      or_compute_hit(xxM, xyM, yyM, n, &e, &T, &pv);
      if (pv > p_thresh) continue;
      int64_t pos;
#pragma omp atomic capture
      pos = found++;
      if (pos < cap) {
        ij_out[2 * pos] = i;
        ij_out[2 * pos + 1] = jr;
        c_out[pos] = c12;
        xxM_out[pos] = xxM;
        xyM_out[pos] = xyM;
        eff_out[pos] = e;
        T_out[pos] = T;
        p_out[pos] = pv;
      }
    }
  }
  free(t12);
  }
  free(X);
  free(Uy);
  free(Ur);
  return found;
}

// Smallest double T with erfc(sqrt(T/2)) <= p_thresh (bisection on the
// reference's own p-value expression).  Used by tests to state thresholds in
// T space.
double or_T_threshold(double p_thresh) {
  if (p_thresh >= 1.0) return 0.0;
  double lo = 0.0, hi = 1.0;
  while (or_chi2_sf_k1(hi) > p_thresh) hi *= 2.0;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (or_chi2_sf_k1(mid) <= p_thresh) hi = mid;
    else lo = mid;
  }
  return hi;
}

// Row-major float matrix X[n][m] (phenotyped samples x k-mers) from packed
// rows: the layout load_matrix_rows builds (src/kgwas_utils.hpp:225-282).
void or_unpack_rows(const uint64_t* bits, int64_t m, int64_t row_words, const int32_t* idxs,
                    int n, float* X) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j) {
    const int a = idxs[j];
    const uint64_t* col = bits + (a >> 6);
    float* xr = X + (size_t)j * (size_t)m;
    for (int64_t i = 0; i < m; ++i)
      xr[i] = (float)((col[(size_t)i * (size_t)row_words] >> (a & 63)) & 1ULL);
  }
}
