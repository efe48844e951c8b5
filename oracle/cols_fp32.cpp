// CPU restatement of SuperGWAS `kgwas --mode cols` arithmetic (FP32) --
// TEST / BASELINE INFRASTRUCTURE ONLY.  Never linked into the product.
//
// It is the CPU baseline that bench.py times beside the GPU kernels ("port"
// kind) and the reference-FP32 error that the FP64 oracle is compared with.
// It follows, on an in-memory n x N row-major float matrix (the layout
// load_matrix_rows produces, src/kgwas_utils.hpp:225-282):
//   src/kgwas.cpp:183-218  serial FP32 Pearson pre-filter + column compaction
//   src/kgwas.cpp:138-166  compute_cols_stats: FP32 U'y, U'X, X'y, |x|^2,
//                          |UX|^2, UX'Uy; xyM = Xy - UyUX, xxM = xx - sUX;
//                          FP64 yyM
//   src/kgwas.cpp:225-259  emit: FP64 compute_hit per k-mer, OpenMP over
//                          1,000,000-k-mer chunks, "%g" text per survivor
//   src/kgwas_utils.hpp:225-282  load_matrix_rows (plain .txt columns): the
//                          ingest baseline beside the packed-row loader
// The reference multiplies with OpenBLAS sgemm/sgemv (Eigen -DEIGEN_USE_BLAS);
// this file uses a cache-blocked FP32 kernel with runtime ISA dispatch
// (target_clones) in their place, so FP32 rounding matches the reference's
// only up to summation order.
#include <fcntl.h>
#include <math.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

namespace {

double now_s() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

// src/kgwas.cpp:44-56 (FP64)
inline void compute_hit(double xxM, double xyM, double yyM, int n, double* eff, double* T,
                        double* pv) {
  const double eps = 1e-12;
  *eff = 0.0;
  *T = 0.0;
  *pv = 1.0;
  if (xxM > eps && yyM > eps) {
    double r2 = (xyM * xyM) / (xxM * yyM);
    if (r2 < 0) r2 = 0;
    if (r2 >= 1) r2 = 1 - 1e-15;
    *T = (double)n * log(1.0 / (1.0 - r2));
    *pv = erfc(sqrt(0.5 * *T));
    *eff = xyM / (xxM * yyM);
  }
}

constexpr int JB = 64;  // k-mer columns per block

// UX[p][N] = Uf' X, Xy[N] = X' yf, xx[N] = colwise |x|^2 over a row-major X.
__attribute__((target_clones("avx512f", "avx2", "default"))) void block_gemm(
    const float* X, int n, int64_t N, int64_t j0, int jb, const float* Uf, int p,
    const float* yf, float* UXb /* p x JB */, float* Xyb, float* xxb) {
  for (int k = 0; k < p; ++k)
    for (int j = 0; j < JB; ++j) UXb[k * JB + j] = 0.f;
  for (int j = 0; j < JB; ++j) {
    Xyb[j] = 0.f;
    xxb[j] = 0.f;
  }
  float xi[JB];
  for (int i = 0; i < n; ++i) {
    const float* row = X + (size_t)i * (size_t)N + j0;
    for (int j = 0; j < JB; ++j) xi[j] = (j < jb) ? row[j] : 0.f;
    const float yi = yf[i];
    for (int j = 0; j < JB; ++j) {
      Xyb[j] += xi[j] * yi;
      xxb[j] += xi[j] * xi[j];
    }
    for (int k = 0; k < p; ++k) {
      const float u = Uf[(size_t)k * n + i];
      float* acc = UXb + k * JB;
      for (int j = 0; j < JB; ++j) acc[j] += u * xi[j];
    }
  }
}

}  // namespace

extern "C" {

// Runs pre-filter + LA + emit on X (n x N, row-major, 0/1 floats).
// Outputs: *n_hits_1d / *n_hits_2d, stage seconds t_stage[0..2]
// (pre-filter, LA, emit); optional per-kept-column xxM/xyM (FP32) and the
// kept original column indices keep_out (size >= N).  Returns the number of
// columns passing the pre-filter.
int64_t cpu_cols_fp32(const float* X_in, int n, int64_t N, const double* y, const double* U,
                      int p, double corr_thr, double log10_thresh, double log10_thresh_2d,
                      int threads, int64_t* n_hits_1d, int64_t* n_hits_2d, double* t_stage,
                      float* xxM_out, float* xyM_out, int64_t* keep_out,
                      int64_t* text_bytes) {
  if (threads < 1) threads = 1;
  // ---- pre-filter (src/kgwas.cpp:183-218), serial FP32 ----
  double t0 = now_s();
  std::vector<float> yf(n);
  for (int i = 0; i < n; ++i) yf[i] = (float)y[i];
  float ysum = 0.f;
  for (int i = 0; i < n; ++i) ysum += yf[i];
  const float ymean = ysum / (float)n;
  std::vector<float> yc(n);
  float yn2 = 0.f;
  for (int i = 0; i < n; ++i) {
    yc[i] = yf[i] - ymean;
    yn2 += yc[i] * yc[i];
  }
  const float ynorm = sqrtf(yn2);
  const float thresh = (float)corr_thr;
  std::vector<int64_t> keep;
  keep.reserve((size_t)N);
  for (int64_t j = 0; j < N; ++j) {
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += X_in[(size_t)i * (size_t)N + j];
    const float mean = s / (float)n;
    float xn2 = 0.f, dot = 0.f;
    for (int i = 0; i < n; ++i) {
      const float d = X_in[(size_t)i * (size_t)N + j] - mean;
      xn2 += d * d;
      dot += d * yc[i];
    }
    const float xnorm = sqrtf(xn2);
    if (xnorm < 1e-12f || ynorm < 1e-12f) continue;
    const float cor = dot / (xnorm * ynorm);
    if (fabsf(cor) >= thresh) keep.push_back(j);
  }
  const int64_t passed = (int64_t)keep.size();
  std::vector<float> Xf;
  const float* X = X_in;
  int64_t Nk = N;
  if (passed < N) {  // src/kgwas.cpp:207-216
    Xf.resize((size_t)n * (size_t)passed);
    for (int64_t c = 0; c < passed; ++c)
      for (int i = 0; i < n; ++i)
        Xf[(size_t)i * (size_t)passed + c] = X_in[(size_t)i * (size_t)N + keep[c]];
    X = Xf.data();
    Nk = passed;
  }
  double t1 = now_s();

  // ---- compute_cols_stats (src/kgwas.cpp:138-166), FP32 ----
  std::vector<float> Uf((size_t)n * p);
  for (size_t i = 0; i < Uf.size(); ++i) Uf[i] = (float)U[i];
  std::vector<float> Uyf(p);
  for (int k = 0; k < p; ++k) {
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += Uf[(size_t)k * n + i] * yf[i];
    Uyf[k] = s;
  }
  std::vector<float> xxM((size_t)Nk), xyM((size_t)Nk);
  omp_set_num_threads(threads);
#pragma omp parallel
  {
    std::vector<float> UXb((size_t)p * JB), Xyb(JB), xxb(JB);
#pragma omp for schedule(dynamic, 4)
    for (int64_t j0 = 0; j0 < Nk; j0 += JB) {
      const int jb = (int)((Nk - j0) < JB ? (Nk - j0) : JB);
      block_gemm(X, n, Nk, j0, jb, Uf.data(), p, yf.data(), UXb.data(), Xyb.data(),
                 xxb.data());
      for (int j = 0; j < jb; ++j) {
        float sux = 0.f, uyux = 0.f;
        for (int k = 0; k < p; ++k) {
          const float v = UXb[(size_t)k * JB + j];
          sux += v * v;
          uyux += v * Uyf[k];
        }
        xyM[(size_t)(j0 + j)] = Xyb[j] - uyux;
        xxM[(size_t)(j0 + j)] = xxb[j] - sux;
      }
    }
  }
  double yy = 0.0, uu = 0.0;  // FP64 yyM (src/kgwas.cpp:160-161)
  for (int i = 0; i < n; ++i) yy += y[i] * y[i];
  for (int k = 0; k < p; ++k) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += U[(size_t)k * n + i] * y[i];
    uu += s * s;
  }
  const double yyM = yy - uu;
  double t2 = now_s();

  // ---- emit (src/kgwas.cpp:225-259) ----
  const double p_thresh = pow(10.0, -log10_thresh);
  const bool do_2d = log10_thresh_2d > 0.0;
  const double p_thresh_2d = do_2d ? pow(10.0, -log10_thresh_2d) : 0.0;
  const int64_t CH = 1000000;
  const int64_t nch = (Nk + CH - 1) / CH;
  std::vector<std::string> buf1((size_t)nch), buf2((size_t)nch);
  int64_t h1 = 0, h2 = 0;
#pragma omp parallel for schedule(static) reduction(+ : h1, h2)
  for (int64_t ci = 0; ci < nch; ++ci) {
    const int64_t s = ci * CH, e = (s + CH < Nk) ? s + CH : Nk;
    std::string o1, o2;
    char line[256];
    for (int64_t b = s; b < e; ++b) {
      double eff, T, pv;
      compute_hit((double)xxM[(size_t)b], (double)xyM[(size_t)b], yyM, n, &eff, &T, &pv);
      if (pv <= p_thresh) {
        const int64_t orig = (passed < N) ? keep[(size_t)b] : b;
        int len = snprintf(line, sizeof(line), "k%lld\t%g\t%g\n", (long long)orig, eff, pv);
        o1.append(line, (size_t)len);
        ++h1;
        if (do_2d && pv <= p_thresh_2d) {
          std::string bits((size_t)n, '0');  // reconstruct_bitstring (src/kgwas.cpp:59-64)
          for (int i = 0; i < n; ++i)
            bits[(size_t)i] = (X[(size_t)i * (size_t)Nk + b] > 0.5f) ? '1' : '0';
          o2 += "k" + std::to_string(orig) + "\t" + bits + "\t";
          len = snprintf(line, sizeof(line), "%g\t%g\n", eff, pv);
          o2.append(line, (size_t)len);
          ++h2;
        }
      }
    }
    buf1[(size_t)ci] = std::move(o1);
    buf2[(size_t)ci] = std::move(o2);
  }
  int64_t bytes = 0;
  for (auto& s : buf1) bytes += (int64_t)s.size();
  for (auto& s : buf2) bytes += (int64_t)s.size();
  double t3 = now_s();

  if (n_hits_1d) *n_hits_1d = h1;
  if (n_hits_2d) *n_hits_2d = h2;
  if (text_bytes) *text_bytes = bytes;
  if (t_stage) {
    t_stage[0] = t1 - t0;
    t_stage[1] = t2 - t1;
    t_stage[2] = t3 - t2;
  }
  for (int64_t b = 0; b < Nk; ++b) {
    if (xxM_out) xxM_out[b] = xxM[(size_t)b];
    if (xyM_out) xyM_out[b] = xyM[(size_t)b];
    if (keep_out) keep_out[b] = (passed < N) ? keep[(size_t)b] : b;
  }
  return passed;
}

// load_matrix_rows restated for plain .txt columns (src/kgwas_utils.hpp:225-282):
// sample k's column bits_dir/col_%06d.txt (accession idxs[k]) is mapped and
// converted to X[k][i] = float(byte_i & 1), OpenMP over samples.  Returns 0, or
// -1 - k when sample k's column is missing or shorter than N.
int cpu_load_columns(const char* bits_dir, const int32_t* idxs, int n, int64_t N, float* X,
                     int threads) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (int k = 0; k < n; ++k) {
    char base[64];
    snprintf(base, sizeof(base), "/col_%06d.txt", idxs[k]);
    const std::string path = std::string(bits_dir) + base;
    const int fd = ::open(path.c_str(), O_RDONLY);
    struct stat st {};
    if (fd < 0 || fstat(fd, &st) != 0 || (int64_t)st.st_size < N) {
      if (fd >= 0) ::close(fd);
#pragma omp critical
      bad = bad ? bad : -1 - k;
      continue;
    }
    void* p = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (p == MAP_FAILED) {
#pragma omp critical
      bad = bad ? bad : -1 - k;
      continue;
    }
    madvise(p, (size_t)st.st_size, MADV_SEQUENTIAL);
    const unsigned char* bytes = (const unsigned char*)p;
    float* row = X + (size_t)k * (size_t)N;
#pragma omp simd
    for (int64_t i = 0; i < N; ++i) row[i] = (float)(bytes[i] & 1u);
    munmap(p, (size_t)st.st_size);
  }
  return bad;
}

}  // extern "C"
