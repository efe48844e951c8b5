// C ABI of libkgwas_hip: null-model set-up in FP64, weight quantisation and
// layout, scan orchestration on one HIP stream (optionally replayed as a HIP
// graph), and the final per-k-mer decisions with the reference's own
// compute_hit arithmetic.
//
// Reference seam replaced (src/kgwas.cpp):
//   183-218  Pearson pre-filter      -> bulk kernel bound test + FP64 host test
//   138-166  compute_cols_stats      -> scan_i8 / scan_f32 + refine_f64
//   44-56    compute_hit             -> host_hit() below (glibc log/erfc)
//   239-259  threshold + chunk order -> decide() + sort by k-mer index
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/kgwas_hip.h"
#include "kg_internal.h"

namespace kg {
hipError_t launch_scan_i8(int NT, int L, int t, int R_req, int grid, hipStream_t s,
                          const ShardDesc* sd, int nshards, int64_t total_blocks,
                          const ScanOut& o, const int64_t* total_dev);
int64_t screen_block_kmers(int t, bool w8);
int screen_rows(int t);
bool screen_count_row(int t);
int screen_max_chunks();
int screen_wgs_per_cu(int t, int nch, bool dd, bool w8);
bool screen_ring_fits(int nch, int W);
hipError_t launch_screen_ring(int W, int nch_max, int grid, hipStream_t s, const ScreenDesc* sd,
                              int nshards, int64_t total_blocks, bool dd);
hipError_t launch_screen(int t, int nch_max, int grid, hipStream_t s, const ScreenDesc* sd,
                         int nshards, int64_t total_blocks, bool dd, bool w8);
int screen_chunk_words();
hipError_t launch_screen_compact(hipStream_t s, const ScrChunk* ch, int ncc, const ScreenDesc* ss,
                                 uint32_t* ccount, int64_t* P, ShardDesc* sd2, int nsh,
                                 int64_t bk, int64_t* total);
int64_t i8_block_kmers(int NT, int L, int t, int R_req);
int i8_tiles_for(int NT, int L, int t, int R_req);
int i8_column_capacity(int NT, int L);
hipError_t launch_scan_i8_res(int NT, int nch_max, int grid, hipStream_t s, const ShardDesc* sd,
                              int nshards, int64_t total_blocks, const ScanOut& o,
                              const int64_t* total_dev);
bool i8_res_fits(int NT, int nch);
int64_t i8_res_block_kmers();
hipError_t launch_scan_f6_res(int NT, int D, int cm, int nch_max, int grid, hipStream_t s,
                              const ShardDesc* sd, int nshards, int64_t total_blocks,
                              const ScanOut& o, const int64_t* total_dev);
bool f6_res_fits(int NT, int nch);
hipError_t launch_rchk_tables(hipStream_t s, ShardDesc* sd3, int nsh, int64_t bk, int64_t* total);
hipError_t set_kernel_attributes();
hipError_t launch_scan_f32(int NT, int nch, int grid, hipStream_t s, const uint64_t* bits,
                           int64_t m, int64_t rw, int64_t row0, const float* wts,
                           const uint64_t* mask, const EpiF32& ep, const ScanOut& o);
hipError_t launch_refine(hipStream_t s, const Cand* cand, const unsigned long long* count,
                         int64_t cap, const RefDesc* rd, double* res, int blocks,
                         const RefHost& hb);
hipError_t launch_zero(hipStream_t s, unsigned long long* p, int n);
hipError_t noalloc_scratch(int device, void** ptr);
size_t setup_max_bytes();
hipError_t launch_setup(hipStream_t s, const void* src, size_t bytes, void* dst,
                        unsigned long long* zero, int nzero);
hipError_t launch_pair_masks(hipStream_t s, const uint64_t* bits, int64_t rw,
                            const uint64_t* mask, int64_t W, int64_t i0, int64_t cnt,
                            uint64_t* out);
hipError_t launch_synth(hipStream_t s, uint64_t* out, uint64_t seed, int64_t kmer0, int64_t m,
                        uint32_t n_acc, int64_t rw, const uint8_t* pop, const uint32_t* thr,
                        uint32_t npop, uint32_t nprof);
}  // namespace kg

using namespace kg;

namespace {

// Device buffer that frees itself.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, b ? b : 16);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

// Page-locked host buffer (bounce buffer for pageable rows; with
// hipHostMallocCoherent, a buffer kernels write into directly through `dev`).
struct HBuf {
  void* p = nullptr;
  void* dev = nullptr;  // device-visible address of p
  size_t bytes = 0;
  unsigned flags = hipHostMallocDefault;  // of the current allocation
  HBuf() = default;
  HBuf(const HBuf&) = delete;
  HBuf& operator=(const HBuf&) = delete;
  ~HBuf() {
    if (p) (void)hipHostFree(p);
  }
  hipError_t ensure(size_t b, unsigned want = hipHostMallocDefault) {
    if (b <= bytes && p && want == flags) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = dev = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&p, b ? b : 16, want);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&dev, p, 0);
    if (e == hipSuccess) {
      bytes = b;
      flags = want;
    } else if (p) {
      (void)hipHostFree(p);
      p = dev = nullptr;
    }
    return e;
  }
  template <class T>
  T* dev_as() const {
    return reinterpret_cast<T*>(dev);
  }
};

// src/kgwas.cpp:44-56
void host_hit(double xxM, double xyM, double yyM, int n, double* eff, double* T, double* pv) {
  const double eps = 1e-12;
  *eff = 0.0;
  *T = 0.0;
  *pv = 1.0;
  if (xxM > eps && yyM > eps) {
    double r2 = (xyM * xyM) / (xxM * yyM);
    if (r2 < 0) r2 = 0;
    if (r2 >= 1) r2 = 1 - 1e-15;
    *T = (double)n * std::log(1.0 / (1.0 - r2));
    *pv = std::erfc(std::sqrt(0.5 * *T));
    *eff = xyM / (xxM * yyM);
  }
}

// Smallest double T with erfc(sqrt(T/2)) <= p_thr (the reference's p-value is
// non-increasing in T, so the report set is T >= T*).
double t_star(double p_thr) {
  if (p_thr >= 1.0) return 0.0;
  auto sf = [](double T) { return T <= 0.0 ? 1.0 : std::erfc(std::sqrt(0.5 * T)); };
  double lo = 0.0, hi = 1.0;
  while (sf(hi) > p_thr && hi < 1e300) hi *= 2.0;
  for (int it = 0; it < 2000; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (!(mid > lo && mid < hi)) break;
    if (sf(mid) <= p_thr) hi = mid;
    else lo = mid;
  }
  return hi;
}

struct Model {
  int n_acc = 0, n = 0, p = 0, t = 0;
  std::vector<int32_t> idxs;
  std::vector<double> U, Y;      // column-major
  // FP64 derived quantities
  double sigma = 0.0, base_coef = 0.0, dmax = 0.0;
  bool drop_d = false, rotated = false;
  std::vector<double> Uc;        // rotated, column-centred U  (n x p)
  std::vector<double> bprime;    // [t][p]
  std::vector<double> rc;        // centred residuals (n x t)
  std::vector<double> rbar, ycn, ymean, yyM;
  std::vector<uint64_t> maskw;   // accession mask, 8 * nchunks_i8 words
  bool mask_full = false;        // every accession of every 512-sample chunk is in the model
  int nch_i8 = 0, nch_f32 = 0;
};

// Full-statistic mode: a k-mer/trait whose bulk values were recomputed in FP64.
struct StatPatch {
  int64_t idx;
  int trait, c;
  double xxM, xyM, S;
};

// A null model with its device-side weights and constants.
// An int8-limb weight set of a model at some limb count: chunk-major MFMA
// operands, column and trait records, epilogue constants.
struct I8Set {
  int NT = 0;
  bool cc = false;  // FP6 digits: the last column slot is the count column
  DBuf wts, colrec, trc;
  EpiI8 epi{};
};

struct ModelDev {
  Model md;
  int NT = 0, nz = 0, nb = 0;
  DBuf wts, mask, rbar, B64, bpfull, colrec, trc;  // int8 path
  // two-limb set for the resident-weight kernel (single trait, when every
  // chunk's weights fit its LDS)
  I8Set res;
  bool res_ok = false;
  // FP6-digit set for the resident kernel (f6_digits digits per column; 0:
  // not built)
  I8Set f6;
  bool f6_ok = false;
  int f6_digits = 0;
  DBuf wf32, uyf;                                                     // FP32 path
  // pre-filter screen: int8 limbs of y - ybar (three for one trait, two per
  // trait otherwise; absent when the model is too wide for the screen's LDS
  // weight table)
  DBuf swts;
  bool screen_ok = false;
  double sw[kMaxTraits] = {}, sh[kMaxTraits] = {};
  EpiI8 epi{};
  EpiF32 epf{};
};

}  // namespace

struct kg_ctx {
  int device = 0;
  int precision = KG_PREC_MIXED;
  int limbs = 3;
  int cus = 256;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // kg_scan streaming: copy stream, copy-done events and copy timing events
  // per staging buffer, page-locked bounce buffers for pageable rows
  hipStream_t cstream = nullptr;
  // descriptor tables of a non-graph int8 scan: page-locked staging and the
  // device copy (one asynchronous upload per scan)
  HBuf hdesc;
  DBuf ddesc;
  // page-locked read-back, written by refine_f64: the counters, and the first
  // kEagerCands candidates with their FP64 results
  HBuf hcnt, hcand, hres;
  hipEvent_t cev[2] = {nullptr, nullptr};
  hipEvent_t ctev[4] = {nullptr, nullptr, nullptr, nullptr};
  HBuf bounce[2];
  std::string err;
  std::vector<std::unique_ptr<ModelDev>> models;
  // scratch
  DBuf counters, cand, res, stage, stage2, st_c, st_xx, st_xy, st_S, sdesc, sdesc1, rdesc;
  int64_t cand_cap = 0;
  size_t counters_len = 0;
  kg_timing timing{};
  int grid_mult = 1;  // int8 scan: persistent workgroups per CU (KG_GRID_MULT)
  int tiles = 0;      // int8 scan: k-mer tiles per wave (0 = the kernel's default)
  int grid_mult_f32 = 8;  // FP32 scan: workgroups per CU
  bool scr_w8 = false;    // screen: the 8-wave workgroup shape (kg_autotune may pick it)
  // one-trait screen of models of 4 or 8 chunks: rows streamed through an LDS
  // ring by workgroups of this many waves (8 or 12; 0: the register-loaded screen)
  int scr_ring = 8;
  bool resident = true;   // single-trait screened scans: the resident-weight kernel
  // its operands: 0 = two int8 limbs; 2 or 3 = FP6 digits per column against
  // FP4 bits (2: with the three-limb recheck pass of the k-mers it leaves open)
  int res_digits = 2;
  void* scratch = nullptr;  // device address of this GPU's allocation-free scratch array
  // pre-filter screen (int8 path): mode, and in auto mode the unscreened
  // scans left before the screen is tried again
  int screen_mode = KG_SCREEN_AUTO;
  int screen_skip = 0;
  // auto mode: device time of the last unscreened scan that could have been
  // screened, and its k-mer count; scans of up to kProbeKmers k-mers measure
  // it now and then (fixed costs decide there, so scaling the survivor scan
  // to all k-mers misjudges them)
  double full_ms = 0.0;
  int64_t full_m = 0;
  int probe_tick = 0;
  int dedup = 0;  // screened scans check each row against the previous one
  std::vector<StatPatch> stat_patch;  // FP64 values of the last full-statistic scan
  DBuf surv, ssdesc, sdesc2, chdesc;  // survivor rows; graph-path screen tables
  DBuf rchk, sdesc3;                  // recheck rows of the two-digit kernel; its graph-path table
  DBuf bmap, scr_aux;                 // screen bitmaps; chunk counts and prefixes
  DBuf pmask;                         // all-pairs scan: mask AND row i, per shard
  DBuf dmap, smult;                   // duplicate-row scans: dup bitmaps, survivor run lengths
  hipEvent_t evs = nullptr;   // end of the screen kernel
  int scr_occ_key = -1, scr_occ = 1;  // cached screen workgroups per CU
  // captured graph of the last batched scan
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<char> graph_key;
};

#define KG_CK(ctx, expr)                                                            \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
      return KG_ERR_DEVICE;                                                         \
    }                                                                               \
  } while (0)

static int fail(kg_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

// ---------------------------------------------------------------------------
// FP64 null-model algebra (all on the host; n x p is at most a few MB).
//
// With m = U'1, a Householder Q maps m to sigma e_0.  U' = U Q spans the same
// column space and U'U'' = UU', so x'x - |U'x|^2 and x'y - (U'x)'(U'y) are
// unchanged.  Writing U'_c = U' - 1 m''/n (column-centred):
//   U''x = U'_c'x + (c/n) sigma e_0
//   xxM  = c - c^2 sigma^2/n^2 - 2 (c/n) sigma z_0 - |z|^2,   z = U'_c'x
//   xyM  = x'r = x'r_c + c mean(r),                            r = y - U U'y
//   S    = x'(y - mean y) = x'r_c + z'b',                      b' = Q'U'y
// The device then only contracts centred columns, and when 1 is in span(U)
// (the intercept of the reference's U.bin) column z_0 is identically zero.
static void prepare_model(Model& md) {
  const int n = md.n, p = md.p, t = md.t;
  std::vector<double> mvec(p, 0.0);
  for (int k = 0; k < p; ++k) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += md.U[(size_t)k * n + j];
    mvec[k] = s;
  }
  double mn2 = 0.0;
  for (double v : mvec) mn2 += v * v;
  const double mn = std::sqrt(mn2);
  std::vector<double> Q((size_t)p * p, 0.0);  // column-major p x p
  for (int k = 0; k < p; ++k) Q[(size_t)k * p + k] = 1.0;
  md.sigma = 0.0;
  md.rotated = false;
  if (p > 0 && mn > 0.0) {
    // Householder: H m = sigma e_0, sigma = -sign(m_0) |m|
    const double sg = (mvec[0] >= 0.0) ? -mn : mn;
    std::vector<double> v(mvec);
    v[0] -= sg;
    double vv = 0.0;
    for (double x : v) vv += x * x;
    if (vv > 0.0) {
      for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b) Q[(size_t)b * p + a] -= 2.0 * v[a] * v[b] / vv;
    }
    md.sigma = sg;
    md.rotated = true;
  }
  md.base_coef = md.sigma * md.sigma / ((double)n * (double)n);
  // U' = U Q, then centre columns
  md.Uc.assign((size_t)n * p, 0.0);
  for (int b = 0; b < p; ++b)
    for (int a = 0; a < p; ++a) {
      const double q = Q[(size_t)b * p + a];
      if (q == 0.0) continue;
      const double* ua = &md.U[(size_t)a * n];
      double* ub = &md.Uc[(size_t)b * n];
      for (int j = 0; j < n; ++j) ub[j] += ua[j] * q;
    }
  for (int b = 0; b < p; ++b) {
    const double mb = (b == 0) ? md.sigma : 0.0;  // m' = sigma e_0
    double* ub = &md.Uc[(size_t)b * n];
    if (md.rotated) {
      for (int j = 0; j < n; ++j) ub[j] -= mb / n;
    }
  }
  // d column (index 0 after rotation): dropped when it is numerically zero
  md.dmax = 0.0;
  md.drop_d = false;
  if (md.rotated) {
    for (int j = 0; j < n; ++j) md.dmax = std::max(md.dmax, std::fabs(md.Uc[j]));
    md.drop_d = md.dmax <= 1e-12;
  }
  // traits
  md.bprime.assign((size_t)t * p, 0.0);
  md.rc.assign((size_t)n * t, 0.0);
  md.rbar.assign(t, 0.0);
  md.ycn.assign(t, 0.0);
  md.ymean.assign(t, 0.0);
  md.yyM.assign(t, 0.0);
  for (int tr = 0; tr < t; ++tr) {
    const double* y = &md.Y[(size_t)tr * n];
    std::vector<double> b(p, 0.0);
    double yy = 0.0, bb = 0.0;
    for (int j = 0; j < n; ++j) yy += y[j] * y[j];
    for (int k = 0; k < p; ++k) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += md.U[(size_t)k * n + j] * y[j];
      b[k] = s;
      bb += s * s;
    }
    md.yyM[tr] = yy - bb;  // src/kgwas.cpp:160-161
    for (int k2 = 0; k2 < p; ++k2) {
      double s = 0.0;
      for (int k = 0; k < p; ++k) s += Q[(size_t)k2 * p + k] * b[k];
      md.bprime[(size_t)tr * p + k2] = s;
    }
    std::vector<double> r(y, y + n);
    for (int k = 0; k < p; ++k) {
      const double* u = &md.U[(size_t)k * n];
      for (int j = 0; j < n; ++j) r[j] -= u[j] * b[k];
    }
    double rs = 0.0;
    for (int j = 0; j < n; ++j) rs += r[j];
    const double rb = rs / n;
    md.rbar[tr] = rb;
    for (int j = 0; j < n; ++j) md.rc[(size_t)tr * n + j] = r[j] - rb;
    double ys = 0.0;
    for (int j = 0; j < n; ++j) ys += y[j];
    const double ym = ys / n;  // src/kgwas.cpp:387-388
    double yc2 = 0.0;
    for (int j = 0; j < n; ++j) yc2 += (y[j] - ym) * (y[j] - ym);
    md.ymean[tr] = ym;
    md.ycn[tr] = std::sqrt(yc2);
  }
}

// Quantise columns to `L` balanced int8 limbs: v = scale * sum_l limb_l / (127 * 254^l).
// *h bounds |e_j| (e_j = col_j - v_j); with E != nullptr, *E = sum_j e_j and
// *hc bounds |e_j - E / n| (the mean-corrected remainder), both with the FP64
// rounding of the remainders and of their sum allowed for.
static void quantize(const double* col, int n, int L, std::vector<int8_t>& limbs, double* w,
                     double* h, double* E = nullptr, double* hc = nullptr) {
  double scale = 0.0;
  for (int j = 0; j < n; ++j) scale = std::max(scale, std::fabs(col[j]));
  limbs.assign((size_t)n * L, 0);
  if (E) *E = 0.0;
  if (hc) *hc = 0.0;
  if (scale == 0.0) {
    for (int l = 0; l < L; ++l) w[l] = 0.0;
    *h = 0.0;
    return;
  }
  double mul = 127.0;
  for (int l = 0; l < L; ++l) {
    w[l] = scale / mul;
    mul *= 254.0;
  }
  for (int j = 0; j < n; ++j) {
    double res = col[j] / scale;
    double m = 127.0;
    for (int l = 0; l < L; ++l) {
      double q = std::nearbyint(res * m);
      if (q > 127.0) q = 127.0;
      if (q < -127.0) q = -127.0;
      limbs[(size_t)j * L + l] = (int8_t)q;
      res -= q / m;
      m *= 254.0;
    }
  }
  // exact representable remainder bound: half of the last limb's step, plus
  // FP64 rounding of the residual recursion
  *h = 0.5 * w[L - 1] + 1e-15 * scale;
  if (E) {
    std::vector<double> e(n);
    double sum = 0.0;
    for (int j = 0; j < n; ++j) {
      double v = 0.0;
      for (int l = L - 1; l >= 0; --l) v += w[l] * (double)limbs[(size_t)j * L + l];
      e[j] = col[j] - v;
      sum += e[j];
    }
    const double mean = sum / n;
    double mx = 0.0;
    for (int j = 0; j < n; ++j) mx = std::max(mx, std::fabs(e[j] - mean));
    *E = sum;
    // e_j carries FP64 rounding of v and of the subtraction (< 1e-15 scale),
    // the sum n such roundings of terms below *h
    *hc = mx + 2e-15 * scale + 4e-16 * (double)n * (*h);
  }
}

// FP6 e2m3 code of the digit d / 8 (|d| <= 15): sign bit 5, then |d| itself
// (0-7: exponent 0, mantissa |d|; 8-15: exponent 1, mantissa |d| - 8).
static inline uint32_t f6_code(int d) { return (d < 0 ? 32u : 0u) | (uint32_t)std::abs(d); }

// Columns as D base-30 digits for the FP6 resident kernel: v = s * sum_l
// d_l 30^(D-1-l) with |d_l| <= 15, s_0 = max|v| / 15, s_{l+1} = s_l / 30 and
// s = s_{D-1} (the remainder after digit l is at most s_l / 2 = 15 s_{l+1}, so
// no later digit clamps).  w[0] = 8 s, since the kernel's digit sums count
// d / 8; *h, *E and *hc as for quantize().
static void quantize_digits(const double* col, int n, int D, std::vector<int8_t>& dig, double* w,
                            double* h, double* E, double* hc) {
  double scale = 0.0;
  for (int j = 0; j < n; ++j) scale = std::max(scale, std::fabs(col[j]));
  dig.assign((size_t)n * D, 0);
  for (int l = 0; l < D; ++l) w[l] = 0.0;
  *E = 0.0;
  *hc = 0.0;
  if (scale == 0.0) {
    *h = 0.0;
    return;
  }
  double sl[3];
  sl[0] = scale / 15.0;
  for (int l = 1; l < D; ++l) sl[l] = sl[l - 1] / 30.0;
  const double s = sl[D - 1];
  std::vector<double> e(n);
  double sum = 0.0;
  for (int j = 0; j < n; ++j) {
    double res = col[j];
    double iv = 0.0;
    for (int l = 0; l < D; ++l) {
      double d = std::nearbyint(res / sl[l]);
      d = std::min(15.0, std::max(-15.0, d));
      dig[(size_t)j * D + l] = (int8_t)d;
      res -= d * sl[l];
      iv = 30.0 * iv + d;
    }
    e[j] = col[j] - s * iv;
    sum += e[j];
  }
  w[0] = 8.0 * s;
  *h = 0.5 * s + 1e-15 * scale;
  const double mean = sum / n;
  double mx = 0.0;
  for (int j = 0; j < n; ++j) mx = std::max(mx, std::fabs(e[j] - mean));
  *E = sum;
  *hc = mx + 2e-15 * scale + 4e-16 * (double)n * (*h);
}

constexpr int kNoFit = -100;  // a secondary weight set that does not fit its kernel

// Weight set at L limbs.  primary: the model's main set (also uploads the mask,
// the FP64 refinement tables and b'; a model too wide for it is an error);
// otherwise kNoFit when the columns need more tiles than L allows.  f6: L FP6
// digits per column in the resident kernel's operand layout instead.
static int build_i8_set(kg_ctx* ctx, ModelDev& M, int L, bool primary, int& NT_out,
                        DBuf& wts_b, DBuf& colrec_b, DBuf& trc_b, EpiI8& ep, bool f6 = false,
                        bool* cc_out = nullptr) {
  Model& md = M.md;
  const int n = md.n, p = md.p, t = md.t;
  // 32385 n < 2^31 keeps every limb sum (|sum| <= 127 n) and the epilogue's
  // int32 arithmetic exact; n is far below this in practice
  if (n > 60000)
    return fail(ctx, KG_ERR_ARG, "more than 60000 phenotyped samples for the int8 kernel");
  std::vector<int> zc;
  for (int k = 0; k < p; ++k)
    if (!(k == 0 && md.rotated && md.drop_d)) zc.push_back(k);
  const int nz = (int)zc.size();
  const int nb = nz + t;  // base columns: t residuals, then nz structure columns
  const int nt_max = f6 ? 6 : ((L == 3) ? 16 : (L == 2 ? 6 : 3));
  int NT = 1;
  while (NT <= nt_max && i8_column_capacity(NT, L) < nb) ++NT;
  if (NT > nt_max) {
    if (!primary) return kNoFit;
    return fail(ctx, KG_ERR_ARG,
                "p + t exceeds the int8 kernel's weight columns for this limb count (at most " +
                    std::to_string(i8_column_capacity(nt_max, L)) + " structure + trait columns)");
  }
  const int NU = 16 * NT;
  const int ncol = i8_column_capacity(NT, L);
  NT_out = NT;
  // FP6 digits: a spare column slot becomes the count column (the last slot)
  const bool cc = f6 && ncol > nb;
  if (cc_out) *cc_out = cc;
  if (primary) {
    M.nz = nz;
    M.nb = nb;
  }
  // base column b -> source column, limb l -> MFMA output row u
  auto src_col = [&](int b) -> const double* {
    return (b < t) ? &md.rc[(size_t)b * n] : &md.Uc[(size_t)zc[b - t] * n];
  };
  auto urow = [&](int b, int l) {
    const int sl = L * (b >> 2) + l;
    return 16 * (sl >> 2) + 4 * (b & 3) + (sl & 3);
  };
  std::vector<int8_t> lim;
  std::vector<double> colw((size_t)nb * L), colh(nb), colE(nb);
  std::vector<int8_t> ucol((size_t)n * NU, 0);  // [sample][u-row]
  for (int b = 0; b < nb; ++b) {
    double h0 = 0.0;
    if (f6)
      quantize_digits(src_col(b), n, L, lim, &colw[(size_t)b * L], &h0, &colE[b], &colh[b]);
    else
      quantize(src_col(b), n, L, lim, &colw[(size_t)b * L], &h0, &colE[b], &colh[b]);
    for (int l = 0; l < L; ++l) {
      const int u = urow(b, l);
      for (int j = 0; j < n; ++j) ucol[(size_t)j * NU + u] = lim[(size_t)j * L + l];
    }
  }
  // count column: digit 8 (the FP6 value 1) for every sample, so its digit-0
  // sum is the allele count; its column record stays zero
  if (cc)
    for (int j = 0; j < n; ++j) ucol[(size_t)j * NU + urow(ncol - 1, 0)] = 8;
  // accession -> sample
  std::vector<int> samp(md.n_acc, -1);
  for (int j = 0; j < n; ++j) samp[md.idxs[j]] = j;
  const int nch = md.nch_i8;
  const size_t frag = (size_t)8 * NT * 64 * 16;  // bytes per chunk
  std::vector<int8_t> W(f6 ? 0 : (size_t)nch * frag, 0);
  if (f6) {
    // [chunk][shift k][tile][lane] operands of 32 six-bit digit codes: the
    // first 16 bytes of every operand, then the last 8 (kF6OperandBytes);
    // element j of lane group g is accession 512 C + 128 g + 32 (j / 8) +
    // 4 (j % 8) + k, at bit 6 j
    const size_t nop = (size_t)nch * 4 * NT * 64;
    W.assign(nop * kF6OperandBytes, 0);
    uint32_t* X = reinterpret_cast<uint32_t*>(W.data());
    uint32_t* Y = X + nop * 4;
    for (int C = 0; C < nch; ++C)
      for (int k = 0; k < 4; ++k)
        for (int tt = 0; tt < NT; ++tt)
          for (int lane = 0; lane < 64; ++lane) {
            const int g = lane >> 4, uc = 16 * tt + (lane & 15);
            uint32_t bits[6] = {0, 0, 0, 0, 0, 0};
            for (int j = 0; j < 32; ++j) {
              const int a = 512 * C + 128 * g + 32 * (j >> 3) + 4 * (j & 7) + k;
              if (a >= md.n_acc || samp[a] < 0) continue;
              const uint64_t code = f6_code(ucol[(size_t)samp[a] * NU + uc]);
              const int b0 = 6 * j;
              bits[b0 >> 5] |= (uint32_t)(code << (b0 & 31));
              if ((b0 & 31) > 26) bits[(b0 >> 5) + 1] |= (uint32_t)(code >> (32 - (b0 & 31)));
            }
            const size_t q = (((size_t)C * 4 + k) * NT + tt) * 64 + lane;
            for (int i = 0; i < 4; ++i) X[q * 4 + i] = bits[i];
            Y[q * 2] = bits[4];
            Y[q * 2 + 1] = bits[5];
          }
  }
  for (int C = 0; C < (f6 ? 0 : nch); ++C)
    for (int jj = 0; jj < 8; ++jj)
      for (int tt = 0; tt < NT; ++tt)
        for (int lane = 0; lane < 64; ++lane) {
          const int i = lane & 15, g = lane >> 4;
          const int uc = 16 * tt + i;
          int8_t* dst = &W[(size_t)C * frag + (((size_t)jj * NT + tt) * 64 + lane) * 16];
          for (int e = 0; e < 16; ++e) {
            // operand byte e of step jj: bit 8 (e % 4) + 4 (jj % 2) + e / 4 of
            // the lane group's word jj / 2 of the chunk (kg_kernels.hip)
            const int a = 512 * C + 128 * g + 32 * (jj >> 1) + 8 * (e & 3) + 4 * (jj & 1) + (e >> 2);
            if (a < md.n_acc && samp[a] >= 0) dst[e] = ucol[(size_t)samp[a] * NU + uc];
          }
        }
  // FP64 refinement columns: all p rotated-centred columns + t residuals, by accession
  const int nbf = p + t;
  std::vector<double> B((size_t)md.n_acc * nbf, 0.0);
  for (int j = 0; j < n; ++j) {
    const int a = md.idxs[j];
    for (int k = 0; k < p; ++k) B[(size_t)a * nbf + k] = md.Uc[(size_t)k * n + j];
    for (int tr = 0; tr < t; ++tr) B[(size_t)a * nbf + p + tr] = md.rc[(size_t)tr * n + j];
  }
  // column records {w0, w1, w2, k_0..k_7, E, h^2, 0}; zero past nb
  std::vector<double> rec((size_t)ncol * kColRec, 0.0);
  double sum_h2 = 0.0, sum_E2 = 0.0;
  std::vector<double> sum_hbp(kMaxTraits, 0.0), sum_Ek(kMaxTraits, 0.0), sum_k2(kMaxTraits, 0.0);
  for (int b = 0; b < nb; ++b) {
    double* r = &rec[(size_t)b * kColRec];
    for (int l = 0; l < L; ++l) r[l] = colw[(size_t)b * L + l];
    r[11] = colE[b];
    r[12] = colh[b] * colh[b];
    if (b < t) {
      r[3 + b] = 1.0;
      sum_Ek[b] += colE[b];
    } else {
      const int k = zc[b - t];
      sum_h2 += colh[b] * colh[b];
      sum_E2 += colE[b] * colE[b];
      for (int tr = 0; tr < t; ++tr) {
        const double bpv = md.bprime[(size_t)tr * p + k];
        r[3 + tr] = bpv;
        sum_hbp[tr] += colh[b] * std::fabs(bpv);
        sum_k2[tr] += bpv * bpv;
        sum_Ek[tr] += colE[b] * bpv;
      }
    }
  }
  // per-trait records {mean r, |y - ybar|, yyM, b'_0 (dropped d column), h of r_c,
  // E of r_c, sum_b E_b k_tr, |k_tr| over the structure columns}
  std::vector<double> trc((size_t)t * kTrcRec, 0.0);
  for (int tr = 0; tr < t; ++tr) {
    double* r = &trc[(size_t)tr * kTrcRec];
    r[0] = md.rbar[tr];
    r[1] = md.ycn[tr];
    r[2] = md.yyM[tr];
    r[3] = (md.rotated && md.drop_d) ? md.bprime[(size_t)tr * p + 0] : 0.0;
    r[4] = colh[tr];
    r[5] = colE[tr];
    r[6] = sum_Ek[tr];
    r[7] = std::sqrt(sum_k2[tr]) * (1.0 + 1e-12);  // norm of the structure weights k_tr
  }
  auto up = [&](DBuf& d, const void* src, size_t bytes) -> hipError_t {
    hipError_t e = d.ensure(bytes);
    if (e != hipSuccess) return e;
    return hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice);
  };
  KG_CK(ctx, up(wts_b, W.data(), W.size()));
  if (primary) {
    KG_CK(ctx, up(M.mask, md.maskw.data(), md.maskw.size() * 8));
    KG_CK(ctx, up(M.rbar, md.rbar.data(), t * 8));
    KG_CK(ctx, up(M.B64, B.data(), B.size() * 8));
    // b' has t * p entries: none when U has no columns (one zero keeps the upload valid)
    std::vector<double> bpf(md.bprime);
    if (bpf.empty()) bpf.push_back(0.0);
    KG_CK(ctx, up(M.bpfull, bpf.data(), bpf.size() * 8));
  }
  KG_CK(ctx, up(colrec_b, rec.data(), rec.size() * 8));
  KG_CK(ctx, up(trc_b, trc.data(), trc.size() * 8));
  ep = EpiI8{};
  ep.n = n;
  ep.nz = nz;
  ep.t = t;
  ep.L = L;
  ep.ncol = ncol;
  ep.has_d = (md.rotated && !md.drop_d) ? 1 : 0;
  ep.sigma = md.sigma;
  ep.base_coef = md.base_coef;
  ep.dmax_drop = (md.rotated && md.drop_d) ? md.dmax : 0.0;
  ep.h_d = ep.has_d ? colh[t] : 0.0;  // structure column 0 is d when it is kept
  ep.inv_n = 1.0 / (double)n;
  ep.colrec = colrec_b.as<double>();
  ep.trc = trc_b.as<double>();
  ep.sum_h2 = sum_h2 * (1.0 + 1e-12);
  for (int tr = 0; tr < kMaxTraits; ++tr) ep.sum_hbp[tr] = sum_hbp[tr] * (1.0 + 1e-12);
  ep.sum_E2 = sum_E2;
  ep.E_d = ep.has_d ? colE[t] : 0.0;
  ep.stat_est = 0;
  return KG_OK;
}

// The FP6-digit resident set at ctx->res_digits digits (one trait; none when
// the context uses int8 limbs or the model does not fit the kernel).
static int build_f6(kg_ctx* ctx, ModelDev& M) {
  M.f6_ok = false;
  M.f6_digits = 0;
  const int D = ctx->res_digits;
  if (M.md.t != 1 || D == 0) return KG_OK;
  const int rc = build_i8_set(ctx, M, D, false, M.f6.NT, M.f6.wts, M.f6.colrec, M.f6.trc,
                              M.f6.epi, true, &M.f6.cc);
  if (rc == kNoFit) return KG_OK;
  if (rc != KG_OK) return rc;
  M.f6_ok = f6_res_fits(M.f6.NT, M.md.nch_i8);
  M.f6_digits = D;
  return KG_OK;
}

static int build_i8(kg_ctx* ctx, ModelDev& M) {
  int rc = build_i8_set(ctx, M, ctx->limbs, true, M.NT, M.wts, M.colrec, M.trc, M.epi);
  if (rc != KG_OK) return rc;
  // the resident-weight kernel's set: two limbs, one trait, every chunk in LDS
  M.res_ok = false;
  if (M.md.t == 1) {
    rc = build_i8_set(ctx, M, 2, false, M.res.NT, M.res.wts, M.res.colrec, M.res.trc, M.res.epi);
    if (rc == KG_OK) M.res_ok = i8_res_fits(M.res.NT, M.md.nch_i8);
    else if (rc != kNoFit) return rc;
  }
  return build_f6(ctx, M);
}

// Pre-filter screen weights: y - ybar of every trait as two balanced int8
// limbs, v_j = w (254 q_0 + q_1), and (for t <= 7) a count row holding 1 for
// every phenotyped accession, in the screen kernel's operand layout
// (ScreenDesc, kg_internal.h).  h bounds |(y_j - exact mean) - v_j|: the
// measured limb residual plus the FP64 rounding of the mean and of y - ybar.
// Digits of y - ybar for the one-trait screen: v = s4 sum_l d_l 30^(4 - l),
// |d_l| <= 15, s0 = max|v| / 15, s_{l+1} = s_l / 30; returns s4 and writes the
// digits [sample][5] (the residual after digit l is at most s_l / 2 = 15 s_{l+1}).
static double quantize_f6(const double* col, int n, std::vector<int8_t>& dig) {
  double scale = 0.0;
  for (int j = 0; j < n; ++j) scale = std::max(scale, std::fabs(col[j]));
  dig.assign((size_t)n * 5, 0);
  if (scale == 0.0) return 0.0;
  double sl[5];
  sl[0] = scale / 15.0;
  for (int l = 1; l < 5; ++l) sl[l] = sl[l - 1] / 30.0;
  for (int j = 0; j < n; ++j) {
    double res = col[j];
    for (int l = 0; l < 5; ++l) {
      double d = std::nearbyint(res / sl[l]);
      d = std::min(15.0, std::max(-15.0, d));
      dig[(size_t)j * 5 + l] = (int8_t)d;
      res -= d * sl[l];
    }
  }
  return sl[4];
}

static int build_screen(kg_ctx* ctx, ModelDev& M) {
  const Model& md = M.md;
  const int n = md.n, t = md.t, nch = md.nch_i8;
  M.screen_ok = false;
  if (t > kMaxTraits || nch > screen_max_chunks()) return KG_OK;
  const int Q = screen_rows(t), KG = 16 / Q;
  const bool count_row = screen_count_row(t);
  const bool f6 = Q == 4;  // one trait: FP6 digits against FP4 bits
  std::vector<std::vector<int8_t>> lim(t);
  std::vector<double> yc(n);
  for (int tr = 0; tr < t; ++tr) {
    const double* y = &md.Y[(size_t)tr * n];
    const double ym = md.ymean[tr];
    double ymax = 0.0;
    for (int j = 0; j < n; ++j) {
      yc[j] = y[j] - ym;
      ymax = std::max(ymax, std::fabs(y[j]));
    }
    // S~ = w * (integer digit combination over the set bits); h bounds each
    // sample's remainder (the FP64 rounding of it allowed for)
    const int SL = f6 ? 5 : 2;
    const double base = f6 ? 30.0 : 254.0;
    double w;
    if (f6) {
      w = quantize_f6(yc.data(), n, lim[tr]);
    } else {
      double wl[2], hq;
      quantize(yc.data(), n, 2, lim[tr], wl, &hq);
      w = wl[1];
    }
    double h = 0.0;
    for (int j = 0; j < n; ++j) {
      double iv = 0.0;
      for (int l = 0; l < SL; ++l) iv = base * iv + lim[tr][(size_t)SL * j + l];
      h = std::max(h, std::fabs(yc[j] - w * iv));
    }
    M.sw[tr] = w;
    M.sh[tr] = h * (1.0 + 1e-9) + 2.3e-16 * ((double)n + 8.0) * (ymax + std::fabs(ym));
  }
  std::vector<int> samp(md.n_acc, -1);
  for (int j = 0; j < n; ++j) samp[md.idxs[j]] = j;
  const size_t per = (size_t)8 * 64 * 16;  // bytes per 512-sample chunk (both forms)
  std::vector<int8_t> W((size_t)nch * per, 0);
  if (f6) {
    // [chunk][shift k 0..3][lane][32 bytes]: lane (g, row) holds row `row` of
    // K-block g; shift k uses rows 8 (k % 2) + {0..4: digits, 5: count}; element
    // j = 8 d + m (6 bits at bit 6 j) is accession 512 C + 128 g + 32 d + 4 m + k
    for (int C = 0; C < nch; ++C)
      for (int k = 0; k < 4; ++k)
        for (int lane = 0; lane < 64; ++lane) {
          const int g = lane >> 4, row = lane & 15;
          if (row / 8 != k % 2) continue;
          const int qn = row % 8;
          if (qn > 5) continue;
          uint32_t bits[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          for (int j = 0; j < 32; ++j) {
            const int a = 512 * C + 128 * g + 32 * (j >> 3) + 4 * (j & 7) + k;
            if (a >= md.n_acc || samp[a] < 0) continue;
            const int d = qn == 5 ? 1 : lim[0][(size_t)samp[a] * 5 + qn];
            const uint64_t code = f6_code(d);
            const int b0 = 6 * j;
            bits[b0 >> 5] |= (uint32_t)(code << (b0 & 31));
            if ((b0 & 31) > 26) bits[(b0 >> 5) + 1] |= (uint32_t)(code >> (32 - (b0 & 31)));
          }
          memcpy(&W[(size_t)C * per + (((size_t)k * 64 + lane) * 32)], bits, 32);
        }
  } else {
    for (int C = 0; C < nch; ++C)
      for (int k = 0; k < 8; ++k)
        for (int lane = 0; lane < 64; ++lane) {
          const int g = lane >> 4, row = lane & 15;
          if (row / Q != k % KG) continue;  // rows of another shift group
          const int qn = row % Q, tr = qn >> 1, l = qn & 1;
          const bool is_count = count_row && qn == 2 * t;
          if (tr >= t && !is_count) continue;
          int8_t* dst = &W[(size_t)C * per + ((size_t)k * 64 + lane) * 16];
          for (int e = 0; e < 16; ++e) {
            // operand byte e = 4 i + q: bit 8 q + k of the lane group's dword i
            const int a = 512 * C + 128 * g + 32 * (e >> 2) + 8 * (e & 3) + k;
            if (a < md.n_acc && samp[a] >= 0)
              dst[e] = is_count ? (int8_t)1 : lim[tr][(size_t)samp[a] * 2 + l];
          }
        }
  }
  KG_CK(ctx, M.swts.ensure(W.size()));
  KG_CK(ctx, hipMemcpy(M.swts.p, W.data(), W.size(), hipMemcpyHostToDevice));
  M.screen_ok = true;
  return KG_OK;
}

static int build_f32(kg_ctx* ctx, ModelDev& M) {
  Model& md = M.md;
  const int n = md.n, p = md.p, t = md.t;
  const int ncol = p + 2 * t;
  const int NT = (ncol + 15) / 16;
  if (NT > 6) return fail(ctx, KG_ERR_ARG, "p + 2t exceeds the FP32 kernel's 96 columns");
  M.NT = NT;
  std::vector<float> colv((size_t)n * 16 * NT, 0.f);  // [sample][col]
  const int NU = 16 * NT;
  EpiF32& ep = M.epf;
  memset(&ep, 0, sizeof(ep));
  std::vector<float> Uyf((size_t)t * (p > 0 ? p : 1), 0.f);
  for (int tr = 0; tr < t; ++tr) {
    const double* y = &md.Y[(size_t)tr * n];
    std::vector<float> yf(n);
    float ys = 0.f;
    for (int j = 0; j < n; ++j) {
      yf[j] = (float)y[j];
      ys += yf[j];
    }
    const float ymean = ys / (float)n;  // src/kgwas.cpp:187-190 in FP32
    float ycs = 0.f, yn2 = 0.f;
    for (int j = 0; j < n; ++j) {
      const float d = yf[j] - ymean;
      ycs += d;
      yn2 += d * d;
      colv[(size_t)j * NU + p + tr] = yf[j];
      colv[(size_t)j * NU + p + t + tr] = d;
    }
    ep.ymean_f[tr] = ymean;
    ep.ycsum_f[tr] = ycs;
    ep.ynorm_f[tr] = std::sqrt(yn2);
    ep.yyM[tr] = md.yyM[tr];
    for (int k = 0; k < p; ++k) {  // src/kgwas.cpp:149 Uyf = Uf' yf
      float s = 0.f;
      for (int j = 0; j < n; ++j) s += (float)md.U[(size_t)k * n + j] * yf[j];
      Uyf[(size_t)tr * p + k] = s;
    }
  }
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < p; ++k) colv[(size_t)j * NU + k] = (float)md.U[(size_t)k * n + j];
  std::vector<int> samp(md.n_acc, -1);
  for (int j = 0; j < n; ++j) samp[md.idxs[j]] = j;
  const int nch = md.nch_f32;
  const size_t frag = (size_t)64 * NT * 64;  // floats per chunk
  std::vector<float> W((size_t)nch * frag, 0.f);
  for (int C = 0; C < nch; ++C)
    for (int e = 0; e < 64; ++e)
      for (int tt = 0; tt < NT; ++tt)
        for (int lane = 0; lane < 64; ++lane) {
          const int i = lane & 15, g = lane >> 4;
          const int a = 256 * C + 64 * g + e;
          if (a < md.n_acc && samp[a] >= 0)
            W[(size_t)C * frag + ((size_t)e * NT + tt) * 64 + lane] =
                colv[(size_t)samp[a] * NU + 16 * tt + i];
        }
  KG_CK(ctx, M.wf32.ensure(W.size() * 4));
  KG_CK(ctx, hipMemcpy(M.wf32.p, W.data(), W.size() * 4, hipMemcpyHostToDevice));
  KG_CK(ctx, M.mask.ensure(md.maskw.size() * 8));
  KG_CK(ctx, hipMemcpy(M.mask.p, md.maskw.data(), md.maskw.size() * 8, hipMemcpyHostToDevice));
  KG_CK(ctx, M.uyf.ensure(Uyf.size() * 4));
  KG_CK(ctx, hipMemcpy(M.uyf.p, Uyf.data(), Uyf.size() * 4, hipMemcpyHostToDevice));
  ep.n = n;
  ep.p = p;
  ep.t = t;
  ep.NU = NU;
  ep.Uyf = M.uyf.as<float>();
  return KG_OK;
}

// ---------------------------------------------------------------------------

extern "C" int kg_create(kg_ctx** out, const kg_opts* opts) {
  if (!out) return KG_ERR_ARG;
  *out = nullptr;
  kg_ctx* ctx = new kg_ctx();
  if (opts) {
    ctx->device = opts->device;
    ctx->precision = opts->precision;
    if (opts->limbs) ctx->limbs = opts->limbs;
    ctx->screen_mode = opts->screen;
    ctx->dedup = opts->dedup;
  }
  if (ctx->dedup != 0 && ctx->dedup != 1) {
    delete ctx;
    return KG_ERR_ARG;
  }
  if (ctx->screen_mode < KG_SCREEN_AUTO || ctx->screen_mode > KG_SCREEN_ON) {
    delete ctx;
    return KG_ERR_ARG;
  }
  if (ctx->precision != KG_PREC_MIXED && ctx->precision != KG_PREC_FP32) {
    delete ctx;
    return KG_ERR_ARG;
  }
  if (ctx->limbs < 1 || ctx->limbs > 3) {
    delete ctx;
    return KG_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= ctx->device || ctx->device < 0) {
    *out = ctx;  // return the context so the caller can read the message
    ctx->err = "no HIP device " + std::to_string(ctx->device) + " (found " +
               std::to_string(ndev) + ")";
    return KG_ERR_DEVICE;
  }
  *out = ctx;
  KG_CK(ctx, hipSetDevice(ctx->device));
  hipDeviceProp_t prop;
  KG_CK(ctx, hipGetDeviceProperties(&prop, ctx->device));
  ctx->cus = prop.multiProcessorCount;
  KG_CK(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  KG_CK(ctx, hipStreamCreateWithFlags(&ctx->cstream, hipStreamNonBlocking));
  for (auto& e : ctx->ev) KG_CK(ctx, hipEventCreate(&e));
  KG_CK(ctx, hipEventCreate(&ctx->evs));
  for (auto& e : ctx->cev) KG_CK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (auto& e : ctx->ctev) KG_CK(ctx, hipEventCreate(&e));
  KG_CK(ctx, set_kernel_attributes());
  KG_CK(ctx, noalloc_scratch(ctx->device, &ctx->scratch));
  if (const char* v = std::getenv("KG_GRID_MULT")) {
    const int g = atoi(v);
    if (g >= 1 && g <= 64) ctx->grid_mult = g;
  }
  return KG_OK;
}

static void drop_graph(kg_ctx* ctx) {
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->gexec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_key.clear();
}

extern "C" void kg_destroy(kg_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  drop_graph(ctx);
  ctx->models.clear();
  if (ctx->cstream) (void)hipStreamSynchronize(ctx->cstream);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->evs) (void)hipEventDestroy(ctx->evs);
  for (auto& e : ctx->cev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->ctev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->cstream) (void)hipStreamDestroy(ctx->cstream);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* kg_last_error(const kg_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

static int add_model(kg_ctx* ctx, int32_t n_acc, const int32_t* idxs, int32_t n, const double* U,
                     int32_t p, const double* Y, int32_t t, int32_t* model_id) {
  if (!ctx) return KG_ERR_ARG;
  if (!ctx->stream) return fail(ctx, KG_ERR_STATE, "context has no device");
  if (n_acc <= 0 || n <= 0 || n > n_acc || p < 0 || t < 1 || t > kMaxTraits || !idxs || !Y ||
      (p > 0 && !U))
    return fail(ctx, KG_ERR_ARG,
                "kg_set_model: bad shape (need 0 < n <= n_acc, p >= 0, 1 <= t <= 8)");
  for (int j = 0; j < n; ++j) {
    if (idxs[j] < 0 || idxs[j] >= n_acc || (j > 0 && idxs[j] <= idxs[j - 1]))
      return fail(ctx, KG_ERR_ARG, "kg_set_model: idxs must be increasing accession positions");
  }
  KG_CK(ctx, hipSetDevice(ctx->device));
  auto M = std::make_unique<ModelDev>();
  Model& md = M->md;
  md.n_acc = n_acc;
  md.n = n;
  md.p = p;
  md.t = t;
  md.idxs.assign(idxs, idxs + n);
  if (p > 0) md.U.assign(U, U + (size_t)n * p);
  md.Y.assign(Y, Y + (size_t)n * t);
  md.nch_i8 = (n_acc + 511) / 512;
  md.nch_f32 = (n_acc + 255) / 256;
  md.maskw.assign((size_t)8 * md.nch_i8, 0ull);
  for (int j = 0; j < n; ++j) md.maskw[idxs[j] >> 6] |= 1ull << (idxs[j] & 63);
  md.mask_full = true;
  for (uint64_t w : md.maskw) md.mask_full = md.mask_full && w == ~0ull;
  prepare_model(md);
  int rc = (ctx->precision == KG_PREC_MIXED) ? build_i8(ctx, *M) : build_f32(ctx, *M);
  if (rc == KG_OK && ctx->precision == KG_PREC_MIXED) rc = build_screen(ctx, *M);
  if (rc != KG_OK) return rc;
  drop_graph(ctx);
  ctx->models.push_back(std::move(M));
  if (model_id) *model_id = (int32_t)ctx->models.size() - 1;
  return KG_OK;
}

extern "C" int kg_set_model(kg_ctx* ctx, int32_t n_acc, const int32_t* idxs, int32_t n,
                            const double* U, int32_t p, const double* Y, int32_t t) {
  if (!ctx) return KG_ERR_ARG;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  ctx->models.clear();
  ctx->screen_skip = 0;
  ctx->full_m = 0;
  ctx->probe_tick = 0;
  return add_model(ctx, n_acc, idxs, n, U, p, Y, t, nullptr);
}

extern "C" int kg_add_model(kg_ctx* ctx, int32_t n_acc, const int32_t* idxs, int32_t n,
                            const double* U, int32_t p, const double* Y, int32_t t,
                            int32_t* model_id) {
  return add_model(ctx, n_acc, idxs, n, U, p, Y, t, model_id);
}

// Final decision for one candidate (k-mer idx, trait) with FP64 values.
struct Decided {
  bool pass;
  bool hit;
  kg_hit h;
};

static Decided decide(const Model& md, int64_t idx, int tr, int c, double xxM, double xyM,
                      double S, bool have_S, const kg_scan_params& prm, bool pairs = false) {
  Decided d{};
  d.pass = true;
  // all-pairs scan: no pre-filter, and a pair with xxM or yyM <= 1e-12 is
  // skipped before its p-value (src/kgwas_2d.cpp:380), even at p_thresh = 1
  if (pairs && !(xxM > 1e-12 && md.yyM[tr] > 1e-12)) return d;
  if (have_S) {  // FP64 pre-filter, src/kgwas.cpp:462-468
    const int n = md.n;
    const double v = (double)c * (double)(n - c) / (double)n;
    const double xcn = std::sqrt(v > 0.0 ? v : 0.0);
    const double ycn = md.ycn[tr];
    if (xcn < 1e-12 || ycn < 1e-12) d.pass = false;
    else if (std::fabs(S / (xcn * ycn)) < prm.corr_threshold) d.pass = false;
  }
  if (!d.pass) return d;
  double eff, T, pv;
  host_hit(xxM, xyM, md.yyM[tr], md.n, &eff, &T, &pv);
  const double p_thr = std::pow(10.0, -prm.log10_thresh);
  if (pv <= p_thr) {
    d.hit = true;
    d.h.idx = idx;
    d.h.trait = tr;
    d.h.c = c;
    d.h.flags = 0;
    if (prm.log10_thresh_2d > 0.0 && pv <= std::pow(10.0, -prm.log10_thresh_2d))
      d.h.flags |= KG_HIT_2D;
    d.h.xxM = xxM;
    d.h.xyM = xyM;
    d.h.effect = eff;
    d.h.T = T;
    d.h.pval = pv;
  }
  return d;
}

// A shard as the host sees it.
struct HShard {
  const uint64_t* bits;
  int64_t m, rw, idx0;
  int model;
  const uint64_t* mask = nullptr;  // device accession mask words (null: the model's)
};

static int check_shard(kg_ctx* ctx, const HShard& s) {
  if (s.model < 0 || s.model >= (int)ctx->models.size())
    return fail(ctx, KG_ERR_ARG, "shard refers to an unknown model");
  const Model& md = ctx->models[s.model]->md;
  if (s.m < 0) return fail(ctx, KG_ERR_ARG, "negative k-mer count");
  if (s.m > 0 && !s.bits) return fail(ctx, KG_ERR_ARG, "null bits");
  if (s.rw % 8 != 0 || s.rw < 8 * (int64_t)md.nch_i8)
    return fail(ctx, KG_ERR_ARG,
                "row_words must be a multiple of 8 and cover all accessions (>= " +
                    std::to_string(8 * md.nch_i8) + ")");
  if (((uintptr_t)s.bits & 15) != 0) return fail(ctx, KG_ERR_ARG, "bits must be 16-byte aligned");
  return KG_OK;
}

// Device tables and launch shape of a screened scan.
struct ScreenRun {
  bool on = false;
  const ScreenDesc* d_ss = nullptr;
  ShardDesc* d_sd2 = nullptr;   // full-kernel table over the survivor lists
  int64_t* d_total = nullptr;   // full-kernel block count, written by scr_scan
  const ScrChunk* d_ch = nullptr;
  int ncc = 0;                  // compaction chunks
  uint32_t* ccount = nullptr;
  int64_t* P = nullptr;
  int64_t sblocks = 0;          // screen blocks
  int sgrid = 0;
  int nch_max = 0;
  bool dd = false;              // with the duplicate-row check
  bool w8 = false;              // the screen's 8-wave shape
  int ring = 0;                 // the LDS-ring screen's waves per workgroup (0: not used)
  bool res = false;             // survivors through the resident-weight kernel
  int res_nt = 0;               // its weight tiles
  int res_d = 0;                // its FP6 digits per column (0: two int8 limbs)
  int res_cm = kCntMasked;      // FP6: how the allele count is formed (kCnt*)
  // recheck pass of the k-mers the two-digit kernel leaves open: the streaming
  // kernel (primary weights, `rchk_L` limbs, two tiles per wave) over the
  // device-built table d_sd3 (d_total3 blocks of bk3 k-mers)
  bool rchk = false;
  ShardDesc* d_sd3 = nullptr;
  int64_t* d_total3 = nullptr;
  int64_t bk3 = 0;
  int64_t rblocks_full = 0;     // its block count if every k-mer went to it (grid bound)
  int64_t bk = 0;               // k-mers per block of the full kernel over the survivors
  int64_t sblocks_full = 0;     // its block count if every k-mer survived (grid bound)
};

// Device part of a batched int8 scan: zero counters (unless the table upload
// did), [screen, fix-up,] bulk kernel(s), refine (with the host read-back).
static int enqueue_i8(kg_ctx* ctx, int NT, int t, int nsh, int64_t total_blocks, bool per_launch,
                      const std::vector<int64_t>& nblk, const ScanOut& o, int ref_blocks,
                      const ShardDesc* d_sd, const ShardDesc* d_sd1, const RefDesc* d_rd,
                      const ScreenRun& sr, bool zeroed, const RefHost& hb) {
  const int L = ctx->limbs;
  const int grid = ctx->cus * ctx->grid_mult;
  if (!zeroed) KG_CK(ctx, launch_zero(ctx->stream, o.counters, (int)ctx->counters_len));
  KG_CK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  if (sr.on) {
    if (sr.ring)
      KG_CK(ctx, launch_screen_ring(sr.ring, sr.nch_max, sr.sgrid, ctx->stream, sr.d_ss, nsh,
                                    sr.sblocks, sr.dd));
    else
      KG_CK(ctx, launch_screen(t, sr.nch_max, sr.sgrid, ctx->stream, sr.d_ss, nsh, sr.sblocks,
                               sr.dd, sr.w8));
    KG_CK(ctx, hipEventRecord(ctx->evs, ctx->stream));
    KG_CK(ctx, launch_screen_compact(ctx->stream, sr.d_ch, sr.ncc, sr.d_ss, sr.ccount, sr.P,
                                     sr.d_sd2, nsh, sr.bk, sr.d_total));
    if (sr.res && sr.res_d > 0) {
      KG_CK(ctx, launch_scan_f6_res(sr.res_nt, sr.res_d, sr.res_cm, sr.nch_max, ctx->cus,
                                    ctx->stream,
                                    sr.d_sd2, nsh, sr.sblocks_full, o, sr.d_total));
      if (sr.rchk) {
        KG_CK(ctx, launch_rchk_tables(ctx->stream, sr.d_sd3, nsh, sr.bk3, sr.d_total3));
        KG_CK(ctx, launch_scan_i8(NT, L, t, 2, grid, ctx->stream, sr.d_sd3, nsh,
                                  sr.rblocks_full, o, sr.d_total3));
      }
    } else if (sr.res)
      KG_CK(ctx, launch_scan_i8_res(sr.res_nt, sr.nch_max, ctx->cus, ctx->stream, sr.d_sd2, nsh,
                                    sr.sblocks_full, o, sr.d_total));
    else
      KG_CK(ctx, launch_scan_i8(NT, L, t, ctx->tiles, grid, ctx->stream, sr.d_sd2, nsh,
                                total_blocks, o, sr.d_total));
  } else if (!per_launch) {
    KG_CK(ctx, launch_scan_i8(NT, L, t, ctx->tiles, grid, ctx->stream, d_sd, nsh, total_blocks,
                              o, nullptr));
  } else {
    for (int s = 0; s < nsh; ++s) {
      if (nblk[s] == 0) continue;
      KG_CK(ctx, launch_scan_i8(NT, L, t, ctx->tiles, grid, ctx->stream, d_sd1 + s, 1, nblk[s],
                                o, nullptr));
    }
  }
  KG_CK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
  KG_CK(ctx, launch_refine(ctx->stream, o.cand, o.counters, o.cap, d_rd,
                           ctx->res.as<double>(), ref_blocks, hb));
  KG_CK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
  return KG_OK;
}

enum { kPerLaunch = 1, kGraph = 2, kPairs = 4 };
constexpr int64_t kEagerCands = 4096;  // candidates read back with the counters
constexpr int64_t kProbeKmers = int64_t(1) << 22;  // auto screen: scans that measure unscreened

// Descriptors of shard s (index `sid`) for the int8 bulk kernel and the FP64
// refinement; its pre-filter passes are counted at `pass`, its blocks start at
// global block blk0.
static void fill_descs(const kg_ctx* ctx, const HShard& sh, int sid, const kg_scan_params& prm,
                       double Ts, double p_thr, bool pairs, bool stats, unsigned long long* pass,
                       int64_t blk0, ShardDesc& D, RefDesc& R) {
  const ModelDev& M = *ctx->models[sh.model];
  const Model& md = M.md;
  D.bits = sh.bits;
  D.m = sh.m;
  D.rw = sh.rw;
  D.blk0 = blk0;
  D.wts = M.wts.p;
  D.mask = sh.mask ? static_cast<const void*>(sh.mask) : M.mask.p;
  D.nchunks = md.nch_i8;
  D.pad = 0;
  D.pass = pass;
  D.id = sid;
  D.pad2 = 0;
  D.rows = nullptr;
  D.mult = nullptr;
  D.rchk = nullptr;
  D.rchk_n = nullptr;
  D.ep = M.epi;
  D.ep.tau = prm.corr_threshold;
  D.ep.tau_trivial = prm.corr_threshold <= 0.0 ? 1 : 0;
  D.ep.all_sig = p_thr >= 1.0 ? 1 : 0;
  D.ep.pairs = pairs ? 1 : 0;
  D.ep.stat_est = stats ? 1 : 0;
  const double Tmax = (double)md.n * std::log(1.0 / (1.0 - (1 - 1e-15)));
  D.ep.r2star = (Ts > Tmax) ? INFINITY : -std::expm1(-Ts / (double)md.n);
  R.bits = sh.bits;
  R.rw = sh.rw;
  R.maskw = sh.mask ? sh.mask : M.mask.as<uint64_t>();
  R.B64 = M.B64.as<double>();
  R.bpfull = M.bpfull.as<double>();
  R.rbar = M.rbar.as<double>();
  R.sigma = md.sigma;
  R.base_coef = md.base_coef;
  R.nwords = (md.n_acc + 63) / 64;
  R.p = md.p;
  R.t = md.t;
  R.n = md.n;
}

// Scans every shard (int8 path), appending hits (idx offset by shard idx0,
// shard index set) and adding pre-filter passes to n_passed[s * t + tr].
// `overlap`, when given, runs once on the host after the first attempt's device
// work is enqueued and before the host waits for it (kg_scan issues the next
// chunk's copy there).
using Overlap = std::function<int()>;

static int scan_i8_shards(kg_ctx* ctx, const std::vector<HShard>& sh, const kg_scan_params& prm,
                          unsigned flags, std::vector<kg_hit>& hits, int64_t* n_passed,
                          bool stats, const Overlap* overlap = nullptr) {
  const int nsh = (int)sh.size();
  const int t = ctx->models[sh[0].model]->md.t;
  const int NT = ctx->models[sh[0].model]->NT;
  for (const auto& s : sh) {
    const ModelDev& M = *ctx->models[s.model];
    if (M.NT != NT || M.md.t != t)
      return fail(ctx, KG_ERR_ARG, "all shards of one scan need the same p, t and limb count");
  }
  const double p_thr = std::pow(10.0, -prm.log10_thresh);
  const double Ts = t_star(p_thr);
  const int64_t bk = i8_block_kmers(NT, ctx->limbs, t, ctx->tiles);
  std::vector<ShardDesc> sd(nsh), sd1(nsh);
  std::vector<RefDesc> rd(nsh);
  std::vector<int64_t> nblk(nsh);
  int64_t total_blocks = 0;
  const bool per_launch = (flags & kPerLaunch) != 0;
  const bool pairs = (flags & kPairs) != 0;
  // pre-filter screen: only where it can drop k-mers (tau > 0) and every
  // shard's model has screen weights; row lists are 32-bit
  bool screen = !stats && !per_launch && !pairs && prm.corr_threshold > 0.0 &&
                ctx->screen_mode != KG_SCREEN_OFF;
  int nch_max = 0;
  int64_t mtot = 0;
  for (const auto& s : sh) {
    const ModelDev& M = *ctx->models[s.model];
    if (!M.screen_ok || s.m > (int64_t)kRowMask || s.rw > (int64_t)INT32_MAX) screen = false;
    nch_max = std::max(nch_max, M.md.nch_i8);
    mtot += s.m;
  }
  // auto mode: unscreened while a measurement said the screen does not pay;
  // small scans also run unscreened once per 32 (or at a new size) to measure
  // the full kernel directly
  const bool auto_scr = screen && ctx->screen_mode == KG_SCREEN_AUTO && !ctx->dedup &&
                        (flags & kGraph) == 0;
  if (auto_scr) {
    const bool same_size = ctx->full_m > 0 && mtot <= 2 * ctx->full_m && 2 * mtot >= ctx->full_m;
    if (ctx->screen_skip > 0) {
      --ctx->screen_skip;
      screen = false;
    } else if (mtot <= kProbeKmers && (!same_size || ++ctx->probe_tick >= 32)) {
      ctx->probe_tick = 0;
      screen = false;
    }
  }
  // counters: [cand][passes][survivors][total][recheck counts][recheck total]
  const size_t base_surv = 1 + (size_t)nsh * t;
  const size_t base_rchk = base_surv + (screen ? (size_t)nsh + 1 : 0);
  const bool dd = screen && ctx->dedup != 0;
  // single trait: the survivors go through the resident-weight kernel (the
  // screen then proves the pre-filter passes): FP6 digits when the context
  // asks for them and every model has that set, else two int8 limbs where every
  // model's weights fit its LDS; two digits come with the recheck pass (not in
  // duplicate-row scans, whose run lengths it does not carry)
  bool use_res = screen && t == 1 && ctx->resident;
  int res_nt = 0, res_d = 0;
  if (use_res && ctx->res_digits > 0 && !(ctx->res_digits == 2 && dd)) {
    res_d = ctx->res_digits;
    for (const auto& s : sh) {
      const ModelDev& M = *ctx->models[s.model];
      if (!M.f6_ok || M.f6_digits != res_d || (res_nt != 0 && M.f6.NT != res_nt)) {
        res_d = 0;
        break;
      }
      res_nt = M.f6.NT;
    }
    if (res_d > 0 && !f6_res_fits(res_nt, nch_max)) res_d = 0;
    if (res_d == 0) res_nt = 0;
  }
  // FP6 allele count: from the count column when every model has one, else a
  // bit count, of the bits themselves when every model covers every accession
  // of its chunks; the shards' own masks (none in screened scans) need the
  // masked count
  int res_cm = kCntMasked;
  if (res_d > 0) {
    bool all_cc = true, all_full = true;
    for (const auto& s : sh) {
      const ModelDev& M = *ctx->models[s.model];
      all_cc = all_cc && M.f6.cc && s.mask == nullptr;
      all_full = all_full && M.md.mask_full && s.mask == nullptr;
    }
    res_cm = all_cc ? kCntColumn : (all_full ? kCntBits : kCntMasked);
  }
  if (use_res && res_d == 0) {
    for (const auto& s : sh) {
      const ModelDev& M = *ctx->models[s.model];
      if (!M.res_ok || (res_nt != 0 && M.res.NT != res_nt)) use_res = false;
      res_nt = M.res.NT;
    }
    if (use_res && !i8_res_fits(res_nt, nch_max)) use_res = false;
  }
  const bool rchk = use_res && res_d == 2;
  ctx->counters_len = base_rchk + (rchk ? (size_t)nsh + 1 : 0);
  KG_CK(ctx, ctx->counters.ensure(sizeof(unsigned long long) * ctx->counters_len));
  unsigned long long* dc = ctx->counters.as<unsigned long long>();
  for (int s = 0; s < nsh; ++s) {
    fill_descs(ctx, sh[s], s, prm, Ts, p_thr, pairs, stats, dc + 1 + (size_t)s * t, total_blocks,
               sd[s], rd[s]);
    nblk[s] = (sh[s].m + bk - 1) / bk;
    total_blocks += nblk[s];
    sd1[s] = sd[s];
    sd1[s].blk0 = 0;
  }
  // screen tables: survivors of shard s go to surv[off_s, off_s + m_s); its
  // bitmap covers its screen blocks; compaction chunks of at most
  // screen_chunk_words() bitmap words never straddle shards
  std::vector<ScreenDesc> ss;
  std::vector<ShardDesc> sd2, sd3;
  std::vector<ScrChunk> chunks;
  ScreenRun sr;
  if (screen) {
    sr.res = use_res;
    sr.res_nt = res_nt;
    sr.res_d = res_d;
    sr.res_cm = res_cm;
    sr.rchk = rchk;
    sr.bk = use_res ? i8_res_block_kmers() : bk;
    sr.bk3 = i8_block_kmers(NT, ctx->limbs, t, 2);
    sr.w8 = ctx->scr_w8;
    // the LDS-ring screen where every model fits it (one trait, 4 or 8 chunks)
    sr.ring = (t == 1) ? ctx->scr_ring : 0;
    for (const auto& s : sh)
      if (sr.ring && !screen_ring_fits(ctx->models[s.model]->md.nch_i8, sr.ring)) sr.ring = 0;
    if (sr.ring && !screen_ring_fits(nch_max, sr.ring)) sr.ring = 0;
    const int64_t sbk = sr.ring ? 64 * (int64_t)sr.ring : screen_block_kmers(t, sr.w8);
    const int64_t cw = screen_chunk_words();
    int64_t bm_words = 0;
    // one trait (the 4-row screen): each 64 k-mers' kept words are followed by
    // their proven words
    const bool ilv = screen_rows(t) == 4;
    const int64_t wpb = (sbk / 32) * (ilv ? 2 : 1);  // bitmap words per screen block
    for (int s = 0; s < nsh; ++s) bm_words += (sh[s].m + sbk - 1) / sbk * wpb;
    KG_CK(ctx, ctx->surv.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(mtot, 1)));
    KG_CK(ctx, ctx->bmap.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(bm_words, 1)));
    if (dd) {
      KG_CK(ctx, ctx->dmap.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(bm_words, 1)));
      KG_CK(ctx, ctx->smult.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(mtot, 1)));
    }
    if (rchk) KG_CK(ctx, ctx->rchk.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(mtot, 1)));
    uint32_t* sv = ctx->surv.as<uint32_t>();
    uint32_t* bmw = ctx->bmap.as<uint32_t>();
    uint32_t* dmw = dd ? ctx->dmap.as<uint32_t>() : nullptr;
    uint32_t* mlw = dd ? ctx->smult.as<uint32_t>() : nullptr;
    ss.resize(nsh);
    sd2 = sd;
    if (rchk) sd3 = sd;
    int64_t off = 0, boff = 0;
    for (int s = 0; s < nsh; ++s) {
      const ModelDev& M = *ctx->models[sh[s].model];
      const Model& md = M.md;
      ScreenDesc& S = ss[s];
      S.bits = sh[s].bits;
      S.m = sh[s].m;
      S.rw = sh[s].rw;
      S.blk0 = sr.sblocks;
      sr.sblocks += (sh[s].m + sbk - 1) / sbk;
      S.wts = M.swts.p;
      S.mask = M.mask.p;
      S.bmap = bmw + boff;
      S.dmap = dd ? dmw + boff : nullptr;
      boff += (sh[s].m + sbk - 1) / sbk * wpb;
      S.list = sv + off;
      S.mult = dd ? mlw + off : nullptr;
      S.nsurv = dc + base_surv + s;
      S.chunk0 = (int)chunks.size();
      const int64_t used_words = (sh[s].m + 31) / 32;
      for (int64_t w0 = 0; w0 < used_words; w0 += cw) {
        ScrChunk c{};
        c.shard = s;
        c.w0 = w0;
        c.w1 = std::min<int64_t>(w0 + cw, used_words);
        chunks.push_back(c);
      }
      S.chunk1 = (int)chunks.size();
      S.nchunks = md.nch_i8;
      S.t = md.t;
      S.n = md.n;
      S.ilv = ilv ? 1 : 0;
      // FP32 screen constants: w and the threshold to nearest (their relative
      // error sits inside the kernel's 1e-5 margin), h rounded up
      for (int tr = 0; tr < kMaxTraits; ++tr) {
        S.wf[tr] = tr < md.t ? (float)M.sw[tr] : 0.0f;
        S.hf[tr] = tr < md.t ? (float)(M.sh[tr] * (1.0 + 1e-6)) : 0.0f;
        S.thrf[tr] = (tr < md.t && md.ycn[tr] >= 1e-12)
                         ? (float)(prm.corr_threshold * prm.corr_threshold * md.ycn[tr] *
                                   md.ycn[tr] / (double)md.n)
                         : -1.0f;
      }
      sd2[s].rows = S.list;
      sd2[s].mult = S.mult;
      sd2[s].m = 0;     // written on the device by the fix-up
      sd2[s].blk0 = 0;
      if (use_res) {  // the resident set, with this scan's thresholds
        const EpiI8& e0 = sd[s].ep;
        const I8Set& rs = res_d > 0 ? M.f6 : M.res;
        EpiI8 e = rs.epi;
        e.tau = e0.tau;
        e.tau_trivial = e0.tau_trivial;
        e.all_sig = e0.all_sig;
        e.pairs = e0.pairs;
        e.stat_est = e0.stat_est;
        e.r2star = e0.r2star;
        sd2[s].wts = rs.wts.p;
        sd2[s].ep = e;
      }
      if (rchk) {
        // shard s's open k-mers go to rchk[off_s, ...) (at most its m_s
        // survivors) and are scanned again by the streaming kernel
        uint32_t* rl = ctx->rchk.as<uint32_t>() + off;
        unsigned long long* rn = dc + base_rchk + s;
        sd2[s].rchk = rl;
        sd2[s].rchk_n = rn;
        sd3[s].rows = rl;
        sd3[s].rchk_n = rn;  // read by rchk_tables
        sd3[s].m = 0;        // written on the device
        sd3[s].blk0 = 0;
        sr.rblocks_full += (sh[s].m + sr.bk3 - 1) / sr.bk3;
      }
      sr.sblocks_full += (sh[s].m + sr.bk - 1) / sr.bk;
      off += sh[s].m;
    }
    sr.on = true;
    sr.dd = dd;
    sr.nch_max = nch_max;
    sr.d_total = reinterpret_cast<int64_t*>(dc + base_surv + nsh);
    if (rchk) sr.d_total3 = reinterpret_cast<int64_t*>(dc + base_rchk + nsh);
    sr.ncc = (int)chunks.size();
    KG_CK(ctx, ctx->scr_aux.ensure(sizeof(uint32_t) * (chunks.size() + 2) +
                                   sizeof(int64_t) * (chunks.size() + 2)));
    sr.P = ctx->scr_aux.as<int64_t>();
    sr.ccount = reinterpret_cast<uint32_t*>(sr.P + chunks.size() + 2);
    const int okey = ((t * 64 + nch_max) * 2 + (dd ? 1 : 0)) * 2 + (sr.w8 ? 1 : 0);
    if (ctx->scr_occ_key != okey) {
      ctx->scr_occ = screen_wgs_per_cu(t, nch_max, dd, sr.w8);
      ctx->scr_occ_key = okey;
    }
    sr.sgrid = sr.ring ? ctx->cus : ctx->cus * ctx->scr_occ;  // the ring: one workgroup per CU
  }
  if (ctx->cand_cap == 0) {
    int64_t mt = 0;
    for (const auto& s : sh) mt += s.m;
    ctx->cand_cap = std::max<int64_t>(1 << 16, std::min<int64_t>(mt / 16 + 1024, 1 << 24));
  }
  KG_CK(ctx, ctx->cand.ensure(sizeof(Cand) * ctx->cand_cap));
  KG_CK(ctx, ctx->res.ensure(sizeof(double) * 4 * ctx->cand_cap));
  KG_CK(ctx, ctx->sdesc.ensure(sizeof(ShardDesc) * nsh));
  KG_CK(ctx, ctx->sdesc1.ensure(sizeof(ShardDesc) * nsh));
  KG_CK(ctx, ctx->rdesc.ensure(sizeof(RefDesc) * nsh));
  ScanOut o{};
  if (stats) {
    const int64_t m = sh[0].m;
    KG_CK(ctx, ctx->st_c.ensure(sizeof(int32_t) * m));
    KG_CK(ctx, ctx->st_xx.ensure(sizeof(double) * m));
    KG_CK(ctx, ctx->st_xy.ensure(sizeof(double) * m * t));
    KG_CK(ctx, ctx->st_S.ensure(sizeof(double) * m * t));
    o.st_c = ctx->st_c.as<int32_t>();
    o.st_xx = ctx->st_xx.as<double>();
    o.st_xy = ctx->st_xy.as<double>();
    o.st_S = ctx->st_S.as<double>();
  }
  const bool use_graph = (flags & kGraph) != 0 && !stats;
  ctx->stat_patch.clear();
  const int ref_blocks = ctx->cus * 2;
  std::vector<unsigned long long> cnt(ctx->counters_len);
  KG_CK(ctx, ctx->hcnt.ensure(sizeof(unsigned long long) * cnt.size(), hipHostMallocCoherent));
  for (int attempt = 0; attempt < 3; ++attempt) {
    o.cand = ctx->cand.as<Cand>();
    o.counters = dc;
    o.cap = ctx->cand_cap;
    // the counters and the first candidates come back through page-locked
    // memory that refine_f64 writes (one synchronisation for the usual scan;
    // more candidates are copied after it)
    const int64_t eager = std::min<int64_t>(ctx->cand_cap, kEagerCands);
    KG_CK(ctx, ctx->hcand.ensure(sizeof(Cand) * eager, hipHostMallocCoherent));
    KG_CK(ctx, ctx->hres.ensure(sizeof(double) * 4 * eager, hipHostMallocCoherent));
    RefHost hb{};
    hb.cand = ctx->hcand.dev_as<Cand>();
    hb.res = ctx->hres.dev_as<double>();
    hb.cnt = ctx->hcnt.dev_as<unsigned long long>();
    hb.eager = eager;
    hb.ncnt = (int)cnt.size();
    // graph key: everything the captured launches depend on
    std::vector<char> key;
    if (use_graph) {
      const size_t kb = sizeof(ShardDesc) * nsh * 2 + sizeof(RefDesc) * nsh + sizeof(o) + 16 +
                        (screen ? (sizeof(ScreenDesc) + sizeof(ShardDesc)) * nsh + 32 +
                                      sizeof(ScrChunk) * chunks.size()
                                : 0) +
                        sizeof(ShardDesc) * sd3.size();
      key.resize(kb);
      char* kp = key.data();
      if (screen) {
        memcpy(kp, ss.data(), sizeof(ScreenDesc) * nsh);
        kp += sizeof(ScreenDesc) * nsh;
        memcpy(kp, sd2.data(), sizeof(ShardDesc) * nsh);
        kp += sizeof(ShardDesc) * nsh;
        // launch shape and the compaction scratch the captured kernels use
        const int64_t sx[4] = {sr.sgrid, sr.nch_max, (int64_t)(uintptr_t)sr.P,
                               (int64_t)(uintptr_t)sr.ccount};
        memcpy(kp, sx, 32);
        kp += 32;
        if (!chunks.empty()) memcpy(kp, chunks.data(), sizeof(ScrChunk) * chunks.size());
        kp += sizeof(ScrChunk) * chunks.size();
        if (!sd3.empty()) memcpy(kp, sd3.data(), sizeof(ShardDesc) * sd3.size());
        kp += sizeof(ShardDesc) * sd3.size();
      }
      memcpy(kp, sd.data(), sizeof(ShardDesc) * nsh);
      kp += sizeof(ShardDesc) * nsh;
      memcpy(kp, sd1.data(), sizeof(ShardDesc) * nsh);
      kp += sizeof(ShardDesc) * nsh;
      memcpy(kp, rd.data(), sizeof(RefDesc) * nsh);
      kp += sizeof(RefDesc) * nsh;
      memcpy(kp, &o, sizeof(o));
      kp += sizeof(o);
      const int64_t extra[2] = {(int64_t)flags,
                                (((int64_t)ctx->grid_mult * 64 + ctx->tiles) * 2 + (ctx->scr_w8 ? 1 : 0)) * 16 +
                                    sr.ring};
      memcpy(kp, extra, 16);
      key.insert(key.end(), reinterpret_cast<const char*>(&hb),
                 reinterpret_cast<const char*>(&hb) + sizeof(hb));
    }
    const bool replay = use_graph && ctx->gexec && key == ctx->graph_key;
    const ShardDesc* d_sd = ctx->sdesc.as<ShardDesc>();
    const ShardDesc* d_sd1 = ctx->sdesc1.as<ShardDesc>();
    const RefDesc* d_rd = ctx->rdesc.as<RefDesc>();
    bool zeroed = false;  // counters zeroed by the table upload
    if (use_graph) {
      if (!replay) {
        KG_CK(ctx, hipMemcpy(ctx->sdesc.p, sd.data(), sizeof(ShardDesc) * nsh, hipMemcpyHostToDevice));
        KG_CK(ctx, hipMemcpy(ctx->sdesc1.p, sd1.data(), sizeof(ShardDesc) * nsh,
                             hipMemcpyHostToDevice));
        KG_CK(ctx, hipMemcpy(ctx->rdesc.p, rd.data(), sizeof(RefDesc) * nsh, hipMemcpyHostToDevice));
        if (screen) {
          KG_CK(ctx, ctx->ssdesc.ensure(sizeof(ScreenDesc) * nsh));
          KG_CK(ctx, ctx->sdesc2.ensure(sizeof(ShardDesc) * nsh));
          KG_CK(ctx, hipMemcpy(ctx->ssdesc.p, ss.data(), sizeof(ScreenDesc) * nsh,
                               hipMemcpyHostToDevice));
          KG_CK(ctx, hipMemcpy(ctx->sdesc2.p, sd2.data(), sizeof(ShardDesc) * nsh,
                               hipMemcpyHostToDevice));
          KG_CK(ctx, ctx->chdesc.ensure(sizeof(ScrChunk) * std::max<size_t>(chunks.size(), 1)));
          if (!chunks.empty())
            KG_CK(ctx, hipMemcpy(ctx->chdesc.p, chunks.data(), sizeof(ScrChunk) * chunks.size(),
                                 hipMemcpyHostToDevice));
          if (!sd3.empty()) {
            KG_CK(ctx, ctx->sdesc3.ensure(sizeof(ShardDesc) * nsh));
            KG_CK(ctx, hipMemcpy(ctx->sdesc3.p, sd3.data(), sizeof(ShardDesc) * nsh,
                                 hipMemcpyHostToDevice));
          }
        }
      }
      if (screen) {
        sr.d_ss = ctx->ssdesc.as<ScreenDesc>();
        sr.d_sd2 = ctx->sdesc2.as<ShardDesc>();
        sr.d_ch = ctx->chdesc.as<ScrChunk>();
        if (!sd3.empty()) sr.d_sd3 = ctx->sdesc3.as<ShardDesc>();
      }
    } else {
      // descriptor tables: packed into one page-locked staging buffer and
      // copied with one asynchronous copy on the scan's stream (the previous
      // scan ended with a synchronisation, so both buffers are free); tables
      // that fit a kernel argument (a single shard's) go up by value in the
      // launch that also zeroes the counters
      auto al = [](size_t b) { return (b + 15) & ~(size_t)15; };
      const size_t b0 = sizeof(ShardDesc) * nsh, b2 = sizeof(RefDesc) * nsh;
      const size_t b3 = screen ? sizeof(ScreenDesc) * nsh : 0, b4 = screen ? b0 : 0;
      const size_t b5 = screen ? sizeof(ScrChunk) * chunks.size() : 0;
      const size_t b6 = sizeof(ShardDesc) * sd3.size();
      const size_t o1 = al(b0), o2 = o1 + al(b0), o3 = o2 + al(b2), o4 = o3 + al(b3),
                   o5 = o4 + al(b4), o6 = o5 + al(std::max<size_t>(b5, 1)), tot = o6 + al(b6);
      KG_CK(ctx, ctx->hdesc.ensure(tot));
      KG_CK(ctx, ctx->ddesc.ensure(tot));
      char* hp = static_cast<char*>(ctx->hdesc.p);
      memcpy(hp, sd.data(), b0);
      memcpy(hp + o1, sd1.data(), b0);
      memcpy(hp + o2, rd.data(), b2);
      if (screen) {
        memcpy(hp + o3, ss.data(), b3);
        memcpy(hp + o4, sd2.data(), b4);
        if (b5) memcpy(hp + o5, chunks.data(), b5);
        if (b6) memcpy(hp + o6, sd3.data(), b6);
      }
      if (tot <= setup_max_bytes() && ctx->counters_len <= (size_t)INT32_MAX) {
        KG_CK(ctx, launch_setup(ctx->stream, hp, tot, ctx->ddesc.p, dc, (int)ctx->counters_len));
        zeroed = true;
      } else {
        KG_CK(ctx, hipMemcpyAsync(ctx->ddesc.p, hp, tot, hipMemcpyHostToDevice, ctx->stream));
      }
      char* dp = static_cast<char*>(ctx->ddesc.p);
      d_sd = reinterpret_cast<const ShardDesc*>(dp);
      d_sd1 = reinterpret_cast<const ShardDesc*>(dp + o1);
      d_rd = reinterpret_cast<const RefDesc*>(dp + o2);
      if (screen) {
        sr.d_ss = reinterpret_cast<const ScreenDesc*>(dp + o3);
        sr.d_sd2 = reinterpret_cast<ShardDesc*>(dp + o4);
        sr.d_ch = reinterpret_cast<const ScrChunk*>(dp + o5);
        if (b6) sr.d_sd3 = reinterpret_cast<ShardDesc*>(dp + o6);
      }
    }
    KG_CK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    if (use_graph) {
      if (!replay) {
        drop_graph(ctx);
        KG_CK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_i8(ctx, NT, t, nsh, total_blocks, per_launch, nblk, o, ref_blocks, d_sd,
                            d_sd1, d_rd, sr, false, hb);
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        if (rc != KG_OK) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        KG_CK(ctx, ec);
        ctx->graph = g;
        KG_CK(ctx, hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0));
        ctx->graph_key = key;
      }
      KG_CK(ctx, hipGraphLaunch(ctx->gexec, ctx->stream));
      KG_CK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    } else {
      int rc = enqueue_i8(ctx, NT, t, nsh, total_blocks, per_launch, nblk, o, ref_blocks, d_sd,
                          d_sd1, d_rd, sr, zeroed, hb);
      if (rc != KG_OK) return rc;
    }
    if (overlap && attempt == 0) {
      const int orc = (*overlap)();
      if (orc != KG_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        return orc;
      }
    }
    KG_CK(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(cnt.data(), ctx->hcnt.p, sizeof(unsigned long long) * cnt.size());
    if ((int64_t)cnt[0] <= ctx->cand_cap) break;
    ctx->cand_cap = (int64_t)cnt[0] + (int64_t)cnt[0] / 4 + 1024;
    KG_CK(ctx, ctx->cand.ensure(sizeof(Cand) * ctx->cand_cap));
    KG_CK(ctx, ctx->res.ensure(sizeof(double) * 4 * ctx->cand_cap));
    drop_graph(ctx);
  }
  const int64_t nc = std::min<int64_t>((int64_t)cnt[0], ctx->cand_cap);
  float ms = 0.f;
  if (use_graph) {
    KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
    ctx->timing.bulk_ms = ms;  // graph: bulk + refine together
    ctx->timing.refine_ms = 0.0;
    ctx->timing.total_ms = ms;
    ctx->timing.screen_ms = 0.0;
  } else {
    ctx->timing.screen_ms = 0.0;
    if (screen) {
      KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ev[1], ctx->evs));
      ctx->timing.screen_ms = ms;
      KG_CK(ctx, hipEventElapsedTime(&ms, ctx->evs, ctx->ev[2]));
    } else {
      KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
    }
    ctx->timing.bulk_ms = ms;
    KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
    ctx->timing.refine_ms = ms;
    KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
    ctx->timing.total_ms = ms;
  }
  ctx->timing.candidates = nc;
  ctx->timing.kmers = mtot;
  ctx->timing.survivors = mtot;
  ctx->timing.stat_refined = 0;
  ctx->timing.rechecked = 0;
  if (rchk)
    for (int s = 0; s < nsh; ++s) ctx->timing.rechecked += (int64_t)cnt[base_rchk + s];
  if (screen) {
    int64_t ns = 0;
    for (int s = 0; s < nsh; ++s) ns += (int64_t)cnt[base_surv + s];
    ctx->timing.survivors = ns;
    // auto: when the screen plus the survivor scan took longer than the full
    // kernel over every k-mer (measured on an unscreened scan of about this
    // size, else its per-survivor time scaled to all k-mers), skip the screen
    // for the next 8 scans, then measure again
    if (auto_scr && ns > 0) {
      double full = ctx->timing.bulk_ms * ((double)mtot / (double)ns);
      if (ctx->full_m > 0 && mtot <= 2 * ctx->full_m && 2 * mtot >= ctx->full_m)
        full = ctx->full_ms * (double)mtot / (double)ctx->full_m;
      if (ctx->timing.screen_ms + ctx->timing.bulk_ms > full) ctx->screen_skip = 8;
    }
  } else if (auto_scr) {
    ctx->full_ms = ctx->timing.bulk_ms;
    ctx->full_m = mtot;
  }
  for (int s = 0; s < nsh; ++s)
    for (int tr = 0; tr < t; ++tr) n_passed[s * t + tr] += (int64_t)cnt[1 + (size_t)s * t + tr];
  if (nc == 0) return KG_OK;
  std::vector<Cand> cands(nc);
  std::vector<double> res((size_t)4 * nc);
  const int64_t ne = std::min<int64_t>(nc, std::min<int64_t>(ctx->cand_cap, kEagerCands));
  memcpy(cands.data(), ctx->hcand.p, sizeof(Cand) * ne);
  memcpy(res.data(), ctx->hres.p, sizeof(double) * 4 * ne);
  if (nc > ne) {
    KG_CK(ctx, hipMemcpy(cands.data() + ne, ctx->cand.as<Cand>() + ne, sizeof(Cand) * (nc - ne),
                         hipMemcpyDeviceToHost));
    KG_CK(ctx, hipMemcpy(res.data() + 4 * ne, ctx->res.as<double>() + 4 * ne,
                         sizeof(double) * 4 * (nc - ne), hipMemcpyDeviceToHost));
  }
  for (int64_t i = 0; i < nc; ++i) {
    const Cand& cd = cands[i];
    const HShard& s = sh[cd.shard];
    const Model& md = ctx->models[s.model]->md;
    const int c = (int)res[4 * i + 0];
    if (stats && (cd.flags & kStatRef))
      ctx->stat_patch.push_back(
          StatPatch{s.idx0 + cd.idx, cd.trait, c, res[4 * i + 1], res[4 * i + 2], res[4 * i + 3]});
    Decided d = decide(md, s.idx0 + cd.idx, cd.trait, c, res[4 * i + 1], res[4 * i + 2],
                       res[4 * i + 3], !pairs, prm, pairs);
    // a duplicate-row scan decided row idx for the cd.mult equal rows idx, idx + 1, ...
    const int mult = cd.mult > 1 ? cd.mult : 1;
    if ((cd.flags & kPfAmb) && d.pass) n_passed[cd.shard * t + cd.trait] += mult;
    if (d.hit) {
      d.h.shard = cd.shard;
      for (int q = 0; q < mult; ++q) {
        hits.push_back(d.h);
        ++d.h.idx;
      }
    }
  }
  return KG_OK;
}

// Plain FP32 path, one shard at a time.
static int scan_f32_shard(kg_ctx* ctx, const HShard& s, int sidx, const kg_scan_params& prm,
                          std::vector<kg_hit>& hits, int64_t* n_passed, bool stats,
                          const Overlap* overlap = nullptr) {
  const ModelDev& M = *ctx->models[s.model];
  const Model& md = M.md;
  const int t = md.t;
  const double p_thr = std::pow(10.0, -prm.log10_thresh);
  const double Ts = t_star(p_thr);
  const double Tmax = (double)md.n * std::log(1.0 / (1.0 - (1 - 1e-15)));
  ctx->counters_len = 1 + (size_t)kMaxTraits;
  KG_CK(ctx, ctx->counters.ensure(sizeof(unsigned long long) * ctx->counters_len));
  if (ctx->cand_cap == 0)
    ctx->cand_cap = std::max<int64_t>(1 << 16, std::min<int64_t>(s.m / 16 + 1024, 1 << 24));
  KG_CK(ctx, ctx->cand.ensure(sizeof(Cand) * ctx->cand_cap));
  ScanOut o{};
  if (stats) {
    KG_CK(ctx, ctx->st_c.ensure(sizeof(int32_t) * s.m));
    KG_CK(ctx, ctx->st_xx.ensure(sizeof(double) * s.m));
    KG_CK(ctx, ctx->st_xy.ensure(sizeof(double) * s.m * t));
    KG_CK(ctx, ctx->st_S.ensure(sizeof(double) * s.m * t));
    o.st_c = ctx->st_c.as<int32_t>();
    o.st_xx = ctx->st_xx.as<double>();
    o.st_xy = ctx->st_xy.as<double>();
    o.st_S = ctx->st_S.as<double>();
  }
  EpiF32 ep = M.epf;
  ep.tau = (float)prm.corr_threshold;
  ep.tau_trivial = prm.corr_threshold <= 0.0 ? 1 : 0;
  ep.all_sig = p_thr >= 1.0 ? 1 : 0;
  ep.Tstar_lo = (Ts > Tmax) ? INFINITY : Ts * (1.0 - 1e-9);
  const int grid = ctx->cus * ctx->grid_mult_f32;
  KG_CK(ctx,
        ctx->hcnt.ensure(sizeof(unsigned long long) * (1 + kMaxTraits), hipHostMallocCoherent));
  unsigned long long* cnt = static_cast<unsigned long long*>(ctx->hcnt.p);
  for (int attempt = 0; attempt < 3; ++attempt) {
    o.cand = ctx->cand.as<Cand>();
    o.counters = ctx->counters.as<unsigned long long>();
    o.cap = ctx->cand_cap;
    KG_CK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    KG_CK(ctx, launch_zero(ctx->stream, o.counters, 1 + kMaxTraits));
    KG_CK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    KG_CK(ctx, launch_scan_f32(M.NT, md.nch_f32, grid, ctx->stream, s.bits, s.m, s.rw, 0,
                               M.wf32.as<float>(),
                               M.mask.as<uint64_t>(), ep, o));
    KG_CK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    KG_CK(ctx, hipMemcpyAsync(cnt, o.counters, sizeof(unsigned long long) * (1 + kMaxTraits),
                              hipMemcpyDeviceToHost, ctx->stream));
    if (overlap && attempt == 0) {
      const int orc = (*overlap)();
      if (orc != KG_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        return orc;
      }
    }
    KG_CK(ctx, hipStreamSynchronize(ctx->stream));
    if ((int64_t)cnt[0] <= ctx->cand_cap) break;
    ctx->cand_cap = (int64_t)cnt[0] + (int64_t)cnt[0] / 4 + 1024;
    KG_CK(ctx, ctx->cand.ensure(sizeof(Cand) * ctx->cand_cap));
  }
  float ms = 0.f;
  KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
  ctx->timing.bulk_ms = ms;
  ctx->timing.refine_ms = 0.0;
  KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[2]));
  ctx->timing.total_ms = ms;
  const int64_t nc = std::min<int64_t>((int64_t)cnt[0], ctx->cand_cap);
  ctx->timing.candidates = nc;
  ctx->timing.kmers = s.m;
  for (int tr = 0; tr < t; ++tr) n_passed[sidx * t + tr] += (int64_t)cnt[1 + tr];
  std::vector<Cand> cands(nc);
  if (nc)
    KG_CK(ctx, hipMemcpy(cands.data(), ctx->cand.p, sizeof(Cand) * nc, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < nc; ++i) {
    const Cand& cd = cands[i];
    Decided d = decide(md, s.idx0 + cd.idx, cd.trait, cd.c, (double)cd.xx32, (double)cd.xy32, 0.0,
                       false, prm);
    if (d.hit) {
      d.h.shard = sidx;
      hits.push_back(d.h);
    }
  }
  return KG_OK;
}

static int scan_dispatch(kg_ctx* ctx, const std::vector<HShard>& sh, const kg_scan_params& prm,
                         unsigned flags, std::vector<kg_hit>& hits, int64_t* n_passed,
                         bool stats, const Overlap* overlap = nullptr) {
  if (ctx->precision == KG_PREC_MIXED)
    return scan_i8_shards(ctx, sh, prm, flags, hits, n_passed, stats, overlap);
  for (size_t i = 0; i < sh.size(); ++i) {
    int rc = scan_f32_shard(ctx, sh[i], (int)i, prm, hits, n_passed, stats,
                            i + 1 == sh.size() ? overlap : nullptr);
    if (rc) return rc;
  }
  return KG_OK;
}

static int finish_hits(kg_ctx* ctx, std::vector<kg_hit>& hits, kg_hit* out, int64_t cap,
                       int64_t* n_out) {
  std::sort(hits.begin(), hits.end(), [](const kg_hit& a, const kg_hit& b) {
    if (a.shard != b.shard) return a.shard < b.shard;
    return a.idx != b.idx ? a.idx < b.idx : a.trait < b.trait;
  });
  if (n_out) *n_out = (int64_t)hits.size();
  if ((int64_t)hits.size() > cap)
    return fail(ctx, KG_ERR_CAPACITY,
                "hit capacity " + std::to_string(cap) + " < " + std::to_string(hits.size()));
  if (!hits.empty()) {
    if (!out) return fail(ctx, KG_ERR_ARG, "null hit buffer");
    memcpy(out, hits.data(), sizeof(kg_hit) * hits.size());
  }
  return KG_OK;
}

static int check_common(kg_ctx* ctx, const kg_scan_params* prm) {
  if (!ctx) return KG_ERR_ARG;
  if (ctx->models.empty()) return fail(ctx, KG_ERR_STATE, "scan before kg_set_model");
  if (!prm) return fail(ctx, KG_ERR_ARG, "null scan params");
  if (!(prm->log10_thresh >= 0.0)) return fail(ctx, KG_ERR_ARG, "log10_thresh must be >= 0");
  return KG_OK;
}

extern "C" int kg_scan_device(kg_ctx* ctx, const uint64_t* d_bits, int64_t m, int64_t row_words,
                              const kg_scan_params* prm, kg_hit* out, int64_t cap,
                              int64_t* n_out, int64_t* n_passed) {
  int rc = check_common(ctx, prm);
  if (rc) return rc;
  HShard s{d_bits, m, row_words, 0, 0};
  rc = check_shard(ctx, s);
  if (rc) return rc;
  KG_CK(ctx, hipSetDevice(ctx->device));
  const int t = ctx->models[0]->md.t;
  std::vector<kg_hit> hits;
  std::vector<int64_t> np(t, 0);
  if (m > 0) {
    rc = scan_dispatch(ctx, {s}, *prm, 0, hits, np.data(), false);
    if (rc) return rc;
  }
  if (n_passed)
    for (int tr = 0; tr < t; ++tr) n_passed[tr] = np[tr];
  return finish_hits(ctx, hits, out, cap, n_out);
}

extern "C" int kg_scan(kg_ctx* ctx, const uint64_t* bits, int64_t m, int64_t row_words,
                       const kg_scan_params* prm, kg_hit* out, int64_t cap, int64_t* n_out,
                       int64_t* n_passed) {
  const auto wall0 = std::chrono::steady_clock::now();
  int rc = check_common(ctx, prm);
  if (rc) return rc;
  if (m > 0 && !bits) return fail(ctx, KG_ERR_ARG, "null bits");
  const Model& md = ctx->models[0]->md;
  if (row_words % 8 != 0 || row_words < 8 * (int64_t)md.nch_i8)
    return fail(ctx, KG_ERR_ARG, "row_words must be a multiple of 8 and cover all accessions");
  KG_CK(ctx, hipSetDevice(ctx->device));
  const int t = md.t;
  // 64 MiB chunks, double-buffered on the device; chunk i+1 is copied on the
  // copy stream while chunk i is scanned on the compute stream
  const size_t rowb = (size_t)row_words * 8;
  const int64_t rows = std::max<int64_t>(1, (int64_t)((64ull << 20) / rowb));
  const int64_t nchunk = (m + rows - 1) / rows;
  const size_t cb = (size_t)std::min<int64_t>(rows, std::max<int64_t>(m, 1)) * rowb;
  KG_CK(ctx, ctx->stage.ensure(cb));
  KG_CK(ctx, ctx->stage2.ensure(cb));
  uint64_t* dbuf[2] = {ctx->stage.as<uint64_t>(), ctx->stage2.as<uint64_t>()};
  // rows in page-locked memory are copied directly, pageable rows through
  // the bounce buffers
  bool pinned = false;
  if (m > 0) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, bits) == hipSuccess && attr.type == hipMemoryTypeHost)
      pinned = true;
    (void)hipGetLastError();  // a pageable pointer may leave an error behind
  }
  if (!pinned && m > 0) {
    KG_CK(ctx, ctx->bounce[0].ensure(cb));
    KG_CK(ctx, ctx->bounce[1].ensure(cb));
  }
  double copy_ms = 0.0;
  int64_t ncopied = 0;
  auto issue = [&](int64_t i) -> int {
    const int b = (int)(i & 1);
    const int64_t r0 = i * rows, mm = std::min<int64_t>(rows, m - r0);
    const size_t bytes = (size_t)mm * rowb;
    const void* src = bits + r0 * row_words;
    if (!pinned) {
      // the bounce buffer's previous copy (chunk i-2) must be done
      KG_CK(ctx, hipEventSynchronize(ctx->cev[b]));
      memcpy(ctx->bounce[b].p, src, bytes);
      src = ctx->bounce[b].p;
    }
    KG_CK(ctx, hipEventRecord(ctx->ctev[2 * b], ctx->cstream));
    KG_CK(ctx, hipMemcpyAsync(dbuf[b], src, bytes, hipMemcpyHostToDevice, ctx->cstream));
    KG_CK(ctx, hipEventRecord(ctx->ctev[2 * b + 1], ctx->cstream));
    KG_CK(ctx, hipEventRecord(ctx->cev[b], ctx->cstream));
    ++ncopied;
    return KG_OK;
  };
  auto copy_time = [&](int b) -> int {
    float ms = 0.f;
    KG_CK(ctx, hipEventSynchronize(ctx->ctev[2 * b + 1]));
    KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ctev[2 * b], ctx->ctev[2 * b + 1]));
    copy_ms += ms;
    return KG_OK;
  };
  std::vector<kg_hit> hits;
  std::vector<int64_t> np(t, 0);
  kg_timing acc{};
  if (nchunk > 0) {
    rc = issue(0);
    if (rc) return rc;
  }
  for (int64_t i = 0; i < nchunk; ++i) {
    const int b = (int)(i & 1);
    const int64_t r0 = i * rows, mm = std::min<int64_t>(rows, m - r0);
    KG_CK(ctx, hipStreamWaitEvent(ctx->stream, ctx->cev[b], 0));
    const Overlap next = [&]() -> int { return (i + 1 < nchunk) ? issue(i + 1) : KG_OK; };
    HShard s{dbuf[b], mm, row_words, r0, 0};
    rc = scan_dispatch(ctx, {s}, *prm, 0, hits, np.data(), false, &next);
    if (rc) {
      (void)hipStreamSynchronize(ctx->cstream);
      return rc;
    }
    rc = copy_time(b);
    if (rc) return rc;
    acc.bulk_ms += ctx->timing.bulk_ms;
    acc.refine_ms += ctx->timing.refine_ms;
    acc.total_ms += ctx->timing.total_ms;
    acc.candidates += ctx->timing.candidates;
    acc.screen_ms += ctx->timing.screen_ms;
    acc.survivors += ctx->timing.survivors;
    acc.rechecked += ctx->timing.rechecked;
  }
  KG_CK(ctx, hipStreamSynchronize(ctx->cstream));
  acc.kmers = m;
  acc.copy_ms = copy_ms;
  acc.wall_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
  ctx->timing = acc;
  if (n_passed)
    for (int tr = 0; tr < t; ++tr) n_passed[tr] = np[tr];
  return finish_hits(ctx, hits, out, cap, n_out);
}

// ---------------------------------------------------------------------------
// Allocation-free scan: kg_scan_device on the GPU's scratch array (counters,
// descriptors, candidates, FP64 results) and on host staging arrays of fixed
// size, with no device or host allocation.  Unscreened, single shard, model 0.
// The calls of a process are serialised by one lock: they share the host
// staging, and calls on one GPU share its scratch array (which keeps the last
// call's contents; the counters are cleared by each call's first kernel).
namespace {
std::mutex g_noalloc_mu;
Cand g_na_cand[kNoAllocCands];
double g_na_res[4 * kNoAllocCands];
kg_hit g_na_hits[kNoAllocCands];
unsigned long long g_na_cnt[16];
}  // namespace

extern "C" int kg_scan_device_noalloc(kg_ctx* ctx, const uint64_t* d_bits, int64_t m,
                                      int64_t row_words, const kg_scan_params* prm, kg_hit* out,
                                      int64_t cap, int64_t* n_out, int64_t* n_passed) {
  int rc = check_common(ctx, prm);
  if (rc) return rc;
  if (ctx->precision != KG_PREC_MIXED)
    return fail(ctx, KG_ERR_ARG, "kg_scan_device_noalloc needs KG_PREC_MIXED");
  const HShard sh{d_bits, m, row_words, 0, 0};
  rc = check_shard(ctx, sh);
  if (rc) return rc;
  if (!ctx->scratch) return fail(ctx, KG_ERR_STATE, "no scratch array on this device");
  const ModelDev& M = *ctx->models[0];
  const Model& md = M.md;
  const int t = md.t;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  KG_CK(ctx, hipSetDevice(ctx->device));
  if (n_out) *n_out = 0;
  int64_t np[kMaxTraits] = {};
  int64_t nh = 0;
  if (m > 0) {
    char* base = static_cast<char*>(ctx->scratch);
    auto* dc = reinterpret_cast<unsigned long long*>(base + kNoAllocCounterOff);
    auto* d_cand = reinterpret_cast<Cand*>(base + kNoAllocCandOff);
    auto* d_res = reinterpret_cast<double*>(base + kNoAllocResOff);
    const double p_thr = std::pow(10.0, -prm->log10_thresh);
    const double Ts = t_star(p_thr);
    // descriptors go up as the argument of the kernel that also clears the counters
    struct {
      ShardDesc D;
      RefDesc R;
    } desc{};
    fill_descs(ctx, sh, 0, *prm, Ts, p_thr, false, false, dc + 1, 0, desc.D, desc.R);
    static_assert(sizeof(desc) % 8 == 0, "descriptor upload in words");
    auto* d_desc = base + kNoAllocDescOff;
    const int64_t bk = i8_block_kmers(M.NT, ctx->limbs, t, ctx->tiles);
    const int64_t blocks = (m + bk - 1) / bk;
    ScanOut o{};
    o.cand = d_cand;
    o.counters = dc;
    o.cap = kNoAllocCands;
    KG_CK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    KG_CK(ctx, launch_setup(ctx->stream, &desc, sizeof(desc), d_desc, dc, 1 + t));
    KG_CK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    KG_CK(ctx, launch_scan_i8(M.NT, ctx->limbs, t, ctx->tiles, ctx->cus * ctx->grid_mult,
                              ctx->stream, reinterpret_cast<const ShardDesc*>(d_desc), 1,
                              blocks, o, nullptr));
    KG_CK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    KG_CK(ctx, launch_refine(ctx->stream, d_cand, dc, kNoAllocCands,
                             reinterpret_cast<const RefDesc*>(d_desc + sizeof(ShardDesc)), d_res,
                             ctx->cus * 2, RefHost{}));
    KG_CK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    KG_CK(ctx, hipMemcpyAsync(g_na_cnt, dc, sizeof(unsigned long long) * (1 + t),
                              hipMemcpyDeviceToHost, ctx->stream));
    KG_CK(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t nc = (int64_t)g_na_cnt[0];
    if (nc > kNoAllocCands)
      return fail(ctx, KG_ERR_SCRATCH,
                  "kg_scan_device_noalloc: " + std::to_string(nc) + " candidates exceed its " +
                      std::to_string(kNoAllocCands) + " (use kg_scan_device)");
    if (nc > 0) {
      KG_CK(ctx, hipMemcpyAsync(g_na_cand, d_cand, sizeof(Cand) * nc, hipMemcpyDeviceToHost,
                                ctx->stream));
      KG_CK(ctx, hipMemcpyAsync(g_na_res, d_res, sizeof(double) * 4 * nc, hipMemcpyDeviceToHost,
                                ctx->stream));
      KG_CK(ctx, hipStreamSynchronize(ctx->stream));
    }
    for (int tr = 0; tr < t; ++tr) np[tr] = (int64_t)g_na_cnt[1 + tr];
    for (int64_t i = 0; i < nc; ++i) {
      const Cand& cd = g_na_cand[i];
      Decided d = decide(md, cd.idx, cd.trait, (int)g_na_res[4 * i], g_na_res[4 * i + 1],
                         g_na_res[4 * i + 2], g_na_res[4 * i + 3], true, *prm);
      if ((cd.flags & kPfAmb) && d.pass) ++np[cd.trait];
      if (d.hit) {
        d.h.shard = 0;
        g_na_hits[nh++] = d.h;
      }
    }
    std::sort(g_na_hits, g_na_hits + nh, [](const kg_hit& a, const kg_hit& b) {
      return a.idx != b.idx ? a.idx < b.idx : a.trait < b.trait;
    });
    float ms = 0.f;
    KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
    ctx->timing = kg_timing{};
    ctx->timing.bulk_ms = ms;
    KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
    ctx->timing.refine_ms = ms;
    KG_CK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
    ctx->timing.total_ms = ms;
    ctx->timing.candidates = nc;
    ctx->timing.kmers = m;
    ctx->timing.survivors = m;
  }
  if (n_passed)
    for (int tr = 0; tr < t; ++tr) n_passed[tr] = np[tr];
  if (n_out) *n_out = nh;
  if (nh > cap)
    return fail(ctx, KG_ERR_CAPACITY,
                "hit capacity " + std::to_string(cap) + " < " + std::to_string(nh));
  if (nh > 0) {
    if (!out) return fail(ctx, KG_ERR_ARG, "null hit buffer");
    memcpy(out, g_na_hits, sizeof(kg_hit) * (size_t)nh);
  }
  return KG_OK;
}

extern "C" int kg_host_alloc(kg_ctx* ctx, size_t bytes, void** h_ptr) {
  if (!ctx || !h_ptr) return KG_ERR_ARG;
  *h_ptr = nullptr;
  KG_CK(ctx, hipSetDevice(ctx->device));
  KG_CK(ctx, hipHostMalloc(h_ptr, bytes ? bytes : 16, hipHostMallocDefault));
  return KG_OK;
}

extern "C" int kg_host_free(kg_ctx* ctx, void* h_ptr) {
  if (!ctx) return KG_ERR_ARG;
  if (!h_ptr) return KG_OK;
  KG_CK(ctx, hipHostFree(h_ptr));
  return KG_OK;
}

extern "C" int kg_scan_shards(kg_ctx* ctx, const kg_shard* shards, int32_t nshards,
                              const kg_scan_params* prm, uint32_t flags, kg_hit* out,
                              int64_t cap, int64_t* n_out, int64_t* n_passed) {
  int rc = check_common(ctx, prm);
  if (rc) return rc;
  if (nshards <= 0 || !shards) return fail(ctx, KG_ERR_ARG, "no shards");
  std::vector<HShard> sh(nshards);
  for (int s = 0; s < nshards; ++s) {
    sh[s] = HShard{shards[s].d_bits, shards[s].m, shards[s].row_words, 0, shards[s].model};
    rc = check_shard(ctx, sh[s]);
    if (rc) return rc;
  }
  KG_CK(ctx, hipSetDevice(ctx->device));
  const int t = ctx->models[sh[0].model]->md.t;
  std::vector<int64_t> np((size_t)nshards * t, 0);
  std::vector<kg_hit> hits;
  unsigned f = 0;
  if (flags & KG_SHARDS_PER_LAUNCH) f |= kPerLaunch;
  if (flags & KG_SHARDS_GRAPH) f |= kGraph;
  rc = scan_dispatch(ctx, sh, *prm, f, hits, np.data(), false);
  if (rc) return rc;
  if (n_passed) memcpy(n_passed, np.data(), sizeof(int64_t) * np.size());
  return finish_hits(ctx, hits, out, cap, n_out);
}

// All-pairs epistasis scan (src/kgwas_2d.cpp:258-401, emit_pairs_self): every
// row pair i < j with i in [i0, i1) is tested as the k-mer x_i AND x_j, with
// no pre-filter (src/kgwas_2d.cpp:362-388).  Shard k of a batch is row
// i = bi + k: rows i+1 .. N-1 under the accession mask (model mask AND row
// i), written on the device, so the bulk kernel, the FP64 refinement and the
// host decision run unchanged.  Batches of consecutive rows bound the shard
// tables and the mask buffer.
constexpr int64_t kPairBatch = int64_t(1) << 28;  // pairs per batch (soft)
constexpr int64_t kPairShards = 32768;            // rows i per batch

extern "C" int kg_scan_pairs(kg_ctx* ctx, const uint64_t* d_bits, int64_t N, int64_t row_words,
                             int64_t i0, int64_t i1, const kg_scan_params* prm, kg_hit* out,
                             int64_t cap, int64_t* n_out) {
  int rc = check_common(ctx, prm);
  if (rc) return rc;
  if (ctx->precision != KG_PREC_MIXED)
    return fail(ctx, KG_ERR_ARG, "kg_scan_pairs needs KG_PREC_MIXED");
  if (N < 0) return fail(ctx, KG_ERR_ARG, "negative row count");
  if (i0 < 0 || i1 < i0 || i1 > N)
    return fail(ctx, KG_ERR_ARG, "row range [i0, i1) must lie in [0, N]");
  rc = check_shard(ctx, HShard{d_bits, N, row_words, 0, 0});
  if (rc) return rc;
  KG_CK(ctx, hipSetDevice(ctx->device));
  const ModelDev& M = *ctx->models[0];
  const int t = M.md.t;
  const int64_t W = 8 * (int64_t)M.md.nch_i8;  // mask words per row i
  kg_scan_params pp = *prm;
  pp.corr_threshold = 0.0;
  pp.log10_thresh_2d = 0.0;
  std::vector<kg_hit> hits;
  kg_timing acc{};
  const int64_t last = std::min<int64_t>(i1, N - 1);  // row N-1 has no partner j > i
  for (int64_t bi = i0; bi < last;) {
    int64_t be = bi, np_ = 0;
    while (be < last && be - bi < kPairShards &&
           (np_ == 0 || np_ + (N - 1 - be) <= kPairBatch)) {
      np_ += N - 1 - be;
      ++be;
    }
    const int64_t cnt = be - bi;
    KG_CK(ctx, ctx->pmask.ensure(sizeof(uint64_t) * (size_t)(cnt * W)));
    uint64_t* pm = ctx->pmask.as<uint64_t>();
    KG_CK(ctx, launch_pair_masks(ctx->stream, d_bits, row_words, M.mask.as<uint64_t>(), W, bi,
                                 cnt, pm));
    std::vector<HShard> sh((size_t)cnt);
    for (int64_t k = 0; k < cnt; ++k) {
      const int64_t i = bi + k;
      sh[k] = HShard{d_bits + (i + 1) * row_words, N - 1 - i, row_words, i + 1, 0};
      sh[k].mask = pm + k * W;
    }
    std::vector<int64_t> passes((size_t)cnt * t, 0);
    std::vector<kg_hit> bh;
    rc = scan_i8_shards(ctx, sh, pp, kPairs, bh, passes.data(), false);
    if (rc) return rc;
    for (kg_hit& h : bh) {
      h.shard = (int32_t)(bi + h.shard);  // row i; h.idx is row j
      hits.push_back(h);
    }
    acc.bulk_ms += ctx->timing.bulk_ms;
    acc.refine_ms += ctx->timing.refine_ms;
    acc.total_ms += ctx->timing.total_ms;
    acc.candidates += ctx->timing.candidates;
    acc.kmers += ctx->timing.kmers;
    bi = be;
  }
  acc.survivors = acc.kmers;
  ctx->timing = acc;
  return finish_hits(ctx, hits, out, cap, n_out);
}

static void fill_tune(const kg_ctx* ctx, kg_tune* out) {
  *out = kg_tune{};
  out->tiles_per_wave = ctx->tiles;
  out->workgroups_per_cu = ctx->grid_mult;
  out->screen_wgs_per_cu = ctx->scr_w8 ? 2 : 1;
  int ring = 0;
  if (!ctx->models.empty() && ctx->precision == KG_PREC_MIXED) {
    const ModelDev& M = *ctx->models[0];
    out->tiles_per_wave = i8_tiles_for(M.NT, ctx->limbs, M.md.t, ctx->tiles);
    out->kmers_per_workgroup = (int32_t)i8_block_kmers(M.NT, ctx->limbs, M.md.t, ctx->tiles);
    // the shape model 0's screened scans use: the ring where it fits
    ring = (M.md.t == 1 && ctx->scr_ring && screen_ring_fits(M.md.nch_i8, ctx->scr_ring))
               ? ctx->scr_ring : 0;
    out->screen_kmers_per_workgroup =
        (int32_t)(ring ? 64 * ring : screen_block_kmers(M.md.t, ctx->scr_w8));
    out->resident_active =
        (ctx->resident && M.screen_ok && (M.res_ok || (M.f6_ok && M.f6_digits == ctx->res_digits)))
            ? 1 : 0;
  }
  out->stream_weights = ctx->resident ? 0 : 1;
  if (ring) out->screen_wgs_per_cu = ring == 8 ? 3 : 4;
  out->resident_form = ctx->res_digits == 0 ? 1 : ctx->res_digits;
}

extern "C" int kg_get_tune(const kg_ctx* ctx, kg_tune* out) {
  if (!ctx || !out) return KG_ERR_ARG;
  fill_tune(ctx, out);
  return KG_OK;
}

extern "C" int kg_set_tune(kg_ctx* ctx, const kg_tune* shape) {
  if (!ctx || !shape) return KG_ERR_ARG;
  if (shape->tiles_per_wave != 0 && shape->tiles_per_wave != 2 && shape->tiles_per_wave != 4)
    return fail(ctx, KG_ERR_ARG, "tiles_per_wave must be 0, 2 or 4");
  if (shape->workgroups_per_cu < 1 || shape->workgroups_per_cu > 64)
    return fail(ctx, KG_ERR_ARG, "workgroups_per_cu must be 1-64");
  if (shape->screen_wgs_per_cu < 0 || shape->screen_wgs_per_cu > 4)
    return fail(ctx, KG_ERR_ARG, "screen_wgs_per_cu must be 0-4");
  if (shape->stream_weights < 0 || shape->stream_weights > 1)
    return fail(ctx, KG_ERR_ARG, "stream_weights must be 0 or 1");
  if (shape->resident_form < 0 || shape->resident_form > 3)
    return fail(ctx, KG_ERR_ARG, "resident_form must be 0-3");
  const int digits = shape->resident_form == 1 ? 0 : (shape->resident_form == 3 ? 3 : 2);
  if (digits != ctx->res_digits) {
    KG_CK(ctx, hipSetDevice(ctx->device));
    ctx->res_digits = digits;
    if (ctx->precision == KG_PREC_MIXED)
      for (auto& M : ctx->models) {
        const int rc = build_f6(ctx, *M);
        if (rc != KG_OK) return rc;
      }
  }
  // 4 tiles is the default wherever it is available; asking for it elsewhere
  // keeps the default (2)
  ctx->tiles = shape->tiles_per_wave == 2 ? 2 : 0;
  ctx->grid_mult = shape->workgroups_per_cu;
  ctx->scr_w8 = shape->screen_wgs_per_cu == 2;
  // 0: the default (the 8-wave ring where it fits, else one 16-wave workgroup)
  ctx->scr_ring = shape->screen_wgs_per_cu == 4 ? 12
                  : (shape->screen_wgs_per_cu == 0 || shape->screen_wgs_per_cu == 3) ? 8 : 0;
  ctx->resident = shape->stream_weights == 0;
  drop_graph(ctx);
  return KG_OK;
}

// Autotuner: times the bulk scan of `sample` synthetic rows (random bits)
// with model 0 for each candidate (k-mer tiles per wave x persistent
// workgroups per CU for the full kernel, then the screen's workgroups per CU)
// and keeps the fastest for this context.
extern "C" int kg_autotune(kg_ctx* ctx, int64_t sample, kg_tune* out) {
  if (!ctx) return KG_ERR_ARG;
  if (ctx->models.empty()) return fail(ctx, KG_ERR_STATE, "autotune before kg_set_model");
  if (sample <= 0) return fail(ctx, KG_ERR_ARG, "autotune sample must be positive");
  KG_CK(ctx, hipSetDevice(ctx->device));
  const ModelDev& M = *ctx->models[0];
  const Model& md = M.md;
  if (ctx->precision == KG_PREC_MIXED) {
    const int64_t rw = 8 * (int64_t)md.nch_i8;
    DBuf rows, pop, thr;
    KG_CK(ctx, rows.ensure((size_t)sample * rw * 8));
    KG_CK(ctx, pop.ensure((size_t)md.n_acc));
    KG_CK(ctx, thr.ensure(sizeof(uint32_t)));
    KG_CK(ctx, hipMemset(pop.p, 0, (size_t)md.n_acc));
    const uint32_t half = 32768u;  // every bit set with probability 1/2 (16-bit threshold)
    KG_CK(ctx, hipMemcpy(thr.p, &half, sizeof(half), hipMemcpyHostToDevice));
    KG_CK(ctx, launch_synth(ctx->stream, rows.as<uint64_t>(), 12345, 0, sample, (uint32_t)md.n_acc,
                            rw, pop.as<uint8_t>(), thr.as<uint32_t>(), 1, 1));
    KG_CK(ctx, hipStreamSynchronize(ctx->stream));
    const kg_scan_params prm{0.2, 6.0, 0.0};
    const int keep_tiles = ctx->tiles, keep_gm = ctx->grid_mult, keep_mode = ctx->screen_mode,
              keep_skip = ctx->screen_skip;
    const bool keep_w8 = ctx->scr_w8;
    const int keep_ring = ctx->scr_ring;
    auto restore = [&]() {
      ctx->tiles = keep_tiles;
      ctx->grid_mult = keep_gm;
      ctx->scr_w8 = keep_w8;
      ctx->scr_ring = keep_ring;
      ctx->screen_mode = keep_mode;
      ctx->screen_skip = keep_skip;
    };
    // best of 3 timed scans after one warm-up; `screen` selects the timing read
    auto time_scan = [&](bool screen, double* best_ms) -> int {
      *best_ms = 1e300;
      for (int rep = 0; rep < 4; ++rep) {
        std::vector<kg_hit> hits;
        std::vector<int64_t> np(md.t, 0);
        HShard s{rows.as<uint64_t>(), sample, rw, 0, 0};
        const int rc = scan_dispatch(ctx, {s}, prm, 0, hits, np.data(), false);
        if (rc) return rc;
        const double ms = screen ? ctx->timing.screen_ms : ctx->timing.bulk_ms;
        if (rep > 0) *best_ms = std::min(*best_ms, ms);
      }
      return KG_OK;
    };
    // the full kernel's shape, on every row (random rows fail the screen)
    const int R_def = i8_tiles_for(M.NT, ctx->limbs, md.t, 0);
    std::vector<int> Rs{0};
    if (i8_tiles_for(M.NT, ctx->limbs, md.t, 2) != R_def) Rs.push_back(2);
    const int gms[3] = {1, 2, 4};
    ctx->screen_mode = KG_SCREEN_OFF;
    double best = 1e300;
    int best_R = keep_tiles, best_gm = keep_gm;
    for (int R : Rs)
      for (int gm : gms) {
        ctx->tiles = R;
        ctx->grid_mult = gm;
        double tmin = 0.0;
        const int rc = time_scan(false, &tmin);
        if (rc) {
          restore();
          return rc;
        }
        // a shape replaces the earlier (default first) one only when it is
        // faster by more than the run-to-run spread of these short scans
        if (tmin < best * 0.98) {
          best = tmin;
          best_R = R;
          best_gm = gm;
        }
      }
    ctx->tiles = best_R;
    ctx->grid_mult = best_gm;
    // the screen's shape (where the model has a screen): the default first
    // (the 8-wave LDS ring where it fits), then the register-loaded shapes
    // (one 16-wave or two 8-wave workgroups per CU) and the 12-wave ring
    bool best_w8 = keep_w8;
    int best_ring = keep_ring;
    if (M.screen_ok) {
      ctx->screen_mode = KG_SCREEN_ON;
      double sbest = 1e300;
      std::vector<std::pair<bool, int>> shapes;  // (w8, ring waves)
      const bool ring_ok = md.t == 1 && screen_ring_fits(md.nch_i8, 8);
      if (ring_ok) shapes.push_back({false, 8});
      shapes.push_back({false, 0});
      shapes.push_back({true, 0});
      if (md.t == 1 && screen_ring_fits(md.nch_i8, 12)) shapes.push_back({false, 12});
      for (const auto& sh : shapes) {
        ctx->scr_w8 = sh.first;
        ctx->scr_ring = sh.second;
        double tmin = 0.0;
        const int rc = time_scan(true, &tmin);
        if (rc) {
          restore();
          return rc;
        }
        if (tmin < sbest * 0.98) {
          sbest = tmin;
          best_w8 = sh.first;
          best_ring = sh.second;
        }
      }
      // a register-loaded shape keeps the ring setting off only when it won
      if (best_ring == 0 && !ring_ok) best_ring = keep_ring;
    }
    ctx->scr_w8 = best_w8;
    ctx->scr_ring = best_ring;
    ctx->screen_mode = keep_mode;
    ctx->screen_skip = keep_skip;
    drop_graph(ctx);
  }
  if (out) fill_tune(ctx, out);
  return KG_OK;
}

extern "C" int kg_stats_device(kg_ctx* ctx, const uint64_t* d_bits, int64_t m,
                               int64_t row_words, int32_t* c, double* xxM, double* xyM,
                               double* S) {
  kg_scan_params prm{0.2, 6.0, 0.0};
  int rc = check_common(ctx, &prm);
  if (rc) return rc;
  HShard s{d_bits, m, row_words, 0, 0};
  rc = check_shard(ctx, s);
  if (rc) return rc;
  if (m == 0) return KG_OK;
  if (!c || !xxM || !xyM || !S) return fail(ctx, KG_ERR_ARG, "null output array");
  KG_CK(ctx, hipSetDevice(ctx->device));
  const int t = ctx->models[0]->md.t;
  std::vector<kg_hit> hits;
  std::vector<int64_t> np(t, 0);
  rc = scan_dispatch(ctx, {s}, prm, 0, hits, np.data(), true);
  if (rc) return rc;
  KG_CK(ctx, hipMemcpy(c, ctx->st_c.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  KG_CK(ctx, hipMemcpy(xxM, ctx->st_xx.p, sizeof(double) * m, hipMemcpyDeviceToHost));
  KG_CK(ctx, hipMemcpy(xyM, ctx->st_xy.p, sizeof(double) * m * t, hipMemcpyDeviceToHost));
  KG_CK(ctx, hipMemcpy(S, ctx->st_S.p, sizeof(double) * m * t, hipMemcpyDeviceToHost));
  // the mixed path's k-mers whose bulk values risk more than the error budget
  // were recomputed in FP64 (refine_f64)
  for (const StatPatch& q : ctx->stat_patch) {
    xxM[q.idx] = q.xxM;
    xyM[q.idx * t + q.trait] = q.xyM;
    S[q.idx * t + q.trait] = q.S;
  }
  ctx->timing.stat_refined = (int64_t)ctx->stat_patch.size();
  ctx->stat_patch.clear();
  return KG_OK;
}

extern "C" int kg_set_screen(kg_ctx* ctx, int32_t mode) {
  if (!ctx) return KG_ERR_ARG;
  if (mode < KG_SCREEN_AUTO || mode > KG_SCREEN_ON)
    return fail(ctx, KG_ERR_ARG, "screen mode must be KG_SCREEN_AUTO, _OFF or _ON");
  ctx->screen_mode = mode;
  ctx->screen_skip = 0;
  ctx->full_m = 0;
  ctx->probe_tick = 0;
  return KG_OK;
}

extern "C" int kg_set_dedup(kg_ctx* ctx, int32_t on) {
  if (!ctx) return KG_ERR_ARG;
  if (on != 0 && on != 1) return fail(ctx, KG_ERR_ARG, "dedup must be 0 or 1");
  ctx->dedup = on;
  return KG_OK;
}

extern "C" int kg_last_timing(const kg_ctx* ctx, kg_timing* out) {
  if (!ctx || !out) return KG_ERR_ARG;
  *out = ctx->timing;
  return KG_OK;
}

extern "C" int kg_synth_bits(kg_ctx* ctx, uint64_t* d_out, uint64_t seed, int64_t kmer0,
                             int64_t m, int32_t n_acc, int64_t row_words, const uint8_t* pop,
                             const uint32_t* thr, int32_t npop, int32_t n_profiles) {
  if (!ctx || !ctx->stream) return KG_ERR_ARG;
  if (m < 0 || n_acc <= 0 || row_words * 64 < n_acc || npop <= 0 || n_profiles <= 0 || !pop ||
      !thr || (m > 0 && !d_out))
    return fail(ctx, KG_ERR_ARG, "kg_synth_bits: bad arguments");
  for (int i = 0; i < n_acc; ++i)
    if (pop[i] >= npop) return fail(ctx, KG_ERR_ARG, "kg_synth_bits: population out of range");
  KG_CK(ctx, hipSetDevice(ctx->device));
  DBuf dpop, dthr;
  KG_CK(ctx, dpop.ensure(n_acc));
  KG_CK(ctx, dthr.ensure((size_t)4 * npop * n_profiles));
  KG_CK(ctx, hipMemcpy(dpop.p, pop, n_acc, hipMemcpyHostToDevice));
  KG_CK(ctx, hipMemcpy(dthr.p, thr, (size_t)4 * npop * n_profiles, hipMemcpyHostToDevice));
  if (m > 0)
    KG_CK(ctx, launch_synth(ctx->stream, d_out, seed, kmer0, m, (uint32_t)n_acc, row_words,
                            dpop.as<uint8_t>(), dthr.as<uint32_t>(), (uint32_t)npop,
                            (uint32_t)n_profiles));
  KG_CK(ctx, hipStreamSynchronize(ctx->stream));
  return KG_OK;
}

extern "C" int kg_dev_alloc(kg_ctx* ctx, size_t bytes, void** d_ptr) {
  if (!ctx || !ctx->stream || !d_ptr) return KG_ERR_ARG;
  KG_CK(ctx, hipSetDevice(ctx->device));
  KG_CK(ctx, hipMalloc(d_ptr, bytes ? bytes : 16));
  return KG_OK;
}

extern "C" int kg_dev_free(kg_ctx* ctx, void* d_ptr) {
  if (!ctx || !ctx->stream) return KG_ERR_ARG;
  if (!d_ptr) return KG_OK;
  KG_CK(ctx, hipSetDevice(ctx->device));
  KG_CK(ctx, hipStreamSynchronize(ctx->stream));
  KG_CK(ctx, hipFree(d_ptr));
  return KG_OK;
}

extern "C" int kg_memcpy(kg_ctx* ctx, void* dst, const void* src, size_t bytes,
                         int32_t direction) {
  if (!ctx || !ctx->stream) return KG_ERR_ARG;
  if (bytes == 0) return KG_OK;
  if (!dst || !src) return fail(ctx, KG_ERR_ARG, "kg_memcpy: null pointer");
  hipMemcpyKind k;
  switch (direction) {
    case 0: k = hipMemcpyHostToDevice; break;
    case 1: k = hipMemcpyDeviceToHost; break;
    case 2: k = hipMemcpyDeviceToDevice; break;
    default: return fail(ctx, KG_ERR_ARG, "kg_memcpy: direction must be 0, 1 or 2");
  }
  KG_CK(ctx, hipSetDevice(ctx->device));
  KG_CK(ctx, hipStreamSynchronize(ctx->stream));
  KG_CK(ctx, hipMemcpy(dst, src, bytes, k));
  return KG_OK;
}
