// C ABI of libkgwas_hip: null-model set-up in FP64, weight quantisation and
// layout, scan orchestration on one HIP stream, and the final per-k-mer
// decisions with the reference's own compute_hit arithmetic.
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
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/kgwas_hip.h"
#include "kg_internal.h"

namespace kg {
hipError_t launch_scan_i8(int NT, int R, int grid, hipStream_t s, const uint64_t* bits,
                          int64_t m, int64_t rw, int64_t row0, const void* wts, const void* mask,
                          const EpiI8& ep, const ScanOut& o);
hipError_t launch_scan_f32(int NT, int grid, hipStream_t s, const uint64_t* bits, int64_t m,
                           int64_t rw, int64_t row0, const float* wts, const uint64_t* mask,
                           const EpiF32& ep, const ScanOut& o);
hipError_t launch_refine(hipStream_t s, const Cand* cand, int64_t ncand, const uint64_t* bits,
                         int64_t rw, int64_t row0, const uint64_t* maskw, int nwords,
                         const double* B64, int nbf, int p, int t, int n, double sigma,
                         double base_coef, const double* bpfull, const double* rbar,
                         double* res);
hipError_t launch_zero(hipStream_t s, unsigned long long* p, int n);
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
  int nch_i8 = 0, nch_f32 = 0;
};

}  // namespace

struct kg_ctx {
  int device = 0;
  int precision = KG_PREC_MIXED;
  int limbs = 3;
  int cus = 256;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::string err;
  bool have_model = false;
  Model md;
  // i8 path
  int NT = 0, nz = 0, nb = 0;
  DBuf wts, mask, colw, colh, bp, bp0, rbar, ycn, yyM;
  DBuf B64, bpfull;
  // f32 path
  DBuf wf32, uyf;
  EpiF32 epf{};
  EpiI8 epi{};
  // scratch
  DBuf counters, cand, res, stage, st_c, st_xx, st_xy, st_S;
  int64_t cand_cap = 0;
  kg_timing timing{};
  int R = 4;
  int grid_mult = 8;
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
static void quantize(const double* col, int n, int L, std::vector<int8_t>& limbs, double* w,
                     double* h) {
  double scale = 0.0;
  for (int j = 0; j < n; ++j) scale = std::max(scale, std::fabs(col[j]));
  limbs.assign((size_t)n * L, 0);
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
}

static int build_i8(kg_ctx* ctx) {
  Model& md = ctx->md;
  const int n = md.n, p = md.p, t = md.t, L = ctx->limbs;
  std::vector<int> zc;
  for (int k = 0; k < p; ++k)
    if (!(k == 0 && md.rotated && md.drop_d)) zc.push_back(k);
  const int nz = (int)zc.size();
  const int nb = nz + t;
  const int NU = ((L * nb + 15) / 16) * 16;
  const int NT = NU / 16;
  if (NT > 8)
    return fail(ctx, KG_ERR_ARG,
                "weight columns exceed the int8 kernel's 128-column tile budget (p + t too "
                "large for this limb count)");
  ctx->NT = NT;
  ctx->nz = nz;
  ctx->nb = nb;
  std::vector<int8_t> lim;
  std::vector<double> colw((size_t)nb * L), colh(nb);
  std::vector<int8_t> ucol((size_t)n * NU, 0);  // [sample][u-col]
  for (int b = 0; b < nb; ++b) {
    const double* src = (b < nz) ? &md.Uc[(size_t)zc[b] * n] : &md.rc[(size_t)(b - nz) * n];
    quantize(src, n, L, lim, &colw[(size_t)b * L], &colh[b]);
    for (int j = 0; j < n; ++j)
      for (int l = 0; l < L; ++l) ucol[(size_t)j * NU + L * b + l] = lim[(size_t)j * L + l];
  }
  // accession -> sample
  std::vector<int> samp(md.n_acc, -1);
  for (int j = 0; j < n; ++j) samp[md.idxs[j]] = j;
  const int nch = md.nch_i8;
  const size_t frag = (size_t)8 * NT * 64 * 16;  // bytes per chunk
  std::vector<int8_t> W((size_t)nch * frag, 0);
  for (int C = 0; C < nch; ++C)
    for (int jj = 0; jj < 8; ++jj)
      for (int tt = 0; tt < NT; ++tt)
        for (int lane = 0; lane < 64; ++lane) {
          const int i = lane & 15, g = lane >> 4;
          const int uc = 16 * tt + i;
          int8_t* dst = &W[(size_t)C * frag + (((size_t)jj * NT + tt) * 64 + lane) * 16];
          for (int e = 0; e < 16; ++e) {
            const int a = 512 * C + 128 * g + 16 * jj + e;
            if (a < md.n_acc && samp[a] >= 0) dst[e] = ucol[(size_t)samp[a] * NU + uc];
          }
        }
  std::vector<double> bp((size_t)t * (nz > 0 ? nz : 1), 0.0), bp0(t, 0.0);
  for (int tr = 0; tr < t; ++tr) {
    for (int b = 0; b < nz; ++b) bp[(size_t)tr * nz + b] = md.bprime[(size_t)tr * p + zc[b]];
    if (md.rotated && md.drop_d) bp0[tr] = md.bprime[(size_t)tr * p + 0];
  }
  // FP64 refinement columns: all p rotated-centred columns + t residuals, by accession
  const int nbf = p + t;
  std::vector<double> B((size_t)md.n_acc * nbf, 0.0);
  for (int j = 0; j < n; ++j) {
    const int a = md.idxs[j];
    for (int k = 0; k < p; ++k) B[(size_t)a * nbf + k] = md.Uc[(size_t)k * n + j];
    for (int tr = 0; tr < t; ++tr) B[(size_t)a * nbf + p + tr] = md.rc[(size_t)tr * n + j];
  }
  auto up = [&](DBuf& d, const void* src, size_t bytes) -> hipError_t {
    hipError_t e = d.ensure(bytes);
    if (e != hipSuccess) return e;
    return hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice);
  };
  KG_CK(ctx, up(ctx->wts, W.data(), W.size()));
  KG_CK(ctx, up(ctx->mask, md.maskw.data(), md.maskw.size() * 8));
  KG_CK(ctx, up(ctx->colw, colw.data(), colw.size() * 8));
  KG_CK(ctx, up(ctx->colh, colh.data(), colh.size() * 8));
  KG_CK(ctx, up(ctx->bp, bp.data(), bp.size() * 8));
  KG_CK(ctx, up(ctx->bp0, bp0.data(), bp0.size() * 8));
  KG_CK(ctx, up(ctx->rbar, md.rbar.data(), t * 8));
  KG_CK(ctx, up(ctx->ycn, md.ycn.data(), t * 8));
  KG_CK(ctx, up(ctx->yyM, md.yyM.data(), t * 8));
  KG_CK(ctx, up(ctx->B64, B.data(), B.size() * 8));
  KG_CK(ctx, up(ctx->bpfull, md.bprime.data(), std::max<size_t>(md.bprime.size(), 1) * 8));
  EpiI8& ep = ctx->epi;
  ep.n = n;
  ep.nz = nz;
  ep.t = t;
  ep.L = L;
  ep.NU = NU;
  ep.has_d = (md.rotated && !md.drop_d) ? 1 : 0;
  ep.sigma = md.sigma;
  ep.base_coef = md.base_coef;
  ep.dmax_drop = (md.rotated && md.drop_d) ? md.dmax : 0.0;
  ep.colw = ctx->colw.as<double>();
  ep.colh = ctx->colh.as<double>();
  ep.bp = ctx->bp.as<double>();
  ep.bp0 = ctx->bp0.as<double>();
  ep.rbar = ctx->rbar.as<double>();
  ep.ycn = ctx->ycn.as<double>();
  ep.yyM = ctx->yyM.as<double>();
  return KG_OK;
}

static int build_f32(kg_ctx* ctx) {
  Model& md = ctx->md;
  const int n = md.n, p = md.p, t = md.t;
  const int ncol = p + 2 * t;
  const int NT = (ncol + 15) / 16;
  if (NT > 6) return fail(ctx, KG_ERR_ARG, "p + 2t exceeds the FP32 kernel's 96 columns");
  ctx->NT = NT;
  std::vector<float> colv((size_t)n * 16 * NT, 0.f);  // [sample][col]
  const int NU = 16 * NT;
  EpiF32& ep = ctx->epf;
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
  KG_CK(ctx, ctx->wf32.ensure(W.size() * 4));
  KG_CK(ctx, hipMemcpy(ctx->wf32.p, W.data(), W.size() * 4, hipMemcpyHostToDevice));
  KG_CK(ctx, ctx->mask.ensure(md.maskw.size() * 8));
  KG_CK(ctx, hipMemcpy(ctx->mask.p, md.maskw.data(), md.maskw.size() * 8, hipMemcpyHostToDevice));
  KG_CK(ctx, ctx->uyf.ensure(Uyf.size() * 4));
  KG_CK(ctx, hipMemcpy(ctx->uyf.p, Uyf.data(), Uyf.size() * 4, hipMemcpyHostToDevice));
  ep.n = n;
  ep.p = p;
  ep.t = t;
  ep.NU = NU;
  ep.Uyf = ctx->uyf.as<float>();
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
  for (auto& e : ctx->ev) KG_CK(ctx, hipEventCreate(&e));
  KG_CK(ctx, ctx->counters.ensure(sizeof(unsigned long long) * (1 + kMaxTraits)));
  if (const char* v = std::getenv("KG_TILES_PER_WAVE")) {
    const int r = atoi(v);
    if (r == 2 || r == 4) ctx->R = r;
  }
  if (const char* v = std::getenv("KG_GRID_MULT")) {
    const int g = atoi(v);
    if (g >= 1 && g <= 64) ctx->grid_mult = g;
  }
  return KG_OK;
}

extern "C" void kg_destroy(kg_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* kg_last_error(const kg_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int kg_set_model(kg_ctx* ctx, int32_t n_acc, const int32_t* idxs, int32_t n,
                            const double* U, int32_t p, const double* Y, int32_t t) {
  if (!ctx) return KG_ERR_ARG;
  if (!ctx->stream) return fail(ctx, KG_ERR_STATE, "context has no device");
  if (n_acc <= 0 || n <= 0 || n > n_acc || p < 0 || t < 1 || t > kMaxTraits || !idxs || !Y ||
      (p > 0 && !U))
    return fail(ctx, KG_ERR_ARG, "kg_set_model: bad shape (need 0 < n <= n_acc, p >= 0, 1 <= t <= 8)");
  for (int j = 0; j < n; ++j) {
    if (idxs[j] < 0 || idxs[j] >= n_acc || (j > 0 && idxs[j] <= idxs[j - 1]))
      return fail(ctx, KG_ERR_ARG, "kg_set_model: idxs must be increasing accession positions");
  }
  KG_CK(ctx, hipSetDevice(ctx->device));
  Model& md = ctx->md;
  md = Model();
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
  prepare_model(md);
  ctx->have_model = false;
  const int rc = (ctx->precision == KG_PREC_MIXED) ? build_i8(ctx) : build_f32(ctx);
  if (rc != KG_OK) return rc;
  ctx->have_model = true;
  return KG_OK;
}

// Final decision for one candidate (k-mer idx, trait) with FP64 values.
struct Decided {
  bool pass;
  bool hit;
  kg_hit h;
};

static Decided decide(const kg_ctx* ctx, int64_t idx, int tr, int c, double xxM, double xyM,
                      double S, bool have_S, const kg_scan_params& prm) {
  const Model& md = ctx->md;
  Decided d{};
  d.pass = true;
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

// One bulk pass over device rows [0, m) (+ refinement), appending hits with
// k-mer indices offset by idx0.
static int scan_core(kg_ctx* ctx, const uint64_t* d_bits, int64_t m, int64_t rw, int64_t idx0,
                     const kg_scan_params& prm, std::vector<kg_hit>& hits, int64_t* n_passed,
                     bool stats) {
  Model& md = ctx->md;
  const int t = md.t;
  const double p_thr = std::pow(10.0, -prm.log10_thresh);
  const double Ts = t_star(p_thr);
  const double Tmax = (double)md.n * std::log(1.0 / (1.0 - (1 - 1e-15)));
  if (ctx->cand_cap == 0) {
    ctx->cand_cap = std::max<int64_t>(1 << 16, std::min<int64_t>(m / 16 + 1024, 1 << 24));
    KG_CK(ctx, ctx->cand.ensure(sizeof(Cand) * ctx->cand_cap));
  }
  ScanOut o{};
  if (stats) {
    KG_CK(ctx, ctx->st_c.ensure(sizeof(int32_t) * m));
    KG_CK(ctx, ctx->st_xx.ensure(sizeof(double) * m));
    KG_CK(ctx, ctx->st_xy.ensure(sizeof(double) * m * t));
    KG_CK(ctx, ctx->st_S.ensure(sizeof(double) * m * t));
    o.st_c = ctx->st_c.as<int32_t>();
    o.st_xx = ctx->st_xx.as<double>();
    o.st_xy = ctx->st_xy.as<double>();
    o.st_S = ctx->st_S.as<double>();
  }
  unsigned long long cnt[1 + kMaxTraits];
  const int grid = ctx->cus * ctx->grid_mult;
  for (int attempt = 0; attempt < 2; ++attempt) {
    o.cand = ctx->cand.as<Cand>();
    o.counters = ctx->counters.as<unsigned long long>();
    o.cap = ctx->cand_cap;
    KG_CK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    KG_CK(ctx, launch_zero(ctx->stream, o.counters, 1 + kMaxTraits));
    KG_CK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    if (ctx->precision == KG_PREC_MIXED) {
      EpiI8 ep = ctx->epi;
      ep.tau = prm.corr_threshold;
      ep.tau_trivial = prm.corr_threshold <= 0.0 ? 1 : 0;
      ep.all_sig = p_thr >= 1.0 ? 1 : 0;
      ep.r2star = (Ts > Tmax) ? INFINITY : -std::expm1(-Ts / (double)md.n);
      KG_CK(ctx, launch_scan_i8(ctx->NT, ctx->R, grid, ctx->stream, d_bits, m, rw, 0,
                                ctx->wts.p, ctx->mask.p, ep, o));
    } else {
      EpiF32 ep = ctx->epf;
      ep.tau = (float)prm.corr_threshold;
      ep.tau_trivial = prm.corr_threshold <= 0.0 ? 1 : 0;
      ep.all_sig = p_thr >= 1.0 ? 1 : 0;
      ep.Tstar_lo = (Ts > Tmax) ? INFINITY : Ts * (1.0 - 1e-9);
      KG_CK(ctx, launch_scan_f32(ctx->NT, grid, ctx->stream, d_bits, m, rw, 0,
                                 ctx->wf32.as<float>(), ctx->mask.as<uint64_t>(), ep, o));
    }
    KG_CK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    KG_CK(ctx, hipMemcpyAsync(cnt, o.counters, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
    KG_CK(ctx, hipStreamSynchronize(ctx->stream));
    if ((int64_t)cnt[0] <= ctx->cand_cap) break;
    ctx->cand_cap = (int64_t)cnt[0] + (int64_t)cnt[0] / 4 + 1024;
    KG_CK(ctx, ctx->cand.ensure(sizeof(Cand) * ctx->cand_cap));
  }
  const int64_t nc = (int64_t)cnt[0];
  float ms_bulk = 0.f;
  KG_CK(ctx, hipEventElapsedTime(&ms_bulk, ctx->ev[1], ctx->ev[2]));
  ctx->timing.bulk_ms = ms_bulk;
  ctx->timing.refine_ms = 0.0;
  ctx->timing.candidates = nc;
  ctx->timing.kmers = m;
  for (int tr = 0; tr < t; ++tr) n_passed[tr] += (int64_t)cnt[1 + tr];

  std::vector<Cand> cands(nc);
  if (nc) {
    KG_CK(ctx, hipMemcpy(cands.data(), ctx->cand.p, sizeof(Cand) * nc, hipMemcpyDeviceToHost));
  }
  if (ctx->precision == KG_PREC_MIXED) {
    std::vector<double> res((size_t)4 * nc);
    if (nc) {
      KG_CK(ctx, ctx->res.ensure(sizeof(double) * 4 * nc));
      KG_CK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
      KG_CK(ctx, launch_refine(ctx->stream, ctx->cand.as<Cand>(), nc, d_bits, rw, 0,
                               ctx->mask.as<uint64_t>(), (md.n_acc + 63) / 64,
                               ctx->B64.as<double>(), md.p + t, md.p, t, md.n, md.sigma,
                               md.base_coef, ctx->bpfull.as<double>(), ctx->rbar.as<double>(),
                               ctx->res.as<double>()));
      KG_CK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
      KG_CK(ctx, hipMemcpyAsync(res.data(), ctx->res.p, sizeof(double) * 4 * nc,
                                hipMemcpyDeviceToHost, ctx->stream));
      KG_CK(ctx, hipStreamSynchronize(ctx->stream));
      float ms_ref = 0.f;
      KG_CK(ctx, hipEventElapsedTime(&ms_ref, ctx->ev[2], ctx->ev[3]));
      ctx->timing.refine_ms = ms_ref;
    }
    for (int64_t i = 0; i < nc; ++i) {
      const Cand& cd = cands[i];
      const int c = (int)res[4 * i + 0];
      Decided d = decide(ctx, idx0 + cd.idx, cd.trait, c, res[4 * i + 1], res[4 * i + 2],
                         res[4 * i + 3], true, prm);
      if ((cd.flags & kPfAmb) && d.pass) n_passed[cd.trait] += 1;
      if (d.hit) hits.push_back(d.h);
    }
  } else {
    for (int64_t i = 0; i < nc; ++i) {
      const Cand& cd = cands[i];
      Decided d = decide(ctx, idx0 + cd.idx, cd.trait, -1, (double)cd.xx32, (double)cd.xy32,
                         0.0, false, prm);
      if (d.hit) hits.push_back(d.h);
    }
    if (!hits.empty()) {  // allele counts of reported k-mers from the bits
      // (the FP32 bulk pass keeps only xxM/xyM in its records)
      std::vector<uint64_t> row(rw);
      for (auto& h : hits) {
        if (h.c >= 0) continue;
        KG_CK(ctx, hipMemcpy(row.data(), d_bits + (h.idx - idx0) * rw, rw * 8,
                             hipMemcpyDeviceToHost));
        int c = 0;
        for (size_t w = 0; w < md.maskw.size() && (int64_t)w < rw; ++w)
          c += __builtin_popcountll(row[w] & md.maskw[w]);
        h.c = c;
      }
    }
  }
  float ms_tot = 0.f;
  KG_CK(ctx, hipEventElapsedTime(&ms_tot, ctx->ev[0],
                                 (ctx->precision == KG_PREC_MIXED && nc) ? ctx->ev[3] : ctx->ev[2]));
  ctx->timing.total_ms = ms_tot;
  return KG_OK;
}

static int check_scan_args(kg_ctx* ctx, const void* bits, int64_t m, int64_t rw,
                           const kg_scan_params* prm) {
  if (!ctx) return KG_ERR_ARG;
  if (!ctx->have_model) return fail(ctx, KG_ERR_STATE, "scan before kg_set_model");
  if (!prm) return fail(ctx, KG_ERR_ARG, "null scan params");
  if (m < 0) return fail(ctx, KG_ERR_ARG, "negative k-mer count");
  if (m > 0 && !bits) return fail(ctx, KG_ERR_ARG, "null bits");
  if (rw % 8 != 0 || rw < 8 * (int64_t)ctx->md.nch_i8)
    return fail(ctx, KG_ERR_ARG,
                "row_words must be a multiple of 8 and cover all accessions (>= " +
                    std::to_string(8 * ctx->md.nch_i8) + ")");
  if (((uintptr_t)bits & 15) != 0) return fail(ctx, KG_ERR_ARG, "bits must be 16-byte aligned");
  if (!(prm->log10_thresh >= 0.0)) return fail(ctx, KG_ERR_ARG, "log10_thresh must be >= 0");
  return KG_OK;
}

static int finish_hits(kg_ctx* ctx, std::vector<kg_hit>& hits, kg_hit* out, int64_t cap,
                       int64_t* n_out) {
  std::sort(hits.begin(), hits.end(), [](const kg_hit& a, const kg_hit& b) {
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

extern "C" int kg_scan_device(kg_ctx* ctx, const uint64_t* d_bits, int64_t m, int64_t row_words,
                              const kg_scan_params* prm, kg_hit* out, int64_t cap,
                              int64_t* n_out, int64_t* n_passed) {
  int rc = check_scan_args(ctx, d_bits, m, row_words, prm);
  if (rc) return rc;
  KG_CK(ctx, hipSetDevice(ctx->device));
  std::vector<kg_hit> hits;
  std::vector<int64_t> np(ctx->md.t, 0);
  if (m > 0) {
    rc = scan_core(ctx, d_bits, m, row_words, 0, *prm, hits, np.data(), false);
    if (rc) return rc;
  }
  if (n_passed)
    for (int tr = 0; tr < ctx->md.t; ++tr) n_passed[tr] = np[tr];
  return finish_hits(ctx, hits, out, cap, n_out);
}

extern "C" int kg_scan(kg_ctx* ctx, const uint64_t* bits, int64_t m, int64_t row_words,
                       const kg_scan_params* prm, kg_hit* out, int64_t cap, int64_t* n_out,
                       int64_t* n_passed) {
  if (!ctx) return KG_ERR_ARG;
  if (!ctx->have_model) return fail(ctx, KG_ERR_STATE, "scan before kg_set_model");
  if (m > 0 && !bits) return fail(ctx, KG_ERR_ARG, "null bits");
  if (row_words % 8 != 0 || row_words < 8 * (int64_t)ctx->md.nch_i8)
    return fail(ctx, KG_ERR_ARG, "row_words must be a multiple of 8 and cover all accessions");
  if (!prm) return fail(ctx, KG_ERR_ARG, "null scan params");
  KG_CK(ctx, hipSetDevice(ctx->device));
  // slices of at most 512 MiB of bits
  const int64_t slice = std::max<int64_t>(1, (int64_t)(512ll << 20) / (row_words * 8));
  const int64_t rows = std::min<int64_t>(slice, std::max<int64_t>(m, 1));
  KG_CK(ctx, ctx->stage.ensure((size_t)rows * row_words * 8));
  std::vector<kg_hit> hits;
  std::vector<int64_t> np(ctx->md.t, 0);
  for (int64_t s = 0; s < m; s += rows) {
    const int64_t mm = std::min<int64_t>(rows, m - s);
    KG_CK(ctx, hipMemcpyAsync(ctx->stage.p, bits + s * row_words, (size_t)mm * row_words * 8,
                              hipMemcpyHostToDevice, ctx->stream));
    int rc = scan_core(ctx, ctx->stage.as<uint64_t>(), mm, row_words, s, *prm, hits, np.data(),
                       false);
    if (rc) return rc;
  }
  if (n_passed)
    for (int tr = 0; tr < ctx->md.t; ++tr) n_passed[tr] = np[tr];
  return finish_hits(ctx, hits, out, cap, n_out);
}

extern "C" int kg_stats_device(kg_ctx* ctx, const uint64_t* d_bits, int64_t m,
                               int64_t row_words, int32_t* c, double* xxM, double* xyM,
                               double* S) {
  kg_scan_params prm{0.2, 6.0, 0.0};
  int rc = check_scan_args(ctx, d_bits, m, row_words, &prm);
  if (rc) return rc;
  if (m == 0) return KG_OK;
  if (!c || !xxM || !xyM || !S) return fail(ctx, KG_ERR_ARG, "null output array");
  KG_CK(ctx, hipSetDevice(ctx->device));
  std::vector<kg_hit> hits;
  std::vector<int64_t> np(ctx->md.t, 0);
  rc = scan_core(ctx, d_bits, m, row_words, 0, prm, hits, np.data(), true);
  if (rc) return rc;
  const int t = ctx->md.t;
  KG_CK(ctx, hipMemcpy(c, ctx->st_c.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  KG_CK(ctx, hipMemcpy(xxM, ctx->st_xx.p, sizeof(double) * m, hipMemcpyDeviceToHost));
  KG_CK(ctx, hipMemcpy(xyM, ctx->st_xy.p, sizeof(double) * m * t, hipMemcpyDeviceToHost));
  KG_CK(ctx, hipMemcpy(S, ctx->st_S.p, sizeof(double) * m * t, hipMemcpyDeviceToHost));
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
  KG_CK(ctx, hipSetDevice(ctx->device));
  DBuf dpop, dthr;
  KG_CK(ctx, dpop.ensure(n_acc));
  KG_CK(ctx, dthr.ensure((size_t)4 * npop * n_profiles));
  KG_CK(ctx, hipMemcpy(dpop.p, pop, n_acc, hipMemcpyHostToDevice));
  KG_CK(ctx, hipMemcpy(dthr.p, thr, (size_t)4 * npop * n_profiles, hipMemcpyHostToDevice));
  for (int i = 0; i < n_acc; ++i)
    if (pop[i] >= npop) return fail(ctx, KG_ERR_ARG, "kg_synth_bits: population out of range");
  if (m > 0)
    KG_CK(ctx, launch_synth(ctx->stream, d_out, seed, kmer0, m, (uint32_t)n_acc, row_words,
                            dpop.as<uint8_t>(), dthr.as<uint32_t>(), (uint32_t)npop,
                            (uint32_t)n_profiles));
  KG_CK(ctx, hipStreamSynchronize(ctx->stream));
  return KG_OK;
}
