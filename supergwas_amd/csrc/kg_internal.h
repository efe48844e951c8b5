// Shared definitions between the host side of libkgwas_hip (model set-up,
// decisions) and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace kg {

constexpr int kMaxTraits = 8;

// Bulk-kernel candidate: a k-mer/trait whose pre-filter or 1D significance is
// not decided "no" by the bulk pass.
struct Cand {
  int64_t idx;     // k-mer row within its shard
  int32_t trait;
  uint32_t flags;  // kPf* / kSig* bits below
  int32_t shard;
  float xx32;      // plain FP32 path: the reference-arithmetic xxM, xyM
  float xy32;
  int32_t mult;    // rows idx .. idx + mult - 1 are this row (duplicate-row scans), else 1
  int32_t c;       // allele count over the (shard-)masked accessions
  int32_t pad;
};
// Survivor lists hold rows below 2^31; a single-trait screen sets the top bit
// of an entry whose pre-filter pass it proved.
constexpr uint32_t kRowMask = 0x7FFFFFFFu;
// Dynamic LDS of the resident-weight int8 kernel (its static LDS, the column
// and trait records, stays below 8 KiB of the CU's 160 KiB).
constexpr size_t kResLdsBytes = 152 * 1024;
// FP6-digit resident weights: per 512-sample chunk, 4 shifts x NT tiles x 64
// lanes of 24-byte A operands (32 six-bit digits), stored as a 16-byte part
// [chunk][shift][tile][lane] followed by an 8-byte part in the same order.
constexpr size_t kF6OperandBytes = 24;
// Allele count of the FP6 resident kernel: a bit count of the masked bits
// (any launch), of the bits themselves (every accession of every chunk is in
// the model: no mask), or the count column of the weights (every model of the
// launch has a spare column slot: no mask, no bit count).
constexpr int kCntMasked = 0;
constexpr int kCntBits = 1;
constexpr int kCntColumn = 2;
constexpr uint32_t kPfPass = 1u;  // pre-filter decided pass
constexpr uint32_t kPfAmb = 2u;   // pre-filter within the error bound of tau
constexpr uint32_t kSigDef = 4u;  // 1D threshold decided pass
constexpr uint32_t kSigAmb = 8u;  // 1D threshold within the error bound
// full-statistic mode: the statistic's estimated quantisation error exceeds the
// bulk budget, so the k-mer's values are recomputed in FP64
constexpr uint32_t kStatRef = 16u;

// Per-scan outputs of the bulk kernels.
struct ScanOut {
  Cand* cand;                    // candidate records (unordered)
  unsigned long long* counters;  // [0] candidates; FP32 path: [1 + tr] pre-filter passes
  int64_t cap;                   // capacity of cand
  // full-statistic mode (null pointers otherwise)
  int32_t* st_c;
  double* st_xx;
  double* st_xy;  // [m][t]
  double* st_S;   // [m][t]
};

// Epilogue constants of the int8-limb kernel (device pointers for arrays).
//
// Weight columns ("base columns") b = 0..nb-1 are ordered residuals first
// (b < t: r_c of trait b), then the projected structure columns (b = t + k:
// z_k).  Column b, limb l sits in MFMA output row
//   u = 16 * (s / 4) + 4 * (b % 4) + s % 4,   s = L * (b / 4) + l,
// so every limb of a column lands in one lane of the 16x16 output tile.
struct EpiI8 {
  int n;            // phenotyped samples
  int nz;           // projected structure columns kept (z columns)
  int t;            // traits
  int L;            // limbs per column
  int ncol;         // column records (4 per lane slot; zero past nb)
  int tau_trivial;  // corr threshold <= 0: every polymorphic k-mer passes
  int all_sig;      // p threshold >= 1: every pre-filter pass is reported
  int has_d;        // z column 0 is the kept intercept-remainder column d
  int pairs;        // all-pairs scan: no pre-filter, and c = n is tested
  int pad_e;
  double sigma;     // (Q'U'1)_0: the rotated U'1 = sigma e_0
  double base_coef; // sigma^2 / n^2  (x'x - |cn m'|^2 = c - c^2 base_coef)
  double dmax_drop; // max|d| of the dropped intercept-remainder column (0 if kept)
  double tau;
  double r2star;    // 1D threshold in r^2
  double h_d;       // quantisation bound per set bit of column d (if kept)
  double inv_n;     // 1 / n
  const double* colrec; // [ncol][kColRec] {w0, w1, w2, k_0..k_7, E, h^2, 0}: staged in LDS
  const double* trc;    // [t][kTrcRec] {mean r, |y - ybar|, yyM, b'_0, h of r_c, E of r_c,
                        //               sum_b E_b k_tr (the mean correction of S),
                        //               |k_tr| over the structure columns}
  double sum_h2;        // sum over z columns of h^2   (|z|^2 bound: 2c' |z| |h| + c'^2 |h|^2)
  double sum_hbp[8];    // per trait: sum over z columns of h |b'|
  double sum_E2;        // sum over z columns of E^2
  double E_d;           // E of column d (if kept)
  int stat_est;         // full-statistic mode: flag k-mers whose estimated error is large
  int pad_s;
};
// Mean correction.  A column v is quantised to v~ = v - e (e = the per-sample
// remainder below the last limb, |e_j| <= h0).  With E = sum_j e_j over the
// phenotyped samples, the kernel adds (c / n) E to x'v~, so its error is
// sum_{j in x} (e_j - E / n) = -sum_{j not in x} (e_j - E / n): at most
// c' h with c' = min(c, n - c) and h = max_j |e_j - E / n|.
//
// Column record: limb weights w_l, then per trait k_tr = b'_tr of a structure
// column, 1 for the trait's own residual column and 0 otherwise, so that
// sum_b z_b k_tr = z'b' + x'r_c = S; then E and h^2 of the column.
constexpr int kColRec = 14;
constexpr int kTrcRec = 8;
constexpr int kMaxCols = 88;      // column records the int8 kernel stages (NT <= 16, L = 3)

// One shard of a batched int8 scan: its rows and its null model.  The
// persistent kernel walks the concatenation of all shards' 256-k-mer blocks.
struct ShardDesc {
  const uint64_t* bits;          // rows [m][rw]
  int64_t m, rw;
  int64_t blk0;                  // first global block of this shard
  const void* wts;               // model weights (int8 limbs, chunk-major)
  const void* mask;              // accession mask words
  int nchunks;                   // 512-sample chunks of the model
  int pad;
  unsigned long long* pass;      // [t] decided pre-filter passes of this shard
  int id;                        // index of this shard in the caller's table
  int pad2;
  const uint32_t* rows;          // k-mer j of the scan is row rows[j] (null: row j)
  const uint32_t* mult;          // k-mer j stands for mult[j] equal rows (null: 1 each)
  // FP6-digit resident kernel with a recheck pass (null otherwise): survivor
  // entries it cannot decide (pre-filter or threshold inside its bounds) are
  // appended here, proof bit kept, and neither counted nor reported; the
  // streaming three-limb kernel then scans this list
  uint32_t* rchk;
  unsigned long long* rchk_n;    // entries appended to rchk
  EpiI8 ep;
};

// Pre-filter screen of one shard: a cheap pass that contracts each k-mer with
// y - mean(y) only (SL int8 limbs per trait: three for one trait, two for
// more) and keeps every k-mer whose pre-filter is not certainly failed for
// some trait.
//
// Weights: [chunk][shift k][lane] 16-byte MFMA A operands.  Lane (g, row),
// operand byte e = 4 i + q holds accession 512 C + 128 g + 32 i + 8 q + k;
// row = Q (k % KG) + SL tr + l for trait tr, limb l, and row Q (k % KG) + SL t
// is the count row (1 for phenotyped accessions; t <= 7); Q = 4, 8 or 16
// rows per shift group, KG = 16 / Q; zero in the other rows.
struct ScreenDesc {
  const uint64_t* bits;
  int64_t m, rw;
  int64_t blk0;                  // first global screen block of this shard
  const void* wts;               // screen weights (above)
  const void* mask;              // accession mask words (as ShardDesc)
  uint32_t* bmap;                // kept k-mers: bit j % 32 of word j / 32
  uint32_t* list;                // survivors: rows within the shard, increasing
  unsigned long long* nsurv;     // survivor count (written by the compaction)
  // duplicate-row scans (null otherwise): bit j of dmap set = row j equals row
  // j - 1 under the accession mask (it is then not a survivor); mult[i] = the
  // run of equal rows that survivor list[i] starts
  uint32_t* dmap;
  uint32_t* mult;
  int chunk0, chunk1;            // this shard's compaction chunks [chunk0, chunk1)
  int nchunks, t, n;
  int ilv;                       // one trait: bmap interleaves kept and proven words
                                 // (scr_word), list entries carry the proof in bit 31
  float wf[kMaxTraits];          // S~ = w_tr (254 A_0 + A_1), A_l = limb sums
  float hf[kMaxTraits];          // |S - S~| <= c h_tr (quantisation + rounding), rounded up
  float thrf[kMaxTraits];        // tau^2 |y - ybar|^2 / n; < 0: trait never passes
};

// Compaction chunk: words [w0, w1) of shard `shard`'s screen bitmap.
struct ScrChunk {
  int shard;
  int pad;
  int64_t w0, w1;
};

// FP64 refinement inputs of one shard.
struct RefDesc {
  const uint64_t* bits;
  int64_t rw;
  const uint64_t* maskw;
  const double* B64;             // [n_acc][p + t]
  const double* bpfull;          // [t][p]
  const double* rbar;            // [t]
  double sigma, base_coef;
  int nwords, p, t, n;
};

// Read-back of a scan written by refine_f64 straight into page-locked host
// memory (device-visible pointers): the first `eager` candidate records with
// their FP64 results, and the `ncnt` counters.  Null pointers: nothing.
struct RefHost {
  Cand* cand;               // [eager]
  double* res;              // [4 * eager]
  unsigned long long* cnt;  // [ncnt]
  int64_t eager;
  int ncnt, pad;
};

// Allocation-free scan (kg_scan_device_noalloc): one scratch array per GPU
// holding the counters, the shard and refinement descriptors, and up to
// kNoAllocCands candidate records with their FP64 results.
constexpr int64_t kNoAllocCands = 65536;
constexpr size_t kNoAllocCounterOff = 0;          // 16 u64 counters
constexpr size_t kNoAllocDescOff = 256;           // ShardDesc, then RefDesc
constexpr size_t kNoAllocCandOff = 4096;          // Cand[kNoAllocCands]
constexpr size_t kNoAllocResOff = kNoAllocCandOff + sizeof(Cand) * (size_t)kNoAllocCands;
constexpr size_t kNoAllocScratchBytes = kNoAllocResOff + 4 * sizeof(double) * (size_t)kNoAllocCands;
constexpr int kNoAllocMaxDevices = 64;
static_assert(kNoAllocDescOff + sizeof(ShardDesc) + sizeof(RefDesc) <= kNoAllocCandOff,
              "descriptor slot of the allocation-free scratch");

// Epilogue constants of the plain FP32 kernel.
struct EpiF32 {
  int n, p, t, NU;
  int tau_trivial, all_sig;
  float tau;
  float ymean_f[kMaxTraits];    // FP32 mean of y (reference pre-filter)
  float ycsum_f[kMaxTraits];    // FP32 sum of y - mean
  float ynorm_f[kMaxTraits];    // FP32 |y - mean|
  double yyM[kMaxTraits];
  double Tstar_lo;              // loose 1D threshold in T (host decides exactly)
  const float* Uyf;             // [t][p] FP32 U'y
};

}  // namespace kg
