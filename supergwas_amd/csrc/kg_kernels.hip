// gfx950 kernels of the 1D k-mer association scan.
//
//  scan_i8   : packed bits x int8-limb weights on v_mfma_i32_16x16x64_i8 (exact
//              int32 sums), FP64 epilogue with a certified error bound per
//              k-mer, wave-ballot compaction of candidates.
//  scan_f32  : the reference's cols-mode arithmetic (src/kgwas.cpp:146-158,
//              183-200) on v_mfma_f32_16x16x4_f32, FP32 epilogue.
//  refine_f64: exact FP64 statistic of listed k-mers from their bits.
//  synth_bits: seeded synthetic rows (synth_core.h).
//
// Sample (K) order.  Both bulk kernels contract over accessions in an order
// chosen so that every lane streams contiguous bits:
//   scan_i8 : accession 512*C + 128*g + 32*(j/2) + 8*(e%4) + 4*(j%2) + e/4
//             <->  chunk C, lane group g = lane/16, MFMA step j (0..7),
//             operand byte e (0..15)
//   scan_f32: accession 256*C + 64*g + e          <->  chunk C, lane group g,
//             MFMA step e (0..63)
// The host lays the weights out in the same order (kg_api.cpp), so the
// contraction is the plain sum over accessions.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <type_traits>

#include "kg_internal.h"
#include "synth_core.h"

namespace kg {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

// Global-address-space view of a pointer read from a descriptor.  Loads
// through a generic pointer become flat loads, which also count against the
// LDS wait counter: every LDS wait of the compute would then wait for the
// loads still in flight as well.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gptr(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gptr_w(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

constexpr int kWaves = 4;        // waves per workgroup (FP32 scan)
constexpr int kRefWaves = 16;    // waves per candidate (refine_f64)
constexpr int kThreads = 64 * kWaves;
// The int8 scan runs one 8-wave workgroup per CU (two waves per SIMD): two
// 4-wave workgroups per CU measured ~40% slower on the MFMA pipe.
constexpr int kI8Waves = 8;
constexpr int kI8Threads = 64 * kI8Waves;

// Bare workgroup barrier for the stage loop.  Every LDS value a wave read in
// the stage has been consumed (so those reads are complete) and each wave has
// waited for its own direct-to-LDS loads with an explicit vmcnt; a fence here
// would make the compiler wait for every load in flight, including the bits
// prefetch meant to overlap the next stage.  The empty asm statements keep
// the compiler from moving memory accesses across the barrier.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// B operand of one MFMA step from a lane's chunk word w: operand byte
// e = 4 d + q takes bit 8 q + sb + d (sb = 0 or 4), so each dword is one shift
// and one mask.  Step js (0..7) of a 512-sample chunk uses word js / 2 and
// sb = 4 (js % 2).
__device__ __forceinline__ v4i unpack_word(uint32_t w, int sb) {
  return v4i{(int)((w >> sb) & 0x01010101u), (int)((w >> (sb + 1)) & 0x01010101u),
             (int)((w >> (sb + 2)) & 0x01010101u), (int)((w >> (sb + 3)) & 0x01010101u)};
}

// Wave-aggregated append of candidate records.  Must be reached by every lane
// of the wave (uniform control flow).
__device__ __forceinline__ void append_cand(bool want, const Cand& rec, const ScanOut& o) {
  const unsigned long long ballot = __ballot(want);
  if (ballot == 0) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)ballot) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(o.counters, (unsigned long long)__popcll(ballot));
  base = __shfl(base, leader);
  if (want) {
    const unsigned long long below = ballot & ((1ull << lane) - 1ull);
    const unsigned long long pos = base + (unsigned long long)__popcll(below);
    if ((long long)pos < o.cap) o.cand[pos] = rec;
  }
}

// ----------------------------------------------------------------------------
// Pre-filter + 1D threshold decision of the int8 path, with error bounds.
// Returns candidate flags (0 = decided "not reported and not counted").
// *pf_def is set when the pre-filter is decided "pass" (counted on device).
// Both tests compare squares (no division or square root); the relative
// margins 3e-12 and 1e-9 dwarf the rounding of the squared forms.
// pf_sure: the screen proved this k-mer's pre-filter pass (single-trait
// screened scans), so S and its bound are not consulted.
__device__ __forceinline__ uint32_t classify_i8(int c, double xxM, double bxx, double S,
                                                double bS, double xyM, double bxy,
                                                const EpiI8& ep, double ycn, double yy,
                                                bool pf_sure, bool* pf_def) {
  *pf_def = false;
  const int n = ep.n;
  if (c <= 0) return 0;                         // x = 0 (src/kgwas_2d.cpp:364-374)
  if (!ep.pairs) {
    if (c >= n) return 0;                       // |x - mean x| = 0 (src/kgwas.cpp:465)
    if (!(ycn >= 1e-12)) return 0;              // |y - mean y| < eps
  }
  uint32_t f;
  if (ep.tau_trivial || pf_sure) {
    f = kPfPass;
  } else {
    // |cor| >= tau  <=>  S^2 >= tau^2 (c (n - c) / n) |y - ybar|^2
    const double thr2 = ep.tau * ep.tau * ((double)c * (double)(n - c) / (double)n) * ycn * ycn;
    const double aS = fabs(S);
    const double lo = aS - bS, hi = aS + bS;
    if (lo > 0.0 && lo * lo >= thr2 * (1.0 + 3e-12)) f = kPfPass;
    else if (hi * hi < thr2 * (1.0 - 3e-12)) return 0;
    else f = kPfAmb;
  }
  *pf_def = (f == kPfPass);
  uint32_t s;
  if (ep.all_sig) {
    s = kSigDef;
  } else if (!(yy > 1e-12)) {
    s = 0;                                      // compute_hit -> p = 1
  } else if (xxM + bxx <= 1e-12) {
    s = 0;
  } else if (xxM - bxx <= 1e-12) {
    s = kSigAmb;
  } else {
    // r^2 = xyM^2 / (xxM yyM) against r2star, cross-multiplied
    const double axy = fabs(xyM);
    const double hi = axy + bxy;
    const double lo = fmax(axy - bxy, 0.0);
    if (lo * lo >= ep.r2star * (1.0 + 1e-9) * ((xxM + bxx) * yy)) s = kSigDef;
    else if (hi * hi < ep.r2star * (1.0 - 1e-9) * ((xxM - bxx) * yy)) s = 0;
    else s = kSigAmb;
  }
  if (s == 0 && f != kPfAmb) return 0;
  return f | s;
}

// ----------------------------------------------------------------------------
// Full-statistic mode: does the bulk statistic of this k-mer risk more than the
// budget of relative error (|dT| / max(T, 1) > 4e-6)?  The mean-corrected
// remainders behave as a sample of c of the n per-sample remainders, so
// dz_k has standard deviation h_k sqrt(c (n - c) / (3 n)) at most; the
// estimate takes kappa = 8 of those (the certified bounds, linear in c, would
// flag most k-mers).  Such k-mers get FP64 values from refine_f64.
__device__ __forceinline__ bool stat_needs_fp64(int c, double xx, double xy, double zh2,
                                                double h_r, double yy, const EpiI8& ep) {
  constexpr double kappa = 8.0, budget = 4e-6;
  const double cd = (double)c, nn = (double)ep.n;
  const double v3 = cd * (nn - cd) / (3.0 * nn);
  if (!(v3 > 0.0) || !(yy > 1e-12)) return false;  // c = 0, c = n: T = 0 either way
  const double exx = 2.0 * kappa * sqrt(zh2 * v3) + kappa * kappa * ep.sum_h2 * v3;
  const double exy = kappa * h_r * sqrt(v3);
  const double lo = xx - exx;
  if (!(lo > 1e-12)) return xx + exx > 1e-12;       // compute_hit's xxM > eps may flip
  const double r2 = xy * xy / (xx * yy);
  const double dr2 = (2.0 * fabs(xy) * exy + exy * exy) / (lo * yy) + r2 * exx / lo;
  const double q = 1.0 - r2 - dr2;
  if (!(q > 1e-15)) return true;                     // near the r^2 clamp
  const double T = -nn * log1p(-fmin(r2, 1.0 - 1e-15));
  return nn * dr2 / q > budget * fmax(T, 1.0);
}

// ----------------------------------------------------------------------------
// Bulk int8-limb scan over a table of shards.  Workgroup = 8 waves (one per
// CU); global block `w` belongs to the shard s with sd[s].blk0 <= w <
// sd[s+1].blk0, and wave v of that block owns the shard's k-mers
// [(b*8 + v) * 16R, +16R) (b = w - blk0) as R tiles of 16.  Persistent:
// blocks are strided over the grid, so one launch covers every shard.
//
// Pipeline per block: weights stream through two LDS stage buffers of one
// 512-sample chunk (8 MFMA steps) each, filled by direct global->LDS loads
// for chunk C+1 while chunk C computes; each lane's bits and accession mask
// for chunk C+1 are delivered the same way into its wave's LDS slot; the next
// block's first bits chunk loads during the block's last chunk, and its first
// weight stage during the last stage (even stage counts) or the epilogue.
// MFMA steps per LDS stage: a whole 512-sample chunk while two stages of
// weights fit in LDS, a half or a quarter chunk for the widest tiles.
__host__ __device__ constexpr int i8_stage_steps(int NT) { return NT <= 8 ? 8 : (NT <= 15 ? 4 : 2); }

template <int NT>
__device__ __forceinline__ void dma_stage(const v4i* __restrict__ src, v4i* dst, int wave,
                                          int lane) {
  // the stage's SS NT blocks of 64 v4i, dealt round-robin to the waves
  constexpr int NB = i8_stage_steps(NT) * NT;
#pragma unroll
  for (int i = 0; i < (NB + kI8Waves - 1) / kI8Waves; ++i) {
    const int qd = i * kI8Waves + wave;
    if (qd < NB)
      __builtin_amdgcn_global_load_lds((const void*)(src + qd * 64 + lane),
                                       (__attribute__((address_space(3))) void*)(dst + qd * 64),
                                       16, 0, 0);
  }
}

// Each lane's 16 bytes of chunk C for every tile, into this wave's LDS slot.
template <int R>
__device__ __forceinline__ void dma_bits(const uint64_t* const (&rowp)[R], int C, v4i* dst,
                                         int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r)
    __builtin_amdgcn_global_load_lds((const void*)(rowp[r] + 8 * C),
                                     (__attribute__((address_space(3))) void*)(dst + r * 64), 16,
                                     0, 0);
}

// The accession-mask slice of lane group g for chunk C (each lane its own copy).
__device__ __forceinline__ void dma_mask(const void* mask, int C, int g, v4i* dst) {
  __builtin_amdgcn_global_load_lds((const void*)(reinterpret_cast<const uint4*>(mask) + 4 * C + g),
                                   (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}

// Rows of block w for this lane, in two steps so that the survivor-list loads
// are in flight long before the addresses are needed: block_idx issues them
// (unconditionally: without a list it reads the accession mask, a valid word
// it then ignores), rows_at turns them into row pointers at first use.
template <int R>
__device__ __forceinline__ void block_idx(const ShardDesc& D, int64_t w, int wave, int col,
                                          int64_t (&k)[R], uint32_t (&ridx)[R]) {
  const int64_t kbase = ((w - D.blk0) * kI8Waves + wave) * (16 * R);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    k[r] = kbase + 16 * r + col;
    if (k[r] > D.m - 1) k[r] = D.m - 1;  // clamp: rows past m are read, never reported
    const uint32_t* src = D.rows ? D.rows + k[r] : reinterpret_cast<const uint32_t*>(D.mask);
    ridx[r] = *gptr(src);
  }
}

template <int R>
__device__ __forceinline__ void rows_at(const ShardDesc& D, const int64_t (&k)[R],
                                        const uint32_t (&ridx)[R], int g,
                                        const uint64_t* (&rowp)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t row = D.rows ? (int64_t)(ridx[r] & kRowMask) : k[r];
    rowp[r] = D.bits + row * D.rw + 2 * g;
  }
}

template <int R>
__device__ __forceinline__ void block_rows(const ShardDesc& D, int64_t w, int wave, int g, int col,
                                           const uint64_t* (&rowp)[R]) {
  int64_t k[R];
  uint32_t ridx[R];
  block_idx<R>(D, w, wave, col, k, ridx);
  rows_at<R>(D, k, ridx, g, rowp);
}

// Lane-swap sums.  swap16(x, y): rows of 16 lanes pair up (0,1), (2,3); the
// even row gets x summed over the pair, the odd row y summed over the pair.
// swap32(x, y): halves of 32 lanes pair up; the low half gets x summed over
// both halves, the high half y.  (v_permlane16_swap / v_permlane32_swap.)
__device__ __forceinline__ uint32_t swap16(uint32_t x, uint32_t y) {
  const auto v = __builtin_amdgcn_permlane16_swap(x, y, false, false);
  return v[0] + v[1];
}
__device__ __forceinline__ uint32_t swap32(uint32_t x, uint32_t y) {
  const auto v = __builtin_amdgcn_permlane32_swap(x, y, false, false);
  return v[0] + v[1];
}
__device__ __forceinline__ float swap16(float x, float y) {
  return __uint_as_float(0u) + [&] {
    const auto v = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false,
                                                    false);
    return __uint_as_float(v[0]) + __uint_as_float(v[1]);
  }();
}
__device__ __forceinline__ float swap32(float x, float y) {
  return __uint_as_float(0u) + [&] {
    const auto v = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false,
                                                    false);
    return __uint_as_float(v[0]) + __uint_as_float(v[1]);
  }();
}
__device__ __forceinline__ double swap16(double x, double y) {
  const uint64_t xb = (uint64_t)__double_as_longlong(x), yb = (uint64_t)__double_as_longlong(y);
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)xb, (uint32_t)yb, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(xb >> 32), (uint32_t)(yb >> 32),
                                                   false, false);
  return __longlong_as_double((long long)(((uint64_t)hi[0] << 32) | lo[0])) +
         __longlong_as_double((long long)(((uint64_t)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ double swap32(double x, double y) {
  const uint64_t xb = (uint64_t)__double_as_longlong(x), yb = (uint64_t)__double_as_longlong(y);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)xb, (uint32_t)yb, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(xb >> 32), (uint32_t)(yb >> 32),
                                                   false, false);
  return __longlong_as_double((long long)(((uint64_t)hi[0] << 32) | lo[0])) +
         __longlong_as_double((long long)(((uint64_t)hi[1] << 32) | lo[1]));
}

// Reduce-scatter over the four lane groups g = lane / 16: v[r] holds lane
// group g's part of tile r; returns the total of tile g % R.
template <int R, typename T>
__device__ __forceinline__ T group_scatter(const T (&v)[R]) {
  if constexpr (R == 4) {
    const T n0 = swap32(v[0], v[2]);  // halves: tiles 0 | 2
    const T n1 = swap32(v[1], v[3]);  //         tiles 1 | 3
    return swap16(n0, n1);            // rows:   even n0, odd n1
  } else {
    static_assert(R == 2, "tiles per wave");
    const T n = swap16(v[0], v[1]);   // even rows tile 0, odd rows tile 1
    return swap32(n, n);
  }
}

// Default tiles per wave of an instantiation: 4 while the accumulators leave
// room, 2 for wide tiles or the multi-trait epilogue.
__host__ __device__ constexpr int i8_tiles(int NT, int TT) { return (TT == 1 && NT <= 6) ? 4 : 2; }

// L = limbs per column; TT = traits the epilogue carries (1, or kMaxTraits);
// R = k-mer tiles of 16 per wave.
// total_dev, when given, holds the block count (written on the device by the
// screen's fixup; total_blocks is then only an upper bound).
//
// RES (resident weights): every chunk of the model's weights sits in dynamic
// LDS for as long as the workgroup scans that model (staged when the block's
// model changes), each lane loads its next chunk of bits straight into
// registers, and the chunk loop has no barrier at all, so the two waves of a
// SIMD never wait for each other.  Used for single-trait screened scans at two
// limbs when nchunks x NT x 8 KiB fits in LDS; the survivor list's top bit
// then carries the screen's proof of the pre-filter pass (pf_sure).
//
// F6 (resident only): the weights are L FP6 e2m3 digits per column (base 30,
// digit d stored as d / 8, |d| <= 15) against FP4 bit operands on
// v_mfma_scale_f32_16x16x128_f8f6f4, as in the screen: shift k (0..3) keeps
// bit k of every nibble of the lane's four dwords, so one MFMA covers 128
// samples; the B scale (2, 1, 1/2, 1/2 for k = 0..3) makes every bit weigh 1,
// so all four shifts accumulate into the same rows, and every sum is exact in
// FP32.  Element j = 8 d + m of lane group g, shift k, is accession
// 512 C + 128 g + 32 d + 4 m + k.  With D.rchk set, k-mers whose decision the
// digit bounds leave open go to the recheck list instead of the candidates.
// CM: the allele count from a bit count of the masked bits, of the unmasked
// bits, or from the count column (kCnt*); the MFMA operands are never masked.
template <int NT, int L, int TT, int R, bool RES, bool F6 = false, int CM = kCntMasked>
__global__ __launch_bounds__(kI8Threads, 2) void scan_i8(const ShardDesc* __restrict__ sd,
                                                        int nshards, int64_t total_blocks,
                                                        ScanOut out,
                                                        const int64_t* __restrict__ total_dev) {
  constexpr int SS = i8_stage_steps(NT);             // MFMA steps per stage
  constexpr int SPC = 8 / SS;                        // stages per 512-sample chunk
  constexpr int FRAG2 = SS * NT * 64;                // v4i per stage
  constexpr int J = (4 * NT) / L;                    // base columns per lane group
  constexpr int JSP = (TT + 4) / 4;                  // slots that may hold r_c or d columns
  static_assert(4 * J <= kMaxCols, "column records exceed the LDS table");
  static_assert(!RES || TT == 1, "resident weights: single-trait epilogue");
  static_assert(!F6 || (RES && (L == 2 || L == 3)), "FP6 digits: resident form, 2 or 3 digits");
  static_assert(F6 || CM == kCntMasked, "count modes: FP6 digits only");
  // the bits need the accession mask only for a popcount of masked bits
  constexpr bool MASKED = CM == kCntMasked;
  typedef int v8i __attribute__((ext_vector_type(8)));
  typedef int v2i __attribute__((ext_vector_type(2)));
  using AccV = std::conditional_t<F6, v4f, v4i>;
  // Separate LDS objects, so that reads of one buffer do not wait for the
  // direct-to-LDS loads still filling another.
  __shared__ v4i buf0[RES ? 1 : FRAG2];
  __shared__ v4i buf1[RES ? 1 : FRAG2];
  __shared__ v4i bitsbuf[RES ? 1 : kI8Waves * (R + 1) * 64];  // [wave][r | mask][lane]
  __shared__ double colrec_lds[4 * J * kColRec];
  __shared__ __attribute__((aligned(16))) float colf_lds[F6 ? 4 * J * 4 : 4];  // F6: {w0, E, k_0}, FP32
  __shared__ double trc_lds[kMaxTraits * kTrcRec];
  extern __shared__ v4i wres[];  // RES: [chunk][step][tile][lane] of the staged model
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15;
  v4i* const mybits = bitsbuf + (RES ? 0 : wave * (R + 1) * 64);
  // this wave's decided pre-filter passes of the current shard (flushed when
  // the shard changes, so no workgroup barrier closes a block)
  uint32_t wpass[TT];
#pragma unroll
  for (int tr = 0; tr < TT; ++tr) wpass[tr] = 0;

  if (total_dev != nullptr) total_blocks = *total_dev;
  int64_t w = blockIdx.x;
  if (w >= total_blocks) return;
  int s = 0;
  while (s + 1 < nshards && sd[s + 1].blk0 <= w) ++s;
  const uint64_t* rowp[R];  // this lane's rows of the current block
  block_rows<R>(sd[s], w, wave, g, col, rowp);
  v4u rcur[R], mcur;        // RES: this lane's bits and mask slice of the next chunk
  if constexpr (RES) {
#pragma unroll
    for (int r = 0; r < R; ++r) rcur[r] = *gptr(reinterpret_cast<const v4u*>(rowp[r]));
    if constexpr (MASKED) mcur = gptr(reinterpret_cast<const v4u*>(sd[s].mask))[g];
  } else {
    dma_bits<R>(rowp, 0, mybits, lane);
    dma_mask(sd[s].mask, 0, g, mybits + R * 64);
    dma_stage<NT>(reinterpret_cast<const v4i*>(sd[s].wts), buf0, wave, lane);
  }
  const double* staged_cr = nullptr;  // model whose epilogue constants are in LDS
  const void* staged_w = nullptr;     // RES: model whose weights are in LDS

  while (w < total_blocks) {
    const ShardDesc& D = sd[s];
    const EpiI8& ep = D.ep;
    const v4i* __restrict__ wts = reinterpret_cast<const v4i*>(D.wts);
    const int nchunks = D.nchunks;
    const int64_t kbase = ((w - D.blk0) * kI8Waves + wave) * (16 * R);
    // the next block's shard and rows, located now: with survivor lists the
    // row pointers come from a load whose latency the chunk loop hides
    const int64_t wn = w + gridDim.x;
    int sn = s;
    while (sn + 1 < nshards && sd[sn + 1].blk0 <= wn) ++sn;
    // (row pointers of the next block: nrowp, formed where first used)
    int64_t nk[R];
    uint32_t nridx[R];
    if (wn < total_blocks) block_idx<R>(sd[sn], wn, wave, col, nk, nridx);
    const uint64_t* nrowp[R];

    AccV acc[R][NT];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[r][t] = AccV{0, 0, 0, 0};
    uint32_t cnt[R];
#pragma unroll
    for (int r = 0; r < R; ++r) cnt[r] = 0;

    if constexpr (RES) {
      if (D.wts != staged_w) {  // workgroup-uniform: every wave is on block w
        // every wave is past the previous model's weights and constants
        __syncthreads();
        const v4i* __restrict__ gw = reinterpret_cast<const v4i*>(D.wts);
        // int8: 8 steps x NT tiles of 16-byte operands per chunk; FP6: 4 shifts
        // x NT tiles of 24-byte operands (16-byte part, then 8-byte part)
        const int nv = F6 ? nchunks * 4 * NT * 64 * 3 / 2 : nchunks * 8 * NT * 64;
        for (int i = threadIdx.x; i < nv; i += kI8Threads) wres[i] = gw[i];
        const double* __restrict__ gcr = ep.colrec;
        for (int i = threadIdx.x; i < ep.ncol * kColRec; i += kI8Threads) colrec_lds[i] = gcr[i];
        for (int i = threadIdx.x; i < ep.t * kTrcRec; i += kI8Threads) trc_lds[i] = ep.trc[i];
        if constexpr (F6) {
          for (int i = threadIdx.x; i < 4 * J; i += kI8Threads) {
            const bool in = i < ep.ncol;
            colf_lds[4 * i + 0] = in ? (float)gcr[i * kColRec + 0] : 0.f;
            colf_lds[4 * i + 1] = in ? (float)gcr[i * kColRec + 11] : 0.f;
            colf_lds[4 * i + 2] = in ? (float)gcr[i * kColRec + 3] : 0.f;
            colf_lds[4 * i + 3] = 0.f;
          }
        }
        __syncthreads();
        staged_w = D.wts;
        staged_cr = ep.colrec;
      }
      // MFMA loop at raised priority: the other wave's epilogue fills its gaps
      if constexpr (F6) __builtin_amdgcn_s_setprio(1);
      for (int C = 0; C < nchunks; ++C) {
        // FP6 operands need no mask: accessions outside the model have zero
        // weights in every row (the count column's included)
        uint32_t wd[R][4];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if constexpr (!MASKED) {
            wd[r][0] = rcur[r].x;
            wd[r][1] = rcur[r].y;
            wd[r][2] = rcur[r].z;
            wd[r][3] = rcur[r].w;
            if constexpr (CM == kCntBits) {
              // one bit count with accumulate per word (the compiler would
              // pair the counts and add them with a three-input add)
              uint32_t cc = cnt[r];
#pragma unroll
              for (int i = 0; i < 4; ++i) asm("v_bcnt_u32_b32 %0, %1, %0" : "+v"(cc) : "v"(wd[r][i]));
              cnt[r] = cc;
            }
          } else {
            wd[r][0] = rcur[r].x & mcur.x;
            wd[r][1] = rcur[r].y & mcur.y;
            wd[r][2] = rcur[r].z & mcur.z;
            wd[r][3] = rcur[r].w & mcur.w;
            cnt[r] += __popc(wd[r][0]) + __popc(wd[r][1]) + __popc(wd[r][2]) + __popc(wd[r][3]);
          }
        }
        // the next chunk loads under this one's MFMAs (the last chunk reloads
        // itself: unconditional loads into the same registers keep every wait
        // where the data is used), issued before the MFMAs
        {
          const int Cn = C + 1 < nchunks ? C + 1 : C;
#pragma unroll
          for (int r = 0; r < R; ++r)
            rcur[r] = *gptr(reinterpret_cast<const v4u*>(rowp[r] + 8 * Cn));
          if constexpr (MASKED) mcur = gptr(reinterpret_cast<const v4u*>(D.mask))[4 * Cn + g];
        }
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (F6) {
          // fragments q = NT k + t (shift k, tile t); fragment q + PD is read
          // while q's MFMAs run, shift k + 1's operands are formed beside
          // shift k's MFMAs
          const v4i* const xc = wres + (size_t)C * (4 * NT * 64);
          const v2i* const yc = reinterpret_cast<const v2i*>(wres + (size_t)nchunks * (4 * NT * 64)) +
                                (size_t)C * (4 * NT * 64);
          constexpr int NQ = 4 * NT;
          constexpr int PD = 4;
          v8i af[PD + 1];
          v4u bf[R], bn[R];
          auto opnd = [&](int r, int k) -> v4u {
            if (k < 3) {
              const uint32_t M = 0x11111111u << k;
              return v4u{wd[r][0] & M, wd[r][1] & M, wd[r][2] & M, wd[r][3] & M};
            }
            return v4u{(wd[r][0] >> 1) & 0x44444444u, (wd[r][1] >> 1) & 0x44444444u,
                       (wd[r][2] >> 1) & 0x44444444u, (wd[r][3] >> 1) & 0x44444444u};
          };
          auto afrag = [&](int q) -> v8i {
            const v4i a0 = xc[q * 64 + lane];
            const v2i a1 = yc[q * 64 + lane];
            return v8i{a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, 0, 0};
          };
#pragma unroll
          for (int r = 0; r < R; ++r) bf[r] = opnd(r, 0);
#pragma unroll
          for (int i = 0; i < PD; ++i) af[i] = afrag(i);
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const int k = q / NT, t = q % NT;
            if (k + 1 < 4) {
#pragma unroll
              for (int r = t; r < R; r += NT) bn[r] = opnd(r, k + 1);
            }
            if (q + PD < NQ) af[(q + PD) % (PD + 1)] = afrag(q + PD);
            // B scale (E8M0): bit values 0.5, 1, 2, 2 -> 1
            constexpr int kSb[4] = {128, 127, 126, 126};
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const v8i b = v8i{(int)bf[r].x, (int)bf[r].y, (int)bf[r].z, (int)bf[r].w, 0, 0, 0, 0};
              acc[r][t] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                  af[q % (PD + 1)], b, acc[r][t], 2, 4, 0, 127, 0, kSb[k]);
            }
            if (t == NT - 1 && k + 1 < 4) {
#pragma unroll
              for (int r = 0; r < R; ++r) bf[r] = bn[r];
            }
            __builtin_amdgcn_sched_barrier(0);
          }
        } else {
        const v4i* const wc = wres + (size_t)C * (8 * NT * 64);
        // the same software pipeline as the streaming form: fragment q + PD is
        // read while q's MFMAs run, step jj + 1's operands are unpacked beside
        // step jj's MFMAs
        constexpr int NQ = 8 * NT;
        constexpr int PD = 4;
        v4i bf[R], bn[R], af[PD + 1];
#pragma unroll
        for (int r = 0; r < R; ++r) bf[r] = unpack_word(wd[r][0], 0);
#pragma unroll
        for (int i = 0; i < PD; ++i) af[i] = wc[i * 64 + lane];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int jj = q / NT, t = q % NT;
          if (jj + 1 < 8) {
#pragma unroll
            for (int r = t; r < R; r += NT) bn[r] = unpack_word(wd[r][(jj + 1) >> 1], 4 * ((jj + 1) & 1));
          }
          if (q + PD < NQ) af[(q + PD) % (PD + 1)] = wc[(q + PD) * 64 + lane];
#pragma unroll
          for (int r = 0; r < R; ++r)
            acc[r][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[q % (PD + 1)], bf[r], acc[r][t],
                                                              0, 0, 0);
          if (t == NT - 1 && jj + 1 < 8) {
#pragma unroll
            for (int r = 0; r < R; ++r) bf[r] = bn[r];
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        }  // int8 limbs
      }
      if constexpr (F6) __builtin_amdgcn_s_setprio(0);
      // the next block's first chunk, under this block's epilogue
      if (wn < total_blocks) {
        rows_at<R>(sd[sn], nk, nridx, g, nrowp);
#pragma unroll
        for (int r = 0; r < R; ++r) rcur[r] = *gptr(reinterpret_cast<const v4u*>(nrowp[r]));
        if constexpr (MASKED) mcur = gptr(reinterpret_cast<const v4u*>(sd[sn].mask))[g];
      }
    }
    if constexpr (!RES) {
    __builtin_amdgcn_s_waitcnt(0);  // stage 0 weights and chunk 0 bits have landed
    __syncthreads();
    uint32_t wd[R][4];  // the current chunk's 128 bits per tile (masked)
    for (int C = 0; C < nchunks; ++C) {
#pragma unroll
      for (int hc = 0; hc < SPC; ++hc) {  // stages of the chunk
        const int S = C * SPC + hc;
        const bool odd = (SPC == 1) ? (C & 1) != 0 : (hc & 1) != 0;
        const v4i* cur = odd ? buf1 : buf0;
        if (S + 1 < nchunks * SPC)
          dma_stage<NT>(wts + (size_t)(S + 1) * FRAG2, odd ? buf0 : buf1, wave, lane);
        else if (!((nchunks * SPC) & 1) && wn < total_blocks)
          // even stage count: the last stage reads buf1 and buf0 was released
          // at the previous barrier, so the next block's first stage loads
          // now, under this stage's MFMAs and the epilogue
          dma_stage<NT>(reinterpret_cast<const v4i*>(sd[sn].wts), buf0, wave, lane);
        if (hc == 0) {
          const v4i mk = mybits[R * 64 + lane];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const v4i v = mybits[r * 64 + lane];
            wd[r][0] = (uint32_t)(v.x & mk.x);
            wd[r][1] = (uint32_t)(v.y & mk.y);
            wd[r][2] = (uint32_t)(v.z & mk.z);
            wd[r][3] = (uint32_t)(v.w & mk.w);
            cnt[r] += __popc(wd[r][0]) + __popc(wd[r][1]) + __popc(wd[r][2]) + __popc(wd[r][3]);
          }
          __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this wave's bits reads are done
          if (C + 1 < nchunks) {
            dma_bits<R>(rowp, C + 1, mybits, lane);
            dma_mask(D.mask, C + 1, g, mybits + R * 64);
          } else if (wn < total_blocks) {  // the next block's first chunk
            rows_at<R>(sd[sn], nk, nridx, g, nrowp);
            dma_bits<R>(nrowp, 0, mybits, lane);
            dma_mask(sd[sn].mask, 0, g, mybits + R * 64);
          }
        }
        // Software pipeline over the stage's SS NT weight fragments q = NT jj + t:
        // fragment q+PD is read from LDS while q's MFMAs run, and step jj+1's
        // bit operands are unpacked beside step jj's MFMAs.  The scheduling
        // barrier after each fragment keeps the compiler from regrouping the
        // LDS reads next to their first use.
        constexpr int NQ = SS * NT;
        constexpr int PD = 4;  // fragments read ahead
        v4i bf[R], bn[R], af[PD + 1];
#pragma unroll
        for (int r = 0; r < R; ++r) bf[r] = unpack_word(wd[r][(hc * SS) >> 1], 0);
#pragma unroll
        for (int i = 0; i < PD; ++i) af[i] = cur[i * 64 + lane];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int jj = q / NT, t = q % NT;
          if (jj + 1 < SS) {  // tile r's next operands beside fragment t = r mod NT
#pragma unroll
            for (int r = t; r < R; r += NT)
              bn[r] = unpack_word(wd[r][(hc * SS + jj + 1) >> 1], 4 * ((jj + 1) & 1));
          }
          if (q + PD < NQ) af[(q + PD) % (PD + 1)] = cur[(q + PD) * 64 + lane];
#pragma unroll
          for (int r = 0; r < R; ++r)
            acc[r][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[q % (PD + 1)], bf[r], acc[r][t],
                                                              0, 0, 0);
          if (t == NT - 1 && jj + 1 < SS) {
#pragma unroll
            for (int r = 0; r < R; ++r) bf[r] = bn[r];
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        // next stage's weights (and next chunk's bits) landed in every wave's
        // share.  After the block's last stage a wave goes straight to its
        // epilogue when the stage count is even: the last stage read buf1,
        // and buf0 (which the next block's first stage is loaded into) was
        // released at the previous barrier, so the early waves' epilogues
        // overlap the late waves' last MFMAs.
        // (No wait after an unbarriered last stage: what is in flight then is
        // the next block's data, which the block's opening wait covers.)
        if (S + 1 < nchunks * SPC || ((nchunks * SPC) & 1)) {
          __builtin_amdgcn_s_waitcnt(0);
          lds_barrier();
        }
      }
    }
    }  // !RES

    // this model's epilogue constants into LDS (when the model changes; the
    // wait also drains the next block's loads, so the epilogue then reads only
    // LDS); workgroup-uniform, and every wave finished the previous block's
    // epilogue before this block's opening barrier (RES: staged with the
    // weights, behind barriers of their own)
    const bool restage = !RES && ep.colrec != staged_cr;
    if (restage) {
      const double* __restrict__ gcr = ep.colrec;
      for (int i = threadIdx.x; i < ep.ncol * kColRec; i += kI8Threads) colrec_lds[i] = gcr[i];
      for (int i = threadIdx.x; i < ep.t * kTrcRec; i += kI8Threads) trc_lds[i] = ep.trc[i];
      __builtin_amdgcn_s_waitcnt(0);
      staged_cr = ep.colrec;
    }

    // next block's first stage, when the stage count is odd (the last stage
    // read buf0, released by the barrier that closed it); with an even count
    // it was loaded at the start of the last stage, and the first bits chunk
    // at the start of the last chunk
    if constexpr (!RES) {
      if (wn < total_blocks && ((nchunks * SPC) & 1))
        dma_stage<NT>(reinterpret_cast<const v4i*>(sd[sn].wts), buf0, wave, lane);
    }
    if (restage) lds_barrier();  // column records staged by every wave

    // ---- epilogue, from registers ----
    // Lane (g, col) holds, for k-mer col of tile r, output rows
    // u = 16 t + 4 g + i: the L limbs of base columns b = 4 j + g, j < J.  Each
    // lane folds its columns into partial sums per tile, then a reduce-scatter
    // over the four lane groups (lane-swap instructions, no LDS) leaves lane
    // (g, col) with the totals of tile g % R: every lane of the wave then owns
    // one k-mer (R = 4) and decides it alone.
    const int t = ep.t, n = ep.n;
    const int rme = g % R;
    const bool primary = g < R;
    const bool est = ep.stat_est != 0;  // full-statistic mode: error estimate per k-mer
    // FP6 digits (one trait): the column passes in FP32 -- z = w0 sum_l
    // 30^(L-1-l) A_l from the exact digit sums, then |z|^2, z'E, the column d
    // and the trait's own residual column, and (without a recheck pass) S; the
    // per-k-mer totals go on in FP64, and the bounds take a relative 1e-5 for
    // these FP32 roundings (under 4e-6 for sums of 32 terms).
    constexpr bool F32E = F6;
    // kCntColumn: the allele count is the digit-0 sum of the count column
    // (the last column slot, b = 4 J - 1: digit 1 for every sample of the
    // model, zero column record), exact in FP32.
    float fzz[R], fzE[R], fz0[R], fsv[R], fS[R], fcn[R];
    const bool f_S = F32E && D.rchk == nullptr;  // S of unproven k-mers decided here
    if constexpr (F32E) {
#pragma unroll
      for (int r = 0; r < R; ++r) fzz[r] = fzE[r] = fz0[r] = fsv[r] = fS[r] = fcn[r] = 0.f;
#pragma unroll
      for (int jc = 0; jc < J; ++jc) {
        const double* cr = colrec_lds + (4 * jc + g) * kColRec;
        const v4f cf4 = *reinterpret_cast<const v4f*>(colf_lds + 4 * (4 * jc + g));
        const float w0f = cf4.x, cEf = cf4.y, kcf = cf4.z;
        const int b = 4 * jc + g;
        const int sl = L * jc;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const float d0 = acc[r][sl >> 2][sl & 3], d1 = acc[r][(sl + 1) >> 2][(sl + 1) & 3];
          if (CM == kCntColumn && jc == J - 1) fcn[r] = g == 3 ? d0 : 0.f;
          float z;
          if (L == 2) {
            z = w0f * fmaf(30.0f, d0, d1);
          } else {
            const float d2 = acc[r][(sl + 2) >> 2][(sl + 2) & 3];
            z = (float)(cr[0] * fma(900.0, (double)d0, fma(30.0, (double)d1, (double)d2)));
          }
          // column slot j < JSP: the trait's residual (b = 0) or column d (b = 1)
          const float zs = (jc < JSP && b < t) ? 0.f : z;
          fzz[r] = fmaf(zs, zs, fzz[r]);
          fzE[r] = fmaf(zs, cEf, fzE[r]);
          if (jc < JSP && b == t) fz0[r] = z;
          if (jc < JSP && b == 0) fsv[r] = z;
          if (f_S) fS[r] = fmaf(z, kcf, fS[r]);
        }
      }
    }
    // this lane's column values z (FP64) of every tile, after which the limb
    // sums are dead: the per-trait passes below then need R registers of
    // partial sums each, not R x t
    double zc[F32E ? 1 : R][F32E ? 1 : J];
#pragma unroll
    for (int jc = 0; jc < (F32E ? 0 : J); ++jc) {
      const double* cr = colrec_lds + (4 * jc + g) * kColRec;
      const double w0 = cr[0], w1 = cr[1], w2 = cr[2];
      const int sl = L * jc;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (F6) {
          // digit sums A_l (exact FP32 multiples of 1/8): z = w0 sum_l 30^(L-1-l) A_l,
          // exact in FP32 for two digits (|30 A_0 + A_1| < 2^21) and in FP64 for three
          const float d0 = acc[r][sl >> 2][sl & 3], d1 = acc[r][(sl + 1) >> 2][(sl + 1) & 3];
          if (L == 2) {
            zc[r][jc] = w0 * (double)fmaf(30.0f, d0, d1);
          } else {
            const float d2 = acc[r][(sl + 2) >> 2][(sl + 2) & 3];
            zc[r][jc] = w0 * fma(900.0, (double)d0, fma(30.0, (double)d1, (double)d2));
          }
          continue;
        }
        const int a0 = acc[r][sl >> 2][sl & 3];
        if (L == 3) {
          // limbs 1 and 2 merged exactly: |a| <= 127 n < 2^23 for n <= 60000, so
          // the 24-bit multiply-add is exact
          const int a12 = __mul24((int)acc[r][(sl + 1) >> 2][(sl + 1) & 3], 254) +
                          (int)acc[r][(sl + 2) >> 2][(sl + 2) & 3];
          zc[r][jc] = fma(w2, (double)a12, w0 * (double)a0);
        } else if (L == 2) {
          zc[r][jc] = fma(w1, (double)(int)acc[r][(sl + 1) >> 2][(sl + 1) & 3], w0 * (double)a0);
        } else {
          zc[r][jc] = w0 * (double)a0;
        }
      }
    }
    // trait-independent sums over the structure columns (b >= t)
    double pzz[R], pz0[R], pzE[R], pzh[R];
#pragma unroll
    for (int r = 0; r < R; ++r) pzz[r] = pz0[r] = pzE[r] = pzh[r] = 0.0;
#pragma unroll
    for (int jc = 0; jc < (F32E ? 0 : J); ++jc) {
      const double* cr = colrec_lds + (4 * jc + g) * kColRec;
      const double cE = cr[11], ch2 = cr[12];
      const int b = 4 * jc + g;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double z = zc[r][jc];
        // columns j < JSP may be a residual column (b < t) or column d (b == t)
        const double zs = (jc < JSP && b < t) ? 0.0 : z;
        pzz[r] = fma(zs, zs, pzz[r]);
        pzE[r] = fma(zs, cE, pzE[r]);
        if (est) pzh[r] = fma(zs * zs, ch2, pzh[r]);
        if (jc < JSP && b == t) pz0[r] = z;
      }
    }
    const int c = CM == kCntColumn ? (int)group_scatter<R>(fcn) : (int)group_scatter<R>(cnt);
    // FP32 column totals (FP6 digits): this lane's k-mer
    float tzE = 0.f, tzz = 0.f, tz0 = 0.f, tsv = 0.f;
    if constexpr (F32E) {
      tzE = group_scatter<R>(fzE);
      tzz = group_scatter<R>(fzz);
      tz0 = group_scatter<R>(fz0);
      tsv = group_scatter<R>(fsv);
    }
    const double cd = (double)c, cn = cd * ep.inv_n;
    // mean correction (kg_internal.h): z_k + (c/n) E_k for every column
    const double kzE = F32E ? (double)tzE : group_scatter<R>(pzE);
    const double kzz = (F32E ? (double)tzz : group_scatter<R>(pzz)) +
                       cn * fma(cn, ep.sum_E2, 2.0 * kzE);
    const double kz0 = (F32E ? (double)tz0 : group_scatter<R>(pz0)) + cn * ep.E_d;
    const double kzh = est ? group_scatter<R>(pzh) : 0.0;
    // FP32 column sums: relative slack for their roundings
    const double f32s = F32E ? 1e-5 : 0.0;

    const int64_t kidx = kbase + 16 * rme + col;
    const bool valid = primary && (kidx < D.m);
    // survivor lists: the row, and (top bit) the screen's proof of the pass
    const uint32_t lrow = (D.rows != nullptr && valid) ? gptr(D.rows)[kidx] : 0u;
    const int64_t krow = (D.rows != nullptr && valid) ? (int64_t)(lrow & kRowMask) : kidx;
    // (single-trait lists only: a multi-trait screen proves nothing)
    const bool pf_sure = (RES || TT == 1) && (lrow & ~kRowMask) != 0u;
    const uint32_t kmult = (D.mult != nullptr && valid) ? gptr(D.mult)[kidx] : 1u;

    // FP6 digits, hits-only scans: a decision in FP32 first.  A k-mer is
    // decided here when its pre-filter pass is proven (or the threshold is
    // trivial) and r^2 < r2star is certain with the bounds of the FP64 pass
    // below, evaluated in FP32 with 3e-5 relative slack for the column sums'
    // and these totals' roundings (each under 1e-6) and a 1e-4 margin on the
    // comparison; degenerate k-mers (c = 0, c = n, |y - ybar| = 0) are
    // decided as classify_i8 decides them.  A wave whose k-mers are all
    // decided skips the FP64 pass: it only counts its passes.
    bool fast_all = false, fpd = false;
    if constexpr (F32E) {
      const double* tc = trc_lds;  // the one trait
      bool fast = !valid;
      if (valid && !ep.pairs && out.st_c == nullptr && !ep.all_sig) {
        if (c <= 0 || c >= n || !(tc[1] >= 1e-12)) {
          fast = true;  // classify_i8: not reported, not counted
        } else if (pf_sure || ep.tau_trivial) {
          if (!(tc[2] > 1e-12)) {
            fast = fpd = true;  // compute_hit gives p = 1
          } else {
            const float cf = (float)c, nf = (float)n;
            const float cnf = cf * (float)ep.inv_n, cmf = fminf(cf, nf - cf);
            const float sh2 = (float)ep.sum_h2, sE2 = (float)ep.sum_E2;
            const float kzzf = tzz + cnf * fmaf(cnf, sE2, 2.f * tzE);
            const float kzzp = fmaxf(kzzf, 0.f);
            float crossf = 0.f, bcrf = 0.f;
            if (ep.has_d) {
              const float sg = (float)ep.sigma;
              crossf = 2.f * cnf * sg * (tz0 + cnf * (float)ep.E_d);
              bcrf = 2.f * cnf * fabsf(sg) * cmf * (float)ep.h_d;
            }
            const float c2b = cf * cf * (float)ep.base_coef;
            const float xxf = cf - c2b - crossf - kzzf;
            float bxxf = 2.f * cmf * sqrtf(kzzp * sh2) + cmf * cmf * sh2 + bcrf +
                         3e-5f * (cf + c2b + kzzp + fabsf(crossf) +
                                  2.f * cnf * sqrtf(kzzp * sE2) + 1.f);
            if (ep.dmax_drop > 0.0) {
              const float e0 = cf * (float)ep.dmax_drop;
              bxxf += 2.f * cnf * fabsf((float)ep.sigma) * e0 + e0 * e0;
            }
            const float svf = tsv + cnf * (float)tc[5], rbf = (float)tc[0];
            const float xyf = fmaf(cf, rbf, svf);
            const float bxyf = cmf * (float)tc[4] + 3e-5f * (fabsf(svf) + cf * fabsf(rbf) + 1.f);
            const float lo = xxf - bxxf, hi = fabsf(xyf) + bxyf;
            if (lo > 1e-6f * cf && hi * hi < (float)ep.r2star * 0.9999f * lo * (float)tc[2])
              fast = fpd = true;
          }
        }
      }
      fast_all = __ballot(!fast) == 0ull;
    }
    if (fast_all) {
      if (D.mult == nullptr) {
        wpass[0] += (uint32_t)__popcll(__ballot(fpd));
      } else {
        uint32_t v = fpd ? kmult : 0u;
        for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
        wpass[0] += v;
      }
    } else {
    // |dz_k| <= c' h_k (c' = min(c, n - c))  =>
    //   |d |z|^2| <= 2c' sum|z_k| h_k + c'^2 sum h_k^2 <= 2c' |z| |h| + c'^2 |h|^2
    const double cm = fmin(cd, (double)ep.n - cd);
    const double bzz = 2.0 * cm * sqrt(kzz * ep.sum_h2) + cm * cm * ep.sum_h2;
    double cross = 0.0, bcross = 0.0;
    if (ep.has_d) {
      cross = 2.0 * cn * ep.sigma * kz0;
      bcross = 2.0 * cn * fabs(ep.sigma) * cm * ep.h_d;
    }
    const double xxM = cd - cd * cd * ep.base_coef - cross - kzz;
    double bxx = bzz + bcross + 1e-11 * (cd + kzz + fabs(cross) + 1.0);
    if (F32E)
      bxx += f32s * (cd + kzz + fabs(cross) + 2.0 * cn * sqrt(fmax(kzz, 0.0) * ep.sum_E2) + 1.0);
    if (ep.dmax_drop > 0.0) {
      const double e0 = cd * ep.dmax_drop;
      bxx += 2.0 * cn * fabs(ep.sigma) * e0 + e0 * e0;
    }
    if (out.st_c != nullptr && valid) {
      out.st_c[kidx] = c;
      out.st_xx[kidx] = xxM;
    }
#pragma unroll
    for (int tr = 0; tr < TT; ++tr) {
      if (tr < t) {
        // S = z'b' + x'r_c (column weights k_tr) and x'r_c of this trait
        double vS[R], vsv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) vS[r] = vsv[r] = 0.0;
#pragma unroll
        for (int jc = 0; jc < (F32E ? 0 : J); ++jc) {
          const double kc = colrec_lds[(4 * jc + g) * kColRec + 3 + tr];
          const bool own = jc < JSP && 4 * jc + g == tr;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            vS[r] = fma(zc[r][jc], kc, vS[r]);
            if (own) vsv[r] = zc[r][jc];
          }
        }
        const double* tc = trc_lds + tr * kTrcRec;
        double S, sv;
        if constexpr (F32E) {
          // without S (recheck pass), an unproven pre-filter stays open: bS = inf
          S = f_S ? (double)group_scatter<R>(fS) + cn * tc[6] : 0.0;
          sv = (double)tsv + cn * tc[5];
        } else {
          S = group_scatter<R>(vS) + cn * tc[6];
          sv = group_scatter<R>(vsv) + cn * tc[5];
        }
        const double rbar = tc[0];
        const double es = cm * tc[4];
        double bS = cm * ep.sum_hbp[tr] + es;
        if (ep.dmax_drop > 0.0) bS += cd * ep.dmax_drop * fabs(tc[3]);
        bS += 1e-12 * (fabs(S) + fabs(sv) + 1.0);
        if (F32E) {
          // FP32 S = sum_k z_k k_k + x'r_c: its rounding is below 1e-5 (|z| |k| +
          // |x'r_c| + |S|), |k| = the norm of the structure columns' weights k_tr
          // (trait record tc[7])
          bS = f_S ? bS + f32s * (sqrt(fmax(kzz, 0.0)) * tc[7] + fabs(S) + fabs(sv) + 1.0)
                   : INFINITY;
        }
        const double xyM = sv + cd * rbar;
        const double bxy = es + (1e-12 + f32s) * (fabs(sv) + cd * fabs(rbar) + 1.0);
        bool pf_def = false;
        uint32_t fl = 0;
        if (valid) fl = classify_i8(c, xxM, bxx, S, bS, xyM, bxy, ep, tc[1], tc[2], pf_sure, &pf_def);
        if constexpr (F6) {
          if (D.rchk != nullptr) {
            // an open pre-filter or threshold goes to the recheck pass, which
            // counts and reports the k-mer itself
            const bool rc = (fl & (kPfAmb | kSigAmb)) != 0u;
            if (rc) {
              fl = 0;
              pf_def = false;
            }
            const unsigned long long bal = __ballot(rc);
            if (bal != 0ull) {
              const int leader = __ffsll((long long)bal) - 1;
              unsigned long long base = 0;
              if (lane == leader) base = atomicAdd(D.rchk_n, (unsigned long long)__popcll(bal));
              base = __shfl(base, leader);
              if (rc) gptr_w(D.rchk)[base + __popcll(bal & ((1ull << lane) - 1ull))] = lrow;
            }
          }
        }
        if (out.st_c != nullptr && valid) {
          out.st_xy[kidx * t + tr] = xyM;
          out.st_S[kidx * t + tr] = S;
          if (est && stat_needs_fp64(c, xxM, xyM, kzh, tc[4], tc[2], ep)) fl |= kStatRef;
        }
        if (D.mult == nullptr) {
          wpass[tr] += (uint32_t)__popcll(__ballot(pf_def));
        } else {  // each k-mer stands for kmult equal rows
          uint32_t v = pf_def ? kmult : 0u;
          for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
          wpass[tr] += v;
        }
        Cand rec;
        rec.idx = krow;
        rec.trait = tr;
        rec.flags = fl;
        rec.shard = D.id;
        rec.xx32 = 0.f;
        rec.xy32 = 0.f;
        rec.mult = (int32_t)kmult;
        rec.c = c;
        rec.pad = 0;
        append_cand(fl != 0, rec, out);
      }
    }
    }  // FP64 pass
    // the shard's passes counted by this wave, when the next block is another
    // shard's (or there is none)
    if (sn != s || wn >= total_blocks) {
#pragma unroll
      for (int tr = 0; tr < TT; ++tr) {
        if (lane == 0 && tr < t && wpass[tr] != 0)
          atomicAdd(&D.pass[tr], (unsigned long long)wpass[tr]);
        wpass[tr] = 0;
      }
    }
    w = wn;
    s = sn;
    if (w < total_blocks) rows_at<R>(sd[s], nk, nridx, g, rowp);
  }
}

// ----------------------------------------------------------------------------
// Pre-filter screen (src/kgwas.cpp:183-218 runs the correlation filter before
// the projections; this is the same order on the device).  Per k-mer it forms
// S~ = x'(y - ybar) from int8 limbs of y - ybar (three for one trait, two per
// trait otherwise) and the allele count c,
// and keeps the k-mer unless |cor| < tau is certain for every trait:
//   (|S~| + c h)^2 < tau^2 |y - ybar|^2 c (n - c) / n  (1 - 1e-5)
// (evaluated in FP32 with every rounding on the safe side).
// The kept rows go to scan_i8, which decides them exactly as it decides an
// unscreened scan, so screening changes no count and no hit.
//
// B operands need no shift: step k (0..7) of a 512-sample chunk masks bit k of
// every byte of the lane's four dwords (one AND per dword), so operand bytes
// are 0 or 2^k (-128 for k = 7).  Each output row sees a single k: the
// weights place shift k in rows Q (k % KG) .. + Q - 1 of accumulator k / KG
// and zero elsewhere, and the epilogue divides each row by its exact power
// of two.  Workgroup: W independent waves (scr_waves); the model's weights
// (8 KiB per chunk) sit in LDS for the whole launch.  Wave v of global block
// w owns the shard's k-mers [(b*W + v) * 16R, +16R), b = w - blk0.  The
// wave's ballot of
// kept k-mers is stored as bitmap words (no atomics); scr_count, scr_scan and
// scr_write then compact the bitmaps into increasing row lists.
constexpr int kScrMaxChunks = 16;
constexpr int kScrChunkWords = 2048;  // bitmap words per compaction chunk
// Weight rows per shift group: 2 limbs per trait plus, for t <= 7, a count
// row (the phenotype mask, so that the MFMA also yields c); t = 8 counts with
// popcounts instead.
__host__ __device__ constexpr bool scr_count_row(int t) { return 2 * t + 1 <= 16; }
__host__ __device__ constexpr int scr_rows(int t) {
  return (2 * t + (scr_count_row(t) ? 1 : 0)) <= 4 ? 4 : ((2 * t + (scr_count_row(t) ? 1 : 0)) <= 8 ? 8 : 16);
}
// k-mer tiles per wave: 4 for the two-accumulator form (Q = 4), else 2; the
// chunks per load group keep a group at 8 KiB of rows per wave.
__host__ __device__ constexpr int scr_tiles(int Q) { return Q == 4 ? 4 : 2; }
// Waves per workgroup: 16 (the whole CU's registers at 128 VGPRs) for Q = 4,
// else 8; one workgroup then owns whole CUs, so a screen launched beside
// another kernel takes whole CUs and leaves whole CUs.
__host__ __device__ constexpr int scr_waves(int Q) { return Q == 4 ? 16 : 8; }
__host__ __device__ constexpr int scr_group_chunks(int R) { return 8 / R; }
// LDS: the accession mask (64 B per chunk), then the weights (8 KiB per chunk).
__host__ __device__ constexpr size_t scr_lds_bytes(int nch) {
  return (size_t)nch * 64 + (size_t)nch * 8 * 64 * 16;
}

// Load group (block w of shard D, chunks C0 .. C0 + DC - 1) of this lane's R
// rows.  Chunks past the model's last are loaded again as the last chunk
// (never used): unconditional loads keep the wait counts exact.
template <int W, int R, int DC>
__device__ __forceinline__ void scr_load(const ScreenDesc& D, int64_t w, int C0, int wave, int g,
                                         int col, v4u (&q)[DC][R]) {
  // screened shards hold fewer than 2^31 rows: 32-bit row indices, and one
  // 32 x 32 -> 64-bit multiply-add per row address
  const uint32_t kbase = (uint32_t)(((w - D.blk0) * W + wave) * (16 * R));
  const uint32_t klast = (uint32_t)(D.m - 1);
  const int nch = D.nchunks;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    uint32_t k = kbase + 16 * r + col;
    if (k > klast) k = klast;  // clamp: rows past m are read, never kept
    const auto* rp =
        gptr(reinterpret_cast<const v4u*>(D.bits + (uint64_t)k * (uint32_t)D.rw)) + g;
#pragma unroll
    for (int u = 0; u < DC; ++u) q[u][r] = rp[4 * min(C0 + u, nch - 1)];
  }
}

// MFMA steps of one load group into the block's accumulators (and, without a
// count row, the popcount of the phenotyped bits).
// Row rotate right by one lane within each row of 16 lanes: lane (g, col)
// receives lane (g, col - 1 mod 16)'s value.
__device__ __forceinline__ uint32_t ror16(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x121, 0xF, 0xF, false);
}

// DD: ne[r] |= bits where this lane's slice of tile r's row differs (under the
// accession mask) from the previous row's: row kbase + 16 r + col - 1 is lane
// col - 1 of the same tile, or lane 15 of tile r - 1 when col = 0 (the wave's
// first row has no previous row here and is never a duplicate).
template <int KG, int NA, int R, int DC, bool CR, bool DD>
__device__ __forceinline__ void scr_group(const v4u (&q)[DC][R], int C0, int nch, const v4i* wl,
                                          const v4u* mk_lds, int g, int lane, v4i (&acc)[R][NA],
                                          uint32_t (&cnt)[R], uint32_t (&ne)[R]) {
#pragma unroll
  for (int u = 0; u < DC; ++u) {
    if (C0 + u < nch) {
      const int C = C0 + u;
      if (!CR || DD) {
        const v4u mk = mk_lds[4 * C + g];
        if (!CR) {
#pragma unroll
          for (int r = 0; r < R; ++r)
            cnt[r] += __popc(q[u][r].x & mk.x) + __popc(q[u][r].y & mk.y) +
                      __popc(q[u][r].z & mk.z) + __popc(q[u][r].w & mk.w);
        }
        if (DD) {
          // tile by tile, so that only the previous tile's rotation stays live
          const bool first_col = (lane & 15) == 0;
          v4u prev = v4u{0u, 0u, 0u, 0u};
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const v4u rr = v4u{ror16(q[u][r].x), ror16(q[u][r].y), ror16(q[u][r].z),
                               ror16(q[u][r].w)};
            const v4u pv = (r > 0 && first_col) ? prev : rr;
            ne[r] |= ((q[u][r].x ^ pv.x) & mk.x) | ((q[u][r].y ^ pv.y) & mk.y) |
                     ((q[u][r].z ^ pv.z) & mk.z) | ((q[u][r].w ^ pv.w) & mk.w);
            prev = rr;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const v4i a = wl[(C * 8 + k) * 64 + lane];
        const uint32_t M = 0x01010101u << k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const v4i b = v4i{(int)(q[u][r].x & M), (int)(q[u][r].y & M), (int)(q[u][r].z & M),
                            (int)(q[u][r].w & M)};
          acc[r][k / KG] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc[r][k / KG], 0, 0, 0);
        }
      }
    }
  }
}

// One-trait screen on FP4 bit operands (v_mfma_scale_f32_16x16x128_f8f6f4,
// A = FP6 e2m3 weight digits, B = FP4 e2m1 bits).  Shift k (0..3) of a
// 512-sample chunk keeps bit k of every nibble of the lane's four dwords with
// one AND (k = 3 is moved to bit 2 first: a set sign bit would read as -0), so
// each MFMA covers 128 samples: element j = 8 d + m of lane group g is sample
// 128 g + 32 d + 4 m + k, with the value 0.5, 1, 2, 2 for k = 0..3.  The
// weights put shift k in rows 8 (k % 2) .. + 7 of accumulator k / 2; every
// product and sum is exact in FP32 (digits d / 8 with |d| <= 15).
template <int R, int DC, bool DD>
__device__ __forceinline__ void scr_group_f6(const v4u (&q)[DC][R], int C0, int nch,
                                             const v4i* wl, const v4u* mk_lds, int g, int lane,
                                             v4f (&acc)[R][2], uint32_t (&ne)[R]) {
  typedef int v8i __attribute__((ext_vector_type(8)));
#pragma unroll
  for (int u = 0; u < DC; ++u) {
    if (C0 + u < nch) {
      const int C = C0 + u;
      if (DD) {
        const v4u mk = mk_lds[4 * C + g];
        const bool first_col = (lane & 15) == 0;
        v4u prev = v4u{0u, 0u, 0u, 0u};
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const v4u rr = v4u{ror16(q[u][r].x), ror16(q[u][r].y), ror16(q[u][r].z),
                             ror16(q[u][r].w)};
          const v4u pv = (r > 0 && first_col) ? prev : rr;
          ne[r] |= ((q[u][r].x ^ pv.x) & mk.x) | ((q[u][r].y ^ pv.y) & mk.y) |
                   ((q[u][r].z ^ pv.z) & mk.z) | ((q[u][r].w ^ pv.w) & mk.w);
          prev = rr;
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const v4i a0 = wl[((C * 4 + k) * 64 + lane) * 2];
        const v4i a1 = wl[((C * 4 + k) * 64 + lane) * 2 + 1];
        const v8i a = v8i{a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, 0, 0};
#pragma unroll
        for (int r = 0; r < R; ++r) {
          v8i b;
          if (k < 3) {
            const uint32_t M = 0x11111111u << k;
            b = v8i{(int)(q[u][r].x & M), (int)(q[u][r].y & M), (int)(q[u][r].z & M),
                    (int)(q[u][r].w & M), 0, 0, 0, 0};
          } else {
            b = v8i{(int)((q[u][r].x >> 1) & 0x44444444u), (int)((q[u][r].y >> 1) & 0x44444444u),
                    (int)((q[u][r].z >> 1) & 0x44444444u), (int)((q[u][r].w >> 1) & 0x44444444u),
                    0, 0, 0, 0};
          }
          acc[r][k >> 1] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
              a, b, acc[r][k >> 1], 2, 4, 0, 127, 0, 127);
        }
      }
    }
  }
}

// Decision of the one-trait FP6 screen for the k-mer lane (g, col) owns after
// the reduce-scatter of R tiles: c from the count row, S~ = w (27000 D_01 +
// 30 D_23 + D_4) from the digit sums; keep unless |cor| < tau is certain, sure
// when |cor| >= tau is proven.  acc[r][0] holds shift 0 (rows 0-7, bit value
// 0.5) or 1 (rows 8-15, value 1), acc[r][1] shift 2 or 3 (value 2); a digit d
// weighs d / 8; lane groups 0, 2 hold digits 0-3, groups 1, 3 digit 4 and the
// count row.  The test runs in FP32: every FP32 value is within 2e-6 relative
// of its exact counterpart (a handful of correctly rounded operations on
// positive terms), so the margin 1e-5 keeps "certainly fails" certain; for the
// proof |S~| is scaled down by 4e-6 and h was rounded up, so lo <= |S| before
// its last rounding, and the roundings of lo^2 and of the threshold product
// total below 5e-7 relative.
// One tile's partials: a0 = shift 0 (lanes g < 2) or 1, a1 = shift 2 or 3.
struct F6Part {
  uint32_t x01, x23, x4, cp;
};
__device__ __forceinline__ F6Part f6_partial(const v4f& a0, const v4f& a1, int g) {
  const float m0 = g < 2 ? 16.0f : 8.0f;
  const bool lo_g = (g & 1) == 0;
  int v[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = (int)fmaf(a0[i], m0, a1[i] * 4.0f);
  // D_01 = 30 D_0 + D_1 and D_23 = 30 D_2 + D_3 are exact in int32 (|.| <= 465 n)
  F6Part p;
  p.x01 = lo_g ? (uint32_t)(30 * v[0] + v[1]) : 0u;
  p.x23 = lo_g ? (uint32_t)(30 * v[2] + v[3]) : 0u;
  p.x4 = lo_g ? 0u : (uint32_t)v[0];
  p.cp = lo_g ? 0u : (uint32_t)v[1];
  return p;
}

template <int R>
__device__ __forceinline__ void f6_finish(const F6Part (&pt)[R], const ScreenDesc& D,
                                          int* c_out, bool* keep, bool* sure) {
  uint32_t x01[R], x23[R], x4[R], cpart[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    x01[r] = pt[r].x01;
    x23[r] = pt[r].x23;
    x4[r] = pt[r].x4;
    cpart[r] = pt[r].cp;
  }
  const int c = (int)group_scatter<R>(cpart);
  const int D01 = (int)group_scatter<R>(x01), D23 = (int)group_scatter<R>(x23);
  const int D4 = (int)group_scatter<R>(x4);
  // the digit combination in int64, so S~ rounds once
  const float sA = (float)((int64_t)D01 * 27000 + (int64_t)D23 * 30 + D4);
  const float cf = (float)c, vv = cf * (float)(D.n - c);
  const float aS = fabsf(D.wf[0] * sA);
  const float hi = aS * (1.0f + 4e-6f) + cf * D.hf[0];
  const float lo = aS * (1.0f - 4e-6f) - cf * D.hf[0];
  *keep = D.thrf[0] >= 0.0f && !(hi * hi < D.thrf[0] * vv * (1.0f - 1e-5f));
  *sure = D.thrf[0] >= 0.0f && lo > 0.0f && lo * lo >= D.thrf[0] * vv * (1.0f + 1e-5f);
  *c_out = c;
}

template <int R>
__device__ __forceinline__ void f6_decide(const v4f (&acc)[R][2], const ScreenDesc& D, int g,
                                          int* c_out, bool* keep, bool* sure) {
  F6Part pt[R];
#pragma unroll
  for (int r = 0; r < R; ++r) pt[r] = f6_partial(acc[r][0], acc[r][1], g);
  f6_finish<R>(pt, D, c_out, keep, sure);
}

// Software pipeline over load groups (block, first chunk): the next group's
// rows (the same block's next chunks, or the next block's first) are loaded
// into a second register set while the current group computes, then copied
// into the current set (the copy is the only wait for them).
//
// W = waves per workgroup; MINB = 2 caps the registers at 128 per wave (four
// waves per SIMD), so that two 8-wave workgroups share a CU.
template <int Q, bool CR, bool DD, int W, int MINB>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(MINB == 2 ? 4 : 1)))
void screen_i8(const ScreenDesc* __restrict__ sd, int nshards, int64_t total_blocks) {
  constexpr int NTH = 64 * W;
  constexpr int KG = 16 / Q;   // shifts sharing an accumulator
  constexpr int NA = 8 / KG;   // accumulators per tile
  constexpr int TT = Q / 2;    // trait slots an instantiation covers
  constexpr int R = scr_tiles(Q);
  constexpr int DC = scr_group_chunks(R);
  extern __shared__ v4i scr_lds[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15;
  const v4u* const mk_lds = reinterpret_cast<const v4u*>(scr_lds);
  // output rows 4 g .. 4 g + 3 of accumulator a: shift a KG + g / (Q / 4),
  // trait slots tb and tb + 1 (limbs 0, 1 each); the count row is limb 0 of
  // slot t.  (Q = 4, one trait: FP6 digits instead, scr_group_f6.)
  const int ksh = g / (Q / 4);
  const int tb = 2 * (g % (Q / 4));

  int64_t w = blockIdx.x;
  if (w >= total_blocks) return;
  int s = 0;
  while (s + 1 < nshards && sd[s + 1].blk0 <= w) ++s;
  // first block of the shard after s (read again only when s moves)
  int64_t next0 = (s + 1 < nshards) ? sd[s + 1].blk0 : total_blocks;
  int C0 = 0;
  const void* staged = nullptr;
  int nch_staged = 0;
  // Q = 4 (one trait): FP6 digits x FP4 bits, FP32 accumulators (exact)
  constexpr bool F6 = Q == 4;
  using AccT = std::conditional_t<F6, v4f, v4i>;
  AccT acc[R][NA];
  uint32_t cnt[R], ne[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    cnt[r] = 0;
    ne[r] = 0;
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[r][a] = AccT{0, 0, 0, 0};
  }
  v4u cq[DC][R], nq[DC][R];
  // first group through the prefetch set as well, so that the current set is
  // complete at the loop head on every path
  scr_load<W, R, DC>(sd[s], w, C0, wave, g, col, nq);
#pragma unroll
  for (int u = 0; u < DC; ++u)
#pragma unroll
    for (int r = 0; r < R; ++r) cq[u][r] = nq[u][r];

  while (true) {
    const ScreenDesc& D = sd[s];
    const int nch = D.nchunks;
    const bool last_group = C0 + DC >= nch;
    const int64_t w2 = last_group ? w + gridDim.x : w;
    const int C02 = last_group ? 0 : C0 + DC;
    int s2 = s;
    while (w2 >= next0 && s2 + 1 < nshards) {
      ++s2;
      next0 = (s2 + 1 < nshards) ? sd[s2 + 1].blk0 : total_blocks;
    }
    if (w2 < total_blocks) scr_load<W, R, DC>(sd[s2], w2, C02, wave, g, col, nq);

    if (C0 == 0) {
      if (D.wts != staged) {  // workgroup-uniform: every wave walks the same groups
        __syncthreads();
        const v4i* __restrict__ gw = reinterpret_cast<const v4i*>(D.wts);
        const v4u* __restrict__ gm = reinterpret_cast<const v4u*>(D.mask);
        v4i* const wd = scr_lds + 4 * nch;
        for (int i = threadIdx.x; i < nch * 512; i += NTH) wd[i] = gw[i];
        for (int i = threadIdx.x; i < nch * 4; i += NTH)
          reinterpret_cast<v4u*>(scr_lds)[i] = gm[i];
        __syncthreads();
        staged = D.wts;
        nch_staged = nch;
      }
    }
    const v4i* const wl = scr_lds + 4 * nch_staged;
    if constexpr (F6)
      scr_group_f6<R, DC, DD>(cq, C0, nch, wl, mk_lds, g, lane, acc, ne);
    else
      scr_group<KG, NA, R, DC, CR, DD>(cq, C0, nch, wl, mk_lds, g, lane, acc, cnt, ne);

    if (last_group) {
      const int64_t kbase = ((w - D.blk0) * W + wave) * (16 * R);
      const int64_t kidx = kbase + 16 * (g % R) + col;
      const bool valid = (g < R) && kidx < D.m;
      const int n = D.n;
      int c = 0;
      bool keep = false, sure = false;
      if constexpr (F6) {
        f6_decide<R>(acc, D, g, &c, &keep, &sure);
      } else {
      // exact row totals over the shifts (every row of accumulator a is a
      // multiple of 2^k, k = a KG + ksh), then per-lane partials
      // 254 A_0 + A_1 of trait slots tb, tb + 1 and of the count row
      int part[R][2];
      uint32_t cpart[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int v[4] = {0, 0, 0, 0};
#pragma unroll
        for (int a = 0; a < NA; ++a) {
          const int kk = a * KG + ksh;
          const int sg = (kk == 7) ? -1 : 1;
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] += sg * (acc[r][a][i] >> kk);
        }
        part[r][0] = 254 * v[0] + v[1];
        part[r][1] = 254 * v[2] + v[3];
        if (CR) cpart[r] = (uint32_t)((tb == D.t) ? v[0] : ((tb + 1 == D.t) ? v[2] : 0));
      }
      c = CR ? (int)group_scatter<R>(cpart) : (int)group_scatter<R>(cnt);
      // Decision in FP32: every FP32 value below is within 2e-6 relative of
      // its exact counterpart (a handful of correctly rounded operations on
      // positive terms), so the margin 1e-5 keeps "certainly fails" certain.
      const float cf = (float)c, vv = cf * (float)(n - c);
#pragma unroll
      for (int tr = 0; tr < TT; ++tr) {
        if (tr < D.t) {
          uint32_t x[R];
#pragma unroll
          for (int r = 0; r < R; ++r)
            x[r] = (uint32_t)((tb == tr) ? part[r][0] : ((tb + 1 == tr) ? part[r][1] : 0));
          const float sA = (float)(int)group_scatter<R>(x);
          const float aS = fabsf(D.wf[tr] * sA);
          const float hi = aS * (1.0f + 4e-6f) + cf * D.hf[tr];
          if (D.thrf[tr] >= 0.0f && !(hi * hi < D.thrf[tr] * vv * (1.0f - 1e-5f))) keep = true;
        }
      }
      }  // int8 limbs
      keep = keep && valid && c > 0 && c < n;
      if (DD) {
        // dup bit 16 r + col: every lane group's slice of row kbase + 16 r + col
        // equals the previous row's; rows past m are never duplicates
        unsigned long long dm = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const unsigned long long b = __ballot(ne[r] != 0u);
          const uint32_t d16 = (uint32_t)((b | (b >> 16) | (b >> 32) | (b >> 48)) & 0xFFFFull);
          dm |= (unsigned long long)(~d16 & 0xFFFFu) << (16 * r);
          ne[r] = 0;
        }
        dm &= ~1ull;
        const int64_t nv = D.m - kbase;
        if (nv < 16 * R) dm &= nv <= 0 ? 0ull : ((1ull << nv) - 1ull);
        keep = keep && ((dm >> lane) & 1ull) == 0;
        if (lane == 0) {
          auto* dw = gptr_w(D.dmap) + (kbase >> 5);
          dw[0] = (uint32_t)dm;
          if (R == 4) dw[1] = (uint32_t)(dm >> 32);
        }
      }
      // lane l owns k-mer kbase + l (lanes < 16 R): the ballot is the bitmap
      const unsigned long long bal = __ballot(keep);
      if constexpr (Q == 4) {
        // kept and proven words of the wave's 64 k-mers in one store
        const unsigned long long pb = __ballot(keep && sure);
        if (lane == 0)
          gptr_w(reinterpret_cast<v4u*>(D.bmap))[kbase >> 6] =
              v4u{(uint32_t)bal, (uint32_t)(bal >> 32), (uint32_t)pb, (uint32_t)(pb >> 32)};
      } else {
        if (lane == 0) {
          auto* bm = gptr_w(D.bmap) + (kbase >> 5);
          bm[0] = (uint32_t)bal;
          if (R == 4) bm[1] = (uint32_t)(bal >> 32);
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cnt[r] = 0;
#pragma unroll
        for (int a = 0; a < NA; ++a) acc[r][a] = AccT{0, 0, 0, 0};
      }
    }
    w = w2;
    C0 = C02;
    s = s2;
    if (w >= total_blocks) break;
#pragma unroll
    for (int u = 0; u < DC; ++u)
#pragma unroll
      for (int r = 0; r < R; ++r) cq[u][r] = nq[u][r];
  }
}

// ----------------------------------------------------------------------------
// One-trait screen with its rows streamed through an LDS ring (the same test
// as screen_i8<4, ...>: five FP6 digits of y - ybar and a count row against
// FP4 bits).  Every load instruction of a wave reads a contiguous KiB -- four
// whole 256-byte row segments -- with the nontemporal hint, straight into LDS
// (global_load_lds: nothing held in registers while in flight).  A linear
// nontemporal stream reads at about 7.0 TB/s on this part, against about
// 6.3 TB/s for the 64-byte row slices the register-loaded screen reads (with
// the hint those slices drop to 5.3 TB/s).
//
// Unit = 16 rows x 4 chunks (4 KiB, one ring slot); a wave has NBUF units in
// flight.  Lane L of load instruction u (rows 4u .. 4u + 3 of the unit)
// fetches the 16 bytes of row 4u + rho, chunk C, part g that belong in slot
// L = 16 C + (4 rho + g + u) % 16, so the operand reads (lane (g, col) takes
// row col, chunk C, part g) hit 16 distinct bank groups in every 16
// consecutive lanes.  A wave block is 4 tiles of 16 k-mers (64 k-mers), a
// workgroup block W of those; units run tile by tile, segment (4 chunks) by
// segment.  Models of 4 or 8 chunks (kRingMaxSeg segments).
//
// WREG (4-chunk models): the model's weight fragments stay in registers,
// loaded from global memory by each wave when its model changes, so the LDS
// serves only the rows (and has room for more ring slots); the LDS weight reads
// would otherwise be 2 KiB per k-mer, since a ring slot holds one tile and
// every fragment would be read once per tile.  Shift k fills rows 8 (k % 2) ..
// + 5 only, so a lane (row i) holds a nonzero fragment for shifts of parity
// i / 8 only: it keeps those two per chunk, each shift accumulates on its own
// (the rows of the other half then collect products that are never read), and
// each tile is reduced to the decision's partials as soon as it is complete.
//
// The wave's bitmap words collect in an LDS stash and go out in batches: a
// global store between the ring's loads would leave the vmcnt wait that
// tracks them unreliable (stores and loads may retire out of order with
// respect to each other), so each batch store is followed by a full wait.
constexpr int kRingMaxSeg = 2;
constexpr int kRingStash = 32;                // bitmap records per wave between stores
constexpr int kRingAux = 2;                   // cache policy of the ring loads: nt
// LDS: the accession mask, the weights (none with WREG), the ring, the stash
__host__ __device__ constexpr size_t ring_lds_bytes(int W, int NBUF, int nch, bool wreg) {
  return (size_t)nch * 64 + (wreg ? 0 : (size_t)nch * 8 * 64 * 16) + (size_t)W * NBUF * 4096 +
         (size_t)W * kRingStash * 40;
}

template <int W, int NBUF, bool DD, bool WREG>
__global__ __launch_bounds__(64 * W) void screen_ring(const ScreenDesc* __restrict__ sd,
                                                      int nshards, int64_t total_blocks,
                                                      int nch_max) {
  extern __shared__ v4i scr_lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // scalar row arithmetic
  const int g = lane >> 4, col = lane & 15;
  const v4u* const mk_lds = reinterpret_cast<const v4u*>(scr_lds);
  v4i* const wl = scr_lds + 4 * nch_max;
  const int nwl = WREG ? 0 : 512 * nch_max;  // v4i of LDS weights
  v4u* const ring = reinterpret_cast<v4u*>(wl + nwl) + (size_t)wave * NBUF * 256;
  // stash: [W][kRingStash] 16-byte words, then their addresses, then (DD) dup
  // words and addresses
  v4u* const st_w = reinterpret_cast<v4u*>(wl + nwl) + (size_t)W * NBUF * 256 +
                    (size_t)wave * kRingStash;
  uint64_t* const st_a = reinterpret_cast<uint64_t*>(reinterpret_cast<v4u*>(wl + nwl) +
                                                     (size_t)W * NBUF * 256 + W * kRingStash) +
                         (size_t)wave * kRingStash;
  uint64_t* const st_d = st_a + (size_t)W * kRingStash;
  uint64_t* const st_da = st_d + (size_t)W * kRingStash;

  int64_t w = blockIdx.x;
  if (w >= total_blocks) return;
  // consumer shard
  int s = 0;
  while (s + 1 < nshards && sd[s + 1].blk0 <= w) ++s;
  // producer position: block, shard, tile, segment; `done` once past the last unit
  int64_t pw = w;
  int ps = s, pt = 0, pg = 0;
  bool pdone = false;
  // this lane's 16 bytes of load u (the KiB of rows 4u .. 4u + 3 of a unit):
  // row 4u + rho, chunk C = lane / 16, part g, so that it lands in slot lane
  int lrow[4], lcol[4];  // row within the unit; offset within the row (8-byte words)
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int x = ((lane & 15) - u) & 15;
    lrow[u] = 4 * u + (x >> 2);
    lcol[u] = 8 * (lane >> 4) + 2 * (x & 3);
  }
  // the producer's shard, reloaded when it changes: rows, row stride, last row,
  // segments, first block, first block of the next shard; per-lane offsets
  const uint64_t* p_bits = nullptr;
  uint32_t p_rw = 0, p_last = 0;
  int p_nseg = 1;
  int64_t p_blk0 = 0, p_next0 = 0;
  uint32_t loff[4];
  auto load_p = [&]() {
    const ScreenDesc& P = sd[ps];
    p_bits = P.bits;
    p_rw = (uint32_t)P.rw;
    p_last = (uint32_t)(P.m - 1);
    p_nseg = P.nchunks / 4;
    p_blk0 = P.blk0;
    p_next0 = (ps + 1 < nshards) ? sd[ps + 1].blk0 : total_blocks;
#pragma unroll
    for (int u = 0; u < 4; ++u) loff[u] = (uint32_t)lrow[u] * p_rw + (uint32_t)lcol[u];
  };
  load_p();
  auto issue = [&](int buf) {
    const uint32_t kbase = (uint32_t)(((pw - p_blk0) * W + wave) * 64 + 16 * pt);
    if (kbase + 15 <= p_last) {  // every row of the unit exists: one add per load
      const uint64_t* rb = p_bits + (uint64_t)kbase * p_rw + 32 * pg;
#pragma unroll
      for (int u = 0; u < 4; ++u)
        __builtin_amdgcn_global_load_lds(
            (const void*)(rb + loff[u]),
            (__attribute__((address_space(3))) void*)(ring + buf * 256 + u * 64), 16, 0, kRingAux);
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        uint32_t k = kbase + (uint32_t)lrow[u];
        if (k > p_last) k = p_last;  // rows past m are read, never kept
        const uint64_t* src = p_bits + (uint64_t)k * p_rw + 32 * pg + lcol[u];
        __builtin_amdgcn_global_load_lds(
            (const void*)src, (__attribute__((address_space(3))) void*)(ring + buf * 256 + u * 64),
            16, 0, kRingAux);
      }
    }
  };
  auto advance = [&]() {
    if (pdone) return;
    if (++pg < p_nseg) return;
    pg = 0;
    if (++pt < 4) return;
    pt = 0;
    pw += gridDim.x;
    if (pw >= total_blocks) {
      pdone = true;  // the last unit is issued again (into a slot never read)
      pw -= gridDim.x;
      pt = 3;
      pg = p_nseg - 1;
      return;
    }
    if (pw >= p_next0) {
      while (ps + 1 < nshards && sd[ps + 1].blk0 <= pw) ++ps;
      load_p();
    }
  };
  int jb = 0;  // ring slot of the unit the consumer takes next
#pragma unroll
  for (int i = 0; i < NBUF - 1; ++i) {
    issue(i);
    advance();
  }

  const void* staged = nullptr;
  int nst = 0;  // stashed records of this wave
  auto flush = [&]() {
    if (lane < nst) {
      *gptr_w(reinterpret_cast<v4u*>(st_a[lane])) = st_w[lane];
      if (DD) *gptr_w(reinterpret_cast<uint64_t*>(st_da[lane])) = st_d[lane];
    }
    __builtin_amdgcn_s_waitcnt(0);  // the stores, and every ring load in flight
    nst = 0;
  };
  v4f acc[WREG ? 1 : 4][2];  // LDS weights: the block's tiles, shifts 0/1 and 2/3
  v4f acc4[WREG ? 4 : 1];     // WREG: the current tile, one accumulator per shift
  F6Part part[4];             // WREG: the block's finished tiles
#pragma unroll
  for (int r = 0; r < (WREG ? 1 : 4); ++r) acc[r][0] = acc[r][1] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < (WREG ? 4 : 1); ++k) acc4[k] = v4f{0.f, 0.f, 0.f, 0.f};
  uint32_t ne[4] = {0, 0, 0, 0};
  typedef int v8i __attribute__((ext_vector_type(8)));
  // WREG: [chunk][shifts 0/1, 2/3] fragments of this lane's row half
  v8i wm[WREG ? 4 : 1][WREG ? 2 : 1];
  const int hrow = (lane & 15) >> 3;  // 1: this lane holds a row of the upper half

  while (true) {
    const ScreenDesc& D = sd[s];
    const int nseg = WREG ? 1 : D.nchunks / 4;
    if (WREG && D.wts != staged) {  // this wave's fragments, straight from global memory
      const auto* gw = gptr(reinterpret_cast<const v4i*>(D.wts));
#pragma unroll
      for (int C = 0; C < 4; ++C)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int k = 2 * h + hrow;
          const v4i a0 = gw[((C * 4 + k) * 64 + lane) * 2];
          const v4i a1 = gw[((C * 4 + k) * 64 + lane) * 2 + 1];
          wm[C][h] = v8i{a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, 0, 0};
        }
      // every load in flight, the ring's included, has landed: the explicit
      // counts of the ring waits then cover ring loads only
      __builtin_amdgcn_s_waitcnt(0);
      asm volatile("" ::: "memory");
      staged = D.wts;
    }
    if (!WREG && D.wts != staged) {  // workgroup-uniform: every wave walks the same blocks
      __syncthreads();
      const v4i* __restrict__ gw = reinterpret_cast<const v4i*>(D.wts);
      const v4u* __restrict__ gm = reinterpret_cast<const v4u*>(D.mask);
      for (int i = threadIdx.x; i < D.nchunks * 512; i += 64 * W) wl[i] = gw[i];
      for (int i = threadIdx.x; i < D.nchunks * 4; i += 64 * W)
        reinterpret_cast<v4u*>(scr_lds)[i] = gm[i];
      __syncthreads();
      staged = D.wts;
    }
    v4u prev[kRingMaxSeg][4];  // DD: previous tile's rotated rows (lane 0: its row 15)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int sg = 0; sg < kRingMaxSeg; ++sg) {
        if (sg < nseg) {
          // the unit NBUF - 1 ahead, into the slot the previous unit was read
          // from: those reads have returned first (lgkmcnt(0)), and the fence
          // keeps the refill below that wait
          __builtin_amdgcn_s_waitcnt(0xC07F);
          asm volatile("" ::: "memory");
          issue(jb == 0 ? NBUF - 1 : jb - 1);
          advance();
          constexpr int K = 4 * (NBUF - 1);
          __builtin_amdgcn_s_waitcnt((K & 15) | (7 << 4) | (15 << 8) | ((K >> 4) << 14));
          asm volatile("" ::: "memory");  // the slot's reads stay after the wait
          const v4u* ub = ring + jb * 256 + 64 * (col >> 2);
          v4u q[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) q[c] = ub[16 * c + ((4 * (col & 3) + g + (col >> 2)) & 15)];
          jb = jb + 1 == NBUF ? 0 : jb + 1;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int C = 4 * sg + c;
            if (DD) {
              const v4u mk = mk_lds[4 * C + g];
              const v4u rr = v4u{ror16(q[c].x), ror16(q[c].y), ror16(q[c].z), ror16(q[c].w)};
              const v4u pv = (t > 0 && col == 0) ? prev[sg][c] : rr;
              ne[t] |= ((q[c].x ^ pv.x) & mk.x) | ((q[c].y ^ pv.y) & mk.y) |
                       ((q[c].z ^ pv.z) & mk.z) | ((q[c].w ^ pv.w) & mk.w);
              prev[sg][c] = rr;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              v8i a;
              if constexpr (WREG) {
                a = wm[c][k >> 1];
              } else {
                const v4i a0 = wl[((C * 4 + k) * 64 + lane) * 2];
                const v4i a1 = wl[((C * 4 + k) * 64 + lane) * 2 + 1];
                a = v8i{a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, 0, 0};
              }
              v8i b;
              if (k < 3) {
                const uint32_t M = 0x11111111u << k;
                b = v8i{(int)(q[c].x & M), (int)(q[c].y & M), (int)(q[c].z & M),
                        (int)(q[c].w & M), 0, 0, 0, 0};
              } else {
                b = v8i{(int)((q[c].x >> 1) & 0x44444444u), (int)((q[c].y >> 1) & 0x44444444u),
                        (int)((q[c].z >> 1) & 0x44444444u), (int)((q[c].w >> 1) & 0x44444444u),
                        0, 0, 0, 0};
              }
              if constexpr (WREG)
                acc4[k] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, acc4[k], 2, 4,
                                                                          0, 127, 0, 127);
              else
                acc[t][k >> 1] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                    a, b, acc[t][k >> 1], 2, 4, 0, 127, 0, 127);
            }
          }
        }
      }
      if constexpr (WREG) {
        // rows 0-7 (lane groups 0, 1): shifts 0 and 2; rows 8-15: shifts 1 and 3
        part[t] = f6_partial(g < 2 ? acc4[0] : acc4[1], g < 2 ? acc4[2] : acc4[3], g);
#pragma unroll
        for (int k = 0; k < 4; ++k) acc4[k] = v4f{0.f, 0.f, 0.f, 0.f};
      }
    }
    // the block's decision (as screen_i8<4, ...>)
    const int64_t kbase = ((w - D.blk0) * W + wave) * 64;
    const int64_t kidx = kbase + 16 * g + col;
    int c = 0;
    bool keep = false, sure = false;
    if constexpr (WREG)
      f6_finish<4>(part, D, &c, &keep, &sure);
    else
      f6_decide<4>(acc, D, g, &c, &keep, &sure);
    keep = keep && kidx < D.m && c > 0 && c < D.n;
    unsigned long long dm = 0;
    if (DD) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const unsigned long long b = __ballot(ne[r] != 0u);
        const uint32_t d16 = (uint32_t)((b | (b >> 16) | (b >> 32) | (b >> 48)) & 0xFFFFull);
        dm |= (unsigned long long)(~d16 & 0xFFFFu) << (16 * r);
        ne[r] = 0;
      }
      dm &= ~1ull;
      const int64_t nv = D.m - kbase;
      if (nv < 64) dm &= nv <= 0 ? 0ull : ((1ull << nv) - 1ull);
      keep = keep && ((dm >> lane) & 1ull) == 0;
    }
    const unsigned long long bal = __ballot(keep);
    const unsigned long long pb = __ballot(keep && sure);
    if (lane == 0) {
      st_w[nst] = v4u{(uint32_t)bal, (uint32_t)(bal >> 32), (uint32_t)pb, (uint32_t)(pb >> 32)};
      st_a[nst] = (uint64_t)(reinterpret_cast<v4u*>(D.bmap) + (kbase >> 6));
      if (DD) {
        st_d[nst] = dm;
        st_da[nst] = (uint64_t)(D.dmap + (kbase >> 5));
      }
    }
    ++nst;
#pragma unroll
    for (int r = 0; r < (WREG ? 1 : 4); ++r) acc[r][0] = acc[r][1] = v4f{0.f, 0.f, 0.f, 0.f};
    w += gridDim.x;
    if (w >= total_blocks) break;
    if (nst == kRingStash) flush();
    while (s + 1 < nshards && sd[s + 1].blk0 <= w) ++s;
  }
  flush();
}

// Physical index of kept-bitmap word w: one trait interleaves each 64 k-mers'
// two kept words with their two proven words (kept lo, kept hi, proven lo,
// proven hi), so the screen writes them in one store.
__device__ __forceinline__ int64_t scr_word(int64_t w, bool ilv) {
  return ilv ? 2 * w - (w & 1) : w;
}

// Compaction of the screen bitmaps into row lists, in three launches over
// chunks of at most kScrChunkWords bitmap words (a chunk lies in one shard):
//   scr_count: kept k-mers per chunk;
//   scr_scan : one workgroup: exclusive prefix P over the chunks, survivor
//              counts per shard, and the full kernel's block table (sd2[s].m,
//              .blk0 in blocks of bk k-mers, and the block total);
//   scr_write: each chunk writes its rows at P[c] - P[chunk0 of its shard].
__global__ __launch_bounds__(256) void scr_count(const ScrChunk* __restrict__ ch,
                                                 const ScreenDesc* __restrict__ ss,
                                                 uint32_t* __restrict__ ccount) {
  __shared__ uint32_t red[4];
  const ScrChunk C = ch[blockIdx.x];
  const uint32_t* bm = ss[C.shard].bmap;
  const bool ilv = ss[C.shard].ilv != 0;
  uint32_t v = 0;
  for (int64_t w = C.w0 + threadIdx.x; w < C.w1; w += 256) v += __popc(bm[scr_word(w, ilv)]);
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) ccount[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(1024) void scr_scan(const uint32_t* __restrict__ ccount, int ncc,
                                                 int64_t* __restrict__ P,
                                                 const ScreenDesc* __restrict__ ss,
                                                 ShardDesc* __restrict__ sd2, int nsh, int64_t bk,
                                                 int64_t* __restrict__ total) {
  __shared__ int64_t part[1024];
  const int tid = threadIdx.x;
  // P[0..ncc]: exclusive prefix of the chunk counts
  int64_t carry = 0;
  for (int base = 0; base < ncc; base += 1024) {
    const int i = base + tid;
    const int64_t v = (i < ncc) ? (int64_t)ccount[i] : 0;
    part[tid] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int64_t add = (tid >= off) ? part[tid - off] : 0;
      __syncthreads();
      part[tid] += add;
      __syncthreads();
    }
    if (i < ncc) P[i] = carry + part[tid] - v;
    const int64_t blk = part[1023];
    __syncthreads();
    carry += blk;
  }
  if (tid == 0) P[ncc] = carry;
  __syncthreads();
  // survivors per shard, then the full kernel's block table
  const int per = (nsh + 1023) / 1024;
  const int s0 = tid * per, s1 = min(nsh, s0 + per);
  int64_t sum = 0;
  for (int s = s0; s < s1; ++s) {
    const int64_t m = P[ss[s].chunk1] - P[ss[s].chunk0];
    *ss[s].nsurv = (unsigned long long)m;
    sum += (m + bk - 1) / bk;
  }
  part[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int64_t add = (tid >= off) ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += add;
    __syncthreads();
  }
  int64_t run = part[tid] - sum;
  for (int s = s0; s < s1; ++s) {
    const int64_t m = P[ss[s].chunk1] - P[ss[s].chunk0];
    sd2[s].m = m;
    sd2[s].blk0 = run;
    run += (m + bk - 1) / bk;
  }
  if (tid == 1023) *total = part[1023];
}

// Wave v of chunk c owns a quarter of its words (at most 512); a pass takes 64
// words, one per lane: a wave prefix sum of their popcounts places each word's
// rows, and each lane writes its own word's rows (lowest bit first), so the
// list is in row order.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t* total) {
  const int lane = __lane_id();
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off);
    if (lane >= off) x += y;
  }
  *total = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
  return x - v;
}

__global__ __launch_bounds__(256) void scr_write(const ScrChunk* __restrict__ ch,
                                                 const ScreenDesc* __restrict__ ss,
                                                 const int64_t* __restrict__ P) {
  static_assert(kScrChunkWords == 4 * 512, "a wave's quarter is 8 passes of 64 words");
  __shared__ uint32_t wcnt[4];
  const ScrChunk C = ch[blockIdx.x];
  const ScreenDesc& D = ss[C.shard];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t per = (C.w1 - C.w0 + 3) / 4;
  const int64_t a0 = min(C.w1, C.w0 + wave * per), a1 = min(C.w1, a0 + per);
  const int nw = (int)(a1 - a0);
  const uint32_t* __restrict__ bm = D.bmap;
  const bool dd = D.mult != nullptr;  // duplicate-row scan: write run lengths too
  const bool ilv = D.ilv != 0;        // one trait: proven passes set the entry's top bit
  uint32_t wv[8];
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int j = 64 * i + lane;
    wv[i] = (j < nw) ? bm[scr_word(a0 + j, ilv)] : 0u;
    v += __popc(wv[i]);
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (lane == 0) wcnt[wave] = v;
  __syncthreads();
  int64_t o = P[blockIdx.x] - P[D.chunk0];
  for (int i = 0; i < wave; ++i) o += wcnt[i];
  uint32_t* __restrict__ list = D.list;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (64 * i >= nw) break;
    const int j = 64 * i + lane;
    const int64_t wi = a0 + j;  // this lane's word
    uint32_t word = wv[i];
    uint32_t total;
    const uint32_t pre = wave_excl_scan((uint32_t)__popc(word), &total);
    const uint32_t pw = (ilv && word != 0u) ? bm[scr_word(wi, true) + 2] : 0u;
    const uint32_t dword = (dd && word != 0u) ? D.dmap[wi] : 0u;
    int64_t pos = o + pre;
    while (word != 0u) {
      const int b = __builtin_ctz(word);
      word &= word - 1u;
      list[pos] = (uint32_t)(32 * wi + b) | (((pw >> b) & 1u) << 31);
      if (dd) {
        // the run of dup bits after this row: to the end of its word, then
        // into the next (a run never crosses a wave's first row, whose dup
        // bit is clear, so two words hold it)
        uint32_t run = b < 31 ? (uint32_t)__builtin_ctz(~(dword >> (b + 1))) : 0u;
        if (run == (uint32_t)(31 - b) && (wi + 1) * 32 < D.m) {
          const uint32_t nx = D.dmap[wi + 1];
          run += nx == 0xFFFFFFFFu ? 32u : (uint32_t)__builtin_ctz(~nx);
        }
        D.mult[pos] = 1u + run;
      }
      ++pos;
    }
    o += total;
  }
}

// ----------------------------------------------------------------------------
// Plain FP32 scan.  Columns: [0,p) U, [p,p+t) y, [p+t,p+2t) y - mean(y).
template <int NT, int R>
__global__ __launch_bounds__(kThreads, 2) void scan_f32(
    const uint64_t* __restrict__ bits, int64_t m, int64_t row_words, int64_t row0,
    int nchunks, const float* __restrict__ wts, const uint64_t* __restrict__ maskw,
    EpiF32 ep, ScanOut out, int64_t nblocks) {
  constexpr int NU = 16 * NT;
  constexpr int FRAG = 64 * NT * 64;  // floats per 256-sample chunk
  extern __shared__ float ldsf[];
  __shared__ unsigned int wg_pass[kMaxTraits];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15;
  const int kk = lane >> 2, q = lane & 3;
  if (threadIdx.x < kMaxTraits) wg_pass[threadIdx.x] = 0;

  for (int64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
    const int64_t kbase = (blk * kWaves + wave) * (16 * R);
    v4f acc[R][NT];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[r][t] = v4f{0.f, 0.f, 0.f, 0.f};
    uint32_t cnt[R];
    const uint64_t* rowp[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cnt[r] = 0;
      int64_t k = kbase + 16 * r + col;
      if (k > m - 1) k = m - 1;
      rowp[r] = bits + k * row_words + g;
    }
    for (int C = 0; C < nchunks; ++C) {
      uint64_t raw[R];
#pragma unroll
      for (int r = 0; r < R; ++r) raw[r] = rowp[r][4 * C];
      const uint64_t mk = maskw[4 * C + g];
      __syncthreads();
      const float* src = wts + (size_t)C * FRAG;
      for (int i = threadIdx.x; i < FRAG / 4; i += kThreads)
        reinterpret_cast<float4*>(ldsf)[i] = reinterpret_cast<const float4*>(src)[i];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        raw[r] &= mk;
        cnt[r] += (uint32_t)__popcll(raw[r]);
      }
      __syncthreads();
#pragma unroll 4
      for (int e = 0; e < 64; ++e) {
        float a[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) a[t] = ldsf[(e * NT + t) * 64 + lane];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const float b = (float)(uint32_t)((raw[r] >> e) & 1ull);
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t], b, acc[r][t], 0, 0, 0);
        }
      }
    }

    __syncthreads();
    float* sc = ldsf + wave * 16 * NU;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      uint32_t cc = cnt[r];
      cc += __shfl_xor(cc, 16);
      cc += __shfl_xor(cc, 32);
#pragma unroll
      for (int t = 0; t < NT; ++t)
        *reinterpret_cast<v4f*>(sc + col * NU + 16 * t + 4 * g) = acc[r][t];
      __syncthreads();
      const int c = (int)__shfl(cc, kk);
      const int64_t kidx = kbase + 16 * r + kk;
      const bool valid = (kidx < m);
      const float* drow = sc + kk * NU;
      for (int tr = 0; tr < ep.t; ++tr) {
        // src/kgwas.cpp:153-158 in FP32 (partial sums over lanes q)
        float sux = 0.f, uyux = 0.f;
        for (int k = q; k < ep.p; k += 4) {
          const float v = drow[k];
          sux += v * v;
          uyux += v * ep.Uyf[tr * ep.p + k];
        }
        sux += __shfl_xor(sux, 1);
        sux += __shfl_xor(sux, 2);
        uyux += __shfl_xor(uyux, 1);
        uyux += __shfl_xor(uyux, 2);
        const float Xy = drow[ep.p + tr];
        const float Syc = drow[ep.p + ep.t + tr];
        const float xyM = Xy - uyux;
        const float xxM = (float)c - sux;  // colwise squaredNorm of 0/1 = c
        // pre-filter, src/kgwas.cpp:187-199 in FP32
        const float fn = (float)ep.n;
        const float mean = (float)c / fn;
        const float xn2 = (float)c * (1.f - mean) * (1.f - mean) + (fn - (float)c) * mean * mean;
        const float xnorm = sqrtf(xn2);
        const float dot = Syc - mean * ep.ycsum_f[tr];
        bool pass = false;
        if (!(xnorm < 1e-12f || ep.ynorm_f[tr] < 1e-12f)) {
          const float cor = dot / (xnorm * ep.ynorm_f[tr]);
          pass = ep.tau_trivial ? true : (fabsf(cor) >= ep.tau);
        }
        pass = pass && valid && q == 0;
        bool cand = false;
        if (pass) {
          if (ep.all_sig) {
            cand = true;
          } else {
            const double xx = (double)xxM, xy = (double)xyM, yy = ep.yyM[tr];
            if (xx > 1e-12 && yy > 1e-12) {
              double r2 = (xy * xy) / (xx * yy);
              if (r2 < 0) r2 = 0;
              if (r2 >= 1) r2 = 1 - 1e-15;
              const double T = (double)ep.n * log(1.0 / (1.0 - r2));
              cand = (T >= ep.Tstar_lo);
            }
          }
        }
        if (out.st_c != nullptr && q == 0 && valid) {
          const int64_t o = kidx - row0;
          if (tr == 0) {
            out.st_c[o] = c;
            out.st_xx[o] = (double)xxM;
          }
          out.st_xy[o * ep.t + tr] = (double)xyM;
          out.st_S[o * ep.t + tr] = (double)dot;
        }
        const unsigned long long pb = __ballot(pass);
        if (lane == 0 && pb) atomicAdd(&wg_pass[tr], (unsigned int)__popcll(pb));
        Cand rec;
        rec.idx = kidx;
        rec.trait = tr;
        rec.flags = kPfPass | kSigAmb;
        rec.shard = 0;
        rec.mult = 1;
        rec.xx32 = xxM;
        rec.xy32 = xyM;
        rec.c = c;
        rec.pad = 0;
        append_cand(cand, rec, out);
      }
      __syncthreads();
    }
  }
  __syncthreads();
  if (threadIdx.x < ep.t && wg_pass[threadIdx.x])
    atomicAdd(&out.counters[1 + threadIdx.x], (unsigned long long)wg_pass[threadIdx.x]);
}

// ----------------------------------------------------------------------------
// Exact FP64 statistic of each candidate.  One workgroup of kRefWaves waves
// per candidate: wave v sums the columns over the set, phenotyped bits of the
// v-th part of the row's words in accession order (lane b owns base column b
// and b + 64), and wave 0 adds the parts in order, so the result is
// deterministic.  A candidate's time is a chain of dependent column loads, 8
// set bits per round, so more waves per candidate mean fewer rounds.  The
// candidate count is read on the device (so the launch can sit in a captured
// graph).  Columns of RefDesc::B64: [0,p) rotated centred U, [p, p+t) centred
// residuals.  res[i] = {c, xxM, xyM, S}.  The first hb.eager candidates and
// results, and the counters, are also written to hb's page-locked host
// buffers, which the host reads after its one synchronisation.
__global__ __launch_bounds__(64 * kRefWaves) void refine_f64(
    const Cand* __restrict__ cand, const unsigned long long* __restrict__ count, int64_t cap,
    const RefDesc* __restrict__ rd, double* __restrict__ res, RefHost hb) {
  __shared__ double pa[kRefWaves][2][64];
  __shared__ int pc[kRefWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the counters are final: the kernels that count ran before this one
  if (blockIdx.x == 0 && hb.cnt)
    for (int i = threadIdx.x; i < hb.ncnt; i += blockDim.x) hb.cnt[i] = count[i];
  int64_t ncand = (int64_t)count[0];
  if (ncand > cap) ncand = cap;
  for (int64_t i = blockIdx.x; i < ncand; i += gridDim.x) {
    const Cand cd = cand[i];
    const RefDesc& R = rd[cd.shard];
    const uint64_t* row = R.bits + cd.idx * R.rw;
    const int nbf = R.p + R.t, p = R.p;
    const int per = (R.nwords + kRefWaves - 1) / kRefWaves;
    const int wb = min(R.nwords, wave * per), we = min(R.nwords, wb + per);
    double a0 = 0.0, a1 = 0.0;
    int c = 0;
    for (int w = wb; w < we; ++w) {
      uint64_t word = row[w] & R.maskw[w];
      c += __popcll(word);
      // set bits in batches of 8: the batch's column loads are all issued
      // before its adds, which still run in accession order
      while (word) {
        int bit[8];
        int nb = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          bit[j] = 0;
          if (word) {
            bit[j] = __ffsll((long long)word) - 1;
            word &= word - 1;
            nb = j + 1;
          }
        }
        double v0[8], v1[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const double* Br = R.B64 + (size_t)(64 * w + bit[j]) * nbf;
          v0[j] = (j < nb && lane < nbf) ? Br[lane] : 0.0;
          v1[j] = (j < nb && lane + 64 < nbf) ? Br[lane + 64] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if (j < nb) {
            a0 += v0[j];
            a1 += v1[j];
          }
        }
      }
    }
    pa[wave][0][lane] = a0;
    pa[wave][1][lane] = a1;
    if (lane == 0) pc[wave] = c;
    __syncthreads();
    if (wave == 0) {
      a0 = pa[0][0][lane];
      a1 = pa[0][1][lane];
      c = pc[0];
      for (int v = 1; v < kRefWaves; ++v) {
        a0 += pa[v][0][lane];
        a1 += pa[v][1][lane];
        c += pc[v];
      }
      const int tr = cd.trait;
      double zz = 0.0, zb = 0.0, z0 = 0.0, sv = 0.0;
      if (lane < p) {
        zz += a0 * a0;
        zb += a0 * R.bpfull[tr * p + lane];
        if (lane == 0) z0 = a0;
      }
      if (lane + 64 < p) {
        zz += a1 * a1;
        zb += a1 * R.bpfull[tr * p + lane + 64];
      }
      if (lane == p + tr) sv = a0;
      if (lane + 64 == p + tr) sv = a1;
      for (int off = 32; off > 0; off >>= 1) {
        zz += __shfl_xor(zz, off);
        zb += __shfl_xor(zb, off);
        z0 += __shfl_xor(z0, off);
        sv += __shfl_xor(sv, off);
      }
      if (lane == 0) {
        const double cdv = (double)c, cn = cdv / (double)R.n;
        const double xxM = cdv - cdv * cdv * R.base_coef - 2.0 * cn * R.sigma * z0 - zz;
        const double r4[4] = {cdv, xxM, sv + cdv * R.rbar[tr], sv + zb};
        for (int k = 0; k < 4; ++k) res[4 * i + k] = r4[k];
        if (i < hb.eager) {
          for (int k = 0; k < 4; ++k) hb.res[4 * i + k] = r4[k];
          hb.cand[i] = cd;
        }
      }
    }
    __syncthreads();  // the partials are rewritten for the next candidate
  }
}

__global__ void zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Accession masks of an all-pairs scan (src/kgwas_2d.cpp:372: x_ij = x_i * x_j):
// out[k][w] = mask[w] & row (i0 + k)[w], w < W, for k < cnt.  The pair (i, j)
// is then row j of the scan under mask k = i - i0.
__global__ void pair_masks(const uint64_t* __restrict__ bits, int64_t rw,
                           const uint64_t* __restrict__ mask, int64_t W, int64_t i0, int64_t cnt,
                           uint64_t* __restrict__ out) {
  const int64_t total = cnt * W;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e / W, w = e - k * W;
    out[e] = mask[w] & bits[(i0 + k) * rw + w];
  }
}

__global__ void synth_bits(uint64_t* __restrict__ out, uint64_t seed, int64_t kmer0, int64_t m,
                           uint32_t n_acc, int64_t row_words, const uint8_t* __restrict__ pop,
                           const uint32_t* __restrict__ thr, uint32_t npop, uint32_t nprof) {
  const int64_t total = m * row_words;
  const uint32_t used = (n_acc + 63u) / 64u;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = i / row_words;
    const uint32_t w = (uint32_t)(i - k * row_words);
    out[i] = (w < used) ? ks_word(seed, (uint64_t)(kmer0 + k), w, n_acc, pop, thr, npop, nprof)
                        : 0ull;
  }
}

// ----------------------------------------------------------------------------
// Launchers (host).  The tile count NT is a template parameter.

template <int NT, int L, int TT, int R>
static hipError_t launch_i8_tr(dim3 grid, hipStream_t s, const ShardDesc* sd, int nshards,
                               int64_t total_blocks, const ScanOut& o, const int64_t* total_dev) {
  auto kfn = scan_i8<NT, L, TT, R, false>;
  hipLaunchKernelGGL(kfn, grid, dim3(kI8Threads), 0, s, sd, nshards, total_blocks, o, total_dev);
  return hipGetLastError();
}

// Resident-weight instantiations: two limbs, one trait, 4 tiles per wave,
// NT = 1..kResMaxTiles; the dynamic LDS holds nchunks x NT x 8 KiB.
constexpr int kResMaxTiles = 4;
template <int NT>
static hipError_t launch_res_nt(dim3 grid, size_t shm, hipStream_t s, const ShardDesc* sd,
                                int nshards, int64_t total_blocks, const ScanOut& o,
                                const int64_t* total_dev) {
  auto kfn = scan_i8<NT, 2, 1, 4, true>;
  hipLaunchKernelGGL(kfn, grid, dim3(kI8Threads), shm, s, sd, nshards, total_blocks, o, total_dev);
  return hipGetLastError();
}

size_t i8_res_lds_bytes(int NT, int nch) { return (size_t)nch * NT * 8 * 64 * 16; }

bool i8_res_fits(int NT, int nch) {
  return NT >= 1 && NT <= kResMaxTiles && i8_res_lds_bytes(NT, nch) <= kResLdsBytes;
}

// k-mers per workgroup block of the resident kernel
int64_t i8_res_block_kmers() { return (int64_t)kI8Waves * 16 * 4; }

// FP6-digit resident instantiations: D = 2 or 3 digits per column, NT = 1..6,
// count mode CM (kCnt*).
constexpr int kF6MaxTiles = 6;
template <int NT, int D, int CM>
static hipError_t launch_f6_nt(dim3 grid, size_t shm, hipStream_t s, const ShardDesc* sd,
                               int nshards, int64_t total_blocks, const ScanOut& o,
                               const int64_t* total_dev) {
  auto kfn = scan_i8<NT, D, 1, 4, true, true, CM>;
  hipLaunchKernelGGL(kfn, grid, dim3(kI8Threads), shm, s, sd, nshards, total_blocks, o, total_dev);
  return hipGetLastError();
}

size_t f6_res_lds_bytes(int NT, int nch) { return (size_t)nch * 4 * NT * 64 * kF6OperandBytes; }

bool f6_res_fits(int NT, int nch) {
  return NT >= 1 && NT <= kF6MaxTiles && f6_res_lds_bytes(NT, nch) <= kResLdsBytes;
}

hipError_t launch_scan_f6_res(int NT, int D, int cm, int nch_max, int grid, hipStream_t s,
                              const ShardDesc* sd, int nshards, int64_t total_blocks,
                              const ScanOut& o, const int64_t* total_dev) {
  if (!f6_res_fits(NT, nch_max) || (D != 2 && D != 3) || cm < kCntMasked || cm > kCntColumn)
    return hipErrorInvalidValue;
  dim3 g((unsigned)(total_blocks < grid ? (total_blocks > 0 ? total_blocks : 1) : grid));
  const size_t shm = f6_res_lds_bytes(NT, nch_max);
  switch ((NT * 4 + D) * 4 + cm) {
#define KG_CASE(X, Y)                                                                       \
  case (X * 4 + Y) * 4 + kCntMasked:                                                        \
    return launch_f6_nt<X, Y, kCntMasked>(g, shm, s, sd, nshards, total_blocks, o, total_dev); \
  case (X * 4 + Y) * 4 + kCntBits:                                                          \
    return launch_f6_nt<X, Y, kCntBits>(g, shm, s, sd, nshards, total_blocks, o, total_dev);   \
  case (X * 4 + Y) * 4 + kCntColumn:                                                        \
    return launch_f6_nt<X, Y, kCntColumn>(g, shm, s, sd, nshards, total_blocks, o, total_dev);
    KG_CASE(1, 2) KG_CASE(2, 2) KG_CASE(3, 2) KG_CASE(4, 2) KG_CASE(5, 2) KG_CASE(6, 2)
    KG_CASE(1, 3) KG_CASE(2, 3) KG_CASE(3, 3) KG_CASE(4, 3) KG_CASE(5, 3) KG_CASE(6, 3)
#undef KG_CASE
    default:
      return hipErrorInvalidValue;
  }
}

// Block table of the recheck pass: shard s scans its rchk_n[s] listed rows in
// blocks of bk k-mers; the total goes to *total (one thread: shard tables are
// small).
__global__ void rchk_tables(ShardDesc* __restrict__ sd3, int nsh, int64_t bk,
                            int64_t* __restrict__ total) {
  if (threadIdx.x != 0) return;
  int64_t b = 0;
  for (int s = 0; s < nsh; ++s) {
    const int64_t m = (int64_t)*sd3[s].rchk_n;
    sd3[s].m = m;
    sd3[s].blk0 = b;
    b += (m + bk - 1) / bk;
  }
  *total = b;
}

hipError_t launch_rchk_tables(hipStream_t s, ShardDesc* sd3, int nsh, int64_t bk, int64_t* total) {
  hipLaunchKernelGGL(rchk_tables, dim3(1), dim3(64), 0, s, sd3, nsh, bk, total);
  return hipGetLastError();
}

hipError_t launch_scan_i8_res(int NT, int nch_max, int grid, hipStream_t s, const ShardDesc* sd,
                              int nshards, int64_t total_blocks, const ScanOut& o,
                              const int64_t* total_dev) {
  if (!i8_res_fits(NT, nch_max)) return hipErrorInvalidValue;
  dim3 g((unsigned)(total_blocks < grid ? (total_blocks > 0 ? total_blocks : 1) : grid));
  const size_t shm = i8_res_lds_bytes(NT, nch_max);
  switch (NT) {
    case 1: return launch_res_nt<1>(g, shm, s, sd, nshards, total_blocks, o, total_dev);
    case 2: return launch_res_nt<2>(g, shm, s, sd, nshards, total_blocks, o, total_dev);
    case 3: return launch_res_nt<3>(g, shm, s, sd, nshards, total_blocks, o, total_dev);
    case 4: return launch_res_nt<4>(g, shm, s, sd, nshards, total_blocks, o, total_dev);
    default: return hipErrorInvalidValue;
  }
}

// Tiles per wave actually used for (NT, t) and a requested count (0 = default).
// Besides the default, single-trait tiles of at most 96 columns at 3 limbs
// can run 2 tiles per wave.
int i8_tiles_for(int NT, int L, int t, int R_req) {
  const int def = i8_tiles(NT, t > 1 ? kMaxTraits : 1);
  if (R_req == 2 && def == 4 && L == 3) return 2;
  return def;
}

// k-mers per workgroup block
int64_t i8_block_kmers(int NT, int L, int t, int R_req) {
  return (int64_t)kI8Waves * 16 * i8_tiles_for(NT, L, t, R_req);
}

// Column capacity of NT tiles at L limbs: 4 lane groups x floor(4 NT / L).
int i8_column_capacity(int NT, int L) { return 4 * ((4 * NT) / L); }

// NT ranges per limb count: 1-3 (L = 1), 1-6 (L = 2), 1-16 (L = 3).
hipError_t launch_scan_i8(int NT, int L, int t, int R_req, int grid, hipStream_t s,
                          const ShardDesc* sd, int nshards, int64_t total_blocks,
                          const ScanOut& o, const int64_t* total_dev) {
  dim3 g((unsigned)(total_blocks < grid ? (total_blocks > 0 ? total_blocks : 1) : grid));
  const int TT = t > 1 ? kMaxTraits : 1;
  const int R = i8_tiles_for(NT, L, t, R_req);
  switch (((NT * 4 + L) * 16 + TT) * 8 + R) {
#define KG_CASE(X, Y)                                                                       \
  case ((X * 4 + Y) * 16 + 1) * 8 + i8_tiles(X, 1):                                         \
    return launch_i8_tr<X, Y, 1, i8_tiles(X, 1)>(g, s, sd, nshards, total_blocks, o, total_dev);       \
  case ((X * 4 + Y) * 16 + kMaxTraits) * 8 + i8_tiles(X, kMaxTraits):                       \
    return launch_i8_tr<X, Y, kMaxTraits, i8_tiles(X, kMaxTraits)>(g, s, sd, nshards,       \
                                                                    total_blocks, o, total_dev);
#define KG_CASE_R2(X)                                                                       \
  case ((X * 4 + 3) * 16 + 1) * 8 + 2:                                                      \
    return launch_i8_tr<X, 3, 1, 2>(g, s, sd, nshards, total_blocks, o, total_dev);
    KG_CASE(1, 1) KG_CASE(2, 1) KG_CASE(3, 1)
    KG_CASE(1, 2) KG_CASE(2, 2) KG_CASE(3, 2) KG_CASE(4, 2) KG_CASE(5, 2) KG_CASE(6, 2)
    KG_CASE(1, 3) KG_CASE(2, 3) KG_CASE(3, 3) KG_CASE(4, 3) KG_CASE(5, 3) KG_CASE(6, 3)
    KG_CASE(7, 3) KG_CASE(8, 3) KG_CASE(9, 3) KG_CASE(10, 3) KG_CASE(11, 3) KG_CASE(12, 3)
    KG_CASE(13, 3) KG_CASE(14, 3) KG_CASE(15, 3) KG_CASE(16, 3)
    KG_CASE_R2(1) KG_CASE_R2(2) KG_CASE_R2(3) KG_CASE_R2(4) KG_CASE_R2(5) KG_CASE_R2(6)
#undef KG_CASE
#undef KG_CASE_R2
    default:
      return hipErrorInvalidValue;
  }
}

// Pre-filter screen over a shard table; nch_max = the largest model's 512-sample
// chunks (sizes the LDS weight table).  grid = persistent workgroups.
// w8: the 8-wave shape (at most 128 registers per wave, two workgroups per
// CU) in place of the default (16 waves for the two-accumulator form).
__host__ __device__ constexpr int scr_waves_w8(int Q, bool w8) { return w8 ? 8 : scr_waves(Q); }
int64_t screen_block_kmers(int t, bool w8) {
  const int Q = scr_rows(t);
  return (int64_t)scr_waves_w8(Q, w8) * 16 * scr_tiles(Q);
}
int screen_max_chunks() { return kScrMaxChunks; }
int screen_rows(int t) { return scr_rows(t); }
bool screen_count_row(int t) { return scr_count_row(t); }

template <int Q, bool CR, bool DD, bool W8>
static const void* screen_fn() {
  return (const void*)screen_i8<Q, CR, DD, scr_waves_w8(Q, W8), W8 ? 2 : 1>;
}

template <bool DD, bool W8>
static const void* screen_kernel_t(int t) {
  if (!scr_count_row(t)) return screen_fn<16, false, DD, W8>();
  switch (scr_rows(t)) {
    case 4: return screen_fn<4, true, DD, W8>();
    case 8: return screen_fn<8, true, DD, W8>();
    default: return screen_fn<16, true, DD, W8>();
  }
}

// The screen instantiation for (t, dd, w8).
static const void* screen_kernel(int t, bool dd, bool w8) {
  if (dd) return w8 ? screen_kernel_t<true, true>(t) : screen_kernel_t<true, false>(t);
  return w8 ? screen_kernel_t<false, true>(t) : screen_kernel_t<false, false>(t);
}

// dd: the duplicate-row check (every ScreenDesc then has dmap and mult).
hipError_t launch_screen(int t, int nch_max, int grid, hipStream_t s, const ScreenDesc* sd,
                         int nshards, int64_t total_blocks, bool dd, bool w8) {
  if (nch_max > kScrMaxChunks || t < 1 || t > kMaxTraits) return hipErrorInvalidValue;
  const void* fn = screen_kernel(t, dd, w8);
  dim3 g((unsigned)(total_blocks < grid ? (total_blocks > 0 ? total_blocks : 1) : grid));
  const size_t shm = scr_lds_bytes(nch_max);
  void* args[] = {(void*)&sd, (void*)&nshards, (void*)&total_blocks};
  const dim3 blk(64 * scr_waves_w8(scr_rows(t), w8));
  return hipLaunchKernel(fn, g, blk, args, shm, s);
}

// The ring screen (one trait): W = 8 or 12 waves per workgroup, one workgroup
// per CU; 12 waves take 4-chunk models only (LDS).
bool screen_ring_fits(int nch, int W) {
  return nch >= 4 && nch % 4 == 0 && nch <= 4 * kRingMaxSeg && (W == 8 || (W == 12 && nch == 4));
}

// Instantiation for (W, nch_max, dd): on 4-chunk models, 8 waves keep the
// weights in registers with four ring slots and 12 waves with three (with the
// duplicate check, whose registers that leaves no room for, the weights in LDS
// with three or two slots); 8-chunk models read the weights from LDS with two
// slots.  No instantiation may spill: scratch stores would share the vmcnt
// counter with the ring's loads and void its waits.
struct RingCfg {
  int nbuf;
  bool wreg;
};
static RingCfg ring_cfg(int W, int nch_max, bool dd) {
  if (W == 8 && nch_max <= 4) return dd ? RingCfg{3, false} : RingCfg{4, true};
  if (W == 12 && !dd) return RingCfg{3, true};
  return RingCfg{2, false};
}
size_t screen_ring_lds(int W, int nch_max, bool dd) {
  const RingCfg c = ring_cfg(W, nch_max, dd);
  return ring_lds_bytes(W, c.nbuf, nch_max, c.wreg);
}

template <int W, int NBUF, bool DD, bool WREG>
static const void* ring_fn() {
  return (const void*)screen_ring<W, NBUF, DD, WREG>;
}

static const void* ring_kernel(int W, int nch_max, bool dd) {
  const RingCfg c = ring_cfg(W, nch_max, dd);
  if (W == 8) {
    if (c.wreg) return ring_fn<8, 4, false, true>();
    if (c.nbuf == 3) return ring_fn<8, 3, true, false>();
    return dd ? ring_fn<8, 2, true, false>() : ring_fn<8, 2, false, false>();
  }
  return dd ? ring_fn<12, 2, true, false>() : ring_fn<12, 3, false, true>();
}

hipError_t launch_screen_ring(int W, int nch_max, int grid, hipStream_t s, const ScreenDesc* sd,
                              int nshards, int64_t total_blocks, bool dd) {
  if (!screen_ring_fits(nch_max, W)) return hipErrorInvalidValue;
  const void* fn = ring_kernel(W, nch_max, dd);
  dim3 g((unsigned)(total_blocks < grid ? (total_blocks > 0 ? total_blocks : 1) : grid));
  void* args[] = {(void*)&sd, (void*)&nshards, (void*)&total_blocks, (void*)&nch_max};
  return hipLaunchKernel(fn, g, dim3(64 * W), args, screen_ring_lds(W, nch_max, dd), s);
}

// Workgroups of the screen one CU holds at once for a model of nch chunks.
int screen_wgs_per_cu(int t, int nch, bool dd, bool w8) {
  int nb = 0;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, screen_kernel(t, dd, w8), 64 * scr_waves_w8(scr_rows(t), w8), scr_lds_bytes(nch));
  return (e != hipSuccess || nb < 1) ? 1 : nb;
}

int screen_chunk_words() { return kScrChunkWords; }

// Bitmap compaction and the full kernel's block table (see scr_scan).
hipError_t launch_screen_compact(hipStream_t s, const ScrChunk* ch, int ncc, const ScreenDesc* ss,
                                 uint32_t* ccount, int64_t* P, ShardDesc* sd2, int nsh,
                                 int64_t bk, int64_t* total) {
  if (ncc > 0) hipLaunchKernelGGL(scr_count, dim3((unsigned)ncc), dim3(256), 0, s, ch, ss, ccount);
  hipLaunchKernelGGL(scr_scan, dim3(1), dim3(1024), 0, s, ccount, ncc, P, ss, sd2, nsh, bk, total);
  if (ncc > 0) hipLaunchKernelGGL(scr_write, dim3((unsigned)ncc), dim3(256), 0, s, ch, ss, P);
  return hipGetLastError();
}

template <int NT>
static hipError_t launch_f32_nt(dim3 grid, size_t shm, hipStream_t s, const uint64_t* bits,
                                int64_t m, int64_t rw, int64_t row0, int nch, const float* wts,
                                const uint64_t* mask, const EpiF32& ep, const ScanOut& o,
                                int64_t nblocks) {
  auto kfn = scan_f32<NT, 4>;
  hipLaunchKernelGGL(kfn, grid, dim3(kThreads), shm, s, bits, m, rw, row0, nch, wts, mask, ep, o,
                     nblocks);
  return hipGetLastError();
}

// nch = the model's 256-sample chunks (the weights hold exactly that many).
hipError_t launch_scan_f32(int NT, int nch, int grid, hipStream_t s, const uint64_t* bits,
                           int64_t m, int64_t rw, int64_t row0, const float* wts,
                           const uint64_t* mask, const EpiF32& ep, const ScanOut& o) {
  const int64_t per_block = (int64_t)kWaves * 16 * 4;
  const int64_t nblocks = (m + per_block - 1) / per_block;
  const size_t shm = (size_t)64 * NT * 64 * 4;
  dim3 g((unsigned)(nblocks < grid ? nblocks : grid));
  switch (NT) {
#define KG_CASE(X) \
  case X:          \
    return launch_f32_nt<X>(g, shm, s, bits, m, rw, row0, nch, wts, mask, ep, o, nblocks);
    KG_CASE(1) KG_CASE(2) KG_CASE(3) KG_CASE(4) KG_CASE(5) KG_CASE(6)
#undef KG_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_refine(hipStream_t s, const Cand* cand, const unsigned long long* count,
                         int64_t cap, const RefDesc* rd, double* res, int blocks,
                         const RefHost& hb) {
  hipLaunchKernelGGL(refine_f64, dim3((unsigned)blocks), dim3(64 * kRefWaves), 0, s, cand, count,
                     cap, rd, res, hb);
  return hipGetLastError();
}

// Descriptor tables passed by value: one small launch writes them to device
// memory and zeroes the scan's counters (in place of a host-to-device copy and
// a separate zeroing launch).
constexpr int kSetupWords = 240;
struct SetupArgs {
  uint64_t* dst;
  unsigned long long* zero;
  int nwords, nzero;
  uint64_t w[kSetupWords];
};

__global__ void setup_tables(SetupArgs a) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  for (int i = i0; i < a.nwords; i += st) a.dst[i] = a.w[i];
  for (int i = i0; i < a.nzero; i += st) a.zero[i] = 0ull;
}

size_t setup_max_bytes() { return sizeof(uint64_t) * kSetupWords; }

hipError_t launch_setup(hipStream_t s, const void* src, size_t bytes, void* dst,
                        unsigned long long* zero, int nzero) {
  if (bytes > setup_max_bytes() || bytes % 8 != 0 || nzero < 0) return hipErrorInvalidValue;
  SetupArgs a{};
  a.dst = static_cast<uint64_t*>(dst);
  a.zero = zero;
  a.nwords = (int)(bytes / 8);
  a.nzero = nzero;
  if (bytes) memcpy(a.w, src, bytes);
  const int work = a.nwords > nzero ? a.nwords : nzero;
  const int blocks = work <= 256 ? 1 : (work + 255) / 256 < 64 ? (work + 255) / 256 : 64;
  hipLaunchKernelGGL(setup_tables, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

// Dynamic-LDS limits of every bulk-kernel instantiation, set once per process
// (outside any stream capture).
template <int NT>
static hipError_t set_attr_nt() {
  hipError_t e = hipSuccess;
  if (NT <= 6)
    e = hipFuncSetAttribute((const void*)scan_f32<(NT <= 6 ? NT : 6), 4>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 64 * NT * 64 * 4);
  return e;
}

hipError_t set_kernel_attributes() {
  hipError_t e = hipSuccess;
  const void* res_fns[] = {(const void*)scan_i8<1, 2, 1, 4, true>,
                           (const void*)scan_i8<2, 2, 1, 4, true>,
                           (const void*)scan_i8<3, 2, 1, 4, true>,
                           (const void*)scan_i8<4, 2, 1, 4, true>};
  for (const void* f : res_fns)
    if (e == hipSuccess)
      e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kResLdsBytes);
#define KG_F6(X, Y)                                                              \
  (const void*)scan_i8<X, Y, 1, 4, true, true, kCntMasked>,                     \
      (const void*)scan_i8<X, Y, 1, 4, true, true, kCntBits>,                   \
      (const void*)scan_i8<X, Y, 1, 4, true, true, kCntColumn>
  const void* f6_fns[] = {KG_F6(1, 2), KG_F6(2, 2), KG_F6(3, 2), KG_F6(4, 2),
                          KG_F6(5, 2), KG_F6(6, 2), KG_F6(1, 3), KG_F6(2, 3),
                          KG_F6(3, 3), KG_F6(4, 3), KG_F6(5, 3), KG_F6(6, 3)};
#undef KG_F6
  for (const void* f : f6_fns)
    if (e == hipSuccess)
      e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kResLdsBytes);
  const int scr = (int)scr_lds_bytes(kScrMaxChunks);
  for (int W : {8, 12})
    for (int nch : {4, 8})
      for (int dd = 0; dd < 2; ++dd)
        if (e == hipSuccess && screen_ring_fits(nch, W))
          e = hipFuncSetAttribute(ring_kernel(W, nch, dd != 0),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)screen_ring_lds(W, nch, dd != 0));
  for (int t = 1; t <= kMaxTraits; ++t)
    for (int v = 0; v < 4; ++v)
      if (e == hipSuccess)
        e = hipFuncSetAttribute(screen_kernel(t, (v & 1) != 0, (v & 2) != 0),
                                hipFuncAttributeMaxDynamicSharedMemorySize, scr);
#define KG_ATTR(X) \
  if (e == hipSuccess) e = set_attr_nt<X>();
  KG_ATTR(1) KG_ATTR(2) KG_ATTR(3) KG_ATTR(4) KG_ATTR(5) KG_ATTR(6) KG_ATTR(7) KG_ATTR(8)
#undef KG_ATTR
  return e;
}

hipError_t launch_pair_masks(hipStream_t s, const uint64_t* bits, int64_t rw,
                            const uint64_t* mask, int64_t W, int64_t i0, int64_t cnt,
                            uint64_t* out) {
  if (cnt <= 0) return hipSuccess;
  const int64_t total = cnt * W;
  const int64_t blocks = std::min<int64_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(pair_masks, dim3((unsigned)blocks), dim3(256), 0, s, bits, rw, mask, W, i0,
                     cnt, out);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// Scratch of the allocation-free scan (kg_scan_device_noalloc): counters,
// descriptor tables, candidate records and their FP64 results, in a
// file-scope device array that each GPU holds for the life of the process.
// The host never passes `scratch` itself to a kernel: its host-side name is a
// stand-in, not a device address.  noalloc_scratch() looks the address up with
// hipGetSymbolAddress once per device (with that device current) and every
// host use goes through that address.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(int device, void** ptr) {
  static std::mutex mu;
  static void* addr[kNoAllocMaxDevices] = {};
  if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

hipError_t launch_zero(hipStream_t s, unsigned long long* p, int n) {
  hipLaunchKernelGGL(zero_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_synth(hipStream_t s, uint64_t* out, uint64_t seed, int64_t kmer0, int64_t m,
                        uint32_t n_acc, int64_t rw, const uint8_t* pop, const uint32_t* thr,
                        uint32_t npop, uint32_t nprof) {
  const int64_t total = m * rw;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 1 << 20) blocks = 1 << 20;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(synth_bits, dim3((unsigned)blocks), dim3(256), 0, s, out, seed, kmer0, m,
                     n_acc, rw, pop, thr, npop, nprof);
  return hipGetLastError();
}

}  // namespace kg
