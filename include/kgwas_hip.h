/* kgwas_hip.h -- C ABI of the MI355X (gfx950) 1D k-mer association scan.
 *
 * The reference has no FFI for this path; the hot path is the in-process seam
 * of `kgwas --mode cols` (src/kgwas.cpp:183-289): Pearson pre-filter
 * (src/kgwas.cpp:183-218), compute_cols_stats (src/kgwas.cpp:138-166),
 * compute_hit + threshold (src/kgwas.cpp:44-56, 239-259).  These entry points
 * replace that seam; the drop-in `kgwas` binary (supergwas_amd/host) keeps the
 * process seam (argv/env/files) and calls them.
 *
 * Conventions: plain C types; every function returns 0 (KG_OK) or a KG_ERR_*
 * code and never throws; kg_last_error() gives the message.  The caller owns
 * host memory, the library owns device memory it allocates.  One context per
 * host thread.  Device pointers passed in (kg_scan_device, kg_stats_device,
 * kg_synth_bits) must be 16-byte aligned device allocations.
 */
#ifndef KGWAS_HIP_H
#define KGWAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KG_OK 0
#define KG_ERR_ARG 1       /* invalid argument / shape */
#define KG_ERR_DEVICE 2    /* HIP runtime error (no device, launch or copy failure) */
#define KG_ERR_STATE 3     /* call order (e.g. scan before kg_set_model) */
#define KG_ERR_CAPACITY 4  /* more hits than `cap`; *n_out holds the required size */
#define KG_ERR_SCRATCH 5   /* kg_scan_device_noalloc: more candidates than its fixed scratch holds */

#define KG_PREC_MIXED 0    /* bits x int8-limb weights, exact int32 MFMA accumulation,
                              FP64 epilogue, FP64 refinement of every reported k-mer */
#define KG_PREC_FP32 1     /* plain FP32: the reference's cols-mode arithmetic
                              (src/kgwas.cpp:146-158) on FP32 MFMA */

#define KG_HIT_2D 1u       /* p <= 10^-log10_thresh_2d (src/kgwas.cpp:251-254) */

/* Pre-filter screen of KG_PREC_MIXED scans (src/kgwas.cpp:183-218 filters
 * before the projections; the screen does the same on the device): a cheap
 * bulk pass drops the k-mers whose pre-filter fails for certain, and the full
 * kernel runs on the rest.  Counts and hits are identical either way.
 * AUTO screens whenever corr_threshold > 0 and keeps screening while the
 * screen plus the survivor scan take less time than the full kernel over every
 * k-mer would (measured per scan; after a loss it skips 8 scans; scans of up
 * to 2^22 k-mers run unscreened at a new size and once per 32 scans to measure
 * the full kernel); OFF never screens; ON always screens where the model
 * allows (n_acc <= 8192). */
#define KG_SCREEN_AUTO 0
#define KG_SCREEN_OFF 1
#define KG_SCREEN_ON 2

typedef struct kg_ctx kg_ctx;

typedef struct kg_opts {
  int32_t device;      /* HIP device ordinal */
  int32_t precision;   /* KG_PREC_MIXED or KG_PREC_FP32 */
  int32_t limbs;       /* int8 limbs per weight column for KG_PREC_MIXED (2 or 3; 0 = 3) */
  int32_t screen;      /* KG_SCREEN_AUTO (0), KG_SCREEN_OFF or KG_SCREEN_ON */
  int32_t dedup;       /* 1: duplicate-row scans (kg_set_dedup); 0 (default): off */
  int32_t reserved[3];
} kg_opts;

/* Thresholds of one scan (src/kgwas.cpp:28-30, 101-112, 230-232). */
typedef struct kg_scan_params {
  double corr_threshold;   /* CORR_THRESHOLD: keep k-mer iff |cor(x, y)| >= this */
  double log10_thresh;     /* report iff p <= 10^-log10_thresh */
  double log10_thresh_2d;  /* KG_HIT_2D iff p <= 10^-log10_thresh_2d; <= 0 disables */
} kg_scan_params;

/* One reported k-mer: the values compute_hit prints (src/kgwas.cpp:37-56). */
typedef struct kg_hit {
  int64_t idx;      /* k-mer row within the scanned batch */
  int32_t trait;    /* phenotype column */
  int32_t c;        /* allele count over phenotyped samples (x'x) */
  uint32_t flags;   /* KG_HIT_2D */
  int32_t shard;    /* shard index (kg_scan_shards), else 0 */
  double xxM;       /* x'x - |U'x|^2 */
  double xyM;       /* x'y - (U'x)'(U'y) */
  double effect;    /* xyM / (xxM * yyM) */
  double T;         /* n ln(1 / (1 - r^2)) */
  double pval;      /* erfc(sqrt(T / 2)) */
} kg_hit;

/* Timings of the last scan on the context's stream (HIP events), ms. */
typedef struct kg_timing {
  double bulk_ms;       /* bulk scan kernel (when screened: the survivor-table fix-up and the
                           full kernel over the survivors) */
  double refine_ms;     /* FP64 refinement kernel (0 for KG_PREC_FP32) */
  double total_ms;      /* first to last device operation of the scan */
  int64_t candidates;   /* k-mers sent to refinement / host decision */
  int64_t kmers;        /* k-mers scanned */
  double copy_ms;       /* kg_scan: host-to-device copies (copy stream, summed) */
  double wall_ms;       /* kg_scan: host wall time of the whole call */
  double screen_ms;     /* pre-filter screen kernel (0 if unscreened) */
  int64_t survivors;    /* k-mers the screen passed to the full kernel (= kmers if unscreened) */
  int64_t stat_refined; /* kg_stats_device (KG_PREC_MIXED): k-mer/trait values recomputed in
                           FP64 because their bulk error estimate exceeded the budget */
  int64_t rechecked;    /* screened scans with two FP6 digits (kg_tune.resident_form 2):
                           survivors whose decision the digit bounds left open, scanned again
                           by the three-limb kernel */
} kg_timing;

/* Device selection, stream creation.  Replaces the OpenBLAS/OpenMP thread
 * set-up of src/kgwas.cpp:143-144, 227-228. */
int kg_create(kg_ctx** out, const kg_opts* opts);
void kg_destroy(kg_ctx* ctx);
/* Message of the last failing call on ctx (never NULL; "" when none). */
const char* kg_last_error(const kg_ctx* ctx);

/* Null model.  n_acc accessions (bits per k-mer); idxs[n] = accession
 * positions that have a phenotype, increasing (build_overlap,
 * src/kgwas_utils.hpp:291-314); U[n*p] column-major FP64 in overlap order
 * (load_Ubin, src/kgwas_utils.hpp:135-157); Y[n*t] column-major FP64
 * phenotypes.  The library copies everything; FP64 residualisation and the
 * weight quantisation happen here. */
int kg_set_model(kg_ctx* ctx, int32_t n_acc, const int32_t* idxs, int32_t n, const double* U,
                 int32_t p, const double* Y, int32_t t);

/* Scan m packed k-mer rows (row_words uint64 words each, row_words a multiple
 * of 8 and >= ceil(n_acc/64); accession a is bit a%64 of word a/64).
 * Reports every k-mer (and trait) that passes the pre-filter and has
 * p <= 10^-log10_thresh, sorted by (idx, trait), into out[0..*n_out).
 * n_passed[t] receives the pre-filter pass count per trait (the `.count`
 * value, src/kgwas.cpp:292-304).  If more than `cap` hits exist, returns
 * KG_ERR_CAPACITY with *n_out = required count.
 * kg_scan: host bits, streamed to the device in 64 MiB chunks on a copy
 * stream, each chunk's copy overlapping the previous chunk's scan;
 * kg_scan_device: bits already on the context's device. */
int kg_scan(kg_ctx* ctx, const uint64_t* bits, int64_t m, int64_t row_words,
            const kg_scan_params* prm, kg_hit* out, int64_t cap, int64_t* n_out,
            int64_t* n_passed);
int kg_scan_device(kg_ctx* ctx, const uint64_t* d_bits, int64_t m, int64_t row_words,
                   const kg_scan_params* prm, kg_hit* out, int64_t cap, int64_t* n_out,
                   int64_t* n_passed);

/* A further null model (its own sample mask, U and phenotypes) in the same
 * context; *model_id receives its index (kg_set_model's model is 0).  Used by
 * batched scans whose shards have different phenotyped samples, as separate
 * kgwas runs over batch_* directories would. */
int kg_add_model(kg_ctx* ctx, int32_t n_acc, const int32_t* idxs, int32_t n, const double* U,
                 int32_t p, const double* Y, int32_t t, int32_t* model_id);

/* One shard of a batched scan: device rows and the model they are tested with. */
typedef struct kg_shard {
  const uint64_t* d_bits;
  int64_t m;
  int64_t row_words;
  int32_t model;
  int32_t pad;
} kg_shard;

#define KG_SHARDS_PER_LAUNCH 1u  /* one bulk launch per shard (comparison mode) */
#define KG_SHARDS_GRAPH 2u       /* capture the device work once as a HIP graph, replay it */

/* Scans many shards of uneven size in one persistent launch (the reference
 * runs one process per batch_* directory: scripts/run_pipeline.sh:243-253).
 * All shards' models must share p, t and the limb count.  Hits carry their
 * shard index and are sorted by (shard, idx, trait); n_passed[s * t + tr] is
 * the per-shard `.count` value. */
int kg_scan_shards(kg_ctx* ctx, const kg_shard* shards, int32_t nshards,
                   const kg_scan_params* prm, uint32_t flags, kg_hit* out, int64_t cap,
                   int64_t* n_out, int64_t* n_passed);

/* All-pairs epistasis scan (replaces emit_pairs_self, src/kgwas_2d.cpp:258-401):
 * for rows i < j of the N device rows at d_bits, with i in [i0, i1) (the
 * reference's --task-id/--num-tasks row range, src/kgwas_2d.cpp:283-285), the
 * k-mer x_i AND x_j (src/kgwas_2d.cpp:372) is tested against model 0 with the
 * statistic and threshold of kg_scan (p <= 10^-log10_thresh), without the
 * pre-filter; corr_threshold and log10_thresh_2d are ignored.  Each reported
 * pair is a kg_hit with shard = i and idx = j, sorted by (i, j, trait)
 * (src/kgwas_2d.cpp:415-418); c is the pair's allele count.  Needs
 * KG_PREC_MIXED. */
int kg_scan_pairs(kg_ctx* ctx, const uint64_t* d_bits, int64_t N, int64_t row_words,
                  int64_t i0, int64_t i1, const kg_scan_params* prm, kg_hit* out, int64_t cap,
                  int64_t* n_out);

/* Full-statistic mode (validation and checksums): the statistic's inputs for
 * every k-mer, written to host arrays: c[m], xxM[m], xyM[m*t], S[m*t]
 * (S = x'(y - mean y), the pre-filter numerator), k-mer-major.  KG_PREC_FP32:
 * the FP32 kernel's values.  KG_PREC_MIXED: the bulk kernel's values, except
 * for the k-mers whose estimated error of T exceeds 4e-6 relative (k-mers
 * almost in the span of U, where x'x - |U'x|^2 cancels): those are recomputed
 * in FP64 on the device (kg_timing.stat_refined counts them). */
int kg_stats_device(kg_ctx* ctx, const uint64_t* d_bits, int64_t m, int64_t row_words,
                    int32_t* c, double* xxM, double* xyM, double* S);

int kg_last_timing(const kg_ctx* ctx, kg_timing* out);

/* Allocation-free variant of kg_scan_device (no reference counterpart; for
 * callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and errors
 * as kg_scan_device for KG_PREC_MIXED, scanned without the pre-filter screen;
 * it allocates no device or host memory: counters, descriptors and up to
 * 65,536 candidate k-mers (those the bulk kernel cannot decide alone or
 * reports) with their FP64 values live in a fixed 5.3 MB array that each GPU
 * holds for the life of the process, and in fixed host staging.  A scan with
 * more candidates returns KG_ERR_SCRATCH (kg_scan_device handles any number).
 * Calls share that array and staging, so the library serialises them (one
 * lock per process). */
int kg_scan_device_noalloc(kg_ctx* ctx, const uint64_t* d_bits, int64_t m, int64_t row_words,
                           const kg_scan_params* prm, kg_hit* out, int64_t cap, int64_t* n_out,
                           int64_t* n_passed);

/* Changes the context's pre-filter screen mode (KG_SCREEN_*) for later scans
 * (no reference counterpart: the reference always filters first). */
int kg_set_screen(kg_ctx* ctx, int32_t mode);

/* Duplicate-row scans (the batches of src/convert_matrix.cpp:366-368 are sorted
 * by bit pattern, so k-mers with equal rows are adjacent): when on, a screened
 * scan compares every row with the previous one under the accession mask
 * inside the pre-filter screen, runs the full kernel on the first row of each
 * run of equal rows only, counts that row once per member of its run and
 * reports every member of a reported run (same values, consecutive idx).
 * Results are identical either way; it needs the screen (corr_threshold > 0,
 * screen mode not KG_SCREEN_OFF) and keeps it on in KG_SCREEN_AUTO.  Runs are
 * detected within each 64 (or 32) k-mer wave tile of the screen (and within a
 * kg_scan chunk).  kg_scan_pairs does not use it.  No reference counterpart
 * (the reference computes every k-mer). */
int kg_set_dedup(kg_ctx* ctx, int32_t on);

/* Launch shape of the int8 scan (no reference counterpart; the reference has
 * one code path).  Every shape gives identical results. */
typedef struct kg_tune {
  int32_t tiles_per_wave;      /* full kernel: k-mer tiles of 16 per wave, 4 or 2 (0: the
                                  model's default, 4 for one trait and at most 6 weight
                                  tiles, else 2; 2 is available at 3 limbs) */
  int32_t workgroups_per_cu;   /* full kernel: persistent 8-wave workgroups per CU (1-64) */
  int32_t screen_wgs_per_cu;   /* pre-filter screen shape: 1 (one 16-wave workgroup per CU for
                                  one trait, one 8-wave workgroup otherwise, rows loaded into
                                  registers), 2 (two 8-wave workgroups of at most 128 registers
                                  per wave), 3 (one trait, models of 4 or 8 512-sample chunks:
                                  one 8-wave workgroup per CU streaming its rows through an LDS
                                  ring with contiguous nontemporal loads), 4 (the same with 12
                                  waves, 4-chunk models); 0 in kg_set_tune: the default (3
                                  where it applies, else 1).  kg_get_tune reports the shape
                                  model 0's scans use */
  int32_t kmers_per_workgroup; /* out: k-mers per full-kernel workgroup block */
  int32_t screen_kmers_per_workgroup; /* out: k-mers per screen workgroup block */
  int32_t stream_weights;      /* 0: single-trait screened scans run their survivors through
                                  the resident-weight kernel (two int8 limbs, every weight
                                  chunk in LDS, the screen's proven pre-filter passes) where
                                  the model's weights fit; 1: always the streaming kernel */
  int32_t resident_active;     /* out: 1 when screened scans of model 0 use the
                                  resident-weight kernel */
  int32_t resident_form;       /* operands of the resident-weight kernel: 1 = two int8 limbs
                                  (shift-and-mask bit unpack), 2 = two FP6 digits per column
                                  against FP4 bit operands (one AND per operand dword), with a
                                  three-limb recheck pass of the k-mers its bounds leave open,
                                  3 = three FP6 digits; 0 in kg_set_tune: the default (2) */
} kg_tune;

/* Autotuner: times the int8 bulk scan of `sample` random rows with model 0 for
 * each candidate (k-mer tiles per wave x persistent workgroups per CU for the
 * full kernel, then the screen's workgroups per CU), keeps the fastest for
 * later scans on ctx and returns it in *out.  The default shape is timed first
 * and is replaced only by one that is more than 2% faster.  A no-op for
 * KG_PREC_FP32.  kg_set_tune sets a shape explicitly; kg_get_tune reads it. */
int kg_autotune(kg_ctx* ctx, int64_t sample, kg_tune* out);
int kg_set_tune(kg_ctx* ctx, const kg_tune* shape);
int kg_get_tune(const kg_ctx* ctx, kg_tune* out);

/* Device memory on the context's device for callers without their own
 * allocator (tests, benchmarks, ctypes hosts).  kg_memcpy direction:
 * 0 = host to device, 1 = device to host, 2 = device to device. */
int kg_dev_alloc(kg_ctx* ctx, size_t bytes, void** d_ptr);
int kg_dev_free(kg_ctx* ctx, void* d_ptr);
int kg_memcpy(kg_ctx* ctx, void* dst, const void* src, size_t bytes, int32_t direction);

/* Page-locked host memory for callers that stream rows with kg_scan: rows in
 * such memory are copied in 64 MiB chunks on a copy stream while the previous
 * chunk is scanned (pageable rows go through the context's own page-locked
 * bounce buffers).  No reference counterpart (the reference reads its input
 * from files into host memory, src/kgwas_utils.hpp:225-282). */
int kg_host_alloc(kg_ctx* ctx, size_t bytes, void** h_ptr);
int kg_host_free(kg_ctx* ctx, void* h_ptr);

/* Seeded synthetic bits on the device (benchmarks and tests; no reference
 * counterpart): rows kmer0 .. kmer0+m-1 of the model in synth_core.h. */
int kg_synth_bits(kg_ctx* ctx, uint64_t* d_out, uint64_t seed, int64_t kmer0, int64_t m,
                  int32_t n_acc, int64_t row_words, const uint8_t* pop, const uint32_t* thr,
                  int32_t npop, int32_t n_profiles);

#ifdef __cplusplus
}
#endif
#endif /* KGWAS_HIP_H */
