// Seeded synthetic k-mer presence bits, shared verbatim by the host (gcc) and
// the device (hipcc) so that both sides produce bit-identical matrices.
//
// Model: accession j belongs to population pop[j]; k-mer i draws a frequency
// profile prof(i) from a table of P profiles, each holding one 16-bit
// threshold per population (thr = floor(freq * 65536)).  Bit (i, j) is set
// iff a 16-bit uniform drawn from a counter-based hash of (seed, i, j) is
// below thr[prof(i) * npop + pop[j]].  Only integer arithmetic is used, so the
// result does not depend on the compiler or the device.
//
// Packed layout (the hot path's input layout): k-mer i occupies row_words
// consecutive uint64 words; accession j is bit (j % 64) of word (j / 64).
// Bits at positions >= n_acc are zero.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define KS_HD __host__ __device__ __forceinline__
#else
#define KS_HD static inline
#endif

// splitmix64 finaliser (Steele, Lea, Flood 2014).
KS_HD uint64_t ks_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Per-k-mer hash: everything about k-mer i derives from it.
KS_HD uint64_t ks_kmer_hash(uint64_t seed, uint64_t kmer) {
  return ks_mix64(ks_mix64(seed ^ 0x6a09e667f3bcc908ULL) + kmer * 0x9e3779b97f4a7c15ULL);
}

KS_HD uint32_t ks_profile(uint64_t kh, uint32_t n_profiles) {
  return (uint32_t)((kh >> 32) % n_profiles);
}

// 64 random bits for accessions 4q .. 4q+3 of the k-mer with hash kh.
KS_HD uint64_t ks_quad(uint64_t kh, uint32_t q) {
  return ks_mix64(kh + (uint64_t)(q + 1) * 0xbb67ae8584caa73bULL);
}

// Word w (accessions 64w .. 64w+63) of k-mer `kmer`.
KS_HD uint64_t ks_word(uint64_t seed, uint64_t kmer, uint32_t w, uint32_t n_acc,
                       const uint8_t* pop, const uint32_t* thr, uint32_t npop,
                       uint32_t n_profiles) {
  const uint64_t kh = ks_kmer_hash(seed, kmer);
  const uint32_t* t = thr + (uint64_t)ks_profile(kh, n_profiles) * npop;
  uint64_t word = 0;
  for (uint32_t q = 0; q < 16; ++q) {
    const uint32_t j0 = 64u * w + 4u * q;
    if (j0 >= n_acc) break;
    const uint64_t h = ks_quad(kh, 16u * w + q);
    for (uint32_t e = 0; e < 4; ++e) {
      const uint32_t j = j0 + e;
      if (j >= n_acc) break;
      const uint32_t u = (uint32_t)((h >> (16 * e)) & 0xffffu);
      if (u < t[pop[j]]) word |= 1ULL << (4 * q + e);
    }
  }
  return word;
}
