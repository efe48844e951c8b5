// Host-side fill of seeded synthetic k-mer bits (same arithmetic as the
// device fill kernel; both include synth_core.h).
#include <stddef.h>
#include "synth_core.h"

// Writes rows kmer0 .. kmer0+m-1 into out[m][row_words]; words past
// ceil(n_acc/64) are zero.
void ks_fill_bits_host(uint64_t seed, int64_t kmer0, int64_t m, uint32_t n_acc,
                       int64_t row_words, const uint8_t* pop, const uint32_t* thr,
                       uint32_t npop, uint32_t n_profiles, uint64_t* out) {
  const uint32_t used = (n_acc + 63u) / 64u;
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    uint64_t* row = out + (size_t)i * (size_t)row_words;
    for (int64_t w = 0; w < row_words; ++w)
      row[w] = (w < (int64_t)used)
                   ? ks_word(seed, (uint64_t)(kmer0 + i), (uint32_t)w, n_acc, pop, thr,
                             npop, n_profiles)
                   : 0ULL;
  }
}
