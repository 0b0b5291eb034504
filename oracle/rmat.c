/*
 * Oracle side of the synthetic RMAT input (test infrastructure).
 *
 * The reference ships no generator; its input is a packed binary edge list
 * (core/type.hpp:27-41, one EdgeUnit<E> per edge, |E| = file size / record
 * size as in core/graph.hpp:781-782).  The product generates the same edge
 * list on the device; this file is the CPU statement of the same definition,
 * used to check the device generator bit for bit.
 *
 * Definition (integer-only so that CPU and GPU agree exactly):
 *   key        = mix64(seed ^ 0xD1B54A32D192ED03)
 *   word(i, j) = mix64(key + (16*i + j + 1) * 0x9E3779B97F4A7C15)   (splitmix64, counter mode)
 *   level l (0 = most significant bit) uses the 32-bit half (l & 1) of word(i, l >> 1):
 *     r <  T_a            -> (src bit, dst bit) = (0, 0)
 *     r <  T_ab           -> (0, 1)
 *     r <  T_abc          -> (1, 0)
 *     otherwise           -> (1, 1)
 *   with T_x = (uint64)(x * 2^32) computed in double on the host.
 *   weight     = ((word(i, 15) >> 40) + 1) * 2^-24, exactly representable in (0, 1].
 *   permutation (optional, applied to both endpoints): three odd-multiply/add
 *   steps modulo 2^scale separated by x ^= x >> ceil(scale/2); every step is a
 *   bijection on [0, 2^scale).
 * Duplicates and self-loops are kept, as the reference's loaders keep them.
 */
#include "oracle.h"
#include <string.h>

#define OG_GOLDEN 0x9E3779B97F4A7C15ULL

static inline uint64_t og_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static inline uint64_t og_word(uint64_t key, uint64_t i, uint64_t j) {
  return og_mix64(key + (16ULL * i + j + 1ULL) * OG_GOLDEN);
}

uint32_t og_rmat_permute(uint32_t x, int scale, uint64_t seed) {
  if (scale <= 0) return x;
  uint64_t mask = (scale >= 32) ? 0xFFFFFFFFULL : ((1ULL << scale) - 1ULL);
  uint64_t k1 = og_mix64(seed ^ 0x632BE59BD9B4E019ULL);
  uint64_t k2 = og_mix64(k1 + OG_GOLDEN);
  uint64_t k3 = og_mix64(k2 + OG_GOLDEN);
  uint64_t m1 = (k1 & mask) | 1ULL, a1 = (k1 >> 32) & mask;
  uint64_t m2 = (k2 & mask) | 1ULL, a2 = (k2 >> 32) & mask;
  uint64_t m3 = (k3 & mask) | 1ULL, a3 = (k3 >> 32) & mask;
  int h = (scale + 1) / 2;
  uint64_t v = x;
  v = (v * m1 + a1) & mask;
  v ^= v >> h;
  v = (v * m2 + a2) & mask;
  v ^= v >> h;
  v = (v * m3 + a3) & mask;
  return (uint32_t)v;
}

void og_rmat_generate(const og_rmat_params *p, void *out) {
  const uint64_t key = og_mix64(p->seed ^ 0xD1B54A32D192ED03ULL);
  const uint64_t ta = (uint64_t)(p->a * 4294967296.0);
  const uint64_t tab = (uint64_t)((p->a + p->b) * 4294967296.0);
  const uint64_t tabc = (uint64_t)((p->a + p->b + p->c) * 4294967296.0);
  const int scale = p->scale;
  const size_t rec = p->weighted ? 12 : 8;
  unsigned char *base = (unsigned char *)out;
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)p->num_edges; ii++) {
    uint64_t i = (uint64_t)ii;
    uint32_t src = 0, dst = 0;
    uint64_t w = 0;
    for (int l = 0; l < scale; l++) {
      if ((l & 1) == 0) w = og_word(key, i, (uint64_t)(l >> 1));
      uint64_t r = (l & 1) ? (w >> 32) : (w & 0xFFFFFFFFULL);
      uint32_t bit = 1u << (scale - 1 - l);
      if (r < ta) {
      } else if (r < tab) {
        dst |= bit;
      } else if (r < tabc) {
        src |= bit;
      } else {
        src |= bit;
        dst |= bit;
      }
    }
    if (p->permute) {
      src = og_rmat_permute(src, scale, p->seed);
      dst = og_rmat_permute(dst, scale, p->seed);
    }
    unsigned char *r = base + i * rec;
    memcpy(r, &src, 4);
    memcpy(r + 4, &dst, 4);
    if (p->weighted) {
      uint64_t k = og_word(key, i, 15) >> 40;
      float wt = (float)(k + 1ULL) * (1.0f / 16777216.0f);
      memcpy(r + 8, &wt, 4);
    }
  }
}
