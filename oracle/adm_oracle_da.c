/*
 * adm_oracle_da.c -- CPU restatement of adamantine's ensemble draws and the
 * EnKF analysis (TEST INFRASTRUCTURE: the checker for the HIP library, never
 * part of the product path).
 *
 *   ensemble draws   source/ensemble_management.cc:18-52, 166-265
 *   EnKF analysis    source/DataAssimilator.cc:95-189 (update_ensemble),
 *                    :403-476 (apply_kalman_gain), :478-502 (calc_H),
 *                    :614-628 (calc_Hx), :672-688 (gaspari_cohn_function),
 *                    :690-752 (calc_sample_covariance_sparse)
 *
 * The draws use std::mt19937 (default seed) and std::normal_distribution<>
 * of the C++ standard library the reference is built with (libstdc++).
 * Their published algorithms are restated here:
 *   - MT19937: Matsumoto & Nishimura, ACM TOMACS 8 (1998), with the C++11
 *     default seed 5489 and the initialisation multiplier 1812433253;
 *   - std::generate_canonical<double, 53> over a 32-bit engine
 *     ([rand.util.canonical]): k = 2 draws, (g0 + g1 * 2^32) / 2^64, a
 *     result of 1 replaced by the largest double below 1 (libstdc++);
 *   - libstdc++'s normal_distribution: Marsaglia's polar method with the
 *     second variate of each pair saved for the next call.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "adm_oracle.h"

typedef struct {
  uint32_t mt[624];
  int index;
} mt19937_t;

static void mt_seed(mt19937_t *g, uint32_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 624; ++i)
    g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->index = 624;
}

static uint32_t mt_next(mt19937_t *g) {
  if (g->index >= 624) {
    for (int i = 0; i < 624; ++i) {
      uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      g->mt[i] = v;
    }
    g->index = 0;
  }
  uint32_t y = g->mt[g->index++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

static double canonical53(mt19937_t *g) {
  const double r = 4294967296.0; /* 2^32 */
  double sum = (double)mt_next(g);
  sum += (double)mt_next(g) * r;
  double x = sum / (r * r);
  if (x >= 1.0) x = nextafter(1.0, 0.0);
  return x;
}

typedef struct {
  mt19937_t g;
  double saved;
  int saved_available;
} normal_t;

static double normal_next(normal_t *n, double mean, double stddev) {
  double ret;
  if (n->saved_available) {
    n->saved_available = 0;
    ret = n->saved;
  } else {
    double x, y, r2;
    do {
      x = 2.0 * canonical53(&n->g) - 1.0;
      y = 2.0 * canonical53(&n->g) - 1.0;
      r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double mult = sqrt(-2 * log(r2) / r2);
    n->saved = x * mult;
    n->saved_available = 1;
    ret = y * mult;
  }
  return ret * stddev + mean;
}

/* get_normal_random_vector (source/ensemble_management.cc:18-52) */
void oracle_normal_random_vector(uint32_t length, uint32_t n_rejected_draws,
                                 double mean, double stddev, double *out) {
  normal_t n;
  memset(&n, 0, sizeof(n));
  mt_seed(&n.g, 5489u);
  for (uint32_t i = 0; i < n_rejected_draws; ++i) (void)normal_next(&n, mean, stddev);
  for (uint32_t i = 0; i < length; ++i) {
    do {
      out[i] = normal_next(&n, mean, stddev);
    } while (out[i] < 0.);
  }
}

/* raw MT19937 output (default seed), for the engine's known answer */
void oracle_mt19937(uint32_t n, uint32_t *out) {
  mt19937_t g;
  mt_seed(&g, 5489u);
  for (uint32_t i = 0; i < n; ++i) out[i] = mt_next(&g);
}
