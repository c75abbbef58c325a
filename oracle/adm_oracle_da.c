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

/* ------------------------------------------------------------------------ */
/* EnKF analysis                                                             */
/* ------------------------------------------------------------------------ */

#include <stdlib.h>

/* DataAssimilator::gaspari_cohn_function (source/DataAssimilator.cc:672-688) */
double oracle_gaspari_cohn(double r) {
  if (r < 1.0)
    return 1. - 5. / 3. * pow(r, 2) + 5. / 8. * pow(r, 3) + 0.5 * pow(r, 4) -
           0.25 * pow(r, 5);
  else if (r < 2)
    return 4. - 5. * r + 5. / 3. * pow(r, 2) + 5. / 8. * pow(r, 3) -
           0.5 * pow(r, 4) + 1. / 12. * pow(r, 5) - 2. / (3. * r);
  return 0.;
}

/* Localisation factor of a (sim, sim) covariance entry at support-point
 * distance d: entries farther than the cut-off are outside the sparsity
 * pattern (update_covariance_sparsity_pattern, :512-612, ArborX sphere
 * query, boundary inclusive) and read as 0; inside it the factor is 1
 * (none), 1 unless d > cutoff (step_function) or GC(2 d / cutoff)
 * (gaspari_cohn) (calc_sample_covariance_sparse :727-745).
 * fn: 0 none, 1 step_function, 2 gaspari_cohn. */
double oracle_localization(int fn, double d, double cutoff) {
  if (!(d <= cutoff)) return 0.0;
  if (fn == 2) return oracle_gaspari_cohn(2.0 * d / cutoff);
  if (fn == 1 && d > cutoff) return 0.0;
  return 1.0;
}

static double point_distance(const double *coords, int dim, int64_t i,
                             int64_t j) {
  double s = 0.0;
  for (int d = 0; d < dim; ++d) {
    const double t = coords[i * dim + d] - coords[j * dim + d];
    s += t * t;
  }
  return sqrt(s);
}

/* calc_sample_covariance_sparse (:690-752), dense output P[n][n]:
 * ens[member][n] member-major; rows/columns >= n_sim are parameters
 * (factor 1); coords[n_sim][dim] support points (NULL: every distance
 * 0). */
void oracle_sample_covariance(int N, int64_t n, int64_t n_sim,
                              const double *ens, const double *coords, int dim,
                              int fn, double cutoff, double *P) {
  double *mean = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; ++i) {
    double s = ens[i];
    for (int k = 1; k < N; ++k) s += ens[(int64_t)k * n + i];
    mean[i] = s / N;
  }
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) {
      double loc = 1.0;
      if (i < n_sim && j < n_sim)
        loc = oracle_localization(
            fn, coords ? point_distance(coords, dim, i, j) : 0.0, cutoff);
      if (loc == 0.0) {
        P[i * n + j] = 0.0;
        continue;
      }
      double e = 0.0;
      for (int k = 0; k < N; ++k)
        e += (ens[(int64_t)k * n + i] - mean[i]) * (ens[(int64_t)k * n + j] - mean[j]);
      e /= (N - 1.0);
      P[i * n + j] = e * loc;
    }
  free(mean);
}

/* Solve A x = b (A m x m, row-major, destroyed) by Gaussian elimination with
 * partial pivoting in long double. */
static void solve_dense(int m, long double *A, long double *b) {
  for (int c = 0; c < m; ++c) {
    int piv = c;
    for (int r = c + 1; r < m; ++r)
      if (fabsl(A[(int64_t)r * m + c]) > fabsl(A[(int64_t)piv * m + c])) piv = r;
    if (piv != c) {
      for (int k = 0; k < m; ++k) {
        long double t = A[(int64_t)c * m + k];
        A[(int64_t)c * m + k] = A[(int64_t)piv * m + k];
        A[(int64_t)piv * m + k] = t;
      }
      long double t = b[c];
      b[c] = b[piv];
      b[piv] = t;
    }
    for (int r = c + 1; r < m; ++r) {
      const long double f = A[(int64_t)r * m + c] / A[(int64_t)c * m + c];
      if (f == 0.0L) continue;
      for (int k = c; k < m; ++k) A[(int64_t)r * m + k] -= f * A[(int64_t)c * m + k];
      b[r] -= f * b[c];
    }
  }
  for (int r = m - 1; r >= 0; --r) {
    long double s = b[r];
    for (int k = r + 1; k < m; ++k) s -= A[(int64_t)r * m + k] * b[k];
    b[r] = s / A[(int64_t)r * m + r];
  }
}

/* apply_kalman_gain (:403-476): out[j] = P H^T (H P H^T + R)^{-1} innov[j]
 * for every member, with P from oracle_sample_covariance and H the
 * observation -> DoF selection of calc_H (:478-502). The reference applies
 * the inverse with GMRES to 1e-10 relative residual; here it is a direct
 * solve in extended precision. R[m][m] row-major; innov[N][m];
 * out[N][n]. */
void oracle_apply_kalman_gain(int N, int64_t n, int64_t n_sim,
                              const double *ens, const double *coords, int dim,
                              int fn, double cutoff, int m,
                              const int64_t *obs_dof, const double *R,
                              const double *innov, double *out) {
  double *P = (double *)malloc(sizeof(double) * n * n);
  oracle_sample_covariance(N, n, n_sim, ens, coords, dim, fn, cutoff, P);
  long double *C = (long double *)malloc(sizeof(long double) * m * m);
  long double *e = (long double *)malloc(sizeof(long double) * m);
  for (int j = 0; j < N; ++j) {
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b)
        C[(int64_t)a * m + b] = (long double)P[obs_dof[a] * n + obs_dof[b]] +
                                R[(int64_t)a * m + b];
    for (int a = 0; a < m; ++a) e[a] = innov[(int64_t)j * m + a];
    solve_dense(m, C, e);
    for (int64_t i = 0; i < n; ++i) {
      long double s = 0.0L;
      for (int a = 0; a < m; ++a) s += (long double)P[i * n + obs_dof[a]] * e[a];
      out[(int64_t)j * n + i] = (double)s;
    }
  }
  free(e);
  free(C);
  free(P);
}

/* update_ensemble (:95-189) with given observation noise:
 * innov_j = noise_j + (y - H x_j), x_j += K innov_j. ens[N][n] in place. */
void oracle_update_ensemble(int N, int64_t n, int64_t n_sim, double *ens,
                            const double *coords, int dim, int fn,
                            double cutoff, int m, const int64_t *obs_dof,
                            const double *y, const double *R,
                            const double *noise) {
  double *innov = (double *)malloc(sizeof(double) * N * m);
  double *shift = (double *)malloc(sizeof(double) * N * n);
  for (int j = 0; j < N; ++j)
    for (int a = 0; a < m; ++a) {
      double v = noise ? noise[(int64_t)j * m + a] : 0.0;
      v += y[a] - ens[(int64_t)j * n + obs_dof[a]];
      innov[(int64_t)j * m + a] = v;
    }
  oracle_apply_kalman_gain(N, n, n_sim, ens, coords, dim, fn, cutoff, m,
                           obs_dof, R, innov, shift);
  for (int64_t k = 0; k < (int64_t)N * n; ++k) ens[k] += shift[k];
  free(shift);
  free(innov);
}

/* fill_noise_vector (:630-670) for `count` consecutive vectors of size m
 * from the DataAssimilator's own default-seeded mt19937 and N(0, 1)
 * distribution (source/DataAssimilator.hh:280-286), whose state persists
 * across calls: *state must point to a zeroed oracle_da_rng_t at first use.
 * Diagonal R: sqrt(R_ii) z_i; otherwise L z with R = L L^T (Cholesky). */
typedef struct {
  normal_t n;
  int seeded;
} oracle_da_rng_t;

size_t oracle_da_rng_size(void) { return sizeof(oracle_da_rng_t); }

void oracle_fill_noise(void *state, int count, int m, const double *R,
                       double *out) {
  oracle_da_rng_t *s = (oracle_da_rng_t *)state;
  if (!s->seeded) {
    memset(s, 0, sizeof(*s));
    mt_seed(&s->n.g, 5489u);
    s->seeded = 1;
  }
  int diagonal = 1;
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b)
      if (a != b && R[(int64_t)a * m + b] != 0.0) diagonal = 0;
  double *L = NULL;
  if (!diagonal) {
    L = (double *)calloc((size_t)m * m, sizeof(double));
    for (int j = 0; j < m; ++j) {
      double d = R[(int64_t)j * m + j];
      for (int k = 0; k < j; ++k) d -= L[(int64_t)j * m + k] * L[(int64_t)j * m + k];
      L[(int64_t)j * m + j] = sqrt(d);
      for (int i = j + 1; i < m; ++i) {
        double v = R[(int64_t)i * m + j];
        for (int k = 0; k < j; ++k) v -= L[(int64_t)i * m + k] * L[(int64_t)j * m + k];
        L[(int64_t)i * m + j] = v / L[(int64_t)j * m + j];
      }
    }
  }
  double *z = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
  for (int c = 0; c < count; ++c) {
    double *v = out + (int64_t)c * m;
    if (diagonal) {
      for (int i = 0; i < m; ++i)
        v[i] = normal_next(&s->n, 0.0, 1.0) * sqrt(R[(int64_t)i * m + i]);
    } else {
      for (int i = 0; i < m; ++i) z[i] = normal_next(&s->n, 0.0, 1.0);
      for (int i = 0; i < m; ++i) {
        double acc = 0.0;
        for (int k = 0; k <= i; ++k) acc += L[(int64_t)i * m + k] * z[k];
        v[i] = acc;
      }
    }
  }
  free(z);
  free(L);
}
