/*
 * adm_oracle.c -- CPU restatement of adamantine's host matrix-free heat path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the HIP library is
 * compared against; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py load it. Nothing in adamantine_amd/ links or calls it.
 *
 * Reference (paths relative to the adamantine source tree):
 *   cell kernel        source/ThermalOperator.templates.hh:479-617
 *   state update       source/ThermalOperator.templates.hh:227-317
 *   1/(rho c_p)        source/ThermalOperator.templates.hh:415-477
 *   face kernel        source/ThermalOperator.templates.hh:619-751, 319-413
 *   lumped mass        source/ThermalOperator.templates.hh:81-164
 *   properties         source/MaterialProperty.hh:529-800,
 *                      source/MaterialProperty.templates.hh:1158-1212
 *   sources            source/GoldakHeatSource.cc:24-148,
 *                      source/ElectronBeamHeatSource.cc:24-158,
 *                      source/CubeHeatSource.cc:45-107
 *   scan path          source/ScanPath.cc:197-275
 *   time stepping      source/ThermalPhysics.templates.hh:43-124, 854-888
 *                      (+ deal.II 9.6 TimeStepping::ExplicitRungeKutta)
 *
 * Lane semantics: one SIMD lane (W = 1); see adm_oracle.h.
 * Floating point: built with -ffp-contract=off; the sum-factorisation sweeps
 * use explicit fma() chains, so the values and gradients at quadrature points
 * follow a fixed operation order.
 */
#include "adm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 6     /* fe_degree + 1 <= 6 */
#define MAXQ 216   /* (fe_degree + 1)^3 */
#define MAXFQ 36

static const double STEFAN_BOLTZMANN = 5.670374419e-8; /* types.hh:158 */

struct oracle {
  int dim, p, n1, nq, nfq;
  int nc[3];
  int64_t nn[3];
  int64_t n_cells, n_dofs;
  int axis[3];
  double h[3], org[3], inv_h[3];
  double xs[MAXN], xq[MAXN], wq[MAXN], wgl[MAXN];
  double S[MAXN][MAXN], Dc[MAXN][MAXN];
  double jxw[MAXQ];
  double fjxw[3][MAXFQ]; /* per normal direction */
  adm_material_desc mat;
  int n_bids;
  int btype[ADM_MAX_BOUNDARY_IDS];
  int adiabatic_only;
  int n_beams;
  adm_beam_desc beams[ADM_MAX_BEAMS];
  adm_segment *segs[ADM_MAX_BEAMS];
  int old_seg[ADM_MAX_BEAMS];
  adm_beam_state bs[ADM_MAX_BEAMS];
  int sources_on;
  uint8_t *active;
  uint8_t *material;
  double *dcos, *dsin;
  double *liquid, *powder, *face_powder;
  double *inv_mass;
};

/* ------------------------------------------------------------------------ */
/* 1-D quadrature and Lagrange bases                                         */
/* ------------------------------------------------------------------------ */

static void legendre(int n, long double x, long double *p, long double *dp) {
  long double p0 = 1.0L, p1 = x;
  if (n == 0) {
    *p = 1.0L;
    *dp = 0.0L;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    long double pk = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
    p0 = p1;
    p1 = pk;
  }
  *p = p1;
  *dp = n * (x * p1 - p0) / (x * x - 1.0L);
}

/* QGauss(n) on [0,1] */
static void gauss(int n, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) /
                         (n + 0.5L));
    for (int it = 0; it < 100; ++it) {
      long double p, dp;
      legendre(n, z, &p, &dp);
      long double dz = p / dp;
      z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    long double p, dp;
    legendre(n, z, &p, &dp);
    long double ww = 2.0L / ((1.0L - z * z) * dp * dp);
    /* ascending order on [0,1] */
    x[n - 1 - i] = (double)(0.5L + 0.5L * z);
    w[n - 1 - i] = (double)(0.5L * ww);
  }
}

/* QGaussLobatto(n) on [0,1] (n >= 2) */
static void gauss_lobatto(int n, double *x, double *w) {
  int m = n - 1;
  x[0] = 0.0;
  x[n - 1] = 1.0;
  long double pe, dpe;
  legendre(m, 1.0L - 1e-30L, &pe, &dpe);
  w[0] = w[n - 1] = (double)(0.5L * 2.0L / (m * (m + 1.0L)));
  for (int i = 1; i < n - 1; ++i) {
    /* interior points: roots of P'_m */
    long double z = -cosl(3.14159265358979323846264338327950288L * i / m);
    for (int it = 0; it < 200; ++it) {
      /* Newton on P'_m using (1-z^2) P'' = 2 z P' - m(m+1) P */
      long double p, dp;
      legendre(m, z, &p, &dp);
      long double d2p = (2.0L * z * dp - m * (m + 1.0L) * p) / (1.0L - z * z);
      long double dz = dp / d2p;
      z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    long double p, dp;
    legendre(m, z, &p, &dp);
    x[i] = (double)(0.5L + 0.5L * z);
    w[i] = (double)(0.5L * 2.0L / (m * (m + 1.0L) * p * p));
  }
}

static double lagrange(int n, const double *pts, int i, double x) {
  double v = 1.0;
  for (int j = 0; j < n; ++j)
    if (j != i) v *= (x - pts[j]) / (pts[i] - pts[j]);
  return v;
}

static double lagrange_deriv(int n, const double *pts, int k, double x) {
  double s = 0.0;
  for (int m = 0; m < n; ++m) {
    if (m == k) continue;
    double t = 1.0 / (pts[k] - pts[m]);
    for (int j = 0; j < n; ++j)
      if (j != k && j != m) t *= (x - pts[j]) / (pts[k] - pts[j]);
    s += t;
  }
  return s;
}

/* ------------------------------------------------------------------------ */
/* Material properties                                                       */
/* ------------------------------------------------------------------------ */

/* MaterialProperty::compute_property_from_table
 * (source/MaterialProperty.templates.hh:1158-1212) */
double oracle_property_from_table(const adm_material_desc *m, int mid, int prop,
                                  int state, double T) {
  const double(*t)[2] = m->table[mid][prop][state];
  if (T <= t[0][0]) return t[0][1];
  int i = 0;
  const int size = ADM_TABLE_SIZE;
  for (; i < size; ++i)
    if (T < t[i][0]) break;
  if (i >= size - 1) return t[size - 1][1];
  return t[i - 1][1] + (T - t[i - 1][0]) * (t[i][1] - t[i - 1][1]) /
                           (t[i][0] - t[i - 1][0]);
}

/* MaterialProperty::compute_material_property, vectorized host form
 * (source/MaterialProperty.hh:529-800) with precomputed temperature powers
 * (source/ThermalOperator.templates.hh:524-530). */
static double property_tp(const adm_material_desc *m, int prop, int mid,
                          const double *r, double T, const double *tp) {
  double value = 0.0;
  const int ns = m->n_states;
  if (m->use_table) {
    for (int s = 0; s < ns; ++s) {
      double pv = oracle_property_from_table(m, mid, prop, s, T);
      value += r[s] * pv;
    }
  } else if (m->single_material_layout) {
    for (int i = 0; i <= m->p_order; ++i)
      for (int s = 0; s < ns; ++s) {
        double c = m->poly[mid][prop][s][i];
        value += r[s] * c * tp[i];
      }
  } else {
    for (int s = 0; s < ns; ++s)
      for (int i = 0; i <= m->p_order; ++i) {
        double c = m->poly[mid][prop][s][i];
        value += r[s] * c * tp[i];
      }
  }
  return value;
}

double oracle_material_property(const adm_material_desc *m, int prop, int mid,
                                const double *r, double T) {
  double tp[ADM_MAX_POLY];
  tp[0] = 1.0;
  for (int i = 1; i <= m->p_order; ++i) tp[i] = tp[i - 1] * T;
  return property_tp(m, prop, mid, r, T, tp);
}

/* update_state_ratios (source/ThermalOperator.templates.hh:227-317) for one
 * lane; powder_io holds the stored powder ratio (SLP). Returns ratios in r. */
static void state_ratios(const adm_material_desc *m, int mid, double T,
                         double *liquid_io, double *powder_io, double *r) {
  const int ns = m->n_states;
  if (ns == 1) {
    r[0] = 1.0;
    return;
  }
  double solidus = m->prop[mid][ADM_SOLIDUS];
  double liquidus = m->prop[mid][ADM_LIQUIDUS];
  double l;
  if (T < solidus)
    l = 0.0;
  else if (T > liquidus)
    l = 1.0;
  else
    l = (T - solidus) / (liquidus - solidus);
  if (ns == 2) {
    r[1] = l;
    r[0] = 1.0 - l;
    if (liquid_io) *liquid_io = l;
    return;
  }
  double p = *powder_io;
  double a = 1.0 - l;
  /* std::min(1 - liquid, powder) */
  p = (p < a) ? p : a;
  r[1] = l;
  r[2] = p;
  r[0] = 1.0 - l - p;
  if (liquid_io) *liquid_io = l;
  *powder_io = p;
}

/* get_inv_rho_cp (source/ThermalOperator.templates.hh:415-477) */
static double inv_rho_cp(const adm_material_desc *m, int mid, const double *r,
                         double T, const double *tp) {
  double density = property_tp(m, ADM_DENSITY, mid, r, T, tp);
  double specific_heat = property_tp(m, ADM_SPECIFIC_HEAT, mid, r, T, tp);
  if (m->n_states > 1) {
    double solidus = m->prop[mid][ADM_SOLIDUS];
    double liquidus = m->prop[mid][ADM_LIQUIDUS];
    double latent = m->prop[mid][ADM_LATENT_HEAT];
    double is_mushy = r[1] * r[0];
    if (is_mushy > 0.0) specific_heat += latent / (liquidus - solidus);
  }
  return 1.0 / (density * specific_heat);
}

/* ------------------------------------------------------------------------ */
/* Scan path and heat sources                                                */
/* ------------------------------------------------------------------------ */

/* ScanPath::update_current_segment_info (source/ScanPath.cc:197-226) */
static int current_segment(const adm_beam_desc *b, double t, int *old_segment,
                           int save, double sp[3], double *st) {
  int cur = old_segment ? *old_segment : 0;
  while (t > b->segments[cur].end_time) ++cur;
  if (cur > 0) {
    *st = b->segments[cur - 1].end_time;
    for (int d = 0; d < 3; ++d) sp[d] = b->segments[cur - 1].end_point[d];
  } else {
    *st = 0.0;
    for (int d = 0; d < 3; ++d) sp[d] = b->segments[cur].end_point[d];
  }
  if (save && old_segment) *old_segment = cur;
  return cur;
}

/* ScanPath::value (source/ScanPath.cc:228-257). deal.II's Tensor division by
 * a scalar multiplies by the reciprocal. */
void oracle_scan_path_position(const adm_beam_desc *b, double t,
                               int *old_segment, double out[3]) {
  const adm_segment *last = &b->segments[b->n_segments - 1];
  if (t > last->end_time) {
    for (int d = 0; d < 3; ++d) out[d] = -DBL_MAX;
    return;
  }
  double sp[3], st;
  int cur = current_segment(b, t, old_segment, 1, sp, &st);
  double inv = 1.0 / (b->segments[cur].end_time - st);
  for (int d = 0; d < 3; ++d)
    out[d] = sp[d] + ((b->segments[cur].end_point[d] - sp[d]) * inv) * (t - st);
}

/* ScanPath::get_power_modifier (source/ScanPath.cc:259-275) */
double oracle_scan_path_power_modifier(const adm_beam_desc *b, double t,
                                       int *old_segment) {
  const adm_segment *last = &b->segments[b->n_segments - 1];
  if (t > last->end_time) return 0.0;
  double sp[3], st;
  int cur = current_segment(b, t, old_segment, 1, sp, &st);
  return b->segments[cur].power_modifier;
}

/* ------------------------------------------------------------------------ */
/* Quaternions of five-axis scan paths (source/Quaternion.cc)                 */
/* ------------------------------------------------------------------------ */

/* Quaternion::operator*=(Quaternion) (source/Quaternion.cc:62-83): the
 * Hamilton product a <- a b. */
void oracle_quaternion_mul(double a[4], const double b[4]) {
  double r = a[0], i = a[1], j = a[2], k = a[3];
  double br = b[0], bi = b[1], bj = b[2], bk = b[3];  /* b may alias a */
  a[0] = r * br - i * bi - j * bj - k * bk;
  a[1] = r * bi + i * br + j * bk - k * bj;
  a[2] = r * bj - i * bk + j * br + k * bi;
  a[3] = r * bk + i * bj - j * bi + k * br;
}

/* Quaternion::operator/=(Quaternion) (source/Quaternion.cc:92-99): a <- a
 * conj(b), b being of unit norm. */
void oracle_quaternion_div(double a[4], const double b[4]) {
  double c[4] = {b[0], -b[1], -b[2], -b[3]};
  oracle_quaternion_mul(a, c);
}

/* Quaternion::pow (source/Quaternion.cc:101-118): the axis-angle power,
 * theta = atan2(|v|, r). */
void oracle_quaternion_pow(double q[4], double e) {
  double pn = sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  double u[3] = {q[1] / pn, q[2] / pn, q[3] / pn};
  double theta = atan2(pn, q[0]);
  q[0] = cos(e * theta);
  for (int d = 0; d < 3; ++d) q[d + 1] = u[d] * sin(e * theta);
}

/* Quaternion::build_rotation_matrices (source/Quaternion.cc:120-150):
 * the rotation matrix, row-major; the inverse is its transpose. */
void oracle_quaternion_matrix(const double q[4], double R[9]) {
  double r = q[0], i = q[1], j = q[2], k = q[3];
  R[0] = 1.0 - 2.0 * (j * j + k * k);
  R[1] = 2.0 * (i * j - k * r);
  R[2] = 2.0 * (i * k + j * r);
  R[3] = 2.0 * (i * j + k * r);
  R[4] = 1.0 - 2.0 * (i * i + k * k);
  R[5] = 2.0 * (j * k - i * r);
  R[6] = 2.0 * (i * k - j * r);
  R[7] = 2.0 * (j * k + i * r);
  R[8] = 1.0 - 2.0 * (i * i + j * j);
}

/* Tensor<2,3> * Point<3> (Quaternion::rotate, source/Quaternion.hh:
 * 135-139); inverse = 1 multiplies by the transpose (inv_rotate). */
void oracle_rotate(const double R[9], int inverse, const double p[3],
                   double out[3]) {
  double r[3];
  for (int a = 0; a < 3; ++a) {
    double acc = 0.0;
    for (int b = 0; b < 3; ++b)
      acc += (inverse ? R[3 * b + a] : R[3 * a + b]) * p[b];
    r[a] = acc;
  }
  for (int a = 0; a < 3; ++a) out[a] = r[a];
}

static int five_axis(const adm_beam_desc *b) {
  const double *q = b->segments[0].end_rotation;
  return q[0] != 0.0 || q[1] != 0.0 || q[2] != 0.0 || q[3] != 0.0;
}

/* ScanPath::get_quaternion (source/ScanPath.cc:290-341): identity on a
 * three-axis path, the last rotation after the end, else the slerp between
 * the segment's start and end rotations on the shorter arc. */
void oracle_scan_path_quaternion(const adm_beam_desc *b, double t,
                                 int *old_segment, double q[4]) {
  q[0] = 1.0;
  q[1] = q[2] = q[3] = 0.0;
  if (!five_axis(b)) return;
  const adm_segment *last = &b->segments[b->n_segments - 1];
  if (t > last->end_time) {
    for (int d = 0; d < 4; ++d) q[d] = last->end_rotation[d];
    return;
  }
  double sp[3], st;
  int cur = current_segment(b, t, old_segment, 0, sp, &st);
  const double *q0 = cur > 0 ? b->segments[cur - 1].end_rotation
                             : b->segments[cur].end_rotation;
  const double *q1 = b->segments[cur].end_rotation;
  if (q0[0] == q1[0] && q0[1] == q1[1] && q0[2] == q1[2] && q0[3] == q1[3]) {
    for (int d = 0; d < 4; ++d) q[d] = q0[d];
    return;
  }
  double duration = b->segments[cur].end_time - st;
  for (int d = 0; d < 4; ++d) q[d] = q1[d];
  if (duration < 1e-12) return;
  double dot = 0.0;
  for (int d = 0; d < 4; ++d) dot += q0[d] * q[d];
  if (dot < 0.0)
    for (int d = 0; d < 4; ++d) q[d] *= -1.0;
  oracle_quaternion_div(q, q0);
  oracle_quaternion_pow(q, (t - st) / duration);
  oracle_quaternion_mul(q, q0);
}

/* HeatSource::update_time for the three source kinds */
void oracle_beam_update(const adm_beam_desc *b, double t, int *old_segment,
                        adm_beam_state *s) {
  memset(s, 0, sizeof(*s));
  s->type = b->type;
  if (b->type == ADM_SOURCE_CUBE) {
    /* source/CubeHeatSource.cc:45-49 */
    s->on = (t > b->start_time) && (t < b->end_time);
    s->cube_value = b->value;
    for (int d = 0; d < 3; ++d) {
      s->cube_min[d] = b->min_point[d];
      s->cube_max[d] = b->max_point[d];
    }
    return;
  }
  double pmod = oracle_scan_path_power_modifier(b, t, old_segment);
  s->on = pmod > 0.0;
  oracle_scan_path_position(b, t, old_segment, s->center);
  if (five_axis(b)) {
    /* source/GoldakHeatSource.cc:31, ElectronBeamHeatSource.cc:31 */
    double q[4];
    oracle_scan_path_quaternion(b, t, old_segment, q);
    oracle_quaternion_matrix(q, s->rotation);
    s->rotated = 1;
  }
  double r2 = pow(b->radius, 2);
  double d = b->depth;
  s->depth = d;
  s->radius_squared = r2;
  if (b->type == ADM_SOURCE_GOLDAK) {
    /* source/GoldakHeatSource.cc:24-44 */
    double pi_over_3_to_1p5 = pow(3.14159265358979323846 / 3.0, 1.5);
    s->alpha = 2.0 * b->absorption_efficiency * b->max_power * pmod /
               (r2 * d * pi_over_3_to_1p5);
    s->inv_radius_squared = 1.0 / r2;
    s->inv_depth_squared = 1.0 / (d * d);
  } else {
    /* source/ElectronBeamHeatSource.cc:24-45 */
    double log_01 = log(0.1);
    s->alpha = -b->absorption_efficiency * b->max_power * pmod * log_01 /
               (3.14159265358979323846 * r2 * d);
    s->inv_depth = 1.0 / d;
    s->inv_radius_squared = 1.0 / r2;
  }
}

/* HeatSource::value at one point, vectorized form with one lane
 * (source/GoldakHeatSource.cc:91-148, ElectronBeamHeatSource.cc:95-158,
 * CubeHeatSource.cc:73-107). In 2D the point is (x, z) and the beam
 * centre components 0 and 1 are used, as in the reference. */
double oracle_beam_value(const adm_beam_state *s, int dim, const double *pt) {
  if (s->type == ADM_SOURCE_CUBE) {
    if (!s->on) return 0.0;
    for (int d = 0; d < dim; ++d)
      if (!(pt[d] >= s->cube_min[d] && pt[d] <= s->cube_max[d])) return 0.0;
    return 1.0 * s->cube_value;
  }
  double rp[3];
  if (dim == 3 && s->rotated) {
    /* the five-axis point (source/GoldakHeatSource.cc:95-103) */
    oracle_rotate(s->rotation, 0, pt, rp);
    pt = rp;
  }
  int zi = dim - 1;
  double z = pt[zi] - s->center[zi];
  if (z + s->depth < 0.0) return 0.0;
  double dx = pt[0] - s->center[0];
  double xpy_squared = dx * dx;
  if (dim == 3) {
    double dy = pt[1] - s->center[1];
    xpy_squared += dy * dy;
  }
  double xpy_area = 5. * s->radius_squared - xpy_squared;
  if (xpy_area < 0.) return 0.0;
  if (s->type == ADM_SOURCE_GOLDAK) {
    double exponent = z * z * s->inv_depth_squared;
    exponent += xpy_squared * s->inv_radius_squared;
    exponent *= -3.0;
    return 1.0 * s->alpha * exp(exponent);
  }
  double log_01 = log(0.1);
  double z_rel = z * s->inv_depth;
  double dist = z_rel * z_rel;
  dist *= -3.0;
  dist -= z_rel;
  dist -= z_rel;
  dist += 1.0;
  double exponent = log_01 * xpy_squared;
  exponent *= s->inv_radius_squared;
  return 1.0 * s->alpha * exp(exponent) * dist;
}

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

/* lumped mass with QGaussLobatto(p+1) over the active cells
 * (ThermalOperator.templates.hh:118-164), inverted; shape functions at their
 * own support points are the identity */
static void compute_inverse_mass(oracle_t *o) {
  double det = 1.0;
  for (int d = 0; d < o->dim; ++d) det *= o->h[d];
  memset(o->inv_mass, 0, sizeof(double) * o->n_dofs);
  for (int64_t c = 0; c < o->n_cells; ++c) {
    if (!o->active[c]) continue;
    int64_t cc[3] = {0, 0, 0}, r = c;
    for (int d = 0; d < o->dim; ++d) {
      cc[d] = r % o->nc[d];
      r /= o->nc[d];
    }
    for (int i = 0; i < o->nq; ++i) {
      int64_t node = 0, stride = 1;
      int ri = i;
      double w = 1.0;
      for (int d = 0; d < o->dim; ++d) {
        int li = ri % o->n1;
        ri /= o->n1;
        node += (cc[d] * o->p + li) * stride;
        stride *= o->nn[d];
        w *= o->wgl[li];
      }
      o->inv_mass[node] += det * w;
    }
  }
  for (int64_t i = 0; i < o->n_dofs; ++i)
    o->inv_mass[i] = o->inv_mass[i] > 0.0 ? 1.0 / o->inv_mass[i] : 0.0;
}

oracle_t *oracle_create(const adm_problem_desc *desc) {
  const adm_mesh_desc *m = &desc->mesh;
  if (m->fe_degree < 1 || m->fe_degree > MAXN - 1 || m->dim < 2 ||
      m->dim > 3)
    return NULL;
  oracle_t *o = (oracle_t *)calloc(1, sizeof(oracle_t));
  o->dim = m->dim;
  o->p = m->fe_degree;
  o->n1 = o->p + 1;
  o->nq = ipow(o->n1, o->dim);
  o->nfq = ipow(o->n1, o->dim - 1);
  if (o->dim == 3) {
    o->axis[0] = 0;
    o->axis[1] = 1;
    o->axis[2] = 2;
  } else {
    o->axis[0] = 0;
    o->axis[1] = 2;
  }
  o->n_cells = 1;
  o->n_dofs = 1;
  for (int d = 0; d < o->dim; ++d) {
    int a = o->axis[d];
    o->nc[d] = m->n_cells[a];
    o->nn[d] = (int64_t)o->p * o->nc[d] + 1;
    o->n_cells *= o->nc[d];
    o->n_dofs *= o->nn[d];
    o->org[d] = m->origin[a];
    o->h[d] = m->extent[a] / o->nc[d];
    o->inv_h[d] = 1.0 / o->h[d];
  }
  /* support points: QGaussLobatto(p+1) (FE_Q default) */
  double wtmp[MAXN];
  gauss_lobatto(o->n1, o->xs, o->wgl);
  if (m->quadrature == 1)
    gauss_lobatto(o->n1, o->xq, o->wq);
  else
    gauss(o->n1, o->xq, o->wq);
  (void)wtmp;
  for (int q = 0; q < o->n1; ++q)
    for (int i = 0; i < o->n1; ++i) {
      o->S[q][i] = lagrange(o->n1, o->xs, i, o->xq[q]);
      o->Dc[q][i] = lagrange_deriv(o->n1, o->xq, i, o->xq[q]);
    }
  double det = 1.0;
  for (int d = 0; d < o->dim; ++d) det *= o->h[d];
  for (int q = 0; q < o->nq; ++q) {
    double w = 1.0;
    int r = q;
    for (int d = 0; d < o->dim; ++d) {
      w *= o->wq[r % o->n1];
      r /= o->n1;
    }
    o->jxw[q] = det * w;
  }
  for (int nd = 0; nd < o->dim; ++nd) {
    double fdet = 1.0;
    for (int d = 0; d < o->dim; ++d)
      if (d != nd) fdet *= o->h[d];
    for (int q = 0; q < o->nfq; ++q) {
      double w = 1.0;
      int r = q;
      for (int d = 0; d < o->dim - 1; ++d) {
        w *= o->wq[r % o->n1];
        r /= o->n1;
      }
      o->fjxw[nd][q] = fdet * w;
    }
  }
  o->mat = desc->material;
  o->n_bids = desc->n_boundary_ids;
  int n_adiabatic = 0;
  for (int i = 0; i < o->n_bids; ++i) {
    o->btype[i] = desc->boundary_type[i];
    if (o->btype[i] & ADM_BC_ADIABATIC) ++n_adiabatic;
  }
  o->adiabatic_only = (n_adiabatic == o->n_bids);
  o->n_beams = desc->n_beams;
  for (int b = 0; b < o->n_beams; ++b) {
    o->beams[b] = desc->beams[b];
    int ns = desc->beams[b].n_segments;
    if (ns > 0) {
      o->segs[b] = (adm_segment *)malloc(sizeof(adm_segment) * ns);
      memcpy(o->segs[b], desc->beams[b].segments, sizeof(adm_segment) * ns);
    }
    o->beams[b].segments = o->segs[b];
    o->old_seg[b] = 0;
  }
  o->active = (uint8_t *)malloc(o->n_cells);
  o->material = (uint8_t *)malloc(o->n_cells);
  o->dcos = (double *)malloc(sizeof(double) * o->n_cells);
  o->dsin = (double *)malloc(sizeof(double) * o->n_cells);
  for (int64_t c = 0; c < o->n_cells; ++c) {
    o->active[c] = desc->cell_active ? desc->cell_active[c] : 1;
    o->material[c] = desc->cell_material ? desc->cell_material[c] : 0;
    o->dcos[c] = desc->deposition_cos ? desc->deposition_cos[c] : 1.0;
    o->dsin[c] = desc->deposition_sin ? desc->deposition_sin[c] : 0.0;
  }
  o->liquid = (double *)calloc(o->n_cells * o->nq, sizeof(double));
  o->powder = (double *)calloc(o->n_cells * o->nq, sizeof(double));
  o->face_powder =
      (double *)calloc(o->n_cells * 2 * o->dim * o->nfq, sizeof(double));
  if (desc->cell_state) oracle_set_cell_state(o, desc->cell_state);

  o->inv_mass = (double *)calloc(o->n_dofs, sizeof(double));
  compute_inverse_mass(o);
  return o;
}

void oracle_destroy(oracle_t *o) {
  if (!o) return;
  for (int b = 0; b < o->n_beams; ++b) free(o->segs[b]);
  free(o->active);
  free(o->material);
  free(o->dcos);
  free(o->dsin);
  free(o->liquid);
  free(o->powder);
  free(o->face_powder);
  free(o->inv_mass);
  free(o);
}

int64_t oracle_n_dofs(const oracle_t *o) { return o->n_dofs; }
int64_t oracle_n_cells(const oracle_t *o) { return o->n_cells; }
int oracle_n_q(const oracle_t *o) { return o->nq; }
int oracle_n_face_q(const oracle_t *o) { return o->nfq; }

void oracle_basis_1d(const oracle_t *o, double *support, double *qpoints,
                     double *qweights, double *shape, double *coll_deriv,
                     double *gl_weights) {
  for (int i = 0; i < o->n1; ++i) {
    if (support) support[i] = o->xs[i];
    if (qpoints) qpoints[i] = o->xq[i];
    if (qweights) qweights[i] = o->wq[i];
    if (gl_weights) gl_weights[i] = o->wgl[i];
    for (int j = 0; j < o->n1; ++j) {
      if (shape) shape[i * o->n1 + j] = o->S[i][j];
      if (coll_deriv) coll_deriv[i * o->n1 + j] = o->Dc[i][j];
    }
  }
}

void oracle_inverse_mass(const oracle_t *o, double *out) {
  memcpy(out, o->inv_mass, sizeof(double) * o->n_dofs);
}

void oracle_set_cell_state(oracle_t *o, const double *cs) {
  const int ns = o->mat.n_states;
  const int nf = 2 * o->dim * o->nfq;
  for (int64_t c = 0; c < o->n_cells; ++c) {
    double l = ns > 1 ? cs[c * ns + 1] : 0.0;
    double p = ns > 2 ? cs[c * ns + 2] : 0.0;
    for (int q = 0; q < o->nq; ++q) {
      o->liquid[c * o->nq + q] = l;
      o->powder[c * o->nq + q] = p;
    }
    for (int q = 0; q < nf; ++q) o->face_powder[c * nf + q] = p;
  }
}

void oracle_get_cell_state(const oracle_t *o, double *cs) {
  const int ns = o->mat.n_states;
  for (int64_t c = 0; c < o->n_cells; ++c) {
    if (ns == 1) {
      cs[c] = 1.0;
      continue;
    }
    double ls = 0.0, ps = 0.0;
    for (int q = 0; q < o->nq; ++q) {
      ls += o->liquid[c * o->nq + q];
      ps += o->powder[c * o->nq + q];
    }
    double l = ls / o->nq;
    if (ns == 2) {
      cs[c * 2 + 1] = l;
      cs[c * 2 + 0] = 1. - l;
    } else {
      double p = ps / o->nq;
      cs[c * 3 + 1] = l;
      cs[c * 3 + 2] = p;
      cs[c * 3 + 0] = 1. - l - p;
    }
  }
}

void oracle_get_qp_state(const oracle_t *o, double *liquid, double *powder) {
  if (liquid) memcpy(liquid, o->liquid, sizeof(double) * o->n_cells * o->nq);
  if (powder) memcpy(powder, o->powder, sizeof(double) * o->n_cells * o->nq);
}

void oracle_set_qp_state(oracle_t *o, const double *liquid,
                         const double *powder) {
  if (liquid) memcpy(o->liquid, liquid, sizeof(double) * o->n_cells * o->nq);
  if (powder) memcpy(o->powder, powder, sizeof(double) * o->n_cells * o->nq);
}

void oracle_get_face_powder(const oracle_t *o, double *fp) {
  memcpy(fp, o->face_powder, sizeof(double) * o->n_cells * 2 * o->dim * o->nfq);
}

void oracle_set_deposition(oracle_t *o, const double *c, const double *s) {
  for (int64_t i = 0; i < o->n_cells; ++i) {
    o->dcos[i] = c[i];
    o->dsin[i] = s[i];
  }
}

void oracle_set_beam_parameters(oracle_t *o, int b, double absorption,
                                double max_power, double radius,
                                double depth) {
  o->beams[b].absorption_efficiency = absorption;
  o->beams[b].max_power = max_power;
  o->beams[b].radius = radius;
  o->beams[b].depth = depth;
}

/* ThermalOperator::set_time (source/ThermalOperator.hh:298-308) */
void oracle_set_time(oracle_t *o, double t) {
  o->sources_on = 0;
  for (int b = 0; b < o->n_beams; ++b) {
    oracle_beam_update(&o->beams[b], t, &o->old_seg[b], &o->bs[b]);
    if (o->bs[b].on) o->sources_on = 1;
  }
}

/* ------------------------------------------------------------------------ */
/* Cell and face kernels                                                     */
/* ------------------------------------------------------------------------ */

/* out = M (or M^T) applied along direction d of an n^dim tensor (x fastest);
 * accumulate adds the contraction to out. */
static void sweep(int dim, int n, int d, const double M[MAXN][MAXN],
                  int transpose, const double *in, double *out,
                  int accumulate) {
  int stride = ipow(n, d);
  int total = ipow(n, dim);
  for (int o = 0; o < total; ++o) {
    int c = (o / stride) % n;
    int base = o - c * stride;
    double a0 = transpose ? M[0][c] : M[c][0];
    double acc = a0 * in[base];
    for (int i = 1; i < n; ++i) {
      double a = transpose ? M[i][c] : M[c][i];
      acc = fma(a, in[base + i * stride], acc);
    }
    out[o] = accumulate ? out[o] + acc : acc;
  }
}

static void cell_coords(const oracle_t *o, int64_t c, int64_t cc[3]) {
  int64_t r = c;
  for (int d = 0; d < o->dim; ++d) {
    cc[d] = r % o->nc[d];
    r /= o->nc[d];
  }
}

static int64_t cell_node(const oracle_t *o, const int64_t cc[3], int i) {
  int64_t node = 0, stride = 1;
  for (int d = 0; d < o->dim; ++d) {
    int li = i % o->n1;
    i /= o->n1;
    node += (cc[d] * o->p + li) * stride;
    stride *= o->nn[d];
  }
  return node;
}

/* face f = 2*d + side; returns boundary type to apply or 0 */
static int face_type(const oracle_t *o, const int64_t cc[3], int f) {
  int d = f / 2, side = f % 2;
  int64_t nb = cc[d] + (side ? 1 : -1);
  unsigned int bid;
  if (nb < 0 || nb >= o->nc[d]) {
    bid = (unsigned int)f; /* colorized ids 2*d + side */
  } else {
    int64_t cn[3] = {cc[0], cc[1], cc[2]};
    cn[d] = nb;
    int64_t c2 = 0, stride = 1;
    for (int k = 0; k < o->dim; ++k) {
      c2 += cn[k] * stride;
      stride *= o->nc[k];
    }
    if (o->active[c2]) return 0; /* interior face of the active domain */
    bid = 0xffffffffu;             /* numbers::internal_face_boundary_id */
  }
  /* Boundary::get_boundary_type (source/Boundary.hh:120-126) */
  int t = (bid < (unsigned int)(o->n_bids - 1)) ? o->btype[bid]
                                                : o->btype[o->n_bids - 1];
  return t & (ADM_BC_RADIATIVE | ADM_BC_CONVECTIVE) ? t : 0;
}

static void apply_cell(oracle_t *o, int64_t c, const double *src,
                       double *dst) {
  const int dim = o->dim, n1 = o->n1, nq = o->nq;
  const adm_material_desc *m = &o->mat;
  const int mid = o->material[c];
  int64_t cc[3] = {0, 0, 0};
  cell_coords(o, c, cc);
  int64_t nodes[MAXQ];
  double u[MAXQ], t1[MAXQ], t2[MAXQ], v[MAXQ], g[3][MAXQ], w[3][MAXQ];
  double val[MAXQ];
  for (int i = 0; i < nq; ++i) {
    nodes[i] = cell_node(o, cc, i);
    u[i] = src[nodes[i]];
  }
  /* values at quadrature points: S along x, then y, then z */
  if (dim == 3) {
    sweep(dim, n1, 0, o->S, 0, u, t1, 0);
    sweep(dim, n1, 1, o->S, 0, t1, t2, 0);
    sweep(dim, n1, 2, o->S, 0, t2, v, 0);
  } else {
    sweep(dim, n1, 0, o->S, 0, u, t1, 0);
    sweep(dim, n1, 1, o->S, 0, t1, v, 0);
  }
  /* reference gradients by collocation derivative */
  for (int d = 0; d < dim; ++d) sweep(dim, n1, d, o->Dc, 0, v, g[d], 0);

  const double cs = o->dcos[c], sn = o->dsin[c];
  for (int q = 0; q < nq; ++q) {
    double T = v[q];
    double tp[ADM_MAX_POLY];
    tp[0] = 1.0;
    for (int i = 1; i <= m->p_order; ++i) tp[i] = tp[i - 1] * T;
    double r[3];
    state_ratios(m, mid, T, &o->liquid[c * nq + q], &o->powder[c * nq + q], r);
    double irc = inv_rho_cp(m, mid, r, T, tp);
    double grad[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; ++d) grad[d] = g[d][q] * o->inv_h[d];
    double flux[3];
    if (dim == 2) {
      flux[0] = grad[0] * property_tp(m, ADM_THERMAL_CONDUCTIVITY_X, mid, r, T, tp);
      flux[1] = grad[1] * property_tp(m, ADM_THERMAL_CONDUCTIVITY_Z, mid, r, T, tp);
    } else {
      double kx = property_tp(m, ADM_THERMAL_CONDUCTIVITY_X, mid, r, T, tp);
      double ky = property_tp(m, ADM_THERMAL_CONDUCTIVITY_Y, mid, r, T, tp);
      flux[0] = (kx * cs * cs + ky * sn * sn) * grad[0] +
                ((kx - ky) * sn * cs) * grad[1];
      flux[1] = ((kx - ky) * sn * cs) * grad[0] +
                (kx * sn * sn + ky * cs * cs) * grad[1];
      flux[2] = grad[2] *
                property_tp(m, ADM_THERMAL_CONDUCTIVITY_Z, mid, r, T, tp);
    }
    /* submit_gradient(-inv_rho_cp * K grad T): test-function gradient maps
     * with the inverse Jacobian, weighted by JxW */
    for (int d = 0; d < dim; ++d)
      w[d][q] = ((-irc) * flux[d]) * o->jxw[q] * o->inv_h[d];
    if (o->sources_on) {
      double pt[3];
      int64_t r2 = q;
      for (int d = 0; d < dim; ++d) {
        int qd = r2 % n1;
        r2 /= n1;
        pt[d] = o->org[d] + (cc[d] + o->xq[qd]) * o->h[d];
      }
      double qs = 0.0;
      for (int b = 0; b < o->n_beams; ++b)
        qs += oracle_beam_value(&o->bs[b], dim, pt);
      qs *= irc;
      val[q] = qs * o->jxw[q];
    }
  }
  /* integrate: collocation derivative transposed, then S^T sweeps */
  sweep(dim, n1, 0, o->Dc, 1, w[0], t1, 0);
  for (int d = 1; d < dim; ++d) sweep(dim, n1, d, o->Dc, 1, w[d], t1, 1);
  if (o->sources_on)
    for (int q = 0; q < nq; ++q) t1[q] += val[q];
  if (dim == 3) {
    sweep(dim, n1, 2, o->S, 1, t1, t2, 0);
    sweep(dim, n1, 1, o->S, 1, t2, t1, 0);
    sweep(dim, n1, 0, o->S, 1, t1, u, 0);
  } else {
    sweep(dim, n1, 1, o->S, 1, t1, t2, 0);
    sweep(dim, n1, 0, o->S, 1, t2, u, 0);
  }

  /* faces of the active domain with non-adiabatic type */
  if (!o->adiabatic_only) {
    for (int f = 0; f < 2 * dim; ++f) {
      int bt = face_type(o, cc, f);
      if (!bt) continue;
      int nd = f / 2, side = f % 2;
      int li_n = side ? o->p : 0;
      /* face node slice (tangential dims in increasing order) */
      int tdims[2], nt = 0;
      for (int d = 0; d < dim; ++d)
        if (d != nd) tdims[nt++] = d;
      const int nf = o->nfq;
      double fu[MAXFQ], ft[MAXFQ], fv[MAXFQ], fval[MAXFQ];
      int fidx[MAXFQ];
      for (int j = 0; j < nf; ++j) {
        int lj[3];
        int rj = j;
        for (int k = 0; k < nt; ++k) {
          lj[tdims[k]] = rj % n1;
          rj /= n1;
        }
        lj[nd] = li_n;
        int idx = 0, st = 1;
        for (int d = 0; d < dim; ++d) {
          idx += lj[d] * st;
          st *= n1;
        }
        fidx[j] = idx;
        fu[j] = src[nodes[idx]];
      }
      if (nt == 2) {
        sweep(2, n1, 0, o->S, 0, fu, ft, 0);
        sweep(2, n1, 1, o->S, 0, ft, fv, 0);
      } else {
        sweep(1, n1, 0, o->S, 0, fu, fv, 0);
      }
      double *fpow = &o->face_powder[(c * 2 * dim + f) * nf];
      for (int q = 0; q < nf; ++q) {
        double T = fv[q];
        double tp[ADM_MAX_POLY];
        tp[0] = 1.0;
        for (int i = 1; i <= m->p_order; ++i) tp[i] = tp[i - 1] * T;
        double r[3];
        state_ratios(m, mid, T, NULL, &fpow[q], r);
        double irc = inv_rho_cp(m, mid, r, T, tp);
        double conv_T = 0.0, conv_h = 0.0, rad_T = 0.0, rad_h = 0.0;
        if (bt & ADM_BC_CONVECTIVE) {
          conv_T = m->prop[mid][ADM_CONVECTION_TEMPERATURE_INFTY];
          conv_h = property_tp(m, ADM_CONVECTION_HEAT_TRANSFER_COEF, mid, r, T,
                               tp);
        }
        if (bt & ADM_BC_RADIATIVE) {
          rad_T = m->prop[mid][ADM_RADIATION_TEMPERATURE_INFTY];
          rad_h = property_tp(m, ADM_EMISSIVITY, mid, r, T, tp) *
                  STEFAN_BOLTZMANN * (T + rad_T) * (T * T + rad_T * rad_T);
        }
        double bv = -irc * (conv_h * (T - conv_T) + rad_h * (T - rad_T));
        fval[q] = bv * o->fjxw[nd][q];
      }
      if (nt == 2) {
        sweep(2, n1, 1, o->S, 1, fval, ft, 0);
        sweep(2, n1, 0, o->S, 1, ft, fv, 0);
      } else {
        sweep(1, n1, 0, o->S, 1, fval, fv, 0);
      }
      for (int j = 0; j < nf; ++j) u[fidx[j]] += fv[j];
    }
  }
  for (int i = 0; i < nq; ++i) dst[nodes[i]] += u[i];
}

/* ThermalOperator::vmult / vmult_add (source/ThermalOperator.templates.hh:
 * 176-225). Cells are visited colour by colour (parity of the cell
 * coordinates), so cells of one colour share no node. */
void oracle_vmult(oracle_t *o, const double *src, double *dst, int add) {
  if (!add) memset(dst, 0, sizeof(double) * o->n_dofs);
  const int ncol = 1 << o->dim;
  for (int color = 0; color < ncol; ++color) {
    int64_t n = o->n_cells;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) {
      if (!o->active[c]) continue;
      int64_t cc[3] = {0, 0, 0};
      cell_coords(o, c, cc);
      int col = 0;
      for (int d = 0; d < o->dim; ++d) col |= (int)(cc[d] & 1) << d;
      if (col != color) continue;
      apply_cell(o, c, src, dst);
    }
  }
}

/* evaluate_thermal_physics_impl, host (source/ThermalPhysics.templates.hh:
 * 43-124): set_time, value = 0, vmult_add, scale by the inverse mass. */
void oracle_evaluate(oracle_t *o, double t, const double *y, double *out) {
  oracle_set_time(o, t);
  memset(out, 0, sizeof(double) * o->n_dofs);
  oracle_vmult(o, y, out, 1);
  for (int64_t i = 0; i < o->n_dofs; ++i) out[i] *= o->inv_mass[i];
}

/* evolve_one_time_step (source/ThermalPhysics.templates.hh:854-888): forward
 * Euler via sadd(1, dt, f); RK3/RK4 via deal.II's ExplicitRungeKutta
 * (compute_stages: Y = y + sum_j dt*a_ij f_j; then y += dt*b_i f_i). */
double oracle_evolve(oracle_t *o, int method, double t, double dt, int n_steps,
                     double *T) {
  const int64_t n = o->n_dofs;
  static const double a3[3][3] = {{0, 0, 0}, {0.5, 0, 0}, {-1.0, 2.0, 0}};
  static const double b3[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
  static const double c3[3] = {0.0, 0.5, 1.0};
  static const double a4[4][4] = {
      {0, 0, 0, 0}, {0.5, 0, 0, 0}, {0.0, 0.5, 0, 0}, {0.0, 0.0, 1.0, 0}};
  static const double b4[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  static const double c4[4] = {0.0, 0.5, 0.5, 1.0};
  double *f = (double *)malloc(sizeof(double) * n * 4);
  double *Y = (double *)malloc(sizeof(double) * n);
  for (int s = 0; s < n_steps; ++s) {
    if (method == ADM_FORWARD_EULER) {
      oracle_evaluate(o, t, T, f);
      for (int64_t i = 0; i < n; ++i) T[i] = 1. * T[i] + dt * f[i];
    } else {
      int ns = method == ADM_RK_THIRD_ORDER ? 3 : 4;
      for (int i = 0; i < ns; ++i) {
        memcpy(Y, T, sizeof(double) * n);
        for (int j = 0; j < i; ++j) {
          double a = dt * (ns == 3 ? a3[i][j] : a4[i][j]);
          double *fj = f + (int64_t)j * n;
          for (int64_t k = 0; k < n; ++k) Y[k] = 1. * Y[k] + a * fj[k];
        }
        double ci = ns == 3 ? c3[i] : c4[i];
        oracle_evaluate(o, t + ci * dt, Y, f + (int64_t)i * n);
      }
      for (int i = 0; i < ns; ++i) {
        double b = dt * (ns == 3 ? b3[i] : b4[i]);
        double *fi = f + (int64_t)i * n;
        for (int64_t k = 0; k < n; ++k) T[k] = 1. * T[k] + b * fi[k];
      }
    }
    t = t + dt;
  }
  free(f);
  free(Y);
  return t;
}

/* Nodes of active cells (1) and the rest (0). */
static uint8_t *node_activity(const oracle_t *o) {
  uint8_t *na = (uint8_t *)calloc(o->n_dofs, 1);
  for (int64_t c = 0; c < o->n_cells; ++c) {
    if (!o->active[c]) continue;
    int64_t cc[3] = {0, 0, 0}, r = c;
    for (int d = 0; d < o->dim; ++d) {
      cc[d] = r % o->nc[d];
      r /= o->nc[d];
    }
    for (int i = 0; i < o->nq; ++i) {
      int64_t node = 0, stride = 1;
      int ri = i;
      for (int d = 0; d < o->dim; ++d) {
        int li = ri % o->n1;
        ri /= o->n1;
        node += (cc[d] * o->p + li) * stride;
        stride *= o->nn[d];
      }
      na[node] = 1;
    }
  }
  return na;
}

/* get_elements_to_activate (source/material_deposition.cc:376-435): for every
 * inactive cell, the last box (in input order) whose closed box intersects
 * the cell's bounding box, or -1. boxes[b] = {min x, y, z, max x, y, z}. */
void oracle_cells_to_activate(const oracle_t *o, int n_boxes,
                              const double *boxes, int32_t *box_of_cell) {
  for (int64_t c = 0; c < o->n_cells; ++c) {
    box_of_cell[c] = -1;
    if (o->active[c]) continue;
    int64_t cc[3] = {0, 0, 0}, r = c;
    for (int d = 0; d < o->dim; ++d) {
      cc[d] = r % o->nc[d];
      r /= o->nc[d];
    }
    for (int b = 0; b < n_boxes; ++b) {
      int hit = 1;
      for (int d = 0; d < o->dim; ++d) {
        const int ax = o->axis[d];
        const double lo = o->org[d] + cc[d] * o->h[d];
        const double hi = o->org[d] + (cc[d] + 1) * o->h[d];
        if (!(lo <= boxes[6 * b + 3 + ax] && boxes[6 * b + ax] <= hi)) hit = 0;
      }
      if (hit) box_of_cell[c] = b;
    }
  }
}

/* add_material_start + add_material_end (source/ThermalPhysics.templates.hh:
 * 595-825): the phase state goes to cell means and back to every quadrature
 * point (set/get_state_to/from_material_properties), the selected cells
 * become active with their box's deposition direction, nodes that no
 * previously active cell held take new_temperature, the others keep their
 * value, and the lumped mass is recomputed. */
void oracle_add_material(oracle_t *o, const int32_t *box_of_cell,
                         const double *box_cos, const double *box_sin,
                         double new_temperature, double *T) {
  const int ns = o->mat.n_states;
  double *cs = (double *)malloc(sizeof(double) * o->n_cells * ns);
  oracle_get_cell_state(o, cs);
  uint8_t *before = node_activity(o);
  for (int64_t c = 0; c < o->n_cells; ++c) {
    const int b = box_of_cell[c];
    if (b < 0 || o->active[c]) continue;
    o->active[c] = 1;
    o->dcos[c] = box_cos[b];
    o->dsin[c] = box_sin[b];
  }
  uint8_t *after = node_activity(o);
  for (int64_t i = 0; i < o->n_dofs; ++i)
    if (after[i] && !before[i]) T[i] = new_temperature;
  oracle_set_cell_state(o, cs);
  compute_inverse_mass(o);
  free(after);
  free(before);
  free(cs);
}
