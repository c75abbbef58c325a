// C-ABI of the MI355X matrix-free heat path (declared in include/adm.h).
//
// The context replaces adamantine's ThermalOperatorDevice + the device
// instantiation of ThermalPhysics for a structured box mesh; see adm.h for the
// reference interface each entry point stands in for.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/adm.h"
#include "adm_context.hpp"
#include "adm_kernels.hpp"
#include "adm_kernels_2d.hpp"
#include "adm_launch.hpp"
#include "adm_plane_kernel.hpp"
#include "adm_step_kernels.hpp"

namespace adm {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define ADM_HIP(call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      throw AdmError(ADM_ERR_HIP, std::string(#call) + ": " +             \
                                      hipGetErrorString(e_));             \
  } while (0)

// ---------------------------------------------------------------------------
// 1-D quadrature (same construction as QGauss / QGaussLobatto: Newton on
// Legendre polynomials in extended precision)
// ---------------------------------------------------------------------------

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

void gauss_points(int n, double *x, double *w) {
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
    x[n - 1 - i] = (double)(0.5L + 0.5L * z);
    w[n - 1 - i] = (double)(0.5L * ww);
  }
}

void gauss_lobatto_points(int n, double *x, double *w) {
  int m = n - 1;
  x[0] = 0.0;
  x[n - 1] = 1.0;
  w[0] = w[n - 1] = (double)(0.5L * 2.0L / (m * (m + 1.0L)));
  for (int i = 1; i < n - 1; ++i) {
    long double z = -cosl(3.14159265358979323846264338327950288L * i / m);
    for (int it = 0; it < 200; ++it) {
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

// ---------------------------------------------------------------------------
// scan path (source/ScanPath.cc:197-275) and HeatSource::update_time
// ---------------------------------------------------------------------------

static int scan_segment(const BeamData &b, double t, int &old_segment,
                        double sp[3], double *st) {
  int cur = old_segment;
  while (t > b.segments[cur].end_time) ++cur;
  if (cur > 0) {
    *st = b.segments[cur - 1].end_time;
    for (int d = 0; d < 3; ++d) sp[d] = b.segments[cur - 1].end_point[d];
  } else {
    *st = 0.0;
    for (int d = 0; d < 3; ++d) sp[d] = b.segments[cur].end_point[d];
  }
  old_segment = cur;
  return cur;
}

// Unit quaternion (r, i, j, k) of a five-axis path (source/Quaternion.cc).
struct Quat {
  double r, i, j, k;
  static Quat of(const double (&q)[4]) { return {q[0], q[1], q[2], q[3]}; }
  bool operator==(const Quat &o) const {
    return r == o.r && i == o.i && j == o.j && k == o.k;
  }
  // Hamilton product, this * o (source/Quaternion.cc:62-83)
  Quat operator*(const Quat &o) const {
    return {r * o.r - i * o.i - j * o.j - k * o.k,
            r * o.i + i * o.r + j * o.k - k * o.j,
            r * o.j - i * o.k + j * o.r + k * o.i,
            r * o.k + i * o.j - j * o.i + k * o.r};
  }
  Quat conj() const { return {r, -i, -j, -k}; }
  double dot(const Quat &o) const {
    double d = 0.0;
    d += r * o.r;
    d += i * o.i;
    d += j * o.j;
    d += k * o.k;
    return d;
  }
  // q^e through the axis-angle form (source/Quaternion.cc:101-118)
  Quat pow(double e) const {
    const double n = std::sqrt(i * i + j * j + k * k);
    const double ux = i / n, uy = j / n, uz = k / n;
    const double th = std::atan2(n, r);
    const double s_ = std::sin(e * th);
    return {std::cos(e * th), ux * s_, uy * s_, uz * s_};
  }
  // row-major rotation matrix (source/Quaternion.cc:120-134)
  void matrix(double (&R)[9]) const {
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
};

static bool five_axis_path(const BeamData &b) {
  const double *q = b.segments[0].end_rotation;
  return q[0] != 0.0 || q[1] != 0.0 || q[2] != 0.0 || q[3] != 0.0;
}

// ScanPath::get_quaternion (source/ScanPath.cc:290-341): the slerp from the
// segment's start rotation to its end rotation on the shorter arc; `cur` and
// `st` come from the segment search at the same time.
static Quat path_rotation(const BeamData &b, double t, int cur, double st) {
  const adm_segment &last = b.segments.back();
  if (t > last.end_time) return Quat::of(last.end_rotation);
  const Quat q1 = Quat::of(b.segments[cur].end_rotation);
  const Quat q0 = Quat::of(b.segments[cur > 0 ? cur - 1 : 0].end_rotation);
  if (q0 == q1) return q0;
  const double duration = b.segments[cur].end_time - st;
  if (duration < 1e-12) return q1;
  Quat q = q1;
  if (q0.dot(q) < 0.0) q = {-q.r, -q.i, -q.j, -q.k};
  return (q * q0.conj()).pow((t - st) / duration) * q0;
}

void beam_update(const BeamData &b, double t, int &old_segment,
                 adm_beam_state &s, double *pmod_out, double *den_out) {
  std::memset(&s, 0, sizeof(s));
  if (pmod_out) *pmod_out = 0.0;
  if (den_out) *den_out = 1.0;
  s.type = b.desc.type;
  if (b.desc.type == ADM_SOURCE_CUBE) {
    s.on = (t > b.desc.start_time) && (t < b.desc.end_time);
    s.cube_value = b.desc.value;
    for (int d = 0; d < 3; ++d) {
      s.cube_min[d] = b.desc.min_point[d];
      s.cube_max[d] = b.desc.max_point[d];
    }
    return;
  }
  const adm_segment &last = b.segments.back();
  double pmod = 0.0;
  if (!(t > last.end_time)) {
    double sp[3], st;
    int cur = scan_segment(b, t, old_segment, sp, &st);
    pmod = b.segments[cur].power_modifier;
    // deal.II's Tensor / scalar multiplies by the reciprocal
    const double inv = 1.0 / (b.segments[cur].end_time - st);
    for (int d = 0; d < 3; ++d)
      s.center[d] =
          sp[d] + ((b.segments[cur].end_point[d] - sp[d]) * inv) * (t - st);
    if (five_axis_path(b)) {
      path_rotation(b, t, cur, st).matrix(s.rotation);
      s.rotated = 1;
    }
  } else {
    for (int d = 0; d < 3; ++d) s.center[d] = -DBL_MAX;
    if (five_axis_path(b)) {
      Quat::of(last.end_rotation).matrix(s.rotation);
      s.rotated = 1;
    }
  }
  s.on = pmod > 0.0;
  const double r2 = std::pow(b.desc.radius, 2);
  const double dd = b.desc.depth;
  s.depth = dd;
  s.radius_squared = r2;
  if (pmod_out) *pmod_out = pmod;
  if (b.desc.type == ADM_SOURCE_GOLDAK) {
    const double pi_over_3_to_1p5 = std::pow(M_PI / 3.0, 1.5);
    s.alpha = 2.0 * b.desc.absorption_efficiency * b.desc.max_power * pmod /
              (r2 * dd * pi_over_3_to_1p5);
    s.inv_radius_squared = 1.0 / r2;
    s.inv_depth_squared = 1.0 / (dd * dd);
    if (den_out) *den_out = r2 * dd * pi_over_3_to_1p5;
  } else {
    const double log_01 = std::log(0.1);
    s.alpha = -b.desc.absorption_efficiency * b.desc.max_power * pmod *
              log_01 / (M_PI * r2 * dd);
    s.inv_depth = 1.0 / dd;
    s.inv_radius_squared = 1.0 / r2;
    if (den_out) *den_out = M_PI * r2 * dd;
  }
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

static bool table_sorted(const double (*t)[2]) {
  for (int i = 1; i < ADM_TABLE_SIZE; ++i)
    if (t[i][0] < t[i - 1][0]) return false;
  return true;
}

// A table has the regular form when its temperatures increase strictly up to
// some entry and the remaining entries repeat that last pair (the padding of
// MaterialProperty::fill_properties, source/MaterialProperty.templates.hh:
// 975-1004).
static bool table_regular(const double (*t)[2]) {
  int n = 1;
  while (n < ADM_TABLE_SIZE && t[n][0] > t[n - 1][0]) ++n;
  for (int k = n; k < ADM_TABLE_SIZE; ++k)
    if (t[k][0] != t[n - 1][0] || t[k][1] != t[n - 1][1]) return false;
  return true;
}

// Merged piecewise-linear form (see MergedTable): in interval j, i.e.
// [B_{j-1}, B_j), each table's index in compute_property_from_table is fixed,
// so the table is (value at B_{j-1}) + (T - B_{j-1}) * slope.
void Context::build_merged_tables() {
  MatConst &mc = kc.mat;
  mc.merged_ok = mc.use_table ? 1 : 0;
  if (!mc.use_table) return;
  for (int mid = 0; mid < mc.n_materials; ++mid) {
    std::vector<double> B;
    for (int pr = 0; pr < ADM_N_STATE_PROPS; ++pr)
      for (int s = 0; s < mc.n_states; ++s) {
        const double(*t)[2] = mc.table[mid][pr][s];
        if (!table_regular(t)) mc.merged_ok = 0;
        for (int k = 0; k < ADM_TABLE_SIZE; ++k) B.push_back(t[k][0]);
      }
    std::sort(B.begin(), B.end());
    B.erase(std::unique(B.begin(), B.end()), B.end());
    if ((int)B.size() > kMergedMax - 1) {
      mc.merged_ok = 0;
      continue;
    }
    MergedTable &mt = mc.merged[mid];
    const int nB = (int)B.size();
    for (int k = 0; k < kMergedMax; ++k)
      mt.B[k] = k < nB ? B[k] : HUGE_VAL;
    for (int pr = 0; pr < ADM_N_STATE_PROPS; ++pr)
      for (int j = 0; j < kMergedMax; ++j)
        for (int s = 0; s < ADM_MAX_STATES; ++s) {
          double A = 0.0, S = 0.0;
          if (s < mc.n_states && j <= nB) {
            const double(*t)[2] = mc.table[mid][pr][s];
            if (j == 0) {
              A = t[0][1];
            } else {
              const double Tl = B[j - 1];
              int i = 0;
              for (int k = 0; k < ADM_TABLE_SIZE; ++k) i += t[k][0] <= Tl;
              if (i == 0) {
                A = t[0][1];
              } else if (i >= ADM_TABLE_SIZE - 1) {
                A = t[ADM_TABLE_SIZE - 1][1];
              } else {
                S = (t[i][1] - t[i - 1][1]) / (t[i][0] - t[i - 1][0]);
                A = (double)((long double)t[i - 1][1] +
                             ((long double)Tl - (long double)t[i - 1][0]) *
                                 (long double)S);
              }
            }
          }
          mt.coef[pr][j][s][0] = A;
          mt.coef[pr][j][s][1] = S;
        }
    // bins of the one-lookup interval search (MergedBins)
    MergedBins &mb = mc.bins[mid];
    std::memset(&mb, 0, sizeof(mb));
    const double lo = B[0], hi = B[nB - 1];
    const double w = (hi - lo) / kMergedBins;
    mb.lo = lo;
    mb.inv_w = w > 0.0 ? 1.0 / w : 0.0;
    const double margin =
        1e-9 * (hi - lo) + 1e-12 * std::max(std::fabs(lo), std::fabs(hi));
    int maxc = 1;
    for (int b = 0; b < kMergedBins; ++b) {
      const double rlo = b == 0 ? -HUGE_VAL : lo + b * w - margin;
      const double rhi = b == kMergedBins - 1 || w == 0.0
                             ? HUGE_VAL
                             : lo + (b + 1) * w + margin;
      int k0 = 0, cnt = 0;
      for (int k = 0; k < nB; ++k) {
        if (B[k] < rlo)
          ++k0;
        else if (B[k] <= rhi)
          ++cnt;
      }
      mb.k0[b] = (uint8_t)k0;
      maxc = std::max(maxc, cnt);
    }
    mb.maxc = maxc;
  }
}

int Context::property_kind() const {
  if (!use_table) return 0;
  return kc.mat.merged_ok ? 2 : 1;
}

// Scratch of adm_vmult_noalloc: the plane kernel's patch buffer in a
// static device array, so that the call allocates nothing. Each device the
// library runs on holds its own copy for the life of the process; host code
// uses the address hipGetSymbolAddress gives on that device, never the
// array's name.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

// The array's address on device dev (dev current), resolved once per device.
static double *scratch_on_device(int dev) {
  static std::mutex mu;
  static double *ptr[64] = {};
  if (dev < 0 || dev >= 64)
    throw AdmError(ADM_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    double *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw AdmError(ADM_ERR_HIP, std::string("scratch address: ") +
                                      hipGetErrorString(e));
    ptr[dev] = p;
  }
  return ptr[dev];
}

Context::Context(const adm_problem_desc &d) {
  if (const char *e = std::getenv("ADM_CELL_ATOMICS"))
    use_coloring = !(e[0] == '1');
  {
    const char *e = std::getenv("ADM_CELL_KERNEL");
    use_plane = d.mesh.fe_degree <= 2 && !(e && std::string(e) == "line");
  }
  if (const char *e = std::getenv("ADM_XCD_ORDER")) xcd_order = e[0] == '1';
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                              dev) == hipSuccess &&
        cus > 0)
      n_cu = cus;
  }
  const adm_mesh_desc &m = d.mesh;
  if (m.dim != 2 && m.dim != 3)
    throw AdmError(ADM_ERR_INVALID, "dim must be 2 or 3");
  dim = m.dim;
  // the plane kernel is three-dimensional
  if (dim == 2) use_plane = false;
  if (m.fe_degree < 1 || m.fe_degree > 5)
    throw AdmError(ADM_ERR_UNSUPPORTED,
                   "fe_degree must be 1 to 5 on the device path");
  if (m.quadrature != 0 && m.quadrature != 1)
    throw AdmError(ADM_ERR_INVALID, "quadrature must be 0 (Gauss) or 1 "
                                    "(Gauss-Lobatto)");
  // dim = 2 uses axes 0 (length) and 2 (height) of the descriptor
  for (int k = 0; k < 3; ++k)
    if ((dim == 3 || k != 1) && (m.n_cells[k] < 1 || !(m.extent[k] > 0.0)))
      throw AdmError(ADM_ERR_INVALID, "invalid mesh size");
  const adm_material_desc &mat = d.material;
  if (mat.n_materials < 1 || mat.n_materials > ADM_MAX_MATERIALS)
    throw AdmError(ADM_ERR_INVALID, "n_materials must be in 1..4");
  if (mat.n_states < 1 || mat.n_states > 3)
    throw AdmError(ADM_ERR_INVALID, "n_states must be 1, 2 or 3");
  if (mat.p_order < 0 || mat.p_order > 4)
    throw AdmError(ADM_ERR_INVALID, "p_order must be in 0..4");
  if (d.n_boundary_ids < 1 || d.n_boundary_ids > ADM_MAX_BOUNDARY_IDS)
    throw AdmError(ADM_ERR_INVALID, "invalid number of boundary ids");
  if (d.n_beams < 0 || d.n_beams > ADM_MAX_BEAMS)
    throw AdmError(ADM_ERR_INVALID, "at most 4 heat sources");

  p = m.fe_degree;
  n1 = p + 1;
  nq = dim == 3 ? n1 * n1 * n1 : n1 * n1;
  nfq = dim == 3 ? n1 * n1 : n1;
  n_cells = 1;
  n_dofs = 1;
  for (int k = 0; k < 3; ++k) {
    const bool flat = dim == 2 && k == 1;
    nc[k] = flat ? 1 : m.n_cells[k];
    nn[k] = flat ? 1 : p * nc[k] + 1;
    n_cells *= nc[k];
    n_dofs *= nn[k];
  }
  // cell and row-segment counts are 32-bit in the kernels
  if (n_cells > (int64_t)INT32_MAX)
    throw AdmError(ADM_ERR_UNSUPPORTED, "more than 2^31 - 1 cells");
  // the plane kernel's gather pass indexes nodes with 32-bit integers
  if (n_dofs >= (int64_t)INT32_MAX) use_plane = false;
  n_states = mat.n_states;
  use_table = mat.use_table != 0;

  // ---- constants ----------------------------------------------------------
  std::memset(&kc, 0, sizeof(kc));
  double xs[kMaxN], wq[kMaxN];
  gauss_lobatto_points(n1, xs, wgl);
  if (m.quadrature == 1)
    gauss_lobatto_points(n1, kc.xq, wq);
  else
    gauss_points(n1, kc.xq, wq);
  for (int q = 0; q < n1; ++q)
    for (int i = 0; i < n1; ++i) {
      kc.S[q][i] = lagrange(n1, xs, i, kc.xq[q]);
      kc.Dc[q][i] = lagrange_deriv(n1, kc.xq, i, kc.xq[q]);
    }
  for (int q = 0; q < n1; ++q)
    for (int i = 0; i < n1; ++i) {
      long double acc = 0.0L;
      for (int k = 0; k < n1; ++k)
        acc += (long double)kc.Dc[q][k] * (long double)kc.S[k][i];
      kc.DcS[q][i] = (double)acc;
    }
  // the plane kernel's Q2 form relies on these equalities of the 1-D
  // tables (they hold for the Gauss and Gauss-Lobatto rules); without them
  // Q2 goes through the line kernel
  if (n1 == 3) {
    const auto &S = kc.S, &D = kc.Dc;
    const bool sym = S[1][0] == 0.0 && S[1][1] == 1.0 && S[1][2] == 0.0 &&
                     S[2][0] == S[0][2] && S[2][2] == S[0][0] &&
                     D[1][1] == 0.0 && D[1][2] == -D[1][0] &&
                     D[2][0] == -D[0][2] && D[2][1] == -D[0][1] &&
                     D[2][2] == -D[0][0];
    if (!sym) {
      use_plane = false;
      plane_tables_ok = false;
    }
  }
  det = 1.0;
  for (int k = 0; k < 3; ++k) {
    if (dim == 2 && k == 1) {
      kc.h[k] = kc.inv_h[k] = 1.0;
      kc.origin[k] = 0.0;
      continue;
    }
    kc.h[k] = m.extent[k] / nc[k];
    kc.inv_h[k] = 1.0 / kc.h[k];
    kc.origin[k] = m.origin[k];
    det *= kc.h[k];
  }
  for (int q = 0; q < n1; ++q) kc.wq[q] = wq[q];
  kc.det = det;
  for (int q = 0; q < nq; ++q) {
    double w = 1.0;
    int r = q;
    for (int k = 0; k < dim; ++k) {
      w *= wq[r % n1];
      r /= n1;
    }
    kc.jxw[q] = det * w;
  }
  for (int nd = 0; nd < 3; ++nd) {
    if (!has_face(2 * nd)) continue;
    double fdet = 1.0;
    for (int k = 0; k < 3; ++k)
      if (k != nd && (dim == 3 || k != 1)) fdet *= kc.h[k];
    for (int q = 0; q < nfq; ++q) {
      double w = 1.0;
      int r = q;
      for (int k = 0; k < dim - 1; ++k) {
        w *= wq[r % n1];
        r /= n1;
      }
      kc.fjxw[nd][q] = fdet * w;
    }
  }
  kc.n_boundary_ids = d.n_boundary_ids;
  int n_adiabatic = 0;
  for (int i = 0; i < d.n_boundary_ids; ++i) {
    kc.boundary_type[i] = d.boundary_type[i];
    if (d.boundary_type[i] & ADM_BC_ADIABATIC) ++n_adiabatic;
  }
  kc.adiabatic_only = n_adiabatic == d.n_boundary_ids;
  kc.log_01 = std::log(0.1);

  MatConst &mc = kc.mat;
  mc.n_materials = mat.n_materials;
  mc.n_states = mat.n_states;
  mc.use_table = mat.use_table;
  mc.p_order = mat.p_order;
  mc.single_layout = mat.single_material_layout;
  std::memcpy(mc.table, mat.table, sizeof(mc.table));
  std::memcpy(mc.poly, mat.poly, sizeof(mc.poly));
  std::memcpy(mc.prop, mat.prop, sizeof(mc.prop));
  for (int mid = 0; mid < ADM_MAX_MATERIALS; ++mid) {
    mc.latent_factor[mid] = mat.prop[mid][ADM_LATENT_HEAT] /
                            (mat.prop[mid][ADM_LIQUIDUS] -
                             mat.prop[mid][ADM_SOLIDUS]);
    for (int pr = 0; pr < ADM_N_STATE_PROPS; ++pr)
      for (int s = 0; s < ADM_MAX_STATES; ++s) {
        const double(*t)[2] = mat.table[mid][pr][s];
        if (mat.use_table && mid < mat.n_materials && s < mat.n_states &&
            !table_sorted(t))
          throw AdmError(ADM_ERR_INVALID,
                         "material tables must have increasing temperatures");
        bool constant = true;
        for (int i = 1; i < ADM_TABLE_SIZE; ++i)
          if (t[i][1] != t[0][1]) constant = false;
        mc.kind[mid][pr][s] = constant ? kTableConstant : kTableGeneral;
      }
    const int kx = ADM_THERMAL_CONDUCTIVITY_X;
    bool iso = true;
    for (int other = kx + 1; other <= ADM_THERMAL_CONDUCTIVITY_Z; ++other) {
      if (mat.use_table) {
        if (std::memcmp(mat.table[mid][kx], mat.table[mid][other],
                        sizeof(mat.table[mid][kx])) != 0)
          iso = false;
      } else {
        if (std::memcmp(mat.poly[mid][kx], mat.poly[mid][other],
                        sizeof(mat.poly[mid][kx])) != 0)
          iso = false;
      }
    }
    mc.isotropic[mid] = iso ? 1 : 0;
  }
  build_merged_tables();

  // ---- per-cell data --------------------------------------------------------
  active.assign(n_cells, 1);
  material.assign(n_cells, 0);
  dep_cos.assign(n_cells, 1.0);
  dep_sin.assign(n_cells, 0.0);
  for (int64_t c = 0; c < n_cells; ++c) {
    if (d.cell_active) active[c] = d.cell_active[c] ? 1 : 0;
    if (d.cell_material) {
      if (d.cell_material[c] >= mat.n_materials)
        throw AdmError(ADM_ERR_INVALID, "cell material id out of range");
      material[c] = d.cell_material[c];
    }
    if (d.deposition_cos) dep_cos[c] = d.deposition_cos[c];
    if (d.deposition_sin) dep_sin[c] = d.deposition_sin[c];
  }
  cell_state.assign((size_t)n_cells * n_states, 0.0);
  for (int64_t c = 0; c < n_cells; ++c) {
    if (d.cell_state) {
      for (int s = 0; s < n_states; ++s)
        cell_state[c * n_states + s] = d.cell_state[c * n_states + s];
    } else {
      cell_state[c * n_states] = 1.0;
    }
  }

  for (int b = 0; b < d.n_beams; ++b) {
    BeamData bd;
    bd.desc = d.beams[b];
    if (bd.desc.type != ADM_SOURCE_CUBE) {
      if (bd.desc.n_segments < 1 || !bd.desc.segments)
        throw AdmError(ADM_ERR_INVALID, "beam without scan path segments");
      bd.segments.assign(bd.desc.segments,
                         bd.desc.segments + bd.desc.n_segments);
      if (five_axis_path(bd))
        for (const adm_segment &sg : bd.segments) {
          const double *q = sg.end_rotation;
          // Quaternion::reinit (source/Quaternion.cc:44-53)
          if (!(std::abs(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] +
                         q[3] * q[3] - 1.0) < 1e-10))
            throw AdmError(ADM_ERR_INVALID,
                           "five-axis scan path: the quaternion is not normed");
        }
    } else if (bd.desc.type < 0 || bd.desc.type > 2) {
      throw AdmError(ADM_ERR_INVALID, "unknown heat source type");
    }
    bd.desc.segments = nullptr;
    beams.push_back(bd);
  }
  old_segment.assign(beams.size(), 0);
  beam_states.resize(ADM_MAX_BEAMS);
  std::memset(beam_states.data(), 0, sizeof(adm_beam_state) * ADM_MAX_BEAMS);

  ADM_HIP(hipGetDevice(&device));
  d_scratch = scratch_on_device(device);
  ADM_HIP(hipMalloc(&d_kc, sizeof(KConst)));
  ADM_HIP(hipMalloc(&d_meta, sizeof(uint32_t) * n_cells));
  ADM_HIP(hipMalloc(&d_powder_slot, sizeof(int32_t) * n_cells));
  ADM_HIP(hipMalloc(&d_cos, sizeof(double) * n_cells));
  ADM_HIP(hipMalloc(&d_sin, sizeof(double) * n_cells));
  ADM_HIP(hipMalloc(&d_wgl, sizeof(double) * kMaxN));
  ADM_HIP(hipMalloc(&d_fn_rank, sizeof(int32_t) * n_dofs));
  ADM_HIP(hipMemcpy(d_wgl, wgl, sizeof(double) * n1, hipMemcpyHostToDevice));
  if (n_states > 1) {
    ADM_HIP(hipMalloc(&d_liquid, sizeof(double) * n_cells * nq));
    ADM_HIP(hipMalloc(&d_liquid_dirty, n_cells));
  }
  upload_constants();
  rebuild_cells();
  set_cell_state(cell_state.data());
}

#define ADM_FREE(p) (void)hipFree(p)
Context::~Context() {
  ADM_FREE(d_kc);
  ADM_FREE(d_meta);
  ADM_FREE(d_powder_slot);
  ADM_FREE(d_cos);
  ADM_FREE(d_sin);
  ADM_FREE(d_wgl);
  ADM_FREE(d_liquid);
  ADM_FREE(d_liquid_dirty);
  ADM_FREE(d_powder);
  ADM_FREE(d_face_powder);
  ADM_FREE(d_faces);
  ADM_FREE(d_face_out);
  ADM_FREE(d_fn_node);
  ADM_FREE(d_fn_offset);
  ADM_FREE(d_fn_slot);
  ADM_FREE(d_fn_rank);
  ADM_FREE(d_beam_table);
  ADM_FREE(d_slot);
  ADM_FREE(d_patch);
  clear_dof_map();
}

void Context::clear_dof_map() {
  ADM_FREE(d_dof_map);
  ADM_FREE(d_map_src);
  ADM_FREE(d_map_dst);
  d_dof_map = nullptr;
  d_map_src = d_map_dst = nullptr;
  n_ext = -1;
}

// The map must be injective into the box nodes and cover every node of
// every active cell (the nodes the operator reads and writes); the other
// nodes hold zeros in the node vectors.
void Context::set_dof_map(int64_t n, const int64_t *map) {
  if (n < 0 || (n > 0 && !map))
    throw AdmError(ADM_ERR_INVALID, "DoF map: negative size or null map");
  std::vector<uint8_t> seen((size_t)n_dofs, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (map[i] < 0 || map[i] >= n_dofs)
      throw AdmError(ADM_ERR_INVALID, "DoF map: entry out of the box nodes");
    if (seen[(size_t)map[i]]++)
      throw AdmError(ADM_ERR_INVALID, "DoF map: a node is mapped twice");
  }
  const std::vector<uint8_t> na = node_activity();
  for (int64_t k = 0; k < n_dofs; ++k)
    if (na[(size_t)k] && !seen[(size_t)k])
      throw AdmError(ADM_ERR_INVALID,
                     "DoF map: a node of an active cell has no DoF");
  ADM_HIP(hipDeviceSynchronize());
  clear_dof_map();
  ADM_HIP(hipMalloc(&d_dof_map, sizeof(int64_t) * std::max<int64_t>(n, 1)));
  ADM_HIP(hipMalloc(&d_map_src, sizeof(double) * n_dofs));
  ADM_HIP(hipMalloc(&d_map_dst, sizeof(double) * n_dofs));
  if (n)
    ADM_HIP(hipMemcpy(d_dof_map, map, sizeof(int64_t) * n,
                      hipMemcpyHostToDevice));
  ADM_HIP(hipMemset(d_map_src, 0, sizeof(double) * n_dofs));
  ADM_HIP(hipMemset(d_map_dst, 0, sizeof(double) * n_dofs));
  n_ext = n;
}

void Context::to_nodes(const double *ext, double *node, hipStream_t s) const {
  if (n_ext <= 0) return;
  hipLaunchKernelGGL(dofs_to_nodes_kernel,
                     dim3((unsigned)((n_ext + 255) / 256)), dim3(256), 0, s,
                     ext, d_dof_map, n_ext, node);
  ADM_HIP(hipGetLastError());
}

void Context::to_dofs(const double *node, double *ext, int add,
                      hipStream_t s) const {
  if (n_ext <= 0) return;
  hipLaunchKernelGGL(nodes_to_dofs_kernel,
                     dim3((unsigned)((n_ext + 255) / 256)), dim3(256), 0, s,
                     node, d_dof_map, n_ext, ext, add);
  ADM_HIP(hipGetLastError());
}

void Context::upload_constants() {
  ADM_HIP(hipMemcpy(d_kc, &kc, sizeof(KConst), hipMemcpyHostToDevice));
}

// Faces of the active domain that carry a boundary term
// (source/ThermalOperator.templates.hh:629-652 with Boundary::get_boundary_type)
uint32_t Context::face_mask(int64_t c) const {
  if (kc.adiabatic_only) return 0;
  const int cc[3] = {(int)(c % nc[0]), (int)((c / nc[0]) % nc[1]),
                     (int)(c / ((int64_t)nc[0] * nc[1]))};
  uint32_t mask = 0;
  for (int f = 0; f < 6; ++f) {
    if (!has_face(f)) continue;
    const int dd = f / 2, side = f % 2;
    const int nb = cc[dd] + (side ? 1 : -1);
    unsigned bid;
    if (nb < 0 || nb >= nc[dd]) {
      bid = (unsigned)face_id(f);
    } else {
      int cn[3] = {cc[0], cc[1], cc[2]};
      cn[dd] = nb;
      const int64_t c2 =
          cn[0] + (int64_t)nc[0] * (cn[1] + (int64_t)nc[1] * cn[2]);
      if (active[c2]) continue;
      bid = 0xffffffffu;
    }
    const int n = kc.n_boundary_ids;
    const int t = bid < (unsigned)(n - 1) ? kc.boundary_type[bid]
                                          : kc.boundary_type[n - 1];
    if (t & (ADM_BC_RADIATIVE | ADM_BC_CONVECTIVE)) mask |= 1u << f;
  }
  return mask;
}

void Context::rebuild_cells() {
  meta.assign(n_cells, 0);
  powder_slot.assign(n_cells, -1);
  n_powder_cells = 0;
  nz_active = 0;
  faces.clear();
  for (int64_t c = 0; c < n_cells; ++c) {
    uint32_t w = 0;
    if (active[c]) {
      w |= kActive;
      nz_active = std::max(nz_active,
                           (int)(c / ((int64_t)nc[0] * nc[1])) + 1);
    }
    w |= (uint32_t)material[c] << kMaterialShift;
    if (active[c]) {
      const uint32_t fm = face_mask(c);
      w |= fm << kFaceShift;
      for (int f = 0; f < 6; ++f) {
        if (!(fm & (1u << f))) continue;
        const int cc[3] = {(int)(c % nc[0]), (int)((c / nc[0]) % nc[1]),
                           (int)(c / ((int64_t)nc[0] * nc[1]))};
        const int dd = f / 2, side = f % 2;
        const bool domain = side ? (cc[dd] == nc[dd] - 1) : (cc[dd] == 0);
        const unsigned bid = domain ? (unsigned)face_id(f) : 0xffffffffu;
        const int nb = kc.n_boundary_ids;
        const int bt = bid < (unsigned)(nb - 1) ? kc.boundary_type[bid]
                                                : kc.boundary_type[nb - 1];
        FaceRef fr;
        fr.cell = (int32_t)c;
        fr.info = f | ((bt & 31) << 3) | ((int)material[c] << 8);
        faces.push_back(fr);
      }
    }
    if (dep_cos[c] != 1.0 || dep_sin[c] != 0.0) w |= kRotated;
    if (n_states == 3 && cell_state[c * 3 + 2] > 0.0) {
      w |= kHasPowder;
      powder_slot[c] = (int32_t)n_powder_cells++;
    }
    meta[c] = w;
  }
  ADM_HIP(hipMemcpy(d_meta, meta.data(), sizeof(uint32_t) * n_cells,
                    hipMemcpyHostToDevice));
  ADM_HIP(hipMemcpy(d_powder_slot, powder_slot.data(),
                    sizeof(int32_t) * n_cells, hipMemcpyHostToDevice));
  ADM_HIP(hipMemcpy(d_cos, dep_cos.data(), sizeof(double) * n_cells,
                    hipMemcpyHostToDevice));
  ADM_HIP(hipMemcpy(d_sin, dep_sin.data(), sizeof(double) * n_cells,
                    hipMemcpyHostToDevice));
  ADM_FREE(d_powder);
  ADM_FREE(d_face_powder);
  ADM_FREE(d_faces);
  d_powder = nullptr;
  d_face_powder = nullptr;
  d_faces = nullptr;
  if (n_powder_cells > 0)
    ADM_HIP(hipMalloc(&d_powder, sizeof(double) * n_powder_cells * nq));
  n_faces = (int64_t)faces.size();
  faces_stale = false;
  face_cap = face_powder_cap = 0;
  if (!faces.empty()) {
    ADM_HIP(hipMalloc(&d_faces, sizeof(FaceRef) * faces.size()));
    ADM_HIP(hipMemcpy(d_faces, faces.data(), sizeof(FaceRef) * faces.size(),
                      hipMemcpyHostToDevice));
    face_cap = faces.size();
    if (n_states == 3) {
      ADM_HIP(hipMalloc(&d_face_powder, sizeof(double) * faces.size() * nfq));
      face_powder_cap = faces.size() * nfq;
    }
  }
  build_face_nodes();
}

// The faces' node lists of the face gather pass: every (face, local node)
// slot, grouped by node (nodes in increasing order, each node's slots in
// increasing face order).
void Context::build_face_nodes() {
  ADM_FREE(d_face_out);
  ADM_FREE(d_fn_node);
  ADM_FREE(d_fn_offset);
  ADM_FREE(d_fn_slot);
  d_face_out = nullptr;
  d_fn_node = nullptr;
  d_fn_offset = nullptr;
  d_fn_slot = nullptr;
  n_fnodes = 0;
  face_out_cap = fn_node_cap = fn_offset_cap = fn_slot_cap = 0;
  ADM_HIP(hipMemset(d_fn_rank, 0xff, sizeof(int32_t) * n_dofs));
  if (faces.empty()) return;
  const int N = p + 1, NF = nfq;
  const int64_t ns[3] = {1, (int64_t)nn[0], (int64_t)nn[0] * nn[1]};
  std::vector<std::pair<int64_t, int32_t>> ent;
  ent.reserve(faces.size() * NF);
  for (size_t f = 0; f < faces.size(); ++f) {
    const int64_t c = faces[f].cell;
    const int fid = faces[f].info & 7, d = fid / 2, side = fid % 2;
    const int64_t cc[3] = {c % nc[0], (c / nc[0]) % nc[1],
                           c / ((int64_t)nc[0] * nc[1])};
    const int64_t base = cc[0] * p + ns[1] * cc[1] * p + ns[2] * cc[2] * p +
                         (side ? p : 0) * ns[d];
    if (dim == 2) {
      // an edge: its nodes along the other axis of the (x, z) plane
      const int64_t ts = d == 0 ? ns[2] : 1;
      for (int q = 0; q < NF; ++q)
        ent.emplace_back(base + q * ts, (int32_t)(f * NF + q));
      continue;
    }
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    for (int q = 0; q < NF; ++q)
      ent.emplace_back(base + (q % N) * ns[t1] + (q / N) * ns[t2],
                       (int32_t)(f * NF + q));
  }
  if (ent.size() > (size_t)INT32_MAX)
    throw AdmError(ADM_ERR_UNSUPPORTED, "too many boundary-face nodes");
  std::stable_sort(ent.begin(), ent.end(),
                   [](const std::pair<int64_t, int32_t> &x,
                      const std::pair<int64_t, int32_t> &y) {
                     return x.first < y.first;
                   });
  std::vector<int64_t> node;
  std::vector<int32_t> off, slot;
  for (size_t e = 0; e < ent.size(); ++e) {
    if (e == 0 || ent[e].first != ent[e - 1].first) {
      node.push_back(ent[e].first);
      off.push_back((int32_t)e);
    }
    slot.push_back(ent[e].second);
  }
  off.push_back((int32_t)ent.size());
  n_fnodes = (int64_t)node.size();
  ADM_HIP(hipMalloc(&d_face_out, sizeof(double) * faces.size() * NF));
  ADM_HIP(hipMalloc(&d_fn_node, sizeof(int64_t) * node.size()));
  ADM_HIP(hipMalloc(&d_fn_offset, sizeof(int32_t) * off.size()));
  ADM_HIP(hipMalloc(&d_fn_slot, sizeof(int32_t) * slot.size()));
  face_out_cap = faces.size() * NF;
  fn_node_cap = node.size();
  fn_offset_cap = off.size();
  fn_slot_cap = slot.size();
  {
    std::vector<int32_t> rank((size_t)n_dofs, -1);
    for (size_t k = 0; k < node.size(); ++k) rank[(size_t)node[k]] = (int32_t)k;
    ADM_HIP(hipMemcpy(d_fn_rank, rank.data(), sizeof(int32_t) * n_dofs,
                      hipMemcpyHostToDevice));
  }
  ADM_HIP(hipMemcpy(d_fn_node, node.data(), sizeof(int64_t) * node.size(),
                    hipMemcpyHostToDevice));
  ADM_HIP(hipMemcpy(d_fn_offset, off.data(), sizeof(int32_t) * off.size(),
                    hipMemcpyHostToDevice));
  ADM_HIP(hipMemcpy(d_fn_slot, slot.data(), sizeof(int32_t) * slot.size(),
                    hipMemcpyHostToDevice));
}

// get_state_from_material_properties: every quadrature point of a cell (and
// of its faces) takes the cell's ratios.
void Context::set_cell_state(const double *cs) {
  if (cs != cell_state.data())
    std::copy(cs, cs + (size_t)n_cells * n_states, cell_state.begin());
  // powder storage follows the cells that hold powder
  bool need_rebuild = false;
  if (n_states == 3) {
    for (int64_t c = 0; c < n_cells; ++c)
      if ((cell_state[c * 3 + 2] > 0.0) != (powder_slot[c] >= 0)) {
        need_rebuild = true;
        break;
      }
  }
  if (need_rebuild) rebuild_cells();
  if (n_states > 1) {
    std::vector<double> lq((size_t)n_cells * nq);
    std::vector<uint8_t> dirty(n_cells, 0);
    for (int64_t c = 0; c < n_cells; ++c) {
      const double l = cell_state[c * n_states + 1];
      for (int q = 0; q < nq; ++q) lq[c * nq + q] = l;
      dirty[c] = l != 0.0;
    }
    ADM_HIP(hipMemcpy(d_liquid, lq.data(), sizeof(double) * lq.size(),
                      hipMemcpyHostToDevice));
    ADM_HIP(hipMemcpy(d_liquid_dirty, dirty.data(), n_cells,
                      hipMemcpyHostToDevice));
  }
  if (n_powder_cells > 0) {
    std::vector<double> pw((size_t)n_powder_cells * nq);
    for (int64_t c = 0; c < n_cells; ++c) {
      const int64_t s = powder_slot[c];
      if (s < 0) continue;
      const double p = cell_state[c * 3 + 2];
      for (int q = 0; q < nq; ++q) pw[s * nq + q] = p;
    }
    ADM_HIP(hipMemcpy(d_powder, pw.data(), sizeof(double) * pw.size(),
                      hipMemcpyHostToDevice));
  }
  sync_faces_host();
  if (n_states == 3 && !faces.empty()) {
    std::vector<double> fpw(faces.size() * nfq);
    for (size_t f = 0; f < faces.size(); ++f)
      for (int q = 0; q < nfq; ++q)
        fpw[f * nfq + q] = cell_state[(int64_t)faces[f].cell * 3 + 2];
    ADM_HIP(hipMemcpy(d_face_powder, fpw.data(), sizeof(double) * fpw.size(),
                      hipMemcpyHostToDevice));
  }
}

void Context::get_qp_state(double *liquid, double *powder) {
  ADM_HIP(hipDeviceSynchronize());
  if (liquid) {
    if (n_states > 1)
      ADM_HIP(hipMemcpy(liquid, d_liquid, sizeof(double) * n_cells * nq,
                        hipMemcpyDeviceToHost));
    else
      std::fill(liquid, liquid + (size_t)n_cells * nq, 0.0);
  }
  if (powder) {
    std::fill(powder, powder + (size_t)n_cells * nq, 0.0);
    if (n_powder_cells > 0) {
      std::vector<double> pw((size_t)n_powder_cells * nq);
      ADM_HIP(hipMemcpy(pw.data(), d_powder, sizeof(double) * pw.size(),
                        hipMemcpyDeviceToHost));
      for (int64_t c = 0; c < n_cells; ++c)
        if (powder_slot[c] >= 0)
          std::copy(pw.begin() + powder_slot[c] * nq,
                    pw.begin() + (powder_slot[c] + 1) * nq, powder + c * nq);
    }
  }
}

void Context::set_qp_state(const double *liquid, const double *powder) {
  if (liquid && n_states > 1) {
    ADM_HIP(hipMemcpy(d_liquid, liquid, sizeof(double) * n_cells * nq,
                      hipMemcpyHostToDevice));
    std::vector<uint8_t> dirty(n_cells, 0);
    for (int64_t c = 0; c < n_cells; ++c)
      for (int q = 0; q < nq; ++q)
        if (liquid[c * nq + q] != 0.0) dirty[c] = 1;
    ADM_HIP(hipMemcpy(d_liquid_dirty, dirty.data(), n_cells,
                      hipMemcpyHostToDevice));
  }
  if (powder && n_states == 3) {
    for (int64_t c = 0; c < n_cells; ++c) {
      bool any = false;
      for (int q = 0; q < nq; ++q) any |= powder[c * nq + q] > 0.0;
      if (any && powder_slot[c] < 0)
        throw AdmError(ADM_ERR_INVALID,
                       "powder can only be set in cells that hold powder");
    }
    if (n_powder_cells > 0) {
      std::vector<double> pw((size_t)n_powder_cells * nq);
      for (int64_t c = 0; c < n_cells; ++c)
        if (powder_slot[c] >= 0)
          std::copy(powder + c * nq, powder + (c + 1) * nq,
                    pw.begin() + powder_slot[c] * nq);
      ADM_HIP(hipMemcpy(d_powder, pw.data(), sizeof(double) * pw.size(),
                        hipMemcpyHostToDevice));
    }
  }
}

// MaterialProperty::set_state (source/MaterialProperty.templates.hh:
// 537-620): cell means of the quadrature-point ratios.
void Context::cell_state_means(double *cs) {
  const int ns = n_states;
  std::vector<double> lq((size_t)n_cells * nq), pw((size_t)n_cells * nq);
  get_qp_state(lq.data(), pw.data());
  for (int64_t c = 0; c < n_cells; ++c) {
    if (ns == 1) {
      cs[c] = 1.0;
      continue;
    }
    double ls = 0.0, ps = 0.0;
    for (int q = 0; q < nq; ++q) {
      ls += lq[c * nq + q];
      ps += pw[c * nq + q];
    }
    const double l = ls / nq;
    if (ns == 2) {
      cs[c * 2 + 1] = l;
      cs[c * 2] = 1. - l;
    } else {
      const double pp = ps / nq;
      cs[c * 3 + 1] = l;
      cs[c * 3 + 2] = pp;
      cs[c * 3] = 1. - l - pp;
    }
  }
}

// Nodes of active cells.
std::vector<uint8_t> Context::node_activity() const {
  std::vector<uint8_t> na((size_t)n_dofs, 0);
  for (int64_t c = 0; c < n_cells; ++c) {
    if (!active[c]) continue;
    const int64_t cx = c % nc[0], cy = (c / nc[0]) % nc[1],
                  cz = c / ((int64_t)nc[0] * nc[1]);
    for (int k = 0; k <= p; ++k)
      for (int j = 0; j <= (dim == 3 ? p : 0); ++j)
        for (int i = 0; i <= p; ++i)
          na[(size_t)(cx * p + i) +
             (size_t)nn[0] * ((cy * p + j) + (size_t)nn[1] * (cz * p + k))] = 1;
  }
  return na;
}

// get_elements_to_activate (source/material_deposition.cc:376-435): for
// every inactive cell, the last box (input order) whose closed box meets the
// cell's bounding box, else -1. boxes[b] = {min x, y, z, max x, y, z}.
void Context::cells_to_activate(int n_boxes, const double *boxes,
                                int32_t *box_of_cell) const {
  for (int64_t c = 0; c < n_cells; ++c) {
    box_of_cell[c] = -1;
    if (active[c]) continue;
    const int64_t cc[3] = {c % nc[0], (c / nc[0]) % nc[1],
                           c / ((int64_t)nc[0] * nc[1])};
    for (int b = 0; b < n_boxes; ++b) {
      bool hit = true;
      for (int d = 0; d < 3; ++d) {
        if (dim == 2 && d == 1) continue;  // boxes are (x, -, z) in 2-D
        const double lo = kc.origin[d] + cc[d] * kc.h[d];
        const double hi = kc.origin[d] + (cc[d] + 1) * kc.h[d];
        if (!(lo <= boxes[6 * b + 3 + d] && boxes[6 * b + d] <= hi)) hit = false;
      }
      if (hit) box_of_cell[c] = b;
    }
  }
}

// add_material_start + add_material_end (source/ThermalPhysics.templates.hh:
// 595-825) on the structured numbering: phase state to cell means and back
// to every quadrature point, activation with the box's deposition
// direction, new_temperature on nodes no previously active cell held, and
// the lumped mass recomputed.
void Context::add_material(const int32_t *box_of_cell, const double *box_cos,
                           const double *box_sin, double new_temperature,
                           double *T, hipStream_t s) {
  if (device_activation_fits()) {
    // boxes named by inactive cells (the others' entries are ignored)
    int nb = 0;
    std::vector<int32_t> boc(box_of_cell, box_of_cell + n_cells);
    for (int64_t c = 0; c < n_cells; ++c) {
      if (active[c]) boc[c] = -1;
      nb = std::max(nb, boc[c] + 1);
    }
    box_of_cell = boc.data();
    activate(nb, nullptr, box_of_cell, box_cos, box_sin, new_temperature, T,
             nullptr, s);
  } else {
    add_material_host(box_of_cell, box_cos, box_sin, new_temperature, T, s);
  }
}

// cells_to_activate and add_material_host in one call (the activation of
// 2-D meshes)
void Context::activate_host(int n_boxes, const double *boxes,
                            const double *box_cos, const double *box_sin,
                            double new_temperature, double *T,
                            int32_t *box_of_cell_out, hipStream_t s) {
  std::vector<int32_t> boc((size_t)n_cells);
  cells_to_activate(n_boxes, boxes, boc.data());
  add_material_host(boc.data(), box_cos, box_sin, new_temperature, T, s);
  if (box_of_cell_out) std::copy(boc.begin(), boc.end(), box_of_cell_out);
}

// the device activation's counters and scans are 32-bit
bool Context::device_activation_fits() const {
  // the device activation's face and node kernels are three-dimensional
  if (dim == 2) return false;
  const int64_t NF = (int64_t)(p + 1) * (p + 1);
  return n_cells * 6 * NF < (int64_t)INT32_MAX && n_dofs < (int64_t)INT32_MAX;
}

// The same on the host (meshes too large for the device activation).
void Context::add_material_host(const int32_t *box_of_cell,
                                const double *box_cos, const double *box_sin,
                                double new_temperature, double *T,
                                hipStream_t s) {
  ADM_HIP(hipStreamSynchronize(s));
  sync_faces_host();
  std::vector<double> cs((size_t)n_cells * n_states);
  cell_state_means(cs.data());
  const std::vector<uint8_t> before = node_activity();
  for (int64_t c = 0; c < n_cells; ++c) {
    const int b = box_of_cell[c];
    if (b < 0 || active[c]) continue;
    active[c] = 1;
    dep_cos[c] = box_cos[b];
    dep_sin[c] = box_sin[b];
  }
  const std::vector<uint8_t> after = node_activity();
  std::vector<int64_t> fresh;
  for (int64_t i = 0; i < n_dofs; ++i)
    if (after[i] && !before[i]) fresh.push_back(i);
  std::copy(cs.begin(), cs.end(), cell_state.begin());
  rebuild_cells();
  set_cell_state(cell_state.data());
  if (T && !fresh.empty()) {
    int64_t *d_idx = nullptr;
    ADM_HIP(hipMalloc(&d_idx, sizeof(int64_t) * fresh.size()));
    ADM_HIP(hipMemcpyAsync(d_idx, fresh.data(), sizeof(int64_t) * fresh.size(),
                           hipMemcpyHostToDevice, s));
    const int64_t n = (int64_t)fresh.size();
    hipLaunchKernelGGL(scatter_value_kernel, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, s, T, d_idx, new_temperature, n);
    ADM_HIP(hipGetLastError());
    ADM_HIP(hipStreamSynchronize(s));
    ADM_HIP(hipFree(d_idx));
  }
  if (sb.inv_mass) {
    compute_inverse_mass(sb.inv_mass, s);
    ADM_HIP(hipStreamSynchronize(s));
  }
}

void Context::set_time(double t) {
  sources_on = false;
  for (size_t b = 0; b < beams.size(); ++b) {
    beam_update(beams[b], t, old_segment[b], beam_states[b]);
    if (beam_states[b].on) sources_on = true;
  }
}

CellArgs Context::cell_args() const {
  CellArgs a;
  std::memset(&a, 0, sizeof(a));
  a.xcd_remap = xcd_order;
  a.meta = d_meta;
  a.powder_slot = d_powder_slot;
  a.powder = d_powder;
  a.face_powder = d_face_powder;
  a.liquid = d_liquid;
  a.liquid_dirty = d_liquid_dirty;
  a.dep_cos = d_cos;
  a.dep_sin = d_sin;
  a.kc = d_kc;
  for (int k = 0; k < 3; ++k) {
    a.nc[k] = nc[k];
    a.nn[k] = nn[k];
  }
  a.n_beams = (int)beams.size();
  a.sources_on = sources_on ? 1 : 0;
  a.n_cells = n_cells;
  for (size_t b = 0; b < beams.size(); ++b) a.beams[b] = beam_states[b];
  return a;
}

FaceArgs Context::face_args() const {
  FaceArgs a;
  std::memset(&a, 0, sizeof(a));
  a.faces = d_faces;
  a.face_powder = d_face_powder;
  a.kc = d_kc;
  for (int k = 0; k < 3; ++k) {
    a.nc[k] = nc[k];
    a.nn[k] = nn[k];
  }
  a.n_faces = n_faces;
  a.face_out = d_face_out;
  a.fn_node = d_fn_node;
  a.fn_offset = d_fn_offset;
  a.fn_slot = d_fn_slot;
  a.n_fnodes = n_fnodes;
  a.fn_rank = d_fn_rank;
  return a;
}

void Context::apply(const double *src, double *dst, int add, hipStream_t s,
                    double *patch) {
  CellArgs a = cell_args();
  a.add = add;
  launch(a, face_args(), src, dst, 1, s, nz_active, patch);
}

// Cell and face kernels of the context's configuration; with n_members > 1
// the arguments carry the member table and the grid's y index is the member.
void Context::launch(const CellArgs &a, const FaceArgs &fa, const double *src,
                     double *dst, unsigned n_members, hipStream_t s,
                     int nz_rows, double *patch, const StageArgs *st) const {
  if (st && !(use_plane && p <= 2))
    throw AdmError(ADM_ERR_INVALID, "a fused stage needs the plane kernel");
  StageArgs stage;
  std::memset(&stage, 0, sizeof(stage));
  if (st) stage = *st;
  // cell layers above the highest active one hold no active cell
  const int lim[3] = {nc[0], nc[1],
                      nz_rows >= 0 ? std::min(nz_rows, nc[2]) : nc[2]};
  const int pk = property_kind();
  const size_t lds = pk == 2 ? sizeof(double) * kMergedLds * kc.mat.n_materials
                             : 0;
  const unsigned M = n_members;
  if (dim == 2) {
    launch_2d(a, fa, src, dst, M, lim, s);
    return;
  }
  // the face term's values into its slots (face_out)
  auto launch_faces = [&]() {
    if (!fa.n_faces) return;
#define ADM_DISPATCH(P_, NS_)                                           \
  if (p == P_ && n_states == NS_) {                                      \
    if (pk == 0)                                                         \
      launch_face<P_, NS_, 0>(fa, src, dst, M, s);                       \
    else if (pk == 1)                                                    \
      launch_face<P_, NS_, 1>(fa, src, dst, M, s);                       \
    else                                                                 \
      launch_face<P_, NS_, 2>(fa, src, dst, M, s);                       \
  }
    ADM_DISPATCH(1, 1)
    ADM_DISPATCH(1, 2)
    ADM_DISPATCH(1, 3)
    ADM_DISPATCH(2, 1)
    ADM_DISPATCH(2, 2)
    ADM_DISPATCH(2, 3)
    ADM_DISPATCH(3, 1)
    ADM_DISPATCH(3, 2)
    ADM_DISPATCH(3, 3)
    ADM_DISPATCH(4, 1)
    ADM_DISPATCH(4, 2)
    ADM_DISPATCH(4, 3)
    ADM_DISPATCH(5, 1)
    ADM_DISPATCH(5, 2)
    ADM_DISPATCH(5, 3)
#undef ADM_DISPATCH
  };
  const bool plane = use_plane && p <= 2;
  if (plane) {
    // one launch, one PX x PY cell patch of a cell layer per wavefront
    const size_t plds =
        pk == 2 ? sizeof(double) * kPlaneMergedLds * kc.mat.n_materials : 0;
    CellArgs ac = a;
    ac.color = -1;
    for (int d = 0; d < 3; ++d) ac.lext[d] = lim[d];
    const int px = p == 1 ? PlaneLanes<1>::PX : PlaneLanes<2>::PX;
    const int py = p == 1 ? PlaneLanes<1>::PY : PlaneLanes<2>::PY;
    ac.plane_npx = (lim[0] + px - 1) / px;
    ac.plane_npy = (lim[1] + py - 1) / py;
    ac.n_rows = ac.plane_npx * ac.plane_npy * lim[2];
    const int64_t slots = p == 1 ? PlaneLanes<1>::patch_nodes
                                 : PlaneLanes<2>::patch_nodes;
    ac.plane_E_stride = (int64_t)ac.n_rows * slots;
    if (patch) {
      if ((size_t)ac.plane_E_stride * M > kNoAllocScratchDoubles)
        throw AdmError(ADM_ERR_UNSUPPORTED,
                       "the mesh needs a patch buffer of " +
                           std::to_string((size_t)ac.plane_E_stride * M) +
                           " doubles, more than the scratch array's " +
                           std::to_string(kNoAllocScratchDoubles));
      ac.plane_E = patch;
    } else {
      ac.plane_E = ensure_patch_buffer((size_t)ac.plane_E_stride * M);
    }
    if (ac.n_rows > 0) {
      // a persistent grid: as many blocks as the chip holds at the
      // kernel's occupancy (shared among the members), each looping over
      // patches; at least 8 blocks, one per XCD
      const int64_t full = (ac.n_rows + cell_waves - 1) / cell_waves;
      const int64_t resident =
          (int64_t)n_cu * 4 * kPlaneWavesPerEU / cell_waves / (int64_t)M;
      const int64_t nb = std::min(full, std::max<int64_t>(8, resident));
#define ADM_DISPATCH(P_, NS_)                                              \
  if (p == P_ && n_states == NS_) {                                         \
    if (pk == 0)                                                            \
      launch_plane<P_, NS_, 0>(ac, src, dst, nb, plds, M, cell_waves, s);   \
    else if (pk == 1)                                                       \
      launch_plane<P_, NS_, 1>(ac, src, dst, nb, plds, M, cell_waves, s);   \
    else                                                                    \
      launch_plane<P_, NS_, 2>(ac, src, dst, nb, plds, M, cell_waves, s);   \
  }
      ADM_DISPATCH(1, 1)
      ADM_DISPATCH(1, 2)
      ADM_DISPATCH(1, 3)
      ADM_DISPATCH(2, 1)
      ADM_DISPATCH(2, 2)
      ADM_DISPATCH(2, 3)
#undef ADM_DISPATCH
    }
    // the face term's slots, then the second pass: every node's sum from
    // the patch buffer and its face slots into dst
    launch_faces();
    // a stage update leaves the nodes above the active cell layers as they
    // are (the gather returns at once there), so its grid stops at them;
    // dst = A src is zero there and is written
    const int64_t n_gather =
        st ? std::min<int64_t>(n_dofs,
                               (int64_t)nn[0] * nn[1] *
                                   std::min<int64_t>(nn[2],
                                                     (int64_t)p * lim[2] + 1))
           : n_dofs;
    const dim3 ggrid((unsigned)((n_gather + 255) / 256), M);
    if (p == 1)
      hipLaunchKernelGGL(plane_gather_kernel<1>, ggrid, dim3(256), 0, s,
                         ac.plane_E, ac.plane_E_stride, dst,
                         a.members ? a.member_stride : (int64_t)0, nn[0],
                         nn[1], nn[2], ac.plane_npx, ac.plane_npy, lim[2],
                         a.add, fa, stage);
    else
      hipLaunchKernelGGL(plane_gather_kernel<2>, ggrid, dim3(256), 0, s,
                         ac.plane_E, ac.plane_E_stride, dst,
                         a.members ? a.member_stride : (int64_t)0, nn[0],
                         nn[1], nn[2], ac.plane_npx, ac.plane_npy, lim[2],
                         a.add, fa, stage);
  }
  // the line kernel runs once per parity color (cells of one color share no
  // node: plain read-add-write of dst), or once with atomics
  const int n_colors = (use_plane && p <= 2) ? 0 : (use_coloring ? 8 : 1);
  const int cpw = cells_per_wave(p);
  for (int col = 0; col < n_colors; ++col) {
    CellArgs ac = a;
    ac.color = use_coloring ? col : -1;
    int64_t cells = 1;
    for (int d = 0; d < 3; ++d) {
      const int b = use_coloring ? (col >> d) & 1 : 0;
      ac.lext[d] =
          use_coloring ? (lim[d] > b ? (lim[d] - b + 1) / 2 : 0) : lim[d];
      cells *= ac.lext[d];
    }
    if (cells == 0) continue;
    ac.nseg = (ac.lext[0] + cpw - 1) / cpw;
    ac.n_rows = ac.nseg * ac.lext[1] * ac.lext[2];
    // one wavefront per row segment, cell_waves per block
    const int64_t nlaunch = (ac.n_rows + cell_waves - 1) / cell_waves;
#define ADM_DISPATCH(P_, NS_)                                          \
  if (p == P_ && n_states == NS_) {                                     \
    if (pk == 0)                                                        \
      launch_cell<P_, NS_, 0>(ac, src, dst, nlaunch, lds, M, cell_waves, \
                              s);                                       \
    else if (pk == 1)                                                   \
      launch_cell<P_, NS_, 1>(ac, src, dst, nlaunch, lds, M, cell_waves, \
                              s);                                       \
    else                                                                \
      launch_cell<P_, NS_, 2>(ac, src, dst, nlaunch, lds, M, cell_waves, \
                              s);                                       \
  }
    ADM_DISPATCH(1, 1)
    ADM_DISPATCH(1, 2)
    ADM_DISPATCH(1, 3)
    ADM_DISPATCH(2, 1)
    ADM_DISPATCH(2, 2)
    ADM_DISPATCH(2, 3)
    ADM_DISPATCH(3, 1)
    ADM_DISPATCH(3, 2)
    ADM_DISPATCH(3, 3)
    ADM_DISPATCH(4, 1)
    ADM_DISPATCH(4, 2)
    ADM_DISPATCH(4, 3)
    ADM_DISPATCH(5, 1)
    ADM_DISPATCH(5, 2)
    ADM_DISPATCH(5, 3)
#undef ADM_DISPATCH
  }
  if (!plane) {
    launch_faces();
    if (fa.n_faces && fa.n_fnodes)
      hipLaunchKernelGGL(face_gather_kernel,
                         dim3((unsigned)((fa.n_fnodes + 255) / 256), M),
                         dim3(256), 0, s, fa, dst);
  }
  ADM_HIP(hipGetLastError());
}

// dim = 2: the cell kernel once per parity color of (x, z) (plain
// read-add-write) or once with atomics, then the face kernel and the face
// gather pass.
void Context::launch_2d(const CellArgs &a, const FaceArgs &fa,
                        const double *src, double *dst, unsigned M,
                        const int (&lim)[3], hipStream_t s) const {
  const int pk = property_kind();
  const int n_colors = use_coloring ? 4 : 1;
  const int cpw = 64 / (p + 1);
  for (int col = 0; col < n_colors; ++col) {
    CellArgs ac = a;
    ac.color = use_coloring ? col : -1;
    const int bx = use_coloring ? col & 1 : 0;
    const int bz = use_coloring ? (col >> 1) & 1 : 0;
    ac.lext[0] = use_coloring ? (lim[0] > bx ? (lim[0] - bx + 1) / 2 : 0)
                              : lim[0];
    ac.lext[1] = 1;
    ac.lext[2] = use_coloring ? (lim[2] > bz ? (lim[2] - bz + 1) / 2 : 0)
                              : lim[2];
    if (ac.lext[0] == 0 || ac.lext[2] == 0) continue;
    ac.nseg = (ac.lext[0] + cpw - 1) / cpw;
    ac.n_rows = ac.nseg * ac.lext[2];
#define ADM_DISPATCH(P_, NS_)                                   \
  if (p == P_ && n_states == NS_) {                              \
    if (pk == 0)                                                 \
      launch_cell_2d<P_, NS_, 0>(ac, src, dst, M, s);            \
    else if (pk == 1)                                            \
      launch_cell_2d<P_, NS_, 1>(ac, src, dst, M, s);            \
    else                                                         \
      launch_cell_2d<P_, NS_, 2>(ac, src, dst, M, s);            \
  }
    ADM_DISPATCH(1, 1)
    ADM_DISPATCH(1, 2)
    ADM_DISPATCH(1, 3)
    ADM_DISPATCH(2, 1)
    ADM_DISPATCH(2, 2)
    ADM_DISPATCH(2, 3)
    ADM_DISPATCH(3, 1)
    ADM_DISPATCH(3, 2)
    ADM_DISPATCH(3, 3)
    ADM_DISPATCH(4, 1)
    ADM_DISPATCH(4, 2)
    ADM_DISPATCH(4, 3)
    ADM_DISPATCH(5, 1)
    ADM_DISPATCH(5, 2)
    ADM_DISPATCH(5, 3)
#undef ADM_DISPATCH
  }
  if (fa.n_faces) {
#define ADM_DISPATCH(P_, NS_)                                   \
  if (p == P_ && n_states == NS_) {                              \
    if (pk == 0)                                                 \
      launch_face_2d<P_, NS_, 0>(fa, src, M, s);                 \
    else if (pk == 1)                                            \
      launch_face_2d<P_, NS_, 1>(fa, src, M, s);                 \
    else                                                         \
      launch_face_2d<P_, NS_, 2>(fa, src, M, s);                 \
  }
    ADM_DISPATCH(1, 1)
    ADM_DISPATCH(1, 2)
    ADM_DISPATCH(1, 3)
    ADM_DISPATCH(2, 1)
    ADM_DISPATCH(2, 2)
    ADM_DISPATCH(2, 3)
    ADM_DISPATCH(3, 1)
    ADM_DISPATCH(3, 2)
    ADM_DISPATCH(3, 3)
    ADM_DISPATCH(4, 1)
    ADM_DISPATCH(4, 2)
    ADM_DISPATCH(4, 3)
    ADM_DISPATCH(5, 1)
    ADM_DISPATCH(5, 2)
    ADM_DISPATCH(5, 3)
#undef ADM_DISPATCH
    if (fa.n_fnodes)
      hipLaunchKernelGGL(face_gather_kernel,
                         dim3((unsigned)((fa.n_fnodes + 255) / 256), M),
                         dim3(256), 0, s, fa, dst);
  }
  ADM_HIP(hipGetLastError());
}

size_t Context::patch_buffer_size(int nz_rows, unsigned n_members) const {
  const int nz = nz_rows >= 0 ? std::min(nz_rows, nc[2]) : nc[2];
  const int px = p == 1 ? PlaneLanes<1>::PX : PlaneLanes<2>::PX;
  const int py = p == 1 ? PlaneLanes<1>::PY : PlaneLanes<2>::PY;
  const size_t slots =
      p == 1 ? PlaneLanes<1>::patch_nodes : PlaneLanes<2>::patch_nodes;
  return (size_t)((nc[0] + px - 1) / px) * ((nc[1] + py - 1) / py) * nz *
         slots * n_members;
}

double *Context::ensure_patch_buffer(size_t n) const {
  if (n > patch_cap) {
    if (d_patch) ADM_HIP(hipFree(d_patch));
    d_patch = nullptr;
    patch_cap = 0;
    ADM_HIP(hipMalloc(&d_patch, sizeof(double) * n));
    patch_cap = n;
  }
  return d_patch;
}

void Context::fill(double *x, double v, int64_t n, hipStream_t s) {
  const int block = 256;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + block - 1) / block)),
                     dim3(block), 0, s, x, v, n);
  ADM_HIP(hipGetLastError());
}

void Context::compute_inverse_mass(double *out, hipStream_t s) {
  const int block = 256;
  const unsigned grid = (unsigned)((n_dofs + block - 1) / block);
  if (dim == 2) {
#define ADM_MASS_2D(P_)                                                   \
  if (p == P_)                                                            \
    hipLaunchKernelGGL(inverse_mass_kernel_2d<P_>, dim3(grid), dim3(block), \
                       0, s, d_meta, d_wgl, det, nc[0], nc[2], n_dofs, out);
    ADM_MASS_2D(1)
    ADM_MASS_2D(2)
    ADM_MASS_2D(3)
    ADM_MASS_2D(4)
    ADM_MASS_2D(5)
#undef ADM_MASS_2D
    ADM_HIP(hipGetLastError());
    return;
  }
  if (p == 1)
    hipLaunchKernelGGL(inverse_mass_kernel<1>, dim3(grid), dim3(block), 0, s,
                       d_meta, d_wgl, det, nc[0], nc[1], nc[2], n_dofs, out);
  else if (p == 2)
    hipLaunchKernelGGL(inverse_mass_kernel<2>, dim3(grid), dim3(block), 0, s,
                       d_meta, d_wgl, det, nc[0], nc[1], nc[2], n_dofs, out);
  else if (p == 3)
    hipLaunchKernelGGL(inverse_mass_kernel<3>, dim3(grid), dim3(block), 0, s,
                       d_meta, d_wgl, det, nc[0], nc[1], nc[2], n_dofs, out);
  else if (p == 4)
    hipLaunchKernelGGL(inverse_mass_kernel<4>, dim3(grid), dim3(block), 0, s,
                       d_meta, d_wgl, det, nc[0], nc[1], nc[2], n_dofs, out);
  else
    hipLaunchKernelGGL(inverse_mass_kernel<5>, dim3(grid), dim3(block), 0, s,
                       d_meta, d_wgl, det, nc[0], nc[1], nc[2], n_dofs, out);
  ADM_HIP(hipGetLastError());
}

void StepBuffers::ensure(int method, int64_t n_) {
  n = n_;
  if (!R) {
    ADM_HIP(hipMalloc(&R, sizeof(double) * n));
    ADM_HIP(hipMemset(R, 0, sizeof(double) * n));
  }
  if (method != ADM_FORWARD_EULER && !Y) {
    ADM_HIP(hipMalloc(&Y, sizeof(double) * n));
    ADM_HIP(hipMalloc(&acc, sizeof(double) * n));
  }
  if (method == ADM_RK_THIRD_ORDER && !F)
    ADM_HIP(hipMalloc(&F, sizeof(double) * n * 3));
}

StepBuffers::~StepBuffers() {
  for (double *q : {R, Y, acc, F, inv_mass})
    if (q) (void)hipFree(q);
}

// One evolve_one_time_step (source/ThermalPhysics.templates.hh:854-888) on
// the device, enqueued on s. rhs(t_stage, x, r) sets the time and adds A(x)
// into the cleared residual r; the vector kernels then apply the inverse
// mass and deal.II's stage combinations (forward Euler as sadd; classic RK4
// and RK3 as TimeStepping::ExplicitRungeKutta orders them).
//
// fused(t_stage, x, st), when given, runs the operator on x and applies the
// stage update st in the plane path's gather pass (the same arithmetic as
// the stage kernels, without the residual vector's round trip).
template <class Rhs, class Fused>
static void explicit_step(int method, double t, double dt, double *T,
                          StepBuffers &b, Rhs &&rhs, Fused &&fused,
                          bool use_fused, hipStream_t s) {
  const int block = 256;
  const int64_t n = b.n;
  const unsigned grid = (unsigned)((n + block - 1) / block);
  StageArgs st;
  std::memset(&st, 0, sizeof(st));
  st.y = T;
  st.inv_mass = b.inv_mass;
  st.n = n;
  if (method == ADM_FORWARD_EULER) {
    if (use_fused) {
      st.kind = kStageFE;
      st.c0 = dt;
      fused(t, T, st);
    } else {
      rhs(t, T, b.R);
      hipLaunchKernelGGL(fe_update_kernel, dim3(grid), dim3(block), 0, s, T,
                         b.R, b.inv_mass, dt, n);
    }
  } else if (method == ADM_RK_FOURTH_ORDER) {
    static const double bw[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    static const double c[4] = {0.0, 0.5, 0.5, 1.0};
    static const double anext[4] = {0.5, 0.5, 1.0, 0.0};
    for (int i = 0; i < 4; ++i) {
      if (use_fused) {
        st.kind = kStageRK4;
        st.acc = b.acc;
        st.Y = b.Y;
        st.c0 = dt * bw[i];
        st.c1 = dt * anext[i];
        st.first = i == 0 ? 1 : 0;
        st.last = i == 3 ? 1 : 0;
        fused(t + c[i] * dt, i == 0 ? T : b.Y, st);
        continue;
      }
      rhs(t + c[i] * dt, i == 0 ? T : b.Y, b.R);
      hipLaunchKernelGGL(rk4_stage_kernel, dim3(grid), dim3(block), 0, s, T,
                         b.R, b.inv_mass, b.acc, b.Y, dt * bw[i],
                         dt * anext[i], i == 0 ? 1 : 0, i == 3 ? 1 : 0, n);
    }
  } else {
    static const double a[3][3] = {{0, 0, 0}, {0.5, 0, 0}, {-1.0, 2.0, 0}};
    static const double bw[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
    static const double c[3] = {0.0, 0.5, 1.0};
    for (int i = 0; i < 3; ++i) {
      const double *an = i + 1 < 3 ? a[i + 1] : a[0];
      if (use_fused) {
        st.kind = kStageRK;
        st.F = b.F;
        st.Y = b.Y;
        st.stage = i;
        st.n_stages = 3;
        for (int j = 0; j < 3; ++j) {
          st.a[j] = dt * an[j];
          st.b[j] = dt * bw[j];
        }
        fused(t + c[i] * dt, i == 0 ? T : b.Y, st);
        continue;
      }
      rhs(t + c[i] * dt, i == 0 ? T : b.Y, b.R);
      hipLaunchKernelGGL(rk_stage_kernel, dim3(grid), dim3(block), 0, s, T,
                         b.R, b.inv_mass, b.F, i, 3, dt * an[0], dt * an[1],
                         dt * an[2], b.Y, dt * bw[0], dt * bw[1], dt * bw[2],
                         n);
    }
  }
  ADM_HIP(hipGetLastError());
}

void Context::ensure_step_buffers(int method) {
  sb.ensure(method, n_dofs);
  if (!sb.inv_mass) {
    ADM_HIP(hipMalloc(&sb.inv_mass, sizeof(double) * n_dofs));
    compute_inverse_mass(sb.inv_mass, nullptr);
    ADM_HIP(hipDeviceSynchronize());
  }
}

void Context::step(int method, double t, double dt, double *T, hipStream_t s) {
  explicit_step(
      method, t, dt, T, sb,
      [&](double ts, const double *x, double *r) {
        set_time(ts);
        apply(x, r, 0, s);
      },
      [&](double ts, const double *x, const StageArgs &st) {
        set_time(ts);
        CellArgs a = cell_args();
        a.add = 0;
        launch(a, face_args(), x, nullptr, 1, s, nz_active, nullptr, &st);
      },
      plane_path(), s);
}

void Context::step_hooked(
    int method, double t, double dt, double *T, hipStream_t s,
    const std::function<void(int, const double *)> &hook) {
  int stage = 0;
  auto at = [&](double ts, const double *x) {
    if (hook) hook(stage, x);
    ++stage;
    for (size_t b = 0; b < old_segment.size(); ++b) old_segment[b] = 0;
    set_time(ts);
  };
  explicit_step(
      method, t, dt, T, sb,
      [&](double ts, const double *x, double *r) {
        at(ts, x);
        apply(x, r, 0, s);
      },
      [&](double ts, const double *x, const StageArgs &st) {
        at(ts, x);
        CellArgs a = cell_args();
        a.add = 0;
        launch(a, face_args(), x, nullptr, 1, s, nz_active, nullptr, &st);
      },
      plane_path(), s);
}

// Stage time fractions of explicit_step's methods.
static int n_stages(int method) {
  return method == ADM_FORWARD_EULER ? 1
                                     : (method == ADM_RK_FOURTH_ORDER ? 4 : 3);
}
static double stage_c(int method, int i) {
  static const double c4[4] = {0.0, 0.5, 0.5, 1.0};
  static const double c3[3] = {0.0, 0.5, 1.0};
  if (method == ADM_FORWARD_EULER) return 0.0;
  return method == ADM_RK_FOURTH_ORDER ? c4[i] : c3[i];
}

// The step loop as a HIP graph: one step is captured on a private stream
// and replayed n_steps times. Only the beam states change between steps;
// they are evaluated on the host for every stage beforehand (the same scan
// path search, in the same order, as the direct loop), uploaded as a table,
// and the cell kernel reads row *d_slot of it, a counter that a one-lane
// kernel advances after every operator application. No memset or copy is
// captured.
void Context::evolve_graph(int method, double t, double dt, int n_steps,
                           double *T, hipStream_t s) {
  const int ns = n_stages(method);
  const size_t rows = (size_t)n_steps * ns;
  if (rows > (size_t)INT32_MAX)
    throw AdmError(ADM_ERR_INVALID, "too many steps for one captured loop");
  ADM_HIP(hipStreamSynchronize(s));
  if (rows > beam_table_rows) {
    ADM_FREE(d_beam_table);
    d_beam_table = nullptr;
    beam_table_rows = 0;
    ADM_HIP(hipMalloc(&d_beam_table,
                      sizeof(adm_beam_state) * ADM_MAX_BEAMS * rows));
    beam_table_rows = rows;
  }
  if (!d_slot) ADM_HIP(hipMalloc(&d_slot, sizeof(int32_t)));
  // beam states of every stage, from the current scan-path position on
  std::vector<adm_beam_state> table(rows * ADM_MAX_BEAMS);
  std::memset(table.data(), 0, sizeof(adm_beam_state) * table.size());
  double tk = t;
  for (int k = 0; k < n_steps; ++k) {
    for (int i = 0; i < ns; ++i) {
      set_time(tk + stage_c(method, i) * dt);
      for (size_t b = 0; b < beams.size(); ++b)
        table[((size_t)k * ns + i) * ADM_MAX_BEAMS + b] = beam_states[b];
    }
    tk = tk + dt;
  }
  ADM_HIP(hipMemcpy(d_beam_table, table.data(),
                    sizeof(adm_beam_state) * table.size(),
                    hipMemcpyHostToDevice));
  // capture one step; the kernels take their beams from the table (and
  // the plane kernel's patch buffer is allocated now, not in the capture)
  if (plane_path()) ensure_patch_buffer(patch_buffer_size(nz_active, 1));
  hipStream_t cs;
  ADM_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const int32_t len = (int32_t)rows;
  hipError_t cap = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (cap != hipSuccess) {
    (void)hipStreamDestroy(cs);
    throw AdmError(ADM_ERR_HIP, std::string("graph capture: ") +
                                    hipGetErrorString(cap));
  }
  try {
    explicit_step(
        method, t, dt, T, sb,
        [&](double, const double *x, double *r) {
          CellArgs a = cell_args();
          a.beam_table = d_beam_table;
          a.beam_slot = d_slot;
          a.beam_table_len = len;
          launch(a, face_args(), x, r, 1, cs, nz_active);
          hipLaunchKernelGGL(advance_slot_kernel, dim3(1), dim3(64), 0, cs,
                             d_slot);
        },
        [&](double, const double *x, const StageArgs &st) {
          // the plane path's fused stage update (its arguments are the
          // same every step)
          CellArgs a = cell_args();
          a.add = 0;
          a.beam_table = d_beam_table;
          a.beam_slot = d_slot;
          a.beam_table_len = len;
          launch(a, face_args(), x, nullptr, 1, cs, nz_active, nullptr, &st);
          hipLaunchKernelGGL(advance_slot_kernel, dim3(1), dim3(64), 0, cs,
                             d_slot);
        },
        plane_path(), cs);
  } catch (...) {
    // end the capture before the stream goes away
    if (hipStreamEndCapture(cs, &graph) == hipSuccess && graph)
      (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cs);
    throw;
  }
  cap = hipStreamEndCapture(cs, &graph);
  if (cap != hipSuccess) {
    (void)hipStreamDestroy(cs);
    throw AdmError(ADM_ERR_HIP, std::string("graph capture: ") +
                                    hipGetErrorString(cap));
  }
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(set_slot_kernel, dim3(1), dim3(64), 0, cs, d_slot, 0);
    for (int k = 0; k < n_steps && e == hipSuccess; ++k)
      e = hipGraphLaunch(exec, cs);
    if (e == hipSuccess) e = hipStreamSynchronize(cs);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess)
    throw AdmError(ADM_ERR_HIP, std::string("graph replay: ") +
                                    hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Ensemble
// ---------------------------------------------------------------------------

Ensemble::Ensemble(const adm_problem_desc &base, int n,
                   const adm_member_desc *md) {
  if (n < 1) throw AdmError(ADM_ERR_INVALID, "n_members < 1");
  if (!md) throw AdmError(ADM_ERR_INVALID, "null member descriptors");
  n_members = n;
  for (int j = 0; j < n; ++j) {
    adm_problem_desc d = base;
    const adm_member_desc &m = md[j];
    for (int b = 0; b < base.n_beams && b < ADM_MAX_BEAMS; ++b) {
      d.beams[b].absorption_efficiency = m.absorption_efficiency[b];
      d.beams[b].max_power = m.max_power[b];
    }
    if (!(m.conductivity_scale > 0.0))
      throw AdmError(ADM_ERR_INVALID, "conductivity_scale must be > 0");
    if (m.conductivity_scale != 1.0) {
      adm_material_desc &mat = d.material;
      for (int mm = 0; mm < ADM_MAX_MATERIALS; ++mm)
        for (int pr = ADM_THERMAL_CONDUCTIVITY_X;
             pr <= ADM_THERMAL_CONDUCTIVITY_Z; ++pr)
          for (int st = 0; st < ADM_MAX_STATES; ++st) {
            for (int i = 0; i < ADM_TABLE_SIZE; ++i)
              mat.table[mm][pr][st][i][1] *= m.conductivity_scale;
            for (int i = 0; i < ADM_MAX_POLY; ++i)
              mat.poly[mm][pr][st][i] *= m.conductivity_scale;
          }
    }
    if (m.cell_active) d.cell_active = m.cell_active;
    if (m.cell_state) d.cell_state = m.cell_state;
    members.emplace_back(new Context(d));
  }
  n_dofs = members[0]->n_dofs;
  auto work = std::make_shared<ActWork>();
  for (auto &m : members) m->act = work;
  refresh_members();
}

// The device table of per-member pointers (rebuilt after a member's cells,
// faces or powder storage change).
void Ensemble::refresh_members() {
  const int n = n_members;
  const Context &c0 = *members[0];
  std::vector<MemberArgs> h(n);
  max_faces = 0;
  max_fnodes = 0;
  for (int j = 0; j < n; ++j) {
    const Context &c = *members[j];
    if (c.property_kind() != c0.property_kind())
      throw AdmError(ADM_ERR_INVALID,
                     "members differ in the form of their material tables");
    const CellArgs a = c.cell_args();
    MemberArgs &m = h[j];
    std::memset(&m, 0, sizeof(m));
    m.meta = a.meta;
    m.powder_slot = a.powder_slot;
    m.powder = a.powder;
    m.face_powder = a.face_powder;
    m.liquid = a.liquid;
    m.liquid_dirty = a.liquid_dirty;
    m.dep_cos = a.dep_cos;
    m.dep_sin = a.dep_sin;
    m.kc = a.kc;
    m.faces = c.d_faces;
    m.n_faces = c.n_faces;
    max_faces = std::max(max_faces, m.n_faces);
    m.face_out = c.d_face_out;
    m.fn_node = c.d_fn_node;
    m.fn_offset = c.d_fn_offset;
    m.fn_slot = c.d_fn_slot;
    m.n_fnodes = c.n_fnodes;
    m.fn_rank = c.d_fn_rank;
    max_fnodes = std::max(max_fnodes, m.n_fnodes);
    for (size_t b = 0; b < c.beams.size(); ++b)
      m.beam_gain[b] =
          c.beams[b].desc.absorption_efficiency * c.beams[b].desc.max_power;
  }
  if (!d_members) ADM_HIP(hipMalloc(&d_members, sizeof(MemberArgs) * n));
  ADM_HIP(hipMemcpy(d_members, h.data(), sizeof(MemberArgs) * n,
                    hipMemcpyHostToDevice));
}

Ensemble::~Ensemble() {
  if (d_members) (void)hipFree(d_members);
  if (d_beam_table) (void)hipFree(d_beam_table);
  if (d_pd_table) (void)hipFree(d_pd_table);
  if (d_slot) (void)hipFree(d_slot);
  if (d_da) (void)hipFree(d_da);
  if (h_da) (void)hipHostFree(h_da);
  if (da_done) (void)hipEventDestroy(da_done);
}

double *Ensemble::da_device(size_t n) {
  if (n > da_cap) {
    if (d_da) {
      ADM_HIP(hipDeviceSynchronize());
      ADM_HIP(hipFree(d_da));
      d_da = nullptr;
    }
    const size_t cap = std::max(n, da_cap * 2);
    ADM_HIP(hipMalloc(&d_da, sizeof(double) * cap));
    da_cap = cap;
  }
  return d_da;
}

// Pinned staging, free to write: waits for the last copy that read it.
double *Ensemble::da_host(size_t n) {
  if (!da_done) ADM_HIP(hipEventCreateWithFlags(&da_done, hipEventDisableTiming));
  ADM_HIP(hipEventSynchronize(da_done));
  if (n > h_da_cap) {
    if (h_da) ADM_HIP(hipHostFree(h_da));
    h_da = nullptr;
    const size_t cap = std::max(n, h_da_cap * 2);
    ADM_HIP(hipHostMalloc(&h_da, sizeof(double) * cap, hipHostMallocDefault));
    h_da_cap = cap;
  }
  return h_da;
}

void Ensemble::set_time(double t) {
  // The scan paths, radii and depths are shared, so the beam positions and
  // power modifiers are member 0's; the kernels rebuild each member's alpha.
  Context &c0 = *members[0];
  c0.sources_on = false;
  for (size_t b = 0; b < c0.beams.size(); ++b) {
    beam_update(c0.beams[b], t, c0.old_segment[b], c0.beam_states[b],
                &pmod[b], &den[b]);
    if (c0.beam_states[b].on) c0.sources_on = true;
  }
}

CellArgs Ensemble::cell_args(int add) const {
  CellArgs a = members[0]->cell_args();
  a.add = add;
  a.members = d_members;
  a.member_stride = n_dofs;
  for (int b = 0; b < ADM_MAX_BEAMS; ++b) {
    a.beam_pmod[b] = pmod[b];
    a.beam_den[b] = den[b];
  }
  return a;
}

FaceArgs Ensemble::face_args() const {
  FaceArgs fa = members[0]->face_args();
  fa.members = d_members;
  fa.member_stride = n_dofs;
  fa.n_faces = max_faces;
  fa.n_fnodes = max_fnodes;
  return fa;
}

int Ensemble::nz_active() const {
  int nz = 0;
  for (const auto &m : members) nz = std::max(nz, m->nz_active);
  return nz;
}

void Ensemble::apply(const double *src, double *dst, int add, hipStream_t s,
                     const StageArgs *st) {
  members[0]->launch(cell_args(add), face_args(), src, dst,
                     (unsigned)n_members, s, nz_active(), nullptr, st);
  ADM_HIP(hipGetLastError());
}

void Ensemble::compute_inverse_mass(double *out, hipStream_t s) {
  for (int j = 0; j < n_members; ++j)
    members[j]->compute_inverse_mass(out + (int64_t)j * n_dofs, s);
}

void Ensemble::ensure_step_buffers(int method) {
  sb.ensure(method, n_dofs * n_members);
  if (!sb.inv_mass) {
    ADM_HIP(hipMalloc(&sb.inv_mass, sizeof(double) * n_dofs * n_members));
    compute_inverse_mass(sb.inv_mass, nullptr);
    ADM_HIP(hipDeviceSynchronize());
  }
}

void Ensemble::step(int method, double t, double dt, double *T,
                    hipStream_t s) {
  explicit_step(
      method, t, dt, T, sb,
      [&](double ts, const double *x, double *r) {
        set_time(ts);
        apply(x, r, 0, s);
      },
      [&](double ts, const double *x, const StageArgs &st) {
        set_time(ts);
        apply(x, nullptr, 0, s, &st);
      },
      members[0]->plane_path(), s);
}

// Context::evolve_graph for every member at once: the shared beam states
// and, for each member's alpha, the shared power modifiers and
// denominators of every stage come from tables the host fills beforehand.
void Ensemble::evolve_graph(int method, double t, double dt, int n_steps,
                            double *T, hipStream_t s) {
  Context &c0 = *members[0];
  const int ns = n_stages(method);
  const size_t rows = (size_t)n_steps * ns;
  if (rows > (size_t)INT32_MAX)
    throw AdmError(ADM_ERR_INVALID, "too many steps for one captured loop");
  ADM_HIP(hipStreamSynchronize(s));
  if (rows > beam_table_rows) {
    ADM_FREE(d_beam_table);
    ADM_FREE(d_pd_table);
    d_beam_table = nullptr;
    d_pd_table = nullptr;
    beam_table_rows = 0;
    ADM_HIP(hipMalloc(&d_beam_table,
                      sizeof(adm_beam_state) * ADM_MAX_BEAMS * rows));
    ADM_HIP(hipMalloc(&d_pd_table, sizeof(double) * 2 * ADM_MAX_BEAMS * rows));
    beam_table_rows = rows;
  }
  if (!d_slot) ADM_HIP(hipMalloc(&d_slot, sizeof(int32_t)));
  std::vector<adm_beam_state> table(rows * ADM_MAX_BEAMS);
  std::memset(table.data(), 0, sizeof(adm_beam_state) * table.size());
  std::vector<double> pd(rows * 2 * ADM_MAX_BEAMS, 0.0);
  double tk = t;
  for (int k = 0; k < n_steps; ++k) {
    for (int i = 0; i < ns; ++i) {
      set_time(tk + stage_c(method, i) * dt);
      const size_t row = (size_t)k * ns + i;
      for (size_t b = 0; b < c0.beams.size(); ++b)
        table[row * ADM_MAX_BEAMS + b] = c0.beam_states[b];
      for (int b = 0; b < ADM_MAX_BEAMS; ++b) {
        pd[row * 2 * ADM_MAX_BEAMS + b] = pmod[b];
        pd[row * 2 * ADM_MAX_BEAMS + ADM_MAX_BEAMS + b] = den[b];
      }
    }
    tk = tk + dt;
  }
  ADM_HIP(hipMemcpy(d_beam_table, table.data(),
                    sizeof(adm_beam_state) * table.size(),
                    hipMemcpyHostToDevice));
  ADM_HIP(hipMemcpy(d_pd_table, pd.data(), sizeof(double) * pd.size(),
                    hipMemcpyHostToDevice));
  const int nz = nz_active();
  if (c0.plane_path())
    c0.ensure_patch_buffer(c0.patch_buffer_size(nz, (unsigned)n_members));
  const int32_t len = (int32_t)rows;
  auto table_args = [&](int add) {
    CellArgs a = cell_args(add);
    a.beam_table = d_beam_table;
    a.beam_slot = d_slot;
    a.beam_table_len = len;
    a.beam_pd_table = d_pd_table;
    return a;
  };
  hipStream_t cs;
  ADM_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipError_t cap = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (cap != hipSuccess) {
    (void)hipStreamDestroy(cs);
    throw AdmError(ADM_ERR_HIP, std::string("graph capture: ") +
                                    hipGetErrorString(cap));
  }
  try {
    explicit_step(
        method, t, dt, T, sb,
        [&](double, const double *x, double *r) {
          c0.launch(table_args(0), face_args(), x, r, (unsigned)n_members, cs,
                    nz);
          hipLaunchKernelGGL(advance_slot_kernel, dim3(1), dim3(64), 0, cs,
                             d_slot);
        },
        [&](double, const double *x, const StageArgs &st) {
          c0.launch(table_args(0), face_args(), x, nullptr,
                    (unsigned)n_members, cs, nz, nullptr, &st);
          hipLaunchKernelGGL(advance_slot_kernel, dim3(1), dim3(64), 0, cs,
                             d_slot);
        },
        c0.plane_path(), cs);
  } catch (...) {
    if (hipStreamEndCapture(cs, &graph) == hipSuccess && graph)
      (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cs);
    throw;
  }
  cap = hipStreamEndCapture(cs, &graph);
  if (cap != hipSuccess) {
    (void)hipStreamDestroy(cs);
    throw AdmError(ADM_ERR_HIP, std::string("graph capture: ") +
                                    hipGetErrorString(cap));
  }
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(set_slot_kernel, dim3(1), dim3(64), 0, cs, d_slot, 0);
    for (int k = 0; k < n_steps && e == hipSuccess; ++k)
      e = hipGraphLaunch(exec, cs);
    if (e == hipSuccess) e = hipStreamSynchronize(cs);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess)
    throw AdmError(ADM_ERR_HIP, std::string("graph replay: ") +
                                    hipGetErrorString(e));
}

void Ensemble::step_hooked(
    int method, double t, double dt, double *T, hipStream_t s,
    const std::function<void(int, const double *)> &hook) {
  int stage = 0;
  Context &c0 = *members[0];
  auto at = [&](double ts, const double *x) {
    if (hook) hook(stage, x);
    ++stage;
    for (size_t b = 0; b < c0.old_segment.size(); ++b) c0.old_segment[b] = 0;
    set_time(ts);
  };
  explicit_step(
      method, t, dt, T, sb,
      [&](double ts, const double *x, double *r) {
        at(ts, x);
        apply(x, r, 0, s);
      },
      [&](double ts, const double *x, const StageArgs &st) {
        at(ts, x);
        apply(x, nullptr, 0, s, &st);
      },
      c0.plane_path(), s);
}

}  // namespace adm

using adm::AdmError;
using adm::Context;

#define ADM_API_BEGIN try {
#define ADM_API_END                                              \
  }                                                              \
  catch (const AdmError &e) {                                    \
    return adm::set_error(e.code, e.what());                     \
  }                                                              \
  catch (const std::exception &e) {                              \
    return adm::set_error(ADM_ERR_INVALID, e.what());            \
  }                                                              \
  return ADM_OK;

#define ADM_CTX(ptr)                                              \
  if (!(ptr)) return adm::set_error(ADM_ERR_STATE, "null context"); \
  Context &ctx = *static_cast<Context *>(ptr);

extern "C" {

int adm_create(const adm_problem_desc *desc, void **out) {
  ADM_API_BEGIN
  if (!desc || !out) return adm::set_error(ADM_ERR_INVALID, "null argument");
  *out = new Context(*desc);
  ADM_API_END
}

int adm_destroy(void *ptr) {
  ADM_API_BEGIN
  delete static_cast<Context *>(ptr);
  ADM_API_END
}

const char *adm_last_error(void) { return adm::g_last_error.c_str(); }

int adm_n_dofs(const void *ptr, int64_t *n) {
  ADM_API_BEGIN
  if (!ptr || !n) return adm::set_error(ADM_ERR_INVALID, "null argument");
  const Context &c = *static_cast<const Context *>(ptr);
  *n = c.n_ext >= 0 ? c.n_ext : c.n_dofs;
  ADM_API_END
}

int adm_n_nodes(const void *ptr, int64_t *n) {
  ADM_API_BEGIN
  if (!ptr || !n) return adm::set_error(ADM_ERR_INVALID, "null argument");
  *n = static_cast<const Context *>(ptr)->n_dofs;
  ADM_API_END
}

int adm_set_dof_map(void *ptr, int64_t n_dofs, const int64_t *dof_to_node) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (n_dofs < 0) {
    ADM_HIP(hipDeviceSynchronize());
    ctx.clear_dof_map();
  } else {
    ctx.set_dof_map(n_dofs, dof_to_node);
  }
  ADM_API_END
}

int adm_sizes(const void *ptr, int64_t *n_cells, int32_t *n_q, int32_t *n_fq) {
  ADM_API_BEGIN
  if (!ptr) return adm::set_error(ADM_ERR_INVALID, "null argument");
  const Context &c = *static_cast<const Context *>(ptr);
  if (n_cells) *n_cells = c.n_cells;
  if (n_q) *n_q = c.nq;
  if (n_fq) *n_fq = c.nfq;
  ADM_API_END
}

int adm_compute_inverse_mass(void *ptr, double *inv_mass, void *stream) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!inv_mass) return adm::set_error(ADM_ERR_INVALID, "null output");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (ctx.n_ext >= 0) {
    ctx.compute_inverse_mass(ctx.d_map_dst, s);
    ctx.to_dofs(ctx.d_map_dst, inv_mass, 0, s);
  } else {
    ctx.compute_inverse_mass(inv_mass, s);
  }
  ADM_API_END
}

int adm_set_time(void *ptr, double t) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  ctx.set_time(t);
  ADM_API_END
}

int adm_vmult(void *ptr, const double *src, double *dst, int add,
              void *stream) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!src || !dst) return adm::set_error(ADM_ERR_INVALID, "null vector");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (ctx.n_ext >= 0) {
    // vectors in the caller's DoF order: to box nodes, apply, back
    ctx.to_nodes(src, ctx.d_map_src, s);
    if (!ctx.plane_path()) ctx.fill(ctx.d_map_dst, 0.0, ctx.n_dofs, s);
    ctx.apply(ctx.d_map_src, ctx.d_map_dst, 0, s);
    ctx.to_dofs(ctx.d_map_dst, dst, add, s);
  } else {
    // the line kernel accumulates into dst: start from zero (the plane
    // kernel's gather pass writes every node)
    if (!add && !ctx.plane_path()) ctx.fill(dst, 0.0, ctx.n_dofs, s);
    ctx.apply(src, dst, add, s);
  }
  ADM_API_END
}

// checks a launch configuration for a context of FE degree p
static int check_launch_config(const adm::Context &c, int32_t mode,
                               int32_t waves) {
  const int p = c.p;
  const int64_t n_dofs = c.n_dofs;
  if (waves < 2 || waves > 4)
    return adm::set_error(ADM_ERR_INVALID, "waves_per_block must be 2..4");
  if (mode < ADM_LAUNCH_LINE_ATOMIC || mode > ADM_LAUNCH_PLANE)
    return adm::set_error(ADM_ERR_INVALID, "unknown launch mode");
  if (mode == ADM_LAUNCH_PLANE && p > 2)
    return adm::set_error(ADM_ERR_UNSUPPORTED,
                          "the plane kernel exists for FE degrees 1 and 2");
  if (mode == ADM_LAUNCH_PLANE && n_dofs >= (int64_t)INT32_MAX)
    return adm::set_error(ADM_ERR_UNSUPPORTED,
                          "the plane kernel takes fewer than 2^31 nodes");
  if (mode == ADM_LAUNCH_PLANE && !c.plane_tables_ok)
    return adm::set_error(ADM_ERR_UNSUPPORTED,
                          "the plane kernel needs the symmetric Q2 tables");
  if (mode == ADM_LAUNCH_PLANE && c.dim == 2)
    return adm::set_error(ADM_ERR_UNSUPPORTED,
                          "the plane kernel is three-dimensional");
  return ADM_OK;
}

int adm_vmult_noalloc(void *ptr, const double *src, double *dst, int add,
                      void *stream) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!src || !dst) return adm::set_error(ADM_ERR_INVALID, "null vector");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // as adm_vmult, with the scratch array as the patch buffer
  if (ctx.n_ext >= 0) {
    ctx.to_nodes(src, ctx.d_map_src, s);
    if (!ctx.plane_path()) ctx.fill(ctx.d_map_dst, 0.0, ctx.n_dofs, s);
    ctx.apply(ctx.d_map_src, ctx.d_map_dst, 0, s, ctx.d_scratch);
    ctx.to_dofs(ctx.d_map_dst, dst, add, s);
  } else {
    if (!add && !ctx.plane_path()) ctx.fill(dst, 0.0, ctx.n_dofs, s);
    ctx.apply(src, dst, add, s, ctx.d_scratch);
  }
  ADM_API_END
}

int adm_set_launch_config(void *ptr, int32_t mode, int32_t waves_per_block) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (int e = check_launch_config(ctx, mode, waves_per_block))
    return e;
  ADM_HIP(hipDeviceSynchronize());
  ctx.set_launch_mode(mode);
  ctx.cell_waves = waves_per_block;
  ADM_API_END
}

// Times `apply` (one operator application into a cleared dst) under every
// launch configuration of `cfg`, keeps the fastest in cfg, and puts back
// the evaluation state of every context in `holders` after each trial.
extern "C++" {
template <class Apply>
static void tune_launch(adm::Context &cfg,
                        const std::vector<adm::Context *> &holders,
                        Apply apply, int reps, hipStream_t s,
                        int32_t *mode_out, int32_t *waves_out,
                        double *ms_out) {
  using namespace adm;
  ADM_HIP(hipStreamSynchronize(s));
  struct Saved {
    void *dev;
    void *copy;
    size_t bytes;
  };
  std::vector<Saved> saved;
  auto keep = [&](void *p, size_t bytes) {
    if (!p || !bytes) return;
    void *c = nullptr;
    ADM_HIP(hipMalloc(&c, bytes));
    ADM_HIP(hipMemcpy(c, p, bytes, hipMemcpyDeviceToDevice));
    saved.push_back({p, c, bytes});
  };
  for (Context *h : holders) {
    keep(h->d_liquid, sizeof(double) * h->n_cells * h->nq);
    keep(h->d_liquid_dirty, (size_t)h->n_cells);
    keep(h->d_powder, sizeof(double) * h->n_powder_cells * h->nq);
    keep(h->d_face_powder, sizeof(double) * h->n_faces * h->nfq);
  }
  auto restore = [&]() {
    for (const Saved &v : saved)
      ADM_HIP(hipMemcpy(v.dev, v.copy, v.bytes, hipMemcpyDeviceToDevice));
  };
  hipEvent_t e0, e1;
  ADM_HIP(hipEventCreate(&e0));
  ADM_HIP(hipEventCreate(&e1));
  double best = 1e300;
  int best_mode = cfg.launch_mode(), best_w = cfg.cell_waves, k = 0;
  const int modes[3] = {ADM_LAUNCH_LINE_COLORED, ADM_LAUNCH_LINE_ATOMIC,
                        ADM_LAUNCH_PLANE};
  for (int mi = 0; mi < 3; ++mi)
    for (int w = 2; w <= 4; ++w, ++k) {
      const int mode = modes[mi];
      if (mode == ADM_LAUNCH_PLANE &&
          (cfg.p > 2 || cfg.n_dofs >= (int64_t)INT32_MAX ||
           !cfg.plane_tables_ok || cfg.dim == 2)) {
        if (ms_out) ms_out[k] = -1.0;
        continue;
      }
      cfg.set_launch_mode(mode);
      cfg.cell_waves = w;
      apply();  // warm-up
      ADM_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; ++r) apply();
      ADM_HIP(hipEventRecord(e1, s));
      ADM_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      ADM_HIP(hipEventElapsedTime(&ms, e0, e1));
      const double per = (double)ms / reps;
      if (ms_out) ms_out[k] = per;
      if (per < best) {
        best = per;
        best_mode = mode;
        best_w = w;
      }
      restore();
    }
  ADM_HIP(hipEventDestroy(e0));
  ADM_HIP(hipEventDestroy(e1));
  for (const Saved &v : saved) ADM_HIP(hipFree(v.copy));
  cfg.set_launch_mode(best_mode);
  cfg.cell_waves = best_w;
  if (mode_out) *mode_out = best_mode;
  if (waves_out) *waves_out = best_w;
}
}  // extern "C++"

int adm_autotune(void *ptr, const double *src, double *dst, int32_t reps,
                 void *stream, int32_t *mode_out, int32_t *waves_out,
                 double *ms_out) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!src || !dst || reps < 1)
    return adm::set_error(ADM_ERR_INVALID, "null vector or reps < 1");
  hipStream_t s = static_cast<hipStream_t>(stream);
  tune_launch(
      ctx, {&ctx},
      [&]() {
        if (!ctx.plane_path()) ctx.fill(dst, 0.0, ctx.n_dofs, s);
        ctx.apply(src, dst, 0, s);
      },
      reps, s, mode_out, waves_out, ms_out);
  ADM_API_END
}

int adm_evolve(void *ptr, int method, double t, double dt, int n_steps,
               double *T, int use_graph, double *t_end, void *stream) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!T) return adm::set_error(ADM_ERR_INVALID, "null solution vector");
  if (method < ADM_FORWARD_EULER || method > ADM_RK_FOURTH_ORDER)
    return adm::set_error(ADM_ERR_INVALID, "unknown time stepping method");
  if (n_steps < 0) return adm::set_error(ADM_ERR_INVALID, "n_steps < 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  ctx.ensure_step_buffers(method);
  // with a DoF map the solution is in the caller's DoF order: one gather
  // into the box-node vector before the steps and one scatter after them
  double *T_ext = nullptr;
  if (ctx.n_ext >= 0) {
    T_ext = T;
    T = ctx.d_map_src;
    ctx.to_nodes(T_ext, T, s);
  }
  if (use_graph && n_steps > 0) {
    ctx.evolve_graph(method, t, dt, n_steps, T, s);
    for (int k = 0; k < n_steps; ++k) t = t + dt;
  } else {
    for (int k = 0; k < n_steps; ++k) {
      ctx.step(method, t, dt, T, s);
      t = t + dt;
    }
  }
  if (T_ext) ctx.to_dofs(T, T_ext, 0, s);
  if (t_end) *t_end = t;
  ADM_API_END
}

int adm_set_cell_state(void *ptr, const double *cell_state) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!cell_state) return adm::set_error(ADM_ERR_INVALID, "null state");
  ADM_HIP(hipDeviceSynchronize());
  ctx.set_cell_state(cell_state);
  ADM_API_END
}

int adm_get_cell_state(void *ptr, double *cs) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!cs) return adm::set_error(ADM_ERR_INVALID, "null state");
  ctx.cell_state_means(cs);
  ADM_API_END
}

int adm_get_cell_data(void *ptr, uint8_t *active, double *cos_,
                      double *sin_) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  for (int64_t c = 0; c < ctx.n_cells; ++c) {
    if (active) active[c] = ctx.active[c] ? 1 : 0;
    if (cos_) cos_[c] = ctx.dep_cos[c];
    if (sin_) sin_[c] = ctx.dep_sin[c];
  }
  ADM_API_END
}

int adm_get_qp_state(void *ptr, double *liquid, double *powder) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  ctx.get_qp_state(liquid, powder);
  ADM_API_END
}

int adm_set_qp_state(void *ptr, const double *liquid, const double *powder) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  ADM_HIP(hipDeviceSynchronize());
  ctx.set_qp_state(liquid, powder);
  ADM_API_END
}

int adm_get_face_state(void *ptr, double *face_powder) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!face_powder) return adm::set_error(ADM_ERR_INVALID, "null output");
  ctx.sync_faces_host();
  const int nf = ctx.nfq, nfc = 2 * ctx.dim;
  std::fill(face_powder, face_powder + (size_t)ctx.n_cells * nfc * nf, 0.0);
  if (ctx.d_face_powder && !ctx.faces.empty()) {
    ADM_HIP(hipDeviceSynchronize());
    std::vector<double> fpw(ctx.faces.size() * nf);
    ADM_HIP(hipMemcpy(fpw.data(), ctx.d_face_powder,
                      sizeof(double) * fpw.size(), hipMemcpyDeviceToHost));
    for (size_t f = 0; f < ctx.faces.size(); ++f) {
      const int64_t c = ctx.faces[f].cell;
      const int fid = ctx.face_id(ctx.faces[f].info & 7);
      for (int q = 0; q < nf; ++q)
        face_powder[(c * nfc + fid) * nf + q] = fpw[f * nf + q];
    }
  }
  ADM_API_END
}

int adm_set_face_state(void *ptr, const double *face_powder) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!face_powder) return adm::set_error(ADM_ERR_INVALID, "null input");
  ctx.sync_faces_host();
  const int nf = ctx.nfq, nfc = 2 * ctx.dim;
  if (ctx.d_face_powder && !ctx.faces.empty()) {
    std::vector<double> fpw(ctx.faces.size() * nf);
    for (size_t f = 0; f < ctx.faces.size(); ++f) {
      const int64_t c = ctx.faces[f].cell;
      const int fid = ctx.face_id(ctx.faces[f].info & 7);
      for (int q = 0; q < nf; ++q)
        fpw[f * nf + q] = face_powder[(c * nfc + fid) * nf + q];
    }
    ADM_HIP(hipDeviceSynchronize());
    ADM_HIP(hipMemcpy(ctx.d_face_powder, fpw.data(),
                      sizeof(double) * fpw.size(), hipMemcpyHostToDevice));
  }
  ADM_API_END
}

int adm_set_deposition(void *ptr, const double *c, const double *s) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!c || !s) return adm::set_error(ADM_ERR_INVALID, "null argument");
  ADM_HIP(hipDeviceSynchronize());
  std::copy(c, c + ctx.n_cells, ctx.dep_cos.begin());
  std::copy(s, s + ctx.n_cells, ctx.dep_sin.begin());
  // keep the quadrature-point state: rebuild only the meta words
  std::vector<double> lq((size_t)ctx.n_cells * ctx.nq),
      pw((size_t)ctx.n_cells * ctx.nq);
  ctx.get_qp_state(lq.data(), pw.data());
  std::vector<double> fpw;
  if (ctx.d_face_powder && !ctx.faces.empty()) {
    fpw.resize(ctx.faces.size() * ctx.nfq);
    ADM_HIP(hipMemcpy(fpw.data(), ctx.d_face_powder,
                      sizeof(double) * fpw.size(), hipMemcpyDeviceToHost));
  }
  ctx.rebuild_cells();
  ctx.set_qp_state(lq.data(), pw.data());
  if (!fpw.empty() && ctx.d_face_powder)
    ADM_HIP(hipMemcpy(ctx.d_face_powder, fpw.data(),
                      sizeof(double) * fpw.size(), hipMemcpyHostToDevice));
  ADM_API_END
}

int adm_set_beam_parameters(void *ptr, int beam, double absorption,
                            double max_power, double radius, double depth) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (beam < 0 || beam >= (int)ctx.beams.size())
    return adm::set_error(ADM_ERR_INVALID, "beam index out of range");
  auto &d = ctx.beams[beam].desc;
  d.absorption_efficiency = absorption;
  d.max_power = max_power;
  d.radius = radius;
  d.depth = depth;
  ADM_API_END
}

int adm_beam_state_at(void *ptr, int beam, double t, adm_beam_state *out) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (beam < 0 || beam >= (int)ctx.beams.size() || !out)
    return adm::set_error(ADM_ERR_INVALID, "invalid beam query");
  adm::beam_update(ctx.beams[beam], t, ctx.old_segment[beam], *out);
  ADM_API_END
}

int adm_normal_random_vector(uint32_t length, uint32_t n_rejected_draws,
                             double mean, double stddev, double *out) {
  ADM_API_BEGIN
  if (!out && length) return adm::set_error(ADM_ERR_INVALID, "null output");
  if (!(stddev >= 0.0))
    return adm::set_error(ADM_ERR_INVALID, "stddev must be >= 0");
  // the reference's own generator and distribution (default seed), so the
  // draws are those of the standard library both are built against
  std::mt19937 prng;
  std::normal_distribution<> normal(mean, stddev);
  for (uint32_t i = 0; i < n_rejected_draws; ++i) (void)normal(prng);
  for (uint32_t i = 0; i < length; ++i) {
    do {
      out[i] = normal(prng);
    } while (out[i] < 0.);
  }
  ADM_API_END
}

#define ADM_ENS(ptr)                                                \
  if (!(ptr)) return adm::set_error(ADM_ERR_STATE, "null ensemble"); \
  adm::Ensemble &ens = *static_cast<adm::Ensemble *>(ptr);

int adm_ensemble_create(const adm_problem_desc *base, int32_t n_members,
                        const adm_member_desc *members, void **out) {
  ADM_API_BEGIN
  if (!base || !members || !out)
    return adm::set_error(ADM_ERR_INVALID, "null argument");
  *out = new adm::Ensemble(*base, n_members, members);
  ADM_API_END
}

int adm_ensemble_destroy(void *ptr) {
  ADM_API_BEGIN
  delete static_cast<adm::Ensemble *>(ptr);
  ADM_API_END
}

int adm_ensemble_n_dofs(const void *ptr, int64_t *n) {
  ADM_API_BEGIN
  if (!ptr || !n) return adm::set_error(ADM_ERR_INVALID, "null argument");
  *n = static_cast<const adm::Ensemble *>(ptr)->n_dofs;
  ADM_API_END
}

int adm_ensemble_compute_inverse_mass(void *ptr, double *inv_mass,
                                      void *stream) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (!inv_mass) return adm::set_error(ADM_ERR_INVALID, "null output");
  ens.compute_inverse_mass(inv_mass, static_cast<hipStream_t>(stream));
  ADM_API_END
}

int adm_ensemble_set_time(void *ptr, double t) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  ens.set_time(t);
  ADM_API_END
}

int adm_ensemble_vmult(void *ptr, const double *src, double *dst, int add,
                       void *stream) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (!src || !dst) return adm::set_error(ADM_ERR_INVALID, "null vector");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!add && !ens.members[0]->plane_path())
    ens.members[0]->fill(dst, 0.0, ens.n_dofs * ens.n_members, s);
  ens.apply(src, dst, add, s);
  ADM_API_END
}

int adm_ensemble_set_launch_config(void *ptr, int32_t mode,
                                   int32_t waves_per_block) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (int e = check_launch_config(*ens.members[0], mode, waves_per_block))
    return e;
  ADM_HIP(hipDeviceSynchronize());
  ens.members[0]->set_launch_mode(mode);
  ens.members[0]->cell_waves = waves_per_block;
  ADM_API_END
}

int adm_ensemble_autotune(void *ptr, const double *src, double *dst,
                          int32_t reps, void *stream, int32_t *mode_out,
                          int32_t *waves_out, double *ms_out) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (!src || !dst || reps < 1)
    return adm::set_error(ADM_ERR_INVALID, "null vector or reps < 1");
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::vector<adm::Context *> holders;
  for (auto &m : ens.members) holders.push_back(m.get());
  adm::Context &c0 = *ens.members[0];
  tune_launch(
      c0, holders,
      [&]() {
        if (!c0.plane_path()) c0.fill(dst, 0.0, ens.n_dofs * ens.n_members, s);
        ens.apply(src, dst, 0, s);
      },
      reps, s, mode_out, waves_out, ms_out);
  ADM_API_END
}

int adm_ensemble_evolve(void *ptr, int method, double t, double dt,
                        int n_steps, double *T, double *t_end, void *stream) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (!T) return adm::set_error(ADM_ERR_INVALID, "null solution vector");
  if (method < ADM_FORWARD_EULER || method > ADM_RK_FOURTH_ORDER)
    return adm::set_error(ADM_ERR_INVALID, "unknown time stepping method");
  if (n_steps < 0) return adm::set_error(ADM_ERR_INVALID, "n_steps < 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  ens.ensure_step_buffers(method);
  for (int k = 0; k < n_steps; ++k) {
    ens.step(method, t, dt, T, s);
    t = t + dt;
  }
  if (t_end) *t_end = t;
  ADM_API_END
}

int adm_ensemble_evolve_graph(void *ptr, int method, double t, double dt,
                              int n_steps, double *T, double *t_end,
                              void *stream) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (!T) return adm::set_error(ADM_ERR_INVALID, "null solution vector");
  if (method < ADM_FORWARD_EULER || method > ADM_RK_FOURTH_ORDER)
    return adm::set_error(ADM_ERR_INVALID, "unknown time stepping method");
  if (n_steps < 0) return adm::set_error(ADM_ERR_INVALID, "n_steps < 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  ens.ensure_step_buffers(method);
  if (n_steps > 0) ens.evolve_graph(method, t, dt, n_steps, T, s);
  for (int k = 0; k < n_steps; ++k) t = t + dt;
  if (t_end) *t_end = t;
  ADM_API_END
}

int adm_ensemble_get_qp_state(void *ptr, int32_t member, double *liquid,
                              double *powder) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (member < 0 || member >= ens.n_members)
    return adm::set_error(ADM_ERR_INVALID, "member index out of range");
  ADM_HIP(hipDeviceSynchronize());
  ens.members[member]->get_qp_state(liquid, powder);
  ADM_API_END
}

int adm_ensemble_set_qp_state(void *ptr, int32_t member, const double *liquid,
                              const double *powder) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (member < 0 || member >= ens.n_members)
    return adm::set_error(ADM_ERR_INVALID, "member index out of range");
  ADM_HIP(hipDeviceSynchronize());
  ens.members[member]->set_qp_state(liquid, powder);
  ADM_API_END
}

int adm_ensemble_get_face_state(void *ptr, int32_t member,
                                double *face_powder) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (member < 0 || member >= ens.n_members)
    return adm::set_error(ADM_ERR_INVALID, "member index out of range");
  return adm_get_face_state(ens.members[member].get(), face_powder);
  ADM_API_END
}

int adm_ensemble_set_face_state(void *ptr, int32_t member,
                                const double *face_powder) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (member < 0 || member >= ens.n_members)
    return adm::set_error(ADM_ERR_INVALID, "member index out of range");
  return adm_set_face_state(ens.members[member].get(), face_powder);
  ADM_API_END
}

int adm_ensemble_set_beam_gain(void *ptr, int32_t member, int32_t beam,
                               double absorption_efficiency,
                               double max_power) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (member < 0 || member >= ens.n_members)
    return adm::set_error(ADM_ERR_INVALID, "member index out of range");
  adm::Context &c = *ens.members[member];
  if (beam < 0 || beam >= (int)c.beams.size())
    return adm::set_error(ADM_ERR_INVALID, "beam index out of range");
  ADM_HIP(hipDeviceSynchronize());
  c.beams[beam].desc.absorption_efficiency = absorption_efficiency;
  c.beams[beam].desc.max_power = max_power;
  ens.refresh_members();
  ADM_API_END
}

int adm_cells_to_activate(void *ptr, int32_t n_boxes, const double *boxes,
                          int32_t *box_of_cell) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (n_boxes < 0 || (n_boxes && !boxes) || !box_of_cell)
    return adm::set_error(ADM_ERR_INVALID, "bad box list");
  ctx.cells_to_activate(n_boxes, boxes, box_of_cell);
  ADM_API_END
}

int adm_add_material(void *ptr, const int32_t *box_of_cell,
                     const double *box_cos, const double *box_sin,
                     double new_temperature, double *T, void *stream) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!box_of_cell || !box_cos || !box_sin)
    return adm::set_error(ADM_ERR_INVALID, "null argument");
  if (ctx.n_ext >= 0)
    return adm::set_error(ADM_ERR_UNSUPPORTED,
                          "activation changes the DoFs: clear the DoF map, "
                          "activate on the box-node vector, set the new map");
  ctx.add_material(box_of_cell, box_cos, box_sin, new_temperature, T,
                   static_cast<hipStream_t>(stream));
  ADM_API_END
}

int adm_activate(void *ptr, int32_t n_boxes, const double *boxes,
                 const double *box_cos, const double *box_sin,
                 double new_temperature, double *T, int32_t *box_of_cell,
                 void *stream) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (n_boxes < 0 || (n_boxes && (!boxes || !box_cos || !box_sin)))
    return adm::set_error(ADM_ERR_INVALID, "bad box list");
  if (ctx.n_ext >= 0)
    return adm::set_error(ADM_ERR_UNSUPPORTED,
                          "activation changes the DoFs: clear the DoF map, "
                          "activate on the box-node vector, set the new map");
  if (ctx.dim == 2) {
    ctx.activate_host(n_boxes, boxes, box_cos, box_sin, new_temperature, T,
                      box_of_cell, static_cast<hipStream_t>(stream));
    return ADM_OK;
  }
  if (!ctx.device_activation_fits())
    return adm::set_error(ADM_ERR_UNSUPPORTED,
                          "mesh too large for the device activation");
  ctx.activate(n_boxes, boxes, nullptr, box_cos, box_sin, new_temperature, T,
               box_of_cell, static_cast<hipStream_t>(stream));
  ADM_API_END
}

int adm_dofs_to_nodes(void *ptr, const double *ext, double *node,
                      void *stream) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!ext || !node) return adm::set_error(ADM_ERR_INVALID, "null vector");
  if (ctx.n_ext < 0) return adm::set_error(ADM_ERR_INVALID, "no DoF map set");
  ctx.to_nodes(ext, node, static_cast<hipStream_t>(stream));
  ADM_API_END
}

int adm_nodes_to_dofs(void *ptr, const double *node, double *ext, int add,
                      void *stream) {
  ADM_API_BEGIN
  ADM_CTX(ptr)
  if (!ext || !node) return adm::set_error(ADM_ERR_INVALID, "null vector");
  if (ctx.n_ext < 0) return adm::set_error(ADM_ERR_INVALID, "no DoF map set");
  ctx.to_dofs(node, ext, add, static_cast<hipStream_t>(stream));
  ADM_API_END
}

int adm_ensemble_cells_to_activate(void *ptr, int32_t member, int32_t n_boxes,
                                   const double *boxes,
                                   int32_t *box_of_cell) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (member < 0 || member >= ens.n_members)
    return adm::set_error(ADM_ERR_INVALID, "member index out of range");
  if (n_boxes < 0 || (n_boxes && !boxes) || !box_of_cell)
    return adm::set_error(ADM_ERR_INVALID, "bad box list");
  ens.members[member]->cells_to_activate(n_boxes, boxes, box_of_cell);
  ADM_API_END
}

int adm_ensemble_add_material(void *ptr, int32_t member,
                              const int32_t *box_of_cell,
                              const double *box_cos, const double *box_sin,
                              double new_temperature, double *T,
                              void *stream) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (member < 0 || member >= ens.n_members)
    return adm::set_error(ADM_ERR_INVALID, "member index out of range");
  if (!box_of_cell || !box_cos || !box_sin)
    return adm::set_error(ADM_ERR_INVALID, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  adm::Context &c = *ens.members[member];
  c.add_material(box_of_cell, box_cos, box_sin, new_temperature,
                 T ? T + (int64_t)member * ens.n_dofs : nullptr, s);
  ens.refresh_members();
  if (ens.sb.inv_mass)
    c.compute_inverse_mass(ens.sb.inv_mass + (int64_t)member * ens.n_dofs, s);
  ADM_API_END
}

int adm_ensemble_activate(void *ptr, int32_t member, int32_t n_boxes,
                          const double *boxes, const double *box_cos,
                          const double *box_sin, double new_temperature,
                          double *T, int32_t *box_of_cell, void *stream) {
  ADM_API_BEGIN
  ADM_ENS(ptr)
  if (member < 0 || member >= ens.n_members)
    return adm::set_error(ADM_ERR_INVALID, "member index out of range");
  if (n_boxes < 0 || (n_boxes && (!boxes || !box_cos || !box_sin)))
    return adm::set_error(ADM_ERR_INVALID, "bad box list");
  adm::Context &c = *ens.members[member];
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (c.dim == 2) {
    c.activate_host(n_boxes, boxes, box_cos, box_sin, new_temperature,
                    T ? T + (int64_t)member * ens.n_dofs : nullptr,
                    box_of_cell, s);
  } else {
    if (!c.device_activation_fits())
      return adm::set_error(ADM_ERR_UNSUPPORTED,
                            "mesh too large for the device activation");
    c.activate(n_boxes, boxes, nullptr, box_cos, box_sin, new_temperature,
               T ? T + (int64_t)member * ens.n_dofs : nullptr, box_of_cell,
               s);
  }
  ens.refresh_members();
  if (ens.sb.inv_mass)
    c.compute_inverse_mass(ens.sb.inv_mass + (int64_t)member * ens.n_dofs, s);
  ADM_API_END
}

}  // extern "C"
