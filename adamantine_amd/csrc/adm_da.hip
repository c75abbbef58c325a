// EnKF analysis of a device ensemble (adamantine's DataAssimilator,
// source/DataAssimilator.cc) and the observation operator
// (source/experimental_data_utils.cc:24-110).
//
// The reference forms the localised sample covariance P of the augmented
// state [T; parameters] (calc_sample_covariance_sparse, :690-752) and applies
// x_j += P H^T (H P H^T + R)^{-1} d_j with one GMRES solve per member
// (apply_kalman_gain, :403-476). Here:
//
//  * only H X (members x observations) leaves the device;
//  * the small dense algebra runs on the host in FP64: C = H P H^T + R
//    (m x m, localised between observations) and E = C^{-1} D by LU, or,
//    when P is not localised and R is diagonal, the N x N ensemble-space
//    form W0 = ((N-1) I + S R^{-1} S^T)^{-1} S R^{-1} D^T (Woodbury);
//  * without localisation the state update is X <- X (I + W'), with
//    W' = (I - 1 1^T / N) W0, an n x N x N product on FP64 MFMA;
//  * with localisation (step function or Gaspari-Cohn) every node forms
//    its anomaly-observation covariances, weights them by the localisation
//    factor of its distance to each observation and applies E.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/adm.h"
#include "adm_context.hpp"

namespace adm {

#define ADM_HIP_DA(call)                                                  \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      throw AdmError(ADM_ERR_HIP, std::string(#call) + ": " +             \
                                      hipGetErrorString(e_));             \
  } while (0)

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

// DataAssimilator::gaspari_cohn_function (source/DataAssimilator.cc:672-688)
static double gaspari_cohn(double r) {
  if (r < 1.0)
    return 1. - 5. / 3. * std::pow(r, 2) + 5. / 8. * std::pow(r, 3) +
           0.5 * std::pow(r, 4) - 0.25 * std::pow(r, 5);
  else if (r < 2)
    return 4. - 5. * r + 5. / 3. * std::pow(r, 2) + 5. / 8. * std::pow(r, 3) -
           0.5 * std::pow(r, 4) + 1. / 12. * std::pow(r, 5) - 2. / (3. * r);
  return 0.;
}

// Localisation factor of a (state, state) covariance entry at support-point
// distance d: outside the cut-off the entry is not in the sparsity pattern
// (update_covariance_sparsity_pattern, :512-612); inside it the factor is 1
// (none, step_function) or GC(2 d / cutoff) (:727-745).
static double localization(int fn, double d, double cutoff) {
  if (!(d <= cutoff)) return 0.0;
  if (fn == ADM_LOCALIZATION_GASPARI_COHN)
    return gaspari_cohn(2.0 * d / cutoff);
  return 1.0;
}

// Support point of a node of the structured numbering.
struct NodeGeometry {
  int p = 1;
  int nc[3] = {1, 1, 1}, nn[3] = {2, 2, 2};
  double h[3] = {1, 1, 1}, origin[3] = {0, 0, 0};
  double gl[kMaxN + 2] = {0, 0, 0, 0, 0};

  explicit NodeGeometry(const Context &c) {
    p = c.p;
    for (int d = 0; d < 3; ++d) {
      nc[d] = c.nc[d];
      nn[d] = c.nn[d];
      h[d] = c.kc.h[d];
      origin[d] = c.kc.origin[d];
    }
    double w[kMaxN + 2];
    gauss_lobatto_points(p + 1, gl, w);
  }
  double axis(int d, int64_t i) const {
    int64_t cell = std::min<int64_t>(i / p, nc[d] - 1);
    const int64_t loc = i - cell * p;
    return origin[d] + (cell + gl[loc]) * h[d];
  }
  void coords(int64_t node, double x[3]) const {
    const int64_t ix = node % nn[0], iy = (node / nn[0]) % nn[1],
                  iz = node / ((int64_t)nn[0] * nn[1]);
    x[0] = axis(0, ix);
    x[1] = axis(1, iy);
    x[2] = axis(2, iz);
  }
};

static double dist3(const double a[3], const double b[3]) {
  double s = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double t = a[d] - b[d];
    s += t * t;
  }
  return std::sqrt(s);
}

// LU with partial pivoting; solves A X = B for nrhs right-hand sides stored
// column-major in B (m x nrhs). A (m x m row-major) is overwritten.
static void lu_solve(int m, std::vector<double> &A, int nrhs,
                     std::vector<double> &B) {
  std::vector<int> piv(m);
  for (int c = 0; c < m; ++c) {
    int pr = c;
    for (int r = c + 1; r < m; ++r)
      if (std::fabs(A[(size_t)r * m + c]) > std::fabs(A[(size_t)pr * m + c]))
        pr = r;
    if (A[(size_t)pr * m + c] == 0.0)
      throw AdmError(ADM_ERR_INVALID, "H P H^T + R is singular");
    piv[c] = pr;
    if (pr != c)
      for (int k = 0; k < m; ++k)
        std::swap(A[(size_t)c * m + k], A[(size_t)pr * m + k]);
    const double inv = 1.0 / A[(size_t)c * m + c];
    for (int r = c + 1; r < m; ++r) {
      const double f = A[(size_t)r * m + c] * inv;
      A[(size_t)r * m + c] = f;
      if (f == 0.0) continue;
      for (int k = c + 1; k < m; ++k)
        A[(size_t)r * m + k] -= f * A[(size_t)c * m + k];
    }
  }
  for (int q = 0; q < nrhs; ++q) {
    double *b = B.data() + (size_t)q * m;
    // P b first (the row swaps moved whole rows, stored multipliers
    // included), then L y = P b
    for (int c = 0; c < m; ++c)
      if (piv[c] != c) std::swap(b[c], b[piv[c]]);
    for (int c = 0; c < m; ++c)
      for (int r = c + 1; r < m; ++r) b[r] -= A[(size_t)r * m + c] * b[c];
    for (int r = m - 1; r >= 0; --r) {
      double s = b[r];
      for (int k = r + 1; k < m; ++k) s -= A[(size_t)r * m + k] * b[k];
      b[r] = s / A[(size_t)r * m + r];
    }
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// HX[j][k] = T[j][obs[k]]
__global__ void gather_obs_kernel(const double *__restrict__ T,
                                  const int64_t *__restrict__ obs, int64_t n,
                                  int m, int N, double *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)m * N) return;
  const int j = (int)(t / m), k = (int)(t % m);
  out[t] = T[(int64_t)j * n + obs[k]];
}

typedef double v4d __attribute__((ext_vector_type(4)));

// X <- X Tm for the member-major ensemble X[N][n] and Tm[NT*16][NT*16]
// (rows/columns >= N zero). One wavefront owns 16 nodes and all members:
// NT output tiles of 16 nodes x 16 members on v_mfma_f64_16x16x4f64,
// K = members in steps of 4. Operand layouts (16x16x4 f64): A lane l holds
// A[row l%16][k l/16]; B lane l holds B[k l/16][col l%16]; D lane l
// register r holds D[row 4r + l/16][col l%16]. Every X value of the
// wavefront's nodes is read before any is written, and wavefronts own
// disjoint nodes, so the update is in place.
template <int NT>
__global__ void __launch_bounds__(256)
    ensemble_transform_kernel(double *__restrict__ X,
                              const double *__restrict__ Tm, int N,
                              int64_t n) {
  constexpr int NP = NT * 16;
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int64_t i0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
  if (i0 >= n) return;
  const int row = lane % 16, kq = lane / 16;
  const int64_t node = i0 + row;
  const bool node_ok = node < n;
  v4d acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k0 = 0; k0 < NP; k0 += 4) {
    const int l = k0 + kq;
    const double a = (node_ok && l < N) ? X[(int64_t)l * n + node] : 0.0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const double b = Tm[(int64_t)l * NP + t * 16 + row];
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
    }
  }
  // all loads of this wavefront's nodes precede the stores below
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int j = t * 16 + row;
    if (j >= N) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t nd = i0 + 4 * r + kq;
      if (nd < n) X[(int64_t)j * n + nd] = acc[t][r];
    }
  }
}

// Localised update: per node i, g_k = rho(i, o_k) (sum_l A[l][i] S[l][k]) /
// (N-1) with A the member anomalies at i, then X[j][i] += sum_k g_k E[k][j].
// S[N][m] observation anomalies, E[m][N] = C^{-1} D, ocoord[m][3].
template <int NMAX>
__global__ void __launch_bounds__(256)
    localized_update_kernel(double *__restrict__ X, int N, int64_t n,
                            const double *__restrict__ S,
                            const double *__restrict__ E,
                            const double *__restrict__ ocoord, int m, int fn,
                            double cutoff, NodeGeometry geo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a[NMAX], acc[NMAX];
  double sum = X[i];
  a[0] = sum;
#pragma unroll
  for (int l = 1; l < NMAX; ++l) {
    if (l < N) {
      a[l] = X[(int64_t)l * n + i];
      sum += a[l];
    }
  }
  const double mean = sum / N;
#pragma unroll
  for (int l = 0; l < NMAX; ++l) {
    if (l < N) a[l] -= mean;
    acc[l] = 0.0;
  }
  const int64_t ix = i % geo.nn[0], iy = (i / geo.nn[0]) % geo.nn[1],
                iz = i / ((int64_t)geo.nn[0] * geo.nn[1]);
  const int64_t id[3] = {ix, iy, iz};
  double x[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int64_t cell = min(id[d] / geo.p, (int64_t)geo.nc[d] - 1);
    x[d] = geo.origin[d] + (cell + geo.gl[id[d] - cell * geo.p]) * geo.h[d];
  }
  const double inv_nm1 = 1.0 / (N - 1.0);
  for (int k = 0; k < m; ++k) {
    double s2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double t = x[d] - ocoord[3 * k + d];
      s2 += t * t;
    }
    const double d = sqrt(s2);
    if (!(d <= cutoff)) continue;
    double rho = 1.0;
    if (fn == ADM_LOCALIZATION_GASPARI_COHN) {
      const double r = 2.0 * d / cutoff;
      if (r < 1.0)
        rho = 1. - 5. / 3. * (r * r) + 5. / 8. * (r * r * r) +
              0.5 * (r * r * r * r) - 0.25 * (r * r * r * r * r);
      else if (r < 2)
        rho = 4. - 5. * r + 5. / 3. * (r * r) + 5. / 8. * (r * r * r) -
              0.5 * (r * r * r * r) + 1. / 12. * (r * r * r * r * r) -
              2. / (3. * r);
      else
        rho = 0.0;
    }
    double g = 0.0;
#pragma unroll
    for (int l = 0; l < NMAX; ++l)
      if (l < N) g += a[l] * S[(int64_t)l * m + k];
    g = g * inv_nm1 * rho;
#pragma unroll
    for (int j = 0; j < NMAX; ++j)
      if (j < N) acc[j] += g * E[(int64_t)k * N + j];
  }
#pragma unroll
  for (int j = 0; j < NMAX; ++j)
    if (j < N) X[(int64_t)j * n + i] = (a[j] + mean) + acc[j];
}

// ---------------------------------------------------------------------------
// analysis
// ---------------------------------------------------------------------------

struct DaState {
  std::mt19937 prng;                     // DataAssimilator::_prng
  std::normal_distribution<> normal;     // _normal_dist_generator
};

static void launch_transform(int NT, double *X, const double *Tm, int N,
                             int64_t n, hipStream_t s) {
  const unsigned grid = (unsigned)((n + 63) / 64);
#define ADM_TR(K)                                                         \
  case K:                                                                 \
    hipLaunchKernelGGL(ensemble_transform_kernel<K>, dim3(grid), dim3(256), \
                       0, s, X, Tm, N, n);                                \
    break;
  switch (NT) {
    ADM_TR(1)
    ADM_TR(2)
    ADM_TR(3)
    ADM_TR(4)
    ADM_TR(5)
    ADM_TR(6)
    ADM_TR(7)
    ADM_TR(8)
    default:
      throw AdmError(ADM_ERR_UNSUPPORTED, "more than 128 ensemble members");
  }
#undef ADM_TR
  ADM_HIP_DA(hipGetLastError());
}

// X <- X Tm on the device (Tm host, N x N row-major), the padded matrix
// staged through the ensemble's pinned workspace; no allocation once the
// workspace has grown to the member count, no stream synchronisation.
void ensemble_transform(Ensemble &ens, double *X, const double *Tm_host,
                        hipStream_t s) {
  const int N = ens.n_members;
  const int64_t n = ens.n_dofs;
  const int NT = (N + 15) / 16;
  const int NP = NT * 16;
  const size_t np2 = (size_t)NP * NP;
  double *Tp = ens.da_host(np2);
  std::fill(Tp, Tp + np2, 0.0);
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) Tp[(size_t)a * NP + b] = Tm_host[(size_t)a * N + b];
  double *d_T = ens.da_device(np2);
  ADM_HIP_DA(hipMemcpyAsync(d_T, Tp, sizeof(double) * np2,
                            hipMemcpyHostToDevice, s));
  ADM_HIP_DA(hipEventRecord(ens.da_done, s));
  launch_transform(NT, X, d_T, N, n, s);
}

void ensemble_analysis(Ensemble &ens, double *T, int n_params, double *params,
                       int m, const int64_t *obs_dof, const double *y,
                       const double *R, const double *noise, int fn,
                       double cutoff, hipStream_t s) {
  const int N = ens.n_members;
  const int64_t n = ens.n_dofs;
  if (N < 2) throw AdmError(ADM_ERR_INVALID, "an analysis needs >= 2 members");
  if (m < 1) throw AdmError(ADM_ERR_INVALID, "no observations");
  if (fn < ADM_LOCALIZATION_NONE || fn > ADM_LOCALIZATION_GASPARI_COHN)
    throw AdmError(ADM_ERR_INVALID, "unknown localization function");
  if (!(cutoff > 0.0)) throw AdmError(ADM_ERR_INVALID, "cutoff must be > 0");
  for (int k = 0; k < m; ++k)
    if (obs_dof[k] < 0 || obs_dof[k] >= n)
      throw AdmError(ADM_ERR_INVALID, "observation DoF out of range");
  const Context &c0 = *ens.members[0];
  const NodeGeometry geo(c0);

  // H X on the host
  std::vector<double> HX((size_t)N * m);
  {
    // workspace: m observation indices (as doubles' storage), then H X
    const size_t n_ws = (size_t)m + (size_t)N * m;
    double *ws = ens.da_device(n_ws);
    double *stage = ens.da_host(n_ws);
    int64_t *d_obs = reinterpret_cast<int64_t *>(ws);
    double *d_hx = ws + m;
    std::memcpy(stage, obs_dof, sizeof(int64_t) * m);
    ADM_HIP_DA(hipMemcpyAsync(d_obs, stage, sizeof(int64_t) * m,
                              hipMemcpyHostToDevice, s));
    const int64_t tot = (int64_t)m * N;
    hipLaunchKernelGGL(gather_obs_kernel, dim3((unsigned)((tot + 255) / 256)),
                       dim3(256), 0, s, T, d_obs, n, m, N, d_hx);
    ADM_HIP_DA(hipGetLastError());
    ADM_HIP_DA(hipMemcpyAsync(stage + m, d_hx, sizeof(double) * HX.size(),
                              hipMemcpyDeviceToHost, s));
    ADM_HIP_DA(hipEventRecord(ens.da_done, s));
    // the host solve needs H X: wait for this copy only
    ADM_HIP_DA(hipEventSynchronize(ens.da_done));
    std::memcpy(HX.data(), stage + m, sizeof(double) * HX.size());
  }
  // anomalies at the observations (the mean as calc_sample_covariance_sparse
  // forms it: sum in member order, then / N) and the perturbed innovations
  // d_j = noise_j + (y - H x_j) (update_ensemble, :130-143)
  std::vector<double> S((size_t)N * m), D((size_t)N * m);
  for (int k = 0; k < m; ++k) {
    double sum = HX[k];
    for (int j = 1; j < N; ++j) sum += HX[(size_t)j * m + k];
    const double mean = sum / N;
    for (int j = 0; j < N; ++j) {
      S[(size_t)j * m + k] = HX[(size_t)j * m + k] - mean;
      double v = noise ? noise[(size_t)j * m + k] : 0.0;
      v += y[k] - HX[(size_t)j * m + k];
      D[(size_t)j * m + k] = v;
    }
  }
  std::vector<double> ocoord((size_t)m * 3);
  for (int k = 0; k < m; ++k) geo.coords(obs_dof[k], &ocoord[(size_t)3 * k]);
  // is P localised anywhere? (every node pair closer than the cut-off and a
  // factor of 1 means no)
  double diag = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double e = c0.kc.h[d] * c0.nc[d];
    diag += e * e;
  }
  diag = std::sqrt(diag);
  bool localized = !(cutoff >= diag);
  if (fn == ADM_LOCALIZATION_GASPARI_COHN &&
      !(gaspari_cohn(2.0 * diag / cutoff) == 1.0))
    localized = true;
  bool r_diag = true;
  for (int a = 0; a < m && r_diag; ++a)
    for (int b = 0; b < m; ++b)
      if (a != b && R[(size_t)a * m + b] != 0.0) {
        r_diag = false;
        break;
      }
  const double nm1 = N - 1.0;
  // W0 = S E / (N-1) (N x N, [l][j]); E = C^{-1} D^T (m x N, column j of D)
  std::vector<double> W0((size_t)N * N, 0.0), E;
  if (!localized && r_diag) {
    // Woodbury: S C^{-1} = (N-1) G^{-1} S R^{-1}, G = (N-1) I + S R^{-1} S^T
    std::vector<double> G((size_t)N * N), B((size_t)N * N);
    for (int a = 0; a < N; ++a)
      for (int b = 0; b < N; ++b) {
        double g = 0.0, bb = 0.0;
        for (int k = 0; k < m; ++k) {
          const double rinv = 1.0 / R[(size_t)k * m + k];
          g += S[(size_t)a * m + k] * rinv * S[(size_t)b * m + k];
          bb += S[(size_t)a * m + k] * rinv * D[(size_t)b * m + k];
        }
        G[(size_t)a * N + b] = g + (a == b ? nm1 : 0.0);
        B[(size_t)b * N + a] = bb;  // column b
      }
    lu_solve(N, G, N, B);
    for (int l = 0; l < N; ++l)
      for (int j = 0; j < N; ++j) W0[(size_t)l * N + j] = B[(size_t)j * N + l];
  } else {
    std::vector<double> Cm((size_t)m * m);
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b) {
        const double loc =
            localization(fn, dist3(&ocoord[3 * a], &ocoord[3 * b]), cutoff);
        double e = 0.0;
        if (loc != 0.0) {
          for (int l = 0; l < N; ++l)
            e += S[(size_t)l * m + a] * S[(size_t)l * m + b];
          e /= nm1;
          e *= loc;
        }
        Cm[(size_t)a * m + b] = e + R[(size_t)a * m + b];
      }
    E = D;  // column-major m x N: column j = d_j
    lu_solve(m, Cm, N, E);
    for (int l = 0; l < N; ++l)
      for (int j = 0; j < N; ++j) {
        double w = 0.0;
        for (int k = 0; k < m; ++k)
          w += S[(size_t)l * m + k] * E[(size_t)j * m + k];
        W0[(size_t)l * N + j] = w / nm1;
      }
  }
  // parameters (never localised): p_j += sum_l Ap[l] W0[l][j]
  if (n_params > 0 && params) {
    std::vector<double> Ap((size_t)N * n_params);
    for (int q = 0; q < n_params; ++q) {
      double sum = params[q];
      for (int j = 1; j < N; ++j) sum += params[(size_t)j * n_params + q];
      const double mean = sum / N;
      for (int j = 0; j < N; ++j)
        Ap[(size_t)j * n_params + q] = params[(size_t)j * n_params + q] - mean;
    }
    for (int j = 0; j < N; ++j)
      for (int q = 0; q < n_params; ++q) {
        double sh = 0.0;
        for (int l = 0; l < N; ++l)
          sh += Ap[(size_t)l * n_params + q] * W0[(size_t)l * N + j];
        params[(size_t)j * n_params + q] += sh;
      }
  }
  if (!localized) {
    // X (I + W'), W' = W0 - 1 (1^T W0) / N
    std::vector<double> Tm((size_t)N * N);
    for (int j = 0; j < N; ++j) {
      double cs = 0.0;
      for (int l = 0; l < N; ++l) cs += W0[(size_t)l * N + j];
      cs /= N;
      for (int l = 0; l < N; ++l)
        Tm[(size_t)l * N + j] = (l == j ? 1.0 : 0.0) + (W0[(size_t)l * N + j] - cs);
    }
    ensemble_transform(ens, T, Tm.data(), s);
    return;
  }
  if (N > 64)
    throw AdmError(ADM_ERR_UNSUPPORTED,
                   "localised analysis supports at most 64 members");
  // E as [m][N] row-major for the kernel
  std::vector<double> Ek((size_t)m * N);
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < N; ++j) Ek[(size_t)k * N + j] = E[(size_t)j * m + k];
  const size_t n_ws = S.size() + Ek.size() + ocoord.size();
  double *d_S = ens.da_device(n_ws);
  double *d_E = d_S + S.size();
  double *d_oc = d_E + Ek.size();
  double *stage = ens.da_host(n_ws);
  std::copy(S.begin(), S.end(), stage);
  std::copy(Ek.begin(), Ek.end(), stage + S.size());
  std::copy(ocoord.begin(), ocoord.end(), stage + S.size() + Ek.size());
  ADM_HIP_DA(hipMemcpyAsync(d_S, stage, sizeof(double) * n_ws,
                            hipMemcpyHostToDevice, s));
  ADM_HIP_DA(hipEventRecord(ens.da_done, s));
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (N <= 16)
    hipLaunchKernelGGL(localized_update_kernel<16>, dim3(grid), dim3(256), 0, s,
                       T, N, n, d_S, d_E, d_oc, m, fn, cutoff, geo);
  else if (N <= 32)
    hipLaunchKernelGGL(localized_update_kernel<32>, dim3(grid), dim3(256), 0, s,
                       T, N, n, d_S, d_E, d_oc, m, fn, cutoff, geo);
  else
    hipLaunchKernelGGL(localized_update_kernel<64>, dim3(grid), dim3(256), 0, s,
                       T, N, n, d_S, d_E, d_oc, m, fn, cutoff, geo);
  ADM_HIP_DA(hipGetLastError());
}

// T[idx[k]] = v[k] (distinct indices)
__global__ void gather_set_kernel(double *__restrict__ T,
                                  const int64_t *__restrict__ idx,
                                  const double *__restrict__ v, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) T[idx[k]] = v[k];
}

// get_expt_to_dof_mapping (source/experimental_data_utils.cc:70-110): the
// nearest support point among the nodes of active cells. Equidistant nodes
// resolve to the lowest node index. The search scans index boxes around the
// point's nearest grid index, in increasing node order, and widens the box
// until it holds every node closer than the best one found: the result is
// the one of a scan over all nodes.
void nearest_dofs(const Context &c, int64_t np, const double *pts,
                  int64_t *out) {
  const NodeGeometry geo(c);
  // nodes of active cells
  std::vector<uint8_t> node_active((size_t)c.n_dofs, 0);
  for (int64_t cell = 0; cell < c.n_cells; ++cell) {
    if (!c.active[cell]) continue;
    const int64_t cx = cell % c.nc[0], cy = (cell / c.nc[0]) % c.nc[1],
                  cz = cell / ((int64_t)c.nc[0] * c.nc[1]);
    for (int k = 0; k <= c.p; ++k)
      for (int j = 0; j <= (c.dim == 3 ? c.p : 0); ++j)
        for (int i = 0; i <= c.p; ++i)
          node_active[(size_t)(cx * c.p + i) +
                      (size_t)c.nn[0] * ((cy * c.p + j) +
                                         (size_t)c.nn[1] * (cz * c.p + k))] = 1;
  }
  // smallest node spacing per axis
  double gap = 1.0;
  for (int k = 0; k < c.p; ++k) gap = std::min(gap, geo.gl[k + 1] - geo.gl[k]);
  for (int64_t q = 0; q < np; ++q) {
    const double *x = pts + 3 * q;
    int64_t i0[3];
    for (int d = 0; d < 3; ++d) {
      const double f = (x[d] - geo.origin[d]) / geo.h[d] * c.p;
      const double cl =
          f == f ? std::min(std::max(f, 0.0), (double)(c.nn[d] - 1)) : 0.0;
      i0[d] = (int64_t)std::llround(cl);
    }
    int64_t r = 1;
    double best = DBL_MAX;
    int64_t best_i = -1;
    for (;;) {
      best = DBL_MAX;
      best_i = -1;
      int64_t lo[3], hi[3];
      bool whole = true;
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::max<int64_t>(i0[d] - r, 0);
        hi[d] = std::min<int64_t>(i0[d] + r, c.nn[d] - 1);
        whole = whole && lo[d] == 0 && hi[d] == c.nn[d] - 1;
      }
      for (int64_t iz = lo[2]; iz <= hi[2]; ++iz)
        for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
          for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
            const int64_t i = ix + c.nn[0] * (iy + (int64_t)c.nn[1] * iz);
            if (!node_active[(size_t)i]) continue;
            double xi[3];
            geo.coords(i, xi);
            const double dd = dist3(x, xi);
            if (dd < best) {
              best = dd;
              best_i = i;
            }
          }
      if (whole) break;
      if (best_i < 0) {
        r *= 2;
        continue;
      }
      // every node within `best` of x lies within R index steps of i0
      int64_t R = 0;
      for (int d = 0; d < 3; ++d)
        R = std::max<int64_t>(
            R, (int64_t)std::ceil(best / (geo.h[d] * gap)) + 1);
      if (R <= r) break;
      r = R;
    }
    out[q] = best_i;
  }
}

}  // namespace adm

using adm::AdmError;

#define ADM_DA_BEGIN try {
#define ADM_DA_END                                               \
  }                                                              \
  catch (const AdmError &e) {                                    \
    return adm::set_error(e.code, e.what());                     \
  }                                                              \
  catch (const std::exception &e) {                              \
    return adm::set_error(ADM_ERR_INVALID, e.what());            \
  }                                                              \
  return ADM_OK;

extern "C" {

int adm_da_create(void **da) {
  ADM_DA_BEGIN
  if (!da) return adm::set_error(ADM_ERR_INVALID, "null argument");
  *da = new adm::DaState();
  ADM_DA_END
}

int adm_da_destroy(void *da) {
  ADM_DA_BEGIN
  delete static_cast<adm::DaState *>(da);
  ADM_DA_END
}

int adm_da_fill_noise(void *da, int32_t count, int32_t n_obs, const double *R,
                      double *out) {
  ADM_DA_BEGIN
  if (!da || !R || !out) return adm::set_error(ADM_ERR_INVALID, "null argument");
  if (count < 0 || n_obs < 1) return adm::set_error(ADM_ERR_INVALID, "bad size");
  adm::DaState &st = *static_cast<adm::DaState *>(da);
  const int m = n_obs;
  bool diagonal = true;
  for (int a = 0; a < m && diagonal; ++a)
    for (int b = 0; b < m; ++b)
      if (a != b && R[(size_t)a * m + b] != 0.0) {
        diagonal = false;
        break;
      }
  std::vector<double> L;
  if (!diagonal) {
    // R = L L^T (dealii::FullMatrix::cholesky)
    L.assign((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j) {
      double d = R[(size_t)j * m + j];
      for (int k = 0; k < j; ++k) d -= L[(size_t)j * m + k] * L[(size_t)j * m + k];
      if (!(d > 0.0))
        return adm::set_error(ADM_ERR_INVALID, "R is not positive definite");
      L[(size_t)j * m + j] = std::sqrt(d);
      for (int i = j + 1; i < m; ++i) {
        double v = R[(size_t)i * m + j];
        for (int k = 0; k < j; ++k) v -= L[(size_t)i * m + k] * L[(size_t)j * m + k];
        L[(size_t)i * m + j] = v / L[(size_t)j * m + j];
      }
    }
  }
  std::vector<double> z((size_t)m);
  for (int c = 0; c < count; ++c) {
    double *v = out + (size_t)c * m;
    if (diagonal) {
      for (int i = 0; i < m; ++i)
        v[i] = st.normal(st.prng) * std::sqrt(R[(size_t)i * m + i]);
    } else {
      for (int i = 0; i < m; ++i) z[i] = st.normal(st.prng);
      for (int i = 0; i < m; ++i) {
        double acc = 0.0;
        for (int k = 0; k <= i; ++k) acc += L[(size_t)i * m + k] * z[k];
        v[i] = acc;
      }
    }
  }
  ADM_DA_END
}

int adm_ensemble_nearest_dofs(void *ens, int32_t member, int64_t n_points,
                              const double *points, int64_t *dofs) {
  ADM_DA_BEGIN
  if (!ens || (!points && n_points) || (!dofs && n_points))
    return adm::set_error(ADM_ERR_INVALID, "null argument");
  adm::Ensemble &e = *static_cast<adm::Ensemble *>(ens);
  if (member < 0 || member >= e.n_members)
    return adm::set_error(ADM_ERR_INVALID, "member index out of range");
  adm::nearest_dofs(*e.members[member], n_points, points, dofs);
  ADM_DA_END
}

int adm_nearest_dofs(void *ctx, int64_t n_points, const double *points,
                     int64_t *dofs) {
  ADM_DA_BEGIN
  if (!ctx || (!points && n_points) || (!dofs && n_points))
    return adm::set_error(ADM_ERR_INVALID, "null argument");
  adm::nearest_dofs(*static_cast<adm::Context *>(ctx), n_points, points, dofs);
  ADM_DA_END
}

int adm_set_values(void *ctx, int64_t n, const int64_t *dofs,
                   const double *values, double *T, void *stream) {
  ADM_DA_BEGIN
  if (!ctx || !T || (n && (!dofs || !values)))
    return adm::set_error(ADM_ERR_INVALID, "null argument");
  adm::Context &c = *static_cast<adm::Context *>(ctx);
  // in input order, so a repeated DoF keeps its last value
  std::vector<int64_t> last((size_t)c.n_dofs, -1);
  for (int64_t k = 0; k < n; ++k) {
    if (dofs[k] < 0 || dofs[k] >= c.n_dofs)
      return adm::set_error(ADM_ERR_INVALID, "DoF index out of range");
    last[(size_t)dofs[k]] = k;
  }
  std::vector<int64_t> idx;
  std::vector<double> val;
  for (int64_t k = 0; k < n; ++k)
    if (last[(size_t)dofs[k]] == k) {
      idx.push_back(dofs[k]);
      val.push_back(values[k]);
    }
  if (idx.empty()) return ADM_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int64_t *d_idx = nullptr;
  double *d_val = nullptr;
  ADM_HIP_DA(hipMalloc(&d_idx, sizeof(int64_t) * idx.size()));
  ADM_HIP_DA(hipMalloc(&d_val, sizeof(double) * idx.size()));
  ADM_HIP_DA(hipMemcpyAsync(d_idx, idx.data(), sizeof(int64_t) * idx.size(),
                            hipMemcpyHostToDevice, s));
  ADM_HIP_DA(hipMemcpyAsync(d_val, val.data(), sizeof(double) * idx.size(),
                            hipMemcpyHostToDevice, s));
  const int64_t m = (int64_t)idx.size();
  hipLaunchKernelGGL(adm::gather_set_kernel, dim3((unsigned)((m + 255) / 256)),
                     dim3(256), 0, s, T, d_idx, d_val, m);
  ADM_HIP_DA(hipGetLastError());
  ADM_HIP_DA(hipStreamSynchronize(s));
  ADM_HIP_DA(hipFree(d_idx));
  ADM_HIP_DA(hipFree(d_val));
  ADM_DA_END
}

int adm_ensemble_analysis(void *ens, double *T, int32_t n_params,
                          double *params, int32_t n_obs,
                          const int64_t *obs_dof, const double *obs,
                          const double *R, const double *noise,
                          int32_t localization, double cutoff, void *stream) {
  ADM_DA_BEGIN
  if (!ens || !T || !obs_dof || !obs || !R)
    return adm::set_error(ADM_ERR_INVALID, "null argument");
  if (n_params < 0 || (n_params > 0 && !params))
    return adm::set_error(ADM_ERR_INVALID, "bad parameter block");
  adm::ensemble_analysis(*static_cast<adm::Ensemble *>(ens), T, n_params,
                         params, n_obs, obs_dof, obs, R, noise, localization,
                         cutoff, static_cast<hipStream_t>(stream));
  ADM_DA_END
}

int adm_ensemble_transform(void *ens, double *T, const double *Tm,
                           void *stream) {
  ADM_DA_BEGIN
  if (!ens || !T || !Tm) return adm::set_error(ADM_ERR_INVALID, "null argument");
  adm::Ensemble &e = *static_cast<adm::Ensemble *>(ens);
  adm::ensemble_transform(e, T, Tm, static_cast<hipStream_t>(stream));
  ADM_DA_END
}

}  // extern "C"
