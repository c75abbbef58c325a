// Discrete adjoint of the explicit time steppers and parameter gradients.
//
// The reference has no adjoint. For a final-time misfit
//   J = 1/2 sum_k w_k (T_final[dof_k] - y_k)^2
// this computes dJ/dT0 and dJ/dtheta for theta = (absorption efficiency of
// beam 0, a factor on every thermal conductivity at its current value,
// radius of beam 0) by the exact transposition of the forward Euler, RK3 or
// classic RK4 stage sequence (the forward run is the one of
// ThermalPhysics::evolve_one_time_step, source/ThermalPhysics.templates.hh:
// 854-888), with the trajectory recomputed from checkpoints.
//
// Linearisation of the operator A(T) of source/ThermalOperator.templates.hh:
// 479-751 at a stage input: per quadrature point, with gradient g, test
// field mu (mu = M^-1 times the stage adjoint) and its gradient gm,
//   flux coefficient    c   = -(1/rho c_p) K gm JxW            (as w of A)
//   value coefficient   c_T = -JxW [d(1/rho c_p) gm.K g + (1/rho c_p) gm.K' g]
//                             + mu d(1/rho c_p) Q JxW
// integrated like A's fluxes and source. d/dT of the mixtures includes the
// liquid ratio (mushy range). The powder ratio carried from one evaluation
// to the next, p = min(p_old, 1 - liquid), has its own adjoint per
// quadrature point (cells and boundary faces), taken backwards through the
// evaluations: where p was clamped to 1 - liquid it feeds the temperature
// of that evaluation, elsewhere the powder before it. The latent-heat switch
// and the source cut-offs are piecewise constant and contribute nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <memory>
#include <tuple>
#include <vector>

#include "../../include/adm.h"
#include "adm_context.hpp"
#include "adm_kernels.hpp"
#include "adm_plane_kernel.hpp"

namespace adm {

#define ADM_HIP_ADJ(call)                                                 \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      throw AdmError(ADM_ERR_HIP, std::string(#call) + ": " +             \
                                      hipGetErrorString(e_));             \
  } while (0)

struct AdjArgs {
  CellArgs c;                 // mesh, constants, beams, row mapping
  const double *T;            // stage input
  const double *mu;           // M^-1 * stage adjoint
  double *out;                // += (dA/dT)^T mu
  const double *powder_old;   // powder before the evaluation [slots][n_q]
  double *powder_bar;         // adjoint of the carried powder [slots][n_q]
                              // (in: after, out: before the evaluation);
                              // NULL: zero after, not stored
  double *partials;           // [3]: d/d eta_0, d/d k-factor, d/d R_0
  double inv_eta0;            // 1 / absorption efficiency of beam 0 (or 0)
  // per-block sums of the partials, [member][part_blocks][3]: this launch's
  // blocks at part_base .. (reduce_partials_kernel adds them to partials in
  // a fixed order: no same-address atomics, the same sums every run)
  double *block_part;
  int32_t part_blocks, part_base;
  // ensemble launch (c.members != NULL): member blockIdx.y's vectors at
  // blockIdx.y * c.member_stride, its powder arrays at pw_off[blockIdx.y],
  // its partials at partials + 3 blockIdx.y, 1/eta_0 from inv_eta
  const int64_t *pw_off;
  const double *inv_eta;
};

// The face term's arguments of an ensemble launch: member blockIdx.y's
// face powder arrays at fpw_off[blockIdx.y].
struct AdjFaceArgs {
  const int64_t *fpw_off;
};

// Mixed property value v with its partials: dT at fixed carried powder and
// dp with respect to the carried powder ratio.
template <int NS>
struct Mix {
  double r[NS], rT[NS], rp[NS];
};

// index of the merged-table interval holding T: #{B_k <= T} (binary
// search over the padded breakpoints)
__device__ __forceinline__ int merged_search(const double *B, double T) {
  int k = 0;
#pragma unroll
  for (int step = kMergedMax / 2; step >= 1; step >>= 1)
    k += (B[k + step - 1] <= T) ? step : 0;
  return k;
}

// value, T- and powder-partials of a mixed property from the merged tables,
// in interval k of T
template <int NS>
__device__ __forceinline__ void merged_vd(const double *B, const double *C,
                                          double T, int k, int prop,
                                          const Mix<NS> &x, double &v,
                                          double &vT, double &vp) {
  const double d = T - B[k > 0 ? k - 1 : 0];
  const double *c = C + ((prop * kMergedMax + k) * ADM_MAX_STATES) * 2;
  v = 0.0;
  vT = 0.0;
  vp = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double xs = __builtin_fma(d, c[2 * s + 1], c[2 * s]);
    v += x.r[s] * xs;
    vT += x.rT[s] * xs + x.r[s] * c[2 * s + 1];
    vp += x.rp[s] * xs;
  }
}

// the same for a mixed polynomial property
template <int NS>
__device__ __forceinline__ void poly_vd(const MatConst &m, int mid, double T,
                                        int prop, const Mix<NS> &x, double &v,
                                        double &vT, double &vp) {
  v = 0.0;
  vT = 0.0;
  vp = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double xs = 0.0, dxs = 0.0, tp = 1.0, tpm = 0.0;
#pragma unroll
    for (int i = 0; i < ADM_MAX_POLY; ++i) {
      if (i > m.p_order) break;
      const double ci = m.poly[mid][prop][s][i];
      xs += ci * tp;
      dxs += i * ci * tpm;
      tpm = tp;
      tp *= T;
    }
    v += x.r[s] * xs;
    vT += x.rT[s] * xs + x.r[s] * dxs;
    vp += x.rp[s] * xs;
  }
}

// update_state_ratios at one point with its partials. The carried powder
// p = min(p_old, 1 - liquid) is returned through `clamped` (p = 1 - liquid,
// dp/dT = -dl) or not (p = p_old).
template <int NS>
__device__ __forceinline__ void ratios_lin(double T, double p_old, double sol,
                                           double liq, Mix<NS> &x, double &dl,
                                           bool &clamped) {
  dl = 0.0;
  clamped = false;
  if constexpr (NS == 1) {
    x.r[0] = 1.0;
    x.rT[0] = 0.0;
    x.rp[0] = 0.0;
  } else {
    const double mushy = (T - sol) / (liq - sol);
    const double l = T < sol ? 0.0 : (T > liq ? 1.0 : mushy);
    dl = (T < sol || T > liq) ? 0.0 : 1.0 / (liq - sol);
    x.r[1] = l;
    x.rT[1] = dl;
    x.rp[1] = 0.0;
    if constexpr (NS == 2) {
      x.r[0] = 1.0 - l;
      x.rT[0] = -dl;
      x.rp[0] = 0.0;
    } else {
      const double am = 1.0 - l;
      clamped = !(p_old < am);
      const double p = clamped ? am : p_old;
      x.r[2] = p;
      x.r[0] = 1.0 - l - p;
      x.rT[2] = 0.0;
      x.rT[0] = -dl;
      x.rp[2] = 1.0;
      x.rp[0] = -1.0;
    }
  }
}

// k: the merged-table interval of T (merged tables only)
template <int NS, int PK>
__device__ __forceinline__ void prop_vd(const MatConst &m, int mid,
                                        const double *mB, const double *mC,
                                        double T, int k, int prop,
                                        const Mix<NS> &x, double &v,
                                        double &vT, double &vp) {
  if constexpr (PK == 2)
    merged_vd<NS>(mB, mC, T, k, prop, x, v, vT, vp);
  else
    poly_vd<NS>(m, mid, T, prop, x, v, vT, vp);
}

// 1/(rho c_p) (get_inv_rho_cp, latent heat in the mushy range) and its
// partials
template <int NS, int PK>
__device__ __forceinline__ void inv_rho_cp_lin(const MatConst &m, int mid,
                                               const double *mB,
                                               const double *mC, double T,
                                               int k, double latent,
                                               const Mix<NS> &x, double &irc,
                                               double &ircT, double &ircp) {
  double rho, rhoT, rhop, cp, cpT, cpp;
  prop_vd<NS, PK>(m, mid, mB, mC, T, k, ADM_DENSITY, x, rho, rhoT, rhop);
  prop_vd<NS, PK>(m, mid, mB, mC, T, k, ADM_SPECIFIC_HEAT, x, cp, cpT, cpp);
  if constexpr (NS > 1)
    if (x.r[1] * x.r[0] > 0.0) cp += latent;
  irc = 1.0 / (rho * cp);
  ircT = -irc * irc * (rhoT * cp + rho * cpT);
  ircp = -irc * irc * (rhop * cp + rho * cpp);
}

// d Q_0 / d R_0 at a point for beam 0 (Goldak or electron beam; the cut-offs
// are held fixed)
__device__ __forceinline__ double beam_value_dR(const adm_beam_state &s,
                                                double q, double x, double y,
                                                double z, double log_01) {
  if (s.type == ADM_SOURCE_CUBE || q == 0.0) return 0.0;
  source_point(s, x, y, z);
  const double dx = x - s.center[0], dy = y - s.center[1];
  const double r2 = dx * dx + dy * dy;
  const double R = sqrt(s.radius_squared);
  const double R3 = R * s.radius_squared;
  if (s.type == ADM_SOURCE_GOLDAK) return q * (-2.0 / R + 6.0 * r2 / R3);
  return q * (-2.0 / R - 2.0 * log_01 * r2 / R3);
}

// (dA/dT)^T mu over the cells: same lane layout as cell_kernel (one x-line
// per lane, (p+1)^2 lanes per cell, one row segment per wavefront), two
// fields (T, mu) evaluated, atomic accumulation (or plain read-add-write
// of the cells of one parity colour per launch), parameter partials summed
// per block into block_part (reduce_partials_kernel adds them up).
template <int P, int NS, int PK>
__global__ void __launch_bounds__(256)
    adjoint_cell_kernel(const AdjArgs aa) {
  constexpr int N = P + 1, NL = N * N, N3 = N * N * N;
  constexpr int CPW = CellLanes<P>::cells_per_wave;
  const CellArgs &a = aa.c;
  // one operator, or member blockIdx.y of an ensemble (a constant-space
  // pointer: a generic one selected against the argument block's fields
  // would make the compiler copy the whole argument block to scratch)
  const ADM_CONSTANT MemberArgs *mb =
      a.members ? as_constant(a.members) + blockIdx.y : nullptr;
  const uint32_t *meta_p = mb ? mb->meta : a.meta;
  const int32_t *pslot_p = mb ? mb->powder_slot : a.powder_slot;
  const double *dcos_p = mb ? mb->dep_cos : a.dep_cos;
  const double *dsin_p = mb ? mb->dep_sin : a.dep_sin;
  const int64_t voff = mb ? (int64_t)blockIdx.y * a.member_stride : 0;
  const int64_t poff = mb ? aa.pw_off[blockIdx.y] : 0;
  const double *T_p = aa.T + voff;
  const double *mu_p = aa.mu + voff;
  double *out_p = aa.out + voff;
  const double *pwo_p = aa.powder_old ? aa.powder_old + poff : nullptr;
  double *pwb_p = aa.powder_bar ? aa.powder_bar + poff : nullptr;
  double *bpart_p =
      aa.block_part +
      3 * ((mb ? (int64_t)blockIdx.y * aa.part_blocks : 0) + aa.part_base +
           blockIdx.x);
  const double inv_eta0 = mb ? aa.inv_eta[blockIdx.y] : aa.inv_eta0;
  const KConst &kc = *(const KConst *)as_global(mb ? mb->kc : a.kc);
  const MatConst &m = kc.mat;
  extern __shared__ double lds[];
  __shared__ adm_beam_state sbeams[ADM_MAX_BEAMS];
  const int n_mat = m.n_materials;
  {
    constexpr int words = sizeof(adm_beam_state) * ADM_MAX_BEAMS / 8;
    const int tid = threadIdx.x;
    if (tid < words)
      reinterpret_cast<uint64_t *>(sbeams)[tid] =
          reinterpret_cast<const uint64_t *>(a.beams)[tid];
    if constexpr (PK == 2) {
      const int total = n_mat * kPlaneMergedLds;
      for (int i = tid; i < total; i += blockDim.x) {
        const int mm = i / kPlaneMergedLds, k = i % kPlaneMergedLds;
        lds[i] = k < kMergedLds ? (&m.merged[mm].B[0])[k]
                                : reinterpret_cast<const double *>(
                                      &m.bins[mm])[k - kMergedLds];
      }
    }
    __syncthreads();
    if (mb && tid < ADM_MAX_BEAMS)
      sbeams[tid].alpha = member_alpha(sbeams[tid], mb->beam_gain[tid],
                                       a.beam_pmod[tid], a.beam_den[tid],
                                       kc.log_01);
    if (mb) __syncthreads();
  }
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int j = lane / NL, L = lane % NL;
  const int la = L % N, lb = L / N;
  const int own = N * la + NL * lb;
  double *xb = lds + (PK == 2 ? n_mat * kPlaneMergedLds : 0) +
               (wave * CPW + (j < CPW ? j : 0)) * (4 * N3);
  double *xb2 = xb + N3, *xb3 = xb + 2 * N3, *xb4 = xb + 3 * N3;
  const int r = blockIdx.x * 4 + wave;
  double pe = 0.0, ps = 0.0, pr = 0.0;
  bool ok = r < a.n_rows;
  RowCell rc;
  if (ok) {
    rc = row_cell(a, r, j, CPW);
    ok = rc.ok;
  }
  int64_t c = 0;
  uint32_t meta = 0;
  if (ok) {
    c = rc.cc[0] + (int64_t)a.nc[0] * (rc.cc[1] + (int64_t)a.nc[1] * rc.cc[2]);
    meta = meta_p[c];
    ok = (meta & kActive) != 0;
  }
  if (ok) {
    const int64_t sy = a.nn[0], sz = (int64_t)a.nn[0] * a.nn[1];
    const int64_t base = (int64_t)rc.cc[0] * P + sy * ((int64_t)rc.cc[1] * P + la) +
                         sz * ((int64_t)rc.cc[2] * P + lb);
    const int mid = (meta >> kMaterialShift) & 3;
    // two fields: values at the points by S along x, y, z
    double V[2][N];
    {
      double u0[N], u1[N], t0[N], t1[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        u0[i] = T_p[base + i];
        u1[i] = mu_p[base + i];
      }
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double s0 = kc.S[q][0] * u0[0], s1 = kc.S[q][0] * u1[0];
#pragma unroll
        for (int k = 1; k < N; ++k) {
          s0 = __builtin_fma(kc.S[q][k], u0[k], s0);
          s1 = __builtin_fma(kc.S[q][k], u1[k], s1);
        }
        t0[q] = s0;
        t1[q] = s1;
      }
#pragma unroll
      for (int q = 0; q < N; ++q) {
        xb[own + q] = t0[q];
        xb2[own + q] = t1[q];
      }
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double s0 = kc.S[la][0] * xb[NL * lb + q];
        double s1 = kc.S[la][0] * xb2[NL * lb + q];
#pragma unroll
        for (int k = 1; k < N; ++k) {
          s0 = __builtin_fma(kc.S[la][k], xb[N * k + NL * lb + q], s0);
          s1 = __builtin_fma(kc.S[la][k], xb2[N * k + NL * lb + q], s1);
        }
        t0[q] = s0;
        t1[q] = s1;
      }
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) {
        xb[own + q] = t0[q];
        xb2[own + q] = t1[q];
      }
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double s0 = kc.S[lb][0] * xb[N * la + q];
        double s1 = kc.S[lb][0] * xb2[N * la + q];
#pragma unroll
        for (int k = 1; k < N; ++k) {
          s0 = __builtin_fma(kc.S[lb][k], xb[N * la + NL * k + q], s0);
          s1 = __builtin_fma(kc.S[lb][k], xb2[N * la + NL * k + q], s1);
        }
        V[0][q] = s0;
        V[1][q] = s1;
      }
    }
    // physical gradients of both fields
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < N; ++q) {
      xb[own + q] = V[0][q];
      xb2[own + q] = V[1][q];
    }
    wave_lds_sync();
    double G[2][3][N];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const double *buf = f == 0 ? xb : xb2;
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double ax = kc.Dc[q][0] * V[f][0];
        double ay = kc.Dc[la][0] * buf[NL * lb + q];
        double az = kc.Dc[lb][0] * buf[N * la + q];
#pragma unroll
        for (int k = 1; k < N; ++k) {
          ax = __builtin_fma(kc.Dc[q][k], V[f][k], ax);
          ay = __builtin_fma(kc.Dc[la][k], buf[N * k + NL * lb + q], ay);
          az = __builtin_fma(kc.Dc[lb][k], buf[N * la + NL * k + q], az);
        }
        G[f][0][q] = ax * kc.inv_h[0];
        G[f][1][q] = ay * kc.inv_h[1];
        G[f][2][q] = az * kc.inv_h[2];
      }
    }
    // per-point coefficients
    const double sol = m.prop[mid][ADM_SOLIDUS], liq = m.prop[mid][ADM_LIQUIDUS];
    const double latent = m.latent_factor[mid];
    const int64_t slot =
        (meta & kHasPowder) ? (int64_t)pslot_p[c] : -1;
    const double *pw = slot >= 0 ? pwo_p + slot * N3 + own : nullptr;
    double *pb = (slot >= 0 && pwb_p) ? pwb_p + slot * N3 + own
                                              : nullptr;
    const bool rotated = (meta & kRotated) != 0;
    double cs = 1.0, sn = 0.0;
    if (rotated) {
      cs = dcos_p[c];
      sn = dsin_p[c];
    }
    const double *mB = lds + mid * kPlaneMergedLds;
    const double *mC = mB + kMergedMax;
    const MergedBins &mbin =
        *reinterpret_cast<const MergedBins *>(mB + kMergedLds);
    const bool iso = m.isotropic[mid] != 0;
    const double py = kc.origin[1] + (rc.cc[1] + kc.xq[la]) * kc.h[1];
    const double pz = kc.origin[2] + (rc.cc[2] + kc.xq[lb]) * kc.h[2];
    double cT[N], cw[3][N];
    // wave-uniform fast path (as the forward kernels): every point of the
    // wavefront fully solid (one material, no powder, below the solidus).
    // The mixtures then reduce to the solid state's value and slope bit for
    // bit (the other states enter with zero ratio and zero ratio slope),
    // which the one-state evaluators give.
    const int mid0 = __builtin_amdgcn_readfirstlane(mid);
    bool warm = false;
    if constexpr (NS > 1) {
#pragma unroll
      for (int q = 0; q < N; ++q) warm |= !(V[0][q] < sol);
    }
    const bool fast = !__any(warm || slot >= 0 || mid != mid0);
    auto point = [&](auto solid_tag, auto q_tag) {
      constexpr bool SOLID = decltype(solid_tag)::value;
      constexpr int q = decltype(q_tag)::value;
      const double T = V[0][q], mu = V[1][q];
      double dl = 0.0;
      bool clamped = false;
      double irc, ircT, ircp = 0.0;
      double kx, kxT, kxp = 0.0, ky, kyT, kyp = 0.0, kz, kzT, kzp = 0.0;
      // the merged-table interval of T, once for every property
      const int kt = PK == 2 ? merged_bin_interval(mB, mbin, T) : 0;
      auto props = [&](const auto &x, auto ns_tag) {
        constexpr int NSX = decltype(ns_tag)::value;
        inv_rho_cp_lin<NSX, PK>(m, mid, mB, mC, T, kt, latent, x, irc, ircT,
                                ircp);
        prop_vd<NSX, PK>(m, mid, mB, mC, T, kt, ADM_THERMAL_CONDUCTIVITY_X, x,
                         kx, kxT, kxp);
        if (iso) {
          ky = kz = kx;
          kyT = kzT = kxT;
          kyp = kzp = kxp;
        } else {
          prop_vd<NSX, PK>(m, mid, mB, mC, T, kt, ADM_THERMAL_CONDUCTIVITY_Y,
                           x, ky, kyT, kyp);
          prop_vd<NSX, PK>(m, mid, mB, mC, T, kt, ADM_THERMAL_CONDUCTIVITY_Z,
                           x, kz, kzT, kzp);
        }
      };
      if constexpr (SOLID) {
        Mix<1> x1;
        x1.r[0] = 1.0;
        x1.rT[0] = 0.0;
        x1.rp[0] = 0.0;
        props(x1, std::integral_constant<int, 1>{});
      } else {
        Mix<NS> x;
        ratios_lin<NS>(T, pw ? pw[q] : 0.0, sol, liq, x, dl, clamped);
        props(x, std::integral_constant<int, NS>{});
      }
      const double g0 = G[0][0][q], g1 = G[0][1][q], g2 = G[0][2][q];
      const double m0 = G[1][0][q], m1 = G[1][1][q], m2 = G[1][2][q];
      // mu-gradient . K g for the conductivities (kx, ky, kz) with the x-y
      // rotation by the deposition angle
      auto mKg_of = [&](double ax, double ay, double az) {
        const double kxx = ax * cs * cs + ay * sn * sn;
        const double kyy = ax * sn * sn + ay * cs * cs;
        const double kxy = (ax - ay) * sn * cs;
        return m0 * (kxx * g0 + kxy * g1) + m1 * (kxy * g0 + kyy * g1) +
               m2 * (az * g2);
      };
      const double kxx = kx * cs * cs + ky * sn * sn;
      const double kyy = kx * sn * sn + ky * cs * cs;
      const double kxy = (kx - ky) * sn * cs;
      const double Km0 = kxx * m0 + kxy * m1, Km1 = kxy * m0 + kyy * m1,
                   Km2 = kz * m2;
      double wxyz = 1.0;
      wxyz *= kc.wq[q];
      wxyz *= kc.wq[la];
      wxyz *= kc.wq[lb];
      const double jxw = kc.det * wxyz;
      const double mKg = mKg_of(kx, ky, kz);
      cw[0][q] = ((-irc) * Km0) * jxw * kc.inv_h[0];
      cw[1][q] = ((-irc) * Km1) * jxw * kc.inv_h[1];
      cw[2][q] = ((-irc) * Km2) * jxw * kc.inv_h[2];
      // partials at fixed carried powder (T) and with respect to it (p)
      double c_T = -jxw * (ircT * mKg + irc * mKg_of(kxT, kyT, kzT));
      double c_p = 0.0;
      if constexpr (NS == 3 && !SOLID)
        if (pw) c_p = -jxw * (ircp * mKg + irc * mKg_of(kxp, kyp, kzp));
      ps += -irc * jxw * mKg;
      if (a.sources_on) {
        const double px = kc.origin[0] + (rc.cc[0] + kc.xq[q]) * kc.h[0];
        double qs = 0.0, q0 = 0.0;
#pragma unroll 1
        for (int b = 0; b < a.n_beams; ++b) {
          const double v = beam_value(sbeams[b], px, py, pz, kc.log_01);
          if (b == 0) q0 = v;
          qs += v;
        }
        c_T += mu * ircT * qs * jxw;
        if constexpr (NS == 3 && !SOLID) c_p += mu * ircp * qs * jxw;
        pe += mu * irc * q0 * inv_eta0 * jxw;
        pr += mu * irc * beam_value_dR(sbeams[0], q0, px, py, pz, kc.log_01) * jxw;
      }
      if constexpr (NS == 3 && !SOLID) {
        if (pw) {
          // the carried powder after this evaluation: its adjoint (later
          // evaluations) plus this evaluation's dependence on it goes to T
          // where it was clamped to 1 - liquid, else to the powder before
          const double tot = c_p + (pb ? pb[q] : 0.0);
          if (clamped) c_T -= dl * tot;
          if (pb) pb[q] = clamped ? 0.0 : tot;
        }
      }
      cT[q] = c_T;
    };
    if (fast)
      static_for<0, N>([&](auto q) {
        point(std::integral_constant<bool, true>{}, q);
      });
    else
      static_for<0, N>([&](auto q) {
        point(std::integral_constant<bool, false>{}, q);
      });
    // integration as in cell_kernel
    double t[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double acc = kc.Dc[0][i] * cw[0][0];
#pragma unroll
      for (int q = 1; q < N; ++q) acc = __builtin_fma(kc.Dc[q][i], cw[0][q], acc);
      t[i] = acc;
    }
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < N; ++q) {
      xb3[own + q] = cw[1][q];
      xb4[own + q] = cw[2][q];
    }
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double ay = kc.Dc[0][la] * xb3[NL * lb + q];
      double az = kc.Dc[0][lb] * xb4[N * la + q];
#pragma unroll
      for (int k = 1; k < N; ++k) {
        ay = __builtin_fma(kc.Dc[k][la], xb3[N * k + NL * lb + q], ay);
        az = __builtin_fma(kc.Dc[k][lb], xb4[N * la + NL * k + q], az);
      }
      t[q] = t[q] + ay + az + cT[q];
    }
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < N; ++q) xb3[own + q] = t[q];
    wave_lds_sync();
    double u[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double acc = kc.S[0][lb] * xb3[N * la + q];
#pragma unroll
      for (int k = 1; k < N; ++k)
        acc = __builtin_fma(kc.S[k][lb], xb3[N * la + NL * k + q], acc);
      u[q] = acc;
    }
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < N; ++q) xb3[own + q] = u[q];
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double acc = kc.S[0][la] * xb3[NL * lb + q];
#pragma unroll
      for (int k = 1; k < N; ++k)
        acc = __builtin_fma(kc.S[k][la], xb3[N * k + NL * lb + q], acc);
      t[q] = acc;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double acc = kc.S[0][i] * t[0];
#pragma unroll
      for (int q = 1; q < N; ++q) acc = __builtin_fma(kc.S[q][i], t[q], acc);
      // colour launches: no other cell of the launch holds the node
      if (a.color >= 0)
        out_p[base + i] += acc;
      else
        unsafeAtomicAdd(out_p + base + i, acc);
    }
  }
  // parameter partials: wavefront sums, then the block's sum (waves in
  // order) into its slot. Same-address FP64 atomics from every wavefront
  // serialise at the memory side and slowed the whole kernel (the node
  // accumulation's atomics with them) by 1.7x.
  __shared__ double wsum[4][3];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    pe += __shfl_xor(pe, off);
    ps += __shfl_xor(ps, off);
    pr += __shfl_xor(pr, off);
  }
  if (lane == 0) {
    wsum[wave][0] = pe;
    wsum[wave][1] = ps;
    wsum[wave][2] = pr;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = wsum[0][threadIdx.x];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v += wsum[w][threadIdx.x];
    bpart_p[threadIdx.x] = v;
  }
}

// partials[3 j + k] += the sum over the blocks of member j (blockIdx.x) of
// block_part, in a fixed order: lane l adds blocks l, l + 64, ..., then a
// fixed butterfly over the lanes.
__global__ void __launch_bounds__(64)
    reduce_partials_kernel(const double *__restrict__ block_part,
                           int32_t part_blocks, double *__restrict__ partials) {
  const int lane = threadIdx.x;
  const double *bp = block_part + 3 * (int64_t)blockIdx.x * part_blocks;
  double v[3] = {0.0, 0.0, 0.0};
  for (int b = lane; b < part_blocks; b += 64)
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] += bp[3 * (int64_t)b + k];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] += __shfl_xor(v[k], off);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) partials[3 * blockIdx.x + k] += v[k];
}

// (dA_face/dT)^T mu over the boundary faces, with the face powder recorded
// before the evaluation and the adjoint of the carried face powder
// (face_powder_bar, NULL: zero, not stored) as in the cell kernel. Lanes as
// face_kernel: (p+1)^2 per face, lane q loading node q, evaluating point q
// (T and mu by the same S chains as the forward face kernel) and, after the
// points' terms are shared, integrating node q.
template <int P, int NS, int PK>
__global__ void __launch_bounds__(256)
    adjoint_face_kernel(const FaceArgs a, const double *__restrict__ T,
                        const double *__restrict__ mu,
                        const double *__restrict__ face_powder_old,
                        double *__restrict__ face_powder_bar,
                        double *__restrict__ out, const AdjFaceArgs af) {
  constexpr int N = P + 1, NF = N * N;
  constexpr int FPW = FaceLanes<P>::faces_per_wave;
  __shared__ double xf[4][2][FPW * NF];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int jf = lane / NF, q = lane % NF;
  const int qa = q % N, qb = q / N;
  // one operator, or member blockIdx.y of an ensemble
  const ADM_CONSTANT MemberArgs *mb =
      a.members ? as_constant(a.members) + blockIdx.y : nullptr;
  const int64_t fi = ((int64_t)blockIdx.x * 4 + wave) * FPW + jf;
  // whole faces (all lanes of a face) agree
  if (jf >= FPW || fi >= (mb ? mb->n_faces : a.n_faces)) return;
  if (mb) {
    const int64_t voff = (int64_t)blockIdx.y * a.member_stride;
    const int64_t foff = af.fpw_off[blockIdx.y];
    T += voff;
    mu += voff;
    out += voff;
    if (face_powder_old) face_powder_old += foff;
    if (face_powder_bar) face_powder_bar += foff;
  }
  const KConst &kc = *(const KConst *)as_global(mb ? mb->kc : a.kc);
  const MatConst &m = kc.mat;
  const FaceRef *faces = mb ? mb->faces : a.faces;
  const int64_t c = faces[fi].cell;
  const int info = faces[fi].info;
  const int f = info & 7, bt = (info >> 3) & 31, mid = (info >> 8) & 3;
  const int d = f / 2, side = f % 2;
  const int cc[3] = {(int)(c % a.nc[0]), (int)((c / a.nc[0]) % a.nc[1]),
                     (int)(c / ((int64_t)a.nc[0] * a.nc[1]))};
  const int64_t ns[3] = {1, (int64_t)a.nn[0],
                         (int64_t)a.nn[0] * (int64_t)a.nn[1]};
  const int64_t base = (int64_t)cc[0] * P + ns[1] * cc[1] * P +
                       ns[2] * cc[2] * P + (side ? P : 0) * ns[d];
  const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
  const int64_t node = base + qa * ns[t1] + qb * ns[t2];
  double *fx = &xf[wave][0][jf * NF], *fy = &xf[wave][1][jf * NF];
  fx[q] = T[node];
  fy[q] = mu[node];
  wave_lds_sync();
  // T and mu at point (qa, qb): along a, then along b
  double ft[N], fm[N];
#pragma unroll
  for (int bb = 0; bb < N; ++bb) {
    double s0 = kc.S[qa][0] * fx[N * bb], s1 = kc.S[qa][0] * fy[N * bb];
#pragma unroll
    for (int k = 1; k < N; ++k) {
      s0 = __builtin_fma(kc.S[qa][k], fx[k + N * bb], s0);
      s1 = __builtin_fma(kc.S[qa][k], fy[k + N * bb], s1);
    }
    ft[bb] = s0;
    fm[bb] = s1;
  }
  double Tq = kc.S[qb][0] * ft[0], gq = kc.S[qb][0] * fm[0];
#pragma unroll
  for (int k = 1; k < N; ++k) {
    Tq = __builtin_fma(kc.S[qb][k], ft[k], Tq);
    gq = __builtin_fma(kc.S[qb][k], fm[k], gq);
  }
  const double sol = m.prop[mid][ADM_SOLIDUS], liq = m.prop[mid][ADM_LIQUIDUS];
  const double latent = m.latent_factor[mid];
  const double *mB = &m.merged[mid].B[0];
  const double *mC = &m.merged[mid].coef[0][0][0][0];
  const double *fp = NS == 3 ? face_powder_old + fi * NF : nullptr;
  double *fpb = (NS == 3 && face_powder_bar) ? face_powder_bar + fi * NF
                                             : nullptr;
  Mix<NS> x;
  double dl;
  bool clamped;
  ratios_lin<NS>(Tq, fp ? fp[q] : 0.0, sol, liq, x, dl, clamped);
  double irc, ircT, ircp;
  const int kt = PK == 2 ? merged_search(mB, Tq) : 0;
  inv_rho_cp_lin<NS, PK>(m, mid, mB, mC, Tq, kt, latent, x, irc, ircT, ircp);
  // s = h (T - T_conv) + eps sigma (T + T_rad)(T^2 + T_rad^2)(T - T_rad)
  double s1 = 0.0, s1T = 0.0, s1p = 0.0;
  if (bt & ADM_BC_CONVECTIVE) {
    const double cT = m.prop[mid][ADM_CONVECTION_TEMPERATURE_INFTY];
    double h, hT, hp;
    prop_vd<NS, PK>(m, mid, mB, mC, Tq, kt, ADM_CONVECTION_HEAT_TRANSFER_COEF,
                    x, h, hT, hp);
    s1 += h * (Tq - cT);
    s1T += hT * (Tq - cT) + h;
    s1p += hp * (Tq - cT);
  }
  if (bt & ADM_BC_RADIATIVE) {
    const double rT = m.prop[mid][ADM_RADIATION_TEMPERATURE_INFTY];
    double e, eT, ep;
    prop_vd<NS, PK>(m, mid, mB, mC, Tq, kt, ADM_EMISSIVITY, x, e, eT, ep);
    const double a1 = Tq + rT, a2 = Tq * Tq + rT * rT;
    const double g = kStefanBoltzmann * a1 * a2 * (Tq - rT);
    const double gT = kStefanBoltzmann *
                      (a2 * (Tq - rT) + a1 * 2.0 * Tq * (Tq - rT) + a1 * a2);
    s1 += e * g;
    s1T += eT * g + e * gT;
    s1p += ep * g;
  }
  const double w = gq * kc.fjxw[d][q];
  double c_T = (-ircT * s1 - irc * s1T) * w;
  if constexpr (NS == 3) {
    const double tot = (-ircp * s1 - irc * s1p) * w + (fpb ? fpb[q] : 0.0);
    if (clamped) c_T -= dl * tot;
    if (fpb) fpb[q] = clamped ? 0.0 : tot;
  }
  wave_lds_sync();
  fx[q] = c_T;
  wave_lds_sync();
  // node (qa, qb): S^T along b, then along a
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double acc = kc.S[0][qb] * fx[i];
#pragma unroll
    for (int k = 1; k < N; ++k)
      acc = __builtin_fma(kc.S[k][qb], fx[i + N * k], acc);
    ft[i] = acc;
  }
  double o = kc.S[0][qa] * ft[0];
#pragma unroll
  for (int i = 1; i < N; ++i) o = __builtin_fma(kc.S[i][qa], ft[i], o);
  unsafeAtomicAdd(out + node, o);
}

// ---------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------

// Up to 32 device-to-device copies in one launch (grid.y: the copy).
constexpr int kSegCopies = 32;
struct SegCopyArgs {
  const double *src[kSegCopies];
  double *dst[kSegCopies];
  int64_t len[kSegCopies];
};
__global__ void __launch_bounds__(256)
    seg_copy_kernel(const SegCopyArgs a) {
  const int k = blockIdx.y;
  const double *__restrict__ src = a.src[k];
  double *__restrict__ dst = a.dst[k];
  const int64_t len = a.len[k];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// device copies (src, dst, length in doubles) in launches of kSegCopies:
// one kernel launch costs the host less than one runtime copy call, so a
// sequence of short copies leaves the device fewer idle gaps
using SegCopy = std::tuple<const double *, double *, int64_t>;
inline void launch_seg_copies(const std::vector<SegCopy> &copies,
                              hipStream_t s) {
  for (size_t b = 0; b < copies.size(); b += kSegCopies) {
    SegCopyArgs a;
    std::memset(&a, 0, sizeof(a));
    const int m = (int)std::min<size_t>(kSegCopies, copies.size() - b);
    int64_t longest = 0;
    for (int k = 0; k < m; ++k) {
      a.src[k] = std::get<0>(copies[b + k]);
      a.dst[k] = std::get<1>(copies[b + k]);
      a.len[k] = std::get<2>(copies[b + k]);
      longest = std::max(longest, a.len[k]);
    }
    if (longest == 0) continue;
    const unsigned gx =
        (unsigned)std::min<int64_t>((longest + 255) / 256, 2048);
    hipLaunchKernelGGL(seg_copy_kernel, dim3(gx, (unsigned)m), dim3(256), 0,
                       s, a);
  }
}

// faces per block of the adjoint face kernel (FaceLanes<P>::faces_per_block)
inline int FaceLanes_per_block(int p) {
  return 4 * (64 / ((p + 1) * (p + 1)));
}

// The vector work of a backward step on the first n_act entries of each
// member's block of n (grid.y: the member): the nodes above the highest
// active cell layer have a zero inverse mass, so their stage adjoints stay
// zero and their T adjoint keeps its value.

// K_i = (dt b_i) Tb for the S stages
__global__ void adj_step_in_kernel(double *__restrict__ K0,
                                   double *__restrict__ K1,
                                   double *__restrict__ K2,
                                   double *__restrict__ K3, double c0,
                                   double c1, double c2, double c3, int S,
                                   const double *__restrict__ Tb, int64_t n,
                                   int64_t n_act) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_act) return;
  const int64_t k = (int64_t)blockIdx.y * n + i;
  const double tb = Tb[k];
  K0[k] = c0 * tb;
  if (S > 1) K1[k] = c1 * tb;
  if (S > 2) K2[k] = c2 * tb;
  if (S > 3) K3[k] = c3 * tb;
}

// mu = 1 * K_s * inv_mass (M^-1 times the stage adjoint); Ybar = 0
__global__ void adj_stage_in_kernel(double *__restrict__ mu,
                                    double *__restrict__ Ybar,
                                    const double *__restrict__ Ks,
                                    const double *__restrict__ inv_mass,
                                    int64_t n, int64_t n_act) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_act) return;
  const int64_t k = (int64_t)blockIdx.y * n + i;
  mu[k] = 1.0 * Ks[k] * inv_mass[k];
  Ybar[k] = 0.0;
}

// Tb += Ybar, then K_j += c_j Ybar for the given K_j (the stage input's
// adjoint to the earlier stages)
__global__ void adj_stage_out_kernel(double *__restrict__ Tb,
                                     const double *__restrict__ Ybar,
                                     double *__restrict__ K0,
                                     double *__restrict__ K1,
                                     double *__restrict__ K2, double c0,
                                     double c1, double c2, int64_t n,
                                     int64_t n_act) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_act) return;
  const int64_t k = (int64_t)blockIdx.y * n + i;
  const double y = Ybar[k];
  Tb[k] = 1. * Tb[k] + 1.0 * y;
  if (K0) K0[k] = 1. * K0[k] + c0 * y;
  if (K1) K1[k] = 1. * K1[k] + c1 * y;
  if (K2) K2[k] = 1. * K2[k] + c2 * y;
}

// ---------------------------------------------------------------------------
// driver
// ---------------------------------------------------------------------------

struct DevBuf {
  double *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  void alloc(size_t n_) {
    if (n_ == 0) return;
    ADM_HIP_ADJ(hipMalloc(&p, sizeof(double) * n_));
    n = n_;
  }
  // at least n_ doubles (the contents are not kept)
  void grow(size_t n_) {
    if (n_ <= n) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    alloc(n_);
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// The vector kernels of one backward step of the explicit method
// (a, b: its tableau; K: its stage adjoints), over `members` blocks of n
// entries of which the first n_act are touched; vjp(i) runs stage i's
// transposed operator on mu into Ybar.
template <class Vjp>
void backward_step(int S, const double (&a)[4][4], const double (&b)[4],
                   double dt, DevBuf (&K)[4], double *Tb, double *mu,
                   double *Ybar, const double *inv_mass, int64_t n,
                   int64_t n_act, int members, hipStream_t s, Vjp &&vjp) {
  const dim3 g((unsigned)((n_act + 255) / 256), (unsigned)members);
  // stage adjoints start at dt b_i Tb; Tb_new = Tb + sum Ybar_i;
  // K_j += dt a_ij Ybar_i for j < i (reverse stage order)
  hipLaunchKernelGGL(adj_step_in_kernel, g, dim3(256), 0, s, K[0].p, K[1].p,
                     K[2].p, K[3].p, dt * b[0], dt * b[1], dt * b[2],
                     dt * b[3], S, Tb, n, n_act);
  for (int i = S - 1; i >= 0; --i) {
    hipLaunchKernelGGL(adj_stage_in_kernel, g, dim3(256), 0, s, mu, Ybar,
                       K[i].p, inv_mass, n, n_act);
    vjp(i);
    double *kj[3] = {nullptr, nullptr, nullptr};
    double cj[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < i; ++j)
      if (a[i][j] != 0.0) {
        kj[j] = K[j].p;
        cj[j] = dt * a[i][j];
      }
    hipLaunchKernelGGL(adj_stage_out_kernel, g, dim3(256), 0, s, Tb, Ybar,
                       kj[0], kj[1], kj[2], cj[0], cj[1], cj[2], n, n_act);
  }
  ADM_HIP_ADJ(hipGetLastError());
}

// nodes in the cell layers up to nz (and the node layer above them)
inline int64_t active_prefix(const Context &c, int nz) {
  const int64_t layers = std::min<int64_t>(c.nn[2], (int64_t)c.p * nz + 1);
  return (int64_t)c.nn[0] * c.nn[1] * layers;
}

// vals[j n_obs + k] = T[j n + dof[k]] (grid.y: member j)
__global__ void gather_obs_kernel(const double *__restrict__ T, int64_t n,
                                  const int64_t *__restrict__ dof, int n_obs,
                                  double *__restrict__ vals) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_obs) return;
  vals[(int64_t)blockIdx.y * n_obs + k] = T[(int64_t)blockIdx.y * n + dof[k]];
}

// Tb[j n + dof[u]] = val[j n_u + u] (grid.y: member j)
__global__ void scatter_obs_kernel(double *__restrict__ Tb, int64_t n,
                                   const int64_t *__restrict__ dof, int n_u,
                                   const double *__restrict__ val) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n_u) return;
  Tb[(int64_t)blockIdx.y * n + dof[u]] = val[(int64_t)blockIdx.y * n_u + u];
}

// The misfits J_j = 1/2 sum_k w_k (T_j[obs_dof_k] - obs_jk)^2 of the
// `members` blocks of n in T (obs: members x n_obs), and Tb = dJ/dT: zero
// but at the observed nodes, where each member's weighted residuals are
// summed in observation order. Only the observed values cross the bus.
inline void misfit_gradient(const double *T, int members, int64_t n,
                            int n_obs, const int64_t *obs_dof,
                            const double *obs, const double *obs_weight,
                            double *J, double *Tb, hipStream_t s) {
  // the distinct observed nodes, in order of first appearance
  std::vector<int64_t> udof;
  std::vector<int> slot((size_t)n_obs);
  {
    std::vector<std::pair<int64_t, int>> order((size_t)n_obs);
    for (int k = 0; k < n_obs; ++k) order[k] = {obs_dof[k], k};
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<int64_t, int> &x,
                        const std::pair<int64_t, int> &y) {
                       return x.first < y.first;
                     });
    for (size_t i = 0; i < order.size(); ++i) {
      if (i == 0 || order[i].first != order[i - 1].first)
        udof.push_back(order[i].first);
      slot[order[i].second] = (int)udof.size() - 1;
    }
  }
  const int n_u = (int)udof.size();
  DevBuf d_dof, d_udof, d_vals;  // int64 tables in double-sized slots
  d_dof.alloc((size_t)n_obs);
  d_udof.alloc((size_t)n_u);
  d_vals.alloc((size_t)members * std::max(n_obs, n_u));
  int64_t *dof_p = reinterpret_cast<int64_t *>(d_dof.p);
  int64_t *udof_p = reinterpret_cast<int64_t *>(d_udof.p);
  ADM_HIP_ADJ(hipMemcpyAsync(dof_p, obs_dof, sizeof(int64_t) * n_obs,
                             hipMemcpyHostToDevice, s));
  ADM_HIP_ADJ(hipMemcpyAsync(udof_p, udof.data(), sizeof(int64_t) * n_u,
                             hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(gather_obs_kernel,
                     dim3((unsigned)((n_obs + 255) / 256), (unsigned)members),
                     dim3(256), 0, s, T, n, dof_p, n_obs, d_vals.p);
  std::vector<double> vals((size_t)members * n_obs);
  ADM_HIP_ADJ(hipMemcpyAsync(vals.data(), d_vals.p,
                             sizeof(double) * vals.size(),
                             hipMemcpyDeviceToHost, s));
  ADM_HIP_ADJ(hipStreamSynchronize(s));
  std::vector<double> uval((size_t)members * n_u, 0.0);
  for (int j = 0; j < members; ++j) {
    double Jv = 0.0;
    for (int k = 0; k < n_obs; ++k) {
      const double w = obs_weight ? obs_weight[k] : 1.0;
      const double r = vals[(size_t)j * n_obs + k] - obs[(size_t)j * n_obs + k];
      Jv += 0.5 * w * r * r;
      uval[(size_t)j * n_u + slot[k]] += w * r;
    }
    J[j] = Jv;
  }
  ADM_HIP_ADJ(hipMemsetAsync(Tb, 0, sizeof(double) * n * members, s));
  ADM_HIP_ADJ(hipMemcpyAsync(d_vals.p, uval.data(),
                             sizeof(double) * uval.size(),
                             hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(scatter_obs_kernel,
                     dim3((unsigned)((n_u + 255) / 256), (unsigned)members),
                     dim3(256), 0, s, Tb, n, udof_p, n_u, d_vals.p);
  ADM_HIP_ADJ(hipGetLastError());
  // the host vectors above are read by the copies before they return
  ADM_HIP_ADJ(hipStreamSynchronize(s));
}

// beam states of the context at time t (segment search from the start)
inline void eval_time_reset(Context &c, double t) {
  for (size_t b = 0; b < c.old_segment.size(); ++b) c.old_segment[b] = 0;
  c.set_time(t);
}

// out += (dA/dT)^T mu at the stage input Y with the phase state (pw, fpw)
// recorded before the evaluation, at the context's current time; the
// adjoints of the carried cell and face powder (pb, fpb; NULL: zero) are
// taken from after to before the evaluation, and the parameter partials are
// added to partials[3].
// An ensemble launch: member j (grid y) reads its vectors at j * stride,
// its powder arrays at pw_off[j] / fpw_off[j] (device), its 1/eta_0 at
// inv_eta[j] (device) and writes its partials at partials + 3 j.
struct VjpBatch {
  const MemberArgs *members;
  int n_members;
  int64_t stride;
  double pmod[ADM_MAX_BEAMS], den[ADM_MAX_BEAMS];
  const int64_t *pw_off, *fpw_off;
  const double *inv_eta;
  int64_t max_faces;
  int nz_active;  // the largest over the members
};

inline void launch_vjp(Context &c, const double *Y, const double *pw,
                       const double *fpw, double *pb, double *fpb,
                       const double *mu, double *out, double *partials,
                       double inv_eta0, DevBuf &bpart, hipStream_t s,
                       const VjpBatch *batch = nullptr) {
  const int pk = c.property_kind();
  if (pk == 1)
    throw AdmError(ADM_ERR_UNSUPPORTED,
                   "the adjoint supports polynomial and merged-table "
                   "properties");
  AdjArgs aa;
  std::memset(&aa, 0, sizeof(aa));
  aa.c = c.cell_args();
  const int cpw = cells_per_wave(c.p);
  aa.T = Y;
  aa.mu = mu;
  aa.out = out;
  aa.powder_old = pw;
  aa.powder_bar = pb;
  aa.partials = partials;
  aa.inv_eta0 = inv_eta0;
  FaceArgs fa = c.face_args();
  AdjFaceArgs af{nullptr};
  unsigned ny = 1;
  if (batch) {
    aa.c.members = batch->members;
    aa.c.member_stride = batch->stride;
    for (int b = 0; b < ADM_MAX_BEAMS; ++b) {
      aa.c.beam_pmod[b] = batch->pmod[b];
      aa.c.beam_den[b] = batch->den[b];
    }
    aa.pw_off = batch->pw_off;
    aa.inv_eta = batch->inv_eta;
    fa.members = batch->members;
    fa.member_stride = batch->stride;
    fa.n_faces = batch->max_faces;
    af.fpw_off = batch->fpw_off;
    ny = (unsigned)batch->n_members;
  }
  // the cell layers that can hold an active cell
  const int nz = batch ? batch->nz_active : c.nz_active;
  const int lim[3] = {c.nc[0], c.nc[1], std::min(nz, c.nc[2])};
  const size_t lds =
      (pk == 2 ? sizeof(double) * kPlaneMergedLds * c.kc.mat.n_materials
               : 0) +
      sizeof(double) * 4 * cpw *
          4 * (c.p + 1) * (c.p + 1) * (c.p + 1);
  const unsigned face_blocks = (unsigned)(
      (fa.n_faces + FaceLanes_per_block(c.p) - 1) / FaceLanes_per_block(c.p));
  // one launch with atomic adds, or (ADM_ADJOINT_LAUNCH=colored in the
  // environment) eight launches, one per parity colour, each adding its
  // nodes' terms with a plain read-add-write (cells of one colour share no
  // node). Measured: equal on the 32-member 96x96x40 workload (0.30 s per
  // 50-step gradient either way), 1.4x slower on one 64x32x16 operator
  // (the colour launches are too small to fill the chip).
  bool colored = false;
  if (const char *e = std::getenv("ADM_ADJOINT_LAUNCH"))
    colored = std::string(e) == "colored";
  const int n_launch = colored ? 8 : 1;
  AdjArgs ac[8];
  unsigned blocks[8];
  for (int col = 0; col < n_launch; ++col) {
    ac[col] = aa;
    ac[col].c.color = colored ? col : -1;
    int64_t cells = 1;
    for (int d = 0; d < 3; ++d) {
      const int b = (col >> d) & 1;
      ac[col].c.lext[d] =
          colored ? (lim[d] > b ? (lim[d] - b + 1) / 2 : 0) : lim[d];
      cells *= ac[col].c.lext[d];
    }
    ac[col].c.nseg = (ac[col].c.lext[0] + cpw - 1) / cpw;
    ac[col].c.n_rows = ac[col].c.nseg * ac[col].c.lext[1] * ac[col].c.lext[2];
    blocks[col] = cells ? (unsigned)((ac[col].c.n_rows + 3) / 4) : 0u;
  }
  // every block's partial sums, launch after launch, then their reduction
  int32_t part_blocks = 0;
  for (int col = 0; col < n_launch; ++col) {
    ac[col].part_base = part_blocks;
    part_blocks += (int32_t)blocks[col];
  }
  bpart.grow((size_t)3 * part_blocks * ny);
  for (int col = 0; col < n_launch; ++col) {
    ac[col].block_part = bpart.p;
    ac[col].part_blocks = part_blocks;
  }
#define ADM_ADJ(P_, NS_, PK_)                                              \
  if (c.p == P_ && c.n_states == NS_ && pk == PK_) {                       \
    for (int col = 0; col < n_launch; ++col)                               \
      if (blocks[col])                                                     \
        hipLaunchKernelGGL((adjoint_cell_kernel<P_, NS_, PK_>),            \
                           dim3(blocks[col], ny), dim3(256), lds, s,       \
                           ac[col]);                                       \
    if (fa.n_faces)                                                        \
      hipLaunchKernelGGL((adjoint_face_kernel<P_, NS_, PK_>),              \
                         dim3(face_blocks, ny), dim3(256), 0, s, fa, Y, mu, \
                         fpw, fpb, out, af);                                \
    launched = true;                                                       \
  }
  bool launched = false;
  ADM_ADJ(1, 1, 0)
  ADM_ADJ(1, 2, 0)
  ADM_ADJ(1, 3, 0)
  ADM_ADJ(2, 1, 0)
  ADM_ADJ(2, 2, 0)
  ADM_ADJ(2, 3, 0)
  ADM_ADJ(1, 1, 2)
  ADM_ADJ(1, 2, 2)
  ADM_ADJ(1, 3, 2)
  ADM_ADJ(2, 1, 2)
  ADM_ADJ(2, 2, 2)
  ADM_ADJ(2, 3, 2)
  ADM_ADJ(3, 1, 0)
  ADM_ADJ(3, 2, 0)
  ADM_ADJ(3, 3, 0)
  ADM_ADJ(3, 1, 2)
  ADM_ADJ(3, 2, 2)
  ADM_ADJ(3, 3, 2)
  ADM_ADJ(4, 1, 0)
  ADM_ADJ(4, 2, 0)
  ADM_ADJ(4, 3, 0)
  ADM_ADJ(4, 1, 2)
  ADM_ADJ(4, 2, 2)
  ADM_ADJ(4, 3, 2)
  ADM_ADJ(5, 1, 0)
  ADM_ADJ(5, 2, 0)
  ADM_ADJ(5, 3, 0)
  ADM_ADJ(5, 1, 2)
  ADM_ADJ(5, 2, 2)
  ADM_ADJ(5, 3, 2)
#undef ADM_ADJ
  if (!launched)
    throw AdmError(ADM_ERR_UNSUPPORTED,
                   "no adjoint kernel for this degree / state count");
  if (part_blocks > 0)
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(ny), dim3(64), 0, s,
                       bpart.p, part_blocks, partials);
  ADM_HIP_ADJ(hipGetLastError());
}

// Butcher tableau of the explicit steppers (explicit_step's stage order)
struct Tableau {
  int ns = 1;
  double a[4][4], b[4], c[4];
  explicit Tableau(int method) {
    std::memset(a, 0, sizeof(a));
    std::memset(b, 0, sizeof(b));
    std::memset(c, 0, sizeof(c));
    if (method == ADM_FORWARD_EULER) {
      ns = 1;
      b[0] = 1.0;
    } else if (method == ADM_RK_FOURTH_ORDER) {
      // classic RK4 (TimeStepping::RK_CLASSIC_FOURTH_ORDER)
      ns = 4;
      const double bb[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
      const double cc[4] = {0.0, 0.5, 0.5, 1.0};
      for (int i = 0; i < 4; ++i) {
        b[i] = bb[i];
        c[i] = cc[i];
      }
      a[1][0] = 0.5;
      a[2][1] = 0.5;
      a[3][2] = 1.0;
    } else if (method == ADM_RK_THIRD_ORDER) {
      // Kutta's third order (TimeStepping::RK_THIRD_ORDER)
      ns = 3;
      const double bb[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
      const double cc[3] = {0.0, 0.5, 1.0};
      for (int i = 0; i < 3; ++i) {
        b[i] = bb[i];
        c[i] = cc[i];
      }
      a[1][0] = 0.5;
      a[2][0] = -1.0;
      a[2][1] = 2.0;
    } else {
      throw AdmError(ADM_ERR_INVALID, "unknown time-stepping method");
    }
  }
};

struct AdjointRun {
  Context &c;
  hipStream_t s;
  int method;
  int64_t n;
  int64_t npw, nfpw;  // powder values of cells and faces
  int n_stages;
  double a[4][4], b[4], cst[4];  // Butcher tableau of the forward method
  DevBuf partials, bpart;
  DevBuf K[4];  // stage adjoints of the backward sweep
  DevBuf pbar, fpbar;  // adjoints of the carried cell and face powder

  AdjointRun(Context &c_, int method_, hipStream_t s_)
      : c(c_), s(s_), method(method_) {
    n = c.n_dofs;
    npw = c.n_powder_cells * c.nq;
    nfpw = (c.n_states == 3) ? c.n_faces * c.nfq : 0;
    const Tableau tb(method);
    n_stages = tb.ns;
    std::memcpy(a, tb.a, sizeof(a));
    std::memcpy(b, tb.b, sizeof(b));
    std::memcpy(cst, tb.c, sizeof(cst));
    for (int i = 0; i < n_stages; ++i) K[i].alloc(n);
    partials.alloc(3);
    pbar.alloc(npw);
    fpbar.alloc(nfpw);
    if (npw) ADM_HIP_ADJ(hipMemsetAsync(pbar.p, 0, sizeof(double) * npw, s));
    if (nfpw)
      ADM_HIP_ADJ(hipMemsetAsync(fpbar.p, 0, sizeof(double) * nfpw, s));
  }

  void eval_time(double t) { eval_time_reset(c, t); }

  // one forward step T -> T' (stage inputs Y_i, stage values K_i); with
  // record != nullptr, stage i's input and the powder before its
  // evaluation go to record[i]
  struct StageRecord {
    double *Y, *pw, *fpw;
  };
  void forward_step(double t, double dt, double *T, const StageRecord *rec) {
    // adm_evolve's step (the gather pass applying the stage updates on the
    // plane path), recording each stage's input and the powder before its
    // evaluation
    std::vector<SegCopy> cp;
    c.step_hooked(method, t, dt, T, s, [&](int i, const double *x) {
      if (!rec) return;
      cp.clear();
      cp.emplace_back(x, rec[i].Y, n);
      if (npw) cp.emplace_back(c.d_powder, rec[i].pw, npw);
      if (nfpw) cp.emplace_back(c.d_face_powder, rec[i].fpw, nfpw);
      launch_seg_copies(cp, s);
    });
    ADM_HIP_ADJ(hipGetLastError());
  }

  void vjp(const StageRecord &rec, double t_eval, const double *mu,
           double *out, double inv_eta0) {
    eval_time(t_eval);
    launch_vjp(c, rec.Y, rec.pw, rec.fpw, pbar.p, fpbar.p, mu, out,
               partials.p, inv_eta0, bpart, s);
  }
};


// ---------------------------------------------------------------------------
// ensemble driver: every member's gradient from one batched forward and one
// batched backward sweep, with material activations inside the window
// ---------------------------------------------------------------------------

struct EnsEvent {
  int member, step;
  std::vector<double> boxes, cos, sin;
  double new_temperature;
};

// the beams of the members at time t (scan-path search from the start)
inline void ens_time(Ensemble &E, double t) {
  Context &c0 = *E.members[0];
  for (size_t b = 0; b < c0.old_segment.size(); ++b) c0.old_segment[b] = 0;
  E.set_time(t);
}

struct EnsAdjoint {
  Ensemble &E;
  hipStream_t s;
  int N;
  int64_t n, nt;  // per member, all members
  Tableau tab;
  int method;
  DevBuf K[4];  // stage adjoints of the backward sweep
  // powder of the cells (fixed slots) and of the faces (capacity per member)
  std::vector<int64_t> pw_off, fpw_off;
  int64_t npw = 0, nfpw = 0;
  DevBuf d_off, d_inv_eta;  // [pw_off | fpw_off] as int64, 1/eta_0
  DevBuf bpart;             // per-block partial sums of the cell term
  VjpBatch batch;

  EnsAdjoint(Ensemble &e, int method_, hipStream_t s_)
      : E(e), s(s_), N(e.n_members), n(e.n_dofs), nt(e.n_dofs * e.n_members),
        tab(method_), method(method_) {
    for (int i = 0; i < tab.ns; ++i) K[i].alloc(nt);
    pw_off.assign(N + 1, 0);
    for (int j = 0; j < N; ++j)
      pw_off[j + 1] = pw_off[j] + E.members[j]->n_powder_cells * E.members[j]->nq;
    npw = pw_off[N];
  }

  // device tables once the face capacities are known
  void finish_tables(const std::vector<int64_t> &fcap) {
    fpw_off.assign(N + 1, 0);
    for (int j = 0; j < N; ++j) fpw_off[j + 1] = fpw_off[j] + fcap[j];
    nfpw = fpw_off[N];
    d_off.alloc(2 * (size_t)N);
    std::vector<double> off(2 * (size_t)N);
    for (int j = 0; j < N; ++j) {
      std::memcpy(&off[j], &pw_off[j], sizeof(int64_t));
      std::memcpy(&off[N + j], &fpw_off[j], sizeof(int64_t));
    }
    ADM_HIP_ADJ(hipMemcpy(d_off.p, off.data(), sizeof(double) * off.size(),
                          hipMemcpyHostToDevice));
    std::vector<double> ie(N);
    for (int j = 0; j < N; ++j) {
      const Context &c = *E.members[j];
      const double eta =
          c.beams.empty() ? 0.0 : c.beams[0].desc.absorption_efficiency;
      ie[j] = eta != 0.0 ? 1.0 / eta : 0.0;
    }
    d_inv_eta.alloc(N);
    ADM_HIP_ADJ(hipMemcpy(d_inv_eta.p, ie.data(), sizeof(double) * N,
                          hipMemcpyHostToDevice));
    batch.n_members = N;
    batch.stride = n;
    batch.pw_off = reinterpret_cast<const int64_t *>(d_off.p);
    batch.fpw_off = reinterpret_cast<const int64_t *>(d_off.p) + N;
    batch.inv_eta = d_inv_eta.p;
  }

  // members' powder <-> a packed copy (cells: fixed slots; faces: the
  // current face count of each member, at its offset)
  // copies (src, dst, length), batched kSegCopies to a launch
  std::vector<SegCopy> copies;
  void flush_copies() {
    launch_seg_copies(copies, s);
    copies.clear();
    ADM_HIP_ADJ(hipGetLastError());
  }
  void copy_powder(double *pw, double *fpw, bool save) {
    for (int j = 0; j < N; ++j) {
      Context &c = *E.members[j];
      const int64_t a = pw_off[j + 1] - pw_off[j];
      if (pw && a)
        copies.emplace_back(save ? c.d_powder : pw + pw_off[j],
                            save ? pw + pw_off[j] : c.d_powder, a);
      const int64_t f = c.n_states == 3 ? c.n_faces * c.nfq : 0;
      if (fpw && f)
        copies.emplace_back(save ? c.d_face_powder : fpw + fpw_off[j],
                            save ? fpw + fpw_off[j] : c.d_face_powder, f);
    }
    flush_copies();
  }

  struct Rec {
    double *Y, *pw, *fpw;
  };
  // one batched step of every member (the stage sequence of
  // AdjointRun::forward_step)
  void forward_step(double t, double dt, double *T, const Rec *rec) {
    // the stage inputs' nodes up to the highest active cell layer (the
    // transposed operator reads no other)
    int nz = 0;
    for (const auto &m : E.members) nz = std::max(nz, m->nz_active);
    const int64_t n_act = active_prefix(*E.members[0], nz);
    E.step_hooked(method, t, dt, T, s, [&](int i, const double *x) {
      if (!rec) return;
      for (int j = 0; j < N; ++j)
        copies.emplace_back(x + (int64_t)j * n, rec[i].Y + (int64_t)j * n,
                            n_act);
      copy_powder(rec[i].pw, rec[i].fpw, true);
    });
    ADM_HIP_ADJ(hipGetLastError());
  }

  void vjp(const Rec &rec, double t_eval, const double *mu, double *out,
           double *pbar, double *fpbar, double *partials) {
    ens_time(E, t_eval);
    batch.members = E.d_members;
    for (int b = 0; b < ADM_MAX_BEAMS; ++b) {
      batch.pmod[b] = E.pmod[b];
      batch.den[b] = E.den[b];
    }
    batch.max_faces = E.max_faces;
    batch.nz_active = 0;
    for (const auto &m : E.members)
      batch.nz_active = std::max(batch.nz_active, m->nz_active);
    launch_vjp(*E.members[0], rec.Y, rec.pw, rec.fpw, pbar, fpbar, mu, out,
               partials, 0.0, bpart, s, &batch);
  }

  void activate(const EnsEvent &e, double *T) {
    Context &c = *E.members[e.member];
    c.activate((int)e.cos.size(), e.boxes.data(), nullptr, e.cos.data(),
               e.sin.data(), e.new_temperature, T + (int64_t)e.member * n,
               nullptr, s);
    E.refresh_members();
    c.compute_inverse_mass(E.sb.inv_mass + (int64_t)e.member * n, s);
  }
};

int ensemble_adjoint(Ensemble &E, int method, double t0, double dt,
                     int n_steps, int checkpoint_every, const double *T0,
                     int n_obs, const int64_t *obs_dof, const double *obs,
                     const double *obs_weight, std::vector<EnsEvent> ev,
                     double *T_final, double *J, double *grad,
                     double *lambda0, hipStream_t s) {
  const int N = E.n_members;
  const int64_t n = E.n_dofs, nt = n * N;
  const int K = checkpoint_every > 0 ? checkpoint_every : n_steps;
  for (const auto &m : E.members)
    if (m->n_ext >= 0)
      throw AdmError(ADM_ERR_UNSUPPORTED, "the adjoint works on box nodes");
  for (const EnsEvent &e : ev)
    if (!E.members[e.member]->device_activation_fits())
      throw AdmError(ADM_ERR_UNSUPPORTED,
                     "mesh too large for the device activation");
  E.ensure_step_buffers(method);
  std::stable_sort(ev.begin(), ev.end(), [](const EnsEvent &a,
                                            const EnsEvent &b) {
    return a.step < b.step;
  });
  EnsAdjoint run(E, method, s);
  const int S = run.tab.ns;
  // segments: every K steps and at every activation step
  std::vector<int> bnd;
  for (int k = 0; k < n_steps; ++k) {
    bool b = k % K == 0;
    for (const EnsEvent &e : ev) b |= e.step == k;
    if (b) bnd.push_back(k);
  }
  const int n_seg = (int)bnd.size();
  bnd.push_back(n_steps);
  int max_len = 0;
  for (int i = 0; i < n_seg; ++i) max_len = std::max(max_len, bnd[i + 1] - bnd[i]);
  std::vector<double> times((size_t)n_steps + 1);
  {
    double t = t0;
    for (int k = 0; k <= n_steps; ++k) {
      times[k] = t;
      t = t + dt;
    }
  }
  // forward sweep: activations, checkpoints, steps
  DevBuf T, cpT, cpP;
  T.alloc(nt);
  cpT.alloc((size_t)nt * n_seg);
  cpP.alloc((size_t)run.npw * n_seg);
  // face powder at each segment start, packed over the members (the face
  // counts of that moment give the offsets)
  std::vector<DevBuf> cpF(n_seg);
  std::vector<std::vector<int64_t>> cpF_off(n_seg);
  std::vector<int64_t> fcap(N, 0);
  for (int j = 0; j < N; ++j)
    fcap[j] = E.members[j]->n_states == 3
                  ? E.members[j]->n_faces * E.members[j]->nfq : 0;
  std::vector<std::unique_ptr<GeomSnap>> snap(ev.size());
  ADM_HIP_ADJ(hipMemcpyAsync(T.p, T0, sizeof(double) * nt,
                             hipMemcpyDeviceToDevice, s));
  size_t next_ev = 0;
  auto cp_faces = [&](int i, bool save) {
    std::vector<int64_t> &off = cpF_off[i];
    if (save) {
      off.assign(N + 1, 0);
      for (int j = 0; j < N; ++j) {
        const Context &c = *E.members[j];
        off[j + 1] = off[j] + (c.n_states == 3 ? c.n_faces * c.nfq : 0);
      }
      cpF[i].alloc((size_t)off[N]);
    }
    for (int j = 0; j < N; ++j) {
      const int64_t f = off[j + 1] - off[j];
      if (!f) continue;
      Context &c = *E.members[j];
      run.copies.emplace_back(save ? c.d_face_powder : cpF[i].p + off[j],
                              save ? cpF[i].p + off[j] : c.d_face_powder, f);
    }
    run.flush_copies();
  };
  for (int i = 0; i < n_seg; ++i) {
    const int k0 = bnd[i], k1 = bnd[i + 1];
    while (next_ev < ev.size() && ev[next_ev].step == k0) {
      const EnsEvent &e = ev[next_ev];
      snap[next_ev].reset(new GeomSnap());
      ADM_HIP_ADJ(hipStreamSynchronize(s));
      Context &c = *E.members[e.member];
      // device arrays copied, the host part logged by the activation
      c.save_geometry(*snap[next_ev], s, false);
      c.geom_log = snap[next_ev].get();
      try {
        run.activate(e, T.p);
      } catch (...) {
        c.geom_log = nullptr;
        throw;
      }
      c.geom_log = nullptr;
      if (c.n_states == 3)
        fcap[e.member] = std::max(fcap[e.member], c.n_faces * c.nfq);
      ++next_ev;
    }
    ADM_HIP_ADJ(hipMemcpyAsync(cpT.p + (int64_t)i * nt, T.p,
                               sizeof(double) * nt, hipMemcpyDeviceToDevice,
                               s));
    // cell powder at its fixed offsets (face offsets are not needed here)
    for (int j = 0; j < N; ++j) {
      Context &c = *E.members[j];
      const int64_t a = run.pw_off[j + 1] - run.pw_off[j];
      if (a)
        run.copies.emplace_back(
            c.d_powder, cpP.p + (int64_t)i * run.npw + run.pw_off[j], a);
    }
    run.flush_copies();
    cp_faces(i, true);
    for (int k = k0; k < k1; ++k) run.forward_step(times[k], dt, T.p, nullptr);
  }
  if (next_ev != ev.size())
    throw AdmError(ADM_ERR_INVALID, "an activation step is out of the window");
  run.finish_tables(fcap);
  // the state at the final time (put back at the end)
  std::vector<std::unique_ptr<GeomSnap>> fin_geom(N);
  std::vector<bool> moved(N, false);
  for (const EnsEvent &e : ev) moved[e.member] = true;
  ADM_HIP_ADJ(hipStreamSynchronize(s));
  for (int j = 0; j < N; ++j)
    if (moved[j]) {
      fin_geom[j].reset(new GeomSnap());
      // the host part comes back by redoing the member's logged
      // activations
      E.members[j]->save_geometry(*fin_geom[j], s, false);
    }
  DevBuf finP, finF;
  finP.alloc(run.npw);
  finF.alloc(run.nfpw);
  run.copy_powder(finP.p, finF.p, true);
  // the members' liquid ratios and liquid flags, packed (member j's at
  // j times the member's length)
  DevBuf finL, finD;
  const int64_t nlq = E.members[0]->d_liquid
                          ? E.members[0]->n_cells * E.members[0]->nq : 0;
  const int64_t ndw = (E.members[0]->n_cells + 7) / 8;  // flags, in doubles
  if (nlq) {
    finL.alloc((size_t)nlq * N);
    finD.alloc((size_t)ndw * N);
    for (int j = 0; j < N; ++j) {
      Context &c = *E.members[j];
      if (!c.d_liquid) continue;  // (the members share the phase model)
      run.copies.emplace_back(c.d_liquid, finL.p + (int64_t)j * nlq, nlq);
      ADM_HIP_ADJ(hipMemcpyAsync(finD.p + (int64_t)j * ndw, c.d_liquid_dirty,
                                 (size_t)c.n_cells, hipMemcpyDeviceToDevice,
                                 s));
    }
    run.flush_copies();
  }
  if (T_final)
    ADM_HIP_ADJ(hipMemcpyAsync(T_final, T.p, sizeof(double) * nt,
                               hipMemcpyDeviceToDevice, s));
  // misfits and their gradients
  DevBuf Tb, mu, Ybar, partials, pbar, fpbar;
  Tb.alloc(nt);
  mu.alloc(nt);
  Ybar.alloc(nt);
  partials.alloc(3 * (size_t)N);
  pbar.alloc(run.npw);
  fpbar.alloc(run.nfpw);
  misfit_gradient(T.p, N, n, n_obs, obs_dof, obs, obs_weight, J, Tb.p, s);
  ADM_HIP_ADJ(hipMemsetAsync(partials.p, 0, sizeof(double) * 3 * N, s));
  if (run.npw) ADM_HIP_ADJ(hipMemsetAsync(pbar.p, 0, sizeof(double) * run.npw, s));
  if (run.nfpw)
    ADM_HIP_ADJ(hipMemsetAsync(fpbar.p, 0, sizeof(double) * run.nfpw, s));
  // per-segment recording
  DevBuf recY, recP, recF;
  recY.alloc((size_t)nt * max_len * S);
  recP.alloc((size_t)run.npw * max_len * S);
  recF.alloc((size_t)run.nfpw * max_len * S);
  std::vector<EnsAdjoint::Rec> rec((size_t)max_len * S);
  for (int64_t k = 0; k < (int64_t)max_len * S; ++k) {
    rec[k].Y = recY.p + k * nt;
    rec[k].pw = run.npw ? recP.p + k * run.npw : nullptr;
    rec[k].fpw = run.nfpw ? recF.p + k * run.nfpw : nullptr;
  }
  int ev_i = (int)ev.size() - 1;
  for (int i = n_seg - 1; i >= 0; --i) {
    const int k0 = bnd[i], k1 = bnd[i + 1];
    ADM_HIP_ADJ(hipMemcpyAsync(T.p, cpT.p + (int64_t)i * nt,
                               sizeof(double) * nt, hipMemcpyDeviceToDevice,
                               s));
    for (int j = 0; j < N; ++j) {
      Context &c = *E.members[j];
      const int64_t a = run.pw_off[j + 1] - run.pw_off[j];
      if (a)
        run.copies.emplace_back(
            cpP.p + (int64_t)i * run.npw + run.pw_off[j], c.d_powder, a);
    }
    run.flush_copies();
    cp_faces(i, false);
    for (int k = k0; k < k1; ++k)
      run.forward_step(times[k], dt, T.p, &rec[(size_t)(k - k0) * S]);
    int nz = 0;
    for (const auto &m : E.members) nz = std::max(nz, m->nz_active);
    const int64_t n_act = active_prefix(*E.members[0], nz);
    for (int k = k1 - 1; k >= k0; --k) {
      const EnsAdjoint::Rec *rk = &rec[(size_t)(k - k0) * S];
      backward_step(S, run.tab.a, run.tab.b, dt, run.K, Tb.p, mu.p, Ybar.p,
                    E.sb.inv_mass, n, n_act, N, s, [&](int st) {
                      run.vjp(rk[st], times[k] + run.tab.c[st] * dt, mu.p,
                              Ybar.p, pbar.p, fpbar.p, partials.p);
                    });
    }
    // the activations at the start of the segment, last first
    for (; ev_i >= 0 && ev[ev_i].step == k0; --ev_i) {
      const EnsEvent &e = ev[ev_i];
      const int j = e.member;
      Context &c = *E.members[j];
      c.activation_adjoint(Tb.p + (int64_t)j * n,
                           run.npw ? pbar.p + run.pw_off[j] : nullptr,
                           run.nfpw ? fpbar.p + run.fpw_off[j] : nullptr, s);
      const int64_t f = run.fpw_off[j + 1] - run.fpw_off[j];
      if (f)
        ADM_HIP_ADJ(hipMemsetAsync(fpbar.p + run.fpw_off[j], 0,
                                   sizeof(double) * f, s));
      ADM_HIP_ADJ(hipStreamSynchronize(s));
      c.restore_geometry(*snap[ev_i], s);
      E.refresh_members();
      c.compute_inverse_mass(E.sb.inv_mass + (int64_t)j * n, s);
    }
  }
  // back to the final state
  ADM_HIP_ADJ(hipStreamSynchronize(s));
  for (int j = 0; j < N; ++j) {
    Context &c = *E.members[j];
    if (moved[j]) {
      for (size_t e = 0; e < ev.size(); ++e)
        if (ev[e].member == j) c.redo_host_geometry(*snap[e]);
      c.restore_geometry(*fin_geom[j], s);
      c.compute_inverse_mass(E.sb.inv_mass + (int64_t)j * n, s);
    }
    if (nlq && c.d_liquid) {
      run.copies.emplace_back(finL.p + (int64_t)j * nlq, c.d_liquid, nlq);
      ADM_HIP_ADJ(hipMemcpyAsync(c.d_liquid_dirty, finD.p + (int64_t)j * ndw,
                                 (size_t)c.n_cells, hipMemcpyDeviceToDevice,
                                 s));
    }
  }
  run.flush_copies();
  E.refresh_members();
  run.copy_powder(finP.p, finF.p, false);
  ens_time(E, times[n_steps]);
  std::vector<double> pp(3 * (size_t)N);
  ADM_HIP_ADJ(hipMemcpyAsync(pp.data(), partials.p, sizeof(double) * 3 * N,
                             hipMemcpyDeviceToHost, s));
  if (lambda0)
    ADM_HIP_ADJ(hipMemcpyAsync(lambda0, Tb.p, sizeof(double) * nt,
                               hipMemcpyDeviceToDevice, s));
  ADM_HIP_ADJ(hipStreamSynchronize(s));
  for (int k = 0; k < 3 * N; ++k) grad[k] = pp[k];
  return ADM_OK;
}

}  // namespace adm

using adm::AdmError;

extern "C" {

int adm_adjoint_gradient(void *ptr, int method, double t0, double dt,
                         int n_steps, int checkpoint_every, const double *T0,
                         int32_t n_obs, const int64_t *obs_dof,
                         const double *obs, const double *obs_weight,
                         double *T_final, double *J, double *grad,
                         double *lambda0, void *stream) {
  try {
    if (!ptr) return adm::set_error(ADM_ERR_STATE, "null context");
    adm::Context &c = *static_cast<adm::Context *>(ptr);
    if (c.dim != 3)
      return adm::set_error(ADM_ERR_UNSUPPORTED,
                            "the adjoint is three-dimensional");
    if (!T0 || !obs_dof || !obs || !J || !grad || n_obs < 1 || n_steps < 1)
      return adm::set_error(ADM_ERR_INVALID, "bad adjoint arguments");
    for (int k = 0; k < n_obs; ++k)
      if (obs_dof[k] < 0 || obs_dof[k] >= c.n_dofs)
        return adm::set_error(ADM_ERR_INVALID, "observation DoF out of range");
    const int K = checkpoint_every > 0 ? checkpoint_every : n_steps;
    hipStream_t s = static_cast<hipStream_t>(stream);
    c.ensure_step_buffers(method);
    adm::AdjointRun run(c, method, s);
    const int64_t n = run.n, npw = run.npw, nfpw = run.nfpw;
    const int n_seg = (n_steps + K - 1) / K;
    // checkpoints: T, cell powder, face powder at the start of every segment
    adm::DevBuf cpT, cpP, cpF, T, Tb, mu, Ybar;
    cpT.alloc((size_t)n * n_seg);
    cpP.alloc((size_t)npw * n_seg);
    cpF.alloc((size_t)nfpw * n_seg);
    T.alloc(n);
    Tb.alloc(n);
    mu.alloc(n);
    Ybar.alloc(n);
    // per-segment recording
    const int S = run.n_stages;
    adm::DevBuf recY, recP, recF;
    recY.alloc((size_t)n * K * S);
    recP.alloc((size_t)npw * K * S);
    recF.alloc((size_t)nfpw * K * S);
    std::vector<adm::AdjointRun::StageRecord> rec((size_t)K * S);
    for (int64_t k = 0; k < (int64_t)K * S; ++k) {
      rec[k].Y = recY.p + k * n;
      rec[k].pw = npw ? recP.p + k * npw : nullptr;
      rec[k].fpw = nfpw ? recF.p + k * nfpw : nullptr;
    }
    // forward sweep
    ADM_HIP_ADJ(hipMemcpyAsync(T.p, T0, sizeof(double) * n,
                               hipMemcpyDeviceToDevice, s));
    std::vector<double> times((size_t)n_steps + 1);
    double t = t0;
    for (int k = 0; k < n_steps; ++k) {
      times[k] = t;
      if (k % K == 0) {
        const int sg = k / K;
        ADM_HIP_ADJ(hipMemcpyAsync(cpT.p + (int64_t)sg * n, T.p,
                                   sizeof(double) * n, hipMemcpyDeviceToDevice,
                                   s));
        if (npw)
          ADM_HIP_ADJ(hipMemcpyAsync(cpP.p + (int64_t)sg * npw, c.d_powder,
                                     sizeof(double) * npw,
                                     hipMemcpyDeviceToDevice, s));
        if (nfpw)
          ADM_HIP_ADJ(hipMemcpyAsync(cpF.p + (int64_t)sg * nfpw,
                                     c.d_face_powder, sizeof(double) * nfpw,
                                     hipMemcpyDeviceToDevice, s));
      }
      // the last segment is recorded now: the backward sweep starts there
      // without recomputing it
      const int k_last = (n_seg - 1) * K;
      run.forward_step(t, dt, T.p,
                       k >= k_last ? &rec[(size_t)(k - k_last) * S] : nullptr);
      t = t + dt;
    }
    times[n_steps] = t;
    // phase state at the final time, restored after the backward sweep
    const int64_t nlq = c.d_liquid ? c.n_cells * c.nq : 0;
    const int64_t ndirty = c.d_liquid_dirty ? (c.n_cells + 7) / 8 : 0;
    adm::DevBuf finP, finF, finL, finD;
    finP.alloc(npw);
    finF.alloc(nfpw);
    finL.alloc(nlq);
    finD.alloc(ndirty);
    auto copy_state = [&](bool save) {
      auto cp = [&](void *dst_save, void *live, size_t bytes) {
        if (!bytes) return;
        ADM_HIP_ADJ(hipMemcpyAsync(save ? dst_save : live,
                                   save ? live : dst_save, bytes,
                                   hipMemcpyDeviceToDevice, s));
      };
      cp(finP.p, c.d_powder, sizeof(double) * npw);
      cp(finF.p, c.d_face_powder, sizeof(double) * nfpw);
      cp(finL.p, c.d_liquid, sizeof(double) * nlq);
      cp(finD.p, c.d_liquid_dirty, ndirty ? (size_t)c.n_cells : 0);
    };
    copy_state(true);
    if (T_final)
      ADM_HIP_ADJ(hipMemcpyAsync(T_final, T.p, sizeof(double) * n,
                                 hipMemcpyDeviceToDevice, s));
    // misfit and its gradient at the final time
    adm::misfit_gradient(T.p, 1, n, n_obs, obs_dof, obs, obs_weight, J, Tb.p,
                         s);
    ADM_HIP_ADJ(hipMemsetAsync(run.partials.p, 0, sizeof(double) * 3, s));
    const double eta0 =
        c.beams.empty() ? 0.0 : c.beams[0].desc.absorption_efficiency;
    const double inv_eta0 = eta0 != 0.0 ? 1.0 / eta0 : 0.0;
    // backward sweep over the segments
    for (int sg = n_seg - 1; sg >= 0; --sg) {
      const int k0 = sg * K, k1 = std::min(n_steps, k0 + K);
      if (sg < n_seg - 1) {
        ADM_HIP_ADJ(hipMemcpyAsync(T.p, cpT.p + (int64_t)sg * n,
                                   sizeof(double) * n,
                                   hipMemcpyDeviceToDevice, s));
        if (npw)
          ADM_HIP_ADJ(hipMemcpyAsync(c.d_powder, cpP.p + (int64_t)sg * npw,
                                     sizeof(double) * npw,
                                     hipMemcpyDeviceToDevice, s));
        if (nfpw)
          ADM_HIP_ADJ(hipMemcpyAsync(c.d_face_powder,
                                     cpF.p + (int64_t)sg * nfpw,
                                     sizeof(double) * nfpw,
                                     hipMemcpyDeviceToDevice, s));
        for (int k = k0; k < k1; ++k)
          run.forward_step(times[k], dt, T.p, &rec[(size_t)(k - k0) * S]);
      }
      const int64_t n_act = adm::active_prefix(c, c.nz_active);
      for (int k = k1 - 1; k >= k0; --k) {
        const adm::AdjointRun::StageRecord *rk = &rec[(size_t)(k - k0) * S];
        adm::backward_step(S, run.a, run.b, dt, run.K, Tb.p, mu.p, Ybar.p,
                           c.sb.inv_mass, n, n_act, 1, s, [&](int i) {
                             run.vjp(rk[i], times[k] + run.cst[i] * dt, mu.p,
                                     Ybar.p, inv_eta0);
                           });
      }
    }
    copy_state(false);
    adm::eval_time_reset(c, times[n_steps]);
    ADM_HIP_ADJ(hipGetLastError());
    double pp[3];
    ADM_HIP_ADJ(hipMemcpyAsync(pp, run.partials.p, sizeof(pp),
                               hipMemcpyDeviceToHost, s));
    if (lambda0)
      ADM_HIP_ADJ(hipMemcpyAsync(lambda0, Tb.p, sizeof(double) * n,
                                 hipMemcpyDeviceToDevice, s));
    ADM_HIP_ADJ(hipStreamSynchronize(s));
    for (int k = 0; k < 3; ++k) grad[k] = pp[k];
  } catch (const AdmError &e) {
    return adm::set_error(e.code, e.what());
  } catch (const std::exception &e) {
    return adm::set_error(ADM_ERR_INVALID, e.what());
  }
  return ADM_OK;
}

int adm_ensemble_adjoint_gradient(
    void *ptr, int method, double t0, double dt, int n_steps,
    int checkpoint_every, const double *T0, int32_t n_obs,
    const int64_t *obs_dof, const double *obs, const double *obs_weight,
    int32_t n_events, const adm_activation_event *events, double *T_final,
    double *J, double *grad, double *lambda0, void *stream) {
  try {
    if (!ptr) return adm::set_error(ADM_ERR_STATE, "null ensemble");
    adm::Ensemble &E = *static_cast<adm::Ensemble *>(ptr);
    if (E.members.empty() || E.members[0]->dim != 3)
      return adm::set_error(ADM_ERR_UNSUPPORTED,
                            "the adjoint is three-dimensional");
    if (!T0 || !obs_dof || !obs || !J || !grad || n_obs < 1 || n_steps < 1 ||
        n_events < 0 || (n_events && !events))
      return adm::set_error(ADM_ERR_INVALID, "bad adjoint arguments");
    for (int k = 0; k < n_obs; ++k)
      if (obs_dof[k] < 0 || obs_dof[k] >= E.n_dofs)
        return adm::set_error(ADM_ERR_INVALID, "observation DoF out of range");
    std::vector<adm::EnsEvent> ev(n_events);
    for (int i = 0; i < n_events; ++i) {
      const adm_activation_event &e = events[i];
      if (e.member < 0 || e.member >= E.n_members || e.step < 0 ||
          e.step >= n_steps || e.n_boxes < 0 ||
          (e.n_boxes && (!e.boxes || !e.box_cos || !e.box_sin)))
        return adm::set_error(ADM_ERR_INVALID, "bad activation event");
      ev[i].member = e.member;
      ev[i].step = e.step;
      ev[i].boxes.assign(e.boxes, e.boxes + 6 * (size_t)e.n_boxes);
      ev[i].cos.assign(e.box_cos, e.box_cos + e.n_boxes);
      ev[i].sin.assign(e.box_sin, e.box_sin + e.n_boxes);
      ev[i].new_temperature = e.new_temperature;
    }
    return adm::ensemble_adjoint(E, method, t0, dt, n_steps, checkpoint_every,
                                 T0, n_obs, obs_dof, obs, obs_weight,
                                 std::move(ev), T_final, J, grad, lambda0,
                                 static_cast<hipStream_t>(stream));
  } catch (const AdmError &e) {
    return adm::set_error(e.code, e.what());
  } catch (const std::exception &e) {
    return adm::set_error(ADM_ERR_INVALID, e.what());
  }
}

int adm_operator_vjp(void *ptr, double t, const double *T, const double *mu,
                     double *out, double *grad, void *stream) {
  try {
    if (!ptr) return adm::set_error(ADM_ERR_STATE, "null context");
    adm::Context &c = *static_cast<adm::Context *>(ptr);
    if (c.dim != 3)
      return adm::set_error(ADM_ERR_UNSUPPORTED,
                            "the adjoint is three-dimensional");
    if (!T || !mu || !out)
      return adm::set_error(ADM_ERR_INVALID, "null vector");
    hipStream_t s = static_cast<hipStream_t>(stream);
    adm::DevBuf partials, bpart;
    partials.alloc(3);
    ADM_HIP_ADJ(hipMemsetAsync(partials.p, 0, sizeof(double) * 3, s));
    ADM_HIP_ADJ(hipMemsetAsync(out, 0, sizeof(double) * c.n_dofs, s));
    adm::eval_time_reset(c, t);
    const double eta0 =
        c.beams.empty() ? 0.0 : c.beams[0].desc.absorption_efficiency;
    adm::launch_vjp(c, T, c.d_powder, c.d_face_powder, nullptr, nullptr, mu,
                    out, partials.p, eta0 != 0.0 ? 1.0 / eta0 : 0.0, bpart,
                    s);
    double pp[3];
    ADM_HIP_ADJ(hipMemcpyAsync(pp, partials.p, sizeof(pp),
                               hipMemcpyDeviceToHost, s));
    ADM_HIP_ADJ(hipStreamSynchronize(s));
    if (grad)
      for (int k = 0; k < 3; ++k) grad[k] = pp[k];
  } catch (const AdmError &e) {
    return adm::set_error(e.code, e.what());
  } catch (const std::exception &e) {
    return adm::set_error(ADM_ERR_INVALID, e.what());
  }
  return ADM_OK;
}

}  // extern "C"
