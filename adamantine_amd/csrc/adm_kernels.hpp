// HIP kernels of the matrix-free heat operator (gfx950).
//
// Cell kernel: one lane per cell. The (p+1)^3 nodal values are gathered into
// registers and the whole sum factorisation (values, collocation gradients,
// transposed integration) runs in registers in FP64 with explicit FMA chains.
// The per-quadrature-point physics (phase state, rho c_p with latent heat,
// k(T), deposition rotation, heat sources) and the boundary-face terms of the
// cell are fused in. Reference: source/ThermalOperator.templates.hh:479-751.
#pragma once

#include <type_traits>

#include "adm_device.hpp"

namespace adm {

// Compile-time loop: f(std::integral_constant<int, I>) for I = B .. E-1, so
// every register-array index is a constant whatever the body size.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// out = M (or M^T) applied along direction D of an N^3 tensor (x fastest).
// The order of operations matches the CPU oracle's sweep().
template <int N, int D, bool TRANSPOSE, bool ACC, int DIMS = 3>
__device__ __forceinline__ void sweep(const double (&M)[kMaxN][kMaxN],
                                      const double *in, double *out) {
  constexpr int stride = D == 0 ? 1 : (D == 1 ? N : N * N);
  constexpr int total = DIMS == 3 ? N * N * N : (DIMS == 2 ? N * N : N);
#pragma unroll
  for (int o = 0; o < total; ++o) {
    const int c = (o / stride) % N;
    const int base = o - c * stride;
    double acc = (TRANSPOSE ? M[0][c] : M[c][0]) * in[base];
#pragma unroll
    for (int i = 1; i < N; ++i)
      acc = __builtin_fma(TRANSPOSE ? M[i][c] : M[c][i], in[base + i * stride],
                          acc);
    out[o] = ACC ? out[o] + acc : acc;
  }
}

// Property evaluator selected by the PK template parameter:
// 0 polynomial, 1 general table, 2 merged table.
template <int PK>
struct EvalOf;
template <>
struct EvalOf<0> {
  using type = PolyEval;
};
template <>
struct EvalOf<1> {
  using type = TableEval;
};
template <>
struct EvalOf<2> {
  using type = MergedEval;
};

// Boundary value at one face quadrature point
// (source/ThermalOperator.templates.hh:691-746), already weighted by JxW.
template <int NS, class Eval>
__device__ __forceinline__ double face_value(const MatConst &m, int mid,
                                             int bt, double T, double &powder,
                                             const Eval &ev,
                                             const double (&r)[NS]) {
  const double irc = inv_rho_cp<NS>(m, mid, r, ev);
  double conv_T = 0.0, conv_h = 0.0, rad_T = 0.0, rad_h = 0.0;
  if (bt & ADM_BC_CONVECTIVE) {
    conv_T = m.prop[mid][ADM_CONVECTION_TEMPERATURE_INFTY];
    conv_h = ev(ADM_CONVECTION_HEAT_TRANSFER_COEF, r);
  }
  if (bt & ADM_BC_RADIATIVE) {
    rad_T = m.prop[mid][ADM_RADIATION_TEMPERATURE_INFTY];
    rad_h = ev(ADM_EMISSIVITY, r) * kStefanBoltzmann * (T + rad_T) *
            (T * T + rad_T * rad_T);
  }
  return -irc * (conv_h * (T - conv_T) + rad_h * (T - rad_T));
}

// Face kernel (source/ThermalOperator.templates.hh:619-751): one lane per
// boundary face of the active domain that carries a convective or radiative
// term; face values from the face's (p+1)^2 nodes, per-face-point powder
// state, JxW-weighted integration, atomic accumulation into dst.
template <int P, int NS, int PK>
__global__ void __launch_bounds__(256)
    face_kernel(const FaceArgs a, const double *__restrict__ src,
                double *__restrict__ dst) {
  constexpr int N = P + 1, NF = N * N;
  const int64_t fi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const MemberArgs *mb = a.members ? a.members + blockIdx.y : nullptr;
  if (fi >= (mb ? mb->n_faces : a.n_faces)) return;
  const KConst &kc = *(mb ? mb->kc : a.kc);
  const MatConst &m = kc.mat;
  const FaceRef fr = (mb ? mb->faces : a.faces)[fi];
  double *face_powder = mb ? mb->face_powder : a.face_powder;
  if (mb) {
    src += (int64_t)blockIdx.y * a.member_stride;
    dst += (int64_t)blockIdx.y * a.member_stride;
  }
  const int64_t c = fr.cell;
  const int f = fr.info & 7, bt = (fr.info >> 3) & 31, mid = (fr.info >> 8) & 3;
  const int d = f / 2, side = f % 2;
  const int cc[3] = {(int)(c % a.nc[0]), (int)((c / a.nc[0]) % a.nc[1]),
                     (int)(c / ((int64_t)a.nc[0] * a.nc[1]))};
  const int64_t ns[3] = {1, (int64_t)a.nn[0],
                         (int64_t)a.nn[0] * (int64_t)a.nn[1]};
  const int64_t base = (int64_t)cc[0] * P + ns[1] * cc[1] * P +
                       ns[2] * cc[2] * P + (side ? P : 0) * ns[d];
  const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
  int64_t node[NF];
  double fu[NF], ft[NF], fv[NF];
#pragma unroll
  for (int j = 0; j < NF; ++j) {
    node[j] = base + (j % N) * ns[t1] + (j / N) * ns[t2];
    fu[j] = src[node[j]];
  }
  sweep<N, 0, false, false, 2>(kc.S, fu, ft);
  sweep<N, 1, false, false, 2>(kc.S, ft, fv);
  double *fp = NS == 3 ? face_powder + fi * NF : nullptr;
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    const double T = fv[q];
    double powder = fp ? fp[q] : 0.0;
    double liquid;
    double r[NS];
    if (state_ratios<NS>(m, mid, T, powder, liquid, r)) fp[q] = powder;
    double val;
    if constexpr (PK == 2) {
      const MergedEval ev(m.merged[mid].B, &m.merged[mid].coef[0][0][0][0], T);
      val = face_value<NS>(m, mid, bt, T, powder, ev, r);
    } else {
      const typename EvalOf<PK>::type ev(m, mid, T);
      val = face_value<NS>(m, mid, bt, T, powder, ev, r);
    }
    fu[q] = val * kc.fjxw[d][q];
  }
  sweep<N, 1, true, false, 2>(kc.S, fu, ft);
  sweep<N, 0, true, false, 2>(kc.S, ft, fv);
#pragma unroll
  for (int j = 0; j < NF; ++j) unsafeAtomicAdd(dst + node[j], fv[j]);
}

// dot product of row r of M with an N-line of V (same chain as sweep())
template <int N>
__device__ __forceinline__ double line_dot(const double (&M)[kMaxN][kMaxN],
                                           int r, const double *V, int base,
                                           int stride) {
  double acc = M[r][0] * V[base];
#pragma unroll
  for (int i = 1; i < N; ++i)
    acc = __builtin_fma(M[r][i], V[base + i * stride], acc);
  return acc;
}

// doubles of a MergedTable staged in LDS by the cell kernel: B and the
// coefficients of density .. thermal_conductivity_z
constexpr int kMergedLds = kMergedMax + 5 * kMergedMax * ADM_MAX_STATES * 2;

// Cell kernel (source/ThermalOperator.templates.hh:479-617).
//
// N = p+1 lanes per cell, lane k of a cell owning node plane z = k and
// quadrature plane qz = k. The x and y contractions and the per-point
// physics of the plane's N*N points run in registers; the z contractions
// (values, z-gradient, transposed z-gradient, transposed values) exchange
// planes through a wave-private LDS buffer. The N lanes of a cell sit in the
// same wavefront, so no workgroup barrier is involved.
template <int P>
struct CellLanes {
  static constexpr int N = P + 1;
  static constexpr int cells_per_wave = 64 / N;
  static constexpr int cells_per_block = 4 * cells_per_wave;
};

// LDS ordering between the lanes of one wavefront
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Merged-table coefficients of one interval j for the properties the cell
// term needs (density, specific heat, conductivities), for states 0 and 2.
struct SolidCoef {
  double b;        // left breakpoint of the interval
  double c[5][2][2];  // [prop][state 0 or 2][value, slope]
};

template <int P, int NS, int PK>
__global__ void __launch_bounds__(256)
    cell_kernel(const CellArgs a, const double *__restrict__ src,
                double *__restrict__ dst) {
  constexpr int N = P + 1, NP = N * N;
  constexpr int CPW = CellLanes<P>::cells_per_wave;
  // ensemble member of this block (uniform), else the single operator
  const MemberArgs *mb = a.members ? a.members + blockIdx.y : nullptr;
  const KConst &kc = *(mb ? mb->kc : a.kc);
  const MatConst &m = kc.mat;
  const uint32_t *meta_p = mb ? mb->meta : a.meta;
  const int32_t *slot_p = mb ? mb->powder_slot : a.powder_slot;
  double *powder_p = mb ? mb->powder : a.powder;
  double *liquid_p = mb ? mb->liquid : a.liquid;
  uint8_t *dirty_p = mb ? mb->liquid_dirty : a.liquid_dirty;
  const double *cos_p = mb ? mb->dep_cos : a.dep_cos;
  const double *sin_p = mb ? mb->dep_sin : a.dep_sin;
  if (mb) {
    src += (int64_t)blockIdx.y * a.member_stride;
    dst += (int64_t)blockIdx.y * a.member_stride;
  }
  extern __shared__ double lds[];
  __shared__ adm_beam_state sbeams[ADM_MAX_BEAMS];
  // one lane copies the beam states with uniform (scalar) loads
  if (threadIdx.x == 0) {
    static_for<0, ADM_MAX_BEAMS>([&](auto bc) {
      constexpr int bb = decltype(bc)::value;
      sbeams[bb] = a.beams[bb];
      if (mb)
        sbeams[bb].alpha = member_alpha(a.beams[bb], mb->beam_gain[bb],
                                        a.beam_pmod[bb], a.beam_den[bb],
                                        kc.log_01);
    });
  }
  if constexpr (PK == 2) {
    const int total = m.n_materials * kMergedLds;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
      const int mm = i / kMergedLds, k = i % kMergedLds;
      lds[i] = (&m.merged[mm].B[0])[k];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int j = lane / N, kz = lane % N;
  const int64_t c = ((int64_t)blockIdx.x * 4 + wave) * CPW + j;
  const bool inside = j < CPW && c < a.n_cells;
  const uint32_t meta = inside ? meta_p[c] : 0u;
  if (!(meta & kActive)) return;  // whole cells (all N lanes) leave together
  // exchange buffer of this cell: [N planes][NP]
  double *xb = lds + (PK == 2 ? m.n_materials * kMergedLds : 0) +
               ((int64_t)wave * CPW + j) * N * NP;
  const int mid = (meta >> kMaterialShift) & 3;
  int cc[3];
  cc[0] = (int)(c % a.nc[0]);
  cc[1] = (int)((c / a.nc[0]) % a.nc[1]);
  cc[2] = (int)(c / ((int64_t)a.nc[0] * a.nc[1]));
  const int64_t sy = a.nn[0], sz = (int64_t)a.nn[0] * a.nn[1];
  const int64_t base = (int64_t)cc[0] * P + sy * cc[1] * P +
                       sz * ((int64_t)cc[2] * P + kz);

  // node plane kz, then S along x and y (the first two sweeps of the oracle)
  double u[NP], t[NP];
#pragma unroll
  for (int jj = 0; jj < N; ++jj)
#pragma unroll
    for (int i = 0; i < N; ++i) u[i + N * jj] = src[base + i + sy * jj];
  sweep<N, 0, false, false, 2>(kc.S, u, t);
  sweep<N, 1, false, false, 2>(kc.S, t, u);
  // S along z gives the values on quadrature plane kz (same chain as the
  // oracle); Dc*S along z gives the z-derivative from the same planes
  double V[NP], gzv[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) xb[kz * NP + i] = u[i];
  wave_lds_sync();
  {
    double srow[N], drow[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      srow[k] = kc.S[kz][k];
      drow[k] = kc.DcS[kz][k] * kc.inv_h[2];
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const double u0 = xb[i];
      double acc = srow[0] * u0, gz = drow[0] * u0;
#pragma unroll
      for (int k = 1; k < N; ++k) {
        const double uk = xb[k * NP + i];
        acc = __builtin_fma(srow[k], uk, acc);
        gz = __builtin_fma(drow[k], uk, gz);
      }
      V[i] = acc;
      gzv[i] = gz;
    }
  }

  const int64_t pslot = (meta & kHasPowder) ? (int64_t)slot_p[c] : -1;
  double *pw = pslot >= 0 ? powder_p + (pslot * N + kz) * NP : nullptr;
  double *lq = nullptr;
  bool dirty = false;
  if constexpr (NS > 1) {
    if (liquid_p) {
      dirty = dirty_p[c] != 0;
      lq = liquid_p + (c * N + kz) * NP;
    }
  }
  const bool rotated = (meta & kRotated) != 0;
  double cs = 1.0, sn = 0.0;
  if (rotated) {
    cs = cos_p[c];
    sn = sin_p[c];
  }
  const bool iso = m.isotropic[mid] != 0;
  const double *mB = lds + mid * kMergedLds;
  const double *mC = mB + kMergedMax;
  bool near = false;
  if (a.sources_on) {
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = kc.origin[d] + cc[d] * kc.h[d];
      hi[d] = kc.origin[d] + (cc[d] + 1) * kc.h[d];
    }
#pragma unroll 1
    for (int b = 0; b < a.n_beams; ++b) near |= beam_may_touch(sbeams[b], lo, hi);
  }
  const double z = kc.origin[2] + (cc[2] + kc.xq[kz]) * kc.h[2];

  // Fast path: every point of the plane below the solidus and inside one
  // merged interval. Then the liquid ratio is 0 and the ratios are
  // (1 - p, 0, p) with p the stored powder, the liquid terms of the
  // reference's mixture are r*X = 0 (dropped exactly), and the table
  // coefficients are loaded once for the plane.
  double tmin = V[0], tmax = V[0];
#pragma unroll
  for (int i = 1; i < NP; ++i) {
    tmin = fmin(tmin, V[i]);
    tmax = fmax(tmax, V[i]);
  }
  bool fast = PK == 2;
  int jint = 0;
  double bj = 0.0;
  if constexpr (PK == 2) {
    if constexpr (NS > 1) fast = tmax < m.prop[mid][ADM_SOLIDUS];
    const MergedEval elo(mB, mC, tmin), ehi(mB, mC, tmax);
    fast = fast && elo.j == ehi.j && !rotated;
    jint = elo.j;
    bj = mB[jint > 0 ? jint - 1 : 0];
  }

  double A[NP], w2v[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) A[i] = 0.0;
  bool plane_liquid = false;

  // flux and source of one point from (1/rho c_p, k_x, k_y, k_z)
  auto finish = [&](auto ic, double T, double irc, double kx, double ky,
                    double kzz, double gx, double gy) {
    constexpr int i = decltype(ic)::value;
    constexpr int qx = i % N, qy = i / N;
    double fx, fy;
    if (rotated) {
      fx = (kx * cs * cs + ky * sn * sn) * gx + ((kx - ky) * sn * cs) * gy;
      fy = ((kx - ky) * sn * cs) * gx + (kx * sn * sn + ky * cs * cs) * gy;
    } else {
      // cos = 1, sin = 0: the reference's rotation gives k_x g_x, k_y g_y
      fx = kx * gx;
      fy = ky * gy;
    }
    const double fz = gzv[i] * kzz;
    const double jxw = kc.jxw[kz * NP + i];
    const double w0 = ((-irc) * fx) * jxw * kc.inv_h[0];
    const double w1 = ((-irc) * fy) * jxw * kc.inv_h[1];
    w2v[i] = ((-irc) * fz) * jxw * kc.inv_h[2];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      A[i - qx + k] = __builtin_fma(kc.Dc[qx][k], w0, A[i - qx + k]);
      A[i - qy * N + k * N] =
          __builtin_fma(kc.Dc[qy][k], w1, A[i - qy * N + k * N]);
    }
    if (near) {
      const double x = kc.origin[0] + (cc[0] + kc.xq[qx]) * kc.h[0];
      const double y = kc.origin[1] + (cc[1] + kc.xq[qy]) * kc.h[1];
      double qs = 0.0;
#pragma unroll 1
      for (int b = 0; b < a.n_beams; ++b)
        qs += beam_value(sbeams[b], x, y, z, kc.log_01);
      qs *= irc;
      A[i] += qs * jxw;
    }
  };

  if (fast) {
    // interval coefficients: [prop][state 0 / 2][value, slope]
    constexpr int s2 = NS == 3 ? 2 : 0;
    double cf[5][2][2];
#pragma unroll
    for (int pr = 0; pr < 5; ++pr)
#pragma unroll
      for (int ss = 0; ss < 2; ++ss) {
        const double *cp =
            mC + ((pr * kMergedMax + jint) * ADM_MAX_STATES + (ss ? s2 : 0)) * 2;
        cf[pr][ss][0] = cp[0];
        cf[pr][ss][1] = cp[1];
      }
    const double latent_mushy = 0.0;  // liquid ratio 0: never mushy
    (void)latent_mushy;
    static_for<0, NP>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int qx = i % N, qy = i / N;
      const double T = V[i];
      const double d = T - bj;
      const double gx = line_dot<N>(kc.Dc, qx, V, i - qx, 1) * kc.inv_h[0];
      const double gy = line_dot<N>(kc.Dc, qy, V, i - qy * N, N) * kc.inv_h[1];
      // ratios (1 - p, 0, p); with p = 0 only the solid term remains
      double p = 0.0;
      if (NS == 3 && pw) p = pw[i];
      auto mix = [&](int pr) {
        const double x0 = __builtin_fma(d, cf[pr][0][1], cf[pr][0][0]);
        double v = (1.0 - 0.0 - p) * x0;
        if constexpr (NS == 3) {
          const double x2 = __builtin_fma(d, cf[pr][1][1], cf[pr][1][0]);
          v += p * x2;
        }
        return v;
      };
      const double irc = 1.0 / (mix(ADM_DENSITY) * mix(ADM_SPECIFIC_HEAT));
      const double kx = mix(ADM_THERMAL_CONDUCTIVITY_X);
      double ky = kx, kzz = kx;
      if (!iso) {
        ky = mix(ADM_THERMAL_CONDUCTIVITY_Y);
        kzz = mix(ADM_THERMAL_CONDUCTIVITY_Z);
      }
      finish(ic, T, irc, kx, ky, kzz, gx, gy);
    });
  } else {
    static_for<0, NP>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int qx = i % N, qy = i / N;
      const double T = V[i];
      const double gx = line_dot<N>(kc.Dc, qx, V, i - qx, 1) * kc.inv_h[0];
      const double gy = line_dot<N>(kc.Dc, qy, V, i - qy * N, N) * kc.inv_h[1];
      double p = pw ? pw[i] : 0.0;
      const double p_old = p;
      double r[NS];
      const double l = state_ratios_select<NS>(m, mid, T, p, r);
      if (pw && p != p_old) pw[i] = p;
      if constexpr (NS > 1) {
        if (lq && (l != 0.0 || dirty)) lq[i] = l;
        plane_liquid |= l != 0.0;
      }
      double irc, kx, ky, kzz;
      auto props = [&](const auto &ev) {
        irc = inv_rho_cp<NS>(m, mid, r, ev);
        kx = ev(ADM_THERMAL_CONDUCTIVITY_X, r);
        if (iso) {
          ky = kx;
          kzz = kx;
        } else {
          ky = ev(ADM_THERMAL_CONDUCTIVITY_Y, r);
          kzz = ev(ADM_THERMAL_CONDUCTIVITY_Z, r);
        }
      };
      if constexpr (PK == 2)
        props(MergedEval(mB, mC, T));
      else
        props(typename EvalOf<PK>::type(m, mid, T));
      finish(ic, T, irc, kx, ky, kzz, gx, gy);
      __builtin_amdgcn_sched_barrier(0);
    });
  }
  if constexpr (NS > 1) {
    // fast planes hold no liquid; a plane that held liquid is cleared
    if (fast && lq && dirty) {
#pragma unroll
      for (int i = 0; i < NP; ++i) lq[i] = 0.0;
    }
    // cell-wide flag over the N lanes of the cell
    wave_lds_sync();
    xb[kz * NP] = plane_liquid ? 1.0 : 0.0;
    wave_lds_sync();
    bool any_liquid = false;
#pragma unroll
    for (int k = 0; k < N; ++k) any_liquid |= xb[k * NP] != 0.0;
    if (lq && kz == 0 && any_liquid != dirty)
      dirty_p[c] = any_liquid ? 1 : 0;
  }
  // node plane kz: S^T along z of the x/y-integrated planes, plus the
  // transposed z-derivative of the z-fluxes (Dc*S)^T
  double out[NP];
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < NP; ++i) xb[kz * NP + i] = A[i];
  wave_lds_sync();
  {
    double scol[N];
#pragma unroll
    for (int k = 0; k < N; ++k) scol[k] = kc.S[k][kz];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      double acc = scol[0] * xb[i];
#pragma unroll
      for (int k = 1; k < N; ++k) acc = __builtin_fma(scol[k], xb[k * NP + i], acc);
      out[i] = acc;
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < NP; ++i) xb[kz * NP + i] = w2v[i];
  wave_lds_sync();
  {
    double dcol[N];
#pragma unroll
    for (int k = 0; k < N; ++k) dcol[k] = kc.DcS[k][kz];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
      for (int k = 0; k < N; ++k) out[i] = __builtin_fma(dcol[k], xb[k * NP + i], out[i]);
  }
  // S^T along y and x
  sweep<N, 1, true, false, 2>(kc.S, out, t);
  sweep<N, 0, true, false, 2>(kc.S, t, out);
#pragma unroll
  for (int jj = 0; jj < N; ++jj)
#pragma unroll
    for (int i = 0; i < N; ++i)
      unsafeAtomicAdd(dst + base + i + sy * jj, out[i + N * jj]);
}

// Lumped Gauss-Lobatto mass, inverted (source/ThermalOperator.templates.hh:
// 81-164): one lane per node sums det * w over its active cells in
// increasing cell order.
template <int P>
__global__ void __launch_bounds__(256)
    inverse_mass_kernel(const uint32_t *__restrict__ meta,
                        const double *__restrict__ wgl, double det, int32_t nc0,
                        int32_t nc1, int32_t nc2, int64_t n_dofs,
                        double *__restrict__ inv_mass) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= n_dofs) return;
  const int64_t nn0 = (int64_t)P * nc0 + 1, nn1 = (int64_t)P * nc1 + 1;
  const int64_t ix = node % nn0, iy = (node / nn0) % nn1, iz = node / (nn0 * nn1);
  int cand_c[3][2], cand_l[3][2], ncand[3];
  const int64_t id[3] = {ix, iy, iz};
  const int32_t nc[3] = {nc0, nc1, nc2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ncand[d] = 0;
    const int64_t cq = id[d] / P, lr = id[d] % P;
    if (lr == 0) {
      if (cq - 1 >= 0) {
        cand_c[d][ncand[d]] = (int)(cq - 1);
        cand_l[d][ncand[d]++] = P;
      }
      if (cq < nc[d]) {
        cand_c[d][ncand[d]] = (int)cq;
        cand_l[d][ncand[d]++] = 0;
      }
    } else {
      cand_c[d][ncand[d]] = (int)cq;
      cand_l[d][ncand[d]++] = (int)lr;
    }
  }
  double sum = 0.0;
  for (int kz = 0; kz < ncand[2]; ++kz)
    for (int ky = 0; ky < ncand[1]; ++ky)
      for (int kx = 0; kx < ncand[0]; ++kx) {
        const int64_t c = cand_c[0][kx] + (int64_t)nc0 *
                                              (cand_c[1][ky] +
                                               (int64_t)nc1 * cand_c[2][kz]);
        if (!(meta[c] & kActive)) continue;
        const double w = wgl[cand_l[0][kx]] * wgl[cand_l[1][ky]] *
                         wgl[cand_l[2][kz]];
        sum += det * w;
      }
  inv_mass[node] = sum > 0.0 ? 1.0 / sum : 0.0;
}

}  // namespace adm
