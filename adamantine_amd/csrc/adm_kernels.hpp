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
  if (fi >= a.n_faces) return;
  const KConst &kc = *a.kc;
  const MatConst &m = kc.mat;
  const FaceRef fr = a.faces[fi];
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
  double *fp = NS == 3 ? a.face_powder + fi * NF : nullptr;
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
template <int P, int NS, int PK>
__global__ void __launch_bounds__(256)
    cell_kernel(const CellArgs a, const double *__restrict__ src,
                double *__restrict__ dst, int64_t n_cells) {
  constexpr int N = P + 1, NQ = N * N * N;
  const KConst &kc = *a.kc;
  const MatConst &m = kc.mat;
  extern __shared__ double lds[];
  // beam states in LDS so that the (rolled) beam loop indexes memory
  __shared__ adm_beam_state sbeams[ADM_MAX_BEAMS];
  if (threadIdx.x < ADM_MAX_BEAMS * (sizeof(adm_beam_state) / 8)) {
    const int b = threadIdx.x / (sizeof(adm_beam_state) / 8);
    const int k = threadIdx.x % (sizeof(adm_beam_state) / 8);
    double v = 0.0;
    static_for<0, ADM_MAX_BEAMS>([&](auto bc) {
      constexpr int bb = decltype(bc)::value;
      if (b == bb) v = reinterpret_cast<const double *>(&a.beams[bb])[k];
    });
    reinterpret_cast<double *>(&sbeams[b])[k] = v;
  }
  if constexpr (PK != 2) __syncthreads();
  if constexpr (PK == 2) {
    const int total = m.n_materials * kMergedLds;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
      const int mm = i / kMergedLds, k = i % kMergedLds;
      lds[i] = (&m.merged[mm].B[0])[k];
    }
    __syncthreads();
  }
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const uint32_t meta = a.meta[c];
  if (!(meta & kActive)) return;
  const int mid = (meta >> kMaterialShift) & 3;
  int cc[3];
  cc[0] = (int)(c % a.nc[0]);
  cc[1] = (int)((c / a.nc[0]) % a.nc[1]);
  cc[2] = (int)(c / ((int64_t)a.nc[0] * a.nc[1]));
  const int64_t sy = a.nn[0], sz = (int64_t)a.nn[0] * a.nn[1];
  const int64_t base = (int64_t)cc[0] * P + sy * cc[1] * P + sz * cc[2] * P;

  double A[NQ], V[NQ];
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int i = 0; i < N; ++i)
        A[i + N * (j + N * k)] = src[base + i + sy * j + sz * k];

  // values at the quadrature points: S along x, y, z
  sweep<N, 0, false, false>(kc.S, A, V);
  sweep<N, 1, false, false>(kc.S, V, A);
  sweep<N, 2, false, false>(kc.S, A, V);

  const int64_t pslot = (meta & kHasPowder) ? (int64_t)a.powder_slot[c] : -1;
  double *pw = pslot >= 0 ? a.powder + pslot * NQ : nullptr;
  double *lq = nullptr;
  bool dirty = false;
  if constexpr (NS > 1) {
    if (a.liquid) {
      dirty = a.liquid_dirty[c] != 0;
      lq = a.liquid + c * NQ;
    }
  }
  bool any_liquid = false;
  double cs = 1.0, sn = 0.0;
  const bool rotated = (meta & kRotated) != 0;
  if (rotated) {
    cs = a.dep_cos[c];
    sn = a.dep_sin[c];
  }
  const bool iso = m.isotropic[mid] != 0;
  const double *mB = lds + mid * kMergedLds;
  const double *mC = mB + kMergedMax;
  // cells no beam can reach get exactly zero source contributions
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

#pragma unroll
  for (int q = 0; q < NQ; ++q) A[q] = 0.0;
  static_for<0, NQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int qx = q % N, qy = (q / N) % N, qz = q / (N * N);
    // reference gradients by the collocation derivative (chains of sweep())
    const double gx = line_dot<N>(kc.Dc, qx, V, q - qx, 1) * kc.inv_h[0];
    const double gy = line_dot<N>(kc.Dc, qy, V, q - qy * N, N) * kc.inv_h[1];
    const double gz =
        line_dot<N>(kc.Dc, qz, V, q - qz * N * N, N * N) * kc.inv_h[2];
    const double T = V[q];
    double powder = pw ? pw[q] : 0.0;
    double liquid;
    double r[NS];
    if (state_ratios<NS>(m, mid, T, powder, liquid, r)) pw[q] = powder;
    if constexpr (NS > 1) {
      // the stored liquid ratio (a function of T) is written where it is
      // non-zero, and everywhere in a cell that held liquid before
      if (lq && (liquid != 0.0 || dirty)) lq[q] = liquid;
      any_liquid |= liquid != 0.0;
    }
    double irc, kx, ky, kz;
    auto props = [&](const auto &ev) {
      irc = inv_rho_cp<NS>(m, mid, r, ev);
      kx = ev(ADM_THERMAL_CONDUCTIVITY_X, r);
      if (iso) {
        ky = kx;
        kz = kx;
      } else {
        ky = ev(ADM_THERMAL_CONDUCTIVITY_Y, r);
        kz = ev(ADM_THERMAL_CONDUCTIVITY_Z, r);
      }
    };
    if constexpr (PK == 2)
      props(MergedEval(mB, mC, T));
    else
      props(typename EvalOf<PK>::type(m, mid, T));
    double fx, fy;
    if (rotated) {
      fx = (kx * cs * cs + ky * sn * sn) * gx + ((kx - ky) * sn * cs) * gy;
      fy = ((kx - ky) * sn * cs) * gx + (kx * sn * sn + ky * cs * cs) * gy;
    } else {
      // cos = 1, sin = 0: the reference's rotation reduces to k_x, k_y exactly
      fx = kx * gx;
      fy = ky * gy;
    }
    const double fz = gz * kz;
    const double jxw = kc.jxw[q];
    const double w0 = ((-irc) * fx) * jxw * kc.inv_h[0];
    const double w1 = ((-irc) * fy) * jxw * kc.inv_h[1];
    const double w2 = ((-irc) * fz) * jxw * kc.inv_h[2];
    // transposed collocation derivative, accumulated per quadrature point
#pragma unroll
    for (int k = 0; k < N; ++k) {
      A[q - qx + k] = __builtin_fma(kc.Dc[qx][k], w0, A[q - qx + k]);
      A[q - qy * N + k * N] =
          __builtin_fma(kc.Dc[qy][k], w1, A[q - qy * N + k * N]);
      A[q - qz * N * N + k * N * N] =
          __builtin_fma(kc.Dc[qz][k], w2, A[q - qz * N * N + k * N * N]);
    }
    if (near) {
      const double x = kc.origin[0] + (cc[0] + kc.xq[qx]) * kc.h[0];
      const double y = kc.origin[1] + (cc[1] + kc.xq[qy]) * kc.h[1];
      const double z = kc.origin[2] + (cc[2] + kc.xq[qz]) * kc.h[2];
      double qs = 0.0;
#pragma unroll 1
      for (int b = 0; b < a.n_beams; ++b)
        qs += beam_value(sbeams[b], x, y, z, kc.log_01);
      qs *= irc;
      A[q] += qs * jxw;
    }
  });
  if constexpr (NS > 1) {
    if (lq && any_liquid != dirty) a.liquid_dirty[c] = any_liquid ? 1 : 0;
  }

  // S^T along z, y, x
  sweep<N, 2, true, false>(kc.S, A, V);
  sweep<N, 1, true, false>(kc.S, V, A);
  sweep<N, 0, true, false>(kc.S, A, V);

#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int i = 0; i < N; ++i)
        unsafeAtomicAdd(dst + base + i + sy * j + sz * k,
                        V[i + N * (j + N * k)]);
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
