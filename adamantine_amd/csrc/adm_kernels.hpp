// HIP kernels of the matrix-free heat operator (gfx950).
//
// Cell kernel: one lane per cell. The (p+1)^3 nodal values are gathered into
// registers and the whole sum factorisation (values, collocation gradients,
// transposed integration) runs in registers in FP64 with explicit FMA chains.
// The per-quadrature-point physics (phase state, rho c_p with latent heat,
// k(T), deposition rotation, heat sources) and the boundary-face terms of the
// cell are fused in. Reference: source/ThermalOperator.templates.hh:479-751.
#pragma once

#include "adm_device.hpp"

namespace adm {

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

// Boundary type of face F of cell (cx,cy,cz): Boundary::get_boundary_type of
// the colorized id 2*d+side on the domain boundary, of the internal face id
// (numbers::internal_face_boundary_id) otherwise.
template <int F>
__device__ __forceinline__ int face_boundary_type(const KConst &kc,
                                                  const int (&cc)[3],
                                                  const int32_t (&nc)[3]) {
  constexpr int d = F / 2, side = F % 2;
  const bool domain = side ? (cc[d] == nc[d] - 1) : (cc[d] == 0);
  const unsigned bid = domain ? (unsigned)F : 0xffffffffu;
  const int n = kc.n_boundary_ids;
  return bid < (unsigned)(n - 1) ? kc.boundary_type[bid]
                                 : kc.boundary_type[n - 1];
}

// Boundary value at one face quadrature point
// (source/ThermalOperator.templates.hh:691-746), already weighted by JxW.
template <int NS, bool TABLE>
__device__ __noinline__ double face_qp(const KConst &kc, int mid, int bt,
                                       double T, double *fp_q, double fjxw) {
  const MatConst &m = kc.mat;
  double tp[ADM_MAX_POLY];
  temperature_powers(m, T, tp);
  double powder = fp_q ? *fp_q : 0.0;
  double liquid;
  double r[NS];
  if (state_ratios<NS>(m, mid, T, powder, liquid, r)) *fp_q = powder;
  const double irc = inv_rho_cp<NS, TABLE>(m, mid, r, T, tp);
  double conv_T = 0.0, conv_h = 0.0, rad_T = 0.0, rad_h = 0.0;
  if (bt & ADM_BC_CONVECTIVE) {
    conv_T = m.prop[mid][ADM_CONVECTION_TEMPERATURE_INFTY];
    conv_h = property<NS, TABLE>(m, ADM_CONVECTION_HEAT_TRANSFER_COEF, mid, r,
                                 T, tp);
  }
  if (bt & ADM_BC_RADIATIVE) {
    rad_T = m.prop[mid][ADM_RADIATION_TEMPERATURE_INFTY];
    rad_h = property<NS, TABLE>(m, ADM_EMISSIVITY, mid, r, T, tp) *
            kStefanBoltzmann * (T + rad_T) * (T * T + rad_T * rad_T);
  }
  const double bv = -irc * (conv_h * (T - conv_T) + rad_h * (T - rad_T));
  return bv * fjxw;
}

// Face term of face F (source/ThermalOperator.templates.hh:619-751),
// added into the cell's nodal output `out`.
template <int P, int NS, bool TABLE, int F>
__device__ __forceinline__ void face_term(const CellArgs &a, const KConst &kc,
                                          const int (&cc)[3], int64_t base,
                                          int mid, int64_t pslot,
                                          const double *__restrict__ src,
                                          double *out) {
  constexpr int N = P + 1, NF = N * N;
  constexpr int d = F / 2, side = F % 2;
  constexpr int t1 = d == 0 ? 1 : 0;
  constexpr int t2 = d == 2 ? 1 : 2;
  constexpr int ln = side ? P : 0;
  const int64_t nstride[3] = {1, (int64_t)a.nn[0],
                              (int64_t)a.nn[0] * (int64_t)a.nn[1]};
  constexpr int lstride[3] = {1, N, N * N};
  double fu[NF], ft[NF], fv[NF];
#pragma unroll
  for (int j = 0; j < NF; ++j) {
    const int i1 = j % N, i2 = j / N;
    fu[j] = src[base + ln * nstride[d] + i1 * nstride[t1] + i2 * nstride[t2]];
  }
  sweep<N, 0, false, false, 2>(kc.S, fu, ft);
  sweep<N, 1, false, false, 2>(kc.S, ft, fv);
  const int bt = face_boundary_type<F>(kc, cc, a.nc);
  double *fp = pslot >= 0 ? a.face_powder + (pslot * 6 + F) * NF : nullptr;
#pragma unroll
  for (int q = 0; q < NF; ++q)
    fu[q] = face_qp<NS, TABLE>(kc, mid, bt, fv[q], fp ? fp + q : nullptr,
                               kc.fjxw[d][q]);
  sweep<N, 1, true, false, 2>(kc.S, fu, ft);
  sweep<N, 0, true, false, 2>(kc.S, ft, fv);
#pragma unroll
  for (int j = 0; j < NF; ++j) {
    const int i1 = j % N, i2 = j / N;
    out[ln * lstride[d] + i1 * lstride[t1] + i2 * lstride[t2]] += fv[j];
  }
}

struct QpFlux {
  double w0, w1, w2, val, liquid;
};

// Per-quadrature-point physics of cell_local_apply
// (source/ThermalOperator.templates.hh:522-611): phase state update,
// 1/(rho c_p), conductivity with the deposition rotation, heat sources.
// Returns the JxW-weighted reference-space flux and source value.
template <int NS, bool TABLE>
__device__ __noinline__ QpFlux qp_physics(const KConst &kc,
                                          const adm_beam_state *beams,
                                          int n_beams, int sources_on, int mid,
                                          double T, double gx, double gy,
                                          double gz, double x, double y,
                                          double z, double jxw, double *pw_q,
                                          int rotated, double cs, double sn) {
  const MatConst &m = kc.mat;
  double tp[ADM_MAX_POLY];
  temperature_powers(m, T, tp);
  double powder = pw_q ? *pw_q : 0.0;
  double liquid;
  double r[NS];
  if (state_ratios<NS>(m, mid, T, powder, liquid, r)) *pw_q = powder;
  const double irc = inv_rho_cp<NS, TABLE>(m, mid, r, T, tp);
  const double kx =
      property<NS, TABLE>(m, ADM_THERMAL_CONDUCTIVITY_X, mid, r, T, tp);
  double ky, kz;
  if (m.isotropic[mid]) {
    ky = kx;
    kz = kx;
  } else {
    ky = property<NS, TABLE>(m, ADM_THERMAL_CONDUCTIVITY_Y, mid, r, T, tp);
    kz = property<NS, TABLE>(m, ADM_THERMAL_CONDUCTIVITY_Z, mid, r, T, tp);
  }
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
  QpFlux out;
  out.liquid = liquid;
  out.w0 = ((-irc) * fx) * jxw * kc.inv_h[0];
  out.w1 = ((-irc) * fy) * jxw * kc.inv_h[1];
  out.w2 = ((-irc) * fz) * jxw * kc.inv_h[2];
  out.val = 0.0;
  if (sources_on) {
    double qs = 0.0;
#pragma unroll
    for (int b = 0; b < ADM_MAX_BEAMS; ++b)
      if (b < n_beams) qs += beam_value(beams[b], x, y, z, kc.log_01);
    qs *= irc;
    out.val = qs * jxw;
  }
  return out;
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

template <int P, int NS, bool TABLE>
__global__ void __launch_bounds__(256)
    cell_kernel(const CellArgs a, const double *__restrict__ src,
                double *__restrict__ dst, int64_t n_cells) {
  constexpr int N = P + 1, NQ = N * N * N;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const uint32_t meta = a.meta[c];
  if (!(meta & kActive)) return;
  const KConst &kc = *a.kc;
  const int mid = (meta >> kMaterialShift) & 3;
  int cc[3];
  cc[0] = (int)(c % a.nc[0]);
  cc[1] = (int)((c / a.nc[0]) % a.nc[1]);
  cc[2] = (int)(c / ((int64_t)a.nc[0] * a.nc[1]));
  const int64_t sy = a.nn[0], sz = (int64_t)a.nn[0] * a.nn[1];
  const int64_t base = (int64_t)cc[0] * P + sy * cc[1] * P + sz * cc[2] * P;

  double A[NQ], B[NQ];
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int i = 0; i < N; ++i)
        A[i + N * (j + N * k)] = src[base + i + sy * j + sz * k];

  // values at the quadrature points: S along x, y, z (V kept in B)
  sweep<N, 0, false, false>(kc.S, A, B);
  sweep<N, 1, false, false>(kc.S, B, A);
  sweep<N, 2, false, false>(kc.S, A, B);
  const double *V = B;

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
  const int rotated = (meta & kRotated) ? 1 : 0;
  if (rotated) {
    cs = a.dep_cos[c];
    sn = a.dep_sin[c];
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) A[q] = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int qx = q % N, qy = (q / N) % N, qz = q / (N * N);
    // reference gradients by the collocation derivative (chains of sweep())
    const double g0 = line_dot<N>(kc.Dc, qx, V, q - qx, 1);
    const double g1 = line_dot<N>(kc.Dc, qy, V, q - qy * N, N);
    const double g2 = line_dot<N>(kc.Dc, qz, V, q - qz * N * N, N * N);
    const double x = kc.origin[0] + (cc[0] + kc.xq[qx]) * kc.h[0];
    const double y = kc.origin[1] + (cc[1] + kc.xq[qy]) * kc.h[1];
    const double z = kc.origin[2] + (cc[2] + kc.xq[qz]) * kc.h[2];
    const QpFlux f = qp_physics<NS, TABLE>(
        kc, a.beams, a.n_beams, a.sources_on, mid, V[q], g0 * kc.inv_h[0],
        g1 * kc.inv_h[1], g2 * kc.inv_h[2], x, y, z, kc.jxw[q],
        pw ? pw + q : nullptr, rotated, cs, sn);
    if constexpr (NS > 1) {
      // the stored liquid ratio (a function of T) is written where it is
      // non-zero, and everywhere in a cell that held liquid before
      if (lq && (f.liquid != 0.0 || dirty)) lq[q] = f.liquid;
      any_liquid |= f.liquid != 0.0;
    }
    // transposed collocation derivative, accumulated per quadrature point
#pragma unroll
    for (int k = 0; k < N; ++k) {
      A[q - qx + k] = __builtin_fma(kc.Dc[qx][k], f.w0, A[q - qx + k]);
      A[q - qy * N + k * N] =
          __builtin_fma(kc.Dc[qy][k], f.w1, A[q - qy * N + k * N]);
      A[q - qz * N * N + k * N * N] =
          __builtin_fma(kc.Dc[qz][k], f.w2, A[q - qz * N * N + k * N * N]);
    }
    if (a.sources_on) A[q] += f.val;
  }
  if constexpr (NS > 1) {
    if (lq && any_liquid != dirty) a.liquid_dirty[c] = any_liquid ? 1 : 0;
  }

  // S^T along z, y, x
  sweep<N, 2, true, false>(kc.S, A, B);
  sweep<N, 1, true, false>(kc.S, B, A);
  sweep<N, 0, true, false>(kc.S, A, B);

  const uint32_t faces = (meta >> kFaceShift) & 63u;
  if (faces) {
    if (faces & 1u) face_term<P, NS, TABLE, 0>(a, kc, cc, base, mid, pslot, src, B);
    if (faces & 2u) face_term<P, NS, TABLE, 1>(a, kc, cc, base, mid, pslot, src, B);
    if (faces & 4u) face_term<P, NS, TABLE, 2>(a, kc, cc, base, mid, pslot, src, B);
    if (faces & 8u) face_term<P, NS, TABLE, 3>(a, kc, cc, base, mid, pslot, src, B);
    if (faces & 16u) face_term<P, NS, TABLE, 4>(a, kc, cc, base, mid, pslot, src, B);
    if (faces & 32u) face_term<P, NS, TABLE, 5>(a, kc, cc, base, mid, pslot, src, B);
  }

#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int i = 0; i < N; ++i)
        unsafeAtomicAdd(dst + base + i + sy * j + sz * k,
                        B[i + N * (j + N * k)]);
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
