// Two-dimensional kernels of the matrix-free heat operator (gfx950): the
// reference instantiates its operators for dim 2 as well
// (source/instantiation.hh:121-179). A 2-D mesh is held as a box of
// n_x x 1 x n_z cells with a single node layer in y (nn[1] == 1), so cell
// and node numbers are the 2-D ones: cell x + n_x z, node i + nn_x k. The
// second coordinate of the reference's 2-D points is the height; the device
// keeps it in slot 2 of the geometry constants (h, inv_h, origin) and face
// ids 4, 5 (the reference's boundary ids 2, 3).
//
// Reference: source/ThermalOperator.templates.hh:479-751 (dim = 2),
// source/GoldakHeatSource.cc:91-148, source/ElectronBeamHeatSource.cc:95-158,
// source/CubeHeatSource.cc:73-107 (dim = 2 points).
#pragma once

#include "adm_kernels.hpp"

namespace adm {

// Heat source at a 2-D point (x, z). As the reference's dim = 2 sources do,
// the beam centre's components 0 and 1 are the point's x and height, the
// cube is tested on its first two components, and there is no five-axis
// rotation.
__device__ __forceinline__ double beam_value_2d(const adm_beam_state &s,
                                                double x, double z_,
                                                double log_01) {
  if (s.type == ADM_SOURCE_CUBE) {
    if (!s.on) return 0.0;
    if (!(x >= s.cube_min[0] && x <= s.cube_max[0])) return 0.0;
    if (!(z_ >= s.cube_min[1] && z_ <= s.cube_max[1])) return 0.0;
    return 1.0 * s.cube_value;
  }
  const double z = z_ - s.center[1];
  if (z + s.depth < 0.0) return 0.0;
  const double dx = x - s.center[0];
  const double xpy_squared = dx * dx;
  const double xpy_area = 5. * s.radius_squared - xpy_squared;
  if (xpy_area < 0.) return 0.0;
  double exponent, factor = 1.0;
  if (s.type == ADM_SOURCE_GOLDAK) {
    exponent = z * z * s.inv_depth_squared;
    exponent += xpy_squared * s.inv_radius_squared;
    exponent *= -3.0;
  } else {
    const double z_rel = z * s.inv_depth;
    double dist = z_rel * z_rel;
    dist *= -3.0;
    dist -= z_rel;
    dist -= z_rel;
    dist += 1.0;
    factor = dist;
    exponent = log_01 * xpy_squared;
    exponent *= s.inv_radius_squared;
  }
  return 1.0 * s.alpha * exp(exponent) * factor;
}

// Cell kernel, dim = 2: p+1 lanes per cell, lane b owning the x-line of
// nodes (and then of quadrature points) at height index b. x contractions
// run in registers, z contractions exchange the cell's lines through a
// wave-private LDS buffer. A wavefront takes a row segment of 64 / (p+1)
// cells along x at one cell row z. Operation order of every contraction is
// the oracle's: values S along x then z, gradients Dc along x and z,
// integration Dc^T along x plus Dc^T along z, the source, then S^T along z
// and x.
template <int P>
struct Cell2dLanes {
  static constexpr int N = P + 1;
  static constexpr int cells_per_wave = 64 / N;
  // exchange buffer of a cell: N x N doubles, one apart from the next
  // cell's (fewer LDS bank conflicts)
  static constexpr int stride = N * N + 1;
  static constexpr int waves_per_block = 4;
};

template <int P, int NS, int PK>
__global__ void __launch_bounds__(256)
    cell_kernel_2d(const CellArgs a, const double *__restrict__ src,
                   double *__restrict__ dst) {
  constexpr int N = P + 1, NQ = N * N;
  constexpr int CPW = Cell2dLanes<P>::cells_per_wave;
  constexpr int ST = Cell2dLanes<P>::stride;
  __shared__ double xs[Cell2dLanes<P>::waves_per_block * CPW * ST];
  __shared__ adm_beam_state sbeams[ADM_MAX_BEAMS];
  const ADM_CONSTANT MemberArgs *mb =
      a.members ? as_constant(a.members) + blockIdx.y : nullptr;
  const ADM_CONSTANT KConst &kc = *as_constant(mb ? mb->kc : a.kc);
  const MatConst &m = *(const MatConst *)&kc.mat;
  const ADM_GLOBAL uint32_t *meta_p = as_global(mb ? mb->meta : a.meta);
  const ADM_GLOBAL int32_t *slot_p =
      as_global(mb ? mb->powder_slot : a.powder_slot);
  ADM_GLOBAL double *powder_p = as_global(mb ? mb->powder : a.powder);
  ADM_GLOBAL double *liquid_p = as_global(mb ? mb->liquid : a.liquid);
  ADM_GLOBAL uint8_t *dirty_p =
      as_global(mb ? mb->liquid_dirty : a.liquid_dirty);
  if (mb) {
    src += (int64_t)blockIdx.y * a.member_stride;
    dst += (int64_t)blockIdx.y * a.member_stride;
  }
  // beams of the launch (or of the captured loop's current stage)
  {
    constexpr int words = sizeof(adm_beam_state) * ADM_MAX_BEAMS / 8;
    const int tid = threadIdx.x;
    const adm_beam_state *bsrc = a.beams;
    int bslot = 0;
    if (a.beam_table) {
      bslot = *as_global(a.beam_slot);
      bslot = bslot < 0 ? 0
                        : (bslot < a.beam_table_len ? bslot
                                                    : a.beam_table_len - 1);
      bsrc = a.beam_table + (int64_t)bslot * ADM_MAX_BEAMS;
    }
    if (tid < words)
      reinterpret_cast<uint64_t *>(sbeams)[tid] =
          reinterpret_cast<const uint64_t *>(bsrc)[tid];
    __syncthreads();
    if (mb && tid < ADM_MAX_BEAMS) {
      double pmod = a.beam_pmod[tid], den = a.beam_den[tid];
      if (a.beam_pd_table) {
        const double *pd =
            a.beam_pd_table + (int64_t)bslot * 2 * ADM_MAX_BEAMS;
        pmod = pd[tid];
        den = pd[ADM_MAX_BEAMS + tid];
      }
      sbeams[tid].alpha = member_alpha(sbeams[tid], mb->beam_gain[tid], pmod,
                                       den, kc.log_01);
    }
    __syncthreads();
  }
  int sources_on = a.sources_on;
  if (a.beam_table) {
    sources_on = 0;
    for (int b = 0; b < a.n_beams; ++b) sources_on |= sbeams[b].on ? 1 : 0;
  }
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int j = lane / N, lb = lane % N;
  const int r = blockIdx.x * (blockDim.x >> 6) + wave;
  // whole cells (all lanes of a cell) leave together
  if (r >= a.n_rows || j >= CPW) return;
  const int seg = r % a.nseg, rz = r / a.nseg;
  const int rx = seg * CPW + j;
  if (rx >= a.lext[0]) return;
  int cx = rx, cz = rz;
  if (a.color >= 0) {
    cx = 2 * rx + (a.color & 1);
    cz = 2 * rz + ((a.color >> 1) & 1);
  }
  const int64_t c = cx + (int64_t)a.nc[0] * cz;
  const uint32_t meta = meta_p[c];
  if (!(meta & kActive)) return;
  const int64_t base = (int64_t)cx * P + (int64_t)a.nn[0] * ((int64_t)cz * P + lb);
  double u[N];
#pragma unroll
  for (int i = 0; i < N; ++i) u[i] = src[base + i];
  const int mid = (meta >> kMaterialShift) & 3;
  const double solidus = m.prop[mid][ADM_SOLIDUS];
  const double liquidus = m.prop[mid][ADM_LIQUIDUS];
  const double latent = m.latent_factor[mid];
  const bool iso = m.isotropic[mid] != 0;
  const int own = N * lb;  // this lane's row of the cell's N x N points
  ADM_GLOBAL double *pw =
      (meta & kHasPowder) ? powder_p + (int64_t)slot_p[c] * NQ + own : nullptr;
  double p_in[N];
#pragma unroll
  for (int q = 0; q < N; ++q) p_in[q] = pw ? pw[q] : 0.0;
  bool dirty = false;
  ADM_GLOBAL double *lq = nullptr;
  if constexpr (NS > 1) {
    if (liquid_p) {
      dirty = dirty_p[c] != 0;
      lq = liquid_p + c * NQ + own;
    }
  }
  double *xb = xs + (wave * CPW + j) * ST;
  const double ihx = kc.inv_h[0], ihz = kc.inv_h[2];

  // values: S along x (registers), then along z (LDS)
  double t[N], V[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double acc = kc.S[q][0] * u[0];
#pragma unroll
    for (int k = 1; k < N; ++k) acc = __builtin_fma(kc.S[q][k], u[k], acc);
    t[q] = acc;
  }
#pragma unroll
  for (int q = 0; q < N; ++q) xb[own + q] = t[q];
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double acc = kc.S[lb][0] * xb[q];
#pragma unroll
    for (int k = 1; k < N; ++k)
      acc = __builtin_fma(kc.S[lb][k], xb[N * k + q], acc);
    V[q] = acc;
  }
  // gradients by the collocation derivative: x in registers, z from the
  // cell's values shared through LDS
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < N; ++q) xb[own + q] = V[q];
  wave_lds_sync();
  double gx[N], gz[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double ax = kc.Dc[q][0] * V[0];
    double az = kc.Dc[lb][0] * xb[q];
#pragma unroll
    for (int k = 1; k < N; ++k) {
      ax = __builtin_fma(kc.Dc[q][k], V[k], ax);
      az = __builtin_fma(kc.Dc[lb][k], xb[N * k + q], az);
    }
    gx[q] = ax * ihx;
    gz[q] = az * ihz;
  }

  // per-point physics
  const double det = kc.det, wb = kc.wq[lb];
  const double pz = kc.origin[2] + (cz + kc.xq[lb]) * kc.h[2];
  double w0[N], w1[N], val[N];
  bool line_liquid = false;
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const double T = V[q];
    double p = p_in[q];
    double r_[NS];
    double l = 0.0;
    // update_state_ratios (source/ThermalOperator.templates.hh:227-317)
    if constexpr (NS == 1) {
      r_[0] = 1.0;
    } else {
      const double mushy = (T - solidus) / (liquidus - solidus);
      l = T < solidus ? 0.0 : (T > liquidus ? 1.0 : mushy);
      if constexpr (NS == 2) {
        r_[1] = l;
        r_[0] = 1.0 - l;
      } else {
        const double am = 1.0 - l;
        p = (p < am) ? p : am;
        r_[1] = l;
        r_[2] = p;
        r_[0] = 1.0 - l - p;
      }
    }
    if (pw && p != p_in[q]) pw[q] = p;
    if constexpr (NS > 1) {
      if (lq && (l != 0.0 || dirty)) lq[q] = l;
      line_liquid |= l != 0.0;
    }
    double irc, kx, kz;
    auto props = [&](const auto &ev) {
      // get_inv_rho_cp (source/ThermalOperator.templates.hh:415-477)
      const double density = ev(ADM_DENSITY, r_);
      double specific_heat = ev(ADM_SPECIFIC_HEAT, r_);
      if constexpr (NS > 1) {
        const double is_mushy = r_[1] * r_[0];
        if (is_mushy > 0.0) specific_heat += latent;
      }
      irc = 1.0 / (density * specific_heat);
      kx = ev(ADM_THERMAL_CONDUCTIVITY_X, r_);
      kz = iso ? kx : ev(ADM_THERMAL_CONDUCTIVITY_Z, r_);
    };
    if constexpr (PK == 2)
      props(MergedEval(m.merged[mid].B, &m.merged[mid].coef[0][0][0][0], T));
    else
      props(typename EvalOf<PK>::type(m, mid, T));
    // dim = 2: the flux is (k_x g_x, k_z g_z) (no deposition rotation)
    const double fx = gx[q] * kx;
    const double fz = gz[q] * kz;
    // jxw as the host tabulates it: det * ((1 * w_x) * w_z)
    double wxz = 1.0;
    wxz *= kc.wq[q];
    wxz *= wb;
    const double jxw = det * wxz;
    w0[q] = ((-irc) * fx) * jxw * ihx;
    w1[q] = ((-irc) * fz) * jxw * ihz;
    val[q] = 0.0;
    if (sources_on) {
      const double px = kc.origin[0] + (cx + kc.xq[q]) * kc.h[0];
      double qs = 0.0;
#pragma unroll 1
      for (int b = 0; b < a.n_beams; ++b)
        qs += beam_value_2d(sbeams[b], px, pz, kc.log_01);
      qs *= irc;
      val[q] = qs * jxw;
    }
  }
  if constexpr (NS > 1) {
    // cell-wide "held liquid" flag over the cell's lanes
    const uint64_t ball = __ballot(line_liquid);
    const bool any_liquid = ((ball >> (j * N)) & ((1ull << N) - 1)) != 0;
    if (lq && lb == 0 && any_liquid != dirty) dirty_p[c] = any_liquid ? 1 : 0;
  }

  // integration: Dc^T along x (registers) plus Dc^T along z (LDS), the
  // source term, then S^T along z (LDS) and x (registers)
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double acc = kc.Dc[0][i] * w0[0];
#pragma unroll
    for (int q = 1; q < N; ++q) acc = __builtin_fma(kc.Dc[q][i], w0[q], acc);
    t[i] = acc;
  }
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < N; ++q) xb[own + q] = w1[q];
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double az = kc.Dc[0][lb] * xb[q];
#pragma unroll
    for (int k = 1; k < N; ++k)
      az = __builtin_fma(kc.Dc[k][lb], xb[N * k + q], az);
    t[q] = t[q] + az;
    if (sources_on) t[q] += val[q];
  }
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < N; ++q) xb[own + q] = t[q];
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < N; ++q) {  // S^T along z: lane = node row lb
    double acc = kc.S[0][lb] * xb[q];
#pragma unroll
    for (int k = 1; k < N; ++k)
      acc = __builtin_fma(kc.S[k][lb], xb[N * k + q], acc);
    t[q] = acc;
  }
  double out[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {  // S^T along x
    double acc = kc.S[0][i] * t[0];
#pragma unroll
    for (int q = 1; q < N; ++q) acc = __builtin_fma(kc.S[q][i], t[q], acc);
    out[i] = acc;
  }
  if (a.color >= 0) {
    // no other cell of this launch touches these nodes
#pragma unroll
    for (int i = 0; i < N; ++i) dst[base + i] += out[i];
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) unsafeAtomicAdd(dst + base + i, out[i]);
  }
}

// Face kernel, dim = 2 (source/ThermalOperator.templates.hh:619-751): a
// face is an edge of p+1 nodes; p+1 lanes per face, 64 / (p+1) faces per
// wavefront. Lane q loads node q of its face (along the face's tangential
// axis), evaluates the temperature at face point q, its powder state and
// boundary value, and, after the weighted values are shared, integrates
// node q into its slot of face_out (added to dst by the face gather pass).
template <int P>
struct Face2dLanes {
  static constexpr int N = P + 1;
  static constexpr int faces_per_wave = 64 / N;
  static constexpr int faces_per_block = 4 * faces_per_wave;
};

template <int P, int NS, int PK>
__global__ void __launch_bounds__(256)
    face_kernel_2d(const FaceArgs a, const double *__restrict__ src) {
  constexpr int N = P + 1;
  constexpr int FPW = Face2dLanes<P>::faces_per_wave;
  __shared__ double xf[4][FPW * N];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int jf = lane / N, q = lane % N;
  const ADM_CONSTANT MemberArgs *mb =
      a.members ? as_constant(a.members) + blockIdx.y : nullptr;
  ADM_GLOBAL double *face_out = as_global(mb ? mb->face_out : a.face_out);
  const int64_t fi = ((int64_t)blockIdx.x * 4 + wave) * FPW + jf;
  if (jf >= FPW || fi >= (mb ? mb->n_faces : a.n_faces)) return;
  const KConst &kc = *(const KConst *)as_global(mb ? mb->kc : a.kc);
  const MatConst &m = kc.mat;
  const ADM_GLOBAL FaceRef *frp = as_global(mb ? mb->faces : a.faces) + fi;
  const int64_t c = frp->cell;
  const int info = frp->info;
  ADM_GLOBAL double *face_powder =
      as_global(mb ? mb->face_powder : a.face_powder);
  if (mb) src += (int64_t)blockIdx.y * a.member_stride;
  // faces 0, 1 (normal x) and 4, 5 (normal z)
  const int f = info & 7, bt = (info >> 3) & 31, mid = (info >> 8) & 3;
  const int d = f / 2, side = f % 2;
  const int64_t nx = a.nn[0];
  const int64_t cx = c % a.nc[0], cz = c / a.nc[0];
  const int64_t base = cx * P + nx * cz * P + (side ? P : 0) * (d == 0 ? 1 : nx);
  const int64_t node = base + q * (d == 0 ? nx : 1);
  double *fx = &xf[wave][jf * N];
  fx[q] = src[node];
  wave_lds_sync();
  double T = kc.S[q][0] * fx[0];
#pragma unroll
  for (int k = 1; k < N; ++k) T = __builtin_fma(kc.S[q][k], fx[k], T);
  ADM_GLOBAL double *fp = NS == 3 ? face_powder + fi * N + q : nullptr;
  double powder = fp ? *fp : 0.0;
  double liquid;
  double r[NS];
  if (state_ratios<NS>(m, mid, T, powder, liquid, r)) *fp = powder;
  double val;
  if constexpr (PK == 2) {
    const MergedEval ev(m.merged[mid].B, &m.merged[mid].coef[0][0][0][0], T);
    val = face_value<NS>(m, mid, bt, T, powder, ev, r);
  } else {
    const typename EvalOf<PK>::type ev(m, mid, T);
    val = face_value<NS>(m, mid, bt, T, powder, ev, r);
  }
  const double w = val * kc.fjxw[d][q];
  wave_lds_sync();
  fx[q] = w;
  wave_lds_sync();
  double out = kc.S[0][q] * fx[0];
#pragma unroll
  for (int k = 1; k < N; ++k) out = __builtin_fma(kc.S[k][q], fx[k], out);
  face_out[fi * N + q] = out;
}

// Lumped Gauss-Lobatto mass, dim = 2, inverted: one lane per node sums
// det * (w_x * w_z) over its active cells in increasing cell order.
template <int P>
__global__ void __launch_bounds__(256)
    inverse_mass_kernel_2d(const uint32_t *__restrict__ meta,
                           const double *__restrict__ wgl, double det,
                           int32_t nc0, int32_t nc2, int64_t n_dofs,
                           double *__restrict__ inv_mass) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= n_dofs) return;
  const int64_t nn0 = (int64_t)P * nc0 + 1;
  const int64_t id[2] = {node % nn0, node / nn0};
  const int32_t nc[2] = {nc0, nc2};
  int cand_c[2][2], cand_l[2][2], ncand[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
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
  for (int kz = 0; kz < ncand[1]; ++kz)
    for (int kx = 0; kx < ncand[0]; ++kx) {
      const int64_t c = cand_c[0][kx] + (int64_t)nc0 * cand_c[1][kz];
      if (!(meta[c] & kActive)) continue;
      const double w = wgl[cand_l[0][kx]] * wgl[cand_l[1][kz]];
      sum += det * w;
    }
  inv_mass[node] = sum > 0.0 ? 1.0 / sum : 0.0;
}

}  // namespace adm
