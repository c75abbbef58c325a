// HIP kernels of the matrix-free heat operator (gfx950).
//
// Cell kernel: (p+1)^2 lanes per cell, one x-line each; the sum
// factorisation runs in FP64 with explicit FMA chains, exchanging lines
// through LDS, and the per-quadrature-point physics (phase state, rho c_p
// with latent heat, k(T), deposition rotation, heat sources) is fused in.
// Boundary faces with a convective/radiative term have their own kernel.
// Reference: source/ThermalOperator.templates.hh:479-751.
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

// The pointers carried in the argument structs are generic; kernels
// re-qualify them so that uniform constants become scalar loads and per-cell
// data global (not flat) loads, which LDS waits do not have to wait for.
#define ADM_GLOBAL __attribute__((address_space(1)))
#define ADM_CONSTANT __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ ADM_GLOBAL T *as_global(T *p) {
  return (ADM_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ const ADM_CONSTANT T *as_constant(const T *p) {
  return (const ADM_CONSTANT T *)p;
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

// cells per wavefront of the cell kernel for degree p
inline int cells_per_wave(int p) {
  return 64 / ((p + 1) * (p + 1));
}

// LDS ordering between the lanes of one wavefront
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Face kernel (source/ThermalOperator.templates.hh:619-751) over the
// boundary faces of the active domain that carry a convective or radiative
// term: (p+1)^2 lanes per face, 64 / (p+1)^2 faces per wavefront. Lane q
// loads node q of its face, evaluates the temperature at face point q (the
// x- then y-sweep of S, in the oracle's order, from the face's nodes shared
// through LDS), the point's powder state and boundary value; the weighted
// values are shared again and lane j integrates node j (S^T along y, then
// x) and adds it to dst atomically.
template <int P>
struct FaceLanes {
  static constexpr int N = P + 1;
  static constexpr int lanes_per_face = N * N;
  static constexpr int faces_per_wave = 64 / (N * N);
  static constexpr int faces_per_block = 4 * faces_per_wave;
};

template <int P, int NS, int PK>
__global__ void __launch_bounds__(256)
    face_kernel(const FaceArgs a, const double *__restrict__ src,
                double *__restrict__ dst) {
  constexpr int N = P + 1, NF = N * N;
  constexpr int FPW = FaceLanes<P>::faces_per_wave;
  __shared__ double xf[4][FPW * NF];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int jf = lane / NF, q = lane % NF;
  const int qa = q % N, qb = q / N;
  const ADM_CONSTANT MemberArgs *mb =
      a.members ? as_constant(a.members) + blockIdx.y : nullptr;
  ADM_GLOBAL double *face_out = as_global(mb ? mb->face_out : a.face_out);
  const int64_t fi = ((int64_t)blockIdx.x * 4 + wave) * FPW + jf;
  // whole faces (all lanes of a face) agree
  if (jf >= FPW || fi >= (mb ? mb->n_faces : a.n_faces)) return;
  const KConst &kc = *(const KConst *)as_global(mb ? mb->kc : a.kc);
  const MatConst &m = kc.mat;
  const ADM_GLOBAL FaceRef *frp = as_global(mb ? mb->faces : a.faces) + fi;
  FaceRef fr;
  fr.cell = frp->cell;
  fr.info = frp->info;
  ADM_GLOBAL double *face_powder =
      as_global(mb ? mb->face_powder : a.face_powder);
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
  // this lane's node (face-local index q = a + N b)
  const int64_t node = base + qa * ns[t1] + qb * ns[t2];
  double *fx = &xf[wave][jf * NF];
  fx[q] = src[node];
  wave_lds_sync();
  // T at point (qa, qb): ft[b] = sum_k S[qa][k] u[k + N b], then along b
  double ft[N];
#pragma unroll
  for (int bb = 0; bb < N; ++bb) {
    double acc = kc.S[qa][0] * fx[N * bb];
#pragma unroll
    for (int k = 1; k < N; ++k)
      acc = __builtin_fma(kc.S[qa][k], fx[k + N * bb], acc);
    ft[bb] = acc;
  }
  double T = kc.S[qb][0] * ft[0];
#pragma unroll
  for (int k = 1; k < N; ++k) T = __builtin_fma(kc.S[qb][k], ft[k], T);
  ADM_GLOBAL double *fp = NS == 3 ? face_powder + fi * NF + q : nullptr;
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
  // node (qa, qb): S^T along y (ft[i] = sum_k S[k][qb] w[i + N k]), then
  // along x
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double acc = kc.S[0][qb] * fx[i];
#pragma unroll
    for (int k = 1; k < N; ++k)
      acc = __builtin_fma(kc.S[k][qb], fx[i + N * k], acc);
    ft[i] = acc;
  }
  double out = kc.S[0][qa] * ft[0];
#pragma unroll
  for (int i = 1; i < N; ++i) out = __builtin_fma(kc.S[i][qa], ft[i], out);
  face_out[fi * NF + q] = out;
}

// Second pass of the face term: one lane per boundary node adds its faces'
// slots, in increasing face order, to dst.
static __global__ void __launch_bounds__(256)
    face_gather_kernel(const FaceArgs a, double *__restrict__ dst) {
  const ADM_CONSTANT MemberArgs *mb =
      a.members ? as_constant(a.members) + blockIdx.y : nullptr;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (mb ? mb->n_fnodes : a.n_fnodes)) return;
  const ADM_GLOBAL double *fo = as_global(mb ? mb->face_out : a.face_out);
  const ADM_GLOBAL int32_t *off = as_global(mb ? mb->fn_offset : a.fn_offset);
  const ADM_GLOBAL int32_t *slot = as_global(mb ? mb->fn_slot : a.fn_slot);
  const int64_t node = as_global(mb ? mb->fn_node : a.fn_node)[k];
  if (mb) dst += (int64_t)blockIdx.y * a.member_stride;
  double sum = 0.0;
  for (int32_t e = off[k]; e < off[k + 1]; ++e) sum += fo[slot[e]];
  dst[node] += sum;
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
// (p+1)^2 lanes per cell: a lane owns one x-line of (p+1) nodes or
// quadrature points, indexed by its two other coordinates (a, b). x
// contractions run in the lane's registers; y and z contractions exchange
// lines through a wave-private LDS buffer (the lanes of a cell share a
// wavefront, so no workgroup barrier is involved). The operation order of
// every contraction is the oracle's: values by S along x, y, z; gradients by
// the collocation derivative Dc; integration by Dc^T along x, y, z, the
// source term, then S^T along z, y, x. A lane does the physics of its line's
// p+1 quadrature points.
//
// A wavefront handles a row segment of 64 / (p+1)^2 cells along x at fixed
// (y, z), so cell coordinates are uniform scalar arithmetic. Uniform constants
// are read into registers before the first global store (the stores could
// otherwise alias them and force reloads); entries indexed by the lane's
// (a, b) come from a block-shared LDS copy.
template <int P>
struct CellLanes {
  static constexpr int N = P + 1;
  static constexpr int lanes_per_cell = N * N;
  static constexpr int cells_per_wave = 64 / (N * N);
  // exchange buffers per cell, N^3 doubles each; consecutive cells'
  // buffers are 5 doubles apart (fewer LDS bank conflicts: 3 % less
  // operator time than no gap on the 128x64x32 Q2 workload)
  static constexpr int buffers = 2;
  static constexpr int stride = buffers * N * N * N + 5;
};

// Block-shared copies of the small constant tables (read by lane-varying
// indices from LDS, which global stores cannot alias).
struct CellConstLds {
  double S[kMaxN][kMaxN], Dc[kMaxN][kMaxN], xq[kMaxN], wq[kMaxN];
  double solidus[ADM_MAX_MATERIALS], liquidus[ADM_MAX_MATERIALS];
  double latent[ADM_MAX_MATERIALS], iso[ADM_MAX_MATERIALS];
};

// Cell coordinates of lane-cell j of row segment r of a launch.
struct RowCell {
  int cc[3];
  bool ok;
};
// Workgroups are dispatched round-robin over the 8 XCDs (block b on XCD
// b % 8). Map block b to a logical block so that each XCD runs one
// contiguous run of the grid: row segments that share node lines (the
// neighbours in y and z) then share that XCD's L2. A bijection on [0, g).
__device__ __forceinline__ int xcd_block(int b, int g) {
  const int q = g >> 3, rem = g & 7, x = b & 7, k = b >> 3;
  return x < rem ? x * (q + 1) + k : rem * (q + 1) + (x - rem) * q + k;
}

__device__ __forceinline__ RowCell row_cell(const CellArgs &a, int r, int j,
                                            int cpw) {
  RowCell rc;
  const int seg = r % a.nseg, yz = r / a.nseg;
  const int ry = yz % a.lext[1], rz = yz / a.lext[1];
  const int rx = seg * cpw + j;
  rc.ok = j < cpw && rx < a.lext[0];
  if (a.color >= 0) {
    rc.cc[0] = 2 * rx + (a.color & 1);
    rc.cc[1] = 2 * ry + ((a.color >> 1) & 1);
    rc.cc[2] = 2 * rz + ((a.color >> 2) & 1);
  } else {
    rc.cc[0] = rx;
    rc.cc[1] = ry;
    rc.cc[2] = rz;
  }
  return rc;
}

template <int P, int NS, int PK>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(P >= 3 ? 2 : 3)))
    cell_kernel(const CellArgs a, const double *__restrict__ src,
                double *__restrict__ dst) {
  constexpr int N = P + 1, NL = N * N, N3 = N * N * N;
  constexpr int CPW = CellLanes<P>::cells_per_wave;
  // ensemble member of this block (uniform), else the single operator
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
  const ADM_GLOBAL double *cos_p = as_global(mb ? mb->dep_cos : a.dep_cos);
  const ADM_GLOBAL double *sin_p = as_global(mb ? mb->dep_sin : a.dep_sin);
  if (mb) {
    src += (int64_t)blockIdx.y * a.member_stride;
    dst += (int64_t)blockIdx.y * a.member_stride;
  }
  extern __shared__ double lds[];
  __shared__ adm_beam_state sbeams[ADM_MAX_BEAMS];
  __shared__ CellConstLds sk;
  const int n_mat = kc.mat.n_materials;
  // block prologue: beams (one 8-byte word per lane), constant tables and
  // merged tables
  {
    constexpr int words = sizeof(adm_beam_state) * ADM_MAX_BEAMS / 8;
    static_assert(sizeof(adm_beam_state) % 8 == 0, "beam state layout");
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
    constexpr int nn = kMaxN * kMaxN;
    if (tid < nn) {
      (&sk.S[0][0])[tid] = (&kc.S[0][0])[tid];
      (&sk.Dc[0][0])[tid] = (&kc.Dc[0][0])[tid];
    } else if (tid < nn + kMaxN) {
      sk.xq[tid - nn] = kc.xq[tid - nn];
      sk.wq[tid - nn] = kc.wq[tid - nn];
    } else if (tid >= 64 && tid < 64 + ADM_MAX_MATERIALS) {
      const int k = tid - 64;
      sk.solidus[k] = kc.mat.prop[k][ADM_SOLIDUS];
      sk.liquidus[k] = kc.mat.prop[k][ADM_LIQUIDUS];
      sk.latent[k] = kc.mat.latent_factor[k];
      sk.iso[k] = kc.mat.isotropic[k] ? 1.0 : 0.0;
    }
    if constexpr (PK == 2) {
      const int total = n_mat * kMergedLds;
      for (int i = tid; i < total; i += blockDim.x) {
        const int mm = i / kMergedLds, k = i % kMergedLds;
        lds[i] = (&kc.mat.merged[mm].B[0])[k];
      }
    }
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
    if (mb) __syncthreads();
  }
  // any beam on (from the table row's states when the beams come from it)
  int sources_on = a.sources_on;
  if (a.beam_table) {
    sources_on = 0;
    for (int b = 0; b < a.n_beams; ++b) sources_on |= sbeams[b].on ? 1 : 0;
  }
  // uniform constants (scalar loads; S and Dc are read from the LDS copy)
  double xq[N], wq[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    xq[q] = kc.xq[q];
    wq[q] = kc.wq[q];
  }
  const double det = kc.det, log_01 = kc.log_01;
  double ih[3], hh[3], org[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ih[d] = kc.inv_h[d];
    hh[d] = kc.h[d];
    org[d] = kc.origin[d];
  }
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int j = lane / NL, L = lane % NL;
  const int la = L % N, lb = L / N;  // the line's y and z indices
  const double wa = sk.wq[la], wb = sk.wq[lb];
  const double xqa = sk.xq[la], xqb = sk.xq[lb];
  const int own = N * la + NL * lb;  // lane's slot in a [z][y][x] buffer
  double *xb = lds + (PK == 2 ? n_mat * kMergedLds : 0) +
               (wave * CPW + (j < CPW ? j : 0)) * CellLanes<P>::stride;
  double *xb2 = xb + N3;
  const int64_t sy = a.nn[0], sz = (int64_t)a.nn[0] * a.nn[1];
  const double *mBall = lds;

  // this wavefront's row segment
  const int bx = a.xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const int r = bx * (blockDim.x >> 6) + wave;
  if (r >= a.n_rows) return;
  const RowCell rc = row_cell(a, r, j, CPW);
  if (!rc.ok) return;  // whole cells (all lanes) leave together
  const int64_t c =
      rc.cc[0] + (int64_t)a.nc[0] * (rc.cc[1] + (int64_t)a.nc[1] * rc.cc[2]);
  const int64_t base = (int64_t)rc.cc[0] * P +
                       sy * ((int64_t)rc.cc[1] * P + la) +
                       sz * ((int64_t)rc.cc[2] * P + lb);
  double u[N];
#pragma unroll
  for (int i = 0; i < N; ++i) u[i] = src[base + i];
  const uint32_t meta = meta_p[c];
  const int32_t slot = slot_p[c];
  {
    if (!(meta & kActive)) return;
      const int mid = (meta >> kMaterialShift) & 3;
      const double solidus = sk.solidus[mid], liquidus = sk.liquidus[mid],
                   latent = sk.latent[mid];
      const bool iso = sk.iso[mid] != 0.0;
      // powder and liquid state of the line's points (loaded early)
      const int64_t pslot = (meta & kHasPowder) ? (int64_t)slot : -1;
      ADM_GLOBAL double *pw =
          pslot >= 0 ? powder_p + pslot * N3 + own : nullptr;
      double p_in[N];
#pragma unroll
      for (int q = 0; q < N; ++q) p_in[q] = pw ? pw[q] : 0.0;
      bool dirty = false;
      ADM_GLOBAL double *lq = nullptr;
      if constexpr (NS > 1) {
        if (liquid_p) {
          dirty = dirty_p[c] != 0;
          lq = liquid_p + c * N3 + own;
        }
      }
      const bool rotated = (meta & kRotated) != 0;
      double cs = 1.0, sn = 0.0;
      if (rotated) {
        cs = cos_p[c];
        sn = sin_p[c];
      }

      // values: S along x (registers), y and z (LDS)
      double t[N], V[N];
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double acc = sk.S[q][0] * u[0];
#pragma unroll
        for (int k = 1; k < N; ++k) acc = __builtin_fma(sk.S[q][k], u[k], acc);
        t[q] = acc;
      }
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) xb[own + q] = t[q];
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double acc = sk.S[la][0] * xb[NL * lb + q];
#pragma unroll
        for (int k = 1; k < N; ++k)
          acc = __builtin_fma(sk.S[la][k], xb[N * k + NL * lb + q], acc);
        t[q] = acc;
      }
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) xb[own + q] = t[q];
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double acc = sk.S[lb][0] * xb[N * la + q];
#pragma unroll
        for (int k = 1; k < N; ++k)
          acc = __builtin_fma(sk.S[lb][k], xb[N * la + NL * k + q], acc);
        V[q] = acc;
      }
      // gradients by the collocation derivative
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) xb[own + q] = V[q];
      wave_lds_sync();
      double gx[N], gy[N], gz[N];
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double ax = sk.Dc[q][0] * V[0];
        double ay = sk.Dc[la][0] * xb[NL * lb + q];
        double az = sk.Dc[lb][0] * xb[N * la + q];
#pragma unroll
        for (int k = 1; k < N; ++k) {
          ax = __builtin_fma(sk.Dc[q][k], V[k], ax);
          ay = __builtin_fma(sk.Dc[la][k], xb[N * k + NL * lb + q], ay);
          az = __builtin_fma(sk.Dc[lb][k], xb[N * la + NL * k + q], az);
        }
        gx[q] = ax * ih[0];
        gy[q] = ay * ih[1];
        gz[q] = az * ih[2];
      }

      // per-point physics
      const double *mB = mBall + mid * kMergedLds;
      const double *mC = mB + kMergedMax;
      bool near = false;
      if (sources_on) {
        double lo[3], hi[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          lo[d] = org[d] + rc.cc[d] * hh[d];
          hi[d] = org[d] + (rc.cc[d] + 1) * hh[d];
        }
#pragma unroll 1
        for (int b = 0; b < a.n_beams; ++b)
          near |= beam_may_touch(sbeams[b], lo, hi);
      }
      const double py = org[1] + (rc.cc[1] + xqa) * hh[1];
      const double pz = org[2] + (rc.cc[2] + xqb) * hh[2];
      int kint[N];
      if constexpr (PK == 2) merged_interval<N>(mB, V, kint);
      double w0[N], w1[N], w2[N], val[N];
      bool line_liquid = false;
#pragma unroll
      for (int q = 0; q < N; ++q) {
        const double T = V[q];
        double p = p_in[q];
        double r_[NS];
        double l = 0.0;
        // update_state_ratios (source/ThermalOperator.templates.hh:227-317),
        // branch-free
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
        double irc, kx, ky, kzz;
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
          if (iso) {
            ky = kx;
            kzz = kx;
          } else {
            ky = ev(ADM_THERMAL_CONDUCTIVITY_Y, r_);
            kzz = ev(ADM_THERMAL_CONDUCTIVITY_Z, r_);
          }
        };
        if constexpr (PK == 2)
          props(MergedEval(mB, mC, T, kint[q]));
        else
          props(typename EvalOf<PK>::type(m, mid, T));
        double fx, fy;
        if (rotated) {
          fx = (kx * cs * cs + ky * sn * sn) * gx[q] +
               ((kx - ky) * sn * cs) * gy[q];
          fy = ((kx - ky) * sn * cs) * gx[q] +
               (kx * sn * sn + ky * cs * cs) * gy[q];
        } else {
          // cos = 1, sin = 0: the reference's rotation gives k_x g_x, k_y g_y
          fx = kx * gx[q];
          fy = ky * gy[q];
        }
        const double fz = gz[q] * kzz;
        // jxw as the host tabulates it: det * ((1 * w_x) * w_y) * w_z
        double wxyz = 1.0;
        wxyz *= wq[q];
        wxyz *= wa;
        wxyz *= wb;
        const double jxw = det * wxyz;
        w0[q] = ((-irc) * fx) * jxw * ih[0];
        w1[q] = ((-irc) * fy) * jxw * ih[1];
        w2[q] = ((-irc) * fz) * jxw * ih[2];
        val[q] = 0.0;
        if (near) {
          const double px = org[0] + (rc.cc[0] + xq[q]) * hh[0];
          double qs = 0.0;
#pragma unroll 1
          for (int b = 0; b < a.n_beams; ++b)
            qs += beam_value(sbeams[b], px, py, pz, log_01);
          qs *= irc;
          val[q] = qs * jxw;
        }
      }
      if constexpr (NS > 1) {
        // cell-wide "held liquid" flag over the cell's lanes
        const uint64_t ball = __ballot(line_liquid);
        const bool any_liquid = ((ball >> (j * NL)) & ((1ull << NL) - 1)) != 0;
        if (lq && L == 0 && any_liquid != dirty)
          dirty_p[c] = any_liquid ? 1 : 0;
      }

      // integration: Dc^T along x (registers), y and z (LDS), the source
      // term, then S^T along z, y, x
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double acc = sk.Dc[0][i] * w0[0];
#pragma unroll
        for (int q = 1; q < N; ++q) acc = __builtin_fma(sk.Dc[q][i], w0[q], acc);
        t[i] = acc;
      }
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) {
        xb[own + q] = w1[q];
        xb2[own + q] = w2[q];
      }
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double ay = sk.Dc[0][la] * xb[NL * lb + q];
        double az = sk.Dc[0][lb] * xb2[N * la + q];
#pragma unroll
        for (int k = 1; k < N; ++k) {
          ay = __builtin_fma(sk.Dc[k][la], xb[N * k + NL * lb + q], ay);
          az = __builtin_fma(sk.Dc[k][lb], xb2[N * la + NL * k + q], az);
        }
        t[q] = t[q] + ay;
        t[q] = t[q] + az;
        if (sources_on) t[q] += val[q];
      }
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) xb[own + q] = t[q];
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) {  // S^T along z: lane = (y = la, node z = lb)
        double acc = sk.S[0][lb] * xb[N * la + q];
#pragma unroll
        for (int k = 1; k < N; ++k)
          acc = __builtin_fma(sk.S[k][lb], xb[N * la + NL * k + q], acc);
        u[q] = acc;
      }
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) xb[own + q] = u[q];
      wave_lds_sync();
#pragma unroll
      for (int q = 0; q < N; ++q) {  // S^T along y: lane = (node y = la, z = lb)
        double acc = sk.S[0][la] * xb[NL * lb + q];
#pragma unroll
        for (int k = 1; k < N; ++k)
          acc = __builtin_fma(sk.S[k][la], xb[N * k + NL * lb + q], acc);
        t[q] = acc;
      }
      double out[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {  // S^T along x
        double acc = sk.S[0][i] * t[0];
#pragma unroll
        for (int q = 1; q < N; ++q) acc = __builtin_fma(sk.S[q][i], t[q], acc);
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
