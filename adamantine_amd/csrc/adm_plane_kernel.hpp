// Plane cell kernel of the matrix-free heat operator for FE degrees 1 and 2
// (gfx950). Reference: source/ThermalOperator.templates.hh:479-617.
//
// Lane layout: p+1 lanes per cell, lane s owning the x-y plane z = s of the
// cell's nodes and, after the z contraction, the x-y plane of quadrature
// points z = s. The x and y contractions run on the lane's (p+1)^2 values in
// registers; only the z direction crosses lanes, through a wave-private LDS
// buffer: one exchange for the evaluation and two for the integration
// (against six in the line-per-lane kernel). A wavefront takes a patch of
// PX x PY cells of one cell layer (7 x 3 for Q2: 63 lanes; 8 x 4 for Q1:
// 64 lanes), so the nodes its cells share along x and y are summed inside
// the wavefront (cross-lane shuffles). Each patch then stores its node sums
// in its own block of a patch buffer, and a second kernel (one lane per
// node) adds the blocks of the patches that share the node, in a fixed
// order, into dst. No atomics: FP64 atomic adds to device memory are
// performed at the memory side on this chip and cost more than the whole
// computation (measured: the loads and atomic adds alone of this layout
// took 84 us per application of the 128x64x32 Q2 workload), and the result
// is the same in every run.
//
// Arithmetic: the values at the quadrature points, and so the phase state,
// equal the CPU oracle's bit for bit (S along x, y, then z with the same fma
// chains); the x and y gradients are the oracle's collocation derivatives,
// the z gradient comes from the nodal planes through DcS = Dc S. The
// integration folds each point row into the node plane at once: the x
// part through DcS^T and S^T along x, the y part through S^T along x and
// DcS^T along y, so only two planes of accumulators are live; the z part
// goes through S^T and DcS^T across the lanes. dst differs from the oracle
// by rounding only (1/(rho c_p) comes from the hardware reciprocal refined
// by two Newton steps).
//
// Wave-uniform fast path: when no cell of the wavefront holds powder and no
// point is at or above the solidus, every point is fully solid, and the
// mixtures reduce to the solid value (the reference adds the other states
// with a zero ratio, which changes no bit), without the mushy-zone division.
// Merged tables find a point's interval with one bin lookup (MergedBins).
// The heat sources run in a second pass over the cells a beam can reach.
#pragma once

#include "adm_kernels.hpp"

namespace adm {

template <int P>
struct PlaneLanes {
  static_assert(P == 1 || P == 2, "plane kernel: FE degree 1 or 2");
  static constexpr int N = P + 1;
  static constexpr int NP = N * N;  // values per lane
  static constexpr int PX = P == 1 ? 8 : 7;
  static constexpr int PY = P == 1 ? 4 : 3;
  static constexpr int cells_per_wave = PX * PY;
  // exchange slot per lane (doubles), even so that slots stay 16-byte
  // aligned for 128-bit LDS accesses
  static constexpr int slot = (NP + 1) & ~1;
  // slots of every lane's cell, including the partial cell of the lanes
  // past the last whole one (Q2: lane 63 addresses cell 21's planes)
  static constexpr int wave_lds = ((64 + N - 1) / N) * N * slot;
  // nodes of a patch: (P PX + 1) x (P PY + 1) per plane, N planes; the
  // patch buffer holds one double per node, [plane][y][x]
  static constexpr int patch_nx = P * PX + 1, patch_ny = P * PY + 1;
  static constexpr int patch_nodes = patch_nx * patch_ny * N;
};

// Doubles of the library's static scratch array (adm_vmult_noalloc), which
// holds the plane kernel's patch buffer there: 40 MiB per device, enough
// for the 128x64x32 Q2 workload's 4.2 million patch-buffer doubles.
constexpr size_t kNoAllocScratchDoubles = (size_t)5 << 20;

// doubles of one material's merged tables staged in LDS by the plane
// kernel: B, the coefficients of density .. thermal_conductivity_z, then
// the MergedBins record
constexpr int kPlaneMergedLds =
    kMergedLds + (int)(sizeof(MergedBins) / sizeof(double));
static_assert(sizeof(MergedBins) % sizeof(double) == 0, "bins layout");

// Patch coordinates of patch index r of a plane launch, npx x npy patches
// per cell layer and nz layers: x fastest, then y, then the cell layer; or
// (ymajor) x fastest, then the cell layer, then y -- the order in which the
// XCD-aware launch hands each XCD one contiguous range, so that every XCD
// takes a slab of patch rows through all the layers and as many of the
// slow layers (powder, melt) as the others.
struct PlanePatch {
  int px, py, pz;
};
__device__ __forceinline__ PlanePatch plane_patch(int npx, int npy, int nz,
                                                  bool ymajor, int r) {
  PlanePatch pp;
  pp.px = r % npx;
  const int t = r / npx;
  if (ymajor) {
    pp.pz = t % nz;
    pp.py = t / nz;
  } else {
    pp.py = t % npy;
    pp.pz = t / npy;
  }
  return pp;
}

// 1/x from the hardware reciprocal and two Newton steps (within an ulp or
// two of the correctly rounded quotient)
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// Interval j = #{B_k <= T} of a merged table by its bins (MergedBins).
__device__ __forceinline__ int merged_bin_interval(const double *B,
                                                   const MergedBins &mb,
                                                   double T) {
  const double x = (T - mb.lo) * mb.inv_w;
  int b = (int)x;  // NaN and T below lo give bin 0 after the clamp
  b = x < 0.0 ? 0 : (b > kMergedBins - 1 ? kMergedBins - 1 : b);
  const int k0 = mb.k0[b];
  int j = k0;
  const int maxc = (int)mb.maxc;
  for (int c = 0; c < maxc; ++c) {
    const int k = k0 + c < kMergedMax - 1 ? k0 + c : kMergedMax - 1;
    j += B[k] <= T ? 1 : 0;
  }
  return j;
}

// A uniform value pinned in scalar registers: loads from constant memory
// (kernel arguments, the constant tables) could otherwise be repeated inside
// the patch loop, and every scalar load there makes the next LDS wait a
// full one (scalar loads and LDS share one counter and scalar loads return
// out of order).
template <class T>
__device__ __forceinline__ T in_sgpr(T x) {
  asm volatile("" : "+s"(x));
  return x;
}

// wavefronts per SIMD the plane kernel is compiled for (VGPR budget)
#ifndef ADM_PLANE_WAVES_PER_EU
#define ADM_PLANE_WAVES_PER_EU 2
#endif
constexpr int kPlaneWavesPerEU = ADM_PLANE_WAVES_PER_EU;

template <int P, int NS, int PK>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(ADM_PLANE_WAVES_PER_EU)))
    plane_kernel(const CellArgs a, const double *__restrict__ src,
                 double *__restrict__ dst) {
  using L = PlaneLanes<P>;
  constexpr int N = L::N, NP = L::NP, N3 = N * N * N;
  constexpr int PX = L::PX, PY = L::PY, CPW = L::cells_per_wave;
  constexpr int SL = L::slot;
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
    constexpr int nn = kMaxN * kMaxN;
    if (tid < nn) {
      (&sk.S[0][0])[tid] = (&kc.S[0][0])[tid];
      (&sk.Dc[0][0])[tid] = (&kc.DcS[0][0])[tid];  // Dc S (z direction)
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
      const int total = n_mat * kPlaneMergedLds;
      for (int i = tid; i < total; i += blockDim.x) {
        const int mm = i / kPlaneMergedLds, k = i % kPlaneMergedLds;
        lds[i] = k < kMergedLds
                     ? (&kc.mat.merged[mm].B[0])[k]
                     : reinterpret_cast<const ADM_CONSTANT double *>(
                           &kc.mat.bins[mm])[k - kMergedLds];
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
  int sources_on = a.sources_on;
  if (a.beam_table) {
    sources_on = 0;
    for (int b = 0; b < a.n_beams; ++b) sources_on |= sbeams[b].on ? 1 : 0;
  }
  // uniform constants, in scalar registers for the whole patch loop:
  // in-plane shape values, collocation derivative, DcS, geometry and the
  // kernel arguments the loop reads. For Q2 the host has checked that the
  // middle Gauss point is the middle node (S row 1 = (0, 1, 0)), that S is
  // symmetric under reflection except S[2][1] and that Dc is antisymmetric
  // with Dc[1][1] = 0, bit for bit: S and Dc then take 8 scalar values
  // instead of 18, and the middle rows' products by 0 and 1 are skipped,
  // which leaves every sum as the generic chain gives it (x + 0 = x for the
  // finite temperatures here).
  double S[N][N], Dc[N][N], DS[N][N];
  if constexpr (P == 2) {
    const double s00 = in_sgpr(kc.S[0][0]), s01 = in_sgpr(kc.S[0][1]),
                 s02 = in_sgpr(kc.S[0][2]), s21 = in_sgpr(kc.S[2][1]);
    const double d00 = in_sgpr(kc.Dc[0][0]), d01 = in_sgpr(kc.Dc[0][1]),
                 d02 = in_sgpr(kc.Dc[0][2]), d10 = in_sgpr(kc.Dc[1][0]);
    S[0][0] = s00;
    S[0][1] = s01;
    S[0][2] = s02;
    S[1][0] = -0.0;
    S[1][1] = 1.0;
    S[1][2] = 0.0;
    S[2][0] = s02;
    S[2][1] = s21;
    S[2][2] = s00;
    Dc[0][0] = d00;
    Dc[0][1] = d01;
    Dc[0][2] = d02;
    Dc[1][0] = d10;
    Dc[1][1] = 0.0;
    Dc[1][2] = -d10;
    Dc[2][0] = -d02;
    Dc[2][1] = -d01;
    Dc[2][2] = -d00;
#pragma unroll
    for (int q = 0; q < N; ++q)
#pragma unroll
      for (int k = 0; k < N; ++k) DS[q][k] = in_sgpr(kc.DcS[q][k]);
  } else {
#pragma unroll
    for (int q = 0; q < N; ++q)
#pragma unroll
      for (int k = 0; k < N; ++k) {
        S[q][k] = in_sgpr(kc.S[q][k]);
        Dc[q][k] = in_sgpr(kc.Dc[q][k]);
        DS[q][k] = in_sgpr(kc.DcS[q][k]);
      }
  }
  const double ih0 = in_sgpr(kc.inv_h[0]), ih1 = in_sgpr(kc.inv_h[1]),
               ih2 = in_sgpr(kc.inv_h[2]);
  const double det = in_sgpr(kc.det);
  double wq[N];
#pragma unroll
  for (int q = 0; q < N; ++q) wq[q] = in_sgpr(kc.wq[q]);
  const int lext0 = in_sgpr(a.lext[0]), lext1 = in_sgpr(a.lext[1]);
  const int nc0 = in_sgpr(a.nc[0]), nc1 = in_sgpr(a.nc[1]);
  const int npx = in_sgpr(a.plane_npx), npy = in_sgpr(a.plane_npy);
  const int64_t sy = in_sgpr((int64_t)a.nn[0]);
  const int64_t sz = in_sgpr((int64_t)a.nn[0] * a.nn[1]);
  src = in_sgpr(src);
  meta_p = in_sgpr(meta_p);
  slot_p = in_sgpr(slot_p);
  powder_p = in_sgpr(powder_p);
  liquid_p = in_sgpr(liquid_p);
  dirty_p = in_sgpr(dirty_p);
  double *const patch_E =
      in_sgpr(a.plane_E + (int64_t)blockIdx.y * a.plane_E_stride);

  const int lane = threadIdx.x % 64;
  // the wavefront's index in the block, as a uniform (scalar) value
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int cw = lane / N, s = lane % N;  // cell of the wavefront, plane
  const int ix = cw % PX, iy = cw / PX;
  // The blocks loop over the patches (persistent grid: the prologue above
  // runs once per block). Blocks are dispatched round-robin over the 8
  // XCDs (block b on XCD b % 8). One operator: each XCD takes one
  // contiguous eighth of the patches in the y-major order, a slab of patch
  // rows through every cell layer, so patches that share node lines share
  // that XCD's L2 and every XCD holds its share of the slow layers (an
  // eighth of the layers in the x-y-z order gave the top one, the powder,
  // to one XCD: 0.120 ms per application of the 128x64x32 Q2 workload,
  // against 0.100 in the plain order and 0.098 in y-major slabs). Batched
  // members (grid y): the plain order, block after block through the
  // patches (64 members of 64x32x16: 3.10 ms per RK4 step against 3.20 in
  // y-major slabs and 3.75 in layer slabs).
  const int W = blockDim.x >> 6;
  int first, stride, end;
  const bool ymajor = a.xcd_remap && gridDim.x >= 8 && gridDim.y == 1;
  const int nzl = in_sgpr(a.n_rows / (npx * npy));
  if (ymajor) {
    const int x = blockIdx.x & 7, k = blockIdx.x >> 3;
    const int kx = (gridDim.x - x + 7) >> 3;  // blocks on XCD x
    first = (int)((int64_t)a.n_rows * x / 8) + k * W + wave;
    end = (int)((int64_t)a.n_rows * (x + 1) / 8);
    stride = kx * W;
  } else {
    first = blockIdx.x * W + wave;
    end = a.n_rows;
    stride = gridDim.x * W;
  }
  first = in_sgpr(first);
  end = in_sgpr(end);
  stride = in_sgpr(stride);
  // a patch's addresses and loads (issued together at the top of the
  // patch: carrying the next patch's loads through the computation was
  // measured slower, at this register pressure)
  struct PatchLoad {
    int cx, cy, cz;
    bool in_mesh;
    int64_t c, base;
    uint32_t meta;
    int32_t pslot;
    uint8_t dirty;
    double u[NP];
  };
  auto load_patch = [&](int r, PatchLoad &pl) {
    const PlanePatch pp =
        plane_patch(npx, npy, nzl, ymajor, r < end ? r : first);
    pl.cx = pp.px * PX + ix;
    pl.cy = pp.py * PY + iy;
    pl.cz = pp.pz;
    pl.in_mesh = r < end && cw < CPW && pl.cx < lext0 && pl.cy < lext1;
    pl.c = pl.in_mesh ? pl.cx + (int64_t)nc0 * (pl.cy + (int64_t)nc1 * pl.cz)
                      : 0;
    // first node of the lane's plane (node 0 for lanes outside the mesh)
    pl.base = pl.in_mesh ? (int64_t)pl.cx * P + sy * ((int64_t)pl.cy * P) +
                               sz * ((int64_t)pl.cz * P + s)
                         : 0;
    pl.meta = meta_p[pl.c];
    pl.pslot = slot_p[pl.c];
    pl.dirty = (NS > 1 && liquid_p) ? dirty_p[pl.c] : 0;
#pragma unroll
    for (int yy = 0; yy < N; ++yy)
#pragma unroll
      for (int xx = 0; xx < N; ++xx)
        pl.u[xx + N * yy] = src[pl.in_mesh ? pl.base + xx + sy * yy : 0];
  };
  if (first >= end) return;  // whole wavefronts
  for (int r = first; r < end; r += stride) {
  wave_lds_sync();  // the previous patch's exchange reads are done
  PatchLoad cur;
  load_patch(r, cur);
  const int cx = cur.cx, cy = cur.cy, cz = cur.cz;
  const bool in_mesh = cur.in_mesh;
  const int64_t c = cur.c;
  const uint32_t meta0 = cur.meta;
  const int32_t pslot32 = cur.pslot;
  double u[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) u[q] = cur.u[q];

  const uint32_t meta = in_mesh ? meta0 : 0u;
  const bool act = (meta & kActive) != 0;
  double *xb = lds + (PK == 2 ? n_mat * kPlaneMergedLds : 0) +
               wave * L::wave_lds;
  double *myslot = xb + lane * SL;
  // slot of plane n of this lane's cell
  auto slot_of = [&](int n) { return xb + (cw * N + n) * SL; };

  const int mid = (meta >> kMaterialShift) & 3;
  const double solidus = sk.solidus[mid], liquidus = sk.liquidus[mid];
  const double latent = sk.latent[mid];
  const bool iso = sk.iso[mid] != 0.0;
  const bool has_powder = act && (meta & kHasPowder) != 0;
  ADM_GLOBAL double *pw =
      has_powder ? powder_p + (int64_t)pslot32 * N3 + NP * s : nullptr;
  bool dirty = false;
  ADM_GLOBAL double *lq = nullptr;
  if constexpr (NS > 1) {
    if (liquid_p && act) {
      dirty = cur.dirty != 0;
      lq = liquid_p + c * N3 + NP * s;
    }
  }
  const bool rotated = act && (meta & kRotated) != 0;

  // ---- values: S along x, then y, in registers ---------------------------
  double v[NP];
  {
    double t[NP];
#pragma unroll
    for (int yy = 0; yy < N; ++yy)
#pragma unroll
      for (int qx = 0; qx < N; ++qx) {
        if (P == 2 && qx == 1) {
          t[qx + N * yy] = u[1 + N * yy];
          continue;
        }
        double acc = S[qx][0] * u[N * yy];
#pragma unroll
        for (int k = 1; k < N; ++k)
          acc = __builtin_fma(S[qx][k], u[k + N * yy], acc);
        t[qx + N * yy] = acc;
      }
#pragma unroll
    for (int qy = 0; qy < N; ++qy)
#pragma unroll
      for (int qx = 0; qx < N; ++qx) {
        if (P == 2 && qy == 1) {
          v[qx + N * qy] = t[qx + N];
          continue;
        }
        double acc = S[qy][0] * t[qx];
#pragma unroll
        for (int k = 1; k < N; ++k)
          acc = __builtin_fma(S[qy][k], t[qx + N * k], acc);
        v[qx + N * qy] = acc;
      }
  }
  // ---- exchange 1: the cell's N planes (contracted in x and y) -----------
#pragma unroll
  for (int q = 0; q < NP; ++q) myslot[q] = v[q];
  wave_lds_sync();
  // z contraction: the value at the lane's point plane (the oracle's chain)
  // and the z derivative from the nodal planes (Dc S)
  double Tq[NP], gz[NP];
  {
    double szv[N], dzv[N];
#pragma unroll
    for (int n = 0; n < N; ++n) {
      szv[n] = sk.S[s][n];
      dzv[n] = sk.Dc[s][n];  // (Dc S)[s][n]
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const double w0 = slot_of(0)[q];
      Tq[q] = szv[0] * w0;
      gz[q] = dzv[0] * w0;
    }
#pragma unroll
    for (int n = 1; n < N; ++n)
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const double wn = slot_of(n)[q];
        Tq[q] = __builtin_fma(szv[n], wn, Tq[q]);
        gz[q] = __builtin_fma(dzv[n], wn, gz[q]);
      }
  }

  // ---- per-point physics, one point row (fixed y) at a time --------------
  bool near = false;
  if (sources_on && act) {
    const int cc[3] = {cx, cy, cz};
    double lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = kc.origin[d] + cc[d] * kc.h[d];
      hi[d] = kc.origin[d] + (cc[d] + 1) * kc.h[d];
    }
#pragma unroll 1
    for (int b = 0; b < a.n_beams; ++b)
      near |= beam_may_touch(sbeams[b], lo, hi);
  }
  const double *mB = lds + mid * kPlaneMergedLds;
  const double *mC = mB + kMergedMax;
  const MergedBins &mbin =
      *reinterpret_cast<const MergedBins *>(mB + kMergedLds);
  // wave-uniform fast path: one material, every point fully solid (no
  // powder, below the solidus)
  const int mid0 = __builtin_amdgcn_readfirstlane(mid);
  bool warm = false;
  if constexpr (NS > 1) {
#pragma unroll
    for (int q = 0; q < NP; ++q) warm |= act && !(Tq[q] < solidus);
  }
  const bool fast =
      !__any(warm || has_powder || (act && mid != mid0));
  // accumulators of the node plane (before the z part): Ap gathers the x
  // and y flux terms and the source, Bp the z flux
  double Ap[NP], Bp[NP];
  bool plane_liquid = false;
  double cs = 1.0, sn = 0.0;
  if (rotated) {
    cs = cos_p[c];
    sn = sin_p[c];
  }
  // JxW of the lane's points as the host tabulates it:
  // det * (((1 * w_x) * w_y) * w_z), w_z = w[s]
  const double wz = sk.wq[s];
  const double ihx2 = ih0 * ih0, ihy2 = ih1 * ih1, ihz2 = ih2 * ih2;
  // SOLID: the fast path (material mid_ for every lane, ratios (1, 0, 0));
  // otherwise the lane's own material and the phase-state update
  auto physics = [&](auto solid_tag, int mid_, const double *mB_,
                     const double *mC_, const MergedBins &mbin_, bool iso_,
                     bool rot_) {
    constexpr bool SOLID = decltype(solid_tag)::value;
    if constexpr (SOLID && NS > 1) {
      // every point solid: a liquid ratio held from the last evaluation
      // goes back to zero
      if (lq && dirty) {
#pragma unroll
        for (int q = 0; q < NP; ++q) lq[q] = 0.0;
      }
    }
    // the stored powder of the lane's points, loaded together (one memory
    // round trip for the plane)
    double pin[SOLID || NS < 3 ? 1 : NP];
    if constexpr (!SOLID && NS == 3) {
#pragma unroll
      for (int q = 0; q < NP; ++q) pin[q] = 0.0;
      if (pw) {
#pragma unroll
        for (int q = 0; q < NP; ++q) pin[q] = pw[q];
      }
    }
#pragma unroll
    for (int qy = 0; qy < N; ++qy) {
      double w0[N], w1[N], w2[N];
#pragma unroll
      for (int qx = 0; qx < N; ++qx) {
        const int q = qx + N * qy;
        const double T = Tq[q];
        // x and y gradients by the collocation derivative (the oracle's
        // chains)
        double gxq = Dc[qx][0] * Tq[N * qy];
        double gyq = Dc[qy][0] * Tq[qx];
#pragma unroll
        for (int k = 1; k < N; ++k) {
          if (!(P == 2 && qx == 1 && k == 1))
            gxq = __builtin_fma(Dc[qx][k], Tq[k + N * qy], gxq);
          if (!(P == 2 && qy == 1 && k == 1))
            gyq = __builtin_fma(Dc[qy][k], Tq[qx + N * k], gyq);
        }
        double r_[NS];
        double l = 0.0;
        if constexpr (SOLID || NS == 1) {
          r_[0] = 1.0;
#pragma unroll
          for (int k = 1; k < NS; ++k) r_[k] = 0.0;
        } else {
          // update_state_ratios (source/ThermalOperator.templates.hh:227-317)
          double p0 = 0.0;
          if constexpr (NS == 3) p0 = pin[q];
          double pv = p0;
          const double mushy = (T - solidus) / (liquidus - solidus);
          l = T < solidus ? 0.0 : (T > liquidus ? 1.0 : mushy);
          if constexpr (NS == 2) {
            r_[1] = l;
            r_[0] = 1.0 - l;
          } else {
            const double am = 1.0 - l;
            pv = (pv < am) ? pv : am;
            r_[1] = l;
            r_[2] = pv;
            r_[0] = 1.0 - l - pv;
          }
          if (pw && pv != p0) pw[q] = pv;
          if (lq && (l != 0.0 || dirty)) lq[q] = l;
          plane_liquid |= l != 0.0;
        }
        double irc, kx, ky, kzz;
        auto props = [&](const auto &ev) {
          // get_inv_rho_cp (source/ThermalOperator.templates.hh:415-477)
          const double density = ev(ADM_DENSITY, r_);
          double specific_heat = ev(ADM_SPECIFIC_HEAT, r_);
          if constexpr (NS > 1 && !SOLID) {
            const double is_mushy = r_[1] * r_[0];
            if (is_mushy > 0.0) specific_heat += latent;
          }
          irc = recip(density * specific_heat);
          kx = ev(ADM_THERMAL_CONDUCTIVITY_X, r_);
          if (iso_) {
            ky = kx;
            kzz = kx;
          } else {
            ky = ev(ADM_THERMAL_CONDUCTIVITY_Y, r_);
            kzz = ev(ADM_THERMAL_CONDUCTIVITY_Z, r_);
          }
        };
        if constexpr (PK == 2) {
          const int j = merged_bin_interval(mB_, mbin_, T);
          if constexpr (SOLID) {
            // the solid value alone: (0 + 1 X_s) + 0 X_l + 0 X_p = X_s
            const double d = T - mB_[j > 0 ? j - 1 : 0];
            const double *cj = mC_ + j * (ADM_MAX_STATES * 2);
            auto solid_ev = [&](int prop, const double(&)[NS]) {
              const double *cf = cj + prop * (kMergedMax * ADM_MAX_STATES * 2);
              return __builtin_fma(d, cf[1], cf[0]);
            };
            props(solid_ev);
          } else {
            props(MergedEval(mB_, mC_, T, j));
          }
        } else {
          props(typename EvalOf<PK>::type(m, mid_, T));
        }
        // -(1/rho c_p) K grad T * JxW, each component with its 1/h
        const double jxw = det * ((wq[qx] * wq[qy]) * wz);
        const double cq = (-irc) * jxw;
        if (rot_) {
          const double gx_ = gxq * ih0, gy_ = gyq * ih1;
          const double fx = (kx * cs * cs + ky * sn * sn) * gx_ +
                            ((kx - ky) * sn * cs) * gy_;
          const double fy = ((kx - ky) * sn * cs) * gx_ +
                            (kx * sn * sn + ky * cs * cs) * gy_;
          w0[qx] = (cq * fx) * ih0;
          w1[qx] = (cq * fy) * ih1;
        } else {
          w0[qx] = (cq * kx) * (gxq * ihx2);
          w1[qx] = (cq * ky) * (gyq * ihy2);
        }
        w2[qx] = (cq * kzz) * (gz[q] * ihz2);
      }
      // fold the row into the node-plane accumulators:
      //   Ap[xx, yy] += S[qy][yy] (DcS^T w0)[xx] + DcS[qy][yy] (S^T w1)[xx]
      //   Bp[xx, yy] += S[qy][yy] (S^T w2)[xx]
#pragma unroll
      for (int xx = 0; xx < N; ++xx) {
        double t0 = DS[0][xx] * w0[0];
        double t1 = S[0][xx] * w1[0];
        double t2 = S[0][xx] * w2[0];
#pragma unroll
        for (int q = 1; q < N; ++q) {
          t0 = __builtin_fma(DS[q][xx], w0[q], t0);
          if (P == 2 && q == 1) {
            if (xx == 1) {
              t1 = t1 + w1[q];
              t2 = t2 + w2[q];
            }
            continue;
          }
          t1 = __builtin_fma(S[q][xx], w1[q], t1);
          t2 = __builtin_fma(S[q][xx], w2[q], t2);
        }
#pragma unroll
        for (int yy = 0; yy < N; ++yy) {
          const int o = xx + N * yy;
          if (qy == 0) {
            Ap[o] = __builtin_fma(DS[qy][yy], t1, S[qy][yy] * t0);
            Bp[o] = S[qy][yy] * t2;
          } else if (P == 2 && qy == 1) {
            if (yy == 1) Ap[o] = Ap[o] + t0;
            Ap[o] = __builtin_fma(DS[qy][yy], t1, Ap[o]);
            if (yy == 1) Bp[o] = Bp[o] + t2;
          } else {
            Ap[o] = __builtin_fma(S[qy][yy], t0, Ap[o]);
            Ap[o] = __builtin_fma(DS[qy][yy], t1, Ap[o]);
            Bp[o] = __builtin_fma(S[qy][yy], t2, Bp[o]);
          }
        }
      }
    }
  };
  if (fast) {
    // the wave's material (uniform: scalar registers, uniform LDS rows)
    const double *fB = lds + mid0 * kPlaneMergedLds;
    physics(std::integral_constant<bool, true>{}, mid0, fB, fB + kMergedMax,
            *reinterpret_cast<const MergedBins *>(fB + kMergedLds),
            __builtin_amdgcn_readfirstlane(sk.iso[mid0] != 0.0 ? 1 : 0) != 0,
            __any(rotated));
  } else {
    physics(std::integral_constant<bool, false>{}, mid, mB, mC, mbin, iso,
            rotated);
  }
  // heat sources, in a pass of their own over the cells that a beam can
  // reach (few), point by point in a loop (one copy of the source code):
  // 1/(rho c_p) of each point again from its updated state, the weighted
  // source values through the lane's exchange slot, then
  // Ap[xx, yy] += S[qy][yy] (S^T val)[xx]
  if (__any(near)) {
    wave_lds_sync();
    if (near) {
#pragma unroll
      for (int q = 0; q < NP; ++q) myslot[q] = Tq[q];
#pragma unroll 1
      for (int q = 0; q < NP; ++q) {
        const int qx = q % N, qy = q / N;
        const double T = myslot[q];
        double r_[NS];
        if constexpr (NS == 1) {
          r_[0] = 1.0;
        } else {
          const double mushy = (T - solidus) / (liquidus - solidus);
          const double l = T < solidus ? 0.0 : (T > liquidus ? 1.0 : mushy);
          if constexpr (NS == 2) {
            r_[1] = l;
            r_[0] = 1.0 - l;
          } else {
            // the powder as this evaluation left it
            const double pv = pw ? pw[q] : 0.0;
            r_[1] = l;
            r_[2] = pv;
            r_[0] = 1.0 - l - pv;
          }
        }
        double irc = 0.0;
        auto rho_cp = [&](const auto &ev) {
          const double density = ev(ADM_DENSITY, r_);
          double specific_heat = ev(ADM_SPECIFIC_HEAT, r_);
          if constexpr (NS > 1) {
            const double is_mushy = r_[1] * r_[0];
            if (is_mushy > 0.0) specific_heat += latent;
          }
          irc = recip(density * specific_heat);
        };
        if constexpr (PK == 2)
          rho_cp(MergedEval(mB, mC, T, merged_bin_interval(mB, mbin, T)));
        else
          rho_cp(typename EvalOf<PK>::type(m, mid, T));
        const double px = kc.origin[0] + (cx + sk.xq[qx]) * kc.h[0];
        const double py = kc.origin[1] + (cy + sk.xq[qy]) * kc.h[1];
        const double pz = kc.origin[2] + (cz + sk.xq[s]) * kc.h[2];
        double qs = 0.0;
        for (int b = 0; b < a.n_beams; ++b)
          qs += beam_value(sbeams[b], px, py, pz, kc.log_01);
        const double jxw = det * ((sk.wq[qx] * sk.wq[qy]) * wz);
        myslot[q] = (qs * irc) * jxw;
      }
#pragma unroll
      for (int qy = 0; qy < N; ++qy)
#pragma unroll
        for (int xx = 0; xx < N; ++xx) {
          double tv = S[0][xx] * myslot[N * qy];
#pragma unroll
          for (int q = 1; q < N; ++q)
            tv = __builtin_fma(S[q][xx], myslot[q + N * qy], tv);
#pragma unroll
          for (int yy = 0; yy < N; ++yy)
            Ap[xx + N * yy] = __builtin_fma(S[qy][yy], tv, Ap[xx + N * yy]);
        }
    }
  }
  if constexpr (NS > 1) {
    // cell-wide "held liquid" flag over the cell's N lanes
    const uint64_t ball = __ballot(plane_liquid);
    const bool any_liquid = ((ball >> (cw * N)) & ((1ull << N) - 1)) != 0;
    if (lq && s == 0 && any_liquid != dirty) dirty_p[c] = any_liquid ? 1 : 0;
  }

  // ---- exchanges 2 and 3: z part, node plane s = sum over point planes --
  // out = sum_qz S[qz][s] Ap_qz + sum_qz (Dc S)[qz][s] Bp_qz
  double out[NP];
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < NP; ++q) myslot[q] = Ap[q];
  wave_lds_sync();
  {
    double sc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) sc[k] = sk.S[k][s];
#pragma unroll
    for (int q = 0; q < NP; ++q) out[q] = sc[0] * slot_of(0)[q];
#pragma unroll
    for (int k = 1; k < N; ++k)
#pragma unroll
      for (int q = 0; q < NP; ++q)
        out[q] = __builtin_fma(sc[k], slot_of(k)[q], out[q]);
  }
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < NP; ++q) myslot[q] = Bp[q];
  wave_lds_sync();
  {
    double dc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) dc[k] = sk.Dc[k][s];  // (Dc S)[k][s]
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
      for (int q = 0; q < NP; ++q)
        out[q] = __builtin_fma(dc[k], slot_of(k)[q], out[q]);
  }
  if (!act) {
#pragma unroll
    for (int q = 0; q < NP; ++q) out[q] = 0.0;
  }

  // ---- shared nodes of the patch, then dst --------------------------------
  // x: node column P of cell (ix, iy) += column 0 of cell (ix + 1, iy)
#pragma unroll
  for (int yy = 0; yy < N; ++yy) {
    const double o = __shfl_down(out[N * yy], N);
    if (ix + 1 < PX) out[P + N * yy] += o;
  }
  // y: node row P += row 0 of cell (ix, iy + 1) (its column P already
  // holds the corner sums)
#pragma unroll
  for (int xx = 0; xx < N; ++xx) {
    const double o = __shfl_down(out[xx], N * PX);
    if (iy + 1 < PY) out[xx + N * P] += o;
  }
  if (in_mesh) {
  // the patch's node sums, every node of the patch once, into its slots
  // of the patch buffer (plain stores; the gather kernel adds the slots of
  // the patches that share a node in a fixed order)
  // patch buffer layout [cell layer][plane][patch y][node y][patch x]
  // [node x]: a node row of one plane runs through the patches of a row
  // contiguously, which the gather pass reads along x
  const PlanePatch pp = plane_patch(npx, npy, nzl, ymajor, r);
  const int64_t rowlen = (int64_t)npx * L::patch_nx;
  double *E = patch_E +
              ((((int64_t)pp.pz * N + s) * npy + pp.py) * L::patch_ny +
               iy * P) * rowlen +
              (int64_t)pp.px * L::patch_nx + ix * P;
#pragma unroll
  for (int yy = 0; yy < N; ++yy)
#pragma unroll
    for (int xx = 0; xx < N; ++xx) {
      if (xx == 0 && ix != 0) continue;  // summed into the left neighbour
      if (yy == 0 && iy != 0) continue;
      E[xx + rowlen * yy] = out[xx + N * yy];
    }
  }
  }  // patches
}

// Second pass of the plane kernel: one lane per node (x fastest; the
// grid's y index is the ensemble member) sums the patch-buffer slots of the
// patches that hold the node -- up to two along each axis -- in increasing
// patch order, and writes dst (or adds to it). Nodes above the launched
// cell layers get no slot: 0, or dst unchanged with add. n_nodes < 2^31.
template <int P>
__global__ void __launch_bounds__(256)
    plane_gather_kernel(const double *__restrict__ E, int64_t E_stride,
                        double *__restrict__ dst, int64_t member_stride,
                        int32_t nn0, int32_t nn1, int32_t nn2, int32_t npx,
                        int32_t npy, int32_t nz, int32_t add,
                        const FaceArgs fa, const StageArgs st) {
  using L = PlaneLanes<P>;
  constexpr int WX = P * L::PX, WY = P * L::PY;
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn0 * nn1 * nn2) return;
  E += (int64_t)blockIdx.y * E_stride;
  dst += (int64_t)blockIdx.y * member_stride;
  const int row = node / nn0, X = node - row * nn0;
  const int Z = row / nn1, Y = row - Z * nn1;
  // above the top node plane of the launch's cell layers no cell or face
  // adds a term: the result is 0 (or dst as it is), and a stage update
  // leaves every vector as it is
  if (Z > P * nz) {
    if (st.kind == kStageNone && !add) dst[node] = 0.0;
    return;
  }
  const int64_t rowlen = (int64_t)npx * L::patch_nx;
  // offsets of the blocks holding the node (unused entries stay 0: in
  // bounds of E)
  int yoff[2] = {0, 0}, zoff[2] = {0, 0}, xoff[2] = {0, 0};
  int nyc = 0, nzc = 0, nxc = 0;
  const int py = Y / WY, ly = Y - py * WY;
  if (ly == 0 && py > 0) yoff[nyc++] = (py - 1) * L::patch_ny + WY;
  if (py < npy) yoff[nyc++] = py * L::patch_ny + ly;
  const int pz = Z / P, lz = Z - pz * P;
  if (lz == 0 && pz > 0 && pz - 1 < nz) zoff[nzc++] = (pz - 1) * (P + 1) + P;
  if (pz < nz) zoff[nzc++] = pz * (P + 1) + lz;
  const int px = X / WX, lx = X - px * WX;
  if (lx == 0 && px > 0) xoff[nxc++] = (px - 1) * L::patch_nx + WX;
  if (px < npx) xoff[nxc++] = px * L::patch_nx + lx;
  // up to 2 x 2 x 2 blocks: every load issued before the sums, which add
  // the blocks present in (z, y, x) order
  double v[8];
#pragma unroll
  for (int kz = 0; kz < 2; ++kz)
#pragma unroll
    for (int ky = 0; ky < 2; ++ky) {
      const double *r =
          E + ((int64_t)zoff[kz < nzc ? kz : 0] * npy * L::patch_ny +
               yoff[ky < nyc ? ky : 0]) *
                  rowlen;
#pragma unroll
      for (int kx = 0; kx < 2; ++kx) {
        const bool here = kz < nzc && ky < nyc && kx < nxc;
        v[4 * kz + 2 * ky + kx] = here ? r[xoff[kx < nxc ? kx : 0]] : 0.0;
      }
    }
  double sum = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (((k >> 2) < nzc) && (((k >> 1) & 1) < nyc) && ((k & 1) < nxc))
      sum += v[k];
  double r = add ? dst[node] + sum : sum;
  // the face term's slots at this node, summed in increasing face order
  // and added after the cell term (as a separate face gather pass would)
  const MemberArgs *mb = fa.members ? fa.members + blockIdx.y : nullptr;
  const int32_t *rank_p = mb ? mb->fn_rank : fa.fn_rank;
  if (rank_p && (mb ? mb->n_faces : fa.n_faces) > 0) {
    const int32_t k = rank_p[node];
    if (k >= 0) {
      const double *fo = mb ? mb->face_out : fa.face_out;
      const int32_t *off = mb ? mb->fn_offset : fa.fn_offset;
      const int32_t *slot = mb ? mb->fn_slot : fa.fn_slot;
      double fs = 0.0;
      for (int32_t e = off[k]; e < off[k + 1]; ++e) fs += fo[slot[e]];
      r += fs;
    }
  }
  // the result, or (a time step on the plane path) the stage update it
  // feeds, without storing it
  if (st.kind == kStageNone)
    dst[node] = r;
  else
    stage_apply(st, (int64_t)blockIdx.y * member_stride + node, r);
}

}  // namespace adm
