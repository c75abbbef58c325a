// Device-side data layout and physics of the matrix-free heat operator.
//
// Everything here is read by the HIP kernels; the host builds it once per
// context. Physics follows the host ThermalOperator of adamantine with one
// SIMD lane per quadrature point (reference paths relative to its tree):
//   state update       source/ThermalOperator.templates.hh:227-317
//   1/(rho c_p)        source/ThermalOperator.templates.hh:415-477
//   properties         source/MaterialProperty.hh:529-800,
//                      source/MaterialProperty.templates.hh:1158-1212
//   heat sources       source/GoldakHeatSource.cc:91-148,
//                      source/ElectronBeamHeatSource.cc:95-158,
//                      source/CubeHeatSource.cc:73-107
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/adm.h"

namespace adm {

constexpr int kMaxN = 6;          // fe_degree + 1 <= 6 on the device path
constexpr int kMaxQ = 216;
constexpr int kMaxFQ = 36;
constexpr double kStefanBoltzmann = 5.670374419e-8;  // source/types.hh:158

// cell meta word bits
constexpr uint32_t kActive = 1u;
constexpr uint32_t kMaterialShift = 1;   // 2 bits
constexpr uint32_t kFaceShift = 3;       // 6 bits: faces with a boundary term
constexpr uint32_t kHasPowder = 1u << 9;
constexpr uint32_t kRotated = 1u << 10;
// set by the device activation on the cells it activated, until the next one
constexpr uint32_t kNewlyActive = 1u << 11;

// per-(material, property, state) table kind
constexpr uint8_t kTableGeneral = 0;
constexpr uint8_t kTableConstant = 1;

// Merged piecewise-linear form of all state-property tables of a material:
// the union of their breakpoints B (sorted, +inf padded) splits the
// temperature axis into intervals j = #{B_k <= T}; inside interval j every
// table is linear and stored as (value at the interval's left breakpoint,
// slope), so one search per quadrature point serves every property and state.
constexpr int kMergedMax = 32;
struct MergedTable {
  double B[kMergedMax];
  double coef[ADM_N_STATE_PROPS][kMergedMax][ADM_MAX_STATES][2];
};
// Uniform bins over [B_0, B_last] that turn the interval search into one
// lookup: bin b = clamp(floor((T - lo) * inv_w)) holds k0[b], the number of
// breakpoints below the bin (widened by a margin far above the rounding of
// the bin computation), and at most maxc breakpoints fall in a widened bin,
// so j = #{B_k <= T} = k0[b] + #{c < maxc : B[k0[b] + c] <= T} exactly.
constexpr int kMergedBins = 128;
struct MergedBins {
  double lo, inv_w, maxc, pad_;
  uint8_t k0[kMergedBins];
};

struct MatConst {
  int32_t n_materials, n_states, use_table, p_order, single_layout, pad_;
  double table[ADM_MAX_MATERIALS][ADM_N_STATE_PROPS][ADM_MAX_STATES]
              [ADM_TABLE_SIZE][2];
  double poly[ADM_MAX_MATERIALS][ADM_N_STATE_PROPS][ADM_MAX_STATES]
             [ADM_MAX_POLY];
  double prop[ADM_MAX_MATERIALS][ADM_N_PROPS];
  // latent_heat / (liquidus - solidus), precomputed with the same division
  double latent_factor[ADM_MAX_MATERIALS];
  uint8_t kind[ADM_MAX_MATERIALS][ADM_N_STATE_PROPS][ADM_MAX_STATES];
  // thermal_conductivity_x/y/z identical for this material
  uint8_t isotropic[ADM_MAX_MATERIALS];
  // 1 if every table of the material has the regular form (strictly
  // increasing temperatures, then repeats of the last pair) and the merged
  // breakpoints fit kMergedMax - 1
  uint8_t merged_ok;
  MergedTable merged[ADM_MAX_MATERIALS];
  MergedBins bins[ADM_MAX_MATERIALS];
};

struct KConst {
  double S[kMaxN][kMaxN];    // shape values  S[q][i] = l_i(x_q)
  double Dc[kMaxN][kMaxN];   // collocation derivative at the quadrature points
  double DcS[kMaxN][kMaxN];  // Dc * S: derivative at the points from nodes
  double xq[kMaxN];
  double wq[kMaxN];          // 1-D quadrature weights
  double det;                // cell Jacobian determinant
  double jxw[kMaxQ];         // det * ((1 * wq[x]) * wq[y]) * wq[z]
  double fjxw[3][kMaxFQ];    // face JxW per normal direction
  double h[3], inv_h[3], origin[3];
  int32_t n_boundary_ids;
  int32_t boundary_type[ADM_MAX_BOUNDARY_IDS];
  int32_t adiabatic_only;
  double log_01;             // std::log(0.1) on the host
  MatConst mat;
};

// One boundary face carrying a convective/radiative term.
struct FaceRef {
  int32_t cell;
  int32_t info;   // face (3 bits) | boundary type << 3 | material << 8
};

// Per-member data of an ensemble (device array, indexed by blockIdx.y). The
// pointers replace the CellArgs / FaceArgs ones of the same name; beam_gain
// is absorption_efficiency * max_power of each beam.
struct MemberArgs {
  const uint32_t *meta;
  const int32_t *powder_slot;
  double *powder;
  double *face_powder;
  double *liquid;
  uint8_t *liquid_dirty;
  const double *dep_cos;
  const double *dep_sin;
  const KConst *kc;
  const FaceRef *faces;
  int64_t n_faces;
  double beam_gain[ADM_MAX_BEAMS];
  // the member's face-node lists (FaceArgs)
  double *face_out;
  const int64_t *fn_node;
  const int32_t *fn_offset;
  const int32_t *fn_slot;
  int64_t n_fnodes;
  const int32_t *fn_rank;
};

// Arguments of the cell kernel (by value).
struct CellArgs {
  const uint32_t *meta;       // [n_cells]
  const int32_t *powder_slot; // [n_cells], -1 without powder storage
  double *powder;             // [n_powder_cells][n_q]
  double *face_powder;        // [n_powder_cells][6][n_face_q]
  double *liquid;             // [n_cells][n_q] (tracked sparsely)
  uint8_t *liquid_dirty;      // [n_cells]
  const double *dep_cos;      // [n_cells]
  const double *dep_sin;
  const KConst *kc;
  int32_t nc[3];
  int32_t nn[3];
  int32_t n_beams;
  int32_t sources_on;
  int32_t add;          // unused by the atomic-accumulation kernel
  int64_t n_cells;
  adm_beam_state beams[ADM_MAX_BEAMS];
  // ensemble launch: members[blockIdx.y] replaces the per-cell pointers and
  // kc, vectors are offset by blockIdx.y * member_stride, and each beam's
  // alpha is rebuilt from the member's gain with the shared power modifier
  // and denominator of GoldakHeatSource / ElectronBeamHeatSource::update_time
  const MemberArgs *members;  // NULL: one operator
  int64_t member_stride;
  double beam_pmod[ADM_MAX_BEAMS];
  double beam_den[ADM_MAX_BEAMS];
  // color >= 0: the launch covers the cells whose coordinate parities are
  // (color & 1, color >> 1 & 1, color >> 2 & 1), lext[d] of them per axis.
  // Such cells share no node, so dst is updated by plain read-add-write;
  // color < 0: every cell, atomic accumulation.
  int32_t color;
  // cells of the launch per axis (every cell, or one color's), row
  // segments per x-row and row segments of the launch
  int32_t lext[3];
  int32_t nseg;
  int32_t n_rows;
  // captured step loops: the beams come from row *beam_slot of beam_table
  // (ADM_MAX_BEAMS states per row, beam_table_len rows) instead of `beams`,
  // and an ensemble's shared power modifiers and denominators from the
  // same row of beam_pd_table (ADM_MAX_BEAMS of each per row) instead of
  // beam_pmod / beam_den
  const adm_beam_state *beam_table;
  const int32_t *beam_slot;
  int32_t beam_table_len;
  const double *beam_pd_table;
  // nonzero: give each XCD a contiguous run of row segments (xcd_block)
  int32_t xcd_remap;
  // plane kernel: patches per x row and per y column of a cell layer (a
  // launch has n_rows = plane_npx * plane_npy * lext[2] patches), and the
  // patch buffer the kernel writes its nodes' sums to (plane_slots doubles
  // per patch, plane_E_stride per ensemble member) for the gather kernel
  int32_t plane_npx, plane_npy;
  double *plane_E;
  int64_t plane_E_stride;
};

struct FaceArgs {
  const FaceRef *faces;
  double *face_powder;  // [n_faces][n_face_q] (SolidLiquidPowder)
  const KConst *kc;
  int32_t nc[3];
  int32_t nn[3];
  int64_t n_faces;      // ensemble: the largest member face count
  const MemberArgs *members;  // NULL: one operator
  int64_t member_stride;
  // The face kernel writes node j of face f to face_out[f * (p+1)^2 + j];
  // the face gather pass then adds, for boundary node k (fn_node[k]), the
  // slots fn_slot[fn_offset[k] .. fn_offset[k+1]) (increasing face order)
  // into dst. Every run adds the same terms in the same order.
  double *face_out;
  const int64_t *fn_node;
  const int32_t *fn_offset;
  const int32_t *fn_slot;
  int64_t n_fnodes;     // ensemble: the largest member count
  // per box node, its index in fn_node or -1: the plane kernel's gather
  // pass adds a node's face slots with its patch sums
  const int32_t *fn_rank;
};

// ---------------------------------------------------------------------------
// material properties
// ---------------------------------------------------------------------------

// MaterialProperty::compute_property_from_table. Tables are sorted (checked
// at context creation), so the index of the first breakpoint above T is the
// number of breakpoints not above T.
__device__ __forceinline__ double table_value(const double (*t)[2], double T) {
  if (T <= t[0][0]) return t[0][1];
  int i = 0;
#pragma unroll
  for (int k = 0; k < ADM_TABLE_SIZE; ++k) i += !(T < t[k][0]) ? 1 : 0;
  if (i >= ADM_TABLE_SIZE - 1) return t[ADM_TABLE_SIZE - 1][1];
  const double t0 = t[i - 1][0], t1 = t[i][0];
  const double v0 = t[i - 1][1], v1 = t[i][1];
  return v0 + (T - t0) * (v1 - v0) / (t1 - t0);
}

// compute_material_property (vectorized host form). States with a zero ratio
// add r*X = +-0 in the reference, so skipping them changes no bit.
template <int NS, bool TABLE>
__device__ __forceinline__ double property(const MatConst &m, int prop, int mid,
                                           const double (&r)[NS], double T,
                                           const double (&tp)[ADM_MAX_POLY]) {
  double value = 0.0;
  if constexpr (TABLE) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (r[s] == 0.0) continue;
      double pv;
      if (m.kind[mid][prop][s] == kTableConstant)
        pv = m.table[mid][prop][s][0][1];
      else
        pv = table_value(m.table[mid][prop][s], T);
      value += r[s] * pv;
    }
  } else {
    if (m.single_layout) {
#pragma unroll
      for (int i = 0; i < ADM_MAX_POLY; ++i) {
        if (i > m.p_order) break;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          if (r[s] == 0.0) continue;
          value += r[s] * m.poly[mid][prop][s][i] * tp[i];
        }
      }
    } else {
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (r[s] == 0.0) continue;
#pragma unroll
        for (int i = 0; i < ADM_MAX_POLY; ++i) {
          if (i > m.p_order) break;
          value += r[s] * m.poly[mid][prop][s][i] * tp[i];
        }
      }
    }
  }
  return value;
}

// temperature powers T^i (source/ThermalOperator.templates.hh:524-530)
__device__ __forceinline__ void temperature_powers(const MatConst &m, double T,
                                                   double (&tp)[ADM_MAX_POLY]) {
  tp[0] = 1.0;
#pragma unroll
  for (int i = 1; i < ADM_MAX_POLY; ++i) tp[i] = i <= m.p_order ? tp[i - 1] * T : 0.0;
}

// Property evaluators: polynomial, general table, merged table.
struct PolyEval {
  const MatConst &m;
  int mid;
  double tp[ADM_MAX_POLY];
  __device__ __forceinline__ PolyEval(const MatConst &m_, int mid_, double T)
      : m(m_), mid(mid_) {
    temperature_powers(m, T, tp);
  }
  template <int NS>
  __device__ __forceinline__ double operator()(int prop,
                                               const double (&r)[NS]) const {
    return property<NS, false>(m, prop, mid, r, 0.0, tp);
  }
};

struct TableEval {
  const MatConst &m;
  int mid;
  double T;
  __device__ __forceinline__ TableEval(const MatConst &m_, int mid_, double T_)
      : m(m_), mid(mid_), T(T_) {}
  template <int NS>
  __device__ __forceinline__ double operator()(int prop,
                                               const double (&r)[NS]) const {
    const double tp[ADM_MAX_POLY] = {1.0, 0.0, 0.0, 0.0, 0.0};
    return property<NS, true>(m, prop, mid, r, T, tp);
  }
};

// B and coef may live in LDS (cell kernel) or global memory (face kernel).
struct MergedEval {
  const double *coef;  // &merged.coef[0][0][0][0]
  int j;
  double d;
  __device__ __forceinline__ MergedEval(const double *B, const double *coef_,
                                        double T)
      : coef(coef_) {
    // branch-free binary search (selects only, so the searches of
    // independent points interleave)
    int k = 0;
#pragma unroll
    for (int step = kMergedMax / 2; step >= 1; step >>= 1)
      k += (B[k + step - 1] <= T) ? step : 0;
    j = k;
    d = T - B[k > 0 ? k - 1 : 0];
  }
  // interval already found (merged_interval)
  __device__ __forceinline__ MergedEval(const double *B, const double *coef_,
                                        double T, int k)
      : coef(coef_), j(k), d(T - B[k > 0 ? k - 1 : 0]) {}
  template <int NS>
  __device__ __forceinline__ double operator()(int prop,
                                               const double (&r)[NS]) const {
    double value = 0.0;
    const double *c = coef + ((prop * kMergedMax + j) * ADM_MAX_STATES) * 2;
    // every state is added, as in the reference (a zero ratio adds +-0)
#pragma unroll
    for (int s = 0; s < NS; ++s)
      value += r[s] * __builtin_fma(d, c[2 * s + 1], c[2 * s]);
    return value;
  }
};

// The interval index of MergedEval for M temperatures at once: the M
// searches advance together, so their table reads overlap.
template <int M>
__device__ __forceinline__ void merged_interval(const double *B,
                                                const double (&T)[M],
                                                int (&k)[M]) {
#pragma unroll
  for (int i = 0; i < M; ++i) k[i] = 0;
#pragma unroll
  for (int step = kMergedMax / 2; step >= 1; step >>= 1)
#pragma unroll
    for (int i = 0; i < M; ++i) k[i] += (B[k[i] + step - 1] <= T[i]) ? step : 0;
}

// update_state_ratios for one quadrature point. Returns true if the powder
// ratio changed.
template <int NS>
__device__ __forceinline__ bool state_ratios(const MatConst &m, int mid,
                                             double T, double &powder,
                                             double &liquid,
                                             double (&r)[NS]) {
  if constexpr (NS == 1) {
    r[0] = 1.0;
    liquid = 0.0;
    return false;
  } else {
    const double solidus = m.prop[mid][ADM_SOLIDUS];
    const double liquidus = m.prop[mid][ADM_LIQUIDUS];
    double l;
    if (T < solidus)
      l = 0.0;
    else if (T > liquidus)
      l = 1.0;
    else
      l = (T - solidus) / (liquidus - solidus);
    liquid = l;
    if constexpr (NS == 2) {
      r[1] = l;
      r[0] = 1.0 - l;
      return false;
    } else {
      const double a = 1.0 - l;
      const double p = (powder < a) ? powder : a;
      const bool changed = p != powder;
      powder = p;
      r[1] = l;
      r[2] = p;
      r[0] = 1.0 - l - p;
      return changed;
    }
  }
}

// update_state_ratios written without branches (selects only), returning
// the liquid ratio; the division of the mushy range is evaluated for every
// point and selected, which gives the same bits as the branchy form.
template <int NS>
__device__ __forceinline__ double state_ratios_select(const MatConst &m,
                                                      int mid, double T,
                                                      double &powder,
                                                      double (&r)[NS]) {
  if constexpr (NS == 1) {
    r[0] = 1.0;
    return 0.0;
  } else {
    const double solidus = m.prop[mid][ADM_SOLIDUS];
    const double liquidus = m.prop[mid][ADM_LIQUIDUS];
    const double mushy = (T - solidus) / (liquidus - solidus);
    const double l = T < solidus ? 0.0 : (T > liquidus ? 1.0 : mushy);
    if constexpr (NS == 2) {
      r[1] = l;
      r[0] = 1.0 - l;
    } else {
      const double a = 1.0 - l;
      const double p = (powder < a) ? powder : a;
      powder = p;
      r[1] = l;
      r[2] = p;
      r[0] = 1.0 - l - p;
    }
    return l;
  }
}

template <int NS, class Eval>
__device__ __forceinline__ double inv_rho_cp(const MatConst &m, int mid,
                                             const double (&r)[NS],
                                             const Eval &ev) {
  const double density = ev(ADM_DENSITY, r);
  double specific_heat = ev(ADM_SPECIFIC_HEAT, r);
  if constexpr (NS > 1) {
    const double is_mushy = r[1] * r[0];
    if (is_mushy > 0.0) specific_heat += m.latent_factor[mid];
  }
  return 1.0 / (density * specific_heat);
}

// ---------------------------------------------------------------------------
// heat sources (one lane per point)
// ---------------------------------------------------------------------------

// alpha of an ensemble member's beam from g = absorption_efficiency *
// max_power, with the operation order of GoldakHeatSource::update_time
// (source/GoldakHeatSource.cc:24-44: 2 eta P pmod / (r^2 d (pi/3)^1.5)) and
// ElectronBeamHeatSource::update_time (source/ElectronBeamHeatSource.cc:
// 24-45: -eta P pmod ln(0.1) / (pi r^2 d)); den is the parenthesised
// denominator. Scaling by 2 and negation are exact, so the result equals the
// single-operator host value bit for bit.
__device__ __forceinline__ double member_alpha(const adm_beam_state &s,
                                               double g, double pmod,
                                               double den, double log_01) {
  if (s.type == ADM_SOURCE_GOLDAK) return (2.0 * g) * pmod / den;
  if (s.type == ADM_SOURCE_ELECTRON_BEAM) return (-g) * pmod * log_01 / den;
  return s.alpha;
}

// The point a beam is evaluated at: rotation * (x, y, z) on a five-axis path
// (source/GoldakHeatSource.cc:95-103; Tensor<2,3> * Point<3> sums over the
// columns in order), the point itself otherwise.
__host__ __device__ __forceinline__ void source_point(const adm_beam_state &s,
                                                     double &x, double &y,
                                                     double &z) {
  if (!s.rotated) return;
  const double *R = s.rotation;
  double r[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double acc = R[3 * a] * x;
    acc += R[3 * a + 1] * y;
    acc += R[3 * a + 2] * z;
    r[a] = acc;
  }
  x = r[0];
  y = r[1];
  z = r[2];
}

__device__ __forceinline__ double beam_value(const adm_beam_state &s, double x,
                                             double y, double z_,
                                             double log_01) {
  if (s.type == ADM_SOURCE_CUBE) {
    if (!s.on) return 0.0;
    if (!(x >= s.cube_min[0] && x <= s.cube_max[0])) return 0.0;
    if (!(y >= s.cube_min[1] && y <= s.cube_max[1])) return 0.0;
    if (!(z_ >= s.cube_min[2] && z_ <= s.cube_max[2])) return 0.0;
    return 1.0 * s.cube_value;
  }
  source_point(s, x, y, z_);
  const double z = z_ - s.center[2];
  if (z + s.depth < 0.0) return 0.0;
  const double dx = x - s.center[0];
  const double dy = y - s.center[1];
  double xpy_squared = dx * dx;
  xpy_squared += dy * dy;
  const double xpy_area = 5. * s.radius_squared - xpy_squared;
  if (xpy_area < 0.) return 0.0;
  double exponent, factor = 1.0;
  if (s.type == ADM_SOURCE_GOLDAK) {
    // alpha * exp(-3 (z^2/d^2 + r^2/R^2))
    exponent = z * z * s.inv_depth_squared;
    exponent += xpy_squared * s.inv_radius_squared;
    exponent *= -3.0;
  } else {
    // alpha * exp(ln(0.1) r^2/R^2) * (-3 zr^2 - 2 zr + 1)
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
  // factor == 1 for Goldak: (alpha * e) * 1 is alpha * e exactly
  return 1.0 * s.alpha * exp(exponent) * factor;
}

// True if beam s can be non-zero somewhere in the box [lo, hi] (the cut-offs
// z + depth < 0 and |xy|^2 > 5 R^2 of the reference, or the cube's box).
__host__ __device__ __forceinline__ bool beam_may_touch(
    const adm_beam_state &s, const double (&lo_)[3], const double (&hi_)[3]) {
  if (s.type == ADM_SOURCE_CUBE) {
    if (!s.on) return false;
    for (int d = 0; d < 3; ++d)
      if (hi_[d] < s.cube_min[d] || lo_[d] > s.cube_max[d]) return false;
    return true;
  }
  double lo[3], hi[3];
  if (s.rotated) {
    // the axis-aligned hull of the rotated box (centre R c, half-widths
    // |R| h), widened so that rounding never culls a touched cell
    double c[3], h[3];
    for (int d = 0; d < 3; ++d) {
      c[d] = 0.5 * (lo_[d] + hi_[d]);
      h[d] = 0.5 * (hi_[d] - lo_[d]);
    }
    source_point(s, c[0], c[1], c[2]);
    for (int a = 0; a < 3; ++a) {
      double w = 0.0;
      for (int b = 0; b < 3; ++b) w += fabs(s.rotation[3 * a + b]) * h[b];
      w = w * (1.0 + 1e-6) + 1e-6 * fabs(c[a]);
      lo[a] = c[a] - w;
      hi[a] = c[a] + w;
    }
  } else {
    for (int d = 0; d < 3; ++d) {
      lo[d] = lo_[d];
      hi[d] = hi_[d];
    }
  }
  if (hi[2] - s.center[2] + s.depth < 0.0) return false;
  double d2 = 0.0;
  for (int d = 0; d < 2; ++d) {
    const double e = s.center[d] < lo[d]
                         ? lo[d] - s.center[d]
                         : (s.center[d] > hi[d] ? s.center[d] - hi[d] : 0.0);
    d2 += e * e;
  }
  // generous margin: a false positive only costs the evaluation
  return !(d2 > 5.0 * s.radius_squared * (1.0 + 1e-6));
}

// ---------------------------------------------------------------------------
// explicit-stepper stage updates (deal.II's stage and accumulation order),
// applied to the operator result r of entry i: by the stand-alone step
// kernels (r read from the residual vector) and by the plane path's gather
// pass (r as it is formed, without storing it)
// ---------------------------------------------------------------------------

// Forward Euler (source/ThermalPhysics.templates.hh:871-876): y = 1*y + dt*f
__device__ __forceinline__ void fe_stage(double *__restrict__ y,
                                         const double *__restrict__ inv_mass,
                                         double dt, int64_t i, double r) {
  const double f = r * inv_mass[i];
  y[i] = 1. * y[i] + dt * f;
}

// Classic RK4 stage, f_i = r*Minv: acc = acc + (dt b_i) f_i (acc starts as
// y); next stage input Y = y + (dt a_{i+1,i}) f_i (the other a_{i+1,j} of
// RK4 are zero, adding +-0); the last stage writes y = acc.
__device__ __forceinline__ void rk4_stage(double *__restrict__ y,
                                          const double *__restrict__ inv_mass,
                                          double *__restrict__ acc,
                                          double *__restrict__ Y, double bcoef,
                                          double acoef, int first, int last,
                                          int64_t i, double r) {
  const double f = r * inv_mass[i];
  const double yi = y[i];
  double s = first ? yi : acc[i];
  s = 1. * s + bcoef * f;
  if (last) {
    y[i] = s;
  } else {
    acc[i] = s;
    Y[i] = 1. * yi + acoef * f;
  }
}

// Generic explicit RK stage (RK3, whose a_{3,1} != 0): stores f_i in
// F[stage][n] and forms the next stage input from all stored stages, or
// the new y after the last.
__device__ __forceinline__ void rk_stage(double *__restrict__ y,
                                         const double *__restrict__ inv_mass,
                                         double *__restrict__ F, int stage,
                                         int n_stages, const double *a,
                                         double *__restrict__ Y,
                                         const double *b, int64_t n, int64_t i,
                                         double r) {
  const double f = r * inv_mass[i];
  F[(int64_t)stage * n + i] = f;
  if (stage + 1 < n_stages) {
    double s = y[i];
    for (int j = 0; j <= stage; ++j) s = 1. * s + a[j] * F[(int64_t)j * n + i];
    Y[i] = s;
  } else {
    double s = y[i];
    for (int j = 0; j < n_stages; ++j) s = 1. * s + b[j] * F[(int64_t)j * n + i];
    y[i] = s;
  }
}

// A stage update for the gather pass to apply instead of writing dst.
struct StageArgs {
  int32_t kind;  // kStageNone, kStageFE, kStageRK4, kStageRK
  int32_t first, last, stage, n_stages;
  double *y;
  const double *inv_mass;
  double *acc, *Y, *F;
  double c0, c1;      // FE: dt; RK4: dt b_i, dt a_{i+1,i}
  double a[3], b[3];  // generic stage: dt a_{i+1,j}, dt b_j
  int64_t n;          // entries of y (every member)
};
constexpr int32_t kStageNone = 0, kStageFE = 1, kStageRK4 = 2, kStageRK = 3;

__device__ __forceinline__ void stage_apply(const StageArgs &st, int64_t i,
                                            double r) {
  if (st.kind == kStageFE)
    fe_stage(st.y, st.inv_mass, st.c0, i, r);
  else if (st.kind == kStageRK4)
    rk4_stage(st.y, st.inv_mass, st.acc, st.Y, st.c0, st.c1, st.first,
              st.last, i, r);
  else
    rk_stage(st.y, st.inv_mass, st.F, st.stage, st.n_stages, st.a, st.Y, st.b,
             st.n, i, r);
}

}  // namespace adm
