// Material deposition on the device: get_elements_to_activate
// (source/material_deposition.cc:376-435) and add_material_start +
// add_material_end (source/ThermalPhysics.templates.hh:595-825) as a handful
// of kernels over the cells and the box nodes, with no copy of the phase
// state to the host.
//
//   1. box_kernel       per cell: the last deposition box whose closed box
//                       meets the closed cell (inactive cells only)
//   2. cell_kernel      per cell: the quadrature-point liquid and powder
//                       ratios to their cell means (every cell, as the
//                       reference's set_state + get_state_from_material_
//                       properties round trip does), activation with the
//                       box's deposition direction, the meta word
//   3. face_count       per cell: the faces with a boundary term (the mask
//                       goes into the meta word) and their count
//   4. node_count       per box node: the face-point slots at the node (for
//                       the face gather pass) and the new-material
//                       temperature on nodes that no previously active cell
//                       held
//   5. two exclusive scans (hipCUB) give the face list and the node list
//      positions; one 24-byte copy to the host sizes the lists
//   6. face_write, node_fill: the face list (cells in increasing order,
//      faces in increasing id, as the host build orders it), the faces'
//      powder at the cell mean, and the node -> slot lists in face order
//   7. the lumped inverse mass (compute_inverse_mass)
//
// Every sum is taken in the order of the host restatement, so the state
// means, the face lists and the mass equal those of a rebuild from scratch
// bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "../../include/adm.h"
#include "adm_context.hpp"

namespace adm {

#define ADM_HIP_ACT(call)                                                 \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      throw AdmError(ADM_ERR_HIP, std::string(#call) + ": " +             \
                                      hipGetErrorString(e_));             \
  } while (0)

namespace {

struct ActGeom {
  int nc[3], nn[3];
  int p, nq, nfq, n_states;
  int64_t n_cells, n_dofs;
};

__device__ __forceinline__ void cell_coords(const ActGeom &g, int64_t c,
                                            int cc[3]) {
  cc[0] = (int)(c % g.nc[0]);
  cc[1] = (int)((c / g.nc[0]) % g.nc[1]);
  cc[2] = (int)(c / ((int64_t)g.nc[0] * g.nc[1]));
}

// get_elements_to_activate: boxes[8 b + 0..5] = {min x, y, z, max x, y, z};
// the bounds are formed without contraction, as the host forms them.
__global__ void box_kernel(ActGeom g, const KConst *__restrict__ kc,
                           const uint32_t *__restrict__ meta, int n_boxes,
                           const double *__restrict__ boxes,
                           int32_t *__restrict__ box_of_cell) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.n_cells) return;
  int32_t hit_box = -1;
  if (!(meta[c] & kActive)) {
    int cc[3];
    cell_coords(g, c, cc);
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      const double o = kc->origin[d], h = kc->h[d];
      lo[d] = __dadd_rn(o, __dmul_rn((double)cc[d], h));
      hi[d] = __dadd_rn(o, __dmul_rn((double)(cc[d] + 1), h));
    }
    for (int b = 0; b < n_boxes; ++b) {
      const double *bx = boxes + 8 * b;
      bool hit = true;
      for (int d = 0; d < 3; ++d)
        if (!(lo[d] <= bx[3 + d] && bx[d] <= hi[d])) hit = false;
      if (hit) hit_box = b;
    }
  }
  box_of_cell[c] = hit_box;
}

// State means and activation. boxes[8 b + 6, 7] = cos, sin of box b.
__global__ void cell_kernel(ActGeom g, uint32_t *__restrict__ meta,
                            const int32_t *__restrict__ box_of_cell,
                            const double *__restrict__ boxes,
                            double *__restrict__ dep_cos,
                            double *__restrict__ dep_sin,
                            double *__restrict__ liquid,
                            uint8_t *__restrict__ liquid_dirty,
                            double *__restrict__ powder,
                            const int32_t *__restrict__ powder_slot,
                            double *__restrict__ cell_powder) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.n_cells) return;
  const int nq = g.nq;
  uint32_t m = meta[c] & ~kNewlyActive;
  // a cell without the "held liquid" flag holds zeros at every point: its
  // mean is zero and nothing changes
  if (g.n_states > 1 && liquid_dirty[c]) {
    double *lq = liquid + c * nq;
    double ls = 0.0;
    for (int q = 0; q < nq; ++q) ls += lq[q];
    const double l = ls / nq;
    for (int q = 0; q < nq; ++q) lq[q] = l;
    liquid_dirty[c] = l != 0.0;
  }
  double pm = 0.0;
  if (g.n_states == 3 && powder_slot[c] >= 0) {
    double *pw = powder + (int64_t)powder_slot[c] * nq;
    double ps = 0.0;
    for (int q = 0; q < nq; ++q) ps += pw[q];
    pm = ps / nq;
    for (int q = 0; q < nq; ++q) pw[q] = pm;
    // a cell whose powder is gone keeps its slot (zeros) but leaves the
    // powder path of the kernels
    m = pm > 0.0 ? (m | kHasPowder) : (m & ~kHasPowder);
  }
  cell_powder[c] = pm;
  const int32_t b = box_of_cell[c];
  if (b >= 0 && !(m & kActive)) {
    const double cs = boxes[8 * b + 6], sn = boxes[8 * b + 7];
    dep_cos[c] = cs;
    dep_sin[c] = sn;
    m |= kActive | kNewlyActive;
    m = (cs != 1.0 || sn != 0.0) ? (m | kRotated) : (m & ~kRotated);
  }
  meta[c] = m;
}

// Faces of the active domain with a boundary term (Context::face_mask).
__device__ __forceinline__ uint32_t face_mask_dev(const ActGeom &g,
                                                  const KConst &kc,
                                                  const uint32_t *meta,
                                                  int64_t c) {
  if (kc.adiabatic_only) return 0;
  int cc[3];
  cell_coords(g, c, cc);
  uint32_t mask = 0;
  for (int f = 0; f < 6; ++f) {
    const int dd = f / 2, side = f % 2;
    const int nb = cc[dd] + (side ? 1 : -1);
    unsigned bid;
    if (nb < 0 || nb >= g.nc[dd]) {
      bid = (unsigned)f;
    } else {
      int cn[3] = {cc[0], cc[1], cc[2]};
      cn[dd] = nb;
      const int64_t c2 =
          cn[0] + (int64_t)g.nc[0] * (cn[1] + (int64_t)g.nc[1] * cn[2]);
      if (meta[c2] & kActive) continue;
      bid = 0xffffffffu;
    }
    const int n = kc.n_boundary_ids;
    const int t = bid < (unsigned)(n - 1) ? kc.boundary_type[bid]
                                          : kc.boundary_type[n - 1];
    if (t & (ADM_BC_RADIATIVE | ADM_BC_CONVECTIVE)) mask |= 1u << f;
  }
  return mask;
}

__global__ void face_count_kernel(ActGeom g, const KConst *__restrict__ kc,
                                  uint32_t *__restrict__ meta,
                                  int32_t *__restrict__ count) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.n_cells) return;
  const uint32_t m = meta[c];
  const uint32_t fm = (m & kActive) ? face_mask_dev(g, *kc, meta, c) : 0u;
  count[c] = __popc(fm);
  // the mask, for the node passes; neighbours read only the activity bit,
  // which this store leaves as it was
  meta[c] = (m & ~(63u << kFaceShift)) | (fm << kFaceShift);
}

// The cells holding box node (i0, i1, i2), in increasing cell index.
// f(cell, local index of the node in it) for every cell holding the node,
// in increasing cell order (z, then y, then x): up to 2 x 2 x 2 cells, in
// unrolled loops (no per-lane arrays, which would live in scratch memory)
template <class F>
__device__ __forceinline__ void for_node_cells(const ActGeom &g, int64_t node,
                                               F &&f) {
  const int i[3] = {(int)(node % g.nn[0]), (int)((node / g.nn[0]) % g.nn[1]),
                    (int)(node / ((int64_t)g.nn[0] * g.nn[1]))};
  int lo[3], cnt[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int c_hi = min(i[d] / g.p, g.nc[d] - 1);
    const int c_lo = (i[d] % g.p == 0 && i[d] > 0) ? i[d] / g.p - 1 : c_hi;
    lo[d] = c_lo;
    cnt[d] = c_hi - c_lo + 1;
  }
#pragma unroll
  for (int z = 0; z < 2; ++z)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        if (z >= cnt[2] || y >= cnt[1] || x >= cnt[0]) continue;
        const int cx = lo[0] + x, cy = lo[1] + y, cz = lo[2] + z;
        const int l[3] = {i[0] - cx * g.p, i[1] - cy * g.p, i[2] - cz * g.p};
        f(cx + (int64_t)g.nc[0] * (cy + (int64_t)g.nc[1] * cz), l);
      }
}

// Face f = 2 d + side of a cell holds the node with local index l when
// l[d] is 0 (side 0) or p (side 1); its slot inside the face is
// l[t1] + (p+1) l[t2], t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2.
__device__ __forceinline__ bool on_face(int p, int f, const int (&l)[3],
                                        int &q) {
  const int d = f / 2, side = f % 2;
  if (l[d] != (side ? p : 0)) return false;
  const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
  q = l[t1] + (p + 1) * l[t2];
  return true;
}

// packed[node] = slots | (slots > 0) << 32; new-material temperature
__global__ void node_count_kernel(ActGeom g, const uint32_t *__restrict__ meta,
                                  uint64_t *__restrict__ packed,
                                  double *__restrict__ T, double new_T) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= g.n_dofs) return;
  uint32_t slots = 0;
  bool before = false, after = false;
  for_node_cells(g, node, [&](int64_t cell, const int (&l)[3]) {
    const uint32_t m = meta[cell];
    if (!(m & kActive)) return;
    after = true;
    if (!(m & kNewlyActive)) before = true;
    const uint32_t fm = (m >> kFaceShift) & 63u;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      int q;
      if ((fm >> f) & 1u) slots += on_face(g.p, f, l, q) ? 1u : 0u;
    }
  });
  packed[node] = (uint64_t)slots | ((uint64_t)(slots > 0) << 32);
  if (T && after && !before) T[node] = new_T;
}

// totals[0] = faces, [1] = face-node slots, [2] = face nodes
__global__ void totals_kernel(ActGeom g, const int32_t *__restrict__ count,
                              const int32_t *__restrict__ start,
                              const uint64_t *__restrict__ packed,
                              const uint64_t *__restrict__ scan,
                              int64_t *__restrict__ totals) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t nc = g.n_cells, nn = g.n_dofs;
  totals[0] = (int64_t)start[nc - 1] + count[nc - 1];
  const uint64_t t = scan[nn - 1] + packed[nn - 1];
  totals[1] = (int64_t)(t & 0xffffffffull);
  totals[2] = (int64_t)(t >> 32);
}

__global__ void face_write_kernel(ActGeom g, const KConst *__restrict__ kc,
                                  const uint32_t *__restrict__ meta,
                                  const int32_t *__restrict__ start,
                                  const double *__restrict__ cell_powder,
                                  FaceRef *__restrict__ faces,
                                  double *__restrict__ face_powder) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.n_cells) return;
  const uint32_t m = meta[c];
  const uint32_t fm = (m >> kFaceShift) & 63u;
  if (!fm) return;
  const int material = (int)((m >> kMaterialShift) & 3u);
  int cc[3];
  cell_coords(g, c, cc);
  int64_t fi = start[c];
  const int nb = kc->n_boundary_ids;
  for (int f = 0; f < 6; ++f) {
    if (!((fm >> f) & 1u)) continue;
    const int dd = f / 2, side = f % 2;
    const bool domain = side ? (cc[dd] == g.nc[dd] - 1) : (cc[dd] == 0);
    const unsigned bid = domain ? (unsigned)f : 0xffffffffu;
    const int bt = bid < (unsigned)(nb - 1) ? kc->boundary_type[bid]
                                            : kc->boundary_type[nb - 1];
    FaceRef fr;
    fr.cell = (int32_t)c;
    fr.info = f | ((bt & 31) << 3) | (material << 8);
    faces[fi] = fr;
    if (face_powder)
      for (int q = 0; q < g.nfq; ++q)
        face_powder[fi * g.nfq + q] = cell_powder[c];
    ++fi;
  }
}

__global__ void node_fill_kernel(ActGeom g, const uint32_t *__restrict__ meta,
                                 const int32_t *__restrict__ start,
                                 const uint64_t *__restrict__ packed,
                                 const uint64_t *__restrict__ scan,
                                 int64_t *__restrict__ fn_node,
                                 int32_t *__restrict__ fn_offset,
                                 int32_t *__restrict__ fn_slot,
                                 int32_t *__restrict__ fn_rank,
                                 int64_t n_fnodes) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= g.n_dofs) return;
  const uint64_t pk = packed[node];
  const uint64_t sc = scan[node];
  const int64_t rank = (int64_t)(sc >> 32);
  fn_rank[node] = (pk >> 32) ? (int32_t)rank : -1;
  if (!(pk >> 32)) return;
  const int32_t off = (int32_t)(sc & 0xffffffffull);
  fn_node[rank] = node;
  fn_offset[rank] = off;
  if (rank == n_fnodes - 1)
    fn_offset[rank + 1] = off + (int32_t)(pk & 0xffffffffull);
  const int NF = (g.p + 1) * (g.p + 1);
  int e = off;
  for_node_cells(g, node, [&](int64_t cell, const int (&l)[3]) {
    const uint32_t m = meta[cell];
    if (!(m & kActive)) return;
    const uint32_t fm = (m >> kFaceShift) & 63u;
    int32_t fi = start[cell];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      if (!((fm >> f) & 1u)) continue;
      int q;
      if (on_face(g.p, f, l, q)) fn_slot[e++] = fi * NF + q;
      ++fi;
    }
  });
}

// Tbar = 0 on the nodes the last activation added
__global__ void fresh_zero_kernel(ActGeom g, const uint32_t *__restrict__ meta,
                                  double *__restrict__ Tbar) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= g.n_dofs) return;
  bool before = false, after = false;
  for_node_cells(g, node, [&](int64_t cell, const int (&)[3]) {
    const uint32_t m = meta[cell];
    if (!(m & kActive)) return;
    after = true;
    if (!(m & kNewlyActive)) before = true;
  });
  if (after && !before) Tbar[node] = 0.0;
}

// pbar of a cell's points <- (sum of them + sum over the cell's faces'
// points of fpbar) / nq: the adjoint of p_q = mean and face p = mean.
// Faces are in increasing cell order; a binary search finds the cell's.
__global__ void transfer_adjoint_kernel(ActGeom g,
                                        const int32_t *__restrict__ slot,
                                        const FaceRef *__restrict__ faces,
                                        int64_t n_faces,
                                        const double *__restrict__ fpbar,
                                        double *__restrict__ pbar) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.n_cells) return;
  if (slot[c] < 0) return;
  double *pb = pbar + (int64_t)slot[c] * g.nq;
  double s = 0.0;
  for (int q = 0; q < g.nq; ++q) s += pb[q];
  if (fpbar && n_faces > 0) {
    int64_t lo = 0, hi = n_faces;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (faces[mid].cell < c) lo = mid + 1; else hi = mid;
    }
    for (int64_t f = lo; f < n_faces && faces[f].cell == c; ++f)
      for (int q = 0; q < g.nfq; ++q) s += fpbar[f * g.nfq + q];
  }
  const double v = s / g.nq;
  for (int q = 0; q < g.nq; ++q) pb[q] = v;
}

template <class T>
void grow(T *&ptr, size_t &cap, size_t n) {
  if (n <= cap && ptr) return;
  if (ptr) ADM_HIP_ACT(hipFree(ptr));
  ptr = nullptr;
  const size_t c = std::max<size_t>(std::max<size_t>(n, 2 * cap), 1);
  ADM_HIP_ACT(hipMalloc(&ptr, sizeof(T) * c));
  cap = c;
}

}  // namespace

ActWork::~ActWork() {
  (void)hipFree(boc);
  (void)hipFree(boxes);
  (void)hipFree(count);
  (void)hipFree(start);
  (void)hipFree(cell_powder);
  (void)hipFree(packed);
  (void)hipFree(scan);
  (void)hipFree(temp);
  (void)hipFree(totals);
  if (h_totals) (void)hipHostFree(h_totals);
  if (h_boc) (void)hipHostFree(h_boc);
  if (h_boxes) (void)hipHostFree(h_boxes);
}

GeomSnap::~GeomSnap() {
  (void)hipFree(meta);
  (void)hipFree(cos);
  (void)hipFree(sin);
  (void)hipFree(faces);
  (void)hipFree(fn_node);
  (void)hipFree(fn_offset);
  (void)hipFree(fn_slot);
  (void)hipFree(fn_rank);
}

void Context::save_geometry(GeomSnap &g, hipStream_t s,
                            bool host_copy) const {
  auto keep = [&](auto *&dst, const auto *src, size_t count) {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
    (void)hipFree(dst);
    dst = nullptr;
    ADM_HIP_ACT(hipMalloc(&dst, sizeof(T) * std::max<size_t>(count, 1)));
    if (count)
      ADM_HIP_ACT(hipMemcpyAsync(dst, src, sizeof(T) * count,
                                 hipMemcpyDeviceToDevice, s));
  };
  int64_t n_slots = 0;
  if (n_fnodes > 0)
    ADM_HIP_ACT(hipMemcpy(&n_slots, d_fn_offset + n_fnodes, sizeof(int32_t),
                          hipMemcpyDeviceToHost));
  n_slots &= 0xffffffff;
  keep(g.meta, d_meta, (size_t)n_cells);
  keep(g.cos, d_cos, (size_t)n_cells);
  keep(g.sin, d_sin, (size_t)n_cells);
  keep(g.faces, d_faces, (size_t)n_faces);
  keep(g.fn_node, d_fn_node, (size_t)n_fnodes);
  keep(g.fn_offset, d_fn_offset, n_fnodes ? (size_t)n_fnodes + 1 : 0);
  keep(g.fn_slot, d_fn_slot, (size_t)n_slots);
  keep(g.fn_rank, d_fn_rank, (size_t)n_dofs);
  g.n_faces = n_faces;
  g.n_fnodes = n_fnodes;
  g.n_slots = n_slots;
  g.nz_active = nz_active;
  g.host_copy = host_copy;
  g.log_cell.clear();
  g.log_old.clear();
  g.log_new.clear();
  if (host_copy) {
    g.active = active;
    g.dep_cos = dep_cos;
    g.dep_sin = dep_sin;
  }
}

void Context::restore_geometry(const GeomSnap &g, hipStream_t s) {
  const int NF = (p + 1) * (p + 1);
  grow(d_faces, face_cap, (size_t)g.n_faces);
  grow(d_face_out, face_out_cap, (size_t)g.n_faces * NF);
  if (n_states == 3) grow(d_face_powder, face_powder_cap, (size_t)g.n_faces * nfq);
  grow(d_fn_node, fn_node_cap, (size_t)g.n_fnodes);
  grow(d_fn_offset, fn_offset_cap, (size_t)g.n_fnodes + 1);
  grow(d_fn_slot, fn_slot_cap, (size_t)g.n_slots);
  auto put = [&](auto *dst, const auto *src, size_t count) {
    if (count)
      ADM_HIP_ACT(hipMemcpyAsync(dst, src, sizeof(*src) * count,
                                 hipMemcpyDeviceToDevice, s));
  };
  put(d_meta, g.meta, (size_t)n_cells);
  put(d_cos, g.cos, (size_t)n_cells);
  put(d_sin, g.sin, (size_t)n_cells);
  put(d_faces, g.faces, (size_t)g.n_faces);
  put(d_fn_node, g.fn_node, (size_t)g.n_fnodes);
  put(d_fn_offset, g.fn_offset, g.n_fnodes ? (size_t)g.n_fnodes + 1 : 0);
  put(d_fn_slot, g.fn_slot, (size_t)g.n_slots);
  put(d_fn_rank, g.fn_rank, (size_t)n_dofs);
  n_faces = g.n_faces;
  n_fnodes = g.n_fnodes;
  nz_active = g.nz_active;
  if (g.host_copy) {
    active = g.active;
    dep_cos = g.dep_cos;
    dep_sin = g.dep_sin;
  } else {
    for (size_t i = 0; i < g.log_cell.size(); ++i) {
      const int64_t c = g.log_cell[i];
      active[c] = 0;
      dep_cos[c] = g.log_old[2 * i];
      dep_sin[c] = g.log_old[2 * i + 1];
    }
  }
  faces_stale = true;
}

void Context::redo_host_geometry(const GeomSnap &g) {
  for (size_t i = 0; i < g.log_cell.size(); ++i) {
    const int64_t c = g.log_cell[i];
    active[c] = 1;
    dep_cos[c] = g.log_new[2 * i];
    dep_sin[c] = g.log_new[2 * i + 1];
  }
  nz_active = g.nz_after;
}

void Context::activation_adjoint(double *Tbar, double *pbar,
                                 const double *fpbar, hipStream_t s) {
  ActGeom g;
  for (int d = 0; d < 3; ++d) {
    g.nc[d] = nc[d];
    g.nn[d] = nn[d];
  }
  g.p = p;
  g.nq = nq;
  g.nfq = nfq;
  g.n_states = n_states;
  g.n_cells = n_cells;
  g.n_dofs = n_dofs;
  if (Tbar)
    hipLaunchKernelGGL(fresh_zero_kernel,
                       dim3((unsigned)((n_dofs + 255) / 256)), dim3(256), 0,
                       s, g, d_meta, Tbar);
  if (pbar && n_states == 3)
    hipLaunchKernelGGL(transfer_adjoint_kernel,
                       dim3((unsigned)((n_cells + 255) / 256)), dim3(256), 0,
                       s, g, d_powder_slot, d_faces, n_faces, fpbar, pbar);
  ADM_HIP_ACT(hipGetLastError());
}

void Context::sync_faces_host() {
  if (!faces_stale) return;
  ADM_HIP_ACT(hipDeviceSynchronize());
  faces.resize((size_t)n_faces);
  if (n_faces)
    ADM_HIP_ACT(hipMemcpy(faces.data(), d_faces, sizeof(FaceRef) * n_faces,
                          hipMemcpyDeviceToHost));
  faces_stale = false;
}

// box_of_cell: host, one entry per cell (NULL: computed on the device from
// boxes); box_of_cell_out: host, filled with the cells' boxes (may be NULL).
void Context::activate(int n_boxes, const double *boxes,
                       const int32_t *box_of_cell, const double *box_cos,
                       const double *box_sin, double new_temperature,
                       double *T, int32_t *box_of_cell_out, hipStream_t s) {
  ActGeom g;
  for (int d = 0; d < 3; ++d) {
    g.nc[d] = nc[d];
    g.nn[d] = nn[d];
  }
  g.p = p;
  g.nq = nq;
  g.nfq = nfq;
  g.n_states = n_states;
  g.n_cells = n_cells;
  g.n_dofs = n_dofs;
  const unsigned bc = (unsigned)((n_cells + 255) / 256);
  const unsigned bn = (unsigned)((n_dofs + 255) / 256);
  const int NF = (p + 1) * (p + 1);
  if (!act) act = std::make_shared<ActWork>();
  ActWork &w = *act;
  if (!w.boc) {
    ADM_HIP_ACT(hipMalloc(&w.boc, sizeof(int32_t) * n_cells));
    ADM_HIP_ACT(hipMalloc(&w.count, sizeof(int32_t) * n_cells));
    ADM_HIP_ACT(hipMalloc(&w.start, sizeof(int32_t) * n_cells));
    ADM_HIP_ACT(hipMalloc(&w.cell_powder, sizeof(double) * n_cells));
    ADM_HIP_ACT(hipMalloc(&w.packed, sizeof(uint64_t) * n_dofs));
    ADM_HIP_ACT(hipMalloc(&w.scan, sizeof(uint64_t) * n_dofs));
    ADM_HIP_ACT(hipMalloc(&w.totals, sizeof(int64_t) * 4));
    ADM_HIP_ACT(hipHostMalloc(&w.h_totals, sizeof(int64_t) * 4,
                              hipHostMallocDefault));
    ADM_HIP_ACT(hipHostMalloc(&w.h_boc, sizeof(int32_t) * n_cells,
                              hipHostMallocDefault));
    size_t b1 = 0, b2 = 0;
    ADM_HIP_ACT(hipcub::DeviceScan::ExclusiveSum(nullptr, b1, w.count, w.start,
                                                 (int)n_cells, s));
    ADM_HIP_ACT(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, w.packed,
                                                 w.scan, (int)n_dofs, s));
    w.temp_bytes = std::max(b1, b2);
    ADM_HIP_ACT(hipMalloc(&w.temp, w.temp_bytes));
  }
  // boxes: 6 bounds, cos, sin, staged in pinned memory (the previous
  // activation's copies out of it finished before that call returned)
  const size_t nbx = (size_t)8 * std::max(n_boxes, 1);
  if (nbx > w.h_boxes_cap) {
    if (w.h_boxes) ADM_HIP_ACT(hipHostFree(w.h_boxes));
    w.h_boxes = nullptr;
    ADM_HIP_ACT(hipHostMalloc(&w.h_boxes, sizeof(double) * nbx,
                              hipHostMallocDefault));
    w.h_boxes_cap = nbx;
  }
  std::fill(w.h_boxes, w.h_boxes + nbx, 0.0);
  for (int b = 0; b < n_boxes; ++b) {
    if (boxes)
      for (int k = 0; k < 6; ++k) w.h_boxes[8 * b + k] = boxes[6 * b + k];
    w.h_boxes[8 * b + 6] = box_cos[b];
    w.h_boxes[8 * b + 7] = box_sin[b];
  }
  grow(w.boxes, w.boxes_cap, nbx);
  ADM_HIP_ACT(hipMemcpyAsync(w.boxes, w.h_boxes, sizeof(double) * nbx,
                             hipMemcpyHostToDevice, s));
  if (box_of_cell) {
    std::memcpy(w.h_boc, box_of_cell, sizeof(int32_t) * n_cells);
    ADM_HIP_ACT(hipMemcpyAsync(w.boc, w.h_boc, sizeof(int32_t) * n_cells,
                               hipMemcpyHostToDevice, s));
  } else {
    hipLaunchKernelGGL(box_kernel, dim3(bc), dim3(256), 0, s, g, d_kc, d_meta,
                       n_boxes, w.boxes, w.boc);
  }
  hipLaunchKernelGGL(cell_kernel, dim3(bc), dim3(256), 0, s, g, d_meta, w.boc,
                     w.boxes, d_cos, d_sin, d_liquid, d_liquid_dirty, d_powder,
                     d_powder_slot, w.cell_powder);
  hipLaunchKernelGGL(face_count_kernel, dim3(bc), dim3(256), 0, s, g, d_kc,
                     d_meta, w.count);
  hipLaunchKernelGGL(node_count_kernel, dim3(bn), dim3(256), 0, s, g, d_meta,
                     w.packed, T, new_temperature);
  ADM_HIP_ACT(hipGetLastError());
  size_t tb = w.temp_bytes;
  ADM_HIP_ACT(hipcub::DeviceScan::ExclusiveSum(w.temp, tb, w.count, w.start,
                                               (int)n_cells, s));
  tb = w.temp_bytes;
  ADM_HIP_ACT(hipcub::DeviceScan::ExclusiveSum(w.temp, tb, w.packed, w.scan,
                                               (int)n_dofs, s));
  hipLaunchKernelGGL(totals_kernel, dim3(1), dim3(64), 0, s, g, w.count,
                     w.start, w.packed, w.scan, w.totals);
  ADM_HIP_ACT(hipGetLastError());
  ADM_HIP_ACT(hipMemcpyAsync(w.h_totals, w.totals, sizeof(int64_t) * 3,
                             hipMemcpyDeviceToHost, s));
  // The cells the host bookkeeping below visits: with boxes, only the cell
  // layers the boxes' z extents meet, with one layer of margin on either
  // side (box_kernel's closed-box test finds no box for a cell outside
  // them: their box_of_cell_out entries are -1); otherwise every cell.
  int64_t c_lo = 0, c_hi = n_cells;
  if (!box_of_cell && boxes && n_boxes > 0) {
    // (the extents' lower and higher ends, whichever bound holds them)
    double zlo = boxes[2], zhi = boxes[2];
    bool finite = true;
    for (int b = 0; b < n_boxes; ++b) {
      const double a0 = boxes[6 * b + 2], a1 = boxes[6 * b + 5];
      finite = finite && std::isfinite(a0) && std::isfinite(a1);
      zlo = std::min(zlo, std::min(a0, a1));
      zhi = std::max(zhi, std::max(a0, a1));
    }
    const double o = kc.origin[2], h = kc.h[2];
    if (finite && h > 0.0) {
      const double f0 = std::floor((zlo - o) / h) - 1.0;
      const double f1 = std::floor((zhi - o) / h) + 1.0;
      const double top = (double)(nc[2] - 1);
      const int64_t z0 = (int64_t)std::min(std::max(f0, 0.0), top);
      const int64_t z1 = (int64_t)std::min(std::max(f1, 0.0), top);
      const int64_t layer = (int64_t)nc[0] * nc[1];
      c_lo = z0 * layer;
      c_hi = z1 >= z0 ? (z1 + 1) * layer : c_lo;
    }
  }
  if (c_hi > c_lo)
    ADM_HIP_ACT(hipMemcpyAsync(w.h_boc + c_lo, w.boc + c_lo,
                               sizeof(int32_t) * (c_hi - c_lo),
                               hipMemcpyDeviceToHost, s));
  ADM_HIP_ACT(hipStreamSynchronize(s));
  const int64_t nf = w.h_totals[0], n_slots = w.h_totals[1],
                n_nodes = w.h_totals[2];
  if (n_slots > (int64_t)INT32_MAX || nf * NF > (int64_t)INT32_MAX)
    throw AdmError(ADM_ERR_UNSUPPORTED, "too many boundary-face nodes");
  // list capacities (a larger list reallocates; hipFree waits for the
  // kernels that still read the old one)
  grow(d_faces, face_cap, (size_t)nf);
  grow(d_face_out, face_out_cap, (size_t)nf * NF);
  if (n_states == 3) grow(d_face_powder, face_powder_cap, (size_t)nf * nfq);
  grow(d_fn_node, fn_node_cap, (size_t)n_nodes);
  grow(d_fn_offset, fn_offset_cap, (size_t)n_nodes + 1);
  grow(d_fn_slot, fn_slot_cap, (size_t)n_slots);
  hipLaunchKernelGGL(face_write_kernel, dim3(bc), dim3(256), 0, s, g, d_kc,
                     d_meta, w.start, w.cell_powder, d_faces,
                     n_states == 3 ? d_face_powder : nullptr);
  hipLaunchKernelGGL(node_fill_kernel, dim3(bn), dim3(256), 0, s, g, d_meta,
                     w.start, w.packed, w.scan, d_fn_node, d_fn_offset,
                     d_fn_slot, d_fn_rank, n_nodes);
  ADM_HIP_ACT(hipGetLastError());
  n_faces = nf;
  n_fnodes = n_nodes;
  faces_stale = true;
  // host bookkeeping of the activity and deposition directions
  GeomSnap *const lg = geom_log;
  geom_log = nullptr;
  if (box_of_cell_out) {
    std::fill(box_of_cell_out, box_of_cell_out + c_lo, -1);
    std::fill(box_of_cell_out + std::max(c_hi, c_lo), box_of_cell_out + n_cells,
              -1);
  }
  for (int64_t c = c_lo; c < c_hi; ++c) {
    const int32_t b = w.h_boc[c];
    if (box_of_cell_out) box_of_cell_out[c] = b;
    if (b < 0 || active[c]) continue;
    if (lg) {
      lg->log_cell.push_back(c);
      lg->log_old.push_back(dep_cos[c]);
      lg->log_old.push_back(dep_sin[c]);
      lg->log_new.push_back(box_cos[b]);
      lg->log_new.push_back(box_sin[b]);
    }
    active[c] = 1;
    dep_cos[c] = box_cos[b];
    dep_sin[c] = box_sin[b];
    nz_active = std::max(nz_active, (int)(c / ((int64_t)nc[0] * nc[1])) + 1);
  }
  if (lg) lg->nz_after = nz_active;
  if (sb.inv_mass) compute_inverse_mass(sb.inv_mass, s);
}

}  // namespace adm
