// Vector kernels of the explicit time steppers (non-template; compiled in
// the API translation unit only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adm_device.hpp"

namespace adm {

// Stage counter of a captured step loop: set before the replays, advanced
// after every operator application inside the graph.
__global__ void set_slot_kernel(int32_t *slot, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) slot[0] = v;
}
__global__ void advance_slot_kernel(int32_t *slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) slot[0] += 1;
}

__global__ void fill_kernel(double *__restrict__ x, double v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// DoF map (adm_set_dof_map): an external vector in the caller's DoF order
// into the box-node vector (node[map[i]] = ext[i]), and back
// (ext[i] = node[map[i]], or += with add).
__global__ void dofs_to_nodes_kernel(const double *__restrict__ ext,
                                     const int64_t *__restrict__ map,
                                     int64_t n_ext, double *__restrict__ node) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_ext) node[map[i]] = ext[i];
}
__global__ void nodes_to_dofs_kernel(const double *__restrict__ node,
                                     const int64_t *__restrict__ map,
                                     int64_t n_ext, double *__restrict__ ext,
                                     int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_ext) return;
  const double v = node[map[i]];
  ext[i] = add ? ext[i] + v : v;
}

// T[idx[k]] = v (values of newly activated nodes)
__global__ void scatter_value_kernel(double *__restrict__ T,
                                     const int64_t *__restrict__ idx, double v,
                                     int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) T[idx[k]] = v;
}

// Forward Euler (source/ThermalPhysics.templates.hh:871-876) on the
// residual R; R is cleared for the next stage.
__global__ void fe_update_kernel(double *__restrict__ T, double *__restrict__ R,
                                 const double *__restrict__ inv_mass, double dt,
                                 int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fe_stage(T, inv_mass, dt, i, R[i]);
  R[i] = 0.0;
}

// Classic RK4 stage of deal.II's ExplicitRungeKutta (rk4_stage) on the
// residual R; R is cleared for the next stage.
__global__ void rk4_stage_kernel(double *__restrict__ y, double *__restrict__ R,
                                 const double *__restrict__ inv_mass,
                                 double *__restrict__ acc,
                                 double *__restrict__ Y, double bcoef,
                                 double acoef, int first, int last,
                                 int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = R[i];
  R[i] = 0.0;
  rk4_stage(y, inv_mass, acc, Y, bcoef, acoef, first, last, i, r);
}

// Generic explicit RK stage (rk_stage, used for RK3) on the residual R.
__global__ void rk_stage_kernel(double *__restrict__ y, double *__restrict__ R,
                                const double *__restrict__ inv_mass,
                                double *__restrict__ F, int stage, int n_stages,
                                double a0, double a1, double a2,
                                double *__restrict__ Y, double b0, double b1,
                                double b2, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = R[i];
  R[i] = 0.0;
  const double a[3] = {a0, a1, a2}, b[3] = {b0, b1, b2};
  rk_stage(y, inv_mass, F, stage, n_stages, a, Y, b, n, i, r);
}

}  // namespace adm
