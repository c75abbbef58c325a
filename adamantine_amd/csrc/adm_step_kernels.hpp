// Vector kernels of the explicit time steppers (non-template; compiled in
// the API translation unit only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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

// Forward Euler (source/ThermalPhysics.templates.hh:871-876):
// T = 1*T + dt*(R*Minv); the residual buffer is cleared for the next stage.
__global__ void fe_update_kernel(double *__restrict__ T, double *__restrict__ R,
                                 const double *__restrict__ inv_mass, double dt,
                                 int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = R[i] * inv_mass[i];
  T[i] = 1. * T[i] + dt * f;
  R[i] = 0.0;
}

// Classic RK4 stage of deal.II's ExplicitRungeKutta with the stage vector
// f_i = R*Minv: acc = acc + (dt b_i) f_i (acc starts as y); next stage input
// Y = y + (dt a_{i+1,i}) f_i (the other a_{i+1,j} of RK4 are zero, adding
// +-0); the last stage writes y = acc.
__global__ void rk4_stage_kernel(double *__restrict__ y, double *__restrict__ R,
                                 const double *__restrict__ inv_mass,
                                 double *__restrict__ acc,
                                 double *__restrict__ Y, double bcoef,
                                 double acoef, int first, int last,
                                 int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = R[i] * inv_mass[i];
  const double yi = y[i];
  double s = first ? yi : acc[i];
  s = 1. * s + bcoef * f;
  if (last) {
    y[i] = s;
  } else {
    acc[i] = s;
    Y[i] = 1. * yi + acoef * f;
  }
  R[i] = 0.0;
}

// Generic explicit RK stage (used for RK3, whose a_{3,1} != 0): stores
// f_i and forms the next stage input from all stored stages.
__global__ void rk_stage_kernel(double *__restrict__ y, double *__restrict__ R,
                                const double *__restrict__ inv_mass,
                                double *__restrict__ F, int stage, int n_stages,
                                double a0, double a1, double a2,
                                double *__restrict__ Y, double b0, double b1,
                                double b2, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = R[i] * inv_mass[i];
  F[(int64_t)stage * n + i] = f;
  R[i] = 0.0;
  if (stage + 1 < n_stages) {
    double s = y[i];
    const double a[3] = {a0, a1, a2};
    for (int j = 0; j <= stage; ++j) s = 1. * s + a[j] * F[(int64_t)j * n + i];
    Y[i] = s;
  } else {
    const double b[3] = {b0, b1, b2};
    double s = y[i];
    for (int j = 0; j < n_stages; ++j) s = 1. * s + b[j] * F[(int64_t)j * n + i];
    y[i] = s;
  }
}

}  // namespace adm
