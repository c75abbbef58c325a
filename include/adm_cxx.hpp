// C++ adapter over the C-ABI of adm.h, with the member names of the
// reference's operator and physics interfaces, so that host C++ written
// against them maps one call to one ABI call:
//   ThermalOperator  <- ThermalOperatorBase (source/ThermalOperatorBase.hh:
//                       14-85) as instantiated for MemorySpace::Default
//   ThermalPhysics   <- ThermalPhysicsInterface (source/
//                       ThermalPhysicsInterface.hh) for the explicit steppers
// Vectors are device pointers of n_dofs doubles in the structured node
// numbering of adm.h; streams are hipStream_t passed as void*. Errors throw
// std::runtime_error with adm_last_error().
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "adm.h"

namespace adm_cxx {

inline void check(int rc) {
  if (rc != ADM_OK) throw std::runtime_error(adm_last_error());
}

class ThermalOperator {
 public:
  explicit ThermalOperator(const adm_problem_desc &desc) {
    check(adm_create(&desc, &ctx_));
    check(adm_n_dofs(ctx_, &n_));
  }
  ~ThermalOperator() {
    if (ctx_) (void)adm_destroy(ctx_);
  }
  ThermalOperator(const ThermalOperator &) = delete;
  ThermalOperator &operator=(const ThermalOperator &) = delete;
  ThermalOperator(ThermalOperator &&o) noexcept
      : ctx_(std::exchange(o.ctx_, nullptr)), n_(o.n_) {}

  // ThermalOperatorBase::m / n: the operator is square on the nodes
  int64_t m() const { return n_; }
  int64_t n() const { return n_; }
  // vmult / vmult_add (source/ThermalOperator.templates.hh:176-225)
  void vmult(double *dst, const double *src, void *stream = nullptr) const {
    check(adm_vmult(ctx_, src, dst, 0, stream));
  }
  void vmult_add(double *dst, const double *src,
                 void *stream = nullptr) const {
    check(adm_vmult(ctx_, src, dst, 1, stream));
  }
  // compute_inverse_mass_matrix (source/ThermalOperator.templates.hh:
  // 118-164): the lumped inverse mass into inv_mass (device, n_dofs)
  void compute_inverse_mass_matrix(double *inv_mass,
                                   void *stream = nullptr) const {
    check(adm_compute_inverse_mass(ctx_, inv_mass, stream));
  }
  // set_time (source/ThermalOperator.hh:286-308)
  void set_time(double t) { check(adm_set_time(ctx_, t)); }
  // get/set_state_from/to_material_properties: per-cell ratios (host,
  // [n_cells][n_states])
  void get_state_from_material_properties(const double *cell_state) {
    check(adm_set_cell_state(ctx_, cell_state));
  }
  void set_state_to_material_properties(double *cell_state) const {
    check(adm_get_cell_state(ctx_, cell_state));
  }
  // set_material_deposition_orientation (host, n_cells each)
  void set_material_deposition_orientation(const double *cos_,
                                           const double *sin_) {
    check(adm_set_deposition(ctx_, cos_, sin_));
  }
  void *handle() const { return ctx_; }

 private:
  void *ctx_ = nullptr;
  int64_t n_ = 0;
};

class ThermalPhysics {
 public:
  ThermalPhysics(const adm_problem_desc &desc, int method)
      : op_(desc), method_(method) {}
  // evolve_one_time_step (source/ThermalPhysics.templates.hh:854-888) on
  // the device solution T; returns the new time
  double evolve_one_time_step(double t, double dt, double *T,
                              void *stream = nullptr) {
    double t_end = t;
    check(adm_evolve(op_.handle(), method_, t, dt, 1, T, 0, &t_end, stream));
    return t_end;
  }
  // update_physics_parameters (source/ThermalPhysics.templates.hh:827-852)
  void update_physics_parameters(int beam, double absorption_efficiency,
                                 double max_power, double radius,
                                 double depth) {
    check(adm_set_beam_parameters(op_.handle(), beam, absorption_efficiency,
                                  max_power, radius, depth));
  }
  ThermalOperator &get_thermal_operator() { return op_; }

 private:
  ThermalOperator op_;
  int method_;
};

}  // namespace adm_cxx
