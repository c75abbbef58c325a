/*
 * adm_oracle.h -- CPU restatement of adamantine's host ThermalOperator /
 * ThermalPhysics path (TEST INFRASTRUCTURE: the checker for the HIP library,
 * never part of the product path).
 *
 * Semantics are those of the host (MemorySpace::Host) classes with one SIMD
 * lane (VectorizedArray width 1): the powder update of
 * source/ThermalOperator.templates.hh:294 and the heat-source cut-off of
 * source/GoldakHeatSource.cc:129-139 act per quadrature point.
 */
#ifndef ADM_ORACLE_H
#define ADM_ORACLE_H

#include <stddef.h>

#include "../include/adm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle_t;

oracle_t *oracle_create(const adm_problem_desc *desc);
void oracle_destroy(oracle_t *o);
int64_t oracle_n_dofs(const oracle_t *o);
int64_t oracle_n_cells(const oracle_t *o);
int oracle_n_q(const oracle_t *o);
int oracle_n_face_q(const oracle_t *o);
void oracle_set_threads(int n);

/* ThermalOperator::compute_inverse_mass_matrix result (n_dofs; 0 on nodes of
 * no active cell). */
void oracle_inverse_mass(const oracle_t *o, double *out);
/* ThermalOperator::set_time */
void oracle_set_time(oracle_t *o, double t);
/* ThermalOperator::vmult (add = 0) / vmult_add (add = 1) */
void oracle_vmult(oracle_t *o, const double *src, double *dst, int add);
/* evaluate_thermal_physics: out = Minv * A(y) at time t */
void oracle_evaluate(oracle_t *o, double t, const double *y, double *out);
/* evolve_one_time_step repeated n_steps times; returns the final time */
double oracle_evolve(oracle_t *o, int method, double t, double dt, int n_steps,
                     double *T);

void oracle_get_qp_state(const oracle_t *o, double *liquid, double *powder);
void oracle_set_qp_state(oracle_t *o, const double *liquid,
                         const double *powder);
void oracle_get_face_powder(const oracle_t *o, double *face_powder);
void oracle_set_cell_state(oracle_t *o, const double *cell_state);
void oracle_get_cell_state(const oracle_t *o, double *cell_state);
void oracle_set_deposition(oracle_t *o, const double *c, const double *s);
void oracle_set_beam_parameters(oracle_t *o, int beam, double absorption,
                                double max_power, double radius, double depth);

/* ScanPath / HeatSource pieces, exposed for the reference's unit tests */
void oracle_beam_update(const adm_beam_desc *b, double t, int *old_segment,
                        adm_beam_state *out);
void oracle_scan_path_position(const adm_beam_desc *b, double t,
                               int *old_segment, double out[3]);
double oracle_scan_path_power_modifier(const adm_beam_desc *b, double t,
                                       int *old_segment);
double oracle_beam_value(const adm_beam_state *s, int dim, const double *pt);
/* Quaternion (source/Quaternion.cc) and ScanPath::get_quaternion */
void oracle_quaternion_mul(double a[4], const double b[4]);
void oracle_quaternion_div(double a[4], const double b[4]);
void oracle_quaternion_pow(double q[4], double e);
void oracle_quaternion_matrix(const double q[4], double R[9]);
void oracle_rotate(const double R[9], int inverse, const double p[3],
                   double out[3]);
void oracle_scan_path_quaternion(const adm_beam_desc *b, double t,
                                 int *old_segment, double q[4]);
/* MaterialProperty::compute_material_property (vectorized host form) */
double oracle_material_property(const adm_material_desc *m, int prop,
                                int material_id, const double *state_ratios,
                                double T);
double oracle_property_from_table(const adm_material_desc *m, int material_id,
                                  int prop, int state, double T);

/* 1-D tables used by the operator (n = fe_degree + 1 entries each) */
void oracle_basis_1d(const oracle_t *o, double *support, double *qpoints,
                     double *qweights, double *shape, double *coll_deriv,
                     double *gl_weights);

/* Material activation (source/material_deposition.cc:376-435,
 * source/ThermalPhysics.templates.hh:595-825). */
void oracle_cells_to_activate(const oracle_t *o, int n_boxes,
                              const double *boxes, int32_t *box_of_cell);
void oracle_add_material(oracle_t *o, const int32_t *box_of_cell,
                         const double *box_cos, const double *box_sin,
                         double new_temperature, double *T);

/* Ensemble draws (adm_oracle_da.c): get_normal_random_vector of
 * source/ensemble_management.cc:18-52 and the raw default-seeded MT19937. */
void oracle_normal_random_vector(uint32_t length, uint32_t n_rejected_draws,
                                 double mean, double stddev, double *out);
void oracle_mt19937(uint32_t n, uint32_t *out);

/* EnKF analysis (adm_oracle_da.c), source/DataAssimilator.cc. fn: 0 none,
 * 1 step_function, 2 gaspari_cohn. Ensembles are [N][n] member-major,
 * rows >= n_sim are augmented parameters. */
double oracle_gaspari_cohn(double r);
double oracle_localization(int fn, double d, double cutoff);
void oracle_sample_covariance(int N, int64_t n, int64_t n_sim,
                              const double *ens, const double *coords, int dim,
                              int fn, double cutoff, double *P);
void oracle_apply_kalman_gain(int N, int64_t n, int64_t n_sim,
                              const double *ens, const double *coords, int dim,
                              int fn, double cutoff, int m,
                              const int64_t *obs_dof, const double *R,
                              const double *innov, double *out);
void oracle_update_ensemble(int N, int64_t n, int64_t n_sim, double *ens,
                            const double *coords, int dim, int fn,
                            double cutoff, int m, const int64_t *obs_dof,
                            const double *y, const double *R,
                            const double *noise);
size_t oracle_da_rng_size(void);
void oracle_fill_noise(void *state, int count, int m, const double *R,
                       double *out);

#ifdef __cplusplus
}
#endif

#endif
