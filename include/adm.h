/*
 * adm.h -- C-ABI of the MI355X-native matrix-free heat-equation path.
 *
 * This header is the drop-in boundary that replaces the *Dev instantiations of
 * adamantine's ThermalOperatorDevice / ThermalPhysics (memory_space device).
 * Every entry point below names the reference interface it replaces
 * (paths relative to the adamantine source tree).
 *
 * Conventions
 *   - every function returns int: 0 = success, non-zero = error code;
 *     adm_last_error() returns a message for the calling thread's last error
 *     (the reference throws std::runtime_error through ASSERT_THROW,
 *     source/utils.hh:70-74; no C++ exception crosses this boundary);
 *   - device buffers are caller-owned (hipMalloc / torch) and passed as plain
 *     pointers; host buffers are marked "host";
 *   - work is enqueued on the hipStream_t passed as `stream` (NULL = default
 *     stream) and is stream-ordered; functions that return data to host
 *     buffers synchronise the stream before returning;
 *   - one context per (mesh, FE degree, material, sources, boundary) like the
 *     reference's ThermalOperator; a context is not thread-safe.
 *
 * DoF numbering: the mesh is the structured box of
 * dealii::GridGenerator::subdivided_hyper_rectangle (source/Geometry.cc:252).
 * Nodes of the FE_Q(p) space (Gauss-Lobatto support points) are numbered
 * lexicographically, x fastest:  node = ix + Nx*(iy + Ny*iz),
 * N_d = p*n_cells_d + 1.  Cells are numbered c = cx + nx*(cy + ny*cz).
 * A deal.II DoF maps to node (round((x - x0)/h_x * p), ...) of its support
 * point, which gives the permutation to deal.II's DoF order.
 */
#ifndef ADM_H
#define ADM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ADM_MAX_MATERIALS 4
#define ADM_TABLE_SIZE 12     /* MaterialProperty::table_size */
#define ADM_N_STATE_PROPS 8   /* thermal StateProperty entries, source/types.hh:40-64 */
#define ADM_N_PROPS 5         /* Property entries, source/types.hh:88-96 */
#define ADM_MAX_STATES 3
#define ADM_MAX_POLY 5        /* p_order <= 4, source/MaterialProperty.hh:543 */
#define ADM_MAX_BEAMS 4
#define ADM_MAX_BOUNDARY_IDS 8

/* StateProperty (source/types.hh:40-49) */
enum adm_state_property {
  ADM_DENSITY = 0,
  ADM_SPECIFIC_HEAT = 1,
  ADM_THERMAL_CONDUCTIVITY_X = 2,
  ADM_THERMAL_CONDUCTIVITY_Y = 3,
  ADM_THERMAL_CONDUCTIVITY_Z = 4,
  ADM_EMISSIVITY = 5,
  ADM_RADIATION_HEAT_TRANSFER_COEF = 6,
  ADM_CONVECTION_HEAT_TRANSFER_COEF = 7
};

/* Property (source/types.hh:88-96) */
enum adm_property {
  ADM_LIQUIDUS = 0,
  ADM_SOLIDUS = 1,
  ADM_LATENT_HEAT = 2,
  ADM_RADIATION_TEMPERATURE_INFTY = 3,
  ADM_CONVECTION_TEMPERATURE_INFTY = 4
};

/* BoundaryType bits (source/Boundary.hh:21-29) */
enum adm_boundary_type {
  ADM_BC_ADIABATIC = 0x1,
  ADM_BC_RADIATIVE = 0x2,
  ADM_BC_CONVECTIVE = 0x4
};

/* heat source kinds (source/ThermalPhysics.templates.hh:390-410) */
enum adm_source_type {
  ADM_SOURCE_GOLDAK = 0,
  ADM_SOURCE_ELECTRON_BEAM = 1,
  ADM_SOURCE_CUBE = 2
};

/* explicit time steppers (source/ThermalPhysics.templates.hh:456-471) */
enum adm_time_stepping {
  ADM_FORWARD_EULER = 0,
  ADM_RK_THIRD_ORDER = 1,
  ADM_RK_FOURTH_ORDER = 2
};

/* error codes */
enum adm_status {
  ADM_OK = 0,
  ADM_ERR_INVALID = 1,   /* invalid argument or descriptor */
  ADM_ERR_HIP = 2,       /* HIP runtime error */
  ADM_ERR_UNSUPPORTED = 3,
  ADM_ERR_STATE = 4      /* call out of order (e.g. vmult before create) */
};

/* Material model: the data of MaterialProperty after fill_properties
 * (source/MaterialProperty.templates.hh:796-1119), i.e. tables padded to 12
 * entries with the last pair, missing tables zero, missing state-independent
 * properties DBL_MAX. */
typedef struct adm_material_desc {
  int32_t n_materials;   /* runtime number of materials (1..4) */
  int32_t n_states;      /* 1 Solid, 2 SolidLiquid, 3 SolidLiquidPowder */
  int32_t use_table;     /* property_format table (1) or polynomial (0) */
  int32_t p_order;       /* polynomial order template parameter (0..4) */
  int32_t single_material_layout; /* 1: the reference's n_materials == 1
                                      instantiation (power-major polynomial
                                      sum, MaterialProperty.hh:572-777);
                                      0: n_materials == -1 (state-major,
                                      MaterialProperty.hh:778-796) */
  int32_t pad_;
  double table[ADM_MAX_MATERIALS][ADM_N_STATE_PROPS][ADM_MAX_STATES]
              [ADM_TABLE_SIZE][2];
  double poly[ADM_MAX_MATERIALS][ADM_N_STATE_PROPS][ADM_MAX_STATES]
             [ADM_MAX_POLY];
  double prop[ADM_MAX_MATERIALS][ADM_N_PROPS];
} adm_material_desc;

/* Structured box mesh (source/Geometry.cc:186-260) and FE space
 * (source/ThermalPhysics.templates.hh:364-369). */
typedef struct adm_mesh_desc {
  int32_t dim;           /* 3, or 2: axes x (length) and z (height); the
                            n_cells, origin and extent entries [1] are then
                            unused, boundary ids are 0..3 and the beam
                            centre's components 0 and 1 are (x, height), as
                            in the reference's dim = 2 instantiations
                            (source/instantiation.hh:121-179) */
  int32_t fe_degree;     /* 1..5 */
  int32_t quadrature;    /* 0 = QGauss(p+1), 1 = QGaussLobatto(p+1) */
  int32_t n_cells[3];    /* length_divisions, width_divisions, height_divisions */
  double origin[3];
  double extent[3];      /* length, width, height */
} adm_mesh_desc;

/* One scan-path segment (source/ScanPath.hh:40-48); end_time is absolute.
 * end_rotation is the unit quaternion (r, i, j, k) of a five-axis event
 * series (source/ScanPath.cc:180-187), all zero on a three-axis path. A path
 * is five-axis when its first segment carries a rotation; then every segment
 * must carry one with |q|^2 within 1e-10 of 1 (Quaternion::reinit,
 * source/Quaternion.cc:44-53). */
typedef struct adm_segment {
  double end_time;
  double power_modifier;
  double end_point[3];
  double end_rotation[4];
} adm_segment;

/* One heat source (source/BeamHeatSourceProperties.hh:55-66,
 * source/CubeHeatSource.cc:31-42). */
typedef struct adm_beam_desc {
  int32_t type;          /* adm_source_type */
  int32_t n_segments;
  double absorption_efficiency;
  double max_power;
  double radius;         /* diameter / 2 */
  double depth;
  /* cube source */
  double start_time, end_time, value;
  double min_point[3], max_point[3];
  const adm_segment *segments;  /* host; n_segments entries */
} adm_beam_desc;

/* A beam evaluated at one time (HeatSource::update_time,
 * source/GoldakHeatSource.cc:24-44, source/ElectronBeamHeatSource.cc:24-45). */
typedef struct adm_beam_state {
  int32_t type;
  int32_t on;            /* power modifier > 0 (is_source_on) */
  double center[3];
  double alpha;
  double depth;
  double inv_depth;
  double inv_depth_squared;
  double radius_squared;
  double inv_radius_squared;
  double cube_value;
  double cube_min[3];
  double cube_max[3];
  /* five-axis path: the source is evaluated at rotation * point
   * (source/GoldakHeatSource.cc:46-57), rotation being the row-major matrix
   * of ScanPath::get_quaternion(t) (source/ScanPath.cc:290-341,
   * source/Quaternion.cc:133-150); rotated = 0 on a three-axis path */
  int32_t rotated;
  int32_t pad_;
  double rotation[9];
} adm_beam_state;

/* Everything the reference's ThermalPhysics/ThermalOperator constructors and
 * setup() gather from Geometry, Boundary, MaterialProperty and the sources.
 * Per-cell arrays are host arrays in cell order, NULL meaning the default. */
typedef struct adm_problem_desc {
  adm_mesh_desc mesh;
  adm_material_desc material;
  int32_t n_boundary_ids;                       /* 2*dim for a box */
  int32_t boundary_type[ADM_MAX_BOUNDARY_IDS];  /* Boundary::_types */
  int32_t n_beams;
  int32_t pad_;
  adm_beam_desc beams[ADM_MAX_BEAMS];
  const uint8_t *cell_active;     /* 1 = FE_Q, 0 = FE_Nothing; NULL: all 1 */
  const uint8_t *cell_material;   /* material_id; NULL: 0 */
  const double *cell_state;       /* [n_cells][n_states] ratios; NULL: solid */
  const double *deposition_cos;   /* NULL: 1 */
  const double *deposition_sin;   /* NULL: 0 */
} adm_problem_desc;

/* ------------------------------------------------------------------------ */
/* Context life cycle                                                        */
/* ------------------------------------------------------------------------ */

/* Replaces the ThermalOperatorDevice constructor + reinit() +
 * set_material_deposition_orientation() + get_state_from_material_properties()
 * as called by ThermalPhysics::setup() (source/ThermalPhysics.templates.hh:
 * 496-524, source/ThermalOperatorDevice.templates.hh:520-579). Uses the
 * current HIP device. */
int adm_create(const adm_problem_desc *desc, void **ctx);

/* Replaces ThermalOperatorDevice::clear() + destruction
 * (source/ThermalOperatorDevice.templates.hh:631-640). */
int adm_destroy(void *ctx);

/* Message of the calling thread's last error (ASSERT_THROW text). */
const char *adm_last_error(void);

/* m() / n() (source/ThermalOperatorBase.hh:26-32): the number of DoFs of
 * the caller's numbering when a DoF map is set (adm_set_dof_map), else the
 * number of nodes of the structured numbering (nodes of inactive cells
 * included; they are inert). */
int adm_n_dofs(const void *ctx, int64_t *n_dofs);

/* Number of nodes of the structured box numbering, (p n_x + 1)(p n_y + 1)
 * (p n_z + 1), whatever the DoF map. */
int adm_n_nodes(const void *ctx, int64_t *n_nodes);

/* DoF map of the caller's vectors, for a drop-in behind deal.II's
 * numbering (ThermalOperatorBase::vmult takes LA::distributed::Vector in
 * the DoFHandler's order, source/ThermalOperatorBase.hh:26-55; DoFs are
 * distributed by ThermalPhysics::setup_dofs,
 * source/ThermalPhysics.templates.hh:508-524, with FE_Nothing on inactive
 * cells, so n_dofs counts only the nodes of active cells). dof_to_node
 * (host, n_dofs entries): DoF i is box node dof_to_node[i] (x fastest).
 * The map must be injective and cover every node of every active cell;
 * n_dofs < 0 clears it. While a map is set, adm_vmult,
 * adm_compute_inverse_mass and adm_evolve take and return vectors of
 * n_dofs entries in the caller's order (one gather into box nodes before,
 * one scatter after, per call); adm_add_material refuses (activation
 * renumbers the DoFs: clear the map, activate on the box-node vector, set
 * the new map); every other entry point keeps box-node indexing. */
int adm_set_dof_map(void *ctx, int64_t n_dofs, const int64_t *dof_to_node);

/* With a DoF map set: box-node vector (adm_n_nodes entries) from a vector
 * in the caller's order (entries of unmapped nodes left as they are), and
 * back (ext = node values of the mapped nodes, or += with add). Device
 * vectors, stream-ordered. */
int adm_dofs_to_nodes(void *ctx, const double *ext, double *node,
                      void *stream);
int adm_nodes_to_dofs(void *ctx, const double *node, double *ext, int add,
                      void *stream);

/* Number of cells and of quadrature points per cell. */
int adm_sizes(const void *ctx, int64_t *n_cells, int32_t *n_q_per_cell,
              int32_t *n_face_q);

/* ------------------------------------------------------------------------ */
/* Operator                                                                  */
/* ------------------------------------------------------------------------ */

/* compute_inverse_mass_matrix() + get_inverse_mass_matrix()
 * (source/ThermalOperator.templates.hh:118-164): Gauss-Lobatto lumped mass
 * over active cells, inverted. Nodes touched by no active cell get 0 (the
 * guard of source/ThermalOperatorDevice.templates.hh:624-627).
 * inv_mass: device, n_dofs doubles. */
int adm_compute_inverse_mass(void *ctx, double *inv_mass, void *stream);

/* set_time(t) (source/ThermalOperator.hh:298-308): evaluates every beam at t
 * for the next vmult. */
int adm_set_time(void *ctx, double t);

/* vmult / vmult_add (source/ThermalOperator.templates.hh:176-225):
 * dst = A(src) (add = 0) or dst += A(src) (add = 1), with A the host
 * ThermalOperator's cell term (diffusion + heat sources, per-quadrature-point
 * 1/(rho c_p)) plus its face term on non-adiabatic boundaries of the active
 * domain. Mutates the per-quadrature-point liquid/powder state like the
 * reference's cell_local_apply. src, dst: device, n_dofs doubles. */
int adm_vmult(void *ctx, const double *src, double *dst, int add,
              void *stream);

/* adm_vmult that allocates nothing. No counterpart in the reference: it is
 * for callers that cannot allocate device memory during the call (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). adm_vmult may grow the plane kernel's patch buffer with
 * hipMalloc on its first call or after an activation; this entry point
 * uses instead a static device array of the library, 40 MiB per device
 * (5 * 2^20 doubles), held until the process ends, and refuses
 * (ADM_ERR_UNSUPPORTED) a mesh whose patch buffer needs more (the
 * 128x64x32 Q2 workload needs 4.2 million doubles). Every call on one
 * device shares that array: calls that overlap on one device (several
 * streams, several host threads) race; calls in one stream, one after the
 * other, are safe (the array needs no clearing: every slot a call reads
 * it writes first). Same arguments and results as adm_vmult. */
int adm_vmult_noalloc(void *ctx, const double *src, double *dst, int add,
                      void *stream);

/* Cell-kernel launch modes (no reference counterpart; the reference's
 * matrix-free loop has no such choice). */
#define ADM_LAUNCH_LINE_ATOMIC 0  /* line kernel, one launch, FP64 atomics */
#define ADM_LAUNCH_LINE_COLORED 1 /* line kernel, 8 parity-colour launches
                                     with plain read-add-write */
#define ADM_LAUNCH_PLANE 2        /* plane kernel (FE degree 1 or 2 only):
                                     one launch, shared nodes summed in the
                                     wavefront, then atomics */
#define ADM_N_LAUNCH_CONFIGS 9    /* 3 modes x 2..4 wavefronts per block */

/* Launch configuration of the cell kernel: one of the ADM_LAUNCH_* modes and
 * 2..4 wavefronts per block (default: the plane kernel for FE degrees 1 and
 * 2, the colour launches of the line kernel otherwise, 4 wavefronts). The
 * plane mode is refused for FE degrees above 2. */
int adm_set_launch_config(void *ctx, int32_t mode, int32_t waves_per_block);

/* Autotuner (no reference counterpart): times `reps` vmults of src into dst
 * (device vectors) under each launch configuration (modes 1, 0, 2, each
 * with 2, 3, 4 wavefronts per block; ms_out, host, may be NULL, receives
 * the ADM_N_LAUNCH_CONFIGS per-vmult times in that order, -1 for the plane
 * mode where the FE degree has none), keeps the fastest and returns it.
 * The phase state is restored afterwards, so later results do not depend
 * on the tuning. */
int adm_autotune(void *ctx, const double *src, double *dst, int32_t reps,
                 void *stream, int32_t *mode_out, int32_t *waves_out,
                 double *ms_out);

/* Repeated evolve_one_time_step (source/ThermalPhysics.templates.hh:854-888):
 * n_steps steps of `method` from time t with step dt on the device solution
 * vector T (n_dofs doubles, updated in place). use_graph = 1 captures one
 * step as a HIP graph and replays it (the beam states of every stage are
 * evaluated on the host beforehand and read from a device table); the
 * result equals the direct launches'. t_end (host, may be NULL) receives
 * the final time. */
int adm_evolve(void *ctx, int method, double t, double dt, int n_steps,
               double *T, int use_graph, double *t_end, void *stream);

/* ------------------------------------------------------------------------ */
/* State                                                                     */
/* ------------------------------------------------------------------------ */

/* get_state_from_material_properties (source/ThermalOperator.templates.hh:
 * 753-884): set every quadrature point (cells and boundary faces) from the
 * per-cell state ratios (host, [n_cells][n_states]). */
int adm_set_cell_state(void *ctx, const double *cell_state);

/* set_state_to_material_properties (source/MaterialProperty.templates.hh:
 * 537-733): cell averages of the quadrature-point ratios
 * (host out, [n_cells][n_states]; inactive cells get zeros). */
int adm_get_cell_state(void *ctx, double *cell_state);

/* Per-quadrature-point liquid and powder ratios (host, [n_cells][n_q];
 * either pointer may be NULL). An inactive cell keeps the state it was given
 * until it is activated; powder reads as zero in cells that hold none. */
int adm_get_qp_state(void *ctx, double *liquid, double *powder);
int adm_set_qp_state(void *ctx, const double *liquid, const double *powder);

/* Per-face-quadrature-point powder ratio of the boundary faces
 * (_face_powder_ratio of source/ThermalOperator.hh:225), host out
 * [n_cells][2 dim][n_face_q]; faces without a boundary term read as zero. */
int adm_get_face_state(void *ctx, double *face_powder);
/* Sets the face powder ratios from the same layout (the entries of faces
 * without a boundary term are ignored): the restart counterpart of
 * adm_get_face_state for checkpoints that keep the face history
 * (the reference re-derives it from the cell state,
 * source/ThermalPhysics.templates.hh:1084-1185). */
int adm_set_face_state(void *ctx, const double *face_powder);

/* set_material_deposition_orientation (source/ThermalOperator.templates.hh:
 * 897-940): per-cell cos/sin (host, n_cells each). */
int adm_set_deposition(void *ctx, const double *cos_, const double *sin_);

/* The per-cell data ThermalPhysics::save_checkpoint serialises besides the
 * state (source/ThermalPhysics.templates.hh:1093-1147): the active flag
 * (fe index 0) and the deposition cos/sin, as they stand after every
 * activation so far (host out, n_cells each; any pointer may be NULL). */
int adm_get_cell_data(void *ctx, uint8_t *active, double *cos_, double *sin_);

/* Beam parameter update (ThermalPhysics::update_physics_parameters,
 * source/ThermalPhysics.templates.hh:827-852): new absorption efficiency,
 * max power, radius and depth for beam `beam`. */
int adm_set_beam_parameters(void *ctx, int beam, double absorption_efficiency,
                            double max_power, double radius, double depth);

/* Evaluate beam `beam` of the context at time t on the host, with the
 * forward-only segment search of ScanPath (source/ScanPath.cc:197-275). */
int adm_beam_state_at(void *ctx, int beam, double t, adm_beam_state *out);

/* ------------------------------------------------------------------------ */
/* Material deposition (cell activation)                                     */
/* ------------------------------------------------------------------------ */

/* get_elements_to_activate (source/material_deposition.cc:376-435): for every
 * inactive cell, the index of the last deposition box (input order) whose
 * closed box meets the cell's bounding box, or -1.
 * boxes: host, n_boxes x {min x, min y, min z, max x, max y, max z};
 * box_of_cell: host out, n_cells. */
int adm_cells_to_activate(void *ctx, int32_t n_boxes, const double *boxes,
                          int32_t *box_of_cell);

/* add_material_start + add_material_end (source/ThermalPhysics.templates.hh:
 * 595-825): cells with box_of_cell >= 0 become active with that box's
 * deposition cos/sin (host, per box); the phase state of every cell goes to
 * its cell mean and back to every quadrature point, as the reference's
 * state transfer does; nodes that no previously active cell held take
 * new_temperature in T (device, n_dofs; may be NULL), the others keep their
 * value; the lumped mass used by adm_evolve is recomputed. */
int adm_add_material(void *ctx, const int32_t *box_of_cell,
                     const double *box_cos, const double *box_sin,
                     double new_temperature, double *T, void *stream);

/* get_elements_to_activate followed by add_material_start + add_material_end
 * (source/material_deposition.cc:376-435, source/ThermalPhysics.templates.hh:
 * 595-825) in one call, on the device: the same cells, state transfer,
 * new-material temperature and mass as adm_cells_to_activate +
 * adm_add_material, without copying the phase state to the host.
 * boxes: host, n_boxes x {min x, y, z, max x, y, z}; box_cos, box_sin: host,
 * n_boxes; T: device, n_dofs (may be NULL); box_of_cell: host out, n_cells
 * (may be NULL): the box that activated each cell, or -1. Enqueued on
 * stream; returns after one small device-to-host copy. A 2-D context takes
 * the boxes' x and z bounds and runs the host path of
 * adm_cells_to_activate + adm_add_material (synchronous). */
int adm_activate(void *ctx, int32_t n_boxes, const double *boxes,
                 const double *box_cos, const double *box_sin,
                 double new_temperature, double *T, int32_t *box_of_cell,
                 void *stream);

/* ------------------------------------------------------------------------ */
/* Ensemble                                                                  */
/* ------------------------------------------------------------------------ */

/* get_normal_random_vector (source/ensemble_management.cc:18-52): `length`
 * draws of N(mean, stddev) from a default-seeded std::mt19937 after
 * discarding n_rejected_draws draws; negative draws are redrawn. out: host.
 * create_database_ensemble (:213-265) calls it once per perturbed key, the
 * k-th key (in input order) discarding first_local_member + k*n_members. */
int adm_normal_random_vector(uint32_t length, uint32_t n_rejected_draws,
                             double mean, double stddev, double *out);

/* What one ensemble member changes in the base problem. The reference builds
 * one database per member by drawing each `ensemble.<key>_stddev` key
 * (source/ensemble_management.cc:166-265) and one ThermalPhysics per member
 * (application/adamantine.hh:1961-2060); these are the keys that can vary
 * here. Every member shares the mesh, FE space, material tables (up to the
 * conductivity scale), boundary types, beam kinds, radii, depths and scan
 * paths of the base problem. */
typedef struct adm_member_desc {
  double absorption_efficiency[ADM_MAX_BEAMS]; /* beam_<b>.absorption_efficiency */
  double max_power[ADM_MAX_BEAMS];             /* beam_<b>.max_power */
  /* multiplies every thermal conductivity table value / polynomial
   * coefficient of every material and state (1 = unchanged) */
  double conductivity_scale;
  const uint8_t *cell_active;  /* host [n_cells]; NULL: the base problem's */
  const double *cell_state;    /* host [n_cells][n_states]; NULL: base */
} adm_member_desc;

/* Creates an ensemble of n_members operators/physics over one mesh. Member
 * vectors are stored member-major: member j's nodes are
 * [j*n_dofs, (j+1)*n_dofs) of every ensemble vector. Replaces the
 * per-member ThermalPhysics objects of run_ensemble
 * (application/adamantine.hh:1961-2060). */
int adm_ensemble_create(const adm_problem_desc *base, int32_t n_members,
                        const adm_member_desc *members, void **ens);
int adm_ensemble_destroy(void *ens);

/* n_dofs per member (identical for all members). */
int adm_ensemble_n_dofs(const void *ens, int64_t *n_dofs);

/* Every member's compute_inverse_mass_matrix (see adm_compute_inverse_mass);
 * inv_mass: device, n_members*n_dofs doubles. */
int adm_ensemble_compute_inverse_mass(void *ens, double *inv_mass,
                                      void *stream);

/* set_time(t) for every member. */
int adm_ensemble_set_time(void *ens, double t);

/* vmult / vmult_add of every member in one launch per kernel:
 * dst_j = A_j(src_j) (+ dst_j). src, dst: device, n_members*n_dofs. */
int adm_ensemble_vmult(void *ens, const double *src, double *dst, int add,
                       void *stream);

/* adm_set_launch_config and adm_autotune for the batched launches of an
 * ensemble (no reference counterpart); src, dst: n_members * n_dofs. */
int adm_ensemble_set_launch_config(void *ens, int32_t mode,
                                   int32_t waves_per_block);
int adm_ensemble_autotune(void *ens, const double *src, double *dst,
                          int32_t reps, void *stream, int32_t *mode_out,
                          int32_t *waves_out, double *ms_out);

/* evolve_one_time_step of every member, n_steps times, batched. This
 * replaces the serial member loop of run_ensemble
 * (application/adamantine.hh:2137-2142). T: device, n_members*n_dofs. */
int adm_ensemble_evolve(void *ens, int method, double t, double dt,
                        int n_steps, double *T, double *t_end, void *stream);

/* adm_ensemble_evolve as replays of one captured step (a HIP graph): the
 * same kernels and results, the beams' states and the members' shared
 * power modifiers of every stage read from tables the host fills before
 * the capture. Same replacement (application/adamantine.hh:2137-2142). */
int adm_ensemble_evolve_graph(void *ens, int method, double t, double dt,
                              int n_steps, double *T, double *t_end,
                              void *stream);

/* Member `member`'s per-quadrature-point liquid and powder ratios (host,
 * [n_cells][n_q]; either pointer may be NULL). */
int adm_ensemble_get_qp_state(void *ens, int32_t member, double *liquid,
                              double *powder);
/* The member's counterparts of adm_set_qp_state, adm_get_face_state and
 * adm_set_face_state (checkpoint and restart of an ensemble with its
 * quadrature-point history). */
int adm_ensemble_set_qp_state(void *ens, int32_t member, const double *liquid,
                              const double *powder);
int adm_ensemble_get_face_state(void *ens, int32_t member,
                                double *face_powder);
int adm_ensemble_set_face_state(void *ens, int32_t member,
                                const double *face_powder);

/* update_physics_parameters of one member (source/ThermalPhysics.templates.hh:
 * 827-852) for beam `beam`'s absorption efficiency and max power, e.g. after
 * an analysis of the augmented state (application/adamantine.hh:2306-2361).
 * Radius, depth and scan path are shared by the members of an ensemble. */
int adm_ensemble_set_beam_gain(void *ens, int32_t member, int32_t beam,
                               double absorption_efficiency, double max_power);

/* Material deposition for one member (see adm_cells_to_activate and
 * adm_add_material; T is the whole ensemble vector, or NULL). Replaces the
 * per-member activation of run_ensemble (application/adamantine.hh:
 * 2070-2115). */
int adm_ensemble_cells_to_activate(void *ens, int32_t member, int32_t n_boxes,
                                   const double *boxes, int32_t *box_of_cell);
int adm_ensemble_add_material(void *ens, int32_t member,
                              const int32_t *box_of_cell,
                              const double *box_cos, const double *box_sin,
                              double new_temperature, double *T, void *stream);
/* adm_activate for one member (T: the whole ensemble vector, or NULL). */
int adm_ensemble_activate(void *ens, int32_t member, int32_t n_boxes,
                          const double *boxes, const double *box_cos,
                          const double *box_sin, double new_temperature,
                          double *T, int32_t *box_of_cell, void *stream);

/* ------------------------------------------------------------------------ */
/* Data assimilation (EnKF)                                                  */
/* ------------------------------------------------------------------------ */

/* DataAssimilator localization_cutoff_function (source/DataAssimilator.cc:
 * 70-90) */
enum adm_localization {
  ADM_LOCALIZATION_NONE = 0,
  ADM_LOCALIZATION_STEP_FUNCTION = 1,
  ADM_LOCALIZATION_GASPARI_COHN = 2
};

/* The observation-noise stream of a DataAssimilator: its default-seeded
 * std::mt19937 _prng and N(0, 1) _normal_dist_generator
 * (source/DataAssimilator.hh:280-286), whose state persists across calls. */
int adm_da_create(void **da);
int adm_da_destroy(void *da);

/* fill_noise_vector (source/DataAssimilator.cc:630-670) for `count`
 * consecutive members: out[c][i] (host) = sqrt(R_ii) z_i for diagonal R,
 * else (L z)_i with R = L L^T. R: host, n_obs x n_obs row-major. */
int adm_da_fill_noise(void *da, int32_t count, int32_t n_obs, const double *R,
                      double *out);

/* get_expt_to_dof_mapping (source/experimental_data_utils.cc:70-110): the
 * nearest DoF support point, among nodes of active cells of `member`, of
 * each point (host points[n_points][3] -> host dofs[n_points]). Equidistant
 * candidates resolve to the lowest node index. */
int adm_ensemble_nearest_dofs(void *ens, int32_t member, int64_t n_points,
                              const double *points, int64_t *dofs);

/* get_expt_to_dof_mapping (source/experimental_data_utils.cc:70-110) for one
 * operator: the nearest support point, among nodes of active cells, of each
 * point (host points[n_points][3] -> host dofs[n_points], -1 if no cell is
 * active). Equidistant candidates resolve to the lowest node index. */
int adm_nearest_dofs(void *ctx, int64_t n_points, const double *points,
                     int64_t *dofs);

/* set_with_experimental_data (source/experimental_data_utils.cc:112-137):
 * T[dofs[k]] = values[k] in input order (a repeated DoF keeps its last
 * value). dofs, values: host, n entries; T: device. */
int adm_set_values(void *ctx, int64_t n, const int64_t *dofs,
                   const double *values, double *T, void *stream);

/* DataAssimilator::update_ensemble (source/DataAssimilator.cc:95-189) on the
 * augmented state [T; parameters]: with d_j = noise_j + (obs - H x_j) and P
 * the localised sample covariance (calc_sample_covariance_sparse, :690-752),
 * x_j += P H^T (H P H^T + R)^{-1} d_j (apply_kalman_gain, :403-476).
 *   T       device, n_members*n_dofs (member-major), updated in place
 *   params  host, n_members*n_params (member-major), updated in place
 *   obs_dof host, n_obs DoFs observed (calc_H, :478-502)
 *   obs     host, n_obs values;  R host, n_obs x n_obs row-major
 *   noise   host, n_members*n_obs perturbations (adm_da_fill_noise) or NULL
 *   localization adm_localization, cutoff its distance (DBL_MAX: none).
 * Parameter rows are never localised, as in the reference. The inverse is
 * a direct solve (the reference iterates GMRES to its tolerance). */
int adm_ensemble_analysis(void *ens, double *T, int32_t n_params,
                          double *params, int32_t n_obs,
                          const int64_t *obs_dof, const double *obs,
                          const double *R, const double *noise,
                          int32_t localization, double cutoff, void *stream);

/* The ensemble-space update kernel by itself: T <- T Tm, i.e. member j
 * becomes sum_l Tm[l][j] T_l (Tm host, n_members x n_members row-major).
 * No reference counterpart; adm_ensemble_analysis applies it with
 * Tm = I + W. */
int adm_ensemble_transform(void *ens, double *T, const double *Tm,
                           void *stream);

/* ------------------------------------------------------------------------ */
/* Discrete adjoint                                                          */
/* ------------------------------------------------------------------------ */

/* No reference counterpart (the reference estimates parameters only through
 * the EnKF). Runs n_steps of `method` (ADM_FORWARD_EULER, ADM_RK_THIRD_ORDER
 * or ADM_RK_FOURTH_ORDER) from T0 at t0, exactly as adm_evolve does, and the
 * transposed stage sequence backwards from checkpoints taken every
 * checkpoint_every steps (<= 0: one checkpoint), for the misfit
 *   J = 1/2 sum_k w_k (T(t_end)[obs_dof_k] - obs_k)^2.
 *   T0, T_final, lambda0  device, n_dofs (T_final, lambda0 may be NULL)
 *   obs_dof, obs, obs_weight  host, n_obs (obs_weight NULL: all ones)
 *   J    host out: the misfit
 *   grad host out[3]: dJ/d(absorption efficiency of beam 0),
 *        dJ/ds for every thermal conductivity scaled by s at s = 1,
 *        dJ/d(radius of beam 0)
 *   lambda0 device out: dJ/dT0.
 * The powder ratio carried between operator evaluations is differentiated
 * exactly (its own adjoint per quadrature point); at return the context's
 * phase state is the one at the final time, as after adm_evolve. Polynomial
 * or tabulated properties with the merged-table form (not the general-table
 * evaluator). Three-dimensional contexts only (a 2-D one: ADM_ERR_UNSUPPORTED,
 * as for adm_ensemble_adjoint_gradient and adm_operator_vjp). */
int adm_adjoint_gradient(void *ctx, int method, double t0, double dt,
                         int n_steps, int checkpoint_every, const double *T0,
                         int32_t n_obs, const int64_t *obs_dof,
                         const double *obs, const double *obs_weight,
                         double *T_final, double *J, double *grad,
                         double *lambda0, void *stream);

/* A material activation inside an adjoint window: before step `step`
 * (0 .. n_steps-1) of the window, member `member` runs adm_activate with
 * these boxes (host arrays). */
typedef struct adm_activation_event {
  int32_t member;
  int32_t step;
  int32_t n_boxes;
  const double *boxes;     /* n_boxes x {min x, y, z, max x, y, z} */
  const double *box_cos;   /* n_boxes */
  const double *box_sin;   /* n_boxes */
  double new_temperature;
} adm_activation_event;

/* adm_adjoint_gradient for every member of an ensemble at once (no
 * reference counterpart; the reference's ensemble loop is
 * application/adamantine.hh:1961-2433): one batched forward sweep (the
 * members' steps in one launch per stage, activations per member as listed
 * in events), one batched backward sweep (the transposed cell and face
 * operators of all members in one launch per stage), checkpoints every
 * checkpoint_every steps and at every activation step. Across an
 * activation the adjoint follows add_material: zero on the nodes it added,
 * the carried powder's adjoint through the cell-mean transfer, and the
 * geometry (active cells, faces, mass) before it.
 *   T0, T_final, lambda0  device, n_members x n_dofs (T_final, lambda0 may
 *                         be NULL)
 *   obs_dof (host, n_obs): the observed nodes, shared; obs (host, n_members
 *   x n_obs): each member's data; obs_weight (host, n_obs, or NULL)
 *   J (host out, n_members): each member's misfit; grad (host out,
 *   n_members x 3): d J_j / d(member j's absorption efficiency of beam 0,
 *   its conductivity factor, the shared radius of beam 0).
 * At return every member is in its final state (geometry and phase state),
 * as after the forward run. */
int adm_ensemble_adjoint_gradient(
    void *ens, int method, double t0, double dt, int n_steps,
    int checkpoint_every, const double *T0, int32_t n_obs,
    const int64_t *obs_dof, const double *obs, const double *obs_weight,
    int32_t n_events, const adm_activation_event *events, double *T_final,
    double *J, double *grad, double *lambda0, void *stream);

/* The linearised operator by itself (no reference counterpart): with the
 * beams at time t and the context's current phase state, out (device,
 * overwritten) = (dA/dT)^T mu at T, A being adm_vmult's operator; grad
 * (host out[3], may be NULL) = mu . dA/dtheta for the parameters of
 * adm_adjoint_gradient. Reads the phase state without updating it; the
 * powder ratio this evaluation would carry forward is differentiated as a
 * function of T (no later evaluation depends on it here). */
int adm_operator_vjp(void *ctx, double t, const double *T, const double *mu,
                     double *out, double *grad, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ADM_H */
