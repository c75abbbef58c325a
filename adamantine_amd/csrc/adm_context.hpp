// Host-side context of one thermal operator (one mesh, FE degree, material,
// heat sources and boundary), the object behind the opaque handle of adm.h.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/adm.h"
#include "adm_device.hpp"

namespace adm {

struct AdmError : std::runtime_error {
  int code;
  AdmError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

int set_error(int code, const std::string &msg);

struct BeamData {
  adm_beam_desc desc;
  std::vector<adm_segment> segments;
};

// HeatSource::update_time on the host. pmod_out / den_out (optional) receive
// the power modifier and the parenthesised denominator of alpha.
void beam_update(const BeamData &b, double t, int &old_segment,
                 adm_beam_state &s, double *pmod_out = nullptr,
                 double *den_out = nullptr);
void gauss_points(int n, double *x, double *w);
void gauss_lobatto_points(int n, double *x, double *w);

// Device vectors of the explicit time steppers: the residual R (kept zero
// between stages), stage input Y, RK4 accumulator, RK3 stage values F and the
// inverse lumped mass, all of length n.
struct StepBuffers {
  int64_t n = 0;
  double *R = nullptr, *Y = nullptr, *acc = nullptr, *F = nullptr;
  double *inv_mass = nullptr;
  StepBuffers() = default;
  StepBuffers(const StepBuffers &) = delete;
  StepBuffers &operator=(const StepBuffers &) = delete;
  ~StepBuffers();
  void ensure(int method, int64_t n_);
};

// Device workspace of the material activation, allocated at the first one
// and kept (cells x 20 bytes + nodes x 16 bytes + scan storage).
struct ActWork {
  int32_t *boc = nullptr;          // box of each cell
  double *boxes = nullptr;         // 8 per box: bounds, cos, sin
  size_t boxes_cap = 0;
  int32_t *count = nullptr, *start = nullptr;  // faces per cell, scanned
  double *cell_powder = nullptr;   // cell means of the powder ratio
  uint64_t *packed = nullptr, *scan = nullptr;  // face slots per node
  void *temp = nullptr;
  size_t temp_bytes = 0;
  int64_t *totals = nullptr;
  int64_t *h_totals = nullptr;     // pinned
  int32_t *h_boc = nullptr;        // pinned
  double *h_boxes = nullptr;       // pinned staging of the boxes
  size_t h_boxes_cap = 0;
  ActWork() = default;
  ActWork(const ActWork &) = delete;
  ActWork &operator=(const ActWork &) = delete;
  ~ActWork();
};

// A context's activation geometry: what adm_activate changes besides the
// temperature and the phase state (the adjoint across an activation goes
// back to the geometry before it).
struct GeomSnap {
  uint32_t *meta = nullptr;
  double *cos = nullptr, *sin = nullptr;
  FaceRef *faces = nullptr;
  int64_t *fn_node = nullptr;
  int32_t *fn_offset = nullptr, *fn_slot = nullptr, *fn_rank = nullptr;
  int64_t n_faces = 0, n_fnodes = 0, n_slots = 0;
  int nz_active = 0;
  // host part: full copies of the activity and deposition directions
  // (host_copy), or the log of the one activation that followed the
  // snapshot -- the cells it activated with their {cos, sin} before and
  // after, and nz_active after -- which undoes (restore_geometry) and
  // redoes (redo_host_geometry) it without copying every cell's entries
  bool host_copy = true;
  std::vector<uint8_t> active;
  std::vector<double> dep_cos, dep_sin;
  std::vector<int64_t> log_cell;
  std::vector<double> log_old, log_new;
  int nz_after = 0;
  GeomSnap() = default;
  GeomSnap(const GeomSnap &) = delete;
  GeomSnap &operator=(const GeomSnap &) = delete;
  ~GeomSnap();
};

struct Context {
  explicit Context(const adm_problem_desc &d);
  ~Context();
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;

  void upload_constants();
  uint32_t face_mask(int64_t c) const;
  void rebuild_cells();
  void build_face_nodes();
  void set_cell_state(const double *cs);
  void get_qp_state(double *liquid, double *powder);
  void set_qp_state(const double *liquid, const double *powder);
  void set_time(double t);
  void cell_state_means(double *cs);
  std::vector<uint8_t> node_activity() const;
  void cells_to_activate(int n_boxes, const double *boxes,
                         int32_t *box_of_cell) const;
  void add_material(const int32_t *box_of_cell, const double *box_cos,
                    const double *box_sin, double new_temperature, double *T,
                    hipStream_t s);
  void add_material_host(const int32_t *box_of_cell, const double *box_cos,
                         const double *box_sin, double new_temperature,
                         double *T, hipStream_t s);
  // the device activation (cells to activate from boxes when box_of_cell
  // is NULL) and whether this mesh fits its 32-bit counters
  void activate(int n_boxes, const double *boxes, const int32_t *box_of_cell,
                const double *box_cos, const double *box_sin,
                double new_temperature, double *T, int32_t *box_of_cell_out,
                hipStream_t s);
  bool device_activation_fits() const;
  // the same on the host, box_of_cell_out (optional) receiving the cells'
  // boxes
  void activate_host(int n_boxes, const double *boxes, const double *box_cos,
                     const double *box_sin, double new_temperature, double *T,
                     int32_t *box_of_cell_out, hipStream_t s);
  // faces <- d_faces after a device activation
  void sync_faces_host();
  // the activation geometry, saved and put back (the inverse mass is the
  // caller's: it recomputes it)
  // host_copy false: the host part is left to the log of the next
  // activation (geom_log) or, with no activation after it, to the logs of
  // the activations redone on top of it
  void save_geometry(GeomSnap &g, hipStream_t s, bool host_copy = true) const;
  void restore_geometry(const GeomSnap &g, hipStream_t s);
  void redo_host_geometry(const GeomSnap &g);
  // set: the next activate() logs its host changes into this snapshot
  GeomSnap *geom_log = nullptr;
  // adjoint of the last activation, with the geometry after it (the cells
  // it activated carry kNewlyActive): Tbar = 0 on the nodes it added, and
  // the carried powder's adjoint through the cell-mean transfer, pbar
  // [slots][nq] (in: after, out: before) plus the adjoint of the new faces'
  // powder fpbar [n_faces][nfq] (read)
  void activation_adjoint(double *Tbar, double *pbar, const double *fpbar,
                          hipStream_t s);
  CellArgs cell_args() const;
  FaceArgs face_args() const;
  void build_merged_tables();
  int property_kind() const;
  // patch: the plane kernel's patch buffer for this call (NULL: the
  // context's own, grown on demand)
  void apply(const double *src, double *dst, int add, hipStream_t s,
             double *patch = nullptr);
  // nz_rows: cell layers (from z = 0) that can hold an active cell; the
  // launches cover only those (-1: all)
  // st: a stage update the plane path's gather pass applies instead of
  // writing dst (dst unused then)
  void launch(const CellArgs &a, const FaceArgs &fa, const double *src,
              double *dst, unsigned n_members, hipStream_t s,
              int nz_rows = -1, double *patch = nullptr,
              const StageArgs *st = nullptr) const;
  void launch_2d(const CellArgs &a, const FaceArgs &fa, const double *src,
                 double *dst, unsigned n_members, const int (&lim)[3],
                 hipStream_t s) const;
  void fill(double *x, double v, int64_t n, hipStream_t s);
  // the plane kernel's patch buffer, grown to n doubles if smaller
  double *ensure_patch_buffer(size_t n) const;
  // doubles of the patch buffer for launches over nz_rows cell layers
  size_t patch_buffer_size(int nz_rows, unsigned n_members) const;
  // the cell term goes through the plane kernel (which writes every node)
  bool plane_path() const { return use_plane && p <= 2; }
  void compute_inverse_mass(double *out, hipStream_t s);
  void ensure_step_buffers(int method);
  void step(int method, double t, double dt, double *T, hipStream_t s);
  // one step as `step`, with hook(i, x) enqueued before stage i's operator
  // evaluation (x: the stage input) and the beams set from the start of
  // their scan paths at every stage time
  void step_hooked(int method, double t, double dt, double *T, hipStream_t s,
                   const std::function<void(int, const double *)> &hook);
  // n_steps steps as replays of one captured step (beams from a table)
  void evolve_graph(int method, double t, double dt, int n_steps, double *T,
                    hipStream_t s);

  int device = 0;
  // dim = 2: an n_x x 1 x n_z box with one node layer in y (nn[1] == 1),
  // geometry constants in slots 0 and 2, faces 0, 1, 4, 5 only
  int dim = 3;
  bool has_face(int f) const { return dim == 3 || f / 2 != 1; }
  // the reference's (colorized) boundary id of face f, also the face's
  // slot in the [cell][2 dim][face point] state layout
  int face_id(int f) const { return dim == 2 && f >= 4 ? f - 2 : f; }
  int p = 1, n1 = 2, nq = 8, nfq = 4;
  int nc[3] = {0, 0, 0}, nn[3] = {0, 0, 0};
  int64_t n_cells = 0, n_dofs = 0;
  int n_states = 1;
  bool use_table = false;
  double det = 1.0;
  double wgl[kMaxN] = {0, 0, 0};
  KConst kc;

  std::vector<uint8_t> active, material;
  std::vector<double> dep_cos, dep_sin, cell_state;
  std::vector<uint32_t> meta;
  std::vector<int32_t> powder_slot;
  int64_t n_powder_cells = 0;
  // 1 + the highest cell layer (z index) holding an active cell, 0 if none
  int nz_active = 0;
  // boundary-term faces: d_faces holds n_faces of them; the host copy
  // `faces` is behind it after a device activation until sync_faces_host
  std::vector<FaceRef> faces;
  int64_t n_faces = 0;
  bool faces_stale = false;
  // allocated lengths of the face and face-node lists
  size_t face_cap = 0, face_out_cap = 0, face_powder_cap = 0;
  size_t fn_node_cap = 0, fn_offset_cap = 0, fn_slot_cap = 0;
  // activation workspace (the members of an ensemble share one; their
  // activations run one after the other)
  std::shared_ptr<ActWork> act;
  // the faces' nodes: node list, per-node offsets into the slot list, and
  // the slots f * (p+1)^2 + j of face_out (FaceArgs)
  int64_t n_fnodes = 0;
  double *d_face_out = nullptr;
  int64_t *d_fn_node = nullptr;
  int32_t *d_fn_offset = nullptr;
  int32_t *d_fn_slot = nullptr;
  int32_t *d_fn_rank = nullptr;  // [n_dofs]: index in fn_node, or -1

  std::vector<BeamData> beams;
  std::vector<int> old_segment;
  std::vector<adm_beam_state> beam_states;
  bool sources_on = false;
  // line kernel: 8 parity colors without atomics, or one atomic launch
  // (ADM_CELL_ATOMICS=1 in the environment at creation)
  bool use_coloring = true;
  // plane kernel (FE degrees 1 and 2; the default there unless
  // ADM_CELL_KERNEL=line is in the environment at creation)
  bool use_plane = false;
  // the 1-D tables have the symmetries the plane kernel's Q2 form uses
  bool plane_tables_ok = true;
  // ADM_LAUNCH_* mode of the fields above
  int launch_mode() const {
    return use_plane ? ADM_LAUNCH_PLANE
                     : (use_coloring ? ADM_LAUNCH_LINE_COLORED
                                     : ADM_LAUNCH_LINE_ATOMIC);
  }
  void set_launch_mode(int mode) {
    use_plane = mode == ADM_LAUNCH_PLANE;
    use_coloring = mode == ADM_LAUNCH_LINE_COLORED;
  }
  // row segments in XCD-contiguous block order (default; ADM_XCD_ORDER=0 in
  // the environment at creation restores plain block order)
  bool xcd_order = true;
  // wavefronts (row segments) per block of the cell kernel, 2..4
  int cell_waves = 4;
  int n_cu = 256;  // compute units of the device (grid sizing)

  // device buffers
  KConst *d_kc = nullptr;
  uint32_t *d_meta = nullptr;
  int32_t *d_powder_slot = nullptr;
  double *d_cos = nullptr, *d_sin = nullptr, *d_wgl = nullptr;
  double *d_liquid = nullptr;
  uint8_t *d_liquid_dirty = nullptr;
  double *d_powder = nullptr, *d_face_powder = nullptr;
  FaceRef *d_faces = nullptr;
  StepBuffers sb;
  // captured step loops (adm_evolve with use_graph): the beam states of
  // every stage and the stage counter the kernels read them by
  adm_beam_state *d_beam_table = nullptr;
  size_t beam_table_rows = 0;
  int32_t *d_slot = nullptr;
  // plane kernel patch buffer (members x patches x nodes per patch)
  mutable double *d_patch = nullptr;
  mutable size_t patch_cap = 0;
  // the library's static scratch array on this context's device
  // (adm_vmult_noalloc), resolved when the context is created
  double *d_scratch = nullptr;
  // DoF map of the caller's vectors (adm_set_dof_map): external DoF i is
  // box node dof_map[i]; n_ext < 0 without a map. map_src / map_dst are
  // box-node vectors the mapped entry points work on.
  int64_t n_ext = -1;
  int64_t *d_dof_map = nullptr;
  double *d_map_src = nullptr, *d_map_dst = nullptr;
  void set_dof_map(int64_t n, const int64_t *map);
  void clear_dof_map();
  void to_nodes(const double *ext, double *node, hipStream_t s) const;
  void to_dofs(const double *node, double *ext, int add, hipStream_t s) const;
};

// Members of an ensemble: one Context each (own state, tables, face lists),
// stepped together by single launches over (cells, members).
struct Ensemble {
  Ensemble(const adm_problem_desc &base, int n, const adm_member_desc *md);
  ~Ensemble();
  Ensemble(const Ensemble &) = delete;
  Ensemble &operator=(const Ensemble &) = delete;

  void set_time(double t);
  void apply(const double *src, double *dst, int add, hipStream_t s,
             const StageArgs *st = nullptr);
  void compute_inverse_mass(double *out, hipStream_t s);
  void ensure_step_buffers(int method);
  void step(int method, double t, double dt, double *T, hipStream_t s);
  // Context::step_hooked for every member at once
  void step_hooked(int method, double t, double dt, double *T, hipStream_t s,
                   const std::function<void(int, const double *)> &hook);
  // n_steps steps as replays of one captured step (Context::evolve_graph)
  void evolve_graph(int method, double t, double dt, int n_steps, double *T,
                    hipStream_t s);
  // the batched launches' arguments and cell layers
  CellArgs cell_args(int add) const;
  FaceArgs face_args() const;
  int nz_active() const;
  void refresh_members();

  int n_members = 0;
  int64_t n_dofs = 0;
  int64_t max_faces = 0;
  int64_t max_fnodes = 0;
  std::vector<std::unique_ptr<Context>> members;
  MemberArgs *d_members = nullptr;
  double pmod[ADM_MAX_BEAMS] = {0, 0, 0, 0};
  double den[ADM_MAX_BEAMS] = {1, 1, 1, 1};
  StepBuffers sb;
  // Analysis workspace, grown on demand and kept: device doubles and pinned
  // host staging. da_done is recorded after the last copy that reads the
  // staging, so it is rewritten only once that copy has finished.
  double *d_da = nullptr;
  size_t da_cap = 0;
  double *h_da = nullptr;
  size_t h_da_cap = 0;
  hipEvent_t da_done = nullptr;
  double *da_device(size_t n_doubles);
  double *da_host(size_t n_doubles);
  // captured step loops: beam states, and power modifiers / denominators,
  // of every stage, and the stage counter
  adm_beam_state *d_beam_table = nullptr;
  double *d_pd_table = nullptr;
  size_t beam_table_rows = 0;
  int32_t *d_slot = nullptr;
};

}  // namespace adm
