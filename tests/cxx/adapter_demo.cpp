// Host C++ program on the adm_cxx adapter: the 32x16x8 Q1 plain workload
// (solid, k = 11.8 + 0.0138 T, Goldak 200 W along x at mid-y on the top
// surface, adiabatic), one vmult at t = 1e-4 and 20 forward-Euler steps of
// 1e-6 s from 300 K. Writes n, the vmult result and the final temperature
// (int64 + 2 n doubles) to argv[1]. Built by adamantine_amd/build.py; run
// by tests/test_gpu_cxx_adapter.py.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "adm_cxx.hpp"

static void hip_check(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s output.bin\n", argv[0]);
    return 2;
  }
  try {
    const double mm = 1e-3, um = 1e-6;
    adm_problem_desc d;
    std::memset(&d, 0, sizeof(d));
    d.mesh.dim = 3;
    d.mesh.fe_degree = 1;
    d.mesh.quadrature = 0;
    const int nc[3] = {32, 16, 8};
    const double ext[3] = {2 * mm, 1 * mm, 0.5 * mm};
    for (int k = 0; k < 3; ++k) {
      d.mesh.n_cells[k] = nc[k];
      d.mesh.extent[k] = ext[k];
    }
    adm_material_desc &m = d.material;
    m.n_materials = 1;
    m.n_states = 1;
    m.use_table = 0;
    m.p_order = 1;
    m.single_material_layout = 1;
    for (int a = 0; a < ADM_MAX_MATERIALS; ++a)
      for (int b = 0; b < ADM_N_PROPS; ++b) m.prop[a][b] = DBL_MAX;
    m.poly[0][ADM_DENSITY][0][0] = 7904.0;
    m.poly[0][ADM_SPECIFIC_HEAT][0][0] = 714.0;
    for (int k = ADM_THERMAL_CONDUCTIVITY_X; k <= ADM_THERMAL_CONDUCTIVITY_Z;
         ++k) {
      m.poly[0][k][0][0] = 11.8;
      m.poly[0][k][0][1] = 0.0138;
    }
    d.n_boundary_ids = 6;
    for (int k = 0; k < 6; ++k) d.boundary_type[k] = ADM_BC_ADIABATIC;
    // scan path: a 1e-9 s point at x = 0, then a line to x = 2 mm at 1 m/s
    // (ScanPath::load_segment_scan_path, source/ScanPath.cc:58-141)
    adm_segment seg[2];
    std::memset(seg, 0, sizeof(seg));
    const double y = 0.5 * mm, z = 0.5 * mm;
    seg[0].end_time = 1e-9;
    seg[0].power_modifier = 0.0;
    seg[0].end_point[0] = 0.0;
    seg[0].end_point[1] = y;
    seg[0].end_point[2] = z;
    seg[1].power_modifier = 1.0;
    seg[1].end_point[0] = 2 * mm;
    seg[1].end_point[1] = y;
    seg[1].end_point[2] = z;
    double len2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double dx = seg[1].end_point[k] - seg[0].end_point[k];
      len2 += dx * dx;
    }
    seg[1].end_time = seg[0].end_time + std::fabs(std::sqrt(len2) / 1.0);
    d.n_beams = 1;
    adm_beam_desc &b = d.beams[0];
    b.type = ADM_SOURCE_GOLDAK;
    b.absorption_efficiency = 0.3;
    b.max_power = 200.0;
    b.radius = 50 * um;
    b.depth = 50 * um;
    b.n_segments = 2;
    b.segments = seg;

    // the scan-path search only moves forward in time (as ScanPath's), so
    // the vmult at t = 1e-4 gets an operator of its own
    adm_cxx::ThermalPhysics phys(d, ADM_FORWARD_EULER);
    adm_cxx::ThermalOperator op(d);
    const int64_t n = op.m();
    std::vector<double> host(n, 300.0), out_v(n), out_t(n);
    double *T = nullptr, *dst = nullptr;
    hip_check(hipMalloc(&T, sizeof(double) * n));
    hip_check(hipMalloc(&dst, sizeof(double) * n));
    hip_check(hipMemcpy(T, host.data(), sizeof(double) * n,
                        hipMemcpyHostToDevice));
    op.set_time(1e-4);
    op.vmult(dst, T);
    hip_check(hipDeviceSynchronize());
    hip_check(hipMemcpy(out_v.data(), dst, sizeof(double) * n,
                        hipMemcpyDeviceToHost));
    double t = 0.0;
    for (int k = 0; k < 20; ++k) t = phys.evolve_one_time_step(t, 1e-6, T);
    hip_check(hipDeviceSynchronize());
    hip_check(hipMemcpy(out_t.data(), T, sizeof(double) * n,
                        hipMemcpyDeviceToHost));
    hip_check(hipFree(T));
    hip_check(hipFree(dst));
    std::FILE *f = std::fopen(argv[1], "wb");
    if (!f) throw std::runtime_error("cannot open the output file");
    std::fwrite(&n, sizeof(n), 1, f);
    std::fwrite(out_v.data(), sizeof(double), n, f);
    std::fwrite(out_t.data(), sizeof(double), n, f);
    std::fclose(f);
    std::printf("n_dofs %lld, t %.9g, max T %.6f\n", (long long)n, t,
                *std::max_element(out_t.begin(), out_t.end()));
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
