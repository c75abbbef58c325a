// One instantiation of the cell and face kernels, selected by -DADM_P,
// -DADM_NS and -DADM_PK at compile time.
#include "adm_kernels.hpp"
#include "adm_kernels_2d.hpp"
#include "adm_launch.hpp"
#if ADM_P <= 2
#include "adm_plane_kernel.hpp"
#endif

namespace adm {

template <int P, int NS, int PK>
void launch_cell(const CellArgs &a, const double *src, double *dst,
                 int64_t n_blocks, size_t lds_bytes, unsigned n_members,
                 int waves, hipStream_t s) {
  const int cpb = waves * CellLanes<P>::cells_per_wave;
  const unsigned grid = (unsigned)n_blocks;
  const size_t xbytes = sizeof(double) * cpb * CellLanes<P>::stride;
  hipLaunchKernelGGL((cell_kernel<P, NS, PK>), dim3(grid, n_members),
                     dim3(64 * waves), lds_bytes + xbytes, s, a, src, dst);
}

#if ADM_P <= 2
template <int P, int NS, int PK>
void launch_plane(const CellArgs &a, const double *src, double *dst,
                  int64_t n_blocks, size_t lds_bytes, unsigned n_members,
                  int waves, hipStream_t s) {
  const size_t xbytes = sizeof(double) * waves * PlaneLanes<P>::wave_lds;
  hipLaunchKernelGGL((plane_kernel<P, NS, PK>), dim3((unsigned)n_blocks,
                     n_members), dim3(64 * waves), lds_bytes + xbytes, s, a,
                     src, dst);
}
template void launch_plane<ADM_P, ADM_NS, ADM_PK>(const CellArgs &,
                                                  const double *, double *,
                                                  int64_t, size_t, unsigned,
                                                  int, hipStream_t);
#endif

template <int P, int NS, int PK>
void launch_face(const FaceArgs &a, const double *src, double *dst,
                 unsigned n_members, hipStream_t s) {
  const int64_t per = FaceLanes<P>::faces_per_block;
  const int64_t grid = (a.n_faces + per - 1) / per;
  hipLaunchKernelGGL((face_kernel<P, NS, PK>), dim3((unsigned)grid, n_members),
                     dim3(256), 0, s, a, src, dst);
}

template <int P, int NS, int PK>
void launch_cell_2d(const CellArgs &a, const double *src, double *dst,
                    unsigned n_members, hipStream_t s) {
  constexpr int waves = Cell2dLanes<P>::waves_per_block;
  const unsigned grid = (unsigned)((a.n_rows + waves - 1) / waves);
  hipLaunchKernelGGL((cell_kernel_2d<P, NS, PK>), dim3(grid, n_members),
                     dim3(64 * waves), 0, s, a, src, dst);
}

template <int P, int NS, int PK>
void launch_face_2d(const FaceArgs &a, const double *src, unsigned n_members,
                    hipStream_t s) {
  const int64_t per = Face2dLanes<P>::faces_per_block;
  const int64_t grid = (a.n_faces + per - 1) / per;
  hipLaunchKernelGGL((face_kernel_2d<P, NS, PK>),
                     dim3((unsigned)grid, n_members), dim3(256), 0, s, a, src);
}

template void launch_cell_2d<ADM_P, ADM_NS, ADM_PK>(const CellArgs &,
                                                    const double *, double *,
                                                    unsigned, hipStream_t);
template void launch_face_2d<ADM_P, ADM_NS, ADM_PK>(const FaceArgs &,
                                                    const double *, unsigned,
                                                    hipStream_t);
template void launch_cell<ADM_P, ADM_NS, ADM_PK>(const CellArgs &,
                                                 const double *, double *,
                                                 int64_t, size_t, unsigned,
                                                 int, hipStream_t);
template void launch_face<ADM_P, ADM_NS, ADM_PK>(const FaceArgs &,
                                                 const double *, double *,
                                                 unsigned, hipStream_t);

}  // namespace adm
