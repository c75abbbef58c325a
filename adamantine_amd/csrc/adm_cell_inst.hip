// One instantiation of the cell kernel, selected by -DADM_P, -DADM_NS and
// -DADM_TABLE at compile time.
#include "adm_kernels.hpp"
#include "adm_launch.hpp"

namespace adm {

template <int P, int NS, bool TABLE>
void launch_cell(const CellArgs &a, const double *src, double *dst,
                 int64_t n_cells, hipStream_t s) {
  const int block = 256;
  const int64_t grid = (n_cells + block - 1) / block;
  hipLaunchKernelGGL((cell_kernel<P, NS, TABLE>), dim3((unsigned)grid),
                     dim3(block), 0, s, a, src, dst, n_cells);
}

template void launch_cell<ADM_P, ADM_NS, (ADM_TABLE != 0)>(
    const CellArgs &, const double *, double *, int64_t, hipStream_t);

}  // namespace adm
