// Host launchers of the templated kernels; each instantiation lives in its own
// translation unit (adm_cell_inst.hip compiled once per configuration).
#pragma once

#include <hip/hip_runtime.h>

#include "adm_device.hpp"

namespace adm {

template <int P, int NS, bool TABLE>
void launch_cell(const CellArgs &a, const double *src, double *dst,
                 int64_t n_cells, hipStream_t s);

}  // namespace adm
