// Host launchers of the templated kernels; each instantiation lives in its own
// translation unit (adm_cell_inst.hip compiled once per configuration).
#pragma once

#include <hip/hip_runtime.h>

#include "adm_device.hpp"

namespace adm {

// PK: 0 polynomial properties, 1 general tables, 2 merged tables.
// n_members > 1 only with CellArgs/FaceArgs::members set (grid y = member).
// waves: wavefronts (row segments) per block of the cell kernel, 2..4.
template <int P, int NS, int PK>
void launch_cell(const CellArgs &a, const double *src, double *dst,
                 int64_t n_blocks, size_t lds_bytes, unsigned n_members,
                 int waves, hipStream_t s);

// The plane kernel (FE degrees 1 and 2 only): one launch of n_blocks
// persistent blocks of `waves` wavefronts looping over PX x PY cell
// patches, their node sums into the patch buffer (no atomics).
template <int P, int NS, int PK>
void launch_plane(const CellArgs &a, const double *src, double *dst,
                  int64_t n_blocks, size_t lds_bytes, unsigned n_members,
                  int waves, hipStream_t s);

template <int P, int NS, int PK>
void launch_face(const FaceArgs &a, const double *src, double *dst,
                 unsigned n_members, hipStream_t s);

// dim = 2: the cell kernel over a.n_rows row segments (4 per block) and the
// face kernel writing face_out
template <int P, int NS, int PK>
void launch_cell_2d(const CellArgs &a, const double *src, double *dst,
                    unsigned n_members, hipStream_t s);
template <int P, int NS, int PK>
void launch_face_2d(const FaceArgs &a, const double *src, unsigned n_members,
                    hipStream_t s);

}  // namespace adm
