// Device-side SBP and Tam-Gottlieb filter stencils with their patch closures, and the tile
// bookkeeping shared by the stage kernels (ajet_kernels.hip) and the artificial-diffusivity
// kernels (ajet_lad.hip).  Each translation unit holds its own copy of the constant tables
// and uploads it (upload_tables / upload_lad_tables).
#pragma once

#include <hip/hip_runtime.h>

#include "ajet_internal.h"

namespace ajet {

static __constant__ double c_sbp1[9][6];
static __constant__ double c_sbp2[9][6];
static __constant__ double c_tg[9][11];

// Interior stencils SBP1(4,0:4), SBP2(4,0:4) and TGfilt(4,0:8) (ModOperator.f90:42, 78, 111-112).
constexpr double D1A = 1.0 / 12.0, D1B = -(2.0 / 3.0), D1D = 2.0 / 3.0, D1E = -(1.0 / 12.0);
constexpr double D2A = -(1.0 / 12.0), D2B = 4.0 / 3.0, D2C = -(5.0 / 2.0), D2D = 4.0 / 3.0, D2E = -(1.0 / 12.0);
constexpr double TG0 = 107.0 / 103680.0, TG1 = 149.0 / 12960.0, TG2 = 1997.0 / 25920.0, TG3 = 3091.0 / 12960.0,
                 TG4 = 3565.0 / 10368.0;

// SBP first-derivative stencil sum (before the metric) at a point whose row code is `row`;
// `c` points at the point itself inside a strided line, stride s.  The interior sum skips
// the zero centre weight, which adds an exact zero in the reference.
__device__ __forceinline__ double d1_sum(uint32_t row, const double* c, int s) {
  if (row == ROW_INTERIOR) return ((D1A * c[-2 * s] + D1B * c[-s]) + D1D * c[s]) + D1E * c[2 * s];
  if (row == ROW_BLANK) return 0.0;
  const int base = row < 4 ? -(int)row : -(int)row + 3;
  const double* v = c + base * s;
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < 6; ++m) acc = acc + c_sbp1[row][m] * v[m * s];
  return acc;
}

__device__ __forceinline__ double d2_sum(uint32_t row, const double* c, int s) {
  if (row == ROW_INTERIOR) return (((D2A * c[-2 * s] + D2B * c[-s]) + D2C * c[0]) + D2D * c[s]) + D2E * c[2 * s];
  if (row == ROW_BLANK) return 0.0;
  const int base = row < 4 ? -(int)row : -(int)row + 3;
  const double* v = c + base * s;
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < 6; ++m) acc = acc + c_sbp2[row][m] * v[m * s];
  return acc;
}

// PerformTGfilter at one non-blanked point (ModTGfilter.f90:19-44, 58-83): the interior
// 9-point filter on var(i-4:i+4), or the 11-point closure row on var(i_start:i_start+10)
// (bc_dir = +1, rows 0-3) or var(i_end-10:i_end) (bc_dir = -1, rows 5-8), summed from the
// first element.
__device__ __forceinline__ double tg_sum(uint32_t row, const double* c, int s) {
  if (row == ROW_INTERIOR)
    return (((((((TG0 * c[-4 * s] + TG1 * c[-3 * s]) + TG2 * c[-2 * s]) + TG3 * c[-s]) + TG4 * c[0]) +
              TG3 * c[s]) + TG2 * c[2 * s]) + TG1 * c[3 * s]) + TG0 * c[4 * s];
  const int base = row < 4 ? -(int)row : -(int)row - 2;
  const double* v = c + base * s;
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < 11; ++m) acc = acc + c_tg[row][m] * v[m * s];
  return acc;
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous run of
// tiles (z fastest) so that tiles sharing halo rows meet in the same L2.  nblk_pad is the
// grid size, a multiple of 8.
__device__ __forceinline__ int xcd_tile(int b, int nblk_pad) { return (b & 7) * (nblk_pad >> 3) + (b >> 3); }

}  // namespace ajet
