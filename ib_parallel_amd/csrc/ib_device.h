// ib_device.h -- device-side index helpers shared by the kernel families.
#pragma once

#include "ib_common.h"

namespace ibg {

// node (i,j), i=2..m, j=2..n; x-face (i,j), i=1..m+1, j=1..n; y-face (i,j),
// i=1..m, j=1..n+1 (the reference's DMDA ranges, i fastest)
__device__ __forceinline__ long ND(const Geom& g, int i, int j) { return (long)(i - 2) + (long)(g.m - 1) * (j - 2); }
__device__ __forceinline__ long FXI(const Geom& g, int i, int j) { return (long)(i - 1) + (long)(g.m + 1) * (j - 1); }
__device__ __forceinline__ long FYI(const Geom& g, int i, int j) { return (long)(i - 1) + (long)g.m * (j - 1); }

// node value with physical-boundary ghosts from a boundary vector, in the
// order the reference overwrites them (left, right, bottom, top: corners come
// from bottom/top), src/operators_fluid.F90:448-451 and :563-566
__device__ __forceinline__ double ext(const Geom& g, const double* a, const double* bc, int i, int j) {
  if (j == 1) return bc[g.bottom + i - 1];
  if (j == g.n + 1) return bc[g.top + i - 1];
  if (i == 1) return bc[g.left + j - 1];
  if (i == g.m + 1) return bc[g.right + j - 1];
  return a[ND(g, i, j)];
}

// curl (C) of s with boundary bc at one face of the unified key space
// (x-faces, then Fx + y-faces): src/operators_fluid.F90:423-471
__device__ __forceinline__ double curl_at(const Geom& g, const double* s, const double* bc, int key) {
  if (key < g.Fx) {
    const int i = key % (g.m + 1) + 1, j = key / (g.m + 1) + 1;
    return ext(g, s, bc, i, j + 1) - ext(g, s, bc, i, j);
  }
  const int r = key - g.Fx;
  const int i = r % g.m + 1, j = r / g.m + 1;
  return ext(g, s, bc, i, j) - ext(g, s, bc, i + 1, j);
}

// A kernel family OP (OP::Args, OP::tile(args, bx, by, nbx)) launched over
// a device table of simulations: blockIdx.z selects the entry.
// The table is read through the constant address space (it does not change
// while a launch runs), so argument loads stay scalar like kernel arguments.
template <class OP>
__global__ void __launch_bounds__(256) k_tab(const TabEntry<typename OP::Args>* __restrict__ t) {
  using E = TabEntry<typename OP::Args>;
  const __attribute__((address_space(4))) E* ct = (const __attribute__((address_space(4))) E*)t;
  const E& e = *(const E*)(ct + blockIdx.z);
  if ((int)blockIdx.x >= e.nbx || (int)blockIdx.y >= e.nby) return;
  OP::tile(e.a, blockIdx.x, blockIdx.y, e.nbx);
}

}  // namespace ibg
