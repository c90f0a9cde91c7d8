// ib_common.h -- shared definitions of the MI355X immersed-boundary library.
//
// Layout in HBM (FP64, Fortran order, i fastest), per grid level k:
//   omega, s, rhs_old  : nodes  i = 2..m,   j = 2..n    -> P  = (m-1)(n-1)
//   qx                 : faces  i = 1..m+1, j = 1..n    -> Fx = (m+1) n
//   qy                 : faces  i = 1..m,   j = 1..n+1  -> Fy = m (n+1)
//   boundary vectors   : [left n+1 | right n+1 | bottom m+1 | top m+1] -> nbc
// Levels are stored back to back (level-major) so one launch can address all
// of them through a level stride.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace ibg {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

#define IB_HIP(expr)                                                                              \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess)                                                                         \
      throw ::ibg::HipError(std::string(#expr) + " -> " + hipGetErrorString(_e) + " at " +      \
                            __FILE__ + ":" + std::to_string(__LINE__));                           \
  } while (0)

#define IB_LAUNCH_CHECK() IB_HIP(hipGetLastError())

constexpr int kMaxStages = 24;
constexpr int kTaps = 49;  // (2*support2+1)^2, support2 = ceiling(3) (src/grid.F90:12,468)

// Grid geometry shared by every kernel.
struct Geom {
  int m, n, L;
  int P, Fx, Fy, nbc;
  int left, right, bottom, top;
};

// Mixed-radix Stockham plan for a complex FFT of length L = 2M (M = N+1),
// used to evaluate two real DST-I of length N at once.
struct FFTPlan {
  int L, M, N;
  int nst;
  int rad[kMaxStages];
  const double2* W;  // W[t] = exp(-2 pi i t / L), device
};

// One torsional body (family 2) as the device sees it.
struct TorsBody {
  int off;      // first entry in the force vector
  int npts;
  double gamma0, deltas, xbf0, ybf0;
  double M, K, C;  // itheta, ktheta, ctheta
};

__host__ __device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// One simulation's entry in a launch over several independent simulations
// (the batched step): its launch arguments and the extent of its own grid.
// Such a launch runs a grid of (max nbx, max nby, entries) workgroups; the
// workgroups past their entry's extent return at once.
template <class A>
struct TabEntry {
  A a;
  int nbx, nby;
};

}  // namespace ibg
