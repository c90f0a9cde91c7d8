// ib_fluid.hip -- staggered-grid stencils of src/operators_fluid.F90 on CDNA4.
//
// One thread per output point, i fastest (coalesced along rows).  Index
// conventions follow the reference DMDAs: node (i,j), i=2..m, j=2..n at
// (i-2)+(m-1)(j-2); x-face (i,j), i=1..m+1, j=1..n at (i-1)+(m+1)(j-1);
// y-face (i,j), i=1..m, j=1..n+1 at (i-1)+m(j-1).  Physical-boundary ghosts of
// face fields are 0 (never written by DMGlobalToLocal); node ghosts come from
// boundary vectors, corners from the bottom/top part (the order in which
// curl_bc and nonlinear overwrite them).
#include "ib_device.h"
#include "ib_kernels.h"

namespace ibg {

struct FillOp {
  using Args = FillArgs;
  static __device__ __forceinline__ void tile(const FillArgs& a, int bx, int, int nbx) {
    long i = (long)bx * 256 + threadIdx.x;
    const long stride = (long)nbx * 256;
    for (; i < a.n; i += stride) a.p[i] = a.v;
  }
};

__global__ void __launch_bounds__(256) k_fill(FillArgs a) { FillOp::tile(a, blockIdx.x, 0, gridDim.x); }

__global__ void k_copy(double* d, const double* s, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = s[i];
}

__global__ void k_sub(double* a, const double* b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) a[i] = a[i] - b[i];
}

static int grid_for(long n) {
  long b = (n + 255) / 256;
  if (b > 65535L * 8) b = 65535L * 8;
  return (int)(b < 1 ? 1 : b);
}

TabEntry<FillArgs> fill_entry(double* p, long n, double v) {
  TabEntry<FillArgs> e{};
  e.a = FillArgs{p, n, v};
  e.nbx = grid_for(n);
  e.nby = 1;
  return e;
}

void fill(hipStream_t st, double* p, long n, double v) {
  if (n <= 0) return;
  const TabEntry<FillArgs> e = fill_entry(p, n, v);
  hipLaunchKernelGGL(k_fill, dim3(e.nbx), dim3(256), 0, st, e.a);
  IB_LAUNCH_CHECK();
}

void copy(hipStream_t st, double* d, const double* s, long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(256), 0, st, d, s, n);
  IB_LAUNCH_CHECK();
}

void sub_inplace(hipStream_t st, double* a, const double* b, long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(256), 0, st, a, b, n);
  IB_LAUNCH_CHECK();
}

// rot_int + rot_bc: src/operators_fluid.F90:203-233
__global__ void k_rot(Geom g, const double* __restrict__ hx, const double* __restrict__ hy, double* __restrict__ v) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.P) return;
  const int i = q % (g.m - 1) + 2, j = q / (g.m - 1) + 2;
  v[q] = hy[FYI(g, i, j)] - hy[FYI(g, i - 1, j)] - hx[FXI(g, i, j)] + hx[FXI(g, i, j - 1)];
}

void rot(hipStream_t st, const Geom& g, const double* hx, const double* hy, double* vort) {
  hipLaunchKernelGGL(k_rot, dim3(ceil_div(g.P, 256)), dim3(256), 0, st, g, hx, hy, vort);
  IB_LAUNCH_CHECK();
}

// coarsify: src/operators_fluid.F90:381-419 (fine odd nodes I,J = 3..m-1)
struct CoarsenOp {
  using Args = CoarsenArgs;
  static __device__ __forceinline__ void tile(const CoarsenArgs& a, int bx, int by, int);
};

__device__ __forceinline__ void CoarsenOp::tile(const CoarsenArgs& a, int bx, int by, int) {
  const Geom& g = a.g;
  const int nx = (g.m - 2) / 2, ny = (g.n - 2) / 2;
  const int q = bx * 256 + threadIdx.x;
  const double* __restrict__ wf = a.wf + by * a.stride;
  double* __restrict__ wc = a.wc + by * a.stride;
  if (q >= nx * ny) return;
  const int I = 3 + 2 * (q % nx), J = 3 + 2 * (q / nx);
  const int ic = g.m / 4 + (I + 1) / 2, jc = g.n / 4 + (J + 1) / 2;
  const double v = wf[ND(g, I, J)] +
                   0.5 * (wf[ND(g, I + 1, J)] + wf[ND(g, I, J + 1)] + wf[ND(g, I - 1, J)] + wf[ND(g, I, J - 1)]) +
                   0.25 * (wf[ND(g, I + 1, J + 1)] + wf[ND(g, I + 1, J - 1)] + wf[ND(g, I - 1, J - 1)] +
                           wf[ND(g, I - 1, J + 1)]);
  wc[ND(g, ic, jc)] = v;
}

__global__ void __launch_bounds__(256) k_coarsify(CoarsenArgs a) { CoarsenOp::tile(a, blockIdx.x, blockIdx.y, 0); }

TabEntry<CoarsenArgs> coarsify_entry(const Geom& g, const double* fine, double* coarse, int batch, long stride) {
  TabEntry<CoarsenArgs> e{};
  e.a = CoarsenArgs{g, fine, coarse, stride};
  e.nbx = ceil_div(((g.m - 2) / 2) * ((g.n - 2) / 2), 256);
  e.nby = batch;
  return e;
}

void coarsify(hipStream_t st, const Geom& g, const double* fine, double* coarse, int batch, long stride) {
  const TabEntry<CoarsenArgs> e = coarsify_entry(g, fine, coarse, batch, stride);
  hipLaunchKernelGGL(k_coarsify, dim3(e.nbx, e.nby), dim3(256), 0, st, e.a);
  IB_LAUNCH_CHECK();
}

// get_bc value of slot `slot` of the boundary vector from a coarse node array
// (src/operators_fluid.F90:475-535, slots of src/grid.F90:1091-1140)
__device__ __forceinline__ double getbc_slot(const Geom& g, const double* sc, int slot, double fac) {
  const int i0 = g.m / 4 + 1, i1 = 3 * g.m / 4 + 1, j0 = g.n / 4 + 1, j1 = 3 * g.n / 4 + 1;
  int side_i, side_j, I;
  bool vert;  // true: left/right (runs along j)
  if (slot < g.right) { vert = true; side_i = i0; I = slot; }
  else if (slot < g.bottom) { vert = true; side_i = i1; I = slot - g.right; }
  else if (slot < g.top) { vert = false; side_j = j0; I = slot - g.bottom; }
  else { vert = false; side_j = j1; I = slot - g.top; }
  double v;
  if (vert) {
    const int j = j0 + I / 2;
    if ((I & 1) == 0) v = sc[ND(g, side_i, j)];
    else v = 0.5 * (sc[ND(g, side_i, j)] + sc[ND(g, side_i, j + 1)]);
  } else {
    const int i = i0 + I / 2;
    if ((I & 1) == 0) v = sc[ND(g, i, side_j)];
    else v = 0.5 * (sc[ND(g, i, side_j)] + sc[ND(g, i + 1, side_j)]);
  }
  return v * fac;
}

struct GetBcOp {
  using Args = GetBcArgs;
  static __device__ __forceinline__ void tile(const GetBcArgs& a, int bx, int by, int) {
    const int slot = bx * 256 + threadIdx.x;
    if (slot >= a.g.nbc) return;
    const int l = by;
    a.bc[l * a.bstride + slot] = getbc_slot(a.g, a.src + l * a.sstride, slot, a.fac);
  }
};

__global__ void __launch_bounds__(256) k_get_bc(GetBcArgs a) { GetBcOp::tile(a, blockIdx.x, blockIdx.y, 0); }

TabEntry<GetBcArgs> get_bc_entry(const Geom& g, const double* src, long src_stride, double* bc, long bc_stride,
                                 int nlev, double fac) {
  TabEntry<GetBcArgs> e{};
  e.a = GetBcArgs{g, src, src_stride, bc, bc_stride, fac};
  e.nbx = ceil_div(g.nbc, 256);
  e.nby = nlev;
  return e;
}

void get_bc(hipStream_t st, const Geom& g, const double* src, long src_stride, double* bc, long bc_stride, int nlev,
            double fac) {
  if (nlev <= 0) return;
  const TabEntry<GetBcArgs> e = get_bc_entry(g, src, src_stride, bc, bc_stride, nlev, fac);
  hipLaunchKernelGGL(k_get_bc, dim3(e.nbx, e.nby), dim3(256), 0, st, e.a);
  IB_LAUNCH_CHECK();
}

// curl_int + curl_bc (src/operators_fluid.F90:423-471) and, in the same
// launch, get_bc(s, 1) for the next finer level.
struct CurlOp {
  using Args = CurlArgs;
  static __device__ __forceinline__ void tile(const CurlArgs& a, int bx, int by, int);
};

__device__ __forceinline__ void CurlOp::tile(const CurlArgs& a, int bx, int, int) {
  const Geom& g = a.g;
  const double* __restrict__ s = a.s;
  const double* __restrict__ sbc = a.sbc;
  double* __restrict__ qx = a.qx;
  double* __restrict__ qy = a.qy;
  double* __restrict__ bc_out = a.bc_out;
  const int q = bx * 256 + threadIdx.x;
  if (q < g.Fx) {
    const int i = q % (g.m + 1) + 1, j = q / (g.m + 1) + 1;
    qx[q] = ext(g, s, sbc, i, j + 1) - ext(g, s, sbc, i, j);
  } else if (q < g.Fx + g.Fy) {
    const int r = q - g.Fx;
    const int i = r % g.m + 1, j = r / g.m + 1;
    qy[r] = ext(g, s, sbc, i, j) - ext(g, s, sbc, i + 1, j);
  } else if (bc_out && q < g.Fx + g.Fy + g.nbc) {
    const int slot = q - g.Fx - g.Fy;
    bc_out[slot] = getbc_slot(g, s, slot, 1.0);
  }
}

__global__ void __launch_bounds__(256) k_curl_getbc(CurlArgs a) { CurlOp::tile(a, blockIdx.x, 0, 0); }

TabEntry<CurlArgs> curl_getbc_entry(const Geom& g, const double* s, const double* sbc, double* qx, double* qy,
                                    double* bc_out) {
  TabEntry<CurlArgs> e{};
  e.a = CurlArgs{g, s, sbc, qx, qy, bc_out};
  e.nbx = ceil_div(g.Fx + g.Fy + (bc_out ? g.nbc : 0), 256);
  e.nby = 1;
  return e;
}

void curl_getbc(hipStream_t st, const Geom& g, const double* s, const double* sbc, double* qx, double* qy,
                double* bc_out) {
  const TabEntry<CurlArgs> e = curl_getbc_entry(g, s, sbc, qx, qy, bc_out);
  hipLaunchKernelGGL(k_curl_getbc, dim3(e.nbx), dim3(256), 0, st, e.a);
  IB_LAUNCH_CHECK();
}

// nonlinear (src/operators_fluid.F90:553-587) + rot (:203-233) + apply_bc of
// lastbc and omegabc with vfac (src/timestepper.F90:184-189) + the AB2/Euler
// right-hand side of fluid_predictor (src/operators_fluid.F90:657-660).

__device__ __forceinline__ double QX(const NLArgs& a, int i, int j) {
  return (j < 1 || j > a.g.n) ? 0.0 : a.qx[FXI(a.g, i, j)] + a.q0x;
}
__device__ __forceinline__ double QY(const NLArgs& a, int i, int j) {
  return (i < 1 || i > a.g.m) ? 0.0 : a.qy[FYI(a.g, i, j)] + 0.0;
}
__device__ __forceinline__ double W_(const NLArgs& a, int i, int j) { return ext(a.g, a.w, a.lastbc, i, j); }
__device__ __forceinline__ double uavg(const NLArgs& a, int i, int j) { return 0.5 * (QX(a, i, j) + QX(a, i, j - 1)); }
__device__ __forceinline__ double vavg(const NLArgs& a, int i, int j) { return 0.5 * (QY(a, i, j) + QY(a, i - 1, j)); }
__device__ __forceinline__ double nlx(const NLArgs& a, int i, int j) {
  return 0.5 * (vavg(a, i, j + 1) * W_(a, i, j + 1) + vavg(a, i, j) * W_(a, i, j));
}
__device__ __forceinline__ double nly(const NLArgs& a, int i, int j) {
  return -0.5 * (uavg(a, i + 1, j) * W_(a, i + 1, j) + uavg(a, i, j) * W_(a, i, j));
}

struct NLOp {
  using Args = NLArgs;
  static __device__ __forceinline__ void tile(const NLArgs& a, int bx, int by, int);
};

__device__ __forceinline__ void NLOp::tile(const NLArgs& a, int bx, int, int) {
  const Geom& g = a.g;
  const int q = bx * 256 + threadIdx.x;
  if (q >= g.P) return;
  const int i = q % (g.m - 1) + 2, j = q / (g.m - 1) + 2;
  const double rhs = nly(a, i, j) - nly(a, i - 1, j) - nlx(a, i, j) + nlx(a, i, j - 1);
  if (a.raw) a.raw[q] = rhs;
  if (!a.r) return;
  // rhsbc = apply_bc(0, lastbc, vfac) then apply_bc(., omegabc, vfac)
  double bc = 0.0;
  const double* B[2] = {a.lastbc, a.omegabc};
  for (int b = 0; b < 2; ++b) {
    const double* v = B[b];
    if (i == 2) bc = bc + a.vfac * v[g.left + j - 1];
    if (i == g.m) bc = bc + a.vfac * v[g.right + j - 1];
    if (j == 2) bc = bc + a.vfac * v[g.bottom + i - 1];
    if (j == g.n) bc = bc + a.vfac * v[g.top + i - 1];
  }
  const int first = a.step ? (*a.step == 1) : a.first;
  const double ro = first ? rhs : a.rhs_old[q];
  a.r[q] = a.con1 * rhs + a.con2 * ro + bc;
  a.rhs_old[q] = rhs;
}

__global__ void __launch_bounds__(256) k_nl_rhs(NLArgs a) { NLOp::tile(a, blockIdx.x, 0, 0); }

void nl_rhs(hipStream_t st, const Geom& g, const double* omega, const double* qx, const double* qy, double q0x,
            const double* lastbc, const double* omegabc, double vfac, double con1, double con2, int first_step,
            double* rhs_old, double* r, double* raw_rhs) {
  NLArgs a{g, omega, qx, qy, q0x, lastbc, omegabc, vfac, con1, con2, first_step, nullptr, rhs_old, r, raw_rhs};
  hipLaunchKernelGGL(k_nl_rhs, dim3(ceil_div(g.P, 256)), dim3(256), 0, st, a);
  IB_LAUNCH_CHECK();
}

TabEntry<NLArgs> nl_rhs_entry(const Geom& g, const double* omega, const double* qx, const double* qy, double q0x,
                              const double* lastbc, const double* omegabc, double vfac, double con1, double con2,
                              const int* step, double* rhs_old, double* r) {
  TabEntry<NLArgs> e{};
  e.a = NLArgs{g, omega, qx, qy, q0x, lastbc, omegabc, vfac, con1, con2, 0, step, rhs_old, r, nullptr};
  e.nbx = ceil_div(g.P, 256);
  e.nby = 1;
  return e;
}

const void* fluid_tab_kernel(int kind) {
  switch (kind) {
    case TK_FILL: return (const void*)k_tab<FillOp>;
    case TK_GET_BC: return (const void*)k_tab<GetBcOp>;
    case TK_COARSIFY: return (const void*)k_tab<CoarsenOp>;
    case TK_CURL_GETBC: return (const void*)k_tab<CurlOp>;
    case TK_NL_RHS: return (const void*)k_tab<NLOp>;
    default: throw HipError("no stencil table kernel of this kind");
  }
}

}  // namespace ibg
