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

__global__ void k_fill(double* p, long n, double v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

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

void fill(hipStream_t st, double* p, long n, double v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, st, p, n, v);
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
__global__ void k_coarsify(Geom g, const double* __restrict__ wf, double* __restrict__ wc, long stride) {
  const int nx = (g.m - 2) / 2, ny = (g.n - 2) / 2;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  wf += blockIdx.y * stride;
  wc += blockIdx.y * stride;
  if (q >= nx * ny) return;
  const int I = 3 + 2 * (q % nx), J = 3 + 2 * (q / nx);
  const int ic = g.m / 4 + (I + 1) / 2, jc = g.n / 4 + (J + 1) / 2;
  const double v = wf[ND(g, I, J)] +
                   0.5 * (wf[ND(g, I + 1, J)] + wf[ND(g, I, J + 1)] + wf[ND(g, I - 1, J)] + wf[ND(g, I, J - 1)]) +
                   0.25 * (wf[ND(g, I + 1, J + 1)] + wf[ND(g, I + 1, J - 1)] + wf[ND(g, I - 1, J - 1)] +
                           wf[ND(g, I - 1, J + 1)]);
  wc[ND(g, ic, jc)] = v;
}

void coarsify(hipStream_t st, const Geom& g, const double* fine, double* coarse, int batch, long stride) {
  const int nt = ((g.m - 2) / 2) * ((g.n - 2) / 2);
  hipLaunchKernelGGL(k_coarsify, dim3(ceil_div(nt, 256), batch), dim3(256), 0, st, g, fine, coarse, stride);
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

__global__ void k_get_bc(Geom g, const double* __restrict__ src, long sstride, double* __restrict__ bc, long bstride,
                         double fac) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= g.nbc) return;
  const int l = blockIdx.y;
  bc[l * bstride + slot] = getbc_slot(g, src + l * sstride, slot, fac);
}

void get_bc(hipStream_t st, const Geom& g, const double* src, long src_stride, double* bc, long bc_stride, int nlev,
            double fac) {
  if (nlev <= 0) return;
  hipLaunchKernelGGL(k_get_bc, dim3(ceil_div(g.nbc, 256), nlev), dim3(256), 0, st, g, src, src_stride, bc, bc_stride,
                     fac);
  IB_LAUNCH_CHECK();
}

// curl_int + curl_bc (src/operators_fluid.F90:423-471) and, in the same
// launch, get_bc(s, 1) for the next finer level.
__global__ void k_curl_getbc(Geom g, const double* __restrict__ s, const double* __restrict__ sbc,
                             double* __restrict__ qx, double* __restrict__ qy, double* __restrict__ bc_out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
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

void curl_getbc(hipStream_t st, const Geom& g, const double* s, const double* sbc, double* qx, double* qy,
                double* bc_out) {
  const int nt = g.Fx + g.Fy + (bc_out ? g.nbc : 0);
  hipLaunchKernelGGL(k_curl_getbc, dim3(ceil_div(nt, 256)), dim3(256), 0, st, g, s, sbc, qx, qy, bc_out);
  IB_LAUNCH_CHECK();
}

// nonlinear (src/operators_fluid.F90:553-587) + rot (:203-233) + apply_bc of
// lastbc and omegabc with vfac (src/timestepper.F90:184-189) + the AB2/Euler
// right-hand side of fluid_predictor (src/operators_fluid.F90:657-660).
struct NLArgs {
  Geom g;
  const double* w;
  const double* qx;
  const double* qy;
  double q0x;
  const double* lastbc;
  const double* omegabc;
  double vfac, con1, con2;
  int first;
  double* rhs_old;
  double* r;
  double* raw;
};

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

__global__ void k_nl_rhs(NLArgs a) {
  const Geom& g = a.g;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
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
  const double ro = a.first ? rhs : a.rhs_old[q];
  a.r[q] = a.con1 * rhs + a.con2 * ro + bc;
  a.rhs_old[q] = rhs;
}

void nl_rhs(hipStream_t st, const Geom& g, const double* omega, const double* qx, const double* qy, double q0x,
            const double* lastbc, const double* omegabc, double vfac, double con1, double con2, int first_step,
            double* rhs_old, double* r, double* raw_rhs) {
  NLArgs a{g, omega, qx, qy, q0x, lastbc, omegabc, vfac, con1, con2, first_step, rhs_old, r, raw_rhs};
  hipLaunchKernelGGL(k_nl_rhs, dim3(ceil_div(g.P, 256)), dim3(256), 0, st, a);
  IB_LAUNCH_CHECK();
}

}  // namespace ibg
