// ib_pressure.hip -- pressure post-processing (calculate_pressure,
// src/operators_pressure.F90:110-437) on every grid level.
//
// D D^T p = D N(q) - D E^T f/dt on the m x n cells of each level, solved
// coarse to fine with boundary values from the next coarser level's
// pressure (the coarsest level: 1/2 |q/del|^2), ddti being a 2-D DST-I over
// the cells; the result is the pressure coefficient 2 (p - 1/2 |u|^2).
// One thread per cell / face / boundary slot; the DST passes reuse the
// row and column kernels with a cell geometry (m+1, n+1).
#include "ib_device.h"
#include "ib_kernels.h"

namespace ibg {

__device__ __forceinline__ long CPI(const Geom& g, int i, int j) { return (long)(i - 1) + (long)g.m * (j - 1); }

// nonlinear_corner (:441-459): node products of the face-averaged flux and
// the circulation; q0x is the uniform freestream flux of the level
__global__ void k_p_corner(Geom g, const double* __restrict__ w, const double* __restrict__ qx,
                           const double* __restrict__ qy, double q0x, double* __restrict__ oqx,
                           double* __restrict__ oqy) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.P) return;
  const int i = q % (g.m - 1) + 2, j = q / (g.m - 1) + 2;
  const double om = w[q];
  oqx[q] = 0.5 * ((qy[FYI(g, i, j)] + 0.0) + (qy[FYI(g, i - 1, j)] + 0.0)) * om;
  oqy[q] = -0.5 * ((qx[FXI(g, i, j)] + q0x) + (qx[FXI(g, i, j - 1)] + q0x)) * om;
}

// nonlinear_corner2edge (:463-504) with the 1/del^2 scaling: face averages of
// the node products, physical-boundary nodes from the bc vectors
__global__ void k_p_edges(Geom g, const double* __restrict__ oqx, const double* __restrict__ oqy,
                          const double* __restrict__ ubc, const double* __restrict__ vbc, double del,
                          double* __restrict__ nlx, double* __restrict__ nly) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < g.Fx) {
    const int i = q % (g.m + 1) + 1, j = q / (g.m + 1) + 1;
    nlx[q] = (0.5 * (ext(g, oqx, ubc, i, j + 1) + ext(g, oqx, ubc, i, j))) / (del * del);
  } else if (q < g.Fx + g.Fy) {
    const int r = q - g.Fx;
    const int i = r % g.m + 1, j = r / g.m + 1;
    nly[r] = (0.5 * (ext(g, oqy, vbc, i + 1, j) + ext(g, oqy, vbc, i, j))) / (del * del);
  }
}

// divergence over cells (:508-518): out = sign * div, or out += div
__global__ void k_p_div(Geom g, const double* __restrict__ hx, const double* __restrict__ hy, double* __restrict__ out,
                        int accumulate) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.m * g.n) return;
  const int i = q % g.m + 1, j = q / g.m + 1;
  const double d = hx[FXI(g, i + 1, j)] - hx[FXI(g, i, j)] + hy[FYI(g, i, j + 1)] - hy[FYI(g, i, j)];
  out[q] = accumulate ? out[q] + d : -d;
}

// coarsify_pressure (:530-563): coarse cell (m/4 + i/2 + 1, n/4 + j/2 + 1)
// from the 9-point sums ("temp") at the even fine cells (i, j), (i+2, j),
// (i, j+2), (i+2, j+2), i = 2..m-4, j = 2..n-4
__device__ __forceinline__ double p_temp(const Geom& g, const double* p1, int i, int j) {
  return p1[CPI(g, i, j)] +
         0.5 * (p1[CPI(g, i + 1, j)] + p1[CPI(g, i, j + 1)] + p1[CPI(g, i - 1, j)] + p1[CPI(g, i, j - 1)]) +
         0.25 * (p1[CPI(g, i + 1, j + 1)] + p1[CPI(g, i + 1, j - 1)] + p1[CPI(g, i - 1, j - 1)] +
                 p1[CPI(g, i - 1, j + 1)]);
}
__global__ void k_p_coarsify(Geom g, const double* __restrict__ p1, double* __restrict__ p2) {
  const int nx = (g.m - 4) / 2, ny = (g.n - 4) / 2;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nx * ny) return;
  const int i = 2 + 2 * (q % nx), j = 2 + 2 * (q / nx);
  const double yy = (p_temp(g, p1, i, j) + 3.0 * p_temp(g, p1, i + 2, j) + 3.0 * p_temp(g, p1, i, j + 2) +
                     9.0 * p_temp(g, p1, i + 2, j + 2)) /
                    16.0;
  p2[CPI(g, g.m / 4 + i / 2 + 1, g.n / 4 + j / 2 + 1)] = yy;
}

// boundary slots [left n | right n | bottom m | top m] of the cell grid
struct PSlots {
  int l, r, b, t;
  __device__ PSlots(const Geom& g) : l(0), r(g.n), b(2 * g.n), t(2 * g.n + g.m) {}
};

// get_bc_phi1 (:707-746): del |qx(i+1)+qx(i)|^2 + |qy(j+1)+qy(j)|^2 on the
// boundary cells of the coarsest level
__global__ void k_p_bc1(Geom g, const double* __restrict__ qx, const double* __restrict__ qy, double q0x, double del,
                        double* __restrict__ pbc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const PSlots sl(g);
  if (s >= 2 * g.m + 2 * g.n) return;
  int i, j;
  if (s < sl.r) { i = 1; j = s - sl.l + 1; }
  else if (s < sl.b) { i = g.m; j = s - sl.r + 1; }
  else if (s < sl.t) { j = 1; i = s - sl.b + 1; }
  else { j = g.n; i = s - sl.t + 1; }
  const double a = (qx[FXI(g, i + 1, j)] + q0x) + (qx[FXI(g, i, j)] + q0x);
  const double c = (qy[FYI(g, i, j + 1)] + 0.0) + (qy[FYI(g, i, j)] + 0.0);
  pbc[s] = del * (a * a + c * c);
}

// get_bc_phi2 (:750-795): boundary values for a finer level from the coarser
// level's pressure pc at the cells around the finer domain
__global__ void k_p_bc2(Geom g, const double* __restrict__ pc, double* __restrict__ pbc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const PSlots sl(g);
  if (s >= 2 * g.m + 2 * g.n) return;
  auto P_ = [&](int i, int j) { return pc[CPI(g, i, j)]; };
  double v;
  if (s >= sl.b) {  // bottom / top: cells i = m/4+1 .. 3m/4
    const bool top = s >= sl.t;
    const int o = s - (top ? sl.t : sl.b);
    const int i = g.m / 4 + 1 + o / 2, side = o & 1;
    const int j = top ? 3 * g.n / 4 + 1 : g.n / 4;
    const int jn = top ? j - 1 : j + 1;
    v = 0.25 * (2 * P_(i, j) + P_(i, jn) + P_(side ? i + 1 : i - 1, j));
  } else {  // left / right: cells j = n/4+1 .. 3n/4
    const bool right = s >= sl.r;
    const int o = s - (right ? sl.r : sl.l);
    const int j = g.n / 4 + 1 + o / 2, side = o & 1;
    const int i = right ? 3 * g.m / 4 + 1 : g.m / 4;
    const int in_ = right ? i - 1 : i + 1;
    v = 0.25 * (2 * P_(i, j) + P_(in_, j) + P_(i, side ? j + 1 : j - 1));
  }
  pbc[s] = v;
}

// apply_bc_phi (:799-815): left, right, bottom, top in that order
__global__ void k_p_applybc(Geom g, double* __restrict__ r, const double* __restrict__ pbc) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.m * g.n) return;
  const int i = q % g.m + 1, j = q / g.m + 1;
  if (i != 1 && i != g.m && j != 1 && j != g.n) return;
  const PSlots sl(g);
  double v = r[q];
  if (i == 1) v = v - pbc[sl.l + j - 1];
  if (i == g.m) v = v - pbc[sl.r + j - 1];
  if (j == 1) v = v - pbc[sl.b + i - 1];
  if (j == g.n) v = v - pbc[sl.t + i - 1];
  r[q] = v;
}

// static_plus_dynamic (:819-831) and the scaling to the pressure coefficient
__global__ void k_p_final(Geom g, const double* __restrict__ p, const double* __restrict__ qx,
                          const double* __restrict__ qy, double q0x, double del, double* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.m * g.n) return;
  const int i = q % g.m + 1, j = q / g.m + 1;
  const double a = (qx[FXI(g, i + 1, j)] + q0x) + (qx[FXI(g, i, j)] + q0x);
  const double c = (qy[FYI(g, i, j + 1)] + 0.0) + (qy[FYI(g, i, j)] + 0.0);
  out[q] = 2.0 * (p[q] - del * (a * a + c * c));
}

__global__ void k_scale_div(int n, const double* __restrict__ x, double d, double* __restrict__ y) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) y[q] = x[q] / d;
}

static dim3 blocks(long n) { return dim3((unsigned)ceil_div(n, 256)); }

void pressure_corner(hipStream_t st, const Geom& g, const double* w, const double* qx, const double* qy, double q0x,
                     double* oqx, double* oqy) {
  hipLaunchKernelGGL(k_p_corner, blocks(g.P), dim3(256), 0, st, g, w, qx, qy, q0x, oqx, oqy);
  IB_LAUNCH_CHECK();
}
void pressure_edges(hipStream_t st, const Geom& g, const double* oqx, const double* oqy, const double* ubc,
                    const double* vbc, double del, double* nlx, double* nly) {
  hipLaunchKernelGGL(k_p_edges, blocks((long)g.Fx + g.Fy), dim3(256), 0, st, g, oqx, oqy, ubc, vbc, del, nlx, nly);
  IB_LAUNCH_CHECK();
}
void pressure_div(hipStream_t st, const Geom& g, const double* hx, const double* hy, double* out, bool accumulate) {
  hipLaunchKernelGGL(k_p_div, blocks((long)g.m * g.n), dim3(256), 0, st, g, hx, hy, out, accumulate ? 1 : 0);
  IB_LAUNCH_CHECK();
}
void pressure_coarsify(hipStream_t st, const Geom& g, const double* fine, double* coarse) {
  const long nt = (long)((g.m - 4) / 2) * ((g.n - 4) / 2);
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_p_coarsify, blocks(nt), dim3(256), 0, st, g, fine, coarse);
  IB_LAUNCH_CHECK();
}
void pressure_bc1(hipStream_t st, const Geom& g, const double* qx, const double* qy, double q0x, double del,
                  double* pbc) {
  hipLaunchKernelGGL(k_p_bc1, blocks(2L * g.m + 2L * g.n), dim3(256), 0, st, g, qx, qy, q0x, del, pbc);
  IB_LAUNCH_CHECK();
}
void pressure_bc2(hipStream_t st, const Geom& g, const double* pc, double* pbc) {
  hipLaunchKernelGGL(k_p_bc2, blocks(2L * g.m + 2L * g.n), dim3(256), 0, st, g, pc, pbc);
  IB_LAUNCH_CHECK();
}
void pressure_applybc(hipStream_t st, const Geom& g, double* r, const double* pbc) {
  hipLaunchKernelGGL(k_p_applybc, blocks((long)g.m * g.n), dim3(256), 0, st, g, r, pbc);
  IB_LAUNCH_CHECK();
}
void pressure_final(hipStream_t st, const Geom& g, const double* p, const double* qx, const double* qy, double q0x,
                    double del, double* out) {
  hipLaunchKernelGGL(k_p_final, blocks((long)g.m * g.n), dim3(256), 0, st, g, p, qx, qy, q0x, del, out);
  IB_LAUNCH_CHECK();
}
void scale_div(hipStream_t st, int n, const double* x, double d, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale_div, blocks(n), dim3(256), 0, st, n, x, d, y);
  IB_LAUNCH_CHECK();
}

}  // namespace ibg
