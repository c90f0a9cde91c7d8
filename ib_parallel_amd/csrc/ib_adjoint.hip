// ib_adjoint.hip -- transposes of the step's stencils for the discrete adjoint.
//
// Every kernel here is the exact transpose of a forward kernel of
// ib_fluid.hip (same index maps and boundary rules), written as a gather so
// that each output is summed in a fixed order without atomics.  The spectral
// passes need no transposes: the 2-D DST-I is a symmetric matrix and the
// spectral tables are diagonal.
#include "ib_device.h"
#include "ib_kernels.h"

namespace ibg {

// Transpose of curl (C with boundary values sbc, src/operators_fluid.F90:
// 423-471): interior nodes get rot(qx, qy); each boundary slot gets the
// faces that read its ghost node (ext(): bottom/top rows own the corners).
// sbar_bc is written (not accumulated).
__global__ void k_curl_adj_bc(Geom g, const double* __restrict__ qx, const double* __restrict__ qy,
                              double* __restrict__ bcb) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= g.nbc) return;
  double v = 0.0;
  const int m = g.m, n = g.n;
  if (slot >= g.bottom) {
    const bool top = slot >= g.top;
    const int i = slot - (top ? g.top : g.bottom) + 1;  // node column 1..m+1
    const int j = top ? n + 1 : 1;
    // qx(i, j) = ext(i, j+1) - ext(i, j) reads node (i, 1) with -1; qx(i, n)
    // reads (i, n+1) with +1; qy(i, j) = ext(i, j) - ext(i+1, j)
    if (top) v = qx[FXI(g, i, n)];
    else v = -qx[FXI(g, i, 1)];
    if (i <= m) v = v + qy[FYI(g, i, j)];
    if (i >= 2) v = v - qy[FYI(g, i - 1, j)];
  } else {
    const bool right = slot >= g.right;
    const int j = slot - (right ? g.right : g.left) + 1;  // node row 1..n+1
    if (j >= 2 && j <= n) {
      const int i = right ? m + 1 : 1;
      v = qx[FXI(g, i, j - 1)] - qx[FXI(g, i, j)];
      v = right ? v - qy[FYI(g, m, j)] : v + qy[FYI(g, 1, j)];
    }
  }
  bcb[slot] = v;
}

void curl_adj_bc(hipStream_t st, const Geom& g, const double* qx, const double* qy, double* bcb) {
  hipLaunchKernelGGL(k_curl_adj_bc, dim3(ceil_div(g.nbc, 256)), dim3(256), 0, st, g, qx, qy, bcb);
  IB_LAUNCH_CHECK();
}

// Transpose of get_bc (src/operators_fluid.F90:475-535) as a gather: the
// coarse nodes get_bc reads form the ring i = i0..i1 x j = j0..j1 of the
// coarse level; node t of a side is read by slot 2t (weight 1) and slots
// 2t - 1, 2t + 1 (weight 1/2).  sc(node) += fac * (its slot sum); a corner
// node gathers from both of its sides.  One thread per ring node.
__device__ __forceinline__ double getbc_side(const double* __restrict__ b, int base, int t, int last) {
  double v = b[base + 2 * t];
  if (2 * t + 1 <= last) v = v + 0.5 * b[base + 2 * t + 1];
  if (t >= 1) v = v + 0.5 * b[base + 2 * t - 1];
  return v;
}

__global__ void k_getbc_adj(Geom g, const double* __restrict__ bcb, double fac, double* __restrict__ sc) {
  const int i0 = g.m / 4 + 1, i1 = 3 * g.m / 4 + 1, j0 = g.n / 4 + 1, j1 = 3 * g.n / 4 + 1;
  const int nh = i1 - i0 + 1, nv = j1 - j0 - 1;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2 * nh + 2 * nv) return;
  int i, j;
  if (q < nh) { i = i0 + q; j = j0; }
  else if (q < 2 * nh) { i = i0 + q - nh; j = j1; }
  else if (q < 2 * nh + nv) { i = i0; j = j0 + 1 + (q - 2 * nh); }
  else { i = i1; j = j0 + 1 + (q - 2 * nh - nv); }
  double acc = 0.0;
  if (i == i0) acc = acc + getbc_side(bcb, g.left, j - j0, g.n);
  if (i == i1) acc = acc + getbc_side(bcb, g.right, j - j0, g.n);
  if (j == j0) acc = acc + getbc_side(bcb, g.bottom, i - i0, g.m);
  if (j == j1) acc = acc + getbc_side(bcb, g.top, i - i0, g.m);
  sc[ND(g, i, j)] += fac * acc;
}

void getbc_adj(hipStream_t st, const Geom& g, const double* bcb, double fac, double* sc) {
  const int nring = 2 * (g.m / 2 + 1) + 2 * (g.n / 2 - 1);
  hipLaunchKernelGGL(k_getbc_adj, dim3(ceil_div(nring, 256)), dim3(256), 0, st, g, bcb, fac, sc);
  IB_LAUNCH_CHECK();
}

// Transpose of apply_bc(v, bc, fac) (src/operators_fluid.F90:539-549):
// bcb[slot] = fac * vb(node the slot is added to), 0 for slots no node
// reads.  Written, not accumulated.
__global__ void k_applybc_adj(Geom g, const double* __restrict__ vb, double fac, double* __restrict__ bcb) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= g.nbc) return;
  double v = 0.0;
  if (slot < g.right) { const int j = slot + 1; if (j >= 2 && j <= g.n) v = vb[ND(g, 2, j)]; }
  else if (slot < g.bottom) { const int j = slot - g.right + 1; if (j >= 2 && j <= g.n) v = vb[ND(g, g.m, j)]; }
  else if (slot < g.top) { const int i = slot - g.bottom + 1; if (i >= 2 && i <= g.m) v = vb[ND(g, i, 2)]; }
  else { const int i = slot - g.top + 1; if (i >= 2 && i <= g.m) v = vb[ND(g, i, g.n)]; }
  bcb[slot] = fac * v;
}

void applybc_adj(hipStream_t st, const Geom& g, const double* vb, double fac, double* bcb) {
  hipLaunchKernelGGL(k_applybc_adj, dim3(ceil_div(g.nbc, 256)), dim3(256), 0, st, g, vb, fac, bcb);
  IB_LAUNCH_CHECK();
}

// Transpose of coarsify (src/operators_fluid.F90:381-419): wf += Co^T(wc
// centre), then the coarse centre nodes (overwritten forward) are zeroed.
// One thread per fine node gathers from the up to 4 coarse centres whose
// 3x3 stencil covers it.
__global__ void k_coarsify_adj(Geom g, double* __restrict__ wc, double* __restrict__ wf) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.P) return;
  const int Ip = q % (g.m - 1) + 2, Jp = q / (g.m - 1) + 2;
  double acc = 0.0;
  for (int dI = -1; dI <= 1; ++dI) {
    const int I = Ip + dI;
    if ((I & 1) == 0 || I < 3 || I > g.m - 1) continue;
    for (int dJ = -1; dJ <= 1; ++dJ) {
      const int J = Jp + dJ;
      if ((J & 1) == 0 || J < 3 || J > g.n - 1) continue;
      const double wgt = (dI == 0 && dJ == 0) ? 1.0 : ((dI == 0 || dJ == 0) ? 0.5 : 0.25);
      const int ic = g.m / 4 + (I + 1) / 2, jc = g.n / 4 + (J + 1) / 2;
      acc += wgt * wc[ND(g, ic, jc)];
    }
  }
  wf[q] += acc;
}

__global__ void k_coarsify_adj_zero(Geom g, double* __restrict__ wc) {
  const int nx = (g.m - 2) / 2, ny = (g.n - 2) / 2;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nx * ny) return;
  const int I = 3 + 2 * (q % nx), J = 3 + 2 * (q / nx);
  wc[ND(g, g.m / 4 + (I + 1) / 2, g.n / 4 + (J + 1) / 2)] = 0.0;
}

void coarsify_adj(hipStream_t st, const Geom& g, double* wc, double* wf) {
  hipLaunchKernelGGL(k_coarsify_adj, dim3(ceil_div(g.P, 256)), dim3(256), 0, st, g, wc, wf);
  IB_LAUNCH_CHECK();
  const int nt = ((g.m - 2) / 2) * ((g.n - 2) / 2);
  hipLaunchKernelGGL(k_coarsify_adj_zero, dim3(ceil_div(nt, 256)), dim3(256), 0, st, g, wc);
  IB_LAUNCH_CHECK();
}

// Transpose of the nonlinear term's Jacobian (nonlinear + rot,
// src/operators_fluid.F90:553-587 and :203-233) at the forward point
// (w, qx + q0x, qy, lastbc).  Stage 1 forms the face-array adjoints of nlx
// and nly from rb (the adjoint of rhs on interior nodes):
//   nlxb(i, j) = rb(i, j + 1) - rb(i, j)   (i = 2..m, j = 1..n)
//   nlyb(i, j) = rb(i, j) - rb(i + 1, j)   (i = 1..m, j = 2..n)
// stored on (m+1) x (n+1) node-shaped arrays (zero outside those ranges).
__device__ __forceinline__ double rbi(const Geom& g, const double* rb, int i, int j) {
  return (i >= 2 && i <= g.m && j >= 2 && j <= g.n) ? rb[ND(g, i, j)] : 0.0;
}
__device__ __forceinline__ long NN(const Geom& g, int i, int j) { return (long)(i - 1) + (long)(g.m + 1) * (j - 1); }

__global__ void k_nl_adj1(Geom g, const double* __restrict__ rb, double* __restrict__ nlxb, double* __restrict__ nlyb) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int NA = (g.m + 1) * (g.n + 1);
  if (q >= NA) return;
  const int i = q % (g.m + 1) + 1, j = q / (g.m + 1) + 1;
  nlxb[q] = (i >= 2 && i <= g.m && j >= 1 && j <= g.n) ? rbi(g, rb, i, j + 1) - rbi(g, rb, i, j) : 0.0;
  nlyb[q] = (i >= 1 && i <= g.m && j >= 2 && j <= g.n) ? rbi(g, rb, i, j) - rbi(g, rb, i + 1, j) : 0.0;
}

// Stage 2, one thread per node (i, j) = 1..m+1 x 1..n+1: with W = ext(w,
// lastbc), uavg, vavg of the forward point,
//   Wb = 0.5 vavg (nlxb(i,j-1) + nlxb(i,j)) - 0.5 uavg (nlyb(i-1,j) + nlyb(i,j))
//   vavgb = 0.5 W (nlxb(i,j-1) + nlxb(i,j)),  uavgb = -0.5 W (nlyb(i-1,j) + nlyb(i,j))
// Wb is added to wb on interior nodes and written to the lastbc slot a
// ghost node reads (lbb); vavgb, uavgb go to node arrays for stage 3.
__global__ void k_nl_adj2(Geom g, const double* __restrict__ w, const double* __restrict__ qx,
                          const double* __restrict__ qy, double q0x, const double* __restrict__ lastbc,
                          const double* __restrict__ nlxb, const double* __restrict__ nlyb, double* __restrict__ wb,
                          double* __restrict__ lbb, double* __restrict__ ub, double* __restrict__ vb) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int NA = (g.m + 1) * (g.n + 1);
  if (q >= NA) return;
  const int i = q % (g.m + 1) + 1, j = q / (g.m + 1) + 1;
  auto QX = [&](int ii, int jj) { return (jj < 1 || jj > g.n) ? 0.0 : qx[FXI(g, ii, jj)] + q0x; };
  auto QY = [&](int ii, int jj) { return (ii < 1 || ii > g.m) ? 0.0 : qy[FYI(g, ii, jj)] + 0.0; };
  const double W = ext(g, w, lastbc, i, j);
  const double uavg = 0.5 * (QX(i, j) + QX(i, j - 1));
  const double vavg = 0.5 * (QY(i, j) + QY(i - 1, j));
  const double sx = (j >= 2 ? nlxb[NN(g, i, j - 1)] : 0.0) + (j <= g.n ? nlxb[NN(g, i, j)] : 0.0);
  const double sy = (i >= 2 ? nlyb[NN(g, i - 1, j)] : 0.0) + (i <= g.m ? nlyb[NN(g, i, j)] : 0.0);
  const double Wb = 0.5 * vavg * sx - 0.5 * uavg * sy;
  vb[q] = 0.5 * W * sx;
  ub[q] = -0.5 * W * sy;
  if (j == 1) lbb[g.bottom + i - 1] += Wb;
  else if (j == g.n + 1) lbb[g.top + i - 1] += Wb;
  else if (i == 1) lbb[g.left + j - 1] += Wb;
  else if (i == g.m + 1) lbb[g.right + j - 1] += Wb;
  else wb[ND(g, i, j)] += Wb;
}

// Stage 3: qxb(i, j) += 0.5 (ub(i, j) + ub(i, j+1)), qyb(i, j) += 0.5 (vb(i, j)
// + vb(i+1, j)), with uavg, vavg taken only where the forward used them
// (uavg at j = 2..n, vavg at i = 2..m).
__global__ void k_nl_adj3(Geom g, const double* __restrict__ ub, const double* __restrict__ vb, double* __restrict__ qxb,
                          double* __restrict__ qyb) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < g.Fx) {
    const int i = q % (g.m + 1) + 1, j = q / (g.m + 1) + 1;
    const double a = (j >= 2 && j <= g.n) ? ub[NN(g, i, j)] : 0.0;
    const double b = (j + 1 >= 2 && j + 1 <= g.n) ? ub[NN(g, i, j + 1)] : 0.0;
    qxb[q] += 0.5 * (a + b);
  } else if (q < g.Fx + g.Fy) {
    const int r = q - g.Fx;
    const int i = r % g.m + 1, j = r / g.m + 1;
    const double a = (i >= 2 && i <= g.m) ? vb[NN(g, i, j)] : 0.0;
    const double b = (i + 1 >= 2 && i + 1 <= g.m) ? vb[NN(g, i + 1, j)] : 0.0;
    qyb[r] += 0.5 * (a + b);
  }
}

void nl_adj(hipStream_t st, const Geom& g, const double* w, const double* qx, const double* qy, double q0x,
            const double* lastbc, const double* rb, double* nlxb, double* nlyb, double* ub, double* vb, double* wb,
            double* lbb, double* qxb, double* qyb) {
  const int NA = (g.m + 1) * (g.n + 1);
  hipLaunchKernelGGL(k_nl_adj1, dim3(ceil_div(NA, 256)), dim3(256), 0, st, g, rb, nlxb, nlyb);
  IB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_nl_adj2, dim3(ceil_div(NA, 256)), dim3(256), 0, st, g, w, qx, qy, q0x, lastbc, nlxb, nlyb, wb,
                     lbb, ub, vb);
  IB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_nl_adj3, dim3(ceil_div(g.Fx + g.Fy, 256)), dim3(256), 0, st, g, ub, vb, qxb, qyb);
  IB_LAUNCH_CHECK();
}

// v += rot(hx, hy) on the interior nodes (rot^T of curl's interior part)
__global__ void k_rot_acc(Geom g, const double* __restrict__ hx, const double* __restrict__ hy, double* __restrict__ v) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.P) return;
  const int i = q % (g.m - 1) + 2, j = q / (g.m - 1) + 2;
  v[q] += hy[FYI(g, i, j)] - hy[FYI(g, i - 1, j)] - hx[FXI(g, i, j)] + hx[FXI(g, i, j - 1)];
}

void rot_acc(hipStream_t st, const Geom& g, const double* hx, const double* hy, double* v) {
  hipLaunchKernelGGL(k_rot_acc, dim3(ceil_div(g.P, 256)), dim3(256), 0, st, g, hx, hy, v);
  IB_LAUNCH_CHECK();
}

// The predictor's right-hand side r = con1 rhs + con2 rold + ..., rold <- rhs,
// backwards: rb = ro + c1 r, ro = c2 r (c1 = con1 [+ con2 on the Euler
// step], c2 = con2 [0 on the Euler step])
__global__ void k_rhs_adj(long n, const double* __restrict__ r, double c1, double c2, double* __restrict__ ro,
                          double* __restrict__ rb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = r[i];
  rb[i] = ro[i] + c1 * v;
  ro[i] = c2 * v;
}

void rhs_adj(hipStream_t st, long n, const double* r, double c1, double c2, double* ro, double* rb) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rhs_adj, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, r, c1, c2, ro, rb);
  IB_LAUNCH_CHECK();
}

// Re terms of one predictor level, per node (summed once at the end):
//   acc += ((wsb - rb / N) ws + (rb / N - dlz) w) / Re,  wb += dlz
// with ws the predicted omega, w the step-start omega, wsb its adjoint, rb =
// D(lam1i o D wsb) and dlz = D(lam1 lam1i o D wsb), N = normalize.
__global__ void k_re_pred(long n, const double* __restrict__ ws, const double* __restrict__ w,
                          const double* __restrict__ wsb, const double* __restrict__ rb, const double* __restrict__ dlz,
                          double invN, double invRe, double* __restrict__ acc, double* __restrict__ wb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rb[i] * invN, d = dlz[i];
  acc[i] += invRe * ((wsb[i] - r) * ws[i] + (r - d) * w[i]);
  wb[i] += d;
}

void re_pred(hipStream_t st, long n, const double* ws, const double* w, const double* wsb, const double* rb,
             const double* dlz, double invN, double invRe, double* acc, double* wb) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_re_pred, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, ws, w, wsb, rb, dlz, invN,
                     invRe, acc, wb);
  IB_LAUNCH_CHECK();
}

// Re term of the corrector, per node: acc -= (wb - ainv(wb)) ainv(v) / Re
__global__ void k_re_corr(long n, const double* __restrict__ wb, const double* __restrict__ aiwb,
                          const double* __restrict__ aiv, double invRe, double* __restrict__ acc) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  acc[i] -= invRe * (wb[i] - aiwb[i]) * aiv[i];
}

void re_corr(hipStream_t st, long n, const double* wb, const double* aiwb, const double* aiv, double invRe,
             double* acc) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_re_corr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, wb, aiwb, aiv, invRe, acc);
  IB_LAUNCH_CHECK();
}

// ---- body-point positions (rigid bodies) ---------------------------------
// With A = W G W^T on the faces the taps touch (G: the level-1 flux of a
// unit face source, W: the weights), T(b, j) = sum_q w[j,q] G(b, tap[j,q])
// and U(i, b) = sum_p w[i,p] G(tap[i,p], b), both stored nf x nF with the
// entry index fastest: TT[j + nf b] = T(b, j), U[i + nf b].
__global__ void k_green_tt(int nf, int nF, const double* __restrict__ w, const int* __restrict__ tapidx,
                           const double* __restrict__ G, double* __restrict__ TT) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)nf * nF) return;
  const int j = (int)(q % nf);
  const long b = q / nf;
  double acc = 0.0;
  for (int t = 0; t < kTaps; ++t) acc += w[j * kTaps + t] * G[(long)tapidx[j * kTaps + t] * nF + b];
  TT[q] = acc;
}

__global__ void k_green_u(int nf, int nF, const double* __restrict__ w, const int* __restrict__ tapidx,
                          const double* __restrict__ G, double* __restrict__ U) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)nf * nF) return;
  const int i = (int)(q % nf);
  const long b = q / nf;
  double acc = 0.0;
  for (int t = 0; t < kTaps; ++t) acc += w[i * kTaps + t] * G[b * nF + tapidx[i * kTaps + t]];
  U[q] = acc;
}

void green_tt_u(hipStream_t st, int nf, int nF, const double* w, const int* tapidx, const double* G, double* TT,
                double* U) {
  const long n = (long)nf * nF;
  if (n <= 0) return;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_green_tt, dim3(blocks), dim3(256), 0, st, nf, nF, w, tapidx, G, TT);
  IB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_green_u, dim3(blocks), dim3(256), 0, st, nf, nF, w, tapidx, G, U);
  IB_LAUNCH_CHECK();
}

// out[k*49 + t] = (q + q0)(face of tap t of entry k): the level-1 flux the
// regularisation reads (q0 on x faces only)
__global__ void k_gather_taps(Geom g, int ntaps, const int* __restrict__ key, const double* __restrict__ qx,
                              const double* __restrict__ qy, double q0x, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntaps) return;
  const int kk = key[t];
  out[t] = kk < g.Fx ? qx[kk] + q0x : qy[kk - g.Fx] + 0.0;
}

void gather_taps(hipStream_t st, const Geom& g, int ntaps, const int* key, const double* qx, const double* qy,
                 double q0x, double* out) {
  if (ntaps <= 0) return;
  hipLaunchKernelGGL(k_gather_taps, dim3(ceil_div(ntaps, 256)), dim3(256), 0, st, g, ntaps, key, qx, qy, q0x, out);
  IB_LAUNCH_CHECK();
}

// Position terms of one step, per force entry i (one workgroup, a wavefront
// per tap): with u the adjoint of rhsf, v = f, hb = curl(ainv(omega_0 bar))
// and qt the taped (q* + q0) at the taps,
//   X(i,t) = -v_i hb(tap) + u_i qt(i,t) - Y(i,t)
//   Y(i,t) = sum_{j>i} s_j T(b, j) + sum_{j<i} s_j U(j, b) + u_i v_i (T(b, i) + U(i, b)),
//   s_j = u_i v_j + v_i u_j,  b = tap (i,t)
// (Y: the tap-weight derivative of <u, A_sym v> with A_sym the upper
// triangle of A mirrored, the matrix num_stat factors), and
//   gx[i] += sum_t dw/dx(i,t) X(i,t),  gy[i] += sum_t dw/dy(i,t) X(i,t).
__global__ void __launch_bounds__(256) k_pos_terms(Geom g, int nf, const int* __restrict__ key,
                                                   const int* __restrict__ tapidx, const double* __restrict__ dwx,
                                                   const double* __restrict__ dwy, const double* __restrict__ TT,
                                                   const double* __restrict__ U, const double* __restrict__ u,
                                                   const double* __restrict__ v, const double* __restrict__ qt,
                                                   const double* __restrict__ hbx, const double* __restrict__ hby,
                                                   double* __restrict__ gx, double* __restrict__ gy) {
  __shared__ double sx[4], sy[4];
  const int i = blockIdx.x;
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  const double ui = u[i], vi = v[i];
  double ax = 0.0, ay = 0.0;
  for (int t = wv; t < kTaps; t += 4) {
    const int q = i * kTaps + t;
    const long b = tapidx[q];
    const double* __restrict__ Tb = TT + b * nf;
    const double* __restrict__ Ub = U + b * nf;
    double y = 0.0;
    for (int j = lane; j < nf; j += 64) {
      const double sj = ui * v[j] + vi * u[j];
      if (j > i) y += sj * Tb[j];
      else if (j < i) y += sj * Ub[j];
      else y += ui * vi * (Tb[j] + Ub[j]);
    }
    for (int off = 32; off > 0; off >>= 1) y += __shfl_xor(y, off);
    const int kk = key[q];
    const double hb = kk < g.Fx ? hbx[kk] : hby[kk - g.Fx];
    const double X = -vi * hb + ui * qt[q] - y;
    ax += dwx[q] * X;
    ay += dwy[q] * X;
  }
  if (lane == 0) {
    sx[wv] = ax;
    sy[wv] = ay;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    gx[i] += ((sx[0] + sx[1]) + sx[2]) + sx[3];
    gy[i] += ((sy[0] + sy[1]) + sy[2]) + sy[3];
  }
}

void pos_terms(hipStream_t st, const Geom& g, int nf, const int* key, const int* tapidx, const double* dwx,
               const double* dwy, const double* TT, const double* U, const double* u, const double* v,
               const double* qt, const double* hbx, const double* hby, double* gx, double* gy) {
  if (nf <= 0) return;
  hipLaunchKernelGGL(k_pos_terms, dim3(nf), dim3(256), 0, st, g, nf, key, tapidx, dwx, dwy, TT, U, u, v, qt, hbx, hby,
                     gx, gy);
  IB_LAUNCH_CHECK();
}

// Elementwise helpers of the adjoint sweep
__global__ void k_axpy(long n, double a, const double* __restrict__ x, double* __restrict__ y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] += a * x[i];
}

void axpy(hipStream_t st, long n, double a, const double* x, double* y) {
  if (n <= 0) return;
  long b = (n + 255) / 256;
  if (b > 65535L * 8) b = 65535L * 8;
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)b), dim3(256), 0, st, n, a, x, y);
  IB_LAUNCH_CHECK();
}

// y = x * t (elementwise product with a table; y may be x)
__global__ void k_mul(long n, const double* x, const double* __restrict__ t, double* y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = x[i] * t[i];
}

void mul(hipStream_t st, long n, const double* x, const double* t, double* y) {
  if (n <= 0) return;
  long b = (n + 255) / 256;
  if (b > 65535L * 8) b = 65535L * 8;
  hipLaunchKernelGGL(k_mul, dim3((unsigned)b), dim3(256), 0, st, n, x, t, y);
  IB_LAUNCH_CHECK();
}

// out[0] += sum_i a[i] * b[i] (one workgroup; fixed order of the partial sums)
__global__ void __launch_bounds__(256) k_dot_acc(long n, const double* __restrict__ a, const double* __restrict__ b,
                                                 double scale, double* __restrict__ out) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) acc += a[i] * b[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] += scale * sh[0];
}

void dot_acc(hipStream_t st, long n, const double* a, const double* b, double scale, double* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_dot_acc, dim3(1), dim3(256), 0, st, n, a, b, scale, out);
  IB_LAUNCH_CHECK();
}

// out[0] += scale * sum_i x[i] (one workgroup, fixed order)
__global__ void __launch_bounds__(256) k_sum_acc(long n, const double* __restrict__ x, double scale,
                                                 double* __restrict__ out) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) acc += x[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] += scale * sh[0];
}

void sum_acc(hipStream_t st, long n, const double* x, double scale, double* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sum_acc, dim3(1), dim3(256), 0, st, n, x, scale, out);
  IB_LAUNCH_CHECK();
}

// y = A_sym x with A_sym(i, j) = A(min(i, j), max(i, j)): the matrix whose
// upper triangle the num_stat Cholesky reads (A column major)
__global__ void k_gemv_symup(int n, const double* __restrict__ A, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int j = 0; j < n; ++j) {
    const int r = min(i, j), c = max(i, j);
    acc += A[(long)c * n + r] * x[j];
  }
  y[i] = acc;
}

void gemv_symup(hipStream_t st, int n, const double* A, const double* x, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gemv_symup, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, A, x, y);
  IB_LAUNCH_CHECK();
}

}  // namespace ibg
