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

// the node value apply_bc adds slot `slot` to (0 for slots no node reads)
__device__ __forceinline__ double applybc_node(const Geom& g, const double* __restrict__ vb, int slot) {
  if (slot < g.right) { const int j = slot + 1; return (j >= 2 && j <= g.n) ? vb[ND(g, 2, j)] : 0.0; }
  if (slot < g.bottom) { const int j = slot - g.right + 1; return (j >= 2 && j <= g.n) ? vb[ND(g, g.m, j)] : 0.0; }
  if (slot < g.top) { const int i = slot - g.bottom + 1; return (i >= 2 && i <= g.m) ? vb[ND(g, i, 2)] : 0.0; }
  const int i = slot - g.top + 1;
  return (i >= 2 && i <= g.m) ? vb[ND(g, i, g.n)] : 0.0;
}

// bcb += apply_bc^T(vb) (the transposed vort2flux's boundary adjoint)
__global__ void k_applybc_adj_acc(Geom g, const double* __restrict__ vb, double* __restrict__ bcb) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= g.nbc) return;
  bcb[slot] += applybc_node(g, vb, slot);
}

void applybc_adj_acc(hipStream_t st, const Geom& g, const double* vb, double* bcb) {
  hipLaunchKernelGGL(k_applybc_adj_acc, dim3(ceil_div(g.nbc, 256)), dim3(256), 0, st, g, vb, bcb);
  IB_LAUNCH_CHECK();
}

// The predictor's boundary terms of one level, backwards: with
// v = apply_bc^T(rb) (rb the adjoint of r),
//   bcb = v,  lbb += vfac v,  racc_bc += -vfac/Re v (lastbc + omegabc)
// (the latter the Re term of vfac, summed once at the end of the sweep).
__global__ void k_pred_bc_adj(Geom g, const double* __restrict__ rb, double vfac, double invRe,
                              const double* __restrict__ lastbc, const double* __restrict__ omegabc,
                              double* __restrict__ bcb, double* __restrict__ lbb, double* __restrict__ racc_bc) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= g.nbc) return;
  const double v = applybc_node(g, rb, slot);
  bcb[slot] = v;
  lbb[slot] += vfac * v;
  racc_bc[slot] += -vfac * invRe * (v * (lastbc[slot] + omegabc[slot]));
}

void pred_bc_adj(hipStream_t st, const Geom& g, const double* rb, double vfac, double invRe, const double* lastbc,
                 const double* omegabc, double* bcb, double* lbb, double* racc_bc) {
  hipLaunchKernelGGL(k_pred_bc_adj, dim3(ceil_div(g.nbc, 256)), dim3(256), 0, st, g, rb, vfac, invRe, lastbc, omegabc,
                     bcb, lbb, racc_bc);
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

// ---- torsional bodies (family 2) in the face-space form -----------------
// d w / d theta of every (entry, tap) of torsional body entries: the points
// rotate about the hinge, x = xbf0 + (j-1) ds cos(gamma0 - theta), y = ybf0 +
// (j-1) ds sin(gamma0 - theta) (update_xb, src/operators_struct.F90:600-620),
// so d x/d theta = (j-1) ds sin(gamma0 - theta), d y/d theta = -(j-1) ds
// cos(gamma0 - theta).  Zero for the entries of rigid bodies.
__global__ void k_dthw(int nf, int nt, const int* __restrict__ tors_of, const TorsBody* __restrict__ tbs,
                       const double* __restrict__ th, const double* __restrict__ dwx, const double* __restrict__ dwy,
                       double* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nf * kTaps) return;
  const int i = q / kTaps;
  const int t = tors_of[i];
  double v = 0.0;
  if (t >= 0) {
    const TorsBody b = tbs[t];
    const int jl = i - b.off;  // 0 .. 2 npts - 1
    const int jj = jl < b.npts ? jl : jl - b.npts;
    const double ph = b.gamma0 - th[t];
    v = dwx[q] * ((double)jj * b.deltas * sin(ph)) - dwy[q] * ((double)jj * b.deltas * cos(ph));
  }
  out[q] = v;
}

void dthw(hipStream_t st, int nf, int nt, const int* tors_of, const TorsBody* tb, const double* th, const double* dwx,
          const double* dwy, double* out) {
  if (nf <= 0) return;
  hipLaunchKernelGGL(k_dthw, dim3(ceil_div(nf * kTaps, 256)), dim3(256), 0, st, nf, nt, tors_of, tb, th, dwx, dwy,
                     out);
  IB_LAUNCH_CHECK();
}

// out[b] = face field at face Fkeys[b] (x faces, then Fx + y faces)
__global__ void k_face_gather(Geom g, int nF, const int* __restrict__ fkeys, const double* __restrict__ hx,
                              const double* __restrict__ hy, double* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nF) return;
  const int k = fkeys[b];
  out[b] = k < g.Fx ? hx[k] : hy[k - g.Fx];
}

void face_gather(hipStream_t st, const Geom& g, int nF, const int* fkeys, const double* hx, const double* hy,
                 double* out) {
  if (nF <= 0) return;
  hipLaunchKernelGGL(k_face_gather, dim3(ceil_div(nF, 256)), dim3(256), 0, st, g, nF, fkeys, hx, hy, out);
  IB_LAUNCH_CHECK();
}

// out[i] = sum_t w[i,t] zf[tapidx[i,t]] (E applied to a face-set vector)
__global__ void k_face_regT(int nf, const double* __restrict__ w, const int* __restrict__ tapidx,
                            const double* __restrict__ zf, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  double acc = 0.0;
  for (int t = 0; t < kTaps; ++t) acc += w[i * kTaps + t] * zf[tapidx[i * kTaps + t]];
  out[i] = acc;
}

void face_regT(hipStream_t st, int nf, const double* w, const int* tapidx, const double* zf, double* out) {
  if (nf <= 0) return;
  hipLaunchKernelGGL(k_face_regT, dim3(ceil_div(nf, 256)), dim3(256), 0, st, nf, w, tapidx, zf, out);
  IB_LAUNCH_CHECK();
}

// out[i] += coef[i] * sum_t wt[i,t] fld[i,t] for entries i0 <= i < i1, with
// fld either a per-tap array (ftap) or a face-set vector read through tapidx
// minus a per-tap array (fset[tapidx] - ftap)
__global__ void k_tap_contract(int i0, int i1, const double* __restrict__ wt, const double* __restrict__ coef,
                               const double* __restrict__ ftap, const double* __restrict__ fset,
                               const int* __restrict__ tapidx, double* __restrict__ out) {
  const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= i1) return;
  double acc = 0.0;
  for (int t = 0; t < kTaps; ++t) {
    const int q = i * kTaps + t;
    const double v = fset ? fset[tapidx[q]] - (ftap ? ftap[q] : 0.0) : ftap[q];
    acc += wt[q] * v;
  }
  out[i] += (coef ? coef[i] : 1.0) * acc;
}

void tap_contract(hipStream_t st, int i0, int i1, const double* wt, const double* coef, const double* ftap,
                  const double* fset, const int* tapidx, double* out) {
  if (i1 <= i0) return;
  hipLaunchKernelGGL(k_tap_contract, dim3(ceil_div(i1 - i0, 256)), dim3(256), 0, st, i0, i1, wt, coef, ftap, fset,
                     tapidx, out);
  IB_LAUNCH_CHECK();
}

// y = G^T x for a square column-major n x n matrix (a wavefront per column)
__global__ void __launch_bounds__(256) k_gemv_t(int n, const double* __restrict__ G, const double* __restrict__ x,
                                                double* __restrict__ y) {
  const int j = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x % 64;
  if (j >= n) return;
  const double* __restrict__ c = G + (long)j * n;
  double acc = 0.0;
  for (int b = lane; b < n; b += 64) acc += c[b] * x[b];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) y[j] = acc;
}

void gemv_t(hipStream_t st, int n, const double* G, const double* x, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gemv_t, dim3(ceil_div(n, 4)), dim3(256), 0, st, n, G, x, y);
  IB_LAUNCH_CHECK();
}

// AT = A^T (n x n, column major)
__global__ void k_transpose(int n, const double* __restrict__ A, double* __restrict__ AT) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 256 threads: 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int i = bx + tx, j = by + r;
    if (i < n && j < n) tile[r][tx] = A[(long)j * n + i];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = by + tx, j = bx + r;
    if (i < n && j < n) AT[(long)j * n + i] = tile[tx][r];
  }
}

void transpose(hipStream_t st, int n, const double* A, double* AT) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_transpose, dim3(ceil_div(n, 32), ceil_div(n, 32)), dim3(256), 0, st, n, A, AT);
  IB_LAUNCH_CHECK();
}

// One torsional body t, one workgroup (theta = th[t], or theta0 with at0):
// the scalars of its block,
//   s[0] = BR^T f, s[1] = BR'^T f, s[2] = BR^T a (a given) ,
// and, if r1 is given, r1 += (sc_b/den)(BR' (BR^T f) + BR (BR'^T f)) - sc BR'
// (the theta derivative of b_times2 applied to f and of the structural
// right-hand side BR sc), and, if zc is given, zc = c2 BR on the body's
// entries (zero elsewhere is left to the caller).  BR (BR_build,
// src/operators_struct.F90:319-338) and BR' = d BR / d theta at theta = th[t];
// sc = -delta ((2/dt) Q - thetad0 - 2 theta0/dt) with Q the Newmark
// predictor of compute_rhs (src/operators_struct.F90:342-398) at the
// consistent rates; c2 = delta / (dt (4M/dt^2 + K)) (compute_dx's
// denominator, :448).
__device__ __forceinline__ double brv(const TorsBody& b, int j, double si, double ci) {
  return j <= b.npts ? (double)(j - 1) * b.deltas * si : -(double)(j - b.npts - 1) * b.deltas * ci;
}
__device__ __forceinline__ double brp(const TorsBody& b, int j, double si, double ci) {
  return j <= b.npts ? -(double)(j - 1) * b.deltas * ci : -(double)(j - b.npts - 1) * b.deltas * si;
}

__global__ void __launch_bounds__(256) k_tors_block(const TorsBody* __restrict__ tbs, const double* __restrict__ th,
                                                    int nt, int t, int at0, double dt, double delta,
                                                    const double* __restrict__ f, const double* __restrict__ a,
                                                    double* __restrict__ r1, double* __restrict__ zc,
                                                    double* __restrict__ s) {
  __shared__ double sh[3][256];
  const TorsBody b = tbs[t];
  const double theta = at0 ? th[3 * nt + t] : th[t];
  const double si = sin(b.gamma0 - theta), ci = cos(b.gamma0 - theta);
  double p0 = 0.0, p1 = 0.0, p2 = 0.0;
  for (int j = threadIdx.x + 1; j <= 2 * b.npts; j += blockDim.x) {
    const double fv = f[b.off + j - 1];
    p0 += brv(b, j, si, ci) * fv;
    p1 += brp(b, j, si, ci) * fv;
    if (a) p2 += brv(b, j, si, ci) * a[b.off + j - 1];
  }
  sh[0][threadIdx.x] = p0;
  sh[1][threadIdx.x] = p1;
  sh[2][threadIdx.x] = p2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
    __syncthreads();
  }
  const double brf = sh[0][0], brpf = sh[1][0];
  if (threadIdx.x == 0 && s) {
    s[0] = brf;
    s[1] = brpf;
    s[2] = sh[2][0];
  }
  const double den = 4.0 / (dt * dt) * b.M + 2.0 / dt * b.C + b.K;
  const double scb = (delta * delta / b.deltas / dt) * (2 * b.deltas / dt);
  const double th0 = th[3 * nt + t], thd0 = th[4 * nt + t], thdd0 = th[5 * nt + t];
  const double Q = (b.M * (4.0 / (dt * dt) * th0 + 4.0 / dt * thd0 + thdd0) + b.C * (thd0 + 2.0 / dt * th0)) / den;
  const double sc = -delta * (2.0 / dt * Q - thd0 - 2.0 / dt * th0);
  const double c2 = delta / (dt * (4.0 / (dt * dt) * b.M + b.K));
  for (int j = threadIdx.x + 1; j <= 2 * b.npts; j += blockDim.x) {
    const double B = brv(b, j, si, ci), Bp = brp(b, j, si, ci);
    if (r1) r1[b.off + j - 1] += scb / den * (Bp * brf + B * brpf) - sc * Bp;
    if (zc) zc[b.off + j - 1] = c2 * B;
  }
}

void tors_block(hipStream_t st, const TorsBody* tb, const double* th, int nt, int t, int at0, double dt, double delta,
                const double* f, const double* a, double* r1, double* zc, double* s) {
  hipLaunchKernelGGL(k_tors_block, dim3(1), dim3(256), 0, st, tb, th, nt, t, at0, dt, delta, f, a, r1, zc, s);
  IB_LAUNCH_CHECK();
}

// Rigid-body point terms of one step on the FSI path (S = W G W^T + B
// solved as it is): with l the adjoint of rhsf, f, hb = curl(ainv(omega_0
// bar)) and qt = q* + q0 at the taps, vT = G W^T f and uT = G^T W^T l on the
// face set,
//   X(i,t) = -f_i (hb(i,t) + uT[tap]) - l_i (vT[tap] - qt(i,t)),
//   gx[i] += sum_t dw/dx(i,t) X(i,t), gy[i] likewise (rigid entries only).
__global__ void k_pos_fsi(int nf, const int* __restrict__ tors_of, const int* __restrict__ tapidx,
                          const double* __restrict__ dwx, const double* __restrict__ dwy, const double* __restrict__ l,
                          const double* __restrict__ f, const double* __restrict__ hb, const double* __restrict__ qt,
                          const double* __restrict__ vT, const double* __restrict__ uT, double* __restrict__ gx,
                          double* __restrict__ gy, double* __restrict__ sx, double* __restrict__ sy) {
  // every entry, torsional ones included: their sum over a torsional body
  // is the gradient with respect to its hinge (a hinge shift translates the
  // body's points and changes neither the arms nor the coupling vector)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  (void)tors_of;
  double ax = 0.0, ay = 0.0;
  for (int t = 0; t < kTaps; ++t) {
    const int q = i * kTaps + t;
    const int b = tapidx[q];
    const double X = -f[i] * (hb[q] + uT[b]) - l[i] * (vT[b] - qt[q]);
    ax += dwx[q] * X;
    ay += dwy[q] * X;
  }
  gx[i] += ax;
  gy[i] += ay;
  if (sx) {  // this step's partials alone
    sx[i] = ax;
    sy[i] = ay;
  }
}

void pos_fsi(hipStream_t st, int nf, const int* tors_of, const int* tapidx, const double* dwx, const double* dwy,
             const double* l, const double* f, const double* hb, const double* qt, const double* vT, const double* uT,
             double* gx, double* gy, double* sx, double* sy) {
  if (nf <= 0) return;
  hipLaunchKernelGGL(k_pos_fsi, dim3(ceil_div(nf, 256)), dim3(256), 0, st, nf, tors_of, tapidx, dwx, dwy, l, f, hb, qt,
                     vT, uT, gx, gy, sx, sy);
  IB_LAUNCH_CHECK();
}

// The weights' share of dJ/d(ds) for torsional body t at this step: its
// points sit at the hinge + (j-1) ds (cos, sin)(gamma0 - theta), so the
// step's position partials (sx, sy of every entry) projected on
// (j-1)(cos, sin)(gamma0 - theta), summed in a fixed order and added to
// acc[t] (one workgroup per body).
__global__ void __launch_bounds__(256) k_ds_weights(const TorsBody* __restrict__ tbs, const double* __restrict__ th,
                                                    const double* __restrict__ sx, const double* __restrict__ sy,
                                                    double* __restrict__ acc) {
  __shared__ double sh[256];
  const int t = blockIdx.x;
  const TorsBody b = tbs[t];
  const double theta = th[t];
  const double c = cos(b.gamma0 - theta), s = sin(b.gamma0 - theta);
  double p = 0.0;
  for (int e = threadIdx.x; e < 2 * b.npts; e += blockDim.x) {
    const int j1 = e < b.npts ? e : e - b.npts;  // j - 1 of the entry's point
    p += (double)j1 * (sx[b.off + e] * c + sy[b.off + e] * s);
  }
  sh[threadIdx.x] = p;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[t] += sh[0];
}

void ds_weights(hipStream_t st, const TorsBody* tb, const double* th, int nt, const double* sx, const double* sy,
                double* acc) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_ds_weights, dim3(nt), dim3(256), 0, st, tb, th, sx, sy, acc);
  IB_LAUNCH_CHECK();
}

// out[0] = sum_i a[i] b[i] over [i0, i1) (written, one workgroup)
__global__ void __launch_bounds__(256) k_dot_set(int i0, int i1, const double* __restrict__ a,
                                                 const double* __restrict__ b, double* __restrict__ out) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) acc += a[i] * b[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

void dot_set(hipStream_t st, int i0, int i1, const double* a, const double* b, double* out) {
  hipLaunchKernelGGL(k_dot_set, dim3(1), dim3(256), 0, st, i0, i1, a, b, out);
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

// S(i, j) = A(min(i, j), max(i, j)) (column major): the symmetric matrix
// whose upper triangle is A's
__global__ void k_symmetrize_upper(int n, const double* __restrict__ A, double* __restrict__ S) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)n * n) return;
  const int i = (int)(q % n), j = (int)(q / n);
  S[q] = A[(long)max(i, j) * n + min(i, j)];
}

void symmetrize_upper(hipStream_t st, int n, const double* A, double* S) {
  if (n <= 0) return;
  const long nn = (long)n * n;
  hipLaunchKernelGGL(k_symmetrize_upper, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, n, A, S);
  IB_LAUNCH_CHECK();
}

void gemv_symup(hipStream_t st, int n, const double* A, const double* x, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gemv_symup, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, A, x, y);
  IB_LAUNCH_CHECK();
}

}  // namespace ibg
