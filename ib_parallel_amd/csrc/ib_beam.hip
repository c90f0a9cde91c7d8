// ib_beam.hip -- family-3 bodies: a deformable plate as a linear
// Euler-Bernoulli beam.  The reference declares the family and leaves its
// routines stubs (src/operators_struct.F90:104-106, 146-148, 376-379,
// 472-474), so the model is new and follows the torsional body's coupling
// (b_times2, compute_rhs, compute_dx, initial_acc) with the angle theta
// replaced by the beam's unknowns u and BR(theta) by Phi = n H:
//  * npts = 2 ne + 1 points along the deck's straight line, ne Hermite-cubic
//    elements with nodes at the even points, the first point clamped; u
//    holds each free node's deflection normal to the line and its slope;
//  * H(p, i): the shape functions at point p; the points sit at X0 + n (H u);
//  * Newmark average acceleration: K_eff = 4/dt^2 M + 2/dt C + K, the
//    generalised force delta/dt Phi^T f;
//  * S gains 2 delta^2/dt^2 Phi K_eff^-1 Phi^T on the beam's entries, and
//    the FSI right side delta Phi (ud0 + 2/dt (u0 - u) - Jr), the normal
//    velocity the points take with f = 0.
// Each kernel is one workgroup: the beam has a few hundred unknowns.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ib_device.h"
#include "ib_kernels.h"

namespace ibg {

namespace {

// z = M (4/dt^2 u0 + 4/dt ud0 + udd0) + C (ud0 + 2/dt u0) - K_eff u
__device__ void beam_z(const BeamDev& b, double dt, double* z) {
  const int nd = b.nd;
  for (int i = threadIdx.x; i < nd; i += blockDim.x) {
    double acc = 0.0;
    for (int k = 0; k < nd; ++k) {
      const size_t ik = (size_t)i * nd + k;
      acc += b.M[ik] * (4.0 / (dt * dt) * b.u0[k] + 4.0 / dt * b.ud0[k] + b.udd0[k]) +
             b.C[ik] * (b.ud0[k] + 2.0 / dt * b.u0[k]) - b.Keff[ik] * b.u[k];
    }
    z[i] = acc;
  }
}

// q = delta/dt Phi^T f on the beam's entries (fn: the normal point forces)
__device__ double beam_q(const BeamDev& b, const double* fn, int i, double dt, double delta) {
  double acc = 0.0;
  for (int p = 0; p < b.np; ++p) acc += b.H[(size_t)p * b.nd + i] * fn[p];
  return acc * (delta / dt);
}

__global__ void __launch_bounds__(256) k_beam_rhs(BeamDev b, double dt, double delta, double* rhsf) {
  extern __shared__ double sh[];
  double* z = sh;          // nd
  double* v = sh + b.nd;   // nd
  beam_z(b, dt, z);
  __syncthreads();
  for (int i = threadIdx.x; i < b.nd; i += blockDim.x) {
    double jr = 0.0;
    for (int k = 0; k < b.nd; ++k) jr += b.Keffi[(size_t)i * b.nd + k] * z[k];
    v[i] = b.ud0[i] + 2.0 / dt * (b.u0[i] - b.u[i]) - 2.0 / dt * jr;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < b.np; p += blockDim.x) {
    double w = 0.0;
    for (int i = 0; i < b.nd; ++i) w += b.H[(size_t)p * b.nd + i] * v[i];
    rhsf[b.off + p] += delta * b.nx * w;
    rhsf[b.off + b.np + p] += delta * b.ny * w;
  }
}

__global__ void __launch_bounds__(256) k_beam_dx(BeamDev b, double dt, double delta, const double* f, double* xb,
                                                 double* yb, double* du_abs) {
  extern __shared__ double sh[];
  double* z = sh;                // nd
  double* fn = sh + b.nd;        // np
  double* red = fn + b.np;       // blockDim.x
  double* un = red + blockDim.x; // nd: the new u
  for (int p = threadIdx.x; p < b.np; p += blockDim.x) fn[p] = b.nx * f[b.off + p] + b.ny * f[b.off + b.np + p];
  beam_z(b, dt, z);
  __syncthreads();
  for (int i = threadIdx.x; i < b.nd; i += blockDim.x) z[i] += beam_q(b, fn, i, dt, delta);
  __syncthreads();
  double mx = 0.0;
  for (int i = threadIdx.x; i < b.nd; i += blockDim.x) {
    double du = 0.0;
    for (int k = 0; k < b.nd; ++k) du += b.Keffi[(size_t)i * b.nd + k] * z[k];
    mx = fmax(mx, fabs(du));
    const double u = b.u[i] + du;
    un[i] = u;
    b.u[i] = u;
    b.ud[i] = -b.ud0[i] + 2.0 / dt * (u - b.u0[i]);
    b.udd[i] = -b.udd0[i] - 4.0 / dt * b.ud0[i] + 4.0 / (dt * dt) * (u - b.u0[i]);
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *du_abs = red[0];
  for (int p = threadIdx.x; p < b.np; p += blockDim.x) {
    double w = 0.0;
    for (int i = 0; i < b.nd; ++i) w += b.H[(size_t)p * b.nd + i] * un[i];
    const double X = b.X0[p] + b.nx * w, Y = b.Y0[p] + b.ny * w;
    xb[b.off + p] = X;
    xb[b.off + b.np + p] = X;
    yb[b.off + p] = Y;
    yb[b.off + b.np + p] = Y;
  }
}

// step 1 only: udd = M^-1 (delta/dt Phi^T f - C ud - K u)
__global__ void __launch_bounds__(256) k_beam_initial_acc(BeamDev b, double dt, double delta, const double* f) {
  extern __shared__ double sh[];
  double* r = sh;            // nd
  double* fn = sh + b.nd;    // np
  for (int p = threadIdx.x; p < b.np; p += blockDim.x) fn[p] = b.nx * f[b.off + p] + b.ny * f[b.off + b.np + p];
  __syncthreads();
  for (int i = threadIdx.x; i < b.nd; i += blockDim.x) {
    double cu = 0.0, ku = 0.0;
    for (int k = 0; k < b.nd; ++k) {
      cu += b.C[(size_t)i * b.nd + k] * b.ud[k];
      ku += b.K[(size_t)i * b.nd + k] * b.u[k];
    }
    r[i] = beam_q(b, fn, i, dt, delta) - (cu + ku);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < b.nd; i += blockDim.x) {
    double a = 0.0;
    for (int k = 0; k < b.nd; ++k) a += b.Minv[(size_t)i * b.nd + k] * r[k];
    b.udd[i] = a;
  }
}

__global__ void __launch_bounds__(256) k_beam_save0(BeamDev b) {
  for (int i = threadIdx.x; i < b.nd; i += blockDim.x) {
    b.u0[i] = b.u[i];
    b.ud0[i] = b.ud[i];
    b.udd0[i] = b.udd[i];
  }
}

// S(j, k) += c n_j n_k G(p_j, p_k) on the beam's 2 np entries (S column-major)
__global__ void k_beam_add_schur(BeamDev b, int nf, double c, double* S) {
  const int n2 = 2 * b.np;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (j >= n2 || k >= n2) return;
  const int pj = j % b.np, pk = k % b.np;
  const double nj = j < b.np ? b.nx : b.ny, nk = k < b.np ? b.nx : b.ny;
  S[(size_t)(b.off + k) * nf + b.off + j] += c * (nj * nk * b.G[(size_t)pj * b.np + pk]);
}

// A^-1 of a dense n x n matrix (row-major) by LU with partial pivoting and
// n triangular solves; false if singular
bool lu_inverse(int n, const std::vector<double>& a_in, std::vector<double>& inv) {
  std::vector<double> a = a_in;
  std::vector<int> piv(n);
  for (int c = 0; c < n; ++c) {
    int r = c;
    for (int i = c + 1; i < n; ++i)
      if (std::fabs(a[(size_t)i * n + c]) > std::fabs(a[(size_t)r * n + c])) r = i;
    piv[c] = r;
    if (a[(size_t)r * n + c] == 0.0) return false;
    if (r != c)
      for (int k = 0; k < n; ++k) std::swap(a[(size_t)r * n + k], a[(size_t)c * n + k]);
    for (int i = c + 1; i < n; ++i) {
      const double l = a[(size_t)i * n + c] / a[(size_t)c * n + c];
      a[(size_t)i * n + c] = l;
      for (int k = c + 1; k < n; ++k) a[(size_t)i * n + k] -= l * a[(size_t)c * n + k];
    }
  }
  inv.assign((size_t)n * n, 0.0);
  std::vector<double> col(n);
  for (int j = 0; j < n; ++j) {
    std::fill(col.begin(), col.end(), 0.0);
    col[j] = 1.0;
    for (int c = 0; c < n; ++c) std::swap(col[c], col[piv[c]]);
    for (int i = 1; i < n; ++i)
      for (int k = 0; k < i; ++k) col[i] -= a[(size_t)i * n + k] * col[k];
    for (int i = n - 1; i >= 0; --i) {
      for (int k = i + 1; k < n; ++k) col[i] -= a[(size_t)i * n + k] * col[k];
      col[i] /= a[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + j] = col[i];
  }
  return true;
}

}  // namespace

bool beam_host_setup(const double* x, const double* y, int npts, double mass, double EI, double damp, double dt,
                     BeamHost* out) {
  BeamHost& b = *out;
  b.np = npts;
  b.ne = (npts - 1) / 2;
  b.nd = 2 * b.ne;
  const double lx = x[npts - 1] - x[0], ly = y[npts - 1] - y[0];
  const double L = std::sqrt(lx * lx + ly * ly), h = L / b.ne;
  b.nx = -ly / L;
  b.ny = lx / L;
  const int nd = b.nd;
  b.M.assign((size_t)nd * nd, 0.0);
  b.K.assign((size_t)nd * nd, 0.0);
  // element matrices of the Hermite cubic on [0, h], dofs (w0, w0', w1, w1')
  const double kf = EI / (h * h * h), mf = mass * h / 420.0;
  const double ke[16] = {12, 6 * h, -12, 6 * h, 6 * h, 4 * h * h, -6 * h, 2 * h * h,
                         -12, -6 * h, 12, -6 * h, 6 * h, 2 * h * h, -6 * h, 4 * h * h};
  const double me[16] = {156, 22 * h, 54, -13 * h, 22 * h, 4 * h * h, 13 * h, -3 * h * h,
                         54, 13 * h, 156, -22 * h, -13 * h, -3 * h * h, -22 * h, 4 * h * h};
  for (int e = 0; e < b.ne; ++e) {
    int g[4];
    for (int a = 0; a < 4; ++a) {
      const int node = e + a / 2;
      g[a] = node == 0 ? -1 : 2 * (node - 1) + (a & 1);  // node 0 clamped
    }
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 4; ++c)
        if (g[a] >= 0 && g[c] >= 0) {
          b.K[(size_t)g[a] * nd + g[c]] += kf * ke[4 * a + c];
          b.M[(size_t)g[a] * nd + g[c]] += mf * me[4 * a + c];
        }
  }
  b.C.resize(b.M.size());
  b.Keff.resize(b.M.size());
  for (size_t i = 0; i < b.M.size(); ++i) {
    b.C[i] = damp * b.M[i];
    b.Keff[i] = 4.0 / (dt * dt) * b.M[i] + 2.0 / dt * b.C[i] + b.K[i];
  }
  if (!lu_inverse(nd, b.Keff, b.Keffi) || !lu_inverse(nd, b.M, b.Minv)) return false;
  b.H.assign((size_t)npts * nd, 0.0);
  for (int p = 0; p < npts; ++p) {
    const int e = std::min(p / 2, b.ne - 1);
    const double s = 0.5 * (double)(p - 2 * e), s2 = s * s, s3 = s2 * s;
    const double N[4] = {1 - 3 * s2 + 2 * s3, h * (s - 2 * s2 + s3), 3 * s2 - 2 * s3, h * (s3 - s2)};
    for (int a = 0; a < 4; ++a) {
      const int node = e + a / 2;
      if (node > 0) b.H[(size_t)p * nd + 2 * (node - 1) + (a & 1)] += N[a];
    }
  }
  // G = H K_eff^-1 H^T
  std::vector<double> T((size_t)nd * npts, 0.0);  // K_eff^-1 H^T, nd x np
  for (int i = 0; i < nd; ++i)
    for (int q = 0; q < npts; ++q) {
      double acc = 0.0;
      for (int k = 0; k < nd; ++k) acc += b.Keffi[(size_t)i * nd + k] * b.H[(size_t)q * nd + k];
      T[(size_t)i * npts + q] = acc;
    }
  b.G.assign((size_t)npts * npts, 0.0);
  for (int p = 0; p < npts; ++p)
    for (int q = 0; q < npts; ++q) {
      double acc = 0.0;
      for (int i = 0; i < nd; ++i) acc += b.H[(size_t)p * nd + i] * T[(size_t)i * npts + q];
      b.G[(size_t)p * npts + q] = acc;
    }
  b.X0.assign(x, x + npts);
  b.Y0.assign(y, y + npts);
  return true;
}

void beam_rhs(hipStream_t st, const BeamDev& b, double dt, double delta, double* rhsf) {
  hipLaunchKernelGGL(k_beam_rhs, dim3(1), dim3(256), sizeof(double) * 2 * b.nd, st, b, dt, delta, rhsf);
  IB_LAUNCH_CHECK();
}

void beam_dx(hipStream_t st, const BeamDev& b, double dt, double delta, const double* f, double* xb, double* yb,
             double* du_abs) {
  hipLaunchKernelGGL(k_beam_dx, dim3(1), dim3(256), sizeof(double) * (2 * b.nd + b.np + 256), st, b, dt, delta, f,
                     xb, yb, du_abs);
  IB_LAUNCH_CHECK();
}

void beam_initial_acc(hipStream_t st, const BeamDev& b, double dt, double delta, const double* f) {
  hipLaunchKernelGGL(k_beam_initial_acc, dim3(1), dim3(256), sizeof(double) * (b.nd + b.np), st, b, dt, delta, f);
  IB_LAUNCH_CHECK();
}

void beam_save0(hipStream_t st, const BeamDev& b) {
  hipLaunchKernelGGL(k_beam_save0, dim3(1), dim3(256), 0, st, b);
  IB_LAUNCH_CHECK();
}

void beam_add_schur(hipStream_t st, const BeamDev& b, int nf, double dt, double delta, double* S) {
  const int n2 = 2 * b.np;
  hipLaunchKernelGGL(k_beam_add_schur, dim3((n2 + 255) / 256, n2), dim3(256), 0, st, b, nf,
                     2.0 * delta * delta / (dt * dt), S);
  IB_LAUNCH_CHECK();
}

}  // namespace ibg
