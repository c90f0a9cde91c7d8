// ib_gmres.hip -- device-resident restarted GMRES for the FSI Schur system.
//
// Follows PETSc 3.11 KSPGMRES as configured by setup_ksp
// (src/variables.F90:565-574): no preconditioner, unpreconditioned residual
// norm, nonzero initial guess, classical Gram-Schmidt without refinement,
// Givens-rotation residual estimate, convergence when
// ||r|| <= max(rtol*||b||, atol), divergence when ||r|| >= 1e4*||b||, happy
// breakdown tolerance 1e-30.  The Krylov basis, Hessenberg matrix and
// rotations stay in HBM; the vectors have nf entries, so every kernel here is
// one workgroup and the host only reads the small status block between
// iterations.  The Schur matvec between the kernels is the expensive part.
#include "ib_kernels.h"

namespace ibg {

constexpr int kGmT = 1024;

__device__ double gm_block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = sh[t] + sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// V0 = b - A x; residual norm, initial tolerance and convergence tests
__global__ void __launch_bounds__(kGmT) k_gm_start(int nf, const double* __restrict__ b, const double* __restrict__ Ax,
                                                   double* __restrict__ V, GmresDev* s, double* grs, double rtol,
                                                   double atol) {
  __shared__ double sh[kGmT];
  double pr = 0.0, pb = 0.0;
  for (int i = threadIdx.x; i < nf; i += blockDim.x) {
    const double r = b[i] - Ax[i];
    V[i] = r;
    pr += r * r;
    pb += b[i] * b[i];
  }
  const double res = sqrt(gm_block_sum(pr, sh));
  const double bn = sqrt(gm_block_sum(pb, sh));
  __shared__ int s_reason;
  if (threadIdx.x == 0) {
    if (s->first) {
      s->bnorm = bn;
      s->rn0 = bn > 0.0 ? bn : res;
      s->ttol = fmax(rtol * s->rn0, atol);
      s->first = 0;
    }
    grs[0] = res;
    s->res = res;
    s->it = 0;
    int reason = 0;
    if (res == 0.0 || res <= s->ttol) reason = 1;
    else if (res >= 1e4 * s->rn0) reason = -1;
    s->reason = reason;
    s_reason = reason;
  }
  __syncthreads();
  if (s_reason == 0) {
    const double inv = res;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) V[i] = V[i] / inv;
  }
}

// wavefront-level sum (64 lanes on CDNA4)
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = warpSize / 2; o > 0; o >>= 1) v += __shfl_down(v, o);
  return v;
}

// One Arnoldi step on w = V[it+1] (already A V[it]): CGS, normalisation,
// Hessenberg column update with Givens rotations (KSPGMRESUpdateHessenberg).
// The it+1 dot products are spread over the workgroup's wavefronts, one dot
// product per wavefront at a time, so no block barrier is needed per product.
__global__ void __launch_bounds__(kGmT) k_gm_arnoldi(int nf, int kmax, double* __restrict__ V, GmresDev* s,
                                                     double* __restrict__ H, double* __restrict__ cc,
                                                     double* __restrict__ ss, double* __restrict__ grs) {
  __shared__ double sh[kGmT];
  __shared__ int s_hap;
  __shared__ double s_tt;
  const int it = s->it;
  const int ld = kmax + 1;
  double* w = V + (long)(it + 1) * nf;
  double* hcol = H + (long)it * ld;
  const int lane = threadIdx.x % warpSize, wave = threadIdx.x / warpSize, nwave = blockDim.x / warpSize;
  // h_j = <w, v_j>, all from the unmodified w (classical Gram-Schmidt)
  for (int j = wave; j <= it; j += nwave) {
    const double* vj = V + (long)j * nf;
    double p = 0.0;
    for (int i = lane; i < nf; i += warpSize) p += w[i] * vj[i];
    p = wave_sum(p);
    if (lane == 0) hcol[j] = p;
  }
  __syncthreads();
  double pn = 0.0;
  for (int i = threadIdx.x; i < nf; i += blockDim.x) {
    double acc = w[i];
    for (int j = 0; j <= it; ++j) acc += -hcol[j] * V[(long)j * nf + i];
    w[i] = acc;
    pn += acc * acc;
  }
  const double tt = sqrt(gm_block_sum(pn, sh));
  if (threadIdx.x == 0) {
    hcol[it + 1] = tt;
    double hapbnd = fabs(tt / grs[it]);
    if (hapbnd > 1e-30) hapbnd = 1e-30;
    const int hap = tt < hapbnd;
    s_hap = hap;
    s_tt = tt;
    double* hh = hcol;
    for (int j = 0; j < it; ++j) {
      const double t0 = hh[j];
      hh[j] = cc[j] * t0 + ss[j] * hh[j + 1];
      hh[j + 1] = -ss[j] * t0 + cc[j] * hh[j + 1];
    }
    double res;
    if (!hap) {
      const double t1 = sqrt(hh[it] * hh[it] + hh[it + 1] * hh[it + 1]);
      cc[it] = hh[it] / t1;
      ss[it] = hh[it + 1] / t1;
      grs[it + 1] = -ss[it] * grs[it];
      grs[it] = cc[it] * grs[it];
      hh[it] = cc[it] * hh[it] + ss[it] * hh[it + 1];
      res = fabs(grs[it + 1]);
    } else {
      res = 0.0;
    }
    s->it = it + 1;
    s->its = s->its + 1;
    s->res = res;
    int reason = 0;
    if (res <= s->ttol) reason = 1;
    else if (res >= 1e4 * s->rn0) reason = -1;
    if (hap && !reason) reason = 1;
    s->reason = reason;
  }
  __syncthreads();
  if (!s_hap) {
    const double inv = 1.0 / s_tt;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) w[i] = w[i] * inv;
  }
}

// KSPGMRESBuildSoln: back substitution on the rotated Hessenberg, x += V y
__global__ void __launch_bounds__(kGmT) k_gm_buildsol(int nf, int kmax, const double* __restrict__ V, GmresDev* s,
                                                      const double* __restrict__ H, const double* __restrict__ grs,
                                                      double* __restrict__ x, double* __restrict__ nrs) {
  const int it = s->it;  // number of Arnoldi vectors used
  if (it <= 0) return;
  const int ld = kmax + 1;
  if (threadIdx.x == 0) {
    const int top = it - 1;
    nrs[top] = grs[top] / H[(long)top * ld + top];
    for (int ii = 1; ii <= top; ++ii) {
      const int k = top - ii;
      double t = grs[k];
      for (int j = k + 1; j <= top; ++j) t = t - H[(long)j * ld + k] * nrs[j];
      nrs[k] = t / H[(long)k * ld + k];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nf; i += blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < it; ++j) acc += nrs[j] * V[(long)j * nf + i];
    x[i] = x[i] + acc;
  }
}

void gm_start(hipStream_t st, int nf, int kmax, const double* b, const double* Ax, double* V, GmresDev* s, double* H,
              double* grs, double rtol, double atol) {
  (void)kmax;
  (void)H;
  hipLaunchKernelGGL(k_gm_start, dim3(1), dim3(kGmT), 0, st, nf, b, Ax, V, s, grs, rtol, atol);
  IB_LAUNCH_CHECK();
}

void gm_arnoldi(hipStream_t st, int nf, int kmax, double* V, GmresDev* s, double* H, double* cc, double* ss,
                double* grs) {
  hipLaunchKernelGGL(k_gm_arnoldi, dim3(1), dim3(kGmT), 0, st, nf, kmax, V, s, H, cc, ss, grs);
  IB_LAUNCH_CHECK();
}

void gm_buildsol(hipStream_t st, int nf, int kmax, const double* V, GmresDev* s, const double* H, const double* grs,
                 double* x, double* scratch) {
  hipLaunchKernelGGL(k_gm_buildsol, dim3(1), dim3(kGmT), 0, st, nf, kmax, V, s, H, grs, x, scratch);
  IB_LAUNCH_CHECK();
}

}  // namespace ibg
