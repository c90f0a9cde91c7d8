// ib_gmres.hip -- device-resident restarted GMRES for the FSI Schur system.
//
// Follows PETSc 3.11 KSPGMRES as configured by setup_ksp
// (src/variables.F90:565-574): no preconditioner, unpreconditioned residual
// norm, nonzero initial guess, classical Gram-Schmidt without refinement,
// Givens-rotation residual estimate, convergence when
// ||r|| <= max(rtol*||b||, atol), divergence when ||r|| >= 1e4*||b||, happy
// breakdown tolerance 1e-30.  The Krylov basis, Hessenberg matrix and
// rotations stay in HBM; the host reads the small status block only.
//
// Every kernel returns at once when the status block already carries a
// convergence reason, so the host may enqueue several iterations ahead of
// the status it has read (the extra matvecs then write only into basis
// columns the solution never uses).
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

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = warpSize / 2; o > 0; o >>= 1) v += __shfl_down(v, o);
  return v;
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

// h_j = <w, v_j> for j = 0..it, w = V[it+1] unmodified (classical
// Gram-Schmidt); one wavefront per dot product, many workgroups.
__global__ void __launch_bounds__(256) k_gm_dots(int nf, int kmax, const double* __restrict__ V,
                                                 const GmresDev* __restrict__ s, double* __restrict__ H) {
  if (s->reason) return;
  const int it = s->it;
  const int j = (blockIdx.x * blockDim.x + threadIdx.x) / warpSize;
  const int lane = threadIdx.x % warpSize;
  if (j > it) return;
  const double* w = V + (long)(it + 1) * nf;
  const double* vj = V + (long)j * nf;
  double p = 0.0;
  for (int i = lane; i < nf; i += warpSize) p += w[i] * vj[i];
  p = wave_sum(p);
  if (lane == 0) H[(long)it * (kmax + 1) + j] = p;
}

// w -= sum_j h_j v_j, ||w||, normalisation and the Givens update of the new
// Hessenberg column (KSPGMRESUpdateHessenberg) on LDS copies.
__global__ void __launch_bounds__(kGmT) k_gm_update(int nf, int kmax, double* __restrict__ V, GmresDev* s,
                                                    double* __restrict__ H, double* __restrict__ cc,
                                                    double* __restrict__ ss, double* __restrict__ grs) {
  __shared__ double sh[kGmT];
  extern __shared__ double dyn[];  // hcol[kmax+1] | cc[kmax] | ss[kmax]
  __shared__ int s_hap, s_skip;
  __shared__ double s_tt;
  if (threadIdx.x == 0) s_skip = s->reason != 0;
  __syncthreads();
  if (s_skip) return;
  const int it = s->it;
  const int ld = kmax + 1;
  double* hcol = dyn;
  double* lcc = dyn + ld;
  double* lss = lcc + kmax;
  double* w = V + (long)(it + 1) * nf;
  double* hg = H + (long)it * ld;
  for (int j = threadIdx.x; j <= it; j += blockDim.x) {
    hcol[j] = hg[j];
    if (j < it) {
      lcc[j] = cc[j];
      lss[j] = ss[j];
    }
  }
  __syncthreads();
  // w_i -= sum_j h_j v_j[i]: the j range is split over `nseg` thread groups
  // (partials in LDS, summed in segment order) to keep more loads in flight
  double* part = sh;  // reuses the reduction buffer: nseg * nf <= kGmT doubles
  const int nseg = nf <= kGmT ? (kGmT / nf < 8 ? kGmT / nf : 8) : 1;
  const int jn = it + 1;
  const int per = (jn + nseg - 1) / nseg;
  if (nseg > 1) {
    const int seg = threadIdx.x / nf, i = threadIdx.x - seg * nf;
    if (seg < nseg) {
      const int j0 = seg * per, j1 = j0 + per < jn ? j0 + per : jn;
      double acc = 0.0;
      for (int j = j0; j < j1; ++j) acc += -hcol[j] * V[(long)j * nf + i];
      part[seg * nf + i] = acc;
    }
    __syncthreads();
  }
  double pn = 0.0;
  for (int i = threadIdx.x; i < nf; i += blockDim.x) {
    double acc = w[i];
    if (nseg > 1) {
      for (int sg = 0; sg < nseg; ++sg) acc += part[sg * nf + i];
    } else {
      for (int j = 0; j < jn; ++j) acc += -hcol[j] * V[(long)j * nf + i];
    }
    w[i] = acc;
    pn += acc * acc;
  }
  __syncthreads();
  const double tt = sqrt(gm_block_sum(pn, sh));
  if (threadIdx.x == 0) {
    hcol[it + 1] = tt;
    double hapbnd = fabs(tt / grs[it]);
    if (hapbnd > 1e-30) hapbnd = 1e-30;
    const int hap = tt < hapbnd;
    s_hap = hap;
    s_tt = tt;
    for (int j = 0; j < it; ++j) {
      const double t0 = hcol[j];
      hcol[j] = lcc[j] * t0 + lss[j] * hcol[j + 1];
      hcol[j + 1] = -lss[j] * t0 + lcc[j] * hcol[j + 1];
    }
    double res;
    if (!hap) {
      const double t1 = sqrt(hcol[it] * hcol[it] + hcol[it + 1] * hcol[it + 1]);
      const double c = hcol[it] / t1, sn = hcol[it + 1] / t1;
      cc[it] = c;
      ss[it] = sn;
      grs[it + 1] = -sn * grs[it];
      grs[it] = c * grs[it];
      hcol[it] = c * hcol[it] + sn * hcol[it + 1];
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
  for (int j = threadIdx.x; j <= it + 1; j += blockDim.x) hg[j] = hcol[j];
  if (!s_hap) {
    const double inv = 1.0 / s_tt;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) w[i] = w[i] * inv;
  }
}

// KSPGMRESBuildSoln: back substitution on the rotated Hessenberg, x += V y.
// Columns of H are staged in LDS in chunks of `bs` columns so the dependent
// chain of the substitution runs on LDS values.
static int bs_chunk(int kmax) {
  int c = 6000 / (kmax + 1) - 2;  // keep (c + 2)(kmax + 1) doubles under 48 KB
  return c > 16 ? 16 : (c < 1 ? 1 : c);
}

__global__ void __launch_bounds__(kGmT) k_gm_buildsol(int nf, int kmax, int kBsChunk, const double* __restrict__ V,
                                                      GmresDev* s, const double* __restrict__ H,
                                                      const double* __restrict__ grs, double* __restrict__ x,
                                                      double* __restrict__ nrs) {
  extern __shared__ double dyn[];  // t[kmax+1] | y[kmax+1] | chunk[kBsChunk][kmax+1]
  const int it = s->it;  // number of Arnoldi vectors used
  if (it <= 0) return;
  const int ld = kmax + 1;
  double* t = dyn;
  double* y = dyn + ld;
  double* chunk = dyn + 2 * ld;
  const int top = it - 1;
  for (int j = threadIdx.x; j <= top; j += blockDim.x) t[j] = grs[j];
  __syncthreads();
  for (int k1 = top; k1 >= 0; k1 -= kBsChunk) {
    const int k0 = k1 - kBsChunk + 1 < 0 ? 0 : k1 - kBsChunk + 1;
    const int nc = k1 - k0 + 1;
    // stage columns k0..k1, rows 0..k1
    for (int q = threadIdx.x; q < nc * (k1 + 1); q += blockDim.x) {
      const int c = q / (k1 + 1), r = q - c * (k1 + 1);
      chunk[c * ld + r] = H[(long)(k0 + c) * ld + r];
    }
    __syncthreads();
    for (int k = k1; k >= k0; --k) {
      const double* col = chunk + (k - k0) * ld;
      if (threadIdx.x == 0) y[k] = t[k] / col[k];
      __syncthreads();
      const double yk = y[k];
      for (int j = threadIdx.x; j < k; j += blockDim.x) t[j] = t[j] - col[j] * yk;
      __syncthreads();
    }
  }
  for (int j = threadIdx.x; j <= top; j += blockDim.x) nrs[j] = y[j];
  for (int i = threadIdx.x; i < nf; i += blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j <= top; ++j) acc += y[j] * V[(long)j * nf + i];
    x[i] = x[i] + acc;
  }
}

// y = S x, S dense nf x nf column major.  A workgroup owns 64 rows (one
// per lane) and splits the columns over its 4 wavefronts; the 4 partial sums
// are added in wavefront order.
__global__ void __launch_bounds__(256) k_gemv(int nf, const double* __restrict__ S, const double* __restrict__ x,
                                              double* __restrict__ y) {
  __shared__ double part[4][64];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const int r = blockIdx.x * 64 + lane;
  const int per = (nf + 3) / 4;
  const int j0 = wv * per, j1 = j0 + per < nf ? j0 + per : nf;
  double acc = 0.0;
  if (r < nf)
    for (int j = j0; j < j1; ++j) acc += S[(long)j * nf + r] * x[j];
  part[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && r < nf) y[r] = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
}

void gemv(hipStream_t st, int nf, const double* S, const double* x, double* y) {
  if (nf <= 0) return;
  hipLaunchKernelGGL(k_gemv, dim3(ceil_div(nf, 64)), dim3(256), 0, st, nf, S, x, y);
  IB_LAUNCH_CHECK();
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
  // kmax+1 dot products at most: one wavefront each, 4 per workgroup
  hipLaunchKernelGGL(k_gm_dots, dim3(ceil_div(kmax + 1, 4)), dim3(256), 0, st, nf, kmax, V, s, H);
  IB_LAUNCH_CHECK();
  const size_t lds = sizeof(double) * (size_t)(3 * kmax + 1);
  hipLaunchKernelGGL(k_gm_update, dim3(1), dim3(kGmT), lds, st, nf, kmax, V, s, H, cc, ss, grs);
  IB_LAUNCH_CHECK();
}

void gm_buildsol(hipStream_t st, int nf, int kmax, const double* V, GmresDev* s, const double* H, const double* grs,
                 double* x, double* scratch) {
  const int c = bs_chunk(kmax);
  const size_t lds = sizeof(double) * (size_t)(c + 2) * (kmax + 1);
  hipLaunchKernelGGL(k_gm_buildsol, dim3(1), dim3(kGmT), lds, st, nf, kmax, c, V, s, H, grs, x, scratch);
  IB_LAUNCH_CHECK();
}

}  // namespace ibg
