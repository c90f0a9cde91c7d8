// ib_subdomain.hip -- device kernels of the sub-domain approach (sub_domain = T).
//
// B, the a_times operator restricted to the sub-domain points, is stored by
// columns after its drop tolerance: column c holds rows[colptr[c] ..
// colptr[c+1]) in ascending order with their values.  P (interp_indices /
// interp_weights) gives each force entry np sub-domain entries and weights.
#include "ib_device.h"
#include "ib_kernels.h"

namespace ibg {

// B(r, c) at the listed rows of a batch of columns: the level-1 flux of the
// batch's streamfunction (column c - c0) regularised at the 49 taps of row
// r, summed in tap order (regT_sd_sparse inside a_times_sd,
// src/operators_fluid.F90:124-199).  One wavefront per listed entry; the
// batch's entries are [colptr[c0], colptr[c0 + nc]).
__global__ void __launch_bounds__(256) k_sd_cols(Geom g, int c0, int nc, const long* __restrict__ colptr,
                                                 const int* __restrict__ rows, const double* __restrict__ w,
                                                 const int* __restrict__ key, const double* __restrict__ s1,
                                                 long s_stride, const double* __restrict__ bc1, long bc_stride,
                                                 double* __restrict__ vals) {
  const long e = ((long)blockIdx.x * blockDim.x + threadIdx.x) / warpSize;
  const int lane = threadIdx.x % warpSize;
  const long base = colptr[c0], total = colptr[c0 + nc] - base;
  if (e >= total) return;
  const long pos = base + e;
  // column of this entry: the last c with colptr[c] <= pos
  int lo = c0, hi = c0 + nc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (colptr[mid] <= pos) lo = mid;
    else hi = mid - 1;
  }
  const int b = lo - c0;
  const int r = rows[pos];
  const double* s = s1 + (long)b * s_stride;
  const double* bc = bc1 + (long)b * bc_stride;
  double prod = 0.0;
  if (lane < kTaps) prod = w[(long)r * kTaps + lane] * curl_at(g, s, bc, key[(long)r * kTaps + lane]);
  double acc = 0.0;
  for (int t = 0; t < kTaps; ++t) acc = acc + __shfl(prod, t);
  if (lane == 0) vals[pos] = acc;
}

void sd_columns(hipStream_t st, const Geom& g, int c0, int nc, long nentries, const long* colptr, const int* rows,
                const double* w, const int* key, const double* s1, long s_stride, const double* bc1, long bc_stride,
                double* vals) {
  if (nentries <= 0) return;
  const long waves = nentries, blocks = (waves * 64 + 255) / 256;
  hipLaunchKernelGGL(k_sd_cols, dim3((unsigned)blocks), dim3(256), 0, st, g, c0, nc, colptr, rows, w, key, s1,
                     s_stride, bc1, bc_stride, vals);
  IB_LAUNCH_CHECK();
}

// B(r, c) as compute_operator reads it: the kept entry, else 0
__device__ __forceinline__ double bget(const long* __restrict__ colptr, const int* __restrict__ rows,
                                       const double* __restrict__ vals, int r, int c) {
  long lo = colptr[c], hi = colptr[c + 1];
  while (lo < hi) {
    const long mid = (lo + hi) >> 1;
    if (rows[mid] < r) lo = mid + 1;
    else hi = mid;
  }
  return (lo < colptr[c + 1] && rows[lo] == r) ? vals[lo] : 0.0;
}

// compute_operator + body_mat (src/operators_struct.F90:838-1004):
// lumat(k, i) = body_mat(k, i) + sum_p w_p(k) I(i, idx_p(k)),
// I(i, c) = sum_q B(idx_q(i), c) w_q(i).  body_mat is the torsional block
// written with the reference's operation order.  One thread per (k, i),
// column major.
__global__ void __launch_bounds__(256) k_sd_lumat(int nf, int np, const int* __restrict__ pidx,
                                                  const double* __restrict__ pw, const long* __restrict__ colptr,
                                                  const int* __restrict__ rows, const double* __restrict__ vals,
                                                  const int* __restrict__ tors_of, const TorsBody* __restrict__ tbs,
                                                  const double* __restrict__ theta, double dt, double delta,
                                                  double* __restrict__ lumat) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nf * nf) return;
  const int k = (int)(t % nf), i = (int)(t / nf);
  double body = 0.0;
  const int tk = tors_of[k];
  if (tk >= 0 && tk == tors_of[i]) {
    const TorsBody b = tbs[tk];
    const double si = sin(b.gamma0 - theta[tk]), ci = cos(b.gamma0 - theta[tk]);
    const int nbi = b.npts;
    const int jj = k - b.off + 1, kk = i - b.off + 1;  // row, column within the body block
    const double ds2 = b.deltas * b.deltas;
    const int im = jj - 1, in = kk - 1;
    double v;
    if (jj <= nbi && kk <= nbi) v = (double)(im * in) * (si * si) * ds2;
    else if (jj <= nbi) v = -((double)(im * (in - nbi)) * si * ci * ds2);
    else if (kk <= nbi) v = -((double)((im - nbi) * in) * si * ci * ds2);
    else v = (double)((im - nbi) * (in - nbi)) * (ci * ci) * ds2;
    const double den = 4.0 / (dt * dt) * b.M + 2.0 / dt * b.C + b.K;
    body = v / den * (delta * delta / b.deltas / dt) * (2 * b.deltas / dt);
  }
  double acc = 0.0;
  for (int p = 0; p < np; ++p) {
    const int c = pidx[(long)k * np + p];
    double ic = 0.0;
    for (int q = 0; q < np; ++q) ic += bget(colptr, rows, vals, pidx[(long)i * np + q], c) * pw[(long)i * np + q];
    acc += pw[(long)k * np + p] * ic;
  }
  lumat[t] = body + acc;
}

void sd_lumat(hipStream_t st, int nf, int np, const int* pidx, const double* pw, const long* colptr, const int* rows,
              const double* vals, const int* tors_of, const TorsBody* tb, const double* theta, double dt,
              double delta, double* lumat) {
  if (nf <= 0) return;
  const long n = (long)nf * nf;
  hipLaunchKernelGGL(k_sd_lumat, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nf, np, pidx, pw, colptr, rows,
                     vals, tors_of, tb, theta, dt, delta, lumat);
  IB_LAUNCH_CHECK();
}

// rhsf(j) = sum_p w_p(j) r(idx_p(j)) + rhsf(j) (src/timestepper.F90:389-393)
__global__ void k_sd_interp(int nf, int np, const int* __restrict__ pidx, const double* __restrict__ pw,
                            const double* __restrict__ r, double* __restrict__ rhsf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nf) return;
  double acc = 0.0;
  for (int p = 0; p < np; ++p) acc += pw[(long)j * np + p] * r[pidx[(long)j * np + p]];
  rhsf[j] = acc + rhsf[j];
}

void sd_interp(hipStream_t st, int nf, int np, const int* pidx, const double* pw, const double* r, double* rhsf) {
  if (nf <= 0) return;
  hipLaunchKernelGGL(k_sd_interp, dim3(ceil_div(nf, 256)), dim3(256), 0, st, nf, np, pidx, pw, r, rhsf);
  IB_LAUNCH_CHECK();
}

// h0(c) = sum_j P(j, c) f(j) over the used sub-domain entries c (the
// transpose of P as uptr/uj/uw: entries of c in ascending j), the
// matmul(indicesserialmat, fb) of src/timestepper.F90:502; other entries 0
__global__ void k_sd_ptf(int nused, const int* __restrict__ used, const int* __restrict__ uptr,
                         const int* __restrict__ uj, const double* __restrict__ uw, const double* __restrict__ f,
                         double* __restrict__ h0) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nused) return;
  double acc = 0.0;
  for (int a = uptr[u]; a < uptr[u + 1]; ++a) acc += uw[a] * f[uj[a]];
  h0[used[u]] = acc;
}

void sd_ptf(hipStream_t st, int nused, const int* used, const int* uptr, const int* uj, const double* uw,
            const double* f, double* h0) {
  if (nused <= 0) return;
  hipLaunchKernelGGL(k_sd_ptf, dim3(ceil_div(nused, 256)), dim3(256), 0, st, nused, used, uptr, uj, uw, f, h0);
  IB_LAUNCH_CHECK();
}

}  // namespace ibg
