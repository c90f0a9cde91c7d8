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
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "ib_device.h"
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

// y = S x, S dense nf x nf column major.  A workgroup owns 16 rows and
// splits the columns into 16 contiguous groups (4 per wavefront: lane & 15
// is the row, lane >> 4 the group), so the matrix streams through many CUs;
// the 16 partial sums of a row are added in group order.
struct GemvOp {
  using Args = GemvArgs;
  static __device__ __forceinline__ void tile(const GemvArgs& a, int bx, int by, int);
};

__device__ __forceinline__ void GemvOp::tile(const GemvArgs& a, int bx, int, int) {
  __shared__ double part[16][16];
  const int nf = a.nf;
  const double* __restrict__ S = a.S;
  const double* __restrict__ x = a.x;
  double* __restrict__ y = a.y;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int rl = lane & 15, grp = wv * 4 + (lane >> 4);
  const int r = bx * 16 + rl;
  const int per = (nf + 15) / 16;
  const int j0 = min(nf, grp * per), j1 = min(nf, j0 + per);
  double acc = 0.0;
  if (r < nf) {
    for (int j = j0; j < j1; j += 8) {
      double sv[8], xv[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int jj = min(j + c, j1 - 1);
        sv[c] = S[(long)jj * nf + r];
        xv[c] = x[jj];
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) acc += (j + c < j1 ? sv[c] : 0.0) * xv[c];
    }
  }
  part[grp][rl] = acc;
  __syncthreads();
  if (threadIdx.x < 16 && r < nf) {
    double a = part[0][rl];
    for (int q = 1; q < 16; ++q) a += part[q][rl];
    y[r] = a;
  }
}

__global__ void __launch_bounds__(256) k_gemv(GemvArgs a) { GemvOp::tile(a, blockIdx.x, 0, 0); }

TabEntry<GemvArgs> gemv_entry(int nf, const double* S, const double* x, double* y) {
  TabEntry<GemvArgs> e{};
  e.a = GemvArgs{nf, S, x, y};
  e.nbx = ceil_div(std::max(nf, 1), 16);
  e.nby = 1;
  return e;
}

const void* gemv_tab_kernel() { return (const void*)k_tab<GemvOp>; }

void gemv(hipStream_t st, int nf, const double* S, const double* x, double* y) {
  if (nf <= 0) return;
  const TabEntry<GemvArgs> e = gemv_entry(nf, S, x, y);
  hipLaunchKernelGGL(k_gemv, dim3(e.nbx), dim3(256), 0, st, e.a);
  IB_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------
// Whole solve in one workgroup for a dense Schur matrix.  The same PETSc
// 3.11 GMRES as the kernels above (CGS, Givens, restart, stopping tests),
// but every iteration stays inside the kernel: S is held in registers
// (MAXC columns of one row per thread, the rest read from L2), w, the new
// Hessenberg column, the rotations and the residual vector in LDS, and the
// Krylov basis in HBM/L2.  Iterations cost a few microseconds instead of
// three dependent launches.
// ------------------------------------------------------------------------
struct GmDenseArgs {
  int nf, kmax, maxit;
  int nvl;     // basis vectors held in LDS (the rest are read from HBM/L2)
  int bsc;     // Hessenberg columns staged per back-substitution chunk
  const double* S;
  const double* b;
  double* x;
  double* V;
  double* H;
  GmresDev* s;
  double rtol, atol;
};

constexpr int kGmRem = 12;  // S entries per thread fetched from L2 in one unrolled batch

// acc += -sum_{j in [q0, q1)} h[j] * v_j[r] in j order, v_j from the LDS cache
// for j < nvl and from HBM/L2 otherwise; loads are issued in batches of 8.
__device__ __forceinline__ double gm_axpy_col(double acc, const double* __restrict__ h, const double* cache,
                                              const double* __restrict__ V, int nf, int nvl, int q0, int q1,
                                              int r) {
  const int qa = min(q1, nvl);
  for (int j = q0; j < qa; j += 8) {
    double v[8], hv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int jj = min(j + c, qa - 1);
      v[c] = cache[(long)jj * nf + r];
      hv[c] = h[jj];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) acc += -(j + c < qa ? hv[c] : 0.0) * v[c];
  }
  for (int j = max(q0, nvl); j < q1; j += 8) {
    double v[8], hv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int jj = min(j + c, q1 - 1);
      v[c] = V[(long)jj * nf + r];
      hv[c] = h[jj];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) acc += -(j + c < q1 ? hv[c] : 0.0) * v[c];
  }
  return acc;
}

// The thread index as a value the compiler cannot see through: per-thread
// addresses derived from it are recomputed in each phase instead of being
// hoisted out of the solve loop and kept live in registers for the whole
// kernel (which would crowd out the register-resident part of S).
__device__ __forceinline__ int gm_opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Cross-lane data movement through DPP (register-level, no LDS crossbar).
// CTRL is the gfx9 dpp_ctrl code; doubles move as two 32-bit halves.
template <int CTRL>
__device__ __forceinline__ double gm_dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// every lane gets the sum over its row of 16 lanes: quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_ror:4, row_ror:8
__device__ __forceinline__ double gm_row_sum16(double v) {
  v += gm_dpp<0xB1>(v);
  v += gm_dpp<0x4E>(v);
  v += gm_dpp<0x124>(v);
  v += gm_dpp<0x128>(v);
  return v;
}

__device__ __forceinline__ double gm_readlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum over the 64 lanes of a wavefront (uniform result, fixed order)
__device__ __forceinline__ double gm_wave_sum(double v) {
  const double rs = gm_row_sum16(v);
  return (gm_readlane(rs, 0) + gm_readlane(rs, 16)) + (gm_readlane(rs, 32) + gm_readlane(rs, 48));
}

// sum over i = sub, sub+16, ... < nf of w[i] v[i], in batches of 8 loads
// per operand with clamped indices (no branches between the loads)
template <bool LDS>
__device__ __forceinline__ double gm_dot16(const double* w, const double* v, int nf, int sub) {
  double p = 0.0;
  for (int i0 = sub; i0 < nf; i0 += 16 * 8) {
    double wv[8], vv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int i = min(i0 + 16 * c, nf - 1);
      wv[c] = w[i];
      vv[c] = v[i];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) p += (i0 + 16 * c < nf ? wv[c] : 0.0) * vv[c];
  }
  return p;
}

template <int CTRL, int RM>
__device__ __forceinline__ double gm_dpp_old(double v, double old) {
  const long long b = __double_as_longlong(v), o = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)b, CTRL, RM, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, RM, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, int RM>
__device__ __forceinline__ void gm_scan_step(double& A, double& B) {
  const double Ap = gm_dpp_old<CTRL, RM>(A, 1.0), Bp = gm_dpp_old<CTRL, RM>(B, 0.0);
  B = A * Bp + B;
  A = A * Ap;
}
// inclusive scan of the affine maps x -> A x + B over the 64 lanes (lane l
// ends with f_l o ... o f_0): row_shr 1/2/4/8 within rows of 16, then
// row_bcast 15 and 31 across rows; lanes a shift leaves without a source
// compose with the identity (1, 0)
__device__ __forceinline__ void gm_affine_scan(double& A, double& B) {
  gm_scan_step<0x111, 0xF>(A, B);
  gm_scan_step<0x112, 0xF>(A, B);
  gm_scan_step<0x114, 0xF>(A, B);
  gm_scan_step<0x118, 0xF>(A, B);
  gm_scan_step<0x142, 0xA>(A, B);
  gm_scan_step<0x143, 0xC>(A, B);
}
// the value of lane l-1 (wave_shr 1); lane 0 gets `first`
__device__ __forceinline__ double gm_shift_up1(double v, double first) { return gm_dpp_old<0x138, 0xF>(v, first); }

// row-segment layout of the matrix-vector product: thread (seg, r) owns row
// r and columns [j0, j1); nseg segments of nf threads
struct GmLayout {
  int nseg, seg, r, j0, j1;
  bool act;
  __device__ GmLayout(int t, int nf, int T) {
    nseg = nf <= T ? min(16, T / nf) : 1;
    seg = t / nf;
    r = t - seg * nf;
    act = nf <= T && seg < nseg;
    const int per = (nf + nseg - 1) / nseg;
    j0 = act ? seg * per : 0;
    j1 = act ? min(nf, j0 + per) : 0;
  }
};

// part[0*nf+i] + part[1*nf+i] + ... + part[(nseg-1)*nf+i] in that order (nseg <= 16)
__device__ __forceinline__ double gm_sum_parts(const double* part, int nf, int i, int nseg) {
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = part[min(q, nseg - 1) * nf + i];
  double acc = v[0];
#pragma unroll
  for (int q = 1; q < 16; ++q) acc += (q < nseg ? v[q] : 0.0);
  return acc;
}

__device__ __forceinline__ double gm_block_sum2(double v, double* red, int T) {
  const int t = gm_opaque(threadIdx.x);
  v = gm_wave_sum(v);
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double tot = 0.0;
  for (int q = 0; q < T / 64; ++q) tot += red[q];
  __syncthreads();
  return tot;
}

// out = S vin (vin, out in LDS); S entries [j0, j0+MAXC) of the thread's row
// segment come from registers, the rest from L2 in unrolled batches.  Ends
// with a barrier.
template <int MAXC>
__device__ __forceinline__ void gm_gemv(const double (&sreg)[MAXC], const double* __restrict__ S, int nf, int T,
                                        const double* vin, double* out, double* part) {
  const GmLayout g(gm_opaque(threadIdx.x), nf, T);
  if (nf <= T) {
    // branch-free (clamped indices, zero weights past the segment) so that
    // all loads of a batch are in flight together
    if (g.act) {
      double acc = 0.0;
#pragma unroll
      for (int c0 = 0; c0 < MAXC; c0 += 8) {
        double vv[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) vv[c] = c0 + c < MAXC ? vin[min(g.j0 + c0 + c, nf - 1)] : 0.0;
#pragma unroll
        for (int c = 0; c < 8; ++c)
          if (c0 + c < MAXC) acc += sreg[c0 + c] * vv[c];  // sreg is 0 past the segment
      }
      for (int jb = g.j0 + MAXC; jb < g.j1; jb += kGmRem) {
        double sv[kGmRem], xv[kGmRem];
#pragma unroll
        for (int c = 0; c < kGmRem; ++c) {
          const int j = min(jb + c, g.j1 - 1);
          sv[c] = S[(long)j * nf + g.r];
          xv[c] = vin[j];
        }
#pragma unroll
        for (int c = 0; c < kGmRem; ++c) acc += (jb + c < g.j1 ? sv[c] : 0.0) * xv[c];
      }
      part[g.seg * nf + g.r] = acc;
    }
    __syncthreads();
    for (int i = gm_opaque(threadIdx.x); i < nf; i += T) out[i] = gm_sum_parts(part, nf, i, g.nseg);
  } else {
    for (int i = gm_opaque(threadIdx.x); i < nf; i += T) {
      double acc = 0.0;
      for (int j = 0; j < nf; ++j) acc += S[(long)j * nf + i] * vin[j];
      out[i] = acc;
    }
  }
  __syncthreads();
}

template <int MAXC>
__global__ void __launch_bounds__(kGmT) k_gmres_dense(GmDenseArgs a) {
  extern __shared__ double dyn[];
  const int nf = a.nf, kmax = a.kmax, ld = kmax + 1, T = kGmT;
  double* vl = dyn;               // nf: the basis vector being multiplied
  double* wl = vl + nf;           // nf: w
  double* hcol = wl + nf;         // kmax + 2
  double* lcc = hcol + kmax + 2;  // kmax
  double* lss = lcc + kmax;       // kmax
  double* grs = lss + kmax;       // kmax + 1
  double* yv = grs + kmax + 1;    // kmax + 1
  double* part = yv + kmax + 1;   // T
  double* cache = part + T;       // basis vectors 0..nvl-1; Hessenberg chunks in the back substitution
  const int nvl = a.nvl;
  __shared__ double red[kGmT / 64];
  __shared__ double s_val[2];
  __shared__ int s_flag[2];

  double sreg[MAXC];  // the first MAXC columns of the thread's row segment, for the whole solve
  {
    const GmLayout g(threadIdx.x, nf, T);
#pragma unroll
    for (int c = 0; c < MAXC; ++c) sreg[c] = (g.j0 + c < g.j1) ? a.S[(long)(g.j0 + c) * nf + g.r] : 0.0;
  }

  int its = 0, reason = 0, first = 1;
  double rn0 = 0.0, ttol = 0.0;
  while (true) {
    // ---- cycle start: r = b - S x
    for (int i = gm_opaque(threadIdx.x); i < nf; i += T) vl[i] = a.x[i];
    __syncthreads();
    gm_gemv<MAXC>(sreg, a.S, nf, T, vl, wl, part);
    double pr = 0.0, pb = 0.0;
    for (int i = gm_opaque(threadIdx.x); i < nf; i += T) {
      const double rv = a.b[i] - wl[i];
      wl[i] = rv;
      pr += rv * rv;
      pb += a.b[i] * a.b[i];
    }
    const double res0 = sqrt(gm_block_sum2(pr, red, T));
    const double bn = sqrt(gm_block_sum2(pb, red, T));
    if (first) {
      rn0 = bn > 0.0 ? bn : res0;
      ttol = fmax(a.rtol * rn0, a.atol);
      first = 0;
    }
    if (res0 == 0.0 || res0 <= ttol) reason = 1;
    else if (res0 >= 1e4 * rn0) reason = -1;
    if (reason) break;
    if (threadIdx.x == 0) grs[0] = res0;
    for (int i = gm_opaque(threadIdx.x); i < nf; i += T) {
      const double v0 = wl[i] / res0;
      vl[i] = v0;
      a.V[i] = v0;
      if (nvl > 0) cache[i] = v0;
    }
    __syncthreads();
    int it = 0;
    while (true) {
      gm_gemv<MAXC>(sreg, a.S, nf, T, vl, wl, part);  // w = S v_it
      // classical Gram-Schmidt coefficients h_j = <w, v_j>, j = 0..it: one
      // row of 16 lanes per coefficient, reduced with DPP
      {
        const int t = gm_opaque(threadIdx.x);
        const int sub = t & 15;
        const int ndots = it + 1;
        for (int j = t >> 4; j < ndots; j += T / 16) {
          double p = 0.0;
          if (j < nvl) p = gm_dot16<true>(wl, cache + (long)j * nf, nf, sub);
          else p = gm_dot16<false>(wl, a.V + (long)j * nf, nf, sub);
          p = gm_row_sum16(p);
          if (sub == 0) hcol[j] = p;
        }
      }
      __syncthreads();
      // w -= sum_j h_j v_j (j range split over the row segments, partials in fixed order), ||w||
      double pn = 0.0;
      {
        const GmLayout g(gm_opaque(threadIdx.x), nf, T);
        const int jn = it + 1;
        if (nf <= T && g.nseg > 1) {
          const int per2 = (jn + g.nseg - 1) / g.nseg;
          if (g.act) {
            const int q0 = min(jn, g.seg * per2), q1 = min(jn, q0 + per2);
            part[g.seg * nf + g.r] = gm_axpy_col(0.0, hcol, cache, a.V, nf, nvl, q0, q1, g.r);
          }
          __syncthreads();
          for (int i = gm_opaque(threadIdx.x); i < nf; i += T) {
            const double acc = wl[i] + gm_sum_parts(part, nf, i, g.nseg);
            wl[i] = acc;
            pn += acc * acc;
          }
        } else {
          for (int i = gm_opaque(threadIdx.x); i < nf; i += T) {
            const double acc = gm_axpy_col(wl[i], hcol, cache, a.V, nf, nvl, 0, jn, i);
            wl[i] = acc;
            pn += acc * acc;
          }
        }
      }
      const double tt = sqrt(gm_block_sum2(pn, red, T));
      if (gm_opaque(threadIdx.x) < 64) {
        const int lane = gm_opaque(threadIdx.x);
        // previous rotations on the new column (KSPGMRESUpdateHessenberg): the
        // running entry obeys carry_(j+1) = -s_j carry_j + c_j h_(j+1), an
        // affine recurrence evaluated 64 rotations at a time by a DPP scan
        double carry = hcol[0];
        for (int base = 0; base < it; base += 64) {
          const int j = base + lane;
          const bool act = j < it;
          const double c = act ? lcc[j] : 1.0, sn = act ? lss[j] : 0.0, hn = act ? hcol[j + 1] : 0.0;
          double A = -sn, B = c * hn;
          gm_affine_scan(A, B);
          const double cnext = A * carry + B;            // carry_(j+1)
          const double cj = gm_shift_up1(cnext, carry);  // carry_j (lane 0: the carry in)
          if (act) hcol[j] = c * cj + sn * hn;
          carry = gm_readlane(cnext, min(63, it - 1 - base));
        }
        if (lane == 0) {
          hcol[it] = carry;
          hcol[it + 1] = tt;
          double hapbnd = fabs(tt / grs[it]);
          if (hapbnd > 1e-30) hapbnd = 1e-30;
          const int hap = tt < hapbnd;
          double res;
          if (!hap) {
            const double t1 = sqrt(hcol[it] * hcol[it] + hcol[it + 1] * hcol[it + 1]);
            const double c = hcol[it] / t1, sn = hcol[it + 1] / t1;
            lcc[it] = c;
            lss[it] = sn;
            grs[it + 1] = -sn * grs[it];
            grs[it] = c * grs[it];
            hcol[it] = c * hcol[it] + sn * hcol[it + 1];
            res = fabs(grs[it + 1]);
          } else {
            res = 0.0;
          }
          int rs = 0;
          if (res <= ttol) rs = 1;
          else if (res >= 1e4 * rn0) rs = -1;
          if (hap && !rs) rs = 1;
          s_flag[0] = rs;
          s_flag[1] = hap;
          s_val[0] = tt;
          s_val[1] = res;
        }
      }
      __syncthreads();
      reason = s_flag[0];
      const int hap = s_flag[1];
      for (int j = gm_opaque(threadIdx.x); j <= it + 1; j += T) a.H[(long)it * ld + j] = hcol[j];
      if (!hap) {
        const double inv = 1.0 / s_val[0];
        const bool keep = it + 1 < nvl;
        for (int i = gm_opaque(threadIdx.x); i < nf; i += T) {
          const double v1 = wl[i] * inv;
          vl[i] = v1;
          a.V[(long)(it + 1) * nf + i] = v1;
          if (keep) cache[(long)(it + 1) * nf + i] = v1;
        }
      }
      ++it;
      ++its;
      __syncthreads();
      if (reason || it >= kmax || its >= a.maxit) break;
    }
    // ---- KSPGMRESBuildSoln: back substitution on the rotated Hessenberg
    // (columns staged through the LDS cache, the serial sweep by one
    // wavefront), then x += V y
    const int top = it - 1;
    for (int j = gm_opaque(threadIdx.x); j <= top; j += T) yv[j] = grs[j];  // right-hand side, overwritten by y
    for (int k1 = top; k1 >= 0; k1 -= a.bsc) {
      const int k0 = k1 - a.bsc + 1 < 0 ? 0 : k1 - a.bsc + 1;
      const int nc = k1 - k0 + 1;
      __syncthreads();
      for (int q = gm_opaque(threadIdx.x); q < nc * (k1 + 1); q += T) {
        const int c = q / (k1 + 1), rr = q - c * (k1 + 1);
        cache[(long)c * ld + rr] = a.H[(long)(k0 + c) * ld + rr];
      }
      __syncthreads();
      if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        for (int k = k1; k >= k0; --k) {
          const double* col = cache + (long)(k - k0) * ld;
          const double yk = yv[k] / col[k];
          for (int j = lane; j < k; j += 64) yv[j] = yv[j] - col[j] * yk;
          if (lane == 0) yv[k] = yk;
          // LDS operations of one wavefront complete in order: the next
          // step's read of y(k-1) sees this step's update
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
    __syncthreads();
    {
      const GmLayout g(gm_opaque(threadIdx.x), nf, T);
      if (nf <= T && g.nseg > 1) {
        const int per2 = (top + 1 + g.nseg - 1) / g.nseg;
        if (g.act) {
          const int q0 = min(top + 1, g.seg * per2), q1 = min(top + 1, q0 + per2);
          double acc = 0.0;  // sum_j y_j v_j[r] over this segment's j range
          for (int j = q0; j < q1; j += 8) {
            double v[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) v[c] = a.V[(long)min(j + c, q1 - 1) * nf + g.r];
#pragma unroll
            for (int c = 0; c < 8; ++c)
              if (j + c < q1) acc += yv[j + c] * v[c];
          }
          part[g.seg * nf + g.r] = acc;
        }
        __syncthreads();
        for (int i = gm_opaque(threadIdx.x); i < nf; i += T) a.x[i] = a.x[i] + gm_sum_parts(part, nf, i, g.nseg);
      } else {
        for (int i = gm_opaque(threadIdx.x); i < nf; i += T) {
          double acc = 0.0;
          for (int j = 0; j <= top; ++j) acc += yv[j] * a.V[(long)j * nf + i];
          a.x[i] = a.x[i] + acc;
        }
      }
    }
    __syncthreads();
    if (reason) break;
    if (its >= a.maxit) {
      reason = -2;
      break;
    }
  }
  if (threadIdx.x == 0) {
    a.s->its = its;
    a.s->reason = reason;
    a.s->rn0 = rn0;
    a.s->ttol = ttol;
  }
}

// dynamic LDS of k_gmres_dense: vl, wl (nf each), hcol (kmax+2), lcc, lss
// (kmax each), grs, yv (kmax+1 each), part (kGmT), then the cache region
size_t gmres_dense_lds(int nf, int kmax, size_t cache_doubles) {
  return sizeof(double) * ((size_t)2 * nf + (size_t)5 * kmax + 4 + kGmT + cache_doubles);
}

bool gmres_dense_fits(int nf, int kmax, size_t max_lds) {
  return gmres_dense_lds(nf, kmax, (size_t)kmax + 1) + 1024 <= max_lds;
}

void gmres_dense(hipStream_t st, int nf, int kmax, int maxit, const double* S, const double* b, double* x, double* V,
                 double* H, GmresDev* s, double rtol, double atol, size_t max_lds) {
  // the cache takes what is left of the LDS: basis vectors during the
  // iterations, Hessenberg columns during the back substitution
  const size_t avail = max_lds > 1024 ? max_lds - 1024 : 0;  // room for the static shared arrays
  const size_t base = gmres_dense_lds(nf, kmax, 0);
  if (base + sizeof(double) * (kmax + 1) > avail) throw HipError("dense GMRES: restart too long for LDS");
  const size_t cache = (avail - base) / sizeof(double);
  const int nvl = (int)std::min<size_t>((size_t)kmax + 1, cache / (size_t)std::max(nf, 1));
  const int bsc = (int)std::max<size_t>(1, std::min<size_t>((size_t)kmax, cache / (size_t)(kmax + 1)));
  GmDenseArgs a{nf, kmax, maxit, nvl, bsc, S, b, x, V, H, s, rtol, atol};
  const size_t lds = gmres_dense_lds(nf, kmax, cache);
  hipLaunchKernelGGL(k_gmres_dense<24>, dim3(1), dim3(kGmT), lds, st, a);
  IB_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------
// The same dense solve spread over P workgroups of one XCD.  Workgroup c
// owns rows [c R, c R + R) of S (in registers: four rows by kCoopColQ
// columns per thread) and those rows of every Krylov vector (in LDS), so
// neither S nor the basis is re-read from L2.  Per Arnoldi step the
// workgroups exchange, through L2 and two grid barriers, the partial CGS
// coefficients of their rows, then the partial squared norm and their rows
// of w.  Every workgroup sums the partials of all workgroups in the same
// order and runs the rotations, tests and back substitution on identical
// values, so all take the same branches; each updates its rows of x.  The
// launch has 8P workgroups, dealt by the dispatcher to the 8 XCDs in turn;
// the P on the XCD the host names take part (coop_rank), so a solve keeps
// to one XCD's CUs and L2 and solves on different XCDs run side by side.
// The exchanged values are device-scope atomic stores and loads (coherent
// without cache write-backs); the barrier waits are bounded: on a timeout
// every workgroup leaves with reason -3.
// ------------------------------------------------------------------------
// Thread layout of the matrix-vector product and of w -= V h: wavefront q
// of the 16 owns local rows 4q .. 4q+3 (R <= 64), lane l the columns
// l, l + 64, ... (nf <= 64 kCoopColQ); each lane's four row partials are
// summed over the wavefront by gc_reduce4.
constexpr int kCoopColQ = 4;
constexpr int kCoopRowsPer16 = 4;  // local rows per 16-lane group in the CGS dots: R <= 64
constexpr long kCoopSpin = 1L << 22;

// The barrier block of a cooperative solve (unsigned words, cleared before
// each launch): word 0 the error flag, 8.. the workgroups' barrier counts,
// kXccSlot.. the XCD ids they publish, kClaimSlot the ranks claimed.
constexpr int kCoopSyncDoubles = 40;
constexpr int kXccSlot = 40;
constexpr int kClaimSlot = 72;
static_assert(kClaimSlot < 2 * kCoopSyncDoubles, "barrier block too small");

// The XCD this wavefront runs on (HW_REG_XCC_ID, bits 3:0).
__device__ __forceinline__ unsigned gw_xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 0xFu; }

// The rank of this workgroup in a cooperative solve of P workgroups, or -1
// if it takes no part.  xcd < 0: workgroups 0, 8, 16, ... (one XCD when the
// dispatcher starts the launch's round on it); xcd >= 0: the first P
// workgroups that run on XCD xcd (XCC_ID), ranked in the order they claim.
// A launch of 8P workgroups puts P on every XCD.
__device__ __forceinline__ int coop_rank(unsigned* sync, int P, int xcd, int* s_c) {
  if (xcd < 0) return blockIdx.x % 8 == 0 && (int)(blockIdx.x / 8) < P ? (int)(blockIdx.x / 8) : -1;
  if (threadIdx.x == 0)
    *s_c = gw_xcc_id() == (unsigned)xcd
               ? (int)__hip_atomic_fetch_add(sync + kClaimSlot, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
               : P;
  __syncthreads();
  const int c = *s_c;
  return c < P ? c : -1;
}

// Before a cooperative launch: the barrier block cleared and the status set
// to reason -3, which stays if no workgroup takes part.
__global__ void k_coop_prep(unsigned* sync, GmresDev* s) {
  if (threadIdx.x < 2 * kCoopSyncDoubles) sync[threadIdx.x] = 0u;
  if (threadIdx.x == 0) {
    s->reason = -3;
    s->its = 0;
  }
}

static void coop_prep(hipStream_t st, double* ws, GmresDev* s) {
  hipLaunchKernelGGL(k_coop_prep, dim3(1), dim3(2 * kCoopSyncDoubles), 0, st, reinterpret_cast<unsigned*>(ws), s);
  IB_LAUNCH_CHECK();
}

__global__ void k_xcc_probe(unsigned* ids) {
  if (threadIdx.x == 0) ids[blockIdx.x] = gw_xcc_id();
}

std::vector<int> gmres_xcc_ids(hipStream_t st) {
  static std::mutex mu;
  static std::map<int, std::vector<int>> known;
  int dev = 0;
  IB_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto it = known.find(dev);
  if (it != known.end()) return it->second;
  constexpr int kProbe = 64;
  unsigned* d = nullptr;
  IB_HIP(hipMalloc(&d, sizeof(unsigned) * kProbe));
  unsigned h[kProbe];
  hipLaunchKernelGGL(k_xcc_probe, dim3(kProbe), dim3(64), 0, st, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d);
  IB_HIP(e);
  std::vector<int> ids;
  for (unsigned v : h) ids.push_back((int)v);
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  known[dev] = ids;
  return ids;
}

struct GmCoopArgs {
  int nf, kmax, maxit, P, R;
  int xcd;        // the XCD whose workgroups take part (-1: workgroups 0, 8, 16, ...)
  int bsc;        // Hessenberg columns per back-substitution chunk
  const double* S;
  const double* b;
  double* x;
  GmresDev* s;
  double rtol, atol;
  unsigned* sync;  // [0] error, [8 + c] barriers passed by workgroup c
  double* dots;    // P x (kmax + 2) partial CGS coefficients
  double* norms;   // P x 4 partial squared norms
  double* wfull;   // nf: w (or r) rows of every workgroup
  double* H;       // P x (kmax + 1) x kmax: one Hessenberg copy per workgroup
};

// A value another workgroup wrote before the last grid barrier.  An atomic
// load is always a vector memory access (served by L2 once the barrier has
// invalidated L1); a plain load of a uniform address could be served by the
// scalar cache, which the barrier does not refresh.
__device__ __forceinline__ double gc_ld(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// A value other workgroups read after the next grid barrier (coherent at
// device scope without a cache write-back).
__device__ __forceinline__ void gc_st(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The sum over q = 0..P-1 (in that order) of p[q * stride], values other
// workgroups wrote before the last grid barrier: the loads go out in
// batches of 8 (clamped indices) before any of them is used, so a batch
// costs one memory round trip instead of one per value.
__device__ __forceinline__ double gc_sum(const double* p, long stride, int P) {
  double h = 0.0;
  for (int q0 = 0; q0 < P; q0 += 8) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = gc_ld(p + (long)min(q0 + q, P - 1) * stride);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q0 + q < P) h += v[q];
  }
  return h;
}

// Every lane of a wavefront holds partial sums v0..v3 of four rows; returns
// in lane l the sum over the 64 lanes of row ((l & 1) << 1) | ((l >> 1) & 1)
// (lanes 0, 1, 2, 3: rows 0, 2, 1, 3).  Pairs exchange half their rows
// (xor 1), then quads (xor 2), then the quads of a row of 16 are added
// (row_ror 4, 8) and the four rows of 16 (xor 16, 32).  Fixed order; the
// whole wavefront must be active.
__device__ __forceinline__ double gc_reduce4(double v0, double v1, double v2, double v3) {
  const int l = gm_opaque(threadIdx.x) & 63;
  const bool odd = l & 1, hi = l & 2;
  double k0 = odd ? v2 : v0, k1 = odd ? v3 : v1;
  const double s0 = odd ? v0 : v2, s1 = odd ? v1 : v3;
  k0 += gm_dpp<0xB1>(s0);
  k1 += gm_dpp<0xB1>(s1);
  double k = hi ? k1 : k0;
  k += gm_dpp<0x4E>(hi ? k0 : k1);
  k += gm_dpp<0x124>(k);
  k += gm_dpp<0x128>(k);
  k += __shfl_xor(k, 16);
  k += __shfl_xor(k, 32);
  return k;
}
// the local row whose sum gc_reduce4 leaves in lane l of wavefront q
__device__ __forceinline__ int gc_reduce4_row(int q, int l) { return 4 * q + (((l & 1) << 1) | ((l >> 1) & 1)); }

// Wait until every vector memory operation this wavefront issued has
// completed; for the device-scope stores above that means they have
// reached the coherent level.  A workgroup-scope release fence does not
// emit this wait on gfx950 (it only orders within the CU), so without it
// another wavefront's flag store could overtake this wavefront's data.
__device__ __forceinline__ void gc_wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Grid barrier number k (1, 2, ...) over the P workgroups.  Each
// workgroup's values are device-scope atomic stores.  Every wavefront
// waits for its own stores to complete (gc_wait_stores) before the
// workgroup barrier, so when wavefront 0 then publishes flag[c] = k every
// store of workgroup c is visible at device scope; wavefront 0 polls the P
// flags (one lane each) until all have reached k, and the others wait at
// the second workgroup barrier, after which their device-scope loads see
// the other workgroups' values.  Bounded: on a timeout (or another
// workgroup's) the error word is set and every workgroup leaves.
__device__ __forceinline__ bool gc_sync(unsigned* sync, int P, int c, unsigned k, int* s_err) {
  gc_wait_stores();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x < 64) {
    unsigned* err = sync;
    unsigned* fl = sync + 8;
    const int lane = threadIdx.x;
    if (lane == 0) __hip_atomic_store(fl + c, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    long n = 0;
    while (true) {
      const unsigned v = lane < P ? __hip_atomic_load(fl + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : k;
      if (__all(v >= k)) break;
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || ++n > kCoopSpin) {
        if (lane == 0) {
          __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          *s_err = 1;
        }
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  __syncthreads();
  return *s_err == 0;
}

__global__ void __launch_bounds__(kGmT) k_gmres_coop(GmCoopArgs a) {
  __shared__ int s_c;
  const int c = coop_rank(a.sync, a.P, a.xcd, &s_c);
  if (c < 0) return;
  extern __shared__ double dyn[];
  const int nf = a.nf, kmax = a.kmax, ld = kmax + 1, T = kGmT, R = a.R;
  const int r0 = c * R, Rc = max(0, min(nf, r0 + R) - r0);
  double* vfull = dyn;                          // nf: the vector being multiplied
  double* Vl = vfull + nf;                      // (kmax + 1) R: local rows of the basis, vector j at Vl + j R
  double* wl = Vl + (size_t)(kmax + 1) * R;     // R: local rows of w
  double* part = wl + R;                        // T
  double* hcol = part + T;                      // kmax + 2
  double* lcc = hcol + kmax + 2;                // kmax
  double* lss = lcc + kmax;                     // kmax
  double* grs = lss + kmax;                     // kmax + 1
  double* yv = grs + kmax + 1;                  // kmax + 1
  double* hrot = yv + kmax + 1;                 // kmax + 2: the new Hessenberg column after the rotations
  __shared__ double red[kGmT / 64];
  __shared__ double s_val[2];
  __shared__ int s_flag[2];
  __shared__ int s_err;
  if (threadIdx.x == 0) s_err = 0;
  double* Hc = a.H + (size_t)c * ld * kmax;
  // S[r0 + 4 wq + k][ln + 64 q] in registers (0 outside the matrix)
  const int wq = threadIdx.x >> 6, ln = threadIdx.x & 63;
  double sreg[4][kCoopColQ];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int q = 0; q < kCoopColQ; ++q) {
      const int row = 4 * wq + k, col = ln + 64 * q;
      sreg[k][q] = row < Rc && col < nf ? a.S[(long)col * nf + r0 + row] : 0.0;
    }
  __syncthreads();

  // wl = S[local rows] vfull (ends with a barrier)
  auto matvec = [&]() {
    const int q4 = gm_opaque(threadIdx.x) >> 6, l = gm_opaque(threadIdx.x) & 63;
    if (4 * q4 < Rc) {  // uniform over the wavefront
      double vv[kCoopColQ];
#pragma unroll
      for (int q = 0; q < kCoopColQ; ++q) vv[q] = vfull[min(l + 64 * q, nf - 1)];
      double acc[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        acc[k] = 0.0;
#pragma unroll
        for (int q = 0; q < kCoopColQ; ++q) acc[k] += sreg[k][q] * vv[q];  // sreg is 0 past the matrix
      }
      const double sum = gc_reduce4(acc[0], acc[1], acc[2], acc[3]);
      const int row = gc_reduce4_row(q4, l);
      if (l < 4 && row < Rc) wl[row] = sum;
    }
    __syncthreads();
  };

  // the previous rotations applied to the new column (KSPGMRESUpdateHessenberg)
  // as an affine-recurrence scan, 64 rotations at a time, by one whole
  // wavefront: hrot[j] for j < it, and hrot[it] = the entry the new rotation
  // acts on; hcol keeps the unrotated CGS coefficients
  auto rotate_prev = [&](int it, int lane) {
    double carry = hcol[0];
    for (int base = 0; base < it; base += 64) {
      const int j = base + lane;
      const bool act = j < it;
      const double cj0 = act ? lcc[j] : 1.0, sj = act ? lss[j] : 0.0, hn = act ? hcol[j + 1] : 0.0;
      double A = -sj, B = cj0 * hn;
      gm_affine_scan(A, B);
      const double cnext = A * carry + B;
      const double cj = gm_shift_up1(cnext, carry);
      if (act) hrot[j] = cj0 * cj + sj * hn;
      carry = gm_readlane(cnext, min(63, it - 1 - base));
    }
    if (lane == 0) hrot[it] = carry;
  };
  // the last wavefront owns no rows of the update w -= V h when R <= 60: it
  // runs the scan while the others update w
  const bool scan_early = 4 * (T / 64 - 1) >= Rc;

  int its = 0, reason = 0, first = 1;
  unsigned bk = 0;  // grid barriers passed
  double rn0 = 0.0, ttol = 0.0;
  while (true) {
    // ---- cycle start: r = b - S x (x gathered from every workgroup's rows)
    for (int i = gm_opaque(threadIdx.x); i < nf; i += T) vfull[i] = gc_ld(a.x + i);
    __syncthreads();
    matvec();
    double pr = 0.0, pb = 0.0;
    for (int r = gm_opaque(threadIdx.x); r < Rc; r += T) {
      const double bv = a.b[r0 + r];
      const double rv = bv - wl[r];
      gc_st(a.wfull + r0 + r, rv);
      pr += rv * rv;
      pb += bv * bv;
    }
    pr = gm_block_sum2(pr, red, T);
    pb = gm_block_sum2(pb, red, T);
    if (threadIdx.x == 0) {
      gc_st(a.norms + c * 4 + 1, pr);
      gc_st(a.norms + c * 4 + 2, pb);
    }
    if (!gc_sync(a.sync, a.P, c, ++bk, &s_err)) break;
    const double sr = gc_sum(a.norms + 1, 4, a.P), sb = gc_sum(a.norms + 2, 4, a.P);
    const double res0 = sqrt(sr), bn = sqrt(sb);
    if (first) {
      rn0 = bn > 0.0 ? bn : res0;
      ttol = fmax(a.rtol * rn0, a.atol);
      first = 0;
    }
    if (res0 == 0.0 || res0 <= ttol) reason = 1;
    else if (res0 >= 1e4 * rn0) reason = -1;
    if (reason) break;
    if (threadIdx.x == 0) grs[0] = res0;
    for (int i = gm_opaque(threadIdx.x); i < nf; i += T) vfull[i] = gc_ld(a.wfull + i) / res0;
    for (int r = gm_opaque(threadIdx.x); r < Rc; r += T) Vl[r] = gc_ld(a.wfull + r0 + r) / res0;
    __syncthreads();
    int it = 0;
    bool ok = true;
    while (true) {
      matvec();  // w = S v_it, local rows
      // partial CGS coefficients <w, v_j>, j = 0..it, over the local rows:
      // 16 lanes per coefficient
      {
        const int t = gm_opaque(threadIdx.x), sub = t & 15;
        double wr[kCoopRowsPer16];  // rows sub, sub + 16, ... of w (R <= 16 kCoopRowsPer16)
#pragma unroll
        for (int q = 0; q < kCoopRowsPer16; ++q) wr[q] = sub + 16 * q < Rc ? wl[sub + 16 * q] : 0.0;
        for (int j = t >> 4; j <= it; j += T / 16) {
          double p = 0.0;
#pragma unroll
          for (int q = 0; q < kCoopRowsPer16; ++q)
            if (sub + 16 * q < Rc) p += wr[q] * Vl[(long)j * R + sub + 16 * q];
          p = gm_row_sum16(p);
          if (sub == 0) gc_st(a.dots + (long)c * (kmax + 2) + j, p);
        }
      }
      if (!gc_sync(a.sync, a.P, c, ++bk, &s_err)) {
        ok = false;
        break;
      }
      for (int j = gm_opaque(threadIdx.x); j <= it; j += T) hcol[j] = gc_sum(a.dots + j, kmax + 2, a.P);
      __syncthreads();
      // w -= sum_j h_j v_j on the local rows (wavefront q4: rows 4 q4 ..
      // 4 q4 + 3, lane l: j = l, l + 64, ...); partial ||w||^2; w rows out
      {
        const int q4 = gm_opaque(threadIdx.x) >> 6, l = gm_opaque(threadIdx.x) & 63;
        double pn = 0.0;
        if (4 * q4 < Rc) {  // uniform over the wavefront
          const int rb = 4 * q4;
          const bool ok1 = rb + 1 < Rc, ok2 = rb + 2 < Rc, ok3 = rb + 3 < Rc;
          double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
          for (int j = l; j <= it; j += 64) {
            const double hj = -hcol[j];
            const double* vj = Vl + (long)j * R + rb;
            acc0 += hj * vj[0];
            acc1 += hj * (ok1 ? vj[1] : 0.0);
            acc2 += hj * (ok2 ? vj[2] : 0.0);
            acc3 += hj * (ok3 ? vj[3] : 0.0);
          }
          const double sum = gc_reduce4(acc0, acc1, acc2, acc3);
          const int row = gc_reduce4_row(q4, l);
          if (l < 4 && row < Rc) {
            const double wn = wl[row] + sum;
            wl[row] = wn;
            gc_st(a.wfull + r0 + row, wn);
            pn = wn * wn;
          }
        } else if (scan_early && q4 == T / 64 - 1) {
          rotate_prev(it, l);
        }
        pn = gm_block_sum2(pn, red, T);
        if (threadIdx.x == 0) gc_st(a.norms + c * 4, pn);
      }
      if (!gc_sync(a.sync, a.P, c, ++bk, &s_err)) {
        ok = false;
        break;
      }
      // this thread's entries of the new vector (full and local rows), loaded
      // with the norm partials so that both take one memory round trip
      const int t0 = gm_opaque(threadIdx.x);
      const double wpre = t0 < nf ? gc_ld(a.wfull + t0) : 0.0;
      const double wpre_l = t0 < Rc ? gc_ld(a.wfull + r0 + t0) : 0.0;
      const double tt = sqrt(gc_sum(a.norms, 4, a.P));
      if (gm_opaque(threadIdx.x) < 64) {
        const int lane = gm_opaque(threadIdx.x);
        if (!scan_early) rotate_prev(it, lane);
        if (lane == 0) {
          hrot[it + 1] = tt;
          // tt < min(|tt / grs[it]|, 1e-30) (KSPGMRESCycle's happy-breakdown
          // test); the quotient is needed only when tt < 1e-30
          const int hap = tt < 1e-30 && tt < fabs(tt / grs[it]);
          double res;
          if (!hap) {
            const double t1 = sqrt(hrot[it] * hrot[it] + hrot[it + 1] * hrot[it + 1]);
            const double cc = hrot[it] / t1, sn = hrot[it + 1] / t1;
            lcc[it] = cc;
            lss[it] = sn;
            grs[it + 1] = -sn * grs[it];
            grs[it] = cc * grs[it];
            hrot[it] = cc * hrot[it] + sn * hrot[it + 1];
            res = fabs(grs[it + 1]);
          } else {
            res = 0.0;
          }
          int rs = 0;
          if (res <= ttol) rs = 1;
          else if (res >= 1e4 * rn0) rs = -1;
          if (hap && !rs) rs = 1;
          s_flag[0] = rs;
          s_flag[1] = hap;
          s_val[0] = tt;
        }
      }
      __syncthreads();
      reason = s_flag[0];
      const int hap = s_flag[1];
      for (int j = gm_opaque(threadIdx.x); j <= it + 1; j += T) Hc[(long)it * ld + j] = hrot[j];
      if (!hap) {
        const double inv = 1.0 / s_val[0];
        if (t0 < nf) vfull[t0] = wpre * inv;
        for (int i = t0 + T; i < nf; i += T) vfull[i] = gc_ld(a.wfull + i) * inv;
        if (t0 < Rc) Vl[(long)(it + 1) * R + t0] = wpre_l * inv;
      }
      ++it;
      ++its;
      __syncthreads();
      if (reason || it >= kmax || its >= a.maxit) break;
    }
    if (!ok) break;
    // ---- KSPGMRESBuildSoln: back substitution on this workgroup's copy of
    // the rotated Hessenberg (chunks staged in LDS, the serial sweep by one
    // wavefront), then x += V y on the local rows
    const int top = it - 1;
    for (int j = gm_opaque(threadIdx.x); j <= top; j += T) yv[j] = grs[j];
    for (int k1 = top; k1 >= 0; k1 -= a.bsc) {
      const int k0 = k1 - a.bsc + 1 < 0 ? 0 : k1 - a.bsc + 1;
      const int nc = k1 - k0 + 1;
      __syncthreads();
      for (int q = gm_opaque(threadIdx.x); q < nc * (k1 + 1); q += T) {
        const int cq = q / (k1 + 1), rr = q - cq * (k1 + 1);
        part[(long)cq * ld + rr] = Hc[(long)(k0 + cq) * ld + rr];
      }
      __syncthreads();
      if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        for (int k = k1; k >= k0; --k) {
          const double* col = part + (long)(k - k0) * ld;
          const double yk = yv[k] / col[k];
          for (int j = lane; j < k; j += 64) yv[j] = yv[j] - col[j] * yk;
          if (lane == 0) yv[k] = yk;
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
    __syncthreads();
    for (int r = gm_opaque(threadIdx.x); r < Rc; r += T) {
      double acc = 0.0;
      for (int j = 0; j <= top; ++j) acc += yv[j] * Vl[(long)j * R + r];
      gc_st(a.x + r0 + r, gc_ld(a.x + r0 + r) + acc);
    }
    if (reason) break;
    if (its >= a.maxit) {
      reason = -2;
      break;
    }
    if (!gc_sync(a.sync, a.P, c, ++bk, &s_err)) break;  // every workgroup's rows of x before the next residual
  }
  if (threadIdx.x == 0) {
    if (s_err) {
      a.s->reason = -3;
      a.s->its = its;
    } else if (c == 0) {
      a.s->its = its;
      a.s->reason = reason;
      a.s->rn0 = rn0;
      a.s->ttol = ttol;
    }
  }
}

// dynamic LDS of k_gmres_coop (doubles): vfull nf, Vl (kmax+1) R, wl R,
// part kGmT, hcol kmax+2, lcc/lss kmax each, grs/yv kmax+1 each, hrot kmax+2
static size_t coop_lds(int nf, int kmax, int R) {
  return sizeof(double) * ((size_t)nf + (size_t)(kmax + 1) * R + R + kGmT + (size_t)6 * kmax + 6);
}

bool gmres_coop_plan(int nf, int kmax, size_t max_lds, int* P, int* R) {
  static const int cand[] = {2, 4, 8, 16, 32};
  if (nf > 64 * kCoopColQ) return false;
  for (int p : cand) {
    const int r = (nf + p - 1) / p;
    if (r < 1 || r > 16 * kCoopRowsPer16 || r > 4 * (kGmT / 64) || (p - 1) * r >= nf) continue;
    if (kGmT / (kmax + 1) < 1) continue;  // back-substitution chunk in the part buffer
    if (coop_lds(nf, kmax, r) + 1024 > max_lds) continue;
    *P = p;
    *R = r;
    return true;
  }
  return false;
}

size_t gmres_coop_ws(int nf, int kmax, int P) {
  return kCoopSyncDoubles + (size_t)P * (kmax + 2) + (size_t)4 * P + (size_t)nf + (size_t)P * (kmax + 1) * kmax;
}

void gmres_coop(hipStream_t st, int nf, int kmax, int maxit, int P, int R, const double* S, const double* b, double* x,
                GmresDev* s, double rtol, double atol, double* ws, int xcd, int nxcd) {
  GmCoopArgs a{};
  a.nf = nf; a.kmax = kmax; a.maxit = maxit; a.P = P; a.R = R; a.xcd = xcd;
  a.bsc = kGmT / (kmax + 1);
  a.S = S; a.b = b; a.x = x; a.s = s; a.rtol = rtol; a.atol = atol;
  a.sync = reinterpret_cast<unsigned*>(ws);
  a.dots = ws + kCoopSyncDoubles;
  a.norms = a.dots + (size_t)P * (kmax + 2);
  a.wfull = a.norms + (size_t)4 * P;
  a.H = a.wfull + nf;
  // barrier words cleared by a kernel (a memset node misbehaves on graph replays)
  coop_prep(st, ws, s);
  // an ordinary launch: the barrier is the kernel's own (no runtime grid
  // synchronisation resources), P workgroups per XCD of which the P on XCD
  // xcd stay (8P and workgroups 0, 8, 16, ... with xcd < 0), one per CU; a
  // workgroup that is not resident in time makes the others leave with
  // reason -3 instead of waiting for ever
  const int nwg = (xcd < 0 ? 8 : std::max(nxcd, 1)) * P;
  hipLaunchKernelGGL(k_gmres_coop, dim3(nwg), dim3(kGmT), coop_lds(nf, kmax, R), st, a);
  IB_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------
// The dense solve over P workgroups of one XCD for larger systems: nf <=
// 1024 and Krylov spaces of up to 1024 vectors (full GMRES of the airfoil +
// flap system, nf = 800, restart 1000).  The same PETSc iteration as
// k_gmres_coop, laid out for up to 32 rows per workgroup and long Krylov
// spaces:
//  * wavefront q owns local rows 2q and 2q+1; lane l holds S[row][l + 64k],
//    k < CQ, in registers (nf <= 64 CQ), and gw_reduce2 sums a row pair over
//    the wavefront;
//  * the local rows of the basis are stored row-major in LDS (row r of
//    vector j at Vl[r NL + j]), so thread j forms the partial CGS coefficient
//    of vector j from conflict-free LDS reads, and the update w -= V h reads
//    consecutive j across the lanes; vectors j >= NL (past the LDS) are kept
//    in HBM in the same layout;
//  * the CGS coefficients are reduced and scattered: workgroup c sums the P
//    partials of coefficients c, c + P, ... in a fixed butterfly order and
//    publishes them, and after a third grid barrier every workgroup loads
//    all of them, so all hold identical h and take the same branches;
//  * the exchange goes through the L2 of the one XCD the workgroups share
//    (checked with HW_REG_XCC_ID at the start; device scope otherwise);
//  * the Givens step needs only entry it of the new column after the
//    previous rotations, which is sum_j alpha_j h_j with alpha updated as
//    [-s_it alpha, c_it] after each step: one more term in the block
//    reduction of the norm, instead of a serial scan over the rotations;
//  * only workgroup 0 keeps the Hessenberg matrix (in HBM, unrotated, with
//    the rotated diagonal beside it).  At the end of a cycle it applies the
//    rotations to every column (one thread per column), runs the back
//    substitution in blocks of kCwBlk columns (diagonal block by one
//    wavefront from LDS, the rows above by all threads) and publishes y;
//    every workgroup then adds V y to its rows of x.
// ------------------------------------------------------------------------
constexpr int kCwBlk = 32;

struct GmCoopWArgs {
  int nf, kmax, maxit, P, R, NL;
  int xcd;  // the XCD whose workgroups take part (-1: workgroups 0, 8, 16, ...)
  int l2x;  // 1: exchange through one XCD's L2 when every workgroup is found there
  const double* S;
  const double* b;
  double* x;
  GmresDev* s;
  double rtol, atol;
  unsigned* sync;  // [0] error, [8 + c] barriers passed by workgroup c
  double* dots;    // P x (kmax + 2) partial CGS coefficients
  double* norms;   // P x 4 partial squared norms
  double* wfull;   // nf: w (or r) rows of every workgroup
  double* Hg;      // kmax x (kmax + 1): the rotated Hessenberg, column it at Hg + it (kmax + 1)
  double* yg;      // kmax + 1: the back substitution's y
  double* hg;      // kmax + 1: the summed CGS coefficients of the step
  double* Rd;      // kmax + 1: diagonal of the rotated Hessenberg matrix
  double* Vg;      // P x R x (kmax + 1 - NL): basis vectors NL.. of each workgroup's rows
};

// Every lane holds partial sums v0, v1 of two rows; returns in lane 0 the
// sum over the 64 lanes of row 0 and in lane 1 that of row 1.  Pairs swap
// a row (xor 1), then the pairs of a quad are added (xor 2), the quads of a
// row of 16 (row_ror 4, 8) and the four rows of 16 (xor 16, 32).  Fixed
// order; the whole wavefront must be active.
__device__ __forceinline__ double gw_reduce2(double v0, double v1) {
  const int l = gm_opaque(threadIdx.x) & 63;
  const bool odd = l & 1;
  double k = odd ? v1 : v0;
  k += gm_dpp<0xB1>(odd ? v0 : v1);
  k += gm_dpp<0x4E>(k);
  k += gm_dpp<0x124>(k);
  k += gm_dpp<0x128>(k);
  k += __shfl_xor(k, 16);
  k += __shfl_xor(k, 32);
  return k;
}

// Butterfly sum over aligned groups of g = 1, 2, 4, ..., 32 consecutive
// lanes (xor 1 and xor 2 through DPP, xor 4, 8, 16 through lane shuffles):
// every lane of a group ends with the same bits, ((p0 + p1) + (p2 + p3)) +
// ... in one fixed order.  The whole wavefront must be active.
__device__ __forceinline__ double gw_group_sum(double p, int g) {
  if (g >= 2) p += gm_dpp<0xB1>(p);
  if (g >= 4) p += gm_dpp<0x4E>(p);
  if (g >= 8) p += __shfl_xor(p, 4);
  if (g >= 16) p += __shfl_xor(p, 8);
  if (g >= 32) p += __shfl_xor(p, 16);
  return p;
}

// Lanes per CGS coefficient in the dot phase of step it: as many
// as the 1024 threads allow, up to 8 (groups never straddle a wavefront).
__device__ __forceinline__ int gw_lanes_per_coef(int it) {
  const int n = it + 1;
  return 8 * n <= kGmT ? 8 : (4 * n <= kGmT ? 4 : (2 * n <= kGmT ? 2 : 1));
}

// Wavefront-wide sum of p[q * stride] over q < P (lane q loads one value,
// one memory round trip; fixed reduction order, so every workgroup that
// runs it gets the same bits).

// Exchange between the workgroups of a cooperative solve.  Loads are
// device-scope atomic loads: they bypass the CU's L1 and are served by L2
// (measured on MI355X: ~90 ns per dependent load, against ~56 ns for an L1
// hit).  Stores: device scope, coherent across XCDs, or with X2 (every
// workgroup on one XCD, checked at the start of the launch) workgroup-scope
// stores, which complete at that XCD's L2 -- the one L2 every load of the
// exchange reads -- instead of at memory.  Each wavefront waits for its
// stores to complete before a barrier either way.
template <bool X2>
__device__ __forceinline__ double xc_ld(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool X2>
__device__ __forceinline__ void xc_st(double* p, double v) {
  if constexpr (X2) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool X2>
__device__ __forceinline__ unsigned xc_ldu(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool X2>
__device__ __forceinline__ void xc_stu(unsigned* p, unsigned v) {
  if constexpr (X2) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Grid barrier number k over the P workgroups, as gc_sync, with the
// exchange stores of X2.
template <bool X2>
__device__ __forceinline__ bool xc_sync(unsigned* sync, int P, int c, unsigned k, int* s_err) {
  gc_wait_stores();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x < 64) {
    unsigned* err = sync;
    unsigned* fl = sync + 8;
    const int lane = threadIdx.x;
    if (lane == 0) xc_stu<X2>(fl + c, k);
    long n = 0;
    while (true) {
      const unsigned v = lane < P ? xc_ldu<X2>(fl + lane) : k;
      if (__all(v >= k)) break;
      if (xc_ldu<X2>(err) != 0u || ++n > kCoopSpin) {
        if (lane == 0) {
          xc_stu<X2>(err, 1u);
          *s_err = 1;
        }
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  __syncthreads();
  return *s_err == 0;
}

// Block-wide sums of two values (every thread gets both), in a fixed order:
// each wavefront's sum (gm_wave_sum), then the 16 wavefront sums by a row of
// 16 lanes (gm_row_sum16).  red holds at least 2 kGmT / 64 doubles; it is
// read after the one workgroup barrier here, so the caller must pass another
// barrier before red is written again.
__device__ __forceinline__ void gw_block_sum_pair(double& a, double& b, double* red) {
  const int t = gm_opaque(threadIdx.x), l = t & 63;
  a = gm_wave_sum(a);
  b = gm_wave_sum(b);
  if (l == 0) {
    red[t >> 6] = a;
    red[kGmT / 64 + (t >> 6)] = b;
  }
  __syncthreads();
  // lanes 0..15 of every wavefront take the 16 partials of a, lanes 16..31
  // those of b: one LDS read per lane, then a sum over each row of 16 lanes
  double v = l < 2 * (kGmT / 64) ? red[l] : 0.0;
  v = gm_row_sum16(v);
  a = gm_readlane(v, 0);
  b = gm_readlane(v, 16);
}

template <bool X2>
__device__ __forceinline__ double gw_wave_gather(const double* p, long stride, int P, int lane) {
  const double v = lane < P ? xc_ld<X2>(p + (long)lane * stride) : 0.0;
  return gm_wave_sum(v);
}

template <int CQ, bool X2>
__device__ __forceinline__ void coopw_solve(const GmCoopWArgs& a, int c, unsigned bk, double* dyn, double* red,
                                            double* s_val, int* s_flag, int* s_err) {
  const int nf = a.nf, kmax = a.kmax, ld = kmax + 1, T = kGmT, R = a.R, NL = a.NL;
  const int r0 = c * R, Rc = max(0, min(nf, r0 + R) - r0);
  const int ldg = ld - NL;  // row stride of the basis vectors kept in HBM
  double* Vgc = a.Vg + (size_t)c * R * (size_t)max(ldg, 0);
  double* wl = dyn;                    // 32: local rows of w
  double* hcol = wl + 32;              // kmax + 2: CGS coefficients (y at a cycle's end)
  double* alpha = hcol + kmax + 2;     // kmax + 2: entry it of the rotated column = sum_j alpha_j h_j
  double* lcc = alpha + kmax + 2;      // kmax
  double* lss = lcc + kmax;            // kmax
  double* grs = lss + kmax;            // kmax + 1 (y of the back substitution in place)
  double* stg = grs + kmax + 1;        // kCwBlk^2: a diagonal block of the back substitution
  double* ub = stg + kCwBlk * kCwBlk;  // nf: the vector of a product, copied once from L2
  double* Vl = ub + nf;                // R NL: local rows of basis vectors 0..NL-1, row-major
  // rows 2 wq and 2 wq + 1 of S, columns ln + 64 k, in registers (0 outside the matrix)
  double s0[CQ], s1[CQ];
  {
    const int wq = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const int ra = 2 * wq, rb = 2 * wq + 1;
#pragma unroll
    for (int k = 0; k < CQ; ++k) {
      const int col = ln + 64 * k;
      s0[k] = ra < Rc && col < nf ? a.S[(long)col * nf + r0 + ra] : 0.0;
      s1[k] = rb < Rc && col < nf ? a.S[(long)col * nf + r0 + rb] : 0.0;
    }
  }
  __syncthreads();

  // S[local rows] u for a vector u the workgroups left in L2: the
  // workgroup copies u into LDS once (ub; one load per thread), then every
  // lane reads the CQ entries it multiplies from there, so the L2 serves u
  // once per workgroup instead of once per wavefront.  Lanes 0 and 1 of
  // wavefront q get rows 2q and 2q + 1 (0 in the other lanes).  Every
  // thread must call it (a workgroup barrier follows the copy); ub is next
  // written by the following call, after at least one more barrier.
  auto matvec_g = [&](const double* u) -> double {
    for (int i = gm_opaque(threadIdx.x); i < nf; i += T) ub[i] = xc_ld<X2>(u + i);
    __syncthreads();
    const int q = gm_opaque(threadIdx.x) >> 6, l = gm_opaque(threadIdx.x) & 63;
    double sum = 0.0;
    if (2 * q < Rc) {  // uniform over the wavefront
      double v[CQ];
#pragma unroll
      for (int k = 0; k < CQ; ++k) v[k] = ub[min(l + 64 * k, nf - 1)];
      double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
      for (int k = 0; k < CQ; ++k) {
        acc0 += s0[k] * v[k];  // s0, s1 are 0 past the matrix
        acc1 += s1[k] * v[k];
      }
      sum = gw_reduce2(acc0, acc1);
    }
    return sum;
  };
  // lanes 0/1 of the row-pair wavefronts store their row of S u times scale
  auto store_rows = [&](double sum, double scale) {
    const int q = gm_opaque(threadIdx.x) >> 6, l = gm_opaque(threadIdx.x) & 63;
    const int row = 2 * q + (l & 1);
    if (2 * q < Rc && l < 2 && row < Rc) wl[row] = sum * scale;
  };
  // acc0/acc1 += sum_{j = l, l + 64, ... < jn} coef[j] V(ra / rb, j)
  auto rowpair = [&](const double* coef, int jn, int ra, int rb, int l, double& acc0, double& acc1) {
    const int ja = min(jn, NL);
    int j = l;
#pragma unroll 8
    for (; j < ja; j += 64) {
      const double h = coef[j];
      acc0 += h * Vl[ra * NL + j];
      acc1 += h * Vl[rb * NL + j];
    }
    for (; j < jn; j += 64) {
      const double h = coef[j];
      acc0 += h * Vgc[(size_t)ra * ldg + (j - NL)];
      acc1 += h * Vgc[(size_t)rb * ldg + (j - NL)];
    }
  };
  int its = 0, reason = 0, first = 1;
  double rn0 = 0.0, ttol = 0.0;
  while (true) {
    // ---- cycle start: r = b - S x (x from every workgroup's rows)
    store_rows(matvec_g(a.x), 1.0);
    __syncthreads();
    double pr = 0.0, pb = 0.0;
    for (int r = gm_opaque(threadIdx.x); r < Rc; r += T) {
      const double bv = a.b[r0 + r];
      const double rv = bv - wl[r];
      xc_st<X2>(a.wfull + r0 + r, rv);
      pr += rv * rv;
      pb += bv * bv;
    }
    pr = gm_block_sum2(pr, red, T);
    pb = gm_block_sum2(pb, red, T);
    if (threadIdx.x == 0) {
      xc_st<X2>(a.norms + c * 4 + 1, pr);
      xc_st<X2>(a.norms + c * 4 + 2, pb);
    }
    if (!xc_sync<X2>(a.sync, a.P, c, ++bk, s_err)) break;
    if (threadIdx.x < 64) {
      const double sr = gw_wave_gather<X2>(a.norms + 1, 4, a.P, threadIdx.x);
      const double sb = gw_wave_gather<X2>(a.norms + 2, 4, a.P, threadIdx.x);
      if (threadIdx.x == 0) {
        s_val[0] = sr;
        s_val[1] = sb;
      }
    }
    __syncthreads();
    const double res0 = sqrt(s_val[0]), bn = sqrt(s_val[1]);
    if (first) {
      rn0 = bn > 0.0 ? bn : res0;
      ttol = fmax(a.rtol * rn0, a.atol);
      first = 0;
    }
    if (res0 == 0.0 || res0 <= ttol) reason = 1;
    else if (res0 >= 1e4 * rn0) reason = -1;
    if (reason) break;
    if (threadIdx.x == 0) {
      grs[0] = res0;
      alpha[0] = 1.0;
    }
    // v_0 = r / res0 and w = S v_0 (the row sums of S r scaled by 1 / res0)
    const double inv0 = 1.0 / res0;
    store_rows(matvec_g(a.wfull), inv0);
    for (int r = gm_opaque(threadIdx.x); r < Rc; r += T) {
      const double v = xc_ld<X2>(a.wfull + r0 + r) * inv0;
      if (NL > 0) Vl[r * NL] = v;
      else Vgc[(size_t)r * ldg] = v;
    }
    __syncthreads();
    int it = 0;
    bool ok = true;
    while (true) {
      // here wl = S v_it on the local rows
      // partial CGS coefficient <w, v_j> over the local rows: g lanes per
      // j, lane sub takes rows sub, sub + g, ...
      const int g = gw_lanes_per_coef(it);
      {
        const int t = gm_opaque(threadIdx.x), j = t / g, sub = t - j * g;
        double p = 0.0;
        if (j <= it) {
          if (j < NL) {
#pragma unroll 8
            for (int r = sub; r < Rc; r += g) p += wl[r] * Vl[r * NL + j];
          } else {
            const double* vg = Vgc + (j - NL);
#pragma unroll 8
            for (int r = sub; r < Rc; r += g) p += wl[r] * vg[(size_t)r * ldg];
          }
        }
        p = gw_group_sum(p, g);
        if (j <= it && sub == 0) xc_st<X2>(a.dots + (long)c * (kmax + 2) + j, p);
      }
      if (!xc_sync<X2>(a.sync, a.P, c, ++bk, s_err)) {
        ok = false;
        break;
      }
      // h_j: workgroup c sums the P partials of j = c, c + P, ... (lane q of
      // a group of P lanes loads partial q; fixed butterfly over the group)
      // and publishes it; after a barrier every workgroup loads all h_j.
      // Each partial is read once, not once per workgroup.
      {
        const int t = gm_opaque(threadIdx.x), jj = t / a.P, q = t - jj * a.P;
        const int j = c + jj * a.P;
        double h = j <= it ? xc_ld<X2>(a.dots + (long)q * (kmax + 2) + j) : 0.0;
        h = gw_group_sum(h, a.P);
        if (j <= it && q == 0) xc_st<X2>(a.hg + j, h);
      }
      if (!xc_sync<X2>(a.sync, a.P, c, ++bk, s_err)) {
        ok = false;
        break;
      }
      for (int j = gm_opaque(threadIdx.x); j <= it; j += T) hcol[j] = xc_ld<X2>(a.hg + j);
      __syncthreads();
      // w -= sum_j h_j v_j on the local rows; partial ||w||^2; w rows out.
      // The same block reduction forms sum_j alpha_j h_j, entry it of the
      // new Hessenberg column after the previous rotations (identical bits
      // in every workgroup: the same alpha, h and order).
      double hv;
      {
        const int q = gm_opaque(threadIdx.x) >> 6, l = gm_opaque(threadIdx.x) & 63;
        double pn = 0.0;
        if (2 * q < Rc) {  // uniform over the wavefront
          const int ra = 2 * q, rb = 2 * q + 1 < Rc ? 2 * q + 1 : 2 * q;
          double acc0 = 0.0, acc1 = 0.0;
          rowpair(hcol, it + 1, ra, rb, l, acc0, acc1);
          const double sum = gw_reduce2(-acc0, -acc1);
          const int row = 2 * q + (l & 1);
          if (l < 2 && row < Rc) {
            const double wn = wl[row] + sum;
            wl[row] = wn;
            xc_st<X2>(a.wfull + r0 + row, wn);
            pn = wn * wn;
          }
        }
        const int t = gm_opaque(threadIdx.x);
        double ah = t <= it ? alpha[t] * hcol[t] : 0.0;  // it < kGmT
        gw_block_sum_pair(pn, ah, red);
        hv = ah;
        if (threadIdx.x == 0) xc_st<X2>(a.norms + c * 4, pn);
      }
      if (!xc_sync<X2>(a.sync, a.P, c, ++bk, s_err)) {
        ok = false;
        break;
      }
      // this thread's local-row entry of the new vector w (unnormalised),
      // the row sums of S w for the next step (from L2, beside the norm
      // partials and the Givens step of wavefront 0)
      const int t0 = gm_opaque(threadIdx.x);
      // wavefront 0 loads the P norm partials before the product's vector
      // (lane q partial q, as gw_wave_gather), so that the two L2 round
      // trips overlap
      const double npart = t0 < a.P ? xc_ld<X2>(a.norms + (long)t0 * 4) : 0.0;
      const double swsum = matvec_g(a.wfull);
      const double wpre_l = t0 < Rc ? ub[r0 + t0] : 0.0;
      if (gm_opaque(threadIdx.x) < 64) {
        const int lane = gm_opaque(threadIdx.x);
        const double tt = sqrt(gm_wave_sum(npart));
        if (lane == 0) {
          // KSPGMRESUpdateHessenberg: tt < min(|tt / grs[it]|, 1e-30) is a
          // happy breakdown (the quotient is needed only when tt < 1e-30)
          const int hap = tt < 1e-30 && tt < fabs(tt / grs[it]);
          double res, diag, cc = 1.0, sn = 0.0;
          if (!hap) {
            const double t1 = sqrt(hv * hv + tt * tt);
            cc = hv / t1;
            sn = tt / t1;
            lcc[it] = cc;
            lss[it] = sn;
            grs[it + 1] = -sn * grs[it];
            grs[it] = cc * grs[it];
            diag = cc * hv + sn * tt;
            res = fabs(grs[it + 1]);
          } else {
            diag = hv;
            res = 0.0;
          }
          if (c == 0) a.Rd[it] = diag;
          int rs = 0;
          if (res <= ttol) rs = 1;
          else if (res >= 1e4 * rn0) rs = -1;
          if (hap && !rs) rs = 1;
          s_flag[0] = rs;
          s_flag[1] = hap;
          s_val[0] = tt;
          s_val[1] = sn;
          red[2 * (kGmT / 64)] = cc;
        }
      }
      __syncthreads();
      reason = s_flag[0];
      const int hap = s_flag[1];
      // workgroup 0 keeps the unrotated column (rows 0..it, then tt); the
      // rotations are applied to all columns at the cycle's end
      if (c == 0) {
        for (int j = gm_opaque(threadIdx.x); j <= it; j += T) a.Hg[(long)it * ld + j] = hcol[j];
        if (threadIdx.x == 0) a.Hg[(long)it * ld + it + 1] = s_val[0];
      }
      if (!hap) {
        // v_(it+1) = w / tt, and S v_(it+1) for the next step; alpha for the
        // next column: every entry times -s_it, then c_it
        const double inv = 1.0 / s_val[0];
        store_rows(swsum, inv);
        if (t0 < Rc) {
          const int j = it + 1;
          if (j < NL) Vl[t0 * NL + j] = wpre_l * inv;
          else Vgc[(size_t)t0 * ldg + (j - NL)] = wpre_l * inv;
        }
        const double sn = s_val[1];
        for (int j = gm_opaque(threadIdx.x); j <= it; j += T) alpha[j] = -sn * alpha[j];
        if (threadIdx.x == 0) alpha[it + 1] = red[2 * (kGmT / 64)];
      }
      ++it;
      ++its;
      __syncthreads();
      if (reason || it >= kmax || its >= a.maxit) break;
    }
    if (!ok) break;
    // ---- KSPGMRESBuildSoln: workgroup 0 solves the rotated triangular
    // system (grs becomes y in place) and publishes y
    const int top = it - 1;
    if (c == 0) {
      // the previous rotations on every column k (thread k; the same
      // recurrence as KSPGMRESUpdateHessenberg), its diagonal from the step
      for (int k = gm_opaque(threadIdx.x); k <= top; k += T) {
        double* col = a.Hg + (long)k * ld;
        double carry = col[0];
        for (int j = 0; j < k; ++j) {
          const double hn = col[j + 1];
          col[j] = lcc[j] * carry + lss[j] * hn;
          carry = -lss[j] * carry + lcc[j] * hn;
        }
        col[k] = a.Rd[k];
      }
      __syncthreads();
      for (int kb1 = top; kb1 >= 0; kb1 -= kCwBlk) {
        const int kb0 = max(0, kb1 - kCwBlk + 1), nb = kb1 - kb0 + 1;
        // column k of the diagonal block: stg[k kCwBlk + i] = H(kb0 + i, kb0 + k), i <= k
        for (int q = gm_opaque(threadIdx.x); q < nb * nb; q += T) {
          const int k = q / nb, i = q - k * nb;
          if (i <= k) stg[k * kCwBlk + i] = a.Hg[(long)(kb0 + k) * ld + kb0 + i];
        }
        __syncthreads();
        if (threadIdx.x < 64) {
          const int lane = threadIdx.x;
          double ti = lane < nb ? grs[kb0 + lane] : 0.0;
          for (int k = nb - 1; k >= 0; --k) {
            const double yk = gm_readlane(ti, k) / stg[k * kCwBlk + k];
            if (lane < k) ti -= stg[k * kCwBlk + lane] * yk;
            if (lane == k) ti = yk;
          }
          if (lane < nb) grs[kb0 + lane] = ti;
        }
        __syncthreads();
        // the rows above the block, columns of the block in descending order
        for (int i = gm_opaque(threadIdx.x); i < kb0; i += T) {
          double t = grs[i];
          for (int k = nb - 1; k >= 0; --k) t -= a.Hg[(long)(kb0 + k) * ld + i] * grs[kb0 + k];
          grs[i] = t;
        }
        __syncthreads();
      }
      for (int j = gm_opaque(threadIdx.x); j <= top; j += T) xc_st<X2>(a.yg + j, grs[j]);
    }
    if (!xc_sync<X2>(a.sync, a.P, c, ++bk, s_err)) break;
    for (int j = gm_opaque(threadIdx.x); j <= top; j += T) hcol[j] = xc_ld<X2>(a.yg + j);
    __syncthreads();
    // x += V y on the local rows
    {
      const int q = gm_opaque(threadIdx.x) >> 6, l = gm_opaque(threadIdx.x) & 63;
      if (2 * q < Rc) {
        const int ra = 2 * q, rb = 2 * q + 1 < Rc ? 2 * q + 1 : 2 * q;
        double acc0 = 0.0, acc1 = 0.0;
        rowpair(hcol, top + 1, ra, rb, l, acc0, acc1);
        const double sum = gw_reduce2(acc0, acc1);
        const int row = 2 * q + (l & 1);
        if (l < 2 && row < Rc) xc_st<X2>(a.x + r0 + row, xc_ld<X2>(a.x + r0 + row) + sum);
      }
    }
    if (reason) break;
    if (its >= a.maxit) {
      reason = -2;
      break;
    }
    if (!xc_sync<X2>(a.sync, a.P, c, ++bk, s_err)) break;  // every workgroup's rows of x before the next residual
  }
  if (threadIdx.x == 0) {
    if (*s_err) {
      a.s->reason = -3;
      a.s->its = its;
    } else if (c == 0) {
      a.s->its = its;
      a.s->reason = reason;
      a.s->rn0 = rn0;
      a.s->ttol = ttol;
      a.s->xmode = X2 ? 1 : 2;
    }
  }
}


// One cooperative solve by the workgroup of rank c (of a.P, all on one XCD
// when the XCD ids they publish at the first barrier agree).
template <int CQ>
__device__ __forceinline__ void coopw_run(const GmCoopWArgs& a, int c, double* dyn, double* red, double* s_val,
                                          int* s_flag, int* s_err) {
  if (threadIdx.x == 0) {
    *s_err = 0;
    __hip_atomic_store(a.sync + kXccSlot + c, gw_xcc_id() + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // barrier 1 at device scope; then every workgroup reads the P XCD ids
  // (the same values, so the same decision everywhere)
  if (xc_sync<false>(a.sync, a.P, c, 1u, s_err)) {
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      const unsigned v =
          lane < a.P ? __hip_atomic_load(a.sync + kXccSlot + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
      const unsigned v0 = __builtin_amdgcn_readfirstlane(v);
      const bool same = a.l2x && __all(lane >= a.P || v == v0);
      if (lane == 0) s_flag[0] = same;
    }
    __syncthreads();
    if (s_flag[0]) coopw_solve<CQ, true>(a, c, 1u, dyn, red, s_val, s_flag, s_err);
    else coopw_solve<CQ, false>(a, c, 1u, dyn, red, s_val, s_flag, s_err);
  } else if (threadIdx.x == 0) {
    a.s->reason = -3;
    a.s->its = 0;
  }
}

template <int CQ>
__global__ void __launch_bounds__(kGmT) k_gmres_coopw(GmCoopWArgs a) {
  __shared__ int s_c;
  const int c = coop_rank(a.sync, a.P, a.xcd, &s_c);
  if (c < 0) return;
  extern __shared__ double dyn[];
  __shared__ double red[2 * (kGmT / 64) + 1];
  __shared__ double s_val[2];
  __shared__ int s_flag[2];
  __shared__ int s_err;
  coopw_run<CQ>(a, c, dyn, red, s_val, s_flag, &s_err);
}

// Up to kCoopMulti solves in one launch, solve k on XCD a[k].xcd: the
// workgroups on that XCD claim its ranks, those on XCDs no solve names, or
// past a solve's P, leave at once.  Every XCD's workgroups belong to this
// launch, so no solve waits for CUs another solve holds.
constexpr int kCoopMulti = 8;
struct GmCoopWMulti {
  int n;
  GmCoopWArgs a[kCoopMulti];
};

template <int CQ>
__global__ void __launch_bounds__(kGmT) k_gmres_coopw_multi(GmCoopWMulti m) {
  __shared__ int s_c, s_k;
  if (threadIdx.x == 0) {
    const int id = (int)gw_xcc_id();
    int k = -1;
    for (int j = 0; j < m.n; ++j)
      if (m.a[j].xcd == id) k = j;
    int c = -1;
    if (k >= 0) {
      c = (int)__hip_atomic_fetch_add(m.a[k].sync + kClaimSlot, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (c >= m.a[k].P) c = -1;
    }
    s_c = c;
    s_k = k;
  }
  __syncthreads();
  const int c = s_c;
  if (c < 0) return;
  const int k = __builtin_amdgcn_readfirstlane(s_k);
  extern __shared__ double dyn[];
  __shared__ double red[2 * (kGmT / 64) + 1];
  __shared__ double s_val[2];
  __shared__ int s_flag[2];
  __shared__ int s_err;
  coopw_run<CQ>(m.a[k], c, dyn, red, s_val, s_flag, &s_err);
}

// the barrier blocks of the launch's solves cleared and their statuses set
// to reason -3 (one workgroup per solve)
__global__ void k_coop_prep_multi(GmCoopWMulti m) {
  const GmCoopWArgs& a = m.a[blockIdx.x];
  if (threadIdx.x < 2 * kCoopSyncDoubles) a.sync[threadIdx.x] = 0u;
  if (threadIdx.x == 0) {
    a.s->reason = -3;
    a.s->its = 0;
  }
}

// dynamic LDS of k_gmres_coopw without the basis (doubles): wl 32,
// hcol/alpha kmax+2 each, lcc/lss kmax each, grs kmax+1, stg kCwBlk^2, ub nf
static size_t coopw_lds_fixed(int nf, int kmax) {
  return 32 + (size_t)5 * kmax + 5 + (size_t)kCwBlk * kCwBlk + (size_t)nf;
}

bool gmres_coopw_plan(int nf, int kmax, size_t max_lds, int ncu, int* P, int* R, int* NL) {
  static const int cand[] = {8, 16, 32};
  if (nf < 2 || nf > 64 * 16 || kmax < 1 || kmax + 1 > kGmT) return false;
  const size_t budget = max_lds > 2048 ? (max_lds - 2048) / sizeof(double) : 0;
  const size_t fixed = coopw_lds_fixed(nf, kmax);
  if (fixed >= budget) return false;
  for (int p : cand) {
    if (8 * p > ncu) continue;
    const int r = (nf + p - 1) / p;
    if (r > 2 * (kGmT / 64) || (p - 1) * r >= nf) continue;
    const long nl = std::min<long>(kmax + 1, (long)((budget - fixed) / r));
    if (nl < std::min(kmax + 1, 64)) continue;
    *P = p;
    *R = r;
    *NL = (int)nl;
    return true;
  }
  return false;
}

size_t gmres_coopw_ws(int nf, int kmax, int P, int R, int NL) {
  return kCoopSyncDoubles + (size_t)P * (kmax + 2) + (size_t)4 * P + (size_t)nf + (size_t)(kmax + 1) * kmax +
         (size_t)3 * (kmax + 1) + (size_t)P * R * (size_t)std::max(0, kmax + 1 - NL);
}

// the kernel arguments of one solve, its workspace laid out in ws
static GmCoopWArgs coopw_args(int nf, int kmax, int maxit, int P, int R, int NL, const double* S, const double* b,
                              double* x, GmresDev* s, double rtol, double atol, double* ws, int xcd) {
  GmCoopWArgs a{};
  a.nf = nf; a.kmax = kmax; a.maxit = maxit; a.P = P; a.R = R; a.NL = NL; a.xcd = xcd;
  a.S = S; a.b = b; a.x = x; a.s = s; a.rtol = rtol; a.atol = atol;
  a.sync = reinterpret_cast<unsigned*>(ws);
  a.dots = ws + kCoopSyncDoubles;
  a.norms = a.dots + (size_t)P * (kmax + 2);
  a.wfull = a.norms + (size_t)4 * P;
  a.Hg = a.wfull + nf;
  a.yg = a.Hg + (size_t)(kmax + 1) * kmax;
  a.hg = a.yg + (kmax + 1);
  a.Rd = a.hg + (kmax + 1);
  a.Vg = a.Rd + (kmax + 1);
  // IB_GMRES_L2_EXCHANGE=0 keeps the device-scope exchange even when every
  // workgroup is on one XCD (the results are the same bits either way)
  const char* e = std::getenv("IB_GMRES_L2_EXCHANGE");
  a.l2x = e && std::atoi(e) == 0 ? 0 : 1;
  return a;
}

// register columns per lane of the kernel instance for nf
static int coopw_cq(int nf) { return nf <= 64 * 8 ? 8 : nf <= 64 * 13 ? 13 : 16; }

void gmres_coopw(hipStream_t st, int nf, int kmax, int maxit, int P, int R, int NL, const double* S, const double* b,
                 double* x, GmresDev* s, double rtol, double atol, double* ws, int xcd, int nxcd) {
  const GmCoopWArgs a = coopw_args(nf, kmax, maxit, P, R, NL, S, b, x, s, rtol, atol, ws, xcd);
  coop_prep(st, ws, s);
  const size_t lds = sizeof(double) * (coopw_lds_fixed(nf, kmax) + (size_t)R * NL);
  // P workgroups per XCD, of which the P on XCD xcd stay, one per CU
  const int nwg = (xcd < 0 ? 8 : std::max(nxcd, 1)) * P;
  const int cq = coopw_cq(nf);
  if (cq == 8) hipLaunchKernelGGL(k_gmres_coopw<8>, dim3(nwg), dim3(kGmT), lds, st, a);
  else if (cq == 13) hipLaunchKernelGGL(k_gmres_coopw<13>, dim3(nwg), dim3(kGmT), lds, st, a);
  else hipLaunchKernelGGL(k_gmres_coopw<16>, dim3(nwg), dim3(kGmT), lds, st, a);
  IB_LAUNCH_CHECK();
}

bool gmres_coopw_multi_ok(int nf_a, int nf_b) { return coopw_cq(nf_a) == coopw_cq(nf_b); }

void gmres_coopw_multi(hipStream_t st, int n, const CoopWJob* jobs, const int* xcds, int nxcd) {
  if (n < 1 || n > kCoopMulti || n > std::max(nxcd, 1))
    throw HipError("gmres_coopw_multi: between 1 and min(8, XCDs) solves per launch");
  GmCoopWMulti m{};
  m.n = n;
  size_t lds = 0;
  int maxP = 0;
  const int cq = coopw_cq(jobs[0].nf);
  for (int k = 0; k < n; ++k) {
    const CoopWJob& j = jobs[k];
    if (coopw_cq(j.nf) != cq) throw HipError("gmres_coopw_multi: solves of different register layouts");
    m.a[k] = coopw_args(j.nf, j.kmax, j.maxit, j.P, j.R, j.NL, j.S, j.b, j.x, j.s, j.rtol, j.atol, j.ws, xcds[k]);
    lds = std::max(lds, sizeof(double) * (coopw_lds_fixed(j.nf, j.kmax) + (size_t)j.R * j.NL));
    maxP = std::max(maxP, j.P);
  }
  hipLaunchKernelGGL(k_coop_prep_multi, dim3(n), dim3(2 * kCoopSyncDoubles), 0, st, m);
  IB_LAUNCH_CHECK();
  const int nwg = std::max(nxcd, 1) * maxP;
  if (cq == 8) hipLaunchKernelGGL(k_gmres_coopw_multi<8>, dim3(nwg), dim3(kGmT), lds, st, m);
  else if (cq == 13) hipLaunchKernelGGL(k_gmres_coopw_multi<13>, dim3(nwg), dim3(kGmT), lds, st, m);
  else hipLaunchKernelGGL(k_gmres_coopw_multi<16>, dim3(nwg), dim3(kGmT), lds, st, m);
  IB_LAUNCH_CHECK();
}

// the kernel also has a few hundred bytes of static shared memory, so its
// dynamic share stops 1 KiB short of the device limit (as gmres_dense assumes)
void gmres_dense_set_max_lds(size_t bytes) {
  if (bytes <= 1024) return;
  IB_HIP(hipFuncSetAttribute((const void*)k_gmres_dense<24>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(bytes - 1024)));
  IB_HIP(hipFuncSetAttribute((const void*)k_gmres_coop, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(bytes - 1024)));
  IB_HIP(hipFuncSetAttribute((const void*)k_gmres_coopw<8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(bytes - 1024)));
  IB_HIP(hipFuncSetAttribute((const void*)k_gmres_coopw<13>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(bytes - 1024)));
  IB_HIP(hipFuncSetAttribute((const void*)k_gmres_coopw<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(bytes - 1024)));
  IB_HIP(hipFuncSetAttribute((const void*)k_gmres_coopw_multi<8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(bytes - 1024)));
  IB_HIP(hipFuncSetAttribute((const void*)k_gmres_coopw_multi<13>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(bytes - 1024)));
  IB_HIP(hipFuncSetAttribute((const void*)k_gmres_coopw_multi<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(bytes - 1024)));
}

// ------------------------------------------------------------------------
// Dense Cholesky for the numerically stationary path (num_stat): A = U^T U
// from the upper triangle of a column-major n x n matrix, in place, and the
// explicit inverse A^-1 = U^-1 U^-T, so that each step's solve is one
// matrix-vector product.
// ------------------------------------------------------------------------

// row k of U: U(k,k) = sqrt(A(k,k)), U(k,j) = A(k,j) / U(k,k) for j > k
__global__ void __launch_bounds__(256) k_chol_pivot(int n, double* __restrict__ A, int k, int* __restrict__ err) {
  const double d = A[(long)k * n + k];
  const double dk = d > 0.0 ? sqrt(d) : 1.0;
  if (threadIdx.x == 0 && !(d > 0.0)) err[0] = 1;
  for (int j = k + 1 + threadIdx.x; j < n; j += blockDim.x) A[(long)j * n + k] = A[(long)j * n + k] / dk;
  __syncthreads();
  if (threadIdx.x == 0) A[(long)k * n + k] = dk;
}

// trailing update A(i,j) -= U(k,i) U(k,j) for k < i <= j (16 x 16 tiles)
__global__ void __launch_bounds__(256) k_chol_update(int n, double* __restrict__ A, int k) {
  const int i = k + 1 + blockIdx.x * 16 + threadIdx.x % 16;
  const int j = k + 1 + blockIdx.y * 16 + threadIdx.x / 16;
  if (i >= n || j >= n || i > j) return;
  A[(long)j * n + i] = A[(long)j * n + i] - A[(long)i * n + k] * A[(long)j * n + k];
}

// column c of A^-1: U^T y = e_c (forward), U x = y (backward); one workgroup per column
__global__ void __launch_bounds__(256) k_chol_inverse(int n, const double* __restrict__ U, double* __restrict__ Ainv) {
  extern __shared__ double y[];
  const int c = blockIdx.x;
  for (int j = threadIdx.x; j < n; j += blockDim.x) y[j] = j == c ? 1.0 : 0.0;
  __syncthreads();
  for (int i = c; i < n; ++i) {
    const double yi = y[i] / U[(long)i * n + i];
    __syncthreads();
    for (int j = i + 1 + threadIdx.x; j < n; j += blockDim.x) y[j] = y[j] - U[(long)j * n + i] * yi;
    if (threadIdx.x == 0) y[i] = yi;
    __syncthreads();
  }
  for (int i = n - 1; i >= 0; --i) {
    const double xi = y[i] / U[(long)i * n + i];
    __syncthreads();
    for (int j = threadIdx.x; j < i; j += blockDim.x) y[j] = y[j] - U[(long)i * n + j] * xi;
    if (threadIdx.x == 0) y[i] = xi;
    __syncthreads();
  }
  for (int j = threadIdx.x; j < n; j += blockDim.x) Ainv[(long)c * n + j] = y[j];
}

void chol_factor(hipStream_t st, int n, double* A, int* err) {
  for (int k = 0; k < n; ++k) {
    hipLaunchKernelGGL(k_chol_pivot, dim3(1), dim3(256), 0, st, n, A, k, err);
    IB_LAUNCH_CHECK();
    const int m = n - k - 1;
    if (m > 0) {
      hipLaunchKernelGGL(k_chol_update, dim3(ceil_div(m, 16), ceil_div(m, 16)), dim3(256), 0, st, n, A, k);
      IB_LAUNCH_CHECK();
    }
  }
}

void chol_inverse(hipStream_t st, int n, const double* U, double* Ainv, size_t max_lds) {
  const size_t lds = sizeof(double) * (size_t)n;
  if (lds > max_lds) throw HipError("num_stat: Schur matrix too large for the LDS inverse");
  if (lds > 65536) IB_HIP(hipFuncSetAttribute((const void*)k_chol_inverse, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds));
  hipLaunchKernelGGL(k_chol_inverse, dim3(n), dim3(256), lds, st, n, U, Ainv);
  IB_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------
// Multi-kernel Arnoldi step for large systems (the matrix-free form and dense
// matrices past the one-workgroup size): orthogonalisation over many
// workgroups, Givens in one wavefront (affine scan), normalisation.
// ------------------------------------------------------------------------

// w -= sum_j h_j v_j on 64 rows per workgroup (16 wavefronts split the j
// range, partials added in wavefront order); partial ||w||^2 per workgroup
__global__ void __launch_bounds__(1024) k_gm_orth(int nf, int kmax, double* __restrict__ V,
                                                  const GmresDev* __restrict__ s, const double* __restrict__ H,
                                                  double* __restrict__ pnorm) {
  __shared__ double part[16][64];
  if (s->reason) return;
  const int it = s->it, ld = kmax + 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = blockIdx.x * 64 + lane;
  const double* h = H + (long)it * ld;
  double* w = V + (long)(it + 1) * nf;
  const int jn = it + 1, per = (jn + 15) / 16;
  const int j0 = min(jn, wv * per), j1 = min(jn, j0 + per);
  double acc = 0.0;
  if (r < nf) {
    for (int j = j0; j < j1; j += 8) {
      double v[8], hv[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int jj = min(j + c, j1 - 1);
        v[c] = V[(long)jj * nf + r];
        hv[c] = h[jj];
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) acc += -(j + c < j1 ? hv[c] : 0.0) * v[c];
    }
  }
  part[wv][lane] = acc;
  __syncthreads();
  if (wv == 0) {
    double pn = 0.0;
    if (r < nf) {
      double a = w[r];
      for (int q = 0; q < 16; ++q) a += part[q][lane];
      w[r] = a;
      pn = a * a;
    }
    pn = gm_wave_sum(pn);
    if (lane == 0) pnorm[blockIdx.x] = pn;
  }
}

// Givens update of the new Hessenberg column (KSPGMRESUpdateHessenberg) by
// wavefront 0: tt from the partial norms (block order), the previous
// rotations as an affine scan, the new rotation, residual estimate and
// stopping tests; advances the status to the next iteration.  Then the
// whole workgroup forms v_(it+1) = w / tt (skipped after a happy breakdown;
// a vector closed at convergence is never used).  Only wavefront 0 reads the
// status: the others take what they need from LDS after the barrier, since
// wavefront 0 rewrites the status while they may not have started.
__global__ void __launch_bounds__(256) k_gm_givens(int kmax, int nf, double* __restrict__ V, GmresDev* s,
                                                   double* __restrict__ H, double* __restrict__ cc,
                                                   double* __restrict__ ss, double* __restrict__ grs,
                                                   const double* __restrict__ pnorm, int nblk) {
  __shared__ double sh_tt;
  __shared__ int sh_go, sh_col;
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    if (s->reason) {
      if (lane == 0) sh_go = 0;
    } else {
      const int it = s->it, ld = kmax + 1;
      double* hcol = H + (long)it * ld;
      double a = 0.0;
      for (int b = 0; b < nblk; ++b) a += pnorm[b];
      const double tt = sqrt(a);
      double carry = hcol[0];
      for (int base = 0; base < it; base += 64) {
        const int j = base + lane;
        const bool act = j < it;
        const double c = act ? cc[j] : 1.0, sn = act ? ss[j] : 0.0, hn = act ? hcol[j + 1] : 0.0;
        double A = -sn, B = c * hn;
        gm_affine_scan(A, B);
        const double cnext = A * carry + B;            // carry_(j+1)
        const double cj = gm_shift_up1(cnext, carry);  // carry_j (lane 0: the carry in)
        if (act) hcol[j] = c * cj + sn * hn;
        carry = gm_readlane(cnext, min(63, it - 1 - base));
      }
      if (lane == 0) {
        hcol[it] = carry;
        hcol[it + 1] = tt;
        double hapbnd = fabs(tt / grs[it]);
        if (hapbnd > 1e-30) hapbnd = 1e-30;
        const int hap = tt < hapbnd;
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
        s->tt = tt;
        s->hap = hap;
        s->it = it + 1;
        s->its = s->its + 1;
        s->res = res;
        int reason = 0;
        if (res <= s->ttol) reason = 1;
        else if (res >= 1e4 * s->rn0) reason = -1;
        if (hap && !reason) reason = 1;
        s->reason = reason;
        sh_go = !reason && !hap;
        sh_col = it + 1;
        sh_tt = tt;
      }
    }
  }
  __syncthreads();
  if (!sh_go) return;
  const double inv = 1.0 / sh_tt;
  double* v = V + (long)sh_col * nf;
  for (int r = threadIdx.x; r < nf; r += blockDim.x) v[r] = v[r] * inv;
}

void gm_start(hipStream_t st, int nf, int kmax, const double* b, const double* Ax, double* V, GmresDev* s, double* H,
              double* grs, double rtol, double atol) {
  (void)kmax;
  (void)H;
  hipLaunchKernelGGL(k_gm_start, dim3(1), dim3(kGmT), 0, st, nf, b, Ax, V, s, grs, rtol, atol);
  IB_LAUNCH_CHECK();
}

void gm_arnoldi(hipStream_t st, int nf, int kmax, double* V, GmresDev* s, double* H, double* cc, double* ss,
                double* grs, double* pnorm) {
  // kmax+1 dot products at most: one wavefront each, 4 per workgroup
  hipLaunchKernelGGL(k_gm_dots, dim3(ceil_div(kmax + 1, 4)), dim3(256), 0, st, nf, kmax, V, s, H);
  IB_LAUNCH_CHECK();
  const int nblk = ceil_div(nf, 64);
  hipLaunchKernelGGL(k_gm_orth, dim3(nblk), dim3(1024), 0, st, nf, kmax, V, s, H, pnorm);
  IB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(256), 0, st, kmax, nf, V, s, H, cc, ss, grs, pnorm, nblk);
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
