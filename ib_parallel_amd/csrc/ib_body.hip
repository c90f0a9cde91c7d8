// ib_body.hip -- immersed-boundary coupling and torsional bodies on CDNA4.
//
// reg (E^T) is a scatter in the reference (VecSetValues ADD_VALUES in tap
// order, src/fsinterface.F90:14-47).  Here it is an atomics-free gather: the
// (entry, tap) pairs are sorted once per weight rebuild by destination face
// with a stable radix sort, run-length encoded, and one thread per touched
// face sums its contributions in tap order (k, then l, then p) -- the
// reference's own addition order, so the result is deterministic run to run.

#include "ib_device.h"
#include "ib_kernels.h"

namespace ibg {

// deltafnc of src/grid.F90:955-1003.  With f32 set, the exponent-less literals
// take their REAL(4) values as in the reference source.  The kernel jumps by
// ~1e-6 at r = 1 and 2 with those literals, and body points on grid lines
// land exactly there, so the weight build evaluates every operation as
// written (no FMA contraction) to pick the same branch as the source order.
__device__ double deltafnc(double x1, double x2, double dr, int f32) {
#pragma clang fp contract(off)
  const double r = fabs(x1 - x2);
  const double r1 = r / dr, r2 = r1 * r1, r3 = r2 * r1, r4 = r3 * r1;
  double d;
#define F(lit) (f32 ? (double)(lit##f) : (lit))
  if (r1 <= 1.0) {
    const double a5 = asin((1.0 / 2.0) * sqrt(3.0) * (2.0 * r1 - 1.0));
    const double a8 = sqrt(1.0 - 12.0 * r2 + 12.0 * r1);
    d = 0.4166666667e-1 * r4 + (F(-.1388888889) + 0.3472222222e-1 * a8) * r3 +
        (-0.7121664902e-1 - 0.5208333333e-1 * a8 + F(0.2405626122) * a5) * r2 +
        (F(-.2405626122) * a5 + F(-.3792313933) + F(.1012731481) * a8) * r1 + 0.8018753741e-1 * a5 -
        0.4195601852e-1 * a8 + F(.6485698427);
  } else if (r1 <= 2.0) {
    const double a6 = asin((1.0 / 2.0) * sqrt(3.0) * (-3.0 + 2.0 * r1));
    const double a9 = sqrt(-23.0 + 36.0 * r1 - 12.0 * r2);
    d = -0.6250000000e-1 * r4 + (F(.4861111111) - 0.1736111111e-1 * a9) * r3 +
        (F(-1.143175026) + 0.7812500000e-1 * a9 + F(-.1202813061) * a6) * r2 +
        (F(.8751991178) + F(.3608439183) * a6 + F(-.1548032407) * a9) * r1 + F(-.2806563809) * a6 +
        0.822848104e-2 + F(.1150173611) * a9;
  } else if (r1 <= 3.0) {
    const double a1 = asin((1.0 / 2.0 * (2.0 * r1 - 5.0)) * sqrt(3.0));
    const double a7 = sqrt(-71.0 - 12.0 * r2 + 60.0 * r1);
    d = 0.2083333333e-1 * r4 + (0.3472222222e-2 * a7 + F(-.2638888889)) * r3 +
        (F(1.214391675) - 0.2604166667e-1 * a7 + 0.2405626122e-1 * a1) * r2 +
        (F(-.1202813061) * a1 + F(-2.449273192) + 0.7262731481e-1 * a7) * r1 + F(.1523563211) * a1 +
        F(1.843201677) - 0.7306134259e-1 * a7;
  } else {
    d = 0.0;
  }
#undef F
  return d / dr;
}

// setup_reg: src/grid.F90:459-528, one thread per (entry, tap)
__global__ void k_setup_reg(Geom g, int nf, const double* __restrict__ xb, const double* __restrict__ yb,
                            const int* __restrict__ comp, double delta, double offx, double offy, int f32,
                            double* __restrict__ w, int* __restrict__ key, int* __restrict__ err,
                            int* __restrict__ changed) {
#pragma clang fp contract(off)
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nf * kTaps) return;
  const int k = idx / kTaps, t = idx - k * kTaps;
  const int l = t / 7 - 3, p = t % 7 - 3;
  const double d2 = 0.5 * delta;
  const int ix = (int)((xb[k] + offx) / delta) + 1;
  const int iy = (int)((yb[k] + offy) / delta) + 1;
  const int ii = ix + p, jj = iy + l;
  const int old = changed ? key[idx] : 0;
  double x, y;
  int kk;
  bool ok;
  if (comp[k] == 1) {
    x = delta * (ix - 1 + p) - offx;
    y = delta * (iy - 1 + l) - offy + d2;
    ok = ii >= 1 && ii <= g.m + 1 && jj >= 1 && jj <= g.n;
    kk = (g.m + 1) * (jj - 1) + (ii - 1);
  } else {
    x = delta * (ix - 1 + p) - offx + d2;
    y = delta * (iy - 1 + l) - offy;
    ok = ii >= 1 && ii <= g.m && jj >= 1 && jj <= g.n + 1;
    kk = g.Fx + g.m * (jj - 1) + (ii - 1);
  }
  if (!ok) {
    *err = 1;
    w[idx] = 0.0;
    key[idx] = 0;
    if (changed && old != 0) *changed = 1;
    return;
  }
  w[idx] = delta * delta * deltafnc(x, xb[k], delta, f32) * deltafnc(y, yb[k], delta, f32);
  key[idx] = kk;
  if (changed && old != kk) *changed = 1;
}

void setup_reg(hipStream_t st, const Geom& g, int nf, const double* xb, const double* yb, const int* comp,
               double delta, double offx, double offy, int f32lit, double* w, int* key, int* err, int* changed) {
  const int nt = nf * kTaps;
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_setup_reg, dim3(ceil_div(nt, 256)), dim3(256), 0, st, g, nf, xb, yb, comp, delta, offx, offy,
                     f32lit, w, key, err, changed);
  IB_LAUNCH_CHECK();
}

// d deltafnc(x1, x2, dr) / d x2 (for the adjoint): the derivative in r1 of
// each branch of deltafnc above, with the same literals, times d r1 / d x2 =
// -sign(x1 - x2) / dr, over dr.
__device__ double deltafnc_dx2(double x1, double x2, double dr, int f32) {
  const double r = fabs(x1 - x2);
  const double r1 = r / dr, r2 = r1 * r1, r3 = r2 * r1;
  const double s3 = sqrt(3.0);
  double dd;
#define F(lit) (f32 ? (double)(lit##f) : (lit))
  if (r1 <= 1.0) {
    const double u = 0.5 * s3 * (2.0 * r1 - 1.0);
    const double a5 = asin(u), da5 = s3 / sqrt(1.0 - u * u);
    const double a8 = sqrt(1.0 - 12.0 * r2 + 12.0 * r1), da8 = (6.0 - 12.0 * r1) / a8;
    const double c3 = F(-.1388888889) + 0.3472222222e-1 * a8;
    const double c2 = -0.7121664902e-1 - 0.5208333333e-1 * a8 + F(0.2405626122) * a5;
    const double c1 = F(-.2405626122) * a5 + F(-.3792313933) + F(.1012731481) * a8;
    dd = 4.0 * 0.4166666667e-1 * r3 + 3.0 * c3 * r2 + 0.3472222222e-1 * da8 * r3 + 2.0 * c2 * r1 +
         (-0.5208333333e-1 * da8 + F(0.2405626122) * da5) * r2 + c1 +
         (F(-.2405626122) * da5 + F(.1012731481) * da8) * r1 + 0.8018753741e-1 * da5 - 0.4195601852e-1 * da8;
  } else if (r1 <= 2.0) {
    const double u = 0.5 * s3 * (-3.0 + 2.0 * r1);
    const double a6 = asin(u), da6 = s3 / sqrt(1.0 - u * u);
    const double a9 = sqrt(-23.0 + 36.0 * r1 - 12.0 * r2), da9 = (18.0 - 12.0 * r1) / a9;
    const double c3 = F(.4861111111) - 0.1736111111e-1 * a9;
    const double c2 = F(-1.143175026) + 0.7812500000e-1 * a9 + F(-.1202813061) * a6;
    const double c1 = F(.8751991178) + F(.3608439183) * a6 + F(-.1548032407) * a9;
    dd = 4.0 * -0.6250000000e-1 * r3 + 3.0 * c3 * r2 - 0.1736111111e-1 * da9 * r3 + 2.0 * c2 * r1 +
         (0.7812500000e-1 * da9 + F(-.1202813061) * da6) * r2 + c1 +
         (F(.3608439183) * da6 + F(-.1548032407) * da9) * r1 + F(-.2806563809) * da6 + F(.1150173611) * da9;
  } else if (r1 <= 3.0) {
    const double u = 0.5 * (2.0 * r1 - 5.0) * s3;
    const double a1 = asin(u), da1 = s3 / sqrt(1.0 - u * u);
    const double a7 = sqrt(-71.0 - 12.0 * r2 + 60.0 * r1), da7 = (30.0 - 12.0 * r1) / a7;
    const double c3 = 0.3472222222e-2 * a7 + F(-.2638888889);
    const double c2 = F(1.214391675) - 0.2604166667e-1 * a7 + 0.2405626122e-1 * a1;
    const double c1 = F(-.1202813061) * a1 + F(-2.449273192) + 0.7262731481e-1 * a7;
    dd = 4.0 * 0.2083333333e-1 * r3 + 3.0 * c3 * r2 + 0.3472222222e-2 * da7 * r3 + 2.0 * c2 * r1 +
         (-0.2604166667e-1 * da7 + 0.2405626122e-1 * da1) * r2 + c1 +
         (F(-.1202813061) * da1 + 0.7262731481e-1 * da7) * r1 + F(.1523563211) * da1 - 0.7306134259e-1 * da7;
  } else {
    dd = 0.0;
  }
#undef F
  const double sg = (x1 - x2) > 0.0 ? 1.0 : ((x1 - x2) < 0.0 ? -1.0 : 0.0);
  return -sg * dd / (dr * dr);
}

// d w / d xb and d w / d yb of every (entry, tap) of setup_reg (weights
// w = delta^2 deltafnc(x, xb) deltafnc(y, yb) at the taps setup_reg chose)
__global__ void k_setup_reg_d(Geom g, int nf, const double* __restrict__ xb, const double* __restrict__ yb,
                              const int* __restrict__ comp, double delta, double offx, double offy, int f32,
                              double* __restrict__ dwx, double* __restrict__ dwy) {
#pragma clang fp contract(off)
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nf * kTaps) return;
  const int k = idx / kTaps, t = idx - k * kTaps;
  const int l = t / 7 - 3, p = t % 7 - 3;
  const double d2 = 0.5 * delta;
  const int ix = (int)((xb[k] + offx) / delta) + 1;
  const int iy = (int)((yb[k] + offy) / delta) + 1;
  double x, y;
  if (comp[k] == 1) {
    x = delta * (ix - 1 + p) - offx;
    y = delta * (iy - 1 + l) - offy + d2;
  } else {
    x = delta * (ix - 1 + p) - offx + d2;
    y = delta * (iy - 1 + l) - offy;
  }
  const double fx = deltafnc(x, xb[k], delta, f32), fy = deltafnc(y, yb[k], delta, f32);
  dwx[idx] = delta * delta * deltafnc_dx2(x, xb[k], delta, f32) * fy;
  dwy[idx] = delta * delta * fx * deltafnc_dx2(y, yb[k], delta, f32);
}

void setup_reg_d(hipStream_t st, const Geom& g, int nf, const double* xb, const double* yb, const int* comp,
                 double delta, double offx, double offy, int f32lit, double* dwx, double* dwy) {
  const int nt = nf * kTaps;
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_setup_reg_d, dim3(ceil_div(nt, 256)), dim3(256), 0, st, g, nf, xb, yb, comp, delta, offx,
                     offy, f32lit, dwx, dwy);
  IB_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------
// The touched-face list of reg (RegCSR): for every face some tap touches, its
// taps in tap order.  Entry k's 49 taps cover the 7 x 7 patch of faces around
// its centre face (tap 24), so the entries touching a face are those centred
// within 3 cells of it, in its component.  Built without sorting the taps:
//   1. the entries sorted by centre key (ties by entry index), one workgroup;
//   2. per tap: every entry whose patch holds the tap's face (7 binary
//      searches, one per patch row, then a check of that entry's own key for
//      the face): their count, the lowest of them (whose tap owns the face)
//      and the tap's rank among them by entry index, which is tap order (an
//      entry touches a face with one tap);
//   3. a scan of the owners (run index) and their counts (offsets) over
//      tiles of 1024 taps;
//   4. each tap writes itself at its owner's offset plus its rank.
// A tap whose stencil left the grid (key clamped to 0, weight 0, the error
// flag set) is matched by no entry and owns nothing.

// centre face (component, i, j) of key f; false outside the grid
struct FaceIJ {
  int x, i, j, ni, nj;
};
__device__ __forceinline__ FaceIJ face_ij(const Geom& g, int f) {
  FaceIJ r;
  r.x = f < g.Fx;
  if (r.x) {
    r.ni = g.m + 1;
    r.nj = g.n;
    r.i = f % r.ni;
    r.j = f / r.ni;
  } else {
    r.ni = g.m;
    r.nj = g.n + 1;
    r.i = (f - g.Fx) % r.ni;
    r.j = (f - g.Fx) / r.ni;
  }
  return r;
}
__device__ __forceinline__ int face_key(const Geom& g, int x, int i, int j) {
  return x ? (g.m + 1) * j + i : g.Fx + g.m * j + i;
}

// 1. bitonic sort of the composite keys (centre key << 32 | entry), n2 a
// power of two <= kSortLds, padded with the largest key; one workgroup
constexpr int kSortLds = 16384;
__global__ void __launch_bounds__(1024) k_sort_centres(int nf, int n2, const int* __restrict__ key,
                                                       int* __restrict__ skey, int* __restrict__ sent) {
  __shared__ long long v[kSortLds];
  for (int q = threadIdx.x; q < n2; q += blockDim.x)
    v[q] = q < nf ? ((long long)key[(long)q * kTaps + kTaps / 2] << 32) | q : 0x7fffffffffffffffLL;
  __syncthreads();
  for (int size = 2; size <= n2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int q = threadIdx.x; q < n2 / 2; q += blockDim.x) {
        const int lo = 2 * q - (q & (stride - 1)), hi = lo + stride;  // pair (lo, hi) of this step
        const bool up = (lo & size) == 0;
        const long long a = v[lo], b = v[hi];
        if ((a > b) == up) {
          v[lo] = b;
          v[hi] = a;
        }
      }
      __syncthreads();
    }
  for (int q = threadIdx.x; q < nf; q += blockDim.x) {
    skey[q] = (int)(v[q] >> 32);
    sent[q] = (int)(v[q] & 0xffffffffLL);
  }
}

// the same sort for larger n2 in global memory, one launch per step
__global__ void k_sort_init(int nf, int n2, const int* __restrict__ key, long long* __restrict__ v) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n2) v[q] = q < nf ? ((long long)key[(long)q * kTaps + kTaps / 2] << 32) | q : 0x7fffffffffffffffLL;
}
__global__ void k_sort_step(int n2, int size, int stride, long long* __restrict__ v) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n2 / 2) return;
  const int lo = 2 * q - (q & (stride - 1)), hi = lo + stride;
  const bool up = (lo & size) == 0;
  const long long a = v[lo], b = v[hi];
  if ((a > b) == up) {
    v[lo] = b;
    v[hi] = a;
  }
}
__global__ void k_sort_split(int nf, const long long* __restrict__ v, int* __restrict__ skey, int* __restrict__ sent) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nf) return;
  skey[q] = (int)(v[q] >> 32);
  sent[q] = (int)(v[q] & 0xffffffffLL);
}

// Visits every entry e whose patch holds face f (e's tap index of f in te):
// fn(e, te) for each, in sorted-centre order.
template <class F>
__device__ __forceinline__ void touching(const Geom& g, int nf, const int* __restrict__ key,
                                         const int* __restrict__ skey, const int* __restrict__ sent, int f, F&& fn) {
  const FaceIJ c = face_ij(g, f);
  for (int l = -3; l <= 3; ++l) {
    const int jc = c.j - l;  // the centre row of entries whose patch row l holds f
    if (jc < 0 || jc >= c.nj) continue;
    const int i_lo = max(c.i - 3, 0), i_hi = min(c.i + 3, c.ni - 1);
    const int k_lo = face_key(g, c.x, i_lo, jc), k_hi = face_key(g, c.x, i_hi, jc);
    int lo = 0, hi = nf;  // first sorted centre >= k_lo
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (skey[mid] < k_lo) lo = mid + 1;
      else hi = mid;
    }
    for (int q = lo; q < nf && skey[q] <= k_hi; ++q) {
      const int e = sent[q];
      const int p = c.i - (skey[q] - face_key(g, c.x, 0, jc));  // f's column offset in e's patch
      const int te = (l + 3) * 7 + (p + 3);
      if (key[(long)e * kTaps + te] == f) fn(e, te);
    }
  }
}

// 2. per tap: the count of entries touching its face (owners only, else 0),
// its rank among them, and the owning tap (-1: the tap is matched by none)
__global__ void k_tap_owner(Geom g, int nf, const int* __restrict__ key, const int* __restrict__ skey,
                            const int* __restrict__ sent, int* __restrict__ own_cnt, int* __restrict__ rank,
                            int* __restrict__ otap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nf * kTaps) return;
  const int k = t / kTaps, f = key[t];
  int n = 0, below = 0, lowest = 0x7fffffff, lowest_te = 0;
  bool self = false;
  touching(g, nf, key, skey, sent, f, [&](int e, int te) {
    ++n;
    below += e < k;
    self = self || e == k;
    if (e < lowest) {
      lowest = e;
      lowest_te = te;
    }
  });
  own_cnt[t] = (self && lowest == k) ? n : 0;
  rank[t] = below;
  otap[t] = self ? lowest * kTaps + lowest_te : -1;
}

// 3. run index and offset of every owner tap, in tap order: tiles of 1024
// taps (one per thread, coalesced); the tile totals of owners and counts,
// their exclusive scan in one workgroup, then each tile's own scan.
// toff[t] = offset (or -1), uniq/cnt/off of each run, nruns.
constexpr int kScanTile = 1024;
// inclusive scan of (a, b) over the workgroup's 1024 threads (Hillis-Steele)
__device__ __forceinline__ void block_scan2(int& a, int& b, int* sa, int* sb) {
  const int tid = threadIdx.x;
  sa[tid] = a;
  sb[tid] = b;
  __syncthreads();
  for (int d = 1; d < kScanTile; d <<= 1) {
    const int xa = tid >= d ? sa[tid - d] : 0, xb = tid >= d ? sb[tid - d] : 0;
    __syncthreads();
    sa[tid] += xa;
    sb[tid] += xb;
    __syncthreads();
  }
  a = sa[tid];
  b = sb[tid];
}
__global__ void __launch_bounds__(kScanTile) k_tap_tile_sum(int ntaps, const int* __restrict__ own_cnt,
                                                           int* __restrict__ tsum) {
  __shared__ int sa[kScanTile], sb[kScanTile];
  const int t = blockIdx.x * kScanTile + threadIdx.x;
  const int c = t < ntaps ? own_cnt[t] : 0;
  int r = c > 0, o = c;
  block_scan2(r, o, sa, sb);
  if (threadIdx.x == kScanTile - 1) {
    tsum[2 * blockIdx.x] = r;
    tsum[2 * blockIdx.x + 1] = o;
  }
}
// one workgroup: exclusive scan of the tile totals in place; nruns
__global__ void __launch_bounds__(kScanTile) k_tap_tile_scan(int ntiles, int* __restrict__ tsum,
                                                            int* __restrict__ nruns) {
  __shared__ int sa[kScanTile], sb[kScanTile];
  int carry_r = 0, carry_o = 0;
  for (int b0 = 0; b0 < ntiles; b0 += kScanTile) {  // uniform trip count
    const int b = b0 + threadIdx.x;
    const int r0 = b < ntiles ? tsum[2 * b] : 0, o0 = b < ntiles ? tsum[2 * b + 1] : 0;
    int r = r0, o = o0;
    block_scan2(r, o, sa, sb);
    if (b < ntiles) {
      tsum[2 * b] = carry_r + r - r0;
      tsum[2 * b + 1] = carry_o + o - o0;
    }
    carry_r += sa[kScanTile - 1];
    carry_o += sb[kScanTile - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *nruns = carry_r;
}
__global__ void __launch_bounds__(kScanTile) k_tap_tile_write(int ntaps, const int* __restrict__ key,
                                                             const int* __restrict__ own_cnt,
                                                             const int* __restrict__ tsum, int* __restrict__ toff,
                                                             int* __restrict__ uniq, int* __restrict__ cnt,
                                                             int* __restrict__ off) {
  __shared__ int sa[kScanTile], sb[kScanTile];
  const int t = blockIdx.x * kScanTile + threadIdx.x;
  const int c = t < ntaps ? own_cnt[t] : 0;
  int r = c > 0, o = c;
  block_scan2(r, o, sa, sb);
  if (t >= ntaps) return;
  if (c > 0) {
    const int rr = tsum[2 * blockIdx.x] + r - 1, oo = tsum[2 * blockIdx.x + 1] + o - c;  // exclusive
    uniq[rr] = key[t];
    cnt[rr] = c;
    off[rr] = oo;
    toff[t] = oo;
  } else {
    toff[t] = -1;
  }
}

// 4. each matched tap at its owner's offset plus its rank
__global__ void k_tap_list(int ntaps, const int* __restrict__ rank, const int* __restrict__ otap,
                           const int* __restrict__ toff, int* __restrict__ taps) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntaps) return;
  const int o = otap[t];
  if (o < 0) return;
  taps[toff[o] + rank[t]] = t;
}

void RegCSR::alloc(int ntaps_, int max_key) {
  release();
  ntaps = ntaps_;
  (void)max_key;
  const int nf = ntaps / kTaps;
  const size_t b = sizeof(int) * (size_t)(ntaps > 0 ? ntaps : 1);
  IB_HIP(hipMalloc(&key, b));
  IB_HIP(hipMalloc(&own, b));
  IB_HIP(hipMalloc(&toff, b));
  IB_HIP(hipMalloc(&rank, b));
  IB_HIP(hipMalloc(&otap, b));
  IB_HIP(hipMalloc(&tsum, sizeof(int) * 2 * (size_t)std::max(1, ceil_div(ntaps, kScanTile))));
  IB_HIP(hipMalloc(&tap_sorted, b));
  IB_HIP(hipMalloc(&uniq, b));
  IB_HIP(hipMalloc(&cnt, b));
  IB_HIP(hipMalloc(&off, b));
  IB_HIP(hipMalloc(&nruns, sizeof(int)));
  IB_HIP(hipMalloc(&skey, sizeof(int) * (size_t)std::max(nf, 1)));
  IB_HIP(hipMalloc(&sent, sizeof(int) * (size_t)std::max(nf, 1)));
  n2 = 1;
  while (n2 < nf) n2 <<= 1;
  if (n2 > kSortLds) IB_HIP(hipMalloc(&sortv, sizeof(long long) * (size_t)n2));
}

void RegCSR::release() {
  int* ps[] = {key, own, toff, rank, otap, tsum, tap_sorted, uniq, cnt, off, nruns, skey, sent};
  for (int* p : ps)
    if (p) (void)hipFree(p);
  if (sortv) (void)hipFree(sortv);
  key = own = toff = rank = otap = tsum = tap_sorted = uniq = cnt = off = nruns = skey = sent = nullptr;
  sortv = nullptr;
}

void RegCSR::build(hipStream_t st, const Geom& g) {
  if (ntaps <= 0) return;
  const int nf = ntaps / kTaps;
  if (n2 <= kSortLds) {
    hipLaunchKernelGGL(k_sort_centres, dim3(1), dim3(1024), 0, st, nf, n2, key, skey, sent);
    IB_LAUNCH_CHECK();
  } else {
    hipLaunchKernelGGL(k_sort_init, dim3(ceil_div(n2, 256)), dim3(256), 0, st, nf, n2, key, sortv);
    IB_LAUNCH_CHECK();
    for (int size = 2; size <= n2; size <<= 1)
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        hipLaunchKernelGGL(k_sort_step, dim3(ceil_div(n2 / 2, 256)), dim3(256), 0, st, n2, size, stride, sortv);
        IB_LAUNCH_CHECK();
      }
    hipLaunchKernelGGL(k_sort_split, dim3(ceil_div(nf, 256)), dim3(256), 0, st, nf, sortv, skey, sent);
    IB_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_tap_owner, dim3(ceil_div(ntaps, 256)), dim3(256), 0, st, g, nf, key, skey, sent, own, rank,
                     otap);
  IB_LAUNCH_CHECK();
  const int ntiles = ceil_div(ntaps, kScanTile);
  hipLaunchKernelGGL(k_tap_tile_sum, dim3(ntiles), dim3(kScanTile), 0, st, ntaps, own, tsum);
  IB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_tap_tile_scan, dim3(1), dim3(kScanTile), 0, st, ntiles, tsum, nruns);
  IB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_tap_tile_write, dim3(ntiles), dim3(kScanTile), 0, st, ntaps, key, own, tsum, toff, uniq, cnt,
                     off);
  IB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_tap_list, dim3(ceil_div(ntaps, 256)), dim3(256), 0, st, ntaps, rank, otap, toff, tap_sorted);
  IB_LAUNCH_CHECK();
}

// reg: one thread per touched face, contributions summed in tap order
struct RegOp {
  using Args = RegArgs;
  static __device__ __forceinline__ void tile(const RegArgs& a, int bx, int, int) {
    const int r = bx * 256 + threadIdx.x;
    if (r >= a.ntaps || r >= *a.nruns) return;
    const int o = a.off[r], c = a.cnt[r], key = a.uniq[r];
    const int* __restrict__ taps = a.taps;
    const double* __restrict__ w = a.w;
    const double* __restrict__ f = a.f;
    double acc = 0.0;
    for (int q = o; q < o + c; ++q) {
      const int tp = taps[q];
      acc = acc + w[tp] * f[tp / kTaps];
    }
    if (key < a.g.Fx) a.hx[key] = acc;
    else a.hy[key - a.g.Fx] = acc;
  }
};

__global__ void __launch_bounds__(256) k_reg(RegArgs a) { RegOp::tile(a, blockIdx.x, 0, 0); }

TabEntry<RegArgs> reg_entry(const Geom& g, const RegCSR& c, const double* w, const double* f, double* hx, double* hy) {
  TabEntry<RegArgs> e{};
  e.a = RegArgs{g, c.ntaps, c.nruns, c.uniq, c.off, c.cnt, c.tap_sorted, w, f, hx, hy};
  e.nbx = ceil_div(std::max(c.ntaps, 1), 256);
  e.nby = 1;
  return e;
}

void reg(hipStream_t st, const Geom& g, const RegCSR& c, const double* w, const double* f, double* hx, double* hy) {
  if (c.ntaps <= 0) return;
  const TabEntry<RegArgs> e = reg_entry(g, c, w, f, hx, hy);
  hipLaunchKernelGGL(k_reg, dim3(e.nbx), dim3(256), 0, st, e.a);
  IB_LAUNCH_CHECK();
}

// regT: src/fsinterface.F90:51-80.  One wavefront per force entry: lane t
// forms w_t * u(face_t) for tap t, then the 49 products are added in tap
// order (the reference's sequential accumulation) through lane shuffles.
__device__ __forceinline__ double taps_in_order(double acc, double prod) {
  for (int t = 0; t < kTaps; ++t) acc = acc + __shfl(prod, t);
  return acc;
}

struct RegTOp {
  using Args = RegTArgs;
  static __device__ __forceinline__ void tile(const RegTArgs& a, int bx, int, int) {
    const int k = (bx * 256 + threadIdx.x) / warpSize;
    const int lane = threadIdx.x % warpSize;
    if (k >= a.nf) return;
    double prod = 0.0;
    if (lane < kTaps) {
      const int kk = a.key[k * kTaps + lane];
      const double v = kk < a.g.Fx ? a.hx[kk] + a.shift_x : a.hy[kk - a.g.Fx];
      prod = a.w[k * kTaps + lane] * v;
    }
    const double acc = taps_in_order(a.base ? a.base[k] : 0.0, prod);
    if (lane == 0) a.out[k] = acc;
  }
};

__global__ void __launch_bounds__(256) k_regT(RegTArgs a) { RegTOp::tile(a, blockIdx.x, 0, 0); }

TabEntry<RegTArgs> regT_entry(const Geom& g, int nf, const double* w, const int* key, const double* hx,
                              const double* hy, double shift_x, const double* base, double* out) {
  TabEntry<RegTArgs> e{};
  e.a = RegTArgs{g, nf, w, key, hx, hy, shift_x, base, out};
  e.nbx = ceil_div(std::max(nf, 1), 4);
  e.nby = 1;
  return e;
}

void regT(hipStream_t st, const Geom& g, int nf, const double* w, const int* key, const double* hx, const double* hy,
          double shift_x, const double* base, double* out) {
  if (nf <= 0) return;
  const TabEntry<RegTArgs> e = regT_entry(g, nf, w, key, hx, hy, shift_x, base, out);
  hipLaunchKernelGGL(k_regT, dim3(e.nbx), dim3(256), 0, st, e.a);
  IB_LAUNCH_CHECK();
}

// Columns of the Schur matrix: S(r, j) = E C s1_j (regT of the level-1 flux,
// curl taken at the tap faces) + b_times2(e_j)(r).  One wavefront per (row r,
// column j): tap-order sum as in k_regT, then y = y1 + y2 as in ab_times
// (src/timestepper.F90:780).
__device__ __forceinline__ double BRv(const TorsBody& b, int j, double si, double ci) {
  return j <= b.npts ? (j - 1) * b.deltas * si : -(j - b.npts - 1) * b.deltas * ci;
}

// b_times2(e_j)(r): the rank-one torsional block (src/operators_struct.F90:14-127)
__device__ __forceinline__ double btimes_entry(int r, int j, const int* __restrict__ tors_of,
                                               const TorsBody* __restrict__ tbs, const double* __restrict__ theta,
                                               double dt, double delta) {
  const int tr = tors_of[r];
  if (tr < 0 || tr != tors_of[j]) return 0.0;
  const TorsBody b = tbs[tr];
  const double si = sin(b.gamma0 - theta[tr]), ci = cos(b.gamma0 - theta[tr]);
  const double den = 4.0 / (dt * dt) * b.M + 2.0 / dt * b.C + b.K;
  const double sc = (delta * delta / b.deltas / dt) * (2 * b.deltas / dt);
  return BRv(b, r - b.off + 1, si, ci) * BRv(b, j - b.off + 1, si, ci) / den * sc;
}

__global__ void k_schur_cols(Geom g, int nf, int c0, const double* __restrict__ w, const int* __restrict__ key,
                             const double* __restrict__ s1, long s_stride, const double* __restrict__ bc1,
                             long bc_stride, const int* __restrict__ tors_of, const TorsBody* __restrict__ tbs,
                             const double* __restrict__ theta, double dt, double delta, double* __restrict__ S) {
  const int r = (blockIdx.x * blockDim.x + threadIdx.x) / warpSize;
  const int lane = threadIdx.x % warpSize;
  if (r >= nf) return;
  const int jl = blockIdx.y, j = c0 + jl;
  const double* s = s1 + (long)jl * s_stride;
  const double* bc = bc1 + (long)jl * bc_stride;
  double prod = 0.0;
  if (lane < kTaps) prod = w[r * kTaps + lane] * curl_at(g, s, bc, key[r * kTaps + lane]);
  double y1 = 0.0;
  for (int t = 0; t < kTaps; ++t) y1 = y1 + __shfl(prod, t);
  if (lane == 0) S[(long)j * nf + r] = 1.0 * y1 + btimes_entry(r, j, tors_of, tbs, theta, dt, delta);
}

void schur_columns(hipStream_t st, const Geom& g, int nf, int c0, int nc, const double* w, const int* key,
                   const double* s1, long s_stride, const double* bc1, long bc_stride, const int* tors_of,
                   const TorsBody* tb, const double* theta, double dt, double delta, double* S) {
  if (nf <= 0 || nc <= 0) return;
  hipLaunchKernelGGL(k_schur_cols, dim3(ceil_div(nf, 4), nc), dim3(256), 0, st, g, nf, c0, w, key, s1, s_stride, bc1,
                     bc_stride, tors_of, tb, theta, dt, delta, S);
  IB_LAUNCH_CHECK();
}

__global__ void k_green_cols(Geom g, int nF, int c0, const int* __restrict__ fkeys, const double* __restrict__ s1,
                             long s_stride, const double* __restrict__ bc1, long bc_stride, double* __restrict__ G) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nF) return;
  const int jl = blockIdx.y;
  G[(long)(c0 + jl) * nF + b] = curl_at(g, s1 + (long)jl * s_stride, bc1 + (long)jl * bc_stride, fkeys[b]);
}

void green_columns(hipStream_t st, const Geom& g, int nF, int c0, int nc, const int* fkeys, const double* s1,
                   long s_stride, const double* bc1, long bc_stride, double* G) {
  if (nF <= 0 || nc <= 0) return;
  hipLaunchKernelGGL(k_green_cols, dim3(ceil_div(nF, 256), nc), dim3(256), 0, st, g, nF, c0, fkeys, s1, s_stride,
                     bc1, bc_stride, G);
  IB_LAUNCH_CHECK();
}

__global__ void k_map_taps(int ntaps, const int* __restrict__ key, const int* __restrict__ fmap,
                           int* __restrict__ tapidx, int* __restrict__ flag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntaps) return;
  const int idx = fmap[key[t]];
  tapidx[t] = idx;
  if (idx < 0) flag[0] = 1;
}

void map_taps(hipStream_t st, int nf, const int* key, const int* fmap, int* tapidx, int* flag) {
  const int ntaps = nf * kTaps;
  if (ntaps <= 0) return;
  hipLaunchKernelGGL(k_map_taps, dim3(ceil_div(ntaps, 256)), dim3(256), 0, st, ntaps, key, fmap, tapidx, flag);
  IB_LAUNCH_CHECK();
}

// T(:, j) = sum_q w[j,q] G(:, tapidx[j,q]) for the columns j = jl[jq] (jl
// NULL: j = jq), jq < nj: one thread per face row b.  A tap outside the face
// set (tapidx -1, which k_map_taps flags) reads face 0: the FSI loop forms S
// before it has read that flag and forms it again after rebuilding the set.
// Work item u = tile * nj + jq covers rows 256 tile .. 256 tile + 255 of
// column j.  The launch has 8 per blocks; the k-th block that the dispatcher
// deals to XCD x (block 8 k + x) takes u = x per + k, so every XCD forms
// whole row tiles: a tile's rows of the G columns reach one L2 and serve
// every column j there.  The 49 loads of a row are issued in two batches
// before their sum, which runs in tap order.
constexpr int kGreenHalf = 25;
__global__ void __launch_bounds__(256) k_green_right(int nF, int nj, long total, long per,
                                                     const double* __restrict__ w,
                                                     const int* __restrict__ tapidx, const double* __restrict__ G,
                                                     const int* __restrict__ jl, double* __restrict__ T) {
  __shared__ long s_col[kTaps];
  __shared__ double s_w[kTaps];
  const long u = (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (u >= total) return;  // the whole block
  const int tile = (int)(u / nj), jq = (int)(u - (long)tile * nj);
  const int j = jl ? jl[jq] : jq;
  if (threadIdx.x < kTaps) {
    s_col[threadIdx.x] = (long)max(tapidx[j * kTaps + threadIdx.x], 0) * nF;
    s_w[threadIdx.x] = w[j * kTaps + threadIdx.x];
  }
  __syncthreads();
  const int b = tile * 256 + threadIdx.x;
  if (b >= nF) return;
  double gv[kGreenHalf];
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < kGreenHalf; ++q) gv[q] = G[s_col[q] + b];
#pragma unroll
  for (int q = 0; q < kGreenHalf; ++q) acc += s_w[q] * gv[q];
#pragma unroll
  for (int q = kGreenHalf; q < kTaps; ++q) gv[q - kGreenHalf] = G[s_col[q] + b];
#pragma unroll
  for (int q = kGreenHalf; q < kTaps; ++q) acc += s_w[q] * gv[q - kGreenHalf];
  T[(long)j * nF + b] = acc;
}

static void green_right(hipStream_t st, int nF, int nj, const double* w, const int* tapidx, const double* G,
                        const int* jl, double* T) {
  const long total = (long)ceil_div(nF, 256) * nj, per = (total + 7) / 8;
  hipLaunchKernelGGL(k_green_right, dim3((unsigned)(8 * per)), dim3(256), 0, st, nF, nj, total, per, w, tapidx, G,
                     jl, T);
  IB_LAUNCH_CHECK();
}

// S(i, j) = sum_p w[i,p] T(tapidx[i,p], j) + b_times2(e_j)(i): one thread per
// row i = il[x] (il NULL: i = x) for x < ni, column j = jl[blockIdx.y] (jl
// NULL: blockIdx.y)
__global__ void __launch_bounds__(256) k_green_left(int nf, int nF, int ni, const int* __restrict__ il,
                                                    const int* __restrict__ jl, const double* __restrict__ w,
                                                    const int* __restrict__ tapidx, const double* __restrict__ T,
                                                    const int* __restrict__ tors_of, const TorsBody* __restrict__ tbs,
                                                    const double* __restrict__ theta, double dt, double delta,
                                                    double* __restrict__ S) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= ni) return;
  const int i = il ? il[x] : x;
  const int j = jl ? jl[blockIdx.y] : blockIdx.y;
  const double* Tj = T + (long)j * nF;
  const int* tp = tapidx + (long)i * kTaps;
  const double* wi = w + (long)i * kTaps;
  // the row's taps in two batches of independent loads, each summed in tap
  // order
  double acc = 0.0;
  {
    int ti[kGreenHalf];
    double wv[kGreenHalf], tv[kGreenHalf];
#pragma unroll
    for (int p = 0; p < kGreenHalf; ++p) {
      ti[p] = tp[p];
      wv[p] = wi[p];
    }
#pragma unroll
    for (int p = 0; p < kGreenHalf; ++p) tv[p] = Tj[max(ti[p], 0)];
#pragma unroll
    for (int p = 0; p < kGreenHalf; ++p) acc += wv[p] * tv[p];
  }
  {
    constexpr int n2 = kTaps - kGreenHalf;
    int ti[n2];
    double wv[n2], tv[n2];
#pragma unroll
    for (int p = 0; p < n2; ++p) {
      ti[p] = tp[kGreenHalf + p];
      wv[p] = wi[kGreenHalf + p];
    }
#pragma unroll
    for (int p = 0; p < n2; ++p) tv[p] = Tj[max(ti[p], 0)];
#pragma unroll
    for (int p = 0; p < n2; ++p) acc += wv[p] * tv[p];
  }
  S[(long)j * nf + i] = 1.0 * acc + btimes_entry(i, j, tors_of, tbs, theta, dt, delta);
}

void schur_from_green(hipStream_t st, int nf, int nF, const double* w, const int* tapidx, const double* G,
                      double* T, const int* tors_of, const TorsBody* tb, const double* theta, double dt,
                      double delta, double* S) {
  if (nf <= 0 || nF <= 0) return;
  green_right(st, nF, nf, w, tapidx, G, nullptr, T);
  hipLaunchKernelGGL(k_green_left, dim3(ceil_div(nf, 256), nf), dim3(256), 0, st, nf, nF, nf, nullptr, nullptr, w,
                     tapidx, T, tors_of, tb, theta, dt, delta, S);
  IB_LAUNCH_CHECK();
}

void schur_from_green_moving(hipStream_t st, int nf, int nF, const double* w, const int* tapidx, const double* G,
                             double* T, const int* tors_of, const TorsBody* tb, const double* theta, double dt,
                             double delta, double* S, int nmov, const int* mov, int nrig, const int* rig) {
  if (nf <= 0 || nF <= 0 || nmov <= 0) return;
  // the moving columns of T
  green_right(st, nF, nmov, w, tapidx, G, mov, T);
  // every row of the moving columns of S
  hipLaunchKernelGGL(k_green_left, dim3(ceil_div(nf, 256), nmov), dim3(256), 0, st, nf, nF, nf, nullptr, mov, w,
                     tapidx, T, tors_of, tb, theta, dt, delta, S);
  IB_LAUNCH_CHECK();
  // the moving rows of the other columns
  if (nrig > 0) {
    hipLaunchKernelGGL(k_green_left, dim3(ceil_div(nmov, 256), nrig), dim3(256), 0, st, nf, nF, nmov, mov, rig, w,
                       tapidx, T, tors_of, tb, theta, dt, delta, S);
    IB_LAUNCH_CHECK();
  }
}

// redistribute2 on the touched faces: finter = 1/wght * frc where wght > 1e-10
struct RedistOp {
  using Args = RegArgs;
  static __device__ __forceinline__ void tile(const RegArgs& a, int bx, int, int) {
    const int r = bx * 256 + threadIdx.x;
    if (r >= a.ntaps || r >= *a.nruns) return;
    const int o = a.off[r], c = a.cnt[r], key = a.uniq[r];
    const int* __restrict__ taps = a.taps;
    const double* __restrict__ w = a.w;
    const double* __restrict__ f = a.f;
    double ws = 0.0, fs = 0.0;
    for (int q = o; q < o + c; ++q) {
      const int tp = taps[q];
      ws = ws + w[tp] * 1.0;
      fs = fs + w[tp] * f[tp / kTaps];
    }
    const double v = ws > 1e-10 ? 1.0 / ws * fs : 0.0;
    if (key < a.g.Fx) a.hx[key] = v;
    else a.hy[key - a.g.Fx] = v;
  }
};

__global__ void __launch_bounds__(256) k_redist(RegArgs a) { RedistOp::tile(a, blockIdx.x, 0, 0); }

TabEntry<RegArgs> redist_entry(const Geom& g, const RegCSR& c, const double* w, const double* f, double* fx_out,
                               double* fy_out) {
  return reg_entry(g, c, w, f, fx_out, fy_out);
}

void redist_faces(hipStream_t st, const Geom& g, const RegCSR& c, const double* w, const double* f, double* fx_out,
                  double* fy_out) {
  if (c.ntaps <= 0) return;
  const TabEntry<RegArgs> e = redist_entry(g, c, w, f, fx_out, fy_out);
  hipLaunchKernelGGL(k_redist, dim3(e.nbx), dim3(256), 0, st, e.a);
  IB_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------
// torsional bodies (family 2), src/operators_struct.F90
// ------------------------------------------------------------------------

__device__ __forceinline__ double BR(const TorsBody& b, int j, double si, double ci) {
  // BR_build (:319-338), j = 1..2*npts
  return j <= b.npts ? (j - 1) * b.deltas * si : -(j - b.npts - 1) * b.deltas * ci;
}

__device__ double block_sum(double v, double* sh) {
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

// b_times2 (:14-127) as the rank-1 product c * BR (BR^T x) / den
__global__ void __launch_bounds__(256) k_btimes(const TorsBody* __restrict__ tbs, const double* __restrict__ theta,
                                                double dt, double delta, const double* __restrict__ x,
                                                double* __restrict__ y) {
  __shared__ double sh[256];
  const TorsBody b = tbs[blockIdx.x];
  const double th = theta[blockIdx.x];
  const double si = sin(b.gamma0 - th), ci = cos(b.gamma0 - th);
  double part = 0.0;
  for (int j = threadIdx.x + 1; j <= 2 * b.npts; j += blockDim.x) part += BR(b, j, si, ci) * x[b.off + j - 1];
  const double dot = block_sum(part, sh);
  const double den = 4.0 / (dt * dt) * b.M + 2.0 / dt * b.C + b.K;
  const double sc = (delta * delta / b.deltas / dt) * (2 * b.deltas / dt);
  for (int j = threadIdx.x + 1; j <= 2 * b.npts; j += blockDim.x)
    y[b.off + j - 1] = y[b.off + j - 1] + BR(b, j, si, ci) * dot / den * sc;
}

void body_btimes(hipStream_t st, int nt, const TorsBody* tb, const double* theta, double dt, double delta,
                 const double* x, double* y) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_btimes, dim3(nt), dim3(256), 0, st, tb, theta, dt, delta, x, y);
  IB_LAUNCH_CHECK();
}

// compute_rhs (:342-398) for torsional entries + regT of q + q0 (rhsf of the
// FSI loop, src/timestepper.F90:322-358); one wavefront per force entry
struct FsiRhsOp {
  using Args = FsiRhsArgs;
  static __device__ __forceinline__ void tile(const FsiRhsArgs& a, int bx, int by, int);
};

__device__ __forceinline__ void FsiRhsOp::tile(const FsiRhsArgs& a, int bx, int, int) {
  const Geom& g = a.g;
  const int nf = a.nf, nt = a.nt;
  const int* __restrict__ tors_of = a.tors_of;
  const TorsBody* __restrict__ tbs = a.tbs;
  const double* __restrict__ th = a.th;
  const double dt = a.dt, delta = a.delta;
  const double* __restrict__ w = a.w;
  const int* __restrict__ key = a.key;
  const double* __restrict__ qx = a.qx;
  const double* __restrict__ qy = a.qy;
  const double q0x = a.q0x;
  double* __restrict__ rhsf = a.rhsf;
  const int k = (bx * 256 + threadIdx.x) / warpSize;
  const int lane = threadIdx.x % warpSize;
  if (k >= nf) return;
  double v = 0.0;
  const int t = tors_of[k];
  if (t >= 0) {
    const TorsBody b = tbs[t];
    const double theta = th[t], thetad = th[nt + t];
    const double theta0 = th[3 * nt + t], thetad0 = th[4 * nt + t], thetadd0 = th[5 * nt + t];
    const double rth = thetad0 + thetad + 2.0 / dt * (theta0 - theta);
    const double den = 4.0 / (dt * dt) * b.M + 2.0 / dt * b.C + b.K;
    const double jr = 2.0 / dt / den *
                      (b.M * (4.0 / (dt * dt) * theta0 + 4.0 / dt * thetad0 + thetadd0) +
                       b.C * (thetad0 + 2.0 / dt * theta0) - (4.0 / (dt * dt) * b.M + 2.0 / dt * b.C + b.K) * theta +
                       0.0 + 0.0);
    const double sc = delta * (rth - jr - thetad);
    const double si = sin(b.gamma0 - theta), ci = cos(b.gamma0 - theta);
    v = BR(b, k - b.off + 1, si, ci) * sc;
  }
  double prod = 0.0;
  if (key && lane < kTaps) {  // no key list: the structural part only
    const int kk = key[k * kTaps + lane];
    const double u = kk < g.Fx ? qx[kk] + q0x : qy[kk - g.Fx] + 0.0;
    prod = w[k * kTaps + lane] * u;
  }
  v = taps_in_order(v, prod);
  if (lane == 0) rhsf[k] = v;
}

__global__ void __launch_bounds__(256) k_fsi_rhs(FsiRhsArgs a) { FsiRhsOp::tile(a, blockIdx.x, 0, 0); }

TabEntry<FsiRhsArgs> fsi_rhs_entry(const Geom& g, int nf, const int* tors_of_entry, const TorsBody* tb, int nt,
                                   const double* th, double dt, double delta, const double* w, const int* key,
                                   const double* qx, const double* qy, double q0x, double* rhsf) {
  TabEntry<FsiRhsArgs> e{};
  e.a = FsiRhsArgs{g, nf, tors_of_entry, tb, nt, th, dt, delta, w, key, qx, qy, q0x, rhsf};
  e.nbx = ceil_div(std::max(nf, 1), 4);
  e.nby = 1;
  return e;
}

void fsi_rhs(hipStream_t st, const Geom& g, int nf, const int* tors_of_entry, const TorsBody* tb, int nt,
             const double* th, double dt, double delta, const double* w, const int* key, const double* qx,
             const double* qy, double q0x, double* rhsf) {
  if (nf <= 0) return;
  const TabEntry<FsiRhsArgs> e =
      fsi_rhs_entry(g, nf, tors_of_entry, tb, nt, th, dt, delta, w, key, qx, qy, q0x, rhsf);
  hipLaunchKernelGGL(k_fsi_rhs, dim3(e.nbx), dim3(256), 0, st, e.a);
  IB_LAUNCH_CHECK();
}

const void* body_tab_kernel(int kind) {
  switch (kind) {
    case TK_REG: return (const void*)k_tab<RegOp>;
    case TK_REDIST: return (const void*)k_tab<RedistOp>;
    case TK_REGT: return (const void*)k_tab<RegTOp>;
    case TK_FSI_RHS: return (const void*)k_tab<FsiRhsOp>;
    default: throw HipError("no body table kernel of this kind");
  }
}

// compute_dx (:402-505) + update_xb (:600-620); one workgroup per body
__global__ void __launch_bounds__(256) k_body_dx(const TorsBody* __restrict__ tbs, double* __restrict__ th, double dt,
                                                 double delta, const double* __restrict__ f, double* __restrict__ xb,
                                                 double* __restrict__ yb, double* __restrict__ dabs) {
  __shared__ double sh[256];
  __shared__ double s_theta;
  const int nt = gridDim.x, t = blockIdx.x;
  const TorsBody b = tbs[t];
  const double theta = th[t];
  const double theta0 = th[3 * nt + t], thetad0 = th[4 * nt + t], thetadd0 = th[5 * nt + t];
  const double si = sin(b.gamma0 - theta), ci = cos(b.gamma0 - theta);
  double part = 0.0;
  for (int j = threadIdx.x + 1; j <= 2 * b.npts; j += blockDim.x) part += BR(b, j, si, ci) * f[b.off + j - 1];
  double brf = block_sum(part, sh);
  if (threadIdx.x == 0) {
    const double jr = 1.0 / (4.0 / (dt * dt) * b.M + 2.0 / dt * b.C + b.K) *
                      (b.M * (4 / (dt * dt) * theta0 + 4 / dt * thetad0 + thetadd0) +
                       b.C * (thetad0 + 2.0 / dt * theta0) - (4 / (dt * dt) * b.M + 2.0 / dt * b.C + b.K) * theta +
                       0.0 + 0.0);
    brf = b.deltas / (4 / (dt * dt) * b.M + b.K) * brf * (delta / b.deltas / dt);
    const double dth = jr + brf;
    const double thn = theta + dth;
    th[t] = thn;
    th[nt + t] = -thetad0 + 2 / dt * (thn - theta0);
    th[2 * nt + t] = -thetadd0 - 4 / dt * thetad0 + 4 / (dt * dt) * (thn - theta0);
    dabs[t] = fabs(dth);
    s_theta = thn;
  }
  __syncthreads();
  const double c = cos(b.gamma0 - s_theta), s = sin(b.gamma0 - s_theta);
  for (int j = threadIdx.x + 1; j <= 2 * b.npts; j += blockDim.x) {
    const int jj = j <= b.npts ? j : j - b.npts;
    xb[b.off + j - 1] = b.xbf0 + (jj - 1) * b.deltas * c;
    yb[b.off + j - 1] = b.ybf0 + (jj - 1) * b.deltas * s;
  }
}

void body_dx(hipStream_t st, int nt, const TorsBody* tb, double* th, double dt, double delta, const double* f,
             double* xb, double* yb, double* dtheta_abs) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_body_dx, dim3(nt), dim3(256), 0, st, tb, th, dt, delta, f, xb, yb, dtheta_abs);
  IB_LAUNCH_CHECK();
}

// initial_acc (:233-294) with Ftint = ktheta*theta (var_update, :207-229)
__global__ void __launch_bounds__(256) k_initial_acc(const TorsBody* __restrict__ tbs, double* __restrict__ th,
                                                     double dt, double delta, const double* __restrict__ f) {
  __shared__ double sh[256];
  const int nt = gridDim.x, t = blockIdx.x;
  const TorsBody b = tbs[t];
  const double theta = th[t], thetad = th[nt + t];
  const double si = sin(b.gamma0 - theta), ci = cos(b.gamma0 - theta);
  double part = 0.0;
  for (int j = threadIdx.x + 1; j <= 2 * b.npts; j += blockDim.x) part += BR(b, j, si, ci) * f[b.off + j - 1];
  double brf = block_sum(part, sh);
  if (threadIdx.x == 0) {
    brf = brf * delta / b.deltas / dt;
    const double ftint = b.K * theta;
    th[2 * nt + t] = 1 / b.M * (-b.C * thetad - ftint + brf * b.deltas + 0.0 + 0.0);
  }
}

void body_initial_acc(hipStream_t st, int nt, const TorsBody* tb, double* th, double dt, double delta,
                      const double* f) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_initial_acc, dim3(nt), dim3(256), 0, st, tb, th, dt, delta, f);
  IB_LAUNCH_CHECK();
}

}  // namespace ibg
