// ib_fft.hip -- unnormalised 2-D DST-I (FFTW RODFT00 x RODFT00) on CDNA4.
//
// Replaces the FFTW-MPI transforms of src/myfft.F90:168-213.  A 2-D DST is a
// row (x) pass and a column (y) pass.  Each pass evaluates two real DST-I of
// length N at once with one complex FFT of length L = 2(N+1): the pair
// (x_a, x_b) is packed as z = x_a + i x_b and extended oddly
// (z_0 = z_M = 0, z_{2M-t} = -z_t), then
//     Z = -2i S(x_a) + 2 S(x_b)   =>   S(x_a) = -Im Z,  S(x_b) = Re Z.
// The FFT is a mixed-radix Stockham autosort run entirely in LDS with two
// ping-pong buffers per transform: radix 4/2/3/5 butterflies live in
// registers, any larger prime factor uses one thread per output.  Fused
// pre/post operations keep the ctci/ainv/predictor spectral products inside
// the column pass (Y_SCALE2, Y_PRED) and apply_bc / the corrector subtraction
// inside the row pass (X_APPLYBC, X_SUBFROM).
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "ib_device.h"
#include "ib_kernels.h"

namespace ibg {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

template <int R>
__device__ __forceinline__ void dft(double2* v);

template <>
__device__ __forceinline__ void dft<2>(double2* v) {
  double2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft<3>(double2* v) {
  const double s3 = 0.86602540378443864676;  // sin(2 pi / 3)
  double2 t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
  double2 mm = make_double2(v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y);
  double2 jd = make_double2(s3 * d.y, -s3 * d.x);  // -i s3 d
  v[0] = cadd(v[0], t);
  v[1] = cadd(mm, jd);
  v[2] = csub(mm, jd);
}

template <>
__device__ __forceinline__ void dft<4>(double2* v) {
  double2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
  double2 c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
  v[0] = cadd(a, c);
  v[2] = csub(a, c);
  v[1] = make_double2(b.x + d.y, b.y - d.x);  // b - i d
  v[3] = make_double2(b.x - d.y, b.y + d.x);  // b + i d
}

template <>
__device__ __forceinline__ void dft<5>(double2* v) {
  const double c1 = 0.30901699437494742410;   // cos(2 pi/5)
  const double c2 = -0.80901699437494742410;  // cos(4 pi/5)
  const double s1 = 0.95105651629515357212;   // sin(2 pi/5)
  const double s2 = 0.58778525229247312917;   // sin(4 pi/5)
  double2 p1 = cadd(v[1], v[4]), m1 = csub(v[1], v[4]);
  double2 p2 = cadd(v[2], v[3]), m2 = csub(v[2], v[3]);
  double2 a1 = make_double2(v[0].x + c1 * p1.x + c2 * p2.x, v[0].y + c1 * p1.y + c2 * p2.y);
  double2 a2 = make_double2(v[0].x + c2 * p1.x + c1 * p2.x, v[0].y + c2 * p1.y + c1 * p2.y);
  double2 b1 = make_double2(s1 * m1.x + s2 * m2.x, s1 * m1.y + s2 * m2.y);
  double2 b2 = make_double2(s2 * m1.x - s1 * m2.x, s2 * m1.y - s1 * m2.y);
  v[0] = make_double2(v[0].x + p1.x + p2.x, v[0].y + p1.y + p2.y);
  v[1] = make_double2(a1.x + b1.y, a1.y - b1.x);  // a1 - i b1
  v[4] = make_double2(a1.x - b1.y, a1.y + b1.x);  // a1 + i b1
  v[2] = make_double2(a2.x + b2.y, a2.y - b2.x);  // a2 - i b2
  v[3] = make_double2(a2.x - b2.y, a2.y + b2.x);  // a2 + i b2
}

template <>
__device__ __forceinline__ void dft<8>(double2* v) {
  const double c = 0.70710678118654752440;  // cos(pi/4)
  double2 e[4] = {v[0], v[2], v[4], v[6]};
  double2 o[4] = {v[1], v[3], v[5], v[7]};
  dft<4>(e);
  dft<4>(o);
  // o[k] *= exp(-2 pi i k / 8)
  const double2 o1 = make_double2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));
  const double2 o2 = make_double2(o[2].y, -o[2].x);
  const double2 o3 = make_double2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// One Stockham stage of radix R over nf transforms of length L stored at
// in[s*L + t]; Ns = product of the radices already applied.
template <int R>
__device__ __forceinline__ void stage_fixed(const double2* __restrict__ in, double2* __restrict__ out, int nf,
                                            int L, int Ns, const double2* __restrict__ W) {
  const int LR = L / R;
  const int tws = LR / Ns;
  const int total = nf * LR;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int s = idx / LR;
    const int j = idx - s * LR;
    const double2* ip = in + s * L;
    double2* op = out + s * L;
    const int k = j % Ns;
    double2 v[R];
    v[0] = ip[j];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(ip[j + r * LR], W[k * r * tws]);
    dft<R>(v);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) op[base + r * Ns] = v[r];
  }
}

// A stage of any other radix R (an odd prime >= 7) as a direct length-R DFT
// using its conjugate symmetry.  Pass 1 applies the inter-stage twiddles and
// folds each input pair in place: p_r = x_r + x_(R-r) at position r and
// m_r = x_r - x_(R-r) at position R-r (r = 1..H, H = (R-1)/2).  Pass 2 forms
// U_q = sum p_r cos(2 pi rq/R) and V_q = sum m_r sin(2 pi rq/R), which give
// both X_q = x_0 + U_q - i V_q and X_(R-q) = x_0 + U_q + i V_q: a quarter of
// the real products of the plain sum.  Each thread holds a tile of 2
// sub-transforms x 4 values of q, so every LDS read feeds several products.
// lanes one transform's tiles of the generic stage may spread over (the
// threads of a one-transform workgroup)
constexpr int kGenSplitLanes = 256;

__device__ void stage_generic(double2* __restrict__ in, double2* __restrict__ out, int nf, int L, int Ns, int R,
                              const double2* __restrict__ W) {
  const int LR = L / R;
  const int tws = LR / Ns;
  const int H = (R - 1) / 2;
  for (int idx = threadIdx.x; idx < nf * LR * H; idx += blockDim.x) {
    const int s = idx / (LR * H);
    const int u = idx - s * (LR * H);
    const int r = 1 + u / LR, j = u % LR;
    double2* x = in + s * L + j;
    double2 xa = x[r * LR], xb = x[(R - r) * LR];
    const int k = j % Ns;
    if (k) {
      xa = cmul(xa, W[k * r * tws]);
      xb = cmul(xb, W[k * (R - r) * tws]);
    }
    x[r * LR] = cadd(xa, xb);
    x[(R - r) * LR] = csub(xa, xb);
  }
  __syncthreads();
  const int nj = (LR + 1) / 2, nq = (H + 4) / 4;  // q = 0..H in groups of 4
  const int ntile = nf * nj * nq;
  // with few tiles per transform, 2 or 4 adjacent lanes share a tile's sum
  // over r (terms r = 1 + part, 1 + part + split, ...) and combine it by
  // lane shuffles.  The split depends on the transform's own tile count
  // only (not on how many transforms share the workgroup, nor on its size),
  // so every tile shape of the DST passes sums in the same order and gives
  // the same bits.
  int lsplit = 0;
  while (lsplit < 2 && ((nj * nq) << (lsplit + 1)) <= kGenSplitLanes) ++lsplit;
  const int split = 1 << lsplit;
  for (int tt = threadIdx.x; tt < (ntile << lsplit); tt += blockDim.x) {
    const int t = tt >> lsplit, part = tt & (split - 1);
    const int s = t / (nj * nq);
    const int u = t - s * (nj * nq);
    const int qt = u / nj;
    const int j0 = 2 * (u - qt * nj), q0 = 4 * qt;
    const int j1 = min(j0 + 1, LR - 1);  // an odd LR repeats the last column (not stored)
    const double2* __restrict__ x = in + s * L;
    // cos/sin of 2 pi rq/R from W[((r q) mod R) LR] = (cos, -sin), the index
    // kept running below L per value of q
    int bq[4], dq[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int qL = min(q0 + c, H) * LR;  // below L
      bq[c] = (int)(((long)(1 + part) * qL) % L);
      dq[c] = (int)(((long)split * qL) % L);
    }
    double2 U0[4], V0[4], U1[4], V1[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) U0[c] = V0[c] = U1[c] = V1[c] = make_double2(0.0, 0.0);
    for (int r = 1 + part; r <= H; r += split) {
      const double2 pa = x[j0 + r * LR], ma = x[j0 + (R - r) * LR];
      const double2 pb = x[j1 + r * LR], mb = x[j1 + (R - r) * LR];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double2 w = W[bq[c]];
        const double cs = w.x, sn = -w.y;
        U0[c] = make_double2(U0[c].x + pa.x * cs, U0[c].y + pa.y * cs);
        V0[c] = make_double2(V0[c].x + ma.x * sn, V0[c].y + ma.y * sn);
        U1[c] = make_double2(U1[c].x + pb.x * cs, U1[c].y + pb.y * cs);
        V1[c] = make_double2(V1[c].x + mb.x * sn, V1[c].y + mb.y * sn);
        bq[c] += dq[c];
        if (bq[c] >= L) bq[c] -= L;
      }
    }
    for (int off = 1; off < split; off <<= 1) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        U0[c].x += __shfl_xor(U0[c].x, off); U0[c].y += __shfl_xor(U0[c].y, off);
        V0[c].x += __shfl_xor(V0[c].x, off); V0[c].y += __shfl_xor(V0[c].y, off);
        U1[c].x += __shfl_xor(U1[c].x, off); U1[c].y += __shfl_xor(U1[c].y, off);
        V1[c].x += __shfl_xor(V1[c].x, off); V1[c].y += __shfl_xor(V1[c].y, off);
      }
    }
    if (part) continue;
    const double2 x0a = x[j0], x0b = x[j1];
    double2* __restrict__ o = out + s * L;
    const int k0 = j0 % Ns, k1 = (j0 + 1) % Ns;
    const bool second = j0 + 1 < LR;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int q = q0 + c;
      if (q <= H) {
        const int ba = (j0 - k0) * R + k0, bb = (j0 + 1 - k1) * R + k1;
        o[ba + q * Ns] = make_double2(x0a.x + U0[c].x + V0[c].y, x0a.y + U0[c].y - V0[c].x);
        if (second) o[bb + q * Ns] = make_double2(x0b.x + U1[c].x + V1[c].y, x0b.y + U1[c].y - V1[c].x);
        if (q > 0) {
          o[ba + (R - q) * Ns] = make_double2(x0a.x + U0[c].x - V0[c].y, x0a.y + U0[c].y + V0[c].x);
          if (second) o[bb + (R - q) * Ns] = make_double2(x0b.x + U1[c].x - V1[c].y, x0b.y + U1[c].y + V1[c].x);
        }
      }
    }
  }
}

// Runs the plan on nf transforms held in a (scratch b) with the twiddle table
// W staged in LDS; returns the buffer that holds the result.  Caller
// synchronises before the call.  GEN = false leaves out the stage of radix
// >= 7 (stage_generic, whose registers set the kernel's occupancy): the host
// launches that form only for plans whose radices are all 2, 3, 4, 5, 8.
// Only the row pass has it: the column pass built without the stage
// measured slower (its two-input form 23.4 against 16.7 us per launch at
// 400 x 400, the same LDS-bound occupancy), so it always keeps the stage.
template <bool GEN>
__device__ __forceinline__ double2* fft_run(double2* a, double2* b, int nf, const FFTPlan& pl, const double2* W) {
  int Ns = 1;
  for (int st = 0; st < pl.nst; ++st) {
    const int R = pl.rad[st];
    switch (R) {
      case 8: stage_fixed<8>(a, b, nf, pl.L, Ns, W); break;
      case 4: stage_fixed<4>(a, b, nf, pl.L, Ns, W); break;
      case 2: stage_fixed<2>(a, b, nf, pl.L, Ns, W); break;
      case 3: stage_fixed<3>(a, b, nf, pl.L, Ns, W); break;
      case 5: stage_fixed<5>(a, b, nf, pl.L, Ns, W); break;
      default:
        if constexpr (GEN) {
          stage_generic(a, b, nf, pl.L, Ns, R, W);
        } else {
          // a plan this form cannot run (the host never launches one): the
          // result is NaN, never a stale buffer
          for (int i = threadIdx.x; i < nf * pl.L; i += blockDim.x) b[i] = make_double2(__builtin_nan(""), __builtin_nan(""));
        }
        break;
    }
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}


// X_POINTROT: rot(E^T e_k) at node (i, j), formed from the 49 taps of force
// entry k (the x-face or y-face patch centred on (ix, iy)); with no weights,
// rot of the unit face source at (ix, iy).
struct PointPatch {
  int comp, ix, iy;
  const double* w;
  __device__ double face(int c, int i, int j) const {
    if (c != comp) return 0.0;
    const int p = i - ix, l = j - iy;
    if (!w) return (p == 0 && l == 0) ? 1.0 : 0.0;
    if (p < -3 || p > 3 || l < -3 || l > 3) return 0.0;
    return w[(l + 3) * 7 + (p + 3)];
  }
  __device__ double rot(int i, int j) const {
    return face(2, i, j) - face(2, i - 1, j) - face(1, i, j) + face(1, i, j - 1);
  }
};


// Stage the twiddle table in LDS (read by every stage of every transform).
__device__ __forceinline__ void load_twiddles(double2* Wl, const FFTPlan& pl) {
  for (int t = threadIdx.x; t < pl.L; t += blockDim.x) Wl[t] = pl.W[t];
}

// Row pass.  Rows j0 = 0..n-2 (node j = j0+2) of length N = m-1.
template <int MODE, bool GEN>
__device__ __forceinline__ void dst_x_tile(const XArgs& a, int bx, int by) {
  extern __shared__ double2 sm[];
  const FFTPlan& pl = a.pl;
  const int L = pl.L, M = pl.M, N = pl.N, rp = a.rp;
  const int nrows = a.g.n - 1;
  double2* A = sm;
  double2* B = sm + rp * L;
  double2* Wl = sm + 2 * rp * L;
  const int bsel = a.strided ? 0 : by;
  const double* __restrict__ in = a.in[bsel] ? a.in[bsel] + (long)by * a.s_in : nullptr;
  double* out = a.out[bsel] + (long)by * a.s_out;
  const double* bcb = a.bc ? a.bc + (long)by * a.s_bc : nullptr;
  const int pair0 = bx * rp;
  PointPatch pt{};
  if (MODE == X_POINTROT) {
    const int k = a.c0 + by;
    // centre tap (p = l = 0) of force entry k, or face k of a face list
    const int ck = a.w ? a.key[k * kTaps + kTaps / 2] : a.key[k];
    const Geom& g = a.g;
    if (ck < g.Fx) { pt.comp = 1; pt.ix = ck % (g.m + 1) + 1; pt.iy = ck / (g.m + 1) + 1; }
    else { pt.comp = 2; pt.ix = (ck - g.Fx) % g.m + 1; pt.iy = (ck - g.Fx) / g.m + 1; }
    pt.w = a.w ? a.w + (long)k * kTaps : nullptr;
    // rows whose nodes see no tap transform to zero: write zeros and leave
    const int j0 = 2 * pair0 + 2, j1 = 2 * (pair0 + rp) + 1;  // node rows of this workgroup
    if (j1 < pt.iy - 3 || j0 > pt.iy + 4) {
      for (int idx = threadIdx.x; idx < rp * N; idx += blockDim.x) {
        const int s = idx / N, t = idx - s * N;
        const int ra = 2 * (pair0 + s), rb = ra + 1;
        if (ra < nrows) out[(long)ra * N + t] = 0.0;
        if (rb < nrows) out[(long)rb * N + t] = 0.0;
      }
      return;
    }
  }
  load_twiddles(Wl, pl);
  for (int idx = threadIdx.x; idx < rp * N; idx += blockDim.x) {
    const int s = idx / N;
    const int t = idx - s * N;
    const int ra = 2 * (pair0 + s), rb = ra + 1;
    double xa = 0.0, xb = 0.0;
    if (MODE == X_POINTROT) {
      const int i = t + 2;
      if (ra < nrows) xa = pt.rot(i, ra + 2);
      if (rb < nrows) xb = pt.rot(i, rb + 2);
    } else if (MODE == X_ROT) {
      // rot_int + rot_bc: vort(i,j) = hy(i,j) - hy(i-1,j) - hx(i,j) + hx(i,j-1)
      const Geom& g = a.g;
      const int i = t + 2;
      if (ra < nrows) {
        const int j = ra + 2;
        xa = a.hy[(i - 1) + g.m * (j - 1)] - a.hy[(i - 2) + g.m * (j - 1)] - a.hx[(i - 1) + (g.m + 1) * (j - 1)] +
             a.hx[(i - 1) + (g.m + 1) * (j - 2)];
      }
      if (rb < nrows) {
        const int j = rb + 2;
        xb = a.hy[(i - 1) + g.m * (j - 1)] - a.hy[(i - 2) + g.m * (j - 1)] - a.hx[(i - 1) + (g.m + 1) * (j - 1)] +
             a.hx[(i - 1) + (g.m + 1) * (j - 2)];
      }
    } else {
      if (ra < nrows) xa = in[(long)ra * N + t];
      if (rb < nrows) xb = in[(long)rb * N + t];
    }
    if (MODE == X_APPLYBC) {
      // apply_bc(vort, sbc, 1.d0), src/operators_fluid.F90:539-549
      const int i = t + 2;
      const Geom& g = a.g;
      const double* bc = bcb;
      if (ra < nrows) {
        const int j = ra + 2;
        if (i == 2) xa = xa + bc[g.left + j - 1];
        if (i == g.m) xa = xa + bc[g.right + j - 1];
        if (j == 2) xa = xa + bc[g.bottom + i - 1];
        if (j == g.n) xa = xa + bc[g.top + i - 1];
      }
      if (rb < nrows) {
        const int j = rb + 2;
        if (i == 2) xb = xb + bc[g.left + j - 1];
        if (i == g.m) xb = xb + bc[g.right + j - 1];
        if (j == 2) xb = xb + bc[g.bottom + i - 1];
        if (j == g.n) xb = xb + bc[g.top + i - 1];
      }
    }
    A[s * L + t + 1] = make_double2(xa, xb);
    A[s * L + L - (t + 1)] = make_double2(-xa, -xb);
  }
  for (int s = threadIdx.x; s < rp; s += blockDim.x) {
    A[s * L] = make_double2(0.0, 0.0);
    A[s * L + M] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  const double2* Z = fft_run<GEN>(A, B, rp, pl, Wl);
  for (int idx = threadIdx.x; idx < rp * N; idx += blockDim.x) {
    const int s = idx / N;
    const int t = idx - s * N;
    const int ra = 2 * (pair0 + s), rb = ra + 1;
    const double2 z = Z[s * L + t + 1];
    if (MODE == X_SUBFROM) {
      if (ra < nrows) out[(long)ra * N + t] = out[(long)ra * N + t] - (-z.y);
      if (rb < nrows) out[(long)rb * N + t] = out[(long)rb * N + t] - z.x;
    } else {
      if (ra < nrows) out[(long)ra * N + t] = -z.y;
      if (rb < nrows) out[(long)rb * N + t] = z.x;
    }
  }
}

// Column pass.  Columns i0 = 0..m-2 (stride m-1) of length N = n-1.
template <int MODE>
__device__ __forceinline__ void dst_y_tile(const YArgs& a, int bx, int by) {
  extern __shared__ double2 sm[];
  const FFTPlan& pl = a.pl;
  const int L = pl.L, M = pl.M, N = pl.N, cp = a.cp;
  const int ncols = a.g.m - 1;
  const int nfft = (MODE == Y_PRED) ? 2 * cp : cp;
  double2* A = sm;
  double2* B = sm + nfft * L;
  double2* Wl = sm + 2 * nfft * L;
  const int bsel = a.strided ? 0 : by;
  const double* __restrict__ in = a.in[bsel] + (long)by * a.s_in;
  const double* __restrict__ in2 = a.in2[bsel] ? a.in2[bsel] + (long)by * a.s_in2 : nullptr;
  double* __restrict__ out = a.out[bsel] + (long)by * a.s_out;
  // workgroups are dealt to the 8 XCDs round-robin: give each XCD a
  // contiguous band of column tiles so neighbouring tiles share its L2
  const int tile = (bx % 8) * a.chunk + bx / 8;
  const int col0 = 2 * tile * cp;
  if (col0 >= ncols) return;
  const int w2 = 2 * cp;  // columns handled by this workgroup
  load_twiddles(Wl, pl);
  for (int idx = threadIdx.x; idx < N * w2; idx += blockDim.x) {
    const int t = idx / w2;
    const int c = idx - t * w2;
    const int col = col0 + c;
    const int s = c >> 1;
    double v = 0.0, v2 = 0.0;
    if (col < ncols) {
      v = in[(long)t * ncols + col];
      if (MODE == Y_PRED) v2 = in2[(long)t * ncols + col];
    }
    double* pa = reinterpret_cast<double*>(&A[s * L + t + 1]);
    double* pb = reinterpret_cast<double*>(&A[s * L + L - (t + 1)]);
    pa[c & 1] = v;
    pb[c & 1] = -v;
    if (MODE == Y_PRED) {
      double* qa = reinterpret_cast<double*>(&A[(s + cp) * L + t + 1]);
      double* qb = reinterpret_cast<double*>(&A[(s + cp) * L + L - (t + 1)]);
      qa[c & 1] = v2;
      qb[c & 1] = -v2;
    }
  }
  for (int s = threadIdx.x; s < nfft; s += blockDim.x) {
    A[s * L] = make_double2(0.0, 0.0);
    A[s * L + M] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  double2* Z = fft_run<true>(A, B, nfft, pl, Wl);
  if (MODE == Y_SCALE2 || MODE == Y_PRED) {
    double2* O = (Z == A) ? B : A;
    (void)O;
    for (int idx = threadIdx.x; idx < cp * N; idx += blockDim.x) {
      const int s = idx / N;
      const int t = idx - s * N;  // frequency index t+1
      const int ca = col0 + 2 * s, cb = ca + 1;
      const double2 za = Z[s * L + t + 1];
      double ya = -za.y, yb = za.x;
      if (MODE == Y_SCALE2) {
        ya = (ca < ncols) ? a.tab[(long)t * ncols + ca] * ya / a.div : 0.0;
        yb = (cb < ncols) ? a.tab[(long)t * ncols + cb] * yb / a.div : 0.0;
      } else {
        const double2 zb = Z[(s + cp) * L + t + 1];
        const double ba = -zb.y, bb = zb.x;
        // lam1i * (dst(rhs terms) + lam1 * dst(omega)), src/operators_fluid.F90:657-661
        ya = (ca < ncols) ? a.tab2[(long)t * ncols + ca] * (ya + a.tab[(long)t * ncols + ca] * ba) : 0.0;
        yb = (cb < ncols) ? a.tab2[(long)t * ncols + cb] * (yb + a.tab[(long)t * ncols + cb] * bb) : 0.0;
      }
      Z[s * L + t + 1] = make_double2(ya, yb);
      Z[s * L + L - (t + 1)] = make_double2(-ya, -yb);
    }
    for (int s = threadIdx.x; s < cp; s += blockDim.x) {
      Z[s * L] = make_double2(0.0, 0.0);
      Z[s * L + M] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    double2* other = (Z == A) ? B : A;
    Z = fft_run<true>(Z, other, cp, pl, Wl);
  }
  for (int idx = threadIdx.x; idx < N * w2; idx += blockDim.x) {
    const int t = idx / w2;
    const int c = idx - t * w2;
    const int col = col0 + c;
    const int s = c >> 1;
    if (col < ncols) {
      const double2 z = Z[s * L + t + 1];
      out[(long)t * ncols + col] = (c & 1) ? z.x : -z.y;
    }
  }
}

template <int MODE, bool GEN>
__global__ void __launch_bounds__(256) k_dst_x(XArgs a) {
  dst_x_tile<MODE, GEN>(a, blockIdx.x, blockIdx.y);
}
template <int MODE>
__global__ void __launch_bounds__(256) k_dst_y(YArgs a) {
  dst_y_tile<MODE>(a, blockIdx.x, blockIdx.y);
}
template <int MODE, bool GEN>
struct DstXOp {
  using Args = XArgs;
  static __device__ __forceinline__ void tile(const XArgs& a, int bx, int by, int) { dst_x_tile<MODE, GEN>(a, bx, by); }
};
template <int MODE>
struct DstYOp {
  using Args = YArgs;
  static __device__ __forceinline__ void tile(const YArgs& a, int bx, int by, int) { dst_y_tile<MODE>(a, bx, by); }
};

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------

static std::vector<int> factorize(int L) {
  std::vector<int> r;
  const int small[] = {8, 4, 2, 3, 5};
  for (int q : small)
    while (L % q == 0) {
      r.push_back(q);
      L /= q;
    }
  for (int p = 7; L > 1; p += 2)
    while (L % p == 0) {
      r.push_back(p);
      L /= p;
    }
  return r;
}

void HostFFTPlan::init(int N) {
  dev.N = N;
  dev.M = N + 1;
  dev.L = 2 * (N + 1);
  std::vector<int> rad = factorize(dev.L);
  if ((int)rad.size() > kMaxStages) throw HipError("FFT length has too many factors");
  dev.nst = (int)rad.size();
  for (int i = 0; i < dev.nst; ++i) dev.rad[i] = rad[i];
  std::vector<double2> w(dev.L);
  for (int t = 0; t < dev.L; ++t) {
    long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)dev.L;
    w[t] = make_double2((double)cosl(ang), (double)-sinl(ang));
  }
  IB_HIP(hipMalloc(&W, sizeof(double2) * dev.L));
  IB_HIP(hipMemcpy(W, w.data(), sizeof(double2) * dev.L, hipMemcpyHostToDevice));
  dev.W = W;
}

void HostFFTPlan::release() {
  if (W) (void)hipFree(W);
  W = nullptr;
}

size_t dst_lds_bytes(const FFTPlan& pl, int nfft) { return (size_t)(2 * nfft + 1) * pl.L * sizeof(double2); }

bool fft_has_generic(const FFTPlan& pl) {
  for (int st = 0; st < pl.nst; ++st) {
    switch (pl.rad[st]) {
      case 8: case 4: case 2: case 3: case 5: break;  // the fixed radices of fft_run
      default: return true;
    }
  }
  return false;
}

namespace {
template <bool GEN>
const void* dst_x_kernel_t(int mode) {
  switch (mode) {
    case X_PLAIN: return (const void*)k_dst_x<X_PLAIN, GEN>;
    case X_APPLYBC: return (const void*)k_dst_x<X_APPLYBC, GEN>;
    case X_SUBFROM: return (const void*)k_dst_x<X_SUBFROM, GEN>;
    case X_ROT: return (const void*)k_dst_x<X_ROT, GEN>;
    case X_POINTROT: return (const void*)k_dst_x<X_POINTROT, GEN>;
    default: throw HipError("bad DST row mode");
  }
}
const void* dst_y_kernel(int mode) {
  switch (mode) {
    case Y_PLAIN: return (const void*)k_dst_y<Y_PLAIN>;
    case Y_SCALE2: return (const void*)k_dst_y<Y_SCALE2>;
    case Y_PRED: return (const void*)k_dst_y<Y_PRED>;
    default: throw HipError("bad DST column mode");
  }
}
template <bool GEN>
const void* dst_x_tab_kernel_t(int mode) {
  switch (mode) {
    case X_PLAIN: return (const void*)k_tab<DstXOp<X_PLAIN, GEN>>;
    case X_APPLYBC: return (const void*)k_tab<DstXOp<X_APPLYBC, GEN>>;
    case X_SUBFROM: return (const void*)k_tab<DstXOp<X_SUBFROM, GEN>>;
    case X_ROT: return (const void*)k_tab<DstXOp<X_ROT, GEN>>;
    default: throw HipError("DST row pass: no table kernel for this mode");
  }
}

const void* dst_x_kernel(int mode, bool gen) { return gen ? dst_x_kernel_t<true>(mode) : dst_x_kernel_t<false>(mode); }
}  // namespace

const void* dst_x_tab_kernel(int mode, bool gen) {
  return gen ? dst_x_tab_kernel_t<true>(mode) : dst_x_tab_kernel_t<false>(mode);
}
const void* dst_y_tab_kernel(int mode) {
  switch (mode) {
    case Y_SCALE2: return (const void*)k_tab<DstYOp<Y_SCALE2>>;
    case Y_PRED: return (const void*)k_tab<DstYOp<Y_PRED>>;
    default: throw HipError("DST column pass: no table kernel for this mode");
  }
}

void dst_set_max_lds(size_t bytes) {
  std::vector<const void*> fns;
  for (bool gen : {false, true}) {
    for (int mode : {X_PLAIN, X_APPLYBC, X_SUBFROM, X_ROT, X_POINTROT}) fns.push_back(dst_x_kernel(mode, gen));
    for (int mode : {X_PLAIN, X_APPLYBC, X_SUBFROM, X_ROT}) fns.push_back(dst_x_tab_kernel(mode, gen));
  }
  for (int mode : {Y_PLAIN, Y_SCALE2, Y_PRED}) fns.push_back(dst_y_kernel(mode));
  for (int mode : {Y_SCALE2, Y_PRED}) fns.push_back(dst_y_tab_kernel(mode));
  for (const void* f : fns) IB_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

TabEntry<XArgs> dst_x_entry(const FFTPlan& pl, const DstLaunch& l, const Geom& g, int rp, size_t max_lds,
                            size_t* lds) {
  TabEntry<XArgs> e{};
  XArgs& a = e.a;
  const bool strided = l.strided;
  if (!strided && l.batch > kMaxBatch) throw HipError("DST row pass: batch > 4 needs strides");
  if (!strided && (l.s_in || l.s_out || l.s_bc)) throw HipError("DST row pass: strides on an unstrided batch");
  a.strided = strided;
  for (int b = 0; b < (strided ? 1 : l.batch); ++b) {
    a.in[b] = l.in[b];
    a.out[b] = l.out[b];
  }
  a.s_in = l.s_in;
  a.s_out = l.s_out;
  a.s_bc = l.s_bc;
  a.c0 = l.c0;
  a.w = l.w;
  a.key = l.key;
  a.bc = l.bc;
  a.hx = l.hx;
  a.hy = l.hy;
  a.g = g;
  a.pl = pl;
  while (rp > 1 && dst_lds_bytes(pl, rp) > max_lds) rp /= 2;
  a.rp = rp;
  const int npairs = ceil_div(g.n - 1, 2);
  e.nbx = ceil_div(npairs, rp);
  e.nby = l.batch;
  *lds = dst_lds_bytes(pl, rp);
  if (*lds > max_lds) throw HipError("DST row pass needs more LDS than the device allows");
  return e;
}

TabEntry<YArgs> dst_y_entry(const FFTPlan& pl, int mode, const DstLaunch& l, const Geom& g, int cp,
                            size_t max_lds, size_t* lds) {
  TabEntry<YArgs> e{};
  YArgs& a = e.a;
  const bool strided = l.strided;
  if (!strided && l.batch > kMaxBatch) throw HipError("DST column pass: batch > 4 needs strides");
  if (!strided && (l.s_in || l.s_in2 || l.s_out)) throw HipError("DST column pass: strides on an unstrided batch");
  a.strided = strided;
  for (int b = 0; b < (strided ? 1 : l.batch); ++b) {
    a.in[b] = l.in[b];
    a.in2[b] = l.in2[b];
    a.out[b] = l.out[b];
  }
  a.s_in = l.s_in;
  a.s_in2 = l.s_in2;
  a.s_out = l.s_out;
  a.tab = l.tab;
  a.tab2 = l.tab2;
  a.div = l.div;
  a.g = g;
  a.pl = pl;
  const int mult = (mode == Y_PRED) ? 2 : 1;
  while (cp > 1 && dst_lds_bytes(pl, mult * cp) > max_lds) cp /= 2;
  a.cp = cp;
  const int npairs = ceil_div(g.m - 1, 2);
  const int ntiles = ceil_div(npairs, cp);
  a.chunk = ceil_div(ntiles, 8);
  e.nbx = 8 * a.chunk;
  e.nby = l.batch;
  *lds = dst_lds_bytes(pl, mult * cp);
  if (*lds > max_lds) throw HipError("DST column pass needs more LDS than the device allows");
  return e;
}

void dst_x(hipStream_t st, const FFTPlan& pl, int mode, const DstLaunch& l, const Geom& g, int rp,
           size_t max_lds) {
  size_t lds = 0;
  const TabEntry<XArgs> e = dst_x_entry(pl, l, g, rp, max_lds, &lds);
  void* args[] = {(void*)&e.a};
  IB_HIP(hipLaunchKernel(dst_x_kernel(mode, fft_has_generic(pl)), dim3(e.nbx, e.nby), dim3(256), args, lds, st));
  IB_LAUNCH_CHECK();
}

void dst_y(hipStream_t st, const FFTPlan& pl, int mode, const DstLaunch& l, const Geom& g, int cp,
           size_t max_lds) {
  size_t lds = 0;
  const TabEntry<YArgs> e = dst_y_entry(pl, mode, l, g, cp, max_lds, &lds);
  void* args[] = {(void*)&e.a};
  IB_HIP(hipLaunchKernel(dst_y_kernel(mode), dim3(e.nbx, e.nby), dim3(256), args, lds, st));
  IB_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------
// Form and tile shapes of the DST passes, picked by timing.  The row pairs
// per workgroup of the row pass (rp) and the column pairs of the column pass
// (cp) decide only how many transforms share a workgroup's LDS, never a
// transform's arithmetic, so every choice gives the same bits.  The
// matrix-core form (ib_dst_dense.hip) sums in another order: it is taken for
// a pass class only when it is at least 10% faster, so that a near tie does
// not flip between processes.  Each candidate that fits runs a few times on
// scratch fields between HIP events; the fastest is kept for the process,
// per grid shape.  Two classes are timed: one field per launch (the time
// step) and a strided batch of 8 fields (the Schur assembly, whose launches
// cover hundreds of columns).
//
// IB_DST_TILES=rp,cp forces the Stockham tiles; IB_DST_TUNE=0 keeps the
// fixed rule (rp 2 up to transform length 512, else 1; cp 1); neither times
// anything.  IB_DST_FORM=fft keeps the Stockham form, =dense takes the
// matrix-core form for both classes (its fastest configuration, or the
// (T, KS) = (1, 8) configuration when nothing is timed), =dense:k the matrix-core
// form in configuration k without timing, =auto (the default) times.
namespace {
std::mutex g_tile_mu;
std::map<std::array<long, 6>, DstTiles> g_tiles;
constexpr int kDenseDefault = 1;
constexpr int kTuneBatch = 8;

template <class F>
float time_launches(hipStream_t st, F&& launch) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  IB_HIP(hipEventCreate(&e0));
  IB_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  try {
    launch();  // warm-up (code objects, caches)
    launch();
    for (int rep = 0; rep < 3; ++rep) {
      IB_HIP(hipEventRecord(e0, st));
      for (int k = 0; k < 4; ++k) launch();
      IB_HIP(hipEventRecord(e1, st));
      IB_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      IB_HIP(hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, ms);
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best / 4.0f;
}

// 0 auto, 1 Stockham only, 2 matrix-core form, 3 matrix-core form in
// configuration *cfg
int form_rule(int* cfg) {
  const char* f = std::getenv("IB_DST_FORM");
  if (!f || !*f || std::strcmp(f, "auto") == 0) return 0;
  if (std::strcmp(f, "fft") == 0) return 1;
  if (std::strcmp(f, "dense") == 0) return 2;
  int k = -1;
  char tail = 0;
  if (std::sscanf(f, "dense:%d%c", &k, &tail) == 1 && k >= 0 && k < kDenseCfgs) {
    *cfg = k;
    return 3;
  }
  throw HipError("IB_DST_FORM must be fft, dense, dense:k (0 <= k < 5) or auto");
}

struct Scratch {
  double* p = nullptr;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
};
}  // namespace

DstTiles dst_tiles(hipStream_t st, const FFTPlan& px, const FFTPlan& py, const DenseMat* dx, const DenseMat* dy,
                   double* tq, const Geom& g, size_t max_lds) {
  DstTiles t;
  t.rp = px.L <= 512 ? 2 : 1;
  t.cp = 1;
  int forced_cfg = kDenseDefault;
  const int rule = form_rule(&forced_cfg);
  const bool dense_ok = dx && dy && dx->S && dy->S && tq;
  if (rule >= 2 && !dense_ok) throw HipError("IB_DST_FORM=dense: no sine matrices");
  const int fixed_form = rule >= 2 ? forced_cfg + 1 : 0;
  if (const char* f = std::getenv("IB_DST_TILES")) {
    int rp = 0, cp = 0;
    if (std::sscanf(f, "%d,%d", &rp, &cp) != 2 || rp < 1 || rp > 8 || cp < 1 || cp > 8)
      throw HipError("IB_DST_TILES must be rp,cp with 1 <= rp, cp <= 8");
    t.rp = rp;  // dst_x_entry / dst_y_entry halve a choice the LDS cannot hold
    t.cp = cp;
    t.xform = t.yform = t.xform_b = t.yform_b = fixed_form;
    t.source = 2;
    return t;
  }
  if (const char* f = std::getenv("IB_DST_TUNE"))
    if (std::strcmp(f, "0") == 0) {
      t.xform = t.yform = t.xform_b = t.yform_b = fixed_form;
      return t;
    }
  if (rule == 3) {  // a forced configuration: the Stockham tiles keep the fixed rule
    t.xform = t.yform = t.xform_b = t.yform_b = fixed_form;
    t.source = 2;
    return t;
  }
  const bool try_dense = dense_ok && rule != 1;
  const std::array<long, 6> key{g.m, g.n, px.L, py.L, (long)max_lds, (long)(rule * 2 + (try_dense ? 1 : 0))};
  std::lock_guard<std::mutex> lock(g_tile_mu);
  auto it = g_tiles.find(key);
  if (it != g_tiles.end()) return it->second;
  const long P = g.P;
  Scratch in, out;
  const size_t bytes = sizeof(double) * (size_t)P * kTuneBatch;
  IB_HIP(hipMalloc(&in.p, bytes));
  IB_HIP(hipMalloc(&out.p, bytes));
  IB_HIP(hipMemsetAsync(in.p, 0, bytes, st));
  // one field per launch: the Stockham tiles, then the matrix-core configurations
  DstLaunch lx;
  lx.in[0] = in.p;
  lx.out[0] = out.p;
  DstLaunch ly = lx;
  ly.tab = in.p;  // any table: the values do not change the time
  ly.div = 1.0;
  float best = 1e30f;
  for (int rp : {1, 2, 4}) {
    if (dst_lds_bytes(px, rp) > max_lds) continue;
    const float ms = time_launches(st, [&] { dst_x(st, px, X_PLAIN, lx, g, rp, max_lds); });
    if (ms < best) {
      best = ms;
      t.rp = rp;
    }
  }
  t.us[0] = best * 1e3f;
  best = 1e30f;
  for (int cp : {1, 2, 4}) {
    if (dst_lds_bytes(py, cp) > max_lds) continue;
    const float ms = time_launches(st, [&] { dst_y(st, py, Y_SCALE2, ly, g, cp, max_lds); });
    if (ms < best) {
      best = ms;
      t.cp = cp;
    }
  }
  t.us[2] = best * 1e3f;
  // a strided batch of kTuneBatch fields
  DstLaunch bx = lx;
  bx.batch = kTuneBatch;
  bx.strided = true;
  bx.s_in = bx.s_out = P;
  DstLaunch by = bx;
  by.tab = in.p;
  by.div = 1.0;
  t.us[4] = time_launches(st, [&] { dst_x(st, px, X_PLAIN, bx, g, t.rp, max_lds); }) * 1e3f / kTuneBatch;
  t.us[6] = time_launches(st, [&] { dst_y(st, py, Y_SCALE2, by, g, t.cp, max_lds); }) * 1e3f / kTuneBatch;
  if (try_dense) {
    // a class is timed only if the matrix-core form could beat the measured
    // Stockham time at the FP64 matrix peak (78.6 TFLOP/s; 2N flops per
    // point and product, two products in the column pass); IB_DST_FORM=dense
    // times every class
    const double peak = 78.6e6;  // flop per microsecond
    const double lb[4] = {2.0 * (g.m - 1) * P / peak, 4.0 * (g.n - 1) * P / peak,
                          2.0 * (g.m - 1) * P / peak, 4.0 * (g.n - 1) * P / peak};
    float best_us[4] = {1e30f, 1e30f, 1e30f, 1e30f};
    int cfg[4] = {kDenseDefault, kDenseDefault, kDenseDefault, kDenseDefault};
    const DstLaunch* ls[4] = {&lx, &ly, &bx, &by};
    for (int q = 0; q < 4; ++q) {
      const float fft_us = t.us[2 * q];
      if (rule != 2 && lb[q] >= fft_us) continue;
      const int nf = q < 2 ? 1 : kTuneBatch;
      for (int c = 0; c < kDenseCfgs; ++c) {
        const float us = 1e3f / nf * time_launches(st, [&] {
          if (q % 2 == 0) dst_dense_x(st, *dx, X_PLAIN, *ls[q], g, c);
          else dst_dense_y(st, *dy, Y_SCALE2, *ls[q], g, c, tq, kMaxBatch);
        });
        if (us < best_us[q]) {
          best_us[q] = us;
          cfg[q] = c;
        }
      }
      t.us[2 * q + 1] = best_us[q];
    }
    const auto pick = [&](int q) {
      if (rule == 2) return cfg[q] + 1;
      return (t.us[2 * q + 1] >= 0.0f && t.us[2 * q + 1] < 0.9f * t.us[2 * q]) ? cfg[q] + 1 : 0;
    };
    t.xform = pick(0);
    t.yform = pick(1);
    t.xform_b = pick(2);
    t.yform_b = pick(3);
  }
  IB_HIP(hipStreamSynchronize(st));
  t.source = 1;
  g_tiles[key] = t;
  return t;
}

}  // namespace ibg
