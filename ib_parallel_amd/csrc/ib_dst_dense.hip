// ib_dst_dense.hip -- the DST-I passes as FP64 matrix-core products.
//
// An unnormalised DST-I of length N (FFTW RODFT00, src/myfft.F90:168-213) is
// the product with the symmetric matrix
//     S[t][k] = 2 sin(pi (t+1)(k+1) / (N+1)),   t, k = 0..N-1.
// The row pass of a field (n-1 rows of length N = m-1, i fastest) is F S;
// the column pass (N = n-1 rows of m-1 values) is S F.  Both run on
// v_mfma_f64_16x16x4f64: a wavefront owns a 16 x 16 block of the product and
// accumulates it over 4-deep slices of the sum, reading its fragments from
// tiles staged in LDS.  A workgroup holds T x T such blocks and splits the
// sum over KS groups of wavefronts, whose partial blocks are added in LDS in
// a fixed order (so the bits do not depend on timing).  The sine operand is
// read as rows S[t][.] in both passes (S is symmetric).
//
// S is built once per transform length from a table of the 2(N+1) sines
// (long double, rounded once) and zero-padded to a multiple of 64 in both
// dimensions: every tile reads it without bounds checks, and the padded
// rows make the sum past t = N vanish whatever the field operand holds
// there (which the loads set to zero anyway).
//
// The fused forms of the Stockham passes are kept: apply_bc and rot in the
// load of the row pass, the corrector's subtraction in its store, the
// spectral scaling (ctci/ainv) and the predictor's combination in the store
// of a first column product, followed by a plain second column product.
//
// Cost: 2N flops per point and pass, against about 16 B of HBM per point for
// the Stockham form, so this form can only win where the Stockham passes are
// bound by latency (small fields, one field per launch), never on large
// batches; the handle times both per grid shape.
#include <cmath>
#include <mutex>
#include <vector>

#include "ib_device.h"
#include "ib_kernels.h"

namespace ibg {

typedef double dvec4 __attribute__((ext_vector_type(4)));

namespace {

enum { D_ROW = 0, D_ROW_APPLYBC, D_ROW_ROT, D_ROW_SUBFROM, D_COL, D_COL_SCALE, D_COL_PRED };

// Batch element b of an operand: p[0] + b*s when strided, else p[b].
struct Operand {
  const double* p[kMaxBatch];
  long s;
  int strided;
};
__device__ __forceinline__ const double* at(const Operand& o, int b) {
  return o.strided ? (o.p[0] ? o.p[0] + (long)b * o.s : nullptr) : o.p[b];
}

struct DenseArgs {
  Operand in, in2, out;
  const double* bc;  // apply_bc vector of batch element b at bc + b * s_bc
  long s_bc;
  const double* hx;  // rot(hx, hy) replaces the input (rot_int + rot_bc)
  const double* hy;
  const double* tab;   // spectral scale / predictor lam1
  const double* tab2;  // predictor lam1i
  double div;
  const double* S;  // N x N sine matrix, leading dimension ld, zero past N
  int ld, N;
  int M, Nc;    // rows and columns of the product
  int nti;      // tiles down the product
  int ntiles;   // tiles in all
  int chunk;    // consecutive tiles per XCD
  Geom g;
};

// Row-pass operand A[r][t]: node (i, j) = (t + 2, r + 2) of the input,
// zero outside the field.
template <int MODE>
__device__ __forceinline__ double row_elem(const DenseArgs& a, const double* in, const double* bc, int r, int t) {
  if (r >= a.M || t >= a.N) return 0.0;
  const Geom& g = a.g;
  const int i = t + 2, j = r + 2;
  if (MODE == D_ROW_ROT) {
    // vort(i,j) = hy(i,j) - hy(i-1,j) - hx(i,j) + hx(i,j-1)
    return a.hy[(i - 1) + g.m * (j - 1)] - a.hy[(i - 2) + g.m * (j - 1)] - a.hx[(i - 1) + (g.m + 1) * (j - 1)] +
           a.hx[(i - 1) + (g.m + 1) * (j - 2)];
  }
  double x = in[(long)r * a.N + t];
  if (MODE == D_ROW_APPLYBC) {
    // apply_bc(vort, sbc, 1.d0), src/operators_fluid.F90:539-549
    if (i == 2) x = x + bc[g.left + j - 1];
    if (i == g.m) x = x + bc[g.right + j - 1];
    if (j == 2) x = x + bc[g.bottom + i - 1];
    if (j == g.n) x = x + bc[g.top + i - 1];
  }
  return x;
}

// Tiles.  A workgroup computes a W x W tile of the product (W = 16T) with
// T*T wavefronts per group (each a 16 x 16 block) and KS groups, group g
// summing its own range of t.  The sum advances in chunks of BK = 32 values
// of t: each group stages its chunk of both operands in LDS with coalesced
// loads (consecutive lanes, consecutive doubles of one row), the loads of
// the next chunk being in flight while the products of this one run.  Tiles
// are kept [r][t] (the row pass's field rows: W rows of BK) or [t][x] (sine
// rows and the column pass's field rows: BK rows of W), padded by two
// doubles per row against bank conflicts.
constexpr int BK = 32;
constexpr int PRT = BK + 2;  // pitch of [r][t] tiles
template <int T>
struct Lay {
  static constexpr int W = 16 * T;                                      // tile edge
  static constexpr int PTX = W + 2;                                     // pitch of [t][x] tiles
  static constexpr int REG = (W * PRT > BK * PTX) ? W * PRT : BK * PTX;  // doubles per operand region
  static constexpr int NT = 64 * T * T;                                 // threads per group
  static constexpr int PER = BK * W / NT;                               // values per thread and operand (8/T)
};

template <int MODE, int T>
__device__ __forceinline__ void load_chunk(const DenseArgs& a, const double* in, const double* in2, const double* bc,
                                           int i0, int j0, int t0, int gt, double* va, double* vb, double* vc) {
  using Y = Lay<T>;
  const double* __restrict__ S = a.S;
  const long ld = a.ld;
  const int c_rt = gt % BK, r_rt = gt / BK;    // [r][t] tile: this thread's column and first row
  const int c_tx = gt % Y::W, r_tx = gt / Y::W;  // [t][x] tile
#pragma unroll
  for (int q = 0; q < Y::PER; ++q) {
    const int t = t0 + r_tx + q * (Y::NT / Y::W);
    if (MODE <= D_ROW_SUBFROM) {
      va[q] = row_elem<MODE>(a, in, bc, i0 + r_rt + q * (Y::NT / BK), t0 + c_rt);  // field rows [r][t]
      vb[q] = S[(long)t * ld + j0 + c_tx];                                           // sine rows [t][j]
    } else {
      va[q] = S[(long)t * ld + i0 + c_tx];  // sine rows [t][i] (S[i][t] = S[t][i])
      const int c = j0 + c_tx;
      const bool ok = t < a.N && c < a.Nc;
      vb[q] = ok ? in[(long)t * a.Nc + c] : 0.0;  // field rows [t][c]
      if (MODE == D_COL_PRED) vc[q] = ok ? in2[(long)t * a.Nc + c] : 0.0;
    }
  }
}

template <int MODE, int T>
__device__ __forceinline__ void store_chunk(double* A, double* B, double* C, int gt, const double* va,
                                            const double* vb, const double* vc) {
  using Y = Lay<T>;
  const int c_rt = gt % BK, r_rt = gt / BK;
  const int c_tx = gt % Y::W, r_tx = gt / Y::W;
#pragma unroll
  for (int q = 0; q < Y::PER; ++q) {
    const int tx = (r_tx + q * (Y::NT / Y::W)) * Y::PTX + c_tx;
    if (MODE <= D_ROW_SUBFROM) A[(r_rt + q * (Y::NT / BK)) * PRT + c_rt] = va[q];
    else A[tx] = va[q];
    B[tx] = vb[q];
    if (MODE == D_COL_PRED) C[tx] = vc[q];
  }
}

template <int MODE, int T, int KS>
__global__ void __launch_bounds__(64 * T * T * KS) k_dst_dense(DenseArgs a) {
  using Y = Lay<T>;
  constexpr bool ROW = MODE <= D_ROW_SUBFROM;
  constexpr bool PRED = MODE == D_COL_PRED;
  constexpr int NR = PRED ? 3 : 2;  // operand regions per group
  extern __shared__ double lds[];
  // workgroups are dealt to the 8 XCDs in turn: each XCD takes a contiguous
  // run of tiles, which walk down one band of product columns (one band of
  // S columns in the row pass, of field columns in the column pass)
  const int tile = (int)(blockIdx.x % 8) * a.chunk + (int)(blockIdx.x / 8);
  if (tile >= a.ntiles) return;
  const int bi = tile % a.nti, bj = tile / a.nti;
  const int b = blockIdx.y;
  const double* in = at(a.in, b);
  const double* in2 = PRED ? at(a.in2, b) : nullptr;
  double* out = const_cast<double*>(at(a.out, b));
  const double* bc = a.bc ? a.bc + (long)b * a.s_bc : nullptr;
  const int grp = threadIdx.x / Y::NT, gt = threadIdx.x % Y::NT;
  const int w = gt >> 6, lane = gt & 63;
  const int wm = w % T, wn = w / T;
  const int li = lane & 15, lk = lane >> 4;
  const int i0 = bi * Y::W, j0 = bj * Y::W;
  double* A = lds + (long)grp * NR * Y::REG;
  double* B = A + Y::REG;
  double* C = B + Y::REG;
  // this group's chunks [c0, c1) of the ceil(N / BK) in the sum
  const int nch = (a.N + BK - 1) / BK;
  const int per = (nch + KS - 1) / KS;
  const int c0 = min(grp * per, nch), c1 = min(c0 + per, nch);
  dvec4 acc = {0.0, 0.0, 0.0, 0.0};
  dvec4 acc2 = {0.0, 0.0, 0.0, 0.0};
  double va[Y::PER], vb[Y::PER], vc[Y::PER];
  if (c0 < c1) load_chunk<MODE, T>(a, in, in2, bc, i0, j0, c0 * BK, gt, va, vb, vc);
  // every group runs `per` rounds so that the workgroup's barriers match
  for (int r = 0; r < per; ++r) {
    const int c = c0 + r;
    __syncthreads();  // the previous round's products have read the tiles
    if (c < c1) store_chunk<MODE, T>(A, B, C, gt, va, vb, vc);
    __syncthreads();
    if (c + 1 < c1) load_chunk<MODE, T>(a, in, in2, bc, i0, j0, (c + 1) * BK, gt, va, vb, vc);
    if (c < c1) {
      // fragment layout of v_mfma_f64_16x16x4f64: lane l holds A[l%16][l/16]
      // and B[l/16][l%16]; the product's D[4r + l/16][l%16] lands in its
      // r-th accumulator
#pragma unroll
      for (int s = 0; s < BK / 4; ++s) {
        const int kt = 4 * s + lk;
        const double av = ROW ? A[(wm * 16 + li) * PRT + kt] : A[kt * Y::PTX + wm * 16 + li];
        const double bv = B[kt * Y::PTX + wn * 16 + li];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        if (PRED) acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, C[kt * Y::PTX + wn * 16 + li], acc2, 0, 0, 0);
      }
    }
  }
  if (KS > 1) {
    // groups 1..KS-1 leave their partial blocks in their own LDS area (its
    // tiles are no longer read); group 0 adds them in group order
    constexpr int NB = PRED ? 2 : 1;
    static_assert(Y::NT * 4 * NB <= NR * Y::REG, "partial blocks exceed a group's LDS area");
    __syncthreads();
    if (grp > 0) {
      double* p = A + (long)gt * 4 * NB;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        p[r] = acc[r];
        if (PRED) p[4 + r] = acc2[r];
      }
    }
    __syncthreads();
    if (grp > 0) return;
    for (int q = 1; q < KS; ++q) {
      const double* p = lds + (long)q * NR * Y::REG + (long)gt * 4 * NB;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        acc[r] = acc[r] + p[r];
        if (PRED) acc2[r] = acc2[r] + p[4 + r];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + wm * 16 + 4 * r + lk, j = j0 + wn * 16 + li;
    if (i >= a.M || j >= a.Nc) continue;
    const long o = (long)i * a.Nc + j;
    const double v = acc[r];
    if (MODE == D_ROW_SUBFROM) out[o] = out[o] - v;
    else if (MODE == D_COL_SCALE) out[o] = a.tab[o] * v / a.div;
    // lam1i * (dst(rhs terms) + lam1 * dst(omega)), src/operators_fluid.F90:657-661
    else if (MODE == D_COL_PRED) out[o] = a.tab2[o] * (v + a.tab[o] * acc2[r]);
    else out[o] = v;
  }
}

// tile configurations (T, KS) = (1,4), (1,8), (1,16), (2,2), (2,4): W = 16T,
// KS groups of T*T wavefronts
// a kernel with the launch it was built for: its LDS bytes, threads and
// tile edge
struct Kern {
  const void* fn;
  size_t lds;
  int threads, w;
};
template <int MODE, int T, int KS>
Kern kern_t() {
  return {(const void*)k_dst_dense<MODE, T, KS>, sizeof(double) * (size_t)KS * (MODE == D_COL_PRED ? 3 : 2) * Lay<T>::REG,
          64 * T * T * KS, Lay<T>::W};
}
template <int MODE>
Kern kern_m(int cfg) {
  switch (cfg) {
    case 0: return kern_t<MODE, 1, 4>();
    case 1: return kern_t<MODE, 1, 8>();
    // the predictor's two-input product does not fit 1024 threads' 128
    // registers without spills: it runs as (1, 8) there
    case 2:
      if constexpr (MODE == D_COL_PRED) return kern_t<MODE, 1, 8>();
      else return kern_t<MODE, 1, 16>();
    case 3: return kern_t<MODE, 2, 2>();
    case 4: return kern_t<MODE, 2, 4>();
    default: throw HipError("dense DST: bad configuration");
  }
}
Kern kern(int mode, int cfg) {
  switch (mode) {
    case D_ROW: return kern_m<D_ROW>(cfg);
    case D_ROW_APPLYBC: return kern_m<D_ROW_APPLYBC>(cfg);
    case D_ROW_ROT: return kern_m<D_ROW_ROT>(cfg);
    case D_ROW_SUBFROM: return kern_m<D_ROW_SUBFROM>(cfg);
    case D_COL: return kern_m<D_COL>(cfg);
    case D_COL_SCALE: return kern_m<D_COL_SCALE>(cfg);
    case D_COL_PRED: return kern_m<D_COL_PRED>(cfg);
    default: throw HipError("dense DST: bad mode");
  }
}

// the LDS limit of every kernel raised to what it needs, once per process
// (the attribute belongs to the function, for every device)
std::once_flag g_lds_once;
void set_lds_limits() {
  for (int mode = D_ROW; mode <= D_COL_PRED; ++mode)
    for (int c = 0; c < kDenseCfgs; ++c) {
      const Kern k = kern(mode, c);
      IB_HIP(hipFuncSetAttribute(k.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)k.lds));
    }
}

int round_up(int v, int q) { return ceil_div(v, q) * q; }

// Batch elements [b0, b0 + nb) of a DstLaunch operand.
Operand view(const double* const* p, bool strided, long s, int b0, int nb) {
  Operand o{};
  o.strided = strided ? 1 : 0;
  o.s = s;
  if (strided) o.p[0] = p[0] ? p[0] + (long)b0 * s : nullptr;
  else
    for (int q = 0; q < nb; ++q) o.p[q] = p[b0 + q];
  return o;
}
Operand strided_op(const double* p, long s) {
  Operand o{};
  o.strided = 1;
  o.s = s;
  o.p[0] = p;
  return o;
}

void launch(hipStream_t st, const DenseMat& S, int mode, DenseArgs& a, int nb, int cfg) {
  if (cfg < 0 || cfg >= kDenseCfgs) throw HipError("dense DST: bad configuration");
  if (nb < 1 || nb > 65535) throw HipError("dense DST: batch out of range");
  std::call_once(g_lds_once, set_lds_limits);
  const Kern k = kern(mode, cfg);
  const int W = k.w;
  a.S = S.S;
  a.ld = S.ld;
  a.N = S.N;
  // sine rows are read up to the chunk end, sine columns (the product's
  // columns in the row pass, its rows in the column pass: N of them) up to
  // the tile end
  if (round_up(S.N, BK) > S.ld || round_up(S.N, W) > S.ld || (a.M != S.N && a.Nc != S.N))
    throw HipError("dense DST: tiles exceed the padding");
  a.nti = ceil_div(a.M, W);
  a.ntiles = a.nti * ceil_div(a.Nc, W);
  a.chunk = ceil_div(a.ntiles, 8);
  void* args[] = {(void*)&a};
  IB_HIP(hipLaunchKernel(k.fn, dim3(8 * a.chunk, nb), dim3(k.threads), args, k.lds, st));
  IB_LAUNCH_CHECK();
}

}  // namespace

void DenseMat::init(int n) {
  release();
  N = n;
  ld = round_up(std::max(n, 1), 64);
  const long L2 = 2L * (n + 1);
  std::vector<double> sn(L2);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (long q = 0; q < L2; ++q) sn[q] = (double)(2.0L * sinl(pi * (long double)q / (long double)(n + 1)));
  std::vector<double> h((size_t)ld * ld, 0.0);
  for (long t = 0; t < n; ++t)
    for (long k = 0; k < n; ++k) h[t * ld + k] = sn[((t + 1) * (k + 1)) % L2];
  IB_HIP(hipMalloc(&S, sizeof(double) * h.size()));
  IB_HIP(hipMemcpy(S, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
}

void DenseMat::release() {
  if (S) (void)hipFree(S);
  S = nullptr;
}

void dst_dense_x(hipStream_t st, const DenseMat& S, int mode, const DstLaunch& l, const Geom& g, int cfg) {
  int dm;
  switch (mode) {
    case X_PLAIN: dm = D_ROW; break;
    case X_APPLYBC: dm = D_ROW_APPLYBC; break;
    case X_ROT: dm = D_ROW_ROT; break;
    case X_SUBFROM: dm = D_ROW_SUBFROM; break;
    default: throw HipError("dense DST row pass: no form for this mode");
  }
  if (S.N != g.m - 1) throw HipError("dense DST row pass: matrix length differs from the rows");
  if (!l.strided && l.batch > kMaxBatch) throw HipError("dense DST row pass: batch > 4 needs strides");
  DenseArgs a{};
  a.in = view(l.in, l.strided, l.s_in, 0, l.batch);
  a.out = view(const_cast<const double* const*>(l.out), l.strided, l.s_out, 0, l.batch);
  a.bc = l.bc;
  a.s_bc = l.s_bc;
  a.hx = l.hx;
  a.hy = l.hy;
  a.g = g;
  a.M = g.n - 1;
  a.Nc = g.m - 1;
  launch(st, S, dm, a, l.batch, cfg);
}

void dst_dense_y(hipStream_t st, const DenseMat& S, int mode, const DstLaunch& l, const Geom& g, int cfg,
                 double* tq, int tq_fields) {
  if (S.N != g.n - 1) throw HipError("dense DST column pass: matrix length differs from the columns");
  if (!l.strided && l.batch > kMaxBatch) throw HipError("dense DST column pass: batch > 4 needs strides");
  DenseArgs a{};
  a.g = g;
  a.M = g.n - 1;
  a.Nc = g.m - 1;
  const long P = (long)a.M * a.Nc;
  if (mode == Y_PLAIN) {
    a.in = view(l.in, l.strided, l.s_in, 0, l.batch);
    a.out = view(const_cast<const double* const*>(l.out), l.strided, l.s_out, 0, l.batch);
    launch(st, S, D_COL, a, l.batch, cfg);
    return;
  }
  if (mode != Y_SCALE2 && mode != Y_PRED) throw HipError("dense DST column pass: no form for this mode");
  if (!tq || tq_fields < 1) throw HipError("dense DST column pass: no intermediate buffer");
  // scaled (or combined) first product into tq, plain second product out,
  // tq_fields batch elements at a time
  for (int b0 = 0; b0 < l.batch; b0 += tq_fields) {
    const int nb = std::min(tq_fields, l.batch - b0);
    DenseArgs a1 = a;
    a1.in = view(l.in, l.strided, l.s_in, b0, nb);
    if (mode == Y_PRED) a1.in2 = view(l.in2, l.strided, l.s_in2, b0, nb);
    a1.out = strided_op(tq, P);
    a1.tab = l.tab;
    a1.tab2 = l.tab2;
    a1.div = l.div;
    launch(st, S, mode == Y_PRED ? D_COL_PRED : D_COL_SCALE, a1, nb, cfg);
    DenseArgs a2 = a;
    a2.in = strided_op(tq, P);
    a2.out = view(const_cast<const double* const*>(l.out), l.strided, l.s_out, b0, nb);
    launch(st, S, D_COL, a2, nb, cfg);
  }
}

}  // namespace ibg
