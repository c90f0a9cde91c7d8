// ib_kernels.h -- host-side launchers of the HIP kernels (one translation
// unit per kernel family; all launches go to the caller's stream).
#pragma once

#include <vector>

#include "ib_common.h"

namespace ibg {

// ---------------------------------------------------------------- DST-I ---
// Row (x) pass modes
enum { X_PLAIN = 0, X_APPLYBC = 1, X_SUBFROM = 2, X_ROT = 3, X_POINTROT = 4 };
// Column (y) pass modes
enum { Y_PLAIN = 0, Y_SCALE2 = 1, Y_PRED = 2 };

constexpr int kMaxBatch = 4;

// A launch covers `batch` arrays: in[b]/out[b] for b < kMaxBatch, or, when a
// stride is non-zero, base + b * stride (any batch size up to 65535).
struct DstLaunch {
  int batch = 1;
  bool strided = false;  // batch b at in[0] + b*s_in, out[0] + b*s_out, ...
  long s_in = 0, s_in2 = 0, s_out = 0, s_bc = 0;
  int c0 = 0;                         // X_POINTROT: force entry of batch 0
  const double* w = nullptr;          // X_POINTROT: weights (49 per entry)
  const int* key = nullptr;           // X_POINTROT: face keys (49 per entry)
  const double* in[kMaxBatch] = {};
  const double* in2[kMaxBatch] = {};  // Y_PRED: second input b
  double* out[kMaxBatch] = {};
  const double* bc = nullptr;         // X_APPLYBC boundary vector
  const double* hx = nullptr;         // X_ROT: node values are rot(hx, hy)
  const double* hy = nullptr;
  const double* tab = nullptr;        // Y_SCALE2 table / Y_PRED lam1
  const double* tab2 = nullptr;       // Y_PRED lam1i
  double div = 1.0;                   // Y_SCALE2 divisor
};

// Builds the Stockham plan (radix list) and twiddle table for L = 2(N+1).
struct HostFFTPlan {
  FFTPlan dev{};
  double2* W = nullptr;
  void init(int N);
  void release();
};

// Device arguments of a row pass: batch b reads in[b] (or in[0] + b*s_in
// when strided) and writes out[b]; the plan of the row length travels with
// them.
struct XArgs {
  const double* in[kMaxBatch];
  double* out[kMaxBatch];
  const double* bc;
  const double* hx;  // X_ROT: the node value is rot(hx, hy)
  const double* hy;
  long s_in, s_out, s_bc;  // strided: batch b at base[0] + b * stride
  int strided;              // 0: batch b at in[b], out[b] (b < kMaxBatch)
  int c0;                  // X_POINTROT: force entry (or face) of batch element 0
  const double* w;         // X_POINTROT: weights, 49 per entry; null: unit face sources
  const int* key;          // X_POINTROT: face keys, 49 per entry (or one per face)
  Geom g;
  int rp;
  FFTPlan pl;
};

// Device arguments of a column pass.
struct YArgs {
  const double* in[kMaxBatch];
  const double* in2[kMaxBatch];
  double* out[kMaxBatch];
  long s_in, s_in2, s_out;  // strided: batch b at base[0] + b * stride
  int strided;              // 0: batch b at in[b], out[b] (b < kMaxBatch)
  const double* tab;
  const double* tab2;
  double div;
  Geom g;
  int cp;
  int chunk;  // column-pair tiles per XCD (XCD-aware tile order)
  FFTPlan pl;
};

// x pass over rows of length m-1 (n-1 rows); rp = row pairs per workgroup
void dst_x(hipStream_t st, const FFTPlan& pl, int mode, const DstLaunch& a, const Geom& g, int rp,
           size_t max_lds);
// y pass over columns of length n-1 (m-1 columns); cp = column pairs per workgroup
void dst_y(hipStream_t st, const FFTPlan& pl, int mode, const DstLaunch& a, const Geom& g, int cp,
           size_t max_lds);
// The DST-I matrix S[t][k] = 2 sin(pi (t+1)(k+1)/(N+1)) in device memory,
// leading dimension ld (N rounded up to 64), zero past N.
struct DenseMat {
  double* S = nullptr;
  int N = 0, ld = 0;
  void init(int N);
  void release();
};
// tile shapes of the matrix-core passes: (T, KS) = (1,4), (1,8), (1,16),
// (2,2), (2,4) -- T x T blocks of 16 x 16 per workgroup, the sum split over
// KS groups of T x T wavefronts
constexpr int kDenseCfgs = 5;
// The DST passes as FP64 matrix-core products (every row mode but
// X_POINTROT, every column mode).  Y_SCALE2 and Y_PRED run as two products
// through tq, which holds tq_fields arrays of (m-1)(n-1) values.
void dst_dense_x(hipStream_t st, const DenseMat& S, int mode, const DstLaunch& a, const Geom& g, int cfg);
void dst_dense_y(hipStream_t st, const DenseMat& S, int mode, const DstLaunch& a, const Geom& g, int cfg,
                 double* tq, int tq_fields);

// Form and tile shapes of the DST passes, timed per grid shape once per
// process.  Stockham form: row pairs per workgroup of the row pass (rp),
// column pairs of the column pass (cp).  Forms: 0 Stockham, k > 0 the
// matrix-core form with configuration k-1; xform/yform for the passes of
// the time step (one field per launch), xform_b/yform_b for the batched
// passes of the Schur assembly.  us: best measured microseconds per field
// of {Stockham, matrix-core} x {row, column} x {step, batched} (index
// 4*batched + 2*column + matrix-core), negative where not timed.
// source: 0 fixed rule, 1 timed, 2 forced by IB_DST_TILES or IB_DST_FORM.
struct DstTiles {
  int rp = 1, cp = 1, source = 0;
  int xform = 0, yform = 0, xform_b = 0, yform_b = 0;
  float us[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
};
// dx/dy (may be null: Stockham only) are the matrices of the row and column
// lengths; tq holds kMaxBatch arrays of (m-1)(n-1) values for the timing.
DstTiles dst_tiles(hipStream_t st, const FFTPlan& px, const FFTPlan& py, const DenseMat* dx, const DenseMat* dy,
                   double* tq, const Geom& g, size_t max_lds);
// LDS bytes a launch needs
size_t dst_lds_bytes(const FFTPlan& pl, int nfft);
// The arguments and grid dst_x / dst_y would launch with, as a table entry,
// and the LDS bytes of that launch; the kernel of a mode for table launches.
TabEntry<XArgs> dst_x_entry(const FFTPlan& pl, const DstLaunch& l, const Geom& g, int rp, size_t max_lds,
                            size_t* lds);
TabEntry<YArgs> dst_y_entry(const FFTPlan& pl, int mode, const DstLaunch& l, const Geom& g, int cp,
                            size_t max_lds, size_t* lds);
// true when a stage of the plan has a radix other than 2, 3, 4, 5, 8: only
// the row-pass kernels built with the generic-radix stage may run such a plan
bool fft_has_generic(const FFTPlan& pl);
const void* dst_x_tab_kernel(int mode, bool gen);
const void* dst_y_tab_kernel(int mode);
// Raises the dynamic-LDS limit of the DST kernels to `bytes` (call once).
void dst_set_max_lds(size_t bytes);

// ------------------------------------------------------------- stencils ---
void fill(hipStream_t st, double* p, long n, double v);
void copy(hipStream_t st, double* dst, const double* src, long n);
// rot (C^T): vort = hy(i,j)-hy(i-1,j)-hx(i,j)+hx(i,j-1)
void rot(hipStream_t st, const Geom& g, const double* hx, const double* hy, double* vort);
// coarsify fine level -> centre of coarse level (batch arrays `stride` apart)
void coarsify(hipStream_t st, const Geom& g, const double* fine, double* coarse, int batch = 1, long stride = 0);
// get_bc of `nlev` coarse levels at once: bc[l] = fac * get_bc(src[l])
void get_bc(hipStream_t st, const Geom& g, const double* src, long src_stride, double* bc, long bc_stride,
            int nlev, double fac);
// curl (C) of s with boundary sbc into qx,qy; if bc_out: bc_out = get_bc(s, 1)
void curl_getbc(hipStream_t st, const Geom& g, const double* s, const double* sbc, double* qx, double* qy,
                double* bc_out);
// nonlinear + rot + apply_bc + AB2/Euler combine for the predictor of one level.
// r = con1*rhs + con2*rold + rhsbc; rhs_old <- rhs.  raw_rhs (optional) gets rhs.
void nl_rhs(hipStream_t st, const Geom& g, const double* omega, const double* qx, const double* qy,
            double q0x, const double* lastbc, const double* omegabc, double vfac, double con1, double con2,
            int first_step, double* rhs_old, double* r, double* raw_rhs);
void sub_inplace(hipStream_t st, double* a, const double* b, long n);

// Device arguments of the stencil kernels, for table launches.
struct FillArgs {
  double* p;
  long n;
  double v;
};
struct GetBcArgs {
  Geom g;
  const double* src;
  long sstride;
  double* bc;
  long bstride;
  double fac;
};
struct CoarsenArgs {
  Geom g;
  const double* wf;
  double* wc;
  long stride;
};
struct CurlArgs {
  Geom g;
  const double* s;
  const double* sbc;
  double* qx;
  double* qy;
  double* bc_out;
};
struct NLArgs {
  Geom g;
  const double* w;
  const double* qx;
  const double* qy;
  double q0x;
  const double* lastbc;
  const double* omegabc;
  double vfac, con1, con2;
  int first;
  const int* step;  // if set, the first step is *step == 1 (first is ignored)
  double* rhs_old;
  double* r;
  double* raw;
};
TabEntry<FillArgs> fill_entry(double* p, long n, double v);
TabEntry<GetBcArgs> get_bc_entry(const Geom& g, const double* src, long src_stride, double* bc, long bc_stride,
                                 int nlev, double fac);
TabEntry<CoarsenArgs> coarsify_entry(const Geom& g, const double* fine, double* coarse, int batch, long stride);
TabEntry<CurlArgs> curl_getbc_entry(const Geom& g, const double* s, const double* sbc, double* qx, double* qy,
                                    double* bc_out);
TabEntry<NLArgs> nl_rhs_entry(const Geom& g, const double* omega, const double* qx, const double* qy, double q0x,
                              const double* lastbc, const double* omegabc, double vfac, double con1, double con2,
                              const int* step, double* rhs_old, double* r);
// kernels for table launches of the stencil and body families
enum { TK_FILL, TK_GET_BC, TK_COARSIFY, TK_CURL_GETBC, TK_NL_RHS, TK_REG, TK_REDIST, TK_REGT, TK_FSI_RHS, TK_GEMV };
const void* fluid_tab_kernel(int kind);

// ---------------------------------------------------- immersed boundary ---
// weights of setup_reg / deltafnc for every (entry, tap); face key per tap:
// x-face index for x entries, Fx + y-face index for y entries; err flag set on
// a stencil leaving the grid (key then clamped to 0, weight 0).
// weights and tap keys of every body point; with `changed`, *changed is set
// to 1 when some tap's key differs from the one key held before the call
void setup_reg(hipStream_t st, const Geom& g, int nf, const double* xb, const double* yb, const int* comp,
               double delta, double offx, double offy, int f32lit, double* w, int* key, int* err,
               int* changed = nullptr);
// The touched faces of a tap list and each face's taps in tap order: nruns
// faces uniq[r], their taps tap_sorted[off[r] .. off[r] + cnt[r]).  key holds
// the tap keys (setup_reg); build() forms the rest from them on the device
// (hand-written: the entries sorted by centre face, each tap's face owned by
// the lowest entry touching it, one scan for the offsets).
struct RegCSR {
  int ntaps = 0;
  int *key = nullptr, *tap_sorted = nullptr;
  int *uniq = nullptr, *cnt = nullptr, *off = nullptr, *nruns = nullptr;
  int *own = nullptr, *toff = nullptr;    // per tap: count if it owns its face, offset if so
  int *rank = nullptr, *otap = nullptr;   // per tap: rank among its face's taps, the owning tap
  int* tsum = nullptr;                    // per tile of 1024 taps: owners and counts (then their scan)
  int *skey = nullptr, *sent = nullptr;   // entries sorted by centre key
  long long* sortv = nullptr;             // sort buffer of larger entry counts
  int n2 = 1;
  void alloc(int ntaps, int max_key);
  void release();
  void build(hipStream_t st, const Geom& g);
};
// reg (E^T): hx/hy[face] = sum over taps of w*f[tap/49] in tap order; faces not
// touched are left as they are (callers zero them).
void reg(hipStream_t st, const Geom& g, const RegCSR& c, const double* w, const double* f, double* hx,
         double* hy);
// regT (E): out[k] = base[k] + sum_t w[k,t]*(field[key] + shift) in tap order;
// shift_x is added to x-face values (q0 of level 1), base may be null (0).
void regT(hipStream_t st, const Geom& g, int nf, const double* w, const int* key, const double* hx,
          const double* hy, double shift_x, const double* base, double* out);
// Columns c0..c0+nc-1 of the Schur matrix (ab_times applied to unit
// vectors): S(r, j) = sum_t w[r,t] * curl(s1_j, bc1_j)(face) + b_times2(e_j)(r),
// with s1_j the level-1 streamfunction of column j; S is nf x nf, column major
void schur_columns(hipStream_t st, const Geom& g, int nf, int c0, int nc, const double* w, const int* key,
                   const double* s1, long s_stride, const double* bc1, long bc_stride, const int* tors_of,
                   const TorsBody* tb, const double* theta, double dt, double delta, double* S);
// Level-1 flux operator on a face set F (the faces near the bodies):
// G(b, a) = curl(s1_a, bc1_a)(F[b]) for columns a = c0..c0+nc-1, where s1_a
// is the level-1 streamfunction of the unit face source F[a]; G is nF x nF,
// column major.
void green_columns(hipStream_t st, const Geom& g, int nF, int c0, int nc, const int* fkeys, const double* s1,
                   long s_stride, const double* bc1, long bc_stride, double* G);
// tapidx[t] = fmap[key[t]] for the nf*49 taps; flag[0] = 1 if any tap's face
// is outside the face set (fmap < 0)
void map_taps(hipStream_t st, int nf, const int* key, const int* fmap, int* tapidx, int* flag);
// The Schur matrix from G: T = G W^T (nF x nf), then
// S(i, j) = sum_p w[i,p] T(tapidx[i,p], j) + b_times2(e_j)(i).
void schur_from_green(hipStream_t st, int nf, int nF, const double* w, const int* tapidx, const double* G,
                      double* T, const int* tors_of, const TorsBody* tb, const double* theta, double dt,
                      double delta, double* S);
// The same after a weight rebuild that moved only the entries mov[0..nmov)
// (the torsional bodies): the columns of T of those entries, and the
// entries of S in one of their rows or columns; T's other columns and S's
// block of the other entries (rig[0..nrig)) are kept as they are.
void schur_from_green_moving(hipStream_t st, int nf, int nF, const double* w, const int* tapidx, const double* G,
                             double* T, const int* tors_of, const TorsBody* tb, const double* theta, double dt,
                             double delta, double* S, int nmov, const int* mov, int nrig, const int* rig);
// y = S x for the dense nf x nf Schur matrix (column major)
void gemv(hipStream_t st, int nf, const double* S, const double* x, double* y);
// Device arguments of reg / redistribute (the CSR of the taps), regT,
// fsi_rhs and gemv, for table launches.
struct RegArgs {
  Geom g;
  int ntaps;
  const int* nruns;
  const int* uniq;
  const int* off;
  const int* cnt;
  const int* taps;
  const double* w;
  const double* f;
  double* hx;
  double* hy;
};
struct RegTArgs {
  Geom g;
  int nf;
  const double* w;
  const int* key;
  const double* hx;
  const double* hy;
  double shift_x;
  const double* base;
  double* out;
};
struct FsiRhsArgs {
  Geom g;
  int nf;
  const int* tors_of;
  const TorsBody* tbs;
  int nt;
  const double* th;
  double dt, delta;
  const double* w;
  const int* key;
  const double* qx;
  const double* qy;
  double q0x;
  double* rhsf;
};
struct GemvArgs {
  int nf;
  const double* S;
  const double* x;
  double* y;
};
TabEntry<RegArgs> reg_entry(const Geom& g, const RegCSR& c, const double* w, const double* f, double* hx, double* hy);
TabEntry<RegArgs> redist_entry(const Geom& g, const RegCSR& c, const double* w, const double* f, double* fx_out,
                               double* fy_out);
TabEntry<RegTArgs> regT_entry(const Geom& g, int nf, const double* w, const int* key, const double* hx,
                              const double* hy, double shift_x, const double* base, double* out);
TabEntry<FsiRhsArgs> fsi_rhs_entry(const Geom& g, int nf, const int* tors_of_entry, const TorsBody* tb, int nt,
                                   const double* th, double dt, double delta, const double* w, const int* key,
                                   const double* qx, const double* qy, double q0x, double* rhsf);
TabEntry<GemvArgs> gemv_entry(int nf, const double* S, const double* x, double* y);
const void* body_tab_kernel(int kind);
const void* gemv_tab_kernel();
// redistribute (src/operators_struct.F90:735-834) in CSR form: writes
// finter = sum(w f)/sum(w) (if sum(w) > 1e-10, else 0) into touched faces.
void redist_faces(hipStream_t st, const Geom& g, const RegCSR& c, const double* w, const double* f,
                  double* fx_out, double* fy_out);

// ------------------------------------------------ sub-domain approach ---
// B(r, c) at the listed rows of columns c0..c0+nc-1 (rows[colptr[c] ..
// colptr[c+1])), from the batch's level-1 streamfunctions; nentries =
// colptr[c0+nc] - colptr[c0]
void sd_columns(hipStream_t st, const Geom& g, int c0, int nc, long nentries, const long* colptr, const int* rows,
                const double* w, const int* key, const double* s1, long s_stride, const double* bc1, long bc_stride,
                double* vals);
// lumat = P B^T P^T + body_mat (nf x nf, column major), B by columns with
// rows ascending
void sd_lumat(hipStream_t st, int nf, int np, const int* pidx, const double* pw, const long* colptr, const int* rows,
              const double* vals, const int* tors_of, const TorsBody* tb, const double* theta, double dt,
              double delta, double* lumat);
// rhsf(j) = sum_p pw(j,p) r(pidx(j,p)) + rhsf(j)
void sd_interp(hipStream_t st, int nf, int np, const int* pidx, const double* pw, const double* r, double* rhsf);
// h0(used[u]) = sum over the (j, w) of used entry u of w f(j)
void sd_ptf(hipStream_t st, int nused, const int* used, const int* uptr, const int* uj, const double* uw,
            const double* f, double* h0);

// ------------------------------------------------------ discrete adjoint ---
// transposes of curl's boundary reads, get_bc, apply_bc, coarsify and of the
// nonlinear term's Jacobian (ib_adjoint.hip)
void curl_adj_bc(hipStream_t st, const Geom& g, const double* qx, const double* qy, double* bcb);
void getbc_adj(hipStream_t st, const Geom& g, const double* bcb, double fac, double* sc);
void applybc_adj(hipStream_t st, const Geom& g, const double* vb, double fac, double* bcb);
void applybc_adj_acc(hipStream_t st, const Geom& g, const double* vb, double* bcb);
void pred_bc_adj(hipStream_t st, const Geom& g, const double* rb, double vfac, double invRe, const double* lastbc,
                 const double* omegabc, double* bcb, double* lbb, double* racc_bc);
void coarsify_adj(hipStream_t st, const Geom& g, double* wc, double* wf);
void nl_adj(hipStream_t st, const Geom& g, const double* w, const double* qx, const double* qy, double q0x,
            const double* lastbc, const double* rb, double* nlxb, double* nlyb, double* ub, double* vb, double* wb,
            double* lbb, double* qxb, double* qyb);
// d w / d xb, d w / d yb of every (entry, tap) of setup_reg (ib_body.hip)
void setup_reg_d(hipStream_t st, const Geom& g, int nf, const double* xb, const double* yb, const int* comp,
                 double delta, double offx, double offy, int f32lit, double* dwx, double* dwy);
// TT[j + nf b] = (G W^T)(b, j), U[i + nf b] = (W G)(i, b)
void green_tt_u(hipStream_t st, int nf, int nF, const double* w, const int* tapidx, const double* G, double* TT,
                double* U);
void gather_taps(hipStream_t st, const Geom& g, int ntaps, const int* key, const double* qx, const double* qy,
                 double q0x, double* out);
void pos_terms(hipStream_t st, const Geom& g, int nf, const int* key, const int* tapidx, const double* dwx,
               const double* dwy, const double* TT, const double* U, const double* u, const double* v,
               const double* qt, const double* hbx, const double* hby, double* gx, double* gy);
// torsional bodies in the face-space form (ib_adjoint.hip)
void dthw(hipStream_t st, int nf, int nt, const int* tors_of, const TorsBody* tb, const double* th, const double* dwx,
          const double* dwy, double* out);
void face_gather(hipStream_t st, const Geom& g, int nF, const int* fkeys, const double* hx, const double* hy,
                 double* out);
void face_regT(hipStream_t st, int nf, const double* w, const int* tapidx, const double* zf, double* out);
void tap_contract(hipStream_t st, int i0, int i1, const double* wt, const double* coef, const double* ftap,
                  const double* fset, const int* tapidx, double* out);
void gemv_t(hipStream_t st, int n, const double* G, const double* x, double* y);
void transpose(hipStream_t st, int n, const double* A, double* AT);
void tors_block(hipStream_t st, const TorsBody* tb, const double* th, int nt, int t, int at0, double dt, double delta,
                const double* f, const double* a, double* r1, double* zc, double* s);
// gx, gy += the step's position partials of every entry; sx, sy (may be
// null) get them alone
void pos_fsi(hipStream_t st, int nf, const int* tors_of, const int* tapidx, const double* dwx, const double* dwy,
             const double* l, const double* f, const double* hb, const double* qt, const double* vT, const double* uT,
             double* gx, double* gy, double* sx, double* sy);
// acc[t] += the step's partials of torsional body t projected on d(points)/d(ds)
void ds_weights(hipStream_t st, const TorsBody* tb, const double* th, int nt, const double* sx, const double* sy,
                double* acc);
void dot_set(hipStream_t st, int i0, int i1, const double* a, const double* b, double* out);
void rot_acc(hipStream_t st, const Geom& g, const double* hx, const double* hy, double* v);
void rhs_adj(hipStream_t st, long n, const double* r, double c1, double c2, double* ro, double* rb);
void re_pred(hipStream_t st, long n, const double* ws, const double* w, const double* wsb, const double* rb,
             const double* dlz, double invN, double invRe, double* acc, double* wb);
void re_corr(hipStream_t st, long n, const double* wb, const double* aiwb, const double* aiv, double invRe,
             double* acc);
void sum_acc(hipStream_t st, long n, const double* x, double scale, double* out);
void axpy(hipStream_t st, long n, double a, const double* x, double* y);
void mul(hipStream_t st, long n, const double* x, const double* t, double* y);
void dot_acc(hipStream_t st, long n, const double* a, const double* b, double scale, double* out);
void gemv_symup(hipStream_t st, int n, const double* A, const double* x, double* y);
void symmetrize_upper(hipStream_t st, int n, const double* A, double* S);

// ---------------------------------------------------- torsional bodies ---
// y[body entries] += b_times2(x); one workgroup per body
void body_btimes(hipStream_t st, int nt, const TorsBody* tb, const double* theta, double dt, double delta,
                 const double* x, double* y);
// rhsf = compute_rhs (torsional entries, 0 elsewhere) + E qh
void fsi_rhs(hipStream_t st, const Geom& g, int nf, const int* tors_of_entry, const TorsBody* tb, int nt,
             const double* th, double dt, double delta, const double* w, const int* key, const double* qx,
             const double* qy, double q0x, double* rhsf);
// compute_dx + update_xb; th = [theta, thetad, thetadd, theta0, thetad0, thetadd0]
void body_dx(hipStream_t st, int nt, const TorsBody* tb, double* th, double dt, double delta, const double* f,
             double* xb, double* yb, double* dtheta_abs);
void body_initial_acc(hipStream_t st, int nt, const TorsBody* tb, double* th, double dt, double delta,
                      const double* f);

// ------------------------------------------------------------------ GMRES ---
struct GmresDev {
  double ttol, rn0, bnorm, res;
  int its, it, reason, first;
  double tt;  // norm of the last orthogonalised vector (multi-kernel iterations)
  int hap;
  int xmode;  // exchange of the last cooperative solve: 1 through one XCD's L2, 2 device scope
};
void gm_start(hipStream_t st, int nf, int kmax, const double* b, const double* Ax, double* V, GmresDev* s,
              double* H, double* grs, double rtol, double atol);
// one Arnoldi step: wavefront-parallel CGS dot products over several
// workgroups, then orthogonalisation, norm and Givens update in one workgroup
// one Arnoldi step after the product w = A v_it: wavefront-parallel CGS dot
// products, multi-workgroup orthogonalisation with partial norms (pnorm
// holds ceil(nf/64) doubles), the Givens update in one wavefront and the
// normalisation of the new basis vector
void gm_arnoldi(hipStream_t st, int nf, int kmax, double* V, GmresDev* s, double* H, double* cc, double* ss,
                double* grs, double* pnorm);
void gm_buildsol(hipStream_t st, int nf, int kmax, const double* V, GmresDev* s, const double* H,
                 const double* grs, double* x, double* scratch);
// The whole GMRES solve with a dense nf x nf Schur matrix in one workgroup;
// writes its, reason (1 converged, -1 diverged, -2 iteration limit) into *s.
void gmres_dense(hipStream_t st, int nf, int kmax, int maxit, const double* S, const double* b, double* x, double* V,
                 double* H, GmresDev* s, double rtol, double atol, size_t max_lds);
void gmres_dense_set_max_lds(size_t bytes);
bool gmres_dense_fits(int nf, int kmax, size_t max_lds);
// The dense solve over P workgroups of one XCD (rows of S and of the basis
// on chip, two grid barriers per step); plan: P and rows per workgroup R,
// false if no split fits.  ws holds gmres_coop_ws doubles.  Status as
// gmres_dense; reason -3: a grid barrier timed out.  xcd: the XCC id whose
// workgroups run the solve (nxcd XCDs on the device), or -1 for workgroups
// 0, 8, 16, ...
bool gmres_coop_plan(int nf, int kmax, size_t max_lds, int* P, int* R);
size_t gmres_coop_ws(int nf, int kmax, int P);
void gmres_coop(hipStream_t st, int nf, int kmax, int maxit, int P, int R, const double* S, const double* b, double* x,
                GmresDev* s, double rtol, double atol, double* ws, int xcd, int nxcd);
// The dense solve over P workgroups of one XCD for nf <= 1024 and up to 1024
// Krylov vectors (S rows in registers, basis rows in LDS up to NL vectors,
// the rest in HBM; three grid barriers per step at most); plan: P, rows per
// workgroup R and NL, false if no split fits.  ws holds gmres_coopw_ws doubles.
// One solve of a fused launch of cooperative solves (gmres_coopw's
// arguments); gmres_coopw_multi runs n <= min(8, nxcd) of them at once,
// solve k on XCD xcds[k], provided gmres_coopw_multi_ok holds between their
// nf (the same register layout).
struct CoopWJob {
  int nf, kmax, maxit, P, R, NL;
  const double* S;
  const double* b;
  double* x;
  GmresDev* s;
  double rtol, atol;
  double* ws;
};
bool gmres_coopw_multi_ok(int nf_a, int nf_b);
void gmres_coopw_multi(hipStream_t st, int n, const CoopWJob* jobs, const int* xcds, int nxcd);
// Family-3 bodies: a deformable plate as a linear Euler-Bernoulli beam of
// ne = (npts - 1) / 2 Hermite elements, clamped at the first point; nd = 2
// ne unknowns (normal deflection and slope per free node).  Matrices
// row-major: M, C, K, K_eff = 4/dt^2 M + 2/dt C + K, their inverses where
// named, H (np x nd) the shape functions at the points, G = H K_eff^-1 H^T
// (np x np); n = (nx, ny) the normal of the deck's line.
struct BeamHost {
  int np = 0, ne = 0, nd = 0;
  double nx = 0, ny = 1;
  std::vector<double> M, C, K, Keff, Keffi, Minv, H, G, X0, Y0;
};
struct BeamDev {
  int off, np, nd;  // entries [off, off + 2 np) of the force vector
  double nx, ny;
  const double *M, *C, *K, *Keff, *Keffi, *Minv, *H, *G, *X0, *Y0;
  double *u, *ud, *udd, *u0, *ud0, *udd0;
};
bool beam_host_setup(const double* x, const double* y, int npts, double mass, double EI, double damp, double dt,
                     BeamHost* out);
// rhsf += delta Phi (ud0 + 2/dt (u0 - u) - 2/dt K_eff^-1 z)  (compute_rhs)
void beam_rhs(hipStream_t st, const BeamDev& b, double dt, double delta, double* rhsf);
// the Newmark update for the force f and the new point positions; *du_abs
// gets max |du|  (compute_dx + update_xb)
void beam_dx(hipStream_t st, const BeamDev& b, double dt, double delta, const double* f, double* xb, double* yb,
             double* du_abs);
// udd = M^-1 (delta/dt Phi^T f - C ud - K u)  (initial_acc, step 1)
void beam_initial_acc(hipStream_t st, const BeamDev& b, double dt, double delta, const double* f);
// (u0, ud0, udd0) = (u, ud, udd) at the start of a step
void beam_save0(hipStream_t st, const BeamDev& b);
// S += 2 delta^2/dt^2 Phi K_eff^-1 Phi^T on the beam's entries (b_times2)
void beam_add_schur(hipStream_t st, const BeamDev& b, int nf, double dt, double delta, double* S);
// The XCC ids of the device's XCDs (sorted), found once per device.
std::vector<int> gmres_xcc_ids(hipStream_t st);
bool gmres_coopw_plan(int nf, int kmax, size_t max_lds, int ncu, int* P, int* R, int* NL);
size_t gmres_coopw_ws(int nf, int kmax, int P, int R, int NL);
void gmres_coopw(hipStream_t st, int nf, int kmax, int maxit, int P, int R, int NL, const double* S, const double* b,
                 double* x, GmresDev* s, double rtol, double atol, double* ws, int xcd, int nxcd);
// pressure post-processing (src/operators_pressure.F90): cell arrays are
// m x n, i fastest; pressure boundary vectors [left n | right n | bottom m |
// top m]
void pressure_corner(hipStream_t st, const Geom& g, const double* w, const double* qx, const double* qy, double q0x,
                     double* oqx, double* oqy);
void pressure_edges(hipStream_t st, const Geom& g, const double* oqx, const double* oqy, const double* ubc,
                    const double* vbc, double del, double* nlx, double* nly);
void pressure_div(hipStream_t st, const Geom& g, const double* hx, const double* hy, double* out, bool accumulate);
void pressure_coarsify(hipStream_t st, const Geom& g, const double* fine, double* coarse);
void pressure_bc1(hipStream_t st, const Geom& g, const double* qx, const double* qy, double q0x, double del,
                  double* pbc);
void pressure_bc2(hipStream_t st, const Geom& g, const double* pc, double* pbc);
void pressure_applybc(hipStream_t st, const Geom& g, double* r, const double* pbc);
void pressure_final(hipStream_t st, const Geom& g, const double* p, const double* qx, const double* qy, double q0x,
                    double del, double* out);
void scale_div(hipStream_t st, int n, const double* x, double d, double* y);
// In-place Cholesky A = U^T U of the upper triangle of a column-major n x n
// matrix (err[0] = 1 if a pivot is not positive), and A^-1 from the factor.
void chol_factor(hipStream_t st, int n, double* A, int* err);
void chol_inverse(hipStream_t st, int n, const double* U, double* Ainv, size_t max_lds);

}  // namespace ibg
