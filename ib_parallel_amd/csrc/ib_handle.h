// ib_handle.h -- one simulation's device-resident state and its time step.
//
// The step mirrors src/timestepper.F90:18-556 for sub_domain = F:
//   1. predictor on every level, coarse to fine (:100-236): fused
//      nonlinear+rot+apply_bc+AB2 kernel, then x pass (2 arrays), fused
//      y-DST / lam1 / lam1i / y-DST column pass, x pass;
//   2. vort2flux (:239-245) and the FSI fixed point (:249-456) with
//      device-resident GMRES on the Schur operator a_times2 + b_times2, or
//      the num_stat solve with the precomputed inverse (:457-484);
//   3. corrector (:488-553): reg, rot, ainv with the subtraction fused into its
//      last row pass, vort2flux.
// All state lives in HBM for the handle's lifetime; the host reads back only
// GMRES/FSI status words.  Included by the C ABI (ib_solver.hip) and by the
// batched form (ib_batch.hip).
#pragma once

#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/ib_gpu.h"
#include "ib_kernels.h"

namespace ibg {

// A host-side range over a phase of the step, seen by `rocprofv3
// --marker-trace` (a call and nothing more when no tool is attached).
// Ranges cover the host's enqueue and waits: a phase whose work runs
// asynchronously ends when its last launch is queued.
struct Range {
  explicit Range(const char* name) { roctxRangePush(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t n_) {
    release();
    n = n_;
    IB_HIP(hipMalloc(&p, sizeof(T) * (n > 0 ? n : 1)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
  operator T*() const { return p; }
};

struct KTimer {
  bool on = false;
  int64_t count = 0;
  double total_ms = 0;
  double total_bytes = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, spare;
  hipEvent_t start_ev = nullptr;
  void begin(hipStream_t st, double bytes) {
    if (!on) return;
    total_bytes += bytes;
    std::pair<hipEvent_t, hipEvent_t> e;
    if (!spare.empty()) {
      e = spare.back();
      spare.pop_back();
    } else {
      IB_HIP(hipEventCreate(&e.first));
      IB_HIP(hipEventCreate(&e.second));
    }
    IB_HIP(hipEventRecord(e.first, st));
    pending.push_back(e);
  }
  void end(hipStream_t st) {
    if (!on) return;
    IB_HIP(hipEventRecord(pending.back().second, st));
  }
  void harvest() {
    for (auto& e : pending) {
      float ms = 0;
      IB_HIP(hipEventElapsedTime(&ms, e.first, e.second));
      total_ms += ms;
      ++count;
      spare.push_back(e);
    }
    pending.clear();
  }
  void destroy() {
    for (auto& e : pending) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    for (auto& e : spare) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    pending.clear();
    spare.clear();
  }
};

// input/sd_precomputed.chd, the sub-domain's B as preprocess_sd writes it
// (src/timestepper.F90:690-695,725-728) and reads it back when the deck sets
// sub_domain_precomputed (:753-766): Fortran unformatted sequential records
// (4-byte length markers before and after each, little-endian), per column
// j = 1..nf_sd one record with its count nz, then one with its nz values and
// their nz row numbers (1-based entries of the sub-domain).  The integers
// are PetscInt: 4 bytes in PETSc's default build, 8 with 64-bit indices; the
// reader takes the size from the first record's length.  The path is
// relative to the working directory, as in the reference;
// IB_SD_PRECOMPUTED_FILE names another.
inline std::string sd_chd_path() {
  const char* e = std::getenv("IB_SD_PRECOMPUTED_FILE");
  return (e && *e) ? std::string(e) : std::string("input/sd_precomputed.chd");
}

// the file is written where its folder exists (a case directory's input)
inline bool sd_chd_writable(const std::string& path) {
  const size_t k = path.find_last_of('/');
  const std::string dir = k == std::string::npos ? std::string(".") : (k == 0 ? std::string("/") : path.substr(0, k));
  struct stat sb;
  return ::stat(dir.c_str(), &sb) == 0 && S_ISDIR(sb.st_mode);
}

inline bool sd_write_chd(const std::string& path, const std::vector<long>& kptr, const std::vector<int>& krow,
                         const std::vector<double>& kval, std::string* why) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    *why = std::strerror(errno);
    return false;
  }
  bool ok = true;
  auto put = [&](const void* p, size_t n) { ok = ok && (n == 0 || std::fwrite(p, 1, n, f) == n); };
  std::vector<int32_t> idx;
  for (size_t k = 0; k + 1 < kptr.size(); ++k) {
    const long a = kptr[k], nz = kptr[k + 1] - kptr[k];
    const int32_t m1 = 4, n32 = (int32_t)nz, m2 = (int32_t)(12 * nz);
    put(&m1, 4);
    put(&n32, 4);
    put(&m1, 4);
    idx.resize(nz);
    for (long j = 0; j < nz; ++j) idx[j] = krow[a + j] + 1;
    put(&m2, 4);
    put(kval.data() + a, 8 * (size_t)nz);
    put(idx.data(), 4 * (size_t)nz);
    put(&m2, 4);
  }
  ok = std::fclose(f) == 0 && ok;
  if (!ok) *why = "write error";
  return ok;
}

inline bool sd_read_chd(const std::string& path, int ncol, std::vector<long>& kptr, std::vector<int>& krow,
                        std::vector<double>& kval, std::string* why) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    *why = std::strerror(errno);
    return false;
  }
  auto bad = [&](const std::string& m) {
    *why = m;
    std::fclose(f);
    return false;
  };
  auto rd = [&](void* p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; };
  kptr.assign((size_t)ncol + 1, 0);
  krow.clear();
  kval.clear();
  std::vector<double> v;
  std::vector<int64_t> r;
  std::vector<int32_t> r4;
  std::vector<std::pair<int64_t, double>> col;
  for (int k = 0; k < ncol; ++k) {
    const std::string at = "column " + std::to_string(k + 1) + " of " + std::to_string(ncol);
    int32_t m1 = 0, e1 = 0;
    if (!rd(&m1, 4)) return bad(at + ": file ends");
    if (m1 != 4 && m1 != 8) return bad(at + ": a count record of " + std::to_string(m1) + " bytes");
    int64_t nz = 0;
    if (m1 == 4) {
      int32_t t = 0;
      if (!rd(&t, 4)) return bad(at + ": file ends");
      nz = t;
    } else if (!rd(&nz, 8)) {
      return bad(at + ": file ends");
    }
    if (!rd(&e1, 4) || e1 != m1) return bad(at + ": record markers do not match");
    if (nz < 0 || nz > ncol) return bad(at + ": count " + std::to_string(nz));
    int32_t m2 = 0, e2 = 0;
    if (!rd(&m2, 4) || (int64_t)m2 != nz * (8 + m1)) return bad(at + ": the value record does not hold nz entries");
    v.resize(nz);
    r.resize(nz);
    if (!rd(v.data(), 8 * (size_t)nz)) return bad(at + ": file ends");
    if (m1 == 4) {
      r4.resize(nz);
      if (!rd(r4.data(), 4 * (size_t)nz)) return bad(at + ": file ends");
      for (int64_t j = 0; j < nz; ++j) r[j] = r4[j];
    } else if (!rd(r.data(), 8 * (size_t)nz)) {
      return bad(at + ": file ends");
    }
    if (!rd(&e2, 4) || e2 != m2) return bad(at + ": record markers do not match");
    col.resize(nz);
    for (int64_t j = 0; j < nz; ++j) {
      if (r[j] < 1 || r[j] > ncol) return bad(at + ": row " + std::to_string(r[j]) + " outside 1.." + std::to_string(ncol));
      col[j] = {r[j] - 1, v[j]};
    }
    std::stable_sort(col.begin(), col.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (const auto& e : col) {
      krow.push_back((int)e.first);
      kval.push_back(e.second);
    }
    kptr[k + 1] = (long)krow.size();
  }
  char extra;
  if (std::fread(&extra, 1, 1, f) == 1) return bad("more data after the " + std::to_string(ncol) + " columns");
  std::fclose(f);
  return true;
}

// Simulations stepped at once from several host threads whose cooperative
// GMRES solves run together, one XCD each (ib_batch): solve() takes over
// member's solve and returns once it is on the device, with st waiting for
// it; false when the group is not stepping (the caller solves alone).
struct SolveGroup {
  virtual ~SolveGroup() = default;
  virtual bool solve(int member, const CoopWJob& job, hipStream_t st) = 0;
};

}  // namespace ibg

using namespace ibg;

struct ib_handle {
  ib_params prm{};
  std::string err;
  int dev = 0;
  hipStream_t st = nullptr;
  size_t max_lds = 65536;
  Geom g{};
  int L = 0;
  double delta = 0, dt = 0, Re = 0, offx = 0, offy = 0, normalize = 0;
  HostFFTPlan px, py;
  DstTiles tiles;  // form and tile shapes of the DST passes (dst_tiles)
  DenseMat dmx, dmy;  // sine matrices of the matrix-core DST form (while a class uses it)
  DBuf<double> tq;    // kMaxBatch fields: the matrix-core column pass's intermediate
  DBuf<double> laminv, lam1, lam1i;
  std::vector<double> vfac, con1, con2, q0x;
  DBuf<double> omega, s, qx, qy, rhs_old;
  DBuf<double> t0, t1, t2;
  DBuf<double> vv, sn, vx, vy, hx, hy;
  DBuf<double> lastbc, obc, sbcA, sbcB, zbc;
  // bodies
  int nb = 0, nf = 0, nbodies = 0, ntors = 0;
  std::vector<int> h_cbody, h_ccomp, h_tors_of;
  std::vector<char> h_moving;  // per entry: a torsional or beam body's
  // family-3 bodies: host matrices, force-vector offsets, device views into beam_mem
  int nbeam = 0;
  std::vector<BeamHost> h_beam;
  std::vector<int> h_beam_off;
  std::vector<BeamDev> beams;
  DBuf<double> beam_mem;
  int nmoving() const { return ntors + nbeam; }
  std::vector<TorsBody> h_tb;
  DBuf<double> xb, yb, f, f_rdst, w, rhsf, gw, opx;
  DBuf<int> comp, tors_of, err_flag;
  RegCSR csr;
  DBuf<TorsBody> tb;
  DBuf<double> th, dabs;
  // GMRES
  int kmax = 30;
  DBuf<double> V, H, cc, ss, grs, nrs, pnorm;
  GmresDev* gst = nullptr;
  // the dense solve over several workgroups (gmres_coop): workgroups, rows
  // per workgroup, workspace (barrier words, partials, w, Hessenberg copies)
  int coopP = 0, coopR = 0;
  // the same for larger systems (gmres_coopw): workgroups, rows per
  // workgroup, basis vectors held in LDS
  int cwP = 0, cwR = 0, cwNL = 0;
  int last_xmode = 0;
  // set by a concurrent batch whose cooperative solves would not all fit on
  // one XCD's CUs together: each solve takes the lock from its launch until
  // its work has completed, so only one is resident at a time
  std::mutex* coop_lock = nullptr;
  // the XCD (XCC id) whose CUs run this handle's cooperative solves, of the
  // device's XCDs xcc; handles take the XCDs in turn as they are created (a
  // batch assigns its own).  -1: workgroups 0, 8, 16, ... (no XCD ids known)
  int coop_xcd = -1;
  std::vector<int> xcc;
  SolveGroup* group = nullptr;  // set by a batch whose solves run together
  int group_member = 0;
  static unsigned next_xcd() {
    static std::atomic<unsigned> n{0};
    return n.fetch_add(1u);
  }
  std::unique_lock<std::mutex> coop_turn() {
    return (coop_lock && (coopP > 0 || cwP > 0)) ? std::unique_lock<std::mutex>(*coop_lock)
                                                 : std::unique_lock<std::mutex>();
  }  // exchange of the last wide cooperative solve (1 one XCD's L2, 2 device scope)
  DBuf<double> coopws;
  GmresDev* hst = nullptr;  // pinned
  double* hbuf = nullptr;   // pinned readback
  int* hflag = nullptr;     // pinned
  int itime = 0, ksp_iter = 0, fsi_iter = 0;
  double err_fsi = 0;
  KTimer tim[5];
  // captured GMRES batches, keyed by (first basis column, batch length)
  std::map<std::pair<int, int>, hipGraphExec_t> gm_graphs;
  // dense Schur matrix and the column-batch buffers that assemble it
  bool assembled = false;
  static constexpr int kAutoAssembleMax = 8192;  // schur_mode 0 assembles up to this many force entries
  int ncol = 0;  // columns per assembly batch
  DBuf<double> Sd, Vb, T0b, T1b, Sb, BCa, BCb;
  // face-space form: face set, its index map, G = Q on the set, T = G W^T
  bool facemode = false;
  static constexpr int kFaceMarginDefault = 4;
  int face_margin = kFaceMarginDefault;
  static constexpr size_t kAutoGreenMax = (size_t)1 << 28;  // schur_mode 0: G up to 2 GiB
  int nF = 0;
  int64_t green_builds = 0;
  DBuf<int> Fkeys, fmap, tapidx, face_flag;
  DBuf<double> Gm, Tm;
  // the face-space S formed again after a weight rebuild: the entries of
  // torsional bodies move, the others keep their weights; tcache: T's
  // columns and S's block of the fixed entries are those of the current G,
  // face set and fixed weights
  DBuf<int> movidx, rigidx;
  int nmov = 0, nrig = 0;
  bool tcache = false;
  // numerically stationary path (num_stat): the inverse of the Cholesky-
  // factored a_times matrix, and a torsional map with no torsional entries
  DBuf<double> Ainv;
  DBuf<int> no_tors;
  // sub-domain approach (sub_domain = T): points, their weights and tap
  // list, B by columns after the drop tolerance, and P with its transpose
  bool sd_on = false;
  ib_subdomain sdc{};
  int sd_np = 4, sd_nx = 0, sd_ny = 0, nr_sd = 0, nb_sd = 0, nf_sd = 0, sd_used = 0;
  std::vector<double> sd_x, sd_y;
  std::vector<int> h_fam, h_npts;  // per body
  DBuf<double> w_sd, rsd, h0sd;
  RegCSR csr_sd;  // csr_sd.key: the face of each sub-domain tap
  DBuf<long> bptr;
  DBuf<int> brow;
  DBuf<double> bval;
  int64_t sd_nnz_eval = 0, sd_nnz_kept = 0;
  DBuf<int> d_pidx, d_used, d_uptr, d_uj;
  DBuf<double> d_pw, d_uw;
  // discrete adjoint (allocated on first use): d/dRe of the spectral tables
  // and of the num_stat matrix, checkpoints, one segment's tape, adjoint
  // state and scratch
  bool adj_ready = false;
  DBuf<double> dlam1i, ll, dA, cw;  // d lam1i/dRe, lam1 lam1i, dA/dRe, objective weights
  DBuf<double> aw, aro, aqx, aqy, aws, af, arf, af2, racc, raccb;
  DBuf<double> aP1, aP2, aD1, aD2, aD3, asb, aFx, aFy, anlx, anly, aub, avb;
  DBuf<double> albb, albk, aobk, asbcb, abct, adot, adJ;
  DBuf<double> c_w, c_ro, c_qx, c_qy, t_w, t_qx, t_qy, t_ws, t_f, fin_w, fin_ro, fin_qx, fin_qy;
  DBuf<double> fin_s, fin_f, fin_fr;
  DBuf<double> cur_w, cur_ws, cur_qx, cur_qy, cur_f, cur_qt;  // the tape slot being reversed (graph replays)
  double* tape_ws = nullptr;  // set: advance copies the predicted omega (all levels) here
  double* tape_qt = nullptr;  // set: advance gathers q* + q0 at the taps here
  // body-point positions (rigid): d w/d xb, d w/d yb, G W^T and W G on the
  // faces the taps touch, per-entry gradient accumulators, (q* + q0) tape
  bool adj_pos = false, want_pos = false;
  DBuf<double> dwx, dwy, TTm, Um, gpx, gpy, t_qt;
  // moving (torsional) bodies in the face-space form: S^T, dG/dRe, face-set
  // vectors, the per-body columns r1 and Z, the body-state tape and
  // checkpoints, and the body adjoints (host; a few scalars per body)
  bool adj_fsi = false;
  bool adj_gm = false;  // rigid bodies, GMRES on the fixed assembled S
  double* tape_th = nullptr;  // set: advance copies th, xb, yb after the FSI loop here (and +6nt, +nf)
  DBuf<double> SdT, dGm, xfv, vTv, yfv, zfv, lfv, dgv, uTv, r1v, Zv, aav, zcv, htv, dthv, cvv, hx2, hy2, fmask, scal;
  DBuf<double> t_th, t_f0, c_th, c_f, fin_th;
  int64_t dG_build = -1;
  std::vector<double> bth, bthd, bthdd, gK, gG0, h_scal;
  // hinge and point-spacing gradients per body of the last adjoint run with
  // position gradients; gDS: the spacing's structural share per torsional
  // body, gdsw its weights' share (device), psx/psy a step's point partials
  std::vector<double> hinge_gx, hinge_gy, hinge_gds, gDS;
  bool hinge_ok = false;
  DBuf<double> psx, psy, gdsw;
  // pressure post-processing (allocated on first use): cell-grid DST plans
  // and table, corner products (two levels), rhs, solution, scratch
  bool pres_ready = false;
  Geom gp{};  // the m x n cells seen by the DST kernels as an (m+1) x (n+1) node grid
  HostFFTPlan pxp, pyp;
  DBuf<double> laminv_ddti, poq, prhs, ppl, pout, ptmp, ptmp2, ppbc, pubc, pvbc, pfdt;

  // ------------------------------------------------------------------ util
  void sync() {
    IB_HIP(hipStreamSynchronize(st));
    for (auto& t : tim) t.harvest();
  }

  // algorithmic bytes of one DST pass launch: arrays read + tables + arrays written
  // algorithmic HBM bytes of one DST pass: every field array of the batch
  // read and written once, spectral tables read once per launch
  double pass_bytes(int nread, int ntab, int nwrite, int batch) const {
    return 8.0 * (double)g.P * ((double)(nread + nwrite) * (double)batch + (double)ntab);
  }
  // timing classes: 0/1 row/column passes of the time step, 2/3 the batched
  // passes that assemble the Schur matrix
  static int tclass(int base, const DstLaunch& a) { return base + (a.strided && a.batch > kMaxBatch ? 2 : 0); }
  void xpass(int mode, const DstLaunch& a) {
    const int cls = tclass(0, a);
    KTimer& t = tim[cls];
    t.begin(st, pass_bytes(mode == X_SUBFROM ? 2 : (mode == X_POINTROT ? 0 : 1), 0, 1, a.batch));
    const int form = cls == 0 ? tiles.xform : tiles.xform_b;
    if (form > 0 && mode != X_POINTROT) dst_dense_x(st, dmx, mode, a, g, form - 1);
    else dst_x(st, px.dev, mode, a, g, tiles.rp, max_lds);
    t.end(st);
  }
  void ypass(int mode, const DstLaunch& a) {
    const int cls = tclass(1, a);
    KTimer& t = tim[cls];
    t.begin(st, pass_bytes(mode == Y_PRED ? 2 : 1, mode == Y_PRED ? 2 : (mode == Y_SCALE2 ? 1 : 0), 1, a.batch));
    const int form = cls == 1 ? tiles.yform : tiles.yform_b;
    if (form > 0) dst_dense_y(st, dmy, mode, a, g, form - 1, tq, (int)(tq.n / (size_t)g.P));
    else dst_y(st, py.dev, mode, a, g, tiles.cp, max_lds);
    t.end(st);
  }

  // s = ctci(w [+ apply_bc(bc)]): src/myfft.F90:177-188
  void ctci_dev(const double* win, double* out, const double* bc) {
    DstLaunch a;
    a.in[0] = win; a.out[0] = t0; a.bc = bc;
    xpass(bc ? X_APPLYBC : X_PLAIN, a);
    DstLaunch b;
    b.in[0] = t0; b.out[0] = t1; b.tab = laminv; b.div = 1.0;
    ypass(Y_SCALE2, b);
    DstLaunch c;
    c.in[0] = t1; c.out[0] = out;
    xpass(X_PLAIN, c);
  }

  // ainv: src/myfft.F90:161-173; with subfrom, target -= ainv(v) (corrector).
  // v == nullptr: the input is rot(hx, hy) (rot_int + rot_bc), formed in the
  // load of the first row pass.
  void ainv_dev(const double* v, double* out, bool subfrom) {
    DstLaunch a;
    a.in[0] = v; a.out[0] = t0;
    a.hx = hx; a.hy = hy;
    xpass(v ? X_PLAIN : X_ROT, a);
    DstLaunch b;
    b.in[0] = t0; b.out[0] = t1; b.tab = lam1i; b.div = normalize;
    ypass(Y_SCALE2, b);
    DstLaunch c;
    c.in[0] = t1; c.out[0] = out;
    xpass(subfrom ? X_SUBFROM : X_PLAIN, c);
  }

  // vort2flux: src/operators_fluid.F90:237-377
  void vort2flux_dev(double* wv, double* sv, double* qxv, double* qyv) {
    const long P = g.P;
    for (int k = 1; k < L; ++k) coarsify(st, g, wv + (k - 1) * P, wv + k * P);
    const double* sbc = zbc;
    for (int k = L - 1; k >= 0; --k) {
      ctci_dev(wv + k * P, sv + k * P, k == L - 1 ? nullptr : sbc);
      double* bout = nullptr;
      if (k > 0) bout = (sbc == sbcA) ? sbcB.p : sbcA.p;
      curl_getbc(st, g, sv + k * P, sbc, qxv + (long)k * g.Fx, qyv + (long)k * g.Fy, bout);
      sbc = bout;
    }
  }

  // ab_times = a_times2 + b_times2: src/operators_fluid.F90:14-120.
  // hx/hy are zero outside the faces the current weights touch and reg
  // rewrites every touched face, so they need no clearing per product; the
  // coarse levels of vv are zero outside the centres coarsify rewrites.
  void schur_dev(const double* x, double* y) {
    reg(st, g, csr, w, x, hx, hy);
    ainv_dev(nullptr, vv, false);
    vort2flux_dev(vv, sn, vx, vy);
    regT(st, g, nf, w, csr.key, vx, vy, 0.0, nullptr, y);
    body_btimes(st, ntors, tb, th, dt, delta, x, y);
  }

  // Sd and the column-batch buffers for batches of up to `need` columns;
  // each column needs L + 3 node arrays and two bc vectors, and a batch is
  // held to about 4 GiB of HBM (larger column counts run in several batches).
  void alloc_assembly(int need) {
    if (nf == 0) return;
    if (!Sd.p) Sd.alloc((size_t)nf * nf);
    if (Vb.p) return;
    const size_t per_col = sizeof(double) * ((size_t)(L + 3) * g.P + 2 * (size_t)g.nbc);
    const size_t batch_budget = (size_t)4 << 30;
    ncol = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(need, 1), batch_budget / per_col));
    ncol = std::min(ncol, 65535);  // grid y limit of the batched launches
    Vb.alloc((size_t)ncol * L * g.P);
    T0b.alloc((size_t)ncol * g.P);
    T1b.alloc((size_t)ncol * g.P);
    Sb.alloc((size_t)ncol * g.P);
    BCa.alloc((size_t)ncol * g.nbc);
    BCb.alloc((size_t)ncol * g.nbc);
    fill(st, Vb, (long)Vb.n, 0.0);  // coarse-level margins stay zero (coarsify rewrites only centres)
  }

  // The column-batch buffers are only needed while columns are formed;
  // forms whose matrix never changes give them back after set-up.
  bool keeps_columns() const { return !sd_on && (facemode || (assembled && nmoving() > 0)); }
  void trim_columns() {
    if (keeps_columns() || !Vb.p) return;
    sync();
    Vb.release(); T0b.release(); T1b.release(); Sb.release(); BCa.release(); BCb.release();
    ncol = 0;
  }

  // One batch of nc operator columns through ainv and the multi-level
  // vort2flux.  Column c of the batch starts from rot(E^T e_(c0+c)) (w given:
  // the 49 taps of force entry c0+c) or from rot of the unit face source
  // keys[c0+c] (w null).  Leaves the level-1 streamfunctions in Sb (stride P)
  // and returns their boundary vectors (stride nbc, or zbc with stride 0 when
  // there is a single level).
  const double* column_batch(int c0, int nc, const double* wts, const int* keys, long* bc_stride,
                             const double* ainv_tab = nullptr) {
    const long P = g.P, nbc = g.nbc, LP = (long)L * P;
    DstLaunch a;
    a.batch = nc; a.strided = true; a.out[0] = T0b; a.s_out = P;
    a.c0 = c0; a.w = wts; a.key = keys;  // the input is formed from the weights or the face list
    xpass(X_POINTROT, a);
    DstLaunch b;
    b.batch = nc; b.strided = true; b.in[0] = T0b; b.s_in = P; b.out[0] = T1b; b.s_out = P;
    b.tab = ainv_tab ? ainv_tab : lam1i;
    b.div = normalize;
    ypass(Y_SCALE2, b);
    DstLaunch c;
    c.batch = nc; c.strided = true; c.in[0] = T1b; c.s_in = P; c.out[0] = Vb; c.s_out = LP;
    xpass(X_PLAIN, c);
    for (int k = 1; k < L; ++k) coarsify(st, g, Vb + (k - 1) * P, Vb + k * P, nc, LP);
    double* bcur = nullptr;  // boundary from the next coarser level (none on the coarsest)
    for (int k = L - 1; k >= 0; --k) {
      DstLaunch d;
      d.batch = nc; d.strided = true; d.in[0] = Vb + k * P; d.s_in = LP; d.out[0] = T0b; d.s_out = P;
      if (bcur) { d.bc = bcur; d.s_bc = nbc; }
      xpass(bcur ? X_APPLYBC : X_PLAIN, d);
      DstLaunch e;
      e.batch = nc; e.strided = true; e.in[0] = T0b; e.s_in = P; e.out[0] = T1b; e.s_out = P; e.tab = laminv;
      e.div = 1.0;
      ypass(Y_SCALE2, e);
      DstLaunch f2;
      f2.batch = nc; f2.strided = true; f2.in[0] = T1b; f2.s_in = P; f2.out[0] = Sb; f2.s_out = P;
      xpass(X_PLAIN, f2);
      if (k > 0) {
        double* bnext = (bcur == BCa.p) ? BCb.p : BCa.p;
        get_bc(st, g, Sb, P, bnext, nbc, nc, 1.0);
        bcur = bnext;
      }
    }
    *bc_stride = bcur ? nbc : 0;
    return bcur ? bcur : zbc.p;
  }

  // Dense Schur matrix S(:, j) = ab_times(e_j) for every force entry j, in
  // batches of ncol columns: ainv and the multi-level vort2flux run on all
  // columns of a batch at once, and the level-1 flux is evaluated at the taps
  // (regT) together with b_times2.  Every launch covers ncol arrays, so the
  // kernels are throughput-bound instead of latency-bound.
  void assemble_schur() {
    alloc_assembly(nf);
    for (int c0 = 0; c0 < nf; c0 += ncol) {
      const int nc = std::min(ncol, nf - c0);
      long sbc = 0;
      const double* bc = column_batch(c0, nc, w, csr.key, &sbc);
      schur_columns(st, g, nf, c0, nc, w, csr.key, Sb, g.P, bc, sbc, tors_of, tb, th, dt, delta, Sd);
    }
  }

  // ---- face-space form (moving bodies): the level-1 flux operator Q =
  // C (multi-level CtC)^-1 ainv C^T does not depend on the body, only E does.
  // G = Q restricted to the faces within face_margin cells of the current
  // taps is computed once by batched unit-source columns; every weight
  // rebuild then forms S = W G W^T + B with two small contractions.  G is
  // recomputed (for a new face set) only when a tap leaves the set.
  // face_grow (set while an adjoint runs, whose reverse sweep returns to
  // earlier body positions): keep every face of the current set and add the
  // new taps' neighbourhoods, so the set is rebuilt once rather than at
  // every change of direction.  S's entries do not depend on the set.
  bool face_grow = false;
  // S = W G W^T + B.  moving: the weights changed only for the torsional
  // bodies' entries since the last call, so with a valid cache only their
  // columns of T and rows and columns of S are formed (the same bits as the
  // full product; IB_SCHUR_FULL=1 forms everything every time)
  void form_schur_face(bool moving) {
    const char* e = std::getenv("IB_SCHUR_FULL");
    const bool full_always = e && std::atoi(e) == 1;
    if (moving && tcache && !full_always && nmov > 0 && nmov < nf) {
      schur_from_green_moving(st, nf, nF, w, tapidx, Gm, Tm, tors_of, tb, th, dt, delta, Sd, nmov, movidx, nrig,
                              rigidx);
    } else {
      schur_from_green(st, nf, nF, w, tapidx, Gm, Tm, tors_of, tb, th, dt, delta, Sd);
      tcache = true;
    }
    // both forms rewrite every row and column of the moving entries
    for (const BeamDev& b : beams) beam_add_schur(st, b, nf, dt, delta, Sd);
  }
  // family 3: the structural parts of the FSI iteration
  void beams_rhs() {
    for (const BeamDev& b : beams) beam_rhs(st, b, dt, delta, rhsf);
  }
  void beams_dx() {
    for (int k = 0; k < nbeam; ++k) beam_dx(st, beams[k], dt, delta, f, xb, yb, dabs + ntors + k);
  }

  void build_face_set() {
    tcache = false;
    // touched faces = unique tap keys of the current weights
    int nr = 0;
    IB_HIP(hipMemcpyAsync(hflag, csr.nruns, sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
    nr = *hflag;
    std::vector<int> uniq(nr);
    if (nr) IB_HIP(hipMemcpy(uniq.data(), csr.uniq, sizeof(int) * nr, hipMemcpyDeviceToHost));
    std::vector<int> fm((size_t)g.Fx + g.Fy, -1);
    std::vector<int> keys;
    if (face_grow && nF > 0 && Fkeys.p) {
      keys.resize(nF);
      IB_HIP(hipMemcpy(keys.data(), Fkeys.p, sizeof(int) * nF, hipMemcpyDeviceToHost));
      for (int q = 0; q < nF; ++q) fm[keys[q]] = q;
    }
    const int r = face_margin;
    for (int key : uniq) {
      const bool xf = key < g.Fx;
      const int ni = xf ? g.m + 1 : g.m, nj = xf ? g.n : g.n + 1, base = xf ? 0 : g.Fx;
      const int i = (key - base) % ni, j = (key - base) / ni;
      for (int dj = -r; dj <= r; ++dj)
        for (int di = -r; di <= r; ++di) {
          const int ii = i + di, jj = j + dj;
          if (ii >= 0 && ii < ni && jj >= 0 && jj < nj && fm[base + (size_t)jj * ni + ii] < 0)
            fm[base + (size_t)jj * ni + ii] = -2;
        }
    }
    for (size_t q = 0; q < fm.size(); ++q)
      if (fm[q] == -2) {
        fm[q] = (int)keys.size();
        keys.push_back((int)q);
      }
    nF = (int)keys.size();
    Fkeys.alloc(std::max(1, nF));
    fmap.alloc(fm.size());
    IB_HIP(hipMemcpy(Fkeys.p, keys.data(), sizeof(int) * nF, hipMemcpyHostToDevice));
    IB_HIP(hipMemcpy(fmap.p, fm.data(), sizeof(int) * fm.size(), hipMemcpyHostToDevice));
  }

  void compute_green() {
    tcache = false;
    if (Gm.n < (size_t)nF * nF) Gm.alloc((size_t)nF * nF);
    if (Tm.n < (size_t)nF * nf) Tm.alloc((size_t)nF * nf);
    for (int c0 = 0; c0 < nF; c0 += ncol) {
      const int nc = std::min(ncol, nF - c0);
      long sbc = 0;
      const double* bc = column_batch(c0, nc, nullptr, Fkeys, &sbc);
      green_columns(st, g, nF, c0, nc, Fkeys, Sb, g.P, bc, sbc, Gm);
    }
    ++green_builds;
  }

  // the Schur matrix for the current weights (and theta): called after every
  // weight rebuild once the stencil check has passed
  void refresh_schur() {
    if (!assembled) return;
    if (!facemode) {
      assemble_schur();
      return;
    }
    IB_HIP(hipMemsetAsync(face_flag.p, 0, sizeof(int), st));
    map_taps(st, nf, csr.key, fmap, tapidx, face_flag);
    IB_HIP(hipMemcpyAsync(hflag, face_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
    if (*hflag) {
      build_face_set();
      compute_green();
      map_taps(st, nf, csr.key, fmap, tapidx, face_flag);
    }
    form_schur_face(false);
  }

  // setup_reg + the sorted tap list; the old touched faces of hx/hy are
  // cleared first (the invariant schur_dev relies on).  With key_changed
  // the sorted list is left to the caller: the kernel sets *key_changed when
  // a tap key moved, and the caller runs csr.build() before the list is read
  // again (the list depends on the keys only)
  void rebuild_weights(int* key_changed = nullptr) {
    fill(st, hx, g.Fx, 0.0);
    fill(st, hy, g.Fy, 0.0);
    setup_reg(st, g, nf, xb, yb, comp, delta, offx, offy, prm.deltafnc_f32_literals, w, csr.key, err_flag,
              key_changed);
    if (!key_changed) csr.build(st, g);
  }

  int read_err_flag() {
    IB_HIP(hipMemcpyAsync(hflag, err_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
    return *hflag;
  }

  void read_gmres() {
    IB_HIP(hipMemcpyAsync(hst, gst, sizeof(GmresDev), hipMemcpyDeviceToHost, st));
    sync();
  }

  // PETSc-style GMRES(kmax) on the Schur operator, x = initial guess / result.
  // Iterations are enqueued in batches of up to kGmBatch between status
  // reads; once the device status holds a reason the remaining Arnoldi
  // kernels of the batch return at once, and the surplus matvecs only write
  // basis columns past the converged one, which the solution never reads.
  static constexpr int kGmBatch = 8;
  // the single-workgroup solve streams the part of S that does not fit in
  // its registers from L2 every iteration; past this size the multi-workgroup
  // matrix-vector product of the graph-replayed iterations is faster
  static constexpr int kFusedGmresMax = 400;

  // nb Arnoldi iterations starting at basis column it0.  Every argument of
  // these launches is fixed for the handle's lifetime except the column
  // pointers, so each (it0, nb) is captured once as a HIP graph and replayed.
  // Kernel timing records events per launch and therefore uses plain
  // launches.
  void apply_schur(const double* x, double* y) {
    if (assembled) gemv(st, nf, Sd, x, y);
    else schur_dev(x, y);
  }
  void gmres_batch_direct(int it0, int nb) {
    for (int q = 0; q < nb; ++q) {
      apply_schur(V + (size_t)(it0 + q) * nf, V + (size_t)(it0 + q + 1) * nf);
      gm_arnoldi(st, nf, kmax, V, gst, H, cc, ss, grs, pnorm);
    }
  }
  void gmres_batch(int it0, int nb) {
    if (prm.launch_mode != 0 || no_capture || tim[0].on || tim[1].on || tim[2].on || tim[3].on || tim[4].on) {
      gmres_batch_direct(it0, nb);
      return;
    }
    auto key = std::make_pair(it0, nb);
    auto itr = gm_graphs.find(key);
    if (itr == gm_graphs.end()) {
      hipGraph_t gr = nullptr;
      hipGraphExec_t ge = nullptr;
      IB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        gmres_batch_direct(it0, nb);
      } catch (...) {
        (void)hipStreamEndCapture(st, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
      }
      IB_HIP(hipStreamEndCapture(st, &gr));
      IB_HIP(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
      IB_HIP(hipGraphDestroy(gr));
      itr = gm_graphs.emplace(key, ge).first;
    }
    IB_HIP(hipGraphLaunch(itr->second, st));
  }
  // the whole solve in one launch on a dense matrix M: several workgroups
  // (gmres_coop) where the split fits, else one workgroup
  bool dense_solve_ok() const {
    if (prm.gmres_form == 3) return false;
    return coopP > 0 || cwP > 0 || (nf <= kFusedGmresMax && gmres_dense_fits(nf, kmax, max_lds));
  }
  void dense_launch(const double* M, const double* b, double* x) {
    if (group && cwP > 0 && coopP == 0) {
      const CoopWJob j{nf, kmax, prm.gmres_max_it, cwP, cwR, cwNL, M, b, x, gst, prm.rtol, prm.atol, coopws};
      if (group->solve(group_member, j, st)) return;
    }
    if (coopP > 0 || cwP > 0) dense_launch_ws(M, b, x, coopws, nullptr);
    else dense_launch_ws(M, b, x, V, H);
  }
  // the same solve with its workspace at ws (the cooperative kernels) or at
  // ws and hw (the one-workgroup kernel's basis and Hessenberg matrix)
  void dense_launch_ws(const double* M, const double* b, double* x, double* ws, double* hw) {
    if (coopP > 0)
      gmres_coop(st, nf, kmax, prm.gmres_max_it, coopP, coopR, M, b, x, gst, prm.rtol, prm.atol, ws, coop_xcd,
                 (int)xcc.size());
    else if (cwP > 0)
      gmres_coopw(st, nf, kmax, prm.gmres_max_it, cwP, cwR, cwNL, M, b, x, gst, prm.rtol, prm.atol, ws, coop_xcd,
                  (int)xcc.size());
    else gmres_dense(st, nf, kmax, prm.gmres_max_it, M, b, x, ws, hw, gst, prm.rtol, prm.atol, max_lds);
  }
  // doubles of workspace dense_launch_ws needs (ws and hw back to back)
  size_t dense_ws_doubles() const {
    if (coopP > 0) return gmres_coop_ws(nf, kmax, coopP);
    if (cwP > 0) return gmres_coopw_ws(nf, kmax, cwP, cwR, cwNL);
    return (size_t)(kmax + 1) * std::max(1, nf) + (size_t)(kmax + 1) * kmax;
  }
  // Operand bytes of a dense GMRES solve of its iterations: per Arnoldi step
  // the matrix (nf^2 doubles), the basis vectors 0..j read twice by
  // classical Gram-Schmidt (dots and update) and the new vector written,
  // with j the step's index in its restart cycle; b and x read, x written.
  // What a solve that streamed its operands from HBM would move; the
  // cooperative solves keep S and the basis on chip.
  double gmres_operand_bytes(int its) const {
    const double n = nf, kc = std::max(kmax, 1);
    const double cyc = std::floor(its / kc), rem = its - cyc * kc;
    const double sum_j1 = cyc * kc * (kc + 1) / 2 + rem * (rem + 1) / 2;  // sum of (j + 1)
    return 8.0 * (its * n * n + 2.0 * n * sum_j1 + its * n + 3.0 * n);
  }
  // the status of a dense solve already copied to hst
  int dense_check() {
    if (cwP > 0 && coopP == 0) last_xmode = hst->xmode;
    if (hst->reason == -3) return fail(IB_E_HIP, "GMRES: a grid barrier of the cooperative solve timed out");
    return hst->reason > 0 ? IB_OK : IB_E_NOCONV_GMRES;
  }
  int dense_status() {
    read_gmres();
    return dense_check();
  }
  int dense_solve(const double* M, const double* b, double* x) {
    auto turn = coop_turn();
    dense_launch(M, b, x);
    return dense_status();
  }
  int gmres_dev(const double* b, double* x, int* its_out) {
    if (assembled && dense_solve_ok()) {
      // dense Schur matrix: the whole solve is one launch.  Algorithmic
      // bytes: S, b read once, x read and written, and one Krylov vector
      // written per iteration (added once the count is known)
      auto turn = coop_turn();
      tim[4].begin(st, 0.0);
      dense_launch(Sd, b, x);
      tim[4].end(st);
      const int rc = dense_status();
      turn = std::unique_lock<std::mutex>();
      if (tim[4].on) tim[4].total_bytes += gmres_operand_bytes(hst->its);
      *its_out = hst->its;
      return rc;
    }
    GmresDev init{};
    init.first = 1;
    *hst = init;
    IB_HIP(hipMemcpyAsync(gst, hst, sizeof(GmresDev), hipMemcpyHostToDevice, st));
    int reason = 0, its = 0;
    const int maxit = prm.gmres_max_it;
    while (true) {
      apply_schur(x, gw);
      gm_start(st, nf, kmax, b, gw, V, gst, H, grs, prm.rtol, prm.atol);
      read_gmres();
      reason = hst->reason;
      if (reason) break;
      int it = 0;
      while (true) {
        const int nb = std::min(kGmBatch, std::min(kmax - it, maxit - its));
        gmres_batch(it, nb);
        read_gmres();
        reason = hst->reason;
        its = hst->its;
        it = hst->it;
        if (reason || it >= kmax || its >= maxit) break;
      }
      gm_buildsol(st, nf, kmax, V, gst, H, grs, x, nrs);
      if (reason) break;
      if (its >= maxit) {
        reason = -2;
        break;
      }
    }
    *its_out = its;
    return reason > 0 ? IB_OK : IB_E_NOCONV_GMRES;
  }

  // num_stat preprocess (src/timestepper.F90:560-598) and the Cholesky
  // factorisation of KSPPREONLY + PCCHOLESKY (src/variables.F90:555-561):
  // column j = a_times(e_j) (no b_times2: every entry is treated as rigid),
  // A = U^T U from its upper triangle, then A^-1 once so each step's solve is
  // a matrix-vector product.
  // Buffers a captured batch graph holds pointers to keep their addresses
  // when the set-up runs again (restart).
  int num_stat_setup() {
    if (!no_tors.p) no_tors.alloc(std::max(nf, 1));
    IB_HIP(hipMemsetAsync(no_tors.p, 0xff, sizeof(int) * std::max(nf, 1), st));  // every entry -1
    alloc_assembly(nf);
    for (int c0 = 0; c0 < nf; c0 += ncol) {
      const int nc = std::min(ncol, nf - c0);
      long sbc = 0;
      const double* bc = column_batch(c0, nc, w, csr.key, &sbc);
      schur_columns(st, g, nf, c0, nc, w, csr.key, Sb, g.P, bc, sbc, no_tors, tb, th, dt, delta, Sd);
    }
    IB_HIP(hipMemsetAsync(err_flag.p, 0, sizeof(int), st));
    chol_factor(st, nf, Sd, err_flag);
    if (read_err_flag()) return fail(IB_E_INVAL, "num_stat: the Schur matrix is not positive definite");
    if (!Ainv.p) Ainv.alloc((size_t)nf * nf);
    chol_inverse(st, nf, Sd, Ainv, max_lds);
    sync();
    trim_columns();
    return IB_OK;
  }

  // ------------------------------------------------ sub-domain approach
  // Set-up of src/grid.F90:320-348 (points), :628-692 (E on the sub-domain),
  // src/variables.F90:425-514 with src/grid.F90:696-751 (which rows of each
  // column are evaluated) and src/timestepper.F90:602-770 (the columns,
  // a_times_sd of a unit entry, and the drop tolerance).
  int sd_setup(const ib_body* bodies) {
    const ib_subdomain& c = sdc;
    if (!(c.delta_sd > 0.0) || !(c.sdxe >= c.sdxi) || !(c.sdye >= c.sdyi))
      return fail(IB_E_INVAL, "sub-domain: need delta_sd > 0 and a non-empty rectangle");
    const int s2n = (int)std::lround(std::sqrt((double)sd_np));
    if (sd_np < 1 || s2n * s2n != sd_np) return fail(IB_E_INVAL, "sub-domain: interp_points must be a square");
    sd_x.clear();
    sd_y.clear();
    nr_sd = 0;
    if (!c.full_rectangular) {
      int nrig = 0;
      for (int b = 0; b < nbodies; ++b)
        if (h_fam[b] == 1) {
          ++nrig;
          sd_x.insert(sd_x.end(), bodies[b].x, bodies[b].x + bodies[b].npts);
          sd_y.insert(sd_y.end(), bodies[b].y, bodies[b].y + bodies[b].npts);
        }
      if (nrig > 1)
        return fail(IB_E_UNSUPPORTED, "sub-domain with body points: the reference includes one rigid body only");
      nr_sd = (int)sd_x.size();
    }
    sd_nx = (int)std::lround((c.sdxe - c.sdxi) / c.delta_sd + 1.0);
    sd_ny = (int)std::lround((c.sdye - c.sdyi) / c.delta_sd + 1.0);
    for (int j = 1; j <= sd_ny; ++j)
      for (int i = 1; i <= sd_nx; ++i) {
        sd_x.push_back(c.sdxi + (i - 1) * c.delta_sd);
        sd_y.push_back(c.sdyi + (j - 1) * c.delta_sd);
      }
    nb_sd = (int)sd_x.size();
    nf_sd = 2 * nb_sd;
    // E on the sub-domain: the body weight kernel over the sub-domain entries
    {
      std::vector<double> ex(nf_sd), ey(nf_sd);
      std::vector<int> ec(nf_sd);
      for (int k = 0; k < nf_sd; ++k) {
        ex[k] = sd_x[k % nb_sd];
        ey[k] = sd_y[k % nb_sd];
        ec[k] = k < nb_sd ? 1 : 2;
      }
      DBuf<double> dx_, dy_;
      DBuf<int> dc_;
      dx_.alloc(nf_sd); dy_.alloc(nf_sd); dc_.alloc(nf_sd);
      IB_HIP(hipMemcpy(dx_.p, ex.data(), sizeof(double) * nf_sd, hipMemcpyHostToDevice));
      IB_HIP(hipMemcpy(dy_.p, ey.data(), sizeof(double) * nf_sd, hipMemcpyHostToDevice));
      IB_HIP(hipMemcpy(dc_.p, ec.data(), sizeof(int) * nf_sd, hipMemcpyHostToDevice));
      w_sd.alloc((size_t)nf_sd * kTaps);
      csr_sd.alloc(nf_sd * kTaps, g.Fx + g.Fy);
      IB_HIP(hipMemsetAsync(err_flag.p, 0, sizeof(int), st));
      setup_reg(st, g, nf_sd, dx_, dy_, dc_, delta, offx, offy, prm.deltafnc_f32_literals, w_sd, csr_sd.key, err_flag);
      if (read_err_flag()) return fail(IB_E_INVAL, "sub-domain point regularisation stencil leaves the grid");
      csr_sd.build(st, g);
    }
    std::vector<long> kptr(nf_sd + 1, 0);
    std::vector<int> krow;
    std::vector<double> kval;
    const std::string chd = sd_chd_path();
    if (c.precomputed) {
      // sub_domain_precomputed: B from the file an earlier run wrote
      // (src/timestepper.F90:753-766); no sparsity windows are formed
      std::string why;
      if (!sd_read_chd(chd, nf_sd, kptr, krow, kval, &why))
        return fail(IB_E_INVAL, "sub-domain: cannot read the precomputed B from " + chd + ": " + why);
      sd_nnz_eval = 0;
    } else {
      const int rc = sd_compute_b(kptr, krow, kval);
      if (rc != IB_OK) return rc;
      // like preprocess_sd (src/timestepper.F90:690-695,725-728): the kept
      // columns go to the file, where a case directory has an input folder
      std::string why;
      if (sd_chd_writable(chd) && !sd_write_chd(chd, kptr, krow, kval, &why))
        return fail(IB_E_INVAL, "sub-domain: cannot write " + chd + ": " + why);
    }
    sd_nnz_kept = (int64_t)krow.size();
    bptr.alloc(nf_sd + 1);
    brow.alloc(std::max<size_t>(krow.size(), 1));
    bval.alloc(std::max<size_t>(kval.size(), 1));
    IB_HIP(hipMemcpy(bptr.p, kptr.data(), sizeof(long) * (nf_sd + 1), hipMemcpyHostToDevice));
    if (!krow.empty()) {
      IB_HIP(hipMemcpy(brow.p, krow.data(), sizeof(int) * krow.size(), hipMemcpyHostToDevice));
      IB_HIP(hipMemcpy(bval.p, kval.data(), sizeof(double) * kval.size(), hipMemcpyHostToDevice));
    }
    rsd.alloc(nf_sd);
    h0sd.alloc(nf_sd);
    d_pidx.alloc((size_t)sd_np * std::max(nf, 1));
    d_pw.alloc((size_t)sd_np * std::max(nf, 1));
    return sd_build_interp();
  }

  // B of the sub-domain (get_sparsity_EEt + preprocess_sd): the rows of
  // each column, the columns by batched unit sources, the drop tolerance;
  // kept entries by column, rows ascending
  int sd_compute_b(std::vector<long>& kptr, std::vector<int>& krow, std::vector<double>& kval) {
    const ib_subdomain& c = sdc;
    // rows of each column: the E E^T windows (half-width sn) of two entries
    // share a column of the window matrix exactly when both entries have the
    // same component and their stencil cells differ by at most 2 sn in x and
    // in y, as long as no window reaches the edge of the grid (checked)
    const int sn = c.support_nz;
    std::vector<int> bix(nb_sd), biy(nb_sd);
    for (int q = 0; q < nb_sd; ++q) {
      bix[q] = (int)((sd_x[q] + offx) / delta) + 1;
      biy[q] = (int)((sd_y[q] + offy) / delta) + 1;
      if (bix[q] - sn - 1 < 0 || bix[q] + sn - 1 > g.m - 1 || biy[q] - sn - 1 < 0 || biy[q] + sn - 1 > g.n - 1)
        return fail(IB_E_UNSUPPORTED, "sub-domain: a sparsity window reaches the edge of the grid");
    }
    std::vector<std::vector<int>> cell((size_t)(g.m + 2) * (g.n + 2));
    for (int q = 0; q < nb_sd; ++q) cell[(size_t)biy[q] * (g.m + 2) + bix[q]].push_back(q);
    std::vector<long> colptr(nf_sd + 1, 0);
    std::vector<std::vector<int>> nbr(nb_sd);
    for (int q = 0; q < nb_sd; ++q) {
      std::vector<int>& N = nbr[q];
      for (int jj = std::max(0, biy[q] - 2 * sn); jj <= std::min(g.n + 1, biy[q] + 2 * sn); ++jj)
        for (int ii = std::max(0, bix[q] - 2 * sn); ii <= std::min(g.m + 1, bix[q] + 2 * sn); ++ii) {
          const std::vector<int>& C = cell[(size_t)jj * (g.m + 2) + ii];
          N.insert(N.end(), C.begin(), C.end());
        }
      std::sort(N.begin(), N.end());
    }
    for (int k = 0; k < nf_sd; ++k) colptr[k + 1] = colptr[k] + 2 * (long)nbr[k % nb_sd].size();
    const long total = colptr[nf_sd];
    std::vector<int> rows(total);
    for (int k = 0; k < nf_sd; ++k) {
      const std::vector<int>& N = nbr[k % nb_sd];
      long o = colptr[k];
      for (int q : N) rows[o++] = q;
      for (int q : N) rows[o++] = q + nb_sd;
    }
    nbr.clear();
    cell.clear();
    // the columns: batched unit sources through ainv + vort2flux, the level-1
    // flux regularised at the listed rows
    DBuf<long> dptr;
    DBuf<int> drows;
    DBuf<double> dvals;
    dptr.alloc(nf_sd + 1);
    drows.alloc(std::max<long>(total, 1));
    dvals.alloc(std::max<long>(total, 1));
    IB_HIP(hipMemcpy(dptr.p, colptr.data(), sizeof(long) * (nf_sd + 1), hipMemcpyHostToDevice));
    IB_HIP(hipMemcpy(drows.p, rows.data(), sizeof(int) * total, hipMemcpyHostToDevice));
    alloc_assembly(nf_sd);
    for (int c0 = 0; c0 < nf_sd; c0 += ncol) {
      const int nc = std::min(ncol, nf_sd - c0);
      long sbc = 0;
      const double* bc = column_batch(c0, nc, w_sd, csr_sd.key, &sbc);
      sd_columns(st, g, c0, nc, colptr[c0 + nc] - colptr[c0], dptr, drows, w_sd, csr_sd.key, Sb, g.P, bc, sbc, dvals);
    }
    std::vector<double> vals(total);
    IB_HIP(hipMemcpyAsync(vals.data(), dvals.p, sizeof(double) * total, hipMemcpyDeviceToHost, st));
    sync();
    // drop tolerance: keep |B(r, j)| > droptol * B(j, j), in row order
    krow.reserve(total / 2);
    kval.reserve(total / 2);
    for (int k = 0; k < nf_sd; ++k) {
      double diag = 0.0;
      for (long a = colptr[k]; a < colptr[k + 1]; ++a)
        if (rows[a] == k) diag = vals[a];
      for (long a = colptr[k]; a < colptr[k + 1]; ++a)
        if (std::fabs(vals[a]) > c.droptol * diag) {
          krow.push_back(rows[a]);
          kval.push_back(vals[a]);
        }
      kptr[k + 1] = (long)krow.size();
    }
    sd_nnz_eval = total;
    return IB_OK;
  }

  // setup_reg_subdomain_indices (src/grid.F90:777-951) for the current body
  // positions: P (np sub-domain entries and weights per force entry,
  // bilinear hat weights, or the entry itself for the rigid body's points
  // when they are sub-domain points) and its transpose over the used
  // entries, for P^T f in ascending force order
  int sd_build_interp() {
    const ib_subdomain& c = sdc;
    std::vector<double> hxb(nf), hyb(nf);
    IB_HIP(hipMemcpyAsync(hxb.data(), xb.p, sizeof(double) * nf, hipMemcpyDeviceToHost, st));
    IB_HIP(hipMemcpyAsync(hyb.data(), yb.p, sizeof(double) * nf, hipMemcpyDeviceToHost, st));
    sync();
    const int s = (int)std::lround(std::sqrt((double)sd_np));
    const int i1 = (sd_np % 2 == 0) ? s / 2 + 1 - s : (s - 1) / 2 + 1 - s;
    std::vector<int> pidx((size_t)sd_np * nf), nn(sd_np), ll(s), pp(s);
    std::vector<double> pw((size_t)sd_np * nf), dist(sd_np), wv(sd_np);
    for (int j = 0; j < nf; ++j) {
      const int comp = h_ccomp[j];
      const double indx = (hxb[j] - c.sdxi) / c.delta_sd + 1.0, indy = (hyb[j] - c.sdyi) / c.delta_sd + 1.0;
      for (int a = 0; a < s; ++a) {
        ll[a] = (sd_np % 2 == 0 ? (int)indx : (int)std::lround(indx)) + i1 + a;
        pp[a] = (sd_np % 2 == 0 ? (int)indy : (int)std::lround(indy)) + i1 + a;
      }
      if (!c.full_rectangular && h_fam[h_cbody[j]] == 1) {
        const int j1 = j + 1, nr = h_npts[h_cbody[j]];
        for (int i = 0; i < sd_np; ++i) nn[i] = comp == 1 ? j1 + 1 : j1 + 1 - nr;
        nn[0] = comp == 1 ? j1 : j1 - nr;
        for (int i = 0; i < sd_np; ++i) dist[i] = 1.0;
        dist[0] = 0.0;
      } else {
        int i = 0;
        for (int a = 0; a < s; ++a)
          for (int b = 0; b < s; ++b) {
            if (ll[a] < 1) return fail(IB_E_INVAL, "Body has exited the subdomain from left");
            if (ll[a] > sd_nx) return fail(IB_E_INVAL, "Body has exited the subdomain from right");
            if (pp[b] < 1) return fail(IB_E_INVAL, "Body has exited the subdomain from bottom");
            if (pp[b] > sd_ny) return fail(IB_E_INVAL, "Body has exited the subdomain from top");
            nn[i] = nr_sd + (pp[b] - 1) * sd_nx + ll[a];
            const double dx = hxb[j] - sd_x[nn[i] - 1], dy = hyb[j] - sd_y[nn[i] - 1];
            dist[i] = std::sqrt(dx * dx + dy * dy);
            ++i;
          }
      }
      bool exact = false;
      for (int i = 0; i < sd_np; ++i) {
        const int q = nn[i] - 1;
        if (q < 0 || q >= nb_sd) return fail(IB_E_INVAL, "sub-domain: interpolation point out of range");
        const double hx_ = std::fabs(sd_x[q] - hxb[j]) / c.delta_sd, hy_ = std::fabs(sd_y[q] - hyb[j]) / c.delta_sd;
        const double fxv = (hx_ <= 1.0 ? 1.0 - hx_ : 0.0) / c.delta_sd, fyv = (hy_ <= 1.0 ? 1.0 - hy_ : 0.0) / c.delta_sd;
        wv[i] = c.delta_sd * c.delta_sd * fxv * fyv;
        if (dist[i] == 0.0) {
          wv[i] = 1.0;
          exact = true;
        }
      }
      for (int i = 0; i < sd_np; ++i) {
        if (exact && dist[i] != 0.0) wv[i] = 0.0;
        pidx[(size_t)j * sd_np + i] = nn[i] - 1 + (comp == 1 ? 0 : nb_sd);
        pw[(size_t)j * sd_np + i] = wv[i];
      }
    }
    // transpose over the used entries (the last of equal indices of an entry wins)
    std::vector<int> used(pidx);
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    std::vector<std::vector<std::pair<int, double>>> tr(used.size());
    for (int j = 0; j < nf; ++j) {
      for (int i = 0; i < sd_np; ++i) {
        const int q = (int)(std::lower_bound(used.begin(), used.end(), pidx[(size_t)j * sd_np + i]) - used.begin());
        auto& T = tr[q];
        if (!T.empty() && T.back().first == j) T.back().second = pw[(size_t)j * sd_np + i];
        else T.push_back({j, pw[(size_t)j * sd_np + i]});
      }
    }
    std::vector<int> uptr(used.size() + 1, 0), uj;
    std::vector<double> uw;
    for (size_t q = 0; q < used.size(); ++q) {
      for (auto& e : tr[q]) {
        uj.push_back(e.first);
        uw.push_back(e.second);
      }
      uptr[q + 1] = (int)uj.size();
    }
    sd_used = (int)used.size();
    d_used.alloc(std::max<size_t>(used.size(), 1));
    d_uptr.alloc(used.size() + 1);
    d_uj.alloc(std::max<size_t>(uj.size(), 1));
    d_uw.alloc(std::max<size_t>(uw.size(), 1));
    IB_HIP(hipMemcpyAsync(d_pidx.p, pidx.data(), sizeof(int) * pidx.size(), hipMemcpyHostToDevice, st));
    IB_HIP(hipMemcpyAsync(d_pw.p, pw.data(), sizeof(double) * pw.size(), hipMemcpyHostToDevice, st));
    IB_HIP(hipMemcpyAsync(d_used.p, used.data(), sizeof(int) * used.size(), hipMemcpyHostToDevice, st));
    IB_HIP(hipMemcpyAsync(d_uptr.p, uptr.data(), sizeof(int) * uptr.size(), hipMemcpyHostToDevice, st));
    if (!uj.empty()) {
      IB_HIP(hipMemcpyAsync(d_uj.p, uj.data(), sizeof(int) * uj.size(), hipMemcpyHostToDevice, st));
      IB_HIP(hipMemcpyAsync(d_uw.p, uw.data(), sizeof(double) * uw.size(), hipMemcpyHostToDevice, st));
    }
    sync();  // the host vectors go out of scope
    return IB_OK;
  }

  // lumat for the current P and theta into Sd (the matrix GMRES runs on)
  void sd_matrix() { sd_lumat(st, nf, sd_np, d_pidx, d_pw, bptr, brow, bval, tors_of, tb, th, dt, delta, Sd); }

  // step 2: FSI fixed point (src/timestepper.F90:249-456).  Returns the
  // step status; IB_E_INVAL when a body stencil leaves the grid.
  int fsi_fixed_point() {
    int status = IB_OK;
    int fsi_count = 0;
    double err_glob = 10.0;
    ksp_iter = 0;
    if (ntors > 0) {
      if (itime == 1) body_initial_acc(st, ntors, tb, th, dt, delta, f);
      IB_HIP(hipMemcpyAsync(th + 3 * ntors, th.p, sizeof(double) * 3 * ntors, hipMemcpyDeviceToDevice, st));
    }
    for (const BeamDev& b : beams) {
      if (itime == 1) beam_initial_acc(st, b, dt, delta, f);
      beam_save0(st, b);
    }
    // moving bodies in the face-space form with a dense solve: the GMRES
    // status, the stencil check of the new weights, the face-set check of the
    // new taps and |d theta| come back in one host synchronisation per
    // iteration
    const bool one_sync = nmoving() > 0 && !sd_on && facemode && dense_solve_ok();
    while (err_glob >= prm.fsi_tol && fsi_count <= prm.fsi_max_count) {
      Range iter_range("fsi iteration");
      ++fsi_count;
      int its = 0;
      int rc = IB_OK;
      if (one_sync) {
        fsi_rhs(st, g, nf, tors_of, tb, ntors, th, dt, delta, w, csr.key, qx, qy, q0x[0], rhsf);
        beams_rhs();
        auto turn = coop_turn();
        tim[4].begin(st, 0.0);
        dense_launch(Sd, rhsf, f);
        tim[4].end(st);
        redist_faces(st, g, csr, w, f, hx, hy);
        regT(st, g, nf, w, csr.key, hx, hy, 0.0, nullptr, f_rdst);
        body_dx(st, ntors, tb, th, dt, delta, f, xb, yb, dabs);
        beams_dx();
        // face_flag[0]: a tap outside the face set; [1]: a tap key moved (the
        // sorted tap list is rebuilt after the synchronisation, only then;
        // the next GMRES solve does not read it)
        IB_HIP(hipMemsetAsync(face_flag.p, 0, 2 * sizeof(int), st));
        rebuild_weights(face_flag.p + 1);
        map_taps(st, nf, csr.key, fmap, tapidx, face_flag);
        IB_HIP(hipMemcpyAsync(hbuf, dabs.p, sizeof(double) * nmoving(), hipMemcpyDeviceToHost, st));
        IB_HIP(hipMemcpyAsync(hflag, err_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
        IB_HIP(hipMemcpyAsync(hflag + 1, face_flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        IB_HIP(hipMemcpyAsync(hst, gst, sizeof(GmresDev), hipMemcpyDeviceToHost, st));
        // S for the new weights goes on the stream before the host waits for
        // the flags, so it runs while the host wakes up; it is formed again
        // below if a tap has left the face set
        IB_HIP(hipEventRecord(ev_flags, st));
        form_schur_face(true);
        {
          Range w_range("wait: solve and body update");
          IB_HIP(hipEventSynchronize(ev_flags));
        }
        turn = std::unique_lock<std::mutex>();
        for (auto& t : tim) t.harvest();  // the timed launches all precede the event
        if (hflag[2]) csr.build(st, g);
        rc = dense_check();
        its = hst->its;
        if (tim[4].on) tim[4].total_bytes += gmres_operand_bytes(its);
        ksp_iter += its;
        if (rc != IB_OK) status = rc;
        if (hflag[0]) {
          err = "body point regularisation stencil leaves the grid";
          return IB_E_INVAL;
        }
        if (hflag[1]) {
          build_face_set();  // synchronises the stream before G is reallocated
          compute_green();
          map_taps(st, nf, csr.key, fmap, tapidx, face_flag);
          form_schur_face(false);
        }
        err_glob = 0.0;
        for (int t = 0; t < nmoving(); ++t) err_glob = std::max(err_glob, hbuf[t]);
        continue;
      }
      if (!sd_on) {
        fsi_rhs(st, g, nf, tors_of, tb, ntors, th, dt, delta, w, csr.key, qx, qy, q0x[0], rhsf);
        beams_rhs();
        rc = gmres_dev(rhsf, f, &its);
        ksp_iter += its;
      } else {
        // src/timestepper.F90:368-405: rhsf = compute_rhs + P E_sd (q + q0),
        // lumat for the current P, GMRES on it (its iterations are not
        // added to ksp_iter in the reference)
        fsi_rhs(st, g, nf, tors_of, tb, ntors, th, dt, delta, w, nullptr, qx, qy, q0x[0], rhsf);
        regT(st, g, nf_sd, w_sd, csr_sd.key, qx, qy, q0x[0], nullptr, rsd);
        sd_interp(st, nf, sd_np, d_pidx, d_pw, rsd, rhsf);
        sd_matrix();
        rc = gmres_dev(rhsf, f, &its);
      }
      if (rc != IB_OK) status = rc;
      // redistribute (output only)
      redist_faces(st, g, csr, w, f, hx, hy);
      regT(st, g, nf, w, csr.key, hx, hy, 0.0, nullptr, f_rdst);
      if (ntors > 0 && sd_on) {
        // the body's E stays as set up (the reference updates only P on the
        // sub-domain, src/timestepper.F90:449-450)
        body_dx(st, ntors, tb, th, dt, delta, f, xb, yb, dabs);
        IB_HIP(hipMemcpyAsync(hbuf, dabs.p, sizeof(double) * ntors, hipMemcpyDeviceToHost, st));
        const int ri = sd_build_interp();
        if (ri != IB_OK) return ri;
        err_glob = 0.0;
        for (int t = 0; t < ntors; ++t) err_glob = std::max(err_glob, hbuf[t]);
      } else if (nmoving() > 0) {
        body_dx(st, ntors, tb, th, dt, delta, f, xb, yb, dabs);
        beams_dx();
        rebuild_weights();
        IB_HIP(hipMemcpyAsync(hbuf, dabs.p, sizeof(double) * nmoving(), hipMemcpyDeviceToHost, st));
        if (read_err_flag()) {
          err = "body point regularisation stencil leaves the grid";
          return IB_E_INVAL;
        }
        refresh_schur();
        err_glob = 0.0;
        for (int t = 0; t < nmoving(); ++t) err_glob = std::max(err_glob, hbuf[t]);
      } else {
        err_glob = 0.0;  // rigid bodies only: positions and weights unchanged
      }
    }
    if (err_glob >= prm.fsi_tol && status == IB_OK) status = IB_E_NOCONV_FSI;
    fsi_iter = fsi_count;
    err_fsi = err_glob;
    return status;
  }

  // ddti (src/myfft.F90:192-203): out = S(laminv_ddti S(in)) on the cells
  void ddti_dev(const double* in, double* out) {
    DstLaunch a;
    a.in[0] = in; a.out[0] = ptmp;
    dst_x(st, pxp.dev, X_PLAIN, a, gp, (pxp.dev.L <= 512 ? 2 : 1), max_lds);
    DstLaunch b;
    b.in[0] = ptmp; b.out[0] = ptmp2; b.tab = laminv_ddti; b.div = 1.0;
    dst_y(st, pyp.dev, Y_SCALE2, b, gp, 1, max_lds);
    DstLaunch c;
    c.in[0] = ptmp2; c.out[0] = out;
    dst_x(st, pxp.dev, X_PLAIN, c, gp, (pxp.dev.L <= 512 ? 2 : 1), max_lds);
  }

  void pressure_setup() {
    if (pres_ready) return;
    const int m = g.m, n = g.n;
    gp = g;
    gp.m = m + 1;
    gp.n = n + 1;
    gp.P = m * n;
    pxp.init(m);
    pyp.init(n);
    if (dst_lds_bytes(pxp.dev, 1) > max_lds || dst_lds_bytes(pyp.dev, 2) > max_lds)
      throw HipError("grid too large for the single-workgroup pressure DST (LDS)");
    // setup_fft's pressure table (src/myfft.F90:123,150-155)
    const double pi = 4.0 * std::atan(1.0);
    const double norm_ddti = 4.0 * (double)(float)((m + 1) * (n + 1));
    std::vector<double> h((size_t)m * n);
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= m; ++i) {
        const double lam =
            2.0 * (std::cos(pi * (double)i / (double)(m + 1)) + std::cos(pi * (double)j / (double)(n + 1)) - 2.0);
        h[(size_t)(i - 1) + (size_t)m * (j - 1)] = 1.0 / lam / norm_ddti;
      }
    const size_t C = (size_t)m * n;
    laminv_ddti.alloc(C);
    IB_HIP(hipMemcpy(laminv_ddti.p, h.data(), sizeof(double) * C, hipMemcpyHostToDevice));
    poq.alloc(4 * (size_t)g.P);
    prhs.alloc(C * L); ppl.alloc(C * L); pout.alloc(C * L); ptmp.alloc(C); ptmp2.alloc(C);
    ppbc.alloc(2 * (size_t)(m + n)); pubc.alloc(g.nbc); pvbc.alloc(g.nbc); pfdt.alloc(std::max(nf, 1));
    pres_ready = true;
  }

  // calculate_pressure (src/operators_pressure.F90:110-437) at the current
  // state into pout (mgridlev blocks of m*n pressure coefficients)
  void pressure_dev() {
    pressure_setup();
    const long P = g.P, C = (long)g.m * g.n;
    // nonlinear terms on the faces of every level, coarse to fine; the bc of
    // level k comes from the corner products of level k+1
    double* oq[2][2] = {{poq.p, poq.p + P}, {poq.p + 2 * P, poq.p + 3 * P}};
    int cur = 0;
    for (int k = L - 1; k >= 0; --k) {
      const double* ubc = zbc.p;
      const double* vbc = zbc.p;
      if (k < L - 1) {
        get_bc(st, g, oq[1 - cur][0], P, pubc, g.nbc, 1, 1.0);
        get_bc(st, g, oq[1 - cur][1], P, pvbc, g.nbc, 1, 1.0);
        ubc = pubc;
        vbc = pvbc;
      }
      pressure_corner(st, g, omega + k * P, qx + (long)k * g.Fx, qy + (long)k * g.Fy, q0x[k], oq[cur][0],
                      oq[cur][1]);
      pressure_edges(st, g, oq[cur][0], oq[cur][1], ubc, vbc, delta * std::pow(2.0, (double)k),
                     vx + (long)k * g.Fx, vy + (long)k * g.Fy);
      cur = 1 - cur;
    }
    // rhs: -div(E^T f/dt) on level 1, coarsified outward, plus div N(q)
    scale_div(st, nf, f, dt, pfdt);
    reg(st, g, csr, w, pfdt, hx, hy);  // hx, hy are zero outside the touched faces
    pressure_div(st, g, hx, hy, prhs, false);
    for (int k = 1; k < L; ++k) {
      fill(st, prhs + k * C, C, 0.0);
      pressure_coarsify(st, g, prhs + (k - 1) * C, prhs + k * C);
    }
    for (int k = 0; k < L; ++k) pressure_div(st, g, vx + (long)k * g.Fx, vy + (long)k * g.Fy, prhs + k * C, true);
    // D D^T p = rhs, coarse to fine
    for (int k = L - 1; k >= 0; --k) {
      if (k == L - 1) {
        const double dk = delta * std::pow(2.0, (double)k);
        pressure_bc1(st, g, qx + (long)k * g.Fx, qy + (long)k * g.Fy, q0x[k], 0.125 / (dk * dk), ppbc);
      } else {
        pressure_bc2(st, g, ppl + (k + 1) * C, ppbc);
      }
      pressure_applybc(st, g, prhs + k * C, ppbc);
      ddti_dev(prhs + k * C, ppl + k * C);
    }
    for (int k = 0; k < L; ++k) {
      const double dk = delta * std::pow(2.0, (double)k);
      pressure_final(st, g, ppl + k * C, qx + (long)k * g.Fx, qy + (long)k * g.Fy, q0x[k], 0.125 / (dk * dk),
                     pout + k * C);
    }
  }

  // ---------------------------------------------------------------- advance
  // step 1: the predictor on every level, coarse to fine (vort2flux follows)
  void predict() {
    const long P = g.P, nbc = g.nbc;
    // lastbc(k) = 0.25 get_bc(omega_{k+1}^n) (src/timestepper.F90:103-117)
    if (L > 1) get_bc(st, g, omega + P, P, lastbc, nbc, L - 1, 0.25);
    fill(st, lastbc + (L - 1) * nbc, nbc, 0.0);
    for (int k = L - 1; k >= 0; --k) {
      const double* ob = zbc;
      if (k < L - 1) {
        get_bc(st, g, omega + (k + 1) * P, P, obc, nbc, 1, 0.25);
        ob = obc;
      }
      nl_rhs(st, g, omega + k * P, qx + (long)k * g.Fx, qy + (long)k * g.Fy, q0x[k], lastbc + k * nbc, ob, vfac[k],
             con1[k], con2[k], itime == 1, rhs_old + k * P, t2, nullptr);
      DstLaunch a;
      a.batch = 2;
      a.in[0] = t2; a.out[0] = t0;
      a.in[1] = omega + k * P; a.out[1] = t1;
      xpass(X_PLAIN, a);
      DstLaunch b;
      b.in[0] = t0; b.in2[0] = t1; b.out[0] = t2; b.tab = lam1 + k * P; b.tab2 = lam1i + k * P;
      ypass(Y_PRED, b);
      DstLaunch c;
      c.in[0] = t2; c.out[0] = omega + k * P;
      xpass(X_PLAIN, c);
    }
  }

  // step 2 for a numerically stationary body (src/timestepper.F90:457-484):
  // rhsf = regT(q + q0), f = A^-1 rhsf, redistribute; nothing moves
  void stat_solve() {
    fsi_rhs(st, g, nf, no_tors, tb, 0, th, dt, delta, w, csr.key, qx, qy, q0x[0], rhsf);
    gemv(st, nf, Ainv, rhsf, f);
    redist_faces(st, g, csr, w, f, hx, hy);
    regT(st, g, nf, w, csr.key, hx, hy, 0.0, nullptr, f_rdst);
  }

  // step 3: corrector (src/timestepper.F90:488-553), then vort2flux; on the
  // sub-domain the source is E_sd^T P^T f (:500-510)
  void correct() {
    if (sd_on) {
      fill(st, hx, g.Fx, 0.0);
      fill(st, hy, g.Fy, 0.0);
      fill(st, h0sd, nf_sd, 0.0);
      sd_ptf(st, sd_used, d_used, d_uptr, d_uj, d_uw, f, h0sd);
      reg(st, g, csr_sd, w_sd, h0sd, hx, hy);
      ainv_dev(nullptr, omega, true);
      // leave hx, hy zero outside the body's faces, as the other forms expect
      fill(st, hx, g.Fx, 0.0);
      fill(st, hy, g.Fy, 0.0);
    } else {
      reg(st, g, csr, w, f, hx, hy);
      ainv_dev(nullptr, omega, true);
    }
    vort2flux_dev(omega, s, qx, qy);
  }

  // The parts of a step whose launches never change for the handle (every
  // pointer and scalar fixed; the Euler/AB2 switch gives two versions) are
  // captured once as HIP graphs and replayed: the predictor, the vort2flux
  // after it, the corrector with its vort2flux, and on the num_stat path
  // without an adjoint tape the whole step.  Plain launches with
  // launch_mode 1 and while kernel timing records per-launch events; the
  // adjoint's tape copies go between the graphs.
  std::map<int, hipGraphExec_t> step_graphs;
  hipEvent_t ev_flags = nullptr;  // the FSI loop's status copies (one_sync path)
  // set for the simulations of a batch, which run on several host threads
  // at once: no stream capture, so that none is open while another thread
  // allocates or copies synchronously
  bool no_capture = false;
  bool step_graphs_ok() const {
    bool timing = false;
    for (const auto& t : tim) timing = timing || t.on;
    return prm.launch_mode == 0 && !timing && !sd_on && !no_capture;
  }
  template <class F>
  void replay(int key, F&& body) {
    auto itr = step_graphs.find(key);
    if (itr == step_graphs.end()) {
      hipGraph_t gr = nullptr;
      hipGraphExec_t ge = nullptr;
      IB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        body();
      } catch (...) {
        (void)hipStreamEndCapture(st, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
      }
      IB_HIP(hipStreamEndCapture(st, &gr));
      IB_HIP(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
      IB_HIP(hipGraphDestroy(gr));
      itr = step_graphs.emplace(key, ge).first;
    }
    IB_HIP(hipGraphLaunch(itr->second, st));
  }

  int advance() {
    Range step_range("ib_advance");
    itime += 1;
    const bool graphs = step_graphs_ok();
    const int first = itime == 1 ? 1 : 0;
    const bool taping = tape_ws || tape_qt || tape_th;
    if (graphs && prm.num_stat && !taping) {
      replay(4 + first, [&] {
        predict();
        vort2flux_dev(omega, s, qx, qy);
        stat_solve();
        correct();
      });
      ksp_iter = 0;
      fsi_iter = 0;
      err_fsi = 0.0;
      sync();
      return IB_OK;
    }
    {
      Range r("predictor + vort2flux");
      if (graphs) replay(first, [&] { predict(); });
      else predict();
      if (tape_ws) copy(st, tape_ws, omega, (long)L * g.P);
      if (graphs) replay(3, [&] { vort2flux_dev(omega, s, qx, qy); });
      else vort2flux_dev(omega, s, qx, qy);
    }

    int status = IB_OK;
    if (prm.num_stat) {
      // no iteration counts are kept on this path
      stat_solve();
      ksp_iter = 0;
      fsi_iter = 0;
      err_fsi = 0.0;
    } else {
      Range r("fsi loop");
      status = fsi_fixed_point();
      if (status == IB_E_INVAL) return status;
    }

    // adjoint tape: q* + q0 at the final taps, the body state after the loop
    if (tape_qt) gather_taps(st, g, nf * kTaps, csr.key, qx, qy, q0x[0], tape_qt);
    if (tape_th) {
      copy(st, tape_th, th, 6 * (long)ntors);
      copy(st, tape_th + 6 * ntors, xb, nf);
      copy(st, tape_th + 6 * ntors + nf, yb, nf);
    }

    Range r("corrector + wait");
    if (graphs) replay(2, [&] { correct(); });
    else correct();
    sync();
    return status;
  }

  // ------------------------------------------------------ discrete adjoint
  // (no reference counterpart; defined in ib_adjoint_host.hip)
  void adjoint_setup(int nsteps, int K);
  // G's derivative in Re on the current face set (columns with d lam1i/dRe)
  void compute_green_d();
  // the adjoint's solves with S^T: x = S^-T b by the LU factors of S when
  // lu_factor has run for the current S, else GMRES on S^T (SdT) from x
  // (zeroed first unless warm) with the deck's tolerances
  int solve_T(const double* b, double* x, bool warm = false);
  // S = P L U (rocSOLVER getrf, partial pivoting) of the current Sd
  void lu_factor();
  void lu_release();
  void* lu_blas = nullptr;  // rocblas_handle on st
  DBuf<double> LUd;
  DBuf<int> lu_piv, lu_info;
  bool lu_ok = false;
  // the body at the end of a taped step's FSI loop: theta (and its history),
  // positions, weights, S (and S^T), the weight derivatives
  void body_restore(const double* bs);
  // set-up of the position gradient (once per handle): the weight
  // derivatives, G on the faces the taps touch (the face set with margin 0),
  // and the two contractions T = G W^T, U = W G
  void adjoint_pos_setup();
  // HBM bytes of one taped step and of one checkpoint
  size_t tape_bytes() const { return sizeof(double) * ((size_t)2 * L * g.P + (size_t)L * (g.Fx + g.Fy) + nf); }
  size_t ckpt_bytes() const { return sizeof(double) * ((size_t)2 * L * g.P + (size_t)L * (g.Fx + g.Fy)); }
  // ckpt_every <= 0: tape the whole run when it fits in half of the free
  // HBM, else checkpoints every ceil(sqrt(nsteps)) steps
  int adjoint_segment(int nsteps, int K);
  void save_state(double* w_, double* ro_, double* qx_, double* qy_);
  void load_state(const double* w_, const double* ro_, const double* qx_, const double* qy_);
  // moving bodies: th, positions and f (GMRES's initial guess) with the flow
  void save_body(double* bs, double* f_);
  void load_body(const double* bs, const double* f_);
  // one taped step: the step-start omega and q, the predicted omega, f
  int taped_advance(int sl);
  // transpose of vort2flux: wb (all levels) and qxb, qyb (all levels) are the
  // adjoints of its outputs; on return wb is the adjoint of its input
  // omega, and qxb, qyb are zero
  void v2f_adj(double* wb, double* qxb, double* qyb);
  // one step backwards; tape slot sl holds the step-start omega and q, the
  // predicted omega and f; (aw, aro, aqx, aqy) enter as the adjoint of the
  // step-end state and leave as that of the step-start state
  int adjoint_step(int sl, bool first);
  // The FSI fixed point of a step with moving (torsional) bodies, backwards.
  // At convergence (f, theta) solve
  //   R1 = S(theta) f - BR(theta) sc - E(theta)(q* + q0) = 0
  //   R2 = Q - theta + c2 BR(theta)^T f = 0        (one per torsional body)
  // (compute_rhs, GMRES, compute_dx; at consistent rates compute_rhs's
  // structural factor sc depends on the history only).  Given the adjoints
  // of f (af) and of theta (the body adjoints carried on the host plus the
  // corrector's E(theta) term) the transposed system is solved through
  // S^T: a = S^-T af, Z_t = S^-T c2 BR_t, then the small system for
  // lambda2 and lambda1 = a - Z lambda2 (left in arf).  History, stiffness
  // and pitch adjoints follow from R1, R2 and the Newmark updates of
  // compute_dx; on the first step also from initial_acc.
  int fsi_adjoint(int sl, bool first, const double* f0, const double* qt);
  int adjoint_run(int nsteps, int Kin, const double* cx, const double* cy, double* J, double* dRe, double* dw0,
                  double* dxb, double* dyb, double* dK, double* dg0);

  // ------------------------------------------------------------------ setup
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }

  int init(const ib_params* p, int nbod, const ib_body* bodies, const ib_subdomain* sd = nullptr) {
    prm = *p;
    if (prm.gmres_restart <= 0) prm.gmres_restart = 30;
    if (prm.gmres_max_it <= 0) prm.gmres_max_it = 10000;
    if (prm.fsi_max_count <= 0) prm.fsi_max_count = 1000;
    const int m = prm.m, n = prm.n;
    L = prm.mgridlev;
    if (m < 8 || n < 8 || m % 4 || n % 4) return fail(IB_E_INVAL, "grid must be divisible by 4 (m, n >= 8)");
    if (L < 1) return fail(IB_E_INVAL, "mgridlev must be >= 1");
    if (prm.sub_domain && !sd) return fail(IB_E_UNSUPPORTED, "sub_domain = T needs ib_create_subdomain");
    if (sd && !prm.sub_domain) return fail(IB_E_INVAL, "sub-domain parameters given with sub_domain = F");
    if (sd && prm.num_stat) return fail(IB_E_INVAL, "num_stat and sub_domain exclude each other");
    if (prm.istart < 0) return fail(IB_E_INVAL, "istart must be >= 0");
    if (prm.gmres_restart > 1000) return fail(IB_E_INVAL, "gmres_restart must be <= 1000");
    kmax = prm.gmres_restart;
    dt = prm.dt; Re = prm.re; offx = prm.offsetx; offy = prm.offsety;
    delta = prm.len / (double)m;
    g.m = m; g.n = n; g.L = L;
    g.P = (m - 1) * (n - 1); g.Fx = (m + 1) * n; g.Fy = m * (n + 1);
    g.nbc = 2 * (m + 1) + 2 * (n + 1);
    g.left = 0; g.right = n + 1; g.bottom = 2 * (n + 1); g.top = 2 * (n + 1) + m + 1;

    // bodies (src/grid.F90:239-433)
    if (nbod < 0) return fail(IB_E_INVAL, "negative body count");
    bool seen_tors = false;
    std::vector<int> nbb(1, 0);
    std::vector<double> hxb, hyb;
    for (int b = 0; b < nbod; ++b) {
      const ib_body& B = bodies[b];
      if (B.npts < 2) return fail(IB_E_INVAL, "a body needs at least 2 points");
      if (B.fam < 1 || B.fam > 3)
        return fail(IB_E_UNSUPPORTED, "body families are 1 (rigid), 2 (torsional) and 3 (beam)");
      if (B.fam == 1 && seen_tors) return fail(IB_E_INVAL, "rigid bodies must precede torsional bodies (src/grid.F90:1226)");
      if (B.fam == 3) {
        seen_tors = true;
        if (B.npts < 3 || B.npts % 2 == 0) return fail(IB_E_INVAL, "a beam needs an odd number of points >= 3");
        if (!(B.itheta > 0) || !(B.ktheta > 0)) return fail(IB_E_INVAL, "a beam needs mass (itheta) and stiffness (ktheta) > 0");
        if (prm.num_stat || sd) return fail(IB_E_UNSUPPORTED, "beams (family 3) need the GMRES path without num_stat or sub-domains");
        BeamHost bh;
        if (!beam_host_setup(B.x, B.y, B.npts, B.itheta, B.ktheta, B.ctheta, dt, &bh))
          return fail(IB_E_INVAL, "beam matrices are singular");
        h_beam.push_back(std::move(bh));
        h_beam_off.push_back(nbb.back());
      }
      if (B.fam == 2) {
        seen_tors = true;
        if (B.itheta <= 0) return fail(IB_E_INVAL, "torsional body needs itheta > 0");
        TorsBody t{};
        t.off = nbb.back();
        t.npts = B.npts;
        t.gamma0 = B.gamma0;
        t.deltas = std::sqrt((B.x[1] - B.x[0]) * (B.x[1] - B.x[0]) + (B.y[1] - B.y[0]) * (B.y[1] - B.y[0]));
        t.xbf0 = B.x[0]; t.ybf0 = B.y[0];
        t.M = B.itheta; t.K = B.ktheta; t.C = B.ctheta;
        h_tb.push_back(t);
      }
      for (int j = 0; j < 2 * B.npts; ++j) {
        const int pnt = j < B.npts ? j : j - B.npts;
        h_cbody.push_back(b);
        h_ccomp.push_back(j < B.npts ? 1 : 2);
        h_tors_of.push_back(B.fam == 2 ? (int)h_tb.size() - 1 : -1);
        h_moving.push_back(B.fam >= 2 ? 1 : 0);
        hxb.push_back(B.x[pnt]);
        hyb.push_back(B.y[pnt]);
      }
      nbb.push_back(nbb.back() + 2 * B.npts);
      nb += B.npts;
      h_fam.push_back(B.fam);
      h_npts.push_back(B.npts);
    }
    nbodies = nbod;
    nf = 2 * nb;
    ntors = (int)h_tb.size();
    nbeam = (int)h_beam.size();

    IB_HIP(hipGetDevice(&dev));
    IB_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    IB_HIP(hipEventCreateWithFlags(&ev_flags, hipEventDisableTiming));
    int lds = 0, optin = 0;
    IB_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess) optin = 0;
    (void)hipGetLastError();
    max_lds = (size_t)std::max(lds, optin);
    if (max_lds > 65536) {
      dst_set_max_lds(max_lds);
      gmres_dense_set_max_lds(max_lds);
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) ncu = 0;
    (void)hipGetLastError();
    if (prm.gmres_form < 0 || prm.gmres_form > 4) return fail(IB_E_INVAL, "gmres_form must be 0 to 4");
    coopP = coopR = 0;
    cwP = cwR = cwNL = 0;
    xcc = gmres_xcc_ids(st);
    if (xcc.empty() || xcc.size() > 8) xcc.clear();  // no usable ids: workgroups 0, 8, 16, ...
    coop_xcd = xcc.empty() ? -1 : xcc[next_xcd() % xcc.size()];
    const bool coop_ok = prm.gmres_form == 0 || prm.gmres_form == 2 || prm.gmres_form == 4;
    if (coop_ok && prm.gmres_form != 4 && nf > 0 && gmres_coop_plan(nf, kmax, max_lds, &coopP, &coopR) &&
        8 * coopP <= ncu)
      coopws.alloc(gmres_coop_ws(nf, kmax, coopP));
    else if (coop_ok && nf > 0 && gmres_coopw_plan(nf, kmax, max_lds, ncu, &cwP, &cwR, &cwNL)) {
      coopP = coopR = 0;
      // IB_COOPW_LDS_VECTORS=k caps the basis vectors kept in LDS (the rest
      // go to HBM); the results do not depend on it
      if (const char* e = std::getenv("IB_COOPW_LDS_VECTORS")) {
        const int cap = std::atoi(e);
        if (cap >= 2 && cap < cwNL) cwNL = cap;
      }
      coopws.alloc(gmres_coopw_ws(nf, kmax, cwP, cwR, cwNL));
    } else {
      coopP = coopR = 0;
      cwP = cwR = cwNL = 0;
    }

    px.init(m - 1);
    py.init(n - 1);
    if (dst_lds_bytes(px.dev, 1) > max_lds || dst_lds_bytes(py.dev, 2) > max_lds)
      return fail(IB_E_UNSUPPORTED, "grid too large for the single-workgroup DST (LDS)");
    // the matrix-core form is timed beside the Stockham passes; its
    // matrices and intermediate stay only if a pass class takes it
    dmx.init(m - 1);
    dmy.init(n - 1);
    tq.alloc((size_t)kMaxBatch * g.P);
    tiles = dst_tiles(st, px.dev, py.dev, &dmx, &dmy, tq, g, max_lds);
    if (tiles.xform == 0 && tiles.xform_b == 0) dmx.release();
    if (tiles.yform == 0 && tiles.yform_b == 0) {
      dmy.release();
      tq.release();
    }

    // setup_fft tables: src/myfft.F90:262-289
    const double pi = 4.0 * std::atan(1.0);
    normalize = 4.0 * (double)(float)(m * n);
    const double del2 = delta * delta;
    vfac.resize(L); con1.resize(L); con2.resize(L); q0x.resize(L);
    for (int k = 0; k < L; ++k) {
      const double del22 = del2 * std::pow(4.0, (double)k);
      vfac[k] = 0.5 * dt / Re / del22 / normalize;
      con1[k] = 1.5 * dt / del22 / normalize;
      con2[k] = -0.5 * dt / del22 / normalize;
      // motion_potential with motion_grid = (-1, 0, ...): q0x = -fac * (-1) (src/variables.F90:299-314)
      const double fac = delta * std::pow(2.0, (double)k);
      q0x[k] = -fac * (-1.0) + 0.0;
    }
    const long P = g.P;
    std::vector<double> hl(P), hl1((size_t)P * L), hl1i((size_t)P * L);
    for (int j = 2; j <= n; ++j)
      for (int i = 2; i <= m; ++i) {
        const double lam = -2.0 * (std::cos(pi * (double)(i - 1) / (double)m) + std::cos(pi * (double)(j - 1) / (double)n) - 2.0);
        const long q = (long)(i - 2) + (long)(m - 1) * (j - 2);
        hl[q] = 1.0 / lam / normalize;
        for (int k = 0; k < L; ++k) {
          const double del22 = del2 * std::pow(4.0, (double)k);
          hl1[(size_t)k * P + q] = (1.0 - 0.5 * dt * lam / del22 / Re) / normalize;
          hl1i[(size_t)k * P + q] = 1.0 / (1.0 + 0.5 * dt * lam / del22 / Re);
        }
      }
    laminv.alloc(P); lam1.alloc((size_t)P * L); lam1i.alloc((size_t)P * L);
    IB_HIP(hipMemcpy(laminv.p, hl.data(), sizeof(double) * P, hipMemcpyHostToDevice));
    IB_HIP(hipMemcpy(lam1.p, hl1.data(), sizeof(double) * P * L, hipMemcpyHostToDevice));
    IB_HIP(hipMemcpy(lam1i.p, hl1i.data(), sizeof(double) * P * L, hipMemcpyHostToDevice));

    // state (initial_condition: all zero, src/variables.F90:262-295)
    omega.alloc((size_t)P * L); s.alloc((size_t)P * L); rhs_old.alloc((size_t)P * L);
    qx.alloc((size_t)g.Fx * L); qy.alloc((size_t)g.Fy * L);
    t0.alloc(P); t1.alloc(P); t2.alloc(P);
    vv.alloc((size_t)P * L); sn.alloc((size_t)P * L); vx.alloc((size_t)g.Fx * L); vy.alloc((size_t)g.Fy * L);
    hx.alloc(g.Fx); hy.alloc(g.Fy);
    lastbc.alloc((size_t)g.nbc * L); obc.alloc(g.nbc); sbcA.alloc(g.nbc); sbcB.alloc(g.nbc); zbc.alloc(g.nbc);
    double* zero_these[] = {omega, s, rhs_old, qx, qy, zbc, vv, hx, hy};
    size_t zero_n[] = {omega.n, s.n, rhs_old.n, qx.n, qy.n, zbc.n, vv.n, hx.n, hy.n};
    for (int i = 0; i < 9; ++i) fill(st, zero_these[i], (long)zero_n[i], 0.0);

    xb.alloc(nf); yb.alloc(nf); f.alloc(nf); f_rdst.alloc(nf); rhsf.alloc(nf); gw.alloc(nf); opx.alloc(nf);
    w.alloc((size_t)nf * kTaps);
    comp.alloc(nf); tors_of.alloc(nf); err_flag.alloc(1);
    if (nf > 0) {
      IB_HIP(hipMemcpy(xb.p, hxb.data(), sizeof(double) * nf, hipMemcpyHostToDevice));
      IB_HIP(hipMemcpy(yb.p, hyb.data(), sizeof(double) * nf, hipMemcpyHostToDevice));
      IB_HIP(hipMemcpy(comp.p, h_ccomp.data(), sizeof(int) * nf, hipMemcpyHostToDevice));
      IB_HIP(hipMemcpy(tors_of.p, h_tors_of.data(), sizeof(int) * nf, hipMemcpyHostToDevice));
      std::vector<int> mv, rg;
      for (int k = 0; k < nf; ++k) (h_moving[k] ? mv : rg).push_back(k);
      nmov = (int)mv.size();
      nrig = (int)rg.size();
      movidx.alloc(std::max(nmov, 1));
      rigidx.alloc(std::max(nrig, 1));
      if (nmov) IB_HIP(hipMemcpy(movidx.p, mv.data(), sizeof(int) * nmov, hipMemcpyHostToDevice));
      if (nrig) IB_HIP(hipMemcpy(rigidx.p, rg.data(), sizeof(int) * nrig, hipMemcpyHostToDevice));
    }
    fill(st, f, nf, 0.0);
    fill(st, f_rdst, nf, 0.0);
    IB_HIP(hipMemsetAsync(err_flag.p, 0, sizeof(int), st));
    tb.alloc(std::max(1, ntors));
    if (ntors > 0) IB_HIP(hipMemcpy(tb.p, h_tb.data(), sizeof(TorsBody) * ntors, hipMemcpyHostToDevice));
    th.alloc((size_t)6 * std::max(1, ntors));
    dabs.alloc(std::max(1, nmoving()));
    fill(st, th, (long)th.n, 0.0);
    if (nbeam > 0) {
      // per beam: M, C, K, K_eff, K_eff^-1, M^-1 (nd^2 each), H (np nd), G
      // (np^2), X0, Y0, and the state u, ud, udd, u0, ud0, udd0
      size_t tot = 0;
      for (const BeamHost& b : h_beam)
        tot += 6 * (size_t)b.nd * b.nd + (size_t)b.np * b.nd + (size_t)b.np * b.np + 2 * (size_t)b.np + 6 * (size_t)b.nd;
      beam_mem.alloc(tot);
      fill(st, beam_mem, (long)tot, 0.0);
      double* at = beam_mem.p;
      auto put = [&](const std::vector<double>& v) {
        double* d = at;
        IB_HIP(hipMemcpy(d, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
        at += v.size();
        return d;
      };
      auto take = [&](size_t n) {
        double* d = at;
        at += n;
        return d;
      };
      for (int k = 0; k < nbeam; ++k) {
        const BeamHost& b = h_beam[k];
        BeamDev d{};
        d.off = h_beam_off[k]; d.np = b.np; d.nd = b.nd; d.nx = b.nx; d.ny = b.ny;
        d.M = put(b.M); d.C = put(b.C); d.K = put(b.K); d.Keff = put(b.Keff); d.Keffi = put(b.Keffi);
        d.Minv = put(b.Minv); d.H = put(b.H); d.G = put(b.G); d.X0 = put(b.X0); d.Y0 = put(b.Y0);
        d.u = take(b.nd); d.ud = take(b.nd); d.udd = take(b.nd);
        d.u0 = take(b.nd); d.ud0 = take(b.nd); d.udd0 = take(b.nd);
        beams.push_back(d);
      }
    }

    V.alloc((size_t)(kmax + 1) * std::max(1, nf));
    H.alloc((size_t)(kmax + 1) * kmax);
    cc.alloc(kmax); ss.alloc(kmax); grs.alloc(kmax + 1); nrs.alloc(kmax + 1);
    pnorm.alloc(std::max(1, ceil_div(std::max(nf, 1), 64)));
    IB_HIP(hipMalloc(&gst, sizeof(GmresDev)));
    IB_HIP(hipHostMalloc(&hst, sizeof(GmresDev), hipHostMallocDefault));
    IB_HIP(hipHostMalloc(&hbuf, sizeof(double) * std::max(64, nmoving() + 1), hipHostMallocDefault));
    IB_HIP(hipHostMalloc(&hflag, 4 * sizeof(int), hipHostMallocDefault));

    csr.alloc(nf * kTaps, g.Fx + g.Fy);
    rebuild_weights();
    if (read_err_flag()) return fail(IB_E_INVAL, "body point regularisation stencil leaves the grid");

    // dense Schur matrix: auto picks it whenever it fits comfortably
    if (prm.schur_mode < 0 || prm.schur_mode > 3) return fail(IB_E_INVAL, "schur_mode must be 0, 1, 2 or 3");
    if (prm.face_margin < -1) return fail(IB_E_INVAL, "face_margin must be >= -1");
    face_margin = prm.face_margin == 0 ? kFaceMarginDefault : std::max(0, prm.face_margin);
    if (sd) {
      // the reference builds B and P at the first advance (src/timestepper.F90:
      // 79-97); neither depends on the flow state
      sd_on = true;
      sdc = *sd;
      if (sdc.droptol == 0.0) sdc.droptol = 7.0e-3;
      if (sdc.interp_points == 0) sdc.interp_points = 4;
      if (sdc.support_nz == 0) sdc.support_nz = 10;
      sd_np = sdc.interp_points;
      if (nf == 0) return fail(IB_E_INVAL, "the sub-domain approach needs bodies");
      assembled = true;  // GMRES runs on lumat, rebuilt every FSI iteration
      Sd.alloc((size_t)nf * nf);
      const int rc = sd_setup(bodies);
      if (rc != IB_OK) return rc;
    } else if (nf > 0 && prm.num_stat) {
      const int rc = num_stat_setup();
      if (rc != IB_OK) return rc;
    } else if (nf > 0) {
      if (prm.schur_mode == 3 || (prm.schur_mode == 0 && nmoving() > 0)) {
        assembled = facemode = true;
      } else if (prm.schur_mode == 2 || (prm.schur_mode == 0 && nf <= kAutoAssembleMax)) {
        assembled = true;
      }
    }
    if (facemode) {
      face_flag.alloc(2);
      tapidx.alloc((size_t)nf * kTaps);
      build_face_set();
      // schur_mode 0 with the default margin: narrower margins (2, then 0
      // cells) before the form gives up G; a narrower set is rebuilt more
      // often as the bodies move
      while (prm.schur_mode == 0 && prm.face_margin == 0 && face_margin > 0 && (size_t)nF * nF > kAutoGreenMax) {
        face_margin = face_margin > 2 ? 2 : 0;
        build_face_set();
      }
      if (prm.schur_mode == 0 && (size_t)nF * nF > kAutoGreenMax) {
        // too many faces near the bodies for G: assemble S directly instead
        facemode = false;
        assembled = nf <= kAutoAssembleMax;
        if (assembled) alloc_assembly(nf);
      } else {
        alloc_assembly(nF + nF / 2);  // room for the face set to grow as bodies move
        compute_green();
        map_taps(st, nf, csr.key, fmap, tapidx, face_flag);
      }
    } else if (assembled && !sd_on) {
      alloc_assembly(nf);
    }
    if (nbeam > 0 && !facemode)
      return fail(IB_E_UNSUPPORTED, "beams (family 3) need the face-space Schur form (schur_mode 0 or 3)");
    if (assembled && !sd_on) {
      if (facemode) {
        form_schur_face(false);
      } else {
        assemble_schur();
      }
    }
    trim_columns();
    itime = prm.istart;
    sync();
    return IB_OK;
  }

  ~ib_handle() {
    if (st) (void)hipStreamSynchronize(st);
    lu_release();
    for (auto& kv : gm_graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto& kv : step_graphs) (void)hipGraphExecDestroy(kv.second);
    if (ev_flags) (void)hipEventDestroy(ev_flags);
    for (auto& t : tim) t.destroy();
    csr.release();
    px.release();
    py.release();
    dmx.release();
    dmy.release();
    pxp.release();
    pyp.release();
    if (gst) (void)hipFree(gst);
    if (hst) (void)hipHostFree(hst);
    if (hbuf) (void)hipHostFree(hbuf);
    if (hflag) (void)hipHostFree(hflag);
    if (st) (void)hipStreamDestroy(st);
  }

  // -------------------------------------------------------------- host I/O
  void h2d(double* d, const double* h, size_t n) {
    IB_HIP(hipMemcpyAsync(d, h, sizeof(double) * n, hipMemcpyHostToDevice, st));
  }
  void d2h(double* h, const double* d, size_t n) {
    IB_HIP(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  }
};
