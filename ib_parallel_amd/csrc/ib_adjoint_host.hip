// ib_adjoint_host.hip -- the discrete adjoint's orchestration (methods of
// ib_handle declared in ib_handle.h): set-up, tapes and checkpoints, the
// reverse step of the fluid, the num_stat solve and the FSI fixed point.
//
// No reference counterpart.  J = sum over the steps of sum over bodies of
// cx_b drag_b + cy_b lift_b (write_force2's fx, fy), divided by the number
// of steps.  Rigid bodies with num_stat: a step is predictor, vort2flux,
// f = A^-1 E(q* + q0) with the symmetric num_stat matrix A, corrector,
// vort2flux; the adjoint runs these backwards with the transposed
// operators (the 2-D DST D is symmetric with D D = normalize I, the
// spectral tables are diagonal, E^T <-> E, rot^T = curl with zero boundary
// values, the stencil transposes of ib_adjoint.hip).  Re enters through
// lam1, lam1i, vfac and A.  With d lam1i/dRe = (lam1i - lam1i^2)/Re and
// d lam1/dRe = (1/normalize - lam1)/Re every spectral Re term reduces to
// node products of arrays the sweep already holds; dA/dRe is assembled
// once by the batched columns with d lam1i/dRe in place of lam1i.
// The steps of the last segment are taped during the forward pass; the
// other segments restart from checkpoints (every K steps) and are
// recomputed once, with a tape, before their adjoint.
#include <cmath>
#include <string>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "ib_handle.h"

void ib_handle::adjoint_setup(int nsteps, int K) {
  const long P = g.P, LP = (long)L * P, nbc = g.nbc;
  if (!adj_ready) {
    const double pi = 4.0 * std::atan(1.0), del2 = delta * delta;
    std::vector<double> h1i((size_t)LP), hll((size_t)LP);
    for (int j = 2; j <= g.n; ++j)
      for (int i = 2; i <= g.m; ++i) {
        const double lam = -2.0 * (std::cos(pi * (double)(i - 1) / (double)g.m) +
                                    std::cos(pi * (double)(j - 1) / (double)g.n) - 2.0);
        const long q = (long)(i - 2) + (long)(g.m - 1) * (j - 2);
        for (int k = 0; k < L; ++k) {
          const double a = 0.5 * dt * lam / (del2 * std::pow(4.0, (double)k));
          const double li = 1.0 / (1.0 + a / Re), l1 = (1.0 - a / Re) / normalize;
          h1i[(size_t)k * P + q] = li * li * a / (Re * Re);
          hll[(size_t)k * P + q] = l1 * li;
        }
      }
    dlam1i.alloc(LP); ll.alloc(LP);
    IB_HIP(hipMemcpy(dlam1i.p, h1i.data(), sizeof(double) * LP, hipMemcpyHostToDevice));
    IB_HIP(hipMemcpy(ll.p, hll.data(), sizeof(double) * LP, hipMemcpyHostToDevice));
    if (!adj_fsi) {
      // dA/dRe: the Schur columns with d lam1i / dRe in the ainv stage (no
      // body block: every body is rigid)
      dA.alloc((size_t)nf * nf);
      alloc_assembly(nf);
      for (int c0 = 0; c0 < nf; c0 += ncol) {
        const int nc = std::min(ncol, nf - c0);
        long sbc = 0;
        const double* bc = column_batch(c0, nc, w, csr.key, &sbc, dlam1i);
        schur_columns(st, g, nf, c0, nc, w, csr.key, Sb, g.P, bc, sbc, adj_gm ? tors_of.p : no_tors.p, tb, th, dt,
                      delta, dA);
      }
      if (adj_gm) {
        // GMRES on the fixed assembled S: its transpose once
        SdT.alloc((size_t)nf * nf);
        transpose(st, nf, Sd, SdT);
      } else {
        // num_stat solves with A's upper triangle mirrored: store dA/dRe that
        // way so each step's product is a plain gemv
        DBuf<double> du;
        du.alloc((size_t)nf * nf);
        copy(st, du, dA, (long)nf * nf);
        symmetrize_upper(st, nf, du, dA);
      }
      sync();
      trim_columns();
    } else {
      const int nt = std::max(ntors, 1);
      SdT.alloc((size_t)nf * nf);
      aav.alloc(nf); zcv.alloc(nf); fmask.alloc(nf); cvv.alloc(nf);
      r1v.alloc((size_t)nt * nf); Zv.alloc((size_t)nt * nf);
      htv.alloc((size_t)nf * kTaps); dthv.alloc((size_t)nf * kTaps);
      dwx.alloc((size_t)nf * kTaps); dwy.alloc((size_t)nf * kTaps);
      hx2.alloc(g.Fx); hy2.alloc(g.Fy);
      scal.alloc((size_t)16 * nt + (size_t)8 * nt * nt);
      h_scal.resize((size_t)16 * nt + (size_t)8 * nt * nt);
      gpx.alloc(nf); gpy.alloc(nf);
      fin_th.alloc((size_t)6 * nt + 2 * (size_t)nf);
    }
    aw.alloc(LP); aro.alloc(LP); aws.alloc(LP); asb.alloc(LP); racc.alloc(P);
    cur_w.alloc(LP); cur_ws.alloc(LP); cur_qx.alloc((size_t)L * g.Fx); cur_qy.alloc((size_t)L * g.Fy);
    cur_f.alloc(nf); cur_qt.alloc((size_t)nf * kTaps);
    raccb.alloc(nbc);
    aqx.alloc((size_t)L * g.Fx); aqy.alloc((size_t)L * g.Fy);
    af.alloc(nf); arf.alloc(nf); af2.alloc(nf); cw.alloc(nf);
    aP1.alloc(P); aP2.alloc(P); aD1.alloc(P); aD2.alloc(P); aD3.alloc(P);
    aFx.alloc(g.Fx); aFy.alloc(g.Fy);
    const long NA = (long)(g.m + 1) * (g.n + 1);
    anlx.alloc(NA); anly.alloc(NA); aub.alloc(NA); avb.alloc(NA);
    albb.alloc((size_t)L * nbc); albk.alloc(nbc); aobk.alloc(nbc); asbcb.alloc(nbc); abct.alloc(nbc);
    adot.alloc(1); adJ.alloc(1);
    fin_w.alloc(LP); fin_ro.alloc(LP); fin_qx.alloc((size_t)L * g.Fx); fin_qy.alloc((size_t)L * g.Fy);
    fin_s.alloc(LP); fin_f.alloc(nf); fin_fr.alloc(nf);
    adj_ready = true;
  }
  const int nseg = (nsteps + K - 1) / K;
  const size_t need_c = (size_t)std::max(nseg - 1, 1), need_t = (size_t)std::min(K, nsteps);
  if (c_w.n < need_c * LP) {
    c_w.alloc(need_c * LP); c_ro.alloc(need_c * LP);
    c_qx.alloc(need_c * L * g.Fx); c_qy.alloc(need_c * L * g.Fy);
  }
  if (t_w.n < need_t * LP) {
    t_w.alloc(need_t * LP); t_ws.alloc(need_t * LP);
    t_qx.alloc(need_t * L * g.Fx); t_qy.alloc(need_t * L * g.Fy); t_f.alloc(need_t * nf);
  }
  if ((want_pos || adj_fsi) && t_qt.n < need_t * nf * kTaps) t_qt.alloc(need_t * nf * kTaps);
  if (adj_fsi) {
    const size_t bs = (size_t)6 * ntors + 2 * (size_t)nf;  // th, xb, yb
    if (t_th.n < need_t * bs) { t_th.alloc(need_t * bs); t_f0.alloc(need_t * nf); }
    if (c_th.n < need_c * bs) { c_th.alloc(need_c * bs); c_f.alloc(need_c * nf); }
  }
}

// G's derivative in Re on the current face set (columns with d lam1i/dRe)
void ib_handle::compute_green_d() {
  if (dGm.n < (size_t)nF * nF) dGm.alloc((size_t)nF * nF);
  alloc_assembly(nF);
  for (int c0 = 0; c0 < nF; c0 += ncol) {
    const int nc = std::min(ncol, nF - c0);
    long sbc = 0;
    const double* bc = column_batch(c0, nc, nullptr, Fkeys, &sbc, dlam1i);
    green_columns(st, g, nF, c0, nc, Fkeys, Sb, g.P, bc, sbc, dGm);
  }
  dG_build = green_builds;
}

// S = P L U of the current Sd for the reverse step's solves with S^T: one
// factorisation per reverse step (moving bodies) or per run (the fixed S of
// rigid bodies) in place of a GMRES solve per right side.  The adjoint
// differentiates the converged forward map, so the exact solve is what it
// needs; GMRES on S^T took ~1900 iterations per force-adjoint solve on the
// airfoil + flap system (cond(S) ~ 1.5e8), most of a reverse step.
void ib_handle::lu_factor() {
  if (!lu_blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) throw HipError("adjoint: rocblas_create_handle failed");
    lu_blas = h;
  }
  rocblas_handle h = static_cast<rocblas_handle>(lu_blas);
  if (rocblas_set_stream(h, st) != rocblas_status_success) throw HipError("adjoint: rocblas_set_stream failed");
  if (LUd.n < (size_t)nf * nf) LUd.alloc((size_t)nf * nf);
  if (lu_piv.n < (size_t)nf) lu_piv.alloc(nf);
  if (!lu_info.p) lu_info.alloc(1);
  copy(st, LUd, Sd, (long)nf * nf);
  if (rocsolver_dgetrf(h, nf, nf, LUd.p, nf, lu_piv.p, lu_info.p) != rocblas_status_success)
    throw HipError("adjoint: rocsolver_dgetrf failed");
  int info = 0;
  IB_HIP(hipMemcpyAsync(&info, lu_info.p, sizeof(int), hipMemcpyDeviceToHost, st));
  sync();
  if (info != 0) throw HipError("adjoint: the Schur matrix is singular (getrf info " + std::to_string(info) + ")");
  lu_ok = true;
}

void ib_handle::lu_release() {
  if (lu_blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(lu_blas));
  lu_blas = nullptr;
  lu_ok = false;
}

int ib_handle::solve_T(const double* b, double* x, bool warm) {
  if (lu_ok) {
    copy(st, x, b, nf);
    if (rocsolver_dgetrs(static_cast<rocblas_handle>(lu_blas), rocblas_operation_transpose, nf, 1, LUd.p, nf,
                         lu_piv.p, x, nf) != rocblas_status_success)
      throw HipError("adjoint: rocsolver_dgetrs failed");
    return IB_OK;
  }
  if (!warm) fill(st, x, nf, 0.0);
  if (dense_solve_ok()) return dense_solve(SdT, b, x);
  GmresDev init{};
  init.first = 1;
  *hst = init;
  IB_HIP(hipMemcpyAsync(gst, hst, sizeof(GmresDev), hipMemcpyHostToDevice, st));
  int reason = 0, its = 0;
  const int maxit = prm.gmres_max_it;
  while (true) {
    gemv(st, nf, SdT, x, gw);
    gm_start(st, nf, kmax, b, gw, V, gst, H, grs, prm.rtol, prm.atol);
    read_gmres();
    reason = hst->reason;
    if (reason) break;
    int it = 0;
    while (true) {
      const int nb = std::min(kGmBatch, std::min(kmax - it, maxit - its));
      for (int q = 0; q < nb; ++q) {
        gemv(st, nf, SdT, V + (size_t)(it + q) * nf, V + (size_t)(it + q + 1) * nf);
        gm_arnoldi(st, nf, kmax, V, gst, H, cc, ss, grs, pnorm);
      }
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
  return reason > 0 ? IB_OK : IB_E_NOCONV_GMRES;
}

// the body at the end of a taped step's FSI loop: theta (and its history),
// positions, weights, S (and S^T), the weight derivatives
void ib_handle::body_restore(const double* bs) {
  copy(st, th, bs, 6 * (long)ntors);
  copy(st, xb, bs + 6 * ntors, nf);
  copy(st, yb, bs + 6 * ntors + nf, nf);
  rebuild_weights();
  refresh_schur();
  if (dG_build != green_builds) compute_green_d();
  if (xfv.n < (size_t)nF) {
    xfv.alloc(nF); vTv.alloc(nF); yfv.alloc(nF); zfv.alloc(nF); lfv.alloc(nF); dgv.alloc(nF); uTv.alloc(nF);
  }
  lu_factor();
  setup_reg_d(st, g, nf, xb, yb, comp, delta, offx, offy, prm.deltafnc_f32_literals, dwx, dwy);
  dthw(st, nf, ntors, tors_of, tb, th, dwx, dwy, dthv);
}

// set-up of the position gradient (once per handle): the weight
// derivatives, G on the faces the taps touch (the face set with margin 0),
// and the two contractions T = G W^T, U = W G
void ib_handle::adjoint_pos_setup() {
  if (adj_pos) return;
  dwx.alloc((size_t)nf * kTaps); dwy.alloc((size_t)nf * kTaps);
  setup_reg_d(st, g, nf, xb, yb, comp, delta, offx, offy, prm.deltafnc_f32_literals, dwx, dwy);
  const int fm_save = face_margin;
  face_margin = 0;
  build_face_set();
  face_margin = fm_save;
  if (!tapidx.p) tapidx.alloc((size_t)nf * kTaps);
  if (!face_flag.p) face_flag.alloc(2);
  alloc_assembly(nF);
  compute_green();
  map_taps(st, nf, csr.key, fmap, tapidx, face_flag);
  TTm.alloc((size_t)nf * nF); Um.alloc((size_t)nf * nF);
  green_tt_u(st, nf, nF, w, tapidx, Gm, TTm, Um);
  sync();
  trim_columns();
  Gm.release(); Tm.release();
  tcache = false;
  gpx.alloc(nf); gpy.alloc(nf);
  adj_pos = true;
}

// ckpt_every <= 0: tape the whole run when it fits in half of the free
// HBM, else checkpoints every ceil(sqrt(nsteps)) steps
int ib_handle::adjoint_segment(int nsteps, int K) {
  if (K > 0) return std::min(K, nsteps);
  size_t fr = 0, tot = 0;
  IB_HIP(hipMemGetInfo(&fr, &tot));
  const size_t have = c_w.n * sizeof(double) + t_w.n * sizeof(double) * 3;  // reusable buffers
  const size_t budget = (fr + have) / 2;
  if ((size_t)nsteps * tape_bytes() <= budget) return nsteps;
  return std::max(1, (int)std::ceil(std::sqrt((double)nsteps)));
}

void ib_handle::save_state(double* w_, double* ro_, double* qx_, double* qy_) {
  const long LP = (long)L * g.P;
  copy(st, w_, omega, LP);
  if (ro_) copy(st, ro_, rhs_old, LP);
  copy(st, qx_, qx, (long)L * g.Fx); copy(st, qy_, qy, (long)L * g.Fy);
}

void ib_handle::load_state(const double* w_, const double* ro_, const double* qx_, const double* qy_) {
  const long LP = (long)L * g.P;
  copy(st, omega, w_, LP); copy(st, rhs_old, ro_, LP);
  copy(st, qx, qx_, (long)L * g.Fx); copy(st, qy, qy_, (long)L * g.Fy);
}

// moving bodies: th, positions and f (GMRES's initial guess) with the flow
void ib_handle::save_body(double* bs, double* f_) {
  copy(st, bs, th, 6 * (long)ntors);
  copy(st, bs + 6 * ntors, xb, nf);
  copy(st, bs + 6 * ntors + nf, yb, nf);
  if (f_) copy(st, f_, f, nf);
}

void ib_handle::load_body(const double* bs, const double* f_) {
  copy(st, th, bs, 6 * (long)ntors);
  copy(st, xb, bs + 6 * ntors, nf);
  copy(st, yb, bs + 6 * ntors + nf, nf);
  if (f_) copy(st, f, f_, nf);
  rebuild_weights();
  refresh_schur();
}

// one taped step: the step-start omega and q, the predicted omega, f
int ib_handle::taped_advance(int sl) {
  const long LP = (long)L * g.P;
  save_state(t_w + sl * LP, nullptr, t_qx + (long)sl * L * g.Fx, t_qy + (long)sl * L * g.Fy);
  tape_ws = t_ws + sl * LP;
  tape_qt = (want_pos || adj_fsi) ? t_qt + (long)sl * nf * kTaps : nullptr;
  if (adj_fsi) {
    tape_th = t_th + (long)sl * (6 * ntors + 2 * nf);
    copy(st, t_f0 + (long)sl * nf, f, nf);
  }
  const int rc = advance();
  tape_ws = nullptr;
  tape_qt = nullptr;
  tape_th = nullptr;
  copy(st, t_f + (long)sl * nf, f, nf);
  return rc;
}

// transpose of vort2flux: wb (all levels) and qxb, qyb (all levels) are the
// adjoints of its outputs; on return wb is the adjoint of its input
// omega, and qxb, qyb are zero
void ib_handle::v2f_adj(double* wb, double* qxb, double* qyb) {
  const long P = g.P;
  fill(st, asb, (long)L * P, 0.0);
  for (int k = 0; k < L; ++k) {
    double* qxk = qxb + (long)k * g.Fx;
    double* qyk = qyb + (long)k * g.Fy;
    rot_acc(st, g, qxk, qyk, asb + k * P);
    if (k < L - 1) curl_adj_bc(st, g, qxk, qyk, asbcb);
    ctci_dev(asb + k * P, aP2, nullptr);
    axpy(st, P, 1.0, aP2, wb + k * P);
    if (k < L - 1) {
      applybc_adj_acc(st, g, aP2, asbcb);
      getbc_adj(st, g, asbcb, 1.0, asb + (k + 1) * P);
    }
  }
  fill(st, qxb, (long)L * g.Fx, 0.0);
  fill(st, qyb, (long)L * g.Fy, 0.0);
  for (int k = L - 1; k >= 1; --k) coarsify_adj(st, g, wb + k * P, wb + (k - 1) * P);
}

// one step backwards; tape slot sl holds the step-start omega and q, the
// predicted omega and f; (aw, aro, aqx, aqy) enter as the adjoint of the
// step-end state and leave as that of the step-start state
int ib_handle::adjoint_step(int sl, bool first) {
  const long P = g.P, LP = (long)L * P, nbc = g.nbc;
  const double* w0 = t_w + sl * LP;
  const double* qx0 = t_qx + (long)sl * L * g.Fx;
  const double* qy0 = t_qy + (long)sl * L * g.Fy;
  const double* ws0 = t_ws + sl * LP;
  const double* f0 = t_f + (long)sl * nf;
  const double* qt = (want_pos || adj_fsi) ? t_qt + (long)sl * nf * kTaps : nullptr;
  const double invRe = 1.0 / Re;
  // with graphs, the slot's tape is copied into fixed buffers first so that
  // the captured launches never change
  const bool graphs = step_graphs_ok();
  if (graphs) {
    copy(st, cur_w, w0, LP);
    copy(st, cur_qx, qx0, (long)L * g.Fx);
    copy(st, cur_qy, qy0, (long)L * g.Fy);
    copy(st, cur_ws, ws0, LP);
    copy(st, cur_f, f0, nf);
    if (qt) copy(st, cur_qt, qt, (long)nf * kTaps);
    w0 = cur_w; qx0 = cur_qx; qy0 = cur_qy; ws0 = cur_ws; f0 = cur_f;
    if (qt) qt = cur_qt;
  }
  if (adj_fsi) body_restore(t_th + (long)sl * (6 * ntors + 2 * nf));

  // the end-of-step vort2flux; corrector omega_0 -= ainv(rot(E^T f)): the
  // adjoint of f (J depends on f through cw), and its Re term
  // -<aw_0 - ainv(aw_0), ainv(v)>/Re
  auto corrector = [&] {
    v2f_adj(aw, aqx, aqy);
    copy(st, af, cw, nf);
    ainv_dev(aw, aP1, false);
    curl_getbc(st, g, aP1, zbc, aFx, aFy, nullptr);
    if (adj_fsi) gather_taps(st, g, nf * kTaps, csr.key, aFx, aFy, 0.0, htv);
    regT(st, g, nf, w, csr.key, aFx, aFy, 0.0, nullptr, arf);
    axpy(st, nf, -1.0, arf, af);
    reg(st, g, csr, w, f0, hx, hy);
    ainv_dev(nullptr, aP2, false);
    re_corr(st, P, aw, aP1, aP2, invRe, racc);
  };
  // f = A^-1 rhsf (A symmetric): the adjoint of rhsf, the Re term
  // -<rhsf bar, dA/dRe f>, and the body-point position terms
  auto stat_solve_adj = [&] {
    gemv(st, nf, Ainv, af, arf);
    gemv(st, nf, dA, f0, af2);
    dot_acc(st, nf, arf, af2, -1.0, adot);
    if (want_pos) pos_terms(st, g, nf, csr.key, tapidx, dwx, dwy, TTm, Um, arf, f0, qt, aFx, aFy, gpx, gpy);
  };
  // rhsf = E(q*_0 + q0): the adjoint of q*_0
  auto qstar_adj = [&] {
    fill(st, aFx, g.Fx, 0.0);
    fill(st, aFy, g.Fy, 0.0);
    reg(st, g, csr, w, arf, aFx, aFy);
    axpy(st, g.Fx, 1.0, aFx, aqx);
    axpy(st, g.Fy, 1.0, aFy, aqy);
  };
  // the vort2flux after the predictor, then the predictor
  // omega*_k = D(lam1i o (D r_k + lam1 o D omega_k)), fine to coarse (the
  // reverse of its coarse-to-fine order)
  auto predictor = [&] {
    v2f_adj(aw, aqx, aqy);
    copy(st, aws, aw, LP);
    fill(st, aw, LP, 0.0);
    fill(st, albb, (long)L * nbc, 0.0);
    for (int k = 0; k < L; ++k) {
      const double* wk = w0 + k * P;
      if (k < L - 1) {
        get_bc(st, g, w0 + (k + 1) * P, P, albk, nbc, 1, 0.25);
        get_bc(st, g, ws0 + (k + 1) * P, P, aobk, nbc, 1, 0.25);
      } else {
        fill(st, albk, nbc, 0.0);
        fill(st, aobk, nbc, 0.0);
      }
      // rbar = D(lam1i o D wsb_k), D(lam1 lam1i o D wsb_k): one x pass, two
      // column passes, one x pass over both
      DstLaunch a;
      a.in[0] = aws + k * P; a.out[0] = aD1;
      xpass(X_PLAIN, a);
      DstLaunch b1;
      b1.in[0] = aD1; b1.out[0] = aD2; b1.tab = lam1i + k * P; b1.div = 1.0;
      ypass(Y_SCALE2, b1);
      DstLaunch b2;
      b2.in[0] = aD1; b2.out[0] = aD3; b2.tab = ll + k * P; b2.div = 1.0;
      ypass(Y_SCALE2, b2);
      DstLaunch c;
      c.batch = 2;
      c.in[0] = aD2; c.out[0] = aP1;
      c.in[1] = aD3; c.out[1] = aP2;
      xpass(X_PLAIN, c);
      re_pred(st, P, ws0 + k * P, wk, aws + k * P, aP1, aP2, 1.0 / normalize, invRe, racc, aw + k * P);
      // the boundary terms vfac (apply_bc(lastbc) + apply_bc(omegabc)) and
      // their Re term
      pred_bc_adj(st, g, aP1, vfac[k], invRe, albk, aobk, asbcb, albb + k * nbc, raccb);
      if (k < L - 1) getbc_adj(st, g, asbcb, 0.25 * vfac[k], aws + (k + 1) * P);
      // r = con1 rhs + con2 rold + ..., rhs_old <- rhs
      rhs_adj(st, P, aP1, first ? con1[k] + con2[k] : con1[k], first ? 0.0 : con2[k], aro + k * P, aD2);
      nl_adj(st, g, wk, qx0 + (long)k * g.Fx, qy0 + (long)k * g.Fy, q0x[k], albk, aD2, anlx, anly, aub, avb,
             aw + k * P, albb + k * nbc, aqx + (long)k * g.Fx, aqy + (long)k * g.Fy);
    }
    // lastbc_k = 0.25 get_bc(omega_(k+1))
    for (int k = 0; k < L - 1; ++k) getbc_adj(st, g, albb + k * nbc, 0.25, aw + (k + 1) * P);
  };

  const int fk = first ? 1 : 0;
  if (!adj_fsi && !adj_gm) {
    auto all = [&] {
      corrector();
      stat_solve_adj();
      qstar_adj();
      predictor();
    };
    if (graphs) replay(100 + 2 * fk + (want_pos ? 1 : 0), all);
    else all();
    return IB_OK;
  }
  if (adj_gm) {
    // rigid bodies, GMRES on the fixed assembled S: rhsf bar = S^-T f bar
    // (a GMRES solve, read back between graphs) and the Re term
    // -<rhsf bar, dS/dRe f>
    if (graphs) replay(120, corrector);
    else corrector();
    // from the previous reverse step's solution (zero before the first)
    const int rc = solve_T(af, arf, true);
    if (rc != IB_OK) return fail(rc, "adjoint: GMRES on S^T did not converge");
    gemv(st, nf, dA, f0, af2);
    dot_acc(st, nf, arf, af2, -1.0, adot);
    if (graphs) replay(122 + fk, [&] {
      qstar_adj();
      predictor();
    });
    else {
      qstar_adj();
      predictor();
    }
    return IB_OK;
  }
  // moving bodies: the FSI part reads scalars back to the host and its
  // face-set products change size with the set, so it runs between graphs
  if (graphs) replay(110, corrector);
  else corrector();
  const int rc = fsi_adjoint(sl, first, f0, qt);
  if (rc != IB_OK) return rc;
  if (graphs) replay(111, qstar_adj);
  else qstar_adj();
  // Re through G: -<W^T lambda1, dG/dRe W^T f>; rigid points: S's weights
  face_gather(st, g, nF, Fkeys, aFx, aFy, lfv);
  gemv(st, nF, dGm, xfv, dgv);
  dot_acc(st, nF, lfv, dgv, -1.0, adot);
  if (want_pos) {
    gemv_t(st, nF, Gm, lfv, uTv);
    pos_fsi(st, nf, tors_of, tapidx, dwx, dwy, arf, f0, htv, qt, vTv, uTv, gpx, gpy, ntors ? psx.p : nullptr,
            ntors ? psy.p : nullptr);
    if (ntors) ds_weights(st, tb, th, ntors, psx, psy, gdsw);
  }
  if (graphs) replay(112 + fk, predictor);
  else predictor();
  return IB_OK;
}

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
int ib_handle::fsi_adjoint(int sl, bool first, const double* f0, const double* qt) {
  const int nt = ntors;
  const double* fstart = t_f0 + (long)sl * nf;
  // f on the face set and G W^T f (hx, hy hold E^T f on the touched faces)
  face_gather(st, g, nF, Fkeys, hx, hy, xfv);
  gemv(st, nF, Gm, xfv, vTv);
  // both solves start from the previous reverse step's solutions (zero
  // before the first): S and its right sides change little from step to
  // step, and the sweep visits the steps in the same order for every
  // segment length, so the results do not depend on the checkpoints
  int rc = solve_T(af, aav, true);
  if (rc != IB_OK) return fail(rc, "adjoint: GMRES on S^T did not converge");
  double* sc = scal;
  const long zs = 16 * (long)nt;  // per (t, s): BR_t^T Z_s at [2], r1_t^T Z_s at [3]
  for (int t = 0; t < nt; ++t) {
    const TorsBody& hb = h_tb[t];
    const int o = hb.off, n2 = 2 * hb.npts;
    fill(st, zcv, nf, 0.0);
    tors_block(st, tb, th, nt, t, 0, dt, delta, f0, aav, nullptr, zcv, sc + 16 * t);
    rc = solve_T(zcv, Zv + (long)t * nf, true);
    if (rc != IB_OK) return fail(rc, "adjoint: GMRES on S^T did not converge");
    double* r1 = r1v + (long)t * nf;
    // W G (E'_t^T f) + E'_t (G W^T f - (q* + q0)) + B'_t f - BR'_t sc
    fill(st, fmask, nf, 0.0);
    copy(st, fmask + o, f0 + o, n2);
    fill(st, hx2, g.Fx, 0.0);
    fill(st, hy2, g.Fy, 0.0);
    reg(st, g, csr, dthv, fmask, hx2, hy2);
    face_gather(st, g, nF, Fkeys, hx2, hy2, yfv);
    gemv(st, nF, Gm, yfv, zfv);
    face_regT(st, nf, w, tapidx, zfv, r1);
    tap_contract(st, o, o + n2, dthv, nullptr, qt, vTv, tapidx, r1);
    tors_block(st, tb, th, nt, t, 0, dt, delta, f0, nullptr, r1, nullptr, nullptr);
    dot_set(st, 0, nf, r1, aav, sc + 16 * t + 3);
    // the corrector's E(theta): sum over the body's entries of f dthw hb
    fill(st, cvv, nf, 0.0);
    tap_contract(st, o, o + n2, dthv, nullptr, htv, nullptr, nullptr, cvv);
    dot_set(st, o, o + n2, f0, cvv, sc + 16 * t + 4);
    if (first) tors_block(st, tb, th, nt, t, 1, dt, delta, fstart, nullptr, nullptr, nullptr, sc + 16 * t + 8);
  }
  for (int t = 0; t < nt; ++t)
    for (int u = 0; u < nt; ++u) {
      double* s2 = sc + zs + 8 * ((long)t * nt + u);
      tors_block(st, tb, th, nt, t, 0, dt, delta, f0, Zv + (long)u * nf, nullptr, nullptr, s2);
      dot_set(st, 0, nf, r1v + (long)t * nf, Zv + (long)u * nf, s2 + 3);
    }
  const long nsc = zs + 8 * (long)nt * nt;
  std::vector<double> hth(6 * (size_t)nt);
  IB_HIP(hipMemcpyAsync(h_scal.data(), scal.p, sizeof(double) * nsc, hipMemcpyDeviceToHost, st));
  IB_HIP(hipMemcpyAsync(hth.data(), th.p, sizeof(double) * 6 * nt, hipMemcpyDeviceToHost, st));
  sync();
  auto S0 = [&](int t, int k) { return h_scal[16 * (size_t)t + k]; };
  auto SZ = [&](int t, int u, int k) { return h_scal[zs + 8 * ((size_t)t * nt + u) + k]; };
  // (D - r1^T Z) lambda2 = gtheta - r1^T a, D = diag(-1 + c2 BR'^T f)
  std::vector<double> A((size_t)nt * nt), rhs(nt), lam2(nt);
  for (int t = 0; t < nt; ++t) {
    const TorsBody& b = h_tb[t];
    const double c2 = delta / (dt * (4.0 / (dt * dt) * b.M + b.K));
    const double gth = bth[t] + 2.0 / dt * bthd[t] + 4.0 / (dt * dt) * bthdd[t] - S0(t, 4);
    rhs[t] = gth - S0(t, 3);
    for (int u = 0; u < nt; ++u) A[(size_t)t * nt + u] = (t == u ? -1.0 + c2 * S0(t, 1) : 0.0) - SZ(t, u, 3);
  }
  // Gaussian elimination with partial pivoting (nt is the number of
  // torsional bodies)
  for (int c = 0; c < nt; ++c) {
    int pr = c;
    for (int r = c + 1; r < nt; ++r)
      if (std::fabs(A[(size_t)r * nt + c]) > std::fabs(A[(size_t)pr * nt + c])) pr = r;
    if (pr != c) {
      for (int k = 0; k < nt; ++k) std::swap(A[(size_t)c * nt + k], A[(size_t)pr * nt + k]);
      std::swap(rhs[c], rhs[pr]);
    }
    for (int r = c + 1; r < nt; ++r) {
      const double m = A[(size_t)r * nt + c] / A[(size_t)c * nt + c];
      for (int k = c; k < nt; ++k) A[(size_t)r * nt + k] -= m * A[(size_t)c * nt + k];
      rhs[r] -= m * rhs[c];
    }
  }
  for (int c = nt - 1; c >= 0; --c) {
    double v = rhs[c];
    for (int k = c + 1; k < nt; ++k) v -= A[(size_t)c * nt + k] * lam2[k];
    lam2[c] = v / A[(size_t)c * nt + c];
  }
  // lambda1 = a - Z lambda2
  copy(st, arf, aav, nf);
  for (int u = 0; u < nt; ++u) axpy(st, nf, -lam2[u], Zv + (long)u * nf, arf);
  // history, stiffness and pitch adjoints
  for (int t = 0; t < nt; ++t) {
    const TorsBody& b = h_tb[t];
    const double M = b.M, Cd = b.C, K = b.K;
    const double den = 4.0 / (dt * dt) * M + 2.0 / dt * Cd + K, kd = 4.0 / (dt * dt) * M + K;
    const double c2 = delta / (dt * kd), scb = 2.0 * delta * delta / (dt * dt);
    const double th0 = hth[3 * nt + t], thd0 = hth[4 * nt + t], thdd0 = hth[5 * nt + t];
    const double Q = (M * (4.0 / (dt * dt) * th0 + 4.0 / dt * thd0 + thdd0) + Cd * (thd0 + 2.0 / dt * th0)) / den;
    const double dQ0 = (4.0 / (dt * dt) * M + 2.0 / dt * Cd) / den, dQ1 = (4.0 / dt * M + Cd) / den, dQ2 = M / den;
    const double ds0 = -delta * (2.0 / dt * dQ0 - 2.0 / dt), ds1 = -delta * (2.0 / dt * dQ1 - 1.0);
    const double ds2 = -delta * 2.0 / dt * dQ2;
    double brl1 = S0(t, 2), r1l1 = S0(t, 3);
    for (int u = 0; u < nt; ++u) {
      brl1 -= lam2[u] * SZ(t, u, 2);
      r1l1 -= lam2[u] * SZ(t, u, 3);
    }
    const double l2 = lam2[t], brf = S0(t, 0), brpf = S0(t, 1);
    double nb0 = -2.0 / dt * bthd[t] - 4.0 / (dt * dt) * bthdd[t] + brl1 * ds0 - l2 * dQ0;
    double nb1 = -bthd[t] - 4.0 / dt * bthdd[t] + brl1 * ds1 - l2 * dQ1;
    double nb2 = -bthdd[t] + brl1 * ds2 - l2 * dQ2;
    gK[t] += scb * brl1 * brf / (den * den) + brl1 * (-delta * 2.0 / dt * (-Q / den)) - l2 * (-Q / den - c2 / kd * brf);
    gG0[t] += r1l1 + c2 * brpf * l2 + S0(t, 4);
    // ds: BR is linear in it, so d/d(ds) of (scb/den) BR (BR^T f) - sc BR
    // and of c2 BR^T f are those terms over ds (sc, Q and c2 do not
    // depend on it); the weights' share is accumulated on the device
    const double sc = -delta * (2.0 / dt * Q - thd0 - 2.0 / dt * th0);
    gDS[t] -= ((2.0 * scb / den * brf - sc) * brl1 + c2 * brf * l2) / b.deltas;
    if (first) {
      // initial_acc: thetadd0 = (-C thetad0 - K theta0 + delta/dt BR(theta0)^T f) / M
      const double brp0 = S0(t, 9);
      nb0 += nb2 * (-K / M + delta / dt * brp0 / M);
      nb1 += nb2 * (-Cd / M);
      gK[t] += nb2 * (-th0 / M);
      gG0[t] += nb2 * (-delta / dt * brp0 / M);
      gDS[t] += nb2 * (delta / dt * S0(t, 8) / M / b.deltas);
      nb2 = 0.0;
    }
    bth[t] = nb0;
    bthd[t] = nb1;
    bthdd[t] = nb2;
  }
  return IB_OK;
}

int ib_handle::adjoint_run(int nsteps, int Kin, const double* cx, const double* cy, double* J, double* dRe, double* dw0,
                double* dxb, double* dyb, double* dK, double* dg0) {
  const bool stat_ok = prm.num_stat && ntors == 0 && !sd_on && nf > 0;
  if (nbeam > 0) return fail(IB_E_UNSUPPORTED, "the adjoint does not cover family-3 (beam) bodies");
  const bool fsi_ok = !prm.num_stat && facemode && !sd_on && nf > 0;
  const bool gm_ok = !prm.num_stat && assembled && !facemode && ntors == 0 && !sd_on && nf > 0;
  if (!stat_ok && !fsi_ok && !gm_ok)
    return fail(IB_E_UNSUPPORTED,
                "the adjoint covers rigid bodies with num_stat = 1 or an assembled Schur matrix, and the "
                "face-space form (schur_mode 3, or 0 with torsional bodies)");
  if (gm_ok && (dxb || dyb))
    return fail(IB_E_UNSUPPORTED, "position gradients need num_stat = 1 or the face-space form");
  if (nsteps < 1) return fail(IB_E_INVAL, "nsteps must be >= 1");
  if (adj_ready && (adj_fsi != fsi_ok || adj_gm != gm_ok)) return fail(IB_E_INVAL, "adjoint set-up of another form");
  adj_fsi = fsi_ok;
  adj_gm = gm_ok;
  const int K = adjoint_segment(nsteps, Kin);
  want_pos = dxb || dyb;
  hinge_ok = false;
  if (want_pos && !adj_fsi) adjoint_pos_setup();
  adjoint_setup(nsteps, K);
  if (want_pos || adj_fsi) {
    fill(st, gpx, nf, 0.0);
    fill(st, gpy, nf, 0.0);
  }
  const int nt = ntors;
  bth.assign(nt, 0.0); bthd.assign(nt, 0.0); bthdd.assign(nt, 0.0); gK.assign(nt, 0.0); gG0.assign(nt, 0.0);
  gDS.assign(nt, 0.0);
  if (adj_fsi && nt > 0) {
    if (psx.n < (size_t)nf) {
      psx.alloc(nf);
      psy.alloc(nf);
    }
    if (gdsw.n < (size_t)nt) gdsw.alloc(nt);
    fill(st, gdsw, nt, 0.0);
  }
  const long LP = (long)L * g.P;
  const long CQX = (long)L * g.Fx, CQY = (long)L * g.Fy, BS = 6 * (long)nt + 2 * (long)nf;
  // objective weights per force entry: write_force2's 2 delta / dt, time mean
  std::vector<double> hw(nf);
  for (int k = 0; k < nf; ++k)
    hw[k] = (h_ccomp[k] == 1 ? cx[h_cbody[k]] : cy[h_cbody[k]]) * 2.0 * delta / dt / (double)nsteps;
  IB_HIP(hipMemcpy(cw.p, hw.data(), sizeof(double) * nf, hipMemcpyHostToDevice));
  fill(st, adJ, 1, 0.0);
  fill(st, adot, 1, 0.0);
  fill(st, racc, g.P, 0.0);
  fill(st, raccb, g.nbc, 0.0);
  // the initial guesses of the reverse sweep's S^T solves
  if (adj_gm) fill(st, arf, nf, 0.0);
  if (adj_fsi) {
    fill(st, aav, nf, 0.0);
    fill(st, Zv, (long)std::max(nt, 1) * nf, 0.0);
  }
  const int it0 = itime;
  const int nseg = (nsteps + K - 1) / K, last0 = (nseg - 1) * K;
  // the reverse sweep moves the bodies: keep the forward end state
  const bool keep_end = nseg > 1 || adj_fsi;
  struct GrowGuard {
    bool& f;
    explicit GrowGuard(bool& x) : f(x) { f = true; }
    ~GrowGuard() { f = false; }
  } grow_guard(face_grow);
  // forward: checkpoints at the starts of the segments before the last,
  // the last segment taped
  {
    Range fwd_range("adjoint: forward pass");
    for (int n = 0; n < nsteps; ++n) {
      int rc;
      if (n >= last0) {
        rc = taped_advance(n - last0);
      } else {
        if (n % K == 0) {
          const long c = n / K;
          save_state(c_w + c * LP, c_ro + c * LP, c_qx + c * CQX, c_qy + c * CQY);
          if (adj_fsi) save_body(c_th + c * BS, c_f + c * nf);
        }
        rc = advance();
      }
      if (rc != IB_OK) return rc;
      dot_acc(st, nf, cw, f, 1.0, adJ);
    }
    if (keep_end) {
      save_state(fin_w, fin_ro, fin_qx, fin_qy);
      copy(st, fin_s, s, LP); copy(st, fin_f, f, nf); copy(st, fin_fr, f_rdst, nf);
      if (adj_fsi) save_body(fin_th, nullptr);
    }
  }
  // backward, segment by segment; the solves with S^T use S's LU factors,
  // formed here for the fixed S of rigid bodies and at every reverse step
  // (body_restore) for moving ones, and dropped when the run ends
  struct LuGuard {
    bool& f;
    explicit LuGuard(bool& x) : f(x) { f = false; }
    ~LuGuard() { f = false; }
  } lu_guard(lu_ok);
  if (adj_gm) lu_factor();
  fill(st, aw, LP, 0.0); fill(st, aro, LP, 0.0);
  fill(st, aqx, CQX, 0.0); fill(st, aqy, CQY, 0.0);
  int status = IB_OK;
  for (int sgi = nseg - 1; sgi >= 0 && status == IB_OK; --sgi) {
    const int n0 = sgi * K, n1 = std::min(nsteps, n0 + K);
    if (sgi < nseg - 1) {
      Range r("adjoint: segment recomputed");
      load_state(c_w + (long)sgi * LP, c_ro + (long)sgi * LP, c_qx + (long)sgi * CQX, c_qy + (long)sgi * CQY);
      if (adj_fsi) load_body(c_th + (long)sgi * BS, c_f + (long)sgi * nf);
      itime = it0 + n0;
      for (int n = n0; n < n1 && status == IB_OK; ++n) status = taped_advance(n - n0);
    }
    for (int n = n1 - 1; n >= n0 && status == IB_OK; --n) {
      Range r("adjoint: reverse step");
      status = adjoint_step(n - n0, it0 + n + 1 == 1);
    }
  }
  // the initial q is vort2flux of the initial omega
  if (dw0 && status == IB_OK) v2f_adj(aw, aqx, aqy);
  sum_acc(st, g.P, racc, 1.0, adot);
  sum_acc(st, g.nbc, raccb, 1.0, adot);
  if (keep_end) {
    // leave the handle in the forward end state
    load_state(fin_w, fin_ro, fin_qx, fin_qy);
    if (adj_fsi) load_body(fin_th, nullptr);
    copy(st, s, fin_s, LP); copy(st, f, fin_f, nf); copy(st, f_rdst, fin_fr, nf);
  }
  itime = it0 + nsteps;
  want_pos = want_pos && status == IB_OK;
  if (status != IB_OK) {
    sync();
    want_pos = false;
    return status;
  }
  double hv[2] = {0.0, 0.0};
  IB_HIP(hipMemcpyAsync(hv, adJ.p, sizeof(double), hipMemcpyDeviceToHost, st));
  IB_HIP(hipMemcpyAsync(hv + 1, adot.p, sizeof(double), hipMemcpyDeviceToHost, st));
  if (dw0) IB_HIP(hipMemcpyAsync(dw0, aw.p, sizeof(double) * LP, hipMemcpyDeviceToHost, st));
  sync();
  if (J) *J = hv[0];
  if (dRe) *dRe = hv[1];
  if (want_pos) {
    // per point: the sum over its x-force and y-force entries (both read
    // the point's coordinates); points in deck order.  The points of a
    // torsional body are placed by its hinge and angle: they get 0, and
    // their sum is the body's hinge gradient (ib_adjoint_hinge_grad)
    std::vector<double> gx(nf), gy(nf);
    d2h(gx.data(), gpx, nf);
    d2h(gy.data(), gpy, nf);
    sync();
    hinge_gx.assign(nbodies, 0.0);
    hinge_gy.assign(nbodies, 0.0);
    hinge_gds.assign(nbodies, 0.0);
    std::vector<double> dsw(std::max(1, ntors), 0.0);
    if (adj_fsi && ntors > 0) d2h(dsw.data(), gdsw, ntors);
    sync();
    for (int t = 0; t < ntors && adj_fsi; ++t) hinge_gds[h_cbody[h_tb[t].off]] = gDS[t] + dsw[t];
    int pt = 0;
    for (int k = 0; k < nf;) {
      int n = 0;
      while (k + n < nf && h_cbody[k + n] == h_cbody[k] && h_ccomp[k + n] == 1) ++n;
      const int body = h_cbody[k];
      const bool tors = adj_fsi && h_tors_of[k] >= 0;  // moving in this form
      for (int j = 0; j < n; ++j) {
        const double px = gx[k + j] + gx[k + n + j], py = gy[k + j] + gy[k + n + j];
        if (tors) {
          hinge_gx[body] += px;
          hinge_gy[body] += py;
        }
        if (dxb) dxb[pt + j] = tors ? 0.0 : px;
        if (dyb) dyb[pt + j] = tors ? 0.0 : py;
      }
      pt += n;
      k += 2 * n;
    }
    hinge_ok = adj_fsi;
  }
  // stiffness and pitch: per body in deck order (rigid bodies 0)
  for (int b = 0; b < nbodies; ++b) {
    if (dK) dK[b] = 0.0;
    if (dg0) dg0[b] = 0.0;
  }
  for (int t = 0; t < nt; ++t) {
    const int b = h_cbody[h_tb[t].off];
    if (dK) dK[b] = gK[t];
    if (dg0) dg0[b] = gG0[t];
  }
  want_pos = false;
  return IB_OK;
}
