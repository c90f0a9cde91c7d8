// ib_solver.hip -- the C ABI of one simulation (include/ib_gpu.h) over the
// device-resident handle of ib_handle.h.
#include <cstdio>
#include <mutex>
#include <string>

#include "ib_handle.h"

// The workspace of ib_op_gmres_noalloc: 16 MiB of device memory per GPU the
// library runs on, held until the process ends (the airfoil + flap system,
// nf = 800 at restart 1000, needs 11 MiB).  Host code never passes the
// array's name to a kernel: it stands for no device address on the host.
constexpr size_t kScratchDoubles = (size_t)2 << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::string g_create_err;

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kScratchDevices] = {};

// the device address of `scratch` on device dev, looked up once per device
// with dev current
double* scratch_on(int dev) {
  if (dev < 0 || dev >= kScratchDevices) throw ibg::HipError("ib_op_gmres_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    int cur = 0;
    IB_HIP(hipGetDevice(&cur));
    if (cur != dev) IB_HIP(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) IB_HIP(hipSetDevice(cur));
    IB_HIP(e);
    g_scratch[dev] = static_cast<double*>(p);
  }
  return g_scratch[dev];
}
}  // namespace

// ------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------

#define API_TRY try {
#define API_CATCH(h)                           \
  }                                            \
  catch (const std::exception& e) {            \
    if (h) (h)->err = e.what();                \
    else g_create_err = e.what();              \
    return IB_E_HIP;                           \
  }

extern "C" {

int ib_create(const ib_params* p, int32_t nbodies, const ib_body* bodies, ib_handle** out) {
  if (!p || !out) return IB_E_INVAL;
  *out = nullptr;
  ib_handle* h = nullptr;
  try {
    h = new ib_handle();
    int rc = h->init(p, nbodies, bodies);
    if (rc != IB_OK) {
      g_create_err = h->err;
      delete h;
      return rc;
    }
  } catch (const std::exception& e) {
    g_create_err = e.what();
    delete h;
    return IB_E_HIP;
  }
  *out = h;
  return IB_OK;
}

int ib_create_subdomain(const ib_params* p, const ib_subdomain* sd, int32_t nbodies, const ib_body* bodies,
                        ib_handle** out) {
  if (!p || !sd || !out) return IB_E_INVAL;
  *out = nullptr;
  ib_handle* h = nullptr;
  try {
    h = new ib_handle();
    int rc = h->init(p, nbodies, bodies, sd);
    if (rc != IB_OK) {
      g_create_err = h->err;
      delete h;
      return rc;
    }
  } catch (const std::exception& e) {
    g_create_err = e.what();
    delete h;
    return IB_E_HIP;
  }
  *out = h;
  return IB_OK;
}

int ib_subdomain_info(ib_handle* h, int32_t* nb_sd, int64_t* nnz_evaluated, int64_t* nnz_kept) {
  if (!h) return IB_E_INVAL;
  if (nb_sd) *nb_sd = h->sd_on ? h->nb_sd : 0;
  if (nnz_evaluated) *nnz_evaluated = h->sd_nnz_eval;
  if (nnz_kept) *nnz_kept = h->sd_nnz_kept;
  return IB_OK;
}

int ib_op_subdomain_matrix(ib_handle* h, double* lumat) {
  if (!h || !lumat || !h->sd_on) return IB_E_INVAL;
  API_TRY
  h->sd_matrix();
  IB_HIP(hipMemcpyAsync(lumat, h->Sd.p, sizeof(double) * (size_t)h->nf * h->nf, hipMemcpyDeviceToHost, h->st));
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_destroy(ib_handle* h) {
  delete h;
  return IB_OK;
}

int ib_get_sizes(const ib_handle* h, ib_sizes* o) {
  if (!h || !o) return IB_E_INVAL;
  o->m = h->g.m; o->n = h->g.n; o->mgridlev = h->L; o->nbc = h->g.nbc;
  o->n_nodes = h->g.P; o->n_xfaces = h->g.Fx; o->n_yfaces = h->g.Fy;
  o->nb = h->nb; o->nf = h->nf; o->nbodies = h->nbodies; o->ntors = h->ntors;
  return IB_OK;
}

int ib_last_error(const ib_handle* h, char* buf, int32_t len) {
  const std::string& e = h ? h->err : g_create_err;
  if (buf && len > 0) std::snprintf(buf, (size_t)len, "%s", e.c_str());
  return (int)e.size();
}

int ib_advance(ib_handle* h, int32_t* itime, ib_step_info* info) {
  if (!h) return IB_E_INVAL;
  API_TRY
  int rc = h->advance();
  if (itime) *itime = h->itime;
  if (info) {
    info->itime = h->itime; info->ksp_iter = h->ksp_iter; info->fsi_iter = h->fsi_iter;
    info->status = rc; info->err_fsi = h->err_fsi;
  }
  return rc;
  API_CATCH(h)
}

int ib_get_level_fields(ib_handle* h, int32_t k, double* omega, double* s, double* qx, double* qy, double* rhs_old) {
  if (!h || k < 1 || k > h->L) return IB_E_INVAL;
  API_TRY
  const size_t P = h->g.P, q = k - 1;
  if (omega) h->d2h(omega, h->omega + q * P, P);
  if (s) h->d2h(s, h->s + q * P, P);
  if (qx) h->d2h(qx, h->qx + q * h->g.Fx, h->g.Fx);
  if (qy) h->d2h(qy, h->qy + q * h->g.Fy, h->g.Fy);
  if (rhs_old) h->d2h(rhs_old, h->rhs_old + q * P, P);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_set_level_fields(ib_handle* h, int32_t k, const double* omega, const double* s, const double* qx,
                        const double* qy, const double* rhs_old) {
  if (!h || k < 1 || k > h->L) return IB_E_INVAL;
  API_TRY
  const size_t P = h->g.P, q = k - 1;
  if (omega) h->h2d(h->omega + q * P, omega, P);
  if (s) h->h2d(h->s + q * P, s, P);
  if (qx) h->h2d(h->qx + q * h->g.Fx, qx, h->g.Fx);
  if (qy) h->h2d(h->qy + q * h->g.Fy, qy, h->g.Fy);
  if (rhs_old) h->h2d(h->rhs_old + q * P, rhs_old, P);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_get_body(ib_handle* h, double* f, double* f_rdst, double* xb, double* yb, double* theta, double* thetad,
                double* thetadd) {
  if (!h) return IB_E_INVAL;
  API_TRY
  const size_t nf = h->nf, nt = h->ntors;
  if (f && nf) h->d2h(f, h->f, nf);
  if (f_rdst && nf) h->d2h(f_rdst, h->f_rdst, nf);
  if (xb && nf) h->d2h(xb, h->xb, nf);
  if (yb && nf) h->d2h(yb, h->yb, nf);
  if (theta && nt) h->d2h(theta, h->th, nt);
  if (thetad && nt) h->d2h(thetad, h->th + nt, nt);
  if (thetadd && nt) h->d2h(thetadd, h->th + 2 * nt, nt);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

// family 3: beam k's state (no reference counterpart; the reference's
// family 3 is a stub)
int ib_get_beam(ib_handle* h, int32_t k, double* u, double* ud, double* udd, int32_t* nd) {
  if (!h || k < 0 || k >= h->nbeam) return IB_E_INVAL;
  API_TRY
  const BeamDev& b = h->beams[k];
  if (nd) *nd = b.nd;
  if (u) h->d2h(u, b.u, b.nd);
  if (ud) h->d2h(ud, b.ud, b.nd);
  if (udd) h->d2h(udd, b.udd, b.nd);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

// calculate_pressure (src/operators_pressure.F90:110-437) at the current
// state; p receives mgridlev blocks of m*n values (write_pressure layout)
int ib_pressure(ib_handle* h, double* p) {
  if (!h || !p) return IB_E_INVAL;
  API_TRY
  h->pressure_dev();
  h->d2h(p, h->pout, (size_t)h->g.m * h->g.n * h->L);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

// read_variables' body part (src/variables.F90:871-895) followed by reg_prep
// (:364-373): f, positions and torsional state, weights rebuilt from the
// positions, and the Schur operator of the active form refreshed.
int ib_set_body(ib_handle* h, const double* f, const double* xb, const double* yb, const double* theta,
                const double* thetad, const double* thetadd) {
  if (!h) return IB_E_INVAL;
  API_TRY
  const size_t nf = h->nf, nt = h->ntors;
  if (f && nf) h->h2d(h->f, f, nf);
  if (xb && nf) h->h2d(h->xb, xb, nf);
  if (yb && nf) h->h2d(h->yb, yb, nf);
  if (theta && nt) h->h2d(h->th, theta, nt);
  if (thetad && nt) h->h2d(h->th + nt, thetad, nt);
  if (thetadd && nt) h->h2d(h->th + 2 * nt, thetadd, nt);
  if (nf) {
    h->rebuild_weights();
    if (h->read_err_flag()) return h->fail(IB_E_INVAL, "body point regularisation stencil leaves the grid");
    if (h->sd_on) {
      const int rc = h->sd_build_interp();
      if (rc != IB_OK) return rc;
    } else if (h->prm.num_stat) {
      const int rc = h->num_stat_setup();
      if (rc != IB_OK) return rc;
    } else if (h->facemode) {
      h->refresh_schur();
    } else if (h->assembled) {
      h->assemble_schur();
    }
    h->trim_columns();
  }
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

// write_force2: src/operators_struct.F90:711-731
int ib_get_forces(ib_handle* h, int32_t which, double* fxo, double* fyo) {
  if (!h || !fxo || !fyo) return IB_E_INVAL;
  API_TRY
  std::vector<double> fv(std::max(1, h->nf));
  if (h->nf) h->d2h(fv.data(), which ? h->f_rdst.p : h->f.p, h->nf);
  h->sync();
  std::vector<double> ax(h->nbodies, 0.0), ay(h->nbodies, 0.0);
  for (int k = 0; k < h->nf; ++k) {
    const int b = h->h_cbody[k];
    if (h->h_ccomp[k] == 1) ax[b] = ax[b] + fv[k] / h->dt;
    else ay[b] = ay[b] + fv[k] / h->dt;
  }
  for (int b = 0; b < h->nbodies; ++b) {
    fxo[b] = 2 * h->delta * ax[b];
    fyo[b] = 2 * h->delta * ay[b];
  }
  return IB_OK;
  API_CATCH(h)
}

int ib_op_dst2d(ib_handle* h, int32_t batch, const double* in, double* out) {
  if (!h || batch < 0) return IB_E_INVAL;
  API_TRY
  const size_t P = h->g.P;
  for (int b = 0; b < batch; ++b) {
    h->h2d(h->t2, in + b * P, P);
    DstLaunch a;
    a.in[0] = h->t2; a.out[0] = h->t0;
    h->xpass(X_PLAIN, a);
    DstLaunch c;
    c.in[0] = h->t0; c.out[0] = h->t1;
    h->ypass(Y_PLAIN, c);
    h->d2h(out + b * P, h->t1, P);
    h->sync();
  }
  return IB_OK;
  API_CATCH(h)
}

int ib_op_ctci(ib_handle* h, const double* omega, double* s) {
  if (!h) return IB_E_INVAL;
  API_TRY
  const size_t P = h->g.P;
  h->h2d(h->t2, omega, P);
  h->ctci_dev(h->t2, h->t2, nullptr);
  h->d2h(s, h->t2, P);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_op_ainv(ib_handle* h, const double* in, double* out) {
  if (!h) return IB_E_INVAL;
  API_TRY
  const size_t P = h->g.P;
  h->h2d(h->t2, in, P);
  h->ainv_dev(h->t2, h->t2, false);
  h->d2h(out, h->t2, P);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_op_vort2flux(ib_handle* h, double* omega, double* s, double* qx, double* qy) {
  if (!h) return IB_E_INVAL;
  API_TRY
  const size_t P = h->g.P, L = h->L;
  h->h2d(h->vv, omega, P * L);
  h->vort2flux_dev(h->vv, h->sn, h->vx, h->vy);
  h->d2h(omega, h->vv, P * L);
  if (s) h->d2h(s, h->sn, P * L);
  if (qx) h->d2h(qx, h->vx, (size_t)h->g.Fx * L);
  if (qy) h->d2h(qy, h->vy, (size_t)h->g.Fy * L);
  fill(h->st, h->vv, (long)(P * L), 0.0);  // restore the zero coarse-level margins
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_op_reg(ib_handle* h, const double* fin, double* hxo, double* hyo) {
  if (!h) return IB_E_INVAL;
  API_TRY
  if (h->nf) h->h2d(h->opx, fin, h->nf);
  reg(h->st, h->g, h->csr, h->w, h->opx, h->hx, h->hy);
  h->d2h(hxo, h->hx, h->g.Fx);
  h->d2h(hyo, h->hy, h->g.Fy);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_op_regT(ib_handle* h, const double* hxi, const double* hyi, const double* f_in, double* f_out) {
  if (!h) return IB_E_INVAL;
  API_TRY
  h->h2d(h->vx, hxi, h->g.Fx);
  h->h2d(h->vy, hyi, h->g.Fy);
  if (h->nf) h->h2d(h->opx, f_in, h->nf);
  regT(h->st, h->g, h->nf, h->w, h->csr.key, h->vx, h->vy, 0.0, h->opx, h->gw);
  if (h->nf) h->d2h(f_out, h->gw, h->nf);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_op_schur_matvec(ib_handle* h, const double* x, double* y) {
  if (!h) return IB_E_INVAL;
  API_TRY
  if (h->nf) h->h2d(h->opx, x, h->nf);
  h->schur_dev(h->opx, h->gw);
  if (h->nf) h->d2h(y, h->gw, h->nf);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_op_schur_matrix(ib_handle* h, double* S) {
  if (!h || (!S && h->nf)) return IB_E_INVAL;
  API_TRY
  if (h->nf) {
    h->assemble_schur();
    IB_HIP(hipMemcpyAsync(S, h->Sd.p, sizeof(double) * (size_t)h->nf * h->nf, hipMemcpyDeviceToHost, h->st));
    h->trim_columns();
  }
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_op_gmres(ib_handle* h, const double* b, double* x, int32_t* its) {
  if (!h) return IB_E_INVAL;
  API_TRY
  if (h->nf) {
    h->h2d(h->rhsf, b, h->nf);
    h->h2d(h->opx, x, h->nf);
  }
  int it = 0;
  int rc = h->gmres_dev(h->rhsf, h->opx, &it);
  if (h->nf) h->d2h(x, h->opx, h->nf);
  h->sync();
  if (its) *its = it;
  return rc;
  API_CATCH(h)
}

// ib_op_gmres's one-launch dense solve on device buffers, its workspace in
// the static `scratch` array instead of memory the library allocates.
int ib_op_gmres_noalloc(ib_handle* h, const double* b, double* x, int32_t* its) {
  if (!h) return IB_E_INVAL;
  if (its) *its = 0;
  if (h->nf == 0) return IB_OK;
  if (!b || !x) return IB_E_INVAL;
  API_TRY
  if (!h->dense_solve_ok())
    return h->fail(IB_E_UNSUPPORTED, "ib_op_gmres_noalloc: no one-launch dense GMRES for this handle (gmres_form 3, "
                                     "or a system too large for it)");
  const size_t need = h->dense_ws_doubles();
  if (need > kScratchDoubles)
    return h->fail(IB_E_INVAL, "ib_op_gmres_noalloc: the solve needs " + std::to_string(need * 8) +
                                   " bytes of workspace, above the static maximum of " +
                                   std::to_string(kScratchDoubles * 8));
  if (!h->assembled)
    return h->fail(IB_E_UNSUPPORTED, "ib_op_gmres_noalloc: the handle's GMRES is matrix-free (no assembled Schur "
                                     "matrix)");
  double* ws = scratch_on(h->dev);
  auto turn = h->coop_turn();
  h->dense_launch_ws(h->Sd, b, x, ws, ws + (size_t)(h->kmax + 1) * h->nf);
  const int rc = h->dense_status();
  if (its) *its = h->hst->its;
  return rc;
  API_CATCH(h)
}

int ib_op_nonlinear_rot(ib_handle* h, int32_t k, const double* omega, const double* qx, const double* qy,
                        const double* bc, double* rhs) {
  if (!h || k < 1 || k > h->L) return IB_E_INVAL;
  API_TRY
  const size_t P = h->g.P;
  h->h2d(h->t0, omega, P);
  h->h2d(h->vx, qx, h->g.Fx);
  h->h2d(h->vy, qy, h->g.Fy);
  h->h2d(h->obc, bc, h->g.nbc);
  nl_rhs(h->st, h->g, h->t0, h->vx, h->vy, h->q0x[k - 1], h->obc, h->zbc, 0.0, 0.0, 0.0, 0, nullptr, nullptr,
         h->t1);
  h->d2h(rhs, h->t1, P);
  h->sync();
  return IB_OK;
  API_CATCH(h)
}

int ib_schur_info(ib_handle* h, int32_t* mode, int32_t* nfaces, int64_t* green_builds) {
  if (!h) return IB_E_INVAL;
  if (mode) *mode = h->facemode ? 3 : (h->assembled ? 2 : 1);
  if (nfaces) *nfaces = h->facemode ? h->nF : 0;
  if (green_builds) *green_builds = h->green_builds;
  return IB_OK;
}

int ib_gmres_info(const ib_handle* h, int32_t* form, int32_t* workgroups, int32_t* exchange) {
  if (!h) return IB_E_INVAL;
  int f = 0, P = 1;
  if (h->assembled && h->dense_solve_ok()) {
    if (h->coopP > 0) f = 2, P = h->coopP;
    else if (h->cwP > 0) f = 3, P = h->cwP;
    else f = 1;
  }
  if (form) *form = f;
  if (workgroups) *workgroups = P;
  if (exchange) *exchange = h->last_xmode;
  return IB_OK;
}

int ib_dst_tiles(const ib_handle* h, int32_t* rp, int32_t* cp, int32_t* source) {
  if (!h) return IB_E_INVAL;
  if (rp) *rp = h->tiles.rp;
  if (cp) *cp = h->tiles.cp;
  if (source) *source = h->tiles.source;
  return IB_OK;
}

int ib_dst_form(const ib_handle* h, int32_t* forms, double* us) {
  if (!h || !forms) return IB_E_INVAL;
  forms[0] = h->tiles.xform;
  forms[1] = h->tiles.yform;
  forms[2] = h->tiles.xform_b;
  forms[3] = h->tiles.yform_b;
  if (us)
    for (int q = 0; q < 8; ++q) us[q] = h->tiles.us[q];
  return IB_OK;
}

int ib_kernel_timing(ib_handle* h, int32_t kclass, int32_t enable) {
  if (!h || kclass < 0 || kclass > 4) return IB_E_INVAL;
  API_TRY
  h->sync();
  h->tim[kclass].on = enable != 0;
  h->tim[kclass].count = 0;
  h->tim[kclass].total_ms = 0;
  h->tim[kclass].total_bytes = 0;
  return IB_OK;
  API_CATCH(h)
}

int ib_adjoint_run(ib_handle* h, int32_t nsteps, int32_t ckpt_every, const double* cx, const double* cy, double* J,
                   double* dJ_dRe, double* dJ_domega0, double* dJ_dxb, double* dJ_dyb, double* dJ_dK,
                   double* dJ_dgamma0) {
  if (!h || !cx || !cy) return IB_E_INVAL;
  API_TRY
  return h->adjoint_run(nsteps, ckpt_every, cx, cy, J, dJ_dRe, dJ_domega0, dJ_dxb, dJ_dyb, dJ_dK, dJ_dgamma0);
  API_CATCH(h)
}

int ib_adjoint_hinge_grad(const ib_handle* h, double* dJ_dxh, double* dJ_dyh, double* dJ_dds) {
  if (!h) return IB_E_INVAL;
  if (!h->hinge_ok) return IB_E_INVAL;
  for (int b = 0; b < h->nbodies; ++b) {
    if (dJ_dxh) dJ_dxh[b] = h->hinge_gx[b];
    if (dJ_dyh) dJ_dyh[b] = h->hinge_gy[b];
    if (dJ_dds) dJ_dds[b] = h->hinge_gds[b];
  }
  return IB_OK;
}

int ib_kernel_timing_read(ib_handle* h, int32_t kclass, int64_t* launches, double* total_ms, double* total_bytes) {
  if (!h || kclass < 0 || kclass > 4) return IB_E_INVAL;
  API_TRY
  h->sync();
  if (launches) *launches = h->tim[kclass].count;
  if (total_ms) *total_ms = h->tim[kclass].total_ms;
  if (total_bytes) *total_bytes = h->tim[kclass].total_bytes;
  return IB_OK;
  API_CATCH(h)
}

}  // extern "C"
