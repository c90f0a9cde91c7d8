/*
 * ib_oracle.h -- CPU restatement of IB_Parallel's advance() path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation and the CPU baseline timed beside it; the product path
 * (libibgpu.so) never links, loads or calls it.
 *
 * The entry points mirror include/ib_gpu.h one to one with an orc_ prefix, so
 * a test can drive the oracle and the GPU library with the same arguments.
 */
#ifndef IB_ORACLE_H
#define IB_ORACLE_H
#include "../include/ib_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_handle orc_handle;

int orc_create(const ib_params* p, int32_t nbodies, const ib_body* bodies, orc_handle** out);
int orc_create_subdomain(const ib_params* p, const ib_subdomain* sd, int32_t nbodies, const ib_body* bodies,
                         orc_handle** out);
int orc_subdomain_info(orc_handle* h, int32_t* nb_sd, int64_t* nnz_evaluated, int64_t* nnz_kept);
int orc_op_subdomain_matrix(orc_handle* h, double* lumat);
int orc_op_subdomain_interp(orc_handle* h, int32_t* idx, double* w, double* sdx, double* sdy);
int orc_destroy(orc_handle* h);
int orc_get_sizes(const orc_handle* h, ib_sizes* out);
int orc_last_error(const orc_handle* h, char* buf, int32_t len);
int orc_advance(orc_handle* h, int32_t* itime, ib_step_info* info);
int orc_get_level_fields(orc_handle* h, int32_t k, double* omega, double* s,
                         double* qx, double* qy, double* rhs_old);
int orc_set_level_fields(orc_handle* h, int32_t k, const double* omega, const double* s,
                         const double* qx, const double* qy, const double* rhs_old);
int orc_set_body(orc_handle* h, const double* f, const double* xb, const double* yb, const double* theta,
                 const double* thetad, const double* thetadd);
int orc_pressure(orc_handle* h, double* p);
int orc_get_body(orc_handle* h, double* f, double* f_rdst, double* xb, double* yb,
                 double* theta, double* thetad, double* thetadd);
int orc_get_forces(orc_handle* h, int32_t which, double* fx, double* fy);
int orc_get_beam(orc_handle* h, int32_t k, double* u, double* ud, double* udd, int32_t* nd);
int orc_op_beam_matrices(orc_handle* h, int32_t k, double* M, double* K, double* H);
int orc_op_beam_newmark(orc_handle* h, int32_t k, int32_t nsteps, const double* fb, double* u, double* ud,
                        double* udd);
int orc_op_dst2d(orc_handle* h, int32_t batch, const double* in, double* out);
int orc_op_ctci(orc_handle* h, const double* omega, double* s);
int orc_op_ainv(orc_handle* h, const double* in, double* out);
int orc_op_vort2flux(orc_handle* h, double* omega, double* s, double* qx, double* qy);
int orc_op_reg(orc_handle* h, const double* f, double* hx, double* hy);
int orc_op_regT(orc_handle* h, const double* hx, const double* hy, const double* f_in, double* f_out);
int orc_op_schur_matvec(orc_handle* h, const double* x, double* y);
int orc_op_schur_matrix(orc_handle* h, double* S);
int orc_op_gmres(orc_handle* h, const double* b, double* x, int32_t* its);
int orc_op_nonlinear_rot(orc_handle* h, int32_t k, const double* omega, const double* qx,
                         const double* qy, const double* bc, double* rhs);

/* oracle-only helpers used by the tests */
int orc_op_ddti(orc_handle* h, const double* rhs, double* p);
int orc_op_coarsify_pressure(orc_handle* h, const double* fine, double* coarse);
/* y = b_times2(x) on torsional entries, 0 elsewhere */
int orc_op_btimes(orc_handle* h, const double* x, double* y);
/* overwrite the torsional state (one entry per torsional body) */
int orc_set_theta(orc_handle* h, const double* theta, const double* thetad, const double* thetadd);
/* one Newmark/FSI body update with force f: theta0 <- theta, compute_dx;
 * dtheta gets the increments */
int orc_op_body_update(orc_handle* h, const double* f, double* dtheta);
int orc_op_rot(orc_handle* h, const double* hx, const double* hy, double* vort);
int orc_op_curl(orc_handle* h, const double* s, const double* sbc, double* qx, double* qy);
int orc_op_coarsify(orc_handle* h, const double* fine, double* coarse);
int orc_op_get_bc(orc_handle* h, const double* coarse, double fac, double* bc);
int orc_op_apply_bc(orc_handle* h, double* vort, const double* bc, double fac);
int orc_op_weights(orc_handle* h, double* weight, int32_t* reg_ix, int32_t* reg_iy);
/* phi(r) of deltafnc for r = |x1-x2|/dr with dr = 1 (src/grid.F90:955-1003) */
double orc_deltafnc(double x1, double x2, double dr, int32_t f32_literals);
/* 1-D unnormalised DST-I (FFTW RODFT00) of `count` vectors of length N */
int orc_dst1d(int32_t N, int32_t count, const double* in, double* out);
int orc_set_threads(int32_t nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
