/*
 * ib_gpu.h -- C ABI of the MI355X immersed-boundary projection time step.
 *
 * One handle = one simulation (nested multi-domain grid + immersed bodies),
 * all state resident in HBM of the current HIP device.  Every entry point
 * returns an int status (IB_OK == 0).  Host buffers passed in are borrowed for
 * the duration of the call only.  Arrays use the reference layout: Fortran
 * column-major, i (x) fastest, with the index ranges of the reference's DMDAs:
 *
 *   omega, s, rhs_old : nodes  i = 2..m,   j = 2..n      (m-1)*(n-1) values
 *   qx (x-face flux)  : faces  i = 1..m+1, j = 1..n      (m+1)*n     values
 *   qy (y-face flux)  : faces  i = 1..m,   j = 1..n+1    m*(n+1)     values
 *   force vector f    : [body1 x(1..n1), body1 y(1..n1), body2 x, body2 y, ...]
 *
 * Reference seams replaced (file:line in nirmaljp6/IB_Parallel):
 *   ib_create           setup (src/grid.F90:117-234) + setup_variables
 *                       (src/variables.F90:84-257) + initial_condition
 *                       (src/variables.F90:262-295)
 *   ib_advance          advance(itime) (src/timestepper.F90:18-556)
 *   ib_get_level_fields / ib_set_level_fields
 *                       the per-level Vecs written by write_variables
 *                       (src/variables.F90:709-803) / read by read_variables
 *                       (src/variables.F90:807-913)
 *   ib_get_body         fvec, f_rdstvec, xbvec, ybvec, theta* (src/grid.F90:61,
 *                       src/variables.F90:205,213-214)
 *   ib_get_forces       write_force2 (src/operators_struct.F90:711-731)
 *   ib_op_*             single operators of src/myfft.F90,
 *                       src/operators_fluid.F90, src/fsinterface.F90 and the
 *                       MatShell callback ab_times (src/timestepper.F90:774-782)
 */
#ifndef IB_GPU_H
#define IB_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  IB_OK = 0,
  IB_E_INVAL = 1,          /* bad parameters / deck (e.g. m or n not divisible by 4) */
  IB_E_NOCONV_GMRES = 2,   /* Krylov solve hit gmres_max_it or diverged */
  IB_E_NOCONV_FSI = 3,     /* FSI fixed point exceeded fsi_max_count */
  IB_E_HIP = 4,            /* HIP runtime error */
  IB_E_UNSUPPORTED = 5     /* option present in the reference deck but not built */
};

/* Mirrors namelist READ_PARAMETERS (src/parameters.F90:55-59) plus solver knobs
 * the reference takes from PETSc defaults. */
typedef struct ib_params {
  int32_t m, n;            /* cells in x and y on every level (multiples of 4) */
  int32_t mgridlev;        /* number of nested grid levels */
  int32_t istart;          /* initial time index (0 = cold start) */
  double dt, re;           /* time step, Reynolds number */
  double len;              /* x-extent of the finest domain; delta = len/m */
  double offsetx, offsety; /* finest domain is [-offsetx, len-offsetx] x [-offsety, ...] */
  double fsi_tol;          /* FSI fixed point tolerance on max|dtheta| */
  double atol, rtol;       /* GMRES absolute / relative tolerance */
  int32_t gmres_restart;   /* PETSc default 30 */
  int32_t gmres_max_it;    /* PETSc default 10000 */
  int32_t fsi_max_count;   /* loop runs while count <= this (reference: 1000) */
  int32_t deltafnc_f32_literals; /* 1: exponent-less literals of deltafnc are
                                    REAL(4) as written (src/grid.F90:972-993) */
  int32_t num_stat;        /* 1: rigid bodies held fixed, the Schur system
                              solved with the precomputed inverse of its
                              Cholesky factor (src/timestepper.F90:457-484) */
  int32_t sub_domain;      /* 1: the sub-domain approach (ib_create_subdomain) */
  int32_t launch_mode;     /* 0: GMRES iterations replayed as captured HIP
                              graphs (default); 1: plain stream launches */
  int32_t schur_mode;      /* GMRES operator: 1 matrix-free ab_times per
                              Krylov iteration; 2 dense Schur matrix assembled
                              from batched ab_times columns whenever the
                              weights change (the column construction of
                              src/timestepper.F90:575-590); 3 dense Schur
                              matrix formed as W G W^T + B from the level-1
                              flux operator G on the faces near the bodies
                              (G recomputed only when a tap leaves that set);
                              0 auto: 3 with torsional bodies, else 2 */
  int32_t face_margin;     /* schur_mode 3: cells of margin around the taps'
                              faces in the face set (0: 4 cells, with
                              schur_mode 0 narrowed to 2 or 0 cells when G
                              would pass 2 GiB; -1: none) */
  int32_t gmres_form;      /* GMRES on a dense Schur matrix: 0 auto (the
                              cooperative multi-workgroup solve where it
                              fits: nf <= 1024 and restart <= 1023; else
                              one workgroup up to nf = 400; else one kernel
                              per Arnoldi stage), 1 one workgroup (nf <=
                              400, else per-stage kernels), 2 cooperative
                              where it fits, 3 per-stage kernels always,
                              4 the cooperative solve for larger systems
                              where it fits (as 0 otherwise) */
  int32_t reserved[2];
} ib_params;

/* One body deck input/body.NNN.inp (src/grid.F90:252-288), angles in radians. */
typedef struct ib_body {
  int32_t npts;
  int32_t fam;             /* 1 rigid, 2 torsional (hinged at point 1) */
  double gamma0, alpha0;   /* equilibrium angle, angle w.r.t. previous body */
  double itheta, ktheta, ctheta; /* inertia, stiffness, damping (fam 2) */
  const double *x, *y;     /* npts coordinates */
} ib_body;

typedef struct ib_step_info {
  int32_t itime;           /* time index after the step */
  int32_t ksp_iter;        /* GMRES iterations summed over FSI iterations */
  int32_t fsi_iter;        /* FSI iterations */
  int32_t status;          /* IB_OK or IB_E_NOCONV_* of this step */
  double err_fsi;          /* last max|dtheta| */
} ib_step_info;

typedef struct ib_sizes {
  int32_t m, n, mgridlev, nbc;
  int32_t n_nodes, n_xfaces, n_yfaces; /* (m-1)(n-1), (m+1)n, m(n+1) */
  int32_t nb, nf, nbodies, ntors;      /* points, forces (=2nb), bodies, torsional bodies */
} ib_sizes;

typedef struct ib_handle ib_handle;

/* --- lifetime (setup: src/grid.F90:117-234, src/variables.F90:84-295) --- */
int ib_create(const ib_params* p, int32_t nbodies, const ib_body* bodies, ib_handle** out);
int ib_destroy(ib_handle* h);
int ib_get_sizes(const ib_handle* h, ib_sizes* out);
/* Copies the message of the last failing call on h (or of ib_create when h is
 * NULL) into buf; returns its length. */
int ib_last_error(const ib_handle* h, char* buf, int32_t len);

/* --- one time step: advance(itime), src/timestepper.F90:18-556.
 * Increments *itime like the reference; info may be NULL. --- */
int ib_advance(ib_handle* h, int32_t* itime, ib_step_info* info);

/* --- state access (level k = 1..mgridlev); any pointer may be NULL --- */
int ib_get_level_fields(ib_handle* h, int32_t k, double* omega, double* s,
                        double* qx, double* qy, double* rhs_old);
int ib_set_level_fields(ib_handle* h, int32_t k, const double* omega, const double* s,
                        const double* qx, const double* qy, const double* rhs_old);
/* theta* have one entry per torsional body, in deck order */
/* calculate_pressure (src/operators_pressure.F90:110-437) at the current
 * state: the pressure coefficient on the m x n cells of every level, p =
 * mgridlev blocks of m*n values, i fastest (the write_pressure layout) */
int ib_pressure(ib_handle* h, double* p);
/* restart: read_variables' body part (src/variables.F90:871-895) followed by
 * reg_prep (:364-373): sets f (the next GMRES initial guess), the body point
 * positions and the torsional state (nt values each; any pointer may be
 * NULL to keep the current values), then rebuilds the regularisation weights
 * from the positions.  With ib_set_level_fields and params.istart this
 * resumes a run from a .var file. */
int ib_set_body(ib_handle* h, const double* f, const double* xb, const double* yb, const double* theta,
                const double* thetad, const double* thetadd);
int ib_get_body(ib_handle* h, double* f, double* f_rdst, double* xb, double* yb,
                double* theta, double* thetad, double* thetadd);
/* write_force2: per body fx = 2*delta*sum(f_x)/dt, fy likewise (drag, lift).
 * which = 0: fvec, 1: f_rdstvec. fx, fy hold nbodies entries. */
int ib_get_forces(ib_handle* h, int32_t which, double* fx, double* fy);
/* Family 3 bodies (deformable: a linear Euler-Bernoulli beam; the
 * reference declares the family and leaves its routines stubs,
 * src/operators_struct.F90:104-106,146-148,376-379,472-474): the state of
 * beam k (0-based, deck order) -- per free node (the first point is
 * clamped) the deflection normal to the deck's line and the slope, 2 ne
 * values each for ne = (npts - 1) / 2 elements; *nd gets 2 ne.  Any
 * pointer may be NULL. */
int ib_get_beam(ib_handle* h, int32_t k, double* u, double* ud, double* udd, int32_t* nd);

/* --- single operators on the handle's grid (host buffers) --- */
/* unnormalised 2-D DST-I (FFTW RODFT00 x RODFT00) of `batch` node arrays,
 * src/myfft.F90:207-216 */
int ib_op_dst2d(ib_handle* h, int32_t batch, const double* in, double* out);
/* ctci: s = dst(laminv*dst(omega)), src/myfft.F90:177-188 */
int ib_op_ctci(ib_handle* h, const double* omega, double* s);
/* ainv: vort = dst(lam1i(:,:,1)*dst(vort)/normalize), src/myfft.F90:161-173 */
int ib_op_ainv(ib_handle* h, const double* in, double* out);
/* vort2flux on all levels, src/operators_fluid.F90:237-377.  omega holds
 * mgridlev node arrays and is updated in place (coarsify); s, qx, qy out. */
int ib_op_vort2flux(ib_handle* h, double* omega, double* s, double* qx, double* qy);
/* reg = E^T (src/fsinterface.F90:14-47): hx, hy are overwritten with E^T f */
int ib_op_reg(ib_handle* h, const double* f, double* hx, double* hy);
/* regT = E (src/fsinterface.F90:51-80): f_out = f_in + E (hx, hy) */
int ib_op_regT(ib_handle* h, const double* hx, const double* hy, const double* f_in, double* f_out);
/* ab_times: y = a_times2(x) + b_times2(x), src/timestepper.F90:774-782 */
int ib_op_schur_matvec(ib_handle* h, const double* x, double* y);
/* The Schur matrix column by column, S(:, j) = ab_times(e_j), written
 * column-major (nf x nf): the construction the reference's preprocess
 * performs with a_times(e_j) before its Cholesky factorisation
 * (src/timestepper.F90:560-598), here for torsional entries too and with all
 * columns of a batch transformed at once.  With schur_mode 2 (or 0 and a
 * small enough body) the time step's GMRES runs on this matrix. */
int ib_op_schur_matrix(ib_handle* h, double* S);
/* KSPSolve(ksp_fsi, AB, b, x) (src/timestepper.F90:363-366): GMRES on the
 * Schur operator with x as the initial guess (in) and solution (out); *its
 * gets the iteration count.  Returns IB_E_NOCONV_GMRES if not converged. */
int ib_op_gmres(ib_handle* h, const double* b, double* x, int32_t* its);
/* No counterpart in the reference: ib_op_gmres for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops).  b and x are DEVICE buffers of nf doubles (x: initial guess in,
 * solution out); the solve is ib_op_gmres's one-launch dense GMRES on the
 * handle's assembled Schur matrix, with the same results, and its workspace
 * is a static 16 MiB device array per GPU (held until the process ends)
 * instead of memory the library allocates.  The call allocates nothing and
 * waits for the solve.  Returns IB_E_INVAL when the solve needs more than
 * 16 MiB of workspace, IB_E_UNSUPPORTED when the handle has no assembled
 * matrix or no one-launch solve (gmres_form 3).  All calls on one GPU share
 * the array: calls that overlap on one GPU (other threads, other streams)
 * race and must be serialised by the caller. */
int ib_op_gmres_noalloc(ib_handle* h, const double* b, double* x, int32_t* its);
/* nonlinear + rot of one level: rhs = C^T N(omega_k, q_k + q0_k, bc),
 * src/operators_fluid.F90:553-587 and :203-233; bc has nbc entries */
int ib_op_nonlinear_rot(ib_handle* h, int32_t k, const double* omega, const double* qx,
                        const double* qy, const double* bc, double* rhs);

/* --- instrumentation (no reference counterpart): per-kernel-class timing
 * with HIP events recorded around each launch on the handle's stream.
 * class 0 = DST row pass, 1 = DST column pass of the time step; 2 and 3 the
 * batched row and column passes that assemble the Schur matrix; 4 the dense
 * GMRES solve (one launch: cooperative or one workgroup).  total_bytes
 * accumulates the algorithmic bytes of the timed launches (DST passes: each
 * FP64 field array read or written once, spectral tables read once per
 * launch; GMRES: the operands of every Arnoldi step -- S, the basis vectors
 * 0..j twice for classical Gram-Schmidt, the new vector written -- plus b
 * and x). --- */
int ib_kernel_timing(ib_handle* h, int32_t kclass, int32_t enable);
/* the Schur form in use (1 matrix-free, 2 assembled by columns, 3 face-space
 * G), the size of the face set and how many times G has been computed */
int ib_schur_info(ib_handle* h, int32_t* mode, int32_t* nfaces, int64_t* green_builds);
/* the tile shapes of the handle's DST passes (no reference counterpart): rp
 * row pairs per workgroup of the row pass, cp column pairs per workgroup of
 * the column pass, and how they were chosen (0 fixed rule, 1 timed at
 * ib_create once per grid shape and process, 2 forced by the environment
 * variable IB_DST_TILES=rp,cp; IB_DST_TUNE=0 selects the fixed rule).  The
 * tiles never change results, only speed; a choice the LDS cannot hold is
 * halved at launch. */
int ib_dst_tiles(const ib_handle* h, int32_t* rp, int32_t* cp, int32_t* source);
/* the form of the handle's DST passes (no reference counterpart): forms[4]
 * for {row, column} passes x {the time step's one-field launches, the Schur
 * assembly's batched launches}, 0 the Stockham FFT form, k > 0 the FP64
 * matrix-core form (the DST-I matrix as a GEMM operand) in tile
 * configuration k-1 (of 5).  us[8] (may be null): the microseconds per field the
 * choice was timed at, us[4*batched + 2*column + matrix_core], negative
 * where not timed.  Chosen at ib_create per grid shape and process: the
 * matrix-core form only when at least 10% faster; IB_DST_FORM=fft|dense
 * forces a form.  The two forms sum in different orders (results agree to
 * rounding, not bit for bit). */
int ib_dst_form(const ib_handle* h, int32_t* forms, double* us);
int ib_kernel_timing_read(ib_handle* h, int32_t kclass, int64_t* launches, double* total_ms,
                          double* total_bytes);
/* the GMRES solve in use on a dense Schur matrix (no reference counterpart):
 * form 0 one kernel per Arnoldi stage (or a matrix-free operator), 1 one
 * workgroup, 2 the cooperative solve for nf <= 256, 3 the cooperative solve
 * for larger systems; workgroups it spans; exchange of the last form-3
 * solve: 0 none yet, 1 through one XCD's L2 (all its workgroups were found
 * on one XCD), 2 device scope (they were not). */
int ib_gmres_info(const ib_handle* h, int32_t* form, int32_t* workgroups, int32_t* exchange);

/* --- sub-domain approach: ib.inp with sub_domain = T (src/timestepper.F90:
 * 79-97,368-405,602-770; src/grid.F90:320-348,572-951; src/variables.F90:
 * 425-514; src/operators_struct.F90:838-1004; src/fsinterface.F90:84-161).
 * The Schur operator on a rectangle of sub-domain points (plus the points of
 * the one stationary rigid body unless full_rectangular) is precomputed once,
 * column j at the rows whose E E^T windows (half-width support_nz) overlap,
 * entries at or below droptol * diagonal dropped; the body points
 * interpolate from the sub-domain points (P, bilinear), and each FSI
 * iteration solves lumat f = rhsf by GMRES with lumat = P B^T P^T + body_mat.
 * params->sub_domain must be 1.  Fields left 0 take the reference's values. --- */
typedef struct ib_subdomain {
  double sdxi, sdxe, sdyi, sdye; /* the rectangle of sub-domain points */
  double delta_sd;               /* their spacing */
  int32_t full_rectangular;      /* 0: the rigid body's points are sub-domain points too */
  int32_t precomputed;           /* sub_domain_precomputed of the deck: 1 reads B from input/sd_precomputed.chd
                                    (src/timestepper.F90:753-766; IB_SD_PRECOMPUTED_FILE names another
                                    path), 0 computes B and writes that file where its folder exists
                                    (:690-695,725-728) */
  double droptol;                /* drop tolerance of B (0: 7e-3, src/grid.F90:14) */
  int32_t interp_points;         /* points of P per body entry, a square (0: 4, src/grid.F90:13) */
  int32_t support_nz;            /* half-width of the sparsity window (0: 10, src/variables.F90:440) */
} ib_subdomain;
int ib_create_subdomain(const ib_params* p, const ib_subdomain* sd, int32_t nbodies, const ib_body* bodies,
                        ib_handle** out);
/* sub-domain points, B entries evaluated, B entries kept after droptol */
int ib_subdomain_info(ib_handle* h, int32_t* nb_sd, int64_t* nnz_evaluated, int64_t* nnz_kept);
/* lumat (compute_operator, src/operators_struct.F90:838-913) at the current
 * body positions, nf x nf column-major */
int ib_op_subdomain_matrix(ib_handle* h, double* lumat);

/* --- batched form (no reference counterpart): many independent simulations
 * of uneven grid size, level count and body size advanced one step
 * together.  Simulation i is created from params[i] and the nbodies[i] body
 * decks that follow those of simulations 0..i-1 in `bodies`.  When every
 * simulation has rigid bodies and num_stat = 1, a step is one replay of a HIP
 * graph whose launches each cover all simulations (mode 1, fused; with
 * launch_mode 1 in params[0] the same launches go to the stream directly,
 * without a graph); any other
 * mix advances each simulation's plain step from its own host thread on its
 * own stream (mode 2, concurrent).  Each simulation's results are those of
 * its plain ib_advance. --- */
typedef struct ib_batch ib_batch;
int ib_batch_create(int32_t nsim, const ib_params* params, const int32_t* nbodies, const ib_body* bodies,
                    ib_batch** out);
int ib_batch_destroy(ib_batch* b);
/* one step of every simulation; itimes[i] and infos[i] (either may be NULL)
 * get simulation i's time index and step record.  Returns IB_OK or the
 * first failing simulation's status. */
int ib_batch_advance(ib_batch* b, int32_t* itimes, ib_step_info* infos);
/* simulation i as a handle for every single-simulation entry point (state
 * access, forces, pressure, operators); owned by the batch, not destroyed by
 * the caller */
int ib_batch_handle(ib_batch* b, int32_t i, ib_handle** h);
int ib_batch_info(const ib_batch* b, int32_t* nsim, int32_t* mode);
/* the DST tile rule of the fused form (no reference counterpart), timed at
 * ib_batch_create over table launches that cover every simulation: 0 each
 * simulation's own tiles (ib_dst_tiles), -1 the fixed rule, k > 0 k pairs
 * per workgroup (halved where the LDS is short); row pass and column pass */
int ib_batch_dst_tiles(const ib_batch* b, int32_t* row_rule, int32_t* col_rule);
/* message of the last failing call on b (or of ib_batch_create when b is NULL) */
int ib_batch_last_error(const ib_batch* b, char* buf, int32_t len);
/* ib_adjoint_run of every simulation, each from its own host thread on its
 * own stream (no reference counterpart): cx, cy, dJ_dK and dJ_dgamma0 run over
 * the bodies of simulation 0, then 1, ... (as in ib_batch_create); J and
 * dJ_dRe get one value per simulation.  Returns IB_OK or the first failing
 * simulation's status. */
int ib_batch_adjoint(ib_batch* b, int32_t nsteps, int32_t ckpt_every, const double* cx, const double* cy, double* J,
                     double* dJ_dRe, double* dJ_dK, double* dJ_dgamma0);

/* --- discrete adjoint (no reference counterpart): advances h by nsteps
 * steps and returns the time-mean weighted force objective
 *   J = (1/nsteps) sum over the steps of sum_b (cx[b] fx_b + cy[b] fy_b)
 * (fx, fy as ib_get_forces(which = 0) reports them after each step), its
 * derivative with respect to Re, and, when dJ_domega0 is not NULL, with
 * respect to the vorticity before the first step (mgridlev blocks of
 * (m-1)(n-1), the layout of ib_op_vort2flux's omega: the state is taken as
 * q = vort2flux(omega) of that vorticity), and, when dJ_dxb / dJ_dyb are not
 * NULL, with respect to the coordinates of every body point (deck order,
 * body after body; the Schur matrix is taken as rebuilt for the moved
 * points).  The gradients are those of the
 * discrete step (the transposed operators run backwards).  The steps are
 * taped in HBM in segments of ckpt_every steps: the last segment during the
 * forward pass, the others recomputed once from checkpoints; ckpt_every <= 0
 * tapes the whole run when that fits in half of the free HBM, else uses
 * segments of ceil(sqrt(nsteps)).  Covered: rigid bodies with num_stat = 1;
 * rigid bodies on GMRES with the assembled Schur matrix (no position
 * gradients); the face-space form (schur_mode = 3, the default with
 * torsional bodies).  Matrix-free GMRES and sub-domains return
 * IB_E_UNSUPPORTED.  h is left in the state after the nsteps steps, as
 * ib_advance would leave it. --- */
int ib_adjoint_run(ib_handle* h, int32_t nsteps, int32_t ckpt_every, const double* cx, const double* cy, double* J,
                   double* dJ_dRe, double* dJ_domega0, double* dJ_dxb, double* dJ_dyb, double* dJ_dK,
                   double* dJ_dgamma0);
/* the gradient of the last ib_adjoint_run's J with respect to the hinge
 * (xbf0, ybf0) and the point spacing ds of every torsional body (no
 * reference counterpart; the hinge is the body's first deck point and ds
 * the distance between its first two, src/grid.F90:252-288; they place its
 * points with the angle, src/operators_struct.F90:600-620): one entry per
 * body in deck order, 0 for the others; any pointer may be NULL.  A hinge
 * shift translates the body's points (the sum of their position partials);
 * ds also scales the arms, so its gradient adds the coupling vector's and
 * the structural block's terms.  Needs a run in the face-space form that
 * asked for position gradients (dJ_dxb or dJ_dyb not NULL); IB_E_INVAL
 * otherwise. */
int ib_adjoint_hinge_grad(const ib_handle* h, double* dJ_dxh, double* dJ_dyh, double* dJ_dds);

#ifdef __cplusplus
}
#endif
#endif /* IB_GPU_H */
