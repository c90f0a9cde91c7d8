/*
 * ib_oracle.cpp -- FP64 CPU restatement of IB_Parallel's advance(itime) path.
 *
 * TEST INFRASTRUCTURE ONLY: parity checker and CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Every routine names the reference routine (file:line in nirmaljp6/IB_Parallel)
 * whose arithmetic and loop order it follows.  Single-process semantics: one
 * "rank" owns the whole grid, so PETSc VecSetValues(ADD_VALUES) is a sequential
 * add in insertion order and every DMDA ghost at a physical boundary stays 0
 * unless the reference writes it explicitly.
 *
 * The unnormalised DST-I (FFTW RODFT00) is computed through a complex FFT of
 * length 2M (M = N+1) of an odd-extended pair of real vectors: with
 * z = x_a + i x_b extended oddly, Z = -2i S(x_a) + 2 S(x_b), so
 * RODFT00(x_a) = -Im Z and RODFT00(x_b) = Re Z.  The FFT is a mixed-radix
 * Stockham autosort (radices 4, 2, 3, 5, 7 and a generic odd radix).
 *
 * Parity pins (tests/test_oracle_*.py): DST-I against scipy.fft.dst golden
 * vectors and brute force; ctci/ainv against the 5-point operator they invert;
 * rot/curl adjointness and C^T C = 5-point Laplacian; reg/regT adjointness;
 * deltafnc partition of unity, first moment and continuity; coarsify/get_bc on
 * constant and bilinear fields; GMRES against a dense LAPACK solve of the
 * matrix assembled column by column from the matvec; torsional B rank-1 and the
 * free-spring closed form.
 */
#include "ib_oracle.h"

#include <algorithm>
#include <functional>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cplx;

/* ------------------------------------------------------------------------ */
/* FFT / DST-I                                                               */
/* ------------------------------------------------------------------------ */

std::vector<int> factorize(int L) {
  std::vector<int> r;
  const int small[] = {4, 2, 3, 5, 7};
  for (int q : small)
    while (L % q == 0) { r.push_back(q); L /= q; }
  for (int p = 11; L > 1; p += 2)
    while (L % p == 0) { r.push_back(p); L /= p; }
  return r;
}

struct FFTPlan {
  int L = 0;
  std::vector<int> rad;
  std::vector<cplx> W; /* W[t] = exp(-2 pi i t / L) */
  void init(int L_) {
    L = L_;
    rad = factorize(L);
    W.resize(L);
    for (int t = 0; t < L; ++t) {
      long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)L;
      W[t] = cplx((double)cosl(a), (double)-sinl(a));
    }
  }
  /* Stockham autosort; result lands in a (tmp is scratch of length L). */
  void exec(cplx* a, cplx* tmp) const {
    cplx* in = a;
    cplx* out = tmp;
    int Ns = 1;
    cplx v[128], V[128];
    std::vector<cplx> vbig, Vbig;
    for (int R : rad) {
      cplx* vv = v;
      cplx* VV = V;
      if (R > 128) { vbig.resize(R); Vbig.resize(R); vv = vbig.data(); VV = Vbig.data(); }
      const int LR = L / R;
      const int tws = L / (Ns * R);
      for (int j = 0; j < LR; ++j) {
        const int k = j % Ns;
        for (int r = 0; r < R; ++r) vv[r] = in[j + r * LR] * W[(k * r * tws) % L];
        for (int s = 0; s < R; ++s) {
          cplx acc = 0.0;
          for (int r = 0; r < R; ++r) acc += vv[r] * W[((r * s) % R) * LR];
          VV[s] = acc;
        }
        const int base = (j / Ns) * Ns * R + k;
        for (int s = 0; s < R; ++s) out[base + s * Ns] = VV[s];
      }
      Ns *= R;
      std::swap(in, out);
    }
    if (in != a) std::memcpy(a, in, sizeof(cplx) * L);
  }
};

/* RODFT00 of length N on two strided real vectors at once (xb may be null). */
struct DST1 {
  int N = 0, M = 0;
  FFTPlan plan;
  void init(int N_) { N = N_; M = N + 1; plan.init(2 * M); }
  void pair(const double* xa, const double* xb, long stride, double* ya, double* yb,
            long ostride, cplx* z, cplx* tmp) const {
    z[0] = 0.0;
    z[M] = 0.0;
    for (int t = 1; t <= N; ++t) {
      cplx c(xa[(long)(t - 1) * stride], xb ? xb[(long)(t - 1) * stride] : 0.0);
      z[t] = c;
      z[2 * M - t] = -c;
    }
    plan.exec(z, tmp);
    for (int k = 1; k <= N; ++k) {
      ya[(long)(k - 1) * ostride] = -z[k].imag();
      if (yb) yb[(long)(k - 1) * ostride] = z[k].real();
    }
  }
};

int g_threads = 1;

/* ------------------------------------------------------------------------ */
/* simulation state                                                          */
/* ------------------------------------------------------------------------ */

struct Body {
  int npts = 0, fam = 1;
  double gamma0 = 0, alpha0 = 0, itheta = 0, ktheta = 0, ctheta = 0;
  double deltas = 0, xbf0 = 0, ybf0 = 0;
  std::vector<double> x, y;
};

/* Family 3, a deformable body.  The reference declares it and leaves every
 * routine a stub (src/operators_struct.F90:104-106, 146-148, 376-379,
 * 472-474: "evaluate udd"), so this is new physics in the torsional body's
 * frame: a linear Euler-Bernoulli beam along the deck's straight line of
 * npts = 2 ne + 1 points, ne Hermite-cubic elements (nodes at the even
 * points), clamped at the first point; u holds each free node's deflection
 * normal to the line and its slope.  Mass per unit length itheta, bending
 * stiffness EI ktheta, damping C = ctheta M.  The coupling repeats the
 * torsional body's with theta -> u and BR -> Phi = n H (H_p,i the shape
 * functions at point p; Phi u moves the points along the normal n):
 * Newmark average acceleration with K_eff = 4/dt^2 M + 2/dt C + K, the
 * generalised force Phi^T f delta/dt, b_times2 = 2 delta^2/dt^2 Phi
 * K_eff^-1 Phi^T, compute_rhs = delta Phi (ud0 + 2/dt (u0 - u) - Jr). */
struct Beam {
  int body = 0, np = 0, ne = 0, nd = 0;
  double L = 0, h = 0, m = 0, EI = 0, c = 0;
  double nx = 0, ny = 1;
  std::vector<double> X0, Y0;
  std::vector<double> M, K, C, Keff, Keffi, Minv; /* nd x nd, row-major */
  std::vector<double> H;                          /* np x nd, row-major */
  std::vector<double> G;                          /* np x np: H K_eff^-1 H^T */
  std::vector<double> u, ud, udd, u0, ud0, udd0;
};

/* inverse of a dense n x n matrix (row-major) by Gauss-Jordan elimination
 * with partial pivoting; false if singular */
bool invert(int n, std::vector<double> a, std::vector<double>& inv) {
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = r;
    if (a[(size_t)piv * n + c] == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)piv * n + k], a[(size_t)c * n + k]);
        std::swap(inv[(size_t)piv * n + k], inv[(size_t)c * n + k]);
      }
    const double d = 1.0 / a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] *= d;
      inv[(size_t)c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double fct = a[(size_t)r * n + c];
      if (fct == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= fct * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= fct * inv[(size_t)c * n + k];
      }
    }
  }
  return true;
}

/* the beam's matrices and shape functions for time step dt */
bool beam_build(Beam& be, const Body& b, double dt) {
  be.np = b.npts;
  be.ne = (b.npts - 1) / 2;
  be.nd = 2 * be.ne;
  const double lx = b.x[b.npts - 1] - b.x[0], ly = b.y[b.npts - 1] - b.y[0];
  be.L = std::sqrt(lx * lx + ly * ly);
  be.h = be.L / be.ne;
  be.nx = -ly / be.L;
  be.ny = lx / be.L;
  be.m = b.itheta;
  be.EI = b.ktheta;
  be.c = b.ctheta;
  be.X0 = b.x;
  be.Y0 = b.y;
  const int nd = be.nd;
  const double h = be.h;
  be.M.assign((size_t)nd * nd, 0.0);
  be.K.assign((size_t)nd * nd, 0.0);
  const double ke[4][4] = {{12, 6 * h, -12, 6 * h},
                           {6 * h, 4 * h * h, -6 * h, 2 * h * h},
                           {-12, -6 * h, 12, -6 * h},
                           {6 * h, 2 * h * h, -6 * h, 4 * h * h}};
  const double me[4][4] = {{156, 22 * h, 54, -13 * h},
                           {22 * h, 4 * h * h, 13 * h, -3 * h * h},
                           {54, 13 * h, 156, -22 * h},
                           {-13 * h, -3 * h * h, -22 * h, 4 * h * h}};
  const double kf = be.EI / (h * h * h), mf = be.m * h / 420.0;
  /* global index of local dof a of element e (node 0 clamped: -1) */
  auto gdof = [](int e, int a) { const int node = e + a / 2; return node == 0 ? -1 : 2 * (node - 1) + a % 2; };
  for (int e = 0; e < be.ne; ++e)
    for (int a = 0; a < 4; ++a)
      for (int bb = 0; bb < 4; ++bb) {
        const int ga = gdof(e, a), gb = gdof(e, bb);
        if (ga < 0 || gb < 0) continue;
        be.K[(size_t)ga * nd + gb] += kf * ke[a][bb];
        be.M[(size_t)ga * nd + gb] += mf * me[a][bb];
      }
  be.C.assign((size_t)nd * nd, 0.0);
  for (size_t i = 0; i < be.C.size(); ++i) be.C[i] = be.c * be.M[i];
  be.Keff.assign((size_t)nd * nd, 0.0);
  for (size_t i = 0; i < be.Keff.size(); ++i) be.Keff[i] = 4.0 / (dt * dt) * be.M[i] + 2.0 / dt * be.C[i] + be.K[i];
  if (!invert(nd, be.Keff, be.Keffi) || !invert(nd, be.M, be.Minv)) return false;
  be.H.assign((size_t)be.np * nd, 0.0);
  for (int p = 0; p < be.np; ++p) {
    const int e = std::min(p / 2, be.ne - 1);
    const double xi = 0.5 * (double)(p - 2 * e);
    const double N[4] = {1 - 3 * xi * xi + 2 * xi * xi * xi, h * (xi - 2 * xi * xi + xi * xi * xi),
                         3 * xi * xi - 2 * xi * xi * xi, h * (-xi * xi + xi * xi * xi)};
    for (int a = 0; a < 4; ++a) {
      const int ga = gdof(e, a);
      if (ga >= 0) be.H[(size_t)p * nd + ga] += N[a];
    }
  }
  be.G.assign((size_t)be.np * be.np, 0.0);
  std::vector<double> t(nd);
  for (int q = 0; q < be.np; ++q) {
    for (int i = 0; i < nd; ++i) {
      double acc = 0.0;
      for (int k = 0; k < nd; ++k) acc += be.Keffi[(size_t)i * nd + k] * be.H[(size_t)q * nd + k];
      t[i] = acc;
    }
    for (int p = 0; p < be.np; ++p) {
      double acc = 0.0;
      for (int i = 0; i < nd; ++i) acc += be.H[(size_t)p * nd + i] * t[i];
      be.G[(size_t)p * be.np + q] = acc;
    }
  }
  be.u.assign(nd, 0.0);
  be.ud = be.u; be.udd = be.u; be.u0 = be.u; be.ud0 = be.u; be.udd0 = be.u;
  return true;
}

}  // namespace

struct orc_handle {
  ib_params prm;
  std::string err;
  int m = 0, n = 0, L = 0;
  double delta = 0, dt = 0, Re = 0, len = 0, offx = 0, offy = 0, pi = 0;
  int P = 0, Fx = 0, Fy = 0, nbc = 0, left = 0, right = 0, bottom = 0, top = 0;
  double normalize = 0;
  std::vector<double> laminv, lam1, lam1i; /* lam1, lam1i: L consecutive P-blocks */
  std::vector<double> vfac, con1, con2;
  DST1 dx, dy;
  DST1 dxp, dyp;                   /* pressure: cells 1..m x 1..n */
  std::vector<double> laminv_ddti; /* m*n */
  /* state, level index 0..L-1 (= reference k-1) */
  std::vector<std::vector<double>> omega, s, qx, qy, rhs_old, q0x, q0y;
  /* bodies */
  std::vector<Body> bdy;
  int nb = 0, nf = 0;
  std::vector<int> nb_ind, nbb_ind; /* nbb_ind[b] = end offset of body b in fvec (nbb_ind[0] = 0) */
  std::vector<int> cbody, ccomp;    /* per force entry: body (0-based), component 1=x 2=y */
  std::vector<double> xb, yb, f, f_rdst;
  std::vector<int> tors; /* torsional bodies in deck order */
  std::vector<double> theta, thetad, thetadd, theta0, thetad0, thetadd0, dtheta, Mt, Jt, Ct, Ftint;
  std::vector<Beam> beams; /* family-3 bodies in deck order */
  /* regularisation (src/grid.F90:36-40) */
  int support2 = 3;
  std::vector<double> weight; /* 49 per force entry */
  std::vector<int> reg_ix, reg_iy;
  int itime = 0;
  int ksp_iter = 0, fsi_iter = 0;
  double err_fsi = 0;

  /* ---------------- index helpers ---------------- */
  inline long nd(int i, int j) const { return (long)(i - 2) + (long)(m - 1) * (j - 2); }
  inline long fx(int i, int j) const { return (long)(i - 1) + (long)(m + 1) * (j - 1); }
  inline long fy(int i, int j) const { return (long)(i - 1) + (long)m * (j - 1); }

  int fail(int code, const std::string& msg) { err = msg; return code; }

  /* ---------------- spectral operators (src/myfft.F90) ---------------- */

  /* dst: src/myfft.F90:207-216; FFTW 2-D RODFT00 on (m-1) x (n-1) */
  void dst2d(double* a) const {
    const int N1 = m - 1, N2 = n - 1;
    /* x pass: rows j (contiguous, length N1), two rows per transform */
#pragma omp parallel num_threads(g_threads)
    {
      std::vector<cplx> z(2 * std::max(dx.M, dy.M)), t(2 * std::max(dx.M, dy.M));
#pragma omp for schedule(static)
      for (int jp = 0; jp < (N2 + 1) / 2; ++jp) {
        int ja = 2 * jp, jb = ja + 1;
        double* ra = a + (long)ja * N1;
        double* rb = jb < N2 ? a + (long)jb * N1 : nullptr;
        dx.pair(ra, rb, 1, ra, rb, 1, z.data(), t.data());
      }
      /* y pass: columns i (stride N1, length N2) */
#pragma omp for schedule(static)
      for (int ip = 0; ip < (N1 + 1) / 2; ++ip) {
        int ia = 2 * ip, ib = ia + 1;
        double* ca = a + ia;
        double* cb = ib < N1 ? a + ib : nullptr;
        dy.pair(ca, cb, N1, ca, cb, N1, z.data(), t.data());
      }
    }
  }

  /* ctci: src/myfft.F90:177-188 */
  void ctci(const double* w, double* out) const {
    std::vector<double> in(w, w + P);
    dst2d(in.data());
    for (int q = 0; q < P; ++q) in[q] = laminv[q] * in[q];
    dst2d(in.data());
    std::memcpy(out, in.data(), sizeof(double) * P);
  }

  /* ainv: src/myfft.F90:161-173 */
  void ainv(double* v) const {
    dst2d(v);
    for (int q = 0; q < P; ++q) v[q] = lam1i[q] * v[q] / normalize;
    dst2d(v);
  }

  /* ---------------- stencils (src/operators_fluid.F90) ---------------- */

  /* rot_int + rot_bc: src/operators_fluid.F90:203-233 */
  void rot(const double* hx, const double* hy, double* vort) const {
    for (int j = 2; j <= n; ++j)
      for (int i = 2; i <= m; ++i)
        vort[nd(i, j)] = hy[fy(i, j)] - hy[fy(i - 1, j)] - hx[fx(i, j)] + hx[fx(i, j - 1)];
  }

  /* node value with boundary ghosts taken from a bc vector, in the order the
   * reference overwrites them (left, right, bottom, top; corners end up from
   * bottom/top): src/operators_fluid.F90:448-451 and :563-566 */
  inline double ext(const double* a, const double* bc, int i, int j) const {
    if (j == 1) return bc[bottom + i - 1];
    if (j == n + 1) return bc[top + i - 1];
    if (i == 1) return bc[left + j - 1];
    if (i == m + 1) return bc[right + j - 1];
    return a[nd(i, j)];
  }

  /* curl_int + curl_bc: src/operators_fluid.F90:423-471 */
  void curl(const double* sv, const double* sbc, double* qxo, double* qyo) const {
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= m + 1; ++i)
        qxo[fx(i, j)] = ext(sv, sbc, i, j + 1) - ext(sv, sbc, i, j);
    for (int j = 1; j <= n + 1; ++j)
      for (int i = 1; i <= m; ++i)
        qyo[fy(i, j)] = ext(sv, sbc, i, j) - ext(sv, sbc, i + 1, j);
  }

  /* coarsify: src/operators_fluid.F90:381-419 (fine odd nodes I,J = 3..m-1) */
  void coarsify(const double* wf, double* wc) const {
    for (int J = 3; J <= n - 1; J += 2)
      for (int I = 3; I <= m - 1; I += 2) {
        int ic = m / 4 + (I + 1) / 2, jc = n / 4 + (J + 1) / 2;
        double v = wf[nd(I, J)] +
                   0.5 * (wf[nd(I + 1, J)] + wf[nd(I, J + 1)] + wf[nd(I - 1, J)] + wf[nd(I, J - 1)]) +
                   0.25 * (wf[nd(I + 1, J + 1)] + wf[nd(I + 1, J - 1)] + wf[nd(I - 1, J - 1)] +
                           wf[nd(I - 1, J + 1)]);
        wc[nd(ic, jc)] = v;
      }
  }

  /* get_bc: src/operators_fluid.F90:475-535 with the slot map of
   * setup_getbc (src/grid.F90:1091-1140) */
  void get_bc(const double* sc, double* bc, double fac) const {
    const int i0 = m / 4 + 1, i1 = 3 * m / 4 + 1, j0 = n / 4 + 1, j1 = 3 * n / 4 + 1;
    for (int i = i0; i <= i1; ++i) {
      int slot = 2 * (i - 1 - m / 4);
      bc[bottom + slot] = sc[nd(i, j0)] * fac;
      bc[top + slot] = sc[nd(i, j1)] * fac;
      if (i != i1) {
        bc[bottom + slot + 1] = (0.5 * (sc[nd(i, j0)] + sc[nd(i + 1, j0)])) * fac;
        bc[top + slot + 1] = (0.5 * (sc[nd(i, j1)] + sc[nd(i + 1, j1)])) * fac;
      }
    }
    for (int j = j0; j <= j1; ++j) {
      int slot = 2 * (j - 1 - n / 4);
      bc[left + slot] = sc[nd(i0, j)] * fac;
      bc[right + slot] = sc[nd(i1, j)] * fac;
      if (j != j1) {
        bc[left + slot + 1] = (0.5 * (sc[nd(i0, j)] + sc[nd(i0, j + 1)])) * fac;
        bc[right + slot + 1] = (0.5 * (sc[nd(i1, j)] + sc[nd(i1, j + 1)])) * fac;
      }
    }
  }

  /* apply_bc: src/operators_fluid.F90:539-549 */
  void apply_bc(double* v, const double* bc, double fac) const {
    for (int j = 2; j <= n; ++j) v[nd(2, j)] = v[nd(2, j)] + fac * bc[left + j - 1];
    for (int j = 2; j <= n; ++j) v[nd(m, j)] = v[nd(m, j)] + fac * bc[right + j - 1];
    for (int i = 2; i <= m; ++i) v[nd(i, 2)] = v[nd(i, 2)] + fac * bc[bottom + i - 1];
    for (int i = 2; i <= m; ++i) v[nd(i, n)] = v[nd(i, n)] + fac * bc[top + i - 1];
  }

  /* nonlinear: src/operators_fluid.F90:553-587.  qhx/qhy are q + q0 on the
   * DMDA interior; their physical-boundary ghosts are 0 (never written). */
  void nonlinear(const double* w, const double* qhx, const double* qhy, const double* bc,
                 double* nlx, double* nly) const {
    auto QX = [&](int i, int j) -> double { return (j < 1 || j > n) ? 0.0 : qhx[fx(i, j)]; };
    auto QY = [&](int i, int j) -> double { return (i < 1 || i > m) ? 0.0 : qhy[fy(i, j)]; };
    auto uavg = [&](int i, int j) { return 0.5 * (QX(i, j) + QX(i, j - 1)); };
    auto vavg = [&](int i, int j) { return 0.5 * (QY(i, j) + QY(i - 1, j)); };
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= m + 1; ++i)
        nlx[fx(i, j)] = 0.5 * (vavg(i, j + 1) * ext(w, bc, i, j + 1) + vavg(i, j) * ext(w, bc, i, j));
    for (int j = 1; j <= n + 1; ++j)
      for (int i = 1; i <= m; ++i)
        nly[fy(i, j)] = -0.5 * (uavg(i + 1, j) * ext(w, bc, i + 1, j) + uavg(i, j) * ext(w, bc, i, j));
  }

  /* fluid_predictor: src/operators_fluid.F90:649-663 (level index k0 = k-1) */
  void predictor(double* w, const double* rhs, const double* rold, const double* rhsbc, int k0) const {
    std::vector<double> a(P), b(w, w + P);
    for (int q = 0; q < P; ++q) a[q] = con1[k0] * rhs[q] + con2[k0] * rold[q] + rhsbc[q];
    dst2d(a.data());
    dst2d(b.data());
    const double* l1 = &lam1[(long)k0 * P];
    const double* l1i = &lam1i[(long)k0 * P];
    for (int q = 0; q < P; ++q) a[q] = l1i[q] * (a[q] + l1[q] * b[q]);
    dst2d(a.data());
    std::memcpy(w, a.data(), sizeof(double) * P);
  }

  /* vort2flux: src/operators_fluid.F90:237-377.  w is modified by coarsify. */
  void vort2flux(std::vector<std::vector<double>>& w, std::vector<std::vector<double>>& so,
                 std::vector<std::vector<double>>& qxo, std::vector<std::vector<double>>& qyo) const {
    for (int k = 1; k < L; ++k) coarsify(w[k - 1].data(), w[k].data());
    std::vector<double> sbc(nbc, 0.0), wk(P);
    ctci(w[L - 1].data(), so[L - 1].data());
    curl(so[L - 1].data(), sbc.data(), qxo[L - 1].data(), qyo[L - 1].data());
    if (L > 1) get_bc(so[L - 1].data(), sbc.data(), 1.0);
    for (int k = L - 2; k >= 0; --k) {
      wk.assign(w[k].begin(), w[k].end());
      apply_bc(wk.data(), sbc.data(), 1.0);
      ctci(wk.data(), so[k].data());
      curl(so[k].data(), sbc.data(), qxo[k].data(), qyo[k].data());
      if (k > 0) get_bc(so[k].data(), sbc.data(), 1.0);
    }
  }

  /* ---------------- IB coupling (src/grid.F90, src/fsinterface.F90) -------- */

  /* deltafnc: src/grid.F90:955-1003 */
  static double deltafnc(double x1, double x2, double dr, bool f32) {
    double r = std::fabs(x1 - x2);
    double r1 = r / dr, r2 = r1 * r1, r3 = r2 * r1, r4 = r3 * r1;
    double d;
    /* Exponent-less literals are default REAL(4) in the reference source;
     * F(x) yields that value (f32) or the decimal value as double. */
#define F(lit) (f32 ? (double)(lit##f) : (lit))
    if (r1 <= 1.0) {
      double a5 = std::asin((1.0 / 2.0) * std::sqrt(3.0) * (2.0 * r1 - 1.0));
      double a8 = std::sqrt(1.0 - 12.0 * r2 + 12.0 * r1);
      d = 0.4166666667e-1 * r4 + (F(-.1388888889) + 0.3472222222e-1 * a8) * r3 +
          (-0.7121664902e-1 - 0.5208333333e-1 * a8 + F(0.2405626122) * a5) * r2 +
          (F(-.2405626122) * a5 + F(-.3792313933) + F(.1012731481) * a8) * r1 + 0.8018753741e-1 * a5 -
          0.4195601852e-1 * a8 + F(.6485698427);
    } else if (r1 <= 2.0) {
      double a6 = std::asin((1.0 / 2.0) * std::sqrt(3.0) * (-3.0 + 2.0 * r1));
      double a9 = std::sqrt(-23.0 + 36.0 * r1 - 12.0 * r2);
      d = -0.6250000000e-1 * r4 + (F(.4861111111) - 0.1736111111e-1 * a9) * r3 +
          (F(-1.143175026) + 0.7812500000e-1 * a9 + F(-.1202813061) * a6) * r2 +
          (F(.8751991178) + F(.3608439183) * a6 + F(-.1548032407) * a9) * r1 + F(-.2806563809) * a6 +
          0.822848104e-2 + F(.1150173611) * a9;
    } else if (r1 <= 3.0) {
      double a1 = std::asin((1.0 / 2.0 * (2.0 * r1 - 5.0)) * std::sqrt(3.0));
      double a7 = std::sqrt(-71.0 - 12.0 * r2 + 60.0 * r1);
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

  /* setup_reg: src/grid.F90:459-528 */
  int setup_reg() {
    const int s2 = support2, nt = (2 * s2 + 1) * (2 * s2 + 1);
    const double d2 = 0.5 * delta;
    const bool f32 = prm.deltafnc_f32_literals != 0;
    weight.assign((size_t)nt * nf, 0.0);
    reg_ix.clear();
    reg_iy.clear();
    for (int k = 0; k < nf; ++k) {
      int ix = (int)((xb[k] + offx) / delta) + 1;
      int iy = (int)((yb[k] + offy) / delta) + 1;
      int next = 0;
      for (int l = -s2; l <= s2; ++l)
        for (int p = -s2; p <= s2; ++p) {
          double x, y;
          if (ccomp[k] == 1) {
            x = delta * (ix - 1 + p) - offx;
            y = delta * (iy - 1 + l) - offy + d2;
            int ii = ix + p, jj = iy + l; /* face (ii, jj) */
            if (ii < 1 || ii > m + 1 || jj < 1 || jj > n)
              return fail(IB_E_INVAL, "body point regularisation stencil leaves the x-face grid");
            reg_ix.push_back((m + 1) * (iy + l - 1) + (ix + p - 1));
          } else {
            x = delta * (ix - 1 + p) - offx + d2;
            y = delta * (iy - 1 + l) - offy;
            int ii = ix + p, jj = iy + l;
            if (ii < 1 || ii > m || jj < 1 || jj > n + 1)
              return fail(IB_E_INVAL, "body point regularisation stencil leaves the y-face grid");
            reg_iy.push_back(m * (iy + l - 1) + (ix + p - 1));
          }
          weight[(size_t)k * nt + next] = delta * delta * deltafnc(x, xb[k], delta, f32) * deltafnc(y, yb[k], delta, f32);
          ++next;
        }
    }
    return IB_OK;
  }

  /* reg: src/fsinterface.F90:14-47; hx, hy accumulate (caller zeroes) */
  void reg(const double* h0, double* hx, double* hy) const {
    const int nt = (2 * support2 + 1) * (2 * support2 + 1);
    size_t itx = 0, ity = 0;
    for (int k = 0; k < nf; ++k)
      for (int t = 0; t < nt; ++t) {
        double v = weight[(size_t)k * nt + t] * h0[k];
        if (ccomp[k] == 1) { hx[reg_ix[itx]] += v; ++itx; }
        else { hy[reg_iy[ity]] += v; ++ity; }
      }
  }

  /* regT: src/fsinterface.F90:51-80; h0 accumulates */
  void regT(const double* hx, const double* hy, double* h0) const {
    const int nt = (2 * support2 + 1) * (2 * support2 + 1);
    size_t itx = 0, ity = 0;
    for (int k = 0; k < nf; ++k)
      for (int t = 0; t < nt; ++t) {
        if (ccomp[k] == 1) { h0[k] = h0[k] + weight[(size_t)k * nt + t] * hx[reg_ix[itx]]; ++itx; }
        else { h0[k] = h0[k] + weight[(size_t)k * nt + t] * hy[reg_iy[ity]]; ++ity; }
      }
  }

  /* ---------------- torsional bodies (src/operators_struct.F90) ---------- */

  /* BR_build: src/operators_struct.F90:319-338 (j = 1..2*npts within body) */
  double BRval(const Body& b, int j, double th) const {
    if (j <= b.npts) return (j - 1) * b.deltas * std::sin(b.gamma0 - th);
    return -(j - b.npts - 1) * b.deltas * std::cos(b.gamma0 - th);
  }

  /* b_times2: src/operators_struct.F90:14-127 (rank-1 dense form evaluated
   * row by row as in MatMult of sol_Jmat) */
  void b_times2(const double* x, double* y) const {
    for (int t = 0; t < (int)tors.size(); ++t) {
      const Body& b = bdy[tors[t]];
      const int off = nbb_ind[tors[t]];
      const int nn = 2 * b.npts;
      const double si = std::sin(b.gamma0 - theta[t]), ci = std::cos(b.gamma0 - theta[t]);
      const double den = 4.0 / std::pow(dt, 2.0) * Mt[t] + 2.0 / dt * Ct[t] + Jt[t];
      const double sc = (delta * delta / b.deltas / dt) * (2 * b.deltas / dt);
      for (int j = 1; j <= nn; ++j) {
        const int im = j - 1;
        double acc = 0.0;
        for (int k = 1; k <= nn; ++k) {
          const int in = k - 1;
          double v;
          if (j <= b.npts) {
            if (k <= b.npts) v = (double)(im * in) * (si * si) * (b.deltas * b.deltas);
            else v = -(double)(im * (in - b.npts)) * si * ci * (b.deltas * b.deltas);
          } else {
            if (k <= b.npts) v = -(double)((im - b.npts) * in) * si * ci * (b.deltas * b.deltas);
            else v = (double)((im - b.npts) * (in - b.npts)) * (ci * ci) * (b.deltas * b.deltas);
          }
          v = v / den;
          acc += v * x[off + k - 1];
        }
        y[off + j - 1] = acc * sc;
      }
    }
    /* family 3: 2 delta^2/dt^2 Phi K_eff^-1 Phi^T, Phi = n H */
    const double cb = 2.0 * delta * delta / (dt * dt);
    for (const Beam& be : beams) {
      const int off = nbb_ind[be.body], np = be.np;
      for (int j = 0; j < 2 * np; ++j) {
        const int pj = j % np;
        const double nj = j < np ? be.nx : be.ny;
        double acc = 0.0;
        for (int k = 0; k < 2 * np; ++k) {
          const int pk = k % np;
          const double nk = k < np ? be.nx : be.ny;
          acc += nj * nk * be.G[(size_t)pj * np + pk] * x[off + k];
        }
        y[off + j] = cb * acc;
      }
    }
  }

  /* family 3: the f-independent part of the Newmark right side,
   * M (4/dt^2 u0 + 4/dt ud0 + udd0) + C (ud0 + 2/dt u0) - K_eff u */
  std::vector<double> beam_z(const Beam& be) const {
    const int nd = be.nd;
    std::vector<double> z(nd);
    for (int i = 0; i < nd; ++i) {
      double acc = 0.0;
      for (int k = 0; k < nd; ++k) {
        const size_t ik = (size_t)i * nd + k;
        acc += be.M[ik] * (4.0 / (dt * dt) * be.u0[k] + 4.0 / dt * be.ud0[k] + be.udd0[k]) +
               be.C[ik] * (be.ud0[k] + 2.0 / dt * be.u0[k]) - be.Keff[ik] * be.u[k];
      }
      z[i] = acc;
    }
    return z;
  }
  /* family 3: the generalised force delta/dt Phi^T f of the beam's entries
   * fb (x forces, then y forces) */
  std::vector<double> beam_q(const Beam& be, const double* fb) const {
    std::vector<double> q(be.nd, 0.0);
    for (int p = 0; p < be.np; ++p) {
      const double fn = be.nx * fb[p] + be.ny * fb[be.np + p];
      for (int i = 0; i < be.nd; ++i) q[i] += be.H[(size_t)p * be.nd + i] * fn;
    }
    for (double& v : q) v *= delta / dt;
    return q;
  }
  static std::vector<double> matvec(int n, const std::vector<double>& A, const std::vector<double>& x) {
    std::vector<double> y(n, 0.0);
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc += A[(size_t)i * n + k] * x[k];
      y[i] = acc;
    }
    return y;
  }
  /* family 3: point positions X0 + n (H u) on both of each point's entries */
  void beam_positions(const Beam& be) {
    const int off = nbb_ind[be.body];
    for (int p = 0; p < be.np; ++p) {
      double w = 0.0;
      for (int i = 0; i < be.nd; ++i) w += be.H[(size_t)p * be.nd + i] * be.u[i];
      const double X = be.X0[p] + be.nx * w, Y = be.Y0[p] + be.ny * w;
      xb[off + p] = X; xb[off + be.np + p] = X;
      yb[off + p] = Y; yb[off + be.np + p] = Y;
    }
  }
  /* family 3: one Newmark update of u for the force fb; returns max |du| */
  double beam_update(Beam& be, const double* fb) {
    std::vector<double> z = beam_z(be), q = beam_q(be, fb);
    for (int i = 0; i < be.nd; ++i) z[i] += q[i];
    const std::vector<double> du = matvec(be.nd, be.Keffi, z);
    double e = 0.0;
    for (int i = 0; i < be.nd; ++i) {
      e = std::max(e, std::fabs(du[i]));
      be.u[i] += du[i];
      be.ud[i] = -be.ud0[i] + 2.0 / dt * (be.u[i] - be.u0[i]);
      be.udd[i] = -be.udd0[i] - 4.0 / dt * be.ud0[i] + 4.0 / (dt * dt) * (be.u[i] - be.u0[i]);
    }
    return e;
  }

  /* compute_rhs: src/operators_struct.F90:342-398 (gravity = dynamics = 0,
   * src/user.F90 compute_gravity/compute_dynamics with g_param = 0) */
  void compute_rhs(double* y) const {
    for (int t = 0; t < (int)tors.size(); ++t) {
      const Body& b = bdy[tors[t]];
      const int off = nbb_ind[tors[t]];
      const double grav = 0.0, dyn = 0.0;
      double rth = thetad0[t] + thetad[t] + 2.0 / dt * (theta0[t] - theta[t]);
      double den = 4.0 / std::pow(dt, 2.0) * Mt[t] + 2.0 / dt * Ct[t] + Jt[t];
      double jr = 2.0 / dt / den *
                  (Mt[t] * (4.0 / (dt * dt) * theta0[t] + 4.0 / dt * thetad0[t] + thetadd0[t]) +
                   Ct[t] * (thetad0[t] + 2.0 / dt * theta0[t]) -
                   (4.0 / (dt * dt) * Mt[t] + 2.0 / dt * Ct[t] + Jt[t]) * theta[t] + grav + dyn);
      double sc = delta * (rth - jr - thetad[t]);
      for (int j = 1; j <= 2 * b.npts; ++j) y[off + j - 1] = BRval(b, j, theta[t]) * sc;
    }
    /* family 3: delta Phi (ud0 + 2/dt (u0 - u) - 2/dt K_eff^-1 z) */
    for (const Beam& be : beams) {
      const int off = nbb_ind[be.body];
      const std::vector<double> jr = matvec(be.nd, be.Keffi, beam_z(be));
      std::vector<double> v(be.nd);
      for (int i = 0; i < be.nd; ++i) v[i] = be.ud0[i] + 2.0 / dt * (be.u0[i] - be.u[i]) - 2.0 / dt * jr[i];
      for (int p = 0; p < be.np; ++p) {
        double w = 0.0;
        for (int i = 0; i < be.nd; ++i) w += be.H[(size_t)p * be.nd + i] * v[i];
        y[off + p] = delta * be.nx * w;
        y[off + be.np + p] = delta * be.ny * w;
      }
    }
  }

  /* compute_dx + update_xb: src/operators_struct.F90:402-505, :600-620 */
  double compute_dx() {
    double e = 0.0;
    for (int t = 0; t < (int)tors.size(); ++t) {
      Body& b = bdy[tors[t]];
      const int off = nbb_ind[tors[t]];
      const double grav = 0.0, dyn = 0.0;
      double jr = 1.0 / (4.0 / (dt * dt) * Mt[t] + 2.0 / dt * Ct[t] + Jt[t]) *
                  (Mt[t] * (4 / (dt * dt) * theta0[t] + 4 / dt * thetad0[t] + thetadd0[t]) +
                   Ct[t] * (thetad0[t] + 2.0 / dt * theta0[t]) -
                   (4 / (dt * dt) * Mt[t] + 2.0 / dt * Ct[t] + Jt[t]) * theta[t] + grav + dyn);
      double brf = 0.0;
      for (int j = 1; j <= 2 * b.npts; ++j) brf += BRval(b, j, theta[t]) * f[off + j - 1];
      brf = b.deltas / (4 / (dt * dt) * Mt[t] + Jt[t]) * brf * (delta / b.deltas / dt);
      dtheta[t] = jr + brf;
      e = std::max(e, std::fabs(dtheta[t]));
      theta[t] = theta[t] + dtheta[t];
      thetad[t] = -thetad0[t] + 2 / dt * (theta[t] - theta0[t]);
      thetadd[t] = -thetadd0[t] - 4 / dt * thetad0[t] + 4 / (dt * dt) * (theta[t] - theta0[t]);
      const double c = std::cos(b.gamma0 - theta[t]), s_ = std::sin(b.gamma0 - theta[t]);
      for (int j = 1; j <= 2 * b.npts; ++j) {
        int jj = j <= b.npts ? j : j - b.npts;
        xb[off + j - 1] = b.xbf0 + (jj - 1) * b.deltas * c;
        yb[off + j - 1] = b.ybf0 + (jj - 1) * b.deltas * s_;
      }
    }
    for (Beam& be : beams) {
      e = std::max(e, beam_update(be, f.data() + nbb_ind[be.body]));
      beam_positions(be);
    }
    return e;
  }

  /* initial_acc: src/operators_struct.F90:233-294 */
  void initial_acc() {
    for (int t = 0; t < (int)tors.size(); ++t) {
      const Body& b = bdy[tors[t]];
      const int off = nbb_ind[tors[t]];
      double brf = 0.0;
      for (int j = 1; j <= 2 * b.npts; ++j) brf += BRval(b, j, theta[t]) * f[off + j - 1];
      brf = brf * delta / b.deltas / dt;
      thetadd[t] = 1 / Mt[t] * (-Ct[t] * thetad[t] - Ftint[t] + brf * b.deltas + 0.0 + 0.0);
    }
    /* family 3: udd = M^-1 (Phi^T f delta/dt - C ud - K u) */
    for (Beam& be : beams) {
      std::vector<double> r = beam_q(be, f.data() + nbb_ind[be.body]);
      const std::vector<double> cu = matvec(be.nd, be.C, be.ud), ku = matvec(be.nd, be.K, be.u);
      for (int i = 0; i < be.nd; ++i) r[i] -= cu[i] + ku[i];
      be.udd = matvec(be.nd, be.Minv, r);
    }
  }

  /* ---------------- Schur matvec and GMRES ---------------- */

  /* a_times2 + ab_times: src/operators_fluid.F90:14-120, src/timestepper.F90:774-782;
   * with_b = false gives a_times (src/operators_fluid.F90:1020-1105), the
   * operator the num_stat preprocess assembles */
  void schur(const double* x, double* y, bool with_b = true) const {
    std::vector<double> hx(Fx, 0.0), hy(Fy, 0.0), y2(nf, 0.0), y1(nf, 0.0);
    reg(x, hx.data(), hy.data());
    if (with_b) b_times2(x, y2.data());
    std::vector<std::vector<double>> vv(L, std::vector<double>(P, 0.0)), sv(L, std::vector<double>(P)),
        vx(L, std::vector<double>(Fx)), vy(L, std::vector<double>(Fy));
    rot(hx.data(), hy.data(), vv[0].data());
    ainv(vv[0].data());
    vort2flux(vv, sv, vx, vy);
    regT(vx[0].data(), vy[0].data(), y1.data());
    for (int k = 0; k < nf; ++k) y[k] = 1.0 * y1[k] + y2[k];
  }

  static double dot(const double* a, const double* b, int N) {
    double s = 0.0;
    for (int i = 0; i < N; ++i) s += a[i] * b[i];
    return s;
  }

  /* PETSc 3.11 KSPGMRES as configured in setup_ksp (src/variables.F90:565-574):
   * PCNONE, unpreconditioned norm, nonzero initial guess, classical
   * Gram-Schmidt without refinement, restart gmres_restart, convergence
   * ||r|| <= max(rtol*||b||, atol) on the Givens residual estimate. */
  int gmres(const double* b, double* x, int* its_out) const {
    return gmres_op(nf, [this](const double* v, double* w) { schur(v, w); }, b, x, its_out);
  }
  int gmres_op(int N, const std::function<void(const double*, double*)>& A, const double* b, double* x,
               int* its_out) const {
    const int kmax = prm.gmres_restart, maxit = prm.gmres_max_it;
    std::vector<std::vector<double>> V(kmax + 1, std::vector<double>(N));
    std::vector<double> H((size_t)(kmax + 1) * kmax, 0.0), cc(kmax), ss(kmax), grs(kmax + 1), w(N);
    auto HH = [&](int r, int c) -> double& { return H[(size_t)c * (kmax + 1) + r]; };
    const double bnorm = std::sqrt(dot(b, b, N));
    double ttol = 0.0, rn0 = 0.0;
    int its = 0;
    int reason = 0; /* 0 iterating, 1 converged, -1 diverged */
    bool first = true;
    while (!reason) {
      A(x, w.data());
      for (int i = 0; i < N; ++i) V[0][i] = b[i] - w[i];
      double res = std::sqrt(dot(V[0].data(), V[0].data(), N));
      grs[0] = res;
      if (first) {
        rn0 = bnorm > 0 ? bnorm : res;
        ttol = std::max(prm.rtol * rn0, prm.atol);
        first = false;
      }
      if (res == 0.0) { reason = 1; break; }
      if (res <= ttol) { reason = 1; break; }
      if (res >= 1e4 * rn0) { reason = -1; break; }
      for (int i = 0; i < N; ++i) V[0][i] /= res;
      int it = 0;
      bool hapend = false;
      while (!reason && it < kmax && its < maxit) {
        A(V[it].data(), V[it + 1].data());
        double lhh[1024];
        std::vector<double> lbig;
        double* lh = lhh;
        if (it + 1 > 1024) { lbig.resize(it + 1); lh = lbig.data(); }
        for (int j = 0; j <= it; ++j) lh[j] = dot(V[it + 1].data(), V[j].data(), N);
        for (int i = 0; i < N; ++i) {
          double acc = V[it + 1][i];
          for (int j = 0; j <= it; ++j) acc += -lh[j] * V[j][i];
          V[it + 1][i] = acc;
        }
        for (int j = 0; j <= it; ++j) HH(j, it) = lh[j];
        double tt = std::sqrt(dot(V[it + 1].data(), V[it + 1].data(), N));
        HH(it + 1, it) = tt;
        double hapbnd = std::fabs(tt / grs[it]);
        if (hapbnd > 1e-30) hapbnd = 1e-30;
        if (tt < hapbnd) hapend = true;
        else for (int i = 0; i < N; ++i) V[it + 1][i] = V[it + 1][i] * (1.0 / tt);
        /* KSPGMRESUpdateHessenberg */
        for (int j = 0; j < it; ++j) {
          double t0 = HH(j, it);
          HH(j, it) = cc[j] * t0 + ss[j] * HH(j + 1, it);
          HH(j + 1, it) = -ss[j] * t0 + cc[j] * HH(j + 1, it);
        }
        if (!hapend) {
          double t1 = std::sqrt(HH(it, it) * HH(it, it) + HH(it + 1, it) * HH(it + 1, it));
          cc[it] = HH(it, it) / t1;
          ss[it] = HH(it + 1, it) / t1;
          grs[it + 1] = -ss[it] * grs[it];
          grs[it] = cc[it] * grs[it];
          HH(it, it) = cc[it] * HH(it, it) + ss[it] * HH(it + 1, it);
          res = std::fabs(grs[it + 1]);
        } else {
          res = 0.0;
        }
        ++it;
        ++its;
        if (res <= ttol) reason = 1;
        else if (res >= 1e4 * rn0) reason = -1;
        if (hapend) { if (!reason) reason = 1; break; }
      }
      /* KSPGMRESBuildSoln */
      if (it > 0) {
        std::vector<double> nrs(it);
        int top = it - 1;
        nrs[top] = grs[top] / HH(top, top);
        for (int ii = 1; ii <= top; ++ii) {
          int k = top - ii;
          double t = grs[k];
          for (int j = k + 1; j <= top; ++j) t = t - HH(k, j) * nrs[j];
          nrs[k] = t / HH(k, k);
        }
        for (int i = 0; i < N; ++i) {
          double acc = 0.0;
          for (int j = 0; j <= top; ++j) acc += nrs[j] * V[j][i];
          x[i] = x[i] + acc;
        }
      }
      if (!reason && its >= maxit) { reason = -2; break; }
    }
    *its_out = its;
    return reason > 0 ? IB_OK : IB_E_NOCONV_GMRES;
  }

  /* ---------------- pressure: src/operators_pressure.F90 ---------------- */
  inline long cp(int i, int j) const { return (long)(i - 1) + (long)m * (j - 1); }

  /* 2-D RODFT00 on the m x n cell array (forward_ddti, src/myfft.F90:120) */
  void dst2d_p(double* a) const {
    std::vector<cplx> z(2 * std::max(dxp.M, dyp.M)), t(2 * std::max(dxp.M, dyp.M));
    for (int jp = 0; jp < (n + 1) / 2; ++jp) {
      int ja = 2 * jp, jb = ja + 1;
      double* ra = a + (long)ja * m;
      double* rb = jb < n ? a + (long)jb * m : nullptr;
      dxp.pair(ra, rb, 1, ra, rb, 1, z.data(), t.data());
    }
    for (int ip = 0; ip < (m + 1) / 2; ++ip) {
      int ia = 2 * ip, ib = ia + 1;
      double* ca = a + ia;
      double* cb = ib < m ? a + ib : nullptr;
      dyp.pair(ca, cb, m, ca, cb, m, z.data(), t.data());
    }
  }
  /* ddti: src/myfft.F90:192-203 */
  void ddti(const double* rhs, double* p) const {
    std::vector<double> a(rhs, rhs + (long)m * n);
    dst2d_p(a.data());
    for (long q = 0; q < (long)m * n; ++q) a[q] = laminv_ddti[q] * a[q];
    dst2d_p(a.data());
    std::memcpy(p, a.data(), sizeof(double) * m * n);
  }
  /* divergence over cells (:508-518) */
  void divergence(const double* hx, const double* hy, double* out) const {
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= m; ++i)
        out[cp(i, j)] = hx[fx(i + 1, j)] - hx[fx(i, j)] + hy[fy(i, j + 1)] - hy[fy(i, j)];
  }
  /* coarsify_pressure (:530-563) with the index ranges of
   * setup_coarsify_pressure (:567-605) for one process */
  void coarsify_pressure(const double* p1, double* p2) const {
    std::vector<double> temp((long)m * n, 0.0);
    for (int j = 2; j <= n - 2; j += 2)
      for (int i = 2; i <= m - 2; i += 2)
        temp[cp(i, j)] = p1[cp(i, j)] +
                         0.5 * (p1[cp(i + 1, j)] + p1[cp(i, j + 1)] + p1[cp(i - 1, j)] + p1[cp(i, j - 1)]) +
                         0.25 * (p1[cp(i + 1, j + 1)] + p1[cp(i + 1, j - 1)] + p1[cp(i - 1, j - 1)] +
                                 p1[cp(i - 1, j + 1)]);
    for (int j = 2; j <= n - 4; j += 2)
      for (int i = 2; i <= m - 4; i += 2) {
        const double yy = (temp[cp(i, j)] + 3.0 * temp[cp(i + 2, j)] + 3.0 * temp[cp(i, j + 2)] +
                           9.0 * temp[cp(i + 2, j + 2)]) /
                          16.0;
        const int id = m / 4 + i / 2 + 1, jd = n / 4 + j / 2 + 1;
        p2[cp(id, jd)] = yy;
      }
  }
  /* calculate_pressure (:110-437): pressure coefficient 2 (p + |u|^2/2 ...)
   * on every level, out = L blocks of m*n (write_pressure layout) */
  void pressure(double* out) const {
    const long C = (long)m * n;
    const int nbcp = 2 * m + 2 * n, lphi = 0, rphi = n, bphi = 2 * n, tphi = 2 * n + m;
    std::vector<std::vector<double>> qhx(L, std::vector<double>(Fx)), qhy(L, std::vector<double>(Fy));
    for (int k = 0; k < L; ++k) {
      const double fac = delta * std::pow(2.0, (double)k);
      for (int q = 0; q < Fx; ++q) qhx[k][q] = qx[k][q] + (-fac * (-1.0) + 0.0);
      for (int q = 0; q < Fy; ++q) qhy[k][q] = qy[k][q] + (-fac * 0.0 + 0.0);
    }
    /* nonlinear terms N(q) = q x omega on faces, coarse to fine */
    std::vector<std::vector<double>> nlx(L, std::vector<double>(Fx)), nly(L, std::vector<double>(Fy));
    std::vector<double> oqx(P), oqy(P), oqx_c(P), oqy_c(P), ubc(nbc, 0.0), vbc(nbc, 0.0);
    for (int k = L - 1; k >= 0; --k) {
      if (k == L - 1) {
        std::fill(ubc.begin(), ubc.end(), 0.0);
        std::fill(vbc.begin(), vbc.end(), 0.0);
      } else {
        get_bc(oqx_c.data(), ubc.data(), 1.0);
        get_bc(oqy_c.data(), vbc.data(), 1.0);
      }
      /* nonlinear_corner (:441-459) */
      for (int j = 2; j <= n; ++j)
        for (int i = 2; i <= m; ++i) {
          oqx[nd(i, j)] = 0.5 * (qhy[k][fy(i, j)] + qhy[k][fy(i - 1, j)]) * omega[k][nd(i, j)];
          oqy[nd(i, j)] = -0.5 * (qhx[k][fx(i, j)] + qhx[k][fx(i, j - 1)]) * omega[k][nd(i, j)];
        }
      /* nonlinear_corner2edge (:463-504) and the 1/del^2 scaling */
      const double del = delta * std::pow(2.0, (double)k);
      for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= m + 1; ++i)
          nlx[k][fx(i, j)] = (0.5 * (ext(oqx.data(), ubc.data(), i, j + 1) + ext(oqx.data(), ubc.data(), i, j))) /
                             (del * del);
      for (int j = 1; j <= n + 1; ++j)
        for (int i = 1; i <= m; ++i)
          nly[k][fy(i, j)] = (0.5 * (ext(oqy.data(), vbc.data(), i + 1, j) + ext(oqy.data(), vbc.data(), i, j))) /
                             (del * del);
      oqx_c = oqx;
      oqy_c = oqy;
    }
    /* rhs: -div(E^T f/dt) on level 1 (rhs_forcing adds the default zero body force) */
    std::vector<std::vector<double>> rhs(L, std::vector<double>(C, 0.0));
    std::vector<double> fdt(nf), hx(Fx, 0.0), hy(Fy, 0.0), dv(C);
    for (int k = 0; k < nf; ++k) fdt[k] = f[k] / dt;
    reg(fdt.data(), hx.data(), hy.data());
    divergence(hx.data(), hy.data(), rhs[0].data());
    for (long q = 0; q < C; ++q) rhs[0][q] = -rhs[0][q];
    for (int k = 1; k < L; ++k) coarsify_pressure(rhs[k - 1].data(), rhs[k].data());
    for (int k = 0; k < L; ++k) {
      divergence(nlx[k].data(), nly[k].data(), dv.data());
      for (long q = 0; q < C; ++q) rhs[k][q] = rhs[k][q] + dv[q];
    }
    /* solve D D^T p = rhs, coarse to fine */
    std::vector<std::vector<double>> pl(L, std::vector<double>(C, 0.0));
    std::vector<double> pbc(nbcp);
    for (int k = L - 1; k >= 0; --k) {
      std::fill(pbc.begin(), pbc.end(), 0.0);
      if (k == L - 1) {
        /* get_bc_phi1 (:707-746) */
        const double del = 0.125 / ((delta * std::pow(2.0, (double)k)) * (delta * std::pow(2.0, (double)k)));
        auto val = [&](int i, int j) {
          const double a = qhx[k][fx(i + 1, j)] + qhx[k][fx(i, j)], b = qhy[k][fy(i, j + 1)] + qhy[k][fy(i, j)];
          return del * (a * a + b * b);
        };
        for (int j = 1; j <= n; ++j) {
          pbc[lphi + j - 1] = val(1, j);
          pbc[rphi + j - 1] = val(m, j);
        }
        for (int i = 1; i <= m; ++i) {
          pbc[bphi + i - 1] = val(i, 1);
          pbc[tphi + i - 1] = val(i, n);
        }
      } else {
        /* get_bc_phi2 (:750-795) from the coarser level's pressure */
        const std::vector<double>& pc = pl[k + 1];
        auto P_ = [&](int i, int j) { return pc[cp(i, j)]; };
        for (int i = m / 4 + 1; i <= 3 * m / 4; ++i) {
          int j = n / 4, sl = bphi + 2 * (i - 1 - m / 4);
          pbc[sl] = 0.25 * (2 * P_(i, j) + P_(i, j + 1) + P_(i - 1, j));
          pbc[sl + 1] = 0.25 * (2 * P_(i, j) + P_(i, j + 1) + P_(i + 1, j));
          j = 3 * n / 4 + 1;
          sl = tphi + 2 * (i - 1 - m / 4);
          pbc[sl] = 0.25 * (2 * P_(i, j) + P_(i, j - 1) + P_(i - 1, j));
          pbc[sl + 1] = 0.25 * (2 * P_(i, j) + P_(i, j - 1) + P_(i + 1, j));
        }
        for (int j = n / 4 + 1; j <= 3 * n / 4; ++j) {
          int i = m / 4, sl = lphi + 2 * (j - 1 - n / 4);
          pbc[sl] = 0.25 * (2 * P_(i, j) + P_(i + 1, j) + P_(i, j - 1));
          pbc[sl + 1] = 0.25 * (2 * P_(i, j) + P_(i + 1, j) + P_(i, j + 1));
          i = 3 * m / 4 + 1;
          sl = rphi + 2 * (j - 1 - n / 4);
          pbc[sl] = 0.25 * (2 * P_(i, j) + P_(i - 1, j) + P_(i, j - 1));
          pbc[sl + 1] = 0.25 * (2 * P_(i, j) + P_(i - 1, j) + P_(i, j + 1));
        }
      }
      /* apply_bc_phi (:799-815) */
      std::vector<double>& r = rhs[k];
      for (int j = 1; j <= n; ++j) r[cp(1, j)] = r[cp(1, j)] - pbc[lphi + j - 1];
      for (int j = 1; j <= n; ++j) r[cp(m, j)] = r[cp(m, j)] - pbc[rphi + j - 1];
      for (int i = 1; i <= m; ++i) r[cp(i, 1)] = r[cp(i, 1)] - pbc[bphi + i - 1];
      for (int i = 1; i <= m; ++i) r[cp(i, n)] = r[cp(i, n)] - pbc[tphi + i - 1];
      ddti(r.data(), pl[k].data());
    }
    /* static_plus_dynamic (:819-831) and the scaling to the pressure coefficient */
    for (int k = 0; k < L; ++k) {
      const double del = 0.125 / ((delta * std::pow(2.0, (double)k)) * (delta * std::pow(2.0, (double)k)));
      for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= m; ++i) {
          const double a = qhx[k][fx(i + 1, j)] + qhx[k][fx(i, j)], b = qhy[k][fy(i, j + 1)] + qhy[k][fy(i, j)];
          const double v = pl[k][cp(i, j)] - del * (a * a + b * b);
          out[(long)k * C + cp(i, j)] = 2.0 * v;
        }
    }
  }

  /* ---------------- num_stat: precomputed matrix + Cholesky ----------------
   * preprocess (src/timestepper.F90:560-598) assembles column j as
   * a_times(e_j); KSPPREONLY + PCCHOLESKY (src/variables.F90:555-561) then
   * factors it.  The Cholesky solver (Elemental, third party, not in the
   * reference tree) reads one triangle; this restatement uses the upper
   * one: A(i, j) = S(min(i, j), max(i, j)), A = U^T U. */
  std::vector<double> cholU;  // n x n column major, upper triangle used
  int preprocess() {
    const int N = nf;
    std::vector<double> S((size_t)N * N), e(N, 0.0);
    for (int j = 0; j < N; ++j) {
      e[j] = 1.0;
      schur(e.data(), S.data() + (size_t)j * N, false);
      e[j] = 0.0;
    }
    cholU.assign((size_t)N * N, 0.0);
    for (int j = 0; j < N; ++j) {
      for (int i = 0; i <= j; ++i) {
        double acc = S[(size_t)j * N + i];
        for (int k = 0; k < i; ++k) acc -= cholU[(size_t)i * N + k] * cholU[(size_t)j * N + k];
        if (i < j) {
          cholU[(size_t)j * N + i] = acc / cholU[(size_t)i * N + i];
        } else {
          if (!(acc > 0.0)) return fail(IB_E_INVAL, "num_stat: the Schur matrix is not positive definite");
          cholU[(size_t)j * N + j] = std::sqrt(acc);
        }
      }
    }
    return IB_OK;
  }
  // x = A^-1 b: U^T y = b, then U x = y
  void chol_solve(const double* b, double* x) const {
    const int N = nf;
    std::vector<double> y(N);
    for (int i = 0; i < N; ++i) {
      double acc = b[i];
      for (int k = 0; k < i; ++k) acc -= cholU[(size_t)i * N + k] * y[k];
      y[i] = acc / cholU[(size_t)i * N + i];
    }
    for (int i = N - 1; i >= 0; --i) {
      double acc = y[i];
      for (int k = i + 1; k < N; ++k) acc -= cholU[(size_t)k * N + i] * x[k];
      x[i] = acc / cholU[(size_t)i * N + i];
    }
  }

  /* redistribute + redistribute2: src/operators_struct.F90:735-834 */
  void redistribute() {
    std::vector<double> ones(nf, 1.0), wx(Fx, 0.0), wy(Fy, 0.0), frx(Fx, 0.0), fry(Fy, 0.0), fix(Fx, 0.0),
        fiy(Fy, 0.0);
    reg(ones.data(), wx.data(), wy.data());
    reg(f.data(), frx.data(), fry.data());
    for (int q = 0; q < Fx; ++q)
      if (wx[q] > 1e-10) fix[q] = 1.0 / wx[q] * frx[q];
    for (int q = 0; q < Fy; ++q)
      if (wy[q] > 1e-10) fiy[q] = 1.0 / wy[q] * fry[q];
    std::fill(f_rdst.begin(), f_rdst.end(), 0.0);
    regT(fix.data(), fiy.data(), f_rdst.data());
  }

  /* ---------------- sub-domain approach ----------------------------------
   * Sub-domain points (src/grid.F90:320-348): the points of the one
   * stationary rigid body unless full_rectangular, then the rectangle, x
   * fastest.  Entry k < nb_sd is the x force of point k, the rest the y
   * forces (collect_subdomain, :437-455). */
  bool sd_on = false;
  ib_subdomain sdc{};
  int sd_nx = 0, sd_ny = 0, nr_sd = 0, nb_sd = 0, nf_sd = 0, sd_np = 4;
  std::vector<double> sd_x, sd_y;
  std::vector<double> wsd;  /* weight_sd: 49 per sub-domain entry */
  std::vector<int> ksd;     /* face of each tap: x-face index, or Fx + y-face index */
  std::vector<std::vector<int>> bcol_row;    /* kept rows of column j of B, in list order */
  std::vector<std::vector<double>> bcol_val;
  int64_t sd_nnz_eval = 0, sd_nnz_kept = 0;
  std::vector<int> pidx;    /* interp_indices: sd_np per force entry (0-based sub-domain entries) */
  std::vector<double> pw;   /* interp_weights */
  std::vector<double> lumat;

  static int fnint(double v) { return (int)std::lround(v); } /* Fortran NINT */

  /* deltafnc2 (src/grid.F90:1007-1029): the 2-point hat */
  static double deltafnc2(double x1, double x2, double dr) {
    const double r1 = std::fabs(x1 - x2) / dr;
    const double d = r1 <= 1.0 ? 1.0 - r1 : 0.0;
    return d / dr;
  }

  int setup_subdomain() {
    const ib_subdomain& c = sdc;
    if (!(c.delta_sd > 0.0) || !(c.sdxe >= c.sdxi) || !(c.sdye >= c.sdyi))
      return fail(IB_E_INVAL, "sub-domain: need delta_sd > 0 and a non-empty rectangle");
    const int s2n = (int)std::lround(std::sqrt((double)sd_np));
    if (s2n * s2n != sd_np) return fail(IB_E_INVAL, "sub-domain: interp_points must be a square");
    nr_sd = 0;
    sd_x.clear();
    sd_y.clear();
    if (!c.full_rectangular) {
      int nrig = 0;
      for (const Body& b : bdy)
        if (b.fam == 1) {
          ++nrig;
          sd_x.insert(sd_x.end(), b.x.begin(), b.x.end());
          sd_y.insert(sd_y.end(), b.y.begin(), b.y.end());
        }
      if (nrig > 1)
        return fail(IB_E_UNSUPPORTED, "sub-domain with body points: the reference includes one rigid body only");
      nr_sd = (int)sd_x.size();
    }
    sd_nx = fnint((c.sdxe - c.sdxi) / c.delta_sd + 1.0);
    sd_ny = fnint((c.sdye - c.sdyi) / c.delta_sd + 1.0);
    for (int j = 1; j <= sd_ny; ++j)
      for (int i = 1; i <= sd_nx; ++i) {
        sd_x.push_back(c.sdxi + (i - 1) * c.delta_sd);
        sd_y.push_back(c.sdyi + (j - 1) * c.delta_sd);
      }
    nb_sd = (int)sd_x.size();
    nf_sd = 2 * nb_sd;
    /* setup_pre_reg_subdomain (src/grid.F90:628-692): E on the sub-domain */
    const int s2 = support2, nt = (2 * s2 + 1) * (2 * s2 + 1);
    const double d2 = 0.5 * delta;
    const bool f32 = prm.deltafnc_f32_literals != 0;
    wsd.assign((size_t)nt * nf_sd, 0.0);
    ksd.assign((size_t)nt * nf_sd, 0);
    std::vector<int> bix(nb_sd), biy(nb_sd);
    for (int k = 0; k < nf_sd; ++k) {
      const int pt = k % nb_sd, comp = k < nb_sd ? 1 : 2;
      const int ix = (int)((sd_x[pt] + offx) / delta) + 1;
      const int iy = (int)((sd_y[pt] + offy) / delta) + 1;
      bix[pt] = ix;
      biy[pt] = iy;
      int next = 0;
      for (int l = -s2; l <= s2; ++l)
        for (int p = -s2; p <= s2; ++p) {
          double x, y;
          const int ii = ix + p, jj = iy + l;
          if (comp == 1) {
            x = delta * (ix - 1 + p) - offx;
            y = delta * (iy - 1 + l) - offy + d2;
            if (ii < 1 || ii > m + 1 || jj < 1 || jj > n)
              return fail(IB_E_INVAL, "sub-domain point regularisation stencil leaves the grid");
            ksd[(size_t)k * nt + next] = (m + 1) * (iy + l - 1) + (ix + p - 1);
          } else {
            x = delta * (ix - 1 + p) - offx + d2;
            y = delta * (iy - 1 + l) - offy;
            if (ii < 1 || ii > m || jj < 1 || jj > n + 1)
              return fail(IB_E_INVAL, "sub-domain point regularisation stencil leaves the grid");
            ksd[(size_t)k * nt + next] = Fx + m * (iy + l - 1) + (ix + p - 1);
          }
          wsd[(size_t)k * nt + next] =
              delta * delta * deltafnc(x, sd_x[pt], delta, f32) * deltafnc(y, sd_y[pt], delta, f32);
          ++next;
        }
    }
    /* get_sparsity_EEt (src/variables.F90:425-514) with the window matrix of
     * setup_subdomain_sparsity (src/grid.F90:696-751): rows i and j of B are
     * evaluated together when their windows of half-width support_nz share
     * a column of that matrix (x entries: (m+1)(iy+l-1)+(ix+p-1); y entries:
     * (m+1)n-1 + m(iy+l-1)+(ix+p-1)) */
    const int sn = c.support_nz, nvel = (m + 1) * n + m * (n + 1);
    std::vector<std::pair<int, int>> colrow;
    colrow.reserve((size_t)nf_sd * (2 * sn + 1) * (2 * sn + 1));
    for (int k = 0; k < nf_sd; ++k) {
      const int pt = k % nb_sd;
      for (int l = -sn; l <= sn; ++l)
        for (int p = -sn; p <= sn; ++p) {
          const int col = k < nb_sd ? (m + 1) * (biy[pt] + l - 1) + (bix[pt] + p - 1)
                                    : (m + 1) * n - 1 + m * (biy[pt] + l - 1) + (bix[pt] + p - 1);
          if (col < 0 || col >= nvel) return fail(IB_E_INVAL, "sub-domain: sparsity window leaves the grid");
          colrow.push_back({col, k});
        }
    }
    std::sort(colrow.begin(), colrow.end());
    std::vector<std::vector<int>> ja(nf_sd);
    for (size_t a = 0; a < colrow.size();) {
      size_t b = a;
      while (b < colrow.size() && colrow[b].first == colrow[a].first) ++b;
      for (size_t u = a; u < b; ++u)
        for (size_t v = a; v < b; ++v) ja[colrow[u].second].push_back(colrow[v].second);
      a = b;
    }
    std::vector<std::vector<int>> rows(nf_sd);
    for (int k = 0; k < nf_sd; ++k) {
      std::vector<int>& J = ja[k];
      std::sort(J.begin(), J.end());
      J.erase(std::unique(J.begin(), J.end()), J.end());
      for (int j : J)
        if ((k < nb_sd) != (j < nb_sd))
          return fail(IB_E_INVAL, "sub-domain: an x window meets a y window (sub-domain at the grid corner)");
      std::vector<int>& R = rows[k];
      if (k < nb_sd) {
        R = J;
        for (int j : J) R.push_back(j + nb_sd);
      } else {
        for (int j : J) R.push_back(j - nb_sd);
        R.insert(R.end(), J.begin(), J.end());
      }
    }
    /* preprocess_sd (src/timestepper.F90:602-770): column j of B is
     * a_times_sd(e_j) (src/operators_fluid.F90:124-199) at its rows; an
     * entry is kept when |B(r, j)| > droptol * B(j, j) */
    const double dtol = c.droptol;
    bcol_row.assign(nf_sd, {});
    bcol_val.assign(nf_sd, {});
    sd_nnz_eval = sd_nnz_kept = 0;
    std::vector<double> hx(Fx), hy(Fy), vals;
    std::vector<std::vector<double>> vv(L, std::vector<double>(P)), sv(L, std::vector<double>(P)),
        vx(L, std::vector<double>(Fx)), vy(L, std::vector<double>(Fy));
    for (int j = 0; j < nf_sd; ++j) {
      std::fill(hx.begin(), hx.end(), 0.0);
      std::fill(hy.begin(), hy.end(), 0.0);
      for (int t = 0; t < nt; ++t) {
        const int key = ksd[(size_t)j * nt + t];
        const double v = wsd[(size_t)j * nt + t] * 1.0;
        if (key < Fx) hx[key] += v;
        else hy[key - Fx] += v;
      }
      for (auto& a : vv) std::fill(a.begin(), a.end(), 0.0);
      rot(hx.data(), hy.data(), vv[0].data());
      ainv(vv[0].data());
      vort2flux(vv, sv, vx, vy);
      const std::vector<int>& R = rows[j];
      vals.assign(R.size(), 0.0);
      double diag = 0.0;
      for (size_t a = 0; a < R.size(); ++a) {
        const int r = R[a];
        double acc = 0.0;
        for (int t = 0; t < nt; ++t) {
          const int key = ksd[(size_t)r * nt + t];
          acc = acc + wsd[(size_t)r * nt + t] * (key < Fx ? vx[0][key] : vy[0][key - Fx]);
        }
        vals[a] = acc;
        if (r == j) diag = acc;
      }
      sd_nnz_eval += (int64_t)R.size();
      for (size_t a = 0; a < R.size(); ++a)
        if (std::fabs(vals[a]) > dtol * diag) {
          bcol_row[j].push_back(R[a]);
          bcol_val[j].push_back(vals[a]);
        }
      sd_nnz_kept += (int64_t)bcol_row[j].size();
    }
    return IB_OK;
  }

  /* setup_reg_subdomain_indices (src/grid.F90:777-951): P, the
   * interpolation of each force entry from sub-domain points (bilinear hat
   * weights), or the entry itself for the points of the rigid body when
   * they are sub-domain points */
  int subdomain_interp() {
    const ib_subdomain& c = sdc;
    const int s = (int)std::lround(std::sqrt((double)sd_np));
    int i1, i2;
    if (sd_np % 2 == 0) i1 = s / 2 + 1 - s;
    else i1 = (s - 1) / 2 + 1 - s;
    i2 = i1 + s - 1;
    (void)i2;
    pidx.assign((size_t)sd_np * nf, 0);
    pw.assign((size_t)sd_np * nf, 0.0);
    std::vector<int> nn(sd_np), ll(s), pp(s);
    std::vector<double> dist(sd_np), w(sd_np);
    for (int j = 0; j < nf; ++j) {
      const int comp = ccomp[j];
      const double indx = (xb[j] - c.sdxi) / c.delta_sd + 1.0, indy = (yb[j] - c.sdyi) / c.delta_sd + 1.0;
      for (int a = 0; a < s; ++a) {
        ll[a] = (sd_np % 2 == 0 ? (int)indx : fnint(indx)) + i1 + a;
        pp[a] = (sd_np % 2 == 0 ? (int)indy : fnint(indy)) + i1 + a;
      }
      if (!c.full_rectangular && bdy[cbody[j]].fam == 1) {
        /* the rigid body is the first body; its points are sub-domain
         * points 1..nr in order (1-based j: nearest_neighbor = j, or j - nr
         * for y entries; the rest are placeholders of weight 0) */
        const int j1 = j + 1, nr = nb_ind[cbody[j]];
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
            const double dx = xb[j] - sd_x[nn[i] - 1], dy = yb[j] - sd_y[nn[i] - 1];
            dist[i] = std::sqrt(dx * dx + dy * dy);
            ++i;
          }
      }
      bool anyexact = false;
      for (int i = 0; i < sd_np; ++i) {
        const int q = nn[i] - 1;
        if (q < 0 || q >= nb_sd) return fail(IB_E_INVAL, "sub-domain: interpolation point out of range");
        w[i] = c.delta_sd * c.delta_sd * deltafnc2(sd_x[q], xb[j], c.delta_sd) * deltafnc2(sd_y[q], yb[j], c.delta_sd);
        if (dist[i] == 0.0) {
          w[i] = 1.0;
          anyexact = true;
        }
      }
      for (int i = 0; i < sd_np; ++i) {
        if (anyexact && dist[i] != 0.0) w[i] = 0.0;
        pidx[(size_t)j * sd_np + i] = nn[i] - 1 + (comp == 1 ? 0 : nb_sd);
        pw[(size_t)j * sd_np + i] = w[i];
      }
    }
    return IB_OK;
  }

  /* B(r, c) as compute_operator reads it: the kept entry, else 0 */
  double bget(int r, int col) const {
    const std::vector<int>& R = bcol_row[col];
    for (size_t a = 0; a < R.size(); ++a)
      if (R[a] == r) return bcol_val[col][a];
    return 0.0;
  }

  /* compute_operator + body_mat (src/operators_struct.F90:838-1004):
   * lumat(k, i) = body_mat(k, i) + sum_p w_p(k) I(i, idx_p(k)),
   * I(i, c) = sum_q w_q(i) B(idx_q(i), c) */
  void compute_operator() {
    lumat.assign((size_t)nf * nf, 0.0);
    for (size_t t = 0; t < tors.size(); ++t) {
      const Body& b = bdy[tors[t]];
      const int off = nbb_ind[tors[t]], nbi = b.npts;
      const double si = std::sin(b.gamma0 - theta[t]), ci = std::cos(b.gamma0 - theta[t]);
      const double den = 4.0 / (dt * dt) * Mt[t] + 2.0 / dt * Ct[t] + Jt[t];
      for (int jj = 1; jj <= 2 * nbi; ++jj)
        for (int kk = 1; kk <= 2 * nbi; ++kk) {
          const double im = jj - 1, in = kk - 1, ds2 = b.deltas * b.deltas;
          double v;
          /* the reference's expressions, evaluated left to right */
          if (jj <= nbi && kk <= nbi) v = (im * in) * (si * si) * ds2;
          else if (jj <= nbi) v = -((im * (in - nbi)) * si * ci * ds2);
          else if (kk <= nbi) v = -(((im - nbi) * in) * si * ci * ds2);
          else v = ((im - nbi) * (in - nbi)) * (ci * ci) * ds2;
          v = v / den * (delta * delta / b.deltas / dt) * (2 * b.deltas / dt);
          lumat[(size_t)(off + kk - 1) * nf + (off + jj - 1)] = v;
        }
    }
    std::vector<double> I(nf);
    for (int i = 0; i < nf; ++i)
      for (int k = 0; k < nf; ++k) {
        double acc = 0.0;
        for (int p = 0; p < sd_np; ++p) {
          const int cidx = pidx[(size_t)k * sd_np + p];
          double ic = 0.0;
          for (int q = 0; q < sd_np; ++q) ic += bget(pidx[(size_t)i * sd_np + q], cidx) * pw[(size_t)i * sd_np + q];
          acc += pw[(size_t)k * sd_np + p] * ic;
        }
        lumat[(size_t)i * nf + k] = lumat[(size_t)i * nf + k] + acc;
      }
  }

  /* rhsf += P (E_sd (q + q0)) (regT_sd_sparse, src/fsinterface.F90:130-161,
   * and the interpolation of src/timestepper.F90:389-393) */
  void subdomain_rhs(const double* qhx, const double* qhy, double* rhsf) const {
    const int nt = (2 * support2 + 1) * (2 * support2 + 1);
    std::vector<double> r(nf_sd, 0.0);
    for (int k = 0; k < nf_sd; ++k)
      for (int t = 0; t < nt; ++t) {
        const int key = ksd[(size_t)k * nt + t];
        r[k] = r[k] + wsd[(size_t)k * nt + t] * (key < Fx ? qhx[key] : qhy[key - Fx]);
      }
    for (int j = 0; j < nf; ++j) {
      double acc = 0.0;
      for (int p = 0; p < sd_np; ++p) acc += pw[(size_t)j * sd_np + p] * r[pidx[(size_t)j * sd_np + p]];
      rhsf[j] = acc + rhsf[j];
    }
  }

  /* corrector source: hx, hy = E_sd^T (P^T f) over the used sub-domain
   * entries in ascending order (src/timestepper.F90:500-510,
   * reg_sd_sparse src/fsinterface.F90:84-126) */
  void subdomain_reg(double* hx, double* hy) const {
    const int nt = (2 * support2 + 1) * (2 * support2 + 1);
    std::vector<int> used(pidx.begin(), pidx.end());
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    std::vector<double> h0(nf_sd, 0.0);
    for (int cidx : used) {
      double acc = 0.0;
      for (int j = 0; j < nf; ++j) {
        double wj = 0.0;
        for (int p = 0; p < sd_np; ++p)
          if (pidx[(size_t)j * sd_np + p] == cidx) wj = pw[(size_t)j * sd_np + p];
        acc += wj * f[j];
      }
      h0[cidx] = acc;
    }
    for (int k : used)
      for (int t = 0; t < nt; ++t) {
        const int key = ksd[(size_t)k * nt + t];
        const double v = wsd[(size_t)k * nt + t] * h0[k];
        if (key < Fx) hx[key] += v;
        else hy[key - Fx] += v;
      }
  }

  /* ---------------- one time step: src/timestepper.F90:18-556 ----------- */
  int advance() {
    itime += 1;
    /* motion_potential/motion_rotation (src/variables.F90:299-360) with
     * motion_grid = (-1, 0, 0, 0, 0) (src/user.F90:13-31) */
    for (int k = 0; k < L; ++k) {
      double fac = delta * std::pow(2.0, (double)k);
      std::fill(q0x[k].begin(), q0x[k].end(), -fac * (-1.0) + 0.0);
      std::fill(q0y[k].begin(), q0y[k].end(), -fac * 0.0 + 0.0);
    }
    /* step 1: predictor on every level, coarse to fine */
    std::vector<std::vector<double>> lastbc(L, std::vector<double>(nbc, 0.0));
    for (int k = 0; k < L - 1; ++k) get_bc(omega[k + 1].data(), lastbc[k].data(), 0.25);
    std::vector<double> omegabc(nbc), rhs(P), rhsbc(P), nlx(Fx), nly(Fy), qhx(Fx), qhy(Fy);
    for (int k = L - 1; k >= 0; --k) {
      for (int q = 0; q < Fx; ++q) qhx[q] = qx[k][q] + q0x[k][q];
      for (int q = 0; q < Fy; ++q) qhy[q] = qy[k][q] + q0y[k][q];
      if (k < L - 1) get_bc(omega[k + 1].data(), omegabc.data(), 0.25);
      else std::fill(omegabc.begin(), omegabc.end(), 0.0);
      nonlinear(omega[k].data(), qhx.data(), qhy.data(), lastbc[k].data(), nlx.data(), nly.data());
      std::fill(rhsbc.begin(), rhsbc.end(), 0.0);
      apply_bc(rhsbc.data(), lastbc[k].data(), vfac[k]);
      apply_bc(rhsbc.data(), omegabc.data(), vfac[k]);
      rot(nlx.data(), nly.data(), rhs.data());
      if (itime == 1) rhs_old[k] = rhs;
      predictor(omega[k].data(), rhs.data(), rhs_old[k].data(), rhsbc.data(), k);
      rhs_old[k] = rhs;
    }
    vort2flux(omega, s, qx, qy);
    for (int q = 0; q < Fx; ++q) qhx[q] = qx[0][q] + q0x[0][q];
    for (int q = 0; q < Fy; ++q) qhy[q] = qy[0][q] + q0y[0][q];

    if (prm.num_stat) {
      /* step 2 for a numerically stationary body (src/timestepper.F90:457-484):
       * rhsf = regT(q + q0), f from the Cholesky factors, redistribute; the
       * bodies do not move and no iteration counts are kept */
      std::vector<double> rhsf(nf, 0.0);
      regT(qhx.data(), qhy.data(), rhsf.data());
      chol_solve(rhsf.data(), f.data());
      redistribute();
      ksp_iter = 0;
      fsi_iter = 0;
      err_fsi = 0.0;
    } else {
    /* step 2: FSI fixed point with GMRES (num_stat = F, sub_domain = F) */
    int status = IB_OK;
    int fsi_count = 0;
    double err_glob = 10.0;
    ksp_iter = 0;
    for (size_t t = 0; t < tors.size(); ++t) {
      const Body& b = bdy[tors[t]];
      Mt[t] = b.itheta;          /* get_M */
      Jt[t] = b.ktheta;          /* get_JK */
      Ct[t] = b.ctheta;          /* get_C */
      Ftint[t] = Jt[t] * theta[t]; /* var_update */
    }
    if (itime == 1) initial_acc();
    theta0 = theta;
    thetad0 = thetad;
    thetadd0 = thetadd;
    for (Beam& be : beams) {
      be.u0 = be.u;
      be.ud0 = be.ud;
      be.udd0 = be.udd;
    }
    std::vector<double> rhsf(nf), yb_(nf);
    while (err_glob >= prm.fsi_tol && fsi_count <= prm.fsi_max_count) {
      ++fsi_count;
      std::fill(rhsf.begin(), rhsf.end(), 0.0);
      std::fill(yb_.begin(), yb_.end(), 0.0);
      compute_rhs(yb_.data());
      for (int k = 0; k < nf; ++k) rhsf[k] += yb_[k];
      int its = 0;
      int rc;
      if (!sd_on) {
        regT(qhx.data(), qhy.data(), rhsf.data());
        rc = gmres(rhsf.data(), f.data(), &its);
        ksp_iter += its;
      } else {
        /* src/timestepper.F90:368-405: lumat for the current P, GMRES on
         * it; the reference does not add these iterations to ksp_iter */
        subdomain_rhs(qhx.data(), qhy.data(), rhsf.data());
        compute_operator();
        rc = gmres_op(
            nf,
            [this](const double* v, double* w) {
              for (int k = 0; k < nf; ++k) {
                double acc = 0.0;
                for (int i = 0; i < nf; ++i) acc += lumat[(size_t)i * nf + k] * v[i];
                w[k] = acc;
              }
            },
            rhsf.data(), f.data(), &its);
      }
      if (rc != IB_OK) status = rc;
      redistribute();
      double e = compute_dx();
      err_glob = e;
      if (err_glob >= prm.fsi_tol)
        for (size_t t = 0; t < tors.size(); ++t) Ftint[t] = Jt[t] * theta[t];
      /* E follows the bodies; on the sub-domain P does instead
       * (reg_prep_sd_sparse, src/variables.F90:377-396) */
      int r2 = sd_on ? subdomain_interp() : setup_reg();
      if (r2 != IB_OK) return r2;
    }
    if (err_glob >= prm.fsi_tol && status == IB_OK) status = IB_E_NOCONV_FSI;
    fsi_iter = fsi_count;
    err_fsi = err_glob;
    step_status = status;
    }

    /* step 3: corrector on level 1 */
    std::vector<double> hx(Fx, 0.0), hy(Fy, 0.0), vort(P);
    if (sd_on) subdomain_reg(hx.data(), hy.data());
    else reg(f.data(), hx.data(), hy.data());
    rot(hx.data(), hy.data(), vort.data());
    ainv(vort.data());
    for (int q = 0; q < P; ++q) omega[0][q] = omega[0][q] - vort[q];
    vort2flux(omega, s, qx, qy);
    return prm.num_stat ? IB_OK : step_status;
  }
  int step_status = IB_OK;

  /* ---------------- setup: src/grid.F90:117-434, src/myfft.F90:237-298 --- */
  int init(const ib_params* p, int nbodies, const ib_body* bodies, const ib_subdomain* sd = nullptr) {
    prm = *p;
    if (prm.gmres_restart <= 0) prm.gmres_restart = 30;
    if (prm.gmres_max_it <= 0) prm.gmres_max_it = 10000;
    if (prm.fsi_max_count <= 0) prm.fsi_max_count = 1000;
    m = prm.m; n = prm.n; L = prm.mgridlev;
    if (m < 8 || n < 8 || m % 4 || n % 4) return fail(IB_E_INVAL, "grid must be divisible by 4 (m, n >= 8)");
    if (L < 1) return fail(IB_E_INVAL, "mgridlev must be >= 1");
    if (prm.sub_domain && !sd) return fail(IB_E_UNSUPPORTED, "sub_domain = T needs the sub-domain parameters (create_subdomain)");
    if (sd && !prm.sub_domain) return fail(IB_E_INVAL, "sub-domain parameters given with sub_domain = F");
    if (sd && prm.num_stat) return fail(IB_E_INVAL, "num_stat and sub_domain exclude each other");
    if (prm.istart < 0) return fail(IB_E_INVAL, "istart must be >= 0");
    pi = 4.0 * std::atan(1.0);
    dt = prm.dt; Re = prm.re; len = prm.len; offx = prm.offsetx; offy = prm.offsety;
    delta = len / (double)m;
    P = (m - 1) * (n - 1); Fx = (m + 1) * n; Fy = m * (n + 1);
    nbc = 2 * (m + 1) + 2 * (n + 1);
    left = 0; right = n + 1; bottom = 2 * (n + 1); top = 2 * (n + 1) + m + 1;
    dx.init(m - 1);
    dy.init(n - 1);
    /* setup_fft tables (src/myfft.F90:262-289) */
    normalize = 4.0 * (double)(float)(m * n);
    const double del2 = delta * delta;
    vfac.resize(L); con1.resize(L); con2.resize(L);
    for (int k = 0; k < L; ++k) {
      double del22 = del2 * std::pow(4.0, (double)k);
      vfac[k] = 0.5 * dt / Re / del22 / normalize;
      con1[k] = 1.5 * dt / del22 / normalize;
      con2[k] = -0.5 * dt / del22 / normalize;
    }
    laminv.resize(P); lam1.resize((size_t)P * L); lam1i.resize((size_t)P * L);
    for (int j = 2; j <= n; ++j)
      for (int i = 2; i <= m; ++i) {
        double lam = -2.0 * (std::cos(pi * (double)(i - 1) / (double)m) + std::cos(pi * (double)(j - 1) / (double)n) - 2.0);
        long q = nd(i, j);
        laminv[q] = 1.0 / lam / normalize;
        for (int k = 0; k < L; ++k) {
          double del22 = del2 * std::pow(4.0, (double)k);
          lam1[(size_t)k * P + q] = (1.0 - 0.5 * dt * lam / del22 / Re) / normalize;
          lam1i[(size_t)k * P + q] = 1.0 / (1.0 + 0.5 * dt * lam / del22 / Re);
        }
      }
    /* pressure tables (src/myfft.F90:123,150-155) */
    dxp.init(m);
    dyp.init(n);
    {
      const double norm_ddti = 4.0 * (double)(float)((m + 1) * (n + 1));
      laminv_ddti.resize((size_t)m * n);
      for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= m; ++i) {
          const double lam = 2.0 * (std::cos(pi * (double)i / (double)(m + 1)) +
                                    std::cos(pi * (double)j / (double)(n + 1)) - 2.0);
          laminv_ddti[cp(i, j)] = 1.0 / lam / norm_ddti;
        }
    }
    /* bodies: setup_geometry (src/grid.F90:239-398) */
    if (nbodies < 0) return fail(IB_E_INVAL, "negative body count");
    bool seen_tors = false;
    nbb_ind.assign(1, 0);
    for (int b = 0; b < nbodies; ++b) {
      Body B;
      B.npts = bodies[b].npts;
      B.fam = bodies[b].fam;
      if (B.npts < 2) return fail(IB_E_INVAL, "a body needs at least 2 points");
      if (B.fam < 1 || B.fam > 3) return fail(IB_E_UNSUPPORTED, "body families are 1 (rigid), 2 (torsional), 3 (beam)");
      if (B.fam == 1 && seen_tors) return fail(IB_E_INVAL, "rigid bodies must precede torsional bodies (src/grid.F90:1226)");
      if (B.fam >= 2) seen_tors = true;
      B.gamma0 = bodies[b].gamma0; B.alpha0 = bodies[b].alpha0;
      B.itheta = bodies[b].itheta; B.ktheta = bodies[b].ktheta; B.ctheta = bodies[b].ctheta;
      B.x.assign(bodies[b].x, bodies[b].x + B.npts);
      B.y.assign(bodies[b].y, bodies[b].y + B.npts);
      B.deltas = std::sqrt((B.x[1] - B.x[0]) * (B.x[1] - B.x[0]) + (B.y[1] - B.y[0]) * (B.y[1] - B.y[0]));
      B.xbf0 = B.x[0]; B.ybf0 = B.y[0];
      if (B.fam == 2 && B.itheta <= 0) return fail(IB_E_INVAL, "torsional body needs itheta > 0");
      if (B.fam == 2) tors.push_back(b);
      if (B.fam == 3) {
        if (B.npts < 3 || B.npts % 2 == 0) return fail(IB_E_INVAL, "a beam needs an odd number of points >= 3");
        if (!(B.itheta > 0) || !(B.ktheta > 0)) return fail(IB_E_INVAL, "a beam needs mass (itheta) and stiffness (ktheta) > 0");
        if (prm.num_stat || prm.sub_domain)
          return fail(IB_E_UNSUPPORTED, "beams (family 3) need the GMRES path without num_stat or sub-domains");
        Beam be;
        be.body = b;
        if (!beam_build(be, B, dt)) return fail(IB_E_INVAL, "beam matrices are singular");
        beams.push_back(be);
      }
      bdy.push_back(B);
      nb_ind.push_back(B.npts);
      nbb_ind.push_back(nbb_ind.back() + 2 * B.npts);
      nb += B.npts;
    }
    nf = 2 * nb;
    /* collect_bodies: src/grid.F90:403-433 */
    xb.resize(nf); yb.resize(nf); cbody.resize(nf); ccomp.resize(nf);
    for (int b = 0; b < nbodies; ++b)
      for (int j = 0; j < 2 * bdy[b].npts; ++j) {
        int k = nbb_ind[b] + j;
        int pnt = j < bdy[b].npts ? j : j - bdy[b].npts;
        cbody[k] = b;
        ccomp[k] = j < bdy[b].npts ? 1 : 2;
        xb[k] = bdy[b].x[pnt];
        yb[k] = bdy[b].y[pnt];
      }
    f.assign(nf, 0.0); f_rdst.assign(nf, 0.0);
    size_t nt = tors.size();
    theta.assign(nt, 0.0); thetad.assign(nt, 0.0); thetadd.assign(nt, 0.0);
    theta0 = theta; thetad0 = theta; thetadd0 = theta; dtheta = theta;
    Mt = theta; Jt = theta; Ct = theta; Ftint = theta;
    /* get_M, get_JK, get_C, var_update (src/operators_struct.F90:131-229) hold
     * deck constants; set them now so single operators work before a step */
    for (size_t t = 0; t < nt; ++t) {
      Mt[t] = bdy[tors[t]].itheta; Jt[t] = bdy[tors[t]].ktheta; Ct[t] = bdy[tors[t]].ctheta;
      Ftint[t] = Jt[t] * theta[t];
    }
    int rc = setup_reg();
    if (rc != IB_OK) return rc;
    /* initial_condition: src/variables.F90:262-295 */
    omega.assign(L, std::vector<double>(P, 0.0));
    s.assign(L, std::vector<double>(P, 0.0));
    rhs_old.assign(L, std::vector<double>(P, 0.0));
    qx.assign(L, std::vector<double>(Fx, 0.0));
    qy.assign(L, std::vector<double>(Fy, 0.0));
    q0x.assign(L, std::vector<double>(Fx, 0.0));
    q0y.assign(L, std::vector<double>(Fy, 0.0));
    itime = prm.istart;
    /* the reference factors at the first advance (src/timestepper.F90:76-78);
     * the weights are the same then, so the factors are the same */
    if (prm.num_stat) {
      rc = preprocess();
      if (rc != IB_OK) return rc;
    }
    /* the reference precomputes B and P at the first advance
     * (src/timestepper.F90:79-97); neither depends on the flow state */
    if (sd) {
      sd_on = true;
      sdc = *sd;
      if (sdc.droptol == 0.0) sdc.droptol = 7.0e-3;
      if (sdc.interp_points == 0) sdc.interp_points = 4;
      if (sdc.support_nz == 0) sdc.support_nz = 10;
      sd_np = sdc.interp_points;
      rc = setup_subdomain();
      if (rc != IB_OK) return rc;
      rc = subdomain_interp();
      if (rc != IB_OK) return rc;
    }
    return IB_OK;
  }
};

/* ------------------------------------------------------------------------ */
/* C ABI                                                                     */
/* ------------------------------------------------------------------------ */

static std::string g_create_err;

extern "C" {

int orc_create(const ib_params* p, int32_t nbodies, const ib_body* bodies, orc_handle** out) {
  if (!p || !out) return IB_E_INVAL;
  orc_handle* h = new (std::nothrow) orc_handle();
  if (!h) return IB_E_INVAL;
  int rc = h->init(p, nbodies, bodies);
  if (rc != IB_OK) {
    g_create_err = h->err;
    delete h;
    *out = nullptr;
    return rc;
  }
  *out = h;
  return IB_OK;
}

int orc_create_subdomain(const ib_params* p, const ib_subdomain* sd, int32_t nbodies, const ib_body* bodies,
                         orc_handle** out) {
  if (!p || !sd || !out) return IB_E_INVAL;
  orc_handle* h = new (std::nothrow) orc_handle();
  if (!h) return IB_E_INVAL;
  int rc = h->init(p, nbodies, bodies, sd);
  if (rc != IB_OK) {
    g_create_err = h->err;
    delete h;
    *out = nullptr;
    return rc;
  }
  *out = h;
  return IB_OK;
}

int orc_subdomain_info(orc_handle* h, int32_t* nb_sd, int64_t* nnz_evaluated, int64_t* nnz_kept) {
  if (!h) return IB_E_INVAL;
  if (nb_sd) *nb_sd = h->sd_on ? h->nb_sd : 0;
  if (nnz_evaluated) *nnz_evaluated = h->sd_nnz_eval;
  if (nnz_kept) *nnz_kept = h->sd_nnz_kept;
  return IB_OK;
}

int orc_op_subdomain_matrix(orc_handle* h, double* lumat) {
  if (!h || !lumat || !h->sd_on) return IB_E_INVAL;
  h->compute_operator();
  std::copy(h->lumat.begin(), h->lumat.end(), lumat);
  return IB_OK;
}

/* the interpolation P: np sub-domain entries and weights per force entry,
 * and the sub-domain point coordinates (oracle-only, for tests) */
int orc_op_subdomain_interp(orc_handle* h, int32_t* idx, double* w, double* sdx, double* sdy) {
  if (!h || !h->sd_on) return IB_E_INVAL;
  if (idx) for (size_t a = 0; a < h->pidx.size(); ++a) idx[a] = h->pidx[a];
  if (w) std::copy(h->pw.begin(), h->pw.end(), w);
  if (sdx) std::copy(h->sd_x.begin(), h->sd_x.end(), sdx);
  if (sdy) std::copy(h->sd_y.begin(), h->sd_y.end(), sdy);
  return IB_OK;
}

int orc_destroy(orc_handle* h) {
  delete h;
  return IB_OK;
}

int orc_get_sizes(const orc_handle* h, ib_sizes* o) {
  o->m = h->m; o->n = h->n; o->mgridlev = h->L; o->nbc = h->nbc;
  o->n_nodes = h->P; o->n_xfaces = h->Fx; o->n_yfaces = h->Fy;
  o->nb = h->nb; o->nf = h->nf; o->nbodies = (int)h->bdy.size(); o->ntors = (int)h->tors.size();
  return IB_OK;
}

int orc_last_error(const orc_handle* h, char* buf, int32_t len) {
  const std::string& e = h ? h->err : g_create_err;
  if (buf && len > 0) {
    std::snprintf(buf, (size_t)len, "%s", e.c_str());
  }
  return (int)e.size();
}

int orc_advance(orc_handle* h, int32_t* itime, ib_step_info* info) {
  int rc = h->advance();
  if (itime) *itime = h->itime;
  if (info) {
    info->itime = h->itime; info->ksp_iter = h->ksp_iter; info->fsi_iter = h->fsi_iter;
    info->status = rc; info->err_fsi = h->err_fsi;
  }
  return rc;
}

int orc_get_level_fields(orc_handle* h, int32_t k, double* omega, double* s, double* qx, double* qy,
                         double* rhs_old) {
  if (k < 1 || k > h->L) return IB_E_INVAL;
  int q = k - 1;
  if (omega) std::memcpy(omega, h->omega[q].data(), sizeof(double) * h->P);
  if (s) std::memcpy(s, h->s[q].data(), sizeof(double) * h->P);
  if (qx) std::memcpy(qx, h->qx[q].data(), sizeof(double) * h->Fx);
  if (qy) std::memcpy(qy, h->qy[q].data(), sizeof(double) * h->Fy);
  if (rhs_old) std::memcpy(rhs_old, h->rhs_old[q].data(), sizeof(double) * h->P);
  return IB_OK;
}

int orc_set_level_fields(orc_handle* h, int32_t k, const double* omega, const double* s, const double* qx,
                         const double* qy, const double* rhs_old) {
  if (k < 1 || k > h->L) return IB_E_INVAL;
  int q = k - 1;
  if (omega) std::memcpy(h->omega[q].data(), omega, sizeof(double) * h->P);
  if (s) std::memcpy(h->s[q].data(), s, sizeof(double) * h->P);
  if (qx) std::memcpy(h->qx[q].data(), qx, sizeof(double) * h->Fx);
  if (qy) std::memcpy(h->qy[q].data(), qy, sizeof(double) * h->Fy);
  if (rhs_old) std::memcpy(h->rhs_old[q].data(), rhs_old, sizeof(double) * h->P);
  return IB_OK;
}

int orc_get_body(orc_handle* h, double* f, double* f_rdst, double* xb, double* yb, double* theta,
                 double* thetad, double* thetadd) {
  size_t nf = h->nf, nt = h->tors.size();
  if (f) std::memcpy(f, h->f.data(), sizeof(double) * nf);
  if (f_rdst) std::memcpy(f_rdst, h->f_rdst.data(), sizeof(double) * nf);
  if (xb) std::memcpy(xb, h->xb.data(), sizeof(double) * nf);
  if (yb) std::memcpy(yb, h->yb.data(), sizeof(double) * nf);
  if (theta && nt) std::memcpy(theta, h->theta.data(), sizeof(double) * nt);
  if (thetad && nt) std::memcpy(thetad, h->thetad.data(), sizeof(double) * nt);
  if (thetadd && nt) std::memcpy(thetadd, h->thetadd.data(), sizeof(double) * nt);
  return IB_OK;
}

/* family 3 (no reference counterpart): beam k's state u, ud, udd (2 ne
 * values each: per free node the normal deflection and the slope); *nd
 * gets 2 ne.  Any pointer may be NULL. */
int orc_get_beam(orc_handle* h, int32_t k, double* u, double* ud, double* udd, int32_t* nd) {
  if (k < 0 || k >= (int)h->beams.size()) return IB_E_INVAL;
  const Beam& be = h->beams[k];
  if (nd) *nd = be.nd;
  if (u) std::memcpy(u, be.u.data(), sizeof(double) * be.nd);
  if (ud) std::memcpy(ud, be.ud.data(), sizeof(double) * be.nd);
  if (udd) std::memcpy(udd, be.udd.data(), sizeof(double) * be.nd);
  return IB_OK;
}

/* calculate_pressure (src/operators_pressure.F90:110-437) at the current
 * state: mgridlev blocks of m*n pressure coefficients (write_pressure) */
int orc_pressure(orc_handle* h, double* p) {
  if (!p) return IB_E_INVAL;
  h->pressure(p);
  return IB_OK;
}

/* oracle-only: ddti (src/myfft.F90:192-203) of one m x n cell array */
int orc_op_ddti(orc_handle* h, const double* rhs, double* p) {
  h->ddti(rhs, p);
  return IB_OK;
}

/* oracle-only: coarsify_pressure (src/operators_pressure.F90:530-563) */
int orc_op_coarsify_pressure(orc_handle* h, const double* fine, double* coarse) {
  h->coarsify_pressure(fine, coarse);
  return IB_OK;
}

/* read_variables' body part (src/variables.F90:871-895) + reg_prep (:364-373) */
int orc_set_body(orc_handle* h, const double* f, const double* xb, const double* yb, const double* theta,
                 const double* thetad, const double* thetadd) {
  size_t nf = h->nf, nt = h->tors.size();
  if (f) std::memcpy(h->f.data(), f, sizeof(double) * nf);
  if (xb) std::memcpy(h->xb.data(), xb, sizeof(double) * nf);
  if (yb) std::memcpy(h->yb.data(), yb, sizeof(double) * nf);
  if (theta && nt) std::memcpy(h->theta.data(), theta, sizeof(double) * nt);
  if (thetad && nt) std::memcpy(h->thetad.data(), thetad, sizeof(double) * nt);
  if (thetadd && nt) std::memcpy(h->thetadd.data(), thetadd, sizeof(double) * nt);
  int rc = h->setup_reg();
  if (rc != IB_OK) return rc;
  if (h->prm.num_stat) rc = h->preprocess();
  if (rc == IB_OK && h->sd_on) rc = h->subdomain_interp();
  return rc;
}

/* write_force2: src/operators_struct.F90:711-731 */
int orc_get_forces(orc_handle* h, int32_t which, double* fxo, double* fyo) {
  const std::vector<double>& fv = which ? h->f_rdst : h->f;
  int nbod = (int)h->bdy.size();
  std::vector<double> ax(nbod, 0.0), ay(nbod, 0.0);
  for (int k = 0; k < h->nf; ++k) {
    if (h->ccomp[k] == 1) ax[h->cbody[k]] = ax[h->cbody[k]] + fv[k] / h->dt;
    else ay[h->cbody[k]] = ay[h->cbody[k]] + fv[k] / h->dt;
  }
  for (int b = 0; b < nbod; ++b) {
    fxo[b] = 2 * h->delta * ax[b];
    fyo[b] = 2 * h->delta * ay[b];
  }
  return IB_OK;
}

int orc_op_dst2d(orc_handle* h, int32_t batch, const double* in, double* out) {
  for (int b = 0; b < batch; ++b) {
    std::memcpy(out + (size_t)b * h->P, in + (size_t)b * h->P, sizeof(double) * h->P);
    h->dst2d(out + (size_t)b * h->P);
  }
  return IB_OK;
}

int orc_op_ctci(orc_handle* h, const double* omega, double* s) {
  h->ctci(omega, s);
  return IB_OK;
}

int orc_op_ainv(orc_handle* h, const double* in, double* out) {
  std::memcpy(out, in, sizeof(double) * h->P);
  h->ainv(out);
  return IB_OK;
}

int orc_op_vort2flux(orc_handle* h, double* omega, double* s, double* qx, double* qy) {
  int L = h->L, P = h->P, Fx = h->Fx, Fy = h->Fy;
  std::vector<std::vector<double>> w(L), sv(L, std::vector<double>(P)), vx(L, std::vector<double>(Fx)),
      vy(L, std::vector<double>(Fy));
  for (int k = 0; k < L; ++k) w[k].assign(omega + (size_t)k * P, omega + (size_t)(k + 1) * P);
  h->vort2flux(w, sv, vx, vy);
  for (int k = 0; k < L; ++k) {
    std::memcpy(omega + (size_t)k * P, w[k].data(), sizeof(double) * P);
    if (s) std::memcpy(s + (size_t)k * P, sv[k].data(), sizeof(double) * P);
    if (qx) std::memcpy(qx + (size_t)k * Fx, vx[k].data(), sizeof(double) * Fx);
    if (qy) std::memcpy(qy + (size_t)k * Fy, vy[k].data(), sizeof(double) * Fy);
  }
  return IB_OK;
}

int orc_op_reg(orc_handle* h, const double* f, double* hx, double* hy) {
  std::fill(hx, hx + h->Fx, 0.0);
  std::fill(hy, hy + h->Fy, 0.0);
  h->reg(f, hx, hy);
  return IB_OK;
}

int orc_op_regT(orc_handle* h, const double* hx, const double* hy, const double* f_in, double* f_out) {
  std::memcpy(f_out, f_in, sizeof(double) * h->nf);
  h->regT(hx, hy, f_out);
  return IB_OK;
}

int orc_op_schur_matvec(orc_handle* h, const double* x, double* y) {
  h->schur(x, y);
  return IB_OK;
}

// Column-by-column construction of the Schur matrix as the reference's
// preprocess does it (src/timestepper.F90:575-590, a_times(e_j) for rigid
// bodies); here column j = ab_times(e_j), which adds b_times2 for torsional
// entries.  Column-major, nf x nf.
int orc_op_schur_matrix(orc_handle* h, double* S) {
  const int nf = h->nf;
  std::vector<double> e(nf, 0.0);
  for (int j = 0; j < nf; ++j) {
    e[j] = 1.0;
    h->schur(e.data(), S + (size_t)j * nf);
    e[j] = 0.0;
  }
  return IB_OK;
}

int orc_op_nonlinear_rot(orc_handle* h, int32_t k, const double* omega, const double* qx, const double* qy,
                         const double* bc, double* rhs) {
  if (k < 1 || k > h->L) return IB_E_INVAL;
  std::vector<double> qhx(h->Fx), qhy(h->Fy), nlx(h->Fx), nly(h->Fy);
  double fac = h->delta * std::pow(2.0, (double)(k - 1));
  for (int q = 0; q < h->Fx; ++q) qhx[q] = qx[q] + (-fac * (-1.0) + 0.0);
  for (int q = 0; q < h->Fy; ++q) qhy[q] = qy[q] + (-fac * 0.0 + 0.0);
  h->nonlinear(omega, qhx.data(), qhy.data(), bc, nlx.data(), nly.data());
  h->rot(nlx.data(), nly.data(), rhs);
  return IB_OK;
}

/* fluid_predictor of level k (src/operators_fluid.F90:649-663): omega_out =
 * S[lam1i_k (S[con1_k rhs + con2_k rhs_old + rhsbc] + lam1_k S[omega])] */
int orc_op_predictor(orc_handle* h, int32_t k, const double* omega, const double* rhs, const double* rhs_old,
                     const double* rhsbc, double* omega_out) {
  if (k < 1 || k > h->L) return IB_E_INVAL;
  std::memcpy(omega_out, omega, sizeof(double) * h->P);
  h->predictor(omega_out, rhs, rhs_old, rhsbc, k - 1);
  return IB_OK;
}

/* redistribute (src/operators_struct.F90:735-834) of a given force vector f
 * at the current weights; the handle's own f is left as it was */
int orc_op_redistribute(orc_handle* h, const double* f, double* f_rdst) {
  std::vector<double> keep = h->f;
  std::memcpy(h->f.data(), f, sizeof(double) * h->nf);
  h->redistribute();
  std::memcpy(f_rdst, h->f_rdst.data(), sizeof(double) * h->nf);
  h->f = keep;
  return IB_OK;
}

/* family 3, test operators: beam k's mass and stiffness matrices (nd x nd,
 * row-major) and shape functions H (np x nd: point p's normal deflection
 * per unit of each unknown) */
int orc_op_beam_matrices(orc_handle* h, int32_t k, double* M, double* K, double* H) {
  if (k < 0 || k >= (int)h->beams.size()) return IB_E_INVAL;
  const Beam& be = h->beams[k];
  if (M) std::memcpy(M, be.M.data(), sizeof(double) * be.M.size());
  if (K) std::memcpy(K, be.K.data(), sizeof(double) * be.K.size());
  if (H) std::memcpy(H, be.H.data(), sizeof(double) * be.H.size());
  return IB_OK;
}

/* family 3, test operator: nsteps steps of beam k alone under the fixed
 * point forces fb (x forces, then y forces, 2 np values), from and back
 * into (u, ud, udd): the step's Newmark update with each step's start state
 * as (u0, ud0, udd0); the handle's own beam state is left as it was */
int orc_op_beam_newmark(orc_handle* h, int32_t k, int32_t nsteps, const double* fb, double* u, double* ud,
                        double* udd) {
  if (k < 0 || k >= (int)h->beams.size() || nsteps < 0) return IB_E_INVAL;
  Beam be = h->beams[k];
  be.u.assign(u, u + be.nd);
  be.ud.assign(ud, ud + be.nd);
  be.udd.assign(udd, udd + be.nd);
  for (int n = 0; n < nsteps; ++n) {
    be.u0 = be.u;
    be.ud0 = be.ud;
    be.udd0 = be.udd;
    h->beam_update(be, fb);
  }
  std::memcpy(u, be.u.data(), sizeof(double) * be.nd);
  std::memcpy(ud, be.ud.data(), sizeof(double) * be.nd);
  std::memcpy(udd, be.udd.data(), sizeof(double) * be.nd);
  return IB_OK;
}

int orc_op_gmres(orc_handle* h, const double* b, double* x, int32_t* its) {
  int it = 0;
  int rc = h->gmres(b, x, &it);
  if (its) *its = it;
  return rc;
}

int orc_op_btimes(orc_handle* h, const double* x, double* y) {
  std::fill(y, y + h->nf, 0.0);
  h->b_times2(x, y);
  return IB_OK;
}

int orc_set_theta(orc_handle* h, const double* theta, const double* thetad, const double* thetadd) {
  size_t nt = h->tors.size();
  for (size_t t = 0; t < nt; ++t) {
    if (theta) h->theta[t] = theta[t];
    if (thetad) h->thetad[t] = thetad[t];
    if (thetadd) h->thetadd[t] = thetadd[t];
  }
  return IB_OK;
}

int orc_op_body_update(orc_handle* h, const double* f, double* dtheta) {
  h->theta0 = h->theta;
  h->thetad0 = h->thetad;
  h->thetadd0 = h->thetadd;
  std::memcpy(h->f.data(), f, sizeof(double) * h->nf);
  h->compute_dx();
  for (size_t t = 0; t < h->tors.size(); ++t) dtheta[t] = h->dtheta[t];
  return IB_OK;
}

int orc_op_rot(orc_handle* h, const double* hx, const double* hy, double* vort) {
  h->rot(hx, hy, vort);
  return IB_OK;
}

int orc_op_curl(orc_handle* h, const double* s, const double* sbc, double* qx, double* qy) {
  h->curl(s, sbc, qx, qy);
  return IB_OK;
}

int orc_op_coarsify(orc_handle* h, const double* fine, double* coarse) {
  h->coarsify(fine, coarse);
  return IB_OK;
}

int orc_op_get_bc(orc_handle* h, const double* coarse, double fac, double* bc) {
  h->get_bc(coarse, bc, fac);
  return IB_OK;
}

int orc_op_apply_bc(orc_handle* h, double* vort, const double* bc, double fac) {
  h->apply_bc(vort, bc, fac);
  return IB_OK;
}

int orc_op_weights(orc_handle* h, double* weight, int32_t* reg_ix, int32_t* reg_iy) {
  if (weight) std::memcpy(weight, h->weight.data(), sizeof(double) * h->weight.size());
  if (reg_ix) for (size_t i = 0; i < h->reg_ix.size(); ++i) reg_ix[i] = h->reg_ix[i];
  if (reg_iy) for (size_t i = 0; i < h->reg_iy.size(); ++i) reg_iy[i] = h->reg_iy[i];
  return IB_OK;
}

double orc_deltafnc(double x1, double x2, double dr, int32_t f32) {
  return orc_handle::deltafnc(x1, x2, dr, f32 != 0);
}

int orc_dst1d(int32_t N, int32_t count, const double* in, double* out) {
  DST1 d;
  d.init(N);
  std::vector<cplx> z(2 * d.M), t(2 * d.M);
  for (int c = 0; c < count; c += 2) {
    const double* xa = in + (size_t)c * N;
    const double* xb = c + 1 < count ? in + (size_t)(c + 1) * N : nullptr;
    double* ya = out + (size_t)c * N;
    double* yb2 = c + 1 < count ? out + (size_t)(c + 1) * N : nullptr;
    d.pair(xa, xb, 1, ya, yb2, 1, z.data(), t.data());
  }
  return IB_OK;
}

int orc_set_threads(int32_t nthreads) {
  g_threads = nthreads > 0 ? nthreads : 1;
  return IB_OK;
}

int orc_get_threads(void) { return g_threads; }

}  // extern "C"
