// ib_batch.hip -- many independent simulations advanced one step together.
//
// Fused form: when every simulation's step has no data-dependent control
// flow (rigid bodies with the num_stat solve), one batched step is a fixed
// sequence of table launches, each covering all simulations that take part
// in it (kernel arguments of simulation e in tab[e], blockIdx.z = e), and
// the whole sequence is captured once as a HIP graph (launch_mode 1 of the
// first simulation: the same launches issued directly each step).  Simulations of
// uneven grid size and level count share each launch: levels are aligned
// from the finest (level index 0) and a simulation joins the launches of
// the levels it has.  Every simulation runs exactly the arithmetic of its
// own plain step (the same kernels, the same operands in the same order),
// so batched and plain results are identical.
//
// Concurrent form (every other mix: torsional bodies, GMRES solves with
// host-side convergence control): one host thread per simulation drives its
// plain step on the simulation's own stream.
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>

#include "ib_handle.h"

namespace {

// One launch of the fused step: a table kernel, its grid and its table.
struct BatchOp {
  const void* fn = nullptr;
  dim3 grid;
  size_t lds = 0;
  size_t off = 0;  // table offset in the arena
  int count = 0;
};

// Host staging of the launch tables; all tables go to one device arena.
struct OpList {
  std::vector<BatchOp> ops;
  std::vector<char> bytes;

  template <class A>
  void add(const void* fn, const std::vector<TabEntry<A>>& es, size_t lds = 0) {
    if (es.empty()) return;
    BatchOp op;
    op.fn = fn;
    op.lds = lds;
    op.count = (int)es.size();
    int gx = 1, gy = 1;
    for (const auto& e : es) {
      gx = std::max(gx, e.nbx);
      gy = std::max(gy, e.nby);
    }
    op.grid = dim3(gx, gy, op.count);
    const size_t align = 64;
    op.off = (bytes.size() + align - 1) / align * align;
    bytes.resize(op.off + sizeof(TabEntry<A>) * es.size());
    std::memcpy(bytes.data() + op.off, es.data(), sizeof(TabEntry<A>) * es.size());
    ops.push_back(op);
  }
};

std::string g_batch_create_err;

// The cooperative GMRES solves of a concurrent batch, run together: each
// simulation's host thread hands its solve over when it reaches one
// (solve) and waits; when every simulation still stepping has handed one
// over, or has finished its step (leave), the last to arrive enqueues the
// waiting solves, up to one per XCD per launch (gmres_coopw_multi), on its
// own stream after the others' streams have reached their solves, and each
// other stream waits for the launch.  Separate launches would not do: a
// launch holds workgroups on every XCD, and those off its own XCD cannot
// start (nor the launch end) while another solve holds that XCD's CUs.
class CoopGroup : public SolveGroup {
 public:
  CoopGroup(int n, const std::vector<int>& xcc) : xcc_(xcc), ready_(n), done_(n) {
    for (int i = 0; i < n; ++i) {
      IB_HIP(hipEventCreateWithFlags(&ready_[i], hipEventDisableTiming));
      IB_HIP(hipEventCreateWithFlags(&done_[i], hipEventDisableTiming));
    }
  }
  ~CoopGroup() override {
    for (hipEvent_t e : ready_) (void)hipEventDestroy(e);
    for (hipEvent_t e : done_) (void)hipEventDestroy(e);
  }
  // n simulations start stepping
  void enter(int n) {
    std::lock_guard<std::mutex> lk(mu_);
    active_ = n;
    pend_.clear();
  }
  // a simulation has finished stepping (st: its stream, idle for it now)
  void leave(hipStream_t st) {
    std::lock_guard<std::mutex> lk(mu_);
    --active_;
    if (!pend_.empty() && (int)pend_.size() == active_) launch_locked(st);
  }
  bool solve(int member, const CoopWJob& job, hipStream_t st) override {
    std::unique_lock<std::mutex> lk(mu_);
    if (active_ <= 0) return false;
    IB_HIP(hipEventRecord(ready_[member], st));
    pend_.push_back({member, job});
    const unsigned long long mine = round_;
    if ((int)pend_.size() == active_) launch_locked(st);
    else cv_.wait(lk, [&] { return round_ != mine; });
    const std::string e = err_;
    lk.unlock();
    if (!e.empty()) throw HipError(e);
    IB_HIP(hipStreamWaitEvent(st, done_[member], 0));
    return true;
  }

 private:
  struct Pending {
    int member;
    CoopWJob job;
  };
  void launch_locked(hipStream_t st) {
    err_.clear();
    try {
      for (const Pending& p : pend_) IB_HIP(hipStreamWaitEvent(st, ready_[p.member], 0));
      const size_t nx = xcc_.size();
      for (size_t k0 = 0; k0 < pend_.size(); k0 += nx) {
        const int n = (int)std::min(nx, pend_.size() - k0);
        std::vector<CoopWJob> jobs;
        for (int k = 0; k < n; ++k) jobs.push_back(pend_[k0 + k].job);
        gmres_coopw_multi(st, n, jobs.data(), xcc_.data(), (int)nx);
      }
      for (const Pending& p : pend_) IB_HIP(hipEventRecord(done_[p.member], st));
    } catch (const std::exception& e) {
      err_ = e.what();
    }
    pend_.clear();
    ++round_;
    cv_.notify_all();
  }
  std::vector<int> xcc_;
  std::vector<hipEvent_t> ready_, done_;
  std::mutex mu_;
  std::condition_variable cv_;
  int active_ = 0;
  unsigned long long round_ = 0;
  std::vector<Pending> pend_;
  std::string err_;
};

}  // namespace

struct ib_batch {
  std::vector<ib_handle*> sims;
  std::mutex coop_mu[8];  // per XCD: turns of the cooperative GMRES solves (concurrent form)
  std::unique_ptr<CoopGroup> group;  // concurrent form: the solves run together
  std::string err;
  int mode = 0;  // 1 fused, 2 concurrent
  hipStream_t st = nullptr;
  OpList plan;
  DBuf<char> arena;
  DBuf<int> step;          // itime of the step being taken, per simulation
  int* hstep = nullptr;    // pinned staging of step
  hipGraphExec_t ge = nullptr;
  bool direct = false;     // launch_mode 1 of simulation 0: plain launches, no graph

  ~ib_batch() {
    if (st) (void)hipStreamSynchronize(st);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (hstep) (void)hipHostFree(hstep);
    arena.release();
    step.release();
    if (st) (void)hipStreamDestroy(st);
    for (ib_handle* h : sims) delete h;
  }

  // rigid bodies and the precomputed num_stat inverse: nothing in the step
  // depends on values read back to the host
  static bool fusable(const ib_handle* h) { return h->ntors == 0 && h->nf > 0 && h->prm.num_stat; }

  // ---------------------------------------------------------- fused plan
  // Tile shapes of the fused DST table launches.  One launch covers every
  // simulation, so it is throughput-bound where one simulation's launch is
  // latency-bound, and the best tiles differ from each handle's own:
  // tune_tiles times a row-pass and a column-pass table launch over all
  // simulations for each rule.  Rules: 0 each handle's tiles, -1 the fixed
  // rule (2 row pairs up to transform length 512, else 1; 1 column pair),
  // k > 0 k pairs (halved where the LDS is short).  Tiles never change
  // results.
  int xrule = 0, yrule = 0;
  // the fused row-pass launches use the kernels with the generic-radix stage
  // only when some simulation's row plan needs it
  bool genx = false;
  int xtile(const ib_handle* h) const {
    return xrule > 0 ? xrule : (xrule < 0 ? (h->px.dev.L <= 512 ? 2 : 1) : h->tiles.rp);
  }
  int ytile(const ib_handle* h) const { return yrule > 0 ? yrule : (yrule < 0 ? 1 : h->tiles.cp); }
  TabEntry<XArgs> xe(const ib_handle* h, const DstLaunch& l, size_t* lds) const {
    if (!genx && fft_has_generic(h->px.dev))
      throw HipError("fused batch: a row plan with a radix >= 7 under the row-pass kernels without that stage");
    return dst_x_entry(h->px.dev, l, h->g, xtile(h), h->max_lds, lds);
  }
  TabEntry<YArgs> ye(const ib_handle* h, int mode, const DstLaunch& l, size_t* lds) const {
    return dst_y_entry(h->py.dev, mode, l, h->g, ytile(h), h->max_lds, lds);
  }

  // best of 3 timings of 2 back-to-back launches of one table op
  template <class A>
  float time_tab(const void* fn, const std::vector<TabEntry<A>>& es, size_t lds) {
    OpList ol;
    ol.add(fn, es, lds);
    DBuf<char> ar;
    ar.alloc(ol.bytes.size());
    IB_HIP(hipMemcpy(ar.p, ol.bytes.data(), ol.bytes.size(), hipMemcpyHostToDevice));
    const BatchOp& op = ol.ops[0];
    const void* tab = ar.p + op.off;
    void* args[] = {(void*)&tab};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    IB_HIP(hipEventCreate(&e0));
    IB_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    try {
      for (int rep = 0; rep < 4; ++rep) {
        IB_HIP(hipEventRecord(e0, st));
        for (int k = 0; k < 2; ++k) IB_HIP(hipLaunchKernel(op.fn, op.grid, dim3(256), args, op.lds, st));
        IB_HIP(hipEventRecord(e1, st));
        IB_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        IB_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0) best = std::min(best, ms);  // the first round warms up
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
  }

  // row pass t0 -> t1 and column pass (scaled by laminv) t1 -> t0 of every
  // simulation (scratch fields of the step) under each rule; the environment's choice
  // (IB_DST_TILES, IB_DST_TUNE=0) stays with the handles' tiles
  void tune_tiles() {
    xrule = yrule = 0;
    if (std::getenv("IB_DST_TILES")) return;
    if (const char* f = std::getenv("IB_DST_TUNE"))
      if (std::strcmp(f, "0") == 0) return;
    size_t t = 0;
    float best = 1e30f;
    int bx = 0;
    for (int r : {0, -1, 1, 2, 4}) {
      xrule = r;
      std::vector<TabEntry<XArgs>> es;
      size_t l = 0;
      for (ib_handle* h : sims) {
        DstLaunch a;
        a.in[0] = h->t0; a.out[0] = h->t1;
        es.push_back(xe(h, a, &t));
        l = std::max(l, t);
      }
      const float ms = time_tab(dst_x_tab_kernel(X_PLAIN, genx), es, l);
      if (ms < best) {
        best = ms;
        bx = r;
      }
    }
    xrule = bx;
    best = 1e30f;
    int by = 0;
    for (int r : {0, -1, 2, 4}) {
      yrule = r;
      std::vector<TabEntry<YArgs>> es;
      size_t l = 0;
      for (ib_handle* h : sims) {
        DstLaunch a;
        a.in[0] = h->t1; a.out[0] = h->t0; a.tab = h->laminv; a.div = 1.0;
        es.push_back(ye(h, Y_SCALE2, a, &t));
        l = std::max(l, t);
      }
      const float ms = time_tab(dst_y_tab_kernel(Y_SCALE2), es, l);
      if (ms < best) {
        best = ms;
        by = r;
      }
    }
    yrule = by;
  }

  template <class A>
  using Tab = std::vector<TabEntry<A>>;

  // vort2flux (src/operators_fluid.F90:237-377) of every simulation: the
  // launch sequence of ib_handle::vort2flux_dev, level by level.  A
  // simulation's coarsest ctci takes its zero boundary vector through the
  // apply_bc form, which adds exact zeros.
  void plan_vort2flux(int Lmax) {
    for (int k = 1; k < Lmax; ++k) {
      Tab<CoarsenArgs> c;
      for (ib_handle* h : sims)
        if (k < h->L) c.push_back(coarsify_entry(h->g, h->omega + (k - 1) * (long)h->g.P, h->omega + k * (long)h->g.P, 1, 0));
      plan.add(fluid_tab_kernel(TK_COARSIFY), c);
    }
    std::vector<const double*> sbc(sims.size());
    for (size_t i = 0; i < sims.size(); ++i) sbc[i] = sims[i]->zbc;
    for (int k = Lmax - 1; k >= 0; --k) {
      Tab<XArgs> x1, x2;
      Tab<YArgs> y1;
      Tab<CurlArgs> cu;
      size_t l1 = 0, l2 = 0, l3 = 0, t;
      for (size_t i = 0; i < sims.size(); ++i) {
        ib_handle* h = sims[i];
        if (k >= h->L) continue;
        const long P = h->g.P;
        DstLaunch a;
        a.in[0] = h->omega + k * P; a.out[0] = h->t0; a.bc = sbc[i];
        x1.push_back(xe(h, a, &t)); l1 = std::max(l1, t);
        DstLaunch b;
        b.in[0] = h->t0; b.out[0] = h->t1; b.tab = h->laminv; b.div = 1.0;
        y1.push_back(ye(h, Y_SCALE2, b, &t)); l2 = std::max(l2, t);
        DstLaunch c;
        c.in[0] = h->t1; c.out[0] = h->s + k * P;
        x2.push_back(xe(h, c, &t)); l3 = std::max(l3, t);
        double* bout = nullptr;
        if (k > 0) bout = (sbc[i] == h->sbcA.p) ? h->sbcB.p : h->sbcA.p;
        cu.push_back(curl_getbc_entry(h->g, h->s + k * P, sbc[i], h->qx + (long)k * h->g.Fx,
                                      h->qy + (long)k * h->g.Fy, bout));
        sbc[i] = bout;
      }
      plan.add(dst_x_tab_kernel(X_APPLYBC, genx), x1, l1);
      plan.add(dst_y_tab_kernel(Y_SCALE2), y1, l2);
      plan.add(dst_x_tab_kernel(X_PLAIN, genx), x2, l3);
      plan.add(fluid_tab_kernel(TK_CURL_GETBC), cu);
    }
  }

  // advance(itime) for num_stat = T, rigid bodies: the launch sequence of
  // ib_handle::advance (src/timestepper.F90:18-556) over all simulations
  void plan_step() {
    int Lmax = 0;
    for (ib_handle* h : sims) Lmax = std::max(Lmax, h->L);
    size_t t;
    // step 1: lastbc (src/timestepper.F90:103-117)
    {
      Tab<GetBcArgs> gb;
      Tab<FillArgs> fl;
      for (ib_handle* h : sims) {
        const long P = h->g.P, nbc = h->g.nbc;
        if (h->L > 1) gb.push_back(get_bc_entry(h->g, h->omega + P, P, h->lastbc, nbc, h->L - 1, 0.25));
        fl.push_back(fill_entry(h->lastbc + (h->L - 1) * nbc, nbc, 0.0));
      }
      plan.add(fluid_tab_kernel(TK_GET_BC), gb);
      plan.add(fluid_tab_kernel(TK_FILL), fl);
    }
    // predictor, coarse to fine
    for (int k = Lmax - 1; k >= 0; --k) {
      Tab<GetBcArgs> gb;
      Tab<NLArgs> nl;
      Tab<XArgs> x1, x2;
      Tab<YArgs> y1;
      size_t l1 = 0, l2 = 0, l3 = 0;
      for (size_t i = 0; i < sims.size(); ++i) {
        ib_handle* h = sims[i];
        if (k >= h->L) continue;
        const long P = h->g.P, nbc = h->g.nbc;
        const double* ob = h->zbc;
        if (k < h->L - 1) {
          gb.push_back(get_bc_entry(h->g, h->omega + (k + 1) * P, P, h->obc, nbc, 1, 0.25));
          ob = h->obc;
        }
        nl.push_back(nl_rhs_entry(h->g, h->omega + k * P, h->qx + (long)k * h->g.Fx, h->qy + (long)k * h->g.Fy,
                                  h->q0x[k], h->lastbc + k * nbc, ob, h->vfac[k], h->con1[k], h->con2[k],
                                  step.p + i, h->rhs_old + k * P, h->t2));
        DstLaunch a;
        a.batch = 2;
        a.in[0] = h->t2; a.out[0] = h->t0;
        a.in[1] = h->omega + k * P; a.out[1] = h->t1;
        x1.push_back(xe(h, a, &t)); l1 = std::max(l1, t);
        DstLaunch b;
        b.in[0] = h->t0; b.in2[0] = h->t1; b.out[0] = h->t2; b.tab = h->lam1 + k * P; b.tab2 = h->lam1i + k * P;
        y1.push_back(ye(h, Y_PRED, b, &t)); l2 = std::max(l2, t);
        DstLaunch c;
        c.in[0] = h->t2; c.out[0] = h->omega + k * P;
        x2.push_back(xe(h, c, &t)); l3 = std::max(l3, t);
      }
      plan.add(fluid_tab_kernel(TK_GET_BC), gb);
      plan.add(fluid_tab_kernel(TK_NL_RHS), nl);
      plan.add(dst_x_tab_kernel(X_PLAIN, genx), x1, l1);
      plan.add(dst_y_tab_kernel(Y_PRED), y1, l2);
      plan.add(dst_x_tab_kernel(X_PLAIN, genx), x2, l3);
    }
    plan_vort2flux(Lmax);
    // step 2, num_stat (src/timestepper.F90:457-484): rhsf = regT(q + q0),
    // f = A^-1 rhsf, redistribute
    {
      Tab<FsiRhsArgs> fr;
      Tab<GemvArgs> gv;
      Tab<RegArgs> rd, rg;
      Tab<RegTArgs> rt;
      for (ib_handle* h : sims) {
        fr.push_back(fsi_rhs_entry(h->g, h->nf, h->no_tors, h->tb, 0, h->th, h->dt, h->delta, h->w, h->csr.key,
                                   h->qx, h->qy, h->q0x[0], h->rhsf));
        gv.push_back(gemv_entry(h->nf, h->Ainv, h->rhsf, h->f));
        rd.push_back(redist_entry(h->g, h->csr, h->w, h->f, h->hx, h->hy));
        rt.push_back(regT_entry(h->g, h->nf, h->w, h->csr.key, h->hx, h->hy, 0.0, nullptr, h->f_rdst));
        rg.push_back(reg_entry(h->g, h->csr, h->w, h->f, h->hx, h->hy));
      }
      plan.add(body_tab_kernel(TK_FSI_RHS), fr);
      plan.add(gemv_tab_kernel(), gv);
      plan.add(body_tab_kernel(TK_REDIST), rd);
      plan.add(body_tab_kernel(TK_REGT), rt);
      // step 3: corrector (src/timestepper.F90:488-553)
      plan.add(body_tab_kernel(TK_REG), rg);
    }
    {
      Tab<XArgs> x1, x2;
      Tab<YArgs> y1;
      size_t l1 = 0, l2 = 0, l3 = 0;
      for (ib_handle* h : sims) {
        DstLaunch a;
        a.in[0] = nullptr; a.out[0] = h->t0; a.hx = h->hx; a.hy = h->hy;
        x1.push_back(xe(h, a, &t)); l1 = std::max(l1, t);
        DstLaunch b;
        b.in[0] = h->t0; b.out[0] = h->t1; b.tab = h->lam1i; b.div = h->normalize;
        y1.push_back(ye(h, Y_SCALE2, b, &t)); l2 = std::max(l2, t);
        DstLaunch c;
        c.in[0] = h->t1; c.out[0] = h->omega;
        x2.push_back(xe(h, c, &t)); l3 = std::max(l3, t);
      }
      plan.add(dst_x_tab_kernel(X_ROT, genx), x1, l1);
      plan.add(dst_y_tab_kernel(Y_SCALE2), y1, l2);
      plan.add(dst_x_tab_kernel(X_SUBFROM, genx), x2, l3);
    }
    plan_vort2flux(Lmax);
  }

  void launch_plan() {
    for (const BatchOp& op : plan.ops) {
      const void* tab = arena.p + op.off;
      void* args[] = {(void*)&tab};
      IB_HIP(hipLaunchKernel(op.fn, op.grid, dim3(256), args, op.lds, st));
    }
  }

  void build_fused() {
    genx = false;
    for (const ib_handle* h : sims) genx = genx || fft_has_generic(h->px.dev);
    tune_tiles();
    plan_step();
    arena.alloc(plan.bytes.size());
    IB_HIP(hipMemcpy(arena.p, plan.bytes.data(), plan.bytes.size(), hipMemcpyHostToDevice));
    if (direct) return;
    hipGraph_t gr = nullptr;
    IB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      launch_plan();
    } catch (...) {
      (void)hipStreamEndCapture(st, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      throw;
    }
    IB_HIP(hipStreamEndCapture(st, &gr));
    IB_HIP(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
    IB_HIP(hipGraphDestroy(gr));
  }

  int init(int32_t nsim, const ib_params* params, const int32_t* nbodies, const ib_body* bodies) {
    if (nsim < 1 || nsim > 65535) {
      err = "nsim must be in 1..65535";
      return IB_E_INVAL;
    }
    long b0 = 0;
    for (int i = 0; i < nsim; ++i) {
      if (nbodies[i] < 0) {
        err = "negative body count";
        return IB_E_INVAL;
      }
      ib_handle* h = new ib_handle();
      sims.push_back(h);
      h->no_capture = true;  // advance and adjoint may run the simulations from several host threads
      const int rc = h->init(&params[i], nbodies[i], bodies + b0);
      if (rc != IB_OK) {
        err = "simulation " + std::to_string(i) + ": " + h->err;
        return rc;
      }
      b0 += nbodies[i];
    }
    IB_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    step.alloc(nsim);
    IB_HIP(hipHostMalloc(&hstep, sizeof(int) * nsim, hipHostMallocDefault));
    bool all = true;
    for (ib_handle* h : sims) all = all && fusable(h);
    mode = all ? 1 : 2;
    if (mode == 2) {
      // the simulations run at once, and each cooperative GMRES solve keeps
      // its P workgroups, one per CU, on one XCD until it ends: simulation i
      // solves on the device's XCD i mod 8, so up to 8 solves run side by
      // side.  When the solves of one XCD together need more than its CUs,
      // two of them could wait for each other's CUs (up to the bounded
      // spin), so they take turns: each holds that XCD's lock from its
      // launch until its work has completed, while the other simulations'
      // kernels keep running
      int ncu = 0;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, sims[0]->dev) != hipSuccess) ncu = 0;
      const std::vector<int>& xcc = sims[0]->xcc;
      // solves run together when every simulation has the cooperative solve
      // for larger systems in the same register layout
      bool together = sims.size() > 1 && !xcc.empty();
      for (ib_handle* h : sims)
        together = together && h->cwP > 0 && h->coopP == 0 && h->xcc == xcc &&
                   gmres_coopw_multi_ok(h->nf, sims[0]->nf);
      if (together) {
        group.reset(new CoopGroup((int)sims.size(), xcc));
        for (size_t i = 0; i < sims.size(); ++i) {
          sims[i]->group = group.get();
          sims[i]->group_member = (int)i;
        }
      }
      const size_t nx = xcc.empty() ? 1 : xcc.size();  // no XCD ids: one shared lock, as one XCD
      long need[8] = {};
      for (size_t i = 0; i < sims.size(); ++i) {
        sims[i]->coop_xcd = xcc.empty() ? -1 : xcc[i % nx];
        need[i % nx] += sims[i]->coopP + sims[i]->cwP;
      }
      for (size_t i = 0; i < sims.size() && !group; ++i)
        if (need[i % nx] > ncu / (long)(xcc.empty() ? 8 : nx)) sims[i]->coop_lock = &coop_mu[i % nx];
    }
    direct = params[0].launch_mode == 1;
    if (mode == 1) {
      // the fused launches run the Stockham passes for every simulation; a
      // handle's own passes follow, so that they round alike
      for (ib_handle* h : sims) h->tiles.xform = h->tiles.yform = 0;
      build_fused();
    }
    return IB_OK;
  }

  int advance(int32_t* itimes, ib_step_info* infos) {
    Range batch_range("ib_batch_advance");
    const int n = (int)sims.size();
    std::vector<int> rc(n, IB_OK);
    if (mode == 1) {
      for (int i = 0; i < n; ++i) hstep[i] = sims[i]->itime + 1;
      IB_HIP(hipMemcpyAsync(step.p, hstep, sizeof(int) * n, hipMemcpyHostToDevice, st));
      if (direct) launch_plan();
      else IB_HIP(hipGraphLaunch(ge, st));
      IB_HIP(hipStreamSynchronize(st));
      for (ib_handle* h : sims) {
        h->itime += 1;
        h->ksp_iter = 0;
        h->fsi_iter = 0;
        h->err_fsi = 0.0;
      }
    } else {
      std::vector<std::string> ex(n);
      std::vector<std::thread> pool;
      pool.reserve(n);
      if (group) group->enter(n);
      for (int i = 0; i < n; ++i)
        pool.emplace_back([&, i]() {
          try {
            rc[i] = sims[i]->advance();
          } catch (const std::exception& e) {
            ex[i] = e.what();
            rc[i] = IB_E_HIP;
          }
          if (group) group->leave(sims[i]->st);
        });
      for (auto& t : pool) t.join();
      for (int i = 0; i < n; ++i)
        if (!ex[i].empty()) sims[i]->err = ex[i];
    }
    int first = IB_OK;
    for (int i = 0; i < n; ++i) {
      ib_handle* h = sims[i];
      if (itimes) itimes[i] = h->itime;
      if (infos) {
        infos[i].itime = h->itime;
        infos[i].ksp_iter = h->ksp_iter;
        infos[i].fsi_iter = h->fsi_iter;
        infos[i].status = rc[i];
        infos[i].err_fsi = h->err_fsi;
      }
      if (rc[i] != IB_OK && first == IB_OK) {
        first = rc[i];
        err = "simulation " + std::to_string(i) + ": " + h->err;
      }
    }
    return first;
  }

  // every simulation's discrete adjoint (ib_adjoint_run) from its own host
  // thread on its own stream; cx, cy, dK, dg0 run over the bodies of
  // simulation 0, then 1, ...
  int adjoint(int nsteps, int ckpt, const double* cx, const double* cy, double* J, double* dRe, double* dK,
              double* dg0) {
    const int n = (int)sims.size();
    std::vector<int> boff(n + 1, 0);
    for (int i = 0; i < n; ++i) boff[i + 1] = boff[i] + sims[i]->nbodies;
    std::vector<int> rc(n, IB_OK);
    std::vector<std::string> ex(n);
    std::vector<std::thread> pool;
    pool.reserve(n);
    if (group) group->enter(n);
    for (int i = 0; i < n; ++i)
      pool.emplace_back([&, i]() {
        try {
          const int o = boff[i];
          rc[i] = sims[i]->adjoint_run(nsteps, ckpt, cx + o, cy + o, J ? J + i : nullptr, dRe ? dRe + i : nullptr,
                                       nullptr, nullptr, nullptr, dK ? dK + o : nullptr, dg0 ? dg0 + o : nullptr);
        } catch (const std::exception& e) {
          ex[i] = e.what();
          rc[i] = IB_E_HIP;
        }
        if (group) group->leave(sims[i]->st);
      });
    for (auto& t : pool) t.join();
    int first = IB_OK;
    for (int i = 0; i < n; ++i) {
      if (!ex[i].empty()) sims[i]->err = ex[i];
      if (rc[i] != IB_OK && first == IB_OK) {
        first = rc[i];
        err = "simulation " + std::to_string(i) + ": " + sims[i]->err;
      }
    }
    return first;
  }
};

extern "C" {

int ib_batch_create(int32_t nsim, const ib_params* params, const int32_t* nbodies, const ib_body* bodies,
                    ib_batch** out) {
  if (!params || !nbodies || !out) return IB_E_INVAL;
  *out = nullptr;
  ib_batch* b = nullptr;
  try {
    b = new ib_batch();
    const int rc = b->init(nsim, params, nbodies, bodies);
    if (rc != IB_OK) {
      g_batch_create_err = b->err;
      delete b;
      return rc;
    }
  } catch (const std::exception& e) {
    g_batch_create_err = e.what();
    delete b;
    return IB_E_HIP;
  }
  *out = b;
  return IB_OK;
}

int ib_batch_destroy(ib_batch* b) {
  delete b;
  return IB_OK;
}

int ib_batch_advance(ib_batch* b, int32_t* itimes, ib_step_info* infos) {
  if (!b) return IB_E_INVAL;
  try {
    return b->advance(itimes, infos);
  } catch (const std::exception& e) {
    b->err = e.what();
    return IB_E_HIP;
  }
}

int ib_batch_adjoint(ib_batch* b, int32_t nsteps, int32_t ckpt_every, const double* cx, const double* cy, double* J,
                     double* dJ_dRe, double* dJ_dK, double* dJ_dgamma0) {
  if (!b || !cx || !cy) return IB_E_INVAL;
  try {
    return b->adjoint(nsteps, ckpt_every, cx, cy, J, dJ_dRe, dJ_dK, dJ_dgamma0);
  } catch (const std::exception& e) {
    b->err = e.what();
    return IB_E_HIP;
  }
}

int ib_batch_handle(ib_batch* b, int32_t i, ib_handle** h) {
  if (!b || !h || i < 0 || i >= (int)b->sims.size()) return IB_E_INVAL;
  *h = b->sims[i];
  return IB_OK;
}

int ib_batch_info(const ib_batch* b, int32_t* nsim, int32_t* mode) {
  if (!b) return IB_E_INVAL;
  if (nsim) *nsim = (int32_t)b->sims.size();
  if (mode) *mode = b->mode;
  return IB_OK;
}

int ib_batch_dst_tiles(const ib_batch* b, int32_t* row_rule, int32_t* col_rule) {
  if (!b) return IB_E_INVAL;
  if (row_rule) *row_rule = b->xrule;
  if (col_rule) *col_rule = b->yrule;
  return IB_OK;
}

int ib_batch_last_error(const ib_batch* b, char* buf, int32_t len) {
  const std::string& e = b ? b->err : g_batch_create_err;
  if (buf && len > 0) std::snprintf(buf, (size_t)len, "%s", e.c_str());
  return (int)e.size();
}

}  // extern "C"
