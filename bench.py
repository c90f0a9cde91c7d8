#!/usr/bin/env python3
"""Throughput of the immersed-boundary projection time step on one MI355X.

A "step" is one call of advance(itime) (src/timestepper.F90:18-556): the
predictor on every grid level, vort2flux, the FSI fixed point with its GMRES
Schur solves, and the corrector.  The headline workload is the two-body
system at 1024x512 cells and 5 levels (the configuration the metric is
quoted on): an elliptic airfoil (chord 1, thickness 0.12, 300 points, rigid)
and a flap (chord 0.3, 100 points) on a torsional spring (K = 0.05, mass
ratio 0.5), Re = 1000, dx = 0.004, dt = 5e-4, GMRES and FSI tolerance 1e-10
(full GMRES, nf = 800), synthetic and generated in
ib_parallel_amd.decks.airfoil_flap_case.  State is resident in HBM before
timing.

Metric: Mgridpt-steps/s = m*n*mgridlev*steps / seconds / 1e6 (FP64).

Also reported:
  roofline      the kernel class with the largest share of the instrumented
                steps (on the headline workload the dense GMRES solve, one
                cooperative launch per FSI iteration; the DST pass classes
                follow in other_classes).  achieved = algorithmic bytes per
                launch over the launch duration measured with HIP events
                recorded around every launch on the library's stream, during
                a second timed region of the same workload right after the
                first (per-launch events need plain stream launches); for
                GMRES the algorithmic bytes are the operands of every Arnoldi
                step (S, the basis read twice by classical Gram-Schmidt, the
                new vector), with the FP64 rate, the time per Arnoldi step
                and what limits it beside them; peak = 8 TB/s HBM.  traffic =
                calibrated FETCH_SIZE + WRITE_SIZE bytes per launch from the
                committed counter summary profiles/pmc_traffic.json (made by
                ib_parallel_amd.pmc), null if it has no entry for the kernel.
  window        a longer run of steps after the timed region (face-set
                rebuilds included), as a steady-state check of the headline.
  cpu_baseline  the CPU oracle (oracle/, a C++ FP64 restatement of the same
                algorithm with the reference's matrix-free GMRES, OpenMP over
                DST rows) on the host cores: on the headline workload a
                bounded sample (capped-GMRES steps that give the per-step
                fixed cost and the cost per Krylov iteration, extrapolated to
                the GPU's iteration count of the timed steps), on all cores
                and on one core.
  batched       the batched form (ib_batch_*) on 64 uneven rigid-ellipse
                simulations against the same simulations as standalone
                handles stepped one after another (run_batched).
  adjoint       the discrete adjoint (ib_adjoint_run) against the plain
                forward steps of the same simulation: a rigid cylinder and
                the hinged plate at 400x400x3 (run_adjoint).
Workloads: airfoil_flap (default, the headline), plate (hinged plate
400x400x4) and cylinder (rigid, 200x200x3).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KCLASS = {0: "k_dst_x (row pass, time step)", 1: "k_dst_y (column pass, time step)",
          2: "k_dst_x (batched row pass, Schur assembly)", 3: "k_dst_y (batched column pass, Schur assembly)",
          4: "k_gmres_coopw (one dense GMRES solve over P workgroups of one XCD; k_gmres_coop for nf <= 256, "
             "k_gmres_dense in one workgroup where no split fits)"}
FP64_PEAK_TFLOPS = 78.6


def gridpts(sim):
    return sim.m * sim.n * sim.L


def make_case(name):
    from ib_parallel_amd import decks
    if name == "plate":
        p, b = decks.plate_case()
        return p, b, "hinged plate FSI 400x400x4, 101 pts, Re=200, dt=1e-3, tol=1e-10"
    if name == "airfoil_flap":
        p, b = decks.airfoil_flap_case()
        return p, b, ("elliptic airfoil (300 pts, rigid) + torsional flap (100 pts) FSI 1024x512x5, Re=1000, "
                      "dt=5e-4, tol=1e-10")
    if name == "cylinder":
        p, b = decks.cylinder_case()
        return p, b, "rigid cylinder 200x200x3, 157 pts, Re=100, dt=4e-3, tol=1e-7"
    raise SystemExit(f"unknown workload {name}")


_HIP = None


def _hip():
    """The HIP runtime through ctypes (device synchronisation and the
    calibration copy; no torch import on the measurement path): the symbols
    are looked up through libibgpu.so's handle, so they come from the runtime
    the library is linked against."""
    global _HIP
    if _HIP is None:
        from ib_parallel_amd import library
        _HIP = library()
    return _HIP


def device_sync():
    if _hip().hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def run_gpu(args):
    from ib_parallel_amd import GpuSim
    p, bodies, desc = make_case(args.workload)
    sim = GpuSim(p, bodies)
    for _ in range(args.warmup):
        sim.advance(allow_noconv=True)
    if args.pmc_copy_mb > 0:
        # the calibration copy of counter runs, between the set-up and the
        # timed steps (ib_parallel_amd.pmc averages the dispatches after it)
        copy_calibration(args.pmc_copy_mb)
    # timed region 1: the headline number, no instrumentation
    its = fsi = 0
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        info = sim.advance(allow_noconv=True)
        its += info.ksp_iter
        fsi += info.fsi_iter
    device_sync()
    t1 = time.perf_counter()
    secs = t1 - t0
    # timed region 2: HIP events around every DST pass launch
    classes = (0, 1, 2, 3, 4)
    for kc in classes:
        sim.kernel_timing(kc, True)
    nprof = max(1, args.steps // 2)
    its_prof = 0
    device_sync()
    t2 = time.perf_counter()
    for _ in range(nprof):
        its_prof += sim.advance(allow_noconv=True).ksp_iter
    device_sync()
    t3 = time.perf_counter()
    timing = {kc: sim.kernel_timing_read(kc) for kc in classes}
    for kc in classes:
        sim.kernel_timing(kc, False)
    secs_prof = t3 - t2

    def summary(kc):
        n_l, ms, by = timing[kc]
        return {"kernel": KCLASS[kc], "launches": n_l, "avg_us": round(ms / max(n_l, 1) * 1e3, 3),
                "bytes_per_launch": round(by / max(n_l, 1)), "share_of_step": round(ms * 1e-3 / secs_prof, 4)}

    dom = max(timing, key=lambda kc: timing[kc][1])
    n_l, ms, by = timing[dom]
    achieved = (by / n_l) / (ms / n_l * 1e-3) / 1e9 if n_l and ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None}
    roof.update(summary(dom))
    tr, tk = traffic_for(KCLASS[dom])
    roof["traffic"] = tr
    if tr is not None:
        roof["traffic_source"] = f"profiles/pmc_traffic.json ({tk}, FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
    if dom == 4 and n_l and ms > 0:
        # per Arnoldi step: 2 nf^2 flops for S v, 4 (j + 1) nf for the
        # Gram-Schmidt dots and update; with the operand bytes B = 8 (its nf^2
        # + 2 nf sum(j + 1) + its nf + 3 nf) per solve that is
        # 2 (B / 8 - its nf - 3 nf)
        nf = sim.nf
        flops = 2.0 * (by / 8.0 - its_prof * nf - 3.0 * nf * n_l)
        form, wgs, xch = sim.gmres_info()
        roof["fp64_tflops"] = round(flops / (ms * 1e-3) / 1e12, 4)
        roof["fp64_frac"] = round(flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 5)
        roof["arnoldi_steps_per_launch"] = round(its_prof / n_l, 1)
        roof["us_per_arnoldi_step"] = round(ms * 1e3 / max(its_prof, 1), 3)
        roof["gmres_form"] = {0: "per-stage kernels", 1: "one workgroup", 2: "cooperative (nf <= 256)",
                              3: "cooperative (nf <= 1024)"}[form]
        roof["workgroups"] = wgs
        roof["exchange"] = {0: "none", 1: "through one XCD's L2", 2: "device scope"}[xch]
        roof["limited_by"] = ("latency: three grid barriers and a few dependent L2 round trips per Arnoldi step "
                              "across the workgroups of one XCD; S and the basis stay on chip, so HBM is idle")
        roof["algorithmic_bytes"] = ("per Arnoldi step: S (nf^2 doubles), basis vectors 0..j read twice by "
                                     "classical Gram-Schmidt, the new vector written; b and x per solve")
    roof["other_classes"] = [summary(kc) for kc in classes if kc != dom]
    roof["timing_region"] = (f"{nprof} further steps with per-launch HIP events (plain launches), "
                             f"{secs_prof / nprof * 1e3:.1f} ms/step")
    window = None
    if args.window_steps > 0:
        b0 = sim.schur_info()[2]
        wits = 0
        device_sync()
        t4 = time.perf_counter()
        for _ in range(args.window_steps):
            wits += sim.advance(allow_noconv=True).ksp_iter
        device_sync()
        t5 = time.perf_counter()
        window = {"steps": args.window_steps, "after_step": args.warmup + args.steps + nprof,
                  "ms_per_step": round((t5 - t4) / args.window_steps * 1e3, 3),
                  "Mgridpt_steps_per_s": round(gridpts(sim) * args.window_steps / (t5 - t4) / 1e6, 3),
                  "gmres_its_per_step": round(wits / args.window_steps, 1),
                  "face_set_rebuilds": sim.schur_info()[2] - b0}
    form, nfaces, builds = sim.schur_info()
    res = {"sim": sim, "desc": desc, "secs": secs, "its": its, "fsi": fsi, "roofline": roof, "window": window,
           "tiles": sim.dst_tiles(),
           "schur": {"form": {1: "matrix-free", 2: "assembled", 3: "face-space"}[form], "faces": nfaces,
                     "green_builds": builds}}
    return res


def run_batched(args):
    """Batched form against the same simulations stepped one after another.

    Workload: ellipse_batch_cases(count) (rigid ellipses, num_stat, grids of
    256..512 cells, 3..5 levels, 100..400 points, uneven).  Batched: one
    ib_batch_advance per step (one graph replay covering every simulation).
    Serial plain: the same simulations created as standalone handles
    (ib_create, each replaying its own captured step graphs as a plain step
    does) and advanced one after another with ib_advance, the plain GPU step
    run serially.  Both are whole-batch rates: sum of m*n*mgridlev over the
    simulations per step."""
    from ib_parallel_amd import GpuBatch, GpuSim, decks
    cases = decks.ellipse_batch_cases(count=args.batch, launch_mode=args.batch_launch_mode)
    t0 = time.perf_counter()
    batch = GpuBatch(cases)
    t_create = time.perf_counter() - t0
    pts = sum(gridpts(s) for s in batch.sims)
    batch.advance()
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        batch.advance()
    device_sync()
    tb = time.perf_counter() - t0
    mode = batch.mode
    tiles = batch.dst_tiles()
    batch.close()
    plain = [GpuSim(p, b) for p, b in decks.ellipse_batch_cases(count=args.batch)]
    for s in plain:
        s.advance()
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        for s in plain:
            s.advance()
    device_sync()
    ts = time.perf_counter() - t0
    for s in plain:
        s.close()
    vb = pts * args.batch_steps / tb / 1e6
    vs = pts * args.batch_steps / ts / 1e6
    res = {
        "workload": (f"{args.batch} rigid ellipses (AR U[1,4], AoA U[0,20] deg, Re U[100,400]), grids 256..512 "
                     "per side, 3..5 levels, 100..400 points, num_stat, PCG64 seed 20240611"),
        "instances": args.batch, "mode": {1: "fused", 2: "concurrent"}[mode],
        "gridpts_per_step": pts, "steps": args.batch_steps,
        "value": round(vb, 3), "unit": "Mgridpt-steps/s", "ms_per_step": round(tb / args.batch_steps * 1e3, 3),
        "serial_plain_value": round(vs, 3),
        "serial_plain_ms_per_step": round(ts / args.batch_steps * 1e3, 3),
        "speedup_vs_serial_plain": round(vb / vs, 2), "create_s": round(t_create, 2),
        "serial_plain": "standalone handles (ib_create), each replaying its captured step graphs",
        "fused_dst_tile_rules": dict(zip(("row", "column"), tiles)),
    }
    return res


def run_c4_batch(args):
    """The two-body system as a batch: args.c4_batch instances of the
    airfoil + flap at 1024x512x5 with the flap stiffness spread over
    [0.02, 0.1] (the concurrent form: each instance's plain step from its
    own host thread, the cooperative GMRES solves of all instances in one
    launch, each on its own XCD) against the same instances as standalone
    handles stepped one after another; then the batch adjoint."""
    import numpy as np
    from ib_parallel_amd import GpuBatch, GpuSim, decks
    nb, n = args.c4_batch, args.c4_batch_steps
    cases = [decks.airfoil_flap_case(k_theta=float(k)) for k in np.linspace(0.02, 0.1, nb)]
    t0 = time.perf_counter()
    batch = GpuBatch(cases)
    t_create = time.perf_counter() - t0
    pts = sum(gridpts(s) for s in batch.sims)
    for _ in range(2):
        batch.advance()
    device_sync()
    t0 = time.perf_counter()
    for _ in range(n):
        batch.advance()
    device_sync()
    tb = time.perf_counter() - t0
    forms = [s.gmres_info() for s in batch.sims]
    adj = None
    if args.adjoint_batch_steps > 0:
        # each instance with the gradient of its time-mean drag (airfoil +
        # flap) with respect to the flap stiffness: ib_batch_adjoint of
        # the same steps (forward pass and reverse sweep, the run taped),
        # after an untimed set-up call (dG/dRe per instance)
        na = args.adjoint_batch_steps
        batch.adjoint(1, [1.0] * (2 * nb), [0.0] * (2 * nb))
        device_sync()
        t0 = time.perf_counter()
        for _ in range(na):
            batch.advance()
        device_sync()
        tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        J, _, dk, _ = batch.adjoint(na, [1.0] * (2 * nb), [0.0] * (2 * nb))
        device_sync()
        ta = time.perf_counter() - t0
        adj = {"steps": na, "forward_ms_per_batch_step": round(tf / na * 1e3, 2),
               "adjoint_ms_per_batch_step": round(ta / na * 1e3, 2), "cost_ratio": round(ta / tf, 3),
               "objective": "mean over the steps of the airfoil + flap drag",
               "dJ_dK": [float(dk[2 * i + 1]) for i in range(nb)]}
    batch.close()
    plain = [GpuSim(p, b) for p, b in cases]
    for _ in range(2):
        for s in plain:
            s.advance()
    device_sync()
    t0 = time.perf_counter()
    for _ in range(n):
        for s in plain:
            s.advance()
    device_sync()
    ts = time.perf_counter() - t0
    for s in plain:
        s.close()
    return {"workload": f"{nb} x airfoil + torsional flap 1024x512x5, flap K_theta in [0.02, 0.1], tol 1e-10",
            "instances": nb, "steps": n, "create_s": round(t_create, 1),
            "value": round(pts * n / tb / 1e6, 3), "unit": "Mgridpt-steps/s",
            "ms_per_batch_step": round(tb / n * 1e3, 2),
            "serial_plain_value": round(pts * n / ts / 1e6, 3),
            "serial_plain_ms_per_step": round(ts / n * 1e3, 2),
            "speedup_vs_serial_plain": round(ts / tb, 2),
            "solves": "each instance's cooperative GMRES solves handed to one fused launch, one XCD per instance",
            "gmres": [{"form": f[0], "workgroups": f[1], "exchange": f[2]} for f in forms],
            "adjoint": adj}


def run_flexible(args):
    """The strongly coupled flexible plate as specified: a beam (family 3) of
    50 Hermite elements on 101 points, mass ratio 0.2, K_B = 0.1, pitch 10
    deg, Re 200, 400x400 cells, 4 levels, dt 1e-3, tolerance 1e-10; plain
    steps after a warm-up."""
    from ib_parallel_amd import GpuSim, decks
    p, b = decks.beam_plate_case()
    sim = GpuSim(p, b)
    for _ in range(5):
        sim.advance()
    device_sync()
    n = args.flex_steps
    its = fsi = 0
    t0 = time.perf_counter()
    for _ in range(n):
        info = sim.advance()
        its += info.ksp_iter
        fsi += info.fsi_iter
    device_sync()
    secs = time.perf_counter() - t0
    st = sim.beam(0)
    out = {"workload": "flexible plate (Euler-Bernoulli beam, 50 elements, 101 points, mass ratio 0.2, K_B 0.1, "
                       "pitch 10 deg) FSI 400x400x4, Re=200, dt=1e-3, tol=1e-10",
           "steps": n, "warmup": 5, "ms_per_step": round(secs / n * 1e3, 3),
           "value": round(gridpts(sim) * n / secs / 1e6, 3), "unit": "Mgridpt-steps/s",
           "gmres_its_per_step": round(its / n, 1), "fsi_its_per_step": round(fsi / n, 2),
           "max_abs_deflection": float(abs(st["u"][0::2]).max())}
    sim.close()
    return out


def _time_adjoint(sim, n, cx, cy, ckpts, **grads):
    """Forward wall time of n plain steps, then the wall time of n adjoint
    steps per segment length in ckpts (None: the whole run taped)."""
    device_sync()
    t0 = time.perf_counter()
    for _ in range(n):
        sim.advance()
    device_sync()
    tf = time.perf_counter() - t0
    # set-up outside the timed regions (tables, dA or dG/dRe for the current
    # face set, tape buffers)
    sim.adjoint(n, cx, cy, ckpt_every=ckpts[0], **grads)
    device_sync()
    ta = []
    rebuilds = []
    for K in ckpts:
        g0 = sim.schur_info()[2]
        t0 = time.perf_counter()
        out = sim.adjoint(n, cx, cy, ckpt_every=K, **grads)
        device_sync()
        ta.append(time.perf_counter() - t0)
        rebuilds.append(sim.schur_info()[2] - g0)
    _time_adjoint.rebuilds = rebuilds
    return tf, ta, out


def run_adjoint(args):
    """Discrete adjoint (ib_adjoint_run) against the plain forward steps.

    rigid: the cylinder (200x200x3, 157 points, Re = 100, dt = 4e-3) on the
    num_stat path, gradients of the time-mean drag + 0.5 lift with respect to
    Re and the initial vorticity; the whole run taped, and a checkpoint every
    10 steps.
    fsi: the hinged plate at 400x400x3 (torsional body, K = 0.1, pitch 10 deg,
    Re = 200, dt = 1e-3, GMRES and FSI tolerance 1e-10, face-space form),
    gradients of the time-mean lift with respect to stiffness, pitch, Re and
    the initial vorticity; a checkpoint every 20 steps, and the whole run
    taped.
    Forward: ib_advance of N steps.  Adjoint: ib_adjoint_run of N steps (its
    forward pass, the recomputation from checkpoints and the reverse sweep).
    cost_ratio = adjoint wall / forward wall."""
    from ib_parallel_amd import GpuSim, decks
    res = {}
    n = args.adjoint_steps
    if n > 0:
        # --adjoint-launch-mode 1: plain launches instead of the step and
        # reverse-step graphs (for kernel traces of the reverse sweep)
        p, b = decks.cylinder_case(num_stat=1, launch_mode=args.adjoint_launch_mode)
        sim = GpuSim(p, b)
        for _ in range(3):
            sim.advance()
        tf, (ta, tc), _ = _time_adjoint(sim, n, [1.0], [0.5], [None, 10])
        res["rigid"] = {"workload": "rigid cylinder 200x200x3, 157 pts, Re=100, dt=4e-3, num_stat; d/dRe and "
                                    "d/domega0 of mean(drag + 0.5 lift)",
                        "steps": n, "forward_ms_per_step": round(tf / n * 1e3, 3),
                        "adjoint_ms_per_step": round(ta / n * 1e3, 3), "cost_ratio": round(ta / tf, 3),
                        "adjoint_ckpt10_ms_per_step": round(tc / n * 1e3, 3), "cost_ratio_ckpt10": round(tc / tf, 3)}
        sim.close()
    if args.adjoint_fsi_steps > 0:
        p, b = decks.plate_case(mgridlev=3)
        sim = GpuSim(p, b)
        for _ in range(2):
            sim.advance()
        n = args.adjoint_fsi_steps
        tf, (tc, ta), out = _time_adjoint(sim, n, [0.0], [1.0], [20, None], body_grad=True)
        res["fsi"] = {"workload": "hinged plate 400x400x3, 101 pts, torsional K=0.1, pitch 10 deg, Re=200, dt=1e-3, "
                                  "tol 1e-10, face-space form; d/dK, d/dpitch, d/dRe, d/domega0 of mean(lift)",
                      "steps": n, "forward_ms_per_step": round(tf / n * 1e3, 3),
                      "adjoint_ckpt20_ms_per_step": round(tc / n * 1e3, 3), "cost_ratio_ckpt20": round(tc / tf, 3),
                      "adjoint_ms_per_step": round(ta / n * 1e3, 3), "cost_ratio": round(ta / tf, 3),
                      "dJ_dK": out[3][0], "dJ_dgamma0": out[4][0], "dJ_dRe": out[1],
                      "face_set_rebuilds": {"ckpt20": _time_adjoint.rebuilds[0], "taped": _time_adjoint.rebuilds[1]},
                      "note": ("a face-set rebuild (a tap leaving the set of faces G covers) also rebuilds dG/dRe "
                               "within the same adjoint call")}
        sim.close()
    return res


def run_adjoint_batch(args):
    """The two-body system as a batch with adjoints: elliptic airfoil (rigid)
    + torsional flap at 1024x512x5 (decks.airfoil_flap_case), args.adjoint_batch
    instances with the flap stiffness spread over [0.02, 0.1], each with the
    gradient of its time-mean airfoil + flap drag with respect to the
    stiffness (ib_batch_adjoint: one host thread and stream per instance).
    Forward: ib_batch_advance of N steps; adjoint: ib_batch_adjoint of N
    steps (forward pass, reverse sweep), the whole run taped."""
    import numpy as np
    from ib_parallel_amd import GpuBatch, decks
    nb = args.adjoint_batch
    ks = np.linspace(0.02, 0.1, nb)
    cases = [decks.airfoil_flap_case(k_theta=float(k)) for k in ks]
    t0 = time.perf_counter()
    batch = GpuBatch(cases)
    t_create = time.perf_counter() - t0
    n = args.adjoint_batch_steps
    batch.advance()
    device_sync()
    t0 = time.perf_counter()
    batch.adjoint(1, [1.0] * (2 * nb), [0.0] * (2 * nb))  # adjoint set-up (dG/dRe per instance)
    device_sync()
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(n):
        batch.advance()
    device_sync()
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    J, dre, dk, dg = batch.adjoint(n, [1.0] * (2 * nb), [0.0] * (2 * nb))
    device_sync()
    ta = time.perf_counter() - t0
    pts = sum(gridpts(s) for s in batch.sims)
    res = {"workload": f"{nb} x (elliptic airfoil 300 pts rigid + torsional flap 100 pts) 1024x512x5, Re=1000, "
                       "dt=5e-4, tol 1e-10, flap K_theta in [0.02, 0.1]; d/dK of mean(airfoil + flap drag)",
           "instances": nb, "steps": n, "create_s": round(t_create, 1), "adjoint_setup_s": round(t_setup, 1),
           "forward_ms_per_step": round(tf / n * 1e3, 2),
           "forward_Mgridpt_steps_per_s": round(pts * n / tf / 1e6, 2),
           "adjoint_ms_per_step": round(ta / n * 1e3, 2), "cost_ratio": round(ta / tf, 3),
           "k_theta": [round(float(k), 4) for k in ks], "dJ_dK": [float(dk[2 * i + 1]) for i in range(nb)],
           "mean_drag": [float(j) for j in J]}
    batch.close()
    return res


def copy_calibration(mb):
    """A device-to-device copy of exactly mb MiB (a plain copy kernel of known
    size for calibrating FETCH_SIZE / WRITE_SIZE in counter runs)."""
    import ctypes
    hip = _hip()
    n = int(mb * 1024 * 1024)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3):  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    device_sync()
    hip.hipFree(a)
    hip.hipFree(b)


def run_dst_compare(args):
    """The two DST forms on one grid (one level, no bodies): a handle with
    the Stockham form and one with the FP64 matrix-core form (IB_DST_FORM),
    then --dst-reps spectral solves ctci (row pass, scaled column pass, row
    pass) through each, host copies included (the kernels are what a trace
    measures).  A third handle reports the autotuner's choice and its
    per-field times.  With --pmc-copy-mb the calibration copy separates the
    create-time timing from the measured passes."""
    import numpy as np
    from ib_parallel_amd import GpuSim, abi
    m, n = (int(v) for v in args.dst_compare.lower().split("x"))
    p = abi.make_params(m, n, 1, 0.01, 100.0, 1.0, 0.5, 0.5)
    old = os.environ.get("IB_DST_FORM")
    sims = {}
    try:
        for form in ("auto", "fft", "dense"):
            # a fixed configuration keeps kernel names alike across profiled processes
            os.environ["IB_DST_FORM"] = f"dense:{args.dst_dense_cfg}" if form == "dense" and args.dst_dense_cfg >= 0 \
                else form
            sims[form] = GpuSim(p, [])
    finally:
        if old is None:
            os.environ.pop("IB_DST_FORM", None)
        else:
            os.environ["IB_DST_FORM"] = old
    forms, us = sims["auto"].dst_form()
    dforms = sims["dense"].dst_form()[0]
    sims.pop("auto").close()
    w = np.random.default_rng(0).standard_normal(sims["fft"].P)
    outs = {f: s.ctci(w) for f, s in sims.items()}  # warm-up
    if args.pmc_copy_mb > 0:
        copy_calibration(args.pmc_copy_mb)
    wall = {}
    for f, s in sims.items():
        device_sync()
        t0 = time.perf_counter()
        for _ in range(args.dst_reps):
            outs[f] = s.ctci(w)
        wall[f] = (time.perf_counter() - t0) / args.dst_reps * 1e3
    d = outs["dense"] - outs["fft"]
    names = ["row step", "column step", "row batched", "column batched"]
    res = {"grid": f"{m}x{n}", "transform_lengths": [m - 1, n - 1], "reps": args.dst_reps,
           "auto_forms": {k: ("stockham" if f == 0 else f"matrix-core cfg {f - 1}") for k, f in zip(names, forms)},
           "us_per_field": {k: {"stockham": round(float(us[2 * i]), 3),
                                "matrix_core": None if us[2 * i + 1] < 0 else round(float(us[2 * i + 1]), 3)}
                            for i, k in enumerate(names)},
           "dense_handle_cfgs": list(dforms),
           "ctci_ms_with_host_copies": {k: round(v, 4) for k, v in wall.items()},
           "dense_vs_stockham_rel": float(np.linalg.norm(d) / np.linalg.norm(outs["fft"]))}
    for s in sims.values():
        s.close()
    return res


def traffic_for(kernel_label):
    """Calibrated HBM bytes per launch of the kernel from the committed
    counter summary (profiles/pmc_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        data = json.load(fh)
    base = kernel_label.split(" ")[0]
    best = None
    for k, v in data.get("kernels", {}).items():
        if k.startswith(base) and (best is None or v["dispatches"] > best[1]["dispatches"]):
            best = (k, v)
    if best is None:
        return None, None
    return round(best[1]["bytes_per_launch"]), best[0]


def _host_cpus():
    """(usable cores, CPU model) of this host."""
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return ncpu, model


def run_cpu(args, gpu_its_per_step, gpu_fsi_per_step):
    """The oracle on the host cores: on all cores and on one core, with the
    warm-up step excluded.

    Headline workload: one oracle step there takes minutes (matrix-free
    GMRES, ~0.07-0.3 s per Krylov iteration on 8-16 cores), so the sample is
    bounded: steps whose GMRES solves stop after K = 4 and K = 40 iterations
    (FSI loop capped at the GPU's 3 iterations; about 10 s on 16 cores).  A
    step then costs
    T(K) = F + s (K + 1) t (s solves, each one residual product plus K
    iterations), which gives the fixed cost F (predictor, vort2flux,
    corrector, FSI right-hand sides and body updates) and the cost t of one
    Krylov iteration; the estimate for a step is F + (its + s) t with the
    GPU's iteration and FSI counts of the timed steps.  The oracle's
    matrix-free GMRES needs more iterations than that (2430 against about
    1400 on the first step), so the estimate favours the CPU.  One core: the
    same estimate scaled by the ratio of one matrix-free Schur product on
    one core and on all cores.
    Other workloads: whole oracle steps from the same cold start."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import binding as ob
    ncpu, model = _host_cpus()
    threads = max(1, min(16, ncpu))  # the box's CPU share is 16 cores
    ob.set_threads(threads)
    p, bodies, _ = make_case(args.workload)
    host = {"nproc": ncpu, "cpu_model": model, "threads": threads}
    if args.workload != "airfoil_flap":
        sim = ob.OracleSim(p, bodies)
        steps = its = 0
        t0 = time.perf_counter()
        while True:
            info = sim.advance(allow_noconv=True)
            steps += 1
            its += info.ksp_iter
            if time.perf_counter() - t0 >= args.cpu_seconds or steps >= args.steps + args.warmup:
                break
        secs = time.perf_counter() - t0
        value = gridpts(sim) * steps / secs / 1e6
        return {
            "value": round(value, 4), "unit": "Mgridpt-steps/s", "cores": threads, "kind": "port", **host,
            "sample": (f"oracle (C++ FP64 restatement, OpenMP {threads} threads) advancing steps 1..{steps} of the "
                       f"same workload from the same cold start in {secs:.1f} s; {its} GMRES iterations "
                       f"({its / steps:.1f}/step vs {gpu_its_per_step:.1f}/step in the GPU timed steps)"),
        }
    import numpy as np
    from ib_parallel_amd import decks
    T = {}
    fsi = {}
    t_begin = time.perf_counter()
    lo, hi = 4, 40
    for K in (lo, hi):
        pk, bk = decks.airfoil_flap_case(gmres_max_it=K, fsi_max_count=2)
        sim = ob.OracleSim(pk, bk)
        if K == lo:
            sim.advance(allow_noconv=True)  # warm-up step, excluded
        t0 = time.perf_counter()
        info = sim.advance(allow_noconv=True)
        T[K] = time.perf_counter() - t0
        fsi[K] = info.fsi_iter
        if K == hi:
            x = np.random.default_rng(0).standard_normal(sim.nf)

            def best_of(n):
                best = float("inf")
                for _ in range(n):
                    t0 = time.perf_counter()
                    sim.schur_matvec(x)
                    best = min(best, time.perf_counter() - t0)
                return best
            t_mv_all = best_of(3)
            ob.set_threads(1)
            t_mv_one = best_of(2)
            ob.set_threads(threads)
        npts = gridpts(sim)
        sim.close()
    s_solves = 0.5 * (fsi[lo] + fsi[hi])
    t_it = max(T[hi] - T[lo], 0.0) / (s_solves * (hi - lo))
    F = T[lo] - (lo + 1) * s_solves * t_it
    est = F + (gpu_its_per_step + gpu_fsi_per_step) * t_it
    scale1 = t_mv_one / t_mv_all
    value = npts / est / 1e6
    return {
        "value": round(value, 6), "unit": "Mgridpt-steps/s", "cores": threads, "kind": "port", **host,
        "est_s_per_step": round(est, 2), "fixed_s_per_step": round(F, 3), "s_per_krylov_iteration": round(t_it, 4),
        "one_core": {"value": round(value / scale1, 6), "est_s_per_step": round(est * scale1, 1),
                     "schur_product_s": round(t_mv_one, 3), "schur_product_s_all_cores": round(t_mv_all, 3)},
        "sample": (f"oracle (C++ FP64 restatement, matrix-free GMRES as the reference, OpenMP {threads} threads) "
                   f"on the same workload: one warm-up step excluded, then one step each with the GMRES solves "
                   f"capped at {lo} and {hi} iterations ({s_solves:g} FSI iterations): {T[lo]:.2f} s and "
                   f"{T[hi]:.2f} s, "
                   f"so {F:.2f} s fixed + {t_it:.3f} s per Krylov iteration, extrapolated to the GPU's "
                   f"{gpu_its_per_step:.1f} iterations and {gpu_fsi_per_step:.1f} solves per step; one core "
                   f"scaled by one Schur product (best of 2: {t_mv_one:.2f} s on 1 core; best of 3: {t_mv_all:.2f} s "
                   f"on {threads}); "
                   f"{time.perf_counter() - t_begin:.0f} s of CPU sampling"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="airfoil_flap", choices=["plate", "cylinder", "airfoil_flap"])
    ap.add_argument("--window-steps", type=int, default=100,
                    help="steps of the steady-state window after the timed region (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=64, help="simulations in the batched measurement (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-only", action="store_true", help="only the batched measurement (one JSON object)")
    ap.add_argument("--batch-launch-mode", type=int, default=0,
                    help="1: the batched step's launches go to the stream directly instead of a graph replay")
    ap.add_argument("--pmc-copy-mb", type=float, default=0.0,
                    help="run a device-to-device copy of this size first (calibrates counter runs)")
    ap.add_argument("--adjoint-steps", type=int, default=40, help="steps of the adjoint measurement (0: skip)")
    ap.add_argument("--adjoint-fsi-steps", type=int, default=40,
                    help="steps of the moving-body adjoint measurement (0: skip)")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint measurement (one JSON object)")
    ap.add_argument("--adjoint-batch", type=int, default=0,
                    help="instances of the airfoil + flap batch adjoint measurement (0: skip; 8 in the workload)")
    ap.add_argument("--adjoint-batch-steps", type=int, default=4)
    ap.add_argument("--c4-batch", type=int, default=8,
                    help="instances of the airfoil + flap batch measurement (0: skip)")
    ap.add_argument("--c4-batch-steps", type=int, default=5)
    ap.add_argument("--c4-batch-only", action="store_true", help="only the airfoil + flap batch measurement")
    ap.add_argument("--flex-steps", type=int, default=20,
                    help="timed steps of the flexible-plate (beam) workload (0: skip)")
    ap.add_argument("--dst-compare", default=None, metavar="MxN",
                    help="only the comparison of the two DST forms on an M x N grid (one JSON object)")
    ap.add_argument("--adjoint-launch-mode", type=int, default=0, choices=[0, 1],
                    help="launch_mode of the rigid adjoint workload (1: no HIP graphs)")
    ap.add_argument("--dst-reps", type=int, default=50)
    ap.add_argument("--dst-dense-cfg", type=int, default=-1,
                    help="tile configuration of the matrix-core handle in --dst-compare (-1: timed)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the library has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    if args.adjoint_only:
        out = {"adjoint": run_adjoint(args)} if args.adjoint_steps > 0 or args.adjoint_fsi_steps > 0 else {}
        if args.adjoint_batch > 0:
            out["adjoint_batch"] = run_adjoint_batch(args)
        print(json.dumps(out), flush=True)
        return
    if args.c4_batch_only:
        print(json.dumps({"c4_batch": run_c4_batch(args)}), flush=True)
        return
    if args.dst_compare:
        print(json.dumps({"dst_compare": run_dst_compare(args)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batched": run_batched(args)}), flush=True)
        return
    g = run_gpu(args)
    sim = g["sim"]
    secs = g["secs"]
    value = gridpts(sim) * args.steps / secs / 1e6
    out = {
        "metric": "Mgridpt-steps/s FP64 at 1e-10 rel err; batched & adjoint vs plain GPU",
        "value": round(value, 4),
        "unit": "Mgridpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": g["desc"], "m": sim.m, "n": sim.n, "mgridlev": sim.L, "nb": sim.nf // 2,
                   "gmres_its_per_step": round(g["its"] / args.steps, 2),
                   "fsi_its_per_step": round(g["fsi"] / args.steps, 2), "form": "plain",
                   "schur_operator": g["schur"],
                   "dst_tiles": {"row_pairs_per_workgroup": g["tiles"][0], "column_pairs_per_workgroup": g["tiles"][1],
                                 "chosen_by": {0: "fixed rule", 1: "timed at create", 2: "IB_DST_TILES"}[g["tiles"][2]]}},
        "roofline": g["roofline"],
    }
    if g["window"] is not None:
        out["window"] = g["window"]
    if args.batch > 0:
        out["batched"] = run_batched(args)
    if args.adjoint_steps > 0 or args.adjoint_fsi_steps > 0:
        out["adjoint"] = run_adjoint(args)
    if args.c4_batch > 0:
        out["c4_batch"] = run_c4_batch(args)
    if args.flex_steps > 0:
        out["flexible_plate"] = run_flexible(args)
    if not args.no_cpu:
        cpu = run_cpu(args, g["its"] / args.steps, g["fsi"] / args.steps)
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu"] = round(value / cpu["value"], 2) if cpu["value"] > 0 else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
