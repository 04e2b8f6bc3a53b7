"""Benchmark of the SIMPLEC linear-algebra hot path on one MI355X.

Workload: the 1026x514 pressure-correction (p') system on the wall-resolved
ogive grid (near-wall aspect ratio 1e4, seeded Gaussian right-hand side),
solved by MSI from a zero guess to a 1e-10 residual drop
(||r||_2 < 1e-10 ||r0||_2, the stopping rule of fb2d9).  One step = one full
solve: factorization (lu2d9) plus the iterations (fb2d9), with the
coefficients and right-hand side already resident in HBM.

    python bench.py [--steps K] [--warmup W] [--mode best|fp64|fp32_ir]

Modes: fp64 = plain FP64 MSI (bit-identical to the reference); fp32_ir =
FP32 factor and sweeps inside FP64 iterative refinement (FP64 residuals and
stopping rule).  "best" runs both; value is the FP32-refined solves/s when
its solution is within 1e-8 relative of the FP64 one, else the FP64 figure,
and speedup_vs_plain_fp64 compares the two.  The batch entry times 256
independent p' systems of uneven size (64s x 32s, s log-uniform on [1, 8])
solved together to the same drop (autotuned placement, graph replays);
batch_1026x514 times 48 systems of the headline size (its matrix, their own
seeded right-hand sides) solved together, with their throughput against
the single plain-FP64 solve; msi2d5_1026x514 times the 5-diagonal p' system
of the same grid (the reference's own p' form) on the dedicated 5-diagonal
kernels and, for comparison, on the 9-diagonal ones through the zero-corner
embedding; pl_loop_258x130 times the device-resident pressure-correction
loop of one SIMPLEC iteration of the internal-flow case (258x130, 20 itp
iterations) against the same loop of reference routines on one core;
tdma_258x130 times the device TDMA of the u and v systems with the lines
solved by the Thomas recurrence and by parallel cyclic reduction.

Prints one JSON line.  gbps = algorithmic bytes of the solve over its time
(FP64: 8 N (18 + 11 + 24 K); refinement: 4 N (18 + 24 K_inner) +
8 N (11 (outer + 1) + 3 outer)); roofline = the dominant kernel's
algorithmic bytes per launch over its average launch duration (HIP events on
the solver's stream around each launch, in a repeat of the timed steps: the
event records would lengthen the timed steps themselves);
cpu_baseline = the reference MSI (compiled Fortran, one core) timed on a
bounded sample of the same solve on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mach-2d_amd"))

import numpy as np  # noqa: E402

METRIC = ("p' solves/s and HBM GB/s to 1e-10 residual drop; learned+FP32 vs plain FP64")
HBM_PEAK_GBS = 8000.0
GAMMA = 1e-10
NITM_CAP = 20000

# algorithmic bytes per point of one launch, in units of the value size
BYTES_PER_POINT = {
    "factor": 18,    # read a (9), write dl (5) and du (4)
    "sweeps": 13,    # forward: read r, dl (5), write w; backward: read w, du (4), write z
    "backward": 6,   # staged sweeps only: the backward sweep alone
    "residual": 14,  # one launch: read z, var, a (9), rhs; write var, r
}


# HBM bytes per launch at 1026x514, measured with rocprofv3 (FETCH_SIZE and
# WRITE_SIZE in separate runs of this script, scaled by 2.2488 and 1.1250:
# the factors a conversion kernel of known size reads;
# profiles/msi_1026x514_hbm_traffic.md).  "residual" is the update and the
# residual of one launch.
MEASURED_TRAFFIC = {
    "sweeps": 84.95e6, "residual": 67.13e6, "factor": 111.72e6,
    "sweeps32": 39.25e6, "residual32": 33.75e6, "factor32": 55.94e6,
}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_system():
    from mach2d import problems as P
    c = P.CFG3
    t = time.time()
    s = P.pcorr_system(c["nx"], c["ny"], seed=c["seed"], aspect=c["aspect"], rhs=c["rhs"])
    log("system %dx%d (q=%.5f, aspect %.3g) built in %.1fs" % (
        s.nx, s.ny, s.q, c["aspect"], time.time() - t))
    return s


def gpu_run(s, mode, steps, warmup):
    import torch

    from mach2d import solvers as S
    dev = S.DeviceMSI(s.nx, s.ny, precision=mode)
    ta = torch.tensor(s.a, device="cuda")
    tr = torch.tensor(s.rhs, device="cuda")
    tv = torch.zeros(s.N, dtype=torch.float64, device="cuda")

    def one():
        tv.zero_()
        dev.factor(ta)
        return dev.solve(ta, tr, tv, GAMMA, NITM_CAP)

    for w in range(warmup):
        st = one()
        log("warmup %d: status %d nit %d (%.2f ms)" % (w, st.status, st.nit, st.device_ms))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sts = [one() for _ in range(steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # per-kernel durations: the same steps again with HIP events around every
    # launch (on the solver's stream); kept out of the timed steps above,
    # whose wall time the event records would lengthen
    dev.enable_kernel_timing(True)
    for _ in range(steps):
        one()
    torch.cuda.synchronize()
    kt = dev.kernel_timing()
    dev.enable_kernel_timing(False)
    st = sts[-1]
    ok = all(x.status == 0 and x.rn < x.gammarp for x in sts)
    log("%s: %d steps in %.3fs, nit %d, status %d, r0 %.3e rn %.3e" % (
        mode, steps, dt, st.nit, st.status, st.r0, st.rn))
    x = tv.cpu().numpy()
    dev.close()
    return dict(dt=dt, st=st, ok=ok, kt=kt, x=x)


def learned_run(s, steps, warmup, x64):
    """FP32-refined MSI from the U-Net guess (inference, energy-minimising
    amplitude) to the zero guess's threshold; timed including the
    inference."""
    import torch

    from mach2d import learned as Lr
    from mach2d import solvers as S
    from mach2d import unet as U
    path = Lr.default_weights()
    if not os.path.exists(path):
        log("learned: no weights file %s" % path)
        return None
    net = U.UNet()
    net.load(path)
    dev = S.DeviceMSI(s.nx, s.ny, precision="fp32_ir")
    ta = torch.tensor(s.a, device="cuda")
    tr = torch.tensor(s.rhs, device="cuda")
    tv = torch.zeros(s.N, dtype=torch.float64, device="cuda")

    def one():
        tv.zero_()
        dev.factor(ta)
        return Lr.guessed_solve(dev, net, ta, tr, tv, s.nx, s.ny, GAMMA, NITM_CAP)

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = [one() for _ in range(steps)]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st, used, rg, rb = res[-1]
    ok = all(r[0].status == 0 and r[0].rn < GAMMA * r[3] for r in res)
    x = tv.cpu().numpy()
    rel = float(np.linalg.norm(x - x64) / np.linalg.norm(x64)) if x64 is not None else None
    # inference alone
    t1 = time.perf_counter()
    for _ in range(3):
        net.guess(ta, tr, s.nx, s.ny, tv)
    torch.cuda.synchronize()
    t_inf = (time.perf_counter() - t1) / 3
    dev.close()
    e = {"solves_per_s": round(1.0 / dt, 4), "ms_per_solve": round(dt * 1e3, 3), "nit": st.nit,
         "outer": st.outer, "guess_used": used, "r_guess_over_r_zero": rg / rb,
         "inference_ms": round(t_inf * 1e3, 3),
         "inference_tflops": round(43.5e-3 / t_inf, 2),
         "correct": bool(ok and (rel is None or rel <= 1e-8))}
    if rel is not None:
        e["rel_error_vs_fp64"] = rel
    log("learned fp32_ir: %s" % e)
    return e


def msi5_run(steps, warmup):
    """The 5-diagonal p' system (slots S, W, P, E, N) of the same grid and
    right-hand-side kind: dedicated 5-diagonal handles (msi2d5: 5 + 4 + 3
    arrays per iteration) in FP64 and FP32 refinement, and the FP64
    9-diagonal handle on its zero-corner embedding (same iterates)."""
    import torch

    from mach2d import problems as P
    from mach2d import solvers as S
    c = P.CFG3
    s = P.pcorr_system5(c["nx"], c["ny"], seed=c["seed"], aspect=c["aspect"], rhs=c["rhs"])
    out, xs = {}, {}
    for name, prec, nd in (("fp64", "fp64", 5), ("fp32_ir", "fp32_ir", 5),
                           ("fp64_embedded_9d", "fp64", 9)):
        a = s.a if nd == 5 else S.embed5(s.a, s.N)
        dev = S.DeviceMSI(s.nx, s.ny, precision=prec, ndiag=nd)
        ta = torch.tensor(a, device="cuda")
        tr = torch.tensor(s.rhs, device="cuda")
        tv = torch.zeros(s.N, dtype=torch.float64, device="cuda")

        def one():
            tv.zero_()
            dev.factor(ta)
            return dev.solve(ta, tr, tv, GAMMA, NITM_CAP)

        for _ in range(warmup):
            one()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sts = [one() for _ in range(steps)]
        torch.cuda.synchronize()
        per = (time.perf_counter() - t0) / steps
        st = sts[-1]
        xs[name] = tv.cpu().numpy()
        if prec == "fp64":
            nbytes = 8 * s.N * (12 + 7 + 18 * st.nit) if nd == 5 else \
                8 * s.N * (18 + 11 + 24 * st.nit)
        else:
            nbytes = 4 * s.N * (12 + 18 * st.nit) + 8 * s.N * (7 * (st.outer + 1) + 3 * st.outer)
        e = {"solves_per_s": round(1.0 / per, 4), "ms_per_solve": round(per * 1e3, 3),
             "nit": st.nit, "outer": st.outer, "gbps": round(nbytes / per / 1e9, 2),
             "correct": bool(all(x.status == 0 and x.rn < x.gammarp for x in sts))}
        out[name] = e
        dev.close()
    out["fp32_ir"]["rel_error_vs_fp64"] = float(
        np.linalg.norm(xs["fp32_ir"] - xs["fp64"]) / np.linalg.norm(xs["fp64"]))
    out["fp32_ir"]["correct"] = bool(out["fp32_ir"]["correct"] and
                                     out["fp32_ir"]["rel_error_vs_fp64"] <= 1e-8)
    out["fp64_embedded_9d"]["same_iterate"] = bool(np.array_equal(xs["fp64"],
                                                                  xs["fp64_embedded_9d"]))
    out["speedup_5d_vs_embedded_9d"] = round(out["fp64"]["solves_per_s"] /
                                             out["fp64_embedded_9d"]["solves_per_s"], 3)
    log("msi2d5 1026x514: %s" % out)
    return out


def pl_loop_run(reps, cpu_itp=3):
    """The internal-flow pressure-correction loop of one SIMPLEC iteration
    (20 itp iterations, p' solves of at most 20 MSI iterations) on a 258x130
    grid, device-resident, against the same loop composed from the reference
    routines (compiled Fortran, one core; `cpu_itp` iterations) when built."""
    from mach2d import workloads as W
    g = W.pl_loop(258, 130, 20, reps)
    out = {"ms_per_itp": g["ms_per_itp"], "itp": g["itp"], "ms": g["ms"]}
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle as O
        import pl_loop_oracle as PL
        ref = O.ref_available()
        f = PL.loop_fields(258, 130, seed=1)
        p_avg = float(np.mean(f["p"]))
        t = time.perf_counter()
        PL.run(O, f, 258, 130, 0, 2e-3, 0.3, p_avg, 0.0, cpu_itp, 20, 1e-2, ref=ref)
        cpu = (time.perf_counter() - t) / cpu_itp * 1e3
        out["cpu_ms_per_itp"] = round(cpu, 3)
        out["cpu_kind"] = "reference" if ref else "port"
        out["speedup_vs_cpu"] = round(cpu / g["ms_per_itp"], 2)
    except Exception as e:  # the checker is optional here
        out["cpu_ms_per_itp"] = None
        log("pl_loop cpu: %s" % e)
    log("pl_loop 258x130: %s" % out)
    return out


def tdma_run(nx=258, ny=130, nitm=5, reps=20):
    """The device TDMA (tdma2d9.f90:38) on seeded nx x ny systems, nitm
    iterations per call, with the lines solved by the Thomas recurrence
    (mode 0, the reference's arithmetic bit for bit) and by parallel cyclic
    reduction (mode 1): ms per call and per iteration for the u and v pair
    (count 2) and a single system, and the largest difference between the
    two modes' solutions relative to the solution's largest magnitude."""
    import torch

    from mach2d import problems as P, solvers, workloads as W
    out = {"nx": nx, "ny": ny, "nitm": nitm}
    try:
        for mode, name in ((0, "thomas"), (1, "pcr")):
            solvers.set_tdma_mode(mode)
            rows = {}
            for count in (2, 1):
                g = W.tdma(nx, ny, nitm, count, reps)
                rows["count%d" % count] = {"ms": g["ms"], "ms_per_iteration": g["ms_per_iteration"]}
            out[name] = rows
        s = P.random_f3_system(nx, ny, seed=0)
        a = torch.tensor(s.a.reshape(-1), device="cuda")
        b = torch.tensor(s.rhs, device="cuda")
        xs = []
        for mode in (0, 1):
            solvers.set_tdma_mode(mode)
            x = torch.zeros(nx * ny, dtype=torch.float64, device="cuda")
            solvers.tdma2d9_device_many(nx, ny, nitm, [a], [b], [x])
            torch.cuda.synchronize()
            xs.append(x)
        out["pcr_max_rel_diff"] = float((xs[1] - xs[0]).abs().max() / xs[0].abs().max())
    finally:
        solvers.set_tdma_mode(0)
    log("tdma %dx%d: %s" % (nx, ny, out))
    return out


# the shipped nozzle settings (parametersNozzleEulerPTC.txt: itmmax, itpmax,
# nitm_u, nitm_p, tolerances, beta, Scarborough h0/mincc/maxcc; constant
# properties Rg, gamma) with chamber conditions of a slow test flow
NOZZLE = dict(modvis=0, coord=1, solver9d=0, itmmax=1, itpmax=20, nitm_u=5, nitm_p=20,
              tol_u=1e-2, tol_p=1e-2, tolm=1e-2, beta=0.0, Rg=287.0, gamma=1.4, P0=1.2e5,
              T0=310.0, Cpref=1004.5, Href=1004.5 * 310.0, Twall=-1.0, h0=0.01, mincc=0.99,
              maxcc=0.9999)


def time_step_run(cpu_steps=5):
    """The full SIMPLEC time step of the internal-flow case (Euler, constant
    properties, the shipped nozzle parameters: itmmax 1, itpmax 20, nitm_u 5,
    nitm_p 20, tolerances 1e-2, Scarborough time step) on the 258x130 nozzle
    from inlet Mach 0.2, 200 time steps from the initial state, each step
    timed, device-resident (m2d_intflow_time_step), u and v by MSI, by
    TDMA and by TDMA with the lines solved by parallel cyclic reduction
    (largest difference of the final fields from the Thomas trajectory
    reported beside it); the first `cpu_steps` steps of the same trajectory composed from
    the reference's routines (compiled Fortran, one core) beside them when
    built."""
    import torch

    from mach2d import problems, simplec, solvers
    c = problems.CFG2
    nx, ny, steps = c["nx"], c["ny"], c["steps"]
    out = {"steps": steps, "mach_in": c["mach_in"], "dt0": c["dt0"]}
    f0 = problems.time_step_fields(nx, ny, 5, NOZZLE, mach_in=c["mach_in"])
    final = {}
    for name, solver9d, tdma_mode in (("msi", 0, 0), ("tdma", 1, 0), ("tdma_pcr", 1, 1)):
        solvers.set_tdma_mode(tdma_mode)
        prm = dict(NOZZLE, solver9d=solver9d)
        dev9 = solvers.DeviceMSI(nx, ny) if solver9d == 0 else None
        dev5 = solvers.DeviceMSI(nx, ny, ndiag=5)

        def fresh():
            d = {k: torch.tensor(np.ascontiguousarray(v).reshape(-1), device="cuda")
                 for k, v in f0.items()}
            torch.cuda.synchronize()   # the copies run on torch's stream
            return d
        # warm-up on the tensors the timed steps use (the p' loop's captured
        # batches are kept per set of field arrays), then the initial state
        # copied back into them
        init, d = fresh(), fresh()
        dt = c["dt0"]
        for _ in range(2):
            dt = simplec.intflow_time_step(dev9, dev5, d, prm, dt)[0]
        for k in d:
            d[k].copy_(init[k])
        torch.cuda.synchronize()
        dt = c["dt0"]
        per, itps, caps, reruns = [], 0, 0, 0
        for _ in range(steps):   # the step returns with its stream idle
            t = time.perf_counter()
            dt, r = simplec.intflow_time_step(dev9, dev5, d, prm, dt)
            per.append(time.perf_counter() - t)
            itps += r["itp"]
            used, rr = simplec.pl_loop_cap_info(dev5)
            caps += used
            reruns += rr
        ok = all(bool(torch.isfinite(d[k]).all()) for k in ("u", "v", "p", "T"))
        out[name] = {"ms_per_step": round(sum(per) / steps * 1e3, 3),
                     "itp_per_step": round(itps / steps, 2),
                     "ms_200_steps": round(sum(per) * 1e3, 1),
                     "ms_first_steps": round(sum(per[:cpu_steps]) * 1e3, 2),
                     "p_solve_iterations_enqueued": round(caps / steps, 2),
                     "p_loop_reruns": reruns,
                     "finite": ok}
        final[name] = d
        for h in (dev9, dev5):
            if h is not None:
                h.close()
    solvers.set_tdma_mode(0)
    # the PCR line solver changes the u, v solutions by rounding only: the
    # trajectories' largest difference after the timed steps, field by field
    out["tdma_pcr"]["max_rel_diff_vs_tdma"] = {
        k: float((final["tdma_pcr"][k] - final["tdma"][k]).abs().max()
                 / final["tdma"][k].abs().max()) for k in ("u", "v", "p", "T")}
    try:
        if cpu_steps < 1:
            raise RuntimeError("CPU leg not asked for")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle as O
        import time_step_oracle as TS
        ref = O.ref_available()
        pf = os.path.join(ROOT, "tests", "golden", "parametersNozzleEulerPTC.txt")
        s, dt, itps = f0, c["dt0"], 0
        t = time.perf_counter()
        for _ in range(cpu_steps):
            s, dt, _r = TS.step(O, s, nx, ny, dict(NOZZLE), dt, ref=ref, paramfile=pf)
            itps += _r["itp"]
        cpu = (time.perf_counter() - t) * 1e3
        out["cpu_ms_first_steps"] = round(cpu, 2)
        out["cpu_steps"] = cpu_steps
        out["cpu_itp_first_steps"] = itps
        out["cpu_kind"] = "reference" if ref else "port"
        out["speedup_msi_vs_cpu_first_steps"] = round(cpu / out["msi"]["ms_first_steps"], 2)
        out["speedup_tdma_vs_cpu_first_steps"] = round(cpu / out["tdma"]["ms_first_steps"], 2)
    except Exception as e:  # the checker is optional here
        log("time step cpu: %s" % e)
    log("time step 258x130: %s" % out)
    return out


def grid_tstep_run(reps=20, nx=1026, ny=514):
    """Once-per-run and once-per-step work either side of the solves on the
    1026x514 grid: the tail of grid_create (centroids, metrics, radii of
    mod_grid.f90) and the Scarborough time step (four mean convergence
    coefficients of mod_tstepper_scarborough.f90), device arrays, HIP events
    on the stream the kernels run on; the oracle (one core) beside it."""
    import ctypes as C

    import torch

    from mach2d import _lib, grid, tstep
    from mach2d import problems as P
    L = _lib.load()
    g = P.ogive_grid(nx, ny, 1.05)
    N = nx * ny
    x = torch.tensor(g.x.reshape(-1), device="cuda")
    y = torch.tensor(g.y.reshape(-1), device="cuda")
    xp, yp = grid.centroids(x, y, nx, ny, 2)
    outs = [torch.empty(N, dtype=torch.float64, device="cuda") for _ in grid.METRICS]
    pin = [C.c_void_p(t.data_ptr()) for t in (x, y, xp, yp)]
    pout = [C.c_void_p(t.data_ptr()) for t in outs]
    i_nx, i_ny = C.byref(C.c_int(nx)), C.byref(C.c_int(ny))
    metrics = lambda: L.m2d_get_metrics(i_nx, i_ny, *pin, *pout)
    full = lambda: grid.grid_create(x, y, nx, ny, kcm=2, coord=1)
    rng = np.random.default_rng(7)
    coefs = [torch.tensor(rng.uniform(-1, 1, k * N), device="cuda")
             for k in (9, 9, 9, 5)]
    dtf = lambda: tstep.dt(1e-3, *coefs, nx, ny, 0.1, 1.0, 10.0)
    res = {}
    for name, f, nbytes in (("metrics", metrics, 19 * 8 * N), ("grid_create", full, None),
                            ("tstepper_dt", dtf, None)):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms": round(ms, 4)}
        if nbytes:
            res[name]["gbps"] = round(nbytes / (ms * 1e-3) / 1e9, 1)
            res[name]["frac_hbm_peak"] = round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        hx, hy = g.x.reshape(-1), g.y.reshape(-1)
        t = time.perf_counter()
        hxp, hyp = O.centroids(hx, hy, nx, ny, 2)
        O.grid_metrics(hx, hy, hxp, hyp, nx, ny)
        O.radius(1, hy, hyp, nx, ny)
        res["grid_create"]["cpu_ms"] = round((time.perf_counter() - t) * 1e3, 2)
        hc = [c.cpu().numpy() for c in coefs]
        t = time.perf_counter()
        O.dt_scarborough(1e-3, *hc, nx, ny, 0.1, 1.0, 10.0)
        res["tstepper_dt"]["cpu_ms"] = round((time.perf_counter() - t) * 1e3, 2)
    except Exception as e:  # the checker is optional here
        log("grid/tstep cpu: %s" % e)
    log("grid/tstep 1026x514: %s" % res)
    return res


def batch_cpu_run(count=256, budget_s=12.0):
    """The 256-system batch on the host: the reference lu2d9 + fb2d9 (compiled
    Fortran) with one system per thread over all host cores, on a bounded
    sample (systems taken in order until the budget is spent, each solved to
    the 1e-10 drop); solves/s of the sample, the batch's total scaled by
    the sample's share of its points."""
    import concurrent.futures as cf
    import threading

    from mach2d import problems as P
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ref = O.ref_available()
    R, path = O.ref_timing() if ref else (None, None)
    sizes = P.cfg4_sizes(count)
    threads = min(16, host_cpu()["usable_cores"])

    def solve(q):
        nx, ny = sizes[q]
        s = P.cfg4_system(q, nx, ny)
        if ref:   # this worker thread has the stack fb2d9 needs
            R.ref_solver_msi9d(nx, ny, 20000, GAMMA, np.ascontiguousarray(s.a),
                                     np.ascontiguousarray(s.rhs), np.zeros(s.N))
        else:
            O.solve_msi9d(s.a, s.rhs, np.zeros(s.N), nx, ny, 20000, GAMMA)
        return nx * ny

    old = threading.stack_size()
    threading.stack_size(256 * 1024 * 1024)   # fb2d9's automatic arrays
    try:
        t0 = time.perf_counter()
        done, pts, q = 0, 0, 0
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            while q < count and time.perf_counter() - t0 < budget_s:
                futs = [ex.submit(solve, k) for k in range(q, min(count, q + threads))]
                for fu in futs:
                    pts += fu.result()
                    done += 1
                q += len(futs)
        dt = time.perf_counter() - t0
    finally:
        threading.stack_size(old)
    total = sum(nx * ny for nx, ny in sizes)
    return {"solves_per_s": round(done / dt, 3), "threads": threads, "systems": done,
            "seconds": round(dt, 2), "kind": "reference" if ref else "port",
            "build": O.REF_FLAGS[path] if ref else "gcc -O2 (the oracle)",
            "batch_seconds_est": round(dt * total / pts, 1)}


def host_cpu():
    """Model name of the host CPU and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "usable_cores": usable}


def cpu_baseline(s, nit, budget_s=15.0):
    """Reference lu2d9 + fb2d9 (compiled Fortran, the -O3 build, one core)
    on a bounded sample: the factorization plus M iterations; solves/s
    extrapolated to the nit iterations the same algorithm needs for the
    1e-10 drop."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    kind = "reference" if O.ref_available() else "port"
    if kind == "reference":
        R, path = O.ref_timing()
        build = O.REF_FLAGS[path] + " (" + os.path.basename(path) + ")"
        lu = lambda: O.ref_lu2d9(s.a, s.nx, s.ny, lib=R)
        fb = lambda dl, du, m: O.ref_fb2d9(s.a, dl, du, s.nx, s.ny, np.zeros(s.N), 0.0, m,
                                           s.rhs, lib=R)
    else:
        build = "gcc -O2 -ffp-contract=off (the oracle's C restatement)"
        lu = lambda: O.lu2d9(s.a, s.nx, s.ny)
        fb = lambda dl, du, m: O.fb2d9(s.a, dl, du, s.nx, s.ny, np.zeros(s.N), 0.0, m, s.rhs)
    t = time.perf_counter()
    dl, du = lu()
    t_lu = time.perf_counter() - t
    t = time.perf_counter()
    fb(dl, du, 2)
    t2 = time.perf_counter() - t
    m = int(max(2, min(nit, (budget_s - t_lu) / max(t2 / 2, 1e-6))))
    t = time.perf_counter()
    fb(dl, du, m)
    t_m = time.perf_counter() - t
    t_iter = t_m / m
    t_solve = t_lu + nit * t_iter
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": 1, "kind": kind,
            "build": build, "cpu": host_cpu(),
            "sample": "lu2d9 + %d fb2d9 iterations of the 1026x514 system (%.1fs); "
                      "solve time = t_lu %.3fs + %d iterations x %.2f ms" % (
                          m, t_lu + t_m, t_lu, nit, t_iter * 1e3)}


def solve_bytes(mode, N, st):
    if mode == "fp64":
        return 8 * N * (18 + 11 + 24 * st.nit)
    return 4 * N * (18 + 24 * st.nit) + 8 * N * (11 * (st.outer + 1) + 3 * st.outer)


def roofline(kt, N, traffic_known=True):
    if not kt:
        return None
    dom = max(kt, key=lambda k: kt[k][0])
    base = dom.replace("32", "")
    size = 4 if dom.endswith("32") else 8
    per_launch = BYTES_PER_POINT[base] * size * N
    avg_ms = kt[dom][0] / kt[dom][1]
    ach = per_launch / (avg_ms * 1e-3) / 1e9
    traffic = MEASURED_TRAFFIC.get(dom) if traffic_known else None
    return {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
            "traffic": int(traffic) if traffic else None,
            "avg_launch_us": round(avg_ms * 1e3, 2), "bytes_per_launch": per_launch}


def batch_run(mode, steps, warmup, count=256, same=None, learned=False):
    """256 independent p' systems (own grid size, stretching and RHS) solved
    together to the 1e-10 drop; with `same` (a system), `count` systems with
    its matrix and their own seeded Gaussian right-hand sides."""
    import torch

    from mach2d import problems as P
    from mach2d import solvers as S
    t = time.time()
    if same is None:
        sizes = P.cfg4_sizes(count)
        systems = P.cfg4_systems(count)
        a = [torch.tensor(x.a.reshape(-1), device="cuda") for x in systems]
        rhs = [torch.tensor(x.rhs, device="cuda") for x in systems]
    else:
        # `count` distinct matrices (each system its own copy in HBM, so no
        # two systems share cached coefficients): the matrix of `same` with
        # the diagonal of its interior rows scaled by a seeded factor in
        # [1, 1.001], and a seeded Gaussian right-hand side per system
        sizes = [(same.nx, same.ny)] * count
        systems = [same] * count
        a0 = torch.tensor(same.a.reshape(-1), device="cuda")
        N = same.N
        interior = torch.tensor(same.rhs != 0.0, device="cuda")
        gen = torch.Generator(device="cuda")
        a, rhs = [], []
        for q in range(count):
            gen.manual_seed(7000 + q)
            r = torch.randn(N, dtype=torch.float64, device="cuda", generator=gen)
            rhs.append(torch.where(interior, r, torch.zeros_like(r)))
            f = 1.0 + 1e-3 * float(np.random.default_rng(8000 + q).uniform())
            aq = a0.clone()
            aq[4 * N:5 * N] = torch.where(interior, aq[4 * N:5 * N] * f, aq[4 * N:5 * N])
            a.append(aq)
    log("batch: %d systems, N = %d in total, built in %.1fs" % (
        count, sum(x.N for x in systems), time.time() - t))
    var = [torch.zeros(x.N, dtype=torch.float64, device="cuda") for x in systems]
    b = S.BatchMSI(sizes, precision=mode)
    b.factor(a)
    layout, ms = b.tune(a, rhs, iters=4)
    log("batch %s: layout %d (%s), ms/iteration per layout %s" % (
        mode, layout, S.BatchMSI.LAYOUTS[layout], [round(v, 3) for v in ms]))

    net = helper = None
    if learned:
        from mach2d import learned as Lr
        from mach2d import unet as U
        if not os.path.exists(Lr.default_weights()):
            b.close()
            return None
        net = U.UNet()
        net.load(Lr.default_weights())
        helper = S.DeviceMSI(sizes[0][0], sizes[0][1], precision="fp64")
    used = [False] * count

    def one():
        for v in var:
            v.zero_()
        b.factor(a)
        if net is None:
            return b.solve(a, rhs, var, GAMMA, NITM_CAP)
        gam = []
        if len(set(sizes)) == 1:   # one network pass over every system
            nx, ny = sizes[0]
            for q, (u_, rg, rb) in enumerate(Lr.prepare_guesses(helper, net, a, rhs, var, nx,
                                                                ny)):
                used[q] = u_
                gam.append(GAMMA * rb / rg)
            return b.solve(a, rhs, var, gam, NITM_CAP)
        for q in range(count):   # guess, energy-minimising amplitude; stop at GAMMA ||b||
            nx, ny = sizes[q]
            used[q], rg, rb = Lr.prepare_guess(helper, net, a[q], rhs[q], var[q], nx, ny)
            gam.append(GAMMA * rb / rg)
        return b.solve(a, rhs, var, gam, NITM_CAP)

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = [one() for _ in range(steps)]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    rc, st = res[-1]
    ok = rc == 0 and all(x.status == 0 and x.rn < x.gammarp for x in st)
    nbytes = sum(solve_bytes(mode, x.N, y) for x, y in zip(systems, st))
    b.close()
    if helper is not None:
        helper.close()
    out = {"mode": mode, "systems": count, "solves_per_s": round(count / dt, 2),
           "ms_per_batch": round(dt * 1e3, 3), "gbps": round(nbytes / dt / 1e9, 2),
           "iterations_max": max(y.nit for y in st), "layout": layout, "correct": bool(ok)}
    if learned:
        out["guesses_used"] = int(sum(used))
    return out


def target_comparison(summary, batch3):
    """The 1026x514 solve to the 1e-10 drop: the refined (FP32 sweeps in
    FP64 refinement) and learned-guess paths against plain FP64 MSI, in
    solves/s, three ways: the same batch of distinct systems in both
    precisions (like for like), one system at a time, and the batched paths
    against one plain FP64 solve at a time."""
    if "fp64" not in summary:
        return None
    plain = summary["fp64"]["solves_per_s"]
    out = {"plain_fp64_single_solves_per_s": plain, "target": 20.0}
    single = {}
    for k in ("fp32_ir", "learned_fp32_ir"):
        v = summary.get(k)
        if v and v.get("correct"):
            single[k] = round(v["solves_per_s"] / plain, 3)
    out["single_vs_single_fp64"] = single
    if batch3 and batch3.get("fp64", {}).get("correct"):
        b64 = batch3["fp64"]["solves_per_s"]
        out["plain_fp64_batched_solves_per_s"] = b64
        same, cross = {}, {}
        for k in ("fp32_ir", "learned_fp32_ir"):
            v = batch3.get(k)
            if v and v.get("correct"):
                same[k] = round(v["solves_per_s"] / b64, 3)
                cross[k] = round(v["solves_per_s"] / plain, 2)
        out["batched_vs_batched_fp64"] = same
        out["batched_vs_single_fp64"] = cross
        if cross:
            best = max(cross, key=lambda k: cross[k])
            out["best_batched_vs_single_fp64"] = {"path": "batch_1026x514." + best,
                                                  "speedup": cross[best]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="best", choices=["best", "fp64", "fp32_ir"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-cfg3", type=int, default=48,
                    help="systems in the batch of 1026x514 p' systems")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    s = make_system()
    modes = ["fp64", "fp32_ir"] if args.mode == "best" else [args.mode]
    runs = {m: gpu_run(s, m, args.steps, args.warmup) for m in modes}
    x64 = runs["fp64"]["x"] if "fp64" in runs else None
    summary = {}
    for m, r in runs.items():
        st = r["st"]
        per = r["dt"] / args.steps
        e = {"solves_per_s": round(1.0 / per, 4), "ms_per_solve": round(per * 1e3, 3),
             "nit": st.nit, "outer": st.outer, "correct": bool(r["ok"]),
             "gbps": round(solve_bytes(m, s.N, st) / per / 1e9, 2)}
        if m != "fp64" and x64 is not None:
            e["rel_error_vs_fp64"] = float(np.linalg.norm(r["x"] - x64) / np.linalg.norm(x64))
            e["correct"] = bool(r["ok"] and e["rel_error_vs_fp64"] <= 1e-8)
        summary[m] = e
    if args.mode == "best":
        lr = learned_run(s, args.steps, args.warmup, x64)
        if lr is not None:
            summary["learned_fp32_ir"] = lr
    msi5 = msi5_run(args.steps, args.warmup) if args.mode == "best" else None
    pl_loop = pl_loop_run(max(1, args.steps)) if args.mode == "best" else None
    grid_ts = grid_tstep_run() if args.mode == "best" else None
    tstep_full = time_step_run() if args.mode == "best" else None
    tdma = tdma_run() if args.mode == "best" else None
    head = "fp32_ir" if "fp32_ir" in summary and summary["fp32_ir"]["correct"] else modes[0]
    r, e = runs[head], summary[head]
    st = r["st"]
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        nit64 = runs["fp64"]["st"].nit if "fp64" in runs else st.nit
        cpu = cpu_baseline(s, nit64)
        log("cpu: %.4g solves/s (%s)" % (cpu["value"], cpu["sample"]))
    batch = batch3 = None
    if not args.no_batch:
        batch = {m: batch_run(m, max(1, args.steps - 1), 1) for m in ("fp64", "fp32_ir")}
        for m, v in batch.items():
            log("batch %s: %s" % (m, v))
        batch3 = {m: batch_run(m, 1, 1, count=args.batch_cfg3, same=s)
                  for m in ("fp64", "fp32_ir")}
        lb = batch_run("fp32_ir", 1, 1, count=args.batch_cfg3, same=s, learned=True)
        if lb is not None:
            batch3["learned_fp32_ir"] = lb
        if not args.no_cpu:
            try:
                bc = batch_cpu_run()
                log("batch cpu: %s" % bc)
                for m, v in batch.items():
                    v["cpu_threads_solves_per_s"] = bc["solves_per_s"]
                batch["cpu"] = bc
            except Exception as e:  # the checker is optional here
                log("batch cpu: %s" % e)
        for m, v in batch3.items():
            log("batch 1026x514 %s: %s" % (m, v))
            if "fp64" in summary:
                v["speedup_vs_plain_fp64"] = round(
                    v["solves_per_s"] / summary["fp64"]["solves_per_s"], 2)
    out = {
        "metric": METRIC,
        "value": e["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": e["ms_per_solve"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "p' MSI solve, 1026x514 wall-resolved ogive grid "
                               "(aspect 1e4), zero guess, 1e-10 residual drop",
                   "mode": head, "nx": s.nx, "ny": s.ny, "nit": st.nit, "outer": st.outer,
                   "gamma": GAMMA, "correct": e["correct"]},
        "gbps": e["gbps"],
        "speedup_vs_plain_fp64": (round(e["solves_per_s"] / summary["fp64"]["solves_per_s"], 3)
                                  if "fp64" in summary else None),
        "modes": summary,
        "kernels_ms": {k: [round(v[0], 3), v[1]] for k, v in r["kt"].items()},
        "roofline": roofline(r["kt"], s.N),
        "cpu_baseline": cpu,
        "batch": batch,
        "batch_1026x514": batch3,
        "msi2d5_1026x514": msi5,
        "pl_loop_258x130": pl_loop,
        "grid_tstep_1026x514": grid_ts,
        "simplec_time_step_258x130": tstep_full,
        "tdma_258x130": tdma,
        "target_comparison": target_comparison(summary, batch3),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
