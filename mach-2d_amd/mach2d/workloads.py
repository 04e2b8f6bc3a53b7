"""Single workloads of the benchmark, one per process, for profiling runs
(rocprofv3 traces one command):

    python -m mach2d.workloads single --mode fp32_ir --solves 2
    python -m mach2d.workloads batch --mode fp64 --count 256
    python -m mach2d.workloads batch --mode fp32_ir --count 32 --same-1026x514 [--layout L]
    python -m mach2d.workloads pl_loop --solves 3
    python -m mach2d.workloads tdma --nx 258 --ny 130 --nitm 5 --count 2 --solves 20
    python -m mach2d.workloads time_step --solver9d 1 --solves 20

single: the 1026x514 p' system (wall-resolved ogive grid, aspect 1e4,
Gaussian right-hand side) solved to a 1e-10 drop from zero, `--solves` times
after one warm-up solve.  batch: `count` systems of uneven size (64s x 32s,
s log-uniform on [1, 8]) or `count` copies of the 1026x514 matrix with their
own right-hand sides, solved together (autotuned layout unless given).
pl_loop: the internal-flow pressure-correction loop (20 itp iterations,
p' solves of at most 20 MSI iterations) on a 258x130 grid.  Prints one JSON
line with the timing and iteration counts.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import problems as P

GAMMA = 1e-10
NITM_CAP = 20000


def single(mode, solves, staged, inner_gamma=None, inner_nitm=25):
    import torch

    from .solvers import DeviceMSI
    s = P.pcorr_system(**{k: v for k, v in P.CFG3.items() if k != "gamma"})
    dev = DeviceMSI(s.nx, s.ny, mode)
    dev.set_staging(staged)
    if mode == "fp32_ir":
        dev.set_inner(1e-4 if inner_gamma is None else inner_gamma, inner_nitm)
    a = torch.tensor(s.a.reshape(-1), device="cuda")
    b = torch.tensor(s.rhs, device="cuda")
    x = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    out = []
    for k in range(solves + 1):
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev.factor(a)
        st = dev.solve(a, b, x, GAMMA, NITM_CAP)
        torch.cuda.synchronize()
        if k:
            out.append((time.perf_counter() - t0, st.nit, st.outer, st.status))
        print("solve %d: %.1f ms, nit %d" % (k, (time.perf_counter() - t0) * 1e3, st.nit),
              file=sys.stderr, flush=True)
    dev.close()
    return {"workload": "single 1026x514", "mode": mode, "staged": staged,
            "inner_gamma": inner_gamma, "inner_nitm": inner_nitm,
            "ms": [round(o[0] * 1e3, 3) for o in out], "nit": out[-1][1],
            "outer": out[-1][2], "status": out[-1][3]}


def batch(mode, count, same, layout, inner_gamma=None, inner_nitm=50):
    import torch

    from .solvers import BatchMSI
    if same:
        s = P.pcorr_system(**{k: v for k, v in P.CFG3.items() if k != "gamma"})
        sizes = [(s.nx, s.ny)] * count
        a0 = torch.tensor(s.a.reshape(-1), device="cuda")
        a = [a0] * count
        mask = torch.tensor(s.rhs != 0.0, device="cuda")
        g = torch.Generator(device="cuda")
        rhs = []
        for q in range(count):
            g.manual_seed(7000 + q)
            r = torch.randn(s.N, dtype=torch.float64, device="cuda", generator=g)
            rhs.append(torch.where(mask, r, torch.zeros_like(r)))
    else:
        sizes = P.cfg4_sizes(count)
        systems = P.cfg4_systems(count)
        a = [torch.tensor(x.a.reshape(-1), device="cuda") for x in systems]
        rhs = [torch.tensor(x.rhs, device="cuda") for x in systems]
    var = [torch.zeros(nx * ny, dtype=torch.float64, device="cuda") for nx, ny in sizes]
    b = BatchMSI(sizes, precision=mode)
    if mode == "fp32_ir":
        b.set_inner(1e-4 if inner_gamma is None else inner_gamma, inner_nitm)
    b.factor(a)
    if layout is None:
        layout, ms = b.tune(a, rhs, iters=4)
    else:
        b.set_tuning(layout, 8)
    print("layout %d" % layout, file=sys.stderr, flush=True)
    for v in var:
        v.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.factor(a)
    rc, st = b.solve(a, rhs, var, GAMMA, NITM_CAP)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    b.close()
    return {"workload": "batch", "mode": mode, "count": count, "same_1026x514": same,
            "inner_gamma": inner_gamma, "inner_nitm": inner_nitm,
            "layout": layout, "ms": round(dt * 1e3, 3), "solves_per_s": round(count / dt, 2),
            "iterations_max": max(x.nit for x in st), "status": rc}


def batch_split(mode, count, groups, layout, inner_nitm=50):
    """`count` copies of the 1026x514 matrix (own right-hand sides) solved as
    `groups` batches at once, each on its own stream, driven from its own
    host thread, so one group's sweeps overlap another group's residuals."""
    import threading

    import torch

    from .solvers import BatchMSI
    s = P.pcorr_system(**{k: v for k, v in P.CFG3.items() if k != "gamma"})
    a0 = torch.tensor(s.a.reshape(-1), device="cuda")
    mask = torch.tensor(s.rhs != 0.0, device="cuda")
    g = torch.Generator(device="cuda")
    per = [count // groups + (1 if q < count % groups else 0) for q in range(groups)]
    jobs = []
    for q, m in enumerate(per):
        rhs = []
        for k in range(m):
            g.manual_seed(7000 + sum(per[:q]) + k)
            r = torch.randn(s.N, dtype=torch.float64, device="cuda", generator=g)
            rhs.append(torch.where(mask, r, torch.zeros_like(r)))
        st = torch.cuda.Stream()
        b = BatchMSI([(s.nx, s.ny)] * m, precision=mode, stream=st.cuda_stream)
        if mode == "fp32_ir":
            b.set_inner(1e-4, inner_nitm)
        b.set_tuning(layout, 8)
        var = [torch.zeros(s.N, dtype=torch.float64, device="cuda") for _ in range(m)]
        jobs.append((b, [a0] * m, rhs, var))
    torch.cuda.synchronize()
    out = [None] * groups

    def run(q):
        b, a, rhs, var = jobs[q]
        b.factor(a)
        out[q] = b.solve(a, rhs, var, GAMMA, NITM_CAP)

    for q in range(groups):  # warm-up (graph capture)
        run(q)
    for _, _, _, var in jobs:
        for v in var:
            v.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    th = [threading.Thread(target=run, args=(q,)) for q in range(groups)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rc = max(o[0] for o in out)
    its = max(x.nit for o in out for x in o[1])
    for b, *_ in jobs:
        b.close()
    return {"workload": "batch split", "mode": mode, "count": count, "groups": groups,
            "layout": layout, "inner_nitm": inner_nitm, "ms": round(dt * 1e3, 3),
            "solves_per_s": round(count / dt, 2), "iterations_max": its, "status": rc}


def pl_loop(nx, ny, itpmax, reps, pl_graph=1):
    """The internal-flow pressure-correction loop (m2d_intflow_pl_loop) on an
    nx x ny grid with seeded fields: ms per call and per itp iteration (the
    same arrays every call, reset to the seeded fields in between)."""
    import numpy as np
    import torch

    from . import simplec
    from .solvers import DeviceMSI
    f = P.flow_fields(nx, ny, seed=1)
    rng = np.random.default_rng(18)
    for k in ("due", "dve", "dun", "dvn"):
        f[k] = rng.uniform(-1e-3, 1e-3, size=nx * ny)
    ap = np.zeros((5, nx * ny))
    ap[2] = 1.0
    f["ap"] = ap.reshape(-1)
    f["bp"] = np.zeros(nx * ny)
    p_avg = float(np.mean(f["p"]))
    dev = DeviceMSI(nx, ny, "fp64", ndiag=5)
    simplec.set_pl_graph(dev, pl_graph)
    base = {k: torch.tensor(np.ascontiguousarray(v), device="cuda") for k, v in f.items()}
    t = {k: v.clone() for k, v in base.items()}
    out = []
    for r in range(reps + 1):
        for k, v in base.items():
            t[k].copy_(v)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        itp, normpl = simplec.intflow_pl_loop(dev, t, modvis=0, dt=2e-3, beta=0.3, p_avg=p_avg,
                                              tolm=0.0, itpmax=itpmax, nitm_p=20, tol_p=1e-2)
        torch.cuda.synchronize()
        if r:
            out.append((time.perf_counter() - t0) * 1e3)
    reads = simplec.pl_loop_info(dev)[0]
    dev.close()
    ms = sorted(out)[len(out) // 2]
    return {"workload": "pl_loop", "nx": nx, "ny": ny, "pl_graph": pl_graph, "itp": itp,
            "ms": round(ms, 3), "ms_per_itp": round(ms / itp, 4), "normpl": normpl,
            "host_reads": reads}


def tdma(nx, ny, nitm, count, reps):
    """tdma2d9 on `count` seeded systems of an nx x ny grid in the same
    launches (the u and v solves of a SIMPLEC iteration: count 2), `reps`
    calls after one warm-up: ms per call and per TDMA iteration."""
    import numpy as np
    import torch

    from .solvers import tdma2d9_device_many
    systems = [P.random_f3_system(nx, ny, seed=q) for q in range(count)]
    a = [torch.tensor(s.a.reshape(-1), device="cuda") for s in systems]
    b = [torch.tensor(s.rhs, device="cuda") for s in systems]
    x = [torch.zeros(nx * ny, dtype=torch.float64, device="cuda") for _ in systems]
    out = []
    for r in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tdma2d9_device_many(nx, ny, nitm, a, b, x)
        if r:
            out.append((time.perf_counter() - t0) * 1e3)
    ms = sorted(out)[len(out) // 2]
    return {"workload": "tdma", "nx": nx, "ny": ny, "nitm": nitm, "count": count,
            "ms": round(ms, 4), "ms_per_iteration": round(ms / nitm, 4)}


def time_step(solver9d, steps, pl_graph=1):
    """The internal-flow time step (m2d_intflow_time_step) on the 258x130
    nozzle from inlet Mach 0.2, `steps` steps after one warm-up step;
    pl_graph 0 enqueues the p' loop's kernels directly instead of replaying
    captured graphs."""
    import numpy as np
    import torch

    from . import simplec
    from .solvers import DeviceMSI
    c = P.CFG2
    nx, ny = c["nx"], c["ny"]
    prm = dict(modvis=0, coord=1, solver9d=solver9d, itmmax=1, itpmax=20, nitm_u=5, nitm_p=20,
               tol_u=1e-2, tol_p=1e-2, tolm=1e-2, beta=0.0, Rg=287.0, gamma=1.4, P0=1.2e5,
               T0=310.0, Cpref=1004.5, Href=1004.5 * 310.0, Twall=-1.0, h0=0.01, mincc=0.99,
               maxcc=0.9999)
    f0 = P.time_step_fields(nx, ny, 5, prm, mach_in=c["mach_in"])
    d = {k: torch.tensor(np.ascontiguousarray(v).reshape(-1), device="cuda")
         for k, v in f0.items()}
    dev9 = DeviceMSI(nx, ny) if solver9d == 0 else None
    dev5 = DeviceMSI(nx, ny, ndiag=5)
    simplec.set_pl_graph(dev5, pl_graph)
    torch.cuda.synchronize()
    dt = c["dt0"]
    dt, r = simplec.intflow_time_step(dev9, dev5, d, prm, dt)
    t0 = time.perf_counter()
    itps = 0
    for _ in range(steps):
        dt, r = simplec.intflow_time_step(dev9, dev5, d, prm, dt)
        itps += r["itp"]
    ms = (time.perf_counter() - t0) * 1e3 / max(steps, 1)
    return {"workload": "time_step", "solver9d": solver9d, "pl_graph": pl_graph, "steps": steps,
            "ms_per_step": round(ms, 3), "itp_per_step": round(itps / max(steps, 1), 2),
            "pl_host_reads": simplec.pl_loop_info(dev5)[0]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["single", "batch", "split", "pl_loop", "tdma",
                                     "time_step"])
    ap.add_argument("--nx", type=int, default=258)
    ap.add_argument("--ny", type=int, default=130)
    ap.add_argument("--nitm", type=int, default=5)
    ap.add_argument("--solver9d", type=int, default=0)
    ap.add_argument("--groups", type=int, default=2)
    ap.add_argument("--mode", default="fp32_ir", choices=["fp64", "fp32_ir"])
    ap.add_argument("--solves", type=int, default=1)
    ap.add_argument("--staged", type=int, default=0)
    ap.add_argument("--pl-graph", dest="pl_graph", type=int, default=1)
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--same-1026x514", dest="same", action="store_true")
    ap.add_argument("--layout", type=int, default=None)
    ap.add_argument("--inner-gamma", dest="inner_gamma", type=float, default=None)
    ap.add_argument("--inner-nitm", dest="inner_nitm", type=int, default=None)
    ap.add_argument("--tdma-mode", dest="tdma_mode", type=int, default=0,
                    help="line solver of the device TDMA: 0 recurrence, 1 cyclic reduction")
    args = ap.parse_args(argv)
    if args.tdma_mode:
        from .solvers import set_tdma_mode
        set_tdma_mode(args.tdma_mode)
    if args.what == "single":
        out = single(args.mode, args.solves, bool(args.staged), args.inner_gamma,
                     args.inner_nitm or 25)
    elif args.what == "pl_loop":
        out = pl_loop(258, 130, 20, max(1, args.solves), args.pl_graph)
    elif args.what == "tdma":
        out = tdma(args.nx, args.ny, args.nitm, args.count, max(1, args.solves))
    elif args.what == "time_step":
        out = time_step(args.solver9d, max(1, args.solves), args.pl_graph)
    elif args.what == "split":
        out = batch_split(args.mode, args.count, args.groups,
                          0 if args.layout is None else args.layout, args.inner_nitm or 50)
    else:
        out = batch(args.mode, args.count, args.same, args.layout, args.inner_gamma,
                    args.inner_nitm or 50)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
