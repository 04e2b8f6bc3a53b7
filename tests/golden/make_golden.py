"""Write the golden vectors of the hot path by running the REFERENCE.

The reference's Fortran sources (src/solvers/*.f90, src/equations/
mach2d_coef.f90) are compiled in place by `make -C oracle ref` into
oracle/_ref/libm2dref.so (with a C-callable shim of ours); this script calls
them on small seeded inputs and stores inputs and outputs as .npz files next
to it.  It needs the reference tree, so it runs only in the build container:

    python tests/golden/make_golden.py

Fixtures (float64, diagonal-major arrays (k, nx*ny)):
  msi_small.npz   lu2d9 factors and fb2d9 iterates (nitm = 1,2,3,5,10 with
                  gamma = 0, and a full solve at gamma = 1e-10) on 6x5, 10x6,
                  34x18 random systems and on 10x6 / 34x18 p' systems
  cfg1.npz        130x66 p' system (the CPU-runnable configuration): fb2d9
                  iterates after 1, 10, 100 iterations and the full 500
  tdma.npz        tdma2d9 after nitm = 1, 2, 5 on random systems whose
                  outside couplings are NOT zero (exercises every branch)
  norms.npz       norm_l1_9d, norm_l1_9d_relative (both branches), norm_l1_b
  embed5.npz      lu2d5 + fb2d5 against lu2d9 + fb2d9 on the zero-corner
                  embedding of the same 5-point system
  solver_msi9d.npz  class_solver_msi9d init + solve (the plugin object)
  coef.npz        every routine of module `coefficients` used by the hot path
                  (face densities; u, v, T, p' coefficients and sources; the
                  SIMPLEC glue: face velocities, SIMPLEC coefficients, the
                  p/ro/u/v/face corrections, boundary-face flags, the mass
                  residual) on
                  12x8 and 20x11 ogive grids (axisymmetric Navier-Stokes,
                  planar Euler, planar Navier-Stokes), a 10x7 Cartesian grid,
                  and a 12x8 case with faces on the axis (r = 0) for the
                  axis branches of get_u_source; outputs start from seeded
                  values so untouched points are checked too
  simplec_pairs_34x18.m2dt  learned-guess training pairs (M2DT format) from
                  12 SIMPLEC time steps of a 34x18 nozzle composed from the
                  reference's routines, targets by its class_solver_msi9d
  grid.npz        get_real_centroids_xy (opt 1, 2), get_metrics and
                  get_radius (coord 0, 1) of mod_grid.f90 (module-private,
                  called through their module symbols) on ogive, nozzle,
                  perturbed skewed and strongly stretched grids
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "mach-2d_amd"))

import oracle as O  # noqa: E402
from mach2d import problems as P  # noqa: E402


def save(name, **arrs):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrs)
    print("%-18s %8.1f KB" % (name, os.path.getsize(path) / 1024))


def iterates(a, nx, ny, rhs, ks, x0=None):
    dl, du = O.ref_lu2d9(a, nx, ny)
    x0 = np.zeros(nx * ny) if x0 is None else x0
    out = {}
    for k in ks:
        out["var_%d" % k] = O.ref_fb2d9(a, dl, du, nx, ny, x0, 0.0, k, rhs)
    return dl, du, out


def msi_small():
    arrs = {}
    cases = [("r6x5", P.random_f3_system(6, 5, seed=11)),
             ("r10x6", P.random_f3_system(10, 6, seed=12)),
             ("r34x18", P.random_f3_system(34, 18, seed=13)),
             ("p10x6", P.pcorr_system(10, 6, seed=14)),
             ("p34x18", P.pcorr_system(34, 18, seed=15))]
    for name, s in cases:
        dl, du, it = iterates(s.a, s.nx, s.ny, s.rhs, [1, 2, 3, 5, 10])
        full = O.ref_fb2d9(s.a, dl, du, s.nx, s.ny, np.zeros(s.N), 1e-10, 500, s.rhs)
        arrs.update({name + "_a": s.a, name + "_rhs": s.rhs, name + "_dl": dl, name + "_du": du,
                     name + "_full": full, name + "_shape": np.array([s.nx, s.ny])})
        arrs.update({name + "_" + k: v for k, v in it.items()})
    arrs["cases"] = np.array([c for c, _ in cases])
    save("msi_small.npz", **arrs)


def cfg1():
    c = P.CFG1
    s = P.pcorr_system(c["nx"], c["ny"], seed=c["seed"], q=c["q"], rhs=c["rhs"])
    dl, du = O.ref_lu2d9(s.a, s.nx, s.ny)
    arrs = dict(a=s.a, rhs=s.rhs, shape=np.array([s.nx, s.ny]))
    for k in (1, 10, 100):
        arrs["var_%d" % k] = O.ref_fb2d9(s.a, dl, du, s.nx, s.ny, np.zeros(s.N), 0.0, k, s.rhs)
    arrs["var_full"] = O.ref_fb2d9(s.a, dl, du, s.nx, s.ny, np.zeros(s.N), c["gamma"],
                                   c["nitm"], s.rhs)
    arrs["dl1"] = dl[0]
    arrs["du1"] = du[0]
    save("cfg1.npz", **arrs)


def tdma():
    arrs = {}
    rng = np.random.default_rng(21)
    for nx, ny in [(6, 5), (10, 6), (34, 18), (3, 7)]:
        a = -rng.uniform(0.05, 1.0, size=(9, nx * ny))
        a[4] = 1.5 * np.sum(np.abs(a[[0, 1, 2, 3, 5, 6, 7, 8]]), axis=0) + 0.1
        b = rng.uniform(-1, 1, nx * ny)
        x0 = rng.uniform(-1, 1, nx * ny)
        name = "t%dx%d" % (nx, ny)
        arrs.update({name + "_a": a, name + "_b": b, name + "_x0": x0,
                     name + "_shape": np.array([nx, ny])})
        for k in (1, 2, 5):
            arrs[name + "_x%d" % k] = O.ref_tdma2d9(a, b, x0, nx, ny, k)
    save("tdma.npz", **arrs)


def norms():
    arrs = {}
    rng = np.random.default_rng(31)
    for nx, ny in [(6, 5), (34, 18)]:
        name = "n%dx%d" % (nx, ny)
        a = rng.uniform(-1, 1, size=(9, nx * ny))
        b = rng.uniform(-1, 1, nx * ny)
        v = rng.uniform(-1, 1, nx * ny)
        arrs.update({name + "_a": a, name + "_b": b, name + "_var": v,
                     name + "_shape": np.array([nx, ny]),
                     name + "_l1": np.array(O.ref_norm_l1_9d(a, b, v, nx, ny, 0.25)),
                     name + "_rel": np.array(O.ref_norm_l1_9d_relative(a, b, v, nx, ny, 0.25)),
                     name + "_rel0": np.array(O.ref_norm_l1_9d_relative(a, np.zeros_like(b), v,
                                                                         nx, ny, 0.25)),
                     name + "_nb": np.array(O.ref_norm_l1_b(b, nx, ny))})
    save("norms.npz", **arrs)


def embed5():
    arrs = {}
    for nx, ny, seed in [(10, 6, 41), (34, 18, 42)]:
        s9 = P.pcorr_system(nx, ny, seed=seed)
        a5 = s9.a[[1, 3, 4, 5, 7]].copy()        # the 5-point part (S, W, P, E, N)
        a9 = np.zeros_like(s9.a)
        a9[[1, 3, 4, 5, 7]] = a5
        dl5, du5 = O.ref_lu2d5(a5, nx, ny)
        dl9, du9 = O.ref_lu2d9(a9, nx, ny)
        x5 = O.ref_fb2d5(a5, dl5, du5, nx, ny, np.zeros(s9.N), 0.0, 7, s9.rhs)
        x9 = O.ref_fb2d9(a9, dl9, du9, nx, ny, np.zeros(s9.N), 0.0, 7, s9.rhs)
        name = "e%dx%d" % (nx, ny)
        arrs.update({name + "_a5": a5, name + "_rhs": s9.rhs, name + "_dl5": dl5,
                     name + "_du5": du5, name + "_x5": x5, name + "_x9": x9,
                     name + "_shape": np.array([nx, ny])})
    save("embed5.npz", **arrs)


def solver_msi9d():
    s = P.pcorr_system(34, 18, seed=51)
    sol0 = np.linspace(-1, 1, s.N)
    sol = O.ref_solver_msi9d(s.a, s.rhs, sol0, s.nx, s.ny, 20, 1e-2)
    save("solver_msi9d.npz", a=s.a, b=s.rhs, sol0=sol0, sol=sol, shape=np.array([s.nx, s.ny]),
         nitm=np.array(20), tol=np.array(1e-2))


# scalars of the time-step routines: constant thermophysical model and
# chamber conditions of the shipped nozzle parameter file
STEP_SCALARS = dict(Cpref=1004.5, Href=837083.3, gamma=1.4, P0=1038350.4, T0=833.33333333)

COEF_CASES = [  # tag, nx, ny, coord, modvis, cartesian, axis faces
    ("ns_axi", 12, 8, 1, 1, False, False),
    ("euler", 12, 8, 0, 0, False, False),
    ("ns_planar", 20, 11, 0, 1, False, False),
    ("cart", 10, 7, 1, 1, True, False),
    ("axis", 12, 8, 1, 1, False, True),
]


def coef_fields(nx, ny, coord, cart, axis, seed):
    f = P.flow_fields(nx, ny, seed=seed, coord=coord, cartesian=cart)
    if axis:  # east/north faces of a few points on the axis
        f["re"][nx * 3 + 4] = 0.0      # point (6, 4) west face / point (5, 4) east face
        f["re"][nx * 5 + 8] = 0.0
        f["rn"][nx * 2 + 6] = 0.0      # south face of point (7, 4) / north of (7, 3)
        f["rn"][nx * 5 + 3] = 0.0
    # inputs of the SIMPLEC glue: u, v systems from the assembly itself (unit
    # diagonal on the fictitious rows), deferred-correction and viscous
    # sources, old face velocities, SIMPLEC coefficients
    rng = np.random.default_rng(seed + 1)
    sc = dict(modvis=1, dt=2e-3, coord=coord)
    for key, name in (("au", "u_coefficients"), ("av", "v_coefficients")):
        a = O.ref_coef(name, nx, ny, sc, f)["a"]
        a[4][a[4] == 0.0] = 1.0
        f[key] = a
    for k in ("cup", "cvp", "sup", "svp", "uea", "vea", "una", "vna", "due", "dve", "dun",
              "dvn"):
        f[k] = rng.uniform(-1.0, 1.0, size=nx * ny)
    return f


def coef():
    arrs = {}
    rng = np.random.default_rng(91)
    for tag, nx, ny, coord, mv, cart, axis in COEF_CASES:
        f = coef_fields(nx, ny, coord, cart, axis, seed=nx * 31 + ny + coord)
        # Twall: prescribed north wall temperature on the axisymmetric cases,
        # adiabatic wall (Twall <= 0) on the planar ones
        sc = dict(beta=0.35, modvis=mv, dt=2e-3, coord=coord, Rg=287.0,
                  Twall=300.0 if coord == 1 else -1.0, **STEP_SCALARS)
        for k, v in f.items():
            arrs["%s_in_%s" % (tag, k)] = v
        arrs["%s_scalars" % tag] = np.array([nx, ny, coord, mv, sc["beta"], sc["dt"]])
        arrs["%s_twall" % tag] = np.array([sc["Twall"]])
        for name, (_, _, outs) in O.COEF.items():
            init = {k: rng.uniform(-1, 1, size=(1, ny) if m == "ny" else (1, nx) if m == "nx"
                                   else (m, nx * ny) if m else (1, 1)) for k, m in outs}
            for k, v in init.items():
                arrs["%s_%s_init_%s" % (tag, name, k)] = v
            if name == "intflow_time_cycle_update":   # private module procedures
                res = O.ref_time_cycle_update(nx, ny, sc, f, {k: v.copy() for k, v in init.items()})
                res = {k: v.reshape(-1) for k, v in res.items()}
            else:
                res = O.ref_coef(name, nx, ny, sc, f, out={k: v.copy() for k, v in init.items()})
            for k, v in res.items():
                arrs["%s_%s_%s" % (tag, name, k)] = v
    save("coef.npz", **arrs)


GRID_CASES = [("ogive", 34, 18), ("ogive", 3, 3), ("ogive", 4, 9), ("nozzle", 22, 12),
              ("perturbed", 17, 12), ("stretched", 41, 23)]


def grid_xy(kind, nx, ny):
    """Corner coordinates (north-east corner of each volume, the reference's
    x, y) of the fixture grids: the g01 ogive, the nozzle, a seeded randomly
    perturbed skewed grid, a strongly wall-stretched ogive."""
    if kind in ("ogive", "stretched"):
        g = P.ogive_grid(nx, ny, 1.05 if kind == "ogive" else 1.3)
    elif kind == "nozzle":
        g = P.nozzle_grid(nx, ny)
    else:
        rng = np.random.default_rng(nx * ny)
        x = np.zeros((ny, nx))
        y = np.zeros((ny, nx))
        ii, jj = np.meshgrid(np.arange(nx - 1, dtype=float), np.arange(ny - 1, dtype=float))
        x[: ny - 1, : nx - 1] = 1.0 * ii + 0.15 * jj + 0.2 * rng.uniform(-1, 1, ii.shape)
        y[: ny - 1, : nx - 1] = -0.1 * ii + 0.8 * jj + 0.2 * rng.uniform(-1, 1, ii.shape)
        return x.reshape(-1), y.reshape(-1)
    return g.x.reshape(-1).copy(), g.y.reshape(-1).copy()


def grid():
    """get_real_centroids_xy (both options), get_metrics (from the area-
    weighted centroids) and get_radius (both coordinate systems) of
    mod_grid.f90, the reference's own module procedures."""
    arrs = {}
    for kind, nx, ny in GRID_CASES:
        tag = "%s_%dx%d" % (kind, nx, ny)
        x, y = grid_xy(kind, nx, ny)
        arrs[tag + "_x"], arrs[tag + "_y"] = x, y
        for opt in (1, 2):
            xp, yp = O.ref_centroids(x, y, nx, ny, opt)
            arrs["%s_xp%d" % (tag, opt)], arrs["%s_yp%d" % (tag, opt)] = xp, yp
        m = O.ref_grid_metrics(x, y, xp, yp, nx, ny)
        for k, v in m.items():
            arrs["%s_m_%s" % (tag, k)] = v
        for coord in (0, 1):
            for k, v in zip(("radius", "re", "rn", "rp"), O.ref_radius(coord, y, yp, nx, ny)):
                arrs["%s_r%d_%s" % (tag, coord, k)] = v
    arrs["cases"] = np.array(["%s_%dx%d" % c for c in GRID_CASES])
    save("grid.npz", **arrs)


def simplec_pairs():
    """Training pairs of the learned guess as the SIMPLEC time step
    generates them (mach2d.learned.simplec_pairs), on a 34x18 nozzle: the
    time step composed from the compiled reference's routines, the p'
    system of every step's last pressure-correction iteration, and its
    solution by the reference's class_solver_msi9d (zero-corner embedding)
    to a 1e-10 drop, written in the M2DT format."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import time_step_oracle as TS
    from mach2d import dataset as D
    from mach2d import learned as Lr
    nx, ny, steps = 34, 18, 12
    pf = os.path.join(HERE, "parametersNozzleEulerPTC.txt")
    fseed, area_c, mach = Lr.variant_params(0, 0)
    prm = dict(Lr.NOZZLE)
    s = P.time_step_fields(nx, ny, fseed, prm, mach_in=mach, area_c=area_c)
    dt, n = 5e-6, nx * ny
    path = os.path.join(HERE, "simplec_pairs_34x18.m2dt")
    with D.Writer(path, nx, ny, 5, 1e-10, D.SOURCE_SIMPLEC) as W:
        for k in range(steps):
            s, dt, r = TS.step(O, s, nx, ny, prm, dt, ref=True, paramfile=pf)
            a5 = np.asarray(s["ap"]).reshape(5, n)
            b = np.asarray(s["bp"]).reshape(n)
            a9 = np.zeros((9, n))
            a9[[1, 3, 4, 5, 7]] = a5
            x = O.ref_solver_msi9d(a9, b, np.zeros(n), nx, ny, 20000, 1e-10)
            rel = np.linalg.norm(O.residual9(a9, x, b, nx, ny)) / np.linalg.norm(b)
            W.add(a5, b, x, fseed, k, r["itp"], rel)
    print("%-18s %8.1f KB" % (os.path.basename(path), os.path.getsize(path) / 1024))


if __name__ == "__main__":
    O.build()
    if not O.ref_available():
        sys.exit("reference library not built (no reference tree?)")
    msi_small()
    cfg1()
    tdma()
    norms()
    embed5()
    solver_msi9d()
    coef()
    grid()
    simplec_pairs()
