"""Synthetic 9-diagonal systems of the MACH-2D pressure-correction kind.

The reference assembles its p' system as a 5-point stencil; the workloads
here use the full 9-point curvilinear diffusion stencil instead, built from
the same grid metrics, so that every MSI diagonal is exercised.

Grid: the reference's g01 power-law ogive (src/user/external_flow/
mod_extflow_geometry.f90:178-292 with mach2d_input/geometry_g01.txt): the
south boundary is the body y = rb (x/lr)^lbd, the north boundary an ellipse
of semi-axes la, lb; points along both follow power laws of exponent 1.5.
Interior grid lines are straight segments from each south point to the
north point of the same index, with node spacing in geometric progression
of ratio q away from the wall (q chosen to reach a target near-wall aspect
ratio for the wall-resolved grids).

Metrics: src/grid/mod_grid.f90:304-558 (get_metrics), centroids by the
simple mean of the corners (mod_grid.f90:236-272, opt = 1), restated with
numpy.

Coefficients: the Navier-Stokes diffusion part of get_u_coefficients
(src/equations/mach2d_coef.f90:160-197) with vle = vln = re = rn = 1; then
off-diagonals are forced to -|a_nb| and a_P = 1.001 * sum |a_nb| (an
M-matrix, diagonally dominant).  Boundary rows follow intflow_set_bcpl
(src/user/internal_flow/mod_intflow.f90:410-457) mapped to 9 slots:
west P=1,E=1; east P=1,W=-1; south P=1,N=-1; north P=1,S=-1; corners P=1.
Every coupling towards a neighbour outside the grid is exactly zero.

Arrays are numpy float64, diagonal-major (9, nx*ny), point index
j*nx + i (0-based; the reference's np = nx*(j-1)+i).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# g01 parameters, mach2d_input/geometry_g01.txt
G01 = dict(la=8.382e-2, lb=5.080e-2, lr=7.620e-2, rb=1.270e-2, lbd=1.0, akn=1.5, aks=1.5)


@dataclass
class Grid:
    nx: int
    ny: int
    x: np.ndarray   # (ny, nx) coordinates of the north-east corner of volume (i, j)
    y: np.ndarray
    q: float


def gp_fractions(m: int, q: float) -> np.ndarray:
    """m+1 fractions 0..1 with interval lengths in geometric progression q."""
    k = np.arange(m + 1, dtype=np.float64)
    if abs(q - 1.0) < 1e-14:
        return k / m
    return (q ** k - 1.0) / (q ** m - 1.0)


def ogive_grid(nx: int, ny: int, q: float = 1.05, geom=None) -> Grid:
    g = dict(G01 if geom is None else geom)
    x = np.zeros((ny, nx))
    y = np.zeros((ny, nx))
    t = np.arange(nx - 1, dtype=np.float64) / (nx - 2)
    # south: body; north: ellipse (mod_extflow_geometry.f90:259-284)
    xs = g["lr"] * t ** g["aks"]
    ys = g["rb"] * (xs / g["lr"]) ** g["lbd"]
    xn = g["la"] * t ** g["akn"] - (g["la"] - g["lr"])
    yn = g["lb"] * np.sqrt(np.maximum(1.0 - ((xn - g["lr"]) / g["la"]) ** 2, 0.0))
    f = gp_fractions(ny - 2, q)              # rows 0 .. ny-2 of corners
    x[: ny - 1, : nx - 1] = xs[None, :] + f[:, None] * (xn - xs)[None, :]
    y[: ny - 1, : nx - 1] = ys[None, :] + f[:, None] * (yn - ys)[None, :]
    return Grid(nx, ny, x, y, q)


def ratio_for_aspect(nx: int, ny: int, aspect: float, geom=None) -> float:
    """Geometric ratio q giving a maximum near-wall cell aspect ratio
    (tangential / wall-normal size) of about `aspect` (bisection)."""
    lo, hi = 1.0 + 1e-9, 1.5
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        g = ogive_grid(nx, ny, mid, geom)
        if max_wall_aspect(g) < aspect:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def max_wall_aspect(g: Grid) -> float:
    x, y = g.x, g.y
    dxi = np.hypot(x[0, 2: g.nx - 1] - x[0, 1: g.nx - 2], y[0, 2: g.nx - 1] - y[0, 1: g.nx - 2])
    deta = np.hypot(x[1, 2: g.nx - 1] - x[0, 2: g.nx - 1], y[1, 2: g.nx - 1] - y[0, 2: g.nx - 1])
    return float(np.max(dxi / np.maximum(deta, 1e-300)))


def metrics(g: Grid) -> dict:
    """mod_grid.f90:304-558 restated (0-based (j, i) arrays of shape (ny, nx))."""
    nx, ny, x, y = g.nx, g.ny, g.x, g.y
    z = lambda: np.zeros((ny, nx))
    xp, yp = z(), z()
    # centroids of real volumes, simple mean (mod_grid.f90:258-271)
    xp[1:-1, 1:-1] = (x[1:-1, 1:-1] + x[1:-1, :-2] + x[:-2, :-2] + x[:-2, 1:-1]) / 4.0
    yp[1:-1, 1:-1] = (y[1:-1, 1:-1] + y[1:-1, :-2] + y[:-2, :-2] + y[:-2, 1:-1]) / 4.0
    xe, ye, xen, yen, xk, yk, xke, yke = z(), z(), z(), z(), z(), z(), z(), z()
    Jp, Je, Jn, alphae, gamman, betae, betan = z(), z(), z(), z(), z(), z(), z()
    # eta-derivatives at east faces, j = 2..ny-1, i = 1..nx-1 (1-based)
    xe[1:-1, :-1] = x[1:-1, :-1] - x[:-2, :-1]
    ye[1:-1, :-1] = y[1:-1, :-1] - y[:-2, :-1]
    alphae[1:-1, :-1] = xe[1:-1, :-1] ** 2 + ye[1:-1, :-1] ** 2
    # csi-derivatives at north faces, j = 1..ny-1, i = 2..nx-1
    xk[:-1, 1:-1] = x[:-1, 1:-1] - x[:-1, :-2]
    yk[:-1, 1:-1] = y[:-1, 1:-1] - y[:-1, :-2]
    gamman[:-1, 1:-1] = xk[:-1, 1:-1] ** 2 + yk[:-1, 1:-1] ** 2
    # csi-derivatives at inner east faces, i = 2..nx-2
    xke[1:-1, 1:-2] = xp[1:-1, 2:-1] - xp[1:-1, 1:-2]
    yke[1:-1, 1:-2] = yp[1:-1, 2:-1] - yp[1:-1, 1:-2]
    # west boundary face (stored at i = 1), one-sided
    fw = (x[1:-1, 0] + x[:-2, 0]) / 2.0
    fe = (x[1:-1, 1] + x[:-2, 1]) / 2.0
    xke[1:-1, 0] = -3.0 * fw + 4.0 * xp[1:-1, 1] - fe
    fw = (y[1:-1, 0] + y[:-2, 0]) / 2.0
    fe = (y[1:-1, 1] + y[:-2, 1]) / 2.0
    yke[1:-1, 0] = -3.0 * fw + 4.0 * yp[1:-1, 1] - fe
    # east boundary face (i = nx-1)
    fw = (x[1:-1, nx - 3] + x[:-2, nx - 3]) / 2.0
    fe = (x[1:-1, nx - 2] + x[:-2, nx - 2]) / 2.0
    xke[1:-1, nx - 2] = fw - 4.0 * xp[1:-1, nx - 2] + 3.0 * fe
    fw = (y[1:-1, nx - 3] + y[:-2, nx - 3]) / 2.0
    fe = (y[1:-1, nx - 2] + y[:-2, nx - 2]) / 2.0
    yke[1:-1, nx - 2] = fw - 4.0 * yp[1:-1, nx - 2] + 3.0 * fe
    # eta-derivatives at inner north faces, j = 2..ny-2
    xen[1:-2, 1:-1] = xp[2:-1, 1:-1] - xp[1:-2, 1:-1]
    yen[1:-2, 1:-1] = yp[2:-1, 1:-1] - yp[1:-2, 1:-1]
    # south boundary face (stored at j = 1)
    fs = (x[0, 1:-1] + x[0, :-2]) / 2.0
    fn = (x[1, 1:-1] + x[1, :-2]) / 2.0
    xen[0, 1:-1] = -3.0 * fs + 4.0 * xp[1, 1:-1] - fn
    fs = (y[0, 1:-1] + y[0, :-2]) / 2.0
    fn = (y[1, 1:-1] + y[1, :-2]) / 2.0
    yen[0, 1:-1] = -3.0 * fs + 4.0 * yp[1, 1:-1] - fn
    # north boundary face (j = ny-1)
    fs = (x[ny - 3, 1:-1] + x[ny - 3, :-2]) / 2.0
    fn = (x[ny - 2, 1:-1] + x[ny - 2, :-2]) / 2.0
    xen[ny - 2, 1:-1] = fs - 4.0 * xp[ny - 2, 1:-1] + 3.0 * fn
    fs = (y[ny - 3, 1:-1] + y[ny - 3, :-2]) / 2.0
    fn = (y[ny - 2, 1:-1] + y[ny - 2, :-2]) / 2.0
    yen[ny - 2, 1:-1] = fs - 4.0 * yp[ny - 2, 1:-1] + 3.0 * fn
    # beta, J at east faces (j = 2..ny-1, i = 1..nx-1) and north faces
    sl = (slice(1, -1), slice(0, -1))
    betae[sl] = xke[sl] * xe[sl] + yke[sl] * ye[sl]
    Je[sl] = 1.0 / (xke[sl] * ye[sl] - xe[sl] * yke[sl])
    sl = (slice(0, -1), slice(1, -1))
    betan[sl] = xk[sl] * xen[sl] + yk[sl] * yen[sl]
    Jn[sl] = 1.0 / (xk[sl] * yen[sl] - xen[sl] * yk[sl])
    # J at volume centres
    fw = (x[1:-1, :-2] + x[:-2, :-2]) / 2.0
    fe = (x[1:-1, 1:-1] + x[:-2, 1:-1]) / 2.0
    fs = (x[:-2, 1:-1] + x[:-2, :-2]) / 2.0
    fn = (x[1:-1, 1:-1] + x[1:-1, :-2]) / 2.0
    xkp, xep = fe - fw, fn - fs
    fw = (y[1:-1, :-2] + y[:-2, :-2]) / 2.0
    fe = (y[1:-1, 1:-1] + y[:-2, 1:-1]) / 2.0
    fs = (y[:-2, 1:-1] + y[:-2, :-2]) / 2.0
    fn = (y[1:-1, 1:-1] + y[1:-1, :-2]) / 2.0
    ykp, yep = fe - fw, fn - fs
    Jp[1:-1, 1:-1] = 1.0 / (xkp * yep - xep * ykp)
    Je, Jn, Jp = np.abs(Je), np.abs(Jn), np.abs(Jp)
    return dict(xp=xp, yp=yp, xe=xe, ye=ye, xen=xen, yen=yen, xk=xk, yk=yk, xke=xke, yke=yke,
                Jp=Jp, Je=Je, Jn=Jn, alphae=alphae, gamman=gamman, betae=betae, betan=betan)


def radii(g: Grid, yp: np.ndarray) -> dict:
    """Axisymmetric radii of a grid (the formulas of get_radius,
    mod_grid.f90:562-723, coord = 1): re, rn at face centres, rp at volume
    centres with the fictitious volumes reflected about the boundary faces.
    (ny, nx) arrays; yp: centroid y, (ny, nx)."""
    ny, nx = g.y.shape
    y = g.y
    re = np.ones((ny, nx))
    rn = np.ones((ny, nx))
    re[1:ny - 1, 0:nx - 1] = (y[1:ny - 1, 0:nx - 1] + y[0:ny - 2, 0:nx - 1]) / 2.0
    rn[0:ny - 1, 1:nx - 1] = (y[0:ny - 1, 1:nx - 1] + y[0:ny - 1, 0:nx - 2]) / 2.0
    rp = np.array(yp, dtype=np.float64, copy=True)
    rp[0, 1:nx - 1] = rn[0, 1:nx - 1] - rp[1, 1:nx - 1] + rn[0, 1:nx - 1]
    rp[ny - 1, 1:nx - 1] = rn[ny - 2, 1:nx - 1] - rp[ny - 2, 1:nx - 1] + rn[ny - 2, 1:nx - 1]
    rp[1:ny - 1, 0] = re[1:ny - 1, 0] - rp[1:ny - 1, 1] + re[1:ny - 1, 0]
    rp[1:ny - 1, nx - 1] = re[1:ny - 1, nx - 2] - rp[1:ny - 1, nx - 2] + re[1:ny - 1, nx - 2]
    rp[0, 0] = 4.0 * y[0, 0] - rp[0, 1] - rp[1, 0] - rp[1, 1]
    rp[0, nx - 1] = 4.0 * y[0, nx - 2] - rp[0, nx - 2] - rp[1, nx - 1] - rp[1, nx - 2]
    rp[ny - 1, 0] = 4.0 * y[ny - 2, 0] - rp[ny - 1, 1] - rp[ny - 2, 0] - rp[ny - 2, 1]
    rp[ny - 1, nx - 1] = (4.0 * y[ny - 2, nx - 2] - rp[ny - 1, nx - 2] - rp[ny - 2, nx - 1] -
                          rp[ny - 2, nx - 2])
    return {"re": re, "rn": rn, "rp": rp}


def diffusion_coefficients(nx: int, ny: int, m: dict, dominance: float = 1.001) -> np.ndarray:
    """9-point p' coefficients (9, nx*ny), see the module docstring."""
    a = np.zeros((9, ny, nx))
    Je, Jn = m["Je"], m["Jn"]
    ce = Je * m["betae"]      # vle*re*Je*betae with vle = re = 1
    cn = Jn * m["betan"]
    ge = Je * m["alphae"]
    gn = Jn * m["gamman"]
    J, I = slice(1, ny - 1), slice(1, nx - 1)
    Jm, Im = slice(0, ny - 2), slice(0, nx - 2)   # south / west neighbours
    # mach2d_coef.f90:160-193 (advection terms dropped)
    a[0, J, I] = (ce[J, Im] + cn[Jm, I]) / 4.0
    a[1, J, I] = -gn[Jm, I] + (ce[J, Im] - ce[J, I]) / 4.0
    a[2, J, I] = -(ce[J, I] + cn[Jm, I]) / 4.0
    a[3, J, I] = -ge[J, Im] + (cn[Jm, I] - cn[J, I]) / 4.0
    a[5, J, I] = -ge[J, I] + (-cn[Jm, I] + cn[J, I]) / 4.0
    a[6, J, I] = (-ce[J, Im] - cn[J, I]) / 4.0
    a[7, J, I] = -gn[J, I] - (ce[J, Im] - ce[J, I]) / 4.0
    a[8, J, I] = (ce[J, I] + cn[J, I]) / 4.0
    nb = [0, 1, 2, 3, 5, 6, 7, 8]
    a[nb, J, I] = -np.abs(a[nb, J, I])
    a[4, J, I] = dominance * np.sum(np.abs(a[nb, J, I]), axis=0)
    # boundary rows (intflow_set_bcpl pattern), corners identity
    a[:, :, 0] = 0.0
    a[4, :, 0] = 1.0
    a[5, 1:-1, 0] = 1.0                 # west: P + E
    a[:, :, nx - 1] = 0.0
    a[4, :, nx - 1] = 1.0
    a[3, 1:-1, nx - 1] = -1.0           # east: P - W
    a[:, 0, 1:-1] = 0.0
    a[4, 0, 1:-1] = 1.0
    a[7, 0, 1:-1] = -1.0                # south: P - N
    a[:, ny - 1, 1:-1] = 0.0
    a[4, ny - 1, 1:-1] = 1.0
    a[1, ny - 1, 1:-1] = -1.0           # north: P - S
    return a.reshape(9, nx * ny)


def outside_mask(nx: int, ny: int) -> np.ndarray:
    """(9, nx*ny) bool: slot couples to a neighbour outside the grid."""
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    w, e, s, n = i == 0, i == nx - 1, j == 0, j == ny - 1
    m = np.stack([s | w, s, s | e, w, np.zeros_like(w), e, n | w, n, n | e])
    return m.reshape(9, nx * ny)


@dataclass
class System:
    nx: int
    ny: int
    a: np.ndarray      # (9, N)
    rhs: np.ndarray    # (N,)
    q: float
    seed: int
    metrics: dict | None = None

    @property
    def N(self) -> int:
        return self.nx * self.ny


def rhs_field(nx: int, ny: int, kind: str, rng) -> np.ndarray:
    """Right-hand side on interior rows (fictitious rows 0), flat: 'uniform'
    on [-1,1], 'gauss' N(0,1), or 'smooth' (six random sine modes plus
    0.1 N(0,1), a mass-imbalance-like field)."""
    b = np.zeros((ny, nx))
    if kind == "uniform":
        b[1:-1, 1:-1] = rng.uniform(-1.0, 1.0, size=(ny - 2, nx - 2))
    elif kind == "gauss":
        b[1:-1, 1:-1] = rng.standard_normal(size=(ny - 2, nx - 2))
    elif kind == "smooth":
        yy, xx = np.mgrid[0:ny - 2, 0:nx - 2]
        f = np.zeros((ny - 2, nx - 2))
        for _ in range(6):
            kx, ky = rng.uniform(0.5, 6.0), rng.uniform(0.5, 4.0)
            p1, p2 = rng.uniform(0.0, 2.0 * np.pi, size=2)
            f += rng.standard_normal() * (np.sin(np.pi * kx * xx / (nx - 2) + p1) *
                                          np.sin(np.pi * ky * yy / (ny - 2) + p2))
        b[1:-1, 1:-1] = f + 0.1 * rng.standard_normal(size=f.shape)
    else:
        raise ValueError(kind)
    return b.reshape(-1)


def pcorr_system(nx: int, ny: int, seed: int = 0, q: float | None = None,
                 aspect: float | None = None, rhs: str = "uniform",
                 keep_metrics: bool = False) -> System:
    """A seeded 9-diagonal p' system on the ogive grid.  rhs is drawn on
    interior rows only (rhs_field kinds); fictitious rows get 0."""
    if q is None:
        q = ratio_for_aspect(nx, ny, aspect) if aspect else 1.05
    g = ogive_grid(nx, ny, q)
    m = metrics(g)
    a = diffusion_coefficients(nx, ny, m)
    b = rhs_field(nx, ny, rhs, np.random.default_rng(seed))
    return System(nx, ny, np.ascontiguousarray(a), b, q, seed,
                  m if keep_metrics else None)


def diffusion_coefficients5(nx: int, ny: int, m: dict, dominance: float = 1.001) -> np.ndarray:
    """5-point p' coefficients (5, nx*ny), slots S, W, P, E, N: the
    orthogonal diffusion terms of the 9-point form (the cross-derivative
    terms dropped), so A has the 5-diagonal structure of the reference's own
    p' system (mach2d_coef.f90:1207-1278); a_P = dominance * sum |a_nb|;
    boundary rows as in diffusion_coefficients, corners identity."""
    a = np.zeros((5, ny, nx))
    ge = m["Je"] * m["alphae"]
    gn = m["Jn"] * m["gamman"]
    J, I = slice(1, ny - 1), slice(1, nx - 1)
    Jm, Im = slice(0, ny - 2), slice(0, nx - 2)
    a[0, J, I] = -np.abs(gn[Jm, I])
    a[1, J, I] = -np.abs(ge[J, Im])
    a[3, J, I] = -np.abs(ge[J, I])
    a[4, J, I] = -np.abs(gn[J, I])
    a[2, J, I] = dominance * np.sum(np.abs(a[[0, 1, 3, 4], J, I]), axis=0)
    a[:, :, 0] = 0.0
    a[2, :, 0] = 1.0
    a[3, 1:-1, 0] = 1.0                 # west: P + E
    a[:, :, nx - 1] = 0.0
    a[2, :, nx - 1] = 1.0
    a[1, 1:-1, nx - 1] = -1.0           # east: P - W
    a[:, 0, 1:-1] = 0.0
    a[2, 0, 1:-1] = 1.0
    a[4, 0, 1:-1] = -1.0                # south: P - N
    a[:, ny - 1, 1:-1] = 0.0
    a[2, ny - 1, 1:-1] = 1.0
    a[0, ny - 1, 1:-1] = -1.0           # north: P - S
    return a.reshape(5, nx * ny)


def pcorr_system5(nx: int, ny: int, seed: int = 0, q: float | None = None,
                  aspect: float | None = None, rhs: str = "uniform") -> System:
    """A seeded 5-diagonal p' system (slots S, W, P, E, N) on the ogive grid;
    `a` has shape (5, N)."""
    if q is None:
        q = ratio_for_aspect(nx, ny, aspect) if aspect else 1.05
    m = metrics(ogive_grid(nx, ny, q))
    a = diffusion_coefficients5(nx, ny, m)
    b = rhs_field(nx, ny, rhs, np.random.default_rng(seed))
    return System(nx, ny, np.ascontiguousarray(a), b, q, seed)


def random_f3_system(nx: int, ny: int, seed: int = 0, dominance: float = 1.5) -> System:
    """Random diagonally dominant 9-diagonal system honouring the zero
    outside-coupling invariant (for tests on odd shapes)."""
    rng = np.random.default_rng(seed)
    a = -rng.uniform(0.05, 1.0, size=(9, nx * ny))
    a[outside_mask(nx, ny)] = 0.0
    nb = [0, 1, 2, 3, 5, 6, 7, 8]
    a[4] = dominance * np.sum(np.abs(a[nb]), axis=0) + 1e-3
    rhs = rng.uniform(-1.0, 1.0, size=nx * ny)
    return System(nx, ny, np.ascontiguousarray(a), rhs, 1.0, seed)


# Workload definitions -------------------------------------------------------
CFG1 = dict(nx=130, ny=66, q=1.05, rhs="uniform", seed=1, gamma=1e-10, nitm=500)
CFG3 = dict(nx=1026, ny=514, aspect=1e4, rhs="gauss", seed=3, gamma=1e-10)


def cfg4_sizes(count: int = 256, seed: int = 4):
    """Real sizes 64s x 32s with s log-uniform on [1, 8], plus fictitious rows."""
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(0.0, math.log(8.0), size=count))
    return [(int(round(64 * v)) + 2, int(round(32 * v)) + 2) for v in s]


def cfg4_system(q: int, nx: int, ny: int) -> System:
    """System q of the uneven batch: its own seeded wall stretching (ratio in
    [1.02, 1.08)) and seeded Gaussian right-hand side."""
    return pcorr_system(nx, ny, seed=100 + q, q=1.02 + 0.06 * ((q * 0.618) % 1.0), rhs="gauss")


def cfg4_systems(count: int = 256):
    """The uneven batch: cfg4_sizes(count) with cfg4_system per member."""
    return [cfg4_system(q, nx, ny) for q, (nx, ny) in enumerate(cfg4_sizes(count))]


def flow_fields(nx: int, ny: int, seed: int = 0, q: float = 1.05, coord: int = 1,
                cartesian: bool = False) -> dict:
    """Seeded inputs of the coefficient assembly (mach2d_coef.f90) on the
    ogive grid (or, with cartesian=True, a uniform orthogonal grid), as flat
    N-vectors in the reference layout: metrics, radii (r = y when coord = 1,
    else 1), densities, viscosities, pressure, velocities and contravariant
    face velocities of both signs."""
    rng = np.random.default_rng(seed)
    if cartesian:
        xs = np.linspace(0.0, 1.0, nx - 1)
        ys = np.linspace(0.5, 1.5, ny - 1)
        x = np.zeros((ny, nx))
        y = np.zeros((ny, nx))
        x[: ny - 1, : nx - 1] = xs[None, :]
        y[: ny - 1, : nx - 1] = ys[:, None]
        g = Grid(nx, ny, x, y, 1.0)
    else:
        g = ogive_grid(nx, ny, q)
    m = metrics(g)
    f = {k: np.ascontiguousarray(v.reshape(-1)) for k, v in m.items()}
    N = nx * ny

    def pos(lo, hi):
        return rng.uniform(lo, hi, size=N)

    def signed(scale):
        return rng.uniform(-scale, scale, size=N)

    # radii at centres and faces: the local y when axisymmetric
    yp = np.abs(f["yp"]) + 0.05
    f["rp"] = yp if coord == 1 else np.ones(N)
    f["re"] = (yp * pos(0.95, 1.05)) if coord == 1 else np.ones(N)
    f["rn"] = (yp * pos(0.95, 1.05)) if coord == 1 else np.ones(N)
    for k in ("ro", "roa", "roe", "ron"):
        f[k] = pos(0.8, 1.2)
    for k in ("vle", "vln", "vlp"):
        f[k] = pos(1e-3, 2e-3)
    f["p"] = pos(0.9e5, 1.1e5)
    f["pl"] = signed(50.0)
    f["g"] = pos(1e-5, 1.3e-5)
    f["de"] = pos(1e-4, 1e-3)
    f["dn"] = pos(1e-4, 1e-3)
    for k in ("u", "v", "ua", "va"):
        f[k] = signed(100.0)
    f["Uce"] = signed(1.0)
    f["Vcn"] = signed(1.0)
    f["T"] = pos(280.0, 320.0)
    f["Ta"] = pos(280.0, 320.0)
    f["cp"] = pos(1000.0, 1010.0)
    f["pa"] = pos(0.9e5, 1.1e5)
    f["ke"] = pos(0.02, 0.03)
    f["kn"] = pos(0.02, 0.03)
    for k in ("ue", "ve", "un", "vn"):
        f[k] = signed(100.0)
    # boundary-face flags (1 on faces of the domain boundary, else 0)
    fbe = np.zeros((ny, nx))
    fbn = np.zeros((ny, nx))
    fbe[:, 0] = fbe[:, nx - 2] = 1.0
    fbn[0, :] = fbn[ny - 2, :] = 1.0
    f["fbe"] = fbe.reshape(-1)
    f["fbn"] = fbn.reshape(-1)
    # internal-flow boundary data (mod_intflow.f90): x of the north-east
    # corner of each volume, inflow pressure correction and temperature per
    # row (ny values)
    f["x"] = np.ascontiguousarray(g.x.reshape(-1))
    f["plin"] = rng.uniform(-50.0, 50.0, size=ny)
    f["Tin"] = rng.uniform(280.0, 320.0, size=ny)
    return f


def nozzle_grid(nx, ny, area_c=2.2):
    """Converging-diverging nozzle, x in [0, 3], wall radius sqrt(A(x)/pi)
    with A(x) = 1 + area_c (x - 1.5)^2 (area_c = 2.2: the workload's
    nozzle), the axis at y = 0; corners as the reference stores them
    (north-east corner of each volume)."""
    xs = np.linspace(0.0, 3.0, nx - 1)
    rw = np.sqrt((1.0 + area_c * (xs - 1.5) ** 2) / np.pi)
    x = np.zeros((ny, nx))
    y = np.zeros((ny, nx))
    fr = np.linspace(0.0, 1.0, ny - 1)
    x[: ny - 1, : nx - 1] = xs[None, :]
    y[: ny - 1, : nx - 1] = fr[:, None] * rw[None, :]
    return Grid(nx, ny, x, y, 1.0)


# The internal-flow time-step workload: the 258x130 nozzle, Euler, inlet
# Mach 0.2 (the initial velocity u_in A_in / A(x) is then supersonic past the
# throat, M 1.2), first time step 5e-6 s (the reference's own step diverges
# from this state within ten steps at 2e-5 s; at 5e-6 s it stays finite for
# 200 steps with MSI and with TDMA for u and v), 200 time steps.
CFG2 = dict(nx=258, ny=130, mach_in=0.2, dt0=5e-6, steps=200)


def inlet_speed(mach, params):
    """Velocity of inlet Mach `mach` at the initial temperature T0 - 2 of
    time_step_fields."""
    return mach * math.sqrt(params["gamma"] * params["Rg"] * (params["T0"] - 2.0))


def time_step_fields(nx, ny, seed, params, coord=1, u_in=10.0, mach_in=None, area_c=2.2):
    """Seeded inputs of one internal-flow time step (m2d_intflow_time_step)
    on the nozzle, `params` giving T0, P0, Rg and gamma: a flow at rest
    temperature and pressure (velocity u_in A_in / A(x) plus 1 % noise, the
    matching face velocities and densities; mach_in, when given, sets u_in
    to that inlet Mach number), the radii of get_radius, the 9-diagonal
    systems as the reference allocates them (a = 0, unit diagonal), the
    5-diagonal one likewise, and the inflow vectors."""
    if mach_in is not None:
        u_in = inlet_speed(mach_in, params)
    g = nozzle_grid(nx, ny, area_c)
    m = metrics(g)
    f = {k: np.ascontiguousarray(v.reshape(-1)) for k, v in m.items()}
    N = nx * ny
    rng = np.random.default_rng(seed)
    xf, yf = g.x.reshape(-1), g.y.reshape(-1)
    f["x"] = np.ascontiguousarray(xf)
    if coord == 1:
        r = radii(g, m["yp"])
        f["re"], f["rn"], f["rp"] = (np.ascontiguousarray(r[k].reshape(-1))
                                     for k in ("re", "rn", "rp"))
    else:
        f["re"], f["rn"], f["rp"] = np.ones(N), np.ones(N), np.ones(N)
    area = 1.0 + area_c * (f["xp"] - 1.5) ** 2
    T0, P0, Rg = params["T0"], params["P0"], params["Rg"]
    f["u"] = u_in * (1.0 + area_c * 2.25) / area * (1.0 + 0.01 * rng.uniform(-1, 1, N))
    f["v"] = np.zeros(N)
    f["T"] = np.full(N, T0 - 2.0)
    f["p"] = np.full(N, 0.99 * P0) * (1.0 + 1e-4 * rng.uniform(-1, 1, N))
    f["ro"] = f["p"] / (Rg * f["T"])
    f["roe"], f["ron"] = f["ro"].copy(), f["ro"].copy()
    ug = f["u"].reshape(ny, nx)
    ue = np.zeros((ny, nx))
    ue[:, :-1] = 0.5 * (ug[:, :-1] + ug[:, 1:])
    f["ue"], f["un"] = ue.reshape(-1), f["u"].copy()
    f["ve"], f["vn"] = np.zeros(N), np.zeros(N)
    f["Uce"] = f["ue"] * f["ye"] - f["ve"] * f["xe"]
    f["Vcn"] = f["vn"] * f["xk"] - f["un"] * f["yk"]
    f["pl"] = np.zeros(N)
    f["g"] = 1.0 / (Rg * f["T"])
    ap = np.zeros((5, N))
    ap[2] = 1.0
    f["ap"], f["bp"] = ap, np.zeros(N)
    for k in ("vle", "vln", "vlp"):
        f[k] = np.zeros(N)
    f["cp"] = np.full(N, 1004.5)
    for k in ("ke", "kn"):
        f[k] = np.full(N, 0.026)
    for k in ("au", "av", "at"):
        a = np.zeros((9, N))
        a[4] = 1.0
        f[k] = a
    for k in ("due", "dve", "dun", "dvn"):
        f[k] = np.zeros(N)
    for k in ("bu", "bv", "bt", "cup", "sup", "cvp", "svp", "de", "dn", "ua", "va", "uea",
              "vea", "una", "vna", "pa", "Ta", "roa"):
        f[k] = np.zeros(N)
    f["pin"] = np.full(ny, 0.99 * P0)
    for k in ("uin", "vin", "Tin", "Mw", "plin", "pina"):
        f[k] = np.zeros(ny)
    f["Tin"] = np.full(ny, T0 - 2.0)
    for k in ("Tbe", "Tbw"):
        f[k] = np.zeros(ny)
    for k in ("Tbn", "Tbs"):
        f[k] = np.zeros(nx)
    f["fbe"] = np.zeros(N)
    f["fbn"] = np.zeros(N)
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in f.items()}
