"""One internal-flow time step (src/mach2d_main.f90:94-404, constant
thermophysical properties) composed from the oracle's routines (or, with
`ref`, the compiled reference's), on host arrays: the checker of
m2d_intflow_time_step."""
import numpy as np

import pl_loop_oracle as PL
from mach2d import problems as P

PARAMS = dict(modvis=0, coord=1, solver9d=0, itmmax=1, itpmax=20, nitm_u=5, nitm_p=20,
              tol_u=1e-2, tol_p=1e-2, tolm=1e-2, beta=0.0, Rg=287.0, gamma=1.4, P0=1.2e5,
              T0=310.0, Cpref=1004.5, Href=1004.5 * 310.0, Twall=-1.0, h0=0.01, mincc=0.99,
              maxcc=0.9999)


nozzle_grid = P.nozzle_grid


def step_fields(nx, ny, seed, coord=1, u_in=10.0, mach_in=None):
    return P.time_step_fields(nx, ny, seed, PARAMS, coord=coord, u_in=u_in, mach_in=mach_in)


def step(O, f, nx, ny, prm, dt, ref=False, paramfile=None):
    """Returns (new fields, next dt, result dict); f is not modified.  With
    ref, `paramfile` gives h0, mincc, maxcc to the reference time stepper
    (the oracle takes prm's)."""
    coef = O.ref_coef if ref else O.coef
    s = {k: np.array(v, copy=True) for k, v in f.items()}
    N = nx * ny
    sc = dict(prm, dt=dt)

    def run(name, outs):
        r = coef(name, nx, ny, sc, s, out={k: s[k] for k in outs})
        s.update({k: np.asarray(v).reshape(s[k].shape) for k, v in r.items()})

    def run_as(name, mapping):   # routine output name -> field name
        r = coef(name, nx, ny, sc, s, out={k: s[v] for k, v in mapping.items()})
        s.update({v: np.asarray(r[k]).reshape(s[v].shape) for k, v in mapping.items()})

    def solve9(a, b, x):
        r = _solve9(a, b, x)
        return r[0] if isinstance(r, tuple) else r

    def _solve9(a, b, x):
        if prm["solver9d"] == 1:
            return (O.ref_tdma2d9 if ref else O.tdma2d9)(s[a], s[b], s[x], nx, ny,
                                                         prm["nitm_u"])
        if ref:
            return O.ref_solver_msi9d(s[a], s[b], s[x], nx, ny, prm["nitm_u"], prm["tol_u"])
        return O.solve_msi9d(s[a], s[b], s[x], nx, ny, prm["nitm_u"], prm["tol_u"])

    nrel = O.ref_norm_l1_9d_relative if ref else O.norm_l1_9d_relative
    for a, b in (("roa", "ro"), ("Ta", "T"), ("pa", "p"), ("ua", "u"), ("va", "v"),
                 ("uea", "ue"), ("vea", "ve"), ("una", "un"), ("vna", "vn")):
        s[a] = s[b].copy()
    tc = ("uin", "vin", "pin", "Tin", "Mw", "plin", "pina", "p")
    if ref:
        r = O.ref_time_cycle_update(nx, ny, sc, s, {k: s[k] for k in tc})
        s.update({k: np.asarray(v).reshape(s[k].shape) for k, v in r.items()})
    else:
        run("intflow_time_cycle_update", tc)
    p_avg = O.get_average(nx, ny, s["p"], ref_=ref)
    s["g"] = 1.0 / (prm["Rg"] * s["T"])
    norm = 0.0
    itp = 0
    normpl = 0.0
    for _ in range(prm["itmmax"]):
        norm = 0.0
        run_as("u_coefficients", {"a": "au"})
        run_as("u_source", {"cup": "cup", "sup": "sup", "b": "bu"})
        run_as("intflow_set_bcu", {"a": "au", "b": "bu"})
        run_as("v_coefficients", {"a": "av"})
        run_as("v_source", {"cvp": "cvp", "svp": "svp", "b": "bv"})
        run_as("intflow_set_bcv", {"a": "av", "b": "bv"})
        run("internal_simplec_coefficients", ("due", "dve", "dun", "dvn", "de", "dn"))
        run("intflow_boundary_simplec", ("due", "dve", "dun", "dvn", "de", "dn"))
        s["u"] = np.asarray(solve9("au", "bu", "u"), dtype=float).copy()
        norm = nrel(s["au"], s["bu"], s["u"], nx, ny, norm)
        s["v"] = np.asarray(solve9("av", "bv", "v"), dtype=float).copy()
        norm = nrel(s["av"], s["bv"], s["v"], nx, ny, norm)
        run("velocities_at_internal_faces", ("ue", "ve", "un", "vn", "Uce", "Vcn"))
        run("intflow_velocities_at_boundary_faces", ("ue", "ve", "un", "vn", "Uce", "Vcn"))
        st, itp, normpl = PL.run(O, s, nx, ny, prm["modvis"], dt, prm["beta"], p_avg,
                                 prm["tolm"], prm["itpmax"], prm["nitm_p"], prm["tol_p"], ref=ref)
        s.update({k: np.asarray(v).reshape(s[k].shape) for k, v in st.items()})
        norm = O.norm_l1_5d(nx, ny, s["pl"], s["bp"], s["ap"], norm, ref_=ref)
        run("u_v_at_real_nodes_with_pl", ("u", "v"))
        run("intflow_extrapolate_u_v_to_fictitious", ("u", "v"))
    rmass = coef("mass_conservation", nx, ny, sc, s)["res"]
    if prm["modvis"] == 0:
        s["T"] = (prm["Href"] - (s["u"] ** 2 + s["v"] ** 2) / 2.0) / prm["Cpref"]
        run("intflow_boundary_T_from_H_conservation", ("T", "Tbe", "Tbw", "Tbn", "Tbs"))
    else:
        run_as("T_coefficients_and_source", {"a": "at", "b": "bt"})
        run_as("intflow_set_bcT", {"a": "at", "b": "bt"})
        s["T"] = np.asarray(solve9("at", "bt", "T"), dtype=float).copy()
        norm = nrel(s["at"], s["bt"], s["T"], nx, ny, norm)
    run("density_at_nodes", ("ro",))
    run("density_at_faces", ("roe", "ron"))
    if ref:
        dtn = O.ref_dt_scarborough(paramfile, dt, s["au"], s["av"], s["at"], s["ap"], nx, ny)
    else:
        dtn = O.dt_scarborough(dt, s["au"], s["av"], s["at"], s["ap"], nx, ny, prm["h0"],
                               prm["mincc"], prm["maxcc"])
    return s, dtn, {"norm": norm, "rmass": rmass, "normpl": normpl, "itp": itp}
