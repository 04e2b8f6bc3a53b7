"""The internal-flow pressure-correction loop (src/mach2d_main.f90:257-309)
composed from the oracle's routines (or, with `ref`, the compiled
reference's), on host arrays: the checker of m2d_intflow_pl_loop."""
import numpy as np

from mach2d import problems as P

IN = ("x", "xp", "xe", "ye", "xk", "yk", "rp", "re", "rn", "Jp",
      "due", "dve", "dun", "dvn", "de", "dn", "g", "roa", "u", "v", "plin")
INOUT = ("ue", "ve", "un", "vn", "Uce", "Vcn", "roe", "ron", "ro", "p", "pl")


def loop_fields(nx, ny, seed, coord=1):
    """Seeded inputs: flow_fields plus SIMPLEC face coefficients, and the
    p' system arrays as the reference allocates them (ap = 0, ap(:,3) = 1,
    bp = 0)."""
    f = P.flow_fields(nx, ny, seed=seed, coord=coord)
    rng = np.random.default_rng(seed + 17)
    for k in ("due", "dve", "dun", "dvn"):
        f[k] = rng.uniform(-1e-3, 1e-3, size=nx * ny)
    ap = np.zeros((5, nx * ny))
    ap[2] = 1.0
    f["ap"] = ap
    f["bp"] = np.zeros(nx * ny)
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in f.items()}


def run(O, f, nx, ny, modvis, dt, beta, p_avg, tolm, itpmax, nitm_p, tol_p, ref=False):
    """Returns (state dict, itp, normpl); f is not modified."""
    coef = O.ref_coef if ref else O.coef
    lu, fb = (O.ref_lu2d5, O.ref_fb2d5) if ref else (O.lu2d5, O.fb2d5)
    s = {k: np.array(f[k], copy=True) for k in INOUT + ("ap", "bp")}
    fld = dict(f)
    s["pl"][:] = 0.0
    normpl = 0.0
    itp = 0
    for itp in range(1, itpmax + 1):
        fld.update(s)
        s["ap"] = coef("p_coefficients", nx, ny, {"dt": dt}, fld, out={"a": s["ap"]})["a"]
        fld.update(s)
        r = coef("intflow_set_bcpl", nx, ny, {}, fld, out={"a": s["ap"], "b": s["bp"]})
        s["ap"], s["bp"] = r["a"], r["b"]
        fld.update(s)
        s["bp"] = coef("p_source", nx, ny, {"dt": dt}, fld, out={"b": s["bp"]})["b"]
        dl, du = lu(s["ap"], nx, ny)
        res = fb(s["ap"], dl, du, nx, ny, s["pl"], tol_p, nitm_p, s["bp"])
        s["pl"] = np.array(res[0] if isinstance(res, tuple) else res, copy=True)
        fld.update(s)
        out = {k: s[k] for k in ("ue", "ve", "un", "vn", "Uce", "Vcn")}
        out = coef("velocity_correction_at_faces_with_pl", nx, ny, {}, fld, out=out)
        s.update(out)
        fld.update(s)
        out = {k: s[k] for k in ("ue", "ve", "un", "vn", "Uce", "Vcn")}
        out = coef("intflow_velocities_at_boundary_faces", nx, ny, {"modvis": modvis}, fld,
                   out=out)
        s.update(out)
        fld.update(s)
        s["p"] = coef("pressure_correction_with_pl", nx, ny, {}, fld, out={"p": s["p"]})["p"]
        fld.update(s)
        s["ro"] = coef("density_correction_with_pl", nx, ny, {}, fld, out={"ro": s["ro"]})["ro"]
        fld.update(s)
        r = coef("density_at_faces", nx, ny, {"beta": beta}, fld,
                 out={"roe": s["roe"], "ron": s["ron"]})
        s["roe"], s["ron"] = r["roe"], r["ron"]
        normpl = float(np.max(np.abs(s["pl"]))) / p_avg
        if normpl < tolm:
            break
    return s, itp, normpl
