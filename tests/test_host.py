"""Host logic without a GPU: workload generators, the C ABI's exported
symbols, the solver factory and argument handling."""
import os
import re

import numpy as np
import pytest

from mach2d import problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_generated_systems_honour_outside_invariant():
    for nx, ny, q in [(10, 6, 1.05), (130, 66, 1.05), (66, 34, 1.2)]:
        s = P.pcorr_system(nx, ny, seed=1, q=q)
        assert np.count_nonzero(s.a[P.outside_mask(nx, ny)]) == 0
        # interior rows: M-matrix with a_P = 1.001 * sum |a_nb|
        a = s.a.reshape(9, ny, nx)[:, 1:-1, 1:-1]
        nb = a[[0, 1, 2, 3, 5, 6, 7, 8]]
        assert np.all(nb <= 0)
        assert np.allclose(a[4], 1.001 * np.abs(nb).sum(axis=0), rtol=1e-15)
        # fictitious rows have zero rhs
        b = s.rhs.reshape(ny, nx)
        assert not b[0].any() and not b[-1].any() and not b[:, 0].any() and not b[:, -1].any()


def test_uniform_grid_has_no_cross_terms():
    """Cartesian grid: betae = betan = 0, so the corners vanish (before the
    M-matrix sign change every corner is +-0)."""
    nx, ny = 12, 9
    g = P.Grid(nx, ny, *np.meshgrid(np.arange(nx) * 0.1, np.arange(ny) * 0.2)[::1], q=1.0)
    m = P.metrics(g)
    assert np.allclose(m["betae"][1:-1, :-1], 0) and np.allclose(m["betan"][:-1, 1:-1], 0)
    a = P.diffusion_coefficients(nx, ny, m).reshape(9, ny, nx)
    assert np.allclose(a[[0, 2, 6, 8], 1:-1, 1:-1], 0)


def test_wall_aspect_ratio_target():
    q = P.ratio_for_aspect(130, 66, 1e3)
    asp = P.max_wall_aspect(P.ogive_grid(130, 66, q))
    assert 0.9e3 < asp < 1.1e3


def test_cfg4_sizes():
    sizes = P.cfg4_sizes()
    assert len(sizes) == 256
    for nx, ny in sizes:
        assert 66 <= nx <= 514 and 34 <= ny <= 258


def _declared_symbols():
    syms = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            syms += re.findall(r"\b(m2d_\w+)\s*\(", txt)
    return sorted(set(syms))


def test_library_exports_every_declared_symbol():
    """Loads libmach2d_hip.so (no GPU call) and resolves every entry point
    include/*.h declares."""
    import ctypes

    from mach2d import _lib
    L = _lib.load()
    declared = _declared_symbols()
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(L, name), name
        ctypes.cast(getattr(L, name), ctypes.c_void_p)
    assert set(declared) == set(_lib.EXPORTS)
    assert b"gfx950" in L.m2d_version()


def test_solvers_init_factory():
    from mach2d import solvers as S
    cfg = dict(solver="MSI", nitm_u=5, nitm_p=20, tol_u=1e-2, tol_p=1e-2)
    s5, s9 = S.solvers_init(10, 6, cfg)
    assert isinstance(s5, S.ClassSolverMSI5d) and isinstance(s9, S.ClassSolverMSI9d)
    assert (s9.nitm, s9.tol, s5.nitm, s5.tol) == (5, 1e-2, 20, 1e-2)
    cfg["solver"] = "TDMA"
    s5, s9 = S.solvers_init(10, 6, cfg)
    assert isinstance(s9, S.ClassSolverTDMA9d) and s9.nitm == 5
    cfg["solver"] = "CG"
    with pytest.raises(SystemExit):
        S.solvers_init(10, 6, cfg)


def test_coefficient_layouts_and_embedding():
    from mach2d import solvers as S
    n = 12
    a = np.arange(9 * n, dtype=np.float64).reshape(9, n)
    assert np.array_equal(S._coef(a, n, 9), a)
    fortran = np.asfortranarray(a.T)          # a(n, 9) as the reference holds it
    assert np.array_equal(S._coef(fortran, n, 9), a)
    with pytest.raises(ValueError):
        S._coef(a[:5], n, 9)
    a5 = np.arange(5 * n, dtype=np.float64).reshape(5, n)
    a9 = S.embed5(a5, n)
    assert np.array_equal(a9[[1, 3, 4, 5, 7]], a5) and not a9[[0, 2, 6, 8]].any()


def test_fortran_dropin_module_compiles(tmp_path):
    """The iso_c_binding drop-in module (same module and routine names as
    the reference) compiles with the image's Fortran compiler."""
    import shutil
    import subprocess
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("no Fortran compiler")
    for name in ("msi2d9_hip.f90", "msi2d5_hip.f90", "coefficients_hip.f90",
                 "intflow_bc_hip.f90", "grid_metrics_hip.f90"):
        src = os.path.join(ROOT, "mach-2d_amd", "fortran", name)
        r = subprocess.run([fc, "-c", "-module-dir", str(tmp_path), "-o",
                            str(tmp_path / (name + ".o")), src], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_fortran_coefficients_module_matches_reference_interface():
    """Every routine of the drop-in `coefficients` module has the reference's
    dummy-argument list (src/equations/mach2d_coef.f90), as the C table
    mach2d.assembly.ROUTINES states it."""
    import re

    from mach2d.assembly import ROUTINES
    text = open(os.path.join(ROOT, "mach-2d_amd", "fortran", "coefficients_hip.f90")).read()
    body = text[text.index("contains"):]
    refname = {k: "get_" + k for k in ROUTINES}
    refname["boundary_faces_flags"] = "initialize_boundary_faces_flags"
    refname["mass_conservation"] = "check_mass_conservation"
    for k, (scal, fields, outs) in ROUTINES.items():
        if k.startswith("intflow_"):
            continue   # module mod_intflow, covered below
        m = re.search(r"subroutine %s\((.*?)\)" % refname[k], body, re.S)
        assert m, k
        got = [a.strip() for a in m.group(1).replace("&", " ").split(",")]
        want = (["nx", "ny"] + [n for n, _ in scal] +
                [f if isinstance(f, str) else f[0] for f in fields] + [n for n, _ in outs])
        assert got == want, k


def test_assembly_tables_match_reference_signatures():
    """The argument lists the C ABI, the Fortran drop-in and the oracle use
    are the reference's: compared with the subroutine statements of
    src/equations/mach2d_coef.f90 where the reference tree is present, and
    the product's table with the oracle's in any case."""
    import re

    import oracle as O
    from mach2d.assembly import ROUTINES
    assert set(ROUTINES) == set(O.COEF)
    for k in ROUTINES:
        assert ROUTINES[k] == O.COEF[k], k
    path = os.path.join(os.environ.get("MACH2D_REFERENCE", "/root/reference"), "src",
                        "equations", "mach2d_coef.f90")
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    text = open(path).read()
    refname = {k: "get_" + k for k in ROUTINES}
    refname["boundary_faces_flags"] = "initialize_boundary_faces_flags"
    refname["mass_conservation"] = "check_mass_conservation"
    intflow = open(os.path.join(os.path.dirname(os.path.dirname(path)), "user",
                                "internal_flow", "mod_intflow.f90")).read()
    # the derived-type argument iflow crosses the C ABI as the fields the
    # routine reads; output names a / b stand for the reference's own
    iflow = {"intflow_set_bcpl": ["plin"], "intflow_set_bcT": ["Tin"],
             "intflow_boundary_T_from_H_conservation": ["Tin"]}
    outname = {"intflow_set_bcu": ("au", "b"), "intflow_set_bcv": ("av", "b"),
               "intflow_set_bcpl": ("ap", "bp"), "intflow_set_bcT": ("at", "bt")}
    for k, (scal, fields, outs) in ROUTINES.items():
        src, rn = (intflow, k) if k.startswith("intflow_") else (text, refname[k])
        m = re.search(r"subroutine\s+%s\s*\((.*?)\)" % rn, src, re.S | re.I)
        assert m, k
        raw = re.sub(r"!.*", "", m.group(1))
        got = [a.strip().lower() for a in raw.replace("&", " ").split(",") if a.strip()]
        if k in iflow:   # Twall (a scalar, listed with the scalars) and the arrays
            at = got.index("iflow")
            got = got[:at] + [w.lower() for w in iflow[k]] + got[at + 1:]
            if k == "intflow_set_bcT":
                got = got[:2] + ["twall"] + got[2:]
        if k == "intflow_time_cycle_update":
            # thermomodel and iflow cross as gamma = gm(T0), Rg, P0, T0 and
            # the inflow vectors of iflow (in/out)
            assert got == ["nx", "ny", "x", "xp", "u", "thermomodel", "iflow", "p"], got
            got = (["nx", "ny", "gamma", "rg", "p0", "t0", "x", "xp", "u"] +
                   ["uin", "vin", "pin", "tin", "mw", "plin", "pina", "p"])
        want = (["nx", "ny"] + [n for n, _ in scal] +
                [f if isinstance(f, str) else f[0] for f in fields] + [n for n, _ in outs])
        if k in outname:
            want = want[:-2] + list(outname[k])
        assert got == [w.lower() for w in want], (k, got)


def test_fortran_intflow_module_matches_reference_interface():
    """The boundary writers of intflow_bc_hip keep the dummy arguments of
    src/user/internal_flow/mod_intflow.f90 (iflow replaced by the fields
    read: plin; Twall and Tin), as the C table mach2d.assembly.ROUTINES
    states them."""
    import re

    from mach2d.assembly import ROUTINES
    text = open(os.path.join(ROOT, "mach-2d_amd", "fortran", "intflow_bc_hip.f90")).read()
    body = text[text.index("contains"):]
    names = [k for k in ROUTINES if k.startswith("intflow_")]
    assert names == ["intflow_set_bcu", "intflow_set_bcv", "intflow_set_bcpl", "intflow_set_bcT",
                     "intflow_velocities_at_boundary_faces", "intflow_boundary_simplec",
                     "intflow_extrapolate_u_v_to_fictitious",
                     "intflow_boundary_T_from_H_conservation", "intflow_time_cycle_update"]
    for k in names:
        scal, fields, outs = ROUTINES[k]
        m = re.search(r"subroutine %s_hip\((.*?)\)" % k, body, re.S)
        assert m, k
        got = [a.strip() for a in m.group(1).replace("&", " ").split(",")]
        want = (["nx", "ny"] + [n for n, _ in scal] +
                [f if isinstance(f, str) else f[0] for f in fields] + [n for n, _ in outs])
        want = [{"a": {"intflow_set_bcpl": "ap", "intflow_set_bcT": "at",
                       "intflow_set_bcv": "av"}.get(k, "au"),
                 "b": {"intflow_set_bcpl": "bp", "intflow_set_bcT": "bt"}.get(k, "b")}.get(w, w)
                for w in want]
        assert got == want, k


def test_fortran_msi2d5_module_keeps_reference_interface():
    """msi2d5_hip.f90 exports module msi2d5 with lu2d5 and fb2d5 taking the
    dummy arguments of src/solvers/msi2d5-v03.f90:42,247 in their order."""
    import re
    text = open(os.path.join(ROOT, "mach-2d_amd", "fortran", "msi2d5_hip.f90")).read()
    assert re.search(r"^module msi2d5\s*$", text, re.M)
    body = text[text.index("contains"):]
    sig = {m.group(1): [a.strip() for a in m.group(2).split(",")]
           for m in re.finditer(r"subroutine (\w+)\((.*?)\)", body)}
    assert sig["lu2d5"] == ["b", "dl", "du", "nij", "ni", "nj"]
    assert sig["fb2d5"] == ["b", "dl", "du", "nj", "ni", "nij", "var", "gamma", "nitm", "rhs"]


def test_generated_5point_systems():
    """pcorr_system5: 5 diagonals (S, W, P, E, N), no coupling across the
    grid edge, strictly dominant interior diagonal, identity corners."""
    s = P.pcorr_system5(40, 23, seed=2)
    a = s.a.reshape(5, 23, 40)
    assert s.a.shape == (5, 40 * 23)
    assert not a[0, 0].any() and not a[4, -1].any()
    assert not a[1, :, 0].any() and not a[3, :, -1].any()
    nb = np.abs(a[[0, 1, 3, 4], 1:-1, 1:-1]).sum(axis=0)
    assert (a[2, 1:-1, 1:-1] > nb).all()
    for j, i in ((0, 0), (0, -1), (-1, 0), (-1, -1)):
        assert a[2, j, i] == 1.0 and not a[[0, 1, 3, 4], j, i].any()


def _ifile(name):
    from mach2d.formats import ClassIfile
    path = os.path.join(ROOT, "tests", "golden", name)
    return path, ClassIfile().init(path, "&").load()


@pytest.mark.parametrize("name", ["parameters.txt", "parametersNozzleEulerPTC.txt"])
def test_ifile_reads_reference_parameter_files(name):
    """Known entries of the shipped parameter files, and solvers_init on
    them (mod_solvers.f90:17-87 reads solver, nitm_u, nitm_p, tol_u, tol_p)."""
    from mach2d import solvers as S
    _, f = _ifile(name)
    assert f.get_value("solver", str) == "MSI"
    assert (f.get_value("nitm_u", int), f.get_value("nitm_p", int)) == (5, 20)
    assert f.get_value("tol_p", float) == 1e-2
    s5, s9 = S.solvers_init(10, 6, f)
    assert (s5.nitm, s5.tol, s9.nitm, s9.tol) == (20, 1e-2, 5, 1e-2)
    with pytest.raises(SystemExit):
        f.get_value("no_such_variable", int)


@pytest.mark.parametrize("name", ["parameters.txt", "parametersNozzleEulerPTC.txt"])
def test_ifile_matches_compiled_reference(orc, name):
    """Every variable the parser finds reads the same (as a string, and as
    an integer or real(8) where the field is one) as class_ifile of the
    compiled reference on the same file."""
    if not orc.ref_available():
        pytest.skip("reference library not built")
    path, f = _ifile(name)
    names = [d[1].strip() for d in f.data]
    assert len(names) > 20
    for n in dict.fromkeys(names):
        assert f.get_value(n, str) == orc.ref_ifile_get(path, "&", n, str), n
        v = f.get_value(n, str).split()[0] if f.get_value(n, str).split() else ""
        if re.fullmatch(r"[+-]?\d+", v):
            assert f.get_value(n, int) == orc.ref_ifile_get(path, "&", n, int), n
        if re.fullmatch(r"[+-]?(\d+\.?\d*|\.\d+)([eEdD][+-]?\d+)?", v):
            assert f.get_value(n, float) == orc.ref_ifile_get(path, "&", n, float), n


def test_residual_rows_match_fortran_formatting(orc):
    """residual_header / residual_row against the Fortran runtime's output
    of the same edit descriptors (mach2d_main.f90:495-517)."""
    from mach2d.formats import residual_header, residual_row
    if not orc.ref_available():
        pytest.skip("reference library not built")
    cases = [(1, 1e-3, 0.5, 2.5e-7, -3.25e-12), (123456, 2.0, 0.0, 1.0, -0.0),
             (7, 1.23456789e-120, 9.99999999e99, -5e-100, 3.0e150),
             (42, -7.7777777e-5, 1.0e-9, 12345.678, 0.1)]
    for c in cases:
        row, head = orc.ref_format_residual(*c)
        assert residual_row(*c) == row, c
        assert residual_header() == head


def test_fortran_tstepper_module_compiles(orc, tmp_path):
    """The drop-in mod_tstepper_scarborough compiles against the reference's
    mod_class_ifile (module files of the reference build)."""
    import shutil
    import subprocess
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    moddir = os.path.join(ROOT, "oracle", "_ref", "mod")
    if not os.path.exists(fc) or not os.path.isdir(moddir):
        pytest.skip("no Fortran compiler or reference module files")
    src = os.path.join(ROOT, "mach-2d_amd", "fortran", "tstepper_scarborough_hip.f90")
    r = subprocess.run([fc, "-c", "-I", moddir, "-module-dir", str(tmp_path), "-o",
                        str(tmp_path / "t.o"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_time_step_entry_rejects_bad_arguments():
    """m2d_intflow_time_step validates its arguments before touching the
    GPU: missing handles, fields or parameters are M2D_ERR_ARG."""
    import ctypes as C

    from mach2d import _lib, simplec
    L = _lib.load()
    f = simplec.TimeStepFields()
    p = simplec.TimeStepParams(itmmax=1, nitm_u=5)
    dt = C.c_double(1e-5)
    r = simplec.TimeStepResult()
    assert L.m2d_intflow_time_step(None, None, C.byref(f), C.byref(p), C.byref(dt),
                                   C.byref(r)) == _lib.M2D_ERR_ARG
    assert L.m2d_intflow_time_step(None, None, None, None, None, None) == _lib.M2D_ERR_ARG
    assert dt.value == 1e-5


def test_tuning_entries_check_their_arguments():
    """The host-side setters added beside the reference's interface check
    their arguments without touching the GPU: the TDMA line solver takes
    modes 0 and 1 only (and the default is restored), the p' solve-cap
    setter and query reject a missing handle or output."""
    import ctypes as C

    from mach2d import _lib
    L = _lib.load()
    assert L.m2d_set_tdma_mode(2) != 0
    assert L.m2d_set_tdma_mode(-1) != 0
    assert L.m2d_set_tdma_mode(1) == 0
    assert L.m2d_set_tdma_mode(0) == 0
    a, b = C.c_int(), C.c_int()
    assert L.m2d_set_pl_solve_cap(None, 4) == _lib.M2D_ERR_ARG
    assert L.m2d_pl_loop_cap_info(None, C.byref(a), C.byref(b)) == _lib.M2D_ERR_ARG
    assert L.m2d_pl_loop_info(None, C.byref(a), C.byref(b)) == _lib.M2D_ERR_ARG


def test_m2dt_round_trip_and_checks(tmp_path):
    """The training-set format: records and metadata come back as written
    (float32), the header carries grid, diagonals, count, gamma and source,
    and a wrong magic, version or size is refused."""
    import struct

    from mach2d import dataset as D
    nx, ny, nd = 7, 5, 5
    N = nx * ny
    rng = np.random.default_rng(3)
    recs = [(rng.standard_normal((nd, N)), rng.standard_normal(N), rng.standard_normal(N))
            for _ in range(4)]
    path = str(tmp_path / "t.m2dt")
    with D.Writer(path, nx, ny, nd, 1e-10) as W:
        for q, (a, b, x) in enumerate(recs):
            W.add(a, b, x, seed=q, step=10 + q, itp=2, rel_residual=1e-11 * q)
    ds = D.Dataset(path)
    assert (len(ds), ds.nx, ds.ny, ds.ndiag, ds.gamma, ds.source) == (4, nx, ny, nd, 1e-10,
                                                                       D.SOURCE_SIMPLEC)
    for q, (a, b, x) in enumerate(recs):
        ga, gb, gx = ds[q]
        assert np.array_equal(ga, a.astype(np.float32)) and np.array_equal(gb, b.astype(np.float32))
        assert np.array_equal(gx, x.astype(np.float32))
        assert list(ds.meta[q]) == [q, 10 + q, 2, 1e-11 * q]
    raw = bytearray(open(path, "rb").read())
    for patch, msg in ((lambda r: r.__setitem__(slice(0, 4), b"XXXX"), "not an M2DT"),
                       (lambda r: r.__setitem__(slice(4, 8), struct.pack("<I", 9)), "version"),
                       (lambda r: r.extend(b"\0" * 8), "size")):
        bad = bytearray(raw)
        patch(bad)
        p2 = str(tmp_path / "bad.m2dt")
        open(p2, "wb").write(bytes(bad))
        with pytest.raises(ValueError, match=msg):
            D.Dataset(p2)


def test_simplec_pairs_fixture_is_solved(orc):
    """The committed training pairs (SIMPLEC p' systems of the reference's
    time step, solved by its class_solver_msi9d): each target's residual on
    its own system is below 1e-10 of ||b|| (float32 storage aside: checked
    in FP64 from the stored values against the metadata's drop)."""
    from mach2d import dataset as D
    path = os.path.join(os.path.dirname(__file__), "golden", "simplec_pairs_34x18.m2dt")
    ds = D.Dataset(path)
    assert len(ds) == 12 and ds.ndiag == 5 and ds.gamma == 1e-10
    for q in range(len(ds)):
        a5, b, x = (np.asarray(v, dtype=np.float64) for v in ds[q])
        a9 = np.zeros((9, ds.N))
        a9[[1, 3, 4, 5, 7]] = a5
        assert ds.meta[q][3] < 1e-10
        # the float32-rounded target still solves the float32-rounded system
        # far better than the zero guess
        r = np.linalg.norm(orc.residual9(a9, x, b, ds.nx, ds.ny)) / np.linalg.norm(b)
        assert r < 1e-5, (q, r)
