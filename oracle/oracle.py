"""ctypes front end of the CPU restatement of T-Flows' Native_Mod.

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
benchmark's CPU-baseline leg, as the checker.  The GPU library never uses it.

Each wrapper mirrors one reference routine (Sources/Process/Cpu/Native_Mod/):
  cg / bicg        Cg.f90:2-257 / Bicg.f90:2-269
  nrms             Normalized_Root_Mean_Square.f90:2-65
  residual         Residual_Vector.f90:2-44
  spmv             the product loop of Cg.f90:177-185
  prec_form/solve  Prec_Form.f90:2-93 / Prec_Solve.f90:2-107
  krylov           textbook BiCG / CGS / BiCGSTAB (Templates, Barrett et al.
                   1994) with Cg's prologue and conventions -- the methods the
                   reference's PETSc path offers (Petsc_Interface.cpp:416-436)
Index arrays keep the reference's 1-based values; vectors are plain arrays of
length nt (Fortran element i is Python element i-1).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PREC = {"none": 0, "jacobi": 1, "icc": 2}
CRIT = {"nrms": 0, "relres": 1}
KRYLOV = {"bicg_textbook": 1, "cgs": 2, "bcgs": 3, "bicgstab": 3}


class _Mat(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nt", ctypes.c_int),
                ("row", ctypes.c_void_p), ("col", ctypes.c_void_p),
                ("dia", ctypes.c_void_p), ("val", ctypes.c_void_p)]


_libs = {}


def lib(omp=False):
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name not in _libs:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I, D = ctypes.c_int, ctypes.c_double
        L.or_nrms.restype = D
        L.or_nrms.argtypes = [P, I, P, P, P]
        L.or_residual.argtypes = [P, I, P, P, P]
        L.or_spmv.argtypes = [P, I, P, P]
        L.or_prec_form.argtypes = [P, I, P, P, I]
        L.or_prec_solve.argtypes = [P, I, P, P, P, P, I]
        for f in (L.or_cg, L.or_bicg):
            f.restype = I
            f.argtypes = [P, P, P, I, I, P, D, P, P]
        L.or_cg_ex.restype = I
        L.or_cg_ex.argtypes = [P, P, P, I, I, P, D, P, P, I]
        L.or_set_sum_order.argtypes = [I]
        L.or_set_progress.argtypes = [I]
        L.or_set_parts.argtypes = [P]
        L.or_lap32_form.argtypes = [P, I, D, P]
        L.or_lap32_spmv.argtypes = [P, I, P, P, P]
        L.or_ir_cg.restype = I
        L.or_ir_cg.argtypes = [P, P, P, I, I, P, D, P, P, I, D, I, I, I, P]
        L.or_krylov.restype = I
        L.or_krylov.argtypes = [P, P, P, I, I, I, P, D, P, P, I, P]
        L.or_transpose_positions.restype = I
        L.or_transpose_positions.argtypes = [P, P]
        L.or_pressure_assemble.argtypes = [I, I, I] + [P] * 14 + [I]
        L.or_vel_max.restype = D
        L.or_vel_max.argtypes = [I, P, P, P]
        L.or_momentum_assemble.argtypes = [I, I, I, P, P, P, P, P, P, P, P]
        L.or_adv_face.restype = D
        L.or_adv_face.argtypes = [I, D, D, P, P, P, P, P, D]
        L.or_rhie_chow_fluxes.argtypes = [I, I, P, P, P, P]
        _libs[name] = L
    return _libs[name]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Matrix:
    """Matrix_Type restated: row/col/dia (1-based int32) and val (FP64).

    The solve scales `val` in place and scales it back afterwards, exactly as
    the reference does (Cg.f90:77-84, 245-250)."""

    def __init__(self, row, col, dia, val, ni=None):
        self.row = np.ascontiguousarray(row, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.dia = np.ascontiguousarray(dia, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.nt = len(self.row) - 1
        self.ni = self.nt if ni is None else int(ni)
        self._s = _Mat(self.ni, self.nt, self.row.ctypes.data, self.col.ctypes.data,
                       self.dia.ctypes.data, self.val.ctypes.data)

    @property
    def ref(self):
        return ctypes.byref(self._s)

    def copy(self):
        return Matrix(self.row.copy(), self.col.copy(), self.dia.copy(),
                      self.val.copy(), self.ni)


def _f64(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.shape == (n,)
    return a


def nrms(A, r, x, norm=None):
    r, x = _f64(r, A.nt), _f64(x, A.nt)
    nv = ctypes.c_double(norm) if norm is not None else None
    return lib().or_nrms(A.ref, A.ni, _p(r), _p(x), ctypes.byref(nv) if nv is not None else None)


def residual(A, b, x):
    b, x = _f64(b, A.nt), _f64(x, A.nt)
    r = np.zeros(A.nt)
    lib().or_residual(A.ref, A.ni, _p(r), _p(b), _p(x))
    return r


def spmv(A, p, omp=False):
    p = _f64(p, A.nt)
    q = np.zeros(A.nt)
    lib(omp).or_spmv(A.ref, A.ni, _p(p), _p(q))
    return q


class _Parts:
    """Sets the ranks of the rows (or_set_parts) for the duration of a call;
    parts=None leaves one rank."""

    def __init__(self, L, A, parts):
        self.L = L
        self.parts = None if parts is None else np.ascontiguousarray(parts, dtype=np.int32)
        if self.parts is not None:
            assert self.parts.shape == (A.nt,)

    def __enter__(self):
        if self.parts is not None:
            self.L.or_set_parts(self.parts.ctypes.data)

    def __exit__(self, *exc):
        self.L.or_set_parts(None)


def prec_form(A, prec, parts=None):
    """Prec_Form.f90:47-89; parts (rank of each row): the per-rank DIC."""
    d = np.zeros(A.nt)
    dinv = np.zeros(A.nt)
    with _Parts(lib(), A, parts):
        lib().or_prec_form(A.ref, A.ni, _p(d), _p(dinv), PREC[prec])
    return d, dinv


def prec_solve(A, d, dinv, b, prec, parts=None):
    """Prec_Solve.f90:44-103; parts as for prec_form."""
    d, dinv, b = _f64(d, A.nt), _f64(dinv, A.nt), _f64(b, A.nt)
    x = np.zeros(A.nt)
    with _Parts(lib(), A, parts):
        lib().or_prec_solve(A.ref, A.ni, _p(d), _p(dinv), _p(x), _p(b), PREC[prec])
    return x


class BreakdownError(ArithmeticError):
    """A textbook method broke down (zero rho or denominator, omega = 0)."""

    def __init__(self, niter, fin_res):
        super().__init__(f"Krylov breakdown at iteration {niter}")
        self.niter, self.fin_res = niter, fin_res


def solve(solver, prec, A, x, b, miter, tol, norm=None, criterion="nrms", omp=False,
          sum_order=0, omegas=None, progress=0, parts=None):
    """Solve_Native.f90:2-46 restated: dispatches 'cg' / 'bicg', plus the
    textbook methods 'bicg_textbook', 'cgs', 'bcgs' ('bicgstab').

    x and b are updated in place like the reference's intent(out)/(inout)
    arguments.  Returns (niter, fin_res).  sum_order=1 switches every global
    reduction to a blocked-pairwise order (not the reference's): it measures
    how far rounding alone moves the answer.  progress > 0 prints a line on
    stderr every that many CG iterations.  parts (the rank owning each row)
    emulates a multi-rank run on the global matrix: "icc" becomes the
    per-rank DIC, and sum_order=11 sums each rank's rows serially and adds
    the rank totals in rank order (contiguous ranks)."""
    assert x.dtype == np.float64 and x.flags.c_contiguous and x.shape == (A.nt,)
    assert b.dtype == np.float64 and b.flags.c_contiguous and b.shape == (A.nt,)
    niter = ctypes.c_int(0)
    fin = ctypes.c_double(0.0)
    nv = ctypes.c_double(norm) if norm is not None else None
    nref = ctypes.byref(nv) if nv is not None else None
    L = lib(omp)
    if sum_order == 11:
        assert parts is not None, "sum order 11 sums per rank: parts needed"
    L.or_set_sum_order(int(sum_order))
    L.or_set_progress(int(progress))
    try:
        with _Parts(L, A, parts):
            return _solve(L, solver, prec, A, x, b, miter, tol, nref, criterion, niter, fin,
                          omegas)
    finally:
        L.or_set_sum_order(0)
        L.or_set_progress(0)


def _solve(L, solver, prec, A, x, b, miter, tol, nref, criterion, niter, fin, omegas=None):
    if solver in KRYLOV:
        if prec not in ("jacobi", "none"):
            raise ValueError(f"{solver}: preconditioner {prec} not offered")
        om = _p(omegas) if omegas is not None else None
        rc = L.or_krylov(A.ref, _p(x), _p(b), KRYLOV[solver], PREC[prec], int(miter),
                         ctypes.byref(niter), float(tol), ctypes.byref(fin), nref,
                         CRIT[criterion], om)
        if rc == 1:
            raise BreakdownError(niter.value, fin.value)
        if rc == -3:
            raise ValueError("textbook BiCG needs a symmetric sparsity pattern")
        if rc != 0:
            raise MemoryError("oracle scratch allocation failed")
        return niter.value, fin.value
    if criterion != "nrms":
        if solver not in ("cg", "bicg"):
            raise ValueError(f"Unknown native solver: {solver}")
        rc = L.or_cg_ex(A.ref, _p(x), _p(b), PREC.get(prec, 0), int(miter),
                        ctypes.byref(niter), float(tol), ctypes.byref(fin), nref,
                        CRIT[criterion])
    elif solver == "cg":
        rc = L.or_cg(A.ref, _p(x), _p(b), PREC.get(prec, 0), int(miter),
                     ctypes.byref(niter), float(tol), ctypes.byref(fin), nref)
    elif solver == "bicg":
        rc = L.or_bicg(A.ref, _p(x), _p(b), PREC.get(prec, 0), int(miter),
                       ctypes.byref(niter), float(tol), ctypes.byref(fin), nref)
    else:
        raise ValueError(f"Unknown native solver: {solver}")
    if rc != 0:
        raise MemoryError("oracle scratch allocation failed")
    return niter.value, fin.value


def lap32_form(A, fn=1.0):
    """FP32 row-excess matrix of the FP32 sweeps: s_i = fl32(sum_j a_ij fn)
    in the diagonal slot, fl32(a_ij fn) elsewhere."""
    v = np.zeros(len(A.col), dtype=np.float32)
    lib().or_lap32_form(A.ref, A.ni, float(fn), _p(v))
    return v


def lap32_spmv(A, v32, x):
    """y = s_i x_i + sum_{j != i} a_ij (x_j - x_i) in FP32, CSR order."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros(A.nt, dtype=np.float32)
    lib().or_lap32_spmv(A.ref, A.ni, _p(np.ascontiguousarray(v32, dtype=np.float32)), _p(x), _p(y))
    return y


def ir_solve(prec, A, x, b, miter, tol, norm=None, criterion="nrms", delta=0.1, check=1,
             force=0, deflate=0):
    """FP32 CG (row-excess product) inside FP64 refinement, restated from the
    library's TFG_FP32_IR form: refinement decisions every `check`
    iterations, when the FP32 residual fell by `delta`, its estimate is below
    tol, `force` iterations passed, or at miter; `deflate` deflates the
    constant vector.  Returns (niter, fin_res, refinements)."""
    niter, fin, nref = ctypes.c_int(0), ctypes.c_double(0.0), ctypes.c_int(0)
    nv = ctypes.c_double(norm) if norm is not None else None
    rc = lib().or_ir_cg(A.ref, _p(x), _p(b), PREC.get(prec, 0), int(miter), ctypes.byref(niter),
                        float(tol), ctypes.byref(fin), ctypes.byref(nv) if nv is not None else None,
                        CRIT[criterion], float(delta), int(check), int(force), int(deflate),
                        ctypes.byref(nref))
    if rc == 1:
        raise BreakdownError(niter.value, fin.value)
    if rc != 0:
        raise MemoryError("oracle scratch allocation failed")
    return niter.value, fin.value, nref.value


def transpose_positions(A):
    """0-based position of the transposed entry of every entry (symmetric
    sparsity pattern), or None when the pattern is not symmetric."""
    t = np.zeros(len(A.col), dtype=np.int32)
    return t if lib().or_transpose_positions(A.ref, _p(t)) == 0 else None


def pressure_assemble(f, singular=False):
    """Serial pressure-system assembly (Rhie_And_Chow.f90:184-205,
    Compute_Pressure.f90:121-178, Set_Singular.f90:28-31) from a face set
    with the fields of tflows_cases.Faces; returns (val, b)."""
    val = np.zeros(len(f.col))
    b = np.zeros(len(f.dia))
    arrs = [np.ascontiguousarray(a) for a in (f.c1, f.c2, f.pos, f.dia, f.diag_bc, f.b_region,
                                               f.region, f.fc, f.fw, f.vm, f.vflux)]
    pp = np.ascontiguousarray(f.pp, dtype=np.float64) if f.pp is not None else None
    pp_cell1 = pp.ctypes.data + 8 * (f.n_bnd + 1) if pp is not None else None
    lib().or_pressure_assemble(len(f.dia), len(f.col), len(f.c1), *[_p(a) for a in arrs],
                               pp_cell1, _p(val), _p(b), 1 if singular else 0)
    return val, b


def vel_max(u, v, w):
    """Compute_Momentum.f90:201-204 (before Global % Max_Real)."""
    u, v, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w))
    return lib().or_vel_max(len(u), _p(u), _p(v), _p(w))


# advection schemes (Numerics_Mod/Advection_Scheme_Code.f90 names) and time
# schemes of the momentum assembly
ADV = {"central": 0, "upwind": 1, "luds": 2, "quick": 3, "smart": 4, "gamma": 5,
       "minmod": 6, "blended": 7, "superbee": 8, "avl_smart": 9}
TD = {"linear": 0, "parabolic": 1, "none": 2}


class _Momentum(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int) for k in ("component", "n_bnd", "adv_scheme", "blend_matrix",
                                              "td_scheme")] +
                [(k, ctypes.c_double) for k in ("blend", "dt", "urf", "p_drop")] +
                [(k, ctypes.c_void_p) for k in ("f", "fw", "fc", "v_flux", "vis_eff", "f_stress",
                                                "sx", "sy", "sz", "dx", "dy", "dz", "wall",
                                                "phi_n", "phi_o", "phi_oo", "gx", "gy", "gz",
                                                "uj_i", "uk_i", "density", "vol", "p_i",
                                                "cell_f", "st")])


MOMENTUM_EXT = ("phi_n", "gx", "gy", "gz", "uj_i", "uk_i", "density")


def adv_face(adv, f, flux, phi, grad, d, pmin=(-1e300, -1e300), pmax=(1e300, 1e300),
             blend=0.8):
    """Face value of one face between cells 1 and 2 by the momentum
    assembly's advection-scheme routine (Advection_Scheme.f90): f is
    Grid % f(s) (the weight of cell 1), flux the face's volume flux, phi the
    two cells' values, grad their gradients (6 values: cell 1 then cell 2),
    d the vector from cell 1 to cell 2, pmin / pmax the cells' bounds."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (phi, grad, d, pmin, pmax)]
    return lib().or_adv_face(ADV[adv], float(f), float(flux), *[_p(a) for a in arrs],
                             float(blend))


def momentum_assemble(f, mf, component=1, adv="central", blend_matrix=True, td="linear",
                      blend=0.8, dt=1e-2, urf=0.7, p_drop=0.5, stress=True, cell_force=True,
                      surface_tension=False):
    """Serial momentum assembly of one component (Compute_Momentum.f90
    :119-366 with Numerics_Mod) on a Faces set f and tflows_cases
    MomentumFields mf; returns (val, b, v_m)."""
    keep = []

    def ptr(a, ext=False):
        a = np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data + ((mf.n_bnd + 1) * a.itemsize if ext else 0)

    m = _Momentum()
    m.component, m.n_bnd, m.adv_scheme = component, mf.n_bnd, ADV[adv]
    m.blend_matrix, m.td_scheme = int(blend_matrix), TD[td]
    m.blend, m.dt, m.urf, m.p_drop = blend, dt, urf, p_drop
    for k in ("f", "fw", "fc", "v_flux", "vis_eff", "sx", "sy", "sz", "dx", "dy", "dz", "phi_o",
              "phi_oo", "vol", "p_i"):
        setattr(m, k, ptr(getattr(mf, k)))
    m.wall = ptr(np.ascontiguousarray(mf.wall, dtype=np.int32))
    for k in MOMENTUM_EXT:
        setattr(m, k, ptr(getattr(mf, k), ext=True))
    m.f_stress = ptr(mf.f_stress) if stress else None
    m.cell_f = ptr(mf.cell_f) if cell_force else None
    m.st = ptr(mf.st) if surface_tension else None
    val = np.zeros(len(f.col))
    b = np.zeros(len(f.dia))
    vm = np.zeros(len(f.dia))
    arrs = [np.ascontiguousarray(a) for a in (f.c1, f.c2, f.pos, f.dia)]
    lib().or_momentum_assemble(len(f.dia), len(f.col), len(f.c1), *[_p(a) for a in arrs],
                               ctypes.byref(m), _p(val), _p(b), _p(vm))
    return val, b, vm


RC_PTRS = ("f", "fc", "sx", "sy", "sz", "dx", "dy", "dz", "u", "v", "w", "p", "px", "py", "pz",
           "v_m")
RC_CHOI = ("u_o", "u_oo", "v_o", "v_oo", "w_o_", "w_oo_", "density", "flux_o", "flux_oo")
RC_GU = ("cell_fx", "cell_fy", "cell_fz", "face_fx", "face_fy", "face_fz")
RC_VOF = ("curv", "col", "col_x", "col_y", "col_z")


class _RhieChow(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_void_p) for k in RC_PTRS] + [("choi", ctypes.c_int)] +
                [(k, ctypes.c_void_p) for k in RC_CHOI] +
                [(k, ctypes.c_double) for k in ("dt", "w_o", "w_oo")] + [("gu", ctypes.c_int)] +
                [(k, ctypes.c_void_p) for k in RC_GU[:3] + RC_GU[3:]] +
                [("vof", ctypes.c_int), ("sigma", ctypes.c_double)] +
                [(k, ctypes.c_void_p) for k in RC_VOF])


def rhie_chow_fluxes(f, rf, v_flux, choi=False, dt=1e-2, w_o=1.0, w_oo=0.0, gu=False, vof=False,
                     sigma=0.07):
    """Serial Rhie-Chow fluxes (Rhie_And_Chow.f90:86-259) of the domain faces
    of Faces set f with tflows_cases.RhieChowFields rf; returns a copy of
    v_flux with those faces replaced."""
    keep = []

    def ptr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data
    r = _RhieChow()
    for k in RC_PTRS + RC_CHOI + RC_GU + RC_VOF:
        setattr(r, k, ptr(getattr(rf, k)))
    r.choi, r.gu, r.vof = int(choi), int(gu), int(vof)
    r.dt, r.w_o, r.w_oo, r.sigma = dt, w_o, w_oo, sigma
    out = np.array(v_flux, dtype=np.float64)
    arrs = [np.ascontiguousarray(a) for a in (f.c1, f.c2)]
    lib().or_rhie_chow_fluxes(len(f.dia), len(f.c1), *[_p(a) for a in arrs], ctypes.byref(r),
                              _p(out))
    return out
