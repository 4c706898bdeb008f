"""Host-side mirror of T-Flows' native solver interface over the tflows_gpu C ABI.

The reference call being replaced is
    Nat % Solve_Native(solver, prec, A, x, b, miter, niter, tol, fin_res, norm)
(Sources/Process/Cpu/Native_Mod/Solve_Native.f90:2-46).  `Native.solve_native`
keeps that argument list: it updates `x` in place and returns
(niter, fin_res).  An unknown solver name raises, as the reference stops with
Message % Error (Solve_Native.f90:38-44).

There is no CPU fallback: if libtflows_gpu.so is missing or no GPU is
present, every call raises.
"""
import ctypes
import weakref
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtflows_gpu.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "tflows_gpu.h")

TFG_OK, TFG_EINVAL, TFG_EHIP, TFG_ENOMEM, TFG_EBREAKDOWN, TFG_ENOTSUP = 0, 1, 2, 3, 4, 5
FP64, FP32_IR = 0, 1
DEFLATE_AUTO, DEFLATE_OFF, DEFLATE_ON = 0, 1, 2
CRIT_NRMS, CRIT_RELRES = 0, 1
PATH_AUTO, PATH_KERNELS, PATH_GRAPH = 0, 1, 2
ORDER_NATURAL, ORDER_RCM = 0, 1


class TFlowsGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tflows_gpu error {code}: {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("criterion", ctypes.c_int),
                ("path", ctypes.c_int), ("inner_max", ctypes.c_int),
                ("inner_tol", ctypes.c_double), ("timing", ctypes.c_int),
                ("true_residual", ctypes.c_int), ("textbook_bicg", ctypes.c_int),
                ("deflation", ctypes.c_int)]


class Report(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int),
                ("refinements", ctypes.c_int), ("true_relres", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("spmv_ms", ctypes.c_double),
                ("spmv_samples", ctypes.c_int), ("deflated", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


SUM, MIN, MAX = 0, 1, 2
# int allreduce(void* ctx, double* values, const int* ops, int count)
ALLREDUCE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_int), ctypes.c_int)
# int exchange(void* ctx, double* d_x)
EXCHANGE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)


def lib():
    """Load libtflows_gpu.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P, I, D, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
        L.tfg_matrix_create.argtypes = [I, I, I, P, P, P, ctypes.POINTER(P)]
        L.tfg_matrix_destroy.argtypes = [P]
        L.tfg_matrix_info.argtypes = [P, P, P, P, P, P, P, P, P, P, P, P]
        L.tfg_solve_native.argtypes = [P, ctypes.c_char_p, I, ctypes.c_char_p, I, P, P, P, I,
                                       P, D, P, P]
        L.tfg_solve_native_ex.argtypes = [P, ctypes.c_char_p, I, ctypes.c_char_p, I, P, P, P,
                                          I, P, D, P, P, I, P, P]
        L.tfg_spmv.argtypes = [P, P, P, P]
        L.tfg_spmv_fp32.argtypes = [P, P, D, P, P]
        L.tfg_prec_apply.argtypes = [P, ctypes.c_char_p, I, P, P, P]
        L.tfg_residual.argtypes = [P, P, P, P, P]
        L.tfg_spmv_bench.argtypes = [P, P, P, P, I, ctypes.POINTER(D)]
        L.tfg_spmv_tune.argtypes = [P, P, ctypes.POINTER(I), P]
        L.tfg_spmv_select.argtypes = [P, I]
        L.tfg_device_alloc.argtypes = [ctypes.POINTER(P), LL]
        L.tfg_device_free.argtypes = [P]
        L.tfg_memcpy_h2d.argtypes = [P, P, LL]
        L.tfg_memcpy_d2h.argtypes = [P, P, LL]
        L.tfg_synchronize.argtypes = [P]
        L.tfg_stream_copy.argtypes = [P, P, LL, I, ctypes.POINTER(D)]
        L.tfg_stream_bench.argtypes = [P, P, LL, I, I, ctypes.POINTER(D)]
        L.tfg_sync_bench.argtypes = [I, I, ctypes.POINTER(D), ctypes.POINTER(D)]
        L.tfg_matrix_set_comm.argtypes = [P, ALLREDUCE_FN, EXCHANGE_FN, P]
        L.tfg_faces_create.argtypes = [P, I, P, P, P, P, P, ctypes.POINTER(P)]
        L.tfg_faces_destroy.argtypes = [P]
        L.tfg_pressure_assemble.argtypes = [P, P, P, P, P, P, P, I]
        L.tfg_pressure_assemble_ex.argtypes = [P, P, P, P, P, P, P, P, I]
        L.tfg_vel_max.argtypes = [P, P, P, P, ctypes.POINTER(ctypes.c_double)]
        L.tfg_range.argtypes = [P, P, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]
        L.tfg_momentum_assemble.argtypes = [P, P, P, P, P]
        L.tfg_matrix_reorder.argtypes = [P, I]
        L.tfg_rhie_chow_fluxes.argtypes = [P, P, P]
        L.tfg_batch_create.argtypes = [I, P, P, P, P, P, ctypes.POINTER(P)]
        L.tfg_batch_destroy.argtypes = [P]
        L.tfg_batch_solve.argtypes = [P, ctypes.c_char_p, I, ctypes.c_char_p, I, P, P, P, I, P,
                                      D, P, P, I, P, P]
        L.tfg_batch_times.argtypes = [P, P]
        L.tfg_batch_info.argtypes = [P, P, P, P, P]
        L.tfg_batch_tuning.argtypes = [P, I, P, P, P, P, P]
        L.tfg_solve_native_noalloc.argtypes = [P, ctypes.c_char_p, I, ctypes.c_char_p, I, P, P,
                                               P, I, P, D, P, P, P, P]
        L.tfg_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def declared_symbols(header=HEADER):
    """Names of the functions include/tflows_gpu.h declares."""
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tfg_\w+)\s*\(", text, re.M)))


def _check(rc):
    if rc != TFG_OK:
        raise TFlowsGpuError(rc, lib().tfg_last_error().decode(errors="replace"))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class DeviceArray:
    """A device buffer owned by the library's allocator."""

    def __init__(self, nbytes_or_array, dtype=np.float64):
        if isinstance(nbytes_or_array, np.ndarray):
            a = np.ascontiguousarray(nbytes_or_array)
            self.dtype, self.shape = a.dtype, a.shape
            nbytes = a.nbytes
        else:
            self.dtype, self.shape = np.dtype(dtype), (int(nbytes_or_array) // np.dtype(dtype).itemsize,)
            nbytes = int(nbytes_or_array)
            a = None
        self.nbytes = max(nbytes, 8)
        p = ctypes.c_void_p()
        _check(lib().tfg_device_alloc(ctypes.byref(p), self.nbytes))
        self.ptr = p
        if a is not None:
            self.upload(a)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        _check(lib().tfg_memcpy_h2d(self.ptr, _ptr(a), a.nbytes))

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(lib().tfg_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().tfg_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def stream_copy(dst, src, nbytes, reps=10):
    """Mean time (ms) of a plain device copy of nbytes (calibration kernel)."""
    ms = ctypes.c_double()
    _check(lib().tfg_stream_copy(dst.ptr, src.ptr, int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value


def stream_bench(dst, src, nbytes, mode, reps=10):
    """Mean time (ms) of a sweep over nbytes: mode 0 copy, 1 read-only (src;
    per-workgroup sums into dst), 2 write-only (dst)."""
    ms = ctypes.c_double()
    _check(lib().tfg_stream_bench(dst.ptr, src.ptr if src is not None else None, int(nbytes),
                                  int(mode), int(reps), ctypes.byref(ms)))
    return ms.value


def sync_bench(grid, reps=1000):
    """(us per grid-wide barrier in one cooperative launch, us per kernel
    boundary of back-to-back empty kernels), `grid` workgroups of 256."""
    b, k = ctypes.c_double(), ctypes.c_double()
    _check(lib().tfg_sync_bench(int(grid), int(reps), ctypes.byref(b), ctypes.byref(k)))
    return b.value, k.value


class Native:
    """Native_Type restated for the GPU: the uploaded CSR structure of a grid
    (Native_Mod.f90:30-56, Create_Native.f90:14-21).  row/col/dia are the
    reference's 1-based Matrix_Type arrays."""

    def __init__(self, row, col, dia, ni=None):
        self.row = np.ascontiguousarray(row, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.dia = np.ascontiguousarray(dia, dtype=np.int32)
        self.nt = len(self.row) - 1
        self.ni = self.nt if ni is None else int(ni)
        self.nnz = len(self.col)
        h = ctypes.c_void_p()
        _check(lib().tfg_matrix_create(self.ni, self.nt, self.nnz, _ptr(self.row), _ptr(self.col),
                                       _ptr(self.dia), ctypes.byref(h)))
        self.h = h
        # handles built on this one (Faces), closed before it: a garbage
        # collector may finalise the two in either order
        self._dependents = weakref.WeakSet()

    def info(self):
        vals = [ctypes.c_int() for _ in range(11)]
        _check(lib().tfg_matrix_info(self.h, *[ctypes.byref(v) for v in vals]))
        return dict(zip(["n", "nnz", "spmv_kind", "spmv_grid", "stage_blocks", "offsets",
                         "patterns", "sym_diagonals", "sym_state", "ordering", "icc_tiles"],
                        [v.value for v in vals]))

    def reorder(self, order):
        """Internal ordering: ORDER_NATURAL or ORDER_RCM (tfg_matrix_reorder)."""
        _check(lib().tfg_matrix_reorder(self.h, int(order)))

    def kinds(self):
        """The SpMV engines this matrix offers (6 needs offset codes, 7 row
        patterns, 8 row patterns on a symmetric structure)."""
        i = self.info()
        return (list(range(6)) + ([6] if i["offsets"] else []) + ([7] if i["patterns"] else []) +
                ([8] if i["sym_diagonals"] else []))

    def close(self):
        for d in list(getattr(self, "_dependents", ())):
            d.close()
        if getattr(self, "h", None):
            lib().tfg_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- Solve_Native ------------------------------------------------------
    def solve_native(self, solver, prec, val, x, b, miter, tol, norm=None, options=None,
                     report=False):
        """Solve A x = b like Solve_Native.  Host numpy arrays: val (nnz),
        x (nt; x0 in, solution out), b (nt).  Returns (niter, fin_res) or,
        with report=True, (niter, fin_res, report_dict)."""
        for name, a, n in (("val", val, self.nnz), ("x", x, self.nt), ("b", b, self.nt)):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                    and a.shape == (n,)):
                raise TypeError(f"{name} must be a contiguous float64 array of length {n}")
        return self._solve(solver, prec, _ptr(val), _ptr(x), _ptr(b), miter, tol, norm, options,
                           report, device=False)

    def value_range(self, dv):
        """(min, max) of a device vector over the handle's nt cells (tfg_range)."""
        lo, hi = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(lib().tfg_range(self.h, dv.ptr, ctypes.byref(lo), ctypes.byref(hi)))
        return lo.value, hi.value

    def vel_max(self, du, dv, dw):
        """max(MICRO, max |u|) over the handle's nt cells (tfg_vel_max)."""
        out = ctypes.c_double(0.0)
        _check(lib().tfg_vel_max(self.h, du.ptr, dv.ptr, dw.ptr, ctypes.byref(out)))
        return out.value

    def solve_native_device(self, solver, prec, dval, dx, db, miter, tol, norm=None,
                            options=None, report=False):
        """Same with DeviceArray operands (device-resident solve)."""
        return self._solve(solver, prec, dval.ptr, dx.ptr, db.ptr, miter, tol, norm, options,
                           report, device=True)

    def _solve(self, solver, prec, pval, px, pb, miter, tol, norm, options, report, device):
        niter = ctypes.c_int(0)
        fin = ctypes.c_double(0.0)
        nv = ctypes.c_double(norm) if norm is not None else None
        opt = options if options is not None else Options()
        rep = Report()
        s, p = solver.encode(), prec.encode()
        rc = lib().tfg_solve_native_ex(self.h, s, len(s), p, len(p), pval, px, pb, int(miter),
                                       ctypes.byref(niter), float(tol), ctypes.byref(fin),
                                       ctypes.byref(nv) if nv is not None else None,
                                       1 if device else 0, ctypes.byref(opt), ctypes.byref(rep))
        _check(rc)
        if report:
            return niter.value, fin.value, rep.as_dict()
        return niter.value, fin.value

    def solve_native_noalloc(self, solver, prec, dval, dx, db, miter, tol, norm=None,
                             options=None, report=False):
        """The allocation-free solve (tfg_solve_native_noalloc) on DeviceArray
        operands: the FP64 CG of solve_native_device with its device work in the
        library's static scratch array; x is iterated in place."""
        niter = ctypes.c_int(0)
        fin = ctypes.c_double(0.0)
        nv = ctypes.c_double(norm) if norm is not None else None
        opt = options if options is not None else Options()
        rep = Report()
        s, p = solver.encode(), prec.encode()
        rc = lib().tfg_solve_native_noalloc(self.h, s, len(s), p, len(p), dval.ptr, dx.ptr,
                                            db.ptr, int(miter), ctypes.byref(niter), float(tol),
                                            ctypes.byref(fin),
                                            ctypes.byref(nv) if nv is not None else None,
                                            ctypes.byref(opt), ctypes.byref(rep))
        _check(rc)
        if report:
            return niter.value, fin.value, rep.as_dict()
        return niter.value, fin.value

    # -- building blocks ----------------------------------------------------
    def spmv(self, dval, dp, dq):
        _check(lib().tfg_spmv(self.h, dval.ptr, dp.ptr, dq.ptr))
        _check(lib().tfg_synchronize(self.h))

    def spmv_fp32(self, dval, fn, dx32, dy32):
        """FP32 row-excess product of the refinement form on val * fn."""
        _check(lib().tfg_spmv_fp32(self.h, dval.ptr, float(fn), dx32.ptr, dy32.ptr))
        _check(lib().tfg_synchronize(self.h))

    def set_comm(self, allreduce, exchange):
        """Multi-rank hooks.  allreduce(values: np.ndarray, ops: list) combines
        in place over the ranks; exchange(d_x: int device address) refreshes
        the buffer cells.  Both return None (or raise: the solve then fails)."""
        def ar(ctx, vals, ops, n):
            try:
                v = np.ctypeslib.as_array(vals, shape=(n,))
                allreduce(v, [ops[i] for i in range(n)])
                return 0
            except Exception:  # an exception must not cross the C frame
                return 1

        def ex(ctx, dx):
            try:
                exchange(dx)
                return 0
            except Exception:
                return 1

        self._comm = (ALLREDUCE_FN(ar), EXCHANGE_FN(ex))  # keep alive
        _check(lib().tfg_matrix_set_comm(self.h, self._comm[0], self._comm[1], None))

    def prec_apply(self, prec, dval, dr, dz):
        """z = M^-1 r with the reference's preconditioner of the unscaled matrix."""
        p = prec.encode()
        _check(lib().tfg_prec_apply(self.h, p, len(p), dval.ptr, dr.ptr, dz.ptr))
        _check(lib().tfg_synchronize(self.h))

    def residual(self, dval, db, dx, dr):
        _check(lib().tfg_residual(self.h, dval.ptr, db.ptr, dx.ptr, dr.ptr))
        _check(lib().tfg_synchronize(self.h))

    SPMV_KINDS = ("stage2048", "stage4096", "wave4", "wave8", "wave16", "pipe2048",
                  "pipe2048c", "pipe_pattern", "sym_dia")

    def spmv_tune(self, dval):
        """Time every SpMV engine on dval; returns (chosen kind, {name: us})."""
        k = ctypes.c_int()
        us = (ctypes.c_float * len(self.SPMV_KINDS))()
        _check(lib().tfg_spmv_tune(self.h, dval.ptr, ctypes.byref(k), us))
        return k.value, {n: t for n, t in zip(self.SPMV_KINDS, list(us)) if t >= 0}

    def spmv_select(self, kind):
        _check(lib().tfg_spmv_select(self.h, int(kind)))

    def spmv_bench(self, dval, dp, dq, reps):
        ms = ctypes.c_double()
        _check(lib().tfg_spmv_bench(self.h, dval.ptr, dp.ptr, dq.ptr, int(reps), ctypes.byref(ms)))
        return ms.value


class Batch:
    """Many independent systems solved together (tfg_batch_*).  systems: a
    list of (row, col, dia) 1-based Matrix_Type arrays."""

    def __init__(self, systems):
        self.n = np.array([len(r) - 1 for r, _, _ in systems], dtype=np.int32)
        self.nnz = np.array([len(c) for _, c, _ in systems], dtype=np.int32)
        self.nsys = len(systems)
        self.row = np.ascontiguousarray(np.concatenate([r for r, _, _ in systems]), dtype=np.int32)
        self.col = np.ascontiguousarray(np.concatenate([c for _, c, _ in systems]), dtype=np.int32)
        self.dia = np.ascontiguousarray(np.concatenate([d for _, _, d in systems]), dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.n)]).astype(np.int64)
        self.nnz_offsets = np.concatenate([[0], np.cumsum(self.nnz)]).astype(np.int64)
        self.h = ctypes.c_void_p()
        _check(lib().tfg_batch_create(self.nsys, _ptr(self.n), _ptr(self.nnz), _ptr(self.row),
                                      _ptr(self.col), _ptr(self.dia), ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().tfg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        v = [ctypes.c_int() for _ in range(4)]
        _check(lib().tfg_batch_info(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("nsys", "n", "nnz", "blocks"), (x.value for x in v)))

    def solve(self, solver, prec, val, x, b, miter, tol, norm=None, options=None, device=False):
        """val/x/b: concatenated host arrays (or DeviceArrays with device=True).
        Returns (niter[nsys], fin_res[nsys], report dict)."""
        niter = np.zeros(self.nsys, dtype=np.int32)
        fin = np.zeros(self.nsys)
        nv = None if norm is None else np.ascontiguousarray(norm, dtype=np.float64)
        opt = options if options is not None else Options()
        rep = Report()
        s, p = solver.encode(), prec.encode()
        if device:
            pv, px, pb = val.ptr, x.ptr, b.ptr
        else:
            for name, a, n in (("val", val, len(self.col)), ("x", x, len(self.dia)),
                               ("b", b, len(self.dia))):
                if not (isinstance(a, np.ndarray) and a.dtype == np.float64
                        and a.flags.c_contiguous and a.shape == (n,)):
                    raise TypeError(f"{name} must be a contiguous float64 array of length {n}")
            pv, px, pb = _ptr(val), _ptr(x), _ptr(b)
        rc = lib().tfg_batch_solve(self.h, s, len(s), p, len(p), pv, px, pb, int(miter),
                                   _ptr(niter), float(tol), _ptr(fin),
                                   _ptr(nv) if nv is not None else None, 1 if device else 0,
                                   ctypes.byref(opt), ctypes.byref(rep))
        _check(rc)
        return niter, fin, rep.as_dict()

    def tuning(self, precision=0):
        """The batch autotuner's choice for a precision: chunk (iterations per
        replayed graph), diag_classes (bit c: systems of size class c on
        diagonals use the diagonal products), rows_per_thread of the vector
        kernels, probes timed, best probe ms."""
        c, d, r, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        _check(lib().tfg_batch_tuning(self.h, int(precision), ctypes.byref(c), ctypes.byref(d),
                                      ctypes.byref(r), ctypes.byref(p), ctypes.byref(ms)))
        return {"chunk": c.value, "diag_classes": d.value, "rows_per_thread": r.value,
                "probes": p.value, "best_probe_ms": ms.value}

    def times(self):
        ms = np.zeros(self.nsys)
        _check(lib().tfg_batch_times(self.h, _ptr(ms)))
        return ms


class Faces:
    """Face structure of a grid for device-side pressure assembly
    (tfg_faces_create).  c1, c2, pos, diag_bc, b_region: the reference's face
    arrays (1-based; c2 <= 0 on the boundary; pos 2 x n_faces)."""

    def __init__(self, nat, c1, c2, pos, diag_bc, b_region):
        self.nat = nat
        arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (c1, c2, pos, diag_bc, b_region)]
        self.n_faces = len(arrs[0])
        self.h = ctypes.c_void_p()
        _check(lib().tfg_faces_create(nat.h, self.n_faces, *[_ptr(a) for a in arrs],
                                      ctypes.byref(self.h)))
        nat._dependents.add(self)

    def assemble(self, dfc, dfw, dvm, dvflux, dval, db, singular=False, dpp=None, n_bnd=0):
        """A % val and b on the device (DeviceArrays); waits for the result.
        dpp: pp % n laid out (-n_bnd:n) (ambient regions: tfg_pressure_assemble_ex)."""
        if dpp is None:
            _check(lib().tfg_pressure_assemble(self.h, dfc.ptr, dfw.ptr, dvm.ptr, dvflux.ptr,
                                               dval.ptr, db.ptr, 1 if singular else 0))
        else:
            _check(lib().tfg_pressure_assemble_ex(self.h, dfc.ptr, dfw.ptr, dvm.ptr, dvflux.ptr,
                                                  dpp.ptr.value + 8 * (n_bnd + 1), dval.ptr, db.ptr,
                                                  1 if singular else 0))
        _check(lib().tfg_synchronize(self.nat.h))

    def momentum(self, m, dval, db, dvm):
        """Momentum matrix, right-hand side and v_m of one component on the
        device (tfg_momentum_assemble); m is a Momentum; waits for the result."""
        _check(lib().tfg_momentum_assemble(self.h, ctypes.byref(m.s), dval.ptr, db.ptr, dvm.ptr))
        _check(lib().tfg_synchronize(self.nat.h))

    def rhie_chow(self, rc, dflux):
        """Rhie-Chow volume fluxes of the domain faces into dflux
        (tfg_rhie_chow_fluxes); rc is a RhieChow; waits for the result."""
        _check(lib().tfg_rhie_chow_fluxes(self.h, ctypes.byref(rc.s), dflux.ptr))
        _check(lib().tfg_synchronize(self.nat.h))

    def close(self):
        if self.h:
            lib().tfg_faces_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


ADV = {"central": 0, "upwind": 1, "luds": 2, "quick": 3, "smart": 4, "gamma": 5,
       "minmod": 6, "blended": 7, "superbee": 8, "avl_smart": 9}
TD = {"linear": 0, "parabolic": 1, "none": 2}


class _MomentumS(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int) for k in ("component", "n_bnd", "adv_scheme", "blend_matrix",
                                              "td_scheme")] +
                [(k, ctypes.c_double) for k in ("blend", "dt", "urf", "p_drop")] +
                [(k, ctypes.c_void_p) for k in ("f", "fw", "fc", "v_flux", "vis_eff", "f_stress",
                                                "sx", "sy", "sz", "dx", "dy", "dz", "wall",
                                                "phi_n", "phi_o", "phi_oo", "gx", "gy", "gz",
                                                "uj_i", "uk_i", "density", "vol", "p_i",
                                                "cell_f", "st")])


class Momentum:
    """tfg_momentum on device copies of a set of momentum fields (the
    attributes of tflows_cases.MomentumFields; the boundary-extended cell
    arrays are laid out (-n_bnd:n) and passed at cell 1)."""

    EXT = ("phi_n", "gx", "gy", "gz", "uj_i", "uk_i", "density")
    FACE = ("f", "fw", "fc", "v_flux", "vis_eff", "sx", "sy", "sz", "dx", "dy", "dz")
    CELL = ("phi_o", "phi_oo", "vol", "p_i")

    def __init__(self, mf, component=1, adv="central", blend_matrix=True, td="linear", blend=0.8,
                 dt=1e-2, urf=0.7, p_drop=0.5, stress=True, cell_force=True,
                 surface_tension=False):
        s = self.s = _MomentumS()
        self.dev = {}
        s.component, s.n_bnd, s.adv_scheme = component, mf.n_bnd, ADV[adv]
        s.blend_matrix, s.td_scheme = int(blend_matrix), TD[td]
        s.blend, s.dt, s.urf, s.p_drop = blend, dt, urf, p_drop
        for k in self.FACE + self.CELL:
            setattr(s, k, self._up(k, getattr(mf, k)))
        s.wall = self._up("wall", np.ascontiguousarray(mf.wall, dtype=np.int32))
        for k in self.EXT:
            setattr(s, k, self._up(k, getattr(mf, k), (mf.n_bnd + 1) * 8))
        s.f_stress = self._up("f_stress", mf.f_stress) if stress else None
        s.cell_f = self._up("cell_f", mf.cell_f) if cell_force else None
        s.st = self._up("st", mf.st) if surface_tension else None

    def _up(self, key, a, offset=0):
        d = DeviceArray(np.ascontiguousarray(a))
        self.dev[key] = d
        return d.ptr.value + offset


RC_PTRS = ("f", "fc", "sx", "sy", "sz", "dx", "dy", "dz", "u", "v", "w", "p", "px", "py", "pz",
           "v_m")
RC_CHOI = ("u_o", "u_oo", "v_o", "v_oo", "w_o_", "w_oo_", "density", "flux_o", "flux_oo")
RC_GU = ("cell_fx", "cell_fy", "cell_fz", "face_fx", "face_fy", "face_fz")
RC_VOF = ("curv", "col", "col_x", "col_y", "col_z")


class _RhieChowS(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_void_p) for k in RC_PTRS] + [("choi", ctypes.c_int)] +
                [(k, ctypes.c_void_p) for k in RC_CHOI] +
                [(k, ctypes.c_double) for k in ("dt", "w_o", "w_oo")] + [("gu", ctypes.c_int)] +
                [(k, ctypes.c_void_p) for k in RC_GU] +
                [("vof", ctypes.c_int), ("sigma", ctypes.c_double)] +
                [(k, ctypes.c_void_p) for k in RC_VOF])


class RhieChow:
    """tfg_rhie_chow on device copies of tflows_cases.RhieChowFields; the
    arrays of a switched-off correction are not uploaded (NULL)."""

    def __init__(self, rf, choi=False, dt=1e-2, w_o=1.0, w_oo=0.0, gu=False, vof=False,
                 sigma=0.07):
        s = self.s = _RhieChowS()
        self.dev = {}
        for k in RC_PTRS + (RC_CHOI if choi else ()) + (RC_GU if gu else ()) + \
                (RC_VOF if vof else ()):
            d = DeviceArray(np.ascontiguousarray(getattr(rf, k), dtype=np.float64))
            self.dev[k] = d
            setattr(s, k, d.ptr.value)
        s.choi, s.gu, s.vof = int(choi), int(gu), int(vof)
        s.dt, s.w_o, s.w_oo, s.sigma = dt, w_o, w_oo, sigma
