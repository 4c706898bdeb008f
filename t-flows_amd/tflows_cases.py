"""Synthetic T-Flows linear systems (ctypes front end of csrc/cases.c).

Each generator returns a `System` holding a 1-based CSR matrix in the layout
of Sources/Process/Cpu/Matrix_Mod/Create_Matrix.f90:87-119 (diagonal first,
other columns ascending, int32 indices) and a right-hand side.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "lib", "libtflows_cases.so")
DEFAULT_SEED = 20261202


class _Sys(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nnz", ctypes.c_int),
                ("row", ctypes.POINTER(ctypes.c_int)), ("col", ctypes.POINTER(ctypes.c_int)),
                ("dia", ctypes.POINTER(ctypes.c_int)), ("val", ctypes.POINTER(ctypes.c_double)),
                ("b", ctypes.POINTER(ctypes.c_double))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"{_LIB} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(_LIB)
        I, D, U64 = ctypes.c_int, ctypes.c_double, ctypes.c_uint64
        S = ctypes.POINTER(_Sys)
        L.tfc_hex_pressure.argtypes = [I, I, I, D, U64, S]
        L.tfc_hex_pressure.restype = I
        L.tfc_upwind.argtypes = [I, U64, S]
        L.tfc_upwind.restype = I
        L.tfc_knn.argtypes = [I, I, I, I, I, U64, S]
        L.tfc_knn.restype = I
        L.tfc_free.argtypes = [S]
        L.tfc_batch_params.argtypes = [I, I, I, D, U64, ctypes.c_void_p, ctypes.c_void_p]
        L.tfc_rng_u01.argtypes = [U64, U64, U64]
        L.tfc_rng_u01.restype = D
        L.tfc_hex_nnz.argtypes = [I, I, I]
        L.tfc_hex_nnz.restype = ctypes.c_int64
        _lib = L
    return _lib


@dataclass
class System:
    name: str
    row: np.ndarray   # int32 n+1, 1-based
    col: np.ndarray   # int32 nnz, 1-based
    dia: np.ndarray   # int32 n,   1-based positions
    val: np.ndarray   # float64 nnz
    b: np.ndarray     # float64 n

    @property
    def n(self):
        return len(self.row) - 1

    @property
    def nnz(self):
        return len(self.col)

    def scipy(self):
        """scipy CSR copy (scipy may sort the indices of its matrix in place, so
        it must not share `val` with this system)."""
        import scipy.sparse as sp
        return sp.csr_matrix((self.val.copy(), self.col - 1, self.row - 1), shape=(self.n, self.n))


def _take(s, name):
    n, nnz = s.n, s.nnz
    out = System(name,
                 np.ctypeslib.as_array(s.row, (n + 1,)).copy(),
                 np.ctypeslib.as_array(s.col, (nnz,)).copy(),
                 np.ctypeslib.as_array(s.dia, (n,)).copy(),
                 np.ctypeslib.as_array(s.val, (nnz,)).copy(),
                 np.ctypeslib.as_array(s.b, (n,)).copy())
    lib().tfc_free(ctypes.byref(s))
    return out


def hex_pressure(nx, ny=None, nz=None, rho_out=1.0, seed=DEFAULT_SEED):
    """Pressure-Poisson system on an nx*ny*nz hex grid of the unit cube.

    rho_out = 1: constant-coefficient Laplacian (a_f = A_f/d_f).
    rho_out > 1: two-phase density, rho = 1 inside a seeded sphere of radius
    0.25 and rho_out outside, harmonic-averaged on faces.  Neumann walls,
    first cell pinned, mean-free U(-1,1) RHS."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    s = _Sys()
    rc = lib().tfc_hex_pressure(nx, ny, nz, float(rho_out), seed, ctypes.byref(s))
    if rc:
        raise ValueError(f"tfc_hex_pressure failed ({rc})")
    return _take(s, f"hex_pressure_{nx}x{ny}x{nz}_rho{rho_out:g}")


def upwind(N, seed=DEFAULT_SEED):
    """Non-symmetric first-order upwind convection-diffusion on N^3 cells."""
    s = _Sys()
    rc = lib().tfc_upwind(N, seed, ctypes.byref(s))
    if rc:
        raise ValueError(f"tfc_upwind failed ({rc})")
    return _take(s, f"upwind_{N}")


def knn(nx, ny, nz, kmin=10, kmax=18, seed=DEFAULT_SEED):
    """Unstructured symmetric kNN system on a jittered lattice, randomly
    permuted."""
    s = _Sys()
    rc = lib().tfc_knn(nx, ny, nz, kmin, kmax, seed, ctypes.byref(s))
    if rc:
        raise ValueError(f"tfc_knn failed ({rc})")
    return _take(s, f"knn_{nx}x{ny}x{nz}_k{kmin}-{kmax}")


def batch_params(count, smin=20, smax=80, rmax=1000.0, seed=DEFAULT_SEED):
    sides = np.zeros(count, dtype=np.int32)
    ratios = np.zeros(count, dtype=np.float64)
    lib().tfc_batch_params(count, smin, smax, rmax, seed, sides.ctypes.data, ratios.ctypes.data)
    return sides, ratios


def hex_nnz(nx, ny, nz):
    return int(lib().tfc_hex_nnz(nx, ny, nz))
