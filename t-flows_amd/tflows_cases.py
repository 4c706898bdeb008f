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


@dataclass
class Faces:
    """Face data of a hex grid in the reference's layout (Grid_Mod faces_c,
    Matrix_Type pos / dia): 1-based cells c1, c2 (c2 <= 0 on the boundary),
    the CSR positions pos (2 x n_faces, 1-based; 0 on the boundary) of the
    pressure matrix, per-face fc, interpolation weight fw, volume flux, the
    boundary flags (diag_bc, b_region: 1 pressure, 2 ambient region) and the
    region number of each face (0 internal, 1..6 the sides x0, x1, y0, y1,
    z0, z1); pp = pp % n over the boundary cells and the cells, laid out as
    Fortran's (-n_bnd:n) (cell c at element c + n_bnd), when a side is
    ambient."""
    row: np.ndarray
    col: np.ndarray
    dia: np.ndarray
    c1: np.ndarray
    c2: np.ndarray
    pos: np.ndarray
    diag_bc: np.ndarray
    b_region: np.ndarray
    fc: np.ndarray
    fw: np.ndarray
    vm: np.ndarray
    vflux: np.ndarray
    region: np.ndarray = None
    pp: np.ndarray = None

    @property
    def n(self):
        return len(self.dia)

    @property
    def n_bnd(self):
        return int(max(0, -self.c2.min())) if len(self.c2) else 0

    @property
    def n_faces(self):
        return len(self.c1)


def hex_faces(nx, ny=None, nz=None, seed=DEFAULT_SEED, shuffle=True, ambient=False):
    """Faces of an nx*ny*nz hex grid (cells x slowest, z fastest, as the
    pressure generators number them) with seeded face data: fc = A/d times
    U(0.5, 2), fw ~ U(0.3, 0.7), v_m ~ U(0.5, 2) * 1e-3, flux ~ U(-1, 1).
    The x = 0 side is a pressure region.  Faces are listed in a seeded
    random order (shuffle) so that assembly order matters.  ambient: the
    y = 0 side is an ambient region too (pp ~ U(-1, 1) on its boundary
    cells) and the boundary faces come first, side by side in region order,
    as T-Flows numbers them (the internal faces after them, shuffled)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = nx * ny * nz
    s = hex_pressure(nx, ny, nz)       # structure: diagonal first, columns ascending
    idx = np.arange(n).reshape(nx, ny, nz)
    pairs = [np.stack([idx[:-1].ravel(), idx[1:].ravel()], 1),
             np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1),
             np.stack([idx[:, :, :-1].ravel(), idx[:, :, 1:].ravel()], 1)]
    internal = np.concatenate(pairs)
    bnd_cells, bnd_reg = [], []
    for axis, size in ((0, nx), (1, ny), (2, nz)):
        for side in (0, size - 1):
            sl = [slice(None)] * 3
            sl[axis] = side
            cells = idx[tuple(sl)].ravel()
            bnd_cells.append(cells)
            bnd_reg.append(np.full(cells.size, 2 * axis + (1 if side == 0 else 2)))
    bnd_cells = np.concatenate(bnd_cells)
    bnd_reg = np.concatenate(bnd_reg)
    ni_, nb = len(internal), len(bnd_cells)
    c1 = np.concatenate([internal[:, 0], bnd_cells]) + 1
    c2 = np.concatenate([internal[:, 1] + 1, -(np.arange(nb) + 1)])
    region = np.concatenate([np.zeros(ni_, int), bnd_reg])
    code = np.zeros(len(c1), int)
    code[region == 1] = 1                      # x = 0: pressure
    if ambient:
        code[region == 3] = 2                  # y = 0: ambient
    flag = code > 0
    # CSR positions of (c1, c2) in row c1 and (c2, c1) in row c2
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(s.row))
    keys = rows * n + (s.col - 1)
    korder = np.argsort(keys, kind="stable")
    skeys = keys[korder]

    def position(r, c):
        q = r.astype(np.int64) * n + c
        return korder[np.searchsorted(skeys, q)]

    pos = np.zeros((len(c1), 2), dtype=np.int32)
    pos[:ni_, 0] = position(internal[:, 0], internal[:, 1]) + 1
    pos[:ni_, 1] = position(internal[:, 1], internal[:, 0]) + 1
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(c1)) if shuffle else np.arange(len(c1))
    if ambient:   # boundary faces first in region order, then the internal ones
        order = np.concatenate([np.arange(ni_, ni_ + nb), order[order < ni_]])
    h = 1.0 / nx
    nf = len(c1)
    f = Faces(row=s.row, col=s.col, dia=s.dia,
              c1=np.ascontiguousarray(c1[order], dtype=np.int32),
              c2=np.ascontiguousarray(c2[order], dtype=np.int32),
              pos=np.ascontiguousarray(pos[order].ravel(), dtype=np.int32),
              diag_bc=np.ascontiguousarray(flag[order], dtype=np.int32),
              b_region=np.ascontiguousarray(code[order], dtype=np.int32),
              fc=h * rng.uniform(0.5, 2.0, nf), fw=rng.uniform(0.3, 0.7, nf),
              vm=1e-3 * rng.uniform(0.5, 2.0, n), vflux=rng.uniform(-1.0, 1.0, nf),
              region=np.ascontiguousarray(region[order], dtype=np.int32))
    if ambient:
        f.pp = rng.uniform(-1.0, 1.0, nb + 1 + n)
        f.pp[nb] = 0.0                         # cell 0: unused
    return f


@dataclass
class MomentumFields:
    """Face and cell data of one momentum assembly on a Faces set, in the
    reference's layout.  Per face (n_faces): interpolation factors f
    (Grid % f) and fw (Grid % fw), M % fc, v_flux % n, face viscosity, face
    stress, the face vectors s (Grid % sx, sy, sz) and d (dx, dy, dz), and
    `wall` = 1 on boundary faces with an INFLOW / WALL / CONVECT / WALLFL
    condition.  Per cell: arrays marked ext hold the boundary cells as well,
    laid out as Fortran's (-n_bnd:n) so that cell c is element c + n_bnd
    (element n_bnd, cell 0, is unused)."""
    n_bnd: int
    f: np.ndarray
    fw: np.ndarray
    fc: np.ndarray
    v_flux: np.ndarray
    vis_eff: np.ndarray
    f_stress: np.ndarray
    sx: np.ndarray
    sy: np.ndarray
    sz: np.ndarray
    dx: np.ndarray
    dy: np.ndarray
    dz: np.ndarray
    wall: np.ndarray
    phi_n: np.ndarray    # ext
    gx: np.ndarray       # ext
    gy: np.ndarray       # ext
    gz: np.ndarray       # ext
    uj_i: np.ndarray     # ext
    uk_i: np.ndarray     # ext
    density: np.ndarray  # ext
    phi_o: np.ndarray
    phi_oo: np.ndarray
    vol: np.ndarray
    p_i: np.ndarray
    cell_f: np.ndarray
    st: np.ndarray


def momentum_fields(faces, seed=DEFAULT_SEED, rho_ratio=1000.0):
    """Seeded momentum data on a Faces set (hex_faces): velocities U(-1, 1),
    gradients U(-1, 1) / h, density 1 or rho_ratio per cell (a random half of
    the boundary cells included), viscosity-like face data ~ 1e-3, time-step
    and force terms of unit size."""
    rng = np.random.default_rng(seed + 17)
    nf = faces.n_faces
    n = faces.n
    nb = int(-faces.c2.min()) if faces.c2.min() < 0 else 0
    h = faces.fc.mean()

    def ext(scale=1.0):
        a = scale * rng.uniform(-1.0, 1.0, nb + 1 + n)
        a[nb] = 0.0
        return a

    dens = np.where(rng.uniform(0, 1, nb + 1 + n) < 0.5, 1.0, rho_ratio)
    dens[nb] = 0.0
    return MomentumFields(
        n_bnd=nb, f=rng.uniform(0.3, 0.7, nf), fw=rng.uniform(0.3, 0.7, nf),
        fc=faces.fc.copy(), v_flux=rng.uniform(-1.0, 1.0, nf) * h * h,
        vis_eff=1e-3 * rng.uniform(0.5, 2.0, nf), f_stress=1e-4 * rng.uniform(-1.0, 1.0, nf),
        sx=h * h * rng.uniform(-1.0, 1.0, nf), sy=h * h * rng.uniform(-1.0, 1.0, nf),
        sz=h * h * rng.uniform(-1.0, 1.0, nf), dx=h * rng.uniform(-1.0, 1.0, nf),
        dy=h * rng.uniform(-1.0, 1.0, nf), dz=h * rng.uniform(-1.0, 1.0, nf),
        wall=np.ascontiguousarray((faces.c2 < 0) & (rng.uniform(0, 1, nf) < 0.7), dtype=np.int32),
        phi_n=ext(), gx=ext(1.0 / h), gy=ext(1.0 / h), gz=ext(1.0 / h), uj_i=ext(1.0 / h),
        uk_i=ext(1.0 / h), density=dens, phi_o=rng.uniform(-1.0, 1.0, n),
        phi_oo=rng.uniform(-1.0, 1.0, n), vol=h ** 3 * rng.uniform(0.5, 2.0, n),
        p_i=rng.uniform(-1.0, 1.0, n) / h, cell_f=rng.uniform(-1.0, 1.0, n),
        st=rng.uniform(-1.0, 1.0, n))


@dataclass
class RhieChowFields:
    """Face and cell data of the Rhie-Chow flux loop on a hex Faces set.
    Per face: Grid % f, A % fc, the area vector s and the centre-to-centre
    vector d (c1 to c2) of the uniform grid; per cell (1..n): velocities,
    pressure and its gradient, v_m, and the fields of Choi's and Gu's
    corrections and of the surface-tension term."""
    f: np.ndarray
    fc: np.ndarray
    sx: np.ndarray
    sy: np.ndarray
    sz: np.ndarray
    dx: np.ndarray
    dy: np.ndarray
    dz: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    p: np.ndarray
    px: np.ndarray
    py: np.ndarray
    pz: np.ndarray
    v_m: np.ndarray
    u_o: np.ndarray
    u_oo: np.ndarray
    v_o: np.ndarray
    v_oo: np.ndarray
    w_o_: np.ndarray
    w_oo_: np.ndarray
    density: np.ndarray
    flux_o: np.ndarray
    flux_oo: np.ndarray
    cell_fx: np.ndarray
    cell_fy: np.ndarray
    cell_fz: np.ndarray
    face_fx: np.ndarray
    face_fy: np.ndarray
    face_fz: np.ndarray
    curv: np.ndarray
    col: np.ndarray
    col_x: np.ndarray
    col_y: np.ndarray
    col_z: np.ndarray


def cell_centres(nx, ny, nz):
    """Centres of the cells of the unit cube's nx*ny*nz hex grid (x slowest,
    z fastest, as the generators number them)."""
    i = np.arange(nx * ny * nz)
    return np.stack([(i // (ny * nz) + 0.5) / nx, ((i // nz) % ny + 0.5) / ny,
                     (i % nz + 0.5) / nz], 1)


def rhie_chow_fields(faces, nx, ny, nz, seed=DEFAULT_SEED):
    """Seeded fields on hex_faces(nx, ny, nz): the grid's own geometry
    (d = centre(c2) - centre(c1), s = area normal, f = 1/2, fc = |s|/|d|)
    and random cell and face data of unit size."""
    rng = np.random.default_rng(seed + 29)
    n, nf = faces.n, faces.n_faces
    xc = cell_centres(nx, ny, nz)
    internal = faces.c2 > 0
    d = np.zeros((nf, 3))
    d[internal] = xc[faces.c2[internal] - 1] - xc[faces.c1[internal] - 1]
    h = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    area = np.array([h[1] * h[2], h[0] * h[2], h[0] * h[1]])
    axis = np.argmax(np.abs(d), 1)
    sv = np.zeros((nf, 3))
    sv[np.arange(nf), axis] = area[axis] * np.where(internal, 1.0, 0.0)
    fc = np.where(internal, np.linalg.norm(sv, axis=1) / np.maximum(np.linalg.norm(d, axis=1), 1e-300), 0.0)

    def cells(scale=1.0):
        return scale * rng.uniform(-1.0, 1.0, n)

    def faces_(scale=1.0):
        return scale * rng.uniform(-1.0, 1.0, nf)
    return RhieChowFields(
        f=np.full(nf, 0.5), fc=fc, sx=sv[:, 0].copy(), sy=sv[:, 1].copy(), sz=sv[:, 2].copy(),
        dx=d[:, 0].copy(), dy=d[:, 1].copy(), dz=d[:, 2].copy(),
        u=cells(), v=cells(), w=cells(), p=cells(), px=cells(), py=cells(), pz=cells(),
        v_m=1e-2 * rng.uniform(0.5, 2.0, n), u_o=cells(), u_oo=cells(), v_o=cells(),
        v_oo=cells(), w_o_=cells(), w_oo_=cells(), density=rng.uniform(1.0, 10.0, n),
        flux_o=faces_(area.max()), flux_oo=faces_(area.max()), cell_fx=cells(), cell_fy=cells(),
        cell_fz=cells(), face_fx=faces_(), face_fy=faces_(), face_fz=faces_(), curv=cells(10.0),
        col=rng.uniform(0.0, 1.0, n), col_x=cells(), col_y=cells(), col_z=cells())
