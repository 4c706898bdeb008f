"""The C-ABI library loads, exports every function include/tflows_gpu.h
declares, and validates its structure input on the host (no GPU needed)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol(gpu):
    names = gpu.declared_symbols()
    assert len(names) >= 15
    L = gpu.lib()
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", gpu.LIB_PATH], capture_output=True,
                         text=True).stdout
    for n in names:
        assert f" T {n}" in out, n


def _create(gpu, ni, nt, row, col, dia):
    h = ctypes.c_void_p()
    rc = gpu.lib().tfg_matrix_create(ni, nt, len(col), row.ctypes.data, col.ctypes.data,
                                     dia.ctypes.data, ctypes.byref(h))
    return rc, gpu.lib().tfg_last_error().decode()


def test_structure_validation_on_host(gpu):
    i32 = lambda a: np.array(a, dtype=np.int32)  # noqa: E731
    # diagonal not first (Create_Matrix.f90:118 asserts it is)
    rc, msg = _create(gpu, 2, 2, i32([1, 3, 4]), i32([2, 1, 2]), i32([2, 3]))
    assert rc == gpu.TFG_EINVAL and "diagonal" in msg
    # column out of range
    rc, msg = _create(gpu, 2, 2, i32([1, 3, 4]), i32([1, 3, 2]), i32([1, 3]))
    assert rc == gpu.TFG_EINVAL and "out of range" in msg
    # nnz does not match row(nt+1)-1
    h = ctypes.c_void_p()
    row, col, dia = i32([1, 2, 3]), i32([1, 2]), i32([1, 2])
    rc = gpu.lib().tfg_matrix_create(2, 2, 5, row.ctypes.data, col.ctypes.data, dia.ctypes.data,
                                     ctypes.byref(h))
    assert rc == gpu.TFG_EINVAL
    # more solved rows than stored rows
    rc, msg = _create(gpu, 3, 2, i32([1, 2, 3]), i32([1, 2]), i32([1, 2]))
    assert rc == gpu.TFG_EINVAL and "sizes" in msg
    # null arguments
    assert gpu.lib().tfg_matrix_create(2, 2, 2, None, None, None, ctypes.byref(h)) == gpu.TFG_EINVAL
    # an empty system, and an empty row (every row holds at least its diagonal)
    rc, msg = _create(gpu, 0, 0, i32([1]), i32([1]), i32([1]))
    assert rc == gpu.TFG_EINVAL and "sizes" in msg
    rc, msg = _create(gpu, 2, 2, i32([1, 2, 2]), i32([1, 1]), i32([1, 2]))
    assert rc == gpu.TFG_EINVAL and ("empty" in msg or "nnz" in msg)
    # the face structure's null arguments are refused before any device call
    assert gpu.lib().tfg_faces_create(None, 0, None, None, None, None, None,
                                      ctypes.byref(h)) == gpu.TFG_EINVAL


def test_batch_validation_on_host(gpu):
    """tfg_batch_create checks every system like tfg_matrix_create and names
    the failing one, before touching a device."""
    i32 = lambda a: np.array(a, dtype=np.int32)  # noqa: E731
    n, nnz = i32([2, 2]), i32([3, 3])
    row = i32([1, 3, 4, 1, 3, 4])
    col = i32([1, 2, 2, 2, 1, 2])      # system 1: diagonal not first in row 1
    dia = i32([1, 3, 1, 3])
    h = ctypes.c_void_p()
    rc = gpu.lib().tfg_batch_create(2, n.ctypes.data, nnz.ctypes.data, row.ctypes.data,
                                    col.ctypes.data, dia.ctypes.data, ctypes.byref(h))
    msg = gpu.lib().tfg_last_error().decode()
    assert rc == gpu.TFG_EINVAL and "system 1" in msg and "diagonal" in msg
    rc = gpu.lib().tfg_batch_create(0, n.ctypes.data, nnz.ctypes.data, row.ctypes.data,
                                    col.ctypes.data, dia.ctypes.data, ctypes.byref(h))
    assert rc == gpu.TFG_EINVAL


def test_python_mirror_checks_operands(cases, gpu):
    s = cases.hex_pressure(3)
    nat = object.__new__(gpu.Native)     # no device handle needed for argument checks
    nat.nt, nat.nnz, nat.h = s.n, s.nnz, None
    with pytest.raises(TypeError):
        nat.solve_native("cg", "jacobi", s.val.astype(np.float32), np.zeros(s.n), s.b, 10, 1e-8)
    with pytest.raises(TypeError):
        nat.solve_native("cg", "jacobi", s.val, np.zeros(s.n + 1), s.b, 10, 1e-8)


def test_header_is_plain_c():
    """The boundary header compiles as C (no C++ or HIP types in signatures)."""
    src = os.path.join(ROOT, "include", "tflows_gpu.h")
    r = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-Werror", "-x", "c", src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
