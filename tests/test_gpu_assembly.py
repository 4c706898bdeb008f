"""Pressure-system assembly on the device (tfg_faces_create /
tfg_pressure_assemble) against the oracle's serial face loop: A % val and b
bit-identical, and the solve on the device-assembled system identical to the
solve on the host-built one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _assemble_gpu(gpu, f, singular):
    nat = gpu.Native(f.row, f.col, f.dia)
    faces = gpu.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, f.b_zero)
    d = {k: gpu.DeviceArray(getattr(f, k)) for k in ("fc", "fw", "vm", "vflux")}
    dval, db = gpu.DeviceArray(np.zeros(len(f.col))), gpu.DeviceArray(np.zeros(f.n))
    faces.assemble(d["fc"], d["fw"], d["vm"], d["vflux"], dval, db, singular=singular)
    return nat, faces, dval, db


@pytest.mark.parametrize("dims,singular", [((12, 10, 8), False), ((12, 10, 8), True),
                                           ((30, 30, 30), True), ((1, 1, 9), False)])
def test_assembly_bitexact(cases, oracle, gpu, dims, singular):
    f = cases.hex_faces(*dims)
    print(f"assembly case {dims} n={f.n} faces={f.n_faces} singular={singular}", flush=True)
    val_o, b_o = oracle.pressure_assemble(f, singular=singular)
    nat, faces, dval, db = _assemble_gpu(gpu, f, singular)
    val, b = dval.download(), db.download()
    assert np.array_equal(val, val_o), np.abs(val - val_o).max()
    assert np.array_equal(b, b_o), np.abs(b - b_o).max()


def test_solve_on_assembled_system(cases, oracle, gpu):
    f = cases.hex_faces(16, 14, 12)
    f.b_zero[:] = 0
    val_o, b_o = oracle.pressure_assemble(f, singular=True)
    nat, faces, dval, db = _assemble_gpu(gpu, f, True)
    dx = gpu.DeviceArray(np.zeros(f.n))
    it_d, res_d = nat.solve_native_device("cg", "jacobi", dval, dx, db, 2000, 1e-8, norm=1.0)
    x = np.zeros(f.n)
    it_h, res_h = nat.solve_native("cg", "jacobi", val_o.copy(), x, b_o.copy(), 2000, 1e-8,
                                   norm=1.0)
    assert (it_d, res_d) == (it_h, res_h) and np.array_equal(dx.download(), x)
    assert res_d < 1e-8


def test_faces_validation(cases, gpu):
    f = cases.hex_faces(4, 3, 3)
    nat = gpu.Native(f.row, f.col, f.dia)
    pos = f.pos.copy()
    k = int(np.nonzero(f.c2 > 0)[0][0])
    pos[2 * k] = f.dia[f.c1[k] - 1]          # points at the diagonal
    with pytest.raises(gpu.TFlowsGpuError) as e:
        gpu.Faces(nat, f.c1, f.c2, pos, f.diag_bc, f.b_zero)
    assert e.value.code == gpu.TFG_EINVAL and "pos" in str(e.value)
