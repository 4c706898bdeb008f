"""Pressure-system assembly on the device (tfg_faces_create /
tfg_pressure_assemble / _ex) against the oracle's serial face loop: A % val
and b bit-identical (pressure and ambient regions), and the solve on the
device-assembled system identical to the solve on the host-built one;
vel_max (tfg_vel_max); momentum assembly and Rhie-Chow fluxes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _assemble_gpu(gpu, f, singular):
    nat = gpu.Native(f.row, f.col, f.dia)
    faces = gpu.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, f.b_region)
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


@pytest.mark.parametrize("singular", [False, True])
def test_ambient_assembly_bitexact(cases, oracle, gpu, singular):
    """Compute_Pressure.f90:171-189 with a pressure and an ambient side
    (corner cells on both: the later region's value)."""
    f = cases.hex_faces(12, 10, 8, ambient=True)
    print(f"ambient assembly n={f.n} faces={f.n_faces} singular={singular}", flush=True)
    val_o, b_o = oracle.pressure_assemble(f, singular=singular)
    nat = gpu.Native(f.row, f.col, f.dia)
    faces = gpu.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, f.b_region)
    d = {k: gpu.DeviceArray(getattr(f, k)) for k in ("fc", "fw", "vm", "vflux", "pp")}
    dval, db = gpu.DeviceArray(np.zeros(len(f.col))), gpu.DeviceArray(np.zeros(f.n))
    faces.assemble(d["fc"], d["fw"], d["vm"], d["vflux"], dval, db, singular=singular,
                   dpp=d["pp"], n_bnd=f.n_bnd)
    assert np.array_equal(dval.download(), val_o)
    assert np.array_equal(db.download(), b_o)
    # without pp % n the ambient faces are refused
    with pytest.raises(gpu.TFlowsGpuError) as e:
        faces.assemble(d["fc"], d["fw"], d["vm"], d["vflux"], dval, db)
    assert e.value.code == gpu.TFG_EINVAL and "ambient" in str(e.value)


def test_vel_max(oracle, gpu, cases):
    """Compute_Momentum.f90:201-204 over the handle's cells."""
    s = cases.hex_pressure(20, 18, 16)
    nat = gpu.Native(s.row, s.col, s.dia)
    rng = np.random.default_rng(8)
    u, v, w = (rng.uniform(-2, 2, s.n) for _ in range(3))
    du, dv, dw = (gpu.DeviceArray(a) for a in (u, v, w))
    assert nat.vel_max(du, dv, dw) == oracle.vel_max(u, v, w)
    z = gpu.DeviceArray(np.zeros(s.n))
    assert nat.vel_max(z, z, z) == 1e-6


def test_pressure_range(cases, gpu):
    """Compute_Pressure.f90:104-109: p_min and p_max over the cells, the
    pressure solve's norm range (a minimum and a maximum: numpy's exactly)."""
    s = cases.hex_pressure(19, 17, 13)
    nat = gpu.Native(s.row, s.col, s.dia)
    p = np.random.default_rng(2).uniform(-3.0, 5.0, s.n)
    p[123] = -7.5
    p[s.n - 1] = 9.25
    assert nat.value_range(gpu.DeviceArray(p)) == (-7.5, 9.25)
    c = np.full(s.n, 0.125)
    assert nat.value_range(gpu.DeviceArray(c)) == (0.125, 0.125)


def test_solve_on_assembled_system(cases, oracle, gpu):
    f = cases.hex_faces(16, 14, 12)
    f.b_region[:] = 0
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
    bad = f.b_region.copy()
    bad[np.flatnonzero(f.c2 < 0)[0]] = 3
    with pytest.raises(gpu.TFlowsGpuError) as e:
        gpu.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, bad)
    assert e.value.code == gpu.TFG_EINVAL and "region" in str(e.value)
    pos = f.pos.copy()
    k = int(np.nonzero(f.c2 > 0)[0][0])
    pos[2 * k] = f.dia[f.c1[k] - 1]          # points at the diagonal
    with pytest.raises(gpu.TFlowsGpuError) as e:
        gpu.Faces(nat, f.c1, f.c2, pos, f.diag_bc, f.b_region)
    assert e.value.code == gpu.TFG_EINVAL and "pos" in str(e.value)


# ---- momentum (tfg_momentum_assemble) ------------------------------------

def _momentum_case(cases, dims, seed=5):
    f = cases.hex_faces(*dims, seed=seed)
    return f, cases.momentum_fields(f, seed=seed)


def _momentum_gpu(gpu, f, mf, **kw):
    nat = gpu.Native(f.row, f.col, f.dia)
    faces = gpu.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, f.b_region)
    m = gpu.Momentum(mf, **kw)
    dval, db, dvm = (gpu.DeviceArray(np.full(k, np.nan)) for k in (len(f.col), f.n, f.n))
    faces.momentum(m, dval, db, dvm)
    return nat, dval.download(), db.download(), dvm.download()


MOMENTUM_CASES = [
    # component, scheme, blend_matrix, time scheme, stress, cell force, surface tension
    (1, "central", True, "linear", True, True, False),
    (2, "upwind", True, "parabolic", False, False, True),
    (3, "luds", False, "linear", True, True, True),
    (1, "quick", True, "none", True, False, False),
    (2, "smart", True, "linear", True, True, False),
    (3, "gamma", True, "parabolic", True, True, False),
    (1, "minmod", False, "linear", False, True, False),
    (2, "blended", True, "linear", True, True, False),
    (3, "superbee", True, "parabolic", True, False, True),
    (1, "avl_smart", True, "linear", True, True, False),
]


@pytest.mark.parametrize("comp,adv,blend,td,stress,cf,st", MOMENTUM_CASES,
                         ids=[f"{c[1]}-u{c[0]}" for c in MOMENTUM_CASES])
def test_momentum_bitexact(cases, oracle, gpu, comp, adv, blend, td, stress, cf, st):
    """val, b and v_m of every scheme, time scheme and optional term equal
    the serial loops bit for bit (faces in a shuffled order)."""
    f, mf = _momentum_case(cases, (11, 9, 7))
    kw = dict(component=comp, adv=adv, blend_matrix=blend, td=td, stress=stress,
              cell_force=cf, surface_tension=st)
    print(f"momentum case {kw} n={f.n} faces={f.n_faces}", flush=True)
    val_o, b_o, vm_o = oracle.momentum_assemble(f, mf, **kw)
    _, val, b, vm = _momentum_gpu(gpu, f, mf, **kw)
    assert np.array_equal(val, val_o), np.abs(val - val_o).max()
    assert np.array_equal(b, b_o), np.abs(b - b_o).max()
    assert np.array_equal(vm, vm_o), np.abs(vm - vm_o).max()


def test_momentum_boundary_faces_numbered_first(cases, oracle, gpu):
    """T-Flows numbers the boundary faces first: the advection loops'
    boundary-then-domain order is then plain face order."""
    f = cases.hex_faces(9, 8, 7, shuffle=False)
    order = np.argsort(f.c2 > 0, kind="stable")
    pos = f.pos.reshape(-1, 2)[order].ravel()
    g = type(f)(**{**f.__dict__, "c1": f.c1[order], "c2": f.c2[order], "pos": pos,
                   "diag_bc": f.diag_bc[order], "b_region": f.b_region[order], "fc": f.fc[order],
                   "fw": f.fw[order], "vflux": f.vflux[order]})
    mf = cases.momentum_fields(g)
    kw = dict(adv="superbee", td="parabolic")
    val_o, b_o, vm_o = oracle.momentum_assemble(g, mf, **kw)
    _, val, b, vm = _momentum_gpu(gpu, g, mf, **kw)
    assert np.array_equal(val, val_o) and np.array_equal(b, b_o) and np.array_equal(vm, vm_o)


def test_momentum_system_solves(cases, oracle, gpu):
    """The device-assembled momentum system (non-symmetric with the implicit
    upwind part) solved on the device with BiCGSTAB, against the oracle's
    BiCGSTAB on the host-assembled one."""
    f, mf = _momentum_case(cases, (14, 12, 10))
    kw = dict(adv="minmod", td="linear")
    val_o, b_o, _ = oracle.momentum_assemble(f, mf, **kw)
    nat, val, b, _ = _momentum_gpu(gpu, f, mf, **kw)
    x = np.zeros(f.n)
    it, res = nat.solve_native("bcgs", "jacobi", val.copy(), x, b.copy(), 200, 1e-10, norm=1.0)
    A = oracle.Matrix(f.row, f.col, f.dia, val_o.copy())
    xo = np.zeros(f.n)
    it_o, res_o = oracle.solve("bcgs", "jacobi", A, xo, b_o.copy(), 200, 1e-10, norm=1.0)
    print(f"momentum solve: gpu {it} its res {res:.3e}; oracle {it_o} its res {res_o:.3e}",
          flush=True)
    assert res < 1e-10 and abs(it - it_o) <= 2
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_momentum_validation(cases, gpu):
    f, mf = _momentum_case(cases, (4, 3, 3))
    nat = gpu.Native(f.row, f.col, f.dia)
    faces = gpu.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, f.b_region)
    dval, db, dvm = (gpu.DeviceArray(np.zeros(k)) for k in (len(f.col), f.n, f.n))
    m = gpu.Momentum(mf)
    m.s.n_bnd = mf.n_bnd - 1                     # below the largest boundary cell
    with pytest.raises(gpu.TFlowsGpuError) as e:
        faces.momentum(m, dval, db, dvm)
    assert e.value.code == gpu.TFG_EINVAL
    m = gpu.Momentum(mf)
    m.s.adv_scheme = 10                          # not a face-value scheme
    with pytest.raises(gpu.TFlowsGpuError) as e:
        faces.momentum(m, dval, db, dvm)
    assert e.value.code == gpu.TFG_ENOTSUP
    m = gpu.Momentum(mf)
    m.s.vol = None
    with pytest.raises(gpu.TFlowsGpuError) as e:
        faces.momentum(m, dval, db, dvm)
    assert e.value.code == gpu.TFG_EINVAL


# ---- Rhie-Chow fluxes (tfg_rhie_chow_fluxes) -------------------------------

@pytest.mark.parametrize("choi,gu,vof,td", [(False, False, False, "linear"),
                                            (True, False, False, "linear"),
                                            (True, True, True, "parabolic"),
                                            (False, True, False, "linear"),
                                            (False, False, True, "linear")])
def test_rhie_chow_bitexact(cases, oracle, gpu, choi, gu, vof, td):
    """Every domain face's flux equals the serial loop's bit for bit; the
    boundary faces keep their values."""
    dims = (13, 11, 9)
    f = cases.hex_faces(*dims)
    rf = cases.rhie_chow_fields(f, *dims)
    w_o, w_oo = (1.0, 0.0) if td == "linear" else (2.0, -0.5)
    kw = dict(choi=choi, dt=3e-3, w_o=w_o, w_oo=w_oo, gu=gu, vof=vof, sigma=0.07)
    print(f"rhie-chow case {dims} {kw}", flush=True)
    base = np.random.default_rng(4).uniform(-1.0, 1.0, f.n_faces)
    ref = oracle.rhie_chow_fluxes(f, rf, base, **kw)
    nat = gpu.Native(f.row, f.col, f.dia)
    faces = gpu.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, f.b_region)
    dflux = gpu.DeviceArray(base)
    faces.rhie_chow(gpu.RhieChow(rf, **kw), dflux)
    assert np.array_equal(dflux.download(), ref)


def test_rhie_chow_then_pressure_system(cases, oracle, gpu):
    """Momentum v_m -> Rhie-Chow fluxes -> pressure system, all on the
    device, against the same chain of serial loops."""
    dims = (10, 9, 8)
    f = cases.hex_faces(*dims)
    mf = cases.momentum_fields(f)
    rf = cases.rhie_chow_fields(f, *dims)
    _, _, vm_o = oracle.momentum_assemble(f, mf)
    nat, _, _, vm = _momentum_gpu(gpu, f, mf)
    assert np.array_equal(vm, vm_o)
    import dataclasses
    rf = dataclasses.replace(rf, v_m=vm_o)
    flux_o = oracle.rhie_chow_fluxes(f, rf, f.vflux)
    g = dataclasses.replace(f, vm=vm_o, vflux=flux_o)
    val_o, b_o = oracle.pressure_assemble(g, singular=True)
    faces = gpu.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, f.b_region)
    dflux = gpu.DeviceArray(f.vflux)
    faces.rhie_chow(gpu.RhieChow(rf), dflux)
    assert np.array_equal(dflux.download(), flux_o)
    d = {k: gpu.DeviceArray(getattr(g, k)) for k in ("fc", "fw", "vm")}
    dval, db = gpu.DeviceArray(np.zeros(len(f.col))), gpu.DeviceArray(np.zeros(f.n))
    faces.assemble(d["fc"], d["fw"], d["vm"], dflux, dval, db, singular=True)
    assert np.array_equal(dval.download(), val_o) and np.array_equal(db.download(), b_o)


def test_rhie_chow_validation(cases, gpu):
    dims = (4, 3, 3)
    f = cases.hex_faces(*dims)
    rf = cases.rhie_chow_fields(f, *dims)
    nat = gpu.Native(f.row, f.col, f.dia)
    faces = gpu.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, f.b_region)
    rc = gpu.RhieChow(rf)
    rc.s.choi = 1                                # its arrays were not given
    with pytest.raises(gpu.TFlowsGpuError) as e:
        faces.rhie_chow(rc, gpu.DeviceArray(np.zeros(f.n_faces)))
    assert e.value.code == gpu.TFG_EINVAL
