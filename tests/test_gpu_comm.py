"""Multi-rank solves through the communication hooks (tfg_matrix_set_comm).

A global system is split into two ranks by x-planes, each with the
reference's layout: owned cells first, then the buffer cells its rows read
(ni < nt).  Two threads drive one handle each on the same GPU; the hooks
combine the two ranks' values in rank order (allreduce) and copy buffer
cells between the handles' device vectors (exchange).  The two ranks must
agree with each other and, together, with the oracle's one-domain solve at
the FP64 parity bar.  ICC is per rank in the reference (each rank's DIC over
its own rows), so its oracle is the global solve with the rows' ranks given:
per-rank DIC, and the reductions summed per rank, then over the ranks.
"""
import ctypes
import threading

import numpy as np
import pytest

import parity_bar
import ranks as rank_split

pytestmark = pytest.mark.gpu


class _Comm:
    """Two in-process ranks: rank-ordered allreduce, buffer-cell exchange."""

    def __init__(self, gpu, ranks):
        self.gpu, self.ranks = gpu, ranks
        self.bar = threading.Barrier(2, timeout=120)
        self.slot = [None, None]
        self.owned = [None, None]

    def allreduce(self, r, values, ops):
        self.slot[r] = values.copy()
        self.bar.wait()
        a, b = self.slot
        out = np.array([a[i] + b[i] if o == self.gpu.SUM else
                        (min(a[i], b[i]) if o == self.gpu.MIN else max(a[i], b[i]))
                        for i, o in enumerate(ops)])
        self.bar.wait()
        values[:] = out

    def exchange(self, r, d_x):
        me = self.ranks[r]
        own = np.zeros(me["ni"])
        L = self.gpu.lib()
        assert L.tfg_memcpy_d2h(own.ctypes.data, ctypes.c_void_p(d_x), own.nbytes) == 0
        self.owned[r] = dict(zip(me["glob"][:me["ni"]], own))
        self.bar.wait()
        other = self.owned[1 - r]
        buf = np.array([other[g] for g in me["glob"][me["ni"]:]])
        self.bar.wait()
        if len(buf):
            assert L.tfg_memcpy_h2d(ctypes.c_void_p(d_x + 8 * me["ni"]), buf.ctypes.data,
                                    buf.nbytes) == 0


def _run(gpu, ranks, prec, miter, tol, norm, criterion, solver="cg"):
    comm = _Comm(gpu, ranks)
    out = [None, None]

    def work(r):
        me = ranks[r]
        nat = gpu.Native(me["row"], me["col"], me["dia"], ni=me["ni"])
        nat.set_comm(lambda v, o: comm.allreduce(r, v, o), lambda d: comm.exchange(r, d))
        x = np.zeros(me["nt"])
        opt = gpu.Options()
        opt.criterion = criterion
        try:
            it, res = nat.solve_native(solver, prec, me["val"].copy(), x, me["b"].copy(), miter,
                                       tol, norm=norm, options=opt)
            out[r] = (it, res, x)
        except Exception as e:   # noqa: BLE001 - reported by the caller
            out[r] = e
            comm.bar.abort()

    th = [threading.Thread(target=work, args=(r,)) for r in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(300)
    return out


@pytest.mark.parametrize("prec,criterion,norm", [("jacobi", 0, None), ("jacobi", 1, None),
                                                 ("none", 0, 1.0)])
def test_two_ranks_match_one_domain(cases, oracle, gpu, prec, criterion, norm):
    nx = 12
    s = cases.hex_pressure(nx, 10, 8, rho_out=100.0)
    ranks = rank_split.split(s, nx)
    print(f"two ranks: ni {ranks[0]['ni']}/{ranks[1]['ni']}, nt {ranks[0]['nt']}/{ranks[1]['nt']}",
          flush=True)
    out = _run(gpu, ranks, prec, 2000, 1e-8, norm, criterion)
    for o in out:
        assert not isinstance(o, Exception), o
    (it0, res0, x0), (it1, res1, x1) = out
    assert it0 == it1 and res0 == res1
    x = np.zeros(s.n)
    for rk, xr in zip(ranks, (x0, x1)):
        x[rk["glob"][:rk["ni"]]] = xr[:rk["ni"]]
    cname = "nrms" if criterion == 0 else "relres"

    def solve(so):
        A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
        xo = np.zeros(s.n)
        it, r = oracle.solve("cg", prec, A, xo, s.b.copy(), 2000, 1e-8, norm=norm, criterion=cname,
                             sum_order=so)
        return it, r, xo

    it_o, res_o, x_o = solve(0)
    env, spread = parity_bar.envelope(lambda so: solve(so)[::2], it_o, x_o)
    print(f"  ranks niter {it0} res {res0:.3e}; oracle niter {it_o} res {res_o:.3e}", flush=True)
    parity_bar.check(f"two ranks {prec} {cname} norm={norm}", it0, x, it_o, x_o, env, spread)
    assert res0 < 1e-8
    # buffer cells of x hold the owners' final values
    for rk, xr in zip(ranks, (x0, x1)):
        assert np.array_equal(xr[rk["ni"]:], x[rk["glob"][rk["ni"]:]])


@pytest.mark.parametrize("solver,criterion,norm", [("cg", 0, None), ("cg", 1, None),
                                                   ("bicg", 0, 1.0)])
def test_two_ranks_icc_match_per_rank_oracle(cases, oracle, gpu, solver, criterion, norm):
    """Prec_Form.f90:68-78 / Prec_Solve.f90:64-92 on each rank's own rows."""
    nx = 12
    s = cases.hex_pressure(nx, 10, 8, rho_out=100.0)
    ranks = rank_split.split(s, nx, local_order=True)
    parts = rank_split.parts(ranks)
    print(f"two ranks icc {solver}: ni {ranks[0]['ni']}/{ranks[1]['ni']}, "
          f"nt {ranks[0]['nt']}/{ranks[1]['nt']}", flush=True)
    out = _run(gpu, ranks, "icc", 2000, 1e-8, norm, criterion, solver)
    for o in out:
        assert not isinstance(o, Exception), o
    (it0, res0, x0), (it1, res1, x1) = out
    assert it0 == it1 and res0 == res1
    x = np.zeros(s.n)
    for rk, xr in zip(ranks, (x0, x1)):
        x[rk["glob"][:rk["ni"]]] = xr[:rk["ni"]]
    cname = "nrms" if criterion == 0 else "relres"

    def solve(so, per_rank=True):
        A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
        xo = np.zeros(s.n)
        it, r = oracle.solve(solver, "icc", A, xo, s.b.copy(), 2000, 1e-8, norm=norm,
                             criterion=cname, sum_order=so, parts=parts if per_rank else None)
        return it, r, xo

    it_o, res_o, x_o = solve(11)     # each rank's serial sums, added in rank order
    env, spread = parity_bar.envelope(lambda so: solve(so)[::2], it_o, x_o)
    print(f"  ranks niter {it0} res {res0:.3e}; oracle niter {it_o} res {res_o:.3e}", flush=True)
    err = parity_bar.check(f"two ranks icc {solver} {cname} norm={norm}", it0, x, it_o, x_o, env,
                           spread)
    assert res0 < 1e-8
    # the one-domain DIC is another preconditioner: its run lies outside the bar
    it_1, _, x_1 = solve(0, per_rank=False)
    far = parity_bar.rel(x, x_1)
    print(f"  one-domain icc: niter {it_1}, x differs by {far:.2e} (per-rank err {err:.2e})",
          flush=True)
    assert (abs(it_1 - it0) > max(2, parity_bar.FACTOR * spread)
            or far > max(parity_bar.X_TOL, parity_bar.FACTOR * env))
    for rk, xr in zip(ranks, (x0, x1)):
        assert np.array_equal(xr[rk["ni"]:], x[rk["glob"][rk["ni"]:]])


def test_buffer_cells_need_hooks(cases, gpu):
    s = cases.hex_pressure(6, 4, 4)
    rk = rank_split.split(s, 6)[0]
    nat = gpu.Native(rk["row"], rk["col"], rk["dia"], ni=rk["ni"])
    with pytest.raises(gpu.TFlowsGpuError) as e:
        nat.solve_native("cg", "jacobi", rk["val"].copy(), np.zeros(rk["nt"]), rk["b"].copy(), 10,
                         1e-8)
    assert e.value.code == gpu.TFG_ENOTSUP
