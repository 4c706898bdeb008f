"""Multi-rank solves through the communication hooks (tfg_matrix_set_comm).

A global system is split into two ranks by x-planes, each with the
reference's layout: owned cells first, then the buffer cells its rows read
(ni < nt).  Two threads drive one handle each on the same GPU; the hooks
combine the two ranks' values in rank order (allreduce) and copy buffer
cells between the handles' device vectors (exchange).  The two ranks must
agree with each other and, together, with the oracle's one-domain solve at
the FP64 parity bar.
"""
import ctypes
import threading

import numpy as np
import pytest

import parity_bar

pytestmark = pytest.mark.gpu


def _split(s, nx):
    """Two ranks of a hex system ordered x slowest: (ni, row, col, dia, val,
    b, glob) per rank, glob = global ids of the local cells."""
    n = s.n
    n0 = n * (nx // 2) // nx
    row0 = s.row - 1
    ranks = []
    for r, (lo, hi) in enumerate(((0, n0), (n0, n))):
        cols = []
        for i in range(lo, hi):
            cols.extend(s.col[row0[i]:row0[i + 1]] - 1)
        cols = np.array(cols)
        buf = np.unique(cols[(cols < lo) | (cols >= hi)])
        glob = np.concatenate([np.arange(lo, hi), buf])
        loc = {g: k for k, g in enumerate(glob)}
        ni, nt = hi - lo, len(glob)
        row, col, val = [1], [], []
        for i in range(lo, hi):
            for j in range(row0[i], row0[i + 1]):
                col.append(loc[s.col[j] - 1] + 1)
                val.append(s.val[j])
            row.append(len(col) + 1)
        for k in range(ni, nt):          # buffer rows: never solved
            col.append(k + 1)
            val.append(1.0)
            row.append(len(col) + 1)
        row = np.array(row, dtype=np.int32)
        col = np.array(col, dtype=np.int32)
        b = np.zeros(nt)
        b[:ni] = s.b[lo:hi]
        ranks.append(dict(ni=ni, nt=nt, row=row, col=col, dia=row[:-1].copy(),
                          val=np.array(val), b=b, glob=glob))
    return ranks


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


def _run(gpu, ranks, prec, miter, tol, norm, criterion):
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
            it, res = nat.solve_native("cg", prec, me["val"].copy(), x, me["b"].copy(), miter, tol,
                                       norm=norm, options=opt)
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
    ranks = _split(s, nx)
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


def test_buffer_cells_need_hooks(cases, gpu):
    s = cases.hex_pressure(6, 4, 4)
    rk = _split(s, 6)[0]
    nat = gpu.Native(rk["row"], rk["col"], rk["dia"], ni=rk["ni"])
    with pytest.raises(gpu.TFlowsGpuError) as e:
        nat.solve_native("cg", "jacobi", rk["val"].copy(), np.zeros(rk["nt"]), rk["b"].copy(), 10,
                         1e-8)
    assert e.value.code == gpu.TFG_ENOTSUP
