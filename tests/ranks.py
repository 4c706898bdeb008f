"""Two MPI ranks of a global hex system, in the layout a rank's matrix has
in the reference: owned cells first (ni rows), then the buffer cells its
rows read (rows ni+1..nt, never solved; here identity rows).

The global system is ordered x slowest, so cutting it between two x-planes
gives each rank a contiguous run of global rows.  With local_order=True each
local row holds its diagonal first and its other columns ascending in the
rank's own numbering, as Create_Matrix.f90:87-119 builds a rank's matrix
(the order the DIC recurrences of Prec_Form / Prec_Solve rely on); with
local_order=False the columns keep the global order.
"""
import numpy as np


def split(s, nx, local_order=False):
    """Per rank: dict(ni, nt, row, col, dia, val, b, glob), 1-based CSR,
    glob = global ids of the local cells (owned, then buffer)."""
    n = s.n
    n0 = n * (nx // 2) // nx
    row0 = s.row - 1
    ranks = []
    for lo, hi in ((0, n0), (n0, n)):
        cols = np.concatenate([s.col[row0[i]:row0[i + 1]] - 1 for i in range(lo, hi)])
        buf = np.unique(cols[(cols < lo) | (cols >= hi)])
        glob = np.concatenate([np.arange(lo, hi), buf])
        loc = {int(g): k for k, g in enumerate(glob)}
        ni, nt = hi - lo, len(glob)
        row, col, val = [1], [], []
        for i in range(lo, hi):
            ent = [(loc[int(s.col[j]) - 1], s.val[j]) for j in range(row0[i], row0[i + 1])]
            if local_order:
                ent = ent[:1] + sorted(ent[1:])   # diagonal first (global rows have it first)
            for c, v in ent:
                col.append(c + 1)
                val.append(v)
            row.append(len(col) + 1)
        for k in range(ni, nt):          # buffer rows: never solved
            col.append(k + 1)
            val.append(1.0)
            row.append(len(col) + 1)
        row = np.array(row, dtype=np.int32)
        b = np.zeros(nt)
        b[:ni] = s.b[lo:hi]
        ranks.append(dict(ni=ni, nt=nt, row=row, col=np.array(col, dtype=np.int32),
                          dia=row[:-1].copy(), val=np.array(val), b=b, glob=glob))
    return ranks


def parts(ranks):
    """Rank of every global row (the ranks' owned rows are contiguous)."""
    return np.concatenate([np.full(rk["ni"], r, dtype=np.int32) for r, rk in enumerate(ranks)])
