"""The parity bar of the FP64 GPU solves against the oracle.

The reference's answer is defined up to the order of its global sums: the
GPU sums its dot products in deterministic trees, the oracle (like the
reference) left to right.  Rounding alone moves a solve by the spread of
equally valid orders, so the oracle is re-run with three other orders
(blocks of 64 and of 1000 combined pairwise, and serial from the last term)
and the bar is:

  * iterations within max(2, the largest spread of those orders), and
  * relative x difference at most max(1e-10, the largest x difference of
    those orders from the reference order).

Every check prints one `PARITY` line (error, envelope, their ratio) so that
the runs can be tabulated.
"""
import numpy as np

ORDERS = (1, 2, 3)
X_TOL = 1e-10


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def envelope(solve_order, it_o, x_o):
    """solve_order(order) -> (niter, x) of the oracle with that summation
    order; returns (x envelope, iteration spread) against (it_o, x_o)."""
    env, spread = 0.0, 0
    for so in ORDERS:
        it, x = solve_order(so)
        env = max(env, rel(x, x_o))
        spread = max(spread, abs(int(it) - int(it_o)))
    return env, spread


def check(name, it_g, x_g, it_o, x_o, env, spread):
    """Assert the bar; returns the relative x error."""
    err = rel(x_g, x_o)
    bar = max(X_TOL, env)
    print(f"PARITY {name}: niter gpu {it_g} oracle {it_o} spread {spread}; x err {err:.2e} "
          f"envelope {env:.2e} bar {bar:.2e} err/bar {err / bar:.2f}", flush=True)
    assert abs(int(it_g) - int(it_o)) <= max(2, spread), (name, it_g, it_o, spread)
    assert err <= bar, (name, err, env)
    return err
