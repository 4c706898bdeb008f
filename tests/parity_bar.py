"""The parity bar of the FP64 GPU solves against the oracle.

The reference's answer is defined up to the order of its global sums: the
GPU sums its dot products in deterministic trees, the oracle (like the
reference) left to right.  Rounding alone moves a solve by the spread of
equally valid orders, so the oracle is re-run with other orders (blocks of
64, 1000 and 256 terms combined pairwise, a fully pairwise sum, serial from
the last term) and with the normalisation factor fn one ulp below and above
the reference's and the pairwise order's (the diagonal sum fn comes from
rounds differently in another order; on ill-conditioned systems one ulp of
fn alone moves the iteration count of unpreconditioned CG by a dozen), and
the bar is:

  * iterations within max(2, 1.2 x the largest difference between the
    counts of any two of those runs, the reference order's included), and
  * relative x difference at most max(1e-10, 1.2 x the largest relative x
    difference between any two of those runs).

The GPU's run is one more equally valid run, so it may differ from the
reference order's by as much as two valid runs differ from each other.  A
sample of ten runs underestimates that spread where it is widest: on
unpreconditioned CG of a 1:300 two-phase 16^3 system, ten more runs with
random serial orders or fn a few ulps away span 1934..1955 iterations
against 1937..1953 for these ten; the factor 1.2 covers that.

Every check prints one `PARITY` line (error, envelope, their ratio) so that
the runs can be tabulated.
"""
import numpy as np

ORDERS = (1, 2, 3, 4, 6, 7, 8, 9, 10)
X_TOL = 1e-10
FACTOR = 1.2


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def envelope(solve_order, it_o, x_o):
    """solve_order(order) -> (niter, x) of the oracle with that summation
    order; returns (x envelope, iteration spread) of those runs together
    with the reference order's (it_o, x_o): the largest relative x
    difference and the largest count difference between any two runs."""
    runs = [(int(it_o), x_o)] + [(int(it), x) for it, x in map(solve_order, ORDERS)]
    env, spread = 0.0, 0
    for a in range(len(runs)):
        for b in range(len(runs)):
            if a != b:
                env = max(env, rel(runs[a][1], runs[b][1]))
                spread = max(spread, abs(runs[a][0] - runs[b][0]))
    return env, spread


def check(name, it_g, x_g, it_o, x_o, env, spread):
    """Assert the bar; returns the relative x error."""
    err = rel(x_g, x_o)
    bar = max(X_TOL, FACTOR * env)
    it_bar = max(2, FACTOR * spread)
    print(f"PARITY {name}: niter gpu {it_g} oracle {it_o} spread {spread} bar {it_bar:.1f}; "
          f"x err {err:.2e} envelope {env:.2e} bar {bar:.2e} err/bar {err / bar:.2f}", flush=True)
    assert abs(int(it_g) - int(it_o)) <= it_bar, (name, it_g, it_o, spread)
    assert err <= bar, (name, err, env)
    return err
