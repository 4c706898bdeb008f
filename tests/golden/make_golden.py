"""Writes tests/golden/cg_small.json: reference answers for small systems
computed with numpy/scipy only (no oracle code involved).

For each case: the generated system's size facts, the k-th iterate of
exact-arithmetic preconditioned CG (Galerkin projection onto the Krylov
space, dense QR) and the dense direct solution.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "t-flows_amd"))
import tflows_cases  # noqa: E402


def galerkin(A, Minv, b, k):
    v = Minv @ b
    K = []
    for _ in range(k):
        K.append(v)
        v = Minv @ (A @ v)
    Q, _ = np.linalg.qr(np.array(K).T)
    return Q @ np.linalg.solve(Q.T @ A @ Q, Q.T @ b)


def main():
    out = {"generator": "tflows_cases.hex_pressure", "cases": []}
    for grid, rho, seed, prec, k in [((4, 4, 4), 1.0, 11, "jacobi", 4),
                                     ((5, 4, 3), 1000.0, 12, "jacobi", 5),
                                     ((3, 5, 4), 30.0, 13, "none", 3)]:
        s = tflows_cases.hex_pressure(*grid, rho_out=rho, seed=seed)
        A = s.scipy().toarray()
        Minv = np.diag(1.0 / np.diag(A)) if prec == "jacobi" else np.eye(s.n)
        out["cases"].append({
            "grid": list(grid), "rho_out": rho, "seed": seed, "prec": prec, "miter": k,
            "nnz": int(s.nnz), "val_sum": float(np.sum(s.val)),
            "x_galerkin": galerkin(A, Minv, s.b, k).tolist(),
            "x_solve": np.linalg.solve(A, s.b).tolist(),
        })
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cg_small.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
