"""Writes the golden fixtures of the H2 action (regression pins).

Inputs are seeded (numpy default_rng); outputs come from the FP64 oracle
(oracle/h2_oracle.cpp, the C++ restatement of FMCA's matrix-free action) and,
for the H2 error, from its dense K*x.  FMCA itself cannot be built here (its
Eigen3 dependency is absent), so these pin this build's planner and action
against later changes; they are not FMCA outputs.

  python tests/golden/make_golden.py [name ...]   (rewrites tests/golden/<name>.npz,
                                                   all cases without names)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

# name, dim, n, m (None: square), kernel, l, deg, eta, nrhs, seed, interpolator
CASES = [
    ("exp3d_n4000_deg2", 3, 4000, None, "EXPONENTIAL", 0.5, 2, 0.8, 2, 101, "tensor_product"),
    ("matern32_2d_n2000_deg4", 2, 2000, None, "MATERN32", 0.3, 4, 0.8, 3, 102, "tensor_product"),
    ("gauss3d_rect_n3000_m2500_deg2", 3, 3000, 2500, "GAUSSIAN", 0.2, 2, 0.8, 2, 103,
     "tensor_product"),
    ("td_matern32_3d_n4000_deg4", 3, 4000, None, "MATERN32", 0.4, 4, 0.8, 3, 104, "total_degree"),
    ("td_exp2d_rect_n3000_m2000_deg4", 2, 3000, 2000, "EXPONENTIAL", 1.0, 4, 0.8, 2, 105,
     "total_degree"),
]
SCHEME = {"tensor_product": "tensor", "total_degree": "total_degree"}


def make(case):
    name, dim, n, m, kernel, l, deg, eta, nrhs, seed, interpolator = case
    rng = np.random.default_rng(seed)
    P = np.asfortranarray(rng.random((dim, n)))
    Pe = P if m is None else np.asfortranarray(rng.random((dim, m)))
    x = np.asfortranarray(rng.uniform(-1, 1, (n, nrhs)))
    A = oracle.OracleH2(P, Pe, kernel, l, deg, eta, scheme=SCHEME[interpolator])
    row, col, lr = A.blocks()
    extra = {}
    if interpolator == "total_degree":
        _, Xi, _, invV = oracle.interp_total_degree(dim, deg)
        extra = dict(Xi=Xi, invV=invV)
    return dict(
        **extra, interpolator=interpolator,
        dim=dim, kernel=kernel, l=l, deg=deg, eta=eta, P=P, Peval=Pe, rhs=x,
        rindices=A.indices(0), cindices=A.indices(1),
        block_row=row, block_col=col, block_lr=lr,
        action=A.action(x),
        dense=oracle.dense_action(P, Pe, kernel, l, x),
    )


def main():
    oracle.build()
    names = sys.argv[1:]
    for case in CASES:
        if names and case[0] not in names:
            continue
        d = make(case)
        path = os.path.join(HERE, case[0] + ".npz")
        np.savez_compressed(path, **d)
        err = np.linalg.norm(d["action"] - d["dense"]) / np.linalg.norm(d["dense"])
        print(f"{path}: {len(d['block_row'])} blocks ({int(d['block_lr'].sum())} low-rank), "
              f"H2 error vs dense {err:.3e}")


if __name__ == "__main__":
    main()
