"""Writes the committed golden fixtures (tests/golden/*.npz) from the FP64
CPU oracle.  The reference ships no golden data for this path; these vectors
freeze the oracle's output on small seeded inputs so that GPU runs (which do
not rebuild the oracle's inputs) and later changes are checked against the
same numbers.  Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "magnetic-force_amd"))

import oracle as O  # noqa: E402
from magforce import configs  # noqa: E402


def f32(a):
    """Inputs are what the FP32 API receives: round them once to float32."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def case_products():
    pos = f32(configs.jittered_lattice(5, 6, 7, seed=11))          # n = 210, not a tile multiple
    rng = np.random.default_rng(12)
    mom = f32(rng.normal(size=pos.shape) * 0.5)
    h = 2.0
    return dict(pos=pos, mom=mom, h=h,
                field=O.field_allpairs(pos, pos, mom, h),
                force=O.force_allpairs(pos, mom, h),
                hessian=O.hessian_allpairs(pos, pos, mom, h))


def case_solve():
    pos = f32(configs.jittered_lattice(6, 6, 6, seed=13))
    h, chi = 2.0, 1.0
    kappa = chi / (1 + chi)
    n = pos.shape[0]
    b = np.tile([0.0, 0.0, 1.0], (n, 1))
    A = np.eye(3 * n) - kappa * O.dense_G(pos, h)
    x = np.linalg.solve(A, b.ravel()).reshape(n, 3)
    F = O.force_allpairs(pos, kappa * x, h)
    return dict(pos=pos, h=h, chi=chi, b=b, x=x, force=F)


def case_example():
    """The reference example's second solve (main.cu:21-61): 2x2x1 particles
    at z = 1, one external dipole (0,1,0) at the origin, h = ext h = 1,
    V = 1, chi = 1, H_const = (0,1,0)."""
    pos = np.array([[i, j, 1.0] for i in range(2) for j in range(2)], dtype=np.float64)
    ext_pos = np.zeros((1, 3))
    ext_m = np.array([[0.0, 1.0, 0.0]])
    h = 1.0
    hconst = np.array([0.0, 1.0, 0.0])
    hext = O.field_allpairs(pos, ext_pos, ext_m, h, hconst)
    text = O.hessian_allpairs(pos, ext_pos, ext_m, h)
    kappa = 0.5
    n = pos.shape[0]
    A = np.eye(3 * n) - kappa * O.dense_G(pos, h)
    x = np.linalg.solve(A, hext.ravel()).reshape(n, 3)
    F = O.force_allpairs(pos, kappa * x, h, text=text)
    return dict(pos=pos, ext_pos=ext_pos, ext_m=ext_m, h=h, hconst=hconst, hext=hext,
                text=text, x=x, force=F)


def main():
    for name, fn in [("products", case_products), ("solve", case_solve),
                     ("example", case_example)]:
        np.savez(os.path.join(HERE, f"{name}.npz"), **fn())
        print("wrote", name)


if __name__ == "__main__":
    main()
