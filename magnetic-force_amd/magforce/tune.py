"""Autotuner for FMM plans.

Chooses the expansion order (6, 8 or 10), the M2L precision (BF16x3, BF16x6,
FP32) and the leaf side (in units of h, >= 2) for a particle set by
measurement: every candidate plan runs the field product on seeded random
moments, its time is taken with HIP events and its error is scored on a
seeded subsample against the FP64 all-pairs field (all sources).  The fastest
candidate within the error target wins.  The leaf side trades near-field
pairs (its cube) against translations (its inverse cube); the reference fixes
its leaf budget at ncrit = 384 (magnetic_force.cu:647-648) and P = 6
(:661).  Every product runs in the library's own kernels; the FP64 reference
is the GPU FP64 check path of the same library.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import (M2L_BF16X1, M2L_BF16X1_CUT, M2L_BF16X3, M2L_BF16X6, M2L_FP32, MODE_FMM,
               Context, Plan, field_allpairs_f64)

PRECISION_NAMES = {M2L_FP32: "fp32", M2L_BF16X6: "bf16x6", M2L_BF16X3: "bf16x3",
                   M2L_BF16X1: "bf16x1", M2L_BF16X1_CUT: "bf16x1cut"}


@dataclass
class Candidate:
    order: int
    m2l_precision: int
    ms_per_product: float
    rel_err: float
    leaf_side: float = 2.0

    def as_dict(self):
        return {"order": self.order, "m2l": PRECISION_NAMES[self.m2l_precision],
                "leaf_side": self.leaf_side, "ms_per_product": self.ms_per_product,
                "rel_err": self.rel_err}


def autotune_fmm(ctx: Context, pos, h: float, target_rel_err: float = 1e-5,
                 samples: int = 1024, reps: int = 2, seed: int = 0,
                 orders=(6, 8, 10), precisions=(M2L_BF16X3, M2L_BF16X6, M2L_FP32),
                 leaf_sides=(2.0,)):
    """Returns (best Candidate or None, list of all Candidates).

    pos: (3, n) float32 CUDA tensor.  The error is the relative L2 error of the
    field product over `samples` seeded particles."""
    import torch
    n = pos.shape[1]
    rng = np.random.default_rng(seed)
    m = torch.tensor(rng.normal(size=(3, n)) * 0.5, dtype=torch.float32, device=pos.device)
    k = min(samples, n)
    idx = torch.tensor(np.sort(rng.choice(n, size=k, replace=False)), dtype=torch.int64,
                       device=pos.device)
    pos64 = pos.double()
    ref = field_allpairs_f64(ctx, pos64[:, idx].contiguous(), pos64, m.double(), h)
    out = torch.empty_like(m)
    cands = []
    for side in leaf_sides:
        for order in orders:
            for prec in precisions:
                plan = Plan(ctx, pos, h, mode=MODE_FMM, order=order, m2l_precision=prec,
                            leaf_side=side)
                plan.field(m, out=out)  # warm-up (first launch of each kernel)
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    plan.field(m, out=out)
                e1.record()
                torch.cuda.synchronize()
                got = out[:, idx].double()
                err = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
                cands.append(Candidate(order, prec, e0.elapsed_time(e1) / reps, err, float(side)))
                plan.close()
    ok = [c for c in cands if c.rel_err <= target_rel_err]
    best = min(ok, key=lambda c: c.ms_per_product) if ok else None
    return best, cands
