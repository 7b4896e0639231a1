"""GPU magnetization solve: CG, GMRES and the reference's nonlinear CG against
the oracle solution (dense LU / FP64 CG to 1e-12), and FP64 refinement."""
import os

import numpy as np
import pytest

from conftest import jittered_block, rel_err

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dev(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(np.asarray(a).T),
                        dtype=dtype or torch.float32, device="cuda").contiguous()


def host(t):
    return t.double().cpu().numpy().T


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


@pytest.mark.parametrize("method", [0, 1])
def test_solve_golden(ctx, method):
    import magforce
    g = np.load(os.path.join(GOLD, "solve.npz"))
    kappa = float(g["chi"]) / (1 + float(g["chi"]))
    plan = magforce.Plan(ctx, dev(g["pos"]), float(g["h"]))
    res = plan.solve(kappa, dev(g["b"]), method=method, rtol=1e-6, true_residual=True)
    assert res.stats["converged"] == 1
    assert res.stats["rel_residual_true"] < 2e-6
    x = host(res.x)
    assert rel_err(x, g["x"]) < 5e-6
    F = host(plan.force(dev(kappa * x)))
    assert rel_err(F, g["force"]) < 1e-5


def test_lattice8k_solve_and_force_vs_oracle(ctx, oracle):
    """The 8192-particle workload: CG to 1e-6 relative residual, then the
    force, scored against the FP64 oracle solved to 1e-12."""
    import magforce
    from magforce import configs
    s = configs.lattice8k()
    pos = s.pos.astype(np.float64)
    b = np.tile(s.hconst, (s.n, 1))
    x_ref, it, rel = oracle.cg_solve(pos, s.h, s.kappa, b, rtol=1e-12)
    assert rel < 1e-12
    F_ref = oracle.force_allpairs(pos, s.kappa * x_ref, s.h)
    plan = magforce.Plan(ctx, dev(pos), s.h)
    res = plan.solve(s.kappa, dev(b), rtol=1e-6)
    assert res.stats["converged"] == 1 and res.stats["matvecs"] < 40
    x = host(res.x)
    assert rel_err(x, x_ref) < 5e-6
    F = host(plan.force(dev(s.kappa * x)))
    assert rel_err(F, F_ref) < 1e-5


def test_gmres_restart_and_warm_start(ctx, oracle):
    import magforce
    pos = f32(jittered_block(7, 7, 5, seed=3))
    n = pos.shape[0]
    kappa = 2.0 / 3.0
    b = np.tile([0.3, -0.2, 1.0], (n, 1))
    x_ref, _, _ = oracle.cg_solve(pos, 2.0, kappa, b, rtol=1e-12)
    plan = magforce.Plan(ctx, dev(pos), 2.0)
    res = plan.solve(kappa, dev(b), method=magforce.SOLVE_GMRES, restart=3, rtol=1e-6)
    assert res.stats["converged"] == 1
    assert rel_err(host(res.x), x_ref) < 5e-6
    # warm start from the solution: converges immediately
    res2 = plan.solve(kappa, dev(b), x0=res.x, rtol=1e-5)
    assert res2.stats["iterations"] <= 1


def test_reference_nlcg_on_device(ctx, oracle):
    """The reference's own nonlinear CG (tolerance on RMS residual)."""
    import magforce
    pos = f32(jittered_block(5, 5, 4, seed=9))
    n = pos.shape[0]
    b = np.tile([0.0, 1.0, 0.0], (n, 1))
    x_ref, _, _ = oracle.cg_solve(pos, 2.0, 0.5, b, rtol=1e-12)
    plan = magforce.Plan(ctx, dev(pos), 2.0)
    res = plan.solve(0.5, dev(b), method=magforce.SOLVE_REF_NLCG, rtol=1e-4, maxit=80)
    assert res.stats["converged"] == 1
    assert rel_err(host(res.x), x_ref) < 1e-3


def test_fp64_refinement_reaches_1e12(ctx, oracle):
    import magforce
    import torch
    pos = f32(jittered_block(8, 8, 6, seed=12))
    n = pos.shape[0]
    b = np.tile([0.0, 0.0, 1.0], (n, 1))
    x_ref, _, _ = oracle.cg_solve(pos, 2.0, 0.5, b, rtol=1e-13)
    plan = magforce.Plan(ctx, dev(pos), 2.0)
    res = plan.solve_f64(0.5, dev(b, torch.float64), rtol=1e-12)
    assert res.stats["converged"] == 1
    assert res.stats["rel_residual_true"] < 1e-12
    assert rel_err(host(res.x), x_ref) < 1e-11


@pytest.mark.parametrize("outer", ["bf16x3", "fp32"])
def test_fmm_refinement_vs_oracle(ctx, oracle, outer):
    """Refinement on an FMM plan: x and r = b - A x in FP64, the residual from
    the plan's FMM product (BF16x3 or FP32 translations), the corrections
    from CG with one-part BF16 translations (about 4e-3 per translation) to
    1e-3.  It lands on the FMM operator's solution, within the p = 10
    truncation of the FP64 solution: the same bar as the plain FMM solve."""
    import magforce
    import torch
    from magforce import configs
    s = configs.lattice8k()
    pos = s.pos.astype(np.float64)
    b = np.tile(s.hconst, (s.n, 1))
    x_ref, _, _ = oracle.cg_solve(pos, s.h, s.kappa, b, rtol=1e-12)
    F_ref = oracle.force_allpairs(pos, s.kappa * x_ref, s.h)
    code = {"bf16x3": magforce.M2L_BF16X3, "fp32": magforce.M2L_FP32}[outer]
    plan = magforce.Plan(ctx, dev(pos), s.h, mode=magforce.MODE_FMM, m2l_precision=code)
    res = plan.solve_f64(s.kappa, dev(b, torch.float64), rtol=1e-6,
                         inner_m2l_precision=magforce.M2L_BF16X1, inner_rtol=1e-3)
    st = res.stats
    print(f"fmm refinement ({outer}): {st}")
    assert st["converged"] == 1 and st["refinements"] >= 1
    assert st["rel_residual_true"] <= 1e-6
    x = host(res.x)
    assert rel_err(x, x_ref) < 5e-6
    F = host(plan.force(dev(s.kappa * x)))
    assert rel_err(F, F_ref) < 1e-5


@pytest.mark.parametrize("relax", ["bf16x1", "bf16x1cut"])
def test_relaxed_cg_vs_oracle(ctx, oracle, relax):
    """CG with relaxed products: once the relative residual falls to 3e-3 the
    remaining products run with one-part BF16 translations (a few 1e-4 of
    the field), or with their degree cut.  An error e made at residual r
    moves the final residual by about e r, so the solution stays on the
    BF16x3 solve's accuracy (within the p = 10 truncation of the FP64
    oracle), and the force with it.  A relax threshold at the tolerance
    itself relaxes nothing, and a plan without a far field (all-pairs)
    ignores the option."""
    import magforce
    from magforce import configs
    code = {"bf16x1": magforce.M2L_BF16X1, "bf16x1cut": magforce.M2L_BF16X1_CUT}[relax]
    s = configs.lattice8k()
    pos = s.pos.astype(np.float64)
    b = np.tile(s.hconst, (s.n, 1))
    x_ref, _, _ = oracle.cg_solve(pos, s.h, s.kappa, b, rtol=1e-12)
    F_ref = oracle.force_allpairs(pos, s.kappa * x_ref, s.h)
    plan = magforce.Plan(ctx, dev(pos), s.h, mode=magforce.MODE_FMM,
                         m2l_precision=magforce.M2L_BF16X3)
    full = plan.solve(s.kappa, dev(b), rtol=1e-6)
    g_before = host(plan.field(full.x))
    res = plan.solve(s.kappa, dev(b), rtol=1e-6, relax_m2l_precision=code,
                     relax_rtol=3e-3, true_residual=True)
    st = res.stats
    print(f"relaxed cg ({relax}): {st} (plain: {full.stats})")
    assert st["converged"] == 1 and st["relaxed_matvecs"] >= 1
    assert st["relaxed_matvecs"] < st["matvecs"]
    assert st["rel_residual_true"] <= 2e-6
    x = host(res.x)
    assert rel_err(x, x_ref) < 5e-6
    assert rel_err(x, host(full.x)) < 2e-6
    F = host(plan.force(dev(s.kappa * x)))
    assert rel_err(F, F_ref) < 1e-5
    # the plan's own precision is back after the solve
    assert rel_err(host(plan.field(full.x)), g_before) == 0.0
    none = plan.solve(s.kappa, dev(b), rtol=1e-6, relax_m2l_precision=code,
                      relax_rtol=1e-7)
    assert none.stats["relaxed_matvecs"] == 0
    assert rel_err(host(none.x), host(full.x)) == 0.0
    ap = magforce.Plan(ctx, dev(pos), s.h)
    res_ap = ap.solve(s.kappa, dev(b), rtol=1e-6, relax_m2l_precision=code)
    assert res_ap.stats["relaxed_matvecs"] == 0


def test_bf16x1_product_accuracy(ctx, oracle):
    """The one-part BF16 translations alone: a few 1e-3 of the far field, so
    the product stays within 1e-2 of the oracle (a cheap, inexact operator for
    inner corrections only)."""
    import magforce
    from magforce import configs
    s = configs.lattice8k()
    pos = s.pos.astype(np.float64)
    rng = np.random.default_rng(5)
    m = f32(rng.normal(size=pos.shape) * 0.4)
    plan = magforce.Plan(ctx, dev(pos), s.h, mode=magforce.MODE_FMM,
                         m2l_precision=magforce.M2L_BF16X1)
    err = rel_err(host(plan.field(dev(m))), oracle.field_allpairs(pos, pos, m, s.h))
    print(f"bf16x1 field error {err:.3e}")
    assert 1e-6 < err < 1e-2


def test_empty_system(ctx):
    import magforce
    import torch
    pos = torch.zeros((3, 0), dtype=torch.float32, device="cuda")
    plan = magforce.Plan(ctx, pos, 2.0)
    res = plan.solve(0.5, pos)
    assert res.stats["converged"] == 1
    assert plan.field(pos).shape == (3, 0)
