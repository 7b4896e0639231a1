"""The workload configurations run end to end on the GPU with their own
generators (magforce.configs), scored against FP64:

  cylinder262k   all-pairs GMRES(30) + force: at 8,192 particles against the
                 CPU oracle; at the full 262,144 on 4,096 seeded samples
                 against the GPU FP64 path (itself pinned to the CPU oracle
                 in test_gpu_allpairs)
  pool8m shape   512x512x32's lattice at 64x64x32 (chi = 2, BF16x3 M2L):
                 force error measured directly against F* = the FP64 force at
                 the FP64 solution (all-pairs refinement to 1e-12)
  droplets96     96 spheres, log-uniform sizes (capped at 16,384 particles),
                 per-droplet chi and H, autotuned order, graph replay; every
                 droplet scored on its own against its FP64 solution
  other h        FMM field and force at h = 1 and h = 3 against the oracle

The bar is the metric's: force relative L2 error <= 1e-5 against FP64.
"""
import numpy as np
import pytest

from conftest import random_moments, rel_err

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(np.asarray(a).T),
                        dtype=dtype or torch.float32, device="cuda").contiguous()


def host(t):
    return t.double().cpu().numpy().T


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def tnorm_rel(a, b):
    import torch
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def test_cylinder_small_gmres30_vs_oracle(ctx, oracle):
    """The cylinder generator at 8,192 sites (the first 8 full layers of the
    radius-32 disc and part of the 9th): GMRES(30) for 30 all-pairs products,
    then the force, against the FP64 oracle's CG solution (1e-12) and force."""
    import magforce
    from magforce import configs
    s = configs.cylinder262k(count=8192)
    pos = s.pos.astype(np.float64)
    b = np.tile(s.hconst, (s.n, 1))
    print(f"cylinder n={s.n} GMRES(30)", flush=True)
    plan = magforce.Plan(ctx, dev(pos), s.h)
    res = plan.solve(s.kappa, dev(b), method=magforce.SOLVE_GMRES, restart=30, maxit=30,
                     rtol=1e-30)
    assert res.stats["matvecs"] >= 30
    x = host(res.x)
    F = host(plan.force(dev(s.kappa * x)))
    x_ref, _, _ = oracle.cg_solve(pos, s.h, s.kappa, b, rtol=1e-12)
    F_ref = oracle.force_allpairs(pos, s.kappa * x_ref, s.h)
    ex, ef = rel_err(x, x_ref), rel_err(F, F_ref)
    print(f"cylinder8k: x_err={ex:.3e} F_err={ef:.3e} res={res.stats['rel_residual']:.3e}")
    assert ex < 1e-6
    assert ef < 1e-5


def test_cylinder262k_gmres30_full(ctx):
    """The full cylinder workload as the bench runs it: GMRES(30) with 30
    all-pairs products and the force, scored on 4,096 seeded particles
    against the FP64 solution and the FP64 force over all sources."""
    import torch
    import magforce
    from magforce import configs
    s = configs.cylinder262k()
    assert s.n == 262144
    pos = dev(s.pos)
    b = magforce.uniform_field(s.n, s.hconst)
    print(f"cylinder n={s.n} GMRES(30) full", flush=True)
    plan = magforce.Plan(ctx, pos, s.h)
    sol = plan.solve(s.kappa, b, method=magforce.SOLVE_GMRES, restart=30, maxit=30, rtol=1e-30)
    F = plan.force(s.kappa * sol.x)
    ref = plan.solve_f64(s.kappa, b.double(), rtol=1e-12, max_refine=6)
    assert ref.stats["converged"] == 1
    rng = np.random.default_rng(12345)
    idx = torch.tensor(np.sort(rng.choice(s.n, size=4096, replace=False)), dtype=torch.int64,
                       device="cuda")
    F_ref = magforce.force_subset_f64(ctx, pos.double(), s.kappa * ref.x, s.h, idx)
    ef = tnorm_rel(F[:, idx].double(), F_ref)
    ex = tnorm_rel(sol.x.double(), ref.x)
    print(f"cylinder262k: x_err={ex:.3e} F_err={ef:.3e}")
    assert ex < 1e-6
    assert ef < 1e-5
    plan.close()


def test_pool8m_shape_bf16x3_force_vs_fp64_solution(ctx):
    """The 8.4M lattice's shape and parameters at 131,072 particles
    (64x64x32 jittered lattice, h = 2, chi = 2, H = (0,0,1)), solved with the
    mixed-precision FMM (BF16x3 translations, CG to 1e-6) and its force.
    F* is the FP64 force at the FP64 solution, so the error of the solve and
    of the force product are measured together, over every particle."""
    import torch
    import magforce
    from magforce import configs
    pos_h = configs.jittered_lattice(64, 64, 32, seed=0)
    h, chi = 2.0, 2.0
    kappa = chi / (1.0 + chi)
    n = pos_h.shape[0]
    pos = dev(pos_h)
    b = magforce.uniform_field(n, [0.0, 0.0, 1.0])
    print(f"pool8m-shape n={n} BF16x3 FMM", flush=True)
    plan = magforce.Plan(ctx, pos, h, mode=magforce.MODE_FMM, m2l_precision=magforce.M2L_BF16X3)
    res = plan.solve(kappa, b, rtol=1e-6)
    assert res.stats["converged"] == 1
    F = plan.force(kappa * res.x)
    ap = magforce.Plan(ctx, pos, h)
    ref = ap.solve_f64(kappa, b.double(), rtol=1e-12, max_refine=8)
    assert ref.stats["converged"] == 1
    F_ref = magforce.force_allpairs_f64(ctx, pos.double(), kappa * ref.x, h)
    ef = tnorm_rel(F.double(), F_ref)
    ex = tnorm_rel(res.x.double(), ref.x)
    print(f"pool8m-shape: x_err={ex:.3e} F_err={ef:.3e} iters={res.stats['iterations']}")
    assert ex < 5e-6
    assert ef < 1e-5
    # relaxed products: those made below a residual of 1e-2 (one more than the
    # bench's 5e-3 relaxes on the pool) on the degree-cut one-part translations
    rel = plan.solve(kappa, b, rtol=1e-6, relax_m2l_precision=magforce.M2L_BF16X1_CUT,
                     relax_rtol=1e-2)
    assert rel.stats["converged"] == 1 and rel.stats["relaxed_matvecs"] >= 1
    Fr = plan.force(kappa * rel.x)
    efr = tnorm_rel(Fr.double(), F_ref)
    exr = tnorm_rel(rel.x.double(), ref.x)
    print(f"pool8m-shape relaxed: x_err={exr:.3e} F_err={efr:.3e} "
          f"relaxed {rel.stats['relaxed_matvecs']} of {rel.stats['matvecs']}")
    assert exr < 5e-6
    assert efr < 1e-5
    plan.close()
    ap.close()


def test_droplets96_batch_autotuned_graph(ctx):
    """96 droplets from the droplets96 generator (seed 7, sizes log-uniform,
    capped at 16,384 particles here), chi_i and unit H_i per droplet, one
    batch plan with the autotuner's order (tuned on the largest droplet),
    batched CG to 1e-6 per system with graph replay, block-diagonal force.
    Every droplet is scored on its own: x and F on up to 128 seeded
    particles against that droplet's FP64 solution (all-pairs refinement)."""
    import torch
    import magforce
    from magforce import configs, tune
    systems = configs.droplets96(nmax=16384)
    assert len(systems) == 96
    sizes = [sy.n for sy in systems]
    off = [0]
    for sy in systems:
        off.append(off[-1] + sy.n)
    print(f"droplets96 n={off[-1]} sizes {min(sizes)}..{max(sizes)}", flush=True)
    pos = dev(np.concatenate([sy.pos for sy in systems]))
    b = dev(np.concatenate([np.tile(np.asarray(sy.hconst, np.float64), (sy.n, 1))
                            for sy in systems]))
    kpart = torch.tensor(np.concatenate([np.full(sy.n, sy.kappa, np.float32) for sy in systems]),
                         device="cuda")
    big = max(systems, key=lambda sy: sy.n)
    best, cands = tune.autotune_fmm(ctx, dev(big.pos), big.h, target_rel_err=1e-5,
                                    precisions=(magforce.M2L_BF16X3,), leaf_sides=(2.0, 3.0))
    assert best is not None
    print(f"droplets96 tuned order {best.order} side {best.leaf_side}: "
          f"{[c.as_dict() for c in cands]}")
    plan = magforce.BatchPlan(ctx, pos, off, systems[0].h, order=best.order,
                              m2l_precision=magforce.M2L_BF16X3, leaf_side=best.leaf_side)
    res = plan.solve([sy.kappa for sy in systems], b, rtol=1e-6, graph_iters=2)
    assert res.stats["converged"] == 1
    assert max(res.stats["sys_rel_residual"]) <= 1e-6
    F = plan.force(res.x * kpart)
    rng = np.random.default_rng(4321)
    worst_x = worst_F = 0.0
    for sy, a, e in zip(systems, off[:-1], off[1:]):
        p = pos[:, a:e].contiguous()
        ap = magforce.Plan(ctx, p, sy.h)
        bs = b[:, a:e].contiguous().double()
        ref = ap.solve_f64(sy.kappa, bs, rtol=1e-12, max_refine=8)
        assert ref.stats["converged"] == 1
        k = min(128, sy.n)
        loc = torch.tensor(np.sort(rng.choice(sy.n, size=k, replace=False)), device="cuda")
        F_ref = magforce.force_subset_f64(ctx, p.double(), sy.kappa * ref.x, sy.h, loc)
        worst_F = max(worst_F, tnorm_rel(F[:, a:e][:, loc].double(), F_ref))
        worst_x = max(worst_x, tnorm_rel(res.x[:, a:e].double(), ref.x))
        ap.close()
    print(f"droplets96: worst droplet x_err={worst_x:.3e} F_err={worst_F:.3e}")
    assert worst_x < 5e-6
    assert worst_F < 1e-5
    plan.close()


@pytest.mark.parametrize("h", [1.0, 3.0])
def test_fmm_at_other_smoothing_lengths(ctx, oracle, h):
    """The FMM's leaf side follows h (2h), so h changes the tree, the leaf
    occupancy (8 particles per leaf at h = 1, 216 at h = 3 on this lattice)
    and the near-field split; field and force at p = 10 against the oracle."""
    import magforce
    from magforce import configs
    pos = f32(configs.jittered_lattice(40, 36, 12, seed=17))
    m = f32(random_moments(pos.shape[0], seed=18, scale=0.4))
    print(f"fmm h={h} n={len(pos)}", flush=True)
    plan = magforce.Plan(ctx, dev(pos), h, mode=magforce.MODE_FMM)
    assert plan.info()["depth"] >= 2
    ef = rel_err(host(plan.field(dev(m))), oracle.field_allpairs(pos, pos, m, h))
    eF = rel_err(host(plan.force(dev(m))), oracle.force_allpairs(pos, m, h))
    print(f"fmm h={h}: field {ef:.3e} force {eF:.3e}")
    assert ef < 5e-6
    assert eF < 5e-6
