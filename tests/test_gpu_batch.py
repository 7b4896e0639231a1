"""Batch plans (mf_plan_create_batch / mf_plan_solve_batch): independent
particle systems in one FMM forest.  Every system's block of a product, and
every system's solution, is scored against the FP64 oracle run on that system
alone.  The systems overlap in space on purpose: a product that leaked between
systems would fail."""
import numpy as np
import pytest

from conftest import jittered_block, random_moments, rel_err

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(np.asarray(a).T), dtype=torch.float32,
                        device="cuda").contiguous()


def host(t):
    return t.double().cpu().numpy().T


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def make_systems():
    """Uneven systems, all centred near the origin (overlapping), one with a
    single particle and one empty."""
    rng = np.random.default_rng(21)
    blocks = [jittered_block(14, 12, 10, seed=1), jittered_block(6, 6, 6, seed=2),
              jittered_block(30, 30, 4, seed=3), np.zeros((1, 3)), np.zeros((0, 3)),
              rng.uniform(-20, 20, size=(700, 3))]
    systems = []
    for b in blocks:
        b = np.asarray(b, dtype=np.float64)
        if len(b):
            b = b - b.mean(axis=0) + rng.uniform(-2, 2, size=3)
        systems.append(f32(b))
    return systems


def offsets_of(systems):
    off = [0]
    for s in systems:
        off.append(off[-1] + len(s))
    return off


@pytest.fixture(scope="module")
def batch_case():
    systems = make_systems()
    off = offsets_of(systems)
    pos = np.concatenate(systems)
    m = f32(random_moments(len(pos), seed=5, scale=0.4))
    return systems, off, pos, m


def test_batch_field_and_force_are_block_diagonal(ctx, oracle, batch_case):
    import magforce
    systems, off, pos, m = batch_case
    plan = magforce.BatchPlan(ctx, dev(pos), off, 2.0)
    G = host(plan.field(dev(m)))
    F = host(plan.force(dev(m)))
    for s, p in enumerate(systems):
        a, b = off[s], off[s + 1]
        if b - a == 0:
            continue
        g_ref = oracle.field_allpairs(p, p, m[a:b], 2.0)
        assert rel_err(G[a:b], g_ref) < 5e-6, s
        if b - a > 1:
            f_ref = oracle.force_allpairs(p, m[a:b], 2.0)
            assert rel_err(F[a:b], f_ref) < 5e-6, s
        else:
            assert np.abs(F[a:b]).max() == 0.0


@pytest.mark.parametrize("m2l", ["fp32", "bf16x3"])
def test_batch_solve_per_system(ctx, oracle, batch_case, m2l):
    """One kappa and one H_ext direction per system, solved together."""
    import magforce
    systems, off, pos, _ = batch_case
    rng = np.random.default_rng(8)
    chis = rng.uniform(0.5, 3.0, size=len(systems))
    kappas = chis / (1.0 + chis)
    hdir = rng.normal(size=(len(systems), 3))
    hdir /= np.linalg.norm(hdir, axis=1, keepdims=True)
    b = np.concatenate([np.tile(hdir[s], (len(p), 1)) for s, p in enumerate(systems)])
    code = {"fp32": magforce.M2L_FP32, "bf16x3": magforce.M2L_BF16X3}[m2l]
    plan = magforce.BatchPlan(ctx, dev(pos), off, 2.0, m2l_precision=code)
    res = plan.solve(kappas, dev(b), rtol=1e-6)
    assert res.stats["converged"] == 1
    assert max(res.stats["sys_rel_residual"]) <= 1e-6
    x = host(res.x)
    for s, p in enumerate(systems):
        a, e = off[s], off[s + 1]
        if e - a == 0:
            continue
        x_ref, _, _ = oracle.cg_solve(p, 2.0, kappas[s], b[a:e], rtol=1e-12)
        assert rel_err(x[a:e], x_ref) < 5e-6, s


@pytest.mark.parametrize("graph_iters,relax", [(0, "bf16x1"), (2, "bf16x1"), (2, "bf16x1cut")])
def test_batch_relaxed_products(ctx, oracle, batch_case, graph_iters, relax):
    """Batched CG with relaxed products (one-part BF16 translations, or
    their degree cut, once the largest relative residual is at 3e-3): every
    system still meets 1e-6 and its oracle solution; with graph replay the
    switch re-captures the iteration."""
    import magforce
    systems, off, pos, _ = batch_case
    rng = np.random.default_rng(18)
    chis = rng.uniform(0.5, 3.0, size=len(systems))
    kappas = chis / (1.0 + chis)
    hdir = rng.normal(size=(len(systems), 3))
    hdir /= np.linalg.norm(hdir, axis=1, keepdims=True)
    b = np.concatenate([np.tile(hdir[s], (len(p), 1)) for s, p in enumerate(systems)])
    plan = magforce.BatchPlan(ctx, dev(pos), off, 2.0, m2l_precision=magforce.M2L_BF16X3)
    res = plan.solve(kappas, dev(b), rtol=1e-6, graph_iters=graph_iters,
                     relax_m2l_precision={"bf16x1": magforce.M2L_BF16X1,
                                          "bf16x1cut": magforce.M2L_BF16X1_CUT}[relax],
                     relax_rtol=3e-3)
    print(f"batch relaxed ({relax}, graph_iters={graph_iters}): {res.stats['matvecs']} products, "
          f"{res.stats['relaxed_matvecs']} relaxed")
    assert res.stats["converged"] == 1
    assert 1 <= res.stats["relaxed_matvecs"] < res.stats["matvecs"]
    assert max(res.stats["sys_rel_residual"]) <= 1e-6
    x = host(res.x)
    for s, p in enumerate(systems):
        a, e = off[s], off[s + 1]
        if e - a == 0:
            continue
        x_ref, _, _ = oracle.cg_solve(p, 2.0, kappas[s], b[a:e], rtol=1e-12)
        assert rel_err(x[a:e], x_ref) < 5e-6, s


def test_batch_matches_single_plan_for_one_system(ctx, oracle):
    """A batch of one system is the ordinary plan."""
    import magforce
    p = f32(jittered_block(16, 16, 8, seed=9))
    m = f32(random_moments(len(p), seed=10))
    one = magforce.BatchPlan(ctx, dev(p), [0, len(p)], 2.0)
    ref = magforce.Plan(ctx, dev(p), 2.0, mode=magforce.MODE_FMM)
    assert rel_err(host(one.field(dev(m))), host(ref.field(dev(m)))) < 1e-7
    assert rel_err(host(one.force(dev(m))), host(ref.force(dev(m)))) < 1e-7


def test_batch_update_positions_and_errors(ctx, oracle, batch_case):
    import magforce
    systems, off, pos, m = batch_case
    plan = magforce.BatchPlan(ctx, dev(pos), off, 2.0)
    moved = pos + 0.05
    plan.update_positions(dev(moved))
    G = host(plan.field(dev(m)))
    a, b = off[0], off[1]
    ref = oracle.field_allpairs(moved[a:b], moved[a:b], m[a:b], 2.0)
    assert rel_err(G[a:b], ref) < 5e-6
    with pytest.raises(magforce.MagforceError):
        plan.update_positions(dev(moved[:-1]))
    with pytest.raises(ValueError):
        magforce.BatchPlan(ctx, dev(pos), [0, 5], 2.0)
    with pytest.raises(ValueError):
        plan.solve([0.5], dev(np.ones_like(pos)))


def test_batch_solve_graph_replay(ctx, oracle, batch_case):
    """Graph mode (one captured CG iteration replayed between convergence
    checks) reaches the same per-system solutions; replays after every
    system has converged leave them frozen."""
    import magforce
    systems, off, pos, _ = batch_case
    kappas = [0.5] * len(systems)
    b = np.tile(np.array([0.0, 0.0, 1.0]), (len(pos), 1))
    plan = magforce.BatchPlan(ctx, dev(pos), off, 2.0)
    eager = plan.solve(kappas, dev(b), rtol=1e-6)
    for gi in (1, 3):
        res = plan.solve(kappas, dev(b), rtol=1e-6, graph_iters=gi)
        assert res.stats["converged"] == 1
        assert eager.stats["iterations"] <= res.stats["iterations"] <= eager.stats["iterations"] + gi
        x = host(res.x)
        for s, p in enumerate(systems):
            a, e = off[s], off[s + 1]
            if e - a == 0:
                continue
            x_ref, _, _ = oracle.cg_solve(p, 2.0, kappas[s], b[a:e], rtol=1e-12)
            assert rel_err(x[a:e], x_ref) < 5e-6, (gi, s)
