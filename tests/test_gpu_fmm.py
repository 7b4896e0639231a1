"""GPU FMM (MODE_FMM plans) against the FP64 oracle: field, force, solve,
degenerate trees.  Tolerances per expansion order p (near-set |o|^2 <= 6,
leaves of side 2h): field/force relative L2 error <= 2e-6 at p = 10,
3e-5 at p = 8, 5e-4 at p = 6."""
import numpy as np
import pytest

from conftest import jittered_block, random_moments, rel_err

pytestmark = pytest.mark.gpu

TOL = {6: 5e-4, 8: 3e-5, 10: 2e-6}


def dev(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(np.asarray(a).T),
                        dtype=dtype or torch.float32, device="cuda").contiguous()


def host(t):
    return t.double().cpu().numpy().T


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def slab(nx, ny, nz, seed=0):
    return f32(jittered_block(nx, ny, nz, seed=seed))


@pytest.fixture(scope="module")
def slab_case(oracle):
    pos = slab(40, 36, 10, seed=5)
    m = f32(random_moments(pos.shape[0], seed=6, scale=0.4))
    field = oracle.field_allpairs(pos, pos, m, 2.0)
    force = oracle.force_allpairs(pos, m, 2.0)
    return pos, m, field, force


@pytest.mark.parametrize("order", [6, 8, 10])
def test_fmm_field_matches_oracle(ctx, slab_case, order):
    import magforce
    pos, m, field, _ = slab_case
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, order=order)
    got = host(plan.field(dev(m)))
    assert rel_err(got, field) < TOL[order]


@pytest.mark.parametrize("order", [6, 8, 10])
def test_fmm_force_matches_oracle(ctx, slab_case, order):
    import magforce
    pos, m, _, force = slab_case
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, order=order)
    got = host(plan.force(dev(m)))
    assert rel_err(got, force) < TOL[order]


# BF16-split M2L: six products match FP32 to ~1e-7 of the far field; three
# products leave ~2^-16 per translation, a few 1e-6 of the total field.
BF16_TOL = {"bf16x6": TOL, "bf16x3": {6: 5e-4, 8: 4e-5, 10: 2e-5}}


@pytest.mark.parametrize("order", [6, 8, 10])
@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3"])
def test_fmm_bf16_m2l_matches_oracle(ctx, slab_case, order, prec):
    import magforce
    pos, m, field, force = slab_case
    code = {"bf16x6": magforce.M2L_BF16X6, "bf16x3": magforce.M2L_BF16X3}[prec]
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, order=order,
                         m2l_precision=code)
    ref32 = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, order=order)
    mm = dev(m)
    g, g32 = host(plan.field(mm)), host(ref32.field(mm))
    F, F32 = host(plan.force(mm)), host(ref32.force(mm))
    assert rel_err(g, field) < BF16_TOL[prec][order]
    assert rel_err(F, force) < BF16_TOL[prec][order]
    if prec == "bf16x6":
        # same truncation as the FP32 plan, rounding differences only
        assert rel_err(g, g32) < 2e-7
        assert rel_err(F, F32) < 2e-7


@pytest.mark.parametrize("order", [6, 8, 10])
def test_fmm_bf16x1_cut_matches_oracle(ctx, slab_case, order):
    """Degree-cut one-part translations (the relaxed products' M2L): the cut
    drops the MFMA blocks whose every term has local plus multipole degree
    above min(order, 7), terms that fall off like rho^(j + n).  Field and force
    stay within a few 1e-4 of the oracle and near the uncut one-part product;
    a wrongly dropped low-degree block would put them off by 1e-2 or more,
    and a cut that dropped nothing would leave them equal to it."""
    import magforce
    pos, m, field, force = slab_case
    cut = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, order=order,
                        m2l_precision=magforce.M2L_BF16X1_CUT)
    one = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, order=order,
                        m2l_precision=magforce.M2L_BF16X1)
    mm = dev(m)
    g, g1 = host(cut.field(mm)), host(one.field(mm))
    F, F1 = host(cut.force(mm)), host(one.force(mm))
    eg, ef = rel_err(g, field), rel_err(F, force)
    dg, df = rel_err(g, g1), rel_err(F, F1)
    print(f"order {order}: cut field {eg:.3e} force {ef:.3e}; uncut field "
          f"{rel_err(g1, field):.3e} force {rel_err(F1, force):.3e}; cut vs uncut {dg:.3e} {df:.3e}")
    assert eg < 3e-4 and ef < 3e-4
    assert 0.0 < dg < 2e-4 and 0.0 < df < 2e-4


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 2, 2), (3, 7, 2), (9, 9, 9)])
def test_fmm_small_and_shallow_trees(ctx, oracle, shape):
    """Trees of one or two levels have no far field (exact near field only);
    the 9^3 cube already has a level-2 interaction list."""
    import magforce
    pos = slab(*shape, seed=11)
    m = f32(random_moments(pos.shape[0], seed=12))
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM)
    assert rel_err(host(plan.field(dev(m))), oracle.field_allpairs(pos, pos, m, 2.0)) < 5e-6
    F = host(plan.force(dev(m)))
    if pos.shape[0] > 1:
        assert rel_err(F, oracle.force_allpairs(pos, m, 2.0)) < 5e-6


def test_fmm_near_table_dense_cloud(ctx, oracle):
    """1000 uniform random particles in a cube of side 7.9 (h = 2): the tree
    has no far field, so every pair goes through the near-field kernels, and
    the pair distances cover q in [0, 2) densely (the full-leaf shape reads
    alpha, beta from the 64-interval cubic table there).  The table is within
    1.5e-7 of the exact coefficients, below FP32 product noise; a wrong
    interval or coefficient shows orders of magnitude above the bound.  The
    moments are all alike, so the field does not cancel."""
    import magforce
    rng = np.random.default_rng(51)
    pos = f32(rng.uniform(0.0, 7.9, size=(1000, 3)))
    m = f32(np.tile([0.2, -0.1, 0.6], (pos.shape[0], 1)) + 0.05 * rng.normal(size=(1000, 3)))
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM)
    assert plan.info()["depth"] < 2
    assert rel_err(host(plan.field(dev(m))), oracle.field_allpairs(pos, pos, m, 2.0)) < 1e-6


def test_fmm_coincident_particles(ctx, oracle):
    """Exact duplicates inside a slab with a far field (r = 0 pairs between
    distinct particles): the field pair takes its q = 0 value (the self term)
    and the force pair is zero (Dipole_B.kern:114-128,
    Dipole_hessian.kern:101-104), in the FMM near field as in the oracle."""
    import magforce
    base = slab(24, 24, 6, seed=61)
    rng = np.random.default_rng(62)
    dup = base[rng.choice(base.shape[0], size=300, replace=False)]
    pos = np.concatenate([base, dup])
    m = f32(random_moments(pos.shape[0], seed=63, scale=0.4))
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM)
    assert plan.info()["depth"] >= 2
    assert rel_err(host(plan.field(dev(m))), oracle.field_allpairs(pos, pos, m, 2.0)) < TOL[10]
    assert rel_err(host(plan.force(dev(m))), oracle.force_allpairs(pos, m, 2.0)) < TOL[10]


def test_fmm_sparse_deep_tree(ctx, oracle):
    """Widely spread clusters: many levels with few boxes each."""
    import magforce
    rng = np.random.default_rng(3)
    centres = rng.uniform(0, 600, size=(12, 3))
    pos = f32(np.concatenate([c + rng.normal(size=(150, 3)) * 2.0 for c in centres]))
    m = f32(random_moments(pos.shape[0], seed=4))
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM)
    assert rel_err(host(plan.field(dev(m))), oracle.field_allpairs(pos, pos, m, 2.0)) < 2e-6
    assert rel_err(host(plan.force(dev(m))), oracle.force_allpairs(pos, m, 2.0)) < 2e-6


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_fmm_lattice_under_sparse_layer(ctx, oracle, prec):
    """A lattice slab under a sparse random layer (the pool's shape, at a size
    the oracle covers).  The layer's leaves hold 1..16 particles and have
    lattice leaves in their near sets, so the sparse-leaf kernels stream many
    windows of sources, across near ranges of very different lengths."""
    import magforce
    lat = slab(24, 24, 8, seed=31)
    rng = np.random.default_rng(32)
    layer = rng.uniform(0.0, 1.0, size=(900, 3)) * np.array([24.0, 24.0, 8.0]) + np.array([0.0, 0.0, 8.0])
    pos = f32(np.concatenate([lat, layer]))
    m = f32(random_moments(pos.shape[0], seed=33))
    code = {"fp32": magforce.M2L_FP32, "bf16x3": magforce.M2L_BF16X3}[prec]
    # p = 10 truncation on this layout (random layer, worst separation ratio
    # 0.61) measures 2.1e-6 in the field; a fault in the sparse-leaf path
    # would show orders of magnitude above either bound
    tol = 5e-6 if prec == "fp32" else BF16_TOL["bf16x3"][10]
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, m2l_precision=code)
    assert rel_err(host(plan.field(dev(m))), oracle.field_allpairs(pos, pos, m, 2.0)) < tol
    assert rel_err(host(plan.force(dev(m))), oracle.force_allpairs(pos, m, 2.0)) < tol


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_near_overlap_matches_serial(ctx, slab_case, prec):
    """The near field on the second stream (the default) and serialized on the
    context stream (serial_near) run the same kernels on the same data, so
    their products are bitwise equal.  Products with different moments are
    issued back to back without host synchronization: a near-field kernel
    that ran past its join would read the next product's moments."""
    import magforce
    pos, m, _, _ = slab_case
    code = {"fp32": magforce.M2L_FP32, "bf16x3": magforce.M2L_BF16X3}[prec]
    a = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, m2l_precision=code)
    b = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, m2l_precision=code,
                      serial_near=True)
    ms = [dev(m), dev(f32(random_moments(pos.shape[0], seed=41))), dev(-m)]
    ga = [a.field(x) for x in ms] + [a.force(x) for x in ms]
    gb = [b.field(x) for x in ms] + [b.force(x) for x in ms]
    for u, v in zip(ga, gb):
        assert np.array_equal(host(u), host(v))


def test_fmm_hconst_and_update_positions(ctx, oracle):
    import magforce
    pos = slab(20, 20, 6, seed=21)
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM)
    pos2 = slab(17, 23, 8, seed=22) + 3.3
    m2 = f32(random_moments(pos2.shape[0], seed=23))
    plan.update_positions(dev(pos2))
    got = host(plan.field(dev(m2), hconst=[0.0, 0.5, 1.0]))
    want = oracle.field_allpairs(pos2, pos2, m2, 2.0, hconst=[0.0, 0.5, 1.0])
    assert rel_err(got, want) < 2e-6


def test_fmm_solve_and_force_vs_oracle(ctx, oracle):
    """The 8192-particle lattice solved with FMM products (CG, 1e-6), force
    by FMM, scored against the FP64 oracle solution (1e-12)."""
    import magforce
    from magforce import configs
    s = configs.lattice8k()
    pos = s.pos.astype(np.float64)
    b = np.tile(s.hconst, (s.n, 1))
    x_ref, _, _ = oracle.cg_solve(pos, s.h, s.kappa, b, rtol=1e-12)
    F_ref = oracle.force_allpairs(pos, s.kappa * x_ref, s.h)
    plan = magforce.Plan(ctx, dev(pos), s.h, mode=magforce.MODE_FMM)
    res = plan.solve(s.kappa, dev(b), rtol=1e-6)
    assert res.stats["converged"] == 1
    x = host(res.x)
    assert rel_err(x, x_ref) < 5e-6
    F = host(plan.force(dev(s.kappa * x)))
    assert rel_err(F, F_ref) < 1e-5


def test_fmm_time_step_warm_start(ctx, oracle):
    """The simulation caller's step (magnetic_force.cu:806-888): the particles
    of the 8192 lattice move by N(0, (0.02 h)^2) per axis, the plan is rebuilt
    in place, and CG starts from the previous solution (the unknown kept
    across Resize, :706-744).  The warm solve meets the same FP64 solution at
    the moved positions as a cold one, in fewer iterations."""
    import magforce
    from magforce import configs
    s = configs.lattice8k()
    pos0 = s.pos.astype(np.float64)
    b = np.tile(s.hconst, (s.n, 1))
    plan = magforce.Plan(ctx, dev(pos0), s.h, mode=magforce.MODE_FMM)
    x0 = plan.solve(s.kappa, dev(b), rtol=1e-6).x
    rng = np.random.default_rng(41)
    pos1 = f32(pos0 + rng.normal(scale=0.02 * s.h, size=pos0.shape))
    plan.update_positions(dev(pos1))
    warm = plan.solve(s.kappa, dev(b), x0=x0, rtol=1e-6)
    cold = plan.solve(s.kappa, dev(b), rtol=1e-6)
    assert warm.stats["converged"] == 1 and cold.stats["converged"] == 1
    # FP64 CG on this system: 5 iterations warm, 7 cold
    assert warm.stats["iterations"] < cold.stats["iterations"]
    x_ref, _, _ = oracle.cg_solve(pos1, s.h, s.kappa, b, rtol=1e-12)
    assert rel_err(host(warm.x), x_ref) < 5e-6
    F = host(plan.force(dev(s.kappa * host(warm.x))))
    F_ref = oracle.force_allpairs(pos1, s.kappa * x_ref, s.h)
    assert rel_err(F, F_ref) < 1e-5


def test_fmm_matches_fp64_gpu_at_100k(ctx):
    """A 100k-particle slab with a random layer above it, against the GPU FP64
    all-pairs field (itself pinned to the CPU oracle in test_gpu_allpairs)."""
    import magforce
    import torch
    rng = np.random.default_rng(9)
    lat = jittered_block(64, 64, 22, seed=9)
    top = np.column_stack([rng.uniform(0, 64, 4000), rng.uniform(0, 64, 4000),
                           rng.uniform(22, 40, 4000)])
    pos = f32(np.concatenate([lat, top]))
    m = f32(random_moments(pos.shape[0], seed=10, scale=0.5))
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM)
    got = host(plan.field(dev(m)))
    want = host(magforce.field_allpairs_f64(ctx, dev(pos, torch.float64), dev(pos, torch.float64),
                                            dev(m, torch.float64), 2.0))
    # random moments: the field is small where it cancels, so the relative
    # error runs above that of a solved (smooth) magnetisation
    assert rel_err(got, want) < 5e-6


@pytest.mark.parametrize("h,side", [(2.0, 3.0), (2.0, 4.0), (1.0, 2.5), (1.0, 6.0)])
def test_fmm_leaf_side(ctx, oracle, h, side):
    """Leaves of side leaf_side * h (> 2h): more pairs in the near field,
    fewer translations; the near sets still hold every pair within 2h (field)
    and 4h (force), and the force near field sorts its boxes into the ones
    that may hold pairs within 4h and the far-only rest for this side."""
    import magforce
    pos = slab(40, 36, 10, seed=71)
    m = f32(random_moments(pos.shape[0], seed=72, scale=0.4))
    base = magforce.Plan(ctx, dev(pos), h, mode=magforce.MODE_FMM)
    plan = magforce.Plan(ctx, dev(pos), h, mode=magforce.MODE_FMM, leaf_side=side)
    bi, pi = base.info(), plan.info()
    print(f"h={h} side={side}: base {bi} plan {pi}")
    assert pi["leaves"] < bi["leaves"]
    assert pi["p2p_field_pairs"] > bi["p2p_field_pairs"]
    ef = rel_err(host(plan.field(dev(m))), oracle.field_allpairs(pos, pos, m, h))
    eF = rel_err(host(plan.force(dev(m))), oracle.force_allpairs(pos, m, h))
    print(f"h={h} side={side}: field {ef:.3e} force {eF:.3e}")
    assert ef < 5e-6 and eF < 5e-6


def test_fmm_leaf_size_budget(ctx, oracle):
    """leaf_size (the reference's ncrit role): the leaf side grows in powers
    of two from 2h while the non-empty leaves hold at most leaf_size particles
    on average.  On a unit lattice at h = 1 (8 per leaf at side 2h), a budget
    of 64 gives side 4h (64 per leaf), 8 keeps 2h."""
    import magforce
    pos = slab(32, 32, 12, seed=73)
    m = f32(random_moments(pos.shape[0], seed=74, scale=0.4))
    p8 = magforce.Plan(ctx, dev(pos), 1.0, mode=magforce.MODE_FMM, leaf_size=8)
    p64 = magforce.Plan(ctx, dev(pos), 1.0, mode=magforce.MODE_FMM, leaf_size=64)
    p4h = magforce.Plan(ctx, dev(pos), 1.0, mode=magforce.MODE_FMM, leaf_side=4.0)
    assert p8.info()["leaves"] == magforce.Plan(ctx, dev(pos), 1.0, mode=magforce.MODE_FMM).info()["leaves"]
    assert p64.info()["leaves"] == p4h.info()["leaves"]
    assert pos.shape[0] / p64.info()["leaves"] <= 64
    ef = rel_err(host(p64.field(dev(m))), oracle.field_allpairs(pos, pos, m, 1.0))
    assert ef < 5e-6


def test_fmm_leaf_side_rejected_below_2h(ctx):
    import magforce
    pos = slab(4, 4, 4, seed=75)
    with pytest.raises(magforce.MagforceError) as e:
        magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, leaf_side=1.5)
    assert e.value.code == magforce.MF_ERR_INVALID_ARG


def test_autotuner_leaf_side_search(ctx, slab_case):
    """With leaf sides in the search, every (side, order, precision) is
    measured and the choice is the fastest within the target."""
    import magforce
    from magforce import tune
    pos, m, field, _ = slab_case
    best, cands = tune.autotune_fmm(ctx, dev(pos), 2.0, target_rel_err=1e-5, samples=512,
                                    orders=(8, 10), precisions=(magforce.M2L_BF16X3,),
                                    leaf_sides=(2.0, 3.0))
    assert len(cands) == 4 and sorted({c.leaf_side for c in cands}) == [2.0, 3.0]
    assert best is not None and best.rel_err <= 1e-5
    assert not [c for c in cands if c.rel_err <= 1e-5 and c.ms_per_product < best.ms_per_product]
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, order=best.order,
                         m2l_precision=best.m2l_precision, leaf_side=best.leaf_side)
    assert rel_err(host(plan.field(dev(m))), field) < 1e-5


def test_autotuner_picks_fastest_within_target(ctx, slab_case):
    """The autotuner's choice meets the error target on an independent check,
    and no faster candidate also met it."""
    import magforce
    from magforce import tune
    pos, m, field, _ = slab_case
    best, cands = tune.autotune_fmm(ctx, dev(pos), 2.0, target_rel_err=1e-5, samples=512)
    assert best is not None and best.rel_err <= 1e-5
    assert len(cands) == 9
    faster_ok = [c for c in cands if c.rel_err <= 1e-5 and c.ms_per_product < best.ms_per_product]
    assert not faster_ok
    plan = magforce.Plan(ctx, dev(pos), 2.0, mode=magforce.MODE_FMM, order=best.order,
                         m2l_precision=best.m2l_precision)
    assert rel_err(host(plan.field(dev(m))), field) < 1e-5
