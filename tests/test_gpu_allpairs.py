"""GPU parity of the all-pairs kernels against the FP64 oracle and the golden
fixtures.  Bar: FP32 path within 1e-5 relative (norm over all particles);
FP64 path within 1e-12.  Inputs are rounded to float32 once and the oracle
sees exactly those values.
"""
import os

import numpy as np
import pytest

from conftest import jittered_block, max_rel_err, random_moments, rel_err

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


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 1, 1), (2, 2, 2), (4, 4, 4), (5, 7, 3),
                                   (9, 9, 9), (16, 16, 8)])
def test_plan_field_matches_oracle(ctx, oracle, shape):
    import magforce
    pos = f32(jittered_block(*shape, seed=sum(shape)))
    m = f32(random_moments(pos.shape[0], seed=len(pos)))
    want = oracle.field_allpairs(pos, pos, m, 2.0)
    plan = magforce.Plan(ctx, dev(pos), 2.0)
    got = host(plan.field(dev(m)))
    assert rel_err(got, want) < 1e-5
    assert max_rel_err(got, want) < 1e-5


def test_field_separate_sources_and_constant(ctx, oracle):
    import magforce
    tpos = f32(jittered_block(7, 5, 3, seed=3))
    spos = f32(jittered_block(3, 4, 11, seed=4) + [1.3, -2.0, 0.5])
    m = f32(random_moments(spos.shape[0], seed=5))
    hc = [0.1, -0.2, 1.0]
    for h in (0.7, 2.0, 3.5):
        want = oracle.field_allpairs(tpos, spos, m, h, hconst=hc)
        got = host(magforce.field_allpairs(ctx, dev(tpos), dev(spos), dev(m), h, hconst=hc))
        assert rel_err(got, want) < 1e-5


def test_coincident_particles(ctx, oracle):
    """Two particles on the same spot (the example's 'singularity' case):
    self-term for the field, no force between them."""
    import magforce
    pos = f32(np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [1.5, 0.2, -0.3], [9.0, 9.0, 9.0]]))
    m = f32(np.array([[0.0, 1.0, 0.0], [0.5, 0.1, 0.2], [0.3, -0.3, 0.9], [1.0, 0.0, 0.0]]))
    plan = magforce.Plan(ctx, dev(pos), 1.0)
    got = host(plan.field(dev(m)))
    assert np.all(np.isfinite(got))
    assert rel_err(got, oracle.field_allpairs(pos, pos, m, 1.0)) < 1e-6
    F = host(plan.force(dev(m)))
    assert np.all(np.isfinite(F))
    assert rel_err(F, oracle.force_allpairs(pos, m, 1.0)) < 1e-5


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 3, 3), (5, 7, 3), (9, 9, 9), (12, 12, 12)])
def test_plan_force_matches_oracle(ctx, oracle, shape):
    import magforce
    pos = f32(jittered_block(*shape, seed=7 + sum(shape)))
    m = f32(random_moments(pos.shape[0], seed=8))
    want = oracle.force_allpairs(pos, m, 2.0)
    plan = magforce.Plan(ctx, dev(pos), 2.0)
    got = host(plan.force(dev(m)))
    if pos.shape[0] == 1:
        assert np.all(got == 0)
        return
    assert rel_err(got, want) < 1e-5
    # the same through the stateless entry point
    got2 = host(magforce.force_allpairs(ctx, dev(pos), dev(m), 2.0))
    assert np.array_equal(got, got2)


def test_force_with_external_tensor(ctx, oracle):
    import magforce
    import torch
    pos = f32(jittered_block(4, 5, 6, seed=21))
    n = pos.shape[0]
    m = f32(random_moments(n, seed=22))
    rng = np.random.default_rng(23)
    T = rng.normal(size=(n, 3, 3)) * 1e-7
    T = f32(0.5 * (T + T.transpose(0, 2, 1)))
    want = oracle.force_allpairs(pos, m, 2.0, text=T)
    # dHextd{x,y,z} columns: text9 rows 3a..3a+2 = components of column a
    text = torch.tensor(np.stack([T[:, r, a] for a in range(3) for r in range(3)]),
                        dtype=torch.float32, device="cuda")
    got = host(magforce.Plan(ctx, dev(pos), 2.0).force(dev(m), text=text))
    assert rel_err(got, want) < 1e-5


def test_hessian_matches_oracle(ctx, oracle):
    import magforce
    tpos = f32(jittered_block(5, 5, 5, seed=31))
    spos = f32(jittered_block(4, 6, 3, seed=32) + [0.4, 0.3, 1.1])
    m = f32(random_moments(spos.shape[0], seed=33))
    want = oracle.hessian_allpairs(tpos, spos, m, 2.0)
    got6 = magforce.hessian_allpairs(ctx, dev(tpos), dev(spos), dev(m), 2.0).double().cpu().numpy()
    xx, yy, zz, xy, xz, yz = got6
    got = np.stack([np.stack([xx, xy, xz], 1), np.stack([xy, yy, yz], 1),
                    np.stack([xz, yz, zz], 1)], 1)
    assert rel_err(got.reshape(-1), want.reshape(-1)) < 1e-5


@pytest.mark.parametrize("layout", ["near", "far_and_near", "large"])
def test_hessian9_matches_oracle(ctx, oracle, layout):
    """The tensor written straight into the nine dHext arrays (T_ab, a-major),
    against the oracle: sources overlapping the targets (every tile mixed),
    a second cluster beyond 4h (tiles culled to the dipole form), and enough
    targets and sources to split the source tiles into chunks."""
    import magforce
    tpos = f32(jittered_block(6, 5, 4, seed=51))
    spos = f32(jittered_block(4, 4, 3, seed=52) + [0.3, 0.2, 0.9])
    if layout == "far_and_near":
        spos = np.concatenate([spos, f32(jittered_block(5, 5, 5, seed=53) + [30.0, -12.0, 4.0])])
    if layout == "large":
        tpos = f32(jittered_block(24, 20, 12, seed=54))
        spos = f32(jittered_block(20, 20, 16, seed=55) + [5.5, -3.0, 2.0])
    m = f32(random_moments(spos.shape[0], seed=56))
    want = oracle.hessian_allpairs(tpos, spos, m, 2.0)
    got = magforce.hessian_allpairs9(ctx, dev(tpos), dev(spos), dev(m), 2.0).double().cpu().numpy()
    got = got.T.reshape(-1, 3, 3)
    assert rel_err(got.reshape(-1), want.reshape(-1)) < 1e-5
    # symmetric, and the same values as the six-component entry point
    assert np.array_equal(got[:, 0, 1], got[:, 1, 0]) and np.array_equal(got[:, 1, 2], got[:, 2, 1])
    six = magforce.hessian_allpairs(ctx, dev(tpos), dev(spos), dev(m), 2.0).double().cpu().numpy()
    assert np.array_equal(six[0], got[:, 0, 0]) and np.array_equal(six[5], got[:, 1, 2])


def test_fp64_path_matches_oracle(ctx, oracle):
    import magforce
    import torch
    pos = jittered_block(6, 5, 7, seed=41)
    m = random_moments(pos.shape[0], seed=42)
    want = oracle.field_allpairs(pos, pos, m, 2.0, hconst=[0, 0, 1])
    got = magforce.field_allpairs_f64(ctx, dev(pos, torch.float64), dev(pos, torch.float64),
                                      dev(m, torch.float64), 2.0, hconst=[0, 0, 1])
    assert rel_err(host(got), want) < 1e-13
    F = magforce.force_allpairs_f64(ctx, dev(pos, torch.float64), dev(m, torch.float64), 2.0)
    # the oracle keeps the reference's rotation frame with its 1e-10 pole
    # patch (Dipole_hessian.kern:116); the GPU uses the exact isotropic form
    assert rel_err(host(F), oracle.force_allpairs(pos, m, 2.0)) < 1e-9


def test_golden_products(ctx):
    import magforce
    g = np.load(os.path.join(GOLD, "products.npz"))
    plan = magforce.Plan(ctx, dev(g["pos"]), float(g["h"]))
    assert rel_err(host(plan.field(dev(g["mom"]))), g["field"]) < 1e-5
    assert rel_err(host(plan.force(dev(g["mom"]))), g["force"]) < 1e-5


def test_lattice8k_field_and_force(ctx, oracle):
    """The 8192-particle jittered lattice (the CPU-runnable workload)."""
    import magforce
    from magforce import configs
    sysm = configs.lattice8k()
    pos = sysm.pos.astype(np.float64)
    m = f32(random_moments(sysm.n, seed=51, scale=0.3))
    plan = magforce.Plan(ctx, dev(pos), sysm.h)
    assert rel_err(host(plan.field(dev(m))), oracle.field_allpairs(pos, pos, m, sysm.h)) < 1e-5
    assert rel_err(host(plan.force(dev(m))), oracle.force_allpairs(pos, m, sysm.h)) < 1e-5


def test_deterministic(ctx):
    import magforce
    pos = f32(jittered_block(10, 10, 10, seed=61))
    m = f32(random_moments(pos.shape[0], seed=62))
    plan = magforce.Plan(ctx, dev(pos), 2.0)
    a = host(plan.field(dev(m)))
    b = host(plan.field(dev(m)))
    assert np.array_equal(a, b)


def test_update_positions_resizes(ctx, oracle):
    import magforce
    pos = f32(jittered_block(4, 4, 4, seed=71))
    plan = magforce.Plan(ctx, dev(pos), 2.0)
    pos2 = f32(jittered_block(5, 3, 7, seed=72))
    m2 = f32(random_moments(pos2.shape[0], seed=73))
    plan.update_positions(dev(pos2))
    assert rel_err(host(plan.field(dev(m2))), oracle.field_allpairs(pos2, pos2, m2, 2.0)) < 1e-5


def test_invalid_arguments_raise(ctx):
    import magforce
    pos = f32(jittered_block(2, 2, 2))
    with pytest.raises(magforce.MagforceError):
        magforce.Plan(ctx, dev(pos), -1.0)
    with pytest.raises(magforce.MagforceError):
        magforce.field_allpairs(ctx, dev(pos), dev(pos), dev(pos), 0.0)


def test_results_ordered_with_torch_stream(ctx):
    """Library kernels run on torch's current stream: a read-back issued
    right after the call (no explicit synchronisation) sees the results."""
    import magforce
    import torch
    pos = f32(jittered_block(40, 40, 20, seed=81))
    m = f32(random_moments(pos.shape[0], seed=82))
    plan = magforce.Plan(ctx, dev(pos), 2.0)
    out = torch.full((3, pos.shape[0]), float("nan"), dtype=torch.float32, device="cuda")
    plan.field(dev(m), out=out)
    first = out.cpu()
    assert not torch.isnan(first).any()
    torch.cuda.synchronize()
    assert torch.equal(first, out.cpu())
