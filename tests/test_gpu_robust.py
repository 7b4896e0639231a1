"""Failure handling on the GPU: every device allocation refused in turn by
the fault-injection hook (mf_set_alloc_hook), and non-finite inputs.  Each
refused allocation must surface as MF_ERR_ALLOC from the call that made it,
leave no half-built plan behind, and leave the context usable: the same
sequence run again once the hook is gone gives the clean run's results
bit for bit."""
import numpy as np
import pytest

from conftest import jittered_block

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(np.asarray(a).T), dtype=torch.float32,
                        device="cuda").contiguous()


POS = jittered_block(12, 12, 8, seed=41)
KAPPA = 0.5


def run_sequence(c, log=None):
    """FMM plan build, field, solve, force, position update, field: returns
    the results; `log` collects the name of each step as it starts."""
    import magforce
    import torch
    step = (lambda name: log.append(name)) if log is not None else (lambda name: None)
    step("plan")
    p = magforce.Plan(c, dev(POS), 2.0, mode=magforce.MODE_FMM)
    m = dev(np.full_like(POS, 0.3))
    try:
        step("field")
        g = p.field(m)
        step("solve")
        b = magforce.uniform_field(p.n, [0.0, 0.0, 1.0])
        x = p.solve(KAPPA, b, rtol=1e-6).x
        step("force")
        F = p.force(KAPPA * x)
        step("update")
        p.update_positions(dev(POS + 0.05))
        step("field2")
        g2 = p.field(m)
        torch.cuda.synchronize()
        return [t.cpu().numpy() for t in (g, x, F, g2)]
    finally:
        p.close()


def test_every_allocation_failure_is_clean():
    import magforce
    clean_ctx = magforce.Context(0)
    ref = run_sequence(clean_ctx)
    clean_ctx.close()
    # the allocations of the sequence on a fresh context
    sizes = []
    magforce.set_alloc_hook(lambda nbytes: sizes.append(nbytes) or False)
    try:
        c = magforce.Context(0)
        run_sequence(c)
    finally:
        magforce.set_alloc_hook(None)
    c.close()
    n = len(sizes)
    print(f"{n} device allocations in the sequence")
    assert n >= 10
    failed_steps = set()
    for k in range(n):
        calls = [0]

        def refuse_kth(nbytes, k=k):
            calls[0] += 1
            return calls[0] == k + 1

        c = magforce.Context(0)
        log = []
        magforce.set_alloc_hook(refuse_kth)
        try:
            with pytest.raises(magforce.MagforceError) as err:
                run_sequence(c, log)
        finally:
            magforce.set_alloc_hook(None)
        assert err.value.code == magforce.MF_ERR_ALLOC, (k, str(err.value))
        assert "refused" in str(err.value)
        failed_steps.add(log[-1])
        # the same context, the hook gone: the clean run's results exactly
        got = run_sequence(c)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (k, log[-1])
        c.close()
    print(f"refusals surfaced in steps {sorted(failed_steps)}")
    assert "plan" in failed_steps


def test_failed_update_leaves_no_half_built_plan(ctx):
    """A position update whose allocation fails leaves the plan refusing
    products (MF_ERR_INVALID_ARG) until an update succeeds."""
    import magforce
    p = magforce.Plan(ctx, dev(POS), 2.0, mode=magforce.MODE_FMM)
    m = dev(np.full_like(POS, 0.3))
    before = p.field(m).cpu().numpy()
    big = np.concatenate([POS, POS + [0.0, 0.0, 40.0], POS + [0.0, 40.0, 0.0]])
    magforce.set_alloc_hook(lambda nbytes: True)
    try:
        with pytest.raises(magforce.MagforceError) as err:
            p.update_positions(dev(big))
    finally:
        magforce.set_alloc_hook(None)
    assert err.value.code == magforce.MF_ERR_ALLOC
    # the plan's particle count is the new one, its buffers are not: the
    # product must refuse rather than read the old-size moments past their end
    with pytest.raises(magforce.MagforceError) as err2:
        p.field(m)
    assert err2.value.code == magforce.MF_ERR_INVALID_ARG
    assert "update failed" in str(err2.value)
    p.update_positions(dev(POS))
    assert np.array_equal(p.field(m).cpu().numpy(), before)
    p.close()


def test_non_finite_inputs(ctx):
    """NaN or Inf: the FMM tree build rejects positions that are not finite,
    h, kappa and the solve tolerances must be finite; an all-pairs product
    (no tree) carries a non-finite input into its outputs, as the
    reference's kernels do, without failing."""
    import magforce
    import torch
    bad = POS.copy()
    bad[5, 1] = np.nan
    with pytest.raises(magforce.MagforceError) as err:
        magforce.Plan(ctx, dev(bad), 2.0, mode=magforce.MODE_FMM)
    assert err.value.code == magforce.MF_ERR_INVALID_ARG and "finite" in str(err.value)
    bad[5, 1] = np.inf
    with pytest.raises(magforce.MagforceError):
        magforce.Plan(ctx, dev(bad), 2.0, mode=magforce.MODE_FMM)
    for h in (float("nan"), float("inf"), 0.0):
        with pytest.raises(magforce.MagforceError) as err:
            magforce.Plan(ctx, dev(POS), h, mode=magforce.MODE_FMM)
        assert err.value.code == magforce.MF_ERR_INVALID_ARG
    p = magforce.Plan(ctx, dev(POS), 2.0, mode=magforce.MODE_FMM)
    b = magforce.uniform_field(p.n, [0.0, 0.0, 1.0])
    for kw in ({"kappa": float("nan")}, {"kappa": 0.5, "rtol": float("nan")},
               {"kappa": 0.5, "rtol": 1e-6, "relax_m2l_precision": magforce.M2L_BF16X1,
                "relax_rtol": float("inf")}):
        kappa = kw.pop("kappa")
        with pytest.raises(magforce.MagforceError) as err:
            p.solve(kappa, b, **kw)
        assert err.value.code == magforce.MF_ERR_INVALID_ARG
    assert p.solve(0.5, b).stats["converged"] == 1  # the plan is unharmed
    p.close()
    ap = magforce.Plan(ctx, dev(bad), 2.0)
    out = ap.field(dev(np.full_like(POS, 0.3)))
    torch.cuda.synchronize()
    assert not torch.isfinite(out).all()
    ap.close()
