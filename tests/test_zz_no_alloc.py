"""The allocation-free field product (mf_field_allpairs_noalloc) against the
allocating one (mf_field_allpairs) and the FP64 oracle, through the C ABI.

Both entry points run the same kernels; at these sizes they pick the same
source-chunk split, so their outputs agree bitwise.  Two calls in a row, the
second smaller, show that nothing the first call left in the shared scratch
array leaks into the second.
"""
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


def _case(shape_t, shape_s, seed, shift):
    tpos = f32(jittered_block(*shape_t, seed=seed))
    spos = f32(jittered_block(*shape_s, seed=seed + 1) + shift)
    m = f32(random_moments(spos.shape[0], seed=seed + 2))
    return tpos, spos, m


def test_noalloc_matches_allocating_entry_point(ctx, oracle):
    import magforce
    hc = [0.0, 0.25, 1.0]
    # first call: 4096 targets, 3000 sources; second: smaller on both sides
    cases = [((16, 16, 16), (10, 15, 20), 11, [0.4, -0.3, 0.2], 2.0),
             ((5, 7, 3), (4, 4, 9), 23, [1.1, 0.0, -0.6], 1.5)]
    for shape_t, shape_s, seed, shift, h in cases:
        tpos, spos, m = _case(shape_t, shape_s, seed, shift)
        print(f"noalloc case n_t={len(tpos)} n_s={len(spos)} h={h}", flush=True)
        got = host(magforce.field_allpairs_noalloc(ctx, dev(tpos), dev(spos), dev(m), h, hconst=hc))
        ref = host(magforce.field_allpairs(ctx, dev(tpos), dev(spos), dev(m), h, hconst=hc))
        want = oracle.field_allpairs(tpos, spos, m, h, hconst=hc)
        assert np.array_equal(got, ref)
        assert rel_err(got, want) < 1e-5


def test_noalloc_into_preallocated_output_repeated(ctx, oracle):
    """The serving-loop shape: the output tensor is allocated once, the entry
    point is called repeatedly on shrinking inputs."""
    import torch
    import magforce
    tpos, spos, m = _case((8, 8, 8), (8, 8, 8), 5, [0.0, 0.0, 0.0])
    out = torch.empty((3, len(tpos)), dtype=torch.float32, device="cuda")
    for n_t in (512, 300, 1):
        print(f"noalloc repeated n_t={n_t}", flush=True)
        magforce.field_allpairs_noalloc(ctx, dev(tpos[:n_t]), dev(spos), dev(m), 2.0, out=out[:, :n_t])
        want = oracle.field_allpairs(tpos[:n_t], spos, m, 2.0)
        assert rel_err(host(out[:, :n_t]), want) < 1e-5


def test_noalloc_empty_and_no_sources(ctx, oracle):
    import magforce
    tpos, spos, m = _case((3, 3, 3), (2, 2, 2), 9, [0.5, 0.5, 0.5])
    hc = [0.3, -0.1, 2.0]
    got = host(magforce.field_allpairs_noalloc(ctx, dev(tpos), dev(spos[:0]), dev(m[:0]), 2.0,
                                               hconst=hc))
    assert np.allclose(got, np.tile(np.float32(hc), (len(tpos), 1)), rtol=0, atol=0)
    empty = magforce.field_allpairs_noalloc(ctx, dev(tpos[:0]), dev(spos), dev(m), 2.0)
    assert empty.shape[1] == 0


def test_noalloc_rejects_sizes_beyond_scratch(ctx):
    """The size check runs on the host before anything is launched."""
    import torch
    import magforce
    cap = magforce.noalloc_scratch_bytes()
    assert cap == 32 << 20
    n = cap // 32  # the target records alone fill the array
    t = torch.zeros((3, n), dtype=torch.float32, device="cuda")
    s = torch.zeros((3, 1), dtype=torch.float32, device="cuda")
    with pytest.raises(magforce.MagforceError) as e:
        magforce.field_allpairs_noalloc(ctx, t, s, s, 2.0)
    assert e.value.code == magforce.MF_ERR_INVALID_ARG
    assert "scratch" in str(e.value)
