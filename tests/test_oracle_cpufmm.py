"""CPU FMM restatement (oracle/mf_cpufmm.c), the reference's fmmtl field
product in FP64: pinned against the FP64 all-pairs oracle and limits the
algorithm fixes.  Test infrastructure for the CPU baseline of bench.py."""
import numpy as np
import pytest

from conftest import jittered_block, random_moments, rel_err


@pytest.fixture(scope="module")
def case(oracle):
    pos = jittered_block(20, 20, 10, seed=81)
    m = random_moments(pos.shape[0], seed=82, scale=0.4)
    return pos, m, oracle.field_allpairs(pos, pos, m, 2.0)


def test_matches_allpairs_with_a_far_field(oracle, case):
    """Leaves of at most 16 points: most box pairs are translated (S2M, M2M,
    M2L, L2L, L2T all run), and P = 6 at theta = 0.5 keeps the field within a
    few 1e-5 of the direct sum."""
    pos, m, ref = case
    got, st = oracle.fmm_field(pos, m, 2.0, ncrit=16)
    assert st["far_box_pairs"] > st["near_box_pairs"] / 10 > 0
    assert st["levels"] >= 4
    assert rel_err(got, ref) < 1e-4


def test_error_falls_with_the_opening_angle(oracle, case):
    """A wrong translation would leave an error that does not fall with theta."""
    pos, m, ref = case
    errs = [rel_err(oracle.fmm_field(pos, m, 2.0, ncrit=16, theta=t)[0], ref)
            for t in (0.7, 0.5, 0.3)]
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-5


def test_no_far_pairs_is_the_direct_sum(oracle, case):
    """theta -> 0 accepts no box pair: every pair is near, the result is the
    all-pairs sum up to its order."""
    pos, m, ref = case
    got, st = oracle.fmm_field(pos, m, 2.0, ncrit=16, theta=1e-6)
    assert st["far_box_pairs"] == 0
    assert rel_err(got, ref) < 1e-13


def test_gap_keeps_smoothed_pairs_near(oracle):
    """Two clusters 5h apart with boxes of radius below h: theta alone would
    translate them, the 3h gap of the reference's MAC keeps them near at
    small leaves; the far field of a distant cluster is the point-dipole
    field, which the expansion reproduces."""
    rng = np.random.default_rng(83)
    a = rng.uniform(-0.5, 0.5, size=(300, 3))
    b = rng.uniform(-0.5, 0.5, size=(300, 3)) + [40.0, 0.0, 0.0]
    pos = np.concatenate([a, b])
    m = random_moments(pos.shape[0], seed=84)
    ref = oracle.field_allpairs(pos, pos, m, 2.0)
    got, st = oracle.fmm_field(pos, m, 2.0, ncrit=64)
    assert st["far_box_pairs"] > 0
    assert rel_err(got, ref) < 1e-5


def test_constant_field_and_single_point(oracle):
    pos = np.array([[1.0, 2.0, 3.0]])
    m = np.array([[0.0, 0.0, 1.0]])
    got, st = oracle.fmm_field(pos, m, 2.0, hconst=[0.0, 0.5, 1.0])
    want = oracle.field_allpairs(pos, pos, m, 2.0, hconst=[0.0, 0.5, 1.0])
    assert st["boxes"] == 1
    assert np.allclose(got, want, rtol=1e-14, atol=0)
