"""Pins the CPU oracle (oracle/mf_oracle.c) against facts that do not come from
the oracle's own formulas: closed forms, a physical derivation, invariants and
brute force.  The reference ships no tests or golden vectors for this path
(its only analytic check, DipoleHessianSpherical.hpp:322-376, is for the FMM
expansions), so these are the pins.  CPU only.
"""
import numpy as np
import pytest
from scipy import integrate

from conftest import jittered_block, random_moments, rel_err

H = 2.0


def m4_spline(r, h=H):
    """Cubic M4 spline with sigma = 1/(pi h^3), written from its textbook form."""
    q = r / h
    s = 1.0 / (np.pi * h ** 3)
    if q < 1:
        return s * (1 - 1.5 * q * q + 0.75 * q ** 3)
    if q < 2:
        return s * 0.25 * (2 - q) ** 3
    return 0.0


def ball_mean(r, h=H, dens=m4_spline):
    return 3.0 / r ** 3 * integrate.quad(lambda s: s * s * dens(s, h), 0, r,
                                         points=[h, 2 * h], limit=200)[0]


# ---- A1 weights -------------------------------------------------------------

@pytest.mark.parametrize("q", [0.05, 0.3, 0.9, 1.0, 1.2, 1.7, 1.99, 2.0, 2.5, 7.0])
def test_weights_are_spline_and_its_ball_average(oracle, q):
    r = q * H
    wa, wd = oracle.weights(r, H)
    assert wd == pytest.approx(m4_spline(r), rel=1e-12, abs=1e-15)
    assert wa == pytest.approx(ball_mean(r), rel=1e-9)


def test_weights_at_origin_and_continuity(oracle):
    sigma = 1 / (np.pi * H ** 3)
    assert oracle.weights(0.0, H) == pytest.approx((sigma, sigma), rel=1e-14)
    for q in (1.0, 2.0):
        a = oracle.weights(q * H * (1 - 1e-12), H)
        b = oracle.weights(q * H, H)
        assert a[0] == pytest.approx(b[0], rel=1e-9)
        assert a[1] == pytest.approx(b[1], rel=1e-9, abs=1e-30)


# ---- A2 field pair -------------------------------------------------------------

def smoothed_field_fd(d, m, h=H, dens=m4_spline, eps=1e-4):
    """B/mu0 of a dipole m smeared with `dens`: H = (m.grad) grad Phi with
    grad Phi = -(r wbar/3) rhat (wbar = ball mean of dens, by quadrature),
    differentiated numerically, plus M = dens m."""
    def gphi(x):
        r = np.linalg.norm(x)
        return -(r * ball_mean(r, h, dens) / 3.0) * x / r
    Hf = np.zeros(3)
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        Hf += m[a] * (gphi(d + e) - gphi(d - e)) / (2 * eps)
    return Hf + dens(np.linalg.norm(d), h) * m


@pytest.mark.parametrize("q", [0.2, 0.7, 1.3, 1.8, 2.6, 5.0])
def test_field_pair_is_field_of_smeared_dipole(oracle, q):
    rng = np.random.default_rng(int(q * 100))
    d = rng.normal(size=3)
    d *= q * H / np.linalg.norm(d)
    m = rng.normal(size=3)
    got = oracle.field_pair(d, m, H)
    want = smoothed_field_fd(d, m)
    assert rel_err(got, want) < 1e-6


def test_field_pair_far_is_point_dipole_and_self_term(oracle):
    rng = np.random.default_rng(0)
    for _ in range(20):
        d = rng.normal(size=3)
        d *= rng.uniform(2.0, 30.0) * H / np.linalg.norm(d)
        m = rng.normal(size=3)
        r = np.linalg.norm(d)
        u = d / r
        want = (3 * u * (u @ m) - m) / (4 * np.pi * r ** 3)
        assert rel_err(oracle.field_pair(d, m, H), want) < 1e-13
    m = np.array([0.3, -1.2, 0.7])
    sigma = 1 / (np.pi * H ** 3)
    assert rel_err(oracle.field_pair(np.zeros(3), m, H), 2.0 / 3.0 * sigma * m) < 1e-14


def test_field_pair_symmetry_and_linearity(oracle):
    rng = np.random.default_rng(5)
    for _ in range(30):
        d = rng.normal(size=3) * rng.uniform(0.1, 3.0)
        K = np.stack([oracle.field_pair(d, e, H) for e in np.eye(3)], axis=1)
        Km = np.stack([oracle.field_pair(-d, e, H) for e in np.eye(3)], axis=1)
        assert np.abs(K - K.T).max() < 1e-15
        assert np.abs(K - Km).max() < 1e-15
        m = rng.normal(size=3)
        assert rel_err(oracle.field_pair(d, m, H), K @ m) < 1e-13


# ---- A4 force tensor --------------------------------------------------------------

def w2(r, h=H):
    """W*W (the density of two smeared particles), radial 3-D convolution."""
    if r == 0:
        return integrate.quad(lambda s: 4 * np.pi * s * s * m4_spline(s, h) ** 2, 0, 2 * h,
                              points=[h], limit=200)[0]

    def inner(s):
        lo, hi = abs(r - s), min(r + s, 2 * h)
        if lo >= hi:
            return 0.0
        return s * m4_spline(s, h) * integrate.quad(lambda t: t * m4_spline(t, h), lo, hi,
                                                    points=[h] if lo < h < hi else None,
                                                    limit=200)[0]
    return 2 * np.pi / r * integrate.quad(inner, 0, 2 * h, points=[h, r], limit=200)[0]


@pytest.mark.parametrize("q", [0.6, 1.4, 2.3, 3.3, 3.9])
def test_fitted_curves_are_force_between_smeared_dipoles(oracle, q):
    """The reference's fitted quartics (Dipole_hessian.kern:53-61) fit
    mu0 grad H of a dipole smeared with W*W; they agree to the fit's own
    accuracy (<= 2.2e-3 of the largest component, worst just above q = 2),
    which a wrong coefficient, sign or interval fails by orders of magnitude."""
    rng = np.random.default_rng(int(q * 10))
    d = rng.normal(size=3)
    d *= q * H / np.linalg.norm(d)
    m = rng.normal(size=3)
    T = oracle.hessian_pair(d, m, H, "iso")

    def Hs(x):
        r = np.linalg.norm(x)
        u = x / r
        wb = 3.0 / r ** 3 * integrate.quad(lambda s: s * s * w2(s), 0, r, limit=100,
                                           points=[H, 2 * H, 3 * H])[0]
        return (wb - w2(r)) * u * (u @ m) - (wb / 3) * m
    eps = 1e-4
    G = np.array([(Hs(d + e) - Hs(d - e)) / (2 * eps) for e in np.eye(3) * eps])
    assert np.abs(T - oracle.MU0 * G).max() / np.abs(T).max() < 4e-3


def test_force_tensor_far_is_gradient_of_dipole_field(oracle):
    rng = np.random.default_rng(2)
    for _ in range(10):
        d = rng.normal(size=3)
        d *= rng.uniform(4.2, 20.0) * H / np.linalg.norm(d)
        m = rng.normal(size=3)
        T = oracle.hessian_pair(d, m, H, "frame")
        eps = 1e-5 * np.linalg.norm(d)
        G = np.array([(oracle.field_pair(d + e, m, H) - oracle.field_pair(d - e, m, H)) / (2 * eps)
                      for e in np.eye(3) * eps])
        assert np.abs(T - oracle.MU0 * G).max() / np.abs(T).max() < 1e-7


def test_force_tensor_frame_equals_isotropic_and_symmetric(oracle):
    rng = np.random.default_rng(7)
    for _ in range(500):
        d = rng.normal(size=3)
        d *= rng.uniform(0.01, 10.0) / np.linalg.norm(d)
        m = rng.normal(size=3)
        A = oracle.hessian_pair(d, m, H, "frame")
        B = oracle.hessian_pair(d, m, H, "iso")
        assert np.abs(A - B).max() <= 1e-8 * np.abs(B).max()
        assert np.abs(A - A.T).max() <= 1e-8 * np.abs(A).max()
    assert np.all(oracle.hessian_pair(np.zeros(3), np.ones(3), H) == 0)


def test_curves_meet_far_field_at_q4(oracle):
    c1, c2 = oracle.curves(4.0)
    # far field in the same units: a h^4 = 3 mu0/(4 pi q^4), (b + 3a) h^4 = -6 mu0/(4 pi q^4)
    assert c1 == pytest.approx(3e-7 / 256, rel=2e-3)
    assert c2 == pytest.approx(-6e-7 / 256, rel=2e-3)
    assert oracle.curves(0.0) == (0.0, 0.0)
    assert oracle.curves(4.0001) == (0.0, 0.0)


# ---- A3 / A5 / A7 all-pairs, force, solve ---------------------------------------

def test_G_symmetric_translation_rotation(oracle):
    pos = jittered_block(4, 4, 3, seed=3)
    n = pos.shape[0]
    rng = np.random.default_rng(0)
    u, v = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    Gu = oracle.field_allpairs(pos, pos, u, H)
    Gv = oracle.field_allpairs(pos, pos, v, H)
    assert abs((v * Gu).sum() - (u * Gv).sum()) < 1e-12 * np.abs(v * Gu).sum()
    Gt = oracle.field_allpairs(pos + [3.5, -2.0, 11.0], pos + [3.5, -2.0, 11.0], u, H)
    assert rel_err(Gt, Gu) < 1e-12
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Gr = oracle.field_allpairs(pos @ Q.T, pos @ Q.T, u @ Q.T, H)
    assert rel_err(Gr, Gu @ Q.T) < 1e-12


def test_dense_G_matches_allpairs(oracle):
    pos = jittered_block(3, 3, 3, seed=4)
    m = random_moments(pos.shape[0], seed=2)
    G = oracle.dense_G(pos, H)
    assert rel_err(G @ m.ravel(), oracle.field_allpairs(pos, pos, m, H).ravel()) < 1e-13
    assert np.abs(G - G.T).max() < 1e-15


def test_two_body_force_textbook(oracle):
    """Two dipoles beyond 4h: F1 = 3 mu0/(4 pi r^5)[(m1.r)m2 + (m2.r)m1 + (m1.m2)r
    - 5 (m1.r)(m2.r) r / r^2], r = x1 - x2; and F2 = -F1."""
    x = np.array([[0.0, 0.0, 0.0], [3.0, 7.0, -6.0]])
    m = np.array([[0.2, -0.5, 0.9], [1.1, 0.3, -0.4]])
    F = oracle.force_allpairs(x, m, H)
    r = x[0] - x[1]
    rn = np.linalg.norm(r)
    m1, m2 = m
    want = 3 * oracle.MU0 / (4 * np.pi * rn ** 5) * (
        (m1 @ r) * m2 + (m2 @ r) * m1 + (m1 @ m2) * r - 5 * (m1 @ r) * (m2 @ r) * r / rn ** 2)
    assert rel_err(F[0], want) < 1e-13
    assert rel_err(F[1], -want) < 1e-13


def test_momentum_conservation(oracle):
    pos = jittered_block(4, 4, 4, seed=6)
    m = random_moments(pos.shape[0], seed=3)
    F = oracle.force_allpairs(pos, m, H)
    # exact antisymmetry up to the reference frame form's 1e-10 pole patch
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).sum()


def test_cg_matches_dense_solve(oracle):
    pos = jittered_block(4, 3, 3, seed=8)
    n = pos.shape[0]
    kappa = 0.5
    b = np.tile([0.0, 0.0, 1.0], (n, 1))
    A = np.eye(3 * n) - kappa * oracle.dense_G(pos, H)
    x_dense = np.linalg.solve(A, b.ravel()).reshape(n, 3)
    x, it, rel = oracle.cg_solve(pos, H, kappa, b, rtol=1e-13)
    assert rel < 1e-13 and it < 100
    assert rel_err(x, x_dense) < 1e-11
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > 0.0  # SPD


def test_reference_nlcg_descends_toward_solution(oracle):
    """The reference's own solver (0.1 RMS tolerance) gets close to the exact
    solution; with a tight tolerance it converges to it."""
    pos = jittered_block(4, 4, 2, seed=9)
    n = pos.shape[0]
    kappa = 0.5
    b = np.tile([0.0, 1.0, 0.0], (n, 1))
    x_ref, _, _ = oracle.cg_solve(pos, H, kappa, b, rtol=1e-13)
    x, hist = oracle.ref_nlcg_solve(pos, H, kappa, b, tol=1e-9, max_iteration=80)
    assert np.all(np.diff(hist) <= 0)
    assert rel_err(x, x_ref) < 1e-6
    x_loose, hist2 = oracle.ref_nlcg_solve(pos, H, kappa, b, tol=0.1)
    assert np.sqrt(2 * hist2[-1] / n) < 0.2


def test_f32_restatements_track_fp64(oracle):
    """The reference's FP32 op order stays within ~1e-4 of FP64 (its
    Horner-form curves cancel near q = 4, so not better)."""
    pos = jittered_block(6, 6, 6, seed=10)
    m = random_moments(pos.shape[0], seed=4)
    f64 = oracle.field_allpairs(pos, pos, m, H)
    f32 = oracle.field_allpairs_f32ref(pos, m, H)
    assert rel_err(f32, f64) < 1e-5
    F64 = oracle.force_allpairs(pos, m, H)
    F32 = oracle.force_allpairs_f32ref(pos, m, H)
    assert rel_err(F32, F64) < 1e-3
