"""CPU pins of the FMM near-field coefficient tables (make_field_table and
make_hess_table in magnetic-force_amd/csrc/mf_allpairs.hip).

Force: against the oracle's fitted curves (oracle/mf_oracle.c mfo_curves,
the reference's Dipole_hessian.kern:53-61 quartic pieces on (0, 1], (1, 2],
(2, 3], (3, 4]).  The lookup below restates the device function
hess_near_curves_tab (mf_pair.hpp): u = q kHessTabN / 4, entry e = ceil(u),
t = u - e, two cubics in t.  The pieces jump at integer q (C1 by 7e-4 of its
value at q = 1), so a table that put an integer q into the upper piece, or an
interval that straddled two pieces, would miss by far more than the bound.

Field: against the oracle's pair field (mfo_field_pair, the reference's
Dipole_B.kern:79-104), from which alpha and beta follow exactly (a moment
across d and one along d).  The lookup restates field_near_coeffs_tab:
u = q kFieldTabN / 2, entry k = floor(u), t = u - k.  q = 1, where the two
forms of the smoothing kernel meet, is an interval boundary.
"""
import ctypes

import numpy as np
import pytest

import magforce

KMU0 = 1.25663706e-6
NTAB = 64
ENTRIES = {0: NTAB + 1, 1: NTAB + 2}  # padding entries at the ends of the range


def table(kind, h):
    L = magforce.lib()
    f = L.mf_internal_near_table
    f.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_float)]
    f.restype = ctypes.c_int
    n = ENTRIES[kind]
    out = np.zeros(8 * n, dtype=np.float32)
    assert f(kind, h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    return out.reshape(n, 8).astype(np.float64)


def cubics(tab, e, t):
    a, b = tab[e, :4], tab[e, 4:]
    return (((a[:, 3] * t + a[:, 2]) * t + a[:, 1]) * t + a[:, 0],
            ((b[:, 3] * t + b[:, 2]) * t + b[:, 1]) * t + b[:, 0])


def lookup(tab, q):
    u = q * NTAB / 4.0
    e = np.ceil(u)
    return cubics(tab, e.astype(int), u - e)


def field_lookup(tab, q):
    u = q * NTAB / 2.0
    k = np.floor(u)
    return cubics(tab, k.astype(int), u - k)


@pytest.mark.parametrize("h", [1.0, 2.0, 3.5])
def test_force_curve_table_matches_oracle(oracle, h):
    tab = table(1, h)
    scale = 1.0 / (3.0 * KMU0 / (4.0 * np.pi) * h ** 4)
    rng = np.random.default_rng(3)
    ints = np.array([1.0, 2.0, 3.0, 4.0])
    q = np.concatenate([rng.uniform(1e-6, 4.0, 4000), ints, ints - 1e-9, ints[:3] + 1e-9,
                        np.linspace(1e-6, 4.0, 4001)])
    ref = np.array([oracle.curves(float(x)) for x in q])
    c1_ref = scale * ref[:, 0]
    dq_ref = -scale * (ref[:, 1] - 3.0 * ref[:, 0]) / 5.0
    c1, dq = lookup(tab, q)
    big1, bigd = np.abs(c1_ref).max(), np.abs(dq_ref).max()
    e1 = np.abs(c1 - c1_ref).max() / big1
    ed = np.abs(dq - dq_ref).max() / bigd
    print(f"h={h}: table error C1 {e1:.2e}, (C2 - 3 C1) / -5 {ed:.2e} of the largest value")
    assert e1 < 1.5e-7 and ed < 1.5e-7
    # at an integer q the lower piece applies (reference: q <= 1 takes piece 0)
    jump = abs(oracle.curves(1.0)[0] - oracle.curves(1.0 + 1e-12)[0]) * scale / big1
    assert jump > 100 * 1.5e-7


def test_force_curve_table_padding():
    """u = 0 (a pair at d = 0) and u just above kHessTabN (rounding at q -> 4)
    read the padding entries, which continue their neighbours' cubics."""
    tab = table(1, 2.0)
    big = np.abs(tab[:, [0, 4]]).max(axis=0)  # largest value of each curve

    def close(x, y):
        return all(abs(a[0] - b[0]) < 1.5e-7 * g for a, b, g in zip(x, y, big))

    assert close(cubics(tab, np.array([0]), np.array([0.0])),
                 cubics(tab, np.array([1]), np.array([-1.0])))
    assert close(cubics(tab, np.array([NTAB + 1]), np.array([-1.0 + 1e-6])),
                 cubics(tab, np.array([NTAB]), np.array([1e-6])))


def field_coeffs_ref(oracle, q, h):
    """alpha, beta of the pair c (3 alpha d (d.m) - beta m), c = 1 / (4 pi), from
    the oracle's pair field: a moment across d gives -c beta m, one along d
    gives c (3 alpha r^2 - beta) m."""
    c = 0.25 / np.pi
    r = q * h
    d = [r, 0.0, 0.0]
    beta = -oracle.field_pair(d, [0.0, 1.0, 0.0], h)[1] / c
    along = oracle.field_pair(d, [1.0, 0.0, 0.0], h)[0] / c
    alpha = (along + beta) / (3.0 * r * r)
    return alpha, beta


@pytest.mark.parametrize("h", [1.0, 2.0, 3.5])
def test_field_coeff_table_matches_oracle(oracle, h):
    tab = table(0, h)
    rng = np.random.default_rng(4)
    q = np.concatenate([rng.uniform(1e-3, 2.0, 3000), [1.0, 1.0 - 1e-9, 1.0 + 1e-9, 0.5, 1.5],
                        np.linspace(1e-3, 2.0 - 1e-9, 2001)])
    ref = np.array([field_coeffs_ref(oracle, float(x), h) for x in q])
    alpha, beta = field_lookup(tab, q)
    ea = np.abs(alpha - ref[:, 0]) / np.maximum(np.abs(ref[:, 0]), (2 * h) ** -5)
    eb = np.abs(beta - ref[:, 1]) / np.maximum(np.abs(ref[:, 1]), (2 * h) ** -3)
    print(f"h={h}: field table error alpha {ea.max():.2e}, beta {eb.max():.2e}")
    assert ea.max() < 1.5e-7 and eb.max() < 1.5e-7
    # q -> 0: the closed form of the inner interval (alpha(0) = 0.8 h^-5)
    a0, b0 = cubics(tab, np.array([0]), np.array([0.0]))
    assert abs(a0[0] - 0.8 / h ** 5) < 1e-6 * 0.8 / h ** 5
    # the padding entry continues the last interval: u rounded up to kFieldTabN
    pad = cubics(tab, np.array([NTAB]), np.array([1e-6]))
    last = cubics(tab, np.array([NTAB - 1]), np.array([1.0 + 1e-6]))
    assert abs(pad[0][0] - last[0][0]) < 1.5e-7 * max(abs(last[0][0]), (2 * h) ** -5)
    assert abs(pad[1][0] - last[1][0]) < 1.5e-7 * max(abs(last[1][0]), (2 * h) ** -3)
