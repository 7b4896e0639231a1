"""CPU pin of the FMM force near-field curve table (make_hess_table in
magnetic-force_amd/csrc/mf_allpairs.hip) against the oracle's fitted curves
(oracle/mf_oracle.c mfo_curves, the reference's Dipole_hessian.kern:53-61
quartic pieces on (0, 1], (1, 2], (2, 3], (3, 4]).

The lookup below restates the device function hess_near_coeffs_tab
(mf_pair.hpp): u = q kHessTabN / 4, kc = clamp(ceil(u), 1, kHessTabN),
t = u - kc, two cubics in t.  The pieces jump at integer q (C1 by 7e-4 of its
value at q = 1), so a table that put an integer q into the upper piece, or an
interval that straddled two pieces, would miss by far more than the bound.
"""
import ctypes

import numpy as np
import pytest

import magforce

KMU0 = 1.25663706e-6
NTAB = 64


def table(kind, h):
    L = magforce.lib()
    f = L.mf_internal_near_table
    f.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_float)]
    f.restype = ctypes.c_int
    n = {0: 64, 1: NTAB}[kind]
    out = np.zeros(8 * n, dtype=np.float32)
    assert f(kind, h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    return out.reshape(n, 8).astype(np.float64)


def lookup(tab, q):
    u = q * NTAB / 4.0
    kc = np.clip(np.ceil(u), 1, NTAB)
    t = u - kc
    k = kc.astype(int) - 1
    a, b = tab[k, :4], tab[k, 4:]
    c1 = ((a[:, 3] * t + a[:, 2]) * t + a[:, 1]) * t + a[:, 0]
    dq = ((b[:, 3] * t + b[:, 2]) * t + b[:, 1]) * t + b[:, 0]
    return c1, dq


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
