"""ctypes wrapper of the CPU oracle (oracle/lib/libmf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package never
imports this module.

Arrays are numpy float64 (n, 3) for positions/moments/fields and (n, 3, 3)
for tensors; the FP32 restatements take and return float32 (n, 3).
Every function restates a reference function (file:line in mf_oracle.h).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libmf_oracle.so")
_lib = None

MU0 = 1.25663706e-6


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        i64, dbl, flt, cint = ctypes.c_int64, ctypes.c_double, ctypes.c_float, ctypes.c_int
        L.mfo_weights.argtypes = [dbl, dbl, dp, dp]
        L.mfo_field_pair.argtypes = [dp, dp, dbl, dp]
        L.mfo_hessian_pair_frame.argtypes = [dp, dp, dbl, dp]
        L.mfo_hessian_pair_iso.argtypes = [dp, dp, dbl, dp]
        L.mfo_curves.argtypes = [dbl, dp, dp]
        L.mfo_field_allpairs.argtypes = [i64, dp, i64, dp, dp, dbl, dp, dp, cint]
        L.mfo_hessian_allpairs.argtypes = [i64, dp, i64, dp, dp, dbl, dp, cint]
        L.mfo_force_allpairs.argtypes = [i64, dp, dp, dbl, dp, dp, cint]
        L.mfo_apply_A.argtypes = [i64, dp, dbl, dbl, dp, dp, cint]
        L.mfo_dense_G.argtypes = [i64, dp, dbl, dp]
        L.mfo_cg_solve.argtypes = [i64, dp, dbl, dbl, dp, dp, dbl, cint, dp, cint]
        L.mfo_cg_solve.restype = cint
        L.mfo_ref_nlcg_solve.argtypes = [i64, dp, dbl, dbl, dp, dp, cint, dbl, cint, dp, cint, cint]
        L.mfo_ref_nlcg_solve.restype = cint
        L.mfo_field_allpairs_f32ref.argtypes = [i64, fp, fp, fp, fp, fp, fp, flt, fp, fp, fp, cint]
        L.mfo_force_allpairs_f32ref.argtypes = [i64, fp, fp, fp, fp, fp, fp, flt, fp, fp, fp, cint]
        L.mfo_fmm_field.argtypes = [i64, dp, dp, dbl, cint, dbl, dbl, dp, dp, cint,
                                    ctypes.POINTER(FmmStats)]
        L.mfo_fmm_field.restype = cint
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def weights(r: float, h: float):
    wa, wd = ctypes.c_double(), ctypes.c_double()
    lib().mfo_weights(r, h, ctypes.byref(wa), ctypes.byref(wd))
    return wa.value, wd.value


def curves(q: float):
    c1, c2 = ctypes.c_double(), ctypes.c_double()
    lib().mfo_curves(q, ctypes.byref(c1), ctypes.byref(c2))
    return c1.value, c2.value


def field_pair(d, m, h):
    d, dp_ = _d(d)
    m, mp = _d(m)
    out = np.zeros(3)
    _, op = _d(out)
    lib().mfo_field_pair(dp_, mp, h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def hessian_pair(d, m, h, form="frame"):
    d, dp_ = _d(d)
    m, mp = _d(m)
    out = np.zeros(9)
    fn = lib().mfo_hessian_pair_frame if form == "frame" else lib().mfo_hessian_pair_iso
    fn(dp_, mp, h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out.reshape(3, 3)


class FmmStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("boxes", "leaves", "levels", "near_box_pairs",
                                               "far_box_pairs", "near_particle_pairs")]


def fmm_field(pos, mom, h, ncrit=384, theta=0.5, gap=None, hconst=None, nthreads=0):
    """The reference's FMM field product restated on the CPU in FP64
    (mf_cpufmm.c): ncrit-split octree, dual traversal, P = 6.  gap defaults to
    the reference's 3h.  Returns (field (n, 3), stats dict)."""
    pos, pp = _d(pos)
    mom, mp = _d(mom)
    out = np.zeros((pos.shape[0], 3))
    hc = None
    if hconst is not None:
        hcv, hc = _d(np.asarray(hconst, dtype=np.float64).reshape(3))
    st = FmmStats()
    lib().mfo_fmm_field(pos.shape[0], pp, mp, h, int(ncrit), float(theta),
                        float(3.0 * h if gap is None else gap), hc,
                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nthreads,
                        ctypes.byref(st))
    return out, {k: getattr(st, k) for k, _ in FmmStats._fields_}


def field_allpairs(tpos, spos, smom, h, hconst=None, nthreads=0):
    tpos, tp = _d(tpos)
    spos, sp = _d(spos)
    smom, mp = _d(smom)
    out = np.zeros((tpos.shape[0], 3))
    hc = None
    if hconst is not None:
        hcv, hc = _d(np.asarray(hconst, dtype=np.float64).reshape(3))
    lib().mfo_field_allpairs(tpos.shape[0], tp, spos.shape[0], sp, mp, h, hc,
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nthreads)
    return out


def hessian_allpairs(tpos, spos, smom, h, nthreads=0):
    tpos, tp = _d(tpos)
    spos, sp = _d(spos)
    smom, mp = _d(smom)
    out = np.zeros((tpos.shape[0], 9))
    lib().mfo_hessian_allpairs(tpos.shape[0], tp, spos.shape[0], sp, mp, h,
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nthreads)
    return out.reshape(-1, 3, 3)


def force_allpairs(pos, mom, h, text=None, nthreads=0):
    pos, pp = _d(pos)
    mom, mp = _d(mom)
    tp = None
    if text is not None:
        tv, tp = _d(np.asarray(text).reshape(-1, 9))
    out = np.zeros((pos.shape[0], 3))
    lib().mfo_force_allpairs(pos.shape[0], pp, mp, h, tp,
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nthreads)
    return out


def apply_A(pos, h, kappa, x, nthreads=0):
    pos, pp = _d(pos)
    x, xp = _d(x)
    y = np.zeros_like(x)
    lib().mfo_apply_A(pos.shape[0], pp, h, kappa, xp,
                      y.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nthreads)
    return y


def dense_G(pos, h):
    pos, pp = _d(pos)
    n = pos.shape[0]
    G = np.zeros((3 * n, 3 * n))
    lib().mfo_dense_G(n, pp, h, G.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return G


def cg_solve(pos, h, kappa, b, x0=None, rtol=1e-12, maxit=1000, nthreads=0):
    pos, pp = _d(pos)
    b, bp = _d(b)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    rel = ctypes.c_double()
    it = lib().mfo_cg_solve(pos.shape[0], pp, h, kappa, bp,
                            x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), rtol, maxit,
                            ctypes.byref(rel), nthreads)
    return x, it, rel.value


def ref_nlcg_solve(pos, h, kappa, hext, x0=None, niter=100, tol=0.1, max_iteration=80,
                   nthreads=0):
    pos, pp = _d(pos)
    hext, hp = _d(hext)
    x = np.zeros_like(hext) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max_iteration + 2)
    nh = lib().mfo_ref_nlcg_solve(pos.shape[0], pp, h, kappa, hp,
                                  x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), niter, tol,
                                  max_iteration,
                                  hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  hist.shape[0], nthreads)
    return x, hist[:nh]


def field_allpairs_f32ref(pos, mom, h, nthreads=0):
    """FP32 in cuda_Apply_G's own operation order (targets == sources)."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    mom = np.ascontiguousarray(mom, dtype=np.float32)
    cols = [np.ascontiguousarray(pos[:, k]) for k in range(3)] + \
           [np.ascontiguousarray(mom[:, k]) for k in range(3)]
    outs = [np.zeros(pos.shape[0], dtype=np.float32) for _ in range(3)]
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    lib().mfo_field_allpairs_f32ref(pos.shape[0], *[P(c) for c in cols], h,
                                    *[P(o) for o in outs], nthreads)
    return np.stack(outs, axis=1)


def force_allpairs_f32ref(pos, mom, h, nthreads=0):
    """FP32 in fitted_force_ofM_kernel's own operation order (no external tensor)."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    mom = np.ascontiguousarray(mom, dtype=np.float32)
    cols = [np.ascontiguousarray(pos[:, k]) for k in range(3)] + \
           [np.ascontiguousarray(mom[:, k]) for k in range(3)]
    outs = [np.zeros(pos.shape[0], dtype=np.float32) for _ in range(3)]
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    lib().mfo_force_allpairs_f32ref(pos.shape[0], *[P(c) for c in cols], h,
                                    *[P(o) for o in outs], nthreads)
    return np.stack(outs, axis=1)
