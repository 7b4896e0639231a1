"""The device 3-vector layer against numpy: the fused C-ABI operations
(mf_vec_axpby / mul / dot / fill, their FP64 variants, mf_vec_convert) and the
drop-in class Libo::cuda_abstract_vector<T> for T = float, double and int
(cuda_abstract_vector.cuh:166-563), driven by lib/vector_check.

Bars: element-wise operations are single roundings of the same operands, so
they equal numpy's float32 / float64 results bitwise; reductions sum in FP64
and match numpy's FP64 sum of the same per-element products to 1e-12
relative (float inputs: each element's product is rounded once to FP32, so
the bar there is 1e-6 of sum |a_i b_i|).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 7, 1001, 65537]


def _load(d, name, dtype, n):
    a = np.fromfile(os.path.join(d, name + ".bin"), dtype=dtype)
    assert a.size == 3 * n, (name, a.size, n)
    return a.reshape(3, n)


@pytest.mark.parametrize("n", SIZES)
def test_abstract_vector_class(tmp_path, n):
    import magforce
    exe = os.path.join(os.path.dirname(magforce.EXAMPLE_PATH), "vector_check")
    print(f"vector_check n={n}", flush=True)
    out = subprocess.run([exe, str(n), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert f"VECTOR_CHECK_OK {n}" in out.stdout
    sc = {}
    for line in out.stdout.splitlines():
        parts = line.split()
        if len(parts) == 2 and parts[0] != "VECTOR_CHECK_OK":
            sc[parts[0]] = float(parts[1])
    d = str(tmp_path)
    for tag, dt in (("f32", np.float32), ("f64", np.float64)):
        a = _load(d, f"{tag}_a", dt, n)
        b = _load(d, f"{tag}_b", dt, n)
        c175, cm05, c3 = dt(1.75), dt(-0.5), dt(3.0)
        np.testing.assert_array_equal(_load(d, f"{tag}_add", dt, n), a + b)
        np.testing.assert_array_equal(_load(d, f"{tag}_sub", dt, n), a - b)
        np.testing.assert_array_equal(_load(d, f"{tag}_neg", dt, n), -a)
        np.testing.assert_array_equal(_load(d, f"{tag}_mul_r", dt, n), a * c175)
        np.testing.assert_array_equal(_load(d, f"{tag}_mul_l", dt, n), b * cm05)
        c = a + b
        np.testing.assert_array_equal(_load(d, f"{tag}_iadd", dt, n), c)
        c = (c - a) - a
        np.testing.assert_array_equal(_load(d, f"{tag}_isub", dt, n), c)
        np.testing.assert_array_equal(_load(d, f"{tag}_imul", dt, n), c * c3)
        np.testing.assert_array_equal(_load(d, f"{tag}_assign_self_add", dt, n), a + a)
        np.testing.assert_array_equal(_load(d, f"{tag}_a_after", dt, n), a)
        if tag == "f32":
            prod = (a * b).astype(np.float64)  # bound covers the FP32 rounding per element
            scale = np.abs(a.astype(np.float64) * b).sum()
            tol = 1e-6 * scale + 1e-30
        else:
            prod = a * b
            scale = np.abs(prod).sum()
            tol = 1e-12 * scale + 1e-300
        want = float(prod.sum())
        assert abs(sc[f"{tag}_dot_fp64"] - want) <= tol
        assert abs(sc[f"{tag}_dot"] - want) <= tol + abs(want) * (1e-7 if tag == "f32" else 1e-15)
        nn = float((a.astype(np.float64) ** 2).sum())
        assert abs(sc[f"{tag}_norm2"] - nn) <= 1e-6 * nn + 1e-30
        assert abs(sc[f"{tag}_norm"] - np.sqrt(nn)) <= 1e-6 * np.sqrt(nn) + 1e-30
        np.testing.assert_array_equal(_load(d, f"{tag}_to_f32", np.float32, n), a.astype(np.float32))
        np.testing.assert_array_equal(_load(d, f"{tag}_to_f64", np.float64, n), a.astype(np.float64))
        np.testing.assert_array_equal(_load(d, f"{tag}_to_i32", np.int32, n),
                                      np.trunc(a).astype(np.int32))
    k = _load(d, "i32_a", np.int32, n)
    np.testing.assert_array_equal(_load(d, "i32_copy", np.int32, n), k)
    np.testing.assert_array_equal(_load(d, "i32_to_f32", np.float32, n), k.astype(np.float32))
    np.testing.assert_array_equal(_load(d, "i32_to_f64", np.float64, n), k.astype(np.float64))
    if n > 0:
        k2 = k.copy()
        k2[:, n - 1] += np.array([5, -6, 7], dtype=np.int32)
        np.testing.assert_array_equal(_load(d, "i32_add_elem", np.int32, n), k2)


def _ptrs(t):
    return [ctypes.c_void_p(t[c].data_ptr()) for c in range(3)]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_vec_capi_ops(ctx, n, dtype):
    """The fused C entry points directly: axpby with and without b and with
    out aliasing a, element-wise mul, dot, fill."""
    import torch
    import magforce
    L = magforce.lib()
    td = torch.float32 if dtype == "f32" else torch.float64
    nd = np.float32 if dtype == "f32" else np.float64
    sfx = "" if dtype == "f32" else "_f64"
    cty = ctypes.c_float if dtype == "f32" else ctypes.c_double
    axpby, mul = getattr(L, "mf_vec_axpby" + sfx), getattr(L, "mf_vec_mul" + sfx)
    dot, fill = getattr(L, "mf_vec_dot" + sfx), getattr(L, "mf_vec_fill" + sfx)
    if dtype == "f64":
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        axpby.argtypes = [vp, i64, cty] + [vp] * 3 + [cty] + [vp] * 6
        mul.argtypes = [vp, i64] + [vp] * 9
        dot.argtypes = [vp, i64] + [vp] * 6 + [ctypes.POINTER(ctypes.c_double)]
        fill.argtypes = [vp, i64, cty, vp]
        for f in (axpby, mul, dot, fill):
            f.restype = ctypes.c_int
    rng = np.random.default_rng(n + 3)
    a_h = rng.normal(size=(3, n)).astype(nd)
    b_h = rng.normal(size=(3, n)).astype(nd)
    a = torch.tensor(a_h, dtype=td, device="cuda")
    b = torch.tensor(b_h, dtype=td, device="cuda")
    o = torch.empty_like(a)
    al, be = nd(0.625), nd(-1.5)
    assert axpby(ctx.h, n, cty(float(al)), *_ptrs(a), cty(float(be)), *_ptrs(b), *_ptrs(o)) == 0
    # out = fma(beta, b, alpha a): alpha a rounded once, then one rounding of the fma
    x1 = (al * a_h).astype(np.float64)
    exact = be.astype(np.float64) * b_h.astype(np.float64) + x1
    got = o.cpu().numpy().astype(np.float64)
    terms = np.abs(x1) + np.abs(be.astype(np.float64) * b_h)
    if dtype == "f32":
        # the FP64 value is exact to 2^-52 of the terms; the FMA rounds it once
        tol = np.abs(exact) * 2.0 ** -24 + terms * 2.0 ** -50
    else:
        # numpy's own sum rounds twice: a few units in the last place of the terms
        tol = terms * 2.0 ** -51
    assert np.all(np.abs(got - exact) <= tol + 1e-300)
    # b = NULL, out aliasing a: a <- alpha a
    assert axpby(ctx.h, n, cty(float(al)), *_ptrs(a), cty(0), None, None, None, *_ptrs(a)) == 0
    np.testing.assert_array_equal(a.cpu().numpy(), al * a_h)
    assert mul(ctx.h, n, *_ptrs(a), *_ptrs(b), *_ptrs(o)) == 0
    np.testing.assert_array_equal(o.cpu().numpy(), (al * a_h) * b_h)
    r = ctypes.c_double(-1.0)
    assert dot(ctx.h, n, *_ptrs(a), *_ptrs(b), ctypes.byref(r)) == 0
    prod = (al * a_h).astype(np.float64) * b_h.astype(np.float64)
    rtol = 1e-6 if dtype == "f32" else 1e-13
    assert abs(r.value - prod.sum()) <= rtol * np.abs(prod).sum() + 1e-300
    if n:
        assert fill(ctx.h, n, cty(2.5), ctypes.c_void_p(o[1].data_ptr())) == 0
        got = o.cpu().numpy()
        assert np.all(got[1] == nd(2.5))
        np.testing.assert_array_equal(got[0], (al * a_h[0]) * b_h[0])


def test_vec_convert_capi(ctx):
    import torch
    import magforce
    L = magforce.lib()
    L.mf_vec_convert.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_int, ctypes.c_void_p]
    L.mf_vec_convert.restype = ctypes.c_int
    x = torch.tensor([-2.75, -0.5, 0.0, 0.5, 3.99, 1e6], dtype=torch.float64, device="cuda")
    i = torch.empty(6, dtype=torch.int32, device="cuda")
    f = torch.empty(6, dtype=torch.float32, device="cuda")
    assert L.mf_vec_convert(ctx.h, 6, 1, ctypes.c_void_p(x.data_ptr()), 2,
                            ctypes.c_void_p(i.data_ptr())) == 0
    assert i.cpu().tolist() == [-2, 0, 0, 0, 3, 1000000]
    assert L.mf_vec_convert(ctx.h, 6, 2, ctypes.c_void_p(i.data_ptr()), 0,
                            ctypes.c_void_p(f.data_ptr())) == 0
    assert f.cpu().tolist() == [-2.0, 0.0, 0.0, 0.0, 3.0, 1e6]
    # an unknown type code is an argument error, raised before any launch
    assert L.mf_vec_convert(ctx.h, 6, 7, ctypes.c_void_p(x.data_ptr()), 0,
                            ctypes.c_void_p(f.data_ptr())) == magforce.MF_ERR_INVALID_ARG
