"""magforce -- Python mirror of the magnetic hot path's host API.

Binds the C ABI of libmagforce.so (include/mf_capi.h) with ctypes.  Device
memory comes from torch tensors on a ROCm device (torch is plumbing here:
every product runs in the library's own HIP kernels).  The mirror keeps the
reference's vocabulary (magnetic_solver/magnetic_force.cuh):

  Plan.field(m)        -> G m            (cuda_magnet_cgOptimizer::G)
  Plan.force(m, text)  -> F              (evaluate_fitted_near_foece)
  Plan.solve(b, x)     -> x              (solve / start / step_iteration)
  field_allpairs(...)  -> H_ext + K m    (cuda_ext_field_kernel)

Particle arrays are torch float32 tensors of shape (3, n) (SoA rows x, y, z),
the layout of cuda_abstract_vector.  If the library is missing, every entry
point raises: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

from . import configs  # noqa: F401  (seeded workload generators)

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libmagforce.so")
EXAMPLE_PATH = os.path.join(PKG_ROOT, "lib", "example")
HEADER_PATH = os.path.join(os.path.dirname(PKG_ROOT), "include", "mf_capi.h")

MF_OK = 0
MF_ERR_INVALID_ARG = 1
MF_ERR_HIP = 2
MF_ERR_ALLOC = 3
MF_ERR_NOT_CONVERGED = 4
MF_ERR_UNSUPPORTED = 5
MF_ERR_NO_DEVICE = 6

MODE_ALLPAIRS = 0
MODE_FMM = 1
M2L_FP32 = 0
M2L_BF16X6 = 1
M2L_BF16X3 = 2
M2L_BF16X1 = 3
M2L_BF16X1_CUT = 4  # BF16X1 cut to the translation terms of low total degree (relaxed products)
SOLVE_CG = 0
SOLVE_GMRES = 1
SOLVE_REF_NLCG = 2


class MagforceError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class PlanOpts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("leaf_size", ctypes.c_int), ("order", ctypes.c_int),
                ("m2l_precision", ctypes.c_int), ("serial_near", ctypes.c_int),
                ("leaf_side", ctypes.c_float), ("reserved", ctypes.c_int * 2)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("mode", ctypes.c_int), ("order", ctypes.c_int),
                ("depth", ctypes.c_int), ("reserved0", ctypes.c_int), ("leaves", ctypes.c_int64),
                ("boxes", ctypes.c_int64), ("p2p_field_pairs", ctypes.c_int64),
                ("p2p_force_pairs", ctypes.c_int64), ("m2l_field_pairs", ctypes.c_int64),
                ("m2l_force_pairs", ctypes.c_int64), ("m2l_field_tile_slots", ctypes.c_int64),
                ("m2l_force_tile_slots", ctypes.c_int64),
                ("m2l_field_operators", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved0"}


class SolveOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("restart", ctypes.c_int), ("maxit", ctypes.c_int),
                ("rtol", ctypes.c_double), ("refine_fp64", ctypes.c_int),
                ("true_residual", ctypes.c_int), ("zero_start", ctypes.c_int),
                ("graph_iters", ctypes.c_int), ("inner_m2l_precision", ctypes.c_int),
                ("inner_rtol", ctypes.c_double), ("relax_m2l_precision", ctypes.c_int),
                ("reserved_solve", ctypes.c_int), ("relax_rtol", ctypes.c_double)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("matvecs", ctypes.c_int),
                ("converged", ctypes.c_int), ("refinements", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("rel_residual_true", ctypes.c_double),
                ("relaxed_matvecs", ctypes.c_int), ("reserved_stats", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_f = ctypes.c_float
_d = ctypes.c_double
_i = ctypes.c_int

_SIGS = {
    "mf_abi_version": ([], _i),
    "mf_status_string": ([_i], ctypes.c_char_p),
    "mf_last_error": ([], ctypes.c_char_p),
    "mf_ctx_create": ([_i, ctypes.POINTER(_vp)], _i),
    "mf_ctx_destroy": ([_vp], _i),
    "mf_ctx_last_error": ([_vp], ctypes.c_char_p),
    "mf_ctx_get_default": ([ctypes.POINTER(_vp)], _i),
    "mf_ctx_set_stream": ([_vp, _vp], _i),
    "mf_ctx_get_stream": ([_vp], _vp),
    "mf_ctx_synchronize": ([_vp], _i),
    "mf_ctx_enable_timing": ([_vp, _i], _i),
    "mf_ctx_kernel_time": ([_vp, ctypes.c_char_p, ctypes.POINTER(_d), ctypes.POINTER(_i64)], _i),
    "mf_ctx_reset_timing": ([_vp], _i),
    "mf_field_allpairs": ([_vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _f, _vp,
                           _vp, _vp, _vp], _i),
    "mf_field_allpairs_f64": ([_vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _d,
                               _vp, _vp, _vp, _vp], _i),
    "mf_hessian_allpairs": ([_vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _f,
                             _vp], _i),
    "mf_hessian_allpairs9": ([_vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _f,
                              _vp], _i),
    "mf_force_allpairs": ([_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _f, _vp, _vp, _vp, _vp], _i),
    "mf_force_allpairs_f64": ([_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp], _i),
    "mf_force_subset_f64": ([_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _d, _i64, _vp, _vp, _vp,
                             _vp], _i),
    "mf_plan_create": ([_vp, _i64, _vp, _vp, _vp, _f, ctypes.POINTER(PlanOpts),
                        ctypes.POINTER(_vp)], _i),
    "mf_plan_update_positions": ([_vp, _i64, _vp, _vp, _vp], _i),
    "mf_plan_destroy": ([_vp], _i),
    "mf_plan_size": ([_vp], _i64),
    "mf_plan_get_info": ([_vp, ctypes.POINTER(PlanInfo)], _i),
    "mf_plan_field": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "mf_plan_force": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "mf_plan_solve": ([_vp, _d, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(SolveOpts),
                       ctypes.POINTER(SolveStats)], _i),
    "mf_plan_solve_f64": ([_vp, _d, _vp, _vp, ctypes.POINTER(SolveOpts),
                           ctypes.POINTER(SolveStats)], _i),
    "mf_plan_create_batch": ([_vp, _i, ctypes.POINTER(_i64), _vp, _vp, _vp, _f,
                              ctypes.POINTER(PlanOpts), ctypes.POINTER(_vp)], _i),
    "mf_plan_solve_batch": ([_vp, ctypes.POINTER(_d), _vp, _vp, _vp, _vp, _vp, _vp,
                             ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats),
                             ctypes.POINTER(_d)], _i),
    "mf_vec_axpby": ([_vp, _i64, _f, _vp, _vp, _vp, _f, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "mf_vec_mul": ([_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "mf_vec_dot": ([_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_d)], _i),
    "mf_vec_fill": ([_vp, _i64, _f, _vp], _i),
    "mf_field_allpairs_noalloc": ([_vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                   _f, _vp, _vp, _vp, _vp], _i),
    "mf_noalloc_scratch_bytes": ([ctypes.POINTER(_i64)], _i),
    "mf_set_alloc_hook": ([_vp, _vp], _i),
}

# mf_alloc_hook: int (*)(uint64_t bytes, void *user)
ALLOC_HOOK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p)
_alloc_hook_ref = None


def lib():
    """Loads libmagforce.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MagforceError(MF_ERR_UNSUPPORTED,
                                f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def header_functions(path: str = HEADER_PATH):
    """Names of every function declared in include/mf_capi.h."""
    import re
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", text)))


def _check(rc: int, ctx=None):
    if rc != MF_OK:
        L = lib()
        msg = (L.mf_ctx_last_error(ctx) if ctx else L.mf_last_error()) or b""
        raise MagforceError(rc, msg.decode(errors="replace"))


def _ptr(t, row=None):
    if t is None:
        return None
    if row is None:
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t[row].data_ptr())


def _soa(t, n=None):
    """(3, n) contiguous float32 CUDA tensor view of a (3, n) or (n, 3) tensor."""
    import torch
    if t.dim() != 2:
        raise ValueError("expected a 2-D tensor")
    if t.shape[0] != 3 and t.shape[1] == 3:
        t = t.t()
    t = t.to(dtype=torch.float32).contiguous()
    if n is not None and t.shape[1] != n:
        raise ValueError(f"expected {n} particles, got {t.shape[1]}")
    return t


class Context:
    """An mf_ctx bound to one HIP device.  By default it adopts torch's current
    stream so torch events and the library's kernels share one stream."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        import torch
        self.device = device
        h = ctypes.c_void_p()
        torch.cuda.set_device(device)
        _check(lib().mf_ctx_create(device, ctypes.byref(h)))
        self.h = h
        if use_torch_stream:
            s = torch.cuda.current_stream(device).cuda_stream
            _check(lib().mf_ctx_set_stream(self.h, ctypes.c_void_p(s)), self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().mf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().mf_ctx_synchronize(self.h), self.h)

    def enable_timing(self, on: bool = True):
        _check(lib().mf_ctx_enable_timing(self.h, 1 if on else 0), self.h)

    def reset_timing(self):
        _check(lib().mf_ctx_reset_timing(self.h), self.h)

    def kernel_time(self, name: str):
        ms, cnt = ctypes.c_double(), ctypes.c_int64()
        _check(lib().mf_ctx_kernel_time(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)),
               self.h)
        return ms.value, cnt.value


def field_allpairs(ctx: Context, tpos, spos, smom, h: float, hconst=None):
    """H_i = hconst + sum_j K(t_i - s_j) m_j  (cuda_ext_field_kernel)."""
    import torch
    t = _soa(tpos)
    s = _soa(spos)
    m = _soa(smom, s.shape[1])
    out = torch.empty_like(t)
    hc = None
    if hconst is not None:
        hc = (ctypes.c_float * 3)(*[float(v) for v in hconst])
    _check(lib().mf_field_allpairs(ctx.h, t.shape[1], _ptr(t, 0), _ptr(t, 1), _ptr(t, 2),
                                   s.shape[1], _ptr(s, 0), _ptr(s, 1), _ptr(s, 2), _ptr(m, 0),
                                   _ptr(m, 1), _ptr(m, 2), float(h), hc, _ptr(out, 0),
                                   _ptr(out, 1), _ptr(out, 2)), ctx.h)
    return out


def field_allpairs_noalloc(ctx: Context, tpos, spos, smom, h: float, hconst=None, out=None):
    """field_allpairs through the allocation-free entry point: its scratch is a
    static device array of noalloc_scratch_bytes() per GPU.  `out` may be a
    preallocated (3, n_t) float32 tensor."""
    import torch
    t = _soa(tpos)
    s = _soa(spos)
    m = _soa(smom, s.shape[1])
    if out is None:
        out = torch.empty_like(t)
    hc = None
    if hconst is not None:
        hc = (ctypes.c_float * 3)(*[float(v) for v in hconst])
    _check(lib().mf_field_allpairs_noalloc(ctx.h, t.shape[1], _ptr(t, 0), _ptr(t, 1), _ptr(t, 2),
                                           s.shape[1], _ptr(s, 0), _ptr(s, 1), _ptr(s, 2),
                                           _ptr(m, 0), _ptr(m, 1), _ptr(m, 2), float(h), hc,
                                           _ptr(out, 0), _ptr(out, 1), _ptr(out, 2)), ctx.h)
    return out


def set_alloc_hook(fn=None):
    """Installs fn(bytes) -> bool as the library's allocation hook (True
    refuses that device allocation: MF_ERR_ALLOC); None removes it."""
    global _alloc_hook_ref
    if fn is None:
        _check(lib().mf_set_alloc_hook(None, None))
        _alloc_hook_ref = None
        return
    hook = ALLOC_HOOK(lambda nbytes, user: 1 if fn(int(nbytes)) else 0)
    _check(lib().mf_set_alloc_hook(ctypes.cast(hook, ctypes.c_void_p), None))
    _alloc_hook_ref = hook  # the C side keeps a pointer to it


def noalloc_scratch_bytes() -> int:
    v = ctypes.c_int64()
    _check(lib().mf_noalloc_scratch_bytes(ctypes.byref(v)))
    return v.value


def field_allpairs_f64(ctx: Context, tpos, spos, smom, h: float, hconst=None):
    """FP64 all-pairs field (the GPU FP64 check path)."""
    import torch

    def soa64(a):
        if a.shape[0] != 3 and a.shape[1] == 3:
            a = a.t()
        return a.to(dtype=torch.float64).contiguous()
    t, s, m = soa64(tpos), soa64(spos), soa64(smom)
    out = torch.empty_like(t)
    hc = None
    if hconst is not None:
        hc = (ctypes.c_double * 3)(*[float(v) for v in hconst])
    _check(lib().mf_field_allpairs_f64(ctx.h, t.shape[1], _ptr(t, 0), _ptr(t, 1), _ptr(t, 2),
                                       s.shape[1], _ptr(s, 0), _ptr(s, 1), _ptr(s, 2), _ptr(m, 0),
                                       _ptr(m, 1), _ptr(m, 2), float(h), hc, _ptr(out, 0),
                                       _ptr(out, 1), _ptr(out, 2)), ctx.h)
    return out


def hessian_allpairs(ctx: Context, tpos, spos, smom, h: float):
    """6-component force tensor (xx, yy, zz, xy, xz, yz) at the targets."""
    import torch
    t = _soa(tpos)
    s = _soa(spos)
    m = _soa(smom, s.shape[1])
    out = torch.empty((6, t.shape[1]), dtype=torch.float32, device=t.device)
    _check(lib().mf_hessian_allpairs(ctx.h, t.shape[1], _ptr(t, 0), _ptr(t, 1), _ptr(t, 2),
                                     s.shape[1], _ptr(s, 0), _ptr(s, 1), _ptr(s, 2), _ptr(m, 0),
                                     _ptr(m, 1), _ptr(m, 2), float(h), _ptr(out)), ctx.h)
    return out


def hessian_allpairs9(ctx: Context, tpos, spos, smom, h: float):
    """The force tensor as nine (3, 3)-ordered rows T_ab, a-major (the text9 /
    dHext layout), shape (9, n_t)."""
    import torch
    t = _soa(tpos)
    s = _soa(spos)
    m = _soa(smom, s.shape[1])
    out = torch.empty((9, t.shape[1]), dtype=torch.float32, device=t.device)
    ptrs = (ctypes.c_void_p * 9)(*[out[k].data_ptr() for k in range(9)])
    _check(lib().mf_hessian_allpairs9(ctx.h, t.shape[1], _ptr(t, 0), _ptr(t, 1), _ptr(t, 2),
                                      s.shape[1], _ptr(s, 0), _ptr(s, 1), _ptr(s, 2), _ptr(m, 0),
                                      _ptr(m, 1), _ptr(m, 2), float(h), ptrs), ctx.h)
    return out


def _text_ptrs(text):
    if text is None:
        return None
    if text.shape[0] != 9:
        raise ValueError("text must have shape (9, n)")
    return (ctypes.c_void_p * 9)(*[text[k].data_ptr() for k in range(9)])


def force_allpairs(ctx: Context, pos, mom, h: float, text=None):
    """F_i = (sum_j B(x_i - x_j)[m_j] + Text_i) m_i  (fitted_force_ofM_kernel)."""
    import torch
    p = _soa(pos)
    m = _soa(mom, p.shape[1])
    out = torch.empty_like(p)
    tp = None
    if text is not None:
        text = text.to(dtype=torch.float32).contiguous()
        tp = _text_ptrs(text)
    _check(lib().mf_force_allpairs(ctx.h, p.shape[1], _ptr(p, 0), _ptr(p, 1), _ptr(p, 2),
                                   _ptr(m, 0), _ptr(m, 1), _ptr(m, 2), float(h), tp, _ptr(out, 0),
                                   _ptr(out, 1), _ptr(out, 2)), ctx.h)
    return out


def force_allpairs_f64(ctx: Context, pos, mom, h: float):
    import torch

    def soa64(a):
        if a.shape[0] != 3 and a.shape[1] == 3:
            a = a.t()
        return a.to(dtype=torch.float64).contiguous()
    p, m = soa64(pos), soa64(mom)
    out = torch.empty_like(p)
    _check(lib().mf_force_allpairs_f64(ctx.h, p.shape[1], _ptr(p, 0), _ptr(p, 1), _ptr(p, 2),
                                       _ptr(m, 0), _ptr(m, 1), _ptr(m, 2), float(h), _ptr(out, 0),
                                       _ptr(out, 1), _ptr(out, 2)), ctx.h)
    return out


def force_subset_f64(ctx: Context, pos, mom, h: float, idx):
    """FP64 force at particles idx (int64 CUDA tensor), all particles as sources."""
    import torch

    def soa64(a):
        if a.shape[0] != 3 and a.shape[1] == 3:
            a = a.t()
        return a.to(dtype=torch.float64).contiguous()
    p, m = soa64(pos), soa64(mom)
    idx = idx.to(dtype=torch.int64).contiguous()
    out = torch.empty((3, idx.numel()), dtype=torch.float64, device=p.device)
    _check(lib().mf_force_subset_f64(ctx.h, p.shape[1], _ptr(p, 0), _ptr(p, 1), _ptr(p, 2),
                                     _ptr(m, 0), _ptr(m, 1), _ptr(m, 2), float(h), idx.numel(),
                                     _ptr(idx), _ptr(out, 0), _ptr(out, 1), _ptr(out, 2)), ctx.h)
    return out


@dataclass
class SolveResult:
    x: object
    stats: dict


class Plan:
    """A particle set with its all-pairs records (mode=MODE_ALLPAIRS) or
    octree (mode=MODE_FMM)."""

    def __init__(self, ctx: Context, pos, h: float, mode: int = MODE_ALLPAIRS,
                 leaf_size: int = 0, order: int = 0, m2l_precision: int = 0,
                 serial_near: bool = False, leaf_side: float = 0.0):
        self.ctx = ctx
        self.h = float(h)
        p = _soa(pos)
        self.n = p.shape[1]
        self.opts = PlanOpts(mode=mode, leaf_size=leaf_size, order=order,
                             m2l_precision=m2l_precision, serial_near=int(bool(serial_near)),
                             leaf_side=float(leaf_side))
        hp = ctypes.c_void_p()
        _check(lib().mf_plan_create(ctx.h, self.n, _ptr(p, 0), _ptr(p, 1), _ptr(p, 2), self.h,
                                    ctypes.byref(self.opts), ctypes.byref(hp)), ctx.h)
        self.h_plan = hp
        self.pos = p

    def close(self):
        if getattr(self, "h_plan", None):
            lib().mf_plan_destroy(self.h_plan)
            self.h_plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        pi = PlanInfo()
        _check(lib().mf_plan_get_info(self.h_plan, ctypes.byref(pi)), self.ctx.h)
        return pi.as_dict()

    def update_positions(self, pos):
        p = _soa(pos)
        _check(lib().mf_plan_update_positions(self.h_plan, p.shape[1], _ptr(p, 0), _ptr(p, 1),
                                              _ptr(p, 2)), self.ctx.h)
        self.n = p.shape[1]
        self.pos = p

    def field(self, mom, hconst=None, out=None):
        import torch
        m = _soa(mom, self.n)
        if out is None:
            out = torch.empty_like(m)
        hc = None if hconst is None else (ctypes.c_float * 3)(*[float(v) for v in hconst])
        _check(lib().mf_plan_field(self.h_plan, _ptr(m, 0), _ptr(m, 1), _ptr(m, 2), hc,
                                   _ptr(out, 0), _ptr(out, 1), _ptr(out, 2)), self.ctx.h)
        return out

    def force(self, mom, text=None, out=None):
        import torch
        m = _soa(mom, self.n)
        if out is None:
            out = torch.empty_like(m)
        tp = None
        if text is not None:
            text = text.to(dtype=torch.float32).contiguous()
            tp = _text_ptrs(text)
        _check(lib().mf_plan_force(self.h_plan, _ptr(m, 0), _ptr(m, 1), _ptr(m, 2), tp,
                                   _ptr(out, 0), _ptr(out, 1), _ptr(out, 2)), self.ctx.h)
        return out

    def solve(self, kappa: float, b, x0=None, method: int = SOLVE_CG, rtol: float = 1e-6,
              maxit: int = 1000, restart: int = 30, true_residual: bool = False,
              relax_m2l_precision=None, relax_rtol: float = 0.0):
        """relax_m2l_precision (CG on FMM plans): the M2L precision of the
        products made once the relative residual is at or below relax_rtol
        (0 = 1e-3); None = every product at the plan's own precision."""
        import torch
        bb = _soa(b, self.n)
        # x0 None: the solver starts from zero itself, saving one product
        x = torch.empty_like(bb) if x0 is None else _soa(x0, self.n).clone()
        opts = SolveOpts(method=method, restart=restart, maxit=maxit, rtol=rtol,
                         true_residual=1 if true_residual else 0,
                         zero_start=1 if x0 is None else 0,
                         relax_m2l_precision=0 if relax_m2l_precision is None
                         else int(relax_m2l_precision) + 1,
                         relax_rtol=float(relax_rtol))
        st = SolveStats()
        _check(lib().mf_plan_solve(self.h_plan, float(kappa), _ptr(bb, 0), _ptr(bb, 1),
                                   _ptr(bb, 2), _ptr(x, 0), _ptr(x, 1), _ptr(x, 2),
                                   ctypes.byref(opts), ctypes.byref(st)), self.ctx.h)
        return SolveResult(x, st.as_dict())

    def solve_f64(self, kappa: float, b, x0=None, rtol: float = 1e-12, max_refine: int = 10,
                  maxit: int = 500, inner_m2l_precision=None, inner_rtol: float = 0.0):
        """Iterative refinement with x and the residual in FP64.  All-pairs
        plans: FP64 all-pairs residual (the FP64 solution).  FMM plans: the
        residual from the plan's FMM product, inner corrections at
        inner_m2l_precision (None = the plan's own)."""
        import torch
        if b.shape[0] != 3 and b.shape[1] == 3:
            b = b.t()
        bb = b.to(dtype=torch.float64).contiguous()
        x = torch.zeros_like(bb) if x0 is None else x0.to(torch.float64).contiguous().clone()
        opts = SolveOpts(method=SOLVE_CG, maxit=maxit, rtol=rtol, refine_fp64=max_refine,
                         inner_m2l_precision=0 if inner_m2l_precision is None
                         else int(inner_m2l_precision) + 1,
                         inner_rtol=float(inner_rtol))
        st = SolveStats()
        _check(lib().mf_plan_solve_f64(self.h_plan, float(kappa), _ptr(bb), _ptr(x),
                                       ctypes.byref(opts), ctypes.byref(st)), self.ctx.h)
        return SolveResult(x, st.as_dict())


class BatchPlan(Plan):
    """Independent particle systems in one FMM plan (mf_plan_create_batch):
    system s owns columns offsets[s]:offsets[s+1] of pos.  field/force are
    block-diagonal; solve() runs the batched CG with one kappa per system."""

    def __init__(self, ctx: Context, pos, offsets, h: float, order: int = 0,
                 m2l_precision: int = 0, serial_near: bool = False, leaf_side: float = 0.0):
        self.ctx = ctx
        self.h = float(h)
        p = _soa(pos)
        self.n = p.shape[1]
        off = [int(v) for v in offsets]
        if off[0] != 0 or off[-1] != self.n:
            raise ValueError("offsets must start at 0 and end at n")
        self.offsets = off
        self.nsys = len(off) - 1
        self.opts = PlanOpts(mode=MODE_FMM, order=order, m2l_precision=m2l_precision,
                             serial_near=int(bool(serial_near)), leaf_side=float(leaf_side))
        arr = (ctypes.c_int64 * len(off))(*off)
        hp = ctypes.c_void_p()
        _check(lib().mf_plan_create_batch(ctx.h, self.nsys, arr, _ptr(p, 0), _ptr(p, 1),
                                          _ptr(p, 2), ctypes.c_float(self.h),
                                          ctypes.byref(self.opts), ctypes.byref(hp)), ctx.h)
        self.h_plan = hp
        self.pos = p

    def solve(self, kappas, b, x0=None, rtol: float = 1e-6, maxit: int = 1000,
              graph_iters: int = 0, relax_m2l_precision=None, relax_rtol: float = 0.0):
        """relax_m2l_precision / relax_rtol as in Plan.solve, the residual being
        the largest relative residual over the systems."""
        import torch
        bb = _soa(b, self.n)
        if len(kappas) != self.nsys:
            raise ValueError("one kappa per system")
        x = torch.empty_like(bb) if x0 is None else _soa(x0, self.n).clone()
        opts = SolveOpts(method=SOLVE_CG, maxit=maxit, rtol=rtol,
                         zero_start=1 if x0 is None else 0, graph_iters=graph_iters,
                         relax_m2l_precision=0 if relax_m2l_precision is None
                         else int(relax_m2l_precision) + 1,
                         relax_rtol=float(relax_rtol))
        st = SolveStats()
        kap = (ctypes.c_double * self.nsys)(*[float(k) for k in kappas])
        res = (ctypes.c_double * self.nsys)()
        _check(lib().mf_plan_solve_batch(self.h_plan, kap, _ptr(bb, 0), _ptr(bb, 1), _ptr(bb, 2),
                                         _ptr(x, 0), _ptr(x, 1), _ptr(x, 2), ctypes.byref(opts),
                                         ctypes.byref(st), res), self.ctx.h)
        d = st.as_dict()
        d["sys_rel_residual"] = list(res)
        return SolveResult(x, d)


def uniform_field(n: int, hconst, device="cuda"):
    """(3, n) float32 tensor holding the constant field in every particle."""
    import torch
    h = torch.tensor([float(v) for v in hconst], dtype=torch.float32, device=device)
    return h.view(3, 1).expand(3, n).contiguous()
