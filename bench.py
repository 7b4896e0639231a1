"""Benchmark of the magnetic hot path on one MI355X.

One step = plan (re)build + magnetization solve of (I - kappa G) x = H_ext to
1e-6 relative residual (CG) + force F = (B(m) + Text) m, with the particle
positions already resident in HBM.  The default workload is the 1,153,433
particle pool (256x256x16 jittered lattice + 104,857 random particles above
it, h = 2, chi = 1, H = (0,0,1)).

Printed JSON line (rank 0): metric/value/unit per the contract, plus
  "error"       : ||F - F*|| / ||F*|| on a seeded subsample, F* the FP64
                  all-pairs force at the FP64 solution (solved to 1e-12),
  "allpairs"    : Gpair-interactions/s of the FP32 all-pairs field kernel on
                  the 262,144-particle cylinder,
  "roofline"    : dominant kernel, algorithmic FLOPs per launch / measured
                  launch time against the FP32 peak,
  "timestep"    : the moving-particle step (plan rebuilt in place, solve
                  warm-started from the last solution): ms/step, rebuild
                  time and CG iterations warm against cold,
  "refined_fmm" : the step with FP64-residual refinement on the FMM plan
                  (inner corrections with one-part BF16 translations),
  "cpu_baseline": the reference's CPU FMM path restated in FP64 (oracle/), one
                  field product of the workload timed on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "magnetic-force_amd"))

FP32_PEAK_TFLOPS = 157.3        # MI355X published FP32 vector / FP32 MFMA peak
BF16_PEAK_TFLOPS = 2500.0       # MI355X published dense BF16 MFMA peak
HBM_PEAK_GBS = 8000.0           # MI355X published HBM3E bandwidth
FIELD_FLOPS_PER_PAIR = 32       # far-branch field pair, FMA = 2, rsqrt = 1
FORCE_FLOPS_PER_PAIR = 64       # far-branch 6-component force tensor pair


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def to_dev(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(np.asarray(a).T), dtype=dtype or torch.float32,
                        device="cuda").contiguous()


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="pool1m")
    p.add_argument("--mode", default="auto", choices=["auto", "fmm", "allpairs"])
    p.add_argument("--rtol", type=float, default=1e-6)
    p.add_argument("--m2l", default="bf16x3", choices=["fp32", "bf16x6", "bf16x3"],
                   help="FMM M2L precision (bf16x3: the mixed-precision FMM)")
    p.add_argument("--relax", default="bf16x1cut", choices=["none", "bf16x1", "bf16x1cut", "bf16x3"],
                   help="M2L precision of the CG products made once the relative residual "
                        "is at or below --relax-rtol (inexact Krylov); none = every product "
                        "at --m2l")
    p.add_argument("--relax-rtol", type=float, default=5e-3,
                   help="relative residual below which products are relaxed; 5e-3 keeps "
                        "the pool's force error at 2.2e-6 (1e-2 relaxes one more product: "
                        "about 1.7 ms faster, 2.5e-6)")
    p.add_argument("--batch-graph-iters", type=int, default=0,
                   help="batch leg: iterations per captured-graph replay between host "
                        "convergence checks (0 = plain launches, the faster here)")
    p.add_argument("--no-batch", action="store_true",
                   help="skip the 96-droplet batch leg")
    p.add_argument("--no-compare", action="store_true",
                   help="skip the FP32-M2L FMM run of the same workload")
    p.add_argument("--no-check", action="store_true")
    p.add_argument("--check-samples", type=int, default=0,
                   help="subsample size of the error check (default 4096; 65536 for n > 2M)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-allpairs", action="store_true")
    p.add_argument("--no-timestep", action="store_true",
                   help="skip the moving-particle time-step leg (warm-started solves)")
    p.add_argument("--profile-only", action="store_true",
                   help="only the timed steps (for rocprofv3 runs)")
    p.add_argument("--copy-calib-mb", type=int, default=0,
                   help="run one device-to-device copy of this many MiB first: a kernel of "
                        "known traffic that scales the FETCH_SIZE/WRITE_SIZE counters")
    p.add_argument("--traffic", default=None,
                   help="JSON with measured HBM bytes per launch of the dominant kernel "
                        "(from the rocprofv3 counter runs under profiles/)")
    return p.parse_args()


def m2l_issued_per_tile_slot(order, m2l):
    """MFMA FLOPs the field M2L issues per tile-slot (64 target boxes, one
    offset).  FP32 and BF16x6: the whole KP x KP operator, 1 or 6 products.
    BF16x3 (degree-split kernel): per 16 x 32 block of the operator, 3
    products where the lowest local and multipole degrees in the block add up
    to at most the order minus 3, 1 product elsewhere, none on pure padding."""
    nd = (order + 1) ** 2
    kp = (nd + 31) // 32 * 32
    if m2l != "bf16x3":
        return 64.0 * kp * kp * 2 * {"fp32": 1, "bf16x6": 6}[m2l]
    prods = 0
    for rb in range(kp // 16):
        for ks in range(kp // 32):
            if 16 * rb >= nd or 32 * ks >= nd:
                continue
            prods += 3 if math.isqrt(16 * rb) + math.isqrt(32 * ks) <= order - 3 else 1
    return float(prods) * 4 * (16 * 16 * 32 * 2)  # 4 column blocks of 16 targets


def relax_kwargs(args):
    """Solve options of the relaxed products (none: every product at --m2l)."""
    import magforce
    if args.relax == "none":
        return {}
    code = {"bf16x1": magforce.M2L_BF16X1, "bf16x1cut": magforce.M2L_BF16X1_CUT,
            "bf16x3": magforce.M2L_BF16X3}[args.relax]
    return {"relax_m2l_precision": code, "relax_rtol": args.relax_rtol}


def relax_info(args):
    return None if args.relax == "none" else {"m2l": args.relax, "below_rel_residual": args.relax_rtol}


class Stepper:
    """Runs plan build + solve + force for one particle system."""

    def __init__(self, ctx, system, mode, rtol, m2l=0, serial_near=False, relax=None):
        import torch
        import magforce
        self.ctx = ctx
        self.relax = relax or {}  # relax_m2l_precision / relax_rtol of the solve
        self.s = system
        self.mode = mode
        self.rtol = rtol
        self.m2l = m2l
        self.serial_near = serial_near
        self.pos = to_dev(system.pos)
        self.b = magforce.uniform_field(system.n, system.hconst)
        self.x = torch.zeros_like(self.b)
        self.F = torch.empty_like(self.b)
        self.plan = None
        self.stats = None

    def step(self):
        import magforce
        if self.plan is None:
            self.plan = magforce.Plan(self.ctx, self.pos, self.s.h, mode=self.mode,
                                      m2l_precision=self.m2l, serial_near=self.serial_near)
        else:
            self.plan.update_positions(self.pos)
        res = self.plan.solve(self.s.kappa, self.b, x0=None, rtol=self.rtol, **self.relax)
        self.x = res.x
        self.stats = res.stats
        self.plan.force(self.s.kappa * self.x, out=self.F)


def measure_allpairs(ctx, log_prefix="allpairs"):
    """Gpairs/s of the FP32 all-pairs field kernel on the 262,144 cylinder."""
    import torch
    import magforce
    from magforce import configs
    s = configs.cylinder262k()
    pos = to_dev(s.pos)
    rng = np.random.default_rng(0)
    m = to_dev(rng.normal(size=(s.n, 3)) * 0.5)
    plan = magforce.Plan(ctx, pos, s.h)
    out = torch.empty_like(m)
    plan.field(m, out=out)
    torch.cuda.synchronize()
    ctx.reset_timing()
    ctx.enable_timing(True)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.field(m, out=out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms, cnt = ctx.kernel_time("field")
    ctx.enable_timing(False)
    pairs = float(s.n) * s.n
    avg_ms = ms / max(cnt, 1)
    res = {"workload": "cylinder262k", "n": s.n, "gpairs_per_s": pairs / (avg_ms * 1e-3) / 1e9,
           "kernel_ms": avg_ms, "wall_ms_per_matvec": wall / reps * 1e3,
           "tflops": pairs * FIELD_FLOPS_PER_PAIR / (avg_ms * 1e-3) / 1e12}
    res["frac_fp32_peak"] = res["tflops"] / FP32_PEAK_TFLOPS
    # the cylinder's own step: GMRES(30) run for 30 all-pairs products, then the
    # force; scored against the FP64 solution (all-pairs FP64 refinement to
    # 1e-12) and FP64 forces on 4096 seeded samples
    b = magforce.uniform_field(s.n, s.hconst)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    sol = plan.solve(s.kappa, b, method=magforce.SOLVE_GMRES, restart=30, maxit=30, rtol=1e-30)
    F = plan.force(s.kappa * sol.x)
    e1.record()
    torch.cuda.synchronize()
    step = {"ms": e0.elapsed_time(e1), "solver": sol.stats}
    ref = plan.solve_f64(s.kappa, b.double(), rtol=1e-12, max_refine=6)
    rng = np.random.default_rng(12345)
    idx = torch.tensor(np.sort(rng.choice(s.n, size=4096, replace=False)), dtype=torch.int64,
                       device="cuda")
    F_ref = magforce.force_subset_f64(ctx, pos.double(), s.kappa * ref.x, s.h, idx)
    step["force_rel_l2"] = float(torch.linalg.norm(F[:, idx].double() - F_ref) /
                                 torch.linalg.norm(F_ref))
    step["x_rel_l2"] = float(torch.linalg.norm(sol.x.double() - ref.x) / torch.linalg.norm(ref.x))
    res["gmres30_step"] = step
    log(f"{log_prefix}: {res}")
    plan.close()
    return res


def time_fp32_fmm(ctx, system, args):
    """The same step with FP32 M2L (the plain FP32 FMM build), same steps."""
    import torch
    import magforce
    st = Stepper(ctx, system, magforce.MODE_FMM, args.rtol, magforce.M2L_FP32)
    for _ in range(max(args.warmup, 1)):
        st.step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        st.step()
    e1.record()
    torch.cuda.synchronize()
    out = {"ms_per_step": e0.elapsed_time(e1) / args.steps, "m2l": "fp32",
           "solver": st.stats}
    st.plan.close()
    log(f"fp32 fmm: {out}")
    return out


def time_refined(ctx, system, args, m2l_code, samples=4096):
    """The same step with the magnetization solved by iterative refinement on
    the FMM plan (mf_plan_solve_f64): x and r = b - A x in FP64, the residual
    from the plan's own FMM product (M2L at the step's precision), the inner
    corrections by CG with one-part BF16 translations to 1e-3.  Reported beside
    the plain mixed-precision CG step, with its own error against FP64."""
    import torch
    import magforce
    s = system
    pos = to_dev(s.pos)
    b64 = magforce.uniform_field(s.n, s.hconst).double()
    F = torch.empty((3, s.n), dtype=torch.float32, device="cuda")
    state = {}

    def step():
        if "plan" not in state:
            state["plan"] = magforce.Plan(ctx, pos, s.h, mode=magforce.MODE_FMM,
                                          m2l_precision=m2l_code)
        else:
            state["plan"].update_positions(pos)
        res = state["plan"].solve_f64(s.kappa, b64, rtol=args.rtol, max_refine=8,
                                      inner_m2l_precision=magforce.M2L_BF16X1, inner_rtol=1e-3)
        state["res"] = res
        state["plan"].force((s.kappa * res.x).float(), out=F)

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    out = {"ms_per_step": e0.elapsed_time(e1) / args.steps, "outer_m2l": args.m2l,
           "inner_m2l": "bf16x1", "inner_rtol": 1e-3, "solver": state["res"].stats}
    if not args.no_check:
        ap = magforce.Plan(ctx, pos, s.h)
        ref = ap.solve_f64(s.kappa, b64, rtol=1e-12, max_refine=6)
        rng = np.random.default_rng(12345)
        idx = torch.tensor(np.sort(rng.choice(s.n, size=min(samples, s.n), replace=False)),
                           dtype=torch.int64, device="cuda")
        F_ref = magforce.force_subset_f64(ctx, pos.double(), s.kappa * ref.x, s.h, idx)
        out["force_rel_l2"] = float(torch.linalg.norm(F[:, idx].double() - F_ref) /
                                    torch.linalg.norm(F_ref))
        out["x_rel_l2"] = float(torch.linalg.norm(state["res"].x - ref.x) / torch.linalg.norm(ref.x))
        ap.close()
    state["plan"].close()
    log(f"refined fmm: {out}")
    return out


def check_error(ctx, st: Stepper, samples: int):
    """||F - F*|| / ||F*|| over a seeded subsample; F* from the FP64 path."""
    import torch
    import magforce
    s = st.s
    t0 = time.perf_counter()
    ap = magforce.Plan(ctx, st.pos, s.h, mode=magforce.MODE_ALLPAIRS)
    b64 = st.b.double()
    ref = ap.solve_f64(s.kappa, b64, rtol=1e-12, max_refine=6)
    x64 = ref.x
    rng = np.random.default_rng(12345)
    k = min(samples, s.n)
    idx = torch.tensor(np.sort(rng.choice(s.n, size=k, replace=False)), dtype=torch.int64,
                       device="cuda")
    pos64 = st.pos.double()
    F_ref = magforce.force_subset_f64(ctx, pos64, s.kappa * x64, s.h, idx)
    F = st.F[:, idx].double()
    x_err = float(torch.linalg.norm(st.x.double() - x64) / torch.linalg.norm(x64))
    err = float(torch.linalg.norm(F - F_ref) / torch.linalg.norm(F_ref))
    rms = torch.sqrt((F_ref ** 2).sum(0).mean())
    err_max = float((F - F_ref).norm(dim=0).max() / rms)
    ap.close()
    out = {"force_rel_l2": err, "force_max_over_rms": err_max, "x_rel_l2": x_err,
           "samples": k, "ref_rel_residual": ref.stats["rel_residual_true"],
           "ref_seconds": time.perf_counter() - t0}
    log(f"check: {out}")
    return out


def check_error_subsample(ctx, st: Stepper, samples: int):
    """Error check for n too large for an FP64 all-pairs solve, on a seeded
    subsample S of the particles:
      force_op_rel_l2  ||F - F64|| / ||F64|| over S, F64 the FP64 all-pairs
                       force (all n sources) at the GPU's own moments: the
                       error of the force product;
      residual_rel_l2  ||r|| / ||b|| over S with r = b - x + kappa G x in FP64
                       (all n sources): the solve's true residual.  With
                       cond(I - kappa G) <= 1 + chi it bounds the solution
                       error, and the force error at F* follows from both."""
    import torch
    import magforce
    s = st.s
    t0 = time.perf_counter()
    rng = np.random.default_rng(12345)
    k = min(samples, s.n)
    idx = torch.tensor(np.sort(rng.choice(s.n, size=k, replace=False)), dtype=torch.int64,
                       device="cuda")
    pos64 = st.pos.double()
    m64 = s.kappa * st.x.double()
    F_ref = magforce.force_subset_f64(ctx, pos64, m64, s.h, idx)
    F = st.F[:, idx].double()
    op_err = float(torch.linalg.norm(F - F_ref) / torch.linalg.norm(F_ref))
    rms = torch.sqrt((F_ref ** 2).sum(0).mean())
    op_max = float((F - F_ref).norm(dim=0).max() / rms)
    g = magforce.field_allpairs_f64(ctx, pos64[:, idx].contiguous(), pos64, m64, s.h)
    bs = st.b[:, idx].double()
    r = bs - st.x[:, idx].double() + g
    res = float(torch.linalg.norm(r) / torch.linalg.norm(bs))
    out = {"method": "subsample", "force_op_rel_l2": op_err, "force_op_max_over_rms": op_max,
           "residual_rel_l2": res, "x_err_bound": (1.0 + s.chi) * res, "samples": k,
           "ref_seconds": time.perf_counter() - t0}
    log(f"check: {out}")
    return out


def run_timestep(ctx, st: Stepper, steps: int, sigma_h: float = 0.02, seed: int = 31):
    """The simulation caller's time step (magnetic_force.cu:806-888 with the
    unknown kept across Resize, :706-744): particles move a little, the plan
    is rebuilt in place for the new positions, and the solve starts from the
    previous step's solution.  Each step moves every particle by a normal
    displacement of sigma_h * h per axis.  The displaced position sets are
    made before the timed region; the timed step is update_positions + warm
    CG to the same rtol + force.  After the loop the last positions are solved
    again from zero, so the warm and cold solutions and iteration counts
    compare on the same system."""
    import torch
    s = st.s
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    poses, p = [], st.pos.clone()
    for _ in range(steps + 1):
        p = p + torch.randn(p.shape, generator=gen, device="cuda") * (sigma_h * s.h)
        poses.append(p)
    x = st.x.clone()
    kappa = s.kappa

    def step(pos):
        nonlocal x
        st.plan.update_positions(pos)
        res = st.plan.solve(kappa, st.b, x0=x, rtol=st.rtol, **st.relax)
        x = res.x
        st.plan.force(kappa * x, out=st.F)
        return res.stats

    step(poses[0])      # warm-up at the first displaced positions
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * steps + 2)]
    iters = []
    ev[0].record()
    for k in range(steps):
        ev[2 * k + 1].record()
        st.plan.update_positions(poses[k + 1])
        ev[2 * k + 2].record()
        res = st.plan.solve(kappa, st.b, x0=x, rtol=st.rtol, **st.relax)
        x = res.x
        st.plan.force(kappa * x, out=st.F)
        iters.append(res.stats)
    ev[-1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[-1]) / steps
    rebuild = [ev[2 * k + 1].elapsed_time(ev[2 * k + 2]) for k in range(steps)]
    cold = st.plan.solve(kappa, st.b, x0=None, rtol=st.rtol, **st.relax)
    torch.cuda.synchronize()
    dx = float(torch.linalg.norm(x - cold.x) / torch.linalg.norm(cold.x))
    out = {"motion": f"N(0, ({sigma_h} h)^2) per axis per step", "steps": steps,
           "ms_per_step": ms, "rebuild_ms": sum(rebuild) / steps,
           "iterations_warm": [int(r["iterations"]) for r in iters],
           "matvecs_warm": [int(r["matvecs"]) for r in iters],
           "iterations_cold": int(cold.stats["iterations"]),
           "matvecs_cold": int(cold.stats["matvecs"]),
           "warm_vs_cold_x_rel_l2": dx}
    st.pos, st.x = poses[-1], x
    log(f"timestep: {out}")
    return out


def run_droplets(ctx, args, m2l_code):
    """The 96 uneven droplets (2,048..262,144 particles each, chi in [0.5, 3],
    random unit H directions) as one batch plan: one step = batch plan
    (re)build + batched CG to rtol per system (one CG iteration captured in a
    hipGraph and replayed) + block-diagonal force.  The expansion order comes
    from the autotuner run on the largest droplet.  Error: per droplet, up to
    128 seeded particles, FP64 force at the GPU moments (that droplet's
    sources) and the FP64 residual b - x + kappa G x."""
    import torch
    import magforce
    from magforce import configs
    systems = configs.droplets96()
    off = [0]
    for sy in systems:
        off.append(off[-1] + sy.n)
    pos = to_dev(np.concatenate([sy.pos for sy in systems]))
    b = to_dev(np.concatenate([np.tile(np.asarray(sy.hconst, np.float64), (sy.n, 1))
                               for sy in systems]))
    kappas = [sy.kappa for sy in systems]
    kpart = torch.tensor(np.concatenate([np.full(sy.n, sy.kappa, np.float32) for sy in systems]),
                         device="cuda")
    state = {"plan": None, "x": None, "F": torch.empty_like(b), "stats": None}
    # order and M2L precision from the autotuner, run on the largest droplet
    # (target 1e-5 relative field error against FP64); --m2l fixes the precision
    from magforce import tune
    big = max(systems, key=lambda sy: sy.n)
    best, cands = tune.autotune_fmm(ctx, to_dev(big.pos), big.h, target_rel_err=1e-5,
                                    precisions=(m2l_code,), leaf_sides=(2.0, 2.5, 3.0))
    order = best.order if best else 10
    leaf_side = best.leaf_side if best else 2.0
    tuned = {"on": big.name, "n": big.n, "chosen": best.as_dict() if best else None,
             "candidates": [c.as_dict() for c in cands]}
    log(f"droplets96 autotune: {tuned}")

    def step():
        if state["plan"] is None:
            state["plan"] = magforce.BatchPlan(ctx, pos, off, systems[0].h, order=order,
                                               m2l_precision=m2l_code, leaf_side=leaf_side)
        else:
            state["plan"].update_positions(pos)
        res = state["plan"].solve(kappas, b, rtol=args.rtol, graph_iters=args.batch_graph_iters,
                                  **relax_kwargs(args))
        state["x"], state["stats"] = res.x, res.stats
        state["plan"].force(res.x * kpart, out=state["F"])

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    st = state["stats"]
    out = {"workload": "droplets96", "systems": len(systems), "n": off[-1],
           "n_min": min(sy.n for sy in systems), "n_max": max(sy.n for sy in systems),
           "ms_per_step": ms, "m2l_precision": args.m2l, "order": order,
           "leaf_side_h": leaf_side, "autotune": tuned,
           "graph_iters": args.batch_graph_iters, "relax": relax_info(args),
           "solver": {"iterations": st["iterations"], "matvecs": st["matvecs"],
                      "relaxed_matvecs": st["relaxed_matvecs"], "converged": st["converged"],
                      "max_sys_rel_residual": max(st["sys_rel_residual"])},
           "plan": state["plan"].info()}
    if not args.no_check:
        t0 = time.perf_counter()
        rng = np.random.default_rng(4321)
        dF = nF = dr = nb = 0.0
        pos64 = pos.double()
        m64 = state["x"].double() * kpart.double()
        for sy, a, e in zip(systems, off[:-1], off[1:]):
            k = min(128, sy.n)
            loc = torch.tensor(np.sort(rng.choice(sy.n, size=k, replace=False)), device="cuda")
            p64, mm = pos64[:, a:e].contiguous(), m64[:, a:e].contiguous()
            F_ref = magforce.force_subset_f64(ctx, p64, mm, sy.h, loc)
            F = state["F"][:, a:e][:, loc].double()
            dF += float(((F - F_ref) ** 2).sum())
            nF += float((F_ref ** 2).sum())
            g = magforce.field_allpairs_f64(ctx, p64[:, loc].contiguous(), p64, mm, sy.h)
            bs = b[:, a:e][:, loc].double()
            r = bs - state["x"][:, a:e][:, loc].double() + g
            dr += float((r ** 2).sum())
            nb += float((bs ** 2).sum())
        out["check"] = {"method": "subsample per droplet (<=128 each)",
                        "force_op_rel_l2": (dF / nF) ** 0.5, "residual_rel_l2": (dr / nb) ** 0.5,
                        "ref_seconds": time.perf_counter() - t0}
    state["plan"].close()
    log(f"droplets96: {out}")
    return out


def run_c1(ctx):
    """The reference's CPU-sized case (8,192 jittered lattice, h = 2, chi = 1):
    GPU all-pairs solve (CG to 1e-6) + force, next to the CPU oracle doing the
    same in FP64 (CG to 1e-6) and the FP32 restatement of the reference's
    kernels timed for one field and one force product."""
    import torch
    import magforce
    from magforce import configs
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    s = configs.lattice8k()
    threads = min(16, os.cpu_count() or 1)
    pos = to_dev(s.pos)
    b = magforce.uniform_field(s.n, s.hconst)
    plan = magforce.Plan(ctx, pos, s.h)
    plan.solve(s.kappa, b, rtol=1e-6)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    sol = plan.solve(s.kappa, b, rtol=1e-6)
    F = plan.force(s.kappa * sol.x)
    e1.record()
    torch.cuda.synchronize()
    gpu_ms = e0.elapsed_time(e1)
    p64 = s.pos.astype(np.float64)
    bh = np.tile(np.asarray(s.hconst, np.float64), (s.n, 1))
    t0 = time.perf_counter()
    x64, it, _ = oracle.cg_solve(p64, s.h, s.kappa, bh, rtol=1e-6, nthreads=threads)
    F64 = oracle.force_allpairs(p64, s.kappa * x64, s.h, nthreads=threads)
    cpu64 = time.perf_counter() - t0
    m32 = (s.kappa * x64).astype(np.float32)
    t0 = time.perf_counter()
    oracle.field_allpairs_f32ref(s.pos, m32, s.h, nthreads=threads)
    t_f32_field = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.force_allpairs_f32ref(s.pos, m32, s.h, nthreads=threads)
    t_f32_force = time.perf_counter() - t0
    err = float(np.linalg.norm(F.double().cpu().numpy().T - F64) / np.linalg.norm(F64))
    out = {"workload": "lattice8k", "n": s.n, "gpu_allpairs_ms": gpu_ms,
           "gpu_solver": sol.stats, "gpu_force_rel_l2_vs_cpu_fp64": err,
           "cpu_fp64_solve_force_s": cpu64, "cpu_fp64_cg_iterations": it,
           "cpu_f32_restatement_field_s": t_f32_field,
           "cpu_f32_restatement_force_s": t_f32_force, "cpu_threads": threads}
    plan.close()
    log(f"c1: {out}")
    return out


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(system, budget_s, matvecs_per_step):
    """The reference's CPU path for the step, restated in FP64 (oracle/mf_cpufmm.c):
    its fmmtl FMM field product (ncrit-split octree with ncrit 384, dual
    traversal with theta 0.5 and the 3h gap, P = 6 terms, smoothed near
    pairs), timed for one product of the whole workload on the host's cores.
    The step is that product times (matvecs + 2): the solve's products plus a
    force product costed at two field products (its pair work is twice the
    field's).  A bounded FP64 all-pairs row sample, extrapolated, is reported
    beside it as a second figure."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(16, os.cpu_count() or 1)  # the box's CPU share for one GPU
    pos = system.pos.astype(np.float64)
    n = system.n
    rng = np.random.default_rng(1)
    m = rng.normal(size=(n, 3)) * 0.5
    t0 = time.perf_counter()
    field, st = oracle.fmm_field(pos, m, system.h, ncrit=384, theta=0.5, hconst=None,
                                 nthreads=threads)
    t_fmm = time.perf_counter() - t0
    # its accuracy on 256 seeded rows against the FP64 all-pairs oracle
    sel = rng.choice(n, size=min(256, n), replace=False)
    ref = oracle.field_allpairs(pos[sel], pos, m, system.h, nthreads=threads)
    err = float(np.linalg.norm(field[sel] - ref) / np.linalg.norm(ref))
    # second figure: FP64 all-pairs rows, extrapolated
    rows, done, t_ap = 64, 0, 0.0
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < budget_s * 0.25:
        s_rows = rng.choice(n, size=rows, replace=False)
        t1 = time.perf_counter()
        oracle.field_allpairs(pos[s_rows], pos, m, system.h, nthreads=threads)
        t_ap += time.perf_counter() - t1
        done += rows
    ap_step_ms = t_ap / max(done, 1) * n * (matvecs_per_step + 2) * 1e3
    out = {"value": t_fmm * (matvecs_per_step + 2) * 1e3, "unit": "ms/step", "cores": threads,
           "kind": "port",
           "sample": (f"one FP64 CPU FMM field product of the whole workload (restatement of the "
                      f"reference's fmmtl path: ncrit 384, theta 0.5, 3h gap, P = 6; "
                      f"{st['leaves']} leaves, {st['far_box_pairs']} far and "
                      f"{st['near_box_pairs']} near box pairs) took {t_fmm:.1f} s on {threads} "
                      f"threads; step = {matvecs_per_step} field products + a force product "
                      f"costed as 2"),
           "fmm_field_product_s": t_fmm, "fmm_rel_err_vs_fp64_allpairs": err, "fmm_tree": st,
           "allpairs_extrapolation": {"value": ap_step_ms, "unit": "ms/step",
                                      "sample": f"{done} FP64 all-pairs target rows x {n} "
                                                f"sources in {t_ap:.1f} s, extrapolated"}}
    out.update(cpu_info())
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the path is "
              f"measured on one GPU's share and multi-GPU sharding is not implemented",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import magforce
    from magforce import configs

    log(f"bench: workload={args.workload} steps={args.steps} warmup={args.warmup}")
    ctx = magforce.Context(0)
    system = configs.WORKLOADS[args.workload]()
    mode = {"fmm": magforce.MODE_FMM, "allpairs": magforce.MODE_ALLPAIRS}.get(args.mode)
    if mode is None:
        try:
            p = magforce.Plan(ctx, to_dev(system.pos[:64]), system.h, mode=magforce.MODE_FMM)
            p.close()
            mode = magforce.MODE_FMM
        except magforce.MagforceError:
            mode = magforce.MODE_ALLPAIRS
    mode_name = "fmm" if mode == magforce.MODE_FMM else "allpairs"
    if args.copy_calib_mb > 0:
        nbytes = args.copy_calib_mb << 20
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy calibration: {nbytes} bytes read + {nbytes} bytes written")
        del a, b
    m2l_code = {"fp32": magforce.M2L_FP32, "bf16x6": magforce.M2L_BF16X6,
                "bf16x3": magforce.M2L_BF16X3}[args.m2l]
    def timed_steps(stepper, timing):
        for w in range(args.warmup):
            stepper.step()
            torch.cuda.synchronize()
            log(f"warmup {w} done: {stepper.stats}")
        ctx.reset_timing()
        ctx.enable_timing(timing)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for k in range(args.steps):
            stepper.step()
        e1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_time(e1) / args.steps
        kt = {k: ctx.kernel_time(k) for k in
              ("field", "force", "p2m", "m2m", "split", "m2l", "m2l_relaxed", "m2l_force", "l2l",
               "l2p_field", "p2p_field", "l2p_force", "p2p_force")}
        ctx.enable_timing(False)
        return ms, wall, kt

    # Per-kernel durations (roofline, breakdown) come from a run with the near
    # field serialized on one stream (serial_near): each kernel then has the
    # GPU to itself.  --profile-only runs only that configuration, so the
    # rocprofv3 averages and the roofline describe the same launches.
    kst = Stepper(ctx, system, mode, args.rtol, m2l_code, serial_near=True,
                  relax=relax_kwargs(args))
    ms_serial, _, ktimes = timed_steps(kst, True)
    log(f"kernel timing (near field serialized): {ms_serial:.2f} ms/step")
    if args.profile_only:
        print(json.dumps({"ms_per_step_serial": ms_serial, "matvecs": kst.stats["matvecs"]}))
        return
    kst.plan.close()
    del kst
    # The measured step: the near field overlaps the M2L (the default).
    st = Stepper(ctx, system, mode, args.rtol, m2l_code, relax=relax_kwargs(args))
    ms_step, wall, _ = timed_steps(st, False)
    log(f"timed: {args.steps} steps, {ms_step:.2f} ms/step (wall {wall / args.steps * 1e3:.2f})")
    matvecs = st.stats["matvecs"]

    n = system.n
    info = st.plan.info()

    def kernel_roofline(name, flops, peak, bound):
        ms, cnt = ktimes[name]
        avg = ms / max(cnt, 1)
        achieved = flops / (avg * 1e-3) / 1e12 if avg > 0 else 0.0
        return {"kernel": name, "bound": bound, "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                "avg_launch_ms": avg, "launches": cnt, "ms_per_step": ms / args.steps,
                "flops_per_launch": flops}

    if mode == magforce.MODE_FMM:
        # P2P: direct pairs x 32 FLOPs against the FP32 vector peak.  M2L:
        # algorithmic work 2 (p+1)^4 FLOPs per box pair (dense (p+1)^2 x (p+1)^2
        # real operator, padding excluded) against the published peak of the
        # MFMA instruction the kernel issues (FP32 MFMA 157.3, BF16 dense 2500
        # TFLOP/s).  The MFMA work actually issued (64-target tiles, KP-padded
        # operators, 3 or 6 BF16 products per FP32 product) and the algorithmic
        # bytes per launch are reported beside it.
        order = info["order"]
        kp = ((order + 1) ** 2 + 31) // 32 * 32
        m2l_flops = float(info["m2l_field_pairs"]) * 2.0 * (order + 1) ** 4
        m2l_peak = FP32_PEAK_TFLOPS if args.m2l == "fp32" else BF16_PEAK_TFLOPS
        m2l = kernel_roofline("m2l", m2l_flops, m2l_peak, "mfma")
        issued = float(info["m2l_field_tile_slots"]) * m2l_issued_per_tile_slot(order, args.m2l)
        m2l["issued_mfma_flops_per_launch"] = issued
        m2l["issued_mfma_frac"] = (issued / (m2l["avg_launch_ms"] * 1e-3) / 1e12 / m2l_peak
                                   if m2l["avg_launch_ms"] > 0 else 0.0)
        # each source multipole read once, each local written once, each distinct
        # offset operator read once; 4 bytes per coefficient (FP32, or two BF16 parts)
        m2l["algorithmic_bytes_per_launch"] = (2.0 * info["boxes"] * (order + 1) ** 2 * 4
                                               + float(info["m2l_field_operators"])
                                               * (order + 1) ** 4 * 4)
        cands = [kernel_roofline("p2p_field", float(info["p2p_field_pairs"]) * FIELD_FLOPS_PER_PAIR,
                                 FP32_PEAK_TFLOPS, "valu"), m2l]
        cands.sort(key=lambda r: -r["ms_per_step"])
        roofline = cands[0]
        roofline["others"] = cands[1:]
        breakdown = {}
        for k in ("p2m", "m2m", "split", "m2l", "m2l_relaxed", "m2l_force", "l2l", "l2p_field",
                  "p2p_field", "l2p_force", "p2p_force"):
            ms, cnt = ktimes[k]
            breakdown[k] = {"ms_total": ms / args.steps, "launches": cnt // max(args.steps, 1)}
        roofline["per_step_breakdown"] = breakdown
        roofline["kernel_timing"] = {"near_field": "serialized (serial_near=1), one kernel at a time",
                                     "ms_per_step_serial": ms_serial}
        roofline["plan"] = info
    else:
        roofline = kernel_roofline("field", float(n) * n * FIELD_FLOPS_PER_PAIR, FP32_PEAK_TFLOPS,
                                   "valu")
    traffic_path = args.traffic or os.path.join(
        ROOT, "profiles", f"traffic_{mode_name}_{system.name}.json")
    if os.path.exists(traffic_path):
        # HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes
        # summarised under profiles/ (one file per workload, one entry per kernel)
        with open(traffic_path) as f:
            t = json.load(f)
        if t.get("workload") == system.name:
            ks = t.get("kernels", {})
            for r in [roofline] + roofline.get("others", []):
                if r["kernel"] in ks:
                    r["traffic"] = ks[r["kernel"]]["hbm_bytes_per_launch"]
                    r["traffic_source"] = os.path.relpath(traffic_path, ROOT)
    for r in [roofline] + roofline.get("others", []):
        if r["kernel"] != "m2l":
            continue
        # The bound follows the measurements: the fraction of HBM bandwidth the
        # counted traffic takes against the fraction of the MFMA peak the issued
        # MFMA work takes.  Under an HBM bound, achieved and peak are restated as
        # algorithmic bytes per launch / launch time against 8 TB/s.
        t_s = r["avg_launch_ms"] * 1e-3
        alg_b = r["algorithmic_bytes_per_launch"]
        if r.get("traffic") and t_s > 0:
            r["traffic_ratio"] = r["traffic"] / alg_b
            r["hbm_frac_counted"] = r["traffic"] / t_s / (HBM_PEAK_GBS * 1e9)
            if r["hbm_frac_counted"] > r["issued_mfma_frac"]:
                r["bound"] = "hbm"
                r["algorithmic_tflops"] = r["achieved"]
                r["achieved"] = alg_b / t_s / 1e9
                r["peak"] = HBM_PEAK_GBS
                r["unit"] = "GB/s"
                r["frac"] = r["achieved"] / HBM_PEAK_GBS

    compare = None
    if mode == magforce.MODE_FMM and args.m2l != "fp32" and not args.no_compare:
        compare = time_fp32_fmm(ctx, system, args)
        compare["speedup"] = compare["ms_per_step"] / ms_step
    refined = None
    if mode == magforce.MODE_FMM and not args.no_compare and n <= 2_000_000:
        refined = time_refined(ctx, system, args, m2l_code)
    batch = None if args.no_batch else run_droplets(ctx, args, m2l_code)
    c1 = None if args.no_cpu else run_c1(ctx)
    allpairs = None if args.no_allpairs else measure_allpairs(ctx)
    if allpairs is not None:
        # plain FP32 all-pairs step at this n, from the measured pair rates
        # (force tensor pairs cost 2x field pairs): an estimate, not a run
        per_field_ms = float(n) * n / (allpairs["gpairs_per_s"] * 1e9) * 1e3
        allpairs["est_step_ms_at_workload_n"] = per_field_ms * (matvecs + 2)
        allpairs["est_speedup_of_value"] = allpairs["est_step_ms_at_workload_n"] / ms_step
    check = None
    if not args.no_check:
        if n > 2_000_000:
            check = check_error_subsample(ctx, st, args.check_samples or 65536)
        else:
            check = check_error(ctx, st, args.check_samples or 4096)
    timestep = None
    if mode == magforce.MODE_FMM and not args.no_timestep:
        timestep = run_timestep(ctx, st, max(args.steps, 3))
    cpu = None if args.no_cpu else cpu_baseline(system, args.cpu_seconds, matvecs)

    line = {
        "metric": "ms/step (M-solve+force) at N=1M, err<=1e-5 vs FP64; Gpairs/s all-pairs",
        "value": ms_step, "unit": "ms/step", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{system.name} (n={n}, h={system.h}, chi={system.chi}, "
                               f"H=(0,0,1), CG rtol {args.rtol}, {mode_name}"
                               + (f", M2L {args.m2l}" + (f", products below residual "
                                  f"{args.relax_rtol:g} at {args.relax})" if args.relax != "none"
                                  else ")") if mode == magforce.MODE_FMM else ")"),
                   "mode": mode_name, "n": n, "matvecs_per_step": matvecs,
                   "m2l_precision": args.m2l if mode == magforce.MODE_FMM else None,
                   "relax": relax_info(args) if mode == magforce.MODE_FMM else None,
                   "solver": st.stats},
        "error": check,
        "fp32_fmm": compare,
        "refined_fmm": refined,
        "batch": batch,
        "c1_lattice8k": c1,
        "allpairs": allpairs,
        "timestep": timestep,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
