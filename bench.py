#!/usr/bin/env python3
"""SISL step throughput at 256x256x80 on one MI355X: mixed precision vs FP64.

Workload: the limited-area case ``lam`` of suetes_amd.workloads (256x256x80,
dx=1 km, stretched SLEVE to 20 km, seeded Gaussian random terrain, N=0.01
background, seeded +-2 m/s winds, 10-cell Davies rim, dt=10 s; GMRES restart
20, tol 1e-10).  One "step" = one full SISL step (departure points, explicit
tendencies, six tricubic gathers, FFSL, two implicit solves, boundaries).

Prints ONE JSON line.  ``value`` = mixed-precision steps/s with the state
resident in HBM; the same line carries the FP64 rate, the mixed-vs-FP64 field
error after the timed steps, the adjoint/forward cost ratio, the roofline of
the dominant kernel and a CPU baseline (the FP64 oracle on a bounded sample).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SISL steps/s at 256²×80, mixed vs FP64 (err≤1e-5); adjoint/forward cost"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def fields_err(a, b):
    import torch
    worst = 0.0
    for k in ("u", "v", "w", "pi", "th_v", "rho"):
        x, y = a[k].double(), b[k].double()
        worst = max(worst, float((x - y).norm() / y.norm().clamp_min(1e-300)))
    return worst


def run_steps(case, n, state=None):
    import torch
    st = case.state if state is None else state
    for i in range(n):
        st = case.stepper.step(st, i * case.dt, case.forcing, case.bc_fn)
    return st


def time_steps(case, steps, warmup):
    import torch
    st = run_steps(case, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = run_steps(case, steps, st)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, st


def kernel_roofline(case, reps=20):
    """Average launch duration of the operator-apply kernel (FP32, the mixed
    path's inner kernel) measured with events on its stream, against its
    algorithmic bytes per launch."""
    import torch
    from suetes_amd import _lib
    sol = case.stepper.implicit_solver
    sol.prepare(case.stepper.bg_precomputed())
    g = case.grid
    n = int(_lib.lib().sisl_krylov_size(C.byref(sol.op64.struct.g)))
    nx, ny, nz = g.nx, g.ny, g.nz
    x = torch.randn(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    L = _lib.lib()
    s = torch.cuda.current_stream()
    fn = L.sisl_helmholtz_apply_f32
    for _ in range(3):
        _lib.check(fn(C.byref(sol.op32.struct), x.data_ptr(), y.data_ptr(), s.cuda_stream), "A")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn(C.byref(sol.op32.struct), x.data_ptr(), y.data_ptr(), s.cuda_stream)
    e1.record(s)
    torch.cuda.synchronize()
    dt = e0.elapsed_time(e1) / reps * 1e-3
    Nu, Nv, Nw, Nm = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1), nx * ny * nz
    # algorithmic bytes: x and y once each (5 fields), 13 3-D coefficient arrays once
    coef = 3 * Nu + 3 * Nv + 6 * Nw + Nm
    nbytes = 4 * (2 * n + coef)
    achieved = nbytes / dt / 1e9
    return {"kernel": "helmholtz_kernel<float> (operator apply A)", "bound": "hbm", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "algorithmic_bytes": nbytes, "avg_launch_us": dt * 1e6}


def copy_reference(nbytes=1 << 29, reps=5):
    """A plain device copy of known size (reads and writes `nbytes` each):
    the scale for the FETCH_SIZE / WRITE_SIZE hardware counters."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    return nbytes


def adjoint_cost(case, nsteps=1):
    """(t_value_and_grad - t_forward) / t_forward over `nsteps` steps
    (core_ad_scaling/run.py:255-258 definition), on the mixed path."""
    import torch

    def fwd(u0):
        st = dict(case.state)
        st["u"] = u0
        st = run_steps(case, nsteps, st)
        return (st["w"] ** 2).mean()
    u0 = case.state["u"].detach().clone()
    with torch.no_grad():
        fwd(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        fwd(u0)
    torch.cuda.synchronize()
    tf = time.perf_counter() - t0
    uu = u0.clone().requires_grad_(True)
    J = fwd(uu)
    torch.autograd.grad(J, uu)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    uu = u0.clone().requires_grad_(True)
    J = fwd(uu)
    (g,) = torch.autograd.grad(J, uu)
    torch.cuda.synchronize()
    tvg = time.perf_counter() - t0
    return (tvg - tf) / tf, tf, tvg


def cpu_baseline(seconds=20.0):
    """The FP64 NumPy oracle, single thread, timing whole steps of the same
    workload on a 48x48x80 sub-grid and scaling by grid points to 256x256x80."""
    try:
        from threadpoolctl import threadpool_limits
    except Exception:  # pragma: no cover
        threadpool_limits = None
    from suetes_amd import workloads as W
    from tests.helpers import oracle_case
    cfg = W.lam(nx=48, ny=48, nz=80)
    g, op, st, s, bc, ext = oracle_case(cfg)
    ctx = threadpool_limits(limits=1) if threadpool_limits else None
    t0 = time.perf_counter()
    n = 0
    try:
        while True:
            s = st.step(s, n * 10.0, ext, bc)
            n += 1
            if time.perf_counter() - t0 > seconds or n >= 3:
                break
    finally:
        if ctx is not None:
            ctx.__exit__(None, None, None)
    dt = (time.perf_counter() - t0) / n
    scale = (48 * 48 * 80) / (256 * 256 * 80)
    return {"value": scale / dt, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} FP64 oracle steps of the same workload on a 48x48x80 sub-grid "
                      f"({dt:.2f} s/step), scaled by grid points to 256x256x80"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--nz", type=int, default=80)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only launch the operator kernel and a plain copy of known size "
                         "(the run that hardware-counter passes profile)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from suetes_amd import workloads as W
    from suetes_amd.cases import build_case

    cfg = W.lam(nx=args.nx, ny=args.nx, nz=args.nz)
    log(f"building workload lam {args.nx}x{args.nx}x{args.nz}")
    if args.kernels_only:
        cmx = build_case(cfg, device="cuda", precision="mixed")
        roof = kernel_roofline(cmx)
        nbytes = copy_reference()
        print(json.dumps({"roofline": roof, "copy_bytes": nbytes}), flush=True)
        return
    c64 = build_case(cfg, device="cuda", precision="fp64")
    cmx = build_case(cfg, device="cuda", precision="mixed")
    log("timing mixed")
    t_mx, s_mx = time_steps(cmx, args.steps, args.warmup)
    log(f"mixed {t_mx*1e3:.1f} ms/step; timing fp64")
    t_64, s_64 = time_steps(c64, args.steps, args.warmup)
    log(f"fp64 {t_64*1e3:.1f} ms/step")
    err = fields_err(s_mx, s_64)
    st_mx = cmx.stepper.implicit_solver.last_stats[-4:]
    st_64 = c64.stepper.implicit_solver.last_stats[-4:]
    roof = kernel_roofline(cmx)
    log(f"roofline {roof['achieved']:.0f} GB/s")
    adj = None
    if not args.no_adjoint:
        try:
            r, tf, tvg = adjoint_cost(cmx, 1)
            adj = {"ratio": r, "t_forward_s": tf, "t_value_and_grad_s": tvg}
            log(f"adjoint/forward {r:.2f}")
        except Exception as e:  # report, do not hide
            adj = {"error": repr(e)[:300]}
    cpu = None if args.no_cpu else cpu_baseline()
    npts = args.nx * args.nx * args.nz
    out = {
        "metric": METRIC, "value": 1.0 / t_mx, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_mx * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 kernels + f64 Krylov/state", "data": "synthetic",
        "config": {"workload": f"lam {args.nx}x{args.nx}x{args.nz} dx=1km dt=10s stretched-SLEVE "
                               "GRF terrain, Davies rim 10, GMRES restart 20 tol 1e-10",
                   "grid": [args.nx, args.nx, args.nz], "dt": 10.0},
        "fp64_steps_per_s": 1.0 / t_64, "fp64_ms_per_step": t_64 * 1e3,
        "speedup_mixed_vs_fp64": t_64 / t_mx,
        "field_err_mixed_vs_fp64": err,
        "grid_point_updates_per_s": npts / t_mx,
        "solver_stats_mixed": st_mx, "solver_stats_fp64": st_64,
        "adjoint_forward_cost": adj,
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
