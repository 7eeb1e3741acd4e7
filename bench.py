#!/usr/bin/env python3
"""SISL step throughput at 256x256x80 on one MI355X: mixed precision vs FP64.

Workload: the limited-area case ``lam`` of suetes_amd.workloads (256x256x80,
dx=1 km, stretched SLEVE to 20 km, seeded Gaussian random terrain, N=0.01
background, seeded +-2 m/s winds with 10 km correlation, 10-cell Davies rim with
equal inflow/outflow weights, dt=10 s; GMRES restart
20, tol 1e-10).  One "step" = one full SISL step (departure points, explicit
tendencies, six tricubic gathers, FFSL, two implicit solves, boundaries).

Prints ONE JSON line.  ``value`` = mixed-precision steps/s with the state
resident in HBM; the same line carries the FP64 rate, the mixed-vs-FP64 field
error after the timed steps, the adjoint/forward cost ratios (mixed and
FP64), per-kernel event timings of one instrumented step with the roofline of
the dominant kernel, and a CPU baseline (the FP64 oracle on a bounded sample).
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


KERNEL_NAMES = {
    "helmholtz": "helmholtz_kernel (operator apply A)",
    "precond": "precond_seg_kernel (column preconditioner M)",
    "krylov_dots": "dots_kernel (CGS projections)",
    "krylov_update": "update_kernel (CGS subtraction + norm)",
    "krylov_normalize": "normalize_gated_kernel (basis vector)",
    "tricubic": "tricubic_kernel (semi-Lagrangian gather)",
    "helmholtz_T": "helmholtz_T_kernel (operator transpose A^T)",
    "precond_T": "precond_T_seg_kernel + faces (preconditioner transpose M^T)",
    "tricubic_adj": "tricubic_adj_kernel (deterministic gather transpose)",
    "helmholtz_h16": "helmholtz_kernel on the FP16 basis column (operator apply A)",
    "helmholtz_T_h16": "helmholtz_T_kernel on the FP16 basis column (operator transpose A^T)",
}


def _timer_table():
    from suetes_amd import _lib
    table = []
    for k in _lib.TIMED:
        for fp32 in (True, False):
            ms, n, nw, by = _lib.timer_read(k, fp32)
            if n == 0:
                continue
            table.append({"kernel": KERNEL_NAMES[k], "type": "fp32" if fp32 else "fp64",
                          "launches": n, "working_launches": nw, "total_ms": ms,
                          "avg_launch_us": 1e3 * ms / n,
                          "algorithmic_bytes_per_working_launch": by / max(nw, 1),
                          "achieved_GBs": by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0})
    table.sort(key=lambda r: -r["total_ms"])
    return table


def instrumented_adjoint(case):
    """Per-kernel launch timing over one value-and-gradient step (forward and
    reverse pass; the transposed kernels appear only here)."""
    import torch
    from suetes_amd import _lib
    st = dict(case.state)
    u0 = st["u"].detach().clone().requires_grad_(True)
    st["u"] = u0
    torch.cuda.synchronize()
    _lib.timer_enable(True)
    out = case.stepper.step(st, 0.0, case.forcing, case.bc_fn)
    torch.autograd.grad((out["w"] ** 2).mean(), u0)
    torch.cuda.synchronize()
    _lib.timer_enable(False)
    return _timer_table()


def instrumented_step(case, state):
    """One more step of the same workload with the library's launch timer on:
    every launch of the hot kernels is bracketed by HIP events on its own
    stream.  Returns per-kernel totals and the roofline of the kernel with
    the largest total time."""
    import torch
    from suetes_amd import _lib
    torch.cuda.synchronize()
    _lib.timer_enable(True)
    case.stepper.step(state, 0.0, case.forcing, case.bc_fn)
    torch.cuda.synchronize()
    _lib.timer_enable(False)
    table = _timer_table()
    return _roofline(table), table


# HBM bytes per launch of the dominant kernels from the two hardware-counter
# passes (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) on
# this workload, scaled by the same counters of a plain device copy of known
# size (read factor 2.00, write factor 1.00); profiles/sisl_step_lam256x256x80.md.
# FP32 operator on the FP16 basis column with bfloat16 coefficients
# (helmholtz2_kernel<float, 4, bf16>, the inner steps after the first of each
# cycle: the bulk of the operator launches of a mixed step), and the launch
# with the fused basis normalisation (the first step of each cycle).
MEASURED_TRAFFIC = {
    "helmholtz_kernel on the FP16 basis column (operator apply A) [fp32]": (99172 * 2.0 + 103075 * 1.0) * 1024,
    "helmholtz_kernel (operator apply A) [fp32]": (129684 * 2.0 + 154647 * 1.0) * 1024,
}


def _roofline(table):
    top = table[0]
    key = f"{top['kernel']} [{top['type']}]"
    roof = {"kernel": key, "bound": "hbm",
            "achieved": top["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": top["achieved_GBs"] / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC.get(key),
            "avg_launch_us": top["avg_launch_us"],
            "algorithmic_bytes_per_launch": top["algorithmic_bytes_per_working_launch"],
            "method": "HIP events around every launch on its stream during one instrumented step; "
                      "achieved = algorithmic bytes of working launches / event time of all launches"}
    return roof


def attainable(case, reps=20):
    """Attainable HBM rates on this device, timed with HIP events on the
    launch stream: a 1 GiB copy and a 1 GiB read with 16-byte accesses, and
    the inner operator launches' access mixes alone (their input and output
    arrays, each touched once at the thread's own column, in their tiles and
    level chunks, no stencil or arithmetic: csrc/probe.hip) against the
    algorithmic bytes of those launches: the launch on the stored FP16 basis
    column (5 FP16 inputs, 13 bfloat16 coefficient arrays, 5 FP32 outputs)
    and the launch with the fused basis normalisation (FP32 inputs and 5 FP16
    basis outputs besides)."""
    import ctypes as C
    import torch
    from suetes_amd import _lib
    L = _lib.lib()
    vp = C.c_void_p
    L.sisl_probe_stream.argtypes = [vp, vp, C.c_int64, C.c_int32, vp]
    L.sisl_probe_operator_streams_f32.argtypes = [vp, vp, vp, vp, C.c_int32, vp]
    L.sisl_probe_operator_streams_h16_f32.argtypes = [vp, vp, vp, C.c_int32, vp]
    s = _lib.stream_ptr()

    def timed(fn):
        for _ in range(3):
            _lib.check(fn(), "probe")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e-3 / reps
    nb = 1 << 30
    a = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    t_copy = timed(lambda: L.sisl_probe_stream(a.data_ptr(), b.data_ptr(), nb, 0, s))
    t_read = timed(lambda: L.sisl_probe_stream(a.data_ptr(), b.data_ptr(), nb, 1, s))
    del a, b
    sol = case.stepper.implicit_solver
    ops = sol.op_inner
    g = ops.struct.g
    n = int(L.sisl_krylov_size(C.byref(g)))
    x = torch.randn(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    bas = torch.empty(n, dtype=torch.float16, device="cuda")
    nu, nv = (g.nx + 1) * g.ny * g.nz, g.nx * (g.ny + 1) * g.nz
    nw, nm = g.nx * g.ny * (g.nz + 1), g.nx * g.ny * g.nz
    alg = 4 * 2 * n + 2 * (3 * nu + 3 * nv + 6 * nw + nm) + 2 * n
    alg16 = (2 + 4) * n + 2 * (3 * nu + 3 * nv + 6 * nw + nm)
    x16 = x.half()
    best = best16 = None
    for kc in (3, 6, 11):
        t = timed(lambda: L.sisl_probe_operator_streams_f32(C.byref(ops.struct), x.data_ptr(), y.data_ptr(),
                                                            bas.data_ptr(), kc, s))
        best = t if best is None else min(best, t)
        t = timed(lambda: L.sisl_probe_operator_streams_h16_f32(C.byref(ops.struct), x16.data_ptr(),
                                                                y.data_ptr(), kc, s))
        best16 = t if best16 is None else min(best16, t)
    return {"copy_GBs": 2 * nb / t_copy / 1e9, "read_GBs": nb / t_read / 1e9,
            "operator_access_mix_GBs": alg16 / best16 / 1e9, "operator_access_mix_us": best16 * 1e6,
            "operator_access_mix_fused_GBs": alg / best / 1e9, "operator_access_mix_fused_us": best * 1e6,
            "method": "HIP events around 20 launches of each probe (csrc/probe.hip); the access-mix "
                      "rate is the operator launch's algorithmic bytes over the probe's time "
                      "(operator_access_mix: the launch on the FP16 basis column; _fused: the launch "
                      "with the fused basis normalisation)"}


def operator_shapes(case):
    """The operator launch shapes in force (levels per thread, unroll; unroll
    3 = two columns per thread) for the launches of the mixed and FP64 steps."""
    import ctypes as C
    from suetes_amd import _lib
    L = _lib.lib()
    g = case.stepper.implicit_solver.op64.struct.g
    out = {}
    for name, fp32, var in (("fp32 A on the FP16 basis column, bf16 coefficients (inner forward)", 1, 14),
                            ("fp32 A^T on the FP16 basis column (inner transposed)", 1, 9),
                            ("fp32 A, fused normalisation, bf16 coefficients (inner forward, first step)", 1, 6),
                            ("fp32 A^T, fused normalisation (inner transposed, first step)", 1, 3),
                            ("fp64 A (FP64 solver; refinement residual)", 0, 0),
                            ("fp64 A^T", 0, 1)):
        kc, un = C.c_int32(), C.c_int32()
        _lib.check(L.sisl_get_operator_shape(C.byref(g), fp32, var, C.byref(kc), C.byref(un)), "shape")
        out[name] = [kc.value, un.value]
    return out


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
    times = []
    for _ in range(3):  # median of three value-and-gradient runs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        uu = u0.clone().requires_grad_(True)
        J = fwd(uu)
        (g,) = torch.autograd.grad(J, uu)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        del J, g, uu
    tvg = sorted(times)[1]
    return (tvg - tf) / tf, tf, tvg


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(repeats=5, sub=48):
    """The FP64 NumPy oracle of the same workload on a sub x sub x 80 sub-grid
    (same generator, seed and settings), one step from the initial state
    timed `repeats` times after an untimed warm-up, median taken
    (cpu_gpu_short_scaling/run.py:135-167 protocol); scaled by grid points
    to 256x256x80.  The BLAS / OpenMP pools get every host core (the
    Gram-Schmidt projections and updates of the oracle's GMRES are BLAS
    calls); NumPy's elementwise stencil work runs on one thread.  The
    sub-grid's Arnoldi steps per time step are reported beside the GPU's at
    full size, since the per-step solver work scales with them."""
    import statistics
    try:
        from threadpoolctl import threadpool_limits
    except Exception:  # pragma: no cover
        threadpool_limits = None
    from suetes_amd import workloads as W
    from tests.helpers import oracle_case
    # every host core this process may use: the CPU share the environment
    # grants (OMP_NUM_THREADS) when it names one, else all cores
    cores = int(os.environ.get("OMP_NUM_THREADS", 0) or 0) or os.cpu_count() or 1
    cfg = W.lam(nx=sub, ny=sub, nz=80)
    g, op, st, s0, bc, ext = oracle_case(cfg)
    ctx = threadpool_limits(limits=cores) if threadpool_limits else None
    times, arnoldi = [], []
    try:
        for rep in range(repeats + 1):
            s = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in s0.items()}
            st.solver.stats = {}
            t0 = time.perf_counter()
            st.step(s, 0.0, ext, bc)
            dt = time.perf_counter() - t0
            if rep:  # the first step is the warm-up
                times.append(dt)
                arnoldi.append(st.solver.stats.get("cycles", 0) * st.solver.solver_restart)
    finally:
        if ctx is not None:
            ctx.__exit__(None, None, None)
    med = statistics.median(times)
    log(f"cpu baseline: {med:.2f} s per {sub}x{sub}x80 step on {cores} threads")
    scale = (sub * sub * 80) / (256 * 256 * 80)
    return {"value": scale / med, "unit": "steps/s", "cores": cores, "kind": "port",
            "cpu_model": _cpu_model(), "os_cpu_count": os.cpu_count(),
            "median_s_per_substep": med, "repeats_s": times,
            "oracle_arnoldi_steps_per_step": arnoldi[-1] if arnoldi else None,
            "sample": f"median of {repeats} FP64 oracle steps (after one warm-up) of the same workload on a "
                      f"{sub}x{sub}x80 sub-grid ({med:.2f} s/step), scaled by grid points to 256x256x80; "
                      f"BLAS pool {cores} threads, elementwise NumPy on one; the sub-grid solve took "
                      f"{arnoldi[-1] if arnoldi else '?'} Arnoldi steps per time step (GPU FP64 at full size: "
                      "see solver_stats_fp64)"}


def terrain_gradient(args):
    """Reverse mode through the terrain-bump case (128x128x60, 50 steps, dx=1
    km, 8 bump amplitudes, J = mean w^2 over a downstream 20x20x10 box): mixed
    and FP64 value-and-gradient with one checkpoint per step, a central
    difference and a Taylor test along one direction (FP64, replayed
    departure points, differentiated edge padding)."""
    import numpy as np
    import torch
    from suetes_amd import workloads as W
    from suetes_amd.cases import terrain_objective
    cfg = W.adjoint(nx=args.nx, ny=args.nx, nz=args.nz)
    n = cfg["steps"] if args.adj_steps is None else args.adj_steps
    a0 = torch.tensor(cfg["terrain"][1]["amps"], dtype=torch.float64, device="cuda")
    out = {"metric": "terrain-control gradient: value+grad time, FD and Taylor agreement",
           "config": {"workload": f"adjoint {args.nx}x{args.nx}x{args.nz} dx=1km dt=10s, 8 Gaussian "
                                  "bumps (sigma 5 km), J = mean w^2 over a 20x20x10 box",
                      "steps": n}}

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    grads = {}
    for prec in ("mixed", "fp64"):
        with torch.no_grad():
            terrain_objective(cfg, a0, 1, precision=prec)  # warm-up (autotune, library load)
            _, tf = timed(lambda: terrain_objective(cfg, a0, n, precision=prec))
        # warm-up of the reverse pass (its kernels' first launches, captured cycles)
        aw = a0.clone().requires_grad_(True)
        torch.autograd.grad(terrain_objective(cfg, aw, 1, precision=prec, checkpoint=True), aw)

        def vg():
            a = a0.clone().requires_grad_(True)
            J = terrain_objective(cfg, a, n, precision=prec, checkpoint=True)
            (g,) = torch.autograd.grad(J, a)
            return float(J), g
        (J, g), tvg = timed(vg)
        grads[prec] = g
        out[prec] = {"J": J, "t_forward_s": tf, "t_value_and_grad_s": tvg,
                     "adjoint_forward_ratio": (tvg - tf) / tf,
                     "peak_mem_GB": torch.cuda.max_memory_allocated() / 1e9}
        log(f"{prec}: forward {tf:.2f} s, value+grad {tvg:.2f} s")
        torch.cuda.reset_peak_memory_stats()
    out["grad_mixed_vs_fp64_rel"] = float((grads["mixed"] - grads["fp64"]).norm() / grads["fp64"].norm())
    out["speedup_value_and_grad_mixed_vs_fp64"] = out["fp64"]["t_value_and_grad_s"] / out["mixed"]["t_value_and_grad_s"]
    # exact-derivative check along one direction
    rec = []
    with torch.no_grad():
        terrain_objective(cfg, a0, n, record=rec)

    def J(a):
        return terrain_objective(cfg, a, n, coords=rec, exact_edges=True)
    a = a0.clone().requires_grad_(True)
    val = J(a)
    (g,) = torch.autograd.grad(val, a)
    d = torch.linspace(-1.0, 1.0, 8, dtype=torch.float64, device="cuda")
    d = d / d.norm()
    dJ = float((g * d).sum())
    h = 0.05
    with torch.no_grad():
        fd = (float(J(a0 + h * d)) - float(J(a0 - h * d))) / (2 * h)
        rem = [abs(float(J(a0 + e * d)) - float(val) - e * dJ) for e in (20.0, 10.0, 5.0)]
    out["fd_check"] = {"direction": "linspace(-1, 1, 8) normalised", "h_m": h, "fd": fd, "adjoint": dJ,
                       "rel_err": abs(fd - dJ) / abs(fd),
                       "taylor_remainders": rem,
                       "taylor_orders": [float(np.log2(rem[i] / rem[i + 1])) for i in range(2)]}
    out["grad_fp64"] = [float(x) for x in grads["fp64"]]
    log(f"fd rel err {out['fd_check']['rel_err']:.2e}, mixed vs fp64 grad {out['grad_mixed_vs_fp64_rel']:.2e}")
    print(json.dumps(out), flush=True)


def ensemble_run(args):
    """The batched ensemble (32 members on seeded uneven grids 96..192 squared
    x 48..80 levels, theta noise 0.5 K, mixed precision): lockstep batched
    steps against the same members stepped one after another, then
    forward + reverse mode over the steps with one checkpoint per ensemble
    step."""
    import torch
    from suetes_amd import workloads as W
    from suetes_amd.cases import build_case
    from suetes_amd.ensemble import Ensemble
    cfgs = W.ensemble_members(M=args.members)
    n = args.ens_steps
    cases = [build_case(c, device="cuda", precision="mixed") for c in cfgs]
    pts = sum(c.grid.nx * c.grid.ny * c.grid.nz for c in cases)
    ens = Ensemble([c.stepper for c in cases])
    forc, bcs = [c.forcing for c in cases], [c.bc_fn for c in cases]
    log(f"ensemble of {len(cases)} members, {pts / 1e6:.1f} M points")
    with torch.no_grad():
        ens.integrate([dict(c.state) for c in cases], 0.0, 1, forc, bcs)  # warm-up
        for c in cases:
            c.stepper.integrate(dict(c.state), 0.0, 1, c.forcing, c.bc_fn)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lock = ens.integrate([dict(c.state) for c in cases], 0.0, n, forc, bcs)
        torch.cuda.synchronize()
        t_lock = time.perf_counter() - t0
        log(f"lockstep {n} steps: {t_lock:.2f} s")
        t0 = time.perf_counter()
        seq = [c.stepper.integrate(dict(c.state), 0.0, n, c.forcing, c.bc_fn) for c in cases]
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t0
        log(f"member by member {n} steps: {t_seq:.2f} s")
        diff = max(float((a[k].double() - b[k].double()).norm() / b[k].double().norm())
                   for a, b in zip(lock, seq) for k in ("u", "v", "w", "pi", "th_v", "rho"))
        # the same lockstep step recorded once as a hipGraph and replayed
        from suetes_amd.steppers import state_to_internal
        start = [state_to_internal(s) for s in lock]
        t1 = n * cases[0].dt
        eager = ens.step_internal(start, t1, forc, bcs)  # reference for the first replay
        eager = [{k: v.clone() for k, v in d.items() if torch.is_tensor(v)} for d in eager]
        cap = ens.capture_step(start, t1, forc, bcs)
        first = cap.advance()
        same = all(torch.equal(first[m][k], eager[m][k]) for m in range(len(cases)) for k in eager[m])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            cap.advance()
        torch.cuda.synchronize()
        t_cap = time.perf_counter() - t0
        cap_stats = cap.stats()
        log(f"captured step x{n}: {t_cap:.2f} s (first replay bit-identical to the eager step: {same})")
        del cap, first, eager, start
    del lock, seq
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    adj = None
    try:
        th0 = [c.state["th_v"].detach().clone().requires_grad_(True) for c in cases]
        sts = []
        for c, t in zip(cases, th0):
            st = dict(c.state)
            st["th_v"] = t
            sts.append(st)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = ens.integrate(sts, 0.0, n, forc, bcs, checkpoint=True)
        J = sum((o["w"] ** 2).mean() for o in out)
        grads = torch.autograd.grad(J, th0)
        torch.cuda.synchronize()
        t_vg = time.perf_counter() - t0
        adj = {"t_value_and_grad_s": t_vg, "adjoint_forward_ratio": (t_vg - t_lock) / t_lock,
               "grad_norm": float(sum(g.norm() ** 2 for g in grads) ** 0.5),
               "peak_mem_GB": torch.cuda.max_memory_allocated() / 1e9}
        log(f"forward + reverse {n} steps: {t_vg:.2f} s")
    except Exception as e:  # report, do not hide
        adj = {"error": repr(e)[:300]}
    out = {"metric": "batched ensemble steps/s (lockstep, one launch per kernel over members)",
           "value": n / t_lock, "unit": "ensemble steps/s", "higher_is_better": True,
           "config": {"workload": f"{len(cases)} members, seeded nx=ny in 96..192, nz in 48..80, dx=1km, "
                                  "theta noise 0.5 K, mixed precision", "steps": n,
                      "member_grids": [[c.grid.nx, c.grid.ny, c.grid.nz] for c in cases]},
           "grid_point_updates_per_s": pts * n / t_lock,
           "t_lockstep_s": t_lock, "t_member_by_member_s": t_seq, "speedup_lockstep": t_seq / t_lock,
           "t_captured_step_s": t_cap, "captured_steps_per_s": n / t_cap,
           "speedup_captured_vs_lockstep": t_lock / t_cap, "captured_first_replay_bit_identical": same,
           "captured_solver_converged": all(s["residual"] <= s["ptol"] for s in cap_stats),
           "max_rel_diff_lockstep_vs_members": diff, "forward_and_adjoint": adj,
           "solver_stats_last": ens.last_stats[-1][:4] if ens.last_stats else None}
    print(json.dumps(out), flush=True)


def literal_variant(args):
    """The limited-area workload exactly as written: white-noise +-2 m/s
    winds and the Davies rim's class default outflow factor 0.01
    (boundaries.py:21,110-113).  Times mixed and FP64 steps and reports the
    mixed-vs-FP64 error after `args.literal_steps` steps (the headline uses
    smooth winds and equal rim weights, whose error stays smooth)."""
    from suetes_amd import workloads as W
    from suetes_amd.cases import build_case
    cfg = W.lam(nx=args.nx, ny=args.nx, nz=args.nz, winds="white", outflow_factor=0.01)
    c64 = build_case(cfg, device="cuda", precision="fp64")
    cmx = build_case(cfg, device="cuda", precision="mixed")
    log("literal workload (white winds, outflow factor 0.01): timing")
    t_mx, s_mx = time_steps(cmx, args.steps, args.warmup)
    t_64, s_64 = time_steps(c64, args.steps, args.warmup)
    done = args.warmup + args.steps
    errs = {}
    a, b = s_64, s_mx
    for i in range(done, args.literal_steps):
        a = c64.stepper.step(a, i * c64.dt, c64.forcing, c64.bc_fn)
        b = cmx.stepper.step(b, i * cmx.dt, cmx.forcing, cmx.bc_fn)
        if (i + 1) % 20 == 0 or i + 1 == args.literal_steps:
            errs[i + 1] = fields_err(b, a)
    umax = float(a["u"].abs().max())
    log(f"literal: mixed {t_mx*1e3:.1f} ms, fp64 {t_64*1e3:.1f} ms, errors {errs}")
    return {"workload": f"lam {args.nx}x{args.nx}x{args.nz}, white U(-2,2) m/s winds, Davies rim 10 with "
                        "outflow_factor 0.01 (the class default)",
            "ms_per_step_mixed": t_mx * 1e3, "ms_per_step_fp64": t_64 * 1e3,
            "speedup_mixed_vs_fp64": t_64 / t_mx,
            "max_rel_l2_mixed_vs_fp64_by_step": errs, "fp64_max_abs_u_at_end": umax}


def acoustic_substep_bytes(nx, ny, nz):
    """Algorithmic HBM bytes of one acoustic substep (csrc/acoustic.hip, FP64):
    every array each of the four kernels must read or write once.
    predictor: pi, rz, 5 u-face and 5 v-face inputs, 2 + 2 outputs;
    column_rhs: 4 mass, 7 w-level, 2 + 2 face inputs, 2 outputs;
    column: 5 mass inputs, 2 outputs; corrector: 2 mass, 7 + 7 face and
    5 w-level inputs, 3 outputs."""
    nm, nu, nv, nw = nx * ny * nz, (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
    return 8 * (15 * nm + 17 * nu + 17 * nv + 15 * nw)


def split_explicit_leg(args, steps=3, warmup=1):
    """The split-explicit core (steppers.py:1076-1737) on the limited-area
    grid (dt 6 s, ns 6: 2 + 3 + 6 acoustic substeps of 1 s per step): FP64
    ms per step, and the acoustic substep (four kernels behind one C call)
    timed with HIP events on its stream against its algorithmic bytes."""
    import torch
    from suetes_amd import workloads as W
    from suetes_amd.cases import build_case
    from suetes_amd.steppers import state_to_internal
    cfg = W.lam(nx=args.nx, ny=args.nx, nz=args.nz)
    cfg["core"] = dict(cfg["core"], dt=6.0, ns=6)
    case = build_case(cfg, device="cuda", core_type="split-explicit")
    log("split-explicit: timing")
    t, s = time_steps(case, steps, warmup)
    st = case.stepper
    kern = st._kern
    ist = state_to_internal(s)
    with torch.no_grad():
        bg = st.bg_precomputed()
        ist["eta_dot"], ist["w"] = st._diagnose_eta_dot(ist)
        slow = st._compute_slow_tendencies(ist, bg, {}, {}, None)
        args_loop = (ist["u"], ist["v"], ist["w"], ist["pi"], ist["u"], ist["v"], ist["w"],
                     slow["u"], slow["v"], slow["w"], slow["pi"])
        n = 20
        dev0 = [torch.zeros_like(x) for x in args_loop[:4]]
        kern.loop(st.dtau_stage3, 2, *dev0, *args_loop[4:])
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        kern.loop(st.dtau_stage3, n, *dev0, *args_loop[4:])
        e1.record()
        torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / n
    nb = acoustic_substep_bytes(args.nx, args.nx, args.nz)
    log(f"split-explicit {t*1e3:.1f} ms/step; acoustic substep {us:.0f} us, {nb / us / 1e3:.0f} GB/s")
    return {"workload": f"lam {args.nx}x{args.nx}x{args.nz}, split-explicit dt 6 s, ns 6 (2+3+6 substeps), "
                        "FP64", "ms_per_step": t * 1e3, "steps": steps, "warmup": warmup,
            "acoustic_substep": {"avg_us": us, "algorithmic_bytes": nb, "achieved_GBs": nb / us / 1e3,
                                 "frac_of_hbm_peak": nb / us / 1e3 / HBM_PEAK_GBS,
                                 "method": "HIP events around 20 substeps of one C call on its stream"},
            "max_abs_w_at_end": float(s["w"].abs().max())}


def trace_summary(path, step=2):
    """Per-kernel breakdown of one step from a rocprofv3 kernel-trace CSV of
    `bench.py --steps 2 --warmup 1 ...` (steps are delimited by the first of
    the four departure-point launches of each step; step 0 is the mixed
    path's warm-up with the autotuning, steps 1-2 its timed steps).  Prints a
    markdown table."""
    import csv
    import re
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()

    codes = {"DF16_": "half", "f": "float", "d": "double", "t": "bf16"}

    def short(name):
        n = re.sub(r"\(.*", "", name).replace("void ", "").replace("unsigned short", "bf16")
        m = re.match(r"_ZN4sisl\d+(\w+?)I((?:DF16_|[fdt]|Li\d+E)+)E", n)
        if m:  # a mangled instantiation: decode its template arguments
            toks = re.findall(r"DF16_|Li\d+E|[fdt]", m.group(2))
            return "sisl::%s<%s>" % (m.group(1), ", ".join(
                codes.get(t, t[2:-1]) for t in toks))
        if n.startswith("sisl::"):
            return n
        if "at::native" in name or "copyBuffer" in name:
            return "torch elementwise / copy"
        return n[:40]
    starts = [i for i, r in enumerate(rows) if "index_velocity_kernel" in r[2]][::4]
    a, b = starts[step], starts[step + 1]
    win = rows[a:b]
    span = (win[-1][1] - win[0][0]) / 1e6
    tot = {}
    for s0, s1, n in win:
        k = short(n)
        c, t = tot.get(k, (0, 0.0))
        tot[k] = (c + 1, t + (s1 - s0) / 1e6)
    busy = sum(t for _, t in tot.values())
    print(f"step {step}: {len(win)} launches, span {span:.1f} ms, kernels busy {busy:.1f} ms")
    print("| kernel | launches | total ms | avg us |")
    print("|---|---:|---:|---:|")
    for k, (c, t) in sorted(tot.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{k}` | {c} | {t:.2f} | {1e3 * t / c:.1f} |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--nz", type=int, default=80)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--err-steps", type=int, default=100,
                    help="continue both paths to this many steps and report the field error "
                         "(0: skip)")
    ap.add_argument("--literal-steps", type=int, default=100,
                    help="also time the workload as written (white +-2 m/s winds, Davies outflow "
                         "factor 0.01) and report its mixed-vs-FP64 error after this many steps "
                         "(0: skip)")
    ap.add_argument("--terrain-gradient", action="store_true",
                    help="run the terrain-control reverse-mode case instead (128x128x60, 50 steps)")
    ap.add_argument("--adj-steps", type=int, default=None)
    ap.add_argument("--ensemble", action="store_true",
                    help="run the batched ensemble case instead (32 uneven members)")
    ap.add_argument("--members", type=int, default=32)
    ap.add_argument("--ens-steps", type=int, default=20)
    ap.add_argument("--no-split-explicit", action="store_true",
                    help="skip the split-explicit core's timing")
    ap.add_argument("--split-explicit-only", action="store_true",
                    help="only time the split-explicit core (the run its profiles come from)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only launch the operator kernel and a plain copy of known size "
                         "(the run that hardware-counter passes profile)")
    ap.add_argument("--trace-summary", default=None,
                    help="print the per-kernel breakdown of one mixed step from a rocprofv3 kernel-trace CSV "
                         "(of a --steps 2 --warmup 1 run) and exit")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from suetes_amd import workloads as W
    from suetes_amd.cases import build_case

    if args.ensemble:
        ensemble_run(args)
        return
    if args.terrain_gradient:
        if args.nx == 256 and args.nz == 80:
            args.nx, args.nz = 128, 60
        terrain_gradient(args)
        return
    if args.split_explicit_only:
        out = {"split_explicit": split_explicit_leg(args), "copy_bytes": copy_reference()}
        print(json.dumps(out), flush=True)
        return
    cfg = W.lam(nx=args.nx, ny=args.nx, nz=args.nz)
    log(f"building workload lam {args.nx}x{args.nx}x{args.nz}")
    if args.kernels_only:
        cmx = build_case(cfg, device="cuda", precision="mixed")
        st = run_steps(cmx, 1)
        roof, table = instrumented_step(cmx, st)
        nbytes = copy_reference()
        print(json.dumps({"roofline": roof, "kernels": table, "copy_bytes": nbytes}), flush=True)
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
    err_long = None
    done = args.warmup + args.steps
    if args.err_steps > done:
        a, b = s_64, s_mx
        for i in range(done, args.err_steps):
            a = c64.stepper.step(a, i * c64.dt, c64.forcing, c64.bc_fn)
            b = cmx.stepper.step(b, i * cmx.dt, cmx.forcing, cmx.bc_fn)
        err_long = {"steps": args.err_steps, "max_rel_l2": fields_err(b, a),
                    "per_field": {k: float((b[k].double() - a[k].double()).norm() / a[k].double().norm())
                                  for k in ("u", "v", "w", "pi", "th_v", "rho")}}
        log(f"mixed vs fp64 after {args.err_steps} steps: {err_long['max_rel_l2']:.2e}")
    roof, table = instrumented_step(cmx, s_mx)
    roof["attainable"] = attainable(cmx)
    if roof["kernel"].startswith("helmholtz_kernel on the FP16") and roof["kernel"].endswith("[fp32]"):
        roof["frac_of_access_mix"] = roof["achieved"] / roof["attainable"]["operator_access_mix_GBs"]
    elif roof["kernel"].startswith("helmholtz_kernel") and roof["kernel"].endswith("[fp32]"):
        roof["frac_of_access_mix"] = roof["achieved"] / roof["attainable"]["operator_access_mix_fused_GBs"]
    log(f"dominant kernel {roof['kernel']}: {roof['achieved']:.0f} GB/s; attainable: copy "
        f"{roof['attainable']['copy_GBs']:.0f}, its access mix {roof['attainable']['operator_access_mix_GBs']:.0f}")
    adj = None
    if not args.no_adjoint:
        try:
            r, tf, tvg = adjoint_cost(cmx, 1)
            r64, tf64, tvg64 = adjoint_cost(c64, 1)
            adj = {"ratio": r, "t_forward_s": tf, "t_value_and_grad_s": tvg,
                   "fp64": {"ratio": r64, "t_forward_s": tf64, "t_value_and_grad_s": tvg64},
                   "speedup_value_and_grad_mixed_vs_fp64": tvg64 / tvg}
            log(f"adjoint/forward {r:.2f} (fp64 {r64:.2f}); value+grad speedup {tvg64 / tvg:.2f}")
            adj["kernels_value_and_grad_mixed"] = instrumented_adjoint(cmx)
        except Exception as e:  # report, do not hide
            adj = {"error": repr(e)[:300]}
    literal = None
    if args.literal_steps > 0:
        literal = literal_variant(args)
    split = None if args.no_split_explicit else split_explicit_leg(args)
    cpu = None if args.no_cpu else cpu_baseline()
    npts = args.nx * args.nx * args.nz
    out = {
        "metric": METRIC, "value": 1.0 / t_mx, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_mx * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f16 Krylov basis, bf16 operator coefficients; f64 refinement)",
        "data": "synthetic",
        "precision": "FP32 arithmetic in the operator, preconditioner, departure points and tricubic; "
                     "storage: FP16 inner Krylov basis (scaled by sqrt(n); FP64 inner products), bfloat16 "
                     "3-D coefficient arrays of the inner operator and preconditioner (forward solves; "
                     "transposed solves read FP32); FP64 outer residual refinement, state and "
                     "explicit tendencies",
        "config": {"workload": f"lam {args.nx}x{args.nx}x{args.nz} dx=1km dt=10s stretched-SLEVE "
                               "GRF terrain, +-2 m/s winds (10 km correlation), Davies rim 10 "
                               "(equal inflow/outflow weights), GMRES restart 20 tol 1e-10",
                   "grid": [args.nx, args.nx, args.nz], "dt": 10.0},
        "fp64_steps_per_s": 1.0 / t_64, "fp64_ms_per_step": t_64 * 1e3,
        "speedup_mixed_vs_fp64": t_64 / t_mx,
        "field_err_mixed_vs_fp64": err,
        "field_err_mixed_vs_fp64_long": err_long,
        "grid_point_updates_per_s": npts / t_mx,
        "solver_stats_mixed": st_mx, "solver_stats_fp64": st_64,
        "adjoint_forward_cost": adj,
        "roofline": roof, "kernels_mixed_step": table, "cpu_baseline": cpu,
        "literal_config": literal,
        "split_explicit": split,
        "operator_launch_shapes": operator_shapes(cmx),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
