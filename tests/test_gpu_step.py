"""GPU parity of the complete SISL step against the FP64 oracle, the mixed
precision path against the FP64 path, and the reverse mode.

Bars: FP64 fields after each step rel-L2 <= 1e-12 vs the oracle (both sides
solve GMRES to 1e-13); mixed path rel-L2 <= 1e-5 vs FP64; gradients pass a
Taylor test (second-order remainder) and central finite differences.
"""
import numpy as np
import pytest
import torch

from tests.helpers import oracle_case, rel_l2, small_bubble, small_lam

pytestmark = pytest.mark.gpu
DEV = "cuda"
FIELDS = ("u", "v", "w", "pi", "th_v", "rho", "eta_dot")


def build(cfg, **kw):
    from suetes_amd.cases import build_case
    return build_case(cfg, device=DEV, **kw)


def tight(cfg):
    cfg["core"].update(solver_tol=1e-14, solver_maxiter=60, solver_restart=20)
    return cfg


def np_state(st):
    return {k: st[k].detach().cpu().numpy() for k in FIELDS}


@pytest.mark.parametrize("cfgf,nsteps", [(small_lam, 3), (small_bubble, 2)])
def test_fp64_step_matches_oracle(cfgf, nsteps):
    cfg = tight(cfgf())
    case = build(cfg)
    go, opo, sto, s_o, bc_o, ext_o = oracle_case(cfg)
    s_p = case.state
    for i in range(nsteps):
        s_o = sto.step(s_o, i * case.dt, ext_o, bc_o)
        s_p = case.stepper.step(s_p, i * case.dt, case.forcing, case.bc_fn)
        got = np_state(s_p)
        for k in FIELDS:
            scale = np.linalg.norm(s_o[k])
            if scale == 0.0:
                assert np.max(np.abs(got[k])) < 1e-14
                continue
            e = rel_l2(got[k], s_o[k])
            if k in ("pi", "th_v", "rho"):  # also the perturbation about the background
                ref = sto.physics.pi_bg if k == "pi" else sto.physics.theta_bg if k == "th_v" else None
                if ref is not None:
                    dp = np.linalg.norm(s_o[k] - ref)
                    if dp > 0:
                        e = max(e, np.linalg.norm(got[k] - s_o[k]) / dp * 1e-3)
            assert e < 1e-12, (i, k, e)


def test_resting_mountain_stays_at_rest_on_gpu():
    cfg = small_lam()
    case = build(cfg)
    st = dict(case.state)
    ph = case.stepper.physics
    for k in ("u", "v"):
        st[k] = torch.zeros_like(st[k])
    st["th_v"], st["pi"] = ph.theta_bg, ph.pi_bg
    c = ph.c
    st["rho"] = c["p0"] / (c["Rd"] * st["th_v"]) * st["pi"] ** (c["cvd"] / c["Rd"])
    for i in range(5):
        st = case.stepper.step(st, i * case.dt, None, lambda s, f: s)
    assert float(st["w"].abs().max()) < 1e-7


def test_mixed_step_tracks_fp64():
    cfg = small_lam()
    c64 = build(cfg)
    c32 = build(cfg, precision="mixed")
    a, b = c64.state, c32.state
    for i in range(6):
        a = c64.stepper.step(a, i * c64.dt, c64.forcing, c64.bc_fn)
        b = c32.stepper.step(b, i * c32.dt, c32.forcing, c32.bc_fn)
    for k in ("u", "v", "w", "pi", "th_v"):
        e = rel_l2(b[k].cpu().numpy(), a[k].cpu().numpy())
        assert e < 1e-5, (k, e)


def _objective(case, nsteps, coords=None):
    """J = 0.5 sum(rho w_m^2 + 0.1 theta'^2) after nsteps, as a function of the
    initial theta (verification/adjoint_gradient_convergence.py:75-125).
    ``coords``: per-step departure points to replay (frozen trajectories)."""
    from suetes_amd.steppers import state_to_internal
    ph = case.stepper.physics

    def J(th0):
        st = dict(case.state)
        st["th_v"] = th0
        st = state_to_internal(st)
        for i in range(nsteps):
            st = case.stepper.step_internal(st, i * case.dt, case.forcing, case.bc_fn,
                                            coords=None if coords is None else coords[i])
        wm = 0.5 * (st["w"][:-1] + st["w"][1:])
        thp = st["th_v"] - ph.i_theta_bg
        return 0.5 * (st["rho"] * wm ** 2 + 0.1 * thp ** 2).sum()
    return J


def _direction(case):
    g = case.grid
    x = g.x_m[:, None, None]
    z = g.z_m[None, None, :]
    d = torch.cos(2 * np.pi * x / (g.nx * g.dx)) * torch.cos(np.pi * z / (g.nz * g.dz)) * torch.ones(
        1, g.ny, 1, dtype=torch.float64, device=x.device)
    return d / d.pow(2).mean().sqrt()


def _grad(J, th0, d):
    t = th0.detach().clone().requires_grad_(True)
    val = J(t)
    (g,) = torch.autograd.grad(val, t)
    return float(val.detach()), float((g * d).sum())


def _case_fd():
    cfg = small_lam(nx=12, ny=10, nz=8)
    cfg["core"].update(solver_tol=1e-13, solver_maxiter=40)
    # The Davies rim switches its relaxation weight on the sign of the
    # (stop-gradient) boundary winds (boundaries.py:106-117): J is not
    # differentiable where a wind crosses zero, and a finite-difference step
    # can cross it.  An outflow factor of 1 makes the weights sign-free.
    cfg["boundary"] = ("davies", dict(sponge_depth=4, dt=10.0, outflow_factor=1.0))
    case = build(cfg)
    case.op.use_stop_grad = False  # differentiate the padded edge values too
    return case


def test_gradient_one_step_fd_and_taylor():
    """One step: departure points do not depend on theta, so with edge
    padding differentiated the adjoint is the exact derivative."""
    case = _case_fd()
    J = _objective(case, 1)
    th0 = case.state["th_v"].detach().clone()
    d = _direction(case)
    val, dJ = _grad(J, th0, d)
    rem = []
    with torch.no_grad():
        for e in (0.4, 0.2, 0.1):
            rem.append(abs(float(J(th0 + e * d)) - val - e * dJ))
        h = 1e-3
        fd = (float(J(th0 + h * d)) - float(J(th0 - h * d))) / (2 * h)
    order = [np.log2(rem[i] / rem[i + 1]) for i in range(2)]
    assert min(order) > 1.9, (rem, order)
    assert abs(fd - dJ) <= 1e-6 * abs(dJ), (fd, dJ)


def test_gradient_three_steps_frozen_trajectories_fd():
    """Three steps with the departure points of the base run replayed: the
    adjoint (tricubic transposes, transposed GMRES, autograd glue) matches
    central finite differences to 1e-6."""
    case = _case_fd()
    from suetes_amd.steppers import state_to_internal
    case.stepper.record_coords = []
    st = state_to_internal(dict(case.state))
    for i in range(3):
        st = case.stepper.step_internal(st, i * case.dt, case.forcing, case.bc_fn)
    coords = case.stepper.record_coords
    case.stepper.record_coords = None
    J = _objective(case, 3, coords)
    th0 = case.state["th_v"].detach().clone()
    d = _direction(case)
    val, dJ = _grad(J, th0, d)
    h = 1e-3
    with torch.no_grad():
        fd = (float(J(th0 + h * d)) - float(J(th0 - h * d))) / (2 * h)
    assert abs(fd - dJ) <= 1e-6 * abs(dJ), (fd, dJ)


def test_checkpointed_integrate_gives_the_same_gradient():
    cfg = small_lam(nx=12, ny=10, nz=8)
    case = build(cfg)
    grads = []
    for ck in (False, True):
        u0 = case.state["u"].detach().clone().requires_grad_(True)
        st = dict(case.state)
        st["u"] = u0
        out = case.stepper.integrate(st, 0.0, 3, case.forcing, case.bc_fn, checkpoint=ck)
        (g,) = torch.autograd.grad((out["w"] ** 2).sum() + (out["th_v"] ** 2).sum() * 1e-3, u0)
        grads.append(g)
    assert torch.equal(grads[0], grads[1])


def test_step_is_reproducible():
    cfg = small_lam()
    case = build(cfg)
    outs = []
    for _ in range(2):
        st = case.stepper.step(dict(case.state), 0.0, case.forcing, case.bc_fn)
        outs.append(torch.cat([st[k].reshape(-1) for k in FIELDS]))
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("name", ["bubble", "schaer"])
def test_bubble_and_schaer_fp64_match_oracle(name):
    """The warm bubble (doubly periodic, flat) on 32^3 and the Schaer slice
    (y-periodic, 3 rows) on 200x3x50, with the workloads' own physics: one
    FP64 step against the oracle, rel-L2 <= 1e-12 (the oracle needs ~10 s
    per step at these sizes)."""
    from suetes_amd import workloads as W
    cfg = tight(W.bubble(n=32) if name == "bubble" else W.schaer(nx=200, nz=50))
    case = build(cfg)
    go, opo, sto, s_o, bc_o, ext_o = oracle_case(cfg)
    s_o = sto.step(s_o, 0.0, ext_o, bc_o)
    got = np_state(case.stepper.step(case.state, 0.0, case.forcing, case.bc_fn))
    # velocity components are measured against the whole velocity field: in
    # the 3-row slice v is solver noise (~1e-9 m/s beside u ~ 10 m/s)
    vel = np.sqrt(sum(np.linalg.norm(s_o[k]) ** 2 for k in ("u", "v", "w")))
    for k in FIELDS:
        scale = vel if k in ("u", "v", "w") else np.linalg.norm(s_o[k])
        if scale == 0.0:
            assert np.max(np.abs(got[k])) < 1e-14, k
            continue
        e = np.linalg.norm(got[k] - s_o[k]) / scale
        assert e < 1e-12, (k, e)


@pytest.mark.parametrize("name,nsteps", [("lam", 100), ("schaer", 100)])
def test_mixed_within_1e5_of_fp64_after_100_steps(name, nsteps):
    """Mixed path (FP32 kernels, FP16 Krylov basis, FP64 refinement) against
    the FP64 path after 100 steps: every field rel-L2 <= 1e-5 (velocity
    components against the whole velocity field, as v is rounding noise in
    the 3-row slice)."""
    from suetes_amd import workloads as W
    cfg = small_lam(nx=48, ny=40, nz=24) if name == "lam" else W.schaer(nx=200, nz=50)
    if name == "lam":  # rim without the sign switch, as the limited-area workload
        cfg["boundary"] = ("davies", dict(sponge_depth=4, dt=10.0, outflow_factor=1.0))
    c64 = build(cfg)
    cmx = build(cfg, precision="mixed")
    a, b = c64.state, cmx.state
    for i in range(nsteps):
        a = c64.stepper.step(a, i * c64.dt, c64.forcing, c64.bc_fn)
        b = cmx.stepper.step(b, i * cmx.dt, cmx.forcing, cmx.bc_fn)
    vel = float(sum(a[k].double().norm() ** 2 for k in ("u", "v", "w")) ** 0.5)
    for k in ("u", "v", "w", "pi", "th_v", "rho"):
        scale = vel if k in ("u", "v", "w") else float(a[k].double().norm())
        e = float((b[k].double() - a[k].double()).norm()) / scale
        assert e < 1e-5, (k, e)


def _terrain_cfg():
    from suetes_amd import workloads as W
    cfg = W.adjoint(nx=24, ny=20, nz=12)
    cfg["core"].update(solver_tol=1e-13, solver_maxiter=40)
    return cfg


def test_terrain_gradient_fd_taylor_and_mixed():
    """Gradient of J = mean w^2 over a downstream box after 3 steps with
    respect to the 8 bump amplitudes (grid, background, initial state and
    operator all depend on them): adjoint vs central differences (1e-6
    relative) and a Taylor test, with replayed departure points and
    differentiated edge padding; the mixed path's gradient within 1e-4 of
    FP64's."""
    from suetes_amd.cases import terrain_objective
    cfg = _terrain_cfg()
    a0 = torch.tensor(cfg["terrain"][1]["amps"], dtype=torch.float64, device=DEV)
    rec = []
    with torch.no_grad():
        terrain_objective(cfg, a0, 3, record=rec)

    def J(a, precision="fp64"):
        return terrain_objective(cfg, a, 3, coords=rec, exact_edges=True, precision=precision)

    a = a0.clone().requires_grad_(True)
    val = J(a)
    (g,) = torch.autograd.grad(val, a)
    d = torch.linspace(-1.0, 1.0, 8, dtype=torch.float64, device=DEV)
    d = d / d.norm()
    dJ = float((g * d).sum())
    h = 0.05
    with torch.no_grad():
        fd = (float(J(a0 + h * d)) - float(J(a0 - h * d))) / (2 * h)
    assert abs(fd - dJ) <= 1e-6 * abs(fd), (fd, dJ)
    rem = []
    with torch.no_grad():
        for eps in (20.0, 10.0, 5.0):
            rem.append(abs(float(J(a0 + eps * d)) - float(val) - eps * dJ))
    orders = [np.log2(rem[i] / rem[i + 1]) for i in range(2)]
    assert min(orders) > 1.8, (rem, orders)
    am = a0.clone().requires_grad_(True)
    (gm,) = torch.autograd.grad(J(am, "mixed"), am)
    assert float((gm - g).norm() / g.norm()) < 1e-4
