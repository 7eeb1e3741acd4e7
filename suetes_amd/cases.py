"""Build a ready-to-step workload on the device (grid, stepper, state, bc)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import geometry as G
from .boundaries import DaviesSponge, XRelaxation
from .operators import CGridOperator3D
from .steppers import build_dynamical_core
from .workloads import CONSTANTS, WORKLOADS, initial_state, make_terrain


def make_transform(spec):
    kind, kw = spec
    if kind == "galchen":
        return G.GalChenSigma()
    if kind == "sleve":
        return G.SleveSimple(**kw)
    if kind == "stretched_sleve":
        return G.StretchedSleveSimple(**kw)
    raise ValueError(kind)


def identity_bc(state, forcing):
    return state


identity_bc.suetes_internal = True
identity_bc.suetes_rho_independent = True


@dataclass
class Case:
    cfg: dict
    grid: object
    op: object
    stepper: object
    state: dict
    bc_fn: object
    forcing: object
    dt: float


def to_device_state(st_np, device):
    """Reference-shaped NumPy dict -> dict of reference-shaped views of
    internal (x-fastest) device tensors."""
    out = {}
    for k, v in st_np.items():
        if isinstance(v, np.ndarray) and v.ndim == 3:
            t = torch.from_numpy(np.ascontiguousarray(np.transpose(v, (2, 1, 0)))).to(device)
            out[k] = G.ref_view(t)
        else:
            out[k] = v
    return out


def build_case(cfg, device="cuda", precision="fp64", physics_suite=None, extra_fields=None,
               core_type="sisl", **core_overrides):
    """Grid, stepper, initial state and boundary function of workload
    ``cfg`` on ``device``.  ``extra_fields``: reference-shaped NumPy arrays
    added to the initial state (e.g. tracers of ``physics_suite``);
    ``core_type``: 'sisl' or 'split-explicit'."""
    if isinstance(cfg, str):
        cfg = WORKLOADS[cfg]()
    gk = cfg["grid"]
    terrain = make_terrain(cfg["terrain"], gk)
    grid = G.RegionalGrid3D(gk["nx"], gk["ny"], gk["nz"], gk["dx"], gk["dy"], gk["dz"], gk["lat_center"],
                            gk["lon_center"], h_func=terrain, transform=make_transform(cfg["transform"]),
                            device=device)
    op = CGridOperator3D(grid, cfg["periodic_axes"])
    st_np = initial_state(cfg, grid.Z_m, grid.x_m, grid.y_m)
    st_np.update(extra_fields or {})
    state = to_device_state(st_np, device)
    forcing = dict(state)
    core = dict(cfg["core"])
    core["precision"] = precision
    core.update(core_overrides)
    stepper, dt = build_dynamical_core(core_type, grid, op, CONSTANTS, state, physics_suite=physics_suite,
                                       **core)
    bc = identity_bc
    if cfg["boundary"] is not None and cfg["boundary"][0] == "davies":
        sponge = DaviesSponge(grid, op, **cfg["boundary"][1])
        bc = sponge.blend
    elif cfg["boundary"] is not None and cfg["boundary"][0] == "xrelax":
        bc = XRelaxation(grid, **cfg["boundary"][1]).blend
    return Case(cfg, grid, op, stepper, state, bc, forcing, dt)


def analytic_background_torch(Z_m, N_bv, c=CONSTANTS, theta0=300.0):
    """theta and hydrostatic pi of euler.py:86-112 on internal (Z, Y, X)
    heights, as torch ops (differentiable in the heights)."""
    g, cp = c["g"], c["cp"]
    if N_bv > 0.0:
        th = theta0 * torch.exp((N_bv ** 2 / g) * Z_m)
        pi_top = 1.0 + (g ** 2 / (cp * theta0 * N_bv ** 2)) * (torch.exp(-N_bv ** 2 * Z_m.max() / g) - 1.0)
    else:
        th = theta0 * torch.ones_like(Z_m)
        pi_top = 1.0 - (g / (cp * theta0)) * Z_m.max()
    dzz = Z_m[1:] - Z_m[:-1]
    thw = 0.5 * (th[1:] + th[:-1])
    dpi = -(g * dzz) / (cp * thw)
    csr = torch.flip(torch.cumsum(torch.flip(dpi, (0,)), dim=0), (0,))
    top = pi_top * torch.ones_like(Z_m[:1])
    return th, torch.cat([top - csr, top], dim=0)


def build_terrain_case(cfg, amps, device="cuda", precision="fp64", **core_overrides):
    """``build_case`` for a terrain-bump workload whose bump amplitudes are the
    torch tensor ``amps``: grid, background, initial state (the analytic
    background over the terrain at rest plus a uniform wind) and the stepper
    are built with torch ops, so anything computed from the trajectory is
    differentiable in ``amps`` (the pattern of the reference's
    experiments/gravity_wave_optimal_topography/run.py:77-150)."""
    from .workloads import BumpTerrain
    gk = cfg["grid"]
    kind, kw = cfg["terrain"]
    if kind != "bumps":
        raise ValueError("build_terrain_case needs a 'bumps' terrain")
    terrain = BumpTerrain(kw["centres"], kw["sigma"], amps)
    grid = G.RegionalGrid3D(gk["nx"], gk["ny"], gk["nz"], gk["dx"], gk["dy"], gk["dz"], gk["lat_center"],
                            gk["lon_center"], h_func=terrain, transform=make_transform(cfg["transform"]),
                            device=device)
    op = CGridOperator3D(grid, cfg["periodic_axes"])
    th, pi = analytic_background_torch(grid.i_Z_m, cfg["N_bv"])
    nx, ny, nz = gk["nx"], gk["ny"], gk["nz"]
    u0 = cfg["init"][1].get("u", 0.0) if cfg["init"][0] == "uniform" else 0.0
    kwf = dict(dtype=torch.float64, device=grid.device)
    rho = CONSTANTS["p0"] / (CONSTANTS["Rd"] * th) * pi ** (CONSTANTS["cvd"] / CONSTANTS["Rd"])
    st = {"u": torch.full((nz, ny, nx + 1), float(u0), **kwf), "v": torch.zeros((nz, ny + 1, nx), **kwf),
          "w": torch.zeros((nz + 1, ny, nx), **kwf), "eta_dot": torch.zeros((nz + 1, ny, nx), **kwf),
          "pi": pi, "th_v": th, "rho": rho}
    state = {k: G.ref_view(v) for k, v in st.items()}
    forcing = dict(state)
    core = dict(cfg["core"])
    core["precision"] = precision
    core.update(core_overrides)
    stepper, dt = build_dynamical_core("sisl", grid, op, CONSTANTS, state, **core)
    bc = identity_bc
    if cfg["boundary"] is not None and cfg["boundary"][0] == "xrelax":
        bc = XRelaxation(grid, **cfg["boundary"][1]).blend
    elif cfg["boundary"] is not None and cfg["boundary"][0] == "davies":
        bc = DaviesSponge(grid, op, **cfg["boundary"][1]).blend
    return Case(cfg, grid, op, stepper, state, bc, forcing, dt)


def terrain_objective(cfg, amps, nsteps, coords=None, precision="fp64", record=None, checkpoint=False,
                      exact_edges=False, device="cuda"):
    """J = mean of w^2 over ``cfg['objective_box']`` (reference (x, y, z)
    slices) after ``nsteps`` steps of the terrain-bump case with amplitudes
    ``amps`` (the objective of the reference's optimal-topography
    experiment, run.py:144-148, on a box).

    ``coords``: per-step departure points to replay (frozen trajectories);
    ``record``: a list that receives the departure points of this run;
    ``checkpoint``: one torch checkpoint per step (departure points kept);
    ``exact_edges``: differentiate the padded edge values as well (no
    stop-gradient on them), which finite differences see."""
    from torch.utils.checkpoint import checkpoint as ckpt
    from .steppers import state_to_internal
    case = build_terrain_case(cfg, amps, device=device, precision=precision)
    if exact_edges:
        case.op.use_stop_grad = False
    stp = case.stepper
    st = state_to_internal(case.state)
    ckpt_on = checkpoint and torch.is_grad_enabled()
    if ckpt_on:
        stp.warm_caches(case.forcing, case.bc_fn)
    stp.record_coords = None if ckpt_on else record  # checkpointed steps record below
    try:
        for i in range(int(nsteps)):
            co = coords[i] if coords is not None else None
            if ckpt_on:
                if co is None:
                    co = stp.departure_points(st)
                    if record is not None:
                        record.append(co)
                st = ckpt(stp.step_internal, st, i * case.dt, case.forcing, case.bc_fn, None, co,
                          use_reentrant=False)
            else:
                st = stp.step_internal(st, i * case.dt, case.forcing, case.bc_fn, coords=co)
    finally:
        stp.record_coords = None
    bx, by, bz = cfg["objective_box"]
    return (st["w"][bz, by, bx] ** 2).mean()
