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


def build_case(cfg, device="cuda", precision="fp64", **core_overrides):
    if isinstance(cfg, str):
        cfg = WORKLOADS[cfg]()
    gk = cfg["grid"]
    terrain = make_terrain(cfg["terrain"], gk)
    grid = G.RegionalGrid3D(gk["nx"], gk["ny"], gk["nz"], gk["dx"], gk["dy"], gk["dz"], gk["lat_center"],
                            gk["lon_center"], h_func=terrain, transform=make_transform(cfg["transform"]),
                            device=device)
    op = CGridOperator3D(grid, cfg["periodic_axes"])
    st_np = initial_state(cfg, grid.Z_m, grid.x_m, grid.y_m)
    state = to_device_state(st_np, device)
    forcing = dict(state)
    core = dict(cfg["core"])
    core["precision"] = precision
    core.update(core_overrides)
    stepper, dt = build_dynamical_core("sisl", grid, op, CONSTANTS, state, **core)
    bc = identity_bc
    if cfg["boundary"] is not None and cfg["boundary"][0] == "davies":
        sponge = DaviesSponge(grid, op, **cfg["boundary"][1])
        bc = sponge.blend
    elif cfg["boundary"] is not None and cfg["boundary"][0] == "xrelax":
        bc = XRelaxation(grid, **cfg["boundary"][1]).blend
    return Case(cfg, grid, op, stepper, state, bc, forcing, dt)
