"""Shared builders: the same workload on the oracle (NumPy) and on the product."""
from __future__ import annotations

import numpy as np
import torch

from oracle import sisl as O
from suetes_amd import workloads as W

C = W.CONSTANTS


def oracle_transform(spec):
    kind, kw = spec
    return {"galchen": O.GalChenSigma, "sleve": O.SleveSimple,
            "stretched_sleve": O.StretchedSleveSimple}[kind](**kw)


def oracle_case(cfg, physics_suite=None, extra_fields=None, core_type="sisl"):
    gk = cfg["grid"]
    terrain = W.make_terrain(cfg["terrain"], gk)
    g = O.Grid(gk["nx"], gk["ny"], gk["nz"], gk["dx"], gk["dy"], gk["dz"], gk["lat_center"],
               gk["lon_center"], h_func=terrain, transform=oracle_transform(cfg["transform"]))
    op = O.Ops(g, cfg["periodic_axes"])
    st = W.initial_state(cfg, g.Z_m, g.x_m, g.y_m)
    st.update({k: v.copy() for k, v in (extra_fields or {}).items()})
    ext = {k: v.copy() for k, v in st.items()}
    stepper, dt = O.build_dynamical_core(core_type, g, op, C, st, physics_suite=physics_suite, **cfg["core"])
    bc = (lambda s, f: s)
    if cfg["boundary"] is not None and cfg["boundary"][0] == "davies":
        sp = O.DaviesSponge(g, op, **cfg["boundary"][1])
        bc = sp.blend
    elif cfg["boundary"] is not None and cfg["boundary"][0] == "xrelax":
        bc = O.XRelaxation(gk["nx"], **cfg["boundary"][1]).blend
    return g, op, stepper, st, bc, ext


def small_lam(nx=20, ny=18, nz=12, dx=1000.0, top=8000.0, seed=0):
    cfg = W.lam(nx=nx, ny=ny, nz=nz, dx=dx, top=top, seed=seed, steps=4)
    cfg["core"]["damp_height"] = 5000.0
    cfg["boundary"] = ("davies", dict(sponge_depth=4, dt=10.0))
    return cfg


def small_bubble(n=16):
    cfg = W.bubble(n=n, dx=50.0)
    return cfg


def small_schaer(nx=40, nz=14):
    """Mountain-wave slice: periodic in y (ny = 3) only, x relaxation."""
    return W.schaer(nx=nx, nz=nz, dx=500.0, top=10000.0, steps=4)


def to_np_ref(t):
    """Internal (Z,Y,X) or reference-view tensor -> reference-shaped NumPy."""
    return t.detach().cpu().numpy()


def rel_l2(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)
