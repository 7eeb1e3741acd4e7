"""Lateral relaxation zones used as the step's ``bc_fn``.

``DaviesSponge`` mirrors suetes/regional3d/boundaries.py:13-121 and
``BenchmarkSponge`` boundaries.py:123-189.  The weights are computed in
float32 with NumPy exactly as the reference computes them in float32.
``blend`` takes reference-shaped dicts (the ``bc_fn`` contract);
``blend_internal`` takes internal (Z, Y, X) dicts and is what the stepper
calls when it is handed ``sponge.blend``.
"""
from __future__ import annotations

import numpy as np
import torch

from .geometry import ref_view, to_internal


def davies_masks(nx, ny, depth, loc):
    """boundaries.py:36-61 (float32 weights)."""
    nxp = nx + 1 if loc == "u" else nx
    nyp = ny + 1 if loc == "v" else ny
    ix = np.arange(nxp, dtype=np.float32)
    iy = np.arange(nyp, dtype=np.float32)
    ox = np.float32(0.0 if loc == "u" else 0.5)
    oy = np.float32(0.0 if loc == "v" else 0.5)
    d = np.float32(depth)
    dw = np.clip((ix + ox) / d, 0.0, 1.0).astype(np.float32)
    de = np.clip((np.float32(nx) - (ix + ox)) / d, 0.0, 1.0).astype(np.float32)
    ds = np.clip((iy + oy) / d, 0.0, 1.0).astype(np.float32)
    dn = np.clip((np.float32(ny) - (iy + oy)) / d, 0.0, 1.0).astype(np.float32)
    hp = np.float32(0.5 * np.pi)

    def wgt(z):
        return (np.cos(hp * z) ** 2).astype(np.float32)
    return {"west": wgt(dw), "east": wgt(de), "south": wgt(ds), "north": wgt(dn)}


class DaviesSponge:
    # blend() computes no field from rho (the stepper may skip transporting
    # rho when the equation of state overwrites it)
    rho_independent = True

    def __init__(self, grid, operators, sponge_depth=30, dt=30.0, tau_bndy_factor=10.0,
                 outflow_factor=0.01):
        self.grid, self.op, self.depth = grid, operators, sponge_depth
        self.outflow_factor = outflow_factor
        self.max_c = dt / (tau_bndy_factor * dt)
        dev = grid.device
        self.masks = {}
        for loc in ("u", "v", "m"):
            m = davies_masks(grid.nx, grid.ny, sponge_depth, loc)
            # internal broadcast shapes: x weights (1,1,X), y weights (1,Y,1)
            self.masks[loc] = {
                "west": torch.from_numpy(m["west"]).to(dev).reshape(1, 1, -1),
                "east": torch.from_numpy(m["east"]).to(dev).reshape(1, 1, -1),
                "south": torch.from_numpy(m["south"]).to(dev).reshape(1, -1, 1),
                "north": torch.from_numpy(m["north"]).to(dev).reshape(1, -1, 1)}

    def get_interior_mask(self):
        out = {}
        for loc in ("u", "v", "m"):
            ms = self.masks[loc]
            comb = torch.maximum(torch.maximum(ms["west"], ms["east"]), torch.maximum(ms["south"], ms["north"]))
            inter = ref_view(1.0 - comb)
            if loc == "m":
                out["th_v"] = inter
                out["w"] = inter
            else:
                out[loc] = inter
        return out

    def blend_internal(self, st, ext):
        # boundaries.py:83-121
        op = self.op
        out = {}
        ub, vb = st["u"].detach(), st["v"].detach()
        for k in st.keys():
            if k in ("u", "v", "th_v", "q", "pi") and k in ext:
                m = self.masks.get(k, self.masks["m"])
                if k == "u":
                    ue, ve = ub, op.avg(op.avg(vb, 1, "v", "m"), 0, "m", "u")
                elif k == "v":
                    ue, ve = op.avg(op.avg(ub, 0, "u", "m"), 1, "m", "v"), vb
                else:
                    ue, ve = op.avg(ub, 0, "u", "m"), op.avg(vb, 1, "v", "m")
                one = torch.ones((), dtype=ue.dtype, device=ue.device)
                of = torch.full((), self.outflow_factor, dtype=ue.dtype, device=ue.device)
                ww = m["west"] * torch.where(ue > 0, one, of)
                we = m["east"] * torch.where(ue < 0, one, of)
                ws = m["south"] * torch.where(ve > 0, one, of)
                wn = m["north"] * torch.where(ve < 0, one, of)
                ct = torch.maximum(torch.maximum(ww, we), torch.maximum(ws, wn))
                out[k] = (1.0 - ct) * st[k] + ct * ext[k]
            else:
                out[k] = st[k]
        return out

    def blend(self, intermediate_state, external_state):
        st = {k: (to_internal(v) if torch.is_tensor(v) and v.dim() == 3 else v)
              for k, v in intermediate_state.items()}
        ex = {k: (to_internal(v) if torch.is_tensor(v) and v.dim() == 3 else v)
              for k, v in external_state.items()}
        out = self.blend_internal(st, ex)
        return {k: (ref_view(v) if torch.is_tensor(v) and v.dim() == 3 else v) for k, v in out.items()}


class XRelaxation:
    """Symmetric cos^2 relaxation over ``cells`` columns at both x ends toward
    the external state, for u, v, th_v, pi, rho: the boundary function of
    the reference's mountain-wave benchmark
    (benchmarks/physical/schaer_mountain_3d/run.py:103-126)."""
    rho_independent = True  # rho is blended on its own; no other field reads it

    def __init__(self, grid, cells=8):
        nx = grid.nx
        idx = np.arange(nx)
        dist = np.minimum(idx, nx - 1 - idx)
        m = np.where(dist < cells, np.cos(0.5 * np.pi * dist / cells) ** 2, 0.0)
        mu = np.concatenate([m, m[-1:]])
        dev = grid.device
        # internal layout: x is the last dimension
        self.mask = torch.from_numpy(m).to(dev).reshape(1, 1, -1)
        self.mask_u = torch.from_numpy(mu).to(dev).reshape(1, 1, -1)

    def blend_internal(self, st, ext):
        out = {}
        for k, v in st.items():
            if k in ("u", "v", "th_v", "pi", "rho") and k in ext:
                m = self.mask_u if k == "u" else self.mask
                out[k] = (1.0 - m) * v + m * ext[k]
            else:
                out[k] = v
        return out

    def blend(self, intermediate_state, external_state):
        st = {k: (to_internal(v) if torch.is_tensor(v) and v.dim() == 3 else v)
              for k, v in intermediate_state.items()}
        ex = {k: (to_internal(v) if torch.is_tensor(v) and v.dim() == 3 else v)
              for k, v in external_state.items()}
        out = self.blend_internal(st, ex)
        return {k: (ref_view(v) if torch.is_tensor(v) and v.dim() == 3 else v) for k, v in out.items()}


class BenchmarkSponge:
    """Symmetric relaxation (boundaries.py:123-189)."""
    rho_independent = True  # rho is blended on its own; no other field reads it

    def __init__(self, nx, ny=None, sponge_depth=10, axes=("x",), blend_vars=None, device=None):
        self.nx, self.ny, self.sponge_depth, self.axes = nx, ny, sponge_depth, axes
        if blend_vars is None:
            blend_vars = (["u", "v", "th_v", "pi", "rho", "q_tr"] if ny is None
                          else ["u", "v", "w", "th_v", "pi", "rho", "q_tr"])
        self.blend_vars = blend_vars

        def w1(n):
            idx = np.arange(n, dtype=np.float32)
            dist = np.minimum(idx, np.float32(n - 1) - idx)
            return np.where(dist < sponge_depth,
                            np.cos(np.float32(0.5 * np.pi) * dist / np.float32(sponge_depth)) ** 2,
                            0.0).astype(np.float32)
        wx = w1(nx) if "x" in axes else np.zeros(nx, np.float32)
        if ny is not None:
            wy = w1(ny) if "y" in axes else np.zeros(ny, np.float32)
            m2 = (1.0 - (1.0 - wx[:, None]) * (1.0 - wy[None, :])).astype(np.float32)
            mm = m2[:, :, None]
            mu = np.pad(mm, ((0, 1), (0, 0), (0, 0)), mode="edge")
            mv = np.pad(mm, ((0, 0), (0, 1), (0, 0)), mode="edge")
        else:
            mm = wx[:, None, None]
            mu = np.pad(mm, ((0, 1), (0, 0), (0, 0)), mode="edge")
            mv = mm
        dev = torch.device(device) if device is not None else None
        # internal layout: (1, Y, X)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(np.transpose(a, (2, 1, 0)))).to(dev)  # noqa
        self.mask_m, self.mask_u, self.mask_v, self.mask_w = t(mm), t(mu), t(mv), t(mm)

    def blend_internal(self, st, ext):
        out = {}
        for k in st.keys():
            if k in self.blend_vars and k in ext:
                m = {"u": self.mask_u, "v": self.mask_v, "w": self.mask_w}.get(k, self.mask_m)
                m = m.to(st[k].device)
                out[k] = (1.0 - m) * st[k] + m * ext[k]
            else:
                out[k] = st[k]
        return out

    def blend(self, state_in, ext_state):
        st = {k: (to_internal(v) if torch.is_tensor(v) and v.dim() == 3 else v) for k, v in state_in.items()}
        ex = {k: (to_internal(v) if torch.is_tensor(v) and v.dim() == 3 else v) for k, v in ext_state.items()}
        out = self.blend_internal(st, ex)
        return {k: (ref_view(v) if torch.is_tensor(v) and v.dim() == 3 else v) for k, v in out.items()}
