"""Staggered-grid averages and differences (CGridOperator3D.avg / diff) as
one HIP launch each, against the tensor-operation path of the same class on
host tensors (reference operators.py:37-132).

Bars: forward within 2 ulp-level rel-L2 (1e-15 FP64, 1e-6 FP32 -- the host
path divides by the spacing where the device multiplies by its reciprocal);
gradients (the gather-form transpose) to the same bars, with the reference's
stop-gradient on repeated edge values and without it."""
import numpy as np
import pytest
import torch

from suetes_amd import geometry as G
from suetes_amd.cases import make_terrain, make_transform
from suetes_amd.operators import CGridOperator3D
from tests.helpers import small_lam

pytestmark = pytest.mark.gpu

COMBOS = [(0, "m", "u"), (1, "m", "v"), (2, "m", "w"), (0, "u", "m"), (1, "v", "m"), (2, "w", "m")]


def _grids():
    cfg = small_lam(nx=14, ny=11, nz=9)
    gk = cfg["grid"]
    out = []
    for dev in ("cpu", "cuda"):
        out.append(G.RegionalGrid3D(gk["nx"], gk["ny"], gk["nz"], gk["dx"], gk["dy"], gk["dz"], gk["lat_center"],
                                    gk["lon_center"], h_func=make_terrain(cfg["terrain"], gk),
                                    transform=make_transform(cfg["transform"]), device=dev))
    return out


def _shape(g, loc):
    return {"m": (g.nz, g.ny, g.nx), "u": (g.nz, g.ny, g.nx + 1), "v": (g.nz, g.ny + 1, g.nx),
            "w": (g.nz + 1, g.ny, g.nx)}[loc]


def _rel(a, b):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


@pytest.mark.parametrize("periodic", [(), (0, 1)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("exact", [False, True])
def test_staggered_operators_match_tensor_path(periodic, dtype, exact):
    gc, gg = _grids()
    for g in (gc, gg):  # map factors in the working type
        g.i_m_factors = {k: v.to(dtype) for k, v in g.i_m_factors.items()}
    oc, og = CGridOperator3D(gc, periodic), CGridOperator3D(gg, periodic)
    oc.use_stop_grad = og.use_stop_grad = not exact
    tol = 1e-15 if dtype == torch.float64 else 1e-6
    rng = np.random.default_rng(3)
    for axis, fr, to in COMBOS:
        for name in ("avg", "diff"):
            f = torch.from_numpy(rng.standard_normal(_shape(gc, fr))).to(dtype)
            fc = f.clone().requires_grad_(True)
            fg = f.cuda().requires_grad_(True)
            yc = getattr(oc, name)(fc, axis, fr, to)
            yg = getattr(og, name)(fg, axis, fr, to)
            assert yg.grad_fn is not None and type(yg.grad_fn).__name__.startswith("_Stagger"), (axis, fr, to)
            assert yc.shape == yg.shape, (name, axis, fr, to)
            assert _rel(yg, yc) <= tol, (name, axis, fr, to, _rel(yg, yc))
            w = torch.from_numpy(rng.standard_normal(tuple(yc.shape))).to(dtype)
            (gcf,) = torch.autograd.grad(yc, fc, w)
            (ggf,) = torch.autograd.grad(yg, fg, w.cuda())
            assert _rel(ggf, gcf) <= 4 * tol, (name, axis, fr, to, _rel(ggf, gcf))
