"""Flux-form semi-Lagrangian sweeps (FluxFormAdvector.advect_1d) as one HIP
launch per sweep, against the tensor-operation path of the same method on
host tensors (reference steppers.py:552-594).

Bars: forward rel-L2 <= 1e-14 (FP64; summation order of the cumulative mass
differs) / 2e-6 (FP32); mass and Courant-number cotangents to the same bars,
including faces whose Courant number is exactly zero (the clamp's ties, as
eta_dot at the bottom and top) and Courant numbers above one; total mass
is conserved on closed and periodic lines."""
import numpy as np
import pytest
import torch

from suetes_amd.steppers import FluxFormAdvector

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


@pytest.mark.parametrize("dim", [0, 1, 2])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_ffsl_sweep_matches_tensor_path(dim, periodic, dtype):
    rng = np.random.default_rng(5 + dim + 3 * periodic)
    shape = [9, 11, 14]
    fshape = list(shape)
    fshape[dim] += 1
    mass = torch.from_numpy(rng.uniform(0.5, 2.0, shape)).to(dtype)
    cfl = torch.from_numpy(rng.uniform(-1.6, 1.6, fshape)).to(dtype)
    if not periodic:  # closed line: zero flux through the end faces (ties of the clamp)
        cfl.narrow(dim, 0, 1).zero_()
        cfl.narrow(dim, fshape[dim] - 1, 1).zero_()
    else:  # faces 0 and n are one physical face on a periodic line
        cfl.narrow(dim, fshape[dim] - 1, 1).copy_(cfl.narrow(dim, 0, 1))
    tol = 1e-14 if dtype == torch.float64 else 2e-6
    mc, cc = mass.clone().requires_grad_(True), cfl.clone().requires_grad_(True)
    mg, cg = mass.cuda().requires_grad_(True), cfl.cuda().requires_grad_(True)
    yc = FluxFormAdvector.advect_1d(mc, cc, dim, periodic)
    yg = FluxFormAdvector.advect_1d(mg, cg, dim, periodic)
    assert type(yg.grad_fn).__name__.startswith("_Ffsl")
    assert _rel(yg, yc) <= tol
    tot = float(mass.double().sum())
    assert abs(float(yg.detach().double().sum()) - tot) <= (1e-12 if dtype == torch.float64 else 1e-5) * tot
    w = torch.from_numpy(rng.standard_normal(shape)).to(dtype)
    gmc, gcc = torch.autograd.grad(yc, (mc, cc), w)
    gmg, gcg = torch.autograd.grad(yg, (mg, cg), w.cuda())
    assert _rel(gmg, gmc) <= 4 * tol
    assert _rel(gcg, gcc) <= 4 * tol
