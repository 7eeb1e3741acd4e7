"""Adjoint of the limited tricubic gather (operators.py:146-209 with
use_limiter=True): clip of the interpolant to the chained min / max of the 8
cell corners, differentiated with the balanced tie rule of JAX's
minimum / maximum (each operand equal to the result gets half the
cotangent; jax.lax.min / max, the rule behind jnp.clip).  JAX is not
importable here, so the tie rule is pinned by its closed form on a field
of exact ties rather than against JAX output.

Bars:
* central differences of the HIP forward vs the adjoint along random
  directions (the limited gather is piecewise linear in the field, so the
  difference is exact away from branch switches; FP64): relative 1e-7,
  with the limiter active at a large share of the points;
* a zero field (every clip and every corner comparison ties): the
  cotangent of one output splits 1/4 to the interpolant's 64 taps
  (Catmull-Rom weights) and 3/4 to the corners with the chain weights
  2^-7 (corner 0) and 2^-(8-q) (corner q >= 1): 1e-15;
* two backward passes bit-identical; FP32 within 1e-5 of FP64;
* a full SISL step with use_limiter=True from a state without ties (seeded
  theta noise) matches central differences (one step, 1e-6).
"""
import numpy as np
import pytest
import torch

from oracle import sisl as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _coords(shape, seed, spread=2.5):
    nz, ny, nx = shape
    g = torch.Generator(device="cpu").manual_seed(seed)
    co = (torch.rand((3,) + shape, generator=g, dtype=torch.float64) * 2 - 1) * spread
    return co.to(DEV)


def _gather(f, co, lim):
    from suetes_amd.operators import Coords, tensor_product_interp_3d
    return tensor_product_interp_3d(f, Coords(co), use_limiter=lim)


@pytest.mark.parametrize("shape", [(8, 10, 12), (5, 9, 70)])
def test_limited_adjoint_matches_central_differences(shape):
    print(f"[case] limited tricubic adjoint vs FD, field {shape}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(3)
    # noise plus a sign-alternating pattern along x: the cubic overshoots its
    # cell corners often, so the clip is active at many points
    alt = torch.tensor([(-1.0) ** i for i in range(shape[2])], dtype=torch.float64)
    f = (torch.randn(shape, generator=g, dtype=torch.float64) + 2.0 * alt).to(DEV)
    co = _coords(shape, 4)
    y = torch.randn(shape, generator=g, dtype=torch.float64).to(DEV)
    with torch.no_grad():
        active = float(((_gather(f, co, True) - _gather(f, co, False)).abs() > 0).double().mean())
    print(f"  limiter active at {active:.1%} of the points", flush=True)
    assert active > 0.005, active
    fl = f.clone().requires_grad_(True)
    (gf,) = torch.autograd.grad((_gather(fl, co, True) * y).sum(), fl)
    for seed in (5, 6):
        d = torch.randn(shape, generator=torch.Generator(device="cpu").manual_seed(seed),
                        dtype=torch.float64).to(DEV)
        h = 1e-7
        with torch.no_grad():
            fd = float(((_gather(f + h * d, co, True) - _gather(f - h * d, co, True)) / (2 * h) * y).sum())
        adj = float((gf * d).sum())
        print(f"  fd {fd:.12e} adjoint {adj:.12e}", flush=True)
        assert abs(fd - adj) <= 1e-7 * abs(adj)
    # deterministic
    fl2 = f.clone().requires_grad_(True)
    (gf2,) = torch.autograd.grad((_gather(fl2, co, True) * y).sum(), fl2)
    assert torch.equal(gf, gf2)
    # FP32 (the mixed path's gathers)
    f32 = f.float().clone().requires_grad_(True)
    (g32,) = torch.autograd.grad((_gather(f32, co.float(), True) * y.float()).sum(), f32)
    assert float((g32.double() - gf).norm() / gf.norm()) < 1e-5


def test_limited_adjoint_tie_rule_closed_form():
    print("[case] limited tricubic adjoint on a field of ties", flush=True)
    shape = (7, 8, 9)
    nz, ny, nx = shape
    f = torch.zeros(shape, dtype=torch.float64, device=DEV, requires_grad=True)
    co = torch.zeros((3,) + shape, dtype=torch.float64, device=DEV)
    K, J, I = 3, 4, 5
    tx, ty, tz = 0.3, 0.55, 0.8
    co[0, K, J, I], co[1, K, J, I], co[2, K, J, I] = tx, ty, tz
    gout = torch.zeros(shape, dtype=torch.float64, device=DEV)
    gout[K, J, I] = 1.0
    (gf,) = torch.autograd.grad((_gather(f, co, True) * gout).sum(), f)
    # Catmull-Rom weights of the four taps, from the oracle's cubic through unit vectors
    def w(t):
        return np.array([O.cubic_weight(*np.eye(4)[q], t) for q in range(4)])
    wx, wy, wz = w(tx), w(ty), w(tz)
    want = np.zeros(shape)
    for c in range(4):
        for b in range(4):
            for a in range(4):
                want[K - 1 + c, J - 1 + b, I - 1 + a] += 0.25 * wz[c] * wy[b] * wx[a]
    chain = [2.0 ** -7] + [2.0 ** -(8 - q) for q in range(1, 8)]
    assert abs(sum(chain) - 1.0) < 1e-15
    for q in range(8):
        ox, oy, oz = q & 1, (q >> 1) & 1, (q >> 2) & 1
        want[K + oz, J + oy, I + ox] += 0.75 * chain[q]
    np.testing.assert_allclose(gf.cpu().numpy(), want, rtol=0, atol=1e-15)


def test_step_with_limiter_differentiates():
    from tests.test_gpu_step import _case_fd, _direction, _grad, _objective
    print("[case] one SISL step with use_limiter=True: adjoint vs FD", flush=True)
    case = _case_fd()
    case.stepper.use_limiter = True
    J = _objective(case, 1)
    # theta' = 0 exactly in this workload, so every limited gather of theta's
    # arrival field would sit on exact ties (where the balanced tie rule, not
    # a derivative, applies): start from seeded theta noise instead
    th_s = case.state["th_v"].detach()
    g = torch.Generator(device="cpu").manual_seed(8)
    th0 = th_s + 0.3 * torch.randn(th_s.shape, generator=g, dtype=torch.float64).to(th_s.device)
    d = _direction(case)
    val, dJ = _grad(J, th0, d)
    with torch.no_grad():
        for h in (1e-4, 1e-5):
            fd = (float(J(th0 + h * d)) - float(J(th0 - h * d))) / (2 * h)
            print(f"  h {h:.0e} fd {fd:.12e} adjoint {dJ:.12e}", flush=True)
    assert abs(fd - dJ) <= 1e-6 * abs(dJ), (fd, dJ)
