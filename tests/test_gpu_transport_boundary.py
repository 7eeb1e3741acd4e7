"""GPU parity of the transport, boundary and solver-diagnostic pieces of the
step against the FP64 oracle, through the product's public calls:

* FluxFormAdvector.advect_3d_split (steppers.py:596-647) on the HIP sweeps
  vs the oracle (rel-L2 <= 1e-13, closed and periodic lateral axes);
* a full step carrying a passive tracer (the registration pattern of
  tests/test_core_formula_regressions.py:33-42), so the FFSL output of both
  rho and the tracer reaches compared fields (rel-L2 <= 1e-12 per field per
  step, 3 steps);
* SemiImplicitSolver3D.relative_residual (steppers.py:498-527) on device
  tensors vs the oracle's, and of the HIP GMRES solution;
* the state swapped in and out through DLPack (NumPy arrays, zero-copy on
  the host side) giving the oracle's step;
* GMRES edge cases: b = 0 returns x0 unchanged (JAX's NaN threshold), a
  non-finite right-hand side is reported as an error.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import sisl as O
from tests.helpers import oracle_case, rel_l2, small_lam

pytestmark = pytest.mark.gpu
DEV = "cuda"
FIELDS = ("u", "v", "w", "pi", "th_v", "rho", "eta_dot")


def _int(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0))).to(DEV)


def _np(t):
    return t.detach().cpu().numpy()


class Tracer:
    """Passive tracer registered through the physics-suite hooks
    (physics/base.py:28-68): no tendencies, no state updates."""
    tracer_keys = ["q_test"]

    def get_explicit_tendencies(self, *a, **k):
        return {}

    def apply_state_updates(self, s):
        return s


def tight(cfg):
    cfg["core"].update(solver_tol=1e-14, solver_maxiter=60, solver_restart=20)
    return cfg


@pytest.mark.parametrize("periodic", [(), (0, 1)])
def test_ffsl_advect_3d_split_matches_oracle(periodic):
    import suetes_amd as S
    from suetes_amd import workloads as W
    from suetes_amd.cases import make_transform
    from tests.helpers import C, oracle_transform
    print(f"[case] ffsl 3-D split vs oracle, periodic={periodic}", flush=True)
    cfg = small_lam()
    gk = cfg["grid"]
    terr = W.make_terrain(cfg["terrain"], gk)
    args = (gk["nx"], gk["ny"], gk["nz"], gk["dx"], gk["dy"], gk["dz"], gk["lat_center"], gk["lon_center"])
    go = O.Grid(*args, h_func=terr, transform=oracle_transform(cfg["transform"]))
    gd = S.RegionalGrid3D(*args, h_func=terr, transform=make_transform(cfg["transform"]), device=DEV)
    st = W.initial_state(cfg, go.Z_m, go.x_m, go.y_m)
    pho = O.Euler(go, O.Ops(go, periodic), C, dt=10.0, initial_era5_state=st)
    phd = S.Euler3D(gd, S.CGridOperator3D(gd, periodic), dict(C), dt=10.0,
                    initial_era5_state={"i_th_v": _int(st["th_v"]), "i_pi": _int(st["pi"])})
    bgo = pho.precompute_bg(pho.bg_reference())
    bgd = phd.precompute_bg(phd.bg_reference())
    rng = np.random.default_rng(2)
    s = {"u": rng.uniform(-30, 30, go.Z_u.shape), "v": rng.uniform(-30, 30, go.Z_v.shape),
         "eta_dot": rng.uniform(-0.02, 0.02, go.Z_w.shape)}
    if periodic:  # faces 0 and n are one face on a periodic axis
        s["u"][-1] = s["u"][0]
        s["v"][:, -1] = s["v"][:, 0]
    s["eta_dot"][:, :, [0, -1]] = 0.0
    q = st["rho"] * (1 + 0.05 * rng.standard_normal(go.Z_m.shape))
    fo = O.FFSL(go, 10.0, periodic).advect_3d_split(q, s, bgo)
    adv = S.FluxFormAdvector(gd, 10.0, periodic)
    fp = adv.advect_3d_split(_int(q), {k: _int(v) for k, v in s.items()}, bgd)
    assert fp.is_cuda
    e = rel_l2(_np(fp).transpose(2, 1, 0), fo)
    print(f"  rel-L2 {e:.2e}", flush=True)
    assert e < 1e-13
    assert rel_l2(fo, q) > 1e-3          # the transport moved mass


def test_tracer_step_matches_oracle():
    """Three FP64 steps with a passive tracer: the FFSL-transported rho
    weights the tracer (steppers.py:879-886, 1049-1060), so the HIP FFSL
    output reaches q_test."""
    from suetes_amd.cases import build_case
    print("[case] lam 20x18x12 with tracer q_test, 3 steps vs oracle", flush=True)
    cfg = tight(small_lam())
    gk = cfg["grid"]
    nx, ny, nz = gk["nx"], gk["ny"], gk["nz"]
    x = np.arange(nx)[:, None, None] / nx
    y = np.arange(ny)[None, :, None] / ny
    z = np.arange(nz)[None, None, :] / nz
    q = 1.0 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.exp(-z) + 0.0 * z
    case = build_case(cfg, device=DEV, physics_suite=Tracer(), extra_fields={"q_test": q})
    go, opo, sto, s_o, bc_o, ext_o = oracle_case(cfg, physics_suite=Tracer(), extra_fields={"q_test": q})
    s_p = case.state
    for i in range(3):
        s_o = sto.step(s_o, i * case.dt, ext_o, bc_o)
        s_p = case.stepper.step(s_p, i * case.dt, case.forcing, case.bc_fn)
        for k in FIELDS + ("q_test",):
            e = rel_l2(_np(s_p[k]), s_o[k])
            assert e < 1e-12, (i, k, e)
        # the tracer actually moved (the comparison is not of an unchanged field)
        assert rel_l2(s_o["q_test"], q) > 1e-6
        assert "tend_q_test_prev" in s_p


def test_relative_residual_matches_oracle_and_gmres_meets_tolerance():
    from suetes_amd.cases import build_case
    print("[case] relative_residual lam 20x18x12", flush=True)
    cfg = small_lam()
    case = build_case(cfg, device=DEV)
    go, opo, sto, s_o, bc_o, ext_o = oracle_case(cfg)
    stp = case.stepper
    solver = stp.implicit_solver
    bg = stp.bg_precomputed()
    bgo = sto.physics.precompute_bg(sto.physics.bg_reference())
    rng = np.random.default_rng(9)
    shapes = {"u": go.Z_u.shape, "v": go.Z_v.shape, "w": go.Z_w.shape, "pi": go.Z_m.shape,
              "eta_dot": go.Z_w.shape}
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "pi": 1e-4, "eta_dot": 1e-4}
    sol = {k: rng.standard_normal(s) * scale[k] for k, s in shapes.items()}
    rhs = {k: rng.standard_normal(s) * scale[k] for k, s in shapes.items()}
    ro = sto.solver.relative_residual(sol, rhs, bgo)
    rp = solver.relative_residual({k: _int(v) for k, v in sol.items()}, {k: _int(v) for k, v in rhs.items()}, bg)
    assert abs(rp - ro) <= 1e-12 * ro, (rp, ro)
    # the HIP GMRES solution meets the reference stopping rule's tolerance in this norm
    solver.solver_tol, solver.solver_maxiter = 1e-10, 60
    x = solver.solve({k: _int(v) for k, v in rhs.items()}, bg)
    rr = solver.relative_residual(x, {k: _int(v) for k, v in rhs.items()}, bg)
    print(f"  oracle {ro:.6e} product {rp:.6e}; solve residual {rr:.2e}", flush=True)
    assert rr < 1e-8


def test_dlpack_state_swap_step_matches_oracle():
    """A caller's NumPy state enters through DLPack (torch.from_dlpack, zero
    copy), is moved to the device, stepped, and leaves through DLPack
    (np.from_dlpack of the host copy)."""
    from suetes_amd.cases import build_case
    from suetes_amd.geometry import ref_view
    print("[case] DLPack state swap, lam 20x18x12, 1 step", flush=True)
    cfg = tight(small_lam())
    case = build_case(cfg, device=DEV)
    go, opo, sto, s_o, bc_o, ext_o = oracle_case(cfg)
    host = {k: np.array(s_o[k]) for k in FIELDS}
    state = {}
    for k, v in host.items():
        t = torch.from_dlpack(v)                                  # shares the NumPy memory
        assert t.data_ptr() == v.__array_interface__["data"][0]
        state[k] = ref_view(t.permute(2, 1, 0).contiguous().to(DEV))
    for k in ("u_prev", "v_prev", "w_prev", "eta_dot_prev", "tend_th_v_prev", "is_first_step"):
        state[k] = case.state[k]
    out = case.stepper.step(state, 0.0, case.forcing, case.bc_fn)
    s_o = sto.step(s_o, 0.0, ext_o, bc_o)
    for k in FIELDS:
        back = np.from_dlpack(out[k].detach().cpu().contiguous())
        e = rel_l2(back, s_o[k])
        assert e < 1e-12, (k, e)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gmres_zero_rhs_returns_x0_and_nonfinite_rhs_is_reported(precision):
    from suetes_amd.cases import build_case
    print(f"[case] GMRES b = 0 / NaN rhs, {precision}", flush=True)
    cfg = small_lam()
    case = build_case(cfg, device=DEV, precision=precision)
    solver = case.stepper.implicit_solver
    solver.prepare(case.stepper.bg_precomputed())
    from suetes_amd import _lib
    n = int(_lib.lib().sisl_krylov_size(C.byref(solver.op64.struct.g)))
    b = torch.zeros(n, dtype=torch.float64, device=DEV)
    x0 = torch.randn(n, dtype=torch.float64, device=DEV)
    x = solver.solve_flat(b, x0)
    assert torch.equal(x, x0)                 # the oracle / JAX: ptol = NaN, no cycle
    xo = O.gmres(lambda v: v, np.zeros(8), x0=np.arange(8.0), tol=1e-6)
    np.testing.assert_array_equal(xo, np.arange(8.0))
    bn = torch.randn(n, dtype=torch.float64, device=DEV)
    bn[n // 3] = float("nan")
    with pytest.raises(RuntimeError, match="non-finite"):
        solver.solve_flat(bn, torch.zeros_like(bn))
