"""Host-side logic of the product on the CPU: the torch glue (grid, background,
tendencies, torch linear operator, FFSL, sponge) against the oracle, and the
C-ABI library exports.  No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import suetes_amd as S
from suetes_amd import _lib
from suetes_amd import workloads as W
from suetes_amd.cases import make_transform
from oracle import sisl as O
from tests.helpers import C, oracle_transform, small_lam, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def both(cfg):
    gk = cfg["grid"]
    terr = W.make_terrain(cfg["terrain"], gk)
    go = O.Grid(gk["nx"], gk["ny"], gk["nz"], gk["dx"], gk["dy"], gk["dz"], gk["lat_center"], gk["lon_center"],
                h_func=terr, transform=oracle_transform(cfg["transform"]))
    gp = S.RegionalGrid3D(gk["nx"], gk["ny"], gk["nz"], gk["dx"], gk["dy"], gk["dz"], gk["lat_center"],
                          gk["lon_center"], h_func=terr, transform=make_transform(cfg["transform"]))
    return go, gp


def tnp(t):
    return t.detach().cpu().numpy()


def test_grid_matches_oracle():
    go, gp = both(small_lam())
    for name in ("Z_m", "Z_u", "Z_v", "Z_w", "dz_m_full", "dz_w_full", "z_xi_w", "z_eta_w", "f_u", "f_v"):
        a, b = tnp(getattr(gp, name)), getattr(go, name)
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-9, err_msg=name)
    for loc in "muvw":
        np.testing.assert_allclose(tnp(gp.m_factors[loc]), go.m_factors[loc], rtol=1e-15)


def _phys(cfg):
    go, gp = both(cfg)
    st = W.initial_state(cfg, go.Z_m, go.x_m, go.y_m)
    opo, opp = O.Ops(go, cfg["periodic_axes"]), S.CGridOperator3D(gp, cfg["periodic_axes"])
    core = cfg["core"]
    kw = dict(dt=core["dt"], N_bv=core["N_bv"], damp_height=5000.0, max_damp=3.0, nu_div_factor=0.1,
              nu_h_factor=0.1)
    pho = O.Euler(go, opo, C, initial_era5_state=st, **kw)
    sti = {k: torch.from_numpy(np.ascontiguousarray(v.transpose(2, 1, 0))) for k, v in st.items()}
    php = S.Euler3D(gp, opp, C, initial_era5_state={"i_th_v": sti["th_v"], "i_pi": sti["pi"]}, **kw)
    return go, gp, pho, php, st


def _int(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0)))


def test_background_and_tendencies_match_oracle():
    go, gp, pho, php, st = _phys(small_lam())
    np.testing.assert_allclose(tnp(php.theta_bg), pho.theta_bg, rtol=1e-14)
    np.testing.assert_allclose(tnp(php.pi_bg), pho.pi_bg, rtol=1e-14)
    np.testing.assert_allclose(tnp(php.tau_damp), pho.tau_damp, rtol=1e-12, atol=1e-15)
    bgo = pho.precompute_bg(pho.bg_reference())
    bgp = php.precompute_bg(php.bg_reference())
    rng = np.random.default_rng(0)
    sp = {"u": rng.standard_normal(go.Z_u.shape), "v": rng.standard_normal(go.Z_v.shape),
          "w": rng.standard_normal(go.Z_w.shape) * 0.1, "pi": rng.standard_normal(go.Z_m.shape) * 1e-4,
          "eta_dot": rng.standard_normal(go.Z_w.shape) * 1e-4}
    thp = rng.standard_normal(go.Z_m.shape)
    sp["th_v"] = pho.theta_bg + thp
    sp["th_v_prime_u"] = O.Ops(go).avg(thp, 0, "m", "u")
    sp["th_v_prime_v"] = O.Ops(go).avg(thp, 1, "m", "v")
    sp["th_v_prime_w"] = O.Ops(go).avg(thp, 2, "m", "w")
    spp = {k: _int(v) for k, v in sp.items()}
    to = pho.get_tendencies(sp, bgo, True)
    tp = php.get_tendencies(spp, bgp, True)
    for k in ("u", "v", "w", "pi", "th_v"):
        assert rel_l2(tnp(tp[k]).transpose(2, 1, 0), to[k]) < 1e-12, k
    Lo = pho.linear_operator(sp, bgo, 10.0)
    Lp = php.linear_operator(spp, bgp, 10.0)
    for k in Lo:
        assert rel_l2(tnp(Lp[k]).transpose(2, 1, 0), Lo[k]) < 1e-13, k
    nlo = pho.nonlinear_pi_tendency(sp, thp, bgo)
    nlp = php.nonlinear_pi_tendency(spp, _int(thp), bgp)
    assert rel_l2(tnp(nlp).transpose(2, 1, 0), nlo) < 1e-12


def test_ffsl_and_sponge_match_oracle():
    cfg = small_lam()
    go, gp, pho, php, st = _phys(cfg)
    bgo = pho.precompute_bg(pho.bg_reference())
    bgp = php.precompute_bg(php.bg_reference())
    rng = np.random.default_rng(1)
    s = {"u": rng.uniform(-30, 30, go.Z_u.shape), "v": rng.uniform(-30, 30, go.Z_v.shape),
         "eta_dot": rng.uniform(-0.02, 0.02, go.Z_w.shape)}
    rho = st["rho"] * (1 + 0.01 * rng.standard_normal(go.Z_m.shape))
    for per in ((), (0, 1)):
        fo = O.FFSL(go, 10.0, per).advect_3d_split(rho, s, bgo)
        fp = S.FluxFormAdvector(gp, 10.0, per).advect_3d_split(_int(rho), {k: _int(v) for k, v in s.items()}, bgp)
        assert rel_l2(tnp(fp).transpose(2, 1, 0), fo) < 1e-13
    opo, opp = O.Ops(go), S.CGridOperator3D(gp)
    spo = O.DaviesSponge(go, opo, sponge_depth=4, dt=10.0)
    spp = S.DaviesSponge(gp, opp, sponge_depth=4, dt=10.0)
    state = {"u": s["u"], "v": s["v"], "th_v": st["th_v"] + 1.0, "pi": st["pi"] * 1.001, "w": st["w"]}
    ext = {k: np.zeros_like(v) + 0.5 for k, v in state.items()}
    bo = spo.blend(state, ext)
    bp = spp.blend({k: torch.from_numpy(v.copy()) for k, v in state.items()},
                   {k: torch.from_numpy(v.copy()) for k, v in ext.items()})
    for k in bo:
        np.testing.assert_allclose(tnp(bp[k]), bo[k], rtol=1e-15, atol=0)
    # x-only relaxation of the mountain-wave benchmark: product vs oracle and
    # its closed form (weight 1 on the edge columns, 0 from `cells` inward)
    state["rho"] = st["rho"]
    ext["rho"] = np.zeros_like(st["rho"]) + 0.5
    xo = O.XRelaxation(go.nx, cells=3).blend(state, ext)
    xp = S.XRelaxation(gp, cells=3).blend({k: torch.from_numpy(v.copy()) for k, v in state.items()},
                                          {k: torch.from_numpy(v.copy()) for k, v in ext.items()})
    for k in xo:
        np.testing.assert_allclose(tnp(xp[k]), xo[k], rtol=1e-15, atol=0)
    np.testing.assert_array_equal(xo["th_v"][0], ext["th_v"][0])
    np.testing.assert_array_equal(xo["u"][-1], ext["u"][-1])
    np.testing.assert_array_equal(xo["pi"][3:-3], state["pi"][3:-3])
    np.testing.assert_array_equal(xo["w"], state["w"])


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "suetes_sisl.h")).read()
    names = sorted(set(re.findall(r"^(?:int|int32_t|int64_t)\s+(sisl_[A-Za-z0-9_]+)\(", hdr, re.M)))
    assert len(names) >= 20
    L = _lib.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTED)


def test_kernels_refuse_cpu_tensors():
    f = torch.zeros(4, 4, 4, dtype=torch.float64)
    with pytest.raises(RuntimeError, match="HIP device"):
        S.tensor_product_interp_3d(f, torch.zeros(3, 4, 4, 4, dtype=torch.float64))


def _dense_solver(solver, bg):
    """Replace the HIP GMRES of ``solver`` by a dense solve of the torch
    operator (tiny grids only), so the solve's autograd rule runs on the CPU."""
    n = int(sum(int(np.prod(s)) for s in _packed_shapes(solver.physics.grid)))
    with torch.no_grad():
        eye = torch.eye(n, dtype=torch.float64)
        cols = [solver.scaled_operator(eye[:, q], {k: (v.detach() if torch.is_tensor(v) else v)
                                                   for k, v in bg.items()}) for q in range(n)]
        A = torch.stack(cols, dim=1)

    def solve_flat(b, x0, transpose=False):
        M = A.T if transpose else A
        return torch.linalg.solve(M, b)

    solver.solve_flat = solve_flat
    solver._bg, solver._bg_id = bg, id(bg)
    return A


def _packed_shapes(g):
    return [(g.nz, g.ny, g.nx + 1), (g.nz, g.ny + 1, g.nx), (g.nz + 1, g.ny, g.nx), (g.nz, g.ny, g.nx),
            (g.nz + 1, g.ny, g.nx)]


def test_solve_adjoint_reaches_terrain_through_the_operator():
    """d/d(bump amplitudes) of <c, A(h)^{-1} b>: the implicit-solve rule
    (transposed solve, then -lambda^T d(A x)/d(inputs) through the torch
    operator) against central differences, with the HIP solve replaced by a
    dense solve of the same operator on a 6x5x4 grid."""
    from suetes_amd.cases import build_terrain_case
    cfg = W.adjoint(nx=6, ny=5, nz=4)
    cfg["terrain"][1]["centres"] = np.array([[-1500.0, 0.0], [1000.0, 500.0]])
    cfg["terrain"][1]["amps"] = np.array([250.0, 150.0])
    rng = np.random.default_rng(3)
    g0 = build_terrain_case(cfg, torch.tensor([250.0, 150.0], dtype=torch.float64), device="cpu").grid
    shapes = _packed_shapes(g0)
    rhs = {k: torch.from_numpy(rng.standard_normal(s)) for k, s in zip(S.steppers.KEYS, shapes)}
    cvec = torch.from_numpy(rng.standard_normal(sum(int(np.prod(s)) for s in shapes)))

    def J(amps):
        case = build_terrain_case(cfg, amps, device="cpu")
        case.op.use_stop_grad = False  # differentiate the padded edge values too
        sol = case.stepper.implicit_solver
        bg = case.stepper.bg_precomputed()
        _dense_solver(sol, bg)
        x = sol.solve(rhs, bg)
        return (S.steppers.pack(x) * cvec).sum()

    a0 = torch.tensor([250.0, 150.0], dtype=torch.float64, requires_grad=True)
    val = J(a0)
    (grad,) = torch.autograd.grad(val, a0)
    assert float(grad.abs().max()) > 0
    for q in range(2):
        h = 1e-2
        e = torch.zeros(2, dtype=torch.float64)
        e[q] = h
        with torch.no_grad():
            fd = (float(J(a0.detach() + e)) - float(J(a0.detach() - e))) / (2 * h)
        assert abs(fd - float(grad[q])) <= 1e-6 * abs(fd), (q, fd, float(grad[q]))


def test_tricubic_plan_holder_releases_after_last_pending_adjoint():
    """The adjoint plan of a coordinate set stays while any gather's adjoint
    is pending and goes with the last one; tokens of gathers whose autograd
    nodes die without a backward (a checkpoint's recomputation) vanish from
    the pending set by themselves."""
    import gc

    from suetes_amd.operators import _PlanShare
    sh = _PlanShare()
    toks = [sh.register() for _ in range(3)]
    recomputed = sh.register()
    sh.plan = object()  # built at the first adjoint
    sh.done(toks[0])
    assert sh.plan is not None and len(sh.pending) == 3
    del recomputed  # its node died without a backward
    gc.collect()
    assert len(sh.pending) == 2
    sh.done(toks[1])
    assert sh.plan is not None
    sh.done(toks[2])
    assert sh.plan is None and len(sh.pending) == 0


def test_mixed_solver_options_are_validated():
    from suetes_amd.steppers import SemiImplicitSolver3D
    for kw in ({"inner_gram_schmidt": "twice"}, {"inner_coefficients": "fp16"}, {"inner_basis": "bf16"},
               {"inner_operand": "fp16"}, {"refine_preconditioner": "bf16"}):
        with pytest.raises(ValueError):
            SemiImplicitSolver3D(None, 1.0, precision="mixed", **kw)


def test_bench_refuses_more_than_one_gpu():
    """One GPU per box: --gpus above 1 exits with a message, before touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


@pytest.mark.parametrize("ny,axes", [(None, ("x",)), (10, ("x", "y")), (10, ("y",))])
def test_benchmark_sponge_matches_oracle(ny, axes):
    """boundaries.py:123-189: masks and blend of the product against the
    oracle (float32 weights, bit for bit)."""
    nx, nz, depth = 14, 5, 4
    so = O.BenchmarkSponge(nx, ny, sponge_depth=depth, axes=axes)
    sp = S.BenchmarkSponge(nx, ny, sponge_depth=depth, axes=axes)
    nyy = ny if ny is not None else 3
    rng = np.random.default_rng(7)
    shapes = {"u": (nx + 1, nyy, nz), "v": (nx, nyy + 1, nz), "w": (nx, nyy, nz + 1), "th_v": (nx, nyy, nz),
              "pi": (nx, nyy, nz), "rho": (nx, nyy, nz), "q_tr": (nx, nyy, nz), "eta_dot": (nx, nyy, nz + 1)}
    if ny is not None:  # w is blended with the mass mask: same shape as the mass points in the reference
        shapes["w"] = (nx, nyy, nz)
    state = {k: rng.standard_normal(s) for k, s in shapes.items()}
    ext = {k: rng.standard_normal(s) for k, s in shapes.items()}
    bo = so.blend(state, ext)
    bp = sp.blend({k: torch.from_numpy(v.copy()) for k, v in state.items()},
                  {k: torch.from_numpy(v.copy()) for k, v in ext.items()})
    assert set(bo) == set(bp)
    for k in bo:
        np.testing.assert_allclose(tnp(bp[k]), bo[k], rtol=1e-15, atol=1e-15, err_msg=k)


def test_dlpack_state_round_trip_is_zero_copy_and_exact():
    """The drop-in swaps state with the caller through DLPack
    (torch.from_dlpack / np.from_dlpack): a NumPy state dict enters as
    tensors that share its memory, the internal layout and back is exact, and
    the results leave as NumPy arrays over the tensors' memory."""
    from suetes_amd.steppers import state_to_internal, state_to_reference
    rng = np.random.default_rng(11)
    nx, ny, nz = 6, 5, 4
    st = {"u": rng.standard_normal((nx + 1, ny, nz)), "v": rng.standard_normal((nx, ny + 1, nz)),
          "w": rng.standard_normal((nx, ny, nz + 1)), "pi": rng.standard_normal((nx, ny, nz)),
          "th_v": rng.standard_normal((nx, ny, nz)), "rho": rng.standard_normal((nx, ny, nz)),
          "eta_dot": rng.standard_normal((nx, ny, nz + 1))}
    tens = {k: torch.from_dlpack(v) for k, v in st.items()}
    for k, v in st.items():
        assert tens[k].data_ptr() == v.__array_interface__["data"][0]   # shared memory
    back = state_to_reference(state_to_internal(tens))
    for k, v in st.items():
        out = np.from_dlpack(back[k].contiguous())
        np.testing.assert_array_equal(out, v)
    # a write through the NumPy view is seen by the tensor (no copy on entry)
    st["pi"][0, 0, 0] = 123.0
    assert float(tens["pi"][0, 0, 0]) == 123.0


def test_rho_transport_skipped_only_when_unobserved():
    """The stepper transports rho when a tracer, a physics suite or an
    unknown boundary function can observe it."""
    from suetes_amd.steppers import SISLStepper3D
    from suetes_amd.cases import identity_bc

    class Ph:
        physics_suite = None

    st = SISLStepper3D.__new__(SISLStepper3D)
    st.physics, st.tracer_keys = Ph(), ["q"]
    sponge = S.BenchmarkSponge(8, 6)
    assert not st._rho_transport_observed(identity_bc, {"rho": 1})
    assert not st._rho_transport_observed(sponge.blend, {"rho": 1})
    assert st._rho_transport_observed(lambda s, f: s, {"rho": 1})
    assert st._rho_transport_observed(identity_bc, {"rho": 1, "q": 2})
    Ph.physics_suite = object()
    assert st._rho_transport_observed(identity_bc, {"rho": 1})
