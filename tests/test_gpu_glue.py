"""Fused explicit tendencies, hyperdiffusion, predictor-corrector and
divergence damping (csrc/tendency.hip) against the FP64 oracle and the
torch restatement of the same arithmetic.

Bars (FP64):
* raw explicit tendencies (tu, tv, tw, tpi, tth) vs the oracle's
  Euler.get_tendencies(is_explicit=True) (euler.py:208-372 with the
  hyperdiffusion of diffusion.py:80-172 switched on): rel-L2 <= 1e-12;
* every output of the three fused groups vs their torch restatement:
  rel-L2 <= 1e-13;
* whole steps with the fused kernels vs the torch ops: rel-L2 <= 1e-12 per
  field after 3 steps (limited area, closed and periodic lateral axes);
* reverse mode: each group's backward (with the edge stop-gradient off, so
  that it is the exact derivative) against central differences of its HIP
  forward along a random direction (the groups are at most quadratic, so
  the central difference is exact up to rounding): relative 1e-9.
"""
import numpy as np
import pytest
import torch

from oracle import sisl as O
from tests.helpers import oracle_case, rel_l2, small_bubble, small_lam

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _np(t):
    return t.detach().cpu().numpy()


def _rt(t):
    """internal (Z, Y, X) tensor -> reference-shaped NumPy"""
    return _np(t).transpose(2, 1, 0)


def _rel(a, b):
    a, b = a.detach().double(), b.detach().double()
    return float((a - b).norm() / b.norm().clamp_min(1e-300))


def _perturbed_state(case, seed=0):
    from suetes_amd.steppers import state_to_internal
    st = state_to_internal(case.state)
    rng = torch.Generator(device="cpu").manual_seed(seed)

    def noise(t, s):
        return t + s * torch.randn(t.shape, generator=rng, dtype=torch.float64).to(t.device)
    ph = case.stepper.physics
    return {"u": noise(st["u"], 2.0), "v": noise(st["v"], 2.0), "w": noise(st["w"], 0.3),
            "pi": noise(st["pi"], 2e-4), "th_v": noise(st["th_v"], 0.5),
            "eta_dot": noise(st["eta_dot"], 2e-4)}


@pytest.mark.parametrize("periodic", [False, True])
def test_fused_explicit_tendencies_match_oracle(periodic):
    from suetes_amd.cases import build_case
    print(f"[case] fused explicit tendencies vs oracle, lam 20x18x12 periodic={periodic}", flush=True)
    cfg = small_lam()
    if periodic:
        cfg["periodic_axes"] = (0, 1)
    case = build_case(cfg, device=DEV)
    go, opo, sto, s_o, bc_o, ext_o = oracle_case(cfg)
    stp = case.stepper
    assert stp.physics.nu_h > 0
    bg = stp.bg_precomputed()
    s = _perturbed_state(case)
    glue = stp._fused(bg)
    n = s["u"].numel() + s["v"].numel() + s["w"].numel() + s["pi"].numel()
    tend = torch.zeros(n, dtype=torch.float64, device=DEV)
    outs = glue.explicit(s["u"], s["v"], s["w"], s["pi"], s["th_v"], s["eta_dot"], tend=tend)
    pho = sto.physics
    bgo = pho.precompute_bg(pho.bg_reference())
    thp = _rt(s["th_v"]) - pho.theta_bg
    spo = {"u": _rt(s["u"]), "v": _rt(s["v"]), "w": _rt(s["w"]), "pi": _rt(s["pi"]) - pho.pi_bg,
           "eta_dot": _rt(s["eta_dot"]), "th_v": _rt(s["th_v"]),
           "th_v_prime_u": opo.avg(thp, 0, "m", "u"), "th_v_prime_v": opo.avg(thp, 1, "m", "v"),
           "th_v_prime_w": opo.avg(thp, 2, "m", "w")}
    to = pho.get_tendencies(spo, bgo, is_explicit=True)
    nu, nv, nw = s["u"].numel(), s["v"].numel(), s["w"].numel()
    parts = {"u": tend[:nu].view(s["u"].shape), "v": tend[nu:nu + nv].view(s["v"].shape),
             "w": tend[nu + nv:nu + nv + nw].view(s["w"].shape), "pi": tend[nu + nv + nw:].view(s["pi"].shape),
             "th_v": outs[6]}
    for k, t in parts.items():
        e = rel_l2(_rt(t), to[k])
        print(f"  {k}: rel-L2 {e:.2e}", flush=True)
        assert e < 1e-12, (k, e)
    # the hyperdiffusion part alone (phys_diff) is non-trivial here
    assert np.linalg.norm(to["phys_diff_th_v"]) > 1e-6 * np.linalg.norm(to["th_v"])
    # the arrival fields against the torch restatement
    ref = stp._explicit_stage_torch(s["u"], s["v"], s["w"], s["pi"], s["th_v"], s["eta_dot"])
    for name, a, b in zip(("u_in", "v_in", "w_in", "pi_in", "th_in", "res", "tth"), outs[:7], ref):
        assert _rel(a, b) < 1e-13, (name, _rel(a, b))


def test_fused_predictor_and_final_match_torch_restatement():
    from suetes_amd.cases import build_case
    from suetes_amd.steppers import pack
    print("[case] fused predictor / final vs torch, lam 20x18x12", flush=True)
    case = build_case(small_lam(), device=DEV)
    stp = case.stepper
    bg = stp.bg_precomputed()
    glue = stp._fused(bg)
    s = _perturbed_state(case, 1)
    g = torch.Generator(device="cpu").manual_seed(5)

    def r(t, sc):
        return (sc * torch.randn(t.shape, generator=g, dtype=torch.float64)).to(DEV)
    ph = stp.physics
    ins = (s["u"] + r(s["u"], 1.0), s["v"] + r(s["v"], 1.0), s["w"] + r(s["w"], 0.1), r(s["pi"], 1e-4),
           r(s["th_v"], 0.3), r(s["th_v"], 1e-3), s["w"], r(s["w"], 1e-3), r(s["pi"], 1e-4),
           s["u"] + r(s["u"], 0.5), s["v"] + r(s["v"], 0.5), s["w"] + r(s["w"], 0.05), r(s["w"], 1e-4),
           s["u"] + r(s["u"], 0.5), s["v"] + r(s["v"], 0.5))
    b, th = glue.predictor(*ins)
    b_t, th_t = stp._predictor_torch(*ins)
    assert _rel(th, th_t) < 1e-13
    assert _rel(b, b_t) < 1e-13, _rel(b, b_t)
    x = pack({"u": ins[0], "v": ins[1], "w": ins[2], "pi": r(s["pi"], 0.1), "eta_dot": r(s["w"], 1e-2)})
    for a, c in zip(glue.final(x), stp._final_torch(x)):
        assert _rel(a, c) < 1e-13, _rel(a, c)
    assert ph.nu_div > 0


@pytest.mark.parametrize("cfgf,periodic", [(small_lam, False), (small_lam, True), (small_bubble, None)])
def test_fused_steps_match_torch_steps(cfgf, periodic):
    from suetes_amd.cases import build_case
    print(f"[case] fused vs torch glue, 3 FP64 steps, {cfgf.__name__} periodic={periodic}", flush=True)
    cfg = cfgf()
    if periodic:
        cfg["periodic_axes"] = (0, 1)
    cfg["core"].update(solver_tol=1e-13, solver_maxiter=60)
    a = build_case(cfg, device=DEV)
    b = build_case(cfg, device=DEV)
    b.stepper.fused_glue = False
    sa, sb = a.state, b.state
    for i in range(3):
        sa = a.stepper.step(sa, i * a.dt, a.forcing, a.bc_fn)
        sb = b.stepper.step(sb, i * b.dt, b.forcing, b.bc_fn)
        assert a.stepper._glue is not None and b.stepper._glue is None
        for k in ("u", "v", "w", "pi", "th_v", "rho", "eta_dot", "tend_th_v_prev"):
            e = _rel(sa[k], sb[k])
            assert e < 1e-12, (i, k, e)


def _adjoint_check(fn_hip, fn_grad, ins, nout, seed):
    """<J d, y> by central differences of the HIP forward vs <d, J^T y> from
    the node's backward, for random d and y."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    ds = [torch.randn(t.shape, generator=g, dtype=torch.float64).to(DEV) * t.abs().mean().clamp_min(1e-3)
          for t in ins]
    with torch.no_grad():
        outs = fn_hip(*ins)[:nout]
    ys = [torch.randn(o.shape, generator=g, dtype=torch.float64).to(DEV) for o in outs]
    h = 1e-4
    with torch.no_grad():
        op = fn_hip(*[t + h * d for t, d in zip(ins, ds)])[:nout]
        om = fn_hip(*[t - h * d for t, d in zip(ins, ds)])[:nout]
    jd = sum(float(((p - m) / (2 * h) * y).sum()) for p, m, y in zip(op, om, ys))
    leaves = [t.detach().clone().requires_grad_(True) for t in ins]
    outs = fn_grad(*leaves)[:nout]
    grads = torch.autograd.grad(outs, leaves, ys, allow_unused=True)
    jty = sum(float((gr * d).sum()) for gr, d in zip(grads, ds) if gr is not None)
    return jd, jty


def test_fused_nodes_backward_is_the_adjoint_of_the_hip_forward():
    from suetes_amd.cases import build_case
    from suetes_amd.glue import ExplicitFn, FinalFn, PredictorFn
    from suetes_amd.steppers import pack
    print("[case] fused node adjoints vs central differences, lam 20x18x12", flush=True)
    case = build_case(small_lam(), device=DEV)
    stp = case.stepper
    glue = stp._fused(stp.bg_precomputed())
    # the step's reverse mode drops the cotangents of repeated edge values as
    # the reference does (operators.py:37-63); the exact derivative, which
    # central differences see, keeps them
    case.op.use_stop_grad = False
    s = _perturbed_state(case, 2)
    ins = (s["u"], s["v"], s["w"], s["pi"], s["th_v"], s["eta_dot"])
    jd, jty = _adjoint_check(lambda *a: glue.explicit(*a), lambda *a: ExplicitFn.apply(glue, *a), ins, 7, 0)
    print(f"  explicit: {jd:.12e} vs {jty:.12e}", flush=True)
    assert abs(jd - jty) <= 1e-9 * abs(jty)
    g = torch.Generator(device="cpu").manual_seed(6)

    def r(t, sc):
        return (sc * torch.randn(t.shape, generator=g, dtype=torch.float64)).to(DEV)
    pins = (s["u"], s["v"], s["w"], r(s["pi"], 1e-4), r(s["th_v"], 0.3), r(s["th_v"], 1e-3), s["w"],
            r(s["w"], 1e-3), r(s["pi"], 1e-4), s["u"] + r(s["u"], 0.5), s["v"] + r(s["v"], 0.5),
            s["w"] + r(s["w"], 0.05), r(s["w"], 1e-4), s["u"] + r(s["u"], 0.5), s["v"] + r(s["v"], 0.5))
    jd, jty = _adjoint_check(lambda *a: glue.predictor(*a), lambda *a: PredictorFn.apply(glue, *a), pins, 2, 1)
    print(f"  predictor: {jd:.12e} vs {jty:.12e}", flush=True)
    assert abs(jd - jty) <= 1e-9 * abs(jty)
    x = pack({"u": s["u"], "v": s["v"], "w": s["w"], "pi": r(s["pi"], 0.1), "eta_dot": r(s["w"], 1e-2)})
    jd, jty = _adjoint_check(lambda a: glue.final(a), lambda a: FinalFn.apply(glue, a), (x,), 4, 2)
    print(f"  final: {jd:.12e} vs {jty:.12e}", flush=True)
    assert abs(jd - jty) <= 1e-9 * abs(jty)


def _explicit_vjp_cases():
    from tests.helpers import small_schaer
    per = small_lam()
    per["periodic_axes"] = (0, 1)
    return [("lam 20x18x12 closed", small_lam()), ("lam 20x18x12 periodic", per),
            ("bubble 16^3", small_bubble()), ("schaer 40x3x14", small_schaer())]


@pytest.mark.parametrize("name,cfg", _explicit_vjp_cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_explicit_stage_adjoint_matches_autodiff_of_the_restatement(name, cfg):
    """The hand-written reverse mode of the explicit stage
    (sisl_explicit_adj_f64) against torch autograd through the torch
    restatement of the same arithmetic (_explicit_stage_torch: the
    reference's array formulas, its stop-gradient on repeated edge values):
    every input cotangent rel-L2 <= 1e-12, for all seven output cotangents
    at once and for each alone."""
    from suetes_amd.cases import build_case
    print(f"[case] explicit stage adjoint vs autodiff, {name}", flush=True)
    case = build_case(cfg, device=DEV)
    stp = case.stepper
    glue = stp._fused(stp.bg_precomputed())
    s = _perturbed_state(case, 3)
    ins = (s["u"], s["v"], s["w"], s["pi"], s["th_v"], s["eta_dot"])
    leaves = [t.detach().clone().requires_grad_(True) for t in ins]
    outs = stp._explicit_stage_torch(*leaves)
    g = torch.Generator(device="cpu").manual_seed(11)
    ys = [torch.randn(o.shape, generator=g, dtype=torch.float64).to(DEV) for o in outs]
    for sel in [list(range(7))] + [[q] for q in range(7)]:
        gsel = [ys[q] if q in sel else None for q in range(7)]
        pairs = [(o, y) for o, y in zip(outs, gsel) if y is not None]
        ref = torch.autograd.grad([o for o, _ in pairs], leaves, [y for _, y in pairs], retain_graph=True,
                                  allow_unused=True)
        got = glue.explicit_adj(ins, gsel + [None])
        for name_i, r, h in zip(("u", "v", "w", "pi", "th", "eta"), ref, got):
            r = torch.zeros_like(h) if r is None else r
            if float(r.norm()) == 0.0:
                assert float(h.norm()) == 0.0, (sel, name_i)
                continue
            e = _rel(h, r)
            assert e <= 1e-12, (sel, name_i, e)
    print("  all cotangent selections within 1e-12", flush=True)


@pytest.mark.parametrize("name,cfg", _explicit_vjp_cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_predictor_and_final_adjoints_match_autodiff_of_the_restatement(name, cfg):
    """The hand-written reverse modes of a predictor-corrector pass and of the
    final group (sisl_predictor_adj_f64, sisl_final_adj_f64) against torch
    autograd through their torch restatements (_predictor_torch,
    _final_torch; the reference's stop-gradient on repeated edge values):
    every input cotangent rel-L2 <= 1e-12."""
    from suetes_amd.cases import build_case
    from suetes_amd.steppers import pack
    print(f"[case] predictor / final adjoints vs autodiff, {name}", flush=True)
    case = build_case(cfg, device=DEV)
    stp = case.stepper
    glue = stp._fused(stp.bg_precomputed())
    s = _perturbed_state(case, 4)
    g = torch.Generator(device="cpu").manual_seed(8)

    def r(t, sc):
        return (sc * torch.randn(t.shape, generator=g, dtype=torch.float64)).to(DEV)
    pins = (s["u"], s["v"], s["w"], r(s["pi"], 1e-4), r(s["th_v"], 0.3), r(s["th_v"], 1e-3), s["w"],
            r(s["w"], 1e-3), r(s["pi"], 1e-4), s["u"] + r(s["u"], 0.5), s["v"] + r(s["v"], 0.5),
            s["w"] + r(s["w"], 0.05), r(s["w"], 1e-4), s["u"] + r(s["u"], 0.5), s["v"] + r(s["v"], 0.5))
    leaves = [t.detach().clone().requires_grad_(True) for t in pins]
    b, th_next = stp._predictor_torch(*leaves)
    yb, yt = r(b, 1.0), r(th_next, 1.0)
    for gb, gt in ((yb, yt), (yb, None), (None, yt)):
        pairs = [(o, y) for o, y in ((b, gb), (th_next, gt)) if y is not None]
        ref = torch.autograd.grad([o for o, _ in pairs], leaves, [y for _, y in pairs], retain_graph=True,
                                  allow_unused=True)
        got = glue.predictor_adj(pins, th_next.detach(), gb, gt)
        for q, (rq, h) in enumerate(zip(ref, got)):
            rq = torch.zeros_like(h) if rq is None else rq
            if float(rq.norm()) == 0.0:
                assert float(h.norm()) == 0.0, (q, gb is None, gt is None)
                continue
            e = _rel(h, rq)
            assert e <= 1e-12, (q, gb is None, gt is None, e)
    x = pack({"u": s["u"], "v": s["v"], "w": s["w"], "pi": r(s["pi"], 0.1), "eta_dot": r(s["w"], 1e-2)})
    xl = x.detach().clone().requires_grad_(True)
    outs = stp._final_torch(xl)
    ys = [r(o, 1.0) for o in outs]
    (ref,) = torch.autograd.grad(outs, [xl], ys)
    got = glue.final_adj(ys, x)
    assert _rel(got, ref) <= 1e-12, _rel(got, ref)
    print("  predictor and final within 1e-12", flush=True)
