"""Fused per-step glue of the SISL step on the HIP kernels of csrc/tendency.hip.

The reference writes the explicit tendencies, the hyperdiffusion, the
predictor-corrector right-hand sides and the divergence damping of a step
(suetes/regional3d/euler.py:162-372, diffusion.py:93-172,
steppers.py:776-1035) as array expressions; their torch restatement in
``steppers.SISLStepper3D`` issues ~740 launches per step.  ``FusedGlue``
runs the same arithmetic in five FP64 kernels:

* ``explicit``: del^4 hyperdiffusion (four launches) and one launch for every
  explicit tendency, combined into the arrival fields u_in .. th_in, the
  kinematic residual, theta's tendency and the packed first guess x0;
* ``predictor``: theta at n+1 and the packed right-hand side of the implicit
  solve (two launches per predictor-corrector pass);
* ``final``: divergence damping and unscaling of the solution (one launch).

Reverse mode: each group is one autograd node whose backward is a
hand-written gather-form kernel set (``sisl_explicit_adj_f64``,
``sisl_predictor_adj_f64``, ``sisl_final_adj_f64``) following the
reference's own rules (the stop-gradient on repeated edge values) exactly.
With the stop-gradient switched off (``use_stop_grad = False`` on the
operators: the exact derivative) a group's backward re-evaluates its torch
restatement on the saved inputs and differentiates it.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib

TEND_FIELDS = ["mm", "mu", "mv", "fu", "fv", "dmdx", "dmdy", "th_bg", "pi_bg", "thu_bg", "thv_bg", "thw_bg",
               "zxu", "zev", "zxm", "zem", "rz", "zxiw", "zetaw", "dzw", "dzm", "cpi", "dthm", "cx", "cy",
               "cxp", "cyp", "czs", "czp"]


class SislTendency(C.Structure):
    _fields_ = ([("g", _lib.SislGrid), ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                 ("cp", C.c_double), ("grav", C.c_double), ("alpha", C.c_double), ("dt", C.c_double),
                 ("nu_h", C.c_double), ("nu_div", C.c_double)]
                + [(f, C.c_void_p) for f in TEND_FIELDS])


def _bind(L):
    if getattr(L, "_glue_bound", False):
        return L
    vp, i64 = C.c_void_p, C.c_int64
    P = C.POINTER(SislTendency)
    L.sisl_hyperdiffusion_workspace_f64.restype = i64
    L.sisl_hyperdiffusion_workspace_f64.argtypes = [C.POINTER(_lib.SislGrid)]
    L.sisl_hyperdiffusion_f64.argtypes = [P, vp, vp, vp, vp, vp, vp, vp]
    L.sisl_explicit_f64.argtypes = [P] + [vp] * 16 + [vp]
    L.sisl_predictor_f64.argtypes = [P] + [vp] * 17 + [vp]
    L.sisl_final_f64.argtypes = [P, vp, vp, vp, vp, vp, vp]
    L.sisl_explicit_adj_workspace_f64.restype = i64
    L.sisl_explicit_adj_workspace_f64.argtypes = [C.POINTER(_lib.SislGrid)]
    L.sisl_explicit_adj_f64.argtypes = [P] + [vp] * 20 + [vp]
    L.sisl_predictor_adj_workspace_f64.restype = i64
    L.sisl_predictor_adj_workspace_f64.argtypes = [C.POINTER(_lib.SislGrid)]
    L.sisl_predictor_adj_f64.argtypes = [P] + [vp] * 25 + [vp]
    L.sisl_final_adj_f64.argtypes = [P] + [vp] * 6 + [vp]
    L._glue_bound = True
    return L


class FusedGlue:
    """Static fields of the fused kernels for one stepper and background
    (built once; the background never changes during a run)."""

    def __init__(self, stepper, bg):
        ph, op, g = stepper.physics, stepper.physics.op, stepper.physics.grid
        self.stepper, self.bg = stepper, bg
        c = ph.c
        dzw = bg["dz_w_full"]
        rz = g.dz / dzw
        zxu = op.diff(g.i_Z_m, 0, "m", "u")
        zev = op.diff(g.i_Z_m, 1, "m", "v")
        mu, mv = g.i_m_factors["u"], g.i_m_factors["v"]
        f = {
            "mm": g.i_m_factors["m"], "mu": mu, "mv": mv, "fu": g.i_f_u, "fv": g.i_f_v,
            "dmdx": g.i_dm_dx_m, "dmdy": g.i_dm_dy_m,
            "th_bg": ph.i_theta_bg, "pi_bg": ph.i_pi_bg,
            "thu_bg": bg["th_v_u"], "thv_bg": bg["th_v_v"], "thw_bg": bg["th_v_w"],
            "zxu": zxu, "zev": zev, "zxm": op.avg(zxu, 0, "u", "m"), "zem": op.avg(zev, 1, "v", "m"),
            "rz": rz, "zxiw": g.i_z_xi_w, "zetaw": g.i_z_eta_w, "dzw": dzw, "dzm": bg["dz_m_full"],
            "cpi": bg["C_pi"],
            "dthm": op.avg(op.diff(bg["th_v"], 2, "m", "w") * rz, 2, "w", "m"),
            "cx": bg["rho_u"] * bg["th_v_u"] * bg["dz_u"] / mu[None],
            "cy": bg["rho_v"] * bg["th_v_v"] * bg["dz_v"] / mv[None],
            "cxp": bg["rho_u"] * bg["dz_u"] / mu[None],
            "cyp": bg["rho_v"] * bg["dz_v"] / mv[None],
            "czs": dzw * bg["rho_w"] * bg["th_v_w"],
            "czp": dzw * bg["rho_w"],
        }
        self.t = {k: v.detach().to(torch.float64).contiguous() for k, v in f.items()}
        _lib.require_device(*self.t.values())
        per = op.periodic_axes
        grid = _lib.SislGrid(g.nx, g.ny, g.nz, int(0 in per), int(1 in per))
        self.struct = SislTendency(grid, float(g.dx), float(g.dy), float(g.dz), float(c["cp"]), float(c["g"]),
                                   float(stepper.alpha), float(stepper.dt), float(ph.nu_h), float(ph.nu_div),
                                   *[self.t[k].data_ptr() for k in TEND_FIELDS])
        self.L = _bind(_lib.lib())
        self.nx, self.ny, self.nz = g.nx, g.ny, g.nz
        self._work = None

    # -- HIP forward ------------------------------------------------------------------
    def _new(self, *shape, like):
        return torch.empty(shape, dtype=torch.float64, device=like.device)

    def explicit(self, u, v, w, pi, th, eta, tend=None):
        """(u_in, v_in, w_in, pi_in, th_in, res, tth, x0) of the step from state
        (u, v, w, pi, th, eta) (internal tensors)."""
        L, s = self.L, _lib.stream_ptr(u.device)
        ins = [t.contiguous() for t in (u, v, w, pi, th, eta)]
        _lib.require_device(*ins)
        u, v, w, pi, th, eta = ins
        lap2 = None
        if self.struct.nu_h != 0.0:
            if self._work is None:
                n = int(L.sisl_hyperdiffusion_workspace_f64(C.byref(self.struct.g)))
                self._work = torch.empty(n, dtype=torch.float64, device=u.device)
            lap2 = torch.empty(4 * pi.numel(), dtype=torch.float64, device=u.device)
            _lib.check(L.sisl_hyperdiffusion_f64(C.byref(self.struct), u.data_ptr(), v.data_ptr(), w.data_ptr(),
                                                 th.data_ptr(), self._work.data_ptr(), lap2.data_ptr(), s),
                       "hyperdiffusion")
        outs = [torch.empty_like(u), torch.empty_like(v), torch.empty_like(w), torch.empty_like(pi),
                torch.empty_like(th), torch.empty_like(w), torch.empty_like(th)]
        x0 = torch.empty(u.numel() + v.numel() + 2 * w.numel() + pi.numel(), dtype=torch.float64, device=u.device)
        _lib.check(L.sisl_explicit_f64(C.byref(self.struct), *[t.data_ptr() for t in ins],
                                       lap2.data_ptr() if lap2 is not None else None,
                                       *[o.data_ptr() for o in outs], x0.data_ptr(),
                                       tend.data_ptr() if tend is not None else None, s), "explicit tendencies")
        return tuple(outs) + (x0,)

    def explicit_adj(self, saved, gouts):
        """Cotangents of (u, v, w, pi, th, eta) from those of the explicit
        stage's outputs (u_in, v_in, w_in, pi_in, th_in, res, tth; None for
        zero), at the saved state."""
        u, v, w, pi, th, eta = [t.contiguous() for t in saved]
        _lib.require_device(u, v, w, pi, th, eta)
        gs = [g.contiguous() if g is not None else None for g in gouts[:7]]
        for g in gs:
            if g is not None:
                _lib.require_device(g)
        L = self.L
        n = int(L.sisl_explicit_adj_workspace_f64(C.byref(self.struct.g)))
        work = torch.empty(n, dtype=torch.float64, device=u.device)
        outs = [torch.empty_like(u), torch.empty_like(v), torch.empty_like(w), torch.empty_like(pi),
                torch.empty_like(th), torch.empty_like(eta)]
        _lib.check(L.sisl_explicit_adj_f64(C.byref(self.struct), *[t.data_ptr() for t in (u, v, w, pi, th, eta)],
                                           *[g.data_ptr() if g is not None else None for g in gs],
                                           *[o.data_ptr() for o in outs], work.data_ptr(),
                                           _lib.stream_ptr(u.device)), "explicit stage adjoint")
        return outs

    def predictor(self, rhs_u, rhs_v, rhs_w, rhs_pi, th_adv, tth, w_n, reta, g_pi, g_u, g_v, g_w, g_eta,
                  m_u, m_v):
        """(b, th_next): the packed right-hand side of one predictor-corrector
        pass and theta at n+1."""
        ins = [t.contiguous() for t in (rhs_u, rhs_v, rhs_w, rhs_pi, th_adv, tth, w_n, reta, g_pi, g_u, g_v,
                                         g_w, g_eta, m_u, m_v)]
        _lib.require_device(*ins)
        th_next = torch.empty_like(ins[4])
        n = rhs_u.numel() + rhs_v.numel() + 2 * rhs_w.numel() + rhs_pi.numel()
        b = torch.empty(n, dtype=torch.float64, device=rhs_u.device)
        _lib.check(self.L.sisl_predictor_f64(C.byref(self.struct), *[t.data_ptr() for t in ins], th_next.data_ptr(),
                                             b.data_ptr(), _lib.stream_ptr(rhs_u.device)), "predictor")
        return b, th_next

    def predictor_adj(self, saved, th_next, gb, gth):
        """Cotangents of the fifteen predictor inputs from those of (b,
        th_next) (None for zero), at the saved inputs and th_next."""
        ins = [t.contiguous() for t in saved]
        g_pi, g_u, g_v, g_eta, m_u, m_v = ins[8], ins[9], ins[10], ins[12], ins[13], ins[14]
        th_next = th_next.contiguous()
        gb = gb.contiguous() if gb is not None else None
        gth = gth.contiguous() if gth is not None else None
        _lib.require_device(g_pi, g_u, g_v, g_eta, m_u, m_v, th_next, *[g for g in (gb, gth) if g is not None])
        L = self.L
        n = int(L.sisl_predictor_adj_workspace_f64(C.byref(self.struct.g)))
        work = torch.empty(n, dtype=torch.float64, device=th_next.device)
        outs = [torch.empty_like(t) for t in ins]
        _lib.check(L.sisl_predictor_adj_f64(C.byref(self.struct), *[t.data_ptr() for t in
                                                                    (g_pi, g_u, g_v, g_eta, m_u, m_v, th_next)],
                                            gb.data_ptr() if gb is not None else None,
                                            gth.data_ptr() if gth is not None else None,
                                            *[o.data_ptr() for o in outs], work.data_ptr(),
                                            _lib.stream_ptr(th_next.device)), "predictor adjoint")
        return outs

    def final_adj(self, gouts, like):
        """Cotangent of the packed solution from those of (u, v, pi, eta)."""
        gs = [g.contiguous() if g is not None else None for g in gouts]
        for g in gs:
            if g is not None:
                _lib.require_device(g)
        gx = torch.empty_like(like)
        work = torch.empty(self.nx * self.ny * self.nz, dtype=torch.float64, device=like.device)
        _lib.check(self.L.sisl_final_adj_f64(C.byref(self.struct), *[g.data_ptr() if g is not None else None
                                                                     for g in gs],
                                             gx.data_ptr(), work.data_ptr(), _lib.stream_ptr(like.device)),
                   "final adjoint")
        return gx

    def final(self, x):
        """(u, v, pi, eta) of the step's result from the packed solution x."""
        nx, ny, nz = self.nx, self.ny, self.nz
        x = x.contiguous()
        _lib.require_device(x)
        kw = dict(dtype=torch.float64, device=x.device)
        u = torch.empty((nz, ny, nx + 1), **kw)
        v = torch.empty((nz, ny + 1, nx), **kw)
        pi = torch.empty((nz, ny, nx), **kw)
        eta = torch.empty((nz + 1, ny, nx), **kw)
        _lib.check(self.L.sisl_final_f64(C.byref(self.struct), x.data_ptr(), u.data_ptr(), v.data_ptr(),
                                         pi.data_ptr(), eta.data_ptr(), _lib.stream_ptr(x.device)), "final")
        return u, v, pi, eta


def _recompute_grads(fn, saved, need, gouts, nout):
    """Cotangents of the inputs of ``fn`` (the torch restatement of a fused
    group) at the saved inputs, for the output cotangents ``gouts``."""
    leaves = [t.detach().requires_grad_(bool(n)) for t, n in zip(saved, need)]
    req = [x for x, n in zip(leaves, need) if n]
    if not req:
        return [None] * len(need)
    with torch.enable_grad():
        outs = fn(*leaves)
    pairs = [(o, g) for o, g in zip(outs[:nout], gouts[:nout]) if g is not None and o.requires_grad]
    if not pairs:
        return [None] * len(need)
    grads = iter(torch.autograd.grad([o for o, _ in pairs], req, [g for _, g in pairs], allow_unused=True))
    return [next(grads) if n else None for n in need]


class ExplicitFn(torch.autograd.Function):
    """Fused explicit stage; outputs (u_in, v_in, w_in, pi_in, th_in, res, tth, x0)."""

    # the state fields (u, v, w, pi, th, eta) each output depends on
    # (euler.py:162-372: u_in, v_in, w_in, pi_in, th_in, res, tth); an output
    # none of whose inputs is differentiated is marked non-differentiable, so
    # that the gathers downstream of it skip their adjoints
    DEPS = ((0, 1, 3, 4), (0, 1, 3, 4), (2, 3, 4), (0, 1, 3, 4, 5), (2, 4), (0, 1, 2, 5), (2, 4))

    @staticmethod
    def forward(ctx, glue, u, v, w, pi, th, eta):
        outs = glue.explicit(u, v, w, pi, th, eta)
        ctx.glue = glue
        ctx.save_for_backward(u, v, w, pi, th, eta)
        need = ctx.needs_input_grad[1:]
        ctx.frozen = [not any(need[q] for q in d) for d in ExplicitFn.DEPS]
        ctx.mark_non_differentiable(outs[-1], *[o for o, f in zip(outs, ctx.frozen) if f])
        return outs

    @staticmethod
    def backward(ctx, *gouts):
        need = ctx.needs_input_grad[1:]
        if not any(need):
            return (None,) * 7
        if not getattr(ctx.glue.stepper.physics.op, "use_stop_grad", True):
            # the exact derivative (edge copies pass their cotangents on) is
            # only in the torch restatement
            g = _recompute_grads(ctx.glue.stepper._explicit_stage_torch, ctx.saved_tensors, need, gouts, 7)
            return (None,) + tuple(g)
        gouts = [None if f else g for g, f in zip(gouts[:7], ctx.frozen)]
        g = ctx.glue.explicit_adj(ctx.saved_tensors, gouts)
        return (None,) + tuple(gi if n else None for gi, n in zip(g, need))


class PredictorFn(torch.autograd.Function):
    """Fused predictor-corrector pass; outputs (b, th_next)."""

    @staticmethod
    def forward(ctx, glue, *ins):
        b, th_next = glue.predictor(*ins)
        ctx.glue = glue
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(*ins, th_next)
        return b, th_next

    @staticmethod
    def backward(ctx, *gouts):
        need = ctx.needs_input_grad[1:]
        saved = ctx.saved_tensors
        ins, th_next = saved[:-1], saved[-1]
        if not any(need) or all(g is None for g in gouts):
            return (None,) * 16
        if not getattr(ctx.glue.stepper.physics.op, "use_stop_grad", True):
            g = _recompute_grads(ctx.glue.stepper._predictor_torch, ins, need, gouts, 2)
            return (None,) + tuple(g)
        g = ctx.glue.predictor_adj(ins, th_next, gouts[0], gouts[1])
        return (None,) + tuple(gi if n else None for gi, n in zip(g, need))


class FinalFn(torch.autograd.Function):
    """Fused divergence damping and unscaling; outputs (u, v, pi, eta)."""

    @staticmethod
    def forward(ctx, glue, x):
        outs = glue.final(x)
        ctx.glue = glue
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(x)
        return outs

    @staticmethod
    def backward(ctx, *gouts):
        if not ctx.needs_input_grad[1] or all(g is None for g in gouts):
            return None, None
        (x,) = ctx.saved_tensors
        if not getattr(ctx.glue.stepper.physics.op, "use_stop_grad", True):
            g = _recompute_grads(ctx.glue.stepper._final_torch, ctx.saved_tensors, ctx.needs_input_grad[1:],
                                 gouts, 4)
            return (None,) + tuple(g)
        return None, ctx.glue.final_adj(gouts, x)
