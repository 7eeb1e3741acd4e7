"""The split-explicit RK3 dynamical core on MI355X (drop-in for the
reference's ``SplitExplicitStepper3D``, suetes/regional3d/steppers.py:1076-1737,
built by ``build_dynamical_core('split-explicit', ...)``, steppers.py:95-119).

A step is three Wicker-Skamarock RK3 stages (dt/3, dt/2, dt).  Each stage
evaluates the slow tendencies at its stage state -- the explicit Euler
tendencies, third-order upwind advection of theta' and of the momentum
components, and the theta' Exner flux -- and then integrates the deviation
from the stage state over ns/3, ns/2 and ns acoustic substeps.  A substep is
a horizontal predictor (half-step winds), an implicit vertical column for
(pi, w) off-centred by alpha and solved by the Thomas algorithm, and a
horizontal corrector with the averaged pressure gradient; divergence
damping acts on the total winds and Rayleigh damping on the total w.

The acoustic loop runs as four FP64 HIP kernels per substep
(csrc/acoustic.hip: predictor, column right-hand side, column sweep,
corrector) behind one C entry point; the slow tendencies are torch ops over
the HIP stagger kernels.  Reverse mode: the loop is one autograd node whose
backward re-evaluates the torch restatement of the loop
(``_acoustic_core_torch``) on the saved inputs and differentiates it.

State dicts at the public API have the reference's keys and shapes;
internally every field is a contiguous (Z, Y, X) tensor.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .operators import DIM
from .steppers import FluxFormAdvector, SISLStepper3D, state_to_internal, state_to_reference

AC_FIELDS = ["mu", "mv", "mm", "mw", "thu", "thv", "thw", "zxu", "zev", "rz", "dzm", "cpi", "cx", "cy",
             "rtw", "zxiw", "zetaw", "tau"]


class SislAcoustic(C.Structure):
    _fields_ = ([("g", _lib.SislGrid), ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                 ("cp", C.c_double), ("alpha", C.c_double), ("nu_div", C.c_double)]
                + [(f, C.c_void_p) for f in AC_FIELDS])


def _bind(L):
    if getattr(L, "_acoustic_bound", False):
        return L
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.sisl_acoustic_workspace_f64.restype = i64
    L.sisl_acoustic_workspace_f64.argtypes = [C.POINTER(_lib.SislGrid)]
    L.sisl_acoustic_loop_f64.argtypes = [C.POINTER(SislAcoustic), i32, C.c_double] + [vp] * 17 + [vp]
    L._acoustic_bound = True
    return L


def _cat_levels(first, middle, last):
    return torch.cat([first.unsqueeze(0), middle, last.unsqueeze(0)], dim=0)


class _SlowKernels:
    """The slow tendencies of a stage on csrc/tendency.hip (forward only):
    the static fields of the fused explicit tendencies, with the
    hyperdiffusion coefficient of the physics' HyperFilter."""

    def __init__(self, stepper, bg):
        from .glue import FusedGlue, _bind as bind_glue
        ph = stepper.physics
        self.glue = FusedGlue(stepper, bg)
        self.glue.struct.nu_h = float(ph.diffusion.nu_h)
        self.L = bind_glue(_lib.lib())
        L = self.L
        if not getattr(L, "_slow_bound", False):
            vp = C.c_void_p
            L.sisl_slow_tendencies_f64.argtypes = [C.POINTER(glue_struct_type())] + [vp] * 8 + [C.c_double,
                                                                                                vp, vp, vp]
            L._slow_bound = True
        self.eps_l = float(ph.c.get("eps_l", 1e-12))
        self.hyper_on = float(ph.diffusion.nu_h) != 0.0
        self._work = None

    def lap2(self, u, v, w, th):
        """del^4 of (u_m, v_m, w_m, theta') of a state (4 mass fields)."""
        L, g = self.L, self.glue
        if self._work is None:
            n = int(L.sisl_hyperdiffusion_workspace_f64(C.byref(g.struct.g)))
            self._work = torch.empty(n, dtype=torch.float64, device=u.device)
        out = torch.empty(4 * th.numel(), dtype=torch.float64, device=u.device)
        _lib.check(L.sisl_hyperdiffusion_f64(C.byref(g.struct), u.data_ptr(), v.data_ptr(), w.data_ptr(),
                                             th.data_ptr(), self._work.data_ptr(), out.data_ptr(),
                                             _lib.stream_ptr(u.device)), "hyperdiffusion")
        return out

    def tendencies(self, cur, lap2s, lap20):
        u, v, w, pi, th, eta = (cur[k].detach().contiguous() for k in ("u", "v", "w", "pi", "th_v", "eta_dot"))
        nz, ny, nx = pi.shape
        nu, nv, nw, nm = u.numel(), v.numel(), w.numel(), pi.numel()
        adv = torch.empty(4 * nm, dtype=torch.float64, device=u.device)
        tend = torch.empty(nu + nv + nw + 2 * nm, dtype=torch.float64, device=u.device)
        ptr = (lambda x: x.data_ptr() if x is not None else None)
        _lib.check(self.L.sisl_slow_tendencies_f64(C.byref(self.glue.struct), u.data_ptr(), v.data_ptr(),
                                                   w.data_ptr(), pi.data_ptr(), th.data_ptr(), eta.data_ptr(),
                                                   ptr(lap2s), ptr(lap20), self.eps_l, adv.data_ptr(),
                                                   tend.data_ptr(), _lib.stream_ptr(u.device)),
                   "slow tendencies")
        o = [0, nu, nu + nv, nu + nv + nw, nu + nv + nw + nm, nu + nv + nw + 2 * nm]
        return {"u": tend[o[0]:o[1]].view(u.shape), "v": tend[o[1]:o[2]].view(v.shape),
                "w": tend[o[2]:o[3]].view(w.shape), "pi": tend[o[3]:o[4]].view(pi.shape),
                "th_v": tend[o[4]:o[5]].view(th.shape)}


def glue_struct_type():
    from .glue import SislTendency
    return SislTendency


class _AcousticKernels:
    """Static fields of the acoustic kernels for one stepper (built once: the
    background is the reference state of the run and never changes)."""

    def __init__(self, stepper, bg):
        ph, op, g = stepper.physics, stepper.physics.op, stepper.physics.grid
        self.stepper, self.bg = stepper, bg
        self.L = _bind(_lib.lib())
        mu, mv = g.i_m_factors["u"], g.i_m_factors["v"]
        f = {"mu": mu, "mv": mv, "mm": g.i_m_factors["m"], "mw": g.i_m_factors["w"],
             "thu": bg["th_v_u"], "thv": bg["th_v_v"], "thw": bg["th_v_w"],
             "zxu": op.diff(g.i_Z_m, 0, "m", "u"), "zev": op.diff(g.i_Z_m, 1, "m", "v"),
             "rz": g.dz / bg["dz_w_full"], "dzm": bg["dz_m_full"], "cpi": bg["C_pi"],
             "cx": bg["rho_u"] * bg["th_v_u"] * bg["dz_u"] / mu[None],
             "cy": bg["rho_v"] * bg["th_v_v"] * bg["dz_v"] / mv[None],
             "rtw": bg["rho_w"] * bg["th_v_w"], "zxiw": g.i_z_xi_w, "zetaw": g.i_z_eta_w,
             "tau": ph.i_tau_damp}
        self.fields = {k: v.detach().to(torch.float64).contiguous() for k, v in f.items()}
        nx, ny, nz = g.nx, g.ny, g.nz
        self.shapes = {"u": (nz, ny, nx + 1), "v": (nz, ny + 1, nx), "w": (nz + 1, ny, nx), "m": (nz, ny, nx)}
        want = {"mu": (ny, nx + 1), "mv": (ny + 1, nx), "mm": (ny, nx), "mw": (ny, nx), "thu": self.shapes["u"],
                "thv": self.shapes["v"], "thw": self.shapes["w"], "zxu": self.shapes["u"], "zev": self.shapes["v"],
                "rz": self.shapes["w"], "dzm": self.shapes["m"], "cpi": self.shapes["m"], "cx": self.shapes["u"],
                "cy": self.shapes["v"], "rtw": self.shapes["w"], "zxiw": self.shapes["w"],
                "zetaw": self.shapes["w"], "tau": self.shapes["w"]}
        for k, shp in want.items():
            if tuple(self.fields[k].shape) != shp or not self.fields[k].is_cuda:
                raise RuntimeError(f"acoustic field {k}: shape {tuple(self.fields[k].shape)}, expected {shp}")
        self.grid = _lib.SislGrid(g.nx, g.ny, g.nz, int(0 in op.periodic_axes), int(1 in op.periodic_axes))
        self.ctx = SislAcoustic(self.grid, g.dx, g.dy, g.dz, ph.c["cp"], stepper.alpha, ph.nu_div,
                                *[self.fields[k].data_ptr() for k in AC_FIELDS])
        self.work_len = int(self.L.sisl_acoustic_workspace_f64(C.byref(self.grid)))
        self._factors = {}

    def factors(self, dtau):
        """Thomas factors of the column matrix for one substep length: the
        sub-diagonal, c'_k and 1 / (main_k - lower_k c'_{k-1})."""
        key = float(dtau)
        if key not in self._factors:
            with torch.no_grad():
                lower, main, upper = self.stepper._column_matrix(self.bg, dtau)
                cp_ = torch.empty_like(main)
                inv = torch.empty_like(main)
                inv[0] = 1.0 / main[0]
                cp_[0] = upper[0] * inv[0]
                for k in range(1, main.shape[0]):
                    inv[k] = 1.0 / (main[k] - lower[k] * cp_[k - 1])
                    cp_[k] = upper[k] * inv[k]
            self._factors[key] = tuple(t.contiguous() for t in (lower, cp_, inv))
        return self._factors[key]

    def loop(self, dtau, n, u0, v0, w0, p0, uc, vc, wc, su, sv, sw, sp):
        """Deviation (u'', v'', w'', pi'') after ``n`` substeps from the
        initial deviation (u0 .. p0), stage state (uc, vc, wc) and slow
        tendencies (su .. sp)."""
        lower, cpr, inv = self.factors(dtau)
        u, v, w, p = (t.detach().contiguous().clone() for t in (u0, v0, w0, p0))
        ub, vb = u0.detach().contiguous(), v0.detach().contiguous()
        ins = [t.detach().contiguous() for t in (uc, vc, wc, su, sv, sw, sp)]
        for t, loc in zip([u, v, w, p, *ins, ub, vb], "uvwm" + "uvw" + "uvwm" + "uv"):
            if not (t.is_cuda and t.dtype == torch.float64) or tuple(t.shape) != self.shapes[loc]:
                raise RuntimeError(f"the acoustic kernels take FP64 device tensors of the grid's shapes "
                                   f"(got {tuple(t.shape)} {t.dtype} for a {loc} field)")
        work = torch.empty(self.work_len, dtype=torch.float64, device=u.device)
        ptrs = [x.data_ptr() for x in (lower, cpr, inv, u, v, w, p, *ins, ub, vb, work)]
        _lib.check(self.L.sisl_acoustic_loop_f64(C.byref(self.ctx), int(n), float(dtau), *ptrs,
                                                 _lib.stream_ptr(u.device)), "sisl_acoustic_loop_f64")
        return u, v, w, p


def _recompute_grads(fn, saved, need, gouts):
    leaves = [t.detach().requires_grad_(bool(n)) for t, n in zip(saved, need)]
    req = [x for x, n in zip(leaves, need) if n]
    if not req:
        return [None] * len(need)
    with torch.enable_grad():
        outs = fn(*leaves)
    pairs = [(o, g) for o, g in zip(outs, gouts) if g is not None and o.requires_grad]
    if not pairs:
        return [None] * len(need)
    grads = iter(torch.autograd.grad([o for o, _ in pairs], req, [g for _, g in pairs], allow_unused=True))
    return [next(grads) if n else None for n in need]


class _AcousticFn(torch.autograd.Function):
    """Acoustic loop of one stage on the HIP kernels; outputs the totals
    (u, v, w, pi) = stage state + deviation."""

    @staticmethod
    def forward(ctx, kern, dtau, n, ui, vi, wi, pii, uc, vc, wc, pic, su, sv, sw, sp):
        du, dv, dw, dp = kern.loop(dtau, n, ui - uc, vi - vc, wi - wc, pii - pic, uc, vc, wc, su, sv, sw, sp)
        ctx.kern, ctx.dtau, ctx.n = kern, dtau, n
        ctx.save_for_backward(ui, vi, wi, pii, uc, vc, wc, pic, su, sv, sw, sp)
        return uc + du, vc + dv, wc + dw, pic + dp

    @staticmethod
    def backward(ctx, *gouts):
        st = ctx.kern.stepper

        def fn(ui, vi, wi, pii, uc, vc, wc, pic, su, sv, sw, sp):
            return st._acoustic_core_torch({"u": ui, "v": vi, "w": wi, "pi": pii},
                                           {"u": uc, "v": vc, "w": wc, "pi": pic},
                                           {"u": su, "v": sv, "w": sw, "pi": sp}, ctx.kern.bg, ctx.dtau, ctx.n)
        g = _recompute_grads(fn, ctx.saved_tensors, ctx.needs_input_grad[3:], gouts)
        return (None, None, None) + tuple(g)


class SplitExplicitStepper3D:
    """Split-explicit RK3 core (steppers.py:1076-1737)."""

    _call_bc = SISLStepper3D._call_bc

    def __init__(self, physics, dt, ns, alpha=0.55):
        # steppers.py:1092-1116 (Python's round: halves to even)
        self.physics, self.dt, self.ns, self.alpha = physics, dt, ns, alpha
        self.ns_stage1 = max(1, round(ns / 3))
        self.ns_stage2 = max(1, round(ns / 2))
        self.ns_stage3 = ns
        self.dtau_stage1 = (dt / 3.0) / float(self.ns_stage1)
        self.dtau_stage2 = (dt / 2.0) / float(self.ns_stage2)
        self.dtau_stage3 = dt / float(self.ns_stage3)
        self.ffsl_advector = FluxFormAdvector(physics.grid, dt, periodic_axes=physics.op.periodic_axes)
        self.tracer_keys = physics.physics_suite.tracer_keys if physics.physics_suite is not None else []
        self._bg = None
        self._ext_cache = None
        # run the acoustic loop on the HIP kernels (csrc/acoustic.hip) for FP64
        # device states, and (forward passes without a physics suite) the slow
        # tendencies on csrc/tendency.hip; False: the torch restatement
        self.fused_acoustic = True
        self.fused_slow = True
        self._kern = None
        self._slow = None

    # -- public API (reference shapes) -----------------------------------------
    def step(self, state, t, forcing, bc_fn, ml_params=None):
        if "eta_dot" not in state:
            state["eta_dot"] = torch.zeros_like(state["w"])
        return state_to_reference(self.step_internal(state_to_internal(state), t, forcing, bc_fn, ml_params))

    def integrate(self, state, t_start, num_steps, forcing, bc_fn):
        s = state_to_internal(state)
        if "eta_dot" not in s:
            s["eta_dot"] = torch.zeros_like(s["w"])
        for i in range(int(num_steps)):
            s = self.step_internal(s, t_start + i * self.dt, forcing, bc_fn)
        return state_to_reference(s)

    def bg_precomputed(self):
        if self._bg is None:
            ph = self.physics
            self._bg = ph.precompute_bg(ph.bg_reference())
        return self._bg

    # -- step (steppers.py:1118-1269) -------------------------------------------
    def step_internal(self, state, t, forcing, bc_fn, ml_params=None):
        ph, op = self.physics, self.physics.op
        c = ph.c
        bg = self.bg_precomputed()
        st = dict(state)
        if "eta_dot" not in st:
            st["eta_dot"] = torch.zeros_like(st["w"])
        st["eta_dot"], st["w"] = self._diagnose_eta_dot(st)
        thp = st["th_v"] - ph.i_theta_bg
        sp = {"u": st["u"], "v": st["v"], "w": st["w"], "th_v": st["th_v"], "pi": st["pi"] - ph.i_pi_bg,
              "eta_dot": st["eta_dot"], "th_v_prime_u": op.avg(thp, 0, "m", "u"),
              "th_v_prime_v": op.avg(thp, 1, "m", "v"), "th_v_prime_w": op.avg(thp, 2, "m", "w")}
        for k in st:
            if k not in sp:
                sp[k] = st[k]
        if self._slow_fused_applies(st):
            return self._step_fused_slow(st, bg, forcing, bc_fn)
        phys_tends, diff_tends = {}, {}
        if ph.physics_suite is not None:
            phys_tends = ph.physics_suite.get_explicit_tendencies(sp, bg, interior_mask=ph.interior_mask,
                                                                  ml_params=ml_params)
        if ph.nu_h > 0.0 or ph.nu_div > 0.0:
            diff_tends = ph.diffusion.get_tendencies(sp, bg)
        # the stages run with the physics suite and nu_h hidden
        # (steppers.py:1186-1191); the hyperdiffusion object keeps its own nu_h
        suite, nu_h = ph.physics_suite, ph.nu_h
        ph.physics_suite, ph.nu_h = None, 0.0
        try:
            s1 = self._compute_slow_tendencies(st, bg, phys_tends, diff_tends, ml_params)
            star = self._acoustic_loop(st, st, s1, bg, self.dtau_stage1, self.ns_stage1)
            star["eta_dot"], star["w"] = self._diagnose_eta_dot(star)
            s2 = self._compute_slow_tendencies(star, bg, phys_tends, diff_tends, ml_params)
            dstar = self._acoustic_loop(st, star, s2, bg, self.dtau_stage2, self.ns_stage2)
            dstar["eta_dot"], dstar["w"] = self._diagnose_eta_dot(dstar)
            s3 = self._compute_slow_tendencies(dstar, bg, phys_tends, diff_tends, ml_params)
            nxt = self._acoustic_loop(st, dstar, s3, bg, self.dtau_stage3, self.ns_stage3)
        finally:
            ph.physics_suite, ph.nu_h = suite, nu_h
        eps = c.get("eps", 1e-15)
        if suite is not None:
            for key in suite.tracer_keys:
                if key in st:
                    rq = self.ffsl_advector.advect_3d_split(st["rho"] * st[key], st, bg)
                    nxt[key] = rq / (nxt["rho"] + eps)
            nxt = state_to_internal(suite.apply_state_updates(state_to_reference(nxt)))
        nxt = self._call_bc(bc_fn, nxt, forcing)
        nxt["eta_dot"], nxt["w"] = self._diagnose_eta_dot(nxt)
        tmass = {}
        if suite is not None:
            tmass = {k: nxt["rho"] * nxt[k] for k in suite.tracer_keys if k in nxt}
        nxt["rho"] = c["p0"] / (c["Rd"] * nxt["th_v"]) * (nxt["pi"] ** (c["cvd"] / c["Rd"]))
        for k, rq in tmass.items():
            nxt[k] = rq / (nxt["rho"] + eps)
        return nxt

    def _slow_fused_applies(self, st):
        ph = self.physics
        if not (self.fused_slow and self._fused_applies(st) and ph.physics_suite is None):
            return False
        return not (torch.is_grad_enabled() and any(torch.is_tensor(v) and v.requires_grad for v in st.values()))

    def _step_fused_slow(self, st, bg, forcing, bc_fn):
        """The step with the slow tendencies on the HIP kernels (no physics
        suite, no tracers, no gradient): the start state's hyperdiffusion is
        the diff_tends term of every stage and, with nu_div > 0, also the
        stage-1 term (the same state)."""
        ph = self.physics
        c = ph.c
        if self._slow is None:
            self._slow = _SlowKernels(self, bg)
        sk = self._slow
        with_start = sk.hyper_on and (ph.nu_h > 0.0 or ph.nu_div > 0.0)
        with_stage = sk.hyper_on and ph.nu_div > 0.0

        def lap2(s):
            return sk.lap2(*(s[k].detach().contiguous() for k in ("u", "v", "w", "th_v")))
        lap0 = lap2(st) if (with_start or with_stage) else None
        nu_h = ph.nu_h
        ph.nu_h = 0.0  # hidden during the stages, as in the torch path
        try:
            s1 = sk.tendencies(st, lap0 if with_stage else None, lap0 if with_start else None)
            star = self._acoustic_loop(st, st, s1, bg, self.dtau_stage1, self.ns_stage1)
            star["eta_dot"], star["w"] = self._diagnose_eta_dot(star)
            s2 = sk.tendencies(star, lap2(star) if with_stage else None, lap0 if with_start else None)
            dstar = self._acoustic_loop(st, star, s2, bg, self.dtau_stage2, self.ns_stage2)
            dstar["eta_dot"], dstar["w"] = self._diagnose_eta_dot(dstar)
            s3 = sk.tendencies(dstar, lap2(dstar) if with_stage else None, lap0 if with_start else None)
            nxt = self._acoustic_loop(st, dstar, s3, bg, self.dtau_stage3, self.ns_stage3)
        finally:
            ph.nu_h = nu_h
        nxt = self._call_bc(bc_fn, nxt, forcing)
        nxt["eta_dot"], nxt["w"] = self._diagnose_eta_dot(nxt)
        nxt["rho"] = c["p0"] / (c["Rd"] * nxt["th_v"]) * (nxt["pi"] ** (c["cvd"] / c["Rd"]))
        return nxt

    # -- steppers.py:1271-1309 ----------------------------------------------------
    def _surface_and_terrain_w(self, u, v):
        """(w_sfc, m_w (u_w z_xi + v_w z_eta)) of the kinematic boundary condition."""
        op, g = self.physics.op, self.physics.grid
        uw = op.avg(op.avg(u, 0, "u", "m"), 2, "m", "w")
        vw = op.avg(op.avg(v, 1, "v", "m"), 2, "m", "w")
        mw = op.m3("w")
        w_sfc = mw[0] * (uw[0] * g.i_z_xi_w[0] + vw[0] * g.i_z_eta_w[0])
        return w_sfc, mw * (uw * g.i_z_xi_w + vw * g.i_z_eta_w)

    def _diagnose_eta_dot(self, state):
        g = self.physics.grid
        w_sfc, terr = self._surface_and_terrain_w(state["u"], state["v"])
        w = state["w"]
        w = _cat_levels(w_sfc, w[1:-1], torch.zeros_like(w_sfc))
        eta = (w - terr) / g.i_dz_w_full
        zero = torch.zeros_like(w_sfc)
        return _cat_levels(zero, eta[1:-1], zero), w

    # -- steppers.py:1310-1356 ----------------------------------------------------
    def _upwind_flux_3rd_order(self, flux, field, axis):
        ph = self.physics
        d = DIM[axis]
        n = field.shape[d]
        if axis in ph.op.periodic_axes:
            q = torch.cat([field.narrow(d, n - 2, 2), field, field.narrow(d, 0, 2)], dim=d)
        else:
            lo, hi = field.narrow(d, 0, 1), field.narrow(d, n - 1, 1)
            q = torch.cat([lo, lo, field, hi, hi], dim=d)
        q_im2, q_im1 = q.narrow(d, 0, n + 1), q.narrow(d, 1, n + 1)
        q_i, q_ip1 = q.narrow(d, 2, n + 1), q.narrow(d, 3, n + 1)
        q4 = (7.0 / 12.0) * (q_i + q_im1) - (1.0 / 12.0) * (q_ip1 + q_im2)
        qb = (1.0 / 12.0) * ((q_ip1 - q_im2) - 3.0 * (q_i - q_im1))
        clean = torch.where(flux.abs() < ph.c.get("eps_l", 1e-12), torch.zeros_like(flux), flux)
        return flux * (q4 + torch.sign(clean) * qb)

    # -- steppers.py:1358-1470 ----------------------------------------------------
    def _compute_slow_tendencies(self, cur, bg, phys_tends, diff_tends, ml_params):
        ph, op, g = self.physics, self.physics.op, self.physics.grid
        thp = cur["th_v"] - ph.i_theta_bg
        sp = {"u": cur["u"], "v": cur["v"], "w": cur["w"], "th_v": cur["th_v"], "pi": cur["pi"] - ph.i_pi_bg,
              "eta_dot": cur["eta_dot"] if "eta_dot" in cur else torch.zeros_like(cur["w"]),
              "th_v_prime_u": op.avg(thp, 0, "m", "u"), "th_v_prime_v": op.avg(thp, 1, "m", "v"),
              "th_v_prime_w": op.avg(thp, 2, "m", "w")}
        for k in cur:
            if k not in sp:
                sp[k] = cur[k]
        tends = dict(ph.get_tendencies(sp, bg, is_explicit=True, ml_params=ml_params))
        tends["pi"] = tends["pi"] + ph.nonlinear_pi_tendency(sp, thp, bg)
        for k in ("u", "v", "w", "th_v"):
            if phys_tends and k in phys_tends:
                tends[k] = tends[k] + phys_tends[k]
            if diff_tends and k in diff_tends:
                tends[k] = tends[k] + diff_tends[k]
        mm, mu, mv = op.m3("m"), op.m3("u"), op.m3("v")
        um = op.avg(cur["u"], 0, "u", "m")
        vm = op.avg(cur["v"], 1, "v", "m")
        wm = op.avg(cur["w"], 2, "w", "m")
        U = sp["u"] / mu
        V = sp["v"] / mv
        W = sp["eta_dot"] * bg["dz_w_full"]
        rz = g.dz / bg["dz_m_full"]
        divU = op.diff(U, 0, "u", "m")
        divV = op.diff(V, 1, "v", "m")
        divW = op.diff(W, 2, "w", "m") * rz

        def adv(q):
            ax = op.diff(self._upwind_flux_3rd_order(U, q, 0), 0, "u", "m")
            ay = op.diff(self._upwind_flux_3rd_order(V, q, 1), 1, "v", "m")
            az = op.diff(self._upwind_flux_3rd_order(W, q, 2), 2, "w", "m") * rz
            return -(mm ** 2) * (ax - q * divU + ay - q * divV) - (az - q * divW)
        tends["th_v"] = tends["th_v"] + adv(thp)
        tends["u"] = tends["u"] + op.avg(adv(um), 0, "m", "u")
        tends["v"] = tends["v"] + op.avg(adv(vm), 1, "m", "v")
        tends["w"] = tends["w"] + op.avg(adv(wm), 2, "m", "w")
        return tends

    # -- steppers.py:1472-1651 ----------------------------------------------------
    def _fused_applies(self, s_cur):
        u = s_cur["u"]
        return self.fused_acoustic and u.is_cuda and u.dtype == torch.float64

    def _acoustic_loop(self, s_init, s_cur, slow, bg, dtau, num_steps):
        if self._fused_applies(s_cur):
            if self._kern is None:
                self._kern = _AcousticKernels(self, bg)
            u, v, w, p = _AcousticFn.apply(self._kern, dtau, num_steps, s_init["u"], s_init["v"], s_init["w"],
                                           s_init["pi"], s_cur["u"], s_cur["v"], s_cur["w"], s_cur["pi"],
                                           slow["u"], slow["v"], slow["w"], slow["pi"])
        else:
            u, v, w, p = self._acoustic_core_torch(s_init, s_cur, slow, bg, dtau, num_steps)
        out = {"u": u, "v": v, "w": w, "pi": p,
               "th_v": s_init["th_v"] + dtau * num_steps * slow["th_v"], "rho": s_cur["rho"],
               "eta_dot": s_cur["eta_dot"] if "eta_dot" in s_cur else torch.zeros_like(w)}
        for k in s_cur:
            if k not in out:
                out[k] = s_cur[k]
        return out

    def _acoustic_core_torch(self, s_init, s_cur, slow, bg, dtau, num_steps):
        """Torch restatement of the loop: totals (u, v, w, pi) after ``num_steps`` substeps."""
        ph, op, g = self.physics, self.physics.op, self.physics.grid
        cp = ph.c["cp"]
        u_pp = s_init["u"] - s_cur["u"]
        v_pp = s_init["v"] - s_cur["v"]
        w_pp = s_init["w"] - s_cur["w"]
        pi_pp = s_init["pi"] - s_cur["pi"]
        ub, vb = u_pp, v_pp
        zxu = op.diff(g.i_Z_m, 0, "m", "u")
        zev = op.diff(g.i_Z_m, 1, "m", "v")

        def hgrad(p):
            gzm = op.avg(op.diff(p, 2, "m", "w") * (g.dz / bg["dz_w_full"]), 2, "w", "m")
            return (op.diff(p, 0, "m", "u") - zxu * op.avg(gzm, 0, "m", "u"),
                    op.diff(p, 1, "m", "v") - zev * op.avg(gzm, 1, "m", "v"))

        def ddamp(up, vp):
            div = op.diff(s_cur["u"] + up, 0, "u", "m") + op.diff(s_cur["v"] + vp, 1, "v", "m")
            return op.diff(div, 0, "m", "u"), op.diff(div, 1, "m", "v")

        def bounds(up, vp):
            nx, ny = up.shape[2] - 1, vp.shape[1] - 1
            if 0 in op.periodic_axes:
                up = torch.cat([up[..., :nx], up[..., :1]], dim=2)
            else:
                up = torch.cat([ub[..., :1], up[..., 1:nx], ub[..., nx:]], dim=2)
            if 1 in op.periodic_axes:
                vp = torch.cat([vp[:, :ny], vp[:, :1]], dim=1)
            else:
                vp = torch.cat([vb[:, :1], vp[:, 1:ny], vb[:, ny:]], dim=1)
            return up, vp

        for _ in range(num_steps):
            u_old, v_old, pi_old = u_pp, v_pp, pi_pp
            gox, goy = hgrad(pi_old)
            dox, doy = ddamp(u_old, v_old)
            ru = slow["u"] - cp * bg["th_v_u"] * gox + ph.nu_div * dox
            rv = slow["v"] - cp * bg["th_v_v"] * goy + ph.nu_div * doy
            u_half, v_half = bounds(u_old + 0.5 * dtau * ru, v_old + 0.5 * dtau * rv)
            mm = op.m3("m")
            fx = (u_half * bg["rho_u"] * bg["th_v_u"] * bg["dz_u"]) / op.m3("u")
            fy = (v_half * bg["rho_v"] * bg["th_v_v"] * bg["dz_v"]) / op.m3("v")
            div_h = mm * (op.diff(fx, 0, "u", "m") / bg["dz_m_full"] + op.diff(fy, 1, "v", "m") / bg["dz_m_full"])
            pi_pp, w_pp = self._solve_vertical_acoustic_column(pi_pp, w_pp, u_half, v_half, div_h, slow["pi"],
                                                               slow["w"], bg, dtau)
            gnx, gny = hgrad(pi_pp)
            dhx, dhy = ddamp(u_half, v_half)
            u_pp = u_old + dtau * (slow["u"] - 0.5 * cp * bg["th_v_u"] * (gox + gnx) + ph.nu_div * dhx)
            v_pp = v_old + dtau * (slow["v"] - 0.5 * cp * bg["th_v_v"] * (goy + gny) + ph.nu_div * dhy)
            u_pp, v_pp = bounds(u_pp, v_pp)
            wt = s_cur["w"] + w_pp
            w_pp = wt / (1.0 + dtau * ph.i_tau_damp) - s_cur["w"]
        return s_cur["u"] + u_pp, s_cur["v"] + v_pp, s_cur["w"] + w_pp, s_cur["pi"] + pi_pp

    # -- steppers.py:1653-1737 ----------------------------------------------------
    def _column_matrix(self, bg, dtau):
        """(lower, main, upper) of the implicit column for pi (mass levels)."""
        a, cp = self.alpha, self.physics.c["cp"]
        thw = bg["th_v_w"]
        gamma = (a * dtau) ** 2 * cp * thw * bg["rho_w"] * bg["th_v_w"] / bg["dz_w_full"]
        zero = torch.zeros_like(gamma[0])
        gamma = _cat_levels(zero, gamma[1:-1], zero)
        r = bg["C_pi"] / bg["dz_m_full"]
        lower = -r * gamma[:-1]
        upper = -r * gamma[1:]
        return lower, 1.0 - lower - upper, upper

    def _solve_vertical_acoustic_column(self, pi_pp, w_pp, u_pp, v_pp, div_h, slow_f_pi, slow_f_w, bg, dtau):
        ph, op, g = self.physics, self.physics.op, self.physics.grid
        a, cp = self.alpha, ph.c["cp"]
        thw, dzw, dzm = bg["th_v_w"], bg["dz_w_full"], bg["dz_m_full"]
        w_sfc, terr = self._surface_and_terrain_w(u_pp, v_pp)
        zero = torch.zeros_like(w_sfc)
        gp = op.diff(pi_pp, 2, "m", "w") * (g.dz / dzw)
        w_exp = w_pp + dtau * (slow_f_w - cp * thw * gp)
        w_avg = _cat_levels(w_sfc, ((1.0 - a) * w_pp + a * w_exp)[1:-1], zero)
        edz = _cat_levels(zero, (w_avg - terr)[1:-1], zero)
        fz = edz * bg["rho_w"] * bg["th_v_w"]
        div_z = (op.diff(fz, 2, "w", "m") * g.dz) / dzm
        rhs = dtau * (slow_f_pi - bg["C_pi"] * (div_h + div_z))
        lower, main, upper = self._column_matrix(bg, dtau)
        dpi = _thomas(lower, main, upper, rhs)
        gd = op.diff(dpi, 2, "m", "w") * (g.dz / dzw)
        w_next = w_exp - a * dtau * cp * thw * gd
        w_next = _cat_levels(w_sfc, w_next[1:-1], zero)
        return pi_pp + dpi, w_next


def _thomas(lower, main, upper, rhs):
    """Tridiagonal solve along the level dimension (dim 0), one column per
    (y, x); differentiable (a torch loop over the levels)."""
    n = rhs.shape[0]
    cps, dps = [upper[0] / main[0]], [rhs[0] / main[0]]
    for k in range(1, n):
        den = main[k] - lower[k] * cps[k - 1]
        cps.append(upper[k] / den)
        dps.append((rhs[k] - lower[k] * dps[k - 1]) / den)
    x = [None] * n
    x[n - 1] = dps[n - 1]
    for k in range(n - 2, -1, -1):
        x[k] = dps[k] - cps[k] * x[k + 1]
    return torch.stack(x, dim=0)
