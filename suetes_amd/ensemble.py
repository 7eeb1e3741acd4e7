"""Batched ensemble of SISL steppers on uneven domains.

An ensemble advances every member one step in lockstep.  Each member runs
its own unchanged ``SISLStepper3D.step_internal`` in a worker thread; the
three hot operations of the step -- departure points (4 per step), tricubic
gathers (6 per step) and implicit solves (2 per step) -- are collected
from every member in turn (one member thread runs at a time), and the last
member issues ONE batched launch for all members (``sisl_departure_batched_*``, ``sisl_tricubic_batched_*``,
``sisl_gmres_batched``).  The per-member glue (tendencies, FFSL, EOS,
boundaries) stays torch ops on device tensors.

The batched tricubic and the batched solve are single autograd Functions over
all members, so the reverse mode issues batched transposed solves as well.
``integrate(..., checkpoint=True)`` checkpoints whole ensemble steps (a
per-member checkpoint could not be recomputed in lockstep): each step is one
autograd node that recomputes and differentiates itself in the backward pass.

Members must use ``precision='mixed'`` (the batched solver is the mixed
FP32/FP16/FP64 one) and share ``dt``; at most 32 members.
"""
from __future__ import annotations

import ctypes as C
import threading

import torch

from . import _lib
from .operators import Coords, _tricubic_adjoint
from .steppers import KEYS, LOC_ID, _operator_inputs, pack, state_to_internal, state_to_reference, unpack

MAX_MEMBERS = 32


class _Baton:
    """Lockstep schedule of the member threads: exactly one member runs at a
    time (no interpreter-lock contention, so the host keeps ahead of the
    device as in a single-threaded loop).  Member m runs until its next batched
    call, deposits its arguments and passes the baton to member m+1; the last
    member runs the batched function and passes the baton back to member 0.
    Every member therefore executes the same call sequence in the order
    0, 1, ..., M-1, batch.  A failure in any member releases and aborts all."""

    def __init__(self, M):
        self.M = M
        self.go = [threading.Semaphore(0) for _ in range(M)]
        self.items = [None] * M
        self.kind = None
        self.result = None
        self.broken = None

    def _check(self):
        if self.broken is not None:
            raise RuntimeError("ensemble step aborted") from self.broken

    def enter(self, m):
        self.go[m].acquire()
        self._check()

    def __call__(self, m, kind, payload, fn):
        self._check()
        if m == 0:
            self.kind = kind
            self.items = [None] * self.M
        elif kind != self.kind:
            err = RuntimeError(f"members diverged: {self.kind} vs {kind}")
            self.fail(err)
            raise err
        self.items[m] = payload
        if m < self.M - 1:
            self.go[m + 1].release()
        else:
            try:
                self.result = fn(self.items)
            except BaseException as e:  # noqa: BLE001 -- propagated to every member
                self.fail(e)
                raise
            self.go[0].release()
        self.go[m].acquire()
        self._check()
        return self.result[m]

    def leave(self, m):
        if m < self.M - 1:
            self.go[m + 1].release()

    def fail(self, exc):
        if self.broken is None:
            self.broken = exc
        for g in self.go:
            g.release()


def run_lockstep(baton, body):
    """Run ``body(m)`` for every member m in its own thread under ``baton``
    and return the results in member order.

    Grad mode and saved-tensor hooks are thread-local in torch, so each member
    thread re-installs the caller's: a non-reentrant checkpoint around the
    ensemble step then sees every tensor the members save (otherwise each
    member step would keep its whole graph).  The baton runs the members one
    at a time in a fixed order, so the tensors are saved in the same order in
    the forward pass and in the checkpoint's recomputation, which matches them
    by position.  The first failure is re-raised in the caller."""
    M = baton.M
    out = [None] * M
    errs = []
    grad = torch.is_grad_enabled()
    hooks = torch._C._autograd._top_saved_tensors_default_hooks(False)

    def work(m):
        pushed = False
        try:
            baton.enter(m)
            torch.set_grad_enabled(grad)
            if hooks is not None:
                torch._C._autograd._push_saved_tensors_default_hooks(*hooks)
                pushed = True
            out[m] = body(m)
            baton.leave(m)
        except BaseException as e:  # noqa: BLE001 -- re-raised in the caller
            errs.append(e)
            baton.fail(e)
        finally:
            if pushed:
                torch._C._autograd._pop_saved_tensors_default_hooks()
    threads = [threading.Thread(target=work, args=(m,)) for m in range(M)]
    for th in threads:
        th.start()
    baton.go[0].release()
    for th in threads:
        th.join()
    if errs:
        raise errs[0]
    return out


class _BatchedTricubic(torch.autograd.Function):
    """Tricubic gathers of M fields in one launch; the adjoint is each
    member's deterministic transpose."""

    @staticmethod
    def forward(ctx, lim, coords_list, *fields):
        # departure points through save_for_backward, so that a checkpointed
        # ensemble step drops them after the forward pass
        ctx.save_for_backward(*[c.data for c in coords_list])
        ctx.shares = [c._share for c in coords_list]
        ctx.toks = [sh.register() for sh in ctx.shares]
        ctx.lim = lim
        fields = [f.contiguous() for f in fields]
        outs = [torch.empty_like(f) for f in fields]
        M = len(fields)
        dims = (C.c_int32 * (3 * M))()
        for m, f in enumerate(fields):
            _lib.require_device(f, coords_list[m].data)
            fz, fy, fx = f.shape
            dims[3 * m], dims[3 * m + 1], dims[3 * m + 2] = fx, fy, fz
        L = _lib.lib()
        fn = L.sisl_tricubic_batched_f64 if fields[0].dtype == torch.float64 else L.sisl_tricubic_batched_f32
        _lib.check(fn(M, dims, _lib.ptr_array(fields), _lib.ptr_array([c.data for c in coords_list]),
                      _lib.ptr_array(outs), int(bool(lim)), _lib.stream_ptr(fields[0].device)),
                   "batched tricubic")
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        if ctx.lim:
            raise NotImplementedError("the adjoint of the limited tricubic is not implemented")
        outs = []
        for g, co, sh, tok in zip(grads, ctx.saved_tensors, ctx.shares, ctx.toks):
            if g is None:
                sh.done(tok)
                outs.append(None)
            else:
                outs.append(_tricubic_adjoint(g, co, sh, tok))
        return (None, None) + tuple(outs)


class _BatchedSolve(torch.autograd.Function):
    """x_m = A_m^{-1} b_m for every member with one batched mixed GMRES; the
    cotangents are one batched transposed solve (x0 receives none)."""

    @staticmethod
    def forward(ctx, ens, *bx):
        M = len(bx) // 2
        ctx.ens = ens
        # members have uneven grids: a missing cotangent becomes zeros of its
        # own member's packed length
        ctx.like = [(b.numel(), b.dtype, b.device) for b in bx[:M]]
        return tuple(ens.solve_flat(list(bx[:M]), list(bx[M:]), transpose=False))

    @staticmethod
    def backward(ctx, *gx):
        ens = ctx.ens
        gx = [g.contiguous() if g is not None else torch.zeros(n, dtype=dt, device=dev)
              for g, (n, dt, dev) in zip(gx, ctx.like)]
        lam = ens.solve_flat(gx, [torch.zeros_like(g) for g in gx], transpose=True)
        return (None,) + tuple(lam) + (None,) * len(gx)


class Ensemble:
    """Lockstep batched stepping of ``steppers`` (one ``SISLStepper3D`` per
    member, each with its own grid)."""

    def __init__(self, steppers):
        if not 1 <= len(steppers) <= MAX_MEMBERS:
            raise ValueError(f"an ensemble holds 1..{MAX_MEMBERS} members")
        if any(s.precision != "mixed" for s in steppers):
            raise ValueError("ensemble members must use precision='mixed'")
        if len({float(s.dt) for s in steppers}) != 1:
            raise ValueError("ensemble members must share dt")
        self.steppers = list(steppers)
        self.M = len(steppers)
        self.dt = steppers[0].dt
        self._workspace = None
        self._ws_key = None
        self.last_stats = []

    # -- batched solve -----------------------------------------------------------
    def _prepare(self, bgs):
        key = tuple(id(b) for b in bgs)
        if self._ws_key == key:
            return
        sols = [s.implicit_solver for s in self.steppers]
        for sol, bg in zip(sols, bgs):
            sol.prepare(bg)
        R = sols[0]._restart()
        if any(s._restart() != R for s in sols):
            raise ValueError("ensemble members must share the (inner) restart length")
        L = _lib.lib()
        grids = (_lib.SislGrid * self.M)(*[s.op64.struct.g for s in sols])
        nbytes = int(L.sisl_gmres_batched_workspace_bytes(grids, self.M, R))
        if nbytes <= 0:
            raise ValueError("batched solver: unsupported member grids")
        dev = sols[0].op64.tensors["cua"].device
        self._workspace = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        ops64 = (_lib.SislOperator * self.M)(*[s.op64.struct for s in sols])
        ops32 = (_lib.SislOperator * self.M)(*[s.op32.struct for s in sols])
        _lib.check(L.sisl_gmres_batched_setup(ops64, ops32, self.M, R, self._workspace.data_ptr(),
                                              _lib.stream_ptr(dev)), "batched solver setup")
        self._ws_key = key
        self._R = R

    def solver_outcome(self):
        """Per-member outcome (cycles, Arnoldi steps, residual, threshold) of
        the last batched solve, read from the device."""
        st = (_lib.SislGmresStats * self.M)()
        _lib.check(_lib.lib().sisl_gmres_batched_outcome(self.M, self._workspace.data_ptr(), st,
                                                         _lib.stream_ptr(self._workspace.device)),
                   "batched solver outcome")
        return [{"cycles": s.cycles, "arnoldi_steps": s.arnoldi_steps, "residual": s.residual, "ptol": s.ptol}
                for s in st]

    def capture_step(self, states, t, forcings, bc_fns):
        """Record one lockstep step as a hipGraph (:class:`CapturedStep`)."""
        return CapturedStep(self, states, t, forcings, bc_fns)

    def solve_flat(self, bs, x0s, transpose=False):
        sols = [s.implicit_solver for s in self.steppers]
        xs = [x.detach().clone().contiguous() for x in x0s]
        bs = [b.detach().contiguous() for b in bs]
        p = sols[0]._params(transpose)
        st = (_lib.SislGmresStats * self.M)()
        L = _lib.lib()
        _lib.check(L.sisl_gmres_batched(self.M, C.byref(p), _lib.ptr_array(bs), _lib.ptr_array(xs),
                                        self._workspace.data_ptr(), st, _lib.stream_ptr(bs[0].device)),
                   "batched GMRES")
        self.last_stats.append([{"transpose": bool(transpose), "cycles": s.cycles,
                                 "arnoldi_steps": s.arnoldi_steps, "residual": s.residual, "ptol": s.ptol}
                                for s in st])
        return xs

    def _batched_solve(self, items):
        bgs = [it[1] for it in items]
        self._prepare(bgs)
        bs, x0s, tmpl = [], [], []
        for m, (rhs_prime, bg, x0) in enumerate(items):
            sol = self.steppers[m].implicit_solver
            if torch.is_grad_enabled() and _operator_inputs(sol):
                raise NotImplementedError("ensemble solves do not propagate operator-input cotangents")
            dzw = bg["dz_w_full"]
            rs = {k: (rhs_prime[k] * sol.pi_scale if k == "pi" else rhs_prime[k]) for k in KEYS}
            if x0 is not None:
                x0s_ = {"u": x0["u"], "v": x0["v"], "w": x0["w"], "pi": x0["pi"] * sol.pi_scale,
                        "eta_dot": x0["eta_dot"] * dzw}
            else:
                x0s_ = rs
            bs.append(pack(rs))
            x0s.append(pack(x0s_))
            tmpl.append((rs, dzw, sol.pi_scale))
        xs = _BatchedSolve.apply(self, *bs, *x0s)
        out = []
        for x, (rs, dzw, ps) in zip(xs, tmpl):
            d = unpack(x, rs)
            out.append({"u": d["u"], "v": d["v"], "w": d["w"], "pi": d["pi"] / ps, "eta_dot": d["eta_dot"] / dzw})
        return out

    def _batched_solve_packed(self, items):
        """The members' solves on packed scaled vectors (the fused step's
        form): one batched solve, packed solutions back."""
        self._prepare([it[2] for it in items])
        for m in range(len(items)):
            if torch.is_grad_enabled() and _operator_inputs(self.steppers[m].implicit_solver):
                raise NotImplementedError("ensemble solves do not propagate operator-input cotangents")
        bs = [it[0].contiguous() for it in items]
        x0s = [it[1].contiguous() for it in items]
        return list(_BatchedSolve.apply(self, *bs, *x0s))

    # -- batched trajectories and gathers ------------------------------------------
    def _batched_departure(self, items):
        loc = items[0][1]
        advs = [self.steppers[m].advector for m in range(self.M)]
        dt_ = advs[0].coord_dtype
        us, vs, es, m2s, outs, scrs, grids = [], [], [], [], [], [], []
        for m, (state, _loc) in enumerate(items):
            a = advs[m]
            g = a.grid
            us.append(state["u"].detach().to(dt_).contiguous())
            vs.append(state["v"].detach().to(dt_).contiguous())
            es.append(state["eta_dot"].detach().to(dt_).contiguous())
            m2s.append(g.i_m_factors[loc].detach().to(dt_).contiguous())
            shape = {"m": (g.nz, g.ny, g.nx), "u": (g.nz, g.ny, g.nx + 1),
                     "v": (g.nz, g.ny + 1, g.nx), "w": (g.nz + 1, g.ny, g.nx)}[loc]
            outs.append(torch.empty((3,) + shape, dtype=dt_, device=us[-1].device))
            scrs.append(torch.empty_like(outs[-1]))
            grids.append(a._cgrid)
        L = _lib.lib()
        fn = L.sisl_departure_batched_f64 if dt_ == torch.float64 else L.sisl_departure_batched_f32
        dx = (C.c_double * self.M)(*[a.grid.dx for a in advs])
        dy = (C.c_double * self.M)(*[a.grid.dy for a in advs])
        with torch.no_grad():
            _lib.check(fn(self.M, (_lib.SislGrid * self.M)(*grids), LOC_ID[loc], _lib.ptr_array(us),
                          _lib.ptr_array(vs), _lib.ptr_array(es), _lib.ptr_array(m2s), dx, dy,
                          float(self.dt), 2, _lib.ptr_array(scrs), _lib.ptr_array(outs),
                          _lib.stream_ptr(us[0].device)), "batched departure points")
        return [Coords(o) for o in outs]

    def _batched_tricubic(self, items):
        lim = bool(items[0][2])
        fields, coords, cast = [], [], []
        for field, co, _lim in items:
            if not isinstance(co, Coords):
                co = Coords(co)
            c32 = co.data.dtype == torch.float32 and field.dtype == torch.float64
            fields.append(field.float() if c32 else field)
            cast.append(c32)
            coords.append(co)
        outs = _BatchedTricubic.apply(lim, coords, *fields)
        return [o.double() if c else o for o, c in zip(outs, cast)]

    # -- stepping --------------------------------------------------------------------
    def _install(self, rdv):
        """Route each member's hot calls through the baton (instance
        attributes shadow the class methods; removed by _uninstall)."""
        for m, stp in enumerate(self.steppers):
            adv, sol = stp.advector, stp.implicit_solver

            def dep(state, loc="m", iterations=2, m=m):
                return rdv(m, "departure:" + loc, (state, loc), self._batched_departure)

            def cub(field, coords, use_limiter=False, m=m):
                return rdv(m, "tricubic", (field, coords, use_limiter), self._batched_tricubic)

            def solve(rhs_prime, bg, x0=None, m=m):
                return rdv(m, "solve", (rhs_prime, bg, x0), self._batched_solve)

            def solve_packed(b, x0, bg, m=m):
                return rdv(m, "solve_packed", (b, x0, bg), self._batched_solve_packed)
            adv.compute_departure_indices = dep
            adv.advect_cubic = cub
            sol.solve = solve
            sol.solve_packed = solve_packed

    def _uninstall(self):
        for stp in self.steppers:
            for obj, name in ((stp.advector, "compute_departure_indices"), (stp.advector, "advect_cubic"),
                              (stp.implicit_solver, "solve"), (stp.implicit_solver, "solve_packed")):
                obj.__dict__.pop(name, None)

    def step_internal(self, states, t, forcings, bc_fns):
        """One lockstep step of every member (internal dicts in and out)."""
        baton = _Baton(self.M)
        self._install(baton)
        stream = torch.cuda.current_stream()

        def body(m):
            with torch.cuda.stream(stream):
                return self.steppers[m].step_internal(states[m], t, forcings[m], bc_fns[m])
        try:
            return run_lockstep(baton, body)
        finally:
            self._uninstall()

    def step(self, states, t, forcings, bc_fns):
        """Reference-shaped states in and out, one per member."""
        ins = [state_to_internal(s) for s in states]
        return [state_to_reference(s) for s in self.step_internal(ins, t, forcings, bc_fns)]

    def integrate(self, states, t_start, num_steps, forcings, bc_fns, checkpoint=False):
        """``num_steps`` lockstep steps.  ``checkpoint=True``: each ensemble
        step is one autograd node that keeps only its input states and, in
        the backward pass, recomputes the step in lockstep and differentiates
        it at once (the activation memory of one step, whatever the number of
        steps).  Gradients flow to the member states; tensors captured by the
        forcings or boundary closures receive none."""
        sts = [state_to_internal(s) for s in states]
        for s in sts:
            if "eta_dot" not in s:
                s["eta_dot"] = torch.zeros_like(s["w"])
        if checkpoint:
            for stp, f, b in zip(self.steppers, forcings, bc_fns):
                stp.warm_caches(f, b)
            self._prepare([stp.bg_precomputed() for stp in self.steppers])
        for i in range(int(num_steps)):
            t = t_start + i * self.dt
            if checkpoint and torch.is_grad_enabled():
                spec, flat = _flatten(sts)
                box = {}
                outs = _EnsembleStep.apply(self, t, forcings, bc_fns, spec, box, *flat)
                sts = _unflatten(box["spec"], outs)
            else:
                sts = self.step_internal(sts, t, forcings, bc_fns)
        return [state_to_reference(s) for s in sts]


class CapturedStep:
    """One lockstep step of an :class:`Ensemble` recorded once as a hipGraph
    and replayed: every member's departure points, gathers, fused glue,
    boundaries and the batched solves of the step, with no host work between
    the launches.  The batched solver records every refinement iteration it
    could take (gated on the device once a member has converged), so each
    replay gives the bits of an eager lockstep step.

    ``states``: internal member dicts of a step that is not a first step
    (history keys present); they are copied into static buffers, which
    :meth:`advance` replaces with each step's result.  Recording needs the
    ensemble to have taken an eager step with the same members, forcings and
    boundary functions (workspaces, plans and tuned launch shapes exist).
    No gradients: the recorded step runs without autograd."""

    def __init__(self, ens, states, t, forcings, bc_fns):
        for s in states:
            if float(s.get("is_first_step", 1.0)) == 1.0:
                raise ValueError("record a step after the first one (history keys in place)")
        self.ens = ens
        self.static = [{k: (v.detach().clone() if torch.is_tensor(v) else v) for k, v in s.items()}
                       for s in states]
        self.graph = torch.cuda.CUDAGraph()
        with torch.no_grad():
            with torch.cuda.graph(self.graph):
                self.out = ens.step_internal(self.static, t, forcings, bc_fns)
        self._pairs = []
        for st, out in zip(self.static, self.out):
            for k, v in out.items():
                if torch.is_tensor(v) and torch.is_tensor(st.get(k)):
                    if st[k].shape != v.shape or st[k].dtype != v.dtype:
                        raise ValueError(f"step output {k!r} does not match its input")
                    self._pairs.append((st[k], v))
        ptrs = {d.data_ptr() for d, _ in self._pairs}
        self._alias = [src.data_ptr() in ptrs for _, src in self._pairs]

    def replay(self):
        """Run the recorded step on the static states; the result (internal
        dicts, valid until the next replay) is returned."""
        self.graph.replay()
        return self.out

    def advance(self):
        """Replay, then make the result the static states of the next step
        (outputs that are inputs of this step, the history keys, are copied
        first)."""
        self.graph.replay()
        with torch.no_grad():
            srcs = [src.clone() if a else src for (_, src), a in zip(self._pairs, self._alias)]
            for (dst, _), src in zip(self._pairs, srcs):
                dst.copy_(src)
        return self.static

    def stats(self):
        """The batched solver's per-member outcome of the last solve of the
        last replay (host read)."""
        return self.ens.solver_outcome()


def _flatten(states):
    """Member state dicts -> (spec, tensors): spec keeps the keys and the
    non-tensor entries."""
    spec, flat = [], []
    for s in states:
        keys = sorted(k for k, v in s.items() if torch.is_tensor(v))
        spec.append((keys, {k: v for k, v in s.items() if not torch.is_tensor(v)}))
        flat.extend(s[k] for k in keys)
    return spec, flat


def _unflatten(spec, flat):
    it = iter(flat)
    out = []
    for keys, extra in spec:
        d = dict(extra)
        for k in keys:
            d[k] = next(it)
        out.append(d)
    return out


class _EnsembleStep(torch.autograd.Function):
    """One lockstep ensemble step as a single autograd node: the forward pass
    runs without a graph and saves the input states; the backward pass
    re-runs the step in lockstep with a graph and differentiates it.

    Recomputing inside the node (rather than with torch's non-reentrant
    checkpoint) keeps the steps apart in the backward pass: autograd orders
    ready nodes by a per-thread sequence number, and the member threads of
    every step start that count afresh, so the nodes of different steps would
    otherwise interleave and several recomputed steps would be alive at once."""

    @staticmethod
    def forward(ctx, ens, t, forcings, bc_fns, spec, box, *flat):
        ctx.ens, ctx.t, ctx.forcings, ctx.bc_fns, ctx.spec = ens, t, forcings, bc_fns, spec
        ctx.save_for_backward(*flat)
        with torch.no_grad():
            res = ens.step_internal(_unflatten(spec, flat), t, forcings, bc_fns)
        ospec, oflat = _flatten(res)
        ins = {id(x) for x in flat}
        oflat = [o.clone() if id(o) in ins else o for o in oflat]  # outputs never alias inputs
        box["spec"] = ospec
        return tuple(oflat)

    @staticmethod
    def backward(ctx, *gouts):
        req = ctx.needs_input_grad[6:]
        ins = [x.detach().requires_grad_(bool(r)) for x, r in zip(ctx.saved_tensors, req)]
        with torch.enable_grad():
            res = ctx.ens.step_internal(_unflatten(ctx.spec, ins), ctx.t, ctx.forcings, ctx.bc_fns)
        _, oflat = _flatten(res)
        outs, gs = [], []
        for o, g in zip(oflat, gouts):
            if g is not None and o.requires_grad:
                outs.append(o)
                gs.append(g)
        targets = [x for x, r in zip(ins, req) if r]
        grads = torch.autograd.grad(outs, targets, gs, allow_unused=True) if outs and targets else []
        it = iter(grads)
        return (None,) * 6 + tuple(next(it) if r else None for r in req)
