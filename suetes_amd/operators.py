"""Stagger-aware C-grid operators and the tricubic semi-Lagrangian interpolator.

``CGridOperator3D`` mirrors suetes/regional3d/operators.py:12-132 on internal
(Z, Y, X) tensors: reference axis 0 (x) is tensor dim 2, axis 1 (y) dim 1,
axis 2 (z) dim 0.  ``tensor_product_interp_3d`` (operators.py:146-209) runs
the HIP gather; its backward is the deterministic segmented transpose.
"""
from __future__ import annotations

import ctypes as C
import weakref

import torch

from . import _lib

DIM = {0: 2, 1: 1, 2: 0}


class CGridOperator3D:
    """diff / avg between staggerings with the reference's edge closures."""

    def __init__(self, grid, periodic_axes=()):
        if set(periodic_axes) - {0, 1}:
            raise ValueError("Only horizontal axes 0 (x) and 1 (y) may be periodic")
        self.grid = grid
        self.periodic_axes = frozenset(periodic_axes)
        self.use_stop_grad = True

    def m3(self, loc):
        return self.grid.i_m_factors[loc][None]

    def _stagger(self, f, axis, from_loc, to_loc, op):
        """avg (op 0) / diff (op 1) on a device tensor in one HIP launch
        (forward and transpose); None when the call needs the tensor path
        (host tensors, mixed dtypes)."""
        if not f.is_cuda or f.dim() != 3 or f.dtype not in (torch.float64, torch.float32):
            return None
        d = DIM[axis]
        n = f.shape[d]
        if n < 2:
            return None
        if axis in self.periodic_axes and from_loc == "m" and to_loc in ("u", "v"):
            pad = 2
        elif from_loc == "m" and to_loc in ("u", "v", "w"):
            pad = 1
        else:
            pad = 0
        m3 = None
        if op == 1 and axis in (0, 1):
            m3 = self.grid.i_m_factors[to_loc]
            if m3.dtype != f.dtype or m3.requires_grad:
                return None
        shape = list(f.shape)
        shape[d] = n + 1 if pad else n - 1
        if m3 is not None and tuple(m3.shape) != tuple(shape[1:]):
            return None
        rd = 1.0 / self.grid.delta[axis] if op == 1 else 0.0
        return _Stagger.apply(f, d, pad, op, rd, m3, not self.use_stop_grad, tuple(shape))

    def _apply_padding(self, f, axis, from_loc, to_loc):
        # operators.py:37-63: duplicate the edge values (detached when use_stop_grad)
        if from_loc == "m" and to_loc in ("u", "v", "w"):
            d = DIM[axis]
            n = f.shape[d]
            lo, hi = f.narrow(d, 0, 1), f.narrow(d, n - 1, 1)
            if self.use_stop_grad:
                lo, hi = lo.detach(), hi.detach()
            return torch.cat([lo, f, hi], dim=d)
        return f

    def diff(self, f, axis, from_loc, to_loc):
        # operators.py:65-103
        out = self._stagger(f, axis, from_loc, to_loc, 1)
        if out is not None:
            return out
        d = DIM[axis]
        delta = self.grid.delta[axis]
        n = f.shape[d]
        der = (f.narrow(d, 1, n - 1) - f.narrow(d, 0, n - 1)) / delta
        if axis in self.periodic_axes and from_loc == "m" and to_loc in ("u", "v"):
            b = (f.narrow(d, 0, 1) - f.narrow(d, n - 1, 1)) / delta
            der = torch.cat([b, der, b], dim=d)
        else:
            der = self._apply_padding(der, axis, from_loc, to_loc)
        if axis in (0, 1):
            der = der * self.m3(to_loc)
        return der

    def avg(self, f, axis, from_loc, to_loc):
        # operators.py:105-132
        out = self._stagger(f, axis, from_loc, to_loc, 0)
        if out is not None:
            return out
        d = DIM[axis]
        n = f.shape[d]
        a = 0.5 * (f.narrow(d, 0, n - 1) + f.narrow(d, 1, n - 1))
        if axis in self.periodic_axes and from_loc == "m" and to_loc in ("u", "v"):
            b = 0.5 * (f.narrow(d, 0, 1) + f.narrow(d, n - 1, 1))
            return torch.cat([b, a, b], dim=d)
        return self._apply_padding(a, axis, from_loc, to_loc)


_STAGGER_FNS = {}


def _stagger_fn(transpose, dtype):
    key = (transpose, dtype)
    fn = _STAGGER_FNS.get(key)
    if fn is None:
        sfx = "f64" if dtype == torch.float64 else "f32"
        fn = getattr(_lib.lib(), f"sisl_stagger_T_{sfx}" if transpose else f"sisl_stagger_{sfx}")
        _STAGGER_FNS[key] = fn
    return fn


def _stagger_launch(transpose, src, out_shape, d, pad, op, rd, m3, exact):
    src = src.contiguous()
    out = torch.empty(out_shape, dtype=src.dtype, device=src.device)
    shape = out_shape if transpose else src.shape
    outer = shape[0] if d >= 1 else 1
    if d == 2:
        outer *= shape[1]
    inner = shape[2] if d <= 1 else 1
    if d == 0:
        inner *= shape[1]
    n = shape[d]
    m3p, m3n = (m3.data_ptr(), m3.numel()) if m3 is not None else (None, 0)
    stream = torch.cuda.current_stream(src.device).cuda_stream
    if transpose:
        rc = _stagger_fn(True, src.dtype)(src.data_ptr(), out.data_ptr(), outer, n, inner, op, pad,
                                          int(exact), rd, m3p, m3n, stream)
    else:
        rc = _stagger_fn(False, src.dtype)(src.data_ptr(), out.data_ptr(), outer, n, inner, op, pad,
                                           rd, m3p, m3n, stream)
    if rc:
        _lib.check(rc, "staggered operator")
    return out


class _Stagger(torch.autograd.Function):
    """One staggered average / difference (with edge closure and map factor)
    as a HIP launch; the backward is its gather-form transpose."""

    @staticmethod
    def forward(ctx, f, d, pad, op, rd, m3, exact, out_shape):
        m3c = m3.contiguous() if m3 is not None else None
        ctx.args = (d, pad, op, rd, exact, tuple(f.shape))
        ctx.m3 = m3c
        return _stagger_launch(False, f, out_shape, d, pad, op, rd, m3c, exact)

    @staticmethod
    def backward(ctx, g):
        d, pad, op, rd, exact, in_shape = ctx.args
        gf = _stagger_launch(True, g, in_shape, d, pad, op, rd, ctx.m3, exact)
        return gf, None, None, None, None, None, None, None


# ---------------------------------------------------------------------------
# tricubic interpolation (HIP)
# ---------------------------------------------------------------------------
class TricubicPlan:
    """Destination-sorted segment map of one coordinate set, built once and
    shared by every adjoint that uses the same departure points."""

    def __init__(self, coords: torch.Tensor, dims):
        L = _lib.lib()
        fz, fy, fx = dims
        n = fx * fy * fz
        dev = coords.device
        i32 = dict(dtype=torch.int32, device=dev)
        self.key = torch.empty(n, **i32)
        self.perm = torch.empty(n, **i32)
        self.start = torch.empty(n + 1, **i32)
        count = torch.empty(n, **i32)
        sums = torch.empty(max(1, L.sisl_scan_blocks(n)), **i32)
        fn = L.sisl_tricubic_adj_plan_f64 if coords.dtype == torch.float64 else L.sisl_tricubic_adj_plan_f32
        _lib.check(fn(fx, fy, fz, coords.data_ptr(), self.key.data_ptr(), self.perm.data_ptr(),
                      self.start.data_ptr(), count.data_ptr(), sums.data_ptr(), _lib.stream_ptr(dev)),
                   "tricubic adjoint plan")
        self.dims = (fx, fy, fz)
        self.frac = None

    def ensure_rows(self, coords: torch.Tensor):
        """Per sorted position: the fractions (tx, ty, tz) and the base x of
        the source point (for the sorted-row transpose)."""
        if self.frac is not None:
            return
        fx, fy, fz = self.dims
        n = fx * fy * fz
        self.frac = torch.empty(4 * n, dtype=coords.dtype, device=coords.device)
        L = _lib.lib()
        fn = L.sisl_tricubic_adj_frac_f64 if coords.dtype == torch.float64 else L.sisl_tricubic_adj_frac_f32
        _lib.check(fn(fx, fy, fz, coords.data_ptr(), self.perm.data_ptr(), self.frac.data_ptr(),
                      _lib.stream_ptr(coords.device)), "tricubic adjoint plan fractions")


class _Pending:
    """Token of one gather whose adjoint has not run yet."""
    __slots__ = ("__weakref__",)


class _PlanShare:
    """The adjoint plan of one coordinate set, shared by the adjoints of every
    field interpolated with it.  ``pending`` holds (weakly) a token per gather
    whose adjoint is still to come; the last adjoint releases the plan, so a
    long checkpointed backward pass holds the plans of one step at a time.
    Tokens of gathers whose autograd nodes die without a backward (a
    checkpoint's recomputation, outputs outside the loss) vanish with them."""
    __slots__ = ("plan", "pending")

    def __init__(self):
        self.plan, self.pending = None, weakref.WeakSet()

    def register(self):
        tok = _Pending()
        self.pending.add(tok)
        return tok

    def done(self, tok):
        self.pending.discard(tok)
        if not self.pending:
            self.plan = None


class Coords:
    """Departure points of one staggering: tensor (3, Z, Y, X) of index-space
    offsets x_d - x_a (component 0 along x) plus a lazily built adjoint plan.
    ``absolute()`` returns the reference's coordinates (steppers.py:360)."""

    def __init__(self, data: torch.Tensor, share: "_PlanShare" = None):
        self.data = data
        self._share = share if share is not None else _PlanShare()

    def absolute(self):
        _, nz, ny, nx = self.data.shape
        dev, dt = self.data.device, self.data.dtype
        idx = [torch.arange(nx, device=dev, dtype=dt).reshape(1, 1, -1),
               torch.arange(ny, device=dev, dtype=dt).reshape(1, -1, 1),
               torch.arange(nz, device=dev, dtype=dt).reshape(-1, 1, 1)]
        return torch.stack([self.data[a] + idx[a] for a in range(3)])

    def plan(self):
        sh = self._share
        if sh.plan is None:
            sh.plan = TricubicPlan(self.data, tuple(self.data.shape[1:]))
        return sh.plan


def _tricubic_fwd(field, coords, use_limiter):
    _lib.require_device(field, coords)
    if coords.dtype != field.dtype or tuple(coords.shape) != (3,) + tuple(field.shape):
        raise ValueError("coords must be (3,)+field.shape with the field's dtype")
    fz, fy, fx = field.shape
    out = torch.empty_like(field)
    L = _lib.lib()
    fn = L.sisl_tricubic_f64 if field.dtype == torch.float64 else L.sisl_tricubic_f32
    _lib.check(fn(fx, fy, fz, field.data_ptr(), coords.data_ptr(), out.data_ptr(),
                  int(bool(use_limiter)), _lib.stream_ptr(field.device)), "tricubic")
    return out


def _tricubic_adjoint(g, data, share, tok):
    """Adjoint of one gather (token ``tok``) with departure offsets ``data``:
    the plan comes from (or goes to) ``share``, which drops it after the last
    pending adjoint."""
    gf = tricubic_transpose(g.contiguous(), Coords(data, share))
    share.done(tok)
    return gf


def _limited_adjoint(g, field, data, share, tok):
    """Adjoint of one limited gather: the interpolant's share of the
    cotangent through the unlimited transpose, the 8 corners' shares
    gathered per cell over the same plan (both deterministic)."""
    g = g.contiguous()
    fz, fy, fx = g.shape
    L = _lib.lib()
    sfx = "f64" if g.dtype == torch.float64 else "f32"
    gi = torch.empty_like(g)
    gc = torch.empty((8,) + tuple(g.shape), dtype=g.dtype, device=g.device)
    s = _lib.stream_ptr(g.device)
    _lib.check(getattr(L, f"sisl_tricubic_limited_adj_split_{sfx}")(
        fx, fy, fz, field.data_ptr(), data.data_ptr(), g.data_ptr(), gi.data_ptr(), gc.data_ptr(), s),
        "limited tricubic adjoint")
    co = Coords(data, share)
    gf = tricubic_transpose(gi, co)
    plan = co.plan()
    _lib.check(getattr(L, f"sisl_tricubic_limited_adj_corners_{sfx}")(
        fx, fy, fz, plan.perm.data_ptr(), plan.start.data_ptr(), gc.data_ptr(), gf.data_ptr(), s),
        "limited tricubic adjoint corners")
    share.done(tok)
    return gf


class _Tricubic(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, coords_obj, use_limiter):
        # the offsets go through save_for_backward (a checkpointed step drops
        # them after its forward pass); only the small plan holder stays.
        # The limited gather's adjoint also needs the field (its clip
        # branches and corner extrema).
        field = field.contiguous()
        if use_limiter:
            ctx.save_for_backward(coords_obj.data, field)
        else:
            ctx.save_for_backward(coords_obj.data)
        ctx.share = coords_obj._share
        ctx.tok = ctx.share.register()
        ctx.use_limiter = use_limiter
        return _tricubic_fwd(field, coords_obj.data, use_limiter)

    @staticmethod
    def backward(ctx, g):
        if ctx.use_limiter:
            data, field = ctx.saved_tensors
            return _limited_adjoint(g, field, data, ctx.share, ctx.tok), None, None
        (data,) = ctx.saved_tensors
        return _tricubic_adjoint(g, data, ctx.share, ctx.tok), None, None


def _tricubic_multi_fwd(fields, data, use_limiter):
    """One launch over the fields sharing the offsets ``data``: FP64 fields
    with FP32 offsets gather in FP32 and return FP64 (mixed precision)."""
    _lib.require_device(data, *fields)
    f0 = fields[0]
    mixed = data.dtype == torch.float32 and f0.dtype == torch.float64
    for f in fields:
        if f.dtype != f0.dtype or tuple(data.shape) != (3,) + tuple(f.shape):
            raise ValueError("fields must share dtype and shape, coords must be (3,)+field.shape")
    if not mixed and data.dtype != f0.dtype:
        raise ValueError("coords must have the fields' dtype (or FP32 offsets with FP64 fields)")
    fz, fy, fx = f0.shape
    outs = [torch.empty_like(f) for f in fields]
    L = _lib.lib()
    sfx = "mixed" if mixed else ("f64" if f0.dtype == torch.float64 else "f32")
    s = _lib.stream_ptr(f0.device)
    lim = int(bool(use_limiter))
    if len(fields) == 1:
        fn = L.sisl_tricubic_mixed if mixed else getattr(L, f"sisl_tricubic_{sfx}")
        _lib.check(fn(fx, fy, fz, f0.data_ptr(), data.data_ptr(), outs[0].data_ptr(), lim, s), "tricubic")
    elif len(fields) == 2:
        _lib.check(getattr(L, f"sisl_tricubic2_{sfx}")(fx, fy, fz, fields[0].data_ptr(), fields[1].data_ptr(),
                                                        data.data_ptr(), outs[0].data_ptr(), outs[1].data_ptr(),
                                                        lim, s), "tricubic (two fields)")
    else:
        raise ValueError("one or two fields per launch")
    return outs


class _TricubicN(torch.autograd.Function):
    """One or two fields gathered at one set of departure points in one
    launch (FP64 fields with FP32 offsets: the mixed path, gathered in FP32).
    Each field's adjoint is the deterministic transpose of its own gather
    over the shared plan (FP32 in the mixed path, as the separate cast,
    gather and widening would be differentiated)."""

    @staticmethod
    def forward(ctx, coords_obj, use_limiter, *fields):
        fields = [f.contiguous() for f in fields]
        data = coords_obj.data
        ctx.set_materialize_grads(False)
        ctx.share = coords_obj._share
        ctx.toks = [ctx.share.register() for _ in fields]
        ctx.use_limiter = use_limiter
        if use_limiter:
            ctx.save_for_backward(data, *fields)
        else:
            ctx.save_for_backward(data)
        return tuple(_tricubic_multi_fwd(fields, data, use_limiter))

    @staticmethod
    def backward(ctx, *gs):
        saved = ctx.saved_tensors
        data, fields = saved[0], saved[1:]
        out = []
        for q, g in enumerate(gs):
            tok = ctx.toks[q]
            if g is None or not ctx.needs_input_grad[2 + q]:
                ctx.share.done(tok)
                out.append(None)
                continue
            mixed = data.dtype == torch.float32 and g.dtype == torch.float64
            gg = g.float() if mixed else g
            if ctx.use_limiter:
                f = fields[q].float() if mixed else fields[q]
                gf = _limited_adjoint(gg, f, data, ctx.share, tok)
            else:
                gf = _tricubic_adjoint(gg, data, ctx.share, tok)
            out.append(gf.double() if mixed else gf)
        return (None, None) + tuple(out)


def tensor_product_interp_pair(f0, f1, coords, use_limiter=False):
    """``tensor_product_interp_3d`` of two fields on the same departure
    points in one launch: the same bits as two separate calls."""
    if not isinstance(coords, Coords):
        coords = Coords(coords)
    return _TricubicN.apply(coords, use_limiter, f0, f1)


def tensor_product_interp_3d(field, coords, use_limiter=False):
    """Tricubic Catmull-Rom gather at departure points (operators.py:146-209).

    ``field``: internal (Z, Y, X) tensor; ``coords``: a :class:`Coords` (or a
    (3, Z, Y, X) tensor) of departure offsets x_d - x_a in index space.  When coords are
    FP32 and the field FP64 the interpolation runs in FP32 and the result is
    returned in FP64 (mixed precision)."""
    if not isinstance(coords, Coords):
        coords = Coords(coords)
    if coords.data.dtype == torch.float32 and field.dtype == torch.float64:
        return _TricubicN.apply(coords, use_limiter, field)[0]
    return _Tricubic.apply(field, coords, use_limiter)


def tricubic_transpose(g, coords, method="rows"):
    """Explicit T^T g (deterministic), the adjoint of the unlimited gather.
    ``method='rows'``: the sorted-row kernel (plan fractions built once per
    coordinate set); ``'segments'``: the per-destination segment gather."""
    if not isinstance(coords, Coords):
        coords = Coords(coords)
    _lib.require_device(g, coords.data)
    plan = coords.plan()
    g = g.contiguous()
    gf = torch.empty_like(g)
    fz, fy, fx = g.shape
    L = _lib.lib()
    sfx = "f64" if g.dtype == torch.float64 else "f32"
    if method == "rows":
        plan.ensure_rows(coords.data)
        gs = torch.empty_like(g)
        _lib.check(getattr(L, f"sisl_tricubic_adj_rows_{sfx}")(
            fx, fy, fz, plan.perm.data_ptr(), plan.start.data_ptr(), plan.frac.data_ptr(),
            g.data_ptr(), gs.data_ptr(), gf.data_ptr(), _lib.stream_ptr(g.device)),
            "tricubic adjoint")
    elif method == "segments":
        _lib.check(getattr(L, f"sisl_tricubic_adj_{sfx}")(
            fx, fy, fz, coords.data.data_ptr(), plan.perm.data_ptr(), plan.start.data_ptr(), g.data_ptr(),
            gf.data_ptr(), _lib.stream_ptr(g.device)), "tricubic adjoint")
    else:
        raise ValueError(f"unknown method {method!r}")
    return gf
