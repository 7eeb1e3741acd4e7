"""Host-side lockstep schedule of the ensemble (CPU, toy members).

The member threads must run in the fixed order 0..M-1 between batched calls,
see the caller's grad mode and saved-tensor hooks (so that a non-reentrant
checkpoint around an ensemble step stores nothing of the members' graphs),
and propagate the first failure to the caller without deadlock."""
import threading

import pytest
import torch
from torch.utils.checkpoint import checkpoint

from suetes_amd.ensemble import _Baton, run_lockstep


class _Scale(torch.autograd.Function):
    """y_m = c_m * x_m for all members at once (stands in for a batched launch)."""

    @staticmethod
    def forward(ctx, cs, *xs):
        ctx.save_for_backward(*xs)
        ctx.cs = cs
        return tuple(c * x for c, x in zip(cs, xs))

    @staticmethod
    def backward(ctx, *gs):
        return (None,) + tuple(c * g for c, g in zip(ctx.cs, gs))


def _toy_step(xs, order=None):
    """Each member: tanh, batched scale, sin * x, batched scale again."""
    M = len(xs)
    baton = _Baton(M)
    cs = [1.0 + 0.5 * m for m in range(M)]

    def batched(items):
        return _Scale.apply(cs, *items)

    def body(m):
        if order is not None:
            order.append(m)
        a = torch.tanh(xs[m]) * xs[m]
        b = baton(m, "scale", a, batched)
        if order is not None:
            order.append(m)
        c = torch.sin(b) * b
        return baton(m, "scale", c, batched)
    return run_lockstep(baton, body)


def _serial_step(xs):
    cs = [1.0 + 0.5 * m for m in range(len(xs))]
    a = [torch.tanh(x) * x for x in xs]
    b = [c * t for c, t in zip(cs, a)]
    d = [torch.sin(t) * t for t in b]
    return [c * t for c, t in zip(cs, d)]


def _inputs(M=4, n=64):
    g = torch.Generator().manual_seed(3)
    return [torch.randn(n + 7 * m, dtype=torch.float64, generator=g).requires_grad_(True) for m in range(M)]


def test_members_run_in_fixed_order():
    order = []
    xs = _inputs()
    with torch.no_grad():
        _toy_step(xs, order)
    assert order == [0, 1, 2, 3, 0, 1, 2, 3]


def test_results_and_gradients_match_serial():
    xs = _inputs()
    ys = _toy_step(xs)
    ref = _serial_step(xs)
    for y, r in zip(ys, ref):
        assert torch.equal(y, r)
    J = sum((y ** 2).sum() for y in ys)
    Jr = sum((r ** 2).sum() for r in ref)
    g = torch.autograd.grad(J, xs)
    gr = torch.autograd.grad(Jr, xs)
    for a, b in zip(g, gr):
        assert torch.allclose(a, b, rtol=1e-14, atol=0)


def test_member_threads_see_the_callers_saved_tensor_hooks():
    packed = {"n": 0, "threads": set()}

    def pack(t):
        packed["n"] += 1
        packed["threads"].add(threading.get_ident())
        return t

    xs = _inputs()
    with torch.autograd.graph.saved_tensors_hooks(pack, lambda t: t):
        _toy_step(xs)
    n_lock = packed["n"]
    packed["n"] = 0
    with torch.autograd.graph.saved_tensors_hooks(pack, lambda t: t):
        _serial_step(xs)
    # the batched stand-in saves its M inputs in each of its 2 calls
    assert n_lock == packed["n"] + 2 * len(xs)
    assert packed["n"] > 0
    assert len(packed["threads"]) > 1   # the members packed from their own threads


@pytest.mark.parametrize("steps", [1, 3])
def test_checkpointed_lockstep_gradient(steps):
    xs = _inputs()

    def fn(*a):
        return tuple(_toy_step(list(a)))
    ys = list(xs)
    for _ in range(steps):
        ys = list(checkpoint(fn, *ys, use_reentrant=False))
    J = sum((y ** 2).sum() for y in ys)
    g = torch.autograd.grad(J, xs)
    ref = list(xs)
    for _ in range(steps):
        ref = _serial_step(ref)
    gr = torch.autograd.grad(sum((r ** 2).sum() for r in ref), xs)
    for a, b in zip(g, gr):
        assert torch.allclose(a, b, rtol=1e-13, atol=0)


def test_failure_in_a_member_reaches_the_caller():
    M = 3
    baton = _Baton(M)

    def body(m):
        if m == 1:
            raise ValueError("member 1 failed")
        return baton(m, "x", m, lambda items: items)
    with pytest.raises(ValueError, match="member 1 failed"):
        run_lockstep(baton, body)


def test_failure_in_the_batched_call_reaches_the_caller():
    baton = _Baton(3)

    def fail(items):
        raise ArithmeticError("batched call failed")
    with pytest.raises(ArithmeticError, match="batched call failed"):
        run_lockstep(baton, lambda m: baton(m, "x", m, fail))


def test_diverging_members_are_reported():
    baton = _Baton(2)
    with pytest.raises(RuntimeError, match="members diverged"):
        run_lockstep(baton, lambda m: baton(m, "a" if m == 0 else "b", m, lambda items: items))


class _ToyEnsemble:
    """Stands in for Ensemble: step_internal over member dicts in lockstep
    (field 'c' passes through unchanged, 'x' is stepped)."""

    def step_internal(self, states, t, forcings, bc_fns):
        xs = _toy_step([s["x"] for s in states])
        return [{"x": x * f, "c": s["c"], "n": s["n"] + 1} for x, s, f in zip(xs, states, forcings)]


@pytest.mark.parametrize("steps", [1, 3])
def test_recomputing_step_node_gradient(steps):
    from suetes_amd.ensemble import _EnsembleStep, _flatten, _unflatten
    ens = _ToyEnsemble()
    xs = _inputs(M=3)
    cs = [torch.full_like(x, 0.5).requires_grad_(True) for x in xs]
    forc = [1.0, 0.9, 1.1]
    sts = [{"x": x, "c": c, "n": 0} for x, c in zip(xs, cs)]
    for i in range(steps):
        spec, flat = _flatten(sts)
        box = {}
        outs = _EnsembleStep.apply(ens, float(i), forc, None, spec, box, *flat)
        sts = _unflatten(box["spec"], outs)
    assert all(s["n"] == steps for s in sts)
    J = sum((s["x"] ** 2).sum() + (s["c"] * s["x"]).sum() for s in sts)
    g = torch.autograd.grad(J, xs + cs)
    ref = [{"x": x, "c": c, "n": 0} for x, c in zip(xs, cs)]
    for i in range(steps):
        ref = ens.step_internal(ref, float(i), forc, None)
    Jr = sum((s["x"] ** 2).sum() + (s["c"] * s["x"]).sum() for s in ref)
    gr = torch.autograd.grad(Jr, xs + cs)
    for a, b in zip(g, gr):
        assert torch.allclose(a, b, rtol=1e-13, atol=1e-300)
