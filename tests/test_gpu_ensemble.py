"""Batched ensemble on uneven grids: one launch per kernel over members,
checked against the same members stepped one by one.

Bars: batched departure points and tricubic gathers are bit-identical to the
single-member kernels; the batched mixed GMRES reaches every member's
stopping threshold and agrees with the single-member mixed solve to 1e-7;
lockstep steps agree with per-member steps to 1e-6 (the batched solver runs
the same algorithm with its own rounding); gradients through checkpointed
lockstep steps agree with per-member gradients to 1e-5.
"""
import numpy as np
import pytest
import torch

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _cases(M=3):
    from suetes_amd import workloads as W
    from suetes_amd.cases import build_case
    cfgs = W.ensemble_members(seed=7, sizes=[(24, 8), (32, 12), (16, 10)][:M])
    out = []
    for c in cfgs:
        c["boundary"] = ("davies", dict(sponge_depth=4, dt=10.0, outflow_factor=1.0))
        c["core"].update(solver_tol=1e-10, solver_maxiter=20)
        out.append(build_case(c, device=DEV, precision="mixed"))
    return out


def _ens(cases):
    from suetes_amd.ensemble import Ensemble
    return Ensemble([c.stepper for c in cases])


def test_member_grids_are_uneven():
    cases = _cases()
    shapes = {(c.grid.nx, c.grid.ny, c.grid.nz) for c in cases}
    assert len(shapes) == len(cases)


def test_batched_departure_and_tricubic_bit_identical():
    from suetes_amd.steppers import state_to_internal
    cases = _cases()
    ens = _ens(cases)
    states = [state_to_internal(c.state) for c in cases]
    for loc in ("u", "v", "w", "m"):
        got = ens._batched_departure([(s, loc) for s in states])
        for c, s, co in zip(cases, states, got):
            ref = c.stepper.advector.compute_departure_indices(s, loc)
            assert torch.equal(co.data, ref.data), loc
    from suetes_amd.operators import tensor_product_interp_3d
    dep = [c.stepper.advector.compute_departure_indices(s, "m") for c, s in zip(cases, states)]
    for lim in (False, True):
        items = [(s["th_v"], co, lim) for s, co in zip(states, dep)]
        got = ens._batched_tricubic(items)
        for (f, co, _), g in zip(items, got):
            assert torch.equal(g, tensor_product_interp_3d(f, co, lim))


def test_batched_solver_matches_single_member_solves():
    cases = _cases()
    ens = _ens(cases)
    bgs = [c.stepper.bg_precomputed() for c in cases]
    ens._prepare(bgs)
    rng = np.random.default_rng(4)
    bs, refs = [], []
    for c in cases:
        sol = c.stepper.implicit_solver
        n = sum(int(np.prod(s)) for s in _packed_shapes(c.grid))
        b = sol.apply_A(torch.from_numpy(rng.standard_normal(n)).to(DEV))
        bs.append(b)
        refs.append(sol.solve_flat(b, torch.zeros_like(b)))
    for transpose in (False, True):
        xs = ens.solve_flat(bs, [torch.zeros_like(b) for b in bs], transpose=transpose)
        for m, (c, x, st) in enumerate(zip(cases, xs, ens.last_stats[-1])):
            assert st["residual"] <= st["ptol"], (m, st)
            sol = c.stepper.implicit_solver
            r = sol.apply_A(x, transpose=transpose) - bs[m]
            assert float(r.norm() / bs[m].norm()) < 1e-8
            if not transpose:
                assert float((x - refs[m]).norm() / refs[m].norm()) < 1e-7


def test_batched_cycle_graph_replays_match_eager_launches():
    """The batched solver's inner cycle (Arnoldi steps with the fused basis
    normalisation, least squares, update of x) replayed from a captured
    graph gives the same bits as eager launches, on every replay, in both
    directions, with eager GPU work and synchronisations between solves."""
    from suetes_amd import _lib
    cases = _cases()
    ens = _ens(cases)
    ens._prepare([c.stepper.bg_precomputed() for c in cases])
    rng = np.random.default_rng(11)
    bs = []
    for c in cases:
        n = sum(int(np.prod(s)) for s in _packed_shapes(c.grid))
        bs.append(c.stepper.implicit_solver.apply_A(torch.from_numpy(rng.standard_normal(n)).to(DEV)))
    L = _lib.lib()
    try:
        for transpose in (False, True):
            L.sisl_set_graphs(0)
            ref = ens.solve_flat(bs, [torch.zeros_like(b) for b in bs], transpose=transpose)
            ref_st = ens.last_stats[-1]
            L.sisl_set_graphs(1)
            for rep in range(3):
                junk = torch.randn(1 << 20, device=DEV).sum()  # eager work between replays
                torch.cuda.synchronize()
                xs = ens.solve_flat(bs, [torch.zeros_like(b) for b in bs], transpose=transpose)
                torch.cuda.synchronize()
                for m, (x, r) in enumerate(zip(xs, ref)):
                    assert torch.equal(x, r), (transpose, rep, m, (x - r).abs().max().item())
                assert [s["arnoldi_steps"] for s in ens.last_stats[-1]] == \
                    [s["arnoldi_steps"] for s in ref_st], (transpose, rep)
                del junk
    finally:
        L.sisl_set_graphs(1)


def _packed_shapes(g):
    return [(g.nz, g.ny, g.nx + 1), (g.nz, g.ny + 1, g.nx), (g.nz + 1, g.ny, g.nx), (g.nz, g.ny, g.nx),
            (g.nz + 1, g.ny, g.nx)]


def test_lockstep_steps_match_member_steps():
    cases = _cases()
    ens = _ens(cases)
    a = [dict(c.state) for c in cases]
    b = [dict(c.state) for c in cases]
    for i in range(2):
        a = ens.step(a, i * ens.dt, [c.forcing for c in cases], [c.bc_fn for c in cases])
        b = [c.stepper.step(s, i * c.dt, c.forcing, c.bc_fn) for c, s in zip(cases, b)]
    for m in range(len(cases)):
        for k in ("u", "v", "w", "pi", "th_v", "rho"):
            e = rel_l2(a[m][k].detach().cpu().numpy(), b[m][k].detach().cpu().numpy())
            assert e < 1e-6, (m, k, e)


def test_checkpointed_lockstep_gradient_matches_members():
    """d/d(initial theta) of sum_m mean(w_m^2) after 2 lockstep steps, one
    checkpoint per ensemble step, against each member's own gradient."""
    cases = _cases()
    ens = _ens(cases)
    th0 = [c.state["th_v"].detach().clone().requires_grad_(True) for c in cases]
    sts = []
    for c, t in zip(cases, th0):
        s = dict(c.state)
        s["th_v"] = t
        sts.append(s)
    out = ens.integrate(sts, 0.0, 2, [c.forcing for c in cases], [c.bc_fn for c in cases], checkpoint=True)
    J = sum((o["w"] ** 2).mean() for o in out)
    grads = torch.autograd.grad(J, th0)
    for m, c in enumerate(cases):
        t = c.state["th_v"].detach().clone().requires_grad_(True)
        s = dict(c.state)
        s["th_v"] = t
        o = c.stepper.integrate(s, 0.0, 2, c.forcing, c.bc_fn)
        (g,) = torch.autograd.grad((o["w"] ** 2).mean(), t)
        assert float((grads[m] - g).norm() / g.norm()) < 1e-5, m


def test_captured_lockstep_step_replays_match_eager_steps():
    """One lockstep step recorded as a hipGraph (departure points, gathers,
    fused glue, boundaries, the batched solves with every refinement
    iteration they could take) and replayed three times, with eager GPU work
    and host synchronisation between the replays, gives the bits of three
    eager lockstep steps; every member's solves converged."""
    from suetes_amd.steppers import state_to_internal
    cases = _cases()
    ens = _ens(cases)
    forc, bcs = [c.forcing for c in cases], [c.bc_fn for c in cases]
    dt = cases[0].stepper.dt
    sts = [state_to_internal(c.state) for c in cases]
    for s in sts:
        s.setdefault("eta_dot", torch.zeros_like(s["w"]))
    with torch.no_grad():
        sts = ens.step_internal(sts, 0.0, forc, bcs)  # first step: history in place
        sts = ens.step_internal(sts, dt, forc, bcs)   # every lazy workspace and plan exists
        eager, cur = [], sts
        for i in range(3):
            cur = ens.step_internal(cur, (2 + i) * dt, forc, bcs)
            eager.append([{k: v.clone() for k, v in d.items() if torch.is_tensor(v)} for d in cur])
        cap = ens.capture_step(sts, 2 * dt, forc, bcs)
        for i in range(3):
            got = cap.advance()
            torch.cuda.synchronize()
            x = torch.randn(4096, device=DEV, dtype=torch.float64)  # eager work between replays
            assert float((x * x).sum()) > 0.0
            for m, (g, e) in enumerate(zip(got, eager[i])):
                for k, v in e.items():
                    assert torch.equal(g[k], v), (i, m, k, float((g[k] - v).abs().max()))
        for st in cap.stats():
            assert st["cycles"] >= 1 and st["residual"] <= st["ptol"], st
