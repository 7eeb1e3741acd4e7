"""GPU parity of each HIP kernel against the FP64 oracle, through the C ABI.

Tolerances: FP64 kernels agree with the oracle to rel-L2 <= 1e-12 (rounding
order differs); FP32 kernels to <= 2e-6; transposes pass dot-product tests to
1e-12 (FP64); the tricubic adjoint is bit-reproducible run to run.
"""
import numpy as np
import pytest
import torch

from oracle import sisl as O
from tests.helpers import oracle_case, rel_l2, small_bubble, small_lam, small_schaer

pytestmark = pytest.mark.gpu
DEV = "cuda"


def cases():
    from suetes_amd.cases import build_case
    return build_case


def to_int(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0))).to(DEV, dtype)


def from_int(t):
    return t.detach().cpu().numpy().transpose(2, 1, 0)


def index_grid(shape):
    return np.stack(np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij"))


def offsets_int(pts, dtype=torch.float64):
    """Absolute index-space points (3, x, y, z) -> offsets x_d - x_a as a
    (3, Z, Y, X) device tensor (the kernels' departure-point format)."""
    off = pts - index_grid(pts.shape[1:])
    return torch.stack([to_int(off[a], dtype) for a in range(3)])


def pack_ref(d):
    return torch.cat([to_int(d[k]).reshape(-1) for k in O.KEYS])


def unpack_to_ref(x, like_ref):
    out, o = {}, 0
    for k in O.KEYS:
        shp = like_ref[k].shape
        n = int(np.prod(shp))
        out[k] = x[o:o + n].view(shp[2], shp[1], shp[0]).cpu().numpy().transpose(2, 1, 0)
        o += n
    return out


@pytest.mark.parametrize("shape", [(9, 7, 6), (10, 7, 6), (9, 8, 6), (9, 7, 7), (1, 5, 4)])
@pytest.mark.parametrize("lim", [False, True])
def test_tricubic_matches_oracle(shape, lim):
    from suetes_amd.operators import tensor_product_interp_3d
    rng = np.random.default_rng(hash(shape) % 1000 + lim)
    f = rng.standard_normal(shape)
    pts = np.stack([rng.uniform(-2, s + 1, size=shape) for s in shape])
    pts[:, 0, 0, 0] = [0.0, 0.0, 0.0]
    pts[:, -1, -1, -1] = [s - 1.0 for s in shape]
    ref = O.tricubic(f, pts, lim)
    co = offsets_int(pts)
    out = tensor_product_interp_3d(to_int(f), co, lim)
    assert rel_l2(from_int(out), ref) < 1e-13
    out32 = tensor_product_interp_3d(to_int(f, torch.float32), co.float(), lim)
    assert rel_l2(from_int(out32), ref) < 2e-6


@pytest.mark.parametrize("lim", [False, True])
def test_tricubic_staged_and_far_departures_match_oracle(lim):
    """Tiles of the staged gather (64 x 4 columns, level chunks) with both
    near departure points (read from the staged planes) and far ones (read
    from global memory), on a grid spanning several tiles."""
    from suetes_amd.operators import tensor_product_interp_3d
    rng = np.random.default_rng(40 + lim)
    shape = (70, 9, 17)
    f = rng.standard_normal(shape)
    base = index_grid(shape)
    near = rng.uniform(-0.9, 0.9, size=(3,) + shape)
    far = rng.uniform(-6.0, 6.0, size=(3,) + shape)
    pick = rng.random(shape) < 0.2
    pts = base + np.where(pick[None], far, near)
    ref = O.tricubic(f, pts, lim)
    co = offsets_int(pts)
    out = tensor_product_interp_3d(to_int(f), co, lim)
    assert rel_l2(from_int(out), ref) < 1e-13
    out32 = tensor_product_interp_3d(to_int(f, torch.float32), co.float(), lim)
    assert rel_l2(from_int(out32), ref) < 2e-6


@pytest.mark.parametrize("lim", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32", "mixed"])
def test_tricubic_pair_and_mixed_storage_bit_identical(prec, lim):
    """Two fields gathered in one launch on shared departure points give the
    same bits as two single gathers, forward and adjoint; the mixed gather
    (FP64 fields, FP32 offsets and arithmetic) gives the same bits as an FP64
    -> FP32 cast, an FP32 gather and the widening of its result.  On a grid
    of several tiles mixing staged and far departures; the FP64 pair also
    against the oracle."""
    from suetes_amd import operators as ops
    rng = np.random.default_rng(71 + lim)
    shape = (70, 9, 17)
    base = index_grid(shape)
    near = rng.uniform(-0.9, 0.9, size=(3,) + shape)
    far = rng.uniform(-6.0, 6.0, size=(3,) + shape)
    pts = base + np.where((rng.random(shape) < 0.2)[None], far, near)
    fa, fb = rng.standard_normal(shape), rng.standard_normal(shape)
    ga, gb = to_int(rng.standard_normal(shape)), to_int(rng.standard_normal(shape))
    fdt = torch.float32 if prec == "f32" else torch.float64
    cdt = torch.float64 if prec == "f64" else torch.float32
    co = offsets_int(pts, cdt)

    def single(f, g):
        x = to_int(f, fdt).requires_grad_(True)
        if prec == "mixed":
            y = ops.tensor_product_interp_3d(x.float(), ops.Coords(co), lim).double()
        else:
            y = ops.tensor_product_interp_3d(x, ops.Coords(co), lim)
        (gx,) = torch.autograd.grad(y, x, g.to(fdt))
        return y.detach(), gx

    ya, gxa = single(fa, ga)
    yb, gxb = single(fb, gb)
    xa, xb = to_int(fa, fdt).requires_grad_(True), to_int(fb, fdt).requires_grad_(True)
    pa, pb = ops.tensor_product_interp_pair(xa, xb, ops.Coords(co), lim)
    assert torch.equal(pa, ya) and torch.equal(pb, yb)
    ha, hb = torch.autograd.grad((pa, pb), (xa, xb), (ga.to(fdt), gb.to(fdt)))
    assert torch.equal(ha, gxa) and torch.equal(hb, gxb)
    # one output unused: its field gets no cotangent, the other the single one
    xa2 = to_int(fa, fdt).requires_grad_(True)
    qa, _ = ops.tensor_product_interp_pair(xa2, to_int(fb, fdt), ops.Coords(co), lim)
    (ka,) = torch.autograd.grad(qa, xa2, ga.to(fdt))
    assert torch.equal(ka, gxa)
    # one field without a gradient: only the other's transpose runs
    from suetes_amd import _lib
    xa3 = to_int(fa, fdt).requires_grad_(True)
    ra, rb = ops.tensor_product_interp_pair(xa3, to_int(fb, fdt), ops.Coords(co), lim)
    _lib.timer_enable(True)
    (la,) = torch.autograd.grad((ra, rb), xa3, (ga.to(fdt), gb.to(fdt)), allow_unused=True)
    torch.cuda.synchronize()
    _lib.timer_enable(False)
    assert torch.equal(la, gxa)
    if not lim:
        assert _lib.timer_read("tricubic_adj", fdt == torch.float32 or prec == "mixed")[1] == 1
    if prec == "mixed":  # the one-field mixed launch
        xm = to_int(fa).requires_grad_(True)
        ym = ops.tensor_product_interp_3d(xm, ops.Coords(co), lim)
        assert ym.dtype == torch.float64 and torch.equal(ym, ya)
        (gm,) = torch.autograd.grad(ym, xm, ga)
        assert torch.equal(gm, gxa)
    if prec == "f64":
        assert rel_l2(from_int(pa), O.tricubic(fa, pts, lim)) < 1e-13
        assert rel_l2(from_int(pb), O.tricubic(fb, pts, lim)) < 1e-13


@pytest.mark.parametrize("checkpointed", [False, True])
def test_tricubic_adjoint_plan_shared_then_released(monkeypatch, checkpointed):
    """Fields gathered at one coordinate set share one adjoint plan, built
    at the first adjoint and released by the last (also when a checkpoint
    recomputes the gathers, whose recomputed nodes never run a backward);
    the gradients are the explicit transposes."""
    from torch.utils.checkpoint import checkpoint
    from suetes_amd import operators as ops
    rng = np.random.default_rng(13)
    shape = (10, 9, 8)
    pts = np.stack([rng.uniform(-1, s, size=shape) for s in shape])
    co = ops.Coords(offsets_int(pts))
    fs = [to_int(rng.standard_normal(shape)).requires_grad_(True) for _ in range(3)]
    gs = [to_int(rng.standard_normal(shape)) for _ in range(3)]
    built = []
    orig = ops.TricubicPlan.__init__

    def counting(self, *a, **k):
        built.append(1)
        orig(self, *a, **k)
    monkeypatch.setattr(ops.TricubicPlan, "__init__", counting)

    def objective(*fields):
        return sum((ops.tensor_product_interp_3d(f, co) * g).sum() for f, g in zip(fields, gs))
    J = checkpoint(objective, *fs, use_reentrant=False) if checkpointed else objective(*fs)
    assert len(co._share.pending) == 3 and co._share.plan is None
    grads = torch.autograd.grad(J, fs)
    assert len(built) == 1
    assert len(co._share.pending) == 0 and co._share.plan is None
    for gr, g in zip(grads, gs):
        assert torch.equal(gr, ops.tricubic_transpose(g, ops.Coords(co.data.clone())))


@pytest.mark.parametrize("method", ["rows", "segments"])
def test_tricubic_adjoint_dot_and_determinism(method):
    from functools import partial

    from suetes_amd.operators import Coords, tensor_product_interp_3d
    from suetes_amd.operators import tricubic_transpose as tt
    tricubic_transpose = partial(tt, method=method)
    rng = np.random.default_rng(11)
    shape = (13, 11, 9)
    pts = np.stack([rng.uniform(-1, s, size=shape) for s in shape])
    pts[:, :2] = np.round(pts[:, :2])  # clamped edges and integer points
    co = Coords(offsets_int(pts))
    f = to_int(rng.standard_normal(shape))
    g = to_int(rng.standard_normal(shape))
    Tf = tensor_product_interp_3d(f, co)
    Ttg = tricubic_transpose(g, co)
    lhs = float((Tf * g).sum())
    rhs = float((f * Ttg).sum())
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    again = tricubic_transpose(g, Coords(co.data.clone()))
    assert torch.equal(Ttg, again)
    # dense check on a tiny grid: column q of T^T is the response to e_q
    small = (4, 3, 5)
    p2 = np.stack([rng.uniform(-1, s, size=small) for s in small])
    c2 = Coords(offsets_int(p2))
    n = int(np.prod(small))
    Tm = np.zeros((n, n))
    for q in range(n):
        e = np.zeros(n)
        e[q] = 1.0
        Tm[:, q] = O.tricubic(e.reshape(small), p2).ravel()
    gg = rng.standard_normal(small)
    ref = (Tm.T @ gg.ravel()).reshape(small)
    assert rel_l2(from_int(tricubic_transpose(to_int(gg), c2)), ref) < 1e-13


def test_tricubic_adjoint_methods_agree():
    """The sorted-row transpose and the segment transpose are the same
    operator (summation order aside), in FP64 and FP32."""
    from suetes_amd.operators import Coords, tricubic_transpose
    rng = np.random.default_rng(12)
    shape = (40, 33, 21)
    pts = np.stack([rng.uniform(-1.5, s + 0.5, size=shape) for s in shape])
    for dt, tol in ((torch.float64, 1e-14), (torch.float32, 1e-6)):
        co = Coords(offsets_int(pts).to(dt))
        g = to_int(rng.standard_normal(shape), dt)
        a = tricubic_transpose(g, co, method="rows")
        b = tricubic_transpose(g, co, method="segments")
        assert rel_l2(from_int(a), from_int(b)) < tol, dt


def test_tricubic_adjoint_rows_clustered_sources():
    """Departure points crowded into a few cells (a tile's staged sources
    overflow LDS and the row kernel reads them from global memory): the row
    transpose still equals the segment transpose and passes the dot test."""
    from suetes_amd.operators import Coords, tensor_product_interp_3d, tricubic_transpose
    rng = np.random.default_rng(14)
    shape = (70, 12, 9)
    pts = np.stack([rng.uniform(20.0, 23.0, size=shape), rng.uniform(4.0, 6.0, size=shape),
                    rng.uniform(3.0, 5.0, size=shape)])
    # a destination next to the cluster sums thousands of terms: the two
    # summation orders differ by ~sqrt(terms) * eps in FP32
    for dt, tol in ((torch.float64, 1e-14), (torch.float32, 2e-5)):
        co = Coords(offsets_int(pts).to(dt))
        f = to_int(rng.standard_normal(shape), dt)
        g = to_int(rng.standard_normal(shape), dt)
        a = tricubic_transpose(g, co, method="rows")
        b = tricubic_transpose(g, co, method="segments")
        assert rel_l2(from_int(a), from_int(b)) < tol, dt
        lhs = float((tensor_product_interp_3d(f, co) * g).double().sum())
        rhs = float((f * a).double().sum())
        assert abs(lhs - rhs) <= (1e-12 if dt == torch.float64 else 1e-4) * max(abs(lhs), 1.0), dt
        assert torch.equal(a, tricubic_transpose(g, Coords(co.data.clone()), method="rows"))


@pytest.mark.parametrize("cfgf", [small_lam, small_bubble])
def test_departure_points_match_oracle(cfgf):
    cfg = cfgf()
    case = cases()(cfg, device=DEV)
    go, opo, sto, st_o, bco, ext = oracle_case(cfg)
    rng = np.random.default_rng(5)
    s = {"u": rng.uniform(-20, 20, go.Z_u.shape), "v": rng.uniform(-20, 20, go.Z_v.shape),
         "eta_dot": rng.uniform(-0.05, 0.05, go.Z_w.shape), "w": np.zeros(go.Z_w.shape)}
    adv_p = case.stepper.advector
    for loc in ("u", "v", "w", "m"):
        ref = sto.advector.departure(s, loc)
        got = adv_p.compute_departure_indices({k: to_int(v) for k, v in s.items()}, loc).data
        got = np.stack([from_int(got[a]) for a in range(3)]) + index_grid(ref.shape[1:])
        assert rel_l2(got, ref) < 1e-13, loc


def _solver_pair(cfg):
    case = cases()(cfg, device=DEV)
    go, opo, sto, st_o, bco, ext = oracle_case(cfg)
    bgo = sto.physics.precompute_bg(sto.physics.bg_reference())
    A_o, M_o = sto.solver.operators(bgo)
    sol = case.stepper.implicit_solver
    sol.prepare(case.stepper.bg_precomputed())
    return case, go, A_o, M_o, sol


def _rand_krylov(go, rng):
    return {"u": rng.standard_normal(go.Z_u.shape), "v": rng.standard_normal(go.Z_v.shape),
            "w": rng.standard_normal(go.Z_w.shape), "pi": rng.standard_normal(go.Z_m.shape),
            "eta_dot": rng.standard_normal(go.Z_w.shape)}


@pytest.mark.parametrize("cfgf", [small_lam, small_bubble])
def test_operator_and_preconditioner_match_oracle(cfgf):
    case, go, A_o, M_o, sol = _solver_pair(cfgf())
    rng = np.random.default_rng(3)
    x = _rand_krylov(go, rng)
    xp = pack_ref(x)
    Ax = unpack_to_ref(sol.apply_A(xp), x)
    ref = A_o(x)
    for k in O.KEYS:
        assert rel_l2(Ax[k], ref[k]) < 1e-12, ("A", k)
    Mx = unpack_to_ref(sol.apply_M(xp), x)
    ref = M_o(x)
    for k in O.KEYS:
        assert rel_l2(Mx[k], ref[k]) < 1e-12, ("M", k)
    # in-place M (used inside the Arnoldi step)
    y = xp.clone()
    from suetes_amd import _lib
    import ctypes as C
    _lib.check(_lib.lib().sisl_precond_apply_f64(C.byref(sol.op64.struct), y.data_ptr(), y.data_ptr(),
                                                 _lib.stream_ptr()), "M")
    assert torch.equal(y, sol.apply_M(xp))
    # FP32 kernels
    sol32 = case.stepper.implicit_solver
    from suetes_amd.steppers import _OperatorSet
    co = sol32.physics.operator_coefficients(case.stepper.bg_precomputed(), sol32.dt, sol32.alpha)
    sol32.op32 = _OperatorSet(sol32.physics.grid, sol32.physics.op.periodic_axes, co, sol32.alpha, torch.float32)
    Ax32 = unpack_to_ref(sol32.apply_A(xp, dtype=torch.float32).double(), x)
    for k in O.KEYS:
        assert rel_l2(Ax32[k], A_o(x)[k]) < 2e-6, ("A32", k)
    Mx32 = unpack_to_ref(sol32.apply_M(xp, dtype=torch.float32).double(), x)
    for k in O.KEYS:
        assert rel_l2(Mx32[k], M_o(x)[k]) < 2e-5, ("M32", k)


@pytest.mark.parametrize("cfgf", [small_lam, small_bubble, small_schaer])
def test_transposes_pass_dot_tests(cfgf):
    case, go, A_o, M_o, sol = _solver_pair(cfgf())
    rng = np.random.default_rng(9)
    x = pack_ref(_rand_krylov(go, rng))
    y = pack_ref(_rand_krylov(go, rng))
    for f, name in ((sol.apply_A, "A"), (sol.apply_M, "M")):
        lhs = float(torch.dot(f(x), y))
        rhs = float(torch.dot(x, f(y, transpose=True)))
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1e-300), name


def test_gmres_solution_matches_oracle():
    cfg = small_lam()
    case, go, A_o, M_o, sol = _solver_pair(cfg)
    rng = np.random.default_rng(21)
    known = _rand_krylov(go, rng)
    b = A_o(known)
    P = O.Packer(b)
    # reference-semantics solve at a moderate tolerance: same cycles, same iterates up to rounding
    for tol, maxiter in ((1e-6, 3), (1e-13, 40)):
        xo = O.gmres(lambda v: P.pack(A_o(P.unpack(v))), P.pack(b), tol=tol, restart=10, maxiter=maxiter,
                     M=lambda v: P.pack(M_o(P.unpack(v))))
        sol.solver_tol, sol.solver_restart, sol.solver_maxiter = tol, 10, maxiter
        sol._bg_id = None
        sol.prepare(case.stepper.bg_precomputed())
        xg = sol.solve_flat(pack_ref(b), torch.zeros_like(pack_ref(b)))
        got = unpack_to_ref(xg, b)
        ref = P.unpack(xo)
        for k in O.KEYS:
            assert rel_l2(got[k], ref[k]) < (1e-11 if tol < 1e-10 else 1e-9), (tol, k)
    # transposed solve: A^T y = c
    c = pack_ref(_rand_krylov(go, rng))
    yT = sol.solve_flat(c, torch.zeros_like(c), transpose=True)
    r = sol.apply_A(yT, transpose=True) - c
    assert float(r.norm() / c.norm()) < 1e-9


@pytest.mark.parametrize("operand", ["basis", "vector"])
@pytest.mark.parametrize("gs", ["single", "iterated"])
@pytest.mark.parametrize("basis", ["fp16", "fp32"])
@pytest.mark.parametrize("coefs", ["bf16", "fp32"])
def test_mixed_solver_reaches_tolerance(basis, coefs, gs, operand):
    if operand == "vector" and not (basis == "fp16" and gs == "single"):
        pytest.skip("the operand choice exists for an FP16 basis with one Gram-Schmidt pass only")
    cfg = small_lam()
    case, go, A_o, M_o, sol = _solver_pair(cfg)
    rng = np.random.default_rng(22)
    b = pack_ref(A_o(_rand_krylov(go, rng)))
    from suetes_amd.steppers import SemiImplicitSolver3D
    ms = SemiImplicitSolver3D(sol.physics, sol.dt, sol.alpha, solver_tol=1e-10, solver_maxiter=30,
                              solver_restart=20, precision="mixed", inner_basis=basis,
                              inner_coefficients=coefs, inner_gram_schmidt=gs, inner_operand=operand)
    ms.prepare(case.stepper.bg_precomputed())
    for transpose in (False, True):
        x = ms.solve_flat(b, torch.zeros_like(b), transpose=transpose)
        st = ms.last_stats[-1]
        assert st["residual"] <= st["ptol"], (transpose, st)
        x64 = sol.solve_flat(b, torch.zeros_like(b), transpose=transpose)
        assert float((x - x64).norm() / x64.norm()) < 1e-8, transpose


@pytest.mark.parametrize("cfgf", [small_lam, small_bubble])
@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_operator_launch_shapes_bit_identical_and_autotune(dtype, transpose, cfgf):
    """Every launch shape of the operator apply (and of its transpose) gives
    the same bits, on a limited-area grid and on a doubly periodic one; the
    autotuner returns a valid shape and leaves results unchanged."""
    import ctypes as C
    from suetes_amd import _lib
    case, go, A_o, M_o, sol = _solver_pair(cfgf())
    ops = sol.op64
    if dtype == torch.float32:
        from suetes_amd.steppers import _OperatorSet
        co = sol.physics.operator_coefficients(case.stepper.bg_precomputed(), sol.dt, sol.alpha)
        ops = _OperatorSet(sol.physics.grid, sol.physics.op.periodic_axes, co, sol.alpha, dtype)
    L = _lib.lib()
    g = ops.struct.g
    n = int(L.sisl_krylov_size(C.byref(g)))
    x = torch.randn(n, dtype=dtype, device=DEV)
    fp32 = int(dtype == torch.float32)

    def apply():
        y = torch.empty_like(x)
        if transpose:
            fn = L.sisl_helmholtz_apply_T_f32 if fp32 else L.sisl_helmholtz_apply_T_f64
            rc = fn(C.byref(ops.struct), x.data_ptr(), y.data_ptr(), None, _lib.stream_ptr())
        else:
            fn = L.sisl_helmholtz_apply_f32 if fp32 else L.sisl_helmholtz_apply_f64
            rc = fn(C.byref(ops.struct), x.data_ptr(), y.data_ptr(), _lib.stream_ptr())
        _lib.check(rc, "A")
        return y

    outs, shapes = [], []
    for chunk in (2, 3, 5, 13, 81):
        for unr in (1, 2, 3):  # 3: two columns per thread
            _lib.check(L.sisl_set_operator_shape(C.byref(g), fp32, chunk, unr), "shape")
            outs.append(apply())
            shapes.append((chunk, unr))
    for y, sh in zip(outs[1:], shapes[1:]):
        assert torch.equal(y, outs[0]), (sh, (y - outs[0]).abs().max().item())
    kc, un = C.c_int32(), C.c_int32()
    out = torch.empty_like(x)
    nb = int(L.sisl_autotune_scratch_bytes(C.byref(g), fp32))
    fused = 256 + 2 * x.numel() if fp32 else 0
    assert nb == ((fused + 255) // 256) * 256 + (512 << 20)
    tune_scr = torch.empty(max(nb, 1), dtype=torch.uint8, device=x.device)
    _lib.check(L.sisl_autotune_operator(C.byref(ops.struct), fp32, x.data_ptr(), out.data_ptr(),
                                        tune_scr.data_ptr() if nb else None, C.byref(kc), C.byref(un),
                                        _lib.stream_ptr()), "autotune")
    assert kc.value >= 2 and un.value in (1, 2, 3)
    assert torch.equal(apply(), outs[0])


@pytest.mark.parametrize("coefs", ["fp32", "bf16"])
@pytest.mark.parametrize("transpose", [False, True])
def test_fp16_operand_launch_matches_fused_and_plain_launches(transpose, coefs):
    """The operator launch on a stored FP16 basis column (input converted on
    load, outputs scaled by *inv, nothing else stored) gives the bits of the
    fused-normalisation launch on the same values in FP32, at every launch
    shape, and (FP32 coefficients) of the plain launch times the scale."""
    import ctypes as C
    from suetes_amd import _lib
    from suetes_amd.steppers import SemiImplicitSolver3D
    case, go, A_o, M_o, sol = _solver_pair(small_lam())
    ms = SemiImplicitSolver3D(sol.physics, sol.dt, sol.alpha, solver_tol=1e-10, solver_maxiter=30,
                              solver_restart=20, precision="mixed", inner_coefficients=coefs)
    ms.prepare(case.stepper.bg_precomputed())
    ops = ms.op_inner
    cb = int(coefs == "bf16")
    L = _lib.lib()
    g = ops.struct.g
    n = int(L.sisl_krylov_size(C.byref(g)))
    gen = torch.Generator(device=DEV).manual_seed(5)
    x16 = (torch.randn(n, device=DEV, generator=gen) * 3.0).to(torch.float16)
    x32 = x16.float()
    inv = torch.tensor([0.3716], dtype=torch.float64, device=DEV)
    basis = torch.empty(n, dtype=torch.float16, device=DEV)
    st = _lib.stream_ptr()
    outs = []
    for chunk, unr in ((2, 1), (5, 2), (13, 1), (4, 3), (81, 3)):
        _lib.check(L.sisl_set_operator_shape(C.byref(g), 1, chunk, unr), "shape")
        y16 = torch.empty(n, dtype=torch.float32, device=DEV)
        yf = torch.empty_like(y16)
        _lib.check(L.sisl_helmholtz_apply_h16_f32(C.byref(ops.struct), x16.data_ptr(), y16.data_ptr(),
                                                  inv.data_ptr(), cb, int(transpose), st), "h16")
        _lib.check(L.sisl_helmholtz_apply_fused_f32(C.byref(ops.struct), x32.data_ptr(), yf.data_ptr(),
                                                    inv.data_ptr(), basis.data_ptr(), 1.0, cb, int(transpose),
                                                    st), "fused")
        assert torch.equal(y16, yf), ((chunk, unr), (y16 - yf).abs().max().item())
        if not cb:
            yp = torch.empty_like(y16)
            fn = L.sisl_helmholtz_apply_T_f32 if transpose else L.sisl_helmholtz_apply_f32
            args = (C.byref(ops.struct), x32.data_ptr(), yp.data_ptr()) + ((None,) if transpose else ())
            _lib.check(fn(*args, st), "plain")
            assert torch.equal(y16, yp * torch.tensor(0.3716, dtype=torch.float32, device=DEV)), (chunk, unr)
        outs.append(y16)
    for y in outs[1:]:
        assert torch.equal(y, outs[0])
    assert bool(torch.isfinite(outs[0]).all()) and float(outs[0].abs().max()) > 0


@pytest.mark.parametrize("transpose", [False, True])
def test_mixed_solve_independent_of_launch_shapes(transpose):
    """The mixed solver (whose inner cycles use the operator kernels with the
    fused basis normalisation and on the stored FP16 basis column, in FP32
    and bfloat16 coefficient storage) gives the same bits whatever launch
    shapes are pinned, with the inner cycle replayed from a captured graph or
    launched eagerly."""
    import ctypes as C
    from suetes_amd import _lib
    from suetes_amd.steppers import SemiImplicitSolver3D
    case, go, A_o, M_o, sol = _solver_pair(small_lam())
    rng = np.random.default_rng(29)
    b = pack_ref(A_o(_rand_krylov(go, rng)))
    ms = SemiImplicitSolver3D(sol.physics, sol.dt, sol.alpha, solver_tol=1e-10, solver_maxiter=30,
                              solver_restart=20, precision="mixed")
    ms.prepare(case.stepper.bg_precomputed())
    L = _lib.lib()
    g = ms.op64.struct.g
    outs = []
    try:
        for graphs in (0, 1):
            L.sisl_set_graphs(graphs)
            for chunk, unr in ((2, 1), (5, 2), (13, 1), (4, 2), (4, 3), (13, 3)):
                for fp32 in (0, 1):
                    _lib.check(L.sisl_set_operator_shape(C.byref(g), fp32, chunk, unr), "shape")
                x = ms.solve_flat(b, torch.zeros_like(b), transpose=transpose)
                torch.cuda.synchronize()
                outs.append(((graphs, chunk, unr), x))
    finally:
        L.sisl_set_graphs(1)
    ref = outs[0][1]
    for tag, x in outs[1:]:
        assert torch.equal(x, ref), (tag, (x - ref).abs().max().item())


def test_bfloat16_inner_operator_autotune_keeps_the_bits():
    """The inner operator's bfloat16 launches are tuned on their own (fp32 =
    2: FP32 arithmetic, bfloat16 3-D coefficients); re-tuning returns a valid
    shape and leaves the mixed solve's bits unchanged."""
    import ctypes as C
    from suetes_amd import _lib
    from suetes_amd.steppers import SemiImplicitSolver3D
    case, go, A_o, M_o, sol = _solver_pair(small_lam())
    rng = np.random.default_rng(31)
    b = pack_ref(A_o(_rand_krylov(go, rng)))
    ms = SemiImplicitSolver3D(sol.physics, sol.dt, sol.alpha, solver_tol=1e-10, solver_maxiter=30,
                              solver_restart=20, precision="mixed")
    ms.prepare(case.stepper.bg_precomputed())
    assert ms.op_inner.coef_bf16
    x1 = ms.solve_flat(b, torch.zeros_like(b))
    L = _lib.lib()
    g = ms.op_inner.struct.g
    n = int(L.sisl_krylov_size(C.byref(g)))
    x = torch.randn(n, dtype=torch.float32, device=DEV)
    y = torch.empty_like(x)
    nb = int(L.sisl_autotune_scratch_bytes(C.byref(g), 1))
    scr = torch.empty(nb, dtype=torch.uint8, device=DEV)
    kc, un = C.c_int32(), C.c_int32()
    _lib.check(L.sisl_autotune_operator(C.byref(ms.op_inner.struct), 2, x.data_ptr(), y.data_ptr(),
                                        scr.data_ptr(), C.byref(kc), C.byref(un), _lib.stream_ptr()),
               "bf16 autotune")
    assert kc.value >= 2 and un.value in (1, 2, 3)
    x2 = ms.solve_flat(b, torch.zeros_like(b))
    torch.cuda.synchronize()
    assert torch.equal(x1, x2)


def test_inner_cycle_graph_replays_match_eager_launches():
    """The mixed solver's inner Arnoldi cycle captured once as a hipGraph and
    replayed: every replay (with eager GPU work and synchronisations in
    between) gives bit-identical solutions to the eager launches, forward
    and transposed."""
    from suetes_amd import _lib
    from suetes_amd.steppers import SemiImplicitSolver3D
    cfg = small_lam()
    case, go, A_o, M_o, sol = _solver_pair(cfg)
    rng = np.random.default_rng(23)
    b = pack_ref(A_o(_rand_krylov(go, rng)))
    ms = SemiImplicitSolver3D(sol.physics, sol.dt, sol.alpha, solver_tol=1e-10, solver_maxiter=30,
                              solver_restart=20, precision="mixed")
    ms.prepare(case.stepper.bg_precomputed())
    L = _lib.lib()
    try:
        for transpose in (False, True):
            L.sisl_set_graphs(0)
            ref = ms.solve_flat(b, torch.zeros_like(b), transpose=transpose)
            cyc = ms.last_stats[-1]["cycles"]
            assert cyc >= 3  # several cycles: the graph is replayed within one solve too
            L.sisl_set_graphs(1)
            for rep in range(3):
                x = ms.solve_flat(b, torch.zeros_like(b), transpose=transpose)
                torch.cuda.synchronize()
                assert torch.equal(x, ref), (transpose, rep)
                junk = torch.randn_like(b) * 2.0 + 1.0  # eager work between replays
                torch.cuda.synchronize()
                del junk
    finally:
        L.sisl_set_graphs(1)
