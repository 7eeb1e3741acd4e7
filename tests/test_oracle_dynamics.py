"""Pins for the oracle's operators, tendencies, solver and transport.

Sources of each assertion are the reference's own tests (file:line cited)
or closed forms / SciPy routines independent of the oracle's code.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as ssl

from oracle import sisl as O

C = {'g': 9.81, 'cp': 1004.0, 'Rd': 287.0, 'cvd': 717.0, 'p0': 100000.0}


def mountain_env(nx=16, ny=14, nz=10, dt=100.0, periodic=()):
    def h(x, y):
        return 1500.0 * np.exp(-(x ** 2 + y ** 2) / 5000.0 ** 2)
    g = O.Grid(nx, ny, nz, 1000.0, 1000.0, 500.0, 45.0, 0.0, h_func=h)
    op = O.Ops(g, periodic)
    ph = O.Euler(g, op, C, dt=dt, N_bv=0.01)
    return g, op, ph


def rand_state(g, rng, scale=1.0):
    return {'u': rng.standard_normal((g.nx + 1, g.ny, g.nz)) * scale,
            'v': rng.standard_normal((g.nx, g.ny + 1, g.nz)) * scale,
            'w': rng.standard_normal((g.nx, g.ny, g.nz + 1)) * scale,
            'pi': rng.standard_normal((g.nx, g.ny, g.nz)) * scale * 1e-3,
            'eta_dot': rng.standard_normal((g.nx, g.ny, g.nz + 1)) * scale * 1e-3}


# --- operators: tests/test_periodic_boundaries.py:19-48 --------------------
def test_periodic_and_open_closures():
    g = O.Grid(8, 4, 2, 1.0, 1.0, 1.0, 0.0, 0.0)
    f = np.arange(8, dtype=np.float64)[:, None, None] * np.ones((8, 4, 2))
    op = O.Ops(g, (0,))
    d = op.diff(f, 0, 'm', 'u')[:, 0, 0]
    a = op.avg(f, 0, 'm', 'u')[:, 0, 0]
    m = g.m_factors['u'][:, 0]
    np.testing.assert_allclose(d[[0, -1]] / m[[0, -1]], -7.0)
    np.testing.assert_allclose(a[[0, -1]], 3.5)
    np.testing.assert_allclose(d[1:-1] / m[1:-1], 1.0)
    op2 = O.Ops(g)
    d2 = op2.diff(f, 0, 'm', 'u')[:, 0, 0]
    np.testing.assert_allclose(d2[[0, -1]] / m[[0, -1]], 1.0)


# --- tendencies at rest: test_dynamical_core_3d.py:201-228, 408-483 -----------
def test_resting_tendencies_vanish_over_terrain():
    g, op, ph = mountain_env()
    bg = ph.precompute_bg(ph.bg_reference())
    zero = {'u': np.zeros_like(g.Z_u), 'v': np.zeros_like(g.Z_v), 'w': np.zeros_like(g.Z_w),
            'pi': np.zeros_like(g.Z_m), 'eta_dot': np.zeros_like(g.Z_w)}
    t = ph.get_tendencies(zero, bg)
    assert np.max(np.abs(t['u'])) < 1e-12 and np.max(np.abs(t['w'])) < 1e-12


def test_schaer_metric_balance_and_divergence():
    # test_dynamical_core_3d.py:386-483 (SleveSimple, Schaer terrain)
    def schaer(x, y):
        return 2000.0 * np.exp(-(x ** 2) / 15000.0 ** 2) * np.cos(np.pi * x / 8000.0) ** 2
    g = O.Grid(150, 5, 50, 1000.0, 1000.0, 400.0, 0.0, 0.0, h_func=schaer,
               transform=O.SleveSimple(4000.0, 15000.0, 1.35))
    op = O.Ops(g)
    ph = O.Euler(g, op, C, dt=100.0, damp_height=g.Lz, N_bv=0.01)
    bg = ph.precompute_bg(ph.bg_reference())
    z = {'u': np.zeros_like(g.Z_u), 'v': np.zeros_like(g.Z_v), 'w': np.zeros_like(g.Z_w),
         'pi': np.zeros_like(g.Z_m), 'eta_dot': np.zeros_like(g.Z_w)}
    assert np.max(np.abs(ph.get_tendencies(z, bg)['u'])) <= 1e-6
    u = 10.0 * np.ones_like(g.Z_u)
    v = np.zeros_like(g.Z_v)
    mu, mv = g.m_factors['u'][..., None], g.m_factors['v'][..., None]
    fx = u * bg['rho_u'] * bg['th_v_u'] * bg['dz_u'] / mu
    fy = v * bg['rho_v'] * bg['th_v_v'] * bg['dz_v'] / mv
    dh = op.diff(fx, 0, 'u', 'm') / bg['dz_m_full'] + op.diff(fy, 1, 'v', 'm') / bg['dz_m_full']
    fz = np.concatenate([np.zeros((g.nx, g.ny, 1)), np.cumsum(-dh * bg['dz_m_full'], axis=2)], axis=2)
    eta = fz / (bg['dz_w_full'] * bg['rho_w'] * bg['th_v_w'])
    t = ph.get_tendencies({'u': u, 'v': v, 'w': np.zeros_like(g.Z_w), 'pi': np.zeros_like(g.Z_m),
                           'eta_dot': eta}, bg)
    assert np.max(np.abs(t['pi'][:, :, :-1])) <= 1e-5
    assert np.min(bg['dz_w_full']) >= 10.0


def test_linear_operator_is_linear_and_matches_tendency_definition():
    g, op, ph = mountain_env()
    bg = ph.precompute_bg(ph.bg_reference())
    rng = np.random.default_rng(0)
    x, y = rand_state(g, rng), rand_state(g, rng)
    a, b = 0.7, -1.3
    xy = {k: a * x[k] + b * y[k] for k in x}
    Lx = ph.linear_operator(x, bg, 100.0)
    Ly = ph.linear_operator(y, bg, 100.0)
    Lxy = ph.linear_operator(xy, bg, 100.0)
    for k in Lx:
        np.testing.assert_allclose(Lxy[k], a * Lx[k] + b * Ly[k], rtol=1e-10, atol=1e-12)
    # interior of L_pi is pi - alpha dt T_pi (euler.py:401)
    t = ph.get_tendencies(x, bg)
    np.testing.assert_allclose(Lx['pi'], x['pi'] - 0.55 * 100.0 * t['pi'], rtol=1e-13)


# --- preconditioner ---------------------------------------------------------
def test_thomas_matches_scipy_solve_banded():
    rng = np.random.default_rng(5)
    n = 23
    lo = -rng.uniform(0, 1, (4, n))
    up = -rng.uniform(0, 1, (4, n))
    ma = 1.0 - lo - up
    r = rng.standard_normal((4, n))
    x = O.thomas(lo, ma, up, r)
    for c in range(4):
        ab = np.zeros((3, n))
        ab[0, 1:] = up[c, :-1]
        ab[1] = ma[c]
        ab[2, :-1] = lo[c, 1:]
        np.testing.assert_allclose(x[c], sla.solve_banded((1, 1), ab, r[c]), rtol=1e-12)


def test_preconditioner_inverts_the_column_operator():
    """On a flat f-free grid, horizontally uniform states with u=v=0 only see
    the vertical couplings, which the column solve inverts exactly."""
    g = O.Grid(4, 3, 12, 1000.0, 1000.0, 400.0, 0.0, 0.0)
    op = O.Ops(g)
    ph = O.Euler(g, op, C, dt=30.0, N_bv=0.012, damp_height=3000.0, max_damp=0.2)
    g.f_u[:] = 0.0
    g.f_v[:] = 0.0
    bg = ph.precompute_bg(ph.bg_reference())
    sol = O.ImplicitSolver(ph, 30.0)
    A, M = sol.operators(bg)
    M.precompute_banded(bg)
    rng = np.random.default_rng(6)
    colw = rng.standard_normal(13)
    colp = rng.standard_normal(12)
    cole = rng.standard_normal(13)
    x = {'u': np.zeros((5, 3, 12)), 'v': np.zeros((4, 4, 12)),
         'w': np.broadcast_to(colw, (4, 3, 13)).copy(),
         'pi': np.broadcast_to(colp, (4, 3, 12)).copy(),
         'eta_dot': np.broadcast_to(cole, (4, 3, 13)).copy()}
    y = M(A(x))
    for k in x:
        np.testing.assert_allclose(y[k], x[k], rtol=1e-10, atol=1e-12)


# --- GMRES: test_dynamical_core_3d.py:230-280, plus SciPy ------------------
@pytest.mark.parametrize('mode', ['jax', 'cgs2'])
def test_gmres_manufactured_solution(mode):
    g, op, ph = mountain_env(nx=12, ny=10, nz=8)
    bg = ph.precompute_bg(ph.bg_reference())
    rng = np.random.default_rng(42)
    known = {'u': rng.uniform(size=(13, 10, 8)) * 0.1, 'v': rng.uniform(size=(12, 11, 8)) * 0.1,
             'w': rng.uniform(size=(12, 10, 9)) * 0.01, 'pi': rng.uniform(size=(12, 10, 8)) * 1e-3,
             'eta_dot': rng.uniform(size=(12, 10, 9)) * 1e-3}
    for k in ('w', 'eta_dot'):
        known[k][:, :, 0] = 0.0
        known[k][:, :, -1] = 0.0
    rhs = ph.linear_operator(known, bg, dt=100.0, alpha=0.55)
    s = O.ImplicitSolver(ph, 100.0, solver_tol=1e-8, solver_maxiter=100, solver_restart=100, gs_mode=mode)
    sol = s.solve(rhs, bg)
    assert s.relative_residual(sol, rhs, bg) < 1e-6
    tight = O.ImplicitSolver(ph, 100.0, solver_tol=1e-13, solver_maxiter=100, solver_restart=100,
                             gs_mode=mode)
    sol = tight.solve(rhs, bg)
    for k in known:
        np.testing.assert_allclose(sol[k], known[k], atol=1e-9 * np.max(np.abs(known[k])) + 1e-14)


def test_gmres_agrees_with_scipy():
    rng = np.random.default_rng(7)
    n = 60
    A = np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    Mi = np.diag(1.0 / np.diag(A))
    b = rng.standard_normal(n)
    x = O.gmres(lambda v: A @ v, b, tol=1e-13, restart=15, maxiter=50, M=lambda v: Mi @ v)
    xs, info = ssl.gmres(A, b, rtol=1e-14, restart=15, maxiter=200)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(x, xs, rtol=1e-8, atol=1e-9)


# --- departure points: test_dynamical_core_3d.py:100-123 --------------------
def test_departure_uniform_wind():
    g, op, ph = mountain_env(nx=32, ny=32, nz=15)
    adv = O.SLAdvector(g, ph, 100.0)
    st = {'u': np.ones((33, 32, 15)) * 10.0, 'v': np.zeros((32, 33, 15)), 'eta_dot': np.zeros((32, 32, 16))}
    d = adv.departure(st, 'm')
    X = np.arange(32)[:, None, None]
    err = np.max(np.abs((X - d[0]) - (10.0 * 100.0 / 1000.0) * g.m_factors['m'][..., None]))
    assert err < 1e-4


def test_departure_linear_strain_third_order():
    """u = -a x (index space): exact x_d = x_a e^{a dt}; midpoint iteration
    error is O(dt^3) (sisl_component_convergence.py:147-197 pattern)."""
    g = O.Grid(40, 3, 3, 1.0, 1.0, 1.0, 0.0, 0.0)
    op = O.Ops(g)
    errs = []
    for dt in (0.4, 0.2, 0.1):
        adv = O.SLAdvector(g, type('P', (), {'op': op})(), dt)
        a = 0.5
        xf = np.arange(41.0) - 20.0
        u = (-a * xf)[:, None, None] * np.ones((41, 3, 3)) / g.m_factors['u'][..., None]
        st = {'u': u, 'v': np.zeros((40, 4, 3)), 'eta_dot': np.zeros((40, 3, 4))}
        d = adv.departure(st, 'u')
        exact = xf * np.exp(a * dt) + 20.0
        errs.append(np.max(np.abs(d[0][15:26, 1, 1] - exact[15:26])))
    r1, r2 = np.log2(errs[0] / errs[1]), np.log2(errs[1] / errs[2])
    assert r1 > 2.7 and r2 > 2.7


# --- FFSL: test_dynamical_core_3d.py:125-179, test_periodic_boundaries.py:51-82
def test_ffsl_closed_box_mass():
    g, op, ph = mountain_env(nx=32, ny=32, nz=15)
    bg = ph.precompute_bg(ph.bg_reference())
    rho = ph.bg_reference()['rho']
    st = {'u': np.ones((33, 32, 15)) * 15.0, 'v': np.ones((32, 33, 15)) * 5.0,
          'eta_dot': np.zeros((32, 32, 16))}
    st['u'][0] = st['u'][-1] = 0.0
    st['v'][:, 0] = st['v'][:, -1] = 0.0
    vol = (g.dx * g.dy / g.m_factors['m'][..., None] ** 2) * g.dz_m_full
    out = O.FFSL(g, 100.0).advect_3d_split(rho, st, bg)
    m0, m1 = np.sum(rho * vol), np.sum(out * vol)
    assert abs(m1 - m0) / m0 < 1e-14


def test_ffsl_periodic_roll_and_fractional():
    s = np.array([1.0, 2.0, 4.0, 8.0, 16.0, 32.0, 64.0, 128.0])
    r = O.FFSL.advect_lines(s, np.ones(9), True)
    l = O.FFSL.advect_lines(s, -np.ones(9), True)
    np.testing.assert_allclose(r, np.roll(s, 1))
    np.testing.assert_allclose(l, np.roll(s, -1))
    s2 = np.array([0.0, 0.0, 1.0, 2.0, 1.0, 0.0, 0.0, 0.0])
    t = O.FFSL.advect_lines(s2, np.full(9, 0.35), True)
    np.testing.assert_allclose(t.sum(), s2.sum(), rtol=1e-12)
    assert t.min() >= 0.0


# --- Davies sponge: test_dynamical_core_3d.py:282-332 -----------------------
def test_davies_sponge_inflow_outflow():
    g, op, ph = mountain_env(nx=32, ny=32, nz=15)
    sp = O.DaviesSponge(g, op, sponge_depth=5, dt=30.0, outflow_factor=0.01)
    nx, ny, nz = 32, 32, 15
    st = {'u': np.ones((33, 32, 15)) * 10.0, 'v': np.zeros((32, 33, 15)), 'w': np.zeros((32, 32, 16)),
          'th_v': np.full((32, 32, 15), 300.0), 'pi': np.ones((32, 32, 15)), 'q': np.zeros((32, 32, 15))}
    ext = {k: (np.zeros_like(v) if k in ('u', 'v', 'w', 'q') else v.copy()) for k, v in st.items()}
    b = sp.blend(st, ext)
    assert np.isclose(b['u'][nx // 2, ny // 2, nz // 2], 10.0)
    assert np.isclose(b['u'][0, ny // 2, nz // 2], 0.0)
    assert np.isclose(b['u'][-1, ny // 2, nz // 2], 0.99 * 10.0)
