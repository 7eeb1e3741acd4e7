"""Full-step pins for the oracle (steppers.py:720-1073).

Follows the reference's own step-level tests: resting balance
(test_dynamical_core_3d.py:485-539), Coriolis/metric regressions and closed-box
budgets (test_core_formula_regressions.py:168-292).
"""
import numpy as np

from oracle import sisl as O

C = {'g': 9.81, 'cp': 1004.0, 'Rd': 287.0, 'cvd': 717.0, 'p0': 100000.0}


def ident(s, f):
    return s


def rest_state(ph):
    g = ph.grid
    st = {'u': np.zeros_like(g.Z_u), 'v': np.zeros_like(g.Z_v), 'w': np.zeros_like(g.Z_w),
          'eta_dot': np.zeros_like(g.Z_w), 'pi': ph.pi_bg.copy(), 'th_v': ph.theta_bg.copy(),
          'rho': ph.bg_reference()['rho']}
    st['u_prev'], st['v_prev'], st['w_prev'] = st['u'], st['v'], st['w']
    st['eta_dot_prev'] = st['eta_dot']
    st['tend_th_v_prev'] = np.zeros_like(st['th_v'])
    st['is_first_step'] = 1.0
    return st


def test_resting_mountain_stays_at_rest():
    def h(x, y):
        return 1500.0 / (1.0 + (x ** 2 + y ** 2) / 30000.0 ** 2)
    g = O.Grid(16, 14, 12, 6000.0, 6000.0, 500.0, 45.0, 5.0, h_func=h)
    op = O.Ops(g)
    ph = O.Euler(g, op, C, dt=100.0, N_bv=0.01)
    st = O.Stepper(ph, 100.0)
    s = rest_state(ph)
    for i in range(8):
        s = st.step(s, i * 100.0, None, ident)
    assert np.max(np.abs(s['w'])) < 1e-7


def _env(dt, latitude=45.0, tracer=False):
    g = O.Grid(8, 8, 5, 2000.0, 2000.0, 500.0, latitude, 0.0, h_func=lambda x, y: np.zeros_like(x + y))
    op = O.Ops(g)

    class Tr:
        tracer_keys = ['q_test']

        def get_explicit_tendencies(self, *a, **k):
            return {}

        def apply_state_updates(self, s):
            return s
    ph = O.Euler(g, op, C, dt=dt, N_bv=0.0, damp_height=1e9, max_damp=0.0, nu_div_factor=0.0,
                 nu_h_factor=0.0, physics_suite=Tr() if tracer else None)
    return g, op, ph


def _base(ph, u=0.0, v=0.0):
    g = ph.grid
    return {'u': np.full((g.nx + 1, g.ny, g.nz), u), 'v': np.full((g.nx, g.ny + 1, g.nz), v),
            'w': np.zeros((g.nx, g.ny, g.nz + 1)), 'eta_dot': np.zeros((g.nx, g.ny, g.nz + 1)),
            'pi': ph.pi_bg.copy(), 'th_v': ph.theta_bg.copy(),
            'rho': C['p0'] / (C['Rd'] * ph.theta_bg) * ph.pi_bg ** (C['cvd'] / C['Rd'])}


def test_nonlinear_pgf_does_not_rescale_coriolis():
    # test_core_formula_regressions.py:168-194
    dt = 0.05
    g, op, ph = _env(dt, latitude=55.0)

    def adv(off):
        s = _base(ph, 8.0, 12.0)
        s['th_v'] = s['th_v'] + off
        s['rho'] = C['p0'] / (C['Rd'] * s['th_v']) * s['pi'] ** (C['cvd'] / C['Rd'])
        return O.Stepper(ph, dt, alpha=0.5, solver_tol=1e-11, solver_maxiter=80,
                         solver_restart=40).step(s, 0.0, None, ident)
    r, w = adv(0.0), adv(30.0)
    err = max(np.max(np.abs(w['u'][1:-1] - r['u'][1:-1])), np.max(np.abs(w['v'][:, 1:-1] - r['v'][:, 1:-1])))
    inc = np.max(np.abs(r['u'] - 8.0))
    assert inc > 1e-8 and err < 1e-5 * inc


def test_full_metric_curvature_tendency():
    # test_core_formula_regressions.py:197-235
    dt = 1e-3
    g, op, ph = _env(dt, latitude=0.0)
    for loc in ('m', 'u', 'v', 'w'):
        g.m_factors[loc] = np.ones_like(g.m_factors[loc])
    op.m3 = {k: v[..., None] for k, v in g.m_factors.items()}
    g.f_u = np.zeros_like(g.f_u)
    g.f_v = np.zeros_like(g.f_v)
    g.dm_dx_m = np.full_like(g.dm_dx_m, 2e-4)
    g.dm_dy_m = np.full_like(g.dm_dy_m, -1e-4)
    s = _base(ph, 10.0, 6.0)
    thp = s['th_v'] - ph.theta_bg
    bg = ph.precompute_bg(ph.bg_reference())
    prime = {'u': s['u'], 'v': s['v'], 'w': s['w'], 'eta_dot': s['eta_dot'], 'pi': s['pi'] - ph.pi_bg,
             'th_v': s['th_v'], 'th_v_prime_u': op.avg(thp, 0, 'm', 'u'),
             'th_v_prime_v': op.avg(thp, 1, 'm', 'v'), 'th_v_prime_w': op.avg(thp, 2, 'm', 'w')}
    ex = ph.get_tendencies(prime, bg, is_explicit=True)
    r = O.Stepper(ph, dt, alpha=0.5, solver_tol=1e-12, solver_maxiter=80, solver_restart=40).step(
        s, 0.0, None, ident)
    eu = np.max(np.abs((r['u'] - s['u'])[1:-1] / dt - ex['u'][1:-1]))
    ev = np.max(np.abs((r['v'] - s['v'])[:, 1:-1] / dt - ex['v'][:, 1:-1]))
    sig = max(np.max(np.abs(ex['u'])), np.max(np.abs(ex['v'])))
    assert sig > 1e-4 and max(eu, ev) < 2e-3 * sig


def test_closed_box_density_and_tracer_budget():
    # test_core_formula_regressions.py:238-292
    dt = 0.02
    g, op, ph = _env(dt, tracer=True)
    s = _base(ph)
    iu = np.arange(g.nx + 1.0)[:, None, None]
    jv = np.arange(g.ny + 1.0)[None, :, None]
    s['u'] = 2.0 * np.sin(np.pi * iu / g.nx) + np.zeros_like(s['u'])
    s['v'] = -1.5 * np.sin(np.pi * jv / g.ny) + np.zeros_like(s['v'])
    x = np.arange(g.nx * 1.0)[:, None, None]
    y = np.arange(g.ny * 1.0)[None, :, None]
    s['q_test'] = 1.0 + 0.2 * np.sin(2 * np.pi * x / g.nx) * np.cos(2 * np.pi * y / g.ny) + np.zeros_like(s['rho'])
    vol = g.dx * g.dy * g.dz_m_full / g.m_factors['m'][..., None] ** 2
    m0, t0 = np.sum(s['rho'] * vol), np.sum(s['rho'] * s['q_test'] * vol)
    st = O.Stepper(ph, dt, alpha=0.5, solver_tol=1e-11, solver_maxiter=80, solver_restart=40)
    bg = ph.precompute_bg(ph.bg_reference())
    tm = st.ffsl.advect_3d_split(s['rho'] * s['q_test'], s, bg)
    r = st.step(s, 0.0, None, ident)
    m1, t1 = np.sum(r['rho'] * vol), np.sum(r['rho'] * r['q_test'] * vol)
    assert abs(m1 - m0) / m0 < 1e-10
    assert abs(t1 - t0) / t0 < 1e-10
    assert np.linalg.norm(r['rho'] * r['q_test'] - tm) / np.linalg.norm(tm) < 1e-12
