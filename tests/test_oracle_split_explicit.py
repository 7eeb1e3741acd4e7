"""Pins of the oracle's split-explicit core (oracle/split_explicit.py,
reference steppers.py:1076-1737): the reference's own regressions
(tests/test_core_formula_regressions.py:129-165, 238-292), closed forms
for the third-order upwind flux, the acoustic column solve against the coupled
off-centred column equations it eliminates, and a resting atmosphere over terrain."""
import numpy as np
import pytest

from oracle import sisl as O
from oracle.split_explicit import SplitExplicitStepper

C = {'g': 9.81, 'cp': 1004.0, 'Rd': 287.0, 'cvd': 717.0, 'p0': 100000.0}


class Tr:
    tracer_keys = ['q_test']

    def get_explicit_tendencies(self, *a, **k):
        return {}

    def apply_state_updates(self, s):
        return s


def _env(dt=0.1, tracer=False, terrain=None, N=0.0, periodic=()):
    g = O.Grid(8, 8, 5, 2000.0, 2000.0, 500.0, 45.0, 0.0,
               h_func=terrain if terrain else (lambda x, y: np.zeros_like(x + y)))
    op = O.Ops(g, periodic)
    ph = O.Euler(g, op, C, dt=dt, N_bv=N, damp_height=1e9, max_damp=0.0, nu_div_factor=0.0,
                 nu_h_factor=0.0, physics_suite=Tr() if tracer else None)
    return g, op, ph


def _base(ph, u=0.0, v=0.0):
    g = ph.grid
    return {'u': np.full((g.nx + 1, g.ny, g.nz), u), 'v': np.full((g.nx, g.ny + 1, g.nz), v),
            'w': np.zeros((g.nx, g.ny, g.nz + 1)), 'eta_dot': np.zeros((g.nx, g.ny, g.nz + 1)),
            'pi': ph.pi_bg.copy(), 'th_v': ph.theta_bg.copy(),
            'rho': C['p0'] / (C['Rd'] * ph.theta_bg) * ph.pi_bg ** (C['cvd'] / C['Rd'])}


def _bg(ph):
    return ph.precompute_bg(ph.bg_reference())


def test_split_continuity_includes_theta_prime_flux():
    # tests/test_core_formula_regressions.py:129-165
    g, op, ph = _env(dt=0.2)
    st = SplitExplicitStepper(ph, dt=0.2, ns=2, alpha=0.5)
    s = _base(ph)
    x = np.arange(g.nx + 1.0)[:, None, None]
    z = np.arange(g.nz * 1.0)[None, None, :]
    s['u'] = 3.0 + 0.15 * x + np.zeros_like(s['u'])
    thp = 4.0 + 0.3 * np.arange(g.nx * 1.0)[:, None, None] + 0.2 * z + np.zeros_like(s['th_v'])
    s['th_v'] = ph.theta_bg + thp
    bg = _bg(ph)
    actual = st._compute_slow_tendencies(s, bg, {}, {}, None)['pi']
    prime = {'u': s['u'], 'v': s['v'], 'w': s['w'], 'eta_dot': s['eta_dot'], 'pi': s['pi'] - ph.pi_bg,
             'th_v': s['th_v'], 'th_v_prime_u': op.avg(thp, 0, 'm', 'u'),
             'th_v_prime_v': op.avg(thp, 1, 'm', 'v'), 'th_v_prime_w': op.avg(thp, 2, 'm', 'w')}
    linear = ph.get_tendencies(prime, bg, is_explicit=True)['pi']
    # independent discretisation of the theta' Exner flux (the reference test's own)
    mu, mv, mm = (g.m_factors[k][..., None] for k in ('u', 'v', 'm'))
    fx = s['u'] * bg['rho_u'] * op.avg(thp, 0, 'm', 'u') * bg['dz_u'] / mu
    fy = s['v'] * bg['rho_v'] * op.avg(thp, 1, 'm', 'v') * bg['dz_v'] / mv
    fz = s['eta_dot'] * bg['dz_w_full'] * bg['rho_w'] * op.avg(thp, 2, 'm', 'w')
    fz[:, :, [0, -1]] = 0.0
    nonlinear = -bg['C_pi'] * (mm * (op.diff(fx, 0, 'u', 'm') / bg['dz_m_full'] +
                                     op.diff(fy, 1, 'v', 'm') / bg['dz_m_full']) +
                               op.diff(fz, 2, 'w', 'm') * g.dz / bg['dz_m_full'])
    err = np.max(np.abs(actual - (linear + nonlinear)))
    sig = np.max(np.abs(nonlinear))
    assert sig > 1e-12 and err < 1e-11 * sig


def test_complete_split_step_conserves_density_and_tracer_mass():
    # tests/test_core_formula_regressions.py:238-292 (split_explicit)
    dt = 0.02
    g, op, ph = _env(dt=dt, tracer=True)
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
    st = SplitExplicitStepper(ph, dt, ns=2, alpha=0.5)
    tm = st.ffsl_advector.advect_3d_split(s['rho'] * s['q_test'], s, _bg(ph))
    r = st.step(s, 0.0, None, lambda a, b: a)
    m1, t1 = np.sum(r['rho'] * vol), np.sum(r['rho'] * r['q_test'] * vol)
    assert abs(m1 - m0) / m0 < 1e-10
    assert abs(t1 - t0) / t0 < 1e-10
    assert np.linalg.norm(r['rho'] * r['q_test'] - tm) / np.linalg.norm(tm) < 1e-12


@pytest.mark.parametrize("periodic", [False, True])
def test_upwind_flux_centred_part_exact_on_cubic_averages_and_bias_is_third_difference(periodic):
    g, op, ph = _env(periodic=(0,) if periodic else ())
    st = SplitExplicitStepper(ph, dt=1.0, ns=1)
    n = g.nx
    i = np.arange(n, dtype=float)[:, None, None] + np.zeros((1, g.ny, g.nz))
    flux = np.ones((n + 1, g.ny, g.nz))
    if not periodic:
        # q holds cell averages of a polynomial P over [i-1/2, i+1/2]; face f
        # lies between cells f-1 and f at x = f - 1/2.  The centred part is the
        # fourth-order finite-volume reconstruction (exact on cubics), the bias
        # is a twelfth of the third difference (a/2 for P = a x^3 + ...).
        a, b, c, d = 0.3, -1.0, 2.0, -5.0
        q = a * (i ** 3 + i / 4) + b * (i ** 2 + 1.0 / 12) + c * i + d
        xf = np.arange(n + 1, dtype=float) - 0.5
        exact = a * xf ** 3 + b * xf ** 2 + c * xf + d
        vp = st._upwind_flux_3rd_order(flux, q, 0)[2:-2, 0, 0]
        vm = -st._upwind_flux_3rd_order(-flux, q, 0)[2:-2, 0, 0]
        np.testing.assert_allclose(0.5 * (vp + vm), exact[2:-2], rtol=0, atol=1e-12)
        np.testing.assert_allclose(0.5 * (vp - vm), a / 2, rtol=0, atol=1e-12)
        # on quadratics the upwind face value is exact for both flow directions
        q2 = b * (i ** 2 + 1.0 / 12) + c * i + d
        e2 = b * xf ** 2 + c * xf + d
        np.testing.assert_allclose(st._upwind_flux_3rd_order(flux, q2, 0)[2:-2, 0, 0], e2[2:-2], rtol=0, atol=1e-12)
        np.testing.assert_allclose(-st._upwind_flux_3rd_order(-flux, q2, 0)[2:-2, 0, 0], e2[2:-2], rtol=0,
                                   atol=1e-12)
        # a flux below eps_l is gated to the centred value
        small = np.full_like(flux, 1e-13)
        np.testing.assert_allclose(st._upwind_flux_3rd_order(small, q, 0)[2:-2, 0, 0], 1e-13 * exact[2:-2],
                                   rtol=1e-12)
    else:
        # periodic: a single Fourier mode's face value has the closed-form symbol
        k = 2 * np.pi / n
        q = np.cos(k * i)
        f = st._upwind_flux_3rd_order(flux, q, 0)
        xf = np.arange(n + 1, dtype=float) - 0.5
        a4 = (7 / 6) * np.cos(k / 2) - (1 / 6) * np.cos(3 * k / 2)
        bias = (1 / 12) * (-2 * np.sin(3 * k / 2) + 6 * np.sin(k / 2))
        np.testing.assert_allclose(f[:, 0, 0], a4 * np.cos(k * xf) + bias * np.sin(k * xf), rtol=0, atol=1e-12)


def test_acoustic_column_solve_satisfies_the_coupled_implicit_equations():
    def h(x, y):
        return 400.0 * np.exp(-(x ** 2 + y ** 2) / 4000.0 ** 2)
    g, op, ph = _env(terrain=h, N=0.01)
    st = SplitExplicitStepper(ph, dt=6.0, ns=3)
    bg = _bg(ph)
    rng = np.random.default_rng(1)
    pi = 1e-4 * rng.standard_normal(g.Z_m.shape)
    w = rng.standard_normal(g.Z_w.shape)
    u = rng.standard_normal(g.Z_u.shape)
    v = rng.standard_normal(g.Z_v.shape)
    div_h = 1e-3 * rng.standard_normal(g.Z_m.shape)
    sf_pi = 1e-5 * rng.standard_normal(g.Z_m.shape)
    sf_w = 1e-2 * rng.standard_normal(g.Z_w.shape)
    dtau = 2.0
    p1, w1 = st._solve_vertical_acoustic_column(pi, w, u, v, div_h, sf_pi, sf_w, bg['th_v_w'], bg['C_pi'],
                                                bg['dz_m_full'], bg['dz_w_full'], bg, dtau)
    # the pair solves the coupled off-centred (alpha) column equations, not
    # only the eliminated tridiagonal system:
    #   w1  = w  + dtau (sf_w  - cp th_w d_z[(1 - alpha) pi + alpha pi1])      (interior levels)
    #   pi1 = pi + dtau (sf_pi - C_pi (div_h + d_z[rho th ((1 - alpha) w + alpha w1 - w_terrain)]))
    a, cp = st.alpha, C['cp']
    dzdz = lambda p: op.diff(p, 2, 'm', 'w') * (g.dz / bg['dz_w_full'])
    w_want = w + dtau * (sf_w - cp * bg['th_v_w'] * dzdz((1 - a) * pi + a * p1))
    np.testing.assert_allclose(w1[:, :, 1:-1], w_want[:, :, 1:-1], rtol=1e-12, atol=1e-12 * np.abs(w).max())
    uw = op.avg(op.avg(u, 0, 'u', 'm'), 2, 'm', 'w')
    vw = op.avg(op.avg(v, 1, 'v', 'm'), 2, 'm', 'w')
    mw = g.m_factors['w'][..., None]
    edz = (1 - a) * w + a * w1 - mw * (uw * g.z_xi_w + vw * g.z_eta_w)
    edz[:, :, [0, -1]] = 0.0
    dz_flux = op.diff(edz * bg['rho_w'] * bg['th_v_w'], 2, 'w', 'm') * g.dz / bg['dz_m_full']
    p_want = pi + dtau * (sf_pi - bg['C_pi'] * (div_h + dz_flux))
    np.testing.assert_allclose(p1, p_want, rtol=1e-12, atol=1e-12 * np.abs(p1 - pi).max())
    # the implicit coupling matters here (a plain forward step misses by far more)
    assert np.abs(p1 - (pi + dtau * (sf_pi - bg['C_pi'] * div_h))).max() > 1e-6 * np.abs(p1 - pi).max()
    # the kinematic surface and rigid lid hold
    uw = op.avg(op.avg(u, 0, 'u', 'm'), 2, 'm', 'w')
    vw = op.avg(op.avg(v, 1, 'v', 'm'), 2, 'm', 'w')
    ws = g.m_factors['w'] * (uw[:, :, 0] * g.z_xi_w[:, :, 0] + vw[:, :, 0] * g.z_eta_w[:, :, 0])
    np.testing.assert_allclose(w1[:, :, 0], ws, rtol=1e-14)
    assert np.all(w1[:, :, -1] == 0.0)


def test_split_explicit_resting_atmosphere_over_terrain_stays_at_rest():
    def h(x, y):
        return 600.0 / (1.0 + (x ** 2 + y ** 2) / 5000.0 ** 2)
    g = O.Grid(10, 9, 8, 2000.0, 2000.0, 500.0, 45.0, 0.0, h_func=h)
    op = O.Ops(g)
    ph = O.Euler(g, op, C, dt=4.0, N_bv=0.01)
    rest = {'pi': ph.pi_bg.copy(), 'th_v': ph.theta_bg.copy()}
    st, dt = O.build_dynamical_core('split-explicit', g, op, C, rest, dt=4.0, ns=3)
    assert isinstance(st, SplitExplicitStepper) and dt == 4.0 and st.ns == 3 and st.alpha == 0.55
    ph = st.physics
    s = {'u': np.zeros_like(g.Z_u), 'v': np.zeros_like(g.Z_v), 'w': np.zeros_like(g.Z_w),
         'eta_dot': np.zeros_like(g.Z_w), 'pi': ph.pi_bg.copy(), 'th_v': ph.theta_bg.copy(),
         'rho': ph.bg_reference()['rho']}
    for i in range(5):
        s = st.step(s, i * 4.0, None, lambda a, b: a)
    assert np.max(np.abs(s['w'])) < 1e-9 and np.max(np.abs(s['u'])) < 1e-9


# --- Fourier-symbol pins of the time discretisation ---------------------------------------
# With g = 0 and N = 0 the background is uniform (theta 300 K, Exner 1), and
# with unit map factors and no Coriolis a mode cos(k x) that is uniform in y
# and z keeps w = 0.  Linearised about it, the grid operators act on
# e^{ikx} as multiplications: the face gradient and the centred divergence
# both by s = 2i sin(k dx / 2) / dx, so every stage of the step is a 2x2
# complex matrix on (u, pi) and the step's amplification follows without
# the grid code.

def _flat_env(nx=16, dx=2000.0, dt=12.0, ns=5, nu_div_factor=0.0):
    c0 = dict(C, g=0.0)
    g = O.Grid(nx, 3, 4, dx, dx, 500.0, 0.0, 0.0)
    for k in g.m_factors:
        g.m_factors[k] = np.ones_like(g.m_factors[k])
    g.dm_dx_m = np.zeros_like(g.dm_dx_m)
    g.dm_dy_m = np.zeros_like(g.dm_dy_m)
    g.f_u = np.zeros_like(g.f_u)
    g.f_v = np.zeros_like(g.f_v)
    op = O.Ops(g, (0, 1))
    ph = O.Euler(g, op, c0, dt=dt, N_bv=0.0, damp_height=1e9, max_damp=0.0, nu_div_factor=nu_div_factor,
                 nu_h_factor=0.0)
    assert np.all(ph.pi_bg == 1.0) and np.all(ph.theta_bg == 300.0)
    return g, ph, SplitExplicitStepper(ph, dt=dt, ns=ns, alpha=0.5)


def _upwind_symbol(t, sign):
    """Face value of e^{ikx} relative to e^{ik x_face} (third-order upwind, t = k dx)."""
    return ((7 / 6) * np.cos(t / 2) - (1 / 6) * np.cos(3 * t / 2)
            + sign * (1j / 6) * (np.sin(3 * t / 2) - 3 * np.sin(t / 2)))


def _acoustic_symbol(s, a, b, dt, ns, adv=0.0, nu_div=0.0):
    """Wicker-Skamarock RK3 with forward-backward acoustic substeps
    (steppers.py:1092-1116, 1472-1651) on d/dt (u, pi) = [[adv, -a s], [-b s, 0]] (u, pi);
    ``adv`` (momentum advection) is a slow term, held at the stage state
    through the substeps; divergence damping nu_div grad(div u) acts on the
    total wind inside the substeps (predictor at the old, corrector at the
    half-step wind).  Python's round (half to even) sets the stage substep
    counts, as in the reference."""
    M = np.array([[adv, -a * s], [-b * s, 0.0]])
    n1, n2, n3 = max(1, round(ns / 3)), max(1, round(ns / 2)), ns

    def acoustic(x0, xc, dtau, n):
        slow = M @ xc
        up, pp = x0 - xc
        for _ in range(n):
            gox = s * pp
            u_half = up + 0.5 * dtau * (slow[0] - a * gox + nu_div * s * s * (xc[0] + up))
            p_new = pp + dtau * (slow[1] - b * s * u_half)
            up = up + dtau * (slow[0] - 0.5 * a * (gox + s * p_new) + nu_div * s * s * (xc[0] + u_half))
            pp = p_new
        return xc + np.array([up, pp])

    G = np.zeros((2, 2), complex)
    for j in range(2):
        x0 = np.eye(2, dtype=complex)[j]
        x1 = acoustic(x0, x0, (dt / 3) / n1, n1)
        x2 = acoustic(x0, x1, (dt / 2) / n2, n2)
        G[:, j] = acoustic(x0, x2, dt / n3, n3)
    return G


@pytest.mark.parametrize("ns,mode,u0,ndf", [(5, 1, 0.0, 0.0), (4, 3, 12.0, 0.0), (7, 2, -9.0, 0.0),
                                           (5, 1, 15.0, 0.0), (6, 3, 7.0, 0.05)])
def test_horizontal_acoustic_mode_follows_the_rk3_forward_backward_symbol(ns, mode, u0, ndf):
    # with a uniform wind u0 the momentum advection (face -> centre average,
    # upwind advection, centre -> face average) is a slow term, so the stage
    # states and the substep counts of all three stages enter
    g, ph, st = _flat_env(ns=ns, nu_div_factor=ndf)
    k = 2 * np.pi * mode / (g.nx * g.dx)
    t = k * g.dx
    s = 2j * np.sin(t / 2) / g.dx
    a, b = C['cp'] * 300.0, C['Rd'] / C['cvd']
    adv = -u0 * np.cos(t / 2) ** 2 * _upwind_symbol(t, np.sign(u0)) * s
    G = _acoustic_symbol(s, a, b, st.dt, ns, adv, ph.nu_div)
    # second-order terms (momentum advection, Exner nonlinearity) scale as
    # eps relative to the response: 1e-7 puts them near rounding
    eps = 1e-7
    xm = g.x_m[:, None, None] + np.zeros((1, g.ny, g.nz))
    xu = g.x_c[:, None, None] + np.zeros((1, g.ny, g.nz))
    s0 = {'u': np.full(g.Z_u.shape, u0), 'v': np.zeros_like(g.Z_v), 'w': np.zeros_like(g.Z_w),
          'eta_dot': np.zeros_like(g.Z_w), 'pi': 1.0 + eps * np.cos(k * xm), 'th_v': ph.theta_bg.copy()}
    s0['rho'] = C['p0'] / (C['Rd'] * s0['th_v']) * s0['pi'] ** (C['cvd'] / C['Rd'])
    r = st.step(dict(s0), 0.0, None, lambda x, f: x)
    r['u'] = r['u'] - u0
    want_p = eps * np.real(G[1, 1] * np.exp(1j * k * xm))
    want_u = eps * np.real(G[0, 1] * np.exp(1j * k * xu))
    assert abs(G[1, 1] - 1) > 0.05 and abs(G[0, 1]) * eps > 1e-9  # a step that moves the wave
    np.testing.assert_allclose(r['pi'] - 1.0, want_p, rtol=0, atol=5e-8 * eps)
    np.testing.assert_allclose(r['u'], want_u, rtol=0, atol=5e-8 * np.abs(want_u).max())
    assert np.abs(r['w']).max() < 1e-12 * np.abs(want_u).max()


@pytest.mark.parametrize("mode", [1, 2, 4])
def test_theta_advection_follows_the_rk3_taylor_polynomial(mode):
    # uniform u0 advects theta' by the third-order upwind operator; to first
    # order in theta' the step multiplies a mode by 1 + z + z^2/2 + z^3/6,
    # z = dt sigma(k), whatever the pressure response (it does not feed back)
    g, ph, st = _flat_env(dt=20.0, ns=6)
    u0 = 12.0
    k = 2 * np.pi * mode / (g.nx * g.dx)
    t = k * g.dx
    z = -st.dt * u0 * _upwind_symbol(t, 1.0) * 2j * np.sin(t / 2) / g.dx
    amp = 1 + z + z ** 2 / 2 + z ** 3 / 6
    eps = 1e-5
    xm = g.x_m[:, None, None] + np.zeros((1, g.ny, g.nz))
    s0 = {'u': np.full(g.Z_u.shape, u0), 'v': np.zeros_like(g.Z_v), 'w': np.zeros_like(g.Z_w),
          'eta_dot': np.zeros_like(g.Z_w), 'pi': ph.pi_bg.copy(), 'th_v': 300.0 + eps * np.cos(k * xm)}
    s0['rho'] = C['p0'] / (C['Rd'] * s0['th_v']) * s0['pi'] ** (C['cvd'] / C['Rd'])
    r = st.step(dict(s0), 0.0, None, lambda x, f: x)
    want = eps * np.real(amp * np.exp(1j * k * xm))
    assert abs(abs(amp) - 1) > 1e-6 and abs(np.angle(amp)) > 1e-3
    np.testing.assert_allclose(r['th_v'] - 300.0, want, rtol=0, atol=1e-6 * eps)


def _terrain_env(nu_div_factor=0.05, max_damp=2.0, damp_height=2500.0):
    def h(x, y):
        return 500.0 * np.exp(-(x ** 2 + (0.5 * y) ** 2) / 5000.0 ** 2)
    g = O.Grid(12, 10, 8, 2000.0, 2000.0, 500.0, 45.0, 0.0, h_func=h)
    g.f_u = np.zeros_like(g.f_u)
    g.f_v = np.zeros_like(g.f_v)
    op = O.Ops(g, (1,))
    ph = O.Euler(g, op, C, dt=12.0, N_bv=0.01, damp_height=damp_height, max_damp=max_damp,
                 nu_div_factor=nu_div_factor, nu_h_factor=0.0)
    return g, ph, SplitExplicitStepper(ph, dt=12.0, ns=4, alpha=0.55)


def _perturbed(g, ph, seed, eps):
    rng = np.random.default_rng(seed)
    s = {'u': eps * rng.standard_normal(g.Z_u.shape), 'v': eps * rng.standard_normal(g.Z_v.shape),
         'w': eps * rng.standard_normal(g.Z_w.shape), 'pi': ph.pi_bg + 1e-3 * eps * rng.standard_normal(g.Z_m.shape),
         'th_v': ph.theta_bg.copy()}
    # periodic axes carry one face twice
    if 0 in ph.op.periodic_axes:
        s['u'][-1] = s['u'][0]
    if 1 in ph.op.periodic_axes:
        s['v'][:, -1] = s['v'][:, 0]
    s['rho'] = C['p0'] / (C['Rd'] * s['th_v']) * s['pi'] ** (C['cvd'] / C['Rd'])
    return s


def test_acoustic_loop_is_independent_of_the_stage_state_in_the_linear_regime():
    # The loop integrates the deviation from the stage state with the fast
    # terms (pressure gradient with its terrain-slope part, the implicit
    # column, divergence damping, Rayleigh damping, the open-boundary faces),
    # while the slow tendencies hold the same fast terms at the stage state.
    # For a linear fast system the two parts add up to the total, so the
    # result does not depend on which stage state the loop starts from -- a
    # property of the splitting that breaks as soon as the loop's fast
    # operator and the slow tendencies' differ.  Slow terms that are not fast
    # (advection, metric terms) are quadratic, so a small amplitude bounds them.
    g, ph, st = _terrain_env()
    bg = ph.precompute_bg(ph.bg_reference())
    eps = 1e-6
    x0 = _perturbed(g, ph, 1, eps)
    x0['eta_dot'], x0['w'] = st._diagnose_eta_dot(x0)
    outs = []
    for seed in (2, 3):
        d = _perturbed(g, ph, seed, eps)
        xc = {k: (x0[k] + d[k] - (ph.pi_bg if k == 'pi' else 0.0)) if k in ('u', 'v', 'w', 'pi') else x0[k]
              for k in x0}
        xc['eta_dot'], xc['w'] = st._diagnose_eta_dot(xc)
        slow = st._compute_slow_tendencies(xc, bg, {}, {}, None)
        outs.append(st._acoustic_loop(x0, xc, slow, bg, 1.5, 4))
    for k in ('u', 'v', 'w', 'pi'):
        ref = x0[k] - (ph.pi_bg if k == 'pi' else 0.0)
        change = np.abs(outs[0][k] - x0[k]).max()
        diff = np.abs(outs[0][k] - outs[1][k]).max()
        assert change > 1e-3 * np.abs(ref).max(), k
        assert diff < 1e-6 * change, (k, diff, change)


def test_rayleigh_damping_acts_above_the_damping_height():
    # a strong sponge (tau dtau >> 1) leaves w near zero where tau > 0 and
    # nowhere else
    out = {}
    for md in (0.0, 1e4):
        g, ph, st = _terrain_env(nu_div_factor=0.0, max_damp=md, damp_height=2500.0)
        s = _perturbed(g, ph, 4, 1.0)
        s['eta_dot'], s['w'] = st._diagnose_eta_dot(s)
        out[md] = st.step(s, 0.0, None, lambda x, f: x)['w']
    tau = ph.tau_damp
    layer = tau * (st.dt / st.ns) > 1e3
    assert layer.any() and (tau == 0).any()
    assert np.abs(out[1e4][layer]).max() < 1e-2 * np.abs(out[0.0][layer]).max()
    free = np.where(tau == 0)
    assert np.abs(out[1e4][free]).max() > 0.1 * np.abs(out[0.0][free]).max()
