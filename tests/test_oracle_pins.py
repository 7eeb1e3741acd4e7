"""Independent pins of oracle functions that no reference test reaches.

Each assertion comes from a closed form or an invariant the mathematics
fixes, chosen so that a plausible slip in the restatement (a dropped metric
term, a wrong sign, a transposed axis, a missing map factor or scale) fails:

* HyperFilter (diffusion.py:80-172): the discrete Laplacian's eigenvalue on a
  flat periodic grid, its map-factor scaling, and second-order convergence of
  the terrain-corrected Cartesian Laplacian to the analytic derivative over
  terrain (including fields of physical height only, whose horizontal
  Cartesian derivatives vanish);
* the factory's background from the initial state (steppers.py:55-62,
  euler.py:76-112): exact recovery of a horizontally uniform theta(z) and the
  closed-form isentropic Exner profile;
* StretchedSleveSimple (transforms.py:54-86): surface and lid values and the
  stretch function against an independent closed form;
* the stereographic latitude and map factor (geometry.py:35-87): centre
  value, mirror symmetry, the meridian closed form, and the map factor as
  the stereographic scale 2 / (1 + cos c);
* grid metrics (geometry.py:221-261): terrain slopes of a linear ramp;
* the kinematic bottom boundary through the implicit solve
  (tests/test_dynamical_core_3d.py:335-382);
* relative_residual (steppers.py:498-527): zero for an exact solve and the
  scaled norm of a known defect;
* BenchmarkSponge (boundaries.py:123-189): weights at the edge, at half depth
  and beyond, the 1 - (1-wx)(1-wy) combination and the staggered padding.
"""
import numpy as np
import pytest

from oracle import sisl as O

C = {'g': 9.81, 'cp': 1004.0, 'Rd': 287.0, 'cvd': 717.0, 'p0': 100000.0}


def unit_map_factors(g, op):
    for loc in ('m', 'u', 'v', 'w'):
        g.m_factors[loc] = np.ones_like(g.m_factors[loc])
    op.m3 = {k: v[..., None] for k, v in g.m_factors.items()}


def flat_bg(g, op):
    ph = O.Euler(g, op, C, dt=10.0, N_bv=0.01)
    return ph, ph.precompute_bg(ph.bg_reference())


# --- HyperFilter: flat periodic grid --------------------------------------------
def lam(k, d):
    """Eigenvalue of the 3-point second difference: (2 - 2 cos k d) / d^2."""
    return (2.0 - 2.0 * np.cos(k * d)) / d ** 2


def periodic_flat(nx=16, ny=12, nz=6, dx=700.0, dy=900.0):
    g = O.Grid(nx, ny, nz, dx, dy, 300.0, 45.0, 0.0)
    op = O.Ops(g, (0, 1))
    unit_map_factors(g, op)
    return g, op


def test_hyperfilter_laplacian_eigenvalue_on_flat_periodic_grid():
    g, op = periodic_flat()
    ph, bg = flat_bg(g, op)
    kx, ky = 2 * np.pi * 3 / (g.nx * g.dx), 2 * np.pi * 2 / (g.ny * g.dy)
    x = g.x_m[:, None, None]
    y = g.y_m[None, :, None]
    zfac = (1.0 + 0.1 * np.arange(g.nz))[None, None, :]   # any z dependence passes through
    f = np.cos(kx * x + 0.3) * np.cos(ky * y - 0.2) * zfac
    hf = O.HyperFilter(g, op, nu_h=1.0, nu_v=0.0)
    L = -(lam(kx, g.dx) + lam(ky, g.dy))
    np.testing.assert_allclose(hf._lap_h(f, bg), L * f, rtol=0, atol=1e-12 * abs(L))
    np.testing.assert_allclose(hf._lap_h(hf._lap_h(f, bg), bg), L * L * f, rtol=0, atol=1e-12 * L * L)


def test_hyperfilter_tendencies_closed_form_per_variable():
    """u, v, w and theta' tendencies of -nu_h del^4 on single modes; u and v
    go through the face-to-mass and mass-to-face averages, which multiply a
    mode by cos(k d / 2) each way."""
    g, op = periodic_flat()
    ph, bg = flat_bg(g, op)
    nu = 3.7e5
    hf = O.HyperFilter(g, op, nu_h=nu, nu_v=0.0)
    kx, ky = 2 * np.pi * 2 / (g.nx * g.dx), 2 * np.pi / (g.ny * g.dy)
    L = lam(kx, g.dx) + lam(ky, g.dy)
    xu, xm = g.x_c[:, None, None], g.x_m[:, None, None]
    yv, ym = g.y_c[None, :, None], g.y_m[None, :, None]
    one_z = np.ones((1, 1, g.nz))
    u = np.cos(kx * xu) * np.cos(ky * ym) * one_z
    v = np.sin(kx * xm) * np.cos(ky * yv) * one_z
    w = np.cos(kx * xm) * np.sin(ky * ym) * np.ones((1, 1, g.nz + 1))
    thp = np.sin(kx * xm + 1.0) * np.sin(ky * ym) * one_z
    sp = {'u': u, 'v': v, 'w': w, 'th_v': bg['th_v'] + thp}
    t = hf.get_tendencies(sp, bg)
    cx2, cy2 = np.cos(kx * g.dx / 2) ** 2, np.cos(ky * g.dy / 2) ** 2
    tol = 1e-11 * nu * L * L
    np.testing.assert_allclose(t['u'], -nu * L * L * cx2 * u, rtol=0, atol=tol)
    np.testing.assert_allclose(t['v'], -nu * L * L * cy2 * v, rtol=0, atol=tol)
    np.testing.assert_allclose(t['w'], -nu * L * L * w, rtol=0, atol=tol)
    np.testing.assert_allclose(t['th_v'], -nu * L * L * thp, rtol=0, atol=tol)


def test_hyperfilter_map_factor_enters_every_horizontal_difference():
    """A constant map factor c scales each horizontal difference by c, so the
    Laplacian by c^2 and the hyper-Laplacian by c^4."""
    g, op = periodic_flat()
    ph, bg = flat_bg(g, op)
    rng = np.random.default_rng(4)
    f = rng.standard_normal(g.Z_m.shape)
    hf = O.HyperFilter(g, op, nu_h=1.0, nu_v=0.0)
    ref = hf._lap_h(hf._lap_h(f, bg), bg)
    c = 1.3
    for loc in ('m', 'u', 'v', 'w'):
        g.m_factors[loc] = np.full_like(g.m_factors[loc], c)
    op.m3 = {k: v[..., None] for k, v in g.m_factors.items()}
    got = hf._lap_h(hf._lap_h(f, bg), bg)
    np.testing.assert_allclose(got, c ** 4 * ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref)))


# --- HyperFilter over terrain: Cartesian derivatives at fixed physical height ----
def terrain_grid(n):
    """x-z slice over a smooth ridge, nz = n / 2, refined together."""
    L, H = 40000.0, 8000.0

    def h(x, y):
        return 600.0 * np.exp(-(x / 8000.0) ** 2) + 0.0 * y
    g = O.Grid(n, 3, n // 2, L / n, L / n, H / (n // 2), 45.0, 0.0, h_func=h)
    op = O.Ops(g)
    unit_map_factors(g, op)
    ph, bg = flat_bg(g, op)
    return g, op, bg, L, H


def interior(a, g, mx, mz):
    return a[mx:g.nx - mx, 1, mz:g.nz - mz]


def test_cartesian_laplacian_over_terrain_converges_at_second_order():
    errs = []
    for n in (32, 64, 128):
        g, op, bg, L, H = terrain_grid(n)
        kx, kz = 2 * np.pi / L * 1.5, np.pi / H
        X, Z = g.x_m[:, None, None] + 0.0 * g.Z_m, g.Z_m
        f = np.sin(kx * X) * np.cos(kz * Z)
        hf = O.HyperFilter(g, op, nu_h=1.0, nu_v=0.0)
        lap = hf._lap_h(f, bg)
        # at fixed physical height only x varies: d2/dx2 = -kx^2 f
        exact = -kx * kx * f
        mx, mz = n // 8, n // 16
        errs.append(np.max(np.abs(interior(lap - exact, g, mx, mz))) / (kx * kx))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert errs[2] < 2e-3, errs
    assert r1 > 3.3 and r2 > 3.6, (errs, r1, r2)


def test_fields_of_height_only_have_no_cartesian_horizontal_laplacian():
    """f = F(z) has zero Cartesian x-derivatives; g = x F(z) has d/dx = F(z)
    and d2/dx2 = 0, which exercises the second-level metric term.  Both
    vanish up to the O(dx^2) truncation; a sign slip in either metric term
    leaves an O(1) remainder."""
    res = []
    for n in (32, 64):
        g, op, bg, L, H = terrain_grid(n)
        hf = O.HyperFilter(g, op, nu_h=1.0, nu_v=0.0)
        Fz = np.cos(np.pi * g.Z_m / H)
        X = g.x_m[:, None, None] + 0.0 * g.Z_m
        mx, mz = n // 8, n // 16
        scale = (np.pi / H) ** 2
        a = np.max(np.abs(interior(hf._lap_h(Fz, bg), g, mx, mz))) / scale
        b = np.max(np.abs(interior(hf._lap_h(X / L * Fz, bg), g, mx, mz))) / scale
        res.append((a, b))
    (a1, b1), (a2, b2) = res
    assert a2 < 1e-3 and b2 < 1e-3, res
    assert a1 / a2 > 3.0 and b1 / b2 > 3.0, res


def test_terrain_metric_terms_are_not_optional():
    """Guard on the previous pin: the uncorrected (coordinate-surface)
    Laplacian of F(z) over the same terrain is O(1), so the vanishing there
    is the metric correction's doing."""
    g, op, bg, L, H = terrain_grid(64)
    Fz = np.cos(np.pi * g.Z_m / H)
    naive = op.diff(op.diff(Fz, 0, 'm', 'u'), 0, 'u', 'm')
    hf = O.HyperFilter(g, op, nu_h=1.0, nu_v=0.0)
    corrected = hf._lap_h(Fz, bg)
    assert np.max(np.abs(interior(naive, g, 8, 4))) > 100.0 * np.max(np.abs(interior(corrected, g, 8, 4)))


# --- factory background from the initial state -----------------------------------
def test_background_from_uniform_theta_profile_is_exact():
    g = O.Grid(10, 8, 12, 1000.0, 1000.0, 400.0, 45.0, 0.0)
    op = O.Ops(g)
    theta_z = 290.0 + 0.004 * g.Z_m + 0.3 * np.sin(g.Z_m / 900.0)
    rng = np.random.default_rng(1)
    # a horizontal pattern with zero mean on every level does not move the mean
    pat = rng.standard_normal((g.nx, g.ny, 1))
    pat -= pat.mean()
    pat = pat - pat.mean(axis=(0, 1), keepdims=True)
    st = {'u': np.zeros_like(g.Z_u), 'v': np.zeros_like(g.Z_v), 'w': np.zeros_like(g.Z_w),
          'pi': 0.9 + 0.0 * g.Z_m - 1e-5 * g.Z_m, 'th_v': theta_z + pat * 0.0,
          'rho': np.ones_like(g.Z_m)}
    stepper, _ = O.build_dynamical_core('sisl', g, op, C, st, dt=10.0)
    np.testing.assert_array_equal(stepper.physics.theta_bg, theta_z)
    pi_top = float(np.mean(st['pi'][:, :, -1]))
    np.testing.assert_allclose(stepper.physics.pi_bg[:, :, -1], pi_top, rtol=0, atol=1e-15)
    # with a zero-mean pattern the exactly rounded mean is still theta(z)
    st2 = dict(st, th_v=theta_z + 0.5 * pat)
    st2 = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in st2.items()}
    stepper2, _ = O.build_dynamical_core('sisl', g, op, C, st2, dt=10.0)
    np.testing.assert_allclose(stepper2.physics.theta_bg, theta_z, rtol=0, atol=1e-12)


def test_isentropic_background_exner_is_linear_in_height():
    """theta = theta0 everywhere: the hydrostatic integral from the lid is
    pi(z) = pi_top + g (z_top - z) / (cp theta0) exactly."""
    g = O.Grid(6, 5, 20, 1000.0, 1000.0, 250.0, 45.0, 0.0)
    op = O.Ops(g)
    th0 = 300.0
    st = {'u': np.zeros_like(g.Z_u), 'v': np.zeros_like(g.Z_v), 'w': np.zeros_like(g.Z_w),
          'pi': np.full_like(g.Z_m, 0.85), 'th_v': np.full_like(g.Z_m, th0), 'rho': np.ones_like(g.Z_m)}
    stepper, _ = O.build_dynamical_core('sisl', g, op, C, st, dt=10.0)
    ztop = g.Z_m[:, :, -1:]
    want = 0.85 + C['g'] * (ztop - g.Z_m) / (C['cp'] * th0)
    np.testing.assert_allclose(stepper.physics.pi_bg, want, rtol=0, atol=1e-14)


def test_analytic_background_profile_is_hydrostatic_and_stable():
    """N_bv > 0 (no initial state): theta = theta0 exp(N^2 z / g), and the
    discrete Exner profile decreases with height at the rate -g / (cp theta)
    to second order in dz."""
    g = O.Grid(4, 4, 80, 1000.0, 1000.0, 125.0, 45.0, 0.0)
    op = O.Ops(g)
    ph = O.Euler(g, op, C, dt=10.0, N_bv=0.01)
    th = ph.theta_bg[0, 0]
    z = g.Z_m[0, 0]
    np.testing.assert_allclose(th, 300.0 * np.exp(1e-4 * z / C['g']), rtol=1e-14)
    dpidz = np.diff(ph.pi_bg[0, 0]) / np.diff(z)
    mid = 0.5 * (th[1:] + th[:-1])
    np.testing.assert_allclose(dpidz, -C['g'] / (C['cp'] * mid), rtol=1e-12)
    # continuous solution: pi(z) - pi(z_top) = g^2/(cp theta0 N^2) (exp(-N^2 z/g) - exp(-N^2 ztop/g))
    N2 = 1e-4
    cont = C['g'] ** 2 / (C['cp'] * 300.0 * N2) * (np.exp(-N2 * z / C['g']) - np.exp(-N2 * z[-1] / C['g']))
    np.testing.assert_allclose(ph.pi_bg[0, 0] - ph.pi_bg[0, 0, -1], cont, rtol=0, atol=2e-7)


# --- vertical transform ---------------------------------------------------------------
def test_stretched_sleve_surface_lid_and_stretch():
    Lz = 20000.0
    tr = O.StretchedSleveSimple(stretch_kappa=np.log(2.0), scale_s=4000.0, n=1.35)
    zeta = np.linspace(0.0, Lz, 11)
    h = np.full_like(zeta, 350.0)
    z = tr(None, zeta, h, Lz)
    assert z[0] == pytest.approx(350.0, abs=1e-12)          # follows the terrain at the surface
    assert abs(z[-1] - Lz) < 1e-6                            # flat lid
    assert np.all(np.diff(z) > 0)
    # flat terrain: Lz (2^eta - 1) for kappa = ln 2 (independent closed form)
    z0 = tr(None, zeta, np.zeros_like(zeta), Lz)
    np.testing.assert_allclose(z0, Lz * (2.0 ** (zeta / Lz) - 1.0), rtol=1e-13, atol=1e-9)
    # kappa -> 0 recovers the unstretched coordinate
    tr0 = O.StretchedSleveSimple(stretch_kappa=1e-7)
    np.testing.assert_allclose(tr0(None, zeta, np.zeros_like(zeta), Lz), zeta, rtol=1e-6, atol=1e-6)
    # the terrain signal decays with height as sinh((Lz - z)/s) / sinh(Lz/s) to the power n
    mid = tr(None, np.array([Lz / 2]), np.array([350.0]), Lz)[0] - \
        tr(None, np.array([Lz / 2]), np.array([0.0]), Lz)[0]
    zs = Lz * (np.sqrt(2.0) - 1.0)
    assert mid == pytest.approx(350.0 * (np.sinh((Lz - zs) / 4000.0) / np.sinh(Lz / 4000.0)) ** 1.35,
                                rel=1e-12)


# --- latitude and map factor ------------------------------------------------------------
def test_latitude_centre_symmetry_and_meridian():
    lat_c = 47.0
    assert O.latitude_deg(np.array(0.0), np.array(0.0), lat_c) == pytest.approx(lat_c, abs=1e-12)
    x = np.array([1.3e5, 4.0e5, 9.0e5])
    y = np.array([-2.0e5, 3.0e4, 6.0e5])
    np.testing.assert_allclose(O.latitude_deg(x, y, lat_c), O.latitude_deg(-x, y, lat_c), rtol=0, atol=1e-12)
    # along the central meridian the stereographic angular distance is 2 atan(r / 2R)
    yy = np.array([-8.0e5, -1.0e5, 2.0e5, 7.0e5])
    c = 2.0 * np.arctan(np.abs(yy) / (2.0 * O.R_EARTH))
    want = lat_c + np.sign(yy) * np.degrees(c)
    np.testing.assert_allclose(O.latitude_deg(np.zeros_like(yy), yy, lat_c), want, rtol=0, atol=1e-10)


def test_map_factor_is_the_stereographic_scale():
    r = np.array([0.0, 1e5, 5e5, 2e6])
    c = 2.0 * np.arctan(r / (2.0 * O.R_EARTH))
    np.testing.assert_allclose(O.map_factor(r, 0.0 * r), 2.0 / (1.0 + np.cos(c)), rtol=1e-14)
    np.testing.assert_allclose(O.map_factor(0.6 * r, 0.8 * r), O.map_factor(r, 0.0 * r), rtol=1e-15)


# --- grid metrics ---------------------------------------------------------------------------
def test_terrain_slopes_of_a_linear_ramp():
    a = 0.02

    def h(x, y):
        return 500.0 + a * x + 0.0 * y
    g = O.Grid(12, 4, 6, 1000.0, 1000.0, 500.0, 45.0, 0.0, h_func=h)
    Lz = g.Lz
    zeta = g.z_c[None, None, :]
    # Gal-Chen: z = h + zeta (Lz - h) / Lz, so dz/dx = a (1 - zeta / Lz) (centred, exact)
    want = a * (1.0 - zeta / Lz) + 0.0 * g.Z_w
    np.testing.assert_allclose(g.z_xi_w[1:-1], want[1:-1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(g.z_xi_w[[0, -1]], 0.5 * want[[0, -1]], rtol=0, atol=1e-12)  # edge-padded
    np.testing.assert_allclose(g.z_eta_w, 0.0, atol=1e-12)
    # extrapolated end spacings: uniform in zeta, so every dz_w equals (Lz - h)/nz * 1
    dz = (Lz - h(g.x_m, 0.0)) / g.nz
    np.testing.assert_allclose(g.dz_w_full, np.broadcast_to(dz[:, None, None], g.dz_w_full.shape), rtol=1e-12)
    np.testing.assert_allclose(g.dz_m_full, np.broadcast_to(dz[:, None, None], g.dz_m_full.shape), rtol=1e-12)


# --- kinematic bottom boundary: tests/test_dynamical_core_3d.py:335-382 ------------------
def test_kinematic_bottom_boundary_survives_the_implicit_solve():
    def h(x, y):
        return 800.0 * np.exp(-(x ** 2 + y ** 2) / 4000.0 ** 2)
    nx, ny, nz = 14, 12, 10
    g = O.Grid(nx, ny, nz, 1000.0, 1000.0, 600.0, 45.0, 0.0, h_func=h)
    op = O.Ops(g)
    ph = O.Euler(g, op, C, dt=20.0, N_bv=0.01)
    bg = ph.precompute_bg(ph.bg_reference())
    rhs = {'u': np.full((nx + 1, ny, nz), 10.0), 'v': np.zeros((nx, ny + 1, nz)),
           'w': np.zeros((nx, ny, nz + 1)), 'pi': np.zeros((nx, ny, nz)),
           'eta_dot': np.zeros((nx, ny, nz + 1))}
    uw = op.avg(op.avg(rhs['u'], 0, 'u', 'm'), 2, 'm', 'w')
    vw = op.avg(op.avg(rhs['v'], 1, 'v', 'm'), 2, 'm', 'w')
    kb = g.m_factors['w'] * (uw[:, :, 0] * g.z_xi_w[:, :, 0] + vw[:, :, 0] * g.z_eta_w[:, :, 0])
    rhs['w'][:, :, 0] = kb
    sol = O.ImplicitSolver(ph, 20.0, solver_tol=1e-10, solver_maxiter=40, solver_restart=30).solve(rhs, bg)
    assert np.max(kb) > 0.5
    np.testing.assert_allclose(np.max(sol['w'][:, :, 0]), np.max(kb), rtol=1e-4)
    # the surface row is an identity row of A: the whole bottom level is preserved
    np.testing.assert_allclose(sol['w'][:, :, 0], kb, rtol=0, atol=1e-8)


# --- relative residual: steppers.py:498-527 -------------------------------------------------
def test_relative_residual_of_exact_and_perturbed_solutions():
    def h(x, y):
        return 300.0 * np.exp(-(x ** 2 + y ** 2) / 3000.0 ** 2)
    g = O.Grid(8, 7, 6, 1000.0, 1000.0, 500.0, 45.0, 0.0, h_func=h)
    op = O.Ops(g)
    ph = O.Euler(g, op, C, dt=20.0, N_bv=0.01)
    bg = ph.precompute_bg(ph.bg_reference())
    solver = O.ImplicitSolver(ph, 20.0)
    rng = np.random.default_rng(3)
    x = {'u': rng.standard_normal(g.Z_u.shape), 'v': rng.standard_normal(g.Z_v.shape),
         'w': rng.standard_normal(g.Z_w.shape), 'pi': 1e-3 * rng.standard_normal(g.Z_m.shape),
         'eta_dot': 1e-3 * rng.standard_normal(g.Z_w.shape)}
    b = ph.linear_operator(x, bg, 20.0, alpha=solver.alpha)
    assert solver.relative_residual(x, b, bg) < 1e-14
    # a defect only in pi is weighted by the solver's pi scale (1000)
    d = 1e-6 * rng.standard_normal(g.Z_m.shape)
    b2 = dict(b, pi=b['pi'] + d)
    bs = sum(float(np.sum((b2[k] * (1000.0 if k == 'pi' else 1.0)) ** 2)) for k in b2)
    want = 1000.0 * np.linalg.norm(d) / np.sqrt(bs)
    assert solver.relative_residual(x, b2, bg) == pytest.approx(want, rel=1e-9)


# --- BenchmarkSponge: boundaries.py:123-189 --------------------------------------------------
def test_benchmark_sponge_weights():
    nx, ny, depth = 24, 16, 6
    sp = O.BenchmarkSponge(nx, ny, sponge_depth=depth, axes=('x', 'y'))
    m = sp.mask_m[:, :, 0].astype(np.float64)
    # an edge column is fully relaxed; half the depth gives cos^2(pi/4) = 1/2; beyond: none
    np.testing.assert_allclose(m[0, ny // 2], 1.0, atol=1e-7)
    np.testing.assert_allclose(m[depth // 2, ny // 2], 0.5, atol=1e-7)
    np.testing.assert_allclose(m[depth:nx - depth, depth:ny - depth], 0.0, atol=0)
    # symmetric in x and y
    np.testing.assert_array_equal(m, m[::-1, :])
    np.testing.assert_array_equal(m, m[:, ::-1])
    # combined as 1 - (1 - wx)(1 - wy): at (depth/2, depth/2) 1 - 1/4
    np.testing.assert_allclose(m[depth // 2, depth // 2], 0.75, atol=1e-7)
    # staggered masks repeat the last mass column / row
    np.testing.assert_array_equal(sp.mask_u[-1], sp.mask_u[-2])
    np.testing.assert_array_equal(sp.mask_v[:, -1], sp.mask_v[:, -2])
    assert sp.mask_u.shape == (nx + 1, ny, 1) and sp.mask_v.shape == (nx, ny + 1, 1)
    # x only (slice): no y relaxation, w not blended
    sx = O.BenchmarkSponge(nx, None, sponge_depth=depth)
    assert 'w' not in sx.blend_vars and sx.mask_m.shape == (nx, 1, 1)
    st = {'w': np.ones((nx, 3, 4)), 'u': np.zeros((nx + 1, 3, 4))}
    ext = {'w': np.zeros((nx, 3, 4)), 'u': np.ones((nx + 1, 3, 4))}
    out = sx.blend(st, ext)
    np.testing.assert_array_equal(out['w'], st['w'])
    np.testing.assert_allclose(out['u'][0], 1.0, atol=1e-7)
    np.testing.assert_allclose(out['u'][nx // 2], 0.0, atol=0)
