"""NumPy FP64 restatement of the reference SISL step (test infrastructure only).

Arrays use the reference layout: shape (x, y, z), z contiguous, staggered
arrays +1 on their axis (u: nx+1, v: ny+1, w / eta_dot: nz+1).
Citations are ``file:line`` in the reference repository.
"""
from __future__ import annotations

import numpy as np

EPS64 = np.finfo(np.float64).eps
R_EARTH = 6371229.0
OMEGA = 7.2921e-5


# ---------------------------------------------------------------------------
# Vertical transforms  (suetes/shared/transforms.py:16-86)
# ---------------------------------------------------------------------------
class GalChenSigma:
    """z = h + zeta (Lz - h) / Lz  (transforms.py:16-22)."""

    def __call__(self, xi, zeta, h, Lz):
        return h + zeta * (Lz - h) / Lz


class SleveSimple:
    """Single-scale SLEVE (transforms.py:38-51)."""

    def __init__(self, scale_s=4000.0, scale_l=15000.0, n=1.35):
        self.ss, self.sl, self.n = scale_s, scale_l, n

    def __call__(self, xi, zeta, h, Lz):
        zs = np.clip(zeta, 0.0, Lz)
        b = np.sinh((Lz - zs) / self.ss) / np.sinh(Lz / self.ss)
        return zs + h * np.clip(b, 1e-8, 1.0) ** self.n


class StretchedSleveSimple:
    """SLEVE with exponential stretching (transforms.py:54-86)."""

    def __init__(self, stretch_kappa=2.5, scale_s=4000.0, n=1.35):
        self.kappa, self.ss, self.n = stretch_kappa, scale_s, n

    def __call__(self, xi, zeta, h, Lz):
        eta = zeta / Lz
        if abs(self.kappa) > 1e-5:
            zst = Lz * (np.exp(self.kappa * eta) - 1.0) / (np.exp(self.kappa) - 1.0)
        else:
            zst = zeta
        zst = np.clip(zst, 0.0, Lz)
        b = np.sinh((Lz - zst) / self.ss) / np.sinh(Lz / self.ss)
        return zst + h * np.clip(b, 1e-8, 1.0) ** self.n


# ---------------------------------------------------------------------------
# Grid  (suetes/regional3d/geometry.py:13-272)
# ---------------------------------------------------------------------------
def map_factor(x, y):
    """m = 1 + (x^2+y^2)/(4R^2)  (geometry.py:35-52)."""
    return 1.0 + (x ** 2 + y ** 2) / (4.0 * R_EARTH ** 2)


def latitude_deg(x, y, lat_c, eps=1e-15):
    """Inverse oblique stereographic latitude (geometry.py:54-87)."""
    phic = np.radians(lat_c)
    rho = np.sqrt(x ** 2 + y ** 2)
    rho = np.where(rho == 0, eps, rho)
    c = 2.0 * np.arctan(rho / (2.0 * R_EARTH))
    lat = np.arcsin(np.cos(c) * np.sin(phic) + (y * np.sin(c) * np.cos(phic)) / rho)
    return np.degrees(lat)


class Grid:
    """Restates RegionalGrid3D (geometry.py:141-272)."""

    def __init__(self, nx, ny, nz, dx, dy, dz, lat_center, lon_center,
                 h_func=None, transform=None):
        self.shape = (nx, ny, nz)
        self.delta = (dx, dy, dz)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.dz = dx, dy, dz
        self.lat_c, self.lon_c = lat_center, lon_center
        Lx, Ly = nx * dx, ny * dy
        self.Lz = nz * dz
        self.x_c = np.arange(nx + 1) * dx - Lx / 2.0
        self.y_c = np.arange(ny + 1) * dy - Ly / 2.0
        self.z_c = np.arange(nz + 1) * dz
        self.x_m = 0.5 * (self.x_c[:-1] + self.x_c[1:])
        self.y_m = 0.5 * (self.y_c[:-1] + self.y_c[1:])
        self.z_m = 0.5 * (self.z_c[:-1] + self.z_c[1:])
        Xm, Ym = np.meshgrid(self.x_m, self.y_m, indexing='ij')
        Xu, Yu = np.meshgrid(self.x_c, self.y_m, indexing='ij')
        Xv, Yv = np.meshgrid(self.x_m, self.y_c, indexing='ij')
        self.m_factors = {'m': map_factor(Xm, Ym), 'u': map_factor(Xu, Yu),
                          'v': map_factor(Xv, Yv), 'w': map_factor(Xm, Ym)}
        self.dm_dx_m = Xm / (2.0 * R_EARTH ** 2)
        self.dm_dy_m = Ym / (2.0 * R_EARTH ** 2)
        self.f_m = 2.0 * OMEGA * np.sin(np.radians(latitude_deg(Xm, Ym, lat_center)))
        self.f_u = 2.0 * OMEGA * np.sin(np.radians(latitude_deg(Xu, Yu, lat_center)))
        self.f_v = 2.0 * OMEGA * np.sin(np.radians(latitude_deg(Xv, Yv, lat_center)))
        self.h_func = h_func if h_func is not None else (lambda x, y: 0.0)
        self.transform_op = transform if transform is not None else GalChenSigma()

        def apply(xc, yc, zc):
            Xi, Yi, Ze = np.meshgrid(xc, yc, zc, indexing='ij')
            return self.transform_op(Xi, Ze, self.h_func(Xi, Yi), self.Lz) + np.zeros_like(Xi)

        self.Z_m = apply(self.x_m, self.y_m, self.z_m)
        self.Z_u = apply(self.x_c, self.y_m, self.z_m)
        self.Z_v = apply(self.x_m, self.y_c, self.z_m)
        self.Z_w = apply(self.x_m, self.y_m, self.z_c)
        self.dz_m_full = self.Z_w[:, :, 1:] - self.Z_w[:, :, :-1]
        bot = self.Z_m[:, :, 0] - (self.Z_m[:, :, 1] - self.Z_m[:, :, 0])
        top = self.Z_m[:, :, -1] + (self.Z_m[:, :, -1] - self.Z_m[:, :, -2])
        zpad = np.concatenate([bot[..., None], self.Z_m, top[..., None]], axis=-1)
        self.dz_w_full = zpad[:, :, 1:] - zpad[:, :, :-1]
        zx = np.pad(self.Z_w, ((1, 1), (0, 0), (0, 0)), mode='edge')
        self.z_xi_w = (zx[2:] - zx[:-2]) / (2.0 * dx)
        zy = np.pad(self.Z_w, ((0, 0), (1, 1), (0, 0)), mode='edge')
        self.z_eta_w = (zy[:, 2:] - zy[:, :-2]) / (2.0 * dy)
        if float(np.min(self.dz_m_full)) <= 0.0:
            raise ValueError("Grid Tangling Detected! Minimum dz is "
                             f"{float(np.min(self.dz_m_full)):.2f} m.")


# ---------------------------------------------------------------------------
# C-grid operators  (suetes/regional3d/operators.py:12-132)
# ---------------------------------------------------------------------------
def _take(f, idx, axis):
    sl = [slice(None)] * f.ndim
    sl[axis] = idx
    return f[tuple(sl)]


class Ops:
    """Restates CGridOperator3D (operators.py:12-132).

    stop_gradient has no effect on forward values, so it is not modelled.
    """

    def __init__(self, grid, periodic_axes=()):
        if set(periodic_axes) - {0, 1}:
            raise ValueError("Only horizontal axes 0 (x) and 1 (y) may be periodic")
        self.grid = grid
        self.periodic_axes = frozenset(periodic_axes)
        self.m3 = {k: v[..., None] for k, v in grid.m_factors.items()}

    @staticmethod
    def _pad(f, axis, from_loc, to_loc):
        # operators.py:37-63
        if from_loc == 'm' and to_loc in ('u', 'v', 'w'):
            return np.concatenate([_take(f, slice(0, 1), axis), f,
                                   _take(f, slice(-1, None), axis)], axis=axis)
        return f

    def diff(self, f, axis, from_loc, to_loc):
        # operators.py:65-103
        delta = self.grid.delta[axis]
        d = np.diff(f, axis=axis) / delta
        if axis in self.periodic_axes and from_loc == 'm' and to_loc in ('u', 'v'):
            b = np.expand_dims((_take(f, 0, axis) - _take(f, -1, axis)) / delta, axis)
            d = np.concatenate([b, d, b], axis=axis)
        else:
            d = self._pad(d, axis, from_loc, to_loc)
        if axis in (0, 1):
            d = d * self.m3[to_loc]
        return d

    def avg(self, f, axis, from_loc, to_loc):
        # operators.py:105-132
        n = f.shape[axis]
        a = 0.5 * (_take(f, slice(0, n - 1), axis) + _take(f, slice(1, n), axis))
        if axis in self.periodic_axes and from_loc == 'm' and to_loc in ('u', 'v'):
            b = np.expand_dims(0.5 * (_take(f, 0, axis) + _take(f, -1, axis)), axis)
            return np.concatenate([b, a, b], axis=axis)
        return self._pad(a, axis, from_loc, to_loc)


# ---------------------------------------------------------------------------
# Interpolation  (operators.py:134-209; steppers.py:362-367)
# ---------------------------------------------------------------------------
def cubic_weight(p0, p1, p2, p3, t):
    """Catmull-Rom cubic in polynomial form (operators.py:134-144)."""
    return ((-0.5 * p0 + 1.5 * p1 - 1.5 * p2 + 0.5 * p3) * t ** 3
            + (p0 - 2.5 * p1 + 2.0 * p2 - 0.5 * p3) * t ** 2
            + (-0.5 * p0 + 0.5 * p2) * t + p1)


def tricubic(field, coords, use_limiter=False):
    """tensor_product_interp_3d (operators.py:146-209)."""
    nx, ny, nz = field.shape
    x = np.clip(coords[0], 0, nx - 1)
    y = np.clip(coords[1], 0, ny - 1)
    z = np.clip(coords[2], 0, nz - 1)
    x0 = np.floor(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    z0 = np.floor(z).astype(np.int64)
    tx, ty, tz = x - x0, y - y0, z - z0

    def st(i, n):
        return (np.maximum(i - 1, 0), i, np.minimum(i + 1, n - 1), np.minimum(i + 2, n - 1))

    ix, iy, iz = st(x0, nx), st(y0, ny), st(z0, nz)

    def iz_(a, b):
        return cubic_weight(*(field[a, b, iz[k]] for k in range(4)), tz)

    def iy_(a):
        return cubic_weight(*(iz_(a, iy[k]) for k in range(4)), ty)

    out = cubic_weight(*(iy_(ix[k]) for k in range(4)), tx)
    if use_limiter:
        x1 = np.minimum(x0 + 1, nx - 1)
        y1 = np.minimum(y0 + 1, ny - 1)
        z1 = np.minimum(z0 + 1, nz - 1)
        fmin = fmax = field[x0, y0, z0]
        for i, j, k in [(x1, y0, z0), (x0, y1, z0), (x1, y1, z0), (x0, y0, z1),
                        (x1, y0, z1), (x0, y1, z1), (x1, y1, z1)]:
            nb = field[i, j, k]
            fmin = np.minimum(fmin, nb)
            fmax = np.maximum(fmax, nb)
        out = np.clip(out, fmin, fmax)
    return out


def map_coordinates_linear_nearest(field, coords):
    """jax.scipy.ndimage.map_coordinates(order=1, mode='nearest').

    Per axis: i = floor(c), t = c - i, value = (1-t) f[clip(i)] + t f[clip(i+1)]
    (JAX's linear index/weight scheme with the 'nearest' index fixer; used at
    steppers.py:367 and steppers.py:579,593).
    """
    shape = field.shape
    terms = []
    for ax, c in enumerate(coords):
        lo = np.floor(c)
        wt = c - lo
        i = lo.astype(np.int64)
        n = shape[ax]
        terms.append(((np.clip(i, 0, n - 1), 1.0 - wt), (np.clip(i + 1, 0, n - 1), wt)))
    out = 0.0
    import itertools
    for combo in itertools.product(*terms):
        idx = tuple(c[0] for c in combo)
        w = 1.0
        for c in combo:
            w = w * c[1]
        out = out + w * field[idx]
    return out


# ---------------------------------------------------------------------------
# Hyperdiffusion  (suetes/regional3d/diffusion.py:80-172)
# ---------------------------------------------------------------------------
class HyperFilter:
    def __init__(self, grid, op, nu_h, nu_v):
        self.grid, self.op, self.nu_h, self.nu_v = grid, op, nu_h, nu_v

    def _lap_h(self, f, bg):
        # diffusion.py:93-129
        op, g = self.op, self.grid
        dfx = op.diff(f, 0, 'm', 'u')
        dfz_m = op.avg(op.diff(f, 2, 'm', 'w') * (g.dz / bg['dz_w_full']), 2, 'w', 'm')
        dfz_u = op.avg(dfz_m, 0, 'm', 'u')
        zxu = op.diff(g.Z_m, 0, 'm', 'u')
        Gx = dfx - zxu * dfz_u
        dfy = op.diff(f, 1, 'm', 'v')
        dfz_v = op.avg(dfz_m, 1, 'm', 'v')
        zev = op.diff(g.Z_m, 1, 'm', 'v')
        Gy = dfy - zev * dfz_v
        dGx = op.diff(Gx, 0, 'u', 'm')
        Gxm = op.avg(Gx, 0, 'u', 'm')
        dGxz = op.avg(op.diff(Gxm, 2, 'm', 'w') * (g.dz / bg['dz_w_full']), 2, 'w', 'm')
        zxm = op.avg(zxu, 0, 'u', 'm')
        dGy = op.diff(Gy, 1, 'v', 'm')
        Gym = op.avg(Gy, 1, 'v', 'm')
        dGyz = op.avg(op.diff(Gym, 2, 'm', 'w') * (g.dz / bg['dz_w_full']), 2, 'w', 'm')
        zem = op.avg(zev, 1, 'v', 'm')
        return (dGx - zxm * dGxz) + (dGy - zem * dGyz)

    def _lap_v(self, f, bg):
        # diffusion.py:131-134
        op, g = self.op, self.grid
        d = op.diff(f, 2, 'm', 'w') * (g.dz / bg['dz_w_full'])
        return op.diff(d, 2, 'w', 'm') * (g.dz / bg['dz_m_full'])

    def get_tendencies(self, sp, bg):
        # diffusion.py:136-172
        op = self.op
        out = {}
        for k in ('u', 'v', 'w', 'th_v'):
            if not (k in sp or (k == 'th_v' and 'th_v_prime_m' in sp)):
                continue
            if k == 'u':
                f = op.avg(sp['u'], 0, 'u', 'm')
            elif k == 'v':
                f = op.avg(sp['v'], 1, 'v', 'm')
            elif k == 'w':
                f = op.avg(sp['w'], 2, 'w', 'm')
            else:
                f = sp['th_v'] - bg['th_v'] if 'th_v' in sp else sp['th_v_prime_m']
            l1h = self._lap_h(f, bg)
            l1v = self._lap_v(f, bg)
            t = -self.nu_h * self._lap_h(l1h, bg) + (-self.nu_v * self._lap_v(l1v, bg))
            if k == 'u':
                out['u'] = op.avg(t, 0, 'm', 'u')
            elif k == 'v':
                out['v'] = op.avg(t, 1, 'm', 'v')
            elif k == 'w':
                out['w'] = op.avg(t, 2, 'm', 'w')
            else:
                out['th_v'] = t
        return out


# ---------------------------------------------------------------------------
# Euler3D  (suetes/regional3d/euler.py:13-439)
# ---------------------------------------------------------------------------
def exact_mean(a):
    """Exactly rounded mean of all elements (math.fsum)."""
    import math
    a = np.asarray(a, dtype=np.float64).ravel()
    return math.fsum(a.tolist()) / a.size


def exact_mean_levels(a):
    """Exactly rounded horizontal mean per z level of an (x, y, z) array."""
    return np.array([exact_mean(a[:, :, k]) for k in range(a.shape[2])])


class Euler:
    def __init__(self, grid, op, c, dt, initial_era5_state=None, N_bv=0.01,
                 damp_height=20000.0, max_damp=0.5, nu_div_factor=0.8,
                 nu_h_factor=0.1, physics_suite=None, interior_mask=None):
        # euler.py:30-124
        self.grid, self.op, self.c, self.dt = grid, op, c, dt
        self.nu_div = nu_div_factor * (grid.dx ** 2) / (4.0 * dt)
        self.nu_h = nu_h_factor * (grid.dx ** 4) / (64.0 * dt)
        self.physics_suite = physics_suite
        self.interior_mask = interior_mask
        Z_m = grid.Z_m
        if initial_era5_state is not None:
            # horizontal means (jnp.mean) taken exactly rounded (math.fsum) so
            # the background does not depend on a summation order
            z1 = exact_mean_levels(Z_m)
            th1 = exact_mean_levels(np.asarray(initial_era5_state['th_v']))
            self.theta_bg = np.interp(Z_m, z1, th1)
            pi_top = exact_mean(np.asarray(initial_era5_state['pi'])[:, :, -1])
        else:
            self.theta_0 = 300.0
            if N_bv > 0.0:
                self.theta_bg = self.theta_0 * np.exp((N_bv ** 2 / c['g']) * Z_m)
            else:
                self.theta_bg = self.theta_0 * np.ones_like(Z_m)
            zmax = np.max(Z_m)
            if N_bv > 0.0:
                pi_top = 1.0 + (c['g'] ** 2 / (c['cp'] * self.theta_0 * N_bv ** 2)) * (
                    np.exp(-N_bv ** 2 * zmax / c['g']) - 1.0)
            else:
                pi_top = 1.0 - (c['g'] / (c['cp'] * self.theta_0)) * zmax
        dzz = Z_m[:, :, 1:] - Z_m[:, :, :-1]
        thw = 0.5 * (self.theta_bg[:, :, 1:] + self.theta_bg[:, :, :-1])
        dpi = -(c['g'] * dzz) / (c['cp'] * thw)
        csr = np.cumsum(dpi[..., ::-1], axis=-1)[..., ::-1]
        top = np.full((grid.nx, grid.ny, 1), pi_top)
        self.pi_bg = np.concatenate([top - csr, top], axis=2)
        with np.errstate(divide='ignore', invalid='ignore'):
            self.tau_damp = np.where(
                grid.Z_w > damp_height,
                max_damp * 0.5 * (1.0 + np.tanh(np.pi * (grid.Z_w - damp_height)
                                                / (grid.Lz - damp_height) - np.pi / 2)),
                0.0)
        self.diffusion = HyperFilter(grid, op, self.nu_h, 0.0)

    def bg_reference(self):
        """Background dict built at steppers.py:769-774."""
        c = self.c
        rho = c['p0'] / (c['Rd'] * self.theta_bg) * (self.pi_bg ** (c['cvd'] / c['Rd']))
        return {'rho': rho, 'pi': self.pi_bg, 'th_v': self.theta_bg}

    def precompute_bg(self, b):
        # euler.py:127-160
        op, g = self.op, self.grid
        th, rho, pi = b['th_v'], b['rho'], b['pi']
        return {
            'th_v_u': op.avg(th, 0, 'm', 'u'), 'th_v_v': op.avg(th, 1, 'm', 'v'),
            'th_v_w': op.avg(th, 2, 'm', 'w'),
            'rho_u': op.avg(rho, 0, 'm', 'u'), 'rho_v': op.avg(rho, 1, 'm', 'v'),
            'rho_w': op.avg(rho, 2, 'm', 'w'),
            'dz_m_full': g.dz_m_full, 'dz_w_full': g.dz_w_full,
            'dz_u': op.avg(g.dz_m_full, 0, 'm', 'u'), 'dz_v': op.avg(g.dz_m_full, 1, 'm', 'v'),
            'C_pi': (self.c['Rd'] / self.c['cvd']) * (pi / (rho * th)),
            'pi_bg': pi, 'th_v': th,
        }

    def nonlinear_pi_tendency(self, sp, thp, bg):
        # euler.py:162-186
        op, g = self.op, self.grid
        tu = op.avg(thp, 0, 'm', 'u')
        tv = op.avg(thp, 1, 'm', 'v')
        tw = op.avg(thp, 2, 'm', 'w')
        mu, mv, mm = g.m_factors['u'][..., None], g.m_factors['v'][..., None], g.m_factors['m'][..., None]
        fx = sp['u'] * bg['rho_u'] * tu * bg['dz_u'] / mu
        fy = sp['v'] * bg['rho_v'] * tv * bg['dz_v'] / mv
        fz = sp['eta_dot'] * bg['dz_w_full'] * bg['rho_w'] * tw
        fz[:, :, 0] = 0.0
        fz[:, :, -1] = 0.0
        dxx = op.diff(fx, 0, 'u', 'm') / bg['dz_m_full']
        dyy = op.diff(fy, 1, 'v', 'm') / bg['dz_m_full']
        dzz = op.diff(fz, 2, 'w', 'm') * g.dz / bg['dz_m_full']
        return -bg['C_pi'] * (mm * (dxx + dyy) + dzz)

    def metric_curvature_tendencies(self, u, v):
        # euler.py:188-206
        op, g = self.op, self.grid
        um = op.avg(u, 0, 'u', 'm')
        vm = op.avg(v, 1, 'v', 'm')
        vau = op.avg(op.avg(v, 1, 'v', 'm'), 0, 'm', 'u')
        uav = op.avg(op.avg(u, 0, 'u', 'm'), 1, 'm', 'v')
        mm = vm * g.dm_dx_m[..., None] - um * g.dm_dy_m[..., None]
        return op.avg(mm, 0, 'm', 'u') * vau, -op.avg(mm, 1, 'm', 'v') * uav

    def get_tendencies(self, sp, bg, is_explicit=False, ml_params=None):
        # euler.py:208-372
        op, g, c = self.op, self.grid, self.c
        u, v, w, pi, eta = sp['u'], sp['v'], sp['w'], sp['pi'], sp['eta_dot']
        if is_explicit:
            thu = bg['th_v_u'] + sp.get('th_v_prime_u', 0.0)
            thv = bg['th_v_v'] + sp.get('th_v_prime_v', 0.0)
            thw = bg['th_v_w'] + sp.get('th_v_prime_w', 0.0)
        else:
            thu, thv, thw = bg['th_v_u'], bg['th_v_v'], bg['th_v_w']
        gx = op.diff(pi, 0, 'm', 'u')
        gy = op.diff(pi, 1, 'm', 'v')
        gzw = op.diff(pi, 2, 'm', 'w') * (g.dz / bg['dz_w_full'])
        zxu = op.diff(g.Z_m, 0, 'm', 'u')
        zev = op.diff(g.Z_m, 1, 'm', 'v')
        gzm = op.avg(gzw, 2, 'w', 'm')
        gzu = op.avg(gzm, 0, 'm', 'u')
        gzv = op.avg(gzm, 1, 'm', 'v')
        tu = -c['cp'] * thu * (gx - zxu * gzu)
        tv = -c['cp'] * thv * (gy - zev * gzv)
        if is_explicit:
            buoy = c['g'] * (sp.get('th_v_prime_w', 0.0) / bg['th_v_w'])
            tw = -c['cp'] * thw * gzw + buoy
        else:
            tw = -c['cp'] * thw * gzw
        vau = op.avg(op.avg(v, 1, 'v', 'm'), 0, 'm', 'u')
        uav = op.avg(op.avg(u, 0, 'u', 'm'), 1, 'm', 'v')
        tu = tu + g.f_u[..., None] * vau
        tv = tv - g.f_v[..., None] * uav
        if is_explicit:
            mu_, mv_ = self.metric_curvature_tendencies(u, v)
            tu = tu + mu_
            tv = tv + mv_
        mu, mv, mm = g.m_factors['u'][..., None], g.m_factors['v'][..., None], g.m_factors['m'][..., None]
        fx = (u * bg['rho_u'] * bg['th_v_u'] * bg['dz_u']) / mu
        fy = (v * bg['rho_v'] * bg['th_v_v'] * bg['dz_v']) / mv
        dxx = op.diff(fx, 0, 'u', 'm') / bg['dz_m_full']
        dyy = op.diff(fy, 1, 'v', 'm') / bg['dz_m_full']
        fz = eta * bg['dz_w_full'] * bg['rho_w'] * bg['th_v_w']
        fz = fz.copy()
        fz[:, :, 0] = 0.0
        fz[:, :, -1] = 0.0
        dzz = (op.diff(fz, 2, 'w', 'm') * g.dz) / bg['dz_m_full']
        tpi = -bg['C_pi'] * (mm * (dxx + dyy) + dzz)
        dthw = op.diff(bg['th_v'], 2, 'm', 'w') * (g.dz / bg['dz_w_full'])
        dthm = op.avg(dthw, 2, 'w', 'm')
        wm = op.avg(w, 2, 'w', 'm')
        tth = -wm * dthm
        if is_explicit:
            pdu = pdv = pdw = pdth = 0.0
            if self.nu_h > 0.0 or self.nu_div > 0.0:
                d = self.diffusion.get_tendencies(sp, bg)
                tu = tu + d.get('u', 0.0)
                tv = tv + d.get('v', 0.0)
                tw = tw + d.get('w', 0.0)
                tth = tth + d.get('th_v', 0.0)
                pdu = pdu + d.get('u', 0.0)
                pdv = pdv + d.get('v', 0.0)
                pdw = pdw + d.get('w', 0.0)
                pdth = pdth + d.get('th_v', 0.0)
            if self.physics_suite is not None:
                pt = self.physics_suite.get_explicit_tendencies(
                    sp, bg, interior_mask=self.interior_mask, ml_params=ml_params)
                if 'u' in pt:
                    tu = tu + pt['u']; pdu = pdu + pt['u']
                if 'v' in pt:
                    tv = tv + pt['v']; pdv = pdv + pt['v']
                if 'w' in pt:
                    tw = tw + pt['w']; pdw = pdw + pt['w']
                if 'th_v' in pt:
                    tth = tth + pt['th_v']; pdth = pdth + pt['th_v']
            return {'u': tu, 'v': tv, 'w': tw, 'pi': tpi, 'th_v': tth,
                    'phys_diff_u': pdu, 'phys_diff_v': pdv, 'phys_diff_w': pdw,
                    'phys_diff_th_v': pdth}
        return {'u': tu, 'v': tv, 'w': tw, 'pi': tpi, 'th_v': tth}

    def linear_operator(self, sp, bg, dt, alpha=0.55):
        # euler.py:374-439
        op, g = self.op, self.grid
        t = self.get_tendencies(sp, bg, is_explicit=False)
        Lu = sp['u'] - alpha * dt * t['u']
        Lv = sp['v'] - alpha * dt * t['v']
        Lw = (1.0 + dt * self.tau_damp) * sp['w'] - alpha * dt * t['w']
        Lpi = sp['pi'] - alpha * dt * t['pi']
        uw = op.avg(op.avg(sp['u'], 0, 'u', 'm'), 2, 'm', 'w')
        vw = op.avg(op.avg(sp['v'], 1, 'v', 'm'), 2, 'm', 'w')
        if 0 in op.periodic_axes:
            Lu[-1] = Lu[0]
        else:
            Lu[0] = sp['u'][0]
            Lu[-1] = sp['u'][-1]
        if 1 in op.periodic_axes:
            Lv[:, -1] = Lv[:, 0]
        else:
            Lv[:, 0] = sp['v'][:, 0]
            Lv[:, -1] = sp['v'][:, -1]
        mw = g.m_factors['w'][..., None]
        Lw[:, :, 0] = sp['w'][:, :, 0]
        Lw[:, :, -1] = sp['w'][:, :, -1]
        Le = alpha * (bg['dz_w_full'] * sp['eta_dot'] + mw * (uw * g.z_xi_w + vw * g.z_eta_w) - sp['w'])
        Le[:, :, 0] = sp['eta_dot'][:, :, 0]
        Le[:, :, -1] = sp['eta_dot'][:, :, -1]
        return {'u': Lu, 'v': Lv, 'w': Lw, 'pi': Lpi, 'eta_dot': Le}


# ---------------------------------------------------------------------------
# Column tridiagonal preconditioner  (steppers.py:125-253)
# ---------------------------------------------------------------------------
def thomas(lower, main, upper, rhs):
    """Tridiagonal solve along the last axis (LAPACK gtsv without pivoting;
    the matrix is diagonally dominant, steppers.py:183-185)."""
    n = rhs.shape[-1]
    cp = np.empty_like(rhs)
    dp = np.empty_like(rhs)
    cp[..., 0] = upper[..., 0] / main[..., 0]
    dp[..., 0] = rhs[..., 0] / main[..., 0]
    for k in range(1, n):
        den = main[..., k] - lower[..., k] * cp[..., k - 1]
        cp[..., k] = upper[..., k] / den
        dp[..., k] = (rhs[..., k] - lower[..., k] * dp[..., k - 1]) / den
    x = np.empty_like(rhs)
    x[..., n - 1] = dp[..., n - 1]
    for k in range(n - 2, -1, -1):
        x[..., k] = dp[..., k] - cp[..., k] * x[..., k + 1]
    return x


class Preconditioner:
    def __init__(self, physics, dt, alpha=0.55, pi_scale=1000.0):
        self.physics, self.dt, self.alpha, self.pi_scale = physics, dt, alpha, pi_scale

    def precompute_banded(self, bg):
        # steppers.py:153-185
        dt, a, cp = self.dt, self.alpha, self.physics.c['cp']
        self.th_v_w, self.dz_w_full, self.rho_w = bg['th_v_w'], bg['dz_w_full'], bg['rho_w']
        self.tau_damp = self.physics.tau_damp
        self.K_w = (a * dt * cp * self.th_v_w) / (self.dz_w_full * (1.0 + dt * self.tau_damp))
        P = self.K_w * self.rho_w * self.th_v_w
        P[:, :, 0] = 0.0
        P[:, :, -1] = 0.0
        self.L_pi = (a * dt * bg['C_pi']) / bg['dz_m_full']
        self.lower = -self.L_pi * P[:, :, :-1]
        self.upper = -self.L_pi * P[:, :, 1:]
        self.main = 1.0 - self.lower - self.upper

    def __call__(self, r):
        # steppers.py:187-253
        ph, op, g = self.physics, self.physics.op, self.physics.grid
        rpi = r['pi'] / self.pi_scale
        rw = r['w']
        uw = op.avg(op.avg(r['u'], 0, 'u', 'm'), 2, 'm', 'w')
        vw = op.avg(op.avg(r['v'], 1, 'v', 'm'), 2, 'm', 'w')
        mw = g.m_factors['w'][..., None]
        kin = mw * (uw * g.z_xi_w + vw * g.z_eta_w)
        wck = rw / (1.0 + self.dt * self.tau_damp) + r['eta_dot'] / self.alpha - kin
        wt = self.rho_w * self.th_v_w * wck
        wt[:, :, 0] = 0.0
        wt[:, :, -1] = 0.0
        div = op.diff(wt, 2, 'w', 'm') * g.dz
        rh = rpi - self.L_pi * div
        pp = thomas(self.lower, self.main, self.upper, rh)
        gp = op.diff(pp, 2, 'm', 'w') * (g.dz / self.dz_w_full)
        a, dt, cp = self.alpha, self.dt, ph.c['cp']
        pw = (rw - a * dt * cp * self.th_v_w * gp) / (1.0 + dt * self.tau_damp)
        pw[:, :, 0] = rw[:, :, 0]
        pw[:, :, -1] = rw[:, :, -1]
        pe = r['eta_dot'] / self.alpha + pw - kin
        pe[:, :, 0] = r['eta_dot'][:, :, 0] * self.dz_w_full[:, :, 0]
        pe[:, :, -1] = r['eta_dot'][:, :, -1] * self.dz_w_full[:, :, -1]
        return {'u': r['u'], 'v': r['v'], 'w': pw, 'pi': pp * self.pi_scale, 'eta_dot': pe}


# ---------------------------------------------------------------------------
# GMRES: restatement of jax.scipy.sparse.linalg.gmres, solve_method='batched'
# (third-party; called at steppers.py:465-481).  JAX is not installed here;
# the algorithm below follows its published source: restarted, left
# preconditioned, every cycle runs all `restart` Arnoldi steps unless a
# breakdown occurs, classical Gram-Schmidt with re-orthogonalisation,
# least squares by normal equations, stop when ||M(b - A x)|| <= ptol.
# ---------------------------------------------------------------------------
KEYS = ('u', 'v', 'w', 'pi', 'eta_dot')


class Packer:
    def __init__(self, template):
        self.shapes = [template[k].shape for k in KEYS]
        self.sizes = [int(np.prod(s)) for s in self.shapes]

    def pack(self, d):
        return np.concatenate([np.asarray(d[k], dtype=np.float64).ravel() for k in KEYS])

    def unpack(self, x):
        out, o = {}, 0
        for k, s, n in zip(KEYS, self.shapes, self.sizes):
            out[k] = x[o:o + n].reshape(s).copy()
            o += n
        return out


def _safe_normalize(x, thresh=None):
    n = np.sqrt(np.dot(x, x))
    if thresh is None:
        thresh = EPS64
    if n > thresh:
        return x / n, n
    return np.zeros_like(x), 0.0


def _cgs(Q, x, xnorm, max_iterations=2, mode='jax'):
    """Iterated classical Gram-Schmidt.

    mode 'jax': JAX's _iterative_classical_gram_schmidt loop condition
    (a second pass runs while ||r|| < ||q||/sqrt(2)); mode 'cgs2': always two
    passes.  Both give the same Krylov basis up to rounding.
    """
    r = np.zeros(Q.shape[0])
    q = x
    qns = xnorm / np.sqrt(2.0)
    k = 0
    while True:
        if mode == 'jax':
            rn = np.sqrt(np.dot(r, r))
            rn = rn if rn > EPS64 else 0.0
            if not (k < max_iterations and rn < qns):
                break
        elif k >= max_iterations:
            break
        h = Q @ q
        q = q - Q.T @ h
        r = r + h
        if k < max_iterations - 1:
            _, qn = _safe_normalize(q)
            qns = qn / np.sqrt(2.0)
        k += 1
    return q, r


def gmres(A, b, x0=None, tol=1e-5, atol=0.0, restart=20, maxiter=None, M=None,
          gs_mode='jax', stats=None):
    """Flat-vector restatement of jax.scipy.sparse.linalg.gmres (batched)."""
    if x0 is None:
        x0 = np.zeros_like(b)
    if M is None:
        M = lambda v: v  # noqa: E731
    size = b.size
    if maxiter is None:
        maxiter = 10 * size
    restart = min(restart, size)
    b_norm = np.sqrt(np.dot(b, b))
    atol = max(tol * b_norm, atol)
    Mb_norm = np.sqrt(np.dot(M(b), M(b)))
    with np.errstate(divide='ignore', invalid='ignore'):
        ptol = Mb_norm * np.minimum(1.0, atol / b_norm)
    residual = M(b - A(x0))
    unit, rnorm = _safe_normalize(residual)
    x = x0.copy()
    k = 0
    cycles = 0
    while k < maxiter and rnorm > ptol:
        V = np.zeros((restart + 1, size))
        V[0] = unit
        H = np.eye(restart, restart + 1)
        j = 0
        breakdown = False
        while j < restart and not breakdown:
            v = M(A(V[j]))
            _, vn0 = _safe_normalize(v)
            v, h = _cgs(V, v, vn0, 2, gs_mode)
            uv, vn1 = _safe_normalize(v, thresh=EPS64 * vn0)
            V[j + 1] = uv
            h = h.copy()
            h[j + 1] = vn1
            H[j, :] = h
            breakdown = vn1 == 0.0
            j += 1
        beta = np.zeros(restart + 1)
        beta[0] = rnorm
        a = H.T
        y = np.linalg.solve(a.T @ a, a.T @ beta)
        x = x + V[:-1].T @ y
        residual = M(b - A(x))
        unit, rnorm = _safe_normalize(residual)
        k += 1
        cycles += 1
    if stats is not None:
        stats['cycles'] = stats.get('cycles', 0) + cycles
        stats['residual'] = rnorm
        stats['ptol'] = ptol
    return x


class ImplicitSolver:
    """SemiImplicitSolver3D (steppers.py:377-527)."""

    def __init__(self, physics, dt, alpha=0.55, solver_tol=1e-4, solver_maxiter=20,
                 solver_restart=20, gs_mode='jax'):
        self.physics, self.dt, self.alpha = physics, dt, alpha
        self.pi_scale = 1000.0
        self.solver_tol, self.solver_maxiter, self.solver_restart = solver_tol, solver_maxiter, solver_restart
        self.gs_mode = gs_mode
        self.stats = {}

    def operators(self, bg):
        pre = Preconditioner(self.physics, self.dt, alpha=self.alpha, pi_scale=self.pi_scale)
        pre.precompute_banded(bg)

        def A_fn(s):
            sp = {'u': s['u'], 'v': s['v'], 'w': s['w'], 'pi': s['pi'] / self.pi_scale,
                  'eta_dot': s['eta_dot'] / bg['dz_w_full']}
            L = self.physics.linear_operator(sp, bg, self.dt, alpha=self.alpha)
            return {'u': L['u'], 'v': L['v'], 'w': L['w'], 'pi': L['pi'] * self.pi_scale,
                    'eta_dot': L['eta_dot']}
        return A_fn, pre

    def solve(self, rhs_prime, bg, x0=None):
        # steppers.py:401-496
        rs = {k: (rhs_prime[k] * self.pi_scale if k == 'pi' else rhs_prime[k]) for k in rhs_prime}
        if x0 is not None:
            x0s = {'u': x0['u'], 'v': x0['v'], 'w': x0['w'], 'pi': x0['pi'] * self.pi_scale,
                   'eta_dot': x0['eta_dot'] * bg['dz_w_full']}
        else:
            x0s = rs
        A_fn, pre = self.operators(bg)
        P = Packer(rs)
        xs = gmres(lambda v: P.pack(A_fn(P.unpack(v))), P.pack(rs), x0=P.pack(x0s),
                   tol=self.solver_tol, maxiter=self.solver_maxiter, restart=self.solver_restart,
                   M=lambda v: P.pack(pre(P.unpack(v))), gs_mode=self.gs_mode, stats=self.stats)
        x = P.unpack(xs)
        return {'u': x['u'], 'v': x['v'], 'w': x['w'], 'pi': x['pi'] / self.pi_scale,
                'eta_dot': x['eta_dot'] / bg['dz_w_full']}

    def relative_residual(self, sol, rhs_prime, bg):
        # steppers.py:498-527
        Ax = self.physics.linear_operator(sol, bg, self.dt, alpha=self.alpha)
        num = den = 0.0
        for k in KEYS:
            s = self.pi_scale if k == 'pi' else 1.0
            r = (Ax[k] - rhs_prime[k]) * s
            num += float(np.sum(r * r))
            den += float(np.sum((rhs_prime[k] * s) ** 2))
        return np.sqrt(num / max(den, np.finfo(np.float64).tiny))


# ---------------------------------------------------------------------------
# Semi-Lagrangian trajectories  (steppers.py:256-374)
# ---------------------------------------------------------------------------
class SLAdvector:
    def __init__(self, grid, physics, dt):
        self.grid, self.physics, self.dt, self.op = grid, physics, dt, physics.op

    def index_velocities(self, u, v, eta, loc='m'):
        # steppers.py:276-316
        op, g = self.op, self.grid
        if loc == 'm':
            ul, vl, wl = op.avg(u, 0, 'u', 'm'), op.avg(v, 1, 'v', 'm'), op.avg(eta, 2, 'w', 'm')
        elif loc == 'u':
            ul = u
            vl = op.avg(op.avg(v, 1, 'v', 'm'), 0, 'm', 'u')
            wl = op.avg(op.avg(eta, 2, 'w', 'm'), 0, 'm', 'u')
        elif loc == 'v':
            ul = op.avg(op.avg(u, 0, 'u', 'm'), 1, 'm', 'v')
            vl = v
            wl = op.avg(op.avg(eta, 2, 'w', 'm'), 1, 'm', 'v')
        else:
            ul = op.avg(op.avg(u, 0, 'u', 'm'), 2, 'm', 'w')
            vl = op.avg(op.avg(v, 1, 'v', 'm'), 2, 'm', 'w')
            wl = eta
        mf = g.m_factors[loc][..., None]
        return (ul * mf) / g.dx, (vl * mf) / g.dy, wl

    def departure(self, state, loc='m', iterations=2):
        # steppers.py:318-360
        ui, vi, wi = self.index_velocities(state['u'], state['v'], state['eta_dot'], loc)
        g = self.grid
        ix = np.arange(g.nx + (1 if loc == 'u' else 0), dtype=np.float64)
        iy = np.arange(g.ny + (1 if loc == 'v' else 0), dtype=np.float64)
        iz = np.arange(g.nz + (1 if loc == 'w' else 0), dtype=np.float64)
        X, Y, Z = np.meshgrid(ix, iy, iz, indexing='ij')
        ax, ay, az = self.dt * ui, self.dt * vi, self.dt * wi
        for _ in range(iterations):
            mid = [X - 0.5 * ax, Y - 0.5 * ay, Z - 0.5 * az]
            ax = self.dt * map_coordinates_linear_nearest(ui, mid)
            ay = self.dt * map_coordinates_linear_nearest(vi, mid)
            az = self.dt * map_coordinates_linear_nearest(wi, mid)
        return np.stack([X - ax, Y - ay, Z - az], axis=0)


# ---------------------------------------------------------------------------
# Flux-form semi-Lagrangian transport  (steppers.py:530-647)
# ---------------------------------------------------------------------------
class FFSL:
    def __init__(self, grid, dt, periodic_axes=()):
        self.grid, self.dt, self.periodic_axes = grid, dt, frozenset(periodic_axes)

    @staticmethod
    def advect_lines(mass, cfl, periodic):
        """advect_1d (steppers.py:552-594) vectorised over leading axes;
        the advected axis is the last one."""
        N = mass.shape[-1]
        if periodic:
            ext = np.concatenate([mass, mass, mass], axis=-1)
            M = np.concatenate([np.zeros(mass.shape[:-1] + (1,)), np.cumsum(ext, axis=-1)], axis=-1)
            idx = np.arange(N + 1, dtype=np.float64) + N - cfl
        else:
            M = np.concatenate([np.zeros(mass.shape[:-1] + (1,)), np.cumsum(mass, axis=-1)], axis=-1)
            idx = np.clip(np.arange(N + 1, dtype=np.float64) - cfl, 0.0, float(N))
        n = M.shape[-1]
        lo = np.floor(idx)
        t = idx - lo
        i = lo.astype(np.int64)
        a = np.take_along_axis(M, np.clip(i, 0, n - 1), axis=-1)
        b = np.take_along_axis(M, np.clip(i + 1, 0, n - 1), axis=-1)
        Md = (1.0 - t) * a + t * b
        return Md[..., 1:] - Md[..., :-1]

    def advect_3d_split(self, field, state, bg):
        # steppers.py:596-647
        g = self.grid
        cx = (state['u'] * g.m_factors['u'][..., None] * self.dt) / g.dx
        cy = (state['v'] * g.m_factors['v'][..., None] * self.dt) / g.dy
        cz = state['eta_dot'] * self.dt
        vol = (g.dx * g.dy / g.m_factors['m'][..., None] ** 2) * bg['dz_m_full']
        m = field * vol
        m = np.moveaxis(self.advect_lines(np.moveaxis(m, 0, -1), np.moveaxis(cx, 0, -1),
                                          0 in self.periodic_axes), -1, 0)
        m = np.moveaxis(self.advect_lines(np.moveaxis(m, 1, -1), np.moveaxis(cy, 1, -1),
                                          1 in self.periodic_axes), -1, 1)
        m = self.advect_lines(m, cz, False)
        return m / vol


# ---------------------------------------------------------------------------
# The SISL step  (steppers.py:650-1073)
# ---------------------------------------------------------------------------
class Stepper:
    def __init__(self, physics, dt, alpha=0.55, use_limiter=False, use_checkpointing=False,
                 solver_tol=1e-4, solver_maxiter=20, solver_restart=20, gs_mode='jax'):
        self.physics, self.dt, self.alpha, self.use_limiter = physics, dt, alpha, use_limiter
        self.advector = SLAdvector(physics.grid, physics, dt)
        self.ffsl = FFSL(physics.grid, dt, periodic_axes=physics.op.periodic_axes)
        self.solver = ImplicitSolver(physics, dt, alpha=alpha, solver_tol=solver_tol,
                                     solver_maxiter=solver_maxiter, solver_restart=solver_restart,
                                     gs_mode=gs_mode)
        self.tracer_keys = physics.physics_suite.tracer_keys if physics.physics_suite is not None else []

    def step(self, state, t, forcing, bc_fn, ml_params=None, record=None):
        ph, op, g, c = self.physics, self.physics.op, self.physics.grid, self.physics.c
        a, dt = self.alpha, self.dt
        if 'eta_dot' not in state:
            state['eta_dot'] = np.zeros_like(state['w'])
        first = float(state.get('is_first_step', 1.0)) == 1.0
        if 'u_prev' in state:
            def ex(k):
                return state[k] if first else 1.5 * state[k] - 0.5 * state[k + '_prev']
            traj = {'u': ex('u'), 'v': ex('v'), 'w': ex('w'), 'eta_dot': ex('eta_dot')}
        else:
            traj = state
        co = {loc: self.advector.departure(traj, loc) for loc in ('u', 'v', 'w', 'm')}
        if record is not None:
            record['coords'] = co
        bg = ph.precompute_bg(ph.bg_reference())
        thp_n = state['th_v'] - ph.theta_bg
        sp_n = {'u': state['u'], 'v': state['v'], 'w': state['w'], 'th_v': state['th_v'],
                'pi': state['pi'] - ph.pi_bg, 'eta_dot': state['eta_dot'],
                'th_v_prime_u': op.avg(thp_n, 0, 'm', 'u'),
                'th_v_prime_v': op.avg(thp_n, 1, 'm', 'v'),
                'th_v_prime_w': op.avg(thp_n, 2, 'm', 'w')}
        for k in ('theta_surf', 'target_th_v'):
            if k in state:
                sp_n[k] = state[k]
        tn = ph.get_tendencies(sp_n, bg, True, ml_params)
        if record is not None:
            record['tends'] = tn
        u_in = state['u'] + dt * ((1.0 - a) * tn['u'] + a * tn.get('phys_diff_u', 0.0))
        v_in = state['v'] + dt * ((1.0 - a) * tn['v'] + a * tn.get('phys_diff_v', 0.0))
        w_in = state['w'] + dt * ((1.0 - a) * tn['w'] + a * tn.get('phys_diff_w', 0.0))
        nl_n = ph.nonlinear_pi_tendency(sp_n, thp_n, bg)
        pi_in = sp_n['pi'] + (1.0 - a) * dt * (tn['pi'] + nl_n)
        th_in = thp_n + (1.0 - a) * dt * tn['th_v']
        tracers_in = {}
        for key in self.tracer_keys:
            if key in state:
                te = tn.get(key, 0.0)
                if 'u_prev' in state and not first:
                    te = 1.5 * tn.get(key, 0.0) - 0.5 * state.get(f'tend_{key}_prev', 0.0)
                tracers_in[key] = state[key] + dt * te
        uw = op.avg(op.avg(state['u'], 0, 'u', 'm'), 2, 'm', 'w')
        vw = op.avg(op.avg(state['v'], 1, 'v', 'm'), 2, 'm', 'w')
        res_n = bg['dz_w_full'] * state['eta_dot'] + uw * g.z_xi_w + vw * g.z_eta_w - state['w']
        lim = self.use_limiter
        R_eta = -(1.0 - a) * tricubic(res_n, co['w'], lim)
        rhs_u = tricubic(u_in, co['u'], lim)
        rhs_v = tricubic(v_in, co['v'], lim)
        if 0 in op.periodic_axes:
            rhs_u[-1] = rhs_u[0]
        else:
            rhs_u[0] = state['u'][0]
            rhs_u[-1] = state['u'][-1]
        if 1 in op.periodic_axes:
            rhs_v[:, -1] = rhs_v[:, 0]
        else:
            rhs_v[:, 0] = state['v'][:, 0]
            rhs_v[:, -1] = state['v'][:, -1]
        rhs_w = tricubic(w_in, co['w'], lim)
        rhs_pi = tricubic(pi_in, co['m'], lim)
        th_adv = tricubic(th_in, co['m'], lim)
        th_next = th_adv + a * dt * tn['th_v'] + ph.theta_bg
        rho_next = self.ffsl.advect_3d_split(state['rho'], state, bg)
        tracers_next = {}
        for key in self.tracer_keys:
            if key in state:
                rq = self.ffsl.advect_3d_split(state['rho'] * tracers_in[key], state, bg)
                tracers_next[key] = rq / (rho_next + c.get('eps', 1e-15))
        R_eta_adv = R_eta.copy()
        R_eta_adv[:, :, 0] = 0.0
        R_eta_adv[:, :, -1] = 0.0
        dthw = op.diff(bg['th_v'], 2, 'm', 'w') * (g.dz / bg['dz_w_full'])
        dthm = op.avg(dthw, 2, 'w', 'm')
        w_m = op.avg(state['w'], 2, 'w', 'm')
        mw = g.m_factors['w'][..., None]
        w_m_next = w_m
        guess = sp_n
        ug, vg = rhs_u, rhs_v
        zxu = op.diff(g.Z_m, 0, 'm', 'u')
        zev = op.diff(g.Z_m, 1, 'm', 'v')
        cad = c['cp'] * a * dt
        for _ in range(2):
            tth1 = tn['th_v'] + (w_m - w_m_next) * dthm
            th_next = th_adv + a * dt * tth1 + ph.theta_bg
            thp1 = th_next - ph.theta_bg
            thu1 = op.avg(thp1, 0, 'm', 'u')
            thv1 = op.avg(thp1, 1, 'm', 'v')
            thw1 = op.avg(thp1, 2, 'm', 'w')
            buoy = a * dt * (c['g'] * (thw1 / bg['th_v_w']))
            pg = guess['pi']
            gx = op.diff(pg, 0, 'm', 'u')
            gy = op.diff(pg, 1, 'm', 'v')
            gzw = op.diff(pg, 2, 'm', 'w') * (g.dz / bg['dz_w_full'])
            gzm = op.avg(gzw, 2, 'w', 'm')
            gzu = op.avg(gzm, 0, 'm', 'u')
            gzv = op.avg(gzm, 1, 'm', 'v')
            nlu = -cad * thu1 * (gx - zxu * gzu)
            nlv = -cad * thv1 * (gy - zev * gzv)
            nlw = -cad * thw1 * gzw
            nld = ph.nonlinear_pi_tendency(guess, thp1, bg)
            mun, mvn = ph.metric_curvature_tendencies(ug, vg)
            rhs_u_k = rhs_u + nlu + a * dt * mun
            rhs_v_k = rhs_v + nlv + a * dt * mvn
            rhs_w_k = rhs_w + buoy + nlw
            rhs_pi_k = rhs_pi + a * dt * nld
            uwg = op.avg(op.avg(ug, 0, 'u', 'm'), 2, 'm', 'w')
            vwg = op.avg(op.avg(vg, 1, 'v', 'm'), 2, 'm', 'w')
            kb = mw[:, :, 0] * (uwg[:, :, 0] * g.z_xi_w[:, :, 0] + vwg[:, :, 0] * g.z_eta_w[:, :, 0])
            rhs_w_k = rhs_w_k.copy()
            rhs_w_k[:, :, 0] = kb
            rhs_w_k[:, :, -1] = 0.0
            rk = {'u': rhs_u_k, 'v': rhs_v_k, 'w': rhs_w_k, 'pi': rhs_pi_k, 'eta_dot': R_eta_adv}
            nxt = self.solver.solve(rk, bg, x0=guess)
            if record is not None:
                record.setdefault('solves', []).append((rk, guess, nxt))
            guess = nxt
            w_m_next = op.avg(nxt['w'], 2, 'w', 'm')
            ug, vg = nxt['u'], nxt['v']
        out = {'u': guess['u'].copy(), 'v': guess['v'].copy(), 'w': guess['w'],
               'pi': guess['pi'] + ph.pi_bg, 'rho': rho_next, 'th_v': th_next,
               'eta_dot': guess['eta_dot']}
        for key in self.tracer_keys:
            if key in tracers_next:
                out[key] = tracers_next[key]
        div = op.diff(out['u'], 0, 'u', 'm') + op.diff(out['v'], 1, 'v', 'm')
        out['u'] = out['u'] + ph.nu_div * dt * op.diff(div, 0, 'm', 'u')
        out['v'] = out['v'] + ph.nu_div * dt * op.diff(div, 1, 'm', 'v')
        if ph.physics_suite is not None:
            out = ph.physics_suite.apply_state_updates(out)
        out = bc_fn(out, forcing)
        tmass = {k: out['rho'] * out[k] for k in self.tracer_keys if k in out}
        out['rho'] = c['p0'] / (c['Rd'] * out['th_v']) * (out['pi'] ** (c['cvd'] / c['Rd']))
        for k, rq in tmass.items():
            out[k] = rq / (out['rho'] + c.get('eps', 1e-15))
        out['u_prev'], out['v_prev'], out['w_prev'] = state['u'], state['v'], state['w']
        out['eta_dot_prev'] = state['eta_dot']
        out['tend_th_v_prev'] = tn['th_v']
        for key in self.tracer_keys:
            out[f'tend_{key}_prev'] = tn.get(key, np.zeros_like(out[key]))
        out['is_first_step'] = 0.0
        return out


def build_dynamical_core(core_type, grid, operators, constants, initial_state,
                         physics_suite=None, interior_mask=None, gs_mode='jax', **kwargs):
    """build_dynamical_core('sisl', ...) (steppers.py:12-88); 'split-explicit'
    goes to oracle.split_explicit (steppers.py:95-119)."""
    if core_type.lower() == 'split-explicit':
        from .split_explicit import build_split_explicit
        return build_split_explicit(grid, operators, constants, initial_state, physics_suite=physics_suite,
                                    interior_mask=interior_mask, **kwargs)
    if core_type.lower() != 'sisl':
        raise ValueError(f"Unknown core_type: '{core_type}'. Use 'sisl' or 'split-explicit'.")
    p = {'dt': 120.0, 'damp_height': 9000.0, 'max_damp': 3.0, 'nu_div_factor': 0.1,
         'nu_h_factor': 0.1, 'N_bv': 0.01, 'alpha': 0.55, 'use_limiter': False,
         'use_checkpointing': False, 'solver_tol': 1e-12, 'solver_maxiter': 100,
         'solver_restart': 100, 'background_state': None}
    p.update(kwargs)
    ref = p['background_state'] if p['background_state'] is not None else initial_state
    physics = Euler(grid, operators, constants, dt=p['dt'], initial_era5_state=ref,
                    damp_height=p['damp_height'], max_damp=p['max_damp'],
                    nu_div_factor=p['nu_div_factor'], nu_h_factor=p['nu_h_factor'],
                    N_bv=p['N_bv'], physics_suite=physics_suite, interior_mask=interior_mask)
    if 'u_prev' not in initial_state:
        initial_state['u_prev'] = initial_state['u']
        initial_state['v_prev'] = initial_state['v']
        initial_state['w_prev'] = initial_state['w']
        initial_state['eta_dot_prev'] = initial_state.get('eta_dot', np.zeros_like(initial_state['w']))
        initial_state['tend_th_v_prev'] = np.zeros_like(initial_state['th_v'])
        initial_state['is_first_step'] = 1.0
        for key in (physics_suite.tracer_keys if physics_suite is not None else []):
            if key in initial_state:
                initial_state[f'tend_{key}_prev'] = np.zeros_like(initial_state[key])
    st = Stepper(physics, p['dt'], alpha=p['alpha'], use_limiter=p['use_limiter'],
                 use_checkpointing=p['use_checkpointing'], solver_tol=p['solver_tol'],
                 solver_maxiter=p['solver_maxiter'], solver_restart=p['solver_restart'],
                 gs_mode=gs_mode)
    return st, p['dt']


# ---------------------------------------------------------------------------
# Lateral boundaries  (suetes/regional3d/boundaries.py:13-121)
# ---------------------------------------------------------------------------
class DaviesSponge:
    def __init__(self, grid, op, sponge_depth=30, dt=30.0, tau_bndy_factor=10.0, outflow_factor=0.01):
        self.grid, self.op, self.depth, self.outflow_factor = grid, op, sponge_depth, outflow_factor
        self.max_c = dt / (tau_bndy_factor * dt)
        self.masks = {loc: davies_masks(grid.nx, grid.ny, sponge_depth, loc) for loc in ('u', 'v', 'm')}

    def blend(self, st, ext):
        # boundaries.py:83-121
        op = self.op
        out = {}
        ub, vb = st['u'], st['v']
        for k in st.keys():
            if k in ('u', 'v', 'th_v', 'q', 'pi') and k in ext:
                m = self.masks.get(k, self.masks['m'])
                if k == 'u':
                    ue, ve = ub, op.avg(op.avg(vb, 1, 'v', 'm'), 0, 'm', 'u')
                elif k == 'v':
                    ue, ve = op.avg(op.avg(ub, 0, 'u', 'm'), 1, 'm', 'v'), vb
                else:
                    ue, ve = op.avg(ub, 0, 'u', 'm'), op.avg(vb, 1, 'v', 'm')
                ww = m['west'] * np.where(ue > 0, 1.0, self.outflow_factor)
                we = m['east'] * np.where(ue < 0, 1.0, self.outflow_factor)
                ws = m['south'] * np.where(ve > 0, 1.0, self.outflow_factor)
                wn = m['north'] * np.where(ve < 0, 1.0, self.outflow_factor)
                ct = np.maximum(np.maximum(ww, we), np.maximum(ws, wn))
                out[k] = (1.0 - ct) * st[k] + ct * ext[k]
            else:
                out[k] = st[k]
        return out


class XRelaxation:
    """Symmetric cos^2 relaxation over ``cells`` columns at both x ends toward
    the exterior state, for u, v, th_v, pi, rho; the boundary function of the
    reference's mountain-wave benchmark
    (benchmarks/physical/schaer_mountain_3d/run.py:103-126)."""

    def __init__(self, nx, cells=8):
        idx = np.arange(nx)
        dist = np.minimum(idx, nx - 1 - idx)
        self.mask = np.where(dist < cells, np.cos(0.5 * np.pi * dist / cells) ** 2, 0.0)[:, None, None]
        self.mask_u = np.pad(self.mask, ((0, 1), (0, 0), (0, 0)), mode='edge')

    def blend(self, st, ext):
        out = {}
        for k, v in st.items():
            if k in ('u', 'v', 'th_v', 'pi', 'rho') and k in ext:
                m = self.mask_u if k == 'u' else self.mask
                out[k] = (1.0 - m) * v + m * ext[k]
            else:
                out[k] = v
        return out


class BenchmarkSponge:
    """Symmetric cos^2 relaxation along x and/or y (boundaries.py:123-189):
    per-axis float32 weights cos^2(pi/2 * d / depth) for distance d from the
    nearer edge below ``depth`` (0 beyond), combined as 1 - (1-wx)(1-wy);
    u / v masks repeat the last mass column / row."""

    def __init__(self, nx, ny=None, sponge_depth=10, axes=('x',), blend_vars=None):
        self.nx, self.ny, self.depth, self.axes = nx, ny, sponge_depth, axes
        if blend_vars is None:
            blend_vars = (['u', 'v', 'th_v', 'pi', 'rho', 'q_tr'] if ny is None
                          else ['u', 'v', 'w', 'th_v', 'pi', 'rho', 'q_tr'])
        self.blend_vars = blend_vars

        def axis_weight(n):
            # boundaries.py:145-158 (float32 index, distance and cosine)
            idx = np.arange(n, dtype=np.float32)
            dist = np.minimum(idx, np.float32(n - 1) - idx)
            arg = np.float32(0.5 * np.pi) * dist / np.float32(sponge_depth)
            return np.where(dist < sponge_depth, np.cos(arg) ** 2, np.float32(0.0)).astype(np.float32)

        wx = axis_weight(nx) if 'x' in axes else np.zeros(nx, np.float32)
        if ny is not None:
            wy = axis_weight(ny) if 'y' in axes else np.zeros(ny, np.float32)
            mm = (1.0 - (1.0 - wx[:, None]) * (1.0 - wy[None, :])).astype(np.float32)[:, :, None]
            self.mask_m = mm
            self.mask_u = np.pad(mm, ((0, 1), (0, 0), (0, 0)), mode='edge')
            self.mask_v = np.pad(mm, ((0, 0), (0, 1), (0, 0)), mode='edge')
        else:
            self.mask_m = wx[:, None, None]
            self.mask_u = np.pad(self.mask_m, ((0, 1), (0, 0), (0, 0)), mode='edge')
            self.mask_v = self.mask_m
        self.mask_w = self.mask_m

    def blend(self, st, ext):
        # boundaries.py:174-189
        out = {}
        for k, v in st.items():
            if k in self.blend_vars and k in ext:
                m = {'u': self.mask_u, 'v': self.mask_v, 'w': self.mask_w}.get(k, self.mask_m)
                out[k] = (1.0 - m) * v + m * ext[k]
            else:
                out[k] = v
        return out


def davies_masks(nx, ny, depth, loc):
    """boundaries.py:36-61 (float32 weights, as in the reference)."""
    nxp = nx + 1 if loc == 'u' else nx
    nyp = ny + 1 if loc == 'v' else ny
    ix = np.arange(nxp, dtype=np.float32)
    iy = np.arange(nyp, dtype=np.float32)
    ox = np.float32(0.0 if loc == 'u' else 0.5)
    oy = np.float32(0.0 if loc == 'v' else 0.5)
    d = np.float32(depth)
    dw = np.clip((ix + ox) / d, 0.0, 1.0).astype(np.float32)
    de = np.clip((np.float32(nx) - (ix + ox)) / d, 0.0, 1.0).astype(np.float32)
    ds = np.clip((iy + oy) / d, 0.0, 1.0).astype(np.float32)
    dn = np.clip((np.float32(ny) - (iy + oy)) / d, 0.0, 1.0).astype(np.float32)
    hp = np.float32(0.5 * np.pi)

    def wgt(z):
        return (np.cos(hp * z) ** 2).astype(np.float32)

    return {'west': wgt(dw)[:, None, None], 'east': wgt(de)[:, None, None],
            'south': wgt(ds)[None, :, None], 'north': wgt(dn)[None, :, None]}
