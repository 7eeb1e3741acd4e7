"""NumPy FP64 restatement of the reference's split-explicit RK3 core
(suetes/regional3d/steppers.py:1076-1737) -- test infrastructure only.

Arrays in the reference layout (x, y, z), z contiguous.  Follows the
reference statement by statement, including what it does to the physics
object during a step: the physics suite and ``nu_h`` are hidden while the
RK stages run (steppers.py:1186-1191), but the hyperdiffusion object keeps
its own coefficient, so with ``nu_div > 0`` the stage tendencies still
contain the hyperdiffusion that is also added from the cached start-of-step
tendencies (steppers.py:1323-1332 with euler.py:337-346).
"""
from __future__ import annotations

import numpy as np

from .sisl import FFSL, Euler, thomas


def _sl(n_axis, axis, start, stop):
    s = [slice(None)] * 3
    s[axis] = slice(start, stop)
    return tuple(s)


class SplitExplicitStepper:
    """SplitExplicitStepper3D (steppers.py:1076-1737)."""

    def __init__(self, physics, dt, ns, alpha=0.55):
        # steppers.py:1092-1116
        self.physics, self.dt, self.ns, self.alpha = physics, dt, ns, alpha
        self.ns_stage1 = max(1, round(ns / 3))
        self.ns_stage2 = max(1, round(ns / 2))
        self.ns_stage3 = ns
        self.dtau_stage1 = (dt / 3.0) / float(self.ns_stage1)
        self.dtau_stage2 = (dt / 2.0) / float(self.ns_stage2)
        self.dtau_stage3 = dt / float(self.ns_stage3)
        self.ffsl_advector = FFSL(physics.grid, dt, periodic_axes=physics.op.periodic_axes)

    # -- steppers.py:1118-1269 ----------------------------------------------------------
    def step(self, state, t, forcing, bc_fn, ml_params=None):
        ph = self.physics
        c = ph.c
        bg_ref = {'rho': c['p0'] / (c['Rd'] * ph.theta_bg) * (ph.pi_bg ** (c['cvd'] / c['Rd'])),
                  'pi': ph.pi_bg, 'th_v': ph.theta_bg}
        bg = ph.precompute_bg(bg_ref)
        state_t = state
        state_t['eta_dot'], state_t['w'] = self._diagnose_eta_dot(state_t)
        phys_tends, diff_tends = {}, {}
        thp = state_t['th_v'] - ph.theta_bg
        op = ph.op
        sp = {'u': state_t['u'], 'v': state_t['v'], 'w': state_t['w'], 'th_v': state_t['th_v'],
              'pi': state_t['pi'] - ph.pi_bg, 'eta_dot': state_t['eta_dot'],
              'th_v_prime_u': op.avg(thp, 0, 'm', 'u'), 'th_v_prime_v': op.avg(thp, 1, 'm', 'v'),
              'th_v_prime_w': op.avg(thp, 2, 'm', 'w')}
        for k in state_t:
            if k not in sp:
                sp[k] = state_t[k]
        if ph.physics_suite is not None:
            phys_tends = ph.physics_suite.get_explicit_tendencies(sp, bg, interior_mask=ph.interior_mask,
                                                                  ml_params=ml_params)
        if ph.nu_h > 0.0 or ph.nu_div > 0.0:
            diff_tends = ph.diffusion.get_tendencies(sp, bg)
        suite, nu_h = ph.physics_suite, ph.nu_h
        ph.physics_suite, ph.nu_h = None, 0.0
        try:
            s1 = self._compute_slow_tendencies(state_t, bg, phys_tends, diff_tends, ml_params)
            star = self._acoustic_loop(state_t, state_t, s1, bg, self.dtau_stage1, self.ns_stage1)
            star['eta_dot'], star['w'] = self._diagnose_eta_dot(star)
            s2 = self._compute_slow_tendencies(star, bg, phys_tends, diff_tends, ml_params)
            dstar = self._acoustic_loop(state_t, star, s2, bg, self.dtau_stage2, self.ns_stage2)
            dstar['eta_dot'], dstar['w'] = self._diagnose_eta_dot(dstar)
            s3 = self._compute_slow_tendencies(dstar, bg, phys_tends, diff_tends, ml_params)
            nxt = self._acoustic_loop(state_t, dstar, s3, bg, self.dtau_stage3, self.ns_stage3)
        finally:
            ph.physics_suite, ph.nu_h = suite, nu_h
        if suite is not None:
            for key in suite.tracer_keys:
                if key in state_t:
                    rq = self.ffsl_advector.advect_3d_split(state_t['rho'] * state_t[key], state_t, bg)
                    nxt[key] = rq / (nxt['rho'] + c.get('eps', 1e-15))
        if ph.physics_suite is not None:
            nxt = ph.physics_suite.apply_state_updates(nxt)
        nxt = bc_fn(nxt, forcing)
        nxt['eta_dot'], nxt['w'] = self._diagnose_eta_dot(nxt)
        tmass = {}
        if ph.physics_suite is not None:
            tmass = {k: nxt['rho'] * nxt[k] for k in ph.physics_suite.tracer_keys if k in nxt}
        nxt['rho'] = c['p0'] / (c['Rd'] * nxt['th_v']) * (nxt['pi'] ** (c['cvd'] / c['Rd']))
        for k, rq in tmass.items():
            nxt[k] = rq / (nxt['rho'] + c.get('eps', 1e-15))
        return nxt

    # -- steppers.py:1271-1309 -----------------------------------------------------------
    def _diagnose_eta_dot(self, state):
        ph, op, g = self.physics, self.physics.op, self.physics.grid
        uw = op.avg(op.avg(state['u'], 0, 'u', 'm'), 2, 'm', 'w')
        vw = op.avg(op.avg(state['v'], 1, 'v', 'm'), 2, 'm', 'w')
        mw = g.m_factors['w'][..., None]
        w_sfc = mw[:, :, 0] * (uw[:, :, 0] * g.z_xi_w[:, :, 0] + vw[:, :, 0] * g.z_eta_w[:, :, 0])
        w = state['w'].copy()
        w[:, :, 0] = w_sfc
        w[:, :, -1] = 0.0
        edz = w - mw * (uw * g.z_xi_w + vw * g.z_eta_w)
        eta = edz / g.dz_w_full
        eta[:, :, 0] = 0.0
        eta[:, :, -1] = 0.0
        return eta, w

    # -- steppers.py:1311-1356 -----------------------------------------------------------
    def _upwind_flux_3rd_order(self, flux, field, axis):
        ph = self.physics
        pad = [(0, 0)] * 3
        pad[axis] = (2, 2)
        q = np.pad(field, pad, mode='wrap' if axis in ph.op.periodic_axes else 'edge')
        q_im2 = q[_sl(3, axis, 0, -3)]
        q_im1 = q[_sl(3, axis, 1, -2)]
        q_i = q[_sl(3, axis, 2, -1)]
        q_ip1 = q[_sl(3, axis, 3, None)]
        q4 = (7.0 / 12.0) * (q_i + q_im1) - (1.0 / 12.0) * (q_ip1 + q_im2)
        qb = (1.0 / 12.0) * ((q_ip1 - q_im2) - 3.0 * (q_i - q_im1))
        clean = np.where(np.abs(flux) < ph.c.get('eps_l', 1e-12), 0.0, flux)
        return flux * (q4 + np.sign(clean) * qb)

    # -- steppers.py:1358-1470 -----------------------------------------------------------
    def _compute_slow_tendencies(self, cur, bg, phys_tends, diff_tends, ml_params):
        ph, op, g = self.physics, self.physics.op, self.physics.grid
        thp = cur['th_v'] - ph.theta_bg
        sp = {'u': cur['u'], 'v': cur['v'], 'w': cur['w'], 'th_v': cur['th_v'], 'pi': cur['pi'] - ph.pi_bg,
              'eta_dot': cur.get('eta_dot', np.zeros_like(cur['w'])),
              'th_v_prime_u': op.avg(thp, 0, 'm', 'u'), 'th_v_prime_v': op.avg(thp, 1, 'm', 'v'),
              'th_v_prime_w': op.avg(thp, 2, 'm', 'w')}
        for k in cur:
            if k not in sp:
                sp[k] = cur[k]
        tends = dict(ph.get_tendencies(sp, bg, is_explicit=True, ml_params=ml_params))
        tends['pi'] = tends['pi'] + ph.nonlinear_pi_tendency(sp, thp, bg)
        for k in ('u', 'v', 'w', 'th_v'):
            if phys_tends and k in phys_tends:
                tends[k] = tends[k] + phys_tends[k]
            if diff_tends and k in diff_tends:
                tends[k] = tends[k] + diff_tends[k]
        mm, mu, mv = g.m_factors['m'][..., None], g.m_factors['u'][..., None], g.m_factors['v'][..., None]
        um = op.avg(cur['u'], 0, 'u', 'm')
        vm = op.avg(cur['v'], 1, 'v', 'm')
        wm = op.avg(cur['w'], 2, 'w', 'm')
        U = sp['u'] / mu
        V = sp['v'] / mv
        W = sp['eta_dot'] * bg['dz_w_full']
        rz = g.dz / bg['dz_m_full']
        divU = op.diff(U, 0, 'u', 'm')
        divV = op.diff(V, 1, 'v', 'm')
        divW = op.diff(W, 2, 'w', 'm') * rz

        def adv(q):
            ax = op.diff(self._upwind_flux_3rd_order(U, q, 0), 0, 'u', 'm')
            ay = op.diff(self._upwind_flux_3rd_order(V, q, 1), 1, 'v', 'm')
            az = op.diff(self._upwind_flux_3rd_order(W, q, 2), 2, 'w', 'm') * rz
            return -(mm ** 2) * (ax - q * divU + ay - q * divV) - (az - q * divW)
        tends['th_v'] = tends['th_v'] + adv(thp)
        tends['u'] = tends['u'] + op.avg(adv(um), 0, 'm', 'u')
        tends['v'] = tends['v'] + op.avg(adv(vm), 1, 'm', 'v')
        tends['w'] = tends['w'] + op.avg(adv(wm), 2, 'm', 'w')
        return tends

    # -- steppers.py:1472-1651 -----------------------------------------------------------
    def _acoustic_loop(self, s_init, s_cur, slow, bg, dtau, num_steps):
        ph, op, g = self.physics, self.physics.op, self.physics.grid
        cp = ph.c['cp']
        u_pp = s_init['u'] - s_cur['u']
        v_pp = s_init['v'] - s_cur['v']
        w_pp = s_init['w'] - s_cur['w']
        pi_pp = s_init['pi'] - s_cur['pi']
        mu, mv, mm = g.m_factors['u'][..., None], g.m_factors['v'][..., None], g.m_factors['m'][..., None]
        zxu = op.diff(g.Z_m, 0, 'm', 'u')
        zev = op.diff(g.Z_m, 1, 'm', 'v')

        def hgrad(p):
            gzm = op.avg(op.diff(p, 2, 'm', 'w') * (g.dz / bg['dz_w_full']), 2, 'w', 'm')
            return (op.diff(p, 0, 'm', 'u') - zxu * op.avg(gzm, 0, 'm', 'u'),
                    op.diff(p, 1, 'm', 'v') - zev * op.avg(gzm, 1, 'm', 'v'))

        def ddamp(up, vp):
            div = op.diff(s_cur['u'] + up, 0, 'u', 'm') + op.diff(s_cur['v'] + vp, 1, 'v', 'm')
            return op.diff(div, 0, 'm', 'u'), op.diff(div, 1, 'm', 'v')

        def bounds(up, vp):
            ub = s_init['u'] - s_cur['u']
            vb = s_init['v'] - s_cur['v']
            up, vp = up.copy(), vp.copy()
            if 0 in op.periodic_axes:
                up[-1] = up[0]
            else:
                up[0], up[-1] = ub[0], ub[-1]
            if 1 in op.periodic_axes:
                vp[:, -1] = vp[:, 0]
            else:
                vp[:, 0], vp[:, -1] = vb[:, 0], vb[:, -1]
            return up, vp

        for _ in range(num_steps):
            u_old, v_old, pi_old = u_pp, v_pp, pi_pp
            gox, goy = hgrad(pi_old)
            dox, doy = ddamp(u_old, v_old)
            ru = slow['u'] - cp * bg['th_v_u'] * gox + ph.nu_div * dox
            rv = slow['v'] - cp * bg['th_v_v'] * goy + ph.nu_div * doy
            u_half = u_old + 0.5 * dtau * ru
            v_half = v_old + 0.5 * dtau * rv
            u_half, v_half = bounds(u_half, v_half)
            fx = (u_half * bg['rho_u'] * bg['th_v_u'] * bg['dz_u']) / mu
            fy = (v_half * bg['rho_v'] * bg['th_v_v'] * bg['dz_v']) / mv
            dxp = op.diff(fx, 0, 'u', 'm') / bg['dz_m_full']
            dyp = op.diff(fy, 1, 'v', 'm') / bg['dz_m_full']
            div_h = mm * (dxp + dyp)
            pi_pp, w_pp = self._solve_vertical_acoustic_column(
                pi_pp, w_pp, u_half, v_half, div_h, slow['pi'], slow['w'], bg['th_v_w'], bg['C_pi'],
                bg['dz_m_full'], bg['dz_w_full'], bg, dtau)
            gnx, gny = hgrad(pi_pp)
            dhx, dhy = ddamp(u_half, v_half)
            u_pp = u_old + dtau * (slow['u'] - 0.5 * cp * bg['th_v_u'] * (gox + gnx) + ph.nu_div * dhx)
            v_pp = v_old + dtau * (slow['v'] - 0.5 * cp * bg['th_v_v'] * (goy + gny) + ph.nu_div * dhy)
            u_pp, v_pp = bounds(u_pp, v_pp)
            wt = s_cur['w'] + w_pp
            w_pp = wt / (1.0 + dtau * ph.tau_damp) - s_cur['w']
        out = {'u': s_cur['u'] + u_pp, 'v': s_cur['v'] + v_pp, 'w': s_cur['w'] + w_pp, 'pi': s_cur['pi'] + pi_pp,
               'th_v': s_init['th_v'] + dtau * num_steps * slow['th_v'], 'rho': s_cur['rho'],
               'eta_dot': s_cur.get('eta_dot', np.zeros_like(w_pp))}
        for k in s_cur:
            if k not in out:
                out[k] = s_cur[k]
        return out

    # -- steppers.py:1653-1737 -----------------------------------------------------------
    def _solve_vertical_acoustic_column(self, pi_pp, w_pp, u_pp, v_pp, div_h, slow_f_pi, slow_f_w, th_v_w_bg,
                                        C_pi, dz_m, dz_w, bg, dtau):
        a, cp = self.alpha, self.physics.c['cp']
        op, g = self.physics.op, self.physics.grid
        uw = op.avg(op.avg(u_pp, 0, 'u', 'm'), 2, 'm', 'w')
        vw = op.avg(op.avg(v_pp, 1, 'v', 'm'), 2, 'm', 'w')
        mw = g.m_factors['w'][..., None]
        w_sfc = mw[:, :, 0] * (uw[:, :, 0] * g.z_xi_w[:, :, 0] + vw[:, :, 0] * g.z_eta_w[:, :, 0])
        gp = op.diff(pi_pp, 2, 'm', 'w') * (g.dz / dz_w)
        w_exp = w_pp + dtau * (slow_f_w - cp * th_v_w_bg * gp)
        w_avg = (1.0 - a) * w_pp + a * w_exp
        w_avg[:, :, 0] = w_sfc
        w_avg[:, :, -1] = 0.0
        edz = w_avg - mw * (uw * g.z_xi_w + vw * g.z_eta_w)
        edz[:, :, 0] = 0.0
        edz[:, :, -1] = 0.0
        fz = edz * bg['rho_w'] * bg['th_v_w']
        div_z = (op.diff(fz, 2, 'w', 'm') * g.dz) / dz_m
        rhs = dtau * (slow_f_pi - C_pi * (div_h + div_z))
        gamma = (a * dtau) ** 2 * cp * th_v_w_bg * bg['rho_w'] * bg['th_v_w'] / dz_w
        gamma[:, :, 0] = 0.0
        gamma[:, :, -1] = 0.0
        lower = -(C_pi / dz_m) * gamma[:, :, :-1]
        upper = -(C_pi / dz_m) * gamma[:, :, 1:]
        main = 1.0 - lower - upper
        dpi = thomas(lower, main, upper, rhs)
        gd = op.diff(dpi, 2, 'm', 'w') * (g.dz / dz_w)
        w_next = w_exp - a * dtau * cp * th_v_w_bg * gd
        w_next[:, :, 0] = w_sfc
        w_next[:, :, -1] = 0.0
        return pi_pp + dpi, w_next


def build_split_explicit(grid, operators, constants, initial_state, physics_suite=None, interior_mask=None,
                         **kwargs):
    """build_dynamical_core('split-explicit', ...) (steppers.py:95-119)."""
    p = {'dt': 40.0, 'ns': 3, 'damp_height': 12000.0, 'max_damp': 3.0, 'nu_div_factor': 0.03,
         'nu_h_factor': 0.03, 'N_bv': 0.01, 'alpha': 0.55, 'background_state': None}
    p.update(kwargs)
    ref = p['background_state'] if p['background_state'] is not None else initial_state
    physics = Euler(grid, operators, constants, dt=p['dt'], initial_era5_state=ref,
                    damp_height=p['damp_height'], max_damp=p['max_damp'], nu_div_factor=p['nu_div_factor'],
                    nu_h_factor=p['nu_h_factor'], N_bv=p['N_bv'], physics_suite=physics_suite,
                    interior_mask=interior_mask)
    return SplitExplicitStepper(physics, dt=p['dt'], ns=p['ns'], alpha=p['alpha']), p['dt']
