"""Euler equations on the terrain-following C grid (device tensors).

Mirrors ``Euler3D`` (suetes/regional3d/euler.py:13-439) and ``HyperFilter``
(suetes/regional3d/diffusion.py:80-172) on internal (Z, Y, X) tensors.  The
explicit tendencies run once per step as torch ops (differentiable); the
implicit operator used inside the Krylov solve is the HIP kernel, with
coefficients precomputed here (``operator_coefficients``).
"""
from __future__ import annotations

import math

import torch


def interp1d(x, xp, fp):
    """``interp`` with NumPy's arithmetic: for xp[j] <= x < xp[j+1],
    slope*(x - xp[j]) + fp[j]; exact node hits return fp[j]; clamped to the
    end values outside [xp[0], xp[-1]]."""
    n = xp.shape[0]
    j = torch.searchsorted(xp.contiguous(), x.contiguous().reshape(-1), right=True).reshape(x.shape) - 1
    jc = torch.clamp(j, 0, n - 2)
    x0, x1, f0, f1 = xp[jc], xp[jc + 1], fp[jc], fp[jc + 1]
    f = ((f1 - f0) / (x1 - x0)) * (x - x0) + f0
    f = torch.where(x == x0, f0, f)
    f = torch.where(j < 0, fp[0].expand_as(f), f)
    return torch.where(j >= n - 1, fp[-1].expand_as(f), f)


def exact_mean(t, dims):
    """Mean over ``dims`` whose value is exactly rounded (math.fsum on the
    host, independent of summation order) and whose gradient is the mean's."""
    import math
    m = t.mean(dim=dims)
    a = t.detach().double().cpu().numpy()
    axes = tuple(d % t.dim() for d in dims)
    keep = [d for d in range(t.dim()) if d not in axes]
    a = a.transpose(keep + list(axes)).reshape([t.shape[d] for d in keep] + [-1])
    flat = a.reshape(-1, a.shape[-1])
    vals = [math.fsum(r.tolist()) / r.size for r in flat]
    ex = torch.tensor(vals, dtype=t.dtype).reshape(m.shape).to(t.device)
    return m + (ex - m).detach()


class HyperFilter:
    """Terrain-corrected 4th-order hyperdiffusion (diffusion.py:80-172)."""

    def __init__(self, grid, operators, nu_h=1e6, nu_v=1e6):
        self.grid, self.op, self.nu_h, self.nu_v = grid, operators, nu_h, nu_v
        self._metric_cache = None

    def _metrics(self, bg):
        """Terrain metrics of the Laplacian (the same in every call of a step):
        dz / dz_w, z_x at u, z_y at v and their mass-point averages.  Kept
        across calls while neither the heights nor the background spacing
        carry a gradient (terrain controls recompute them every call)."""
        op, g = self.op, self.grid
        zm, dzw = g.i_Z_m, bg["dz_w_full"]
        c = self._metric_cache
        if c is not None and c[0] is zm and c[1] is dzw:
            return c[2]
        rz = g.dz / dzw
        zxu = op.diff(zm, 0, "m", "u")
        zev = op.diff(zm, 1, "m", "v")
        vals = (rz, zxu, zev, op.avg(zxu, 0, "u", "m"), op.avg(zev, 1, "v", "m"))
        if not (zm.requires_grad or dzw.requires_grad):
            self._metric_cache = (zm, dzw, vals)
        return vals

    def _cartesian_horizontal_laplacian(self, f, bg):
        op = self.op
        rz, zxu, zev, zxm, zem = self._metrics(bg)
        dfx = op.diff(f, 0, "m", "u")
        dfz_m = op.avg(op.diff(f, 2, "m", "w") * rz, 2, "w", "m")
        dfz_u = op.avg(dfz_m, 0, "m", "u")
        Gx = dfx - zxu * dfz_u
        dfy = op.diff(f, 1, "m", "v")
        dfz_v = op.avg(dfz_m, 1, "m", "v")
        Gy = dfy - zev * dfz_v
        dGx = op.diff(Gx, 0, "u", "m")
        Gxm = op.avg(Gx, 0, "u", "m")
        dGxz = op.avg(op.diff(Gxm, 2, "m", "w") * rz, 2, "w", "m")
        dGy = op.diff(Gy, 1, "v", "m")
        Gym = op.avg(Gy, 1, "v", "m")
        dGyz = op.avg(op.diff(Gym, 2, "m", "w") * rz, 2, "w", "m")
        return (dGx - zxm * dGxz) + (dGy - zem * dGyz)

    def _vertical_laplacian(self, f, bg):
        op, g = self.op, self.grid
        d = op.diff(f, 2, "m", "w") * (g.dz / bg["dz_w_full"])
        return op.diff(d, 2, "w", "m") * (g.dz / bg["dz_m_full"])

    def get_tendencies(self, sp, bg):
        op = self.op
        out = {}
        for k in ("u", "v", "w", "th_v"):
            if not (k in sp or (k == "th_v" and "th_v_prime_m" in sp)):
                continue
            if k == "u":
                f = op.avg(sp["u"], 0, "u", "m")
            elif k == "v":
                f = op.avg(sp["v"], 1, "v", "m")
            elif k == "w":
                f = op.avg(sp["w"], 2, "w", "m")
            else:
                f = sp["th_v"] - bg["th_v"] if "th_v" in sp else sp["th_v_prime_m"]
            l1h = self._cartesian_horizontal_laplacian(f, bg)
            t = -self.nu_h * self._cartesian_horizontal_laplacian(l1h, bg)
            if self.nu_v != 0.0:
                l1v = self._vertical_laplacian(f, bg)
                t = t - self.nu_v * self._vertical_laplacian(l1v, bg)
            if k == "u":
                out["u"] = op.avg(t, 0, "m", "u")
            elif k == "v":
                out["v"] = op.avg(t, 1, "m", "v")
            elif k == "w":
                out["w"] = op.avg(t, 2, "m", "w")
            else:
                out["th_v"] = t
        return out


class Euler3D:
    """Background state, tendencies and linear operator (euler.py:13-439)."""

    def __init__(self, grid, operators, constants, dt, initial_era5_state=None, N_bv=0.01,
                 damp_height=20000.0, max_damp=0.5, nu_div_factor=0.8, nu_h_factor=0.1,
                 physics_suite=None, interior_mask=None):
        self.grid, self.op, self.c, self.dt = grid, operators, constants, dt
        self.nu_div = nu_div_factor * (grid.dx ** 2) / (4.0 * dt)
        self.nu_h = nu_h_factor * (grid.dx ** 4) / (64.0 * dt)
        self.physics_suite = physics_suite
        self.interior_mask = interior_mask
        c = constants
        Zm = grid.i_Z_m
        if initial_era5_state is not None:
            th = initial_era5_state["i_th_v"] if "i_th_v" in initial_era5_state else \
                initial_era5_state["th_v"].permute(2, 1, 0)
            pi = initial_era5_state["i_pi"] if "i_pi" in initial_era5_state else \
                initial_era5_state["pi"].permute(2, 1, 0)
            th = th.to(Zm)
            pi = pi.to(Zm)
            z1 = exact_mean(Zm, (1, 2))
            th1 = exact_mean(th, (1, 2))
            self.i_theta_bg = interp1d(Zm, z1, th1)
            pi_top = exact_mean(pi[-1], (0, 1))
        else:
            self.theta_0 = 300.0
            if N_bv > 0.0:
                self.i_theta_bg = self.theta_0 * torch.exp((N_bv ** 2 / c["g"]) * Zm)
            else:
                self.i_theta_bg = self.theta_0 * torch.ones_like(Zm)
            zmax = Zm.max()
            if N_bv > 0.0:
                pi_top = 1.0 + (c["g"] ** 2 / (c["cp"] * self.theta_0 * N_bv ** 2)) * (
                    torch.exp(-N_bv ** 2 * zmax / c["g"]) - 1.0)
            else:
                pi_top = 1.0 - (c["g"] / (c["cp"] * self.theta_0)) * zmax
        dzz = Zm[1:] - Zm[:-1]
        thw = 0.5 * (self.i_theta_bg[1:] + self.i_theta_bg[:-1])
        dpi = -(c["g"] * dzz) / (c["cp"] * thw)
        # hydrostatic integration from the lid, summed sequentially on the host
        dev = dpi.device
        csr = torch.flip(torch.cumsum(torch.flip(dpi.cpu(), (0,)), dim=0), (0,)).to(dev)
        top = pi_top * torch.ones_like(Zm[:1])
        self.i_pi_bg = torch.cat([top - csr, top], dim=0)
        Zw = grid.i_Z_w
        if grid.Lz != damp_height:
            prof = max_damp * 0.5 * (1.0 + torch.tanh(math.pi * (Zw - damp_height)
                                                       / (grid.Lz - damp_height) - math.pi / 2))
        else:
            prof = torch.zeros_like(Zw)
        self.i_tau_damp = torch.where(Zw > damp_height, prof, torch.zeros_like(Zw))
        self.diffusion = HyperFilter(grid, operators, nu_h=self.nu_h, nu_v=0.0)

    # reference-shaped views
    @property
    def theta_bg(self):
        return self.i_theta_bg.permute(2, 1, 0)

    @property
    def pi_bg(self):
        return self.i_pi_bg.permute(2, 1, 0)

    @property
    def tau_damp(self):
        return self.i_tau_damp.permute(2, 1, 0)

    def bg_reference(self):
        """steppers.py:769-774."""
        c = self.c
        rho = c["p0"] / (c["Rd"] * self.i_theta_bg) * (self.i_pi_bg ** (c["cvd"] / c["Rd"]))
        return {"rho": rho, "pi": self.i_pi_bg, "th_v": self.i_theta_bg}

    def precompute_bg(self, b):
        # euler.py:127-160
        op, g = self.op, self.grid
        th, rho, pi = b["th_v"], b["rho"], b["pi"]
        return {
            "th_v_u": op.avg(th, 0, "m", "u"), "th_v_v": op.avg(th, 1, "m", "v"),
            "th_v_w": op.avg(th, 2, "m", "w"),
            "rho_u": op.avg(rho, 0, "m", "u"), "rho_v": op.avg(rho, 1, "m", "v"),
            "rho_w": op.avg(rho, 2, "m", "w"),
            "dz_m_full": g.i_dz_m_full, "dz_w_full": g.i_dz_w_full,
            "dz_u": op.avg(g.i_dz_m_full, 0, "m", "u"), "dz_v": op.avg(g.i_dz_m_full, 1, "m", "v"),
            "C_pi": (self.c["Rd"] / self.c["cvd"]) * (pi / (rho * th)),
            "pi_bg": pi, "th_v": th,
        }

    def nonlinear_pi_tendency(self, sp, thp, bg):
        # euler.py:162-186
        op, g = self.op, self.grid
        tu = op.avg(thp, 0, "m", "u")
        tv = op.avg(thp, 1, "m", "v")
        tw = op.avg(thp, 2, "m", "w")
        mu, mv, mm = op.m3("u"), op.m3("v"), op.m3("m")
        fx = sp["u"] * bg["rho_u"] * tu * bg["dz_u"] / mu
        fy = sp["v"] * bg["rho_v"] * tv * bg["dz_v"] / mv
        fz = sp["eta_dot"] * bg["dz_w_full"] * bg["rho_w"] * tw
        fz = torch.cat([torch.zeros_like(fz[:1]), fz[1:-1], torch.zeros_like(fz[:1])], dim=0)
        dxx = op.diff(fx, 0, "u", "m") / bg["dz_m_full"]
        dyy = op.diff(fy, 1, "v", "m") / bg["dz_m_full"]
        dzz = op.diff(fz, 2, "w", "m") * g.dz / bg["dz_m_full"]
        return -bg["C_pi"] * (mm * (dxx + dyy) + dzz)

    def metric_curvature_tendencies(self, u, v):
        # euler.py:188-206
        op, g = self.op, self.grid
        um = op.avg(u, 0, "u", "m")
        vm = op.avg(v, 1, "v", "m")
        vau = op.avg(vm, 0, "m", "u")
        uav = op.avg(um, 1, "m", "v")
        mm = vm * g.i_dm_dx_m[None] - um * g.i_dm_dy_m[None]
        return op.avg(mm, 0, "m", "u") * vau, -op.avg(mm, 1, "m", "v") * uav

    def get_tendencies(self, sp, bg, is_explicit=False, ml_params=None):
        # euler.py:208-372
        op, g, c = self.op, self.grid, self.c
        u, v, w, pi, eta = sp["u"], sp["v"], sp["w"], sp["pi"], sp["eta_dot"]
        if is_explicit:
            thu = bg["th_v_u"] + sp.get("th_v_prime_u", 0.0)
            thv = bg["th_v_v"] + sp.get("th_v_prime_v", 0.0)
            thw = bg["th_v_w"] + sp.get("th_v_prime_w", 0.0)
        else:
            thu, thv, thw = bg["th_v_u"], bg["th_v_v"], bg["th_v_w"]
        rz = g.dz / bg["dz_w_full"]
        gx = op.diff(pi, 0, "m", "u")
        gy = op.diff(pi, 1, "m", "v")
        gzw = op.diff(pi, 2, "m", "w") * rz
        zxu = op.diff(g.i_Z_m, 0, "m", "u")
        zev = op.diff(g.i_Z_m, 1, "m", "v")
        gzm = op.avg(gzw, 2, "w", "m")
        gzu = op.avg(gzm, 0, "m", "u")
        gzv = op.avg(gzm, 1, "m", "v")
        tu = -c["cp"] * thu * (gx - zxu * gzu)
        tv = -c["cp"] * thv * (gy - zev * gzv)
        if is_explicit:
            buoy = c["g"] * (sp.get("th_v_prime_w", 0.0) / bg["th_v_w"])
            tw = -c["cp"] * thw * gzw + buoy
        else:
            tw = -c["cp"] * thw * gzw
        vau = op.avg(op.avg(v, 1, "v", "m"), 0, "m", "u")
        uav = op.avg(op.avg(u, 0, "u", "m"), 1, "m", "v")
        tu = tu + g.i_f_u[None] * vau
        tv = tv - g.i_f_v[None] * uav
        if is_explicit:
            mu_, mv_ = self.metric_curvature_tendencies(u, v)
            tu = tu + mu_
            tv = tv + mv_
        mu, mv, mm = op.m3("u"), op.m3("v"), op.m3("m")
        fx = (u * bg["rho_u"] * bg["th_v_u"] * bg["dz_u"]) / mu
        fy = (v * bg["rho_v"] * bg["th_v_v"] * bg["dz_v"]) / mv
        dxx = op.diff(fx, 0, "u", "m") / bg["dz_m_full"]
        dyy = op.diff(fy, 1, "v", "m") / bg["dz_m_full"]
        fz = eta * bg["dz_w_full"] * bg["rho_w"] * bg["th_v_w"]
        fz = torch.cat([torch.zeros_like(fz[:1]), fz[1:-1], torch.zeros_like(fz[:1])], dim=0)
        dzz = (op.diff(fz, 2, "w", "m") * g.dz) / bg["dz_m_full"]
        tpi = -bg["C_pi"] * (mm * (dxx + dyy) + dzz)
        dthw = op.diff(bg["th_v"], 2, "m", "w") * rz
        dthm = op.avg(dthw, 2, "w", "m")
        wm = op.avg(w, 2, "w", "m")
        tth = -wm * dthm
        if is_explicit:
            pdu = pdv = pdw = pdth = 0.0
            if self.nu_h > 0.0 or self.nu_div > 0.0:
                d = self.diffusion.get_tendencies(sp, bg)
                tu, tv, tw, tth = tu + d["u"], tv + d["v"], tw + d["w"], tth + d["th_v"]
                pdu, pdv, pdw, pdth = d["u"], d["v"], d["w"], d["th_v"]
            if self.physics_suite is not None:
                pt = self.physics_suite.get_explicit_tendencies(
                    sp, bg, interior_mask=self.interior_mask, ml_params=ml_params)
                if "u" in pt:
                    tu, pdu = tu + pt["u"], pdu + pt["u"]
                if "v" in pt:
                    tv, pdv = tv + pt["v"], pdv + pt["v"]
                if "w" in pt:
                    tw, pdw = tw + pt["w"], pdw + pt["w"]
                if "th_v" in pt:
                    tth, pdth = tth + pt["th_v"], pdth + pt["th_v"]
            return {"u": tu, "v": tv, "w": tw, "pi": tpi, "th_v": tth, "phys_diff_u": pdu,
                    "phys_diff_v": pdv, "phys_diff_w": pdw, "phys_diff_th_v": pdth}
        return {"u": tu, "v": tv, "w": tw, "pi": tpi, "th_v": tth}

    def linear_operator(self, sp, bg, dt, alpha=0.55):
        """Torch restatement of euler.py:374-439 (test reference and the
        metric-cotangent path of the implicit solve)."""
        op, g = self.op, self.grid
        t = self.get_tendencies(sp, bg, is_explicit=False)
        Lu = sp["u"] - alpha * dt * t["u"]
        Lv = sp["v"] - alpha * dt * t["v"]
        Lw = (1.0 + dt * self.i_tau_damp) * sp["w"] - alpha * dt * t["w"]
        Lpi = sp["pi"] - alpha * dt * t["pi"]
        uw = op.avg(op.avg(sp["u"], 0, "u", "m"), 2, "m", "w")
        vw = op.avg(op.avg(sp["v"], 1, "v", "m"), 2, "m", "w")
        if 0 in op.periodic_axes:
            Lu = torch.cat([Lu[:, :, :-1], Lu[:, :, :1]], dim=2)
        else:
            Lu = torch.cat([sp["u"][:, :, :1], Lu[:, :, 1:-1], sp["u"][:, :, -1:]], dim=2)
        if 1 in op.periodic_axes:
            Lv = torch.cat([Lv[:, :-1], Lv[:, :1]], dim=1)
        else:
            Lv = torch.cat([sp["v"][:, :1], Lv[:, 1:-1], sp["v"][:, -1:]], dim=1)
        Lw = torch.cat([sp["w"][:1], Lw[1:-1], sp["w"][-1:]], dim=0)
        mw = op.m3("w")
        Le = alpha * (bg["dz_w_full"] * sp["eta_dot"] + mw * (uw * g.i_z_xi_w + vw * g.i_z_eta_w) - sp["w"])
        Le = torch.cat([sp["eta_dot"][:1], Le[1:-1], sp["eta_dot"][-1:]], dim=0)
        return {"u": Lu, "v": Lv, "w": Lw, "pi": Lpi, "eta_dot": Le}

    def operator_coefficients(self, bg, dt, alpha):
        """Coefficient arrays of the HIP operator A and preconditioner M
        (layout and meaning: include/suetes_sisl.h, SislOperator)."""
        op, g, c = self.op, self.grid, self.c
        cp = c["cp"]
        tau = self.i_tau_damp
        dzw = bg["dz_w_full"]
        zxu = op.diff(g.i_Z_m, 0, "m", "u")
        zev = op.diff(g.i_Z_m, 1, "m", "v")
        mu, mv, mm = g.i_m_factors["u"], g.i_m_factors["v"], g.i_m_factors["m"]
        mw = g.i_m_factors["w"]
        rt = bg["rho_w"] * bg["th_v_w"]
        co = {
            "cua": alpha * dt * cp * bg["th_v_u"] / 1000.0,
            "zxu": zxu,
            "cx": bg["rho_u"] * bg["th_v_u"] * bg["dz_u"] / mu[None],
            "cva": alpha * dt * cp * bg["th_v_v"] / 1000.0,
            "zev": zev,
            "cy": bg["rho_v"] * bg["th_v_v"] * bg["dz_v"] / mv[None],
            "dw": 1.0 + dt * tau,
            "cw": alpha * dt * cp * bg["th_v_w"] / dzw / 1000.0,
            "cz": rt,
            "kx": mw[None] * g.i_z_xi_w,
            "ky": mw[None] * g.i_z_eta_w,
            "rdzw": 1.0 / dzw,
            "cpm": 1000.0 * alpha * dt * bg["C_pi"] / bg["dz_m_full"],
            "mudx": mu / g.dx,
            "mvdy": mv / g.dy,
            "afu": alpha * dt * g.i_f_u,
            "afv": alpha * dt * g.i_f_v,
            "mmx": mm * mm / g.dx,
            "mmy": mm * mm / g.dy,
            "idw": 1.0 / (1.0 + dt * tau),
            "gw": alpha * dt * cp * bg["th_v_w"] / ((1.0 + dt * tau) * dzw),
            "dzw0": dzw[0],
            "dzwn": dzw[-1],
        }
        # column matrix (steppers.py:153-185) and its Thomas factorisation
        Kw = (alpha * dt * cp * bg["th_v_w"]) / (dzw * (1.0 + dt * tau))
        P = Kw * rt
        P = torch.cat([torch.zeros_like(P[:1]), P[1:-1], torch.zeros_like(P[:1])], dim=0)
        lpi = (alpha * dt * bg["C_pi"]) / bg["dz_m_full"]
        lower = -lpi * P[:-1]
        upper = -lpi * P[1:]
        main = 1.0 - lower - upper
        nz = lower.shape[0]
        den = torch.empty_like(main)
        cpr = torch.empty_like(main)
        den[0] = main[0]
        cpr[0] = upper[0] / den[0]
        for k in range(1, nz):
            den[k] = main[k] - lower[k] * cpr[k - 1]
            cpr[k] = upper[k] / den[k]
        co.update({"lpi": lpi, "lower": lower, "invden": 1.0 / den, "cprime": cpr})
        return co
