"""Synthetic workloads of the SISL step (seeded, NumPy in / NumPy out).

Each workload fixes the grid, terrain, background, initial state, stepper
options and boundary treatment of one benchmark/parity case:

* ``bubble``   -- dry warm bubble, 64^3, dx=dz=50 m, doubly periodic, at rest,
                  isentropic theta0=300 K; cos^2 bubble +2 K, r=250 m, centre
                  (1600,1600,350) m from the corner; dt=1 s, restart 20, tol 1e-10.
* ``schaer``   -- x-z mountain-wave slice run as a 3-D grid with ny=3 (the
                  reference's own idiom, benchmarks/physical/schaer_mountain_3d),
                  400x3x100, dx=250 m, top 20 km, stretched SLEVE,
                  h(x)=250 m exp(-(x/5 km)^2) cos^2(pi x/4 km), U=10 m/s,
                  N=0.01 1/s, Rayleigh sponge above 12 km, 8-column x
                  relaxation, dt=8 s.
* ``lam``      -- limited-area step, 256x256x80, dx=1 km, stretched SLEVE to
                  20 km, terrain = seeded Gaussian random field (std 400 m,
                  20 km correlation, clipped >= 0), N=0.01 background, seeded
                  +-2 m/s wind noise (10 km correlation), 10-cell Davies rim
                  (equal inflow/outflow weights), dt=10 s.
* ``adjoint``  -- 128x128x60, dx=1 km, terrain = 8 seeded Gaussian bumps
                  (sigma=5 km, <= 300 m) whose amplitudes are the control.

Grid arrays are accepted from either the torch grid (``grid.Z_m`` views) or
any object exposing NumPy arrays with the reference shapes.
"""
from __future__ import annotations

import math

import numpy as np

CONSTANTS = {"g": 9.81, "cp": 1004.0, "Rd": 287.0, "cvd": 717.0, "p0": 100000.0}


def _np(a):
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _is_torch(a):
    return hasattr(a, "detach")


def analytic_background(Z_m, N_bv, c=CONSTANTS, theta0=300.0):
    """theta and hydrostatic pi of euler.py:86-112 (NumPy, reference shapes)."""
    Z = _np(Z_m)
    if N_bv > 0.0:
        th = theta0 * np.exp((N_bv ** 2 / c["g"]) * Z)
        pi_top = 1.0 + (c["g"] ** 2 / (c["cp"] * theta0 * N_bv ** 2)) * (np.exp(-N_bv ** 2 * Z.max() / c["g"]) - 1.0)
    else:
        th = theta0 * np.ones_like(Z)
        pi_top = 1.0 - (c["g"] / (c["cp"] * theta0)) * Z.max()
    dzz = Z[:, :, 1:] - Z[:, :, :-1]
    thw = 0.5 * (th[:, :, 1:] + th[:, :, :-1])
    dpi = -(c["g"] * dzz) / (c["cp"] * thw)
    csr = np.cumsum(dpi[..., ::-1], axis=-1)[..., ::-1]
    top = np.full(Z.shape[:2] + (1,), pi_top)
    return th, np.concatenate([top - csr, top], axis=2)


def eos_rho(th, pi, c=CONSTANTS):
    return c["p0"] / (c["Rd"] * th) * pi ** (c["cvd"] / c["Rd"])


def rest_state(shape, th, pi):
    nx, ny, nz = shape
    return {"u": np.zeros((nx + 1, ny, nz)), "v": np.zeros((nx, ny + 1, nz)),
            "w": np.zeros((nx, ny, nz + 1)), "eta_dot": np.zeros((nx, ny, nz + 1)),
            "pi": pi.copy(), "th_v": th.copy(), "rho": eos_rho(th, pi)}


# ---------------------------------------------------------------------------
class GaussianRandomTerrain:
    """Seeded Gaussian random field as a continuous function of (x, y):
    white noise on a 2 km lattice convolved with a Gaussian kernel
    (kernel sigma = corr/2, so the covariance decays as exp(-r^2/corr^2)),
    then shifted/scaled to zero mean and ``std`` on the mass grid and
    clipped at zero.  Works on NumPy arrays and on torch tensors."""

    def __init__(self, Lx, Ly, seed=0, std=400.0, corr=20000.0, spacing=2000.0, x_m=None, y_m=None):
        self.sig = corr / 2.0
        m = 3.0 * self.sig
        self.xs = np.arange(-Lx / 2 - m, Lx / 2 + m + spacing, spacing)
        self.ys = np.arange(-Ly / 2 - m, Ly / 2 + m + spacing, spacing)
        rng = np.random.default_rng(seed)
        self.W = rng.standard_normal((self.xs.size, self.ys.size))
        self.mu, self.scale = 0.0, 1.0
        if x_m is not None:
            X, Y = np.meshgrid(_np(x_m), _np(y_m), indexing="ij")
            raw = self._raw(X, Y)
            self.mu = float(raw.mean())
            self.scale = std / float(raw.std())

    def _raw(self, X, Y):
        gx = np.exp(-((X.reshape(-1, 1) - self.xs[None, :]) ** 2) / (2 * self.sig ** 2))
        gy = np.exp(-((Y.reshape(-1, 1) - self.ys[None, :]) ** 2) / (2 * self.sig ** 2))
        return np.einsum("pi,ij,pj->p", gx, self.W, gy).reshape(X.shape)

    def __call__(self, x, y):
        if _is_torch(x):
            import torch
            h = self(_np(x), _np(y))
            return torch.from_numpy(h).to(device=x.device, dtype=x.dtype)
        X, Y = np.broadcast_arrays(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64))
        # evaluate on the unique (x, y) pairs only (the grid repeats them in z)
        flat = np.stack([X.reshape(-1), Y.reshape(-1)], axis=1)
        uniq, inv = np.unique(flat, axis=0, return_inverse=True)
        h = np.maximum((self._raw(uniq[:, 0], uniq[:, 1]) - self.mu) * self.scale, 0.0)
        return h[inv.reshape(-1)].reshape(X.shape)


class SchaerTerrain:
    """h(x) = h0 exp(-(x/a)^2) cos^2(pi x / lam)."""

    def __init__(self, h0=250.0, a=5000.0, lam=4000.0):
        self.h0, self.a, self.lam = h0, a, lam

    def __call__(self, x, y):
        if _is_torch(x):
            import torch
            return self.h0 * torch.exp(-(x / self.a) ** 2) * torch.cos(math.pi * x / self.lam) ** 2
        return self.h0 * np.exp(-(x / self.a) ** 2) * np.cos(np.pi * x / self.lam) ** 2


class BumpTerrain:
    """Sum of Gaussian bumps with amplitudes ``amps`` (a torch tensor keeps the
    terrain differentiable with respect to the amplitudes)."""

    def __init__(self, centres, sigma=5000.0, amps=None):
        self.centres = np.asarray(centres, dtype=np.float64)
        self.sigma = sigma
        self.amps = amps

    def __call__(self, x, y):
        if _is_torch(x):
            import torch
            h = torch.zeros_like(x)
            for b in range(self.centres.shape[0]):
                cx, cy = float(self.centres[b, 0]), float(self.centres[b, 1])
                h = h + self.amps[b] * torch.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * self.sigma ** 2))
            return h
        amps = _np(self.amps)
        h = np.zeros(np.broadcast(x, y).shape)
        for b in range(self.centres.shape[0]):
            cx, cy = self.centres[b]
            h = h + amps[b] * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * self.sigma ** 2))
        return h


# ---------------------------------------------------------------------------
WORKLOADS = {}


def workload(fn):
    WORKLOADS[fn.__name__] = fn
    return fn


@workload
def bubble(n=64, dx=50.0, seed=0):
    """Config of the warm-bubble parity case (see module docstring)."""
    return {
        "name": "bubble", "grid": dict(nx=n, ny=n, nz=n, dx=dx, dy=dx, dz=dx, lat_center=45.0,
                                       lon_center=0.0),
        "terrain": None, "transform": ("galchen", {}), "periodic_axes": (0, 1),
        "core": dict(dt=1.0, N_bv=0.0, nu_div_factor=0.05, nu_h_factor=0.05, damp_height=7500.0,
                     max_damp=0.05, alpha=0.55, solver_tol=1e-10, solver_maxiter=20, solver_restart=20),
        "N_bv": 0.0, "steps": 60, "boundary": None,
        "init": ("bubble", dict(dtheta=2.0, radius=250.0 * dx / 50.0,
                                centre=(n * dx / 2.0, n * dx / 2.0, 350.0 * dx / 50.0))),
    }


@workload
def lam(nx=256, ny=256, nz=80, dx=1000.0, top=20000.0, seed=0, steps=100, winds="smooth",
        outflow_factor=1.0):
    """Config of the 3-D limited-area case (see module docstring).

    ``winds='smooth'`` (default): +-2 m/s noise with a 10 km horizontal
    correlation; ``'white'``: independent U(-2, 2) per face.

    ``outflow_factor`` of the Davies rim: 1.0 (default) relaxes inflow and
    outflow points alike; the reference's class default 0.01 weakens the
    relaxation where the boundary wind points outwards.  That switch is a
    step function of the wind's sign, so the trajectory depends
    discontinuously on rounding-size differences wherever a rim wind
    crosses zero (at 256x256x80 an FP32-kernel run leaves the FP64 run by
    1e-4 at step 40 and 0.3 at step 100, versus 2.8e-6 without the switch),
    and with white-noise winds it also feeds a growing mode at the closed
    faces (max |u| 4 -> 65 m/s in 100 FP64 steps)."""
    init = (("smooth_winds", dict(seed=seed + 1, amp=2.0, corr=10000.0)) if winds == "smooth"
            else ("winds", dict(seed=seed + 1, amp=2.0)))
    return {
        "name": "lam", "grid": dict(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, dz=top / nz, lat_center=45.0,
                                    lon_center=0.0),
        "terrain": ("grf", dict(seed=seed, std=400.0, corr=20000.0)),
        "transform": ("stretched_sleve", dict(stretch_kappa=2.5, scale_s=4000.0, n=1.35)),
        "periodic_axes": (),
        "core": dict(dt=10.0, N_bv=0.01, alpha=0.55, solver_tol=1e-10, solver_maxiter=10,
                     solver_restart=20),
        "N_bv": 0.01, "steps": steps,
        "boundary": ("davies", dict(sponge_depth=10, dt=10.0, outflow_factor=outflow_factor)),
        "init": init,
    }


@workload
def schaer(nx=400, nz=100, dx=250.0, top=20000.0, steps=500):
    """Config of the mountain-wave slice (see module docstring)."""
    return {
        "name": "schaer", "grid": dict(nx=nx, ny=3, nz=nz, dx=dx, dy=dx, dz=top / nz, lat_center=0.0,
                                       lon_center=0.0),
        "terrain": ("schaer", dict(h0=250.0, a=5000.0, lam=4000.0)),
        "transform": ("stretched_sleve", dict(stretch_kappa=2.5, scale_s=4000.0, n=1.35)),
        "periodic_axes": (1,),
        # stepper options and boundary function of the reference's benchmark
        # (benchmarks/physical/schaer_mountain_3d/run.py:103-136): alpha 0.5,
        # no hyperdiffusion, Rayleigh damping 0.5 above 12 km, 8-column
        # symmetric x relaxation; solver tolerance 1e-10, restart 20
        "core": dict(dt=8.0, N_bv=0.01, alpha=0.5, damp_height=12000.0, max_damp=0.5,
                     nu_div_factor=0.0, nu_h_factor=0.0, solver_tol=1e-10, solver_maxiter=10,
                     solver_restart=20),
        "N_bv": 0.01, "steps": steps, "boundary": ("xrelax", dict(cells=8)),
        "init": ("uniform", dict(u=10.0)),
    }


@workload
def adjoint(nx=128, ny=128, nz=60, dx=1000.0, top=15000.0, seed=3, steps=50):
    """Config of the reverse-mode case (see module docstring)."""
    rng = np.random.default_rng(seed)
    L = nx * dx
    centres = rng.uniform(-0.3 * L, 0.1 * L, size=(8, 2))
    amps = rng.uniform(100.0, 300.0, size=8)
    return {
        "name": "adjoint", "grid": dict(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, dz=top / nz, lat_center=45.0,
                                        lon_center=0.0),
        "terrain": ("bumps", dict(centres=centres, sigma=5000.0, amps=amps)),
        "transform": ("stretched_sleve", dict(stretch_kappa=2.5, scale_s=4000.0, n=1.35)),
        "periodic_axes": (),
        # stepper options and x relaxation of the reference's optimal-topography
        # experiment (experiments/gravity_wave_optimal_topography/run.py:70-140)
        "core": dict(dt=10.0, N_bv=0.01, alpha=0.55, nu_div_factor=0.0, nu_h_factor=0.0,
                     damp_height=12000.0, max_damp=0.5, solver_tol=1e-10, solver_maxiter=10,
                     solver_restart=20),
        "N_bv": 0.01, "steps": steps, "boundary": ("xrelax", dict(cells=10)),
        "init": ("uniform", dict(u=10.0)),
        "objective_box": (slice(int(0.55 * nx), int(0.55 * nx) + 20), slice(ny // 2 - 10, ny // 2 + 10),
                          slice(0, 10)),
    }


def ensemble_members(M=32, seed=5, steps=20, nx_range=(96, 192), nz_range=(48, 80), sizes=None):
    """Configs of the batched-ensemble case: ``M`` limited-area members on
    seeded uneven grids (nx = ny in ``nx_range`` by 16, nz in ``nz_range`` by
    8, dx = 1 km, top 20 km; or the explicit (n, nz) list ``sizes``), each
    with its own terrain and wind seeds and seeded theta noise of 0.5 K."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(M if sizes is None else len(sizes)):
        if sizes is not None:
            n, nz = sizes[m]
        else:
            n = int(rng.integers(nx_range[0] // 16, nx_range[1] // 16 + 1)) * 16
            nz = int(rng.integers(nz_range[0] // 8, nz_range[1] // 8 + 1)) * 8
        cfg = lam(nx=n, ny=n, nz=nz, seed=seed * 100 + m, steps=steps)
        cfg["name"] = f"ensemble[{m}]"
        kind, kw = cfg["init"]
        cfg["init"] = (kind, dict(kw, theta_noise=0.5, noise_seed=seed * 1000 + m))
        out.append(cfg)
    return out


def make_terrain(spec, grid_kw):
    if spec is None:
        return None
    kind, kw = spec
    nx, ny, dx, dy = grid_kw["nx"], grid_kw["ny"], grid_kw["dx"], grid_kw["dy"]
    if kind == "grf":
        Lx, Ly = nx * dx, ny * dy
        x_m = (np.arange(nx) + 0.5) * dx - Lx / 2
        y_m = (np.arange(ny) + 0.5) * dy - Ly / 2
        return GaussianRandomTerrain(Lx, Ly, x_m=x_m, y_m=y_m, **kw)
    if kind == "schaer":
        return SchaerTerrain(**kw)
    if kind == "bumps":
        return BumpTerrain(**kw)
    raise ValueError(kind)


def initial_state(cfg, Z_m, x_m, y_m, z_m=None, c=CONSTANTS):
    """Reference-shaped NumPy initial state for workload ``cfg``."""
    th, pi = analytic_background(Z_m, cfg["N_bv"], c)
    g = cfg["grid"]
    shape = (g["nx"], g["ny"], g["nz"])
    st = rest_state(shape, th, pi)
    kind, kw = cfg["init"]
    if kind == "bubble":
        Z = _np(Z_m)
        xm, ym = _np(x_m), _np(y_m)
        X = (xm + g["nx"] * g["dx"] / 2.0)[:, None, None]
        Y = (ym + g["ny"] * g["dy"] / 2.0)[None, :, None]
        cx, cy, cz = kw["centre"]
        r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2)
        bub = np.where(r <= kw["radius"], kw["dtheta"] * np.cos(0.5 * np.pi * r / kw["radius"]) ** 2, 0.0)
        st["th_v"] = th + bub
        st["rho"] = eos_rho(st["th_v"], pi, c)
    elif kind == "winds":
        rng = np.random.default_rng(kw["seed"])
        st["u"] = rng.uniform(-kw["amp"], kw["amp"], size=st["u"].shape)
        st["v"] = rng.uniform(-kw["amp"], kw["amp"], size=st["v"].shape)
    elif kind == "smooth_winds":
        # horizontally correlated noise (Gaussian random field of the given
        # correlation length), the same at every level, scaled to max |.| = amp
        xm, ym = _np(x_m).reshape(-1), _np(y_m).reshape(-1)
        if xm.size != g["nx"]:
            xm = _np(x_m)[:, 0, 0]
            ym = _np(y_m)[0, :, 0]
        xu = np.concatenate([xm - g["dx"] / 2.0, xm[-1:] + g["dx"] / 2.0])
        yv = np.concatenate([ym - g["dy"] / 2.0, ym[-1:] + g["dy"] / 2.0])
        Lx, Ly = g["nx"] * g["dx"], g["ny"] * g["dy"]
        for key, xs, ys, sd in (("u", xu, ym, kw["seed"]), ("v", xm, yv, kw["seed"] + 1)):
            fld = GaussianRandomTerrain(Lx, Ly, seed=sd, corr=kw["corr"], spacing=kw["corr"] / 4.0)
            X, Y = np.meshgrid(xs, ys, indexing="ij")
            raw = fld._raw(X, Y)
            raw = raw - raw.mean()
            raw *= kw["amp"] / np.abs(raw).max()
            st[key] = np.repeat(raw[:, :, None], g["nz"], axis=2)
    elif kind == "uniform":
        st["u"] = np.full_like(st["u"], kw.get("u", 0.0))
        st["v"] = np.full_like(st["v"], kw.get("v", 0.0))
    if kw.get("theta_noise", 0.0) > 0.0:
        rng = np.random.default_rng(kw.get("noise_seed", 0))
        st["th_v"] = st["th_v"] + rng.normal(0.0, kw["theta_noise"], size=st["th_v"].shape)
        st["rho"] = eos_rho(st["th_v"], st["pi"], c)
    return st
