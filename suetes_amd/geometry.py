"""Grid geometry on the device: map factors, Coriolis, terrain-following heights.

Mirrors ``RegionalGrid3D`` (suetes/regional3d/geometry.py:127-295) and the
vertical transforms (suetes/shared/transforms.py:16-139).  All arrays are
torch tensors in the internal layout: a reference field of shape
(NX, NY, NZ) is stored as (NZ, NY, NX) with x fastest; 2-D fields as (NY, NX).
Reference-shaped views are exposed under the reference attribute names
(``grid.Z_m`` has shape (nx, ny, nz)); the ``i_`` prefixed attributes are the
internal tensors.  Everything is differentiable in torch, so a terrain
control (h_func parameters) propagates gradients into the metrics.
"""
from __future__ import annotations

import math

import torch

R_EARTH = 6371229.0
OMEGA = 7.2921e-5


def ref_view(t: torch.Tensor) -> torch.Tensor:
    """Internal (Z,Y,X) or (Y,X) tensor -> reference-shaped view."""
    return t.permute(*reversed(range(t.dim())))


def to_internal(t: torch.Tensor) -> torch.Tensor:
    """Reference-shaped (x,y,z) tensor -> contiguous internal tensor (no copy
    when ``t`` is a view produced by :func:`ref_view`)."""
    return t.permute(*reversed(range(t.dim()))).contiguous()


# ---------------------------------------------------------------------------
# vertical transforms (transforms.py:11-139)
# ---------------------------------------------------------------------------
class GalChenSigma:
    """z = h + zeta (Lz - h)/Lz (transforms.py:16-22)."""

    def __call__(self, xi, zeta, h, Lz):
        return h + zeta * (Lz - h) / Lz


class HybridSigma:
    """transforms.py:25-35."""

    def __init__(self, scale_height=5000.0):
        self.scale_height = scale_height

    def __call__(self, xi, zeta, h, Lz):
        return zeta + h * (1.0 - zeta / Lz) * torch.exp(-zeta / self.scale_height)


class SleveSimple:
    """transforms.py:38-51."""

    def __init__(self, scale_s=4000.0, scale_l=15000.0, n=1.35):
        self.ss, self.sl, self.n = scale_s, scale_l, n

    def __call__(self, xi, zeta, h, Lz):
        zs = torch.clamp(zeta, 0.0, Lz)
        b = torch.sinh((Lz - zs) / self.ss) / math.sinh(Lz / self.ss)
        return zs + h * torch.clamp(b, 1e-8, 1.0) ** self.n


class StretchedSleveSimple:
    """transforms.py:54-86."""

    def __init__(self, stretch_kappa=2.5, scale_s=4000.0, n=1.35):
        self.kappa, self.ss, self.n = stretch_kappa, scale_s, n

    def __call__(self, xi, zeta, h, Lz):
        eta = zeta / Lz
        if abs(self.kappa) > 1e-5:
            zst = Lz * (torch.exp(self.kappa * eta) - 1.0) / (math.exp(self.kappa) - 1.0)
        else:
            zst = zeta
        zst = torch.clamp(zst, 0.0, Lz)
        b = torch.sinh((Lz - zst) / self.ss) / math.sinh(Lz / self.ss)
        return zst + h * torch.clamp(b, 1e-8, 1.0) ** self.n


class ObliqueStereographic:
    """Conformal projection (geometry.py:13-124): map factor and latitude."""

    def __init__(self, lat_center, lon_center, R_earth=R_EARTH, eps=1e-15):
        self.phi_c = math.radians(lat_center)
        self.lam_c = math.radians(lon_center)
        self.R = R_earth
        self.eps = eps

    def get_map_factor(self, x, y):
        return 1.0 + (x ** 2 + y ** 2) / (4.0 * self.R ** 2)

    def get_lat_lon(self, x, y):
        rho = torch.sqrt(x ** 2 + y ** 2)
        rho = torch.where(rho == 0, torch.full_like(rho, self.eps), rho)
        c = 2.0 * torch.atan(rho / (2.0 * self.R))
        lat = torch.asin(torch.cos(c) * math.sin(self.phi_c)
                         + (y * torch.sin(c) * math.cos(self.phi_c)) / rho)
        lon = self.lam_c + torch.atan2(x * torch.sin(c), rho * math.cos(self.phi_c) * torch.cos(c)
                                       - y * math.sin(self.phi_c) * torch.sin(c))
        return torch.rad2deg(lat), torch.rad2deg(lon)


class RegionalGrid3D:
    """Terrain-following Arakawa C grid (geometry.py:141-272).

    ``h_func(x, y)`` receives torch tensors of logical (centred) coordinates
    and returns the terrain height (a tensor or a Python scalar).
    """

    def __init__(self, nx, ny, nz, dx, dy, dz, lat_center, lon_center, h_func=None,
                 transform=None, eps=1e-15, device=None, dtype=torch.float64):
        dev = torch.device(device) if device is not None else torch.device("cpu")
        kw = dict(dtype=dtype, device=dev)
        self.device, self.dtype = dev, dtype
        self.shape = (nx, ny, nz)
        self.delta = (dx, dy, dz)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.lat_c, self.lon_c = lat_center, lon_center
        Lx, Ly = nx * dx, ny * dy
        self.Lz = nz * dz
        self.x_c = torch.arange(nx + 1, **kw) * dx - Lx / 2.0
        self.y_c = torch.arange(ny + 1, **kw) * dy - Ly / 2.0
        self.z_c = torch.arange(nz + 1, **kw) * dz
        self.x_m = 0.5 * (self.x_c[:-1] + self.x_c[1:])
        self.y_m = 0.5 * (self.y_c[:-1] + self.y_c[1:])
        self.z_m = 0.5 * (self.z_c[:-1] + self.z_c[1:])

        def mesh2(xs, ys):  # internal (Y, X)
            Y, X = torch.meshgrid(ys, xs, indexing="ij")
            return X, Y

        Xm, Ym = mesh2(self.x_m, self.y_m)
        Xu, Yu = mesh2(self.x_c, self.y_m)
        Xv, Yv = mesh2(self.x_m, self.y_c)
        self.proj = ObliqueStereographic(lat_center, lon_center, eps=eps)
        self.i_m_factors = {
            "m": self.proj.get_map_factor(Xm, Ym), "u": self.proj.get_map_factor(Xu, Yu),
            "v": self.proj.get_map_factor(Xv, Yv), "w": self.proj.get_map_factor(Xm, Ym)}
        self.i_dm_dx_m = Xm / (2.0 * self.proj.R ** 2)
        self.i_dm_dy_m = Ym / (2.0 * self.proj.R ** 2)
        lat_m, _ = self.proj.get_lat_lon(Xm, Ym)
        lat_u, _ = self.proj.get_lat_lon(Xu, Yu)
        lat_v, _ = self.proj.get_lat_lon(Xv, Yv)
        self.i_f_m = 2.0 * OMEGA * torch.sin(torch.deg2rad(lat_m))
        self.i_f_u = 2.0 * OMEGA * torch.sin(torch.deg2rad(lat_u))
        self.i_f_v = 2.0 * OMEGA * torch.sin(torch.deg2rad(lat_v))
        self.h_func = h_func if h_func is not None else (lambda x, y: 0.0)
        self.transform_op = transform if transform is not None else GalChenSigma()

        def apply(xs, ys, zs):
            Ze, Yi, Xi = torch.meshgrid(zs, ys, xs, indexing="ij")
            z = self.transform_op(Xi, Ze, self.h_func(Xi, Yi), self.Lz)
            return z + torch.zeros_like(Xi)

        self.i_Z_m = apply(self.x_m, self.y_m, self.z_m)
        self.i_Z_u = apply(self.x_c, self.y_m, self.z_m)
        self.i_Z_v = apply(self.x_m, self.y_c, self.z_m)
        self.i_Z_w = apply(self.x_m, self.y_m, self.z_c)
        Zm, Zw = self.i_Z_m, self.i_Z_w
        self.i_dz_m_full = Zw[1:] - Zw[:-1]
        bot = Zm[0] - (Zm[1] - Zm[0])
        top = Zm[-1] + (Zm[-1] - Zm[-2])
        zpad = torch.cat([bot[None], Zm, top[None]], dim=0)
        self.i_dz_w_full = zpad[1:] - zpad[:-1]
        zx = torch.cat([Zw[:, :, :1], Zw, Zw[:, :, -1:]], dim=2)
        self.i_z_xi_w = (zx[:, :, 2:] - zx[:, :, :-2]) / (2.0 * self.dx)
        zy = torch.cat([Zw[:, :1], Zw, Zw[:, -1:]], dim=1)
        self.i_z_eta_w = (zy[:, 2:] - zy[:, :-2]) / (2.0 * self.dy)
        dz_min = float(self.i_dz_m_full.detach().min())
        if dz_min <= 0.0:
            raise ValueError(f"Grid Tangling Detected! Minimum dz is {dz_min:.2f} m.\n"
                             "The 3D topography is too steep. Smooth the terrain or increase Lz.")

    # reference-shaped views --------------------------------------------------
    @property
    def m_factors(self):
        return {k: ref_view(v) for k, v in self.i_m_factors.items()}

    def __getattr__(self, name):
        if name in ("Z_m", "Z_u", "Z_v", "Z_w", "dz_m_full", "dz_w_full", "z_xi_w", "z_eta_w",
                    "f_m", "f_u", "f_v", "dm_dx_m", "dm_dy_m"):
            return ref_view(self.__dict__["i_" + name])
        raise AttributeError(name)

    def to(self, device):
        """Move every tensor attribute to ``device`` (returns self)."""
        dev = torch.device(device)
        for k, v in list(self.__dict__.items()):
            if torch.is_tensor(v):
                self.__dict__[k] = v.to(dev)
            elif isinstance(v, dict):
                self.__dict__[k] = {a: (b.to(dev) if torch.is_tensor(b) else b) for a, b in v.items()}
        self.device = dev
        return self
