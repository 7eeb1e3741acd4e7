// Fused explicit tendencies, hyperdiffusion and the predictor-corrector
// right-hand side of one SISL step (FP64, the state's precision).
//
// The reference evaluates these as ~740 array expressions per step
// (suetes/regional3d/euler.py:208-372 explicit tendencies,
// suetes/regional3d/diffusion.py:93-172 terrain-following hyperdiffusion,
// suetes/regional3d/steppers.py:742-1003 step glue).  Here they are:
//
//   hyper_faces / hyper_mass  the terrain-corrected Cartesian horizontal
//                             Laplacian of four mass fields (u_m, v_m, w_m,
//                             theta'), applied twice for del^4: a face pass
//                             (Gx at u, Gy at v) and a mass pass;
//   explicit_kernel           every explicit tendency and the arrival-point
//                             fields of the step (u_in .. th_in, the kinematic
//                             residual, theta's tendency, the packed x0);
//   thnext_kernel, rhs_kernel the predictor-corrector: theta at n+1 and the
//                             packed right-hand side of the implicit solve;
//   final_kernel              divergence damping and unscaling of the
//                             solution.
//
// Staggered operators follow CGridOperator3D (operators.py:37-132): a face
// value between mass points a and b is their difference / average; at a
// closed boundary the first / last interior face value is repeated, at a
// periodic one the face between the last and the first cell is used at
// both ends; horizontal differences carry the map factor of the target
// location.  Each thread evaluates one output point from its neighbours
// (the inputs stay in L2 across neighbouring threads), in the reference's
// operation order, so results agree with the array formulation to rounding.
#include "common.h"

namespace sisl {

struct TendCtx {
  int nx, ny, nz, px, py;
  double rdx, rdy, rdz, dz;
  double cp, grav, alpha, dt, nu_h, nu_div;
  const double *mm, *mu, *mv, *fu, *fv, *dmdx, *dmdy;
  const double *th_bg, *pi_bg, *thu_bg, *thv_bg, *thw_bg, *zxu, *zev, *zxm, *zem, *rz, *zxiw, *zetaw;
  const double *dzw, *dzm, *cpi, *dthm, *cx, *cy, *cxp, *cyp, *czs, *czp;

  // element offsets (internal layout, x fastest)
  __device__ __forceinline__ int m(int i, int j, int k) const { return (k * ny + j) * nx + i; }
  __device__ __forceinline__ int u(int i, int j, int k) const { return (k * ny + j) * (nx + 1) + i; }
  __device__ __forceinline__ int v(int i, int j, int k) const { return (k * (ny + 1) + j) * nx + i; }
  __device__ __forceinline__ int w(int i, int j, int k) const { return (k * ny + j) * nx + i; }
  __device__ __forceinline__ int m2(int i, int j) const { return j * nx + i; }
  __device__ __forceinline__ int u2(int i, int j) const { return j * (nx + 1) + i; }
  __device__ __forceinline__ int v2(int i, int j) const { return j * nx + i; }
  __device__ __forceinline__ int nm() const { return nx * ny * nz; }
  __device__ __forceinline__ int nu() const { return (nx + 1) * ny * nz; }
  __device__ __forceinline__ int nv() const { return nx * (ny + 1) * nz; }
  __device__ __forceinline__ int nw() const { return nx * ny * (nz + 1); }
};

// Mass points (a, b) whose difference / average gives face f (0..n) along an
// axis of n mass points (operators.py:37-103: closed ends repeat the first /
// last interior face, periodic ends use the face between cell n-1 and 0).
__device__ __forceinline__ void face_pair(int f, int n, int per, int &a, int &b) {
  if (per && (f == 0 || f == n)) {
    a = n - 1;
    b = 0;
    return;
  }
  const int g = min(max(f, 1), n - 1);
  a = g - 1;
  b = g;
}
// The mass level below w level kk whose vertical difference gives it (closed ends).
__device__ __forceinline__ int wlev(int kk, int nz) { return min(max(kk, 1), nz - 1); }

// diff_z(F, m->w) * rz at w level kk of column c (F a mass field).
__device__ __forceinline__ double dzw_rz(const TendCtx &t, const double *F, int i, int j, int kk) {
  const int g = wlev(kk, t.nz);
  return ((F[t.m(i, j, g)] - F[t.m(i, j, g - 1)]) * t.rdz) * t.rz[t.w(i, j, kk)];
}
// avg_z(diff_z(F) * rz) at mass level k (the vertical derivative on mass points).
__device__ __forceinline__ double dz_mass(const TendCtx &t, const double *F, int i, int j, int k) {
  return 0.5 * (dzw_rz(t, F, i, j, k) + dzw_rz(t, F, i, j, k + 1));
}

// ---- hyperdiffusion (diffusion.py:93-129, twice: del^4) ----------------------------

// Base fields of the four tendencies (diffusion.py:136-160): u_m, v_m, w_m, theta'.
__global__ void __launch_bounds__(256) hyper_base_kernel(TendCtx t, const double *__restrict__ U,
                                                         const double *__restrict__ V,
                                                         const double *__restrict__ W,
                                                         const double *__restrict__ TH, double *__restrict__ F) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = t.nm();
  if (p >= N) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  F[p] = 0.5 * (U[t.u(i, j, k)] + U[t.u(i + 1, j, k)]);
  F[N + p] = 0.5 * (V[t.v(i, j, k)] + V[t.v(i, j + 1, k)]);
  F[2 * N + p] = 0.5 * (W[t.w(i, j, k)] + W[t.w(i, j, k + 1)]);
  F[3 * N + p] = TH[p] - t.th_bg[p];
}

// Face pass: Gx = dF/dx - z_x dF/dz at u faces, Gy likewise at v faces, for
// nf mass fields F (stride nm) into GX (stride nu), GY (stride nv).
__global__ void __launch_bounds__(256) hyper_faces_kernel(TendCtx t, int nf, const double *__restrict__ F,
                                                          double *__restrict__ GX, double *__restrict__ GY) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * t.nz) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int N = t.nm(), Nu = t.nu(), Nv = t.nv();
  if (j < t.ny) {
    int a, b;
    face_pair(i, t.nx, t.px, a, b);
    const double mf = t.mu[t.u2(i, j)], zx = t.zxu[t.u(i, j, k)];
    for (int f = 0; f < nf; ++f) {
      const double *Ff = F + (size_t)f * N;
      const double dfx = ((Ff[t.m(b, j, k)] - Ff[t.m(a, j, k)]) * t.rdx) * mf;
      const double dfz = 0.5 * (dz_mass(t, Ff, a, j, k) + dz_mass(t, Ff, b, j, k));
      GX[(size_t)f * Nu + t.u(i, j, k)] = dfx - zx * dfz;
    }
  }
  if (i < t.nx) {
    int a, b;
    face_pair(j, t.ny, t.py, a, b);
    const double mf = t.mv[t.v2(i, j)], ze = t.zev[t.v(i, j, k)];
    for (int f = 0; f < nf; ++f) {
      const double *Ff = F + (size_t)f * N;
      const double dfy = ((Ff[t.m(i, b, k)] - Ff[t.m(i, a, k)]) * t.rdy) * mf;
      const double dfz = 0.5 * (dz_mass(t, Ff, i, a, k) + dz_mass(t, Ff, i, b, k));
      GY[(size_t)f * Nv + t.v(i, j, k)] = dfy - ze * dfz;
    }
  }
}

// Mass pass: L = (dGx/dx - z_x dGx_m/dz) + (dGy/dy - z_y dGy_m/dz).
__device__ __forceinline__ double gx_m(const TendCtx &t, const double *G, int i, int j, int k) {
  return 0.5 * (G[t.u(i, j, k)] + G[t.u(i + 1, j, k)]);
}
__device__ __forceinline__ double gy_m(const TendCtx &t, const double *G, int i, int j, int k) {
  return 0.5 * (G[t.v(i, j, k)] + G[t.v(i, j + 1, k)]);
}
__global__ void __launch_bounds__(256) hyper_mass_kernel(TendCtx t, int nf, const double *__restrict__ GX,
                                                         const double *__restrict__ GY, double *__restrict__ L) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = t.nm();
  if (p >= N) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  const int Nu = t.nu(), Nv = t.nv();
  const double mmv = t.mm[t.m2(i, j)];
  const int g0 = wlev(k, t.nz), g1 = wlev(k + 1, t.nz);
  const double rz0 = t.rz[t.w(i, j, k)], rz1 = t.rz[t.w(i, j, k + 1)];
  for (int f = 0; f < nf; ++f) {
    const double *Gx = GX + (size_t)f * Nu, *Gy = GY + (size_t)f * Nv;
    const double dGx = ((Gx[t.u(i + 1, j, k)] - Gx[t.u(i, j, k)]) * t.rdx) * mmv;
    const double ex0 = ((gx_m(t, Gx, i, j, g0) - gx_m(t, Gx, i, j, g0 - 1)) * t.rdz) * rz0;
    const double ex1 = ((gx_m(t, Gx, i, j, g1) - gx_m(t, Gx, i, j, g1 - 1)) * t.rdz) * rz1;
    const double dGxz = 0.5 * (ex0 + ex1);
    const double dGy = ((Gy[t.v(i, j + 1, k)] - Gy[t.v(i, j, k)]) * t.rdy) * mmv;
    const double ey0 = ((gy_m(t, Gy, i, j, g0) - gy_m(t, Gy, i, j, g0 - 1)) * t.rdz) * rz0;
    const double ey1 = ((gy_m(t, Gy, i, j, g1) - gy_m(t, Gy, i, j, g1 - 1)) * t.rdz) * rz1;
    const double dGyz = 0.5 * (ey0 + ey1);
    L[(size_t)f * N + p] = (dGx - t.zxm[p] * dGxz) + (dGy - t.zem[p] * dGyz);
  }
}

// ---- explicit tendencies and arrival fields (euler.py:208-372, steppers.py:776-850) --

struct ExplicitIO {
  const double *u, *v, *w, *pi, *th, *eta;
  const double *lap2;  // del^4 of (u_m, v_m, w_m, theta'), 4 mass fields; null when nu_h = 0
  double *u_in, *v_in, *w_in, *pi_in, *th_in, *res, *tth;
  double *x0;    // optional: packed u | v | w | pi' * 1000 | eta * dz_w (the first solve's x0)
  double *tend;  // optional: raw tendencies tu | tv | tw | tpi (the arrays' own layouts)
};

__device__ __forceinline__ double pip(const TendCtx &t, const double *PI, int q) { return PI[q] - t.pi_bg[q]; }
__device__ __forceinline__ double thp(const TendCtx &t, const double *TH, int q) { return TH[q] - t.th_bg[q]; }
// d pi' / dz at w level kk (diff_z(pi') * rz), and its mass-point average
__device__ __forceinline__ double gzw_pi(const TendCtx &t, const double *PI, int i, int j, int kk) {
  const int g = wlev(kk, t.nz);
  return ((pip(t, PI, t.m(i, j, g)) - pip(t, PI, t.m(i, j, g - 1))) * t.rdz) * t.rz[t.w(i, j, kk)];
}
__device__ __forceinline__ double gzm_pi(const TendCtx &t, const double *PI, int i, int j, int k) {
  return 0.5 * (gzw_pi(t, PI, i, j, k) + gzw_pi(t, PI, i, j, k + 1));
}
// the same for a perturbation field given directly (no background to subtract)
__device__ __forceinline__ double gzw_p(const TendCtx &t, const double *P, int i, int j, int kk) {
  const int g = wlev(kk, t.nz);
  return ((P[t.m(i, j, g)] - P[t.m(i, j, g - 1)]) * t.rdz) * t.rz[t.w(i, j, kk)];
}
__device__ __forceinline__ double gzm_p(const TendCtx &t, const double *P, int i, int j, int k) {
  return 0.5 * (gzw_p(t, P, i, j, k) + gzw_p(t, P, i, j, k + 1));
}
__device__ __forceinline__ double um_at(const TendCtx &t, const double *U, int i, int j, int k) {
  return 0.5 * (U[t.u(i, j, k)] + U[t.u(i + 1, j, k)]);
}
__device__ __forceinline__ double vm_at(const TendCtx &t, const double *V, int i, int j, int k) {
  return 0.5 * (V[t.v(i, j, k)] + V[t.v(i, j + 1, k)]);
}
// map-curvature mass term vm dm/dx - um dm/dy (euler.py:188-206)
__device__ __forceinline__ double curv_m(const TendCtx &t, const double *U, const double *V, int i, int j, int k) {
  return vm_at(t, V, i, j, k) * t.dmdx[t.m2(i, j)] - um_at(t, U, i, j, k) * t.dmdy[t.m2(i, j)];
}
// -nu_h del^4 of base field f at mass point q
__device__ __forceinline__ double hyp(const TendCtx &t, const double *L2, int f, int q) {
  return L2 ? (-t.nu_h) * L2[(size_t)f * t.nm() + q] : 0.0;
}

// Divergence of a horizontal flux pair (face arrays FXc * FX, FYc * FY) and of a
// vertical flux at mass point (i, j, k): -C_pi (m (dxx + dyy) + dzz)
// (euler.py:298-318 and 162-186).  `thx`, `thy`, `thz` multiply the fluxes when
// given (theta' at the faces), else 1.
template <typename FU, typename FV, typename FZ>
__device__ __forceinline__ double pi_div(const TendCtx &t, int i, int j, int k, FU fxu, FV fyv, FZ fzw) {
  const int q = t.m(i, j, k);
  const double mmv = t.mm[t.m2(i, j)];
  const double dzm = t.dzm[q];
  const double dxx = (((fxu(i + 1) - fxu(i)) * t.rdx) * mmv) / dzm;
  const double dyy = (((fyv(j + 1) - fyv(j)) * t.rdy) * mmv) / dzm;
  const double dzz = (((fzw(k + 1) - fzw(k)) * t.rdz) * t.dz) / dzm;
  return (-t.cpi[q]) * (mmv * (dxx + dyy) + dzz);
}

__global__ void __launch_bounds__(256) explicit_kernel(TendCtx t, ExplicitIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * (t.nz + 1)) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const double a = t.alpha, dt = t.dt, oma = 1.0 - t.alpha;
  const double *U = io.u, *V = io.v, *W = io.w, *PI = io.pi, *TH = io.th, *E = io.eta;
  const int Nu = t.nu(), Nv = t.nv(), Nw = t.nw(), N = t.nm();
  // ---- u face (i, j, k) ----
  if (j < ny && k < nz) {
    int ia, ib;
    face_pair(i, nx, t.px, ia, ib);
    const int qu = t.u(i, j, k);
    const double gx = ((pip(t, PI, t.m(ib, j, k)) - pip(t, PI, t.m(ia, j, k))) * t.rdx) * t.mu[t.u2(i, j)];
    const double gzu = 0.5 * (gzm_pi(t, PI, ia, j, k) + gzm_pi(t, PI, ib, j, k));
    const double thu = t.thu_bg[qu] + 0.5 * (thp(t, TH, t.m(ia, j, k)) + thp(t, TH, t.m(ib, j, k)));
    const double vau = 0.5 * (vm_at(t, V, ia, j, k) + vm_at(t, V, ib, j, k));
    double tu = ((-t.cp) * thu) * (gx - t.zxu[qu] * gzu);
    tu = tu + t.fu[t.u2(i, j)] * vau;
    tu = tu + (0.5 * (curv_m(t, U, V, ia, j, k) + curv_m(t, U, V, ib, j, k))) * vau;
    const double pd = io.lap2 ? 0.5 * (hyp(t, io.lap2, 0, t.m(ia, j, k)) + hyp(t, io.lap2, 0, t.m(ib, j, k))) : 0.0;
    if (io.lap2) tu = tu + pd;
    io.u_in[qu] = U[qu] + dt * (oma * tu + a * pd);
    if (io.tend) io.tend[qu] = tu;
    if (io.x0) io.x0[qu] = U[qu];
  }
  // ---- v face (i, j, k) ----
  if (i < nx && k < nz) {
    int ja, jb;
    face_pair(j, ny, t.py, ja, jb);
    const int qv = t.v(i, j, k);
    const double gy = ((pip(t, PI, t.m(i, jb, k)) - pip(t, PI, t.m(i, ja, k))) * t.rdy) * t.mv[t.v2(i, j)];
    const double gzv = 0.5 * (gzm_pi(t, PI, i, ja, k) + gzm_pi(t, PI, i, jb, k));
    const double thv = t.thv_bg[qv] + 0.5 * (thp(t, TH, t.m(i, ja, k)) + thp(t, TH, t.m(i, jb, k)));
    const double uav = 0.5 * (um_at(t, U, i, ja, k) + um_at(t, U, i, jb, k));
    double tv = ((-t.cp) * thv) * (gy - t.zev[qv] * gzv);
    tv = tv - t.fv[t.v2(i, j)] * uav;
    tv = tv + (-(0.5 * (curv_m(t, U, V, i, ja, k) + curv_m(t, U, V, i, jb, k)))) * uav;
    const double pd = io.lap2 ? 0.5 * (hyp(t, io.lap2, 1, t.m(i, ja, k)) + hyp(t, io.lap2, 1, t.m(i, jb, k))) : 0.0;
    if (io.lap2) tv = tv + pd;
    io.v_in[qv] = V[qv] + dt * (oma * tv + a * pd);
    if (io.tend) io.tend[Nu + qv] = tv;
    if (io.x0) io.x0[Nu + qv] = V[qv];
  }
  // ---- w level (i, j, k), k = 0..nz ----
  if (i < nx && j < ny) {
    const int qw = t.w(i, j, k);
    const int g = wlev(k, nz);
    const double thpw = 0.5 * (thp(t, TH, t.m(i, j, g - 1)) + thp(t, TH, t.m(i, j, g)));
    const double thw = t.thw_bg[qw] + thpw;
    const double buoy = t.grav * (thpw / t.thw_bg[qw]);
    double tw = ((-t.cp) * thw) * gzw_pi(t, PI, i, j, k) + buoy;
    const double pd = io.lap2 ? 0.5 * (hyp(t, io.lap2, 2, t.m(i, j, g - 1)) + hyp(t, io.lap2, 2, t.m(i, j, g))) : 0.0;
    if (io.lap2) tw = tw + pd;
    io.w_in[qw] = W[qw] + dt * (oma * tw + a * pd);
    if (io.tend) io.tend[Nu + Nv + qw] = tw;
    // kinematic residual dz_w eta + u_w z_xi + v_w z_eta - w (steppers.py:837-842)
    const double uw = 0.5 * (um_at(t, U, i, j, g - 1) + um_at(t, U, i, j, g));
    const double vw = 0.5 * (vm_at(t, V, i, j, g - 1) + vm_at(t, V, i, j, g));
    io.res[qw] = ((t.dzw[qw] * E[qw] + uw * t.zxiw[qw]) + vw * t.zetaw[qw]) - W[qw];
    if (io.x0) {
      io.x0[Nu + Nv + qw] = W[qw];
      io.x0[Nu + Nv + Nw + N + qw] = E[qw] * t.dzw[qw];
    }
  }
  // ---- mass point (i, j, k) ----
  if (i < nx && j < ny && k < nz) {
    const int q = t.m(i, j, k);
    // linear Exner tendency: background fluxes (euler.py:298-318)
    const double tpi = pi_div(
        t, i, j, k, [&](int f) { return U[t.u(f, j, k)] * t.cx[t.u(f, j, k)]; },
        [&](int f) { return V[t.v(i, f, k)] * t.cy[t.v(i, f, k)]; },
        [&](int kk) { return (kk == 0 || kk == nz) ? 0.0 : E[t.w(i, j, kk)] * t.czs[t.w(i, j, kk)]; });
    // theta' Exner flux (euler.py:162-186): theta' averaged to the faces
    auto thpu = [&](int f) {
      int ia, ib;
      face_pair(f, nx, t.px, ia, ib);
      return 0.5 * (thp(t, TH, t.m(ia, j, k)) + thp(t, TH, t.m(ib, j, k)));
    };
    auto thpv = [&](int f) {
      int ja, jb;
      face_pair(f, ny, t.py, ja, jb);
      return 0.5 * (thp(t, TH, t.m(i, ja, k)) + thp(t, TH, t.m(i, jb, k)));
    };
    const double nl = pi_div(
        t, i, j, k, [&](int f) { return U[t.u(f, j, k)] * thpu(f) * t.cxp[t.u(f, j, k)]; },
        [&](int f) { return V[t.v(i, f, k)] * thpv(f) * t.cyp[t.v(i, f, k)]; },
        [&](int kk) {
          if (kk == 0 || kk == nz) return 0.0;
          const double tw_ = 0.5 * (thp(t, TH, t.m(i, j, kk - 1)) + thp(t, TH, t.m(i, j, kk)));
          return E[t.w(i, j, kk)] * tw_ * t.czp[t.w(i, j, kk)];
        });
    const double wm = 0.5 * (W[t.w(i, j, k)] + W[t.w(i, j, k + 1)]);
    double tth = (-wm) * t.dthm[q];
    if (io.lap2) tth = tth + hyp(t, io.lap2, 3, q);
    const double pp = pip(t, PI, q);
    io.pi_in[q] = pp + (oma * dt) * (tpi + nl);
    io.th_in[q] = thp(t, TH, q) + (oma * dt) * tth;
    io.tth[q] = tth;
    if (io.tend) io.tend[Nu + Nv + Nw + q] = tpi;
    if (io.x0) io.x0[Nu + Nv + Nw + q] = pp * 1000.0;
  }
}

// ---- slow tendencies of the split-explicit core (steppers.py:1358-1470) --------------
//
// adv_kernel: the flux-form advection terms adv(q) of theta', u_m, v_m and w_m
// at mass points, with the third-order upwind-biased face values
// (steppers.py:1310-1356: fourth-order centred part plus a twelfth of the third
// difference signed by the face flux, the sign gated to 0 below eps_l; edge
// padding on closed axes, wrap on periodic ones).  slow_kernel: the explicit
// tendencies of the stage state (as explicit_kernel) plus the hyperdiffusion of
// the stage and of the step's start state and the averaged advection terms.

struct AdvIO {
  const double *u, *v, *w, *th, *eta;
  double eps_l;
  double *adv;  // 4 mass fields: adv(theta'), adv(u_m), adv(v_m), adv(w_m)
};

__device__ __forceinline__ int pad_index(int idx, int n, int per) {
  return per ? ((idx % n) + n) % n : min(max(idx, 0), n - 1);
}
__device__ __forceinline__ double upwind3(double flux, double qm2, double qm1, double q0, double qp1, double eps) {
  const double q4 = (7.0 / 12.0) * (q0 + qm1) - (1.0 / 12.0) * (qp1 + qm2);
  const double qb = (1.0 / 12.0) * ((qp1 - qm2) - 3.0 * (q0 - qm1));
  const double sg = fabs(flux) < eps ? 0.0 : (flux > 0.0 ? 1.0 : (flux < 0.0 ? -1.0 : 0.0));
  return flux * (q4 + sg * qb);
}
// the advected mass field F at (i, j, k): 0 theta', 1 u_m, 2 v_m, 3 w_m
template <int F>
__device__ __forceinline__ double adv_q(const TendCtx &t, const AdvIO &io, int i, int j, int k) {
  if constexpr (F == 0) return io.th[t.m(i, j, k)] - t.th_bg[t.m(i, j, k)];
  else if constexpr (F == 1) return 0.5 * (io.u[t.u(i, j, k)] + io.u[t.u(i + 1, j, k)]);
  else if constexpr (F == 2) return 0.5 * (io.v[t.v(i, j, k)] + io.v[t.v(i, j + 1, k)]);
  else return 0.5 * (io.w[t.w(i, j, k)] + io.w[t.w(i, j, k + 1)]);
}
template <int F>
__device__ __forceinline__ double adv_flux_x(const TendCtx &t, const AdvIO &io, int f, int j, int k) {
  const double U = io.u[t.u(f, j, k)] / t.mu[t.u2(f, j)];
  return upwind3(U, adv_q<F>(t, io, pad_index(f - 2, t.nx, t.px), j, k),
                 adv_q<F>(t, io, pad_index(f - 1, t.nx, t.px), j, k), adv_q<F>(t, io, pad_index(f, t.nx, t.px), j, k),
                 adv_q<F>(t, io, pad_index(f + 1, t.nx, t.px), j, k), io.eps_l);
}
template <int F>
__device__ __forceinline__ double adv_flux_y(const TendCtx &t, const AdvIO &io, int i, int f, int k) {
  const double V = io.v[t.v(i, f, k)] / t.mv[t.v2(i, f)];
  return upwind3(V, adv_q<F>(t, io, i, pad_index(f - 2, t.ny, t.py), k),
                 adv_q<F>(t, io, i, pad_index(f - 1, t.ny, t.py), k), adv_q<F>(t, io, i, pad_index(f, t.ny, t.py), k),
                 adv_q<F>(t, io, i, pad_index(f + 1, t.ny, t.py), k), io.eps_l);
}
template <int F>
__device__ __forceinline__ double adv_flux_z(const TendCtx &t, const AdvIO &io, int i, int j, int kk) {
  const double W = io.eta[t.w(i, j, kk)] * t.dzw[t.w(i, j, kk)];
  return upwind3(W, adv_q<F>(t, io, i, j, pad_index(kk - 2, t.nz, 0)), adv_q<F>(t, io, i, j, pad_index(kk - 1, t.nz, 0)),
                 adv_q<F>(t, io, i, j, pad_index(kk, t.nz, 0)), adv_q<F>(t, io, i, j, pad_index(kk + 1, t.nz, 0)),
                 io.eps_l);
}
// adv(q) = -m^2 (ax - q divU + ay - q divV) - (az - q divW) at (i, j, k)
template <int F>
__device__ __forceinline__ double adv_term(const TendCtx &t, const AdvIO &io, int i, int j, int k) {
  const double mmv = t.mm[t.m2(i, j)];
  const double rzm = t.dz / t.dzm[t.m(i, j, k)];
  const double q = adv_q<F>(t, io, i, j, k);
  const double ax = ((adv_flux_x<F>(t, io, i + 1, j, k) - adv_flux_x<F>(t, io, i, j, k)) * t.rdx) * mmv;
  const double ay = ((adv_flux_y<F>(t, io, i, j + 1, k) - adv_flux_y<F>(t, io, i, j, k)) * t.rdy) * mmv;
  const double az = ((adv_flux_z<F>(t, io, i, j, k + 1) - adv_flux_z<F>(t, io, i, j, k)) * t.rdz) * rzm;
  const double divU = ((io.u[t.u(i + 1, j, k)] / t.mu[t.u2(i + 1, j)] - io.u[t.u(i, j, k)] / t.mu[t.u2(i, j)]) *
                       t.rdx) * mmv;
  const double divV = ((io.v[t.v(i, j + 1, k)] / t.mv[t.v2(i, j + 1)] - io.v[t.v(i, j, k)] / t.mv[t.v2(i, j)]) *
                       t.rdy) * mmv;
  const double divW = ((io.eta[t.w(i, j, k + 1)] * t.dzw[t.w(i, j, k + 1)] -
                        io.eta[t.w(i, j, k)] * t.dzw[t.w(i, j, k)]) * t.rdz) * rzm;
  return (-(mmv * mmv)) * (((ax - q * divU) + ay) - q * divV) - (az - q * divW);
}

__global__ void __launch_bounds__(256) adv_kernel(TendCtx t, AdvIO io) {
  const int p = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int N = t.nm();
  if (p >= N) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  io.adv[p] = adv_term<0>(t, io, i, j, k);
  io.adv[N + p] = adv_term<1>(t, io, i, j, k);
  io.adv[2 * N + p] = adv_term<2>(t, io, i, j, k);
  io.adv[3 * N + p] = adv_term<3>(t, io, i, j, k);
}

struct SlowIO {
  const double *u, *v, *w, *pi, *th, *eta;
  const double *lap2s, *lap20;  // del^4 of the stage / start state (4 mass fields each); null: omitted
  const double *adv;            // output of adv_kernel
  double *tend;                 // tu | tv | tw | tpi | tth
};

__global__ void __launch_bounds__(256) slow_kernel(TendCtx t, SlowIO io) {
  const int p = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * (t.nz + 1)) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const double *U = io.u, *V = io.v, *W = io.w, *PI = io.pi, *TH = io.th, *E = io.eta;
  const int Nu = t.nu(), Nv = t.nv(), Nw = t.nw(), N = t.nm();
  const double *A = io.adv;
  if (j < ny && k < nz) {
    int ia, ib;
    face_pair(i, nx, t.px, ia, ib);
    const int qu = t.u(i, j, k);
    const double gx = ((pip(t, PI, t.m(ib, j, k)) - pip(t, PI, t.m(ia, j, k))) * t.rdx) * t.mu[t.u2(i, j)];
    const double gzu = 0.5 * (gzm_pi(t, PI, ia, j, k) + gzm_pi(t, PI, ib, j, k));
    const double thu = t.thu_bg[qu] + 0.5 * (thp(t, TH, t.m(ia, j, k)) + thp(t, TH, t.m(ib, j, k)));
    const double vau = 0.5 * (vm_at(t, V, ia, j, k) + vm_at(t, V, ib, j, k));
    double tu = ((-t.cp) * thu) * (gx - t.zxu[qu] * gzu);
    tu = tu + t.fu[t.u2(i, j)] * vau;
    tu = tu + (0.5 * (curv_m(t, U, V, ia, j, k) + curv_m(t, U, V, ib, j, k))) * vau;
    if (io.lap2s) tu = tu + 0.5 * (hyp(t, io.lap2s, 0, t.m(ia, j, k)) + hyp(t, io.lap2s, 0, t.m(ib, j, k)));
    if (io.lap20) tu = tu + 0.5 * (hyp(t, io.lap20, 0, t.m(ia, j, k)) + hyp(t, io.lap20, 0, t.m(ib, j, k)));
    tu = tu + 0.5 * (A[N + t.m(ia, j, k)] + A[N + t.m(ib, j, k)]);
    io.tend[qu] = tu;
  }
  if (i < nx && k < nz) {
    int ja, jb;
    face_pair(j, ny, t.py, ja, jb);
    const int qv = t.v(i, j, k);
    const double gy = ((pip(t, PI, t.m(i, jb, k)) - pip(t, PI, t.m(i, ja, k))) * t.rdy) * t.mv[t.v2(i, j)];
    const double gzv = 0.5 * (gzm_pi(t, PI, i, ja, k) + gzm_pi(t, PI, i, jb, k));
    const double thv = t.thv_bg[qv] + 0.5 * (thp(t, TH, t.m(i, ja, k)) + thp(t, TH, t.m(i, jb, k)));
    const double uav = 0.5 * (um_at(t, U, i, ja, k) + um_at(t, U, i, jb, k));
    double tv = ((-t.cp) * thv) * (gy - t.zev[qv] * gzv);
    tv = tv - t.fv[t.v2(i, j)] * uav;
    tv = tv + (-(0.5 * (curv_m(t, U, V, i, ja, k) + curv_m(t, U, V, i, jb, k)))) * uav;
    if (io.lap2s) tv = tv + 0.5 * (hyp(t, io.lap2s, 1, t.m(i, ja, k)) + hyp(t, io.lap2s, 1, t.m(i, jb, k)));
    if (io.lap20) tv = tv + 0.5 * (hyp(t, io.lap20, 1, t.m(i, ja, k)) + hyp(t, io.lap20, 1, t.m(i, jb, k)));
    tv = tv + 0.5 * (A[2 * N + t.m(i, ja, k)] + A[2 * N + t.m(i, jb, k)]);
    io.tend[Nu + qv] = tv;
  }
  if (i < nx && j < ny) {
    const int qw = t.w(i, j, k);
    const int g = wlev(k, nz);
    const double thpw = 0.5 * (thp(t, TH, t.m(i, j, g - 1)) + thp(t, TH, t.m(i, j, g)));
    const double thw = t.thw_bg[qw] + thpw;
    const double buoy = t.grav * (thpw / t.thw_bg[qw]);
    double tw = ((-t.cp) * thw) * gzw_pi(t, PI, i, j, k) + buoy;
    if (io.lap2s) tw = tw + 0.5 * (hyp(t, io.lap2s, 2, t.m(i, j, g - 1)) + hyp(t, io.lap2s, 2, t.m(i, j, g)));
    if (io.lap20) tw = tw + 0.5 * (hyp(t, io.lap20, 2, t.m(i, j, g - 1)) + hyp(t, io.lap20, 2, t.m(i, j, g)));
    tw = tw + 0.5 * (A[3 * N + t.m(i, j, g - 1)] + A[3 * N + t.m(i, j, g)]);
    io.tend[Nu + Nv + qw] = tw;
  }
  if (i < nx && j < ny && k < nz) {
    const int q = t.m(i, j, k);
    const double tpi = pi_div(
        t, i, j, k, [&](int f) { return U[t.u(f, j, k)] * t.cx[t.u(f, j, k)]; },
        [&](int f) { return V[t.v(i, f, k)] * t.cy[t.v(i, f, k)]; },
        [&](int kk) { return (kk == 0 || kk == nz) ? 0.0 : E[t.w(i, j, kk)] * t.czs[t.w(i, j, kk)]; });
    auto thpu = [&](int f) {
      int ia, ib;
      face_pair(f, nx, t.px, ia, ib);
      return 0.5 * (thp(t, TH, t.m(ia, j, k)) + thp(t, TH, t.m(ib, j, k)));
    };
    auto thpv = [&](int f) {
      int ja, jb;
      face_pair(f, ny, t.py, ja, jb);
      return 0.5 * (thp(t, TH, t.m(i, ja, k)) + thp(t, TH, t.m(i, jb, k)));
    };
    const double nl = pi_div(
        t, i, j, k, [&](int f) { return U[t.u(f, j, k)] * thpu(f) * t.cxp[t.u(f, j, k)]; },
        [&](int f) { return V[t.v(i, f, k)] * thpv(f) * t.cyp[t.v(i, f, k)]; },
        [&](int kk) {
          if (kk == 0 || kk == nz) return 0.0;
          const double tw_ = 0.5 * (thp(t, TH, t.m(i, j, kk - 1)) + thp(t, TH, t.m(i, j, kk)));
          return E[t.w(i, j, kk)] * tw_ * t.czp[t.w(i, j, kk)];
        });
    const double wm = 0.5 * (W[t.w(i, j, k)] + W[t.w(i, j, k + 1)]);
    double tth = (-wm) * t.dthm[q];
    if (io.lap2s) tth = tth + hyp(t, io.lap2s, 3, q);
    if (io.lap20) tth = tth + hyp(t, io.lap20, 3, q);
    tth = tth + A[q];
    io.tend[Nu + Nv + Nw + q] = tpi + nl;
    io.tend[Nu + Nv + Nw + N + q] = tth;
  }
}

// ---- predictor-corrector (steppers.py:911-998) --------------------------------------

struct PredictorIO {
  const double *rhs_u, *rhs_v, *rhs_w, *rhs_pi, *th_adv, *tth, *w_n, *reta;
  const double *g_pi, *g_u, *g_v, *g_w, *g_eta;  // current guess (physical units)
  const double *m_u, *m_v;                       // winds of the metric and kinematic terms
  double *th_next, *b;                           // theta at n+1; packed right-hand side
};

// theta at n+1 from the current guess's w (steppers.py:911-919).
__global__ void __launch_bounds__(256) thnext_kernel(TendCtx t, PredictorIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= t.nm()) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  const double wm = 0.5 * (io.w_n[t.w(i, j, k)] + io.w_n[t.w(i, j, k + 1)]);
  const double wmn = 0.5 * (io.g_w[t.w(i, j, k)] + io.g_w[t.w(i, j, k + 1)]);
  const double tth1 = io.tth[p] + (wm - wmn) * t.dthm[p];
  io.th_next[p] = (io.th_adv[p] + (t.alpha * t.dt) * tth1) + t.th_bg[p];
}

// Packed right-hand side u | v | w | pi * 1000 | eta of the implicit solve
// (steppers.py:920-968): the nonlinear pressure-gradient and buoyancy
// corrections with theta at n+1, the map-curvature terms of the current
// winds, the kinematic bottom row and the theta' Exner flux of the guess.
__global__ void __launch_bounds__(256) rhs_kernel(TendCtx t, PredictorIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * (t.nz + 1)) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const double adt = t.alpha * t.dt;
  const double cad = (t.cp * t.alpha) * t.dt;
  const int Nu = t.nu(), Nv = t.nv(), Nw = t.nw();
  const double *TN = io.th_next;
  auto tp1 = [&](int q) { return TN[q] - t.th_bg[q]; };
  if (j < ny && k < nz) {
    int ia, ib;
    face_pair(i, nx, t.px, ia, ib);
    const int qu = t.u(i, j, k);
    const double thu1 = 0.5 * (tp1(t.m(ia, j, k)) + tp1(t.m(ib, j, k)));
    const double gx = ((io.g_pi[t.m(ib, j, k)] - io.g_pi[t.m(ia, j, k)]) * t.rdx) * t.mu[t.u2(i, j)];
    const double gzu = 0.5 * (gzm_p(t, io.g_pi, ia, j, k) + gzm_p(t, io.g_pi, ib, j, k));
    const double nlu = ((-cad) * thu1) * (gx - t.zxu[qu] * gzu);
    const double vau = 0.5 * (vm_at(t, io.m_v, ia, j, k) + vm_at(t, io.m_v, ib, j, k));
    const double mun = (0.5 * (curv_m(t, io.m_u, io.m_v, ia, j, k) + curv_m(t, io.m_u, io.m_v, ib, j, k))) * vau;
    io.b[qu] = (io.rhs_u[qu] + nlu) + adt * mun;
  }
  if (i < nx && k < nz) {
    int ja, jb;
    face_pair(j, ny, t.py, ja, jb);
    const int qv = t.v(i, j, k);
    const double thv1 = 0.5 * (tp1(t.m(i, ja, k)) + tp1(t.m(i, jb, k)));
    const double gy = ((io.g_pi[t.m(i, jb, k)] - io.g_pi[t.m(i, ja, k)]) * t.rdy) * t.mv[t.v2(i, j)];
    const double gzv = 0.5 * (gzm_p(t, io.g_pi, i, ja, k) + gzm_p(t, io.g_pi, i, jb, k));
    const double nlv = ((-cad) * thv1) * (gy - t.zev[qv] * gzv);
    const double uav = 0.5 * (um_at(t, io.m_u, i, ja, k) + um_at(t, io.m_u, i, jb, k));
    const double mvn =
        (-(0.5 * (curv_m(t, io.m_u, io.m_v, i, ja, k) + curv_m(t, io.m_u, io.m_v, i, jb, k)))) * uav;
    io.b[Nu + qv] = (io.rhs_v[qv] + nlv) + adt * mvn;
  }
  if (i < nx && j < ny) {
    const int qw = t.w(i, j, k);
    double bw;
    if (k == 0) {  // kinematic bottom: m_w (u_w z_xi + v_w z_eta) of the current winds
      const double uw = 0.5 * (um_at(t, io.m_u, i, j, 0) + um_at(t, io.m_u, i, j, 1));
      const double vw = 0.5 * (vm_at(t, io.m_v, i, j, 0) + vm_at(t, io.m_v, i, j, 1));
      bw = t.mm[t.m2(i, j)] * (uw * t.zxiw[qw] + vw * t.zetaw[qw]);
    } else if (k == nz) {
      bw = 0.0;
    } else {
      const double thw1 = 0.5 * (tp1(t.m(i, j, k - 1)) + tp1(t.m(i, j, k)));
      const double buoy = adt * (t.grav * (thw1 / t.thw_bg[qw]));
      const double nlw = ((-cad) * thw1) * gzw_p(t, io.g_pi, i, j, k);
      bw = (io.rhs_w[qw] + buoy) + nlw;
    }
    io.b[Nu + Nv + qw] = bw;
    // eta row: -(1 - alpha) times the advected kinematic residual, ends zero
    const double re = (k == 0 || k == nz) ? 0.0 : (-(1.0 - t.alpha)) * io.reta[qw];
    io.b[Nu + Nv + Nw + t.nm() + qw] = re;
  }
  if (i < nx && j < ny && k < nz) {
    const int q = t.m(i, j, k);
    auto thu = [&](int f) {
      int ia, ib;
      face_pair(f, nx, t.px, ia, ib);
      return 0.5 * (tp1(t.m(ia, j, k)) + tp1(t.m(ib, j, k)));
    };
    auto thv = [&](int f) {
      int ja, jb;
      face_pair(f, ny, t.py, ja, jb);
      return 0.5 * (tp1(t.m(i, ja, k)) + tp1(t.m(i, jb, k)));
    };
    const double nld = pi_div(
        t, i, j, k, [&](int f) { return io.g_u[t.u(f, j, k)] * thu(f) * t.cxp[t.u(f, j, k)]; },
        [&](int f) { return io.g_v[t.v(i, f, k)] * thv(f) * t.cyp[t.v(i, f, k)]; },
        [&](int kk) {
          if (kk == 0 || kk == nz) return 0.0;
          const double tw_ = 0.5 * (tp1(t.m(i, j, kk - 1)) + tp1(t.m(i, j, kk)));
          return io.g_eta[t.w(i, j, kk)] * tw_ * t.czp[t.w(i, j, kk)];
        });
    io.b[Nu + Nv + Nw + q] = (io.rhs_pi[q] + adt * nld) * 1000.0;
  }
}

// ---- divergence damping and unscaling (steppers.py:1003-1035) ----------------------

struct FinalIO {
  const double *x;  // packed solution u | v | w | pi * 1000 | eta * dz_w
  double *u, *v, *pi, *eta;
};

__global__ void __launch_bounds__(256) final_kernel(TendCtx t, FinalIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * (t.nz + 1)) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const int Nu = t.nu(), Nv = t.nv(), Nw = t.nw(), N = t.nm();
  const double *XU = io.x, *XV = io.x + Nu;
  const double kd = t.nu_div * t.dt;
  auto div = [&](int ii, int jj) {
    const double mmv = t.mm[t.m2(ii, jj)];
    return ((XU[t.u(ii + 1, jj, k)] - XU[t.u(ii, jj, k)]) * t.rdx) * mmv +
           ((XV[t.v(ii, jj + 1, k)] - XV[t.v(ii, jj, k)]) * t.rdy) * mmv;
  };
  if (j < ny && k < nz) {
    int ia, ib;
    face_pair(i, nx, t.px, ia, ib);
    const int qu = t.u(i, j, k);
    const double dd = ((div(ib, j) - div(ia, j)) * t.rdx) * t.mu[t.u2(i, j)];
    io.u[qu] = XU[qu] + kd * dd;
  }
  if (i < nx && k < nz) {
    int ja, jb;
    face_pair(j, ny, t.py, ja, jb);
    const int qv = t.v(i, j, k);
    const double dd = ((div(i, jb) - div(i, ja)) * t.rdy) * t.mv[t.v2(i, j)];
    io.v[qv] = XV[qv] + kd * dd;
  }
  if (i < nx && j < ny) {
    const int qw = t.w(i, j, k);
    io.eta[qw] = io.x[Nu + Nv + Nw + N + qw] / t.dzw[qw];
    if (k < nz) {
      const int q = t.m(i, j, k);
      io.pi[q] = io.x[Nu + Nv + Nw + q] / 1000.0 + t.pi_bg[q];
    }
  }
}

static bool make_ctx(const SislTendency *d, TendCtx &t) {
  if (!d) return false;
  const SislGrid &g = d->g;
  if (g.nx < 2 || g.ny < 2 || g.nz < 2) return false;
  if ((int64_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1) * 4 >= (int64_t(1) << 31)) return false;
  t.nx = g.nx; t.ny = g.ny; t.nz = g.nz; t.px = g.periodic_x != 0; t.py = g.periodic_y != 0;
  t.rdx = 1.0 / d->dx; t.rdy = 1.0 / d->dy; t.rdz = 1.0 / d->dz; t.dz = d->dz;
  t.cp = d->cp; t.grav = d->grav; t.alpha = d->alpha; t.dt = d->dt; t.nu_h = d->nu_h; t.nu_div = d->nu_div;
  t.mm = d->mm; t.mu = d->mu; t.mv = d->mv; t.fu = d->fu; t.fv = d->fv; t.dmdx = d->dmdx; t.dmdy = d->dmdy;
  t.th_bg = d->th_bg; t.pi_bg = d->pi_bg; t.thu_bg = d->thu_bg; t.thv_bg = d->thv_bg; t.thw_bg = d->thw_bg;
  t.zxu = d->zxu; t.zev = d->zev; t.zxm = d->zxm; t.zem = d->zem; t.rz = d->rz; t.zxiw = d->zxiw;
  t.zetaw = d->zetaw; t.dzw = d->dzw; t.dzm = d->dzm; t.cpi = d->cpi; t.dthm = d->dthm; t.cx = d->cx;
  t.cy = d->cy; t.cxp = d->cxp; t.cyp = d->cyp; t.czs = d->czs; t.czp = d->czp;
  const void *ptrs[] = {t.mm, t.mu, t.mv, t.fu, t.fv, t.dmdx, t.dmdy, t.th_bg, t.pi_bg, t.thu_bg, t.thv_bg,
                        t.thw_bg, t.zxu, t.zev, t.zxm, t.zem, t.rz, t.zxiw, t.zetaw, t.dzw, t.dzm, t.cpi,
                        t.dthm, t.cx, t.cy, t.cxp, t.cyp, t.czs, t.czp};
  for (const void *q : ptrs)
    if (!q) return false;
  return true;
}

static unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---- reverse mode of the explicit stage (vector-Jacobian products) -----------------
//
// Cotangents of (u, v, w, pi, theta, eta_dot) from those of the stage's outputs
// (u_in, v_in, w_in, pi_in, th_in, the kinematic residual, theta's tendency),
// following the reference's autodiff of the same array expressions: the values
// that the staggered operators repeat at closed edges (m -> u, v faces 0 and n,
// m -> w levels 0 and nz) are stop-gradient copies (operators.py:37-63), so a
// cotangent arriving there is dropped; on a periodic axis faces 0 and n both
// pass theirs to cells n - 1 and 0.  Every kernel is in gather form: a thread
// sums the contributions to its own elements in a fixed order, so gradients
// are reproducible.  Five passes (plus four for the hyperdiffusion):
//   adjA (faces, w levels): the cotangents of the face / w-level intermediates
//                           (theta' averages, pressure gradients, dpi/dz
//                           averages, u_m / v_m averages, curvature, del^4);
//   adjB (mass points):     their transposes to mass-point intermediates and
//                           the Exner-divergence coefficients;
//   lapT faces / mass (x2): the transposed terrain-following Laplacian;
//   adjD (faces, w levels): cotangents of u, v, w, eta_dot;
//   adjE (mass points):     cotangents of pi and theta.

// Face f of n is a stop-gradient copy (closed axis, f == 0 or f == n).
__device__ __forceinline__ bool face_frozen(int f, int n, int per) { return !per && (f == 0 || f == n); }

// Transpose of a face quantity X built from the mass pair face_pair(f) = (a, b):
// the sum over the live faces of wa X(f) where a == c and wb X(f) where b == c.
template <class FX>
__device__ __forceinline__ double face_gather(int c, int n, int per, double wa, double wb, FX X) {
  double s = 0.0;
  if (!face_frozen(c, n, per)) s += wb * X(c);          // face c: (c - 1, c), or periodic face 0
  if (!face_frozen(c + 1, n, per)) s += wa * X(c + 1);  // face c + 1: (c, c + 1), or periodic face n
  if (per && c == 0) s += wb * X(n);                    // periodic face n: (n - 1, 0)
  if (per && c == n - 1) s += wa * X(0);                // periodic face 0: (n - 1, 0)
  return s;
}
// The same along z for m -> w quantities (w levels 0 and nz are frozen).
template <class FX>
__device__ __forceinline__ double level_gather(int c, int nz, double wa, double wb, FX X) {
  double s = 0.0;
  if (c >= 1) s += wb * X(c);
  if (c + 1 <= nz - 1) s += wa * X(c + 1);
  return s;
}
__device__ __forceinline__ double ldz(const double *p, int q) { return p ? p[q] : 0.0; }

struct ExplicitAdjIO {
  const double *u, *v, *w, *pi, *th, *eta;                        // saved inputs
  const double *gu_in, *gv_in, *gw_in, *gpi_in, *gth_in, *gres, *gtth;  // output cotangents (null: zero)
  int hyper;                                                       // the stage included -nu_h del^4
  double *cu, *cv, *cw;  // adjA: 6 u-face, 6 v-face, 5 w-level arrays
  double *cm;            // adjB: gpp, ggzm, gum, gvm, gwm, 4 del^4 cotangents, cx, cy, cz, gthp
  const double *fb;      // del^4 transposed: cotangents of the four base fields (null: none)
  double *tu, *tv, *tw;  // adjD: theta' flux cotangents at u, v faces and w levels
  double *gu, *gv, *gw, *gpi, *gth, *geta;  // input cotangents
};

__global__ void __launch_bounds__(256) adjA_kernel(TendCtx t, ExplicitAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * (t.nz + 1)) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const double dt = t.dt, oma = 1.0 - t.alpha;
  const double *U = io.u, *V = io.v, *PI = io.pi, *TH = io.th;
  const int Nu = t.nu(), Nv = t.nv(), Nw = t.nw();
  if (j < ny && k < nz) {  // u face (i, j, k)
    const int qu = t.u(i, j, k);
    double c[6] = {0, 0, 0, 0, 0, 0};
    const double gin = ldz(io.gu_in, qu);
    if (!face_frozen(i, nx, t.px) && gin != 0.0) {
      int ia, ib;
      face_pair(i, nx, t.px, ia, ib);
      const double gx = ((pip(t, PI, t.m(ib, j, k)) - pip(t, PI, t.m(ia, j, k))) * t.rdx) * t.mu[t.u2(i, j)];
      const double gzu = 0.5 * (gzm_pi(t, PI, ia, j, k) + gzm_pi(t, PI, ib, j, k));
      const double thu = t.thu_bg[qu] + 0.5 * (thp(t, TH, t.m(ia, j, k)) + thp(t, TH, t.m(ib, j, k)));
      const double vau = 0.5 * (vm_at(t, V, ia, j, k) + vm_at(t, V, ib, j, k));
      const double cvu = 0.5 * (curv_m(t, U, V, ia, j, k) + curv_m(t, U, V, ib, j, k));
      const double g = gin * (dt * oma);
      const double gc = g * ((-t.cp) * thu);
      c[0] = g * ((-t.cp) * (gx - t.zxu[qu] * gzu));  // -> theta' (average)
      c[1] = gc * (t.rdx * t.mu[t.u2(i, j)]);           // -> pi' (difference)
      c[2] = gc * (-t.zxu[qu]);                         // -> dpi/dz at mass points (average)
      c[3] = g * (t.fu[t.u2(i, j)] + cvu);              // -> v_m (average)
      c[4] = g * vau;                                   // -> curvature term (average)
      c[5] = io.hyper ? gin * dt : 0.0;                 // -> -nu_h del^4 u_m (average)
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) io.cu[(size_t)r * Nu + qu] = c[r];
  }
  if (i < nx && k < nz) {  // v face (i, j, k)
    const int qv = t.v(i, j, k);
    double c[6] = {0, 0, 0, 0, 0, 0};
    const double gin = ldz(io.gv_in, qv);
    if (!face_frozen(j, ny, t.py) && gin != 0.0) {
      int ja, jb;
      face_pair(j, ny, t.py, ja, jb);
      const double gy = ((pip(t, PI, t.m(i, jb, k)) - pip(t, PI, t.m(i, ja, k))) * t.rdy) * t.mv[t.v2(i, j)];
      const double gzv = 0.5 * (gzm_pi(t, PI, i, ja, k) + gzm_pi(t, PI, i, jb, k));
      const double thv = t.thv_bg[qv] + 0.5 * (thp(t, TH, t.m(i, ja, k)) + thp(t, TH, t.m(i, jb, k)));
      const double uav = 0.5 * (um_at(t, U, i, ja, k) + um_at(t, U, i, jb, k));
      const double cvv = 0.5 * (curv_m(t, U, V, i, ja, k) + curv_m(t, U, V, i, jb, k));
      const double g = gin * (dt * oma);
      const double gc = g * ((-t.cp) * thv);
      c[0] = g * ((-t.cp) * (gy - t.zev[qv] * gzv));
      c[1] = gc * (t.rdy * t.mv[t.v2(i, j)]);
      c[2] = gc * (-t.zev[qv]);
      c[3] = g * (-t.fv[t.v2(i, j)] - cvv);  // -> u_m (average)
      c[4] = g * (-uav);
      c[5] = io.hyper ? gin * dt : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) io.cv[(size_t)r * Nv + qv] = c[r];
  }
  if (i < nx && j < ny) {  // w level (i, j, k), k = 0..nz
    const int qw = t.w(i, j, k);
    double c[5] = {0, 0, 0, 0, 0};
    if (k >= 1 && k <= nz - 1) {
      const double gin = ldz(io.gw_in, qw), gr = ldz(io.gres, qw);
      const double thpw = 0.5 * (thp(t, TH, t.m(i, j, k - 1)) + thp(t, TH, t.m(i, j, k)));
      const double thw = t.thw_bg[qw] + thpw;
      const double gzw = gzw_pi(t, PI, i, j, k);
      const double g = gin * (dt * oma);
      c[0] = g * ((-t.cp) * gzw + t.grav / t.thw_bg[qw]);         // -> theta' (average)
      c[1] = g * ((-t.cp) * thw) * (t.rdz * t.rz[qw]);            // -> pi' (difference)
      c[2] = io.hyper ? gin * dt : 0.0;                            // -> -nu_h del^4 w_m (average)
      c[3] = gr * t.zxiw[qw];                                      // -> u_m (average)
      c[4] = gr * t.zetaw[qw];                                     // -> v_m (average)
    }
#pragma unroll
    for (int r = 0; r < 5; ++r) io.cw[(size_t)r * Nw + qw] = c[r];
  }
}

__global__ void __launch_bounds__(256) adjB_kernel(TendCtx t, ExplicitAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = t.nm();
  if (p >= N) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const int Nu = t.nu(), Nv = t.nv(), Nw = t.nw();
  const double dt = t.dt, oma = 1.0 - t.alpha;
  auto CU = [&](int r) { return [&, r](int f) { return io.cu[(size_t)r * Nu + t.u(f, j, k)]; }; };
  auto CV = [&](int r) { return [&, r](int f) { return io.cv[(size_t)r * Nv + t.v(i, f, k)]; }; };
  auto CW = [&](int r) { return [&, r](int kk) { return io.cw[(size_t)r * Nw + t.w(i, j, kk)]; }; };
  const double gpin = ldz(io.gpi_in, p), gthin = ldz(io.gth_in, p);
  const double gpp = gpin + face_gather(i, nx, t.px, -1.0, 1.0, CU(1)) + face_gather(j, ny, t.py, -1.0, 1.0, CV(1)) +
                     level_gather(k, nz, -1.0, 1.0, CW(1));
  const double ggzm = face_gather(i, nx, t.px, 0.5, 0.5, CU(2)) + face_gather(j, ny, t.py, 0.5, 0.5, CV(2));
  const double gcv = face_gather(i, nx, t.px, 0.5, 0.5, CU(4)) + face_gather(j, ny, t.py, 0.5, 0.5, CV(4));
  const double gvm = face_gather(i, nx, t.px, 0.5, 0.5, CU(3)) + gcv * t.dmdx[t.m2(i, j)] +
                     level_gather(k, nz, 0.5, 0.5, CW(4));
  const double gum = face_gather(j, ny, t.py, 0.5, 0.5, CV(3)) - gcv * t.dmdy[t.m2(i, j)] +
                     level_gather(k, nz, 0.5, 0.5, CW(3));
  const double gt = ldz(io.gtth, p) + gthin * (oma * dt);
  const double gwm = (-gt) * t.dthm[p];
  const double gthp = gthin + face_gather(i, nx, t.px, 0.5, 0.5, CU(0)) + face_gather(j, ny, t.py, 0.5, 0.5, CV(0)) +
                      level_gather(k, nz, 0.5, 0.5, CW(0));
  const double mmv = t.mm[t.m2(i, j)], dzm = t.dzm[p];
  const double g = gpin * (oma * dt) * (-t.cpi[p]);
  double *cm = io.cm;
  cm[p] = gpp;
  cm[(size_t)N + p] = ggzm;
  cm[2 * (size_t)N + p] = gum;
  cm[3 * (size_t)N + p] = gvm;
  cm[4 * (size_t)N + p] = gwm;
  if (io.hyper) {  // cotangents of del^4 of the four base fields (hyp = -nu_h del^4)
    cm[5 * (size_t)N + p] = (-t.nu_h) * face_gather(i, nx, t.px, 0.5, 0.5, CU(5));
    cm[6 * (size_t)N + p] = (-t.nu_h) * face_gather(j, ny, t.py, 0.5, 0.5, CV(5));
    cm[7 * (size_t)N + p] = (-t.nu_h) * level_gather(k, nz, 0.5, 0.5, CW(2));
    cm[8 * (size_t)N + p] = (-t.nu_h) * gt;
  }
  cm[9 * (size_t)N + p] = ((g * mmv) * (mmv * t.rdx)) / dzm;  // Exner-divergence coefficients
  cm[10 * (size_t)N + p] = ((g * mmv) * (mmv * t.rdy)) / dzm;
  cm[11 * (size_t)N + p] = (g * (t.rdz * t.dz)) / dzm;
  cm[12 * (size_t)N + p] = gthp;
}

// Transposed terrain-following Laplacian (the adjoint of hyper_faces then
// hyper_mass, with the stop-gradient edges): lapT_faces takes the mass-point
// cotangents LB of the Laplacian to the cotangents GX, GY of the face fields
// Gx, Gy; lapT_mass takes those to the cotangents FB of the input field.
__global__ void __launch_bounds__(256) lapT_faces_kernel(TendCtx t, int nf, const double *__restrict__ LB,
                                                         double *__restrict__ GX, double *__restrict__ GY) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * t.nz) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int nx = t.nx, ny = t.ny, nz = t.nz, N = t.nm(), Nu = t.nu(), Nv = t.nv();
  for (int f = 0; f < nf; ++f) {
    const double *L = LB + (size_t)f * N;
    // cotangent of the vertical difference EX at w level kk of column (ci, cj),
    // from dGz = avg_z(EX) weighted by -z_x (-z_y) at the two masses around kk
    auto ebar = [&](int ci, int cj, int kk, const double *zm) {
      if (kk < 1 || kk > nz - 1) return 0.0;
      const int q0 = t.m(ci, cj, kk - 1), q1 = t.m(ci, cj, kk);
      return 0.5 * ((-zm[q0]) * L[q0]) + 0.5 * ((-zm[q1]) * L[q1]);
    };
    // cotangent of the mass-point average Gm of column (ci, cj) at level k
    auto gmbar = [&](int ci, int cj, const double *zm) {
      return t.rdz * (ebar(ci, cj, k, zm) * t.rz[t.w(ci, cj, k)] - ebar(ci, cj, k + 1, zm) * t.rz[t.w(ci, cj, k + 1)]);
    };
    if (j < ny) {  // u face (i, j, k)
      double s = 0.0;
      if (i >= 1) s += L[t.m(i - 1, j, k)] * (t.rdx * t.mm[t.m2(i - 1, j)]);
      if (i <= nx - 1) s -= L[t.m(i, j, k)] * (t.rdx * t.mm[t.m2(i, j)]);
      if (i >= 1) s += 0.5 * gmbar(i - 1, j, t.zxm);
      if (i <= nx - 1) s += 0.5 * gmbar(i, j, t.zxm);
      GX[(size_t)f * Nu + t.u(i, j, k)] = s;
    }
    if (i < nx) {  // v face (i, j, k)
      double s = 0.0;
      if (j >= 1) s += L[t.m(i, j - 1, k)] * (t.rdy * t.mm[t.m2(i, j - 1)]);
      if (j <= ny - 1) s -= L[t.m(i, j, k)] * (t.rdy * t.mm[t.m2(i, j)]);
      if (j >= 1) s += 0.5 * gmbar(i, j - 1, t.zem);
      if (j <= ny - 1) s += 0.5 * gmbar(i, j, t.zem);
      GY[(size_t)f * Nv + t.v(i, j, k)] = s;
    }
  }
}

__global__ void __launch_bounds__(256) lapT_mass_kernel(TendCtx t, int nf, const double *__restrict__ GX,
                                                        const double *__restrict__ GY, double *__restrict__ FB) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = t.nm();
  if (p >= N) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  const int nx = t.nx, ny = t.ny, nz = t.nz, Nu = t.nu(), Nv = t.nv();
  for (int f = 0; f < nf; ++f) {
    const double *Gx = GX + (size_t)f * Nu, *Gy = GY + (size_t)f * Nv;
    double s = face_gather(i, nx, t.px, -1.0, 1.0,
                           [&](int ff) { return Gx[t.u(ff, j, k)] * (t.rdx * t.mu[t.u2(ff, j)]); }) +
               face_gather(j, ny, t.py, -1.0, 1.0,
                           [&](int ff) { return Gy[t.v(i, ff, k)] * (t.rdy * t.mv[t.v2(i, ff)]); });
    // cotangent of the mass-point vertical derivative dfz_m at level kk of this column
    auto dfzbar = [&](int kk) {
      return face_gather(i, nx, t.px, 0.5, 0.5, [&](int ff) { return (-t.zxu[t.u(ff, j, kk)]) * Gx[t.u(ff, j, kk)]; }) +
             face_gather(j, ny, t.py, 0.5, 0.5, [&](int ff) { return (-t.zev[t.v(i, ff, kk)]) * Gy[t.v(i, ff, kk)]; });
    };
    const double dc = dfzbar(k);
    const double dlo = k >= 1 ? dfzbar(k - 1) : 0.0, dhi = k + 1 <= nz - 1 ? dfzbar(k + 1) : 0.0;
    // D at w level kk (interior) carries 0.5 (dfz_m(kk - 1) + dfz_m(kk)); F(k) enters D(k) with +, D(k + 1) with -
    if (k >= 1) s += t.rdz * ((0.5 * (dlo + dc)) * t.rz[t.w(i, j, k)]);
    if (k + 1 <= nz - 1) s -= t.rdz * ((0.5 * (dc + dhi)) * t.rz[t.w(i, j, k + 1)]);
    FB[(size_t)f * N + p] = s;
  }
}

__global__ void __launch_bounds__(256) adjD_kernel(TendCtx t, ExplicitAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * (t.nz + 1)) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int nx = t.nx, ny = t.ny, nz = t.nz, N = t.nm();
  const double *U = io.u, *V = io.v, *E = io.eta, *TH = io.th;
  const double *cm = io.cm, *fb = io.fb;
  auto mget = [&](int r, int q) { return cm[(size_t)r * N + q]; };
  if (j < ny && k < nz) {  // u face (i, j, k)
    const int qu = t.u(i, j, k);
    auto gumT = [&](int c) {
      const int q = t.m(c, j, k);
      return mget(2, q) + (fb ? fb[q] : 0.0);
    };
    double s = ldz(io.gu_in, qu);
    double um = 0.0;
    if (i >= 1) um += gumT(i - 1);
    if (i <= nx - 1) um += gumT(i);
    s += 0.5 * um;
    const double cf = (i >= 1 ? mget(9, t.m(i - 1, j, k)) : 0.0) - (i <= nx - 1 ? mget(9, t.m(i, j, k)) : 0.0);
    int ia, ib;
    face_pair(i, nx, t.px, ia, ib);
    const double thpu = 0.5 * (thp(t, TH, t.m(ia, j, k)) + thp(t, TH, t.m(ib, j, k)));
    s += cf * (t.cx[qu] + thpu * t.cxp[qu]);
    io.gu[qu] = s;
    io.tu[qu] = face_frozen(i, nx, t.px) ? 0.0 : cf * (U[qu] * t.cxp[qu]);
  }
  if (i < nx && k < nz) {  // v face (i, j, k)
    const int qv = t.v(i, j, k);
    auto gvmT = [&](int c) {
      const int q = t.m(i, c, k);
      return mget(3, q) + (fb ? fb[(size_t)N + q] : 0.0);
    };
    double s = ldz(io.gv_in, qv);
    double vm = 0.0;
    if (j >= 1) vm += gvmT(j - 1);
    if (j <= ny - 1) vm += gvmT(j);
    s += 0.5 * vm;
    const double cf = (j >= 1 ? mget(10, t.m(i, j - 1, k)) : 0.0) - (j <= ny - 1 ? mget(10, t.m(i, j, k)) : 0.0);
    int ja, jb;
    face_pair(j, ny, t.py, ja, jb);
    const double thpv = 0.5 * (thp(t, TH, t.m(i, ja, k)) + thp(t, TH, t.m(i, jb, k)));
    s += cf * (t.cy[qv] + thpv * t.cyp[qv]);
    io.gv[qv] = s;
    io.tv[qv] = face_frozen(j, ny, t.py) ? 0.0 : cf * (V[qv] * t.cyp[qv]);
  }
  if (i < nx && j < ny) {  // w level (i, j, k), k = 0..nz
    const int qw = t.w(i, j, k);
    auto gwmT = [&](int c) {
      const int q = t.m(i, j, c);
      return mget(4, q) + (fb ? fb[2 * (size_t)N + q] : 0.0);
    };
    const double gr = ldz(io.gres, qw);
    double s = ldz(io.gw_in, qw) - gr;
    double wm = 0.0;
    if (k >= 1) wm += gwmT(k - 1);
    if (k <= nz - 1) wm += gwmT(k);
    io.gw[qw] = s + 0.5 * wm;
    double e = gr * t.dzw[qw];
    double tz = 0.0;
    if (k >= 1 && k <= nz - 1) {
      const double cf = mget(11, t.m(i, j, k - 1)) - mget(11, t.m(i, j, k));
      const double thpw = 0.5 * (thp(t, TH, t.m(i, j, k - 1)) + thp(t, TH, t.m(i, j, k)));
      e += cf * (t.czs[qw] + thpw * t.czp[qw]);
      tz = cf * (E[qw] * t.czp[qw]);
    }
    io.geta[qw] = e;
    io.tw[qw] = tz;
  }
}

__global__ void __launch_bounds__(256) adjE_kernel(TendCtx t, ExplicitAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = t.nm();
  if (p >= N) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const double *cm = io.cm;
  // cotangent of d pi'/dz at w level kk (interior) from the mass-point
  // averages at kk - 1 and kk, times the difference's factor rdz rz(kk)
  auto gw = [&](int kk) {
    const double avg = 0.5 * (cm[(size_t)N + t.m(i, j, kk - 1)] + cm[(size_t)N + t.m(i, j, kk)]);
    return avg * (t.rdz * t.rz[t.w(i, j, kk)]);
  };
  double s = cm[p];
  if (k >= 1) s += gw(k);
  if (k + 1 <= nz - 1) s -= gw(k + 1);
  io.gpi[p] = s;
  double th = cm[12 * (size_t)N + p] + (io.fb ? io.fb[3 * (size_t)N + p] : 0.0);
  th += face_gather(i, nx, t.px, 0.5, 0.5, [&](int f) { return io.tu[t.u(f, j, k)]; });
  th += face_gather(j, ny, t.py, 0.5, 0.5, [&](int f) { return io.tv[t.v(i, f, k)]; });
  th += level_gather(k, nz, 0.5, 0.5, [&](int kk) { return io.tw[t.w(i, j, kk)]; });
  io.gth[p] = th;
}


// ---- reverse mode of one predictor-corrector pass and of the final group ----------
//
// Predictor (steppers.py:911-968 through the reference's autodiff): from the
// cotangents of the packed right-hand side b and of theta at n+1, the
// cotangents of the fifteen inputs of sisl_predictor_f64.  The kinematic
// bottom row of b is built from stop-gradient edge values and the top row is
// zero, so neither passes anything back.  Passes: predA (faces, w levels),
// predB (mass), predD (faces, w levels), predE (mass), predF (w levels).
struct PredictorAdjIO {
  const double *g_pi, *g_u, *g_v, *g_eta, *m_u, *m_v;  // saved inputs the derivatives need
  const double *th_next;                               // saved output
  const double *gb, *gth;                              // cotangents of b, th_next (null: zero)
  double *cu, *cv, *cw;  // predA: 5 u-face, 5 v-face, 2 w-level arrays
  double *cm;            // predB: tp, gp, ggzm, gum, gvm, cx, cy, cz
  double *tu, *tv, *tw;  // predD: theta' flux cotangents
  // input cotangents (sisl_predictor_f64 order)
  double *grhs_u, *grhs_v, *grhs_w, *grhs_pi, *gth_adv, *gtth, *gw_n, *greta, *gg_pi, *gg_u, *gg_v, *gg_w,
      *gg_eta, *gm_u, *gm_v;
};

__global__ void __launch_bounds__(256) predA_kernel(TendCtx t, PredictorAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * (t.nz + 1)) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const int Nu = t.nu(), Nv = t.nv(), Nw = t.nw();
  const double adt = t.alpha * t.dt, cad = (t.cp * t.alpha) * t.dt;
  const double *TN = io.th_next;
  auto tp1 = [&](int q) { return TN[q] - t.th_bg[q]; };
  if (j < ny && k < nz) {  // u face
    const int qu = t.u(i, j, k);
    double c[5] = {0, 0, 0, 0, 0};
    const double gb = ldz(io.gb, qu);
    io.grhs_u[qu] = gb;
    if (!face_frozen(i, nx, t.px) && gb != 0.0) {
      int ia, ib;
      face_pair(i, nx, t.px, ia, ib);
      const double thu1 = 0.5 * (tp1(t.m(ia, j, k)) + tp1(t.m(ib, j, k)));
      const double gx = ((io.g_pi[t.m(ib, j, k)] - io.g_pi[t.m(ia, j, k)]) * t.rdx) * t.mu[t.u2(i, j)];
      const double gzu = 0.5 * (gzm_p(t, io.g_pi, ia, j, k) + gzm_p(t, io.g_pi, ib, j, k));
      const double vau = 0.5 * (vm_at(t, io.m_v, ia, j, k) + vm_at(t, io.m_v, ib, j, k));
      const double cvu = 0.5 * (curv_m(t, io.m_u, io.m_v, ia, j, k) + curv_m(t, io.m_u, io.m_v, ib, j, k));
      const double gc = gb * ((-cad) * thu1);
      c[0] = gb * ((-cad) * (gx - t.zxu[qu] * gzu));  // -> theta' at n+1 (average)
      c[1] = gc * (t.rdx * t.mu[t.u2(i, j)]);         // -> guess pi' (difference)
      c[2] = gc * (-t.zxu[qu]);                       // -> dpi/dz at mass points (average)
      c[3] = (gb * adt) * cvu;                        // -> v_m of the metric winds (average)
      c[4] = (gb * adt) * vau;                        // -> curvature term (average)
    }
#pragma unroll
    for (int r = 0; r < 5; ++r) io.cu[(size_t)r * Nu + qu] = c[r];
  }
  if (i < nx && k < nz) {  // v face
    const int qv = t.v(i, j, k);
    double c[5] = {0, 0, 0, 0, 0};
    const double gb = ldz(io.gb, Nu + qv);
    io.grhs_v[qv] = gb;
    if (!face_frozen(j, ny, t.py) && gb != 0.0) {
      int ja, jb;
      face_pair(j, ny, t.py, ja, jb);
      const double thv1 = 0.5 * (tp1(t.m(i, ja, k)) + tp1(t.m(i, jb, k)));
      const double gy = ((io.g_pi[t.m(i, jb, k)] - io.g_pi[t.m(i, ja, k)]) * t.rdy) * t.mv[t.v2(i, j)];
      const double gzv = 0.5 * (gzm_p(t, io.g_pi, i, ja, k) + gzm_p(t, io.g_pi, i, jb, k));
      const double uav = 0.5 * (um_at(t, io.m_u, i, ja, k) + um_at(t, io.m_u, i, jb, k));
      const double cvv = 0.5 * (curv_m(t, io.m_u, io.m_v, i, ja, k) + curv_m(t, io.m_u, io.m_v, i, jb, k));
      const double gc = gb * ((-cad) * thv1);
      c[0] = gb * ((-cad) * (gy - t.zev[qv] * gzv));
      c[1] = gc * (t.rdy * t.mv[t.v2(i, j)]);
      c[2] = gc * (-t.zev[qv]);
      c[3] = (gb * adt) * (-cvv);  // -> u_m (average)
      c[4] = (gb * adt) * (-uav);
    }
#pragma unroll
    for (int r = 0; r < 5; ++r) io.cv[(size_t)r * Nv + qv] = c[r];
  }
  if (i < nx && j < ny) {  // w level
    const int qw = t.w(i, j, k);
    double c[2] = {0, 0};
    const bool in = k >= 1 && k <= nz - 1;
    const double gb = in ? ldz(io.gb, Nu + Nv + qw) : 0.0;
    io.grhs_w[qw] = gb;
    io.greta[qw] = in ? (-(1.0 - t.alpha)) * ldz(io.gb, Nu + Nv + Nw + t.nm() + qw) : 0.0;
    if (in && gb != 0.0) {
      const double thw1 = 0.5 * (tp1(t.m(i, j, k - 1)) + tp1(t.m(i, j, k)));
      const double gzw = gzw_p(t, io.g_pi, i, j, k);
      c[0] = gb * (adt * (t.grav / t.thw_bg[qw]) + (-cad) * gzw);  // -> theta' (average)
      c[1] = gb * ((-cad) * thw1) * (t.rdz * t.rz[qw]);             // -> guess pi' (difference)
    }
    io.cw[qw] = c[0];
    io.cw[(size_t)Nw + qw] = c[1];
  }
}

__global__ void __launch_bounds__(256) predB_kernel(TendCtx t, PredictorAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = t.nm();
  if (p >= N) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const int Nu = t.nu(), Nv = t.nv(), Nw = t.nw();
  const double adt = t.alpha * t.dt;
  auto CU = [&](int r) { return [&, r](int f) { return io.cu[(size_t)r * Nu + t.u(f, j, k)]; }; };
  auto CV = [&](int r) { return [&, r](int f) { return io.cv[(size_t)r * Nv + t.v(i, f, k)]; }; };
  auto CW = [&](int r) { return [&, r](int kk) { return io.cw[(size_t)r * Nw + t.w(i, j, kk)]; }; };
  double *cm = io.cm;
  cm[p] = ldz(io.gth, p) + face_gather(i, nx, t.px, 0.5, 0.5, CU(0)) + face_gather(j, ny, t.py, 0.5, 0.5, CV(0)) +
          level_gather(k, nz, 0.5, 0.5, CW(0));
  cm[(size_t)N + p] = face_gather(i, nx, t.px, -1.0, 1.0, CU(1)) + face_gather(j, ny, t.py, -1.0, 1.0, CV(1)) +
                      level_gather(k, nz, -1.0, 1.0, CW(1));
  cm[2 * (size_t)N + p] = face_gather(i, nx, t.px, 0.5, 0.5, CU(2)) + face_gather(j, ny, t.py, 0.5, 0.5, CV(2));
  const double gcv = face_gather(i, nx, t.px, 0.5, 0.5, CU(4)) + face_gather(j, ny, t.py, 0.5, 0.5, CV(4));
  cm[3 * (size_t)N + p] = face_gather(j, ny, t.py, 0.5, 0.5, CV(3)) - gcv * t.dmdy[t.m2(i, j)];
  cm[4 * (size_t)N + p] = face_gather(i, nx, t.px, 0.5, 0.5, CU(3)) + gcv * t.dmdx[t.m2(i, j)];
  const double gpb = ldz(io.gb, Nu + Nv + Nw + p);
  io.grhs_pi[p] = gpb * 1000.0;
  const double mmv = t.mm[t.m2(i, j)], dzm = t.dzm[p];
  const double g = ((gpb * 1000.0) * adt) * (-t.cpi[p]);
  cm[5 * (size_t)N + p] = ((g * mmv) * (mmv * t.rdx)) / dzm;
  cm[6 * (size_t)N + p] = ((g * mmv) * (mmv * t.rdy)) / dzm;
  cm[7 * (size_t)N + p] = (g * (t.rdz * t.dz)) / dzm;
}

__global__ void __launch_bounds__(256) predD_kernel(TendCtx t, PredictorAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * (t.nz + 1)) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int nx = t.nx, ny = t.ny, nz = t.nz, N = t.nm();
  const double *cm = io.cm, *TN = io.th_next;
  auto tp1 = [&](int q) { return TN[q] - t.th_bg[q]; };
  auto mget = [&](int r, int q) { return cm[(size_t)r * N + q]; };
  if (j < ny && k < nz) {  // u face
    const int qu = t.u(i, j, k);
    double um = 0.0;
    if (i >= 1) um += mget(3, t.m(i - 1, j, k));
    if (i <= nx - 1) um += mget(3, t.m(i, j, k));
    io.gm_u[qu] = 0.5 * um;
    const double cf = (i >= 1 ? mget(5, t.m(i - 1, j, k)) : 0.0) - (i <= nx - 1 ? mget(5, t.m(i, j, k)) : 0.0);
    int ia, ib;
    face_pair(i, nx, t.px, ia, ib);
    const double thu1 = 0.5 * (tp1(t.m(ia, j, k)) + tp1(t.m(ib, j, k)));
    io.gg_u[qu] = cf * (thu1 * t.cxp[qu]);
    io.tu[qu] = face_frozen(i, nx, t.px) ? 0.0 : cf * (io.g_u[qu] * t.cxp[qu]);
  }
  if (i < nx && k < nz) {  // v face
    const int qv = t.v(i, j, k);
    double vm = 0.0;
    if (j >= 1) vm += mget(4, t.m(i, j - 1, k));
    if (j <= ny - 1) vm += mget(4, t.m(i, j, k));
    io.gm_v[qv] = 0.5 * vm;
    const double cf = (j >= 1 ? mget(6, t.m(i, j - 1, k)) : 0.0) - (j <= ny - 1 ? mget(6, t.m(i, j, k)) : 0.0);
    int ja, jb;
    face_pair(j, ny, t.py, ja, jb);
    const double thv1 = 0.5 * (tp1(t.m(i, ja, k)) + tp1(t.m(i, jb, k)));
    io.gg_v[qv] = cf * (thv1 * t.cyp[qv]);
    io.tv[qv] = face_frozen(j, ny, t.py) ? 0.0 : cf * (io.g_v[qv] * t.cyp[qv]);
  }
  if (i < nx && j < ny) {  // w level
    const int qw = t.w(i, j, k);
    double ge = 0.0, tz = 0.0;
    if (k >= 1 && k <= nz - 1) {
      const double cf = mget(7, t.m(i, j, k - 1)) - mget(7, t.m(i, j, k));
      const double thw1 = 0.5 * (tp1(t.m(i, j, k - 1)) + tp1(t.m(i, j, k)));
      ge = cf * (thw1 * t.czp[qw]);
      tz = cf * (io.g_eta[qw] * t.czp[qw]);
    }
    io.gg_eta[qw] = ge;
    io.tw[qw] = tz;
  }
}

__global__ void __launch_bounds__(256) predE_kernel(TendCtx t, PredictorAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = t.nm();
  if (p >= N) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const double *cm = io.cm;
  double th = cm[p];
  th += face_gather(i, nx, t.px, 0.5, 0.5, [&](int f) { return io.tu[t.u(f, j, k)]; });
  th += face_gather(j, ny, t.py, 0.5, 0.5, [&](int f) { return io.tv[t.v(i, f, k)]; });
  th += level_gather(k, nz, 0.5, 0.5, [&](int kk) { return io.tw[t.w(i, j, kk)]; });
  io.gth_adv[p] = th;
  io.gtth[p] = (t.alpha * t.dt) * th;
  auto gw = [&](int kk) {
    const double avg = 0.5 * (cm[2 * (size_t)N + t.m(i, j, kk - 1)] + cm[2 * (size_t)N + t.m(i, j, kk)]);
    return avg * (t.rdz * t.rz[t.w(i, j, kk)]);
  };
  double s = cm[(size_t)N + p];
  if (k >= 1) s += gw(k);
  if (k + 1 <= nz - 1) s -= gw(k + 1);
  io.gg_pi[p] = s;
}

// w_n and the guess's w enter theta at n+1 through their mass-point averages
__global__ void __launch_bounds__(256) predF_kernel(TendCtx t, PredictorAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= t.nw()) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  const double adt = t.alpha * t.dt;
  auto T = [&](int c) {
    const int q = t.m(i, j, c);
    return (adt * io.gth_adv[q]) * t.dthm[q];
  };
  double s = 0.0;
  if (k >= 1) s += T(k - 1);
  if (k <= t.nz - 1) s += T(k);
  io.gw_n[p] = 0.5 * s;
  io.gg_w[p] = -(0.5 * s);
}

// Final group (steppers.py:1003-1035): from the cotangents of u, v (after
// divergence damping), pi and eta, the cotangent of the packed solution x.
struct FinalAdjIO {
  const double *gu, *gv, *gpi, *geta;  // null: zero
  double *gdiv;                        // mass scratch
  double *gx;                          // packed cotangent
};

__global__ void __launch_bounds__(256) finA_kernel(TendCtx t, FinalAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = t.nm();
  if (p >= N) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  const double kd = t.nu_div * t.dt;
  const double s =
      face_gather(i, t.nx, t.px, -1.0, 1.0,
                  [&](int f) { return (kd * ldz(io.gu, t.u(f, j, k))) * (t.rdx * t.mu[t.u2(f, j)]); }) +
      face_gather(j, t.ny, t.py, -1.0, 1.0,
                  [&](int f) { return (kd * ldz(io.gv, t.v(i, f, k))) * (t.rdy * t.mv[t.v2(i, f)]); });
  io.gdiv[p] = s;
}

__global__ void __launch_bounds__(256) finB_kernel(TendCtx t, FinalAdjIO io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * (t.nz + 1)) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const int Nu = t.nu(), Nv = t.nv(), Nw = t.nw(), N = t.nm();
  if (j < ny && k < nz) {
    const int qu = t.u(i, j, k);
    double s = ldz(io.gu, qu);
    if (i >= 1) s += io.gdiv[t.m(i - 1, j, k)] * (t.rdx * t.mm[t.m2(i - 1, j)]);
    if (i <= nx - 1) s -= io.gdiv[t.m(i, j, k)] * (t.rdx * t.mm[t.m2(i, j)]);
    io.gx[qu] = s;
  }
  if (i < nx && k < nz) {
    const int qv = t.v(i, j, k);
    double s = ldz(io.gv, qv);
    if (j >= 1) s += io.gdiv[t.m(i, j - 1, k)] * (t.rdy * t.mm[t.m2(i, j - 1)]);
    if (j <= ny - 1) s -= io.gdiv[t.m(i, j, k)] * (t.rdy * t.mm[t.m2(i, j)]);
    io.gx[Nu + qv] = s;
  }
  if (i < nx && j < ny) {
    const int qw = t.w(i, j, k);
    io.gx[Nu + Nv + qw] = 0.0;  // w of the solution leaves through its own path
    io.gx[Nu + Nv + Nw + N + qw] = ldz(io.geta, qw) / t.dzw[qw];
    if (k < nz) {
      const int q = t.m(i, j, k);
      io.gx[Nu + Nv + Nw + q] = ldz(io.gpi, q) / 1000.0;
    }
  }
}

}  // namespace sisl

extern "C" {

int64_t sisl_hyperdiffusion_workspace_f64(const SislGrid *g) {
  if (!g) return -1;
  const int64_t N = (int64_t)g->nx * g->ny * g->nz;
  const int64_t Nu = (int64_t)(g->nx + 1) * g->ny * g->nz, Nv = (int64_t)g->nx * (g->ny + 1) * g->nz;
  return 4 * (2 * N + Nu + Nv);  // elements: base fields, first Laplacian, Gx, Gy
}

int sisl_hyperdiffusion_f64(const SislTendency *d, const double *u, const double *v, const double *w,
                            const double *th, double *work, double *lap2, void *stream) {
  sisl::TendCtx t;
  if (!sisl::make_ctx(d, t) || !u || !v || !w || !th || !work || !lap2) return SISL_ERR_ARG;
  hipStream_t s = sisl::as_stream(stream);
  const int64_t N = (int64_t)t.nx * t.ny * t.nz;
  const int64_t Nu = (int64_t)(t.nx + 1) * t.ny * t.nz;
  double *F = work, *L1 = work + 4 * N, *GX = work + 8 * N, *GY = GX + 4 * Nu;
  const unsigned bm = sisl::nblocks(N), bf = sisl::nblocks((int64_t)(t.nx + 1) * (t.ny + 1) * t.nz);
  hipLaunchKernelGGL(sisl::hyper_base_kernel, dim3(bm), dim3(sisl::kBlock), 0, s, t, u, v, w, th, F);
  hipLaunchKernelGGL(sisl::hyper_faces_kernel, dim3(bf), dim3(sisl::kBlock), 0, s, t, 4, (const double *)F, GX, GY);
  hipLaunchKernelGGL(sisl::hyper_mass_kernel, dim3(bm), dim3(sisl::kBlock), 0, s, t, 4, (const double *)GX,
                     (const double *)GY, L1);
  hipLaunchKernelGGL(sisl::hyper_faces_kernel, dim3(bf), dim3(sisl::kBlock), 0, s, t, 4, (const double *)L1, GX, GY);
  hipLaunchKernelGGL(sisl::hyper_mass_kernel, dim3(bm), dim3(sisl::kBlock), 0, s, t, 4, (const double *)GX,
                     (const double *)GY, lap2);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

int sisl_explicit_f64(const SislTendency *d, const double *u, const double *v, const double *w, const double *pi,
                      const double *th, const double *eta, const double *lap2, double *u_in, double *v_in,
                      double *w_in, double *pi_in, double *th_in, double *res, double *tth, double *x0,
                      double *tend, void *stream) {
  sisl::TendCtx t;
  if (!sisl::make_ctx(d, t) || !u || !v || !w || !pi || !th || !eta || !u_in || !v_in || !w_in || !pi_in ||
      !th_in || !res || !tth)
    return SISL_ERR_ARG;
  if (!lap2 && t.nu_h != 0.0) return SISL_ERR_ARG;
  sisl::ExplicitIO io{u, v, w, pi, th, eta, t.nu_h != 0.0 ? lap2 : nullptr, u_in, v_in, w_in, pi_in, th_in,
                      res, tth, x0, tend};
  const int64_t n = (int64_t)(t.nx + 1) * (t.ny + 1) * (t.nz + 1);
  hipLaunchKernelGGL(sisl::explicit_kernel, dim3(sisl::nblocks(n)), dim3(sisl::kBlock), 0, sisl::as_stream(stream),
                     t, io);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

int sisl_slow_tendencies_f64(const SislTendency *d, const double *u, const double *v, const double *w,
                             const double *pi, const double *th, const double *eta, const double *lap2s,
                             const double *lap20, double eps_l, double *adv, double *tend, void *stream) {
  sisl::TendCtx t;
  if (!sisl::make_ctx(d, t) || !u || !v || !w || !pi || !th || !eta || !adv || !tend) return SISL_ERR_ARG;
  if ((lap2s || lap20) && t.nu_h == 0.0) return SISL_ERR_ARG;
  hipStream_t s = sisl::as_stream(stream);
  const int64_t N = (int64_t)t.nx * t.ny * t.nz;
  const int64_t span = (int64_t)(t.nx + 1) * (t.ny + 1) * (t.nz + 1);
  auto xgrid = [](int64_t n) { return dim3((sisl::nblocks(n) + sisl::kXcds - 1) / sisl::kXcds * sisl::kXcds); };
  sisl::AdvIO a{u, v, w, th, eta, eps_l, adv};
  hipLaunchKernelGGL(sisl::adv_kernel, xgrid(N), dim3(sisl::kBlock), 0, s, t, a);
  sisl::SlowIO io{u, v, w, pi, th, eta, lap2s, lap20, adv, tend};
  hipLaunchKernelGGL(sisl::slow_kernel, xgrid(span), dim3(sisl::kBlock), 0, s, t, io);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

int sisl_predictor_f64(const SislTendency *d, const double *rhs_u, const double *rhs_v, const double *rhs_w,
                       const double *rhs_pi, const double *th_adv, const double *tth, const double *w_n,
                       const double *reta, const double *g_pi, const double *g_u, const double *g_v,
                       const double *g_w, const double *g_eta, const double *m_u, const double *m_v,
                       double *th_next, double *b, void *stream) {
  sisl::TendCtx t;
  if (!sisl::make_ctx(d, t)) return SISL_ERR_ARG;
  const void *req[] = {rhs_u, rhs_v, rhs_w, rhs_pi, th_adv, tth, w_n, reta, g_pi, g_u, g_v, g_w, g_eta, m_u, m_v,
                       th_next, b};
  for (const void *q : req)
    if (!q) return SISL_ERR_ARG;
  sisl::PredictorIO io{rhs_u, rhs_v, rhs_w, rhs_pi, th_adv, tth, w_n, reta, g_pi, g_u, g_v, g_w, g_eta,
                       m_u, m_v, th_next, b};
  hipStream_t s = sisl::as_stream(stream);
  const int64_t N = (int64_t)t.nx * t.ny * t.nz;
  const int64_t n = (int64_t)(t.nx + 1) * (t.ny + 1) * (t.nz + 1);
  hipLaunchKernelGGL(sisl::thnext_kernel, dim3(sisl::nblocks(N)), dim3(sisl::kBlock), 0, s, t, io);
  hipLaunchKernelGGL(sisl::rhs_kernel, dim3(sisl::nblocks(n)), dim3(sisl::kBlock), 0, s, t, io);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

int sisl_final_f64(const SislTendency *d, const double *x, double *u, double *v, double *pi, double *eta,
                   void *stream) {
  sisl::TendCtx t;
  if (!sisl::make_ctx(d, t) || !x || !u || !v || !pi || !eta) return SISL_ERR_ARG;
  sisl::FinalIO io{x, u, v, pi, eta};
  const int64_t n = (int64_t)(t.nx + 1) * (t.ny + 1) * (t.nz + 1);
  hipLaunchKernelGGL(sisl::final_kernel, dim3(sisl::nblocks(n)), dim3(sisl::kBlock), 0, sisl::as_stream(stream),
                     t, io);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

int64_t sisl_explicit_adj_workspace_f64(const SislGrid *g) {
  if (!g) return -1;
  const int64_t N = (int64_t)g->nx * g->ny * g->nz;
  const int64_t Nu = (int64_t)(g->nx + 1) * g->ny * g->nz, Nv = (int64_t)g->nx * (g->ny + 1) * g->nz;
  const int64_t Nw = (int64_t)g->nx * g->ny * (g->nz + 1);
  // face / w-level intermediates (also the transposed Laplacian's face fields),
  // 13 mass arrays, two 4-field mass arrays of the transposed del^4, theta' fluxes
  return 6 * Nu + 6 * Nv + 5 * Nw + 13 * N + 8 * N + Nu + Nv + Nw;
}

int sisl_explicit_adj_f64(const SislTendency *d, const double *u, const double *v, const double *w,
                          const double *pi, const double *th, const double *eta, const double *gu_in,
                          const double *gv_in, const double *gw_in, const double *gpi_in, const double *gth_in,
                          const double *gres, const double *gtth, double *gu, double *gv, double *gw, double *gpi,
                          double *gth, double *geta, double *work, void *stream) {
  sisl::TendCtx t;
  if (!sisl::make_ctx(d, t) || !u || !v || !w || !pi || !th || !eta || !gu || !gv || !gw || !gpi || !gth ||
      !geta || !work)
    return SISL_ERR_ARG;
  hipStream_t s = sisl::as_stream(stream);
  const int64_t N = (int64_t)t.nx * t.ny * t.nz;
  const int64_t Nu = (int64_t)(t.nx + 1) * t.ny * t.nz, Nv = (int64_t)t.nx * (t.ny + 1) * t.nz;
  const int64_t Nw = (int64_t)t.nx * t.ny * (t.nz + 1);
  double *cu = work, *cv = cu + 6 * Nu, *cw = cv + 6 * Nv, *cm = cw + 5 * Nw;
  double *l1 = cm + 13 * N, *fb = l1 + 4 * N, *tu = fb + 4 * N, *tv = tu + Nu, *tw = tv + Nv;
  const int hyper = t.nu_h != 0.0;
  sisl::ExplicitAdjIO io{u, v, w, pi, th, eta, gu_in, gv_in, gw_in, gpi_in, gth_in, gres, gtth, hyper,
                         cu, cv, cw, cm, hyper ? fb : nullptr, tu, tv, tw, gu, gv, gw, gpi, gth, geta};
  const int64_t span = (int64_t)(t.nx + 1) * (t.ny + 1) * (t.nz + 1);
  const unsigned bs = sisl::nblocks(span), bm = sisl::nblocks(N);
  const unsigned bf = sisl::nblocks((int64_t)(t.nx + 1) * (t.ny + 1) * t.nz);
  hipLaunchKernelGGL(sisl::adjA_kernel, dim3(bs), dim3(sisl::kBlock), 0, s, t, io);
  hipLaunchKernelGGL(sisl::adjB_kernel, dim3(bm), dim3(sisl::kBlock), 0, s, t, io);
  if (hyper) {  // del^4 transposed: the face arrays of the first two passes are free again
    const double *lb = cm + 5 * N;
    hipLaunchKernelGGL(sisl::lapT_faces_kernel, dim3(bf), dim3(sisl::kBlock), 0, s, t, 4, lb, cu, cv);
    hipLaunchKernelGGL(sisl::lapT_mass_kernel, dim3(bm), dim3(sisl::kBlock), 0, s, t, 4, (const double *)cu,
                       (const double *)cv, l1);
    hipLaunchKernelGGL(sisl::lapT_faces_kernel, dim3(bf), dim3(sisl::kBlock), 0, s, t, 4, (const double *)l1, cu,
                       cv);
    hipLaunchKernelGGL(sisl::lapT_mass_kernel, dim3(bm), dim3(sisl::kBlock), 0, s, t, 4, (const double *)cu,
                       (const double *)cv, fb);
  }
  hipLaunchKernelGGL(sisl::adjD_kernel, dim3(bs), dim3(sisl::kBlock), 0, s, t, io);
  hipLaunchKernelGGL(sisl::adjE_kernel, dim3(bm), dim3(sisl::kBlock), 0, s, t, io);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

int64_t sisl_predictor_adj_workspace_f64(const SislGrid *g) {
  if (!g) return -1;
  const int64_t N = (int64_t)g->nx * g->ny * g->nz;
  const int64_t Nu = (int64_t)(g->nx + 1) * g->ny * g->nz, Nv = (int64_t)g->nx * (g->ny + 1) * g->nz;
  const int64_t Nw = (int64_t)g->nx * g->ny * (g->nz + 1);
  return 5 * Nu + 5 * Nv + 2 * Nw + 8 * N + Nu + Nv + Nw;
}

int sisl_predictor_adj_f64(const SislTendency *d, const double *g_pi, const double *g_u, const double *g_v,
                           const double *g_eta, const double *m_u, const double *m_v, const double *th_next,
                           const double *gb, const double *gth, double *grhs_u, double *grhs_v, double *grhs_w,
                           double *grhs_pi, double *gth_adv, double *gtth, double *gw_n, double *greta,
                           double *gg_pi, double *gg_u, double *gg_v, double *gg_w, double *gg_eta, double *gm_u,
                           double *gm_v, double *work, void *stream) {
  sisl::TendCtx t;
  if (!sisl::make_ctx(d, t)) return SISL_ERR_ARG;
  const void *req[] = {g_pi, g_u, g_v, g_eta, m_u, m_v, th_next, grhs_u, grhs_v, grhs_w, grhs_pi, gth_adv,
                       gtth, gw_n, greta, gg_pi, gg_u, gg_v, gg_w, gg_eta, gm_u, gm_v, work};
  for (const void *q : req)
    if (!q) return SISL_ERR_ARG;
  hipStream_t s = sisl::as_stream(stream);
  const int64_t N = (int64_t)t.nx * t.ny * t.nz;
  const int64_t Nu = (int64_t)(t.nx + 1) * t.ny * t.nz, Nv = (int64_t)t.nx * (t.ny + 1) * t.nz;
  const int64_t Nw = (int64_t)t.nx * t.ny * (t.nz + 1);
  double *cu = work, *cv = cu + 5 * Nu, *cw = cv + 5 * Nv, *cm = cw + 2 * Nw;
  double *tu = cm + 8 * N, *tv = tu + Nu, *tw = tv + Nv;
  sisl::PredictorAdjIO io{g_pi, g_u, g_v, g_eta, m_u, m_v, th_next, gb, gth, cu, cv, cw, cm, tu, tv, tw,
                          grhs_u, grhs_v, grhs_w, grhs_pi, gth_adv, gtth, gw_n, greta, gg_pi, gg_u, gg_v, gg_w,
                          gg_eta, gm_u, gm_v};
  const int64_t span = (int64_t)(t.nx + 1) * (t.ny + 1) * (t.nz + 1);
  const unsigned bs = sisl::nblocks(span), bm = sisl::nblocks(N), bw = sisl::nblocks(Nw);
  hipLaunchKernelGGL(sisl::predA_kernel, dim3(bs), dim3(sisl::kBlock), 0, s, t, io);
  hipLaunchKernelGGL(sisl::predB_kernel, dim3(bm), dim3(sisl::kBlock), 0, s, t, io);
  hipLaunchKernelGGL(sisl::predD_kernel, dim3(bs), dim3(sisl::kBlock), 0, s, t, io);
  hipLaunchKernelGGL(sisl::predE_kernel, dim3(bm), dim3(sisl::kBlock), 0, s, t, io);
  hipLaunchKernelGGL(sisl::predF_kernel, dim3(bw), dim3(sisl::kBlock), 0, s, t, io);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

int sisl_final_adj_f64(const SislTendency *d, const double *gu, const double *gv, const double *gpi,
                       const double *geta, double *gx, double *work, void *stream) {
  sisl::TendCtx t;
  if (!sisl::make_ctx(d, t) || !gx || !work) return SISL_ERR_ARG;
  hipStream_t s = sisl::as_stream(stream);
  const int64_t N = (int64_t)t.nx * t.ny * t.nz;
  const int64_t span = (int64_t)(t.nx + 1) * (t.ny + 1) * (t.nz + 1);
  sisl::FinalAdjIO io{gu, gv, gpi, geta, work, gx};
  hipLaunchKernelGGL(sisl::finA_kernel, dim3(sisl::nblocks(N)), dim3(sisl::kBlock), 0, s, t, io);
  hipLaunchKernelGGL(sisl::finB_kernel, dim3(sisl::nblocks(span)), dim3(sisl::kBlock), 0, s, t, io);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

}  // extern "C"


