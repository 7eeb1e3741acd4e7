// Acoustic substeps of the split-explicit core (FP64).
//
// The reference integrates the deviation (u'', v'', w'', pi'') of each RK3
// stage from its stage state over ns/3, ns/2 and ns substeps
// (suetes/regional3d/steppers.py:1472-1737).  One substep here is four
// launches:
//
//   predictor  (one thread per u / v face)  pressure gradient of pi'' with
//              its terrain-slope part, divergence damping of the total wind,
//              the half-step winds and the lateral boundary faces;
//   column_rhs (one thread per mass point)  the explicit w, its off-centred
//              average with the kinematic surface value, the vertical and
//              horizontal mass-flux divergences and the right-hand side of
//              the implicit pi column;
//   column     (one thread per column)      the Thomas sweep with the
//              precomputed factors of the column matrix (one set per
//              substep length), pi'' += dpi in place;
//   corrector  (one thread per u / v face and w level)  the averaged
//              pressure gradient and damped half-step divergence for the
//              winds, the implicit w update and Rayleigh damping of w.
//
// Staggered operators follow CGridOperator3D (operators.py:37-132): a face
// value between mass points a and b is their difference / average; at a
// closed boundary the first / last interior face value is repeated, at a
// periodic one the face between the last and the first cell is used at
// both ends; horizontal differences carry the map factor of the target
// location.  Threads read their neighbours in the internal x-fastest layout;
// the block grids are padded to a multiple of the 8 XCDs and renumbered so
// that each XCD runs one contiguous slab of levels, and the rows and levels
// a thread reads around its own are then in its XCD's L2.
#include "common.h"

namespace sisl {

struct AcCtx {
  int nx, ny, nz, px, py;
  double rdx, rdy, rdz, dz, cp, alpha, nu_div, dtau;
  const double *mu, *mv, *mm, *mw, *thu, *thv, *thw, *zxu, *zev, *rz, *dzm, *cpi, *cx, *cy, *rtw;
  const double *zxiw, *zetaw, *tau;

  __device__ __forceinline__ int m(int i, int j, int k) const { return (k * ny + j) * nx + i; }
  __device__ __forceinline__ int u(int i, int j, int k) const { return (k * ny + j) * (nx + 1) + i; }
  __device__ __forceinline__ int v(int i, int j, int k) const { return (k * (ny + 1) + j) * nx + i; }
  __device__ __forceinline__ int w(int i, int j, int k) const { return (k * ny + j) * nx + i; }
  __device__ __forceinline__ int m2(int i, int j) const { return j * nx + i; }
  __device__ __forceinline__ int u2(int i, int j) const { return j * (nx + 1) + i; }
  __device__ __forceinline__ int v2(int i, int j) const { return j * nx + i; }
};

// Mass points (a, b) whose difference / average gives face f (0..n).
__device__ __forceinline__ void ac_face_pair(int f, int n, int per, int &a, int &b) {
  if (per && (f == 0 || f == n)) {
    a = n - 1;
    b = 0;
    return;
  }
  const int g = min(max(f, 1), n - 1);
  a = g - 1;
  b = g;
}
// The mass level above the pair whose difference / average gives w level kk.
__device__ __forceinline__ int ac_wlev(int kk, int nz) { return min(max(kk, 1), nz - 1); }

// d/dz of a mass field at w level kk, times dz / dz_w.
__device__ __forceinline__ double ac_dzw(const AcCtx &t, const double *F, int i, int j, int kk) {
  const int g = ac_wlev(kk, t.nz);
  return ((F[t.m(i, j, g)] - F[t.m(i, j, g - 1)]) * t.rdz) * t.rz[t.w(i, j, kk)];
}
// The vertical derivative averaged to mass level k.
__device__ __forceinline__ double ac_dzm(const AcCtx &t, const double *F, int i, int j, int k) {
  return 0.5 * (ac_dzw(t, F, i, j, k) + ac_dzw(t, F, i, j, k + 1));
}
// Horizontal divergence of the total wind (UC + U, VC + V) at mass point (c, j, k).
__device__ __forceinline__ double ac_div(const AcCtx &t, const double *UC, const double *U, const double *VC,
                                         const double *V, int c, int j, int k) {
  const double mmc = t.mm[t.m2(c, j)];
  const double dxu = (((UC[t.u(c + 1, j, k)] + U[t.u(c + 1, j, k)]) - (UC[t.u(c, j, k)] + U[t.u(c, j, k)])) *
                      t.rdx) * mmc;
  const double dyv = (((VC[t.v(c, j + 1, k)] + V[t.v(c, j + 1, k)]) - (VC[t.v(c, j, k)] + V[t.v(c, j, k)])) *
                      t.rdy) * mmc;
  return dxu + dyv;
}
// Pressure gradient at u face (i, j, k) between mass points a, b: dp/dx - z_x dp/dz.
__device__ __forceinline__ double ac_gx(const AcCtx &t, const double *P, int ie, int a, int b, int j, int k) {
  return ((P[t.m(b, j, k)] - P[t.m(a, j, k)]) * t.rdx) * t.mu[t.u2(ie, j)] -
         t.zxu[t.u(ie, j, k)] * (0.5 * (ac_dzm(t, P, a, j, k) + ac_dzm(t, P, b, j, k)));
}
__device__ __forceinline__ double ac_gy(const AcCtx &t, const double *P, int je, int a, int b, int i, int k) {
  return ((P[t.m(i, b, k)] - P[t.m(i, a, k)]) * t.rdy) * t.mv[t.v2(i, je)] -
         t.zev[t.v(i, je, k)] * (0.5 * (ac_dzm(t, P, i, a, k) + ac_dzm(t, P, i, b, k)));
}

// ---- predictor (steppers.py:1575-1603) ------------------------------------------
__global__ void __launch_bounds__(256) ac_predictor_kernel(
    AcCtx t, const double *__restrict__ P, const double *__restrict__ U, const double *__restrict__ V,
    const double *__restrict__ UC, const double *__restrict__ VC, const double *__restrict__ SU,
    const double *__restrict__ SV, const double *__restrict__ UB, const double *__restrict__ VB,
    double *__restrict__ UH, double *__restrict__ VH, double *__restrict__ GX, double *__restrict__ GY) {
  const int p = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * t.nz) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  if (j < t.ny) {
    // a periodic axis carries face 0 twice: the last face takes face 0's value
    const int ie = (t.px && i == t.nx) ? 0 : i;
    int a, b;
    ac_face_pair(ie, t.nx, t.px, a, b);
    const int f = t.u(ie, j, k);
    const double gox = ac_gx(t, P, ie, a, b, j, k);
    double ru = SU[f] - t.cp * t.thu[f] * gox;
    if (t.nu_div != 0.0) {
      const double dox = ((ac_div(t, UC, U, VC, V, b, j, k) - ac_div(t, UC, U, VC, V, a, j, k)) * t.rdx) *
                         t.mu[t.u2(ie, j)];
      ru = ru + t.nu_div * dox;
    }
    double uh = U[f] + 0.5 * t.dtau * ru;
    if (!t.px && (i == 0 || i == t.nx)) uh = UB[t.u(i, j, k)];
    UH[t.u(i, j, k)] = uh;
    GX[t.u(i, j, k)] = gox;
  }
  if (i < t.nx) {
    const int je = (t.py && j == t.ny) ? 0 : j;
    int a, b;
    ac_face_pair(je, t.ny, t.py, a, b);
    const int f = t.v(i, je, k);
    const double goy = ac_gy(t, P, je, a, b, i, k);
    double rv = SV[f] - t.cp * t.thv[f] * goy;
    if (t.nu_div != 0.0) {
      const double doy = ((ac_div(t, UC, U, VC, V, i, b, k) - ac_div(t, UC, U, VC, V, i, a, k)) * t.rdy) *
                         t.mv[t.v2(i, je)];
      rv = rv + t.nu_div * doy;
    }
    double vh = V[f] + 0.5 * t.dtau * rv;
    if (!t.py && (j == 0 || j == t.ny)) vh = VB[t.v(i, j, k)];
    VH[t.v(i, j, k)] = vh;
    GY[t.v(i, j, k)] = goy;
  }
}

// ---- column right-hand side (steppers.py:1604-1614, 1653-1710) -----------------
// u_w, v_w: the half-step winds averaged to mass points and then to w level kk.
__device__ __forceinline__ double ac_uw(const AcCtx &t, const double *UH, int i, int j, int kk) {
  const int g = ac_wlev(kk, t.nz);
  const double lo = 0.5 * (UH[t.u(i, j, g - 1)] + UH[t.u(i + 1, j, g - 1)]);
  const double hi = 0.5 * (UH[t.u(i, j, g)] + UH[t.u(i + 1, j, g)]);
  return 0.5 * (lo + hi);
}
__device__ __forceinline__ double ac_vw(const AcCtx &t, const double *VH, int i, int j, int kk) {
  const int g = ac_wlev(kk, t.nz);
  const double lo = 0.5 * (VH[t.v(i, j, g - 1)] + VH[t.v(i, j + 1, g - 1)]);
  const double hi = 0.5 * (VH[t.v(i, j, g)] + VH[t.v(i, j + 1, g)]);
  return 0.5 * (lo + hi);
}
// Kinematic surface value of w from the half-step winds.
__device__ __forceinline__ double ac_wsfc(const AcCtx &t, const double *UH, const double *VH, int i, int j) {
  return t.mw[t.m2(i, j)] * (ac_uw(t, UH, i, j, 0) * t.zxiw[t.w(i, j, 0)] + ac_vw(t, VH, i, j, 0) * t.zetaw[t.w(i, j, 0)]);
}
// Explicit w at level kk and the vertical mass flux rho th (w_avg - terrain part) there.
__device__ __forceinline__ double ac_wexp(const AcCtx &t, const double *P, const double *W, const double *SW, int i,
                                          int j, int kk) {
  const int q = t.w(i, j, kk);
  return W[q] + t.dtau * (SW[q] - t.cp * t.thw[q] * ac_dzw(t, P, i, j, kk));
}
__device__ __forceinline__ double ac_fz(const AcCtx &t, const double *P, const double *W, const double *SW,
                                        const double *UH, const double *VH, int i, int j, int kk, double &we) {
  we = ac_wexp(t, P, W, SW, i, j, kk);
  if (kk == 0 || kk == t.nz) return 0.0;
  const int q = t.w(i, j, kk);
  const double wa = (1.0 - t.alpha) * W[q] + t.alpha * we;
  const double edz = wa - t.mw[t.m2(i, j)] * (ac_uw(t, UH, i, j, kk) * t.zxiw[q] + ac_vw(t, VH, i, j, kk) * t.zetaw[q]);
  return edz * t.rtw[q];
}

__global__ void __launch_bounds__(256) ac_column_rhs_kernel(
    AcCtx t, const double *__restrict__ P, const double *__restrict__ W, const double *__restrict__ SW,
    const double *__restrict__ SP, const double *__restrict__ UH, const double *__restrict__ VH,
    double *__restrict__ WE, double *__restrict__ R) {
  const int p = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= t.nx * t.ny * t.nz) return;
  const int i = p % t.nx, j = (p / t.nx) % t.ny, k = p / (t.nx * t.ny);
  double we0, we1;
  const double f0 = ac_fz(t, P, W, SW, UH, VH, i, j, k, we0);
  const double f1 = ac_fz(t, P, W, SW, UH, VH, i, j, k + 1, we1);
  WE[t.w(i, j, k)] = we0;
  if (k == t.nz - 1) WE[t.w(i, j, t.nz)] = we1;
  const int c = t.m(i, j, k);
  const double dzm = t.dzm[c];
  const double mmc = t.mm[t.m2(i, j)];
  const double div_z = (((f1 - f0) * t.rdz) * t.dz) / dzm;
  const double fxr = UH[t.u(i + 1, j, k)] * t.cx[t.u(i + 1, j, k)];
  const double fxl = UH[t.u(i, j, k)] * t.cx[t.u(i, j, k)];
  const double fyr = VH[t.v(i, j + 1, k)] * t.cy[t.v(i, j + 1, k)];
  const double fyl = VH[t.v(i, j, k)] * t.cy[t.v(i, j, k)];
  const double dxp = (((fxr - fxl) * t.rdx) * mmc) / dzm;
  const double dyp = (((fyr - fyl) * t.rdy) * mmc) / dzm;
  const double div_h = mmc * (dxp + dyp);
  R[c] = t.dtau * (SP[c] - t.cpi[c] * (div_h + div_z));
}

// ---- column sweep (Thomas with precomputed factors), pi'' += dpi -----------------
__global__ void __launch_bounds__(64) ac_column_kernel(AcCtx t, const double *__restrict__ lower,
                                                       const double *__restrict__ cprime,
                                                       const double *__restrict__ invden, double *__restrict__ R,
                                                       double *__restrict__ P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int n2 = t.nx * t.ny;
  if (p >= n2) return;
  double d = R[p] * invden[p];
  R[p] = d;
  for (int k = 1; k < t.nz; ++k) {
    const int q = k * n2 + p;
    d = (R[q] - lower[q] * d) * invden[q];
    R[q] = d;
  }
  double x = d;
  P[(t.nz - 1) * n2 + p] += x;
  for (int k = t.nz - 2; k >= 0; --k) {
    const int q = k * n2 + p;
    x = R[q] - cprime[q] * x;
    R[q] = x;
    P[q] += x;
  }
}

// ---- corrector, implicit w and Rayleigh damping (steppers.py:1615-1651, 1711-1737) --
__global__ void __launch_bounds__(256) ac_corrector_kernel(
    AcCtx t, const double *__restrict__ P, const double *__restrict__ D, const double *__restrict__ UH,
    const double *__restrict__ VH, const double *__restrict__ UC, const double *__restrict__ VC,
    const double *__restrict__ WC, const double *__restrict__ SU, const double *__restrict__ SV,
    const double *__restrict__ GX, const double *__restrict__ GY, const double *__restrict__ UB,
    const double *__restrict__ VB, const double *__restrict__ WE, double *__restrict__ U,
    double *__restrict__ V, double *__restrict__ W) {
  const int p = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int sx = t.nx + 1, sy = t.ny + 1;
  if (p >= sx * sy * (t.nz + 1)) return;
  const int i = p % sx, j = (p / sx) % sy, k = p / (sx * sy);
  // U and V are updated in place, each face by the one thread that reads its
  // old value; on a periodic axis the thread of face 0 also writes the last
  // face (the same face), whose own thread does nothing
  if (j < t.ny && k < t.nz && !(t.px && i == t.nx)) {
    int a, b;
    ac_face_pair(i, t.nx, t.px, a, b);
    const int f = t.u(i, j, k);
    const double gnx = ac_gx(t, P, i, a, b, j, k);
    double r = SU[f] - 0.5 * t.cp * t.thu[f] * (GX[f] + gnx);
    if (t.nu_div != 0.0) {
      const double dhx = ((ac_div(t, UC, UH, VC, VH, b, j, k) - ac_div(t, UC, UH, VC, VH, a, j, k)) * t.rdx) *
                         t.mu[t.u2(i, j)];
      r = r + t.nu_div * dhx;
    }
    double un = U[f] + t.dtau * r;
    if (!t.px && (i == 0 || i == t.nx)) un = UB[f];
    U[f] = un;
    if (t.px && i == 0) U[t.u(t.nx, j, k)] = un;
  }
  if (i < t.nx && k < t.nz && !(t.py && j == t.ny)) {
    int a, b;
    ac_face_pair(j, t.ny, t.py, a, b);
    const int f = t.v(i, j, k);
    const double gny = ac_gy(t, P, j, a, b, i, k);
    double r = SV[f] - 0.5 * t.cp * t.thv[f] * (GY[f] + gny);
    if (t.nu_div != 0.0) {
      const double dhy = ((ac_div(t, UC, UH, VC, VH, i, b, k) - ac_div(t, UC, UH, VC, VH, i, a, k)) * t.rdy) *
                         t.mv[t.v2(i, j)];
      r = r + t.nu_div * dhy;
    }
    double vn = V[f] + t.dtau * r;
    if (!t.py && (j == 0 || j == t.ny)) vn = VB[f];
    V[f] = vn;
    if (t.py && j == 0) V[t.v(i, t.ny, k)] = vn;
  }
  if (i < t.nx && j < t.ny) {
    const int q = t.w(i, j, k);
    double wn;
    if (k == 0) {
      wn = ac_wsfc(t, UH, VH, i, j);
    } else if (k == t.nz) {
      wn = 0.0;
    } else {
      wn = WE[q] - t.alpha * t.dtau * t.cp * t.thw[q] * ac_dzw(t, D, i, j, k);
    }
    const double wt = WC[q] + wn;
    W[q] = wt / (1.0 + t.dtau * t.tau[q]) - WC[q];
  }
}

}  // namespace sisl

using namespace sisl;

extern "C" int64_t sisl_acoustic_workspace_f64(const SislGrid *g) {
  if (!g || g->nx < 2 || g->ny < 2 || g->nz < 2) return -1;
  const int64_t nu = (int64_t)(g->nx + 1) * g->ny * g->nz, nv = (int64_t)g->nx * (g->ny + 1) * g->nz;
  const int64_t nw = (int64_t)g->nx * g->ny * (g->nz + 1), nm = (int64_t)g->nx * g->ny * g->nz;
  return 2 * nu + 2 * nv + nw + nm;
}

extern "C" int sisl_acoustic_loop_f64(const SislAcoustic *a, int32_t nsteps, double dtau, const double *lower,
                                      const double *cprime, const double *invden, double *u, double *v,
                                      double *w, double *pi, const double *uc, const double *vc,
                                      const double *wc, const double *su, const double *sv, const double *sw,
                                      const double *sp, const double *ub, const double *vb, double *work,
                                      void *stream) {
  if (!a || nsteps < 0) return SISL_ERR_ARG;
  const SislGrid &g = a->g;
  if (g.nx < 2 || g.ny < 2 || g.nz < 2) return SISL_ERR_ARG;
  const int64_t nu = (int64_t)(g.nx + 1) * g.ny * g.nz, nv = (int64_t)g.nx * (g.ny + 1) * g.nz;
  const int64_t nw = (int64_t)g.nx * g.ny * (g.nz + 1), nm = (int64_t)g.nx * g.ny * g.nz;
  const int64_t span = (int64_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  if (span > INT32_MAX) return SISL_ERR_TOO_LARGE;
  const void *need[] = {lower, cprime, invden, u, v, w, pi, uc, vc, wc, su, sv, sw, sp, ub, vb, work,
                        a->mu, a->mv, a->mm, a->mw, a->thu, a->thv, a->thw, a->zxu, a->zev, a->rz, a->dzm,
                        a->cpi, a->cx, a->cy, a->rtw, a->zxiw, a->zetaw, a->tau};
  for (const void *q : need)
    if (!q) return SISL_ERR_ARG;
  AcCtx t;
  t.nx = g.nx; t.ny = g.ny; t.nz = g.nz; t.px = g.periodic_x; t.py = g.periodic_y;
  t.rdx = 1.0 / a->dx; t.rdy = 1.0 / a->dy; t.rdz = 1.0 / a->dz; t.dz = a->dz;
  t.cp = a->cp; t.alpha = a->alpha; t.nu_div = a->nu_div; t.dtau = dtau;
#define AC_P(name) t.name = static_cast<const double *>(a->name)
  AC_P(mu); AC_P(mv); AC_P(mm); AC_P(mw); AC_P(thu); AC_P(thv); AC_P(thw); AC_P(zxu); AC_P(zev); AC_P(rz);
  AC_P(dzm); AC_P(cpi); AC_P(cx); AC_P(cy); AC_P(rtw); AC_P(zxiw); AC_P(zetaw); AC_P(tau);
#undef AC_P
  double *UH = work, *VH = UH + nu, *GX = VH + nv, *GY = GX + nu, *WE = GY + nv, *R = WE + nw;
  (void)nm;
  hipStream_t s = as_stream(stream);
  const int64_t n_faces = (int64_t)(g.nx + 1) * (g.ny + 1) * g.nz;
  // block counts rounded up to a multiple of the XCD count (xcd_tile)
  auto xgrid = [](int64_t n) { return dim3((grid_for(n) + kXcds - 1) / kXcds * kXcds); };
  for (int step = 0; step < nsteps; ++step) {
    hipLaunchKernelGGL(ac_predictor_kernel, xgrid(n_faces), dim3(kBlock), 0, s, t, pi, u, v, uc, vc, su, sv, ub,
                       vb, UH, VH, GX, GY);
    hipLaunchKernelGGL(ac_column_rhs_kernel, xgrid(nm), dim3(kBlock), 0, s, t, pi, w, sw, sp, UH, VH, WE, R);
    hipLaunchKernelGGL(ac_column_kernel, dim3(grid_for((int64_t)g.nx * g.ny, 64)), dim3(64), 0, s, t, lower,
                       cprime, invden, R, pi);
    hipLaunchKernelGGL(ac_corrector_kernel, xgrid(span), dim3(kBlock), 0, s, t, pi, R, UH, VH, uc, vc, wc, su, sv,
                       GX, GY, ub, vb, WE, u, v, w);
    SISL_CHECK_LAUNCH();
  }
  return SISL_OK;
}
