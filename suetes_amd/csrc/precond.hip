// (c) Column tridiagonal preconditioner M and its transpose.
//
// VerticalPreconditioner (suetes/regional3d/steppers.py:125-253): per column,
// build the Helmholtz right-hand side, one Thomas solve, back-substitute w
// and eta_dot.  One lane owns one (i, j) column and sweeps k in registers;
// consecutive lanes own consecutive x columns, so every level's loads are
// coalesced along x.  The Thomas factorisation (pivots, c') depends only on
// the background state and is precomputed once per stepper.
#include "common.h"

namespace sisl {

template <typename T>
__global__ void __launch_bounds__(256) precond_kernel(OpT<T> o, const T *__restrict__ r,
                                                      T *__restrict__ z,
                                                      const int *__restrict__ stop) {
  if (stop && *stop) return;
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (int64_t)nx * ny) return;
  const int i = (int)(col % nx), j = (int)(col / nx);
  const T *RU = r + s.ou, *RV = r + s.ov, *RW = r + s.ow, *RP = r + s.om, *RE = r + s.oe;
  T *ZW = z + s.ow, *ZP = z + s.om, *ZE = z + s.oe;
  const T a = o.alpha;
  auto um = [&](int k) { return (T)0.5 * (RU[s.u.at(i, j, k)] + RU[s.u.at(i + 1, j, k)]); };
  auto vm = [&](int k) { return (T)0.5 * (RV[s.v.at(i, j, k)] + RV[s.v.at(i, j + 1, k)]); };
  const int64_t w0 = s.w.at(i, j, 0), wn = s.w.at(i, j, nz);
  const T rw0 = RW[w0], rwn = RW[wn], re0 = RE[w0], ren = RE[wn];
  // Sweep up: Helmholtz RHS (steppers.py:201-223) and Thomas forward elimination.
  T wt_lo = 0, dprev = 0;
  T um_k = um(0), vm_k = vm(0);
  for (int k = 0; k < nz; ++k) {
    T wt_hi = 0;
    if (k + 1 < nz) {
      const int64_t w = s.w.at(i, j, k + 1);
      const T um1 = um(k + 1), vm1 = vm(k + 1);
      const T kin = o.kx[w] * ((T)0.5 * (um_k + um1)) + o.ky[w] * ((T)0.5 * (vm_k + vm1));
      const T re = RE[w];
      const T wck = RW[w] * o.idw[w] + re / a - kin;
      wt_hi = o.cz[w] * wck;
      ZE[w] = re / a - kin;  // partial eta_dot back-substitution
      um_k = um1;
      vm_k = vm1;
    }
    const int64_t m = s.m.at(i, j, k);
    const T rh = RP[m] * (T)1e-3 - o.lpi[m] * (wt_hi - wt_lo);
    const T d = (rh - o.lower[m] * dprev) * o.invden[m];
    ZP[m] = d;
    dprev = d;
    wt_lo = wt_hi;
  }
  // Sweep down: back substitution, w and eta_dot (steppers.py:230-245).
  T xhi = 0;
  for (int k = nz - 1; k >= 0; --k) {
    const int64_t m = s.m.at(i, j, k);
    const T x = ZP[m] - o.cprime[m] * xhi;
    ZP[m] = x * (T)1000;
    if (k + 1 <= nz - 1) {
      const int64_t w = s.w.at(i, j, k + 1);
      const T pw = RW[w] * o.idw[w] - o.gw[w] * (xhi - x);
      ZW[w] = pw;
      ZE[w] += pw;
    }
    xhi = x;
  }
  ZW[w0] = rw0;
  ZW[wn] = rwn;
  const int64_t q2 = s.m2(i, j);
  ZE[w0] = re0 * o.dzw0[q2];
  ZE[wn] = ren * o.dzwn[q2];
}

// Transpose: column sweeps of the reversed chain; the u/v cotangents of the
// kinematic term leave through mass-point scratch and are gathered to faces.
template <typename T>
__global__ void __launch_bounds__(256) precond_T_kernel(OpT<T> o, const T *__restrict__ zb,
                                                        T *__restrict__ rb,
                                                        T *__restrict__ umb,
                                                        T *__restrict__ vmb,
                                                        const int *__restrict__ stop) {
  if (stop && *stop) return;
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (int64_t)nx * ny) return;
  const int i = (int)(col % nx), j = (int)(col / nx);
  const T *ZW = zb + s.ow, *ZP = zb + s.om, *ZE = zb + s.oe;
  T *RW = rb + s.ow, *RP = rb + s.om, *RE = rb + s.oe;
  const T a = o.alpha;
  // U^T y = pbar_p (forward), storing y in RP.
  T yprev = 0;
  for (int k = 0; k < nz; ++k) {
    const int64_t m = s.m.at(i, j, k);
    T pp = (T)1000 * ZP[m];
    if (k >= 1) {
      const int64_t w = s.w.at(i, j, k);
      pp -= o.gw[w] * (ZW[w] + ZE[w]);
    }
    if (k + 1 <= nz - 1) {
      const int64_t w = s.w.at(i, j, k + 1);
      pp += o.gw[w] * (ZW[w] + ZE[w]);
    }
    const T y = (k == 0) ? pp : pp - o.cprime[s.m.at(i, j, k - 1)] * yprev;
    RP[m] = y;
    yprev = y;
  }
  // L^T x = y (backward), then the Helmholtz-RHS transpose per w level.
  T xhi = 0, divb_hi = 0, umb_hi = 0, vmb_hi = 0;
  for (int k = nz - 1; k >= 0; --k) {
    const int64_t m = s.m.at(i, j, k);
    const T y = RP[m];
    const T x = (k == nz - 1) ? y * o.invden[m] : (y - o.lower[s.m.at(i, j, k + 1)] * xhi) * o.invden[m];
    RP[m] = x * (T)1e-3;
    const T divb = -o.lpi[m] * x;
    if (k + 1 <= nz - 1) {
      const int64_t w = s.w.at(i, j, k + 1);
      const T wckb = o.cz[w] * (divb - divb_hi);
      const T ze = ZE[w];
      RW[w] = o.idw[w] * (ZW[w] + ze + wckb);
      RE[w] = (ze + wckb) / a;
      const T kinb = -ze - wckb;
      const T uwb = o.kx[w] * kinb, vwb = o.ky[w] * kinb;
      umb[s.m.at(i, j, k + 1)] = umb_hi + (T)0.5 * uwb;
      vmb[s.m.at(i, j, k + 1)] = vmb_hi + (T)0.5 * vwb;
      umb_hi = (T)0.5 * uwb;
      vmb_hi = (T)0.5 * vwb;
    }
    xhi = x;
    divb_hi = divb;
  }
  umb[s.m.at(i, j, 0)] = umb_hi;
  vmb[s.m.at(i, j, 0)] = vmb_hi;
  const int64_t w0 = s.w.at(i, j, 0), wn = s.w.at(i, j, nz);
  const int64_t q2 = s.m2(i, j);
  RW[w0] = ZW[w0];
  RW[wn] = ZW[wn];
  RE[w0] = ZE[w0] * o.dzw0[q2];
  RE[wn] = ZE[wn] * o.dzwn[q2];
}

template <typename T>
__global__ void __launch_bounds__(256) precond_T_faces_kernel(OpT<T> o, const T *__restrict__ zb,
                                                              const T *__restrict__ umb,
                                                              const T *__restrict__ vmb,
                                                              T *__restrict__ rb,
                                                              const int *__restrict__ stop) {
  if (stop && *stop) return;
  const Stag &s = o.s;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.ow) return;
  const int nx = s.nx, ny = s.ny;
  T out = zb[q];
  if (q < s.ov) {
    const int64_t p = q;
    const int i = (int)(p % (nx + 1)), j = (int)((p / (nx + 1)) % ny), k = (int)(p / ((int64_t)(nx + 1) * ny));
    T acc = 0;
    if (i - 1 >= 0) acc += umb[s.m.at(i - 1, j, k)];
    if (i <= nx - 1) acc += umb[s.m.at(i, j, k)];
    out += (T)0.5 * acc;
  } else {
    const int64_t p = q - s.ov;
    const int i = (int)(p % nx), j = (int)((p / nx) % (ny + 1)), k = (int)(p / ((int64_t)nx * (ny + 1)));
    T acc = 0;
    if (j - 1 >= 0) acc += vmb[s.m.at(i, j - 1, k)];
    if (j <= ny - 1) acc += vmb[s.m.at(i, j, k)];
    out += (T)0.5 * acc;
  }
  rb[q] = out;
}

template <typename T>
int launch_M(const OpT<T> &o, const T *r, T *z, hipStream_t st, const int *stop) {
  if (r != z) {
    if (hipMemcpyAsync(z, r, sizeof(T) * o.s.ow, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return SISL_ERR_LAUNCH;
  }
  const int64_t cols = (int64_t)o.s.nx * o.s.ny;
  hipLaunchKernelGGL(precond_kernel<T>, dim3(grid_for(cols)), dim3(kBlock), 0, st, o, r, z, stop);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int launch_MT(const OpT<T> &o, const T *zb, T *rb, T *scratch, hipStream_t st, const int *stop) {
  const int64_t cols = (int64_t)o.s.nx * o.s.ny;
  const int64_t nm = o.s.m.n();
  hipLaunchKernelGGL(precond_T_kernel<T>, dim3(grid_for(cols)), dim3(kBlock), 0, st, o, zb, rb,
                     scratch, scratch + nm, stop);
  hipLaunchKernelGGL(precond_T_faces_kernel<T>, dim3(grid_for(o.s.ow)), dim3(kBlock), 0, st, o, zb,
                     (const T *)scratch, (const T *)(scratch + nm), rb, stop);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template int launch_M<double>(const OpT<double> &, const double *, double *, hipStream_t, const int *);
template int launch_M<float>(const OpT<float> &, const float *, float *, hipStream_t, const int *);
template int launch_MT<double>(const OpT<double> &, const double *, double *, double *, hipStream_t,
                               const int *);
template int launch_MT<float>(const OpT<float> &, const float *, float *, float *, hipStream_t,
                              const int *);

}  // namespace sisl

extern "C" {

int sisl_precond_apply_f64(const SislOperator *op, const double *r, double *z, void *s) {
  if (!op || !r || !z) return SISL_ERR_ARG;
  auto o = sisl::OpT<double>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::launch_M(o, r, z, sisl::as_stream(s));
}
int sisl_precond_apply_f32(const SislOperator *op, const float *r, float *z, void *s) {
  if (!op || !r || !z) return SISL_ERR_ARG;
  auto o = sisl::OpT<float>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::launch_M(o, r, z, sisl::as_stream(s));
}
int sisl_precond_apply_T_f64(const SislOperator *op, const double *r, double *z, double *scr,
                             void *s) {
  if (!op || !r || !z || !scr || r == z) return SISL_ERR_ARG;
  auto o = sisl::OpT<double>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::launch_MT(o, r, z, scr, sisl::as_stream(s));
}
int sisl_precond_apply_T_f32(const SislOperator *op, const float *r, float *z, float *scr,
                             void *s) {
  if (!op || !r || !z || !scr || r == z) return SISL_ERR_ARG;
  auto o = sisl::OpT<float>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::launch_MT(o, r, z, scr, sisl::as_stream(s));
}

}  // extern "C"
