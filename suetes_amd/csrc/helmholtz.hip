// (b) Matrix-free operator A of the semi-implicit solve and its transpose.
//
// A is Euler3D.linear_operator (suetes/regional3d/euler.py:374-439) seen
// through the solver's scaling (steppers.py:439-461): packed Krylov vectors
// u | v | w | pi*1000 | eta_dot*dz_w.  One thread per output element; all
// staggered neighbours come through the cache hierarchy.  Every coefficient
// that the reference recomputes per call (background averages, metrics,
// map factors) is precomputed once per stepper by the host layer.
#include "common.h"

namespace sisl {

template <typename T>
struct AView {
  const OpT<T> &o;
  const T *U, *V, *W, *P, *E;
  __device__ AView(const OpT<T> &op, const T *x)
      : o(op), U(x + op.s.ou), V(x + op.s.ov), W(x + op.s.ow), P(x + op.s.om), E(x + op.s.oe) {}

  // d(pi)/dz at a w level times the 1/dz_w metric, with the edge closure of
  // op.diff(m->w) (operators.py:37-63, euler.py:252).
  __device__ __forceinline__ T gzw(int i, int j, int k) const {
    const Stag &s = o.s;
    int ka = k, kb = k - 1;
    if (k == 0) { ka = 1; kb = 0; }
    else if (k == s.nz) { ka = s.nz - 1; kb = s.nz - 2; }
    return (P[s.m.at(i, j, ka)] - P[s.m.at(i, j, kb)]) * o.rdzw[s.w.at(i, j, k)];
  }
  __device__ __forceinline__ T gzm(int i, int j, int k) const {
    return (T)0.5 * (gzw(i, j, k) + gzw(i, j, k + 1));
  }
  __device__ __forceinline__ T um(int i, int j, int k) const {
    const Stag &s = o.s;
    return (T)0.5 * (U[s.u.at(i, j, k)] + U[s.u.at(i + 1, j, k)]);
  }
  __device__ __forceinline__ T vm(int i, int j, int k) const {
    const Stag &s = o.s;
    return (T)0.5 * (V[s.v.at(i, j, k)] + V[s.v.at(i, j + 1, k)]);
  }
};

// Launch geometry of the point-wise kernels: block (64, 4) threads over
// (i, j); blockIdx.z enumerates (segment, k).  No integer division per thread.
constexpr int kTX = 64, kTY = 4;

template <typename T>
__global__ void __launch_bounds__(256) helmholtz_kernel(OpT<T> o, const T *__restrict__ x,
                                                        T *__restrict__ y,
                                                        const int *__restrict__ stop) {
  if (stop && *stop) return;
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int seg = blockIdx.z / (nz + 1), k = blockIdx.z - seg * (nz + 1);
  const int i = blockIdx.x * kTX + threadIdx.x, j = blockIdx.y * kTY + threadIdx.y;
  AView<T> a(o, x);
  if (seg == 0) {  // ---- L_u (euler.py:398, 411-415) ----
    if (i > nx || j >= ny || k >= nz) return;
    const int64_t p = s.u.at(i, j, k);
    T out;
    if (!s.px && (i == 0 || i == nx)) {
      out = a.U[p];
    } else {
      const int f = (i == nx) ? 0 : i;           // periodic: face nx is face 0
      const int im = f == 0 ? nx - 1 : f - 1;
      const int64_t pf = s.u.at(f, j, k);
      const T dpi = a.P[s.m.at(f, j, k)] - a.P[s.m.at(im, j, k)];
      const T gzu = (T)0.5 * (a.gzm(im, j, k) + a.gzm(f, j, k));
      const T vau = (T)0.5 * (a.vm(im, j, k) + a.vm(f, j, k));
      out = a.U[pf] + o.cua[pf] * (o.mudx[s.u2(f, j)] * dpi - o.zxu[pf] * gzu) - o.afu[s.u2(f, j)] * vau;
    }
    y[s.ou + p] = out;
  } else if (seg == 1) {  // ---- L_v (euler.py:399, 416-420) ----
    if (i >= nx || j > ny || k >= nz) return;
    const int64_t p = s.v.at(i, j, k);
    T out;
    if (!s.py && (j == 0 || j == ny)) {
      out = a.V[p];
    } else {
      const int f = (j == ny) ? 0 : j;
      const int jm = f == 0 ? ny - 1 : f - 1;
      const int64_t pf = s.v.at(i, f, k);
      const T dpi = a.P[s.m.at(i, f, k)] - a.P[s.m.at(i, jm, k)];
      const T gzv = (T)0.5 * (a.gzm(i, jm, k) + a.gzm(i, f, k));
      const T uav = (T)0.5 * (a.um(i, jm, k) + a.um(i, f, k));
      out = a.V[pf] + o.cva[pf] * (o.mvdy[s.v2(i, f)] * dpi - o.zev[pf] * gzv) + o.afv[s.v2(i, f)] * uav;
    }
    y[s.ov + p] = out;
  } else if (seg == 2) {  // ---- L_w (euler.py:400, 426-427) ----
    if (i >= nx || j >= ny) return;
    const int64_t p = s.w.at(i, j, k);
    T out;
    if (k == 0 || k == nz) {
      out = a.W[p];
    } else {
      out = o.dw[p] * a.W[p] + o.cw[p] * (a.P[s.m.at(i, j, k)] - a.P[s.m.at(i, j, k - 1)]);
    }
    y[s.ow + p] = out;
  } else if (seg == 3) {  // ---- L_pi (euler.py:298-318, 401) ----
    if (i >= nx || j >= ny || k >= nz) return;
    const int64_t p = s.m.at(i, j, k);
    const T fx = a.U[s.u.at(i + 1, j, k)] * o.cx[s.u.at(i + 1, j, k)] - a.U[s.u.at(i, j, k)] * o.cx[s.u.at(i, j, k)];
    const T fy = a.V[s.v.at(i, j + 1, k)] * o.cy[s.v.at(i, j + 1, k)] - a.V[s.v.at(i, j, k)] * o.cy[s.v.at(i, j, k)];
    const T ftop = (k + 1 == nz) ? (T)0 : a.E[s.w.at(i, j, k + 1)] * o.cz[s.w.at(i, j, k + 1)];
    const T fbot = (k == 0) ? (T)0 : a.E[s.w.at(i, j, k)] * o.cz[s.w.at(i, j, k)];
    const int64_t q2 = s.m2(i, j);
    y[s.om + p] = a.P[p] + o.cpm[p] * (o.mmx[q2] * fx + o.mmy[q2] * fy + (ftop - fbot));
  } else {  // ---- L_eta (euler.py:429-437) ----
    if (i >= nx || j >= ny) return;
    const int64_t p = s.w.at(i, j, k);
    T out;
    if (k == 0 || k == nz) {
      out = a.E[p] * o.rdzw[p];
    } else {
      const T uw = (T)0.5 * (a.um(i, j, k - 1) + a.um(i, j, k));
      const T vw = (T)0.5 * (a.vm(i, j, k - 1) + a.vm(i, j, k));
      out = o.alpha * (a.E[p] + o.kx[p] * uw + o.ky[p] * vw - a.W[p]);
    }
    y[s.oe + p] = out;
  }
}

inline dim3 point_grid(const Stag &s, int segments) {
  return dim3((unsigned)((s.nx + 1 + kTX - 1) / kTX), (unsigned)((s.ny + 1 + kTY - 1) / kTY),
              (unsigned)(segments * (s.nz + 1)));
}

// ---- transpose ---------------------------------------------------------------
// Phase 1 (per mass point): cotangents of the mass-point intermediates
// gz_m, v_m, u_m.  Phase 2 (per packed element): gather into A^T y.

template <typename T>
struct ATView {
  const OpT<T> &o;
  const T *U, *V, *W, *P, *E;  // cotangents of L_u .. L_eta
  __device__ ATView(const OpT<T> &op, const T *y)
      : o(op), U(y + op.s.ou), V(y + op.s.ov), W(y + op.s.ow), P(y + op.s.om), E(y + op.s.oe) {}
  // effective cotangent of the formula at u face f (periodic faces 0 / nx merged)
  __device__ __forceinline__ T ut(int f, int j, int k) const {
    const Stag &s = o.s;
    if (s.px) {
      if (f == s.nx) f = 0;
      if (f == 0) return U[s.u.at(0, j, k)] + U[s.u.at(s.nx, j, k)];
      return U[s.u.at(f, j, k)];
    }
    if (f == 0 || f == s.nx) return (T)0;
    return U[s.u.at(f, j, k)];
  }
  __device__ __forceinline__ T vt(int i, int f, int k) const {
    const Stag &s = o.s;
    if (s.py) {
      if (f == s.ny) f = 0;
      if (f == 0) return V[s.v.at(i, 0, k)] + V[s.v.at(i, s.ny, k)];
      return V[s.v.at(i, f, k)];
    }
    if (f == 0 || f == s.ny) return (T)0;
    return V[s.v.at(i, f, k)];
  }
  __device__ __forceinline__ int ufold(int f) const { return (o.s.px && f == o.s.nx) ? 0 : f; }
  __device__ __forceinline__ int vfold(int f) const { return (o.s.py && f == o.s.ny) ? 0 : f; }
};

template <typename T>
__global__ void __launch_bounds__(256) helmholtz_T1_kernel(OpT<T> o, const T *__restrict__ y,
                                                           T *__restrict__ gzm_b,
                                                           T *__restrict__ um_b,
                                                           T *__restrict__ vm_b,
                                                           const int *__restrict__ stop) {
  if (stop && *stop) return;
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int i = blockIdx.x * kTX + threadIdx.x, j = blockIdx.y * kTY + threadIdx.y, k = blockIdx.z;
  if (i >= nx || j >= ny || k >= nz) return;
  const int64_t p = s.m.at(i, j, k);
  ATView<T> t(o, y);
  const int fl = i, fr = t.ufold(i + 1);
  const T ul = t.ut(fl, j, k), ur = t.ut(i + 1, j, k);
  const int64_t pl = s.u.at(fl, j, k), pr = s.u.at(fr, j, k);
  const int gl = j, gr = t.vfold(j + 1);
  const T vl = t.vt(i, gl, k), vr = t.vt(i, j + 1, k);
  const int64_t ql = s.v.at(i, gl, k), qr = s.v.at(i, gr, k);
  T g = (T)-0.5 * (ul * o.cua[pl] * o.zxu[pl] + ur * o.cua[pr] * o.zxu[pr]);
  g += (T)-0.5 * (vl * o.cva[ql] * o.zev[ql] + vr * o.cva[qr] * o.zev[qr]);
  T vmb = (T)-0.5 * (ul * o.afu[s.u2(fl, j)] + ur * o.afu[s.u2(fr, j)]);
  T umb = (T)0.5 * (vl * o.afv[s.v2(i, gl)] + vr * o.afv[s.v2(i, gr)]);
  // L_eta interior levels k and k+1 read u_m/v_m at mass level k
  T ek = 0, ek1 = 0, kxk = 0, kxk1 = 0, kyk = 0, kyk1 = 0;
  if (k >= 1) { const int64_t w = s.w.at(i, j, k); ek = t.E[w]; kxk = o.kx[w]; kyk = o.ky[w]; }
  if (k + 1 <= nz - 1) { const int64_t w = s.w.at(i, j, k + 1); ek1 = t.E[w]; kxk1 = o.kx[w]; kyk1 = o.ky[w]; }
  umb += (T)0.5 * o.alpha * (ek * kxk + ek1 * kxk1);
  vmb += (T)0.5 * o.alpha * (ek * kyk + ek1 * kyk1);
  gzm_b[p] = g;
  um_b[p] = umb;
  vm_b[p] = vmb;
}

template <typename T>
__global__ void __launch_bounds__(256) helmholtz_T2_kernel(OpT<T> o, const T *__restrict__ y,
                                                           const T *__restrict__ gzm_b,
                                                           const T *__restrict__ um_b,
                                                           const T *__restrict__ vm_b,
                                                           T *__restrict__ xo,
                                                           const int *__restrict__ stop) {
  if (stop && *stop) return;
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int seg = blockIdx.z / (nz + 1), k = blockIdx.z - seg * (nz + 1);
  const int i = blockIdx.x * kTX + threadIdx.x, j = blockIdx.y * kTY + threadIdx.y;
  ATView<T> t(o, y);
  T out;
  int64_t q;
  if (seg == 0) {  // ---- U bar ----
    if (i > nx || j >= ny || k >= nz) return;
    const int64_t p = s.u.at(i, j, k);
    q = s.ou + p;
    if (s.px) out = (i == 0) ? t.U[p] + t.U[s.u.at(nx, j, k)] : (i == nx ? (T)0 : t.U[p]);
    else out = t.U[p];
    T fl = 0;
    if (i - 1 >= 0) { const int64_t m = s.m.at(i - 1, j, k); fl += o.cpm[m] * o.mmx[s.m2(i - 1, j)] * t.P[m]; }
    if (i <= nx - 1) { const int64_t m = s.m.at(i, j, k); fl -= o.cpm[m] * o.mmx[s.m2(i, j)] * t.P[m]; }
    out += o.cx[p] * fl;
    T ub = 0;
    if (i - 1 >= 0) ub += um_b[s.m.at(i - 1, j, k)];
    if (i <= nx - 1) ub += um_b[s.m.at(i, j, k)];
    out += (T)0.5 * ub;
  } else if (seg == 1) {  // ---- V bar ----
    if (i >= nx || j > ny || k >= nz) return;
    const int64_t p = s.v.at(i, j, k);
    q = s.ov + p;
    if (s.py) out = (j == 0) ? t.V[p] + t.V[s.v.at(i, ny, k)] : (j == ny ? (T)0 : t.V[p]);
    else out = t.V[p];
    T fl = 0;
    if (j - 1 >= 0) { const int64_t m = s.m.at(i, j - 1, k); fl += o.cpm[m] * o.mmy[s.m2(i, j - 1)] * t.P[m]; }
    if (j <= ny - 1) { const int64_t m = s.m.at(i, j, k); fl -= o.cpm[m] * o.mmy[s.m2(i, j)] * t.P[m]; }
    out += o.cy[p] * fl;
    T vb = 0;
    if (j - 1 >= 0) vb += vm_b[s.m.at(i, j - 1, k)];
    if (j <= ny - 1) vb += vm_b[s.m.at(i, j, k)];
    out += (T)0.5 * vb;
  } else if (seg == 2) {  // ---- W bar ----
    if (i >= nx || j >= ny) return;
    const int64_t p = s.w.at(i, j, k);
    q = s.ow + p;
    if (k == 0 || k == nz) out = t.W[p];
    else out = o.dw[p] * t.W[p] - o.alpha * t.E[p];
  } else if (seg == 3) {  // ---- P bar ----
    if (i >= nx || j >= ny || k >= nz) return;
    const int64_t p = s.m.at(i, j, k);
    q = s.om + p;
    out = t.P[p];
    {  // x faces i (P is the upper cell) and i+1 (P is the lower cell)
      const int64_t a = s.u.at(i, j, k), b = s.u.at(t.ufold(i + 1), j, k);
      out += t.ut(i, j, k) * o.cua[a] * o.mudx[s.u2(i, j)];
      out -= t.ut(i + 1, j, k) * o.cua[b] * o.mudx[s.u2(t.ufold(i + 1), j)];
    }
    {
      const int64_t a = s.v.at(i, j, k), b = s.v.at(i, t.vfold(j + 1), k);
      out += t.vt(i, j, k) * o.cva[a] * o.mvdy[s.v2(i, j)];
      out -= t.vt(i, j + 1, k) * o.cva[b] * o.mvdy[s.v2(i, t.vfold(j + 1))];
    }
    if (k >= 1) { const int64_t w = s.w.at(i, j, k); out += o.cw[w] * t.W[w]; }
    if (k + 1 <= nz - 1) { const int64_t w = s.w.at(i, j, k + 1); out -= o.cw[w] * t.W[w]; }
    // gz_w cotangent at level kk: 0.5 (gzm_b(kk-1) + gzm_b(kk))
    auto gzwb = [&](int kk) {
      T r = 0;
      if (kk - 1 >= 0) r += gzm_b[s.m.at(i, j, kk - 1)];
      if (kk <= nz - 1) r += gzm_b[s.m.at(i, j, kk)];
      return (T)0.5 * r * o.rdzw[s.w.at(i, j, kk)];
    };
    // levels whose upper operand is pi(k)
    if (k >= 1 && k <= nz - 1) out += gzwb(k);
    if (k == 1) out += gzwb(0);
    if (k == nz - 1) out += gzwb(nz);
    // levels whose lower operand is pi(k)
    if (k + 1 >= 1 && k + 1 <= nz - 1) out -= gzwb(k + 1);
    if (k == 0) out -= gzwb(0);
    if (k == nz - 2) out -= gzwb(nz);
  } else {  // ---- E bar ----
    if (i >= nx || j >= ny) return;
    const int64_t p = s.w.at(i, j, k);
    q = s.oe + p;
    if (k == 0 || k == nz) {
      out = t.E[p] * o.rdzw[p];
    } else {
      out = o.alpha * t.E[p] +
            o.cz[p] * (o.cpm[s.m.at(i, j, k - 1)] * t.P[s.m.at(i, j, k - 1)] - o.cpm[s.m.at(i, j, k)] * t.P[s.m.at(i, j, k)]);
    }
  }
  xo[q] = out;
}

template <typename T>
int launch_A(const OpT<T> &o, const T *x, T *y, hipStream_t st, const int *stop) {
  hipLaunchKernelGGL(helmholtz_kernel<T>, point_grid(o.s, 5), dim3(kTX, kTY), 0, st, o, x, y, stop);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int launch_AT(const OpT<T> &o, const T *y, T *x, T *scratch, hipStream_t st, const int *stop) {
  const int64_t nm = o.s.m.n();
  T *g = scratch, *um = scratch + nm, *vm = scratch + 2 * nm;
  const dim3 g1((unsigned)((o.s.nx + kTX - 1) / kTX), (unsigned)((o.s.ny + kTY - 1) / kTY),
                (unsigned)o.s.nz);
  hipLaunchKernelGGL(helmholtz_T1_kernel<T>, g1, dim3(kTX, kTY), 0, st, o, y, g, um, vm, stop);
  hipLaunchKernelGGL(helmholtz_T2_kernel<T>, point_grid(o.s, 5), dim3(kTX, kTY), 0, st, o, y,
                     (const T *)g, (const T *)um, (const T *)vm, x, stop);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template int launch_A<double>(const OpT<double> &, const double *, double *, hipStream_t, const int *);
template int launch_A<float>(const OpT<float> &, const float *, float *, hipStream_t, const int *);
template int launch_AT<double>(const OpT<double> &, const double *, double *, double *, hipStream_t,
                               const int *);
template int launch_AT<float>(const OpT<float> &, const float *, float *, float *, hipStream_t,
                              const int *);

}  // namespace sisl

extern "C" {

int64_t sisl_krylov_size(const SislGrid *g) { return g ? sisl::Stag::make(*g).n : -1; }

int sisl_helmholtz_apply_f64(const SislOperator *op, const double *x, double *y, void *s) {
  if (!op || !x || !y) return SISL_ERR_ARG;
  auto o = sisl::OpT<double>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::launch_A(o, x, y, sisl::as_stream(s));
}
int sisl_helmholtz_apply_f32(const SislOperator *op, const float *x, float *y, void *s) {
  if (!op || !x || !y) return SISL_ERR_ARG;
  auto o = sisl::OpT<float>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::launch_A(o, x, y, sisl::as_stream(s));
}
int sisl_helmholtz_apply_T_f64(const SislOperator *op, const double *x, double *y, double *scr,
                               void *s) {
  if (!op || !x || !y || !scr || x == y) return SISL_ERR_ARG;
  auto o = sisl::OpT<double>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::launch_AT(o, x, y, scr, sisl::as_stream(s));
}
int sisl_helmholtz_apply_T_f32(const SislOperator *op, const float *x, float *y, float *scr,
                               void *s) {
  if (!op || !x || !y || !scr || x == y) return SISL_ERR_ARG;
  auto o = sisl::OpT<float>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::launch_AT(o, x, y, scr, sisl::as_stream(s));
}

}  // extern "C"
