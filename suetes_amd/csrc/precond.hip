// (c) Column tridiagonal preconditioner M and its transpose.
//
// VerticalPreconditioner (suetes/regional3d/steppers.py:125-253): per column,
// build the Helmholtz right-hand side, one Thomas solve, back-substitute w
// and eta_dot.  Consecutive lanes own consecutive x columns, so every level's
// loads are coalesced along x.  The Thomas factorisation (pivots, c') depends
// only on the background state and is precomputed once per stepper.
#include "common.h"

namespace sisl {

// Forward M, segment-parallel.  A workgroup owns 64 consecutive columns of
// one row j; wave `sg` owns levels [sg*L, sg*L + L) of those columns and keeps
// them in registers.  With the Thomas factors fixed, forward and backward
// substitution are first-order affine recurrences: each segment runs its
// part with a zero incoming value while accumulating the product of the
// recurrence coefficients, the segment ends are composed through LDS, and a
// final multiply-add gives the exact sequential result up to rounding.  Every
// input element is read once and every output element written once.
constexpr int kMaxSeg = 8;  // at most 8 segments (512 threads): columns up to 256 levels

// Body for launch block `blk` of `nblk`; blockDim.y may exceed the segments
// the column needs (the extra waves carry identity compositions).
template <typename T, int L>
__device__ __forceinline__ void precond_seg_body(const OpT<T> &o, const T *__restrict__ r,
                                                 T *__restrict__ z, int blk, int nblk) {
  __shared__ T la[kMaxSeg][64], lb[kMaxSeg][64];
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int nbx = (nx + 63) / 64;
  const int tile = xcd_tile(blk, nblk);
  const int bx = tile % nbx, j = tile / nbx;
  const int tx = threadIdx.x, sg = threadIdx.y, S = blockDim.y;
  const int i = bx * 64 + tx;
  const int ka = sg * L, kb = min(ka + L, nz);
  const bool act = i < nx && ka < nz;
  const T *RU = r + s.ou, *RV = r + s.ov, *RW = r + s.ow, *RP = r + s.om, *RE = r + s.oe;
  T *ZU = z + s.ou, *ZV = z + s.ov, *ZW = z + s.ow, *ZP = z + s.om, *ZE = z + s.oe;
  const T ra = (T)1 / o.alpha, half = (T)0.5;
  const uint32_t pm = (uint32_t)nx * ny, pu = (uint32_t)(nx + 1) * ny, pv = (uint32_t)nx * (ny + 1);
  const uint32_t mc = (uint32_t)j * nx + i, uc = (uint32_t)j * (nx + 1) + i, vc = (uint32_t)j * nx + i;
  auto um = [&](int k) { const uint32_t q = uc + (uint32_t)k * pu; return half * (ldg(RU, q) + ldg(RU, q + 1)); };
  auto vm = [&](int k) { const uint32_t q = vc + (uint32_t)k * pv; return half * (ldg(RV, q) + ldg(RV, q + nx)); };
  T d[L], B[L], zep[L], rwi[L];
  T dl = 0, bl = 1;
  if (act) {
    // ---- Helmholtz right-hand side (steppers.py:201-223) + local forward sweep ----
    T um_k = um(ka), vm_k = vm(ka), wt_lo = 0;
    if (ka >= 1) {
      const uint32_t w = mc + (uint32_t)ka * pm;
      const T kin = ldg(o.kx, w) * (half * (um(ka - 1) + um_k)) + ldg(o.ky, w) * (half * (vm(ka - 1) + vm_k));
      zep[0] = ldg(RE, w) * ra - kin;
      rwi[0] = ldg(RW, w) * ldg(o.idw, w);
      wt_lo = ldg(o.cz, w) * (rwi[0] + zep[0]);
    }
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int k = ka + t;
      if (k < kb) {
        T wt_hi = 0;
        if (k + 1 < nz) {
          const uint32_t w = mc + (uint32_t)(k + 1) * pm;
          const T um1 = um(k + 1), vm1 = vm(k + 1);
          const T kin = ldg(o.kx, w) * (half * (um_k + um1)) + ldg(o.ky, w) * (half * (vm_k + vm1));
          const T zp = ldg(RE, w) * ra - kin, rw = ldg(RW, w) * ldg(o.idw, w);
          wt_hi = ldg(o.cz, w) * (rw + zp);
          if (t + 1 < L) { zep[t + 1] = zp; rwi[t + 1] = rw; }
          um_k = um1;
          vm_k = vm1;
        }
        const uint32_t m = mc + (uint32_t)k * pm;
        const T rh = ldg(RP, m) * (T)1e-3 - ldg(o.lpi, m) * (wt_hi - wt_lo);
        const T lo = ldg(o.lower, m), iv = ldg(o.invden, m);
        dl = (rh - lo * dl) * iv;
        bl = bl * (-lo * iv);
        d[t] = dl;
        B[t] = bl;
        wt_lo = wt_hi;
      }
    }
  }
  la[sg][tx] = dl;
  lb[sg][tx] = bl;
  __syncthreads();
  T din = 0;
  for (int t = 0; t < sg; ++t) din = la[t][tx] + lb[t][tx] * din;
  __syncthreads();
  // ---- local backward sweep: x = d - c' x_above ----
  T xl = 0, gl = 1;
  if (act) {
#pragma unroll
    for (int t = L - 1; t >= 0; --t) {
      const int k = ka + t;
      if (k < kb) {
        const T dk = d[t] + B[t] * din;
        const T cp = ldg(o.cprime, mc + (uint32_t)k * pm);
        xl = dk - cp * xl;
        gl = -cp * gl;
        d[t] = xl;
        B[t] = gl;
      }
    }
  }
  la[sg][tx] = xl;
  lb[sg][tx] = gl;
  __syncthreads();
  T xin = 0;
  for (int t = S - 1; t > sg; --t) xin = la[t][tx] + lb[t][tx] * xin;
  T xtop = 0;
  if (act) {
#pragma unroll
    for (int t = 0; t < L; ++t)
      if (ka + t < kb) {
        d[t] = d[t] + B[t] * xin;
        xtop = d[t];
      }
  }
  __syncthreads();
  la[sg][tx] = xtop;
  __syncthreads();
  if (!act) return;
  const T xbelow = sg > 0 ? la[sg - 1][tx] : (T)0;
  // ---- outputs: pi, w and eta_dot back-substitution (steppers.py:230-245) ----
#pragma unroll
  for (int t = 0; t < L; ++t) {
    const int k = ka + t;
    if (k < kb) {
      const uint32_t m = mc + (uint32_t)k * pm;
      stg(ZP, m, d[t] * (T)1000);
      if (k >= 1) {
        const T xm1 = t == 0 ? xbelow : d[t > 0 ? t - 1 : 0];
        const T pw = rwi[t] - ldg(o.gw, m) * (d[t] - xm1);
        stg(ZW, m, pw);
        stg(ZE, m, zep[t] + pw);
      }
    }
  }
  const uint32_t q2 = (uint32_t)s.m2(i, j);
  if (ka == 0) {
    stg(ZW, mc, ldg(RW, mc));
    stg(ZE, mc, ldg(RE, mc) * ldg(o.dzw0, q2));
  }
  if (kb == nz) {
    const uint32_t w = mc + (uint32_t)nz * pm;
    stg(ZW, w, ldg(RW, w));
    stg(ZE, w, ldg(RE, w) * ldg(o.dzwn, q2));
  }
  // ---- u, v pass through unchanged (steppers.py:247-253) ----
  if (z != r) {
    for (int k = ka; k < kb; ++k) {
      const uint32_t qu = uc + (uint32_t)k * pu, qv = vc + (uint32_t)k * pv;
      stg(ZU, qu, ldg(RU, qu));
      if (i == nx - 1) stg(ZU, qu + 1, ldg(RU, qu + 1));
      stg(ZV, qv, ldg(RV, qv));
      if (j == ny - 1) stg(ZV, qv + nx, ldg(RV, qv + nx));
    }
  }
}

template <typename T, int L>
__global__ void __launch_bounds__(512) precond_seg_kernel(OpT<T> o, const T *__restrict__ r,
                                                           T *__restrict__ z,
                                                           const int *__restrict__ stop) {
  if (stop && *stop) return;  // uniform over the launch
  precond_seg_body<T, L>(o, r, z, blockIdx.x, gridDim.x);
}

// Transpose: column sweeps of the reversed chain; the u/v cotangents of the
// kinematic term leave through mass-point scratch and are gathered to faces.
template <typename T>
__device__ __forceinline__ void precond_T_body(const OpT<T> &o, const T *__restrict__ zb,
                                               T *__restrict__ rb, T *__restrict__ umb,
                                               T *__restrict__ vmb, int64_t col) {
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
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
__global__ void __launch_bounds__(256) precond_T_kernel(OpT<T> o, const T *__restrict__ zb,
                                                        T *__restrict__ rb,
                                                        T *__restrict__ umb,
                                                        T *__restrict__ vmb,
                                                        const int *__restrict__ stop) {
  if (stop && *stop) return;
  precond_T_body(o, zb, rb, umb, vmb, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T>
__device__ __forceinline__ void precond_T_faces_body(const OpT<T> &o, const T *__restrict__ zb,
                                                     const T *__restrict__ umb,
                                                     const T *__restrict__ vmb,
                                                     T *__restrict__ rb, int64_t q) {
  const Stag &s = o.s;
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
__global__ void __launch_bounds__(256) precond_T_faces_kernel(OpT<T> o, const T *__restrict__ zb,
                                                              const T *__restrict__ umb,
                                                              const T *__restrict__ vmb,
                                                              T *__restrict__ rb,
                                                              const int *__restrict__ stop) {
  if (stop && *stop) return;
  precond_T_faces_body(o, zb, umb, vmb, rb, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Levels per segment: the smallest supported L with at most kMaxSeg
// segments per column.
template <typename T>
static int launch_M_seg(const OpT<T> &o, const T *r, T *z, hipStream_t st, const int *stop) {
  const int nz = o.s.nz;
  static const int sizes[] = {2, 4, 8, 10, 12, 16, 24, 32};
  int L = 0;
  for (int c : sizes)
    if ((nz + c - 1) / c <= kMaxSeg) { L = c; break; }
  if (!L) return SISL_ERR_ARG;  // taller than kMaxSeg * 32 levels
  const int S = (nz + L - 1) / L;
  const dim3 grid((unsigned)(((o.s.nx + 63) / 64) * o.s.ny)), block(64, S);
  const Stag &g = o.s;
  const double bytes = (double)sizeof(T) * (g.n + 7.0 * g.w.n() + 5.0 * g.m.n() + 2.0 * g.nx * g.ny +
                                            (r != z ? (double)(g.u.n() + g.v.n()) : 0.0));
  const int tid = timer_begin(SISL_T_PRECOND, sizeof(T) == 4, bytes, stop, nullptr, st);
  switch (L) {
#define SISL_M_CASE(c) \
  case c: hipLaunchKernelGGL((precond_seg_kernel<T, c>), grid, block, 0, st, o, r, z, stop); break;
    SISL_M_CASE(2) SISL_M_CASE(4) SISL_M_CASE(8) SISL_M_CASE(10) SISL_M_CASE(12) SISL_M_CASE(16)
    SISL_M_CASE(24) SISL_M_CASE(32)
#undef SISL_M_CASE
  }
  timer_end(tid, st);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int launch_M(const OpT<T> &o, const T *r, T *z, hipStream_t st, const int *stop) {
  if ((uint64_t)o.s.n * sizeof(T) >= (1ull << 32)) return SISL_ERR_ARG;  // 32-bit offsets
  return launch_M_seg(o, r, z, st, stop);
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

// ---- batched (ensemble) launches -------------------------------------------

int blocks_M(const Stag &s) { return ((s.nx + 63) / 64) * s.ny; }
int seg_levels(int nz) {
  static const int sizes[] = {2, 4, 8, 10, 12, 16, 24, 32};
  for (int c : sizes)
    if ((nz + c - 1) / c <= kMaxSeg) return c;
  return 0;
}

template <typename T, int L>
__global__ void __launch_bounds__(512) precond_seg_batched_kernel(const BOp<T> *__restrict__ tab, int j) {
  const BOp<T> &b = tab[blockIdx.y];
  if (b.stop && b.stop[j * b.sj]) return;
  const int nb = ((b.o.s.nx + 63) / 64) * b.o.s.ny;
  if ((int)blockIdx.x >= nb) return;
  precond_seg_body<T, L>(b.o, b.x + j * b.xj, b.y, blockIdx.x, nb);
}

template <typename T>
__global__ void __launch_bounds__(256) precond_T_batched_kernel(const BOp<T> *__restrict__ tab, int j) {
  const BOp<T> &b = tab[blockIdx.y];
  if (b.stop && b.stop[j * b.sj]) return;
  const int64_t nm = b.o.s.m.n();
  precond_T_body(b.o, b.x + j * b.xj, b.y, b.scr, b.scr + nm,
                 (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T>
__global__ void __launch_bounds__(256) precond_T_faces_batched_kernel(const BOp<T> *__restrict__ tab, int j) {
  const BOp<T> &b = tab[blockIdx.y];
  if (b.stop && b.stop[j * b.sj]) return;
  const int64_t nm = b.o.s.m.n();
  precond_T_faces_body(b.o, b.x + j * b.xj, (const T *)b.scr, (const T *)(b.scr + nm), b.y,
                       (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T>
int launch_M_batched(const BOp<T> *tab, int M, int maxb, int L, int S, int j, hipStream_t st) {
  const dim3 grid(maxb, M), block(64, S);
  switch (L) {
#define SISL_MB_CASE(c) \
  case c: hipLaunchKernelGGL((precond_seg_batched_kernel<T, c>), grid, block, 0, st, tab, j); break;
    SISL_MB_CASE(2) SISL_MB_CASE(4) SISL_MB_CASE(8) SISL_MB_CASE(10) SISL_MB_CASE(12) SISL_MB_CASE(16)
    SISL_MB_CASE(24) SISL_MB_CASE(32)
#undef SISL_MB_CASE
    default: return SISL_ERR_ARG;
  }
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}
template <typename T>
int launch_MT_batched(const BOp<T> *tab, int M, int maxcols, int maxfaces, int j, hipStream_t st) {
  hipLaunchKernelGGL(precond_T_batched_kernel<T>, dim3(grid_for(maxcols), M), dim3(kBlock), 0, st, tab, j);
  hipLaunchKernelGGL(precond_T_faces_batched_kernel<T>, dim3(grid_for(maxfaces), M), dim3(kBlock), 0, st,
                     tab, j);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}
template int launch_M_batched<double>(const BOp<double> *, int, int, int, int, int, hipStream_t);
template int launch_M_batched<float>(const BOp<float> *, int, int, int, int, int, hipStream_t);
template int launch_MT_batched<double>(const BOp<double> *, int, int, int, int, hipStream_t);
template int launch_MT_batched<float>(const BOp<float> *, int, int, int, int, hipStream_t);

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
