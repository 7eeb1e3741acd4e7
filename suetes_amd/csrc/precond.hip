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
// Waves per SIMD the segment kernels ask the compiler for: 8 for FP32 with
// segments of at most 10 levels (64 VGPRs, four 512-thread workgroups per CU,
// so a 256 x 256 grid's 1024 workgroups run in one round instead of 1.33;
// measured M 110 -> 97 us, M^T 145 -> 113 us at 256 x 256 x 80), no
// constraint otherwise (FP64 and long segments would spill).
template <typename T, int L> struct SegWaves { static constexpr int v = (sizeof(T) == 4 && L <= 10) ? 8 : 1; };
// Workgroups of the segment-parallel kernels: one per 64 columns of a row.
__host__ __device__ inline int blocks_M_dev(const Stag &s) { return ((s.nx + 63) / 64) * s.ny; }

// Body for launch block `blk` of `nblk`; blockDim.y may exceed the segments
// the column needs (the extra waves carry identity compositions).
template <typename T, int L, typename TC = T>
__device__ __forceinline__ void precond_seg_body(const OpT<T> &o, const T *__restrict__ r,
                                                 T *__restrict__ z, int blk, int nblk) {
  __shared__ T la[kMaxSeg][64], lb[kMaxSeg][64];
  const Stag &s = o.s;
  // 3-D coefficient arrays (T or bfloat16 storage)
  const TC *__restrict__ k_kx = reinterpret_cast<const TC *>(o.kx);
  const TC *__restrict__ k_ky = reinterpret_cast<const TC *>(o.ky);
  const TC *__restrict__ k_idw = reinterpret_cast<const TC *>(o.idw);
  const TC *__restrict__ k_cz = reinterpret_cast<const TC *>(o.cz);
  const TC *__restrict__ k_lpi = reinterpret_cast<const TC *>(o.lpi);
  const TC *__restrict__ k_lower = reinterpret_cast<const TC *>(o.lower);
  const TC *__restrict__ k_invden = reinterpret_cast<const TC *>(o.invden);
  const TC *__restrict__ k_cprime = reinterpret_cast<const TC *>(o.cprime);
  const TC *__restrict__ k_gw = reinterpret_cast<const TC *>(o.gw);
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
      const T kin = ldc<T>(k_kx, w) * (half * (um(ka - 1) + um_k)) + ldc<T>(k_ky, w) * (half * (vm(ka - 1) + vm_k));
      zep[0] = ldg(RE, w) * ra - kin;
      rwi[0] = ldg(RW, w) * ldc<T>(k_idw, w);
      wt_lo = ldc<T>(k_cz, w) * (rwi[0] + zep[0]);
    }
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int k = ka + t;
      if (k < kb) {
        T wt_hi = 0;
        if (k + 1 < nz) {
          const uint32_t w = mc + (uint32_t)(k + 1) * pm;
          const T um1 = um(k + 1), vm1 = vm(k + 1);
          const T kin = ldc<T>(k_kx, w) * (half * (um_k + um1)) + ldc<T>(k_ky, w) * (half * (vm_k + vm1));
          const T zp = ldg(RE, w) * ra - kin, rw = ldg(RW, w) * ldc<T>(k_idw, w);
          wt_hi = ldc<T>(k_cz, w) * (rw + zp);
          if (t + 1 < L) { zep[t + 1] = zp; rwi[t + 1] = rw; }
          um_k = um1;
          vm_k = vm1;
        }
        const uint32_t m = mc + (uint32_t)k * pm;
        const T rh = ldg(RP, m) * (T)1e-3 - ldc<T>(k_lpi, m) * (wt_hi - wt_lo);
        const T lo = ldc<T>(k_lower, m), iv = ldc<T>(k_invden, m);
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
        const T cp = ldc<T>(k_cprime, mc + (uint32_t)k * pm);
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
        const T pw = rwi[t] - ldc<T>(k_gw, m) * (d[t] - xm1);
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

template <typename T, int L, typename TC>
__global__ void __launch_bounds__(512, (SegWaves<T, L>::v)) precond_seg_kernel(OpT<T> o, const T *__restrict__ r,
                                                           T *__restrict__ z,
                                                           const int *__restrict__ stop) {
  if (stop && *stop) return;  // uniform over the launch
  precond_seg_body<T, L, TC>(o, r, z, blockIdx.x, gridDim.x);
}

// Transpose, segment-parallel like the forward solve.  M = B_out * T^{-1} *
// B_in per column, with T = L U the Thomas factorisation (U unit upper with
// super-diagonal c', L lower with pivots 1/invden and sub-diagonal `lower`).
// The reversed chain: p = B_out^T zbar (the pi, w, eta_dot back-substitution
// transposed), U^T y = p (forward recurrence y_k = p_k - c'_{k-1} y_{k-1}),
// L^T x = y (backward recurrence x_k = invden_k (y_k - lower_{k+1} x_{k+1})),
// then B_in^T x: the Helmholtz right-hand side transposed into w, eta_dot
// and mass-point cotangents of the averaged u, v (umb, vmb), which the faces
// kernel gathers to the u, v faces.  Wave `sg` owns levels [sg*L, sg*L + L)
// of 64 columns of row j; both recurrences compose across segments through
// LDS as in the forward solve.
template <typename T, int L, typename TC = T>
__device__ __forceinline__ void precond_T_seg_body(const OpT<T> &o, const T *__restrict__ zb,
                                                   T *__restrict__ rb, T *__restrict__ umb,
                                                   T *__restrict__ vmb, int blk, int nblk) {
  __shared__ T la[kMaxSeg][64], lb[kMaxSeg][64];
  const Stag &s = o.s;
  // 3-D coefficient arrays (T or bfloat16 storage)
  const TC *__restrict__ k_kx = reinterpret_cast<const TC *>(o.kx);
  const TC *__restrict__ k_ky = reinterpret_cast<const TC *>(o.ky);
  const TC *__restrict__ k_idw = reinterpret_cast<const TC *>(o.idw);
  const TC *__restrict__ k_cz = reinterpret_cast<const TC *>(o.cz);
  const TC *__restrict__ k_lpi = reinterpret_cast<const TC *>(o.lpi);
  const TC *__restrict__ k_lower = reinterpret_cast<const TC *>(o.lower);
  const TC *__restrict__ k_invden = reinterpret_cast<const TC *>(o.invden);
  const TC *__restrict__ k_cprime = reinterpret_cast<const TC *>(o.cprime);
  const TC *__restrict__ k_gw = reinterpret_cast<const TC *>(o.gw);
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int nbx = (nx + 63) / 64;
  const int tile = xcd_tile(blk, nblk);
  const int bx = tile % nbx, j = tile / nbx;
  const int tx = threadIdx.x, sg = threadIdx.y, S = blockDim.y;
  const int i = bx * 64 + tx;
  const int ka = sg * L, kb = min(ka + L, nz);
  const bool act = i < nx && ka < nz;
  const T *ZW = zb + s.ow, *ZP = zb + s.om, *ZE = zb + s.oe;
  T *RW = rb + s.ow, *RP = rb + s.om, *RE = rb + s.oe;
  const T ra = (T)1 / o.alpha, half = (T)0.5;
  const uint32_t pm = (uint32_t)nx * ny;
  const uint32_t mc = (uint32_t)j * nx + i;
  // zs[t] = zbar_w + zbar_eta at w level ka + t, ze[t] = zbar_eta there
  T d[L], B[L], zs[L], ze[L];
  T dl = 0, bl = 1;
  if (act) {
    // ---- p_k = 1000 zbar_pi - g_k + g_{k+1}, g = gw (zbar_w + zbar_eta) on interior w levels;
    //      U^T y = p, local forward sweep ----
    T g_lo = 0;
    if (ka >= 1) {
      const uint32_t w = mc + (uint32_t)ka * pm;
      const T e = ldg(ZE, w);
      zs[0] = ldg(ZW, w) + e;
      ze[0] = e;
      g_lo = ldc<T>(k_gw, w) * zs[0];
    }
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int k = ka + t;
      if (k < kb) {
        T g_hi = 0;
        if (k + 1 < nz) {
          const uint32_t w = mc + (uint32_t)(k + 1) * pm;
          const T e = ldg(ZE, w), sum = ldg(ZW, w) + e;
          g_hi = ldc<T>(k_gw, w) * sum;
          if (t + 1 < L) { zs[t + 1] = sum; ze[t + 1] = e; }
        }
        const uint32_t m = mc + (uint32_t)k * pm;
        const T pp = (T)1000 * ldg(ZP, m) - g_lo + g_hi;
        const T a = k >= 1 ? -ldc<T>(k_cprime, m - pm) : (T)0;
        dl = pp + a * dl;
        bl = a * bl;
        d[t] = dl;
        B[t] = bl;
        g_lo = g_hi;
      }
    }
  }
  la[sg][tx] = dl;
  lb[sg][tx] = bl;
  __syncthreads();
  T din = 0;
  for (int t = 0; t < sg; ++t) din = la[t][tx] + lb[t][tx] * din;
  __syncthreads();
  // ---- L^T x = y, local backward sweep ----
  T xl = 0, gl = 1;
  if (act) {
#pragma unroll
    for (int t = L - 1; t >= 0; --t) {
      const int k = ka + t;
      if (k < kb) {
        const uint32_t m = mc + (uint32_t)k * pm;
        const T y = d[t] + B[t] * din;
        const T iv = ldc<T>(k_invden, m);
        const T c = k + 1 < nz ? -iv * ldc<T>(k_lower, m + pm) : (T)0;
        xl = iv * y + c * xl;
        gl = c * gl;
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
  la[sg][tx] = xtop;       // x at the segment's top level, for the segment above
  lb[sg][tx] = act ? d[0] : (T)0;  // x at the segment's bottom level, for the segment below
  __syncthreads();
  if (!act) return;
  const T xbelow = sg > 0 ? la[sg - 1][tx] : (T)0;
  const T xabove = kb < nz ? lb[sg + 1][tx] : (T)0;
  const uint32_t q2 = (uint32_t)s.m2(i, j);
  // ---- outputs: pi; then per w level k in [ka, kb] the right-hand-side transpose
  //      (w, eta_dot owned for k < kb; the top face only feeds umb, vmb) ----
  T divb_lo = ka >= 1 ? -ldc<T>(k_lpi, mc + (uint32_t)(ka - 1) * pm) * xbelow : (T)0;
  T uw_lo = 0, vw_lo = 0;
#pragma unroll
  for (int t = 0; t <= L; ++t) {
    const int k = ka + t;
    if (k <= kb) {
      const uint32_t m = mc + (uint32_t)k * pm;  // mass index of level k == w index of level k
      const T xk = k < kb ? d[t < L ? t : L - 1] : xabove;
      if (k < kb) stg(RP, m, xk * (T)1e-3);
      const T divb_hi = k < nz ? -ldc<T>(k_lpi, m) * xk : (T)0;
      T uw = 0, vw = 0;
      if (k == 0) {
        stg(RW, m, ldg(ZW, m));
        stg(RE, m, ldg(ZE, m) * ldg(o.dzw0, q2));
      } else if (k == nz) {
        stg(RW, m, ldg(ZW, m));
        stg(RE, m, ldg(ZE, m) * ldg(o.dzwn, q2));
      } else {
        T zsum, e;
        if (k < kb && t < L) {
          zsum = zs[t < L ? t : L - 1];
          e = ze[t < L ? t : L - 1];
        } else {
          e = ldg(ZE, m);
          zsum = ldg(ZW, m) + e;
        }
        const T wckb = ldc<T>(k_cz, m) * (divb_lo - divb_hi);
        if (k < kb) {
          stg(RW, m, ldc<T>(k_idw, m) * (zsum + wckb));
          stg(RE, m, (e + wckb) * ra);
        }
        const T kinb = -e - wckb;
        uw = ldc<T>(k_kx, m) * kinb;
        vw = ldc<T>(k_ky, m) * kinb;
      }
      if (k > ka) {
        stg(umb, m - pm, half * uw_lo + half * uw);
        stg(vmb, m - pm, half * vw_lo + half * vw);
      }
      uw_lo = uw;
      vw_lo = vw;
      divb_lo = divb_hi;
    }
  }
}

template <typename T, int L, typename TC>
__global__ void __launch_bounds__(512, (SegWaves<T, L>::v)) precond_T_seg_kernel(OpT<T> o, const T *__restrict__ zb,
                                                             T *__restrict__ rb,
                                                             T *__restrict__ umb,
                                                             T *__restrict__ vmb,
                                                             const int *__restrict__ stop) {
  if (stop && *stop) return;
  precond_T_seg_body<T, L, TC>(o, zb, rb, umb, vmb, blockIdx.x, gridDim.x);
}

// u, v faces of the transpose: zbar passes through and each face gathers half
// of the mass-point cotangent of its (one or two) neighbouring columns.  One
// thread per cell (i, j, k) of a 64x4 tile of one level: it does the u face i
// and the v face j, and the cells of the last column / row also the east /
// north domain faces, so that every thread has work when nx and ny are
// multiples of 64 and 4.
constexpr int kFTY = 4;
__host__ __device__ inline int faces_blocks(const Stag &s) {
  return ((s.nx + 63) / 64) * ((s.ny + kFTY - 1) / kFTY) * s.nz;
}
template <typename T>
__device__ __forceinline__ void precond_T_faces_body(const OpT<T> &o, const T *__restrict__ zb,
                                                     const T *__restrict__ umb,
                                                     const T *__restrict__ vmb,
                                                     T *__restrict__ rb, int blk) {
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny;
  const int nbx = (nx + 63) / 64, nby = (ny + kFTY - 1) / kFTY;
  const int bx = blk % nbx, r = blk / nbx, by = r % nby, k = r / nby;
  const int i = bx * 64 + threadIdx.x, j = by * kFTY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const T half = (T)0.5;
  const uint32_t m = ((uint32_t)k * ny + j) * nx + i;  // mass index of (i, j, k)
  const T uc = ldg(umb, m), vc = ldg(vmb, m);
  {
    const uint32_t q = ((uint32_t)k * ny + j) * (nx + 1) + i;
    T acc = 0;
    if (i >= 1) acc += ldg(umb, m - 1);
    acc += uc;
    stg(rb + s.ou, q, ldg(zb + s.ou, q) + half * acc);
    if (i == nx - 1) stg(rb + s.ou, q + 1, ldg(zb + s.ou, q + 1) + half * uc);
  }
  {
    const uint32_t q = ((uint32_t)k * (ny + 1) + j) * nx + i;
    T acc = 0;
    if (j >= 1) acc += ldg(vmb, m - nx);
    acc += vc;
    stg(rb + s.ov, q, ldg(zb + s.ov, q) + half * acc);
    if (j == ny - 1) stg(rb + s.ov, q + nx, ldg(zb + s.ov, q + nx) + half * vc);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) precond_T_faces_kernel(OpT<T> o, const T *__restrict__ zb,
                                                              const T *__restrict__ umb,
                                                              const T *__restrict__ vmb,
                                                              T *__restrict__ rb,
                                                              const int *__restrict__ stop) {
  if (stop && *stop) return;
  precond_T_faces_body(o, zb, umb, vmb, rb, blockIdx.x);
}

// Levels per segment: the smallest supported L with at most kMaxSeg
// segments per column.
template <typename T, typename TC>
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
  // packed vector in and out, 5 w-level and 4 mass-level coefficient arrays (T or bfloat16), w and
  // eta_dot ends, 2-D end factors
  const double sc = (double)sizeof(TC);
  const double bytes = (double)sizeof(T) * (g.n + 2.0 * g.w.n() + g.m.n() + 2.0 * g.nx * g.ny +
                                            (r != z ? (double)(g.u.n() + g.v.n()) : 0.0)) +
                       sc * (5.0 * g.w.n() + 4.0 * g.m.n());
  const int tid = timer_begin(SISL_T_PRECOND, sizeof(T) == 4, bytes, stop, nullptr, st);
  switch (L) {
#define SISL_M_CASE(c) \
  case c: hipLaunchKernelGGL((precond_seg_kernel<T, c, TC>), grid, block, 0, st, o, r, z, stop); break;
    SISL_M_CASE(2) SISL_M_CASE(4) SISL_M_CASE(8) SISL_M_CASE(10) SISL_M_CASE(12) SISL_M_CASE(16)
    SISL_M_CASE(24) SISL_M_CASE(32)
#undef SISL_M_CASE
  }
  timer_end(tid, st);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int launch_M(const OpT<T> &o, const T *r, T *z, hipStream_t st, const int *stop, int cb) {
  if ((uint64_t)o.s.n * sizeof(T) >= (1ull << 32)) return SISL_ERR_ARG;  // 32-bit offsets
  if (cb) {
    if constexpr (sizeof(T) == 4) return launch_M_seg<T, bf16_t>(o, r, z, st, stop);
    return SISL_ERR_ARG;
  }
  return launch_M_seg<T, T>(o, r, z, st, stop);
}

template <typename T, typename TC>
static int launch_MT_tc(const OpT<T> &o, const T *zb, T *rb, T *scratch, hipStream_t st, const int *stop) {
  const int L = seg_levels(o.s.nz);
  if (!L) return SISL_ERR_ARG;
  const int S = (o.s.nz + L - 1) / L;
  const Stag &g = o.s;
  const int64_t nm = g.m.n();
  T *umb = scratch, *vmb = scratch + nm;
  const dim3 grid((unsigned)blocks_M(g)), block(64, S);
  // column part: zbar_w, zbar_eta, zbar_pi; gw, cz, idw, kx, ky (w), cprime, invden, lower, lpi
  // (mass), dzw0, dzwn (2-D) in; w, eta_dot, pi outputs and umb, vmb out.
  // faces: zbar_u, zbar_v, umb, vmb in; u, v outputs out.
  const double nw = (double)g.w.n(), nmf = (double)nm, nuv = (double)(g.u.n() + g.v.n());
  const double bytes = (double)sizeof(T) * ((2 * nw + nmf) + 2.0 * g.nx * g.ny + (2 * nw + 3 * nmf) +
                                            (nuv + 2 * nmf) + nuv) +
                       (double)sizeof(TC) * (5 * nw + 4 * nmf);
  const int tid = timer_begin(SISL_T_PRECOND_T, sizeof(T) == 4, bytes, stop, nullptr, st);
  switch (L) {
#define SISL_MT_CASE(c) \
  case c: hipLaunchKernelGGL((precond_T_seg_kernel<T, c, TC>), grid, block, 0, st, o, zb, rb, umb, vmb, stop); break;
    SISL_MT_CASE(2) SISL_MT_CASE(4) SISL_MT_CASE(8) SISL_MT_CASE(10) SISL_MT_CASE(12) SISL_MT_CASE(16)
    SISL_MT_CASE(24) SISL_MT_CASE(32)
#undef SISL_MT_CASE
    default: return SISL_ERR_ARG;
  }
  hipLaunchKernelGGL(precond_T_faces_kernel<T>, dim3((unsigned)faces_blocks(g)), dim3(64, kFTY), 0, st, o,
                     zb, (const T *)umb, (const T *)vmb, rb, stop);
  timer_end(tid, st);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int launch_MT(const OpT<T> &o, const T *zb, T *rb, T *scratch, hipStream_t st, const int *stop, int cb) {
  if ((uint64_t)o.s.n * sizeof(T) >= (1ull << 32)) return SISL_ERR_ARG;  // 32-bit offsets
  if (cb) {
    if constexpr (sizeof(T) == 4) return launch_MT_tc<T, bf16_t>(o, zb, rb, scratch, st, stop);
    return SISL_ERR_ARG;
  }
  return launch_MT_tc<T, T>(o, zb, rb, scratch, st, stop);
}

// ---- batched (ensemble) launches -------------------------------------------

int blocks_M(const Stag &s) { return blocks_M_dev(s); }
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
  precond_seg_body<T, L, T>(b.o, b.x + j * b.xj, b.y, blockIdx.x, nb);
}

template <typename T, int L>
__global__ void __launch_bounds__(512) precond_T_seg_batched_kernel(const BOp<T> *__restrict__ tab, int j) {
  const BOp<T> &b = tab[blockIdx.y];
  if (b.stop && b.stop[j * b.sj]) return;
  const int nb = blocks_M_dev(b.o.s);
  if ((int)blockIdx.x >= nb) return;
  const int64_t nm = b.o.s.m.n();
  precond_T_seg_body<T, L, T>(b.o, b.x + j * b.xj, b.y, b.scr, b.scr + nm, blockIdx.x, nb);
}

template <typename T>
__global__ void __launch_bounds__(256) precond_T_faces_batched_kernel(const BOp<T> *__restrict__ tab, int j) {
  const BOp<T> &b = tab[blockIdx.y];
  if (b.stop && b.stop[j * b.sj]) return;
  if ((int)blockIdx.x >= faces_blocks(b.o.s)) return;
  const int64_t nm = b.o.s.m.n();
  precond_T_faces_body(b.o, b.x + j * b.xj, (const T *)b.scr, (const T *)(b.scr + nm), b.y, blockIdx.x);
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
int launch_MT_batched(const BOp<T> *tab, int M, int maxb, int L, int S, int maxf, int j, hipStream_t st) {
  const dim3 grid(maxb, M), block(64, S);
  switch (L) {
#define SISL_MTB_CASE(c) \
  case c: hipLaunchKernelGGL((precond_T_seg_batched_kernel<T, c>), grid, block, 0, st, tab, j); break;
    SISL_MTB_CASE(2) SISL_MTB_CASE(4) SISL_MTB_CASE(8) SISL_MTB_CASE(10) SISL_MTB_CASE(12) SISL_MTB_CASE(16)
    SISL_MTB_CASE(24) SISL_MTB_CASE(32)
#undef SISL_MTB_CASE
    default: return SISL_ERR_ARG;
  }
  hipLaunchKernelGGL(precond_T_faces_batched_kernel<T>, dim3(maxf, M), dim3(64, kFTY), 0, st, tab, j);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}
int blocks_MT_faces(const Stag &s) { return faces_blocks(s); }
template int launch_M_batched<double>(const BOp<double> *, int, int, int, int, int, hipStream_t);
template int launch_M_batched<float>(const BOp<float> *, int, int, int, int, int, hipStream_t);
template int launch_MT_batched<double>(const BOp<double> *, int, int, int, int, int, int, hipStream_t);
template int launch_MT_batched<float>(const BOp<float> *, int, int, int, int, int, int, hipStream_t);

template int launch_M<double>(const OpT<double> &, const double *, double *, hipStream_t, const int *, int);
template int launch_M<float>(const OpT<float> &, const float *, float *, hipStream_t, const int *, int);
template int launch_MT<double>(const OpT<double> &, const double *, double *, double *, hipStream_t,
                               const int *, int);
template int launch_MT<float>(const OpT<float> &, const float *, float *, float *, hipStream_t,
                              const int *, int);

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
