// GMRES for the semi-implicit solve: Krylov reductions and the restart driver.
//
// Restatement of jax.scipy.sparse.linalg.gmres (solve_method='batched'),
// called at suetes/regional3d/steppers.py:465-481: left preconditioned,
// restarted; with early_exit = 0 every cycle runs all `restart` Arnoldi
// steps unless a breakdown occurs; Gram-Schmidt is classical with one
// re-orthogonalisation pass; the least squares uses the normal equations of
// the Hessenberg matrix; the loop stops when ||M(b - A x)|| <= ptol.
//
// Memory traffic per Arnoldi step j (j+1 basis vectors): one fused pass for
// the projections, one fused "subtract + re-project" pass, one fused
// "subtract + norm" pass, one normalisation pass.  Inner products accumulate
// in FP64 whatever the basis type; reductions are two-stage and fixed-order,
// so results are bit-reproducible run to run.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

namespace sisl {

// min(1, r) with NaN kept, as jnp.minimum / np.minimum do (std::fmin and
// fmin return the non-NaN operand): with b = 0 and atol = 0 the stopping
// threshold is NaN and no cycle runs, so x0 is returned unchanged.
__host__ __device__ inline double nan_min1(double r) { return r != r ? r : (r < 1.0 ? r : 1.0); }

constexpr int kRedGrid = 1024;  // blocks of the reduction passes
constexpr int kMaxRestart = 64;
constexpr int kPStride = kMaxRestart + 2;  // partial sums per block

// Device-side scalar state of one solve.
struct Small {
  double H[kMaxRestart * (kMaxRestart + 1)];  // row j = Arnoldi step j (JAX layout)
  double h1[kMaxRestart + 2];                 // first projection (+ w.w in slot nv)
  double h2[kMaxRestart + 2];                 // re-orthogonalisation projection
  double nn1a[1], nn1b[1];                    // ||w||^2 after pass 1 / pass 2
  double inv, target;
  double g[kMaxRestart + 1], cs[kMaxRestart], sn[kMaxRestart];
  double y[kMaxRestart];
  double scal[4];
  int broken;   // cycle stopped (breakdown or early exit): later kernels skip
  int steps;    // Arnoldi steps completed in this cycle
  int gate2;    // second Gram-Schmidt pass required
  int pad;
  // device-controlled outer loop (batched solver)
  int done, outer, steps_total, pad2;
  double b_norm, Mb_norm, ptol, rnorm;
  // per-vector basis scales (inner FP16 steps whose operator reads the
  // basis): unit basis vector v = vsc[v] * V[v]
  double vsc[kMaxRestart + 2];
};

__device__ void decide_body(Small *sm, int nv);
__device__ void finalize_body(Small *sm, int j, int R, int early_exit, double vs = 0.0);
// finalize_body for restarts up to kSmallRestart with the Hessenberg column
// and the rotations held in registers (fully unrolled, predicated on j):
// the same arithmetic in the same order, without the private-memory round
// trips of a dynamically indexed array
constexpr int kSmallRestart = 16;
__device__ void finalize_small(Small *sm, int j, int R, int early_exit, double vs);

// ---- block reduction helpers -------------------------------------------------
// One value per thread -> partials[slot][block] (fixed-order, deterministic).
__device__ __forceinline__ void block_store_slot(double v, int slot, double *partials, int blk) {
  __shared__ double red1[kBlock / 64];
  const double s = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red1[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) t += red1[q];
    partials[(int64_t)slot * kRedGrid + blk] = t;
  }
  __syncthreads();
}

// kGrp values per thread -> partials[v0 + r][block] for r < cnt.
template <int G>
__device__ __forceinline__ void block_store_group(double (&acc)[G], int v0, int cnt, double *partials,
                                                  int blk) {
  __shared__ double red[kBlock / 64][G];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < G; ++r) {
    const double s = wave_sum(acc[r]);
    if (lane == 0) red[wid][r] = s;
  }
  __syncthreads();
  if (threadIdx.x < cnt) {
    double t = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) t += red[q][threadIdx.x];
    partials[(int64_t)(v0 + threadIdx.x) * kRedGrid + blk] = t;
  }
  __syncthreads();
}

// out[v] = sum over blocks of partials[v][b], fixed order.  grid = nv blocks.
__global__ void __launch_bounds__(256) reduce_partials_kernel(const double *__restrict__ partials,
                                                              double *__restrict__ out,
                                                              const int *__restrict__ stop,
                                                              const int *__restrict__ gate) {
  if ((stop && *stop) || (gate && !*gate)) return;
  __shared__ double red[kBlock / 64];
  const int v = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < kRedGrid; b += blockDim.x) s += partials[(int64_t)v * kRedGrid + b];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int q = 0; q < kBlock / 64; ++q) t += red[q];
    out[v] = t;
  }
}

// Reduction tail of a projection / update pass: one workgroup sums the
// per-block partial sums of `nout` slots (wave q takes slots q, q + 16, ...;
// lane l the blocks l, l + 64, ... in order, then a fixed butterfly -- a
// fixed order, so results are reproducible) and then runs the scalar
// bookkeeping that follows the pass: the second-pass decision (mode 1) or
// the Arnoldi step's Hessenberg column, breakdown test and Givens update
// (mode 2), which also records the stop flag that the next step's kernels
// read.  One launch replaces the separate reduction, decision, bookkeeping
// and flag-copy launches of the step.  A skipped pass (stopped cycle, or no
// second Gram-Schmidt pass) reduces nothing but still runs its bookkeeping.
constexpr int kTailThreads = 1024;
struct Tail {
  double *out = nullptr;  // reduced slots
  int nout = 0;
  int mode = 0;
  Small *sm = nullptr;
  int nv = 0, j = 0, R = 0, early_exit = 0;
  int *next_flag = nullptr;
  double vs = 0.0;  // > 0: the update pass stored the next basis column (finalize_body)
};

// SMALL: the pass's restart is at most kSmallRestart (finalize_small; no
// private memory in the kernel).
template <bool SMALL>
__global__ void __launch_bounds__(kTailThreads) tail_kernel(const double *__restrict__ partials, Tail t,
                                                            const int *__restrict__ stop,
                                                            const int *__restrict__ gate) {
  const bool skipped = (stop && *stop) || (gate && !*gate);
  if (!skipped) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int v = wv; v < t.nout; v += nw) {
      double acc = 0;
      for (int b = lane; b < kRedGrid; b += 64) acc += partials[(int64_t)v * kRedGrid + b];
      acc = wave_sum(acc);
      if (lane == 0) t.out[v] = acc;
    }
  }
  if (!t.mode) return;
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (t.mode == 1) {
    decide_body(t.sm, t.nv);
  } else {
    if constexpr (SMALL) finalize_small(t.sm, t.j, t.R, t.early_exit, t.vs);
    else finalize_body(t.sm, t.j, t.R, t.early_exit, t.vs);
    *t.next_flag = t.sm->broken;
  }
}

// Elements per lane per access: 16-byte loads of the working vectors (2
// doubles or 4 floats); a basis stored in FP16 is read 4 halves (8 bytes)
// at a time alongside FP32 working vectors.
template <typename TV> struct Vec;
template <> struct Vec<double> { static constexpr int N = 2; };
template <> struct Vec<float> { static constexpr int N = 4; };
template <> struct Vec<_Float16> { static constexpr int N = 4; };

template <int Bytes> struct Raw;
template <> struct Raw<16> { using T = uint4; };
template <> struct Raw<8> { using T = uint2; };

// Basis storage scale: a basis kept in FP16 holds sqrt(n) * v (entries O(1)
// for a unit vector, far from FP16 overflow and underflow); other bases
// hold v itself.
template <typename TB>
inline double basis_scale(int64_t n) {
  return sizeof(TB) == 2 ? std::sqrt((double)n) : 1.0;
}

template <int N, typename T>
__device__ __forceinline__ void vload(const T *__restrict__ p, int64_t q, T (&a)[N]) {
  using R = typename Raw<sizeof(T) * N>::T;
  union { R v; T e[N]; } u;
  u.v = reinterpret_cast<const R *>(p)[q];
#pragma unroll
  for (int k = 0; k < N; ++k) a[k] = u.e[k];
}
template <int N, typename T>
__device__ __forceinline__ void vstore(T *__restrict__ p, int64_t q, const T (&a)[N]) {
  using R = typename Raw<sizeof(T) * N>::T;
  union { R v; T e[N]; } u;
#pragma unroll
  for (int k = 0; k < N; ++k) u.e[k] = a[k];
  reinterpret_cast<R *>(p)[q] = u.v;
}

// Vectors are visited in groups of 8 whose loads carry no branches (indices
// past nv are clamped to the last vector and their contribution masked), so
// each wave keeps 8 x 1 KB in flight per group.  Group loops are runtime
// loops, which keeps register use (and so occupancy) independent of nv.
constexpr int kGrp = 8;

// Uniform early exit of a pass: skip when *stop != 0, or when a gate exists
// and *gate == 0 (the second Gram-Schmidt pass).
__device__ __forceinline__ bool pass_skipped(const int *stop, const int *gate) {
  return (stop && *stop) || (gate && !*gate);
}

// Inner FP16 Arnoldi steps whose operator reads the stored basis column:
// `scale[v]` replaces the common basis scale vinv for basis vector v, and
// the update pass stores the orthogonalised vector, pre-scaled by
// vs / sqrt(*wn2) (wn2: its squared norm before the projection), into the
// basis column `next` instead of writing it back to w.
template <typename TB> struct BasisOut {
  const double *scale = nullptr;
  TB *next = nullptr;
  const double *wn2 = nullptr;
  double vs = 1.0;
};

// partial[v][block] = vinv * sum over the block's elements of V[v] . w for
// every v in [v0, v0+8) of every group, plus (slot nv) w . w.  FP64
// accumulation; vinv undoes the basis storage scale.
template <typename TB, typename TV>
__device__ __forceinline__ void dots_body(const TB *__restrict__ V, int64_t ldv, int nv,
                                          const TV *__restrict__ w, int64_t n,
                                          double *__restrict__ partials, double vinv,
                                          const int *__restrict__ stop,
                                          const int *__restrict__ gate, int blk, int nblk,
                                          const double *__restrict__ vsc = nullptr) {
  if (pass_skipped(stop, gate)) return;
  constexpr int N = Vec<TV>::N;
  const int64_t nq = n / N;
  const int64_t stride = (int64_t)nblk * blockDim.x;
  const int64_t t0 = (int64_t)blk * blockDim.x + threadIdx.x;
  const int64_t tail = n - nq * N;
  {  // w . w
    double acc[1] = {0.0};
    for (int64_t q = t0; q < nq; q += stride) {
      TV we[N];
      vload<N>(w, q, we);
#pragma unroll
      for (int k = 0; k < N; ++k) acc[0] += (double)we[k] * (double)we[k];
    }
    if (t0 < tail) acc[0] += (double)w[nq * N + t0] * (double)w[nq * N + t0];
    block_store_slot(acc[0], nv, partials, blk);
  }
#pragma unroll 1
  for (int g = 0; g * kGrp < nv; ++g) {
    double acc[kGrp];
#pragma unroll
    for (int r = 0; r < kGrp; ++r) acc[r] = 0;
    const TB *base[kGrp];
#pragma unroll
    for (int r = 0; r < kGrp; ++r) base[r] = V + (int64_t)min(g * kGrp + r, nv - 1) * ldv;
    for (int64_t q = t0; q < nq; q += stride) {
      TV we[N];
      TB x[kGrp][N];
      vload<N>(w, q, we);
#pragma unroll
      for (int r = 0; r < kGrp; ++r) vload<N>(base[r], q, x[r]);
#pragma unroll
      for (int r = 0; r < kGrp; ++r)
#pragma unroll
        for (int k = 0; k < N; ++k) acc[r] += (double)x[r][k] * (double)we[k];
    }
    if (t0 < tail) {
      const int64_t e = nq * N + t0;
#pragma unroll
      for (int r = 0; r < kGrp; ++r) acc[r] += (double)base[r][e] * (double)w[e];
    }
    const int cnt = min(kGrp, nv - g * kGrp);
#pragma unroll
    for (int r = 0; r < kGrp; ++r) acc[r] *= vsc ? vsc[min(g * kGrp + r, nv - 1)] : vinv;
    block_store_group(acc, g * kGrp, cnt, partials, blk);
  }
}

template <typename TB, typename TV>
__global__ void __launch_bounds__(256) dots_kernel(const TB *__restrict__ V, int64_t ldv, int nv,
                                                   const TV *__restrict__ w, int64_t n,
                                                   double *__restrict__ partials, double vinv,
                                                   const int *__restrict__ stop,
                                                   const int *__restrict__ gate,
                                                   const double *__restrict__ vsc) {
  dots_body(V, ldv, nv, w, n, partials, vinv, stop, gate, blockIdx.x, gridDim.x, vsc);
}

// Single-sweep variant for nv <= 8*G: w is read once, all G groups of
// projections and w . w accumulate together (8*G + 1 FP64 accumulators per
// lane).  Same partial-slot layout as dots_kernel.
template <typename TB, typename TV, int G>
__device__ __forceinline__ void dots1_body(const TB *__restrict__ V, int64_t ldv, int nv,
                                           const TV *__restrict__ w, int64_t n,
                                           double *__restrict__ partials, double vinv,
                                           const int *__restrict__ stop,
                                           const int *__restrict__ gate, int blk, int nblk,
                                           const double *__restrict__ vsc = nullptr) {
  if (pass_skipped(stop, gate)) return;
  constexpr int N = Vec<TV>::N;
  const int64_t nq = n / N;
  const int64_t stride = (int64_t)nblk * blockDim.x;
  const int64_t t0 = (int64_t)blk * blockDim.x + threadIdx.x;
  double acc[G][kGrp];
  double ww = 0;
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int r = 0; r < kGrp; ++r) acc[g][r] = 0;
  for (int64_t q = t0; q < nq; q += stride) {
    TV we[N];
    vload<N>(w, q, we);
#pragma unroll
    for (int k = 0; k < N; ++k) ww += (double)we[k] * (double)we[k];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      TB x[kGrp][N];
#pragma unroll
      for (int r = 0; r < kGrp; ++r) vload<N>(V + (int64_t)min(g * kGrp + r, nv - 1) * ldv, q, x[r]);
#pragma unroll
      for (int r = 0; r < kGrp; ++r)
#pragma unroll
        for (int k = 0; k < N; ++k) acc[g][r] += (double)x[r][k] * (double)we[k];
    }
  }
  if (t0 < n - nq * N) {
    const int64_t e = nq * N + t0;
    const double we = (double)w[e];
    ww += we * we;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int r = 0; r < kGrp; ++r) acc[g][r] += (double)V[(int64_t)min(g * kGrp + r, nv - 1) * ldv + e] * we;
  }
  block_store_slot(ww, nv, partials, blk);
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (g * kGrp >= nv) break;
#pragma unroll
    for (int r = 0; r < kGrp; ++r) acc[g][r] *= vsc ? vsc[min(g * kGrp + r, nv - 1)] : vinv;
    block_store_group(acc[g], g * kGrp, min(kGrp, nv - g * kGrp), partials, blk);
  }
}

template <typename TB, typename TV, int G>
__global__ void __launch_bounds__(256) dots1_kernel(const TB *__restrict__ V, int64_t ldv, int nv,
                                                    const TV *__restrict__ w, int64_t n,
                                                    double *__restrict__ partials, double vinv,
                                                    const int *__restrict__ stop,
                                                    const int *__restrict__ gate,
                                                    const double *__restrict__ vsc) {
  dots1_body<TB, TV, G>(V, ldv, nv, w, n, partials, vinv, stop, gate, blockIdx.x, gridDim.x, vsc);
}

// Exactly NV basis vectors (no clamped group members): the same sums in the
// same order as dots1_body, so the same bits, with no redundant loads of the
// last vector -- faster than the group of 8 for short Arnoldi prefixes.
template <typename TB, typename TV, int NV>
__global__ void __launch_bounds__(256) dotsx_kernel(const TB *__restrict__ V, int64_t ldv,
                                                    const TV *__restrict__ w, int64_t n,
                                                    double *__restrict__ partials, double vinv,
                                                    const int *__restrict__ stop,
                                                    const int *__restrict__ gate,
                                                    const double *__restrict__ vsc) {
  if (pass_skipped(stop, gate)) return;
  constexpr int N = Vec<TV>::N;
  const int64_t nq = n / N;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc[NV];
  double ww = 0;
#pragma unroll
  for (int r = 0; r < NV; ++r) acc[r] = 0;
  for (int64_t q = t0; q < nq; q += stride) {
    TV we[N];
    vload<N>(w, q, we);
#pragma unroll
    for (int k = 0; k < N; ++k) ww += (double)we[k] * (double)we[k];
    TB x[NV][N];
#pragma unroll
    for (int r = 0; r < NV; ++r) vload<N>(V + (int64_t)r * ldv, q, x[r]);
#pragma unroll
    for (int r = 0; r < NV; ++r)
#pragma unroll
      for (int k = 0; k < N; ++k) acc[r] += (double)x[r][k] * (double)we[k];
  }
  if (t0 < n - nq * N) {
    const int64_t e = nq * N + t0;
    const double we = (double)w[e];
    ww += we * we;
#pragma unroll
    for (int r = 0; r < NV; ++r) acc[r] += (double)V[(int64_t)r * ldv + e] * we;
  }
  block_store_slot(ww, NV, partials, blockIdx.x);
#pragma unroll
  for (int g = 0; g * kGrp < NV; ++g) {
    double a[kGrp];
#pragma unroll
    for (int r = 0; r < kGrp; ++r) {
      const int v = g * kGrp + r;
      const int vc = v < NV ? v : 0;
      a[r] = v < NV ? acc[vc] * (vsc ? vsc[vc] : vinv) : 0.0;
    }
    block_store_group(a, g * kGrp, min(kGrp, NV - g * kGrp), partials, blockIdx.x);
  }
}

// Store of the orthogonalised elements: back into w (slot 0 <- w . w), or
// with `bo.next` pre-scaled into the next basis column (slot 0 <- the
// squared norm of the stored values).
template <int N, typename TB, typename TV>
__device__ __forceinline__ void update_store(TV *__restrict__ w, int64_t q, const double (&s)[N],
                                             const BasisOut<TB> &bo, double pre, double &ww) {
  if (bo.next) {
    TB sb[N];
#pragma unroll
    for (int k = 0; k < N; ++k) sb[k] = (TB)(float)(s[k] * pre);
    vstore<N>(bo.next, q, sb);
#pragma unroll
    for (int k = 0; k < N; ++k) ww += (double)sb[k] * (double)sb[k];
  } else {
    TV sn[N];
#pragma unroll
    for (int k = 0; k < N; ++k) sn[k] = (TV)s[k];
    vstore<N>(w, q, sn);
#pragma unroll
    for (int k = 0; k < N; ++k) ww += (double)sn[k] * (double)sn[k];
  }
}
template <typename TB, typename TV>
__device__ __forceinline__ void update_store1(TV *__restrict__ w, int64_t e, double sd, const BasisOut<TB> &bo,
                                              double pre, double &ww) {
  if (bo.next) {
    const TB sb = (TB)(float)(sd * pre);
    bo.next[e] = sb;
    ww += (double)sb * (double)sb;
  } else {
    const TV sn = (TV)sd;
    w[e] = sn;
    ww += (double)sn * (double)sn;
  }
}
// pre-scale of the next basis column: vs / ||w|| before the projection
template <typename TB>
__device__ __forceinline__ double next_prescale(const BasisOut<TB> &bo) {
  if (!bo.next) return 0.0;
  const double w2 = *bo.wn2;
  return w2 > 0.0 ? bo.vs / sqrt(w2) : 0.0;
}

// w <- w - sum_v h[v] vinv V[v] (h read from device memory), slot 0 <- w . w
// (or, with bo.next, the result stored into the next basis column)
template <typename TB, typename TV>
__device__ __forceinline__ void update_body(const TB *__restrict__ V, int64_t ldv, int nv,
                                            const double *__restrict__ h, TV *__restrict__ w,
                                            int64_t n, double *__restrict__ partials, double vinv,
                                            const int *__restrict__ stop,
                                            const int *__restrict__ gate, int blk, int nblk,
                                            BasisOut<TB> bo = BasisOut<TB>{}) {
  if (pass_skipped(stop, gate)) return;
  constexpr int N = Vec<TV>::N;
  __shared__ double hs[kMaxRestart + 8];
  const int nvr = ((nv + kGrp - 1) / kGrp) * kGrp;
  for (int v = threadIdx.x; v < nvr; v += blockDim.x) hs[v] = v < nv ? h[v] * (bo.scale ? bo.scale[v] : vinv) : 0.0;
  __syncthreads();
  const double pre = next_prescale(bo);
  const int64_t nq = n / N;
  const int64_t stride = (int64_t)nblk * blockDim.x;
  const int64_t t0 = (int64_t)blk * blockDim.x + threadIdx.x;
  double ww = 0;
  for (int64_t q = t0; q < nq; q += stride) {
    TV we[N];
    vload<N>(w, q, we);
    double s[N];
#pragma unroll
    for (int k = 0; k < N; ++k) s[k] = (double)we[k];
#pragma unroll 1
    for (int g = 0; g * kGrp < nv; ++g) {
      TB x[kGrp][N];
#pragma unroll
      for (int r = 0; r < kGrp; ++r) vload<N>(V + (int64_t)min(g * kGrp + r, nv - 1) * ldv, q, x[r]);
#pragma unroll
      for (int r = 0; r < kGrp; ++r) {
        const double hv = hs[g * kGrp + r];  // zero past nv
#pragma unroll
        for (int k = 0; k < N; ++k) s[k] -= hv * (double)x[r][k];
      }
    }
    update_store<N>(w, q, s, bo, pre, ww);
  }
  if (t0 < n - nq * N) {  // scalar tail
    const int64_t e = nq * N + t0;
    double sd = (double)w[e];
    for (int v = 0; v < nv; ++v) sd -= hs[v] * (double)V[(int64_t)v * ldv + e];
    update_store1(w, e, sd, bo, pre, ww);
  }
  block_store_slot(ww, 0, partials, blk);
}

template <typename TB, typename TV>
__global__ void __launch_bounds__(256) update_kernel(const TB *__restrict__ V, int64_t ldv, int nv,
                                                     const double *__restrict__ h,
                                                     TV *__restrict__ w, int64_t n,
                                                     double *__restrict__ partials, double vinv,
                                                     const int *__restrict__ stop,
                                                     const int *__restrict__ gate, BasisOut<TB> bo) {
  update_body(V, ldv, nv, h, w, n, partials, vinv, stop, gate, blockIdx.x, gridDim.x, bo);
}

// update_kernel for exactly NV basis vectors (same bits, no clamped loads).
template <typename TB, typename TV, int NV>
__global__ void __launch_bounds__(256) updatex_kernel(const TB *__restrict__ V, int64_t ldv,
                                                      const double *__restrict__ h,
                                                      TV *__restrict__ w, int64_t n,
                                                      double *__restrict__ partials, double vinv,
                                                      const int *__restrict__ stop,
                                                      const int *__restrict__ gate, BasisOut<TB> bo) {
  if (pass_skipped(stop, gate)) return;
  constexpr int N = Vec<TV>::N;
  __shared__ double hs[NV];
  for (int v = threadIdx.x; v < NV; v += blockDim.x) hs[v] = h[v] * (bo.scale ? bo.scale[v] : vinv);
  __syncthreads();
  const double pre = next_prescale(bo);
  const int64_t nq = n / N;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double ww = 0;
  for (int64_t q = t0; q < nq; q += stride) {
    TV we[N];
    vload<N>(w, q, we);
    double sd[N];
#pragma unroll
    for (int k = 0; k < N; ++k) sd[k] = (double)we[k];
    TB x[NV][N];
#pragma unroll
    for (int r = 0; r < NV; ++r) vload<N>(V + (int64_t)r * ldv, q, x[r]);
#pragma unroll
    for (int r = 0; r < NV; ++r) {
      const double hv = hs[r];
#pragma unroll
      for (int k = 0; k < N; ++k) sd[k] -= hv * (double)x[r][k];
    }
    update_store<N>(w, q, sd, bo, pre, ww);
  }
  if (t0 < n - nq * N) {
    const int64_t e = nq * N + t0;
    double sd = (double)w[e];
    for (int v = 0; v < NV; ++v) sd -= hs[v] * (double)V[(int64_t)v * ldv + e];
    update_store1(w, e, sd, bo, pre, ww);
  }
  block_store_slot(ww, 0, partials, blockIdx.x);
}

// JAX _iterative_classical_gram_schmidt loop condition: a second pass runs
// when ||r|| < ||q||/sqrt(2) after the first (norms below eps count as 0).
__device__ void decide_body(Small *sm, int nv) {
  if (sm->broken) {
    sm->gate2 = 0;
    return;
  }
  double r2 = 0;
  for (int v = 0; v < nv; ++v) r2 += sm->h1[v] * sm->h1[v];
  double rn = sqrt(r2), qn = sqrt(sm->nn1a[0]);
  const double eps = 2.220446049250313e-16;
  if (!(rn > eps)) rn = 0.0;
  if (!(qn > eps)) qn = 0.0;
  sm->gate2 = rn < qn / sqrt(2.0) ? 1 : 0;
}

// Arnoldi bookkeeping for step j (one thread): JAX's H row, breakdown test
// (v_norm_1 <= eps * v_norm_0), Givens residual estimate for early exit.
//
// vs > 0: the update pass stored the orthogonalised vector pre-scaled by
// vs / vn0 as basis column j + 1 and nn1a holds the squared norm q^2 of the
// stored values, so that ||w|| = q vn0 / vs, the unit vector is V[j+1] / q
// (scale 1 / q, which the next operator launch applies to its input).
__device__ void finalize_body(Small *sm, int j, int R, int early_exit, double vs) {
  if (sm->broken) {
    sm->inv = 0.0;
    return;
  }
  const int two = sm->gate2;
  double h[kMaxRestart + 2];
  for (int v = 0; v <= j; ++v) h[v] = sm->h1[v] + (two ? sm->h2[v] : 0.0);
  const double vn0 = sqrt(sm->h1[j + 1]);  // ||w||^2 before projection sits in slot nv
  const double q = vs > 0.0 ? sqrt(sm->nn1a[0]) : 0.0;
  double vn1 = vs > 0.0 ? q * vn0 / vs : sqrt(two ? sm->nn1b[0] : sm->nn1a[0]);
  const double thresh = 2.220446049250313e-16 * vn0;
  if (!(vn1 > thresh)) vn1 = 0.0;
  h[j + 1] = vn1;
  double *row = sm->H + (int64_t)j * (R + 1);
  for (int v = 0; v <= R; ++v) row[v] = v <= j + 1 ? h[v] : 0.0;
  if (vs > 0.0) {
    sm->vsc[j + 1] = vn1 > 0.0 ? 1.0 / q : 0.0;
    sm->inv = sm->vsc[j + 1];
  } else {
    sm->inv = vn1 > 0.0 ? 1.0 / vn1 : 0.0;
  }
  sm->steps = j + 1;
  if (vn1 == 0.0) sm->broken = 1;
  // Givens rotations on a copy of the new column
  for (int v = 0; v < j; ++v) {
    const double a = sm->cs[v] * h[v] + sm->sn[v] * h[v + 1];
    const double b = -sm->sn[v] * h[v] + sm->cs[v] * h[v + 1];
    h[v] = a;
    h[v + 1] = b;
  }
  const double r = hypot(h[j], h[j + 1]);
  const double c = r > 0.0 ? h[j] / r : 1.0, s = r > 0.0 ? h[j + 1] / r : 0.0;
  sm->cs[j] = c;
  sm->sn[j] = s;
  sm->g[j + 1] = -s * sm->g[j];
  sm->g[j] = c * sm->g[j];
  if (early_exit && fabs(sm->g[j + 1]) <= sm->target) sm->broken = 1;
}

__device__ void finalize_small(Small *sm, int j, int R, int early_exit, double vs) {
  constexpr int HM = kSmallRestart + 2;
  if (sm->broken) {
    sm->inv = 0.0;
    return;
  }
  const int two = sm->gate2;
  double h[HM], cs[HM], sn[HM];
#pragma unroll
  for (int v = 0; v < HM; ++v) {
    h[v] = v <= j ? sm->h1[v] + (two ? sm->h2[v] : 0.0) : 0.0;
    cs[v] = v < j ? sm->cs[v] : 0.0;
    sn[v] = v < j ? sm->sn[v] : 0.0;
  }
  const double vn0 = sqrt(sm->h1[j + 1]);
  const double q = vs > 0.0 ? sqrt(sm->nn1a[0]) : 0.0;
  double vn1 = vs > 0.0 ? q * vn0 / vs : sqrt(two ? sm->nn1b[0] : sm->nn1a[0]);
  const double thresh = 2.220446049250313e-16 * vn0;
  if (!(vn1 > thresh)) vn1 = 0.0;
#pragma unroll
  for (int v = 0; v < HM; ++v)
    if (v == j + 1) h[v] = vn1;
  double *row = sm->H + (int64_t)j * (R + 1);
#pragma unroll
  for (int v = 0; v < HM; ++v)
    if (v <= R) row[v] = v <= j + 1 ? h[v] : 0.0;
  if (vs > 0.0) {
    const double sc = vn1 > 0.0 ? 1.0 / q : 0.0;
    sm->vsc[j + 1] = sc;
    sm->inv = sc;
  } else {
    sm->inv = vn1 > 0.0 ? 1.0 / vn1 : 0.0;
  }
  sm->steps = j + 1;
  int broken = vn1 == 0.0 ? 1 : 0;
  // Givens rotations on a copy of the new column
#pragma unroll
  for (int v = 0; v < HM - 1; ++v)
    if (v < j) {
      const double a = cs[v] * h[v] + sn[v] * h[v + 1];
      const double b = -sn[v] * h[v] + cs[v] * h[v + 1];
      h[v] = a;
      h[v + 1] = b;
    }
  double hj = 0.0, hj1 = 0.0;
#pragma unroll
  for (int v = 0; v < HM; ++v) {
    if (v == j) hj = h[v];
    if (v == j + 1) hj1 = h[v];
  }
  const double r = hypot(hj, hj1);
  const double c = r > 0.0 ? hj / r : 1.0, s = r > 0.0 ? hj1 / r : 0.0;
  sm->cs[j] = c;
  sm->sn[j] = s;
  const double gj = sm->g[j];
  const double g1 = -s * gj;
  sm->g[j + 1] = g1;
  sm->g[j] = c * gj;
  if (early_exit && fabs(g1) <= sm->target) broken = 1;
  if (broken) sm->broken = 1;
}

// V[j+1] = vscale * w * inv (basis storage) and, when `cur` is given, the
// unscaled working copy cur = w * inv that the next operator apply reads;
// nothing once the cycle has stopped (those columns are outside the least
// squares).
template <typename TB, typename TV>
__device__ __forceinline__ void normalize_gated_body(const TV *__restrict__ w, TB *__restrict__ out,
                                                     TV *__restrict__ cur, int64_t n, double vscale,
                                                     const Small *__restrict__ sm,
                                                     const int *__restrict__ stop_before, int blk,
                                                     int nblk) {
  if (*stop_before) return;
  const double inv = sm->inv;
  constexpr int N = Vec<TV>::N;
  const int64_t nq = n / N;
  const int64_t stride = (int64_t)nblk * blockDim.x;
  const int64_t t0 = (int64_t)blk * blockDim.x + threadIdx.x;
  for (int64_t q = t0; q < nq; q += stride) {
    TV a[N];
    TB o[N];
    vload<N>(w, q, a);
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const double u = inv != 0.0 ? (double)a[k] * inv : 0.0;
      a[k] = (TV)u;
      o[k] = (TB)(u * vscale);
    }
    vstore<N>(out, q, o);
    if (cur) vstore<N>(cur, q, a);
  }
  if (t0 < n - nq * N) {
    const int64_t e = nq * N + t0;
    const double u = inv != 0.0 ? (double)w[e] * inv : 0.0;
    out[e] = (TB)(u * vscale);
    if (cur) cur[e] = (TV)u;
  }
}

template <typename TB, typename TV>
__global__ void normalize_gated_kernel(const TV *__restrict__ w, TB *__restrict__ out,
                                       TV *__restrict__ cur, int64_t n, double vscale,
                                       const Small *__restrict__ sm,
                                       const int *__restrict__ stop_before) {
  normalize_gated_body(w, out, cur, n, vscale, sm, stop_before, blockIdx.x, gridDim.x);
}

// snapshot of the stop flag taken before a step's finalize

// u = src / ||src|| safe-normalised (JAX _safe_normalize, thresh eps):
// out = vscale * u (basis storage), cur = u when given.
template <typename TS, typename TB, typename TC>
__global__ void unit_kernel(const TS *__restrict__ src, TB *__restrict__ out, TC *__restrict__ cur,
                            int64_t n, double vscale, const double *__restrict__ norm2) {
  const double nr = sqrt(*norm2);
  const double inv = nr > 2.220446049250313e-16 ? 1.0 / nr : 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    const double u = inv != 0.0 ? (double)src[e] * inv : 0.0;
    out[e] = (TB)(u * vscale);
    if (cur) cur[e] = (TC)u;
  }
}

// ||w||^2 (FP64, fixed-order partials into slot 0) and the FP32 copy of w:
// the unnormalised start vector of an inner cycle, which the cycle's first
// operator launch normalises (one read of w instead of a norm pass and a
// separate normalisation pass).
__device__ __forceinline__ void norm2_to_f32_body(const double *__restrict__ w, float *__restrict__ out,
                                                  int64_t n, double *__restrict__ partials, int blk, int nblk) {
  const int64_t stride = (int64_t)nblk * blockDim.x;
  double acc = 0;
  const int64_t nq = n / 2;
  for (int64_t q = (int64_t)blk * blockDim.x + threadIdx.x; q < nq; q += stride) {
    double e[2];
    vload<2>(w, q, e);
    acc += e[0] * e[0] + e[1] * e[1];
    float2 f;
    f.x = (float)e[0];
    f.y = (float)e[1];
    reinterpret_cast<float2 *>(out)[q] = f;
  }
  if (blk == 0 && threadIdx.x == 0 && (n & 1)) {
    const double v = w[n - 1];
    acc += v * v;
    out[n - 1] = (float)v;
  }
  block_store_slot(acc, 0, partials, blk);
}
__global__ void __launch_bounds__(256) norm2_to_f32_kernel(const double *__restrict__ w,
                                                           float *__restrict__ out, int64_t n,
                                                           double *__restrict__ partials) {
  norm2_to_f32_body(w, out, n, partials, blockIdx.x, gridDim.x);
}

// x += sum_v y[v] vinv V[v]   (x is FP64)
template <typename TV>
__device__ __forceinline__ void combine_body(const TV *__restrict__ V, int64_t ldv, int nv,
                                             const double *__restrict__ y, double *__restrict__ x,
                                             int64_t n, double vinv, int blk, int nblk,
                                             const double *__restrict__ vsc = nullptr) {
  constexpr int N = Vec<TV>::N;
  __shared__ double ys[kMaxRestart + 8];
  const int nvr = ((nv + kGrp - 1) / kGrp) * kGrp;
  for (int v = threadIdx.x; v < nvr; v += blockDim.x) ys[v] = v < nv ? y[v] * (vsc ? vsc[v] : vinv) : 0.0;
  __syncthreads();
  const int64_t nq = n / N;
  const int64_t stride = (int64_t)nblk * blockDim.x;
  const int64_t t0 = (int64_t)blk * blockDim.x + threadIdx.x;
  for (int64_t q = t0; q < nq; q += stride) {
    double s[N];
#pragma unroll
    for (int k = 0; k < N; ++k) s[k] = 0;
#pragma unroll 1
    for (int g = 0; g * kGrp < nv; ++g) {
      TV xv[kGrp][N];
#pragma unroll
      for (int r = 0; r < kGrp; ++r) vload<N>(V + (int64_t)min(g * kGrp + r, nv - 1) * ldv, q, xv[r]);
#pragma unroll
      for (int r = 0; r < kGrp; ++r)
#pragma unroll
        for (int k = 0; k < N; ++k) s[k] += ys[g * kGrp + r] * (double)xv[r][k];
    }
#pragma unroll
    for (int k = 0; k < N; ++k) x[q * N + k] += s[k];
  }
  if (t0 < n - nq * N) {
    const int64_t e = nq * N + t0;
    double sd = 0;
    for (int v = 0; v < nv; ++v) sd += ys[v] * (double)V[(int64_t)v * ldv + e];
    x[e] += sd;
  }
}

// nvp (optional): the number of basis vectors, read from device memory
template <typename TV>
__global__ void __launch_bounds__(256) combine_kernel(const TV *__restrict__ V, int64_t ldv, int nv,
                                                      const double *__restrict__ y,
                                                      double *__restrict__ x, int64_t n,
                                                      double vinv, const int *__restrict__ nvp,
                                                      const double *__restrict__ vsc) {
  combine_body(V, ldv, nvp ? *nvp : nv, y, x, n, vinv, blockIdx.x, gridDim.x, vsc);
}

// Least squares of one cycle, min || beta e1 - H^T y || over the `steps`
// columns, by the normal equations with partial pivoting (JAX's
// _gmres_batched solve of the small system), by one workgroup: H is staged
// in LDS, each entry of G = H H^T is one thread's dot product (fixed order),
// the elimination updates the rows below the pivot in parallel, the back
// substitution is one thread's.  Same arithmetic as the sequential
// algorithm.  All threads of the block call it.
constexpr int kLsqThreads = 256;
__device__ void lstsq_block(Small *sm, int R, double beta) {
  __shared__ double Hs[kMaxRestart * (kMaxRestart + 1)];
  __shared__ double G[kMaxRestart * kMaxRestart];
  __shared__ double r[kMaxRestart], f[kMaxRestart], y[kMaxRestart];
  __shared__ int piv;
  const int t = threadIdx.x, nt = blockDim.x;
  const int s = sm->steps;
  for (int q = t; q < s * (R + 1); q += nt) Hs[q] = sm->H[q];
  __syncthreads();
  for (int q = t; q < s * s; q += nt) {
    const int a = q / s, b = q % s;
    double acc = 0;
    for (int l = 0; l <= s; ++l) acc += Hs[a * (R + 1) + l] * Hs[b * (R + 1) + l];
    G[q] = acc;
  }
  for (int a = t; a < s; a += nt) r[a] = Hs[a * (R + 1)] * beta;
  __syncthreads();
  for (int c = 0; c < s; ++c) {
    if (t == 0) {
      int p = c;
      for (int q = c + 1; q < s; ++q)
        if (fabs(G[q * s + c]) > fabs(G[p * s + c])) p = q;
      piv = p;
    }
    __syncthreads();
    const int p = piv;
    if (p != c) {
      for (int q = t; q < s; q += nt) {
        const double tq = G[c * s + q];
        G[c * s + q] = G[p * s + q];
        G[p * s + q] = tq;
      }
      if (t == 0) { const double tr = r[c]; r[c] = r[p]; r[p] = tr; }
    }
    __syncthreads();
    const double d = G[c * s + c];
    for (int q = c + 1 + t; q < s; q += nt) f[q] = d != 0.0 ? G[q * s + c] / d : 0.0;
    __syncthreads();
    const int w = s - c;  // columns c..s-1
    for (int e = t; e < (s - c - 1) * w; e += nt) {
      const int q = c + 1 + e / w, k2 = c + e % w;
      if (f[q] != 0.0) G[q * s + k2] -= f[q] * G[c * s + k2];
    }
    for (int q = c + 1 + t; q < s; q += nt)
      if (f[q] != 0.0) r[q] -= f[q] * r[c];
    __syncthreads();
  }
  if (t == 0) {
    for (int c = s - 1; c >= 0; --c) {
      double acc = r[c];
      for (int k2 = c + 1; k2 < s; ++k2) acc -= G[c * s + k2] * y[k2];
      const double d = G[c * s + c];
      y[c] = d != 0.0 ? acc / d : 0.0;
    }
    sm->steps_total += s;
  }
  __syncthreads();
  for (int c = t; c < s; c += nt) sm->y[c] = y[c];
}

// Single-solve driver: the cycle's least squares on the device, so that the
// host does not wait for the Hessenberg matrix (one synchronisation per
// cycle, for the residual norm, remains).
__global__ void __launch_bounds__(kLsqThreads) lstsq_kernel(Small *sm, int R, double beta) {
  lstsq_block(sm, R, beta);
}

__global__ void zero_steps_total_kernel(Small *sm) { sm->steps_total = 0; }

// vinv0: the scale of basis vector 0 (stored by the cycle's first operator
// launch as vs * unit vector)
__global__ void reset_cycle_kernel(Small *sm, int R, double g0, double target, int *flags, double vinv0) {
  const int t = threadIdx.x;
  for (int q = t; q < R * (R + 1); q += blockDim.x) {
    const int r = q / (R + 1), c = q % (R + 1);
    sm->H[q] = r == c ? 1.0 : 0.0;
  }
  if (t == 0) {
    sm->broken = 0;
    sm->steps = 0;
    sm->gate2 = 0;
    sm->inv = g0 > 0.0 ? 1.0 / g0 : 0.0;  // normalises the cycle's start vector (fused first step)
    for (int v = 0; v <= R; ++v) sm->g[v] = 0.0;
    sm->g[0] = g0;
    sm->target = target;
    sm->vsc[0] = vinv0;
    flags[0] = 0;  // stop flag of step 0
  }
}

// ---- host-side dispatch -------------------------------------------------------

static void launch_tail(const double *partials, const Tail *tl, const int *stop, const int *gate, hipStream_t s) {
  if (!tl) return;
  if (tl->R <= kSmallRestart)
    hipLaunchKernelGGL(tail_kernel<true>, dim3(1), dim3(kTailThreads), 0, s, partials, *tl, stop, gate);
  else
    hipLaunchKernelGGL(tail_kernel<false>, dim3(1), dim3(kTailThreads), 0, s, partials, *tl, stop, gate);
}

template <typename TB, typename TV>
static void dots(const TB *V, int64_t ldv, int nv, const TV *w, int64_t n, double *partials,
                 hipStream_t s, double vinv = 1.0, const int *stop = nullptr,
                 const int *gate = nullptr, const Tail *tl = nullptr, const double *vsc = nullptr) {
  const double bytes = (double)n * ((double)nv * sizeof(TB) + sizeof(TV));
  const int tid = timer_begin(SISL_T_KRYLOV_DOTS, sizeof(TV) == 4, bytes, stop, gate, s);
  const dim3 gr(kRedGrid), bl(kBlock);
  // exact-count variants where they measured faster (short prefixes, and
  // 9 / 10 vectors, which the two-group variant would load as 16)
#define SISL_DOTSX(K) hipLaunchKernelGGL((dotsx_kernel<TB, TV, K>), gr, bl, 0, s, V, ldv, w, n, partials, vinv, stop, gate, vsc)
  if (nv <= 0)
    hipLaunchKernelGGL((dots_kernel<TB, TV>), gr, bl, 0, s, V, ldv, nv, w, n, partials, vinv, stop, gate, vsc);
  else if (nv == 1) SISL_DOTSX(1);
  else if (nv == 2) SISL_DOTSX(2);
  else if (nv == 3) SISL_DOTSX(3);
  else if (nv == 4) SISL_DOTSX(4);
  else if (nv == 5) SISL_DOTSX(5);  // 64 / 75 us against 82 us of the group of 8
  else if (nv == 6) SISL_DOTSX(6);  // (7 and 8 exact: 87 / 96 us, slower)
  else if (nv == 9) SISL_DOTSX(9);
  else if (nv == 10) SISL_DOTSX(10);
  else if (nv <= kGrp)
    hipLaunchKernelGGL((dots1_kernel<TB, TV, 1>), gr, bl, 0, s, V, ldv, nv, w, n, partials, vinv, stop, gate, vsc);
  else if (nv <= 2 * kGrp)
    hipLaunchKernelGGL((dots1_kernel<TB, TV, 2>), gr, bl, 0, s, V, ldv, nv, w, n, partials, vinv, stop, gate, vsc);
  else if (nv <= 3 * kGrp)
    hipLaunchKernelGGL((dots1_kernel<TB, TV, 3>), gr, bl, 0, s, V, ldv, nv, w, n, partials, vinv, stop, gate, vsc);
  else if (nv <= 4 * kGrp)
    hipLaunchKernelGGL((dots1_kernel<TB, TV, 4>), gr, bl, 0, s, V, ldv, nv, w, n, partials, vinv, stop, gate, vsc);
  else
    hipLaunchKernelGGL((dots_kernel<TB, TV>), gr, bl, 0, s, V, ldv, nv, w, n, partials, vinv, stop, gate, vsc);
  timer_end(tid, s);
  launch_tail(partials, tl, stop, gate, s);
}
template <typename TB, typename TV>
static void update(const TB *V, int64_t ldv, int nv, const double *h, TV *w, int64_t n,
                   double *partials, hipStream_t s, double vinv = 1.0, const int *stop = nullptr,
                   const int *gate = nullptr, const Tail *tl = nullptr, BasisOut<TB> bo = BasisOut<TB>{}) {
  const double bytes = (double)n * ((double)nv * sizeof(TB) + sizeof(TV) + (bo.next ? sizeof(TB) : sizeof(TV)));
  const int tid = timer_begin(SISL_T_KRYLOV_UPDATE, sizeof(TV) == 4, bytes, stop, gate, s);
#define SISL_UPDX(K) hipLaunchKernelGGL((updatex_kernel<TB, TV, K>), dim3(kRedGrid), dim3(kBlock), 0, s, V, ldv, h, w, n, partials, vinv, stop, gate, bo)
  if (nv == 1) SISL_UPDX(1);
  else if (nv == 2) SISL_UPDX(2);
  else if (nv == 3) SISL_UPDX(3);
  else if (nv == 4) SISL_UPDX(4);
  else if (nv == 5) SISL_UPDX(5);   // (6 to 8 exact: 88 / 100 / 105 us, slower than the group form)
  else if (nv == 9) SISL_UPDX(9);    // the two-group form would load 16 vectors
  else if (nv == 10) SISL_UPDX(10);
  else
    hipLaunchKernelGGL((update_kernel<TB, TV>), dim3(kRedGrid), dim3(kBlock), 0, s, V, ldv, nv, h, w,
                       n, partials, vinv, stop, gate, bo);
  timer_end(tid, s);
  launch_tail(partials, tl, stop, gate, s);
}
template <typename TB>
static void combine(const TB *V, int64_t ldv, int nv, const double *y, double *x, int64_t n,
                    double vinv, hipStream_t s, const int *nvp = nullptr, const double *vsc = nullptr) {
  hipLaunchKernelGGL(combine_kernel<TB>, dim3(kRedGrid), dim3(kBlock), 0, s, V, ldv, nv, y, x, n,
                     vinv, nvp, vsc);
}
static void reduce(const double *partials, int nslots, double *out, hipStream_t s,
                   const int *stop = nullptr, const int *gate = nullptr) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(nslots), dim3(kBlock), 0, s, partials, out, stop, gate);
}

static inline int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }
// leading dimension of the basis: rows start 256-byte aligned for vector loads
static inline int64_t basis_ld(int64_t n) { return (n + 63) & ~(int64_t)63; }

// Workspace carve-up (byte offsets).
struct Layout {
  int64_t V, w, t, scr, partials, small, flags, V32, w32, t32, scr32, cur32, total;
};

static Layout layout(const Stag &s, int R, bool mixed) {
  Layout L{};
  int64_t o = 0;
  const int64_t n = s.n, nm = s.m.n();
  auto take = [&](int64_t bytes) { const int64_t r = o; o += align256(bytes); return r; };
  L.small = take(sizeof(Small));
  L.flags = take(sizeof(int) * (kMaxRestart + 2));
  L.partials = take(sizeof(double) * kPStride * kRedGrid);
  L.w = take(sizeof(double) * basis_ld(n));
  L.t = take(sizeof(double) * basis_ld(n));
  L.scr = take(sizeof(double) * 2 * nm);
  if (!mixed) {
    L.V = take(sizeof(double) * basis_ld(n) * (R + 1));
  } else {
    L.V32 = take(sizeof(float) * basis_ld(n) * (R + 1));  // FP32 basis, or FP16 in its first half
    L.cur32 = take(sizeof(float) * basis_ld(n));
    L.w32 = take(sizeof(float) * basis_ld(n));
    L.t32 = take(sizeof(float) * basis_ld(n));
    L.scr32 = take(sizeof(float) * 2 * nm);
  }
  L.total = o;
  return L;
}

// Least squares of JAX's _lstsq (normal equations) on the leading s steps.
// One restart cycle's Arnoldi process.  TV: type of the operator applies and
// working vectors; TB: basis storage type (TV, or FP16 under the FP32
// working type).  With an FP16 basis, `cur` holds the newest basis vector
// unscaled in TV for the next operator apply.
template <typename TV, typename TB = TV>
struct Krylov {
  const OpT<TV> &op;
  int transpose;
  TB *V;
  TV *w, *t, *scr;
  double *partials;
  Small *sm;
  int *flags;   // flags[j]: stop flag snapshot before step j
  int64_t n;
  hipStream_t s;
  TV *cur = nullptr;
  int cb = 0;  // 1: the operator's 3-D coefficient arrays are bfloat16
  int single_cgs = 0;  // 1: one Gram-Schmidt pass per step (no second pass)
  // 1 (FP16 basis, one Gram-Schmidt pass): steps after the first apply the
  // operator to the stored FP16 basis column, and each step's update pass
  // stores the next column (with its own scale) instead of writing w back
  int in16 = 0;
  int64_t ldv() const { return basis_ld(n); }
  double vscale() const { return basis_scale<TB>(n); }
  const TV *basis_in(int j) const {
    if constexpr (sizeof(TB) == sizeof(TV)) return reinterpret_cast<const TV *>(V) + (int64_t)j * ldv();
    else return cur;
  }

  // out <- (M A) in  (or M^T A^T); kernels skip when *stop != 0; `fuse`:
  // the operator launch also normalises `in` into a basis column
  int apply(const TV *in, TV *out, const int *stop, const Fuse *fuse = nullptr) {
    int rc;
    Fuse f = fuse ? *fuse : Fuse{};
    f.cb = cb;
    if (!transpose) {
      if ((rc = launch_A<TV>(op, in, out, s, stop, &f))) return rc;
      return launch_M<TV>(op, out, out, s, stop, cb);
    }
    if ((rc = launch_AT<TV>(op, in, t, scr, s, stop, &f))) return rc;
    return launch_MT<TV>(op, t, out, scr, s, stop, cb);
  }

  // One Arnoldi step (JAX _kth_arnoldi_iteration) with JAX's iterated CGS.
  //
  // FP32 inner solver: the normalisation v_j = w_{j-1} / h_{j,j-1} is fused
  // into the operator launch of step j (which reads the unnormalised w_{j-1},
  // scales its outputs and stores basis column j), and the Arnoldi vector
  // alternates between the buffers w (even j) and cur (odd j).  Step 0 reads
  // the unnormalised residual that starts the cycle.  With in16, step j > 0
  // reads basis column j itself (stored by step j-1's update pass, unit
  // after the scale vsc[j]) and the Arnoldi vector always sits in w, so the
  // operator applies to exactly the basis vector the projections and the
  // solution update use.  FP64 solver: a separate normalisation pass after
  // the step.
  int arnoldi(int j, int R, int early_exit) {
    int rc;
    const int64_t ld = ldv();
    int *stop = flags + j;
    const double vs = vscale(), vinv = 1.0 / vs;
    constexpr bool fused = sizeof(TV) == 4;
    TV *wj = w;
    if constexpr (fused) {
      // step 0 reads the cycle's unnormalised start vector (cur), later
      // steps the previous unnormalised Arnoldi vector; every launch scales
      // by 1 / norm and stores basis column j
      wj = (in16 || j % 2 == 0) ? w : cur;
      Fuse f;
      f.inv = &sm->inv;
      if (in16 && j > 0) {
        f.nb = 4;  // input: basis column j in FP16, scaled by *inv = vsc[j]
        if ((rc = apply(reinterpret_cast<const TV *>(V + (int64_t)j * ld), wj, stop, &f))) return rc;
      } else {
        const TV *src = j % 2 == 0 ? cur : w;
        f.vout = V + (int64_t)j * ld;
        f.nb = sizeof(TB) == 2 ? 1 : 2;
        f.vs = vs;
        if ((rc = apply(src, wj, stop, &f))) return rc;
      }
    } else {
      if ((rc = apply(basis_in(j), w, stop))) return rc;
    }
    const int nv = j + 1;
    const int *gate = &sm->gate2;
    // each pass is followed by one tail launch: its reduction and the bookkeeping after it
    Tail t{sm->h1, nv + 1, 0, sm, nv, j, R, early_exit, flags + j + 1};
    const double *vsc = in16 ? sm->vsc : nullptr;
    dots<TB, TV>(V, ld, nv, wj, n, partials, s, vinv, stop, nullptr, &t, vsc);  // h1[0..nv-1], ||w||^2 in h1[nv]
    if (single_cgs) {
      // one pass: gate2 stays 0 from the cycle reset, so the bookkeeping reads h1, nn1a
      t.out = sm->nn1a; t.nout = 1; t.mode = 2;
      BasisOut<TB> bo;
      if (in16) {
        bo.scale = sm->vsc;
        bo.next = V + (int64_t)(j + 1) * ld;
        bo.wn2 = &sm->h1[nv];
        bo.vs = vs;
        t.vs = vs;
      }
      update<TB, TV>(V, ld, nv, sm->h1, wj, n, partials, s, vinv, stop, nullptr, &t, bo);
    } else {
      t.out = sm->nn1a; t.nout = 1; t.mode = 1;                             // + second-pass decision
      update<TB, TV>(V, ld, nv, sm->h1, wj, n, partials, s, vinv, stop, nullptr, &t);
      t.out = sm->h2; t.nout = nv; t.mode = 0;
      dots<TB, TV>(V, ld, nv, wj, n, partials, s, vinv, stop, gate, &t);
      t.out = sm->nn1b; t.nout = 1; t.mode = 2;                             // + Arnoldi bookkeeping
      update<TB, TV>(V, ld, nv, sm->h2, wj, n, partials, s, vinv, stop, gate, &t);
    }
    if constexpr (fused) {
      SISL_CHECK_LAUNCH();
      return SISL_OK;
    }
    const double nb = (double)n * (sizeof(TV) + sizeof(TB) + (cur ? sizeof(TV) : 0));
    const int tid = timer_begin(SISL_T_KRYLOV_NORMALIZE, sizeof(TV) == 4, nb, stop, nullptr, s);
    hipLaunchKernelGGL((normalize_gated_kernel<TB, TV>), dim3(kRedGrid), dim3(kBlock), 0, s,
                       (const TV *)w, V + (int64_t)(j + 1) * ld, cur, n, vs, (const Small *)sm,
                       (const int *)stop);
    timer_end(tid, s);
    SISL_CHECK_LAUNCH();
    return SISL_OK;
  }
};

// squared norm of a vector into *out (device)
template <typename TV>
static void norm2(const TV *x, int64_t n, double *partials, double *out, hipStream_t s) {
  dots<TV, TV>(x, n, 0, x, n, partials, s);
  reduce(partials, 1, out, s);
}

static int sync_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) != hipSuccess) return SISL_ERR_LAUNCH;
  if (hipStreamSynchronize(s) != hipSuccess) return SISL_ERR_LAUNCH;
  return SISL_OK;
}

// FP64 GMRES (reference semantics unless early_exit).
static int gmres64(const OpT<double> &op, const SislGmresParams &p, const double *b, double *x,
                   char *ws, SislGmresStats *st, hipStream_t s) {
  const int R = p.restart;
  const Layout L = layout(op.s, R, false);
  Krylov<double> K{op, p.transpose, (double *)(ws + L.V), (double *)(ws + L.w), (double *)(ws + L.t),
                   (double *)(ws + L.scr), (double *)(ws + L.partials), (Small *)(ws + L.small),
                   (int *)(ws + L.flags), op.s.n, s};
  const int64_t n = op.s.n;
  Small *sm = K.sm;
  double hs[4];
  int rc;
  // ||b||, ||M b||  (jax gmres: atol = max(tol*||b||, atol); ptol = ||Mb|| min(1, atol/||b||))
  norm2<double>(b, n, K.partials, &sm->scal[0], s);
  hipMemcpyAsync(K.w, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
  if (!p.transpose) {
    if ((rc = launch_M<double>(op, K.w, K.w, s))) return rc;
    norm2<double>(K.w, n, K.partials, &sm->scal[1], s);
  } else {
    if ((rc = launch_MT<double>(op, K.w, K.t, K.scr, s))) return rc;
    norm2<double>(K.t, n, K.partials, &sm->scal[1], s);
  }
  // residual M (b - A x), unit vector into V[0]
  Fuse rb;  // the operator launch stores b - A x
  rb.res = b;
  auto residual = [&](double *scal_slot) -> int {
    int r2;
    if (!p.transpose) {
      if ((r2 = launch_A<double>(op, x, K.w, s, nullptr, &rb))) return r2;
      if ((r2 = launch_M<double>(op, K.w, K.w, s))) return r2;
    } else {
      if ((r2 = launch_AT<double>(op, x, K.t, K.scr, s, nullptr, &rb))) return r2;
      if ((r2 = launch_MT<double>(op, K.t, K.w, K.scr, s))) return r2;
    }
    norm2<double>(K.w, n, K.partials, scal_slot, s);
    hipLaunchKernelGGL((unit_kernel<double, double, double>), dim3(kRedGrid), dim3(kBlock), 0, s,
                       (const double *)K.w, K.V, (double *)nullptr, n, 1.0,
                       (const double *)scal_slot);
    SISL_CHECK_LAUNCH();
    return SISL_OK;
  };
  if ((rc = residual(&sm->scal[2]))) return rc;
  if ((rc = sync_copy(hs, sm->scal, sizeof(double) * 3, s))) return rc;
  const double b_norm = std::sqrt(hs[0]);
  const double Mb_norm = std::sqrt(hs[1]);
  const double atol = std::fmax(p.tol * b_norm, p.atol);
  const double ptol = Mb_norm * nan_min1(atol / b_norm);  // NaN when b = 0 and atol = 0: no cycle
  double rnorm = std::sqrt(hs[2]);
  if (!(rnorm > 2.220446049250313e-16) && !std::isnan(rnorm)) rnorm = 0.0;
  hipLaunchKernelGGL(zero_steps_total_kernel, dim3(1), dim3(1), 0, s, sm);
  int k = 0;
  while (k < p.maxiter && rnorm > ptol) {
    hipLaunchKernelGGL(reset_cycle_kernel, dim3(1), dim3(kBlock), 0, s, sm, R, rnorm,
                       p.early_exit ? ptol : 0.0, K.flags, 1.0);
    for (int j = 0; j < R; ++j)
      if ((rc = K.arnoldi(j, R, p.early_exit))) return rc;
    hipLaunchKernelGGL(lstsq_kernel, dim3(1), dim3(kLsqThreads), 0, s, sm, R, rnorm);
    combine<double>(K.V, K.ldv(), R, sm->y, x, n, 1.0, s, &sm->steps);
    if ((rc = residual(&sm->scal[2]))) return rc;
    if ((rc = sync_copy(hs, sm->scal, sizeof(double) * 3, s))) return rc;
    rnorm = std::sqrt(hs[2]);
    if (!(rnorm > 2.220446049250313e-16) && !std::isnan(rnorm)) rnorm = 0.0;
    ++k;
  }
  if (st) {
    int steps_total = 0;
    if ((rc = sync_copy(&steps_total, &sm->steps_total, sizeof(int), s))) return rc;
    st->cycles = k;
    st->arnoldi_steps = steps_total;
    st->residual = rnorm;
    st->ptol = ptol;
    st->b_norm = b_norm;
  }
  return std::isfinite(rnorm) ? SISL_OK : SISL_ERR_NOT_FINITE;
}

// ---- hipGraph of the inner Arnoldi cycle ----------------------------------------
// One inner cycle is R Arnoldi steps of ~10 launches each with fixed
// arguments (workspace and coefficient pointers, step index), gated on the
// device by the cycle's stop flags and free of host synchronisation.  It is
// captured once -- on a private stream of default priority, since the
// caller's stream may be the legacy default stream -- and replayed on the
// caller's stream for every later cycle with the same workspace, operator,
// restart length, direction, coefficient storage and operator launch shapes
// (a change of shapes retires every cached graph).  The captured launches
// are the ones the eager path makes (no memory-set nodes); with launch timing
// on, the eager path runs so that every launch is timed.
static std::mutex g_graph_mu;
static bool g_graphs_on = true;
// A cached graph and an event recorded after its latest replay.
struct CachedGraph {
  hipGraphExec_t exec;
  hipEvent_t done;
};
static std::map<std::string, CachedGraph> g_graph_cache;
static uint64_t g_graph_generation = 0;  // shape generation of the cached graphs
constexpr size_t kMaxGraphs = 64;

// Destroy every cached graph once its latest replay has finished (waits on
// that replay's event only, not on the device).  Caller holds g_graph_mu.
static void purge_graphs() {
  for (auto &e : g_graph_cache) {
    hipEventSynchronize(e.second.done);
    hipGraphExecDestroy(e.second.exec);
    hipEventDestroy(e.second.done);
  }
  g_graph_cache.clear();
}
static std::map<int, hipStream_t> g_capture_streams;

static hipStream_t capture_stream() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  auto it = g_capture_streams.find(dev);
  if (it != g_capture_streams.end()) return it->second;
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return nullptr;
  g_capture_streams[dev] = cs;
  return cs;
}

template <typename T>
static void key_add(std::string &k, const T &v) {
  k.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

template <typename TB>
static std::string cycle_key(const Krylov<float, TB> &K, int R) {
  std::string k;
  int dev = 0;
  hipGetDevice(&dev);
  key_add(k, dev);
  key_add(k, R);
  key_add(k, (int)sizeof(TB));
  key_add(k, K.transpose);
  key_add(k, K.cb);
  key_add(k, K.single_cgs);
  key_add(k, K.in16);
  key_add(k, K.n);
  const OpT<float> &o = K.op;
  const int dims[5] = {o.s.nx, o.s.ny, o.s.nz, o.s.px, o.s.py};
  key_add(k, dims);
  key_add(k, o.alpha);
  key_add(k, shape_generation());  // the captured operator launches embed the current shapes
  const void *ptrs[] = {K.V, K.w, K.t, K.scr, K.partials, K.sm, K.flags, K.cur,
                        o.cua, o.zxu, o.cx, o.cva, o.zev, o.cy, o.dw, o.cw, o.cz, o.kx, o.ky, o.rdzw,
                        o.cpm, o.mudx, o.mvdy, o.afu, o.afv, o.mmx, o.mmy, o.idw, o.gw, o.lpi, o.lower,
                        o.invden, o.cprime, o.dzw0, o.dzwn};
  key_add(k, ptrs);
  return k;
}

// R Arnoldi steps of one inner cycle: replay of the captured graph, or the
// eager launches.
template <typename TB>
static int run_inner_cycle(Krylov<float, TB> &K, int R) {
  bool use_graph;
  {
    std::lock_guard<std::mutex> g(g_graph_mu);
    use_graph = g_graphs_on;
  }
  if (!use_graph || timer_on()) {
    for (int j = 0; j < R; ++j) {
      const int rc = K.arnoldi(j, R, 1);
      if (rc) return rc;
    }
    return SISL_OK;
  }
  const std::string key = cycle_key(K, R);
  bool cached;
  {
    std::lock_guard<std::mutex> g(g_graph_mu);
    const uint64_t gen = shape_generation();
    if (gen != g_graph_generation || g_graph_cache.size() >= kMaxGraphs) {
      purge_graphs();  // graphs of older shapes never match again; bound the cache
      g_graph_generation = gen;
    }
    cached = g_graph_cache.count(key) != 0;
  }
  if (!cached) {
    hipGraphExec_t exec = nullptr;
    hipStream_t cs;
    {
      std::lock_guard<std::mutex> g(g_graph_mu);
      cs = capture_stream();
    }
    if (!cs) return SISL_ERR_LAUNCH;
    const hipStream_t orig = K.s;
    K.s = cs;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      K.s = orig;
      return SISL_ERR_LAUNCH;
    }
    int rc = SISL_OK;
    for (int j = 0; j < R && rc == SISL_OK; ++j) rc = K.arnoldi(j, R, 1);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &graph);
    K.s = orig;
    if (rc || ec != hipSuccess || !graph) {
      if (graph) hipGraphDestroy(graph);
      return rc ? rc : SISL_ERR_LAUNCH;
    }
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) return SISL_ERR_LAUNCH;
    hipEvent_t done = nullptr;
    if (hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess) {
      hipGraphExecDestroy(exec);
      return SISL_ERR_LAUNCH;
    }
    std::lock_guard<std::mutex> g(g_graph_mu);
    auto it = g_graph_cache.find(key);
    if (it != g_graph_cache.end()) {  // another thread cached the same cycle meanwhile
      hipGraphExecDestroy(exec);
      hipEventDestroy(done);
    } else {
      g_graph_cache[key] = CachedGraph{exec, done};
    }
  }
  // launch under the lock, so that no purge destroys the graph in between;
  // the event marks this replay for a later purge
  std::lock_guard<std::mutex> g(g_graph_mu);
  auto it = g_graph_cache.find(key);
  if (it == g_graph_cache.end()) {  // purged meanwhile: run the cycle eagerly
    for (int j = 0; j < R; ++j) {
      const int rc = K.arnoldi(j, R, 1);
      if (rc) return rc;
    }
    return SISL_OK;
  }
  if (hipGraphLaunch(it->second.exec, K.s) != hipSuccess) return SISL_ERR_LAUNCH;
  if (hipEventRecord(it->second.done, K.s) != hipSuccess) return SISL_ERR_LAUNCH;
  return SISL_OK;
}

// Mixed: FP64 residual refinement around FP32 GMRES cycles with early exit;
// TB is the inner basis storage type (float, or FP16 scaled by sqrt(n)).
//
// opr (optional, FP32 coefficients): the refinement stores its FP64 residual
// b - A x rounded to FP32 (the operator launch's FP32 residual epilogue) and
// applies the preconditioner in FP32 to it; z = M r then serves the stopping
// test (its norm, relative error ~1e-7) and is the next cycle's start vector
// as before.  Without it the refinement applies M in FP64.
template <typename TB>
static int gmres_mixed(const OpT<double> &op64, const OpT<float> &op32, const SislGmresParams &p,
                       const double *b, double *x, char *ws, SislGmresStats *st, hipStream_t s,
                       const OpT<float> *opr = nullptr) {
  const int R = p.restart;
  const Layout L = layout(op64.s, R, true);
  const int64_t n = op64.s.n;
  Small *sm = (Small *)(ws + L.small);
  double *partials = (double *)(ws + L.partials);
  double *w64 = (double *)(ws + L.w), *t64 = (double *)(ws + L.t), *scr64 = (double *)(ws + L.scr);
  Krylov<float, TB> K{op32, p.transpose, (TB *)(ws + L.V32), (float *)(ws + L.w32),
                      (float *)(ws + L.t32), (float *)(ws + L.scr32), partials, sm,
                      (int *)(ws + L.flags), n, s};
  K.cur = (float *)(ws + L.cur32);  // second Arnoldi buffer (and the FP16 path's unit residual)
  K.cb = p.coef_bf16 ? 1 : 0;
  K.single_cgs = p.inner_single_cgs ? 1 : 0;
  K.in16 = (sizeof(TB) == 2 && K.single_cgs && p.inner_basis_operand) ? 1 : 0;
  int rc;
  double hs[4];
  auto M64 = [&](double *in, double *out) -> int {
    if (!p.transpose) return launch_M<double>(op64, in, out, s);
    return launch_MT<double>(op64, in, out, scr64, s);
  };
  norm2<double>(b, n, partials, &sm->scal[0], s);
  hipMemcpyAsync(t64, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
  if ((rc = M64(t64, w64))) return rc;
  norm2<double>(w64, n, partials, &sm->scal[1], s);
  if ((rc = sync_copy(hs, sm->scal, sizeof(double) * 2, s))) return rc;
  const double b_norm = std::sqrt(hs[0]);
  const double Mb_norm = std::sqrt(hs[1]);
  const double atol = std::fmax(p.tol * b_norm, p.atol);
  const double ptol = Mb_norm * nan_min1(atol / b_norm);
  const int max_outer = p.max_outer > 0 ? p.max_outer : p.maxiter;
  Fuse rb;  // the FP64 operator launch stores b - A x
  rb.res = b;
  Fuse rb32 = rb;  // ... rounded to FP32 into K.t (free outside the inner cycles)
  rb32.nb = 5;
  rb32.vout = K.t;
  hipLaunchKernelGGL(zero_steps_total_kernel, dim3(1), dim3(1), 0, s, sm);
  int outer = 0;
  double rnorm = 0;
  for (;;) {
    if (opr) {
      // r = b - A x (formed in FP64, stored in FP32), z = M r (FP32) into the
      // next cycle's start vector, ||z||^2
      if (!p.transpose) {
        if ((rc = launch_A<double>(op64, x, t64, s, nullptr, &rb32))) return rc;
        if ((rc = launch_M<float>(*opr, K.t, K.cur, s))) return rc;
      } else {
        if ((rc = launch_AT<double>(op64, x, t64, scr64, s, nullptr, &rb32))) return rc;
        if ((rc = launch_MT<float>(*opr, K.t, K.cur, K.scr, s))) return rc;
      }
      norm2<float>(K.cur, n, partials, &sm->scal[2], s);
    } else {
      // r = b - A x (FP64, stored by the operator launch), z = M r (FP64)
      if (!p.transpose) {
        if ((rc = launch_A<double>(op64, x, t64, s, nullptr, &rb))) return rc;
      } else {
        if ((rc = launch_AT<double>(op64, x, t64, scr64, s, nullptr, &rb))) return rc;
      }
      if ((rc = M64(t64, w64))) return rc;
      // ||z||^2 and the FP32 start vector of the next cycle in one pass over z
      hipLaunchKernelGGL(norm2_to_f32_kernel, dim3(kRedGrid), dim3(kBlock), 0, s, (const double *)w64, K.cur, n,
                         partials);
      reduce(partials, 1, &sm->scal[2], s);
    }
    if ((rc = sync_copy(hs, sm->scal, sizeof(double) * 3, s))) return rc;
    rnorm = std::sqrt(hs[2]);
    if (!(rnorm > 2.220446049250313e-16) && !std::isnan(rnorm)) rnorm = 0.0;
    if (!(rnorm > ptol) || outer >= max_outer) break;
    const double target = std::fmax(p.inner_rtol * rnorm, 0.5 * ptol);
    hipLaunchKernelGGL(reset_cycle_kernel, dim3(1), dim3(kBlock), 0, s, sm, R, rnorm, target, K.flags,
                       1.0 / K.vscale());
    if ((rc = run_inner_cycle(K, R))) return rc;
    hipLaunchKernelGGL(lstsq_kernel, dim3(1), dim3(kLsqThreads), 0, s, sm, R, rnorm);
    combine<TB>(K.V, K.ldv(), R, sm->y, x, n, 1.0 / K.vscale(), s, &sm->steps, K.in16 ? sm->vsc : nullptr);
    ++outer;
  }
  if (st) {
    int steps_total = 0;
    if ((rc = sync_copy(&steps_total, &sm->steps_total, sizeof(int), s))) return rc;
    st->cycles = outer;
    st->arnoldi_steps = steps_total;
    st->residual = rnorm;
    st->ptol = ptol;
    st->b_norm = b_norm;
  }
  return std::isfinite(rnorm) ? SISL_OK : SISL_ERR_NOT_FINITE;
}

// ---- batched (ensemble) mixed GMRES -------------------------------------------
//
// M members on their own grids advance in lockstep: every pass of the solver
// (operator, preconditioner, projections, update, normalisation, reductions,
// small per-member bookkeeping) is one launch over all members (member =
// blockIdx.y, or blockIdx.x for one-block kernels).  The outer refinement is
// controlled on the device (convergence test, the Hessenberg least squares by
// one thread per member), so the host only reads the members' done flags once
// per outer iteration.  Precision as gmres_mixed with the FP16 basis.

using BT = _Float16;

struct KMem {  // one member's solver state (device table entry)
  BT *V;
  float *w, *cur, *t, *scr;
  double *w64, *t64, *scr64, *x64, *b64;
  double *partials;
  Small *sm;
  int *flags;
  int64_t n, ldv;
  double vscale, vinv;
};

enum { kSelH1 = 0, kSelNn1a = 1, kSelH2 = 2, kSelNn1b = 3, kSelScal0 = 4 };

__device__ __forceinline__ double *small_slot(Small *sm, int sel) {
  switch (sel) {
    case kSelH1: return sm->h1;
    case kSelNn1a: return sm->nn1a;
    case kSelH2: return sm->h2;
    case kSelNn1b: return sm->nn1b;
    default: return sm->scal + (sel - kSelScal0);
  }
}

// Arnoldi-step gating of member m at step j: stop = flags[j]; `gated`
// additionally requires the second Gram-Schmidt pass.  The Arnoldi vector of
// step j lives in w (even j) or cur (odd j): the operator launch of step j
// reads the previous, unnormalised one and normalises it on the fly.
__device__ __forceinline__ float *kb_arnoldi_vec(const KMem &k, int j) { return (j & 1) ? k.cur : k.w; }
template <int G>
__global__ void __launch_bounds__(256) kb_dots1_kernel(const KMem *__restrict__ km, int nv, int j,
                                                      int gated) {
  const KMem &k = km[blockIdx.y];
  dots1_body<BT, float, G>(k.V, k.ldv, nv, kb_arnoldi_vec(k, j), k.n, k.partials, k.vinv, k.flags + j,
                           gated ? &k.sm->gate2 : nullptr, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) kb_dots_kernel(const KMem *__restrict__ km, int nv, int j,
                                                     int gated) {
  const KMem &k = km[blockIdx.y];
  dots_body<BT, float>(k.V, k.ldv, nv, kb_arnoldi_vec(k, j), k.n, k.partials, k.vinv, k.flags + j,
                       gated ? &k.sm->gate2 : nullptr, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) kb_update_kernel(const KMem *__restrict__ km, int nv, int j,
                                                       int sel_h, int gated) {
  const KMem &k = km[blockIdx.y];
  update_body<BT, float>(k.V, k.ldv, nv, small_slot(k.sm, sel_h), kb_arnoldi_vec(k, j), k.n, k.partials, k.vinv,
                         k.flags + j, gated ? &k.sm->gate2 : nullptr, blockIdx.x, gridDim.x);
}
// out[v] = sum over blocks of partials[v][b]; grid (nslots, M).  mode 0:
// Arnoldi step j (stop = flags[j]); mode 1: outer loop (skip when done).
__global__ void __launch_bounds__(256) kb_reduce_kernel(const KMem *__restrict__ km, int j, int sel,
                                                       int gated, int mode) {
  const KMem &k = km[blockIdx.y];
  if (mode == 0 && k.flags[j]) return;
  if (mode == 1 && k.sm->done) return;
  if (gated && !k.sm->gate2) return;
  __shared__ double red[kBlock / 64];
  const int v = blockIdx.x;
  double sacc = 0;
  for (int b = threadIdx.x; b < kRedGrid; b += blockDim.x) sacc += k.partials[(int64_t)v * kRedGrid + b];
  sacc = wave_sum(sacc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sacc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int q = 0; q < kBlock / 64; ++q) t += red[q];
    small_slot(k.sm, sel)[v] = t;
  }
}
__global__ void kb_snapshot_kernel(const KMem *__restrict__ km, int j) {
  const KMem &k = km[blockIdx.x];
  if (threadIdx.x == 0) k.flags[j] = k.sm->broken;
}
__global__ void kb_decide_kernel(const KMem *__restrict__ km, int nv) {
  if (threadIdx.x == 0) decide_body(km[blockIdx.x].sm, nv);
}
__global__ void kb_finalize_kernel(const KMem *__restrict__ km, int j, int R) {
  if (threadIdx.x == 0) finalize_body(km[blockIdx.x].sm, j, R, 1);
}
// ||vec||^2 into scal[slot]: vec 0 = b64, 1 = w64 (outer loop, skipped when done)
__global__ void __launch_bounds__(256) kb_norm2_kernel(const KMem *__restrict__ km, int which) {
  const KMem &k = km[blockIdx.y];
  if (k.sm->done) return;
  const double *v = which == 0 ? k.b64 : k.w64;
  dots_body<double, double>(v, k.n, 0, v, k.n, k.partials, 1.0, nullptr, nullptr, blockIdx.x, gridDim.x);
}
__global__ void kb_residual_kernel(const KMem *__restrict__ km) {  // t64 = b64 - t64
  const KMem &k = km[blockIdx.y];
  if (k.sm->done) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < k.n; e += stride)
    k.t64[e] = k.b64[e] - k.t64[e];
}
// t = b64 - t64 rounded to FP32 (the FP32 refinement's residual)
__global__ void kb_residual32_kernel(const KMem *__restrict__ km) {
  const KMem &k = km[blockIdx.y];
  if (k.sm->done) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < k.n; e += stride)
    k.t[e] = (float)(k.b64[e] - k.t64[e]);
}
// ||cur||^2 partials (the FP32 refinement's preconditioned residual)
__global__ void __launch_bounds__(256) kb_norm2_cur_kernel(const KMem *__restrict__ km) {
  const KMem &k = km[blockIdx.y];
  if (k.sm->done) return;
  dots_body<float, float>(k.cur, k.n, 0, k.cur, k.n, k.partials, 1.0, nullptr, nullptr, blockIdx.x, gridDim.x);
}
// ||w64||^2 partials and the FP32 copy cur = w64, the next cycle's
// unnormalised start vector (its first operator launch normalises it).
__global__ void __launch_bounds__(256) kb_norm2_f32_kernel(const KMem *__restrict__ km) {
  const KMem &k = km[blockIdx.y];
  if (k.sm->done) return;
  norm2_to_f32_body(k.w64, k.cur, k.n, k.partials, blockIdx.x, gridDim.x);
}
// First kernel of a solve: clear the previous solve's done flag.
__global__ void kb_begin_kernel(const KMem *__restrict__ km) {
  if (threadIdx.x == 0) km[blockIdx.x].sm->done = 0;
}
// Start of the solve: norms in scal[0], scal[1] -> ptol (jax gmres semantics).
__global__ void kb_init_kernel(const KMem *__restrict__ km, double tol, double atol) {
  if (threadIdx.x != 0) return;
  Small *sm = km[blockIdx.x].sm;
  const double b_norm = sqrt(sm->scal[0]), Mb_norm = sqrt(sm->scal[1]);
  const double at = fmax(tol * b_norm, atol);
  sm->b_norm = b_norm;
  sm->Mb_norm = Mb_norm;
  sm->ptol = Mb_norm * nan_min1(at / b_norm);
  sm->outer = 0;
  sm->steps_total = 0;
  sm->done = 0;
}
// Outer iteration: residual norm in scal[2]; convergence test; new cycle.
__global__ void kb_reset_kernel(const KMem *__restrict__ km, int R, double inner_rtol, int max_outer) {
  Small *sm = km[blockIdx.x].sm;
  int *flags = km[blockIdx.x].flags;
  const int t = threadIdx.x;
  __shared__ int fin;
  if (t == 0) {
    fin = sm->done;
    if (!fin) {
      double rn = sqrt(sm->scal[2]);
      if (!(rn > 2.220446049250313e-16) && !isnan(rn)) rn = 0.0;
      sm->rnorm = rn;
      if (!(rn > sm->ptol) || sm->outer >= max_outer) fin = 1;
      sm->done = fin;
    }
  }
  __syncthreads();
  for (int q = t; q < R * (R + 1); q += blockDim.x) {
    const int r = q / (R + 1), c = q % (R + 1);
    sm->H[q] = r == c ? 1.0 : 0.0;
  }
  for (int q = t; q <= R; q += blockDim.x) flags[q] = fin;
  if (t == 0) {
    sm->broken = fin;
    sm->steps = 0;
    sm->gate2 = 0;
    sm->inv = (!fin && sm->rnorm > 0.0) ? 1.0 / sm->rnorm : 0.0;  // normalises the start vector
    for (int v = 0; v <= R; ++v) sm->g[v] = 0.0;
    sm->g[0] = sm->rnorm;
    sm->target = fmax(inner_rtol * sm->rnorm, 0.5 * sm->ptol);
  }
}
// Least squares of the cycle (lstsq_block) on the device, one workgroup per
// member; y into sm->y.
// One member's solve outcome (a copy of its Small record's outer-loop fields).
struct EnsOutcome {
  int done, outer, steps_total, pad;
  double b_norm, Mb_norm, ptol, rnorm;
};
// *all_done = 1 when every member's refinement has finished (one read per
// outer iteration instead of one per member).
__global__ void kb_all_done_kernel(const KMem *__restrict__ km, int M, int *__restrict__ all_done) {
  if (threadIdx.x != 0) return;
  int all = 1;
  for (int m = 0; m < M; ++m) all = all && km[m].sm->done;
  *all_done = all;
}
__global__ void kb_outcome_kernel(const KMem *__restrict__ km, int M, EnsOutcome *__restrict__ out) {
  for (int m = threadIdx.x; m < M; m += blockDim.x) {
    const Small *sm = km[m].sm;
    out[m] = EnsOutcome{sm->done, sm->outer, sm->steps_total, 0, sm->b_norm, sm->Mb_norm, sm->ptol, sm->rnorm};
  }
}
__global__ void __launch_bounds__(kLsqThreads) kb_lstsq_kernel(const KMem *__restrict__ km, int R) {
  Small *sm = km[blockIdx.x].sm;
  if (sm->done) return;  // uniform over the block
  lstsq_block(sm, R, sm->rnorm);
  if (threadIdx.x == 0) sm->outer += 1;
}
__global__ void __launch_bounds__(256) kb_combine_kernel(const KMem *__restrict__ km) {
  const KMem &k = km[blockIdx.y];
  if (k.sm->done) return;
  combine_body<BT>(k.V, k.ldv, k.sm->steps, k.sm->y, k.x64, k.n, k.vinv, blockIdx.x, gridDim.x);
}
// Copy between the caller's vectors and the members' workspace copies.
struct PtrTable {
  const double *src[kMaxMembers];
  double *dst[kMaxMembers];
  int64_t n[kMaxMembers];
};
__global__ void kb_copy_kernel(PtrTable t) {
  const int m = blockIdx.y;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < t.n[m]; e += stride)
    t.dst[m][e] = t.src[m][e];
}

// Host-side plan of one batched workspace (set up once per operator set).
struct EnsPlan {
  int M, R, L, S;
  int maxA, maxM, maxCols, maxFaces;
  int64_t maxN;
  std::vector<int64_t> n;
  KMem *km;                  // device table
  // inner cycle, step parity q = j % 2 (stop flags[j]): the operator launch
  // reads the previous unnormalised Arnoldi vector src_q (cur for even
  // steps, w for odd ones), normalises it into basis column j and writes
  // dst_q (the other buffer); the preconditioner runs in place on dst_q
  BOp<float> *A32[2], *M32[2];    // src_q -> dst_q; in place dst_q
  BOp<float> *AT32[2], *MT32[2];  // transpose: src_q -> t; t -> dst_q
  BOp<double> *A64, *M64;    // outer: x64 -> t64 (skip when done); t64 -> w64
  BOp<double> *AT64, *MT64;
  BOp<double> *Mb64, *MTb64; // start: b64 -> w64 (t64 for the transpose)
  BOp<float> *Mr32, *MTr32;  // refinement in FP32: t (FP32 residual) -> cur (skip when done)
  int *all_done;             // 1 when every member's refinement has finished
  EnsOutcome *outcome;       // the members' outcomes, gathered after a solve
  std::vector<Small *> sm;
  std::vector<KMem> kmh;     // host copy of the device table
};
static std::mutex g_ens_mu;
static std::map<const void *, EnsPlan> g_ens;

constexpr int kEnsTables = 16;  // 8 FP32 inner tables (2 parities x A, M, AT, MT), 6 FP64 outer ones, 2 FP32 refinement ones
static int64_t ens_head_bytes(int M) {
  return align256(sizeof(KMem) * M) + kEnsTables * align256(sizeof(BOp<double>) * M) + align256(sizeof(int)) +
         align256(sizeof(EnsOutcome) * M);
}
static int64_t ens_member_bytes(const Stag &s, int R) {
  const Layout L = layout(s, R, true);
  return align256(L.total) + 2 * align256(sizeof(double) * s.n);
}

static int ens_setup(const SislOperator *ops64, const SislOperator *ops32, int M, int R, char *ws,
                     hipStream_t st) {
  EnsPlan P;
  P.M = M;
  P.R = R;
  P.L = 0;
  P.maxA = P.maxM = P.maxCols = P.maxFaces = 1;
  P.maxN = 1;
  std::vector<KMem> km(M);
  std::vector<BOp<float>> a32[2], m32[2], at32[2], mt32[2];
  for (int q = 0; q < 2; ++q) {
    a32[q].resize(M);
    m32[q].resize(M);
    at32[q].resize(M);
    mt32[q].resize(M);
  }
  std::vector<BOp<double>> a64(M), m64(M), at64(M), mt64(M), mb64(M), mtb64(M);
  std::vector<BOp<float>> mr32(M), mtr32(M);
  int64_t off = ens_head_bytes(M);
  int S = 1;
  for (int m = 0; m < M; ++m) {
    const OpT<double> o64 = OpT<double>::make(ops64[m]);
    const OpT<float> o32 = OpT<float>::make(ops32[m]);
    if (!o64.complete() || !o32.complete()) return SISL_ERR_ARG;
    const Stag &s = o64.s;
    if ((uint64_t)s.n * sizeof(double) >= (1ull << 32)) return SISL_ERR_ARG;
    const int Lm = seg_levels(s.nz);
    if (!Lm) return SISL_ERR_ARG;
    P.L = std::max(P.L, Lm);
    const Layout Lay = layout(s, R, true);
    char *base = ws + off;
    KMem &k = km[m];
    k.V = (BT *)(base + Lay.V32);
    k.w = (float *)(base + Lay.w32);
    k.cur = (float *)(base + Lay.cur32);
    k.t = (float *)(base + Lay.t32);
    k.scr = (float *)(base + Lay.scr32);
    k.w64 = (double *)(base + Lay.w);
    k.t64 = (double *)(base + Lay.t);
    k.scr64 = (double *)(base + Lay.scr);
    k.partials = (double *)(base + Lay.partials);
    k.sm = (Small *)(base + Lay.small);
    k.flags = (int *)(base + Lay.flags);
    k.x64 = (double *)(base + align256(Lay.total));
    k.b64 = (double *)(base + align256(Lay.total) + align256(sizeof(double) * s.n));
    k.n = s.n;
    k.ldv = basis_ld(s.n);
    k.vscale = basis_scale<BT>(s.n);
    k.vinv = 1.0 / k.vscale;
    P.sm.push_back(k.sm);
    P.n.push_back(s.n);
    for (int q = 0; q < 2; ++q) {
      float *src = q == 0 ? k.cur : k.w, *dst = q == 0 ? k.w : k.cur;
      a32[q][m] = BOp<float>{o32, src, 0, dst, k.scr, k.flags, 1, &k.sm->inv, k.V, k.ldv, k.vscale};
      m32[q][m] = BOp<float>{o32, dst, 0, dst, k.scr, k.flags, 1, nullptr, nullptr, 0, 1.0};
      at32[q][m] = BOp<float>{o32, src, 0, k.t, k.scr, k.flags, 1, &k.sm->inv, k.V, k.ldv, k.vscale};
      mt32[q][m] = BOp<float>{o32, k.t, 0, dst, k.scr, k.flags, 1, nullptr, nullptr, 0, 1.0};
    }
    const int *dn = &k.sm->done;
    a64[m] = BOp<double>{o64, k.x64, 0, k.t64, k.scr64, dn, 0, nullptr, nullptr, 0, 1.0};
    m64[m] = BOp<double>{o64, k.t64, 0, k.w64, k.scr64, dn, 0, nullptr, nullptr, 0, 1.0};
    at64[m] = BOp<double>{o64, k.x64, 0, k.t64, k.scr64, dn, 0, nullptr, nullptr, 0, 1.0};
    mt64[m] = BOp<double>{o64, k.t64, 0, k.w64, k.scr64, dn, 0, nullptr, nullptr, 0, 1.0};
    mb64[m] = BOp<double>{o64, k.b64, 0, k.w64, k.scr64, nullptr, 0, nullptr, nullptr, 0, 1.0};
    mr32[m] = BOp<float>{o32, k.t, 0, k.cur, k.scr, dn, 0, nullptr, nullptr, 0, 1.0};
    mtr32[m] = BOp<float>{o32, k.t, 0, k.cur, k.scr, dn, 0, nullptr, nullptr, 0, 1.0};
    mtb64[m] = BOp<double>{o64, k.b64, 0, k.w64, k.scr64, nullptr, 0, nullptr, nullptr, 0, 1.0};
    P.maxA = std::max(P.maxA, blocks_A(s));
    P.maxM = std::max(P.maxM, blocks_M(s));
    P.maxCols = std::max(P.maxCols, s.nx * s.ny);
    P.maxFaces = std::max(P.maxFaces, blocks_MT_faces(s));
    P.maxN = std::max(P.maxN, s.n);
    off += ens_member_bytes(s, R);
  }
  for (int m = 0; m < M; ++m) S = std::max(S, (OpT<double>::make(ops64[m]).s.nz + P.L - 1) / P.L);
  P.S = S;
  // tables at the workspace head
  char *h = ws;
  P.km = (KMem *)h;
  h += align256(sizeof(KMem) * M);
  BOp<float> **f32[] = {&P.A32[0], &P.A32[1], &P.M32[0], &P.M32[1], &P.AT32[0], &P.AT32[1], &P.MT32[0], &P.MT32[1],
                        &P.Mr32, &P.MTr32};
  std::vector<BOp<float>> *v32[] = {&a32[0], &a32[1], &m32[0], &m32[1], &at32[0], &at32[1], &mt32[0], &mt32[1],
                                    &mr32, &mtr32};
  for (int q = 0; q < 10; ++q) {
    *f32[q] = (BOp<float> *)h;
    h += align256(sizeof(BOp<double>) * M);
  }
  BOp<double> **f64[] = {&P.A64, &P.M64, &P.AT64, &P.MT64, &P.Mb64, &P.MTb64};
  std::vector<BOp<double>> *v64[] = {&a64, &m64, &at64, &mt64, &mb64, &mtb64};
  for (int q = 0; q < 6; ++q) {
    *f64[q] = (BOp<double> *)h;
    h += align256(sizeof(BOp<double>) * M);
  }
  P.kmh = km;
  P.all_done = (int *)h;
  h += align256(sizeof(int));
  P.outcome = (EnsOutcome *)h;
  h += align256(sizeof(EnsOutcome) * M);
  if (hipMemcpyAsync(P.km, km.data(), sizeof(KMem) * M, hipMemcpyHostToDevice, st) != hipSuccess)
    return SISL_ERR_LAUNCH;
  for (int q = 0; q < 10; ++q)
    if (hipMemcpyAsync(*f32[q], v32[q]->data(), sizeof(BOp<float>) * M, hipMemcpyHostToDevice, st) != hipSuccess)
      return SISL_ERR_LAUNCH;
  for (int q = 0; q < 6; ++q)
    if (hipMemcpyAsync(*f64[q], v64[q]->data(), sizeof(BOp<double>) * M, hipMemcpyHostToDevice, st) != hipSuccess)
      return SISL_ERR_LAUNCH;
  if (hipStreamSynchronize(st) != hipSuccess) return SISL_ERR_LAUNCH;
  std::lock_guard<std::mutex> g(g_ens_mu);
  g_ens[ws] = std::move(P);
  return SISL_OK;
}

static void kb_dots_dispatch(const EnsPlan &P, int nv, int j, int gated, hipStream_t st) {
  const dim3 gr(kRedGrid, P.M), bl(kBlock);
  if (nv <= kGrp) hipLaunchKernelGGL(kb_dots1_kernel<1>, gr, bl, 0, st, P.km, nv, j, gated);
  else if (nv <= 2 * kGrp) hipLaunchKernelGGL(kb_dots1_kernel<2>, gr, bl, 0, st, P.km, nv, j, gated);
  else if (nv <= 3 * kGrp) hipLaunchKernelGGL(kb_dots1_kernel<3>, gr, bl, 0, st, P.km, nv, j, gated);
  else if (nv <= 4 * kGrp) hipLaunchKernelGGL(kb_dots1_kernel<4>, gr, bl, 0, st, P.km, nv, j, gated);
  else hipLaunchKernelGGL(kb_dots_kernel, gr, bl, 0, st, P.km, nv, j, gated);
}

static int ens_arnoldi(const EnsPlan &P, int j, int transpose, int single_cgs, hipStream_t st) {
  int rc;
  const int M = P.M, R = P.R, nv = j + 1;
  const int q = j & 1;
  hipLaunchKernelGGL(kb_snapshot_kernel, dim3(M), dim3(64), 0, st, (const KMem *)P.km, j);
  if (!transpose) {
    if ((rc = launch_A_batched<float>(P.A32[q], M, P.maxA, j, st, 1))) return rc;
    if ((rc = launch_M_batched<float>(P.M32[q], M, P.maxM, P.L, P.S, j, st))) return rc;
  } else {
    if ((rc = launch_AT_batched<float>(P.AT32[q], M, P.maxA, j, st, 1))) return rc;
    if ((rc = launch_MT_batched<float>(P.MT32[q], M, P.maxM, P.L, P.S, P.maxFaces, j, st))) return rc;
  }
  kb_dots_dispatch(P, nv, j, 0, st);
  hipLaunchKernelGGL(kb_reduce_kernel, dim3(nv + 1, M), dim3(kBlock), 0, st, (const KMem *)P.km, j, kSelH1, 0, 0);
  hipLaunchKernelGGL(kb_update_kernel, dim3(kRedGrid, M), dim3(kBlock), 0, st, (const KMem *)P.km, nv, j, kSelH1, 0);
  hipLaunchKernelGGL(kb_reduce_kernel, dim3(1, M), dim3(kBlock), 0, st, (const KMem *)P.km, j, kSelNn1a, 0, 0);
  if (!single_cgs) {  // one pass: gate2 stays 0 from the cycle reset, the bookkeeping reads h1, nn1a
    hipLaunchKernelGGL(kb_decide_kernel, dim3(M), dim3(64), 0, st, (const KMem *)P.km, nv);
    kb_dots_dispatch(P, nv, j, 1, st);
    hipLaunchKernelGGL(kb_reduce_kernel, dim3(nv, M), dim3(kBlock), 0, st, (const KMem *)P.km, j, kSelH2, 1, 0);
    hipLaunchKernelGGL(kb_update_kernel, dim3(kRedGrid, M), dim3(kBlock), 0, st, (const KMem *)P.km, nv, j, kSelH2, 1);
    hipLaunchKernelGGL(kb_reduce_kernel, dim3(1, M), dim3(kBlock), 0, st, (const KMem *)P.km, j, kSelNn1b, 1, 0);
  }
  hipLaunchKernelGGL(kb_finalize_kernel, dim3(M), dim3(64), 0, st, (const KMem *)P.km, j, R);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

// One inner cycle of the batched solver: R Arnoldi steps, the members' least
// squares and the update of x.  Its launches have fixed arguments (the plan's
// tables in the workspace, the step indices, the launch geometry of the
// member grids) and are gated on the device by each member's stop and done
// flags, so the cycle is captured once per workspace layout, direction and
// Gram-Schmidt form (on the private capture stream, as the single-member
// cycle) and replayed on the caller's stream for every later cycle.
static std::map<std::string, CachedGraph> g_ens_graphs;  // guarded by g_graph_mu
constexpr size_t kMaxEnsGraphs = 16;

static int ens_cycle_eager(const EnsPlan &P, int transpose, int single_cgs, hipStream_t st) {
  int rc;
  for (int j = 0; j < P.R; ++j)
    if ((rc = ens_arnoldi(P, j, transpose, single_cgs, st))) return rc;
  hipLaunchKernelGGL(kb_lstsq_kernel, dim3(P.M), dim3(kLsqThreads), 0, st, (const KMem *)P.km, P.R);
  hipLaunchKernelGGL(kb_combine_kernel, dim3(kRedGrid, P.M), dim3(kBlock), 0, st, (const KMem *)P.km);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

// Whether work issued to stream st is being captured into a graph.
static bool stream_capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive;
}

static int ens_cycle(const EnsPlan &P, int transpose, int single_cgs, hipStream_t st) {
  bool use_graph;
  {
    std::lock_guard<std::mutex> g(g_graph_mu);
    use_graph = g_graphs_on;
  }
  // inside a caller's capture the cycle's launches become part of its graph
  if (!use_graph || timer_on() || stream_capturing(st)) return ens_cycle_eager(P, transpose, single_cgs, st);
  std::string key;
  int dev = 0;
  hipGetDevice(&dev);
  key_add(key, dev);
  key_add(key, (const void *)P.km);
  const int geo[11] = {P.M, P.R, P.L, P.S, P.maxA, P.maxM, P.maxCols, P.maxFaces, transpose, single_cgs, 0};
  key_add(key, geo);
  key_add(key, P.maxN);
  bool cached;
  {
    std::lock_guard<std::mutex> g(g_graph_mu);
    if (g_ens_graphs.size() >= kMaxEnsGraphs && !g_ens_graphs.count(key)) {
      for (auto &e : g_ens_graphs) {  // bound the cache: retire all once their replays finished
        hipEventSynchronize(e.second.done);
        hipGraphExecDestroy(e.second.exec);
        hipEventDestroy(e.second.done);
      }
      g_ens_graphs.clear();
    }
    cached = g_ens_graphs.count(key) != 0;
  }
  if (!cached) {
    hipStream_t cs;
    {
      std::lock_guard<std::mutex> g(g_graph_mu);
      cs = capture_stream();
    }
    if (!cs) return SISL_ERR_LAUNCH;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) return SISL_ERR_LAUNCH;
    const int rc = ens_cycle_eager(P, transpose, single_cgs, cs);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &graph);
    if (rc || ec != hipSuccess || !graph) {
      if (graph) hipGraphDestroy(graph);
      return rc ? rc : SISL_ERR_LAUNCH;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) return SISL_ERR_LAUNCH;
    hipEvent_t done = nullptr;
    if (hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess) {
      hipGraphExecDestroy(exec);
      return SISL_ERR_LAUNCH;
    }
    std::lock_guard<std::mutex> g(g_graph_mu);
    if (g_ens_graphs.count(key)) {
      hipGraphExecDestroy(exec);
      hipEventDestroy(done);
    } else {
      g_ens_graphs[key] = CachedGraph{exec, done};
    }
  }
  std::lock_guard<std::mutex> g(g_graph_mu);
  auto it = g_ens_graphs.find(key);
  if (it == g_ens_graphs.end()) return ens_cycle_eager(P, transpose, single_cgs, st);
  if (hipGraphLaunch(it->second.exec, st) != hipSuccess) return SISL_ERR_LAUNCH;
  if (hipEventRecord(it->second.done, st) != hipSuccess) return SISL_ERR_LAUNCH;
  return SISL_OK;
}

static int ens_outcome(const EnsPlan &P, SislGmresStats *stats, hipStream_t st);

static int ens_solve(const void *ws, const SislGmresParams &p, const double *const *b, double *const *x,
                     SislGmresStats *stats, hipStream_t st) {
  EnsPlan P;
  {
    std::lock_guard<std::mutex> g(g_ens_mu);
    auto it = g_ens.find(ws);
    if (it == g_ens.end()) return SISL_ERR_ARG;
    P = it->second;
  }
  if (p.restart != P.R) return SISL_ERR_ARG;
  const int M = P.M, R = P.R;
  int rc;
  PtrTable tin{}, tout{};
  const std::vector<KMem> &km = P.kmh;
  for (int m = 0; m < M; ++m) {
    if (!b[m] || !x[m]) return SISL_ERR_ARG;
    tin.src[m] = b[m];
    tin.dst[m] = km[m].b64;
    tin.n[m] = km[m].n;
    tout.src[m] = x[m];
    tout.dst[m] = km[m].x64;
    tout.n[m] = km[m].n;
  }
  const dim3 cg(grid_for(P.maxN), M);
  hipLaunchKernelGGL(kb_copy_kernel, cg, dim3(kBlock), 0, st, tin);
  hipLaunchKernelGGL(kb_copy_kernel, cg, dim3(kBlock), 0, st, tout);
  // ||b||, ||M b||, ptol
  hipLaunchKernelGGL(kb_begin_kernel, dim3(M), dim3(64), 0, st, (const KMem *)P.km);
  hipLaunchKernelGGL(kb_norm2_kernel, dim3(kRedGrid, M), dim3(kBlock), 0, st, (const KMem *)P.km, 0);
  hipLaunchKernelGGL(kb_reduce_kernel, dim3(1, M), dim3(kBlock), 0, st, (const KMem *)P.km, 0, kSelScal0 + 0, 0, 1);
  if (!p.transpose) {
    if ((rc = launch_M_batched<double>(P.Mb64, M, P.maxM, P.L, P.S, 0, st))) return rc;
  } else {
    if ((rc = launch_MT_batched<double>(P.MTb64, M, P.maxM, P.L, P.S, P.maxFaces, 0, st))) return rc;
  }
  hipLaunchKernelGGL(kb_norm2_kernel, dim3(kRedGrid, M), dim3(kBlock), 0, st, (const KMem *)P.km, 1);
  hipLaunchKernelGGL(kb_reduce_kernel, dim3(1, M), dim3(kBlock), 0, st, (const KMem *)P.km, 0, kSelScal0 + 1, 0, 1);
  hipLaunchKernelGGL(kb_init_kernel, dim3(M), dim3(64), 0, st, (const KMem *)P.km, p.tol, p.atol);
  const int max_outer = p.max_outer > 0 ? p.max_outer : p.maxiter;
  // Under a caller's stream capture (a graph-captured time step) nothing is
  // read back: every outer iteration the eager loop could run is recorded,
  // and those after a member has finished do nothing on the device (every
  // launch of the outer loop and of the cycle is gated by the member's done
  // and stop flags), so a replay gives the eager solve's bits.  The members'
  // outcomes stay on the device (sisl_gmres_batched_outcome).
  const bool capturing = stream_capturing(st);
  for (int outer = 0;; ++outer) {
    // residual M (b - A x) in FP64, its norm, the first basis vector
    if (!p.transpose) {
      if ((rc = launch_A_batched<double>(P.A64, M, P.maxA, 0, st))) return rc;
    } else {
      if ((rc = launch_AT_batched<double>(P.AT64, M, P.maxA, 0, st))) return rc;
    }
    if (p.op_refine) {
      // as the single-member solver: the residual stored in FP32, the
      // preconditioner in FP32 into the next cycle's start vector
      hipLaunchKernelGGL(kb_residual32_kernel, dim3(kRedGrid, M), dim3(kBlock), 0, st, (const KMem *)P.km);
      if (!p.transpose) {
        if ((rc = launch_M_batched<float>(P.Mr32, M, P.maxM, P.L, P.S, 0, st))) return rc;
      } else {
        if ((rc = launch_MT_batched<float>(P.MTr32, M, P.maxM, P.L, P.S, P.maxFaces, 0, st))) return rc;
      }
      hipLaunchKernelGGL(kb_norm2_cur_kernel, dim3(kRedGrid, M), dim3(kBlock), 0, st, (const KMem *)P.km);
    } else {
      hipLaunchKernelGGL(kb_residual_kernel, dim3(kRedGrid, M), dim3(kBlock), 0, st, (const KMem *)P.km);
      if (!p.transpose) {
        if ((rc = launch_M_batched<double>(P.M64, M, P.maxM, P.L, P.S, 0, st))) return rc;
      } else {
        if ((rc = launch_MT_batched<double>(P.MT64, M, P.maxM, P.L, P.S, P.maxFaces, 0, st))) return rc;
      }
      hipLaunchKernelGGL(kb_norm2_f32_kernel, dim3(kRedGrid, M), dim3(kBlock), 0, st, (const KMem *)P.km);
    }
    hipLaunchKernelGGL(kb_reduce_kernel, dim3(1, M), dim3(kBlock), 0, st, (const KMem *)P.km, 0, kSelScal0 + 2, 0, 1);
    hipLaunchKernelGGL(kb_reset_kernel, dim3(M), dim3(kBlock), 0, st, (const KMem *)P.km, R, p.inner_rtol,
                       max_outer);
    if (capturing) {
      if (outer > max_outer) break;
    } else {
      // one read per outer iteration: whether every member has finished
      hipLaunchKernelGGL(kb_all_done_kernel, dim3(1), dim3(64), 0, st, (const KMem *)P.km, M, P.all_done);
      int all = 0;
      if (hipMemcpyAsync(&all, P.all_done, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return SISL_ERR_LAUNCH;
      if (all || outer > max_outer) break;
    }
    if ((rc = ens_cycle(P, p.transpose, p.inner_single_cgs, st))) return rc;
  }
  for (int m = 0; m < M; ++m) {
    tout.src[m] = km[m].x64;
    tout.dst[m] = x[m];
  }
  hipLaunchKernelGGL(kb_copy_kernel, cg, dim3(kBlock), 0, st, tout);
  SISL_CHECK_LAUNCH();
  if (capturing) {
    if (stats)
      for (int m = 0; m < M; ++m) stats[m] = SislGmresStats{-1, -1, 0.0, 0.0, 0.0};
    return SISL_OK;
  }
  return ens_outcome(P, stats, st);
}

// The members' outcomes of the last solve on workspace P, gathered on the
// device and read in one copy.
static int ens_outcome(const EnsPlan &P, SislGmresStats *stats, hipStream_t st) {
  const int M = P.M;
  hipLaunchKernelGGL(kb_outcome_kernel, dim3(1), dim3(64), 0, st, (const KMem *)P.km, M, P.outcome);
  std::vector<EnsOutcome> oc(M);
  if (hipMemcpyAsync(oc.data(), P.outcome, sizeof(EnsOutcome) * M, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return SISL_ERR_LAUNCH;
  bool finite = true;
  for (int m = 0; m < M; ++m) {
    const EnsOutcome &hs = oc[m];
    finite = finite && std::isfinite(hs.rnorm);
    if (stats) {
      stats[m].cycles = hs.outer;
      stats[m].arnoldi_steps = hs.steps_total;
      stats[m].residual = hs.rnorm;
      stats[m].ptol = hs.ptol;
      stats[m].b_norm = hs.b_norm;
    }
  }
  return finite ? SISL_OK : SISL_ERR_NOT_FINITE;
}

}  // namespace sisl

extern "C" {

int sisl_set_graphs(int32_t on) {
  std::lock_guard<std::mutex> g(sisl::g_graph_mu);
  sisl::g_graphs_on = on != 0;
  return SISL_OK;
}

int64_t sisl_gmres_workspace_bytes_f64(const SislGrid *g, int32_t restart) {
  if (!g || restart < 1 || restart > sisl::kMaxRestart) return -1;
  return sisl::layout(sisl::Stag::make(*g), restart, false).total;
}
int64_t sisl_gmres_workspace_bytes_mixed(const SislGrid *g, int32_t restart) {
  if (!g || restart < 1 || restart > sisl::kMaxRestart) return -1;
  return sisl::layout(sisl::Stag::make(*g), restart, true).total;
}

int sisl_gmres_f64(const SislOperator *op, const SislGmresParams *p, const double *b, double *x,
                   void *ws, SislGmresStats *st, void *stream) {
  if (!op || !p || !b || !x || !ws || p->restart < 1 || p->restart > sisl::kMaxRestart ||
      p->maxiter < 0)
    return SISL_ERR_ARG;
  auto o = sisl::OpT<double>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::gmres64(o, *p, b, x, (char *)ws, st, sisl::as_stream(stream));
}

int sisl_gmres_mixed(const SislOperator *op64, const SislOperator *op32, const SislGmresParams *p,
                     const double *b, double *x, void *ws, SislGmresStats *st, void *stream) {
  if (!op64 || !op32 || !p || !b || !x || !ws || p->restart < 1 ||
      p->restart > sisl::kMaxRestart || p->maxiter < 0)
    return SISL_ERR_ARG;
  auto o64 = sisl::OpT<double>::make(*op64);
  auto o32 = sisl::OpT<float>::make(*op32);
  if (!o64.complete() || !o32.complete()) return SISL_ERR_ARG;
  sisl::OpT<float> orf{};
  if (p->op_refine) {
    orf = sisl::OpT<float>::make(*p->op_refine);
    if (!orf.complete() || orf.s.n != o64.s.n) return SISL_ERR_ARG;
  }
  const sisl::OpT<float> *opr = p->op_refine ? &orf : nullptr;
  if (p->basis_fp16)
    return sisl::gmres_mixed<_Float16>(o64, o32, *p, b, x, (char *)ws, st, sisl::as_stream(stream), opr);
  return sisl::gmres_mixed<float>(o64, o32, *p, b, x, (char *)ws, st, sisl::as_stream(stream), opr);
}

int64_t sisl_gmres_batched_workspace_bytes(const SislGrid *grids, int32_t M, int32_t restart) {
  if (!grids || M < 1 || M > sisl::kMaxMembers || restart < 1 || restart > sisl::kMaxRestart) return -1;
  int64_t t = sisl::ens_head_bytes(M);
  for (int m = 0; m < M; ++m) t += sisl::ens_member_bytes(sisl::Stag::make(grids[m]), restart);
  return t;
}

int sisl_gmres_batched_setup(const SislOperator *ops64, const SislOperator *ops32, int32_t M,
                             int32_t restart, void *ws, void *stream) {
  if (!ops64 || !ops32 || !ws || M < 1 || M > sisl::kMaxMembers || restart < 1 ||
      restart > sisl::kMaxRestart)
    return SISL_ERR_ARG;
  return sisl::ens_setup(ops64, ops32, M, restart, (char *)ws, sisl::as_stream(stream));
}

int sisl_gmres_batched(int32_t M, const SislGmresParams *p, const double *const *b, double *const *x,
                       void *ws, SislGmresStats *stats, void *stream) {
  if (!p || !b || !x || !ws || M < 1 || M > sisl::kMaxMembers) return SISL_ERR_ARG;
  return sisl::ens_solve(ws, *p, b, x, stats, sisl::as_stream(stream));
}

int sisl_gmres_batched_outcome(int32_t M, void *ws, SislGmresStats *stats, void *stream) {
  if (!ws || !stats || M < 1 || M > sisl::kMaxMembers) return SISL_ERR_ARG;
  sisl::EnsPlan P;
  {
    std::lock_guard<std::mutex> g(sisl::g_ens_mu);
    auto it = sisl::g_ens.find(ws);
    if (it == sisl::g_ens.end() || it->second.M != M) return SISL_ERR_ARG;
    P = it->second;
  }
  return sisl::ens_outcome(P, stats, sisl::as_stream(stream));
}

int sisl_dot_f64(const double *x, const double *y, int64_t n, double *out, double *partials,
                 void *stream) {
  if (!x || !y || !out || !partials || n < 0) return SISL_ERR_ARG;
  hipStream_t s = sisl::as_stream(stream);
  sisl::dots<double>(x, n, 1, y, n, partials, s);
  sisl::reduce(partials, 1, out, s);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

}  // extern "C"
