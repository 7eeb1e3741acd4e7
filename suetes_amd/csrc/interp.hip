// (a) Semi-Lagrangian interpolation: tricubic gather, its deterministic
// transpose, index velocities and departure points.
//
// Reference: suetes/regional3d/operators.py:134-209 (tricubic),
//            suetes/regional3d/steppers.py:276-367 (trajectories).
#include <algorithm>

#include "common.h"

namespace sisl {

// Catmull-Rom weights of the reference's cubic_weight (operators.py:134-144)
// written per sample: f(t) = w0 p0 + w1 p1 + w2 p2 + w3 p3.
template <typename T>
__device__ __forceinline__ void cr_weights(T t, T w[4]) {
  const T t2 = t * t, t3 = t2 * t;
  w[0] = (T)-0.5 * t3 + t2 - (T)0.5 * t;
  w[1] = (T)1.5 * t3 - (T)2.5 * t2 + (T)1.0;
  w[2] = (T)-1.5 * t3 + (T)2.0 * t2 + (T)0.5 * t;
  w[3] = (T)0.5 * t3 - (T)0.5 * t2;
}

// Departure points are stored as offsets d = x_d - x_a from the arrival
// index X (exact in FP32 too: |d| is a few cells, while absolute indices up
// to 256 would keep only ~1.5e-5 cells in FP32).  Split X + d into base
// index and fraction, clamped to [0, n-1] like the reference's coordinates
// (operators.py:169-174).
// The floor is bounded to [-(n+1), n+1] before the integer conversion, so a
// far or non-finite offset clamps to the nearer edge instead of overflowing.
template <typename T>
__device__ __forceinline__ T bounded_floor(T fl, int n) {
  return fmax(fmin(fl, (T)(n + 1)), (T)(-(n + 1)));
}

template <typename T>
__device__ __forceinline__ int offset_split(int X, T d, int n, T &t) {
  const T fl = floor(d);
  int b = X + (int)bounded_floor(fl, n);
  t = d - fl;
  if (b < 0) {
    b = 0;
    t = (T)0;
  } else if (b > n - 1 || (b == n - 1 && t > (T)0)) {
    b = n - 1;
    t = (T)0;
  }
  return b;
}

// Tricubic of one point from a tap accessor F(a, b, c), a, b, c in 0..3
// relative to the stencil origin (x0-1, y0-1, z0-1); the accessor applies the
// reference's index clamping (or reads a staged copy that holds clamped
// values).  Cascade z -> y -> x as the reference does.
template <typename T, typename Fetch>
__device__ __forceinline__ T tricubic_apply(const T wx[4], const T wy[4], const T wz[4], int lim, Fetch F) {
  T acc_x = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    T acc_y = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      T acc_z = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) acc_z += wz[c] * F(a, b, c);
      acc_y += wy[b] * acc_z;
    }
    acc_x += wx[a] * acc_y;
  }
  if (lim) {  // clip to the 8 cell corners (operators.py:188-207)
    T lo = F(1, 1, 1), hi = lo;
#pragma unroll
    for (int q = 1; q < 8; ++q) {
      const T nb = F(1 + (q & 1), 1 + ((q >> 1) & 1), 1 + ((q >> 2) & 1));
      lo = nb < lo ? nb : lo;
      hi = nb > hi ? nb : hi;
    }
    acc_x = acc_x < lo ? lo : acc_x;
    acc_x = acc_x > hi ? hi : acc_x;
  }
  return acc_x;
}

// tricubic_apply reading the taps from global memory at clamped indices
// (ix: columns, iy: row offsets y * nx, iz: levels).  The z sums are formed
// one stencil plane at a time (the loop is kept rolled, so only one plane's
// 16 taps are in flight), each in the order of tricubic_apply: the same bits.
template <typename T, typename TF>
__device__ __forceinline__ T tricubic_global(const TF *__restrict__ f, const int ix[4], const int iy[4],
                                             const int iz[4], uint32_t plane, const T wx[4], const T wy[4],
                                             const T wz[4], int lim) {
  T az[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) az[a][b] = 0;
#pragma unroll 1
  for (int c = 0; c < 4; ++c) {
    const uint32_t zo = (uint32_t)iz[c] * plane;
    const T w = wz[c];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) az[a][b] += w * (T)ldg(f, zo + (uint32_t)iy[b] + (uint32_t)ix[a]);
  }
  T acc_x = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    T acc_y = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) acc_y += wy[b] * az[a][b];
    acc_x += wx[a] * acc_y;
  }
  if (lim) {
    auto F = [&](int a, int b, int c) {
      return (T)ldg(f, (uint32_t)iz[c] * plane + (uint32_t)iy[b] + (uint32_t)ix[a]);
    };
    T lo = F(1, 1, 1), hi = lo;
#pragma unroll
    for (int q = 1; q < 8; ++q) {
      const T nb = F(1 + (q & 1), 1 + ((q >> 1) & 1), 1 + ((q >> 2) & 1));
      lo = nb < lo ? nb : lo;
      hi = nb > hi ? nb : hi;
    }
    acc_x = acc_x < lo ? lo : acc_x;
    acc_x = acc_x > hi ? hi : acc_x;
  }
  return acc_x;
}

template <typename T, typename Fetch>
__device__ __forceinline__ T tricubic_point(T tx, T ty, T tz, int lim, Fetch F) {
  T wx[4], wy[4], wz[4];
  cr_weights(tx, wx);
  cr_weights(ty, wy);
  cr_weights(tz, wz);
  return tricubic_apply(wx, wy, wz, lim, F);
}

// Staged tricubic gather.  A block owns a 64 x 4 tile of (x, y) columns and
// marches a chunk of levels.  A ring of 6 LDS planes holds levels k-2 .. k+3
// of the field over the tile plus halo ([x0-2, x0+70) x [y0-2, y0+7)), each
// staged value being the field at the clamped index, so that every stencil
// within the staged box reads LDS with the reference's clamping built in.
// One new plane is staged per level (coalesced along x).  Points whose
// stencil leaves the box (displacement beyond ~1 cell) gather from global
// memory; both paths perform the same arithmetic in the same order.
//
// NF fields that share one set of departure points are gathered together:
// the offsets, their split and the weights are formed once per point, each
// field has its own ring.  TF is the storage type of the fields and outputs,
// T the arithmetic (and offset) type: FP64 fields gathered in FP32 are
// rounded to FP32 as they are staged (as a separate conversion would) and
// the FP32 results widened on the store, so every variant gives the same
// bits as one field at a time in T.
constexpr int kIX = 64, kIY = 4, kSX = kIX + 8, kSY = kIY + 5, kRing = 6;
// Elements of one staged plane, padded to a multiple of 64 dwords: the lanes
// of a wave read the same tap at their own columns, and lanes whose stencils
// start on different planes (vertical departures of either sign) then hit
// the banks of their columns instead of colliding at a shifted bank.
constexpr int kPlane = ((kSY * kSX + 63) / 64) * 64;
// Staged elements per thread and plane.
constexpr int kStageN = (kSY * kSX + kIX * kIY - 1) / (kIX * kIY);

template <typename TF, int NF>
struct FieldSet {
  const TF *f[NF];
  TF *o[NF];
};

template <typename T, typename TF, int NF>
__device__ __forceinline__ void tricubic_multi_body(const Dims &d, const FieldSet<TF, NF> &fs,
                                                    const T *__restrict__ co, int lim, int kchunk, int blk,
                                                    int nblk) {
  __shared__ T ring[NF][kRing * kPlane];
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int nbx = (nx + kIX - 1) / kIX, nby = (ny + kIY - 1) / kIY;
  const int tile = xcd_tile(blk, nblk);
  const int bx = tile % nbx, by = (tile / nbx) % nby, bz = tile / (nbx * nby);
  const int tid = threadIdx.y * kIX + threadIdx.x;
  const int xs = bx * kIX - 2, ys = by * kIY - 2;
  const int X = bx * kIX + threadIdx.x, Y = by * kIY + threadIdx.y;
  const bool own = X < nx && Y < ny;
  const int k0 = bz * kchunk, k1 = min(k0 + kchunk, nz);
  const uint32_t plane = (uint32_t)nx * ny;
  const uint32_t n = plane * (uint32_t)nz;
  // the clamped source offset within a level of each element this thread
  // stages (the same for every level), -1 past the plane
  int goff[kStageN];
#pragma unroll
  for (int r = 0; r < kStageN; ++r) {
    const int q = tid + r * kIX * kIY;
    const int ly = q / kSX, lx = q - ly * kSX;
    const int gx = min(max(xs + lx, 0), nx - 1), gy = min(max(ys + ly, 0), ny - 1);
    goff[r] = q < kSY * kSX ? gy * nx + gx : -1;
  }
  auto stage = [&](int kk) {
    const int slot = ((kk % kRing) + kRing) % kRing;
    const uint32_t zoff = (uint32_t)min(max(kk, 0), nz - 1) * plane;
#pragma unroll
    for (int r = 0; r < kStageN; ++r)
      if (goff[r] >= 0) {
#pragma unroll
        for (int fi = 0; fi < NF; ++fi)
          ring[fi][slot * kPlane + tid + r * kIX * kIY] = (T)ldg(fs.f[fi], zoff + (uint32_t)goff[r]);
      }
  };
  for (int kk = k0 - 2; kk <= k0 + 2; ++kk) stage(kk);
  for (int k = k0; k < k1; ++k) {
    __syncthreads();  // level k-1 is done with plane k-3
    stage(k + 3);
    __syncthreads();
    if (!own) continue;
    const uint32_t p = (uint32_t)k * plane + (uint32_t)Y * nx + X;
    T tx, ty, tz;
    const int x0 = offset_split(X, ldg(co, p), nx, tx);
    const int y0 = offset_split(Y, ldg(co, n + p), ny, ty);
    const int z0 = offset_split(k, ldg(co, 2 * n + p), nz, tz);
    const bool staged = x0 - 1 >= xs && x0 + 2 < xs + kSX && y0 - 1 >= ys && y0 + 2 < ys + kSY &&
                        z0 - 1 >= k - 2 && z0 + 2 <= k + 3;
    T wx[4], wy[4], wz[4];
    cr_weights(tx, wx);
    cr_weights(ty, wy);
    cr_weights(tz, wz);
    T r[NF];
    if (staged) {
      // one LDS base per stencil plane; taps are constant offsets from it
      int sl[4];
      int slot = (z0 - 1) % kRing;
      if (slot < 0) slot += kRing;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        sl[c] = slot;
        slot = slot + 1 == kRing ? 0 : slot + 1;
      }
#pragma unroll
      for (int fi = 0; fi < NF; ++fi) {
        // one field at a time: without the scheduling barrier the compiler
        // hoists every field's taps into registers at once (occupancy 2)
        if (fi) __builtin_amdgcn_sched_barrier(0);
        const T *zb[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) zb[c] = &ring[fi][sl[c] * kPlane + (y0 - 1 - ys) * kSX + (x0 - 1 - xs)];
        r[fi] = tricubic_apply(wx, wy, wz, lim, [&](int a, int b, int c) { return zb[c][b * kSX + a]; });
      }
    } else {
      int ix[4], iy[4], iz[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ix[q] = min(max(x0 - 1 + q, 0), nx - 1);
        iy[q] = min(max(y0 - 1 + q, 0), ny - 1) * nx;
        iz[q] = min(max(z0 - 1 + q, 0), nz - 1);
      }
#pragma unroll
      for (int fi = 0; fi < NF; ++fi) r[fi] = tricubic_global(fs.f[fi], ix, iy, iz, plane, wx, wy, wz, lim);
    }
#pragma unroll
    for (int fi = 0; fi < NF; ++fi) stg(fs.o[fi], p, (TF)r[fi]);
  }
}

template <typename T>
__device__ __forceinline__ void tricubic_body(const Dims &d, const T *__restrict__ f,
                                              const T *__restrict__ co, T *__restrict__ out, int lim,
                                              int kchunk, int blk, int nblk) {
  FieldSet<T, 1> fs;
  fs.f[0] = f;
  fs.o[0] = out;
  tricubic_multi_body<T, T, 1>(d, fs, co, lim, kchunk, blk, nblk);
}

template <typename T, typename TF, int NF>
__global__ void __launch_bounds__(256) tricubic_multi_kernel(Dims d, FieldSet<TF, NF> fs,
                                                             const T *__restrict__ co, int lim, int kchunk) {
  tricubic_multi_body<T, TF, NF>(d, fs, co, lim, kchunk, blockIdx.x, gridDim.x);
}

template <typename T>
__global__ void __launch_bounds__(256) tricubic_kernel(Dims d, const T *__restrict__ f,
                                                       const T *__restrict__ co,
                                                       T *__restrict__ out, int lim, int kchunk) {
  tricubic_body(d, f, co, out, lim, kchunk, blockIdx.x, gridDim.x);
}

// Launch geometry of the staged gather for one field: ~8 level chunks per
// column keep enough blocks in flight.
__host__ __device__ inline int tricubic_chunk(int fnz) { return (fnz + 7) / 8 > 1 ? (fnz + 7) / 8 : 1; }
__host__ __device__ inline int tricubic_blocks(int fnx, int fny, int fnz) {
  const int kc = tricubic_chunk(fnz);
  return ((fnx + kIX - 1) / kIX) * ((fny + kIY - 1) / kIY) * ((fnz + kc - 1) / kc);
}

// Batched gather: member m = blockIdx.y, its table passed by value (no
// device-side descriptor upload, so a captured graph replays it as is).
template <typename T>
struct TricubicBatch {
  int M, lim;
  Dims d[kMaxMembers];
  const T *f[kMaxMembers];
  const T *co[kMaxMembers];
  T *out[kMaxMembers];
};

template <typename T>
__global__ void __launch_bounds__(256) tricubic_batched_kernel(TricubicBatch<T> tb) {
  const int m = blockIdx.y;
  const Dims &d = tb.d[m];
  const int nb = tricubic_blocks(d.nx, d.ny, d.nz);
  if ((int)blockIdx.x >= nb) return;
  tricubic_body(d, tb.f[m], tb.co[m], tb.out[m], tb.lim, tricubic_chunk(d.nz), blockIdx.x, nb);
}

// ---- deterministic transpose ------------------------------------------------
// Plan: sort output points by the base cell of their stencil (counting sort),
// then every destination cell gathers from the <= 7x7x7 base cells whose
// clamped taps can reach it, visiting points in increasing index order.

template <typename T>
__global__ void adj_key_kernel(Dims d, const T *__restrict__ co, int32_t *__restrict__ key,
                               int32_t *__restrict__ count) {
  const int64_t n = d.n();
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int X = (int)(p % d.nx), Y = (int)((p / d.nx) % d.ny), Z = (int)(p / ((int64_t)d.nx * d.ny));
  T t;
  const int x0 = offset_split(X, co[p], d.nx, t);
  const int y0 = offset_split(Y, co[n + p], d.ny, t);
  const int z0 = offset_split(Z, co[2 * n + p], d.nz, t);
  const int32_t k = (int32_t)d.at(x0, y0, z0);
  key[p] = k;
  atomicAdd(&count[k], 1);
}

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

// Exclusive block-local scan of count -> start, block totals -> sums.
__global__ void __launch_bounds__(256) scan_tile_kernel(const int32_t *__restrict__ in,
                                                        int32_t *__restrict__ out,
                                                        int32_t *__restrict__ sums, int64_t n) {
  __shared__ int32_t wsum[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int32_t run = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t i = base + q;
    const int32_t x = i < n ? in[i] : 0;
    v[q] = run;
    run += x;
  }
  // inclusive scan of per-thread totals across the block
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t incl = run;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  int32_t wofs = 0;
  for (int q = 0; q < wid; ++q) wofs += wsum[q];
  const int32_t excl = incl - run + wofs;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t i = base + q;
    if (i < n) out[i] = v[q] + excl;
  }
  if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = excl + run;
}

// Single-block exclusive scan of the block totals (nb <= 256*64).
__global__ void __launch_bounds__(256) scan_sums_kernel(int32_t *__restrict__ sums, int nb,
                                                        int32_t *__restrict__ total) {
  __shared__ int32_t wsum[kBlock / 64];
  constexpr int kI = 64;
  const int base = threadIdx.x * kI;
  int32_t run = 0;
  for (int q = 0; q < kI; ++q) {
    const int i = base + q;
    if (i < nb) run += sums[i];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t incl = run;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  int32_t wofs = 0;
  for (int q = 0; q < wid; ++q) wofs += wsum[q];
  int32_t acc = incl - run + wofs;
  for (int q = 0; q < kI; ++q) {
    const int i = base + q;
    if (i < nb) {
      const int32_t x = sums[i];
      sums[i] = acc;
      acc += x;
    }
  }
  if (threadIdx.x == kBlock - 1) *total = acc;
}

__global__ void scan_add_kernel(int32_t *__restrict__ out, const int32_t *__restrict__ sums,
                                int64_t n, int32_t *__restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] += sums[i / kScanTile];
  count[i] = 0;  // becomes the scatter cursor
}

__global__ void adj_scatter_kernel(int64_t n, const int32_t *__restrict__ key,
                                   const int32_t *__restrict__ start, int32_t *__restrict__ cur,
                                   int32_t *__restrict__ perm) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t k = key[p];
  const int32_t pos = atomicAdd(&cur[k], 1);
  perm[start[k] + pos] = (int32_t)p;
}

// Sort each (small) segment by point index so accumulation order is fixed.
__global__ void adj_segsort_kernel(int64_t n, const int32_t *__restrict__ start,
                                   int32_t *__restrict__ perm) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t s = start[k], e = start[k + 1];
  for (int32_t a = s + 1; a < e; ++a) {
    const int32_t x = perm[a];
    int32_t b = a - 1;
    while (b >= s && perm[b] > x) {
      perm[b + 1] = perm[b];
      --b;
    }
    perm[b + 1] = x;
  }
}

// (base, tap) pairs whose clamped tap index equals i on an axis of length n.
__device__ __forceinline__ int tap_candidates(int i, int n, int *b, int *t) {
  int c = 0;
  if (i + 1 <= n - 1) { b[c] = i + 1; t[c] = 0; ++c; }
  if (i == 0) { b[c] = 0; t[c] = 0; ++c; }
  b[c] = i; t[c] = 1; ++c;
  if (i - 1 >= 0) { b[c] = i - 1; t[c] = 2; ++c; }
  if (i == n - 1) { b[c] = n - 1; t[c] = 2; ++c; }
  if (i - 2 >= 0) { b[c] = i - 2; t[c] = 3; ++c; }
  if (i == n - 1) {
    if (n - 2 >= 0 && n - 2 != i - 2) { b[c] = n - 2; t[c] = 3; ++c; }
    if (n - 1 != i - 2) { b[c] = n - 1; t[c] = 3; ++c; }
  }
  return c;
}

template <typename T>
__global__ void __launch_bounds__(256) tricubic_adj_kernel(Dims d, const T *__restrict__ co,
                                                           const int32_t *__restrict__ perm,
                                                           const int32_t *__restrict__ start,
                                                           const T *__restrict__ g,
                                                           T *__restrict__ gf) {
  const int64_t n = d.n();
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int i = (int)(q % d.nx);
  const int j = (int)((q / d.nx) % d.ny);
  const int k = (int)(q / ((int64_t)d.nx * d.ny));
  int bx[8], txa[8], by[8], tya[8], bz[8], tza[8];
  const int nxc = tap_candidates(i, d.nx, bx, txa);
  const int nyc = tap_candidates(j, d.ny, by, tya);
  const int nzc = tap_candidates(k, d.nz, bz, tza);
  T acc = 0;
  for (int a = 0; a < nxc; ++a)
    for (int b = 0; b < nyc; ++b)
      for (int c = 0; c < nzc; ++c) {
        const int64_t key = d.at(bx[a], by[b], bz[c]);
        const int32_t s = start[key], e = start[key + 1];
        for (int32_t r = s; r < e; ++r) {
          const int64_t p = perm[r];
          const int X = (int)(p % d.nx), Y = (int)((p / d.nx) % d.ny);
          const int Z = (int)(p / ((int64_t)d.nx * d.ny));
          T tx, ty, tz;
          offset_split(X, co[p], d.nx, tx);
          offset_split(Y, co[n + p], d.ny, ty);
          offset_split(Z, co[2 * n + p], d.nz, tz);
          T wx[4], wy[4], wz[4];
          cr_weights(tx, wx);
          cr_weights(ty, wy);
          cr_weights(tz, wz);
          acc += wx[txa[a]] * (wy[tya[b]] * (wz[tza[c]] * g[p]));
        }
      }
  gf[q] = acc;
}

// ---- trajectories ---------------------------------------------------------

// Index-space velocities at staggering `loc` (steppers.py:276-316).
template <typename T>
__device__ __forceinline__ void index_velocity_body(const Stag &s, int loc, const T *__restrict__ u,
                                                   const T *__restrict__ v, const T *__restrict__ e,
                                                   const T *__restrict__ m2, T rdx, T rdy,
                                                   T *__restrict__ out, const Dims &L, int64_t p) {
  const int64_t n = L.n();
  if (p >= n) return;
  const int i = (int)(p % L.nx);
  const int j = (int)((p / L.nx) % L.ny);
  const int k = (int)(p / ((int64_t)L.nx * L.ny));
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  // column averages to mass points
  auto um = [&](int ii, int jj, int kk) { return (T)0.5 * (u[s.u.at(ii, jj, kk)] + u[s.u.at(ii + 1, jj, kk)]); };
  auto vm = [&](int ii, int jj, int kk) { return (T)0.5 * (v[s.v.at(ii, jj, kk)] + v[s.v.at(ii, jj + 1, kk)]); };
  auto em = [&](int ii, int jj, int kk) { return (T)0.5 * (e[s.w.at(ii, jj, kk)] + e[s.w.at(ii, jj, kk + 1)]); };
  T ul, vl, wl, mf;
  if (loc == 0) {
    ul = um(i, j, k); vl = vm(i, j, k); wl = em(i, j, k); mf = m2[s.m2(i, j)];
  } else if (loc == 1) {
    ul = u[s.u.at(i, j, k)];
    // op.avg(m->u) pads with the averaged edge values (operators.py:57-63,
    // 121-132): faces 0 and nx repeat faces 1 and nx-1 unless x is periodic.
    if (s.px && (i == 0 || i == nx)) {
      vl = (T)0.5 * (vm(0, j, k) + vm(nx - 1, j, k));
      wl = (T)0.5 * (em(0, j, k) + em(nx - 1, j, k));
    } else {
      const int f = min(max(i, 1), nx - 1);
      vl = (T)0.5 * (vm(f - 1, j, k) + vm(f, j, k));
      wl = (T)0.5 * (em(f - 1, j, k) + em(f, j, k));
    }
    mf = m2[s.u2(i, j)];
  } else if (loc == 2) {
    vl = v[s.v.at(i, j, k)];
    if (s.py && (j == 0 || j == ny)) {
      ul = (T)0.5 * (um(i, 0, k) + um(i, ny - 1, k));
      wl = (T)0.5 * (em(i, 0, k) + em(i, ny - 1, k));
    } else {
      const int f = min(max(j, 1), ny - 1);
      ul = (T)0.5 * (um(i, f - 1, k) + um(i, f, k));
      wl = (T)0.5 * (em(i, f - 1, k) + em(i, f, k));
    }
    mf = m2[s.v2(i, j)];
  } else {
    wl = e[s.w.at(i, j, k)];
    const int f = min(max(k, 1), nz - 1);
    ul = (T)0.5 * (um(i, j, f - 1) + um(i, j, f));
    vl = (T)0.5 * (vm(i, j, f - 1) + vm(i, j, f));
    mf = m2[s.m2(i, j)];
  }
  out[p] = (ul * mf) * rdx;
  out[n + p] = (vl * mf) * rdy;
  out[2 * n + p] = wl;
}

template <typename T>
__global__ void index_velocity_kernel(Stag s, int loc, const T *__restrict__ u,
                                      const T *__restrict__ v, const T *__restrict__ e,
                                      const T *__restrict__ m2, T rdx, T rdy,
                                      T *__restrict__ out, Dims L) {
  index_velocity_body(s, loc, u, v, e, m2, rdx, rdy, out, L,
                      (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// map_coordinates(order=1, mode='nearest') of one field at X + d
// (steppers.py:367): per axis i = floor(x), t = x - i, indices clamped.
template <typename T>
__device__ __forceinline__ void lin_split(int X, T d, int n, int &i0, int &i1, T &w1) {
  const T fl = floor(d);
  const int i = X + (int)bounded_floor(fl, n);
  w1 = d - fl;
  i0 = min(max(i, 0), n - 1);
  i1 = min(max(i + 1, 0), n - 1);
}

template <typename T>
__device__ __forceinline__ T trilinear_nearest(const T *__restrict__ f, Dims d, int X, int Y, int Z,
                                               T dx, T dy, T dz) {
  int x0, x1, y0, y1, z0, z1;
  T wx1, wy1, wz1;
  lin_split(X, dx, d.nx, x0, x1, wx1);
  lin_split(Y, dy, d.ny, y0, y1, wy1);
  lin_split(Z, dz, d.nz, z0, z1, wz1);
  const T wx0 = (T)1 - wx1, wy0 = (T)1 - wy1, wz0 = (T)1 - wz1;
  T r = 0;
  r += f[d.at(x0, y0, z0)] * (wx0 * wy0 * wz0);
  r += f[d.at(x0, y0, z1)] * (wx0 * wy0 * wz1);
  r += f[d.at(x0, y1, z0)] * (wx0 * wy1 * wz0);
  r += f[d.at(x0, y1, z1)] * (wx0 * wy1 * wz1);
  r += f[d.at(x1, y0, z0)] * (wx1 * wy0 * wz0);
  r += f[d.at(x1, y0, z1)] * (wx1 * wy0 * wz1);
  r += f[d.at(x1, y1, z0)] * (wx1 * wy1 * wz0);
  r += f[d.at(x1, y1, z1)] * (wx1 * wy1 * wz1);
  return r;
}

// Midpoint trajectory iteration (steppers.py:318-360); writes the departure
// offsets -alpha = x_d - x_a.
template <typename T>
__device__ __forceinline__ void departure_body(const Dims &L, const T *__restrict__ vel, T dt,
                                              int iters, T *__restrict__ out, int64_t p) {
  const int64_t n = L.n();
  if (p >= n) return;
  const int X = (int)(p % L.nx), Y = (int)((p / L.nx) % L.ny), Z = (int)(p / ((int64_t)L.nx * L.ny));
  T ax = dt * vel[p], ay = dt * vel[n + p], az = dt * vel[2 * n + p];
  for (int it = 0; it < iters; ++it) {
    const T mx = (T)-0.5 * ax, my = (T)-0.5 * ay, mz = (T)-0.5 * az;
    const T um = trilinear_nearest(vel, L, X, Y, Z, mx, my, mz);
    const T vm = trilinear_nearest(vel + n, L, X, Y, Z, mx, my, mz);
    const T wm = trilinear_nearest(vel + 2 * n, L, X, Y, Z, mx, my, mz);
    ax = dt * um; ay = dt * vm; az = dt * wm;
  }
  out[p] = -ax;
  out[n + p] = -ay;
  out[2 * n + p] = -az;
}

template <typename T>
__global__ void departure_kernel(Dims L, const T *__restrict__ vel, T dt, int iters,
                                 T *__restrict__ out) {
  departure_body(L, vel, dt, iters, out, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Dims of staggering loc (0=m, 1=u, 2=v, 3=w) of a grid.
inline __host__ __device__ Dims loc_dims(const SislGrid &g, int loc) {
  return Dims{g.nx + (loc == 1 ? 1 : 0), g.ny + (loc == 2 ? 1 : 0), g.nz + (loc == 3 ? 1 : 0)};
}

// Batched trajectories: member m = blockIdx.y, table by value.
template <typename T>
struct DepartureBatch {
  int M, loc, iters;
  T dt;
  SislGrid g[kMaxMembers];
  const T *u[kMaxMembers], *v[kMaxMembers], *e[kMaxMembers], *m2[kMaxMembers];
  T *scr[kMaxMembers], *out[kMaxMembers];
  T rdx[kMaxMembers], rdy[kMaxMembers];
};

template <typename T>
__global__ void index_velocity_batched_kernel(DepartureBatch<T> db) {
  const int m = blockIdx.y;
  const Stag s = Stag::make(db.g[m]);
  index_velocity_body(s, db.loc, db.u[m], db.v[m], db.e[m], db.m2[m], db.rdx[m], db.rdy[m], db.scr[m],
                      loc_dims(db.g[m], db.loc), (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T>
__global__ void departure_batched_kernel(DepartureBatch<T> db) {
  const int m = blockIdx.y;
  departure_body(loc_dims(db.g[m], db.loc), (const T *)db.scr[m], db.dt, db.iters, db.out[m],
                 (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T>
int tricubic(int fnx, int fny, int fnz, const T *f, const T *co, T *out, int lim, void *st) {
  if (!f || !co || !out || fnx < 1 || fny < 1 || fnz < 1) return SISL_ERR_ARG;
  Dims d{fnx, fny, fnz};
  if ((uint64_t)d.n() * 3 * sizeof(T) >= (1ull << 32)) return SISL_ERR_ARG;  // 32-bit offsets
  const int kc = tricubic_chunk(fnz);
  const int nb = tricubic_blocks(fnx, fny, fnz);
  const int tid = timer_begin(SISL_T_TRICUBIC, sizeof(T) == 4, (double)sizeof(T) * 5.0 * d.n(),
                              nullptr, nullptr, as_stream(st));
  hipLaunchKernelGGL(tricubic_kernel<T>, dim3(nb), dim3(kIX, kIY), 0, as_stream(st), d, f, co, out,
                     lim, kc);
  timer_end(tid, as_stream(st));
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

// NF fields on one set of departure points; TF storage, T arithmetic.
template <typename T, typename TF, int NF>
int tricubic_multi(int fnx, int fny, int fnz, const TF *const *f, const T *co, TF *const *out, int lim,
                   void *st) {
  if (!f || !co || !out || fnx < 1 || fny < 1 || fnz < 1) return SISL_ERR_ARG;
  FieldSet<TF, NF> fs;
  for (int q = 0; q < NF; ++q) {
    if (!f[q] || !out[q]) return SISL_ERR_ARG;
    fs.f[q] = f[q];
    fs.o[q] = out[q];
  }
  Dims d{fnx, fny, fnz};
  if ((uint64_t)d.n() * 3 * sizeof(T) >= (1ull << 32) || (uint64_t)d.n() * sizeof(TF) >= (1ull << 32))
    return SISL_ERR_ARG;  // 32-bit offsets
  const int kc = tricubic_chunk(fnz);
  const int nb = tricubic_blocks(fnx, fny, fnz);
  const double bytes = ((double)NF * 2.0 * sizeof(TF) + 3.0 * sizeof(T)) * d.n();
  const int tid = timer_begin(SISL_T_TRICUBIC, sizeof(T) == 4, bytes, nullptr, nullptr, as_stream(st));
  hipLaunchKernelGGL((tricubic_multi_kernel<T, TF, NF>), dim3(nb), dim3(kIX, kIY), 0, as_stream(st), d, fs, co,
                     lim, kc);
  timer_end(tid, as_stream(st));
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int tricubic_plan(int fnx, int fny, int fnz, const T *co, int32_t *key, int32_t *perm,
                  int32_t *start, int32_t *count, int32_t *sums, void *st) {
  if (!co || !key || !perm || !start || !count || !sums) return SISL_ERR_ARG;
  Dims d{fnx, fny, fnz};
  const int64_t n = d.n();
  const int nb = (int)((n + kScanTile - 1) / kScanTile);
  if (nb > kBlock * 64 || n >= (int64_t)1 << 31) return SISL_ERR_ARG;
  hipStream_t s = as_stream(st);
  if (hipMemsetAsync(count, 0, sizeof(int32_t) * n, s) != hipSuccess) return SISL_ERR_LAUNCH;
  hipLaunchKernelGGL(adj_key_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, d, co, key, count);
  hipLaunchKernelGGL(scan_tile_kernel, dim3(nb), dim3(kBlock), 0, s, count, start, sums, n);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kBlock), 0, s, sums, nb, start + n);
  hipLaunchKernelGGL(scan_add_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, start, sums, n, count);
  hipLaunchKernelGGL(adj_scatter_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, n, key, start,
                     count, perm);
  hipLaunchKernelGGL(adj_segsort_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, n, start, perm);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int tricubic_adj(int fnx, int fny, int fnz, const T *co, const int32_t *perm, const int32_t *start,
                 const T *g, T *gf, void *st) {
  if (!co || !perm || !start || !g || !gf) return SISL_ERR_ARG;
  Dims d{fnx, fny, fnz};
  hipLaunchKernelGGL(tricubic_adj_kernel<T>, dim3(grid_for(d.n())), dim3(kBlock), 0,
                     as_stream(st), d, co, perm, start, g, gf);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

// ---- adjoint over sorted rows -------------------------------------------------
// The plan's perm lists the points sorted by base cell key = at(bx, by, bz)
// (x fastest), so the sources whose base lies in x-range [x0, x1] of one
// (by, bz) row occupy one contiguous range of sorted positions.  A
// destination (i, j, k) can only receive from bases in [i-2, i+1] x [j-2, j+1]
// x [k-2, k+1] (clamped to the grid), i.e. from 16 such ranges.  Per sorted
// position the plan extension keeps the source's 12 Catmull-Rom weights and
// its base x; each call gathers g into sorted order once; the transpose then
// sums, per destination, weight x weight x weight x g over its 16 ranges in a
// fixed order (deterministic).

// weight of index i of an axis of length n in the 4-tap stencil of base b
template <typename T>
__device__ __forceinline__ T tap_weight(int i, int b, int n, const T *__restrict__ w) {
  if (b >= 1 && b + 2 <= n - 1) return w[i - b + 1];  // no clamping: exactly one tap
  T s = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t)
    if (min(max(b - 1 + t, 0), n - 1) == i) s += w[t];
  return s;
}

// Sorted-row plan: per sorted position r the source's fractions (tx, ty, tz)
// and base x, as four values of T (the base exactly representable).
template <typename T>
__global__ void __launch_bounds__(256) adj_frac_kernel(Dims d, const T *__restrict__ co,
                                                       const int32_t *__restrict__ perm, T *__restrict__ fr) {
  const int64_t n = d.n();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t p = perm[r];
  const int X = (int)(p % d.nx), Y = (int)((p / d.nx) % d.ny), Z = (int)(p / ((int64_t)d.nx * d.ny));
  T tx, ty, tz;
  const int bx = offset_split(X, co[p], d.nx, tx);
  offset_split(Y, co[n + p], d.ny, ty);
  offset_split(Z, co[2 * n + p], d.nz, tz);
  fr[4 * r] = tx;
  fr[4 * r + 1] = ty;
  fr[4 * r + 2] = tz;
  fr[4 * r + 3] = (T)bx;
}

template <typename T>
__global__ void __launch_bounds__(256) adj_gather_kernel(int64_t n, const int32_t *__restrict__ perm,
                                                         const T *__restrict__ g, T *__restrict__ gs) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) gs[r] = g[perm[r]];
}

// Contribution of sorted source (fractions f[0..2], base x f[3], value gv) to
// destination (i, j, k) from its (by, bz) row: the product of the three
// tap weights (clamped taps summed), accumulated as wx (wy (wz g)).
template <typename T>
__device__ __forceinline__ T adj_term(const T *f, T gv, int i, int j, int k, int by, int bz, Dims d) {
  T wx[4], wy[4], wz[4];
  cr_weights(f[0], wx);
  cr_weights(f[1], wy);
  cr_weights(f[2], wz);
  const T a = tap_weight(i, (int)f[3], d.nx, wx);
  const T b = tap_weight(j, by, d.ny, wy);
  const T c = tap_weight(k, bz, d.nz, wz);
  return a * (b * (c * gv));
}

// Transpose gather over the sorted-row plan.  One workgroup per 64 x 4 tile
// of destinations on one level k.  The sources of the 4 x 7 (bz, by) rows
// that can reach the tile (bases x in [i0-2, i0+64]) are staged in LDS, each
// with the work that does not depend on the destination column done once:
// its effective x weights for the four columns bx-1..bx+2 (clamped taps
// summed) and, for each of the tile's four rows j, wy(j) (wz(k) g).  Every
// destination then sums, in a fixed order (bz, by, sorted position), wx times
// that product over the sources of the 16 rows and 4 bases around it -- the
// same products in the same order as the plain loop, which a tile whose
// sources do not fit in LDS runs from global memory.
constexpr int kAdjTX = 64, kAdjTY = 4, kAdjRows = 4 * (kAdjTY + 3), kAdjW = kAdjTX + 4;
template <typename T> struct AdjCap { static constexpr int v = 2048; };
template <> struct AdjCap<double> { static constexpr int v = 1024; };

template <typename T>
__global__ void __launch_bounds__(256) tricubic_adj_tile_kernel(Dims d, const int32_t *__restrict__ start,
                                                                const T *__restrict__ fr,
                                                                const T *__restrict__ gs,
                                                                T *__restrict__ gf) {
  constexpr int cap = AdjCap<T>::v;
  __shared__ T sx[cap * 4];        // effective x weights of columns bx-1 .. bx+2
  __shared__ T sy[cap * kAdjTY];   // wy(j0 + t) (wz(k) g) for the tile's rows
  __shared__ int sb[cap];          // base x
  __shared__ int st[kAdjRows][kAdjW];  // per row: sorted index of the first source of bases xa..xb+1
  __shared__ int off[kAdjRows + 1];
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int nbx = (nx + kAdjTX - 1) / kAdjTX, nby = (ny + kAdjTY - 1) / kAdjTY;
  const int tile = (int)blockIdx.x;
  const int i0 = (tile % nbx) * kAdjTX, j0 = ((tile / nbx) % nby) * kAdjTY, k = tile / (nbx * nby);
  static_assert(kAdjTX * kAdjTY == 256, "the staging loops assume 256 threads");
  const int tid = threadIdx.y * kAdjTX + threadIdx.x;
  const int xa = max(i0 - 2, 0), xb = min(i0 + kAdjTX, nx - 1);
  const int wspan = xb + 2 - xa;  // start entries per row: bases xa..xb and the end of xb
  // row slot q = dz * (kAdjTY + 3) + dy: bz = k - 2 + dz, by = j0 - 2 + dy.
  // Every load of a thread is issued before any is used (the loops below have
  // fixed trip counts and are unrolled): the staging is latency-bound.
  constexpr int kStE = kAdjRows * kAdjW, kStPer = (kStE + 255) / 256;
  {
    int sv[kStPer];
#pragma unroll
    for (int u = 0; u < kStPer; ++u) {
      const int e = tid + 256 * u;
      const int q = e / kAdjW, x = e % kAdjW;
      const int bz = k - 2 + q / (kAdjTY + 3), by = j0 - 2 + q % (kAdjTY + 3);
      sv[u] = (e < kStE && x < wspan && bz >= 0 && bz < nz && by >= 0 && by < ny)
                  ? start[((uint32_t)bz * ny + by) * nx + xa + x]
                  : 0;
    }
#pragma unroll
    for (int u = 0; u < kStPer; ++u) {
      const int e = tid + 256 * u;
      if (e < kStE) st[e / kAdjW][e % kAdjW] = sv[u];
    }
  }
  __syncthreads();
  if (tid == 0) {
    int c[kAdjRows];
#pragma unroll
    for (int q = 0; q < kAdjRows; ++q) c[q] = st[q][wspan - 1] - st[q][0];
    int o = 0;
#pragma unroll
    for (int q = 0; q < kAdjRows; ++q) {
      off[q] = o;
      o += c[q];
    }
    off[kAdjRows] = o;
  }
  __syncthreads();
  const int total = off[kAdjRows];
  const bool staged = total <= cap;
  if (staged) {
    constexpr int kPer = cap / 256;
    T f[kPer][4], gv[kPer];
    int qs[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = tid + 256 * u;
      qs[u] = -1;
      if (e < total) {
        int lo = 0, hi = kAdjRows - 1;  // the row q with off[q] <= e < off[q + 1]
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (off[mid] <= e) lo = mid;
          else hi = mid - 1;
        }
        qs[u] = lo;
        const uint32_t r = (uint32_t)(st[lo][0] + (e - off[lo]));
#pragma unroll
        for (int c = 0; c < 4; ++c) f[u][c] = fr[4 * r + c];
        gv[u] = gs[r];
      }
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = tid + 256 * u, q = qs[u];
      if (q < 0) continue;
      const int bz = k - 2 + q / (kAdjTY + 3), by = j0 - 2 + q % (kAdjTY + 3);
      T wx[4], wy[4], wz[4];
      cr_weights(f[u][0], wx);
      cr_weights(f[u][1], wy);
      cr_weights(f[u][2], wz);
      const int bx = (int)f[u][3];
      sb[e] = bx;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int ii = bx - 1 + t;
        sx[4 * e + t] = (ii >= 0 && ii < nx) ? tap_weight(ii, bx, nx, wx) : (T)0;
      }
      const T zg = tap_weight(k, bz, nz, wz) * gv[u];
#pragma unroll
      for (int t = 0; t < kAdjTY; ++t) {
        const int jj = j0 + t;
        sy[kAdjTY * e + t] = (jj < ny && jj >= by - 1 && jj <= by + 2) ? tap_weight(jj, by, ny, wy) * zg : (T)0;
      }
    }
  }
  __syncthreads();
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int x0 = max(i - 2, 0), x1 = min(i + 1, nx - 1);
  T acc = 0;
  for (int bz = max(k - 2, 0); bz <= min(k + 1, nz - 1); ++bz)
    for (int by = max(j - 2, 0); by <= min(j + 1, ny - 1); ++by) {
      const int q = (bz - k + 2) * (kAdjTY + 3) + (by - j0 + 2);
      const int ra = st[q][x0 - xa], re = st[q][x1 + 1 - xa];
      if (staged) {
        const int base = off[q] - st[q][0];
        for (int r = ra; r < re; ++r) {
          const int e = base + r;
          acc += sx[4 * e + (i - sb[e] + 1)] * sy[kAdjTY * e + threadIdx.y];
        }
      } else {
        for (int r = ra; r < re; ++r) {
          T f[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) f[c] = fr[4 * (int64_t)r + c];
          acc += adj_term(f, gs[r], i, j, k, by, bz, d);
        }
      }
    }
  gf[((uint32_t)k * ny + j) * nx + i] = acc;
}

template <typename T>
int tricubic_adj_frac(int fnx, int fny, int fnz, const T *co, const int32_t *perm, T *fr, void *st) {
  if (!co || !perm || !fr) return SISL_ERR_ARG;
  Dims d{fnx, fny, fnz};
  if (d.n() >= (int64_t)1 << 29) return SISL_ERR_ARG;  // 4n entries with 32-bit indices
  hipLaunchKernelGGL(adj_frac_kernel<T>, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(st), d, co, perm, fr);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int tricubic_adj_rows(int fnx, int fny, int fnz, const int32_t *perm, const int32_t *start, const T *fr,
                      const T *g, T *gs, T *gf, void *st) {
  if (!perm || !start || !fr || !g || !gs || !gf || g == gf) return SISL_ERR_ARG;
  Dims d{fnx, fny, fnz};
  const int64_t n = d.n();
  if (n >= (int64_t)1 << 29) return SISL_ERR_ARG;
  hipStream_t hs = as_stream(st);
  // g in, gs out and in, gf out; plan: perm, start, 4 values per point
  const double bytes = (double)n * (8.0 * sizeof(T) + 2.0 * 4.0);
  const int tid = timer_begin(SISL_T_TRICUBIC_ADJ, sizeof(T) == 4, bytes, nullptr, nullptr, hs);
  hipLaunchKernelGGL(adj_gather_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, hs, n, perm, g, gs);
  const int nbx = (fnx + kAdjTX - 1) / kAdjTX, nby = (fny + kAdjTY - 1) / kAdjTY;
  hipLaunchKernelGGL(tricubic_adj_tile_kernel<T>, dim3((unsigned)(nbx * nby * fnz)), dim3(kAdjTX, kAdjTY), 0, hs,
                     d, start, fr, (const T *)gs, gf);
  timer_end(tid, hs);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int departure(const SislGrid *g, int loc, const T *u, const T *v, const T *e, const T *m2,
              double dx, double dy, double dt, int iters, T *scratch3, T *out, void *st) {
  if (!g || !u || !v || !e || !m2 || !scratch3 || !out || loc < 0 || loc > 3) return SISL_ERR_ARG;
  Stag s = Stag::make(*g);
  Dims L = loc == 0 ? s.m : loc == 1 ? s.u : loc == 2 ? s.v : s.w;
  hipStream_t hs = as_stream(st);
  hipLaunchKernelGGL(index_velocity_kernel<T>, dim3(grid_for(L.n())), dim3(kBlock), 0, hs, s, loc,
                     u, v, e, m2, (T)(1.0 / dx), (T)(1.0 / dy), scratch3, L);
  hipLaunchKernelGGL(departure_kernel<T>, dim3(grid_for(L.n())), dim3(kBlock), 0, hs, L,
                     (const T *)scratch3, (T)dt, iters, out);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int tricubic_batched(int M, const int32_t *dims, const T *const *f, const T *const *co, T *const *out,
                     int lim, void *st) {
  if (M < 1 || M > kMaxMembers || !dims || !f || !co || !out) return SISL_ERR_ARG;
  TricubicBatch<T> tb;
  tb.M = M;
  tb.lim = lim;
  int maxb = 1;
  double bytes = 0;
  for (int m = 0; m < M; ++m) {
    const Dims d{dims[3 * m], dims[3 * m + 1], dims[3 * m + 2]};
    if (d.nx < 1 || d.ny < 1 || d.nz < 1 || !f[m] || !co[m] || !out[m]) return SISL_ERR_ARG;
    if ((uint64_t)d.n() * 3 * sizeof(T) >= (1ull << 32)) return SISL_ERR_ARG;
    tb.d[m] = d;
    tb.f[m] = f[m];
    tb.co[m] = co[m];
    tb.out[m] = out[m];
    maxb = max(maxb, tricubic_blocks(d.nx, d.ny, d.nz));
    bytes += (double)sizeof(T) * 5.0 * d.n();
  }
  hipStream_t hs = as_stream(st);
  const int tid = timer_begin(SISL_T_TRICUBIC, sizeof(T) == 4, bytes, nullptr, nullptr, hs);
  hipLaunchKernelGGL(tricubic_batched_kernel<T>, dim3(maxb, M), dim3(kIX, kIY), 0, hs, tb);
  timer_end(tid, hs);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int departure_batched(int M, const SislGrid *g, int loc, const T *const *u, const T *const *v,
                      const T *const *e, const T *const *m2, const double *dx, const double *dy, double dt,
                      int iters, T *const *scr, T *const *out, void *st) {
  if (M < 1 || M > kMaxMembers || !g || !u || !v || !e || !m2 || !dx || !dy || !scr || !out || loc < 0 ||
      loc > 3)
    return SISL_ERR_ARG;
  DepartureBatch<T> db;
  db.M = M;
  db.loc = loc;
  db.iters = iters;
  db.dt = (T)dt;
  int64_t maxn = 1;
  for (int m = 0; m < M; ++m) {
    if (!u[m] || !v[m] || !e[m] || !m2[m] || !scr[m] || !out[m] || g[m].nx < 2 || g[m].nz < 2)
      return SISL_ERR_ARG;
    db.g[m] = g[m];
    db.u[m] = u[m];
    db.v[m] = v[m];
    db.e[m] = e[m];
    db.m2[m] = m2[m];
    db.scr[m] = scr[m];
    db.out[m] = out[m];
    db.rdx[m] = (T)(1.0 / dx[m]);
    db.rdy[m] = (T)(1.0 / dy[m]);
    maxn = std::max<int64_t>(maxn, loc_dims(g[m], loc).n());
  }
  hipStream_t hs = as_stream(st);
  const dim3 grid(grid_for(maxn), M);
  hipLaunchKernelGGL(index_velocity_batched_kernel<T>, grid, dim3(kBlock), 0, hs, db);
  hipLaunchKernelGGL(departure_batched_kernel<T>, grid, dim3(kBlock), 0, hs, db);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

// ---- adjoint of the limited gather ------------------------------------------------
// The limiter is clip(f, lo, hi) = min(max(f, lo), hi) with lo / hi the
// chained binary minimum / maximum of the 8 cell corners in the reference's
// order (operators.py:188-207).  JAX differentiates a binary min / max with
// the balanced rule: the cotangent goes to the operand equal to the result,
// split in half when both are.  Per output point the split kernel recomputes
// the unclipped interpolant and both chains and writes the interpolant's
// share of the cotangent (gi, transposed by the unlimited adjoint) and the
// eight corners' shares (gc, corner q = ox + 2 oy + 4 oz at stride n); the
// corner kernel gathers the corner shares per destination cell over the
// plan's base-cell buckets in a fixed order (deterministic).

// d min(a, b) / da and / db, and the same for max, with balanced ties
template <typename T>
__device__ __forceinline__ void dmin(T a, T b, T &da, T &db) {
  da = a < b ? (T)1 : (a == b ? (T)0.5 : (T)0);
  db = b < a ? (T)1 : (a == b ? (T)0.5 : (T)0);
}
template <typename T>
__device__ __forceinline__ void dmax(T a, T b, T &da, T &db) {
  da = a > b ? (T)1 : (a == b ? (T)0.5 : (T)0);
  db = b > a ? (T)1 : (a == b ? (T)0.5 : (T)0);
}

template <typename T>
__global__ void __launch_bounds__(256) tricubic_limit_split_kernel(Dims d, const T *__restrict__ f,
                                                                   const T *__restrict__ co,
                                                                   const T *__restrict__ g,
                                                                   T *__restrict__ gi, T *__restrict__ gc) {
  const int64_t n = d.n();
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int X = (int)(p % nx), Y = (int)((p / nx) % ny), Z = (int)(p / ((int64_t)nx * ny));
  T tx, ty, tz;
  const int x0 = offset_split(X, co[p], nx, tx);
  const int y0 = offset_split(Y, co[n + p], ny, ty);
  const int z0 = offset_split(Z, co[2 * n + p], nz, tz);
  int ix[4], iy[4], iz[4];
  for (int q = 0; q < 4; ++q) {
    ix[q] = min(max(x0 - 1 + q, 0), nx - 1);
    iy[q] = min(max(y0 - 1 + q, 0), ny - 1);
    iz[q] = min(max(z0 - 1 + q, 0), nz - 1);
  }
  auto F = [&](int a, int b, int c) { return f[d.at(ix[a], iy[b], iz[c])]; };
  const T fi = tricubic_point(tx, ty, tz, 0, F);
  T v[8], cmin[8], cmax[8];
  for (int q = 0; q < 8; ++q) v[q] = F(1 + (q & 1), 1 + ((q >> 1) & 1), 1 + ((q >> 2) & 1));
  cmin[0] = cmax[0] = v[0];
  for (int q = 1; q < 8; ++q) {
    cmin[q] = v[q] < cmin[q - 1] ? v[q] : cmin[q - 1];
    cmax[q] = v[q] > cmax[q - 1] ? v[q] : cmax[q - 1];
  }
  const T lo = cmin[7], hi = cmax[7];
  const T m = fi < lo ? lo : fi;  // max(f, lo)
  T dm_f, dm_lo, do_m, do_hi;
  dmax(fi, lo, dm_f, dm_lo);
  dmin(m, hi, do_m, do_hi);
  const T gp = g[p];
  const T gm = gp * do_m;
  gi[p] = gm * dm_f;
  T glo = gm * dm_lo, ghi = gp * do_hi;
  T gv[8];
  for (int q = 7; q >= 1; --q) {  // back through the chains
    T a_, b_, c_, e_;
    dmin(cmin[q - 1], v[q], a_, b_);
    dmax(cmax[q - 1], v[q], c_, e_);
    gv[q] = glo * b_ + ghi * e_;
    glo *= a_;
    ghi *= c_;
  }
  gv[0] = glo + ghi;
  for (int q = 0; q < 8; ++q) gc[q * n + p] = gv[q];
}

// (base, corner offset) pairs whose clamped corner is cell i on an axis of n cells
__device__ __forceinline__ int corner_candidates(int i, int n, int *b, int *o) {
  int c = 0;
  b[c] = i; o[c] = 0; ++c;
  if (i >= 1) { b[c] = i - 1; o[c] = 1; ++c; }
  if (i == n - 1) { b[c] = i; o[c] = 1; ++c; }  // x1 = min(x0 + 1, n - 1) = x0
  return c;
}

template <typename T>
__global__ void __launch_bounds__(256) tricubic_limit_corner_kernel(Dims d, const int32_t *__restrict__ perm,
                                                                    const int32_t *__restrict__ start,
                                                                    const T *__restrict__ gc,
                                                                    T *__restrict__ gf) {
  const int64_t n = d.n();
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int X = (int)(p % d.nx), Y = (int)((p / d.nx) % d.ny), Z = (int)(p / ((int64_t)d.nx * d.ny));
  int bx[3], ox[3], by[3], oy[3], bz[3], oz[3];
  const int cx = corner_candidates(X, d.nx, bx, ox);
  const int cy = corner_candidates(Y, d.ny, by, oy);
  const int cz = corner_candidates(Z, d.nz, bz, oz);
  T acc = 0;
  for (int c = 0; c < cz; ++c)
    for (int b = 0; b < cy; ++b)
      for (int a = 0; a < cx; ++a) {
        const int64_t key = d.at(bx[a], by[b], bz[c]);
        const int q = ox[a] + 2 * oy[b] + 4 * oz[c];
        const int32_t s = start[key], e = start[key + 1];
        for (int32_t r = s; r < e; ++r) acc += gc[q * n + perm[r]];
      }
  gf[p] += acc;
}

template <typename T>
int tricubic_limited_split(int fnx, int fny, int fnz, const T *f, const T *co, const T *g, T *gi, T *gc,
                           void *st) {
  if (!f || !co || !g || !gi || !gc || fnx < 1 || fny < 1 || fnz < 1) return SISL_ERR_ARG;
  Dims d{fnx, fny, fnz};
  if (8 * d.n() >= ((int64_t)1 << 31)) return SISL_ERR_ARG;
  hipLaunchKernelGGL(tricubic_limit_split_kernel<T>, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(st), d, f,
                     co, g, gi, gc);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T>
int tricubic_limited_corners(int fnx, int fny, int fnz, const int32_t *perm, const int32_t *start, const T *gc,
                             T *gf, void *st) {
  if (!perm || !start || !gc || !gf || fnx < 1 || fny < 1 || fnz < 1) return SISL_ERR_ARG;
  Dims d{fnx, fny, fnz};
  if (8 * d.n() >= ((int64_t)1 << 31)) return SISL_ERR_ARG;
  hipLaunchKernelGGL(tricubic_limit_corner_kernel<T>, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(st), d,
                     perm, start, gc, gf);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

}  // namespace sisl

extern "C" {

int sisl_tricubic_limited_adj_split_f64(int32_t a, int32_t b, int32_t c, const double *field, const double *co,
                                        const double *gout, double *gi, double *gc, void *s) {
  return sisl::tricubic_limited_split<double>(a, b, c, field, co, gout, gi, gc, s);
}
int sisl_tricubic_limited_adj_split_f32(int32_t a, int32_t b, int32_t c, const float *field, const float *co,
                                        const float *gout, float *gi, float *gc, void *s) {
  return sisl::tricubic_limited_split<float>(a, b, c, field, co, gout, gi, gc, s);
}
int sisl_tricubic_limited_adj_corners_f64(int32_t a, int32_t b, int32_t c, const int32_t *perm,
                                          const int32_t *start, const double *gc, double *gf, void *s) {
  return sisl::tricubic_limited_corners<double>(a, b, c, perm, start, gc, gf, s);
}
int sisl_tricubic_limited_adj_corners_f32(int32_t a, int32_t b, int32_t c, const int32_t *perm,
                                          const int32_t *start, const float *gc, float *gf, void *s) {
  return sisl::tricubic_limited_corners<float>(a, b, c, perm, start, gc, gf, s);
}

int sisl_tricubic_batched_f64(int32_t M, const int32_t *dims, const double *const *f,
                              const double *const *co, double *const *out, int32_t lim, void *s) {
  return sisl::tricubic_batched<double>(M, dims, f, co, out, lim, s);
}
int sisl_tricubic_batched_f32(int32_t M, const int32_t *dims, const float *const *f,
                              const float *const *co, float *const *out, int32_t lim, void *s) {
  return sisl::tricubic_batched<float>(M, dims, f, co, out, lim, s);
}
int sisl_departure_batched_f64(int32_t M, const SislGrid *g, int32_t loc, const double *const *u,
                               const double *const *v, const double *const *e,
                               const double *const *m2, const double *dx, const double *dy, double dt,
                               int32_t it, double *const *scr, double *const *out, void *s) {
  return sisl::departure_batched<double>(M, g, loc, u, v, e, m2, dx, dy, dt, it, scr, out, s);
}
int sisl_departure_batched_f32(int32_t M, const SislGrid *g, int32_t loc, const float *const *u,
                               const float *const *v, const float *const *e,
                               const float *const *m2, const double *dx, const double *dy, double dt,
                               int32_t it, float *const *scr, float *const *out, void *s) {
  return sisl::departure_batched<float>(M, g, loc, u, v, e, m2, dx, dy, dt, it, scr, out, s);
}

int32_t sisl_scan_blocks(int64_t n) { return (int32_t)((n + sisl::kScanTile - 1) / sisl::kScanTile); }

int sisl_tricubic_f64(int32_t a, int32_t b, int32_t c, const double *f, const double *co,
                      double *out, int32_t lim, void *s) {
  return sisl::tricubic<double>(a, b, c, f, co, out, lim, s);
}
int sisl_tricubic_f32(int32_t a, int32_t b, int32_t c, const float *f, const float *co,
                      float *out, int32_t lim, void *s) {
  return sisl::tricubic<float>(a, b, c, f, co, out, lim, s);
}
int sisl_tricubic2_f64(int32_t a, int32_t b, int32_t c, const double *f0, const double *f1, const double *co,
                       double *o0, double *o1, int32_t lim, void *s) {
  const double *f[2] = {f0, f1};
  double *o[2] = {o0, o1};
  return sisl::tricubic_multi<double, double, 2>(a, b, c, f, co, o, lim, s);
}
int sisl_tricubic2_f32(int32_t a, int32_t b, int32_t c, const float *f0, const float *f1, const float *co,
                       float *o0, float *o1, int32_t lim, void *s) {
  const float *f[2] = {f0, f1};
  float *o[2] = {o0, o1};
  return sisl::tricubic_multi<float, float, 2>(a, b, c, f, co, o, lim, s);
}
int sisl_tricubic_mixed(int32_t a, int32_t b, int32_t c, const double *f0, const float *co, double *o0,
                        int32_t lim, void *s) {
  const double *f[1] = {f0};
  double *o[1] = {o0};
  return sisl::tricubic_multi<float, double, 1>(a, b, c, f, co, o, lim, s);
}
int sisl_tricubic2_mixed(int32_t a, int32_t b, int32_t c, const double *f0, const double *f1, const float *co,
                         double *o0, double *o1, int32_t lim, void *s) {
  const double *f[2] = {f0, f1};
  double *o[2] = {o0, o1};
  return sisl::tricubic_multi<float, double, 2>(a, b, c, f, co, o, lim, s);
}
int sisl_tricubic_adj_plan_f64(int32_t a, int32_t b, int32_t c, const double *co, int32_t *key,
                               int32_t *perm, int32_t *start, int32_t *count, int32_t *sums,
                               void *s) {
  return sisl::tricubic_plan<double>(a, b, c, co, key, perm, start, count, sums, s);
}
int sisl_tricubic_adj_plan_f32(int32_t a, int32_t b, int32_t c, const float *co, int32_t *key,
                               int32_t *perm, int32_t *start, int32_t *count, int32_t *sums,
                               void *s) {
  return sisl::tricubic_plan<float>(a, b, c, co, key, perm, start, count, sums, s);
}
int sisl_tricubic_adj_f64(int32_t a, int32_t b, int32_t c, const double *co, const int32_t *perm,
                          const int32_t *start, const double *g, double *gf, void *s) {
  return sisl::tricubic_adj<double>(a, b, c, co, perm, start, g, gf, s);
}
int sisl_tricubic_adj_f32(int32_t a, int32_t b, int32_t c, const float *co, const int32_t *perm,
                          const int32_t *start, const float *g, float *gf, void *s) {
  return sisl::tricubic_adj<float>(a, b, c, co, perm, start, g, gf, s);
}
int sisl_departure_f64(const SislGrid *g, int32_t loc, const double *u, const double *v,
                       const double *e, const double *m2, double dx, double dy, double dt,
                       int32_t it, double *scr, double *out, void *s) {
  return sisl::departure<double>(g, loc, u, v, e, m2, dx, dy, dt, it, scr, out, s);
}
int sisl_departure_f32(const SislGrid *g, int32_t loc, const float *u, const float *v,
                       const float *e, const float *m2, double dx, double dy, double dt,
                       int32_t it, float *scr, float *out, void *s) {
  return sisl::departure<float>(g, loc, u, v, e, m2, dx, dy, dt, it, scr, out, s);
}

int sisl_tricubic_adj_frac_f64(int32_t a, int32_t b, int32_t c, const double *co, const int32_t *perm,
                               double *frac, void *s) {
  return sisl::tricubic_adj_frac<double>(a, b, c, co, perm, frac, s);
}
int sisl_tricubic_adj_frac_f32(int32_t a, int32_t b, int32_t c, const float *co, const int32_t *perm,
                               float *frac, void *s) {
  return sisl::tricubic_adj_frac<float>(a, b, c, co, perm, frac, s);
}
int sisl_tricubic_adj_rows_f64(int32_t a, int32_t b, int32_t c, const int32_t *perm, const int32_t *start,
                               const double *frac, const double *g, double *g_sorted, double *gf, void *s) {
  return sisl::tricubic_adj_rows<double>(a, b, c, perm, start, frac, g, g_sorted, gf, s);
}
int sisl_tricubic_adj_rows_f32(int32_t a, int32_t b, int32_t c, const int32_t *perm, const int32_t *start,
                               const float *frac, const float *g, float *g_sorted, float *gf, void *s) {
  return sisl::tricubic_adj_rows<float>(a, b, c, perm, start, frac, g, g_sorted, gf, s);
}

}  // extern "C"
