// Shared device helpers for the SISL kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "suetes_sisl.h"

#define SISL_CHECK_LAUNCH()                                        \
  do {                                                             \
    hipError_t e_ = hipGetLastError();                             \
    if (e_ != hipSuccess) return SISL_ERR_LAUNCH;                  \
  } while (0)

namespace sisl {

constexpr int kBlock = 256;
constexpr int kMaxMembers = 32;  // members of one batched (ensemble) launch

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t b = (n + block - 1) / block;
  return (unsigned)(b < 1 ? 1 : b);
}

// Field extents in the internal (x-fastest) layout.
struct Dims {
  int nx, ny, nz;
  __host__ __device__ int64_t n() const { return (int64_t)nx * ny * nz; }
  __host__ __device__ int64_t at(int i, int j, int k) const {
    return ((int64_t)k * ny + j) * nx + i;
  }
};

// Staggered extents of the C grid.
struct Stag {
  Dims u, v, w, m;
  int64_t ou, ov, ow, om, oe, n;  // offsets of the packed Krylov vector
  int nx, ny, nz, px, py;
  __host__ __device__ static Stag make(const SislGrid &g) {
    Stag s;
    s.nx = g.nx; s.ny = g.ny; s.nz = g.nz; s.px = g.periodic_x; s.py = g.periodic_y;
    s.u = Dims{g.nx + 1, g.ny, g.nz};
    s.v = Dims{g.nx, g.ny + 1, g.nz};
    s.w = Dims{g.nx, g.ny, g.nz + 1};
    s.m = Dims{g.nx, g.ny, g.nz};
    s.ou = 0;
    s.ov = s.ou + s.u.n();
    s.ow = s.ov + s.v.n();
    s.om = s.ow + s.w.n();
    s.oe = s.om + s.m.n();
    s.n = s.oe + s.w.n();
    return s;
  }
  __host__ __device__ int64_t u2(int i, int j) const { return (int64_t)j * (nx + 1) + i; }
  __host__ __device__ int64_t v2(int i, int j) const { return (int64_t)j * nx + i; }
  __host__ __device__ int64_t m2(int i, int j) const { return (int64_t)j * nx + i; }
};

template <typename T>
struct OpT {
  Stag s;
  T alpha;
  const T *cua, *zxu, *cx, *cva, *zev, *cy, *dw, *cw, *cz, *kx, *ky, *rdzw, *cpm;
  const T *mudx, *mvdy, *afu, *afv, *mmx, *mmy;
  const T *idw, *gw, *lpi, *lower, *invden, *cprime, *dzw0, *dzwn;
  static OpT make(const SislOperator &o) {
    OpT r;
    r.s = Stag::make(o.g);
    r.alpha = (T)o.alpha;
#define SISL_P(name) r.name = static_cast<const T *>(o.name)
    SISL_P(cua); SISL_P(zxu); SISL_P(cx); SISL_P(cva); SISL_P(zev); SISL_P(cy);
    SISL_P(dw); SISL_P(cw); SISL_P(cz); SISL_P(kx); SISL_P(ky); SISL_P(rdzw); SISL_P(cpm);
    SISL_P(mudx); SISL_P(mvdy); SISL_P(afu); SISL_P(afv); SISL_P(mmx); SISL_P(mmy);
    SISL_P(idw); SISL_P(gw); SISL_P(lpi); SISL_P(lower); SISL_P(invden); SISL_P(cprime);
    SISL_P(dzw0); SISL_P(dzwn);
#undef SISL_P
    return r;
  }
  bool complete() const {
    const void *p[] = {cua, zxu, cx, cva, zev, cy, dw, cw, cz, kx, ky, rdzw, cpm, mudx, mvdy,
                       afu, afv, mmx, mmy, idw, gw, lpi, lower, invden, cprime, dzw0, dzwn};
    for (auto q : p)
      if (!q) return false;
    return s.nx >= 2 && s.ny >= 1 && s.nz >= 2;
  }
};

// The hardware hands consecutive workgroups of a launch to the 8 XCDs in
// turn.  Renumber them so that each XCD receives one contiguous range of
// logical tiles: neighbouring tiles then share an L2 and their halos hit it.
constexpr int kXcds = 8;
__device__ inline int xcd_tile(int b, int nblocks) {
  if (nblocks % kXcds) return b;
  return (b % kXcds) * (nblocks / kXcds) + b / kXcds;
}

// Load/store through a wave-uniform base pointer and a 32-bit element index.
// The byte offset is formed in 32 bits so that the access compiles to the
// scalar-base + 32-bit-vector-offset form (no 64-bit address per array in
// vector registers).  Callers guarantee every array stays below 4 GiB.
template <typename T>
__device__ __forceinline__ T ldg(const T *base, uint32_t idx) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + idx * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void stg(T *base, uint32_t idx, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(T)) = v;
}

// Coefficient storage: the working type T, or bfloat16 bits (uint16_t) for
// the inner solver of the mixed path; loads convert to T.
using bf16_t = uint16_t;
template <typename T, typename TC>
__device__ __forceinline__ T ldc(const TC *base, uint32_t idx) {
  if constexpr (sizeof(TC) == 2) {
    return (T)__uint_as_float((uint32_t)ldg(base, idx) << 16);
  } else {
    return (T)ldg(base, idx);
  }
}

// Wave64 sum reduction through DPP/shuffles.
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace sisl

namespace sisl {
// One member's operator application in a batched launch (device table entry):
// input x + j * xj, output y, transpose scratch scr (2 mass arrays), skipped
// when stop is set and stop[j * sj] != 0.  Operator launches with the fused
// basis normalisation (FP32 operator, FP16 basis) also read the scale *inv
// and store the basis column vout + j * vld (vs: the basis storage scale).
template <typename T>
struct BOp {
  OpT<T> o;
  const T *x;
  int64_t xj;
  T *y;
  T *scr;
  const int *stop;
  int sj;
  const double *inv;
  void *vout;
  int64_t vld;
  double vs;
};
// Batched launchers: member m = blockIdx.y over M members; `tab` in device
// memory; `maxb` = the largest per-member block count (host-computed);
// nb = 1: the operator launch does the fused basis normalisation.
template <typename T>
int launch_A_batched(const BOp<T> *tab, int M, int maxb, int j, hipStream_t s, int nb = 0);
template <typename T>
int launch_AT_batched(const BOp<T> *tab, int M, int maxb, int j, hipStream_t s, int nb = 0);
template <typename T>
int launch_M_batched(const BOp<T> *tab, int M, int maxb, int L, int S, int j, hipStream_t s);
template <typename T>
int launch_MT_batched(const BOp<T> *tab, int M, int maxb, int L, int S, int maxf, int j, hipStream_t s);
// Host-side launch geometry of one member (for maxb).
int blocks_A(const Stag &s);
int blocks_M(const Stag &s);
int blocks_MT_faces(const Stag &s);
int seg_levels(int nz);  // levels per segment L of the preconditioner (0: too tall)
// Generation of the operator launch shapes (advances when a shape is tuned or
// pinned): a captured launch sequence is valid for one generation.
uint64_t shape_generation();

// Optional launch timing (timing.hip): returns a record id or -1 when off.
bool timer_on();
int timer_begin(int kind, bool fp32, double bytes, const int *stop, const int *gate, hipStream_t s);
void timer_end(int id, hipStream_t s);

// Launchers shared between translation units (explicitly instantiated for
// float and double next to their kernels).
// `stop` (device int, optional): the kernels return at once when *stop != 0
// (used to retire the remaining Arnoldi steps of a converged cycle).
// Fused basis normalisation of an operator launch (FP32 operator only): the
// input is the unnormalised Arnoldi vector, the outputs are scaled by *inv
// and the launch stores the basis column vs * x * (*inv) into vout
// (nb = 1: FP16, nb = 2: FP32).  nb = 4 (FP32 operator): the input pointer
// holds FP16 values (a stored basis column), the outputs are scaled by
// *inv and nothing else is stored.  nb = 0: a plain apply.
struct Fuse {
  const double *inv = nullptr;
  void *vout = nullptr;
  int nb = 0;
  double vs = 1.0;
  int cb = 0;  // 1: the 3-D coefficient arrays are bfloat16 (FP32 operator only)
  const void *res = nullptr;  // packed vector b: store b - A x (nb 0; nb 5: in FP32 into vout)
};
template <typename T>
int launch_A(const OpT<T> &op, const T *x, T *y, hipStream_t s, const int *stop = nullptr,
             const Fuse *fuse = nullptr);
template <typename T>
int launch_AT(const OpT<T> &op, const T *x, T *y, T *scratch, hipStream_t s, const int *stop = nullptr,
              const Fuse *fuse = nullptr);
// cb = 1: the 3-D coefficient arrays of `op` are bfloat16 (FP32 operator only)
template <typename T>
int launch_M(const OpT<T> &op, const T *r, T *z, hipStream_t s, const int *stop = nullptr, int cb = 0);
template <typename T>
int launch_MT(const OpT<T> &op, const T *r, T *z, T *scratch, hipStream_t s, const int *stop = nullptr,
              int cb = 0);
}  // namespace sisl
