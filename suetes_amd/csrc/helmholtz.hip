// (b) Matrix-free operator A of the semi-implicit solve and its transpose.
//
// A is Euler3D.linear_operator (suetes/regional3d/euler.py:374-439) seen
// through the solver's scaling (steppers.py:439-461): packed Krylov vectors
// u | v | w | pi*1000 | eta_dot*dz_w.  Every coefficient that the reference
// recomputes per call (background averages, metrics, map factors) is
// precomputed once per stepper by the host layer.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.h"

namespace sisl {

// Launch geometry (operator and transpose): one thread per (i, j) column of a
// 64 x 4 tile, marching a chunk of `kchunk` levels upwards and carrying the
// vertical neighbours in registers; it writes all five packed outputs of its
// column, so every input array is read once per launch (plus tile halos, which
// the XCD-aware tile order keeps in L2).
constexpr int kTX = 64, kTY = 4;

// Blocks of the forward operator at the default chunk (8 level chunks).
__device__ __forceinline__ int blocks_A_dev(const Stag &s) {
  const int kc = (s.nz + 1 + 7) / 8;
  return ((s.nx + kTX - 1) / kTX) * ((s.ny + kTY - 1) / kTY) * ((s.nz + 1 + kc - 1) / kc);
}

// Fused basis normalisation of the inner (FP32) Krylov solver.  With NB = 1, 2
// the input is the unnormalised Arnoldi vector w: the outputs are scaled by
// 1 / ||w|| (A is linear) and every thread also stores the elements of the
// normalised basis column vs * w / ||w|| that it owns -- the input elements at
// its own output positions (NB = 1: FP16 storage, NB = 2: the working type).
// This replaces a separate normalisation pass over w.
//
// NB = 3 (FP64 only) is the residual epilogue of the FP64 refinement instead:
// every output is stored as res - (A x), res the packed right-hand side.
// NB = 5 (FP64 only): the same residual, formed in FP64 and stored rounded to
// FP32 into vout (float, packed layout) instead of y -- the input of the
// FP32 preconditioner that forms the next inner cycle's start vector.
template <typename T, int NB> struct BasisT { using type = T; };
template <typename T> struct BasisT<T, 1> { using type = _Float16; };
// NB = 4 (FP32 only): the input is a basis column stored in FP16 (read and
// converted to T), the outputs are scaled by *inv, and nothing else is
// stored -- the operator launch of an Arnoldi step whose vector the previous
// step's update pass already wrote to the basis.
template <typename T, int NB> struct InT { using type = T; };
template <typename T> struct InT<T, 4> { using type = _Float16; };
template <typename T, typename TI>
__device__ __forceinline__ T ldx(const TI *base, uint32_t idx) {
  return (T)ldg(base, idx);
}
#define FUSED_NORMALISATION_SETUP                                                   \
  using BT = typename BasisT<T, NB>::type;                                          \
  constexpr bool kNorm = NB == 1 || NB == 2;                                        \
  constexpr bool kScale = kNorm || NB == 4;                                         \
  T osc = 1, bsc = 1;                                                               \
  BT *BU = nullptr, *BV = nullptr, *BW = nullptr, *BP = nullptr, *BE = nullptr;     \
  if constexpr (kScale) {                                                           \
    const double iv = *inv;                                                         \
    osc = (T)iv;                                                                    \
    bsc = (T)(iv * vs);                                                             \
  }                                                                                 \
  if constexpr (kNorm) {                                                            \
    BT *b_ = static_cast<BT *>(vout);                                               \
    BU = b_ + s.ou; BV = b_ + s.ov; BW = b_ + s.ow; BP = b_ + s.om; BE = b_ + s.oe; \
  }                                                                                 \
  auto so = [&](T v) -> T {                                                         \
    if constexpr (kScale) return v * osc;                                           \
    else return v;                                                                  \
  };                                                                                \
  auto bput = [&](BT *b, uint32_t q, T v) {                                         \
    if constexpr (kNorm) stg(b, q, (BT)(v * bsc));                                  \
  };

// Body of the forward operator for launch block `blk` of `nblk` (a member of
// a batched launch passes its own block index and count).
template <typename T, int UNR, int NB = 0, typename TC = T>
__device__ __forceinline__ void helmholtz_body(const OpT<T> &o, const T *__restrict__ x,
                                               T *__restrict__ y, int kchunk, int blk, int nblk,
                                               const double *inv = nullptr, void *vout = nullptr,
                                               double vs = 1.0, const T *__restrict__ res = nullptr) {
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int nbx = (nx + kTX - 1) / kTX, nby = (ny + kTY - 1) / kTY;
  const int tile = xcd_tile(blk, nblk);
  const int bx = tile % nbx, by = (tile / nbx) % nby, bz = tile / (nbx * nby);
  const int i = bx * kTX + threadIdx.x, j = by * kTY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int k0 = bz * kchunk, k1 = min(k0 + kchunk, nz + 1);
  const int il = i == 0 ? nx - 1 : i - 1, jd = j == 0 ? ny - 1 : j - 1;
  using TI = typename InT<T, NB>::type;  // input element type
  const TI *xi = reinterpret_cast<const TI *>(x);
  const TI *U = xi + s.ou, *V = xi + s.ov, *W = xi + s.ow, *P = xi + s.om, *E = xi + s.oe;
  T *YU = y + s.ou, *YV = y + s.ov, *YW = y + s.ow, *YP = y + s.om, *YE = y + s.oe;
  FUSED_NORMALISATION_SETUP
  // output store (NB = 3: the residual res - A x)
  auto put = [&](T *Y_, uint32_t q, T v) {
    if constexpr (NB == 3 || NB == 5) v = ldg(res + (Y_ - y), q) - v;
    if constexpr (NB == 5) stg(static_cast<float *>(vout) + (Y_ - y), q, (float)v);
    else stg(Y_, q, so(v));
  };
  const bool uform = s.px || i > 0, vform = s.py || j > 0;  // else face 0 is the identity
  const T mudx = ldg(o.mudx, (uint32_t)s.u2(i, j)), afu = ldg(o.afu, (uint32_t)s.u2(i, j));
  const T mvdy = ldg(o.mvdy, (uint32_t)s.v2(i, j)), afv = ldg(o.afv, (uint32_t)s.v2(i, j));
  const T mmx = ldg(o.mmx, (uint32_t)s.m2(i, j)), mmy = ldg(o.mmy, (uint32_t)s.m2(i, j));
  const T half = (T)0.5, alpha = o.alpha;
  // coefficient arrays as restrict locals: loads of level k+1 may move above stores of level k
  const TC *__restrict__ c_cua = reinterpret_cast<const TC *>(o.cua);
  const TC *__restrict__ c_zxu = reinterpret_cast<const TC *>(o.zxu);
  const TC *__restrict__ c_cx = reinterpret_cast<const TC *>(o.cx);
  const TC *__restrict__ c_cva = reinterpret_cast<const TC *>(o.cva);
  const TC *__restrict__ c_zev = reinterpret_cast<const TC *>(o.zev);
  const TC *__restrict__ c_cy = reinterpret_cast<const TC *>(o.cy);
  const TC *__restrict__ c_dw = reinterpret_cast<const TC *>(o.dw);
  const TC *__restrict__ c_cw = reinterpret_cast<const TC *>(o.cw);
  const TC *__restrict__ c_cz = reinterpret_cast<const TC *>(o.cz);
  const TC *__restrict__ c_kx = reinterpret_cast<const TC *>(o.kx);
  const TC *__restrict__ c_ky = reinterpret_cast<const TC *>(o.ky);
  const TC *__restrict__ c_rdzw = reinterpret_cast<const TC *>(o.rdzw);
  const TC *__restrict__ c_cpm = reinterpret_cast<const TC *>(o.cpm);
  // 32-bit element offsets at level 0 and the plane strides: own column
  // (i, j), left column (il, j), lower column (i, jd).  Mass and w points
  // share the (i, j) plane layout.
  const uint32_t pm = (uint32_t)nx * ny, pu = (uint32_t)(nx + 1) * ny, pv = (uint32_t)nx * (ny + 1);
  const uint32_t mc = (uint32_t)j * nx + i, ml = (uint32_t)j * nx + il, md = (uint32_t)jd * nx + i;
  const uint32_t uc = (uint32_t)j * (nx + 1) + i, ud = (uint32_t)jd * (nx + 1) + i;
  const uint32_t vc = (uint32_t)j * nx + i, vl = (uint32_t)j * nx + il;
  const uint32_t ux = (uint32_t)j * (nx + 1) + nx;  // east face of row j
  const uint32_t vy = (uint32_t)ny * nx + i;         // north face of column i
  // carried state at level k: P at k and k-1, gz_w at k (three columns),
  // u_m / v_m of the own column at k-1, eta_dot and its flux at w level k
  T pc0 = 0, pl0 = 0, pd0 = 0, pcm = 0, plm = 0, pdm = 0, gc0 = 0, gl0 = 0, gd0 = 0;
  T umc_p = 0, vmc_p = 0;
  if (k0 < nz) {
    const uint32_t q = (uint32_t)k0 * pm;
    pc0 = ldx<T>(P, mc + q); pl0 = ldx<T>(P, ml + q); pd0 = ldx<T>(P, md + q);
  }
  if (k0 >= 1) {
    const uint32_t q = (uint32_t)(k0 - 1) * pm;
    pcm = ldx<T>(P, mc + q); plm = ldx<T>(P, ml + q); pdm = ldx<T>(P, md + q);
    const uint32_t qu = (uint32_t)(k0 - 1) * pu, qv = (uint32_t)(k0 - 1) * pv;
    umc_p = half * (ldx<T>(U, uc + qu) + ldx<T>(U, uc + qu + 1));
    vmc_p = half * (ldx<T>(V, vc + qv) + ldx<T>(V, vc + qv + nx));
  }
  if (k0 == 0) {
    gc0 = (ldx<T>(P, mc + pm) - pc0) * ldc<T>(c_rdzw, mc);
    gl0 = (ldx<T>(P, ml + pm) - pl0) * ldc<T>(c_rdzw, ml);
    gd0 = (ldx<T>(P, md + pm) - pd0) * ldc<T>(c_rdzw, md);
  } else if (k0 < nz) {
    const uint32_t q = (uint32_t)k0 * pm;
    gc0 = (pc0 - pcm) * ldc<T>(c_rdzw, mc + q);
    gl0 = (pl0 - plm) * ldc<T>(c_rdzw, ml + q);
    gd0 = (pd0 - pdm) * ldc<T>(c_rdzw, md + q);
  }
  T e0 = ldx<T>(E, mc + (uint32_t)k0 * pm);
  T fz0 = (k0 >= 1 && k0 < nz) ? e0 * ldc<T>(c_cz, mc + (uint32_t)k0 * pm) : (T)0;
#pragma unroll UNR
  for (int k = k0; k < k1; ++k) {
    const uint32_t qm = (uint32_t)k * pm, qu = (uint32_t)k * pu, qv = (uint32_t)k * pv;
    const uint32_t wk = mc + qm;  // own column at w level k (same plane layout)
    const T wv = ldx<T>(W, wk);
    if (k == nz) {  // top w level: both rows are identities
      put(YW, wk, wv);
      put(YE, wk, e0 * ldc<T>(c_rdzw, wk));
      bput(BW, wk, wv);
      bput(BE, wk, e0);
      break;
    }
    // ---- mass level k: gz_m at the three columns ----
    T gc1, gl1, gd1, pc1 = 0, pl1 = 0, pd1 = 0;
    if (k + 1 < nz) {
      const uint32_t q1 = qm + pm;
      pc1 = ldx<T>(P, mc + q1); pl1 = ldx<T>(P, ml + q1); pd1 = ldx<T>(P, md + q1);
      gc1 = (pc1 - pc0) * ldc<T>(c_rdzw, mc + q1);
      gl1 = (pl1 - pl0) * ldc<T>(c_rdzw, ml + q1);
      gd1 = (pd1 - pd0) * ldc<T>(c_rdzw, md + q1);
    } else {  // top closure: gz_w(nz) repeats the last interior difference
      const uint32_t q1 = qm + pm;
      gc1 = (pc0 - pcm) * ldc<T>(c_rdzw, mc + q1);
      gl1 = (pl0 - plm) * ldc<T>(c_rdzw, ml + q1);
      gd1 = (pd0 - pdm) * ldc<T>(c_rdzw, md + q1);
    }
    const T gzc = half * (gc0 + gc1), gzl = half * (gl0 + gl1), gzd = half * (gd0 + gd1);
    const uint32_t au = uc + qu, av = vc + qv;
    const T u0 = ldx<T>(U, au), u1 = ldx<T>(U, au + 1), v0 = ldx<T>(V, av), v1 = ldx<T>(V, av + nx);
    const T um_k = half * (u0 + u1), vm_k = half * (v0 + v1);
    // ---- L_u at face i (euler.py:398, 411-415) ----
    {
      T out = u0;
      if (uform) {
        const uint32_t al = vl + qv;
        const T vml = half * (ldx<T>(V, al) + ldx<T>(V, al + nx));
        out = u0 + ldc<T>(c_cua, au) * (mudx * (pc0 - pl0) - ldc<T>(c_zxu, au) * (half * (gzl + gzc))) -
              afu * (half * (vml + vm_k));
      }
      put(YU, au, out);
      bput(BU, au, u0);
      if (i == nx - 1 && !s.px) {                      // closed east face: identity
        put(YU, au + 1, u1);
        bput(BU, au + 1, u1);
      }
      if (i == 0 && s.px) {                            // periodic: face nx repeats face 0
        put(YU, ux + qu, out);
        if constexpr (kNorm) bput(BU, ux + qu, ldx<T>(U, ux + qu));
      }
    }
    // ---- L_v at face j (euler.py:399, 416-420) ----
    {
      T out = v0;
      if (vform) {
        const uint32_t ad = ud + qu;
        const T umd = half * (ldx<T>(U, ad) + ldx<T>(U, ad + 1));
        out = v0 + ldc<T>(c_cva, av) * (mvdy * (pc0 - pd0) - ldc<T>(c_zev, av) * (half * (gzd + gzc))) +
              afv * (half * (umd + um_k));
      }
      put(YV, av, out);
      bput(BV, av, v0);
      if (j == ny - 1 && !s.py) {
        put(YV, av + nx, v1);
        bput(BV, av + nx, v1);
      }
      if (j == 0 && s.py) {
        put(YV, vy + qv, out);
        if constexpr (kNorm) bput(BV, vy + qv, ldx<T>(V, vy + qv));
      }
    }
    // ---- L_pi (euler.py:298-318, 401) ----
    const uint32_t w1 = wk + pm;
    const T e1 = ldx<T>(E, w1);
    const T fz1 = (k + 1 < nz) ? e1 * ldc<T>(c_cz, w1) : (T)0;
    {
      const T fx = u1 * ldc<T>(c_cx, au + 1) - u0 * ldc<T>(c_cx, au);
      const T fy = v1 * ldc<T>(c_cy, av + nx) - v0 * ldc<T>(c_cy, av);
      put(YP, wk, pc0 + ldc<T>(c_cpm, wk) * (mmx * fx + mmy * fy + (fz1 - fz0)));
      bput(BP, wk, pc0);
    }
    // ---- w level k: L_w (euler.py:400, 426-427), L_eta (euler.py:429-437) ----
    if (k == 0) {
      put(YW, wk, wv);
      put(YE, wk, e0 * ldc<T>(c_rdzw, wk));
    } else {
      put(YW, wk, ldc<T>(c_dw, wk) * wv + ldc<T>(c_cw, wk) * (pc0 - pcm));
      put(YE, wk, alpha * (e0 + ldc<T>(c_kx, wk) * (half * (umc_p + um_k)) +
                              ldc<T>(c_ky, wk) * (half * (vmc_p + vm_k)) - wv));
    }
    bput(BW, wk, wv);
    bput(BE, wk, e0);
    // k+1 becomes k
    gc0 = gc1; gl0 = gl1; gd0 = gd1;
    pcm = pc0; plm = pl0; pdm = pd0;
    pc0 = pc1; pl0 = pl1; pd0 = pd1;
    umc_p = um_k; vmc_p = vm_k;
    e0 = e1; fz0 = fz1;
  }
}

template <typename T, int UNR, int NB, typename TC>
__global__ void __launch_bounds__(256) helmholtz_kernel(OpT<T> o, const T *__restrict__ x,
                                                        T *__restrict__ y, int kchunk,
                                                        const int *__restrict__ stop,
                                                        const double *__restrict__ inv, void *vout,
                                                        double vs, const T *__restrict__ res) {
  if (stop && *stop) return;
  helmholtz_body<T, UNR, NB, TC>(o, x, y, kchunk, blockIdx.x, gridDim.x, inv, vout, vs, res);
}


// ---- two columns per thread ------------------------------------------------------
// The forward operator with each thread owning the adjacent columns (i, i+1),
// i even (nx even): column i+1 takes every left-neighbour value (pi and its
// vertical difference, the v-face average, the u face between the two) from
// column i's registers, and the pair's loads of mass-point, w-level and
// v-face arrays are single two-element vector loads (u-face rows have an odd
// stride, so those stay scalar).  Each output is the same expression of the
// same operands as in helmholtz_body, so the results are the same bits.
template <typename T> struct Vec2T;
template <> struct Vec2T<float> { using type = float2; };
template <> struct Vec2T<double> { using type = double2; };
template <typename T>
__device__ __forceinline__ void ldg2(const T *base, uint32_t idx, T &a, T &b) {
  using V = typename Vec2T<T>::type;
  const V v = *reinterpret_cast<const V *>(reinterpret_cast<const char *>(base) + idx * (uint32_t)sizeof(T));
  a = v.x;
  b = v.y;
}
// a pair of input elements, converted to T (FP16 input: one 32-bit load)
template <typename T, typename TI>
__device__ __forceinline__ void ldx2(const TI *base, uint32_t idx, T &a, T &b) {
  if constexpr (sizeof(TI) == 2) {
    typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
    const h2_t v = *reinterpret_cast<const h2_t *>(reinterpret_cast<const char *>(base) + idx * 2u);
    a = (T)v.x;
    b = (T)v.y;
  } else {
    ldg2<T>(base, idx, a, b);
  }
}
template <typename T>
__device__ __forceinline__ void stg2(T *base, uint32_t idx, T a, T b) {
  using V = typename Vec2T<T>::type;
  V v;
  v.x = a;
  v.y = b;
  *reinterpret_cast<V *>(reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(T)) = v;
}
// a pair of coefficients (bfloat16 storage: one 32-bit load, low half first)
template <typename T, typename TC>
__device__ __forceinline__ void ldc2(const TC *base, uint32_t idx, T &a, T &b) {
  if constexpr (sizeof(TC) == 2) {
    const uint32_t v = *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(base) + idx * 2u);
    a = (T)__uint_as_float(v << 16);
    b = (T)__uint_as_float(v & 0xffff0000u);
  } else {
    ldg2<T>(reinterpret_cast<const T *>(base), idx, a, b);
  }
}
// a pair of basis elements (FP16 or the working type).  FP16 elements are
// stored one by one: packing the two conversions into one register lets the
// compiler fuse the scaling multiply and the conversion into mixed-precision
// instructions that round once, not as the one-column kernel does (a product
// rounded to FP32, then to FP16), and a few elements per column then differ
// in their last bit.
template <typename B>
__device__ __forceinline__ void stb2(B *base, uint32_t idx, B a, B b) {
  if constexpr (sizeof(B) == 2) {
    stg(base, idx, a);
    stg(base, idx + 1, b);
  } else {
    stg2<B>(base, idx, a, b);
  }
}

constexpr int kTX2 = 2 * kTX;  // columns per tile row of the two-column form

template <typename T, int NB = 0, typename TC = T>
__device__ __forceinline__ void helmholtz2_body(const OpT<T> &o, const T *__restrict__ x,
                                                T *__restrict__ y, int kchunk, int blk, int nblk,
                                                const double *inv = nullptr, void *vout = nullptr,
                                                double vs = 1.0, const T *__restrict__ res = nullptr) {
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int nbx = (nx + kTX2 - 1) / kTX2, nby = (ny + kTY - 1) / kTY;
  const int tile = xcd_tile(blk, nblk);
  const int bx = tile % nbx, by = (tile / nbx) % nby, bz = tile / (nbx * nby);
  const int i = bx * kTX2 + 2 * threadIdx.x, j = by * kTY + threadIdx.y;  // columns i, i + 1
  if (i >= nx || j >= ny) return;
  const int k0 = bz * kchunk, k1 = min(k0 + kchunk, nz + 1);
  const int il = i == 0 ? nx - 1 : i - 1, jd = j == 0 ? ny - 1 : j - 1;
  using TI = typename InT<T, NB>::type;  // input element type
  const TI *xi = reinterpret_cast<const TI *>(x);
  const TI *U = xi + s.ou, *V = xi + s.ov, *W = xi + s.ow, *P = xi + s.om, *E = xi + s.oe;
  T *YU = y + s.ou, *YV = y + s.ov, *YW = y + s.ow, *YP = y + s.om, *YE = y + s.oe;
  FUSED_NORMALISATION_SETUP
  auto put = [&](T *Y_, uint32_t q, T v) {
    if constexpr (NB == 3 || NB == 5) v = ldg(res + (Y_ - y), q) - v;
    if constexpr (NB == 5) stg(static_cast<float *>(vout) + (Y_ - y), q, (float)v);
    else stg(Y_, q, so(v));
  };
  auto put2 = [&](T *Y_, uint32_t q, T va, T vb) {
    if constexpr (NB == 3 || NB == 5) {
      T ra, rb;
      ldg2(res + (Y_ - y), q, ra, rb);
      va = ra - va;
      vb = rb - vb;
    }
    if constexpr (NB == 5) stg2<float>(static_cast<float *>(vout) + (Y_ - y), q, (float)va, (float)vb);
    else stg2(Y_, q, so(va), so(vb));
  };
  auto bput2 = [&](BT *b, uint32_t q, T va, T vb) {
    if constexpr (kNorm) stb2<BT>(b, q, (BT)(va * bsc), (BT)(vb * bsc));
  };
  const bool uform0 = s.px || i > 0, vform = s.py || j > 0;  // column i + 1 > 0 always has its formula
  const T mudx0 = ldg(o.mudx, (uint32_t)s.u2(i, j)), mudx1 = ldg(o.mudx, (uint32_t)s.u2(i + 1, j));
  const T afu0 = ldg(o.afu, (uint32_t)s.u2(i, j)), afu1 = ldg(o.afu, (uint32_t)s.u2(i + 1, j));
  T mvdy[2], afv[2], mmx[2], mmy[2];
  ldg2(o.mvdy, (uint32_t)s.v2(i, j), mvdy[0], mvdy[1]);
  ldg2(o.afv, (uint32_t)s.v2(i, j), afv[0], afv[1]);
  ldg2(o.mmx, (uint32_t)s.m2(i, j), mmx[0], mmx[1]);
  ldg2(o.mmy, (uint32_t)s.m2(i, j), mmy[0], mmy[1]);
  const T half = (T)0.5, alpha = o.alpha;
  const TC *__restrict__ c_cua = reinterpret_cast<const TC *>(o.cua);
  const TC *__restrict__ c_zxu = reinterpret_cast<const TC *>(o.zxu);
  const TC *__restrict__ c_cx = reinterpret_cast<const TC *>(o.cx);
  const TC *__restrict__ c_cva = reinterpret_cast<const TC *>(o.cva);
  const TC *__restrict__ c_zev = reinterpret_cast<const TC *>(o.zev);
  const TC *__restrict__ c_cy = reinterpret_cast<const TC *>(o.cy);
  const TC *__restrict__ c_dw = reinterpret_cast<const TC *>(o.dw);
  const TC *__restrict__ c_cw = reinterpret_cast<const TC *>(o.cw);
  const TC *__restrict__ c_cz = reinterpret_cast<const TC *>(o.cz);
  const TC *__restrict__ c_kx = reinterpret_cast<const TC *>(o.kx);
  const TC *__restrict__ c_ky = reinterpret_cast<const TC *>(o.ky);
  const TC *__restrict__ c_rdzw = reinterpret_cast<const TC *>(o.rdzw);
  const TC *__restrict__ c_cpm = reinterpret_cast<const TC *>(o.cpm);
  const uint32_t pm = (uint32_t)nx * ny, pu = (uint32_t)(nx + 1) * ny, pv = (uint32_t)nx * (ny + 1);
  const uint32_t mc = (uint32_t)j * nx + i, ml = (uint32_t)j * nx + il, md = (uint32_t)jd * nx + i;
  const uint32_t uc = (uint32_t)j * (nx + 1) + i, ud = (uint32_t)jd * (nx + 1) + i;
  const uint32_t vc = (uint32_t)j * nx + i, vl = (uint32_t)j * nx + il;
  const uint32_t ux = (uint32_t)j * (nx + 1) + nx;
  const uint32_t vy = (uint32_t)ny * nx + i;
  // carried state per column (index 0: column i, 1: column i + 1); the left
  // column of column i is carried separately (pl*, gl0), that of i + 1 is column i
  T pc0[2] = {0, 0}, pd0[2] = {0, 0}, pcm[2] = {0, 0}, pdm[2] = {0, 0}, gc0[2] = {0, 0}, gd0[2] = {0, 0};
  T pl0 = 0, plm = 0, gl0 = 0;
  T umc_p[2] = {0, 0}, vmc_p[2] = {0, 0};
  if (k0 < nz) {
    const uint32_t q = (uint32_t)k0 * pm;
    ldx2<T>(P, mc + q, pc0[0], pc0[1]);
    pl0 = ldx<T>(P, ml + q);
    ldx2<T>(P, md + q, pd0[0], pd0[1]);
  }
  if (k0 >= 1) {
    const uint32_t q = (uint32_t)(k0 - 1) * pm;
    ldx2<T>(P, mc + q, pcm[0], pcm[1]);
    plm = ldx<T>(P, ml + q);
    ldx2<T>(P, md + q, pdm[0], pdm[1]);
    const uint32_t qu = (uint32_t)(k0 - 1) * pu, qv = (uint32_t)(k0 - 1) * pv;
    const T ua = ldx<T>(U, uc + qu), ub = ldx<T>(U, uc + qu + 1), ue = ldx<T>(U, uc + qu + 2);
    umc_p[0] = half * (ua + ub);
    umc_p[1] = half * (ub + ue);
    T va[2], vb[2];
    ldx2<T>(V, vc + qv, va[0], va[1]);
    ldx2<T>(V, vc + qv + nx, vb[0], vb[1]);
    vmc_p[0] = half * (va[0] + vb[0]);
    vmc_p[1] = half * (va[1] + vb[1]);
  }
  if (k0 == 0) {
    T p1[2], d1[2], r[2], rd[2];
    ldx2<T>(P, mc + pm, p1[0], p1[1]);
    ldx2<T>(P, md + pm, d1[0], d1[1]);
    ldc2<T>(c_rdzw, mc, r[0], r[1]);
    ldc2<T>(c_rdzw, md, rd[0], rd[1]);
    for (int c = 0; c < 2; ++c) {
      gc0[c] = (p1[c] - pc0[c]) * r[c];
      gd0[c] = (d1[c] - pd0[c]) * rd[c];
    }
    gl0 = (ldx<T>(P, ml + pm) - pl0) * ldc<T>(c_rdzw, ml);
  } else if (k0 < nz) {
    const uint32_t q = (uint32_t)k0 * pm;
    T r[2], rd[2];
    ldc2<T>(c_rdzw, mc + q, r[0], r[1]);
    ldc2<T>(c_rdzw, md + q, rd[0], rd[1]);
    for (int c = 0; c < 2; ++c) {
      gc0[c] = (pc0[c] - pcm[c]) * r[c];
      gd0[c] = (pd0[c] - pdm[c]) * rd[c];
    }
    gl0 = (pl0 - plm) * ldc<T>(c_rdzw, ml + q);
  }
  T e0[2], fz0[2] = {0, 0};
  ldx2<T>(E, mc + (uint32_t)k0 * pm, e0[0], e0[1]);
  if (k0 >= 1 && k0 < nz) {
    T cz[2];
    ldc2<T>(c_cz, mc + (uint32_t)k0 * pm, cz[0], cz[1]);
    fz0[0] = e0[0] * cz[0];
    fz0[1] = e0[1] * cz[1];
  }
  for (int k = k0; k < k1; ++k) {
    const uint32_t qm = (uint32_t)k * pm, qu = (uint32_t)k * pu, qv = (uint32_t)k * pv;
    const uint32_t wk = mc + qm;
    T wv[2];
    ldx2<T>(W, wk, wv[0], wv[1]);
    if (k == nz) {
      T r[2];
      ldc2<T>(c_rdzw, wk, r[0], r[1]);
      put2(YW, wk, wv[0], wv[1]);
      put2(YE, wk, e0[0] * r[0], e0[1] * r[1]);
      bput2(BW, wk, wv[0], wv[1]);
      bput2(BE, wk, e0[0], e0[1]);
      break;
    }
    // ---- mass level k: gz_m of the columns ----
    T gc1[2], gd1[2], pc1[2] = {0, 0}, pd1[2] = {0, 0}, gl1, pl1 = 0;
    {
      const uint32_t q1 = qm + pm;
      T r[2], rd[2];
      ldc2<T>(c_rdzw, mc + q1, r[0], r[1]);
      ldc2<T>(c_rdzw, md + q1, rd[0], rd[1]);
      const T rl = ldc<T>(c_rdzw, ml + q1);
      if (k + 1 < nz) {
        ldx2<T>(P, mc + q1, pc1[0], pc1[1]);
        pl1 = ldx<T>(P, ml + q1);
        ldx2<T>(P, md + q1, pd1[0], pd1[1]);
        for (int c = 0; c < 2; ++c) {
          gc1[c] = (pc1[c] - pc0[c]) * r[c];
          gd1[c] = (pd1[c] - pd0[c]) * rd[c];
        }
        gl1 = (pl1 - pl0) * rl;
      } else {  // top closure: gz_w(nz) repeats the last interior difference
        for (int c = 0; c < 2; ++c) {
          gc1[c] = (pc0[c] - pcm[c]) * r[c];
          gd1[c] = (pd0[c] - pdm[c]) * rd[c];
        }
        gl1 = (pl0 - plm) * rl;
      }
    }
    const T gzc0 = half * (gc0[0] + gc1[0]), gzc1 = half * (gc0[1] + gc1[1]);
    const T gzl = half * (gl0 + gl1);
    const T gzd0 = half * (gd0[0] + gd1[0]), gzd1 = half * (gd0[1] + gd1[1]);
    const uint32_t au = uc + qu, av = vc + qv;
    const T ua = ldx<T>(U, au), ub = ldx<T>(U, au + 1), ue = ldx<T>(U, au + 2);  // faces i, i+1, i+2
    T v0[2], v1[2];
    ldx2<T>(V, av, v0[0], v0[1]);
    ldx2<T>(V, av + nx, v1[0], v1[1]);
    const T um0 = half * (ua + ub), um1 = half * (ub + ue);
    const T vm0 = half * (v0[0] + v1[0]), vm1 = half * (v0[1] + v1[1]);
    // ---- L_u at faces i and i+1 ----
    {
      T out0 = ua;
      if (uform0) {
        const uint32_t al = vl + qv;
        const T vml = half * (ldx<T>(V, al) + ldx<T>(V, al + nx));
        out0 = ua + ldc<T>(c_cua, au) * (mudx0 * (pc0[0] - pl0) - ldc<T>(c_zxu, au) * (half * (gzl + gzc0))) -
               afu0 * (half * (vml + vm0));
      }
      const T out1 = ub + ldc<T>(c_cua, au + 1) * (mudx1 * (pc0[1] - pc0[0]) -
                                                   ldc<T>(c_zxu, au + 1) * (half * (gzc0 + gzc1))) -
                     afu1 * (half * (vm0 + vm1));
      put(YU, au, out0);
      put(YU, au + 1, out1);
      if constexpr (kNorm) {
        stg(BU, au, (BT)(ua * bsc));
        stg(BU, au + 1, (BT)(ub * bsc));
      }
      if (i + 1 == nx - 1 && !s.px) {  // closed east face: identity
        put(YU, au + 2, ue);
        if constexpr (kNorm) stg(BU, au + 2, (BT)(ue * bsc));
      }
      if (i == 0 && s.px) {  // periodic: face nx repeats face 0
        put(YU, ux + qu, out0);
        if constexpr (kNorm) stg(BU, ux + qu, (BT)(ldx<T>(U, ux + qu) * bsc));
      }
    }
    // ---- L_v at face j of both columns ----
    {
      T out[2] = {v0[0], v0[1]};
      if (vform) {
        const uint32_t ad = ud + qu;
        const T da = ldx<T>(U, ad), db = ldx<T>(U, ad + 1), de = ldx<T>(U, ad + 2);
        const T umd[2] = {half * (da + db), half * (db + de)};
        const T umk[2] = {um0, um1};
        const T gzc[2] = {gzc0, gzc1}, gzd[2] = {gzd0, gzd1};
        T cva[2], zev[2];
        ldc2<T>(c_cva, av, cva[0], cva[1]);
        ldc2<T>(c_zev, av, zev[0], zev[1]);
        for (int c = 0; c < 2; ++c)
          out[c] = v0[c] + cva[c] * (mvdy[c] * (pc0[c] - pd0[c]) - zev[c] * (half * (gzd[c] + gzc[c]))) +
                   afv[c] * (half * (umd[c] + umk[c]));
      }
      put2(YV, av, out[0], out[1]);
      bput2(BV, av, v0[0], v0[1]);
      if (j == ny - 1 && !s.py) {
        put2(YV, av + nx, v1[0], v1[1]);
        bput2(BV, av + nx, v1[0], v1[1]);
      }
      if (j == 0 && s.py) {
        put2(YV, vy + qv, out[0], out[1]);
        if constexpr (kNorm) {
          T na, nb_;
          ldx2<T>(V, vy + qv, na, nb_);
          bput2(BV, vy + qv, na, nb_);
        }
      }
    }
    // ---- L_pi ----
    const uint32_t w1 = wk + pm;
    T e1[2], fz1[2] = {0, 0};
    ldx2<T>(E, w1, e1[0], e1[1]);
    if (k + 1 < nz) {
      T cz[2];
      ldc2<T>(c_cz, w1, cz[0], cz[1]);
      fz1[0] = e1[0] * cz[0];
      fz1[1] = e1[1] * cz[1];
    }
    {
      const T cxa = ldc<T>(c_cx, au), cxb = ldc<T>(c_cx, au + 1), cxe = ldc<T>(c_cx, au + 2);
      const T fx0 = ub * cxb - ua * cxa, fx1 = ue * cxe - ub * cxb;
      T cya[2], cyb[2], cpm[2];
      ldc2<T>(c_cy, av, cya[0], cya[1]);
      ldc2<T>(c_cy, av + nx, cyb[0], cyb[1]);
      ldc2<T>(c_cpm, wk, cpm[0], cpm[1]);
      const T fy0 = v1[0] * cyb[0] - v0[0] * cya[0], fy1 = v1[1] * cyb[1] - v0[1] * cya[1];
      put2(YP, wk, pc0[0] + cpm[0] * (mmx[0] * fx0 + mmy[0] * fy0 + (fz1[0] - fz0[0])),
           pc0[1] + cpm[1] * (mmx[1] * fx1 + mmy[1] * fy1 + (fz1[1] - fz0[1])));
      bput2(BP, wk, pc0[0], pc0[1]);
    }
    // ---- w level k ----
    if (k == 0) {
      T r[2];
      ldc2<T>(c_rdzw, wk, r[0], r[1]);
      put2(YW, wk, wv[0], wv[1]);
      put2(YE, wk, e0[0] * r[0], e0[1] * r[1]);
    } else {
      T dw[2], cw[2], kx[2], ky[2];
      ldc2<T>(c_dw, wk, dw[0], dw[1]);
      ldc2<T>(c_cw, wk, cw[0], cw[1]);
      ldc2<T>(c_kx, wk, kx[0], kx[1]);
      ldc2<T>(c_ky, wk, ky[0], ky[1]);
      const T umk[2] = {um0, um1}, vmk[2] = {vm0, vm1};
      T yw[2], ye[2];
      for (int c = 0; c < 2; ++c) {
        yw[c] = dw[c] * wv[c] + cw[c] * (pc0[c] - pcm[c]);
        ye[c] = alpha * (e0[c] + kx[c] * (half * (umc_p[c] + umk[c])) + ky[c] * (half * (vmc_p[c] + vmk[c])) - wv[c]);
      }
      put2(YW, wk, yw[0], yw[1]);
      put2(YE, wk, ye[0], ye[1]);
    }
    bput2(BW, wk, wv[0], wv[1]);
    bput2(BE, wk, e0[0], e0[1]);
    // k+1 becomes k
    for (int c = 0; c < 2; ++c) {
      gc0[c] = gc1[c];
      gd0[c] = gd1[c];
      pcm[c] = pc0[c];
      pdm[c] = pd0[c];
      pc0[c] = pc1[c];
      pd0[c] = pd1[c];
      e0[c] = e1[c];
      fz0[c] = fz1[c];
    }
    gl0 = gl1;
    plm = pl0;
    pl0 = pl1;
    umc_p[0] = um0;
    umc_p[1] = um1;
    vmc_p[0] = vm0;
    vmc_p[1] = vm1;
  }
}

template <typename T, int NB, typename TC>
__global__ void __launch_bounds__(256) helmholtz2_kernel(OpT<T> o, const T *__restrict__ x,
                                                         T *__restrict__ y, int kchunk,
                                                         const int *__restrict__ stop,
                                                         const double *__restrict__ inv, void *vout,
                                                         double vs, const T *__restrict__ res) {
  if (stop && *stop) return;
  helmholtz2_body<T, NB, TC>(o, x, y, kchunk, blockIdx.x, gridDim.x, inv, vout, vs, res);
}

// The two-column form needs even nx and even offsets of the packed vector's
// components (aligned pair loads and stores).
__host__ __device__ inline bool two_column_ok(const Stag &s) {
  return s.nx % 2 == 0 && s.ov % 2 == 0 && s.ow % 2 == 0 && s.om % 2 == 0 && s.oe % 2 == 0;
}

// ---- fused transpose -----------------------------------------------------------
// x = A^T y in gather form, one thread per (i, j) column marching a chunk of
// levels like the forward operator.  Per mass level the thread needs the
// cotangents of three mass-point intermediates of the forward formulas:
//   gzm_b = -1/2 sum over the adjacent u / v faces of (formula cotangent)
//           * cua zxu (cva zev)                  (own cell, levels k-1..k+1)
//   um_b  = 1/2 (adjacent v faces) * afv + alpha/2 (eta_dot rows k, k+1) * kx
//                                                (own cell and the one to the left)
//   vm_b  = -1/2 (adjacent u faces) * afu + alpha/2 (eta_dot rows k, k+1) * ky
//                                                (own cell and the one below)
// The own-cell values are carried up the column; the neighbours' are formed
// from loads that neighbouring lanes also make (cache hits).  Outputs per
// level: u face i, v face j (and the east / north faces at the domain edge),
// pi, w and eta_dot of the column.
template <typename T, int UNR, int NB = 0, typename TC = T>
__device__ __forceinline__ void helmholtz_T_body(const OpT<T> &o, const T *__restrict__ y,
                                                 T *__restrict__ x, int kchunk, int blk, int nblk,
                                                 const double *inv = nullptr, void *vout = nullptr,
                                                 double vs = 1.0, const T *__restrict__ res = nullptr) {
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int nbx = (nx + kTX - 1) / kTX, nby = (ny + kTY - 1) / kTY;
  const int tile = xcd_tile(blk, nblk);
  const int bx = tile % nbx, by = (tile / nbx) % nby, bz = tile / (nbx * nby);
  const int i = bx * kTX + threadIdx.x, j = by * kTY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int k0 = bz * kchunk, k1 = min(k0 + kchunk, nz + 1);
  using TI = typename InT<T, NB>::type;  // input element type
  const TI *yi = reinterpret_cast<const TI *>(y);
  const TI *YU = yi + s.ou, *YV = yi + s.ov, *YW = yi + s.ow, *YP = yi + s.om, *YE = yi + s.oe;
  T *XU = x + s.ou, *XV = x + s.ov, *XW = x + s.ow, *XP = x + s.om, *XE = x + s.oe;
  FUSED_NORMALISATION_SETUP
  // output store (NB = 3: the residual res - A^T y)
  auto put = [&](T *X_, uint32_t q, T v) {
    if constexpr (NB == 3 || NB == 5) v = ldg(res + (X_ - x), q) - v;
    if constexpr (NB == 5) stg(static_cast<float *>(vout) + (X_ - x), q, (float)v);
    else stg(X_, q, so(v));
  };
  const bool px = s.px, py = s.py;
  const bool hl = i >= 1, hd = j >= 1;  // the cell to the left / below exists
  const T half = (T)0.5, mhalf = (T)-0.5, alpha = o.alpha, ha = half * o.alpha;
  const TC *__restrict__ c_cua = reinterpret_cast<const TC *>(o.cua);
  const TC *__restrict__ c_zxu = reinterpret_cast<const TC *>(o.zxu);
  const TC *__restrict__ c_cx = reinterpret_cast<const TC *>(o.cx);
  const TC *__restrict__ c_cva = reinterpret_cast<const TC *>(o.cva);
  const TC *__restrict__ c_zev = reinterpret_cast<const TC *>(o.zev);
  const TC *__restrict__ c_cy = reinterpret_cast<const TC *>(o.cy);
  const TC *__restrict__ c_dw = reinterpret_cast<const TC *>(o.dw);
  const TC *__restrict__ c_cw = reinterpret_cast<const TC *>(o.cw);
  const TC *__restrict__ c_cz = reinterpret_cast<const TC *>(o.cz);
  const TC *__restrict__ c_kx = reinterpret_cast<const TC *>(o.kx);
  const TC *__restrict__ c_ky = reinterpret_cast<const TC *>(o.ky);
  const TC *__restrict__ c_rdzw = reinterpret_cast<const TC *>(o.rdzw);
  const TC *__restrict__ c_cpm = reinterpret_cast<const TC *>(o.cpm);
  const uint32_t pm = (uint32_t)nx * ny, pu = (uint32_t)(nx + 1) * ny, pv = (uint32_t)nx * (ny + 1);
  const uint32_t mc = (uint32_t)j * nx + i;
  const uint32_t uc = (uint32_t)j * (nx + 1) + i, vc = (uint32_t)j * nx + i;
  const uint32_t u0r = (uint32_t)j * (nx + 1);     // face 0 of row j
  const uint32_t v0c = (uint32_t)i;                // face 0 of column i
  // folded face indices of the face after the cell (periodic: face nx is face 0)
  const int fr = (px && i + 1 == nx) ? 0 : i + 1;
  const int gr = (py && j + 1 == ny) ? 0 : j + 1;
  // 2-D coefficients: own faces, the faces of the left cell's rows, the lower row
  const T mudx_i = ldg(o.mudx, (uint32_t)s.u2(i, j)), mudx_r = ldg(o.mudx, (uint32_t)s.u2(fr, j));
  const T afu_i = ldg(o.afu, (uint32_t)s.u2(i, j)), afu_r = ldg(o.afu, (uint32_t)s.u2(fr, j));
  const T mvdy_j = ldg(o.mvdy, (uint32_t)s.v2(i, j)), mvdy_r = ldg(o.mvdy, (uint32_t)s.v2(i, gr));
  const T afv_j = ldg(o.afv, (uint32_t)s.v2(i, j)), afv_r = ldg(o.afv, (uint32_t)s.v2(i, gr));
  const T afv_lj = hl ? ldg(o.afv, (uint32_t)s.v2(i - 1, j)) : (T)0;
  const T afv_lr = hl ? ldg(o.afv, (uint32_t)s.v2(i - 1, gr)) : (T)0;
  const T afu_dj = hd ? ldg(o.afu, (uint32_t)s.u2(i, j - 1)) : (T)0;
  const T afu_dr = hd ? ldg(o.afu, (uint32_t)s.u2(fr, j - 1)) : (T)0;
  const T mmx_c = ldg(o.mmx, (uint32_t)s.m2(i, j)), mmy_c = ldg(o.mmy, (uint32_t)s.m2(i, j));
  const T mmx_l = hl ? ldg(o.mmx, (uint32_t)s.m2(i - 1, j)) : (T)0;
  const T mmy_d = hd ? ldg(o.mmy, (uint32_t)s.m2(i, j - 1)) : (T)0;
  // Formula cotangent of u face f of row `row` (level offset q): faces 0 / nx are
  // identities unless periodic, where both carry the merged cotangent.
  auto ut = [&](int f, uint32_t row, uint32_t q) -> T {
    if (f == 0 || f == nx) return px ? ldx<T>(YU, row + q) + ldx<T>(YU, row + nx + q) : (T)0;
    return ldx<T>(YU, row + f + q);
  };
  auto vt = [&](int ii, int f, uint32_t q) -> T {
    if (f == 0 || f == ny) return py ? ldx<T>(YV, (uint32_t)ii + q) + ldx<T>(YV, (uint32_t)ny * nx + ii + q) : (T)0;
    return ldx<T>(YV, (uint32_t)f * nx + ii + q);
  };
  // gzm_b of the own cell at mass level k from the four face cotangents
  auto gzmb = [&](T uti, T utr, T vtj, T vtr, int k) -> T {
    const uint32_t qu = (uint32_t)k * pu, qv = (uint32_t)k * pv;
    const uint32_t al = uc + qu, ar = u0r + fr + qu, bl = vc + qv, br = (uint32_t)gr * nx + i + qv;
    T g = mhalf * (uti * ldc<T>(c_cua, al) * ldc<T>(c_zxu, al) + utr * ldc<T>(c_cua, ar) * ldc<T>(c_zxu, ar));
    g += mhalf * (vtj * ldc<T>(c_cva, bl) * ldc<T>(c_zev, bl) + vtr * ldc<T>(c_cva, br) * ldc<T>(c_zev, br));
    return g;
  };
  // carried state
  T ut_i = 0, ut_r = 0, vt_j = 0, vt_r = 0;  // own formula cotangents at level k
  T gm1 = 0, g0 = 0;                          // gzm_b at k-1 and k
  T cpmP_m1 = 0;                              // cpm * pibar at mass level k-1
  T e0 = 0, e0l = 0, e0d = 0;                 // eta_dot cotangent at w level k (interior only)
  T kx0 = 0, kx0l = 0, ky0 = 0, ky0d = 0;     // kx, ky at w level k
  T cwW0 = 0;                                 // cw * wbar at w level k (interior only)
  if (k0 < nz) {
    const uint32_t qu = (uint32_t)k0 * pu, qv = (uint32_t)k0 * pv;
    ut_i = ut(i, u0r, qu); ut_r = ut(i + 1, u0r, qu);
    vt_j = vt(i, j, qv); vt_r = vt(i, j + 1, qv);
    g0 = gzmb(ut_i, ut_r, vt_j, vt_r, k0);
  }
  if (k0 >= 1) {
    const int km = k0 - 1;
    const uint32_t qu = (uint32_t)km * pu, qv = (uint32_t)km * pv;
    gm1 = gzmb(ut(i, u0r, qu), ut(i + 1, u0r, qu), vt(i, j, qv), vt(i, j + 1, qv), km);
    const uint32_t m = mc + (uint32_t)km * pm;
    cpmP_m1 = ldc<T>(c_cpm, m) * ldx<T>(YP, m);
  }
  if (k0 >= 1 && k0 <= nz - 1) {
    const uint32_t w = mc + (uint32_t)k0 * pm;
    e0 = ldx<T>(YE, w);
    if (hl) e0l = ldx<T>(YE, w - 1);
    if (hd) e0d = ldx<T>(YE, w - nx);
    cwW0 = ldc<T>(c_cw, w) * ldx<T>(YW, w);
  }
  if (k0 <= nz) {
    const uint32_t w = mc + (uint32_t)k0 * pm;
    kx0 = ldc<T>(c_kx, w); ky0 = ldc<T>(c_ky, w);
    if (hl) kx0l = ldc<T>(c_kx, w - 1);
    if (hd) ky0d = ldc<T>(c_ky, w - nx);
  }
#pragma unroll UNR
  for (int k = k0; k < k1; ++k) {
    const uint32_t qm = (uint32_t)k * pm, qu = (uint32_t)k * pu, qv = (uint32_t)k * pv;
    const uint32_t wk = mc + qm;  // own column at w level k and at mass level k
    if (k == nz) {  // top w level: identity rows
      const T yw = ldx<T>(YW, wk), ye = ldx<T>(YE, wk);
      put(XW, wk, yw);
      put(XE, wk, ye * ldc<T>(c_rdzw, wk));
      bput(BW, wk, yw);
      bput(BE, wk, ye);
      break;
    }
    // ---- level k+1 data (carried into the next iteration) ----
    T ut_i1 = 0, ut_r1 = 0, vt_j1 = 0, vt_r1 = 0, g1 = 0;
    T e1 = 0, e1l = 0, e1d = 0, cwW1 = 0;
    const uint32_t w1 = wk + pm;
    if (k + 1 < nz) {
      const uint32_t qu1 = qu + pu, qv1 = qv + pv;
      ut_i1 = ut(i, u0r, qu1); ut_r1 = ut(i + 1, u0r, qu1);
      vt_j1 = vt(i, j, qv1); vt_r1 = vt(i, j + 1, qv1);
      g1 = gzmb(ut_i1, ut_r1, vt_j1, vt_r1, k + 1);
      e1 = ldx<T>(YE, w1);
      if (hl) e1l = ldx<T>(YE, w1 - 1);
      if (hd) e1d = ldx<T>(YE, w1 - nx);
      cwW1 = ldc<T>(c_cw, w1) * ldx<T>(YW, w1);
    }
    const T kx1 = ldc<T>(c_kx, w1), ky1 = ldc<T>(c_ky, w1);
    const T kx1l = hl ? ldc<T>(c_kx, w1 - 1) : (T)0, ky1d = hd ? ldc<T>(c_ky, w1 - nx) : (T)0;
    // ---- mass-point intermediates at level k ----
    const T umb = half * (vt_j * afv_j + vt_r * afv_r) + ha * (e0 * kx0 + e1 * kx1);
    const T vmb = mhalf * (ut_i * afu_i + ut_r * afu_r) + ha * (e0 * ky0 + e1 * ky1);
    T umb_l = 0, vmb_d = 0;
    if (hl) umb_l = half * (vt(i - 1, j, qv) * afv_lj + vt(i - 1, j + 1, qv) * afv_lr) + ha * (e0l * kx0l + e1l * kx1l);
    if (hd) vmb_d = mhalf * (ut(i, u0r - (nx + 1), qu) * afu_dj + ut(i + 1, u0r - (nx + 1), qu) * afu_dr) +
                    ha * (e0d * ky0d + e1d * ky1d);
    const T yp = ldx<T>(YP, wk);
    const T cpmP = ldc<T>(c_cpm, wk) * yp;
    const T cpmPl = hl ? ldc<T>(c_cpm, wk - 1) * ldx<T>(YP, wk - 1) : (T)0;
    const T cpmPd = hd ? ldc<T>(c_cpm, wk - nx) * ldx<T>(YP, wk - nx) : (T)0;
    // ---- u face i (and the east face) ----
    {
      const uint32_t au = uc + qu;
      const T yu = ldx<T>(YU, au);
      bput(BU, au, yu);
      T idp = yu;
      if (px && i == 0) idp = yu + ldx<T>(YU, au + nx);
      T fl = 0;
      if (hl) fl += cpmPl * mmx_l;
      fl -= cpmP * mmx_c;
      T out = idp + ldc<T>(c_cx, au) * fl;
      T ub = 0;
      if (hl) ub += umb_l;
      ub += umb;
      put(XU, au, out + half * ub);
      if (i == nx - 1) {
        const T yn = ldx<T>(YU, au + 1);
        bput(BU, au + 1, yn);
        const T idn = px ? (T)0 : yn;
        put(XU, au + 1, idn + ldc<T>(c_cx, au + 1) * (cpmP * mmx_c) + half * umb);
      }
    }
    // ---- v face j (and the north face) ----
    {
      const uint32_t av = vc + qv;
      const T yv = ldx<T>(YV, av);
      bput(BV, av, yv);
      T idp = yv;
      if (py && j == 0) idp = yv + ldx<T>(YV, av + (uint32_t)ny * nx);
      T fl = 0;
      if (hd) fl += cpmPd * mmy_d;
      fl -= cpmP * mmy_c;
      T out = idp + ldc<T>(c_cy, av) * fl;
      T vb = 0;
      if (hd) vb += vmb_d;
      vb += vmb;
      put(XV, av, out + half * vb);
      if (j == ny - 1) {
        const T yn = ldx<T>(YV, av + nx);
        bput(BV, av + nx, yn);
        const T idn = py ? (T)0 : yn;
        put(XV, av + nx, idn + ldc<T>(c_cy, av + nx) * (cpmP * mmy_c) + half * vmb);
      }
    }
    // ---- pi ----
    {
      const uint32_t al = uc + qu, ar = u0r + fr + qu, bl = vc + qv, br = (uint32_t)gr * nx + i + qv;
      T out = yp;
      out += ut_i * ldc<T>(c_cua, al) * mudx_i;
      out -= ut_r * ldc<T>(c_cua, ar) * mudx_r;
      out += vt_j * ldc<T>(c_cva, bl) * mvdy_j;
      out -= vt_r * ldc<T>(c_cva, br) * mvdy_r;
      if (k >= 1) out += cwW0;
      if (k + 1 <= nz - 1) out -= cwW1;
      // pressure-gradient levels: gz_w(kk) = (pi(kk) - pi(kk-1)) rdzw(kk) with the edge closures
      if (k >= 1) out += half * (gm1 + g0) * ldc<T>(c_rdzw, wk);
      if (k == 1) out += half * gm1 * ldc<T>(c_rdzw, mc);
      if (k == nz - 1) out += half * g0 * ldc<T>(c_rdzw, mc + (uint32_t)nz * pm);
      if (k + 1 <= nz - 1) out -= half * (g0 + g1) * ldc<T>(c_rdzw, w1);
      if (k == 0) out -= half * g0 * ldc<T>(c_rdzw, mc);
      if (k == nz - 2) out -= half * g1 * ldc<T>(c_rdzw, mc + (uint32_t)nz * pm);
      put(XP, wk, out);
      bput(BP, wk, yp);
    }
    // ---- w level k: w and eta_dot ----
    {
      const T yw = ldx<T>(YW, wk);
      bput(BW, wk, yw);
      if (k == 0) {
        const T ye = ldx<T>(YE, wk);
        bput(BE, wk, ye);
        put(XW, wk, yw);
        put(XE, wk, ye * ldc<T>(c_rdzw, wk));
      } else {
        bput(BE, wk, e0);
        put(XW, wk, ldc<T>(c_dw, wk) * yw - alpha * e0);
        put(XE, wk, alpha * e0 + ldc<T>(c_cz, wk) * (cpmP_m1 - cpmP));
      }
    }
    // k+1 becomes k
    ut_i = ut_i1; ut_r = ut_r1; vt_j = vt_j1; vt_r = vt_r1;
    gm1 = g0; g0 = g1;
    cpmP_m1 = cpmP;
    e0 = e1; e0l = e1l; e0d = e1d;
    kx0 = kx1; kx0l = kx1l; ky0 = ky1; ky0d = ky1d;
    cwW0 = cwW1;
  }
}

template <typename T, int UNR, int NB, typename TC>
__global__ void __launch_bounds__(256) helmholtz_T_kernel(OpT<T> o, const T *__restrict__ y,
                                                          T *__restrict__ x, int kchunk,
                                                          const int *__restrict__ stop,
                                                          const double *__restrict__ inv, void *vout,
                                                          double vs, const T *__restrict__ res) {
  if (stop && *stop) return;
  helmholtz_T_body<T, UNR, NB, TC>(o, y, x, kchunk, blockIdx.x, gridDim.x, inv, vout, vs, res);
}

// The transpose with two adjacent columns (i, i+1) per thread, i even: the
// left-cell quantities of column i+1 (its intermediate u_m cotangent, the
// Exner-flux term, the eta_dot cotangent and k_x at the w levels, the face
// between the two cells) are column i's own, and the pair's loads of mass,
// w-level and v-face arrays are two-element vector loads.  Each output is the
// same sequence of operations on the same operands as in helmholtz_T_body.
template <typename T, int NB = 0, typename TC = T>
__device__ __forceinline__ void helmholtz_T2_body(const OpT<T> &o, const T *__restrict__ y,
                                                  T *__restrict__ x, int kchunk, int blk, int nblk,
                                                  const double *inv = nullptr, void *vout = nullptr,
                                                  double vs = 1.0, const T *__restrict__ res = nullptr) {
  const Stag &s = o.s;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int nbx = (nx + kTX2 - 1) / kTX2, nby = (ny + kTY - 1) / kTY;
  const int tile = xcd_tile(blk, nblk);
  const int bx = tile % nbx, by = (tile / nbx) % nby, bz = tile / (nbx * nby);
  const int i = bx * kTX2 + 2 * threadIdx.x, j = by * kTY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int k0 = bz * kchunk, k1 = min(k0 + kchunk, nz + 1);
  using TI = typename InT<T, NB>::type;  // input element type
  const TI *yi = reinterpret_cast<const TI *>(y);
  const TI *YU = yi + s.ou, *YV = yi + s.ov, *YW = yi + s.ow, *YP = yi + s.om, *YE = yi + s.oe;
  T *XU = x + s.ou, *XV = x + s.ov, *XW = x + s.ow, *XP = x + s.om, *XE = x + s.oe;
  FUSED_NORMALISATION_SETUP
  auto put = [&](T *X_, uint32_t q, T v) {
    if constexpr (NB == 3 || NB == 5) v = ldg(res + (X_ - x), q) - v;
    if constexpr (NB == 5) stg(static_cast<float *>(vout) + (X_ - x), q, (float)v);
    else stg(X_, q, so(v));
  };
  auto put2 = [&](T *X_, uint32_t q, T va, T vb) {
    if constexpr (NB == 3 || NB == 5) {
      T ra, rb;
      ldg2(res + (X_ - x), q, ra, rb);
      va = ra - va;
      vb = rb - vb;
    }
    if constexpr (NB == 5) stg2<float>(static_cast<float *>(vout) + (X_ - x), q, (float)va, (float)vb);
    else stg2(X_, q, so(va), so(vb));
  };
  const bool px = s.px, py = s.py;
  const bool hl0 = i >= 1, hd = j >= 1;  // column i+1 always has a left cell (column i)
  const T half = (T)0.5, mhalf = (T)-0.5, alpha = o.alpha, ha = half * o.alpha;
  const TC *__restrict__ c_cua = reinterpret_cast<const TC *>(o.cua);
  const TC *__restrict__ c_zxu = reinterpret_cast<const TC *>(o.zxu);
  const TC *__restrict__ c_cx = reinterpret_cast<const TC *>(o.cx);
  const TC *__restrict__ c_cva = reinterpret_cast<const TC *>(o.cva);
  const TC *__restrict__ c_zev = reinterpret_cast<const TC *>(o.zev);
  const TC *__restrict__ c_cy = reinterpret_cast<const TC *>(o.cy);
  const TC *__restrict__ c_dw = reinterpret_cast<const TC *>(o.dw);
  const TC *__restrict__ c_cw = reinterpret_cast<const TC *>(o.cw);
  const TC *__restrict__ c_cz = reinterpret_cast<const TC *>(o.cz);
  const TC *__restrict__ c_kx = reinterpret_cast<const TC *>(o.kx);
  const TC *__restrict__ c_ky = reinterpret_cast<const TC *>(o.ky);
  const TC *__restrict__ c_rdzw = reinterpret_cast<const TC *>(o.rdzw);
  const TC *__restrict__ c_cpm = reinterpret_cast<const TC *>(o.cpm);
  const uint32_t pm = (uint32_t)nx * ny, pu = (uint32_t)(nx + 1) * ny, pv = (uint32_t)nx * (ny + 1);
  const uint32_t mc = (uint32_t)j * nx + i;
  const uint32_t uc = (uint32_t)j * (nx + 1) + i, vc = (uint32_t)j * nx + i;
  const uint32_t u0r = (uint32_t)j * (nx + 1);
  // folded face after each cell (periodic: face nx is face 0); column i's is i+1
  const int fr0 = i + 1, fr1 = (px && i + 2 == nx) ? 0 : i + 2;
  const int gr = (py && j + 1 == ny) ? 0 : j + 1;
  const T mudx_a = ldg(o.mudx, (uint32_t)s.u2(i, j)), mudx_b = ldg(o.mudx, (uint32_t)s.u2(fr0, j));
  const T mudx_c = ldg(o.mudx, (uint32_t)s.u2(fr1, j));
  const T afu_a = ldg(o.afu, (uint32_t)s.u2(i, j)), afu_b = ldg(o.afu, (uint32_t)s.u2(fr0, j));
  const T afu_c = ldg(o.afu, (uint32_t)s.u2(fr1, j));
  const T mudx_i[2] = {mudx_a, mudx_b}, mudx_r[2] = {mudx_b, mudx_c};
  const T afu_i[2] = {afu_a, afu_b}, afu_r[2] = {afu_b, afu_c};
  T mvdy_j[2], mvdy_r[2], afv_j[2], afv_r[2];
  ldg2(o.mvdy, (uint32_t)s.v2(i, j), mvdy_j[0], mvdy_j[1]);
  ldg2(o.mvdy, (uint32_t)s.v2(i, gr), mvdy_r[0], mvdy_r[1]);
  ldg2(o.afv, (uint32_t)s.v2(i, j), afv_j[0], afv_j[1]);
  ldg2(o.afv, (uint32_t)s.v2(i, gr), afv_r[0], afv_r[1]);
  const T afv_lj0 = hl0 ? ldg(o.afv, (uint32_t)s.v2(i - 1, j)) : (T)0;
  const T afv_lr0 = hl0 ? ldg(o.afv, (uint32_t)s.v2(i - 1, gr)) : (T)0;
  T afu_dj[2] = {0, 0}, afu_dr[2] = {0, 0};
  if (hd) {
    const T da = ldg(o.afu, (uint32_t)s.u2(i, j - 1)), db = ldg(o.afu, (uint32_t)s.u2(fr0, j - 1));
    const T dc = ldg(o.afu, (uint32_t)s.u2(fr1, j - 1));
    afu_dj[0] = da; afu_dr[0] = db; afu_dj[1] = db; afu_dr[1] = dc;
  }
  T mmx_c[2], mmy_c[2], mmy_d[2] = {0, 0};
  ldg2(o.mmx, (uint32_t)s.m2(i, j), mmx_c[0], mmx_c[1]);
  ldg2(o.mmy, (uint32_t)s.m2(i, j), mmy_c[0], mmy_c[1]);
  const T mmx_l0 = hl0 ? ldg(o.mmx, (uint32_t)s.m2(i - 1, j)) : (T)0;
  if (hd) ldg2(o.mmy, (uint32_t)s.m2(i, j - 1), mmy_d[0], mmy_d[1]);
  const T mmx_l[2] = {mmx_l0, mmx_c[0]};
  const int frs[2] = {fr0, fr1};
  auto ut = [&](int f, uint32_t row, uint32_t q) -> T {
    if (f == 0 || f == nx) return px ? ldx<T>(YU, row + q) + ldx<T>(YU, row + nx + q) : (T)0;
    return ldx<T>(YU, row + f + q);
  };
  auto vt = [&](int ii, int f, uint32_t q) -> T {
    if (f == 0 || f == ny) return py ? ldx<T>(YV, (uint32_t)ii + q) + ldx<T>(YV, (uint32_t)ny * nx + ii + q) : (T)0;
    return ldx<T>(YV, (uint32_t)f * nx + ii + q);
  };
  // the pair (ii = i, i+1) of v-face cotangents of face row f
  auto vt2 = [&](int f, uint32_t q, T &a, T &b) {
    if (f == 0 || f == ny) {
      if (py) {
        T a0, b0, a1, b1;
        ldx2<T>(YV, (uint32_t)i + q, a0, b0);
        ldx2<T>(YV, (uint32_t)ny * nx + i + q, a1, b1);
        a = a0 + a1;
        b = b0 + b1;
      } else {
        a = (T)0;
        b = (T)0;
      }
      return;
    }
    ldx2<T>(YV, (uint32_t)f * nx + i + q, a, b);
  };
  auto gzmb = [&](int c, T uti, T utr, T vtj, T vtr, int k) -> T {
    const uint32_t qu = (uint32_t)k * pu, qv = (uint32_t)k * pv;
    const uint32_t al = uc + c + qu, ar = u0r + frs[c] + qu, bl = vc + c + qv, br = (uint32_t)gr * nx + i + c + qv;
    T g = mhalf * (uti * ldc<T>(c_cua, al) * ldc<T>(c_zxu, al) + utr * ldc<T>(c_cua, ar) * ldc<T>(c_zxu, ar));
    g += mhalf * (vtj * ldc<T>(c_cva, bl) * ldc<T>(c_zev, bl) + vtr * ldc<T>(c_cva, br) * ldc<T>(c_zev, br));
    return g;
  };
  T ut_i[2] = {0, 0}, ut_r[2] = {0, 0}, vt_j[2] = {0, 0}, vt_r[2] = {0, 0};
  T gm1[2] = {0, 0}, g0[2] = {0, 0}, cpmP_m1[2] = {0, 0};
  T e0[2] = {0, 0}, e0l0 = 0, e0d[2] = {0, 0};
  T kx0[2] = {0, 0}, kx0l0 = 0, ky0[2] = {0, 0}, ky0d[2] = {0, 0};
  T cwW0[2] = {0, 0};
  auto load_faces = [&](uint32_t qu, uint32_t qv, T (&uti)[2], T (&utr)[2], T (&vtj)[2], T (&vtr)[2]) {
    const T fa = ut(i, u0r, qu), fb = ut(i + 1, u0r, qu), fc = ut(i + 2, u0r, qu);
    uti[0] = fa; utr[0] = fb; uti[1] = fb; utr[1] = fc;
    vt2(j, qv, vtj[0], vtj[1]);
    vt2(j + 1, qv, vtr[0], vtr[1]);
  };
  if (k0 < nz) {
    load_faces((uint32_t)k0 * pu, (uint32_t)k0 * pv, ut_i, ut_r, vt_j, vt_r);
    for (int c = 0; c < 2; ++c) g0[c] = gzmb(c, ut_i[c], ut_r[c], vt_j[c], vt_r[c], k0);
  }
  if (k0 >= 1) {
    const int km = k0 - 1;
    T a[2], b[2], cc[2], d[2];
    load_faces((uint32_t)km * pu, (uint32_t)km * pv, a, b, cc, d);
    for (int c = 0; c < 2; ++c) gm1[c] = gzmb(c, a[c], b[c], cc[c], d[c], km);
    const uint32_t m = mc + (uint32_t)km * pm;
    T cp[2], yp[2];
    ldc2<T>(c_cpm, m, cp[0], cp[1]);
    ldx2<T>(YP, m, yp[0], yp[1]);
    cpmP_m1[0] = cp[0] * yp[0];
    cpmP_m1[1] = cp[1] * yp[1];
  }
  if (k0 >= 1 && k0 <= nz - 1) {
    const uint32_t w = mc + (uint32_t)k0 * pm;
    ldx2<T>(YE, w, e0[0], e0[1]);
    if (hl0) e0l0 = ldx<T>(YE, w - 1);
    if (hd) ldx2<T>(YE, w - nx, e0d[0], e0d[1]);
    T cw[2], yw[2];
    ldc2<T>(c_cw, w, cw[0], cw[1]);
    ldx2<T>(YW, w, yw[0], yw[1]);
    cwW0[0] = cw[0] * yw[0];
    cwW0[1] = cw[1] * yw[1];
  }
  if (k0 <= nz) {
    const uint32_t w = mc + (uint32_t)k0 * pm;
    ldc2<T>(c_kx, w, kx0[0], kx0[1]);
    ldc2<T>(c_ky, w, ky0[0], ky0[1]);
    if (hl0) kx0l0 = ldc<T>(c_kx, w - 1);
    if (hd) ldc2<T>(c_ky, w - nx, ky0d[0], ky0d[1]);
  }
  for (int k = k0; k < k1; ++k) {
    const uint32_t qm = (uint32_t)k * pm, qu = (uint32_t)k * pu, qv = (uint32_t)k * pv;
    const uint32_t wk = mc + qm;
    if (k == nz) {  // top w level: identity rows
      T yw[2], ye[2], r[2];
      ldx2<T>(YW, wk, yw[0], yw[1]);
      ldx2<T>(YE, wk, ye[0], ye[1]);
      ldc2<T>(c_rdzw, wk, r[0], r[1]);
      put2(XW, wk, yw[0], yw[1]);
      put2(XE, wk, ye[0] * r[0], ye[1] * r[1]);
      bput(BW, wk, yw[0]); bput(BW, wk + 1, yw[1]);
      bput(BE, wk, ye[0]); bput(BE, wk + 1, ye[1]);
      break;
    }
    // ---- level k+1 data ----
    T ut_i1[2] = {0, 0}, ut_r1[2] = {0, 0}, vt_j1[2] = {0, 0}, vt_r1[2] = {0, 0}, g1[2] = {0, 0};
    T e1[2] = {0, 0}, e1l0 = 0, e1d[2] = {0, 0}, cwW1[2] = {0, 0};
    const uint32_t w1 = wk + pm;
    if (k + 1 < nz) {
      load_faces(qu + pu, qv + pv, ut_i1, ut_r1, vt_j1, vt_r1);
      for (int c = 0; c < 2; ++c) g1[c] = gzmb(c, ut_i1[c], ut_r1[c], vt_j1[c], vt_r1[c], k + 1);
      ldx2<T>(YE, w1, e1[0], e1[1]);
      if (hl0) e1l0 = ldx<T>(YE, w1 - 1);
      if (hd) ldx2<T>(YE, w1 - nx, e1d[0], e1d[1]);
      T cw[2], yw[2];
      ldc2<T>(c_cw, w1, cw[0], cw[1]);
      ldx2<T>(YW, w1, yw[0], yw[1]);
      cwW1[0] = cw[0] * yw[0];
      cwW1[1] = cw[1] * yw[1];
    }
    T kx1[2], ky1[2], ky1d[2] = {0, 0};
    ldc2<T>(c_kx, w1, kx1[0], kx1[1]);
    ldc2<T>(c_ky, w1, ky1[0], ky1[1]);
    const T kx1l0 = hl0 ? ldc<T>(c_kx, w1 - 1) : (T)0;
    if (hd) ldc2<T>(c_ky, w1 - nx, ky1d[0], ky1d[1]);
    // ---- mass-point intermediates at level k ----
    T umb[2], vmb[2], vmb_d[2] = {0, 0};
    for (int c = 0; c < 2; ++c) {
      umb[c] = half * (vt_j[c] * afv_j[c] + vt_r[c] * afv_r[c]) + ha * (e0[c] * kx0[c] + e1[c] * kx1[c]);
      vmb[c] = mhalf * (ut_i[c] * afu_i[c] + ut_r[c] * afu_r[c]) + ha * (e0[c] * ky0[c] + e1[c] * ky1[c]);
    }
    T umb_l0 = 0;
    if (hl0) umb_l0 = half * (vt(i - 1, j, qv) * afv_lj0 + vt(i - 1, j + 1, qv) * afv_lr0) + ha * (e0l0 * kx0l0 + e1l0 * kx1l0);
    if (hd) {
      const uint32_t row = u0r - (nx + 1);
      const T da = ut(i, row, qu), db = ut(i + 1, row, qu), dc = ut(i + 2, row, qu);
      vmb_d[0] = mhalf * (da * afu_dj[0] + db * afu_dr[0]) + ha * (e0d[0] * ky0d[0] + e1d[0] * ky1d[0]);
      vmb_d[1] = mhalf * (db * afu_dj[1] + dc * afu_dr[1]) + ha * (e0d[1] * ky0d[1] + e1d[1] * ky1d[1]);
    }
    const T umb_l[2] = {umb_l0, umb[0]};
    T yp[2], cpm[2], cpmP[2], cpmPd[2] = {0, 0};
    ldx2<T>(YP, wk, yp[0], yp[1]);
    ldc2<T>(c_cpm, wk, cpm[0], cpm[1]);
    cpmP[0] = cpm[0] * yp[0];
    cpmP[1] = cpm[1] * yp[1];
    const T cpmPl0 = hl0 ? ldc<T>(c_cpm, wk - 1) * ldx<T>(YP, wk - 1) : (T)0;
    if (hd) {
      T cd[2], pd[2];
      ldc2<T>(c_cpm, wk - nx, cd[0], cd[1]);
      ldx2<T>(YP, wk - nx, pd[0], pd[1]);
      cpmPd[0] = cd[0] * pd[0];
      cpmPd[1] = cd[1] * pd[1];
    }
    const T cpmPl[2] = {cpmPl0, cpmP[0]};
    // ---- u faces i, i+1 (and the east face) ----
    for (int c = 0; c < 2; ++c) {
      const bool hl = c == 1 || hl0;
      const uint32_t au = uc + c + qu;
      const T yu = ldx<T>(YU, au);
      bput(BU, au, yu);
      T idp = yu;
      if (px && i + c == 0) idp = yu + ldx<T>(YU, au + nx);
      T fl = 0;
      if (hl) fl += cpmPl[c] * mmx_l[c];
      fl -= cpmP[c] * mmx_c[c];
      T out = idp + ldc<T>(c_cx, au) * fl;
      T ub = 0;
      if (hl) ub += umb_l[c];
      ub += umb[c];
      put(XU, au, out + half * ub);
      if (c == 1 && i + 1 == nx - 1) {
        const T yn = ldx<T>(YU, au + 1);
        bput(BU, au + 1, yn);
        const T idn = px ? (T)0 : yn;
        put(XU, au + 1, idn + ldc<T>(c_cx, au + 1) * (cpmP[1] * mmx_c[1]) + half * umb[1]);
      }
    }
    // ---- v face j of both columns (and the north face) ----
    {
      const uint32_t av = vc + qv;
      T yv[2];
      ldx2<T>(YV, av, yv[0], yv[1]);
      bput(BV, av, yv[0]); bput(BV, av + 1, yv[1]);
      T idp[2] = {yv[0], yv[1]};
      if (py && j == 0) {
        T pa, pb;
        ldx2<T>(YV, av + (uint32_t)ny * nx, pa, pb);
        idp[0] = yv[0] + pa;
        idp[1] = yv[1] + pb;
      }
      T cyv[2];
      ldc2<T>(c_cy, av, cyv[0], cyv[1]);
      T o2[2];
      for (int c = 0; c < 2; ++c) {
        T fl = 0;
        if (hd) fl += cpmPd[c] * mmy_d[c];
        fl -= cpmP[c] * mmy_c[c];
        T out = idp[c] + cyv[c] * fl;
        T vb = 0;
        if (hd) vb += vmb_d[c];
        vb += vmb[c];
        o2[c] = out + half * vb;
      }
      put2(XV, av, o2[0], o2[1]);
      if (j == ny - 1) {
        T yn[2], cyn[2];
        ldx2<T>(YV, av + nx, yn[0], yn[1]);
        bput(BV, av + nx, yn[0]); bput(BV, av + nx + 1, yn[1]);
        ldc2<T>(c_cy, av + nx, cyn[0], cyn[1]);
        T on[2];
        for (int c = 0; c < 2; ++c) {
          const T idn = py ? (T)0 : yn[c];
          on[c] = idn + cyn[c] * (cpmP[c] * mmy_c[c]) + half * vmb[c];
        }
        put2(XV, av + nx, on[0], on[1]);
      }
    }
    // ---- pi ----
    {
      T r0[2], rn[2], r1[2], rk[2];
      ldc2<T>(c_rdzw, wk, rk[0], rk[1]);
      ldc2<T>(c_rdzw, mc, r0[0], r0[1]);
      ldc2<T>(c_rdzw, mc + (uint32_t)nz * pm, rn[0], rn[1]);
      ldc2<T>(c_rdzw, w1, r1[0], r1[1]);
      T outp[2];
      for (int c = 0; c < 2; ++c) {
        const uint32_t al = uc + c + qu, ar = u0r + frs[c] + qu, bl = vc + c + qv, br = (uint32_t)gr * nx + i + c + qv;
        T out = yp[c];
        out += ut_i[c] * ldc<T>(c_cua, al) * mudx_i[c];
        out -= ut_r[c] * ldc<T>(c_cua, ar) * mudx_r[c];
        out += vt_j[c] * ldc<T>(c_cva, bl) * mvdy_j[c];
        out -= vt_r[c] * ldc<T>(c_cva, br) * mvdy_r[c];
        if (k >= 1) out += cwW0[c];
        if (k + 1 <= nz - 1) out -= cwW1[c];
        if (k >= 1) out += half * (gm1[c] + g0[c]) * rk[c];
        if (k == 1) out += half * gm1[c] * r0[c];
        if (k == nz - 1) out += half * g0[c] * rn[c];
        if (k + 1 <= nz - 1) out -= half * (g0[c] + g1[c]) * r1[c];
        if (k == 0) out -= half * g0[c] * r0[c];
        if (k == nz - 2) out -= half * g1[c] * rn[c];
        outp[c] = out;
      }
      put2(XP, wk, outp[0], outp[1]);
      bput(BP, wk, yp[0]); bput(BP, wk + 1, yp[1]);
    }
    // ---- w level k: w and eta_dot ----
    {
      T yw[2];
      ldx2<T>(YW, wk, yw[0], yw[1]);
      bput(BW, wk, yw[0]); bput(BW, wk + 1, yw[1]);
      if (k == 0) {
        T ye[2], r[2];
        ldx2<T>(YE, wk, ye[0], ye[1]);
        ldc2<T>(c_rdzw, wk, r[0], r[1]);
        bput(BE, wk, ye[0]); bput(BE, wk + 1, ye[1]);
        put2(XW, wk, yw[0], yw[1]);
        put2(XE, wk, ye[0] * r[0], ye[1] * r[1]);
      } else {
        T dw[2], cz[2];
        ldc2<T>(c_dw, wk, dw[0], dw[1]);
        ldc2<T>(c_cz, wk, cz[0], cz[1]);
        bput(BE, wk, e0[0]); bput(BE, wk + 1, e0[1]);
        put2(XW, wk, dw[0] * yw[0] - alpha * e0[0], dw[1] * yw[1] - alpha * e0[1]);
        put2(XE, wk, alpha * e0[0] + cz[0] * (cpmP_m1[0] - cpmP[0]), alpha * e0[1] + cz[1] * (cpmP_m1[1] - cpmP[1]));
      }
    }
    // k+1 becomes k
    for (int c = 0; c < 2; ++c) {
      ut_i[c] = ut_i1[c]; ut_r[c] = ut_r1[c]; vt_j[c] = vt_j1[c]; vt_r[c] = vt_r1[c];
      gm1[c] = g0[c]; g0[c] = g1[c];
      cpmP_m1[c] = cpmP[c];
      e0[c] = e1[c]; e0d[c] = e1d[c];
      kx0[c] = kx1[c]; ky0[c] = ky1[c]; ky0d[c] = ky1d[c];
      cwW0[c] = cwW1[c];
    }
    e0l0 = e1l0;
    kx0l0 = kx1l0;
  }
}

template <typename T, int NB, typename TC>
__global__ void __launch_bounds__(256) helmholtz_T2_kernel(OpT<T> o, const T *__restrict__ y,
                                                           T *__restrict__ x, int kchunk,
                                                           const int *__restrict__ stop,
                                                           const double *__restrict__ inv, void *vout,
                                                           double vs, const T *__restrict__ res) {
  if (stop && *stop) return;
  helmholtz_T2_body<T, NB, TC>(o, y, x, kchunk, blockIdx.x, gridDim.x, inv, vout, vs, res);
}

// Launch shape of the operator and of its transpose: levels per thread
// (chunk) and the unroll of the level loop (1, 2), or 3: two adjacent
// columns per thread (even nx; otherwise unroll 1).  Every shape computes
// bit-identical results; the autotuner below picks one per grid, element type
// and direction by timing, the default is 8 chunks per column without
// unrolling.
struct AShape {
  int kc, unr;
};
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int, int>, AShape> g_tune;  // (nx, ny, nz, bytes, variant)
static uint64_t g_tune_generation = 0;  // advanced on every change of g_tune

// Captured launch sequences embed the shapes in force when they were
// recorded: they are valid while the generation is unchanged.
uint64_t shape_generation() {
  std::lock_guard<std::mutex> g(g_tune_mu);
  return g_tune_generation;
}

// variant = transpose + 2 * (fused basis normalisation) + 4 * (bfloat16
// coefficients) + 8 * (FP16 basis column as the input, in place of the fused
// normalisation); a variant not tuned on its own uses the fused-normalisation
// shape (FP16 input), then the FP32 coefficients' shape (bfloat16).
static AShape a_shape(const Stag &s, int bytes, int variant = 0) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto find = [&](int v, AShape &out) {
    auto it = g_tune.find({s.nx, s.ny, s.nz, bytes, v});
    if (it == g_tune.end()) return false;
    out = it->second;
    return true;
  };
  AShape sh{(s.nz + 1 + 7) / 8, 1};
  if (find(variant, sh)) return sh;
  if (variant & 8) {
    variant = (variant & 7) | 2;
    if (find(variant, sh)) return sh;
  }
  if (variant & 4) find(variant - 4, sh);
  return sh;
}

// Launch of kernel K<T, UNR, NB> (operator or transpose) in a given shape;
// NB = 1, 2 (fused basis normalisation) and 4 (FP16 basis column as the
// input) exist for the FP32 operator only, NB = 3 (residual epilogue) for the
// FP64 one.
template <typename T, int UNR, int NB, typename TC>
struct AKern {
  static void go(dim3 g, hipStream_t st, const OpT<T> &o, const T *x, T *y, int kc, const int *stop,
                 const Fuse &f) {
    hipLaunchKernelGGL((helmholtz_kernel<T, UNR, NB, TC>), g, dim3(kTX, kTY), 0, st, o, x, y, kc, stop, f.inv,
                       f.vout, f.vs, static_cast<const T *>(f.res));
  }
};
template <typename T, int UNR, int NB, typename TC>
struct ATKern {
  static void go(dim3 g, hipStream_t st, const OpT<T> &o, const T *x, T *y, int kc, const int *stop,
                 const Fuse &f) {
    hipLaunchKernelGGL((helmholtz_T_kernel<T, UNR, NB, TC>), g, dim3(kTX, kTY), 0, st, o, x, y, kc, stop,
                       f.inv, f.vout, f.vs, static_cast<const T *>(f.res));
  }
};
template <template <typename, int, int, typename> class K, typename T, typename TC>
static int launch_shape_tc(dim3 g, const OpT<T> &o, const T *x, T *y, hipStream_t st, const int *stop, int kc,
                           int unr, const Fuse &f) {
  if (f.res) {
    if constexpr (sizeof(T) == 8) {
      if (f.nb == 5) {
        if (unr == 2) K<T, 2, 5, TC>::go(g, st, o, x, y, kc, stop, f);
        else K<T, 1, 5, TC>::go(g, st, o, x, y, kc, stop, f);
        return SISL_OK;
      }
      if (f.nb) return SISL_ERR_ARG;
      if (unr == 2) K<T, 2, 3, TC>::go(g, st, o, x, y, kc, stop, f);
      else K<T, 1, 3, TC>::go(g, st, o, x, y, kc, stop, f);
      return SISL_OK;
    }
    return SISL_ERR_ARG;
  }
  if (f.nb == 0) {
    if (unr == 2) K<T, 2, 0, TC>::go(g, st, o, x, y, kc, stop, f);
    else K<T, 1, 0, TC>::go(g, st, o, x, y, kc, stop, f);
    return SISL_OK;
  }
  if constexpr (sizeof(T) == 4) {
    if (f.nb == 1) {
      if (unr == 2) K<T, 2, 1, TC>::go(g, st, o, x, y, kc, stop, f);
      else K<T, 1, 1, TC>::go(g, st, o, x, y, kc, stop, f);
      return SISL_OK;
    }
    if (f.nb == 2) {
      if (unr == 2) K<T, 2, 2, TC>::go(g, st, o, x, y, kc, stop, f);
      else K<T, 1, 2, TC>::go(g, st, o, x, y, kc, stop, f);
      return SISL_OK;
    }
    if (f.nb == 4) {
      if (unr == 2) K<T, 2, 4, TC>::go(g, st, o, x, y, kc, stop, f);
      else K<T, 1, 4, TC>::go(g, st, o, x, y, kc, stop, f);
      return SISL_OK;
    }
  }
  return SISL_ERR_ARG;
}
template <template <typename, int, int, typename> class K, typename T>
static int launch_shape(const OpT<T> &o, const T *x, T *y, hipStream_t st, const int *stop, AShape sh,
                        const Fuse &f) {
  const int kc = sh.kc;
  const dim3 g((unsigned)(((o.s.nx + kTX - 1) / kTX) * ((o.s.ny + kTY - 1) / kTY) *
                          ((o.s.nz + 1 + kc - 1) / kc)));
  if (f.cb) {
    if constexpr (sizeof(T) == 4) return launch_shape_tc<K, T, bf16_t>(g, o, x, y, st, stop, kc, sh.unr, f);
    return SISL_ERR_ARG;
  }
  return launch_shape_tc<K, T, T>(g, o, x, y, st, stop, kc, sh.unr, f);
}

// The two-column forward kernel (shape unroll value 3) in its variants.
template <typename T, int NB, typename TC>
static void go2(dim3 g, hipStream_t st, const OpT<T> &o, const T *x, T *y, int kc, const int *stop, const Fuse &f) {
  hipLaunchKernelGGL((helmholtz2_kernel<T, NB, TC>), g, dim3(kTX, kTY), 0, st, o, x, y, kc, stop, f.inv, f.vout,
                     f.vs, static_cast<const T *>(f.res));
}
template <typename T, typename TC>
static int launch2_tc(dim3 g, const OpT<T> &o, const T *x, T *y, hipStream_t st, const int *stop, int kc,
                      const Fuse &f) {
  if (f.res) {
    if constexpr (sizeof(T) == 8) {
      if (f.nb == 5) {
        go2<T, 5, TC>(g, st, o, x, y, kc, stop, f);
        return SISL_OK;
      }
      if (f.nb) return SISL_ERR_ARG;
      go2<T, 3, TC>(g, st, o, x, y, kc, stop, f);
      return SISL_OK;
    }
    return SISL_ERR_ARG;
  }
  if (f.nb == 0) {
    go2<T, 0, TC>(g, st, o, x, y, kc, stop, f);
    return SISL_OK;
  }
  if constexpr (sizeof(T) == 4) {
    if (f.nb == 1) {
      go2<T, 1, TC>(g, st, o, x, y, kc, stop, f);
      return SISL_OK;
    }
    if (f.nb == 2) {
      go2<T, 2, TC>(g, st, o, x, y, kc, stop, f);
      return SISL_OK;
    }
    if (f.nb == 4) {
      go2<T, 4, TC>(g, st, o, x, y, kc, stop, f);
      return SISL_OK;
    }
  }
  return SISL_ERR_ARG;
}

// The scaled variants read *inv on the device, the fused ones also store
// into vout: refused on the host when those pointers are missing.
static bool fuse_args_ok(const Fuse &f) {
  if ((f.nb == 1 || f.nb == 2 || f.nb == 4) && !f.inv) return false;
  if ((f.nb == 1 || f.nb == 2) && !f.vout) return false;
  if (f.nb == 5) return f.res && f.vout;  // residual stored in FP32
  return f.nb == 0 || f.nb == 1 || f.nb == 2 || f.nb == 4;
}

template <typename T>
static int launch_A_shape(const OpT<T> &o, const T *x, T *y, hipStream_t st, const int *stop, AShape sh,
                          const Fuse &f = Fuse{}) {
  if (!fuse_args_ok(f)) return SISL_ERR_ARG;
  if (sh.unr == 3) {
    if (two_column_ok(o.s)) {
      const int kc = sh.kc;
      const dim3 g((unsigned)(((o.s.nx + kTX2 - 1) / kTX2) * ((o.s.ny + kTY - 1) / kTY) *
                              ((o.s.nz + 1 + kc - 1) / kc)));
      if (f.cb) {
        if constexpr (sizeof(T) == 4) return launch2_tc<T, bf16_t>(g, o, x, y, st, stop, kc, f);
        return SISL_ERR_ARG;
      }
      return launch2_tc<T, T>(g, o, x, y, st, stop, kc, f);
    }
    sh.unr = 1;  // odd nx: one column per thread
  }
  return launch_shape<AKern, T>(o, x, y, st, stop, sh, f);
}

template <typename T>
int launch_A(const OpT<T> &o, const T *x, T *y, hipStream_t st, const int *stop, const Fuse *fuse) {
  if ((uint64_t)o.s.n * sizeof(T) >= (1ull << 32)) return SISL_ERR_ARG;  // 32-bit offsets
  const Fuse f = fuse ? *fuse : Fuse{};
  const AShape sh = a_shape(o.s, (int)sizeof(T), (f.nb == 4 ? 8 : f.nb && f.nb != 5 ? 2 : 0) + (f.cb ? 4 : 0));
  const Stag &s = o.s;
  const double sc = f.cb ? 2.0 : (double)sizeof(T);  // bytes per coefficient element
  const double si = f.nb == 4 ? 2.0 : (double)sizeof(T);  // bytes per input element
  const double so = f.nb == 5 ? 4.0 : (double)sizeof(T);  // bytes per output element
  const double bytes = (so + si) * s.n + sc * (3.0 * s.u.n() + 3.0 * s.v.n() + 6.0 * s.w.n() + s.m.n()) +
                       (f.nb == 1 ? 2.0 : f.nb == 2 ? (double)sizeof(T) : 0.0) * s.n +
                       (f.res ? (double)sizeof(T) * s.n : 0.0);
  const int tid = timer_begin(f.nb == 4 ? SISL_T_HELMHOLTZ_H16 : SISL_T_HELMHOLTZ, sizeof(T) == 4, bytes, stop,
                              nullptr, st);
  const int rc = launch_A_shape(o, x, y, st, stop, sh, f);
  timer_end(tid, st);
  if (rc) return rc;
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

template <typename T, int NB, typename TC>
static void goT2(dim3 g, hipStream_t st, const OpT<T> &o, const T *y, T *x, int kc, const int *stop,
                 const Fuse &f) {
  hipLaunchKernelGGL((helmholtz_T2_kernel<T, NB, TC>), g, dim3(kTX, kTY), 0, st, o, y, x, kc, stop, f.inv,
                     f.vout, f.vs, static_cast<const T *>(f.res));
}
template <typename T, typename TC>
static int launchT2_tc(dim3 g, const OpT<T> &o, const T *y, T *x, hipStream_t st, const int *stop, int kc,
                       const Fuse &f) {
  if (f.res) {
    if constexpr (sizeof(T) == 8) {
      if (f.nb == 5) {
        goT2<T, 5, TC>(g, st, o, y, x, kc, stop, f);
        return SISL_OK;
      }
      if (f.nb) return SISL_ERR_ARG;
      goT2<T, 3, TC>(g, st, o, y, x, kc, stop, f);
      return SISL_OK;
    }
    return SISL_ERR_ARG;
  }
  if (f.nb == 0) {
    goT2<T, 0, TC>(g, st, o, y, x, kc, stop, f);
    return SISL_OK;
  }
  if constexpr (sizeof(T) == 4) {
    if (f.nb == 1) {
      goT2<T, 1, TC>(g, st, o, y, x, kc, stop, f);
      return SISL_OK;
    }
    if (f.nb == 2) {
      goT2<T, 2, TC>(g, st, o, y, x, kc, stop, f);
      return SISL_OK;
    }
    if (f.nb == 4) {
      goT2<T, 4, TC>(g, st, o, y, x, kc, stop, f);
      return SISL_OK;
    }
  }
  return SISL_ERR_ARG;
}

template <typename T>
static int launch_AT_shape(const OpT<T> &o, const T *y, T *x, hipStream_t st, const int *stop, AShape sh,
                           const Fuse &f = Fuse{}) {
  if (!fuse_args_ok(f)) return SISL_ERR_ARG;
  if (sh.unr == 3) {
    if (two_column_ok(o.s)) {
      const int kc = sh.kc;
      const dim3 g((unsigned)(((o.s.nx + kTX2 - 1) / kTX2) * ((o.s.ny + kTY - 1) / kTY) *
                              ((o.s.nz + 1 + kc - 1) / kc)));
      if (f.cb) {
        if constexpr (sizeof(T) == 4) return launchT2_tc<T, bf16_t>(g, o, y, x, st, stop, kc, f);
        return SISL_ERR_ARG;
      }
      return launchT2_tc<T, T>(g, o, y, x, st, stop, kc, f);
    }
    sh.unr = 1;  // odd nx: one column per thread
  }
  return launch_shape<ATKern, T>(o, y, x, st, stop, sh, f);
}

template <typename T>
int launch_AT(const OpT<T> &o, const T *y, T *x, T * /*scratch: unused*/, hipStream_t st, const int *stop,
              const Fuse *fuse) {
  if ((uint64_t)o.s.n * sizeof(T) >= (1ull << 32)) return SISL_ERR_ARG;  // 32-bit offsets
  const Fuse f = fuse ? *fuse : Fuse{};
  const AShape sh = a_shape(o.s, (int)sizeof(T), (f.nb == 4 ? 9 : f.nb && f.nb != 5 ? 3 : 1) + (f.cb ? 4 : 0));
  const Stag &s = o.s;
  // same arrays as the forward operator: y in, x out, 13 coefficient arrays (+ the basis column)
  const double sc = f.cb ? 2.0 : (double)sizeof(T);  // bytes per coefficient element
  const double si = f.nb == 4 ? 2.0 : (double)sizeof(T);  // bytes per input element
  const double so = f.nb == 5 ? 4.0 : (double)sizeof(T);  // bytes per output element
  const double bytes = (so + si) * s.n + sc * (3.0 * s.u.n() + 3.0 * s.v.n() + 6.0 * s.w.n() + s.m.n()) +
                       (f.nb == 1 ? 2.0 : f.nb == 2 ? (double)sizeof(T) : 0.0) * s.n +
                       (f.res ? (double)sizeof(T) * s.n : 0.0);
  const int tid = timer_begin(f.nb == 4 ? SISL_T_HELMHOLTZ_T_H16 : SISL_T_HELMHOLTZ_T, sizeof(T) == 4, bytes, stop,
                              nullptr, st);
  const int rc = launch_AT_shape(o, y, x, st, stop, sh, f);
  timer_end(tid, st);
  if (rc) return rc;
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

// Time every launch shape of the operator and of its transpose (FP32 also
// with the fused FP16 basis normalisation and with an FP16 basis column as
// the input) on the caller's buffers and keep the fastest of each for this
// grid and element type.
__global__ void set_one_kernel(double *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 1.0;
}

// Bytes of caller scratch the autotuner needs: FP32 times the fused-
// normalisation variants, which read a scale (1.0) from device memory and
// store an FP16 basis column; FP64 needs none.
//
// Every timed launch follows a copy through a flush region larger than the
// 256 MiB last-level cache: launched back to back on the same buffers, a
// shape's inputs would stay partly cache-resident, which the solver never
// sees (its other passes stream gigabytes between two operator launches),
// and the ranking of near-equal shapes followed the cache instead.
constexpr int64_t kTuneFlush = 512ll << 20;
static int64_t fused_scratch_bytes(const Stag &s, bool fp32) { return fp32 ? 256 + 2 * (int64_t)s.n : 0; }
int64_t autotune_scratch_bytes(const Stag &s, bool fp32) {
  return ((fused_scratch_bytes(s, fp32) + 255) & ~(int64_t)255) + kTuneFlush;
}
__global__ void __launch_bounds__(256) tune_flush_kernel(const uint4 *__restrict__ a, uint4 *__restrict__ b,
                                                         int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) b[q] = a[q];
}

template <typename T>
static int autotune_A(const OpT<T> &o, const T *x, T *y, void *scratch, hipStream_t st, int32_t *best_kc,
                      int32_t *best_unr, bool cb = false) {
  if ((uint64_t)o.s.n * sizeof(T) >= (1ull << 32)) return SISL_ERR_ARG;
  const int levels = o.s.nz + 1;
  std::vector<AShape> cand;
  for (int parts : {1, 2, 4, 6, 8, 12, 16})
    for (int unr : {1, 2, 3}) {
      const int kc = (levels + parts - 1) / parts;
      if (unr == 3 && !two_column_ok(o.s)) continue;
      if (kc >= 2) cand.push_back({kc, unr});
    }
  // FP32: 0, 1 plain, 2, 3 fused normalisation, 4, 5 FP16 basis input
  // (operator, transpose); FP64: 0, 1
  const int nvar = sizeof(T) == 4 ? 6 : 2;
  // bfloat16 coefficients (cb): the inner solver's variants only
  const int var0 = cb ? 2 : 0;
  if (cb && sizeof(T) != 4) return SISL_ERR_ARG;
  // fused variants: scale 1 from device memory, basis column into the caller's scratch
  Fuse fz;
  fz.cb = cb ? 1 : 0;
  if (sizeof(T) == 4) {
    if (!scratch) return SISL_ERR_ARG;
    hipLaunchKernelGGL(set_one_kernel, dim3(1), dim3(64), 0, st, static_cast<double *>(scratch));
    SISL_CHECK_LAUNCH();
    fz.inv = static_cast<const double *>(scratch);
    fz.vout = static_cast<char *>(scratch) + 256;
    fz.nb = 1;
    fz.vs = 1.0;
  }
  if (!scratch) return SISL_ERR_ARG;
  char *flush = static_cast<char *>(scratch) + ((fused_scratch_bytes(o.s, sizeof(T) == 4) + 255) & ~(int64_t)255);
  const int64_t nf = kTuneFlush / 2 / 16;
  auto flush_cache = [&]() {
    hipLaunchKernelGGL(tune_flush_kernel, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<const uint4 *>(flush), reinterpret_cast<uint4 *>(flush + kTuneFlush / 2), nf);
  };
  constexpr int kReps = 3;
  hipEvent_t ev[2 * kReps];
  int made = 0;
  for (; made < 2 * kReps; ++made)
    if (hipEventCreate(&ev[made]) != hipSuccess) break;
  if (made < 2 * kReps) {
    for (int q = 0; q < made; ++q) hipEventDestroy(ev[q]);
    return SISL_ERR_LAUNCH;
  }
  // the FP16-input variants read the caller's x as FP16 (its first half)
  Fuse fz16 = fz;
  fz16.nb = 4;
  AShape best[6];
  for (int v = 0; v < 6; ++v) best[v] = {(levels + 7) / 8, 1};
  int rc = SISL_OK;
  for (int var = var0; var < nvar && rc == SISL_OK; ++var) {
    const int tr = var & 1;
    const Fuse f = var >= 4 ? fz16 : var >= 2 ? fz : Fuse{};
    // two timing rounds over all candidates, the faster of each candidate's
    // two times counts (one round alone let clock noise pick a shape ~20%
    // slower than the best)
    std::vector<float> t(cand.size(), 1e30f);
    for (int round = 0; round < 2 && rc == SISL_OK; ++round)
      for (size_t ci = 0; ci < cand.size(); ++ci) {
        const AShape &c = cand[ci];
        auto run = [&]() {
          return tr ? launch_AT_shape(o, x, y, st, nullptr, c, f) : launch_A_shape(o, x, y, st, nullptr, c, f);
        };
        if ((rc = run())) break;  // warm-up (code objects, first-touch)
        for (int r = 0; r < kReps; ++r) {
          flush_cache();
          hipEventRecord(ev[2 * r], st);
          run();
          hipEventRecord(ev[2 * r + 1], st);
        }
        if (hipEventSynchronize(ev[2 * kReps - 1]) != hipSuccess) { rc = SISL_ERR_LAUNCH; break; }
        float ms = 0;
        for (int r = 0; r < kReps; ++r) {
          float e = 0;
          hipEventElapsedTime(&e, ev[2 * r], ev[2 * r + 1]);
          ms += e;
        }
        t[ci] = std::min(t[ci], ms);
      }
    float best_ms = 1e30f;
    for (size_t ci = 0; ci < cand.size(); ++ci)
      if (t[ci] < best_ms) { best_ms = t[ci]; best[var] = cand[ci]; }
  }
  for (int q = 0; q < 2 * kReps; ++q) hipEventDestroy(ev[q]);
  if (rc) return rc;
  SISL_CHECK_LAUNCH();
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (int var = var0; var < nvar; ++var)
      g_tune[{o.s.nx, o.s.ny, o.s.nz, (int)sizeof(T), (var >= 4 ? var + 4 : var) + (cb ? 4 : 0)}] = best[var];
    ++g_tune_generation;
  }
  if (best_kc) *best_kc = best[var0].kc;
  if (best_unr) *best_unr = best[var0].unr;
  return SISL_OK;
}

// ---- batched (ensemble) launches -------------------------------------------

int blocks_A(const Stag &s) {
  const int kc = (s.nz + 1 + 7) / 8;
  return ((s.nx + kTX - 1) / kTX) * ((s.ny + kTY - 1) / kTY) * ((s.nz + 1 + kc - 1) / kc);
}
// NB = 1: fused basis normalisation (scale *b.inv, FP16 basis column j).
// One column per thread: the members' grids are small, and the two-column
// form's halved thread count ran the 32-member ensemble's operator slower
// (1.19 vs 0.99 ms per launch).
template <typename T, int NB>
__global__ void __launch_bounds__(256) helmholtz_batched_kernel(const BOp<T> *__restrict__ tab, int j) {
  const BOp<T> &b = tab[blockIdx.y];
  if (b.stop && b.stop[j * b.sj]) return;
  const int nb = blocks_A_dev(b.o.s);
  if ((int)blockIdx.x >= nb) return;
  void *vout = nullptr;
  if constexpr (NB == 1) vout = static_cast<_Float16 *>(b.vout) + j * b.vld;
  helmholtz_body<T, 1, NB, T>(b.o, b.x + j * b.xj, b.y, (b.o.s.nz + 1 + 7) / 8, blockIdx.x, nb, b.inv, vout,
                              b.vs);
}

template <typename T, int NB>
__global__ void __launch_bounds__(256) helmholtz_T_batched_kernel(const BOp<T> *__restrict__ tab, int j) {
  const BOp<T> &b = tab[blockIdx.y];
  if (b.stop && b.stop[j * b.sj]) return;
  const int nb = blocks_A_dev(b.o.s);
  if ((int)blockIdx.x >= nb) return;
  void *vout = nullptr;
  if constexpr (NB == 1) vout = static_cast<_Float16 *>(b.vout) + j * b.vld;
  helmholtz_T_body<T, 1, NB, T>(b.o, b.x + j * b.xj, b.y, (b.o.s.nz + 1 + 7) / 8, blockIdx.x, nb, b.inv, vout,
                                b.vs);
}

template <typename T>
int launch_A_batched(const BOp<T> *tab, int M, int maxb, int j, hipStream_t st, int nb) {
  if (nb) {
    if constexpr (sizeof(T) == 4)
      hipLaunchKernelGGL((helmholtz_batched_kernel<T, 1>), dim3(maxb, M), dim3(kTX, kTY), 0, st, tab, j);
    else
      return SISL_ERR_ARG;
  } else {
    hipLaunchKernelGGL((helmholtz_batched_kernel<T, 0>), dim3(maxb, M), dim3(kTX, kTY), 0, st, tab, j);
  }
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}
template <typename T>
int launch_AT_batched(const BOp<T> *tab, int M, int maxb, int j, hipStream_t st, int nb) {
  if (nb) {
    if constexpr (sizeof(T) == 4)
      hipLaunchKernelGGL((helmholtz_T_batched_kernel<T, 1>), dim3(maxb, M), dim3(kTX, kTY), 0, st, tab, j);
    else
      return SISL_ERR_ARG;
  } else {
    hipLaunchKernelGGL((helmholtz_T_batched_kernel<T, 0>), dim3(maxb, M), dim3(kTX, kTY), 0, st, tab, j);
  }
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}
template int launch_A_batched<double>(const BOp<double> *, int, int, int, hipStream_t, int);
template int launch_A_batched<float>(const BOp<float> *, int, int, int, hipStream_t, int);
template int launch_AT_batched<double>(const BOp<double> *, int, int, int, hipStream_t, int);
template int launch_AT_batched<float>(const BOp<float> *, int, int, int, hipStream_t, int);

template int launch_A<double>(const OpT<double> &, const double *, double *, hipStream_t, const int *,
                              const Fuse *);
template int launch_A<float>(const OpT<float> &, const float *, float *, hipStream_t, const int *,
                             const Fuse *);
template int launch_AT<double>(const OpT<double> &, const double *, double *, double *, hipStream_t,
                               const int *, const Fuse *);
template int launch_AT<float>(const OpT<float> &, const float *, float *, float *, hipStream_t,
                              const int *, const Fuse *);

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
int sisl_set_operator_shape(const SislGrid *g, int32_t fp32, int32_t chunk, int32_t unroll) {
  if (!g || chunk < 2 || unroll < 1 || unroll > 3) return SISL_ERR_ARG;
  std::lock_guard<std::mutex> lk(sisl::g_tune_mu);
  for (int var = 0; var < 16; ++var) sisl::g_tune[{g->nx, g->ny, g->nz, fp32 ? 4 : 8, var}] = sisl::AShape{chunk, unroll};
  ++sisl::g_tune_generation;
  return SISL_OK;
}

int sisl_get_operator_shape(const SislGrid *g, int32_t fp32, int32_t variant, int32_t *chunk, int32_t *unroll) {
  if (!g || !chunk || !unroll || variant < 0 || variant > 15) return SISL_ERR_ARG;
  const sisl::AShape sh = sisl::a_shape(sisl::Stag::make(*g), fp32 ? 4 : 8, variant);
  *chunk = sh.kc;
  *unroll = sh.unr;
  return SISL_OK;
}
int64_t sisl_autotune_scratch_bytes(const SislGrid *g, int32_t fp32) {
  if (!g) return -1;
  return sisl::autotune_scratch_bytes(sisl::Stag::make(*g), fp32 != 0);
}
int sisl_autotune_operator(const SislOperator *op, int32_t fp32, void *x, void *y, void *scratch,
                           int32_t *chunk, int32_t *unroll, void *s) {
  if (!op || !x || !y || x == y) return SISL_ERR_ARG;
  if (fp32 < 0 || fp32 > 2) return SISL_ERR_ARG;
  if (fp32) {
    auto o = sisl::OpT<float>::make(*op);
    if (!o.complete()) return SISL_ERR_ARG;
    return sisl::autotune_A(o, (const float *)x, (float *)y, scratch, sisl::as_stream(s), chunk, unroll,
                            fp32 == 2);
  }
  auto o = sisl::OpT<double>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::autotune_A(o, (const double *)x, (double *)y, scratch, sisl::as_stream(s), chunk, unroll);
}
int sisl_helmholtz_apply_fused_f32(const SislOperator *op, const float *x, float *y, const double *inv,
                                   void *basis, double vs, int32_t coef_bf16, int32_t transpose, void *s) {
  if (!op || !x || !y || !inv || !basis || x == y) return SISL_ERR_ARG;
  auto o = sisl::OpT<float>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  sisl::Fuse f;
  f.inv = inv;
  f.vout = basis;
  f.nb = 1;
  f.vs = vs;
  f.cb = coef_bf16 ? 1 : 0;
  if (transpose) return sisl::launch_AT(o, x, y, (float *)nullptr, sisl::as_stream(s), nullptr, &f);
  return sisl::launch_A(o, x, y, sisl::as_stream(s), nullptr, &f);
}
int sisl_helmholtz_apply_h16_f32(const SislOperator *op, const void *x16, float *y, const double *inv,
                                 int32_t coef_bf16, int32_t transpose, void *s) {
  if (!op || !x16 || !y || !inv || x16 == (const void *)y) return SISL_ERR_ARG;
  auto o = sisl::OpT<float>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  sisl::Fuse f;
  f.inv = inv;
  f.nb = 4;
  f.cb = coef_bf16 ? 1 : 0;
  const float *x = static_cast<const float *>(x16);  // read as FP16 by the NB = 4 kernels
  if (transpose) return sisl::launch_AT(o, x, y, (float *)nullptr, sisl::as_stream(s), nullptr, &f);
  return sisl::launch_A(o, x, y, sisl::as_stream(s), nullptr, &f);
}
int sisl_helmholtz_apply_T_f64(const SislOperator *op, const double *x, double *y, double *scr,
                               void *s) {
  if (!op || !x || !y || x == y) return SISL_ERR_ARG;
  auto o = sisl::OpT<double>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::launch_AT(o, x, y, scr, sisl::as_stream(s));
}
int sisl_helmholtz_apply_T_f32(const SislOperator *op, const float *x, float *y, float *scr,
                               void *s) {
  if (!op || !x || !y || x == y) return SISL_ERR_ARG;
  auto o = sisl::OpT<float>::make(*op);
  if (!o.complete()) return SISL_ERR_ARG;
  return sisl::launch_AT(o, x, y, scr, sisl::as_stream(s));
}

}  // extern "C"
