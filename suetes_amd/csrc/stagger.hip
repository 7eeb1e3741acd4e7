// Staggered-grid averages and differences of the C-grid operators, and their
// transposes.
//
// CGridOperator3D.avg / diff (suetes/regional3d/operators.py:65-132 with the
// edge closures of operators.py:37-63): along one axis of a contiguous array
// viewed as (outer, n, inner),
//   value(a) = 0.5 (f[a] + f[a+1])          (avg)
//            = (f[a+1] - f[a]) * rd         (diff, rd = 1 / spacing)
// for a in [0, n-2], then one of three closures:
//   pad 0: out = value                      (n - 1 points; e.g. u -> m)
//   pad 1: out = [value(0), value, value(n-2)]   (edges repeated; m -> u, v, w)
//   pad 2: out = [wrap, value, wrap]        (periodic m -> u, v; wrap is the
//                                            avg / diff of f[n-1] and f[0])
// and an optional multiply by a map factor m3 over the last two dimensions of
// the output (horizontal differences).  One launch replaces the 3-5 tensor
// operations of each call (slice arithmetic, scale, concatenation,
// map factor), with the same arithmetic in the same order.
//
// The transpose gathers, for every input element, the (one or two) outputs
// that read it.  `exact` = 0 follows the reference's stop-gradient on the
// repeated edge values (pad 1: those outputs pass no cotangent back).
#include <algorithm>

#include "common.h"

namespace sisl {

template <typename T>
struct StagArgs {
  uint32_t outer, inner;
  int n, op, pad, exact;
  T rd;
  const T *m3;
  uint32_t m3n;
};

template <typename T>
__global__ void __launch_bounds__(256) stag_kernel(const T *__restrict__ f, T *__restrict__ out,
                                                   StagArgs<T> a, uint32_t total) {
  const int n = a.n, nout = a.pad ? n + 1 : n - 1;
  const uint32_t inner = a.inner;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    const uint32_t q = e % inner, r = e / inner;
    const int i = (int)(r % (uint32_t)nout);
    const uint32_t o = r / (uint32_t)nout;
    const uint32_t base = o * (uint32_t)n * inner + q;
    int lo, hi;  // output i = value over (f[lo], f[hi])
    if (a.pad == 2 && (i == 0 || i == n)) {
      lo = n - 1;
      hi = 0;
    } else {
      int v = a.pad ? i - 1 : i;
      v = min(max(v, 0), n - 2);
      lo = v;
      hi = v + 1;
    }
    const T x0 = ldg(f, base + (uint32_t)lo * inner), x1 = ldg(f, base + (uint32_t)hi * inner);
    T val = a.op == 0 ? (T)0.5 * (x1 + x0) : (x1 - x0) * a.rd;
    if (a.m3) val = val * ldg(a.m3, e % a.m3n);
    stg(out, e, val);
  }
}

// g: cotangent of the output (outer, n_out, inner); gf: of f (outer, n, inner).
template <typename T>
__global__ void __launch_bounds__(256) stag_T_kernel(const T *__restrict__ g, T *__restrict__ gf,
                                                     StagArgs<T> a, uint32_t total) {
  const int n = a.n, nout = a.pad ? n + 1 : n - 1;
  const uint32_t inner = a.inner;
  // coefficient of f[lo] and of f[hi] in value(lo..hi)
  const T c_lo = a.op == 0 ? (T)0.5 : -a.rd, c_hi = a.op == 0 ? (T)0.5 : a.rd;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    const uint32_t q = e % inner, r = e / inner;
    const int i = (int)(r % (uint32_t)n);
    const uint32_t o = r / (uint32_t)n;
    const uint32_t ob = o * (uint32_t)nout * inner + q;  // output (o, 0, q)
    // cotangent of output index k after the map factor
    auto go = [&](int k) -> T {
      const uint32_t idx = ob + (uint32_t)k * inner;
      T v = ldg(g, idx);
      if (a.m3) v = v * ldg(a.m3, idx % a.m3n);
      return v;
    };
    // cotangent of value(v), v in [0, n-2]
    auto gv = [&](int v) -> T {
      if (a.pad == 0) return go(v);
      T s = go(v + 1);
      if (a.pad == 1 && a.exact) {
        if (v == 0) s += go(0);
        if (v == n - 2) s += go(n);
      }
      return s;
    };
    T acc = 0;
    if (i >= 1) acc += c_hi * gv(i - 1);
    if (i <= n - 2) acc += c_lo * gv(i);
    if (a.pad == 2 && (i == 0 || i == n - 1)) {
      const T gw = go(0) + go(n);
      acc += (i == 0 ? c_hi : c_lo) * gw;  // wrap = value over (f[n-1], f[0])
    }
    stg(gf, e, acc);
  }
}

template <typename T>
static int stag_launch(bool transpose, const T *in, T *out, int64_t outer, int32_t n, int64_t inner,
                       int32_t op, int32_t pad, int32_t exact, double rd, const T *m3, int64_t m3n,
                       void *s) {
  if (!in || !out || n < 2 || outer < 1 || inner < 1 || op < 0 || op > 1 || pad < 0 || pad > 2)
    return SISL_ERR_ARG;
  const int64_t nout = pad ? n + 1 : n - 1;
  const int64_t total = outer * (transpose ? n : nout) * inner;
  const int64_t biggest = outer * (n + 1) * inner;
  if (biggest * (int64_t)sizeof(T) >= (1ll << 32)) return SISL_ERR_ARG;  // 32-bit offsets
  if (m3 && (m3n < 1 || (outer * nout * inner) % m3n != 0)) return SISL_ERR_ARG;
  StagArgs<T> a{(uint32_t)outer, (uint32_t)inner, n, op, pad, exact, (T)rd, m3, (uint32_t)(m3 ? m3n : 1)};
  const unsigned blocks = (unsigned)std::min<int64_t>((total + kBlock - 1) / kBlock, 8192);
  if (transpose)
    hipLaunchKernelGGL(stag_T_kernel<T>, dim3(blocks), dim3(kBlock), 0, as_stream(s), in, out, a,
                       (uint32_t)total);
  else
    hipLaunchKernelGGL(stag_kernel<T>, dim3(blocks), dim3(kBlock), 0, as_stream(s), in, out, a,
                       (uint32_t)total);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

}  // namespace sisl

extern "C" {

int sisl_stagger_f64(const double *f, double *out, int64_t outer, int32_t n, int64_t inner, int32_t op,
                     int32_t pad, double rd, const double *m3, int64_t m3n, void *stream) {
  return sisl::stag_launch<double>(false, f, out, outer, n, inner, op, pad, 0, rd, m3, m3n, stream);
}
int sisl_stagger_f32(const float *f, float *out, int64_t outer, int32_t n, int64_t inner, int32_t op,
                     int32_t pad, double rd, const float *m3, int64_t m3n, void *stream) {
  return sisl::stag_launch<float>(false, f, out, outer, n, inner, op, pad, 0, rd, m3, m3n, stream);
}
int sisl_stagger_T_f64(const double *g, double *gf, int64_t outer, int32_t n, int64_t inner, int32_t op,
                       int32_t pad, int32_t exact, double rd, const double *m3, int64_t m3n,
                       void *stream) {
  return sisl::stag_launch<double>(true, g, gf, outer, n, inner, op, pad, exact, rd, m3, m3n, stream);
}
int sisl_stagger_T_f32(const float *g, float *gf, int64_t outer, int32_t n, int64_t inner, int32_t op,
                       int32_t pad, int32_t exact, double rd, const float *m3, int64_t m3n,
                       void *stream) {
  return sisl::stag_launch<float>(true, g, gf, outer, n, inner, op, pad, exact, rd, m3, m3n, stream);
}

}  // extern "C"
