// Flux-form semi-Lagrangian transport along one axis, and its transpose.
//
// FluxFormAdvector.advect_1d (suetes/regional3d/steppers.py:552-594): for
// every line of N cells with face Courant numbers c_f (f = 0..N),
//   M_p   = sum_{j < p} m_j                      (cumulative mass, p = 0..N)
//   x_f   = min(max(f - c_f, 0), N)              (closed line), or
//           f + N - c_f on the line repeated three times (periodic; the
//           cumulative mass of the repetition is q T + M_r for p = qN + r,
//           T = M_N)
//   Md_f  = (1 - t) M_a + t M_b,  a = floor(x_f), t = x_f - a, b = a + 1
//           (indices clamped to the cumulative-mass range)
//   out_i = Md_{i+1} - Md_i.
// One thread per line walks it in order (the lines of a y or z sweep are
// adjacent in memory across threads, so every access is coalesced); the
// cumulative mass goes to a scratch line.  The transpose walks the faces in
// order, accumulates the cumulative-mass cotangents into scratch bins (one
// thread per line: a fixed order, deterministic), takes their suffix sums
// for the mass cotangent, and forms the Courant-number cotangent
// -(M_b - M_a) times the derivative of the clamp, with ties at the clamp
// bounds split in half like the tensor operators' maximum / minimum.
#include <mutex>

#include "common.h"

namespace sisl {

template <typename T>
struct FfslLine {
  const T *c;     // cfl line base (stride `inner`)
  T *M;           // cumulative-mass scratch line (stride `inner`)
  uint32_t inner;
  int n, periodic;
  T total;
  __device__ __forceinline__ T cfl(int f) const { return ldg(c, (uint32_t)f * inner); }
  __device__ __forceinline__ int nm() const { return periodic ? 3 * n + 1 : n + 1; }
  __device__ __forceinline__ T Mv(int p) const {
    if (!periodic) return ldg(M, (uint32_t)p * inner);
    const int q = p / n, r = p - q * n;
    return (T)q * total + ldg(M, (uint32_t)r * inner);
  }
  // interpolation position of face f and its derivative with respect to c_f
  __device__ __forceinline__ T pos(int f, T &dx_dc) const {
    const T cf = cfl(f);
    if (periodic) {
      dx_dc = (T)-1;
      return (T)f + (T)n - cf;
    }
    const T x = (T)f - cf;
    const T d1 = x > (T)0 ? (T)1 : (x == (T)0 ? (T)0.5 : (T)0);   // d max(x, 0) / dx
    const T y = x > (T)0 ? x : (T)0;
    const T d2 = y < (T)n ? (T)1 : (y == (T)n ? (T)0.5 : (T)0);   // d min(y, N) / dy
    dx_dc = -(d1 * d2);
    return y < (T)n ? y : (T)n;
  }
};

template <typename T>
__global__ void __launch_bounds__(256) ffsl_kernel(const T *__restrict__ mass, const T *__restrict__ cfl,
                                                   T *__restrict__ out, T *__restrict__ Ms, uint32_t lines,
                                                   int n, uint32_t inner, int periodic) {
  const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
  if (L >= lines) return;
  const uint32_t o = L / inner, q = L - o * inner;
  const uint32_t mb = o * (uint32_t)n * inner + q, fb = o * (uint32_t)(n + 1) * inner + q;
  FfslLine<T> ln{cfl + fb, Ms + fb, inner, n, periodic, (T)0};
  T acc = 0;
  stg(ln.M, 0u, acc);
  for (int i = 0; i < n; ++i) {
    acc += ldg(mass + mb, (uint32_t)i * inner);
    stg(ln.M, (uint32_t)(i + 1) * inner, acc);
  }
  ln.total = acc;
  const int top = ln.nm() - 1;
  auto Md = [&](int f) -> T {
    T dxdc;
    const T x = ln.pos(f, dxdc);
    const T lo = floor(x);
    const T t = x - lo;
    const int li = (int)fmin(fmax(lo, (T)-1), (T)(top + 1));  // in range before the conversion
    const int a = min(max(li, 0), top), b = min(max(li + 1, 0), top);
    return ((T)1 - t) * ln.Mv(a) + t * ln.Mv(b);
  };
  T prev = Md(0);
  for (int i = 0; i < n; ++i) {
    const T next = Md(i + 1);
    stg(out + mb, (uint32_t)i * inner, next - prev);
    prev = next;
  }
}

// gout: cotangent of out; gm, gc: cotangents of mass and cfl.  S holds two
// scratch lines per line: the cumulative mass and its cotangent bins.
template <typename T>
__global__ void __launch_bounds__(256) ffsl_T_kernel(const T *__restrict__ mass, const T *__restrict__ cfl,
                                                     const T *__restrict__ gout, T *__restrict__ gm,
                                                     T *__restrict__ gc, T *__restrict__ Ms,
                                                     T *__restrict__ Gs, uint32_t lines, int n,
                                                     uint32_t inner, int periodic) {
  const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
  if (L >= lines) return;
  const uint32_t o = L / inner, q = L - o * inner;
  const uint32_t mb = o * (uint32_t)n * inner + q, fb = o * (uint32_t)(n + 1) * inner + q;
  FfslLine<T> ln{cfl + fb, Ms + fb, inner, n, periodic, (T)0};
  T *G = Gs + fb;  // bins 0..n: cotangent of M_p (bin n also carries T for the periodic repetition)
  T acc = 0;
  stg(ln.M, 0u, acc);
  stg(G, 0u, (T)0);
  for (int i = 0; i < n; ++i) {
    acc += ldg(mass + mb, (uint32_t)i * inner);
    stg(ln.M, (uint32_t)(i + 1) * inner, acc);
    stg(G, (uint32_t)(i + 1) * inner, (T)0);
  }
  ln.total = acc;
  const int top = ln.nm() - 1;
  auto bin_add = [&](int p, T v) {
    if (!periodic) {
      stg(G, (uint32_t)p * inner, ldg(G, (uint32_t)p * inner) + v);
      return;
    }
    const int qq = p / n, r = p - qq * n;
    if (r) stg(G, (uint32_t)r * inner, ldg(G, (uint32_t)r * inner) + v);  // M_0 = 0 has no cotangent
    if (qq) stg(G, (uint32_t)n * inner, ldg(G, (uint32_t)n * inner) + (T)qq * v);
  };
  for (int f = 0; f <= n; ++f) {
    T g = 0;
    if (f >= 1) g += ldg(gout + mb, (uint32_t)(f - 1) * inner);
    if (f <= n - 1) g -= ldg(gout + mb, (uint32_t)f * inner);
    T dxdc;
    const T x = ln.pos(f, dxdc);
    const T lo = floor(x);
    const T t = x - lo;
    const int li = (int)fmin(fmax(lo, (T)-1), (T)(top + 1));
    const int a = min(max(li, 0), top), b = min(max(li + 1, 0), top);
    const T Ma = ln.Mv(a), Mb = ln.Mv(b);
    stg(gc + fb, (uint32_t)f * inner, g * (Mb - Ma) * dxdc);
    bin_add(a, g * ((T)1 - t));
    bin_add(b, g * t);
  }
  // gm_j = sum_{p > j} G_p
  T suf = 0;
  for (int j = n - 1; j >= 0; --j) {
    suf += ldg(G, (uint32_t)(j + 1) * inner);
    stg(gm + mb, (uint32_t)j * inner, suf);
  }
}

template <typename T>
static int ffsl_launch(bool transpose, const T *mass, const T *cfl, const T *gout, T *out, T *gc, T *scratch,
                       int64_t outer, int32_t n, int64_t inner, int32_t periodic, void *s) {
  if (!mass || !cfl || !out || !scratch || n < 1 || outer < 1 || inner < 1) return SISL_ERR_ARG;
  if (transpose && (!gout || !gc)) return SISL_ERR_ARG;
  const int64_t lines = outer * inner, faces = outer * (n + 1) * inner;
  if (2 * faces * (int64_t)sizeof(T) >= (1ll << 32)) return SISL_ERR_ARG;  // 32-bit offsets
  const unsigned blocks = (unsigned)((lines + kBlock - 1) / kBlock);
  if (transpose)
    hipLaunchKernelGGL(ffsl_T_kernel<T>, dim3(blocks), dim3(kBlock), 0, as_stream(s), mass, cfl, gout, out, gc,
                       scratch, scratch + faces, (uint32_t)lines, n, (uint32_t)inner, periodic);
  else
    hipLaunchKernelGGL(ffsl_kernel<T>, dim3(blocks), dim3(kBlock), 0, as_stream(s), mass, cfl, out, scratch,
                       (uint32_t)lines, n, (uint32_t)inner, periodic);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

// ---- allocation-free sweep ----------------------------------------------------
// Fixed scratch of sisl_ffsl_1d_f64_noalloc: the cumulative-mass lines of
// one FP64 sweep, outer * (n + 1) * inner <= kFfslScratchMax elements
// (64 MiB, enough for a 256 x 256 x 80 field along any axis).  Every device
// the library runs on holds its own copy until the process ends.  In host
// code `scratch` names a host-side stand-in, never its device memory: the
// device address comes from hipGetSymbolAddress, once per device.
constexpr int64_t kFfslScratchMax = int64_t(1) << 23;
static __device__ __align__(256) double scratch[kFfslScratchMax];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call on each device, with that device current).
static int scratch_on_current_device(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return SISL_ERR_LAUNCH;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SISL_ERR_LAUNCH;
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return SISL_OK;
}

}  // namespace sisl

extern "C" {

int sisl_ffsl_1d_f64_noalloc(const double *mass, const double *cfl, double *out, int64_t outer, int32_t n,
                             int64_t inner, int32_t periodic, void *stream) {
  if (!mass || !cfl || !out || n < 1 || outer < 1 || inner < 1) return SISL_ERR_ARG;
  if (outer > sisl::kFfslScratchMax || inner > sisl::kFfslScratchMax ||
      outer * (int64_t)(n + 1) * inner > sisl::kFfslScratchMax)
    return SISL_ERR_TOO_LARGE;
  double *ms = nullptr;
  if (const int rc = sisl::scratch_on_current_device(&ms)) return rc;
  // every cumulative-mass element is written by its line's thread before it
  // is read, so the scratch needs no clearing between calls
  return sisl::ffsl_launch<double>(false, mass, cfl, nullptr, out, nullptr, ms, outer, n, inner, periodic,
                                   stream);
}

int sisl_ffsl_1d_f64(const double *mass, const double *cfl, double *out, double *scratch, int64_t outer,
                     int32_t n, int64_t inner, int32_t periodic, void *stream) {
  return sisl::ffsl_launch<double>(false, mass, cfl, nullptr, out, nullptr, scratch, outer, n, inner, periodic,
                                   stream);
}
int sisl_ffsl_1d_f32(const float *mass, const float *cfl, float *out, float *scratch, int64_t outer,
                     int32_t n, int64_t inner, int32_t periodic, void *stream) {
  return sisl::ffsl_launch<float>(false, mass, cfl, nullptr, out, nullptr, scratch, outer, n, inner, periodic,
                                  stream);
}
int sisl_ffsl_1d_T_f64(const double *mass, const double *cfl, const double *gout, double *gmass, double *gcfl,
                       double *scratch, int64_t outer, int32_t n, int64_t inner, int32_t periodic,
                       void *stream) {
  return sisl::ffsl_launch<double>(true, mass, cfl, gout, gmass, gcfl, scratch, outer, n, inner, periodic,
                                   stream);
}
int sisl_ffsl_1d_T_f32(const float *mass, const float *cfl, const float *gout, float *gmass, float *gcfl,
                       float *scratch, int64_t outer, int32_t n, int64_t inner, int32_t periodic,
                       void *stream) {
  return sisl::ffsl_launch<float>(true, mass, cfl, gout, gmass, gcfl, scratch, outer, n, inner, periodic,
                                  stream);
}

}  // extern "C"
