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
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace sisl {

constexpr int kRedGrid = 1024;  // blocks of the reduction passes
constexpr int kMaxRestart = 64;
constexpr int kPStride = kMaxRestart + 2;  // partial sums per block

// Device-side scalar state of one solve.
struct Small {
  double H[kMaxRestart * (kMaxRestart + 1)];  // row j = Arnoldi step j (JAX layout)
  double h1[kMaxRestart + 2];
  double h2[kMaxRestart + 2];
  double nn0, nn1, inv, target;
  double g[kMaxRestart + 1], cs[kMaxRestart], sn[kMaxRestart];
  double y[kMaxRestart];
  double scal[4];
  int broken, steps, pad0, pad1;
};

// ---- block reduction helpers -------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_store_partials(double (&acc)[NV], int nv, double *partials) {
  __shared__ double red[kBlock / 64][NV];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (v < nv) {
      const double s = wave_sum(acc[v]);
      if (lane == 0) red[wid][v] = s;
    }
  }
  __syncthreads();
  for (int v = threadIdx.x; v < nv; v += blockDim.x) {
    double s = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) s += red[q][v];
    partials[(int64_t)v * kRedGrid + blockIdx.x] = s;
  }
}

// out[v] = sum over blocks of partials[v][b], fixed order.  grid = nv blocks.
__global__ void __launch_bounds__(256) reduce_partials_kernel(const double *__restrict__ partials,
                                                              double *__restrict__ out) {
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

// acc[0..nv-1] = V[v] . w ; acc[nv] = w . w
template <typename TV, int MAXV>
__global__ void __launch_bounds__(256) dots_kernel(const TV *__restrict__ V, int64_t ldv, int nv,
                                                   const TV *__restrict__ w, int64_t n,
                                                   double *__restrict__ partials) {
  double acc[MAXV + 1];
#pragma unroll
  for (int v = 0; v <= MAXV; ++v) acc[v] = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    const double we = (double)w[e];
#pragma unroll
    for (int v = 0; v < MAXV; ++v)
      if (v < nv) acc[v] += (double)V[(int64_t)v * ldv + e] * we;
    acc[MAXV] += we * we;
  }
  const double ww = acc[MAXV];  // w.w goes to slot nv
#pragma unroll
  for (int v = 0; v <= MAXV; ++v)
    if (v == nv) acc[v] = ww;
  block_store_partials<MAXV + 1>(acc, nv + 1, partials);
}

// w <- w - sum_v h[v] V[v]; then acc[v] = V[v] . w (if reproj), acc[last] = w . w
template <typename TV, int MAXV>
__global__ void __launch_bounds__(256) update_kernel(const TV *__restrict__ V, int64_t ldv, int nv,
                                                     const double *__restrict__ h,
                                                     TV *__restrict__ w, int64_t n, int reproj,
                                                     double *__restrict__ partials) {
  __shared__ double hv[MAXV];
  for (int v = threadIdx.x; v < MAXV; v += blockDim.x) hv[v] = v < nv ? h[v] : 0.0;
  __syncthreads();
  double acc[MAXV + 1];
#pragma unroll
  for (int v = 0; v <= MAXV; ++v) acc[v] = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    double vals[MAXV];
    double s = (double)w[e];
#pragma unroll
    for (int v = 0; v < MAXV; ++v) {
      if (v < nv) {
        vals[v] = (double)V[(int64_t)v * ldv + e];
        s -= hv[v] * vals[v];
      } else {
        vals[v] = 0;
      }
    }
    const TV sn = (TV)s;
    w[e] = sn;
    const double sd = (double)sn;
    if (reproj) {
#pragma unroll
      for (int v = 0; v < MAXV; ++v)
        if (v < nv) acc[v] += vals[v] * sd;
    }
    acc[MAXV] += sd * sd;
  }
  const int nout = reproj ? nv : 0;
  const double ww = acc[MAXV];
#pragma unroll
  for (int v = 0; v <= MAXV; ++v)
    if (v == nout) acc[v] = ww;
  block_store_partials<MAXV + 1>(acc, nout + 1, partials);
}

// Arnoldi bookkeeping for step j (one thread): JAX's H row, breakdown test
// (v_norm_1 <= eps * v_norm_0), Givens residual estimate for early exit.
__global__ void finalize_kernel(Small *sm, int j, int R, int early_exit) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (sm->broken) {
    sm->inv = 0.0;
    return;
  }
  double h[kMaxRestart + 2];
  for (int v = 0; v <= j; ++v) h[v] = sm->h1[v] + sm->h2[v];
  const double vn0 = sqrt(sm->nn0);
  double vn1 = sqrt(sm->nn1);
  const double thresh = 2.220446049250313e-16 * vn0;
  if (!(vn1 > thresh)) vn1 = 0.0;
  h[j + 1] = vn1;
  double *row = sm->H + (int64_t)j * (R + 1);
  for (int v = 0; v <= R; ++v) row[v] = v <= j + 1 ? h[v] : 0.0;
  sm->inv = vn1 > 0.0 ? 1.0 / vn1 : 0.0;
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

// V[j+1] = w * inv (exact zero once the cycle stopped)
template <typename TV>
__global__ void normalize_kernel(const TV *__restrict__ w, TV *__restrict__ out, int64_t n,
                                 const Small *__restrict__ sm) {
  const double inv = sm->inv;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride)
    out[e] = inv != 0.0 ? (TV)((double)w[e] * inv) : (TV)0;
}

// out = src * scal[idx] safe-normalised (JAX _safe_normalize, thresh eps)
template <typename TS, typename TD>
__global__ void unit_kernel(const TS *__restrict__ src, TD *__restrict__ out, int64_t n,
                            const double *__restrict__ norm2) {
  const double nr = sqrt(*norm2);
  const double inv = nr > 2.220446049250313e-16 ? 1.0 / nr : 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride)
    out[e] = inv != 0.0 ? (TD)((double)src[e] * inv) : (TD)0;
}

// t = b - t
template <typename T>
__global__ void residual_kernel(const T *__restrict__ b, T *__restrict__ t, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride)
    t[e] = b[e] - t[e];
}

// x += scale * sum_v y[v] V[v]
template <typename TV, int MAXV>
__global__ void __launch_bounds__(256) combine_kernel(const TV *__restrict__ V, int64_t ldv, int nv,
                                                      const double *__restrict__ y,
                                                      double *__restrict__ x, int64_t n) {
  double yv[MAXV];
#pragma unroll
  for (int v = 0; v < MAXV; ++v) yv[v] = v < nv ? y[v] : 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    double s = 0;
#pragma unroll
    for (int v = 0; v < MAXV; ++v)
      if (v < nv) s += yv[v] * (double)V[(int64_t)v * ldv + e];
    x[e] += s;
  }
}

__global__ void reset_cycle_kernel(Small *sm, int R, double g0, double target) {
  const int t = threadIdx.x;
  for (int q = t; q < R * (R + 1); q += blockDim.x) {
    const int r = q / (R + 1), c = q % (R + 1);
    sm->H[q] = r == c ? 1.0 : 0.0;
  }
  if (t == 0) {
    sm->broken = 0;
    sm->steps = 0;
    sm->inv = 0.0;
    for (int v = 0; v <= R; ++v) sm->g[v] = 0.0;
    sm->g[0] = g0;
    sm->target = target;
  }
}

// ---- host-side dispatch -------------------------------------------------------

template <typename TV, int MAXV>
static void dots_launch(const TV *V, int64_t ldv, int nv, const TV *w, int64_t n, double *partials,
                        hipStream_t s) {
  hipLaunchKernelGGL((dots_kernel<TV, MAXV>), dim3(kRedGrid), dim3(kBlock), 0, s, V, ldv, nv, w, n,
                     partials);
}
template <typename TV>
static void dots(const TV *V, int64_t ldv, int nv, const TV *w, int64_t n, double *partials,
                 hipStream_t s) {
  if (nv <= 8) dots_launch<TV, 8>(V, ldv, nv, w, n, partials, s);
  else if (nv <= 16) dots_launch<TV, 16>(V, ldv, nv, w, n, partials, s);
  else if (nv <= 24) dots_launch<TV, 24>(V, ldv, nv, w, n, partials, s);
  else if (nv <= 32) dots_launch<TV, 32>(V, ldv, nv, w, n, partials, s);
  else if (nv <= 48) dots_launch<TV, 48>(V, ldv, nv, w, n, partials, s);
  else dots_launch<TV, 64>(V, ldv, nv, w, n, partials, s);
}
template <typename TV, int MAXV>
static void update_launch(const TV *V, int64_t ldv, int nv, const double *h, TV *w, int64_t n,
                          int reproj, double *partials, hipStream_t s) {
  hipLaunchKernelGGL((update_kernel<TV, MAXV>), dim3(kRedGrid), dim3(kBlock), 0, s, V, ldv, nv, h, w,
                     n, reproj, partials);
}
template <typename TV>
static void update(const TV *V, int64_t ldv, int nv, const double *h, TV *w, int64_t n, int reproj,
                   double *partials, hipStream_t s) {
  if (nv <= 8) update_launch<TV, 8>(V, ldv, nv, h, w, n, reproj, partials, s);
  else if (nv <= 16) update_launch<TV, 16>(V, ldv, nv, h, w, n, reproj, partials, s);
  else if (nv <= 24) update_launch<TV, 24>(V, ldv, nv, h, w, n, reproj, partials, s);
  else if (nv <= 32) update_launch<TV, 32>(V, ldv, nv, h, w, n, reproj, partials, s);
  else if (nv <= 48) update_launch<TV, 48>(V, ldv, nv, h, w, n, reproj, partials, s);
  else update_launch<TV, 64>(V, ldv, nv, h, w, n, reproj, partials, s);
}
template <typename TV, int MAXV>
static void combine_launch(const TV *V, int64_t ldv, int nv, const double *y, double *x, int64_t n,
                           hipStream_t s) {
  hipLaunchKernelGGL((combine_kernel<TV, MAXV>), dim3(kRedGrid), dim3(kBlock), 0, s, V, ldv, nv, y, x,
                     n);
}
template <typename TV>
static void combine(const TV *V, int64_t ldv, int nv, const double *y, double *x, int64_t n,
                    hipStream_t s) {
  if (nv <= 8) combine_launch<TV, 8>(V, ldv, nv, y, x, n, s);
  else if (nv <= 16) combine_launch<TV, 16>(V, ldv, nv, y, x, n, s);
  else if (nv <= 32) combine_launch<TV, 32>(V, ldv, nv, y, x, n, s);
  else combine_launch<TV, 64>(V, ldv, nv, y, x, n, s);
}

static inline int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

// Workspace carve-up (byte offsets).
struct Layout {
  int64_t V, w, t, scr, partials, small, V32, w32, t32, scr32, total;
};

static Layout layout(const Stag &s, int R, bool mixed) {
  Layout L{};
  int64_t o = 0;
  const int64_t n = s.n, nm = s.m.n();
  auto take = [&](int64_t bytes) { const int64_t r = o; o += align256(bytes); return r; };
  L.small = take(sizeof(Small));
  L.partials = take(sizeof(double) * kPStride * kRedGrid);
  L.w = take(sizeof(double) * n);
  L.t = take(sizeof(double) * n);
  L.scr = take(sizeof(double) * 3 * nm);
  if (!mixed) {
    L.V = take(sizeof(double) * n * (R + 1));
  } else {
    L.V32 = take(sizeof(float) * n * (R + 1));
    L.w32 = take(sizeof(float) * n);
    L.t32 = take(sizeof(float) * n);
    L.scr32 = take(sizeof(float) * 3 * nm);
  }
  L.total = o;
  return L;
}

// Least squares of JAX's _lstsq (normal equations) on the leading s steps.
static void host_lstsq(const double *H, int R, int s, double beta, double *y) {
  std::vector<double> G((size_t)s * s), r(s);
  for (int a = 0; a < s; ++a) {
    for (int b = 0; b < s; ++b) {
      double acc = 0;
      for (int l = 0; l <= s; ++l) acc += H[a * (R + 1) + l] * H[b * (R + 1) + l];
      G[(size_t)a * s + b] = acc;
    }
    r[a] = H[a * (R + 1) + 0] * beta;
  }
  // Gaussian elimination with partial pivoting
  std::vector<int> piv(s);
  for (int c = 0; c < s; ++c) {
    int p = c;
    for (int q = c + 1; q < s; ++q)
      if (std::fabs(G[(size_t)q * s + c]) > std::fabs(G[(size_t)p * s + c])) p = q;
    if (p != c) {
      for (int q = 0; q < s; ++q) std::swap(G[(size_t)c * s + q], G[(size_t)p * s + q]);
      std::swap(r[c], r[p]);
    }
    const double d = G[(size_t)c * s + c];
    for (int q = c + 1; q < s; ++q) {
      const double f = d != 0.0 ? G[(size_t)q * s + c] / d : 0.0;
      if (f == 0.0) continue;
      for (int k = c; k < s; ++k) G[(size_t)q * s + k] -= f * G[(size_t)c * s + k];
      r[q] -= f * r[c];
    }
  }
  for (int c = s - 1; c >= 0; --c) {
    double acc = r[c];
    for (int k = c + 1; k < s; ++k) acc -= G[(size_t)c * s + k] * y[k];
    const double d = G[(size_t)c * s + c];
    y[c] = d != 0.0 ? acc / d : 0.0;
  }
}

__global__ void copy_nn0_kernel(Small *sm, int nv) {
  sm->nn0 = sm->h1[nv];
  sm->h1[nv] = 0.0;
  sm->h2[nv] = 0.0;
}

template <typename TV>
struct Krylov {
  const OpT<TV> &op;
  int transpose;
  TV *V, *w, *t, *scr;
  double *partials;
  Small *sm;
  int64_t n;
  hipStream_t s;

  // w <- (M A) V[j]  (or M^T A^T)
  int apply(const TV *in, TV *out) {
    int rc;
    if (!transpose) {
      if ((rc = launch_A<TV>(op, in, out, s))) return rc;
      return launch_M<TV>(op, out, out, s);
    }
    if ((rc = launch_AT<TV>(op, in, t, scr, s))) return rc;
    return launch_MT<TV>(op, t, out, scr, s);
  }

  // One Arnoldi step with CGS2 (JAX _kth_arnoldi_iteration).
  int arnoldi(int j, int R, int early_exit) {
    int rc;
    TV *Vj = V + (int64_t)j * n;
    if ((rc = apply(Vj, w))) return rc;
    const int nv = j + 1;
    dots<TV>(V, n, nv, w, n, partials, s);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(nv + 1), dim3(kBlock), 0, s, partials, sm->h1);
    // nn0 sits in h1[nv]; copy it into place with the finalize (reads h1[nv])
    update<TV>(V, n, nv, sm->h1, w, n, 1, partials, s);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(nv + 1), dim3(kBlock), 0, s, partials, sm->h2);
    update<TV>(V, n, nv, sm->h2, w, n, 0, partials, s);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kBlock), 0, s, partials, &sm->nn1);
    hipLaunchKernelGGL(copy_nn0_kernel, dim3(1), dim3(1), 0, s, sm, nv);
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(64), 0, s, sm, j, R, early_exit);
    hipLaunchKernelGGL(normalize_kernel<TV>, dim3(kRedGrid), dim3(kBlock), 0, s, (const TV *)w,
                       V + (int64_t)(j + 1) * n, n, (const Small *)sm);
    SISL_CHECK_LAUNCH();
    return SISL_OK;
  }

};

// squared norm of a vector into *out (device), via the dots pass with nv = 0
template <typename TV>
static void norm2(const TV *x, int64_t n, double *partials, double *out, hipStream_t s) {
  dots<TV>(x, n, 0, x, n, partials, s);
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kBlock), 0, s, partials, out);
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
                   op.s.n, s};
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
  // initial residual M (b - A x0)
  auto residual = [&](double *scal_slot) -> int {
    int r2;
    if (!p.transpose) {
      if ((r2 = launch_A<double>(op, x, K.w, s))) return r2;
      hipLaunchKernelGGL(residual_kernel<double>, dim3(kRedGrid), dim3(kBlock), 0, s, b, K.w, n);
      if ((r2 = launch_M<double>(op, K.w, K.w, s))) return r2;
    } else {
      if ((r2 = launch_AT<double>(op, x, K.t, K.scr, s))) return r2;
      hipLaunchKernelGGL(residual_kernel<double>, dim3(kRedGrid), dim3(kBlock), 0, s, b, K.t, n);
      if ((r2 = launch_MT<double>(op, K.t, K.w, K.scr, s))) return r2;
    }
    norm2<double>(K.w, n, K.partials, scal_slot, s);
    hipLaunchKernelGGL((unit_kernel<double, double>), dim3(kRedGrid), dim3(kBlock), 0, s,
                       (const double *)K.w, K.V, n, (const double *)scal_slot);
    SISL_CHECK_LAUNCH();
    return SISL_OK;
  };
  if ((rc = residual(&sm->scal[2]))) return rc;
  if ((rc = sync_copy(hs, sm->scal, sizeof(double) * 3, s))) return rc;
  const double b_norm = std::sqrt(hs[0]);
  const double Mb_norm = std::sqrt(hs[1]);
  const double atol = std::fmax(p.tol * b_norm, p.atol);
  const double ptol = Mb_norm * std::fmin(1.0, atol / b_norm);  // NaN when b = 0: no cycle
  double rnorm = std::sqrt(hs[2]);
  if (!(rnorm > 2.220446049250313e-16)) rnorm = 0.0;
  std::vector<double> H((size_t)R * (R + 1));
  double y[kMaxRestart];
  int k = 0, steps_total = 0;
  while (k < p.maxiter && rnorm > ptol) {
    hipLaunchKernelGGL(reset_cycle_kernel, dim3(1), dim3(kBlock), 0, s, sm, R, rnorm,
                       p.early_exit ? ptol : 0.0);
    for (int j = 0; j < R; ++j)
      if ((rc = K.arnoldi(j, R, p.early_exit))) return rc;
    int steps = 0;
    if ((rc = sync_copy(H.data(), sm->H, sizeof(double) * R * (R + 1), s))) return rc;
    if ((rc = sync_copy(&steps, &sm->steps, sizeof(int), s))) return rc;
    steps_total += steps;
    host_lstsq(H.data(), R, steps, rnorm, y);
    if (hipMemcpyAsync(sm->y, y, sizeof(double) * steps, hipMemcpyHostToDevice, s) != hipSuccess)
      return SISL_ERR_LAUNCH;
    combine<double>(K.V, n, steps, sm->y, x, n, s);
    if ((rc = residual(&sm->scal[2]))) return rc;
    if ((rc = sync_copy(hs, sm->scal, sizeof(double) * 3, s))) return rc;
    rnorm = std::sqrt(hs[2]);
    if (!(rnorm > 2.220446049250313e-16) && !std::isnan(rnorm)) rnorm = 0.0;
    ++k;
  }
  if (st) {
    st->cycles = k;
    st->arnoldi_steps = steps_total;
    st->residual = rnorm;
    st->ptol = ptol;
    st->b_norm = b_norm;
  }
  return SISL_OK;
}

// Mixed: FP64 residual refinement around FP32 GMRES cycles with early exit.
static int gmres_mixed(const OpT<double> &op64, const OpT<float> &op32, const SislGmresParams &p,
                       const double *b, double *x, char *ws, SislGmresStats *st, hipStream_t s) {
  const int R = p.restart;
  const Layout L = layout(op64.s, R, true);
  const int64_t n = op64.s.n;
  Small *sm = (Small *)(ws + L.small);
  double *partials = (double *)(ws + L.partials);
  double *w64 = (double *)(ws + L.w), *t64 = (double *)(ws + L.t), *scr64 = (double *)(ws + L.scr);
  Krylov<float> K{op32, p.transpose, (float *)(ws + L.V32), (float *)(ws + L.w32),
                  (float *)(ws + L.t32), (float *)(ws + L.scr32), partials, sm, n, s};
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
  const double ptol = Mb_norm * std::fmin(1.0, atol / b_norm);
  const int max_outer = p.max_outer > 0 ? p.max_outer : p.maxiter;
  std::vector<double> H((size_t)R * (R + 1));
  double y[kMaxRestart];
  int outer = 0, steps_total = 0;
  double rnorm = 0;
  for (;;) {
    // r = b - A x (FP64), z = M r (FP64)
    if (!p.transpose) {
      if ((rc = launch_A<double>(op64, x, t64, s))) return rc;
    } else {
      if ((rc = launch_AT<double>(op64, x, t64, scr64, s))) return rc;
    }
    hipLaunchKernelGGL(residual_kernel<double>, dim3(kRedGrid), dim3(kBlock), 0, s, b, t64, n);
    if ((rc = M64(t64, w64))) return rc;
    norm2<double>(w64, n, partials, &sm->scal[2], s);
    hipLaunchKernelGGL((unit_kernel<double, float>), dim3(kRedGrid), dim3(kBlock), 0, s,
                       (const double *)w64, K.V, n, (const double *)&sm->scal[2]);
    if ((rc = sync_copy(hs, sm->scal, sizeof(double) * 3, s))) return rc;
    rnorm = std::sqrt(hs[2]);
    if (!(rnorm > 2.220446049250313e-16) && !std::isnan(rnorm)) rnorm = 0.0;
    if (!(rnorm > ptol) || outer >= max_outer) break;
    const double target = std::fmax(p.inner_rtol * rnorm, 0.5 * ptol);
    hipLaunchKernelGGL(reset_cycle_kernel, dim3(1), dim3(kBlock), 0, s, sm, R, rnorm, target);
    for (int j = 0; j < R; ++j)
      if ((rc = K.arnoldi(j, R, 1))) return rc;
    int steps = 0;
    if ((rc = sync_copy(H.data(), sm->H, sizeof(double) * R * (R + 1), s))) return rc;
    if ((rc = sync_copy(&steps, &sm->steps, sizeof(int), s))) return rc;
    steps_total += steps;
    host_lstsq(H.data(), R, steps, rnorm, y);
    if (hipMemcpyAsync(sm->y, y, sizeof(double) * steps, hipMemcpyHostToDevice, s) != hipSuccess)
      return SISL_ERR_LAUNCH;
    combine<float>(K.V, n, steps, sm->y, x, n, s);
    ++outer;
  }
  if (st) {
    st->cycles = outer;
    st->arnoldi_steps = steps_total;
    st->residual = rnorm;
    st->ptol = ptol;
    st->b_norm = b_norm;
  }
  return SISL_OK;
}

}  // namespace sisl

extern "C" {

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
  return sisl::gmres_mixed(o64, o32, *p, b, x, (char *)ws, st, sisl::as_stream(stream));
}

int sisl_dot_f64(const double *x, const double *y, int64_t n, double *out, double *partials,
                 void *stream) {
  if (!x || !y || !out || !partials || n < 0) return SISL_ERR_ARG;
  hipStream_t s = sisl::as_stream(stream);
  sisl::dots<double>(x, n, 1, y, n, partials, s);
  hipLaunchKernelGGL(sisl::reduce_partials_kernel, dim3(1), dim3(sisl::kBlock), 0, s, partials, out);
  SISL_CHECK_LAUNCH();
  return SISL_OK;
}

}  // extern "C"
