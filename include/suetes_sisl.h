/*
 * suetes_sisl.h -- C ABI of the MI355X-native SISL dynamical-core kernels.
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream
 * (passed as void*), allocates no device memory (all buffers are
 * caller-owned; the one exception is the measurement timer, whose
 * sisl_timer_enable keeps a 4 MiB launch record of its own) and returns 0 or
 * a negative SislStatus.  Launches are ordered on the caller's stream; the
 * process-wide tables (tuned launch shapes, captured solver cycles, the
 * timer record) are guarded by locks, so calls on different streams with
 * their own buffers may run concurrently.  Suffix _f64 / _f32 is the element
 * type of every array argument; Krylov inner products, norms and the
 * Hessenberg least squares are FP64 in both variants.
 *
 * Array layout ("internal layout"): a field of reference shape (NX, NY, NZ)
 * is stored with x fastest, i.e. element (i, j, k) at ((k*NY)+j)*NX+i.  The
 * reference stores z fastest (suetes/regional3d/geometry.py:221-224); the
 * host layer permutes once at the trajectory boundary.  Staggered fields:
 * u (nx+1, ny, nz), v (nx, ny+1, nz), w / eta_dot (nx, ny, nz+1),
 * mass points (nx, ny, nz).  A "Krylov vector" is the packed concatenation
 * u | v | w | pi | eta_dot in that order, each in the internal layout, in
 * the solver's scaled units (pi x 1000, eta_dot x dz_w_full; reference
 * suetes/regional3d/steppers.py:413-424,445-461).
 */
#ifndef SUETES_SISL_H
#define SUETES_SISL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum SislStatus {
  SISL_OK = 0,
  SISL_ERR_ARG = -1,      /* null pointer / bad size / bad option          */
  SISL_ERR_LAUNCH = -2,   /* a HIP launch or runtime call failed           */
  SISL_ERR_NOT_FINITE = -3,/* GMRES produced a non-finite residual         */
  SISL_ERR_TOO_LARGE = -4 /* input exceeds the fixed scratch of an
                             allocation-free entry point                    */
};

/* Grid descriptor shared by every entry point. */
typedef struct SislGrid {
  int32_t nx, ny, nz;          /* mass-point counts                           */
  int32_t periodic_x, periodic_y;
} SislGrid;

/*
 * Coefficients of the linear operator A (euler.py:374-439 through
 * steppers.py:439-461) and of the column preconditioner M
 * (steppers.py:153-253), precomputed once per stepper by the host layer.
 * Pointers are device pointers of the entry point's element type.
 * Shapes: [u] (nx+1,ny,nz) [v] (nx,ny+1,nz) [w] (nx,ny,nz+1) [m] (nx,ny,nz);
 * 2-D: [u2] (nx+1,ny) [v2] (nx,ny+1) [m2] (nx,ny).
 */
typedef struct SislOperator {
  SislGrid g;
  double alpha;        /* semi-implicit off-centring                         */
  /* A: pressure-gradient, Coriolis, divergence, kinematic couplings */
  const void *cua;     /* [u]  alpha dt cp th_v_u / 1000                     */
  const void *zxu;     /* [u]  dz/dxi at u (includes m_u)                    */
  const void *cx;      /* [u]  rho_u th_v_u dz_u / m_u                       */
  const void *cva;     /* [v]  alpha dt cp th_v_v / 1000                     */
  const void *zev;     /* [v]  dz/deta at v (includes m_v)                   */
  const void *cy;      /* [v]  rho_v th_v_v dz_v / m_v                       */
  const void *dw;      /* [w]  1 + dt tau                                    */
  const void *cw;      /* [w]  alpha dt cp th_v_w / dz_w / 1000              */
  const void *cz;      /* [w]  rho_w th_v_w                                  */
  const void *kx;      /* [w]  m_w z_xi_w                                    */
  const void *ky;      /* [w]  m_w z_eta_w                                   */
  const void *rdzw;    /* [w]  1 / dz_w_full                                 */
  const void *cpm;     /* [m]  1000 alpha dt C_pi / dz_m                     */
  const void *mudx;    /* [u2] m_u / dx                                      */
  const void *mvdy;    /* [v2] m_v / dy                                      */
  const void *afu;     /* [u2] alpha dt f_u                                  */
  const void *afv;     /* [v2] alpha dt f_v                                  */
  const void *mmx;     /* [m2] m_m^2 / dx                                    */
  const void *mmy;     /* [m2] m_m^2 / dy                                    */
  /* M: column Helmholtz solve */
  const void *idw;     /* [w]  1 / (1 + dt tau)                              */
  const void *gw;      /* [w]  alpha dt cp th_v_w / ((1 + dt tau) dz_w)      */
  const void *lpi;     /* [m]  alpha dt C_pi / dz_m                          */
  const void *lower;   /* [m]  sub-diagonal of the column matrix             */
  const void *invden;  /* [m]  1 / Thomas pivot                              */
  const void *cprime;  /* [m]  Thomas super-diagonal ratio c'                */
  const void *dzw0;    /* [m2] dz_w_full at the surface w level              */
  const void *dzwn;    /* [m2] dz_w_full at the top w level                  */
} SislOperator;

/* Number of elements in a packed Krylov vector for grid g. */
int64_t sisl_krylov_size(const SislGrid *g);

/* ---- (a) semi-Lagrangian interpolation ------------------------------- */

/* tensor_product_interp_3d (suetes/regional3d/operators.py:146-209):
 * Catmull-Rom tricubic of `field` (dims fnx,fny,fnz) at the departure points
 * of every output point, optional 8-corner clip.  `coords` holds departure
 * OFFSETS in index space, d = x_d - x_a (x, y, z components, each
 * n = fnx*fny*fnz long, in the internal layout): the point of output (i,j,k)
 * is (i + d_x, j + d_y, k + d_z), clamped to the field like the reference's
 * coordinates.  Offsets keep full precision in FP32. */
int sisl_tricubic_f64(int32_t fnx, int32_t fny, int32_t fnz, const double *field,
                      const double *coords, double *out, int32_t use_limiter, void *stream);
int sisl_tricubic_f32(int32_t fnx, int32_t fny, int32_t fnz, const float *field,
                      const float *coords, float *out, int32_t use_limiter, void *stream);
/* Several fields on one set of departure points (the step's pairs that
 * share one: pi' and theta' at mass points, the kinematic residual and w at
 * w levels; steppers.py:853-875): offsets, their split and the weights are
 * formed once per point.  Same bits as one tensor_product_interp_3d per
 * field.  _mixed: FP64 fields and outputs, FP32 offsets and arithmetic (each
 * field value rounded to FP32 as it is read, each result widened on the
 * store: the same bits as converting the field, gathering in FP32 and
 * widening the result). */
int sisl_tricubic2_f64(int32_t fnx, int32_t fny, int32_t fnz, const double *f0, const double *f1,
                       const double *coords, double *out0, double *out1, int32_t use_limiter, void *stream);
int sisl_tricubic2_f32(int32_t fnx, int32_t fny, int32_t fnz, const float *f0, const float *f1,
                       const float *coords, float *out0, float *out1, int32_t use_limiter, void *stream);
int sisl_tricubic_mixed(int32_t fnx, int32_t fny, int32_t fnz, const double *field, const float *coords,
                        double *out, int32_t use_limiter, void *stream);
int sisl_tricubic2_mixed(int32_t fnx, int32_t fny, int32_t fnz, const double *f0, const double *f1,
                         const float *coords, double *out0, double *out1, int32_t use_limiter, void *stream);

/* Workspace of the deterministic tricubic adjoint, int32 device arrays:
 * key[n], perm[n], start[n+1], count[n], block_sums[sisl_scan_blocks(n)]. */
int32_t sisl_scan_blocks(int64_t n);
/* Build the destination-sorted segment map of one coordinate set
 * (reusable by every adjoint call that shares the coordinates). */
int sisl_tricubic_adj_plan_f64(int32_t fnx, int32_t fny, int32_t fnz, const double *coords,
                               int32_t *key, int32_t *perm, int32_t *start, int32_t *count,
                               int32_t *block_sums, void *stream);
int sisl_tricubic_adj_plan_f32(int32_t fnx, int32_t fny, int32_t fnz, const float *coords,
                               int32_t *key, int32_t *perm, int32_t *start, int32_t *count,
                               int32_t *block_sums, void *stream);
/* Transpose of sisl_tricubic (no limiter): gfield = T^T gout, accumulated
 * per destination in a fixed order (bit-reproducible). gfield is overwritten. */
int sisl_tricubic_adj_f64(int32_t fnx, int32_t fny, int32_t fnz, const double *coords,
                          const int32_t *perm, const int32_t *start, const double *gout,
                          double *gfield, void *stream);
int sisl_tricubic_adj_f32(int32_t fnx, int32_t fny, int32_t fnz, const float *coords,
                          const int32_t *perm, const int32_t *start, const float *gout,
                          float *gfield, void *stream);
/* Faster transpose over the same plan: per sorted position the source's
 * fractions (tx, ty, tz) and base x index as four values (frac[4n]), built
 * once per coordinate set; then gf = T^T g with g gathered into sorted order
 * (g_sorted[n], scratch) and, per destination, a fixed-order sum over the 16
 * (y, z) rows of candidate bases (sources staged in LDS per 64 x 4 tile) --
 * deterministic, the same operator as sisl_tricubic_adj_* up to summation
 * order. */
int sisl_tricubic_adj_frac_f64(int32_t fnx, int32_t fny, int32_t fnz, const double *coords,
                               const int32_t *perm, double *frac, void *stream);
int sisl_tricubic_adj_frac_f32(int32_t fnx, int32_t fny, int32_t fnz, const float *coords,
                               const int32_t *perm, float *frac, void *stream);
int sisl_tricubic_adj_rows_f64(int32_t fnx, int32_t fny, int32_t fnz, const int32_t *perm,
                               const int32_t *start, const double *frac, const double *g,
                               double *g_sorted, double *gf, void *stream);
int sisl_tricubic_adj_rows_f32(int32_t fnx, int32_t fny, int32_t fnz, const int32_t *perm,
                               const int32_t *start, const float *frac, const float *g,
                               float *g_sorted, float *gf, void *stream);

/* Adjoint of the LIMITED gather (use_limiter = 1; operators.py:188-207,
 * clip to the chained min / max of the 8 cell corners), with the balanced
 * tie rule of JAX's minimum / maximum (half the cotangent to each operand
 * that equals the result).  split: per output point, the interpolant's
 * share gi[n] (to be transposed by sisl_tricubic_adj_rows_* / _adj_*) and
 * the corners' shares gc[8 n] (corner q = ox + 2 oy + 4 oz at stride n);
 * corners: gf += the corner shares gathered per cell over the plan
 * (deterministic). */
int sisl_tricubic_limited_adj_split_f64(int32_t fnx, int32_t fny, int32_t fnz, const double *field,
                                        const double *coords, const double *gout, double *gi, double *gc,
                                        void *stream);
int sisl_tricubic_limited_adj_split_f32(int32_t fnx, int32_t fny, int32_t fnz, const float *field,
                                        const float *coords, const float *gout, float *gi, float *gc,
                                        void *stream);
int sisl_tricubic_limited_adj_corners_f64(int32_t fnx, int32_t fny, int32_t fnz, const int32_t *perm,
                                          const int32_t *start, const double *gc, double *gf, void *stream);
int sisl_tricubic_limited_adj_corners_f32(int32_t fnx, int32_t fny, int32_t fnz, const int32_t *perm,
                                          const int32_t *start, const float *gc, float *gf, void *stream);

/* _get_index_velocities + compute_departure_indices
 * (suetes/regional3d/steppers.py:276-360): departure points at staggering
 * loc (0=m, 1=u, 2=v, 3=w) from trajectory winds u, v, eta_dot, written as
 * offsets x_d - x_a (3 x loc size).  m2 is the 2-D map factor at loc;
 * scratch3 holds 3 x loc size elements. */
int sisl_departure_f64(const SislGrid *g, int32_t loc, const double *u, const double *v,
                       const double *eta, const double *m2, double dx, double dy, double dt,
                       int32_t iterations, double *scratch3, double *coords_out, void *stream);
int sisl_departure_f32(const SislGrid *g, int32_t loc, const float *u, const float *v,
                       const float *eta, const float *m2, double dx, double dy, double dt,
                       int32_t iterations, float *scratch3, float *coords_out, void *stream);

/* ---- (b) operator apply ---------------------------------------------- */

/* y = A x on packed Krylov vectors (Euler3D.linear_operator in the solver's
 * scaled units, suetes/regional3d/steppers.py:439-461). */
int sisl_helmholtz_apply_f64(const SislOperator *op, const double *x, double *y, void *stream);
int sisl_helmholtz_apply_f32(const SislOperator *op, const float *x, float *y, void *stream);
/* Autotuner of the operator apply: times every launch shape (levels per
 * thread, loop unroll) on the caller's packed buffers x (input, any finite
 * values) and y (scratch output) and keeps the fastest for this grid and
 * element type (fp32 = 1: the _f32 entry points; fp32 = 2: FP32 with the
 * op's 3-D coefficient arrays in bfloat16, the mixed solver's inner
 * operator, whose fused launches are tuned on their own) for the rest of the
 * process.  All shapes give bit-identical results.  `scratch`: caller-owned
 * device memory of sisl_autotune_scratch_bytes(g, fp32) bytes: the
 * fused-normalisation variants' scale and basis column, and a 512 MiB region
 * copied before every timed launch so that no candidate runs from the
 * last-level cache.  Optional outputs: the chosen chunk and unroll (of the
 * plain operator; of its fused launch for fp32 = 2).  Synchronises on its own
 * timing events. */
int64_t sisl_autotune_scratch_bytes(const SislGrid *g, int32_t fp32);
int sisl_autotune_operator(const SislOperator *op, int32_t fp32, void *x, void *y, void *scratch,
                           int32_t *chunk, int32_t *unroll, void *stream);
/* Pin the operator launch shape for grid g and element type (chunk >= 2
 * levels per thread, unroll 1 or 2, or 3: two adjacent columns per thread,
 * which needs an even nx and otherwise runs unroll 1) of the operator and of
 * its transpose, with and without the fused basis normalisation, e.g. to
 * reproduce a tuned choice.  Every shape gives the same bits. */
int sisl_set_operator_shape(const SislGrid *g, int32_t fp32, int32_t chunk, int32_t unroll);
/* The launch shape in force for grid g, element type and variant
 * (transpose + 2 * fused normalisation + 4 * bfloat16 coefficients). */
int sisl_get_operator_shape(const SislGrid *g, int32_t fp32, int32_t variant, int32_t *chunk, int32_t *unroll);
/* y = A^T x (the transpose used by the GMRES adjoint solve,
 * steppers.py:473-481).  scratch: not used (may be NULL); kept so that the
 * transposes of A and M share one signature shape. */
int sisl_helmholtz_apply_T_f64(const SislOperator *op, const double *x, double *y,
                               double *scratch, void *stream);
int sisl_helmholtz_apply_T_f32(const SislOperator *op, const float *x, float *y,
                               float *scratch, void *stream);
/* The inner FP32 solver's operator launch with the fused basis
 * normalisation (no reference counterpart of its own: it is
 * Euler3D.linear_operator, euler.py:374-439, applied to the Arnoldi vector
 * of the mixed GMRES): y = A (x s) or A^T (x s) with s = *inv, and
 * basis = FP16(vs * s * x) stored from the elements each thread owns.
 * coef_bf16 = 1: the op's 3-D coefficient arrays are bfloat16.  Honours the
 * pinned / tuned launch shapes (all give the same bits). */
int sisl_helmholtz_apply_fused_f32(const SislOperator *op, const float *x, float *y, const double *inv,
                                   void *basis, double vs, int32_t coef_bf16, int32_t transpose,
                                   void *stream);
/* The inner FP32 solver's operator launch on a stored FP16 basis column
 * (inner_basis_operand; no reference counterpart of its own: it is
 * Euler3D.linear_operator, euler.py:374-439, applied to the Arnoldi vector
 * of the mixed GMRES): y = A (x16 s) or A^T (x16 s) with s = *inv, x16 the
 * packed vector in FP16, converted to FP32 on load; y = the plain FP32
 * launch on the converted values, times s.  coef_bf16 as above. */
int sisl_helmholtz_apply_h16_f32(const SislOperator *op, const void *x16, float *y, const double *inv,
                                 int32_t coef_bf16, int32_t transpose, void *stream);

/* ---- (c) column preconditioner ----------------------------------------- */

/* z = M r (VerticalPreconditioner.__call__, steppers.py:187-253).
 * In place (z == r) is allowed. */
int sisl_precond_apply_f64(const SislOperator *op, const double *r, double *z, void *stream);
int sisl_precond_apply_f32(const SislOperator *op, const float *r, float *z, void *stream);
/* z = M^T r (jax.linear_transpose(M_fn), steppers.py:474). scratch: 2 mass
 * arrays. z must not alias r. */
int sisl_precond_apply_T_f64(const SislOperator *op, const double *r, double *z,
                             double *scratch, void *stream);
int sisl_precond_apply_T_f32(const SislOperator *op, const float *r, float *z,
                             float *scratch, void *stream);

/* ---- GMRES (jax.scipy.sparse.linalg.gmres, batched, steppers.py:465-481) */

typedef struct SislGmresParams {
  double tol;              /* relative tolerance                              */
  double atol;             /* absolute tolerance                              */
  int32_t restart;         /* Krylov dimension per cycle (<= 64)              */
  int32_t maxiter;         /* maximum restart cycles                          */
  int32_t transpose;       /* 0: solve A x = b with M;  1: A^T x = b with M^T */
  int32_t early_exit;      /* 0: every cycle runs `restart` Arnoldi steps
                              (reference semantics); 1: stop a cycle once the
                              Givens residual estimate meets the target       */
  double inner_rtol;       /* mixed solver: inner FP32 cycle target (relative)*/
  int32_t max_outer;       /* mixed solver: refinement iterations             */
  int32_t basis_fp16;      /* mixed solver: keep the inner Krylov basis in
                              FP16 (scaled by sqrt(n)) instead of FP32        */
  int32_t coef_bf16;       /* mixed solver: the FP32 operator's 3-D coefficient
                              arrays are bfloat16 (2-D ones stay FP32)        */
  int32_t inner_single_cgs;/* mixed solver: one classical Gram-Schmidt pass per
                              inner Arnoldi step (0: iterated CGS as in the
                              FP64 solver, a second pass when the first
                              removed most of the vector)                     */
  int32_t inner_basis_operand; /* mixed solver with basis_fp16 and
                              inner_single_cgs: the inner operator reads the
                              stored FP16 basis column and each update pass
                              stores the next column (one scale per column)
                              instead of writing an FP32 Arnoldi vector back
                              (0: FP32 Arnoldi vectors between the passes)   */
  const struct SislOperator *op_refine; /* mixed solver, optional (NULL: the
                              refinement applies M in FP64): an operator with
                              FP32 coefficient arrays; the refinement forms
                              b - A x in FP64, stores it rounded to FP32 and
                              applies M (M^T) of this operator in FP32; the
                              result is the stopping test's preconditioned
                              residual and the next cycle's start vector.
                              Batched solver: non-NULL selects the same FP32
                              refinement with each member's FP32 operator    */
} SislGmresParams;

typedef struct SislGmresStats {
  int32_t cycles;          /* restart cycles (outer iterations for mixed)     */
  int32_t arnoldi_steps;   /* total Arnoldi steps executed                    */
  double residual;         /* final ||M(b - A x)||                            */
  double ptol;             /* stopping threshold                              */
  double b_norm;
} SislGmresStats;

/* Bytes of device workspace the solvers need (caller allocates, 256-B aligned). */
int64_t sisl_gmres_workspace_bytes_f64(const SislGrid *g, int32_t restart);
int64_t sisl_gmres_workspace_bytes_mixed(const SislGrid *g, int32_t restart);

/* FP64 restarted left-preconditioned GMRES. x: in = x0, out = solution. */
int sisl_gmres_f64(const SislOperator *op64, const SislGmresParams *p, const double *b,
                   double *x, void *workspace, SislGmresStats *stats, void *stream);
/* Mixed precision: FP64 outer residual refinement (A, M, norms in FP64),
 * FP32 inner GMRES (A, M, basis in FP32; inner products in FP64). */
int sisl_gmres_mixed(const SislOperator *op64, const SislOperator *op32,
                     const SislGmresParams *p, const double *b, double *x, void *workspace,
                     SislGmresStats *stats, void *stream);

/* ---- staggered-grid operators (explicit tendencies) --------------------- */

/* CGridOperator3D.avg (op 0) / diff (op 1) along one axis
 * (suetes/regional3d/operators.py:65-132, edge closures operators.py:37-63):
 * f is contiguous (outer, n, inner); out is (outer, n_out, inner) with
 * n_out = n - 1 (pad 0), or n + 1 with the edge values repeated (pad 1,
 * m -> u, v, w) or the periodic wrap value at both ends (pad 2).  diff
 * multiplies by rd = 1 / spacing; m3 (optional, m3n elements) multiplies
 * the output by a map factor indexed by (output flat index mod m3n). */
int sisl_stagger_f64(const double *f, double *out, int64_t outer, int32_t n, int64_t inner, int32_t op,
                     int32_t pad, double rd, const double *m3, int64_t m3n, void *stream);
int sisl_stagger_f32(const float *f, float *out, int64_t outer, int32_t n, int64_t inner, int32_t op,
                     int32_t pad, double rd, const float *m3, int64_t m3n, void *stream);
/* Transpose: gf = (cotangent of f) from g (cotangent of out).  exact = 0:
 * the repeated edge values of pad 1 pass nothing back (the reference's
 * stop-gradient); exact = 1: they do. */
int sisl_stagger_T_f64(const double *g, double *gf, int64_t outer, int32_t n, int64_t inner, int32_t op,
                       int32_t pad, int32_t exact, double rd, const double *m3, int64_t m3n,
                       void *stream);
int sisl_stagger_T_f32(const float *g, float *gf, int64_t outer, int32_t n, int64_t inner, int32_t op,
                       int32_t pad, int32_t exact, double rd, const float *m3, int64_t m3n,
                       void *stream);

/* ---- fused explicit tendencies and predictor-corrector (FP64) ----------- */

/*
 * Static fields of the explicit tendencies (the background and the grid's
 * metrics; precomputed once per stepper by the host layer).  Device
 * pointers to FP64 arrays in the internal layout; [u2] [v2] [m2] are 2-D
 * (per column), [u] [v] [w] [m] 3-D as in SislOperator.
 */
typedef struct SislTendency {
  SislGrid g;
  double dx, dy, dz;      /* grid spacings                                    */
  double cp, grav;        /* constants['cp'], constants['g']                  */
  double alpha, dt;       /* off-centring, time step                          */
  double nu_h, nu_div;    /* hyperdiffusion, divergence-damping coefficients  */
  const double *mm;       /* [m2] map factor at mass points                   */
  const double *mu, *mv;  /* [u2] [v2] map factors at u / v faces             */
  const double *fu, *fv;  /* [u2] [v2] Coriolis parameter                     */
  const double *dmdx, *dmdy; /* [m2] map-factor derivatives                   */
  const double *th_bg, *pi_bg;            /* [m] background theta, Exner     */
  const double *thu_bg, *thv_bg, *thw_bg; /* [u] [v] [w] its averages        */
  const double *zxu, *zev;  /* [u] [v] terrain slopes (differences of Z_m)    */
  const double *zxm, *zem;  /* [m] their mass-point averages                  */
  const double *rz;         /* [w] dz / dz_w_full                             */
  const double *zxiw, *zetaw; /* [w] z_xi_w, z_eta_w                          */
  const double *dzw, *dzm;  /* [w] dz_w_full, [m] dz_m_full                   */
  const double *cpi;        /* [m] C_pi                                       */
  const double *dthm;       /* [m] background d theta / dz at mass points     */
  const double *cx, *cy;    /* [u] [v] rho th dz / m (linear Exner flux)      */
  const double *cxp, *cyp;  /* [u] [v] rho dz / m (theta' Exner flux)         */
  const double *czs, *czp;  /* [w] dz_w rho_w th_w, dz_w rho_w                */
} SislTendency;

/* HyperFilter (suetes/regional3d/diffusion.py:93-172): the terrain-corrected
 * Cartesian horizontal Laplacian applied twice to u_m, v_m, w_m and theta'
 * of the state (u, v, w, th); lap2 receives the four mass fields (4 x N).
 * work: sisl_hyperdiffusion_workspace_f64(g) elements. */
int64_t sisl_hyperdiffusion_workspace_f64(const SislGrid *g);
int sisl_hyperdiffusion_f64(const SislTendency *t, const double *u, const double *v, const double *w,
                            const double *th, double *work, double *lap2, void *stream);
/* Euler3D.get_tendencies(is_explicit=True) (suetes/regional3d/euler.py:
 * 208-372) with nonlinear_pi_tendency and metric_curvature_tendencies
 * (euler.py:162-206), combined into the step's arrival fields
 * (suetes/regional3d/steppers.py:776-850): u_in, v_in, w_in, pi_in (pi'),
 * th_in (theta'), the kinematic residual res (w points) and theta's
 * tendency tth.  lap2: output of sisl_hyperdiffusion_f64 (NULL when
 * nu_h = 0).  Optional: x0 (packed u | v | w | pi' 1000 | eta dz_w, the
 * first solve's initial guess) and tend (raw tendencies tu | tv | tw | tpi
 * in their own layouts, N_u + N_v + N_w + N elements). */
int sisl_explicit_f64(const SislTendency *t, const double *u, const double *v, const double *w,
                      const double *pi, const double *th, const double *eta, const double *lap2,
                      double *u_in, double *v_in, double *w_in, double *pi_in, double *th_in,
                      double *res, double *tth, double *x0, double *tend, void *stream);
/* Reverse mode of the explicit stage above including its hyperdiffusion
 * (the vector-Jacobian product that the reference's autodiff forms through
 * euler.py:162-372 and diffusion.py:93-172, with the stop-gradient on the
 * edge values the staggered operators repeat, operators.py:37-63): from the
 * cotangents of u_in, v_in, w_in, pi_in, th_in, res and tth (any may be NULL
 * for zero) at the saved state (u .. eta), the cotangents gu .. geta of the
 * state.  Gather form, deterministic.  work:
 * sisl_explicit_adj_workspace_f64(g) elements. */
int64_t sisl_explicit_adj_workspace_f64(const SislGrid *g);
int sisl_explicit_adj_f64(const SislTendency *t, const double *u, const double *v, const double *w,
                          const double *pi, const double *th, const double *eta, const double *gu_in,
                          const double *gv_in, const double *gw_in, const double *gpi_in,
                          const double *gth_in, const double *gres, const double *gtth, double *gu,
                          double *gv, double *gw, double *gpi, double *gth, double *geta, double *work,
                          void *stream);
/* One predictor-corrector pass (steppers.py:911-968): theta at n+1 from the
 * advected theta', its tendency and the guess's w; then the packed
 * right-hand side b = u | v | w | pi 1000 | eta of the implicit solve from
 * the advected fields (rhs_u / rhs_v with their lateral faces already set),
 * the raw advected kinematic residual reta, the guess (g_*: physical units)
 * and the winds of the metric and kinematic-bottom terms (m_u, m_v). */
int sisl_predictor_f64(const SislTendency *t, const double *rhs_u, const double *rhs_v,
                       const double *rhs_w, const double *rhs_pi, const double *th_adv,
                       const double *tth, const double *w_n, const double *reta, const double *g_pi,
                       const double *g_u, const double *g_v, const double *g_w, const double *g_eta,
                       const double *m_u, const double *m_v, double *th_next, double *b, void *stream);
/* Divergence damping of the solved winds and the solution's unscaling
 * (steppers.py:1003-1035): from the packed solution x, u and v after
 * damping, pi = x_pi / 1000 + pi_bg and eta = x_eta / dz_w. */
int sisl_final_f64(const SislTendency *t, const double *x, double *u, double *v, double *pi, double *eta,
                   void *stream);
/* Reverse mode of sisl_predictor_f64 (steppers.py:911-968 through the
 * reference's autodiff, stop-gradient edge values as in
 * sisl_explicit_adj_f64; the kinematic bottom row passes nothing back): from
 * the cotangents gb of b and gth of th_next (NULL for zero) at the saved
 * guess, metric winds and th_next, the cotangents of the fifteen inputs in
 * sisl_predictor_f64's order.  work: sisl_predictor_adj_workspace_f64(g)
 * elements. */
int64_t sisl_predictor_adj_workspace_f64(const SislGrid *g);
int sisl_predictor_adj_f64(const SislTendency *t, const double *g_pi, const double *g_u,
                           const double *g_v, const double *g_eta, const double *m_u, const double *m_v,
                           const double *th_next, const double *gb, const double *gth, double *grhs_u,
                           double *grhs_v, double *grhs_w, double *grhs_pi, double *gth_adv,
                           double *gtth, double *gw_n, double *greta, double *gg_pi, double *gg_u,
                           double *gg_v, double *gg_w, double *gg_eta, double *gm_u, double *gm_v,
                           double *work, void *stream);
/* Reverse mode of sisl_final_f64: the cotangent gx of the packed solution
 * from those of u, v, pi and eta (NULL for zero; the solution's w is not an
 * output of the group, its slot is zero).  work: N elements. */
int sisl_final_adj_f64(const SislTendency *t, const double *gu, const double *gv, const double *gpi,
                       const double *geta, double *gx, double *work, void *stream);

/* SplitExplicitStepper3D._compute_slow_tendencies (suetes/regional3d/
 * steppers.py:1358-1470) with _upwind_flux_3rd_order (steppers.py:1310-1356)
 * on the stage state (u, v, w, pi, th, eta): the explicit tendencies as in
 * sisl_explicit_f64 with the theta' Exner flux, plus -nu_h del^4 of the stage
 * (lap2s) and of the step's start state (lap20) -- each an output of
 * sisl_hyperdiffusion_f64, NULL to omit, t->nu_h their coefficient -- plus
 * the advection of theta', u_m, v_m and w_m averaged to their points.
 * eps_l: the flux magnitude below which the upwind bias is off (1e-12).
 * adv: 4 N elements of work; tend: tu | tv | tw | tpi | tth
 * (N_u + N_v + N_w + 2 N elements). */
int sisl_slow_tendencies_f64(const SislTendency *t, const double *u, const double *v, const double *w,
                             const double *pi, const double *th, const double *eta, const double *lap2s,
                             const double *lap20, double eps_l, double *adv, double *tend, void *stream);

/* ---- split-explicit core: acoustic substeps (FP64) -------------------- */

/*
 * Static fields of the acoustic substeps of SplitExplicitStepper3D
 * (suetes/regional3d/steppers.py:1472-1737), precomputed once per stepper
 * from the reference background; layouts as in SislTendency.
 */
typedef struct SislAcoustic {
  SislGrid g;
  double dx, dy, dz;      /* grid spacings                                    */
  double cp, alpha;       /* constants['cp'], acoustic off-centring            */
  double nu_div;          /* divergence-damping coefficient                   */
  const double *mu, *mv;  /* [u2] [v2] map factors at u / v faces             */
  const double *mm, *mw;  /* [m2] map factors at mass points / w columns      */
  const double *thu, *thv, *thw; /* [u] [v] [w] background theta averages     */
  const double *zxu, *zev;  /* [u] [v] terrain slopes (differences of Z_m)    */
  const double *rz;         /* [w] dz / dz_w_full                             */
  const double *dzm;        /* [m] dz_m_full                                  */
  const double *cpi;        /* [m] C_pi                                       */
  const double *cx, *cy;    /* [u] [v] rho th dz / m (horizontal Exner flux)  */
  const double *rtw;        /* [w] rho_w th_w (vertical Exner flux)           */
  const double *zxiw, *zetaw; /* [w] z_xi_w, z_eta_w                          */
  const double *tau;        /* [w] Rayleigh damping rate                      */
} SislAcoustic;

/* Elements of the work buffer of sisl_acoustic_loop_f64 (-1: bad grid). */
int64_t sisl_acoustic_workspace_f64(const SislGrid *g);
/* SplitExplicitStepper3D._acoustic_loop (steppers.py:1472-1651) with
 * _solve_vertical_acoustic_column (steppers.py:1653-1737): nsteps substeps
 * of length dtau on the deviation (u, v, w, pi) = (u'', v'', w'', pi'') from
 * the stage state (uc, vc, wc), updated in place.  su, sv, sw, sp: slow
 * tendencies of the stage; ub, vb: the deviation's values at the start of
 * the loop (its lateral boundary faces); lower, cprime, invden [m]: Thomas
 * factors of the column matrix for this dtau (sub-diagonal, c'_k and
 * 1 / (main_k - lower_k c'_{k-1})).  work: sisl_acoustic_workspace_f64(g)
 * elements.  Needs nx, ny, nz >= 2. */
int sisl_acoustic_loop_f64(const SislAcoustic *a, int32_t nsteps, double dtau, const double *lower,
                           const double *cprime, const double *invden, double *u, double *v, double *w,
                           double *pi, const double *uc, const double *vc, const double *wc, const double *su,
                           const double *sv, const double *sw, const double *sp, const double *ub,
                           const double *vb, double *work, void *stream);

/* ---- flux-form semi-Lagrangian transport ------------------------------- */

/* FluxFormAdvector.advect_1d (suetes/regional3d/steppers.py:552-594) along
 * one axis of a contiguous (outer, n, inner) array: mass (outer, n, inner),
 * face Courant numbers cfl (outer, n + 1, inner), out like mass; periodic =
 * 1 wraps the line.  scratch: outer * (n + 1) * inner elements. */
int sisl_ffsl_1d_f64(const double *mass, const double *cfl, double *out, double *scratch, int64_t outer,
                     int32_t n, int64_t inner, int32_t periodic, void *stream);
int sisl_ffsl_1d_f32(const float *mass, const float *cfl, float *out, float *scratch, int64_t outer,
                     int32_t n, int64_t inner, int32_t periodic, void *stream);
/* Transpose: cotangents of mass and of cfl from the cotangent of out (the
 * clamp of a closed line passes half a cotangent at its bounds, as the
 * tensor operators' maximum / minimum).  scratch: 2 * outer * (n+1) * inner. */
int sisl_ffsl_1d_T_f64(const double *mass, const double *cfl, const double *gout, double *gmass,
                       double *gcfl, double *scratch, int64_t outer, int32_t n, int64_t inner,
                       int32_t periodic, void *stream);
int sisl_ffsl_1d_T_f32(const float *mass, const float *cfl, const float *gout, float *gmass, float *gcfl,
                       float *scratch, int64_t outer, int32_t n, int64_t inner, int32_t periodic,
                       void *stream);

/* Allocation-free variant of sisl_ffsl_1d_f64.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The cumulative-mass scratch is a fixed device array inside
 * the library (64 MiB per device, held until the process ends):
 * outer * (n + 1) * inner must not exceed 8,388,608 elements, else
 * SISL_ERR_TOO_LARGE.  Calls that overlap on one device (e.g. on two
 * streams) share that scratch and race; serialise them. */
int sisl_ffsl_1d_f64_noalloc(const double *mass, const double *cfl, double *out, int64_t outer, int32_t n,
                             int64_t inner, int32_t periodic, void *stream);

/* ---- launch graphs ------------------------------------------------------ */

/* on = 1 (default): the mixed solver captures each inner Arnoldi cycle into
 * a hipGraph once (per workspace, operator, restart length, direction and
 * coefficient storage) and replays it for every later cycle, and so does the
 * batched solver (per workspace layout, direction and Gram-Schmidt form);
 * on = 0: eager launches (bit-identical results).  No counterpart in the
 * reference. */
int sisl_set_graphs(int32_t on);

/* ---- kernel timing (measurement runs) ------------------------------------ */

/* Timed kernels: operator apply, column preconditioner, the Krylov
 * projection / update / normalisation passes, the tricubic gather, the
 * transposes of the operator, the preconditioner and the gather, and the
 * operator and its transpose applied to a stored FP16 basis column. */
enum SislTimedKernel {
  SISL_T_HELMHOLTZ = 0, SISL_T_PRECOND = 1, SISL_T_KRYLOV_DOTS = 2, SISL_T_KRYLOV_UPDATE = 3,
  SISL_T_KRYLOV_NORMALIZE = 4, SISL_T_TRICUBIC = 5, SISL_T_HELMHOLTZ_T = 6, SISL_T_PRECOND_T = 7,
  SISL_T_TRICUBIC_ADJ = 8, SISL_T_HELMHOLTZ_H16 = 9, SISL_T_HELMHOLTZ_T_H16 = 10, SISL_T_COUNT = 11
};
/* on = 1: clear the record and bracket every later launch of a timed kernel
 * with HIP events on its stream; on = 0: stop recording (the record stays).
 * The first enable allocates the record's device flags (4 MiB, kept for the
 * process): a measurement facility, not for allocation-free callers. */
int sisl_timer_enable(int32_t on);
/* Totals of one kernel (fp32 = 1: FP32 working-type variant) since the last
 * enable: event time of all launches, number of launches, launches that did
 * work (not retired by a device stop flag), algorithmic bytes of those. */
int sisl_timer_read(int32_t kernel, int32_t fp32, double *total_ms, int64_t *launches,
                    int64_t *working_launches, double *bytes);

/* Bandwidth probes for the roofline (measurement only, no reference
 * counterpart).  sisl_probe_stream: mode 0 copies `bytes` (a multiple of 16)
 * from in to out with 16-byte accesses, mode 1 only reads them (out: 4
 * bytes).  sisl_probe_operator_streams_f32: the memory traffic of the inner
 * solver's operator launch (the FP32 operator with bfloat16 3-D
 * coefficients and the fused FP16 basis column) without its stencil or
 * arithmetic: each input array read once at the thread's own column, the
 * same outputs, tiles (64 x 4), level chunk and tile order.
 * sisl_probe_operator_streams_h16_f32: the same for the launch on a stored
 * FP16 basis column (x16: FP16 input arrays; no basis outputs). */
int sisl_probe_stream(const void *in, void *out, int64_t bytes, int32_t mode, void *stream);
int sisl_probe_operator_streams_f32(const SislOperator *op, const float *x, float *y, void *basis,
                                    int32_t chunk, void *stream);
int sisl_probe_operator_streams_h16_f32(const SislOperator *op, const void *x16, float *y, int32_t chunk,
                                        void *stream);

/* ---- batched ensemble (one launch per kernel over members) --------------- */

/* At most 32 members per batched call, each on its own grid (uneven sizes). */

/* tensor_product_interp_3d of one field per member: member m has field dims
 * dims[3m..3m+2] = (fnx, fny, fnz) and its own field / offsets / output. */
int sisl_tricubic_batched_f64(int32_t nmembers, const int32_t *dims, const double *const *field,
                              const double *const *coords, double *const *out, int32_t use_limiter,
                              void *stream);
int sisl_tricubic_batched_f32(int32_t nmembers, const int32_t *dims, const float *const *field,
                              const float *const *coords, float *const *out, int32_t use_limiter,
                              void *stream);
/* Departure points at staggering loc of every member (grids[m], dx[m], dy[m]). */
int sisl_departure_batched_f64(int32_t nmembers, const SislGrid *grids, int32_t loc,
                               const double *const *u, const double *const *v,
                               const double *const *eta, const double *const *m2, const double *dx,
                               const double *dy, double dt, int32_t iterations,
                               double *const *scratch3, double *const *coords_out, void *stream);
int sisl_departure_batched_f32(int32_t nmembers, const SislGrid *grids, int32_t loc,
                               const float *const *u, const float *const *v,
                               const float *const *eta, const float *const *m2, const double *dx,
                               const double *dy, double dt, int32_t iterations,
                               float *const *scratch3, float *const *coords_out, void *stream);
/* Mixed-precision GMRES of every member's system (as sisl_gmres_mixed with
 * the FP16 basis), all members in lockstep: every solver pass is one launch;
 * the refinement loop and the Hessenberg least squares run on the device.
 * Workspace: caller-owned, sized by _workspace_bytes; _setup records the
 * members' operators in it (once per operator set); then any number of
 * solves (p->transpose selects A^T / M^T). x: in = x0, out = solution. */
int64_t sisl_gmres_batched_workspace_bytes(const SislGrid *grids, int32_t nmembers, int32_t restart);
int sisl_gmres_batched_setup(const SislOperator *ops64, const SislOperator *ops32, int32_t nmembers,
                             int32_t restart, void *workspace, void *stream);
int sisl_gmres_batched(int32_t nmembers, const SislGmresParams *p, const double *const *b,
                       double *const *x, void *workspace, SislGmresStats *stats, void *stream);
/* Under a stream capture of the caller (a graph-captured time step)
 * sisl_gmres_batched reads nothing back: it records every outer iteration
 * the refinement could run (max_outer + 1), of which those after a member
 * has converged do nothing on the device, so a replay gives the same bits
 * as the eager solve; its stats are then set to cycles = -1.  This reads
 * the members' outcomes of the last solve on the workspace (e.g. after a
 * replay).  No counterpart in the reference. */
int sisl_gmres_batched_outcome(int32_t nmembers, void *workspace, SislGmresStats *stats, void *stream);

/* Packed-vector helpers used by the host layer. out = sum over d of x*y (FP64). */
int sisl_dot_f64(const double *x, const double *y, int64_t n, double *out_dev, double *partials,
                 void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SUETES_SISL_H */
