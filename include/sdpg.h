/*
 * sdpg.h — C-ABI of the MI355X Newton engine for SDPA-GMP.
 *
 * Drop-in boundary: the reference has no plugin/FFI; its boundary is the C++
 * class Newton (sdpa_newton.h:49-168), constructed once (sdpa_main.cpp:214),
 * given a dense Schur matrix (sdpa_main.cpp:221) and the formula table
 * (sdpa_main.cpp:225), then called as Newton::Mehrotra twice per IPM
 * iteration (sdpa_main.cpp:309, 342).  These entry points are what a host
 * Newton::Mehrotra calls, in the reference order:
 *   Make_gVec -> [PREDICTOR] Make_bMat -> compute_DyVec ([PREDICTOR] Cholesky
 *   + solve) -> compute_DzMat -> compute_DxMat   (sdpa_newton.cpp:1375-1414).
 *
 * Conventions: every function returns int; 0 = success, > 0 = a LAPACK-style
 * info (only sdpg_potrf: B not positive definite, which the reference maps to
 * FAILURE, sdpa_linear.cpp:622-630), < 0 = an API error (SDPG_E*).  All
 * buffers are caller-owned host memory.  A context owns its device buffers
 * and one HIP stream; calls on one context must not overlap.
 *
 * Multiple-precision exchange format ("packed mp"): L+2 32-bit words per
 * element, L = sdpg_limbs(prec_bits):
 *   word 0  int32 binary exponent e,  word 1  sign (0 = +, 1 = -),
 *   words 2..L+1  mantissa limbs, least significant first,
 *   value = (-1)^sign * M * 2^(e - 32L), bit 32L-1 of M set; zero: M = 0,
 *   e = SDPG_EZERO, sign = 0.
 * A "dense linear space" (X, Z, dX, ...) is the concatenation of the SDP
 * blocks, each n_l x n_l column-major (sdpa_struct.h:111-133 DenseMatrix
 * de_ele[i + n*j]), followed by the LP diagonal entries.
 */
#ifndef SDPG_H
#define SDPG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDPG_OK 0
#define SDPG_EARG (-1)
#define SDPG_EHIP (-2)
#define SDPG_ESTATE (-3)
#define SDPG_ENOMEM (-4)
#define SDPG_EPREC (-5)
#define SDPG_ENODEV (-6)

#define SDPG_EZERO (-(1 << 29))

/* sdpg_get / sdpg_put selectors */
#define SDPG_X 1
#define SDPG_Z 2
#define SDPG_INVZ 3
#define SDPG_Y 4
#define SDPG_RP 5      /* primal residual vector b - A.X            (m)   */
#define SDPG_RD 6      /* dual residual matrix C - Z - sum y_i A_i  (dls) */
#define SDPG_GVEC 7    /* right-hand side g                         (m)   */
#define SDPG_B 8       /* Schur matrix before factorisation         (m*m) */
#define SDPG_L 9       /* Cholesky factor, lower triangle valid     (m*m) */
#define SDPG_DY 10     /* (m)   */
#define SDPG_DZ 11     /* (dls) */
#define SDPG_DX 12     /* (dls) */
#define SDPG_RZINV 13  /* r_zinv of the last sdpg_gvec call (dls) */
#define SDPG_INVCHOLX 14 /* inv(chol(X)) of the last sdpg_factor_point, lower (dls) */
#define SDPG_INVCHOLZ 15 /* inv(chol(Z)) of the last sdpg_factor_point, lower (dls) */
#define SDPG_X0 16     /* initial point of an initial-point file (-id / -is), put once (dls) */
#define SDPG_Z0 17

/* sdpg_inner selectors: scalar inner products over the dense linear space / vectors */
#define SDPG_IP_XZ 1    /* X . Z   */
#define SDPG_IP_XDZ 2   /* X . dZ  */
#define SDPG_IP_DXZ 3   /* dX . Z  */
#define SDPG_IP_DXDZ 4  /* dX . dZ */
#define SDPG_IP_CX 5    /* C . X   */
#define SDPG_IP_CDX 6   /* C . dX  */
#define SDPG_IP_BY 7    /* b . y   */
#define SDPG_IP_BDY 8   /* b . dy  */
#define SDPG_IP_TRX 9   /* trace X (SDP diagonals and LP entries) */
#define SDPG_IP_TRZ 10  /* trace Z */
#define SDPG_IP_X0Z 11  /* X0 . Z  (SolveInfo::update with an initial point, sdpa_parts.cpp:747-796) */
#define SDPG_IP_XZ0 12  /* X . Z0  */

/* GEMM forms (sdpg_set_gemm) */
#define SDPG_GEMM_PLAIN 0  /* multi-limb integer MAD on the vector ALUs */
#define SDPG_GEMM_OZAKI 1  /* INT8 slices on MFMA, exact INT32 accumulation */
#define SDPG_GEMM_ELEMENT (-1) /* one output per thread, floating limb MAC per term */

/* phase codes (SolveInfo::phaseType, sdpa_parts.h) */
#define SDPG_NOINFO 0
#define SDPG_PFEAS 1
#define SDPG_DFEAS 2
#define SDPG_PDFEAS 3

/* limbs L for a requested precision in bits (GMP mpf rounding of precision) */
int sdpg_limbs(int prec_bits);

typedef struct sdpg_ctx sdpg_ctx;

/*
 * Newton::Newton + Newton::initialize (sdpa_newton.cpp:61-120) and
 * initialize_dense_bMat (sdpa_newton.cpp:218-222): m constraints, n_sdp SDP
 * blocks of the given sizes, n_lp LP variables (sdpa_main.cpp:114-131 block
 * typing).  Device = HIP device ordinal.
 */
int sdpg_ctx_create(sdpg_ctx **ctx, int prec_bits, int m, int n_sdp, const int *sdp_sizes, int n_lp, int device);
void sdpg_ctx_destroy(sdpg_ctx *ctx);

/*
 * Input data (InputData::A, C; sdpa_io.cpp:626-751 setElement).  con = -1 for
 * C (already negated as in sdpa_io.cpp:661), 0..m-1 for A_con.  Entries are
 * 0-based (row, col) as read; dense != 0 means the block was switched to
 * dense storage (SparseMatrix::changeToDense, sdpa_struct.cpp:484-525) and val
 * holds n*n column-major values (row/col ignored).  Each (con, block) may be
 * set once.
 */
int sdpg_set_block(sdpg_ctx *ctx, int con, int sdp_block, int nnz, const int *row, const int *col, const uint32_t *val,
                   int dense);
int sdpg_set_lp(sdpg_ctx *ctx, int con, int nnz, const int *pos, const uint32_t *val);
int sdpg_set_b(sdpg_ctx *ctx, const uint32_t *b);
/*
 * Formula table of Newton::computeFormula_SDP (sdpa_newton.cpp:694-819):
 * use_formula[i*n_sdp + l] in {1 = F1, 2 = F2, 3 = F3}, 0 where A_i has no
 * block l.
 */
int sdpg_set_formula(sdpg_ctx *ctx, const int *use_formula);
/* freeze the input data: build the pair lists (sdpa_newton.cpp:1052-1054 rule) and upload */
int sdpg_commit(sdpg_ctx *ctx);

/*
 * GEMM form used by the Schur precompute, the Cholesky panel / trailing updates
 * and dX: SDPG_GEMM_PLAIN (limb MADs on the vector ALUs), SDPG_GEMM_OZAKI (INT8
 * slices on MFMA), SDPG_GEMM_ELEMENT (one output per thread, the Rgemm order).
 * slices (Ozaki only): 0 = default 4L-4, or a multiple of 4 in [max(8, 4L-12), 4L] (8S-2 fraction
 * bits per row / column of the block fixed point); other counts return SDPG_EARG.  sdpg_tune_slices
 * picks the count for a requested precision.
 */
int sdpg_set_gemm(sdpg_ctx *ctx, int mode, int slices);

/* current point (Solutions xMat, zMat, invzMat, yVec) */
int sdpg_set_point(sdpg_ctx *ctx, const uint32_t *X, const uint32_t *Z, const uint32_t *invZ, const uint32_t *y);
/*
 * Residuals::compute (sdpa_dataset.cpp:625-646) and the objective inner
 * products of SolveInfo (sdpa_parts.cpp:747-760): Rp = b - A.X,
 * Rd = C - Z - sum y_i A_i, their max-abs norms, objP = C.X, objD = b.y.
 * Any output pointer may be NULL.
 */
int sdpg_residuals(sdpg_ctx *ctx, uint32_t *normRp, uint32_t *normRd, uint32_t *objP, uint32_t *objD);
/* Newton::Make_gVec incl. compute_rMat (sdpa_newton.cpp:821-895); corrector != 0 for CORRECTOR */
int sdpg_gvec(sdpg_ctx *ctx, int corrector, const uint32_t *mu, const uint32_t *beta, int phase);
/* Newton::Make_bMat, dense SDP + LP parts (sdpa_newton.cpp:1007-1096, 1190-1225, 1251-1272) */
int sdpg_schur(sdpg_ctx *ctx);
/* Lal::choleskyFactorWithAdjust -> Rpotrf("Lower") (sdpa_linear.cpp:608-631); *info > 0: not PD */
int sdpg_potrf(sdpg_ctx *ctx, int *info);
/* Lal::solveSystems: Rtrsv(L,N) then Rtrsv(L,T) on g (sdpa_linear.cpp:683-695) -> dy */
int sdpg_potrs(sdpg_ctx *ctx);
/* Newton::compute_DzMat (sdpa_newton.cpp:1346-1356) */
int sdpg_dz(sdpg_ctx *ctx, int phase);
/* Newton::compute_DxMat incl. getSymmetrize (sdpa_newton.cpp:1358-1373) */
int sdpg_dx(sdpg_ctx *ctx);

/*
 * Solutions::update (sdpa_dataset.cpp:120-133): X += alphaP dX, y += alphaD dy,
 * Z += alphaD dZ on the device (the step lengths are packed mp scalars).
 */
int sdpg_update(sdpg_ctx *ctx, const uint32_t *alphaP, const uint32_t *alphaD);
/*
 * Solutions::computeInverse (sdpa_dataset.cpp:96-118): invCholX = inv(chol(X)),
 * invCholZ = inv(chol(Z)), invZ = invCholZ^T invCholZ (Jal::getInvChol /
 * getInvCholAndInv, sdpa_jordan.cpp:123-171), blocked Cholesky and recursive
 * triangular inverse on the device.  *info = 1 when X or Z is not positive
 * definite (the reference's "cannot move").
 */
int sdpg_factor_point(sdpg_ctx *ctx, int *info);
/* scalar inner products of StepLength / DirectionParameter / SolveInfo (sdpa_parts.cpp:362-860) */
int sdpg_inner(sdpg_ctx *ctx, int what, uint32_t *out);
/*
 * Jal::getMinEigen(lMat, xMat) (sdpa_jordan.cpp:55-88): smallest eigenvalue of
 * lMat xMat lMat^T over all blocks, which 0 = (invCholX, dX), 1 = (invCholZ, dZ).
 * Blocks > 20: Lanczos (Lal::getMinEigen, sdpa_linear.cpp:27-124) with device
 * matrix-vector products; blocks <= 20: the product on the device, the
 * symmetric eigenproblem (Rsyev there) on the host.
 */
int sdpg_min_eigen(sdpg_ctx *ctx, int which, uint32_t *out);

/* copy a device quantity out / in (packed mp); sizes per selector comment */
int sdpg_get(sdpg_ctx *ctx, int what, uint32_t *out);
int sdpg_put(sdpg_ctx *ctx, int what, const uint32_t *in);

/*
 * accumulated device milliseconds: [0] schur, [1] potrf, [2] potrs, [3] gvec,
 * [4] dz+dx, [5] residuals, [6] factor_point, [7] min_eigen
 */
int sdpg_timers(sdpg_ctx *ctx, double *ms, int n);
int sdpg_reset_timers(sdpg_ctx *ctx);
int sdpg_sync(sdpg_ctx *ctx);
/*
 * Per-kernel device time (no reference counterpart; measurement support):
 * on = 1 clears and starts recording events around each launch of the product
 * and Cholesky block-column kernels, on = 0 stops.  out (may be NULL) receives, for each
 * kind k with 3k + 2 < n (n = 12 for all four) = 0 Ozaki MFMA product (short K),
 * 1 plain limb product, 2 Cholesky block column (diagonal block and panel rows),
 * 3 Ozaki MFMA product (long K):
 * out[3k] launches, out[3k+1] milliseconds, out[3k+2] algorithmic work
 * (k = 0, 3: INT8 operations; 1, 2: mp multiply-adds).  The events sit in the
 * engine's stream between the kernels and lengthen the wall time a little.
 */
int sdpg_profile(sdpg_ctx *ctx, int on, double *out, int n);
/*
 * Rgemm (mplapack/Rgemm_omp.cpp:41-127 -> Rgemm_NN_omp / Rgemm_NT_omp) on host
 * buffers, one call = one device round trip: C = op(A) op(B) (op_c 0),
 * C - op(A) op(B) (1) or C + op(A) op(B) (2); column-major packed mp, op(X) = X^T
 * when trans != 0.  lower != 0 writes only C[i, j] with i >= j (the Rsyrk half
 * used by the Cholesky update, mplapack/Rsyrk.cpp:157-177).  mode / slices as
 * sdpg_set_gemm; *ms (may be NULL) receives the device time of the product.
 */
int sdpg_rgemm(int prec_bits, int device, int transa, int transb, int m, int n, int k, const uint32_t *A, int lda,
               const uint32_t *B, int ldb, uint32_t *C, int ldc, int op_c, int mode, int slices, int lower,
               double *ms);
/*
 * No counterpart in the reference: sdpg_rgemm for callers that cannot allocate device memory
 * (code under HIP graph capture of the surrounding work, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as sdpg_rgemm (the same kernels run), but the operand
 * copies, the staging buffer and the product workspace are carved from a static device array of
 * 128 MiB per device, held from the first use until the process ends; the call makes no hipMalloc,
 * hipFree, stream or event, and runs on the null stream.  Sizes whose operands and workspace do not
 * fit in the array return SDPG_EARG (before any device work; C is untouched).  Calls on one device
 * share its array: the library serializes them with a per-device lock, so calls that overlap from
 * several host threads wait for one another.  *ms (may be NULL) receives the host-measured
 * milliseconds of the product.
 */
int sdpg_rgemm_noalloc(int prec_bits, int device, int transa, int transb, int m, int n, int k, const uint32_t *A,
                       int lda, const uint32_t *B, int ldb, uint32_t *C, int ldc, int op, int mode, int slices,
                       int lower, double *ms);
/*
 * Slice-count autotuner of the Ozaki form (no reference counterpart: the reference's precision is a
 * run parameter, sdpa_parts.cpp:199-200, and its GEMMs run at the full mpf width).  For the limb
 * count of prec_bits it measures every slice count built (multiples of 4 in [max(8, 4L-12), 4L]):
 * effective bits of the product on calibration tiles against a product at L+2 limbs, and device time
 * on two representative products; *slices receives the fastest count whose effective bits reach
 * target_bits + 8 (target_bits 0 = prec_bits).  table (may be NULL) receives 4 doubles per count:
 * slices, effective bits, ms of a 512^3 product, ms of a 1024 x 1024 x 32 lower update.  Cached per
 * (device, limb count, target).  SDPG_EPREC when no count reaches the target or the reference limb
 * count is not built.
 */
int sdpg_tune_slices(int prec_bits, int device, int target_bits, int *slices, double *table, int ntable);

/*
 * Measurement support (no reference counterpart): the sustained device rate of 32 x 32 -> 64-bit limb
 * multiply-adds with carry (v_mad_u64_u32 + v_addc_co_u32), the integer-VALU peak that the plain limb
 * kernels are quoted against; about 0.2 s of device time.
 */
int sdpg_valu_peak(int device, double *limb_mac_per_s);

/*
 * Batched Newton hot path over many small problems (no single reference counterpart: the reference
 * runs one problem per process; each problem's step is Newton::Mehrotra's Make_bMat ->
 * choleskyFactorWithAdjust -> solveSystems, sdpa_newton.cpp:1375-1414, with sdpa_linear.cpp:608-631 and
 * 683-695).  A batch holds n committed contexts of one precision on one device whose constraint blocks
 * are all dense and that have no LP block (at most 4 SDP blocks each); the contexts stay usable on
 * their own.  sdpg_batch_newton takes one request per context: SDPG_BATCH_NONE, SDPG_BATCH_PREDICTOR
 * (Schur build, Cholesky, solve: what sdpg_schur, sdpg_potrf, sdpg_potrs do) or SDPG_BATCH_CORRECTOR
 * (solve only, with the factor of that context's last batched predictor), after the context's
 * sdpg_gvec; it writes dy of every requested context and info[p] (Rpotrf's info for a predictor, 0
 * otherwise; a context with info > 0 gets no dy).  Follow with each context's sdpg_dz / sdpg_dx.  The
 * batched factor replaces the context's own: after a batched predictor, the context's sdpg_potrs is
 * not valid until its next sdpg_potrf.  With sdpg_batch_set_graph(b, 1) the three phase kernels run as
 * one HIP graph, captured on the first call.  sdpg_batch_time gives the device time of the last
 * call's phases in ms.  Calls on one batch are not thread-safe.
 */
typedef struct sdpg_batch sdpg_batch;
#define SDPG_BATCH_NONE 0
#define SDPG_BATCH_PREDICTOR 1
#define SDPG_BATCH_CORRECTOR 2
int sdpg_batch_create(sdpg_batch **batch, int n, sdpg_ctx *const *ctxs);
void sdpg_batch_destroy(sdpg_batch *batch);
int sdpg_batch_newton(sdpg_batch *batch, const int *req, int *info);
int sdpg_batch_set_graph(sdpg_batch *batch, int on);
int sdpg_batch_time(sdpg_batch *batch, double *ms);
/*
 * The other phases of an IPM iteration, batched the same way (one request per context, nonzero = run
 * it for that context), each the counterpart of the per-context call named:
 *   sdpg_batch_gvec        sdpg_gvec; req 1 predictor, 2 corrector; mu, beta: n packed scalars;
 *                          phase: n phase values (Newton::Make_gVec, sdpa_newton.cpp:821-895)
 *   sdpg_batch_dz / _dx    sdpg_dz / sdpg_dx (sdpa_newton.cpp:1346-1373)
 *   sdpg_batch_update      sdpg_update with n packed alphaP, alphaD (sdpa_dataset.cpp:120-133)
 *   sdpg_batch_factor_point sdpg_factor_point; info: n ints (sdpa_dataset.cpp:96-118)
 *   sdpg_batch_inner       sdpg_inner for the selectors SDPG_IP_XZ .. SDPG_IP_TRZ at once: out holds
 *                          10 packed values per context, selector s at index s - 1
 *   sdpg_batch_residuals   sdpg_residuals: out holds 4 packed values per context (max |Rp|, max |Rd|,
 *                          C.X, b.y) (sdpa_parts.cpp:747-790)
 *   sdpg_batch_min_eigen   sdpg_min_eigen (sdpa_jordan.cpp:55-88): out holds 1 packed value per
 *                          context.  Blocks of at most 20 take the smallest eigenvalue (Rsyev in the
 *                          reference) from a double-precision eigenvector refined by its Rayleigh
 *                          quotient; larger blocks follow the reference's Lanczos process with its
 *                          tridiagonal eigenproblem solved in double and refined the same way.
 * sdpg_batch_wait_contexts(b, 0): the calls no longer wait for the contexts' own streams (the caller
 * guarantees that no context work is pending).  sdpg_batch_timers: accumulated device ms of the batched
 * calls by kind (newton, gvec, dz, dx, update, factor_point, inner, residuals, min_eigen; n <= 9).
 */
int sdpg_batch_gvec(sdpg_batch *batch, const int *req, const uint32_t *mu, const uint32_t *beta, const int *phase);
int sdpg_batch_dz(sdpg_batch *batch, const int *req, const int *phase);
int sdpg_batch_dx(sdpg_batch *batch, const int *req);
int sdpg_batch_update(sdpg_batch *batch, const int *req, const uint32_t *alphaP, const uint32_t *alphaD);
int sdpg_batch_factor_point(sdpg_batch *batch, const int *req, int *info);
int sdpg_batch_inner(sdpg_batch *batch, const int *req, uint32_t *out);
int sdpg_batch_residuals(sdpg_batch *batch, const int *req, uint32_t *out);
int sdpg_batch_min_eigen(sdpg_batch *batch, const int *req, int which, uint32_t *out);
int sdpg_batch_wait_contexts(sdpg_batch *batch, int on);
int sdpg_batch_timers(sdpg_batch *batch, double *ms, int n);

/* last error message of this thread */
const char *sdpg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
