/*
 * orc — CPU oracle: C restatement of the SDPA-GMP 7.1.3 dense-Schur solver.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Each routine names the reference file:line it follows.  The mplapack
 * kernels (Rgemm, Rpotrf, Rpotrf2, Rsyrk, Rtrsm, Rtrsv, Rtrmv, Rtrmm, Rgemv,
 * Rdot, Raxpy, Rscal) are restated from the reference-BLAS/LAPACK loop
 * orders used by the mplapack sources.  Two eigen routines are replaced by
 * textbook methods of equal accuracy (their outputs are mathematically
 * determined, only rounding differs):
 *   Rsyev("N","L") (used for blocks <= 20, sdpa_linear.cpp:126-158) -> cyclic
 *   Jacobi rotations on the lower triangle;
 *   Rsteqr("I") (tridiagonal eigenpairs inside the Lanczos step length,
 *   sdpa_linear.cpp:84) -> implicit symmetric QL with Wilkinson shifts.
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef ompf F;

static double now_s(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) {
        fprintf(stderr, "orc: out of memory\n");
        exit(2);
    }
    return p;
}

/* scalar helpers */
static void Fd(F *r, double x) { ompf_set_d(r, x); }
static F Fv(double x) {
    F r;
    ompf_set_d(&r, x);
    return r;
}
static int Flt(const F *a, const F *b) { return ompf_cmp(a, b) < 0; }
static int Fgt(const F *a, const F *b) { return ompf_cmp(a, b) > 0; }

/* ================= data structures (sdpa_struct.h:72-197) ================= */

typedef struct {
    int blk, n, dense, nnz, cap, effect;
    int *ri, *ci;
    F *v;  /* COO values (sparse) */
    F *de; /* n*n column-major (dense) */
} SpBlk;

typedef struct {
    int nsdp;
    SpBlk *s;
    int nlp;
    int *lpi;
    F *lpv;
} SLS;

typedef struct {
    int nsdp;
    int *n;
    F **b;
    int nlp;
    F *lp;
} DLS;

typedef struct {
    int m, nBlock, nsdp, nlp, nDim;
    int *blockStruct, *blockType, *blockNumber, *sdpStruct;
    F *b;
    SLS C;
    SLS *A;
    int *sdpNCon, **sdpCon, **sdpBIdx;
    int *lpNCon, **lpCon, **lpBIdx;
} Prob;

/* ================= dense linear space helpers ================= */

static void dls_init(DLS *d, const Prob *P) {
    d->nsdp = P->nsdp;
    d->n = (int *)xcalloc(P->nsdp, sizeof(int));
    d->b = (F **)xcalloc(P->nsdp, sizeof(F *));
    for (int l = 0; l < P->nsdp; l++) {
        d->n[l] = P->sdpStruct[l];
        d->b[l] = (F *)xcalloc((size_t)d->n[l] * d->n[l], sizeof(F));
    }
    d->nlp = P->nlp;
    d->lp = (F *)xcalloc(P->nlp, sizeof(F));
}
static void dls_free(DLS *d) {
    for (int l = 0; l < d->nsdp; l++) free(d->b[l]);
    free(d->b);
    free(d->n);
    free(d->lp);
}
static void dls_zero(DLS *d) {
    for (int l = 0; l < d->nsdp; l++)
        for (int i = 0; i < d->n[l] * d->n[l]; i++) ompf_set_zero(&d->b[l][i]);
    for (int i = 0; i < d->nlp; i++) ompf_set_zero(&d->lp[i]);
}
static void dls_copy(DLS *r, const DLS *a) {
    for (int l = 0; l < a->nsdp; l++)
        for (int i = 0; i < a->n[l] * a->n[l]; i++) ompf_set(&r->b[l][i], &a->b[l][i]);
    for (int i = 0; i < a->nlp; i++) ompf_set(&r->lp[i], &a->lp[i]);
}
/* DenseMatrix::setIdentity (sdpa_struct.cpp:839) */
static void dls_identity(DLS *d, const F *s) {
    dls_zero(d);
    for (int l = 0; l < d->nsdp; l++)
        for (int i = 0; i < d->n[l]; i++) ompf_set(&d->b[l][i + d->n[l] * i], s);
    for (int i = 0; i < d->nlp; i++) ompf_set(&d->lp[i], s);
}

/* Rscal: x[i] = a*x[i] */
static void Rscal(int n, const F *a, F *x) {
    for (int i = 0; i < n; i++) ompf_mul(&x[i], a, &x[i]);
}
/* Raxpy: y[i] += a*x[i] (mplapack/Raxpy_omp.cpp:55-61) */
static void Raxpy(int n, const F *a, const F *x, int incx, F *y, int incy) {
    F t;
    for (int i = 0; i < n; i++) {
        ompf_mul(&t, a, &x[(size_t)i * incx]);
        ompf_add(&y[(size_t)i * incy], &y[(size_t)i * incy], &t);
    }
}
/* Rdot: sequential sum (mplapack/Rdot_omp.cpp:48-63) */
static F Rdot(int n, const F *x, int incx, const F *y, int incy) {
    F s, t;
    ompf_set_zero(&s);
    for (int i = 0; i < n; i++) {
        ompf_mul(&t, &x[(size_t)i * incx], &y[(size_t)i * incy]);
        ompf_add(&s, &s, &t);
    }
    return s;
}

/* multiply(DLS, DLS, scalar): Rcopy + Rscal (sdpa_linear.cpp:1713) */
static void dls_scale(DLS *r, const DLS *a, const F *s) {
    for (int l = 0; l < a->nsdp; l++) {
        int len = a->n[l] * a->n[l];
        for (int i = 0; i < len; i++) ompf_mul(&r->b[l][i], s, &a->b[l][i]);
    }
    for (int i = 0; i < a->nlp; i++) ompf_mul(&r->lp[i], &a->lp[i], s);
}
/* plus(DLS, DLS, DLS, scalar): r = a + s*b (sdpa_linear.cpp:1755, 1161) */
static void dls_axpy(DLS *r, const DLS *a, const DLS *b, const F *s) {
    F one = Fv(1.0);
    if (!s) s = &one;
    for (int l = 0; l < a->nsdp; l++) {
        int len = a->n[l] * a->n[l];
        if (r != a)
            for (int i = 0; i < len; i++) ompf_set(&r->b[l][i], &a->b[l][i]);
        Raxpy(len, s, b->b[l], 1, r->b[l], 1);
    }
    for (int i = 0; i < a->nlp; i++) {
        F t;
        ompf_mul(&t, &b->lp[i], s);
        ompf_add(&r->lp[i], &a->lp[i], &t);
    }
}
/* getInnerProduct(DLS, DLS) (sdpa_linear.cpp:1630) */
static F dls_inner(const DLS *a, const DLS *b) {
    F r, t;
    ompf_set_zero(&r);
    for (int l = 0; l < a->nsdp; l++) {
        t = Rdot(a->n[l] * a->n[l], a->b[l], 1, b->b[l], 1);
        ompf_add(&r, &r, &t);
    }
    for (int i = 0; i < a->nlp; i++) {
        ompf_mul(&t, &a->lp[i], &b->lp[i]);
        ompf_add(&r, &r, &t);
    }
    return r;
}

/* ================= mplapack restatements ================= */

/* Rgemm_NN_omp (mplapack/Rgemm_NN_omp.cpp:36-74): C = alpha*A*B + beta*C, beta in {0,1} */
static void gemm_nn(int m, int n, int k, const F *alpha, const F *A, int lda, const F *B, int ldb, int beta1, F *C,
                    int ldc) {
    if (!beta1)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) ompf_set_zero(&C[i + (size_t)j * ldc]);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++) {
        F temp, templ;
        for (int l = 0; l < k; l++) {
            ompf_mul(&temp, alpha, &B[l + (size_t)j * ldb]);
            for (int i = 0; i < m; i++) {
                ompf_mul(&templ, &temp, &A[i + (size_t)l * lda]);
                ompf_add(&C[i + (size_t)j * ldc], &C[i + (size_t)j * ldc], &templ);
            }
        }
    }
}
/* Rgemm_NT_omp (mplapack/Rgemm_NT_omp.cpp:36-76, serial): C = alpha*A*B^T + beta*C */
static void gemm_nt(int m, int n, int k, const F *alpha, const F *A, int lda, const F *B, int ldb, int beta1, F *C,
                    int ldc) {
    if (!beta1)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) ompf_set_zero(&C[i + (size_t)j * ldc]);
    F temp, templ;
    for (int j = 0; j < n; j++)
        for (int l = 0; l < k; l++) {
            ompf_mul(&temp, alpha, &B[j + (size_t)l * ldb]);
            for (int i = 0; i < m; i++) {
                ompf_mul(&templ, &temp, &A[i + (size_t)l * lda]);
                ompf_add(&C[i + (size_t)j * ldc], &C[i + (size_t)j * ldc], &templ);
            }
        }
}
/* Rsyrk("L","N") with beta = 1 (mplapack/Rsyrk.cpp:157-177) */
static void syrk_ln(int n, int k, const F *alpha, const F *A, int lda, F *C, int ldc) {
    F temp, t;
    for (int j = 0; j < n; j++)
        for (int l = 0; l < k; l++) {
            if (ompf_sgn(&A[j + (size_t)l * lda]) != 0) {
                ompf_mul(&temp, alpha, &A[j + (size_t)l * lda]);
                for (int i = j; i < n; i++) {
                    ompf_mul(&t, &temp, &A[i + (size_t)l * lda]);
                    ompf_add(&C[i + (size_t)j * ldc], &C[i + (size_t)j * ldc], &t);
                }
            }
        }
}
/* Rtrsm("R","L","T","N", alpha = 1): B := B * inv(A^T) (mplapack/Rtrsm.cpp:263-285) */
static void trsm_rltn(int m, int n, const F *A, int lda, F *B, int ldb) {
    F one = Fv(1.0), temp, t;
    for (int k = 0; k < n; k++) {
        ompf_div(&temp, &one, &A[k + (size_t)k * lda]);
        for (int i = 0; i < m; i++) ompf_mul(&B[i + (size_t)k * ldb], &temp, &B[i + (size_t)k * ldb]);
        for (int j = k + 1; j < n; j++) {
            if (ompf_sgn(&A[j + (size_t)k * lda]) != 0) {
                ompf_set(&temp, &A[j + (size_t)k * lda]);
                for (int i = 0; i < m; i++) {
                    ompf_mul(&t, &temp, &B[i + (size_t)k * ldb]);
                    ompf_sub(&B[i + (size_t)j * ldb], &B[i + (size_t)j * ldb], &t);
                }
            }
        }
    }
}
/* Rtrsm("L","L","N","N", alpha = 1): B := inv(A) * B (reference dtrsm, Left/Lower/NoTrans) */
static void trsm_llnn(int m, int n, const F *A, int lda, F *B, int ldb) {
    F t;
    for (int j = 0; j < n; j++)
        for (int k = 0; k < m; k++) {
            F *bkj = &B[k + (size_t)j * ldb];
            if (ompf_sgn(bkj) != 0) {
                ompf_div(bkj, bkj, &A[k + (size_t)k * lda]);
                for (int i = k + 1; i < m; i++) {
                    ompf_mul(&t, bkj, &A[i + (size_t)k * lda]);
                    ompf_sub(&B[i + (size_t)j * ldb], &B[i + (size_t)j * ldb], &t);
                }
            }
        }
}
/* Rpotrf2 lower, recursive (mplapack/Rpotrf2.cpp:32-151); returns info */
static int potrf2_l(int n, F *a, int lda) {
    if (n == 0) return 0;
    if (n == 1) {
        if (ompf_sgn(&a[0]) <= 0) return 1;
        ompf_sqrt(&a[0], &a[0]);
        return 0;
    }
    int n1 = n / 2, n2 = n - n1;
    int info = potrf2_l(n1, a, lda);
    if (info) return info;
    trsm_rltn(n2, n1, a, lda, &a[n1], lda);
    F mone = Fv(-1.0);
    syrk_ln(n2, n1, &mone, &a[n1], lda, &a[n1 + (size_t)n1 * lda], lda);
    info = potrf2_l(n2, &a[n1 + (size_t)n1 * lda], lda);
    if (info) return info + n1;
    return 0;
}
/* Rpotrf("Lower") blocked, nb = 64 (mplapack/Rpotrf.cpp:32-154, iMlaenv.cpp:222-228) */
static int potrf_l(int n, F *a, int lda) {
    const int nb = 64;
    if (nb <= 1 || nb >= n) return potrf2_l(n, a, lda);
    F mone = Fv(-1.0);
    for (int j = 0; j < n; j += nb) {
        int jb = nb < n - j ? nb : n - j;
        syrk_ln(jb, j, &mone, &a[j], lda, &a[j + (size_t)j * lda], lda);
        int info = potrf2_l(jb, &a[j + (size_t)j * lda], lda);
        if (info) return info + j;
        if (j + jb < n) {
            gemm_nt(n - j - jb, jb, j, &mone, &a[j + jb], lda, &a[j], lda, 1, &a[(j + jb) + (size_t)j * lda], lda);
            trsm_rltn(n - j - jb, jb, &a[j + (size_t)j * lda], lda, &a[(j + jb) + (size_t)j * lda], lda);
        }
    }
    return 0;
}
/* Rtrsv("L","N","NonUnit") (mplapack/Rtrsv.cpp) */
static void trsv_ln(int n, const F *A, int lda, F *x) {
    F t;
    for (int j = 0; j < n; j++) {
        if (ompf_sgn(&x[j]) != 0) {
            ompf_div(&x[j], &x[j], &A[j + (size_t)j * lda]);
            F temp = x[j];
            for (int i = j + 1; i < n; i++) {
                ompf_mul(&t, &temp, &A[i + (size_t)j * lda]);
                ompf_sub(&x[i], &x[i], &t);
            }
        }
    }
}
/* Rtrsv("L","T","NonUnit") */
static void trsv_lt(int n, const F *A, int lda, F *x) {
    F t;
    for (int j = n - 1; j >= 0; j--) {
        F temp = x[j];
        for (int i = n - 1; i > j; i--) {
            ompf_mul(&t, &A[i + (size_t)j * lda], &x[i]);
            ompf_sub(&temp, &temp, &t);
        }
        ompf_div(&temp, &temp, &A[j + (size_t)j * lda]);
        x[j] = temp;
    }
}
/* Rtrmv("L","T","N"): x := A^T x */
static void trmv_lt(int n, const F *A, int lda, F *x) {
    F t;
    for (int j = 0; j < n; j++) {
        F temp = x[j];
        ompf_mul(&temp, &temp, &A[j + (size_t)j * lda]);
        for (int i = j + 1; i < n; i++) {
            ompf_mul(&t, &A[i + (size_t)j * lda], &x[i]);
            ompf_add(&temp, &temp, &t);
        }
        x[j] = temp;
    }
}
/* Rtrmv("L","N","N"): x := A x */
static void trmv_ln(int n, const F *A, int lda, F *x) {
    F t;
    for (int j = n - 1; j >= 0; j--) {
        if (ompf_sgn(&x[j]) != 0) {
            F temp = x[j];
            for (int i = n - 1; i > j; i--) {
                ompf_mul(&t, &temp, &A[i + (size_t)j * lda]);
                ompf_add(&x[i], &x[i], &t);
            }
            ompf_mul(&x[j], &x[j], &A[j + (size_t)j * lda]);
        }
    }
}
/* Rgemv("N"), beta = 0: y = A x */
static void gemv_n(int m, int n, const F *A, int lda, const F *x, F *y) {
    F t;
    for (int i = 0; i < m; i++) ompf_set_zero(&y[i]);
    for (int j = 0; j < n; j++) {
        if (ompf_sgn(&x[j]) != 0) {
            F temp = x[j];
            for (int i = 0; i < m; i++) {
                ompf_mul(&t, &temp, &A[i + (size_t)j * lda]);
                ompf_add(&y[i], &y[i], &t);
            }
        }
    }
}
/* Rtrmm("L","L","T","N", alpha = 1): B := A^T B (reference dtrmm) */
static void trmm_lltn(int m, int n, const F *A, int lda, F *B, int ldb) {
    F t;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            F temp = B[i + (size_t)j * ldb];
            ompf_mul(&temp, &temp, &A[i + (size_t)i * lda]);
            for (int k = i + 1; k < m; k++) {
                ompf_mul(&t, &A[k + (size_t)i * lda], &B[k + (size_t)j * ldb]);
                ompf_add(&temp, &temp, &t);
            }
            B[i + (size_t)j * ldb] = temp;
        }
}

/* ================= eigen routines ================= */

/* smallest eigenvalue of the symmetric matrix whose lower triangle is a (n x n), cyclic Jacobi */
static F min_eig_jacobi(int n, const F *a) {
    F *A = (F *)xcalloc((size_t)n * n, sizeof(F));
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) A[i + (size_t)j * n] = (i >= j) ? a[i + (size_t)j * n] : a[j + (size_t)i * n];
    F one = Fv(1.0), two = Fv(2.0), t1, t2, th, t, c, s, tau, app, aqq, apq, akp, akq;
    F eps;
    ompf_mul_2exp(&eps, &one, -(long)ompf_prec_bits() - 8);
    for (int sweep = 0; sweep < 100; sweep++) {
        /* off-diagonal norm vs diagonal */
        F off, dg;
        ompf_set_zero(&off);
        ompf_set_zero(&dg);
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                ompf_mul(&t, &A[i + (size_t)j * n], &A[i + (size_t)j * n]);
                if (i != j)
                    ompf_add(&off, &off, &t);
                else
                    ompf_add(&dg, &dg, &t);
            }
        ompf_mul(&t, &dg, &eps);
        ompf_mul(&t, &t, &eps);
        if (ompf_cmp(&off, &t) <= 0 || ompf_sgn(&off) == 0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                apq = A[p + (size_t)q * n];
                if (ompf_sgn(&apq) == 0) continue;
                app = A[p + (size_t)p * n];
                aqq = A[q + (size_t)q * n];
                /* theta = (aqq - app) / (2 apq); t = sgn(theta)/(|theta| + sqrt(theta^2+1)) */
                ompf_sub(&t1, &aqq, &app);
                ompf_mul(&t2, &two, &apq);
                ompf_div(&th, &t1, &t2);
                ompf_mul(&t1, &th, &th);
                ompf_add(&t1, &t1, &one);
                ompf_sqrt(&t1, &t1);
                ompf_abs(&t2, &th);
                ompf_add(&t1, &t1, &t2);
                ompf_div(&t, &one, &t1);
                if (ompf_sgn(&th) < 0) ompf_neg(&t, &t);
                ompf_mul(&t1, &t, &t);
                ompf_add(&t1, &t1, &one);
                ompf_sqrt(&t1, &t1);
                ompf_div(&c, &one, &t1);
                ompf_mul(&s, &t, &c);
                ompf_add(&t1, &one, &c);
                ompf_div(&tau, &s, &t1);
                /* app -= t apq ; aqq += t apq ; apq = 0 */
                ompf_mul(&t1, &t, &apq);
                ompf_sub(&A[p + (size_t)p * n], &app, &t1);
                ompf_add(&A[q + (size_t)q * n], &aqq, &t1);
                ompf_set_zero(&A[p + (size_t)q * n]);
                ompf_set_zero(&A[q + (size_t)p * n]);
                for (int k = 0; k < n; k++) {
                    if (k == p || k == q) continue;
                    akp = A[k + (size_t)p * n];
                    akq = A[k + (size_t)q * n];
                    /* akp' = akp - s (akq + tau akp); akq' = akq + s (akp - tau akq) */
                    ompf_mul(&t1, &tau, &akp);
                    ompf_add(&t1, &t1, &akq);
                    ompf_mul(&t1, &s, &t1);
                    ompf_sub(&t2, &akp, &t1);
                    A[k + (size_t)p * n] = t2;
                    A[p + (size_t)k * n] = t2;
                    ompf_mul(&t1, &tau, &akq);
                    ompf_sub(&t1, &akp, &t1);
                    ompf_mul(&t1, &s, &t1);
                    ompf_add(&t2, &akq, &t1);
                    A[k + (size_t)q * n] = t2;
                    A[q + (size_t)k * n] = t2;
                }
            }
    }
    F mn = A[0];
    for (int i = 1; i < n; i++)
        if (Flt(&A[i + (size_t)i * n], &mn)) mn = A[i + (size_t)i * n];
    free(A);
    return mn;
}

/*
 * Eigen-decomposition of the symmetric tridiagonal matrix (d[0..n), e[0..n-1))
 * by the implicit QL method with Wilkinson shifts; on return d holds the
 * eigenvalues in ascending order and z (n x n, column-major, leading dim ldz)
 * the eigenvectors.  Returns 0 or >0 when it does not converge.
 */
static int tridiag_ql(int n, F *d, F *e_in, F *z, int ldz) {
    F *e = (F *)xcalloc(n + 1, sizeof(F));
    for (int i = 0; i < n - 1; i++) e[i] = e_in[i];
    ompf_set_zero(&e[n - 1]);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            if (i == j)
                Fd(&z[i + (size_t)j * ldz], 1.0);
            else
                ompf_set_zero(&z[i + (size_t)j * ldz]);
        }
    F one = Fv(1.0), two = Fv(2.0), eps, g, r, s, c, p, f, b, t1, t2, dd;
    ompf_mul_2exp(&eps, &one, -(long)ompf_prec_bits());
    int info = 0;
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        for (;;) {
            for (m = l; m < n - 1; m++) {
                ompf_abs(&t1, &d[m]);
                ompf_abs(&t2, &d[m + 1]);
                ompf_add(&dd, &t1, &t2);
                ompf_abs(&t1, &e[m]);
                ompf_mul(&t2, &eps, &dd);
                if (ompf_cmp(&t1, &t2) <= 0) break;
            }
            if (m == l) break;
            if (iter++ == 60 * n) {
                info = l + 1;
                goto done;
            }
            /* g = (d[l+1]-d[l])/(2 e[l]); r = hypot(g,1); g = d[m]-d[l]+e[l]/(g + sign(r,g)) */
            ompf_sub(&t1, &d[l + 1], &d[l]);
            ompf_mul(&t2, &two, &e[l]);
            ompf_div(&g, &t1, &t2);
            ompf_mul(&t1, &g, &g);
            ompf_add(&t1, &t1, &one);
            ompf_sqrt(&r, &t1);
            if (ompf_sgn(&g) < 0) ompf_neg(&r, &r);
            ompf_add(&t1, &g, &r);
            ompf_div(&t1, &e[l], &t1);
            ompf_sub(&g, &d[m], &d[l]);
            ompf_add(&g, &g, &t1);
            Fd(&s, 1.0);
            Fd(&c, 1.0);
            ompf_set_zero(&p);
            int i;
            int early = 0;
            for (i = m - 1; i >= l; i--) {
                ompf_mul(&f, &s, &e[i]);
                ompf_mul(&b, &c, &e[i]);
                /* r = hypot(f, g) */
                ompf_mul(&t1, &f, &f);
                ompf_mul(&t2, &g, &g);
                ompf_add(&t1, &t1, &t2);
                ompf_sqrt(&r, &t1);
                e[i + 1] = r;
                if (ompf_sgn(&r) == 0) {
                    ompf_sub(&d[i + 1], &d[i + 1], &p);
                    ompf_set_zero(&e[m]);
                    early = 1;
                    break;
                }
                ompf_div(&s, &f, &r);
                ompf_div(&c, &g, &r);
                ompf_sub(&g, &d[i + 1], &p);
                ompf_sub(&t1, &d[i], &g);
                ompf_mul(&t1, &t1, &s);
                ompf_mul(&t2, &two, &c);
                ompf_mul(&t2, &t2, &b);
                ompf_add(&r, &t1, &t2);
                ompf_mul(&p, &s, &r);
                ompf_add(&d[i + 1], &g, &p);
                ompf_mul(&t1, &c, &r);
                ompf_sub(&g, &t1, &b);
                for (int k = 0; k < n; k++) {
                    F *zk1 = &z[k + (size_t)(i + 1) * ldz], *zk = &z[k + (size_t)i * ldz];
                    f = *zk1;
                    ompf_mul(&t1, &s, zk);
                    ompf_mul(&t2, &c, &f);
                    ompf_add(zk1, &t1, &t2);
                    ompf_mul(&t1, &c, zk);
                    ompf_mul(&t2, &s, &f);
                    ompf_sub(zk, &t1, &t2);
                }
            }
            if (early) continue;
            ompf_sub(&d[l], &d[l], &p);
            e[l] = g;
            ompf_set_zero(&e[m]);
        }
    }
    /* sort ascending with vectors */
    for (int i = 0; i < n - 1; i++) {
        int k = i;
        for (int j = i + 1; j < n; j++)
            if (Flt(&d[j], &d[k])) k = j;
        if (k != i) {
            F t = d[k];
            d[k] = d[i];
            d[i] = t;
            for (int q = 0; q < n; q++) {
                F u = z[q + (size_t)i * ldz];
                z[q + (size_t)i * ldz] = z[q + (size_t)k * ldz];
                z[q + (size_t)k * ldz] = u;
            }
        }
    }
done:
    free(e);
    return info;
}

/*
 * Lal::getMinEigen — Lanczos estimate of the minimum eigenvalue of
 * L X L^T (sdpa_linear.cpp:27-124), control flow restated as is.
 */
static F lanczos_min_eig(int nDim, const F *lMat, const F *xMat) {
    F min = Fv(1.0e+51), min_old = Fv(1.0e+52), min_min = Fv(1.0e+50);
    F error = Fv(1.0e+10), one = Fv(1.0);
    F alpha, beta, value, t1, t2, t3;
    F *diagVec = (F *)xcalloc(nDim, sizeof(F)), *diagVec2 = (F *)xcalloc(nDim, sizeof(F));
    F *q = (F *)xcalloc(nDim, sizeof(F)), *qold = (F *)xcalloc(nDim, sizeof(F));
    F *r = (F *)xcalloc(nDim, sizeof(F)), *w = (F *)xcalloc(nDim, sizeof(F)), *tmp = (F *)xcalloc(nDim, sizeof(F));
    F *out = (F *)xcalloc(nDim, sizeof(F)), *bb = (F *)xcalloc(nDim, sizeof(F));
    F *Q = (F *)xcalloc((size_t)nDim * nDim, sizeof(F));
    for (int i = 0; i < nDim; i++) {
        diagVec[i] = min_min;
        ompf_set_zero(&diagVec2[i]);
        ompf_set_zero(&q[i]);
        r[i] = one;
    }
    F nD;
    ompf_set_si(&nD, nDim);
    ompf_sqrt(&beta, &nD);
    F sqrtn;
    ompf_sqrt(&sqrtn, &nD);
    F sqrtn10 = sqrtn;
    F ten = Fv(10.0);
    ompf_add(&sqrtn10, &sqrtn, &ten);
    F nD9;
    ompf_set_si(&nD9, nDim + 9);
    F sqrtn9;
    ompf_sqrt(&sqrtn9, &nD9);
    F c1e16 = Fv(1.0e-16), c1e5 = Fv(1.0e-5), c1e8 = Fv(1.0e-8), c1e2 = Fv(1.0e-2), c1e4 = Fv(1.0e-4);
    int k = 0, kk = 0;
    for (;;) {
        /* while condition */
        F kf;
        ompf_set_si(&kf, k);
        if (!(k < nDim)) break;
        if (!(ompf_cmp(&kf, &sqrtn10) < 0)) break;
        if (!(Fgt(&beta, &c1e16))) break;
        ompf_sub(&t1, &min, &min_old);
        ompf_abs(&t1, &t1);
        ompf_abs(&t2, &min);
        ompf_mul(&t2, &c1e5, &t2);
        ompf_add(&t2, &t2, &c1e8);
        int condA = Fgt(&t1, &t2);
        ompf_mul(&t1, &error, &beta);
        ompf_abs(&t1, &t1);
        ompf_abs(&t3, &min);
        ompf_mul(&t3, &c1e2, &t3);
        ompf_add(&t3, &t3, &c1e4);
        int condB = Fgt(&t1, &t3);
        if (!(condA || condB)) break;

        for (int i = 0; i < nDim; i++) qold[i] = q[i];
        ompf_div(&value, &one, &beta);
        for (int i = 0; i < nDim; i++) ompf_mul(&q[i], &value, &r[i]); /* Rcopy + Rscal */
        for (int i = 0; i < nDim; i++) w[i] = q[i];
        trmv_lt(nDim, lMat, nDim, w);
        gemv_n(nDim, nDim, xMat, nDim, w, tmp);
        for (int i = 0; i < nDim; i++) w[i] = tmp[i];
        trmv_ln(nDim, lMat, nDim, w);
        alpha = Rdot(nDim, q, 1, w, 1);
        diagVec[k] = alpha;
        /* r = w - alpha q ; r = r - beta qold */
        F malpha, mbeta;
        ompf_neg(&malpha, &alpha);
        for (int i = 0; i < nDim; i++) r[i] = w[i];
        Raxpy(nDim, &malpha, q, 1, r, 1);
        ompf_neg(&mbeta, &beta);
        Raxpy(nDim, &mbeta, qold, 1, r, 1);

        F kf2, sqk;
        ompf_set_si(&kf2, kk);
        ompf_set_si(&t1, k);
        ompf_sqrt(&sqk, &t1);
        if (ompf_cmp(&kf2, &sqk) >= 0 || k == nDim - 1 || ompf_cmp(&kf, &sqrtn9) > 0) {
            kk = 0;
            for (int i = 0; i < nDim; i++) {
                out[i] = diagVec[i];
                bb[i] = diagVec2[i];
            }
            out[nDim - 1] = diagVec[k];
            ompf_set_zero(&bb[nDim - 1]);
            int kp1 = k + 1;
            int info = tridiag_ql(kp1, out, bb, Q, nDim);
            if (info > 0) {
                fprintf(stderr, " rLanczos :: cannot converge %d\n", info);
                break;
            }
            min_old = min;
            min = out[0];
            error = Q[k]; /* component k of the eigenvector of the smallest eigenvalue */
        }
        value = Rdot(nDim, r, 1, r, 1);
        ompf_sqrt(&beta, &value);
        diagVec2[k] = beta;
        ++k;
        ++kk;
    }
    free(diagVec);
    free(diagVec2);
    free(q);
    free(qold);
    free(r);
    free(w);
    free(tmp);
    free(out);
    free(bb);
    free(Q);
    /* return min - |error*beta| */
    ompf_mul(&t1, &error, &beta);
    ompf_abs(&t1, &t1);
    F res;
    ompf_sub(&res, &min, &t1);
    return res;
}

/* Jal::getMinEigen(lMat, xMat) (sdpa_jordan.cpp:55-88) */
static F jal_min_eigen(const DLS *lMat, const DLS *xMat) {
    F mn = Fv(1.0E50), value;
    F one = Fv(1.0);
    for (int l = 0; l < xMat->nsdp; l++) {
        int n = xMat->n[l];
        if (n > 20) {
            value = lanczos_min_eig(n, lMat->b[l], xMat->b[l]);
        } else {
            F *w2 = (F *)xcalloc((size_t)n * n, sizeof(F)), *w1 = (F *)xcalloc((size_t)n * n, sizeof(F));
            gemm_nt(n, n, n, &one, xMat->b[l], n, lMat->b[l], n, 0, w2, n); /* multiply_tran */
            gemm_nn(n, n, n, &one, lMat->b[l], n, w2, n, 0, w1, n);
            value = min_eig_jacobi(n, w1);
            free(w1);
            free(w2);
        }
        if (Flt(&value, &mn)) mn = value;
    }
    for (int l = 0; l < xMat->nlp; l++) {
        ompf_mul(&value, &xMat->lp[l], &lMat->lp[l]);
        ompf_mul(&value, &value, &lMat->lp[l]);
        if (Flt(&value, &mn)) mn = value;
    }
    return mn;
}

/* ================= sparse operations ================= */

/* getInnerProduct(SparseMatrix, DenseMatrix), _OPENMP branch (sdpa_linear.cpp:235-254) */
static F sp_inner(const SpBlk *a, const F *B) {
    int n = a->n;
    if (a->dense) return Rdot(n * n, a->de, 1, B, 1);
    F ret, temp;
    ompf_set_zero(&ret);
    for (int idx = 0; idx < a->nnz; idx++) {
        int i = a->ri[idx], j = a->ci[idx];
        ompf_set_zero(&temp);
        if (i == j) {
            ompf_add(&temp, &temp, &a->v[idx]);
            ompf_mul(&temp, &temp, &B[i + (size_t)n * j]);
        } else {
            ompf_add(&temp, &temp, &B[i + (size_t)n * j]);
            ompf_add(&temp, &temp, &B[j + (size_t)n * i]);
            ompf_mul(&temp, &temp, &a->v[idx]);
        }
        ompf_add(&ret, &ret, &temp);
    }
    return ret;
}
/* getInnerProduct(SparseLinearSpace, DenseLinearSpace) (sdpa_linear.cpp:1676-1711) */
static F sls_inner(const SLS *a, const DLS *b) {
    F ret, t;
    ompf_set_zero(&ret);
    for (int l = 0; l < a->nsdp; l++) {
        t = sp_inner(&a->s[l], b->b[a->s[l].blk]);
        ompf_add(&ret, &ret, &t);
    }
    for (int l = 0; l < a->nlp; l++) {
        ompf_mul(&t, &a->lpv[l], &b->lp[a->lpi[l]]);
        ompf_add(&ret, &ret, &t);
    }
    return ret;
}
/* plus(DenseMatrix, DenseMatrix, SparseMatrix, scalar) in place: ret += s*A (sdpa_linear.cpp:1278-1372) */
static void sp_plus_inplace(F *ret, const SpBlk *a, const F *s) {
    int n = a->n;
    if (a->dense) {
        Raxpy(n * n, s, a->de, 1, ret, 1);
        return;
    }
    F value;
    for (int idx = 0; idx < a->nnz; idx++) {
        int i = a->ri[idx], j = a->ci[idx];
        ompf_mul(&value, &a->v[idx], s);
        if (i != j) {
            ompf_add(&ret[i + (size_t)n * j], &ret[i + (size_t)n * j], &value);
            ompf_add(&ret[j + (size_t)n * i], &ret[j + (size_t)n * i], &value);
        } else {
            ompf_add(&ret[i + (size_t)n * i], &ret[i + (size_t)n * i], &value);
        }
    }
}
/* multiply(DenseMatrix, SparseMatrix, DenseMatrix, NULL): ret = A*B (sdpa_linear.cpp:877-955) */
static void sp_mul_dense(F *ret, const SpBlk *a, const F *B) {
    int n = a->n;
    F one = Fv(1.0);
    if (a->dense) {
        gemm_nn(n, n, n, &one, a->de, n, B, n, 0, ret, n);
        return;
    }
    for (int i = 0; i < n * n; i++) ompf_set_zero(&ret[i]);
    for (int idx = 0; idx < a->nnz; idx++) {
        int i = a->ri[idx], j = a->ci[idx];
        F value = a->v[idx];
        if (i != j) {
            Raxpy(n, &value, B + j, n, ret + i, n);
            Raxpy(n, &value, B + i, n, ret + j, n);
        } else {
            Raxpy(n, &value, B + j, n, ret + j, n);
        }
    }
}

/* ================= input (sdpa_io.cpp:123-142, 410-751; sdpa_struct.cpp:484-525, 1259-1291) ================= */

typedef struct {
    FILE *fp;
} Rd;

static int skip_to_num(FILE *fp) {
    int c;
    while ((c = fgetc(fp)) != EOF) {
        if ((c >= '0' && c <= '9') || c == '+' || c == '-') {
            ungetc(c, fp);
            return 1;
        }
    }
    return 0;
}
static int read_int(FILE *fp, int *v) {
    if (!skip_to_num(fp)) return 0;
    return fscanf(fp, "%d", v) == 1;
}
static int read_F(FILE *fp, F *v) {
    if (!skip_to_num(fp)) return 0;
    char buf[512];
    int n = 0, c;
    while ((c = fgetc(fp)) != EOF) {
        if ((c >= '0' && c <= '9') || c == '+' || c == '-' || c == '.' || c == 'e' || c == 'E') {
            if (n < 510) buf[n++] = (char)c;
        } else {
            ungetc(c, fp);
            break;
        }
    }
    buf[n] = 0;
    return ompf_set_str(v, buf) > 0;
}

static void spblk_push(SpBlk *s, int i, int j, const F *v) {
    if (s->nnz == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 8;
        s->ri = (int *)realloc(s->ri, sizeof(int) * s->cap);
        s->ci = (int *)realloc(s->ci, sizeof(int) * s->cap);
        s->v = (F *)realloc(s->v, sizeof(F) * s->cap);
    }
    s->ri[s->nnz] = i;
    s->ci[s->nnz] = j;
    s->v[s->nnz] = *v;
    s->nnz++;
    s->effect += (i == j) ? 1 : 2;
}

static SpBlk *sls_block(SLS *a, int blk, int n) {
    for (int k = 0; k < a->nsdp; k++)
        if (a->s[k].blk == blk) return &a->s[k];
    /* keep blocks sorted by block index (setBlockStruct sorts SDP_index) */
    a->s = (SpBlk *)realloc(a->s, sizeof(SpBlk) * (a->nsdp + 1));
    int pos = a->nsdp;
    while (pos > 0 && a->s[pos - 1].blk > blk) {
        a->s[pos] = a->s[pos - 1];
        pos--;
    }
    memset(&a->s[pos], 0, sizeof(SpBlk));
    a->s[pos].blk = blk;
    a->s[pos].n = n;
    a->nsdp++;
    return &a->s[pos];
}
static void sls_lp_set(SLS *a, int pos, const F *v) {
    for (int k = 0; k < a->nlp; k++)
        if (a->lpi[k] == pos) {
            a->lpv[k] = *v;
            return;
        }
    a->lpi = (int *)realloc(a->lpi, sizeof(int) * (a->nlp + 1));
    a->lpv = (F *)realloc(a->lpv, sizeof(F) * (a->nlp + 1));
    int p = a->nlp;
    while (p > 0 && a->lpi[p - 1] > pos) {
        a->lpi[p] = a->lpi[p - 1];
        a->lpv[p] = a->lpv[p - 1];
        p--;
    }
    a->lpi[p] = pos;
    a->lpv[p] = *v;
    a->nlp++;
}
/* SparseMatrix::changeToDense (sdpa_struct.cpp:484-525) */
static void spblk_maybe_dense(SpBlk *s) {
    int n = s->n;
    if ((double)s->nnz < (double)(n * n) * 0.20) return;
    s->dense = 1;
    s->de = (F *)xcalloc((size_t)n * n, sizeof(F));
    for (int idx = 0; idx < s->nnz; idx++) {
        int i = s->ri[idx], j = s->ci[idx];
        s->de[i + (size_t)n * j] = s->v[idx];
        s->de[j + (size_t)n * i] = s->v[idx];
    }
    s->nnz = s->effect = n * n;
}

static int read_problem(const char *path, Prob *P) {
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    memset(P, 0, sizeof(*P));
    char line[4096];
    for (;;) {
        if (!fgets(line, sizeof line, fp)) {
            fclose(fp);
            return -2;
        }
        if (line[0] == '*' || line[0] == '"') continue;
        sscanf(line, "%d", &P->m);
        break;
    }
    if (fscanf(fp, "%d", &P->nBlock) != 1) {
        fclose(fp);
        return -3;
    }
    P->blockStruct = (int *)xcalloc(P->nBlock, sizeof(int));
    P->blockType = (int *)xcalloc(P->nBlock, sizeof(int));
    P->blockNumber = (int *)xcalloc(P->nBlock, sizeof(int));
    for (int l = 0; l < P->nBlock; l++) read_int(fp, &P->blockStruct[l]);
    /* block typing (sdpa_main.cpp:114-131) */
    int nsdp = 0, nlp = 0;
    for (int i = 0; i < P->nBlock; i++) {
        if (P->blockStruct[i] >= 2) {
            P->blockType[i] = 1;
            P->blockNumber[i] = nsdp++;
        } else if (P->blockStruct[i] < 0) {
            P->blockType[i] = 3;
            P->blockStruct[i] = -P->blockStruct[i];
            P->blockNumber[i] = nlp;
            nlp += P->blockStruct[i];
        } else if (P->blockStruct[i] == 1) {
            P->blockType[i] = 3;
            P->blockNumber[i] = nlp;
            nlp += 1;
        } else {
            fclose(fp);
            return -4;
        }
    }
    P->nsdp = nsdp;
    P->nlp = nlp;
    P->sdpStruct = (int *)xcalloc(nsdp, sizeof(int));
    for (int i = 0, k = 0; i < P->nBlock; i++)
        if (P->blockType[i] == 1) P->sdpStruct[k++] = P->blockStruct[i];
    P->nDim = nlp;
    for (int l = 0; l < nsdp; l++) P->nDim += P->sdpStruct[l];
    P->b = (F *)xcalloc(P->m, sizeof(F));
    for (int k = 0; k < P->m; k++) read_F(fp, &P->b[k]);
    P->A = (SLS *)xcalloc(P->m, sizeof(SLS));
    for (;;) {
        int k, l, i, j;
        F v;
        if (!read_int(fp, &k)) break;
        if (!read_int(fp, &l)) break;
        if (!read_int(fp, &i)) break;
        if (!read_int(fp, &j)) break;
        if (!read_F(fp, &v)) break;
        if (k < 0 || k > P->m || l < 1 || l > P->nBlock) {
            fclose(fp);
            return -5;
        }
        SLS *tgt = (k == 0) ? &P->C : &P->A[k - 1];
        F val = v;
        if (k == 0) ompf_neg(&val, &v); /* C = -F0 (sdpa_io.cpp:661) */
        if (P->blockType[l - 1] == 1) {
            int l2 = P->blockNumber[l - 1];
            SpBlk *s = sls_block(tgt, l2, P->sdpStruct[l2]);
            if (i < 1 || j < 1 || i > s->n || j > s->n) {
                fclose(fp);
                return -6;
            }
            spblk_push(s, i - 1, j - 1, &val);
        } else {
            if (i != j) {
                fclose(fp);
                return -7;
            }
            sls_lp_set(tgt, P->blockNumber[l - 1] + i - 1, &val);
        }
    }
    fclose(fp);
    /* changeToDense on C and all A_k (sdpa_main.cpp:197-200) */
    for (int l = 0; l < P->C.nsdp; l++) spblk_maybe_dense(&P->C.s[l]);
    for (int k = 0; k < P->m; k++)
        for (int l = 0; l < P->A[k].nsdp; l++) spblk_maybe_dense(&P->A[k].s[l]);
    /* InputData::initialize_index (sdpa_dataset.cpp:254-325, 395-474) */
    P->sdpNCon = (int *)xcalloc(nsdp, sizeof(int));
    P->sdpCon = (int **)xcalloc(nsdp, sizeof(int *));
    P->sdpBIdx = (int **)xcalloc(nsdp, sizeof(int *));
    for (int i = 0; i < P->m; i++)
        for (int k = 0; k < P->A[i].nsdp; k++) P->sdpNCon[P->A[i].s[k].blk]++;
    for (int l = 0; l < nsdp; l++) {
        P->sdpCon[l] = (int *)xcalloc(P->sdpNCon[l], sizeof(int));
        P->sdpBIdx[l] = (int *)xcalloc(P->sdpNCon[l], sizeof(int));
        P->sdpNCon[l] = 0;
    }
    for (int i = 0; i < P->m; i++)
        for (int k = 0; k < P->A[i].nsdp; k++) {
            int l = P->A[i].s[k].blk;
            P->sdpCon[l][P->sdpNCon[l]] = i;
            P->sdpBIdx[l][P->sdpNCon[l]] = k;
            P->sdpNCon[l]++;
        }
    P->lpNCon = (int *)xcalloc(nlp, sizeof(int));
    P->lpCon = (int **)xcalloc(nlp, sizeof(int *));
    P->lpBIdx = (int **)xcalloc(nlp, sizeof(int *));
    for (int i = 0; i < P->m; i++)
        for (int k = 0; k < P->A[i].nlp; k++) P->lpNCon[P->A[i].lpi[k]]++;
    for (int l = 0; l < nlp; l++) {
        P->lpCon[l] = (int *)xcalloc(P->lpNCon[l], sizeof(int));
        P->lpBIdx[l] = (int *)xcalloc(P->lpNCon[l], sizeof(int));
        P->lpNCon[l] = 0;
    }
    for (int i = 0; i < P->m; i++)
        for (int k = 0; k < P->A[i].nlp; k++) {
            int l = P->A[i].lpi[k];
            P->lpCon[l][P->lpNCon[l]] = i;
            P->lpBIdx[l][P->lpNCon[l]] = k;
            P->lpNCon[l]++;
        }
    return 0;
}

static void free_sls(SLS *a) {
    for (int l = 0; l < a->nsdp; l++) {
        free(a->s[l].ri);
        free(a->s[l].ci);
        free(a->s[l].v);
        free(a->s[l].de);
    }
    free(a->s);
    free(a->lpi);
    free(a->lpv);
}
static void free_problem(Prob *P) {
    free_sls(&P->C);
    for (int k = 0; k < P->m; k++) free_sls(&P->A[k]);
    free(P->A);
    free(P->b);
    for (int l = 0; l < P->nsdp; l++) {
        free(P->sdpCon[l]);
        free(P->sdpBIdx[l]);
    }
    for (int l = 0; l < P->nlp; l++) {
        free(P->lpCon[l]);
        free(P->lpBIdx[l]);
    }
    free(P->sdpNCon);
    free(P->sdpCon);
    free(P->sdpBIdx);
    free(P->lpNCon);
    free(P->lpCon);
    free(P->lpBIdx);
    free(P->blockStruct);
    free(P->blockType);
    free(P->blockNumber);
    free(P->sdpStruct);
}

/* ================= parameters (sdpa_parts.cpp:136-225) ================= */

void orc_param_default(orc_param *p) {
    p->maxIteration = 200;
    p->epsilonStar = 1.0e-30;
    p->lambdaStar = 1.0e+4;
    p->omegaStar = 2.0;
    p->lowerBound = -1.0e+5;
    p->upperBound = 1.0e+5;
    p->betaStar = 0.1;
    p->betaBar = 0.3;
    p->gammaStar = 0.9;
    p->epsilonDash = 1.0e-30;
    p->precision = 200;
}

int orc_param_read(orc_param *p, const char *path) {
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    int ok = 1;
    ok &= fscanf(fp, "%d%*[^\n]", &p->maxIteration) == 1;
    ok &= fscanf(fp, "%lf%*[^\n]", &p->epsilonStar) == 1;
    ok &= fscanf(fp, "%lf%*[^\n]", &p->lambdaStar) == 1;
    ok &= fscanf(fp, "%lf%*[^\n]", &p->omegaStar) == 1;
    ok &= fscanf(fp, "%lf%*[^\n]", &p->lowerBound) == 1;
    ok &= fscanf(fp, "%lf%*[^\n]", &p->upperBound) == 1;
    ok &= fscanf(fp, "%lf%*[^\n]", &p->betaStar) == 1;
    ok &= fscanf(fp, "%lf%*[^\n]", &p->betaBar) == 1;
    ok &= fscanf(fp, "%lf%*[^\n]", &p->gammaStar) == 1;
    ok &= fscanf(fp, "%lf%*[^\n]", &p->epsilonDash) == 1;
    ok &= fscanf(fp, "%d%*[^\n]", &p->precision) == 1;
    fclose(fp);
    return ok ? 0 : -2;
}

/* ================= solver state ================= */

enum { noINFO, pFEAS, dFEAS, pdFEAS, pdINF, pFEAS_dINF, pINF_dFEAS, pdOPT, pUNBD, dUNBD };
enum { F1 = 1, F2 = 2, F3 = 3 };

typedef struct {
    int m;
    F *bMat; /* m x m column-major */
    F *gVec, *dy;
    DLS dX, dZ, r_zinv, x_rd_zinv;
    int *useFormula; /* m * nsdp */
} Newton;

typedef struct {
    DLS X, Z, invCholX, invCholZ, invZ;
    F *y;
} Sol;

typedef struct {
    F *primalVec;
    DLS dualMat;
    F normPrimalVec, normDualMat;
} Res;

typedef struct {
    DLS DLS1, DLS2;
} Work;

/* Newton::computeFormula_SDP (sdpa_newton.cpp:694-819), KAPPA = 1.2 (sdpa_main.cpp:27) */
static void compute_formula(Newton *N, const Prob *P) {
    int m = P->m, nb = P->nsdp;
    int *up = (int *)xcalloc((size_t)m * (nb ? nb : 1), sizeof(int));
    for (int l = 0; l < nb; l++)
        for (int k1 = 0; k1 < P->sdpNCon[l]; k1++) {
            int i = P->sdpCon[l][k1], ib = P->sdpBIdx[l][k1];
            int inz = P->A[i].s[ib].effect, u = inz;
            for (int k2 = 0; k2 < P->sdpNCon[l]; k2++) {
                int j = P->sdpCon[l][k2], jb = P->sdpBIdx[l][k2];
                int jnz = P->A[j].s[jb].effect;
                if (jnz < inz)
                    u += jnz;
                else if (jnz == inz && j < i)
                    u += jnz;
            }
            up[i * nb + l] = u;
        }
    F kappa = Fv(1.2), one = Fv(1.0), two = Fv(2.0);
    for (int l = 0; l < nb; l++)
        for (int k = 0; k < P->sdpNCon[l]; k++) {
            int i = P->sdpCon[l][k], ib = P->sdpBIdx[l][k];
            const SpBlk *s = &P->A[i].s[ib];
            F inz, n, u, f1, f2, f3, t, t2;
            ompf_set_si(&inz, s->effect);
            ompf_set_si(&n, s->n);
            ompf_set_si(&u, up[i * nb + l]);
            /* f1 = K n inz + n^3 + K up */
            ompf_mul(&f1, &kappa, &n);
            ompf_mul(&f1, &f1, &inz);
            ompf_mul(&t, &n, &n);
            ompf_mul(&t, &t, &n);
            ompf_add(&f1, &f1, &t);
            ompf_mul(&t, &kappa, &u);
            ompf_add(&f1, &f1, &t);
            /* f2 = K n inz + K (n+1) up */
            ompf_mul(&f2, &kappa, &n);
            ompf_mul(&f2, &f2, &inz);
            ompf_add(&t, &n, &one);
            ompf_mul(&t, &kappa, &t);
            ompf_mul(&t, &t, &u);
            ompf_add(&f2, &f2, &t);
            /* f3 = K (2 K inz + 1) up / K */
            ompf_mul(&t, &two, &kappa);
            ompf_mul(&t, &t, &inz);
            ompf_add(&t, &t, &one);
            ompf_mul(&t2, &kappa, &t);
            ompf_mul(&t2, &t2, &u);
            ompf_div(&f3, &t2, &kappa);
            int fm;
            if (s->dense) {
                fm = Flt(&f1, &f2) ? F1 : F2;
            } else {
                if (Flt(&f1, &f2) && Flt(&f1, &f3))
                    fm = F1;
                else if (Flt(&f2, &f3))
                    fm = F2;
                else
                    fm = F3;
            }
            N->useFormula[i * nb + l] = fm;
        }
    free(up);
}

/* Newton::calF2 (sdpa_newton.cpp:899-937) */
static void calF2(F *ret, const F *Fm, F *G, const F *X, const SpBlk *Aj, int *hasG, int n) {
    F one = Fv(1.0);
    if (!Aj->dense) {
        F value2, t;
        ompf_set_zero(ret);
        for (int idx = 0; idx < Aj->nnz; idx++) {
            int alpha = Aj->ri[idx], beta = Aj->ci[idx];
            value2 = Rdot(n, X + alpha, n, Fm + (size_t)n * beta, 1);
            ompf_mul(&t, &Aj->v[idx], &value2);
            ompf_add(ret, ret, &t);
            if (alpha != beta) {
                value2 = Rdot(n, X + beta, n, Fm + (size_t)n * alpha, 1);
                ompf_mul(&t, &Aj->v[idx], &value2);
                ompf_add(ret, ret, &t);
            }
        }
    } else {
        if (!*hasG) {
            gemm_nn(n, n, n, &one, X, n, Fm, n, 0, G, n);
            *hasG = 1;
        }
        *ret = sp_inner(Aj, G);
    }
}

/* Newton::calF3 (sdpa_newton.cpp:939-1005) */
static void calF3(F *ret, const F *X, const F *invZ, const SpBlk *Ai, const SpBlk *Aj, int n) {
    F local, sum, plu;
    ompf_set_zero(&local);
    for (int i1 = 0; i1 < Aj->nnz; i1++) {
        int alpha = Aj->ri[i1], beta = Aj->ci[i1];
        const F *value1 = &Aj->v[i1];
        ompf_set_zero(&sum);
        for (int i2 = 0; i2 < Ai->nnz; i2++) {
            int gamma = Ai->ri[i2], delta = Ai->ci[i2];
            ompf_mul(&plu, &Ai->v[i2], &invZ[delta + (size_t)n * beta]);
            ompf_mul(&plu, &plu, &X[alpha + (size_t)n * gamma]);
            ompf_add(&sum, &sum, &plu);
            if (gamma != delta) {
                ompf_mul(&plu, &Ai->v[i2], &invZ[gamma + (size_t)n * beta]);
                ompf_mul(&plu, &plu, &X[alpha + (size_t)n * delta]);
                ompf_add(&sum, &sum, &plu);
            }
        }
        ompf_mul(&plu, value1, &sum);
        ompf_add(&local, &local, &plu);
        if (alpha == beta) continue;
        ompf_set_zero(&sum);
        for (int i2 = 0; i2 < Ai->nnz; i2++) {
            int gamma = Ai->ri[i2], delta = Ai->ci[i2];
            ompf_mul(&plu, &Ai->v[i2], &invZ[delta + (size_t)n * alpha]);
            ompf_mul(&plu, &plu, &X[beta + (size_t)n * gamma]);
            ompf_add(&sum, &sum, &plu);
            if (gamma != delta) {
                ompf_mul(&plu, &Ai->v[i2], &invZ[gamma + (size_t)n * alpha]);
                ompf_mul(&plu, &plu, &X[beta + (size_t)n * delta]);
                ompf_add(&sum, &sum, &plu);
            }
        }
        ompf_mul(&plu, value1, &sum);
        ompf_add(&local, &local, &plu);
    }
    ompf_set_zero(ret);
    ompf_add(ret, ret, &local);
}

/* Make_bMat -> compute_bMat_dense_SDP + compute_bMat_dense_LP (sdpa_newton.cpp:1007-1096, 1190-1225, 1251-1272) */
static void make_bmat(Newton *N, const Prob *P, const Sol *S, Work *W) {
    int m = P->m, nb = P->nsdp;
    for (size_t i = 0; i < (size_t)m * m; i++) ompf_set_zero(&N->bMat[i]);
    F one = Fv(1.0);
    for (int l = 0; l < nb; l++) {
        int n = P->sdpStruct[l];
        const F *X = S->X.b[l], *invZ = S->invZ.b[l];
        F *work1 = W->DLS1.b[l], *work2 = W->DLS2.b[l];
        for (int k1 = 0; k1 < P->sdpNCon[l]; k1++) {
            int i = P->sdpCon[l][k1], ib = P->sdpBIdx[l][k1];
            const SpBlk *Ai = &P->A[i].s[ib];
            int inz = Ai->effect;
            int fm = N->useFormula[i * nb + l];
            int hasG = 0;
            if (fm == F1) {
                sp_mul_dense(work1, Ai, invZ);
                gemm_nn(n, n, n, &one, X, n, work1, n, 0, work2, n);
            } else if (fm == F2) {
                sp_mul_dense(work1, Ai, invZ);
            }
            for (int k2 = 0; k2 < P->sdpNCon[l]; k2++) {
                int j = P->sdpCon[l][k2], jb = P->sdpBIdx[l][k2];
                const SpBlk *Aj = &P->A[j].s[jb];
                int jnz = Aj->effect;
                if (inz < jnz || (inz == jnz && i < j)) continue;
                F value;
                if (fm == F1)
                    value = sp_inner(Aj, work2);
                else if (fm == F2)
                    calF2(&value, work1, work2, X, Aj, &hasG, n);
                else
                    calF3(&value, X, invZ, Ai, Aj, n);
                if (i != j) {
                    ompf_add(&N->bMat[i + (size_t)m * j], &N->bMat[i + (size_t)m * j], &value);
                    ompf_add(&N->bMat[j + (size_t)m * i], &N->bMat[j + (size_t)m * i], &value);
                } else {
                    ompf_add(&N->bMat[i + (size_t)m * i], &N->bMat[i + (size_t)m * i], &value);
                }
            }
        }
    }
    for (int l = 0; l < P->nlp; l++) {
        const F *x = &S->X.lp[l], *iz = &S->invZ.lp[l];
        for (int k1 = 0; k1 < P->lpNCon[l]; k1++) {
            int i = P->lpCon[l][k1], ib = P->lpBIdx[l][k1];
            const F *Ai = &P->A[i].lpv[ib];
            for (int k2 = k1; k2 < P->lpNCon[l]; k2++) {
                int j = P->lpCon[l][k2], jb = P->lpBIdx[l][k2];
                const F *Aj = &P->A[j].lpv[jb];
                F value;
                ompf_mul(&value, x, iz);
                ompf_mul(&value, &value, Ai);
                ompf_mul(&value, &value, Aj);
                if (i != j) {
                    ompf_add(&N->bMat[i + (size_t)m * j], &N->bMat[i + (size_t)m * j], &value);
                    ompf_add(&N->bMat[j + (size_t)m * i], &N->bMat[j + (size_t)m * i], &value);
                } else {
                    ompf_add(&N->bMat[i + (size_t)m * i], &N->bMat[i + (size_t)m * i], &value);
                }
            }
        }
    }
}

/* ns_jordan_triple_product: ret = A*B*C (sdpa_jordan.cpp:336-340) */
static void triple(DLS *ret, const DLS *A, const DLS *B, const DLS *C, DLS *work) {
    F one = Fv(1.0);
    for (int l = 0; l < A->nsdp; l++) {
        int n = A->n[l];
        gemm_nn(n, n, n, &one, A->b[l], n, B->b[l], n, 0, work->b[l], n);
        gemm_nn(n, n, n, &one, work->b[l], n, C->b[l], n, 0, ret->b[l], n);
    }
    for (int l = 0; l < A->nlp; l++) {
        ompf_mul(&work->lp[l], &A->lp[l], &B->lp[l]);
        ompf_mul(&ret->lp[l], &work->lp[l], &C->lp[l]);
    }
}

/* Lal::getSymmetrize (sdpa_linear.cpp:457-482) */
static void symmetrize(DLS *d) {
    F one = Fv(1.0), half = Fv(0.5);
    for (int l = 0; l < d->nsdp; l++) {
        int n = d->n[l];
        F *a = d->b[l];
        for (int idx = 0; idx < n - 1; idx++) {
            int i1 = idx + idx * n + 1, i2 = idx + (idx + 1) * n, len = n - 1 - idx;
            Raxpy(len, &one, &a[i2], n, &a[i1], 1);
            Rscal(len, &half, &a[i1]);
            for (int t = 0; t < len; t++) a[i2 + (size_t)t * n] = a[i1 + t];
        }
    }
}

/* multi_InnerProductToA (sdpa_dataset.cpp:477-485) */
static void multi_inner(const Prob *P, const DLS *X, F *ret) {
    for (int i = 0; i < P->m; i++) ret[i] = sls_inner(&P->A[i], X);
}
/* multi_plusToA (sdpa_dataset.cpp:488-493) */
static void multi_plus(const Prob *P, const F *y, DLS *ret) {
    dls_zero(ret);
    for (int i = 0; i < P->m; i++) {
        const SLS *a = &P->A[i];
        for (int l = 0; l < a->nsdp; l++) sp_plus_inplace(ret->b[a->s[l].blk], &a->s[l], &y[i]);
        for (int l = 0; l < a->nlp; l++) {
            F t;
            ompf_mul(&t, &a->lpv[l], &y[i]);
            ompf_add(&ret->lp[a->lpi[l]], &ret->lp[a->lpi[l]], &t);
        }
    }
}
/* C + s*dense with C sparse: plus(DLS, DLS, SLS) ret = a + C (scalar NULL) */
static void dls_plus_sls(DLS *r, const DLS *a, const SLS *c) {
    F one = Fv(1.0);
    if (r != a) dls_copy(r, a);
    for (int l = 0; l < c->nsdp; l++) sp_plus_inplace(r->b[c->s[l].blk], &c->s[l], &one);
    for (int l = 0; l < c->nlp; l++) ompf_add(&r->lp[c->lpi[l]], &r->lp[c->lpi[l]], &c->lpv[l]);
}

static F max_abs_vec(const F *v, int n) {
    F r, t;
    ompf_set_zero(&r);
    for (int i = 0; i < n; i++) {
        ompf_abs(&t, &v[i]);
        if (Fgt(&t, &r)) r = t;
    }
    return r;
}
static F max_abs_dls(const DLS *d) {
    F r, t;
    ompf_set_zero(&r);
    for (int l = 0; l < d->nsdp; l++)
        for (int i = 0; i < d->n[l] * d->n[l]; i++) {
            ompf_abs(&t, &d->b[l][i]);
            if (Fgt(&t, &r)) r = t;
        }
    for (int i = 0; i < d->nlp; i++) {
        ompf_abs(&t, &d->lp[i]);
        if (Fgt(&t, &r)) r = t;
    }
    return r;
}

/* Residuals::compute (sdpa_dataset.cpp:625-646) */
static void residuals(Res *R, const Prob *P, const Sol *S) {
    F mone = Fv(-1.0), one = Fv(1.0);
    multi_inner(P, &S->X, R->primalVec);
    for (int i = 0; i < P->m; i++) ompf_mul(&R->primalVec[i], &mone, &R->primalVec[i]);
    for (int i = 0; i < P->m; i++) {
        F t;
        ompf_mul(&t, &one, &P->b[i]);
        ompf_add(&R->primalVec[i], &R->primalVec[i], &t);
    }
    multi_plus(P, S->y, &R->dualMat);
    dls_scale(&R->dualMat, &R->dualMat, &mone);
    dls_plus_sls(&R->dualMat, &R->dualMat, &P->C);
    dls_axpy(&R->dualMat, &R->dualMat, &S->Z, &mone);
    R->normPrimalVec = max_abs_vec(R->primalVec, P->m);
    R->normDualMat = max_abs_dls(&R->dualMat);
}

/* Jal::getInvChol (sdpa_jordan.cpp:123-150); Lal::getCholesky + getInvLowTriangularMatrix */
static int inv_chol(DLS *invChol, const DLS *A, DLS *work) {
    F one = Fv(1.0);
    for (int l = 0; l < A->nsdp; l++) {
        int n = A->n[l];
        F *L = work->b[l];
        for (int i = 0; i < n * n; i++) L[i] = A->b[l][i];
        if (potrf_l(n, L, n) != 0) return 0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < j; i++) ompf_set_zero(&L[i + (size_t)n * j]);
        F *R = invChol->b[l];
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                if (i == j)
                    R[i + (size_t)n * j] = one;
                else
                    ompf_set_zero(&R[i + (size_t)n * j]);
            }
        trsm_llnn(n, n, L, n, R, n);
    }
    for (int l = 0; l < A->nlp; l++) {
        if (ompf_sgn(&A->lp[l]) < 0) return 0;
        F t;
        ompf_sqrt(&t, &A->lp[l]);
        ompf_div(&invChol->lp[l], &one, &t);
    }
    return 1;
}
static int inv_chol_and_inv(DLS *invChol, DLS *inv, const DLS *A, DLS *work) {
    if (!inv_chol(invChol, A, work)) return 0;
    F one = Fv(1.0);
    for (int l = 0; l < A->nsdp; l++) {
        int n = A->n[l];
        for (int i = 0; i < n * n; i++) inv->b[l][i] = invChol->b[l][i];
        trmm_lltn(n, n, invChol->b[l], n, inv->b[l], n);
    }
    for (int l = 0; l < A->nlp; l++) ompf_div(&inv->lp[l], &one, &A->lp[l]);
    return 1;
}

/* ================= dump of Newton inputs / outputs ================= */

static void dump_arr(const char *dir, const char *name, const F *v, int n, int L) {
    char path[1024];
    snprintf(path, sizeof path, "%s/%s.bin", dir, name);
    FILE *fp = fopen(path, "wb");
    if (!fp) return;
    int32_t hdr[2] = {n, L};
    fwrite(hdr, sizeof hdr, 1, fp);
    int32_t *sg = (int32_t *)xcalloc(n, 4), *ex = (int32_t *)xcalloc(n, 4);
    uint32_t *lb = (uint32_t *)xcalloc((size_t)n * L, 4);
    for (int i = 0; i < n; i++) ompf_export32(&v[i], L, &sg[i], &ex[i], &lb[(size_t)i * L]);
    fwrite(sg, 4, n, fp);
    fwrite(ex, 4, n, fp);
    fwrite(lb, 4, (size_t)n * L, fp);
    free(sg);
    free(ex);
    free(lb);
    fclose(fp);
}
static void dump_dls(const char *dir, const char *name, const DLS *d, int L) {
    int tot = d->nlp;
    for (int l = 0; l < d->nsdp; l++) tot += d->n[l] * d->n[l];
    F *v = (F *)xcalloc(tot, sizeof(F));
    int o = 0;
    for (int l = 0; l < d->nsdp; l++)
        for (int i = 0; i < d->n[l] * d->n[l]; i++) v[o++] = d->b[l][i];
    for (int i = 0; i < d->nlp; i++) v[o++] = d->lp[i];
    dump_arr(dir, name, v, tot, L);
    free(v);
}

/* ================= the main loop (sdpa_main.cpp:43-461) ================= */

static const char *phase_str(int v) {
    switch (v) {
    case noINFO: return "noINFO    ";
    case pFEAS: return "pFEAS     ";
    case dFEAS: return "dFEAS     ";
    case pdFEAS: return "pdFEAS    ";
    case pdINF: return "pdINF     ";
    case pFEAS_dINF: return "pFEAS_dINF";
    case pINF_dFEAS: return "pINF_dFEAS";
    case pdOPT: return "pdOPT     ";
    case pUNBD: return "pUNBD     ";
    case dUNBD: return "dUNBD     ";
    }
    return "phase error";
}
/* Phase::reverse with REVERSE_PRIMAL_DUAL (sdpa_parts.cpp:1013-1048) */
static int phase_reverse(int v) {
    switch (v) {
    case pFEAS: return dFEAS;
    case dFEAS: return pFEAS;
    case pFEAS_dINF: return pINF_dFEAS;
    case pINF_dFEAS: return pFEAS_dINF;
    case pUNBD: return dUNBD;
    case dUNBD: return pUNBD;
    }
    return v;
}

typedef struct {
    F objValPrimal, objValDual, rho, etaPrimal, etaDual;
} SolveInfo;

/* Phase::updateCheck (sdpa_parts.cpp:932-1011), REVERSE_PRIMAL_DUAL branch */
static int phase_update(int *phase, const Res *R, const SolveInfo *si, const orc_param *prm) {
    F acc = Fv(prm->epsilonDash), NONZERO = Fv(1.0e-6), one = Fv(1.0), two = Fv(2.0);
    int v = noINFO;
    if (ompf_cmp(&R->normPrimalVec, &acc) <= 0) v = (ompf_cmp(&R->normDualMat, &acc) <= 0) ? pdFEAS : pFEAS;
    if (v == noINFO && ompf_cmp(&R->normDualMat, &acc) <= 0) v = dFEAS;
    *phase = v;
    F thr;
    ompf_add(&thr, &one, &NONZERO);
    if (v == pdFEAS) {
        F a, b, mean, gap, dom, t;
        ompf_abs(&a, &si->objValPrimal);
        ompf_abs(&b, &si->objValDual);
        ompf_add(&mean, &a, &b);
        ompf_div(&mean, &mean, &two);
        ompf_sub(&gap, &si->objValPrimal, &si->objValDual);
        ompf_abs(&gap, &gap);
        dom = Flt(&mean, &one) ? one : mean;
        ompf_div(&t, &gap, &dom);
        if (ompf_cmp_d(&t, prm->epsilonStar) <= 0) {
            *phase = pdOPT;
            return 0;
        }
    }
    if (v == noINFO && Fgt(&si->rho, &thr)) {
        *phase = pdINF;
        return 0;
    }
    if (v == pFEAS) {
        if (ompf_cmp_d(&si->objValPrimal, -prm->upperBound) <= 0) {
            *phase = pUNBD;
            return 0;
        }
        if (Fgt(&si->rho, &thr)) {
            *phase = pFEAS_dINF;
            return 0;
        }
    }
    if (v == dFEAS) {
        if (ompf_cmp_d(&si->objValDual, -prm->lowerBound) >= 0) {
            *phase = dUNBD;
            return 0;
        }
        if (Fgt(&si->rho, &thr)) {
            *phase = pINF_dFEAS;
            return 0;
        }
    }
    return 1;
}

static void fmt(char *buf, int sz, const F *v, int digits) { ompf_get_str(buf, sz, v, digits); }
static void fmt_signed(char *buf, int sz, const F *v, int digits) {
    char t[128];
    ompf_get_str(t, sizeof t, v, digits);
    if (t[0] == '-')
        snprintf(buf, sz, "%s", t);
    else
        snprintf(buf, sz, "+%s", t);
}

static void print_iter(FILE *fp, int it, const F *mu, const F *thP, const F *thD, const SolveInfo *si, const F *aP,
                       const F *aD, const F *beta) {
    if (!fp) return;
    char s1[64], s2[64], s3[64], s4[64], s5[64], s6[64], s7[64], s8[64];
    F t1, t2;
    ompf_neg(&t1, &si->objValDual);
    ompf_neg(&t2, &si->objValPrimal);
    fmt(s1, 64, mu, 2);
    fmt(s2, 64, thD, 2);
    fmt(s3, 64, thP, 2);
    fmt_signed(s4, 64, &t1, 3);
    fmt_signed(s5, 64, &t2, 3);
    fmt(s6, 64, aD, 2);
    fmt(s7, 64, aP, 2);
    fmt(s8, 64, beta, 3);
    fprintf(fp, "%2d %s %s %s %s %s %s %s %s\n", it, s1, s2, s3, s4, s5, s6, s7, s8);
    fflush(fp);
}

typedef struct {
    Prob P;
    orc_param prm;
    Newton N;
    Sol S;
    Res cur, init;
    Work W;
    int m;
} Ctx;

static void newton_alloc(Newton *N, const Prob *P) {
    N->m = P->m;
    N->bMat = (F *)xcalloc((size_t)P->m * P->m, sizeof(F));
    N->gVec = (F *)xcalloc(P->m, sizeof(F));
    N->dy = (F *)xcalloc(P->m, sizeof(F));
    dls_init(&N->dX, P);
    dls_init(&N->dZ, P);
    dls_init(&N->r_zinv, P);
    dls_init(&N->x_rd_zinv, P);
    N->useFormula = (int *)xcalloc((size_t)P->m * (P->nsdp ? P->nsdp : 1), sizeof(int));
}
static void newton_free(Newton *N) {
    free(N->bMat);
    free(N->gVec);
    free(N->dy);
    dls_free(&N->dX);
    dls_free(&N->dZ);
    dls_free(&N->r_zinv);
    dls_free(&N->x_rd_zinv);
    free(N->useFormula);
}

/* Newton::Mehrotra (sdpa_newton.cpp:1375-1414). Returns 0 on Cholesky failure. */
static int mehrotra(Ctx *c, int predictor, const F *mu, const F *beta, int phase, double *t_bmat, double *t_chol,
                    const char *dump_dir, int dump_L) {
    Prob *P = &c->P;
    Newton *N = &c->N;
    Sol *S = &c->S;
    Work *W = &c->W;
    int m = P->m;
    F mone = Fv(-1.0), target;
    /* compute_rMat (sdpa_newton.cpp:821-838) */
    ompf_mul(&target, beta, mu);
    dls_scale(&N->r_zinv, &S->invZ, &target);
    dls_axpy(&N->r_zinv, &N->r_zinv, &S->X, &mone);
    if (!predictor) {
        triple(&W->DLS1, &N->dX, &N->dZ, &S->invZ, &W->DLS2);
        dls_axpy(&N->r_zinv, &N->r_zinv, &W->DLS1, &mone);
    }
    /* Make_gVec (sdpa_newton.cpp:840-895) */
    if (phase == pFEAS || phase == noINFO) {
        if (!predictor) {
            dls_axpy(&W->DLS1, &N->r_zinv, &N->x_rd_zinv, &mone);
        } else {
            triple(&N->x_rd_zinv, &S->X, &c->cur.dualMat, &S->invZ, &W->DLS2);
            dls_axpy(&W->DLS1, &N->r_zinv, &N->x_rd_zinv, &mone);
        }
    } else {
        dls_copy(&W->DLS1, &N->r_zinv);
    }
    multi_inner(P, &W->DLS1, N->gVec);
    for (int i = 0; i < m; i++) ompf_mul(&N->gVec[i], &mone, &N->gVec[i]);
    {
        F one = Fv(1.0), t;
        for (int i = 0; i < m; i++) {
            ompf_mul(&t, &one, &c->cur.primalVec[i]);
            ompf_add(&N->gVec[i], &N->gVec[i], &t);
        }
    }
    if (predictor) {
        double t0 = now_s();
        make_bmat(N, P, S, W);
        double t1 = now_s();
        if (t_bmat) *t_bmat += t1 - t0;
        if (dump_dir) {
            dump_dls(dump_dir, "X", &S->X, dump_L);
            dump_dls(dump_dir, "Z", &S->Z, dump_L);
            dump_dls(dump_dir, "invZ", &S->invZ, dump_L);
            dump_arr(dump_dir, "y", S->y, m, dump_L);
            dump_arr(dump_dir, "Rp", c->cur.primalVec, m, dump_L);
            dump_dls(dump_dir, "Rd", &c->cur.dualMat, dump_L);
            dump_arr(dump_dir, "mu", mu, 1, dump_L);
            dump_arr(dump_dir, "beta", beta, 1, dump_L);
            dump_arr(dump_dir, "B", N->bMat, m * m, dump_L);
            dump_arr(dump_dir, "g", N->gVec, m, dump_L);
            char path[1024];
            snprintf(path, sizeof path, "%s/phase.txt", dump_dir);
            FILE *fp = fopen(path, "w");
            if (fp) {
                fprintf(fp, "%d\n", phase);
                fclose(fp);
            }
        }
        t0 = now_s();
        int info = potrf_l(m, N->bMat, m);
        t1 = now_s();
        if (t_chol) *t_chol += t1 - t0;
        if (info > 0) {
            fprintf(stderr, "cholesky miss condition :: not positive definite :: info = %d\n", info);
            return 0;
        }
        if (dump_dir) dump_arr(dump_dir, "L", N->bMat, m * m, dump_L);
    }
    /* solveSystems (sdpa_linear.cpp:683-695) */
    for (int i = 0; i < m; i++) N->dy[i] = N->gVec[i];
    trsv_ln(m, N->bMat, m, N->dy);
    trsv_lt(m, N->bMat, m, N->dy);
    /* compute_DzMat (sdpa_newton.cpp:1346-1356) */
    multi_plus(P, N->dy, &N->dZ);
    dls_scale(&N->dZ, &N->dZ, &mone);
    if (phase == pFEAS || phase == noINFO) dls_axpy(&N->dZ, &N->dZ, &c->cur.dualMat, NULL);
    /* compute_DxMat (sdpa_newton.cpp:1358-1373) */
    triple(&W->DLS1, &S->X, &N->dZ, &S->invZ, &W->DLS2);
    dls_axpy(&N->dX, &N->r_zinv, &W->DLS1, &mone);
    symmetrize(&N->dX);
    if (predictor && dump_dir) {
        dump_arr(dump_dir, "dy", N->dy, m, dump_L);
        dump_dls(dump_dir, "dX", &N->dX, dump_L);
        dump_dls(dump_dir, "dZ", &N->dZ, dump_L);
    }
    return 1;
}

static int run(Ctx *c, const orc_param *prm, FILE *logf, const char *dump_dir, const int *dump_iters, int n_dump,
               int dump_L, orc_result *res, int stop_after, int warm_only) {
    Prob *P = &c->P;
    int m = P->m, nDim = P->nDim;
    c->prm = *prm;
    newton_alloc(&c->N, P);
    compute_formula(&c->N, P);
    Sol *S = &c->S;
    dls_init(&S->X, P);
    dls_init(&S->Z, P);
    dls_init(&S->invCholX, P);
    dls_init(&S->invCholZ, P);
    dls_init(&S->invZ, P);
    S->y = (F *)xcalloc(m, sizeof(F));
    dls_init(&c->W.DLS1, P);
    dls_init(&c->W.DLS2, P);
    F lam = Fv(prm->lambdaStar), one = Fv(1.0), t;
    /* Solutions::initialize (sdpa_dataset.cpp:35-60) */
    dls_identity(&S->X, &lam);
    dls_identity(&S->Z, &lam);
    ompf_sqrt(&t, &lam);
    ompf_div(&t, &one, &t);
    dls_identity(&S->invCholX, &t);
    dls_identity(&S->invCholZ, &t);
    ompf_div(&t, &one, &lam);
    dls_identity(&S->invZ, &t);
    /* Residuals */
    c->init.primalVec = (F *)xcalloc(m, sizeof(F));
    dls_init(&c->init.dualMat, P);
    residuals(&c->init, P, S);
    c->cur.primalVec = (F *)xcalloc(m, sizeof(F));
    dls_init(&c->cur.dualMat, P);
    for (int i = 0; i < m; i++) c->cur.primalVec[i] = c->init.primalVec[i];
    dls_copy(&c->cur.dualMat, &c->init.dualMat);
    c->cur.normPrimalVec = c->init.normPrimalVec;
    c->cur.normDualMat = c->init.normDualMat;
    /* StepLength, DirectionParameter, Switch, AverageComplementarity, Ratio, SolveInfo, Phase */
    F alphaP = Fv(0.0), alphaD = Fv(0.0);
    F beta = Fv(prm->betaStar);
    int reductionOn = 1;
    F mu_initial, mu_current;
    ompf_mul(&mu_initial, &lam, &lam);
    mu_current = mu_initial;
    F acc = Fv(prm->epsilonDash);
    F thetaP = Flt(&c->init.normPrimalVec, &acc) ? Fv(0.0) : Fv(1.0);
    F thetaD = Flt(&c->init.normDualMat, &acc) ? Fv(0.0) : Fv(1.0);
    SolveInfo si;
    F nD;
    ompf_set_si(&nD, nDim);
    F omega = Fv(prm->omegaStar);
    Fd(&si.rho, 1.0);
    ompf_mul(&si.etaPrimal, &omega, &nD);
    ompf_mul(&si.etaPrimal, &si.etaPrimal, &mu_initial);
    si.etaDual = si.etaPrimal;
    si.objValPrimal = sls_inner(&P->C, &S->X);
    si.objValDual = Rdot(m, P->b, 1, S->y, 1);
    int phase = noINFO;
    phase_update(&phase, &c->init, &si, prm);

    double t_bmat = 0, t_chol = 0;
    double t_start = now_s();
    int pIteration = 0;
    if (logf) fprintf(logf, "   mu      thetaP  thetaD  objP      objD      alphaP  alphaD  beta \n");
    F c09 = Fv(0.9), cone = Fv(1.0), alphaBD = Fv(100.0), inv_alphaBD, xi = Fv(3.0), c095 = Fv(0.95);
    ompf_div(&inv_alphaBD, &cone, &alphaBD);
    F gammaStar = Fv(prm->gammaStar), betaStar = Fv(prm->betaStar), betaBar = Fv(prm->betaBar), nu = Fv(2.0);
    F c1em6 = Fv(1.0e-6), c1em4 = Fv(1.0e-4), five = Fv(5.0), c1em10 = Fv(1.0e-10), zero = Fv(0.0);
    int rc = 0;
    while (phase_update(&phase, &c->cur, &si, prm) && pIteration < prm->maxIteration) {
        if (stop_after >= 0 && pIteration >= stop_after) break;
        /* Switch::MehrotraPredictor, DirectionParameter::MehrotraPredictor (sdpa_parts.cpp:559-567, 605-612) */
        reductionOn = (phase == noINFO || phase == pFEAS || phase == dFEAS);
        if (phase == pdFEAS)
            ompf_set_zero(&beta);
        else
            beta = reductionOn ? betaBar : nu;
        const char *dd = NULL;
        char dpath[1024];
        if (dump_dir)
            for (int q = 0; q < n_dump; q++)
                if (dump_iters[q] == pIteration) {
                    snprintf(dpath, sizeof dpath, "%s/iter%03d", dump_dir, pIteration);
                    char cmd[1100];
                    snprintf(cmd, sizeof cmd, "mkdir -p '%s'", dpath);
                    if (system(cmd) != 0) break;
                    dd = dpath;
                }
        if (!mehrotra(c, 1, &mu_current, &beta, phase, &t_bmat, &t_chol, dd, dump_L)) {
            rc = 1;
            break;
        }
        if (warm_only) {
            /* timing mode: predictor only */
        }
        /* StepLength::MehrotraPredictor (sdpa_parts.cpp:362-410) */
        alphaP = c09;
        alphaD = c09;
        if (phase == noINFO || phase == dFEAS) {
            if (Fgt(&alphaP, &cone)) alphaP = cone;
        } else {
            F inc = sls_inner(&P->C, &c->N.dX);
            if (ompf_sgn(&inc) > 0) {
                if (Fgt(&alphaP, &alphaD)) alphaP = alphaD;
                if (Fgt(&alphaP, &cone)) alphaP = cone;
            }
        }
        if (phase == noINFO || phase == pFEAS) {
            if (Fgt(&alphaD, &cone)) alphaD = cone;
        } else {
            F inc = Rdot(m, P->b, 1, c->N.dy, 1);
            if (ompf_sgn(&inc) < 0) {
                if (Fgt(&alphaD, &alphaP)) alphaD = alphaP;
                if (Fgt(&alphaD, &cone)) alphaD = cone;
            }
        }
        /* DirectionParameter::MehrotraCorrector (sdpa_parts.cpp:614-645) */
        {
            F xv = dls_inner(&S->X, &c->N.dZ), uz = dls_inner(&c->N.dX, &S->Z), uv = dls_inner(&c->N.dX, &c->N.dZ);
            F muT, t1, t2;
            ompf_mul(&t1, &alphaP, &uz);
            ompf_mul(&t2, &alphaD, &xv);
            ompf_add(&t1, &t1, &t2);
            ompf_mul(&t2, &alphaP, &alphaD);
            ompf_mul(&t2, &t2, &uv);
            ompf_add(&t1, &t1, &t2);
            ompf_div(&t1, &t1, &nD);
            ompf_add(&muT, &mu_current, &t1);
            ompf_div(&beta, &muT, &mu_current);
            if (Flt(&beta, &cone)) ompf_mul(&beta, &beta, &beta);
            if (phase == pdFEAS) {
                if (Flt(&beta, &betaStar)) beta = betaStar;
                if (Fgt(&beta, &cone)) beta = cone;
            } else {
                if (Flt(&beta, &betaBar)) beta = betaBar;
            }
        }
        if (phase == pdFEAS && (Fgt(&beta, &five) || Flt(&si.objValPrimal, &si.objValDual))) break;
        mehrotra(c, 0, &mu_current, &beta, phase, NULL, NULL, NULL, dump_L);
        /* StepLength::MehrotraCorrector (sdpa_parts.cpp:412-523) */
        {
            F mnx = jal_min_eigen(&S->invCholX, &c->N.dX), t1;
            ompf_neg(&t1, &mnx);
            if (Fgt(&t1, &inv_alphaBD)) {
                ompf_div(&alphaP, &cone, &mnx);
                ompf_neg(&alphaP, &alphaP);
            } else
                alphaP = alphaBD;
            F mnz = jal_min_eigen(&S->invCholZ, &c->N.dZ);
            ompf_neg(&t1, &mnz);
            if (Fgt(&t1, &inv_alphaBD)) {
                ompf_div(&alphaD, &cone, &mnz);
                ompf_neg(&alphaD, &alphaD);
            } else
                alphaD = alphaBD;
            ompf_mul(&alphaP, &gammaStar, &alphaP);
            ompf_mul(&alphaD, &gammaStar, &alphaD);
            if (phase == noINFO || phase == dFEAS) {
                if (Fgt(&alphaP, &cone)) alphaP = cone;
            } else {
                F inc = sls_inner(&P->C, &c->N.dX);
                if (ompf_sgn(&inc) > 0) {
                    if (Fgt(&alphaP, &alphaD)) alphaP = alphaD;
                    if (Fgt(&alphaP, &cone)) alphaP = cone;
                }
            }
            if (phase == noINFO || phase == pFEAS) {
                if (Fgt(&alphaD, &cone)) alphaD = cone;
            } else {
                F inc = Rdot(m, P->b, 1, c->N.dy, 1);
                if (ompf_sgn(&inc) < 0) {
                    if (Fgt(&alphaD, &alphaP)) alphaD = alphaP;
                    if (Fgt(&alphaD, &cone)) alphaD = cone;
                }
            }
            if (reductionOn && (phase == noINFO || phase == pFEAS || phase == dFEAS)) {
                F xv = dls_inner(&S->X, &c->N.dZ), uz = dls_inner(&c->N.dX, &S->Z),
                  uv = dls_inner(&c->N.dX, &c->N.dZ);
                F thetaMax, muNew, a1, a2, tt;
#define THETAMAX()                                                                                                     \
    do {                                                                                                               \
        ompf_sub(&a1, &cone, &alphaP);                                                                                 \
        ompf_mul(&a1, &a1, &thetaP);                                                                                   \
        ompf_sub(&a2, &cone, &alphaD);                                                                                 \
        ompf_mul(&a2, &a2, &thetaD);                                                                                   \
        thetaMax = Fgt(&a1, &a2) ? a1 : a2;                                                                            \
    } while (0)
#define MUNEW()                                                                                                        \
    do {                                                                                                               \
        ompf_mul(&a1, &alphaP, &uz);                                                                                   \
        ompf_mul(&a2, &alphaD, &xv);                                                                                   \
        ompf_add(&a1, &a1, &a2);                                                                                       \
        ompf_mul(&a2, &alphaP, &alphaD);                                                                               \
        ompf_mul(&a2, &a2, &uv);                                                                                       \
        ompf_add(&a1, &a1, &a2);                                                                                       \
        ompf_div(&a1, &a1, &nD);                                                                                       \
        ompf_add(&muNew, &mu_current, &a1);                                                                            \
    } while (0)
                THETAMAX();
                MUNEW();
                for (;;) {
                    ompf_mul(&tt, &thetaMax, &thetaMax);
                    ompf_mul(&tt, &tt, &mu_initial);
                    ompf_mul(&a1, &xi, &muNew);
                    if (!Fgt(&tt, &a1)) break;
                    F amax = Fgt(&alphaP, &alphaD) ? alphaP : alphaD;
                    ompf_mul(&amax, &c095, &amax);
                    if (Fgt(&alphaP, &amax)) alphaP = amax;
                    if (Fgt(&alphaD, &amax)) alphaD = amax;
                    THETAMAX();
                    MUNEW();
                    if (Flt(&alphaP, &c1em6) && Flt(&alphaD, &c1em6)) break;
                }
#undef THETAMAX
#undef MUNEW
            }
            if (phase == pdFEAS) {
                F oD = Rdot(m, P->b, 1, S->y, 1), oP = sls_inner(&P->C, &S->X);
                F iD = Rdot(m, P->b, 1, c->N.dy, 1), iP = sls_inner(&P->C, &c->N.dX);
                F mr, t2;
                ompf_mul(&iD, &iD, &alphaD);
                ompf_mul(&iP, &iP, &alphaP);
                ompf_sub(&mr, &oD, &oP);
                ompf_sub(&t2, &iP, &iD);
                if (ompf_sgn(&t2) != 0) {
                    ompf_div(&mr, &mr, &t2);
                    if (ompf_sgn(&mr) > 0 && Flt(&mr, &cone)) {
                        ompf_mul(&alphaP, &alphaP, &mr);
                        ompf_mul(&alphaD, &alphaD, &mr);
                    }
                }
            }
        }
        print_iter(logf, pIteration, &mu_current, &thetaP, &thetaD, &si, &alphaP, &alphaD, &beta);
        /* Solutions::update (sdpa_dataset.cpp:120-143) */
        dls_axpy(&S->X, &S->X, &c->N.dX, &alphaP);
        Raxpy(m, &alphaD, c->N.dy, 1, S->y, 1);
        dls_axpy(&S->Z, &S->Z, &c->N.dZ, &alphaD);
        if (Flt(&alphaP, &c1em4) && Flt(&alphaD, &c1em4)) {
            if (logf) fprintf(logf, "Step length is too small. \ncannot move\n");
            rc = 2;
            break;
        }
        if (!inv_chol(&S->invCholX, &S->X, &c->W.DLS1) ||
            !inv_chol_and_inv(&S->invCholZ, &S->invZ, &S->Z, &c->W.DLS2)) {
            if (logf) fprintf(logf, "cannot move\n");
            rc = 3;
            break;
        }
        /* RatioInitResCurrentRes::update (sdpa_parts.cpp:690-697) */
        if (reductionOn) {
            F a;
            ompf_sub(&a, &cone, &alphaP);
            ompf_mul(&a, &a, &thetaP);
            ompf_abs(&thetaP, &a);
            ompf_sub(&a, &cone, &alphaD);
            ompf_mul(&a, &a, &thetaD);
            ompf_abs(&thetaD, &a);
        }
        /* AverageComplementarity::update */
        mu_current = dls_inner(&S->X, &S->Z);
        ompf_div(&mu_current, &mu_current, &nD);
        residuals(&c->cur, P, S);
        /* update_exact */
        if (Fgt(&c->init.normPrimalVec, &c1em10))
            ompf_div(&thetaP, &c->cur.normPrimalVec, &c->init.normPrimalVec);
        else
            thetaP = zero;
        if (Fgt(&c->init.normDualMat, &c1em10))
            ompf_div(&thetaD, &c->cur.normDualMat, &c->init.normDualMat);
        else
            thetaD = zero;
        /* SolveInfo::update(lambda, ...) (sdpa_parts.cpp:774-830) */
        {
            si.objValPrimal = sls_inner(&P->C, &S->X);
            si.objValDual = Rdot(m, P->b, 1, S->y, 1);
            F pr = thetaP, du = thetaD, om = omega, x0z0, xz, x0z, xz0, a1, a2, a3;
            ompf_set_zero(&si.rho);
            ompf_mul(&x0z0, &nD, &mu_initial);
            ompf_mul(&xz, &nD, &mu_current);
            ompf_set_zero(&x0z);
            ompf_set_zero(&xz0);
            for (int l = 0; l < P->nsdp; l++) {
                int n = P->sdpStruct[l];
                for (int i = 0; i < n; i++) {
                    ompf_mul(&a1, &lam, &S->Z.b[l][i * n + i]);
                    ompf_add(&x0z, &x0z, &a1);
                    ompf_mul(&a1, &lam, &S->X.b[l][i * n + i]);
                    ompf_add(&xz0, &xz0, &a1);
                }
            }
            for (int l = 0; l < P->nlp; l++) {
                ompf_mul(&a1, &lam, &S->Z.lp[l]);
                ompf_add(&x0z, &x0z, &a1);
                ompf_mul(&a1, &lam, &S->X.lp[l]);
                ompf_add(&xz0, &xz0, &a1);
            }
            int pok = ompf_cmp(&c->cur.normPrimalVec, &acc) <= 0, dok = ompf_cmp(&c->cur.normDualMat, &acc) <= 0;
            if (pok && Flt(&xz0, &si.etaPrimal)) si.etaPrimal = xz0;
            if (dok && Flt(&x0z, &si.etaDual)) si.etaDual = x0z;
            if (!pok && dok) {
                ompf_sub(&a1, &cone, &pr);
                ompf_mul(&a1, &a1, &om);
                ompf_add(&a1, &pr, &a1);
                ompf_mul(&a1, &a1, &si.etaDual);
                ompf_add(&a1, &a1, &xz);
                ompf_mul(&a2, &pr, &x0z);
                ompf_div(&si.rho, &a2, &a1);
            }
            if (pok && !dok) {
                ompf_sub(&a1, &cone, &du);
                ompf_mul(&a1, &a1, &om);
                ompf_add(&a1, &du, &a1);
                ompf_mul(&a1, &a1, &si.etaPrimal);
                ompf_add(&a1, &a1, &xz);
                ompf_mul(&a2, &du, &xz0);
                ompf_div(&si.rho, &a2, &a1);
            }
            if (!pok && !dok) {
                /* (du xz0 + pr x0z) / ((pr du + om (pr (1-du) + (1-pr) du)) x0z0 + xz) */
                ompf_mul(&a1, &du, &xz0);
                ompf_mul(&a2, &pr, &x0z);
                ompf_add(&a1, &a1, &a2);
                ompf_sub(&a2, &cone, &du);
                ompf_mul(&a2, &pr, &a2);
                ompf_sub(&a3, &cone, &pr);
                ompf_mul(&a3, &a3, &du);
                ompf_add(&a2, &a2, &a3);
                ompf_mul(&a2, &om, &a2);
                ompf_mul(&a3, &pr, &du);
                ompf_add(&a2, &a3, &a2);
                ompf_mul(&a2, &a2, &x0z0);
                ompf_add(&a2, &a2, &xz);
                ompf_div(&si.rho, &a1, &a2);
            }
        }
        pIteration++;
    }
    double t_end = now_s();
    residuals(&c->cur, P, S);
    /* printLastInfo (sdpa_io.cpp:809-994) */
    print_iter(logf, pIteration, &mu_current, &thetaP, &thetaD, &si, &alphaP, &alphaD, &beta);
    int rphase = phase_reverse(phase);
    if (res) {
        memset(res, 0, sizeof(*res));
        res->iterations = pIteration;
        res->phase = phase;
        snprintf(res->phase_name, sizeof res->phase_name, "%s", phase_str(rphase));
        F a, b, mean, gap, relgap, two = Fv(2.0), g2;
        ompf_abs(&a, &si.objValPrimal);
        ompf_abs(&b, &si.objValDual);
        ompf_add(&mean, &a, &b);
        ompf_div(&mean, &mean, &two);
        ompf_sub(&gap, &si.objValPrimal, &si.objValDual);
        ompf_abs(&gap, &gap);
        if (Flt(&mean, &cone))
            relgap = gap;
        else
            ompf_div(&relgap, &gap, &mean);
        ompf_mul(&g2, &mu_current, &nD);
        res->mu = ompf_get_d(&mu_current);
        res->relgap = ompf_get_d(&relgap);
        res->gap = ompf_get_d(&g2);
        F t1, t2;
        ompf_neg(&t1, &si.objValDual);
        ompf_neg(&t2, &si.objValPrimal);
        ompf_get_str(res->objP, sizeof res->objP, &t1, 41);
        ompf_get_str(res->objD, sizeof res->objD, &t2, 41);
        ompf_get_str(res->objP70, sizeof res->objP70, &t1, 70);
        ompf_get_str(res->objD70, sizeof res->objD70, &t2, 70);
        ompf_get_str(res->relgap_s, sizeof res->relgap_s, &relgap, 17);
        ompf_get_str(res->pfeas_s, sizeof res->pfeas_s, &c->cur.normDualMat, 17);
        ompf_get_str(res->dfeas_s, sizeof res->dfeas_s, &c->cur.normPrimalVec, 17);
        res->t_bmat = t_bmat;
        res->t_chol = t_chol;
        res->t_total = t_end - t_start;
    }
    if (logf) {
        char buf[128];
        fprintf(logf, "\nphase.value = %s\n", phase_str(rphase));
        fprintf(logf, "   Iteration = %d\n", pIteration);
        if (res) {
            fprintf(logf, "objValPrimal = %s\n", res->objP);
            fprintf(logf, "objValDual   = %s\n", res->objD);
            fprintf(logf, "relative gap = %s\n", res->relgap_s);
            fprintf(logf, "p.feas.error = %s\n", res->pfeas_s);
            fprintf(logf, "d.feas.error = %s\n", res->dfeas_s);
        }
        ompf_get_str(buf, sizeof buf, &mu_current, 17);
        fprintf(logf, "          mu = %s\n", buf);
        fprintf(logf, "Make bMat    = %.3f s\nCholesky bMat= %.3f s\n", t_bmat, t_chol);
    }
    return rc;
}

static void ctx_free(Ctx *c) {
    newton_free(&c->N);
    dls_free(&c->S.X);
    dls_free(&c->S.Z);
    dls_free(&c->S.invCholX);
    dls_free(&c->S.invCholZ);
    dls_free(&c->S.invZ);
    free(c->S.y);
    dls_free(&c->W.DLS1);
    dls_free(&c->W.DLS2);
    free(c->init.primalVec);
    dls_free(&c->init.dualMat);
    free(c->cur.primalVec);
    dls_free(&c->cur.dualMat);
    free_problem(&c->P);
}

int orc_solve(const char *data_path, const orc_param *param, int prec_override, const char *log_path,
              const char *dump_dir, const int *dump_iters, int n_dump, int dump_L, orc_result *res) {
    orc_param prm;
    if (param)
        prm = *param;
    else
        orc_param_default(&prm);
    if (prec_override > 0) prm.precision = prec_override;
    ompf_set_prec_bits(prm.precision);
    Ctx c;
    memset(&c, 0, sizeof c);
    int rc = read_problem(data_path, &c.P);
    if (rc) return rc;
    FILE *logf = NULL;
    if (log_path) logf = (strcmp(log_path, "-") == 0) ? stdout : fopen(log_path, "w");
    rc = run(&c, &prm, logf, dump_dir, dump_iters, n_dump, dump_L > 0 ? dump_L : 8, res, -1, 0);
    if (logf && logf != stdout) fclose(logf);
    ctx_free(&c);
    return rc;
}

/* read a dump file written by dump_arr into v (count must match) */
static int load_arr(const char *dir, const char *name, F *v, int n) {
    char path[1024];
    snprintf(path, sizeof path, "%s/%s.bin", dir, name);
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    int32_t hdr[2];
    if (fread(hdr, sizeof hdr, 1, fp) != 1 || hdr[0] != n) {
        fclose(fp);
        return -2;
    }
    int L = hdr[1];
    int32_t *sg = (int32_t *)xcalloc(n, 4), *ex = (int32_t *)xcalloc(n, 4);
    uint32_t *lb = (uint32_t *)xcalloc((size_t)n * L, 4);
    int ok = fread(sg, 4, n, fp) == (size_t)n && fread(ex, 4, n, fp) == (size_t)n &&
             fread(lb, 4, (size_t)n * L, fp) == (size_t)n * L;
    fclose(fp);
    if (ok)
        for (int i = 0; i < n; i++) ompf_import32(&v[i], L, sg[i], ex[i], &lb[(size_t)i * L]);
    free(sg);
    free(ex);
    free(lb);
    return ok ? 0 : -3;
}
static int load_dls(const char *dir, const char *name, DLS *d) {
    int tot = d->nlp;
    for (int l = 0; l < d->nsdp; l++) tot += d->n[l] * d->n[l];
    F *v = (F *)xcalloc(tot, sizeof(F));
    int rc = load_arr(dir, name, v, tot);
    if (!rc) {
        int o = 0;
        for (int l = 0; l < d->nsdp; l++)
            for (int i = 0; i < d->n[l] * d->n[l]; i++) d->b[l][i] = v[o++];
        for (int i = 0; i < d->nlp; i++) d->lp[i] = v[o++];
    }
    free(v);
    return rc;
}

/*
 * One predictor Newton step (Make_gVec, Make_bMat, Cholesky, solve, dZ, dX) of
 * the reference on the point stored in in_dir (X, Z, invZ, y, mu, beta,
 * phase.txt as written by a dump), at precision prec_bits; the residuals are
 * recomputed from the point (Residuals::compute).  Outputs go to out_dir in
 * the dump format with out_L interchange limbs.
 */
int orc_newton_step(const char *data_path, int prec_bits, const char *in_dir, const char *out_dir, int out_L) {
    ompf_set_prec_bits(prec_bits);
    Ctx c;
    memset(&c, 0, sizeof c);
    int rc = read_problem(data_path, &c.P);
    if (rc) return rc;
    Prob *P = &c.P;
    int m = P->m;
    newton_alloc(&c.N, P);
    compute_formula(&c.N, P);
    dls_init(&c.S.X, P);
    dls_init(&c.S.Z, P);
    dls_init(&c.S.invCholX, P);
    dls_init(&c.S.invCholZ, P);
    dls_init(&c.S.invZ, P);
    c.S.y = (F *)xcalloc(m, sizeof(F));
    dls_init(&c.W.DLS1, P);
    dls_init(&c.W.DLS2, P);
    c.init.primalVec = (F *)xcalloc(m, sizeof(F));
    dls_init(&c.init.dualMat, P);
    c.cur.primalVec = (F *)xcalloc(m, sizeof(F));
    dls_init(&c.cur.dualMat, P);
    F mu, beta;
    int phase = 0;
    if ((rc = load_dls(in_dir, "X", &c.S.X)) || (rc = load_dls(in_dir, "Z", &c.S.Z)) ||
        (rc = load_dls(in_dir, "invZ", &c.S.invZ)) || (rc = load_arr(in_dir, "y", c.S.y, m)) ||
        (rc = load_arr(in_dir, "mu", &mu, 1)) || (rc = load_arr(in_dir, "beta", &beta, 1))) {
        ctx_free(&c);
        return -100 + rc;
    }
    {
        char path[1024];
        snprintf(path, sizeof path, "%s/phase.txt", in_dir);
        FILE *fp = fopen(path, "r");
        if (!fp || fscanf(fp, "%d", &phase) != 1) phase = 0;
        if (fp) fclose(fp);
    }
    residuals(&c.cur, P, &c.S);
    char cmd[1200];
    snprintf(cmd, sizeof cmd, "mkdir -p '%s'", out_dir);
    if (system(cmd) != 0) {
        ctx_free(&c);
        return -200;
    }
    dump_arr(out_dir, "Rp", c.cur.primalVec, m, out_L);
    dump_dls(out_dir, "Rd", &c.cur.dualMat, out_L);
    int ok = mehrotra(&c, 1, &mu, &beta, phase, NULL, NULL, out_dir, out_L);
    ctx_free(&c);
    return ok ? 0 : 1;
}

int orc_time_hotpath(const char *data_path, const orc_param *param, int prec_override, int warm_iters, int reps,
                     int threads, double *t_bmat, double *t_chol) {
    orc_param prm;
    if (param)
        prm = *param;
    else
        orc_param_default(&prm);
    if (prec_override > 0) prm.precision = prec_override;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    ompf_set_prec_bits(prm.precision);
    Ctx c;
    memset(&c, 0, sizeof c);
    int rc = read_problem(data_path, &c.P);
    if (rc) return rc;
    /* advance warm_iters iterations, then time `reps` Make_bMat + Cholesky on that iterate */
    rc = run(&c, &prm, NULL, NULL, NULL, 0, 8, NULL, warm_iters, 0);
    double tb = 0, tc = 0;
    F *Bsave = (F *)xcalloc((size_t)c.P.m * c.P.m, sizeof(F));
    for (int r = 0; r < reps; r++) {
        double t0 = now_s();
        make_bmat(&c.N, &c.P, &c.S, &c.W);
        double t1 = now_s();
        int info = potrf_l(c.P.m, c.N.bMat, c.P.m);
        double t2 = now_s();
        (void)info;
        tb += t1 - t0;
        tc += t2 - t1;
    }
    free(Bsave);
    *t_bmat = tb / (reps > 0 ? reps : 1);
    *t_chol = tc / (reps > 0 ? reps : 1);
    ctx_free(&c);
    return rc;
}

/* packed sdpg words (L+2 per element: e, sign 0/1, limbs) of a DLS from a raw file */
static int load_packed_dls(const char *path, int L, DLS *d) {
    int64_t tot = d->nlp;
    for (int l = 0; l < d->nsdp; l++) tot += (int64_t)d->n[l] * d->n[l];
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    uint32_t *w = (uint32_t *)xcalloc((size_t)tot * (L + 2), 4);
    size_t got = fread(w, 4, (size_t)tot * (L + 2), fp);
    fclose(fp);
    if (got != (size_t)tot * (L + 2)) {
        free(w);
        return -2;
    }
    int64_t o = 0;
    for (int l = 0; l < d->nsdp; l++)
        for (int64_t i = 0; i < (int64_t)d->n[l] * d->n[l]; i++, o++) {
            const uint32_t *e = &w[o * (L + 2)];
            int z = e[2 + L - 1] == 0;
            ompf_import32(&d->b[l][i], L, z ? 0 : ((e[1] & 1) ? -1 : 1), (int32_t)e[0], e + 2);
        }
    for (int i = 0; i < d->nlp; i++, o++) {
        const uint32_t *e = &w[o * (L + 2)];
        int z = e[2 + L - 1] == 0;
        ompf_import32(&d->lp[i], L, z ? 0 : ((e[1] & 1) ? -1 : 1), (int32_t)e[0], e + 2);
    }
    free(w);
    return 0;
}

/*
 * CPU time of the hot path on a given iterate: Make_bMat on (X, invZ) read from
 * packed files (L limbs), then the reference's Cholesky (potrf_l, the Rpotrf
 * structure) on the leading chol_n x chol_n block of the Schur matrix (chol_n <= 0:
 * all of it).  Used to time a bounded sample of a large workload.
 */
int orc_time_point(const char *data_path, int prec_bits, const char *x_path, const char *invz_path, int L,
                   int chol_n, int threads, double *t_bmat, double *t_chol) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    ompf_set_prec_bits(prec_bits);
    Ctx c;
    memset(&c, 0, sizeof c);
    int rc = read_problem(data_path, &c.P);
    if (rc) return rc;
    Prob *P = &c.P;
    int m = P->m;
    newton_alloc(&c.N, P);
    compute_formula(&c.N, P);
    dls_init(&c.S.X, P);
    dls_init(&c.S.Z, P);
    dls_init(&c.S.invCholX, P);
    dls_init(&c.S.invCholZ, P);
    dls_init(&c.S.invZ, P);
    c.S.y = (F *)xcalloc(m, sizeof(F));
    dls_init(&c.W.DLS1, P);
    dls_init(&c.W.DLS2, P);
    c.init.primalVec = (F *)xcalloc(m, sizeof(F));
    dls_init(&c.init.dualMat, P);
    c.cur.primalVec = (F *)xcalloc(m, sizeof(F));
    dls_init(&c.cur.dualMat, P);
    if ((rc = load_packed_dls(x_path, L, &c.S.X)) || (rc = load_packed_dls(invz_path, L, &c.S.invZ))) {
        ctx_free(&c);
        return -100 + rc;
    }
    double t0 = now_s();
    make_bmat(&c.N, P, &c.S, &c.W);
    double t1 = now_s();
    int n = (chol_n > 0 && chol_n < m) ? chol_n : m;
    (void)potrf_l(n, c.N.bMat, m);
    double t2 = now_s();
    *t_bmat = t1 - t0;
    *t_chol = t2 - t1;
    ctx_free(&c);
    return 0;
}

/* ================= arithmetic probes for tests ================= */

int orc_arith(int prec_bits, int op, int L, int32_t as, int32_t ae, const uint32_t *al, int32_t bs, int32_t be,
              const uint32_t *bl, int32_t *rs, int32_t *re, uint32_t *rl) {
    ompf_set_prec_bits(prec_bits);
    F a, b, r;
    ompf_import32(&a, L, as, ae, al);
    ompf_import32(&b, L, bs, be, bl);
    switch (op) {
    case 0: ompf_add(&r, &a, &b); break;
    case 1: ompf_sub(&r, &a, &b); break;
    case 2: ompf_mul(&r, &a, &b); break;
    case 3:
        if (ompf_sgn(&b) == 0) return -1;
        ompf_div(&r, &a, &b);
        break;
    case 4:
        if (ompf_sgn(&a) < 0) return -1;
        ompf_sqrt(&r, &a);
        break;
    default: return -2;
    }
    ompf_export32(&r, 2 * (ompf_prec_limbs() + 1), rs, re, rl);
    return 0;
}

int orc_parse(int prec_bits, const char *text, int L, int32_t *rs, int32_t *re, uint32_t *rl) {
    ompf_set_prec_bits(prec_bits);
    F r;
    if (ompf_set_str(&r, text) <= 0) return -1;
    ompf_export32(&r, L, rs, re, rl);
    return 0;
}

int orc_format(int prec_bits, int L, int32_t s, int32_t e, const uint32_t *l, int digits, char *buf, int bufsz) {
    ompf_set_prec_bits(prec_bits);
    F a;
    ompf_import32(&a, L, s, e, l);
    ompf_get_str(buf, bufsz, &a, digits);
    return 0;
}

/* ================= structure and factorisation probes for tests ================= */

/*
 * The formula table of Newton::computeFormula_SDP (sdpa_newton.cpp:694-819) for a
 * problem read at prec_bits: out[i * nsdp + l] in {1, 2, 3}, 0 where A_i has no block l.
 */
int orc_formula(const char *data_path, int prec_bits, int *m, int *nsdp, int *out, int cap) {
    ompf_set_prec_bits(prec_bits);
    Prob P;
    int rc = read_problem(data_path, &P);
    if (rc) return rc;
    *m = P.m;
    *nsdp = P.nsdp;
    size_t need = (size_t)P.m * P.nsdp;
    if (need > (size_t)cap) {
        free_problem(&P);
        return -10;
    }
    Newton N;
    memset(&N, 0, sizeof N);
    N.m = P.m;
    N.useFormula = (int *)xcalloc(need ? need : 1, sizeof(int));
    compute_formula(&N, &P);
    memcpy(out, N.useFormula, need * sizeof(int));
    free(N.useFormula);
    free_problem(&P);
    return 0;
}

/*
 * Rpotrf("Lower") (mplapack/Rpotrf.cpp:32-154 through potrf_l) of an n x n column-major
 * matrix given in the packed exchange format of the engine (L+2 words per element:
 * exponent, sign 0/1, L limbs); *info as Rpotrf2 reports it (0, or the first
 * non-positive pivot, 1-based).  lo (may be NULL) receives the factor's lower
 * triangle in the same format (strict upper part zero).
 */
int orc_potrf(int prec_bits, int n, int L, const uint32_t *a, uint32_t *lo, int *info) {
    ompf_set_prec_bits(prec_bits);
    F *A = (F *)xcalloc((size_t)n * n, sizeof(F));
    for (size_t t = 0; t < (size_t)n * n; t++) {
        const uint32_t *w = a + t * (size_t)(L + 2);
        int zero = 1;
        for (int k = 0; k < L; k++)
            if (w[2 + k]) zero = 0;
        ompf_import32(&A[t], L, zero ? 0 : ((w[1] & 1u) ? -1 : 1), (int32_t)w[0], w + 2);
    }
    *info = potrf_l(n, A, n);
    if (lo) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                uint32_t *w = lo + ((size_t)i + (size_t)n * j) * (size_t)(L + 2);
                int32_t s = 0, e = 0;
                if (i >= j) {
                    ompf_export32(&A[i + (size_t)n * j], L, &s, &e, w + 2);
                } else {
                    for (int k = 0; k < L; k++) w[2 + k] = 0;
                }
                w[0] = (uint32_t)(s == 0 ? -(1 << 29) : e);
                w[1] = s < 0 ? 1u : 0u;
            }
    }
    free(A);
    return 0;
}

/*
 * Bounded CPU sample of Make_bMat for a problem whose SDP block 0 takes formula F1 for every
 * constraint (compute_bMat_dense_SDP, sdpa_newton.cpp:1007-1096): for nsample constraints spread over
 * the block, the F1 precompute work1 = A_i invZ, work2 = X work1 (the reference's OpenMP Rgemm NN,
 * `threads` threads) and the inner products A_j . work2 with nsample partners each (serial, as the
 * reference's Rdot).  Returns seconds per constraint's precompute and per pair's inner product, from
 * which the caller scales to the whole Make_bMat.
 */
int orc_time_f1_sample(const char *data_path, int prec_bits, const char *x_path, const char *invz_path, int L,
                       int nsample, int threads, double *t_per_con, double *t_per_pair) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    ompf_set_prec_bits(prec_bits);
    Ctx c;
    memset(&c, 0, sizeof c);
    int rc = read_problem(data_path, &c.P);
    if (rc) return rc;
    Prob *P = &c.P;
    if (P->nsdp < 1 || P->sdpNCon[0] < 1) {
        free_problem(P);
        return -30;
    }
    dls_init(&c.S.X, P);
    dls_init(&c.S.invZ, P);
    dls_init(&c.W.DLS1, P);
    dls_init(&c.W.DLS2, P);
    if ((rc = load_packed_dls(x_path, L, &c.S.X)) || (rc = load_packed_dls(invz_path, L, &c.S.invZ))) {
        ctx_free(&c);
        return -100 + rc;
    }
    const int n = P->sdpStruct[0], nc = P->sdpNCon[0];
    const F *X = c.S.X.b[0], *invZ = c.S.invZ.b[0];
    F *work1 = c.W.DLS1.b[0], *work2 = c.W.DLS2.b[0];
    F one = Fv(1.0), acc;
    ompf_set_zero(&acc);
    double tg = 0, td = 0;
    long pairs = 0;
    for (int s = 0; s < nsample; s++) {
        int k1 = (int)((long)s * nc / nsample);
        int i = P->sdpCon[0][k1], ib = P->sdpBIdx[0][k1];
        double t0 = now_s();
        sp_mul_dense(work1, &P->A[i].s[ib], invZ);
        gemm_nn(n, n, n, &one, X, n, work1, n, 0, work2, n);
        double t1 = now_s();
        for (int q = 0; q < nsample; q++) {
            int k2 = (int)((long)q * nc / nsample);
            F v = sp_inner(&P->A[P->sdpCon[0][k2]].s[P->sdpBIdx[0][k2]], work2);
            ompf_add(&acc, &acc, &v);
            pairs++;
        }
        double t2 = now_s();
        tg += t1 - t0;
        td += t2 - t1;
    }
    *t_per_con = tg / nsample;
    *t_per_pair = pairs ? td / pairs : 0;
    ctx_free(&c);
    return 0;
}
