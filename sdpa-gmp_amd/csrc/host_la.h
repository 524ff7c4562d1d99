// host_la.h — host-side multiple-precision linear algebra for the IPM control.
//
// The step-length eigen routines and the update of the iterate stay on the host
// (as in the reference, sdpa_parts.cpp:298-523 and sdpa_dataset.cpp:96-143),
// in the same fixed-limb arithmetic as the device (mpx::mp<L>).
//   - potrf_lower / inv_lower / trmm_lltn: Jal::getInvChol / getInvCholAndInv
//     (sdpa_jordan.cpp:123-171): Rpotrf + Rtrsm(L,L,N,N) with identity + Rtrmm(L,L,T,N);
//   - lanczos_min_eig: Lal::getMinEigen (sdpa_linear.cpp:27-124), control flow as in the
//     reference; its tridiagonal eigenproblem (Rsteqr there) is solved by implicit QL
//     with Wilkinson shifts;
//   - jacobi_min_eig: Lal::getMinEigenValue (Rsyev "N","L", sdpa_linear.cpp:126-158)
//     for blocks <= 20, by cyclic Jacobi rotations.
#pragma once
#include <algorithm>
#include <vector>

#include "mpx.h"

namespace sdpg {
namespace hla {

using mpx::mp;

template <int L>
struct M {  // helpers on scalars
    static mp<L> d(double x) { return mpx::from_double<L>(x); }
    static mp<L> i(long long x) { return mpx::from_int<L>(x); }
};

template <int L>
inline mp<L> operator+(const mp<L> &a, const mp<L> &b) { return mpx::add<L>(a, b); }
template <int L>
inline mp<L> operator-(const mp<L> &a, const mp<L> &b) { return mpx::sub<L>(a, b); }
template <int L>
inline mp<L> operator*(const mp<L> &a, const mp<L> &b) { return mpx::mul<L>(a, b); }
template <int L>
inline mp<L> operator/(const mp<L> &a, const mp<L> &b) { return mpx::div<L>(a, b); }
template <int L>
inline mp<L> operator-(const mp<L> &a) { return mpx::neg<L>(a); }
template <int L>
inline bool operator<(const mp<L> &a, const mp<L> &b) { return mpx::cmp<L>(a, b) < 0; }
template <int L>
inline bool operator>(const mp<L> &a, const mp<L> &b) { return mpx::cmp<L>(a, b) > 0; }
template <int L>
inline bool operator<=(const mp<L> &a, const mp<L> &b) { return mpx::cmp<L>(a, b) <= 0; }
template <int L>
inline bool operator>=(const mp<L> &a, const mp<L> &b) { return mpx::cmp<L>(a, b) >= 0; }
template <int L>
inline mp<L> abs(const mp<L> &a) { return mpx::absv<L>(a); }
template <int L>
inline mp<L> sqrt(const mp<L> &a) { return mpx::sqrtv<L>(a); }

template <int L>
mp<L> dot(int n, const mp<L> *x, int incx, const mp<L> *y, int incy) {
    mp<L> s = mpx::zero<L>();
    for (int i = 0; i < n; i++) s = s + x[(size_t)i * incx] * y[(size_t)i * incy];
    return s;
}

// Rpotrf lower (unblocked right-looking; same flops as the blocked reference); returns info
template <int L>
int potrf_lower(int n, mp<L> *a) {
    for (int j = 0; j < n; j++) {
        mp<L> d = a[j + (size_t)n * j];
        if (mpx::sgn<L>(d) <= 0) return j + 1;
        d = sqrt(d);
        a[j + (size_t)n * j] = d;
        mp<L> r = mpx::recip<L>(d);
        for (int i = j + 1; i < n; i++) a[i + (size_t)n * j] = r * a[i + (size_t)n * j];
#pragma omp parallel for schedule(static) if (n > 64)
        for (int c = j + 1; c < n; c++) {
            mp<L> acj = a[c + (size_t)n * j];
            for (int i = c; i < n; i++) a[i + (size_t)n * c] = a[i + (size_t)n * c] - a[i + (size_t)n * j] * acj;
        }
    }
    for (int j = 0; j < n; j++)
        for (int i = 0; i < j; i++) a[i + (size_t)n * j] = mpx::zero<L>();
    return 0;
}

// R = inv(Lw) for lower-triangular Lw (Rtrsm Left/Lower/NoTrans on the identity)
template <int L>
void inv_lower(int n, const mp<L> *Lw, mp<L> *R) {
    for (size_t k = 0; k < (size_t)n * n; k++) R[k] = mpx::zero<L>();
#pragma omp parallel for schedule(dynamic) if (n > 64)
    for (int j = 0; j < n; j++) {
        R[j + (size_t)n * j] = M<L>::i(1);
        for (int k = j; k < n; k++) {
            mp<L> bkj = R[k + (size_t)n * j];
            if (mpx::is_zero<L>(bkj)) continue;
            bkj = bkj / Lw[k + (size_t)n * k];
            R[k + (size_t)n * j] = bkj;
            for (int i = k + 1; i < n; i++) R[i + (size_t)n * j] = R[i + (size_t)n * j] - bkj * Lw[i + (size_t)n * k];
        }
    }
}

// B := A^T B, A lower triangular (Rtrmm Left/Lower/Transpose/NonUnit)
template <int L>
void trmm_lltn(int n, const mp<L> *A, mp<L> *B) {
#pragma omp parallel for schedule(static) if (n > 64)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            mp<L> t = B[i + (size_t)n * j] * A[i + (size_t)n * i];
            for (int k = i + 1; k < n; k++) t = t + A[k + (size_t)n * i] * B[k + (size_t)n * j];
            B[i + (size_t)n * j] = t;
        }
}

// C = A * B (n x n)
template <int L>
void gemm_nn(int n, const mp<L> *A, const mp<L> *B, mp<L> *C) {
#pragma omp parallel for schedule(static) if (n > 32)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            mp<L> s = mpx::zero<L>();
            for (int k = 0; k < n; k++) s = s + A[i + (size_t)n * k] * B[k + (size_t)n * j];
            C[i + (size_t)n * j] = s;
        }
}

// smallest eigenvalue of the symmetric matrix whose lower triangle is in a, cyclic Jacobi
template <int L>
mp<L> jacobi_min_eig(int n, const mp<L> *a) {
    std::vector<mp<L>> A((size_t)n * n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) A[i + (size_t)n * j] = (i >= j) ? a[i + (size_t)n * j] : a[j + (size_t)n * i];
    const mp<L> one = M<L>::i(1), two = M<L>::i(2);
    const mp<L> eps2 = mpx::ldexp2<L>(one, -2 * (32 * L + 8));
    for (int sweep = 0; sweep < 100; sweep++) {
        mp<L> off = mpx::zero<L>(), dg = mpx::zero<L>();
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                mp<L> t = A[i + (size_t)n * j] * A[i + (size_t)n * j];
                if (i != j)
                    off = off + t;
                else
                    dg = dg + t;
            }
        if (mpx::is_zero<L>(off) || off <= dg * eps2) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                mp<L> apq = A[p + (size_t)n * q];
                if (mpx::is_zero<L>(apq)) continue;
                mp<L> app = A[p + (size_t)n * p], aqq = A[q + (size_t)n * q];
                mp<L> th = (aqq - app) / (two * apq);
                mp<L> t = one / (abs(th) + sqrt(th * th + one));
                if (mpx::sgn<L>(th) < 0) t = -t;
                mp<L> c = one / sqrt(t * t + one), s = t * c, tau = s / (one + c);
                mp<L> tap = t * apq;
                A[p + (size_t)n * p] = app - tap;
                A[q + (size_t)n * q] = aqq + tap;
                A[p + (size_t)n * q] = mpx::zero<L>();
                A[q + (size_t)n * p] = mpx::zero<L>();
                for (int k = 0; k < n; k++) {
                    if (k == p || k == q) continue;
                    mp<L> akp = A[k + (size_t)n * p], akq = A[k + (size_t)n * q];
                    mp<L> np = akp - s * (akq + tau * akp);
                    mp<L> nq = akq + s * (akp - tau * akq);
                    A[k + (size_t)n * p] = np;
                    A[p + (size_t)n * k] = np;
                    A[k + (size_t)n * q] = nq;
                    A[q + (size_t)n * k] = nq;
                }
            }
    }
    mp<L> mn = A[0];
    for (int i = 1; i < n; i++)
        if (A[i + (size_t)n * i] < mn) mn = A[i + (size_t)n * i];
    return mn;
}

// Symmetric tridiagonal (d[0..n), e[0..n-1)) -> ascending eigenvalues in d, eigenvectors in z
// (column-major, leading dimension ldz).  Implicit QL, Wilkinson shift.  Returns 0 or >0.
template <int L>
int tridiag_ql(int n, mp<L> *d, const mp<L> *e_in, mp<L> *z, int ldz) {
    std::vector<mp<L>> e(n + 1, mpx::zero<L>());
    for (int i = 0; i < n - 1; i++) e[i] = e_in[i];
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) z[i + (size_t)j * ldz] = (i == j) ? M<L>::i(1) : mpx::zero<L>();
    const mp<L> one = M<L>::i(1), two = M<L>::i(2);
    const mp<L> eps = mpx::ldexp2<L>(one, -32 * L);
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        for (;;) {
            for (m = l; m < n - 1; m++) {
                mp<L> dd = abs(d[m]) + abs(d[m + 1]);
                if (abs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (iter++ == 60 * n) return l + 1;
            mp<L> g = (d[l + 1] - d[l]) / (two * e[l]);
            mp<L> r = sqrt(g * g + one);
            if (mpx::sgn<L>(g) < 0) r = -r;
            g = d[m] - d[l] + e[l] / (g + r);
            mp<L> s = one, c = one, p = mpx::zero<L>();
            bool early = false;
            for (int i = m - 1; i >= l; i--) {
                mp<L> f = s * e[i], b = c * e[i];
                r = sqrt(f * f + g * g);
                e[i + 1] = r;
                if (mpx::is_zero<L>(r)) {
                    d[i + 1] = d[i + 1] - p;
                    e[m] = mpx::zero<L>();
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + two * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; k++) {
                    mp<L> &zk1 = z[k + (size_t)(i + 1) * ldz];
                    mp<L> &zk = z[k + (size_t)i * ldz];
                    f = zk1;
                    zk1 = s * zk + c * f;
                    zk = c * zk - s * f;
                }
            }
            if (early) continue;
            d[l] = d[l] - p;
            e[l] = g;
            e[m] = mpx::zero<L>();
        }
    }
    for (int i = 0; i < n - 1; i++) {
        int k = i;
        for (int j = i + 1; j < n; j++)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            std::swap(d[k], d[i]);
            for (int q = 0; q < n; q++) std::swap(z[q + (size_t)i * ldz], z[q + (size_t)k * ldz]);
        }
    }
    return 0;
}

// x := L^T x  (Rtrmv L,T,N)
template <int L>
void trmv_lt(int n, const mp<L> *A, mp<L> *x) {
    for (int j = 0; j < n; j++) {
        mp<L> t = x[j] * A[j + (size_t)n * j];
        for (int i = j + 1; i < n; i++) t = t + A[i + (size_t)n * j] * x[i];
        x[j] = t;
    }
}
// x := L x  (Rtrmv L,N,N)
template <int L>
void trmv_ln(int n, const mp<L> *A, mp<L> *x) {
    for (int j = n - 1; j >= 0; j--) {
        if (mpx::is_zero<L>(x[j])) continue;
        mp<L> t = x[j];
        for (int i = n - 1; i > j; i--) x[i] = x[i] + t * A[i + (size_t)n * j];
        x[j] = x[j] * A[j + (size_t)n * j];
    }
}
// y = A x (Rgemv N)
template <int L>
void gemv_n(int n, const mp<L> *A, const mp<L> *x, mp<L> *y) {
    for (int i = 0; i < n; i++) y[i] = mpx::zero<L>();
    for (int j = 0; j < n; j++) {
        if (mpx::is_zero<L>(x[j])) continue;
        mp<L> t = x[j];
        for (int i = 0; i < n; i++) y[i] = y[i] + t * A[i + (size_t)n * j];
    }
}

// Lal::getMinEigen (sdpa_linear.cpp:27-124): Lanczos estimate of min eig of lMat xMat lMat^T
template <int L>
mp<L> lanczos_min_eig(int nDim, const mp<L> *lMat, const mp<L> *xMat) {
    using MM = M<L>;
    mp<L> min = MM::d(1.0e+51), min_old = MM::d(1.0e+52), min_min = MM::d(1.0e+50);
    mp<L> error = MM::d(1.0e+10), one = MM::i(1);
    std::vector<mp<L>> diagVec(nDim, min_min), diagVec2(nDim, mpx::zero<L>()), q(nDim, mpx::zero<L>()),
        qold(nDim), r(nDim, one), w(nDim), tmp(nDim), out(nDim), bb(nDim), Q((size_t)nDim * nDim);
    mp<L> nD = MM::i(nDim);
    mp<L> beta = sqrt(nD);
    const mp<L> sqrtn10 = sqrt(nD) + MM::i(10), sqrtn9 = sqrt(MM::i(nDim + 9));
    const mp<L> c16 = MM::d(1.0e-16), c5 = MM::d(1.0e-5), c8 = MM::d(1.0e-8), c2 = MM::d(1.0e-2), c4 = MM::d(1.0e-4);
    int k = 0, kk = 0;
    for (;;) {
        mp<L> kf = MM::i(k);
        if (!(k < nDim)) break;
        if (!(kf < sqrtn10)) break;
        if (!(beta > c16)) break;
        bool condA = abs(min - min_old) > c5 * abs(min) + c8;
        bool condB = abs(error * beta) > c2 * abs(min) + c4;
        if (!(condA || condB)) break;
        qold = q;
        mp<L> value = one / beta;
        for (int i = 0; i < nDim; i++) q[i] = value * r[i];
        w = q;
        trmv_lt<L>(nDim, lMat, w.data());
        gemv_n<L>(nDim, xMat, w.data(), tmp.data());
        w = tmp;
        trmv_ln<L>(nDim, lMat, w.data());
        mp<L> alpha = dot<L>(nDim, q.data(), 1, w.data(), 1);
        diagVec[k] = alpha;
        mp<L> ma = -alpha, mb = -beta;
        for (int i = 0; i < nDim; i++) r[i] = w[i] + ma * q[i];
        for (int i = 0; i < nDim; i++) r[i] = r[i] + mb * qold[i];
        if (MM::i(kk) >= sqrt(MM::i(k)) || k == nDim - 1 || kf > sqrtn9) {
            kk = 0;
            out = diagVec;
            bb = diagVec2;
            out[nDim - 1] = diagVec[k];
            bb[nDim - 1] = mpx::zero<L>();
            int info = tridiag_ql<L>(k + 1, out.data(), bb.data(), Q.data(), nDim);
            if (info > 0) break;
            min_old = min;
            min = out[0];
            error = Q[k];
        }
        value = dot<L>(nDim, r.data(), 1, r.data(), 1);
        beta = sqrt(value);
        diagVec2[k] = beta;
        ++k;
        ++kk;
    }
    return min - abs(error * beta);
}

}  // namespace hla
}  // namespace sdpg
