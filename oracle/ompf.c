/*
 * ompf — GMP-mpf-semantics multiple precision for the CPU oracle.
 * TEST INFRASTRUCTURE ONLY (see ompf.h).
 *
 * Restates the published mpf algorithms of GMP 6.2.1 (mpf/mul.c,
 * mpf/add.c, mpf/sub.c, mpf/div.c, mpf/sqrt.c, mpf/set_str.c): limb-granular
 * exponent, truncation to prec+1 limbs.  The big-integer kernels below
 * (schoolbook multiply, Knuth algorithm D division, Newton integer square
 * root) are textbook methods written for this file.
 */
#include "ompf.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static int PREC = 5; /* limbs, GMP's _mp_prec */

void ompf_set_prec_bits(int bits) {
    int b = bits < 53 ? 53 : bits;
    PREC = (b + 127) / 64;
    if (PREC + 1 > OMPF_MAXL) {
        fprintf(stderr, "ompf: precision %d bits exceeds OMPF_MAXL\n", bits);
        abort();
    }
}
int ompf_prec_limbs(void) { return PREC; }
int ompf_prec_bits(void) { return PREC * 64 - 64; }

/* ---------------- big-integer helpers (little-endian uint64 limbs) ---------------- */

static uint64_t bn_add_n(uint64_t *r, const uint64_t *a, const uint64_t *b, int n) {
    uint64_t c = 0;
    for (int i = 0; i < n; i++) {
        u128 s = (u128)a[i] + b[i] + c;
        r[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    return c;
}
static uint64_t bn_sub_n(uint64_t *r, const uint64_t *a, const uint64_t *b, int n) {
    uint64_t br = 0;
    for (int i = 0; i < n; i++) {
        uint64_t ai = a[i], bi = b[i];
        uint64_t d = ai - bi - br;
        br = (ai < bi) || (ai == bi && br) ? 1 : 0;
        r[i] = d;
    }
    return br;
}
static int bn_cmp_n(const uint64_t *a, const uint64_t *b, int n) {
    for (int i = n - 1; i >= 0; i--) {
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    }
    return 0;
}
/* r[0..an+bn) = a*b */
static void bn_mul(uint64_t *r, const uint64_t *a, int an, const uint64_t *b, int bn) {
    memset(r, 0, sizeof(uint64_t) * (an + bn));
    for (int j = 0; j < bn; j++) {
        uint64_t c = 0, bj = b[j];
        for (int i = 0; i < an; i++) {
            u128 t = (u128)a[i] * bj + r[i + j] + c;
            r[i + j] = (uint64_t)t;
            c = (uint64_t)(t >> 64);
        }
        r[an + j] = c;
    }
}
/* a[0..n) *= m, returns carry limb */
static uint64_t bn_mul_1(uint64_t *a, int n, uint64_t m) {
    uint64_t c = 0;
    for (int i = 0; i < n; i++) {
        u128 t = (u128)a[i] * m + c;
        a[i] = (uint64_t)t;
        c = (uint64_t)(t >> 64);
    }
    return c;
}
/* a[0..n) /= d, returns remainder */
static uint64_t bn_divrem_1(uint64_t *a, int n, uint64_t d) {
    u128 r = 0;
    for (int i = n - 1; i >= 0; i--) {
        u128 cur = (r << 64) | a[i];
        a[i] = (uint64_t)(cur / d);
        r = cur % d;
    }
    return (uint64_t)r;
}
static int bn_clz(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

/*
 * Knuth algorithm D: q[0..nn-dn] = floor(n / d), d[dn-1] != 0, nn >= dn.
 * n and d are not modified.
 */
static void bn_divq(uint64_t *q, const uint64_t *n, int nn, const uint64_t *d, int dn) {
    if (dn == 1) {
        memcpy(q, n, sizeof(uint64_t) * nn);
        bn_divrem_1(q, nn, d[0]);
        return;
    }
    uint64_t un[2 * OMPF_MAXL * 4 + 8], vn[2 * OMPF_MAXL * 4 + 8];
    int s = bn_clz(d[dn - 1]);
    for (int i = dn - 1; i > 0; i--) vn[i] = s ? (d[i] << s) | (d[i - 1] >> (64 - s)) : d[i];
    vn[0] = d[0] << s;
    un[nn] = s ? n[nn - 1] >> (64 - s) : 0;
    for (int i = nn - 1; i > 0; i--) un[i] = s ? (n[i] << s) | (n[i - 1] >> (64 - s)) : n[i];
    un[0] = n[0] << s;
    for (int j = nn - dn; j >= 0; j--) {
        u128 num = ((u128)un[j + dn] << 64) | un[j + dn - 1];
        u128 qhat = num / vn[dn - 1];
        u128 rhat = num - qhat * vn[dn - 1];
        while (qhat >> 64 || qhat * vn[dn - 2] > ((rhat << 64) | un[j + dn - 2])) {
            qhat--;
            rhat += vn[dn - 1];
            if (rhat >> 64) break;
        }
        /* multiply and subtract */
        uint64_t borrow = 0, carry = 0;
        for (int i = 0; i < dn; i++) {
            u128 p = qhat * vn[i] + carry;
            carry = (uint64_t)(p >> 64);
            uint64_t pl = (uint64_t)p;
            uint64_t t = un[i + j] - pl - borrow;
            borrow = (un[i + j] < pl) || (un[i + j] == pl && borrow) ? 1 : 0;
            un[i + j] = t;
        }
        uint64_t top = un[j + dn];
        uint64_t t = top - carry - borrow;
        int neg = (top < carry) || (top == carry && borrow);
        un[j + dn] = t;
        if (neg) { /* add back */
            qhat--;
            uint64_t c = 0;
            for (int i = 0; i < dn; i++) {
                u128 sm = (u128)un[i + j] + vn[i] + c;
                un[i + j] = (uint64_t)sm;
                c = (uint64_t)(sm >> 64);
            }
            un[j + dn] += c;
        }
        q[j] = (uint64_t)qhat;
    }
}

/* integer square root floor(sqrt(N)), N has nn limbs (top limb non-zero); result in r (returns limbs) */
static int bn_isqrt(uint64_t *r, const uint64_t *N, int nn) {
    int bits = nn * 64 - bn_clz(N[nn - 1]);
    int xb = (bits + 1) / 2 + 1; /* x0 = 2^xb > sqrt(N) */
    int xn = (xb + 64) / 64;
    uint64_t x[4 * OMPF_MAXL + 4], y[4 * OMPF_MAXL + 4], qt[4 * OMPF_MAXL + 4];
    memset(x, 0, sizeof(x));
    x[xb / 64] = (uint64_t)1 << (xb % 64);
    for (int it = 0; it < 4000; it++) {
        int xl = xn;
        while (xl > 1 && x[xl - 1] == 0) xl--;
        /* y = (x + N/x) / 2 */
        memset(qt, 0, sizeof(qt));
        int ql;
        if (nn >= xl) {
            bn_divq(qt, N, nn, x, xl);
            ql = nn - xl + 1;
        } else {
            ql = 1;
            qt[0] = 0;
        }
        int m = (ql > xl ? ql : xl) + 1;
        memset(y, 0, sizeof(uint64_t) * m);
        uint64_t a[4 * OMPF_MAXL + 4], b[4 * OMPF_MAXL + 4];
        memset(a, 0, sizeof(uint64_t) * m);
        memset(b, 0, sizeof(uint64_t) * m);
        memcpy(a, x, sizeof(uint64_t) * xl);
        memcpy(b, qt, sizeof(uint64_t) * ql);
        bn_add_n(y, a, b, m);
        for (int i = 0; i < m; i++) y[i] = (y[i] >> 1) | (i + 1 < m ? y[i + 1] << 63 : 0);
        /* if y >= x stop */
        int mm = m > xn ? m : xn;
        uint64_t xa[4 * OMPF_MAXL + 4], ya[4 * OMPF_MAXL + 4];
        memset(xa, 0, sizeof(uint64_t) * mm);
        memset(ya, 0, sizeof(uint64_t) * mm);
        memcpy(xa, x, sizeof(uint64_t) * xn);
        memcpy(ya, y, sizeof(uint64_t) * m);
        if (bn_cmp_n(ya, xa, mm) >= 0) break;
        memset(x, 0, sizeof(x));
        memcpy(x, y, sizeof(uint64_t) * m);
        xn = m;
    }
    int xl = xn;
    while (xl > 1 && x[xl - 1] == 0) xl--;
    memcpy(r, x, sizeof(uint64_t) * xl);
    return xl;
}

/* ---------------- ompf core ---------------- */

void ompf_set_zero(ompf *r) {
    r->size = 0;
    r->exp = 0;
}

/* value = t[0..tn) * B^(exp - tn); strip, truncate to prec+1 limbs, store */
static void from_buf(ompf *r, const uint64_t *t, int tn, int32_t exp, int neg) {
    while (tn > 0 && t[tn - 1] == 0) {
        tn--;
        exp--;
    }
    if (tn == 0) {
        ompf_set_zero(r);
        return;
    }
    int keep = PREC + 1;
    const uint64_t *p = t;
    if (tn > keep) {
        p += tn - keep;
        tn = keep;
    }
    while (p[0] == 0) {
        p++;
        tn--;
    }
    memmove(r->d, p, sizeof(uint64_t) * tn);
    r->size = neg ? -tn : tn;
    r->exp = exp;
}

/* value = M * 2^e2 with M = m[0..mn) */
static void from_bin(ompf *r, const uint64_t *m, int mn, long e2, int neg) {
    long q = e2 >= 0 ? e2 / 64 : -((-e2 + 63) / 64);
    int rem = (int)(e2 - q * 64);
    uint64_t t[8 * OMPF_MAXL + 8];
    for (int i = 0; i <= mn; i++) {
        uint64_t lo = i < mn ? m[i] : 0;
        uint64_t lower = i > 0 ? m[i - 1] : 0;
        t[i] = rem ? (lo << rem) | (lower >> (64 - rem)) : lo;
    }
    from_buf(r, t, mn + 1, (int32_t)(q + mn + 1), neg);
}

void ompf_set(ompf *r, const ompf *a) {
    if (r == a) return;
    int n = abs(a->size);
    from_buf(r, a->d, n, a->exp, a->size < 0);
}

void ompf_set_d(ompf *r, double x) {
    if (x == 0.0 || x != x) {
        ompf_set_zero(r);
        return;
    }
    int neg = x < 0;
    int e;
    double f = frexp(fabs(x), &e); /* x = f * 2^e, f in [0.5,1) */
    uint64_t m = (uint64_t)ldexp(f, 53);
    from_bin(r, &m, 1, (long)e - 53, neg);
}

void ompf_set_si(ompf *r, long x) {
    if (x == 0) {
        ompf_set_zero(r);
        return;
    }
    uint64_t m = x < 0 ? (uint64_t)(-(x + 1)) + 1 : (uint64_t)x;
    from_buf(r, &m, 1, 1, x < 0);
}

double ompf_get_d(const ompf *a) {
    int n = abs(a->size);
    if (!n) return 0.0;
    double v = 0.0;
    int k0 = n - 3 < 0 ? 0 : n - 3;
    for (int k = k0; k < n; k++) v += ldexp((double)a->d[k], 64 * (a->exp - n + k));
    return a->size < 0 ? -v : v;
}

int ompf_sgn(const ompf *a) { return a->size > 0 ? 1 : (a->size < 0 ? -1 : 0); }

void ompf_neg(ompf *r, const ompf *a) {
    ompf_set(r, a);
    r->size = -r->size;
}
void ompf_abs(ompf *r, const ompf *a) {
    ompf_set(r, a);
    if (r->size < 0) r->size = -r->size;
}

static int cmp_abs(const ompf *a, const ompf *b) {
    int na = abs(a->size), nb = abs(b->size);
    if (!na || !nb) return (na > 0) - (nb > 0);
    if (a->exp != b->exp) return a->exp < b->exp ? -1 : 1;
    int ia = na - 1, ib = nb - 1;
    while (ia >= 0 && ib >= 0) {
        if (a->d[ia] != b->d[ib]) return a->d[ia] < b->d[ib] ? -1 : 1;
        ia--;
        ib--;
    }
    while (ia >= 0) {
        if (a->d[ia]) return 1;
        ia--;
    }
    while (ib >= 0) {
        if (b->d[ib]) return -1;
        ib--;
    }
    return 0;
}

int ompf_cmp(const ompf *a, const ompf *b) {
    int sa = ompf_sgn(a), sb = ompf_sgn(b);
    if (sa != sb) return sa < sb ? -1 : 1;
    if (sa == 0) return 0;
    int c = cmp_abs(a, b);
    return sa > 0 ? c : -c;
}

int ompf_cmp_d(const ompf *a, double x) {
    ompf t;
    ompf_set_d(&t, x);
    return ompf_cmp(a, &t);
}

/* exact signed add of stored limbs, then truncation (see header) */
static void addsub(ompf *r, const ompf *a, const ompf *b, int bneg) {
    int sa = a->size, sb = bneg ? -b->size : b->size;
    if (sb == 0) {
        ompf_set(r, a);
        return;
    }
    if (sa == 0) {
        from_buf(r, b->d, abs(sb), b->exp, sb < 0);
        return;
    }
    const ompf *u = a, *v = b;
    int su = sa, sv = sb;
    if (u->exp < v->exp) {
        const ompf *t = u;
        u = v;
        v = t;
        int ts = su;
        su = sv;
        sv = ts;
    }
    int nu = abs(su), nv = abs(sv);
    int ediff = u->exp - v->exp;
    if (ediff > PREC + 2) {
        from_buf(r, u->d, nu, u->exp, su < 0);
        return;
    }
    int lo_u = u->exp - nu, lo_v = v->exp - nv;
    int lo = lo_u < lo_v ? lo_u : lo_v;
    int hi = u->exp + 1;
    int n = hi - lo;
    uint64_t A[4 * OMPF_MAXL + 8], Bv[4 * OMPF_MAXL + 8], C[4 * OMPF_MAXL + 8];
    memset(A, 0, sizeof(uint64_t) * n);
    memset(Bv, 0, sizeof(uint64_t) * n);
    memcpy(A + (lo_u - lo), u->d, sizeof(uint64_t) * nu);
    memcpy(Bv + (lo_v - lo), v->d, sizeof(uint64_t) * nv);
    int negu = su < 0, negv = sv < 0;
    int neg;
    if (negu == negv) {
        bn_add_n(C, A, Bv, n);
        neg = negu;
    } else {
        int c = bn_cmp_n(A, Bv, n);
        if (c == 0) {
            ompf_set_zero(r);
            return;
        }
        if (c > 0) {
            bn_sub_n(C, A, Bv, n);
            neg = negu;
        } else {
            bn_sub_n(C, Bv, A, n);
            neg = negv;
        }
    }
    from_buf(r, C, n, hi, neg);
}

void ompf_add(ompf *r, const ompf *a, const ompf *b) { addsub(r, a, b, 0); }
void ompf_sub(ompf *r, const ompf *a, const ompf *b) { addsub(r, a, b, 1); }

void ompf_mul(ompf *r, const ompf *a, const ompf *b) {
    int na = abs(a->size), nb = abs(b->size);
    if (!na || !nb) {
        ompf_set_zero(r);
        return;
    }
    const uint64_t *pa = a->d, *pb = b->d;
    if (na > PREC) {
        pa += na - PREC;
        na = PREC;
    }
    if (nb > PREC) {
        pb += nb - PREC;
        nb = PREC;
    }
    uint64_t t[2 * OMPF_MAXL + 2];
    bn_mul(t, pa, na, pb, nb);
    from_buf(r, t, na + nb, a->exp + b->exp, (a->size < 0) != (b->size < 0));
}

void ompf_addmul(ompf *r, const ompf *a, const ompf *b) {
    ompf t;
    ompf_mul(&t, a, b);
    ompf_add(r, r, &t);
}
void ompf_submul(ompf *r, const ompf *a, const ompf *b) {
    ompf t;
    ompf_mul(&t, a, b);
    ompf_sub(r, r, &t);
}

void ompf_div(ompf *r, const ompf *a, const ompf *b) {
    int na = abs(a->size), nb = abs(b->size);
    if (!nb) {
        fprintf(stderr, "ompf_div: division by zero\n");
        abort();
    }
    if (!na) {
        ompf_set_zero(r);
        return;
    }
    const uint64_t *pa = a->d, *pb = b->d;
    if (nb > PREC + 1) {
        pb += nb - (PREC + 1);
        nb = PREC + 1;
    }
    if (na > PREC + 1) {
        pa += na - (PREC + 1);
        na = PREC + 1;
    }
    int s = PREC + 2 + nb - na;
    if (s < 0) s = 0;
    int nn = na + s;
    uint64_t N[4 * OMPF_MAXL + 8], Q[4 * OMPF_MAXL + 8];
    memset(N, 0, sizeof(uint64_t) * nn);
    memcpy(N + s, pa, sizeof(uint64_t) * na);
    int qn = nn - nb + 1;
    bn_divq(Q, N, nn, pb, nb);
    /* value = Q * B^((a.exp - na - s) - (b.exp - nb)) */
    int32_t lowpos = (a->exp - na - s) - (b->exp - nb);
    from_buf(r, Q, qn, lowpos + qn, (a->size < 0) != (b->size < 0));
}

void ompf_sqrt(ompf *r, const ompf *a) {
    int na = abs(a->size);
    if (a->size < 0) {
        fprintf(stderr, "ompf_sqrt: negative argument\n");
        abort();
    }
    if (!na) {
        ompf_set_zero(r);
        return;
    }
    const uint64_t *pa = a->d;
    if (na > PREC + 1) {
        pa += na - (PREC + 1);
        na = PREC + 1;
    }
    long e = (long)a->exp - na; /* value = A * B^e */
    int o = (e % 2 != 0) ? 1 : 0;
    int t = (2 * (PREC + 2) - na - o + 1) / 2;
    if (t < 0) t = 0;
    int nn = na + 2 * t + o;
    uint64_t N[4 * OMPF_MAXL + 8], S[4 * OMPF_MAXL + 8];
    memset(N, 0, sizeof(uint64_t) * nn);
    memcpy(N + 2 * t + o, pa, sizeof(uint64_t) * na);
    int sn = bn_isqrt(S, N, nn);
    long lowpos = (e - o - 2 * t) / 2;
    from_buf(r, S, sn, (int32_t)(lowpos + sn), 0);
}

void ompf_mul_2exp(ompf *r, const ompf *a, long k) {
    int n = abs(a->size);
    if (!n) {
        ompf_set_zero(r);
        return;
    }
    uint64_t m[OMPF_MAXL];
    memcpy(m, a->d, sizeof(uint64_t) * n);
    from_bin(r, m, n, 64L * (a->exp - n) + k, a->size < 0);
}

/* ---------------- decimal conversion ---------------- */

#define DEC_MAXL 160

int ompf_set_str(ompf *r, const char *s) {
    const char *p = s;
    int neg = 0;
    if (*p == '+' || *p == '-') {
        neg = (*p == '-');
        p++;
    }
    uint64_t D[DEC_MAXL];
    int dn = 1;
    D[0] = 0;
    int ndig = 0, frac = 0, seen_dot = 0;
    while ((*p >= '0' && *p <= '9') || (*p == '.' && !seen_dot)) {
        if (*p == '.') {
            seen_dot = 1;
            p++;
            continue;
        }
        uint64_t c = bn_mul_1(D, dn, 10);
        if (c) {
            if (dn >= DEC_MAXL - 1) return 0;
            D[dn++] = c;
        }
        uint64_t add = (uint64_t)(*p - '0');
        for (int i = 0; i < dn && add; i++) {
            uint64_t o = D[i];
            D[i] += add;
            add = D[i] < o ? 1 : 0;
        }
        if (add) D[dn++] = add;
        ndig++;
        if (seen_dot) frac++;
        p++;
    }
    if (ndig == 0) return 0;
    long e10 = 0;
    if (*p == 'e' || *p == 'E') {
        const char *q = p + 1;
        int eneg = 0;
        if (*q == '+' || *q == '-') {
            eneg = (*q == '-');
            q++;
        }
        if (*q >= '0' && *q <= '9') {
            long ev = 0;
            while (*q >= '0' && *q <= '9') {
                ev = ev * 10 + (*q - '0');
                if (ev > 100000) ev = 100000;
                q++;
            }
            e10 = eneg ? -ev : ev;
            p = q;
        }
    }
    e10 -= frac;
    while (dn > 1 && D[dn - 1] == 0) dn--;
    if (dn == 1 && D[0] == 0) {
        ompf_set_zero(r);
        return (int)(p - s);
    }
    if (e10 >= 0) {
        for (long i = 0; i < e10; i++) {
            uint64_t c = bn_mul_1(D, dn, 10);
            if (c) {
                if (dn >= DEC_MAXL - 1) return 0;
                D[dn++] = c;
            }
        }
        from_buf(r, D, dn, dn, neg);
    } else {
        long k = -e10;
        int sl = PREC + 3 + (int)((k * 3.33 + 63) / 64) + 1;
        if (dn + sl > DEC_MAXL) return 0;
        uint64_t N[DEC_MAXL];
        memset(N, 0, sizeof(uint64_t) * sl);
        memcpy(N + sl, D, sizeof(uint64_t) * dn);
        int nn = dn + sl;
        long left = k;
        while (left > 0) {
            int step = left > 19 ? 19 : (int)left;
            uint64_t pw = 1;
            for (int i = 0; i < step; i++) pw *= 10;
            bn_divrem_1(N, nn, pw);
            left -= step;
        }
        from_buf(r, N, nn, nn - sl, neg);
    }
    return (int)(p - s);
}

void ompf_get_str(char *buf, size_t bufsz, const ompf *a, int digits) {
    int n = abs(a->size);
    if (!n) {
        snprintf(buf, bufsz, "0.%0*de+00", digits > 1 ? digits - 1 : 1, 0);
        return;
    }
    double approx = fabs(ompf_get_d(a));
    int d10 = (approx > 0 && isfinite(approx)) ? (int)floor(log10(approx)) : 0;
    for (int attempt = 0; attempt < 3; attempt++) {
        long k = (long)digits - 1 - d10; /* S = round(|a| * 10^k) */
        uint64_t M[DEC_MAXL];
        int mn = n;
        memcpy(M, a->d, sizeof(uint64_t) * n);
        long e2 = 64L * (a->exp - n);
        if (k >= 0) {
            for (long i = 0; i < k; i++) {
                uint64_t c = bn_mul_1(M, mn, 10);
                if (c) M[mn++] = c;
            }
        } else {
            long kk = -k;
            int sl = (int)((kk * 3.33 + 63) / 64) + 2;
            memmove(M + sl, M, sizeof(uint64_t) * mn);
            memset(M, 0, sizeof(uint64_t) * sl);
            mn += sl;
            e2 -= 64L * sl;
            while (kk > 0) {
                int step = kk > 19 ? 19 : (int)kk;
                uint64_t pw = 1;
                for (int i = 0; i < step; i++) pw *= 10;
                bn_divrem_1(M, mn, pw);
                kk -= step;
            }
        }
        /* S = round(M * 2^e2), half away from zero */
        if (e2 < 0) {
            long sh = -e2;
            /* keep one extra bit for rounding */
            long sh1 = sh - 1;
            long limbs = sh1 / 64;
            int bits = (int)(sh1 % 64);
            if (limbs >= mn) {
                mn = 1;
                M[0] = 0;
            } else {
                memmove(M, M + limbs, sizeof(uint64_t) * (mn - limbs));
                mn -= (int)limbs;
                if (bits) {
                    for (int i = 0; i < mn; i++) M[i] = (M[i] >> bits) | (i + 1 < mn ? M[i + 1] << (64 - bits) : 0);
                }
            }
            int rb = (int)(M[0] & 1);
            for (int i = 0; i < mn; i++) M[i] = (M[i] >> 1) | (i + 1 < mn ? M[i + 1] << 63 : 0);
            if (rb) {
                for (int i = 0; i < mn; i++) {
                    if (++M[i]) break;
                }
            }
        } else if (e2 > 0) {
            for (long i = 0; i < e2; i++) {
                uint64_t c = bn_mul_1(M, mn, 2);
                if (c) M[mn++] = c;
            }
        }
        while (mn > 1 && M[mn - 1] == 0) mn--;
        /* decimal digits of S */
        char tmp[2048];
        int tl = 0;
        uint64_t W[DEC_MAXL];
        memcpy(W, M, sizeof(uint64_t) * mn);
        int wn = mn;
        while (!(wn == 1 && W[0] == 0)) {
            uint64_t rem = bn_divrem_1(W, wn, 10);
            tmp[tl++] = (char)('0' + rem);
            while (wn > 1 && W[wn - 1] == 0) wn--;
            if (tl > 2000) break;
        }
        if (tl == 0) tmp[tl++] = '0';
        if (tl == digits + 1) { /* rounding overflowed or estimate low */
            d10++;
            continue;
        }
        if (tl < digits) {
            d10--;
            continue;
        }
        char out[2100];
        int o = 0;
        if (a->size < 0) out[o++] = '-';
        out[o++] = tmp[tl - 1];
        if (digits > 1) {
            out[o++] = '.';
            for (int i = tl - 2; i >= 0; i--) out[o++] = tmp[i];
        }
        out[o] = 0;
        snprintf(buf, bufsz, "%se%c%02d", out, d10 < 0 ? '-' : '+', d10 < 0 ? -d10 : d10);
        return;
    }
    snprintf(buf, bufsz, "%.*e", digits - 1, ompf_get_d(a));
}

/* ---------------- interchange ---------------- */

void ompf_export32(const ompf *a, int L, int32_t *sign, int32_t *e, uint32_t *limbs) {
    int n = abs(a->size);
    memset(limbs, 0, sizeof(uint32_t) * L);
    if (!n) {
        *sign = 0;
        *e = 0;
        return;
    }
    int c = bn_clz(a->d[n - 1]);
    /* shift whole mantissa left by c bits into buf (n limbs) */
    uint64_t buf[OMPF_MAXL + 1];
    for (int i = n - 1; i >= 0; i--) {
        uint64_t lower = i > 0 ? a->d[i - 1] : 0;
        buf[i] = c ? (a->d[i] << c) | (lower >> (64 - c)) : a->d[i];
    }
    /* 32-bit limbs, most significant first, taken from top */
    for (int j = 0; j < L; j++) {
        int idx32 = 2 * n - 1 - j; /* 32-bit limb index within buf */
        uint32_t v = 0;
        if (idx32 >= 0) {
            uint64_t w = buf[idx32 / 2];
            v = (idx32 & 1) ? (uint32_t)(w >> 32) : (uint32_t)w;
        }
        limbs[L - 1 - j] = v;
    }
    *sign = a->size < 0 ? -1 : 1;
    *e = 64 * a->exp - c;
}

void ompf_import32(ompf *r, int L, int32_t sign, int32_t e, const uint32_t *limbs) {
    if (sign == 0) {
        ompf_set_zero(r);
        return;
    }
    uint64_t m[OMPF_MAXL * 2 + 2];
    int mn = (L + 1) / 2;
    for (int i = 0; i < mn; i++) {
        uint64_t lo = limbs[2 * i];
        uint64_t hi = (2 * i + 1 < L) ? limbs[2 * i + 1] : 0;
        m[i] = lo | (hi << 32);
    }
    from_bin(r, m, mn, (long)e - 32L * L, sign < 0);
}
