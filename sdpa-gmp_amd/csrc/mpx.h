// mpx.h — fixed-limb multiple-precision floating point, shared by host and device.
//
// An element is  value = (-1)^s * M * 2^(e - 32L)  with M = sum_k m[k] 2^(32k)
// (base-2^32 limbs, little-endian) normalised so that bit 32L-1 of M is set,
// i.e. |value| in [2^(e-1), 2^e).  Zero has M = 0, e = EZERO, s = 0.
// L = 32-bit limbs of mantissa; the precision contract is GMP's mpf at the
// same requested bits (reference mpf_class, sdpa_algebra.h:9-12): a request of
// p bits maps to prec = (max(53,p)+127)/64 limbs of 64 bits (gmp-impl.h
// __GMPF_BITS_TO_PREC), i.e. L = 2*prec (a 256-bit request runs with L = 10).
//
// Arithmetic rounds to nearest (ties away) from one guard limb; products
// skip partial-product columns more than two limbs below the guard limb, which
// bounds their error by L*2^-32 guard units.  All loops have compile-time trip
// counts and all limb shifts go through barrel-shifter stages, so no register
// array is ever indexed dynamically on the GPU.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MPX_HD __host__ __device__ __forceinline__
#else
#define MPX_HD inline
#endif

namespace mpx {

constexpr int32_t EZERO = -(1 << 29);

template <int L>
struct mp {
    int32_t e;
    uint32_t s;
    uint32_t m[L];
};

// Requested precision (bits) -> limbs: GMP's mpf keeps prec = (max(53,p)+127)/64 limbs
// of 64 bits per operand (results carry one more), so its guaranteed precision for a
// p-bit request is about 64*prec - 63 bits; L = 2*prec 32-bit limbs (256 -> 10) matches it.
inline int limbs_for_bits(int bits) {
    int b = bits < 53 ? 53 : bits;
    int prec = (b + 127) / 64;
    return 2 * prec;
}

MPX_HD int clz32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return x ? __builtin_clz(x) : 32;
#else
    return x ? __builtin_clz(x) : 32;
#endif
}

// (hi:lo) >> r for 0 <= r < 32
MPX_HD uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t r) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, r);
#else
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> r);
#endif
}

// lo64 += a*b with the carry out of the low 64 bits added into hi32: one v_mad_u64_u32 and one
// v_addc_co_u32 on the GPU, the carry passing through an SGPR pair.  On this target a VALU write
// of an SGPR needs two wait states before a VALU reads it (the compiler pads its own carry chains
// with s_nop for the same reason), so every helper below issues its multiply-adds first and its
// carry-adds after them: in groups of three or four independent multiply-adds each carry-add is
// at least two instructions behind the multiply-add that produced its carry, and the groups of
// one or two pad the gap with s_nop.  Not volatile: the compiler may interleave independent
// groups.
MPX_HD void mac(uint64_t &lo, uint32_t &hi, uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t c;
    asm("v_mad_u64_u32 %0, %2, %3, %4, %0\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32 %1, %2, 0, %1, %2"
        : "+v"(lo), "+v"(hi), "=&s"(c)
        : "v"(a), "v"(b));
#else
    uint64_t p = (uint64_t)a * b;
    lo += p;
    hi += (lo < p);
#endif
}

// two independent multiply-adds (lo0 += a0*b0, lo1 += a1*b1, carries into hi0 / hi1)
MPX_HD void mac2(uint64_t &lo0, uint32_t &hi0, uint32_t a0, uint32_t b0, uint64_t &lo1, uint32_t &hi1, uint32_t a1,
                 uint32_t b1) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t c0, c1;
    asm("v_mad_u64_u32 %0, %4, %6, %7, %0\n\t"
        "v_mad_u64_u32 %1, %5, %8, %9, %1\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32 %2, %4, 0, %2, %4\n\t"
        "v_addc_co_u32 %3, %5, 0, %3, %5"
        : "+v"(lo0), "+v"(lo1), "+v"(hi0), "+v"(hi1), "=&s"(c0), "=&s"(c1)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1));
#else
    mac(lo0, hi0, a0, b0);
    mac(lo1, hi1, a1, b1);
#endif
}

// three independent multiply-adds
MPX_HD void mac3(uint64_t &lo0, uint32_t &hi0, uint32_t a0, uint32_t b0, uint64_t &lo1, uint32_t &hi1, uint32_t a1,
                 uint32_t b1, uint64_t &lo2, uint32_t &hi2, uint32_t a2, uint32_t b2) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t c0, c1, c2;
    asm("v_mad_u64_u32 %0, %6, %9, %10, %0\n\t"
        "v_mad_u64_u32 %1, %7, %11, %12, %1\n\t"
        "v_mad_u64_u32 %2, %8, %13, %14, %2\n\t"
        "v_addc_co_u32 %3, %6, 0, %3, %6\n\t"
        "v_addc_co_u32 %4, %7, 0, %4, %7\n\t"
        "v_addc_co_u32 %5, %8, 0, %5, %8"
        : "+v"(lo0), "+v"(lo1), "+v"(lo2), "+v"(hi0), "+v"(hi1), "+v"(hi2), "=&s"(c0), "=&s"(c1), "=&s"(c2)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1), "v"(a2), "v"(b2));
#else
    mac(lo0, hi0, a0, b0);
    mac(lo1, hi1, a1, b1);
    mac(lo2, hi2, a2, b2);
#endif
}

// four independent multiply-adds
MPX_HD void mac4(uint64_t &lo0, uint32_t &hi0, uint32_t a0, uint32_t b0, uint64_t &lo1, uint32_t &hi1, uint32_t a1,
                 uint32_t b1, uint64_t &lo2, uint32_t &hi2, uint32_t a2, uint32_t b2, uint64_t &lo3, uint32_t &hi3,
                 uint32_t a3, uint32_t b3) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t c0, c1, c2, c3;
    asm("v_mad_u64_u32 %0, %8, %12, %13, %0\n\t"
        "v_mad_u64_u32 %1, %9, %14, %15, %1\n\t"
        "v_mad_u64_u32 %2, %10, %16, %17, %2\n\t"
        "v_mad_u64_u32 %3, %11, %18, %19, %3\n\t"
        "v_addc_co_u32 %4, %8, 0, %4, %8\n\t"
        "v_addc_co_u32 %5, %9, 0, %5, %9\n\t"
        "v_addc_co_u32 %6, %10, 0, %6, %10\n\t"
        "v_addc_co_u32 %7, %11, 0, %7, %11"
        : "+v"(lo0), "+v"(lo1), "+v"(lo2), "+v"(lo3), "+v"(hi0), "+v"(hi1), "+v"(hi2), "+v"(hi3), "=&s"(c0),
          "=&s"(c1), "=&s"(c2), "=&s"(c3)
        : "v"(a0), "v"(b0), "v"(a1), "v"(b1), "v"(a2), "v"(b2), "v"(a3), "v"(b3));
#else
    mac(lo0, hi0, a0, b0);
    mac(lo1, hi1, a1, b1);
    mac(lo2, hi2, a2, b2);
    mac(lo3, hi3, a3, b3);
#endif
}

// Column sums of a[0..N) x b[0..N) for the columns S0..2N-2 added into lo/hi[s - S0]: the products of
// each a[p] go to the adjacent columns p + q, p + q + 1, ... in groups of four.
template <int N, int S0>
MPX_HD void col_products(const uint32_t (&a)[N], const uint32_t (&b)[N], uint64_t (&lo)[2 * N - 1 - S0],
                         uint32_t (&hi)[2 * N - 1 - S0]) {
#pragma unroll
    for (int p = 0; p < N; p++) {
        const int q0 = S0 > p ? S0 - p : 0;
#pragma unroll
        for (int q = q0; q < N; q += 4) {
            const int c = p + q - S0, rem = N - q;
            if (rem >= 4)
                mac4(lo[c], hi[c], a[p], b[q], lo[c + 1], hi[c + 1], a[p], b[q + 1], lo[c + 2], hi[c + 2], a[p],
                     b[q + 2], lo[c + 3], hi[c + 3], a[p], b[q + 3]);
            else if (rem == 3)
                mac3(lo[c], hi[c], a[p], b[q], lo[c + 1], hi[c + 1], a[p], b[q + 1], lo[c + 2], hi[c + 2], a[p],
                     b[q + 2]);
            else if (rem == 2)
                mac2(lo[c], hi[c], a[p], b[q], lo[c + 1], hi[c + 1], a[p], b[q + 1]);
            else
                mac(lo[c], hi[c], a[p], b[q]);
        }
    }
}

// The same column sums written into lo/hi (no initialisation needed): each column starts with its
// first product as a plain 64-bit product (it cannot carry), a[0] b[q] for the columns q < N and
// a[p] b[N-1] for the columns p + N - 1 above; the other products accumulate as in col_products.
template <int N, int S0>
MPX_HD void col_products_set(const uint32_t (&a)[N], const uint32_t (&b)[N], uint64_t (&lo)[2 * N - 1 - S0],
                             uint32_t (&hi)[2 * N - 1 - S0]) {
#pragma unroll
    for (int q = S0; q < N; q++) {
        lo[q - S0] = (uint64_t)a[0] * b[q];
        hi[q - S0] = 0;
    }
#pragma unroll
    for (int p = 1; p < N; p++) {
        lo[p + N - 1 - S0] = (uint64_t)a[p] * b[N - 1];
        hi[p + N - 1 - S0] = 0;
    }
#pragma unroll
    for (int p = 1; p < N; p++) {
        const int q0 = S0 > p ? S0 - p : 0;
#pragma unroll
        for (int q = q0; q < N - 1; q += 4) {
            const int c = p + q - S0, rem = N - 1 - q;
            if (rem >= 4)
                mac4(lo[c], hi[c], a[p], b[q], lo[c + 1], hi[c + 1], a[p], b[q + 1], lo[c + 2], hi[c + 2], a[p],
                     b[q + 2], lo[c + 3], hi[c + 3], a[p], b[q + 3]);
            else if (rem == 3)
                mac3(lo[c], hi[c], a[p], b[q], lo[c + 1], hi[c + 1], a[p], b[q + 1], lo[c + 2], hi[c + 2], a[p],
                     b[q + 2]);
            else if (rem == 2)
                mac2(lo[c], hi[c], a[p], b[q], lo[c + 1], hi[c + 1], a[p], b[q + 1]);
            else
                mac(lo[c], hi[c], a[p], b[q]);
        }
    }
}

template <int L>
MPX_HD bool is_zero(const mp<L> &a) {
    return a.m[L - 1] == 0;
}

template <int L>
MPX_HD mp<L> zero() {
    mp<L> r;
    r.e = EZERO;
    r.s = 0;
#pragma unroll
    for (int k = 0; k < L; k++) r.m[k] = 0;
    return r;
}

// x >>= d bits (N limbs); bits shifted out are OR-ed into sticky
template <int N>
MPX_HD void shr(uint32_t (&x)[N], uint32_t d, uint32_t &sticky) {
    if (d >= 32u * N) {
#pragma unroll
        for (int k = 0; k < N; k++) {
            sticky |= x[k];
            x[k] = 0;
        }
        return;
    }
    uint32_t q = d >> 5, r = d & 31;
#pragma unroll
    for (int st = 16; st >= 1; st >>= 1) {
        if (st < N) {
            bool take = (q & st) != 0;
            uint32_t lost = 0;
#pragma unroll
            for (int k = 0; k < st; k++) lost |= x[k];
            sticky |= take ? lost : 0u;
#pragma unroll
            for (int k = 0; k < N; k++) {
                uint32_t v = (k + st < N) ? x[k + st] : 0u;
                x[k] = take ? v : x[k];
            }
        }
    }
    if (r) {
        sticky |= x[0] & ((1u << r) - 1u);
#pragma unroll
        for (int k = 0; k < N; k++) x[k] = funnel(k + 1 < N ? x[k + 1] : 0u, x[k], r);
    }
}

// x <<= d bits (N limbs), d < 32N
template <int N>
MPX_HD void shl(uint32_t (&x)[N], uint32_t d) {
    uint32_t q = d >> 5, r = d & 31;
#pragma unroll
    for (int st = 16; st >= 1; st >>= 1) {
        if (st < N) {
            bool take = (q & st) != 0;
#pragma unroll
            for (int k = N - 1; k >= 0; k--) {
                uint32_t v = (k - st >= 0) ? x[k - st] : 0u;
                x[k] = take ? v : x[k];
            }
        }
    }
    if (r) {
#pragma unroll
        for (int k = N - 1; k >= 0; k--) x[k] = funnel(x[k], k > 0 ? x[k - 1] : 0u, 32 - r);
    }
}

template <int N>
MPX_HD uint32_t lead_zeros(const uint32_t (&x)[N]) {
    uint32_t lz = 0;
    bool found = false;
#pragma unroll
    for (int k = N - 1; k >= 0; k--) {
        uint32_t c = clz32(x[k]);
        lz += found ? 0u : c;
        found = found || (x[k] != 0);
    }
    return lz;
}

// Round the (L+1)-limb normalised buffer x (x[0] guard) into r.
template <int L>
MPX_HD void round_store(mp<L> &r, const uint32_t (&x)[L + 1], int32_t e, uint32_t s) {
    uint32_t up = x[0] >> 31;
    uint32_t c = up;
#pragma unroll
    for (int k = 0; k < L; k++) {
        uint32_t v = x[k + 1] + c;
        c = (c && v == 0) ? 1u : 0u;
        r.m[k] = v;
    }
    if (c) {  // mantissa overflowed to 2^(32L)
        r.m[L - 1] = 0x80000000u;
        e += 1;
    }
    r.e = e;
    r.s = s;
}

template <int L>
MPX_HD int cmp_abs(const mp<L> &a, const mp<L> &b) {
    bool za = is_zero(a), zb = is_zero(b);
    if (za || zb) return (za ? 0 : 1) - (zb ? 0 : 1);
    if (a.e != b.e) return a.e < b.e ? -1 : 1;
    int res = 0;
#pragma unroll
    for (int k = L - 1; k >= 0; k--) {
        int c = (a.m[k] < b.m[k]) ? -1 : (a.m[k] > b.m[k] ? 1 : 0);
        res = res ? res : c;
    }
    return res;
}

template <int L>
MPX_HD int sgn(const mp<L> &a) {
    return is_zero(a) ? 0 : (a.s ? -1 : 1);
}

template <int L>
MPX_HD int cmp(const mp<L> &a, const mp<L> &b) {
    int sa = sgn(a), sb = sgn(b);
    if (sa != sb) return sa < sb ? -1 : 1;
    if (sa == 0) return 0;
    int c = cmp_abs(a, b);
    return sa > 0 ? c : -c;
}

template <int L>
MPX_HD mp<L> neg(mp<L> a) {
    if (!is_zero(a)) a.s ^= 1u;
    return a;
}
template <int L>
MPX_HD mp<L> absv(mp<L> a) {
    a.s = 0;
    return a;
}

// |b| > |a| for non-zero a, b: the exponents decide, and at equal exponents the borrow out of
// the mantissa subtraction a - b (one carry chain instead of a limb-by-limb comparison)
template <int L>
MPX_HD bool gt_abs(const mp<L> &b, const mp<L> &a) {
    uint32_t br = 0;
#pragma unroll
    for (int k = 0; k < L; k++) {
        uint64_t t = (uint64_t)a.m[k] - b.m[k] - br;
        br = (uint32_t)(t >> 63);
    }
    return b.e > a.e || (b.e == a.e && br != 0);
}

// a + b, rounded to nearest.  One code path for every sign and zero case (no divergence between
// lanes that add and lanes that subtract): x = the larger magnitude, y = the smaller (a zero has the
// smallest exponent, so it always lands in y and aligns to nothing); y is aligned to x with a
// sticky bit; W = x + y or x - y (two's complement, borrow-in from the sticky bit) over L + 2 words
// [guard, L limbs, top]; one left shift by the leading zeros of W (31 after a carry into the top
// word, 32 without, 32 + the cancelled bits after a subtraction) normalises it, and the result has
// exponent e_x + 32 - lz.
template <int L>
MPX_HD mp<L> add(mp<L> a, mp<L> b) {
    // a zero orders by the zero exponent whatever exponent it carries
    if (is_zero(a)) a.e = EZERO;
    if (is_zero(b)) b.e = EZERO;
    const bool sw = gt_abs<L>(b, a);
    const int32_t ex = sw ? b.e : a.e, ey = sw ? a.e : b.e;
    const uint32_t sx = sw ? b.s : a.s, sy = sw ? a.s : b.s;
    uint32_t W[L + 2], y[L + 1];
    W[0] = 0;
    y[0] = 0;
#pragma unroll
    for (int k = 0; k < L; k++) {
        W[k + 1] = sw ? b.m[k] : a.m[k];
        y[k + 1] = sw ? a.m[k] : b.m[k];
    }
    W[L + 1] = 0;
    uint32_t sticky = 0;
    shr<L + 1>(y, (uint32_t)(ex - ey), sticky);
    // subtraction: W - y - [sticky != 0] = W + ~y + [sticky == 0]; the top word ends at 0 there, since
    // |x| >= |y| keeps the difference non-negative
    const uint32_t neg = (sx != sy) ? 0xFFFFFFFFu : 0u;
    uint32_t c = neg ? (sticky ? 0u : 1u) : 0u;
#pragma unroll
    for (int k = 0; k <= L + 1; k++) {
        const uint64_t t = (uint64_t)W[k] + ((k <= L ? y[k] : 0u) ^ neg) + c;
        W[k] = (uint32_t)t;
        c = (uint32_t)(t >> 32);
    }
    const uint32_t lz = lead_zeros<L + 2>(W);
    if (lz >= 32u * (L + 2)) return zero<L>();
    shl<L + 2>(W, lz);
    uint32_t x[L + 1];
#pragma unroll
    for (int k = 0; k <= L; k++) x[k] = W[k + 1];
    mp<L> r;
    round_store<L>(r, x, ex + 32 - (int32_t)lz, sx);
    return r;
}

// add() with an early exit when either operand is zero: for accumulations that start from zero, where
// the exit is taken by whole waves at once
template <int L>
MPX_HD mp<L> add_z(const mp<L> &a, const mp<L> &b) {
    if (is_zero(a)) return b;
    if (is_zero(b)) return a;
    return add<L>(a, b);
}

template <int L>
MPX_HD mp<L> sub(const mp<L> &a, const mp<L> &b) {
    return add<L>(a, neg<L>(b));
}

// Carry-propagate NC column sums (column c = lo[c] + hi[c] 2^64) into NC + 1 words P[0..NC].
template <int NC>
MPX_HD void col_carry(const uint64_t (&lo)[NC], const uint32_t (&hi)[NC], uint32_t *P) {
    uint64_t cl = 0;  // carry into the next column, < 2^64
#pragma unroll
    for (int c = 0; c < NC; c++) {
        uint64_t t = lo[c] + cl;
        uint32_t cy = t < cl ? 1u : 0u;
        P[c] = (uint32_t)t;
        cl = (t >> 32) | ((uint64_t)(hi[c] + cy) << 32);
    }
    P[NC] = (uint32_t)cl;
}

// Product of mantissas: P[0..2L) top columns (columns below L-3 are skipped and read as zero).
// Every column has its own 96-bit accumulator, so consecutive multiply-adds are independent
// (issued in pairs by mac2); the carries run once at the end.
template <int L>
MPX_HD void mant_mul(const uint32_t (&a)[L], const uint32_t (&b)[L], uint32_t (&P)[2 * L]) {
    constexpr int S0 = (L >= 3) ? L - 3 : 0;
    constexpr int NC = 2 * L - 1 - S0;
    uint64_t lo[NC];
    uint32_t hi[NC];
    col_products_set<L, S0>(a, b, lo, hi);
#pragma unroll
    for (int s = 0; s < S0; s++) P[s] = 0;
    col_carry<NC>(lo, hi, P + S0);
}

// a * b, rounded to nearest; one code path (a zero operand selects zero at the end)
template <int L>
MPX_HD mp<L> mul(const mp<L> &a, const mp<L> &b) {
    uint32_t P[2 * L];
    mant_mul<L>(a.m, b.m, P);
    uint32_t x[L + 1];
#pragma unroll
    for (int k = 0; k <= L; k++) x[k] = P[L - 1 + k];
    // the product of two normalised mantissas has its top bit at 2^(64L-1) or 2^(64L-2)
    const bool top = (x[L] & 0x80000000u) != 0;
#pragma unroll
    for (int k = L; k >= 0; k--) x[k] = top ? x[k] : funnel(x[k], k > 0 ? x[k - 1] : P[L - 2 >= 0 ? L - 2 : 0], 31);
    mp<L> r;
    round_store<L>(r, x, a.e + b.e - (top ? 0 : 1), a.s ^ b.s);
    return (is_zero(a) || is_zero(b)) ? zero<L>() : r;
}

// a * b - c * d rounded once to nearest.  Both products keep their top L + 2 words (the low product
// columns skipped weigh below 2^-(32L+26) of the larger product for L = 10), the one with the smaller
// exponent is aligned to the other with a sticky bit, one two's-complement pass over L + 3 words forms
// the sum or difference (negated when a difference of products with close exponents comes out
// negative), and one normalising shift and one rounding give the result.  Compared with
// sub(mul(a, b), mul(c, d)) this skips two normalisations and roundings, and under cancellation it
// is more accurate (no rounding error of the products is amplified).
template <int L>
MPX_HD mp<L> fms(const mp<L> &a, const mp<L> &b, const mp<L> &c, const mp<L> &d) {
    constexpr int N = L + 2;
    // product columns from L - 2 up (the lowest kept word); the skipped ones weigh below
    // 2 (L - 2) 2^-(32L+30) of the product
    constexpr int S0 = (L >= 2) ? L - 2 : 0, NC = 2 * L - 1 - S0;
    uint32_t P[2 * L], Q[2 * L];
    {
        uint64_t lo[NC], lo2[NC];
        uint32_t hi[NC], hi2[NC];
        col_products_set<L, S0>(a.m, b.m, lo, hi);
        col_products_set<L, S0>(c.m, d.m, lo2, hi2);
#pragma unroll
        for (int s = 0; s < S0; s++) P[s] = Q[s] = 0;
        col_carry<NC>(lo, hi, P + S0);
        col_carry<NC>(lo2, hi2, Q + S0);
    }
    // product value = (top N words) * 2^(e - 32N); a zero product orders by the zero exponent
    const int32_t ep = (is_zero(a) || is_zero(b)) ? EZERO : a.e + b.e;
    const int32_t eq = (is_zero(c) || is_zero(d)) ? EZERO : c.e + d.e;
    const uint32_t sp = a.s ^ b.s, sq = c.s ^ d.s ^ 1u;
    const bool sw = eq > ep;
    const int32_t ex = sw ? eq : ep, ey = sw ? ep : eq;
    const uint32_t sx = sw ? sq : sp, sy = sw ? sp : sq;
    uint32_t W[N + 1], y[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        W[k] = sw ? Q[L - 2 + k] : P[L - 2 + k];
        y[k] = sw ? P[L - 2 + k] : Q[L - 2 + k];
    }
    W[N] = 0;
    uint32_t sticky = 0;
    shr<N>(y, (uint32_t)(ex - ey), sticky);
    const uint32_t neg = (sx != sy) ? 0xFFFFFFFFu : 0u;
    uint32_t cy = neg ? (sticky ? 0u : 1u) : 0u;
#pragma unroll
    for (int k = 0; k <= N; k++) {
        const uint64_t t = (uint64_t)W[k] + ((k < N ? y[k] : 0u) ^ neg) + cy;
        W[k] = (uint32_t)t;
        cy = (uint32_t)(t >> 32);
    }
    // magnitude and sign of the result
    const uint32_t wneg = W[N] >> 31, wm = 0u - wneg;
    cy = wneg;
#pragma unroll
    for (int k = 0; k <= N; k++) {
        const uint64_t t = (uint64_t)(W[k] ^ wm) + cy;
        W[k] = (uint32_t)t;
        cy = (uint32_t)(t >> 32);
    }
    const uint32_t lz = lead_zeros<N + 1>(W);
    if (lz >= 32u * (N + 1)) return zero<L>();
    shl<N + 1>(W, lz);
    uint32_t x[L + 1];
#pragma unroll
    for (int k = 0; k <= L; k++) x[k] = W[k + 2];
    mp<L> r;
    round_store<L>(r, x, ex + 32 - (int32_t)lz, sx ^ wneg);
    return r;
}

// value * 2^k
template <int L>
MPX_HD mp<L> ldexp2(mp<L> a, int k) {
    if (!is_zero(a)) a.e += k;
    return a;
}

template <int L>
MPX_HD mp<L> from_double(double x) {
    mp<L> r = zero<L>();
    if (x == 0.0 || x != x) return r;
    uint64_t bits;
    __builtin_memcpy(&bits, &x, 8);
    uint32_t s = (uint32_t)(bits >> 63);
    int ex = (int)((bits >> 52) & 0x7ff);
    uint64_t mant = bits & ((1ull << 52) - 1);
    if (ex == 0) {  // subnormal: normalise so that bit 52 is set
        int lz = __builtin_clzll(mant) - 11;
        mant <<= lz;
        ex = 1 - lz;
    } else {
        mant |= (1ull << 52);
    }
    // value = mant * 2^(ex - 1075); mant has 53 bits -> top bit at 52
    uint64_t top = mant << 11;  // bit 63 set
    r.m[L - 1] = (uint32_t)(top >> 32);
    if (L >= 2) r.m[L - 2] = (uint32_t)top;
    r.e = ex - 1075 + 53;
    r.s = s;
    return r;
}

template <int L>
MPX_HD mp<L> from_int(long long v) {
    mp<L> r = zero<L>();
    if (v == 0) return r;
    uint64_t u = v < 0 ? (uint64_t)(-(v + 1)) + 1 : (uint64_t)v;
    int lz = __builtin_clzll(u);
    u <<= lz;
    r.m[L - 1] = (uint32_t)(u >> 32);
    if (L >= 2) r.m[L - 2] = (uint32_t)u;
    r.e = 64 - lz;
    r.s = v < 0 ? 1u : 0u;
    return r;
}

template <int L>
MPX_HD double to_double(const mp<L> &a) {
    if (is_zero(a)) return 0.0;
    uint64_t top = ((uint64_t)a.m[L - 1] << 32) | (L >= 2 ? a.m[L - 2] : 0u);
    int ex = a.e - 64;
    if (ex > 1100) ex = 1100;
    if (ex < -1200) ex = -1200;
    double v = __builtin_ldexp((double)top, ex);
    return a.s ? -v : v;
}

// 1/b by Newton iteration x <- x + x(1 - b x)
template <int L>
MPX_HD mp<L> recip(const mp<L> &b) {
    // b = B * 2^(e-32L), B in [2^(32L-1), 2^(32L)); work with bn = b * 2^-e in [1/2,1)
    mp<L> bn = b;
    bn.e = 0;
    bn.s = 0;
    mp<L> x = from_double<L>(1.0 / to_double<L>(bn));
    const mp<L> one = from_int<L>(1);
    constexpr int iters = (32 * L > 400) ? 5 : ((32 * L > 200) ? 4 : 3);
#pragma unroll
    for (int it = 0; it < iters; it++) {
        mp<L> t = sub<L>(one, mul<L>(bn, x));
        x = add<L>(x, mul<L>(x, t));
    }
    x.e -= b.e;
    x.s = b.s;
    return x;
}

template <int L>
MPX_HD mp<L> div(const mp<L> &a, const mp<L> &b) {
    if (is_zero(a)) return zero<L>();
    mp<L> x = recip<L>(b);
    mp<L> q = mul<L>(a, x);
    // one correction step: q += x (a - b q)
    mp<L> r = sub<L>(a, mul<L>(b, q));
    return add<L>(q, mul<L>(x, r));
}

// sqrt(a), a >= 0, via y ~ 1/sqrt(a): y <- y + y(1 - a y^2)/2
template <int L>
MPX_HD mp<L> sqrtv(const mp<L> &a) {
    if (is_zero(a)) return zero<L>();
    mp<L> an = a;
    int ev = a.e & ~1;  // even part of the exponent
    an.e = a.e - ev;    // an in [1/2, 2)
    an.s = 0;
    mp<L> y = from_double<L>(1.0 / __builtin_sqrt(to_double<L>(an)));
    const mp<L> one = from_int<L>(1);
    constexpr int iters = (32 * L > 400) ? 5 : ((32 * L > 200) ? 4 : 3);
#pragma unroll
    for (int it = 0; it < iters; it++) {
        mp<L> t = sub<L>(one, mul<L>(an, mul<L>(y, y)));
        y = add<L>(y, ldexp2<L>(mul<L>(y, t), -1));
    }
    mp<L> s = mul<L>(an, y);
    // correction: s += y (an - s^2) / 2
    mp<L> r = sub<L>(an, mul<L>(s, s));
    s = add<L>(s, ldexp2<L>(mul<L>(y, r), -1));
    s.e += ev / 2;
    return s;
}

// Change the limb count: keeps the most significant limbs (truncates when narrowing).
template <int LO, int L>
MPX_HD mp<LO> resize(const mp<L> &a) {
    mp<LO> r;
    r.e = a.e;
    r.s = a.s;
#pragma unroll
    for (int k = 0; k < LO; k++) {
        int src = k - LO + L;
        r.m[k] = (src >= 0 && src < L) ? a.m[src] : 0u;
    }
    return r;
}

// one Newton step y <- y + y (1 - d y^2) / 2 for 1/sqrt(d) at LW limbs
template <int LW>
MPX_HD mp<LW> rsqrt_step(const mp<LW> &d, const mp<LW> &y) {
    const mp<LW> one = from_int<LW>(1);
    mp<LW> t = sub<LW>(one, mul<LW>(d, mul<LW>(y, y)));
    return add<LW>(y, ldexp2<LW>(mul<LW>(y, t), -1));
}

// 1/sqrt(a), a > 0, by Newton iteration with precision doubling: a 53-bit double start,
// then steps at 4, 8 and L limbs (about 106, 212, 424 correct bits), plus one more step
// at L limbs when 32L > 400.  Used for the Cholesky pivots, where the chain of m
// dependent square roots is the serial critical path.
template <int L>
MPX_HD mp<L> rsqrt(const mp<L> &a) {
    mp<L> an = a;
    int ev = a.e & ~1;
    an.e = a.e - ev;  // an in [1/2, 2)
    an.s = 0;
    constexpr int LA = L < 4 ? L : 4;
    constexpr int LB = L < 8 ? L : 8;
    mp<LA> yA = from_double<LA>(1.0 / __builtin_sqrt(to_double<L>(an)));
    yA = rsqrt_step<LA>(resize<LA, L>(an), yA);
    mp<LB> yB = rsqrt_step<LB>(resize<LB, L>(an), resize<LB, LA>(yA));
    mp<L> y = rsqrt_step<L>(an, resize<L, LB>(yB));
    if (32 * L > 400) y = rsqrt_step<L>(an, y);
    y.e -= ev / 2;
    return y;
}

// ---- fixed-point reciprocal square root for the Cholesky pivots ----
// Fixed point with 2 integer bits: an N-word unsigned X stands for X / 2^(32N-2).

// r = (a * b) >> (32N - 2), product columns below N-2 skipped (error below 2^-(32N-36));
// column-parallel accumulators as in mant_mul
template <int N>
MPX_HD void fx2_mulhi(const uint32_t (&a)[N], const uint32_t (&b)[N], uint32_t (&r)[N]) {
    constexpr int S0 = N >= 2 ? N - 2 : 0;
    constexpr int NC = 2 * N - 1 - S0;
    uint64_t lo[NC];
    uint32_t hi[NC];
    col_products_set<N, S0>(a, b, lo, hi);
    uint32_t P[2 * N + 1];
#pragma unroll
    for (int s = 0; s < S0; s++) P[s] = 0;
    col_carry<NC>(lo, hi, P + S0);
    P[2 * N] = 0;
#pragma unroll
    for (int k = 0; k < N; k++) r[k] = (P[N + k] << 2) | (P[N + k - 1] >> 30);
}

// one Newton step y <- y + y (1 - d y^2) / 2 at N words (d, y in the 2-integer-bit format)
template <int N>
MPX_HD void fx2_rsqrt_step(const uint32_t (&d)[N], uint32_t (&y)[N]) {
    uint32_t u[N], v[N], t[N], w[N];
    fx2_mulhi<N>(y, y, u);
    fx2_mulhi<N>(d, u, v);
    // t = |1 - v|, one = 2^(32N-2)
    uint32_t neg = 0;
    {
        uint32_t br = 0;
#pragma unroll
        for (int k = 0; k < N; k++) {
            uint32_t one = (k == N - 1) ? 0x40000000u : 0u;
            uint64_t x = (uint64_t)one - v[k] - br;
            t[k] = (uint32_t)x;
            br = (uint32_t)(x >> 63);
        }
        neg = br;  // v > 1
        if (neg) {
            uint32_t c = 1;
#pragma unroll
            for (int k = 0; k < N; k++) {
                uint32_t x = ~t[k] + c;
                c = (c && x == 0) ? 1u : 0u;
                t[k] = x;
            }
        }
    }
    fx2_mulhi<N>(y, t, w);
    // w /= 2
#pragma unroll
    for (int k = 0; k < N; k++) w[k] = (w[k] >> 1) | (k + 1 < N ? (w[k + 1] << 31) : 0u);
    uint32_t c = 0;
    if (!neg) {
#pragma unroll
        for (int k = 0; k < N; k++) {
            uint64_t x = (uint64_t)y[k] + w[k] + c;
            y[k] = (uint32_t)x;
            c = (uint32_t)(x >> 32);
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) {
            uint64_t x = (uint64_t)y[k] - w[k] - c;
            y[k] = (uint32_t)x;
            c = (uint32_t)(x >> 63);
        }
    }
}

// widen / narrow an N-word fixed value to M words (keeps the most significant words)
template <int M, int N>
MPX_HD void fx2_resize(const uint32_t (&a)[N], uint32_t (&r)[M]) {
#pragma unroll
    for (int k = 0; k < M; k++) {
        int src = k - M + N;
        r[k] = (src >= 0 && src < N) ? a[src] : 0u;
    }
}

// 1/sqrt(a), a > 0, by Newton iteration in 2-integer-bit fixed point with precision doubling:
// a 53-bit double start, then steps at 4, 8 and L+1 words (106, 212 and the full width).
// Used on the serial pivot chain of the Cholesky factorisation, where floating-point adds
// (variable alignment shifts) would dominate.  sq (optional) receives sqrt(a) = a / sqrt(a).
template <int L>
MPX_HD mp<L> rsqrt_fx(const mp<L> &a, mp<L> *sq = nullptr) {
    constexpr int N = L + 1;
    const int ev = a.e & ~1;
    const int sh = (a.e - ev) ? 31 : 30;  // dn = a 2^-ev in [1/2, 2)
    uint32_t d[N];
    // d = M << sh over N words (M = a.m, L words)
    d[0] = a.m[0] << sh;
#pragma unroll
    for (int k = 1; k < N; k++) d[k] = (k < L ? (a.m[k] << sh) : 0u) | (a.m[k - 1] >> (32 - sh));
    double dn = 0;
    {
        uint64_t top = ((uint64_t)d[N - 1] << 32) | d[N - 2];
        dn = (double)top * 0x1p-62;
    }
    double y0 = 1.0 / __builtin_sqrt(dn);
    uint32_t y4[4];
    {
        uint64_t Y = (uint64_t)(y0 * 0x1p62);  // y0 < 2: fits 64 bits
        y4[3] = (uint32_t)(Y >> 32);
        y4[2] = (uint32_t)Y;
        y4[1] = 0;
        y4[0] = 0;
    }
    constexpr int NA = N < 4 ? N : 4;
    constexpr int NB = N < 8 ? N : 8;
    uint32_t dA[NA], yA[NA];
    fx2_resize<NA, N>(d, dA);
    fx2_resize<NA, 4>(y4, yA);
    fx2_rsqrt_step<NA>(dA, yA);
    uint32_t dB[NB], yB[NB];
    fx2_resize<NB, N>(d, dB);
    fx2_resize<NB, NA>(yA, yB);
    fx2_rsqrt_step<NB>(dB, yB);
    uint32_t y[N];
    fx2_resize<N, NB>(yB, y);
    fx2_rsqrt_step<N>(d, y);
    if (32 * L > 400) fx2_rsqrt_step<N>(d, y);
    // y in (1/sqrt(2), sqrt(2)]: normalise the N words (top bit at 2^(32N-1) or 2^(32N-2) or lower)
    if (sq) {  // sqrt(a) = a * (1/sqrt(a)) in the same fixed point
        uint32_t s[N];
        fx2_mulhi<N>(d, y, s);
        uint32_t lzs = lead_zeros<N>(s);
        shl<N>(s, lzs);
        round_store<L>(*sq, s, 2 - (int32_t)lzs + ev / 2, 0u);
    }
    uint32_t lz = lead_zeros<N>(y);
    shl<N>(y, lz);
    mp<L> r;
    round_store<L>(r, y, 2 - (int32_t)lz - ev / 2, 0u);
    return r;
}

}  // namespace mpx
