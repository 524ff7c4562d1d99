// chol_fx.h — the block-column elimination of the blocked Cholesky factorisation in block fixed point,
// shared by the device kernel (one workgroup, LDS arrays) and a serial host build (host arrays) that the
// CPU tests check against exact rational arithmetic.
//
// The reference factors B by Rpotrf: per block of columns a Cholesky of the diagonal block with a square
// root and a division per pivot (mplapack/Rpotrf2.cpp:86-97), the panel by a triangular solve
// (mplapack/Rtrsm.cpp:263-285) and the trailing matrix by Rsyrk (mplapack/Rsyrk.cpp:157-177).  Here one
// kernel eliminates a whole block column (the nb x nb diagonal block and the panel rows below it) by the
// division-free Bareiss recurrence, and only the trailing update stays a GEMM.
//
// Arithmetic.  Each row r of the block column carries a binary scale 2^f_r and its entries are stored as
// fixed-point numbers V_rc = A_rc 2^-(f_r + f_c) (row r against column c; panel rows use their own scale
// against the block's column scales).  With the scales taken from the diagonals (2^(2 f_r) ~ A_rr), a
// positive definite matrix has |V_rc| <= sqrt(V_rr V_cc) < 2, so a fixed point with 3 integer bits and
// 32L - 4 fraction bits holds every entry.  Step j of the elimination does, for every remaining entry at
// once,
//     V_rc <- (V_jj V_rc - V_rj V_cj) 2^-(2 d_j + d_r + d_c)
// where d_r = floor((h_r + 1) / 2) renormalises row r from the leading exponent h_r of its current
// diagonal (so V_rr 2^-(2 d_r) lies in [1/2, 2)): the normalisation of the inputs is folded into one shift
// of the result, and a step costs two fixed-point products and one shift per entry (no alignment of
// floating-point operands, no per-entry normalisation or rounding).  Panel rows carry their diagonal
// along (V_pp <- V_jj V_pp - V_pj^2, the diagonal of the trailing matrix as this block leaves it) only to
// set their scale.  In terms of the Schur complement S^(j) after j pivots, the block holds
// gamma_j S^(j) scaled by the row scales, with gamma_j = prod_{k<j} Vh_k the product of the normalised
// pivots Vh_k = V_kk 2^-(2 d_k) (each in [1/2, 2)).  After the chain, with
// rho_c = 1/sqrt(gamma_c Vh_c) (formed for all c in parallel),
//     L_rc = V_rc 2^(f_r^(c) - d_c) rho_c,     1 / L_cc = 2^-(f_c^(c) + d_c) gamma_c rho_c,
// V_rc the entry as the pivot column c held it and f_r^(c) the row scale before step c.
//
// Precision: products are truncated to the fraction width, 2^-(32L-4) absolute against entries of order
// one after the diagonal scaling (the backward error of a Cholesky factorisation is measured against the
// same scale, sqrt(B_ii B_jj)).  The final L entries are rounded once (one mp product).
//
// Failures.  A pivot V_jj <= 0 is a failed pivot of Rpotrf (info = its 1-based index; the scale is
// positive, so the sign is the pivot's).  An entry outside the fixed-point range cannot occur in a
// positive semi-definite block column; if one does, B is not positive definite and info receives the
// first pivot the elimination has not yet confirmed (for a panel row: the first pivot after the block).
#pragma once
#include <math.h>

#include "mpx.h"

namespace mpx {

constexpr int32_t FX_HZERO = -(1 << 29);  // leading exponent of a zero fixed-point value

// The fixed point: N = fx_limbs<L>() signed limbs of 29 bits, x_0 .. x_{N-1} (int32 words), value
// (sum_k x_k 2^(29 k)) 2^-F with F = fx_frac<L>() = 29 N - 3 >= 32 L.  Canonical form: 0 <= x_k < 2^29 for
// k < N - 1 and the top limb signed, so |value| < 8 <=> -2^29 <= x_{N-1} < 2^29.  Products of two limbs are
// below 2^58 in magnitude, so a column of partial products (at most 2N of them) sums exactly in a signed
// 64-bit accumulator with one v_mad_i64_i32 per partial product and no carry handling, and a b - c d is one
// set of column sums (the limbs of c negated).
constexpr uint32_t FX_M29 = (1u << 29) - 1;
template <int L>
MPX_HD constexpr int fx_limbs() {
    return (32 * L + 3 + 28) / 29;
}
template <int L>
MPX_HD constexpr int fx_frac() {
    return 29 * fx_limbs<L>() - 3;
}

// renormalisation exponent of a row from the leading exponent of its diagonal: V 2^-(2d) in [1/2, 2)
MPX_HD int32_t fx_delta(int32_t h) {
    return h == FX_HZERO ? 0 : ((h + 1) >> 1);
}

// floor(log2 |value|) of a canonical value, FX_HZERO for zero; for a negative value that of the complement
// -value - 2^-F (within one of the magnitude's).  The top two limbs decide it unless both are zero (the
// value below 2^(29 (N - 2) - F), rare: then the other limbs are scanned).
template <int N>
MPX_HD int32_t fx_lead(const int32_t (&x)[N]) {
    const uint32_t m = x[N - 1] < 0 ? FX_M29 : 0u;
    const uint32_t t1 = (uint32_t)(x[N - 1] < 0 ? ~x[N - 1] : x[N - 1]), t0 = (uint32_t)x[N - 2] ^ m;
    if (t1 | t0) {
        return t1 ? 29 * (N - 1) + 31 - (int32_t)clz32(t1) - (29 * N - 3)
                  : 29 * (N - 2) + 31 - (int32_t)clz32(t0) - (29 * N - 3);
    }
    int32_t pos = -1;
    bool found = false;
#pragma unroll
    for (int k = N - 3; k >= 0; k--) {
        const uint32_t v = (uint32_t)x[k] ^ m;
        const int32_t p = 29 * k + 31 - (int32_t)clz32(v);
        pos = (!found && v) ? p : pos;
        found = found || v;
    }
    return found ? pos - (29 * N - 3) : FX_HZERO;
}

// true when the value is <= 0 (a failed pivot)
template <int N>
MPX_HD bool fx_nonpos(const int32_t (&x)[N]) {
    uint32_t o = 0;
#pragma unroll
    for (int k = 0; k < N; k++) o |= (uint32_t)x[k];
    return o == 0 || x[N - 1] < 0;
}

// x <- -x in canonical form
template <int N>
MPX_HD void fx_negate(int32_t (&x)[N]) {
    int32_t br = 0;
#pragma unroll
    for (int k = 0; k < N - 1; k++) {
        const int32_t v = -x[k] - br;
        x[k] = (int32_t)((uint32_t)v & FX_M29);
        br = v < 0 ? 1 : 0;
    }
    x[N - 1] = -x[N - 1] - br;
}

// x = a 2^-sh in the fixed point, truncated; returns true when |a 2^-sh| >= 8 (out of range)
template <int L>
MPX_HD bool fx_from_mp(const mp<L> &a, int32_t sh, int32_t (&x)[fx_limbs<L>()]) {
    constexpr int N = fx_limbs<L>(), W = (29 * N + 31) / 32;
#pragma unroll
    for (int k = 0; k < N; k++) x[k] = 0;
    if (is_zero<L>(a)) return false;
    if (a.e - sh > 3) return true;
    // X = M 2^t with t = e - 32L - sh + F <= 3 + F - 32L < 32: the bits of M moved to their place in W words
    const int32_t t = a.e - 32 * L - sh + fx_frac<L>();
    uint32_t y[W];
#pragma unroll
    for (int k = 0; k < W; k++) y[k] = k < L ? a.m[k] : 0u;
    if (t >= 0) {
        shl<W>(y, (uint32_t)t);
    } else {
        uint32_t sticky = 0;
        shr<W>(y, (uint32_t)(-t < 32 * W ? -t : 32 * W), sticky);
    }
    // repack into 29-bit limbs (positions fixed at compile time)
#pragma unroll
    for (int k = 0; k < N; k++) {
        const int b = 29 * k, w = b >> 5, o = b & 31;
        uint32_t v = y[w] >> o;
        if (o > 3 && w + 1 < W) v |= y[w + 1] << (32 - o);
        x[k] = (int32_t)(v & FX_M29);
    }
    if (a.s & 1u) fx_negate<N>(x);
    return false;
}

// the mp value of x 2^E, rounded to nearest
template <int L>
MPX_HD mp<L> fx_to_mp(const int32_t (&xin)[fx_limbs<L>()], int32_t E) {
    constexpr int N = fx_limbs<L>(), W = (29 * N + 31) / 32 + 1;
    int32_t x[N];
#pragma unroll
    for (int k = 0; k < N; k++) x[k] = xin[k];
    const uint32_t sgn = x[N - 1] < 0 ? 1u : 0u;
    if (sgn) fx_negate<N>(x);
    // the magnitude as W 32-bit words
    uint32_t y[W];
#pragma unroll
    for (int w = 0; w < W; w++) y[w] = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
        const int b = 29 * k, w = b >> 5, o = b & 31;
        y[w] |= (uint32_t)x[k] << o;
        if (o > 3 && w + 1 < W) y[w + 1] |= (uint32_t)x[k] >> (32 - o);
    }
    const uint32_t lz = lead_zeros<W>(y);
    if (lz >= 32u * W) return zero<L>();
    shl<W>(y, lz);
    // top L words and a guard word (the word below it, W >= L + 1)
    uint32_t g[L + 1];
#pragma unroll
    for (int k = 0; k <= L; k++) g[k] = y[W - 1 - L + k];
    mp<L> r;
    round_store<L>(r, g, E - fx_frac<L>() + 32 * W - (int32_t)lz, sgn);
    return r;
}

// acc_k += x_k y_k (signed 32 x 32 -> 64 bits) for k < 4: four independent v_mad_i64_i32 issued together (the
// carry-out SGPRs are not read), so that consecutive multiply-adds never wait on one another's result
MPX_HD void smac4(int64_t &a0, int32_t x0, int32_t y0, int64_t &a1, int32_t x1, int32_t y1, int64_t &a2, int32_t x2,
                  int32_t y2, int64_t &a3, int32_t x3, int32_t y3) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t c0, c1, c2, c3;
    asm("v_mad_i64_i32 %0, %4, %8, %9, %0\n\t"
        "v_mad_i64_i32 %1, %5, %10, %11, %1\n\t"
        "v_mad_i64_i32 %2, %6, %12, %13, %2\n\t"
        "v_mad_i64_i32 %3, %7, %14, %15, %3"
        : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "=&s"(c0), "=&s"(c1), "=&s"(c2), "=&s"(c3)
        : "v"(x0), "v"(y0), "v"(x1), "v"(y1), "v"(x2), "v"(y2), "v"(x3), "v"(y3));
#else
    a0 += (int64_t)x0 * y0;
    a1 += (int64_t)x1 * y1;
    a2 += (int64_t)x2 * y2;
    a3 += (int64_t)x3 * y3;
#endif
}

// col[p + q - S0] += x[p] y[q] for p >= 1, q <= N - 2, p + q >= S0 (the partial products a column does not
// start with): row p's products go to adjacent columns, four at a time
template <int N, int S0>
MPX_HD void fx_rows(const int32_t (&x)[N], const int32_t (&y)[N], int64_t (&col)[2 * N - 1 - S0]) {
#pragma unroll
    for (int p = 1; p < N; p++) {
        const int q0 = S0 > p ? S0 - p : 0;
#pragma unroll
        for (int q = q0; q < N - 1; q += 4) {
            const int cc = p + q - S0, rem = N - 1 - q;
            if (rem >= 4) {
                smac4(col[cc], x[p], y[q], col[cc + 1], x[p], y[q + 1], col[cc + 2], x[p], y[q + 2], col[cc + 3], x[p],
                      y[q + 3]);
            } else {
#pragma unroll
                for (int u = 0; u < rem; u++) col[cc + u] += (int64_t)x[p] * y[q + u];
            }
        }
    }
}

// r = the N limbs of U (canonical limbs of a signed integer whose limbs beyond U are ext: all ones for a
// negative one) taken from bit rb (0 <= rb < 29) up; returns true when the bits above r's top limb are not the
// sign extension (|r| >= 8).
template <int N, int M>
MPX_HD bool fx_extract(const uint32_t (&U)[M], uint32_t rb, uint32_t ext, int32_t (&r)[N]) {
#pragma unroll
    for (int k = 0; k < N - 1; k++) r[k] = (int32_t)(((U[k] >> rb) | (U[k + 1] << (29 - rb))) & FX_M29);
    const uint32_t top = ((U[N - 1] >> rb) | ((N < M ? U[N] : ext) << (29 - rb))) & FX_M29;
    uint32_t bad = ((N < M ? U[N] : ext) >> rb) ^ (ext >> rb);
#pragma unroll
    for (int i = N + 1; i < M; i++) bad |= U[i] ^ ext;
    r[N - 1] = (int32_t)top - (ext ? (int32_t)(1u << 29) : 0);
    return bad != 0;
}

// r = (a b - c d) 2^-s in the fixed point, a > 0, for SMIN <= s <= 4; the partial products of columns below
// S0 are dropped (fx_tier).  Returns true when |r| >= 8 (out of range; r is then meaningless).
template <int L, int S0, int SMIN>
MPX_HD bool fx_upd(const int32_t (&a)[fx_limbs<L>()], const int32_t (&b)[fx_limbs<L>()],
                   const int32_t (&c)[fx_limbs<L>()], const int32_t (&d)[fx_limbs<L>()], int32_t s,
                   int32_t (&r)[fx_limbs<L>()]) {
    constexpr int N = fx_limbs<L>(), NC = 2 * N - 1 - S0;
    int64_t col[NC];
    {
        int32_t cn[N];
#pragma unroll
        for (int k = 0; k < N; k++) cn[k] = -c[k];
        // each column starts with its first partial products (a[0] b[q] for the columns q < N, a[p] b[N-1] for
        // the columns N-1+p above); the rest accumulate row by row, four independent multiply-adds (adjacent
        // columns) at a time
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int cc = k + S0, p0 = cc < N ? 0 : cc - N + 1;
            col[k] = (int64_t)a[p0] * b[cc - p0] + (int64_t)cn[p0] * d[cc - p0];
        }
        fx_rows<N, S0>(a, b, col);
        fx_rows<N, S0>(cn, d, col);
    }
    // carry: canonical limbs w (columns S0 ..), signed remainder cy
    uint32_t w[NC + 2];
    int64_t cy = 0;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const int64_t v = col[k] + cy;
        w[k] = (uint32_t)v & FX_M29;
        cy = v >> 29;
    }
    w[NC] = (uint32_t)cy & FX_M29;
    w[NC + 1] = (uint32_t)(cy >> 29) & FX_M29;
    const uint32_t ext = cy < 0 ? FX_M29 : 0u;  // the limbs above (|cy| < 2^58)
    // result = W 2^(29 S0 - 2F - s) in units of 2^-F: floor(W / 2^off) with off = F + s - 29 S0
    const int32_t off = fx_frac<L>() + s - 29 * S0;
    constexpr int NU = NC + 2;
    // the limb offset q = off / 29 lies in [QMIN, QMAX] for SMIN <= s <= 4
    constexpr int OMIN = fx_frac<L>() + SMIN - 29 * S0, OMAX = fx_frac<L>() + 4 - 29 * S0;
    constexpr int QMIN = OMIN >= 0 ? OMIN / 29 : 0, QMAX = OMAX / 29, QR = QMAX - QMIN;
    if (OMIN >= 0 || off >= 0) {
        const uint32_t q = (uint32_t)off / 29u, rb = (uint32_t)off - 29u * q, qd = q - QMIN;
        // U[i] = w[i + q] (the sign extension beyond)
        uint32_t U[NU + 1];
#pragma unroll
        for (int i = 0; i <= NU; i++) U[i] = i + QMIN < NU ? w[i + QMIN] : ext;
#pragma unroll
        for (int lg = 4; lg >= 0; lg--) {
            const int st = 1 << lg;
            if (st <= QR) {
                const bool take = (qd & st) != 0;
#pragma unroll
                for (int i = 0; i <= NU; i++) U[i] = take ? (i + st <= NU ? U[i + st] : ext) : U[i];
            }
        }
        return fx_extract<N, NU + 1>(U, rb, ext, r);
    }
    // left shift by -off (a step in which the pivot and the row had both been renormalised after deep
    // cancellation; only with every partial product kept, S0 = 0): W moved up by Q limbs, then down by the rest
    const uint32_t sh = (uint32_t)(-off);
    if (sh >= 29u * N) {
        uint32_t o = 0;
#pragma unroll
        for (int k = 0; k < NU; k++) o |= w[k];
#pragma unroll
        for (int k = 0; k < N; k++) r[k] = 0;
        return o != 0;
    }
    const uint32_t Q = (sh + 28u) / 29u, rb = 29u * Q - sh;
    constexpr int NV = NU + N + 1;
    uint32_t U[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) U[i] = i < NU ? w[i] : ext;
#pragma unroll
    for (int lg = 4; lg >= 0; lg--) {
        const int st = 1 << lg;
        if (st <= N) {
            const bool take = (Q & st) != 0;
#pragma unroll
            for (int i = NV - 1; i >= 0; i--) U[i] = take ? (i >= st ? U[i - st] : 0u) : U[i];
        }
    }
    return fx_extract<N, NV>(U, rb, ext, r);
}

// Lowest product column for a shift s: the columns below S0 weigh at most 2 S0 2^(29 (S0 - 1) + 58) units of
// 2^-2F, kept below 2^-10 of the result's last bit 2^(F + s): S0 = N - 2 for s >= -10 (tier 0), N - 4 for
// s >= -68 (tier 1), else every column (tier 2).
template <int L>
MPX_HD int fx_tier(int32_t s) {
    return s >= -10 ? 0 : (s >= -68 ? 1 : 2);
}
template <int L>
MPX_HD bool fx_upd_tier(int tier, const int32_t (&a)[fx_limbs<L>()], const int32_t (&b)[fx_limbs<L>()],
                        const int32_t (&c)[fx_limbs<L>()], const int32_t (&d)[fx_limbs<L>()], int32_t s,
                        int32_t (&r)[fx_limbs<L>()]) {
    constexpr int N = fx_limbs<L>(), T0 = N - 2 > 0 ? N - 2 : 0, T1 = N - 4 > 0 ? N - 4 : 0;
    if (tier == 0) return fx_upd<L, T0, -10>(a, b, c, d, s, r);
    if (tier == 1) return fx_upd<L, T1, -68>(a, b, c, d, s, r);
    return fx_upd<L, 0, -(1 << 20)>(a, b, c, d, s, r);
}

// The product tier of a wave: the deepest any of its lanes needs (one code path per wave on the device;
// per entry on the host).
MPX_HD int uniform_tier(int t) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __any(t == 2) ? 2 : (__any(t == 1) ? 1 : 0);
#else
    return t;
#endif
}

// *p = max(*p, v) (an LDS atomic on the device)
MPX_HD void fx_amax(int32_t *p, int32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicMax(p, v);
#else
    *p = v > *p ? v : *p;
#endif
}

// One workgroup's share of a block column: the nb x nb diagonal block (lower triangle) and R panel rows,
// in caller-provided memory (LDS on the device, host arrays in the serial build).
//   V     the entries, SW = N | 1 words each (an odd stride: consecutive entries fall in distinct LDS
//         banks; each word at a constant offset from the entry's first): entry (r, c) at (r + LDV c) SW,
//         rows 0..nb-1 the block, nb..nb+R-1 the panel rows, columns 0..nb-1; the panel rows' diagonals
//         follow at (LDV nb + p) SW
//   E     the work of a step, one int per item (row | col << 16; col = 0xFFFF for a panel diagonal):
//         the panel diagonals first, then the columns from the last to the second, each with its block
//         entries and its panel entries; step j updates the items whose column exceeds j (a panel diagonal
//         at every step but the last), a prefix of this order, so that the items still active fill whole
//         waves (the device kernel keeps its one item per lane in registers and needs no E)
//   hist  (nb + 1) rows of LDV ints, all FX_HZERO at first: hist[j * LDV + r] sets row r's renormalisation
//         entering step j (d_r = fx_delta of it): for a block row the leading exponent h of its diagonal;
//         for a panel row the larger of that and 2 h + 1 over its entries, so that a panel row whose own
//         pivot (in a later block) fails is still scaled by its entries
//   f0    LDV ints: the initial row scale exponents; pm  R ints: the panel rows' entry exponents at set-up
//   Fpre  nb rows of LDV ints: Fpre[c * LDV + r] = f_r^(c)
//   Gm, Rh  (L+1) planes of nb words each (mp values): normalised pivots / their prefix products, rho
// Set-up order (each stage after the previous one has finished in every thread): block_scale for the
// block rows, panel_scale0 for the panel rows; panel_rowmax for every panel entry; panel_scale for the panel
// rows; load_entry for every entry of the block column (r >= c), load_pdiag for the panel diagonals and
// set_item for every item.
template <int L>
struct ColElim {
    static constexpr int N = fx_limbs<L>(), SW = N | 1;
    uint32_t *V, *Gm, *Rh;
    int32_t *E, *hist, *f0, *pm, *Fpre;
    int nb, R, LDV;

    MPX_HD void ld(int idx, int32_t (&x)[N]) const {
        const uint32_t *p = V + idx * SW;
#pragma unroll
        for (int q = 0; q < N; q++) x[q] = (int32_t)p[q];
    }
    MPX_HD void st(int idx, const int32_t (&x)[N]) const {
        uint32_t *p = V + idx * SW;
#pragma unroll
        for (int q = 0; q < N; q++) p[q] = (uint32_t)x[q];
    }
    MPX_HD int didx(int p) const {  // panel diagonal p
        return LDV * nb + p;
    }
    MPX_HD mp<L> gld(const uint32_t *A, int t) const {
        mp<L> x;
#pragma unroll
        for (int q = 0; q < L; q++) x.m[q] = A[q * nb + t];
        const int32_t hh = (int32_t)A[L * nb + t];
        x.e = hh >> 1;
        x.s = (uint32_t)hh & 1u;
        return x;
    }
    MPX_HD void gst(uint32_t *A, int t, const mp<L> &x) const {
#pragma unroll
        for (int q = 0; q < L; q++) A[q * nb + t] = x.m[q];
        A[L * nb + t] = (uint32_t)((x.e << 1) | (int32_t)(x.s & 1u));
    }

    // ---- set-up ----
    // row scale exponent from a diagonal entry: 2^(2f) ~ |a|
    MPX_HD static int32_t scale_of(const mp<L> &a) {
        return is_zero<L>(a) ? 0 : (a.e >> 1);
    }
    MPX_HD void block_scale(int r, const mp<L> &arr) const {
        f0[r] = scale_of(arr);
    }
    MPX_HD void panel_scale0(int p, const mp<L> &app) const {
        f0[nb + p] = is_zero<L>(app) ? FX_HZERO : (app.e >> 1);
        pm[p] = FX_HZERO;
    }
    // entry (nb + p, c): |a 2^-f_c| < 2^(e - f_c)
    MPX_HD void panel_rowmax(int p, int c, const mp<L> &a) const {
        if (!is_zero<L>(a)) fx_amax(pm + p, a.e - f0[c]);
    }
    // the panel row's scale: its diagonal's, raised to its largest entry (so every entry is at most 1)
    MPX_HD void panel_scale(int p) const {
        const int32_t f = f0[nb + p] > pm[p] ? f0[nb + p] : pm[p];
        f0[nb + p] = f == FX_HZERO ? 0 : f;
    }
    // store entry (r, c) from its mp value; true when out of range
    MPX_HD bool load_entry(int r, int c, const mp<L> &a) const {
        int32_t x[N];
        const bool ovf = fx_from_mp<L>(a, f0[r] + f0[c], x);
        st(r + LDV * c, x);
        if (r == c)
            hist[r] = fx_lead<N>(x);
        else if (r >= nb)
            fx_amax(hist + r, 2 * fx_lead<N>(x) + 1);
        return ovf;
    }
    MPX_HD bool load_pdiag(int p, const mp<L> &a) const {
        int32_t x[N];
        const bool ovf = fx_from_mp<L>(a, 2 * f0[nb + p], x);
        st(didx(p), x);
        fx_amax(hist + nb + p, fx_lead<N>(x));
        return ovf;
    }
    // items: nb (nb - 1) / 2 block entries, R (nb - 1) panel entries (columns 1..nb-1), R panel diagonals
    MPX_HD int nitems() const {
        return nb * (nb - 1) / 2 + R * (nb - 1) + R;
    }
    // item t: the R panel diagonals, then column c = nb-1 down to 1 with its nb - c block entries (rows
    // c..nb-1) and its R panel entries
    MPX_HD int item_code(int t) const {
        if (t < R) return (nb + t) | (0xFFFF << 16);
        t -= R;
        // the m = nb - 1 - c columns above column c hold m (m + 1) / 2 + m R items: the largest m with that <= t
        const float b = 2.0f * R + 1.0f;
        int m = (int)((sqrtf(b * b + 8.0f * t) - b) * 0.5f);
        while (m > 0 && m * (m + 1) / 2 + m * R > t) m--;
        while ((m + 1) * (m + 2) / 2 + (m + 1) * R <= t) m++;
        if (m > nb - 2) m = nb - 2;
        const int c = nb - 1 - m, i = t - (m * (m + 1) / 2 + m * R);
        const int row = i < nb - c ? c + i : nb + (i - (nb - c));
        return row | (c << 16);
    }
    MPX_HD void set_item(int t) const {
        E[t] = item_code(t);
    }
    // entry t of the block column as loaded at set-up: the block's lower triangle column by column
    // (nb (nb + 1) / 2), the panel entries (R nb, column by column), the R panel diagonals.  Returns the
    // row and sets the column (-1 for a panel diagonal).
    MPX_HD int entry_code(int t, int &col) const {
        const int nbl = nb * (nb + 1) / 2;
        if (t < nbl) {
            // columns before c hold c nb - c (c - 1) / 2 entries: the largest c with that <= t
            const float b = 2.0f * nb + 1.0f;
            int c = (int)((b - sqrtf(b * b - 8.0f * t)) * 0.5f);
            while (c > 0 && c * nb - c * (c - 1) / 2 > t) c--;
            while ((c + 1) * nb - (c + 1) * c / 2 <= t) c++;
            col = c;
            return c + t - (c * nb - c * (c - 1) / 2);
        }
        t -= nbl;
        if (t < R * nb) {
            col = t / R;
            return nb + t % R;
        }
        col = -1;
        return nb + (t - R * nb);
    }

    // ---- elimination ----
    // pivot of step j: failed (<= 0)?
    MPX_HD bool pivot_fails(int j) const {
        int32_t a[N];
        ld(j + LDV * j, a);
        return fx_nonpos<N>(a);
    }
    // item t at step j with the pivot a = V_jj and dj its renormalisation exponent.  Returns 0 (also for an
    // item not updated at this step), 1 (a block entry out of range) or 2 (a panel entry out of range).
    MPX_HD int update(int j, int t, const int32_t (&a)[N], int32_t dj) const {
        return update_item(j, E[t], a, dj);
    }
    // the same for the item whose code (item_code) is it
    MPX_HD int update_item(int j, int32_t it, const int32_t (&a)[N], int32_t dj) const {
        if (!item_active(j, it)) return 0;
        ItemOps o;
        item_load(j, it, o);
        return item_update(j, it, a, dj, o);
    }
    // the same in two parts, so that a caller can request the operands with the pivot: is the item updated at
    // step j; its operands (entry, the two pivot-column entries, the two row exponents); the update
    struct ItemOps {
        int32_t b[N], c[N], d[N], hr, hc;
    };
    MPX_HD bool item_active(int j, int32_t it) const {
        const int col = (int)((uint32_t)it >> 16);
        return col > j && !(col == 0xFFFF && j == nb - 1);
    }
    MPX_HD void item_load(int j, int32_t it, ItemOps &o) const {
        const int row = it & 0xFFFF, col = (int)((uint32_t)it >> 16);
        const bool pd = col == 0xFFFF;
        ld(pd ? didx(row - nb) : row + LDV * col, o.b);
        ld(row + LDV * j, o.c);
        ld(pd ? row + LDV * j : col + LDV * j, o.d);
        o.hr = hist[j * LDV + row];
        o.hc = pd ? o.hr : hist[j * LDV + col];
    }
    MPX_HD int item_update(int j, int32_t it, const int32_t (&a)[N], int32_t dj, const ItemOps &o) const {
        const int row = it & 0xFFFF, col = (int)((uint32_t)it >> 16);
        int32_t *hn = hist + (j + 1) * LDV;
        const bool pd = col == 0xFFFF;
        int32_t r[N];
        const int32_t s = 2 * dj + fx_delta(o.hr) + fx_delta(o.hc);
        const bool ovf = fx_upd_tier<L>(uniform_tier(fx_tier<L>(s)), a, o.b, o.c, o.d, s, r);
        st(pd ? didx(row - nb) : row + LDV * col, r);
        if (pd || row == col)
            fx_amax(hn + row, fx_lead<N>(r));
        else if (row >= nb)
            fx_amax(hn + row, 2 * fx_lead<N>(r) + 1);
        return ovf ? (row < nb ? 1 : 2) : 0;
    }

    // ---- after the chain ----
    // f_r^(c) alone (c <= r for a block row), into Fpre
    MPX_HD void prefix_scale(int r, int c) const {
        int32_t f = f0[r];
        for (int q = 0; q < c; q++) f += fx_delta(hist[q * LDV + r]);
        Fpre[c * LDV + r] = f;
    }
    // row r: f_r^(c) for every column c it needs (block rows: c <= r; panel rows: all)
    MPX_HD void prefix_scales(int r) const {
        int32_t f = f0[r];
        const int cmax = r < nb ? r : nb - 1;
        for (int c = 0; c <= cmax; c++) {
            Fpre[c * LDV + r] = f;
            f += fx_delta(hist[c * LDV + r]);
        }
    }
    // normalised pivot Vh_c = V_cc 2^-(2 d_c) as an mp value
    MPX_HD mp<L> pivot_mp(int c) const {
        int32_t x[N];
        ld(c + LDV * c, x);
        return fx_to_mp<L>(x, -2 * fx_delta(hist[c * LDV + c]));
    }
    // the same into Gm[c]
    MPX_HD void pivot_value(int c) const {
        gst(Gm, c, pivot_mp(c));
    }
    // rho_c = 1/sqrt(gamma_c Vh_c) into Rh[c] (Gm holds the exclusive prefix products gamma_c, Vh in vh);
    // returns 1 / L_cc
    MPX_HD mp<L> finish_pivot(int c, const mp<L> &vh) const {
        const mp<L> g = gld(Gm, c);
        const mp<L> rho = rsqrt_fx<L>(mul<L>(g, vh));
        gst(Rh, c, rho);
        return ldexp2<L>(mul<L>(g, rho), -(Fpre[c * LDV + c] + fx_delta(hist[c * LDV + c])));
    }
    // L_rc
    MPX_HD mp<L> factor_entry(int r, int c) const {
        int32_t x[N];
        ld(r + LDV * c, x);
        return mul<L>(fx_to_mp<L>(x, Fpre[c * LDV + r] - fx_delta(hist[c * LDV + c])), gld(Rh, c));
    }
};

}  // namespace mpx
