// mpx_probe.cpp — host-only C entry points exposing the mpx arithmetic to tests
// (no GPU needed): element ops, decimal parsing and formatting at limb count L.
#include <string.h>

#include <string>
#include <vector>

#include "chol_fx.h"
#include "mpx.h"
#include "mpx_host.h"

template <int L>
static mpx::mp<L> ld(const unsigned *w) {
    mpx::mp<L> x;
    x.e = (int)w[0];
    x.s = w[1] & 1u;
    for (int k = 0; k < L; k++) x.m[k] = w[2 + k];
    if (x.m[L - 1] == 0) x = mpx::zero<L>();
    return x;
}
template <int L>
static void st(const mpx::mp<L> &x, unsigned *w) {
    w[0] = (unsigned)x.e;
    w[1] = x.s;
    for (int k = 0; k < L; k++) w[2 + k] = x.m[k];
}
template <int L>
static int op_L(int op, const unsigned *a, const unsigned *b, unsigned *r) {
    mpx::mp<L> x = ld<L>(a), y = ld<L>(b), z;
    switch (op) {
    case 0: z = mpx::add<L>(x, y); break;
    case 1: z = mpx::sub<L>(x, y); break;
    case 2: z = mpx::mul<L>(x, y); break;
    case 3: z = mpx::div<L>(x, y); break;
    case 4: z = mpx::sqrtv<L>(x); break;
    case 5: z = mpx::recip<L>(x); break;
    case 6: z = mpx::rsqrt<L>(x); break;
    case 7: z = mpx::rsqrt_fx<L>(x); break;
    default: return -1;
    }
    st<L>(z, r);
    return 0;
}
template <int L>
static int fms_L(const unsigned *a, const unsigned *b, const unsigned *c, const unsigned *d, unsigned *r) {
    st<L>(mpx::fms<L>(ld<L>(a), ld<L>(b), ld<L>(c), ld<L>(d)), r);
    return 0;
}
template <int L>
static int parse_L(const char *s, unsigned *r) {
    mpx::mp<L> x;
    if (mpx::parse_decimal<L>(s, x) <= 0) return -1;
    st<L>(x, r);
    return 0;
}
template <int L>
static int fmt_L(const unsigned *a, int digits, char *buf, int n) {
    std::string s = mpx::format_decimal<L>(ld<L>(a), digits);
    strncpy(buf, s.c_str(), n - 1);
    buf[n - 1] = 0;
    return 0;
}
// Serial build of the block-column elimination of the blocked Cholesky (chol_fx.h), in the order the device
// kernel follows: A holds the block column, (nb + R) x nb column-major (entry (r, c) at r + (nb + R) c; the
// block's upper triangle is not read), pd the R panel rows' diagonals; Lout receives the factor entries in
// the same layout and rdiag the reciprocal pivots 1 / L_cc.  info: 0, or the 1-based pivot reported (the
// failing pivot, or for an entry out of the fixed-point range the first pivot not yet confirmed, nb + 1 for
// a panel row).
template <int L>
static int potrf_col_L(int nb, int R, const unsigned *A, const unsigned *pd, unsigned *Lout, unsigned *rdiag,
                       int *info) {
    const int LDV = nb + R;
    using CE = mpx::ColElim<L>;
    std::vector<uint32_t> V((size_t)(LDV * nb + R) * CE::SW), Gm((size_t)(L + 1) * nb), Rh((size_t)(L + 1) * nb);
    std::vector<int32_t> E((size_t)nb * nb + (size_t)R * nb + 1), hist((size_t)(nb + 1) * LDV, mpx::FX_HZERO),
        f0(LDV), pm(R ? R : 1), Fpre((size_t)nb * LDV);
    const CE ce{V.data(), Gm.data(), Rh.data(), E.data(), hist.data(), f0.data(), pm.data(), Fpre.data(), nb, R, LDV};
    auto val = [&](int r, int c) { return ld<L>(A + (size_t)(r + LDV * c) * (L + 2)); };
    auto pdv = [&](int p) { return ld<L>(pd + (size_t)p * (L + 2)); };
    *info = 0;
    for (int r = 0; r < nb; r++) ce.block_scale(r, val(r, r));
    for (int p = 0; p < R; p++) ce.panel_scale0(p, pdv(p));
    for (int p = 0; p < R; p++)
        for (int c = 0; c < nb; c++) ce.panel_rowmax(p, c, val(nb + p, c));
    for (int p = 0; p < R; p++) ce.panel_scale(p);
    bool ovf_blk = false, ovf_pan = false;
    for (int c = 0; c < nb; c++)
        for (int r = c; r < LDV; r++) {
            const bool o = ce.load_entry(r, c, val(r, c));
            if (r < nb)
                ovf_blk |= o;
            else
                ovf_pan |= o;
        }
    for (int p = 0; p < R; p++) ovf_pan |= ce.load_pdiag(p, pdv(p));
    for (int t = 0; t < ce.nitems(); t++) ce.set_item(t);
    if (ovf_blk) {
        *info = 1;
        return 0;
    }
    for (int j = 0; j < nb; j++) {
        if (ce.pivot_fails(j)) {
            *info = j + 1;
            return 0;
        }
        if (j == nb - 1) break;
        int32_t a[CE::N];
        ce.ld(j + LDV * j, a);
        const int32_t dj = mpx::fx_delta(hist[(size_t)j * LDV + j]);
        bool blk = false;
        for (int t = 0, n = ce.nitems(); t < n; t++) {
            const int o = ce.update(j, t, a, dj);
            blk |= o == 1;
            ovf_pan |= o == 2;
        }
        if (blk) {
            *info = j + 2;
            return 0;
        }
    }
    if (ovf_pan) {
        *info = nb + 1;
        return 0;
    }
    for (int r = 0; r < LDV; r++) ce.prefix_scales(r);
    std::vector<mpx::mp<L>> vh(nb);
    for (int c = 0; c < nb; c++) {
        ce.pivot_value(c);
        vh[c] = ce.gld(Gm.data(), c);
    }
    mpx::mp<L> g = mpx::from_int<L>(1);
    for (int c = 0; c < nb; c++) {
        ce.gst(Gm.data(), c, g);
        g = mpx::mul<L>(g, vh[c]);
    }
    for (int c = 0; c < nb; c++) st<L>(ce.finish_pivot(c, vh[c]), rdiag + (size_t)c * (L + 2));
    for (int c = 0; c < nb; c++)
        for (int r = c; r < LDV; r++) st<L>(ce.factor_entry(r, c), Lout + (size_t)(r + LDV * c) * (L + 2));
    return 0;
}
template <int L>
static int fxupd_L(const unsigned *a, const unsigned *b, const unsigned *c, const unsigned *d, int s, unsigned *r,
                   int *h) {
    constexpr int N = mpx::fx_limbs<L>();
    int32_t A[N], B[N], C[N], D[N], Rr[N];
    for (int k = 0; k < N; k++) {
        A[k] = (int32_t)a[k];
        B[k] = (int32_t)b[k];
        C[k] = (int32_t)c[k];
        D[k] = (int32_t)d[k];
    }
    const bool ovf = mpx::fx_upd_tier<L>(mpx::fx_tier<L>(s), A, B, C, D, s, Rr);
    for (int k = 0; k < N; k++) r[k] = (unsigned)Rr[k];
    *h = ovf ? 0 : mpx::fx_lead<N>(Rr);
    return ovf ? 1 : 0;
}
#define DISPATCH(L, CALL) \
    switch (L) {          \
    case 4: return CALL(4); \
    case 8: return CALL(8); \
    case 10: return CALL(10); \
    case 12: return CALL(12); \
    case 16: return CALL(16); \
    case 18: return CALL(18); \
    }                     \
    return -2;

extern "C" int mpx_probe_op(int L, int op, const unsigned *a, const unsigned *b, unsigned *r) {
#define C_(LL) op_L<LL>(op, a, b, r)
    DISPATCH(L, C_)
#undef C_
}
// r = a * b - c * d (mpx::fms)
extern "C" int mpx_probe_fms(int L, const unsigned *a, const unsigned *b, const unsigned *c, const unsigned *d,
                             unsigned *r) {
#define C_(LL) fms_L<LL>(a, b, c, d, r)
    DISPATCH(L, C_)
#undef C_
}
extern "C" int mpx_probe_parse(int L, const char *s, unsigned *r) {
#define C_(LL) parse_L<LL>(s, r)
    DISPATCH(L, C_)
#undef C_
}
extern "C" int mpx_probe_format(int L, const unsigned *a, int digits, char *buf, int n) {
#define C_(LL) fmt_L<LL>(a, digits, buf, n)
    DISPATCH(L, C_)
#undef C_
}
extern "C" int mpx_probe_potrf_col(int L, int nb, int R, const unsigned *A, const unsigned *pd, unsigned *Lout,
                                   unsigned *rdiag, int *info) {
#define C_(LL) potrf_col_L<LL>(nb, R, A, pd, Lout, rdiag, info)
    DISPATCH(L, C_)
#undef C_
}
// r = (a b - c d) 2^-s in the block-column elimination's fixed point (chol_fx.h fx_upd); returns 1 when out
// of range
extern "C" int mpx_probe_fx_upd(int L, const unsigned *a, const unsigned *b, const unsigned *c, const unsigned *d,
                                int s, unsigned *r, int *h) {
#define C_(LL) fxupd_L<LL>(a, b, c, d, s, r, h)
    DISPATCH(L, C_)
#undef C_
}
extern "C" int mpx_probe_fx_limbs(int L) {
#define C_(LL) mpx::fx_limbs<LL>()
    DISPATCH(L, C_)
#undef C_
}
extern "C" int mpx_probe_limbs(int bits) { return mpx::limbs_for_bits(bits); }
// the block-column kernel's enumerations for nb block columns and R panel rows: items[t] = item_code(t)
// (t < nitems), entries[2 t], entries[2 t + 1] = the row and column of entry_code(t) (t < nb (nb + 1) / 2 +
// R (nb + 1)); returns the item count
extern "C" int mpx_probe_col_enum(int nb, int R, int *items, int *entries) {
    const mpx::ColElim<10> ce{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nb, R, nb + R};
    const int n = ce.nitems();
    for (int t = 0; t < n; t++) items[t] = ce.item_code(t);
    for (int t = 0, ne = nb * (nb + 1) / 2 + R * (nb + 1); t < ne; t++) {
        int col = 0;
        entries[2 * t] = ce.entry_code(t, col);
        entries[2 * t + 1] = col;
    }
    return n;
}
