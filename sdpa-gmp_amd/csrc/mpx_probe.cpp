// mpx_probe.cpp — host-only C entry points exposing the mpx arithmetic to tests
// (no GPU needed): element ops, decimal parsing and formatting at limb count L.
#include <string.h>

#include <string>

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
#define DISPATCH(L, CALL) \
    switch (L) {          \
    case 4: return CALL(4); \
    case 8: return CALL(8); \
    case 10: return CALL(10); \
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
extern "C" int mpx_probe_limbs(int bits) { return mpx::limbs_for_bits(bits); }
