// mpx_host.h — host-only conversions for mpx::mp<L>: decimal text in and out.
//
// Parsing is exact whenever the decimal text denotes a dyadic rational whose
// odd part fits the mantissa (the inputs written by gen_problems.py are of that
// kind); otherwise it is correctly truncated then rounded from 64 extra bits.
// Matches the role of gmp_fscanf("%Fe") in the reference reader
// (sdpa_io.cpp:405-751) and gmp_fprintf("%Fe") in its printer (sdpa_io.cpp:753-994).
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "mpx.h"

namespace mpx {

// little-endian base-2^32 natural numbers
struct nat {
    std::vector<uint32_t> d;
    bool is_zero() const {
        for (uint32_t v : d)
            if (v) return false;
        return true;
    }
    void trim() {
        while (!d.empty() && d.back() == 0) d.pop_back();
    }
    void mul_small(uint32_t m, uint32_t add = 0) {
        uint64_t c = add;
        for (auto &v : d) {
            uint64_t t = (uint64_t)v * m + c;
            v = (uint32_t)t;
            c = t >> 32;
        }
        if (c) d.push_back((uint32_t)c);
    }
    uint32_t div_small(uint32_t q) {  // returns remainder
        uint64_t r = 0;
        for (size_t i = d.size(); i-- > 0;) {
            uint64_t cur = (r << 32) | d[i];
            d[i] = (uint32_t)(cur / q);
            r = cur % q;
        }
        trim();
        return (uint32_t)r;
    }
    int bits() const {
        for (size_t i = d.size(); i-- > 0;)
            if (d[i]) return (int)(32 * i + 32 - __builtin_clz(d[i]));
        return 0;
    }
    bool bit(int k) const {
        size_t w = (size_t)k / 32;
        return w < d.size() && ((d[w] >> (k % 32)) & 1u);
    }
    bool any_below(int k) const {  // any bit set below bit k
        for (int w = 0; w < k / 32 && w < (int)d.size(); w++)
            if (d[w]) return true;
        if (k % 32 && (size_t)(k / 32) < d.size()) return (d[k / 32] & ((1u << (k % 32)) - 1)) != 0;
        return false;
    }
    void shl_bits(int s) {
        int w = s / 32, r = s % 32;
        std::vector<uint32_t> o(d.size() + w + 1, 0);
        for (size_t i = 0; i < d.size(); i++) {
            o[i + w] |= d[i] << r;
            if (r) o[i + w + 1] |= d[i] >> (32 - r);
        }
        d.swap(o);
        trim();
    }
};

// value = N * 2^e2 rounded to L limbs (nearest, ties away)
template <int L>
mp<L> from_nat(const nat &N, long e2, bool negative) {
    mp<L> r = zero<L>();
    int nb = N.bits();
    if (nb == 0) return r;
    // take the top 32L bits, round with bit below
    int drop = nb - 32 * L;  // bits to drop (may be negative: shift left)
    for (int k = 0; k < L; k++) {
        uint32_t v = 0;
        for (int b = 0; b < 32; b++) {
            int src = drop + 32 * k + b;
            if (src >= 0 && N.bit(src)) v |= 1u << b;
        }
        r.m[k] = v;
    }
    r.e = (int32_t)(nb + e2);
    r.s = negative ? 1u : 0u;
    if (drop > 0 && N.bit(drop - 1)) {  // round up
        uint32_t c = 1;
        for (int k = 0; k < L && c; k++) {
            r.m[k] += 1;
            c = (r.m[k] == 0);
        }
        if (c) {
            r.m[L - 1] = 0x80000000u;
            r.e += 1;
        }
    }
    return r;
}

// Parse [+-]digits[.digits][(e|E)[+-]digits]; returns characters consumed (0 on error).
template <int L>
int parse_decimal(const char *s, mp<L> &out) {
    const char *p = s;
    bool neg = false;
    if (*p == '+' || *p == '-') {
        neg = (*p == '-');
        p++;
    }
    nat D;
    D.d.push_back(0);
    int ndig = 0, frac = 0;
    bool dot = false;
    while ((*p >= '0' && *p <= '9') || (*p == '.' && !dot)) {
        if (*p == '.') {
            dot = true;
            p++;
            continue;
        }
        D.mul_small(10, (uint32_t)(*p - '0'));
        ndig++;
        if (dot) frac++;
        p++;
    }
    if (!ndig) return 0;
    long e10 = 0;
    if (*p == 'e' || *p == 'E') {
        const char *q = p + 1;
        bool en = false;
        if (*q == '+' || *q == '-') {
            en = (*q == '-');
            q++;
        }
        if (*q >= '0' && *q <= '9') {
            long v = 0;
            while (*q >= '0' && *q <= '9') {
                v = v * 10 + (*q - '0');
                if (v > 100000) v = 100000;
                q++;
            }
            e10 = en ? -v : v;
            p = q;
        }
    }
    e10 -= frac;
    D.trim();
    if (D.is_zero()) {
        out = zero<L>();
        return (int)(p - s);
    }
    if (e10 >= 0) {
        for (long i = 0; i < e10; i++) D.mul_small(10);
        out = from_nat<L>(D, 0, neg);
    } else {
        long k = -e10;  // value = D / (2^k 5^k)
        // exact path: strip factors of 5 while divisible
        nat T = D;
        long k5 = 0;
        while (k5 < k) {
            nat U = T;
            if (U.div_small(5) != 0) break;
            T = U;
            k5++;
        }
        if (k5 == k) {
            out = from_nat<L>(T, -k, neg);
        } else {
            // general path: floor(D * 2^S / 5^k) with S leaving 64 spare bits
            int S = 32 * L + 64 + (int)(k * 2.33) + 8 - D.bits();
            if (S < 0) S = 0;
            nat Q = D;
            Q.shl_bits(S);
            bool inexact = false;
            for (long i = 0; i < k; i++)
                if (Q.div_small(5)) inexact = true;
            if (inexact) {  // keep a sticky bit so the rounding is not biased down
                Q.shl_bits(1);
                Q.d[0] |= 1u;
                S += 1;
            }
            out = from_nat<L>(Q, -(long)S - k, neg);
        }
    }
    return (int)(p - s);
}

// Decimal text with `digits` significant digits: "[-]d.ddd...e[+-]xx"
template <int L>
std::string format_decimal(const mp<L> &a, int digits) {
    char buf[64];
    if (is_zero(a)) {
        std::string r = "0.";
        r.append(digits > 1 ? digits - 1 : 1, '0');
        return r + "e+00";
    }
    nat M;
    M.d.assign(a.m, a.m + L);
    long e2 = (long)a.e - 32L * L;  // |a| = M * 2^e2
    double approx = to_double<L>(absv<L>(a));
    int d10 = (approx > 0) ? (int)__builtin_floor(__builtin_log10(approx)) : 0;
    if (!(approx > 0) || approx != approx || approx > 1e300) {
        // very large/small: estimate from the binary exponent
        d10 = (int)__builtin_floor((double)(a.e - 1) * 0.30102999566398120);
    }
    for (int attempt = 0; attempt < 4; attempt++) {
        long k = digits - 1 - d10;  // S = round(M * 2^e2 * 10^k)
        nat T = M;
        long ee = e2;
        if (k >= 0) {
            for (long i = 0; i < k; i++) T.mul_small(5);
            ee += k;  // 10^k = 5^k 2^k
        } else {
            long kk = -k;
            int extra = (int)(kk * 2.33) + 40;
            T.shl_bits(extra);
            ee -= extra;
            for (long i = 0; i < kk; i++) T.div_small(5);
            ee -= kk;
        }
        // S = round(T * 2^ee)
        if (ee < 0) {
            int sh = (int)(-ee);
            bool half = sh >= 1 && T.bit(sh - 1);
            nat U;
            int nb = T.bits();
            U.d.assign((nb > sh) ? (size_t)((nb - sh + 31) / 32) : 1, 0);
            for (int b = sh; b < nb; b++)
                if (T.bit(b)) U.d[(b - sh) / 32] |= 1u << ((b - sh) % 32);
            T = U;
            if (half) {
                uint64_t c = 1;
                for (auto &v : T.d) {
                    uint64_t t = (uint64_t)v + c;
                    v = (uint32_t)t;
                    c = t >> 32;
                    if (!c) break;
                }
                if (c) T.d.push_back(1);
            }
        } else if (ee > 0) {
            T.shl_bits((int)ee);
        }
        T.trim();
        std::string ds;
        nat W = T;
        if (W.d.empty()) ds = "0";
        while (!W.d.empty()) {
            ds.push_back((char)('0' + W.div_small(10)));
        }
        if ((int)ds.size() == digits + 1) {
            d10++;
            continue;
        }
        if ((int)ds.size() < digits) {
            d10--;
            continue;
        }
        std::string out;
        if (a.s) out.push_back('-');
        out.push_back(ds.back());
        if (digits > 1) {
            out.push_back('.');
            for (int i = (int)ds.size() - 2; i >= 0; i--) out.push_back(ds[i]);
        }
        snprintf(buf, sizeof buf, "e%c%02d", d10 < 0 ? '-' : '+', d10 < 0 ? -d10 : d10);
        return out + buf;
    }
    snprintf(buf, sizeof buf, "%.*e", digits - 1, to_double<L>(a));
    return buf;
}

}  // namespace mpx
