/*
 * ompf — multiple-precision floating point for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is part of the oracle: a CPU
 * restatement of the SDPA-GMP Newton path used by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py.  Product code must not
 * include, link or call anything under oracle/.
 *
 * Restates the arithmetic contract of GMP 6.2.1 `mpf_t`, which the
 * reference reaches through `mpf_class` (sdpa_algebra.h:9-12,
 * mplapack/mpblas_gmp.h:36-38):
 *   - 64-bit limbs, value = 0.d[size-1] d[size-2] ... d[0] * B^exp, B=2^64,
 *     top limb non-zero, sign carried in the sign of `size`;
 *   - precision p bits -> prec = (max(53,p) + 127) / 64 limbs
 *     (gmp-impl.h __GMPF_BITS_TO_PREC), results keep at most prec+1 limbs;
 *   - every operation truncates (rounds toward zero); mul truncates its
 *     operands to prec limbs first (gmp mpf/mul.c).
 * Add and sub are computed exactly on the stored limbs and then truncated to
 * prec+1 limbs; GMP's own window truncations inside mpf/add.c and mpf/sub.c
 * are not reproduced limb for limb, so results agree with GMP to the last
 * limb (about 2^-64*prec relative), not bit for bit.
 */
#ifndef OMPF_H
#define OMPF_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMPF_MAXL 20 /* prec+1 <= 20 limbs: up to 1152 requested bits */

typedef struct {
    int32_t size; /* +-(number of limbs); 0 means zero */
    int32_t exp;  /* exponent in limbs */
    uint64_t d[OMPF_MAXL];
} ompf;

/* global precision (like mpf_set_default_prec) */
void ompf_set_prec_bits(int bits);
int ompf_prec_limbs(void); /* GMP's _mp_prec */
int ompf_prec_bits(void);  /* __GMPF_PREC_TO_BITS(prec) */

void ompf_set_zero(ompf *r);
void ompf_set(ompf *r, const ompf *a);
void ompf_set_d(ompf *r, double x);
void ompf_set_si(ompf *r, long x);
double ompf_get_d(const ompf *a);
/* decimal text -> value; exact when the text denotes a dyadic rational that
   fits the precision.  Returns number of characters consumed (0 on error). */
int ompf_set_str(ompf *r, const char *s);

void ompf_add(ompf *r, const ompf *a, const ompf *b);
void ompf_sub(ompf *r, const ompf *a, const ompf *b);
void ompf_mul(ompf *r, const ompf *a, const ompf *b);
void ompf_div(ompf *r, const ompf *a, const ompf *b);
void ompf_sqrt(ompf *r, const ompf *a);
void ompf_neg(ompf *r, const ompf *a);
void ompf_abs(ompf *r, const ompf *a);
void ompf_mul_2exp(ompf *r, const ompf *a, long k); /* a * 2^k (k may be <0) */
int ompf_cmp(const ompf *a, const ompf *b);
int ompf_cmp_d(const ompf *a, double x);
int ompf_sgn(const ompf *a);

/* r += a*b (two truncating ops, like mpf_class `r += a*b`) */
void ompf_addmul(ompf *r, const ompf *a, const ompf *b);
void ompf_submul(ompf *r, const ompf *a, const ompf *b);

/* Decimal output "d.ddddde[+-]xx" with `digits` significant digits,
   correctly rounded (half away from zero) from the exact stored value. */
void ompf_get_str(char *buf, size_t bufsz, const ompf *a, int digits);

/* Export to the interchange format used by tests: sign, binary exponent e and
   L little-endian 32-bit limbs M with value = sign * M * 2^(e-32L) and the top
   bit of M set (truncated to L limbs).  Zero: sign 0, M = 0. */
void ompf_export32(const ompf *a, int L, int32_t *sign, int32_t *e, uint32_t *limbs);
/* Import from the same format (exact when 32L <= 64(prec+1) bits). */
void ompf_import32(ompf *r, int L, int32_t sign, int32_t e, const uint32_t *limbs);

#ifdef __cplusplus
}
#endif
#endif
