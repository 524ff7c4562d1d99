// limbs.h — the limb counts instantiated by this build.
// A p-bit request runs with GMP's mpf precision: prec = (max(53,p)+127)/64 limbs of 64
// bits (gmp-impl.h __GMPF_BITS_TO_PREC), i.e. L = 2*prec 32-bit limbs; 256 bits -> L = 10.
#pragma once
#ifdef SDPG_FAST_BUILD
#define SDPG_FOR_EACH_L(X) X(10)
#else
#define SDPG_FOR_EACH_L(X) X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18)
#endif
