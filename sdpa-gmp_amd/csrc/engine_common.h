// engine_common.h — device views and shared helpers of the Newton engine.
//
// Device layout ("limb-planar"): an array of N multiple-precision elements is
// one allocation of (L+1)*N 32-bit words; word k*N + i is limb k of element i
// (k < L) and word L*N + i is its head (exponent << 1 | sign).  Consecutive
// threads touching consecutive elements therefore read consecutive words of
// every limb plane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpx.h"

namespace sdpg {

using mpx::mp;

struct PV {  // planar view: base pointer, plane stride (elements)
    uint32_t *p;
    int64_t n;
};

template <int L>
__device__ __forceinline__ mp<L> ldm(PV v, int64_t i) {
    mp<L> x;
#pragma unroll
    for (int k = 0; k < L; k++) x.m[k] = v.p[(int64_t)k * v.n + i];
    int32_t h = (int32_t)v.p[(int64_t)L * v.n + i];
    x.e = h >> 1;
    x.s = (uint32_t)h & 1u;
    return x;
}

template <int L>
__device__ __forceinline__ void stm(PV v, int64_t i, const mp<L> &x) {
#pragma unroll
    for (int k = 0; k < L; k++) v.p[(int64_t)k * v.n + i] = x.m[k];
    v.p[(int64_t)L * v.n + i] = (uint32_t)((x.e << 1) | (int32_t)(x.s & 1u));
}

template <int L>
__device__ __forceinline__ mp<L> fma_acc(const mp<L> &acc, const mp<L> &a, const mp<L> &b) {
    return mpx::add<L>(acc, mpx::mul<L>(a, b));
}

}  // namespace sdpg
