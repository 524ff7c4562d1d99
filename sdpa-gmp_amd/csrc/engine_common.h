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

#include <vector>

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

// A bump allocator over device memory that the caller owns (the no-allocation GEMM entry point
// hands it a static device array): take() returns 256-byte aligned pieces and never frees, so a
// workspace that draws from it calls no hipMalloc / hipFree.
struct Arena {
    uint8_t *base = nullptr;
    size_t cap = 0, used = 0;
    void *take(size_t bytes) {
        size_t at = (used + 255) & ~(size_t)255;
        if (at > cap || bytes > cap - at) return nullptr;
        used = at + bytes;
        return base + at;
    }
};

// Per-kernel device time over a measured region: events recorded on the launching stream around
// each launch of the kernels of interest, summed after the region ends.
// kinds: 0 Ozaki MFMA product kernel (short K), 1 plain limb product kernel, 2 Cholesky block column
// (diagonal block and panel rows), 3 Ozaki MFMA product kernel (long K)
struct KProf {
    static constexpr int NK = 4;
    bool on = false;
    struct Rec {
        int kind;
        hipEvent_t a, b;
        double work;
    };
    std::vector<Rec> live;
    std::vector<hipEvent_t> pool;
    long long cnt[NK] = {0};
    double ms[NK] = {0}, work[NK] = {0};
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    hipEvent_t begin(hipStream_t st) {
        if (!on) return nullptr;
        hipEvent_t e = get();
        (void)hipEventRecord(e, st);
        return e;
    }
    void end(hipStream_t st, hipEvent_t a, int kind, double w) {
        if (!on || !a) return;
        hipEvent_t b = get();
        (void)hipEventRecord(b, st);
        live.push_back(Rec{kind, a, b, w});
        if (live.size() > 4096) flush();
    }
    void flush() {
        for (Rec &r : live) {
            float t = 0;
            if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
                ms[r.kind] += t;
                cnt[r.kind] += 1;
                work[r.kind] += r.work;
            }
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        live.clear();
    }
    void reset() {
        flush();
        for (int k = 0; k < NK; k++) {
            cnt[k] = 0;
            ms[k] = work[k] = 0;
        }
    }
    ~KProf() {
        for (Rec &r : live) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

// mp values in LDS planes: word q of element t at sh[q * plane + t], head at sh[L * plane + t]
template <int L>
__device__ __forceinline__ mp<L> lds_ld(const uint32_t *sh, int plane, int t) {
    mp<L> x;
#pragma unroll
    for (int q = 0; q < L; q++) x.m[q] = sh[q * plane + t];
    int32_t h = (int32_t)sh[L * plane + t];
    x.e = h >> 1;
    x.s = (uint32_t)h & 1u;
    return x;
}
template <int L>
__device__ __forceinline__ void lds_st(uint32_t *sh, int plane, int t, const mp<L> &x) {
#pragma unroll
    for (int q = 0; q < L; q++) sh[q * plane + t] = x.m[q];
    sh[L * plane + t] = (uint32_t)((x.e << 1) | (int32_t)(x.s & 1u));
}

template <int L>
__device__ __forceinline__ mp<L> fma_acc(const mp<L> &acc, const mp<L> &a, const mp<L> &b) {
    return mpx::add_z<L>(acc, mpx::mul<L>(a, b));
}

}  // namespace sdpg
