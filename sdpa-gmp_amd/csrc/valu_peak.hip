// valu_peak.hip — measures the sustained rate of the plain form's inner operation, one
// 32x32->64-bit limb multiply-accumulate with carry (v_mad_u64_u32 + v_addc_co_u32), on
// every CU; the result is the "integer-VALU peak" the plain GEMM's utilisation is quoted against.
//   valu_peak [iterations]   -> prints {"limb_mac_per_s": ..., "seconds": ...}
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "mpx.h"

__global__ __launch_bounds__(256) void k_peak(uint32_t *out, int iters, uint32_t seed) {
    uint32_t a0 = seed ^ threadIdx.x, a1 = a0 * 3u + 1u, a2 = a0 * 5u + 7u, a3 = a0 * 7u + 3u;
    uint32_t b0 = a0 + 11u, b1 = a1 + 13u;
    uint64_t l0 = 0, l1 = 0, l2 = 0, l3 = 0, l4 = 0, l5 = 0, l6 = 0, l7 = 0;
    uint32_t h0 = 0, h1 = 0, h2 = 0, h3 = 0, h4 = 0, h5 = 0, h6 = 0, h7 = 0;
    for (int i = 0; i < iters; i++) {
        mpx::mac4(l0, h0, a0, b0, l1, h1, a1, b0, l2, h2, a2, b0, l3, h3, a3, b0);
        mpx::mac4(l4, h4, a0, b1, l5, h5, a1, b1, l6, h6, a2, b1, l7, h7, a3, b1);
    }
    uint64_t s = l0 ^ l1 ^ l2 ^ l3 ^ l4 ^ l5 ^ l6 ^ l7;
    uint32_t hs = h0 + h1 + h2 + h3 + h4 + h5 + h6 + h7;
    out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)s ^ (uint32_t)(s >> 32) ^ hs;
}

int main(int argc, char **argv) {
    int iters = argc > 1 ? atoi(argv[1]) : 20000;
    int dev = 0, cus = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        fprintf(stderr, "no device\n");
        return 1;
    }
    cus = prop.multiProcessorCount;
    int blocks = cus * 8, threads = 256;
    uint32_t *out = nullptr;
    if (hipMalloc(&out, sizeof(uint32_t) * blocks * threads) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    k_peak<<<blocks, threads>>>(out, 100, 1u);
    hipDeviceSynchronize();
    hipEventRecord(e0, 0);
    k_peak<<<blocks, threads>>>(out, iters, 2u);
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    double macs = 8.0 * iters * (double)blocks * threads;
    printf("{\"limb_mac_per_s\": %.6e, \"seconds\": %.6f, \"cus\": %d, \"clock_khz\": %d}\n", macs / (ms * 1e-3),
           ms * 1e-3, cus, prop.clockRate);
    hipFree(out);
    return 0;
}
