// gemm_plain.h — plain limb form of the multiple-precision GEMM.
//
// 1. Block fixed point (k_fix_rows): each row of op(A) and each column of op(B)
//    gets one exponent e_r (its largest element).  Element x becomes the integer
//    a = round(x * 2^(W - e_r)), W = 32L - 2, stored offset-binary as
//    a' = a + O, O = 2^(32L-1), so every operand is an unsigned L-limb integer.
//    The row sums RA_r = sum_k a'_rk are kept (L+1 limbs).
// 2. MAC kernel (k_gemm_plain): 16x16 outputs per workgroup, K staged through LDS
//    in chunks of 16.  Each thread accumulates sum_k a'_ik b'_kj into 96-bit column
//    accumulators (v_mad_u64_u32 + v_addc per limb product), skipping partial-product
//    columns below L-2.
// 3. Epilogue: S = sum a'b' - O (RA_i + RB_j) + K O^2 = sum_k a_ik b_kj exactly (up to the
//    skipped columns), renormalised to one limb-float with exponent e_i + e_j - 2W,
//    then C = S, C - S or C + S.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine_common.h"
#include "mpx.h"

namespace sdpg {

struct GemmArgs;

// Fixed-point conversion of the "rows" of an operand.
//  which = 0: rows i of op(A) over k;  which = 1: columns j of op(B) over k.
//  out layout per batch instance: limb l, k, r  ->  out[((l * K) + k) * R + r]
template <int L>
__global__ void k_fix_rows(GemmArgs g, int which, int R, int nb, uint32_t *out, int32_t *eo, uint32_t *rsum) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int b = blockIdx.y;
    if (r >= R || b >= nb) return;
    const int K = g.K;
    auto at = [&](int k) -> int64_t {
        if (which == 0)
            return g.offA + (int64_t)b * g.sA + (g.transA ? (k + (int64_t)g.lda * r) : (r + (int64_t)g.lda * k));
        return g.offB + (int64_t)b * g.sB + (g.transB ? (r + (int64_t)g.ldb * k) : (k + (int64_t)g.ldb * r));
    };
    PV src = which == 0 ? g.A : g.B;
    int32_t emax = mpx::EZERO;
    for (int k = 0; k < K; k++) {
        int32_t h = (int32_t)src.p[(int64_t)L * src.n + at(k)];
        uint32_t top = src.p[(int64_t)(L - 1) * src.n + at(k)];
        int32_t e = h >> 1;
        if (top != 0 && e > emax) emax = e;
    }
    if (emax == mpx::EZERO) emax = 0;
    uint32_t rs[L + 1];
#pragma unroll
    for (int q = 0; q <= L; q++) rs[q] = 0;
    int64_t base = (int64_t)b * L * K * R;
    for (int k = 0; k < K; k++) {
        mpx::mp<L> x = ldm<L>(src, at(k));
        uint32_t a[L];
        if (mpx::is_zero<L>(x)) {
#pragma unroll
            for (int q = 0; q < L; q++) a[q] = 0;
        } else {
            uint32_t t[L + 1];
            t[0] = 0;
#pragma unroll
            for (int q = 0; q < L; q++) t[q + 1] = x.m[q];
            uint32_t sticky = 0;
            uint32_t d = 2u + (uint32_t)(emax - x.e);
            mpx::shr<L + 1>(t, d, sticky);
            uint32_t c = t[0] >> 31;  // round to nearest
#pragma unroll
            for (int q = 0; q < L; q++) {
                uint32_t v = t[q + 1] + c;
                c = (c && v == 0) ? 1u : 0u;
                a[q] = v;
            }
            if (x.s) {  // two's complement negate
                uint32_t br = 1;
#pragma unroll
                for (int q = 0; q < L; q++) {
                    uint32_t v = ~a[q] + br;
                    br = (br && v == 0) ? 1u : 0u;
                    a[q] = v;
                }
            }
        }
        a[L - 1] += 0x80000000u;  // offset binary: a' = a + 2^(32L-1)
        uint32_t c = 0;
#pragma unroll
        for (int q = 0; q < L; q++) {
            uint64_t s = (uint64_t)rs[q] + a[q] + c;
            rs[q] = (uint32_t)s;
            c = (uint32_t)(s >> 32);
            out[base + ((int64_t)q * K + k) * R + r] = a[q];
        }
        rs[L] += c;
    }
    eo[(int64_t)b * R + r] = emax;
#pragma unroll
    for (int q = 0; q <= L; q++) rsum[((int64_t)b * (L + 1) + q) * R + r] = rs[q];
}

constexpr int PT = 16;  // output tile edge
constexpr int PK = 16;  // K chunk

template <int L>
__global__ __launch_bounds__(256) void k_gemm_plain(GemmArgs g, const uint32_t *__restrict__ af,
                                                    const uint32_t *__restrict__ bf, const int32_t *__restrict__ ea,
                                                    const int32_t *__restrict__ eb, const uint32_t *__restrict__ ra,
                                                    const uint32_t *__restrict__ rb, int nbA, int nbB) {
    constexpr int S0 = (L >= 2) ? L - 2 : 0;
    constexpr int NC = 2 * L - 1 - S0;  // accumulated columns S0 .. 2L-2
    __shared__ uint32_t As[L][PK][PT];
    __shared__ uint32_t Bs[L][PK][PT];
    const int M = g.M, N = g.N, K = g.K;
    int tx = threadIdx.x % PT, ty = threadIdx.x / PT;
    int i0 = blockIdx.x * PT, j0 = blockIdx.y * PT;
    int b = blockIdx.z;
    if (g.lower && i0 + PT - 1 < j0) return;  // tile strictly above the diagonal
    int bA = nbA > 1 ? b : 0, bB = nbB > 1 ? b : 0;
    const uint32_t *Ab = af + (int64_t)bA * L * K * M;
    const uint32_t *Bb = bf + (int64_t)bB * L * K * N;
    int i = i0 + tx, j = j0 + ty;
    uint64_t lo[NC];
    uint32_t hi[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        lo[c] = 0;
        hi[c] = 0;
    }
    for (int k0 = 0; k0 < K; k0 += PK) {
        // cooperative tile loads: 256 threads, L*PK*PT words per operand
        for (int w = threadIdx.x; w < L * PK * PT; w += 256) {
            int q = w / (PK * PT), rem = w % (PK * PT), kk = rem / PT, rr = rem % PT;
            int k = k0 + kk;
            uint32_t va = 0x80000000u * (q == L - 1), vb = va;  // padding = offset zero
            if (k < K && i0 + rr < M) va = Ab[((int64_t)q * K + k) * M + i0 + rr];
            if (k < K && j0 + rr < N) vb = Bb[((int64_t)q * K + k) * N + j0 + rr];
            As[q][kk][rr] = va;
            Bs[q][kk][rr] = vb;
        }
        __syncthreads();
        int kc = (K - k0) < PK ? (K - k0) : PK;
        for (int kk = 0; kk < kc; kk++) {
            uint32_t a[L], bq[L];
#pragma unroll
            for (int q = 0; q < L; q++) {
                a[q] = As[q][kk][tx];
                bq[q] = Bs[q][kk][ty];
            }
#pragma unroll
            for (int s = S0; s <= 2 * L - 2; s++) {
#pragma unroll
                for (int p = 0; p < L; p++) {
                    int q = s - p;
                    if (q >= 0 && q < L) mpx::mac(lo[s - S0], hi[s - S0], a[p], bq[q]);
                }
            }
        }
        __syncthreads();
    }
    if (i >= M || j >= N) return;
    if (g.lower && i < j) return;
    // ---- epilogue: limbs S0 .. 2L+1 of  T + K*O^2 - O*(RA_i + RB_j) ----
    constexpr int NB = 2 * L + 2 - S0;
    int64_t buf[NB];  // signed limb accumulators (value < 2^34 in magnitude per limb before carry)
    {
        uint64_t clo = 0;
        uint32_t chi = 0;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            // (chi:clo) += (hi[c]:lo[c]) ; emit low 32 bits; shift right 32
            uint64_t s = clo + lo[c];
            uint32_t cy = s < clo;
            uint32_t h2 = chi + hi[c] + cy;
            buf[c] = (int64_t)(uint32_t)s;
            clo = (s >> 32) | ((uint64_t)h2 << 32);
            chi = 0;
        }
        buf[NC] = (int64_t)(uint32_t)clo;
        buf[NC + 1] = (int64_t)(uint32_t)(clo >> 32);
#pragma unroll
        for (int c = NC + 2; c < NB; c++) buf[c] = 0;
    }
    // + K * 2^(64L-2): K << 30 at limb 2L-1
    {
        uint64_t kk = (uint64_t)K << 30;
        buf[2 * L - 1 - S0] += (int64_t)(uint32_t)kk;
        buf[2 * L - S0] += (int64_t)(uint32_t)(kk >> 32);
    }
    // - (RA_i + RB_j) * 2^(32L-1)
    {
        uint32_t u[L + 2];
        uint32_t c = 0;
        int bAi = nbA > 1 ? b : 0, bBj = nbB > 1 ? b : 0;
#pragma unroll
        for (int q = 0; q <= L; q++) {
            uint64_t s = (uint64_t)ra[((int64_t)bAi * (L + 1) + q) * M + i] + rb[((int64_t)bBj * (L + 1) + q) * N + j] + c;
            u[q] = (uint32_t)s;
            c = (uint32_t)(s >> 32);
        }
        u[L + 1] = c;
        // shift left by 31 bits: w[t] = (u[t] << 31) | (u[t-1] >> 1), t = 0..L+2
#pragma unroll
        for (int t = 0; t <= L + 2; t++) {
            uint32_t cur = t <= L + 1 ? u[t] : 0u;
            uint32_t prev = t >= 1 ? u[t - 1] : 0u;
            uint32_t w = (cur << 31) | (prev >> 1);
            int pos = L - 1 + t - S0;
            if (pos >= 0 && pos < NB) buf[pos] -= (int64_t)w;
        }
    }
    // carry-normalise to 32-bit limbs (two's complement, top limb carries the sign)
    uint32_t x[NB];
    {
        int64_t carry = 0;
#pragma unroll
        for (int c = 0; c < NB; c++) {
            int64_t v = buf[c] + carry;
            x[c] = (uint32_t)v;
            carry = v >> 32;  // arithmetic shift
        }
    }
    uint32_t neg = x[NB - 1] >> 31;
    if (neg) {
        uint32_t br = 1;
#pragma unroll
        for (int c = 0; c < NB; c++) {
            uint32_t v = ~x[c] + br;
            br = (br && v == 0) ? 1u : 0u;
            x[c] = v;
        }
    }
    mpx::mp<L> res;
    uint32_t lz = mpx::lead_zeros<NB>(x);
    if (lz >= 32u * NB) {
        res = mpx::zero<L>();
    } else {
        mpx::shl<NB>(x, lz);
        uint32_t y[L + 1];
#pragma unroll
        for (int q = 0; q <= L; q++) y[q] = x[NB - 1 - L + q];
        int bAi = nbA > 1 ? b : 0, bBj = nbB > 1 ? b : 0;
        int32_t E = ea[(int64_t)bAi * M + i] + eb[(int64_t)bBj * N + j] - 2 * (32 * L - 2);
        int32_t e = 32 * (2 * L + 2) - (int32_t)lz + E;
        mpx::round_store<L>(res, y, e, neg);
    }
    int64_t cidx = g.offC + (int64_t)b * g.sC + i + (int64_t)g.ldc * j;
    if (g.op == 1)
        res = mpx::sub<L>(ldm<L>(g.C, cidx), res);
    else if (g.op == 2)
        res = mpx::add<L>(ldm<L>(g.C, cidx), res);
    stm<L>(g.C, cidx, res);
}

template <int L>
struct PlainWork {
    uint32_t *af = nullptr, *bf = nullptr, *ra = nullptr, *rb = nullptr;
    int32_t *ea = nullptr, *eb = nullptr;
    size_t cap_af = 0, cap_bf = 0, cap_ra = 0, cap_rb = 0, cap_ea = 0, cap_eb = 0;
    ~PlainWork() {
        hipFree(af);
        hipFree(bf);
        hipFree(ra);
        hipFree(rb);
        hipFree(ea);
        hipFree(eb);
    }
    template <class T>
    static int grow(T **p, size_t &cap, size_t need) {
        if (need <= cap) return 0;
        if (*p) hipFree(*p);
        *p = nullptr;
        size_t n = need + need / 4 + 64;
        if (hipMalloc((void **)p, sizeof(T) * n) != hipSuccess) {
            cap = 0;
            return -4;
        }
        cap = n;
        return 0;
    }
    int run(const GemmArgs &g, hipStream_t st) {
        int nbA = g.sA ? g.batch : 1, nbB = g.sB ? g.batch : 1;
        size_t KK = (size_t)(g.K > 0 ? g.K : 1);
        int rc = 0;
        rc |= grow(&af, cap_af, (size_t)nbA * L * KK * g.M);
        rc |= grow(&bf, cap_bf, (size_t)nbB * L * KK * g.N);
        rc |= grow(&ra, cap_ra, (size_t)nbA * (L + 1) * g.M);
        rc |= grow(&rb, cap_rb, (size_t)nbB * (L + 1) * g.N);
        rc |= grow(&ea, cap_ea, (size_t)nbA * g.M);
        rc |= grow(&eb, cap_eb, (size_t)nbB * g.N);
        if (rc) return -4;
        k_fix_rows<L><<<dim3((g.M + 63) / 64, nbA), 64, 0, st>>>(g, 0, g.M, nbA, af, ea, ra);
        k_fix_rows<L><<<dim3((g.N + 63) / 64, nbB), 64, 0, st>>>(g, 1, g.N, nbB, bf, eb, rb);
        dim3 grd((g.M + PT - 1) / PT, (g.N + PT - 1) / PT, g.batch);
        k_gemm_plain<L><<<grd, 256, 0, st>>>(g, af, bf, ea, eb, ra, rb, nbA, nbB);
        return hipGetLastError() == hipSuccess ? 0 : -2;
    }
};

}  // namespace sdpg
