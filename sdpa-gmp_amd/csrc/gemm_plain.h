// gemm_plain.h — plain limb form of the multiple-precision GEMM.
//
// 1. Block fixed point (k_fix_rows): each row of op(A) and each column of op(B)
//    gets one exponent e_r (its largest element).  Element x becomes the integer
//    a = round(x * 2^(W - e_r)), W = 32L - 2, stored offset-binary as
//    a' = a + O, O = 2^(32L-1), so every operand is an unsigned L-limb integer.
//    The row sums RA_r = sum_k a'_rk are kept (L+1 limbs).
// 2. MAC kernel (k_gemm_plain): 16x16 outputs per workgroup, K staged through LDS
//    in chunks of 16.  Each thread accumulates sum_k a'_ik b'_kj into 96-bit column
//    accumulators (v_mad_u64_u32 + v_addc per limb product, independent columns issued in
//    pairs), skipping partial-product columns below L-2.
// 3. Epilogue: S = sum a'b' - O (RA_i + RB_j) + K O^2 = sum_k a_ik b_kj exactly (up to the
//    skipped columns), renormalised to one limb-float with exponent e_i + e_j - 2W,
//    then C = S, C - S or C + S.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "engine_common.h"
#include "mpx.h"

namespace sdpg {

struct GemmArgs;

// Offset-binary block fixed point: finish sum_k a_k b_k from the column accumulators of
// sum_k a'_k b'_k (columns S0 = L-2 .. 2L-2, 96 bits each), the offset sums RA = sum a',
// RB = sum b' (L+1 limbs each) and the count K:  S = T - O (RA + RB) + K O^2, O = 2^(32L-1);
// returns S * 2^E as a limb float (round to nearest).
template <int L>
__device__ __forceinline__ mpx::mp<L> fx_finish(const uint64_t (&lo)[2 * L - 1 - ((L >= 2) ? L - 2 : 0)],
                                                const uint32_t (&hi)[2 * L - 1 - ((L >= 2) ? L - 2 : 0)], int K,
                                                const uint32_t (&RA)[L + 1], const uint32_t (&RB)[L + 1], int32_t E) {
    constexpr int S0 = (L >= 2) ? L - 2 : 0;
    constexpr int NC = 2 * L - 1 - S0;
    constexpr int NB = 2 * L + 2 - S0;
    int64_t buf[NB];
    {
        uint64_t clo = 0;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            uint64_t s = clo + lo[c];
            uint32_t cy = s < clo;
            uint32_t h2 = hi[c] + cy;
            buf[c] = (int64_t)(uint32_t)s;
            clo = (s >> 32) | ((uint64_t)h2 << 32);
        }
        buf[NC] = (int64_t)(uint32_t)clo;
        buf[NC + 1] = (int64_t)(uint32_t)(clo >> 32);
#pragma unroll
        for (int c = NC + 2; c < NB; c++) buf[c] = 0;
    }
    {  // + K * 2^(64L-2): K << 30 at limb 2L-1
        uint64_t kk = (uint64_t)(uint32_t)K << 30;
        buf[2 * L - 1 - S0] += (int64_t)(uint32_t)kk;
        buf[2 * L - S0] += (int64_t)(uint32_t)(kk >> 32);
    }
    {  // - (RA + RB) * 2^(32L-1)
        uint32_t u[L + 2];
        uint32_t c = 0;
#pragma unroll
        for (int q = 0; q <= L; q++) {
            uint64_t s = (uint64_t)RA[q] + RB[q] + c;
            u[q] = (uint32_t)s;
            c = (uint32_t)(s >> 32);
        }
        u[L + 1] = c;
#pragma unroll
        for (int t = 0; t <= L + 2; t++) {
            uint32_t cur = t <= L + 1 ? u[t] : 0u;
            uint32_t prev = t >= 1 ? u[t - 1] : 0u;
            uint32_t w = (cur << 31) | (prev >> 1);
            int pos = L - 1 + t - S0;
            if (pos >= 0 && pos < NB) buf[pos] -= (int64_t)w;
        }
    }
    uint32_t x[NB];
    {
        int64_t carry = 0;
#pragma unroll
        for (int c = 0; c < NB; c++) {
            int64_t v = buf[c] + carry;
            x[c] = (uint32_t)v;
            carry = v >> 32;
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
    if (lz >= 32u * NB) return mpx::zero<L>();
    mpx::shl<NB>(x, lz);
    uint32_t y[L + 1];
#pragma unroll
    for (int q = 0; q <= L; q++) y[q] = x[NB - 1 - L + q];
    int32_t e = 32 * (2 * L + 2) - (int32_t)lz + E;
    mpx::round_store<L>(res, y, e, neg);
    return res;
}

// x -> offset-binary fixed a' = round(x * 2^(W - emax)) + 2^(32L-1), with W = 32L - head
// (|x| < 2^emax); two's complement before the offset.  Returns the L limbs in a.
template <int L>
__device__ __forceinline__ void fx_convert(const mpx::mp<L> &x, int32_t emax, int head, uint32_t (&a)[L],
                                           bool offset = true) {
    if (mpx::is_zero<L>(x)) {
#pragma unroll
        for (int q = 0; q < L; q++) a[q] = 0;
    } else {
        uint32_t t[L + 1];
        t[0] = 0;
#pragma unroll
        for (int q = 0; q < L; q++) t[q + 1] = x.m[q];
        uint32_t sticky = 0;
        mpx::shr<L + 1>(t, (uint32_t)head + (uint32_t)(emax - x.e), sticky);
        uint32_t c = t[0] >> 31;
#pragma unroll
        for (int q = 0; q < L; q++) {
            uint32_t v = t[q + 1] + c;
            c = (c && v == 0) ? 1u : 0u;
            a[q] = v;
        }
        if (x.s) {
            uint32_t br = 1;
#pragma unroll
            for (int q = 0; q < L; q++) {
                uint32_t v = ~a[q] + br;
                br = (br && v == 0) ? 1u : 0u;
                a[q] = v;
            }
        }
    }
    if (offset) a[L - 1] += 0x80000000u;
}

// Element (r, k) of the operand seen as R "rows" over K: which = 0 -> op(A)[r, k],
// which = 1 -> op(B)[k, r].  contig() tells whether r (true) or k is the memory-contiguous index,
// so that consecutive threads walk consecutive addresses.

// pass 1: per-row largest exponent e_r (EZERO for an all-zero row: a zero's head holds EZERO),
// written once per row.  Only the head word is read.  Rows contiguous in memory: a block covers
// 256 / KS rows with KS threads per row striding over k (coalesced across the rows), reduced in
// LDS; k contiguous: a wave per row.
constexpr int FXM_KS = 8;
// Both operands in one launch: blockIdx.z = 1 selects op(B) with R2 rows and eo2 (same grid shape).
template <int L>
__global__ __launch_bounds__(256) void k_fix_max(GemmArgs g, int which, int R, int32_t *eo, int R2 = 0,
                                                 int32_t *eo2 = nullptr) {
    if (blockIdx.z) {
        which = 1;
        R = R2;
        eo = eo2;
    }
    int b = blockIdx.y;
    PV src = which == 0 ? g.A : g.B;
    const int K = g.K;
    const uint32_t *head = src.p + (int64_t)L * src.n;
    if (fx_rcontig(g, which)) {
        __shared__ int32_t red[256];
        constexpr int RPB = 256 / FXM_KS;
        int rl = threadIdx.x % RPB, ks = threadIdx.x / RPB;
        int r = blockIdx.x * RPB + rl;
        int32_t e = mpx::EZERO;
        if (r < R) {
#pragma unroll 4
            for (int k = ks; k < K; k += FXM_KS) {
                int32_t x = ((int32_t)head[fx_at(g, which, b, r, k)]) >> 1;
                e = x > e ? x : e;
            }
        }
        red[threadIdx.x] = e;
        __syncthreads();
        if (ks == 0 && r < R) {
#pragma unroll
            for (int q = 1; q < FXM_KS; q++) {
                int32_t x = red[rl + q * RPB];
                e = x > e ? x : e;
            }
            eo[(int64_t)b * R + r] = e;
        }
    } else {
        int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        int lane = threadIdx.x & 63;
        int32_t e = mpx::EZERO;
        if (r < R)
#pragma unroll 8
            for (int k = lane; k < K; k += 64) {
                int32_t x = ((int32_t)head[fx_at(g, which, b, r, k)]) >> 1;
                e = x > e ? x : e;
            }
        for (int off = 32; off >= 1; off >>= 1) {
            int32_t o = __shfl_xor(e, off, 64);
            e = o > e ? o : e;
        }
        if (r < R && lane == 0) eo[(int64_t)b * R + r] = e;
    }
}

// grid of k_fix_max for R rows
inline dim3 fix_max_grid(const GemmArgs &g, int which, int R, int nb) {
    bool rc = which == 0 ? !g.transA : g.transB;
    return rc ? dim3((unsigned)((R + 256 / FXM_KS - 1) / (256 / FXM_KS)), nb) : dim3((unsigned)((R + 3) / 4), nb);
}

// op(B) is op(A)^T on the same data (the Cholesky trailing update P P^T): its columns are
// op(A)'s rows, so their exponents and fixed-point / slice forms are the same
inline bool gemm_b_is_at(const GemmArgs &g) {
    return g.B.p == g.A.p && g.offB == g.offA && g.ldb == g.lda && g.transB != g.transA && g.N == g.M &&
           g.sA == g.sB;
}

// pass 2: a' = round(x * 2^(W - e_r)) + 2^(32L-1), W = 32L - 2; out[((q * K) + k) * R + r] per batch
template <int L>
__global__ void k_fix_conv(GemmArgs g, int which, int R, const int32_t *eo, uint32_t *out, int R2 = 0,
                           const int32_t *eo2 = nullptr, uint32_t *out2 = nullptr) {
    if (blockIdx.z) {  // both operands in one launch (op(B) in z = 1)
        which = 1;
        R = R2;
        eo = eo2;
        out = out2;
    }
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int b = blockIdx.y;
    int K = g.K;
    if (t >= (int64_t)R * K) return;
    bool rc = fx_rcontig(g, which);
    int r = rc ? (int)(t % R) : (int)(t / K);
    int k = rc ? (int)(t / R) : (int)(t % K);
    PV src = which == 0 ? g.A : g.B;
    int32_t emax = eo[(int64_t)b * R + r];
    if (emax == mpx::EZERO) emax = 0;
    uint32_t a[L];
    fx_convert<L>(ldm<L>(src, fx_at(g, which, b, r, k)), emax, 2, a);
    int64_t base = (int64_t)b * L * K * R;
#pragma unroll
    for (int q = 0; q < L; q++) out[base + ((int64_t)q * K + k) * R + r] = a[q];
}

static __global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

constexpr int PT = 16;  // output tile edge
constexpr int PK = 16;  // K chunk

// The offset sums RA_i = sum_k a'_ik and RB_j = sum_k b'_kj are accumulated from the LDS tiles by
// the threads of tile rows ty = 0 (A rows) and ty = 1 (B columns).
template <int L>
__global__ __launch_bounds__(256) void k_gemm_plain(GemmArgs g, const uint32_t *__restrict__ af,
                                                    const uint32_t *__restrict__ bf, const int32_t *__restrict__ ea,
                                                    const int32_t *__restrict__ eb, int nbA, int nbB) {
    constexpr int S0 = (L >= 2) ? L - 2 : 0;
    constexpr int NC = 2 * L - 1 - S0;
    __shared__ uint32_t As[L][PK][PT];
    __shared__ uint32_t Bs[L][PK][PT];
    __shared__ uint32_t Ssum[2][PT][L + 1];
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
    uint32_t rsum[L + 1];
#pragma unroll
    for (int q = 0; q <= L; q++) rsum[q] = 0;
    for (int k0 = 0; k0 < K; k0 += PK) {
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
        if (ty < 2) {
            for (int kk = 0; kk < kc; kk++) {
                uint32_t c = 0;
#pragma unroll
                for (int q = 0; q < L; q++) {
                    uint64_t sm = (uint64_t)rsum[q] + (ty == 0 ? As[q][kk][tx] : Bs[q][kk][tx]) + c;
                    rsum[q] = (uint32_t)sm;
                    c = (uint32_t)(sm >> 32);
                }
                rsum[L] += c;
            }
        }
        for (int kk = 0; kk < kc; kk++) {
            uint32_t a[L], bq[L];
#pragma unroll
            for (int q = 0; q < L; q++) {
                a[q] = As[q][kk][tx];
                bq[q] = Bs[q][kk][ty];
            }
            // column accumulators are independent: issue the multiply-adds row by row in pairs on
            // adjacent columns (mpx::col_products), not column by column as a dependent chain
            mpx::col_products<L, S0>(a, bq, lo, hi);
        }
        __syncthreads();
    }
    if (ty < 2) {
#pragma unroll
        for (int q = 0; q <= L; q++) Ssum[ty][tx][q] = rsum[q];
    }
    __syncthreads();
    if (i >= M || j >= N) return;
    if (g.lower && i < j) return;
    uint32_t ras[L + 1], rbs[L + 1];
#pragma unroll
    for (int q = 0; q <= L; q++) {
        ras[q] = Ssum[0][tx][q];
        rbs[q] = Ssum[1][ty][q];
    }
    int32_t E = ea[(int64_t)bA * M + i] + eb[(int64_t)bB * N + j];
    E = (E < -(1 << 28) ? 0 : E) - 2 * (32 * L - 2);
    mpx::mp<L> res = fx_finish<L>(lo, hi, K, ras, rbs, E);
    int64_t cidx = g.offC + (int64_t)b * g.sC + i + (int64_t)g.ldc * j;
    if (g.op == 1)
        res = mpx::sub<L>(ldm<L>(g.C, cidx), res);
    else if (g.op == 2)
        res = mpx::add<L>(ldm<L>(g.C, cidx), res);
    stm<L>(g.C, cidx, res);
}

// outputs a GEMM call computes (the lower triangle only when g.lower), times the batch
inline double gemm_outputs(const GemmArgs &g) {
    double o = (double)g.M * g.N;
    if (g.lower) {
        double mn = g.M < g.N ? g.M : g.N;
        o = mn * (mn + 1) / 2 + (g.M > g.N ? (double)(g.M - g.N) * g.N : 0.0);
    }
    return o * g.batch;
}

template <int L>
struct PlainWork {
    KProf *prof = nullptr;
    uint32_t *af = nullptr, *bf = nullptr;
    int32_t *ea = nullptr, *eb = nullptr;
    size_t cap_af = 0, cap_bf = 0, cap_ea = 0, cap_eb = 0;
    Arena *arena = nullptr;  // when set, the workspace is carved from it (no hipMalloc / hipFree)
    ~PlainWork() {
        if (arena) return;
        hipFree(af);
        hipFree(bf);
        hipFree(ea);
        hipFree(eb);
    }
    template <class T>
    int grow(T **p, size_t &cap, size_t need) {
        if (need <= cap) return 0;
        if (arena) {
            *p = (T *)arena->take(sizeof(T) * (need ? need : 1));
            cap = *p ? need : 0;
            return *p ? 0 : -4;
        }
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
    OpCache<uint32_t> kept;  // fixed-point forms and exponents of operands that never change
    int run(const GemmArgs &g, hipStream_t st, int keep = 0) {
        int nbA = g.sA ? g.batch : 1, nbB = g.sB ? g.batch : 1;
        size_t KK = (size_t)(g.K > 0 ? g.K : 1);
        const bool same = gemm_b_is_at(g);
        const bool keepA = (keep & GEMM_KEEP_A) != 0, keepB = (keep & GEMM_KEEP_B) != 0 && !same;
        const size_t need_a = (size_t)nbA * L * KK * g.M, need_b = (size_t)nbB * L * KK * g.N;
        if ((!keepA && (need_a > cap_af || (size_t)nbA * g.M > cap_ea)) ||
            (!keepB && (need_b > cap_bf || (size_t)nbB * g.N > cap_eb)))
            if (hipStreamSynchronize(st) != hipSuccess) return -2;  // queued kernels may still read the old buffers
        uint32_t *fa = af, *fb = bf;
        int32_t *qa = ea, *qb = eb;
        bool fillA = true, fillB = !same;
        if (keepA) {
            auto *e = kept.get(op_key(g, 0, 0), need_a, (size_t)nbA * g.M, fillA);
            if (!e) return -4;
            fa = e->data;
            qa = e->exps;
        } else {
            if (grow(&af, cap_af, need_a) | grow(&ea, cap_ea, (size_t)nbA * g.M)) return -4;
            fa = af;
            qa = ea;
        }
        if (keepB) {
            auto *e = kept.get(op_key(g, 1, 0), need_b, (size_t)nbB * g.N, fillB);
            if (!e) return -4;
            fb = e->data;
            qb = e->exps;
        } else if (!same) {
            if (grow(&bf, cap_bf, need_b) | grow(&eb, cap_eb, (size_t)nbB * g.N)) return -4;
            fb = bf;
            qb = eb;
        }
        int64_t ta = (int64_t)g.M * g.K, tb = (int64_t)g.N * g.K;
        dim3 ga((unsigned)((ta + 255) / 256), nbA), gb((unsigned)((tb + 255) / 256), nbB);
        if (fillA && fillB && nbA == nbB && fx_rcontig_host(g, 0) == fx_rcontig_host(g, 1)) {
            // both operands in one launch per pass
            dim3 gm = fix_max_grid(g, 0, std::max(g.M, g.N), nbA);
            k_fix_max<L><<<dim3(gm.x, gm.y, 2), 256, 0, st>>>(g, 0, g.M, qa, g.N, qb);
            k_fix_conv<L><<<dim3(std::max(ga.x, gb.x), nbA, 2), 256, 0, st>>>(g, 0, g.M, qa, fa, g.N, qb, fb);
            fillA = fillB = false;
        }
        if (fillA) {
            k_fix_max<L><<<fix_max_grid(g, 0, g.M, nbA), 256, 0, st>>>(g, 0, g.M, qa);
            k_fix_conv<L><<<ga, 256, 0, st>>>(g, 0, g.M, qa, fa);
        }
        if (fillB) {
            k_fix_max<L><<<fix_max_grid(g, 1, g.N, nbB), 256, 0, st>>>(g, 1, g.N, qb);
            k_fix_conv<L><<<gb, 256, 0, st>>>(g, 1, g.N, qb, fb);
        }
        if (same) {
            fb = fa;
            qb = qa;
        }
        dim3 grd((g.M + PT - 1) / PT, (g.N + PT - 1) / PT, g.batch);
        hipEvent_t ev = prof ? prof->begin(st) : nullptr;
        k_gemm_plain<L><<<grd, 256, 0, st>>>(g, fa, fb, qa, qb, nbA, same ? nbA : nbB);
        if (prof) prof->end(st, ev, 1, gemm_outputs(g) * (double)g.K);
        return hipGetLastError() == hipSuccess ? 0 : -2;
    }
};

}  // namespace sdpg
