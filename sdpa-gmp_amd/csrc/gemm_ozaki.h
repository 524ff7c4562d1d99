// gemm_ozaki.h — Ozaki (INT8 slice, MFMA) form of the multiple-precision GEMM.
//
// 1. Block fixed point as in the plain form: row r of op(A) gets the exponent e_r of its
//    largest element and x becomes the integer Y = round(x * 2^(8S - 2 - e_r)),
//    |Y| < 2^(8S-2); the same for the columns of op(B).
// 2. Slicing (k_oz_slice): Y = sum_{p<S} y_p 256^(S-1-p) with balanced digits
//    y_p in [-128, 127] (y_0 in [-64, 64]).  The digits are stored as int8 planes
//    s[chunk][p][row][16]: 16 consecutive k of one row for one slice.
// 3. Products (k_gemm_oz): for the output digit position d = p + q the slice products
//    T_d = sum_{p+q=d} sum_k a_p[i,k] b_q[k,j] are exact INT32 sums; every
//    v_mfma_i32_16x16x64_i8 issues four (p, q) pairs at once (lane group g = lane / 16 carries
//    pair 4u + g, each over the same 16 k of a K chunk).  Pairs with
//    p + q >= S lie below 2^-(8S-8) of the row/column scale and are not formed
//    (S(S+1)/2 slice products instead of S^2).
// 4. Recombination: sum_d T_d 256^-d is carried digit by digit from d = S-1 up to 0
//    into an exact fixed-point accumulator (S/4 fraction words + a 64-bit integer part)
//    held in registers across K chunks, then rounded once to a limb float with
//    exponent e_i + e_j - 12 and applied to C (set / subtract / add), all in full width.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine_common.h"
#include "mpx.h"

namespace sdpg {

typedef int oz_v4i __attribute__((ext_vector_type(4)));
typedef int oz_v16i __attribute__((ext_vector_type(16)));

constexpr int OZ_T = 32;   // output tile edge of one workgroup (2 x 2 waves of 16 x 16)
constexpr int OZ_KC = 16;  // k per chunk (one 16-byte operand per lane and slice)

// slice planes: ((((b * nch + ch) * S + p) * Rp + r) * 16 + kk), Rp = R rounded up to OZ_T
template <int L, int S>
__global__ void k_oz_slice(GemmArgs g, int which, int R, int Rp, int nch, const int32_t *eo, int8_t *out) {
    static_assert(S % 4 == 0 && S <= 4 * L, "slice count");
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int b = blockIdx.y;
    const int Kp = nch * OZ_KC;
    if (t >= (int64_t)Rp * Kp) return;
    int kk = (int)(t % OZ_KC);
    int64_t rest = t / OZ_KC;
    int r = (int)(rest % Rp);
    int ch = (int)(rest / Rp);
    int k = ch * OZ_KC + kk;
    int8_t *o = out + ((((int64_t)b * nch + ch) * S) * Rp + r) * OZ_KC + kk;
    const int64_t pstride = (int64_t)Rp * OZ_KC;
    if (r >= R || k >= g.K) {
#pragma unroll
        for (int p = 0; p < S; p++) o[p * pstride] = 0;
        return;
    }
    PV src = which == 0 ? g.A : g.B;
    int32_t emax = eo[(int64_t)b * R + r];
    if (emax == mpx::EZERO) emax = 0;
    uint32_t a[L];
    // Y = round(x * 2^(8S - 2 - emax)) in two's complement over L limbs
    fx_convert<L>(ldm<L>(src, fx_at(g, which, b, r, k)), emax, 32 * L - 8 * S + 2, a, false);
    uint32_t carry = 0;
#pragma unroll
    for (int u = 0; u < S; u++) {  // u = weight 256^u, digit index p = S-1-u
        int v = (int)((a[u >> 2] >> (8 * (u & 3))) & 255u) + (int)carry;
        int dg = v;
        carry = 0;
        if (v >= 128) {
            dg = v - 256;
            carry = 1;
        }
        o[(S - 1 - u) * pstride] = (int8_t)dg;
    }
}

// one output digit position d of the Horner recombination (compile-time d, S-1 down to 0)
template <int S, int D>
struct OzDigit {
    static __device__ __forceinline__ void run(const oz_v4i *__restrict__ lA, const oz_v4i *__restrict__ lB, int ar,
                                               int bc, int g, uint32_t (&R)[4][S / 4], int64_t (&c)[4]) {
        oz_v4i acc = {0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < (D + 4) / 4; u++) {
            int p = 4 * u + g;
            int pc = p <= D ? p : D;
            oz_v4i av = lA[pc * OZ_T + ar];
            oz_v4i bv = lB[(D - pc) * OZ_T + bc];
            if (p > D) {
                av = oz_v4i{0, 0, 0, 0};
                bv = oz_v4i{0, 0, 0, 0};
            }
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bv, acc, 0, 0, 0);
        }
        constexpr int w = D / 4, sh = 8 * (3 - D % 4);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int64_t v = (int64_t)acc[e] + c[e] + (int64_t)((R[e][w] >> sh) & 255u);
            R[e][w] = (R[e][w] & ~(255u << sh)) | (((uint32_t)v & 255u) << sh);
            c[e] = v >> 8;
        }
        OzDigit<S, D - 1>::run(lA, lB, ar, bc, g, R, c);
    }
};
template <int S>
struct OzDigit<S, -1> {
    static __device__ __forceinline__ void run(const oz_v4i *, const oz_v4i *, int, int, int, uint32_t (&)[4][S / 4],
                                               int64_t (&)[4]) {}
};

template <int L, int S>
__global__ __launch_bounds__(256) void k_gemm_oz(GemmArgs g, const int8_t *__restrict__ as,
                                                 const int8_t *__restrict__ bs, const int32_t *__restrict__ ea,
                                                 const int32_t *__restrict__ eb, int nbA, int nbB, int Mp, int Np,
                                                 int nch) {
    constexpr int NW = S / 4;
    static_assert(NW + 2 >= L + 1, "accumulator narrower than the result");
    __shared__ oz_v4i lA[S * OZ_T], lB[S * OZ_T];
    const int M = g.M, N = g.N;
    const int i0 = blockIdx.x * OZ_T, j0 = blockIdx.y * OZ_T, b = blockIdx.z;
    if (g.lower && i0 + OZ_T - 1 < j0) return;  // tile strictly above the diagonal
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv & 1, wc = wv >> 1, grp = lane >> 4;
    const int ar = wr * 16 + (lane & 15), bc = wc * 16 + (lane & 15);
    const int bA = nbA > 1 ? b : 0, bB = nbB > 1 ? b : 0;
    uint32_t R[4][NW];
    int64_t RI[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        RI[e] = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) R[e][w] = 0;
    }
    const oz_v4i *sa = (const oz_v4i *)as, *sb = (const oz_v4i *)bs;
    for (int ch = 0; ch < nch; ch++) {
        const oz_v4i *pa = sa + (((int64_t)bA * nch + ch) * S) * Mp + i0;
        const oz_v4i *pb = sb + (((int64_t)bB * nch + ch) * S) * Np + j0;
        for (int t = tid; t < S * OZ_T; t += 256) {
            int p = t / OZ_T, r = t % OZ_T;
            lA[t] = pa[(int64_t)p * Mp + r];
            lB[t] = pb[(int64_t)p * Np + r];
        }
        __syncthreads();
        int64_t c[4];
#pragma unroll
        for (int e = 0; e < 4; e++) c[e] = 0;
        OzDigit<S, S - 1>::run(lA, lB, ar, bc, grp, R, c);
#pragma unroll
        for (int e = 0; e < 4; e++) RI[e] += c[e];
        __syncthreads();
    }
    int32_t Ebase = -12;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        int i = i0 + wr * 16 + 4 * grp + e;
        int j = j0 + wc * 16 + (lane & 15);
        if (i >= M || j >= N) continue;
        if (g.lower && i < j) continue;
        // Z = RI * 2^(32 NW) + sum_w R[w] 2^(32 (NW-1-w)), two's complement over NW + 2 words
        constexpr int NX = NW + 2;
        uint32_t x[NX];
#pragma unroll
        for (int w = 0; w < NW; w++) x[w] = R[e][NW - 1 - w];
        x[NW] = (uint32_t)RI[e];
        x[NW + 1] = (uint32_t)((uint64_t)RI[e] >> 32);
        uint32_t neg = x[NX - 1] >> 31;
        if (neg) {
            uint32_t br = 1;
#pragma unroll
            for (int q = 0; q < NX; q++) {
                uint32_t v = ~x[q] + br;
                br = (br && v == 0) ? 1u : 0u;
                x[q] = v;
            }
        }
        mpx::mp<L> res = mpx::zero<L>();
        uint32_t lz = mpx::lead_zeros<NX>(x);
        if (lz < 32u * NX) {
            mpx::shl<NX>(x, lz);
            uint32_t y[L + 1];
#pragma unroll
            for (int q = 0; q <= L; q++) y[q] = x[NX - 1 - L + q];
            int32_t E = ea[(int64_t)bA * M + i] + eb[(int64_t)bB * N + j];
            E = (E < -(1 << 28) ? 0 : E) + Ebase - 8 * (S - 1);
            int32_t ex = 32 * NX - (int32_t)lz + E;
            mpx::round_store<L>(res, y, ex, neg);
        }
        int64_t cidx = g.offC + (int64_t)b * g.sC + i + (int64_t)g.ldc * j;
        if (g.op == 1)
            res = mpx::sub<L>(ldm<L>(g.C, cidx), res);
        else if (g.op == 2)
            res = mpx::add<L>(ldm<L>(g.C, cidx), res);
        stm<L>(g.C, cidx, res);
    }
}

// slice counts instantiated per limb count: the default keeps 8S = 32L - 32 fraction bits of every
// row / column (288 for 320-bit limbs), the wide one 8S = 32L.  The register-resident accumulator
// (16 outputs x S/4 words per lane) bounds the form to L <= OZ_LMAX.
constexpr int OZ_LMAX = 12;
template <int L>
struct OzSlices {
    static constexpr int DEF = 4 * L - 4;
    static constexpr int WIDE = 4 * L;
};

template <int L>
struct OzakiWork {
    KProf *prof = nullptr;
    int8_t *as = nullptr, *bs = nullptr;
    int32_t *ea = nullptr, *eb = nullptr;
    size_t cap_as = 0, cap_bs = 0, cap_ea = 0, cap_eb = 0;
    ~OzakiWork() {
        (void)hipFree(as);
        (void)hipFree(bs);
        (void)hipFree(ea);
        (void)hipFree(eb);
    }
    template <class T>
    static int grow(T **p, size_t &cap, size_t need) {
        if (need <= cap) return 0;
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        size_t n = need + need / 4 + 256;
        if (hipMalloc((void **)p, sizeof(T) * n) != hipSuccess) {
            cap = 0;
            return -4;
        }
        cap = n;
        return 0;
    }
    static int slices_ok(int s) { return L <= OZ_LMAX && (s == OzSlices<L>::DEF || s == OzSlices<L>::WIDE); }
    template <int S>
    int run_s(const GemmArgs &g, hipStream_t st) {
        int nbA = g.sA ? g.batch : 1, nbB = g.sB ? g.batch : 1;
        int nch = (g.K + OZ_KC - 1) / OZ_KC;
        int Mp = (g.M + OZ_T - 1) / OZ_T * OZ_T, Np = (g.N + OZ_T - 1) / OZ_T * OZ_T;
        size_t need_a = (size_t)nbA * nch * S * Mp * OZ_KC, need_b = (size_t)nbB * nch * S * Np * OZ_KC;
        if (need_a > cap_as || need_b > cap_bs || (size_t)nbA * g.M > cap_ea || (size_t)nbB * g.N > cap_eb)
            if (hipStreamSynchronize(st) != hipSuccess) return -2;  // queued kernels may still read the old buffers
        int rc = 0;
        rc |= grow(&as, cap_as, need_a);
        rc |= grow(&bs, cap_bs, need_b);
        rc |= grow(&ea, cap_ea, (size_t)nbA * g.M);
        rc |= grow(&eb, cap_eb, (size_t)nbB * g.N);
        if (rc) return -4;
        const bool same = gemm_b_is_at(g);
        k_fix_max<L><<<fix_max_grid(g, 0, g.M, nbA), 256, 0, st>>>(g, 0, g.M, ea);
        if (!same) k_fix_max<L><<<fix_max_grid(g, 1, g.N, nbB), 256, 0, st>>>(g, 1, g.N, eb);
        int64_t sa_n = (int64_t)Mp * nch * OZ_KC, sb_n = (int64_t)Np * nch * OZ_KC;
        k_oz_slice<L, S><<<dim3((unsigned)((sa_n + 255) / 256), nbA), 256, 0, st>>>(g, 0, g.M, Mp, nch, ea, as);
        if (!same)
            k_oz_slice<L, S><<<dim3((unsigned)((sb_n + 255) / 256), nbB), 256, 0, st>>>(g, 1, g.N, Np, nch, eb, bs);
        dim3 grd(Mp / OZ_T, Np / OZ_T, g.batch);
        hipEvent_t ev = prof ? prof->begin(st) : nullptr;
        k_gemm_oz<L, S><<<grd, 256, 0, st>>>(g, as, same ? as : bs, ea, same ? ea : eb, nbA, same ? nbA : nbB, Mp,
                                            same ? Mp : Np, nch);
        // algorithmic INT8 operations: 2 per multiply-add, S(S+1)/2 slice pairs per mp multiply-add
        if (prof) prof->end(st, ev, 0, gemm_outputs(g) * (double)g.K * 2.0 * (S * (S + 1) / 2));
        return hipGetLastError() == hipSuccess ? 0 : -2;
    }
    int run(const GemmArgs &g, hipStream_t st, int slices) {
        if constexpr (L > OZ_LMAX) {
            (void)g;
            (void)st;
            (void)slices;
            return -5;
        } else {
            if (slices <= 0) slices = OzSlices<L>::DEF;
            if (slices == OzSlices<L>::WIDE) return run_s<OzSlices<L>::WIDE>(g, st);
            if (slices == OzSlices<L>::DEF) return run_s<OzSlices<L>::DEF>(g, st);
            return -1;
        }
    }
};

}  // namespace sdpg
