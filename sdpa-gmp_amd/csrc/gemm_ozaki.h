// gemm_ozaki.h — Ozaki (INT8 slice, MFMA) form of the multiple-precision GEMM.
//
// 1. Block fixed point as in the plain form: row r of op(A) gets the exponent e_r of its
//    largest element and x becomes the integer Y = round(x * 2^(8S - 2 - e_r)),
//    |Y| < 2^(8S-2); the same for the columns of op(B).
// 2. Slicing (k_oz_slice; k_oz_fixslice does exponents and slices in one pass for K <= 64):
//    Y = sum_{p<S} y_p 256^(S-1-p) with balanced digits y_p in [-128, 127] (y_0 in [-64, 64]).
//    The digits are stored as int8 planes s[chunk][p][row][KC]: KC consecutive k of one row for one
//    slice.
// 3. Products: for the output digit position d = p + q the slice products
//    T_d = sum_{p+q=d} sum_k a_p[i,k] b_q[k,j] are exact INT32 sums.  Pairs with
//    p + q >= S lie below 2^-(8S-8) of the row/column scale and are not formed
//    (S(S+1)/2 slice products instead of S^2).  Two kernels:
//    - k_gemm_oz (K <= 64, the Cholesky panels and trailing updates; every K when S > 40): KC = 16;
//      every v_mfma_i32_16x16x64_i8 issues four (p, q) pairs of one d at once (lane group
//      g = lane / 16 carries pair (4u + g, q' - g), each over the same 16 k) into a per-digit INT32
//      accumulator that lives across chunks and is folded every 64 chunks;
//    - k_gemm_ozk (long K: the F1 products and Schur inner products): KC = 64; one MFMA per
//      (p, q) over 64 k into a per-digit INT32 accumulator that lives across chunks and is
//      carried every 32 chunks.
// 4. Recombination: sum_d T_d 256^-d is carried digit by digit from d = S-1 up to 0
//    into an exact fixed-point accumulator (S/4 fraction words + a 64-bit integer part)
//    held in registers across K chunks, then rounded once to a limb float with
//    exponent e_i + e_j - 12 and applied to C (set / subtract / add), all in full width.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "engine_common.h"
#include "mpx.h"

namespace sdpg {

typedef int oz_v4i __attribute__((ext_vector_type(4)));
typedef int oz_v16i __attribute__((ext_vector_type(16)));

constexpr int OZ_T = 32;   // output tile edge of one workgroup (2 x 2 waves of 16 x 16)
constexpr int OZ_KC = 16;  // k per chunk (one 16-byte operand per lane and slice)
constexpr int OZ_LONGK_MIN = 64;  // K above which the long-K kernel (k_gemm_ozk) is used
constexpr int OZ_FOLD = 64;       // chunks of the short-K kernel between folds of its digit sums
constexpr int OZ_QB = 8;          // column fragments per block of the short-K kernel's pair loop

// Balanced base-256 digits of x in the block fixed point of its row: Y = round(x * 2^(8S - 2 - emax))
// (|Y| < 2^(8S-2), two's complement over L limbs) = sum_u d_u 256^u, d_u in [-128, 127]; the digit of
// slice p is d_(S-1-p).  emax is the row's largest exponent (the zero exponent for an all-zero row).
// Z = Y + sum_u 128 * 256^u lies in [0, 2^(8S)) and its plain base-256 digits are d_u + 128 (the
// balanced representation is unique), so one multi-word addition and an XOR with 0x80 per byte give
// every digit: byte u % 4 of zw[u / 4] is d_u as a signed byte.
template <int L, int S>
__device__ __forceinline__ void oz_digit_words(const mpx::mp<L> &x, int32_t emax, uint32_t (&zw)[S / 4]) {
    if (emax == mpx::EZERO) emax = 0;
    uint32_t a[L];
    fx_convert<L>(x, emax, 32 * L - 8 * S + 2, a, false);
    uint32_t c = 0;
#pragma unroll
    for (int t = 0; t < S / 4; t++) {
        const uint64_t v = (uint64_t)a[t] + 0x80808080u + c;
        zw[t] = (uint32_t)v ^ 0x80808080u;
        c = (uint32_t)(v >> 32);
    }
}

// slice planes: ((((b * nch + ch) * S + p) * Rp + r) * KC + kk), Rp = R rounded up to OZ_T; KC = k per
// chunk (16 for the short-K kernel, 64 for the long-K kernel).  One element per thread (consecutive
// threads take consecutive k, one coalesced read per limb plane); each group of four lanes (four
// consecutive k of one row) transposes its digit words in registers, so that lane q holds the four
// k-consecutive bytes of digits 4w + q, and writes 4 bytes per store: S/4 stores per lane.
template <int L, int S, int KC>
__global__ void k_oz_slice(GemmArgs g, int which, int R, int Rp, int nch, const int32_t *eo, int8_t *out, int R2 = 0,
                           int Rp2 = 0, const int32_t *eo2 = nullptr, int8_t *out2 = nullptr) {
    static_assert(S % 4 == 0 && S <= 4 * L && KC % 4 == 0, "slice count");
    if (blockIdx.z) {  // both operands in one launch (op(B) in z = 1)
        which = 1;
        R = R2;
        Rp = Rp2;
        eo = eo2;
        out = out2;
    }
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int b = blockIdx.y;
    // whole waves leave here (Rp * KC is a multiple of 64), so every quad below is complete
    if (t >= (int64_t)Rp * nch * KC) return;
    int kk = (int)(t % KC);
    int64_t rest = t / KC;
    int r = (int)(rest % Rp);
    int ch = (int)(rest / Rp);
    int k = ch * KC + kk;
    uint32_t zw[S / 4];
    if (r < R && k < g.K) {
        PV src = which == 0 ? g.A : g.B;
        oz_digit_words<L, S>(ldm<L>(src, fx_at(g, which, b, r, k)), eo[(int64_t)b * R + r], zw);
    } else {
#pragma unroll
        for (int w = 0; w < S / 4; w++) zw[w] = 0;
    }
    const int q = kk & 3;
    int8_t *ob = out + ((((int64_t)b * nch + ch) * S) * Rp + r) * KC + (kk - q);
    const int64_t pstride = (int64_t)Rp * KC;
    // 4 x 4 byte transpose over the quad: lanes 0, 1 take the low halves of their own word and of
    // lane q ^ 2's, lanes 2, 3 the high halves; then bytes 0 and 2 pair with lane q ^ 1's
    const uint32_t sel1 = q < 2 ? 0x01000504u : 0x07060302u, sel2 = (q & 1) ? 0x07030501u : 0x02060004u;
#pragma unroll
    for (int w = 0; w < S / 4; w++) {
        uint32_t x = zw[w];
        uint32_t y = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, true);  // quad_perm 2,3,0,1
        x = __builtin_amdgcn_perm(x, y, sel1);
        y = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, true);  // quad_perm 1,0,3,2
        x = __builtin_amdgcn_perm(x, y, sel2);
        const int p = S - 1 - (4 * w + q);  // digit u = 4w + q is slice S-1-u
        *(uint32_t *)(ob + p * pstride) = x;
    }
}
// workgroups of 256 threads of k_oz_slice for n slice bytes per plane set (rows x padded K)
inline unsigned oz_slice_blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

// K <= 64 (at most four 16-k chunks, the Cholesky products): row exponents and slices in one pass, the
// same values as k_fix_max followed by k_oz_slice.  A workgroup of 1024 threads covers RW = 32 rows for
// K <= 32, 16 rows for K <= 64: thread (row rl, k = kt) reads its element (consecutive threads take
// consecutive rows, which are contiguous for column-major operands), the row's largest exponent is reduced
// in LDS, each thread writes its S digits into an LDS copy of the output block (per chunk and slice the
// block's RW rows x 16 k are contiguous in the slice planes), and the block leaves in 16-byte stores.
// blockIdx.z = 1: the second operand (R2, Rp2, ...).
template <int L, int S>
__global__ __launch_bounds__(1024) void k_oz_fixslice(GemmArgs g, int which, int R, int Rp, int nch, int32_t *eo,
                                                      int8_t *out, int R2 = 0, int Rp2 = 0, int32_t *eo2 = nullptr,
                                                      int8_t *out2 = nullptr) {
    static_assert(S % 4 == 0 && S <= 4 * L, "slice count");
    if (blockIdx.z) {
        which = 1;
        R = R2;
        Rp = Rp2;
        eo = eo2;
        out = out2;
    }
    const int RW = g.K <= 32 ? 32 : 16, KP = 1024 / RW;  // rows per workgroup, k per row
    const int r0 = blockIdx.x * RW, b = blockIdx.y;
    if (r0 >= Rp) return;  // past this operand's rows (the grid covers the larger operand)
    __shared__ int32_t red[32 * 33];      // [row][KP + 1]
    __shared__ uint4 sout[2 * S * 32];  // [ch][p][row][16 k] as 16-byte rows (nch RW <= 64)
    const int rl = threadIdx.x % RW, kt = threadIdx.x / RW, r = r0 + rl;
    const bool live = r < R && kt < g.K;
    const mpx::mp<L> x = live ? ldm<L>(which == 0 ? g.A : g.B, fx_at(g, which, b, r, kt)) : mpx::zero<L>();
    red[rl * (KP + 1) + kt] = live ? x.e : mpx::EZERO;
    __syncthreads();
    if (kt == 0) {
        int32_t e = mpx::EZERO;
#pragma unroll 8
        for (int q = 0; q < KP; q++) e = red[rl * (KP + 1) + q] > e ? red[rl * (KP + 1) + q] : e;
        red[rl * (KP + 1) + KP] = e;
        if (r < R) eo[(int64_t)b * R + r] = e;
    }
    __syncthreads();
    const int ch = kt >> 4, kk = kt & 15;
    if (ch < nch) {
        uint32_t zw[S / 4];
        if (live) {
            oz_digit_words<L, S>(x, red[rl * (KP + 1) + KP], zw);
        } else {
#pragma unroll
            for (int w = 0; w < S / 4; w++) zw[w] = 0;
        }
        int8_t *so = (int8_t *)sout;
#pragma unroll
        for (int p = 0; p < S; p++) {
            const int u = S - 1 - p;
            so[((ch * S + p) * RW + rl) * 16 + kk] = (int8_t)(zw[u >> 2] >> (8 * (u & 3)));
        }
    }
    __syncthreads();
    // per (ch, p): RW rows x 16 bytes at ((((b * nch + ch) * S + p) * Rp + r0) * 16
    const int nvec = nch * S * RW;
    for (int v = threadIdx.x; v < nvec; v += blockDim.x) {
        const int row = v % RW, cp = v / RW;  // cp = ch * S + p
        uint4 *dst = (uint4 *)(out + ((((int64_t)b * nch) * S + cp) * Rp + r0 + row) * 16);
        *dst = sout[cp * RW + row];
    }
}

// Both kernels form the transposed tile, MFMA(B fragment, A fragment) = (A B)^T, so that lane l holds
// C[i0 + l % 16, j0 + 4 (l / 16) + e]: the 16 lanes of a group touch 16 consecutive rows of a
// column-major C (coalesced 64-byte accesses per limb plane in the read-modify-write of C).
//
// Epilogue of one output: the exact fixed-point sum (RI * 2^(32 NW) + sum_w R[w] 2^(32 (NW-1-w)),
// two's complement over NW + 2 words) is rounded once to a limb float with exponent
// e_i + e_j - 12 - 8 (S-1) and applied to C[i, j] (set / subtract / add) in full width.
template <int L, int S>
__device__ __forceinline__ mpx::mp<L> oz_value(const uint32_t (&R)[S / 4], int64_t RI, int32_t ei, int32_t ej) {
    constexpr int NW = S / 4;
    constexpr int NX = NW + 2;
    uint32_t x[NX];
#pragma unroll
    for (int w = 0; w < NW; w++) x[w] = R[NW - 1 - w];
    x[NW] = (uint32_t)RI;
    x[NW + 1] = (uint32_t)((uint64_t)RI >> 32);
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
        // the top L + 1 words of the accumulator; below an accumulator narrower than that (S < 4L - 4:
        // fewer fraction bits than the result carries) the result's low words are zero
        uint32_t y[L + 1];
#pragma unroll
        for (int q = 0; q <= L; q++) y[q] = NX - 1 - L + q >= 0 ? x[NX - 1 - L + q >= 0 ? NX - 1 - L + q : 0] : 0u;
        int32_t E = ei + ej;
        E = (E < -(1 << 28) ? 0 : E) - 12 - 8 * (S - 1);
        int32_t ex = 32 * NX - (int32_t)lz + E;
        mpx::round_store<L>(res, y, ex, neg);
    }
    return res;
}

// Exact fold of the per-digit INT32 sums into the fixed-point accumulator of each output: sum_d T_d 256^-d
// carried from d = S-1 up, digit d in byte 3 - d % 4 of word d / 4 of R, the carry out of d = 0 into
// the integer part RI; the sums restart at zero.
template <int S>
__device__ __forceinline__ void oz_fold(oz_v4i (&acc)[S], uint32_t (&R)[4][S / 4], int64_t (&RI)[4]) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
        int32_t c = 0;
#pragma unroll
        for (int d = S - 1; d >= 0; d--) {
            const int sh = 8 * (3 - d % 4);
            const int32_t v = acc[d][e] + c + (int32_t)((R[e][d / 4] >> sh) & 255u);
            R[e][d / 4] = (R[e][d / 4] & ~(255u << sh)) | (((uint32_t)v & 255u) << sh);
            c = v >> 8;
        }
        RI[e] += c;
    }
#pragma unroll
    for (int d = 0; d < S; d++) acc[d] = oz_v4i{0, 0, 0, 0};
}

// s_waitcnt vmcnt(n) alone (expcnt and lgkmcnt at their maxima), gfx9 encoding
constexpr int ozk_wait_vm(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }

// LDS bytes of k_gemm_oz: two buffers of S x OZ_T 16-byte fragments per operand
constexpr size_t oz_lds_bytes(int S) { return (size_t)4 * S * OZ_T * 16; }

// Short-K product kernel: one workgroup per 32 x 32 output tile, 2 x 2 waves of 16 x 16, the
// slices of the tile's rows and columns staged in LDS chunk by chunk, double-buffered: the next chunk's
// slices are copied global -> LDS by buffer loads (no registers) while the current chunk's MFMAs run.
// Pair schedule: the MFMA numbered (u, q') gives lane group g = lane / 16 the slice pair
// (p, q) = (4u + g, q' - g), so the four groups of a wave all add into the same digit position
// d = p + q = 4u + q' and that MFMA accumulates into the INT32 digit sum T_d (groups with q < 0 feed a
// zero fragment).  Every pair with p + q < S occurs once (g = p mod 4, u = p / 4, q' = q + g).  A wave
// therefore reads its S/4 row fragments (p = g mod 4) and each column fragment once per chunk, and
// keeps all S digit sums across the chunks: a digit sum over one chunk is bounded by
// S * 16 * 2^14 (|digits| <= 128) <= 2^23.6 for S <= 48, so the sums are folded into the exact
// fixed-point accumulator every OZ_FOLD chunks (and once at the end: the only fold for K <= 64).
template <int L, int S>
__global__ __launch_bounds__(256) void k_gemm_oz(GemmArgs g, const int8_t *__restrict__ as,
                                                 const int8_t *__restrict__ bs, const int32_t *__restrict__ ea,
                                                 const int32_t *__restrict__ eb, int nbA, int nbB, int Mp, int Np,
                                                 int nch) {
    constexpr int NW = S / 4, NU = (S + 3) / 4;
    static_assert(S <= 48, "digit sums of OZ_FOLD chunks must stay below 2^31");
    constexpr int NF = S * OZ_T;  // 16-byte fragments of one operand and chunk (a multiple of 64)
    extern __shared__ oz_v4i ozsh[];  // [2][NF] of A, then [2][NF] of B (oz_lds_bytes)
    const int i0 = blockIdx.x * OZ_T, j0 = blockIdx.y * OZ_T, b = blockIdx.z;
    if (g.lower && i0 + OZ_T - 1 < j0) return;  // tile strictly above the diagonal
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv & 1, wc = wv >> 1, grp = lane >> 4;
    const int ar = wr * 16 + (lane & 15), bc = wc * 16 + (lane & 15);
    const int bA = nbA > 1 ? b : 0, bB = nbB > 1 ? b : 0;
    // chunk ch's slices into LDS buffer bf: fragment t of an operand (slice t / OZ_T, tile row t % OZ_T)
    // at ((ch S + slice) Mp + i0 + row) 16 bytes from this batch entry's planes; a wave moves 64
    // consecutive fragments (a wave-contiguous LDS block), the chunk offset in the scalar offset
    const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(as + (int64_t)bA * nch * S * Mp * 16), (short)0, (int)0xFFFFFFFFu, 0x00020000);
    const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(bs + (int64_t)bB * nch * S * Np * 16), (short)0, (int)0xFFFFFFFFu, 0x00020000);
    auto stage = [&](int ch, int bf) {
#pragma unroll
        for (int g0 = 64 * wv; g0 < NF; g0 += 256) {
            const int t = g0 + lane;
            const int voA = ((t / OZ_T) * Mp + i0 + t % OZ_T) * 16, voB = ((t / OZ_T) * Np + j0 + t % OZ_T) * 16;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rA, (__attribute__((address_space(3))) void *)&ozsh[bf * NF + g0], 16, voA, ch * S * Mp * 16, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rB, (__attribute__((address_space(3))) void *)&ozsh[(2 + bf) * NF + g0], 16, voB, ch * S * Np * 16, 0,
                0);
        }
    };
    stage(0, 0);
    // The four outputs of this lane, C[i, j0c + e]: their exponents and, for op != SET, the raw words
    // of their old C values are requested before the products (at clamped, always valid addresses,
    // without a branch, and decoded only at the end) so that the loads' latency hides under them.
    const int i = i0 + ar, j0c = j0 + wc * 16 + 4 * grp;
    const int ic = i < g.M ? i : g.M - 1;
    const int32_t ei = ea[(int64_t)bA * g.M + ic];
    // (for S > 40 the C words are read in the epilogue: the registers go to the S digit sums)
    constexpr bool CPRE = S <= 40;
    int32_t ej[4];
    uint32_t craw[4][L + 1];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int jc = j0c + e < g.N ? j0c + e : g.N - 1;
        ej[e] = eb[(int64_t)bB * g.N + jc];
        if (CPRE && g.op != GEMM_SET) {
            const int64_t ci = g.offC + (int64_t)b * g.sC + ic + (int64_t)g.ldc * jc;
#pragma unroll
            for (int q = 0; q <= L; q++) craw[e][q] = g.C.p[(int64_t)q * g.C.n + ci];
        }
    }
    oz_v4i acc[S];
#pragma unroll
    for (int d = 0; d < S; d++) acc[d] = oz_v4i{0, 0, 0, 0};
    uint32_t R[4][NW];
    int64_t RI[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        RI[e] = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) R[e][w] = 0;
    }
    for (int ch = 0; ch < nch; ch++) {
        if (ch > 0 && ch % OZ_FOLD == 0) oz_fold<S>(acc, R, RI);
        // this chunk's slices have landed (this wave's copies, then every wave's); the other buffer was
        // last read in the chunk before, which every wave has finished: the next chunk's copies go there
        __builtin_amdgcn_s_waitcnt(ozk_wait_vm(0));
        __syncthreads();
        if (ch + 1 < nch) stage(ch + 1, (ch + 1) & 1);
        const oz_v4i *lA = ozsh + (ch & 1) * NF, *lB = ozsh + (2 + (ch & 1)) * NF;
        oz_v4i af[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int p = 4 * u + grp;
            af[u] = lA[(p < S ? p : S - 1) * OZ_T + ar];
        }
        // column fragments in blocks of OZ_QB, block b + 1 read while block b's MFMAs run; the scheduling
        // fences keep the compiler from hoisting every block's reads (S fragments would not fit
        // beside the S digit sums for the larger S)
        constexpr int NQB = (S + OZ_QB - 1) / OZ_QB;
        oz_v4i bcur[OZ_QB], bnxt[OZ_QB];
#pragma unroll
        for (int x = 0; x < OZ_QB; x++) {
            const int q = x - grp;
            bcur[x] = lB[(q >= 0 && x < S ? q : 0) * OZ_T + bc];
        }
#pragma unroll
        for (int qb = 0; qb < NQB; qb++) {
            if (qb + 1 < NQB) {
#pragma unroll
                for (int x = 0; x < OZ_QB; x++) {
                    const int qq = (qb + 1) * OZ_QB + x, q = qq - grp;
                    bnxt[x] = lB[(q >= 0 && qq < S ? q : 0) * OZ_T + bc];
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int x = 0; x < OZ_QB; x++) {
                const int qq = qb * OZ_QB + x;
                if (qq >= S) break;
                const oz_v4i bv = qq - grp >= 0 ? bcur[x] : oz_v4i{0, 0, 0, 0};
#pragma unroll
                for (int u = 0; u < NU; u++) {
                    if (4 * u + qq < S)
                        acc[4 * u + qq] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bv, af[u], acc[4 * u + qq], 0, 0, 0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int x = 0; x < OZ_QB; x++) bcur[x] = bnxt[x];
        }
    }
    oz_fold<S>(acc, R, RI);
#pragma unroll
    for (int e = 0; e < 4; e++) {
        // the four outputs interleave where registers allow; for the larger S one after another
        if constexpr (S > 40) __builtin_amdgcn_sched_barrier(0);
        const int j = j0c + e;
        mpx::mp<L> res = oz_value<L, S>(R[e], RI[e], ei, ej[e]);
        if (g.op != GEMM_SET) {
            if constexpr (!CPRE) {
                const int jc = j < g.N ? j : g.N - 1;
                const int64_t ci = g.offC + (int64_t)b * g.sC + ic + (int64_t)g.ldc * jc;
#pragma unroll
                for (int q = 0; q <= L; q++) craw[e][q] = g.C.p[(int64_t)q * g.C.n + ci];
            }
            mpx::mp<L> cold;
#pragma unroll
            for (int q = 0; q < L; q++) cold.m[q] = craw[e][q];
            cold.e = (int32_t)craw[e][L] >> 1;
            cold.s = craw[e][L] & 1u;
            res = mpx::add<L>(cold, g.op == GEMM_SUB ? mpx::neg<L>(res) : res);
        }
        // every output is formed (padded rows and columns from zero slices and clamped C words), so the
        // four chains interleave; only the stores are predicated
        if (i < g.M && j < g.N && !(g.lower && i < j))
            stm<L>(g.C, g.offC + (int64_t)b * g.sC + i + (int64_t)g.ldc * j, res);
    }
}

// ---- long-K form ----
// For K beyond one short chunk the per-chunk digit carries of k_gemm_oz (a VALU pass per 16 k)
// dominate.  k_gemm_ozk keeps one INT32 accumulator per output digit position d across chunks of 64 k
// (each v_mfma_i32_16x16x64_i8 forms one slice pair (p, q) over 64 k and adds into digit p + q), and
// carries the digit sums into the fixed-point accumulator only every OZK_CARRY chunks: a digit sum
// over one chunk is bounded by S * 64 * 2^14 < 2^25.2, so 32 chunks stay below 2^31.
// One wave per 16 x 16 output tile, one wave per SIMD (the S digit accumulators and the S column
// fragments of the chunk fill the register file), so the wave hides its own load latency:
//  - the next chunk's first 28 column fragments are copied global -> LDS by buffer loads (28 KB per
//    wave) while the current chunk runs from registers; in the next chunk each moves LDS -> registers
//    a few batches before its first use, so register pressure grows with the batches; the last
//    column fragments are first used late in a chunk and load from global memory at its start;
//  - batches run by descending row slice p (batch p forms the pairs (p, q <= S-1-p)); the row
//    fragments stream through a ring of registers, each requested several batches ahead, the last
//    batches of a chunk requesting the next chunk's first ones.
// Tile mapping: the tile grid of every batch entry is cut into units of BR x BC tiles (16 x 8 when the
// grid is large); unit u runs on XCD u mod 8 (workgroups are dispatched to XCDs by id mod 8), whose
// waves of one unit run together and share its BR row and BC column fragments in that XCD's L2
// (without this, each wave streams its row fragments from HBM: about 8 TB/s at the MFMA rate).
constexpr int OZK_KC = 64;
constexpr int OZK_CARRY = 32;
constexpr int OZK_SMAX = 48;  // register budget: S accumulators (AGPRs) + S column fragments per lane
constexpr int OZK_SMIN = 12;  // at least 4 column fragments prefetched into LDS

// row-fragment prefetch distance (batches): the first batches of a chunk form 1, 2, 3, ... pairs, so a
// distance of 8 requests their fragments only about 36 MFMAs ahead; 16 (1 % faster on the dense rank-2
// products); 4 for the wide slice counts, whose column fragments take more registers
template <int S>
struct OzkPD {
    static constexpr int value = S > 36 ? 4 : 16;
};

// workgroups of k_gemm_ozk for a tm x tn tile grid per batch entry: units of BR x BC tiles, every XCD
// (id mod 8) running whole units
inline int64_t ozk_workgroups(int tm, int tn, int batch) {
    const int BR = tm < 16 ? tm : 16, BC = (128 / BR) < tn ? 128 / BR : tn;
    const int64_t units = (int64_t)((tm + BR - 1) / BR) * ((tn + BC - 1) / BC) * batch;
    return 8 * (int64_t)BR * BC * ((units + 7) / 8);
}

template <int S>
__device__ __forceinline__ void ozk_carry(oz_v4i (&acc)[S], uint32_t (*sR)[64], int lane) {
    constexpr int NW = S / 4;
    int64_t c[4] = {0, 0, 0, 0};
#pragma unroll
    for (int w = NW - 1; w >= 0; w--) {
        uint32_t Rw[4];
#pragma unroll
        for (int e = 0; e < 4; e++) Rw[e] = sR[e * NW + w][lane];
#pragma unroll
        for (int D = 4 * w + 3; D >= 4 * w; D--) {
            const int sh = 8 * (3 - D % 4);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                int64_t v = (int64_t)acc[D][e] + c[e] + (int64_t)((Rw[e] >> sh) & 255u);
                Rw[e] = (Rw[e] & ~(255u << sh)) | (((uint32_t)v & 255u) << sh);
                c[e] = v >> 8;
            }
            acc[D] = oz_v4i{0, 0, 0, 0};
        }
#pragma unroll
        for (int e = 0; e < 4; e++) sR[e * NW + w][lane] = Rw[e];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
        uint64_t ri = ((uint64_t)sR[4 * NW + 2 * e + 1][lane] << 32) | sR[4 * NW + 2 * e][lane];
        ri = (uint64_t)((int64_t)ri + c[e]);
        sR[4 * NW + 2 * e][lane] = (uint32_t)ri;
        sR[4 * NW + 2 * e + 1][lane] = (uint32_t)(ri >> 32);
    }
}

// s_waitcnt lgkmcnt(0) alone (vmcnt and expcnt at their maxima), gfx9 encoding
constexpr int OZK_WAIT_LGKM0 = 0xC07F;
// column fragments q < OzkQL of the next chunk are prefetched into LDS (at most 28 KB per wave, so
// that four waves fit in a CU's LDS with their accumulators); the others, first used by the last
// batches of the descending order, are loaded at the start of their chunk
template <int S>
struct OzkQL {
    // LDS per wave: QL KB of prefetch + (S + 8) 256-byte accumulator rows, at most 40 KB (four waves
    // per CU in its 160 KB)
    static constexpr int lds_cap = (40960 - (S + 8) * 256) / 1024;
    static constexpr int value = (S - 8 < 28 ? S - 8 : 28) < lds_cap ? (S - 8 < 28 ? S - 8 : 28) : lds_cap;
};
constexpr int OZK_LA = 4;  // batches between reading a column fragment from LDS and its first use

template <int L, int S>
__global__ __launch_bounds__(64) void k_gemm_ozk(GemmArgs g, const int8_t *__restrict__ as,
                                                const int8_t *__restrict__ bs, const int32_t *__restrict__ ea,
                                                const int32_t *__restrict__ eb, int nbA, int nbB, int Mp, int Np,
                                                int nch) {
    constexpr int NW = S / 4, PD = OzkPD<S>::value, QL = OzkQL<S>::value;
    static_assert(S <= OZK_SMAX && QL > 0, "long-K form register budget");
    // LDS per wave (4 waves per CU): the next chunk's first QL column fragments (1 KB each) and the
    // fixed-point accumulators (touched only at carries and in the epilogue)
    __shared__ oz_v4i sB[QL][64];
    __shared__ uint32_t sR[4 * NW + 8][64];
    // tile of this wave: see ozk_workgroups (workgroup id -> XCD id % 8, its k-th unit, slot in the unit)
    const int tm = Mp / 16, tn = Np / 16;
    const int BR = tm < 16 ? tm : 16, BC = (128 / BR) < tn ? 128 / BR : tn, SU = BR * BC;
    const int nsr = (tm + BR - 1) / BR, nsc = (tn + BC - 1) / BC;
    const int id = blockIdx.x, local = id >> 3;
    const int unit = 8 * (local / SU) + (id & 7), w = local % SU;
    const int b = unit / (nsr * nsc), su = unit % (nsr * nsc);
    if (b >= g.batch) return;
    const int ti = (su % nsr) * BR + w % BR, tj = (su / nsr) * BC + w / BR;
    if (ti >= tm || tj >= tn) return;
    const int i0 = ti * 16, j0 = tj * 16;
    if (g.lower && i0 + 15 < j0) return;
    const int lane = threadIdx.x, grp = lane >> 4;
    const int bA = nbA > 1 ? b : 0, bB = nbB > 1 ? b : 0;
#pragma unroll
    for (int q = 0; q < 4 * NW + 8; q++) sR[q][lane] = 0;
    oz_v4i acc[S];
#pragma unroll
    for (int d = 0; d < S; d++) acc[d] = oz_v4i{0, 0, 0, 0};
    // a (chunk, slice) plane holds Rp rows of 64 bytes (4 fragments of 16); lane l reads row l % 16,
    // bytes 16 (l / 16) .. +16 of the chunk.  Buffer loads: the batch base in the resource (SGPRs),
    // the lane's byte offset in one VGPR and the plane offset in an SGPR, so a load in flight holds
    // no address registers (the driver keeps one batch's planes below 4 GiB).
    const int64_t pbA = (int64_t)Mp * OZK_KC, pbB = (int64_t)Np * OZK_KC;  // plane bytes
    const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(as + (int64_t)bA * nch * S * pbA), (short)0, (int)0xFFFFFFFFu, 0x00020000);
    const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(bs + (int64_t)bB * nch * S * pbB), (short)0, (int)0xFFFFFFFFu, 0x00020000);
    const int voA = ((i0 + (lane & 15)) * 4 + grp) * 16, voB = ((j0 + (lane & 15)) * 4 + grp) * 16;
    const uint32_t psA = (uint32_t)pbA, psB = (uint32_t)pbB;
    auto ldA = [&](int plane) -> oz_v4i {
        return __builtin_amdgcn_raw_buffer_load_b128(rA, voA, (int)((uint32_t)plane * psA), 0);
    };
    auto ldB = [&](int plane) -> oz_v4i {
        return __builtin_amdgcn_raw_buffer_load_b128(rB, voB, (int)((uint32_t)plane * psB), 0);
    };
    auto ldB_lds = [&](int plane, int q) {
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rB, (__attribute__((address_space(3))) void *)&sB[q][0], 16, voB,
                                                 (int)((uint32_t)plane * psB), 0, 0);
    };
#pragma unroll
    for (int q = 0; q < QL; q++) ldB_lds(q, q);
    // the row-fragment ring: slot k % PD holds the row fragment of batch k = S-1-p; the first PD of a
    // chunk are requested during the last PD batches of the chunk before (chunk 0's here, after the
    // column prefetch: at the top of every chunk only these PD loads follow the prefetch)
    __builtin_amdgcn_sched_barrier(0);
    oz_v4i ap[PD];
#pragma unroll
    for (int t = 0; t < PD; t++) ap[t] = ldA(S - 1 - t);
    // one chunk of 64 k
    auto run_chunk = [&](int ch) {
        const int c0 = ch * S;
        oz_v4i bq[S];
        // the early column fragments were prefetched into LDS during the previous chunk (wait for those
        // loads alone, before this chunk issues its own)
        __builtin_amdgcn_s_waitcnt(ozk_wait_vm(PD));
        // the late column fragments from global memory
#pragma unroll
        for (int q = QL; q < S; q++) bq[q] = ldB(c0 + q);
#pragma unroll
        for (int q = 0; q < OZK_LA && q < QL; q++) bq[q] = sB[q][lane];
        // batches by descending row slice p (batch p forms the pairs (p, q <= S-1-p), first using column
        // fragment S-1-p); column fragments come from LDS OZK_LA batches before their first use, and
        // once the last one is read the next chunk's prefetch into LDS starts; the row fragments
        // stream through the ring, each requested PD batches ahead (the last PD batches request the
        // next chunk's first PD; the last chunk re-reads one of its own planes instead)
#pragma unroll
        for (int p = S - 1; p >= 0; p--) {
            {
                const int qn = S - 1 - p + OZK_LA;  // column fragment read from LDS at this batch
                if (qn < QL) bq[qn] = sB[qn][lane];
                if (qn == QL && ch + 1 < nch) {
                    __builtin_amdgcn_s_waitcnt(OZK_WAIT_LGKM0);
#pragma unroll
                    for (int q = 0; q < QL; q++) ldB_lds(c0 + S + q, q);
                }
            }
            const int t = (S - 1 - p) % PD;
            const oz_v4i a = ap[t];
            if (p - PD >= 0)
                ap[t] = ldA(c0 + p - PD);
            else
                ap[t] = ldA(ch + 1 < nch ? c0 + S + (S - 1 - t) : c0);
            // keep the request here: left to itself the scheduler sinks every load next to its first use
            // (the register file is nearly full), and the ring's lookahead is lost
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q <= S - 1 - p; q++)
                acc[p + q] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bq[q], a, acc[p + q], 0, 0, 0);
        }
    };
    // chunks in carry groups of OZK_CARRY, two per loop iteration: with one, the allocator renames the
    // digit accumulators between the loop's entry and its back edge and moves all of them every chunk
    for (int cg = 0; cg < nch; cg += OZK_CARRY) {
        const int cg1 = cg + OZK_CARRY < nch ? cg + OZK_CARRY : nch;
        int ch = cg;
        for (; ch + 1 < cg1; ch += 2) {
            run_chunk(ch);
            run_chunk(ch + 1);
        }
        if (ch < cg1) run_chunk(ch);
        ozk_carry<S>(acc, sR, lane);
    }
    // the lane's row exponent and its four column exponents are requested together (clamped indices);
    // every output is formed and only the stores (and for SUB / ADD the old C reads) are predicated
    const int i = i0 + (lane & 15), ic = i < g.M ? i : g.M - 1;
    const int32_t ei = ea[(int64_t)bA * g.M + ic];
    int32_t ej[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int j = j0 + 4 * grp + e, jc = j < g.N ? j : g.N - 1;
        ej[e] = eb[(int64_t)bB * g.N + jc];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
        uint32_t R[NW];
#pragma unroll
        for (int q = 0; q < NW; q++) R[q] = sR[e * NW + q][lane];
        const int64_t RI = (int64_t)(((uint64_t)sR[4 * NW + 2 * e + 1][lane] << 32) | sR[4 * NW + 2 * e][lane]);
        mpx::mp<L> res = oz_value<L, S>(R, RI, ei, ej[e]);
        const int j = j0 + 4 * grp + e;
        if (i < g.M && j < g.N && !(g.lower && i < j)) {
            const int64_t ci = g.offC + (int64_t)b * g.sC + i + (int64_t)g.ldc * j;
            if (g.op == GEMM_SUB)
                res = mpx::sub<L>(ldm<L>(g.C, ci), res);
            else if (g.op == GEMM_ADD)
                res = mpx::add<L>(ldm<L>(g.C, ci), res);
            stm<L>(g.C, ci, res);
        }
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
    // the slice counts built (multiples of 4, at least 8): 4L - 12, 4L - 8, 4L - 4, 4L; a request
    // below the full limb width (a 200-bit request runs on 320-bit limbs) needs fewer than 4L - 4
    static constexpr int MIN = 4 * L - 12 >= 8 ? 4 * L - 12 : 8;
    static constexpr bool built(int s) { return s % 4 == 0 && s >= MIN && s <= WIDE; }
};

template <int L>
struct OzakiWork {
    KProf *prof = nullptr;
    int8_t *as = nullptr, *bs = nullptr;
    int32_t *ea = nullptr, *eb = nullptr;
    size_t cap_as = 0, cap_bs = 0, cap_ea = 0, cap_eb = 0;
    Arena *arena = nullptr;  // when set, the workspace is carved from it (no hipMalloc / hipFree)
    ~OzakiWork() {
        if (arena) return;
        (void)hipFree(as);
        (void)hipFree(bs);
        (void)hipFree(ea);
        (void)hipFree(eb);
    }
    template <class T>
    int grow(T **p, size_t &cap, size_t need) {
        if (need <= cap) return 0;
        if (arena) {
            *p = (T *)arena->take(sizeof(T) * (need ? need : 1));
            cap = *p ? need : 0;
            return *p ? 0 : -4;
        }
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
    static int slices_ok(int s) { return L <= OZ_LMAX && OzSlices<L>::built(s); }
    OpCache<int8_t> kept;  // slices and exponents of operands that never change
    template <int S>
    int run_s(const GemmArgs &g, hipStream_t st, int keep) {
        // long K: one slice pair per MFMA over 64-k chunks (k_gemm_ozk); otherwise 16-k chunks, four slice
        // pairs of one digit per MFMA (k_gemm_oz); both keep per-digit sums across chunks
        int Mp = (g.M + OZ_T - 1) / OZ_T * OZ_T, Np = (g.N + OZ_T - 1) / OZ_T * OZ_T;
        // the long-K kernel addresses one batch's slice planes with 32-bit buffer offsets
        const int64_t nch64 = (g.K + OZK_KC - 1) / OZK_KC;
        const bool longk = S <= OZK_SMAX && S >= OZK_SMIN && g.K > OZ_LONGK_MIN &&
                           nch64 * S * (int64_t)std::max(Mp, Np) * OZK_KC < ((int64_t)1 << 32);
        const int KC = longk ? OZK_KC : OZ_KC;
        int nbA = g.sA ? g.batch : 1, nbB = g.sB ? g.batch : 1;
        int nch = (g.K + KC - 1) / KC;
        size_t need_a = (size_t)nbA * nch * S * Mp * KC, need_b = (size_t)nbB * nch * S * Np * KC;
        const bool same = gemm_b_is_at(g);
        const bool keepA = (keep & GEMM_KEEP_A) != 0, keepB = (keep & GEMM_KEEP_B) != 0 && !same;
        if ((!keepA && (need_a > cap_as || (size_t)nbA * g.M > cap_ea)) ||
            (!keepB && (need_b > cap_bs || (size_t)nbB * g.N > cap_eb)))
            if (hipStreamSynchronize(st) != hipSuccess) return -2;  // queued kernels may still read the old buffers
        // row exponents and slices of each operand: kept copies (filled once) or the scratch buffers
        int8_t *pa = as, *pb = bs;
        int32_t *qa = ea, *qb = eb;
        bool fillA = true, fillB = !same;
        const int form = 1000 * S + KC;
        if (keepA) {
            auto *e = kept.get(op_key(g, 0, form), need_a, (size_t)nbA * g.M, fillA);
            if (!e) return -4;
            pa = e->data;
            qa = e->exps;
        } else {
            int rc = grow(&as, cap_as, need_a) | grow(&ea, cap_ea, (size_t)nbA * g.M);
            if (rc) return -4;
            pa = as;
            qa = ea;
        }
        if (keepB) {
            auto *e = kept.get(op_key(g, 1, form), need_b, (size_t)nbB * g.N, fillB);
            if (!e) return -4;
            pb = e->data;
            qb = e->exps;
        } else if (!same) {
            int rc = grow(&bs, cap_bs, need_b) | grow(&eb, cap_eb, (size_t)nbB * g.N);
            if (rc) return -4;
            pb = bs;
            qb = eb;
        }
        int64_t sa_n = (int64_t)Mp * nch * KC, sb_n = (int64_t)Np * nch * KC;
        if (!longk && g.K <= 64 && (fillA || fillB)) {
            // exponents and slices in one pass (both operands in one launch when both are needed)
            const int RW = g.K <= 32 ? 32 : 16;  // rows per workgroup (k_oz_fixslice)
            if (fillA && fillB && nbA == nbB) {
                k_oz_fixslice<L, S><<<dim3((unsigned)(std::max(Mp, Np) / RW), nbA, 2), 1024, 0, st>>>(
                    g, 0, g.M, Mp, nch, qa, pa, g.N, Np, qb, pb);
            } else {
                if (fillA)
                    k_oz_fixslice<L, S><<<dim3((unsigned)(Mp / RW), nbA), 1024, 0, st>>>(g, 0, g.M, Mp, nch, qa, pa);
                if (fillB)
                    k_oz_fixslice<L, S><<<dim3((unsigned)(Np / RW), nbB), 1024, 0, st>>>(g, 1, g.N, Np, nch, qb, pb);
            }
            fillA = fillB = false;
        }
        if (fillA && fillB && nbA == nbB && fx_rcontig_host(g, 0) == fx_rcontig_host(g, 1)) {
            // both operands in one launch per pass
            dim3 gm = fix_max_grid(g, 0, std::max(g.M, g.N), nbA);
            k_fix_max<L><<<dim3(gm.x, gm.y, 2), 256, 0, st>>>(g, 0, g.M, qa, g.N, qb);
            dim3 gs(oz_slice_blocks(std::max(sa_n, sb_n)), nbA, 2);
            if (longk)
                k_oz_slice<L, S, OZK_KC><<<gs, 256, 0, st>>>(g, 0, g.M, Mp, nch, qa, pa, g.N, Np, qb, pb);
            else
                k_oz_slice<L, S, OZ_KC><<<gs, 256, 0, st>>>(g, 0, g.M, Mp, nch, qa, pa, g.N, Np, qb, pb);
            fillA = fillB = false;
        }
        if (fillA) {
            k_fix_max<L><<<fix_max_grid(g, 0, g.M, nbA), 256, 0, st>>>(g, 0, g.M, qa);
            if (longk)
                k_oz_slice<L, S, OZK_KC><<<dim3(oz_slice_blocks(sa_n), nbA), 256, 0, st>>>(g, 0, g.M, Mp, nch, qa, pa);
            else
                k_oz_slice<L, S, OZ_KC><<<dim3(oz_slice_blocks(sa_n), nbA), 256, 0, st>>>(g, 0, g.M, Mp, nch, qa, pa);
        }
        if (fillB) {
            k_fix_max<L><<<fix_max_grid(g, 1, g.N, nbB), 256, 0, st>>>(g, 1, g.N, qb);
            if (longk)
                k_oz_slice<L, S, OZK_KC><<<dim3(oz_slice_blocks(sb_n), nbB), 256, 0, st>>>(g, 1, g.N, Np, nch, qb, pb);
            else
                k_oz_slice<L, S, OZ_KC><<<dim3(oz_slice_blocks(sb_n), nbB), 256, 0, st>>>(g, 1, g.N, Np, nch, qb, pb);
        }
        if (same) {
            pb = pa;
            qb = qa;
        }
        hipEvent_t ev = prof ? prof->begin(st) : nullptr;
        if (longk) {
            const int64_t nwg = ozk_workgroups(Mp / 16, Np / 16, g.batch);
            if constexpr (S <= OZK_SMAX && S >= OZK_SMIN)
                k_gemm_ozk<L, S><<<dim3((unsigned)nwg, 1), 64, 0, st>>>(g, pa, pb, qa, qb, nbA, same ? nbA : nbB, Mp,
                                                                        same ? Mp : Np, nch);
        } else {
            // the kernel addresses one batch entry's slice planes with 32-bit buffer offsets
            if ((int64_t)nch * S * std::max(Mp, Np) * 16 >= ((int64_t)1 << 31)) return -6;
            dim3 grd(Mp / OZ_T, Np / OZ_T, g.batch);
            static const bool lds_set = hipFuncSetAttribute((const void *)k_gemm_oz<L, S>,
                                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                                            (int)oz_lds_bytes(S)) == hipSuccess;
            if (!lds_set) return -2;
            k_gemm_oz<L, S><<<grd, 256, oz_lds_bytes(S), st>>>(g, pa, pb, qa, qb, nbA, same ? nbA : nbB, Mp,
                                                                  same ? Mp : Np, nch);
        }
        // algorithmic INT8 operations: 2 per multiply-add, S(S+1)/2 slice pairs per mp multiply-add
        if (prof) prof->end(st, ev, longk ? 3 : 0, gemm_outputs(g) * (double)g.K * 2.0 * (S * (S + 1) / 2));
        return hipGetLastError() == hipSuccess ? 0 : -2;
    }
    int run(const GemmArgs &g, hipStream_t st, int slices, int keep = 0) {
        if constexpr (L > OZ_LMAX) {
            (void)g;
            (void)st;
            (void)slices;
            (void)keep;
            return -5;
        } else {
            if (slices <= 0) slices = OzSlices<L>::DEF;
            if (!OzSlices<L>::built(slices)) return -1;
            if (slices == 4 * L) return run_s<4 * L>(g, st, keep);
            if (slices == 4 * L - 4) return run_s<4 * L - 4>(g, st, keep);
            if constexpr (4 * L - 8 >= 8)
                if (slices == 4 * L - 8) return run_s<4 * L - 8>(g, st, keep);
            if constexpr (4 * L - 12 >= 8)
                if (slices == 4 * L - 12) return run_s<4 * L - 12>(g, st, keep);
            return -1;
        }
    }
};

}  // namespace sdpg
