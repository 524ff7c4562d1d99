// batch.h — the Newton hot path of many small problems in lockstep: one persistent kernel per phase
// (Make_bMat, the Cholesky factorisation of the Schur matrix, the two triangular solves) over a table of
// problem descriptors, each problem's work selected by a request flag in device memory, so that the
// three launches form one hipGraph replayed every IPM iteration.
//
// The problems are contexts of the single-problem engine (engine_impl.h): the kernels read each
// problem's point (X, inv(Z)), its dense constraint blocks and right-hand side g from that context's
// device arrays and write B (then its Cholesky factor, lower triangle) and dy there, so the context
// continues with sdpg_dz / sdpg_dx as after its own sdpg_schur / sdpg_potrf / sdpg_potrs.
//
// Scope: problems whose constraint blocks are all dense (the batch workload: dense N(0,1) constraint
// matrices), evaluated by formula F1 whatever formula the context chose (the formulas differ in cost,
// not in the value of B), at most BMAXB SDP blocks and no LP block, blocks up to n = nmax and m up to
// mmax within the LDS bound of the Schur kernel.
#pragma once

namespace sdpg {

constexpr int BMAXB = 4;        // SDP blocks per problem
constexpr int BATCH_MMAX = 4096;  // constraints per problem
constexpr int BATCH_MMAX_LDS = 256;  // constraints per problem held in LDS by the fixed-point combinations

struct BatchProb {
    int32_t m, nsdp;
    int32_t n[BMAXB];
    int64_t blkoff[BMAXB];
    const Desc *desc;  // m x nsdp constraint descriptors (index i * nsdp + l)
    PV X, invZ, dval, B, g, dy;
    // the context's arrays of the other phases of an iteration: the point and its inverse factors, the
    // residuals, C (dense), b, the right-hand-side matrices of Make_gVec and two block workspaces
    PV Z, y, b, Rp, Rd, C, rzinv, xrd, dls, dX, dZ, icX, icZ, W1, W2;
    int64_t cs_off;  // 2m values of the batch's scratch: r_j = 1/sqrt(c_j d_j), then 1/L_jj
    // block fixed point of the Schur build (batch arrays): the constraint matrices (element, row and
    // matrix offsets of constraint 0; constraint i adds i Nd, i nsum, i nsdp), X by rows and inv(Z) by
    // columns (element and row offsets)
    int64_t fa_el, fa_row, fa_mat, fp_el, fp_row;
    int32_t nsum, Nd, blk0;  // blk0: index of the problem's block 0 in the block task list
    int32_t rowoff[BMAXB];  // first row of block l among the problem's rows (sum of the sizes before it)
};

// a block task: SDP block l of problem p
struct BatchBlk {
    int32_t p, l;
};

// barrier after global writes that other threads of the workgroup read next
__device__ __forceinline__ void wg_sync() {
    __threadfence_block();
    __syncthreads();
}

// request flags: 0 nothing, 1 PREDICTOR (Schur build, Cholesky, solve), 2 CORRECTOR (solve only, with the
// factor of the last predictor; Newton::Mehrotra, sdpa_newton.cpp:1375-1414)
enum { BATCH_NONE = 0, BATCH_PREDICTOR = 1, BATCH_CORRECTOR = 2 };

// Make_bMat in the plain limb form (compute_bMat_dense_SDP with formula F1, sdpa_newton.cpp:1007-1096): one
// constraint i of one problem per task; for every block l, work1 = A_i inv(Z), work2 = X work1
// (sdpa_linear.cpp:856-955) and B_ij += A_j . work2 for the partners j <= i of the block (the pair rule
// for equal NonZeroEffect, sdpa_newton.cpp:1052-1054).  Every sum of products is exact: the operands are
// in block fixed point (offset binary, W = 32L - 2 fraction bits under one exponent per row, column or
// matrix), the limb products accumulate in 96-bit columns and one rounding gives each result.  A_i is
// fixed by rows (work1) and by matrix (inner products) once for the batch, X by rows and inv(Z) by
// columns per predictor (k_fx_point), work1 by columns and work2 by matrix in LDS.  Tasks are dealt
// round-robin to a grid that fills the device.
// Memory: work1 and work2 as limb floats in a global scratch slice of the workgroup (ws, nmax^2 values per
// workgroup, L2-resident); in LDS the fixed-point work1 or work2 (planes of nmax^2), the column
// exponents and offset sums of work1, the exponent and offset sum of work2 and the row of B (71 KB at
// n = 40, m = 120, L = 10: two workgroups of 512 per CU).  The 8 lanes of a partner add their column
// sums by cross-lane shuffles.
constexpr int BS_THREADS = 512;
constexpr int SCHUR_PK = 2;  // packed work2 values per thread: n(n+1)/2 <= 820 <= 2 x 512 for n <= 40
constexpr int FX_HEAD = 2;

template <int L>
struct FxCfg {
    static constexpr int S0 = (L >= 2) ? L - 2 : 0, NC = 2 * L - 1 - S0, W = 32 * L - FX_HEAD;
};

template <int L>
__device__ __forceinline__ void fx_sum_add(uint32_t (&s)[L + 1], const uint32_t (&a)[L]) {
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < L; q++) {
        const uint64_t t = (uint64_t)s[q] + a[q] + c;
        s[q] = (uint32_t)t;
        c = (uint32_t)(t >> 32);
    }
    s[L] += c;
}
template <int L>
__device__ __forceinline__ void fx_sum_add2(uint32_t (&s)[L + 1], const uint32_t *a) {
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q <= L; q++) {
        const uint64_t t = (uint64_t)s[q] + a[q] + c;
        s[q] = (uint32_t)t;
        c = (uint32_t)(t >> 32);
    }
}
// sum over k = k0, k0 + step, ... < K of the limb products of the operands ld(k, a, b) delivers, into
// the column accumulators: the operands of the next term are requested before the current term's
// multiply-adds (two register sets, no copies), so a load's latency overlaps a term's arithmetic
template <int L, class LD>
__device__ __forceinline__ void fx_mac_pipelined(int K, int step, int k0, LD &&ld,
                                                 uint64_t (&lo)[FxCfg<L>::NC], uint32_t (&hi)[FxCfg<L>::NC]) {
    constexpr int S0 = FxCfg<L>::S0;
    if (k0 >= K) return;
    uint32_t a0[L], b0[L], a1[L], b1[L];
    ld(k0, a0, b0);
    int k = k0;
    for (; k + step < K; k += 2 * step) {
        ld(k + step, a1, b1);
        mpx::col_products<L, S0>(a0, b0, lo, hi);
        if (k + 2 * step < K) ld(k + 2 * step, a0, b0);
        mpx::col_products<L, S0>(a1, b1, lo, hi);
    }
    if (k < K) mpx::col_products<L, S0>(a0, b0, lo, hi);
}

// exponent of a block's product: the sum of the operand exponents, 0 when either side is all zero
__device__ __forceinline__ int32_t fx_exp(int32_t ea, int32_t eb, int W) {
    const int32_t e = ea + eb;
    return (e < -(1 << 28) ? 0 : e) - 2 * W;
}

// the batch's block fixed-point arrays that the phases other than the Schur build read
struct FxArrays {
    const uint32_t *fa_g;   // constraint matrices by matrix (L words per element)
    const int32_t *fa_ge;   // their exponents (per constraint and block)
    const uint32_t *fa_gs;  // their offset sums (L + 1 words)
    const uint32_t *fa_cs;  // per element of a problem's blocks: offset sum over the constraints (L + 1)
    uint32_t *fd_v;         // a dense space (X or DLS) by block: fixed values
    int32_t *fd_e;          // its exponents (per block task)
    uint32_t *fd_s;         // its offset sums (per block task, L + 1)
};

// the 64 lanes of a wave add their 96-bit column sums (every lane ends with the total)
template <int L>
__device__ __forceinline__ void fx_wave_sum(uint64_t (&lo)[FxCfg<L>::NC], uint32_t (&hi)[FxCfg<L>::NC]) {
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) {
#pragma unroll
        for (int z = 0; z < FxCfg<L>::NC; z++) {
            const uint32_t o0 = __shfl_xor((uint32_t)lo[z], h, 64);
            const uint32_t o1 = __shfl_xor((uint32_t)(lo[z] >> 32), h, 64);
            const uint32_t oh = __shfl_xor(hi[z], h, 64);
            const uint64_t v = (uint64_t)o0 | ((uint64_t)o1 << 32);
            const uint64_t sm = lo[z] + v;
            hi[z] += oh + (sm < v ? 1u : 0u);
            lo[z] = sm;
        }
    }
}

// sum_i s_i A_i over the m constraints for block l of problem P, exactly in block fixed point: w_i =
// s_i 2^(exponent of A_i's block) goes to fixed point under one exponent (LDS sw: m x L words, planar),
// each element sums w_i a_i[e] over the constraints in 96-bit columns and is rounded once; f(e, value)
// receives the elements.  sred: L + 1 planes of blockDim.x words; se: blockDim.x ints.
template <int L, class F>
__device__ void wg_fx_combo(const BatchProb &P, int l, PV sv, const FxArrays &fx, uint32_t *sw, uint32_t *sred,
                            int32_t *se, F &&f) {
    using C = FxCfg<L>;
    const int tid = threadIdx.x, nt = blockDim.x, m = P.m, nsdp = P.nsdp, n = P.n[l], nn = n * n;
    auto wv = [&](int i) -> mp<L> {
        mp<L> x = ldm<L>(sv, i);
        const int32_t ge = fx.fa_ge[P.fa_mat + (int64_t)i * nsdp + l];
        if (is_zero<L>(x) || ge == EZERO) return zero<L>();
        x.e += ge;
        return x;
    };
    int32_t e = EZERO;
    for (int i = tid; i < m; i += nt) {
        const mp<L> x = wv(i);
        if (!is_zero<L>(x) && x.e > e) e = x.e;
    }
    se[tid] = e;
    __syncthreads();
    for (int h = nt / 2; h >= 1; h >>= 1) {
        if (tid < h) se[tid] = max(se[tid], se[tid + h]);
        __syncthreads();
    }
    const int32_t ew = se[0];
    uint32_t sm[L + 1];
#pragma unroll
    for (int q = 0; q <= L; q++) sm[q] = 0;
    for (int i = tid; i < m; i += nt) {
        uint32_t a[L];
        fx_convert<L>(wv(i), ew, FX_HEAD, a);
#pragma unroll
        for (int q = 0; q < L; q++) sw[q * m + i] = a[q];
        fx_sum_add<L>(sm, a);
    }
#pragma unroll
    for (int q = 0; q <= L; q++) sred[q * nt + tid] = sm[q];
    __syncthreads();
    for (int h = nt / 2; h >= 1; h >>= 1) {
        if (tid < h) {
            uint32_t a[L + 1], b[L + 1];
#pragma unroll
            for (int q = 0; q <= L; q++) a[q] = sred[q * nt + tid], b[q] = sred[q * nt + tid + h];
            fx_sum_add2<L>(a, b);
#pragma unroll
            for (int q = 0; q <= L; q++) sred[q * nt + tid] = a[q];
        }
        __syncthreads();
    }
    uint32_t ra[L + 1];
#pragma unroll
    for (int q = 0; q <= L; q++) ra[q] = sred[q * nt];
    const int32_t E = fx_exp(ew, 0, C::W);
    const int64_t el0 = P.fa_el + P.blkoff[l];
    for (int o = tid; o < nn; o += nt) {
        uint64_t lo[C::NC];
        uint32_t hi[C::NC];
#pragma unroll
        for (int z = 0; z < C::NC; z++) lo[z] = 0, hi[z] = 0;
        auto ld = [&](int i, uint32_t (&a)[L], uint32_t (&b)[L]) {
            const uint32_t *pb = fx.fa_g + (el0 + (int64_t)i * P.Nd + o) * L;
#pragma unroll
            for (int q = 0; q < L; q++) a[q] = sw[q * m + i], b[q] = pb[q];
        };
        fx_mac_pipelined<L>(m, 1, 0, ld, lo, hi);
        uint32_t rb[L + 1];
#pragma unroll
        for (int q = 0; q <= L; q++) rb[q] = fx.fa_cs[(P.fp_el + P.blkoff[l] + o) * (L + 1) + q];
        f(o, fx_finish<L>(lo, hi, m, ra, rb, E));
    }
    wg_sync();
}

template <int L>
__global__ __launch_bounds__(BS_THREADS) void k_batch_schur(const BatchProb *__restrict__ probs,
                                                            const int2 *__restrict__ tasks, int ntasks,
                                                            const int32_t *__restrict__ req, int nmax, int mmax,
                                                            const uint32_t *__restrict__ fa_r, const int32_t *__restrict__ fa_re,
                                                            const uint32_t *__restrict__ fa_rs, const uint32_t *__restrict__ fa_g,
                                                            const int32_t *__restrict__ fa_ge, const uint32_t *__restrict__ fa_ps,
                                                            const uint32_t *__restrict__ fx_v, const int32_t *__restrict__ fx_e,
                                                            const uint32_t *__restrict__ fx_s, const uint32_t *__restrict__ fz_v,
                                                            const int32_t *__restrict__ fz_e, const uint32_t *__restrict__ fz_s,
                                                            PV ws) {
    using C = FxCfg<L>;
    constexpr int NC = C::NC, W = C::W;
    extern __shared__ uint32_t shm[];
    const int PW = nmax * nmax;
    uint32_t *R2 = shm, *sFs = R2 + PW * L, *sGs = sFs + nmax * (L + 1), *sRed = sGs + (L + 1),
             *sB = sRed + 64 * (L + 1);
    int32_t *sFe = (int32_t *)(sB + mmax * (L + 1)), *sGe = sFe + nmax, *sMx = sGe + 1, *sIdx = sMx + BS_THREADS / 64;
    const int64_t wb = (int64_t)blockIdx.x * PW;  // this workgroup's work1 / work2 limb floats in ws
    const int tid = threadIdx.x, nt = BS_THREADS;
    const int q8 = tid & 7, slot = tid >> 3, nslot = BS_THREADS / 8;
    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const int2 tk = tasks[t];
        if (tk.x < 0 || req[tk.x] != BATCH_PREDICTOR) continue;  // the whole workgroup skips together
        const BatchProb &P = probs[tk.x];
        const int i = tk.y, m = P.m, nsdp = P.nsdp;
        for (int j = tid; j <= i; j += nt) lds_st<L>(sB, mmax, j, zero<L>());
        for (int l = 0; l < nsdp; l++) {
            if (P.desc[(int64_t)i * nsdp + l].kind == 0) continue;
            const int n = P.n[l], nn = n * n;
            const int64_t ael = P.fa_el + (int64_t)i * P.Nd + P.blkoff[l];  // A_i's block elements
            const int64_t arw = P.fa_row + (int64_t)i * P.nsum + P.rowoff[l];
            const int64_t pel = P.fp_el + P.blkoff[l], prw = P.fp_row + P.rowoff[l];
            __syncthreads();
            // work1 = A_i inv(Z)
            for (int o = tid; o < nn; o += nt) {
                const int r = o % n, c = o / n;
                uint64_t lo[NC];
                uint32_t hi[NC];
#pragma unroll
                for (int z = 0; z < NC; z++) lo[z] = 0, hi[z] = 0;
                auto ld = [&](int k, uint32_t (&a)[L], uint32_t (&b)[L]) {
                    const uint32_t *pa = fa_r + (ael + r + (int64_t)n * k) * L, *pb = fz_v + (pel + k + (int64_t)n * c) * L;
#pragma unroll
                    for (int q = 0; q < L; q++) a[q] = pa[q], b[q] = pb[q];
                };
                fx_mac_pipelined<L>(n, 1, 0, ld, lo, hi);
                uint32_t ra[L + 1], rb[L + 1];
#pragma unroll
                for (int q = 0; q <= L; q++) ra[q] = fa_rs[(arw + r) * (L + 1) + q], rb[q] = fz_s[(prw + c) * (L + 1) + q];
                stm<L>(ws, wb + o, fx_finish<L>(lo, hi, n, ra, rb, fx_exp(fa_re[arw + r], fz_e[prw + c], W)));
            }
            __syncthreads();
            // work1 by columns: exponents, fixed point, offset sums
            for (int c = tid; c < n; c += nt) {
                int32_t e = EZERO;
                for (int r = 0; r < n; r++) {
                    const mp<L> v = ldm<L>(ws, wb + r + n * c);
                    if (!is_zero<L>(v) && v.e > e) e = v.e;
                }
                sFe[c] = e;
            }
            __syncthreads();
            for (int o = tid; o < nn; o += nt) {
                uint32_t a[L];
                fx_convert<L>(ldm<L>(ws, wb + o), sFe[o / n], FX_HEAD, a);
#pragma unroll
                for (int q = 0; q < L; q++) R2[q * PW + o] = a[q];
            }
            __syncthreads();
            for (int c = tid; c < n; c += nt) {
                uint32_t sm[L + 1];
#pragma unroll
                for (int q = 0; q <= L; q++) sm[q] = 0;
                for (int r = 0; r < n; r++) {
                    uint32_t a[L];
#pragma unroll
                    for (int q = 0; q < L; q++) a[q] = R2[q * PW + r + n * c];
                    fx_sum_add<L>(sm, a);
                }
#pragma unroll
                for (int q = 0; q <= L; q++) sFs[c * (L + 1) + q] = sm[q];
            }
            __syncthreads();
            // work2 = X work1
            for (int o = tid; o < nn; o += nt) {
                const int r = o % n, c = o / n;
                uint64_t lo[NC];
                uint32_t hi[NC];
#pragma unroll
                for (int z = 0; z < NC; z++) lo[z] = 0, hi[z] = 0;
                auto ld = [&](int k, uint32_t (&a)[L], uint32_t (&b)[L]) {
                    const uint32_t *pa = fx_v + (pel + r + (int64_t)n * k) * L;
#pragma unroll
                    for (int q = 0; q < L; q++) a[q] = pa[q], b[q] = R2[q * PW + k + n * c];
                };
                fx_mac_pipelined<L>(n, 1, 0, ld, lo, hi);
                uint32_t ra[L + 1], rb[L + 1];
#pragma unroll
                for (int q = 0; q <= L; q++) ra[q] = fx_s[(prw + r) * (L + 1) + q], rb[q] = sFs[c * (L + 1) + q];
                stm<L>(ws, wb + o, fx_finish<L>(lo, hi, n, ra, rb, fx_exp(fx_e[prw + r], sFe[c], W)));
            }
            __syncthreads();
            // work2 + work2^T on and below the diagonal, packed column by column (t -> element a + n b, a >= b):
            // for symmetric A_j, A_j . work2 = sum over the lower triangle of A_j[a,b] (work2[a,b] + work2[b,a])
            // with the diagonal once; then by matrix: exponent, fixed point, offset sum
            const int np = n * (n + 1) / 2;
            for (int b = tid; b < n; b += nt) {
                const int base = b * n - b * (b - 1) / 2;
                for (int a = b; a < n; a++) sIdx[base + a - b] = a + n * b;
            }
            __syncthreads();
            mp<L> pv[SCHUR_PK];
            int32_t e = EZERO;
#pragma unroll
            for (int u = 0; u < SCHUR_PK; u++) {
                const int t = tid + u * nt;
                pv[u] = zero<L>();
                if (t < np) {
                    const int x = sIdx[t], a = x % n, b = x / n;
                    pv[u] = a == b ? ldm<L>(ws, wb + x) : add<L>(ldm<L>(ws, wb + x), ldm<L>(ws, wb + b + (int64_t)n * a));
                    if (!is_zero<L>(pv[u]) && pv[u].e > e) e = pv[u].e;
                }
            }
            for (int h = 32; h >= 1; h >>= 1) e = max(e, __shfl_xor(e, h, 64));
            if ((tid & 63) == 0) sMx[tid >> 6] = e;
            __syncthreads();
            if (tid == 0) {
                int32_t m8 = sMx[0];
                for (int w = 1; w < BS_THREADS / 64; w++) m8 = max(m8, sMx[w]);
                sGe[0] = m8;
            }
            __syncthreads();
            const int32_t eg = sGe[0];
#pragma unroll
            for (int u = 0; u < SCHUR_PK; u++) {
                const int t = tid + u * nt;
                if (t < np) {
                    uint32_t a[L];
                    fx_convert<L>(pv[u], eg, FX_HEAD, a);
#pragma unroll
                    for (int q = 0; q < L; q++) R2[q * PW + t] = a[q];
                }
            }
            __syncthreads();
            if (tid < 64) {
                uint32_t sm[L + 1];
#pragma unroll
                for (int q = 0; q <= L; q++) sm[q] = 0;
                for (int o = tid; o < np; o += 64) {
                    uint32_t a[L];
#pragma unroll
                    for (int q = 0; q < L; q++) a[q] = R2[q * PW + o];
                    fx_sum_add<L>(sm, a);
                }
                // lanes 0..63 -> sRed rows, then lane 0 adds them up
#pragma unroll
                for (int q = 0; q <= L; q++) sRed[tid * (L + 1) + q] = sm[q];
            }
            __syncthreads();
            if (tid == 0) {
                uint32_t sm[L + 1];
#pragma unroll
                for (int q = 0; q <= L; q++) sm[q] = sRed[q];
                for (int u = 1; u < 64; u++) fx_sum_add2<L>(sm, sRed + u * (L + 1));
#pragma unroll
                for (int q = 0; q <= L; q++) sGs[q] = sm[q];
            }
            __syncthreads();
            // B_ij += A_j . work2 for the partners j <= i: 8 threads per partner over the packed entries
            uint32_t gsum[L + 1];
#pragma unroll
            for (int q = 0; q <= L; q++) gsum[q] = sGs[q];
            for (int j0 = 0; j0 <= i; j0 += nslot) {
                const int j = j0 + slot;
                const bool on = j <= i && P.desc[(int64_t)j * nsdp + l].kind != 0;
                uint64_t lo[NC];
                uint32_t hi[NC];
#pragma unroll
                for (int z = 0; z < NC; z++) lo[z] = 0, hi[z] = 0;
                if (on) {
                    const int64_t jel = P.fa_el + (int64_t)j * P.Nd + P.blkoff[l];
                    auto ld = [&](int e, uint32_t (&a)[L], uint32_t (&b)[L]) {
                        const uint32_t *pa = fa_g + (jel + sIdx[e]) * L;
#pragma unroll
                        for (int q = 0; q < L; q++) a[q] = pa[q], b[q] = R2[q * PW + e];
                    };
                    fx_mac_pipelined<L>(np, 8, q8, ld, lo, hi);
                }
                // the 8 lanes of a partner (consecutive lanes of one wave) add their column sums
#pragma unroll
                for (int h = 1; h < 8; h <<= 1) {
#pragma unroll
                    for (int z = 0; z < NC; z++) {
                        const uint32_t o0 = __shfl_xor((uint32_t)lo[z], h, 64);
                        const uint32_t o1 = __shfl_xor((uint32_t)(lo[z] >> 32), h, 64);
                        const uint32_t oh = __shfl_xor(hi[z], h, 64);
                        const uint64_t v = (uint64_t)o0 | ((uint64_t)o1 << 32);
                        const uint64_t sm = lo[z] + v;
                        hi[z] += oh + (sm < v ? 1u : 0u);
                        lo[z] = sm;
                    }
                }
                if (q8 == 0 && on) {
                    const int64_t jm = P.fa_mat + (int64_t)j * nsdp + l;
                    uint32_t ra[L + 1];
#pragma unroll
                    for (int q = 0; q <= L; q++) ra[q] = fa_ps[jm * (L + 1) + q];
                    const mp<L> v = fx_finish<L>(lo, hi, np, ra, gsum, fx_exp(fa_ge[jm], eg, W));
                    lds_st<L>(sB, mmax, j, add<L>(lds_ld<L>(sB, mmax, j), v));
                }
            }
        }
        __syncthreads();
        for (int j = tid; j <= i; j += nt) {
            const mp<L> v = lds_ld<L>(sB, mmax, j);
            stm<L>(P.B, i + (int64_t)m * j, v);
            stm<L>(P.B, j + (int64_t)m * i, v);
        }
        __syncthreads();
    }
}

// LDS words of k_batch_schur
template <int L>
inline size_t schur_lds_words(int nmax, int mmax) {
    const size_t PW = (size_t)nmax * nmax;
    return PW * L + (size_t)nmax * (L + 1) + (L + 1) + 64 * (L + 1) + (size_t)mmax * (L + 1) + nmax + 1 +
           BS_THREADS / 64 + (size_t)nmax * (nmax + 1) / 2;
}

// the constraint matrices in block fixed point, once per batch: per task (p, i) and block, by rows
// (exponent, fixed value and offset sum per row) and by matrix (one exponent, fixed values, offset sum)
template <int L>
__global__ __launch_bounds__(256) void k_fx_constraints(const BatchProb *__restrict__ probs,
                                                        const int2 *__restrict__ tasks, int ntasks, uint32_t *fa_r,
                                                        int32_t *fa_re, uint32_t *fa_rs, uint32_t *fa_g,
                                                        int32_t *fa_ge, uint32_t *fa_gs, uint32_t *fa_ps) {
    __shared__ int32_t se[256];
    __shared__ uint32_t ssum[256 * (L + 1)], psum[256 * (L + 1)];
    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const int2 tk = tasks[t];
        if (tk.x < 0) continue;
        const BatchProb &P = probs[tk.x];
        const int i = tk.y;
        for (int l = 0; l < P.nsdp; l++) {
            const Desc d = P.desc[(int64_t)i * P.nsdp + l];
            const int n = P.n[l], nn = n * n;
            const int64_t el = P.fa_el + (int64_t)i * P.Nd + P.blkoff[l];
            const int64_t rw = P.fa_row + (int64_t)i * P.nsum + P.rowoff[l];
            const int64_t mt = P.fa_mat + (int64_t)i * P.nsdp + l;
            auto val = [&](int r, int k) -> mp<L> {
                return d.kind == 2 ? ldm<L>(P.dval, d.off + r + (int64_t)n * k) : zero<L>();
            };
            // by rows
            for (int r = threadIdx.x; r < n; r += blockDim.x) {
                int32_t e = EZERO;
                for (int k = 0; k < n; k++) {
                    const mp<L> v = val(r, k);
                    if (!is_zero<L>(v) && v.e > e) e = v.e;
                }
                uint32_t sm[L + 1];
#pragma unroll
                for (int q = 0; q <= L; q++) sm[q] = 0;
                for (int k = 0; k < n; k++) {
                    uint32_t a[L];
                    fx_convert<L>(val(r, k), e, FX_HEAD, a);
#pragma unroll
                    for (int q = 0; q < L; q++) fa_r[(el + r + (int64_t)n * k) * L + q] = a[q];
                    fx_sum_add<L>(sm, a);
                }
                fa_re[rw + r] = e;
#pragma unroll
                for (int q = 0; q <= L; q++) fa_rs[(rw + r) * (L + 1) + q] = sm[q];
            }
            // by matrix
            int32_t e = EZERO;
            for (int o = threadIdx.x; o < nn; o += blockDim.x) {
                const mp<L> v = val(o % n, o / n);
                if (!is_zero<L>(v) && v.e > e) e = v.e;
            }
            se[threadIdx.x] = e;
            __syncthreads();
            for (int h = blockDim.x / 2; h >= 1; h >>= 1) {
                if ((int)threadIdx.x < h) se[threadIdx.x] = max(se[threadIdx.x], se[threadIdx.x + h]);
                __syncthreads();
            }
            e = se[0];
            uint32_t sm[L + 1], sp[L + 1];  // all elements; the lower triangle with the diagonal
#pragma unroll
            for (int q = 0; q <= L; q++) sm[q] = 0, sp[q] = 0;
            for (int o = threadIdx.x; o < nn; o += blockDim.x) {
                uint32_t a[L];
                fx_convert<L>(val(o % n, o / n), e, FX_HEAD, a);
#pragma unroll
                for (int q = 0; q < L; q++) fa_g[(el + o) * L + q] = a[q];
                fx_sum_add<L>(sm, a);
                if (o % n >= o / n) fx_sum_add<L>(sp, a);
            }
#pragma unroll
            for (int q = 0; q <= L; q++) ssum[threadIdx.x * (L + 1) + q] = sm[q], psum[threadIdx.x * (L + 1) + q] = sp[q];
            __syncthreads();
            if (threadIdx.x == 0) {
                for (int u = 1; u < (int)blockDim.x; u++) {
                    fx_sum_add2<L>(sm, ssum + u * (L + 1));
                    fx_sum_add2<L>(sp, psum + u * (L + 1));
                }
                fa_ge[mt] = e;
#pragma unroll
                for (int q = 0; q <= L; q++) fa_gs[mt * (L + 1) + q] = sm[q], fa_ps[mt * (L + 1) + q] = sp[q];
            }
            __syncthreads();
        }
    }
}

// X by rows and inv(Z) by columns in block fixed point, per predictor (one workgroup per block task)
template <int L>
__global__ __launch_bounds__(256) void k_fx_point(const BatchProb *__restrict__ probs, const BatchBlk *__restrict__ blks,
                                                  int nblk, const int32_t *__restrict__ req, uint32_t *fx_v,
                                                  int32_t *fx_e, uint32_t *fx_s, uint32_t *fz_v, int32_t *fz_e,
                                                  uint32_t *fz_s) {
    for (int t = blockIdx.x; t < nblk; t += gridDim.x) {
        const BatchBlk bk = blks[t];
        if (req[bk.p] != BATCH_PREDICTOR) continue;
        const BatchProb &P = probs[bk.p];
        const int n = P.n[bk.l];
        const int64_t o = P.blkoff[bk.l], el = P.fp_el + o, rw = P.fp_row + P.rowoff[bk.l];
        for (int u = threadIdx.x; u < 2 * n; u += blockDim.x) {
            const bool isx = u < n;
            const int r = isx ? u : u - n;  // row of X, or column of inv(Z)
            auto at = [&](int k) -> int64_t { return isx ? r + (int64_t)n * k : k + (int64_t)n * r; };
            const PV S = isx ? P.X : P.invZ;
            int32_t e = EZERO;
            for (int k = 0; k < n; k++) {
                const mp<L> v = ldm<L>(S, o + at(k));
                if (!is_zero<L>(v) && v.e > e) e = v.e;
            }
            uint32_t sm[L + 1];
#pragma unroll
            for (int q = 0; q <= L; q++) sm[q] = 0;
            uint32_t *dst = isx ? fx_v : fz_v;
            for (int k = 0; k < n; k++) {
                uint32_t a[L];
                fx_convert<L>(ldm<L>(S, o + at(k)), e, FX_HEAD, a);
#pragma unroll
                for (int q = 0; q < L; q++) dst[(el + at(k)) * L + q] = a[q];
                fx_sum_add<L>(sm, a);
            }
            (isx ? fx_e : fz_e)[rw + r] = e;
            uint32_t *ds = isx ? fx_s : fz_s;
#pragma unroll
            for (int q = 0; q <= L; q++) ds[(rw + r) * (L + 1) + q] = sm[q];
        }
    }
}

// Rpotrf("Lower") of one problem's Schur matrix per workgroup (choleskyFactorWithAdjust,
// sdpa_linear.cpp:608-631), unblocked and division-free (the Bareiss form of L D L^T, in mp arithmetic with
// one rounding per update): with c_j > 0 a running scale, At = c_j S^(j) and step j updates every remaining entry of the lower
// triangle at once, At_ic <- 2^-e_j (d_j At_ic - At_ij At_cj) (d_j = At_jj, e_j its exponent); the true
// pivot is d_j / c_j and c_{j+1} = 2^-e_j c_j d_j.  After the chain, r_j = 1/sqrt(c_j d_j) gives
// L_ic = At_ic r_c (so L_jj = d_j r_j) and 1/L_jj = c_j r_j.  info = the first non-positive pivot
// (1-based, as Rpotrf2's info), and the problem's solve is skipped.
constexpr int BPOTRF_THREADS = 1024;
template <int L>
__global__ __launch_bounds__(BPOTRF_THREADS) void k_batch_potrf(const BatchProb *__restrict__ probs, int nprob,
                                                     const int32_t *__restrict__ req, int32_t *__restrict__ info,
                                                     PV cs) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int p = blockIdx.x; p < nprob; p += gridDim.x) {
        if (req[p] != BATCH_PREDICTOR) continue;
        const BatchProb &P = probs[p];
        const int m = P.m;
        const PV B = P.B;
        auto at = [&](int i, int j) -> int64_t { return i + (int64_t)m * j; };
        int fail = 0;
        for (int j = 0; j < m; j++) {
            const mp<L> d = ldm<L>(B, at(j, j));  // the same value in every thread: a uniform exit
            if (sgn<L>(d) <= 0) {
                fail = j + 1;
                break;
            }
            const int n1 = m - 1 - j, cnt = n1 * (n1 + 1) / 2;
            for (int t = tid; t < cnt; t += nt) {
                // lower triangle of the trailing block, row-major: t = ip (ip + 1) / 2 + cp, cp <= ip
                int ip = (int)((__builtin_sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
                while (ip > 0 && ip * (ip + 1) / 2 > t) ip--;
                while ((ip + 1) * (ip + 2) / 2 <= t) ip++;
                const int i = j + 1 + ip, c = j + 1 + (t - ip * (ip + 1) / 2);
                mp<L> v = fms<L>(d, ldm<L>(B, at(i, c)), ldm<L>(B, at(i, j)), ldm<L>(B, at(c, j)));
                if (!is_zero<L>(v)) v.e -= d.e;
                stm<L>(B, at(i, c), v);
            }
            __threadfence_block();
            __syncthreads();
        }
        if (tid == 0) info[p] = fail;
        if (fail) continue;
        // running scales c_j = prod_{k<j} 2^-e_k d_k by a doubling scan over the two halves of the
        // problem's scratch (f_k first, then partial products), then r_j and 1/L_jj in parallel
        int cur = 0;
        for (int j = tid; j < m; j += nt) {
            mp<L> d = ldm<L>(B, at(j, j));
            d.e = 0;
            stm<L>(cs, P.cs_off + j, d);
        }
        __threadfence_block();
        __syncthreads();
        for (int s = 1; s < m; s <<= 1) {
            for (int j = tid; j < m; j += nt) {
                mp<L> x = ldm<L>(cs, P.cs_off + (int64_t)cur * m + j);
                if (j >= s) x = mul<L>(x, ldm<L>(cs, P.cs_off + (int64_t)cur * m + j - s));
                stm<L>(cs, P.cs_off + (int64_t)(1 - cur) * m + j, x);
            }
            cur = 1 - cur;
            __threadfence_block();
            __syncthreads();
        }
        // c_j = (inclusive product up to j - 1); read every c_j before the scratch is overwritten
        mp<L> cj[(BATCH_MMAX + BPOTRF_THREADS - 1) / BPOTRF_THREADS];
        for (int j = tid, q = 0; j < m; j += nt, q++)
            cj[q] = j == 0 ? from_int<L>(1) : ldm<L>(cs, P.cs_off + (int64_t)cur * m + j - 1);
        __syncthreads();
        for (int j = tid, q = 0; j < m; j += nt, q++) {
            const mp<L> c = cj[q], d = ldm<L>(B, at(j, j));
            const mp<L> r = mpx::rsqrt_fx<L>(mul<L>(c, d));
            stm<L>(cs, P.cs_off + m + j, mul<L>(c, r));  // 1 / L_jj
            stm<L>(cs, P.cs_off + j, r);
        }
        __threadfence_block();
        __syncthreads();
        const int tot = m * (m + 1) / 2;
        for (int t = tid; t < tot; t += nt) {
            // column-major lower triangle: column c from the sizes of the columns before it
            int c = (int)(((2.0f * m + 1.0f) - __builtin_sqrtf((2.0f * m + 1.0f) * (2.0f * m + 1.0f) - 8.0f * t)) * 0.5f);
            auto first = [&](int cc) { return cc * m - cc * (cc - 1) / 2; };
            while (c > 0 && first(c) > t) c--;
            while (c + 1 < m && first(c + 1) <= t) c++;
            const int i = c + (t - first(c));
            stm<L>(B, at(i, c), mul<L>(ldm<L>(B, at(i, c)), ldm<L>(cs, P.cs_off + c)));
        }
        __syncthreads();
    }
}

// Lal::solveSystems (sdpa_linear.cpp:683-695) of one problem per workgroup: L y = g then L^T dy = y by
// columns (y_i = y_i / L_ii, then every later entry loses L_ki y_i), with the stored 1/L_ii; skipped for
// a problem whose factorisation failed.
template <int L>
__global__ __launch_bounds__(256) void k_batch_potrs(const BatchProb *__restrict__ probs, int nprob,
                                                     const int32_t *__restrict__ req,
                                                     const int32_t *__restrict__ info, PV cs, int mmax) {
    extern __shared__ uint32_t shm[];
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int p = blockIdx.x; p < nprob; p += gridDim.x) {
        if (req[p] == BATCH_NONE || info[p] != 0) continue;
        const BatchProb &P = probs[p];
        const int m = P.m;
        auto at = [&](int i, int j) -> int64_t { return i + (int64_t)m * j; };
        for (int i = tid; i < m; i += nt) lds_st<L>(shm, mmax, i, ldm<L>(P.g, i));
        __syncthreads();
        for (int i = 0; i < m; i++) {
            const mp<L> yi = mul<L>(lds_ld<L>(shm, mmax, i), ldm<L>(cs, P.cs_off + m + i));
            __syncthreads();
            if (tid == 0) lds_st<L>(shm, mmax, i, yi);
            for (int k = i + 1 + tid; k < m; k += nt)
                lds_st<L>(shm, mmax, k, sub<L>(lds_ld<L>(shm, mmax, k), mul<L>(ldm<L>(P.B, at(k, i)), yi)));
            __syncthreads();
        }
        for (int i = m - 1; i >= 0; i--) {
            const mp<L> yi = mul<L>(lds_ld<L>(shm, mmax, i), ldm<L>(cs, P.cs_off + m + i));
            __syncthreads();
            if (tid == 0) lds_st<L>(shm, mmax, i, yi);
            for (int k = tid; k < i; k += nt)
                lds_st<L>(shm, mmax, k, sub<L>(lds_ld<L>(shm, mmax, k), mul<L>(ldm<L>(P.B, at(i, k)), yi)));
            __syncthreads();
        }
        for (int i = tid; i < m; i += nt) stm<L>(P.dy, i, lds_ld<L>(shm, mmax, i));
        __syncthreads();
    }
}

// ---------------------------------------------------------------- the other phases of an iteration
//
// One workgroup per SDP block of a problem (a "block task"); n <= 40 in the batch workload, so every
// operand of a block stays in the L2-resident context arrays and the products are element-form loops.
// Requests per problem in device memory select the problems of a call; scalars per problem (the
// target beta mu of Make_gVec, the step lengths) come in a planar array sc (BSC_N values per problem)
// and results go out per block task in a planar array bo (BO_N values per task).

enum { BSC_TARGET = 0, BSC_ALPHAP = 1, BSC_ALPHAD = 2, BSC_N = 3 };
enum { BO_XZ, BO_XDZ, BO_DXZ, BO_DXDZ, BO_TRX, BO_TRZ, BO_CX, BO_CDX, BO_BY, BO_BDY, BO_MAXRD, BO_MAXRP, BO_EIG, BO_N };

// C = op(A) op(B) for one n x n block in block fixed point: op(A) by rows and op(B) by columns (one
// exponent each, offset binary), exact 96-bit column sums over k and one rounding per output.  The
// fixed operands, offset sums and exponents go to the workgroup's global scratch slice scr
// (fx_mm_words(n) words, L2-resident).  The caller synchronises before C is read by other threads.
template <int L>
__host__ __device__ constexpr int64_t fx_mm_words(int n) {
    return 2 * (int64_t)n * n * L + 2 * (int64_t)n * (L + 1) + 2 * (int64_t)n;
}
template <int L>
__device__ void wg_mm_fx(int n, PV A, int64_t ao, bool ta, PV B, int64_t bo, bool tb, PV C, int64_t co, uint32_t *scr) {
    using Cf = FxCfg<L>;
    uint32_t *Av = scr, *Bv = Av + (int64_t)n * n * L, *As = Bv + (int64_t)n * n * L, *Bs = As + n * (L + 1);
    int32_t *Ae = (int32_t *)(Bs + n * (L + 1)), *Be = Ae + n;
    const int tid = threadIdx.x, nt = blockDim.x, nn = n * n;
    for (int u = tid; u < 2 * n; u += nt) {
        const bool isA = u < n;
        const int r = isA ? u : u - n;  // row of op(A), or column of op(B)
        // op(A)[r, k] = ta ? A[k, r] : A[r, k]; op(B)[k, c] = tb ? B[c, k] : B[k, c] (column-major)
        auto at = [&](int k) -> int64_t {
            return isA ? ao + (ta ? k + (int64_t)n * r : r + (int64_t)n * k)
                       : bo + (tb ? r + (int64_t)n * k : k + (int64_t)n * r);
        };
        const PV S = isA ? A : B;
        int32_t e = EZERO;
        for (int k = 0; k < n; k++) {
            const mp<L> v = ldm<L>(S, at(k));
            if (!is_zero<L>(v) && v.e > e) e = v.e;
        }
        uint32_t sm[L + 1];
#pragma unroll
        for (int q = 0; q <= L; q++) sm[q] = 0;
        for (int k = 0; k < n; k++) {
            uint32_t a[L];
            fx_convert<L>(ldm<L>(S, at(k)), e, FX_HEAD, a);
            uint32_t *dst = isA ? Av + (r + (int64_t)n * k) * L : Bv + (k + (int64_t)n * r) * L;
#pragma unroll
            for (int q = 0; q < L; q++) dst[q] = a[q];
            fx_sum_add<L>(sm, a);
        }
        (isA ? Ae : Be)[r] = e;
        uint32_t *ds = (isA ? As : Bs) + r * (L + 1);
#pragma unroll
        for (int q = 0; q <= L; q++) ds[q] = sm[q];
    }
    wg_sync();
    for (int o = tid; o < nn; o += nt) {
        const int r = o % n, c = o / n;
        uint64_t lo[Cf::NC];
        uint32_t hi[Cf::NC];
#pragma unroll
        for (int z = 0; z < Cf::NC; z++) lo[z] = 0, hi[z] = 0;
        auto ld = [&](int k, uint32_t (&a)[L], uint32_t (&b)[L]) {
            const uint32_t *pa = Av + (r + (int64_t)n * k) * L, *pb = Bv + (k + (int64_t)n * c) * L;
#pragma unroll
            for (int q = 0; q < L; q++) a[q] = pa[q], b[q] = pb[q];
        };
        fx_mac_pipelined<L>(n, 1, 0, ld, lo, hi);
        uint32_t ra[L + 1], rb[L + 1];
#pragma unroll
        for (int q = 0; q <= L; q++) ra[q] = As[r * (L + 1) + q], rb[q] = Bs[c * (L + 1) + q];
        stm<L>(C, co + o, fx_finish<L>(lo, hi, n, ra, rb, fx_exp(Ae[r], Be[c], Cf::W)));
    }
    wg_sync();  // the scratch is reused by the next product
}

// workgroup sum / max |.| of one value per thread (blockDim.x a power of two; sh: L+1 planes of blockDim.x)
template <int L>
__device__ mp<L> wg_sum(mp<L> v, uint32_t *sh) {
    const int t = threadIdx.x, nt = blockDim.x;
    __syncthreads();
    lds_st<L>(sh, nt, t, v);
    __syncthreads();
    for (int h = nt / 2; h >= 1; h >>= 1) {
        if (t < h) lds_st<L>(sh, nt, t, add<L>(lds_ld<L>(sh, nt, t), lds_ld<L>(sh, nt, t + h)));
        __syncthreads();
    }
    const mp<L> r = lds_ld<L>(sh, nt, 0);
    __syncthreads();
    return r;
}
template <int L>
__device__ mp<L> wg_maxabs(mp<L> v, uint32_t *sh) {
    const int t = threadIdx.x, nt = blockDim.x;
    __syncthreads();
    lds_st<L>(sh, nt, t, absv<L>(v));
    __syncthreads();
    for (int h = nt / 2; h >= 1; h >>= 1) {
        if (t < h) {
            const mp<L> a = lds_ld<L>(sh, nt, t), b = lds_ld<L>(sh, nt, t + h);
            if (cmp_abs<L>(b, a) > 0) lds_st<L>(sh, nt, t, b);
        }
        __syncthreads();
    }
    const mp<L> r = lds_ld<L>(sh, nt, 0);
    __syncthreads();
    return r;
}

#define BATCH_BLOCK_LOOP                                                 \
    for (int t_ = blockIdx.x; t_ < nblk; t_ += gridDim.x)                \
        if (const BatchBlk bk = blks[t_]; req[bk.p] != 0)

// Make_gVec's matrices for one block (compute_rMat and the infeasible-phase term, sdpa_newton.cpp:821-895):
// r_zinv = (beta mu) inv(Z) - X [- dX dZ inv(Z) for the corrector]; in the infeasible phases (iarg 1)
// the predictor forms x_rd_zinv = X Rd inv(Z) and DLS = r_zinv - x_rd_zinv, otherwise DLS = r_zinv.
// req: 1 predictor, 2 corrector.
template <int L>
__global__ __launch_bounds__(256) void k_bb_gvec(const BatchProb *__restrict__ probs, const BatchBlk *__restrict__ blks,
                                                 int nblk, const int32_t *__restrict__ req,
                                                 const int32_t *__restrict__ iarg, PV sc, uint32_t *fxs, int64_t fxn) {
    uint32_t *scr = fxs + (int64_t)blockIdx.x * fxn;  // this workgroup's scratch of the fixed-point products
    BATCH_BLOCK_LOOP {
        const BatchProb &P = probs[bk.p];
        const int n = P.n[bk.l], nn = n * n;
        const int64_t o = P.blkoff[bk.l];
        const mp<L> tg = ldm<L>(sc, (int64_t)bk.p * BSC_N + BSC_TARGET);
        const bool corr = req[bk.p] == 2, infeas = iarg[bk.p] != 0;
        for (int e = threadIdx.x; e < nn; e += blockDim.x)
            stm<L>(P.rzinv, o + e, add<L>(mul<L>(tg, ldm<L>(P.invZ, o + e)), neg<L>(ldm<L>(P.X, o + e))));
        if (corr) {
            wg_mm_fx<L>(n, P.dX, o, false, P.dZ, o, false, P.W1, o, scr);
            wg_sync();
            wg_mm_fx<L>(n, P.W1, o, false, P.invZ, o, false, P.W2, o, scr);
            wg_sync();
            for (int e = threadIdx.x; e < nn; e += blockDim.x)
                stm<L>(P.rzinv, o + e, sub<L>(ldm<L>(P.rzinv, o + e), ldm<L>(P.W2, o + e)));
        }
        if (infeas && !corr) {
            wg_mm_fx<L>(n, P.X, o, false, P.Rd, o, false, P.W1, o, scr);
            wg_sync();
            wg_mm_fx<L>(n, P.W1, o, false, P.invZ, o, false, P.xrd, o, scr);
            wg_sync();
        }
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const mp<L> r = ldm<L>(P.rzinv, o + e);
            stm<L>(P.dls, o + e, infeas ? sub<L>(r, ldm<L>(P.xrd, o + e)) : r);
        }
        wg_sync();
    }
}

// a dense space by block in block fixed point (mode 0: DLS, mode 1: X), for the constraint-wise inner
// products: one exponent per block, the fixed values and their offset sum
template <int L>
__global__ __launch_bounds__(256) void k_fx_dspace(const BatchProb *__restrict__ probs, const BatchBlk *__restrict__ blks,
                                                   int nblk, const int32_t *__restrict__ req, int mode, FxArrays fx) {
    __shared__ int32_t se[256];
    __shared__ uint32_t sred[(L + 1) * 256];
    for (int t = blockIdx.x; t < nblk; t += gridDim.x) {
        const BatchBlk bk = blks[t];
        if (req[bk.p] == 0) continue;
        const BatchProb &P = probs[bk.p];
        const int n = P.n[bk.l], nn = n * n, tid = threadIdx.x, nt = blockDim.x;
        const int64_t o = P.blkoff[bk.l];
        const PV D = mode == 0 ? P.dls : P.X;
        int32_t e = EZERO;
        for (int u = tid; u < nn; u += nt) {
            const mp<L> v = ldm<L>(D, o + u);
            if (!is_zero<L>(v) && v.e > e) e = v.e;
        }
        se[tid] = e;
        __syncthreads();
        for (int h = nt / 2; h >= 1; h >>= 1) {
            if (tid < h) se[tid] = max(se[tid], se[tid + h]);
            __syncthreads();
        }
        e = se[0];
        uint32_t sm[L + 1];
#pragma unroll
        for (int q = 0; q <= L; q++) sm[q] = 0;
        for (int u = tid; u < nn; u += nt) {
            uint32_t a[L];
            fx_convert<L>(ldm<L>(D, o + u), e, FX_HEAD, a);
#pragma unroll
            for (int q = 0; q < L; q++) fx.fd_v[(P.fp_el + o + u) * L + q] = a[q];
            fx_sum_add<L>(sm, a);
        }
#pragma unroll
        for (int q = 0; q <= L; q++) sred[q * nt + tid] = sm[q];
        __syncthreads();
        for (int h = nt / 2; h >= 1; h >>= 1) {
            if (tid < h) {
                uint32_t a[L + 1], b[L + 1];
#pragma unroll
                for (int q = 0; q <= L; q++) a[q] = sred[q * nt + tid], b[q] = sred[q * nt + tid + h];
                fx_sum_add2<L>(a, b);
#pragma unroll
                for (int q = 0; q <= L; q++) sred[q * nt + tid] = a[q];
            }
            __syncthreads();
        }
        if (tid == 0) {
            fx.fd_e[t] = e;
#pragma unroll
            for (int q = 0; q <= L; q++) fx.fd_s[(int64_t)t * (L + 1) + q] = sred[q * nt];
        }
        __syncthreads();
    }
}

// constraint-wise inner products A_i . D over the dense blocks, one wave per constraint (p, i), exactly in
// block fixed point per block (A_i by matrix, D by block from k_fx_dspace) and one rounding per block:
//   mode 0: g_i = -(A_i . DLS) + Rp_i   (Make_gVec, sdpa_newton.cpp:869-895)
//   mode 1: Rp_i = b_i - A_i . X        (Residuals::compute, sdpa_parts.cpp:760-790)
template <int L>
__global__ __launch_bounds__(256) void k_bc_inner(const BatchProb *__restrict__ probs, const int2 *__restrict__ tasks,
                                                  int ntasks, const int32_t *__restrict__ req, int mode, FxArrays fx) {
    using C = FxCfg<L>;
    const int lane = threadIdx.x & 63;
    for (int t0 = blockIdx.x * 4; t0 < ntasks; t0 += gridDim.x * 4) {
        const int t = t0 + (threadIdx.x >> 6);
        int2 tk = make_int2(0, 0);
        bool on = false;
        if (t < ntasks) {
            tk = tasks[t];
            on = tk.x >= 0 && req[tk.x] != 0;
        }
        if (!on) continue;  // whole waves: the wave's lanes share t
        const BatchProb &P = probs[tk.x];
        mp<L> ip = zero<L>();
        for (int l = 0; l < P.nsdp; l++) {
            const int n = P.n[l], nn = n * n;
            const int64_t jm = P.fa_mat + (int64_t)tk.y * P.nsdp + l, bt = P.blk0 + l;
            const int64_t ael = P.fa_el + (int64_t)tk.y * P.Nd + P.blkoff[l], del = P.fp_el + P.blkoff[l];
            uint64_t lo[C::NC];
            uint32_t hi[C::NC];
#pragma unroll
            for (int z = 0; z < C::NC; z++) lo[z] = 0, hi[z] = 0;
            auto ld = [&](int e, uint32_t (&a)[L], uint32_t (&b)[L]) {
                const uint32_t *pa = fx.fa_g + (ael + e) * L, *pb = fx.fd_v + (del + e) * L;
#pragma unroll
                for (int q = 0; q < L; q++) a[q] = pa[q], b[q] = pb[q];
            };
            fx_mac_pipelined<L>(nn, 64, lane, ld, lo, hi);
            fx_wave_sum<L>(lo, hi);
            uint32_t ra[L + 1], rb[L + 1];
#pragma unroll
            for (int q = 0; q <= L; q++) ra[q] = fx.fa_gs[jm * (L + 1) + q], rb[q] = fx.fd_s[bt * (L + 1) + q];
            ip = add<L>(ip, fx_finish<L>(lo, hi, nn, ra, rb, fx_exp(fx.fa_ge[jm], fx.fd_e[bt], C::W)));
        }
        if (lane == 0) {
            if (mode == 0)
                stm<L>(P.g, tk.y, add<L>(neg<L>(ip), ldm<L>(P.Rp, tk.y)));
            else
                stm<L>(P.Rp, tk.y, sub<L>(ldm<L>(P.b, tk.y), ip));
        }
    }
}

// per element of every problem's blocks: the offset sum over the constraints of the fixed-point
// constraint values (by matrix), once per batch (k_fx_combo's column sums)
template <int L>
__global__ __launch_bounds__(256) void k_fx_colsum(const BatchProb *__restrict__ probs, const BatchBlk *__restrict__ blks,
                                                   int nblk, FxArrays fx, uint32_t *fa_cs) {
    for (int t = blockIdx.x; t < nblk; t += gridDim.x) {
        const BatchBlk bk = blks[t];
        const BatchProb &P = probs[bk.p];
        const int n = P.n[bk.l], nn = n * n;
        for (int o = threadIdx.x; o < nn; o += blockDim.x) {
            uint32_t sm[L + 1];
#pragma unroll
            for (int q = 0; q <= L; q++) sm[q] = 0;
            for (int i = 0; i < P.m; i++) {
                uint32_t a[L];
                const uint32_t *pa = fx.fa_g + (P.fa_el + (int64_t)i * P.Nd + P.blkoff[bk.l] + o) * L;
#pragma unroll
                for (int q = 0; q < L; q++) a[q] = pa[q];
                fx_sum_add<L>(sm, a);
            }
#pragma unroll
            for (int q = 0; q <= L; q++) fa_cs[(P.fp_el + P.blkoff[bk.l] + o) * (L + 1) + q] = sm[q];
        }
    }
}

// dZ = -sum_i dy_i A_i (+ Rd in the infeasible phases, iarg 1) (compute_DzMat, sdpa_newton.cpp:1346-1356)
// (sum_i dy_i A_i in block fixed point, wg_fx_combo)
template <int L>
__global__ __launch_bounds__(256) void k_bb_dz(const BatchProb *__restrict__ probs, const BatchBlk *__restrict__ blks,
                                               int nblk, const int32_t *__restrict__ req,
                                               const int32_t *__restrict__ iarg, FxArrays fx) {
    __shared__ uint32_t sw[L * BATCH_MMAX_LDS], sred[(L + 1) * 256];
    __shared__ int32_t se[256];
    BATCH_BLOCK_LOOP {
        const BatchProb &P = probs[bk.p];
        const int64_t o = P.blkoff[bk.l];
        const bool useR = iarg[bk.p] != 0;
        wg_fx_combo<L>(P, bk.l, P.dy, fx, sw, sred, se, [&](int e, const mp<L> &acc) {
            mp<L> v = neg<L>(acc);
            if (useR) v = add<L>(v, ldm<L>(P.Rd, o + e));
            stm<L>(P.dZ, o + e, v);
        });
        wg_sync();
    }
}

// dX = sym(r_zinv - X dZ inv(Z)) (compute_DxMat incl. getSymmetrize, sdpa_newton.cpp:1358-1373)
template <int L>
__global__ __launch_bounds__(256) void k_bb_dx(const BatchProb *__restrict__ probs, const BatchBlk *__restrict__ blks,
                                               int nblk, const int32_t *__restrict__ req, uint32_t *fxs, int64_t fxn) {
    uint32_t *scr = fxs + (int64_t)blockIdx.x * fxn;
    BATCH_BLOCK_LOOP {
        const BatchProb &P = probs[bk.p];
        const int n = P.n[bk.l], nn = n * n;
        const int64_t o = P.blkoff[bk.l];
        wg_mm_fx<L>(n, P.X, o, false, P.dZ, o, false, P.W1, o, scr);
        wg_sync();
        wg_mm_fx<L>(n, P.W1, o, false, P.invZ, o, false, P.W2, o, scr);
        wg_sync();
        for (int e = threadIdx.x; e < nn; e += blockDim.x)
            stm<L>(P.dX, o + e, add<L>(ldm<L>(P.rzinv, o + e), neg<L>(ldm<L>(P.W2, o + e))));
        wg_sync();
        const mp<L> half = ldexp2<L>(from_int<L>(1), -1);
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const int i = e % n, j = e / n;
            if (i <= j) continue;
            const mp<L> v = mul<L>(half, add<L>(ldm<L>(P.dX, o + e), ldm<L>(P.dX, o + j + (int64_t)n * i)));
            stm<L>(P.dX, o + e, v);
            stm<L>(P.dX, o + j + (int64_t)n * i, v);
        }
        wg_sync();
    }
}

// Solutions::update (sdpa_dataset.cpp:120-133): X += alphaP dX, Z += alphaD dZ; the task of block 0
// also does y += alphaD dy
template <int L>
__global__ __launch_bounds__(256) void k_bb_update(const BatchProb *__restrict__ probs,
                                                   const BatchBlk *__restrict__ blks, int nblk,
                                                   const int32_t *__restrict__ req, PV sc) {
    BATCH_BLOCK_LOOP {
        const BatchProb &P = probs[bk.p];
        const int n = P.n[bk.l], nn = n * n;
        const int64_t o = P.blkoff[bk.l];
        const mp<L> aP = ldm<L>(sc, (int64_t)bk.p * BSC_N + BSC_ALPHAP), aD = ldm<L>(sc, (int64_t)bk.p * BSC_N + BSC_ALPHAD);
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            stm<L>(P.X, o + e, add<L>(ldm<L>(P.X, o + e), mul<L>(aP, ldm<L>(P.dX, o + e))));
            stm<L>(P.Z, o + e, add<L>(ldm<L>(P.Z, o + e), mul<L>(aD, ldm<L>(P.dZ, o + e))));
        }
        if (bk.l == 0)
            for (int i = threadIdx.x; i < P.m; i += blockDim.x)
                stm<L>(P.y, i, add<L>(ldm<L>(P.y, i), mul<L>(aD, ldm<L>(P.dy, i))));
        wg_sync();
    }
}

// Inverse Cholesky factor of one block (Jal::getInvChol, sdpa_jordan.cpp:123-171) by the division-free
// elimination of the Schur kernel above applied to [A | I]: step j updates the trailing lower triangle
// of At (in W) and, with the same pivot, the identity-started Xt (in Out): Xt_ic <- 2^-e_j (d_j Xt_ic -
// At_ij Xt_jc) for c <= j < i, and Xt_ii <- 2^-e_j d_j Xt_ii, so that Xt_ii ends as the running scale c_i
// and inv(L)_ic = Xt_ic / sqrt(c_i d_i).  The inverse costs no steps beyond the factorisation's n.
// Returns false (uniformly) when a pivot is not positive.
template <int L>
__device__ bool wg_chol_inv(int n, PV Src, PV W, PV S, int64_t o, PV Out) {
    const int nn = n * n;
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        stm<L>(W, o + e, ldm<L>(Src, o + e));
        stm<L>(Out, o + e, (e % n == e / n) ? from_int<L>(1) : zero<L>());
    }
    wg_sync();
    auto at = [&](int i, int j) -> int64_t { return o + i + (int64_t)n * j; };
    for (int j = 0; j < n; j++) {
        const mp<L> d = ldm<L>(W, at(j, j));
        if (sgn<L>(d) <= 0) return false;
        const int n1 = n - 1 - j, nA = n1 * (n1 + 1) / 2, nX = n1 * (j + 1), tot = nA + nX + n1;
        for (int t = threadIdx.x; t < tot; t += blockDim.x) {
            int i, c;
            PV U;
            mp<L> v;
            if (t < nA) {
                int ip = (int)((__builtin_sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
                while (ip > 0 && ip * (ip + 1) / 2 > t) ip--;
                while ((ip + 1) * (ip + 2) / 2 <= t) ip++;
                i = j + 1 + ip;
                c = j + 1 + (t - ip * (ip + 1) / 2);
                U = W;
                v = fms<L>(d, ldm<L>(W, at(i, c)), ldm<L>(W, at(i, j)), ldm<L>(W, at(c, j)));
            } else if (t < nA + nX) {
                const int u = t - nA;
                i = j + 1 + u / (j + 1);
                c = u % (j + 1);
                U = Out;
                v = fms<L>(d, ldm<L>(Out, at(i, c)), ldm<L>(W, at(i, j)), ldm<L>(Out, at(j, c)));
            } else {
                i = c = j + 1 + (t - nA - nX);
                U = Out;
                v = mul<L>(d, ldm<L>(Out, at(i, i)));
            }
            if (!is_zero<L>(v)) v.e -= d.e;
            stm<L>(U, at(i, c), v);
        }
        wg_sync();
    }
    // r_i = 1/sqrt(c_i d_i) with c_i = Xt_ii, then inv(L)_ic = Xt_ic r_i
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        stm<L>(S, o + i, mpx::rsqrt_fx<L>(mul<L>(ldm<L>(Out, at(i, i)), ldm<L>(W, at(i, i)))));
    wg_sync();
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e % n, c = e / n;
        if (i >= c) stm<L>(Out, o + e, mul<L>(ldm<L>(Out, o + e), ldm<L>(S, o + i)));
    }
    wg_sync();
    return true;
}

// Solutions::computeInverse (sdpa_dataset.cpp:96-118): inverse Cholesky factors of X and Z and
// inv(Z) = icZ^T icZ; info[p] = 1 when X or Z is not positive definite
template <int L>
__global__ __launch_bounds__(256) void k_bb_factor(const BatchProb *__restrict__ probs,
                                                   const BatchBlk *__restrict__ blks, int nblk,
                                                   const int32_t *__restrict__ req, int32_t *__restrict__ info,
                                                   uint32_t *fxs, int64_t fxn) {
    uint32_t *scr = fxs + (int64_t)blockIdx.x * fxn;
    BATCH_BLOCK_LOOP {
        const BatchProb &P = probs[bk.p];
        const int n = P.n[bk.l];
        const int64_t o = P.blkoff[bk.l];
        bool ok = wg_chol_inv<L>(n, P.X, P.W1, P.W2, o, P.icX);
        if (ok) ok = wg_chol_inv<L>(n, P.Z, P.W1, P.W2, o, P.icZ);
        if (!ok) {
            if (threadIdx.x == 0) info[bk.p] = 1;
            wg_sync();
            continue;
        }
        wg_mm_fx<L>(n, P.icZ, o, true, P.icZ, o, false, P.invZ, o, scr);
        wg_sync();
    }
}

// the scalar inner products of StepLength / DirectionParameter / SolveInfo (sdpa_parts.cpp:362-860),
// per block (X.Z, X.dZ, dX.Z, dX.dZ, tr X, tr Z, C.X, C.dX over both triangles); the task of block 0
// adds b.y and b.dy
template <int L>
__global__ __launch_bounds__(256) void k_bb_inner(const BatchProb *__restrict__ probs,
                                                  const BatchBlk *__restrict__ blks, int nblk,
                                                  const int32_t *__restrict__ req, PV bo) {
    __shared__ uint32_t sh[(L + 1) * 256];
    BATCH_BLOCK_LOOP {
        const BatchProb &P = probs[bk.p];
        const int n = P.n[bk.l], nn = n * n;
        const int64_t o = P.blkoff[bk.l];
        mp<L> a[8];
        for (int k = 0; k < 8; k++) a[k] = zero<L>();
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const mp<L> x = ldm<L>(P.X, o + e), z = ldm<L>(P.Z, o + e), dx = ldm<L>(P.dX, o + e),
                        dz = ldm<L>(P.dZ, o + e), c = ldm<L>(P.C, o + e);
            a[0] = add<L>(a[0], mul<L>(x, z));
            a[1] = add<L>(a[1], mul<L>(x, dz));
            a[2] = add<L>(a[2], mul<L>(dx, z));
            a[3] = add<L>(a[3], mul<L>(dx, dz));
            if (e % n == e / n) {
                a[4] = add<L>(a[4], x);
                a[5] = add<L>(a[5], z);
            }
            a[6] = add<L>(a[6], mul<L>(c, x));
            a[7] = add<L>(a[7], mul<L>(c, dx));
        }
        for (int k = 0; k < 8; k++) {
            const mp<L> v = wg_sum<L>(a[k], sh);
            if (threadIdx.x == 0) stm<L>(bo, (int64_t)t_ * BO_N + k, v);
        }
        if (bk.l == 0) {
            mp<L> by = zero<L>(), bdy = zero<L>();
            for (int i = threadIdx.x; i < P.m; i += blockDim.x) {
                const mp<L> bi = ldm<L>(P.b, i);
                by = add<L>(by, mul<L>(bi, ldm<L>(P.y, i)));
                bdy = add<L>(bdy, mul<L>(bi, ldm<L>(P.dy, i)));
            }
            by = wg_sum<L>(by, sh);
            bdy = wg_sum<L>(bdy, sh);
            if (threadIdx.x == 0) {
                stm<L>(bo, (int64_t)t_ * BO_N + BO_BY, by);
                stm<L>(bo, (int64_t)t_ * BO_N + BO_BDY, bdy);
            }
        }
        __syncthreads();
    }
}

// Residuals::compute, dual part (sdpa_parts.cpp:760-790): Rd = C - Z - sum_i y_i A_i with max |Rd| per
// block and C.X; the task of block 0 adds max |Rp| (Rp from k_bc_inner mode 1) and b.y
template <int L>
__global__ __launch_bounds__(256) void k_bb_resid(const BatchProb *__restrict__ probs,
                                                  const BatchBlk *__restrict__ blks, int nblk,
                                                  const int32_t *__restrict__ req, PV bo, FxArrays fx) {
    __shared__ uint32_t sh[(L + 1) * 256], sw[L * BATCH_MMAX_LDS];
    __shared__ int32_t se[256];
    BATCH_BLOCK_LOOP {
        const BatchProb &P = probs[bk.p];
        const int n = P.n[bk.l], nn = n * n;
        const int64_t o = P.blkoff[bk.l];
        // Rd = C - Z - sum_i y_i A_i (the sum in block fixed point)
        wg_fx_combo<L>(P, bk.l, P.y, fx, sw, sh, se, [&](int e, const mp<L> &acc) {
            stm<L>(P.Rd, o + e, sub<L>(sub<L>(ldm<L>(P.C, o + e), ldm<L>(P.Z, o + e)), acc));
        });
        mp<L> mx = zero<L>(), cx = zero<L>();
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const mp<L> c = ldm<L>(P.C, o + e), v = ldm<L>(P.Rd, o + e);
            if (cmp_abs<L>(v, mx) > 0) mx = absv<L>(v);
            cx = add<L>(cx, mul<L>(c, ldm<L>(P.X, o + e)));
        }
        mx = wg_maxabs<L>(mx, sh);
        cx = wg_sum<L>(cx, sh);
        if (threadIdx.x == 0) {
            stm<L>(bo, (int64_t)t_ * BO_N + BO_MAXRD, mx);
            stm<L>(bo, (int64_t)t_ * BO_N + BO_CX, cx);
        }
        if (bk.l == 0) {
            mp<L> rp = zero<L>(), by = zero<L>();
            for (int i = threadIdx.x; i < P.m; i += blockDim.x) {
                const mp<L> v = ldm<L>(P.Rp, i);
                if (cmp_abs<L>(v, rp) > 0) rp = absv<L>(v);
                by = add<L>(by, mul<L>(ldm<L>(P.b, i), ldm<L>(P.y, i)));
            }
            rp = wg_maxabs<L>(rp, sh);
            by = wg_sum<L>(by, sh);
            if (threadIdx.x == 0) {
                stm<L>(bo, (int64_t)t_ * BO_N + BO_MAXRP, rp);
                stm<L>(bo, (int64_t)t_ * BO_N + BO_BY, by);
            }
        }
        __syncthreads();
    }
}

// ---- step length: minimum eigenvalue of lMat xMat lMat^T per block (Jal::getMinEigen, sdpa_jordan.cpp:55-88)

// eigenvalues of the symmetric tridiagonal (d, e) in double by implicit QL with Wilkinson shifts (d in:
// diagonal, out: eigenvalues; e: off-diagonal e[0..n-2], destroyed); returns the index of the smallest,
// or -1 without convergence
__device__ inline int tql_values(int n, double *d, double *e) {
    e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        for (;;) {
            for (m = l; m < n - 1; m++) {
                const double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= 1.1102230246251565e-16 * dd) break;
            }
            if (m == l) break;
            if (iter++ == 60 * n) return -1;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = sqrt(g * g + 1.0);
            g = d[m] - d[l] + e[l] / (g + (g < 0 ? -r : r));
            double s = 1.0, c = 1.0, p = 0.0;
            bool early = false;
            for (int i = m - 1; i >= l; i--) {
                const double f = s * e[i], b = c * e[i];
                r = sqrt(f * f + g * g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    int k = 0;
    for (int i = 1; i < n; i++)
        if (d[i] < d[k]) k = i;
    return k;
}

// unit eigenvector x of the tridiagonal (a: diagonal, b: off-diagonal) for its eigenvalue lam: two steps
// of inverse iteration, (T - sigma I) x_new = x by the tridiagonal elimination (sigma just below lam)
__device__ inline void tri_inverse_iteration(int n, const double *a, const double *b, double lam, double *x, double *w) {
    double scale = 0.0;
    for (int i = 0; i < n; i++) scale = fmax(scale, fabs(a[i]) + (i + 1 < n ? fabs(b[i]) : 0.0));
    const double sigma = lam - 1e-12 * fmax(scale, 1e-300);
    for (int i = 0; i < n; i++) x[i] = 1.0;
    for (int it = 0; it < 2; it++) {
        // forward elimination: w holds the modified super-diagonal
        double den = a[0] - sigma;
        if (den == 0.0) den = 1e-300;
        w[0] = (n > 1 ? b[0] : 0.0) / den;
        x[0] = x[0] / den;
        for (int i = 1; i < n; i++) {
            den = (a[i] - sigma) - b[i - 1] * w[i - 1];
            if (den == 0.0) den = 1e-300;
            w[i] = (i + 1 < n ? b[i] : 0.0) / den;
            x[i] = (x[i] - b[i - 1] * x[i - 1]) / den;
        }
        for (int i = n - 2; i >= 0; i--) x[i] -= w[i] * x[i + 1];
        double nr = 0.0;
        for (int i = 0; i < n; i++) nr += x[i] * x[i];
        nr = 1.0 / sqrt(nr);
        for (int i = 0; i < n; i++) x[i] *= nr;
    }
}

constexpr int BEIG_KMAX = 20;  // Lanczos steps of a block of at most 40 (the loop ends by sqrt(n) + 10)
constexpr int BEIG_NMAX = 40;
constexpr int BEIG_JMAX = 20;  // blocks taking the full eigenvalue problem (Rsyev in the reference)
constexpr int BEIG_LD = BEIG_JMAX + 1;

// one sweep of the parallel cyclic Jacobi method on the symmetric matrix a (leading dimension BEIG_LD,
// padded to an even order N with a zero row and column), eigenvector columns accumulated in v: N - 1
// rounds of N/2 disjoint rotations in round-robin order (index N - 1 fixed, the others rotating), each
// round applied as A <- A J (columns), then A <- J^T A (rows), V <- V J.  cs: N/2 (c, s) pairs.
__device__ inline void jacobi_sweep(int N, double *a, double *v, double *cs, int *pr) {
    const int tid = threadIdx.x, half = N / 2;
    for (int round = 0; round < N - 1; round++) {
        if (tid < half) {
            // pair tid of this round: players (round + tid) mod (N-1) and (round - tid) mod (N-1) for
            // tid > 0, and (round, N - 1) for tid = 0
            int p, q;
            if (tid == 0) {
                p = round;
                q = N - 1;
            } else {
                p = (round + tid) % (N - 1);
                q = (round - tid + (N - 1)) % (N - 1);
            }
            if (p > q) {
                const int t = p;
                p = q;
                q = t;
            }
            double c = 1.0, s = 0.0;
            const double apq = a[p + q * BEIG_LD];
            if (apq != 0.0) {
                const double th = (a[q + q * BEIG_LD] - a[p + p * BEIG_LD]) / (2.0 * apq);
                double t = 1.0 / (fabs(th) + sqrt(th * th + 1.0));
                if (th < 0) t = -t;
                c = 1.0 / sqrt(t * t + 1.0);
                s = t * c;
            }
            cs[2 * tid] = c;
            cs[2 * tid + 1] = s;
            pr[2 * tid] = p;
            pr[2 * tid + 1] = q;
        }
        __syncthreads();
        for (int u = tid; u < half * N; u += blockDim.x) {  // columns: A <- A J, V <- V J
            const int k = u / N, r = u % N;
            const int p = pr[2 * k], q = pr[2 * k + 1];
            const double c = cs[2 * k], s = cs[2 * k + 1];
            const double arp = a[r + p * BEIG_LD], arq = a[r + q * BEIG_LD];
            a[r + p * BEIG_LD] = c * arp - s * arq;
            a[r + q * BEIG_LD] = s * arp + c * arq;
            const double vrp = v[r + p * BEIG_LD], vrq = v[r + q * BEIG_LD];
            v[r + p * BEIG_LD] = c * vrp - s * vrq;
            v[r + q * BEIG_LD] = s * vrp + c * vrq;
        }
        __syncthreads();
        for (int u = tid; u < half * N; u += blockDim.x) {  // rows: A <- J^T A
            const int k = u / N, r = u % N;
            const int p = pr[2 * k], q = pr[2 * k + 1];
            const double c = cs[2 * k], s = cs[2 * k + 1];
            const double apr = a[p + r * BEIG_LD], aqr = a[q + r * BEIG_LD];
            a[p + r * BEIG_LD] = c * apr - s * aqr;
            a[q + r * BEIG_LD] = s * apr + c * aqr;
        }
        __syncthreads();
        if (tid < half) {  // the annihilated pair is zero up to rounding: make it exact
            const int p = pr[2 * tid], q = pr[2 * tid + 1];
            a[p + q * BEIG_LD] = 0.0;
            a[q + p * BEIG_LD] = 0.0;
        }
        __syncthreads();
    }
}

// which 0: (icX, dX), 1: (icZ, dZ).  M = lMat xMat lMat^T is formed in W2.  Blocks of n <= 20: the smallest
// eigenvalue of M (Rsyev, Lal::getMinEigenValue, sdpa_linear.cpp:126-158) from the eigenvector of a parallel
// Jacobi solve in double refined by its Rayleigh quotient in multiple precision.  Larger blocks: the
// Lanczos process of Lal::getMinEigen (sdpa_linear.cpp:27-124) with its stopping rule, in multiple
// precision; its tridiagonal eigenproblem (Rsteqr there) is solved in double (eigenvalues by QL, the
// smallest one's eigenvector by inverse iteration) and the eigenvalue refined by the Rayleigh quotient in
// multiple precision.
template <int L>
__global__ __launch_bounds__(256) void k_bb_mineig(const BatchProb *__restrict__ probs,
                                                   const BatchBlk *__restrict__ blks, int nblk,
                                                   const int32_t *__restrict__ req, int which, PV bo, uint32_t *fxs,
                                                   int64_t fxn) {
    uint32_t *scr = fxs + (int64_t)blockIdx.x * fxn;
    __shared__ uint32_t sh[(L + 1) * 256];
    __shared__ uint32_t vq[(L + 1) * BEIG_NMAX], vqo[(L + 1) * BEIG_NMAX], vr[(L + 1) * BEIG_NMAX],
        vw[(L + 1) * BEIG_NMAX], tal[(L + 1) * BEIG_KMAX], tbe[(L + 1) * BEIG_KMAX], tsc[(L + 1) * 4];
    __shared__ double dA[BEIG_LD * BEIG_LD], dV[BEIG_LD * BEIG_LD], dcs[BEIG_LD + 1];
    __shared__ int dpr[BEIG_LD + 1];
    __shared__ double td[BEIG_KMAX], te[BEIG_KMAX], tx[BEIG_KMAX], tw[BEIG_KMAX];
    __shared__ int sflag;
    BATCH_BLOCK_LOOP {
        const BatchProb &P = probs[bk.p];
        const int n = P.n[bk.l], nn = n * n;
        const int64_t o = P.blkoff[bk.l];
        const PV Lm = which ? P.icZ : P.icX, Xm = which ? P.dZ : P.dX;
        wg_mm_fx<L>(n, Xm, o, false, Lm, o, true, P.W1, o, scr);
        wg_sync();
        wg_mm_fx<L>(n, Lm, o, false, P.W1, o, false, P.W2, o, scr);
        wg_sync();
        const PV M = P.W2;
        mp<L> result;
        if (n <= BEIG_JMAX) {
            const int N = n + (n & 1);
            for (int e = threadIdx.x; e < BEIG_LD * BEIG_LD; e += blockDim.x) {
                const int r = e % BEIG_LD, c = e / BEIG_LD;
                dA[e] = (r < n && c < n) ? to_double<L>(ldm<L>(M, o + r + (int64_t)n * c)) : 0.0;
                dV[e] = (r == c) ? 1.0 : 0.0;
            }
            __syncthreads();
            for (int sweep = 0; sweep < 40; sweep++) {
                // off-diagonal against diagonal mass (1e-26: rotations leave about n^2 u^2 |a|^2 off the
                // diagonal; the Rayleigh quotient squares the eigenvector's error of about 1e-13)
                double off = 0.0, dg = 0.0;
                if (threadIdx.x < 64) {
                    for (int e = threadIdx.x; e < N * N; e += 64) {
                        const int r = e % N, c = e / N;
                        const double x = dA[r + c * BEIG_LD];
                        (r == c ? dg : off) += x * x;
                    }
                    for (int h = 32; h >= 1; h >>= 1) {
                        off += __shfl_xor(off, h, 64);
                        dg += __shfl_xor(dg, h, 64);
                    }
                    if (threadIdx.x == 0) sflag = (off == 0.0 || off <= dg * 1e-26) ? 1 : 0;
                }
                __syncthreads();
                if (sflag) break;
                jacobi_sweep(N, dA, dV, dcs, dpr);
            }
            if (threadIdx.x == 0) {
                int k = 0;
                for (int i = 1; i < n; i++)
                    if (dA[i + i * BEIG_LD] < dA[k + k * BEIG_LD]) k = i;
                sflag = k;
            }
            __syncthreads();
            const int k = sflag;
            // Rayleigh quotient of the eigenvector v: (v^T M v) / (v^T v)
            mp<L> num = zero<L>(), den = zero<L>();
            if (threadIdx.x < n) {
                const int a = threadIdx.x;
                mp<L> s = zero<L>();
                for (int b = 0; b < n; b++)
                    s = add<L>(s, mul<L>(ldm<L>(M, o + a + (int64_t)n * b), from_double<L>(dV[b + k * BEIG_LD])));
                const mp<L> va = from_double<L>(dV[a + k * BEIG_LD]);
                num = mul<L>(va, s);
                den = mul<L>(va, va);
            }
            num = wg_sum<L>(num, sh);
            den = wg_sum<L>(den, sh);
            result = mpx::div<L>(num, den);
        } else {
            const mp<L> one = from_int<L>(1), zr = zero<L>();
            mp<L> mn = from_double<L>(1.0e+51), mn_old = from_double<L>(1.0e+52), error = from_double<L>(1.0e+10);
            const mp<L> nD = from_int<L>(n);
            mp<L> beta = sqrtv<L>(nD);
            const mp<L> c16 = from_double<L>(1.0e-16), c5 = from_double<L>(1.0e-5), c8 = from_double<L>(1.0e-8),
                        c2 = from_double<L>(1.0e-2), c4 = from_double<L>(1.0e-4);
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                lds_st<L>(vq, BEIG_NMAX, i, zr);
                lds_st<L>(vr, BEIG_NMAX, i, one);
            }
            __syncthreads();
            // matrix-vector products: row i split over MV_PARTS threads, partial sums through sh
            constexpr int MV_PARTS = 6;
            int k = 0, kk = 0;
            for (;;) {
                // k < sqrt(n) + 10 exactly: k < 10 or (k - 10)^2 < n
                if (!(k < n) || !(k < 10 || (k - 10) * (k - 10) < n) || !(cmp<L>(beta, c16) > 0) || k >= BEIG_KMAX) break;
                const bool condA = cmp<L>(absv<L>(sub<L>(mn, mn_old)), add<L>(mul<L>(c5, absv<L>(mn)), c8)) > 0;
                const bool condB = cmp<L>(absv<L>(mul<L>(error, beta)), add<L>(mul<L>(c2, absv<L>(mn)), c4)) > 0;
                if (!(condA || condB)) break;
                const mp<L> value = mpx::div<L>(one, beta);
                for (int i = threadIdx.x; i < n; i += blockDim.x) {
                    lds_st<L>(vqo, BEIG_NMAX, i, lds_ld<L>(vq, BEIG_NMAX, i));
                    lds_st<L>(vq, BEIG_NMAX, i, mul<L>(value, lds_ld<L>(vr, BEIG_NMAX, i)));
                }
                __syncthreads();
                if (threadIdx.x < n * MV_PARTS) {
                    const int i = threadIdx.x / MV_PARTS, part = threadIdx.x % MV_PARTS;
                    mp<L> s = zero<L>();
                    for (int b = part; b < n; b += MV_PARTS)
                        s = add<L>(s, mul<L>(ldm<L>(M, o + i + (int64_t)n * b), lds_ld<L>(vq, BEIG_NMAX, b)));
                    lds_st<L>(sh, 256, threadIdx.x, s);
                }
                __syncthreads();
                for (int i = threadIdx.x; i < n; i += blockDim.x) {
                    mp<L> s = lds_ld<L>(sh, 256, i * MV_PARTS);
                    for (int part = 1; part < MV_PARTS; part++) s = add<L>(s, lds_ld<L>(sh, 256, i * MV_PARTS + part));
                    lds_st<L>(vw, BEIG_NMAX, i, s);
                }
                __syncthreads();
                mp<L> part = zero<L>();
                for (int i = threadIdx.x; i < n; i += blockDim.x)
                    part = add<L>(part, mul<L>(lds_ld<L>(vq, BEIG_NMAX, i), lds_ld<L>(vw, BEIG_NMAX, i)));
                const mp<L> alpha = wg_sum<L>(part, sh);
                if (threadIdx.x == 0) lds_st<L>(tal, BEIG_KMAX, k, alpha);
                for (int i = threadIdx.x; i < n; i += blockDim.x) {
                    mp<L> r = add<L>(lds_ld<L>(vw, BEIG_NMAX, i), mul<L>(neg<L>(alpha), lds_ld<L>(vq, BEIG_NMAX, i)));
                    r = add<L>(r, mul<L>(neg<L>(beta), lds_ld<L>(vqo, BEIG_NMAX, i)));
                    lds_st<L>(vr, BEIG_NMAX, i, r);
                }
                __syncthreads();
                // kk >= sqrt(k) and k > sqrt(n + 9), exactly in integers
                if (kk * kk >= k || k == n - 1 || k * k > n + 9) {
                    kk = 0;
                    // T = tridiag(alpha_0..alpha_k; beta_0..beta_{k-1})
                    if (threadIdx.x == 0) {
                        const int nt = k + 1;
                        for (int i = 0; i < nt; i++) {
                            td[i] = to_double<L>(lds_ld<L>(tal, BEIG_KMAX, i));
                            te[i] = i + 1 < nt ? to_double<L>(lds_ld<L>(tbe, BEIG_KMAX, i)) : 0.0;
                        }
                        const int j = tql_values(nt, td, te);
                        sflag = j;
                        if (j >= 0) {
                            const double lam = td[j];
                            for (int i = 0; i < nt; i++) {
                                td[i] = to_double<L>(lds_ld<L>(tal, BEIG_KMAX, i));
                                te[i] = i + 1 < nt ? to_double<L>(lds_ld<L>(tbe, BEIG_KMAX, i)) : 0.0;
                            }
                            tri_inverse_iteration(nt, td, te, lam, tx, tw);
                            // Rayleigh quotient of the eigenvector in multiple precision; error = its last component
                            mp<L> num = zero<L>(), den = zero<L>();
                            for (int i = 0; i < nt; i++) {
                                const mp<L> vi = from_double<L>(tx[i]);
                                mp<L> tv = mul<L>(lds_ld<L>(tal, BEIG_KMAX, i), vi);
                                if (i > 0) tv = add<L>(tv, mul<L>(lds_ld<L>(tbe, BEIG_KMAX, i - 1), from_double<L>(tx[i - 1])));
                                if (i + 1 < nt) tv = add<L>(tv, mul<L>(lds_ld<L>(tbe, BEIG_KMAX, i), from_double<L>(tx[i + 1])));
                                num = add<L>(num, mul<L>(vi, tv));
                                den = add<L>(den, mul<L>(vi, vi));
                            }
                            lds_st<L>(tsc, 4, 0, mpx::div<L>(num, den));
                            lds_st<L>(tsc, 4, 1, mpx::div<L>(from_double<L>(tx[k]), sqrtv<L>(den)));
                        }
                    }
                    __syncthreads();
                    if (sflag < 0) break;
                    mn_old = mn;
                    mn = lds_ld<L>(tsc, 4, 0);
                    error = lds_ld<L>(tsc, 4, 1);
                    __syncthreads();
                }
                part = zero<L>();
                for (int i = threadIdx.x; i < n; i += blockDim.x) {
                    const mp<L> r = lds_ld<L>(vr, BEIG_NMAX, i);
                    part = add<L>(part, mul<L>(r, r));
                }
                beta = sqrtv<L>(wg_sum<L>(part, sh));
                if (threadIdx.x == 0) lds_st<L>(tbe, BEIG_KMAX, k, beta);
                __syncthreads();
                ++k;
                ++kk;
            }
            result = sub<L>(mn, absv<L>(mul<L>(error, beta)));
        }
        if (threadIdx.x == 0) stm<L>(bo, (int64_t)t_ * BO_N + BO_EIG, result);
        wg_sync();
    }
}
#undef BATCH_BLOCK_LOOP

// kinds of batched calls whose device time is accumulated (sdpg_batch_timers)
enum { BT_NEWTON, BT_GVEC, BT_DZ, BT_DX, BT_UPDATE, BT_FACTOR, BT_INNER, BT_RESID, BT_EIG, BT_N };

struct BatchBase {
    virtual ~BatchBase() {}
    virtual int newton(const int *req, int *info) = 0;
    virtual int set_graph(int on) = 0;
    virtual int timing(double *ms) = 0;
    virtual int wait_contexts(int on) = 0;
    virtual int timers(double *ms, int n) = 0;
    virtual int gvec(const int *req, const uint32_t *mu, const uint32_t *beta, const int *phase) = 0;
    virtual int dz(const int *req, const int *phase) = 0;
    virtual int dx(const int *req) = 0;
    virtual int update(const int *req, const uint32_t *alphaP, const uint32_t *alphaD) = 0;
    virtual int factor_point(const int *req, int *info) = 0;
    virtual int inner(const int *req, uint32_t *out) = 0;
    virtual int residuals(const int *req, uint32_t *out) = 0;
    virtual int min_eigen(const int *req, int which, uint32_t *out) = 0;
};

template <int L>
struct Batch : BatchBase {
    std::vector<Engine<L> *> eng;
    int device = 0, nprob = 0, ntasks = 0, nmax = 0, mmax = 0, grid_schur = 0, grid_prob = 0;
    size_t shm_schur = 0, shm_potrs = 0;
    hipStream_t st = nullptr;
    BatchProb *d_probs = nullptr;
    int2 *d_tasks = nullptr;
    int32_t *d_req = nullptr, *d_info = nullptr;
    uint32_t *cs = nullptr;
    int64_t cs_n = 0;
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0;
    bool wait_ctx = true;
    // the other phases: block tasks, per-problem arguments, scalars in, per-task results out
    std::vector<BatchBlk> blks;
    std::vector<int> first_blk;  // per problem: index of its block 0 in blks
    BatchBlk *d_blks = nullptr;
    int32_t *d_iarg = nullptr, *d_binfo = nullptr;
    uint32_t *d_sc = nullptr, *d_bo = nullptr;
    int64_t sc_n = 0, bo_n = 0;
    std::vector<uint32_t> h_sc, h_bo;
    std::vector<int32_t> h_iarg;
    int grid_blk = 0;
    double tm[BT_N] = {0};
    // block fixed point of the Schur build (k_fx_constraints, k_fx_point)
    uint32_t *fa_r = nullptr, *fa_rs = nullptr, *fa_g = nullptr, *fa_gs = nullptr, *fx_v = nullptr, *fx_s = nullptr,
             *fz_v = nullptr, *fz_s = nullptr;
    int32_t *fa_re = nullptr, *fa_ge = nullptr, *fx_e = nullptr, *fz_e = nullptr;
    uint32_t *ws_schur = nullptr;  // per workgroup of k_batch_schur: work1 / work2 as limb floats
    int64_t ws_n = 0;
    uint32_t *fa_cs = nullptr, *fd_v = nullptr, *fd_s = nullptr;  // FxArrays of the other phases
    uint32_t *fa_ps = nullptr;  // offset sums of the constraint matrices' lower triangles (packed inner products)
    int32_t *fd_e = nullptr;
    FxArrays fxa() const { return FxArrays{fa_g, fa_ge, fa_gs, fa_cs, fd_v, fd_e, fd_s}; }
    uint32_t *fxs = nullptr;  // per workgroup of the block-task kernels: the fixed-point products' scratch
    int64_t fxs_n = 0;

    ~Batch() override {
        if (st) (void)hipStreamSynchronize(st);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        (void)hipFree(d_probs);
        (void)hipFree(d_tasks);
        (void)hipFree(d_req);
        (void)hipFree(d_info);
        (void)hipFree(cs);
        (void)hipFree(d_blks);
        (void)hipFree(d_iarg);
        (void)hipFree(d_binfo);
        (void)hipFree(d_sc);
        (void)hipFree(d_bo);
        (void)hipFree(ws_schur);
        (void)hipFree(fa_cs);
        (void)hipFree(fa_ps);
        (void)hipFree(fxs);
        (void)hipFree(fd_v);
        (void)hipFree(fd_s);
        (void)hipFree(fd_e);
        for (void *q : {(void *)fa_r, (void *)fa_rs, (void *)fa_g, (void *)fa_gs, (void *)fx_v, (void *)fx_s,
                        (void *)fz_v, (void *)fz_s, (void *)fa_re, (void *)fa_ge, (void *)fx_e, (void *)fz_e})
            (void)hipFree(q);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    }

    int create(EngineBase *const *es, int n) {
        if (n <= 0) return SDPG_EARG;
        device = es[0]->device;
        int64_t fa_el_n = 0, fa_row_n = 0, fa_mat_n = 0, fp_el_n = 0, fp_row_n = 0;
        std::vector<BatchProb> hp;
        std::vector<int2> tasks;
        int64_t csn = 0;
        for (int p = 0; p < n; p++) {
            auto *e = dynamic_cast<Engine<L> *>(es[p]);
            if (!e || !e->committed || e->device != device) {
                set_err("sdpg_batch_create: contexts must be committed, of one precision and on one device");
                return SDPG_EARG;
            }
            if (e->nlp > 0 || e->nsdp > BMAXB || e->nsdp < 1) {
                set_err("sdpg_batch_create: the batched form takes 1 to 4 SDP blocks and no LP block");
                return SDPG_EARG;
            }
            for (int l = 0; l < e->nsdp; l++)
                if (e->sizes[l] > BEIG_NMAX) {
                    set_err("sdpg_batch_create: the batched form takes blocks of at most 40");
                    return SDPG_EARG;
                }
            BatchProb b{};
            b.m = e->m;
            b.nsdp = e->nsdp;
            for (int l = 0; l < e->nsdp; l++) {
                b.n[l] = e->sizes[l];
                b.blkoff[l] = e->blkoff[l];
                nmax = std::max(nmax, e->sizes[l]);
            }
            for (int i = 0; i < e->m; i++)
                for (int l = 0; l < e->nsdp; l++) {
                    if (e->hdesc[(size_t)i * e->nsdp + l].kind == 1) {
                        set_err("sdpg_batch_create: the batched form takes dense constraint blocks only");
                        return SDPG_EARG;
                    }
                }
            if (e->m > BATCH_MMAX_LDS) {
                set_err("sdpg_batch_create: the batched form takes at most 256 constraints per problem");
                return SDPG_EARG;
            }
            mmax = std::max(mmax, e->m);
            b.desc = e->d_desc;
            b.X = e->X.pv();
            b.invZ = e->invZ.pv();
            b.dval = e->dval.pv();
            b.B = e->Bm.pv();
            b.g = e->g.pv();
            b.dy = e->dy.pv();
            b.Z = e->Z.pv();
            b.y = e->y.pv();
            b.b = e->b.pv();
            b.Rp = e->Rp.pv();
            b.Rd = e->Rd.pv();
            b.C = e->Cfull.pv();
            b.rzinv = e->rzinv.pv();
            b.xrd = e->xrd.pv();
            b.dls = e->DLS1.pv();
            b.dX = e->dX.pv();
            b.dZ = e->dZ.pv();
            b.icX = e->icX.pv();
            b.icZ = e->icZ.pv();
            b.W1 = e->W1.pv();
            b.W2 = e->W2.pv();
            b.blk0 = (int32_t)blks.size();
            for (int l = 0; l < e->nsdp; l++) blks.push_back(BatchBlk{p, l});
            b.Nd = (int32_t)e->Nd;
            b.nsum = 0;
            for (int l = 0; l < e->nsdp; l++) {
                b.rowoff[l] = b.nsum;
                b.nsum += e->sizes[l];
            }
            b.fa_el = fa_el_n;
            b.fa_row = fa_row_n;
            b.fa_mat = fa_mat_n;
            b.fp_el = fp_el_n;
            b.fp_row = fp_row_n;
            fa_el_n += (int64_t)e->m * e->Nd;
            fa_row_n += (int64_t)e->m * b.nsum;
            fa_mat_n += (int64_t)e->m * e->nsdp;
            fp_el_n += e->Nd;
            fp_row_n += b.nsum;
            b.cs_off = csn;
            csn += 2 * (int64_t)e->m;
            hp.push_back(b);
            for (int i = 0; i < e->m; i++) tasks.push_back(int2{p, i});
            eng.push_back(e);
        }
        nprob = n;
        // XCD-aware order of the constraint tasks: the workgroups of a launch go round-robin to the 8 XCDs
        // (workgroup id mod 8), so position t of the task list runs on XCD t mod 8.  Each problem goes
        // whole to one XCD (largest problems first, each to the least loaded XCD), so that the
        // constraint matrices its tasks stream for the inner products stay in that XCD's L2; within a
        // problem the tasks run from the largest i down.  Shorter XCD lists are padded with empty tasks.
        {
            constexpr int NX = 8;
            auto tcost = [&](int q, int i) {
                int64_t c = 0;
                for (int l = 0; l < hp[q].nsdp; l++) c += (int64_t)hp[q].n[l] * hp[q].n[l] * (2 * hp[q].n[l] + i);
                return c;
            };
            std::vector<int> order(n);
            std::vector<int64_t> pcost(n, 0);
            for (int q = 0; q < n; q++) {
                order[q] = q;
                for (int i = 0; i < hp[q].m; i++) pcost[q] += tcost(q, i);
            }
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pcost[a] > pcost[b]; });
            std::vector<int64_t> load(NX, 0);
            std::vector<std::vector<int2>> xt(NX);
            for (int q : order) {
                const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                load[x] += pcost[q];
                for (int i = hp[q].m - 1; i >= 0; i--) xt[x].push_back(int2{q, i});
            }
            size_t len = 0;
            for (auto &v : xt) len = std::max(len, v.size());
            tasks.assign(len * NX, int2{-1, 0});
            for (int x = 0; x < NX; x++)
                for (size_t k = 0; k < xt[x].size(); k++) tasks[k * NX + x] = xt[x][k];
        }
        ntasks = (int)tasks.size();
        size_t lds_max = 65536;
        {
            int v = 0;
            if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
                lds_max = (size_t)v;
        }
        shm_schur = 4 * schur_lds_words<L>(nmax, mmax);
        shm_potrs = (size_t)(L + 1) * 4 * mmax;
        if (shm_schur > lds_max) {
            set_err("sdpg_batch_create: blocks too large for the batched form at this precision");
            return SDPG_EARG;
        }
        int cus = 0;
        if (hipSetDevice(device) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
            return SDPG_EHIP;
        // a multiple of 8 workgroups, so that workgroup w and its tasks w + k grid stay on XCD w mod 8
        grid_schur = std::max(8, (std::min(ntasks, cus * std::max(1, (int)(lds_max / shm_schur))) / 8) * 8);
        grid_prob = std::max(1, std::min(nprob, 4 * cus));
        cs_n = std::max<int64_t>(csn, 1);
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
            hipMalloc((void **)&d_probs, sizeof(BatchProb) * n) != hipSuccess ||
            hipMalloc((void **)&d_tasks, sizeof(int2) * std::max(1, ntasks)) != hipSuccess ||
            hipMalloc((void **)&d_req, sizeof(int32_t) * n) != hipSuccess ||
            hipMalloc((void **)&d_info, sizeof(int32_t) * n) != hipSuccess ||
            hipMalloc((void **)&cs, sizeof(uint32_t) * (L + 1) * cs_n) != hipSuccess ||
            hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            set_err("sdpg_batch_create: out of device memory");
            return SDPG_ENOMEM;
        }
        first_blk.assign(n, 0);
        for (int t = (int)blks.size() - 1; t >= 0; t--)
            if (blks[t].l == 0) first_blk[blks[t].p] = t;
        sc_n = (int64_t)n * BSC_N;
        bo_n = (int64_t)blks.size() * BO_N;
        grid_blk = std::max(1, std::min((int)blks.size(), 8 * cus));
        h_sc.assign((size_t)(L + 1) * sc_n, 0);
        h_bo.assign((size_t)(L + 1) * bo_n, 0);
        h_iarg.assign(n, 0);
        if (hipMalloc((void **)&d_blks, sizeof(BatchBlk) * blks.size()) != hipSuccess ||
            hipMalloc((void **)&d_iarg, sizeof(int32_t) * n) != hipSuccess ||
            hipMalloc((void **)&d_binfo, sizeof(int32_t) * n) != hipSuccess ||
            hipMalloc((void **)&d_sc, sizeof(uint32_t) * h_sc.size()) != hipSuccess ||
            hipMalloc((void **)&d_bo, sizeof(uint32_t) * h_bo.size()) != hipSuccess) {
            set_err("sdpg_batch_create: out of device memory");
            return SDPG_ENOMEM;
        }
        if (hipMemcpy(d_blks, blks.data(), sizeof(BatchBlk) * blks.size(), hipMemcpyHostToDevice) != hipSuccess)
            return SDPG_EHIP;
        // block fixed point of the Schur build: the constraint matrices (filled once below), X and inv(Z)
        auto dm = [&](void **ptr, size_t bytes) { return hipMalloc(ptr, std::max<size_t>(bytes, 4)) == hipSuccess; };
        if (!dm((void **)&fa_r, 4 * (size_t)L * fa_el_n) || !dm((void **)&fa_re, 4 * (size_t)fa_row_n) ||
            !dm((void **)&fa_rs, 4 * (size_t)(L + 1) * fa_row_n) || !dm((void **)&fa_g, 4 * (size_t)L * fa_el_n) ||
            !dm((void **)&fa_ge, 4 * (size_t)fa_mat_n) || !dm((void **)&fa_gs, 4 * (size_t)(L + 1) * fa_mat_n) ||
            !dm((void **)&fx_v, 4 * (size_t)L * fp_el_n) || !dm((void **)&fx_e, 4 * (size_t)fp_row_n) ||
            !dm((void **)&fx_s, 4 * (size_t)(L + 1) * fp_row_n) || !dm((void **)&fz_v, 4 * (size_t)L * fp_el_n) ||
            !dm((void **)&fz_e, 4 * (size_t)fp_row_n) || !dm((void **)&fz_s, 4 * (size_t)(L + 1) * fp_row_n) ||
            !dm((void **)&ws_schur, 4 * (size_t)(L + 1) * (ws_n = (int64_t)grid_schur * nmax * nmax)) ||
            !dm((void **)&fa_cs, 4 * (size_t)(L + 1) * fp_el_n) || !dm((void **)&fa_ps, 4 * (size_t)(L + 1) * fa_mat_n) || !dm((void **)&fd_v, 4 * (size_t)L * fp_el_n) ||
            !dm((void **)&fd_s, 4 * (size_t)(L + 1) * blks.size()) || !dm((void **)&fd_e, 4 * blks.size()) ||
            !dm((void **)&fxs, 4 * (size_t)grid_blk * (fxs_n = fx_mm_words<L>(nmax)))) {
            set_err("sdpg_batch_create: out of device memory");
            return SDPG_ENOMEM;
        }
        if (hipFuncSetAttribute((const void *)k_batch_schur<L>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shm_schur) != hipSuccess ||
            hipMemcpy(d_probs, hp.data(), sizeof(BatchProb) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_tasks, tasks.data(), sizeof(int2) * ntasks, hipMemcpyHostToDevice) != hipSuccess)
            return SDPG_EHIP;
        for (Engine<L> *e : eng) HIPCHK(hipStreamSynchronize(e->st));  // the constraint values are in place
        k_fx_constraints<L><<<std::max(1, std::min(ntasks, 8 * cus)), 256, 0, st>>>(d_probs, d_tasks, ntasks, fa_r,
                                                                                   fa_re, fa_rs, fa_g, fa_ge, fa_gs,
                                                                                   fa_ps);
        HIPCHK(hipGetLastError());
        k_fx_colsum<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), fxa(), fa_cs);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        return SDPG_OK;
    }

    // the three phases on the batch's stream (what the graph holds)
    void launch() {
        k_fx_point<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, fx_v, fx_e, fx_s, fz_v, fz_e,
                                                fz_s);
        k_batch_schur<L><<<grid_schur, BS_THREADS, shm_schur, st>>>(d_probs, d_tasks, ntasks, d_req, nmax, mmax, fa_r,
                                                                    fa_re, fa_rs, fa_g, fa_ge, fa_ps, fx_v, fx_e, fx_s,
                                                                    fz_v, fz_e, fz_s, PV{ws_schur, ws_n});
        k_batch_potrf<L><<<grid_prob, BPOTRF_THREADS, 0, st>>>(d_probs, nprob, d_req, d_info, PV{cs, cs_n});
        k_batch_potrs<L><<<grid_prob, 256, shm_potrs, st>>>(d_probs, nprob, d_req, d_info, PV{cs, cs_n}, mmax);
    }

    // the contexts' own work (point updates, right-hand sides) is complete before the batch reads it
    int begin(const int *req) {
        if (hipSetDevice(device) != hipSuccess) return SDPG_EHIP;
        if (wait_ctx)
            for (Engine<L> *e : eng)
                if (hipStreamSynchronize(e->st) != hipSuccess) return SDPG_EHIP;
        HIPCHK(hipMemcpyAsync(d_req, req, sizeof(int32_t) * nprob, hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(e0, st));
        return SDPG_OK;
    }
    // after the launches: results back, wait, device time of the call
    int end(int kind, void *dst = nullptr, const void *src = nullptr, size_t bytes = 0) {
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e1, st));
        if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) tm[kind] += ms;
        return SDPG_OK;
    }
    // packed scalar (L+2 words) <-> planar arrays
    static void put_planar(std::vector<uint32_t> &h, int64_t N, int64_t i, const uint32_t *w) {
        for (int k = 0; k < L; k++) h[(size_t)k * N + i] = w[2 + k];
        h[(size_t)L * N + i] = (uint32_t)(((int32_t)w[0] << 1) | (int32_t)(w[1] & 1u));
    }
    static mp<L> get_planar(const std::vector<uint32_t> &h, int64_t N, int64_t i) {
        mp<L> x;
        for (int k = 0; k < L; k++) x.m[k] = h[(size_t)k * N + i];
        const int32_t hd = (int32_t)h[(size_t)L * N + i];
        x.e = hd >> 1;
        x.s = (uint32_t)hd & 1u;
        return x;
    }
    static mp<L> unpack_w(const uint32_t *w) {
        mp<L> x;
        x.e = (int32_t)w[0];
        x.s = w[1] & 1u;
        for (int k = 0; k < L; k++) x.m[k] = w[2 + k];
        if (x.m[L - 1] == 0) x = zero<L>();
        return x;
    }
    static void pack_w(const mp<L> &x, uint32_t *w) {
        w[0] = (uint32_t)x.e;
        w[1] = x.s;
        for (int k = 0; k < L; k++) w[2 + k] = x.m[k];
    }
    int upload_sc() {
        HIPCHK(hipMemcpyAsync(d_sc, h_sc.data(), sizeof(uint32_t) * h_sc.size(), hipMemcpyHostToDevice, st));
        return SDPG_OK;
    }
    int upload_iarg() {
        HIPCHK(hipMemcpyAsync(d_iarg, h_iarg.data(), sizeof(int32_t) * nprob, hipMemcpyHostToDevice, st));
        return SDPG_OK;
    }
    static bool infeasible(int phase) { return phase == SDPG_PFEAS || phase == SDPG_NOINFO; }
    int grid_con() const { return std::max(1, std::min((ntasks + 3) / 4, 8 * 256)); }

    int gvec(const int *req, const uint32_t *mu, const uint32_t *beta, const int *phase) override {
        const int W = L + 2;
        for (int p = 0; p < nprob; p++) {
            if (!req[p]) continue;
            const mp<L> t = mul<L>(unpack_w(beta + (size_t)p * W), unpack_w(mu + (size_t)p * W));
            uint32_t w[L + 2];
            pack_w(t, w);
            put_planar(h_sc, sc_n, (int64_t)p * BSC_N + BSC_TARGET, w);
            h_iarg[p] = infeasible(phase[p]) ? 1 : 0;
        }
        int rc;
        if ((rc = begin(req)) || (rc = upload_sc()) || (rc = upload_iarg())) return rc;
        k_bb_gvec<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, d_iarg, PV{d_sc, sc_n}, fxs,
                                               fxs_n);
        k_fx_dspace<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, 0, fxa());
        k_bc_inner<L><<<grid_con(), 256, 0, st>>>(d_probs, d_tasks, ntasks, d_req, 0, fxa());
        return end(BT_GVEC);
    }
    int dz(const int *req, const int *phase) override {
        for (int p = 0; p < nprob; p++) h_iarg[p] = req[p] && infeasible(phase[p]) ? 1 : 0;
        int rc;
        if ((rc = begin(req)) || (rc = upload_iarg())) return rc;
        k_bb_dz<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, d_iarg, fxa());
        return end(BT_DZ);
    }
    int dx(const int *req) override {
        int rc;
        if ((rc = begin(req))) return rc;
        k_bb_dx<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, fxs, fxs_n);
        return end(BT_DX);
    }
    int update(const int *req, const uint32_t *alphaP, const uint32_t *alphaD) override {
        const int W = L + 2;
        for (int p = 0; p < nprob; p++) {
            if (!req[p]) continue;
            put_planar(h_sc, sc_n, (int64_t)p * BSC_N + BSC_ALPHAP, alphaP + (size_t)p * W);
            put_planar(h_sc, sc_n, (int64_t)p * BSC_N + BSC_ALPHAD, alphaD + (size_t)p * W);
        }
        int rc;
        if ((rc = begin(req)) || (rc = upload_sc())) return rc;
        k_bb_update<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, PV{d_sc, sc_n});
        return end(BT_UPDATE);
    }
    int factor_point(const int *req, int *info) override {
        int rc;
        if ((rc = begin(req))) return rc;
        HIPCHK(hipMemsetAsync(d_binfo, 0, sizeof(int32_t) * nprob, st));
        k_bb_factor<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, d_binfo, fxs, fxs_n);
        return end(BT_FACTOR, info, d_binfo, sizeof(int32_t) * nprob);
    }
    int fetch_bo(int kind) { return end(kind, h_bo.data(), d_bo, sizeof(uint32_t) * h_bo.size()); }
    mp<L> bo_at(int t, int slot) const { return get_planar(h_bo, bo_n, (int64_t)t * BO_N + slot); }

    // out: per problem the ten selectors SDPG_IP_XZ .. SDPG_IP_TRZ (packed, selector - 1)
    int inner(const int *req, uint32_t *out) override {
        int rc;
        if ((rc = begin(req))) return rc;
        k_bb_inner<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, PV{d_bo, bo_n});
        if ((rc = fetch_bo(BT_INNER))) return rc;
        const int W = L + 2;
        static const int slot_of[10] = {BO_XZ, BO_XDZ, BO_DXZ, BO_DXDZ, BO_CX, BO_CDX, BO_BY, BO_BDY, BO_TRX, BO_TRZ};
        for (int p = 0; p < nprob; p++) {
            if (!req[p]) continue;
            const int t0 = first_blk[p], nb = eng[p]->nsdp;
            for (int q = 0; q < 10; q++) {
                const int sl = slot_of[q];
                mp<L> v = bo_at(t0, sl);
                if (sl != BO_BY && sl != BO_BDY)
                    for (int l = 1; l < nb; l++) v = add<L>(v, bo_at(t0 + l, sl));
                pack_w(v, out + ((size_t)p * 10 + q) * W);
            }
        }
        return SDPG_OK;
    }
    // out: per problem max |Rp|, max |Rd|, C.X, b.y (packed)
    int residuals(const int *req, uint32_t *out) override {
        int rc;
        if ((rc = begin(req))) return rc;
        k_fx_dspace<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, 1, fxa());
        k_bc_inner<L><<<grid_con(), 256, 0, st>>>(d_probs, d_tasks, ntasks, d_req, 1, fxa());
        k_bb_resid<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, PV{d_bo, bo_n}, fxa());
        if ((rc = fetch_bo(BT_RESID))) return rc;
        const int W = L + 2;
        for (int p = 0; p < nprob; p++) {
            if (!req[p]) continue;
            const int t0 = first_blk[p], nb = eng[p]->nsdp;
            mp<L> rd = bo_at(t0, BO_MAXRD), cx = bo_at(t0, BO_CX);
            for (int l = 1; l < nb; l++) {
                const mp<L> v = bo_at(t0 + l, BO_MAXRD);
                if (cmp_abs<L>(v, rd) > 0) rd = v;
                cx = add<L>(cx, bo_at(t0 + l, BO_CX));
            }
            pack_w(bo_at(t0, BO_MAXRP), out + ((size_t)p * 4 + 0) * W);
            pack_w(rd, out + ((size_t)p * 4 + 1) * W);
            pack_w(cx, out + ((size_t)p * 4 + 2) * W);
            pack_w(bo_at(t0, BO_BY), out + ((size_t)p * 4 + 3) * W);
        }
        return SDPG_OK;
    }
    int min_eigen(const int *req, int which, uint32_t *out) override {
        int rc;
        if ((rc = begin(req))) return rc;
        k_bb_mineig<L><<<grid_blk, 256, 0, st>>>(d_probs, d_blks, (int)blks.size(), d_req, which, PV{d_bo, bo_n}, fxs,
                                                 fxs_n);
        if ((rc = fetch_bo(BT_EIG))) return rc;
        const int W = L + 2;
        for (int p = 0; p < nprob; p++) {
            if (!req[p]) continue;
            const int t0 = first_blk[p], nb = eng[p]->nsdp;
            mp<L> mn = from_double<L>(1.0E50);
            for (int l = 0; l < nb; l++) {
                const mp<L> v = bo_at(t0 + l, BO_EIG);
                if (cmp<L>(v, mn) < 0) mn = v;
            }
            pack_w(mn, out + (size_t)p * W);
        }
        return SDPG_OK;
    }
    int wait_contexts(int on) override {
        wait_ctx = on != 0;
        return SDPG_OK;
    }
    int timers(double *ms, int n) override {
        for (int k = 0; k < n && k < BT_N; k++) ms[k] = tm[k];
        return SDPG_OK;
    }

    int newton(const int *req, int *info) override {
        if (hipSetDevice(device) != hipSuccess) return SDPG_EHIP;
        // the contexts' own work (point updates, right-hand sides) is complete before the batch reads it
        if (wait_ctx)
            for (Engine<L> *e : eng)
                if (hipStreamSynchronize(e->st) != hipSuccess) return SDPG_EHIP;
        HIPCHK(hipMemcpyAsync(d_req, req, sizeof(int32_t) * nprob, hipMemcpyHostToDevice, st));
        // a problem that only solves keeps the info of its factorisation (zero: it is still in the run)
        HIPCHK(hipMemsetAsync(d_info, 0, sizeof(int32_t) * nprob, st));
        HIPCHK(hipEventRecord(e0, st));
        if (use_graph) {
            if (!gexec) {
                hipGraph_t graph = nullptr;
                HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                launch();
                HIPCHK(hipStreamEndCapture(st, &graph));
                hipError_t ge = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIPCHK(ge);
            }
            HIPCHK(hipGraphLaunch(gexec, st));
        } else {
            launch();
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipMemcpyAsync(info, d_info, sizeof(int32_t) * nprob, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
            last_ms = ms;
            tm[BT_NEWTON] += ms;
        }
        return SDPG_OK;
    }
    int set_graph(int on) override {
        use_graph = on != 0;
        return SDPG_OK;
    }
    int timing(double *ms) override {
        *ms = last_ms;
        return SDPG_OK;
    }
};

template <int L>
static int batch_create_L(EngineBase *const *es, int n, BatchBase **out) {
    auto *b = new Batch<L>();
    int rc = b->create(es, n);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return SDPG_OK;
}

}  // namespace sdpg
