// engine_impl.h — MI355X Newton engine for SDPA-GMP: kernels and Engine<L>.
//
// Included by engine.hip (the C-ABI, no instantiation) and by engine_inst.hip, which is
// compiled once per limb count (-DSDPG_INST_L=<L>) so that the instantiations build in
// parallel.
//
// Hot path of one IPM iteration (reference Newton::Mehrotra,
// sdpa_newton.cpp:1375-1414) on the GPU in limb-planar fixed-limb
// multiple precision: right-hand side (compute_rMat + Make_gVec), Schur
// complement (Make_bMat: F1/F2/F3 + LP), Cholesky (Rpotrf), triangular solves
// (Rtrsv x2), dZ (multi_plusToA) and dX (triple product + getSymmetrize), plus
// the residuals of Residuals::compute.  Dense products go through the GEMM
// engine of gemm.hip (plain limb MAD or Ozaki INT8 MFMA).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "chol_fx.h"
#include "engine_common.h"
#include "gemm.h"
#include "host_la.h"
#include "limbs.h"
#include "mpx.h"
#include "sdpg.h"

using namespace mpx;

namespace sdpg {

// last error message of the calling thread (defined in engine.hip)
void set_err(const std::string &s);

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            set_err(std::string("HIP error ") + hipGetErrorString(e_) + " at " #x);            \
            return SDPG_EHIP;                                                                  \
        }                                                                                      \
    } while (0)

static inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
static inline double host_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------------ kernels

template <int L>
__global__ void k_unpack(const uint32_t *__restrict__ src, PV dst, int64_t off, int64_t cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const uint32_t *s = src + i * (L + 2);
    mp<L> x;
    x.e = (int32_t)s[0];
    x.s = s[1] & 1u;
#pragma unroll
    for (int k = 0; k < L; k++) x.m[k] = s[2 + k];
    if (x.m[L - 1] == 0) x = zero<L>();
    stm<L>(dst, off + i, x);
}

template <int L>
__global__ void k_pack(PV src, int64_t off, int64_t cnt, uint32_t *__restrict__ dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    mp<L> x = ldm<L>(src, off + i);
    uint32_t *d = dst + i * (L + 2);
    d[0] = (uint32_t)x.e;
    d[1] = x.s;
#pragma unroll
    for (int k = 0; k < L; k++) d[2 + k] = x.m[k];
}

template <int L>
__global__ void k_zero(PV dst, int64_t off, int64_t cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    stm<L>(dst, off + i, zero<L>());
}

// out = a*X + b*Y  (a, b scalars; Y may be null => b ignored)
template <int L>
__global__ void k_lincomb(PV out, PV X, PV Y, mp<L> a, mp<L> b, int useY, int64_t cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    mp<L> r = mul<L>(a, ldm<L>(X, i));
    if (useY) r = add<L>(r, mul<L>(b, ldm<L>(Y, i)));
    stm<L>(out, i, r);
}

// out = X + s*Y    (Raxpy form: y += s*x as in mplapack/Raxpy_omp.cpp)
template <int L>
__global__ void k_axpy(PV out, PV X, PV Y, mp<L> s, int64_t cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    stm<L>(out, i, add<L>(ldm<L>(X, i), mul<L>(s, ldm<L>(Y, i))));
}

// Lal::getSymmetrize on one n x n block at offset off (sdpa_linear.cpp:457-482):
// lower = (lower + upper) / 2, upper = lower
template <int L>
__global__ void k_symmetrize(PV A, int64_t off, int n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * n) return;
    int i = (int)(t % n), j = (int)(t / n);
    if (i <= j) return;  // thread owns the strictly-lower (i, j)
    const mp<L> half = ldexp2<L>(from_int<L>(1), -1);
    mp<L> lo = ldm<L>(A, off + i + (int64_t)n * j);
    mp<L> up = ldm<L>(A, off + j + (int64_t)n * i);
    mp<L> v = mul<L>(half, add<L>(lo, up));
    stm<L>(A, off + i + (int64_t)n * j, v);
    stm<L>(A, off + j + (int64_t)n * i, v);
}

// Constraint descriptor of block (i, l)
struct Desc {
    int32_t kind;  // 0 absent, 1 sparse, 2 dense
    int32_t n;
    int64_t off;  // sparse: first entry index; dense: first value index
    int32_t nnz;  // sparse stored entries
    int32_t effect;
    int64_t blkoff;  // offset of block l inside a dense linear space
};

// block-tree sum of one mp value per thread (blockDim.x a power of two, LDS planes of blockDim.x)
template <int L>
__device__ mp<L> block_sum(mp<L> v, uint32_t *sh) {
    const int t = threadIdx.x, nt = blockDim.x;
    lds_st<L>(sh, nt, t, v);
    __syncthreads();
    for (int h = nt / 2; h >= 1; h >>= 1) {
        if (t < h) lds_st<L>(sh, nt, t, add<L>(lds_ld<L>(sh, nt, t), lds_ld<L>(sh, nt, t + h)));
        __syncthreads();
    }
    return lds_ld<L>(sh, nt, 0);
}

// one wave-wide sum (64 lanes, LDS planes of the block)
template <int L>
__device__ mp<L> wave_sum64(mp<L> v, uint32_t *sh) {
    const int t = threadIdx.x, lane = t & 63;
    lds_st<L>(sh, blockDim.x, t, v);
    __syncthreads();
    for (int h = 32; h >= 1; h >>= 1) {
        if (lane < h) lds_st<L>(sh, blockDim.x, t, add<L>(lds_ld<L>(sh, blockDim.x, t), lds_ld<L>(sh, blockDim.x, t + h)));
        __syncthreads();
    }
    return lds_ld<L>(sh, blockDim.x, t - lane);
}

// For each constraint c: ip_c = A_c . D (SDP blocks + LP), one wave per constraint (lanes stride
// over the stored entries, then a tree sum), then
//       mode 0: out_c = -ip + R_c   (Make_gVec: gVec = -(A.DLS1) + primalVec)
//       mode 1: out_c =  b_c - ip   (Residuals: primalVec = b - A.X)
//       mode 2: out_c =  ip          (objective C.X)
// Sparse blocks follow the _OPENMP branch of sdpa_linear.cpp:235-254 per entry (v D_ii, or
// v (D_ij + D_ji)), dense blocks Rdot over n^2 values; the summation order is the tree's.
template <int L>
__global__ __launch_bounds__(256) void k_con_inner(int ncon, int nsdp, const Desc *desc, const int32_t *ri,
                                                   const int32_t *ci, PV sv, PV dv, const int64_t *lp_ptr,
                                                   const int32_t *lp_pos, PV lp_val, int64_t lpbase, PV D, PV R,
                                                   PV out, int mode, int whole_block) {
    __shared__ uint32_t sh[(L + 1) * 256];
    // whole_block: the workgroup works on constraint blockIdx.x; otherwise one wave per constraint
    const int lane = whole_block ? threadIdx.x : (threadIdx.x & 63);
    const int G = whole_block ? blockDim.x : 64;
    const int c = whole_block ? blockIdx.x : blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    mp<L> acc = zero<L>();
    if (c < ncon) {
        for (int l = 0; l < nsdp; l++) {
            Desc d = desc[(int64_t)c * nsdp + l];
            if (d.kind == 1) {
                for (int t = lane; t < d.nnz; t += G) {
                    int64_t e = d.off + t;
                    int i = ri[e], j = ci[e];
                    mp<L> v = ldm<L>(sv, e);
                    mp<L> temp;
                    if (i == j) {
                        temp = mul<L>(v, ldm<L>(D, d.blkoff + i + (int64_t)d.n * j));
                    } else {
                        temp = add<L>(ldm<L>(D, d.blkoff + i + (int64_t)d.n * j),
                                      ldm<L>(D, d.blkoff + j + (int64_t)d.n * i));
                        temp = mul<L>(temp, v);
                    }
                    acc = add<L>(acc, temp);
                }
            } else if (d.kind == 2) {
                int64_t nn = (int64_t)d.n * d.n;
                for (int64_t k = lane; k < nn; k += G)
                    acc = fma_acc<L>(acc, ldm<L>(dv, d.off + k), ldm<L>(D, d.blkoff + k));
            }
        }
        for (int64_t t = lp_ptr[c] + lane; t < lp_ptr[c + 1]; t += G)
            acc = fma_acc<L>(acc, ldm<L>(lp_val, t), ldm<L>(D, lpbase + lp_pos[t]));
    }
    mp<L> ret = whole_block ? block_sum<L>(acc, sh) : wave_sum64<L>(acc, sh);
    if (c >= ncon || lane != 0) return;
    mp<L> o;
    if (mode == 0) {
        o = add<L>(neg<L>(ret), ldm<L>(R, c));
    } else if (mode == 1) {
        o = sub<L>(ldm<L>(R, c), ret);
    } else {
        o = ret;
    }
    stm<L>(out, c, o);
}

// Scatter sum_i s_i A_i into the dense linear space (multi_plusToA, sdpa_dataset.cpp:488-493).
// pos_ptr/pos_con/pos_ent: per position, the sparse entries touching it (ordered by constraint);
// dense constraints of the position's block: dcon_ptr/dcon (per block), value offset dcon_off.
// mode 0 (dZ):  out = -acc + (useR ? R : 0)        (compute_DzMat)
// mode 1 (Rd):  out = -acc + Cfull - Z             (Residuals::compute)
template <int L>
__global__ void k_scatter(int64_t npos, const int64_t *pos_ptr, const int32_t *pos_con, const int64_t *pos_ent,
                          PV sv, PV svals_dense, const int32_t *pos_blk, const int64_t *pos_in_blk,
                          const int64_t *dcon_ptr, const int32_t *dcon, const int64_t *dcon_off, PV s, PV R, int useR,
                          PV Cfull, PV Z, PV out, int mode) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npos) return;
    mp<L> acc = zero<L>();
    for (int64_t q = pos_ptr[t]; q < pos_ptr[t + 1]; q++) {
        mp<L> v = mul<L>(ldm<L>(sv, pos_ent[q]), ldm<L>(s, pos_con[q]));
        acc = add<L>(acc, v);
    }
    int b = pos_blk[t];
    if (b >= 0) {
        int64_t k = pos_in_blk[t];
        for (int64_t q = dcon_ptr[b]; q < dcon_ptr[b + 1]; q++) {
            acc = add<L>(acc, mul<L>(ldm<L>(s, dcon[q]), ldm<L>(svals_dense, dcon_off[q] + k)));
        }
    }
    mp<L> o;
    if (mode == 0) {
        o = neg<L>(acc);
        if (useR) o = add<L>(o, ldm<L>(R, t));
    } else {
        o = sub<L>(sub<L>(ldm<L>(Cfull, t), ldm<L>(Z, t)), acc);
    }
    stm<L>(out, t, o);
}

// max |x_i| over cnt elements: per-block partial maxima (one per workgroup)
template <int L>
__global__ void k_maxabs(PV X, int64_t cnt, PV part) {
    __shared__ uint32_t sbuf[256 * (L + 2)];
    mp<L> best = zero<L>();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x) {
        mp<L> v = absv<L>(ldm<L>(X, i));
        if (cmp_abs<L>(v, best) > 0) best = v;
    }
    uint32_t *me = sbuf + threadIdx.x * (L + 2);
    me[0] = (uint32_t)best.e;
    me[1] = best.s;
#pragma unroll
    for (int k = 0; k < L; k++) me[2 + k] = best.m[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        mp<L> b = zero<L>();
        for (int t = 0; t < (int)blockDim.x; t++) {
            uint32_t *o = sbuf + t * (L + 2);
            mp<L> v;
            v.e = (int32_t)o[0];
            v.s = o[1];
#pragma unroll
            for (int k = 0; k < L; k++) v.m[k] = o[2 + k];
            if (cmp_abs<L>(v, b) > 0) b = v;
        }
        stm<L>(part, blockIdx.x, b);
    }
}

// ---- Schur complement ----

// F_c = A_c * invZ for sparse A_c by the symmetric row expansion (Lal::multiply(D,S,D),
// sdpa_linear.cpp:877-955): row r of F = sum over expansion entries (r, q, v) of v * invZ[q, :]
// grid: x over n*n outputs, y over the constraint list
template <int L>
__global__ void k_spmm_rows(int n, const int64_t *exp_rowptr, const int32_t *exp_col, const int64_t *exp_ent, PV sv,
                            PV invZ, int64_t zoff, PV F, int64_t fstride, const int32_t *sp_list) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y;
    if (t >= (int64_t)n * n) return;
    int r = (int)(t % n), c = (int)(t / n);
    int64_t base = (int64_t)k * (n + 1);
    mp<L> acc = zero<L>();
    for (int64_t q = exp_rowptr[base + r]; q < exp_rowptr[base + r + 1]; q++) {
        acc = fma_acc<L>(acc, ldm<L>(sv, exp_ent[q]), ldm<L>(invZ, zoff + exp_col[q] + (int64_t)n * c));
    }
    stm<L>(F, (int64_t)sp_list[k] * fstride + r + (int64_t)n * c, acc);
}

struct Pair {
    int32_t i, j;    // owner i (formula of i), partner j
    int32_t slot;    // workspace slot of F_i / G_i
    int32_t form;    // 1, 2, 3
};

// F1/F2 pairs whose partner A_j is dense, served from the Schur inner-product GEMM:
// value = Cs[idx] = A_j . G_i (sdpa_newton.cpp:897 with the dense branch of getInnerProduct)
struct GvPair {
    int32_t i, j;
    int64_t idx;  // slot_i + M * rank_j in the GEMM output
};

// B[i,j] += value, B[j,i] += value (i != j) — one thread per pair of one block
template <int L>
__device__ __forceinline__ void bacc(PV B, int m, int i, int j, const mp<L> &v) {
    int64_t a = i + (int64_t)m * j;
    stm<L>(B, a, add_z<L>(ldm<L>(B, a), v));
    if (i != j) {
        int64_t b = j + (int64_t)m * i;
        stm<L>(B, b, add_z<L>(ldm<L>(B, b), v));
    }
}

// A (block) . D by one thread (sparse: the _OPENMP branch of sdpa_linear.cpp:235-254; dense: Rdot)
template <int L>
__device__ mp<L> blk_inner(const Desc &d, const int32_t *ri, const int32_t *ci, PV sv, PV dv, PV D, int64_t doff) {
    mp<L> ret = zero<L>();
    if (d.kind == 1) {
        for (int t = 0; t < d.nnz; t++) {
            int64_t e = d.off + t;
            int i = ri[e], j = ci[e];
            mp<L> v = ldm<L>(sv, e);
            mp<L> temp;
            if (i == j) {
                temp = mul<L>(v, ldm<L>(D, doff + i + (int64_t)d.n * j));
            } else {
                temp = add<L>(ldm<L>(D, doff + i + (int64_t)d.n * j), ldm<L>(D, doff + j + (int64_t)d.n * i));
                temp = mul<L>(temp, v);
            }
            ret = add_z<L>(ret, temp);
        }
    } else if (d.kind == 2) {
        int64_t nn = (int64_t)d.n * d.n;
        for (int64_t k = 0; k < nn; k++) ret = fma_acc<L>(ret, ldm<L>(dv, d.off + k), ldm<L>(D, doff + k));
    }
    return ret;
}

// F1: value = A_j . G_i (Newton::calF1, sdpa_newton.cpp:897)
// F2: sparse A_j: sum v * Rdot(n, X[alpha,:], F_i[:,beta]) (+mirror)  (Newton::calF2, sdpa_newton.cpp:899-937)
//     dense A_j : A_j . G_i with G_i = X F_i
// F3: Newton::calF3 (sdpa_newton.cpp:939-1005)
template <int L>
__global__ __launch_bounds__(64) void k_pairs(int npairs, const Pair *pairs, int l, int nsdp, const Desc *desc,
                                              const int32_t *ri, const int32_t *ci, PV sv, PV dv, PV W,
                                              int64_t wstride, PV Fw, PV X, PV invZ, int64_t boff, PV B, int m) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    Pair p = pairs[t];
    Desc dj = desc[(int64_t)p.j * nsdp + l];
    int n = dj.n;
    mp<L> value = zero<L>();
    if (p.form == 1 || (p.form == 2 && dj.kind == 2)) {
        value = blk_inner<L>(dj, ri, ci, sv, dv, W, (int64_t)p.slot * wstride);
    } else if (p.form == 2) {
        int64_t foff = (int64_t)p.slot * wstride;
        for (int e = 0; e < dj.nnz; e++) {
            int64_t ee = dj.off + e;
            int alpha = ri[ee], beta = ci[ee];
            mp<L> v1 = ldm<L>(sv, ee);
            mp<L> v2 = zero<L>();
            for (int k = 0; k < n; k++)
                v2 = fma_acc<L>(v2, ldm<L>(X, boff + alpha + (int64_t)n * k), ldm<L>(Fw, foff + k + (int64_t)n * beta));
            value = add_z<L>(value, mul<L>(v1, v2));
            if (alpha != beta) {
                v2 = zero<L>();
                for (int k = 0; k < n; k++)
                    v2 = fma_acc<L>(v2, ldm<L>(X, boff + beta + (int64_t)n * k),
                                    ldm<L>(Fw, foff + k + (int64_t)n * alpha));
                value = add_z<L>(value, mul<L>(v1, v2));
            }
        }
    } else {
        Desc di = desc[(int64_t)p.i * nsdp + l];
        mp<L> local = zero<L>();
        for (int i1 = 0; i1 < dj.nnz; i1++) {
            int64_t e1 = dj.off + i1;
            int alpha = ri[e1], beta = ci[e1];
            mp<L> v1 = ldm<L>(sv, e1);
            mp<L> sum = zero<L>();
            for (int i2 = 0; i2 < di.nnz; i2++) {
                int64_t e2 = di.off + i2;
                int gamma = ri[e2], delta = ci[e2];
                mp<L> v2 = ldm<L>(sv, e2);
                mp<L> plu = mul<L>(mul<L>(v2, ldm<L>(invZ, boff + delta + (int64_t)n * beta)),
                                   ldm<L>(X, boff + alpha + (int64_t)n * gamma));
                sum = add_z<L>(sum, plu);
                if (gamma != delta) {
                    plu = mul<L>(mul<L>(v2, ldm<L>(invZ, boff + gamma + (int64_t)n * beta)),
                                 ldm<L>(X, boff + alpha + (int64_t)n * delta));
                    sum = add_z<L>(sum, plu);
                }
            }
            local = add_z<L>(local, mul<L>(v1, sum));
            if (alpha == beta) continue;
            sum = zero<L>();
            for (int i2 = 0; i2 < di.nnz; i2++) {
                int64_t e2 = di.off + i2;
                int gamma = ri[e2], delta = ci[e2];
                mp<L> v2 = ldm<L>(sv, e2);
                mp<L> plu = mul<L>(mul<L>(v2, ldm<L>(invZ, boff + delta + (int64_t)n * alpha)),
                                   ldm<L>(X, boff + beta + (int64_t)n * gamma));
                sum = add_z<L>(sum, plu);
                if (gamma != delta) {
                    plu = mul<L>(mul<L>(v2, ldm<L>(invZ, boff + gamma + (int64_t)n * alpha)),
                                 ldm<L>(X, boff + beta + (int64_t)n * delta));
                    sum = add_z<L>(sum, plu);
                }
            }
            local = add_z<L>(local, mul<L>(v1, sum));
        }
        value = local;
    }
    bacc<L>(B, m, p.i, p.j, value);
}

// Packed lower triangle of count n x n matrices (column-major packed index
// t = b n - b (b - 1) / 2 + (a - b) for a >= b): dst[t] = src[a, b] (sum = 0), or
// src[a, a] on the diagonal and src[a, b] + src[b, a] below it (sum = 1).  For symmetric A,
// A . G = sum_t packA[t] packG[t] with the second form for G, which halves the inner-product length.
// One workgroup per T x T tile (a-block >= b-block) of one matrix: every thread reads its own element
// (a row of the tile is T consecutive words of each limb plane), and the mirror tile, read the same
// way, is transposed through LDS, so the transposed operand costs no strided HBM reads.  T = 32
// (128-byte rows) while the mirror tile fits 64 KB of LDS, else 16.
template <int L>
struct SymPackTile {
    static constexpr int T = (L + 1) * 32 * 33 * 4 <= 65536 ? 32 : 16;
};
template <int L>
__global__ __launch_bounds__(SymPackTile<L>::T *SymPackTile<L>::T) void k_sym_pack(
    PV src, int64_t soff, int64_t sstride, PV dst, int64_t doff, int64_t dstride, int n, int sum) {
    constexpr int T = SymPackTile<L>::T;
    __shared__ uint32_t t2[L + 1][T][T + 1];
    // lower tile (ia, ib), ia >= ib, from the linear tile index
    int r = blockIdx.x, ia = 0;
    while (r > ia) {
        r -= ia + 1;
        ia++;
    }
    const int ib = r;
    const int c = blockIdx.y, tx = threadIdx.x % T, ty = threadIdx.x / T;
    const int64_t so = soff + (int64_t)c * sstride;
    const int a0 = ia * T, b0 = ib * T;
    const int a = a0 + tx, b = b0 + ty;
    const bool in = a < n && b < n;
    mp<L> v = zero<L>();
    if (in) v = ldm<L>(src, so + a + (int64_t)n * b);
    if (sum) {
        const int bb = b0 + tx, aa = a0 + ty;  // t2[ty][tx] = src[b0 + tx, a0 + ty]
        if (aa < n && bb < n) {
            mp<L> w = ldm<L>(src, so + bb + (int64_t)n * aa);
#pragma unroll
            for (int q = 0; q < L; q++) t2[q][ty][tx] = w.m[q];
            t2[L][ty][tx] = (uint32_t)((w.e << 1) | (int32_t)(w.s & 1u));
        }
        __syncthreads();
    }
    if (!in || a < b) return;
    if (sum && a != b) {  // src[b, a] = t2[tx][ty]
        mp<L> w;
#pragma unroll
        for (int q = 0; q < L; q++) w.m[q] = t2[q][tx][ty];
        int32_t h = (int32_t)t2[L][tx][ty];
        w.e = h >> 1;
        w.s = (uint32_t)h & 1u;
        v = add<L>(v, w);
    }
    int64_t pt = (int64_t)b * n - (int64_t)b * (b - 1) / 2 + (a - b);
    stm<L>(dst, doff + (int64_t)c * dstride + pt, v);
}

// launch shape of k_sym_pack for n x n matrices: lower T x T tiles of T^2 threads
template <int L>
inline unsigned sym_pack_tiles(int n) {
    constexpr int T = SymPackTile<L>::T;
    unsigned nb = (unsigned)((n + T - 1) / T);
    return nb * (nb + 1) / 2;
}
template <int L>
constexpr int sym_pack_threads() {
    return SymPackTile<L>::T * SymPackTile<L>::T;
}

// dense-partner F1/F2 pairs: B[i,j] += Cs[idx] (and the mirror)
template <int L>
__global__ void k_pairs_gv(int npairs, const GvPair *pairs, PV Cs, PV B, int m) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    GvPair p = pairs[t];
    bacc<L>(B, m, p.i, p.j, ldm<L>(Cs, p.idx));
}

// LP part (compute_bMat_dense_LP, sdpa_newton.cpp:1190-1225): pairs within one LP variable
struct LPPair {
    int32_t i, j;
    int32_t lp;        // LP index
    int32_t pad;
    int64_t ei, ej;    // value indices
};
template <int L>
__global__ void k_pairs_lp(int npairs, const LPPair *pairs, PV lpv, PV X, PV invZ, int64_t lpbase, PV B, int m) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    LPPair p = pairs[t];
    mp<L> v = mul<L>(ldm<L>(X, lpbase + p.lp), ldm<L>(invZ, lpbase + p.lp));
    v = mul<L>(v, ldm<L>(lpv, p.ei));
    v = mul<L>(v, ldm<L>(lpv, p.ej));
    bacc<L>(B, m, p.i, p.j, v);
}

// ---- F1 inner products in block fixed point ----
// Constraint values of the sparse blocks, fixed once per problem: one exponent per (i, l)
// block (its largest entry), offset-binary L-limb integers (W = 32L - 2) per entry, and the
// offset sum RA = sum a' (L+1 limbs).  One thread per descriptor.
template <int L>
__global__ void k_fix_entries(int ndesc, const Desc *desc, PV sv, uint32_t *afx, int64_t sn, int32_t *ea,
                              uint32_t *ra) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ndesc) return;
    Desc d = desc[t];
    int32_t emax = mpx::EZERO;
    if (d.kind == 1)
        for (int e = 0; e < d.nnz; e++) {
            mp<L> x = ldm<L>(sv, d.off + e);
            if (!is_zero<L>(x) && x.e > emax) emax = x.e;
        }
    if (emax == mpx::EZERO) emax = 0;
    uint32_t rs[L + 1];
#pragma unroll
    for (int q = 0; q <= L; q++) rs[q] = 0;
    if (d.kind == 1)
        for (int e = 0; e < d.nnz; e++) {
            uint32_t a[L];
            fx_convert<L>(ldm<L>(sv, d.off + e), emax, 2, a);
            uint32_t c = 0;
#pragma unroll
            for (int q = 0; q < L; q++) {
                uint64_t s = (uint64_t)rs[q] + a[q] + c;
                rs[q] = (uint32_t)s;
                c = (uint32_t)(s >> 32);
                afx[(int64_t)q * sn + d.off + e] = a[q];
            }
            rs[L] += c;
        }
    ea[t] = emax;
#pragma unroll
    for (int q = 0; q <= L; q++) ra[(int64_t)q * ndesc + t] = rs[q];
}

// G_s -> symmetric fixed form Gs[p,q] = G[p,q] + G[q,p] (p != q), G[p,p] on the diagonal,
// exact integer sums of W = 32L - 3 fixed values (one exponent per slot), offset-binary.
// One workgroup per slot (slot = blockIdx.x + s0).
template <int L>
__global__ __launch_bounds__(256) void k_fix_G(int n, int s0, PV G, int64_t wstride, uint32_t *gfx, int64_t gplane,
                                               int32_t *eg) {
    __shared__ int32_t smax[256];
    int s = s0 + blockIdx.x;
    int64_t base = (int64_t)s * wstride;
    int64_t nn = (int64_t)n * n;
    int32_t em = mpx::EZERO;
    for (int64_t t = threadIdx.x; t < nn; t += blockDim.x) {
        mp<L> x = ldm<L>(G, base + t);
        if (!is_zero<L>(x) && x.e > em) em = x.e;
    }
    smax[threadIdx.x] = em;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && smax[threadIdx.x + w] > smax[threadIdx.x]) smax[threadIdx.x] = smax[threadIdx.x + w];
        __syncthreads();
    }
    int32_t emax = smax[0] == mpx::EZERO ? 0 : smax[0];
    for (int64_t t = threadIdx.x; t < nn; t += blockDim.x) {
        int p = (int)(t % n), q = (int)(t / n);
        if (p > q) continue;
        uint32_t a[L], b[L];
        fx_convert<L>(ldm<L>(G, base + p + (int64_t)n * q), emax, 3, a, false);
        if (p != q) {
            fx_convert<L>(ldm<L>(G, base + q + (int64_t)n * p), emax, 3, b, false);
            uint32_t c = 0;
#pragma unroll
            for (int k = 0; k < L; k++) {
                uint64_t sum = (uint64_t)a[k] + b[k] + c;
                a[k] = (uint32_t)sum;
                c = (uint32_t)(sum >> 32);
            }
        }
        a[L - 1] += 0x80000000u;
#pragma unroll
        for (int k = 0; k < L; k++) {
            gfx[(int64_t)k * gplane + base + p + (int64_t)n * q] = a[k];
            gfx[(int64_t)k * gplane + base + q + (int64_t)n * p] = a[k];
        }
    }
    if (threadIdx.x == 0) eg[s] = emax;
}

// value = A_j . G_i = sum over the stored entries (p, q, v) of A_j of v * Gs_i[p, q]
// (Newton::calF1 -> getInnerProduct(Sparse, Dense), sdpa_linear.cpp:235-254), exact in
// offset-binary fixed point: multi-limb MADs into 96-bit columns, one normalisation per pair.
template <int L>
__global__ __launch_bounds__(128) void k_pairs_fx(int npairs, const Pair *pairs, int l, int nsdp, const Desc *desc,
                                                  const int32_t *ri, const int32_t *ci, const uint32_t *afx,
                                                  int64_t sn, const int32_t *ea, const uint32_t *ra, int ndesc,
                                                  const uint32_t *gfx, int64_t gplane, int64_t wstride,
                                                  const int32_t *eg, PV B, int m) {
    constexpr int S0 = (L >= 2) ? L - 2 : 0;
    constexpr int NC = 2 * L - 1 - S0;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    Pair p = pairs[t];
    int di = p.j * nsdp + l;
    Desc dj = desc[di];
    int n = dj.n;
    int64_t gb = (int64_t)p.slot * wstride;
    uint64_t lo[NC];
    uint32_t hi[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        lo[c] = 0;
        hi[c] = 0;
    }
    uint32_t rg[L + 1];
#pragma unroll
    for (int q = 0; q <= L; q++) rg[q] = 0;
    for (int e = 0; e < dj.nnz; e++) {
        int64_t ee = dj.off + e;
        int64_t gi = gb + ri[ee] + (int64_t)n * ci[ee];
        uint32_t a[L], g[L];
#pragma unroll
        for (int q = 0; q < L; q++) {
            a[q] = afx[(int64_t)q * sn + ee];
            g[q] = gfx[(int64_t)q * gplane + gi];
        }
        uint32_t c = 0;
#pragma unroll
        for (int q = 0; q < L; q++) {
            uint64_t s = (uint64_t)rg[q] + g[q] + c;
            rg[q] = (uint32_t)s;
            c = (uint32_t)(s >> 32);
        }
        rg[L] += c;
        mpx::col_products<L, S0>(a, g, lo, hi);
    }
    uint32_t ras[L + 1];
#pragma unroll
    for (int q = 0; q <= L; q++) ras[q] = ra[(int64_t)q * ndesc + di];
    int32_t E = ea[di] + eg[p.slot] - (32 * L - 2) - (32 * L - 3);
    mp<L> v = fx_finish<L>(lo, hi, dj.nnz, ras, rg, E);
    bacc<L>(B, m, p.i, p.j, v);
}

// mirror: B[j,i] = B[i,j] is maintained by bacc; nothing else needed.

// ---- Cholesky (Rpotrf lower), blocked right-looking ----
// info <- v unless a smaller failing pivot is already there (0 = none); several workgroups may report
__device__ __forceinline__ void info_min(int *info, int v) {
    int cur = __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (cur == 0 || v < cur) {
        if (__hip_atomic_compare_exchange_strong(info, &cur, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            break;
    }
}

// Block b of a Cholesky factor (n x n in A, blocks of NBK) by one workgroup, inf the pivot info seen at its
// start (the same in every thread): the block's factor L_kk (written by k_potrf_col to Linv at b NBK^2, leading
// dimension nb) is copied into A, then replaced in Linv by its inverse: with U = L_kk diag(1 / L_cc) (unit lower
// triangular; reciprocal pivots in rdiag), inv(U) by the row operations Y_i <- Y_i - U_ij Y_j (j = 0 .. nb-2,
// i > j) on Y = I, then inv(L_kk)_ic = Y_ic / L_ii.  When a pivot failed (inf > 0), only a block before the
// failing pivot's block is copied, and no inverse is formed.  LDS: U and Y as (L+1) planes of NBK^2 words each.
template <int L>
__device__ void blkinv_block(PV A, int64_t off, int ld, int n, int NBK, PV rdiag, int64_t roff, PV Linv, int b,
                             int inf, uint32_t *shm) {
    const int tid = threadIdx.x, nt = blockDim.x, k = b * NBK, nb = min(NBK, n - k);
    if (inf != 0 && k >= ((inf - 1) / NBK) * NBK) return;
    const int P = NBK * NBK;
    const int64_t lo = (int64_t)b * P;
    uint32_t *sU = shm, *sY = shm + (size_t)(L + 1) * P;
    for (int t = tid; t < nb * nb; t += nt) {
        const int i = t % nb, j = t / nb;
        if (i < j) continue;
        const mp<L> l = ldm<L>(Linv, lo + t);
        stm<L>(A, off + (k + i) + (int64_t)ld * (k + j), l);
        if (inf == 0) lds_st<L>(sU, P, t, i > j ? mul<L>(l, ldm<L>(rdiag, roff + k + j)) : zero<L>());
    }
    if (inf != 0) return;
    for (int t = tid; t < nb * nb; t += nt) {
        const int i = t % nb, j = t / nb;
        if (i < j) lds_st<L>(sU, P, t, zero<L>());
        lds_st<L>(sY, P, t, i == j ? from_int<L>(1) : zero<L>());
    }
    __syncthreads();
    for (int j = 0; j < nb - 1; j++) {
        const int n1 = nb - 1 - j;
        for (int t = tid; t < n1 * (j + 1); t += nt) {
            const int i = j + 1 + t % n1, c = t / n1;
            const mp<L> y = lds_ld<L>(sY, P, i + nb * c);
            lds_st<L>(sY, P, i + nb * c, sub<L>(y, mul<L>(lds_ld<L>(sU, P, i + nb * j), lds_ld<L>(sY, P, j + nb * c))));
        }
        __syncthreads();
    }
    for (int t = tid; t < nb * nb; t += nt) {
        const int i = t % nb, c = t / nb;
        stm<L>(Linv, lo + t, i >= c ? mul<L>(ldm<L>(rdiag, roff + k + i), lds_ld<L>(sY, P, t)) : zero<L>());
    }
}

// k_blkinv: blkinv_block for blocks b0, b0 + 1, ... (one workgroup each)
template <int L>
__global__ __launch_bounds__(256) void k_blkinv(PV A, int64_t off, int ld, int n, int NBK, PV rdiag, int64_t roff,
                                                 PV Linv, const int *info, int b0) {
    extern __shared__ uint32_t shm[];
    __shared__ int s_inf;
    if (threadIdx.x == 0) s_inf = __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    blkinv_block<L>(A, off, ld, n, NBK, rdiag, roff, Linv, b0 + (int)blockIdx.x, s_inf, shm);
}

// LDS words of k_potrf_col for a workgroup of R panel rows
template <int L>
constexpr size_t potrf_col_words(int nb, int R) {
    const size_t ldv = (size_t)nb + R, r1 = R ? R : 1;
    return ((size_t)ldv * nb + R) * ColElim<L>::SW + 2 * (size_t)(L + 1) * nb + (size_t)(nb + 1) * ldv + ldv + r1 +
           (size_t)nb * ldv;
}
// items (entries updated by the steps) and entries (loaded at set-up) of a workgroup of R panel rows
constexpr int potrf_col_items(int nb, int R) {
    return nb * (nb - 1) / 2 + R * nb;
}
constexpr int potrf_col_entries(int nb, int R) {
    return nb * (nb + 1) / 2 + R * (nb + 1);
}
// threads of a k_potrf_col workgroup: one lane per item and per entry, and two more waves (the normalised
// pivots and their prefix products, a lane per row); 0 when R rows do not fit 1024 threads or one wave
constexpr int potrf_col_threads(int nb, int R) {
    if (nb + R > 64) return 0;
    const int it = potrf_col_items(nb, R), en = potrf_col_entries(nb, R);
    const int w = (it + 63) / 64 + 2 > (en + 63) / 64 ? (it + 63) / 64 + 2 : (en + 63) / 64;
    return 64 * w <= 1024 ? 64 * w : 0;
}

// k_potrf_col: eliminates the block column at (k, k) of the n x n matrix A (Rpotrf lower): the nb x nb
// diagonal block (nb <= 32) and the r rows below it, in block fixed point (chol_fx.h, ColElim: one division-free
// step per pivot, two fixed-point products per entry).  Workgroup w takes the panel rows [w R, w R + R) and
// runs the diagonal block's elimination itself (the same arithmetic in every workgroup, so they exchange
// nothing); it writes its panel rows' factor entries over A.  Workgroup 0 also writes the block's factor
// (lower triangle, leading dimension nb) to Lblk at (k / NBK) NBK^2, not over A (other workgroups may still be
// reading the block), and the reciprocal pivots 1 / L_jj; k_blkinv moves the block into A and inverts it.  Each
// thread holds at most one entry at set-up and one item in the steps (blockDim.x = potrf_col_threads(nb, R));
// the last two waves hold no item and form the normalised pivots, their prefix products and the row scales
// while the others update.  The reference factors
// the diagonal block with a square root and a division per pivot and solves the panel
// (mplapack/Rpotrf2.cpp:86-97, mplapack/Rtrsm.cpp:263-285); the factor is the same to rounding.  info: the
// first failing pivot (1-based, as Rpotrf2's), or for an entry out of the fixed-point range (impossible for a
// positive definite matrix) the first pivot not yet confirmed; once info is set, later launches do nothing.
template <int L>
__global__ __launch_bounds__(1024) void k_potrf_col(PV A, int64_t off, int ld, int k, int nb, int r, int R, PV rdiag,
                                                    int64_t roff, PV Lblk, int NBK, int *info) {
    extern __shared__ uint32_t shm[];
    __shared__ int s_flag, s_ovf;
    const int tid = threadIdx.x, nt = blockDim.x, wg = blockIdx.x;
    const int64_t lbo = (int64_t)(k / NBK) * NBK * NBK;
    const int p0 = wg * R, Rw = r > p0 ? min(R, r - p0) : 0;
    const int LDV = nb + Rw, R1 = Rw ? Rw : 1;
    using CE = ColElim<L>;
    uint32_t *sV = shm, *sG = sV + (size_t)(LDV * nb + Rw) * CE::SW, *sRh = sG + (size_t)(L + 1) * nb;
    int32_t *hist = (int32_t *)(sRh + (size_t)(L + 1) * nb), *f0 = hist + (nb + 1) * LDV, *pm = f0 + LDV,
            *Fpre = pm + R1;
    const CE ce{sV, sG, sRh, nullptr, hist, f0, pm, Fpre, nb, Rw, LDV};
    // element (i, j) of the block column (rows 0..nb-1 the block, nb + p this workgroup's panel row p)
    auto at = [&](int i, int j) -> int64_t {
        const int gi = i < nb ? i : nb + p0 + (i - nb);
        return off + (int64_t)(k + gi) + (int64_t)ld * (k + j);
    };
    // this thread's entry (set-up) and item (steps)
    const int nent = potrf_col_entries(nb, Rw), nit = ce.nitems();
    int ecol = 0;
    const int erow = tid < nent ? ce.entry_code(tid, ecol) : 0;
    mp<L> ev;
    if (tid < nent) ev = ldm<L>(A, ecol >= 0 ? at(erow, ecol) : at(erow, nb + p0 + (erow - nb)));
    const bool has_item = tid < nit;
    const int32_t item = has_item ? ce.item_code(tid) : 0;
    if (tid == 0) {
        s_flag = __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_ovf = 0;
    }
    for (int t = tid; t < (nb + 1) * LDV; t += nt) hist[t] = mpx::FX_HZERO;
    __syncthreads();
    bool ok = s_flag == 0;
    // row scales: the block rows' from their diagonals, the panel rows' from their diagonals and entries
    if (ok && tid < nent) {
        if (erow < nb && erow == ecol) ce.block_scale(erow, ev);
        if (ecol < 0) ce.panel_scale0(erow - nb, ev);
    }
    __syncthreads();
    if (ok && tid < nent && erow >= nb && ecol >= 0) ce.panel_rowmax(erow - nb, ecol, ev);
    __syncthreads();
    if (ok && tid < Rw) ce.panel_scale(tid);
    __syncthreads();
    if (ok && tid < nent) {
        const bool o = ecol >= 0 ? ce.load_entry(erow, ecol, ev) : ce.load_pdiag(erow - nb, ev);
        if (o) atomicOr(&s_ovf, erow < nb ? 1 : 2);
    }
    __syncthreads();
    // the last two waves hold no item: at step j one forms the normalised pivot Vh_j (into sRh) and, lane r for
    // row r, the row scale f_r^(j) (into Fpre), the other the prefix product gamma_j = gamma_{j-1} Vh_{j-1}
    // (into sG), each well within the time of an update
    const bool vwave = tid >= nt - 128 && tid < nt - 64, gwave = tid >= nt - 64;
    const int vl = tid - (nt - 128);  // lane in the pivot wave: its row (LDV <= 64)
    int32_t fr = vwave && vl < LDV ? f0[vl] : 0;
    if (ok) {
        mp<L> g = from_int<L>(1);  // gamma_j in the last wave
        for (int j = 0; j < nb; j++) {
            // the pivot (the same for every thread: scalar registers, and so is the failure exit), its row
            // exponent and the flags, read together
            int32_t a[CE::N];
            ce.ld(j + LDV * j, a);
            const int32_t hjj = hist[j * LDV + j], ovf0 = s_ovf;
#pragma unroll
            for (int q = 0; q < CE::N; q++) a[q] = __builtin_amdgcn_readfirstlane(a[q]);
            if ((ovf0 & 1) || mpx::fx_nonpos<CE::N>(a)) {
                if (tid == 0) info_min(info, k + j + 1);
                ok = false;
                break;
            }
            if (gwave) {
                if (j > 0) g = mul<L>(g, ce.gld(sRh, j - 1));
                if (tid == nt - 64) ce.gst(sG, j, g);
            }
            if (vwave && vl < LDV) {
                if (j > 0) fr += mpx::fx_delta(hist[(j - 1) * LDV + vl]);
                Fpre[j * LDV + vl] = fr;  // f_r^(j)
            }
            if (j == nb - 1) break;
            const int32_t dj = __builtin_amdgcn_readfirstlane(mpx::fx_delta(hjj));
            int ov = 0;
            if (has_item) ov = ce.update_item(j, item, a, dj);
            if (vwave) {
                const mp<L> vh = fx_to_mp<L>(a, -2 * dj);
                if (tid == nt - 128) ce.gst(sRh, j, vh);
            }
            if (ov) atomicOr(&s_ovf, ov);
            __syncthreads();
        }
    }
    if (ok) {
        // rho_c and 1 / L_cc (gamma_c in sG and the row scales in Fpre from the steps)
        __syncthreads();
        mp<L> rinv;
        if (tid < nb) rinv = ce.finish_pivot(tid, ce.pivot_mp(tid));  // rho_c into sRh
        __syncthreads();
        if (s_ovf & 2) {
            if (tid == 0) info_min(info, k + nb + 1);
        } else {
            for (int t = tid; t < Rw * nb; t += nt) {
                const int p = t % Rw, c = t / Rw;
                stm<L>(A, at(nb + p, c), ce.factor_entry(nb + p, c));
            }
        }
        if (wg == 0) {
            if (tid < nb) stm<L>(rdiag, roff + k + tid, rinv);
            for (int t = tid; t < nb * nb; t += nt) {
                const int i = t % nb, c = t / nb;
                if (i >= c) stm<L>(Lblk, lbo + t, ce.factor_entry(i, c));
            }
        }
    }
}

// dst[i + ldd*j] = src[i + lds*j] for an r x c sub-matrix
template <int L>
__global__ void k_copy2d(PV dst, int64_t doff, int ldd, PV src, int64_t soff, int lds, int r, int c) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)r * c) return;
    int i = (int)(t % r), j = (int)(t / r);
    stm<L>(dst, doff + i + (int64_t)ldd * j, ldm<L>(src, soff + i + (int64_t)lds * j));
}

// Lal::solveSystems (sdpa_linear.cpp:683-695): L y = g then L^T x = y, blocked by the
// factorisation's diagonal blocks.  Diagonal step: x_k = inv(L_kk) x_k (trans: inv(L_kk)^T x_k)
// with the stored block inverse; off-diagonal step: the remaining entries of x lose the block's
// contribution.  One workgroup per diagonal step; 8 threads share each output's dot product.
template <int L>
__global__ __launch_bounds__(256) void k_trsv_diag(PV linv, int64_t loff, int nb, PV x, int64_t xo, int trans) {
    __shared__ uint32_t sx[(L + 1) * 32];
    __shared__ uint32_t sp[(L + 1) * 256];
    const int tid = threadIdx.x, i = tid >> 3, q = tid & 7;
    if (tid < nb) lds_st<L>(sx, 32, tid, ldm<L>(x, xo + tid));
    __syncthreads();
    mp<L> acc = zero<L>();
    if (i < nb)
        for (int j = q; j < nb; j += 8) {
            if (trans ? (j < i) : (j > i)) continue;
            mp<L> a = ldm<L>(linv, loff + (trans ? (j + (int64_t)nb * i) : (i + (int64_t)nb * j)));
            acc = add<L>(acc, mul<L>(a, lds_ld<L>(sx, 32, j)));
        }
    lds_st<L>(sp, 256, tid, acc);
    __syncthreads();
    for (int h = 4; h >= 1; h >>= 1) {
        if (q < h) lds_st<L>(sp, 256, tid, add<L>(lds_ld<L>(sp, 256, tid), lds_ld<L>(sp, 256, tid + h)));
        __syncthreads();
    }
    if (q == 0 && i < nb) stm<L>(x, xo + i, lds_ld<L>(sp, 256, tid));
}

// forward (trans 0): x[r] -= sum_t A[r, k0+t] x[k0+t] for r >= k0+nb;
// backward (trans 1): x[r] -= sum_t A[k0+t, r] x[k0+t] for r < k0
template <int L>
__global__ __launch_bounds__(256) void k_trsv_update(PV A, int m, int k0, int nb, PV x, int trans) {
    __shared__ uint32_t sx[(L + 1) * 32];
    __shared__ uint32_t sp[(L + 1) * 256];
    const int tid = threadIdx.x, q = tid & 7;
    const int r0 = trans ? 0 : k0 + nb;
    const int nrows = trans ? k0 : m - k0 - nb;
    const int r = blockIdx.x * 32 + (tid >> 3);
    if (tid < nb) lds_st<L>(sx, 32, tid, ldm<L>(x, k0 + tid));
    __syncthreads();
    mp<L> acc = zero<L>();
    if (r < nrows) {
        int row = r0 + r;
        for (int t = q; t < nb; t += 8) {
            int64_t idx = trans ? ((int64_t)(k0 + t) + (int64_t)m * row) : ((int64_t)row + (int64_t)m * (k0 + t));
            acc = add<L>(acc, mul<L>(ldm<L>(A, idx), lds_ld<L>(sx, 32, t)));
        }
    }
    lds_st<L>(sp, 256, tid, acc);
    __syncthreads();
    for (int h = 4; h >= 1; h >>= 1) {
        if (q < h) lds_st<L>(sp, 256, tid, add<L>(lds_ld<L>(sp, 256, tid), lds_ld<L>(sp, 256, tid + h)));
        __syncthreads();
    }
    if (q == 0 && r < nrows) {
        int row = r0 + r;
        stm<L>(x, row, sub<L>(ldm<L>(x, row), lds_ld<L>(sp, 256, tid)));
    }
}

// LP part of dX / r_zinv / x_rd_zinv : out = a*x*y*z style triple on diagonal entries
template <int L>
__global__ void k_lp_triple(PV out, PV A, PV Bm, PV C, int64_t base, int nlp) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlp) return;
    mp<L> v = mul<L>(ldm<L>(A, base + t), ldm<L>(Bm, base + t));
    v = mul<L>(v, ldm<L>(C, base + t));
    stm<L>(out, base + t, v);
}

// ---- reductions, matrix-vector products and point factorisation (Solutions / Jal / Lal) ----

// part[blockIdx.x] = sum over this block's share of x_i * y_i (y.p == nullptr: of x_i); when idx is
// given, element t is x[xo + idx[t]]
template <int L>
__global__ __launch_bounds__(256) void k_dot_part(PV x, int64_t xo, PV y, int64_t yo, const int64_t *idx, int64_t n,
                                                  PV part) {
    __shared__ uint32_t sh[(L + 1) * 256];
    mp<L> acc = zero<L>();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t ix = idx ? idx[i] : i;
        mp<L> a = ldm<L>(x, xo + ix);
        acc = add<L>(acc, y.p ? mul<L>(a, ldm<L>(y, yo + i)) : a);
    }
    mp<L> r = block_sum<L>(acc, sh);
    if (threadIdx.x == 0) stm<L>(part, blockIdx.x, r);
}

// y[yo + j] = sum_i A[ao + i + n j] x[xo + i] over i in [0, n) (tri 0), [j, n) (tri 1) or [0, j] (tri 2);
// one wave per output
template <int L>
__global__ __launch_bounds__(256) void k_gemv_t(PV A, int64_t ao, int n, int tri, PV x, int64_t xo, PV y, int64_t yo) {
    __shared__ uint32_t sh[(L + 1) * 256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = blockIdx.x * 4 + w;
    mp<L> acc = zero<L>();
    if (j < n) {
        int i0 = tri == 1 ? j : 0, i1 = tri == 2 ? j + 1 : n;
        for (int i = i0 + lane; i < i1; i += 64)
            acc = add<L>(acc, mul<L>(ldm<L>(A, ao + i + (int64_t)n * j), ldm<L>(x, xo + i)));
    }
    uint32_t *s = sh;  // per-wave tree inside the block planes
    lds_st<L>(s, 256, threadIdx.x, acc);
    __syncthreads();
    for (int h = 32; h >= 1; h >>= 1) {
        if (lane < h)
            lds_st<L>(s, 256, threadIdx.x, add<L>(lds_ld<L>(s, 256, threadIdx.x), lds_ld<L>(s, 256, threadIdx.x + h)));
        __syncthreads();
    }
    if (lane == 0 && j < n) stm<L>(y, yo + j, lds_ld<L>(s, 256, threadIdx.x));
}

// dst[do + j + n i] = src[so + i + n j]
template <int L>
__global__ void k_transpose(PV src, int64_t so, PV dst, int64_t dof, int n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * n) return;
    int i = (int)(t % n), j = (int)(t / n);
    stm<L>(dst, dof + j + (int64_t)n * i, ldm<L>(src, so + t));
}

// LP part of Jal::getInvChol / getInvCholAndInv (sdpa_jordan.cpp:140-171): 1/sqrt(x), 1/sqrt(z), 1/z
template <int L>
__global__ void k_lp_inverse(PV X, PV Z, PV icx, PV icz, PV iz, int64_t base, int nlp, int *info) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlp) return;
    mp<L> x = ldm<L>(X, base + t), z = ldm<L>(Z, base + t);
    if (sgn<L>(x) < 0 || sgn<L>(z) < 0) {
        atomicMax(info, 1);
        return;
    }
    mp<L> rx = mpx::rsqrt<L>(x), rz = mpx::rsqrt<L>(z);
    stm<L>(icx, base + t, rx);
    stm<L>(icz, base + t, rz);
    stm<L>(iz, base + t, mul<L>(rz, rz));
}

// LP part of Jal::getMinEigen(lMat, xMat) (sdpa_jordan.cpp:80-86): x * l * l per entry
template <int L>
__global__ void k_lp_triple_val(PV Xm, PV Lm, int64_t base, int nlp, PV out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlp) return;
    mp<L> l = ldm<L>(Lm, base + t);
    stm<L>(out, t, mul<L>(mul<L>(ldm<L>(Xm, base + t), l), l));
}

// ------------------------------------------------------------------ Engine

struct EngineBase {
    virtual ~EngineBase() {}
    int nlimb = 0, prec = 0, m = 0, nsdp = 0, nlp = 0, device = 0;
    std::vector<int> sizes;
    std::vector<int64_t> blkoff;
    int64_t Nd = 0;  // dense linear space elements
    bool committed = false;
    int gemm_mode = SDPG_GEMM_PLAIN, gemm_slices = 0;
    double timers[8] = {0};
    // staged input
    struct Blk {
        bool set = false, dense = false;
        std::vector<int> r, c;
        std::vector<uint32_t> v;  // packed mp
    };
    std::vector<Blk> blk;  // (m+1)*nsdp, index (con+1)*nsdp + l
    std::vector<std::vector<int>> lp_pos;
    std::vector<std::vector<uint32_t>> lp_val;  // per con (m+1)
    std::vector<uint32_t> bvec;
    std::vector<int> formula;

    virtual int commit() = 0;
    virtual int set_point(const uint32_t *X, const uint32_t *Z, const uint32_t *invZ, const uint32_t *y) = 0;
    virtual int residuals(uint32_t *nrp, uint32_t *nrd, uint32_t *objP, uint32_t *objD) = 0;
    virtual int gvec(int corrector, const uint32_t *mu, const uint32_t *beta, int phase) = 0;
    virtual int schur() = 0;
    virtual int potrf(int *info) = 0;
    virtual int potrs() = 0;
    virtual int dz(int phase) = 0;
    virtual int dx() = 0;
    virtual int factor_point(int *info) = 0;
    virtual int update(const uint32_t *alphaP, const uint32_t *alphaD) = 0;
    virtual int inner(int what, uint32_t *out) = 0;
    virtual int min_eigen(int which, uint32_t *out) = 0;
    virtual int get(int what, uint32_t *out) = 0;
    virtual int put(int what, const uint32_t *in) = 0;
    virtual int sync() = 0;
    virtual bool slices_ok(int s) const = 0;
    virtual int profile(int on, double *out, int n) = 0;
};

template <int L>
struct Engine : EngineBase {
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // device planar arrays
    struct Arr {
        uint32_t *p = nullptr;
        int64_t n = 0;
        PV pv() const { return PV{p, n}; }
    };
    Arr X, Z, invZ, Rd, rzinv, xrd, dX, dZ, W1, W2, DLS1, Cfull;
    Arr X0, Z0;  // initial point of an initial-point file (allocated on first put)
    Arr y, Rp, g, dy, b, rdiag, scal;
    Arr Bm;
    Arr sval, dval, lpval;  // constraint data
    Arr Fw, Gw;             // Schur workspaces
    Arr Csw;                // Schur inner products with dense partners (slots x dense constraints)
    Arr dpk;                // packed lower triangles of those dense constraints
    uint32_t *afx = nullptr, *rax = nullptr, *gfx = nullptr;  // fixed-point constraint values, G slots
    int32_t *eax = nullptr, *egx = nullptr;
    int64_t sn_total = 0, gplane = 0;
    int64_t wstride = 0, wslots = 0;
    // index arrays
    Desc *d_desc = nullptr, *d_cdesc = nullptr;
    int32_t *d_ri = nullptr, *d_ci = nullptr;
    int64_t *d_lp_ptr = nullptr, *d_clp_ptr = nullptr;
    int32_t *d_lp_pos = nullptr;
    int64_t lp_c_base = 0;  // C's LP values start
    // scatter index
    int64_t *d_pos_ptr = nullptr, *d_pos_ent = nullptr, *d_pos_in_blk = nullptr, *d_dcon_ptr = nullptr,
            *d_dcon_off = nullptr;
    int32_t *d_pos_con = nullptr, *d_pos_blk = nullptr, *d_dcon = nullptr;
    // Schur plan per block
    struct BlockPlan {
        std::vector<int> f1_sparse, f1_dense, f2_sparse, f2_dense;  // constraints needing F (and G)
        std::vector<int> slot_of;  // constraint -> slot (or -1)
        int nslots = 0;
        int32_t *d_slot_cons = nullptr;  // per slot: sparse expansion slot index (for spmm)
        int n_sp_slots = 0;
        int32_t *d_sp_list = nullptr;    // slots of sparse constraints, in order
        std::vector<int> sp_slots, dn_slots, g_slots;  // slots by kind; g_slots: need G
        std::vector<int> dn_cons;
        Pair *d_pairs = nullptr, *d_pairs_fx = nullptr;
        int npairs = 0, npairs_fx = 0;
        // dense partners: Cs = Gw^T [vec A_j] over the dense constraints of the block (GEMM, K = n^2)
        GvPair *d_pairs_gv = nullptr;
        int npairs_gv = 0, gv_M = 0, gv_N = 0;
        bool gv_lower = false;
        int64_t gv_boff = 0;  // dval offset of the first dense constraint of the block
        int64_t gv_pkoff = 0;  // offset of the block's packed dense constraints in dpk
        // runs of dense slots (first slot, dval offset, count) with consecutive slots and values
        std::vector<std::array<int64_t, 3>> dn_runs;
        int64_t *d_exp_rowptr = nullptr, *d_exp_ent = nullptr;
        int32_t *d_exp_col = nullptr, *d_exp_slot = nullptr;
    };
    std::vector<BlockPlan> plan;
    LPPair *d_lppairs = nullptr;
    int nlppairs = 0;
    int *d_info = nullptr;
    std::vector<void *> allocs;
    KProf kprof;
    GemmEngine<L> gemm;

    ~Engine() override {
        if (st) hipStreamSynchronize(st);
        for (void *p : allocs) hipFree(p);
        if (stage) hipFree(stage);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (st) hipStreamDestroy(st);
    }

    template <class T>
    int dalloc(T **p, int64_t count) {
        *p = nullptr;
        if (count <= 0) count = 1;
        hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
        if (e != hipSuccess) {
            set_err(std::string("hipMalloc failed: ") + hipGetErrorString(e));
            return SDPG_ENOMEM;
        }
        allocs.push_back(*p);
        return 0;
    }
    int aalloc(Arr &a, int64_t n) {
        a.n = n > 0 ? n : 1;
        return dalloc(&a.p, (int64_t)(L + 1) * a.n);
    }
    template <class T>
    int upload(T **p, const std::vector<T> &v) {
        int rc = dalloc(p, (int64_t)v.size());
        if (rc) return rc;
        if (!v.empty()) {
            HIPCHK(hipMemcpyAsync(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));  // the host vector may be released by the caller
        }
        return 0;
    }
    // staging buffer for packed transfers (grown on demand, reused afterwards)
    uint32_t *stage = nullptr;
    size_t stage_cap = 0;
    int ensure_stage(size_t words) {
        if (words <= stage_cap) return 0;
        HIPCHK(hipStreamSynchronize(st));
        if (stage) HIPCHK(hipFree(stage));
        stage = nullptr;
        size_t n = words + words / 4 + 1024;
        HIPCHK(hipMalloc((void **)&stage, sizeof(uint32_t) * n));
        stage_cap = n;
        return 0;
    }
    // packed host -> planar device (synchronous with respect to the host buffer)
    int put_packed(Arr &a, int64_t off, const uint32_t *src, int64_t cnt) {
        if (cnt <= 0) return 0;
        int rc = ensure_stage((size_t)(L + 2) * cnt);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipMemcpyAsync(stage, src, sizeof(uint32_t) * (L + 2) * cnt, hipMemcpyHostToDevice, st));
        k_unpack<L><<<cdiv(cnt, 256), 256, 0, st>>>(stage, a.pv(), off, cnt);
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    int get_packed(const Arr &a, int64_t off, uint32_t *dst, int64_t cnt) {
        if (cnt <= 0) return 0;
        int rc = ensure_stage((size_t)(L + 2) * cnt);
        if (rc) return rc;
        k_pack<L><<<cdiv(cnt, 256), 256, 0, st>>>(a.pv(), off, cnt, stage);
        HIPCHK(hipMemcpyAsync(dst, stage, sizeof(uint32_t) * (L + 2) * cnt, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    void zero_arr(Arr &a, int64_t off, int64_t cnt) {
        if (cnt > 0) k_zero<L><<<cdiv(cnt, 256), 256, 0, st>>>(a.pv(), off, cnt);
    }
    static mp<L> unpack1(const uint32_t *w) {
        mp<L> x;
        x.e = (int32_t)w[0];
        x.s = w[1] & 1u;
        for (int k = 0; k < L; k++) x.m[k] = w[2 + k];
        if (x.m[L - 1] == 0) x = zero<L>();
        return x;
    }
    static void pack1(const mp<L> &x, uint32_t *w) {
        w[0] = (uint32_t)x.e;
        w[1] = x.s;
        for (int k = 0; k < L; k++) w[2 + k] = x.m[k];
    }
    void tstart() { hipEventRecord(ev0, st); }
    int tstop(int slot) {
        hipEventRecord(ev1, st);
        HIPCHK(hipEventSynchronize(ev1));
        float ms = 0;
        hipEventElapsedTime(&ms, ev0, ev1);
        timers[slot] += ms;
        return 0;
    }

    int init() {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
        {
            int v = 0;
            if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
                lds_max = (size_t)v;
        }
        blkoff.resize(nsdp + 1);
        Nd = 0;
        for (int l = 0; l < nsdp; l++) {
            blkoff[l] = Nd;
            Nd += (int64_t)sizes[l] * sizes[l];
        }
        blkoff[nsdp] = Nd;
        Nd += nlp;
        blk.assign((size_t)(m + 1) * nsdp, Blk());
        lp_pos.assign(m + 1, {});
        lp_val.assign(m + 1, {});
        return gemm.init(st, &kprof);
    }

    int commit() override {
        const int W = L + 2;
        // ---- constraint data: sparse entries, dense values, descriptors ----
        std::vector<int32_t> ri, ci;
        std::vector<uint32_t> sv, dvv;
        std::vector<Desc> desc((size_t)(m + 1) * nsdp);
        int64_t sn = 0, dn = 0;
        // block-major: the dense values of one block are contiguous in constraint order (C first), so
        // the dense constraints of a block form one strided operand of the batched and Schur GEMMs
        for (int l = 0; l < nsdp; l++)
            for (int c = -1; c < m; c++) {
                Blk &B = blk[(size_t)(c + 1) * nsdp + l];
                Desc d{};
                d.n = sizes[l];
                d.blkoff = blkoff[l];
                if (!B.set) {
                    d.kind = 0;
                } else if (B.dense) {
                    d.kind = 2;
                    d.off = dn;
                    d.nnz = sizes[l] * sizes[l];
                    d.effect = d.nnz;
                    dvv.insert(dvv.end(), B.v.begin(), B.v.end());
                    dn += (int64_t)sizes[l] * sizes[l];
                } else {
                    d.kind = 1;
                    d.off = sn;
                    d.nnz = (int)B.r.size();
                    int eff = 0;
                    for (size_t t = 0; t < B.r.size(); t++) {
                        ri.push_back(B.r[t]);
                        ci.push_back(B.c[t]);
                        eff += (B.r[t] == B.c[t]) ? 1 : 2;
                    }
                    d.effect = eff;
                    sv.insert(sv.end(), B.v.begin(), B.v.end());
                    sn += (int64_t)B.r.size();
                }
                desc[(size_t)(c + 1) * nsdp + l] = d;
            }
        int rc;
        if ((rc = aalloc(sval, sn))) return rc;
        if ((rc = put_packed(sval, 0, sv.data(), sn))) return rc;
        if ((rc = aalloc(dval, dn))) return rc;
        if ((rc = put_packed(dval, 0, dvv.data(), dn))) return rc;
        if ((rc = upload(&d_ri, ri))) return rc;
        if ((rc = upload(&d_ci, ci))) return rc;
        std::vector<Desc> cdesc(desc.begin(), desc.begin() + nsdp), adesc(desc.begin() + nsdp, desc.end());
        if ((rc = upload(&d_cdesc, cdesc))) return rc;
        if ((rc = upload(&d_desc, adesc))) return rc;
        hdesc = adesc;
        // ---- LP data: constraints 0..m-1 then C ----
        std::vector<int64_t> lptr(m + 1, 0), clptr(2, 0);
        std::vector<int32_t> lpos;
        std::vector<uint32_t> lv;
        for (int c = 0; c < m; c++) {
            lptr[c] = (int64_t)lpos.size();
            for (size_t t = 0; t < lp_pos[c + 1].size(); t++) lpos.push_back(lp_pos[c + 1][t]);
            lv.insert(lv.end(), lp_val[c + 1].begin(), lp_val[c + 1].end());
        }
        lptr[m] = (int64_t)lpos.size();
        lp_c_base = (int64_t)lpos.size();
        clptr[0] = lp_c_base;
        for (size_t t = 0; t < lp_pos[0].size(); t++) lpos.push_back(lp_pos[0][t]);
        lv.insert(lv.end(), lp_val[0].begin(), lp_val[0].end());
        clptr[1] = (int64_t)lpos.size();
        if ((rc = upload(&d_lp_ptr, lptr))) return rc;
        if ((rc = upload(&d_clp_ptr, clptr))) return rc;
        if ((rc = upload(&d_lp_pos, lpos))) return rc;
        if ((rc = aalloc(lpval, (int64_t)lpos.size()))) return rc;
        if ((rc = put_packed(lpval, 0, lv.data(), (int64_t)lpos.size()))) return rc;
        // ---- scatter index over the dense linear space (A_i only, ordered by constraint) ----
        {
            std::vector<std::vector<std::pair<int, int64_t>>> lists(Nd);
            std::vector<int64_t> dcon_ptr(nsdp + 1, 0), dcon_off;
            std::vector<int32_t> dcon;
            for (int c = 0; c < m; c++)
                for (int l = 0; l < nsdp; l++) {
                    const Desc &d = desc[(size_t)(c + 1) * nsdp + l];
                    if (d.kind == 1) {
                        for (int t = 0; t < d.nnz; t++) {
                            int64_t e = d.off + t;
                            int i = ri[e], j = ci[e];
                            lists[blkoff[l] + i + (int64_t)sizes[l] * j].push_back({c, e});
                            if (i != j) lists[blkoff[l] + j + (int64_t)sizes[l] * i].push_back({c, e});
                        }
                    }
                }
            for (int l = 0; l < nsdp; l++) {
                dcon_ptr[l] = (int64_t)dcon.size();
                for (int c = 0; c < m; c++) {
                    const Desc &d = desc[(size_t)(c + 1) * nsdp + l];
                    if (d.kind == 2) {
                        dcon.push_back(c);
                        dcon_off.push_back(d.off);
                    }
                }
            }
            dcon_ptr[nsdp] = (int64_t)dcon.size();
            // LP positions
            for (int c = 0; c < m; c++)
                for (int64_t t = lptr[c]; t < lptr[c + 1]; t++) lists[blkoff[nsdp] + lpos[t]].push_back({c, -1 - t});
            std::vector<int64_t> pptr(Nd + 1, 0), pent, pinb(Nd, 0);
            std::vector<int32_t> pcon, pblk(Nd, -1);
            for (int64_t q = 0; q < Nd; q++) {
                pptr[q] = (int64_t)pcon.size();
                for (auto &pr : lists[q]) {
                    pcon.push_back(pr.first);
                    pent.push_back(pr.second);
                }
            }
            pptr[Nd] = (int64_t)pcon.size();
            for (int l = 0; l < nsdp; l++)
                for (int64_t k = 0; k < (int64_t)sizes[l] * sizes[l]; k++) {
                    pblk[blkoff[l] + k] = l;
                    pinb[blkoff[l] + k] = k;
                }
            // LP entries reference the lp value array: encode with a separate value view (offset by sn)
            // by copying LP values after the sparse values into a combined scatter-value array.
            bool has_lp_entries = false;
            for (auto &e : pent)
                if (e < 0) has_lp_entries = true;
            if (has_lp_entries) {
                // combined array = sval ++ lpval(A part)
                Arr comb;
                if ((rc = aalloc(comb, sn + lptr[m]))) return rc;
                std::vector<uint32_t> allv(sv);
                allv.insert(allv.end(), lv.begin(), lv.begin() + lptr[m] * W);
                if ((rc = put_packed(comb, 0, allv.data(), sn + lptr[m]))) return rc;
                for (auto &e : pent)
                    if (e < 0) e = sn + (-1 - e);
                scat_vals = comb;
            } else {
                scat_vals = sval;
            }
            if ((rc = upload(&d_pos_ptr, pptr))) return rc;
            if ((rc = upload(&d_pos_con, pcon))) return rc;
            if ((rc = upload(&d_pos_ent, pent))) return rc;
            if ((rc = upload(&d_pos_blk, pblk))) return rc;
            if ((rc = upload(&d_pos_in_blk, pinb))) return rc;
            if ((rc = upload(&d_dcon_ptr, dcon_ptr))) return rc;
            if ((rc = upload(&d_dcon, dcon))) return rc;
            if ((rc = upload(&d_dcon_off, dcon_off))) return rc;
        }
        // ---- full C as a dense linear space ----
        {
            std::vector<mp<L>> cf(Nd, zero<L>());
            for (int l = 0; l < nsdp; l++) {
                const Blk &B = blk[l];
                if (!B.set) continue;
                int n = sizes[l];
                if (B.dense) {
                    for (int64_t k = 0; k < (int64_t)n * n; k++) cf[blkoff[l] + k] = unpack1(&B.v[k * W]);
                } else {
                    for (size_t t = 0; t < B.r.size(); t++) {
                        mp<L> v = unpack1(&B.v[t * W]);
                        int i = B.r[t], j = B.c[t];
                        cf[blkoff[l] + i + (int64_t)n * j] = v;
                        cf[blkoff[l] + j + (int64_t)n * i] = v;
                    }
                }
            }
            for (size_t t = 0; t < lp_pos[0].size(); t++) cf[blkoff[nsdp] + lp_pos[0][t]] = unpack1(&lp_val[0][t * W]);
            std::vector<uint32_t> pk(Nd * W);
            for (int64_t q = 0; q < Nd; q++) pack1(cf[q], &pk[q * W]);
            if ((rc = aalloc(Cfull, Nd))) return rc;
            if ((rc = put_packed(Cfull, 0, pk.data(), Nd))) return rc;
        }
        // ---- work arrays ----
        Arr *dls[] = {&X, &Z, &invZ, &Rd, &rzinv, &xrd, &dX, &dZ, &W1, &W2, &DLS1};
        for (Arr *a : dls)
            if ((rc = aalloc(*a, Nd))) return rc;
        Arr *vec[] = {&y, &Rp, &g, &dy, &b, &rdiag};
        for (Arr *a : vec)
            if ((rc = aalloc(*a, m))) return rc;
        if ((rc = aalloc(scal, 4096))) return rc;
        if ((rc = aalloc(Bm, (int64_t)m * m))) return rc;
        if ((rc = dalloc(&d_info, 1))) return rc;
        for (Arr *a : dls) zero_arr(*a, 0, Nd);
        if ((rc = alloc_factor_space())) return rc;
        if (!bvec.empty() && (rc = put_packed(b, 0, bvec.data(), m))) return rc;
        // ---- Schur plan: pair rule (sdpa_newton.cpp:1052-1054) and workspace slots ----
        plan.assign(nsdp, BlockPlan());
        int64_t maxslots = 0, maxn2 = 1, maxcs = 0;
        for (int l = 0; l < nsdp; l++) {
            BlockPlan &P = plan[l];
            int n = sizes[l];
            maxn2 = std::max<int64_t>(maxn2, (int64_t)n * n);
            std::vector<int> cons;
            for (int c = 0; c < m; c++)
                if (desc[(size_t)(c + 1) * nsdp + l].kind) cons.push_back(c);
            P.slot_of.assign(m, -1);
            std::vector<Pair> pairs, pairs_fx;
            std::vector<GvPair> pairs_gv;
            std::vector<std::array<int, 4>> gvt;  // (i, j, slot_i, rank_j) of the dense-partner pairs
            std::vector<int> slot_con;
            // rank of each dense constraint of the block (its column in the Schur inner-product GEMM)
            std::vector<int> drank(m, -1);
            int ndense = 0;
            for (int c : cons)
                if (desc[(size_t)(c + 1) * nsdp + l].kind == 2) drank[c] = ndense++;
            for (int i : cons) {
                const Desc &di = desc[(size_t)(i + 1) * nsdp + l];
                int f = formula.empty() ? 1 : formula[(size_t)i * nsdp + l];
                if (f < 1 || f > 3) f = 1;
                // F2 (calF2, sdpa_newton.cpp:899-937) sums v * X[alpha,:] F[:,beta] per entry of A_j; that is
                // A_j . (X F_i), so F2 owners also get G_i = X F_i from the batched GEMM and share the F1
                // inner-product kernels (n^3 extra multiply-adds per F2 constraint instead of nnz_j * n per pair).
                bool needG = (f == 1 || f == 2);
                if (f == 1 || f == 2) {
                    P.slot_of[i] = P.nslots++;
                    slot_con.push_back(i);
                    if (di.kind == 1)
                        P.sp_slots.push_back(P.slot_of[i]);
                    else
                        P.dn_slots.push_back(P.slot_of[i]);
                    if (needG) P.g_slots.push_back(P.slot_of[i]);
                }
                for (int j : cons) {
                    const Desc &dj = desc[(size_t)(j + 1) * nsdp + l];
                    if (di.effect < dj.effect || (di.effect == dj.effect && i < j)) continue;
                    if ((f == 1 || f == 2) && dj.kind == 1)
                        pairs_fx.push_back(Pair{i, j, P.slot_of[i], f});
                    else if ((f == 1 || f == 2) && dj.kind == 2)
                        gvt.push_back({i, j, P.slot_of[i], drank[j]});
                    else
                        pairs.push_back(Pair{i, j, P.slot_of[i], f});
                }
            }
            // k_pairs runs one pair per thread: pairs in 8 x 8 tiles of (i, j), so that a wave's reads of
            // X / inv(Z) and its updates of B[i, j] and B[j, i] fall in 8 runs of 8 consecutive entries each way
            // (any order gives the same B: every pair adds to its own entries)
            std::stable_sort(pairs.begin(), pairs.end(), [](const Pair &x, const Pair &y) {
                const int64_t kx = ((int64_t)(x.i >> 3) << 32) | ((int64_t)(x.j >> 3) << 6) | ((x.i & 7) << 3) | (x.j & 7);
                const int64_t ky = ((int64_t)(y.i >> 3) << 32) | ((int64_t)(y.j >> 3) << 6) | ((y.i & 7) << 3) | (y.j & 7);
                return kx < ky;
            });
            P.npairs = (int)pairs.size();
            if ((rc = upload(&P.d_pairs, pairs))) return rc;
            P.npairs_fx = (int)pairs_fx.size();
            if ((rc = upload(&P.d_pairs_fx, pairs_fx))) return rc;
            P.npairs_gv = (int)gvt.size();
            if (P.npairs_gv) {
                // GEMM rows: the slots (G_s), columns: the dense constraints in rank order, whose values are
                // consecutive n^2 runs from gv_boff
                P.gv_M = P.nslots;
                P.gv_N = ndense;
                for (const auto &q : gvt) pairs_gv.push_back(GvPair{q[0], q[1], (int64_t)q[2] + (int64_t)q[3] * P.gv_M});
                P.gv_boff = -1;
                for (int c : cons) {
                    const Desc &dc = desc[(size_t)(c + 1) * nsdp + l];
                    if (dc.kind != 2) continue;
                    if (P.gv_boff < 0) P.gv_boff = dc.off;
                    if (dc.off != P.gv_boff + (int64_t)drank[c] * n * n) {
                        set_err("commit: dense constraint values are not contiguous per block");
                        return SDPG_ESTATE;
                    }
                }
                // the lower triangle suffices when every pair has slot_i >= rank_j
                P.gv_lower = true;
                for (const GvPair &q : pairs_gv)
                    if (q.idx % P.gv_M < q.idx / P.gv_M) P.gv_lower = false;
                maxcs = std::max<int64_t>(maxcs, (int64_t)P.gv_M * P.gv_N);
            }
            if ((rc = upload(&P.d_pairs_gv, pairs_gv))) return rc;
            P.dn_cons.clear();
            for (int s : P.dn_slots) P.dn_cons.push_back(slot_con[s]);
            P.dn_runs.clear();
            for (size_t t = 0; t < P.dn_slots.size(); t++) {
                int64_t off = desc[(size_t)(P.dn_cons[t] + 1) * nsdp + l].off;
                if (!P.dn_runs.empty()) {
                    auto &r = P.dn_runs.back();
                    if (P.dn_slots[t] == r[0] + r[2] && off == r[1] + r[2] * (int64_t)n * n) {
                        r[2]++;
                        continue;
                    }
                }
                P.dn_runs.push_back({(int64_t)P.dn_slots[t], off, 1});
            }
            // symmetric row expansion of sparse slots (for F = A invZ)
            std::vector<int64_t> rowptr;
            std::vector<int32_t> col, slotlist;
            std::vector<int64_t> ent;
            int k = 0;
            for (int s : P.sp_slots) {
                int c = slot_con[s];
                const Desc &d = desc[(size_t)(c + 1) * nsdp + l];
                std::vector<std::vector<std::pair<int, int64_t>>> rows(n);
                for (int t = 0; t < d.nnz; t++) {
                    int64_t e = d.off + t;
                    int i = ri[e], j = ci[e];
                    rows[i].push_back({j, e});
                    if (i != j) rows[j].push_back({i, e});
                }
                for (int r = 0; r < n; r++) {
                    rowptr.push_back((int64_t)col.size());
                    for (auto &pr : rows[r]) {
                        col.push_back(pr.first);
                        ent.push_back(pr.second);
                    }
                }
                rowptr.push_back((int64_t)col.size());  // sentinel of this slot (n+1 entries per slot)
                slotlist.push_back(k++);
            }
            P.n_sp_slots = k;
            if ((rc = upload(&P.d_exp_rowptr, rowptr))) return rc;
            if ((rc = upload(&P.d_exp_col, col))) return rc;
            if ((rc = upload(&P.d_exp_ent, ent))) return rc;
            if ((rc = upload(&P.d_exp_slot, slotlist))) return rc;
            std::vector<int32_t> spl(P.sp_slots.begin(), P.sp_slots.end());
            if ((rc = upload(&P.d_sp_list, spl))) return rc;
            maxslots = std::max<int64_t>(maxslots, P.nslots);
        }
        wstride = maxn2;
        wslots = std::max<int64_t>(maxslots, 1);
        if (maxcs && (rc = aalloc(Csw, maxcs))) return rc;
        {  // packed lower triangles of the dense constraints that feed the Schur inner-product GEMM
            int64_t tot = 0;
            for (int l = 0; l < nsdp; l++)
                if (plan[l].npairs_gv) {
                    plan[l].gv_pkoff = tot;
                    tot += (int64_t)sizes[l] * (sizes[l] + 1) / 2 * plan[l].gv_N;
                }
            if (tot) {
                if ((rc = aalloc(dpk, tot))) return rc;
                for (int l = 0; l < nsdp; l++) {
                    const BlockPlan &P = plan[l];
                    if (!P.npairs_gv) continue;
                    int n = sizes[l];
                    k_sym_pack<L><<<dim3(sym_pack_tiles<L>(n), P.gv_N), sym_pack_threads<L>(), 0, st>>>(
                        dval.pv(), P.gv_boff, (int64_t)n * n, dpk.pv(), P.gv_pkoff, (int64_t)n * (n + 1) / 2, n, 0);
                }
                HIPCHK(hipGetLastError());
            }
        }
        if ((rc = aalloc(Fw, wstride * wslots))) return rc;
        if ((rc = aalloc(Gw, wstride * wslots))) return rc;
        gplane = wstride * wslots;
        if ((rc = dalloc(&gfx, (int64_t)L * gplane))) return rc;
        if ((rc = dalloc(&egx, wslots))) return rc;
        {  // fixed-point constraint values for the F1 inner products
            int ndesc = m * nsdp;
            sn_total = std::max<int64_t>(sn, 1);
            if ((rc = dalloc(&afx, (int64_t)L * sn_total))) return rc;
            if ((rc = dalloc(&eax, std::max(ndesc, 1)))) return rc;
            if ((rc = dalloc(&rax, (int64_t)(L + 1) * std::max(ndesc, 1)))) return rc;
            if (ndesc)
                k_fix_entries<L><<<cdiv(ndesc, 64), 64, 0, st>>>(ndesc, d_desc, sval.pv(), afx, sn_total, eax, rax);
            HIPCHK(hipGetLastError());
        }
        // LP pairs
        {
            std::vector<LPPair> lpp;
            for (int q = 0; q < nlp; q++) {
                std::vector<std::pair<int, int64_t>> cl;
                for (int c = 0; c < m; c++)
                    for (int64_t t = lptr[c]; t < lptr[c + 1]; t++)
                        if (lpos[t] == q) cl.push_back({c, t});
                for (size_t a = 0; a < cl.size(); a++)
                    for (size_t bb = a; bb < cl.size(); bb++)
                        lpp.push_back(LPPair{cl[a].first, cl[bb].first, q, 0, cl[a].second, cl[bb].second});
            }
            nlppairs = (int)lpp.size();
            if ((rc = upload(&d_lppairs, lpp))) return rc;
        }
        HIPCHK(hipStreamSynchronize(st));
        // release staging memory
        blk.clear();
        blk.shrink_to_fit();
        committed = true;
        return 0;
    }
    Arr scat_vals;

    int set_point(const uint32_t *Xp, const uint32_t *Zp, const uint32_t *iZp, const uint32_t *yp) override {
        int rc;
        if (Xp && (rc = put_packed(X, 0, Xp, Nd))) return rc;
        if (Zp && (rc = put_packed(Z, 0, Zp, Nd))) return rc;
        if (iZp && (rc = put_packed(invZ, 0, iZp, Nd))) return rc;
        if (yp && (rc = put_packed(y, 0, yp, m))) return rc;
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }

    // per-block dense products via the GEMM engine
    int block_gemm(Arr &C, const Arr &A, const Arr &B, int l) {
        int n = sizes[l];
        int64_t o = blkoff[l];
        return gemm.gemm(n, n, n, A.pv(), o, n, false, B.pv(), o, n, false, C.pv(), o, n, GEMM_SET, 1, 0, 0, 0,
                         gemm_mode, gemm_slices);
    }
    // out = A*B*C per block (ns_jordan_triple_product); tmp receives A*B
    int triple(Arr &out, const Arr &A, const Arr &B, const Arr &C, Arr &tmp) {
        int rc;
        for (int l = 0; l < nsdp; l++) {
            if ((rc = block_gemm(tmp, A, B, l))) return rc;
            if ((rc = block_gemm(out, tmp, C, l))) return rc;
        }
        if (nlp) k_lp_triple<L><<<cdiv(nlp, 128), 128, 0, st>>>(out.pv(), A.pv(), B.pv(), C.pv(), blkoff[nsdp], nlp);
        return 0;
    }

    int maxabs(const Arr &A, int64_t cnt, mp<L> &out) {
        int nb = (int)std::min<int64_t>(64, std::max<int64_t>(1, cdiv(cnt, 256)));
        k_maxabs<L><<<nb, 256, 0, st>>>(A.pv(), cnt, scal.pv());
        std::vector<uint32_t> h((size_t)nb * (L + 2));
        int rc = get_packed(scal, 0, h.data(), nb);
        if (rc) return rc;
        out = zero<L>();
        for (int i = 0; i < nb; i++) {
            mp<L> v = unpack1(&h[(size_t)i * (L + 2)]);
            if (cmp_abs<L>(v, out) > 0) out = v;
        }
        return 0;
    }

    int residuals(uint32_t *nrp, uint32_t *nrd, uint32_t *objP, uint32_t *objD) override {
        tstart();
        // Rp = b - A.X
        k_con_inner<L><<<cdiv(m, 4), 256, 0, st>>>(m, nsdp, d_desc, d_ri, d_ci, sval.pv(), dval.pv(), d_lp_ptr,
                                                   d_lp_pos, lpval.pv(), blkoff[nsdp], X.pv(), b.pv(), Rp.pv(), 1, 0);
        // Rd = C - Z - sum y_i A_i
        k_scatter<L><<<cdiv(Nd, 128), 128, 0, st>>>(Nd, d_pos_ptr, d_pos_con, d_pos_ent, scat_vals.pv(), dval.pv(),
                                                     d_pos_blk, d_pos_in_blk, d_dcon_ptr, d_dcon, d_dcon_off, y.pv(),
                                                     Rd.pv(), 0, Cfull.pv(), Z.pv(), Rd.pv(), 1);
        int rc = tstop(5);
        if (rc) return rc;
        mp<L> v;
        if (nrp) {
            if ((rc = maxabs(Rp, m, v))) return rc;
            pack1(v, nrp);
        }
        if (nrd) {
            if ((rc = maxabs(Rd, Nd, v))) return rc;
            pack1(v, nrd);
        }
        if (objP) {  // C . X
            k_con_inner<L><<<1, 256, 0, st>>>(1, nsdp, d_cdesc, d_ri, d_ci, sval.pv(), dval.pv(), d_clp_ptr,
                                              d_lp_pos, lpval.pv(), blkoff[nsdp], X.pv(), b.pv(), PV{scal.p, scal.n}, 2,
                                              1);
            std::vector<uint32_t> h(L + 2);
            if ((rc = get_packed(scal, 0, h.data(), 1))) return rc;
            memcpy(objP, h.data(), sizeof(uint32_t) * (L + 2));
        }
        if (objD) {
            if ((rc = dot(b, 0, &y, 0, nullptr, m, v))) return rc;
            pack1(v, objD);
        }
        return 0;
    }

    int gvec(int corrector, const uint32_t *mu, const uint32_t *beta, int phase) override {
        tstart();
        mp<L> target = mul<L>(unpack1(beta), unpack1(mu));
        mp<L> mone = from_int<L>(-1);
        int rc;
        // compute_rMat: r_zinv = (beta mu) invZ - X  [- dX dZ invZ]
        k_lincomb<L><<<cdiv(Nd, 256), 256, 0, st>>>(rzinv.pv(), invZ.pv(), X.pv(), target, mone, 1, Nd);
        if (corrector) {
            if ((rc = triple(DLS1, dX, dZ, invZ, W2))) return rc;
            k_axpy<L><<<cdiv(Nd, 256), 256, 0, st>>>(rzinv.pv(), rzinv.pv(), DLS1.pv(), mone, Nd);
        }
        // Make_gVec: DLS1 = r_zinv - x_rd_zinv (infeasible phases), gVec = -(A . DLS1) + Rp
        if (phase == SDPG_PFEAS || phase == SDPG_NOINFO) {
            if (!corrector)
                if ((rc = triple(xrd, X, Rd, invZ, W2))) return rc;
            k_axpy<L><<<cdiv(Nd, 256), 256, 0, st>>>(DLS1.pv(), rzinv.pv(), xrd.pv(), mone, Nd);
        } else {
            HIPCHK(hipMemcpyAsync(DLS1.p, rzinv.p, sizeof(uint32_t) * (L + 1) * Nd, hipMemcpyDeviceToDevice, st));
        }
        k_con_inner<L><<<cdiv(m, 4), 256, 0, st>>>(m, nsdp, d_desc, d_ri, d_ci, sval.pv(), dval.pv(), d_lp_ptr,
                                                   d_lp_pos, lpval.pv(), blkoff[nsdp], DLS1.pv(), Rp.pv(), g.pv(), 0, 0);
        return tstop(3);
    }

    int schur() override {
        tstart();
        zero_arr(Bm, 0, (int64_t)m * m);
        int rc;
        for (int l = 0; l < nsdp; l++) {
            BlockPlan &P = plan[l];
            int n = sizes[l];
            int64_t o = blkoff[l];
            // F_i = A_i invZ  (sparse: row expansion; dense: GEMM)
            if (P.n_sp_slots) {
                dim3 grid(cdiv((int64_t)n * n, 128), P.n_sp_slots);
                k_spmm_rows<L><<<grid, 128, 0, st>>>(n, P.d_exp_rowptr, P.d_exp_col, P.d_exp_ent, sval.pv(),
                                                     invZ.pv(), o, Fw.pv(), wstride, P.d_sp_list);
            }
            // dense constraints: F_s = A_c invZ, one batched GEMM per run of consecutive slots whose values
            // are consecutive in dval (one run when every constraint of the block is dense)
            for (const auto &r : P.dn_runs) {
                if ((rc = gemm.gemm(n, n, n, dval.pv(), r[1], n, false, invZ.pv(), o, n, false, Fw.pv(),
                                    r[0] * wstride, n, GEMM_SET, (int)r[2], (int64_t)n * n, 0, wstride, gemm_mode,
                                    gemm_slices, false, GEMM_KEEP_A)))
                    return rc;
            }
            // G_i = X F_i for slots needing it: one batched GEMM with X shared
            if (!P.g_slots.empty()) {
                // slots needing G are contiguous ranges in practice; batch over maximal runs
                size_t a = 0;
                while (a < P.g_slots.size()) {
                    size_t e = a + 1;
                    while (e < P.g_slots.size() && P.g_slots[e] == P.g_slots[e - 1] + 1) e++;
                    int s0 = P.g_slots[a], cnt = (int)(e - a);
                    if ((rc = gemm.gemm(n, n, n, X.pv(), o, n, false, Fw.pv(), (int64_t)s0 * wstride, n, false,
                                        Gw.pv(), (int64_t)s0 * wstride, n, GEMM_SET, cnt, 0, wstride, wstride,
                                        gemm_mode, gemm_slices)))
                        return rc;
                    a = e;
                }
            }
            if (P.npairs_fx) {
                // fixed symmetric G for the F1 slots (runs of consecutive slots)
                size_t a = 0;
                while (a < P.g_slots.size()) {
                    size_t e = a + 1;
                    while (e < P.g_slots.size() && P.g_slots[e] == P.g_slots[e - 1] + 1) e++;
                    k_fix_G<L><<<(int)(e - a), 256, 0, st>>>(n, P.g_slots[a], Gw.pv(), wstride, gfx, gplane, egx);
                    a = e;
                }
                k_pairs_fx<L><<<cdiv(P.npairs_fx, 128), 128, 0, st>>>(
                    P.npairs_fx, P.d_pairs_fx, l, nsdp, d_desc, d_ri, d_ci, afx, sn_total, eax, rax, m * nsdp, gfx,
                    gplane, wstride, egx, Bm.pv(), m);
            }
            if (P.npairs)
                k_pairs<L><<<cdiv(P.npairs, 64), 64, 0, st>>>(P.npairs, P.d_pairs, l, nsdp, d_desc, d_ri, d_ci,
                                                               sval.pv(), dval.pv(), Gw.pv(), wstride, Fw.pv(), X.pv(),
                                                               invZ.pv(), o, Bm.pv(), m);
            if (P.npairs_gv) {
                // inner products with the dense partners as one GEMM over packed lower triangles (F is no
                // longer needed, its slots take the packed G):  Cs[s, r] = packG_s . packA_r
                const int np = n * (n + 1) / 2;
                k_sym_pack<L><<<dim3(sym_pack_tiles<L>(n), P.gv_M), sym_pack_threads<L>(), 0, st>>>(
                    Gw.pv(), 0, wstride, Fw.pv(), 0, wstride, n, 1);
                if ((rc = gemm.gemm(P.gv_M, P.gv_N, np, Fw.pv(), 0, (int)wstride, true, dpk.pv(), P.gv_pkoff, np, false,
                                    Csw.pv(), 0, P.gv_M, GEMM_SET, 1, 0, 0, 0, gemm_mode, gemm_slices, P.gv_lower,
                                    GEMM_KEEP_B)))
                    return rc;
                k_pairs_gv<L><<<cdiv(P.npairs_gv, 64), 64, 0, st>>>(P.npairs_gv, P.d_pairs_gv, Csw.pv(), Bm.pv(), m);
            }
        }
        if (nlppairs)
            k_pairs_lp<L><<<cdiv(nlppairs, 64), 64, 0, st>>>(nlppairs, d_lppairs, lpval.pv(), X.pv(), invZ.pv(),
                                                             blkoff[nsdp], Bm.pv(), m);
        HIPCHK(hipGetLastError());
        return tstop(0);
    }
    std::vector<Desc> hdesc;  // host copy of the A descriptors (index c*nsdp + l)

    // Cholesky workspaces: inverses of every diagonal block (LinvB for the Schur matrix, LinvW for
    // the blocks of X and Z), the panel below the current block, pivots of X / Z blocks
    Arr LinvB, LinvW, rdW;
    Arr icX, icZ, Wc, Wt;        // inverse Cholesky factors of X and Z, factor / inverse workspaces
    Arr lq, lqold, lr, lw, ltmp;  // Lanczos vectors
    int64_t *d_diag_idx = nullptr;
    int64_t ndiag = 0, maxn = 1;
    size_t lds_max = 65536;
    int NBK = 32;
    int col_rmax = 4;
    bool chol_pairs = true;  // Ozaki form: Cholesky trailing updates in pairs of block columns (SDPG_CHOL_PAIRS=0: not)
    int col_r_fixed = 0;     // panel rows per block-column workgroup when > 0 (else chosen from the panel)

    int alloc_factor_space() {
        // diagonal block edge of the blocked Cholesky factorisations (SDPG_NBK = 8, 16 or 32 overrides)
        NBK = 32;
        if (const char *ev = getenv("SDPG_NBK")) {
            const int want = atoi(ev);
            if (want == 8 || want == 16 || want == 32) NBK = want;
        }
        while (NBK > 8 && std::max(potrf_col_words<L>(NBK, 4), (size_t)2 * (L + 1) * NBK * NBK) * 4 + 256 > lds_max)
            NBK /= 2;
        if (const char *ev = getenv("SDPG_CHOL_PAIRS")) chol_pairs = atoi(ev) != 0;
        // panel rows per workgroup of the block-column kernel fixed (SDPG_COL_R; tests of the wide workgroups)
        if (const char *ev = getenv("SDPG_COL_R")) col_r_fixed = std::max(0, atoi(ev));
        // panel rows per workgroup of the block-column kernel: as many as the LDS holds beside its static words
        col_rmax = 64;
        while (col_rmax > 1 && potrf_col_words<L>(NBK, col_rmax) * 4 + 256 > lds_max) col_rmax--;
        HIPCHK(hipFuncSetAttribute((const void *)k_potrf_col<L>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(potrf_col_words<L>(NBK, col_rmax) * 4)));
        HIPCHK(hipFuncSetAttribute((const void *)k_blkinv<L>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)((size_t)2 * (L + 1) * NBK * NBK * 4)));
        maxn = 1;
        for (int l = 0; l < nsdp; l++) maxn = std::max<int64_t>(maxn, sizes[l]);
        int rc;
        int64_t nbl_m = (m + NBK - 1) / NBK, nbl_n = (maxn + NBK - 1) / NBK;
        if ((rc = aalloc(LinvB, nbl_m * NBK * NBK))) return rc;
        if ((rc = aalloc(LinvW, nbl_n * NBK * NBK))) return rc;
        if ((rc = aalloc(rdW, maxn))) return rc;
        if ((rc = aalloc(icX, Nd)) || (rc = aalloc(icZ, Nd))) return rc;
        if ((rc = aalloc(Wc, maxn * maxn)) || (rc = aalloc(Wt, maxn * maxn))) return rc;
        Arr *lv[] = {&lq, &lqold, &lr, &lw, &ltmp};
        for (Arr *v : lv)
            if ((rc = aalloc(*v, maxn))) return rc;
        std::vector<int64_t> di;
        for (int l = 0; l < nsdp; l++)
            for (int i = 0; i < sizes[l]; i++) di.push_back(blkoff[l] + i + (int64_t)sizes[l] * i);
        for (int q = 0; q < nlp; q++) di.push_back(blkoff[nsdp] + q);
        ndiag = (int64_t)di.size();
        if ((rc = upload(&d_diag_idx, di))) return rc;
        return 0;
    }

    // Rpotrf("Lower") of the n x n matrix A[off + i + ld j], blocked right-looking: per block of NBK columns
    // one k_potrf_col launch (the diagonal block and the panel below it; the block's factor to linv at block
    // index k / NBK); in the Ozaki form the trailing updates go in pairs of block columns: after the first of a
    // pair only the next block column is updated (a GEMM of NBK columns), after the second the trailing matrix
    // with both panels at once (a lower-triangular GEMM with K = 2 NBK: half as many passes over the trailing
    // matrix, whose per-output work does not grow with K in that form; the plain form's does).  Then k_blkinv
    // moves the blocks' factors into A and leaves their inverses in linv.  Pivot failures set d_info (checked
    // by the caller).
    int chol(PV A, int64_t off, int n, int ld, PV rd, int64_t roff, Arr &linv) {
        int rc;
        int pend = 0;  // columns of the previous block column not yet applied to the trailing matrix
        for (int k = 0; k < n; k += NBK) {
            const int nb = std::min(NBK, n - k), r = n - k - nb;
            // panel rows per workgroup: 4 while at most 256 workgroups hold the panel (the diagonal block's
            // elimination is repeated in every workgroup), more for larger panels, within the LDS and with
            // one item per thread
            int R = r > 0 ? std::min(col_rmax, col_r_fixed > 0 ? col_r_fixed : std::max(4, (int)cdiv(r, 256))) : 0;
            while (R > 1 && !potrf_col_threads(nb, R)) R--;
            R = std::min(R, r);
            const int G = r > 0 ? (int)cdiv(r, R) : 1, nt = potrf_col_threads(nb, R);
            if (!nt) {
                set_err("k_potrf_col: a block column of " + std::to_string(nb) + " columns does not fit a workgroup");
                return SDPG_EARG;
            }
            const size_t sh = potrf_col_words<L>(nb, R) * 4;
            hipEvent_t ev = kprof.begin(st);
            k_potrf_col<L><<<G, nt, sh, st>>>(A, off, ld, k, nb, r, R, rd, roff, linv.pv(), NBK, d_info);
            kprof.end(st, ev, 2, (double)nb * nb * nb / 6.0 + (double)r * nb * nb / 2.0);
            if (r == 0) break;
            const int64_t c22 = off + (k + nb) + (int64_t)ld * (k + nb);
            if (pend == 0 && r > NBK && chol_pairs && gemm_mode == SDPG_GEMM_OZAKI) {
                // the next block column: A[k+nb.., k+nb .. k+nb+NBK) -= P P[0 .. NBK)^T (its diagonal block's
                // lower triangle)
                const int64_t a21 = off + (k + nb) + (int64_t)ld * k;
                if ((rc = gemm.gemm(r, NBK, nb, A, a21, ld, false, A, a21, ld, true, A, c22, ld, GEMM_SUB, 1, 0, 0,
                                    0, gemm_mode, gemm_slices, true)))
                    return rc;
                pend = nb;
            } else {
                // trailing update A22 -= P P^T on the lower triangle (Rsyrk / Rgemm N,T), P = the panels of
                // this block column and of a pending one
                const int64_t a21 = off + (k + nb) + (int64_t)ld * (k - pend);
                if ((rc = gemm.gemm(r, r, pend + nb, A, a21, ld, false, A, a21, ld, true, A, c22, ld, GEMM_SUB, 1, 0,
                                    0, 0, gemm_mode, gemm_slices, true)))
                    return rc;
                pend = 0;
            }
        }
        // the diagonal blocks into A and their inverses (triangular solves, inverse factors), one launch
        const int nblk = (int)cdiv(n, NBK);
        k_blkinv<L><<<nblk, 256, (size_t)2 * (L + 1) * NBK * NBK * 4, st>>>(A, off, ld, n, NBK, rd, roff, linv.pv(),
                                                                             d_info, 0);
        HIPCHK(hipGetLastError());
        return 0;
    }
    int reset_info() {
        k_fill_i32<<<1, 1, 0, st>>>(d_info, 1, 0);
        HIPCHK(hipGetLastError());
        return 0;
    }
    int read_info(int *info) {
        int hinfo = 0;
        HIPCHK(hipMemcpyAsync(&hinfo, d_info, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        *info = hinfo;
        return 0;
    }

    int potrf(int *info) override {
        tstart();
        int rc;
        if ((rc = reset_info())) return rc;
        if ((rc = chol(Bm.pv(), 0, m, m, rdiag.pv(), 0, LinvB))) return rc;
        if ((rc = read_info(info))) return rc;
        return tstop(1);
    }

    // Lal::getInvLowTriangularMatrix: out[oo + i + ldo j] = inv(L) for the Cholesky factor in
    // Lm[lo + i + ld j] whose diagonal-block inverses are in linv (block index s / NBK).  Recursive
    // on halves: inv([A 0; C B]) = [inv(A) 0; -inv(B) C inv(A)  inv(B)].  The strictly upper part of
    // out must be zero on entry.
    int trtri(PV Lm, int64_t lo, int ld, int s, int len, Arr &linv, PV out, int64_t oo, int ldo) {
        int rc;
        if (len <= NBK) {
            k_copy2d<L><<<cdiv((int64_t)len * len, 256), 256, 0, st>>>(out, oo + s + (int64_t)ldo * s, ldo, linv.pv(),
                                                                     (int64_t)(s / NBK) * NBK * NBK, len, len, len);
            return 0;
        }
        int nblk = (len + NBK - 1) / NBK;
        int n1 = ((nblk + 1) / 2) * NBK, n2 = len - n1;
        if ((rc = trtri(Lm, lo, ld, s, n1, linv, out, oo, ldo))) return rc;
        if ((rc = trtri(Lm, lo, ld, s + n1, n2, linv, out, oo, ldo))) return rc;
        // Wt = C inv(A)   (n2 x n1)
        if ((rc = gemm.gemm(n2, n1, n1, Lm, lo + (s + n1) + (int64_t)ld * s, ld, false, out, oo + s + (int64_t)ldo * s,
                            ldo, false, Wt.pv(), 0, n2, GEMM_SET, 1, 0, 0, 0, gemm_mode, gemm_slices)))
            return rc;
        // out21 = 0 - inv(B) Wt
        return gemm.gemm(n2, n1, n2, out, oo + (s + n1) + (int64_t)ldo * (s + n1), ldo, false, Wt.pv(), 0, n2, false,
                         out, oo + (s + n1) + (int64_t)ldo * s, ldo, GEMM_SUB, 1, 0, 0, 0, gemm_mode, gemm_slices);
    }

    // Solutions::computeInverse (sdpa_dataset.cpp:96-118): invCholX = inv(chol(X)),
    // invCholZ = inv(chol(Z)), invZ = invCholZ^T invCholZ (Jal::getInvChol / getInvCholAndInv,
    // sdpa_jordan.cpp:123-171); *info = 1 when X or Z is not positive definite.
    int factor_point(int *info) override {
        tstart();
        int rc;
        if ((rc = reset_info())) return rc;
        for (int l = 0; l < nsdp; l++) {
            int n = sizes[l];
            int64_t o = blkoff[l];
            for (int which = 0; which < 2; which++) {
                Arr &src = which ? Z : X;
                Arr &ic = which ? icZ : icX;
                k_copy2d<L><<<cdiv((int64_t)n * n, 256), 256, 0, st>>>(Wc.pv(), 0, n, src.pv(), o, n, n, n);
                if ((rc = chol(Wc.pv(), 0, n, n, rdW.pv(), 0, LinvW))) return rc;
                zero_arr(ic, o, (int64_t)n * n);
                if ((rc = trtri(Wc.pv(), 0, n, 0, n, LinvW, ic.pv(), o, n))) return rc;
            }
            if ((rc = gemm.gemm(n, n, n, icZ.pv(), o, n, true, icZ.pv(), o, n, false, invZ.pv(), o, n, GEMM_SET, 1, 0,
                                0, 0, gemm_mode, gemm_slices)))
                return rc;
        }
        if (nlp)
            k_lp_inverse<L><<<cdiv(nlp, 128), 128, 0, st>>>(X.pv(), Z.pv(), icX.pv(), icZ.pv(), invZ.pv(),
                                                            blkoff[nsdp], nlp, d_info);
        HIPCHK(hipGetLastError());
        int hinfo = 0;
        if ((rc = read_info(&hinfo))) return rc;
        *info = hinfo ? 1 : 0;
        return tstop(6);
    }

    // Solutions::update (sdpa_dataset.cpp:120-133): X += alphaP dX, y += alphaD dy, Z += alphaD dZ
    int update(const uint32_t *alphaP, const uint32_t *alphaD) override {
        mp<L> aP = unpack1(alphaP), aD = unpack1(alphaD);
        k_axpy<L><<<cdiv(Nd, 256), 256, 0, st>>>(X.pv(), X.pv(), dX.pv(), aP, Nd);
        k_axpy<L><<<cdiv(m, 256), 256, 0, st>>>(y.pv(), y.pv(), dy.pv(), aD, m);
        k_axpy<L><<<cdiv(Nd, 256), 256, 0, st>>>(Z.pv(), Z.pv(), dZ.pv(), aD, Nd);
        HIPCHK(hipGetLastError());
        return 0;
    }

    // sum of x_i y_i (or of x[idx[i]]) with a two-level reduction; the partial sums are added on the host
    int dot(const Arr &xa, int64_t xo, const Arr *ya, int64_t yo, const int64_t *idx, int64_t n, mp<L> &out) {
        int nb = (int)std::min<int64_t>(256, std::max<int64_t>(1, cdiv(n, 1024)));
        k_dot_part<L><<<nb, 256, 0, st>>>(xa.pv(), xo, ya ? ya->pv() : PV{nullptr, 0}, yo, idx, n, scal.pv());
        HIPCHK(hipGetLastError());
        std::vector<uint32_t> h((size_t)nb * (L + 2));
        int rc = get_packed(scal, 0, h.data(), nb);
        if (rc) return rc;
        out = zero<L>();
        for (int i = 0; i < nb; i++) out = add<L>(out, unpack1(&h[(size_t)i * (L + 2)]));
        return 0;
    }

    // scalar inner products of the step-length / SolveInfo logic (sdpa_parts.cpp:362-860)
    int inner(int what, uint32_t *out) override {
        mp<L> v = zero<L>();
        int rc = 0;
        switch (what) {
        case SDPG_IP_XZ: rc = dot(X, 0, &Z, 0, nullptr, Nd, v); break;
        case SDPG_IP_XDZ: rc = dot(X, 0, &dZ, 0, nullptr, Nd, v); break;
        case SDPG_IP_DXZ: rc = dot(dX, 0, &Z, 0, nullptr, Nd, v); break;
        case SDPG_IP_DXDZ: rc = dot(dX, 0, &dZ, 0, nullptr, Nd, v); break;
        case SDPG_IP_BY: rc = dot(b, 0, &y, 0, nullptr, m, v); break;
        case SDPG_IP_BDY: rc = dot(b, 0, &dy, 0, nullptr, m, v); break;
        case SDPG_IP_TRX: rc = dot(X, 0, nullptr, 0, d_diag_idx, ndiag, v); break;
        case SDPG_IP_TRZ: rc = dot(Z, 0, nullptr, 0, d_diag_idx, ndiag, v); break;
        case SDPG_IP_X0Z:
        case SDPG_IP_XZ0:
            if (!X0.p || !Z0.p) {
                set_err("sdpg_inner: no initial point (sdpg_put SDPG_X0 / SDPG_Z0)");
                return SDPG_ESTATE;
            }
            rc = what == SDPG_IP_X0Z ? dot(X0, 0, &Z, 0, nullptr, Nd, v) : dot(X, 0, &Z0, 0, nullptr, Nd, v);
            break;
        case SDPG_IP_CX:
        case SDPG_IP_CDX: {
            const Arr &M_ = what == SDPG_IP_CX ? X : dX;
            k_con_inner<L><<<1, 256, 0, st>>>(1, nsdp, d_cdesc, d_ri, d_ci, sval.pv(), dval.pv(), d_clp_ptr, d_lp_pos,
                                              lpval.pv(), blkoff[nsdp], M_.pv(), b.pv(), PV{scal.p, scal.n}, 2, 1);
            std::vector<uint32_t> h(L + 2);
            if ((rc = get_packed(scal, 0, h.data(), 1))) return rc;
            v = unpack1(h.data());
            break;
        }
        default: set_err("sdpg_inner: unknown selector"); return SDPG_EARG;
        }
        if (rc) return rc;
        pack1(v, out);
        return 0;
    }

    // Lal::getMinEigen (sdpa_linear.cpp:27-124) for one block: Lanczos on lMat xMat lMat^T with the
    // matrix-vector products on the device; the tridiagonal eigenproblem (Rsteqr there) on the host
    int lanczos(int n, int64_t o, const Arr &Lm, const Arr &Xm, mp<L> &res) {
        using hla::operator+;
        using hla::operator-;
        using hla::operator*;
        using hla::operator/;
        using hla::operator<;
        using hla::operator>;
        using hla::operator>=;
        using MM = hla::M<L>;
        int rc;
        // U = lMat^T so that lMat x is a column-contiguous product
        k_transpose<L><<<cdiv((int64_t)n * n, 256), 256, 0, st>>>(Lm.pv(), o, Wt.pv(), 0, n);
        mp<L> mn = MM::d(1.0e+51), mn_old = MM::d(1.0e+52), mn_min = MM::d(1.0e+50);
        mp<L> error = MM::d(1.0e+10), one = MM::i(1);
        std::vector<mp<L>> diagVec(n, mn_min), diagVec2(n, zero<L>()), out(n), bb(n);
        // the loop ends by k = sqrt(n) + 10, so the tridiagonal eigenvectors need that many rows
        const int ldq = std::min(n, (int)std::sqrt((double)n) + 12);
        std::vector<mp<L>> Q((size_t)ldq * ldq);
        mp<L> nD = MM::i(n);
        mp<L> beta = hla::sqrt(nD);
        const mp<L> sqrtn10 = hla::sqrt(nD) + MM::i(10), sqrtn9 = hla::sqrt(MM::i(n + 9));
        const mp<L> c16 = MM::d(1.0e-16), c5 = MM::d(1.0e-5), c8 = MM::d(1.0e-8), c2 = MM::d(1.0e-2),
                    c4 = MM::d(1.0e-4);
        zero_arr(lq, 0, n);
        {
            std::vector<uint32_t> ones((size_t)n * (L + 2));
            for (int i = 0; i < n; i++) pack1(one, &ones[(size_t)i * (L + 2)]);
            if ((rc = put_packed(lr, 0, ones.data(), n))) return rc;
        }
        int k = 0, kk = 0;
        const mp<L> zr = zero<L>();
        for (;;) {
            mp<L> kf = MM::i(k);
            if (!(k < n) || !(kf < sqrtn10) || !(beta > c16)) break;
            bool condA = hla::abs(mn - mn_old) > c5 * hla::abs(mn) + c8;
            bool condB = hla::abs(error * beta) > c2 * hla::abs(mn) + c4;
            if (!(condA || condB)) break;
            HIPCHK(hipMemcpyAsync(lqold.p, lq.p, sizeof(uint32_t) * (L + 1) * lq.n, hipMemcpyDeviceToDevice, st));
            mp<L> value = one / beta;
            k_lincomb<L><<<cdiv(n, 256), 256, 0, st>>>(lq.pv(), lr.pv(), lr.pv(), value, zr, 0, n);
            // w = lMat^T q ; tmp = xMat w ; w = lMat tmp
            k_gemv_t<L><<<cdiv(n, 4), 256, 0, st>>>(Lm.pv(), o, n, 1, lq.pv(), 0, lw.pv(), 0);
            k_gemv_t<L><<<cdiv(n, 4), 256, 0, st>>>(Xm.pv(), o, n, 0, lw.pv(), 0, ltmp.pv(), 0);
            k_gemv_t<L><<<cdiv(n, 4), 256, 0, st>>>(Wt.pv(), 0, n, 2, ltmp.pv(), 0, lw.pv(), 0);
            mp<L> alpha;
            if ((rc = dot(lq, 0, &lw, 0, nullptr, n, alpha))) return rc;
            diagVec[k] = alpha;
            // r = w - alpha q - beta qold
            k_lincomb<L><<<cdiv(n, 256), 256, 0, st>>>(lr.pv(), lw.pv(), lq.pv(), one, mpx::neg<L>(alpha), 1, n);
            k_axpy<L><<<cdiv(n, 256), 256, 0, st>>>(lr.pv(), lr.pv(), lqold.pv(), mpx::neg<L>(beta), n);
            if (MM::i(kk) >= hla::sqrt(MM::i(k)) || k == n - 1 || kf > sqrtn9) {
                kk = 0;
                out = diagVec;
                bb = diagVec2;
                out[n - 1] = diagVec[k];
                bb[n - 1] = zr;
                int info = hla::tridiag_ql<L>(k + 1, out.data(), bb.data(), Q.data(), ldq);
                if (info > 0) break;
                mn_old = mn;
                mn = out[0];
                error = Q[k];
            }
            mp<L> rr;
            if ((rc = dot(lr, 0, &lr, 0, nullptr, n, rr))) return rc;
            beta = hla::sqrt(rr);
            diagVec2[k] = beta;
            ++k;
            ++kk;
        }
        res = mn - hla::abs(error * beta);
        return 0;
    }

    // Jal::getMinEigen(lMat, xMat) (sdpa_jordan.cpp:55-88): which 0 = (invCholX, dX), 1 = (invCholZ, dZ)
    int min_eigen(int which, uint32_t *out) override {
        tstart();
        using hla::operator<;
        const Arr &Lm = which ? icZ : icX;
        const Arr &Xm = which ? dZ : dX;
        mp<L> mn = mpx::from_double<L>(1.0E50);
        int rc;
        for (int l = 0; l < nsdp; l++) {
            int n = sizes[l];
            int64_t o = blkoff[l];
            mp<L> v;
            if (n > 20) {
                if ((rc = lanczos(n, o, Lm, Xm, v))) return rc;
            } else {
                // Lal::let(DLS2 = xMat lMat^T), DLS1 = lMat DLS2, Rsyev on DLS1 (host, cyclic Jacobi)
                if ((rc = gemm.gemm(n, n, n, Xm.pv(), o, n, false, Lm.pv(), o, n, true, Wt.pv(), 0, n, GEMM_SET, 1, 0, 0,
                                    0, gemm_mode, gemm_slices)))
                    return rc;
                if ((rc = gemm.gemm(n, n, n, Lm.pv(), o, n, false, Wt.pv(), 0, n, false, Wc.pv(), 0, n, GEMM_SET, 1, 0,
                                    0, 0, gemm_mode, gemm_slices)))
                    return rc;
                std::vector<uint32_t> h((size_t)n * n * (L + 2));
                if ((rc = get_packed(Wc, 0, h.data(), (int64_t)n * n))) return rc;
                std::vector<mp<L>> w((size_t)n * n);
                for (size_t q = 0; q < w.size(); q++) w[q] = unpack1(&h[q * (L + 2)]);
                v = hla::jacobi_min_eig<L>(n, w.data());
            }
            if (v < mn) mn = v;
        }
        if (nlp) {
            k_lp_triple_val<L><<<cdiv(nlp, 128), 128, 0, st>>>(Xm.pv(), Lm.pv(), blkoff[nsdp], nlp, Wt.pv());
            std::vector<uint32_t> h((size_t)nlp * (L + 2));
            if ((rc = get_packed(Wt, 0, h.data(), nlp))) return rc;
            for (int q = 0; q < nlp; q++) {
                mp<L> v = unpack1(&h[(size_t)q * (L + 2)]);
                if (v < mn) mn = v;
            }
        }
        pack1(mn, out);
        return tstop(7);
    }

    int potrs() override {
        tstart();
        HIPCHK(hipMemcpyAsync(dy.p, g.p, sizeof(uint32_t) * (L + 1) * dy.n, hipMemcpyDeviceToDevice, st));
        const int nblk = (m + NBK - 1) / NBK;
        for (int kb = 0; kb < nblk; kb++) {  // L y = g
            int k0 = kb * NBK, nb = std::min(NBK, m - k0);
            k_trsv_diag<L><<<1, 256, 0, st>>>(LinvB.pv(), (int64_t)kb * NBK * NBK, nb, dy.pv(), k0, 0);
            if (k0 + nb < m) k_trsv_update<L><<<cdiv(m - k0 - nb, 32), 256, 0, st>>>(Bm.pv(), m, k0, nb, dy.pv(), 0);
        }
        for (int kb = nblk - 1; kb >= 0; kb--) {  // L^T x = y
            int k0 = kb * NBK, nb = std::min(NBK, m - k0);
            k_trsv_diag<L><<<1, 256, 0, st>>>(LinvB.pv(), (int64_t)kb * NBK * NBK, nb, dy.pv(), k0, 1);
            if (k0 > 0) k_trsv_update<L><<<cdiv(k0, 32), 256, 0, st>>>(Bm.pv(), m, k0, nb, dy.pv(), 1);
        }
        HIPCHK(hipGetLastError());
        return tstop(2);
    }

    int dz(int phase) override {
        tstart();
        int useR = (phase == SDPG_PFEAS || phase == SDPG_NOINFO) ? 1 : 0;
        k_scatter<L><<<cdiv(Nd, 128), 128, 0, st>>>(Nd, d_pos_ptr, d_pos_con, d_pos_ent, scat_vals.pv(), dval.pv(),
                                                     d_pos_blk, d_pos_in_blk, d_dcon_ptr, d_dcon, d_dcon_off, dy.pv(),
                                                     Rd.pv(), useR, Cfull.pv(), Z.pv(), dZ.pv(), 0);
        HIPCHK(hipGetLastError());
        return tstop(4);
    }

    int dx() override {
        tstart();
        int rc;
        if ((rc = triple(W1, X, dZ, invZ, W2))) return rc;
        mp<L> mone = from_int<L>(-1);
        k_axpy<L><<<cdiv(Nd, 256), 256, 0, st>>>(dX.pv(), rzinv.pv(), W1.pv(), mone, Nd);
        for (int l = 0; l < nsdp; l++) {
            int n = sizes[l];
            k_symmetrize<L><<<cdiv((int64_t)n * n, 256), 256, 0, st>>>(dX.pv(), blkoff[l], n);
        }
        HIPCHK(hipGetLastError());
        return tstop(4);
    }

    Arr *sel(int what, int64_t &cnt) {
        switch (what) {
        case SDPG_X: cnt = Nd; return &X;
        case SDPG_Z: cnt = Nd; return &Z;
        case SDPG_INVZ: cnt = Nd; return &invZ;
        case SDPG_Y: cnt = m; return &y;
        case SDPG_RP: cnt = m; return &Rp;
        case SDPG_RD: cnt = Nd; return &Rd;
        case SDPG_GVEC: cnt = m; return &g;
        case SDPG_B: cnt = (int64_t)m * m; return &Bm;
        case SDPG_L: cnt = (int64_t)m * m; return &Bm;
        case SDPG_DY: cnt = m; return &dy;
        case SDPG_DZ: cnt = Nd; return &dZ;
        case SDPG_DX: cnt = Nd; return &dX;
        case SDPG_RZINV: cnt = Nd; return &rzinv;
        case SDPG_INVCHOLX: cnt = Nd; return &icX;
        case SDPG_INVCHOLZ: cnt = Nd; return &icZ;
        case SDPG_X0: cnt = Nd; return &X0;
        case SDPG_Z0: cnt = Nd; return &Z0;
        }
        return nullptr;
    }
    int get(int what, uint32_t *out) override {
        int64_t cnt = 0;
        Arr *a = sel(what, cnt);
        if (!a || !a->p) return SDPG_EARG;
        return get_packed(*a, 0, out, cnt);
    }
    int put(int what, const uint32_t *in) override {
        int64_t cnt = 0;
        Arr *a = sel(what, cnt);
        if (!a) return SDPG_EARG;
        int rc = 0;
        if ((what == SDPG_X0 || what == SDPG_Z0) && !a->p && (rc = aalloc(*a, Nd))) return rc;
        rc = put_packed(*a, 0, in, cnt);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    int sync() override {
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    bool slices_ok(int s) const override { return OzakiWork<L>::slices_ok(s) != 0; }
    // on: 1 start (reset), 0 stop; out (n >= 3 * KProf::NK): count, ms, work per kernel kind
    int profile(int on, double *out, int n) override {
        HIPCHK(hipStreamSynchronize(st));
        kprof.flush();
        if (out)
            for (int k = 0; k < KProf::NK && 3 * k + 2 < n; k++) {
                out[3 * k] = (double)kprof.cnt[k];
                out[3 * k + 1] = kprof.ms[k];
                out[3 * k + 2] = kprof.work[k];
            }
        if (on) kprof.reset();
        kprof.on = on != 0;
        return 0;
    }
};

// Standalone multiple-precision GEMM on host buffers (packed words), one call = one stream:
// C = op(A) op(B), C - op(A) op(B) or C + op(A) op(B), column-major as Rgemm.
// Rgemm on host buffers.  scratch == nullptr: device copies, staging and the GEMM workspace are
// allocated for the call and freed after it, on a stream of its own.  scratch != nullptr: all of
// them are carved from the scratch_bytes of device memory at scratch (the static array of the
// no-allocation entry point), the work runs on the null stream, and the call makes no hipMalloc,
// hipFree, stream or event; sizes that do not fit return SDPG_EARG before any device work.
template <int L>
static int rgemm_L(int device, int transa, int transb, int m, int n, int k, const uint32_t *A, int lda,
                   const uint32_t *B, int ldb, uint32_t *C, int ldc, int op, int mode, int slices, int lower,
                   double *ms, uint8_t *scratch = nullptr, size_t scratch_bytes = 0) {
    const int W = L + 2;
    int64_t na = (int64_t)lda * (transa ? m : k), nb = (int64_t)ldb * (transb ? k : n), nc = (int64_t)ldc * n;
    if (na <= 0) na = 1;
    if (nb <= 0) nb = 1;
    hipStream_t st = nullptr;
    uint32_t *dA = nullptr, *dB = nullptr, *dC = nullptr, *stg = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SDPG_OK;
    int64_t stg_n = std::max(na, std::max(nb, nc)) * W;
    const bool fixed = scratch != nullptr;
    Arena arena{scratch, scratch_bytes, 0};
    auto fail = [&](const char *what) {
        set_err(std::string(fixed ? "sdpg_rgemm_noalloc: " : "sdpg_rgemm: ") + what);
        rc = SDPG_EHIP;
    };
    {
        GemmEngine<L> ge;
        if (hipSetDevice(device) != hipSuccess) {
            set_err("sdpg_rgemm: bad device");
            return SDPG_EARG;
        }
        if (fixed) {
            ge.plain.arena = &arena;
            ge.ozaki.arena = &arena;
            dA = (uint32_t *)arena.take(sizeof(uint32_t) * (L + 1) * na);
            dB = (uint32_t *)arena.take(sizeof(uint32_t) * (L + 1) * nb);
            dC = (uint32_t *)arena.take(sizeof(uint32_t) * (L + 1) * nc);
            stg = (uint32_t *)arena.take(sizeof(uint32_t) * stg_n);
            if (!dA || !dB || !dC || !stg) {
                set_err("sdpg_rgemm_noalloc: operands exceed the " + std::to_string(scratch_bytes >> 20) +
                        " MiB static device scratch");
                return SDPG_EARG;
            }
        } else {
            if (hipStreamCreate(&st) != hipSuccess) {
                fail("stream");
                return rc;
            }
            if (hipMalloc((void **)&dA, sizeof(uint32_t) * (L + 1) * na) != hipSuccess ||
                hipMalloc((void **)&dB, sizeof(uint32_t) * (L + 1) * nb) != hipSuccess ||
                hipMalloc((void **)&dC, sizeof(uint32_t) * (L + 1) * nc) != hipSuccess ||
                hipMalloc((void **)&stg, sizeof(uint32_t) * stg_n) != hipSuccess) {
                set_err("sdpg_rgemm: out of device memory");
                rc = SDPG_ENOMEM;
            }
        }
        PV pa{dA, na}, pb{dB, nb}, pc{dC, nc};
        auto upload = [&](const uint32_t *src, int64_t cnt, PV dst) {
            if (rc) return;
            if (hipMemcpyAsync(stg, src, sizeof(uint32_t) * W * cnt, hipMemcpyHostToDevice, st) != hipSuccess)
                return fail("copy");
            k_unpack<L><<<cdiv(cnt, 256), 256, 0, st>>>(stg, dst, 0, cnt);
            if (hipStreamSynchronize(st) != hipSuccess) fail("unpack");
        };
        if (k > 0) {
            upload(A, (int64_t)lda * (transa ? m : k), pa);
            upload(B, (int64_t)ldb * (transb ? k : n), pb);
        }
        upload(C, nc, pc);
        if (!rc) {
            ge.init(st);
            double t0 = 0;
            if (fixed) {
                t0 = host_ms();
            } else {
                (void)hipEventCreate(&e0);
                (void)hipEventCreate(&e1);
                (void)hipEventRecord(e0, st);
            }
            int g = ge.gemm(m, n, k, pa, 0, lda, transa != 0, pb, 0, ldb, transb != 0, pc, 0, ldc, op, 1, 0, 0, 0,
                            mode, slices, lower != 0);
            if (fixed) {
                if (hipStreamSynchronize(st) != hipSuccess) g = -2;
                if (ms) *ms = host_ms() - t0;
            } else {
                (void)hipEventRecord(e1, st);
            }
            if (g == -1) {
                set_err("sdpg_rgemm: slice count not available at this precision");
                rc = SDPG_EARG;
            } else if (g == -5) {
                set_err("sdpg_rgemm: the Ozaki form is built for up to 384-bit limbs");
                rc = SDPG_EPREC;
            } else if (g == -4 && fixed) {
                set_err("sdpg_rgemm_noalloc: operands and product workspace exceed the " +
                        std::to_string(scratch_bytes >> 20) + " MiB static device scratch");
                rc = SDPG_EARG;
            } else if (g) {
                fail("gemm launch");
            }
        }
        if (!rc) {
            k_pack<L><<<cdiv(nc, 256), 256, 0, st>>>(pc, 0, nc, stg);
            if (hipMemcpyAsync(C, stg, sizeof(uint32_t) * W * nc, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                fail("download");
            float t = 0;
            if (!rc && !fixed && ms && hipEventElapsedTime(&t, e0, e1) == hipSuccess) *ms = t;
        }
    }
    if (fixed) return rc;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dC);
    (void)hipFree(stg);
    (void)hipStreamDestroy(st);
    return rc;
}

// Device time of the Ozaki product with S slices on two representative shapes (operands resident,
// one stream, 3 untimed and 10 timed runs each, averaged): a 512 x 512 x 512 product (long-K kernel)
// and the lower-triangle update C -= P P^T with P 1024 x 32 (short-K kernel, the Cholesky's update).
// host holds packed random values for the operands (at least 1024 * 1024 elements).
template <int L>
static int tune_time_L(int device, int S, const uint32_t *host, double *ms_long, double *ms_short) {
    const int W = L + 2;
    const int64_t n1 = 512, n2 = 1024, k2 = 32;
    const int64_t cnt[4] = {n1 * n1, n1 * n1, n2 * k2, n2 * n2};  // A, C1, P, C2
    uint32_t *d[4] = {nullptr, nullptr, nullptr, nullptr}, *stg = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SDPG_OK;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&st) != hipSuccess) return SDPG_EHIP;
    {
        GemmEngine<L> ge;
        ge.init(st);
        bool ok = hipMalloc((void **)&stg, sizeof(uint32_t) * W * n2 * n2) == hipSuccess &&
                  hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
        for (int q = 0; q < 4 && ok; q++) {
            ok = hipMalloc((void **)&d[q], sizeof(uint32_t) * (L + 1) * cnt[q]) == hipSuccess &&
                 hipMemcpyAsync(stg, host, sizeof(uint32_t) * W * cnt[q], hipMemcpyHostToDevice, st) == hipSuccess;
            if (ok) k_unpack<L><<<cdiv(cnt[q], 256), 256, 0, st>>>(stg, PV{d[q], cnt[q]}, 0, cnt[q]);
            ok = ok && hipStreamSynchronize(st) == hipSuccess;
        }
        if (!ok) rc = SDPG_ENOMEM;
        PV A{d[0], cnt[0]}, C1{d[1], cnt[1]}, P{d[2], cnt[2]}, C2{d[3], cnt[3]};
        for (int shape = 0; shape < 2 && !rc; shape++) {
            float tot = 0;
            for (int r = 0; r < 13 && !rc; r++) {
                if (r == 3) (void)hipEventRecord(e0, st);
                int g = shape == 0 ? ge.gemm((int)n1, (int)n1, (int)n1, A, 0, (int)n1, false, A, 0, (int)n1, false, C1, 0,
                                             (int)n1, GEMM_SET, 1, 0, 0, 0, SDPG_GEMM_OZAKI, S)
                                   : ge.gemm((int)n2, (int)n2, (int)k2, P, 0, (int)n2, false, P, 0, (int)n2, true, C2, 0,
                                             (int)n2, GEMM_SUB, 1, 0, 0, 0, SDPG_GEMM_OZAKI, S, true);
                if (g) rc = g == -1 ? SDPG_EARG : SDPG_EHIP;
            }
            if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&tot, e0, e1) != hipSuccess))
                rc = SDPG_EHIP;
            if (!rc) *(shape == 0 ? ms_long : ms_short) = tot / 10.0;
        }
    }
    for (int q = 0; q < 4; q++) (void)hipFree(d[q]);
    (void)hipFree(stg);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    return rc;
}

}  // namespace sdpg

#include "batch.h"

namespace sdpg {

// per-limb-count entry points, defined in engine_inst.hip
#define SDPG_DECL_L_(LL)                                                                                      \
    EngineBase *new_engine_##LL();                                                                          \
    int init_engine_##LL(EngineBase *e);                                                                    \
    int rgemm_##LL(int device, int transa, int transb, int m, int n, int k, const uint32_t *A, int lda,      \
                   const uint32_t *B, int ldb, uint32_t *C, int ldc, int op, int mode, int slices, int lower, \
                   double *ms, uint8_t *scratch, size_t scratch_bytes);                                     \
    int tune_time_##LL(int device, int S, const uint32_t *host, double *ms_long, double *ms_short);        \
    int batch_create_##LL(EngineBase *const *es, int n, BatchBase **out);
SDPG_FOR_EACH_L(SDPG_DECL_L_)
#undef SDPG_DECL_L_

}  // namespace sdpg
