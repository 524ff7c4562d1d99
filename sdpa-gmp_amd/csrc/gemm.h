// gemm.h — multiple-precision GEMM engine (templated on limb count L).
//
// C = op(A) * op(B) (SET), C -= op(A) op(B) (SUB) or C += (ADD), batched with element
// strides, optionally restricted to the lower triangle of C.  This is the GEMM
// the reference reaches through Rgemm_NN_omp / Rgemm_NT_omp / Rsyrk
// (mplapack/Rgemm_NN_omp.cpp:36-74, mplapack/Rgemm_NT_omp.cpp:36-76,
// mplapack/Rsyrk.cpp:157-177); it serves the F1 precompute, the Cholesky
// trailing update and the dX / r_zinv triple products.
//
// Forms:
//  - element form (reference): one output per thread, floating limb multiply-add
//    per term;
//  - plain limb form (gemm_plain.h): rows of op(A) and columns of op(B) are
//    brought to block fixed point (one exponent per row / column), products are
//    exact multi-limb integer MADs (v_mad_u64_u32) into 96-bit column
//    accumulators, renormalised once per output;
//  - Ozaki form (gemm_ozaki.h): the same block fixed-point operands are cut into
//    signed 8-bit slices and every slice-pair product runs as an INT8 MFMA with
//    exact INT32 accumulation; slice products are recombined into limbs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "engine_common.h"
#include "mpx.h"

namespace sdpg {

enum { GEMM_SET = 0, GEMM_SUB = 1, GEMM_ADD = 2 };
enum { GEMM_KEEP_A = 1, GEMM_KEEP_B = 2 };

struct GemmArgs {
    int M, N, K;
    PV A;
    int64_t offA;
    int lda;
    int transA;
    PV B;
    int64_t offB;
    int ldb;
    int transB;
    PV C;
    int64_t offC;
    int ldc;
    int op;
    int batch;
    int64_t sA, sB, sC;
    int lower;  // only C[i,j] with i >= j
};

__device__ __forceinline__ int64_t idxA(const GemmArgs &g, int i, int k, int b) {
    return g.offA + (int64_t)b * g.sA + (g.transA ? (k + (int64_t)g.lda * i) : (i + (int64_t)g.lda * k));
}
__device__ __forceinline__ int64_t idxB(const GemmArgs &g, int k, int j, int b) {
    return g.offB + (int64_t)b * g.sB + (g.transB ? (j + (int64_t)g.ldb * k) : (k + (int64_t)g.ldb * j));
}
__device__ __forceinline__ int64_t idxC(const GemmArgs &g, int i, int j, int b) {
    return g.offC + (int64_t)b * g.sC + i + (int64_t)g.ldc * j;
}

// element form: sequential sum over k (the Rgemm summation order)
template <int L>
__global__ void k_gemm_elem(GemmArgs g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int b = blockIdx.z;
    if (i >= g.M || j >= g.N) return;
    if (g.lower && i < j) return;
    mpx::mp<L> acc = mpx::zero<L>();
    for (int k = 0; k < g.K; k++)
        acc = mpx::add<L>(acc, mpx::mul<L>(ldm<L>(g.A, idxA(g, i, k, b)), ldm<L>(g.B, idxB(g, k, j, b))));
    int64_t c = idxC(g, i, j, b);
    if (g.op == GEMM_SUB)
        acc = mpx::sub<L>(ldm<L>(g.C, c), acc);
    else if (g.op == GEMM_ADD)
        acc = mpx::add<L>(ldm<L>(g.C, c), acc);
    stm<L>(g.C, c, acc);
}

// Operands whose data never change while a context lives (the constraint matrices A_i) can keep
// their converted form (block fixed point or INT8 slices, with the row exponents) between calls.
// OpKey identifies such an operand and its conversion; OpCache owns the converted copies.
struct OpKey {
    const void *p;
    int64_t off, stride;
    int ld, trans, R, K, nb, form;
    bool operator==(const OpKey &o) const {
        return p == o.p && off == o.off && stride == o.stride && ld == o.ld && trans == o.trans && R == o.R &&
               K == o.K && nb == o.nb && form == o.form;
    }
};
template <class T>
struct OpCache {
    struct Ent {
        OpKey key;
        T *data;
        int32_t *exps;
    };
    std::vector<Ent> ents;
    ~OpCache() {
        for (Ent &e : ents) {
            (void)hipFree(e.data);
            (void)hipFree(e.exps);
        }
    }
    // the entry for key (fresh = true when it was just allocated and still has to be filled)
    Ent *get(const OpKey &key, size_t n_data, size_t n_exps, bool &fresh) {
        for (Ent &e : ents)
            if (e.key == key) {
                fresh = false;
                return &e;
            }
        Ent e{key, nullptr, nullptr};
        if (hipMalloc((void **)&e.data, sizeof(T) * (n_data ? n_data : 1)) != hipSuccess) return nullptr;
        if (hipMalloc((void **)&e.exps, sizeof(int32_t) * (n_exps ? n_exps : 1)) != hipSuccess) {
            (void)hipFree(e.data);
            return nullptr;
        }
        ents.push_back(e);
        fresh = true;
        return &ents.back();
    }
};
inline OpKey op_key(const GemmArgs &g, int which, int form) {
    return which == 0 ? OpKey{g.A.p, g.offA, g.sA, g.lda, g.transA, g.M, g.K, g.sA ? g.batch : 1, form}
                      : OpKey{g.B.p, g.offB, g.sB, g.ldb, g.transB, g.N, g.K, g.sB ? g.batch : 1, form};
}

// element (r, k) of an operand seen as R rows over K (which = 0: op(A)[r, k]; 1: op(B)[k, r])
__device__ __forceinline__ int64_t fx_at(const GemmArgs &g, int which, int b, int r, int k) {
    return which == 0 ? idxA(g, r, k, b) : idxB(g, k, r, b);
}
// true when consecutive r are consecutive in memory
__device__ __forceinline__ bool fx_rcontig(const GemmArgs &g, int which) {
    return which == 0 ? !g.transA : g.transB;
}
inline bool fx_rcontig_host(const GemmArgs &g, int which) { return which == 0 ? !g.transA : g.transB; }

}  // namespace sdpg

#include "gemm_plain.h"
#include "gemm_ozaki.h"

namespace sdpg {

template <int L>
struct GemmEngine {
    hipStream_t st = nullptr;
    PlainWork<L> plain;
    OzakiWork<L> ozaki;
    int init(hipStream_t s, KProf *prof = nullptr) {
        st = s;
        plain.prof = prof;
        ozaki.prof = prof;
        return 0;
    }
    // mode: 0 plain, 1 ozaki, -1 element (reference) form; keep: GEMM_KEEP_A / GEMM_KEEP_B mark an
    // operand whose data never change, whose converted form is then kept and reused
    int gemm(int M, int N, int K, PV A, int64_t offA, int lda, bool transA, PV B, int64_t offB, int ldb, bool transB,
             PV C, int64_t offC, int ldc, int op, int batch, int64_t sA, int64_t sB, int64_t sC, int mode, int slices,
             bool lower = false, int keep = 0) {
        if (M <= 0 || N <= 0 || batch <= 0) return 0;
        GemmArgs g{M, N, K, A, offA, lda, transA ? 1 : 0, B, offB, ldb, transB ? 1 : 0, C, offC, ldc, op,
                   batch, sA, sB, sC, lower ? 1 : 0};
        if (K <= 0) {
            if (op != GEMM_SET) return 0;
        }
        if (mode == -1 || K <= 0) {
            dim3 blk(16, 16, 1), grd((M + 15) / 16, (N + 15) / 16, batch);
            k_gemm_elem<L><<<grd, blk, 0, st>>>(g);
            return hipGetLastError() == hipSuccess ? 0 : -2;
        }
        if (mode == 1) return ozaki.run(g, st, slices, keep);
        return plain.run(g, st, keep);
    }
};

}  // namespace sdpg
