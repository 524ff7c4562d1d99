// engine.hip — C-ABI of the MI355X Newton engine (include/sdpg.h).  The engine itself
// (kernels, Engine<L>) is in engine_impl.h, instantiated per limb count in engine_inst.hip.
#include "engine_impl.h"

#include <mutex>

namespace sdpg {
static thread_local std::string g_err;
void set_err(const std::string &s) { g_err = s; }

// Device scratch of the no-allocation GEMM entry point (sdpg_rgemm_noalloc): operand copies,
// staging and the product workspace of one call are carved from this array.  Each device that
// loads the library holds its own copy until the process ends.  In host code the array's name is
// not a device address: its per-device address comes from hipGetSymbolAddress, taken once per
// device with that device current, and only that address is handed to kernels and copies.
constexpr size_t SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

constexpr int SCRATCH_MAXDEV = 64;
static std::mutex g_scratch_mu[SCRATCH_MAXDEV];  // one call at a time per device (they share its array)
static uint8_t *g_scratch_ptr[SCRATCH_MAXDEV];

// the device address of this device's scratch array (looked up on the first call on the device);
// the caller holds g_scratch_mu[device]
static int scratch_address(int device, uint8_t **out) {
    if (!g_scratch_ptr[device]) {
        if (hipSetDevice(device) != hipSuccess) {
            set_err("sdpg_rgemm_noalloc: bad device");
            return SDPG_EARG;
        }
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_err(std::string("sdpg_rgemm_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return SDPG_EHIP;
        }
        g_scratch_ptr[device] = (uint8_t *)p;
    }
    *out = g_scratch_ptr[device];
    return SDPG_OK;
}
}  // namespace sdpg

// ------------------------------------------------------------------ C-ABI

struct sdpg_ctx {
    sdpg::EngineBase *eng;
};

using namespace sdpg;

extern "C" {

const char *sdpg_last_error(void) { return g_err.c_str(); }

int sdpg_limbs(int prec_bits) { return mpx::limbs_for_bits(prec_bits); }

static EngineBase *make_engine(int L) {
#define MK_(LL) \
    if (L == LL) return new_engine_##LL();
    SDPG_FOR_EACH_L(MK_)
#undef MK_
    return nullptr;
}

int sdpg_ctx_create(sdpg_ctx **ctx, int prec_bits, int m, int n_sdp, const int *sdp_sizes, int n_lp, int device) {
    if (!ctx || m <= 0 || n_sdp < 0 || n_lp < 0 || (n_sdp > 0 && !sdp_sizes)) {
        set_err("sdpg_ctx_create: bad arguments");
        return SDPG_EARG;
    }
    *ctx = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err("sdpg_ctx_create: no HIP device");
        return SDPG_ENODEV;
    }
    if (device < 0 || device >= ndev) {
        set_err("sdpg_ctx_create: bad device ordinal");
        return SDPG_EARG;
    }
    int L = mpx::limbs_for_bits(prec_bits);
    EngineBase *e = make_engine(L);
    if (!e) {
        set_err("sdpg_ctx_create: unsupported precision (limbs " + std::to_string(L) + ")");
        return SDPG_EPREC;
    }
    e->nlimb = L;
    e->prec = prec_bits;
    e->m = m;
    e->nsdp = n_sdp;
    e->nlp = n_lp;
    e->device = device;
    e->sizes.assign(sdp_sizes, sdp_sizes + n_sdp);
    for (int s : e->sizes)
        if (s <= 0) {
            delete e;
            set_err("sdpg_ctx_create: bad block size");
            return SDPG_EARG;
        }
    int rc = 0;
#define IN_(LL) \
    if (L == LL) rc = init_engine_##LL(e);
    SDPG_FOR_EACH_L(IN_)
#undef IN_
    if (rc) {
        delete e;
        return rc;
    }
    *ctx = new sdpg_ctx{e};
    return SDPG_OK;
}

void sdpg_ctx_destroy(sdpg_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->eng->device);
    delete ctx->eng;
    delete ctx;
}

#define CTX_CHECK(need_commit)                                                   \
    if (!ctx || !ctx->eng) {                                                     \
        set_err("null context");                                                 \
        return SDPG_EARG;                                                        \
    }                                                                            \
    EngineBase *E = ctx->eng;                                                    \
    if ((need_commit) != E->committed) {                                         \
        set_err(need_commit ? "context not committed" : "context already committed"); \
        return SDPG_ESTATE;                                                      \
    }                                                                            \
    hipSetDevice(E->device);

int sdpg_set_block(sdpg_ctx *ctx, int con, int sdp_block, int nnz, const int *row, const int *col, const uint32_t *val,
                   int dense) {
    CTX_CHECK(false);
    if (con < -1 || con >= E->m || sdp_block < 0 || sdp_block >= E->nsdp || nnz < 0 || (nnz && !val)) {
        set_err("sdpg_set_block: bad arguments");
        return SDPG_EARG;
    }
    int n = E->sizes[sdp_block];
    auto &B = E->blk[(size_t)(con + 1) * E->nsdp + sdp_block];
    if (B.set) {
        set_err("sdpg_set_block: block already set");
        return SDPG_ESTATE;
    }
    const int W = E->nlimb + 2;
    if (dense) {
        if (nnz != n * n) {
            set_err("sdpg_set_block: dense block needs n*n values");
            return SDPG_EARG;
        }
        B.dense = true;
        B.v.assign(val, val + (size_t)nnz * W);
    } else {
        if (nnz && (!row || !col)) {
            set_err("sdpg_set_block: missing indices");
            return SDPG_EARG;
        }
        for (int t = 0; t < nnz; t++)
            if (row[t] < 0 || row[t] >= n || col[t] < 0 || col[t] >= n) {
                set_err("sdpg_set_block: index out of range");
                return SDPG_EARG;
            }
        B.dense = false;
        B.r.assign(row, row + nnz);
        B.c.assign(col, col + nnz);
        B.v.assign(val, val + (size_t)nnz * W);
    }
    B.set = true;
    return SDPG_OK;
}

int sdpg_set_lp(sdpg_ctx *ctx, int con, int nnz, const int *pos, const uint32_t *val) {
    CTX_CHECK(false);
    if (con < -1 || con >= E->m || nnz < 0 || (nnz && (!pos || !val))) {
        set_err("sdpg_set_lp: bad arguments");
        return SDPG_EARG;
    }
    for (int t = 0; t < nnz; t++)
        if (pos[t] < 0 || pos[t] >= E->nlp) {
            set_err("sdpg_set_lp: position out of range");
            return SDPG_EARG;
        }
    const int W = E->nlimb + 2;
    E->lp_pos[con + 1].assign(pos, pos + nnz);
    E->lp_val[con + 1].assign(val, val + (size_t)nnz * W);
    return SDPG_OK;
}

int sdpg_set_b(sdpg_ctx *ctx, const uint32_t *b) {
    CTX_CHECK(false);
    if (!b) return SDPG_EARG;
    E->bvec.assign(b, b + (size_t)E->m * (E->nlimb + 2));
    return SDPG_OK;
}

int sdpg_set_formula(sdpg_ctx *ctx, const int *use_formula) {
    CTX_CHECK(false);
    if (!use_formula) return SDPG_EARG;
    E->formula.assign(use_formula, use_formula + (size_t)E->m * E->nsdp);
    return SDPG_OK;
}

int sdpg_commit(sdpg_ctx *ctx) {
    CTX_CHECK(false);
    if (E->bvec.empty()) {
        set_err("sdpg_commit: b not set");
        return SDPG_ESTATE;
    }
    return E->commit();
}

int sdpg_set_gemm(sdpg_ctx *ctx, int mode, int slices) {
    if (!ctx || !ctx->eng) return SDPG_EARG;
    if (mode < -1 || mode > 1 || slices < 0) {
        set_err("sdpg_set_gemm: bad mode");
        return SDPG_EARG;
    }
    if (mode == SDPG_GEMM_OZAKI && !ctx->eng->slices_ok(slices ? slices : 4 * ctx->eng->nlimb - 4)) {
        set_err(ctx->eng->nlimb > OZ_LMAX
                    ? "sdpg_set_gemm: the Ozaki form is built for up to 384-bit limbs"
                    : "sdpg_set_gemm: slice count not available at this precision (use 0, 4L-4 or 4L)");
        return ctx->eng->nlimb > OZ_LMAX ? SDPG_EPREC : SDPG_EARG;
    }
    ctx->eng->gemm_mode = mode;
    ctx->eng->gemm_slices = slices;
    return SDPG_OK;
}

int sdpg_set_point(sdpg_ctx *ctx, const uint32_t *X, const uint32_t *Z, const uint32_t *invZ, const uint32_t *y) {
    CTX_CHECK(true);
    return E->set_point(X, Z, invZ, y);
}
int sdpg_residuals(sdpg_ctx *ctx, uint32_t *normRp, uint32_t *normRd, uint32_t *objP, uint32_t *objD) {
    CTX_CHECK(true);
    return E->residuals(normRp, normRd, objP, objD);
}
int sdpg_gvec(sdpg_ctx *ctx, int corrector, const uint32_t *mu, const uint32_t *beta, int phase) {
    CTX_CHECK(true);
    if (!mu || !beta) return SDPG_EARG;
    return E->gvec(corrector, mu, beta, phase);
}
int sdpg_schur(sdpg_ctx *ctx) {
    CTX_CHECK(true);
    return E->schur();
}
int sdpg_potrf(sdpg_ctx *ctx, int *info) {
    CTX_CHECK(true);
    if (!info) return SDPG_EARG;
    return E->potrf(info);
}
int sdpg_potrs(sdpg_ctx *ctx) {
    CTX_CHECK(true);
    return E->potrs();
}
int sdpg_dz(sdpg_ctx *ctx, int phase) {
    CTX_CHECK(true);
    return E->dz(phase);
}
int sdpg_dx(sdpg_ctx *ctx) {
    CTX_CHECK(true);
    return E->dx();
}
int sdpg_factor_point(sdpg_ctx *ctx, int *info) {
    CTX_CHECK(true);
    if (!info) return SDPG_EARG;
    return E->factor_point(info);
}
int sdpg_update(sdpg_ctx *ctx, const uint32_t *alphaP, const uint32_t *alphaD) {
    CTX_CHECK(true);
    if (!alphaP || !alphaD) return SDPG_EARG;
    return E->update(alphaP, alphaD);
}
int sdpg_inner(sdpg_ctx *ctx, int what, uint32_t *out) {
    CTX_CHECK(true);
    if (!out) return SDPG_EARG;
    return E->inner(what, out);
}
int sdpg_min_eigen(sdpg_ctx *ctx, int which, uint32_t *out) {
    CTX_CHECK(true);
    if (!out || (which != 0 && which != 1)) return SDPG_EARG;
    return E->min_eigen(which, out);
}
int sdpg_get(sdpg_ctx *ctx, int what, uint32_t *out) {
    CTX_CHECK(true);
    if (!out) return SDPG_EARG;
    return E->get(what, out);
}
int sdpg_put(sdpg_ctx *ctx, int what, const uint32_t *in) {
    CTX_CHECK(true);
    if (!in) return SDPG_EARG;
    return E->put(what, in);
}
int sdpg_timers(sdpg_ctx *ctx, double *ms, int n) {
    if (!ctx || !ctx->eng || !ms) return SDPG_EARG;
    for (int i = 0; i < n && i < 8; i++) ms[i] = ctx->eng->timers[i];
    return SDPG_OK;
}
int sdpg_reset_timers(sdpg_ctx *ctx) {
    if (!ctx || !ctx->eng) return SDPG_EARG;
    for (double &t : ctx->eng->timers) t = 0;
    return SDPG_OK;
}
int sdpg_sync(sdpg_ctx *ctx) {
    if (!ctx || !ctx->eng) return SDPG_EARG;
    return ctx->eng->sync();
}

int sdpg_profile(sdpg_ctx *ctx, int on, double *out, int n) {
    if (!ctx || !ctx->eng) return SDPG_EARG;
    hipSetDevice(ctx->eng->device);
    return ctx->eng->profile(on, out, n);
}

int sdpg_rgemm(int prec_bits, int device, int transa, int transb, int m, int n, int k, const uint32_t *A, int lda,
               const uint32_t *B, int ldb, uint32_t *C, int ldc, int op, int mode, int slices, int lower,
               double *ms) {
    if (m < 0 || n < 0 || k < 0 || op < 0 || op > 2 || mode < -1 || mode > 1 || slices < 0 || !C ||
        (k > 0 && (!A || !B)) || ldc < std::max(1, m) || lda < std::max(1, transa ? k : m) ||
        ldb < std::max(1, transb ? n : k)) {
        set_err("sdpg_rgemm: bad arguments");
        return SDPG_EARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err("sdpg_rgemm: no HIP device");
        return SDPG_ENODEV;
    }
    if (device < 0 || device >= ndev) {
        set_err("sdpg_rgemm: bad device ordinal");
        return SDPG_EARG;
    }
    if (m == 0 || n == 0) return SDPG_OK;
    int L = mpx::limbs_for_bits(prec_bits);
#define RG_(LL)                                                                                                    \
    if (L == LL) return rgemm_##LL(device, transa, transb, m, n, k, A, lda, B, ldb, C, ldc, op, mode, slices, lower, \
                                   ms, nullptr, 0);
    SDPG_FOR_EACH_L(RG_)
#undef RG_
    set_err("sdpg_rgemm: unsupported precision");
    return SDPG_EPREC;
}

int sdpg_rgemm_noalloc(int prec_bits, int device, int transa, int transb, int m, int n, int k, const uint32_t *A,
                       int lda, const uint32_t *B, int ldb, uint32_t *C, int ldc, int op, int mode, int slices,
                       int lower, double *ms) {
    if (m < 0 || n < 0 || k < 0 || op < 0 || op > 2 || mode < -1 || mode > 1 || slices < 0 || !C ||
        (k > 0 && (!A || !B)) || ldc < std::max(1, m) || lda < std::max(1, transa ? k : m) ||
        ldb < std::max(1, transb ? n : k)) {
        set_err("sdpg_rgemm_noalloc: bad arguments");
        return SDPG_EARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err("sdpg_rgemm_noalloc: no HIP device");
        return SDPG_ENODEV;
    }
    if (device < 0 || device >= ndev || device >= SCRATCH_MAXDEV) {
        set_err("sdpg_rgemm_noalloc: bad device ordinal");
        return SDPG_EARG;
    }
    if (m == 0 || n == 0) return SDPG_OK;
    int L = mpx::limbs_for_bits(prec_bits);
    std::lock_guard<std::mutex> lock(g_scratch_mu[device]);
    uint8_t *scr = nullptr;
    int rc = scratch_address(device, &scr);
    if (rc) return rc;
#define RGN_(LL)                                                                                                   \
    if (L == LL) return rgemm_##LL(device, transa, transb, m, n, k, A, lda, B, ldb, C, ldc, op, mode, slices, lower, \
                                   ms, scr, SCRATCH_BYTES);
    SDPG_FOR_EACH_L(RGN_)
#undef RGN_
    set_err("sdpg_rgemm_noalloc: unsupported precision");
    return SDPG_EPREC;
}

}  // extern "C"
