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

// Sustained rate of the plain form's inner operation, one 32 x 32 -> 64-bit limb multiply-add with its
// carry (v_mad_u64_u32 + v_addc_co_u32): eight independent chains per lane, 8 workgroups of 256 per
// CU.  This is the integer-VALU peak that the plain limb kernels and the pivot chain are quoted
// against (a measurement, not a specification figure).
__global__ __launch_bounds__(256) void k_valu_peak(uint32_t *out, int iters, uint32_t seed) {
    uint32_t a0 = seed ^ threadIdx.x, a1 = a0 * 3u + 1u, a2 = a0 * 5u + 7u, a3 = a0 * 7u + 3u;
    uint32_t b0 = a0 + 11u, b1 = a1 + 13u;
    uint64_t l0 = 0, l1 = 0, l2 = 0, l3 = 0, l4 = 0, l5 = 0, l6 = 0, l7 = 0;
    uint32_t h0 = 0, h1 = 0, h2 = 0, h3 = 0, h4 = 0, h5 = 0, h6 = 0, h7 = 0;
    for (int i = 0; i < iters; i++) {
        mpx::mac4(l0, h0, a0, b0, l1, h1, a1, b0, l2, h2, a2, b0, l3, h3, a3, b0);
        mpx::mac4(l4, h4, a0, b1, l5, h5, a1, b1, l6, h6, a2, b1, l7, h7, a3, b1);
    }
    uint64_t s = l0 ^ l1 ^ l2 ^ l3 ^ l4 ^ l5 ^ l6 ^ l7;
    out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)s ^ (uint32_t)(s >> 32) ^ (h0 + h1 + h2 + h3 + h4 + h5 + h6 + h7);
}

// ------------------------------------------------------------------ C-ABI

struct sdpg_ctx {
    sdpg::EngineBase *eng;
};

struct sdpg_batch {
    sdpg::BatchBase *b;
    int n, device;
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
                    : "sdpg_set_gemm: slice count not available at this precision (0, or a multiple of 4 in [4L-12, 4L])");
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

// ---- slice-count autotuner ----
// The Ozaki form keeps 8S - 2 fraction bits of every row / column, so its product error relative to
// K max|A_i.| max|B_.j| falls by 2^8 per slice; the MFMA work grows as S(S+1)/2.  For a requested
// precision the tuner measures, for every slice count built at that limb count, the effective bits
// -log2(max |C - exact| / (K a_i b_j)) on calibration products (full-width random mantissas, exponents
// spread over 40 binades, K = 32 as in the Cholesky updates and K = 300 as in the long products)
// against the plain form at L + 2 limbs (error below 2^-(32L+54), exact at these scales), times each
// count on a 512 x 512 x 512 product and a 1024 x 1024 x 32 lower-triangle update (operands resident
// on the device, averaged over 10 runs), and keeps the fastest count whose effective bits reach the
// request plus 8 guard bits.  Results are cached per
// (device, limb count, target).
namespace {
struct TuneKey {
    int device, L, target;
};
std::mutex g_tune_mu;
std::vector<std::pair<TuneKey, std::vector<double>>> g_tune_cache;  // rows of (S, bits, ms_long, ms_short)

int rgemm_any(int L, int device, int ta, int tb, int m, int n, int k, const uint32_t *A, int lda, const uint32_t *B,
              int ldb, uint32_t *C, int ldc, int op, int mode, int slices, int lower, double *ms) {
#define RGA_(LL) \
    if (L == LL) return rgemm_##LL(device, ta, tb, m, n, k, A, lda, B, ldb, C, ldc, op, mode, slices, lower, ms, nullptr, 0);
    SDPG_FOR_EACH_L(RGA_)
#undef RGA_
    return SDPG_EPREC;
}

bool limb_count_built(int L) {
#define LB_(LL) \
    if (L == LL) return true;
    SDPG_FOR_EACH_L(LB_)
#undef LB_
    return false;
}

constexpr int TL = 20;  // host comparison width (limbs): holds both operands' mantissas exactly

mpx::mp<TL> widen(const uint32_t *w, int L) {
    mpx::mp<TL> x = mpx::zero<TL>();
    bool nz = false;
    for (int q = 0; q < L; q++) nz |= w[2 + q] != 0;
    if (!nz) return x;
    x.e = (int32_t)w[0];
    x.s = w[1] & 1u;
    for (int q = 0; q < L; q++) x.m[TL - L + q] = w[2 + q];
    return x;
}

double log2abs(const mpx::mp<TL> &x) {  // -inf for zero
    if (mpx::is_zero<TL>(x)) return -1e300;
    return (double)x.e - 1.0 + std::log2((double)x.m[TL - 1] / 2147483648.0);
}

struct Rng {
    uint64_t s;
    uint64_t next() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    }
};

// packed r x c column-major matrix of random full-width values with exponents in [-20, 20]
std::vector<uint32_t> calib_matrix(Rng &g, int r, int c, int L) {
    const int W = L + 2;
    std::vector<uint32_t> v((size_t)r * c * W);
    for (size_t t = 0; t < (size_t)r * c; t++) {
        uint32_t *w = &v[t * W];
        w[0] = (uint32_t)((int)(g.next() % 41) - 20);
        w[1] = (uint32_t)(g.next() & 1);
        for (int q = 0; q < L; q++) w[2 + q] = (uint32_t)g.next();
        w[L + 1] |= 0x80000000u;
    }
    return v;
}

std::vector<uint32_t> repad(const std::vector<uint32_t> &v, int L, int L2) {
    const size_t n = v.size() / (L + 2);
    std::vector<uint32_t> o(n * (L2 + 2), 0);
    for (size_t t = 0; t < n; t++) {
        o[t * (L2 + 2)] = v[t * (L + 2)];
        o[t * (L2 + 2) + 1] = v[t * (L + 2) + 1];
        for (int q = 0; q < L; q++) o[t * (L2 + 2) + 2 + (L2 - L) + q] = v[t * (L + 2) + 2 + q];
    }
    return o;
}

// effective bits of the Ozaki product with S slices on one calibration shape (C = A B^T if tb)
int calib_bits(int L, int device, int S, int m, int n, int k, int tb, uint64_t seed, double *bits) {
    Rng g{seed};
    const int W = L + 2, L2 = L + 2, W2 = L2 + 2;
    std::vector<uint32_t> A = calib_matrix(g, m, k, L), B = tb ? calib_matrix(g, n, k, L) : calib_matrix(g, k, n, L);
    std::vector<uint32_t> C((size_t)m * n * W, 0), R((size_t)m * n * W2, 0);
    for (size_t t = 0; t < (size_t)m * n; t++) C[t * W] = (uint32_t)SDPG_EZERO, R[t * W2] = (uint32_t)SDPG_EZERO;
    const int ldb = tb ? n : k;
    int rc = rgemm_any(L, device, 0, tb, m, n, k, A.data(), m, B.data(), ldb, C.data(), m, 0, SDPG_GEMM_OZAKI, S, 0,
                       nullptr);
    if (rc) return rc;
    std::vector<uint32_t> A2 = repad(A, L, L2), B2 = repad(B, L, L2);
    rc = rgemm_any(L2, device, 0, tb, m, n, k, A2.data(), m, B2.data(), ldb, R.data(), m, 0, SDPG_GEMM_PLAIN, 0, 0,
                   nullptr);
    if (rc) return rc;
    // row maxima of A, column maxima of op(B) (log2)
    std::vector<double> la(m, -1e300), lb(n, -1e300);
    for (int i = 0; i < m; i++)
        for (int q = 0; q < k; q++) la[i] = std::max(la[i], log2abs(widen(&A[((size_t)i + (size_t)m * q) * W], L)));
    for (int j = 0; j < n; j++)
        for (int q = 0; q < k; q++) {
            const size_t t = tb ? ((size_t)j + (size_t)n * q) : ((size_t)q + (size_t)k * j);
            lb[j] = std::max(lb[j], log2abs(widen(&B[t * W], L)));
        }
    double worst = -1e300;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            const size_t t = (size_t)i + (size_t)m * j;
            const mpx::mp<TL> d = mpx::sub<TL>(widen(&C[t * W], L), widen(&R[t * W2], L2));
            worst = std::max(worst, log2abs(d) - (std::log2((double)k) + la[i] + lb[j]));
        }
    *bits = -worst;
    return 0;
}

int time_slices(int L, int device, int S, const std::vector<uint32_t> &host, double *ms_long, double *ms_short) {
#define TT_(LL) \
    if (L == LL) return tune_time_##LL(device, S, host.data(), ms_long, ms_short);
    SDPG_FOR_EACH_L(TT_)
#undef TT_
    return SDPG_EPREC;
}
}  // namespace

int sdpg_tune_slices(int prec_bits, int device, int target_bits, int *slices, double *table, int ntable) {
    if (!slices || prec_bits <= 0 || target_bits < 0) {
        set_err("sdpg_tune_slices: bad arguments");
        return SDPG_EARG;
    }
    const int L = mpx::limbs_for_bits(prec_bits);
    const int target = target_bits ? target_bits : prec_bits;
    if (L > OZ_LMAX || !limb_count_built(L)) {
        set_err("sdpg_tune_slices: no Ozaki form at this precision");
        return SDPG_EPREC;
    }
    if (!limb_count_built(L + 2)) {
        set_err("sdpg_tune_slices: the calibration reference needs the " + std::to_string(L + 2) + "-limb build");
        return SDPG_EPREC;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        set_err("sdpg_tune_slices: bad device ordinal");
        return SDPG_ENODEV;
    }
    std::lock_guard<std::mutex> lock(g_tune_mu);
    const std::vector<double> *rows = nullptr;
    for (auto &e : g_tune_cache)
        if (e.first.device == device && e.first.L == L && e.first.target == target) rows = &e.second;
    if (!rows) {
        std::vector<double> r;
        Rng g{77};
        const std::vector<uint32_t> host = calib_matrix(g, 1024, 1024, L);  // operands of the timed products
        for (int S = std::max(8, 4 * L - 12); S <= 4 * L; S += 4) {
            double b1 = 0, b2 = 0, tl = 0, ts = 0;
            int rc = calib_bits(L, device, S, 64, 64, 32, 1, 1000 + S, &b1);
            if (!rc) rc = calib_bits(L, device, S, 48, 40, 300, 0, 2000 + S, &b2);
            if (!rc) rc = time_slices(L, device, S, host, &tl, &ts);
            if (rc) return rc;
            r.insert(r.end(), {(double)S, std::min(b1, b2), tl, ts});
        }
        g_tune_cache.push_back({TuneKey{device, L, target}, r});
        rows = &g_tune_cache.back().second;
    }
    // the fastest qualifying count; counts within 3 % of one another are ties, settled by the smaller
    // count (less MFMA work, rows in increasing S)
    int best = 0;
    double best_t = 0;
    for (size_t q = 0; q + 3 < rows->size(); q += 4) {
        const double S = (*rows)[q], bits = (*rows)[q + 1], t = (*rows)[q + 2] + (*rows)[q + 3];
        if (bits >= target + 8 && (!best || t < 0.97 * best_t)) {
            best = (int)S;
            best_t = t;
        }
    }
    if (table)
        for (int q = 0; q < ntable && q < (int)rows->size(); q++) table[q] = (*rows)[q];
    if (!best) {
        set_err("sdpg_tune_slices: no slice count built at this limb count reaches the target precision");
        return SDPG_EPREC;
    }
    *slices = best;
    return SDPG_OK;
}

int sdpg_valu_peak(int device, double *limb_mac_per_s) {
    if (!limb_mac_per_s) return SDPG_EARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        set_err("sdpg_valu_peak: bad device ordinal");
        return SDPG_ENODEV;
    }
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) return SDPG_EHIP;
    const int blocks = prop.multiProcessorCount * 8, threads = 256, iters = 20000;
    uint32_t *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SDPG_OK;
    float ms = 0;
    if (hipMalloc((void **)&out, sizeof(uint32_t) * blocks * threads) != hipSuccess || hipStreamCreate(&st) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = SDPG_EHIP;
    } else {
        k_valu_peak<<<blocks, threads, 0, st>>>(out, 100, 1u);
        (void)hipEventRecord(e0, st);
        k_valu_peak<<<blocks, threads, 0, st>>>(out, iters, 2u);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0)
            rc = SDPG_EHIP;
        else
            *limb_mac_per_s = 8.0 * iters * (double)blocks * threads / (ms * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(out);
    if (rc) set_err("sdpg_valu_peak: device error");
    return rc;
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

int sdpg_batch_create(sdpg_batch **batch, int n, sdpg_ctx *const *ctxs) {
    if (!batch || n <= 0 || !ctxs) {
        set_err("sdpg_batch_create: bad arguments");
        return SDPG_EARG;
    }
    *batch = nullptr;
    std::vector<EngineBase *> es(n);
    for (int p = 0; p < n; p++) {
        if (!ctxs[p] || !ctxs[p]->eng) {
            set_err("sdpg_batch_create: null context");
            return SDPG_EARG;
        }
        es[p] = ctxs[p]->eng;
        if (es[p]->nlimb != es[0]->nlimb || es[p]->device != es[0]->device) {
            set_err("sdpg_batch_create: contexts must share one precision and one device");
            return SDPG_EARG;
        }
    }
    if (hipSetDevice(es[0]->device) != hipSuccess) {
        set_err("sdpg_batch_create: bad device");
        return SDPG_EHIP;
    }
    const int L = es[0]->nlimb;
    BatchBase *b = nullptr;
    int rc = SDPG_EPREC;
#define BC_(LL) \
    if (L == LL) rc = batch_create_##LL(es.data(), n, &b);
    SDPG_FOR_EACH_L(BC_)
#undef BC_
    if (rc) return rc;
    *batch = new sdpg_batch{b, n, es[0]->device};
    return SDPG_OK;
}

void sdpg_batch_destroy(sdpg_batch *batch) {
    if (!batch) return;
    (void)hipSetDevice(batch->device);
    delete batch->b;
    delete batch;
}

int sdpg_batch_newton(sdpg_batch *batch, const int *req, int *info) {
    if (!batch || !req || !info) {
        set_err("sdpg_batch_newton: bad arguments");
        return SDPG_EARG;
    }
    for (int p = 0; p < batch->n; p++)
        if (req[p] < SDPG_BATCH_NONE || req[p] > SDPG_BATCH_CORRECTOR) {
            set_err("sdpg_batch_newton: bad request");
            return SDPG_EARG;
        }
    return batch->b->newton(req, info);
}

int sdpg_batch_set_graph(sdpg_batch *batch, int on) {
    if (!batch) return SDPG_EARG;
    return batch->b->set_graph(on);
}

int sdpg_batch_time(sdpg_batch *batch, double *ms) {
    if (!batch || !ms) return SDPG_EARG;
    return batch->b->timing(ms);
}

int sdpg_batch_wait_contexts(sdpg_batch *batch, int on) {
    if (!batch) return SDPG_EARG;
    return batch->b->wait_contexts(on);
}

int sdpg_batch_timers(sdpg_batch *batch, double *ms, int n) {
    if (!batch || !ms || n < 0) return SDPG_EARG;
    return batch->b->timers(ms, n);
}

static int batch_args(sdpg_batch *batch, const int *req, const char *what) {
    if (!batch || !req) {
        set_err(std::string(what) + ": bad arguments");
        return SDPG_EARG;
    }
    return SDPG_OK;
}

int sdpg_batch_gvec(sdpg_batch *batch, const int *req, const uint32_t *mu, const uint32_t *beta, const int *phase) {
    int rc = batch_args(batch, req, "sdpg_batch_gvec");
    if (rc) return rc;
    if (!mu || !beta || !phase) return SDPG_EARG;
    for (int p = 0; p < batch->n; p++)
        if (req[p] < SDPG_BATCH_NONE || req[p] > SDPG_BATCH_CORRECTOR) {
            set_err("sdpg_batch_gvec: bad request");
            return SDPG_EARG;
        }
    return batch->b->gvec(req, mu, beta, phase);
}

int sdpg_batch_dz(sdpg_batch *batch, const int *req, const int *phase) {
    int rc = batch_args(batch, req, "sdpg_batch_dz");
    if (rc) return rc;
    if (!phase) return SDPG_EARG;
    return batch->b->dz(req, phase);
}

int sdpg_batch_dx(sdpg_batch *batch, const int *req) {
    int rc = batch_args(batch, req, "sdpg_batch_dx");
    if (rc) return rc;
    return batch->b->dx(req);
}

int sdpg_batch_update(sdpg_batch *batch, const int *req, const uint32_t *alphaP, const uint32_t *alphaD) {
    int rc = batch_args(batch, req, "sdpg_batch_update");
    if (rc) return rc;
    if (!alphaP || !alphaD) return SDPG_EARG;
    return batch->b->update(req, alphaP, alphaD);
}

int sdpg_batch_factor_point(sdpg_batch *batch, const int *req, int *info) {
    int rc = batch_args(batch, req, "sdpg_batch_factor_point");
    if (rc) return rc;
    if (!info) return SDPG_EARG;
    return batch->b->factor_point(req, info);
}

int sdpg_batch_inner(sdpg_batch *batch, const int *req, uint32_t *out) {
    int rc = batch_args(batch, req, "sdpg_batch_inner");
    if (rc) return rc;
    if (!out) return SDPG_EARG;
    return batch->b->inner(req, out);
}

int sdpg_batch_residuals(sdpg_batch *batch, const int *req, uint32_t *out) {
    int rc = batch_args(batch, req, "sdpg_batch_residuals");
    if (rc) return rc;
    if (!out) return SDPG_EARG;
    return batch->b->residuals(req, out);
}

int sdpg_batch_min_eigen(sdpg_batch *batch, const int *req, int which, uint32_t *out) {
    int rc = batch_args(batch, req, "sdpg_batch_min_eigen");
    if (rc) return rc;
    if (!out || which < 0 || which > 1) return SDPG_EARG;
    return batch->b->min_eigen(req, which, out);
}

}  // extern "C"
