// host_solver.cpp — host driver: SDPA-GMP's IPM control on the host, Newton on the GPU.
//
// Follows the reference main loop pinpal() (sdpa_main.cpp:43-461) and the host
// control classes of sdpa_parts.cpp (StepLength, DirectionParameter, Switch,
// AverageComplementarity, RatioInitResCurrentRes, SolveInfo, Phase) and
// sdpa_dataset.cpp (Solutions::initialize / update / computeInverse); the
// Newton system (Newton::Mehrotra, sdpa_newton.cpp:1375-1414) is a sequence of
// C-ABI calls into the device engine (include/sdpg.h).
#include <stdio.h>
#include <stdlib.h>
#include <ctype.h>
#include <string.h>
#include <sys/time.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "host_la.h"
#include "limbs.h"
#include "mpx.h"
#include "mpx_host.h"
#include "sdpa_amd.h"
#include "sdpg.h"

namespace sdpg {
namespace host {

constexpr int BATCH_TIMERS = 9;  // kinds of sdpg_batch_timers

using mpx::mp;
using namespace sdpg::hla;

static double now_s() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

// param.sdpa (sdpa_parts.cpp:136-225)
struct Parameter {
    int maxIteration = 200;
    double epsilonStar = 1.0e-30, lambdaStar = 1.0e+4, omegaStar = 2.0, lowerBound = -1.0e+5, upperBound = 1.0e+5;
    double betaStar = 0.1, betaBar = 0.3, gammaStar = 0.9, epsilonDash = 1.0e-30;
    int precision = 200;
    // print formats of the solution (param.sdpa lines 12-15; defaults of sdpa_parts.cpp:177-180)
    std::string xPrint = "%+50.40Fe", XPrint = "%+50.40Fe", YPrint = "%+50.40Fe", infPrint = "%+50.40Fe";
    // Parameter::setDefaultParameter (sdpa_parts.cpp:136-181): 0 default, 1 unstable but fast, 2 stable
    // but slow (-pt)
    void preset(int type) {
        *this = Parameter();
        if (type == 2) {
            maxIteration = 1000;
            lambdaStar = 1.0e+2;
            betaStar = 0.2;
            betaBar = 0.4;
            gammaStar = 0.5;
            precision = 300;
        } else if (type == 1) {
            maxIteration = 100;
            lambdaStar = 1.0e+2;
            betaStar = 0.01;
            betaBar = 0.02;
            gammaStar = 0.98;
            precision = 100;
        }
    }
    bool read(const char *path) {
        FILE *fp = fopen(path, "r");
        if (!fp) return false;
        bool ok = true;
        ok &= fscanf(fp, "%d%*[^\n]", &maxIteration) == 1;
        ok &= fscanf(fp, "%lf%*[^\n]", &epsilonStar) == 1;
        ok &= fscanf(fp, "%lf%*[^\n]", &lambdaStar) == 1;
        ok &= fscanf(fp, "%lf%*[^\n]", &omegaStar) == 1;
        ok &= fscanf(fp, "%lf%*[^\n]", &lowerBound) == 1;
        ok &= fscanf(fp, "%lf%*[^\n]", &upperBound) == 1;
        ok &= fscanf(fp, "%lf%*[^\n]", &betaStar) == 1;
        ok &= fscanf(fp, "%lf%*[^\n]", &betaBar) == 1;
        ok &= fscanf(fp, "%lf%*[^\n]", &gammaStar) == 1;
        ok &= fscanf(fp, "%lf%*[^\n]", &epsilonDash) == 1;
        ok &= fscanf(fp, "%d%*[^\n]", &precision) == 1;
        std::string *pf[4] = {&xPrint, &XPrint, &YPrint, &infPrint};
        for (std::string *f : pf) {
            char buf[64];
            if (fscanf(fp, "%63s%*[^\n]", buf) != 1) break;  // older files end after the precision
            *f = buf;
        }
        fclose(fp);
        return ok;
    }
};

enum { noINFO, pFEAS, dFEAS, pdFEAS, pdINF, pFEAS_dINF, pINF_dFEAS, pdOPT, pUNBD, dUNBD };

static const char *phase_str(int v) {
    switch (v) {
    case noINFO: return "noINFO    ";
    case pFEAS: return "pFEAS     ";
    case dFEAS: return "dFEAS     ";
    case pdFEAS: return "pdFEAS    ";
    case pdINF: return "pdINF     ";
    case pFEAS_dINF: return "pFEAS_dINF";
    case pINF_dFEAS: return "pINF_dFEAS";
    case pdOPT: return "pdOPT     ";
    case pUNBD: return "pUNBD     ";
    case dUNBD: return "dUNBD     ";
    }
    return "phase error";
}
// Phase::reverse under REVERSE_PRIMAL_DUAL (sdpa_parts.cpp:1013-1048)
static int phase_reverse(int v) {
    switch (v) {
    case pFEAS: return dFEAS;
    case dFEAS: return pFEAS;
    case pFEAS_dINF: return pINF_dFEAS;
    case pINF_dFEAS: return pFEAS_dINF;
    case pUNBD: return dUNBD;
    case dUNBD: return pUNBD;
    }
    return v;
}

// ----- input (sdpa_io.cpp:123-142, 410-751) -----
struct RawEntry {
    int k, l, i, j;
    std::string v;
};

struct RawProblem {
    int m = 0, nBlock = 0;
    std::vector<int> blockStruct;
    std::vector<std::string> b;
    std::vector<RawEntry> ent;
};

static int skip_to_num(FILE *fp) {
    int c;
    while ((c = fgetc(fp)) != EOF) {
        if ((c >= '0' && c <= '9') || c == '+' || c == '-') {
            ungetc(c, fp);
            return 1;
        }
    }
    return 0;
}
static bool read_int(FILE *fp, int *v) { return skip_to_num(fp) && fscanf(fp, "%d", v) == 1; }
static bool read_tok(FILE *fp, std::string &s) {
    if (!skip_to_num(fp)) return false;
    s.clear();
    int c;
    while ((c = fgetc(fp)) != EOF) {
        if ((c >= '0' && c <= '9') || c == '+' || c == '-' || c == '.' || c == 'e' || c == 'E')
            s.push_back((char)c);
        else {
            ungetc(c, fp);
            break;
        }
    }
    return !s.empty();
}

// The reference picks the sparse reader for a data file whose name ends in "-s" when the name is given
// positionally (sdpa_main.cpp:620-623; -ds / -dd choose explicitly); the same rule holds here.
static bool name_is_sparse(const char *path) {
    size_t n = strlen(path);
    return n >= 2 && path[n - 1] == 's' && path[n - 2] == '-';
}

// a decimal token whose mantissa digits are all zero ("0", "-0.000e+3")
static bool token_is_zero(const std::string &t) {
    for (char c : t) {
        if (c == 'e' || c == 'E') break;
        if (c >= '1' && c <= '9') return false;
    }
    return true;
}

static int read_raw(const char *path, RawProblem &R, int dense = -1) {
    if (dense < 0) dense = name_is_sparse(path) ? 0 : 1;
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    char line[8192];
    for (;;) {
        if (!fgets(line, sizeof line, fp)) {
            fclose(fp);
            return -2;
        }
        if (line[0] == '*' || line[0] == '"') continue;
        if (sscanf(line, "%d", &R.m) != 1) {
            fclose(fp);
            return -2;
        }
        break;
    }
    if (fscanf(fp, "%d", &R.nBlock) != 1 || R.nBlock <= 0) {
        fclose(fp);
        return -3;
    }
    R.blockStruct.resize(R.nBlock);
    for (int l = 0; l < R.nBlock; l++)
        if (!read_int(fp, &R.blockStruct[l])) {
            fclose(fp);
            return -3;
        }
    R.b.resize(R.m);
    for (int k = 0; k < R.m; k++)
        if (!read_tok(fp, R.b[k])) {
            fclose(fp);
            return -4;
        }
    if (dense) {
        // dense format (sdpa_io.cpp:696-751): C then A_1 .. A_m, every block in full (an SDP block as n x n
        // values by rows, an LP block as its diagonal); entries with i <= j and a nonzero value are kept
        for (int k = 0; k <= R.m; k++)
            for (int l = 0; l < R.nBlock; l++) {
                const int bs = R.blockStruct[l];
                if (bs >= 2) {
                    for (int i = 0; i < bs; i++)
                        for (int j = 0; j < bs; j++) {
                            std::string t;
                            if (!read_tok(fp, t)) {
                                fclose(fp);
                                return -5;
                            }
                            if (i <= j && !token_is_zero(t)) R.ent.push_back(RawEntry{k, l + 1, i + 1, j + 1, t});
                        }
                } else {
                    const int nl = bs < 0 ? -bs : 1;
                    for (int j = 0; j < nl; j++) {
                        std::string t;
                        if (!read_tok(fp, t)) {
                            fclose(fp);
                            return -5;
                        }
                        if (!token_is_zero(t)) R.ent.push_back(RawEntry{k, l + 1, j + 1, j + 1, t});
                    }
                }
            }
        fclose(fp);
        return 0;
    }
    for (;;) {
        RawEntry e;
        if (!read_int(fp, &e.k) || !read_int(fp, &e.l) || !read_int(fp, &e.i) || !read_int(fp, &e.j) ||
            !read_tok(fp, e.v))
            break;
        R.ent.push_back(std::move(e));
    }
    fclose(fp);
    return 0;
}

// initial-point file (sdpa_io.cpp:150-274): y with the sign flipped, then Z and X (target 1 = Z,
// otherwise X in the sparse form; Z before X in the dense form)
struct RawInit {
    std::vector<std::string> y;
    std::vector<RawEntry> ent;  // k = 1 for Z, 2 for X; l = SDP block number (1-based) or nsdp + LP position
};

static int read_init(const char *path, int sparse, int m, const std::vector<int> &blockStruct, RawInit &I) {
    if (sparse < 0) sparse = name_is_sparse(path) ? 1 : 0;
    FILE *fp = fopen(path, "r");
    if (!fp) return -31;
    I.y.resize(m);
    for (int k = 0; k < m; k++)
        if (!read_tok(fp, I.y[k])) {
            fclose(fp);
            return -32;
        }
    std::vector<int> sdp;
    int nlp = 0;
    for (int b : blockStruct) {
        if (b >= 2) sdp.push_back(b);
        else nlp += b < 0 ? -b : 1;
    }
    const int nsdp = (int)sdp.size();
    if (sparse) {
        for (;;) {
            RawEntry e;
            if (!read_int(fp, &e.k) || !read_int(fp, &e.l) || !read_int(fp, &e.i) || !read_int(fp, &e.j) ||
                !read_tok(fp, e.v))
                break;
            e.k = e.k == 1 ? 1 : 2;
            if (e.l > nsdp) e.l = nsdp + 1;  // the LP part: position i - 1 (sdpa_io.cpp:196-204)
            I.ent.push_back(std::move(e));
        }
    } else {
        for (int tgt = 1; tgt <= 2; tgt++) {
            for (int l = 0; l < nsdp; l++)
                for (int i = 0; i < sdp[l]; i++)
                    for (int j = 0; j < sdp[l]; j++) {
                        std::string t;
                        if (!read_tok(fp, t)) {
                            fclose(fp);
                            return -33;
                        }
                        if (i <= j && !token_is_zero(t)) I.ent.push_back(RawEntry{tgt, l + 1, i + 1, j + 1, t});
                    }
            for (int j = 0; j < nlp; j++) {
                std::string t;
                if (!read_tok(fp, t)) {
                    fclose(fp);
                    return -33;
                }
                if (!token_is_zero(t)) I.ent.push_back(RawEntry{tgt, nsdp + 1, j + 1, j + 1, t});
            }
        }
    }
    fclose(fp);
    return 0;
}

// The device side of one problem's loop.  This one makes the context's own calls; a batched run gives
// each problem a port that meets the other running problems' calls and issues one batched call per kind.
struct Exec {
    virtual ~Exec() {}
    // Newton::Mehrotra (sdpa_newton.cpp:1375-1414) after the control's parameters: Make_gVec, then
    // (predictor) Make_bMat and the Cholesky factorisation, the solve, dZ and dX; *info > 0 when the
    // factorisation fails (and nothing after it runs)
    virtual int step(sdpg_ctx *c, bool predictor, const uint32_t *mu, const uint32_t *beta, int phase, int *info) {
        *info = 0;
        int rc = sdpg_gvec(c, predictor ? 0 : 1, mu, beta, phase);
        if (rc) return rc;
        if (predictor) {
            if ((rc = sdpg_schur(c)) || (rc = sdpg_potrf(c, info))) return rc;
            if (*info > 0) return 0;
        }
        if ((rc = sdpg_potrs(c)) || (rc = sdpg_dz(c, phase))) return rc;
        return sdpg_dx(c);
    }
    virtual int inner(sdpg_ctx *c, int what, uint32_t *out) { return sdpg_inner(c, what, out); }
    // the step-length minimum eigenvalues for (X, dX) and (Z, dZ)
    virtual int min_eigen2(sdpg_ctx *c, uint32_t *x, uint32_t *z) {
        int rc = sdpg_min_eigen(c, 0, x);
        return rc ? rc : sdpg_min_eigen(c, 1, z);
    }
    virtual int update(sdpg_ctx *c, const uint32_t *aP, const uint32_t *aD) { return sdpg_update(c, aP, aD); }
    virtual int factor_point(sdpg_ctx *c, int *info) { return sdpg_factor_point(c, info); }
    virtual int residuals(sdpg_ctx *c, uint32_t *nrp, uint32_t *nrd) {
        return sdpg_residuals(c, nrp, nrd, nullptr, nullptr);
    }
};
static Exec g_direct;

template <int L>
struct Solver {
    using T = mp<L>;
    static constexpr int W = L + 2;
    // problem
    int m = 0, nsdp = 0, nlp = 0, nDim = 0;
    std::vector<int> sdpStruct, blockType, blockNumber, blockStructAbs;
    std::vector<T> b;
    struct SpB {
        int blk = 0, n = 0;
        bool dense = false;
        std::vector<int> r, c;
        std::vector<T> v;
        std::vector<T> de;
        int effect = 0;
    };
    struct SLS {
        std::vector<SpB> s;
        std::vector<int> lpi;
        std::vector<T> lpv;
        SpB *find(int blk) {
            for (auto &x : s)
                if (x.blk == blk) return &x;
            return nullptr;
        }
    };
    SLS C;
    std::vector<SLS> A;
    // dense linear spaces
    std::vector<int64_t> off;
    int64_t Nd = 0;
    std::vector<T> X, Z, invZ, y;  // initial point only; the iterates live on the device
    std::vector<int> useFormula;
    sdpg_ctx *ctx = nullptr;
    int used_slices = 0;  // Ozaki slice count in use (0: another GEMM form)
    const RawInit *init = nullptr;  // initial-point file (-id / -is), if any
    Parameter prm;
    FILE *logf = nullptr;
    std::string err;
    Exec *ex = &g_direct;

    ~Solver() {
        if (ctx) sdpg_ctx_destroy(ctx);
    }

    T parse(const std::string &s) {
        T v;
        if (mpx::parse_decimal<L>(s.c_str(), v) <= 0) v = mpx::zero<L>();
        return v;
    }

    int load(const RawProblem &R) {
        m = R.m;
        blockType.assign(R.nBlock, 0);
        blockNumber.assign(R.nBlock, 0);
        blockStructAbs = R.blockStruct;
        // block typing (sdpa_main.cpp:114-131)
        for (int i = 0; i < R.nBlock; i++) {
            int bs = R.blockStruct[i];
            if (bs >= 2) {
                blockType[i] = 1;
                blockNumber[i] = nsdp++;
                sdpStruct.push_back(bs);
            } else if (bs < 0) {
                blockType[i] = 3;
                blockStructAbs[i] = -bs;
                blockNumber[i] = nlp;
                nlp += -bs;
            } else if (bs == 1) {
                blockType[i] = 3;
                blockNumber[i] = nlp;
                nlp += 1;
            } else {
                err = "block struct";
                return -10;
            }
        }
        nDim = nlp;
        for (int s : sdpStruct) nDim += s;
        b.resize(m);
        for (int k = 0; k < m; k++) b[k] = parse(R.b[k]);
        A.assign(m, SLS());
        for (const RawEntry &e : R.ent) {
            if (e.k < 0 || e.k > m || e.l < 1 || e.l > R.nBlock) {
                err = "entry out of range";
                return -11;
            }
            SLS &t = (e.k == 0) ? C : A[e.k - 1];
            T v = parse(e.v);
            if (e.k == 0) v = mpx::neg<L>(v);  // C = -F0 (sdpa_io.cpp:661)
            if (blockType[e.l - 1] == 1) {
                int l2 = blockNumber[e.l - 1];
                SpB *s = t.find(l2);
                if (!s) {
                    SpB nb;
                    nb.blk = l2;
                    nb.n = sdpStruct[l2];
                    size_t pos = 0;
                    while (pos < t.s.size() && t.s[pos].blk < l2) pos++;
                    t.s.insert(t.s.begin() + pos, nb);
                    s = t.find(l2);
                }
                if (e.i < 1 || e.j < 1 || e.i > s->n || e.j > s->n) {
                    err = "out of range in input data";
                    return -12;
                }
                s->r.push_back(e.i - 1);
                s->c.push_back(e.j - 1);
                s->v.push_back(v);
                s->effect += (e.i == e.j) ? 1 : 2;
            } else {
                if (e.i != e.j) {
                    err = "LP entry off the diagonal";
                    return -13;
                }
                int pos = blockNumber[e.l - 1] + e.i - 1;
                bool found = false;
                for (size_t q = 0; q < t.lpi.size(); q++)
                    if (t.lpi[q] == pos) {
                        t.lpv[q] = v;
                        found = true;
                    }
                if (!found) {
                    size_t q = 0;
                    while (q < t.lpi.size() && t.lpi[q] < pos) q++;
                    t.lpi.insert(t.lpi.begin() + q, pos);
                    t.lpv.insert(t.lpv.begin() + q, v);
                }
            }
        }
        // SparseMatrix::changeToDense (sdpa_struct.cpp:484-525)
        auto densify = [&](SpB &s) {
            int n = s.n;
            if ((double)s.r.size() < (double)(n * n) * 0.20) return;
            s.dense = true;
            s.de.assign((size_t)n * n, mpx::zero<L>());
            for (size_t t = 0; t < s.r.size(); t++) {
                s.de[s.r[t] + (size_t)n * s.c[t]] = s.v[t];
                s.de[s.c[t] + (size_t)n * s.r[t]] = s.v[t];
            }
            s.effect = n * n;
        };
        for (auto &s : C.s) densify(s);
        for (auto &a : A)
            for (auto &s : a.s) densify(s);
        off.assign(nsdp + 1, 0);
        Nd = 0;
        for (int l = 0; l < nsdp; l++) {
            off[l] = Nd;
            Nd += (int64_t)sdpStruct[l] * sdpStruct[l];
        }
        off[nsdp] = Nd;
        Nd += nlp;
        return 0;
    }

    // Newton::computeFormula_SDP (sdpa_newton.cpp:694-819), KAPPA = 1.2 (sdpa_main.cpp:27).
    // The comparisons are evaluated exactly with kappa equal to the double 1.2
    // (= 5404319552844595 / 2^52), the value the reference converts to mpf.
    void compute_formula() {
        useFormula.assign((size_t)m * nsdp, 0);
        const __int128 K = 5404319552844595LL, P52 = ((__int128)1) << 52;
        for (int l = 0; l < nsdp; l++) {
            std::vector<int> cons;
            for (int i = 0; i < m; i++)
                if (A[i].find(l)) cons.push_back(i);
            for (int i : cons) {
                SpB *si = A[i].find(l);
                long long inz = si->effect, up = inz;
                for (int j : cons) {
                    long long jnz = A[j].find(l)->effect;
                    if (jnz < inz || (jnz == inz && j < i)) up += jnz;
                }
                __int128 n = si->n;
                __int128 f1 = K * n * inz + n * n * n * P52 + K * up;
                __int128 f2 = K * n * inz + K * (n + 1) * up;
                __int128 f3 = (2 * K * inz + P52) * up;
                int fm;
                if (si->dense)
                    fm = (f1 < f2) ? 1 : 2;
                else if (f1 < f2 && f1 < f3)
                    fm = 1;
                else if (f2 < f3)
                    fm = 2;
                else
                    fm = 3;
                useFormula[(size_t)i * nsdp + l] = fm;
            }
        }
    }

    static void pack(const std::vector<T> &v, std::vector<uint32_t> &out) {
        out.resize(v.size() * W);
        for (size_t i = 0; i < v.size(); i++) {
            out[i * W] = (uint32_t)v[i].e;
            out[i * W + 1] = v[i].s;
            for (int k = 0; k < L; k++) out[i * W + 2 + k] = v[i].m[k];
        }
    }
    static T unpack1(const uint32_t *w) {
        T x;
        x.e = (int32_t)w[0];
        x.s = w[1] & 1u;
        for (int k = 0; k < L; k++) x.m[k] = w[2 + k];
        if (x.m[L - 1] == 0) x = mpx::zero<L>();
        return x;
    }
    static void unpack(const std::vector<uint32_t> &in, std::vector<T> &v) {
        for (size_t i = 0; i < v.size(); i++) v[i] = unpack1(&in[i * W]);
    }
    static std::vector<uint32_t> pack1(const T &x) {
        std::vector<T> v(1, x);
        std::vector<uint32_t> o;
        pack(v, o);
        return o;
    }

    int setup_device(int device, int gemm_mode, int slices) {
        std::vector<int> sizes(sdpStruct);
        int rc = sdpg_ctx_create(&ctx, prm.precision, m, nsdp, sizes.data(), nlp, device);
        if (rc) return rc;
        auto push = [&](int con, SLS &s) -> int {
            for (auto &blk : s.s) {
                std::vector<uint32_t> pk;
                if (blk.dense) {
                    pack(blk.de, pk);
                    int r = sdpg_set_block(ctx, con, blk.blk, blk.n * blk.n, nullptr, nullptr, pk.data(), 1);
                    if (r) return r;
                } else {
                    pack(blk.v, pk);
                    int r = sdpg_set_block(ctx, con, blk.blk, (int)blk.r.size(), blk.r.data(), blk.c.data(),
                                           pk.data(), 0);
                    if (r) return r;
                }
            }
            if (!s.lpi.empty()) {
                std::vector<uint32_t> pk;
                pack(s.lpv, pk);
                int r = sdpg_set_lp(ctx, con, (int)s.lpi.size(), s.lpi.data(), pk.data());
                if (r) return r;
            }
            return 0;
        };
        if ((rc = push(-1, C))) return rc;
        for (int i = 0; i < m; i++)
            if ((rc = push(i, A[i]))) return rc;
        std::vector<uint32_t> pb;
        pack(b, pb);
        if ((rc = sdpg_set_b(ctx, pb.data()))) return rc;
        if ((rc = sdpg_set_formula(ctx, useFormula.data()))) return rc;
        if ((rc = sdpg_commit(ctx))) return rc;
        // Ozaki form with the slice count left open: the autotuned count for the requested precision
        // (the default 4L - 4 when the tuner cannot run, e.g. without the L + 2 limb build)
        if (gemm_mode == SDPG_GEMM_OZAKI && slices == 0) {
            int tuned = 0;
            if (sdpg_tune_slices(prm.precision, device, prm.precision, &tuned, nullptr, 0) == SDPG_OK) slices = tuned;
        }
        used_slices = gemm_mode == SDPG_GEMM_OZAKI ? (slices ? slices : 4 * L - 4) : 0;
        return sdpg_set_gemm(ctx, gemm_mode, slices);
    }

    std::string fmt(const T &v, int digits) { return mpx::format_decimal<L>(v, digits); }

    // one value in a printf-style exponent format of the parameter file ("%+50.40Fe": flags, width,
    // digits after the point); the reference passes these formats to gmp_fprintf with an mpf value
    // (sdpa_struct.cpp:93-137, 706-740)
    std::string fmt_param(const T &v, const std::string &pf) {
        size_t q = 1;
        bool plus = false, space = false, left = false;
        while (q < pf.size() && strchr("+- 0#", pf[q])) {
            plus |= pf[q] == '+';
            space |= pf[q] == ' ';
            left |= pf[q] == '-';
            q++;
        }
        int width = 0, prec = 6;
        while (q < pf.size() && isdigit((unsigned char)pf[q])) width = width * 10 + (pf[q++] - '0');
        if (q < pf.size() && pf[q] == '.') {
            prec = 0;
            q++;
            while (q < pf.size() && isdigit((unsigned char)pf[q])) prec = prec * 10 + (pf[q++] - '0');
        }
        std::string body = mpx::format_decimal<L>(mpx::absv<L>(v), prec + 1);
        std::string sgn = mpx::sgn<L>(v) < 0 ? "-" : (plus ? "+" : (space ? " " : ""));
        std::string out = sgn + body;
        if ((int)out.size() < width) out = left ? out + std::string(width - out.size(), ' ')
                                               : std::string(width - out.size(), ' ') + out;
        return out;
    }

    // the solution printout of printLastInfo (sdpa_io.cpp:1167-1219, REVERSE_PRIMAL_DUAL): xVec = -y,
    // xMat = Z and yMat = X block by block (DenseMatrix::display rows; LP blocks as one row)
    void print_solution(FILE *f) {
        std::vector<uint32_t> buf;
        auto fetch = [&](int what, int64_t cnt, std::vector<T> &v) {
            buf.assign((size_t)cnt * W, 0);
            v.assign(cnt, mpx::zero<L>());
            if (sdpg_get(ctx, what, buf.data()) == 0)
                for (int64_t i = 0; i < cnt; i++) v[i] = unpack1(&buf[(size_t)i * W]);
        };
        std::vector<T> yv, Zv, Xv;
        fetch(SDPG_Y, m, yv);
        fetch(SDPG_Z, Nd, Zv);
        fetch(SDPG_X, Nd, Xv);
        fprintf(f, "xVec = \n");
        if (prm.xPrint == "NOPRINT") {
            fprintf(f, "NOPRINT\n");
        } else {
            fprintf(f, "{");
            for (int j = 0; j < m; j++)
                fprintf(f, "%s%s", fmt_param(mpx::neg<L>(yv[j]), prm.xPrint).c_str(), j + 1 < m ? "," : "}\n");
            if (m == 0) fprintf(f, "  }\n");
        }
        auto space = [&](const char *name, const std::vector<T> &D, const std::string &pf) {
            fprintf(f, "%s = \n", name);
            if (pf == "NOPRINT") {
                fprintf(f, "NOPRINT\n");
                return;
            }
            fprintf(f, "{\n");
            for (size_t bl = 0; bl < blockType.size(); bl++) {
                if (blockType[bl] == 1) {
                    const int l = blockNumber[bl], n = sdpStruct[l];
                    const T *a = &D[off[l]];
                    fprintf(f, "{");
                    for (int i = 0; i < n - 1; i++) {
                        fprintf(f, "%s{", i == 0 ? " " : "  ");
                        for (int j = 0; j < n; j++)
                            fprintf(f, "%s%s", fmt_param(a[i + (size_t)n * j], pf).c_str(), j + 1 < n ? "," : " },\n");
                    }
                    if (n > 1) fprintf(f, "  {");
                    for (int j = 0; j < n; j++)
                        fprintf(f, "%s%s", fmt_param(a[(n - 1) + (size_t)n * j], pf).c_str(), j + 1 < n ? "," : " }");
                    fprintf(f, n > 1 ? "   }\n" : "\n");
                } else {
                    const int n = blockStructAbs[bl], o = blockNumber[bl];
                    fprintf(f, "{");
                    for (int j = 0; j < n; j++)
                        fprintf(f, "%s%s", fmt_param(D[off[nsdp] + o + j], pf).c_str(), j + 1 < n ? "," : "}\n");
                    if (n <= 0) fprintf(f, "  }\n");
                }
            }
            fprintf(f, "}\n");
        };
        space("xMat", Zv, prm.XPrint);
        space("yMat", Xv, prm.YPrint);
    }
    std::string fmt_signed(const T &v, int digits) {
        std::string s = fmt(v, digits);
        return (s[0] == '-') ? s : "+" + s;
    }

    int upload_point() {
        std::vector<uint32_t> px, pz, piz, py;
        pack(X, px);
        pack(Z, pz);
        pack(invZ, piz);
        pack(y, py);
        return sdpg_set_point(ctx, px.data(), pz.data(), piz.data(), py.data());
    }

    // Residuals::compute on the device; returns the norms
    int residuals(T &nrp, T &nrd) {
        std::vector<uint32_t> a(W), c(W);
        int rc = ex->residuals(ctx, a.data(), c.data());
        if (rc) return rc;
        nrp = unpack1(a.data());
        nrd = unpack1(c.data());
        return 0;
    }

    // Newton::Mehrotra (sdpa_newton.cpp:1375-1414) as C-ABI calls; false on Cholesky failure
    int mehrotra(bool predictor, const T &mu, const T &beta, int phase, bool &ok) {
        ok = true;
        auto pm = pack1(mu), pbeta = pack1(beta);
        int info = 0;
        int rc = ex->step(ctx, predictor, pm.data(), pbeta.data(), phase, &info);
        if (rc) return rc;
        if (info > 0) {
            if (logf) fprintf(logf, "cholesky miss condition :: not positive definite :: info = %d\n", info);
            ok = false;
        }
        return 0;
    }
    T ip(int what) {
        std::vector<uint32_t> w(W);
        int rc = ex->inner(ctx, what, w.data());
        if (rc) {
            ip_err = rc;
            return mpx::zero<L>();
        }
        return unpack1(w.data());
    }
    int ip_err = 0;

    // SolveInfo (sdpa_parts.cpp:713-830)
    struct SolveInfo {
        T objValPrimal, objValDual, rho, etaPrimal, etaDual;
    };

    bool phase_update(int &phase, const T &nrp, const T &nrd, const SolveInfo &si) {
        T acc = M<L>::d(prm.epsilonDash), NONZERO = M<L>::d(1.0e-6), one = M<L>::i(1), two = M<L>::i(2);
        int v = noINFO;
        if (nrp <= acc) v = (nrd <= acc) ? pdFEAS : pFEAS;
        if (v == noINFO && nrd <= acc) v = dFEAS;
        phase = v;
        T thr = one + NONZERO;
        if (v == pdFEAS) {
            T mean = (abs(si.objValPrimal) + abs(si.objValDual)) / two;
            T gap = abs(si.objValPrimal - si.objValDual);
            T dom = (mean < one) ? one : mean;
            if (gap / dom <= M<L>::d(prm.epsilonStar)) {
                phase = pdOPT;
                return false;
            }
        }
        if (v == noINFO && si.rho > thr) {
            phase = pdINF;
            return false;
        }
        if (v == pFEAS) {
            if (si.objValPrimal <= M<L>::d(-prm.upperBound)) {
                phase = pUNBD;
                return false;
            }
            if (si.rho > thr) {
                phase = pFEAS_dINF;
                return false;
            }
        }
        if (v == dFEAS) {
            if (si.objValDual >= M<L>::d(-prm.lowerBound)) {
                phase = dUNBD;
                return false;
            }
            if (si.rho > thr) {
                phase = pINF_dFEAS;
                return false;
            }
        }
        return true;
    }

    void print_iter(int it, const T &mu, const T &thP, const T &thD, const SolveInfo &si, const T &aP, const T &aD,
                    const T &beta) {
        if (!logf) return;
        fprintf(logf, "%2d %s %s %s %s %s %s %s %s\n", it, fmt(mu, 2).c_str(), fmt(thD, 2).c_str(),
                fmt(thP, 2).c_str(), fmt_signed(-si.objValDual, 3).c_str(), fmt_signed(-si.objValPrimal, 3).c_str(),
                fmt(aD, 2).c_str(), fmt(aP, 2).c_str(), fmt(beta, 3).c_str());
        fflush(logf);
    }

    // pinpal main loop (sdpa_main.cpp:214-398); stop_after >= 0 ends the loop early (bench warm-up)
    int run(sdpa_amd_result *res, int stop_after) {
        using MM = M<L>;
        compute_formula();
        const T lam = MM::d(prm.lambdaStar), one = MM::i(1);
        // initial point (sdpa_main.cpp / Solutions::initialize): X = Z = lambda I, y = 0
        auto identity = [&](std::vector<T> &d, const T &v) {
            d.assign(Nd, mpx::zero<L>());
            for (int l = 0; l < nsdp; l++) {
                int n = sdpStruct[l];
                for (int i = 0; i < n; i++) d[off[l] + i + (size_t)n * i] = v;
            }
            for (int q = 0; q < nlp; q++) d[off[nsdp] + q] = v;
        };
        identity(X, lam);
        identity(Z, lam);
        identity(invZ, one / lam);
        y.assign(m, mpx::zero<L>());
        if (init) {
            // an initial-point file overwrites the entries it lists (IO::read, sdpa_io.cpp:150-274, after
            // Solutions::initialize, sdpa_main.cpp:234-243); y is stored with the sign flipped
            for (int k = 0; k < m; k++) y[k] = mpx::neg<L>(parse(init->y[k]));
            for (const RawEntry &e : init->ent) {
                std::vector<T> &D = e.k == 1 ? Z : X;
                const T v = parse(e.v);
                if (e.l <= nsdp) {
                    const int l = e.l - 1, n = sdpStruct[l];
                    if (e.i < 1 || e.j < 1 || e.i > n || e.j > n) {
                        err = "out of range in the initial point";
                        return -34;
                    }
                    D[off[l] + (e.i - 1) + (size_t)n * (e.j - 1)] = v;
                    D[off[l] + (e.j - 1) + (size_t)n * (e.i - 1)] = v;
                } else {
                    if (e.i < 1 || e.i > nlp) {
                        err = "out of range in the initial point";
                        return -34;
                    }
                    D[off[nsdp] + e.i - 1] = v;
                }
            }
        }
        int rc = upload_point();
        if (rc) return rc;
        if (init) {  // initPt_xMat / initPt_zMat (sdpa_main.cpp:246-249)
            std::vector<uint32_t> px, pz;
            pack(X, px);
            pack(Z, pz);
            if ((rc = sdpg_put(ctx, SDPG_X0, px.data())) || (rc = sdpg_put(ctx, SDPG_Z0, pz.data()))) return rc;
        }
        X.clear();
        Z.clear();
        invZ.clear();
        X.shrink_to_fit();
        Z.shrink_to_fit();
        invZ.shrink_to_fit();
        {  // Solutions::initialize / computeInverse: the inverse factors of the initial point, on the device
            int info = 0;
            if ((rc = ex->factor_point(ctx, &info))) return rc;
            if (info) {
                err = "the initial point is not positive definite";
                return -35;
            }
        }
        T initNrp, initNrd;
        if ((rc = residuals(initNrp, initNrd))) return rc;
        T curNrp = initNrp, curNrd = initNrd;
        T alphaP = MM::i(0), alphaD = MM::i(0), beta = MM::d(prm.betaStar);
        bool reductionOn = true;
        // AverageComplementarity: lambda*^2, or X . Z / n for an initial-point file (sdpa_parts.cpp:634-645)
        T mu_initial = init ? ip(SDPG_IP_XZ) / MM::i(nDim) : lam * lam, mu_current = mu_initial;
        T acc = MM::d(prm.epsilonDash);
        T thetaP = (initNrp < acc) ? MM::i(0) : one;
        T thetaD = (initNrd < acc) ? MM::i(0) : one;
        SolveInfo si;
        const T nD = MM::i(nDim), omega = MM::d(prm.omegaStar);
        si.rho = one;
        si.etaPrimal = omega * nD * mu_initial;
        si.etaDual = si.etaPrimal;
        si.objValPrimal = ip(SDPG_IP_CX);
        si.objValDual = ip(SDPG_IP_BY);
        int phase = noINFO;
        phase_update(phase, initNrp, initNrd, si);
        if (logf) fprintf(logf, "   mu      thetaP  thetaD  objP      objD      alphaP  alphaD  beta \n");
        const T c09 = MM::d(0.9), alphaBD = MM::d(100.0), inv_alphaBD = one / alphaBD, xi = MM::d(3.0),
                c095 = MM::d(0.95);
        const T gammaStar = MM::d(prm.gammaStar), betaStar = MM::d(prm.betaStar), betaBar = MM::d(prm.betaBar),
                nu = MM::d(2.0), c1em6 = MM::d(1.0e-6), c1em4 = MM::d(1.0e-4), five = MM::d(5.0),
                c1em10 = MM::d(1.0e-10);
        int pIteration = 0, status = 0;
        double t0 = now_s();
        while (phase_update(phase, curNrp, curNrd, si) && pIteration < prm.maxIteration) {
            if (stop_after >= 0 && pIteration >= stop_after) break;
            reductionOn = (phase == noINFO || phase == pFEAS || phase == dFEAS);
            if (phase == pdFEAS)
                beta = mpx::zero<L>();
            else
                beta = reductionOn ? betaBar : nu;
            bool ok;
            if ((rc = mehrotra(true, mu_current, beta, phase, ok))) return rc;
            if (!ok) {
                status = 1;
                break;
            }
            // StepLength::MehrotraPredictor (sdpa_parts.cpp:362-410)
            alphaP = c09;
            alphaD = c09;
            auto clamp_steps = [&]() {
                if (phase == noINFO || phase == dFEAS) {
                    if (alphaP > one) alphaP = one;
                } else {
                    T inc = ip(SDPG_IP_CDX);
                    if (mpx::sgn<L>(inc) > 0) {
                        if (alphaP > alphaD) alphaP = alphaD;
                        if (alphaP > one) alphaP = one;
                    }
                }
                if (phase == noINFO || phase == pFEAS) {
                    if (alphaD > one) alphaD = one;
                } else {
                    T inc = ip(SDPG_IP_BDY);
                    if (mpx::sgn<L>(inc) < 0) {
                        if (alphaD > alphaP) alphaD = alphaP;
                        if (alphaD > one) alphaD = one;
                    }
                }
            };
            clamp_steps();
            // DirectionParameter::MehrotraCorrector (sdpa_parts.cpp:614-645)
            {
                T xv = ip(SDPG_IP_XDZ), uz = ip(SDPG_IP_DXZ), uv = ip(SDPG_IP_DXDZ);
                T muT = mu_current + (alphaP * uz + alphaD * xv + alphaP * alphaD * uv) / nD;
                beta = muT / mu_current;
                if (beta < one) beta = beta * beta;
                if (phase == pdFEAS) {
                    if (beta < betaStar) beta = betaStar;
                    if (beta > one) beta = one;
                } else {
                    if (beta < betaBar) beta = betaBar;
                }
            }
            if (phase == pdFEAS && (beta > five || si.objValPrimal < si.objValDual)) break;
            if ((rc = mehrotra(false, mu_current, beta, phase, ok))) return rc;
            // StepLength::MehrotraCorrector (sdpa_parts.cpp:412-523)
            {
                T mnx, mnz;
                {
                    std::vector<uint32_t> wx(W), wz(W);
                    if ((rc = ex->min_eigen2(ctx, wx.data(), wz.data()))) return rc;
                    mnx = unpack1(wx.data());
                    mnz = unpack1(wz.data());
                }
                alphaP = (-mnx > inv_alphaBD) ? -(one / mnx) : alphaBD;
                alphaD = (-mnz > inv_alphaBD) ? -(one / mnz) : alphaBD;
                alphaP = gammaStar * alphaP;
                alphaD = gammaStar * alphaD;
                clamp_steps();
                if (reductionOn && (phase == noINFO || phase == pFEAS || phase == dFEAS)) {
                    T xv = ip(SDPG_IP_XDZ), uz = ip(SDPG_IP_DXZ), uv = ip(SDPG_IP_DXDZ);
                    auto thetaMaxF = [&]() {
                        T a1 = (one - alphaP) * thetaP, a2 = (one - alphaD) * thetaD;
                        return (a1 > a2) ? a1 : a2;
                    };
                    auto muNewF = [&]() { return mu_current + (alphaP * uz + alphaD * xv + alphaP * alphaD * uv) / nD; };
                    T thetaMax = thetaMaxF(), muNew = muNewF();
                    while (thetaMax * thetaMax * mu_initial > xi * muNew) {
                        T amax = c095 * ((alphaP > alphaD) ? alphaP : alphaD);
                        if (alphaP > amax) alphaP = amax;
                        if (alphaD > amax) alphaD = amax;
                        thetaMax = thetaMaxF();
                        muNew = muNewF();
                        if (alphaP < c1em6 && alphaD < c1em6) break;
                    }
                }
                if (phase == pdFEAS) {
                    T oD = ip(SDPG_IP_BY), oP = ip(SDPG_IP_CX);
                    T iD = ip(SDPG_IP_BDY) * alphaD, iP = ip(SDPG_IP_CDX) * alphaP;
                    T den = iP - iD;
                    if (!mpx::is_zero<L>(den)) {
                        T mr = (oD - oP) / den;
                        if (mpx::sgn<L>(mr) > 0 && mr < one) {
                            alphaP = alphaP * mr;
                            alphaD = alphaD * mr;
                        }
                    }
                }
            }
            print_iter(pIteration, mu_current, thetaP, thetaD, si, alphaP, alphaD, beta);
            // Solutions::update (sdpa_dataset.cpp:120-143), on the device
            if ((rc = ex->update(ctx, pack1(alphaP).data(), pack1(alphaD).data()))) return rc;
            if (alphaP < c1em4 && alphaD < c1em4) {
                if (logf) fprintf(logf, "Step length is too small. \ncannot move\n");
                status = 2;
                break;
            }
            {
                int info = 0;
                if ((rc = ex->factor_point(ctx, &info))) return rc;
                if (info) {
                    if (logf) fprintf(logf, "cannot move\n");
                    status = 2;
                    break;
                }
            }
            if (reductionOn) {
                thetaP = abs((one - alphaP) * thetaP);
                thetaD = abs((one - alphaD) * thetaD);
            }
            mu_current = ip(SDPG_IP_XZ) / nD;
            if ((rc = residuals(curNrp, curNrd))) return rc;
            thetaP = (initNrp > c1em10) ? curNrp / initNrp : MM::i(0);
            thetaD = (initNrd > c1em10) ? curNrd / initNrd : MM::i(0);
            // SolveInfo::update (sdpa_parts.cpp:747-830): with an initial-point file against that point,
            // otherwise against lambda I
            {
                si.objValPrimal = ip(SDPG_IP_CX);
                si.objValDual = ip(SDPG_IP_BY);
                T pr = thetaP, du = thetaD, x0z0 = nD * mu_initial, xz = nD * mu_current;
                // X0 . Z and X . Z0; for lambda I the reference sums lambda * diagonal entries
                T x0z = init ? ip(SDPG_IP_X0Z) : lam * ip(SDPG_IP_TRZ);
                T xz0 = init ? ip(SDPG_IP_XZ0) : lam * ip(SDPG_IP_TRX);
                si.rho = mpx::zero<L>();
                bool pok = curNrp <= acc, dok = curNrd <= acc;
                if (pok && xz0 < si.etaPrimal) si.etaPrimal = xz0;
                if (dok && x0z < si.etaDual) si.etaDual = x0z;
                if (!pok && dok) si.rho = pr * x0z / ((pr + (one - pr) * omega) * si.etaDual + xz);
                if (pok && !dok) si.rho = du * xz0 / ((du + (one - du) * omega) * si.etaPrimal + xz);
                if (!pok && !dok)
                    si.rho = (du * xz0 + pr * x0z) /
                             ((pr * du + omega * (pr * (one - du) + (one - pr) * du)) * x0z0 + xz);
            }
            if (ip_err) return ip_err;
            pIteration++;
        }
        if (ip_err) return ip_err;
        double t1 = now_s();
        if ((rc = residuals(curNrp, curNrd))) return rc;
        print_iter(pIteration, mu_current, thetaP, thetaD, si, alphaP, alphaD, beta);
        int rphase = phase_reverse(phase);
        T mean = (abs(si.objValPrimal) + abs(si.objValDual)) / MM::i(2);
        T gap = abs(si.objValPrimal - si.objValDual);
        T relgap = (mean < one) ? gap : gap / mean;
        if (res) {
            memset(res, 0, sizeof(*res));
            res->iterations = pIteration;
            res->status = status;
            snprintf(res->phase, sizeof res->phase, "%s", phase_str(rphase));
            snprintf(res->objP, sizeof res->objP, "%s", fmt(-si.objValDual, 41).c_str());
            snprintf(res->objD, sizeof res->objD, "%s", fmt(-si.objValPrimal, 41).c_str());
            snprintf(res->objP70, sizeof res->objP70, "%s", fmt(-si.objValDual, 70).c_str());
            snprintf(res->objD70, sizeof res->objD70, "%s", fmt(-si.objValPrimal, 70).c_str());
            snprintf(res->relgap, sizeof res->relgap, "%s", fmt(relgap, 17).c_str());
            snprintf(res->pfeas, sizeof res->pfeas, "%s", fmt(curNrd, 17).c_str());
            snprintf(res->dfeas, sizeof res->dfeas, "%s", fmt(curNrp, 17).c_str());
            res->mu = mpx::to_double<L>(mu_current);
            double tm[8] = {0};
            sdpg_timers(ctx, tm, 8);
            res->t_schur_ms = tm[0];
            res->t_chol_ms = tm[1];
            res->t_solve_ms = tm[2];
            res->t_gvec_ms = tm[3];
            res->t_dxdz_ms = tm[4];
            res->t_res_ms = tm[5];
            res->t_inv_ms = tm[6];
            res->t_eig_ms = tm[7];
            res->t_total_s = t1 - t0;
            res->limbs = L;
            res->slices = used_slices;
        }
        if (logf) {
            fprintf(logf, "\nphase.value = %s\n", phase_str(rphase));
            fprintf(logf, "   Iteration = %d\n", pIteration);
            fprintf(logf, "          mu = %s\n", fmt(mu_current, 17).c_str());
            fprintf(logf, "relative gap = %s\n", fmt(relgap, 17).c_str());
            fprintf(logf, "objValPrimal = %s\n", fmt(-si.objValDual, 41).c_str());
            fprintf(logf, "objValDual   = %s\n", fmt(-si.objValPrimal, 41).c_str());
            fprintf(logf, "p.feas.error = %s\n", fmt(curNrd, 17).c_str());
            fprintf(logf, "d.feas.error = %s\n", fmt(curNrp, 17).c_str());
            // relative eps: 2^-(mpf precision bits) of the request (Rlamch "E", mplapack/Rlamch.cpp:36-45)
            const int pbits = 64 * ((std::max(53, prm.precision) + 127) / 64) - 64;
            fprintf(logf, "relative eps = %s\n", fmt(mpx::ldexp2<L>(MM::i(1), -pbits), 17).c_str());
            fprintf(logf, "total time   = %.3f\n", t1 - t0);
            if (res)
                fprintf(logf, "Make bMat    = %.3f ms (device)\nCholesky bMat= %.3f ms (device)\nmain loop    = %.3f s\n",
                        res->t_schur_ms, res->t_chol_ms, res->t_total_s);
            fprintf(logf, "\n\nParameters are\n** Parameters **\n");
            fprintf(logf, "maxIteration =    %d\nepsilonStar  = %8.3e\nlambdaStar   = %8.3e\nomegaStar    = %8.3e\n",
                    prm.maxIteration, prm.epsilonStar, prm.lambdaStar, prm.omegaStar);
            fprintf(logf, "lowerBound   = %8.3e\nupperBound   = %8.3e\nbetaStar     = %8.3e\nbetaBar      = %8.3e\n",
                    prm.lowerBound, prm.upperBound, prm.betaStar, prm.betaBar);
            fprintf(logf, "gammaStar    = %8.3e\nepsilonDash  = %8.3e\nprecision    =    %d\n", prm.gammaStar,
                    prm.epsilonDash, prm.precision);
            fprintf(logf, "xPrint       = %s \nXPrint       = %s \nYPrint       = %s \ninfPrint     = %s \n\n",
                    prm.xPrint.c_str(), prm.XPrint.c_str(), prm.YPrint.c_str(), prm.infPrint.c_str());
            print_solution(logf);
        }
        return 0;
    }
};

template <int L>
static int solve_L(const RawProblem &R, const Parameter &prm, const char *log_path, int gemm_mode, int slices,
                   int device, sdpa_amd_result *res, const RawInit *init = nullptr) {
    Solver<L> S;
    S.prm = prm;
    S.init = init;
    int rc = S.load(R);
    if (rc) return rc;
    S.compute_formula();
    if ((rc = S.setup_device(device, gemm_mode, slices))) return rc;
    if (log_path) S.logf = strcmp(log_path, "-") == 0 ? stdout : fopen(log_path, "w");
    rc = S.run(res, -1);
    if (S.logf && S.logf != stdout) fclose(S.logf);
    return rc;
}

template <int L>
static int bench_L(const RawProblem &R, const Parameter &prm, int gemm_mode, int slices, int device, int warm_iters,
                   int warmup, int steps, double *ms_schur, double *ms_chol, double *ms_wall, double *kstats,
                   const char *dump_prefix) {
    Solver<L> S;
    S.prm = prm;
    int rc = S.load(R);
    if (rc) return rc;
    S.compute_formula();
    if ((rc = S.setup_device(device, gemm_mode, slices))) return rc;
    if ((rc = S.run(nullptr, warm_iters))) return rc;
    if (dump_prefix) {  // the iterate (X, invZ) for a CPU replay of the same step
        const int W = L + 2;
        std::vector<uint32_t> buf((size_t)S.Nd * W);
        const int sel[2] = {SDPG_X, SDPG_INVZ};
        const char *nm[2] = {"X.bin", "invZ.bin"};
        for (int q = 0; q < 2; q++) {
            if ((rc = sdpg_get(S.ctx, sel[q], buf.data()))) return rc;
            std::string path = std::string(dump_prefix) + nm[q];
            FILE *f = fopen(path.c_str(), "wb");
            if (!f) return -21;
            size_t wr = fwrite(buf.data(), sizeof(uint32_t), buf.size(), f);
            fclose(f);
            if (wr != buf.size()) return -21;
        }
    }
    // the device holds the point of iteration warm_iters; time [schur + potrf] on it
    for (int w = 0; w < warmup; w++) {
        int info = 0;
        if ((rc = sdpg_schur(S.ctx)) || (rc = sdpg_potrf(S.ctx, &info))) return rc;
    }
    // timed steps with no per-kernel events in the stream (the wall time and the schur / potrf timers)
    sdpg_reset_timers(S.ctx);
    double t0 = now_s();
    for (int s = 0; s < steps; s++) {
        int info = 0;
        if ((rc = sdpg_schur(S.ctx)) || (rc = sdpg_potrf(S.ctx, &info))) return rc;
    }
    sdpg_sync(S.ctx);
    double t1 = now_s();
    double tm[8] = {0};
    sdpg_timers(S.ctx, tm, 8);
    *ms_schur = tm[0] / (steps > 0 ? steps : 1);
    *ms_chol = tm[1] / (steps > 0 ? steps : 1);
    *ms_wall = 1e3 * (t1 - t0) / (steps > 0 ? steps : 1);
    // the same steps again with events around the product and diagonal-block kernels (per-kernel times)
    if (kstats) {
        sdpg_profile(S.ctx, 1, nullptr, 0);
        for (int s = 0; s < steps; s++) {
            int info = 0;
            if ((rc = sdpg_schur(S.ctx)) || (rc = sdpg_potrf(S.ctx, &info))) return rc;
        }
        sdpg_sync(S.ctx);
        sdpg_profile(S.ctx, 0, kstats, 12);
    }
    return 0;
}

// The meeting point of the lockstep solves: every problem's thread posts its next device call and waits;
// the last of the running problems to arrive issues one batched call per kind of call posted
// (sdpg_batch_*) for all of them and wakes the others.  A problem whose loop has ended leaves, which may
// complete a meeting.  The loop's calls come in fixed runs, so a meeting does a whole run and keeps the
// results the problem asks for next: a step (Make_gVec, Newton, dZ, dX) is followed by the inner
// products and, for a corrector, by the step-length eigenvalues; an update by the inverse factors, the
// inner products and the residuals.  Each problem's cached results stay valid until it posts its next
// step or update, so an iteration costs three meetings.
struct Meeting {
    enum { K_STEP, K_INNER, K_EIG, K_UPDATE, K_FACTOR, K_RESID };
    struct Post {
        int kind = 0;
        bool pred = false;
        int phase = 0, what = 0;
        const uint32_t *a = nullptr, *b = nullptr;
        uint32_t *o1 = nullptr, *o2 = nullptr;
        int *info = nullptr;
    };
    // per problem: results of its last meeting that it has not asked for yet
    struct Cache {
        bool ip = false, eig = false, fac = false, res = false;
        int info = 0;
    };
    std::mutex mu;
    std::condition_variable cv;
    sdpg_batch *batch = nullptr;
    const int n, W;
    int active, arrived = 0, rc = 0;
    uint64_t gen = 0;
    std::vector<Post> post;
    std::vector<Cache> cache;
    std::vector<char> posted;
    std::vector<uint32_t> ipc;  // per problem, the ten inner products (SDPG_IP_XZ .. SDPG_IP_TRZ)
    std::vector<uint32_t> eigc;  // per problem, the two step-length eigenvalues
    std::vector<uint32_t> resc;  // per problem, max |Rp|, max |Rd|, C.X, b.y
    std::vector<int> req, req2, phase, info;
    std::vector<uint32_t> sa, sb, out;
    long long meetings = 0, served = 0;

    Meeting(sdpg_batch *b, int n_, int W_)
        : batch(b), n(n_), W(W_), active(n_), post(n_), cache(n_), posted(n_, 0), ipc((size_t)n_ * 10 * W_),
          eigc((size_t)n_ * 2 * W_), resc((size_t)n_ * 4 * W_), req(n_), req2(n_), phase(n_), info(n_),
          sa((size_t)n_ * W_), sb((size_t)n_ * W_), out((size_t)n_ * 10 * W_) {}

    int meet(int p, const Post &q) {
        std::unique_lock<std::mutex> lk(mu);
        post[p] = q;
        posted[p] = 1;
        arrived++;
        const uint64_t g = gen;
        if (arrived == active)
            fire();
        else
            cv.wait(lk, [&] { return gen != g; });
        return rc;
    }
    void leave() {
        std::lock_guard<std::mutex> lk(mu);
        active--;
        if (active > 0 && arrived == active) fire();
    }
    bool select(int kind) {
        bool any = false;
        for (int p = 0; p < n; p++) {
            req[p] = posted[p] && post[p].kind == kind;
            any |= req[p] != 0;
        }
        return any;
    }
    void note(int r) {
        if (r && !rc) rc = r;
    }
    bool any(const std::vector<int> &r) const {
        for (int v : r)
            if (v) return true;
        return false;
    }
    void fetch_inner(const std::vector<int> &r) {
        note(sdpg_batch_inner(batch, r.data(), out.data()));
        for (int p = 0; p < n; p++)
            if (r[p]) {
                memcpy(&ipc[(size_t)p * 10 * W], &out[(size_t)p * 10 * W], sizeof(uint32_t) * 10 * W);
                cache[p].ip = true;
            }
    }
    void fetch_eig(const std::vector<int> &r) {
        note(sdpg_batch_min_eigen(batch, r.data(), 0, out.data()));
        for (int p = 0; p < n; p++)
            if (r[p]) memcpy(&eigc[(size_t)p * 2 * W], &out[(size_t)p * W], sizeof(uint32_t) * W);
        note(sdpg_batch_min_eigen(batch, r.data(), 1, out.data()));
        for (int p = 0; p < n; p++)
            if (r[p]) {
                memcpy(&eigc[((size_t)p * 2 + 1) * W], &out[(size_t)p * W], sizeof(uint32_t) * W);
                cache[p].eig = true;
            }
    }
    void fetch_factor(const std::vector<int> &r) {
        note(sdpg_batch_factor_point(batch, r.data(), info.data()));
        for (int p = 0; p < n; p++)
            if (r[p]) {
                cache[p].info = info[p];
                cache[p].fac = true;
            }
    }
    void fetch_resid(const std::vector<int> &r) {
        note(sdpg_batch_residuals(batch, r.data(), out.data()));
        for (int p = 0; p < n; p++)
            if (r[p]) {
                memcpy(&resc[(size_t)p * 4 * W], &out[(size_t)p * 4 * W], sizeof(uint32_t) * 4 * W);
                cache[p].res = true;
            }
    }
    void fire() {  // with mu held
        if (select(K_STEP)) {
            for (int p = 0; p < n; p++) {
                if (!req[p]) continue;
                req[p] = post[p].pred ? SDPG_BATCH_PREDICTOR : SDPG_BATCH_CORRECTOR;
                phase[p] = post[p].phase;
                memcpy(&sa[(size_t)p * W], post[p].a, sizeof(uint32_t) * W);
                memcpy(&sb[(size_t)p * W], post[p].b, sizeof(uint32_t) * W);
            }
            note(sdpg_batch_gvec(batch, req.data(), sa.data(), sb.data(), phase.data()));
            note(sdpg_batch_newton(batch, req.data(), info.data()));
            for (int p = 0; p < n; p++) {
                req2[p] = req[p] && !(req[p] == SDPG_BATCH_PREDICTOR && info[p] > 0);
                if (req[p]) *post[p].info = req[p] == SDPG_BATCH_PREDICTOR ? info[p] : 0;
            }
            note(sdpg_batch_dz(batch, req2.data(), phase.data()));
            note(sdpg_batch_dx(batch, req2.data()));
            fetch_inner(req2);  // the step-length and corrector-parameter inner products come next
            for (int p = 0; p < n; p++) req2[p] = req2[p] && req[p] == SDPG_BATCH_CORRECTOR;
            if (any(req2)) fetch_eig(req2);  // after a corrector, the step-length eigenvalues
        }
        if (select(K_UPDATE)) {
            for (int p = 0; p < n; p++) {
                if (!req[p]) continue;
                memcpy(&sa[(size_t)p * W], post[p].a, sizeof(uint32_t) * W);
                memcpy(&sb[(size_t)p * W], post[p].b, sizeof(uint32_t) * W);
            }
            note(sdpg_batch_update(batch, req.data(), sa.data(), sb.data()));
            // then the inverse factors, mu (X.Z), the residuals and SolveInfo's inner products
            fetch_factor(req);
            fetch_inner(req);
            fetch_resid(req);
        }
        if (select(K_INNER)) {
            fetch_inner(req);
            for (int p = 0; p < n; p++)
                if (req[p]) memcpy(post[p].o1, &ipc[((size_t)p * 10 + post[p].what - 1) * W], sizeof(uint32_t) * W);
        }
        if (select(K_EIG)) {
            fetch_eig(req);
            for (int p = 0; p < n; p++)
                if (req[p]) {
                    memcpy(post[p].o1, &eigc[(size_t)p * 2 * W], sizeof(uint32_t) * W);
                    memcpy(post[p].o2, &eigc[((size_t)p * 2 + 1) * W], sizeof(uint32_t) * W);
                    cache[p].eig = false;
                }
        }
        if (select(K_FACTOR)) {
            fetch_factor(req);
            for (int p = 0; p < n; p++)
                if (req[p]) {
                    *post[p].info = cache[p].info;
                    cache[p].fac = false;
                }
        }
        if (select(K_RESID)) {
            fetch_resid(req);
            for (int p = 0; p < n; p++)
                if (req[p]) {
                    memcpy(post[p].o1, &resc[(size_t)p * 4 * W], sizeof(uint32_t) * W);
                    memcpy(post[p].o2, &resc[((size_t)p * 4 + 1) * W], sizeof(uint32_t) * W);
                    cache[p].res = false;
                }
        }
        meetings++;
        served += arrived;
        std::fill(posted.begin(), posted.end(), 0);
        arrived = 0;
        gen++;
        cv.notify_all();
    }
};

// one problem's side of the meetings (its cache entries are touched by its own thread, or by the thread
// running a meeting while this one waits in it)
struct Port : Exec {
    Meeting *mt;
    int p;
    Port(Meeting *m, int p_) : mt(m), p(p_) {}
    void invalidate() { mt->cache[p] = Meeting::Cache(); }
    int step(sdpg_ctx *, bool predictor, const uint32_t *mu, const uint32_t *beta, int phase, int *info) override {
        invalidate();
        Meeting::Post q;
        q.kind = Meeting::K_STEP;
        q.pred = predictor;
        q.phase = phase;
        q.a = mu;
        q.b = beta;
        q.info = info;
        return mt->meet(p, q);
    }
    int inner(sdpg_ctx *, int what, uint32_t *out) override {
        if (what < SDPG_IP_XZ || what > SDPG_IP_TRZ) return SDPG_EARG;  // no initial-point products here
        if (mt->cache[p].ip) {
            memcpy(out, &mt->ipc[((size_t)p * 10 + what - 1) * mt->W], sizeof(uint32_t) * mt->W);
            return 0;
        }
        Meeting::Post q;
        q.kind = Meeting::K_INNER;
        q.what = what;
        q.o1 = out;
        return mt->meet(p, q);
    }
    int min_eigen2(sdpg_ctx *, uint32_t *x, uint32_t *z) override {
        if (mt->cache[p].eig) {
            memcpy(x, &mt->eigc[(size_t)p * 2 * mt->W], sizeof(uint32_t) * mt->W);
            memcpy(z, &mt->eigc[((size_t)p * 2 + 1) * mt->W], sizeof(uint32_t) * mt->W);
            mt->cache[p].eig = false;
            return 0;
        }
        Meeting::Post q;
        q.kind = Meeting::K_EIG;
        q.o1 = x;
        q.o2 = z;
        return mt->meet(p, q);
    }
    int update(sdpg_ctx *, const uint32_t *aP, const uint32_t *aD) override {
        invalidate();
        Meeting::Post q;
        q.kind = Meeting::K_UPDATE;
        q.a = aP;
        q.b = aD;
        return mt->meet(p, q);
    }
    int factor_point(sdpg_ctx *, int *info) override {
        if (mt->cache[p].fac) {
            *info = mt->cache[p].info;
            mt->cache[p].fac = false;
            return 0;
        }
        Meeting::Post q;
        q.kind = Meeting::K_FACTOR;
        q.info = info;
        return mt->meet(p, q);
    }
    int residuals(sdpg_ctx *, uint32_t *nrp, uint32_t *nrd) override {
        if (mt->cache[p].res) {
            memcpy(nrp, &mt->resc[(size_t)p * 4 * mt->W], sizeof(uint32_t) * mt->W);
            memcpy(nrd, &mt->resc[((size_t)p * 4 + 1) * mt->W], sizeof(uint32_t) * mt->W);
            mt->cache[p].res = false;
            return 0;
        }
        Meeting::Post q;
        q.kind = Meeting::K_RESID;
        q.o1 = nrp;
        q.o2 = nrd;
        return mt->meet(p, q);
    }
};

template <int L>
static int batch_solve_L(const std::vector<RawProblem> &R, const Parameter &prm, int device, int use_graph,
                         sdpa_amd_result *res, double *latency, double *stats) {
    const int n = (int)R.size();
    std::vector<std::unique_ptr<Solver<L>>> S(n);
    std::vector<int> rcs(n, 0);
    // set-up over a few host threads: read, formula table, committed context
    {
        std::atomic<int> next{0};
        std::vector<std::thread> th;
        const int nth = std::max(1, std::min(n, 16));
        for (int t = 0; t < nth; t++)
            th.emplace_back([&] {
                for (int p; (p = next++) < n;) {
                    S[p].reset(new Solver<L>());
                    S[p]->prm = prm;
                    int rc = S[p]->load(R[p]);
                    if (!rc) {
                        S[p]->compute_formula();
                        rc = S[p]->setup_device(device, SDPG_GEMM_PLAIN, 0);
                    }
                    rcs[p] = rc;
                }
            });
        for (auto &t : th) t.join();
    }
    for (int p = 0; p < n; p++)
        if (rcs[p]) return rcs[p];
    std::vector<sdpg_ctx *> ctxs(n);
    for (int p = 0; p < n; p++) ctxs[p] = S[p]->ctx;
    sdpg_batch *b = nullptr;
    int rc = sdpg_batch_create(&b, n, ctxs.data());
    if (rc) return rc;
    sdpg_batch_set_graph(b, use_graph);
    sdpg_batch_wait_contexts(b, 0);  // every device call of the loops goes through the batch
    Meeting mt(b, n, L + 2);
    std::vector<std::unique_ptr<Port>> ports(n);
    for (int p = 0; p < n; p++) {
        ports[p].reset(new Port(&mt, p));
        S[p]->ex = ports[p].get();
    }
    // one host thread per problem: its own IPM control; the device calls meet
    const double t0 = now_s();
    {
        std::vector<std::thread> th;
        th.reserve(n);
        for (int p = 0; p < n; p++)
            th.emplace_back([&, p] {
                rcs[p] = S[p]->run(res ? &res[p] : nullptr, -1);
                if (latency) latency[p] = now_s() - t0;
                mt.leave();
            });
        for (auto &t : th) t.join();
    }
    const double wall = now_s() - t0;
    if (stats) {
        stats[0] = wall;
        stats[1] = (double)mt.meetings;
        stats[2] = mt.meetings ? (double)mt.served / (double)mt.meetings : 0.0;
        double tm[BATCH_TIMERS] = {0};
        sdpg_batch_timers(b, tm, BATCH_TIMERS);
        for (int k = 0; k < BATCH_TIMERS; k++) stats[3 + k] = tm[k];
    }
    sdpg_batch_destroy(b);
    if (mt.rc) return mt.rc;
    for (int p = 0; p < n; p++)
        if (rcs[p]) return rcs[p];
    return 0;
}

}  // namespace host
}  // namespace sdpg

using namespace sdpg::host;

#define DISPATCH_ONE_(LL) \
    if (L_ == LL) return CALL_(LL);
#define DISPATCH_L(L, CALL)               \
    {                                     \
        const int L_ = (L);               \
        SDPG_FOR_EACH_L(DISPATCH_ONE_)    \
        return SDPG_EPREC;                \
    }

extern "C" int sdpa_amd_run(const sdpa_amd_options *opt, sdpa_amd_result *res) {
    if (!opt || !opt->data_path) return -20;
    Parameter prm;
    prm.preset(opt->preset);
    if (opt->param_path && !prm.read(opt->param_path)) return -20;
    if (opt->prec_override > 0) prm.precision = opt->prec_override;
    RawProblem R;
    int rc = read_raw(opt->data_path, R, opt->data_sparse < 0 ? -1 : (opt->data_sparse ? 0 : 1));
    if (rc) return rc;
    RawInit I;
    if (opt->init_path && (rc = read_init(opt->init_path, opt->init_sparse, R.m, R.blockStruct, I))) return rc;
    int L = sdpg_limbs(prm.precision);
#define CALL_(LL) \
    solve_L<LL>(R, prm, opt->log_path, opt->gemm_mode, opt->slices, opt->device, res, opt->init_path ? &I : nullptr)
    DISPATCH_L(L, CALL_)
#undef CALL_
}

extern "C" int sdpa_amd_solve(const char *data_path, const char *param_path, int prec_override, const char *log_path,
                              int gemm_mode, int slices, int device, sdpa_amd_result *res) {
    sdpa_amd_options o;
    memset(&o, 0, sizeof o);
    o.data_path = data_path;
    o.data_sparse = -1;
    o.init_sparse = -1;
    o.param_path = param_path;
    o.prec_override = prec_override;
    o.log_path = log_path;
    o.gemm_mode = gemm_mode;
    o.slices = slices;
    o.device = device;
    return sdpa_amd_run(&o, res);
}

namespace sdpg {
namespace host {
template <int L>
static int ctx_from_file_L(const RawProblem &R, const Parameter &prm, int device, int gemm_mode, sdpg_ctx **out,
                           int *m, long long *nd, int *formula) {
    Solver<L> S;
    S.prm = prm;
    int rc = S.load(R);
    if (rc) return rc;
    S.compute_formula();
    if ((rc = S.setup_device(device, gemm_mode, 0))) return rc;
    *out = S.ctx;
    S.ctx = nullptr;  // ownership passes to the caller
    if (m) *m = S.m;
    if (nd) *nd = S.Nd;
    if (formula) std::copy(S.useFormula.begin(), S.useFormula.end(), formula);
    return 0;
}
}  // namespace host
}  // namespace sdpg

extern "C" int sdpa_amd_ctx_from_file(const char *data_path, int prec_bits, int device, int gemm_mode, sdpg_ctx **out,
                                      int *m, long long *nd, int *formula) {
    Parameter prm;
    prm.precision = prec_bits;
    RawProblem R;
    int rc = read_raw(data_path, R);
    if (rc) return rc;
    int L = sdpg_limbs(prec_bits);
#define CALL_(LL) ctx_from_file_L<LL>(R, prm, device, gemm_mode, out, m, nd, formula)
    DISPATCH_L(L, CALL_)
#undef CALL_
}

extern "C" int sdpa_amd_formula(const char *data_path, int *m, int *n_sdp, int *formula, int cap) {
    RawProblem R;
    int rc = read_raw(data_path, R);
    if (rc) return rc;
    // the table depends only on the problem's structure (stored entries, block sizes), not on
    // the precision: read the values at 256 bits
    Solver<10> S;
    if ((rc = S.load(R))) return rc;
    S.compute_formula();
    if (m) *m = S.m;
    if (n_sdp) *n_sdp = S.nsdp;
    if (formula) {
        if ((int64_t)S.useFormula.size() > (int64_t)cap) return SDPG_EARG;
        std::copy(S.useFormula.begin(), S.useFormula.end(), formula);
    }
    return 0;
}

extern "C" int sdpa_amd_bench(const char *data_path, const char *param_path, int prec_override, int gemm_mode,
                              int slices, int device, int warm_iters, int warmup, int steps, double *ms_schur,
                              double *ms_chol, double *ms_wall, double *kstats, const char *dump_prefix) {
    Parameter prm;
    if (param_path && !prm.read(param_path)) return -20;
    if (prec_override > 0) prm.precision = prec_override;
    RawProblem R;
    int rc = read_raw(data_path, R);
    if (rc) return rc;
    int L = sdpg_limbs(prm.precision);
#define CALL_(LL)                                                                                                 \
    bench_L<LL>(R, prm, gemm_mode, slices, device, warm_iters, warmup, steps, ms_schur, ms_chol, ms_wall, kstats, \
                dump_prefix)
    DISPATCH_L(L, CALL_)
#undef CALL_
}

extern "C" int sdpa_amd_batch_solve(const char *const *data_paths, int n, const char *param_path, int prec_override,
                                    int device, int use_graph, sdpa_amd_result *res, double *latency_s,
                                    double *stats) {
    if (!data_paths || n <= 0) return SDPG_EARG;
    Parameter prm;
    if (param_path && !prm.read(param_path)) return -20;
    if (prec_override > 0) prm.precision = prec_override;
    std::vector<RawProblem> R(n);
    for (int p = 0; p < n; p++) {
        int rc = read_raw(data_paths[p], R[p]);
        if (rc) return rc;
    }
    int L = sdpg_limbs(prm.precision);
#define CALL_(LL) batch_solve_L<LL>(R, prm, device, use_graph, res, latency_s, stats)
    DISPATCH_L(L, CALL_)
#undef CALL_
}
