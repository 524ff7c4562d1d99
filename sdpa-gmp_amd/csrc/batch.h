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

constexpr int BMAXB = 4;  // SDP blocks per problem

struct BatchProb {
    int32_t m, nsdp;
    int32_t n[BMAXB];
    int64_t blkoff[BMAXB];
    const Desc *desc;  // m x nsdp constraint descriptors (index i * nsdp + l)
    PV X, invZ, dval, B, g, dy;
    int64_t cs_off;  // 2m values of the batch's scratch: r_j = 1/sqrt(c_j d_j), then 1/L_jj
};

// request flags: 0 nothing, 1 PREDICTOR (Schur build, Cholesky, solve), 2 CORRECTOR (solve only, with the
// factor of the last predictor; Newton::Mehrotra, sdpa_newton.cpp:1375-1414)
enum { BATCH_NONE = 0, BATCH_PREDICTOR = 1, BATCH_CORRECTOR = 2 };

// Make_bMat for one constraint i of one problem per task (compute_bMat_dense_SDP with formula F1,
// sdpa_newton.cpp:1007-1096): for every block l of constraint i, work1 = A_i inv(Z) and work2 = X work1
// (sdpa_linear.cpp:856-955) in LDS, then B_ij += A_j . work2 for the partners j <= i of the block (the pair
// rule for equal NonZeroEffect, sdpa_newton.cpp:1052-1054); B_ij and B_ji are written once, after the
// blocks.  Element form: one rounded product and one rounded sum per term.  Tasks are dealt round-robin to
// the workgroups of a grid that fills the device.
// LDS: work1 and work2 as planar n x n arrays (plane stride PW = max(nmax^2, BS_THREADS)), then the row
// of B (mmax values); the partial sums of the inner products reuse work1's planes.
constexpr int BS_THREADS = 512;
template <int L>
__global__ __launch_bounds__(BS_THREADS) void k_batch_schur(const BatchProb *__restrict__ probs,
                                                            const int2 *__restrict__ tasks, int ntasks,
                                                            const int32_t *__restrict__ req, int nmax, int mmax) {
    extern __shared__ uint32_t shm[];
    const int PW = max(nmax * nmax, BS_THREADS);
    uint32_t *sF = shm, *sG = shm + (L + 1) * PW, *sB = shm + 2 * (L + 1) * PW, *sP = sF;
    const int tid = threadIdx.x, nt = BS_THREADS;
    // inner products: BS_THREADS / 8 partners at a time, 8 threads each
    const int q = tid & 7, slot = tid >> 3, nslot = BS_THREADS / 8;
    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const int2 tk = tasks[t];
        if (req[tk.x] != BATCH_PREDICTOR) continue;  // the whole workgroup skips together
        const BatchProb &P = probs[tk.x];
        const int i = tk.y, m = P.m, nsdp = P.nsdp;
        for (int j = tid; j <= i; j += nt) lds_st<L>(sB, mmax, j, zero<L>());
        for (int l = 0; l < nsdp; l++) {
            const Desc di = P.desc[(int64_t)i * nsdp + l];
            if (di.kind == 0) continue;
            const int n = P.n[l], nn = n * n;
            const int64_t bo = P.blkoff[l];
            __syncthreads();
            // work1 = A_i inv(Z)
            for (int o = tid; o < nn; o += nt) {
                const int a = o % n, b = o / n;
                mp<L> acc = zero<L>();
                for (int c = 0; c < n; c++)
                    acc = add<L>(acc, mul<L>(ldm<L>(P.dval, di.off + a + (int64_t)n * c),
                                             ldm<L>(P.invZ, bo + c + (int64_t)n * b)));
                lds_st<L>(sF, PW, o, acc);
            }
            __syncthreads();
            // work2 = X work1
            for (int o = tid; o < nn; o += nt) {
                const int a = o % n, b = o / n;
                mp<L> acc = zero<L>();
                for (int c = 0; c < n; c++)
                    acc = add<L>(acc, mul<L>(ldm<L>(P.X, bo + a + (int64_t)n * c), lds_ld<L>(sF, PW, c + n * b)));
                lds_st<L>(sG, PW, o, acc);
            }
            __syncthreads();
            // B_ij += A_j . work2 (Rdot over n^2, sdpa_linear.cpp:334-337) for the partners j <= i
            for (int j0 = 0; j0 <= i; j0 += nslot) {
                const int j = j0 + slot;
                mp<L> part = zero<L>();
                if (j <= i) {
                    const Desc dj = P.desc[(int64_t)j * nsdp + l];
                    if (dj.kind != 0)
                        for (int e = q; e < nn; e += 8)
                            part = add<L>(part, mul<L>(ldm<L>(P.dval, dj.off + e), lds_ld<L>(sG, PW, e)));
                }
                lds_st<L>(sP, PW, tid, part);
                __syncthreads();
                if (q == 0 && j <= i) {
                    mp<L> s = lds_ld<L>(sP, PW, tid);
                    for (int u = 1; u < 8; u++) s = add<L>(s, lds_ld<L>(sP, PW, tid + u));
                    lds_st<L>(sB, mmax, j, add<L>(lds_ld<L>(sB, mmax, j), s));
                }
                __syncthreads();
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

// Rpotrf("Lower") of one problem's Schur matrix per workgroup (choleskyFactorWithAdjust,
// sdpa_linear.cpp:608-631), unblocked and division-free as the diagonal blocks of the single-problem
// engine: with c_j > 0 a running scale, At = c_j S^(j) and step j updates every remaining entry of the lower
// triangle at once, At_ic <- 2^-e_j (d_j At_ic - At_ij At_cj) (d_j = At_jj, e_j its exponent); the true
// pivot is d_j / c_j and c_{j+1} = 2^-e_j c_j d_j.  After the chain, r_j = 1/sqrt(c_j d_j) gives
// L_ic = At_ic r_c (so L_jj = d_j r_j) and 1/L_jj = c_j r_j.  info = the first non-positive pivot
// (1-based, as Rpotrf2's info), and the problem's solve is skipped.
template <int L>
__global__ __launch_bounds__(256) void k_batch_potrf(const BatchProb *__restrict__ probs, int nprob,
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
        // running scales (sequential, one product per pivot), then r_j and 1/L_jj in parallel
        if (tid == 0) {
            mp<L> c = from_int<L>(1);
            for (int j = 0; j < m; j++) {
                const mp<L> d = ldm<L>(B, at(j, j));
                stm<L>(cs, P.cs_off + j, c);
                c = mul<L>(c, d);
                c.e -= d.e;
            }
        }
        __threadfence_block();
        __syncthreads();
        for (int j = tid; j < m; j += nt) {
            const mp<L> c = ldm<L>(cs, P.cs_off + j), d = ldm<L>(B, at(j, j));
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

struct BatchBase {
    virtual ~BatchBase() {}
    virtual int newton(const int *req, int *info) = 0;
    virtual int set_graph(int on) = 0;
    virtual int timing(double *ms) = 0;
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

    ~Batch() override {
        if (st) (void)hipStreamSynchronize(st);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        (void)hipFree(d_probs);
        (void)hipFree(d_tasks);
        (void)hipFree(d_req);
        (void)hipFree(d_info);
        (void)hipFree(cs);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    }

    int create(EngineBase *const *es, int n) {
        if (n <= 0) return SDPG_EARG;
        device = es[0]->device;
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
            mmax = std::max(mmax, e->m);
            b.desc = e->d_desc;
            b.X = e->X.pv();
            b.invZ = e->invZ.pv();
            b.dval = e->dval.pv();
            b.B = e->Bm.pv();
            b.g = e->g.pv();
            b.dy = e->dy.pv();
            b.cs_off = csn;
            csn += 2 * (int64_t)e->m;
            hp.push_back(b);
            for (int i = 0; i < e->m; i++) tasks.push_back(int2{p, i});
            eng.push_back(e);
        }
        nprob = n;
        // largest tasks first (the constraints of the problems with the largest blocks), dealt round-robin
        std::stable_sort(tasks.begin(), tasks.end(), [&](const int2 &a, const int2 &b) {
            auto cost = [&](const int2 &t) {
                int64_t c = 0;
                for (int l = 0; l < hp[t.x].nsdp; l++) c += (int64_t)hp[t.x].n[l] * hp[t.x].n[l] * (hp[t.x].n[l] + t.y);
                return c;
            };
            return cost(a) > cost(b);
        });
        ntasks = (int)tasks.size();
        size_t lds_max = 65536;
        {
            int v = 0;
            if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
                lds_max = (size_t)v;
        }
        shm_schur = (size_t)(L + 1) * 4 * (2 * (size_t)std::max(nmax * nmax, BS_THREADS) + mmax);
        shm_potrs = (size_t)(L + 1) * 4 * mmax;
        if (shm_schur > lds_max) {
            set_err("sdpg_batch_create: blocks too large for the batched form at this precision");
            return SDPG_EARG;
        }
        int cus = 0;
        if (hipSetDevice(device) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
            return SDPG_EHIP;
        grid_schur = std::max(1, std::min(ntasks, cus * std::max(1, (int)(lds_max / shm_schur))));
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
        if (hipFuncSetAttribute((const void *)k_batch_schur<L>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shm_schur) != hipSuccess ||
            hipMemcpy(d_probs, hp.data(), sizeof(BatchProb) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_tasks, tasks.data(), sizeof(int2) * ntasks, hipMemcpyHostToDevice) != hipSuccess)
            return SDPG_EHIP;
        return SDPG_OK;
    }

    // the three phases on the batch's stream (what the graph holds)
    void launch() {
        k_batch_schur<L><<<grid_schur, BS_THREADS, shm_schur, st>>>(d_probs, d_tasks, ntasks, d_req, nmax, mmax);
        k_batch_potrf<L><<<grid_prob, 256, 0, st>>>(d_probs, nprob, d_req, d_info, PV{cs, cs_n});
        k_batch_potrs<L><<<grid_prob, 256, shm_potrs, st>>>(d_probs, nprob, d_req, d_info, PV{cs, cs_n}, mmax);
    }

    int newton(const int *req, int *info) override {
        if (hipSetDevice(device) != hipSuccess) return SDPG_EHIP;
        // the contexts' own work (point updates, right-hand sides) is complete before the batch reads it
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
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) last_ms = ms;
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
