/*
 * sdpa_amd.h — host driver of the MI355X SDPA-GMP path.
 *
 * sdpa_amd_solve restates the reference driver pinpal() (sdpa_main.cpp:43-461):
 * it reads an SDPA sparse-format problem and a param.sdpa file, runs the
 * Mehrotra predictor-corrector interior-point loop with the Newton system on
 * the GPU (sdpg.h) and the step length / update / stopping tests on the host,
 * and writes the reference's iteration log and final summary.
 */
#ifndef SDPA_AMD_H
#define SDPA_AMD_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int iterations;
    int status;           /* 0 loop ended normally, 1 Cholesky failure, 2 cannot move */
    char phase[16];       /* as printed (phase.value after REVERSE_PRIMAL_DUAL) */
    char objP[96];        /* printed objValPrimal, 41 significant digits */
    char objD[96];        /* printed objValDual */
    char relgap[48];      /* relative gap, 17 digits */
    char pfeas[48], dfeas[48];
    double mu;
    double t_schur_ms;    /* Make bMat, device time, all iterations */
    double t_chol_ms;     /* Cholesky bMat, device time */
    double t_solve_ms, t_gvec_ms, t_dxdz_ms, t_res_ms;
    double t_total_s;     /* wall clock of the main loop */
    int limbs;
    double t_inv_ms;      /* Solutions::computeInverse (sdpg_factor_point), device time */
    double t_eig_ms;      /* step-length minimum eigenvalues (sdpg_min_eigen), device time */
    char objP70[112];     /* objValPrimal with 70 significant digits */
    char objD70[112];
    int slices;           /* Ozaki slice count used (autotuned when the caller passes 0), 0 for other forms */
} sdpa_amd_result;

/*
 * Options of one run, the command line of sdpa_main.cpp:501-677 (-ds / -dd data, -is / -id initial point,
 * -p parameter file, -pt preset, -o output).
 */
typedef struct {
    const char *data_path;  /* problem file */
    int data_sparse;        /* 1 sparse (-ds), 0 dense (-dd), -1 by the name (a name ending in "-s" is sparse) */
    const char *init_path;  /* initial point (sdpa_io.cpp:150-274); NULL: X = Z = lambda* I, y = 0 */
    int init_sparse;        /* 1 (-is), 0 (-id), -1 by the name */
    const char *param_path; /* param.sdpa; NULL: the preset */
    int preset;             /* Parameter::setDefaultParameter: 0 default, 1 unstable but fast, 2 stable but slow */
    int prec_override;      /* > 0 replaces the precision */
    const char *log_path;   /* NULL, "-" (stdout) or a file: iteration log, summary, parameters and solution */
    int gemm_mode, slices, device;
} sdpa_amd_options;

/* pinpal (sdpa_main.cpp:43-461) with the options above; returns 0, < 0 on input or device errors */
int sdpa_amd_run(const sdpa_amd_options *opt, sdpa_amd_result *res);

/*
 * param_path may be NULL (defaults of Parameter::setDefaultParameter,
 * sdpa_parts.cpp:165-178).  prec_override > 0 replaces the precision line.
 * log_path: NULL, "-" (stdout) or a file.  The data file's format follows its name (sdpa_amd_run
 * chooses explicitly).  gemm_mode: SDPG_GEMM_PLAIN or
 * SDPG_GEMM_OZAKI; slices 0 with the Ozaki form: the count sdpg_tune_slices picks for the run's
 * precision.  Returns 0 on success, < 0 on input or device errors.
 */
int sdpa_amd_solve(const char *data_path, const char *param_path, int prec_override, const char *log_path,
                   int gemm_mode, int slices, int device, sdpa_amd_result *res);

/*
 * Replay benchmark of the hot path on one iterate of the solve: runs the solver
 * for `warm_iters` iterations, then times `steps` executions of
 * [Make_bMat + Cholesky] (sdpg_schur + sdpg_potrf) on that iterate after
 * `warmup` untimed ones, with nothing but the step's kernels in the stream.
 * Returns per-step device milliseconds of each phase and the wall time.
 * kstats (may be NULL, else 12 doubles): sdpg_profile figures over a second
 * pass of `steps` steps with events around the product and block-column
 * kernels.  dump_prefix (may be NULL): the iterate is written to
 * <dump_prefix>X.bin and <dump_prefix>invZ.bin (packed mp, raw uint32) so that a
 * CPU implementation can time the same step on the same data.
 */
int sdpa_amd_bench(const char *data_path, const char *param_path, int prec_override, int gemm_mode, int slices,
                   int device, int warm_iters, int warmup, int steps, double *ms_schur, double *ms_chol,
                   double *ms_wall, double *kstats, const char *dump_prefix);

/*
 * The batched form: n problems solved in lockstep (no single reference counterpart; each problem follows
 * pinpal, sdpa_main.cpp:43-461, as sdpa_amd_solve does).  Every problem runs the IPM loop on its own host
 * thread with its own context; every device call of its loop (the Newton step, dZ and dX, the inner
 * products, the step-length eigenvalues, the update, the inverse factors, the residuals) meets those of
 * the other running problems in one batched call per kind (sdpg_batch_*), and a problem whose loop ends
 * leaves the meetings.  The problems must be dense-constraint SDPs without LP
 * blocks (sdpg_batch_create).  use_graph: the batched phases as a HIP graph.  res (may be NULL, else
 * n results), latency_s (may be NULL, else n): seconds from the start of the loops to each problem's end.
 * stats (may be NULL, else 12 doubles): wall seconds of the loops, meetings (rounds of batched calls),
 * mean problems per meeting, then the device milliseconds of the batched calls by kind as
 * sdpg_batch_timers reports them (9 values).  Returns 0, or the first error.
 */
int sdpa_amd_batch_solve(const char *const *data_paths, int n, const char *param_path, int prec_override,
                         int device, int use_graph, sdpa_amd_result *res, double *latency_s, double *stats);

/*
 * Read a problem, compute its formula table (Newton::computeFormula_SDP,
 * sdpa_newton.cpp:694-819) and return a committed engine context
 * (sdpa_main.cpp:197-225 set-up).  *m and *nd receive the number of
 * constraints and the size of a dense linear space; formula (may be NULL,
 * else m*n_sdp ints) receives the table.  The caller destroys the context.
 */
struct sdpg_ctx;
int sdpa_amd_ctx_from_file(const char *data_path, int prec_bits, int device, int gemm_mode, struct sdpg_ctx **out,
                           int *m, long long *nd, int *formula);

/*
 * Newton::computeFormula_SDP (sdpa_newton.cpp:694-819) of a problem file, on the host
 * only (no device): *m, *n_sdp and formula[i * n_sdp + l] in {1 = F1, 2 = F2, 3 = F3},
 * 0 where A_i has no block l (cap = capacity of formula in ints).  This is the table
 * that sdpa_amd_solve and sdpa_amd_ctx_from_file upload (sdpg_set_formula).
 */
int sdpa_amd_formula(const char *data_path, int *m, int *n_sdp, int *formula, int cap);

#ifdef __cplusplus
}
#endif
#endif
