/*
 * orc — CPU oracle for the SDPA-GMP Newton path.
 *
 * TEST INFRASTRUCTURE ONLY.  Used by tests/, by the smoke check and by the
 * CPU-baseline leg of bench.py.  Product code must not include, link or run
 * anything under oracle/.
 *
 * A plain-C restatement of the reference solver's dense-Schur path
 * (sdpa_main.cpp:43-461 pinpal, sdpa_newton.cpp, sdpa_parts.cpp,
 * sdpa_dataset.cpp, sdpa_linear.cpp, sdpa_jordan.cpp and the mplapack
 * routines they call) on top of ompf (GMP mpf semantics).
 */
#ifndef ORC_H
#define ORC_H
#include <stdint.h>
#include "ompf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* param.sdpa (sdpa_parts.cpp:136-225) */
typedef struct {
    int maxIteration;
    double epsilonStar, lambdaStar, omegaStar, lowerBound, upperBound;
    double betaStar, betaBar, gammaStar, epsilonDash;
    int precision;
} orc_param;

void orc_param_default(orc_param *p); /* sdpa_parts.cpp:165-178 */
int orc_param_read(orc_param *p, const char *path);

typedef struct {
    int iterations;
    int phase; /* orc_phase value after the loop (before reversal) */
    char phase_name[16]; /* as printed (after REVERSE_PRIMAL_DUAL) */
    double mu, relgap, gap;
    char objP[80], objD[80]; /* printed objValPrimal / objValDual (40 digits) */
    char relgap_s[80], pfeas_s[80], dfeas_s[80];
    double t_bmat, t_chol, t_total; /* seconds: Make bMat, Cholesky bMat, main loop */
    char objP70[112], objD70[112];  /* the same objectives with 70 significant digits */
} orc_result;

/*
 * Solve an SDPA sparse-format problem.  prec_override > 0 replaces the
 * parameter file's precision.  log_path (may be NULL) receives the
 * per-iteration log in the printOneIteration format.
 * dump_dir (may be NULL): for each iteration k listed in dump_iters
 * (n_dump entries), the predictor's Newton inputs and outputs are written
 * there as raw interchange files (see orc.c, dump_newton).
 */
int orc_solve(const char *data_path, const orc_param *param, int prec_override, const char *log_path,
              const char *dump_dir, const int *dump_iters, int n_dump, int dump_L, orc_result *res);

/*
 * One predictor Newton step on a dumped point (X, Z, invZ, y, mu, beta,
 * phase) at prec_bits; writes Rp, Rd, g, B, L, dy, dX, dZ (and the inputs) to
 * out_dir.  Returns 0, 1 on Cholesky failure, < 0 on I/O errors.
 */
int orc_newton_step(const char *data_path, int prec_bits, const char *in_dir, const char *out_dir, int out_L);

/*
 * Time the reference hot path (Make_bMat + Cholesky of B) on a problem at
 * the default starting point X = Z = lambda*I advanced `warm_iters` IPM
 * iterations; runs `reps` repetitions and returns seconds per repetition of
 * each phase.  Used as the CPU baseline.
 */
int orc_time_hotpath(const char *data_path, const orc_param *param, int prec_override, int warm_iters, int reps,
                     int threads, double *t_bmat, double *t_chol);
/*
 * Time Make_bMat on an iterate (X, invZ) given as packed files of L-limb
 * words, then the Cholesky of the leading chol_n x chol_n block of B
 * (chol_n <= 0: all of it): a bounded CPU sample of a large workload.
 */
int orc_time_point(const char *data_path, int prec_bits, const char *x_path, const char *invz_path, int L,
                   int chol_n, int threads, double *t_bmat, double *t_chol);

/*
 * Bounded CPU sample of Make_bMat for an all-F1 block 0: seconds per constraint's F1 precompute
 * (OpenMP Rgemm NN with `threads` threads) and per pair's inner product (serial), on nsample
 * constraints x nsample partners of the iterate (X, invZ) in packed files of L limbs.
 */
int orc_time_f1_sample(const char *data_path, int prec_bits, const char *x_path, const char *invz_path, int L,
                       int nsample, int threads, double *t_per_con, double *t_per_pair);

/* arithmetic probe for tests: op codes 0 add,1 sub,2 mul,3 div,4 sqrt,5 set_str(a as text) */
int orc_arith(int prec_bits, int op, int L, int32_t as, int32_t ae, const uint32_t *al, int32_t bs, int32_t be,
              const uint32_t *bl, int32_t *rs, int32_t *re, uint32_t *rl);
int orc_parse(int prec_bits, const char *text, int L, int32_t *rs, int32_t *re, uint32_t *rl);
int orc_format(int prec_bits, int L, int32_t s, int32_t e, const uint32_t *l, int digits, char *buf, int bufsz);

/* formula table of Newton::computeFormula_SDP: out[i * nsdp + l] in {1,2,3}, 0 = no block */
int orc_formula(const char *data_path, int prec_bits, int *m, int *nsdp, int *out, int cap);
/* Rpotrf("Lower") of a packed n x n matrix (engine exchange format); *info = first failing pivot */
int orc_potrf(int prec_bits, int n, int L, const uint32_t *a, uint32_t *lo, int *info);

#ifdef __cplusplus
}
#endif
#endif
