/*
 * fednl_hip.h -- C ABI of the MI355X FedNL round engine (libfednl_hip.so).
 *
 * Plain pointers and sizes only. Host arrays are copied in/out by the
 * library; all per-round state stays resident in device memory.
 *
 * Index conventions follow dopt (burlachenkok/ufednl):
 *   - flat index of (i, j) in a d x d matrix is j*LDA + i with
 *     LDA = roundUp(d, 4)  (dopt/dopt/linalg_matrices/include/MatrixNMD.h:44-81, 997-1000);
 *   - the upper triangle (i <= j) is enumerated column by column; the
 *     position of (i, j) in that list is j*(j+1)/2 + i
 *     (dopt/dopt/math_routines/include/SpecialMathRoutinesForMatrix.h:35-64).
 *     "Packed upper" arrays hold S = d*(d+1)/2 values in that order.
 */
#ifndef FEDNL_HIP_H
#define FEDNL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compressor codes, as dopt's enum Compressor
 * (dopt/bin_fednl_local/include/UsedTypes.h:17-25). */
enum fednl_compressor {
    FEDNL_IDENTICAL = 0,
    FEDNL_RANDK = 1,
    FEDNL_TOPK = 2,
    FEDNL_RANDSEQK = 3,
    FEDNL_TOPLEK = 4,
    FEDNL_NATURAL = 5
};

/* Return codes. Non-negative values are success; the negative codes mirror
 * dopt's TrainReturnCodes (dopt/bin_fednl_local/include/Args.h:53-64) and
 * add device-side failures below -8. */
enum fednl_status {
    FEDNL_OK = 0,
    FEDNL_E_MISSING_ARGUMENT = -1,
    FEDNL_E_WRONG_ARGUMENT = -2,
    FEDNL_E_DATASET_NOT_AVAILABLE = -3,
    FEDNL_E_DATASET_PARSE = -4,
    FEDNL_E_SPLIT = -5,
    FEDNL_E_SAVE = -6,
    FEDNL_E_INTERNAL = -7,
    FEDNL_E_UNKNOWN_DEBUG_MODE = -8,
    FEDNL_E_DEVICE = -20,       /* HIP runtime error (see fednl_last_error) */
    FEDNL_E_UNSUPPORTED = -21   /* size or mode this build does not handle */
};

/* Options of one FedNL problem. Fill with fednl_default_opts(). */
typedef struct fednl_opts {
    double lambda;        /* L2 regulariser (--lambda) */
    int compressor;       /* enum fednl_compressor (--compressor) */
    double k_mult;        /* k = (int)(k_mult * d) (--k_compressor_as_d_mult, TrainLoop.h:362) */
    int alpha_mode;       /* 0: explicit alpha; 1/2: theoretical option 1/2 (TrainLoop.h:409-518) */
    double alpha;         /* used when alpha_mode == 0 (--alpha_step_size) */
    int option_b;         /* 1: FedNL option b (shift = mean Lk); 0: option a (shift = lambda) */
    int workers;          /* client worker count of the reference run (--workers); selects
                             the TopLEK random stream of each client (TrainLoop.h:664-686) */
    int tracking;         /* 1: compute f(x_r) every round (--tracking) */
    int hessian_mode;     /* A^T D A contraction: 0 FP64 MFMA (default); 1 FP32 MFMA;
                             2 BF16 hi/lo split with FP32 accumulation. Gradient, Delta,
                             Lk, aggregation and the Newton solve stay FP64. */
    int line_search;      /* 1: FedNL-LS, Armijo backtracking on the Newton step from t = 1
                             (--line_search, TrainLoop.h:344-355, 1546-1640) */
    int max_line_search_steps;  /* trial points per round (--max_line_search_steps, default 5) */
    double c_line_search;       /* sufficient-decrease constant c (--c_line_search, default 0.5) */
    double gamma_line_search;   /* step shrink factor in (0, 1) (--gamma_line_search, default 0.5) */
    int clients_per_round;      /* 0: FedNL; tau in [1, n]: FedNL-PP with tau clients drawn per round
                                   (bin_fednl_local_pp, --clients-per-round, TrainLoop.h:279-287) */
    unsigned client_selection_seed;  /* FedNL-PP client sampler seed (default 1234, Args.h:244) */
    int algorithm;              /* enum fednl_algorithm (--algorithm, TrainLoop.h:297-299) */
    int hessian_form;           /* enum fednl_hessian_form: how the FP64 A^T D A is computed */
    double global_lr;           /* GD step size: --global_lr, or 2 / (mu + L) with
                                   --theoretical_global_lr, times --global_step_size_multiplier
                                   (TrainLoop.h:358,403-407,521) */
} fednl_opts;

/* FEDNL_ALG_FEDNL: FedNL / FedNL-LS / FedNL-PP (--algorithm fednl1);
 * FEDNL_ALG_GD: gradient descent on the clients' averaged gradient,
 * x -= global_lr * (sum_c grad f_c(x)) / n (--algorithm gd; LocalWorkerForGD.h,
 * TrainLoop.h:993-1052). GD uses neither the compressor nor the Hessians and
 * runs without line search and outside batches. */
enum fednl_algorithm { FEDNL_ALG_FEDNL = 0, FEDNL_ALG_GD = 1 };

/* FEDNL_HESS_DENSE: every client's A^T D A on FP64 MFMA over the padded
 * design. FEDNL_HESS_SPARSE (FP64 mode, binary designs: every feature 0 or
 * 1): each upper entry H_c[i,j] is the sum of D_s over the samples holding
 * both features, in ascending sample order; with 0/1 features and +-1 labels
 * every nonzero term of the reference's sum (logistic_regression.h:365-547)
 * is exactly D_s, so this is the bits of its sequential sum, at a cost that
 * follows the number of co-occurring feature pairs instead of m d^2.
 * FEDNL_HESS_AUTO (0, the default) takes the sparse form in FP64 mode when
 * the design is binary, those pairs number at most a tenth of m d(d+1)/2
 * and the per-segment lists fit in LDS; FEDNL_HESS_FORM=dense|sparse in the
 * environment sets it instead. No counterpart in the reference. */
enum fednl_hessian_form { FEDNL_HESS_AUTO = 0, FEDNL_HESS_DENSE = 1, FEDNL_HESS_SPARSE = 2 };

void fednl_default_opts(fednl_opts *o);

/* Last error message of the calling thread ("" if none). */
const char *fednl_last_error(void);

/* Creates a problem: n clients, m samples each, d features (the intercept,
 * if wanted, is already a column of `features`). features: n*m*d doubles
 * (client-major, then sample-major, row of d features); labels: n*m values
 * in {-1, +1}. Client c owns samples [c*m, (c+1)*m) as in dopt's split
 * without reshuffling (TrainLoop.h:212-250). Computes the initial learned
 * Hessians H_c(x0 = 0) and their mean on the device (TrainLoop.h:603-643,732-735).
 * Replaces the per-client setup of train() (TrainLoop.h:600-700). */
int fednl_create(const fednl_opts *opts, int n, int m, int d, const double *features,
                 const double *labels, void **handle);

void fednl_destroy(void *handle);

/* Sets the GD step size of an engine made with FEDNL_ALG_GD, e.g. the
 * theoretical 2 / (mu + L) once L is known (TrainLoop.h:306-316 computes L
 * from the clients' data, :403-407,521 the step); takes effect from the next
 * enqueued round. */
int fednl_set_global_lr(void *handle, double global_lr);

/* Sizes: k actually used, S = d(d+1)/2, LDA of flat indices. */
int fednl_sizes(void *handle, int *k, int64_t *S, int *lda);

/* Alpha actually used (explicit or from the theory formulas). */
double fednl_alpha(void *handle);

/* Enqueues `rounds` FedNL rounds on the engine's stream and waits for them.
 * One round = dopt's client loop (LocalWorkerForFedNL_*.h) for every client
 * plus the master update (TrainLoop.h:1054-1640). A round whose Cholesky
 * factorization fails solves the system by Gauss-Jordan elimination instead,
 * as dopt's master does (TrainLoop.h:1410-1420, 1476-1486), and records the
 * failure in the trace (chol_ok = 0). */
int fednl_run(void *handle, int rounds);

/* Enqueues rounds without waiting (graph replay when enabled). */
int fednl_enqueue(void *handle, int rounds);
int fednl_sync(void *handle);

/* Per-round trace since creation: ||grad f(x_r)||, f(x_r) (0 without
 * tracking or line search), ||x_r||, seconds since the first round was
 * enqueued (from device events), factorization flag (1 ok, 0: the round used
 * the Gauss-Jordan fallback). Arrays sized >= *count. */
int fednl_trace(void *handle, double *grad_norm, double *fval, double *x_norm, double *seconds,
                int *chol_ok, int *count);

/* FedNL-LS: trial points evaluated in each round since creation (1..max;
 * 0 when line search is off). Arrays sized >= *count. */
int fednl_trace_line_search(void *handle, int *trials, int *count);

/* Per round: clients that took part (n, or tau for FedNL-PP) and the Hessian
 * items they sent (k each; TopLEK: its per-client counts). The inputs of
 * dopt's received-bytes counters (TrainLoop.h:1104-1287). Arrays sized >= *count. */
int fednl_trace_clients(void *handle, int *participants, double *items, int *count);

/* FedNL-PP: the clients selected in the last round (n flags 0/1), the first
 * tau entries of dopt::shuffle of 0..n-1 (TrainLoop.h:1155-1157). */
int fednl_pp_selection(void *handle, int *selected);

int fednl_get_x(void *handle, double *x);

/* Outputs of the last round, for parity checks: averaged gradient (d; for
 * FedNL-PP the full gradient at x_r),
 * per-client Lk (n), selected counts (n), selected list positions (n*k,
 * ascending), values sent to the master (n*k), flat indices (n*k), Newton
 * step (d). Any pointer may be NULL. */
int fednl_last_round(void *handle, double *grad, double *lk, int *cnt, uint32_t *pos, double *vals,
                     uint32_t *flat_idx, double *dx);

/* The last round's Hessian differences Delta_c = H_c(x_r) - learned_c as the
 * client pass wrote them (n*S packed upper), the input of every compressor
 * (LocalWorkerForFedNL_TopK.h:70-71). For parity checks: the selection,
 * message values and learned update of a round can be recomputed from it. */
int fednl_last_delta(void *handle, double *delta);

/* Whole state for lockstep checks: x (d), learned Hessians (n*S packed),
 * master Hessian estimate H (S packed), RandK/RandSeqK client generator
 * states (n), TopLEK stream states (workers). Any pointer may be NULL. */
int fednl_get_state(void *handle, double *x, double *learned, double *H, uint32_t *rng,
                    uint32_t *toplek_rng);
int fednl_set_state(void *handle, const double *x, const double *learned, const double *H,
                    const uint32_t *rng, const uint32_t *toplek_rng);

/* The tile shapes the engine picked when it was created, by timing each
 * candidate on its own data: rows per wave of the margin pass (4 or 8; the
 * environment variable FEDNL_MARGIN_ROWS=4|8 sets it instead) and, in the
 * FP64 Hessian mode, the A^T D A tile's sample depth per LDS stage
 * (16 or 32; FEDNL_HESS_CHUNK=16|32; hessian_chunk may be NULL). The
 * candidates of each compute the same bits. No counterpart in the reference. */
int fednl_tuning(void *handle, int *margin_rows, int *hessian_chunk);

/* The Hessian form the engine runs (FEDNL_HESS_DENSE or FEDNL_HESS_SPARSE),
 * or a negative status. No counterpart in the reference. */
int fednl_hessian_form(void *handle);

/* Per-kernel device time (milliseconds, summed over rounds since the last
 * reset) measured with HIP events on the engine stream when enabled. */
int fednl_kernel_timing(void *handle, int enable);
/* The same for a subset: bit q times the kernel fednl_kernel_times lists at
 * index q (the events around each timed launch cost device time). */
int fednl_kernel_timing_mask(void *handle, unsigned mask);
int fednl_kernel_times(void *handle, double *ms, int *launches, int max_kernels, const char **names);

/* Standalone compressor on identical inputs (parity of index selection).
 * delta: nclients*S packed upper values (host). For TopK/TopLEK the
 * selection is dopt's heap scan (SpecialMathRoutinesForMatrix.h:229-511);
 * RandK/RandSeqK draw from rng[c] (one RandomGenIntegerLinear per client,
 * updated in place); TopLEK draws from toplek_rng[c % workers] (updated in
 * place), with delta_param = k/S as the TopLEK worker passes it
 * (LocalWorkerForFedNL_TopLEK.h:19,58). out_pos: nclients*k list positions,
 * ascending (RandSeqK: ascending too); out_cnt: counts. TopK, TopLEK and
 * RandK take k <= 32768 (above 8192, or for S above 2^20, the heap and the
 * selection bitmap live in global memory); FEDNL_E_UNSUPPORTED beyond. */
int fednl_select(int compressor, int nclients, int d, int k, const double *delta, uint32_t *rng,
                 uint32_t *toplek_rng, int workers, double delta_param, uint32_t *out_pos,
                 int *out_cnt);

/* Standalone Newton solve: Cholesky of (H_up + shift*I) and the two
 * triangular solves (TrainLoop.h:1449-1466, CholeskyFactorization.h:153-272,
 * ElementarySolvers.h:72-226). H_up: S packed upper. Returns 1 when the
 * factorization succeeded, 0 otherwise, negative on error. */
int fednl_newton_solve(int d, const double *H_up, double shift, const double *b, double *dx);

/* No counterpart in the reference: fednl_newton_solve for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). H_up (S packed upper), b and dx (d
 * doubles) and ok (one int) are DEVICE pointers; the solve is enqueued on
 * `stream` (a hipStream_t; NULL for the null stream) and the call returns
 * without waiting. The same kernels as fednl_newton_solve give the same dx;
 * *ok becomes 1 when the factorization succeeded, else 0 (dx = 0). The call
 * allocates nothing: its workspace (descriptor, shift, status and the
 * 512 x 512 Cholesky workspace with its block-column handoff buffers,
 * 2 MiB + 138 KiB) is a device array of the
 * library that each GPU holds until the process ends. d <= 512, larger
 * returns FEDNL_E_UNSUPPORTED. Calls on one GPU share that array: calls that
 * overlap on one GPU (other streams, other threads) race. */
int fednl_newton_solve_noalloc(int d, const double *H_up, double shift, const double *b, double *dx, int *ok,
                               void *stream);

/* Diagnostic: raw outputs of three v_mfma_f64_16x16x4f64 probes (768 doubles)
 * used by the tests to pin the lane layout the Hessian kernel relies on.
 * No counterpart in the reference. */
int fednl_probe_mfma_f64(double *out768);

/* Batches: independent problems stepped together, each round kernel launched
 * once for all members (one workgroup grid over a member/block map), the
 * round optionally captured once into a HIP graph (use_graph) and replayed.
 * Members keep their own state, trace and results (fednl_trace, fednl_get_x
 * ... on each engine); while in a batch an engine must not be stepped on its
 * own. fednl_batch_sync / fednl_batch_run return after all enqueued rounds.
 * No single counterpart in the reference: dopt runs one problem per process
 * (TrainLoop.h:103), this replaces running many bin_fednl_local instances. */
int fednl_batch_create(void *const *engines, int count, int use_graph, void **batch);
void fednl_batch_destroy(void *batch);
int fednl_batch_enqueue(void *batch, int rounds);
int fednl_batch_sync(void *batch);
int fednl_batch_run(void *batch, int rounds);

/* Serving: solves that arrive over time, each a whole FedNL problem run to
 * ||grad f||^2 <= tol (or max_rounds). fednl_serve runs them in real time:
 * a request is admitted at or after arrival_s (seconds from the call), in
 * arrival order, while the device bytes of the admitted problems stay within
 * hbm_budget_bytes (fednl_footprint, plus, for a binary design in the sparse
 * Hessian form, its sample lists once read: the budget can be passed by one
 * request's lists); admitted problems are stepped together
 * as one batch, rounds_per_step rounds at a time, members joining and leaving
 * between steps (continuous batching). Results per request, in input order;
 * x_out (d doubles, optional) receives the final iterate. Returns FEDNL_OK or
 * the first device error. No counterpart in the reference. */
typedef struct fednl_request {
    double arrival_s;
    int n, m, d;
    const double *features;  /* n*m*d, as for fednl_create */
    const double *labels;    /* n*m */
    fednl_opts opts;
    int max_rounds;
    double tol;
    double *x_out;
} fednl_request;

typedef struct fednl_result {
    int status;         /* 0 converged, 2 round budget spent, < 0 error, 1 not run */
    int rounds;         /* rounds run */
    int rounds_to_tol;  /* first round with ||grad f||^2 <= tol, -1 if none */
    double admitted_s, finished_s, grad_norm;
} fednl_result;

size_t fednl_footprint(const fednl_opts *opts, int n, int m, int d);
int fednl_serve(const fednl_request *reqs, int count, double hbm_budget_bytes, int rounds_per_step, int use_graph,
                fednl_result *out);

/* Client Hessians H_c(x) for all clients (n*S packed upper, host), computed
 * by the engine's Hessian mode at the given x (d values); the engine's
 * iterate and learned Hessians are unchanged. Used to state the error of the
 * reduced-precision modes against FP64 (logistic_regression.h:365-547 is the
 * FP64 reference of this quantity). */
int fednl_hessians_at(void *handle, const double *x, double *out);

/* Diagnostic: one v_mfma_f32_16x16x4f32 (a32, b32: 64 per-lane values) and
 * one v_mfma_f32_16x16x32_bf16 (abf, bbf: 64 x 8 per-lane values, rounded to
 * BF16), 64 x 4 outputs each; pins the lane layouts the reduced-precision
 * Hessian kernels rely on. No counterpart in the reference. */
int fednl_probe_mfma_mixed(const float *a32, const float *b32, float *out32, const float *abf,
                           const float *bbf, float *outbf);

/* Drop-in for dopt's shared-library entry point
 * (dopt/bin_fednl_local/exported/FedNLLocalShared.h:8-17,
 *  dopt/bin_fednl_local/src/mainShared.cpp:15-39): `command` holds the
 * bin_fednl_local flags (without "train"), split on single spaces. The
 * callback receives (d, ||grad f(x_last)||, ||grad||^2/(2 mu), 2||grad||^2/mu)
 * (TrainLoop.h:1854-1859) once, on success. Returns dopt's TrainReturnCodes. */
typedef void (*ResultCallbackLocal)(int xDimension, double discrWithOptSolution,
                                    double discrWithOptValue, double L2NormInLastIterate);
int doptExecuteLocal(const char *command, ResultCallbackLocal resultCallback);

/* The CLI entry (bin_fednl_local main, dopt/bin_fednl_local/src/main.cpp:25-59). */
int fednl_cli_main(int argc, char **argv);

/* FedNL-PP drop-ins: bin_fednl_local_pp's main (dopt/bin_fednl_local_pp/src/main.cpp:22-56)
 * and its shared-library entry (dopt/bin_fednl_local_pp/exported/FedNLPPLocalShared.h:8-17,
 * src/mainShared.cpp:15-39): the bin_fednl_local flags without the line
 * search, plus --clients-per-round. The callback receives d once, on success
 * (TrainLoop.h:1870-1872). */
typedef void (*ResultCallbackLocalPP)(int xDimension);
int doptExecuteLocalPP(const char *command, ResultCallbackLocalPP resultCallback);
int fednl_cli_pp_main(int argc, char **argv);

/* The CLI's LIBSVM reader, as train() reads the training set
 * (TrainLoop.h:123-170 over dataset_loader_fs.h:381-883): N samples of d
 * attributes (+1 intercept column, last, when add_intercept), labels mapped
 * to +-1. With X == NULL only *N and *d are set; else X (N x d, row-major)
 * and y (N) are filled. Returns 0, FEDNL_E_MISSING_ARGUMENT for a null
 * pointer, or -4 (the CLI's dataset parse error) as train() would. */
int fednl_cli_load_libsvm(const char *path, int add_intercept, int *N, int *d, double *X, double *y);

/* The CLI's CRC-32 (dopt/dopt/system/src/digest/Crc.cpp:74-90, seed
 * 0xFFFFFFFF, no final inversion), as printed for the dataset
 * (Reports.h:25-33) and the last iterate (TrainLoop.h:1767-1768). */
uint32_t fednl_cli_crc32(const void *buf, size_t size);

#ifdef __cplusplus
}
#endif

#endif /* FEDNL_HIP_H */
