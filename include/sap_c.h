/*
 * sap_c.h -- C ABI of the MI355X-native SaP (SPIKE) solver.
 *
 * One entry point per member of the reference's public surface
 * (spikegpu/SaPLibrary sap/solver.h), with plain pointers and sizes:
 *
 *   sap::Options / Options::Options()      -> sap_options / sap_default_options
 *                                              (sap/solver.h:52-81, :232-256)
 *   sap::Stats                             -> sap_stats        (sap/solver.h:89-140)
 *   Solver(int numPartitions, Options)     -> sap_create       (sap/solver.h:157-158, :302-328)
 *   ~Solver()                              -> sap_destroy      (sap/solver.h:159-168)
 *   Solver::setup(BandedMatrix)            -> sap_setup_banded (sap/solver.h:339-404,
 *                                              sap/precond.h:2103-2233)
 *   Solver::setup(cusp::csr_matrix)        -> sap_setup_csr    (sap/solver.h:339-404,
 *                                              sap/precond.h:1656-2101, :2240-2363)
 *   SpmvCusp::operator()(v, Av)            -> sap_csr_spmv     (sap/spmv.h:83-129)
 *   Graph::reorder + Graph::dropOff        -> sap_reorder_csr  (sap/graph.h:446-747,
 *     (host side of transformToBandedMatrix)                    sap/precond.h:1656-1776)
 *   Graph::assembleOffDiagMatrices,        -> sap_reordering_assemble_var
 *     secondLevelReordering, partitionedRCM,   (sap/graph.h:756-1206, :1593-1757)
 *     assembleBandedMatrix
 *   Solver::update(entries)                -> sap_update       (sap/solver.h:421-462,
 *                                              sap/precond.h:751-900)
 *   Solver::solve(spmv, b, x)              -> sap_solve        (sap/solver.h:477-562)
 *   Solver::getStats()                     -> sap_get_stats    (sap/solver.h:182)
 *   Solver::getMonitorCode()               -> sap_monitor_code (sap/solver.h:183-188)
 *   Solver::getMonitorMessage()            -> sap_monitor_message (sap/solver.h:189-194)
 *   MVBanded::operator()(v, Av)            -> sap_band_spmv    (sap/spmv.h:27-74)
 *   Precond::operator()(v, z)              -> sap_precond_apply (sap/precond.h:1370-1386)
 *
 * Conventions
 *   - All array arguments named d_* are device pointers owned by the caller.
 *     The solver copies what it keeps, except the banded matrix given to
 *     sap_setup_banded: like the reference's MVBanded functor (sap/spmv.h:73)
 *     the built-in SpMV keeps a non-owning pointer to it, so it must stay
 *     alive until the last sap_solve that uses the built-in SpMV.
 *   - Banded input: row-major band, A(i,j) at d_band[i*(2k+1) + k + (j-i)]
 *     (sap/banded_matrix.h:74-92).
 *   - CSR input: zero-based row offsets (n+1), column indices and FP64
 *     values (nnz), the layout of cusp::csr_matrix<int,double>.  The
 *     built-in CSR SpMV keeps non-owning pointers to them, as for the band.
 *   - Return codes mirror sap::system_error::Reason (sap/exception.h:16-23):
 *     0 OK, -1 zero pivot, -2 negative DB weight, -3 illegal update,
 *     -4 illegal solve, -5 singular matrix; plus -6 out of device memory,
 *     -7 HIP runtime error, -8 bad argument, -9 unsupported option.
 *     sap_solve returns 1 (converged) or 0 (not converged) when it runs.
 *   - No exception crosses this boundary.  A solver object is not
 *     thread-safe; all work is queued on one HIP stream (sap_set_stream).
 */
#ifndef SAP_C_H
#define SAP_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAP_OK                 0
#define SAP_ERR_ZERO_PIVOT    -1
#define SAP_ERR_NEG_DB_WEIGHT -2
#define SAP_ERR_ILLEGAL_UPDATE -3
#define SAP_ERR_ILLEGAL_SOLVE -4
#define SAP_ERR_SINGULAR      -5
#define SAP_ERR_OOM           -6
#define SAP_ERR_HIP           -7
#define SAP_ERR_BADARG        -8
#define SAP_ERR_UNSUPPORTED   -9

/* sap::KrylovSolverType (sap/common.h:71-82), same order */
enum { SAP_BiCGStab_C = 0, SAP_GMRES_C, SAP_CG_C, SAP_CR_C,
       SAP_BiCGStab1, SAP_BiCGStab2, SAP_BiCGStab, SAP_MINRES };
/* sap::FactorizationMethod (sap/common.h:84-87) */
enum { SAP_LU_UL = 0, SAP_LU_only = 1 };
/* sap::PreconditionerType (sap/common.h:92-96) */
enum { SAP_Spike = 0, SAP_Block = 1, SAP_None = 2 };
/* the PrecValueType template argument of sap::Solver */
enum { SAP_PREC_FLOAT = 0, SAP_PREC_DOUBLE = 1 };

/* Field-for-field mirror of sap::Options (sap/solver.h:52-81) followed by
 * the options this implementation adds. */
typedef struct sap_options {
    int    solverType;          /* default SAP_BiCGStab2                           */
    int    maxNumIterations;    /* default 100                                     */
    double relTol;              /* default 1e-6                                    */
    double absTol;              /* default 0                                       */
    int    testDB;
    int    isSPD;
    int    saveMem;
    int    performReorder;      /* default 1                                       */
    int    performDB;           /* default 1                                       */
    int    dbFirstStageOnly;
    int    applyScaling;        /* default 1                                       */
    int    maxBandwidth;        /* default INT_MAX                                 */
    int    gpuCount;            /* default 1; only 1 is supported                  */
    double dropOffFraction;     /* default 0                                       */
    int    factMethod;          /* default SAP_LU_only                             */
    int    precondType;         /* default SAP_Spike                               */
    int    safeFactorization;   /* default 0                                       */
    int    variableBandwidth;   /* default 1                                       */
    int    trackReordering;     /* default 0                                       */
    int    useBCR;              /* default 0; 1: block cyclic reduction (Block)    */
    int    ilu_level;           /* default -1; only -1 is supported                */
    /* ---- additions ---- */
    int    precValueType;       /* SAP_PREC_FLOAT (default) or SAP_PREC_DOUBLE     */
    int    refine;              /* outer FP64 refinement steps around the Krylov
                                   solve; 0 (default) runs the reference's single
                                   right-preconditioned solve                       */
    double refineTol;           /* relative tolerance of each inner solve (1e-6)   */
    int    profile;             /* 1: time every kernel family with HIP events     */
    int    batch;               /* nrhs > 1 with BiCGStab2 and the built-in SpMV:
                                   1 (default) solves the columns together (one
                                   launch per kernel family for all of them, the
                                   scalars on the device), 0 one after another     */
    int    graph;               /* batched solves: 1 (default) replays each
                                   BiCGStab(2) iteration from a captured HIP graph
                                   (off while profiling)                            */
    int    fuse;                /* sap_batch_solve: 1 (default) runs the sweeps of all
                                   systems as one launch when every system's
                                   preconditioner is one sweep pass (Block, or one
                                   partition); 0 one launch per system             */
    int    autotune;            /* 1: on the first setup of each half-bandwidth and
                                   precision, time the candidate conversion tile
                                   widths and sweep staging depths and keep the
                                   fastest for the process (results unchanged);
                                   0 (default): fixed choices                      */
} sap_options;

/* Field-for-field mirror of sap::Stats (sap/solver.h:89-140), times in
 * milliseconds as the reference reports them, followed by additions. */
typedef struct sap_stats {
    double timeSetup;
    double timeUpdate;
    double timeSolve;
    double time_DB;
    double time_DB_pre;
    double time_DB_first;
    double time_DB_second;
    double time_DB_post;
    double d_p1;
    double d_p1_ori;
    double diag_dom;
    double diag_dom_ori;
    double time_reorder;
    double time_dropOff;
    double time_cpu_assemble;
    double time_transfer;
    double time_toBanded;
    double time_offDiags;
    double time_bandLU;
    double time_bandUL;
    double time_fullLU;
    double time_assembly;
    double time_shuffle;
    double time_bcr_lu;
    double time_bcr_sweep_deflation;
    double time_bcr_mat_mul_deflation;
    double time_bcr_sweep_inflation;
    double time_bcr_mv_inflation;
    int    bandwidthReorder;
    int    bandwidthDB;
    int    bandwidth;
    double nuKf;
    double flops_LU;
    int    numPartitions;
    double actualDropOff;
    float  numIterations;
    double rhsNorm;
    double residualNorm;
    double relResidualNorm;
    int    actual_nnz;
    /* ---- additions ---- */
    int    numApplies;          /* preconditioner applications in the last solve  */
    int    numSpmv;             /* SpMV calls in the last solve                    */
    int    numRefine;           /* outer refinement steps in the last solve        */
} sap_stats;

/* Kernel-family timing collected when sap_options.profile = 1. */
typedef struct sap_kernel_profile {
    const char* name;
    int         launches;
    double      total_ms;
} sap_kernel_profile;

typedef struct sap_solver sap_solver;

/* y = A*v for a user-supplied operator (the SPMV functor of sap/spmv.h:38-56,
 * examples/sequence/driver_seq.cu:82-96).  Called on the solver's stream. */
typedef void (*sap_spmv_fn)(void* ctx, const double* d_v, double* d_Av, int n);

void        sap_default_options(sap_options* opts);
sap_solver* sap_create(int numPartitions, const sap_options* opts);
void        sap_destroy(sap_solver* s);
int         sap_set_stream(sap_solver* s, void* hip_stream);

int         sap_setup_banded(sap_solver* s, int n, int k, const double* d_band);

/* Solver::setup with a CSR matrix (sap/solver.h:339-404).  With
 * performReorder the host runs DB (performDB; applyScaling scales), RCM
 * (when maxBandwidth > 64) and drop-off (dropOffFraction / maxBandwidth)
 * as transformToBandedMatrix does (sap/precond.h:1656-1776), the band of
 * the reordered matrix is factored, and every preconditioner application
 * permutes and scales its vectors (sap/precond.h:1393-1406, :1535-1621).
 * Without performReorder the matrix is converted to a band of half-width
 * max|i-j| as is (sap/precond.h:2240-2363).
 * With variableBandwidth (and more than one partition) the preconditioner
 * is the variable-bandwidth path's truncated-SPIKE operator, factored
 * LU_only with every pivot boosted, on the constant-width band of the
 * reordered matrix.
 * Returns SAP_ERR_SINGULAR when DB finds no perfect matching,
 * SAP_ERR_NEG_DB_WEIGHT on a negative reduced cost.  With testDB only the
 * DB test runs (see sap_reordering_dominance). */
int         sap_setup_csr(sap_solver* s, int n, int nnz, const int* d_row_offsets,
                          const int* d_col_indices, const double* d_values);

/* Solver::update: refactors with new values of the setup matrix, same
 * pattern and same reordering (sap/solver.h:421-462).  d_entries holds count
 * doubles in the layout of the last setup: the CSR values in their order
 * (the reference's entries), or, for a banded setup, the row-major band
 * (an extension: the reference tracks only CSR setups).  Needs
 * trackReordering; SAP_ERR_ILLEGAL_UPDATE before a setup or without it.
 * Returns 1 when done and 0 when count does not match the setup matrix
 * (the reference's false).  The built-in SpMV then uses d_entries
 * (non-owning, like the setup matrix). */
int         sap_update(sap_solver* s, const double* d_entries, int64_t count);

/* Solves A x = b.  d_x holds the initial guess on entry and the solution on
 * return.  spmv == NULL uses the built-in SpMV of the setup matrix (banded or
 * CSR).
 * nrhs > 1 solves the columns of b/x (leading dimension n): with the
 * options' batch = 1, BiCGStab2 and the built-in SpMV all columns together
 * (each column's iterates are those of solving it alone; without
 * refinement bit for bit), otherwise one after another.  Stats then hold
 * the largest iteration count and relative residual over the columns; the
 * monitor code and message are those of the first column that did not
 * converge (of the last column when all did). */
int         sap_solve(sap_solver* s, sap_spmv_fn spmv, void* ctx,
                      const double* d_b, double* d_x, int nrhs);

/* sap_solve without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (serving loops that
 * call solve repeatedly, harnesses that forbid hipMalloc).  sap_solve
 * allocates the solver's Krylov vector pool on its first call; this variant
 * takes the pool (32 vectors of n doubles) and the reduction partials from
 * a static device array of the library instead: 32 * SAP_NOALLOC_MAX_N +
 * 4096 doubles (32 MiB + 32 KiB), one copy per GPU, held until the process
 * ends.  n above SAP_NOALLOC_MAX_N returns SAP_ERR_BADARG, profile = 1
 * returns SAP_ERR_UNSUPPORTED.  The setup's buffers are used as by
 * sap_solve; the array keeps the last call's contents (every vector is
 * written before it is read).  All calls on one GPU share the array, so
 * calls that overlap on one GPU (any solver, any stream) race.  Several
 * right-hand sides are solved one after another; results are those of
 * sap_solve with batch = 0, bit for bit. */
#define SAP_NOALLOC_MAX_N 131072
int         sap_solve_noalloc(sap_solver* s, sap_spmv_fn spmv, void* ctx,
                              const double* d_b, double* d_x, int nrhs);

/* Many independent banded systems solved together.  No counterpart in the
 * reference (its kernels index right-hand sides by block, but a Precond
 * applies one vector at a time, sap/precond.h:1370-1386): each system gets
 * its own partitions and preconditioner (as sap_setup_banded builds it),
 * and all right-hand sides of all systems are one batch of BiCGStab(2)
 * solves (options as for sap_solve with several right-hand sides; only
 * BiCGStab2).  Each column's iterates are those of solving it alone.
 *   sap_batch_create: nsys systems, numPartitions[s] partitions each; NULL
 *     on bad options.
 *   sap_batch_setup_banded: host arrays of nsys sizes, half-bandwidths and
 *     device pointers to the row-major bands (kept non-owning, as by
 *     sap_setup_banded).
 *   sap_batch_solve: d_b[s], d_x[s] device pointers to the column-major
 *     n[s] x nrhs[s] right-hand sides and initial guesses / solutions;
 *     returns 1 when every column converged, 0 otherwise.
 *   sap_batch_result: column c (systems in order, nrhs[s] columns each) of
 *     the last solve: monitor code, iterations, relative residual.
 *   sap_batch_times: wall times of the last setup and solve (ms). */
typedef struct sap_batch sap_batch;
sap_batch*  sap_batch_create(int nsys, const int* numPartitions, const sap_options* opts);
void        sap_batch_destroy(sap_batch* b);
int         sap_batch_set_stream(sap_batch* b, void* hip_stream);
int         sap_batch_setup_banded(sap_batch* b, const int* n, const int* k, const double* const* d_bands);
int         sap_batch_solve(sap_batch* b, const double* const* d_b, double* const* d_x, const int* nrhs);
int         sap_batch_result(const sap_batch* b, int column, int* code, float* iterations, double* rel_residual);
int         sap_batch_times(const sap_batch* b, double* setup_ms, double* solve_ms);
const char* sap_batch_last_error(const sap_batch* b);

int         sap_get_stats(const sap_solver* s, sap_stats* out);
int         sap_monitor_code(const sap_solver* s);
const char* sap_monitor_message(const sap_solver* s);
const char* sap_last_error(const sap_solver* s);

/* z = M^{-1} v through the preconditioner built by the last setup. */
int         sap_precond_apply(sap_solver* s, const double* d_v, double* d_z);

/* y = A x for a row-major FP64 band on the given HIP stream (NULL: default). */
int         sap_band_spmv(int n, int k, const double* d_band, const double* d_x,
                          double* d_y, void* hip_stream);

/* y = A x for an FP64 CSR matrix on the given HIP stream (NULL: default). */
int         sap_csr_spmv(int n, int nnz, const int* d_row_offsets, const int* d_col_indices,
                         const double* d_values, const double* d_x, double* d_y, void* hip_stream);

/* The host reordering of a CSR setup on its own (host arrays in and out, no
 * GPU): DB, RCM and drop-off with the options' performDB, applyScaling,
 * dbFirstStageOnly, maxBandwidth and dropOffFraction, in the precision of
 * precValueType.  NULL on failure with *err set (SAP_ERR_SINGULAR,
 * SAP_ERR_NEG_DB_WEIGHT, SAP_ERR_BADARG, SAP_ERR_OOM).
 *   db_perm[i]   new row of original row i after DB      (dbRowPerm)
 *   row_perm[i]  new row of original row i, DB then RCM  (m_optPerm)
 *   col_reord[j] original column of new column j         (m_optReordering)
 *   row_scale, col_scale: DB scalings (all 1 without DB)
 *   row_offsets (n+1), col_indices, values (nnz): the reordered, scaled and
 *   dropped-off matrix whose band is factored. */
typedef struct sap_reordering sap_reordering;
sap_reordering* sap_reorder_csr(int n, const int* h_row_offsets, const int* h_col_indices,
                                const double* h_values, const sap_options* opts, int* err);
int         sap_reordering_info(const sap_reordering* r, int* k, int* k_db, int* k_reorder,
                                double* drop_off, int* nnz);
int         sap_reordering_get(const sap_reordering* r, int* db_perm, int* row_perm, int* col_reord,
                               double* row_scale, double* col_scale, int* row_offsets,
                               int* col_indices, double* values);
void        sap_reordering_free(sap_reordering* r);
/* With the options' testDB (Graph::reorder's DB test, sap/graph.h:468-490,
 * :537-561): the reordering stops after DB (no RCM, no drop-off; k is the
 * bandwidth after DB) and these are the diagonal-dominance statistics of
 * d_i = |a_ii| / sum_{j != i} |a_ij|: the minimum and the 1st percentile
 * (entry n/100 of the sorted d) before and after DB (0 without testDB).
 * A sap_setup_csr with testDB runs the same test, fills these Stats fields
 * and builds no preconditioner (a following sap_solve returns
 * SAP_ERR_ILLEGAL_SOLVE). */
int         sap_reordering_dominance(const sap_reordering* r, double* diag_dom_ori, double* d_p1_ori,
                                     double* diag_dom, double* d_p1);

/* The variable-bandwidth assembly of the reordered matrix into
 * numPartitions partitions (clamped to max(n/(k+1), 1) as
 * transformToBandedMatrix does, sap/precond.h:1777-1783), on the host:
 * Graph::assembleOffDiagMatrices, secondLevelReordering + partitionedRCM and
 * assembleBandedMatrix (sap/graph.h:756-1206, :1593-1757).  Returns the
 * partition count used (1: no variable-bandwidth assembly, as the reference
 * falls back to the constant band) or a negative code. */
int         sap_reordering_assemble_var(sap_reordering* r, int numPartitions);
/* Its results (NULL skips an array); P partitions, k the half-bandwidth:
 *   ks (P), boffsets (P+1), first_rows (P-1), second_reorder / second_perm (n),
 *   widths_left / widths_right (P-1), perms_left / perms_right ((P-1) k),
 *   wv / offdiags (2 k^2 (P-1)), ks_col / ks_row (n), and the diagonal blocks
 *   after the second-level permutation as CSR (n+1, nnz_d, nnz_d). */
int         sap_reordering_var_get(const sap_reordering* r, int* ks, int64_t* boffsets, int* first_rows,
                                   int* second_reorder, int* second_perm, int* widths_left,
                                   int* widths_right, int* perms_left, int* perms_right, double* wv,
                                   double* offdiags, int* ks_col, int* ks_row, int* diag_row_offsets,
                                   int* diag_col_indices, double* diag_values);

/* Matrix Market input (host only, no GPU), the file format the reference's
 * drivers read with cusp::io::read_matrix_market_file into a CSR matrix
 * (examples/matrix_market/driver_mm.cu:135-140).  Coordinate and array
 * files; real, integer and pattern fields; general, symmetric and
 * skew-symmetric storage (the other triangle is generated).  The CSR is
 * zero-based, sorted by (row, column), duplicates kept as separate entries.
 * NULL on failure, with *err set (SAP_ERR_BADARG for a malformed or missing
 * file, SAP_ERR_OOM) and the reason copied into msg (msg_len bytes). */
typedef struct sap_mm_matrix sap_mm_matrix;
sap_mm_matrix* sap_mm_read(const char* path, int* err, char* msg, int msg_len);
int         sap_mm_info(const sap_mm_matrix* m, int* rows, int* cols, int64_t* nnz);
/* copies (rows + 1) offsets, nnz column indices and nnz values (NULL skips) */
int         sap_mm_get_csr(const sap_mm_matrix* m, int* row_offsets, int* col_indices, double* values);
void        sap_mm_free(sap_mm_matrix* m);

/* The autotuner's choices so far (options.autotune = 1), one entry per
 * (family, half-bandwidth, element size): family 0 = conversion tile width
 * (columns per workgroup), 1 = sweep staging depth (LDS block buffers).
 * Returns the number of entries (copies at most max_entries). */
typedef struct sap_tune_entry {
    int    family;
    int    k;
    int    elem_bytes;
    int    value;
    int    default_value;
    int    candidates;
    double best_ms;
    double default_ms;
} sap_tune_entry;
int         sap_autotune_entries(sap_tune_entry* out, int max_entries);

/* The learned option selector's network (no counterpart in the
 * reference): a fully connected FP32 network n_in -> n_hidden -> n_hidden
 * -> n_hidden -> n_out with ReLU, trained on the GPU by hand-written
 * kernels (forward, backward of the masked mean squared error, Adam).
 * Parameters are packed per layer (weights n_in_l x n_out_l row-major, then
 * biases); d_* are device float arrays (X: batch x n_in, Y/T/M: batch x
 * n_out, M = 1 where a target is known, 0 elsewhere).
 *   sap_mlp_create: He-initialised weights from seed (NULL on bad sizes).
 *   sap_mlp_gradient: the gradient of the loss with respect to every
 *     parameter (host copy) and the loss, without updating.
 *   sap_mlp_train: epochs full-batch Adam steps at learning rate lr; the
 *     loss before each step into h_losses (may be NULL). */
typedef struct sap_mlp sap_mlp;
sap_mlp*    sap_mlp_create(int n_in, int n_hidden, int n_out, uint64_t seed);
void        sap_mlp_destroy(sap_mlp* m);
int64_t     sap_mlp_num_params(const sap_mlp* m);
int         sap_mlp_get_params(const sap_mlp* m, float* h_out);
int         sap_mlp_set_params(sap_mlp* m, const float* h_in);
int         sap_mlp_forward(sap_mlp* m, const float* d_X, int batch, float* d_Y);
int         sap_mlp_gradient(sap_mlp* m, const float* d_X, const float* d_T, const float* d_M, int batch,
                             float* h_grad, float* h_loss);
int         sap_mlp_train(sap_mlp* m, const float* d_X, const float* d_T, const float* d_M, int batch, int epochs,
                          float lr, float* h_losses);
const char* sap_mlp_last_error(const sap_mlp* m);

/* Measurement: the best of reps timed launches (after one warm-up) of the
 * library's streaming-read kernel over n doubles (n even, 16-byte aligned
 * device pointer), in ms on the null stream: the device's practical read
 * bandwidth, beside which the band-streaming kernels are quoted. */
int         sap_stream_read_ms(const double* d_x, int64_t n, int reps, double* best_ms);

/* Kernel-family timing since the last sap_reset_profile (profile = 1). */
int         sap_get_profile(const sap_solver* s, sap_kernel_profile* out, int max_entries);
void        sap_reset_profile(sap_solver* s);

/* Diagnostics for tests: copies of the factored band (column-major,
 * (2k+1) x n, as FP64; after a variable-bandwidth setup the partitions' own
 * bands one after the other, partition p (2 K_p + 1) x n_p in second-level
 * order, K_p = max(1, the partition's half-bandwidth)), of the K x K spike
 * tips (2 k^2 (P-1)) and of the Schur blocks S_i = I - W_i V_i after
 * factorization (k^2 (P-1)); host out. */
int         sap_get_factors(const sap_solver* s, double* h_out, int64_t count);
int         sap_get_spikes(const sap_solver* s, double* h_wv, double* h_schur, int64_t count);
int         sap_num_partitions(const sap_solver* s);

/* Synthetic workload generators (xoshiro256** per row, seeded by splitmix64):
 * off-diagonals U[-1,1), diagonal = dominance * row off-diagonal abs-sum. */
int         sap_gen_banded(int64_t n, int k, double dominance, uint64_t seed,
                           double* d_band, void* hip_stream);
int         sap_gen_vector(int64_t n, uint64_t seed, double* d_x, void* hip_stream);

#ifdef __cplusplus
}
#endif
#endif /* SAP_C_H */
