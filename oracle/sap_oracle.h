/*
 * CPU oracle for the SaP (SPIKE) preconditioned Krylov solver.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  It is a plain-loop restatement of the reference
 * algorithm (spikegpu/SaPLibrary, sap/precond.h, sap/bicgstab2.h,
 * sap/minres.h, sap/monitor.h) used by tests/ to check the HIP path, by
 * __graft_entry__.smoke() as the checker, and by bench.py's CPU-baseline leg.
 *
 * Conventions (identical to the shipped C ABI in include/sap_c.h):
 *   - input banded matrix: row-major band, A(i,j) at band[i*(2K+1) + K + (j-i)]
 *   - partitions: the first N mod P partitions hold N/P + 1 rows
 */
#ifndef SAP_ORACLE_H
#define SAP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Options subset mirrored from sap::Options (sap/solver.h:52-81). */
typedef struct orc_options {
    int    solver;            /* 0 = BiCGStab(2), 1 = BiCGStab(1), 2 = MINRES        */
    int    max_iters;         /* maxNumIterations                                    */
    double rel_tol;           /* relTol                                              */
    double abs_tol;           /* absTol                                              */
    int    fact_method;       /* 0 = LU_UL, 1 = LU_only  (sap/common.h:84-87)        */
    int    precond_type;      /* 0 = Spike, 1 = Block, 2 = None (common.h:92-96)     */
    int    safe_fact;         /* safeFactorization                                   */
    int    prec_double;       /* 1: preconditioner in FP64, 0: FP32                  */
    int    refine;            /* outer FP64 refinement sweeps around the Krylov solve
                                 (0 = the reference's single Krylov solve)            */
    double refine_tol;        /* relative tolerance of each inner Krylov solve        */
} orc_options;

typedef struct orc_stats {
    float  num_iterations;
    double rhs_norm;
    double residual_norm;
    double rel_residual_norm;
    int    code;
    int    num_partitions;
    int    bandwidth;
    double time_setup;
    double time_solve;
    double time_bandLU;
    double time_bandUL;
    double time_assembly;
    double time_fullLU;
    int    num_applies;
    int    num_spmv;
    int    num_refine;
} orc_stats;

/* ---- workload generators (xoshiro256** seeded per row by splitmix64) ---- */
void   orc_gen_banded(int64_t n, int k, double dominance, uint64_t seed, double* band);
void   orc_gen_vector(int64_t n, uint64_t seed, double* x);

/* ---- banded FP64 SpMV on the row-major band ---- */
void   orc_band_spmv(int64_t n, int k, const double* band, const double* x, double* y);

/* ---- FP64 dense-band LU with partial pivoting (LAPACK gbtrf/gbtrs scheme) ---- */
int    orc_band_solve_pp(int64_t n, int k, const double* band, const double* b, double* x);

/* ---- the SaP solver ---- */
void*  orc_create(int num_partitions, const orc_options* opts);
void   orc_destroy(void* s);
/* returns 0 on success, -1 Zero_pivoting (sap/exception.h:16-23) */
int    orc_setup_banded(void* s, int64_t n, int k, const double* band);
/* returns 1 converged / 0 not converged, -4 Illegal_solve */
int    orc_solve(void* s, const double* b, double* x);
void   orc_get_stats(void* s, orc_stats* st);
const char* orc_monitor_message(void* s);
/* z = M^{-1} v (FP64 in/out, preconditioner precision inside) */
int    orc_precond_apply(void* s, const double* v, double* z);
int    orc_num_partitions(void* s);
/* copy of the factored band (column-major, (2K+1)*N) as double */
int    orc_get_factors(void* s, double* out);
/* copy of R (4K^2 (P-1)) after the reduced LU, and of the spike tips WV (2K^2 (P-1)) */
int    orc_get_reduced(void* s, double* R_out, double* WV_out);

#ifdef __cplusplus
}
#endif
#endif
