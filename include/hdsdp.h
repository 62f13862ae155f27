/*
 * hdsdp.h -- public solver API of the MI355X HDSDP build.
 *
 * Same names, argument meaning and return codes as the reference's public API
 * (interface/hdsdp.h:40-116) and user-data API (interface/hdsdp_user_data.h:18-23),
 * so a caller of the reference links against this library unchanged.
 * Extensions (not in the reference) are grouped at the end and marked.
 */
#ifndef HDSDP_AMD_HDSDP_H
#define HDSDP_AMD_HDSDP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* interface/hdsdp.h:40-46 */
typedef enum {
    HDSDP_RETCODE_OK,
    HDSDP_RETCODE_FAILED,
    HDSDP_RETCODE_MEMORY
} hdsdp_retcode;

/* interface/hdsdp.h:48-62 */
typedef enum {
    HDSDP_UNKNOWN,
    HDSDP_DUAL_FEASIBLE,
    HDSDP_DUAL_OPTIMAL,
    HDSDP_PRIMAL_DUAL_OPTIMAL,
    HDSDP_MAXITER,
    HDSDP_SUSPECT_INFEAS_OR_UNBOUNDED,
    HDSDP_INFEAS_OR_UNBOUNDED,
    HDSDP_TIMELIMIT,
    HDSDP_USER_INTERRUPT,
    HDSDP_INTERNAL_ERROR,
    HDSDP_NUMERICAL
} hdsdp_status;

/* interface/def_hdsdp_conic.h:22-32 */
typedef enum {
    HDSDP_CONETYPE_UNKNOWN,
    HDSDP_CONETYPE_LP,
    HDSDP_CONETYPE_BOUND,
    HDSDP_CONETYPE_SCALAR_BOUND,
    HDSDP_CONETYPE_DENSE_SDP,
    HDSDP_CONETYPE_SPARSE_SDP,
    HDSDP_CONETYPE_SOCP
} cone_type;

typedef struct hdsdp_solver_internal hdsdp;
typedef struct hdsdp_user_data user_data;

#define HDSDP_INFINITY 1e+30

/* interface/hdsdp.h:69-89 */
#define INT_PARAM_MAXITER       0
#define INT_PARAM_CORRECTORA    1
#define INT_PARAM_CORRECTORB    2
#define INT_PARAM_THREADS       3
#define INT_PARAM_PSDP          4

#define DBL_PARAM_RELFEASTOL    0
#define DBL_PARAM_RELOPTTOL     1
#define DBL_PARAM_ABSFEASTOL    2
#define DBL_PARAM_ABSOPTTOL     3
#define DBL_PARAM_TIMELIMIT     4
#define DBL_PARAM_POTRHOVAL     5
#define DBL_PARAM_HSDGAMMA      6
#define DBL_PARAM_DUALBOX_LOW   7
#define DBL_PARAM_DUALBOX_UP    8
#define DBL_PARAM_BARMUSTART    9
#define DBL_PARAM_DUALSTART    10
#define DBL_PARAM_POBJSTART    11
#define DBL_PARAM_TRXESTIMATE  12
#define DBL_PARAM_PRECORDACC   13

/* Solver API: interface/hdsdp.h:105-116 */
hdsdp_retcode HDSDPCreate( hdsdp **pHSolver );
hdsdp_retcode HDSDPInit( hdsdp *HSolver, int nRows, int nCones );
hdsdp_retcode HDSDPSetCone( hdsdp *HSolver, int iCone, void *userCone );
void HDSDPSetDualObjective( hdsdp *HSolver, double *dObj );
hdsdp_retcode HDSDPOptimize( hdsdp *HSolver, int dOptOnly );
void HDSDPSetIntParam( hdsdp *HSolver, int intParam, int intParamVal );
void HDSDPSetDblParam( hdsdp *HSolver, int intParam, double dblParamVal );
hdsdp_retcode HDSDPGetRowDual( hdsdp *HSolver, double *pObjVal, double *dObjVal, double *dualVal );
void HDSDPGetConeValues( hdsdp *HSolver, int iCone, double *conePrimal, double *coneDual, double *coneAuxi );
hdsdp_retcode HDSDPCheckSolution( hdsdp *HSolver, double diErrors[6] );
void HDSDPClear( hdsdp *HSolver );
void HDSDPDestroy( hdsdp **pHSolver );

/* User data: interface/hdsdp_user_data.h:18-23.  SDP block = CSC with m + 1
 * columns over packed-lower row indices; column 0 is C, column i + 1 is A_i. */
hdsdp_retcode HUserDataCreate( user_data **pHdata );
void HUserDataSetConeData( user_data *Hdata, cone_type cone, int nRow, int nCol,
                           int *coneMatBeg, int *coneMatIdx, double *coneMatElem );
cone_type HUserDataChooseCone( user_data *Hdata );
void HUserDataClear( user_data *Hdata );
void HUserDataDestroy( user_data **pHdata );

/* ------------------------------------------------------------------------
 * Extensions (no counterpart in the reference)
 * ------------------------------------------------------------------------ */

/* Factor-form rank-one block: A_i = sign[i] * f_i f_i^T with f_i column i of the
 * n x m column-major `factor`; C is given as one CSC column (packed-lower rows).
 * The CSC interface cannot express this at n = 2000, m = 20000 (4e10 packed
 * entries overflow its int32 pointers). */
void HUserDataSetRankOneConeData( user_data *Hdata, int nRow, int nCol,
                                  int cNnz, int *cIdx, double *cElem,
                                  double *factor, double *sign );

/* Raw parameter write without the reference setter's index filter
 * (interface/hdsdp.c:596-615 drops INT_PARAM_THREADS/PSDP, DBL_PARAM_PRECORDACC). */
void HDSDPSetIntParamAny( hdsdp *HSolver, int intParam, int intParamVal );
void HDSDPSetDblParamAny( hdsdp *HSolver, int dblParam, double dblParamVal );

/* Arithmetic of the Schur complement path: 0 = FP64 throughout (default);
 * 1 = mixed (dense and rank-one Schur blocks from bf16 split products with FP32
 * accumulation, FP32-equivalent; FP32 Cholesky of M as preconditioner of an FP64
 * CG on the exact Schur operator; FP64 rebuild of M when that system fails). */
#define HDSDP_PRECISION_FP64  0
#define HDSDP_PRECISION_MIXED 1
void HDSDPSetPrecision( hdsdp *HSolver, int precisionMode );

/* Per-iteration log, one record per printed log line:
 * {phase, iter, pObj, dObj, inf, mu, step, prox, seconds}, phase 0 = infeasible
 * dual (Phase A), 1 = self-dual embedding, 2 = feasible potential reduction. */
#define HDSDP_LOG_FIELDS 9
int HDSDPGetIterLog( hdsdp *HSolver, int maxRecords, double *records );

/* Solver summary: status, iteration count, barrier mu, wall time of the last
 * HDSDPOptimize and the time spent inside it up to the first logged relative
 * gap (pObj - dObj) / (|pObj| + |dObj| + 1) <= gapTarget (negative if never). */
void HDSDPGetSummary( hdsdp *HSolver, double gapTarget, int *status, int *nIter,
                      double *mu, double *solveSeconds, double *secondsToGap );

/* DIMACS errors computed at the end of the last HDSDPOptimize. */
void HDSDPGetDimacsErrors( hdsdp *HSolver, double errs[6] );

/* Silence the iteration log (the reference prints unconditionally). */
void HDSDPSetVerbose( hdsdp *HSolver, int verbose );

/* Device-time profile of the last HDSDPOptimize (slots HDSDP_PROF_* of
 * hdsdp_gpu.h); HDSDPSetProfiling(1) before HDSDPOptimize turns it on.
 * HDSDPGetProfile returns the number of slots filled. */
void HDSDPSetProfiling( hdsdp *HSolver, int enable );
int HDSDPGetProfile( hdsdp *HSolver, double *ms, double *counts, int nSlots );

/* Pre-solve and device set-up without the interior-point iterations (what
 * HDSDPOptimize does first; HDSDPOptimize skips it when it has been done), so
 * that HDSDPCenter can work on a problem that was never solved. */
hdsdp_retcode HDSDPSetup( hdsdp *HSolver );

/* The centred dual point and its adjoint.  For mu > 0, y_mu(b, C) maximises
 * b^T y + mu log det(C - sum_i y_i A_i) (the dual box of the interior-point
 * method is not part of it).  All values are in the units of the data given
 * to the solver.  Callers: differentiation through the solve
 * (hdsdp_amd.adjoint), which needs a smooth map of (b, C). */

/* Barrier parameter and dual iterate of the last HDSDPOptimize. */
hdsdp_retcode HDSDPGetBarrierPoint( hdsdp *HSolver, double *mu, double *y );

/* Damped Newton from y (in/out, S(y) positive definite) to y_mu, until
 * ||b - mu A(S^-1)|| <= tol (1 + ||b||) or maxIter steps (FAILED then).
 * Leaves the dual factor and Schur matrix of the final y on the device. */
hdsdp_retcode HDSDPCenter( hdsdp *HSolver, double mu, double *y, int maxIter, double tol,
                           int *nIter, double *relRes );

/* Gradients of a loss L(y_mu) from dLdy at the centred y: with
 * M_ij = <A_i, S^-1 A_j S^-1>, M lambda = dLdy, dLdb = lambda / mu and
 * dLdC = S^-1 (sum_i lambda_i A_i) S^-1 (n x n, full, column-major).  Reuses
 * the factor and Schur matrix that HDSDPCenter left for this (mu, y). */
hdsdp_retcode HDSDPAdjoint( hdsdp *HSolver, double mu, const double *y, const double *dLdy,
                            double *dLdb, double *dLdC );

#ifdef __cplusplus
}
#endif

#endif /* HDSDP_AMD_HDSDP_H */
