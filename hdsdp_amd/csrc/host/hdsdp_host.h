/*
 * hdsdp_host.h -- internal structures of the C host.
 *
 * The host keeps the reference's control logic: the public API
 * (interface/hdsdp.c), the dual-scaling IPM (interface/hdsdp_algo.c), presolve of
 * the coefficient matrices (linalg/hdsdp_sdpdata.c), the scalar bound cone
 * (interface/hdsdp_conic_bound.c) and the KKT bookkeeping
 * (interface/hdsdp_schur.c).  All O(n^2)-and-up arithmetic of the dense SDP
 * block goes through the hdsdp_dev_* ABI (include/hdsdp_gpu.h).
 */
#ifndef HDSDP_HOST_H
#define HDSDP_HOST_H

#include <assert.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hdsdp.h"
#include "hdsdp_gpu.h"

#define HMAX(x, y) ((x) > (y) ? (x) : (y))
#define HMIN(x, y) ((x) < (y) ? (x) : (y))
#define HCALL(expr) do { retcode = (expr); if (retcode != HDSDP_RETCODE_OK) goto exit_cleanup; } while (0)
#define HALLOC(var, type, n) do { (var) = (type *) calloc((size_t) (n) > 0 ? (size_t) (n) : 1, sizeof(type)); \
                                  if (!(var)) { retcode = HDSDP_RETCODE_MEMORY; goto exit_cleanup; } } while (0)
#define HFREE(var) do { free(var); (var) = NULL; } while (0)

/* Packed-lower column-major index of (i, j), i >= j (interface/hdsdp_utils.h:48-51). */
#define PK_IDX(n, i, j) ((int64_t) (2 * (int64_t) (n) - (j) - 1) * (j) / 2 + (i))

#define KKT_TYPE_INFEASIBLE  0
#define KKT_TYPE_CORRECTOR   1
#define KKT_TYPE_HOMOGENEOUS 2
#define KKT_TYPE_PRIMAL      3

#define BUFFER_DUALVAR   0
#define BUFFER_DUALCHECK 1
#define BUFFER_DUALSTEP  2

#define ABS_NORM 1
#define FRO_NORM 2

#define NUM_INT_PARAM 20
#define NUM_DBL_PARAM 20

/* Feature slots (interface/def_hdsdp.h:22-55). */
#define INT_FEATURE_I_NULLOBJ      0
#define INT_FEATURE_I_MANYCONES    1
#define INT_FEATURE_I_NOPINTERIOR  2
#define INT_FEATURE_I_NODINTERIOR  3
#define INT_FEATURE_I_VERYDENSE    4
#define INT_FEATURE_I_IMPTRACE     5
#define INT_FEATURE_I_IMPYBOUND    6
#define INT_FEATURE_N_SUMCONEDIMS  7
#define INT_FEATURE_N_MAXCONEDIM   8
#define INT_FEATURE_N_CONES        9
#define INT_FEATURE_N_ROWS         10
#define INT_FEATURE_N_SPSDPCONES   11
#define INT_FEATURE_N_DSSDPCONES   12
#define INT_FEATURE_N_LPCONES      13
#define INT_FEATURE_N_BNDCONES     14
#define INT_FEATURE_N_ZEORMATS     15
#define INT_FEATURE_N_SPMATS       16
#define INT_FEATURE_N_DSMATS       17
#define INT_FEATURE_N_SPR1MATS     18
#define INT_FEATURE_N_DSR1MATS     19

#define DBL_FEATURE_OBJFRONORM     0
#define DBL_FEATURE_OBJONENORM     1
#define DBL_FEATURE_RHSFRONORM     2
#define DBL_FEATURE_RHSONENORM     3
#define DBL_FEATURE_RHSINFNORM     4
#define DBL_FEATURE_OBJSCALING     5
#define DBL_FEATURE_RHSSCALING     6
#define DBL_FEATURE_DATAFRONORM    7
#define DBL_FEATURE_DATAONENORM    8
#define DBL_FEATURE_IMPYBOUNDUP    9
#define DBL_FEATURE_IMPYBOUNDLOW   10
#define DBL_FEATURE_IMPTRACEX      11

/* ---- user data (interface/def_hdsdp_user_data.h:22-31) ------------------ */
struct hdsdp_user_data {
    cone_type cone;
    int nConicRow;
    int nConicCol;
    int *coneMatBeg;
    int *coneMatIdx;
    double *coneMatElem;
    /* factor-form rank-one extension */
    int isFactorForm;
    int cNnz;
    int *cIdx;
    double *cElem;
    double *factor;
    double *factorSign;
};

/* ---- one presolved coefficient matrix (host copy) ----------------------- */
typedef struct {
    int type;          /* HDSDP_COEFF_* */
    int n;
    int nnz;           /* SPARSE: triplets; SPR1: factor nonzeros */
    int *row;          /* SPARSE rows / SPR1 indices */
    int *col;          /* SPARSE columns */
    double *val;       /* SPARSE values / DENSE packed / SPR1 values / DSR1 factor */
    double sign;       /* rank-one sign (scale folded in) */
    int ownsVal;       /* 0 when `val` aliases caller memory */
} sdp_coeff;

/* ---- dense SDP cone, host half (interface/def_hdsdp_conic.h:107-152) ----- */
typedef struct {
    int nRow;
    int nCol;
    double dualResidual;
    double dualPerturb;
    sdp_coeff *sdpObj;
    sdp_coeff *sdpRow;     /* nRow entries */
    int sdpConeStats[5];
    hdsdp_dev *dev;
} sdp_cone;

/* ---- scalar bound cone (interface/def_hdsdp_conic.h:229-244) ------------ */
typedef struct {
    int nRow;
    double *dualUpper, *dualLower;
    double *dualUpperInverse, *dualLowerInverse;
    double *dualUpperStep, *dualLowerStep;
    double *dualUpperChecker, *dualLowerChecker;
    double dBoundLow, dBoundUp;
} bound_cone;

/* ---- KKT vectors held on the host (interface/def_hdsdp_schur.h:32-68) --- */
typedef struct {
    int nRow;
    double *dASinvVec;
    double *dASinvRdSinvVec;
    double *dASinvCSinvVec;
    double *dBoundDiag;
    double dCSinvCSinv, dCSinvRdSinv, dCSinv, dTraceSinv;
} host_kkt;

#define HDSDP_MAX_LOG 4096

/* ---- solver (interface/def_hdsdp.h:57-135) ------------------------------- */
struct hdsdp_solver_internal {
    char modelFeatures[200];
    int nRows;
    double *rowRHS;
    int nCones;
    user_data **userCones;
    sdp_cone *sdp;
    bound_cone *bnd;
    host_kkt *kkt;
    int precision;
    int verbose;
    int profiling;

    double *dRowDual;
    double *dRowDualStep;
    double dBarHsdTau;
    double dBarHsdTauStep;
    double *dMinvRowRHS;
    double *dMinvASinv;
    double *dMinvASinvRdSinv;
    double *dMinvASinvCSinv;
    double *dHAuxiVec1;
    double *dHAuxiVec2;

    int whichMethod;
    int nIterCount;
    int nSmallStep;
    double dBarrierMu;
    double dProxNorm;
    double dPotentialVal;
    double dDStep;
    double dResidual;
    double dPerturb;
    double *dPInfeasUpper;
    double *dPInfeasLower;

    double dAllConeDims;
    double pObjVal, dObjVal;
    double pObjInternal, dObjInternal;
    double dObjImprove;
    double pInfeas, dInfeas;
    double comp;

    double *dAccRowDualMaker, *dAccRowDualStepMaker;
    double dAccBarrierMaker;
    double *dInaccRowDualMaker, *dInaccRowDualStepMaker;
    double dInaccBarrierMaker;

    double dTimeBegin;
    double dTimeEnd;
    hdsdp_status HStatus;

    /* centring / adjoint: the point (mu, y) whose dual factor and Schur matrix
     * the device holds (adjoint.c) */
    int adjReady;
    double adjMu;
    double *adjY;

    int HIntParams[NUM_INT_PARAM];
    double HDblParams[NUM_DBL_PARAM];
    int HIntFeatures[20];
    double HDblFeatures[20];
    double dErrs[6];

    int nLog;
    double *iterLog;
};

#define set_int_param(s, p, v) ((s)->HIntParams[p] = (v))
#define set_dbl_param(s, p, v) ((s)->HDblParams[p] = (v))
#define get_int_param(s, p) ((s)->HIntParams[p])
#define get_dbl_param(s, p) ((s)->HDblParams[p])
#define set_int_feature(s, f, v) ((s)->HIntFeatures[f] = (v))
#define set_dbl_feature(s, f, v) ((s)->HDblFeatures[f] = (v))
#define get_int_feature(s, f) ((s)->HIntFeatures[f])
#define get_dbl_feature(s, f) ((s)->HDblFeatures[f])

#define hlog(s, ...) do { if ((s)->verbose) { printf(__VA_ARGS__); } } while (0)

/* ---- host modules --------------------------------------------------------- */
double hdsdp_timestamp( void );

/* sdp_coeff.c */
hdsdp_retcode coeff_set_from_csc( sdp_coeff *A, int n, int nnz, const int *packedIdx, const double *elem );
hdsdp_retcode coeff_set_rank_one( sdp_coeff *A, int n, double sign, const double *factor );
hdsdp_retcode coeff_presolve_rank_one( sdp_coeff *A );
void coeff_scale( sdp_coeff *A, double alpha );
double coeff_norm( const sdp_coeff *A, int type );
int coeff_rank( const sdp_coeff *A );
int coeff_nnz( const sdp_coeff *A );
int coeff_is_eye( const sdp_coeff *A, double *mult );
int coeff_is_unit_col( const sdp_coeff *A, int *col );
void coeff_clear( sdp_coeff *A );
void coeff_view( const sdp_coeff *A, hdsdp_coeff_view *v );

/* cone_sdp.c */
hdsdp_retcode sdp_cone_create( sdp_cone **pcone, user_data *ud, int precision );
hdsdp_retcode sdp_cone_presolve( sdp_cone *cone );
void sdp_cone_destroy( sdp_cone **pcone );
double sdp_cone_obj_norm( sdp_cone *cone, int type );
double sdp_cone_coeff_norm( sdp_cone *cone, int type );
hdsdp_retcode sdp_cone_scal( sdp_cone *cone, double s );
void sdp_cone_features( sdp_cone *cone, const double *rowRHS, int iF[20], double dF[20] );
hdsdp_retcode sdp_cone_update( sdp_cone *cone, double tau, const double *y );
hdsdp_retcode sdp_cone_ratio_test( sdp_cone *cone, double dtau, const double *dy, double ada,
                                   int whichBuffer, double *maxStep );
hdsdp_retcode sdp_cone_schur( sdp_cone *cone, host_kkt *kkt, int typeKKT );
hdsdp_retcode sdp_cone_interior( sdp_cone *cone, double tau, const double *y, int *isInterior );
hdsdp_retcode sdp_cone_interior_expert( sdp_cone *cone, double cCoef, double aScal, const double *aCoef,
                                        double eye, int whichBuffer, int *isInterior );
hdsdp_retcode sdp_cone_barrier( sdp_cone *cone, double tau, const double *y, int whichBuffer, double *logdet );
hdsdp_retcode sdp_cone_axpy_check( sdp_cone *cone, double step, int whichBuffer, int *isInterior );
hdsdp_retcode sdp_cone_primal( sdp_cone *cone, double mu, const double *y, const double *dy, int *ok );

/* cone_bound.c */
hdsdp_retcode bnd_cone_create( bound_cone **pcone, int nRow, double lo, double up );
void bnd_cone_destroy( bound_cone **pcone );
void bnd_cone_update( bound_cone *c, double tau, const double *y );
void bnd_cone_ratio_test( bound_cone *c, double dtau, const double *dy, int whichBuffer, double *maxStep );
void bnd_cone_schur( bound_cone *c, host_kkt *kkt, int typeKKT );
int bnd_cone_interior( bound_cone *c, double tau, const double *y );
int bnd_cone_interior_expert( bound_cone *c, double cCoef, double aScal, const double *aCoef, int whichBuffer );
hdsdp_retcode bnd_cone_barrier( bound_cone *c, double tau, const double *y, int whichBuffer, double *logdet );
int bnd_cone_axpy_check( bound_cone *c, double step, int whichBuffer );
void bnd_cone_primal( bound_cone *c, double mu, const double *y, const double *dy, double *xl, double *xu );

/* kkt.c */
hdsdp_retcode kkt_create( host_kkt **pkkt, int nRow );
void kkt_destroy( host_kkt **pkkt );
hdsdp_retcode kkt_init( host_kkt *kkt, sdp_cone *cone );
hdsdp_retcode kkt_build( hdsdp *s, int typeKKT );
hdsdp_retcode kkt_build_bound( hdsdp *s, int typeKKT );
hdsdp_retcode kkt_factorize( hdsdp *s );
hdsdp_retcode kkt_solve( hdsdp *s, const double *rhs, double *sol );
hdsdp_retcode kkt_regularize( hdsdp *s, double reg );

/* ipm.c */
hdsdp_retcode hdsdp_solve( hdsdp *s, int dOnly );
void hdsdp_record_log( hdsdp *s, int phase );
/* X of the dual method's last accurate (else inaccurate) barrier point, left in the device's
 * primal buffer; *ok = 0 when the recovery step is infeasible (HDSDPGetConeValues,
 * interface/hdsdp.c:711-745) */
hdsdp_retcode hdsdp_recover_primal( hdsdp *s, int *ok );

#endif /* HDSDP_HOST_H */
