/*
 * hdsdp_gpu.h -- C ABI of the MI355X dual-scaling hot path.
 *
 * One `hdsdp_dev` holds one dense SDP block resident in HBM: its coefficient
 * matrices, the three dual buffers S (DUALVAR), trial S (DUALCHECK) and dS
 * (DUALSTEP), their Cholesky factors, S^-1, the Schur matrix M and its
 * preconditioner.  Only m-vectors and scalars cross the host boundary.
 *
 * Each entry point replaces one slot of the reference's plugin tables; the
 * citation names the slot and the implementation it stands in for.  Plain
 * pointers and sizes only.  Buffer ids are the reference's
 * (interface/hdsdp_conic.h:24-26): 0 = DUALVAR, 1 = DUALCHECK, 2 = DUALSTEP.
 * Schur types (interface/hdsdp_conic.h:16-19): 0 = INFEASIBLE, 1 = CORRECTOR,
 * 2 = HOMOGENEOUS, 3 = PRIMAL.  All host arrays are FP64.
 *
 * Errors: HDSDP_RETCODE_FAILED for any device error or unsupported request;
 * a matrix that is not positive definite is not an error (isPsd = 0), as in
 * linalg/hdsdp_linsolver.c:535-545.
 */
#ifndef HDSDP_AMD_HDSDP_GPU_H
#define HDSDP_AMD_HDSDP_GPU_H

#include "hdsdp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Coefficient matrix kinds after presolve (linalg/def_hdsdp_sdpdata.h:25-33). */
#define HDSDP_COEFF_ZERO   0
#define HDSDP_COEFF_SPARSE 1
#define HDSDP_COEFF_DENSE  2
#define HDSDP_COEFF_SPR1   3
#define HDSDP_COEFF_DSR1   4

/* One presolved coefficient matrix (linalg/def_hdsdp_sdpdata.h:78-141).
 *   SPARSE: nnz lower triplets (row >= col, 0-based) in row/col/val.
 *   DENSE : val = packed-lower column-major n(n+1)/2 entries.
 *   SPR1  : A = sign a a^T, a has nnz entries at rows `row`, values `val`.
 *   DSR1  : A = sign a a^T, val = dense a (n entries).
 *   ZERO  : nothing. */
typedef struct {
    int type;
    int nnz;
    const int *row;
    const int *col;
    const double *val;
    double sign;
} hdsdp_coeff_view;

typedef struct hdsdp_dev hdsdp_dev;

/* ---- lifecycle ---------------------------------------------------------- */

/* Allocate device state for an n x n block with m constraints.  precision:
 * HDSDP_PRECISION_FP64 or HDSDP_PRECISION_MIXED (include/hdsdp.h).
 * Replaces sdpDenseConeCreateImpl + the dense-direct HFpLinsysCreate calls
 * (interface/hdsdp_conic_sdp.c:1319-1333, 100-111) and HKKTIAllocDenseKKT
 * (interface/hdsdp_schur.c:11-44). */
hdsdp_retcode hdsdp_dev_create( hdsdp_dev **pdev, int nRow, int nCol, int precision );

/* Upload m + 1 presolved coefficients (coeffs[0] = C, coeffs[i + 1] = A_i).
 * Replaces the data half of sdpDenseConeProcDataImpl/PresolveImpl
 * (interface/hdsdp_conic_sdp.c:1352-1514). */
hdsdp_retcode hdsdp_dev_load( hdsdp_dev *dev, const hdsdp_coeff_view *coeffs );

/* Rescale C in place (coneScal slot: interface/hdsdp_conic_sdp.c:1600-1604). */
hdsdp_retcode hdsdp_dev_scale_obj( hdsdp_dev *dev, double dScal );

void hdsdp_dev_destroy( hdsdp_dev **pdev );

/* ---- dual buffers (hdsdp_cone slots for DENSE_SDP) ---------------------- */

/* target = dACoefScal * sum_i dACoef[i] A_i + dCCoef * C + dEyeCoef * I.
 * The caller folds the cone perturbation into dEyeCoef.
 * Replaces sdpDenseConeIUpdateBuffer (interface/hdsdp_conic_sdp.c:343-402). */
hdsdp_retcode hdsdp_dev_update_buffer( hdsdp_dev *dev, double dCCoef, double dACoefScal,
                                       const double *dACoef, double dEyeCoef, int whichBuffer );

/* Cholesky of buffer DUALVAR (into the dual factor) or DUALCHECK (into the
 * checker factor); *isPsd = 1 iff positive definite.  Replaces
 * HFpLinsysPsdCheck on lapack_flinsys (linalg/hdsdp_linsolver.c:525-560). */
hdsdp_retcode hdsdp_dev_factorize( hdsdp_dev *dev, int whichBuffer, int *isPsd );

/* DUALCHECK: checker = S + dStep * dS;  DUALVAR: S += dStep * dS; then factorize.
 * Replaces sdpDenseConeAddStepToBufferAndCheck (interface/hdsdp_conic_sdp.c:2266-2295). */
hdsdp_retcode hdsdp_dev_axpy_check( hdsdp_dev *dev, double dStep, int whichBuffer, int *isPsd );

/* 2 * sum log L_ii of the factor of `whichBuffer`.  Replaces the factor half of
 * sdpDenseConeGetBarrier (interface/hdsdp_conic_sdp.c:2185-2224). */
hdsdp_retcode hdsdp_dev_logdet( hdsdp_dev *dev, int whichBuffer, double *logdet );

/* Largest step a with S_w + a dS psd from the current DUALSTEP buffer and the
 * factor of `whichBuffer`: Lanczos on L^-1 (-dS) L^-T (coneRatioTest slot;
 * interface/hdsdp_conic_sdp.c:1638-1682, linalg/hdsdp_lanczos.c:161-292). */
hdsdp_retcode hdsdp_dev_ratio_test( hdsdp_dev *dev, int whichBuffer, double *maxStep );

/* ---- Schur complement (coneBuildSchur slot) ----------------------------- */

/* Build with W = S^-1 from the dual factor (typeKKT INFEASIBLE, CORRECTOR) or
 * W = X, the primal method's iterate (PRIMAL; every constraint SPARSE or SPR1,
 * as the reference's strategy M5 requires, after hdsdp_dev_psdp_init):
 *   M_ij = <A_i, W A_j W> (lower triangle, not for CORRECTOR),
 *   dASinv[i] = <A_i, W>, dASinvRdSinv[i] = Rd <A_i, W W>,
 *   *dTraceSinv = tr(W) when Rd != 0 (not for CORRECTOR), else 0.
 * HOMOGENEOUS (2) builds the same; the self-dual embedding's extra terms then come
 * from hdsdp_dev_hsd.
 * Replaces sdpDenseConeGetKKT and its strategies M2-M5
 * (interface/hdsdp_conic_sdp.c:683-1052, 1722-1872). */
hdsdp_retcode hdsdp_dev_schur( hdsdp_dev *dev, int typeKKT, double dualResidual,
                               double *dASinv, double *dASinvRdSinv, double *dTraceSinv );

/* The self-dual embedding's terms of a HOMOGENEOUS build, from its W = S^-1 (call
 * after hdsdp_dev_schur with typeKKT 0 or 2, before the dual factor changes): with
 * B = W C W, dASinvCSinv[i] = <A_i, B> (m entries), *dCSinv = <C, W>,
 * *dCSinvCSinv = <C, B>, *dCSinvRdSinv = dualResidual tr(B).
 * Replaces sdpDenseConeIGetHSDComponents (interface/hdsdp_conic_sdp.c:983-1029) and
 * the dASinvCSinvVec entries of strategies M1-M5 (interface/hdsdp_conic_sdp.c:737-760
 * and the alike blocks of M2-M5).  HDSDP_RETCODE_FAILED when the last build was PRIMAL. */
hdsdp_retcode hdsdp_dev_hsd( hdsdp_dev *dev, double dualResidual, double *dASinvCSinv,
                             double *dCSinv, double *dCSinvCSinv, double *dCSinvRdSinv );

/* ---- Schur solve (hdsdp_linsys_fp slots of the KKT solver) -------------- */

/* CG parameters (conjGradLinSolverSetParam, linalg/hdsdp_linsolver.c:735-760). */
hdsdp_retcode hdsdp_dev_kkt_params( hdsdp_dev *dev, double relTol, double absTol,
                                    int maxIter, int nRestartFreq );
/* M_ii += diag[i] (bound cone: interface/hdsdp_conic_bound.c:224-229). */
hdsdp_retcode hdsdp_dev_kkt_add_diag( hdsdp_dev *dev, const double *diag );
/* 1 once the Schur system has switched to the direct indefinite solver (a Cholesky
 * factorisation of M, or the CG preconditioned by it, failed; the switch is for good, as
 * HFpLinsysSwitchToIndefinite's, linalg/hdsdp_linsolver.c:1236-1266, 1367-1425), else 0.
 * hdsdp_dev_kkt_factorize and hdsdp_dev_kkt_solve make the switch themselves; the
 * host reads it to log it as the reference does. */
int hdsdp_dev_kkt_uses_ldl( hdsdp_dev *dev );
/* HKKTRegularize (interface/hdsdp_schur.c:348-373). */
hdsdp_retcode hdsdp_dev_kkt_regularize( hdsdp_dev *dev, double dKKTReg );
/* Preconditioner build: Jacobi diagonal or Cholesky of M once switched
 * (conjGradLinSolverNumeric, linalg/hdsdp_linsolver.c:768-833). */
hdsdp_retcode hdsdp_dev_kkt_factorize( hdsdp_dev *dev );
/* PCG on M with the reference's restart and Jacobi-to-Cholesky switch
 * (conjGradSolve, linalg/hdsdp_linsolver.c:855-997).  rhs and sol may alias. */
hdsdp_retcode hdsdp_dev_kkt_solve( hdsdp_dev *dev, const double *rhs, double *sol );

/* Copy the Schur matrix M (m x m, full symmetric, bound-cone diagonal and
 * regularization included) to the host; the matrix the CG works on
 * (interface/hdsdp_schur.c:293-326 exports only its vectors). */
hdsdp_retcode hdsdp_dev_get_schur( hdsdp_dev *dev, double *M );

/* Jacobi-preconditioned CG on a caller-held Schur matrix without any device
 * allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  M: m x m FP64, column-major, both
 * triangles; rhs, sol: device m-vectors (may alias); stream: a hipStream_t or
 * NULL.  The iteration is hdsdp_dev_kkt_solve's while it uses the Jacobi
 * preconditioner (same kernels, same order: the results are identical); the
 * CG vectors live in a static device array of the library (about 4 MiB per
 * GPU, held until the process ends), so m <= 65536, and calls that overlap on
 * one GPU race.  Returns HDSDP_RETCODE_FAILED for m outside 1..65536, and when
 * the Jacobi iteration would switch to the Cholesky preconditioner (that needs
 * m^2 device memory: use hdsdp_dev_kkt_solve).  A right-hand side below the
 * tolerance leaves sol as it was, as conjGradSolve does.  *nIter: iterations. */
hdsdp_retcode hdsdp_dev_kkt_solve_noalloc( const double *M, int m, const double *rhs, double *sol,
                                           double relTol, double absTol, int maxIter, int nRestartFreq,
                                           void *stream, int *nIter );

/* ---- primal recovery and solution checks -------------------------------- */

/* X = mu (S^-1 + S^-1 (sum dy_i A_i) S^-1) with S = C - A^T y (+ perturbation)
 * factored into the checker; *ok = 0 if S is not psd.  Replaces
 * sdpDenseConeGetPrimal (interface/hdsdp_conic_sdp.c:2328-2379). */
hdsdp_retcode hdsdp_dev_primal( hdsdp_dev *dev, double mu, const double *y, const double *dy, int *ok );
/* From the last recovered X: AX[i] = <A_i, X>, <C, X>, <X, S>, tr(X)
 * (coneATimesX / coneTraceCX, interface/hdsdp_conic_sdp.c:2453-2471). */
hdsdp_retcode hdsdp_dev_primal_stats( hdsdp_dev *dev, double *AX, double *traceCX,
                                      double *XdotS, double *traceX );
/* Copy the last X (n x n, full) or S (n x n, symmetrized) to the host. */
hdsdp_retcode hdsdp_dev_get_primal( hdsdp_dev *dev, double *X );
hdsdp_retcode hdsdp_dev_get_dual( hdsdp_dev *dev, double *S );

/* ---- primal method (interface/hdsdp_psdp.c) ------------------------------ */

/* Start the primal iterate at the X of the last hdsdp_dev_primal and keep a
 * copy of it as the fixed scaling Xs; the primal Lanczos solver starts afresh.
 * Replaces the device half of HPSDPInit (interface/hdsdp_psdp.c:78-140). */
hdsdp_retcode hdsdp_dev_psdp_init( hdsdp_dev *dev );
/* AX[i] = <A_i, X> and rhs[i] = <A_i, X S X> - mu AX[i] with S the current
 * dual matrix (DUALVAR), kept for the step (interface/hdsdp_psdp.c:180-206). */
hdsdp_retcode hdsdp_dev_psdp_rhs( hdsdp_dev *dev, double mu, double *AX, double *rhs );
/* After the dual ratio test left dS in DUALSTEP: dX = X - (X S X + Xs dS Xs) / mu;
 * on the first call X is factored (FAILED if not positive definite); then the
 * largest step a with X + a dX psd, by the primal Lanczos solver on
 * L_X^-1 (-dX) L_X^-T (interface/hdsdp_psdp.c:208-231, 58-66). */
hdsdp_retcode hdsdp_dev_psdp_direction( hdsdp_dev *dev, double mu, int first, double *maxStep );
/* X += alpha dX and X factored again (*isPsd); <C, X> and <X, S> with the dual
 * matrix of hdsdp_dev_psdp_rhs (interface/hdsdp_psdp.c:237-271). */
hdsdp_retcode hdsdp_dev_psdp_step( hdsdp_dev *dev, double alpha, int *isPsd, double *traceCX,
                                   double *XdotS );

/* ---- adjoint (no counterpart in the reference) -------------------------- */

/* Direct FP64 solve of M x = rhs with the current Schur matrix: FP64 Gram
 * blocks (rebuilt if they hold bf16 split products), FP64 Cholesky, two steps of
 * iterative refinement.  FAILED if M is not positive definite.
 * Callers: the centring Newton steps and the adjoint, which need solutions to
 * working accuracy rather than to the interior-point CG tolerance. */
hdsdp_retcode hdsdp_dev_kkt_solve_exact( hdsdp_dev *dev, const double *rhs, double *sol );

/* Differentiates the barrier optimality conditions b = mu A(S^-1),
 * S = C - sum_i y_i A_i, at the current dual factor and Schur matrix:
 * solves M lambda = dLdy exactly -- reusing the FP64 factor of the Schur
 * matrix last factored on the device (the final centring step's) as the
 * preconditioner of iterative refinement on the current M when it is close
 * enough, factoring M otherwise -- then dLdb = lambda / mu and
 * dLdC = S^-1 (sum_i lambda_i A_i) S^-1 (n x n full).  In the form
 * -mu S^-1 (sum_i l_i A_i) S^-1 this is l = -lambda / mu.
 * Callers: HDSDPAdjoint (autograd through the solver, hdsdp_amd.adjoint). */
hdsdp_retcode hdsdp_dev_adjoint( hdsdp_dev *dev, double mu, const double *dLdy,
                                 double *dLdb, double *dLdC );

/* ---- statistics ---------------------------------------------------------- */

/* Counters for measurement: [0] Schur builds, [1] Cholesky of S, [2] CG
 * iterations, [3] Lanczos matvecs, [4] Cholesky of M, [5] last CG iterations,
 * [6] adjoint solves that refined with the factor of an earlier M instead of
 * factoring again. */
hdsdp_retcode hdsdp_dev_stats( hdsdp_dev *dev, double *stats, int nStats );

/* Device-time profile, measured with events on the library's own stream.
 * Slot k accumulates milliseconds (ms[k]) and the number of timed regions
 * (counts[k]) of:
 *   HDSDP_PROF_SCHUR     whole Schur build (S^-1, traces, M)
 *   HDSDP_PROF_PAIRS     the pair-gather Schur kernel alone (one launch per build)
 *   HDSDP_PROF_MFACTOR   Cholesky of M and its triangular inverse
 *   HDSDP_PROF_CG        one preconditioned CG solve
 *   HDSDP_PROF_SFACTOR   Cholesky of a dual buffer
 *   HDSDP_PROF_LANCZOS   one Lanczos step-length call
 *   HDSDP_PROF_BUFFER    one dual-buffer assembly
 *   HDSDP_PROF_PRIMAL    primal recovery
 *   HDSDP_PROF_GRAM      the Gram products of a Schur build (dense: T_p = L^-1 A_p L^-T
 *                        and <T_p, T_q>; rank-one: W F and F^T W F), FP64 or bf16 split
 *   HDSDP_PROF_GRAM16    the same products of the builds made from bf16 split products
 *                        only (mixed precision; the builds made in FP64 are left out)
 * Profiling is off by default: enabling it adds one event synchronisation per
 * timed region.  No counterpart in the reference (measurement only). */
#define HDSDP_PROF_SCHUR   0
#define HDSDP_PROF_PAIRS   1
#define HDSDP_PROF_MFACTOR 2
#define HDSDP_PROF_CG      3
#define HDSDP_PROF_SFACTOR 4
#define HDSDP_PROF_LANCZOS 5
#define HDSDP_PROF_BUFFER  6
#define HDSDP_PROF_PRIMAL  7
#define HDSDP_PROF_GRAM    8
#define HDSDP_PROF_GRAM16  9
#define HDSDP_PROF_SLOTS   10
hdsdp_retcode hdsdp_dev_set_profiling( hdsdp_dev *dev, int enable );
int hdsdp_dev_get_profile( hdsdp_dev *dev, double *ms, double *counts, int nSlots );

/* Name of the implementation ("hip-gfx950" for the product library). */
const char *hdsdp_dev_backend_name( void );

#ifdef __cplusplus
}
#endif

#endif /* HDSDP_AMD_HDSDP_GPU_H */
