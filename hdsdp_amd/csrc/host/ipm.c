/*
 * ipm.c -- dual-scaling interior-point driver (interface/hdsdp_algo.c).
 *
 * Phase A: infeasible-start dual method with adaptive residual reduction and
 * the infeasible corrector (interface/hdsdp_algo.c:959-1203, 666-957).
 * Phase B: dual potential reduction with barrier choice, ray detection and the
 * feasible corrector (interface/hdsdp_algo.c:1205-1846).
 * Self-dual embedding: run when phase A stops with suspected infeasibility
 * (interface/hdsdp_algo.c:262-545, 1859-1875).
 * The host keeps every control decision of the reference; each matrix
 * operation is one call into the SDP cone (cone_sdp.c -> device ABI).
 *
 * Implicit state the reference relies on, preserved here: the dual factor and
 * the bound-cone slacks always describe the point of the last interior check,
 * and dObjInternal is refreshed only when a log line is printed.
 */
#include "hdsdp_host.h"

#include <sys/time.h>

#define ALGO_DUAL_INFEAS    (0)
#define ALGO_DUAL_HSD       (1)
#define ALGO_DUAL_POTENTIAL (2)
#define ALGO_PRIMAL         (3)

double hdsdp_timestamp( void ) {
    struct timeval t;
    gettimeofday(&t, NULL);
    return 1e-06 * t.tv_usec + t.tv_sec;
}

/* Sum of an m-vector product, kept as a plain loop so both the device and
 * the CPU builds of the host see the same host arithmetic. */
static double vdot( int n, const double *x, const double *y ) {
    double s = 0.0;
    for ( int i = 0; i < n; ++i ) s += x[i] * y[i];
    return s;
}

static void set_start( hdsdp *s ) {
    memset(s->dRowDual, 0, sizeof(double) * s->nRows);
    s->dBarHsdTau = 1.0;
    double objFroNorm = get_dbl_feature(s, DBL_FEATURE_OBJFRONORM) * get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
    objFroNorm = HMAX(objFroNorm, 100.0);
    s->dResidual = -objFroNorm * get_dbl_param(s, DBL_PARAM_DUALSTART);
    s->pInfeas = 1.0 + get_dbl_feature(s, DBL_FEATURE_RHSFRONORM);
    s->pObjInternal = get_dbl_param(s, DBL_PARAM_POBJSTART);
    s->dBarrierMu = s->pObjInternal - s->dObjInternal - s->dResidual * get_dbl_param(s, DBL_PARAM_TRXESTIMATE);
    s->dBarrierMu = s->dBarrierMu / s->dAllConeDims;
    hlog(s, "Initialize with dual residual %3.1e\n", -s->dResidual);
    s->sdp->dualResidual = s->dResidual;
}

static void reset_start( hdsdp *s ) {
    memset(s->dRowDual, 0, sizeof(double) * s->nRows);
    s->dBarHsdTau = 1.0;
    double objFroNorm = get_dbl_feature(s, DBL_FEATURE_OBJFRONORM);
    s->pObjInternal = 1e+15;
    s->dResidual = -HMAX(objFroNorm, 1e+02);
    s->dResidual = s->dResidual * 1e+06;
    s->dResidual = HMAX(s->dResidual, -1e+15);
    hlog(s, "Reset with dual residual %3.1e\n", -s->dResidual);
    s->sdp->dualResidual = s->dResidual;
}

static void print_header( hdsdp *s, int method ) {
    if ( method == ALGO_DUAL_HSD ) {
        hlog(s, "HDSDP starts. Using self-dual method \n\n");
        hlog(s, "    %5s  %15s  %15s  %8s  %8s  %5s  %6s   %5s \n",
             "nIter", "pObj", "dObj", "dInf", "Mu", "Step", "Tau", "T [H]");
    } else if ( method == ALGO_DUAL_INFEAS ) {
        hlog(s, "HDSDP starts. Using infeasible dual method \n\n");
        hlog(s, "    %5s  %15s  %15s  %8s  %8s  %5s  %6s   %5s \n",
             "nIter", "pObj", "dObj", "dInf", "Mu", "Step", "|P|", "T [D]");
    } else {
        hlog(s, "HDSDP re-starts. Using feasible dual method \n\n");
        hlog(s, "    %5s  %15s  %15s  %8s  %8s  %5s  %6s   %5s \n",
             "nIter", "pObj", "dObj", "pInf", "Mu", "Step", "|P|", "T [P]");
    }
}

void hdsdp_record_log( hdsdp *s, int phase ) {
    if ( !s->iterLog || s->nLog >= HDSDP_MAX_LOG ) return;
    double *r = s->iterLog + (size_t) s->nLog * HDSDP_LOG_FIELDS;
    r[0] = phase;
    r[1] = s->nIterCount + 1;
    r[2] = ( phase == ALGO_DUAL_HSD ) ? HDSDP_INFINITY : s->pObjVal;
    r[3] = s->dObjVal;
    r[4] = ( phase == ALGO_DUAL_POTENTIAL || phase == ALGO_PRIMAL ) ? s->pInfeas : s->dInfeas;
    r[5] = s->dBarrierMu;
    r[6] = s->dDStep;
    r[7] = ( phase == ALGO_DUAL_HSD ) ? s->dBarHsdTau : s->dProxNorm;
    r[8] = hdsdp_timestamp() - s->dTimeBegin;
    s->nLog += 1;
}

/* interface/hdsdp_algo.c:151-193 */
static void print_log( hdsdp *s, int method ) {
    double rhsScal = get_dbl_feature(s, DBL_FEATURE_RHSSCALING);
    double objScal = get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
    double elapsed = hdsdp_timestamp() - s->dTimeBegin;
    double nSumCones = (double) get_int_feature(s, INT_FEATURE_N_SUMCONEDIMS);
    double pdObjScal = 1.0 / (rhsScal * objScal * s->dBarHsdTau);
    s->dInfeas = sqrt(nSumCones) * fabs(s->dResidual) / (rhsScal * s->dBarHsdTau);
    s->dObjInternal = vdot(s->nRows, s->rowRHS, s->dRowDual);
    s->dObjVal = s->dObjInternal * pdObjScal;
    s->pObjVal = s->pObjInternal * pdObjScal;
    s->comp = s->pObjVal - s->dObjVal;
    if ( method == ALGO_DUAL_HSD ) {
        hlog(s, "    %5d  %+15.8e  %+15.8e  %8.2e  %8.2e  %5.2f  %5.1e  %4.1f\n", s->nIterCount + 1,
             HDSDP_INFINITY, s->dObjVal, s->dInfeas, s->dBarrierMu, s->dDStep, s->dBarHsdTau, elapsed);
    } else {
        hlog(s, "    %5d  %+15.8e  %+15.8e  %8.2e  %8.2e  %5.2f  %5.1e  %4.1f \n", s->nIterCount + 1,
             s->pObjVal, s->dObjVal, method == ALGO_DUAL_POTENTIAL ? s->pInfeas : s->dInfeas,
             s->dBarrierMu, s->dDStep, s->dProxNorm, elapsed);
    }
    hdsdp_record_log(s, method);
}

/* interface/hdsdp_algo.c:195-215 */
static int check_interior( hdsdp *s, const double *y ) {
    int isInterior = 0;
    if ( sdp_cone_interior(s->sdp, 1.0, y, &isInterior) != HDSDP_RETCODE_OK ) return 0;
    if ( !isInterior ) return 0;
    if ( s->whichMethod != ALGO_DUAL_HSD ) {
        isInterior = bnd_cone_interior(s->bnd, 1.0, y);
    }
    return isInterior;
}

static hdsdp_retcode log_barrier( hdsdp *s, double tau, const double *y, int whichBuffer, double *val ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    double logdet = 0.0, barrier = 0.0;
    HCALL(sdp_cone_barrier(s->sdp, tau, y, whichBuffer, &logdet));
    barrier -= logdet;
    if ( s->whichMethod != ALGO_DUAL_HSD ) {
        HCALL(bnd_cone_barrier(s->bnd, tau, y, whichBuffer, &logdet));
        barrier -= logdet;
    }
    *val = barrier;
exit_cleanup:
    return retcode;
}

static void set_residual( hdsdp *s, double resi ) {
    s->sdp->dualResidual = resi;
}

static hdsdp_retcode ratio_sdp_bnd( hdsdp *s, const double *dy, double ada, int whichBuffer,
                                    int withBound, double *maxStep ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    double step = 0.0, best = HDSDP_INFINITY;
    HCALL(sdp_cone_ratio_test(s->sdp, 0.0, dy, ada, whichBuffer, &step));
    best = HMIN(best, step);
    if ( withBound ) {
        bnd_cone_ratio_test(s->bnd, 0.0, dy, whichBuffer, &step);
        best = HMIN(best, step);
    }
    *maxStep = best;
exit_cleanup:
    return retcode;
}

/* interface/hdsdp_algo.c:547-664 */
static int prox_measure( hdsdp *s ) {
    int isPFeasible = 0;
    int m = s->nRows;
    double dTraceSinv = s->kkt->dTraceSinv;
    double pObjNew = s->dObjInternal;
    double acc = get_dbl_param(s, DBL_PARAM_PRECORDACC);
    double *a1 = s->dHAuxiVec1, *a2 = s->dHAuxiVec2;

    for ( int i = 0; i < m; ++i ) {
        a1[i] = s->dMinvRowRHS[i] / s->dBarrierMu - s->dMinvASinv[i];
        a2[i] = s->rowRHS[i] / s->dBarrierMu - s->kkt->dASinvVec[i];
    }
    s->dProxNorm = vdot(m, a1, a2);
    if ( s->dProxNorm < 0.0 ) {
        s->dProxNorm = 1.0;
        return 0;
    }
    s->dProxNorm = sqrt(s->dProxNorm);
    for ( int i = 0; i < m; ++i ) a2[i] = a1[i] - s->dRowDual[i];

    isPFeasible = bnd_cone_interior_expert(s->bnd, 1.0, 1.0, a2, BUFFER_DUALCHECK);
    if ( isPFeasible ) {
        if ( sdp_cone_interior_expert(s->sdp, 1.0, 1.0, a2, -s->dResidual, BUFFER_DUALCHECK,
                                      &isPFeasible) != HDSDP_RETCODE_OK ) {
            isPFeasible = 0;
        }
    }
    if ( isPFeasible ) {
        double relGap = 0.0;
        if ( s->whichMethod == ALGO_DUAL_INFEAS ) {
            for ( int i = 0; i < m; ++i ) {
                relGap += a1[i] * (s->kkt->dASinvRdSinvVec[i] + s->kkt->dASinvVec[i]);
            }
            relGap += dTraceSinv * s->dResidual;
        } else {
            for ( int i = 0; i < m; ++i ) relGap += a1[i] * s->kkt->dASinvVec[i];
        }
        relGap += s->dAllConeDims;
        pObjNew += relGap * s->dBarrierMu;
        if ( relGap < 0 ) {
            isPFeasible = ( relGap < -1.0 ) ? -1 : 0;
        } else {
            s->pObjInternal = pObjNew;
            /* the first bound is overwritten by the second, as in the reference */
            double inAcc = HMIN(acc * 1000, 5e-03);
            inAcc = HMAX(acc, 1e-04);
            bnd_cone_primal(s->bnd, s->dBarrierMu, s->dRowDual, a1, s->dPInfeasLower, s->dPInfeasUpper);
            double pInfeas = 0.0;
            for ( int i = 0; i < m; ++i ) {
                pInfeas = HMAX(pInfeas, fabs(s->dPInfeasUpper[i] - s->dPInfeasLower[i]));
            }
            s->pInfeas = ( pInfeas < 1e-16 ) ? 0.0 : pInfeas;
            if ( pInfeas < 1.0 ) {
                if ( relGap * s->dBarrierMu > inAcc * (fabs(s->dObjInternal) + 1.0) ) {
                    s->dInaccBarrierMaker = s->dBarrierMu;
                    memcpy(s->dInaccRowDualMaker, s->dRowDual, sizeof(double) * m);
                    memcpy(s->dInaccRowDualStepMaker, a1, sizeof(double) * m);
                } else if ( relGap * s->dBarrierMu > acc * (fabs(s->dObjInternal) + 1.0) ) {
                    s->dAccBarrierMaker = s->dBarrierMu;
                    memcpy(s->dAccRowDualMaker, s->dRowDual, sizeof(double) * m);
                    memcpy(s->dAccRowDualStepMaker, a1, sizeof(double) * m);
                }
            }
        }
    }
    return isPFeasible;
}

/* interface/hdsdp_algo.c:666-738 */
static hdsdp_retcode adaptive_residual( hdsdp *s, double *resiRate ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    double alphaC = 0.0, alphaInf = 0.0, ada = 0.0, maxStep = HDSDP_INFINITY;
    for ( int i = 0; i < s->nRows; ++i ) s->dHAuxiVec1[i] = -s->dMinvASinv[i];
    HCALL(ratio_sdp_bnd(s, s->dHAuxiVec1, 0.0, BUFFER_DUALVAR, 1, &maxStep));
    alphaC = HMIN(0.98 * maxStep, 1.0);
    maxStep = alphaC;

    int isInterior = 0;
    while ( !isInterior && alphaC > 1e-02 * maxStep ) {
        HCALL(sdp_cone_axpy_check(s->sdp, alphaC, BUFFER_DUALCHECK, &isInterior));
        /* the bound cone's answer replaces the SDP cone's, as in the reference */
        isInterior = bnd_cone_axpy_check(s->bnd, alphaC, BUFFER_DUALCHECK);
        if ( !isInterior ) alphaC *= 0.8;
    }

    HCALL(ratio_sdp_bnd(s, s->dMinvASinvRdSinv, 1.0, BUFFER_DUALCHECK, 0, &maxStep));
    alphaInf = maxStep;
    ada = alphaInf / alphaC;
    ada = HMIN(0.98 * ada, 1.0);
    if ( s->dProxNorm < 1.0 ) {
        ada = HMAX(0.9, ada);
    } else if ( s->dProxNorm < 10.0 ) {
        ada = HMAX(0.3, ada);
    } else if ( s->dProxNorm < 50.0 ) {
        ada = HMAX(0.1, ada);
    }
    *resiRate = ada;
exit_cleanup:
    return retcode;
}

static hdsdp_retcode ratio_test( hdsdp *s, double ada, double *maxDist ) {
    return ratio_sdp_bnd(s, s->dRowDualStep, ada, BUFFER_DUALVAR, 1, maxDist);
}

/* interface/hdsdp_algo.c:776-957 */
static hdsdp_retcode infeasible_corrector( hdsdp *s ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    int m = s->nRows;
    int nMaxCorr = get_int_param(s, INT_PARAM_CORRECTORA);
    double barrier = 0.0, newBarrier = 0.0, maxStep = 0.0;
    double alphaC = 0.0, alphaInf = 0.0, ada = 0.0, adaMax = 0.8;

    if ( !check_interior(s, s->dRowDual) ) {
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    HCALL(log_barrier(s, 0.0, NULL, BUFFER_DUALVAR, &barrier));

    for ( int nCorr = 0; nCorr < nMaxCorr; ++nCorr ) {
        if ( !s->dResidual ) break;
        HCALL(kkt_build(s, KKT_TYPE_CORRECTOR));
        HCALL(kkt_build_bound(s, KKT_TYPE_CORRECTOR));
        memcpy(s->dMinvASinv, s->kkt->dASinvVec, sizeof(double) * m);
        memcpy(s->dMinvASinvRdSinv, s->kkt->dASinvRdSinvVec, sizeof(double) * m);
        HCALL(kkt_solve(s, s->dMinvASinv, s->dMinvASinv));
        if ( adaMax ) {
            HCALL(kkt_solve(s, s->dMinvASinvRdSinv, s->dMinvASinvRdSinv));
        }
        for ( int i = 0; i < m; ++i ) s->dRowDualStep[i] = -s->dMinvASinv[i];

        ratio_test(s, 0.0, &maxStep);  /* unchecked, as in the reference */
        maxStep = HMIN(0.8 * maxStep, 1.0);
        int isInterior = 0;
        while ( 1 ) {
            for ( int i = 0; i < m; ++i ) s->dHAuxiVec1[i] = s->dRowDual[i] + maxStep * s->dRowDualStep[i];
            isInterior = check_interior(s, s->dHAuxiVec1);
            if ( !isInterior ) maxStep *= 0.5;
            if ( isInterior || maxStep < 5e-03 ) break;
        }
        if ( maxStep < 5e-03 ) {
            isInterior = check_interior(s, s->dRowDual);
            assert(isInterior);
            break;
        }
        HCALL(log_barrier(s, 0.0, NULL, BUFFER_DUALVAR, &newBarrier));
        if ( newBarrier > barrier ) {
            maxStep *= 0.5;
            for ( int i = 0; i < m; ++i ) s->dHAuxiVec1[i] = s->dRowDual[i] + maxStep * s->dRowDualStep[i];
            isInterior = check_interior(s, s->dHAuxiVec1);
            barrier = -HDSDP_INFINITY;
            assert(isInterior);
        }
        alphaC = maxStep;

        double step = 0.0;
        maxStep = HDSDP_INFINITY;
        HCALL(sdp_cone_ratio_test(s->sdp, 0.0, s->dMinvASinvRdSinv, 1.0, BUFFER_DUALVAR, &step));
        maxStep = HMIN(maxStep, step);
        bnd_cone_ratio_test(s->bnd, 0.0, s->dMinvASinvRdSinv, BUFFER_DUALVAR, &step);
        maxStep = HMIN(maxStep, step);
        alphaInf = maxStep;
        ada = alphaInf / alphaC;
        ada = HMIN(1.0, adaMax * ada);

        double resi = s->dResidual;
        for ( int iTry = 0; ; ++iTry ) {
            s->dResidual = resi * (1 - alphaC * ada);
            set_residual(s, resi * (1 - alphaC * ada));
            for ( int i = 0; i < m; ++i ) {
                s->dHAuxiVec1[i] = s->dRowDual[i] + alphaC * (ada * s->dMinvASinvRdSinv[i] - s->dMinvASinv[i]);
            }
            if ( check_interior(s, s->dHAuxiVec1) ) break;
            ada *= 0.8;
        }
        if ( alphaC * ada < 5e-04 ) {
            adaMax = 0.0;
        } else if ( alphaC * ada < 0.1 ) {
            adaMax = adaMax * 0.9;
        }
        if ( alphaC * ada > 0.8 ) {
            s->dBarrierMu *= 0.8;
            adaMax = HMIN(adaMax * 2.0, 0.9);
        } else if ( alphaC * ada > 0.3 ) {
            s->dBarrierMu *= 0.95;
            adaMax = HMIN(adaMax * 2.0, 0.8);
        }
        memcpy(s->dRowDual, s->dHAuxiVec1, sizeof(double) * m);
        if ( adaMax == 0.0 ) break;
        barrier = newBarrier;
    }
exit_cleanup:
    return retcode;
}

static int time_up( hdsdp *s ) {
    return hdsdp_timestamp() - s->dTimeBegin >= get_dbl_param(s, DBL_PARAM_TIMELIMIT);
}

/* interface/hdsdp_algo.c:959-1203 */
static hdsdp_retcode phase_a( hdsdp *s ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    int m = s->nRows;
    s->whichMethod = ALGO_DUAL_INFEAS;
    int nMaxIter = get_int_param(s, INT_PARAM_MAXITER);
    int allowReset = 1;
    if ( get_int_feature(s, INT_FEATURE_I_MANYCONES) || get_int_feature(s, INT_FEATURE_I_IMPTRACE) ||
         get_int_feature(s, INT_FEATURE_I_VERYDENSE) ) {
        allowReset = 0;
    }
    double nSumCones = (double) get_int_feature(s, INT_FEATURE_N_SUMCONEDIMS);
    double dFeasTol = HMAX(get_dbl_param(s, DBL_PARAM_ABSFEASTOL),
                           get_dbl_param(s, DBL_PARAM_RELFEASTOL) * (1 + get_dbl_feature(s, DBL_FEATURE_OBJONENORM)));
    dFeasTol = dFeasTol * get_dbl_feature(s, DBL_FEATURE_OBJSCALING) / sqrt(nSumCones);
    double ada = 0.0;

    set_start(s);
    int isInterior = 0, pObjFound = 0, pObjType = 0;
    HCALL(sdp_cone_interior(s->sdp, s->dBarHsdTau, s->dRowDual, &isInterior));
    if ( isInterior ) isInterior = bnd_cone_interior(s->bnd, s->dBarHsdTau, s->dRowDual);
    if ( !isInterior ) {
        hlog(s, "Initial point is not in the cone. Adding slack value.\n");
        reset_start(s);
    }
    print_header(s, ALGO_DUAL_INFEAS);

    while ( 1 ) {
        if ( s->nIterCount == 3 && !pObjFound && allowReset ) {
            hlog(s, "Increasing dual infeasibility \n");
            reset_start(s);
            HCALL(sdp_cone_interior(s->sdp, s->dBarHsdTau, s->dRowDual, &isInterior));
            if ( !isInterior ) {
                retcode = HDSDP_RETCODE_FAILED;
                goto exit_cleanup;
            }
        }
        HCALL(kkt_build(s, KKT_TYPE_INFEASIBLE));
        HCALL(kkt_build_bound(s, KKT_TYPE_INFEASIBLE));
        memcpy(s->dMinvASinv, s->kkt->dASinvVec, sizeof(double) * m);
        memcpy(s->dMinvASinvRdSinv, s->kkt->dASinvRdSinvVec, sizeof(double) * m);
        HCALL(kkt_factorize(s));
        HCALL(kkt_solve(s, s->rowRHS, s->dMinvRowRHS));
        HCALL(kkt_solve(s, s->dMinvASinv, s->dMinvASinv));
        HCALL(kkt_solve(s, s->dMinvASinvRdSinv, s->dMinvASinvRdSinv));

        pObjType = prox_measure(s);
        if ( pObjType < 0 ) {
            s->HStatus = HDSDP_SUSPECT_INFEAS_OR_UNBOUNDED;
        } else {
            pObjFound += pObjType;
        }
        if ( pObjType == 1 && s->dProxNorm < 2.0 ) s->dBarrierMu *= 0.7;

        double target = s->pObjInternal - s->dObjInternal;
        target = target - s->dResidual * get_dbl_param(s, DBL_PARAM_TRXESTIMATE);
        target = target / (5.0 * s->dAllConeDims);
        if ( s->dProxNorm < 1.0 ) {
            s->dBarrierMu = s->dBarrierMu * 0.005;
        } else if ( s->dProxNorm < 5.0 ) {
            s->dBarrierMu = s->dBarrierMu * 0.01;
            s->dBarrierMu = HMAX(s->dBarrierMu, target * 0.1);
        } else if ( s->dProxNorm < 10.0 ) {
            s->dBarrierMu = s->dBarrierMu * 0.1;
            s->dBarrierMu = HMAX(s->dBarrierMu, target * 0.8);
        } else {
            s->dBarrierMu = s->dBarrierMu * 0.95;
            s->dBarrierMu = HMAX(s->dBarrierMu, target);
        }

        HCALL(adaptive_residual(s, &ada));
        for ( int i = 0; i < m; ++i ) {
            s->dRowDualStep[i] = s->dMinvRowRHS[i] / s->dBarrierMu - s->dMinvASinv[i] +
                                 ada * s->dMinvASinvRdSinv[i];
        }
        HCALL(ratio_test(s, ada, &s->dDStep));
        s->dDStep = HMIN(0.95 * s->dDStep, 1.0);
        if ( s->dDStep < 1e-03 ) s->nSmallStep += 1;

        for ( int i = 0; i < m; ++i ) s->dRowDual[i] += s->dDStep * s->dRowDualStep[i];
        s->dResidual = s->dResidual * (1.0 - ada * s->dDStep);
        set_residual(s, s->dResidual);

        HCALL(infeasible_corrector(s));
        print_log(s, ALGO_DUAL_INFEAS);

        if ( fabs(s->dResidual) < dFeasTol ) {
            s->HStatus = HDSDP_DUAL_FEASIBLE;
            break;
        }
        if ( s->nSmallStep > 3 ) {
            s->HStatus = HDSDP_SUSPECT_INFEAS_OR_UNBOUNDED;
            break;
        }
        if ( s->HStatus == HDSDP_SUSPECT_INFEAS_OR_UNBOUNDED ) break;
        if ( time_up(s) ) {
            s->HStatus = HDSDP_TIMELIMIT;
            break;
        }
        s->nIterCount += 1;
        if ( s->nIterCount >= nMaxIter ) {
            s->HStatus = HDSDP_MAXITER;
            break;
        }
    }
exit_cleanup:
    return retcode;
}

/* interface/hdsdp_algo.c:1234-1331 */
static hdsdp_retcode choose_barrier( hdsdp *s, int pObjType ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    int m = s->nRows;
    double rho = get_dbl_param(s, DBL_PARAM_POTRHOVAL);
    double absGap = s->pObjInternal - s->dObjInternal;
    double upper = absGap / s->dAllConeDims;
    double lower = upper / rho;
    double maxStep = HDSDP_INFINITY, pStep = 0.0, step = 0.0;

    if ( pObjType > 0 ) {
        for ( int i = 0; i < m; ++i ) s->dHAuxiVec1[i] = -s->dMinvRowRHS[i] / s->dBarrierMu;
        HCALL(sdp_cone_ratio_test(s->sdp, 0.0, s->dHAuxiVec1, 0.0, BUFFER_DUALCHECK, &step));
        maxStep = HMIN(maxStep, step);
        bnd_cone_ratio_test(s->bnd, 0.0, s->dHAuxiVec1, BUFFER_DUALCHECK, &step);
        maxStep = HMIN(maxStep, step);
        maxStep = HMIN(maxStep * 0.97, 1e+05);
        s->dBarrierMu = s->dBarrierMu / (1.0 + maxStep);
    } else {
        for ( int i = 0; i < m; ++i ) {
            s->dHAuxiVec2[i] = -s->dMinvRowRHS[i] / s->dBarrierMu + s->dMinvASinv[i];
        }
        HCALL(sdp_cone_ratio_test(s->sdp, 0.0, s->dHAuxiVec2, 0.0, BUFFER_DUALVAR, &step));
        maxStep = HMIN(maxStep, step);
        bnd_cone_ratio_test(s->bnd, 0.0, s->dHAuxiVec2, BUFFER_DUALVAR, &step);
        pStep = HMIN(maxStep, step);
        if ( pStep < 1.0 ) pStep = 0.97 * pStep;

        int isInterior = 0;
        for ( int iTry = 0; ; ++iTry ) {
            HCALL(sdp_cone_axpy_check(s->sdp, pStep, BUFFER_DUALCHECK, &isInterior));
            isInterior = bnd_cone_axpy_check(s->bnd, pStep, BUFFER_DUALCHECK);
            if ( !isInterior ) {
                pStep = ( iTry > 2 ) ? pStep * 0.97 : pStep * 0.5;
            } else {
                break;
            }
            if ( pStep < 1e-05 ) {
                retcode = HDSDP_RETCODE_FAILED;
                goto exit_cleanup;
            }
        }
        for ( int i = 0; i < m; ++i ) s->dHAuxiVec1[i] = -pStep * s->dMinvRowRHS[i] / s->dBarrierMu;
        HCALL(sdp_cone_ratio_test(s->sdp, 0.0, s->dHAuxiVec1, 0.0, BUFFER_DUALCHECK, &step));
        maxStep = HMIN(maxStep, step);
        bnd_cone_ratio_test(s->bnd, 0.0, s->dHAuxiVec1, BUFFER_DUALCHECK, &step);
        maxStep = HMIN(maxStep, step);
        maxStep = HMIN(maxStep * 0.97, 1e+05);
        s->dBarrierMu = pStep * s->dBarrierMu / (1.0 + maxStep) +
                        (1.0 - pStep) * (s->pObjInternal - s->dObjInternal) / s->dAllConeDims;
    }
    s->dBarrierMu = HMAX(s->dBarrierMu, lower);
    s->dBarrierMu = HMIN(s->dBarrierMu, upper);
exit_cleanup:
    return retcode;
}

/* interface/hdsdp_algo.c:1333-1363 */
static void feasible_build_step( hdsdp *s ) {
    int m = s->nRows;
    for ( int i = 0; i < m; ++i ) {
        s->dRowDualStep[i] = s->dMinvRowRHS[i] / s->dBarrierMu - s->dMinvASinv[i];
    }
    double prox = 0.0;
    for ( int i = 0; i < m; ++i ) {
        s->dHAuxiVec1[i] = s->rowRHS[i] / s->dBarrierMu - s->kkt->dASinvVec[i];
        prox += s->dHAuxiVec1[i] * s->dRowDualStep[i];
    }
    if ( prox < 0.0 ) {
        s->dProxNorm = 1e+02;
        return;
    }
    s->dProxNorm = sqrt(prox);
    if ( s->dProxNorm < 0.1 ) {
        s->dBarrierMu = 0.1 * s->dBarrierMu;
        feasible_build_step(s);
    }
}

static double get_potential( hdsdp *s, double rho, const double *y, int whichBuffer ) {
    double pot = 0.0, logdet = 0.0;
    sdp_cone_barrier(s->sdp, 0.0, NULL, whichBuffer, &logdet);
    pot -= logdet;
    bnd_cone_barrier(s->bnd, 0.0, NULL, whichBuffer, &logdet);
    pot -= logdet;
    pot += rho * log(s->pObjInternal - vdot(s->nRows, s->rowRHS, y));
    return pot;
}

/* interface/hdsdp_algo.c:1388-1455 */
static hdsdp_retcode reduce_potential( hdsdp *s ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    int m = s->nRows;
    double maxStep = 0.0, step = 0.0, potNow = 0.0, potNew = 0.0, reqDecrease = 0.0;
    double rho = (s->pObjVal - s->dObjVal) / s->dBarrierMu;
    double minStepTol = 0.0;
    if ( s->nSmallStep >= 2 ) minStepTol = 0.5;
    if ( s->dProxNorm < 0.5 ) reqDecrease = 0.05;

    HCALL(ratio_test(s, 0.0, &maxStep));
    step = HMIN(maxStep * 0.95, 1.0);
    potNow = get_potential(s, rho, s->dRowDual, BUFFER_DUALVAR);
    for ( int iTry = 0; ; ++iTry ) {
        for ( int i = 0; i < m; ++i ) s->dHAuxiVec1[i] = s->dRowDual[i] + step * s->dRowDualStep[i];
        if ( !check_interior(s, s->dHAuxiVec1) ) {
            step *= 0.33;
            continue;
        }
        potNew = get_potential(s, rho, s->dHAuxiVec1, BUFFER_DUALVAR);
        if ( potNew <= potNow - reqDecrease || step * s->dProxNorm <= 0.001 || step < minStepTol ) {
            memcpy(s->dRowDual, s->dHAuxiVec1, sizeof(double) * m);
            break;
        }
        if ( step < 1e-04 ) {
            if ( !check_interior(s, s->dRowDual) ) retcode = HDSDP_RETCODE_FAILED;
            break;
        }
        step *= 0.3;
    }
    s->dDStep = step;
    s->dPotentialVal = potNew;
exit_cleanup:
    return retcode;
}

static double get_barrier( hdsdp *s, const double *y, int whichBuffer ) {
    double val = 0.0, logdet = 0.0;
    sdp_cone_barrier(s->sdp, 0.0, NULL, whichBuffer, &logdet);
    val += logdet;
    bnd_cone_barrier(s->bnd, 0.0, NULL, whichBuffer, &logdet);
    val += logdet;
    val = vdot(s->nRows, s->rowRHS, y) + s->dBarrierMu * val;
    return -val;
}

/* interface/hdsdp_algo.c:1480-1613 */
static hdsdp_retcode feasible_corrector( hdsdp *s ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    int m = s->nRows;
    double shrink = s->dAllConeDims / (s->dAllConeDims + sqrt(s->dAllConeDims));
    int nMaxCorr = get_int_param(s, INT_PARAM_CORRECTORB);
    if ( nMaxCorr == 0 ) return retcode;
    if ( s->dProxNorm < 0.1 ) nMaxCorr = 0;
    if ( s->dDStep < 1e-02 ) nMaxCorr = 0;
    if ( s->dDStep < 0.1 && s->dBarrierMu < 1e-05 ) {
        nMaxCorr = 0;
        set_int_param(s, INT_PARAM_CORRECTORB, 0);
    }
    if ( s->dDStep < 1e-03 ) {
        nMaxCorr = 0;
        set_int_param(s, INT_PARAM_CORRECTORB, 0);
    }
    if ( s->dBarrierMu < 1e-06 ) {
        nMaxCorr = 0;
        set_int_param(s, INT_PARAM_CORRECTORB, 0);
    }
    double barNow = 0.0, barNew = 0.0;
    double bDotMinvB = vdot(m, s->dMinvRowRHS, s->rowRHS);
    double bDotMinvASinv = 0.0, bDotCorr = 0.0;

    for ( int nCorr = 0; nCorr < nMaxCorr; ++nCorr ) {
        if ( s->dBarrierMu < 1e-05 ) break;
        HCALL(kkt_build(s, KKT_TYPE_CORRECTOR));
        HCALL(kkt_build_bound(s, KKT_TYPE_CORRECTOR));
        memcpy(s->dMinvASinv, s->kkt->dASinvVec, sizeof(double) * m);
        HCALL(kkt_solve(s, s->dMinvASinv, s->dMinvASinv));
        bDotMinvASinv = vdot(m, s->rowRHS, s->dMinvASinv);
        if ( bDotMinvASinv > 0 && bDotMinvB > 0 ) s->dBarrierMu = bDotMinvB / bDotMinvASinv;
        s->dBarrierMu *= shrink;
        for ( int i = 0; i < m; ++i ) {
            s->dHAuxiVec1[i] = s->dMinvRowRHS[i] / s->dBarrierMu - s->dMinvASinv[i];
            bDotCorr += s->rowRHS[i] * s->dHAuxiVec1[i];  /* accumulates across rounds, as in the reference */
        }
        double maxStep = HDSDP_INFINITY, step = 0.0;
        barNow = get_barrier(s, s->dRowDual, BUFFER_DUALVAR);
        HCALL(sdp_cone_ratio_test(s->sdp, 0.0, s->dHAuxiVec1, 0.0, BUFFER_DUALVAR, &step));
        maxStep = HMIN(maxStep, step);
        bnd_cone_ratio_test(s->bnd, 0.0, s->dHAuxiVec1, BUFFER_DUALVAR, &step);
        maxStep = HMIN(maxStep, step);
        maxStep = HMIN(maxStep * 0.95, maxStep);
        maxStep = HMIN(maxStep, get_dbl_param(s, DBL_PARAM_POTRHOVAL) / s->dProxNorm);

        for ( int iTry = 0; ; ++iTry ) {
            for ( int i = 0; i < m; ++i ) s->dHAuxiVec2[i] = s->dRowDual[i] + maxStep * s->dHAuxiVec1[i];
            if ( !check_interior(s, s->dHAuxiVec2) ) {
                maxStep *= 0.5;
                continue;
            }
            barNew = get_barrier(s, s->dHAuxiVec2, BUFFER_DUALVAR);
            if ( maxStep < 1e-04 || barNew <= barNow - fabs(0.05 * bDotCorr * maxStep) ) {
                break;
            } else {
                double t = 2 * (barNew - barNow + bDotCorr * maxStep) / (maxStep * maxStep);
                if ( (bDotCorr / t < maxStep) && (bDotCorr / t > 0) ) {
                    maxStep = bDotCorr / t;
                } else {
                    maxStep *= 0.5;
                }
            }
        }
        if ( maxStep < 1e-04 ) {
            int isInterior = check_interior(s, s->dRowDual);
            assert(isInterior);
            (void) isInterior;
            break;
        } else {
            memcpy(s->dRowDual, s->dHAuxiVec2, sizeof(double) * m);
        }
    }
exit_cleanup:
    return retcode;
}

/* interface/hdsdp_algo.c:1615-1655 */
static int primal_infeas_check( hdsdp *s, int forceDetect ) {
    int isRay = 0, isInterior = 0;
    if ( s->pInfeas >= get_dbl_feature(s, DBL_FEATURE_RHSFRONORM) || forceDetect ||
         (s->pInfeas > 0.01 * get_dbl_feature(s, DBL_FEATURE_RHSONENORM) && s->dBarrierMu < 1e-03) ) {
        if ( s->dObjVal < 0.0 ) return isRay;
        double nrm = sqrt(vdot(s->nRows, s->dRowDual, s->dRowDual));
        for ( int i = 0; i < s->nRows; ++i ) s->dHAuxiVec1[i] = s->dRowDual[i] / nrm;
        if ( sdp_cone_interior_expert(s->sdp, 0.0, -1.0, s->dHAuxiVec1, 1e-08, BUFFER_DUALCHECK,
                                      &isInterior) != HDSDP_RETCODE_OK ) {
            isInterior = 0;
        }
        if ( isInterior ) isRay = 1;
    }
    return isRay;
}

/* The primal method, HPSDPInit + HPSDPOptimize (interface/hdsdp_psdp.c:78-295), entered
 * from the dual method once its gap is below 10% (interface/hdsdp_algo.c:1803-1810): from
 * the recovered primal X, at most five primal-dual steps on the Schur matrix of W = X
 * (built and factored once), each solving M dy = A(X S X) - mu A X, the dual step length
 * by the dual ratio test on dS = -sum dy_i A_i, the primal direction
 * dX = X - (X S X + Xs dS Xs) / mu with Xs the starting X, the primal step by the primal
 * Lanczos solver, both steps at 95% capped at 1; mu = min(mu, <X, S> / (5 n^2)).  The
 * loop test reads the dual method's last complementarity (interface/hdsdp_psdp.c:
 * 287-289), which these steps do not update. */
static hdsdp_retcode psdp_optimize( hdsdp *s ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    const int m = s->nRows, n = s->sdp->nCol;
    hdsdp_dev *dev = s->sdp->dev;
    const double relOpt = get_dbl_param(s, DBL_PARAM_RELOPTTOL), absOpt = get_dbl_param(s, DBL_PARAM_ABSOPTTOL);
    const double pdScal = get_dbl_feature(s, DBL_FEATURE_OBJSCALING) * get_dbl_feature(s, DBL_FEATURE_RHSSCALING);
    double *AX = NULL, *rhs = NULL, mu = s->dBarrierMu;
    int ok = 0;
    /* one SDP cone and no dual infeasibility (interface/hdsdp_psdp.c:85-90) */
    if ( s->dInfeas > 0.0 ) {
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    HALLOC(AX, double, m);
    HALLOC(rhs, double, m);
    /* X from the dual method's recovery (HDSDPGetConeValues, interface/hdsdp_psdp.c:131) */
    HCALL(hdsdp_recover_primal(s, &ok));
    if ( !ok ) {
        hlog(s, "Recovery step is infeasible\n");
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    HCALL(hdsdp_dev_psdp_init(dev));
    hlog(s, "Primal solver is initialized. \n");
    HCALL(kkt_build(s, KKT_TYPE_PRIMAL));
    hlog(s, "KKT Build up: %f s", hdsdp_timestamp() - s->dTimeBegin);
    HCALL(kkt_factorize(s));
    hlog(s, "KKT Factorize: %f s", hdsdp_timestamp() - s->dTimeBegin);
    for ( int it = 0; it < 5; ++it ) {
        double pInf = 0.0, maxD = 0.0, maxP = 0.0, pObj = 0.0, XdotS = 0.0;
        int isPsd = 0, isInterior = 0;
        HCALL(hdsdp_dev_psdp_rhs(dev, mu, AX, rhs));
        for ( int i = 0; i < m; ++i ) pInf += (s->rowRHS[i] - AX[i]) * (s->rowRHS[i] - AX[i]);
        pInf = sqrt(pInf);
        HCALL(kkt_solve(s, rhs, s->dRowDualStep));
        HCALL(sdp_cone_ratio_test(s->sdp, 0.0, s->dRowDualStep, 1.0, BUFFER_DUALVAR, &maxD));
        HCALL(hdsdp_dev_psdp_direction(dev, mu, it == 0, &maxP));
        maxP = HMIN(0.95 * maxP, 1.0);
        maxD = HMIN(0.95 * maxD, 1.0);
        HCALL(hdsdp_dev_psdp_step(dev, maxP, &isPsd, &pObj, &XdotS));
        for ( int i = 0; i < m; ++i ) s->dRowDual[i] += maxD * s->dRowDualStep[i];
        HCALL(sdp_cone_update(s->sdp, 1.0, s->dRowDual));
        HCALL(sdp_cone_interior(s->sdp, 1.0, s->dRowDual, &isInterior));
        if ( !isInterior || !isPsd ) {
            retcode = HDSDP_RETCODE_FAILED;
            goto exit_cleanup;
        }
        double dObj = vdot(m, s->rowRHS, s->dRowDual);
        mu = HMIN(mu, (XdotS / n) / (5.0 * n));
        s->pObjVal = pObj;
        s->dObjVal = dObj;
        s->pInfeas = pInf;
        s->dBarrierMu = mu;
        s->dDStep = maxD;
        s->nIterCount += 1;
        hlog(s, "    %5d  %+15.8e  %+15.8e  %8.2e  %8.2e  %5.2f %5.2f     %4.1f \n", s->nIterCount + 1, s->pObjVal,
             s->dObjVal, s->pInfeas, s->dBarrierMu, maxP, s->dDStep, hdsdp_timestamp() - s->dTimeBegin);
        hdsdp_record_log(s, ALGO_PRIMAL);
        if ( s->comp < (fabs(s->pObjVal) + fabs(s->dObjVal) + 1.0) * relOpt && s->comp < absOpt / pdScal ) break;
    }
exit_cleanup:
    free(AX);
    free(rhs);
    return retcode;
}

/* interface/hdsdp_algo.c:1657-1846 */
static hdsdp_retcode phase_b( hdsdp *s ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    int m = s->nRows;
    s->whichMethod = ALGO_DUAL_POTENTIAL;
    int nMaxIter = get_int_param(s, INT_PARAM_MAXITER);
    int iPrimalMethod = get_int_param(s, INT_PARAM_PSDP);
    double absFeas = get_dbl_param(s, DBL_PARAM_ABSFEASTOL);
    double absOpt = get_dbl_param(s, DBL_PARAM_ABSOPTTOL);
    double relFeas = get_dbl_param(s, DBL_PARAM_RELFEASTOL);
    double relOpt = get_dbl_param(s, DBL_PARAM_RELOPTTOL);
    double objOneNorm = get_dbl_feature(s, DBL_FEATURE_OBJONENORM);
    double objScal = get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
    double pdScal = objScal * get_dbl_feature(s, DBL_FEATURE_RHSSCALING);
    double nSumCones = (double) get_int_feature(s, INT_FEATURE_N_SUMCONEDIMS);
    int nIterInternal = 0;
    double dFeasTol = HMIN(absFeas, relFeas * (1.0 + objOneNorm));
    dFeasTol = dFeasTol * objScal / sqrt(nSumCones);

    /* PhaseB_Initialize (interface/hdsdp_algo.c:1205-1232) */
    if ( s->dBarHsdTau != 1.0 || s->HStatus != HDSDP_DUAL_FEASIBLE ) {
        s->dResidual = s->dResidual / s->dBarHsdTau;
        for ( int i = 0; i < m; ++i ) s->dRowDual[i] = s->dRowDual[i] / s->dBarHsdTau;
        if ( !check_interior(s, s->dRowDual) ) hlog(s, "Incumbent dual solution is infeasible \n");
        s->HStatus = HDSDP_INTERNAL_ERROR;
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    if ( fabs(s->dResidual) > dFeasTol ) {
        hlog(s, "Dual infeasibility from previous algorithm exceeds user-specified tolerance \n");
    }
    s->dPerturb = -10.0 * s->dResidual;
    s->dResidual = 0.0;
    s->sdp->dualResidual = 0.0;
    assert(s->dPerturb >= 0.0);
    s->sdp->dualPerturb = s->dPerturb;
    if ( s->dPerturb != 0.0 ) check_interior(s, s->dRowDual);

    int pObjType = 0, nopObjFound = 0, isRayFound = 0, forceDetect = 1;
    double pObjStart = s->pObjInternal;
    print_header(s, ALGO_DUAL_POTENTIAL);

    while ( 1 ) {
        nIterInternal += 1;
        if ( nIterInternal > 10 ) forceDetect = 0;
        HCALL(kkt_build(s, KKT_TYPE_INFEASIBLE));
        HCALL(kkt_build_bound(s, KKT_TYPE_INFEASIBLE));
        if ( s->dBarrierMu > 1.0 ) {
            HCALL(kkt_regularize(s, 1e-06));
        }
        memcpy(s->dMinvASinv, s->kkt->dASinvVec, sizeof(double) * m);
        memcpy(s->dMinvASinvRdSinv, s->kkt->dASinvRdSinvVec, sizeof(double) * m);
        HCALL(kkt_factorize(s));
        HCALL(kkt_solve(s, s->rowRHS, s->dMinvRowRHS));
        HCALL(kkt_solve(s, s->dMinvASinv, s->dMinvASinv));

        pObjType = prox_measure(s);
        if ( pObjType < 0 ) {
            s->HStatus = HDSDP_SUSPECT_INFEAS_OR_UNBOUNDED;
        } else {
            if ( pObjType ) nopObjFound = 0;
            else nopObjFound += 1;
        }
        HCALL(choose_barrier(s, pObjType));
        feasible_build_step(s);
        isRayFound = primal_infeas_check(s, forceDetect);
        if ( isRayFound ) {
            hlog(s, "HDSDP detects a dual improving ray \n");
            s->HStatus = HDSDP_INFEAS_OR_UNBOUNDED;
            break;
        }
        HCALL(reduce_potential(s));
        if ( s->dDStep < 1e-03 ) s->nSmallStep += 1;
        HCALL(feasible_corrector(s));
        print_log(s, ALGO_DUAL_POTENTIAL);

        if ( s->comp < (fabs(s->pObjVal) + fabs(s->dObjVal) + 1.0) * relOpt && s->comp < absOpt / pdScal ) {
            s->HStatus = HDSDP_PRIMAL_DUAL_OPTIMAL;
            break;
        }
        if ( s->comp < (fabs(s->pObjVal) + fabs(s->dObjVal) + 1.0) * 0.1 && iPrimalMethod ) {
            retcode = psdp_optimize(s);
            break;
        }
        if ( s->nSmallStep > 3 ) {
            s->HStatus = HDSDP_NUMERICAL;
            break;
        }
        if ( s->HStatus == HDSDP_SUSPECT_INFEAS_OR_UNBOUNDED ) break;
        if ( time_up(s) ) {
            s->HStatus = HDSDP_TIMELIMIT;
            break;
        }
        s->nIterCount += 1;
        if ( s->nIterCount >= nMaxIter ) {
            s->HStatus = HDSDP_MAXITER;
            break;
        }
        if ( nopObjFound >= 10 && s->pObjInternal != pObjStart ) {
            s->HStatus = HDSDP_NUMERICAL;
            break;
        }
    }
exit_cleanup:
    return retcode;
}

/* interface/hdsdp_algo.c:73-82: the self-dual start (used when the embedding runs first) */
static void set_start_hsd( hdsdp *s, int dOnly ) {
    memset(s->dRowDual, 0, sizeof(double) * s->nRows);
    s->dBarHsdTau = 1.0;
    double objFroNorm = get_dbl_feature(s, DBL_FEATURE_OBJFRONORM) * get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
    objFroNorm = HMAX(objFroNorm, 100.0);
    s->dBarrierMu = 1e+08;
    s->dResidual = dOnly ? -objFroNorm * get_dbl_param(s, DBL_PARAM_DUALSTART) : -objFroNorm * 1e+01;
    hlog(s, "Initialize with dual residual %3.1e\n", -s->dResidual);
    s->sdp->dualResidual = s->dResidual;
}

/* interface/hdsdp_algo.c:262-313 (HDSDP_HSD_BuildStep): the tau step from the bordered
 * system and dy = d1 (tau + dtau) / mu + d4 dtau - d2 + d3, with d1 = M^-1 b,
 * d2 = M^-1 A S^-1, d3 = M^-1 A S^-1 Rd S^-1, d4 = M^-1 A S^-1 C S^-1 */
static void hsd_build_step( hdsdp *s ) {
    int m = s->nRows;
    const double *b = s->rowRHS, *d1 = s->dMinvRowRHS, *d2 = s->dMinvASinv, *d3 = s->dMinvASinvRdSinv,
                 *d4 = s->dMinvASinvCSinv, *aCS = s->kkt->dASinvCSinvVec;
    double *dd1 = s->dHAuxiVec1, *dy = s->dRowDualStep;
    double mu = s->dBarrierMu, tau = s->dBarHsdTau;
    s->dBarHsdTauStep = 0.0;
    double oldObj = s->dObjInternal;
    s->dObjInternal = vdot(m, b, s->dRowDual);
    s->dObjImprove = s->dObjInternal - oldObj;
    for ( int i = 0; i < m; ++i ) dd1[i] = b[i] - mu * aCS[i];
    double num = -s->dObjInternal + mu / tau + mu * (s->kkt->dCSinv - s->kkt->dCSinvRdSinv);
    double den = mu * s->kkt->dCSinvCSinv + mu / (tau * tau);
    double tauOverMu = tau / mu;
    for ( int i = 0; i < m; ++i ) {
        num -= dd1[i] * (d1[i] * tauOverMu - d2[i] + d3[i]);
        den += dd1[i] * (d1[i] / mu + d4[i]);
    }
    double dtau = ( fabs(den) < 1e-12 ) ? 0.0 : num / den;
    s->dBarHsdTauStep = dtau;
    for ( int i = 0; i < m; ++i ) dy[i] = d1[i] * (tau + dtau) / mu + d4[i] * dtau - d2[i] + d3[i];
}

/* interface/hdsdp_algo.c:315-352 (HDSDP_HSD_RatioTest): tau and the SDP cone, no bound cone */
static hdsdp_retcode hsd_ratio_test( hdsdp *s, double *maxDist ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    double maxStep = HDSDP_INFINITY, step = s->dBarHsdTau / s->dBarHsdTauStep;
    if ( step < 0.0 ) maxStep = HMIN(maxStep, -step);
    HCALL(sdp_cone_ratio_test(s->sdp, s->dBarHsdTauStep, s->dRowDualStep, 1.0, BUFFER_DUALVAR, &step));
    maxStep = HMIN(maxStep, step);
    if ( maxStep < 1e-02 ) {
        s->nSmallStep += 1;
        if ( s->nSmallStep > 2 ) hlog(s, "HDSDP stagates at the cone boundary. \n");
    }
exit_cleanup:
    *maxDist = maxStep;
    return retcode;
}

/* interface/hdsdp_algo.c:354-545 (HDSDP_PhaseA_BarHsdSolve): the self-dual embedding
 * (serves as a certificate of dual (in)feasibility; no primal solution is extracted) */
static hdsdp_retcode phase_hsd( hdsdp *s, int dOnly ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    int m = s->nRows;
    s->whichMethod = ALGO_DUAL_HSD;
    int nMaxIter = get_int_param(s, INT_PARAM_MAXITER);
    double gamma = get_dbl_param(s, DBL_PARAM_HSDGAMMA);
    double absFeas = get_dbl_param(s, DBL_PARAM_ABSFEASTOL), absOpt = get_dbl_param(s, DBL_PARAM_ABSOPTTOL);
    double relFeas = get_dbl_param(s, DBL_PARAM_RELFEASTOL), relOpt = get_dbl_param(s, DBL_PARAM_RELOPTTOL);
    double timeLimit = get_dbl_param(s, DBL_PARAM_TIMELIMIT);
    double objOneNorm = get_dbl_feature(s, DBL_FEATURE_OBJONENORM);
    double objScal = get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
    double nSumCones = (double) get_int_feature(s, INT_FEATURE_N_SUMCONEDIMS);
    if ( !dOnly ) {
        absOpt = 1e+20;
        relOpt = 1e+20;
    }
    double feasTol = HMIN(absFeas, relFeas * (1.0 + objOneNorm));
    feasTol = feasTol * objScal / sqrt(nSumCones);
    absOpt = absOpt * 1e-04;
    relOpt = absOpt * 1e-04;  /* from the absolute tolerance, as interface/hdsdp_algo.c:401 sets it */
    if ( s->HStatus == HDSDP_UNKNOWN ) set_start_hsd(s, dOnly);
    print_header(s, ALGO_DUAL_HSD);

    while ( 1 ) {
        int isInterior = 0;
        HCALL(sdp_cone_interior(s->sdp, s->dBarHsdTau, s->dRowDual, &isInterior));
        if ( !isInterior ) {
            if ( s->nIterCount == 0 ) {
                hlog(s, "Initial point is not in the cone. Adding slack value.\n");
                s->dResidual *= 100.0;
                reset_start(s);
                s->nIterCount += 1;
                continue;
            }
            hlog(s, "Iteration %d is not in the cone. \n", s->nIterCount);
            retcode = HDSDP_RETCODE_FAILED;
            goto exit_cleanup;
        }
        /* HKKTBuildUp over the SDP cone (no bound cone in the embedding); the
         * HKKTRegularize(.., 0.0) that follows adds nothing */
        HCALL(kkt_build(s, KKT_TYPE_HOMOGENEOUS));
        memcpy(s->dMinvASinv, s->kkt->dASinvVec, sizeof(double) * m);
        memcpy(s->dMinvASinvRdSinv, s->kkt->dASinvRdSinvVec, sizeof(double) * m);
        memcpy(s->dMinvASinvCSinv, s->kkt->dASinvCSinvVec, sizeof(double) * m);
        HCALL(kkt_factorize(s));
        HCALL(kkt_solve(s, s->rowRHS, s->dMinvRowRHS));
        HCALL(kkt_solve(s, s->dMinvASinv, s->dMinvASinv));
        HCALL(kkt_solve(s, s->dMinvASinvRdSinv, s->dMinvASinvRdSinv));
        HCALL(kkt_solve(s, s->dMinvASinvCSinv, s->dMinvASinvCSinv));
        hsd_build_step(s);
        HCALL(hsd_ratio_test(s, &s->dDStep));
        if ( s->dDStep > 1.0 ) {
            s->dDStep = HMIN(0.7 * s->dDStep, 1.0);
        } else if ( s->dDStep > 0.5 ) {
            s->dDStep = HMIN(0.5 * s->dDStep, 1.0);
        } else if ( s->dDStep > 0.2 ) {
            s->dDStep = HMIN(0.3 * s->dDStep, 1.0);
        } else {
            s->dDStep = HMIN(0.2 * s->dDStep, 1.0);
        }
        print_log(s, ALGO_DUAL_HSD);
        s->dBarHsdTau += s->dDStep * s->dBarHsdTauStep;
        for ( int i = 0; i < m; ++i ) s->dRowDual[i] += s->dDStep * s->dRowDualStep[i];
        s->dResidual = s->dResidual * (1.0 - s->dDStep);
        set_residual(s, s->dResidual);
        double target = 0.0;
        if ( s->dBarrierMu > 1e-12 ) {
            if ( s->dDStep > 0.8 && s->dBarHsdTau > 1.0 ) {
                target = 0.1 * s->dBarrierMu;
            } else {
                target = gamma * s->dBarrierMu;
            }
            target = HMAX(target, -0.1 * s->dResidual / s->dBarHsdTau);
            s->dBarrierMu = HMIN(s->dBarrierMu, target);
        } else {
            target = 0.8 * s->dBarrierMu;
            s->dBarrierMu = HMIN(s->dBarrierMu, target);
        }
        if ( fabs(s->dResidual) < feasTol * s->dBarHsdTau && s->dBarrierMu < absOpt &&
             (s->dBarrierMu < relOpt * (1 + 2.0 * fabs(s->dObjVal)) &&
              fabs(s->dObjImprove) < 1e-05 * (fabs(s->dObjInternal) + 1.0)) ) {
            s->HStatus = dOnly ? HDSDP_DUAL_OPTIMAL : HDSDP_DUAL_FEASIBLE;
            break;
        }
        if ( s->dBarHsdTau <= 1e-10 ) {
            s->HStatus = HDSDP_SUSPECT_INFEAS_OR_UNBOUNDED;
            break;
        }
        if ( hdsdp_timestamp() - s->dTimeBegin >= timeLimit ) {
            s->HStatus = HDSDP_TIMELIMIT;
            break;
        }
        s->nIterCount += 1;
        if ( s->nIterCount >= nMaxIter ) {
            s->HStatus = HDSDP_MAXITER;
            break;
        }
    }
exit_cleanup:
    if ( retcode != HDSDP_RETCODE_OK ) s->HStatus = HDSDP_NUMERICAL;
    return retcode;
}

/* interface/hdsdp_algo.c:1859-1900 */
hdsdp_retcode hdsdp_solve( hdsdp *s, int dOnly ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    HCALL(phase_a(s));
    if ( s->HStatus == HDSDP_SUSPECT_INFEAS_OR_UNBOUNDED ) {
        hlog(s, "\nInfeasible method stops due to suspected infeasibility \n");
        HCALL(phase_hsd(s, dOnly));
    } else if ( s->HStatus == HDSDP_DUAL_FEASIBLE ) {
        hlog(s, "\nInfeasible method finds a dual feasible solution \n");
        HCALL(phase_b(s));
    }
exit_cleanup:
    return retcode;
}
