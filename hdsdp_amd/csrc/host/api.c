/*
 * api.c -- public solver API (interface/hdsdp.c, interface/hdsdp_user_data.c).
 *
 * Defaults, statistics, automatic parameter adjustment and the DIMACS solution
 * check follow interface/hdsdp.c:29-402 and 617-906.  One dense SDP cone plus
 * the internal scalar bound cone is supported; other cone layouts are refused
 * with HDSDP_RETCODE_FAILED.
 */
#include "hdsdp_host.h"
#include "eig_small.h"

/* ---- user data (interface/hdsdp_user_data.c:38-121) ---------------------- */

hdsdp_retcode HUserDataCreate( user_data **pHdata ) {
    if ( !pHdata ) return HDSDP_RETCODE_FAILED;
    user_data *ud = (user_data *) calloc(1, sizeof(user_data));
    if ( !ud ) return HDSDP_RETCODE_MEMORY;
    *pHdata = ud;
    return HDSDP_RETCODE_OK;
}

void HUserDataSetConeData( user_data *Hdata, cone_type cone, int nRow, int nCol,
                           int *coneMatBeg, int *coneMatIdx, double *coneMatElem ) {
    Hdata->cone = cone;
    Hdata->nConicRow = nRow;
    Hdata->nConicCol = nCol;
    Hdata->coneMatBeg = coneMatBeg;
    Hdata->coneMatIdx = coneMatIdx;
    Hdata->coneMatElem = coneMatElem;
    Hdata->isFactorForm = 0;
}

void HUserDataSetRankOneConeData( user_data *Hdata, int nRow, int nCol, int cNnz, int *cIdx, double *cElem,
                                  double *factor, double *sign ) {
    Hdata->cone = HDSDP_CONETYPE_DENSE_SDP;
    Hdata->nConicRow = nRow;
    Hdata->nConicCol = nCol;
    Hdata->isFactorForm = 1;
    Hdata->cNnz = cNnz;
    Hdata->cIdx = cIdx;
    Hdata->cElem = cElem;
    Hdata->factor = factor;
    Hdata->factorSign = sign;
}

/* interface/hdsdp_user_data.c:73-98 */
cone_type HUserDataChooseCone( user_data *Hdata ) {
    if ( Hdata->cone != HDSDP_CONETYPE_DENSE_SDP ) return Hdata->cone;
    if ( Hdata->isFactorForm ) return HDSDP_CONETYPE_DENSE_SDP;
    int nz = 0;
    for ( int i = 0; i < Hdata->nConicRow; ++i ) {
        nz += ( Hdata->coneMatBeg[i + 2] - Hdata->coneMatBeg[i + 1] > 0 );
    }
    return ( nz > 0.3 * Hdata->nConicRow ) ? HDSDP_CONETYPE_DENSE_SDP : HDSDP_CONETYPE_SPARSE_SDP;
}

void HUserDataClear( user_data *Hdata ) {
    if ( Hdata ) memset(Hdata, 0, sizeof(user_data));
}

void HUserDataDestroy( user_data **pHdata ) {
    if ( !pHdata ) return;
    free(*pHdata);
    *pHdata = NULL;
}

/* ---- parameters (interface/hdsdp.c:378-402) ------------------------------ */

static void default_params( hdsdp *s ) {
    set_int_param(s, INT_PARAM_MAXITER, 500);
    set_int_param(s, INT_PARAM_CORRECTORA, 12);
    set_int_param(s, INT_PARAM_CORRECTORB, 12);
    set_int_param(s, INT_PARAM_THREADS, 12);
    set_int_param(s, INT_PARAM_PSDP, 1);
    set_dbl_param(s, DBL_PARAM_ABSOPTTOL, 1e-08);
    set_dbl_param(s, DBL_PARAM_ABSFEASTOL, 1e-08);
    set_dbl_param(s, DBL_PARAM_RELOPTTOL, 1e-10);
    set_dbl_param(s, DBL_PARAM_RELFEASTOL, 1e-10);
    set_dbl_param(s, DBL_PARAM_TIMELIMIT, 3600.0);
    set_dbl_param(s, DBL_PARAM_POTRHOVAL, 4.0);
    set_dbl_param(s, DBL_PARAM_HSDGAMMA, 0.5);
    set_dbl_param(s, DBL_PARAM_DUALBOX_UP, 1e+07);
    set_dbl_param(s, DBL_PARAM_DUALBOX_LOW, -1e+07);
    set_dbl_param(s, DBL_PARAM_BARMUSTART, 1e+05);
    set_dbl_param(s, DBL_PARAM_POBJSTART, 1e+10);
    set_dbl_param(s, DBL_PARAM_DUALSTART, 1e+05);
    set_dbl_param(s, DBL_PARAM_TRXESTIMATE, 1e+08);
    set_dbl_param(s, DBL_PARAM_PRECORDACC, 1e-08);
}

/* interface/hdsdp.c:29-110 */
static void get_statistics( hdsdp *s ) {
    int n = s->sdp->nCol;
    set_int_feature(s, INT_FEATURE_N_MAXCONEDIM, n);
    set_int_feature(s, INT_FEATURE_N_SUMCONEDIMS, n);
    set_int_feature(s, INT_FEATURE_N_ROWS, s->nRows);
    set_int_feature(s, INT_FEATURE_N_CONES, 1);
    s->dAllConeDims = (double) n + 2 * s->nRows;

    double objOne = sdp_cone_obj_norm(s->sdp, ABS_NORM);
    double dataOne = sdp_cone_coeff_norm(s->sdp, ABS_NORM);
    double t = sdp_cone_obj_norm(s->sdp, FRO_NORM);
    double objFro = sqrt(t * t);
    t = sdp_cone_coeff_norm(s->sdp, FRO_NORM);
    double dataFro = sqrt(t * t);
    set_dbl_feature(s, DBL_FEATURE_OBJFRONORM, objFro);
    set_dbl_feature(s, DBL_FEATURE_DATAFRONORM, dataFro);
    set_dbl_feature(s, DBL_FEATURE_OBJONENORM, objOne);
    set_dbl_feature(s, DBL_FEATURE_DATAONENORM, dataOne);

    double rhsOne = 0.0, rhsFro = 0.0, rhsInf = 0.0;
    for ( int i = 0; i < s->nRows; ++i ) {
        rhsOne += fabs(s->rowRHS[i]);
        rhsFro += s->rowRHS[i] * s->rowRHS[i];
        rhsInf = HMAX(rhsInf, fabs(s->rowRHS[i]));
    }
    set_dbl_feature(s, DBL_FEATURE_RHSONENORM, rhsOne);
    set_dbl_feature(s, DBL_FEATURE_RHSFRONORM, sqrt(rhsFro));
    set_dbl_feature(s, DBL_FEATURE_RHSINFNORM, rhsInf);
    if ( objFro == 0.0 ) set_int_feature(s, INT_FEATURE_I_NULLOBJ, 1);
    set_int_feature(s, INT_FEATURE_N_DSSDPCONES, 1);
    set_int_feature(s, INT_FEATURE_N_SPSDPCONES, 0);
    set_int_feature(s, INT_FEATURE_N_LPCONES, 0);
}

/* interface/hdsdp.c:130-261 */
static void adjust_cone_params( hdsdp *s ) {
    sdp_cone_features(s->sdp, s->rowRHS, s->HIntFeatures, s->HDblFeatures);
    int impTrace = get_int_feature(s, INT_FEATURE_I_IMPTRACE);
    int noPInt = get_int_feature(s, INT_FEATURE_I_NOPINTERIOR);
    int noDInt = get_int_feature(s, INT_FEATURE_I_NODINTERIOR);
    int dense = get_int_feature(s, INT_FEATURE_I_VERYDENSE);
    int impDual = get_int_feature(s, INT_FEATURE_I_IMPYBOUND);
    int noObj = get_int_feature(s, INT_FEATURE_I_NULLOBJ);
    memset(s->modelFeatures, 0, sizeof(s->modelFeatures));
    int hit = (impTrace + noPInt + noDInt + dense + noObj) || impDual;
    if ( hit ) strcat(s->modelFeatures, "This is a ");
    if ( dense ) {
        set_int_param(s, INT_PARAM_CORRECTORA, 4);
        set_dbl_param(s, DBL_PARAM_DUALSTART, 1.0);
        set_dbl_param(s, DBL_PARAM_DUALBOX_UP, 1e+04);
        set_dbl_param(s, DBL_PARAM_DUALBOX_LOW, -1e+04);
        strcat(s->modelFeatures, "dense ");
    }
    if ( impTrace ) {
        set_dbl_param(s, DBL_PARAM_DUALSTART, 1e+03);
        set_dbl_param(s, DBL_PARAM_TRXESTIMATE, get_dbl_feature(s, DBL_FEATURE_IMPTRACEX));
        set_dbl_param(s, DBL_PARAM_POBJSTART, 1e+08);
        set_dbl_param(s, DBL_PARAM_POTRHOVAL, 5.0);
        set_dbl_param(s, DBL_PARAM_DUALBOX_UP, 1e+06);
        set_dbl_param(s, DBL_PARAM_DUALBOX_LOW, -1e+06);
        strcat(s->modelFeatures, "trace-implied ");
    }
    if ( noPInt ) {
        set_dbl_param(s, DBL_PARAM_DUALBOX_UP, 1e+04);
        set_dbl_param(s, DBL_PARAM_DUALBOX_LOW, -1e+04);
        set_dbl_param(s, DBL_PARAM_DUALSTART, 1e+03);
        set_dbl_param(s, DBL_PARAM_PRECORDACC, 1e-07);
        strcat(s->modelFeatures, "no-primal interior ");
    }
    if ( impDual ) {
        double up = get_dbl_param(s, DBL_PARAM_DUALBOX_UP), lo = get_dbl_param(s, DBL_PARAM_DUALBOX_LOW);
        int isUp = 0, isLo = 0;
        if ( get_dbl_feature(s, DBL_FEATURE_IMPYBOUNDUP) ) {
            isUp = 1;
            set_dbl_param(s, DBL_PARAM_DUALBOX_UP, HMIN(up, get_dbl_feature(s, DBL_FEATURE_IMPYBOUNDUP)));
        }
        if ( get_dbl_feature(s, DBL_FEATURE_IMPYBOUNDLOW) ) {
            isLo = 1;
            set_dbl_param(s, DBL_PARAM_DUALBOX_LOW, HMAX(lo, get_dbl_feature(s, DBL_FEATURE_IMPYBOUNDLOW)));
        }
        if ( isUp && isLo ) {
            set_dbl_param(s, DBL_PARAM_DUALSTART, 1e+02);
            set_dbl_param(s, DBL_PARAM_POBJSTART, 1e+05);
        } else {
            set_dbl_param(s, DBL_PARAM_DUALSTART, 1e+05);
            set_dbl_param(s, DBL_PARAM_POBJSTART, 1e+10);
            set_int_param(s, INT_PARAM_CORRECTORA, 12);
            set_int_param(s, INT_PARAM_CORRECTORB, 12);
        }
        set_dbl_param(s, DBL_PARAM_ABSOPTTOL, 1e-01);
        set_dbl_param(s, DBL_PARAM_RELOPTTOL, 1e-04);
        set_dbl_param(s, DBL_PARAM_PRECORDACC, 1e-05);
        strcat(s->modelFeatures, "dual-bounded ");
    }
    if ( noDInt ) {
        set_dbl_param(s, DBL_PARAM_DUALBOX_UP, 1.0);
        set_dbl_param(s, DBL_PARAM_DUALBOX_LOW, -1.0);
        if ( s->dAllConeDims > 100000 ) {
            set_dbl_param(s, DBL_PARAM_DUALSTART, 1e+00);
            set_dbl_param(s, DBL_PARAM_ABSFEASTOL, 1e-04);
            set_dbl_param(s, DBL_PARAM_RELFEASTOL, 1e-05);
        } else {
            set_dbl_param(s, DBL_PARAM_DUALBOX_UP, 1e+01);
            set_dbl_param(s, DBL_PARAM_DUALBOX_LOW, -1e+01);
            set_dbl_param(s, DBL_PARAM_ABSFEASTOL, 1e-05);
            set_dbl_param(s, DBL_PARAM_RELFEASTOL, 1e-07);
        }
        set_dbl_param(s, DBL_PARAM_PRECORDACC, 1e-05);
        strcat(s->modelFeatures, "no-dual interior ");
    }
    if ( noObj ) {
        set_dbl_param(s, DBL_PARAM_DUALSTART, 1.0);
        set_dbl_param(s, DBL_PARAM_DUALBOX_UP, 1.0);
        set_dbl_param(s, DBL_PARAM_DUALBOX_LOW, -1.0);
        strcat(s->modelFeatures, "no objective ");
    }
    if ( hit ) strcat(s->modelFeatures, "SDP problem\n");
}

/* interface/hdsdp.c:263-376 */
static hdsdp_retcode adjust_params( hdsdp *s ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    double objOne = get_dbl_feature(s, DBL_FEATURE_OBJONENORM);
    double rhsInf = get_dbl_feature(s, DBL_FEATURE_RHSINFNORM);
    double objScal = 1.0, rhsScal = 1.0;
    if ( objOne > 1e+10 ) objScal = 1e-08;
    else if ( objOne > 1e+08 ) objScal = 1e-06;
    else if ( objOne > 1e+05 ) objScal = 1e-05;
    if ( rhsInf > 1e+10 ) rhsScal = 1e-08;
    else if ( rhsInf > 1e+08 ) rhsScal = 1e-06;
    set_dbl_feature(s, DBL_FEATURE_OBJSCALING, objScal);
    set_dbl_feature(s, DBL_FEATURE_RHSSCALING, rhsScal);
    HCALL(sdp_cone_scal(s->sdp, objScal));
    for ( int i = 0; i < s->nRows; ++i ) s->rowRHS[i] = rhsScal * s->rowRHS[i];
    hlog(s, "    Scale cone objective by %5.1e \n", objScal);
    hlog(s, "    Scale rhs by %5.1e \n", rhsScal);

    int maxDim = get_int_feature(s, INT_FEATURE_N_MAXCONEDIM);
    int nCorrA = (int) (s->nRows - 2) / maxDim;
    if ( get_int_feature(s, INT_FEATURE_N_SUMCONEDIMS) < 100 && nCorrA == 0 ) nCorrA = 1;
    if ( nCorrA >= 1 ) nCorrA += 1;
    nCorrA = nCorrA * nCorrA;
    if ( s->nRows < 2000 && nCorrA > 10 ) nCorrA = 10;
    int nCorrB = nCorrA;
    if ( maxDim >= 5 * s->nRows ) {
        nCorrB = 0;
        nCorrA = 2;
    } else if ( maxDim >= s->nRows ) {
        nCorrB = HMIN(nCorrB, 2);
        nCorrA = 4;
    } else {
        nCorrA = 6;
    }
    if ( s->nRows > 20 * maxDim ) {
        nCorrB = HMAX(nCorrB, 12);
        nCorrA = 12;
    } else if ( s->nRows > 5 * maxDim ) {
        nCorrB = HMAX(nCorrB, 10);
        nCorrA = 10;
    } else if ( s->nRows > 2 * maxDim ) {
        nCorrB = HMAX(nCorrB, 8);
        nCorrA = 8;
    }
    nCorrB = HMIN(nCorrB, 12);
    nCorrA = HMAX(nCorrA, 2);
    set_int_param(s, INT_PARAM_CORRECTORA, nCorrA);
    set_int_param(s, INT_PARAM_CORRECTORB, nCorrB);
    adjust_cone_params(s);
exit_cleanup:
    return retcode;
}

/* ---- solver lifecycle (interface/hdsdp.c:472-615) ------------------------- */

hdsdp_retcode HDSDPCreate( hdsdp **pHSolver ) {
    if ( !pHSolver ) return HDSDP_RETCODE_FAILED;
    hdsdp *s = (hdsdp *) calloc(1, sizeof(hdsdp));
    if ( !s ) return HDSDP_RETCODE_MEMORY;
    s->verbose = 1;
    *pHSolver = s;
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode HDSDPInit( hdsdp *s, int nRows, int nCones ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    s->nRows = nRows;
    s->nCones = nCones;
    default_params(s);
    if ( nCones != 1 ) {
        /* one dense SDP block per solve in this build */
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    HALLOC(s->rowRHS, double, nRows);
    HALLOC(s->userCones, user_data *, nCones);
    HCALL(kkt_create(&s->kkt, nRows));
    HALLOC(s->dRowDual, double, nRows);
    HALLOC(s->dRowDualStep, double, nRows);
    HALLOC(s->dMinvRowRHS, double, nRows);
    HALLOC(s->dMinvASinv, double, nRows);
    HALLOC(s->dMinvASinvRdSinv, double, nRows);
    HALLOC(s->dMinvASinvCSinv, double, nRows);
    HALLOC(s->dHAuxiVec1, double, nRows);
    HALLOC(s->dHAuxiVec2, double, nRows);
    HALLOC(s->dPInfeasUpper, double, nRows);
    HALLOC(s->dPInfeasLower, double, nRows);
    HALLOC(s->dAccRowDualMaker, double, nRows);
    HALLOC(s->dAccRowDualStepMaker, double, nRows);
    HALLOC(s->dInaccRowDualMaker, double, nRows);
    HALLOC(s->dInaccRowDualStepMaker, double, nRows);
    HALLOC(s->iterLog, double, (size_t) HDSDP_MAX_LOG * HDSDP_LOG_FIELDS);
    s->dBarrierMu = 1e+10;
    s->comp = HDSDP_INFINITY;
    s->dAccBarrierMaker = -1.0;
    s->dInaccBarrierMaker = -1.0;
    s->HStatus = HDSDP_UNKNOWN;
exit_cleanup:
    return retcode;
}

hdsdp_retcode HDSDPSetCone( hdsdp *s, int iCone, void *userCone ) {
    if ( iCone < 0 || iCone >= s->nCones ) return HDSDP_RETCODE_FAILED;
    user_data *ud = (user_data *) userCone;
    if ( HUserDataChooseCone(ud) != HDSDP_CONETYPE_DENSE_SDP ) return HDSDP_RETCODE_FAILED;
    s->userCones[iCone] = ud;
    return HDSDP_RETCODE_OK;
}

void HDSDPSetDualObjective( hdsdp *s, double *dObj ) {
    memcpy(s->rowRHS, dObj, sizeof(double) * s->nRows);
}

void HDSDPSetIntParam( hdsdp *s, int p, int v ) {
    if ( p >= 3 || p < 0 ) return;
    s->HIntParams[p] = v;
}

void HDSDPSetDblParam( hdsdp *s, int p, double v ) {
    if ( p >= 13 || p < 0 ) return;
    s->HDblParams[p] = v;
}

void HDSDPSetIntParamAny( hdsdp *s, int p, int v ) {
    if ( p < 0 || p >= NUM_INT_PARAM ) return;
    s->HIntParams[p] = v;
}

void HDSDPSetDblParamAny( hdsdp *s, int p, double v ) {
    if ( p < 0 || p >= NUM_DBL_PARAM ) return;
    s->HDblParams[p] = v;
}

void HDSDPSetPrecision( hdsdp *s, int mode ) {
    s->precision = ( mode == HDSDP_PRECISION_MIXED ) ? HDSDP_PRECISION_MIXED : HDSDP_PRECISION_FP64;
}

void HDSDPSetVerbose( hdsdp *s, int verbose ) {
    s->verbose = verbose;
}

void HDSDPSetProfiling( hdsdp *s, int enable ) {
    s->profiling = enable ? 1 : 0;
}

int HDSDPGetProfile( hdsdp *s, double *ms, double *counts, int nSlots ) {
    if ( !s->sdp || !s->sdp->dev ) return 0;
    return hdsdp_dev_get_profile(s->sdp->dev, ms, counts, nSlots);
}

/* interface/hdsdp.c:617-689 */
/* Pre-solve: cone and device set-up, rank-one detection, statistics, scaling and
 * parameter adjustment, bound cone (interface/hdsdp.c:617-689). */
static hdsdp_retcode presolve( hdsdp *s ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    hlog(s, "Pre-solver starts \n");
    HCALL(sdp_cone_create(&s->sdp, s->userCones[0], s->precision));
    if ( s->profiling ) HCALL(hdsdp_dev_set_profiling(s->sdp->dev, 1));
    HCALL(sdp_cone_presolve(s->sdp));
    HCALL(kkt_init(s->kkt, s->sdp));
    get_statistics(s);
    HCALL(adjust_params(s));
    HCALL(bnd_cone_create(&s->bnd, s->nRows, get_dbl_param(s, DBL_PARAM_DUALBOX_LOW),
                          get_dbl_param(s, DBL_PARAM_DUALBOX_UP)));
    hlog(s, "    Dual box low: %+5.1e\n", get_dbl_param(s, DBL_PARAM_DUALBOX_LOW));
    hlog(s, "    Dual box up : %+5.1e\n", get_dbl_param(s, DBL_PARAM_DUALBOX_UP));
    hlog(s, "Pre-solver ends. Elapsed time: %3.1f seconds \n", hdsdp_timestamp() - s->dTimeBegin);
    hlog(s, "%s", s->modelFeatures);
exit_cleanup:
    return retcode;
}

hdsdp_retcode HDSDPSetup( hdsdp *s ) {
    if ( !s->userCones || !s->userCones[0] ) return HDSDP_RETCODE_FAILED;
    if ( s->sdp ) return HDSDP_RETCODE_OK;
    s->dTimeBegin = hdsdp_timestamp();
    return presolve(s);
}

hdsdp_retcode HDSDPOptimize( hdsdp *s, int dOptOnly ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    hlog(s, "\nHDSDP: software for semi-definite programming (MI355X build, backend %s)\n",
         hdsdp_dev_backend_name());
    hlog(s, "---------------------------------------------\n");
    s->dTimeBegin = hdsdp_timestamp();
    s->nLog = 0;
    if ( !s->userCones || !s->userCones[0] ) {
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    if ( !s->sdp ) HCALL(presolve(s));
    s->adjReady = 0;

    retcode = hdsdp_solve(s, dOptOnly);
    s->dTimeEnd = hdsdp_timestamp();
    hlog(s, "\nOptimization time: %3.1f seconds\n", s->dTimeEnd - s->dTimeBegin);
    if ( s->HStatus != HDSDP_INFEAS_OR_UNBOUNDED && s->HStatus != HDSDP_SUSPECT_INFEAS_OR_UNBOUNDED ) {
        /* the solve's return code is replaced by the check's (interface/hdsdp.c:673-679) */
        retcode = HDSDPCheckSolution(s, s->dErrs);
    }
    hlog(s, "DIMACS error metric:\n    %5.2e %5.2e %5.2e %5.2e %5.2e %5.2e \n",
         s->dErrs[0], s->dErrs[1], s->dErrs[2], s->dErrs[3], s->dErrs[4], s->dErrs[5]);
    hlog(s, "\nSDP Status: %d \n  pObj %+15.10e\n  dObj %+15.10e\n", (int) s->HStatus, s->pObjVal, s->dObjVal);
exit_cleanup:
    if ( s->dTimeEnd == 0.0 ) s->dTimeEnd = hdsdp_timestamp();
    return retcode;
}

hdsdp_retcode HDSDPGetRowDual( hdsdp *s, double *pObjVal, double *dObjVal, double *dualVal ) {
    if ( pObjVal ) *pObjVal = s->pObjVal;
    if ( dObjVal ) *dObjVal = s->dObjVal;
    if ( dualVal ) memcpy(dualVal, s->dRowDual, sizeof(double) * s->nRows);
    return HDSDP_RETCODE_OK;
}

/* Recover X from the more accurate snapshot, else the less accurate one
 * (interface/hdsdp.c:711-745). */
hdsdp_retcode hdsdp_recover_primal( hdsdp *s, int *ok ) {
    double mu = s->dAccBarrierMaker;
    double *y = s->dAccRowDualMaker, *dy = s->dAccRowDualStepMaker;
    if ( mu <= 0.0 ) {
        mu = s->dInaccBarrierMaker;
        y = s->dInaccRowDualMaker;
        dy = s->dInaccRowDualStepMaker;
    }
    return sdp_cone_primal(s->sdp, mu, y, dy, ok);
}

void HDSDPGetConeValues( hdsdp *s, int iCone, double *conePrimal, double *coneDual, double *coneAuxi ) {
    (void) coneAuxi;
    if ( iCone != 0 || !s->sdp ) return;
    int ok = 0;
    if ( conePrimal ) {
        if ( hdsdp_recover_primal(s, &ok) == HDSDP_RETCODE_OK && ok ) {
            hdsdp_dev_get_primal(s->sdp->dev, conePrimal);
        } else if ( s->verbose ) {
            printf("Recovery step is infeasible\n");
        }
    }
    if ( coneDual ) hdsdp_dev_get_dual(s->sdp->dev, coneDual);
}

/* Largest eigenvalue of the symmetric n x n matrix X (full storage): Lanczos with full
 * reorthogonalisation from the normalised ones vector, at most HDSDP_EIG_SMALL_MAX steps
 * (exact up to rounding for n <= HDSDP_EIG_SMALL_MAX; beyond, the top Ritz value of that
 * Krylov space), then the top eigenvalue of the tridiagonal matrix. */
static double sym_lambda_max( int n, const double *X ) {
    const int kmax = n < HDSDP_EIG_SMALL_MAX ? n : HDSDP_EIG_SMALL_MAX;
    double *V = (double *) calloc((size_t) n * (kmax + 1), sizeof(double));
    double *w = (double *) calloc((size_t) n, sizeof(double));
    double alpha[HDSDP_EIG_SMALL_MAX], beta[HDSDP_EIG_SMALL_MAX], ev = 0.0;
    int k = 0;
    if ( !V || !w ) goto done;
    for ( int i = 0; i < n; ++i ) V[i] = 1.0 / sqrt((double) n);
    for ( k = 0; k < kmax; ++k ) {
        const double *v = V + (size_t) k * n;
        for ( int i = 0; i < n; ++i ) {
            double acc = 0.0;
            for ( int j = 0; j < n; ++j ) acc += X[i + (size_t) j * n] * v[j];
            w[i] = acc;
        }
        double a = 0.0;
        for ( int i = 0; i < n; ++i ) a += w[i] * v[i];
        alpha[k] = a;
        /* full reorthogonalisation against every Lanczos vector, twice */
        for ( int pass = 0; pass < 2; ++pass )
            for ( int q = 0; q <= k; ++q ) {
                const double *u = V + (size_t) q * n;
                double c = 0.0;
                for ( int i = 0; i < n; ++i ) c += w[i] * u[i];
                for ( int i = 0; i < n; ++i ) w[i] -= c * u[i];
            }
        double b = 0.0;
        for ( int i = 0; i < n; ++i ) b += w[i] * w[i];
        b = sqrt(b);
        if ( k + 1 == kmax || b <= 1e-14 * fabs(a) + 1e-300 ) {
            ++k;
            break;
        }
        beta[k] = b;
        for ( int i = 0; i < n; ++i ) V[(size_t) (k + 1) * n + i] = w[i] / b;
    }
    if ( hdsdp_eig_tridiag_top(k, alpha, beta, 1, &ev, NULL) != 0 ) ev = alpha[0];
done:
    free(V);
    free(w);
    return ev;
}

/* interface/hdsdp.c:753-906 */
hdsdp_retcode HDSDPCheckSolution( hdsdp *s, double dErrs[6] ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    double *AX = NULL;
    for ( int k = 0; k < 6; ++k ) dErrs[k] = 1.0;
    if ( s->dInaccBarrierMaker < 0.0 && s->dAccBarrierMaker < 0.0 ) {
        s->HStatus = HDSDP_NUMERICAL;
        return HDSDP_RETCODE_OK;
    }
    int m = s->nRows;
    double dualInfeas = s->dPerturb * sqrt(get_int_feature(s, INT_FEATURE_N_SUMCONEDIMS));
    double primalScal = get_dbl_feature(s, DBL_FEATURE_RHSSCALING);
    double dualScal = get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
    double dualObj = 0.0, primalObj = 0.0, compl = 0.0, traceX = 0.0, maxEig = HDSDP_INFINITY;
    HALLOC(AX, double, m);
    for ( int i = 0; i < m; ++i ) dualObj += s->rowRHS[i] * s->dRowDual[i];

    int ok = 0;
    HCALL(hdsdp_recover_primal(s, &ok));
    if ( ok ) {
        HCALL(hdsdp_dev_primal_stats(s->sdp->dev, AX, &primalObj, &compl, &traceX));
        /* The reference takes the largest eigenvalue of X here (fds_syev with m = 1:
         * dsyevr, il = iu = n, linalg/dense_opts.c:53-73): a positive trace makes it
         * positive (lambda_max >= tr X / n) and the error zero; otherwise it is computed. */
        if ( traceX > 0.0 ) {
            maxEig = traceX / s->sdp->nCol;
        } else {
            const int n = s->sdp->nCol;
            double *X = NULL;
            HALLOC(X, double, (size_t) n * n);
            if ( hdsdp_dev_get_primal(s->sdp->dev, X) == HDSDP_RETCODE_OK ) maxEig = sym_lambda_max(n, X);
            free(X);
        }
    } else {
        hlog(s, "Recovery step is infeasible\n");
    }
    dualObj = dualObj / (primalScal * dualScal);
    primalObj = primalObj / (primalScal * dualScal);
    double pInf = 0.0;
    for ( int i = 0; i < m; ++i ) {
        double r = AX[i] - s->rowRHS[i];
        pInf += r * r;
    }
    pInf = sqrt(pInf) / get_dbl_feature(s, DBL_FEATURE_RHSSCALING);
    dualInfeas = dualInfeas / get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
    double rhsOne = get_dbl_feature(s, DBL_FEATURE_RHSONENORM);
    dErrs[0] = pInf / (1.0 + rhsOne);
    dErrs[1] = ( maxEig < 0.0 ) ? -maxEig / (1.0 + rhsOne) : 0.0;
    dErrs[2] = dualInfeas / (1.0 + get_dbl_feature(s, DBL_FEATURE_OBJONENORM));
    dErrs[3] = 0.0;
    dErrs[4] = (primalObj - dualObj) / (fabs(primalObj) + fabs(dualObj) + 1.0);
    dErrs[5] = compl / (fabs(primalObj) + fabs(dualObj) + 1.0);
    double maxErr = 0.0;
    for ( int k = 0; k < 6; ++k ) maxErr = HMAX(maxErr, fabs(dErrs[k]));
    s->pObjVal = primalObj;
    s->dObjVal = dualObj;
    if ( maxErr > 1e-02 ) {
        if ( s->dAccBarrierMaker < 0.0 ) {
            s->HStatus = HDSDP_NUMERICAL;
        } else {
            hlog(s, "\nDealing with primal solution\n");
            s->dAccBarrierMaker = -1.0;
            free(AX);
            return HDSDPCheckSolution(s, dErrs);
        }
    } else {
        s->HStatus = HDSDP_PRIMAL_DUAL_OPTIMAL;
    }
exit_cleanup:
    free(AX);
    return retcode;
}

void HDSDPClear( hdsdp *s ) {
    if ( !s ) return;
    sdp_cone_destroy(&s->sdp);
    bnd_cone_destroy(&s->bnd);
    kkt_destroy(&s->kkt);
    free(s->rowRHS);
    free(s->userCones);
    free(s->dRowDual); free(s->dRowDualStep);
    free(s->dMinvRowRHS); free(s->dMinvASinv); free(s->dMinvASinvRdSinv); free(s->dMinvASinvCSinv);
    free(s->dHAuxiVec1); free(s->dHAuxiVec2);
    free(s->dPInfeasUpper); free(s->dPInfeasLower);
    free(s->dAccRowDualMaker); free(s->dAccRowDualStepMaker);
    free(s->dInaccRowDualMaker); free(s->dInaccRowDualStepMaker);
    free(s->iterLog);
    free(s->adjY);
    int verbose = s->verbose, profiling = s->profiling;
    memset(s, 0, sizeof(hdsdp));
    s->verbose = verbose;
    s->profiling = profiling;
}

void HDSDPDestroy( hdsdp **pHSolver ) {
    if ( !pHSolver || !*pHSolver ) return;
    HDSDPClear(*pHSolver);
    free(*pHSolver);
    *pHSolver = NULL;
}

void HDSDPGetDimacsErrors( hdsdp *s, double errs[6] ) {
    memcpy(errs, s->dErrs, sizeof(double) * 6);
}

int HDSDPGetIterLog( hdsdp *s, int maxRecords, double *records ) {
    int n = HMIN(maxRecords, s->nLog);
    if ( records && n > 0 ) memcpy(records, s->iterLog, sizeof(double) * HDSDP_LOG_FIELDS * n);
    return s->nLog;
}

void HDSDPGetSummary( hdsdp *s, double gapTarget, int *status, int *nIter, double *mu,
                      double *solveSeconds, double *secondsToGap ) {
    if ( status ) *status = (int) s->HStatus;
    if ( nIter ) *nIter = s->nIterCount;
    if ( mu ) *mu = s->dBarrierMu;
    if ( solveSeconds ) *solveSeconds = s->dTimeEnd - s->dTimeBegin;
    if ( secondsToGap ) {
        *secondsToGap = -1.0;
        for ( int k = 0; k < s->nLog; ++k ) {
            const double *r = s->iterLog + (size_t) k * HDSDP_LOG_FIELDS;
            double gap = (r[2] - r[3]) / (fabs(r[2]) + fabs(r[3]) + 1.0);
            if ( gap <= gapTarget ) {
                *secondsToGap = r[8];
                break;
            }
        }
    }
}
