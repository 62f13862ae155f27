/*
 * cone_sdp.c -- host half of the dense SDP cone.
 *
 * Keeps the cone's scalar state (dual residual Rd, diagonal perturbation) and
 * the presolved coefficients, and forwards every matrix operation to the
 * device ABI.  Slot semantics follow interface/hdsdp_conic_sdp.c:
 *   update         S = tau C - A^T y - Rd I (+ perturbation)   (1612-1629)
 *   ratio test     dS = dtau C - A^T dy + ada Rd I, Lanczos     (1638-1682)
 *   interior check update + Cholesky of S                       (2105-2141)
 *   barrier        2 sum log diag(L)                            (2185-2224)
 *   schur          M, A S^-1, Rd A S^-2                         (1722-1804)
 */
#include "hdsdp_host.h"

hdsdp_retcode sdp_cone_create( sdp_cone **pcone, user_data *ud, int precision ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    sdp_cone *cone = NULL;
    HALLOC(cone, sdp_cone, 1);
    *pcone = cone;
    cone->nRow = ud->nConicRow;
    cone->nCol = ud->nConicCol;
    int n = cone->nCol, m = cone->nRow;
    HALLOC(cone->sdpObj, sdp_coeff, 1);
    HALLOC(cone->sdpRow, sdp_coeff, m);

    if ( ud->isFactorForm ) {
        HCALL(coeff_set_from_csc(cone->sdpObj, n, ud->cNnz, ud->cIdx, ud->cElem));
        for ( int i = 0; i < m; ++i ) {
            HCALL(coeff_set_rank_one(&cone->sdpRow[i], n, ud->factorSign[i], ud->factor + (size_t) i * n));
        }
    } else {
        /* column 0 = C, column i + 1 = A_i (interface/hdsdp_conic_sdp.c:1367-1378) */
        const int *beg = ud->coneMatBeg;
        HCALL(coeff_set_from_csc(cone->sdpObj, n, beg[1], ud->coneMatIdx, ud->coneMatElem));
        for ( int i = 0; i < m; ++i ) {
            HCALL(coeff_set_from_csc(&cone->sdpRow[i], n, beg[i + 2] - beg[i + 1],
                                     ud->coneMatIdx + beg[i + 1], ud->coneMatElem + beg[i + 1]));
        }
    }
    memset(cone->sdpConeStats, 0, sizeof(cone->sdpConeStats));
    cone->sdpConeStats[cone->sdpObj->type] += 1;
    for ( int i = 0; i < m; ++i ) cone->sdpConeStats[cone->sdpRow[i].type] += 1;
    HCALL(hdsdp_dev_create(&cone->dev, m, n, precision));

exit_cleanup:
    return retcode;
}

hdsdp_retcode sdp_cone_presolve( sdp_cone *cone ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    hdsdp_coeff_view *views = NULL;
    /* rank-one detection (interface/hdsdp_conic_sdp.c:1486-1514) */
    HCALL(coeff_presolve_rank_one(cone->sdpObj));
    for ( int i = 0; i < cone->nRow; ++i ) {
        cone->sdpConeStats[cone->sdpRow[i].type] -= 1;
        HCALL(coeff_presolve_rank_one(&cone->sdpRow[i]));
        cone->sdpConeStats[cone->sdpRow[i].type] += 1;
    }
    HALLOC(views, hdsdp_coeff_view, cone->nRow + 1);
    coeff_view(cone->sdpObj, &views[0]);
    for ( int i = 0; i < cone->nRow; ++i ) coeff_view(&cone->sdpRow[i], &views[i + 1]);
    HCALL(hdsdp_dev_load(cone->dev, views));
exit_cleanup:
    free(views);
    return retcode;
}

void sdp_cone_destroy( sdp_cone **pcone ) {
    if ( !pcone || !*pcone ) return;
    sdp_cone *cone = *pcone;
    hdsdp_dev_destroy(&cone->dev);
    if ( cone->sdpObj ) coeff_clear(cone->sdpObj);
    free(cone->sdpObj);
    if ( cone->sdpRow ) {
        for ( int i = 0; i < cone->nRow; ++i ) coeff_clear(&cone->sdpRow[i]);
    }
    free(cone->sdpRow);
    free(cone);
    *pcone = NULL;
}

double sdp_cone_obj_norm( sdp_cone *cone, int type ) {
    return coeff_norm(cone->sdpObj, type);
}

/* interface/hdsdp_conic_sdp.c:1564-1580 */
double sdp_cone_coeff_norm( sdp_cone *cone, int type ) {
    double nrm = 0.0;
    for ( int i = 0; i < cone->nRow; ++i ) {
        double t = coeff_norm(&cone->sdpRow[i], type);
        nrm += ( type == ABS_NORM ) ? t : t * t;
    }
    return ( type == ABS_NORM ) ? nrm : sqrt(nrm);
}

hdsdp_retcode sdp_cone_scal( sdp_cone *cone, double s ) {
    coeff_scale(cone->sdpObj, s);
    return hdsdp_dev_scale_obj(cone->dev, s);
}

/* interface/hdsdp_conic_sdp.c:2560-2654 */
void sdp_cone_features( sdp_cone *cone, const double *rowRHS, int iF[20], double dF[20] ) {
    iF[INT_FEATURE_N_ZEORMATS] = cone->sdpConeStats[HDSDP_COEFF_ZERO];
    iF[INT_FEATURE_N_DSMATS] = cone->sdpConeStats[HDSDP_COEFF_DENSE];
    iF[INT_FEATURE_N_SPMATS] = cone->sdpConeStats[HDSDP_COEFF_SPARSE];
    iF[INT_FEATURE_N_SPR1MATS] = cone->sdpConeStats[HDSDP_COEFF_SPR1];
    iF[INT_FEATURE_N_DSR1MATS] = cone->sdpConeStats[HDSDP_COEFF_DSR1];

    int noPrimalInterior = 0, impliedTrace = 0;
    double impliedTraceX = 0.0;
    for ( int i = 0; i < cone->nRow; ++i ) {
        if ( coeff_rank(&cone->sdpRow[i]) == 1 &&
             fabs(rowRHS[i]) < 1e-03 * coeff_norm(&cone->sdpRow[i], FRO_NORM) ) {
            noPrimalInterior = 1;
        }
    }
    if ( noPrimalInterior ) iF[INT_FEATURE_I_NOPINTERIOR] = 1;

    for ( int i = 0; i < cone->nRow; ++i ) {
        double mult = 0.0;
        if ( coeff_is_eye(&cone->sdpRow[i], &mult) && rowRHS[i] / mult > 0.0 ) {
            impliedTrace = 1;
            break;
        }
    }
    /* The trace found above is discarded before the diagonal test, as in the
     * reference; only diag(X) = d constraints contribute to the estimate. */
    int nMark = HMAX(cone->nRow, cone->nCol);
    int *unitCol = (int *) calloc((size_t) nMark, sizeof(int));
    impliedTraceX = 0.0;
    if ( unitCol && !impliedTrace ) {
        for ( int i = 0; i < cone->nRow; ++i ) {
            int c = 0;
            if ( coeff_is_unit_col(&cone->sdpRow[i], &c) && !unitCol[c] ) {
                unitCol[c] = 1;
                impliedTraceX += rowRHS[i];
            }
        }
    }
    int nSumUnit = 0;
    if ( unitCol ) {
        for ( int i = 0; i < cone->nRow; ++i ) nSumUnit += unitCol[i];
    }
    if ( nSumUnit == cone->nCol ) impliedTrace = 1;
    free(unitCol);
    if ( impliedTrace ) {
        iF[INT_FEATURE_I_IMPTRACE] = 1;
        dF[DBL_FEATURE_IMPTRACEX] = impliedTraceX;
    }
    if ( cone->sdpConeStats[HDSDP_COEFF_DENSE] >= 0.7 * cone->nRow ) {
        iF[INT_FEATURE_I_VERYDENSE] = 1;
    }
}

static hdsdp_retcode sdp_cone_buffer( sdp_cone *cone, double cCoef, double aScal, const double *aCoef,
                                      double eye, int whichBuffer ) {
    if ( whichBuffer != BUFFER_DUALSTEP ) eye += cone->dualPerturb;
    return hdsdp_dev_update_buffer(cone->dev, cCoef, aScal, aCoef, eye, whichBuffer);
}

hdsdp_retcode sdp_cone_update( sdp_cone *cone, double tau, const double *y ) {
    return sdp_cone_buffer(cone, tau, -1.0, y, -cone->dualResidual, BUFFER_DUALVAR);
}

hdsdp_retcode sdp_cone_ratio_test( sdp_cone *cone, double dtau, const double *dy, double ada,
                                   int whichBuffer, double *maxStep ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    HCALL(sdp_cone_buffer(cone, dtau, -1.0, dy, ada * cone->dualResidual, BUFFER_DUALSTEP));
    HCALL(hdsdp_dev_ratio_test(cone->dev, whichBuffer, maxStep));
exit_cleanup:
    return retcode;
}

hdsdp_retcode sdp_cone_schur( sdp_cone *cone, host_kkt *kkt, int typeKKT ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    double *asinv = NULL, *asinvrd = NULL;
    double traceSinv = 0.0;
    int m = cone->nRow;
    HALLOC(asinv, double, m);
    HALLOC(asinvrd, double, m);
    HCALL(hdsdp_dev_schur(cone->dev, typeKKT, cone->dualResidual, asinv, asinvrd, &traceSinv));
    for ( int i = 0; i < m; ++i ) {
        kkt->dASinvVec[i] += asinv[i];
        kkt->dASinvRdSinvVec[i] += asinvrd[i];
    }
    kkt->dTraceSinv += traceSinv;
    if ( typeKKT == KKT_TYPE_HOMOGENEOUS ) {
        /* the embedding's C terms (interface/hdsdp_conic_sdp.c:737-760, 983-1029) */
        double cs = 0.0, csc = 0.0, csrd = 0.0;
        HCALL(hdsdp_dev_hsd(cone->dev, cone->dualResidual, asinv, &cs, &csc, &csrd));
        for ( int i = 0; i < m; ++i ) kkt->dASinvCSinvVec[i] += asinv[i];
        kkt->dCSinv += cs;
        kkt->dCSinvCSinv += csc;
        if ( cone->dualResidual ) kkt->dCSinvRdSinv += csrd;
    }
exit_cleanup:
    free(asinv);
    free(asinvrd);
    return retcode;
}

hdsdp_retcode sdp_cone_interior( sdp_cone *cone, double tau, const double *y, int *isInterior ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    HCALL(sdp_cone_update(cone, tau, y));
    HCALL(hdsdp_dev_factorize(cone->dev, BUFFER_DUALVAR, isInterior));
exit_cleanup:
    return retcode;
}

hdsdp_retcode sdp_cone_interior_expert( sdp_cone *cone, double cCoef, double aScal, const double *aCoef,
                                        double eye, int whichBuffer, int *isInterior ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    HCALL(sdp_cone_buffer(cone, cCoef, aScal, aCoef, eye, whichBuffer));
    HCALL(hdsdp_dev_factorize(cone->dev, whichBuffer == BUFFER_DUALVAR ? BUFFER_DUALVAR : BUFFER_DUALCHECK,
                              isInterior));
exit_cleanup:
    return retcode;
}

hdsdp_retcode sdp_cone_barrier( sdp_cone *cone, double tau, const double *y, int whichBuffer, double *logdet ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    if ( y ) {
        int isPsd = 0;
        HCALL(sdp_cone_update(cone, tau, y));
        HCALL(hdsdp_dev_factorize(cone->dev, BUFFER_DUALVAR, &isPsd));
        if ( !isPsd ) {  /* Numeric (dpotrf) failure is an error here */
            retcode = HDSDP_RETCODE_FAILED;
            goto exit_cleanup;
        }
    }
    HCALL(hdsdp_dev_logdet(cone->dev, whichBuffer, logdet));
exit_cleanup:
    return retcode;
}

hdsdp_retcode sdp_cone_axpy_check( sdp_cone *cone, double step, int whichBuffer, int *isInterior ) {
    return hdsdp_dev_axpy_check(cone->dev, step, whichBuffer, isInterior);
}

hdsdp_retcode sdp_cone_primal( sdp_cone *cone, double mu, const double *y, const double *dy, int *ok ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    /* the recovery factors C - A^T y (+ perturbation) in the checker */
    HCALL(sdp_cone_buffer(cone, 1.0, -1.0, y, 0.0, BUFFER_DUALCHECK));
    HCALL(hdsdp_dev_primal(cone->dev, mu, y, dy, ok));
exit_cleanup:
    return retcode;
}
