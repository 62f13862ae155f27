/*
 * cone_bound.c -- scalar box l <= y <= u as the slacks s_u = tau u - y and
 * s_l = -tau l + y (interface/hdsdp_conic_bound.c:36-418).  O(m), host side.
 */
#include "hdsdp_host.h"

hdsdp_retcode bnd_cone_create( bound_cone **pcone, int nRow, double lo, double up ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    bound_cone *c = NULL;
    HALLOC(c, bound_cone, 1);
    *pcone = c;
    c->nRow = nRow;
    HALLOC(c->dualUpper, double, nRow);
    HALLOC(c->dualLower, double, nRow);
    HALLOC(c->dualUpperInverse, double, nRow);
    HALLOC(c->dualLowerInverse, double, nRow);
    HALLOC(c->dualUpperStep, double, nRow);
    HALLOC(c->dualLowerStep, double, nRow);
    HALLOC(c->dualUpperChecker, double, nRow);
    HALLOC(c->dualLowerChecker, double, nRow);
    c->dBoundLow = lo;
    c->dBoundUp = up;
    if ( !(lo < up) ) retcode = HDSDP_RETCODE_FAILED;
exit_cleanup:
    return retcode;
}

void bnd_cone_destroy( bound_cone **pcone ) {
    if ( !pcone || !*pcone ) return;
    bound_cone *c = *pcone;
    free(c->dualUpper); free(c->dualLower);
    free(c->dualUpperInverse); free(c->dualLowerInverse);
    free(c->dualUpperStep); free(c->dualLowerStep);
    free(c->dualUpperChecker); free(c->dualLowerChecker);
    free(c);
    *pcone = NULL;
}

static void bnd_buffer( bound_cone *c, double cCoef, double aScal, const double *aCoef, int which,
                        double **pl, double **pu ) {
    double *l, *u;
    if ( which == BUFFER_DUALVAR ) { l = c->dualLower; u = c->dualUpper; }
    else if ( which == BUFFER_DUALCHECK ) { l = c->dualLowerChecker; u = c->dualUpperChecker; }
    else { l = c->dualLowerStep; u = c->dualUpperStep; }
    for ( int i = 0; i < c->nRow; ++i ) {
        u[i] = cCoef * c->dBoundUp + aScal * aCoef[i];
        l[i] = -cCoef * c->dBoundLow - aScal * aCoef[i];
    }
    if ( pl ) *pl = l;
    if ( pu ) *pu = u;
}

void bnd_cone_update( bound_cone *c, double tau, const double *y ) {
    bnd_buffer(c, tau, -1.0, y, BUFFER_DUALVAR, NULL, NULL);
}

/* Step bound: the most negative ratio step/slack decides; a direction that
 * never leaves the box reports 100 (interface/hdsdp_conic_bound.c:155-192). */
void bnd_cone_ratio_test( bound_cone *c, double dtau, const double *dy, int whichBuffer, double *maxStep ) {
    bnd_buffer(c, dtau, -1.0, dy, BUFFER_DUALSTEP, NULL, NULL);
    double *l = ( whichBuffer == BUFFER_DUALVAR ) ? c->dualLower : c->dualLowerChecker;
    double *u = ( whichBuffer == BUFFER_DUALVAR ) ? c->dualUpper : c->dualUpperChecker;
    double minRatio = 0.0;
    for ( int i = 0; i < c->nRow; ++i ) {
        minRatio = HMIN(minRatio, c->dualLowerStep[i] / l[i]);
        minRatio = HMIN(minRatio, c->dualUpperStep[i] / u[i]);
    }
    double step = 1.0 / minRatio;
    *maxStep = ( step > 0.0 ) ? 100.0 : -step;
}

/* interface/hdsdp_conic_bound.c:199-241 */
void bnd_cone_schur( bound_cone *c, host_kkt *kkt, int typeKKT ) {
    for ( int i = 0; i < c->nRow; ++i ) {
        c->dualLowerInverse[i] = 1.0 / c->dualLower[i];
        c->dualUpperInverse[i] = 1.0 / c->dualUpper[i];
    }
    for ( int i = 0; i < kkt->nRow; ++i ) {
        kkt->dASinvVec[i] -= c->dualLowerInverse[i];
        kkt->dASinvVec[i] += c->dualUpperInverse[i];
    }
    if ( typeKKT == KKT_TYPE_CORRECTOR ) return;
    for ( int i = 0; i < c->nRow; ++i ) {
        kkt->dBoundDiag[i] = c->dualLowerInverse[i] * c->dualLowerInverse[i] +
                             c->dualUpperInverse[i] * c->dualUpperInverse[i];
    }
}

int bnd_cone_interior( bound_cone *c, double tau, const double *y ) {
    bnd_cone_update(c, tau, y);
    for ( int i = 0; i < c->nRow; ++i ) if ( c->dualLower[i] <= 0.0 ) return 0;
    for ( int i = 0; i < c->nRow; ++i ) if ( c->dualUpper[i] <= 0.0 ) return 0;
    return 1;
}

int bnd_cone_interior_expert( bound_cone *c, double cCoef, double aScal, const double *aCoef, int whichBuffer ) {
    double *l, *u;
    bnd_buffer(c, cCoef, aScal, aCoef, whichBuffer, &l, &u);
    if ( whichBuffer == BUFFER_DUALSTEP ) { l = c->dualLowerChecker; u = c->dualUpperChecker; }
    for ( int i = 0; i < c->nRow; ++i ) if ( l[i] <= 0.0 ) return 0;
    for ( int i = 0; i < c->nRow; ++i ) if ( u[i] <= 0.0 ) return 0;
    return 1;
}

hdsdp_retcode bnd_cone_barrier( bound_cone *c, double tau, const double *y, int whichBuffer, double *logdet ) {
    double *l = ( whichBuffer == BUFFER_DUALVAR ) ? c->dualLower : c->dualLowerChecker;
    double *u = ( whichBuffer == BUFFER_DUALVAR ) ? c->dualUpper : c->dualUpperChecker;
    if ( y ) bnd_cone_update(c, tau, y);
    double s = 0.0;
    for ( int i = 0; i < c->nRow; ++i ) s += log(l[i]) + log(u[i]);
    if ( s != s ) return HDSDP_RETCODE_FAILED;
    *logdet = s;
    return HDSDP_RETCODE_OK;
}

int bnd_cone_axpy_check( bound_cone *c, double step, int whichBuffer ) {
    double *l, *u;
    if ( whichBuffer == BUFFER_DUALVAR ) {
        l = c->dualLower; u = c->dualUpper;
    } else {
        l = c->dualLowerChecker; u = c->dualUpperChecker;
        memcpy(c->dualLowerChecker, c->dualLower, sizeof(double) * c->nRow);
        memcpy(c->dualUpperChecker, c->dualUpper, sizeof(double) * c->nRow);
    }
    for ( int i = 0; i < c->nRow; ++i ) {
        l[i] += step * c->dualLowerStep[i];
        u[i] += step * c->dualUpperStep[i];
        if ( l[i] <= 0.0 || u[i] <= 0.0 ) return 0;
    }
    return 1;
}

/* interface/hdsdp_conic_bound.c:400-418 */
void bnd_cone_primal( bound_cone *c, double mu, const double *y, const double *dy, double *xl, double *xu ) {
    for ( int i = 0; i < c->nRow; ++i ) {
        double ds = -dy[i];
        double su = c->dBoundUp - y[i];
        double sl = y[i] - c->dBoundLow;
        xl[i] = mu * ( 1.0 / sl - ds / (sl * sl) );
        xu[i] = mu * ( 1.0 / su + ds / (su * su) );
    }
}
