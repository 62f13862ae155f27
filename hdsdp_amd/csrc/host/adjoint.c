/*
 * adjoint.c -- the centred dual point at a fixed barrier parameter and its adjoint
 * (no counterpart in the reference).
 *
 * For the solver's data (b, C, A_1..A_m) and mu > 0, the centred dual point
 * y_mu(b, C) maximises b^T y + mu log det S(y), S(y) = C - sum_i y_i A_i, i.e. it
 * solves the barrier optimality conditions
 *     F(y) = b - mu A(S^-1) = 0,   A(W)_i = <A_i, W>.
 * The interior-point method's dual box is not part of this map.  dF/dy = -mu M with
 * M_ij = <A_i, S^-1 A_j S^-1> (the Schur matrix), dF/db = I and, along a symmetric
 * direction dC, dF/dC = mu A(S^-1 dC S^-1).  For a loss L(y_mu) with g = dL/dy:
 *     M lambda = g,   dL/db = lambda / mu,   dL/dC = S^-1 (sum_i lambda_i A_i) S^-1.
 *
 * Units: everything here is in the units of the data handed to the solver.  The
 * device holds the pre-solve scaled cone (C' = objScal C, b' = rhsScal b), so
 * y' = objScal y, S' = objScal S, A(S'^-1) = A(S^-1) / objScal and
 * M' = M / objScal^2; the conversions below follow from these.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hdsdp_host.h"

static double dot( int n, const double *x, const double *y ) {
    double s = 0.0;
    for ( int i = 0; i < n; ++i ) s += x[i] * y[i];
    return s;
}

/* Dual factor and Schur matrix at y; g = b - mu A(S^-1).  *isPd = 0 when S(y) is not
 * positive definite (nothing else is computed then). */
static hdsdp_retcode centre_eval( hdsdp *s, double mu, const double *y, double *g, double *work,
                                  int *isPd ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    hdsdp_dev *dev = s->sdp->dev;
    const int m = s->nRows;
    const double objScal = get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
    const double rhsScal = get_dbl_feature(s, DBL_FEATURE_RHSSCALING);
    double *asinv = work, *asinvrd = work + m, traceSinv = 0.0;
    s->adjReady = 0;
    for ( int i = 0; i < m; ++i ) work[i] = objScal * y[i];
    HCALL(hdsdp_dev_update_buffer(dev, 1.0, -1.0, work, 0.0, 0));
    HCALL(hdsdp_dev_factorize(dev, 0, isPd));
    if ( !*isPd ) goto exit_cleanup;
    HCALL(hdsdp_dev_schur(dev, 0, 0.0, asinv, asinvrd, &traceSinv));
    for ( int i = 0; i < m; ++i ) g[i] = s->rowRHS[i] / rhsScal - mu * objScal * asinv[i];
    memcpy(s->adjY, y, sizeof(double) * m);
    s->adjMu = mu;
    s->adjReady = 1;
exit_cleanup:
    return retcode;
}

hdsdp_retcode HDSDPGetBarrierPoint( hdsdp *s, double *mu, double *y ) {
    if ( !s->sdp || s->nIterCount <= 0 ) return HDSDP_RETCODE_FAILED;
    const double objScal = get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
    const double rhsScal = get_dbl_feature(s, DBL_FEATURE_RHSSCALING);
    if ( mu ) *mu = s->dBarrierMu / (objScal * rhsScal);
    if ( y )
        for ( int i = 0; i < s->nRows; ++i ) y[i] = s->dRowDual[i] / objScal;
    return HDSDP_RETCODE_OK;
}

/* Damped Newton on F(y) = 0 (Nesterov & Nemirovskii's step 1 / (1 + delta) for the
 * self-concordant barrier -b^T y / mu - log det S(y) while the Newton decrement
 * delta = sqrt(dy^T M dy) exceeds 1/4, full steps after), until
 * ||F(y)|| <= tol (1 + ||b||), or until full steps stop reducing the residual: at
 * small mu, S is near-singular (cond ~ 1/mu) and A(S^-1) carries a rounding error of
 * about eps cond(S), below which F cannot be resolved.  The iterate with the
 * smallest residual is kept; FAILED if that residual is above tolFloor. */
hdsdp_retcode HDSDPCenter( hdsdp *s, double mu, double *y, int maxIter, double tol, int *nIter,
                           double *relRes ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    const int m = s->nRows;
    const double tolFloor = 1e-05;
    double *g = NULL, *dy = NULL, *yt = NULL, *work = NULL, *ybest = NULL;
    int it = 0, isPd = 0, stall = 0;
    double res = HDSDP_INFINITY, best = HDSDP_INFINITY, prev = HDSDP_INFINITY;
    if ( !s->sdp || !(mu > 0.0) ) return HDSDP_RETCODE_FAILED;
    if ( !s->adjY ) HALLOC(s->adjY, double, m);
    HALLOC(g, double, m);
    HALLOC(dy, double, m);
    HALLOC(yt, double, m);
    HALLOC(work, double, 2 * m);
    HALLOC(ybest, double, m);
    const double objScal = get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
    const double rhsScal = get_dbl_feature(s, DBL_FEATURE_RHSSCALING);
    double bnorm = 0.0;
    for ( int i = 0; i < m; ++i ) bnorm += (s->rowRHS[i] / rhsScal) * (s->rowRHS[i] / rhsScal);
    bnorm = sqrt(bnorm);

    HCALL(centre_eval(s, mu, y, g, work, &isPd));
    if ( !isPd ) {
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    for ( it = 0;; ++it ) {
        res = sqrt(dot(m, g, g)) / (1.0 + bnorm);
        if ( res < best ) {
            best = res;
            memcpy(ybest, y, sizeof(double) * m);
        }
        if ( res <= tol || it >= maxIter ) break;
        /* after the quadratic phase: three full steps without halving the residual */
        stall = (res > 0.5 * prev) ? stall + 1 : 0;
        if ( stall >= 3 ) break;
        prev = res;
        /* M dy = g / mu in the data's units: M' dy = g / (mu objScal^2) */
        for ( int i = 0; i < m; ++i ) work[i] = g[i] / (mu * objScal * objScal);
        HCALL(hdsdp_dev_kkt_solve_exact(s->sdp->dev, work, dy));
        double delta = sqrt(fmax(dot(m, dy, g) / mu, 0.0));
        double alpha = (delta > 0.25) ? 1.0 / (1.0 + delta) : 1.0;
        if ( alpha < 1.0 ) {
            stall = 0;
            prev = HDSDP_INFINITY;
        }
        if ( getenv("HDSDP_TRACE_CENTER") )
            fprintf(stderr, "[hdsdp] centre step %d: residual %.3e decrement %.3e step %.3e\n", it, res, delta, alpha);
        for ( int k = 0;; ++k ) {
            for ( int i = 0; i < m; ++i ) yt[i] = y[i] + alpha * dy[i];
            HCALL(centre_eval(s, mu, yt, g, work, &isPd));
            if ( isPd ) break;
            if ( k >= 30 ) {
                retcode = HDSDP_RETCODE_FAILED;
                goto exit_cleanup;
            }
            alpha *= 0.5;
        }
        memcpy(y, yt, sizeof(double) * m);
    }
    if ( best < res ) {
        /* back to the best iterate, leaving its factor and Schur matrix on the device */
        memcpy(y, ybest, sizeof(double) * m);
        HCALL(centre_eval(s, mu, y, g, work, &isPd));
        res = sqrt(dot(m, g, g)) / (1.0 + bnorm);
    }
    if ( res > HMAX(tol, tolFloor) ) {
        retcode = HDSDP_RETCODE_FAILED;
        goto exit_cleanup;
    }
    /* the Newton system at the centred point, solved but not stepped: its FP64 factor stays
     * on the device with the Schur matrix of (mu, y), and HDSDPAdjoint at the same point
     * solves with it instead of factoring again */
    for ( int i = 0; i < m; ++i ) work[i] = g[i] / (mu * objScal * objScal);
    HCALL(hdsdp_dev_kkt_solve_exact(s->sdp->dev, work, dy));
exit_cleanup:
    if ( nIter ) *nIter = it;
    if ( relRes ) *relRes = res;
    free(ybest);
    free(g);
    free(dy);
    free(yt);
    free(work);
    return retcode;
}

hdsdp_retcode HDSDPAdjoint( hdsdp *s, double mu, const double *y, const double *dLdy, double *dLdb,
                            double *dLdC ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    const int m = s->nRows;
    double *g = NULL, *work = NULL;
    if ( !s->sdp || !(mu > 0.0) ) return HDSDP_RETCODE_FAILED;
    if ( !s->adjY ) HALLOC(s->adjY, double, m);
    if ( !s->adjReady || s->adjMu != mu || memcmp(s->adjY, y, sizeof(double) * m) != 0 ) {
        int isPd = 0;
        HALLOC(g, double, m);
        HALLOC(work, double, 2 * m);
        HCALL(centre_eval(s, mu, y, g, work, &isPd));
        if ( !isPd ) {
            retcode = HDSDP_RETCODE_FAILED;
            goto exit_cleanup;
        }
    }
    {
        /* lambda' = M'^-1 dLdy = objScal^2 lambda: dLdb = lambda / mu = lambda' / (objScal^2 mu)
         * and S^-1 (sum lambda_i A_i) S^-1 = S'^-1 (sum lambda'_i A_i) S'^-1 */
        const double objScal = get_dbl_feature(s, DBL_FEATURE_OBJSCALING);
        HCALL(hdsdp_dev_adjoint(s->sdp->dev, mu * objScal * objScal, dLdy, dLdb, dLdC));
    }
exit_cleanup:
    free(g);
    free(work);
    return retcode;
}
