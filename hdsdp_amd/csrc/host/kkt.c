/*
 * kkt.c -- Schur-system bookkeeping (interface/hdsdp_schur.c).
 *
 * M itself, its preconditioner and the CG iterations live in the device
 * state of the SDP block; the host keeps the right-hand-side vectors that the
 * IPM reads (dASinvVec, dASinvRdSinvVec) and adds the bound cone's terms.
 */
#include "hdsdp_host.h"

#define KKT_ACCURACY (1e-12)

hdsdp_retcode kkt_create( host_kkt **pkkt, int nRow ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    host_kkt *kkt = NULL;
    HALLOC(kkt, host_kkt, 1);
    *pkkt = kkt;
    kkt->nRow = nRow;
    HALLOC(kkt->dASinvVec, double, nRow);
    HALLOC(kkt->dASinvRdSinvVec, double, nRow);
    HALLOC(kkt->dASinvCSinvVec, double, nRow);
    HALLOC(kkt->dBoundDiag, double, nRow);
exit_cleanup:
    return retcode;
}

void kkt_destroy( host_kkt **pkkt ) {
    if ( !pkkt || !*pkkt ) return;
    host_kkt *kkt = *pkkt;
    free(kkt->dASinvVec);
    free(kkt->dASinvRdSinvVec);
    free(kkt->dASinvCSinvVec);
    free(kkt->dBoundDiag);
    free(kkt);
    *pkkt = NULL;
}

/* CG accuracy by problem size (interface/hdsdp_schur.c:21-35). */
hdsdp_retcode kkt_init( host_kkt *kkt, sdp_cone *cone ) {
    double acc = KKT_ACCURACY;
    int maxIter = -1;
    if ( kkt->nRow > 20000 ) {
        acc = KKT_ACCURACY * 100.0;
        maxIter = 500;
    } else if ( kkt->nRow > 15000 ) {
        acc = KKT_ACCURACY * 50.0;
        maxIter = 450;
    } else if ( kkt->nRow > 5000 ) {
        acc = KKT_ACCURACY * 5.0;
        maxIter = 120;
    }
    return hdsdp_dev_kkt_params(cone->dev, 5.0 * acc, acc, maxIter, -1);
}

/* HKKTClean + cone build (interface/hdsdp_schur.c:141-165, 256-268). */
hdsdp_retcode kkt_build( hdsdp *s, int typeKKT ) {
    host_kkt *kkt = s->kkt;
    memset(kkt->dASinvVec, 0, sizeof(double) * kkt->nRow);
    memset(kkt->dASinvRdSinvVec, 0, sizeof(double) * kkt->nRow);
    memset(kkt->dASinvCSinvVec, 0, sizeof(double) * kkt->nRow);
    kkt->dTraceSinv = 0.0;
    kkt->dCSinv = kkt->dCSinvCSinv = kkt->dCSinvRdSinv = 0.0;
    return sdp_cone_schur(s->sdp, kkt, typeKKT);
}

/* HKKTBuildUpExtraCone with the scalar bound cone. */
hdsdp_retcode kkt_build_bound( hdsdp *s, int typeKKT ) {
    bnd_cone_schur(s->bnd, s->kkt, typeKKT);
    if ( typeKKT == KKT_TYPE_CORRECTOR ) return HDSDP_RETCODE_OK;
    return hdsdp_dev_kkt_add_diag(s->sdp->dev, s->kkt->dBoundDiag);
}

/* HFpLinsysNumeric / HFpLinsysSolve (linalg/hdsdp_linsolver.c:1367-1425) log the
 * switch of the Schur system to the indefinite solver */
hdsdp_retcode kkt_factorize( hdsdp *s ) {
    int before = hdsdp_dev_kkt_uses_ldl(s->sdp->dev);
    hdsdp_retcode rc = hdsdp_dev_kkt_factorize(s->sdp->dev);
    if ( !before && hdsdp_dev_kkt_uses_ldl(s->sdp->dev) ) hlog(s, "KKT system is almost indefinite. Switch to LDL. \n");
    return rc;
}

hdsdp_retcode kkt_solve( hdsdp *s, const double *rhs, double *sol ) {
    int before = hdsdp_dev_kkt_uses_ldl(s->sdp->dev);
    hdsdp_retcode rc = hdsdp_dev_kkt_solve(s->sdp->dev, rhs, sol);
    if ( !before && hdsdp_dev_kkt_uses_ldl(s->sdp->dev) ) hlog(s, "KKT system is unstable. Switch to LDL. \n");
    return rc;
}

hdsdp_retcode kkt_regularize( hdsdp *s, double reg ) {
    return hdsdp_dev_kkt_regularize(s->sdp->dev, reg);
}
