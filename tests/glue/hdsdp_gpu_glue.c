/*
 * hdsdp_gpu_glue.c -- TEST SHIM: the reference's own host solver driving the
 * MI355X device ABI (include/hdsdp_gpu.h) through its plugin tables.
 *
 * Linked into one executable with the reference's unmodified object files
 * (interface/*.c, linalg/*.c, external/hdsdp_cs.c, built by oracle/Makefile.ref)
 * and with hdsdp_amd/lib/libhdsdp_amd.so as it is built for the product.
 * Two link-time wraps install the device slots without touching a reference
 * source file:
 *
 *   HConeSetData (interface/hdsdp_conic.c:50-184) -- after the reference has
 *     filled the DENSE_SDP slot table (interface/hdsdp_conic.c:117-146), the
 *     slots that touch the dual matrices or the Schur matrix are pointed at the
 *     forwarders below.  Data-only slots (create, process, norms, statistics,
 *     symbolic Schur pattern, start / residual / perturbation scalars, and the
 *     traces of a host X handed back in) stay the reference's.
 *
 *   HKKTInit (interface/hdsdp_schur.c:181-254) -- after the reference has
 *     created its dense Schur solver, its hdsdp_linsys_fp table
 *     (linalg/def_hdsdp_linsolver.h:34-57) is pointed at a GPU kind: what a
 *     `case HDSDP_LINSYS_GPU:` in HFpLinsysCreate (linalg/hdsdp_linsolver.c:1268-1336)
 *     would set.  The CG parameters the reference chose (interface/hdsdp_schur.c:21-35)
 *     are handed to the device.
 *
 * Data flow.  M stays in device memory.  The host Schur matrix the reference
 * keeps (kktMatElem) receives only M's diagonal from each build, so that the
 * reference's own bound-cone terms (interface/hdsdp_conic_bound.c:224-229) and
 * HKKTRegularize (interface/hdsdp_schur.c:348-373) act on the true diagonal; the
 * GPU kind's numeric slot then sets the device diagonal to the host's, bit for
 * bit, with two additions (-d, then x).
 *
 * Supported: one DENSE_SDP cone, a dense Schur matrix, the dual-scaling
 * iterations (KKT types INFEASIBLE and CORRECTOR) and the self-dual embedding
 * (HOMOGENEOUS).  PSDP (KKT_TYPE_PRIMAL) returns HDSDP_RETCODE_FAILED.  Device precision:
 * HDSDP_GLUE_PRECISION=0 (FP64, default) or 1 (mixed).
 *
 * At exit one line "GLUE {...}" reports how many calls went to the device.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hdsdp.h"               /* the reference's interface/hdsdp.h */
#include "def_hdsdp_conic.h"
#include "hdsdp_conic.h"
#include "hdsdp_conic_sdp.h"
#include "def_hdsdp_schur.h"
#include "hdsdp_schur.h"
#include "def_hdsdp_linsolver.h"
#include "hdsdp_linsolver.h"
#include "def_hdsdp_sdpdata.h"

/* The device ABI is declared over the reference's hdsdp_retcode: take the
 * reference's public header in place of this library's copy of it. */
#define HDSDP_AMD_HDSDP_H
#define HDSDP_PRECISION_FP64 0
#define HDSDP_PRECISION_MIXED 1
#include "hdsdp_gpu.h"

/* the GPU Schur solver kind, next after the reference's last (HDSDP_LINSYS_DENSE_INDEFINITE) */
#define HDSDP_LINSYS_GPU ((linsys_type) (HDSDP_LINSYS_DENSE_INDEFINITE + 1))

extern hdsdp_retcode __real_HConeSetData( hdsdp_cone *HCone, user_data *usrData );
extern hdsdp_retcode __real_HKKTInit( hdsdp_kkt *HKKT, int nRow, int nCones, hdsdp_cone **cones );

typedef struct {
    hdsdp_cone_sdp_dense *host;  /* the reference's cone: coefficients and scalars */
    hdsdp_dev *dev;
    double *mDiag;               /* diagonal of M as it stands on the device */
    double *mFull;               /* m x m staging for the diagonal read-back */
} glue_cone;

typedef struct {
    glue_cone *g;
    int nCol;
    double *delta;
} glue_kkt;

static glue_cone *gCone = NULL;  /* one device cone per process */
static long gCalls[8];           /* update, factorize, ratio, schur, kktfac, kktsolve, primal, regdiag */

/* Shadow mode (HDSDP_GLUE_SHADOW=1, or 2 to print each difference): every forwarded slot also runs the reference's
 * own implementation at the same iterate, the difference is recorded and the
 * reference's answer is handed back, so the iterates stay the reference's and each
 * device call is checked against the reference at exactly the reference's inputs. */
static int gShadow = 0;
static hdsdp_cone gRef;          /* the DENSE_SDP slot table as the reference filled it */
static hdsdp_linsys_fp gRefLin;  /* the reference's Schur solver kind */
enum { SH_RATIO, SH_LOGDET, SH_ASINV, SH_ASINVRD, SH_TRACE, SH_M, SH_SOLVE, SH_X, SH_HSD, SH_N };
static const char *gShName[SH_N] = {"ratio_test", "logdet", "asinv", "asinvrd", "trace_sinv", "schur_m",
                                    "kkt_solve", "primal_x", "hsd_terms"};
static double gShMax[SH_N];      /* largest relative difference, device vs reference */
static int gShAt[SH_N];          /* the call at which it occurred */
static long gShCount[SH_N];
static long gShFlips;            /* interior-check answers that disagree */

static int gShTrace = 0;         /* HDSDP_GLUE_SHADOW=2: print every difference above 1e-9 */

static void shadow_note( int kind, double diff ) {
    gShCount[kind] += 1;
    if ( gShTrace && !(diff <= 1e-9) ) printf("[shadow] %s call %ld: %.3e\n", gShName[kind], gShCount[kind], diff);
    if ( !(diff <= gShMax[kind]) ) {  /* NaN counts as worse */
        gShMax[kind] = diff;
        gShAt[kind] = (int) gShCount[kind];
    }
}

/* relative distance of two vectors: ||a - b||_inf / max(||b||_inf, tiny) */
static double rel_dist( int n, const double *a, const double *b ) {
    double num = 0.0, den = 0.0;
    for ( int i = 0; i < n; ++i ) {
        num = fmax(num, fabs(a[i] - b[i]));
        den = fmax(den, fabs(b[i]));
    }
    return num / fmax(den, 1e-300);
}

static double rel_scalar( double a, double b ) {
    if ( a == b ) return 0.0;
    return fabs(a - b) / fmax(fabs(b), 1e-300);
}

#define GCALL(x)                                                                      \
    do {                                                                              \
        if ( (retcode = (x)) != HDSDP_RETCODE_OK ) {                                  \
            fprintf(stderr, "[glue] %s failed at %s:%d\n", #x, __FILE__, __LINE__);   \
            goto exit_cleanup;                                                        \
        }                                                                             \
    } while ( 0 )

static glue_cone *glue_of( void *cone ) {
    if ( !gCone || gCone->host != (hdsdp_cone_sdp_dense *) cone ) {
        fprintf(stderr, "[glue] slot called on a cone without device state\n");
        abort();
    }
    return gCone;
}

/* void slots of the reference cannot return an error: stop the run loudly */
static void glue_must( hdsdp_retcode rc, const char *what ) {
    if ( rc != HDSDP_RETCODE_OK ) {
        fprintf(stderr, "[glue] %s failed on the device\n", what);
        abort();
    }
}

/* A presolved coefficient (linalg/def_hdsdp_sdpdata.h:35-141) as the device ABI's view. */
static hdsdp_retcode glue_view( sdp_coeff *A, hdsdp_coeff_view *v ) {
    memset(v, 0, sizeof(*v));
    switch ( A->dataType ) {
        case SDP_COEFF_ZERO:
            v->type = HDSDP_COEFF_ZERO;
            return HDSDP_RETCODE_OK;
        case SDP_COEFF_SPARSE: {
            sdp_coeff_sparse *sp = (sdp_coeff_sparse *) A->dataMat;
            v->type = HDSDP_COEFF_SPARSE;
            v->nnz = sp->nTriMatElem;
            v->row = sp->triMatRow;
            v->col = sp->triMatCol;
            v->val = sp->triMatElem;
            return HDSDP_RETCODE_OK;
        }
        case SDP_COEFF_DENSE: {
            sdp_coeff_dense *ds = (sdp_coeff_dense *) A->dataMat;
            v->type = HDSDP_COEFF_DENSE;
            v->nnz = ds->nSDPCol * (ds->nSDPCol + 1) / 2;
            v->val = ds->dsMatElem;
            return HDSDP_RETCODE_OK;
        }
        case SDP_COEFF_SPR1: {
            sdp_coeff_spr1 *r1 = (sdp_coeff_spr1 *) A->dataMat;
            v->type = HDSDP_COEFF_SPR1;
            v->nnz = r1->nSpR1FactorElem;
            v->row = r1->spR1MatIdx;
            v->val = r1->spR1MatElem;
            v->sign = r1->spR1FactorSign;
            return HDSDP_RETCODE_OK;
        }
        case SDP_COEFF_DSR1: {
            sdp_coeff_dsr1 *r1 = (sdp_coeff_dsr1 *) A->dataMat;
            v->type = HDSDP_COEFF_DSR1;
            v->nnz = r1->nSDPCol;
            v->val = r1->r1MatFactor;
            v->sign = r1->r1FactorSign;
            return HDSDP_RETCODE_OK;
        }
    }
    return HDSDP_RETCODE_FAILED;
}

/* ---- cone slots (interface/hdsdp_conic_sdp.c) ------------------------------ */

/* conePresolveData: the reference's presolve (rank-one detection, KKT ordering),
 * then the presolved coefficients go to the device once. */
static hdsdp_retcode glue_presolve( void *cone ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    hdsdp_cone_sdp_dense *c = (hdsdp_cone_sdp_dense *) cone;
    hdsdp_coeff_view *views = NULL;
    GCALL(sdpDenseConePresolveImpl(c));
    if ( gCone ) {
        fprintf(stderr, "[glue] one device cone per process\n");
        return HDSDP_RETCODE_FAILED;
    }
    gCone = (glue_cone *) calloc(1, sizeof(glue_cone));
    if ( !gCone ) return HDSDP_RETCODE_MEMORY;
    gCone->host = c;
    const char *p = getenv("HDSDP_GLUE_PRECISION");
    int precision = ( p && atoi(p) == 1 ) ? HDSDP_PRECISION_MIXED : HDSDP_PRECISION_FP64;
    GCALL(hdsdp_dev_create(&gCone->dev, c->nRow, c->nCol, precision));
    views = (hdsdp_coeff_view *) calloc((size_t) c->nRow + 1, sizeof(hdsdp_coeff_view));
    gCone->mDiag = (double *) calloc((size_t) c->nRow, sizeof(double));
    gCone->mFull = (double *) malloc(sizeof(double) * (size_t) c->nRow * (size_t) c->nRow);
    if ( !views || !gCone->mDiag || !gCone->mFull ) {
        retcode = HDSDP_RETCODE_MEMORY;
        goto exit_cleanup;
    }
    GCALL(glue_view(c->sdpObj, &views[0]));
    for ( int i = 0; i < c->nRow; ++i ) GCALL(glue_view(c->sdpRow[i], &views[i + 1]));
    GCALL(hdsdp_dev_load(gCone->dev, views));
    printf("[glue] DENSE_SDP cone n = %d, m = %d on the device (precision %d)\n", c->nCol, c->nRow, precision);
exit_cleanup:
    free(views);
    return retcode;
}

static void glue_destroy( void **pcone ) {
    if ( gCone && pcone && *pcone == (void *) gCone->host ) {
        hdsdp_dev_destroy(&gCone->dev);
        free(gCone->mDiag);
        free(gCone->mFull);
        free(gCone);
        gCone = NULL;
    }
    sdpDenseConeDestroyImpl((hdsdp_cone_sdp_dense **) pcone);
}

/* coneScal (1600-1604): the host copy of C keeps the reference's norms and traces */
static void glue_scal( void *cone, double s ) {
    sdpDenseConeScal((hdsdp_cone_sdp_dense *) cone, s);
    glue_must(hdsdp_dev_scale_obj(glue_of(cone)->dev, s), "scale C");
}

/* sdpDenseConeIUpdateBuffer (343-402): B = eye I + aScal A^T y + cCoef C, the
 * perturbation added to every buffer but the step */
static hdsdp_retcode glue_buffer( glue_cone *g, double cCoef, double aScal, double *y, double eye, int which ) {
    if ( which != BUFFER_DUALSTEP ) eye += g->host->dualPerturb;
    gCalls[0] += 1;
    return hdsdp_dev_update_buffer(g->dev, cCoef, aScal, y, eye, which);
}

/* coneUpdate (1612-1629) */
static void glue_update( void *cone, double tau, double *y ) {
    glue_cone *g = glue_of(cone);
    glue_must(glue_buffer(g, tau, -1.0, y, -g->host->dualResidual, BUFFER_DUALVAR), "update S");
    if ( gShadow ) gRef.coneUpdate(cone, tau, y);
}

/* coneRatioTest (1638-1682) */
static hdsdp_retcode glue_ratio_test( void *cone, double dTau, double *dy, double ada, int which, double *maxStep ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    glue_cone *g = glue_of(cone);
    GCALL(glue_buffer(g, dTau, -1.0, dy, ada * g->host->dualResidual, BUFFER_DUALSTEP));
    gCalls[2] += 1;
    GCALL(hdsdp_dev_ratio_test(g->dev, which, maxStep));
    if ( gShadow ) {
        double dev = *maxStep;
        GCALL(gRef.coneRatioTest(cone, dTau, dy, ada, which, maxStep));
        shadow_note(SH_RATIO, rel_scalar(dev, *maxStep));
    }
exit_cleanup:
    return retcode;
}

/* coneBuildSchur (1722-1804): M on the device, the vectors and tr S^-1 added to the
 * host KKT as the reference's cone adds them, M's diagonal into the host diagonal. */
static hdsdp_retcode glue_schur( void *cone, void *kkt, int typeKKT ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    glue_cone *g = glue_of(cone);
    hdsdp_kkt *K = (hdsdp_kkt *) kkt;
    const int m = g->host->nRow;
    double *asinv = NULL, *asinvrd = NULL, *acs = NULL, tr = 0.0, cs = 0.0, csc = 0.0, csrd = 0.0;
    const int hsd = ( typeKKT == KKT_TYPE_HOMOGENEOUS );
    if ( typeKKT != KKT_TYPE_INFEASIBLE && typeKKT != KKT_TYPE_CORRECTOR && !hsd ) {
        fprintf(stderr, "[glue] Schur type %d is not forwarded\n", typeKKT);
        return HDSDP_RETCODE_FAILED;
    }
    asinv = (double *) calloc((size_t) m, sizeof(double));
    asinvrd = (double *) calloc((size_t) m, sizeof(double));
    acs = (double *) calloc((size_t) m, sizeof(double));
    if ( !asinv || !asinvrd || !acs ) {
        retcode = HDSDP_RETCODE_MEMORY;
        goto exit_cleanup;
    }
    gCalls[3] += 1;
    GCALL(hdsdp_dev_schur(g->dev, typeKKT, g->host->dualResidual, asinv, asinvrd, &tr));
    if ( typeKKT != KKT_TYPE_CORRECTOR ) {
        GCALL(hdsdp_dev_get_schur(g->dev, g->mFull));
        for ( int i = 0; i < m; ++i ) g->mDiag[i] = g->mFull[(size_t) i * m + i];
    }
    /* the embedding's C terms (sdpDenseConeIGetHSDComponents and the dASinvCSinvVec
     * entries, interface/hdsdp_conic_sdp.c:737-760, 983-1029) */
    if ( hsd ) GCALL(hdsdp_dev_hsd(g->dev, g->host->dualResidual, acs, &cs, &csc, &csrd));
    if ( gShadow ) {
        /* the reference's build adds into zeroed vectors and, for a new M, a zeroed matrix */
        for ( int i = 0; i < m; ++i ) {
            if ( K->dASinvVec[i] != 0.0 || K->dASinvRdSinvVec[i] != 0.0 ) {
                fprintf(stderr, "[glue] shadow mode needs the SDP cone's build first\n");
                retcode = HDSDP_RETCODE_FAILED;
                goto exit_cleanup;
            }
        }
        const double tr0 = K->dTraceSinv, cs0 = K->dCSinv, csc0 = K->dCSinvCSinv;
        GCALL(gRef.coneBuildSchur(cone, kkt, typeKKT));
        if ( hsd ) {
            shadow_note(SH_HSD, rel_dist(m, acs, K->dASinvCSinvVec));
            shadow_note(SH_HSD, rel_scalar(cs0 + cs, K->dCSinv));
            shadow_note(SH_HSD, rel_scalar(csc0 + csc, K->dCSinvCSinv));
        }
        shadow_note(SH_ASINV, rel_dist(m, asinv, K->dASinvVec));
        if ( g->host->dualResidual != 0.0 ) shadow_note(SH_ASINVRD, rel_dist(m, asinvrd, K->dASinvRdSinvVec));
        if ( typeKKT != KKT_TYPE_CORRECTOR ) {
            shadow_note(SH_TRACE, rel_scalar(tr0 + tr, K->dTraceSinv));
            /* the reference fills one triangle of its dense M: compare where it wrote */
            double num = 0.0, den = 0.0;
            for ( int j = 0; j < m; ++j ) {
                for ( int i = 0; i < m; ++i ) {
                    const double h = K->kktMatElem[(size_t) j * m + i];
                    if ( h == 0.0 ) continue;
                    num = fmax(num, fabs(g->mFull[(size_t) j * m + i] - h));
                    den = fmax(den, fabs(h));
                }
            }
            shadow_note(SH_M, num / fmax(den, 1e-300));
        }
        goto exit_cleanup;
    }
    for ( int i = 0; i < m; ++i ) {
        K->dASinvVec[i] += asinv[i];
        K->dASinvRdSinvVec[i] += asinvrd[i];
    }
    K->dTraceSinv += tr;
    if ( typeKKT != KKT_TYPE_CORRECTOR ) {
        for ( int i = 0; i < m; ++i ) *K->kktDiag[i] += g->mDiag[i];
    }
    if ( hsd ) {
        for ( int i = 0; i < m; ++i ) K->dASinvCSinvVec[i] += acs[i];
        K->dCSinv += cs;
        K->dCSinvCSinv += csc;
        if ( g->host->dualResidual ) K->dCSinvRdSinv += csrd;
    }
exit_cleanup:
    free(asinv);
    free(asinvrd);
    free(acs);
    return retcode;
}

static hdsdp_retcode glue_schur_fixed( void *cone, void *kkt, int typeKKT, int strategy ) {
    (void) cone; (void) kkt; (void) strategy;
    fprintf(stderr, "[glue] fixed-strategy Schur (type %d) is not forwarded\n", typeKKT);
    return HDSDP_RETCODE_FAILED;
}

/* coneInteriorCheck (2105-2113) */
static hdsdp_retcode glue_interior( void *cone, double tau, double *y, int *isInterior ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    glue_cone *g = glue_of(cone);
    GCALL(glue_buffer(g, tau, -1.0, y, -g->host->dualResidual, BUFFER_DUALVAR));
    gCalls[1] += 1;
    GCALL(hdsdp_dev_factorize(g->dev, BUFFER_DUALVAR, isInterior));
    if ( gShadow ) {
        int dev = *isInterior;
        GCALL(gRef.coneInteriorCheck(cone, tau, y, isInterior));
        gShFlips += ( dev != *isInterior );
    }
exit_cleanup:
    return retcode;
}

/* coneInteriorCheckExpert (2125-2141) */
static hdsdp_retcode glue_interior_expert( void *cone, double cCoef, double aScal, double *y, double eye, int which,
                                           int *isInterior ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    glue_cone *g = glue_of(cone);
    GCALL(glue_buffer(g, cCoef, aScal, y, eye, which));
    gCalls[1] += 1;
    GCALL(hdsdp_dev_factorize(g->dev, which == BUFFER_DUALVAR ? BUFFER_DUALVAR : BUFFER_DUALCHECK, isInterior));
    if ( gShadow ) {
        int dev = *isInterior;
        GCALL(gRef.coneInteriorCheckExpert(cone, cCoef, aScal, y, eye, which, isInterior));
        gShFlips += ( dev != *isInterior );
    }
exit_cleanup:
    return retcode;
}

/* coneGetBarrier (2185-2224) */
static hdsdp_retcode glue_barrier( void *cone, double tau, double *y, int which, double *logdet ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    glue_cone *g = glue_of(cone);
    if ( y ) {
        int isPsd = 0;
        GCALL(glue_buffer(g, tau, -1.0, y, -g->host->dualResidual, BUFFER_DUALVAR));
        gCalls[1] += 1;
        GCALL(hdsdp_dev_factorize(g->dev, BUFFER_DUALVAR, &isPsd));
        if ( !isPsd && !gShadow ) {
            /* the reference's HFpLinsysNumeric fails on a matrix dpotrf rejects */
            retcode = HDSDP_RETCODE_FAILED;
            goto exit_cleanup;
        }
        gShFlips += ( gShadow && !isPsd );
    }
    GCALL(hdsdp_dev_logdet(g->dev, which, logdet));
    if ( gShadow ) {
        double dev = *logdet;
        GCALL(gRef.coneGetBarrier(cone, tau, y, which, logdet));
        shadow_note(SH_LOGDET, rel_scalar(dev, *logdet));
    }
exit_cleanup:
    return retcode;
}

/* coneAxpyBufferAndCheck (2266-2295) */
static hdsdp_retcode glue_axpy_check( void *cone, double step, int which, int *isInterior ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    gCalls[1] += 1;
    GCALL(hdsdp_dev_axpy_check(glue_of(cone)->dev, step, which, isInterior));
    if ( gShadow ) {
        int dev = *isInterior;
        GCALL(gRef.coneAxpyBufferAndCheck(cone, step, which, isInterior));
        gShFlips += ( dev != *isInterior );
    }
exit_cleanup:
    return retcode;
}

/* conePRecover (2328-2379): X on the device, copied into the caller's n x n array */
static void glue_primal( void *cone, double mu, double *y, double *dy, double *X, double *aux ) {
    (void) aux;
    glue_cone *g = glue_of(cone);
    int ok = 0;
    gCalls[6] += 1;
    glue_must(glue_buffer(g, 1.0, -1.0, y, 0.0, BUFFER_DUALCHECK), "recovery buffer");
    glue_must(hdsdp_dev_primal(g->dev, mu, y, dy, &ok), "primal recovery");
    if ( gShadow ) {
        const int n = g->host->nCol;
        double *Xd = (double *) malloc(sizeof(double) * (size_t) n * n);
        if ( !Xd ) abort();
        if ( ok ) glue_must(hdsdp_dev_get_primal(g->dev, Xd), "primal read-back");
        gRef.conePRecover(cone, mu, y, dy, X, aux);
        if ( ok ) shadow_note(SH_X, rel_dist(n * n, Xd, X));
        else gShFlips += 1;
        free(Xd);
        return;
    }
    if ( !ok ) {
        printf("Recovery step is infeasible\n");
        return;
    }
    glue_must(hdsdp_dev_get_primal(g->dev, X), "primal read-back");
}

/* coneDRecover (2427-2436) */
static void glue_dual( void *cone, double *S, double *dummy ) {
    (void) dummy;
    if ( gShadow ) {
        gRef.coneDRecover(cone, S, dummy);
        return;
    }
    glue_must(hdsdp_dev_get_dual(glue_of(cone)->dev, S), "dual read-back");
}

extern hdsdp_retcode __wrap_HConeSetData( hdsdp_cone *HCone, user_data *usrData ) {
    hdsdp_retcode retcode = __real_HConeSetData(HCone, usrData);
    if ( retcode != HDSDP_RETCODE_OK || HCone->cone != HDSDP_CONETYPE_DENSE_SDP ) return retcode;
    const char *sh = getenv("HDSDP_GLUE_SHADOW");
    gShadow = ( sh && atoi(sh) >= 1 );
    gShTrace = ( sh && atoi(sh) == 2 );
    gRef = *HCone;
    HCone->conePresolveData = glue_presolve;
    HCone->coneDestroyData = glue_destroy;
    HCone->coneScal = glue_scal;
    HCone->coneUpdate = glue_update;
    HCone->coneRatioTest = glue_ratio_test;
    HCone->coneBuildSchur = glue_schur;
    HCone->coneBuildSchurFixed = glue_schur_fixed;
    HCone->coneInteriorCheck = glue_interior;
    HCone->coneInteriorCheckExpert = glue_interior_expert;
    HCone->coneGetBarrier = glue_barrier;
    HCone->coneAxpyBufferAndCheck = glue_axpy_check;
    HCone->conePRecover = glue_primal;
    HCone->coneDRecover = glue_dual;
    return retcode;
}

/* ---- the GPU kind of hdsdp_linsys_fp (linalg/def_hdsdp_linsolver.h:34-57) ---- */

static hdsdp_retcode gk_create( void **pchol, int nCol ) {
    (void) pchol; (void) nCol;
    return HDSDP_RETCODE_FAILED;  /* installed by HKKTInit's wrap, never created by kind */
}

static void gk_set_param( void *chol, void *params ) {
    glue_kkt *k = (glue_kkt *) chol;
    iterative_params *p = (iterative_params *) params;
    glue_must(hdsdp_dev_kkt_params(k->g->dev, p->relTol, p->absTol, p->maxIter, p->nRestartFreq), "CG parameters");
}

static hdsdp_retcode gk_symbolic( void *chol, int *beg, int *idx ) {
    (void) chol; (void) beg; (void) idx;
    return HDSDP_RETCODE_OK;
}

/* conjGradLinSolverNumeric (768-778): the host diagonal goes back to M on the
 * device, then the preconditioner is built there.  The device only adds to its
 * diagonal (M_ii += v_i, one IEEE addition), so the diagonal is first cleared by
 * adding -d (d + (-d) = 0 exactly) and then set by adding x (0 + x = x): the device
 * diagonal equals the host's bit for bit. */
static hdsdp_retcode gk_numeric_dev( glue_kkt *k, double *elem ) {
    glue_cone *g = k->g;
    const int m = k->nCol;
    int changed = 0;
    for ( int i = 0; i < m; ++i ) changed |= ( elem[(size_t) i * m + i] != g->mDiag[i] );
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    if ( changed ) {
        gCalls[7] += 1;
        for ( int i = 0; i < m; ++i ) k->delta[i] = -g->mDiag[i];
        GCALL(hdsdp_dev_kkt_add_diag(g->dev, k->delta));
        for ( int i = 0; i < m; ++i ) g->mDiag[i] = k->delta[i] = elem[(size_t) i * m + i];
        GCALL(hdsdp_dev_kkt_add_diag(g->dev, k->delta));
    }
    gCalls[4] += 1;
    GCALL(hdsdp_dev_kkt_factorize(g->dev));
exit_cleanup:
    return retcode;
}

static hdsdp_retcode gk_numeric( void *chol, int *beg, int *idx, double *elem ) {
    (void) beg; (void) idx;
    return gk_numeric_dev((glue_kkt *) chol, elem);
}

static hdsdp_retcode gk_psd_check( void *chol, int *beg, int *idx, double *elem, int *isPsd ) {
    (void) chol; (void) beg; (void) idx; (void) elem; (void) isPsd;
    return HDSDP_RETCODE_FAILED;  /* as the CG kind (conjGradLinSolverPsdCheck) */
}

static void gk_noop_solve( void *chol, int nRhs, double *rhs, double *sol ) {
    (void) chol; (void) nRhs; (void) rhs; (void) sol;
}

/* conjGradLinSolverSolveN (999-1016): sol = NULL overwrites rhs */
static hdsdp_retcode gk_solve_dev( glue_kkt *k, int nRhs, double *rhs, double *sol ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    for ( int r = 0; r < nRhs; ++r ) {
        double *b = rhs + (size_t) r * k->nCol;
        double *x = sol ? sol + (size_t) r * k->nCol : b;
        gCalls[5] += 1;
        GCALL(hdsdp_dev_kkt_solve(k->g->dev, b, x));
    }
exit_cleanup:
    return retcode;
}

static hdsdp_retcode gk_solve( void *chol, int nRhs, double *rhs, double *sol ) {
    return gk_solve_dev((glue_kkt *) chol, nRhs, rhs, sol);
}

/* Shadow mode: the reference's CG solver stays the Schur solver; the device factors
 * and solves beside it on the same matrix (its diagonal brought to the host's). */
static glue_kkt *gShadowKkt = NULL;

static hdsdp_retcode gk_shadow_numeric( void *chol, int *beg, int *idx, double *elem ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    GCALL(gk_numeric_dev(gShadowKkt, elem));
    GCALL(gRefLin.cholNumeric(chol, beg, idx, elem));
exit_cleanup:
    return retcode;
}

static hdsdp_retcode gk_shadow_solve( void *chol, int nRhs, double *rhs, double *sol ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    const int m = gShadowKkt->nCol;
    double *xd = (double *) malloc(sizeof(double) * (size_t) m * nRhs);
    if ( !xd ) return HDSDP_RETCODE_MEMORY;
    GCALL(gk_solve_dev(gShadowKkt, nRhs, rhs, xd));
    GCALL(gRefLin.cholSolve(chol, nRhs, rhs, sol));
    shadow_note(SH_SOLVE, rel_dist(m * nRhs, xd, sol ? sol : rhs));
exit_cleanup:
    free(xd);
    return retcode;
}

static hdsdp_retcode gk_get_diag( void *chol, double *diag ) {
    (void) chol; (void) diag;
    return HDSDP_RETCODE_FAILED;
}

static void gk_invert( void *chol, double *a, double *b ) {
    (void) chol; (void) a; (void) b;
}

static void gk_destroy( void **pchol ) {
    if ( !pchol || !*pchol ) return;
    glue_kkt *k = (glue_kkt *) *pchol;
    free(k->delta);
    free(k);
    *pchol = NULL;
}

extern hdsdp_retcode __wrap_HKKTInit( hdsdp_kkt *HKKT, int nRow, int nCones, hdsdp_cone **cones ) {
    hdsdp_retcode retcode = __real_HKKTInit(HKKT, nRow, nCones, cones);
    if ( retcode != HDSDP_RETCODE_OK ) return retcode;
    if ( !gCone ) return retcode;  /* no device cone: the reference's own solver stays */
    if ( nCones != 1 || HKKT->isKKTSparse || HKKT->kktM->LinType != HDSDP_LINSYS_DENSE_ITERATIVE ||
         cones[0]->coneData != (void *) gCone->host ) {
        fprintf(stderr, "[glue] only one dense SDP cone with a dense Schur matrix is forwarded\n");
        return HDSDP_RETCODE_FAILED;
    }
    hdsdp_linsys_fp *L = HKKT->kktM;
    iterative_linsys *cg = (iterative_linsys *) L->chol;
    glue_kkt *k = (glue_kkt *) calloc(1, sizeof(glue_kkt));
    if ( !k ) return HDSDP_RETCODE_MEMORY;
    k->g = gCone;
    k->nCol = nRow;
    k->delta = (double *) calloc((size_t) nRow, sizeof(double));
    if ( !k->delta ) {
        free(k);
        return HDSDP_RETCODE_MEMORY;
    }
    /* the parameters HKKTIAllocDenseKKT set on the CG kind, handed to the device */
    iterative_params p = cg->params;
    if ( gShadow ) {
        gRefLin = *L;
        gShadowKkt = k;
        L->cholNumeric = gk_shadow_numeric;
        L->cholSolve = gk_shadow_solve;
        gk_set_param(k, &p);
        printf("[glue] shadow mode: the reference's own slots run beside the device and steer\n");
        return retcode;
    }
    L->cholDestroy(&L->chol);
    L->LinType = HDSDP_LINSYS_GPU;
    L->chol = k;
    L->cholCreate = gk_create;
    L->cholSetParam = gk_set_param;
    L->cholSymbolic = gk_symbolic;
    L->cholNumeric = gk_numeric;
    L->cholPsdCheck = gk_psd_check;
    L->cholFSolve = gk_noop_solve;
    L->cholBSolve = gk_noop_solve;
    L->cholSolve = gk_solve;
    L->cholGetDiag = gk_get_diag;
    L->cholInvert = gk_invert;
    L->cholDestroy = gk_destroy;
    L->cholSetParam(L->chol, &p);
    printf("[glue] Schur solver: GPU kind (relTol %g, absTol %g, maxIter %d, restart %d)\n", p.relTol, p.absTol,
           p.maxIter, p.nRestartFreq);
    return retcode;
}

__attribute__((destructor)) static void glue_report( void ) {
    if ( gShadow ) {
        printf("SHADOW {");
        for ( int k = 0; k < SH_N; ++k )
            printf("\"%s\": [%.6e, %d, %ld], ", gShName[k], gShMax[k], gShAt[k], gShCount[k]);
        printf("\"interior_flips\": %ld}\n", gShFlips);
    }
    printf("GLUE {\"updates\": %ld, \"factorizations\": %ld, \"ratio_tests\": %ld, \"schur_builds\": %ld, "
           "\"kkt_factorizations\": %ld, \"kkt_solves\": %ld, \"primal_recoveries\": %ld, "
           "\"diagonal_updates\": %ld}\n",
           gCalls[0], gCalls[1], gCalls[2], gCalls[3], gCalls[4], gCalls[5], gCalls[6], gCalls[7]);
    fflush(stdout);
}
