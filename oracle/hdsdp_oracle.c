/*
 * hdsdp_oracle.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the
 * reference's numerical hot path behind the hdsdp_dev_* ABI
 * (include/hdsdp_gpu.h).  Linked with the host sources into a test library
 * (oracle/Makefile) to check the host logic and the device kernels; the
 * product library never links this file.
 *
 * Every routine restates one reference routine in plain C loops (no BLAS):
 *   buffer assembly      interface/hdsdp_conic_sdp.c:343-402, linalg/hdsdp_sdpdata.c:589-683
 *   Cholesky / PSD check linalg/hdsdp_linsolver.c:493-560 (dpotrf 'L')
 *   S^-1                 linalg/hdsdp_linsolver.c:647-664 (dpotri + symmetrize)
 *   Schur M_ij           interface/hdsdp_conic_sdp.c:683-1052, linalg/hdsdp_sdpdata.c:985-2106
 *   CG on M              linalg/hdsdp_linsolver.c:698-1069
 *   Lanczos step length  linalg/hdsdp_lanczos.c:33-292, interface/hdsdp_conic_sdp.c:462-505, 1638-1682
 *   primal recovery      interface/hdsdp_conic_sdp.c:2328-2379
 *   self-dual terms      interface/hdsdp_conic_sdp.c:737-760, 983-1029
 *   LDL^T fallback       linalg/hdsdp_linsolver.c:1115-1189, 1236-1266, 1367-1425
 *
 * Pinning: see tests/test_oracle.py (closed forms, brute-force traces on tiny
 * inputs, numpy/scipy factorizations) and the reference golden traces.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hdsdp_gpu.h"
#include "eig_small.h"

#define OMAX(a, b) ((a) > (b) ? (a) : (b))
#define OMIN(a, b) ((a) < (b) ? (a) : (b))
#define FE(A, n, i, j) (A)[(size_t) (j) * (n) + (i)]
#define PK(n, i, j) ((size_t) ((2 * (int64_t) (n) - (j) - 1) * (j) / 2 + (i)))

typedef struct {
    int type, nnz;
    int *row, *col;
    double *val;
    double sign;
} ocoeff;

typedef struct {
    int n, maxDim, nComputed;
    double *v, *w, *z1, *z2, *V, *H, *Y, *U, *d, *warm;
} olanczos;

struct hdsdp_dev {
    int m, n, precision;
    ocoeff *A;          /* m + 1: C first */
    double *S[3];       /* buffers: DUALVAR, DUALCHECK, DUALSTEP (n x n, lower valid) */
    double *L[2];       /* Cholesky factors of DUALVAR / DUALCHECK */
    double *W, *W2, *X, *tmp, *tmp2, *vecn;
    /* Schur system */
    double *M, *Mchol, *jacobi;
    double *r, *rnew, *p, *z, *Mp, *x, *rhsbuf;
    int useJacobi, paramUseJacobi, maxIter, nRestartFreq;
    double absTol, relTol;
    /* the indefinite solver the Schur system switches to for good once its Cholesky
     * factorisation or its Cholesky-preconditioned CG fails */
    int useLDL;
    double *Mldl;
    int *ldlPiv;
    olanczos lz;
    /* primal method (interface/hdsdp_psdp.c): the iterate is X; Xs its starting value,
     * dXp the direction, XSX the buffer X S X (+ Xs dS Xs), Sold the dual matrix of the
     * iteration, LX the Cholesky factor of X; lzP the primal method's own Lanczos solver */
    double *Xs, *dXp, *XSX, *Sold, *LX;
    olanczos lzP;
    int psdpReady;
    int wFromDual;  /* W holds S^-1 of the last Schur build (not X) */
    double stats[6];
};

/* ---------------------------------------------------------------- helpers */

static double odot( size_t n, const double *x, const double *y ) {
    double s = 0.0;
    for ( size_t i = 0; i < n; ++i ) s += x[i] * y[i];
    return s;
}

static double onrm2( size_t n, const double *x ) { return sqrt(odot(n, x, x)); }

/* y = alpha * A x with A symmetric, lower triangle valid (dsymv 'L'). */
static void osymv( int n, double alpha, const double *A, const double *x, double *y ) {
    for ( int i = 0; i < n; ++i ) y[i] = 0.0;
    for ( int j = 0; j < n; ++j ) {
        double xj = x[j], acc = 0.0;
        const double *col = A + (size_t) j * n;
        y[j] += col[j] * xj;
        for ( int i = j + 1; i < n; ++i ) {
            y[i] += col[i] * xj;
            acc += col[i] * x[i];
        }
        y[j] += acc;
    }
    for ( int i = 0; i < n; ++i ) y[i] *= alpha;
}

/* In-place lower Cholesky, column-major; returns 0 or the failing column + 1. */
static int ochol( int n, double *A ) {
    for ( int j = 0; j < n; ++j ) {
        double *cj = A + (size_t) j * n;
        for ( int k = 0; k < j; ++k ) {
            const double *ck = A + (size_t) k * n;
            double ljk = ck[j];
            if ( ljk == 0.0 ) continue;
            for ( int i = j; i < n; ++i ) cj[i] -= ljk * ck[i];
        }
        double d = cj[j];
        if ( !(d > 0.0) ) return j + 1;
        d = sqrt(d);
        cj[j] = d;
        double inv = 1.0 / d;
        for ( int i = j + 1; i < n; ++i ) cj[i] *= inv;
    }
    return 0;
}

/* Solve L x = b (fwd) or L^T x = b (bwd) in place, one right-hand side. */
static void otrsv( int n, const double *L, double *x, int trans ) {
    if ( !trans ) {
        for ( int j = 0; j < n; ++j ) {
            x[j] /= FE(L, n, j, j);
            double xj = x[j];
            for ( int i = j + 1; i < n; ++i ) x[i] -= FE(L, n, i, j) * xj;
        }
    } else {
        for ( int j = n - 1; j >= 0; --j ) {
            double s = x[j];
            for ( int i = j + 1; i < n; ++i ) s -= FE(L, n, i, j) * x[i];
            x[j] = s / FE(L, n, j, j);
        }
    }
}

static void osymmetrize_lower( int n, double *A ) {
    for ( int j = 0; j < n; ++j )
        for ( int i = j + 1; i < n; ++i ) FE(A, n, j, i) = FE(A, n, i, j);
}

static void otranspose( int n, double *A ) {
    for ( int j = 0; j < n; ++j )
        for ( int i = j + 1; i < n; ++i ) {
            double t = FE(A, n, i, j);
            FE(A, n, i, j) = FE(A, n, j, i);
            FE(A, n, j, i) = t;
        }
}

/* buffer += a * A (lower triangle), linalg/hdsdp_sdpdata.c:589-683 */
static void oadd_coeff( const ocoeff *A, double a, int n, double *buf ) {
    switch ( A->type ) {
        case HDSDP_COEFF_SPARSE:
            for ( int k = 0; k < A->nnz; ++k ) FE(buf, n, A->row[k], A->col[k]) += a * A->val[k];
            break;
        case HDSDP_COEFF_DENSE: {
            const double *p = A->val;
            for ( int j = 0; j < n; ++j ) {
                double *c = buf + (size_t) j * n + j;
                for ( int i = 0; i < n - j; ++i ) c[i] += a * p[i];
                p += n - j;
            }
            break;
        }
        case HDSDP_COEFF_SPR1: {
            double as = a * A->sign;
            for ( int p = 0; p < A->nnz; ++p )
                for ( int q = 0; q <= p; ++q )
                    FE(buf, n, A->row[p], A->row[q]) += as * A->val[p] * A->val[q];
            break;
        }
        case HDSDP_COEFF_DSR1: {
            double as = a * A->sign;
            for ( int j = 0; j < n; ++j ) {
                double fj = as * A->val[j];
                if ( fj == 0.0 ) continue;
                for ( int i = j; i < n; ++i ) FE(buf, n, i, j) += A->val[i] * fj;
            }
            break;
        }
        default:
            break;
    }
}

/* <A, B> for symmetric full B (kkt3AdotB, linalg/hdsdp_sdpdata.c:1275-1359). */
static double otrace_AB( const ocoeff *A, int n, const double *B, double *aux ) {
    double s = 0.0;
    switch ( A->type ) {
        case HDSDP_COEFF_SPARSE:
            for ( int k = 0; k < A->nnz; ++k ) {
                double t = A->val[k] * FE(B, n, A->row[k], A->col[k]);
                s += ( A->row[k] == A->col[k] ) ? 0.5 * t : t;
            }
            return 2.0 * s;
        case HDSDP_COEFF_DENSE: {
            const double *p = A->val;
            for ( int j = 0; j < n; ++j ) {
                const double *bc = B + (size_t) j * n + j;
                s += 0.5 * p[0] * bc[0];
                for ( int i = 1; i < n - j; ++i ) s += p[i] * bc[i];
                p += n - j;
            }
            return 2.0 * s;
        }
        case HDSDP_COEFF_SPR1:
            for ( int c = 0; c < A->nnz; ++c ) {
                s += 0.5 * A->val[c] * A->val[c] * FE(B, n, A->row[c], A->row[c]);
                for ( int r = c + 1; r < A->nnz; ++r ) s += A->val[r] * A->val[c] * FE(B, n, A->row[r], A->row[c]);
            }
            return 2.0 * A->sign * s;
        case HDSDP_COEFF_DSR1:
            osymv(n, 1.0, B, A->val, aux);
            return A->sign * odot(n, A->val, aux);
        default:
            return 0.0;
    }
}

/* trace(A W B W) for A, B in {SPARSE, SPR1} (KKT5 pairs, linalg/hdsdp_sdpdata.c:1711-1927). */
static double opair_sparse( const ocoeff *A, const ocoeff *B, int n, const double *W ) {
    if ( A->type == HDSDP_COEFF_SPR1 && B->type == HDSDP_COEFF_SPR1 ) {
        double q = 0.0;
        for ( int p = 0; p < A->nnz; ++p )
            for ( int r = 0; r < B->nnz; ++r ) q += A->val[p] * B->val[r] * FE(W, n, A->row[p], B->row[r]);
        return q * q * A->sign * B->sign;
    }
    if ( A->type == HDSDP_COEFF_SPR1 ) { const ocoeff *t = A; A = B; B = t; }
    /* A sparse triplets */
    double res = 0.0;
    for ( int a = 0; a < A->nnz; ++a ) {
        int ar = A->row[a], ac = A->col[a];
        double tb = 0.0;
        if ( B->type == HDSDP_COEFF_SPARSE ) {
            for ( int b = 0; b < B->nnz; ++b ) {
                int br = B->row[b], bc = B->col[b];
                tb += B->val[b] * FE(W, n, ar, br) * FE(W, n, ac, bc);
                if ( br != bc ) tb += B->val[b] * FE(W, n, ar, bc) * FE(W, n, ac, br);
            }
        } else {  /* SPR1 */
            for ( int p = 0; p < B->nnz; ++p ) {
                for ( int q = 0; q < p; ++q ) {
                    double bv = B->val[p] * B->val[q];
                    tb += bv * FE(W, n, ar, B->row[p]) * FE(W, n, ac, B->row[q]);
                    tb += bv * FE(W, n, ar, B->row[q]) * FE(W, n, ac, B->row[p]);
                }
                tb += B->val[p] * B->val[p] * FE(W, n, ar, B->row[p]) * FE(W, n, ac, B->row[p]);
            }
            tb *= B->sign;
        }
        res += ( ar == ac ) ? 0.5 * A->val[a] * tb : A->val[a] * tb;
    }
    return 2.0 * res;
}

/* C = A * B for n x n column-major. */
static void ogemm( int n, const double *A, const double *B, double *C ) {
    memset(C, 0, sizeof(double) * (size_t) n * n);
    for ( int j = 0; j < n; ++j )
        for ( int k = 0; k < n; ++k ) {
            double b = FE(B, n, k, j);
            if ( b == 0.0 ) continue;
            const double *ak = A + (size_t) k * n;
            double *cj = C + (size_t) j * n;
            for ( int i = 0; i < n; ++i ) cj[i] += ak[i] * b;
        }
}

/* B = W A W for one coefficient (n x n full). */
static void osandwich( const ocoeff *A, int n, const double *W, double *tmpA, double *tmpB, double *B ) {
    memset(tmpA, 0, sizeof(double) * (size_t) n * n);
    oadd_coeff(A, 1.0, n, tmpA);
    osymmetrize_lower(n, tmpA);
    ogemm(n, W, tmpA, tmpB);
    ogemm(n, tmpB, W, B);
}

/* ------------------------------------------------------------- lifecycle */

static void lz_alloc( olanczos *lz, int n ) {
    lz->n = n;
    lz->maxDim = 30;
    lz->nComputed = 0;
    lz->v = (double *) calloc(n, sizeof(double));
    lz->w = (double *) calloc(n, sizeof(double));
    lz->z1 = (double *) calloc(n, sizeof(double));
    lz->z2 = (double *) calloc(n, sizeof(double));
    lz->warm = (double *) calloc(n, sizeof(double));
    lz->V = (double *) calloc((size_t) n * (lz->maxDim + 1), sizeof(double));
    lz->H = (double *) calloc((size_t) (lz->maxDim + 1) * (lz->maxDim + 1), sizeof(double));
    lz->Y = (double *) calloc((size_t) lz->maxDim * 2, sizeof(double));
    lz->U = (double *) calloc((size_t) lz->maxDim * lz->maxDim, sizeof(double));
    lz->d = (double *) calloc(lz->maxDim, sizeof(double));
}

static void lz_free( olanczos *lz ) {
    free(lz->v); free(lz->w); free(lz->z1); free(lz->z2); free(lz->warm);
    free(lz->V); free(lz->H); free(lz->Y); free(lz->U); free(lz->d);
}

hdsdp_retcode hdsdp_dev_create( hdsdp_dev **pdev, int nRow, int nCol, int precision ) {
    hdsdp_dev *d = (hdsdp_dev *) calloc(1, sizeof(hdsdp_dev));
    if ( !d ) return HDSDP_RETCODE_MEMORY;
    d->m = nRow;
    d->n = nCol;
    d->precision = precision;
    size_t nn = (size_t) nCol * nCol, mm = (size_t) nRow * nRow;
    for ( int k = 0; k < 3; ++k ) d->S[k] = (double *) calloc(nn, sizeof(double));
    for ( int k = 0; k < 2; ++k ) d->L[k] = (double *) calloc(nn, sizeof(double));
    d->W = (double *) calloc(nn, sizeof(double));
    d->W2 = (double *) calloc(nn, sizeof(double));
    d->X = (double *) calloc(nn, sizeof(double));
    d->tmp = (double *) calloc(nn, sizeof(double));
    d->tmp2 = (double *) calloc(nn, sizeof(double));
    d->vecn = (double *) calloc((size_t) nCol * 2, sizeof(double));
    d->M = (double *) calloc(mm, sizeof(double));
    d->Mchol = (double *) calloc(mm, sizeof(double));
    d->jacobi = (double *) calloc(nRow, sizeof(double));
    d->r = (double *) calloc(nRow, sizeof(double));
    d->rnew = (double *) calloc(nRow, sizeof(double));
    d->p = (double *) calloc(nRow, sizeof(double));
    d->z = (double *) calloc(nRow, sizeof(double));
    d->Mp = (double *) calloc(nRow, sizeof(double));
    d->x = (double *) calloc(nRow, sizeof(double));
    d->rhsbuf = (double *) calloc(nRow, sizeof(double));
    /* CG defaults (linalg/hdsdp_linsolver.c:718-728) */
    d->absTol = 1e-06;
    d->relTol = 1e-06;
    d->paramUseJacobi = 1;
    d->useJacobi = 1;
    d->maxIter = OMAX(50, nRow / 20);
    d->nRestartFreq = 20;
    olanczos *lz = &d->lz;
    lz_alloc(lz, nCol);
    if ( !d->M || !d->Mchol || !d->W || !lz->V ) return HDSDP_RETCODE_MEMORY;
    *pdev = d;
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_load( hdsdp_dev *d, const hdsdp_coeff_view *cv ) {
    int n = d->n;
    d->A = (ocoeff *) calloc((size_t) d->m + 1, sizeof(ocoeff));
    for ( int k = 0; k <= d->m; ++k ) {
        ocoeff *A = &d->A[k];
        A->type = cv[k].type;
        A->nnz = cv[k].nnz;
        A->sign = cv[k].sign;
        size_t nv = 0;
        if ( A->type == HDSDP_COEFF_SPARSE || A->type == HDSDP_COEFF_SPR1 ) nv = A->nnz;
        else if ( A->type == HDSDP_COEFF_DENSE ) nv = (size_t) n * (n + 1) / 2;
        else if ( A->type == HDSDP_COEFF_DSR1 ) nv = n;
        if ( nv ) {
            A->val = (double *) malloc(nv * sizeof(double));
            memcpy(A->val, cv[k].val, nv * sizeof(double));
        }
        if ( A->type == HDSDP_COEFF_SPARSE || A->type == HDSDP_COEFF_SPR1 ) {
            A->row = (int *) malloc(sizeof(int) * A->nnz);
            memcpy(A->row, cv[k].row, sizeof(int) * A->nnz);
        }
        if ( A->type == HDSDP_COEFF_SPARSE ) {
            A->col = (int *) malloc(sizeof(int) * A->nnz);
            memcpy(A->col, cv[k].col, sizeof(int) * A->nnz);
        }
    }
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_scale_obj( hdsdp_dev *d, double s ) {
    ocoeff *C = &d->A[0];
    size_t nv = 0;
    if ( C->type == HDSDP_COEFF_SPR1 || C->type == HDSDP_COEFF_DSR1 ) { C->sign *= s; return HDSDP_RETCODE_OK; }
    if ( C->type == HDSDP_COEFF_SPARSE ) nv = C->nnz;
    if ( C->type == HDSDP_COEFF_DENSE ) nv = (size_t) d->n * (d->n + 1) / 2;
    for ( size_t k = 0; k < nv; ++k ) C->val[k] *= s;
    return HDSDP_RETCODE_OK;
}

void hdsdp_dev_destroy( hdsdp_dev **pdev ) {
    if ( !pdev || !*pdev ) return;
    hdsdp_dev *d = *pdev;
    if ( d->A ) {
        for ( int k = 0; k <= d->m; ++k ) { free(d->A[k].val); free(d->A[k].row); free(d->A[k].col); }
        free(d->A);
    }
    free(d->Mldl);
    free(d->ldlPiv);
    for ( int k = 0; k < 3; ++k ) free(d->S[k]);
    for ( int k = 0; k < 2; ++k ) free(d->L[k]);
    free(d->W); free(d->W2); free(d->X); free(d->tmp); free(d->tmp2); free(d->vecn);
    free(d->M); free(d->Mchol); free(d->jacobi);
    free(d->r); free(d->rnew); free(d->p); free(d->z); free(d->Mp); free(d->x); free(d->rhsbuf);
    lz_free(&d->lz);
    lz_free(&d->lzP);
    free(d->Xs); free(d->dXp); free(d->XSX); free(d->Sold); free(d->LX);
    free(d);
    *pdev = NULL;
}

const char *hdsdp_dev_backend_name( void ) { return "cpu-oracle"; }

hdsdp_retcode hdsdp_dev_stats( hdsdp_dev *d, double *stats, int nStats ) {
    for ( int k = 0; k < nStats && k < 6; ++k ) stats[k] = d->stats[k];
    return HDSDP_RETCODE_OK;
}

/* ------------------------------------------------------------ dual buffers */

hdsdp_retcode hdsdp_dev_update_buffer( hdsdp_dev *d, double cCoef, double aScal, const double *aCoef,
                                       double eye, int which ) {
    int n = d->n;
    double *t = d->S[which];
    memset(t, 0, sizeof(double) * (size_t) n * n);
    for ( int i = 0; i < d->m; ++i ) {
        double a = aScal * aCoef[i];
        if ( a != 0.0 ) oadd_coeff(&d->A[i + 1], a, n, t);
    }
    if ( cCoef != 0.0 ) oadd_coeff(&d->A[0], cCoef, n, t);
    if ( eye != 0.0 ) for ( int j = 0; j < n; ++j ) FE(t, n, j, j) += eye;
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_factorize( hdsdp_dev *d, int which, int *isPsd ) {
    int n = d->n;
    int f = ( which == 0 ) ? 0 : 1;
    memcpy(d->L[f], d->S[f], sizeof(double) * (size_t) n * n);
    d->stats[1] += 1;
    *isPsd = ( ochol(n, d->L[f]) == 0 );
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_axpy_check( hdsdp_dev *d, double step, int which, int *isPsd ) {
    size_t nn = (size_t) d->n * d->n;
    int f = ( which == 0 ) ? 0 : 1;
    if ( f == 1 ) memcpy(d->S[1], d->S[0], sizeof(double) * nn);
    for ( size_t k = 0; k < nn; ++k ) d->S[f][k] += step * d->S[2][k];
    return hdsdp_dev_factorize(d, f, isPsd);
}

hdsdp_retcode hdsdp_dev_logdet( hdsdp_dev *d, int which, double *logdet ) {
    int n = d->n, f = ( which == 0 ) ? 0 : 1;
    double s = 0.0;
    for ( int j = 0; j < n; ++j ) s += log(FE(d->L[f], n, j, j));
    *logdet = 2.0 * s;
    return HDSDP_RETCODE_OK;
}

/* ---------------------------------------------------------------- Lanczos */

static void lz_prepare( int n, double *vVec ) {
    srand(n);
    for ( int i = 0; i < n; ++i ) {
        srand(rand());
        vVec[i] = sqrt(sqrt((rand() % 1627))) * (rand() % 2 - 0.5);
    }
}

static void lz_perturb( int n, double *vVec ) {
    srand(n);
    for ( int i = 0; i < n; ++i ) {
        srand(rand());
        vVec[i] += 1e-03 * sqrt(sqrt((rand() % 1627))) * (rand() % 2 - 0.5);
    }
}

static double lz_normalize( int n, double *a ) {
    double nrm = onrm2(n, a);
    if ( nrm > 0.0 ) {
        double inv = 1.0 / nrm;
        for ( int i = 0; i < n; ++i ) a[i] *= inv;
    } else {
        nrm = 0.0;
        memset(a, 0, sizeof(double) * n);
    }
    return nrm;
}

/* x -> L^-1 (-D) L^-T x: the dual cone's operator with the factor of a dual buffer and
 * D = dS (interface/hdsdp_conic_sdp.c:462-505), the primal method's with the factor of X
 * and D = the primal direction (interface/hdsdp_psdp.c:58-66); D's lower triangle. */
static void lz_mvec( hdsdp_dev *d, const double *L, const double *D, const double *in, double *out ) {
    int n = d->n;
    memcpy(out, in, sizeof(double) * n);
    otrsv(n, L, out, 1);
    osymv(n, -1.0, D, out, d->vecn);
    memcpy(out, d->vecn, sizeof(double) * n);
    otrsv(n, L, out, 0);
    d->stats[3] += 1;
}

static hdsdp_retcode lanczos_run( hdsdp_dev *d, olanczos *lz, const double *L, const double *D, double *maxStep );

hdsdp_retcode hdsdp_dev_ratio_test( hdsdp_dev *d, int which, double *maxStep ) {
    int n = d->n, f = ( which == 0 ) ? 0 : 1;
    if ( n == 1 ) {
        double ds = d->S[2][0], s = d->S[f][0];
        *maxStep = ( ds > 0.0 ) ? HDSDP_INFINITY : (-s / ds);
        return HDSDP_RETCODE_OK;
    }
    return lanczos_run(d, &d->lz, d->L[f], d->S[2], maxStep);
}

/* HLanczosSolve (linalg/hdsdp_lanczos.c:161-292) */
static hdsdp_retcode lanczos_run( hdsdp_dev *d, olanczos *lz, const double *L, const double *D, double *maxStep ) {
    int n = d->n;
    int nV = n, nH = lz->maxDim + 1;
#define HH(i, j) lz->H[(size_t) nH * (j) + (i)]
#define VV(i, j) lz->V[(size_t) nV * (j) + (i)]
    if ( lz->nComputed == 0 ) {
        lz_prepare(n, lz->v);
    } else {
        memset(lz->w, 0, sizeof(double) * n);
        memset(lz->z1, 0, sizeof(double) * n);
        memset(lz->z2, 0, sizeof(double) * n);
        memset(lz->V, 0, sizeof(double) * (size_t) n * (lz->maxDim + 1));
        memset(lz->H, 0, sizeof(double) * (size_t) nH * nH);
        memset(lz->Y, 0, sizeof(double) * (size_t) lz->maxDim * 2);
        memset(lz->d, 0, sizeof(double) * lz->maxDim);
        memset(lz->U, 0, sizeof(double) * (size_t) lz->maxDim * lz->maxDim);
        memcpy(lz->v, lz->warm, sizeof(double) * n);
        lz_perturb(n, lz->v);
    }
    lz_normalize(n, lz->v);
    memcpy(lz->V, lz->v, sizeof(double) * n);
    int check = OMIN(lz->maxDim / 5, 3);
    static int trace = -1;
    if ( trace < 0 ) trace = getenv("HDSDP_ORACLE_LZ_TRACE") != NULL;
    if ( trace ) printf("Lanczos call %d\n", lz->nComputed);
    for ( int k = 0; k < lz->maxDim; ++k ) {
        lz_mvec(d, L, D, lz->v, lz->w);
        if ( k > 0 ) {
            double neg = -HH(k, k - 1);
            for ( int i = 0; i < n; ++i ) lz->w[i] += neg * VV(i, k - 1);
        }
        double alp = -odot(n, lz->w, &VV(0, k));
        for ( int i = 0; i < n; ++i ) lz->w[i] += alp * VV(i, k);
        HH(k, k) = -alp;
        if ( trace ) printf("Lanczos Alp value: %f \n", -alp);
        memcpy(lz->v, lz->w, sizeof(double) * n);
        double normPres = lz_normalize(n, lz->v);
        if ( normPres > 0.0 ) {
            memcpy(&VV(0, k + 1), lz->v, sizeof(double) * n);
            HH(k + 1, k) = HH(k, k + 1) = normPres;
        }
        if ( (k + 1) % check == 0 || k > lz->maxDim - 1 || normPres == 0.0 ) {
            int kp1 = k + 1;
            /* projected matrix: the leading (k+1) x (k+1) Lanczos tridiagonal */
            for ( int i = 0; i < kp1; ++i ) {
                lz->U[i] = HH(i, i);
                if ( i + 1 < kp1 ) lz->U[kp1 + i] = HH(i + 1, i);
            }
            if ( kp1 < 2 || hdsdp_eig_tridiag_top(kp1, lz->U, lz->U + kp1, 2, lz->d, lz->Y) != 0 )
                return HDSDP_RETCODE_FAILED;
            double resi = fabs(HH(kp1, k) * lz->Y[kp1 + k]);
            if ( trace ) printf("Lanczos outer resi value: %f \n", resi);
            if ( resi < 1e-04 || k >= lz->maxDim - 1 ) {
                double e1 = lz->d[1], e2 = lz->d[0];
                for ( int i = 0; i < n; ++i ) {
                    double s = 0.0;
                    for ( int c = 0; c < kp1; ++c ) s += VV(i, c) * lz->Y[kp1 + c];
                    lz->z1[i] = s;
                }
                lz_mvec(d, L, D, lz->z1, lz->z2);
                memcpy(lz->warm, lz->z2, sizeof(double) * n);
                for ( int i = 0; i < n; ++i ) lz->z2[i] += -e1 * lz->z1[i];
                double r1 = onrm2(n, lz->z2);
                for ( int i = 0; i < n; ++i ) {
                    double s = 0.0;
                    for ( int c = 0; c < kp1; ++c ) s += VV(i, c) * lz->Y[c];
                    lz->z2[i] = s;
                }
                lz_mvec(d, L, D, lz->z2, lz->z1);
                for ( int i = 0; i < n; ++i ) lz->z1[i] += -e1 * lz->z2[i];
                double r2 = onrm2(n, lz->z1);
                double diff = e1 - e2 - r2;
                double gamma = ( diff > 0 ) ? diff : 1e-16;
                double r1sq = r1 * r1 / gamma;
                gamma = OMIN(r1, r1sq);
                if ( trace ) printf("Lanczos bound k=%d e1=%.12e e2=%.12e r1=%.6e r2=%.6e gamma=%.6e\n", k, e1, e2, r1, r2, gamma);
                if ( gamma < 1e-03 || gamma + e1 <= 0.5 ) {
                    *maxStep = ( gamma + e1 <= 0.0 ) ? HDSDP_INFINITY : 1.0 / (gamma + e1);
                    break;
                } else {
                    if ( normPres == 0.0 ) return HDSDP_RETCODE_FAILED;
                    *maxStep = 1.0 / (gamma + e1);
                }
            }
        }
    }
#undef HH
#undef VV
    lz->nComputed += 1;
    return HDSDP_RETCODE_OK;
}

/* ------------------------------------------------------------------ Schur */

/* W = S^-1 from the dual factor (dpotri + symmetrize). */
static void oinverse( hdsdp_dev *d, const double *L, double *W ) {
    int n = d->n;
    memset(W, 0, sizeof(double) * (size_t) n * n);
    for ( int j = 0; j < n; ++j ) {
        double *e = d->tmp;
        memset(e, 0, sizeof(double) * n);
        e[j] = 1.0;
        otrsv(n, L, e, 0);
        otrsv(n, L, e, 1);
        memcpy(W + (size_t) j * n, e, sizeof(double) * n);
    }
    for ( int j = 0; j < n; ++j )
        for ( int i = j + 1; i < n; ++i ) {
            double v = 0.5 * (FE(W, n, i, j) + FE(W, n, j, i));
            FE(W, n, i, j) = FE(W, n, j, i) = v;
        }
}

static int is_sparse_kind( int t ) { return t == HDSDP_COEFF_SPARSE || t == HDSDP_COEFF_SPR1; }

hdsdp_retcode hdsdp_dev_schur( hdsdp_dev *d, int typeKKT, double Rd, double *asinv, double *asinvrd,
                               double *traceSinv ) {
    int n = d->n, m = d->m;
    size_t nn = (size_t) n * n;
    /* HOMOGENEOUS (2) builds what INFEASIBLE builds; its extra terms: hdsdp_dev_hsd */
    d->wFromDual = ( typeKKT != 3 );
    if ( typeKKT == 3 ) {
        /* PRIMAL: strategy M5 with W := X (interface/hdsdp_conic_sdp.c:1741-1745, 1813-1820);
         * no DENSE or DSR1 constraint (linalg/hdsdp_sdpdata.c:1961-1965, 1995-1999) */
        if ( !d->psdpReady ) return HDSDP_RETCODE_FAILED;
        for ( int i = 1; i <= m; ++i )
            if ( !is_sparse_kind(d->A[i].type) && d->A[i].type != HDSDP_COEFF_ZERO ) return HDSDP_RETCODE_FAILED;
        memcpy(d->W, d->X, sizeof(double) * nn);
    } else {
        oinverse(d, d->L[0], d->W);
    }
    ogemm(n, d->W, d->W, d->W2);
    *traceSinv = 0.0;
    for ( int i = 0; i < m; ++i ) {
        asinv[i] = otrace_AB(&d->A[i + 1], n, d->W, d->vecn);
        asinvrd[i] = ( Rd != 0.0 ) ? Rd * otrace_AB(&d->A[i + 1], n, d->W2, d->vecn) : 0.0;
    }
    d->stats[0] += 1;
    if ( typeKKT == 1 ) return HDSDP_RETCODE_OK;
    if ( Rd != 0.0 ) for ( int j = 0; j < n; ++j ) *traceSinv += FE(d->W, n, j, j);

    /* M_ij = <A_i, W A_j W>: pair formulas between sparse kinds, sandwich B = W A W otherwise */
    double *B = (double *) malloc(sizeof(double) * nn);
    for ( int j = 0; j < m; ++j ) {
        const ocoeff *Aj = &d->A[j + 1];
        if ( Aj->type == HDSDP_COEFF_ZERO ) {
            for ( int i = j; i < m; ++i ) FE(d->M, m, i, j) = 0.0;
            continue;
        }
        int haveB = !is_sparse_kind(Aj->type);
        if ( haveB ) osandwich(Aj, n, d->W, d->tmp, d->tmp2, B);
        for ( int i = j; i < m; ++i ) {
            const ocoeff *Ai = &d->A[i + 1];
            double v;
            if ( Ai->type == HDSDP_COEFF_ZERO ) v = 0.0;
            else if ( haveB ) v = otrace_AB(Ai, n, B, d->vecn);
            else if ( is_sparse_kind(Ai->type) ) v = opair_sparse(Ai, Aj, n, d->W);
            else {
                osandwich(Ai, n, d->W, d->tmp, d->tmp2, d->X);
                v = otrace_AB(Aj, n, d->X, d->vecn);
            }
            FE(d->M, m, i, j) = v;
        }
    }
    free(B);
    return HDSDP_RETCODE_OK;
}

/* The self-dual embedding's terms (interface/hdsdp_conic_sdp.c:983-1029,
 * sdpDenseConeIGetHSDComponents, and the dASinvCSinvVec entries strategies M1-M5 add
 * for KKT_TYPE_HOMOGENEOUS, 737-760 and the alike blocks): with B = W C W and W = S^-1
 * of the last build, dASinvCSinv[i] = <A_i, B>, dCSinv = <C, W>, dCSinvCSinv = <C, B>,
 * dCSinvRdSinv = Rd tr(B). */
hdsdp_retcode hdsdp_dev_hsd( hdsdp_dev *d, double Rd, double *aSinvCSinv, double *cSinv, double *cSinvCSinv,
                             double *cSinvRdSinv ) {
    int n = d->n, m = d->m;
    if ( !d->wFromDual ) return HDSDP_RETCODE_FAILED;
    double *B = (double *) malloc(sizeof(double) * (size_t) n * n);
    if ( !B ) return HDSDP_RETCODE_MEMORY;
    osandwich(&d->A[0], n, d->W, d->tmp, d->tmp2, B);
    for ( int i = 0; i < m; ++i ) aSinvCSinv[i] = otrace_AB(&d->A[i + 1], n, B, d->vecn);
    *cSinv = otrace_AB(&d->A[0], n, d->W, d->vecn);
    *cSinvCSinv = otrace_AB(&d->A[0], n, B, d->vecn);
    double tr = 0.0;
    for ( int j = 0; j < n; ++j ) tr += FE(B, n, j, j);
    *cSinvRdSinv = Rd * tr;
    free(B);
    return HDSDP_RETCODE_OK;
}

/* -------------------------------------------------------------- Schur solve */

hdsdp_retcode hdsdp_dev_kkt_params( hdsdp_dev *d, double relTol, double absTol, int maxIter, int nRestart ) {
    if ( absTol > 0 ) d->absTol = absTol;
    if ( relTol > 0 ) d->relTol = relTol;
    if ( maxIter > 0 ) d->maxIter = maxIter;
    if ( nRestart > 0 ) d->nRestartFreq = nRestart;
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_kkt_add_diag( hdsdp_dev *d, const double *diag ) {
    for ( int i = 0; i < d->m; ++i ) FE(d->M, d->m, i, i) += diag[i];
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_kkt_regularize( hdsdp_dev *d, double reg ) {
    double mn = HDSDP_INFINITY;
    for ( int i = 0; i < d->m; ++i ) mn = OMIN(FE(d->M, d->m, i, i), mn);
    reg = reg * mn;
    reg = OMIN(reg, 1e-05);
    if ( reg < 1e-14 ) reg = 0.0;
    for ( int i = 0; i < d->m; ++i ) FE(d->M, d->m, i, i) += reg;
    return HDSDP_RETCODE_OK;
}

static hdsdp_retcode build_precond( hdsdp_dev *d ) {
    int m = d->m;
    if ( d->useJacobi ) {
        for ( int i = 0; i < m; ++i ) d->jacobi[i] = FE(d->M, m, i, i);
        return HDSDP_RETCODE_OK;
    }
    memcpy(d->Mchol, d->M, sizeof(double) * (size_t) m * m);
    d->stats[4] += 1;
    return ( ochol(m, d->Mchol) == 0 ) ? HDSDP_RETCODE_OK : HDSDP_RETCODE_FAILED;
}

/* Symmetric indefinite LDL^T with Bunch-Kaufman diagonal pivoting on the lower triangle
 * (Bunch & Kaufman 1977, as LAPACK's dsytf2 'L' states it), the factorisation of the
 * reference's indefinite Schur solver (lapackIndefiniteLinSolverNumeric,
 * linalg/hdsdp_linsolver.c:1115-1136, dsytrf).  ipiv as LAPACK's, 1-based: k + 1 for a 1 x 1
 * pivot with row k interchanged, -(k + 1) on both rows of a 2 x 2 pivot.  Returns 0, or
 * the column + 1 of a zero pivot. */
static int oldl_factor( int n, double *A, int *ipiv ) {
    const double alpha = (1.0 + sqrt(17.0)) / 8.0;
    int info = 0, k = 0;
    while ( k < n ) {
        int kstep = 1, kp = k;
        double absakk = fabs(FE(A, n, k, k)), colmax = 0.0;
        int imax = k;
        for ( int i = k + 1; i < n; ++i ) {
            if ( fabs(FE(A, n, i, k)) > colmax ) { colmax = fabs(FE(A, n, i, k)); imax = i; }
        }
        if ( OMAX(absakk, colmax) == 0.0 ) {
            if ( !info ) info = k + 1;
        } else {
            if ( absakk < alpha * colmax ) {
                double rowmax = 0.0;
                for ( int j = k; j < imax; ++j ) rowmax = OMAX(rowmax, fabs(FE(A, n, imax, j)));
                for ( int i = imax + 1; i < n; ++i ) rowmax = OMAX(rowmax, fabs(FE(A, n, i, imax)));
                if ( absakk >= alpha * colmax * (colmax / rowmax) ) {
                    kp = k;
                } else if ( fabs(FE(A, n, imax, imax)) >= alpha * rowmax ) {
                    kp = imax;
                } else {
                    kp = imax;
                    kstep = 2;
                }
            }
            int kk = k + kstep - 1;
            if ( kp != kk ) {
                double t;
                for ( int i = kp + 1; i < n; ++i ) { t = FE(A, n, i, kk); FE(A, n, i, kk) = FE(A, n, i, kp); FE(A, n, i, kp) = t; }
                for ( int j = kk + 1; j < kp; ++j ) { t = FE(A, n, j, kk); FE(A, n, j, kk) = FE(A, n, kp, j); FE(A, n, kp, j) = t; }
                t = FE(A, n, kk, kk); FE(A, n, kk, kk) = FE(A, n, kp, kp); FE(A, n, kp, kp) = t;
                if ( kstep == 2 ) { t = FE(A, n, k + 1, k); FE(A, n, k + 1, k) = FE(A, n, kp, k); FE(A, n, kp, k) = t; }
            }
            if ( kstep == 1 ) {
                double r1 = 1.0 / FE(A, n, k, k);
                for ( int j = k + 1; j < n; ++j ) {
                    double ajk = FE(A, n, j, k);
                    for ( int i = j; i < n; ++i ) FE(A, n, i, j) -= r1 * FE(A, n, i, k) * ajk;
                }
                for ( int i = k + 1; i < n; ++i ) FE(A, n, i, k) *= r1;
            } else if ( k < n - 2 ) {
                double d21 = FE(A, n, k + 1, k);
                double d11 = FE(A, n, k + 1, k + 1) / d21, d22 = FE(A, n, k, k) / d21;
                double t = 1.0 / (d11 * d22 - 1.0);
                d21 = t / d21;
                for ( int j = k + 2; j < n; ++j ) {
                    double wk = d21 * (d11 * FE(A, n, j, k) - FE(A, n, j, k + 1));
                    double wkp1 = d21 * (d22 * FE(A, n, j, k + 1) - FE(A, n, j, k));
                    for ( int i = j; i < n; ++i ) FE(A, n, i, j) -= FE(A, n, i, k) * wk + FE(A, n, i, k + 1) * wkp1;
                    FE(A, n, j, k) = wk;
                    FE(A, n, j, k + 1) = wkp1;
                }
            }
        }
        if ( kstep == 1 ) {
            ipiv[k] = kp + 1;
        } else {
            ipiv[k] = ipiv[k + 1] = -(kp + 1);
        }
        k += kstep;
    }
    return info;
}

/* b <- A^-1 b from oldl_factor's L, D and ipiv (LAPACK's dsytrs 'L' steps:
 * lapackIndefiniteLinSolverSolveN, linalg/hdsdp_linsolver.c:1170-1189) */
static void oldl_solve( int n, const double *A, const int *ipiv, double *b ) {
    double t;
    int k = 0;
    while ( k < n ) {
        if ( ipiv[k] > 0 ) {
            int kp = ipiv[k] - 1;
            if ( kp != k ) { t = b[k]; b[k] = b[kp]; b[kp] = t; }
            for ( int i = k + 1; i < n; ++i ) b[i] -= FE(A, n, i, k) * b[k];
            b[k] /= FE(A, n, k, k);
            k += 1;
        } else {
            int kp = -ipiv[k] - 1;
            if ( kp != k + 1 ) { t = b[k + 1]; b[k + 1] = b[kp]; b[kp] = t; }
            for ( int i = k + 2; i < n; ++i ) b[i] -= FE(A, n, i, k) * b[k] + FE(A, n, i, k + 1) * b[k + 1];
            double akm1k = FE(A, n, k + 1, k);
            double akm1 = FE(A, n, k, k) / akm1k, ak = FE(A, n, k + 1, k + 1) / akm1k;
            double denom = akm1 * ak - 1.0;
            double bkm1 = b[k] / akm1k, bk = b[k + 1] / akm1k;
            b[k] = (ak * bkm1 - bk) / denom;
            b[k + 1] = (akm1 * bk - bkm1) / denom;
            k += 2;
        }
    }
    k = n - 1;
    while ( k >= 0 ) {
        if ( ipiv[k] > 0 ) {
            for ( int i = k + 1; i < n; ++i ) b[k] -= FE(A, n, i, k) * b[i];
            int kp = ipiv[k] - 1;
            if ( kp != k ) { t = b[k]; b[k] = b[kp]; b[kp] = t; }
            k -= 1;
        } else {
            for ( int i = k + 1; i < n; ++i ) {
                b[k] -= FE(A, n, i, k) * b[i];
                b[k - 1] -= FE(A, n, i, k - 1) * b[i];
            }
            int kp = -ipiv[k] - 1;
            if ( kp != k ) { t = b[k]; b[k] = b[kp]; b[kp] = t; }
            k -= 2;
        }
    }
}

/* HFpLinsysSwitchToIndefinite (linalg/hdsdp_linsolver.c:1236-1266): from now on the
 * Schur system is factored by LDL^T and solved directly */
static hdsdp_retcode ldl_factor( hdsdp_dev *d ) {
    int m = d->m;
    if ( !d->Mldl ) {
        d->Mldl = (double *) malloc(sizeof(double) * (size_t) m * m);
        d->ldlPiv = (int *) malloc(sizeof(int) * (size_t) m);
        if ( !d->Mldl || !d->ldlPiv ) return HDSDP_RETCODE_MEMORY;
    }
    d->useLDL = 1;
    memcpy(d->Mldl, d->M, sizeof(double) * (size_t) m * m);
    d->stats[4] += 1;
    return ( oldl_factor(m, d->Mldl, d->ldlPiv) == 0 ) ? HDSDP_RETCODE_OK : HDSDP_RETCODE_FAILED;
}

/* HFpLinsysNumeric (linalg/hdsdp_linsolver.c:1367-1381): a Cholesky factorisation that
 * fails switches the system to the indefinite solver */
hdsdp_retcode hdsdp_dev_kkt_factorize( hdsdp_dev *d ) {
    if ( d->useLDL ) return ldl_factor(d);
    hdsdp_retcode rc = build_precond(d);
    if ( rc == HDSDP_RETCODE_FAILED && !d->useJacobi ) rc = ldl_factor(d);
    return rc;
}

int hdsdp_dev_kkt_uses_ldl( hdsdp_dev *d ) {
    return d->useLDL;
}

static void apply_precond( hdsdp_dev *d, double *v ) {
    if ( d->useJacobi ) {
        for ( int i = 0; i < d->m; ++i ) v[i] = v[i] / d->jacobi[i];
    } else {
        otrsv(d->m, d->Mchol, v, 0);
        otrsv(d->m, d->Mchol, v, 1);
    }
}

/* conjGradSolve, linalg/hdsdp_linsolver.c:855-997 */
static hdsdp_retcode cg_solve( hdsdp_dev *d, const double *rhs, double *sol ) {
    int m = d->m, iter = 0, status = 0; /* 0 ok, 2 maxiter */
    d->useJacobi = d->paramUseJacobi;
    int restart = OMAX(d->nRestartFreq, 20);
    double *r = d->r, *rn = d->rnew, *p = d->p, *z = d->z, *Mp = d->Mp, *x = d->x;
    memset(x, 0, sizeof(double) * m);
    memcpy(r, rhs, sizeof(double) * m);
    double rnorm = onrm2(m, r), bnorm = onrm2(m, rhs);
    double tol = OMIN(d->absTol, bnorm * d->relTol);
    tol = OMAX(tol, 0.1 * d->absTol);
    if ( rnorm < tol ) {
        /* the reference returns here without writing the solution vector */
        return HDSDP_RETCODE_OK;
    }
    memcpy(p, r, sizeof(double) * m);
    apply_precond(d, p);
    memcpy(z, p, sizeof(double) * m);
    osymv(m, 1.0, d->M, p, Mp);
    for ( iter = 0; iter < d->maxIter; ++iter ) {
        double rz = odot(m, z, r);
        double pMp = odot(m, p, Mp);
        double alpha = rz / pMp;
        for ( int i = 0; i < m; ++i ) x[i] += alpha * p[i];
        d->stats[2] += 1;
        if ( iter % restart == 5 && d->useJacobi ) {
            memcpy(r, rhs, sizeof(double) * m);
            osymv(m, 1.0, d->M, x, p);
            for ( int i = 0; i < m; ++i ) r[i] -= p[i];
            memcpy(p, r, sizeof(double) * m);
            apply_precond(d, p);
            osymv(m, 1.0, d->M, p, Mp);
            memcpy(z, r, sizeof(double) * m);
            apply_precond(d, z);
            continue;
        }
        memcpy(rn, r, sizeof(double) * m);
        for ( int i = 0; i < m; ++i ) rn[i] += -alpha * Mp[i];
        memcpy(z, rn, sizeof(double) * m);
        apply_precond(d, z);
        double beta = odot(m, rn, z) / rz;
        for ( int i = 0; i < m; ++i ) p[i] = z[i] + beta * p[i];
        osymv(m, 1.0, d->M, p, Mp);
        memcpy(r, rn, sizeof(double) * m);
        rnorm = onrm2(m, r);
        if ( rnorm != rnorm ) return HDSDP_RETCODE_FAILED;
        if ( iter > 20 && rnorm > 0.01 * bnorm ) {
            status = 2;
            break;
        }
        if ( rnorm < tol ) {
            status = 0;
            break;
        }
    }
    d->stats[5] = iter;
    if ( iter >= d->maxIter || status == 2 ) {
        if ( d->useJacobi ) {
            d->useJacobi = 0;
            d->paramUseJacobi = 0;
            if ( build_precond(d) != HDSDP_RETCODE_OK ) return HDSDP_RETCODE_FAILED;
            if ( cg_solve(d, rhs, sol) != HDSDP_RETCODE_OK ) return HDSDP_RETCODE_FAILED;
            return HDSDP_RETCODE_OK;
        }
        return HDSDP_RETCODE_FAILED;
    }
    memcpy(sol, x, sizeof(double) * m);
    return HDSDP_RETCODE_OK;
}

/* HFpLinsysSolve (linalg/hdsdp_linsolver.c:1410-1425): a CG solve that fails switches
 * the system to the indefinite solver, which then solves this right-hand side */
hdsdp_retcode hdsdp_dev_kkt_solve( hdsdp_dev *d, const double *rhs, double *sol ) {
    int m = d->m;
    memcpy(d->rhsbuf, rhs, sizeof(double) * m);
    if ( !d->useLDL ) {
        hdsdp_retcode rc = cg_solve(d, d->rhsbuf, sol);
        if ( rc != HDSDP_RETCODE_FAILED ) return rc;
        memcpy(d->rhsbuf, rhs, sizeof(double) * m);
        if ( ldl_factor(d) != HDSDP_RETCODE_OK ) return HDSDP_RETCODE_FAILED;
    }
    oldl_solve(m, d->Mldl, d->ldlPiv, d->rhsbuf);
    memcpy(sol, d->rhsbuf, sizeof(double) * m);
    return HDSDP_RETCODE_OK;
}

/* hdsdp_dev_kkt_solve_noalloc on host memory: the Jacobi-preconditioned CG of
 * cg_solve above (conjGradSolve, linalg/hdsdp_linsolver.c:855-997) on a caller-held
 * full symmetric M, FAILED where cg_solve would switch to the Cholesky
 * preconditioner.  The pointers are host pointers here. */
hdsdp_retcode hdsdp_dev_kkt_solve_noalloc( const double *M, int m, const double *rhs, double *sol, double relTol,
                                           double absTol, int maxIter, int nRestartFreq, void *stream, int *nIter ) {
    (void) stream;
    if ( nIter ) *nIter = 0;
    if ( m <= 0 || m > 65536 || !M || !rhs || !sol || !(relTol > 0.0) || !(absTol > 0.0) || maxIter <= 0 )
        return HDSDP_RETCODE_FAILED;
    double *buf = (double *) calloc((size_t) 7 * m, sizeof(double));
    if ( !buf ) return HDSDP_RETCODE_MEMORY;
    double *jac = buf, *r = buf + m, *rn = buf + 2 * m, *p = buf + 3 * m, *z = buf + 4 * m, *Mp = buf + 5 * m;
    double *x = buf + 6 * m;
    hdsdp_retcode rc = HDSDP_RETCODE_FAILED;
    int restart = OMAX(nRestartFreq, 20), iter = 0, converged = 0;
    for ( int i = 0; i < m; ++i ) jac[i] = FE(M, m, i, i);
    memcpy(r, rhs, sizeof(double) * m);
    double rnorm = onrm2(m, r), bnorm = onrm2(m, rhs);
    double tol = OMIN(absTol, bnorm * relTol);
    tol = OMAX(tol, 0.1 * absTol);
    if ( rnorm < tol ) {
        rc = HDSDP_RETCODE_OK; /* solution left unwritten, as the reference */
        goto done;
    }
    for ( int i = 0; i < m; ++i ) p[i] = z[i] = r[i] / jac[i];
    osymv(m, 1.0, M, p, Mp);
    for ( iter = 0; iter < maxIter; ++iter ) {
        double rz = odot(m, z, r), pMp = odot(m, p, Mp), alpha = rz / pMp;
        for ( int i = 0; i < m; ++i ) x[i] += alpha * p[i];
        if ( iter % restart == 5 ) {
            osymv(m, 1.0, M, x, p);
            for ( int i = 0; i < m; ++i ) r[i] = rhs[i] - p[i];
            for ( int i = 0; i < m; ++i ) p[i] = z[i] = r[i] / jac[i];
            osymv(m, 1.0, M, p, Mp);
            continue;
        }
        for ( int i = 0; i < m; ++i ) rn[i] = r[i] - alpha * Mp[i];
        for ( int i = 0; i < m; ++i ) z[i] = rn[i] / jac[i];
        double beta = odot(m, rn, z) / rz;
        for ( int i = 0; i < m; ++i ) p[i] = z[i] + beta * p[i];
        osymv(m, 1.0, M, p, Mp);
        memcpy(r, rn, sizeof(double) * m);
        rnorm = onrm2(m, r);
        if ( rnorm != rnorm ) goto done;
        if ( iter > 20 && rnorm > 0.01 * bnorm ) break;
        if ( rnorm < tol ) {
            converged = 1;
            break;
        }
    }
    if ( nIter ) *nIter = iter;
    if ( converged ) {
        memcpy(sol, x, sizeof(double) * m);
        rc = HDSDP_RETCODE_OK;
    }
done:
    free(buf);
    return rc;
}

/* Direct FP64 solve of M x = rhs: Cholesky of M, then two steps of iterative
 * refinement on the residual with M. */
hdsdp_retcode hdsdp_dev_kkt_solve_exact( hdsdp_dev *d, const double *rhs, double *sol ) {
    const int m = d->m;
    memcpy(d->Mchol, d->M, sizeof(double) * (size_t) m * m);
    d->stats[4] += 1;
    if ( ochol(m, d->Mchol) != 0 ) return HDSDP_RETCODE_FAILED;
    double *r = (double *) calloc((size_t) 2 * m, sizeof(double));
    if ( !r ) return HDSDP_RETCODE_MEMORY;
    double *Mx = r + m;
    memcpy(sol, rhs, sizeof(double) * m);
    otrsv(m, d->Mchol, sol, 0);
    otrsv(m, d->Mchol, sol, 1);
    for ( int it = 0; it < 2; ++it ) {
        osymv(m, 1.0, d->M, sol, Mx);
        for ( int i = 0; i < m; ++i ) r[i] = rhs[i] - Mx[i];
        otrsv(m, d->Mchol, r, 0);
        otrsv(m, d->Mchol, r, 1);
        for ( int i = 0; i < m; ++i ) sol[i] += r[i];
    }
    free(r);
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_get_schur( hdsdp_dev *d, double *M ) {
    int m = d->m;
    for ( int j = 0; j < m; ++j )
        for ( int i = j; i < m; ++i ) FE(M, m, i, j) = FE(M, m, j, i) = FE(d->M, m, i, j);
    return HDSDP_RETCODE_OK;
}

/* --------------------------------------------------------- primal recovery */

hdsdp_retcode hdsdp_dev_primal( hdsdp_dev *d, double mu, const double *y, const double *dy, int *ok ) {
    (void) y;
    int n = d->n;
    hdsdp_dev_factorize(d, 1, ok);
    if ( !*ok ) return HDSDP_RETCODE_OK;
    hdsdp_dev_update_buffer(d, 0.0, 1.0, dy, 0.0, 2);
    const double *L = d->L[1];
    double *dS = d->S[2], *aux = d->tmp, *X = d->X;
    osymmetrize_lower(n, dS);
    /* aux = L^-1 dS; transpose; X = L^-1 aux */
    memcpy(aux, dS, sizeof(double) * (size_t) n * n);
    for ( int j = 0; j < n; ++j ) otrsv(n, L, aux + (size_t) j * n, 0);
    otranspose(n, aux);
    memcpy(X, aux, sizeof(double) * (size_t) n * n);
    for ( int j = 0; j < n; ++j ) otrsv(n, L, X + (size_t) j * n, 0);
    for ( int c = 0; c < n; ++c ) {
        FE(X, n, c, c) += 1.0;
        for ( int r = c + 1; r < n; ++r ) {
            double t = 0.5 * (FE(X, n, c, r) + FE(X, n, r, c));
            FE(X, n, c, r) = FE(X, n, r, c) = t;
        }
    }
    memcpy(aux, X, sizeof(double) * (size_t) n * n);
    for ( int j = 0; j < n; ++j ) otrsv(n, L, aux + (size_t) j * n, 1);
    otranspose(n, aux);
    memcpy(X, aux, sizeof(double) * (size_t) n * n);
    for ( int j = 0; j < n; ++j ) otrsv(n, L, X + (size_t) j * n, 1);
    for ( int c = 0; c < n; ++c ) {
        FE(X, n, c, c) *= mu;
        for ( int r = c + 1; r < n; ++r ) {
            double t = 0.5 * (FE(X, n, c, r) + FE(X, n, r, c));
            FE(X, n, c, r) = FE(X, n, r, c) = mu * t;
        }
    }
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_primal_stats( hdsdp_dev *d, double *AX, double *traceCX, double *XdotS, double *traceX ) {
    int n = d->n;
    for ( int i = 0; i < d->m; ++i ) AX[i] = otrace_AB(&d->A[i + 1], n, d->X, d->vecn);
    *traceCX = otrace_AB(&d->A[0], n, d->X, d->vecn);
    memcpy(d->tmp, d->S[0], sizeof(double) * (size_t) n * n);
    osymmetrize_lower(n, d->tmp);
    *XdotS = odot((size_t) n * n, d->X, d->tmp);
    double t = 0.0;
    for ( int j = 0; j < n; ++j ) t += FE(d->X, n, j, j);
    *traceX = t;
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_get_primal( hdsdp_dev *d, double *X ) {
    memcpy(X, d->X, sizeof(double) * (size_t) d->n * d->n);
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_get_dual( hdsdp_dev *d, double *S ) {
    memcpy(S, d->S[0], sizeof(double) * (size_t) d->n * d->n);
    osymmetrize_lower(d->n, S);
    return HDSDP_RETCODE_OK;
}

/* ------------------------------------------------------- primal method */

/* X S X with X full and S's lower triangle (denseMatTripMultiplypy,
 * interface/hdsdp_psdp.c:21-52): SX = S X column by column, then XSX(i, j) += X_i . SX_j. */
static void otrip( int n, const double *S, const double *X, double *SX, double *XSX ) {
    for ( int j = 0; j < n; ++j ) osymv(n, 1.0, S, X + (size_t) j * n, SX + (size_t) j * n);
    for ( int j = 0; j < n; ++j ) {
        const double *sxj = SX + (size_t) j * n;
        for ( int i = 0; i < j; ++i ) {
            double v = odot(n, sxj, X + (size_t) i * n);
            FE(XSX, n, j, i) += v;
            FE(XSX, n, i, j) += v;
        }
        FE(XSX, n, j, j) += odot(n, sxj, X + (size_t) j * n);
    }
}

hdsdp_retcode hdsdp_dev_psdp_init( hdsdp_dev *d ) {
    size_t nn = (size_t) d->n * d->n;
    if ( !d->Xs ) {
        d->Xs = (double *) calloc(nn, sizeof(double));
        d->dXp = (double *) calloc(nn, sizeof(double));
        d->XSX = (double *) calloc(nn, sizeof(double));
        d->Sold = (double *) calloc(nn, sizeof(double));
        d->LX = (double *) calloc(nn, sizeof(double));
        lz_alloc(&d->lzP, d->n);
        if ( !d->Xs || !d->dXp || !d->XSX || !d->Sold || !d->LX || !d->lzP.V ) return HDSDP_RETCODE_MEMORY;
    }
    memcpy(d->Xs, d->X, sizeof(double) * nn);
    d->lzP.nComputed = 0;
    d->psdpReady = 1;
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_psdp_rhs( hdsdp_dev *d, double mu, double *AX, double *rhs ) {
    int n = d->n, m = d->m;
    size_t nn = (size_t) n * n;
    if ( !d->psdpReady ) return HDSDP_RETCODE_FAILED;
    memcpy(d->Sold, d->S[0], sizeof(double) * nn);
    osymmetrize_lower(n, d->Sold);
    for ( int i = 0; i < m; ++i ) AX[i] = otrace_AB(&d->A[i + 1], n, d->X, d->vecn);
    memset(d->XSX, 0, sizeof(double) * nn);
    otrip(n, d->Sold, d->X, d->tmp, d->XSX);
    for ( int i = 0; i < m; ++i ) rhs[i] = otrace_AB(&d->A[i + 1], n, d->XSX, d->vecn) - mu * AX[i];
    return HDSDP_RETCODE_OK;
}

hdsdp_retcode hdsdp_dev_psdp_direction( hdsdp_dev *d, double mu, int first, double *maxStep ) {
    int n = d->n;
    size_t nn = (size_t) n * n;
    if ( !d->psdpReady ) return HDSDP_RETCODE_FAILED;
    osymmetrize_lower(n, d->S[2]);
    otrip(n, d->S[2], d->Xs, d->tmp, d->XSX);
    for ( size_t e = 0; e < nn; ++e ) d->dXp[e] = d->X[e] - d->XSX[e] / mu;
    if ( first ) {
        memcpy(d->LX, d->X, sizeof(double) * nn);
        if ( ochol(n, d->LX) != 0 ) return HDSDP_RETCODE_FAILED;
    }
    return lanczos_run(d, &d->lzP, d->LX, d->dXp, maxStep);
}

hdsdp_retcode hdsdp_dev_psdp_step( hdsdp_dev *d, double alpha, int *isPsd, double *traceCX, double *XdotS ) {
    int n = d->n;
    size_t nn = (size_t) n * n;
    if ( !d->psdpReady ) return HDSDP_RETCODE_FAILED;
    for ( size_t e = 0; e < nn; ++e ) d->X[e] += alpha * d->dXp[e];
    memcpy(d->LX, d->X, sizeof(double) * nn);
    *isPsd = ( ochol(n, d->LX) == 0 );
    *traceCX = otrace_AB(&d->A[0], n, d->X, d->vecn);
    *XdotS = odot(nn, d->X, d->Sold);
    return HDSDP_RETCODE_OK;
}

/* ---------------------------------------------------------------- adjoint */

hdsdp_retcode hdsdp_dev_adjoint( hdsdp_dev *d, double mu, const double *dLdy, double *dLdb, double *dLdC ) {
    int m = d->m, n = d->n;
    double *lam = (double *) calloc(m, sizeof(double));
    hdsdp_retcode rc = hdsdp_dev_kkt_solve_exact(d, dLdy, lam);
    if ( rc != HDSDP_RETCODE_OK ) { free(lam); return rc; }
    for ( int i = 0; i < m; ++i ) dLdb[i] = lam[i] / mu;
    oinverse(d, d->L[0], d->W);
    memset(d->tmp2, 0, sizeof(double) * (size_t) n * n);
    for ( int i = 0; i < m; ++i ) if ( lam[i] != 0.0 ) oadd_coeff(&d->A[i + 1], lam[i], n, d->tmp2);
    osymmetrize_lower(n, d->tmp2);
    ogemm(n, d->W, d->tmp2, d->tmp);
    ogemm(n, d->tmp, d->W, dLdC);
    for ( int j = 0; j < n; ++j )
        for ( int i = j + 1; i < n; ++i ) {
            double v = 0.5 * (FE(dLdC, n, i, j) + FE(dLdC, n, j, i));
            FE(dLdC, n, i, j) = FE(dLdC, n, j, i) = v;
        }
    free(lam);
    return HDSDP_RETCODE_OK;
}

/* Profiling slots exist for ABI completeness; the oracle reports no device time. */
hdsdp_retcode hdsdp_dev_set_profiling( hdsdp_dev *d, int enable ) {
    (void) d;
    (void) enable;
    return HDSDP_RETCODE_OK;
}

int hdsdp_dev_get_profile( hdsdp_dev *d, double *ms, double *counts, int nSlots ) {
    (void) d;
    for ( int k = 0; k < nSlots && k < HDSDP_PROF_SLOTS; ++k ) {
        if ( ms ) ms[k] = 0.0;
        if ( counts ) counts[k] = 0.0;
    }
    return nSlots < HDSDP_PROF_SLOTS ? nSlots : HDSDP_PROF_SLOTS;
}
