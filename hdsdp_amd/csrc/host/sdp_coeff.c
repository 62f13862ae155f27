/*
 * sdp_coeff.c -- presolve of SDP coefficient matrices on the host.
 *
 * Classification (linalg/hdsdp_sdpdata.c:2262-2286): no entries -> ZERO, more
 * than 30% of the packed lower triangle -> DENSE, else SPARSE.  Rank-one
 * detection (linalg/hdsdp_sdpdata.c:2314-2399 with linalg/sparse_opts.c:151-214
 * and linalg/dense_opts.c:174-226) turns an exact s a a^T into SPR1 (at most n/2
 * factor nonzeros) or DSR1, with a scaled to unit norm and the scale folded into
 * the sign (linalg/hdsdp_sdpdata.c:880-899).  Norms follow
 * linalg/hdsdp_sdpdata.c:198-309.
 */
#include "hdsdp_host.h"

static double vec_nrm2( int n, const double *x ) {
    double s = 0.0;
    for ( int i = 0; i < n; ++i ) s += x[i] * x[i];
    return sqrt(s);
}

static double vec_nrm1( int n, const double *x ) {
    double s = 0.0;
    for ( int i = 0; i < n; ++i ) s += fabs(x[i]);
    return s;
}

/* Map ascending packed-lower indices to (row, col): linalg/sparse_opts.c:122-139. */
static void packed_to_triplet( int n, int nnz, const int *pidx, int *row, int *col ) {
    int64_t j = 0, colEnd = n;  /* first packed index past column j */
    for ( int k = 0; k < nnz; ++k ) {
        while ( (int64_t) pidx[k] >= colEnd ) {
            j += 1;
            colEnd += n - j;
        }
        row[k] = (int) (pidx[k] - colEnd + n);
        col[k] = (int) j;
    }
}

hdsdp_retcode coeff_set_from_csc( sdp_coeff *A, int n, int nnz, const int *packedIdx, const double *elem ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    memset(A, 0, sizeof(sdp_coeff));
    A->n = n;
    int64_t nPack = (int64_t) n * (n + 1) / 2;
    if ( nnz == 0 ) {
        A->type = HDSDP_COEFF_ZERO;
        return retcode;
    }
    /* Entries are processed in ascending packed order (the reference sorts a
     * column that is not ascending before decompressing it). */
    int *order = NULL;
    int ascending = 1;
    for ( int k = 1; k < nnz; ++k ) {
        if ( packedIdx[k] < packedIdx[k - 1] ) { ascending = 0; break; }
    }
    int *pidx = (int *) packedIdx;
    const double *pval = elem;
    double *sortedVal = NULL;
    if ( !ascending ) {
        HALLOC(order, int, nnz);
        HALLOC(pidx, int, nnz);
        HALLOC(sortedVal, double, nnz);
        for ( int k = 0; k < nnz; ++k ) order[k] = k;
        /* insertion sort is enough for the column sizes the API sees when unsorted */
        for ( int k = 1; k < nnz; ++k ) {
            int key = order[k], p = k - 1;
            while ( p >= 0 && packedIdx[order[p]] > packedIdx[key] ) { order[p + 1] = order[p]; --p; }
            order[p + 1] = key;
        }
        for ( int k = 0; k < nnz; ++k ) { pidx[k] = packedIdx[order[k]]; sortedVal[k] = elem[order[k]]; }
        pval = sortedVal;
    }

    if ( (double) nnz > 0.3 * (double) nPack ) {
        A->type = HDSDP_COEFF_DENSE;
        A->nnz = nnz;
        HALLOC(A->val, double, nPack);
        A->ownsVal = 1;
        for ( int k = 0; k < nnz; ++k ) A->val[pidx[k]] = pval[k];
    } else {
        A->type = HDSDP_COEFF_SPARSE;
        A->nnz = nnz;
        HALLOC(A->row, int, nnz);
        HALLOC(A->col, int, nnz);
        HALLOC(A->val, double, nnz);
        A->ownsVal = 1;
        packed_to_triplet(n, nnz, pidx, A->row, A->col);
        memcpy(A->val, pval, sizeof(double) * nnz);
    }

exit_cleanup:
    if ( !ascending ) {
        free(order);
        free(pidx);
        free(sortedVal);
    }
    return retcode;
}

static void rank_one_normalize( int nnz, double *vals, double *sign, double *full, int n ) {
    double aScal = vec_nrm2(nnz, vals);
    *sign *= aScal * aScal;
    double inv = 1.0 / aScal;
    for ( int i = 0; i < nnz; ++i ) vals[i] *= inv;
    if ( full ) {
        for ( int i = 0; i < n; ++i ) full[i] *= inv;
    }
}

/* Replace A by the rank-one form sign * a a^T (a dense length n), choosing
 * SPR1/DSR1 by the count of |a_i| > 1e-10 (linalg/hdsdp_sdpdata.c:2327-2390). */
static hdsdp_retcode coeff_make_rank_one( sdp_coeff *A, double sign, const double *a ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    int n = A->n, nz = 0;
    for ( int i = 0; i < n; ++i ) nz += ( fabs(a[i]) > 1e-10 );
    if ( A->ownsVal ) HFREE(A->val);
    HFREE(A->row);
    HFREE(A->col);
    if ( nz > 0.5 * n ) {
        A->type = HDSDP_COEFF_DSR1;
        A->nnz = n;
        HALLOC(A->val, double, n);
        A->ownsVal = 1;
        memcpy(A->val, a, sizeof(double) * n);
        A->sign = sign;
        rank_one_normalize(n, A->val, &A->sign, NULL, n);
    } else {
        A->type = HDSDP_COEFF_SPR1;
        A->nnz = nz;
        HALLOC(A->row, int, nz);
        HALLOC(A->val, double, nz);
        A->ownsVal = 1;
        A->sign = sign;
        for ( int i = 0, k = 0; i < n; ++i ) {
            if ( fabs(a[i]) > 1e-10 ) { A->row[k] = i; A->val[k] = a[i]; ++k; }
        }
        rank_one_normalize(nz, A->val, &A->sign, NULL, n);
    }
exit_cleanup:
    return retcode;
}

/* Exact rank-one test of lower triplets, first entry must be diagonal
 * (linalg/sparse_opts.c:151-214, including its k == n rejection). */
static int sparse_rank_one( int n, int nnz, const int *Ai, const int *Aj, const double *Ax, double *sgn, double *a ) {
    int i = Ai[0], j = Aj[0];
    double v = Ax[0];
    if ( i != j ) return 0;
    if ( nnz == 1 ) {
        *sgn = Ax[0];
        a[i] = 1.0;
        return 1;
    }
    double s = ( v > 0 ) ? 1.0 : -1.0;
    v = sqrt(fabs(v));
    int k = 0, anz = 0;
    for ( k = 0; k < nnz; ++k ) {
        if ( Aj[k] > i ) break;
        a[Ai[k]] = Ax[k] / v;
        anz += 1;
    }
    if ( nnz != anz * (anz + 1) / 2 ) return 0;
    if ( k == n ) return 0;
    double eps = 0.0;
    for ( int q = 0; q < nnz; ++q ) {
        double prod = a[Ai[q]] * a[Aj[q]];
        eps += ( s == 1.0 ) ? fabs(Ax[q] - prod) : fabs(Ax[q] + prod);
    }
    if ( eps > 1e-10 ) return 0;
    *sgn = s;
    return 1;
}

/* Exact rank-one test of a packed dense matrix anchored at its first nonzero
 * diagonal entry (linalg/dense_opts.c:174-226; the anchor column is read with
 * the same packed addressing for every row). */
static int dense_rank_one( int n, const double *A, double *sgn, double *a ) {
    int i;
    int64_t k = 0;
    for ( i = 0; i < n; ++i ) {
        if ( A[k] != 0 ) break;
        k += n - i;
    }
    if ( i == n ) return 0;
    double s = ( A[k] > 0 ) ? 1.0 : -1.0;
    double v = sqrt(fabs(A[k]));
    int64_t colStart = (int64_t) (2 * (int64_t) n - i - 1) * i / 2;
    for ( int r = 0; r < n; ++r ) a[r] = A[colStart + r] / v;
    double eps = 0.0;
    int64_t id = 0;
    for ( int c = 0; c < n; ++c ) {
        const double *p = A + id;
        for ( int r = 0; r < n - c; ++r ) {
            double prod = a[c] * a[c + r];
            eps += ( s == 1.0 ) ? fabs(p[r] - prod) : fabs(p[r] + prod);
        }
        id += n - c;
        if ( eps > 1e-10 ) return 0;
    }
    *sgn = s;
    return 1;
}

hdsdp_retcode coeff_presolve_rank_one( sdp_coeff *A ) {
    hdsdp_retcode retcode = HDSDP_RETCODE_OK;
    double *a = NULL;
    if ( A->type != HDSDP_COEFF_SPARSE && A->type != HDSDP_COEFF_DENSE ) return retcode;
    HALLOC(a, double, A->n);
    double sgn = 0.0;
    int isRankOne = ( A->type == HDSDP_COEFF_SPARSE )
        ? sparse_rank_one(A->n, A->nnz, A->row, A->col, A->val, &sgn, a)
        : dense_rank_one(A->n, A->val, &sgn, a);
    if ( isRankOne ) {
        HCALL(coeff_make_rank_one(A, sgn, a));
    }
exit_cleanup:
    free(a);
    return retcode;
}

hdsdp_retcode coeff_set_rank_one( sdp_coeff *A, int n, double sign, const double *factor ) {
    memset(A, 0, sizeof(sdp_coeff));
    A->n = n;
    A->type = HDSDP_COEFF_DENSE;  /* placeholder until coeff_make_rank_one */
    return coeff_make_rank_one(A, sign, factor);
}

void coeff_scale( sdp_coeff *A, double alpha ) {
    int64_t nPack = (int64_t) A->n * (A->n + 1) / 2;
    switch ( A->type ) {
        case HDSDP_COEFF_SPARSE:
            for ( int k = 0; k < A->nnz; ++k ) A->val[k] *= alpha;
            break;
        case HDSDP_COEFF_DENSE:
            for ( int64_t k = 0; k < nPack; ++k ) A->val[k] *= alpha;
            break;
        case HDSDP_COEFF_SPR1:
        case HDSDP_COEFF_DSR1:
            A->sign *= alpha;
            break;
        default:
            break;
    }
}

double coeff_norm( const sdp_coeff *A, int type ) {
    double nrm = 0.0;
    int n = A->n;
    switch ( A->type ) {
        case HDSDP_COEFF_SPARSE:
            for ( int k = 0; k < A->nnz; ++k ) {
                double x = A->val[k];
                int diag = ( A->row[k] == A->col[k] );
                nrm += ( type == FRO_NORM ) ? ( diag ? x * x : 2.0 * x * x )
                                            : ( diag ? fabs(x) : 2.0 * fabs(x) );
            }
            return ( type == FRO_NORM ) ? sqrt(nrm) : nrm;
        case HDSDP_COEFF_DENSE: {
            const double *p = A->val;
            for ( int c = 0; c < n; ++c ) {
                int len = n - c - 1;
                if ( type == FRO_NORM ) {
                    double cn = vec_nrm2(len, p + 1);
                    nrm += p[0] * p[0] + 2.0 * cn * cn;
                } else {
                    nrm += fabs(p[0]) + 2.0 * vec_nrm1(len, p + 1);
                }
                p += len + 1;
            }
            return ( type == FRO_NORM ) ? sqrt(nrm) : nrm;
        }
        case HDSDP_COEFF_SPR1:
        case HDSDP_COEFF_DSR1: {
            double s = ( type == FRO_NORM ) ? vec_nrm2(A->nnz, A->val) : vec_nrm1(A->nnz, A->val);
            return s * s * fabs(A->sign);
        }
        default:
            return 0.0;
    }
}

int coeff_rank( const sdp_coeff *A ) {
    if ( A->type == HDSDP_COEFF_ZERO ) return 0;
    if ( A->type == HDSDP_COEFF_SPR1 || A->type == HDSDP_COEFF_DSR1 ) return 1;
    return A->n;
}

int coeff_nnz( const sdp_coeff *A ) {
    switch ( A->type ) {
        case HDSDP_COEFF_SPARSE: return A->nnz;
        case HDSDP_COEFF_DENSE: return A->n * (A->n + 1) / 2;
        case HDSDP_COEFF_SPR1: return A->nnz * (A->nnz + 1) / 2;
        case HDSDP_COEFF_DSR1: return A->n * (A->n + 1) / 2;
        default: return 0;
    }
}

int coeff_is_eye( const sdp_coeff *A, double *mult ) {
    if ( A->type != HDSDP_COEFF_SPARSE || A->nnz != A->n ) return 0;
    double d = A->val[0];
    for ( int k = 0; k < A->nnz; ++k ) {
        if ( A->row[k] != A->col[k] || A->val[k] != d ) return 0;
    }
    if ( mult ) *mult = d;
    return 1;
}

int coeff_is_unit_col( const sdp_coeff *A, int *col ) {
    if ( A->type == HDSDP_COEFF_SPR1 && A->nnz == 1 && A->val[0] == 1.0 && A->sign == 1.0 ) {
        if ( col ) *col = A->row[0];
        return 1;
    }
    return 0;
}

void coeff_clear( sdp_coeff *A ) {
    if ( !A ) return;
    if ( A->ownsVal ) free(A->val);
    free(A->row);
    free(A->col);
    memset(A, 0, sizeof(sdp_coeff));
}

void coeff_view( const sdp_coeff *A, hdsdp_coeff_view *v ) {
    v->type = A->type;
    v->nnz = A->nnz;
    v->row = A->row;
    v->col = A->col;
    v->val = A->val;
    v->sign = A->sign;
}
