/*
 * eig_small.c -- eigenpairs of a small dense symmetric matrix.
 *
 * Replaces the dsyevr call that the Lanczos step-length routine makes on its
 * (k+1) x (k+1) projected matrix, k < 31 (linalg/dense_opts.c:53-73,
 * linalg/hdsdp_lanczos.c:222-233): the two largest eigenvalues in ascending
 * order and their eigenvectors.  Method: cyclic Jacobi rotations
 * (Golub & Van Loan, Matrix Computations, 4th ed., sec. 8.5.2), which is
 * accurate to a few ulps of ||A|| at these sizes.
 */
#include "hdsdp_host.h"
#include "eig_small.h"

int hdsdp_eig_small_top( int n, const double *A, int nTop, double *evals, double *evecs ) {
    if ( n < 1 || nTop < 1 || nTop > n || n > HDSDP_EIG_SMALL_MAX ) return -1;
    double a[HDSDP_EIG_SMALL_MAX * HDSDP_EIG_SMALL_MAX];
    double v[HDSDP_EIG_SMALL_MAX * HDSDP_EIG_SMALL_MAX];
    for ( int j = 0; j < n; ++j ) {
        for ( int i = 0; i < n; ++i ) {
            a[i + j * n] = 0.5 * (A[i + j * n] + A[j + i * n]);
            v[i + j * n] = ( i == j ) ? 1.0 : 0.0;
        }
    }
    for ( int sweep = 0; sweep < 100; ++sweep ) {
        double off = 0.0, diag = 0.0;
        for ( int j = 0; j < n; ++j ) {
            diag += a[j + j * n] * a[j + j * n];
            for ( int i = j + 1; i < n; ++i ) off += a[i + j * n] * a[i + j * n];
        }
        if ( off <= 1e-32 * diag || off == 0.0 ) break;
        for ( int p = 0; p < n - 1; ++p ) {
            for ( int q = p + 1; q < n; ++q ) {
                double apq = a[p + q * n];
                if ( apq == 0.0 ) continue;
                double app = a[p + p * n], aqq = a[q + q * n];
                double theta = (aqq - app) / (2.0 * apq);
                double t = ( theta >= 0 ? 1.0 : -1.0 ) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for ( int k = 0; k < n; ++k ) {
                    double akp = a[k + p * n], akq = a[k + q * n];
                    a[k + p * n] = c * akp - s * akq;
                    a[k + q * n] = s * akp + c * akq;
                }
                for ( int k = 0; k < n; ++k ) {
                    double apk = a[p + k * n], aqk = a[q + k * n];
                    a[p + k * n] = c * apk - s * aqk;
                    a[q + k * n] = s * apk + c * aqk;
                }
                for ( int k = 0; k < n; ++k ) {
                    double vkp = v[k + p * n], vkq = v[k + q * n];
                    v[k + p * n] = c * vkp - s * vkq;
                    v[k + q * n] = s * vkp + c * vkq;
                }
            }
        }
    }
    /* select the nTop largest, reported in ascending order */
    int idx[HDSDP_EIG_SMALL_MAX];
    for ( int i = 0; i < n; ++i ) idx[i] = i;
    for ( int i = 1; i < n; ++i ) {
        int key = idx[i], p = i - 1;
        while ( p >= 0 && a[idx[p] + idx[p] * n] > a[key + key * n] ) { idx[p + 1] = idx[p]; --p; }
        idx[p + 1] = key;
    }
    for ( int t = 0; t < nTop; ++t ) {
        int c = idx[n - nTop + t];
        evals[t] = a[c + c * n];
        if ( evecs ) {
            /* sign convention: largest-magnitude component positive */
            int big = 0;
            for ( int k = 1; k < n; ++k ) if ( fabs(v[k + c * n]) > fabs(v[big + c * n]) ) big = k;
            double sg = ( v[big + c * n] < 0 ) ? -1.0 : 1.0;
            for ( int k = 0; k < n; ++k ) evecs[k + t * n] = sg * v[k + c * n];
        }
    }
    return 0;
}
