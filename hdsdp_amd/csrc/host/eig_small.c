/*
 * eig_small.c -- top eigenpairs of a small symmetric tridiagonal matrix.
 *
 * Replaces the dsyevr call that the Lanczos step-length routine makes on its
 * (k+1) x (k+1) projected matrix (linalg/dense_opts.c:53-73,
 * linalg/hdsdp_lanczos.c:222-233), which is the Lanczos tridiagonal: the two
 * largest eigenvalues in ascending order and their eigenvectors.
 * Eigenvalues: bisection on Sturm counts of T - xI (the count of negative
 * pivots of its LDL^T factorisation), to a few ulps.  Eigenvectors: inverse
 * iteration with the partially pivoted tridiagonal LU, orthogonalised against
 * the previous vector when the two eigenvalues are close (Golub & Van Loan,
 * Matrix Computations, 4th ed., secs. 8.4.1 and 8.2.2; LAPACK dstebz/dstein).
 */
#include <float.h>
#include <math.h>
#include <string.h>

#include "eig_small.h"

/* number of eigenvalues of T below x */
static int sturm_count( int n, const double *d, const double *e2, double x, double pivmin ) {
    int cnt = 0;
    double q = d[0] - x;
    if ( fabs(q) < pivmin ) q = -pivmin;
    if ( q < 0.0 ) ++cnt;
    for ( int i = 1; i < n; ++i ) {
        q = d[i] - x - e2[i - 1] / q;
        if ( fabs(q) < pivmin ) q = -pivmin;
        if ( q < 0.0 ) ++cnt;
    }
    return cnt;
}

/* Solve (T - lam I) x = b in place (b -> x) by Gaussian elimination with
 * partial pivoting; zero pivots are replaced by tiny. */
static void tridiag_shifted_solve( int n, const double *d, const double *e, double lam, double tiny, double *b ) {
    double D[HDSDP_EIG_SMALL_MAX], DL[HDSDP_EIG_SMALL_MAX], DU[HDSDP_EIG_SMALL_MAX], DU2[HDSDP_EIG_SMALL_MAX];
    for ( int i = 0; i < n; ++i ) D[i] = d[i] - lam;
    for ( int i = 0; i + 1 < n; ++i ) {
        DL[i] = e[i];
        DU[i] = e[i];
        DU2[i] = 0.0;
    }
    for ( int i = 0; i + 1 < n; ++i ) {
        if ( fabs(D[i]) >= fabs(DL[i]) ) {
            if ( D[i] == 0.0 ) D[i] = tiny;
            double f = DL[i] / D[i];
            D[i + 1] -= f * DU[i];
            b[i + 1] -= f * b[i];
            DU2[i] = 0.0;
        } else {
            double f = D[i] / DL[i];
            D[i] = DL[i];
            double t = D[i + 1];
            D[i + 1] = DU[i] - f * t;
            if ( i + 2 < n ) {
                DU2[i] = DU[i + 1];
                DU[i + 1] = -f * DU2[i];
            }
            DU[i] = t;
            t = b[i];
            b[i] = b[i + 1];
            b[i + 1] = t - f * b[i + 1];
        }
    }
    if ( D[n - 1] == 0.0 ) D[n - 1] = tiny;
    b[n - 1] /= D[n - 1];
    if ( n > 1 ) b[n - 2] = (b[n - 2] - DU[n - 2] * b[n - 1]) / D[n - 2];
    for ( int i = n - 3; i >= 0; --i ) b[i] = (b[i] - DU[i] * b[i + 1] - DU2[i] * b[i + 2]) / D[i];
}

static double normalize( int n, double *x ) {
    double s = 0.0;
    for ( int i = 0; i < n; ++i ) s += x[i] * x[i];
    s = sqrt(s);
    if ( s > 0.0 ) for ( int i = 0; i < n; ++i ) x[i] /= s;
    return s;
}

int hdsdp_eig_tridiag_top( int n, const double *d, const double *e, int nTop, double *evals, double *evecs ) {
    if ( n < 1 || nTop < 1 || nTop > n || n > HDSDP_EIG_SMALL_MAX ) return -1;
    double e2[HDSDP_EIG_SMALL_MAX];
    double lo = d[0], hi = d[0], emax2 = 0.0;
    for ( int i = 0; i < n; ++i ) {
        double r = ( i > 0 ? fabs(e[i - 1]) : 0.0 ) + ( i + 1 < n ? fabs(e[i]) : 0.0 );
        if ( d[i] - r < lo ) lo = d[i] - r;
        if ( d[i] + r > hi ) hi = d[i] + r;
        if ( i + 1 < n ) {
            e2[i] = e[i] * e[i];
            if ( e2[i] > emax2 ) emax2 = e2[i];
        }
    }
    const double tnorm = fmax(fabs(lo), fabs(hi));
    const double pivmin = DBL_MIN * fmax(1.0, emax2);
    lo -= 2.0 * DBL_EPSILON * tnorm + pivmin;
    hi += 2.0 * DBL_EPSILON * tnorm + pivmin;
    for ( int t = 0; t < nTop; ++t ) {
        const int k = n - nTop + t;  /* k-th smallest, 0-based */
        double a = lo, b = hi;
        for ( int it = 0; it < 200; ++it ) {
            double mid = 0.5 * (a + b);
            if ( b - a <= 2.0 * DBL_EPSILON * fmax(fabs(a), fabs(b)) + pivmin || mid == a || mid == b ) break;
            if ( sturm_count(n, d, e2, mid, pivmin) > k ) b = mid;
            else a = mid;
        }
        evals[t] = 0.5 * (a + b);
    }
    if ( !evecs ) return 0;
    const double tiny = DBL_EPSILON * fmax(tnorm, DBL_MIN);
    for ( int t = 0; t < nTop; ++t ) {
        double *x = evecs + (size_t) t * n;
        for ( int i = 0; i < n; ++i ) x[i] = 1.0 + 0.25 * sin(1.0 + 0.7 * i);
        for ( int it = 0; it < 3; ++it ) {
            tridiag_shifted_solve(n, d, e, evals[t], tiny, x);
            /* keep clustered vectors orthogonal to the ones already computed */
            for ( int s = 0; s < t; ++s ) {
                if ( fabs(evals[t] - evals[s]) > 1e-3 * fmax(tnorm, DBL_MIN) ) continue;
                const double *y = evecs + (size_t) s * n;
                double dp = 0.0;
                for ( int i = 0; i < n; ++i ) dp += x[i] * y[i];
                for ( int i = 0; i < n; ++i ) x[i] -= dp * y[i];
            }
            if ( normalize(n, x) == 0.0 ) x[t % n] = 1.0;
        }
        int big = 0;
        for ( int i = 1; i < n; ++i ) if ( fabs(x[i]) > fabs(x[big]) ) big = i;
        if ( x[big] < 0.0 ) for ( int i = 0; i < n; ++i ) x[i] = -x[i];
    }
    return 0;
}
