#ifndef HDSDP_EIG_SMALL_H
#define HDSDP_EIG_SMALL_H

#ifdef __cplusplus
extern "C" {
#endif

#define HDSDP_EIG_SMALL_MAX 64

/* nTop largest eigenvalues of the symmetric tridiagonal n x n matrix with
 * diagonal d[0..n-1] and off-diagonal e[0..n-2] (n <= 64), ascending in
 * evals[0..nTop-1]; unit eigenvectors as columns of evecs (n x nTop, may be
 * NULL), largest-magnitude component positive.  Returns 0 on success, -1 on
 * invalid sizes. */
int hdsdp_eig_tridiag_top( int n, const double *d, const double *e, int nTop, double *evals, double *evecs );

#ifdef __cplusplus
}
#endif

#endif
