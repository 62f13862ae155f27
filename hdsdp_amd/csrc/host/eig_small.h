#ifndef HDSDP_EIG_SMALL_H
#define HDSDP_EIG_SMALL_H

#ifdef __cplusplus
extern "C" {
#endif

#define HDSDP_EIG_SMALL_MAX 64

/* nTop largest eigenvalues of the symmetric n x n column-major A (n <= 64),
 * ascending in evals[0..nTop-1], eigenvectors as columns of evecs (n x nTop,
 * may be NULL).  Returns 0 on success, -1 on invalid sizes. */
int hdsdp_eig_small_top( int n, const double *A, int nTop, double *evals, double *evecs );

#ifdef __cplusplus
}
#endif

#endif
