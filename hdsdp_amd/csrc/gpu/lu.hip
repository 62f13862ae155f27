// lu.hip -- LU factorisation with partial pivoting of the Schur matrix and its direct
// solve: the indefinite solver the Schur system switches to once its Cholesky
// factorisation, or the CG preconditioned by it, fails (the reference's LDL^T switch,
// linalg/hdsdp_linsolver.c:1236-1266, 1367-1425).  Right-looking and blocked by 64
// columns: a one-workgroup panel kernel (pivot search, row swaps, scaling, rank-1
// updates inside the panel), the panel's row swaps on the other columns, the unit-lower
// triangular solve of the row block, and the trailing update on the FP64 MFMA GEMM.
#include "dev_common.h"

namespace hdk {
namespace {

constexpr int LU_NB = 64;         // panel width
constexpr int LU_THREADS = 1024;  // one workgroup: panel and solve kernels
constexpr int LU_SOLVE_B = 256;   // diagonal block of the triangular solves

// larger |value| wins; ties go to the smaller row (the first maximum, as idamax)
__device__ inline void take_max(double &v, int &i, double v2, int i2) {
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

__global__ __launch_bounds__(LU_THREADS) void lu_panel_kernel(int m, int k0, int nb, double *A, long lda, int *ipiv,
                                                              int *info) {
    __shared__ double wv[LU_THREADS / 64];
    __shared__ int wi[LU_THREADS / 64];
    __shared__ double piv;
    __shared__ int prow;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int j = 0; j < nb; ++j) {
        const int c = k0 + j;
        double *col = A + (long) c * lda;
        double v = -1.0;
        int vi = m;
        for (int r = c + tid; r < m; r += LU_THREADS) take_max(v, vi, fabs(col[r]), r);
        for (int off = 32; off > 0; off >>= 1) {
            const double v2 = __shfl_xor(v, off);
            const int i2 = __shfl_xor(vi, off);
            take_max(v, vi, v2, i2);
        }
        if (lane == 0) {
            wv[wid] = v;
            wi[wid] = vi;
        }
        __syncthreads();
        if (tid == 0) {
            double bv = wv[0];
            int bi = wi[0];
            for (int w = 1; w < LU_THREADS / 64; ++w) take_max(bv, bi, wv[w], wi[w]);
            if (bi >= m) bi = c;
            ipiv[c] = bi;
            prow = bi;
            piv = col[bi];
            if (!(bv > 0.0) && *info == 0) *info = c + 1;  // zero (or NaN) pivot column
        }
        __syncthreads();
        const int p = prow;
        if (p != c) {
            for (int jj = tid; jj < nb; jj += LU_THREADS) {
                double *cj = A + (long) (k0 + jj) * lda;
                const double t = cj[c];
                cj[c] = cj[p];
                cj[p] = t;
            }
        }
        __syncthreads();
        const double pv = piv;
        if (pv != 0.0) {
            const double rp = 1.0 / pv;
            for (int r = c + 1 + tid; r < m; r += LU_THREADS) col[r] *= rp;
        }
        __syncthreads();
        for (int r = c + 1 + tid; r < m; r += LU_THREADS) {
            const double l = col[r];
            for (int jj = j + 1; jj < nb; ++jj) {
                double *cj = A + (long) (k0 + jj) * lda;
                cj[r] -= l * cj[c];
            }
        }
        __syncthreads();
    }
}

// the panel's row interchanges on every column outside it (one column per thread)
__global__ void lu_laswp_kernel(int m, int k0, int nb, double *A, long lda, const int *ipiv) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m || (j >= k0 && j < k0 + nb)) return;
    double *cj = A + (long) j * lda;
    for (int t = 0; t < nb; ++t) {
        const int c = k0 + t, p = ipiv[c];
        if (p != c) {
            const double x = cj[c];
            cj[c] = cj[p];
            cj[p] = x;
        }
    }
}

// A12 <- L11^-1 A12 (L11 the panel's unit lower nb x nb block): one column per thread,
// the columns of the workgroup staged in LDS (row-major by thread, conflict-free)
constexpr int LU_TRSM_T = 128;
__global__ __launch_bounds__(LU_TRSM_T) void lu_trsm_kernel(int m, int k0, int nb, double *A, long lda) {
    __shared__ double L[LU_NB * LU_NB];
    __shared__ double X[LU_NB * LU_TRSM_T];
    const double *Ad = A + k0 + (long) k0 * lda;
    for (int e = threadIdx.x; e < LU_NB * LU_NB; e += blockDim.x) {
        const int r = e % LU_NB, cc = e / LU_NB;
        L[e] = (r < nb && cc < nb) ? Ad[r + (long) cc * lda] : 0.0;
    }
    const int j0 = k0 + nb + blockIdx.x * LU_TRSM_T;
    // stage the block's columns: consecutive threads read consecutive rows of a column
    for (int e = threadIdx.x; e < LU_NB * LU_TRSM_T; e += blockDim.x) {
        const int r = e % LU_NB, t = e / LU_NB;
        X[r * LU_TRSM_T + t] = (r < nb && j0 + t < m) ? A[k0 + r + (long) (j0 + t) * lda] : 0.0;
    }
    __syncthreads();
    const int t = threadIdx.x;
    for (int cc = 0; cc < nb; ++cc) {
        const double xc = X[cc * LU_TRSM_T + t];
        for (int r = cc + 1; r < nb; ++r) X[r * LU_TRSM_T + t] -= L[r + cc * LU_NB] * xc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < LU_NB * LU_TRSM_T; e += blockDim.x) {
        const int r = e % LU_NB, tt = e / LU_NB;
        if (r < nb && j0 + tt < m) A[k0 + r + (long) (j0 + tt) * lda] = X[r * LU_TRSM_T + tt];
    }
}

// x = U^-1 L^-1 (P b) with P b = b[perm]; one workgroup, x in global memory, the
// current diagonal block of the solution staged in LDS.
__global__ __launch_bounds__(LU_THREADS) void lu_solve_kernel(int m, const double *A, long lda, const int *perm,
                                                              const double *b, double *x) {
    __shared__ double xs[LU_SOLVE_B];
    const int tid = threadIdx.x;
    for (int i = tid; i < m; i += LU_THREADS) x[i] = b[perm[i]];
    __syncthreads();
    for (int kb = 0; kb < m; kb += LU_SOLVE_B) {
        const int nbk = min(LU_SOLVE_B, m - kb);
        if (tid < nbk) xs[tid] = x[kb + tid];
        __syncthreads();
        for (int c = 0; c < nbk; ++c) {
            if (tid > c && tid < nbk) xs[tid] -= A[kb + tid + (long) (kb + c) * lda] * xs[c];
            __syncthreads();
        }
        if (tid < nbk) x[kb + tid] = xs[tid];
        for (int i = kb + nbk + tid; i < m; i += LU_THREADS) {
            double acc = 0.0;
            for (int c = 0; c < nbk; ++c) acc += A[i + (long) (kb + c) * lda] * xs[c];
            x[i] -= acc;
        }
        __syncthreads();
    }
    for (int kb = ((m - 1) / LU_SOLVE_B) * LU_SOLVE_B; kb >= 0; kb -= LU_SOLVE_B) {
        const int nbk = min(LU_SOLVE_B, m - kb);
        if (tid < nbk) xs[tid] = x[kb + tid];
        __syncthreads();
        for (int c = nbk - 1; c >= 0; --c) {
            if (tid == c) xs[c] /= A[kb + c + (long) (kb + c) * lda];
            __syncthreads();
            if (tid < c) xs[tid] -= A[kb + tid + (long) (kb + c) * lda] * xs[c];
        }
        __syncthreads();
        if (tid < nbk) x[kb + tid] = xs[tid];
        for (int i = tid; i < kb; i += LU_THREADS) {
            double acc = 0.0;
            for (int c = 0; c < nbk; ++c) acc += A[i + (long) (kb + c) * lda] * xs[c];
            x[i] -= acc;
        }
        __syncthreads();
    }
}

}  // namespace

void lu_factor(hipStream_t s, int m, double *A, int lda, int *ipiv, int *info) {
    for (int k0 = 0; k0 < m; k0 += LU_NB) {
        const int nb = std::min(LU_NB, m - k0);
        lu_panel_kernel<<<1, LU_THREADS, 0, s>>>(m, k0, nb, A, lda, ipiv, info);
        lu_laswp_kernel<<<(m + 255) / 256, 256, 0, s>>>(m, k0, nb, A, lda, ipiv);
        const int rest = m - k0 - nb;
        if (rest > 0) {
            lu_trsm_kernel<<<(rest + LU_TRSM_T - 1) / LU_TRSM_T, LU_TRSM_T, 0, s>>>(m, k0, nb, A, lda);
            gemm<double>(s, false, false, false, rest, rest, nb, -1.0, A + (k0 + nb) + (long) k0 * lda, lda, 0,
                         A + k0 + (long) (k0 + nb) * lda, lda, 0, 1.0, A + (k0 + nb) + (long) (k0 + nb) * lda, lda, 0,
                         1);
        }
    }
}

void lu_solve(hipStream_t s, int m, const double *A, int lda, const int *perm, const double *b, double *x) {
    lu_solve_kernel<<<1, LU_THREADS, 0, s>>>(m, A, lda, perm, b, x);
}

}  // namespace hdk
