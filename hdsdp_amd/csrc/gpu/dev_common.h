// dev_common.h -- shared declarations of the gfx950 kernels and launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <cstdio>

#define HIP_OK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "[hdsdp_amd] HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                \
            return HDSDP_RETCODE_FAILED;                                                \
        }                                                                               \
    } while (0)

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace hdk {

constexpr int TILE = 64;   // GEMM output tile (rows and columns)
constexpr int KSTEP = 16;  // GEMM k-slab staged through LDS
constexpr int NB = 64;     // Cholesky / triangular-inverse block

// C = alpha * op(A) * op(B) + beta * C over `batch` problems (column-major).
// lower != 0 restricts the output to i >= j (tiles above the diagonal skipped).
// tri = 1 when op(B) is lower triangular, 2 when op(A) is: the k ranges that
// multiply the zero triangle are skipped (the stored zeros must be present).
template <typename T>
void gemm(hipStream_t s, bool transA, bool transB, bool lower, int M, int N, int K, T alpha,
          const T *A, int lda, long strideA, const T *B, int ldb, long strideB, T beta, T *C,
          int ldc, long strideC, int batch = 1, int tri = 0);

// gemm with the k range split into parts whose partial products go to work (P x M x N)
// and are summed in a fixed order, for products with few output tiles and long k
// (128-wide tiles fill the chip only with enough of them).  Falls back to gemm when
// work is null or too small for two parts.
// number of k parts gemm_splitk uses for this shape (1: no split)
int splitk_parts(int M, int N, int K, bool lower);
template <typename T>
void gemm_splitk(hipStream_t s, bool transA, bool transB, bool lower, int M, int N, int K, T alpha, const T *A,
                 int lda, const T *B, int ldb, T beta, T *C, int ldc, T *work, size_t workLen);

// C = alpha * X^T Y + beta * C from bf16 split products on the matrix cores: each FP32
// operand split into `terms` bf16 terms (2: three products, ~16 bits; 3: six products,
// ~24 bits, FP32-equivalent), FP32 accumulation flushed into FP64 every 256 k.
// X: K x M (ldx), Y: K x N (ldy), column-major FP32; lower restricts C to i >= j.
// tri = 1 when Y(k, j) = 0 for k > j, 2 when X(k, i) = 0 for k > i (those k are skipped).
template <typename Tout>
void gemm_bf16s_tn(hipStream_t s, int terms, bool lower, int M, int N, int K, double alpha, const float *X, int ldx,
                   long sX, const float *Y, int ldy, long sY, double beta, Tout *C, int ldc, long sC, int batch = 1,
                   int tri = 0);

// A second stream and a ring of events for look-ahead factorisations.
constexpr int LA_EVENTS = 16;
struct LookAhead {
    hipStream_t s2;
    hipEvent_t ev[LA_EVENTS];
};

// In-place lower Cholesky of the n x n matrix A (lda), blocked by NB.
// work: scratch of workLen >= POTRF_WORK elements (the leaf writes the NB x NB inverse
// of every diagonal block, padded with the identity, whatever n is); info: device int,
// 0 on success else failing column + 1.  Returns false, launching nothing, when
// workLen is too small.  With la, the bulk of each trailing update runs on la->s2
// while the next block column is factored on s; s has waited for s2 when the call
// returns.
constexpr size_t POTRF_WORK = (size_t) NB * NB;
template <typename T>
bool potrf(hipStream_t s, int n, T *A, int lda, T *work, size_t workLen, int *info, const LookAhead *la = nullptr);

// Linv = L^-1 for the lower-triangular n x n factor L (upper of Linv zeroed).
// work: n * n scratch.
template <typename T>
void trtri(hipStream_t s, int n, const T *L, int ldl, T *Linv, int ldi, T *work);

// In-place LU with partial pivoting, P A = L U (L unit lower, U upper, both in A),
// blocked by 64 columns.  ipiv (device, m ints): row interchanged with row c at step c,
// 0-based; info (device int, cleared by the caller): 0, or the first zero-pivot column + 1.
void lu_factor(hipStream_t s, int m, double *A, int lda, int *ipiv, int *info);
// x = A^-1 b from lu_factor's factors and perm, the composed interchanges (P b = b[perm]).
// x and b are distinct m-vectors.
void lu_solve(hipStream_t s, int m, const double *A, int lda, const int *perm, const double *b, double *x);

// y[j] = alpha * sum_{i in rows(j)} A[i + j*lda] x[i] + beta * y[j] for j < ncols,
// rows(j) = [0, m) (tri 0), [j, m) (tri 1, lower), [0, j] (tri 2, upper).
template <typename T>
void coldot(hipStream_t s, int m, int ncols, int tri, T alpha, const T *A, int lda, const T *x, T beta, T *y);

// y = M x for symmetric M read only in its lower triangle (one pass, tile partials
// added in a fixed order); work: symv_lower_work(m) doubles.
constexpr int SYMV_R = 512, SYMV_C = 256;
size_t symv_lower_work(int m);
void symv_lower(hipStream_t s, int m, const double *M, long ldm, const double *x, double *y, double *work);

// A[upper] = A[lower]^T for the n x n matrix.
template <typename T>
void mirror_lower(hipStream_t s, int n, T *A, int lda);
// B = A^T (n x n).
template <typename T>
void transpose(hipStream_t s, int n, const T *A, int lda, T *B, int ldb);
// Zero the strictly upper triangle.
template <typename T>
void zero_upper(hipStream_t s, int n, T *A, int lda);

void fill(hipStream_t s, double *x, long n, double v);
void fill_f(hipStream_t s, float *x, long n, float v);
void copy_lower(hipStream_t s, int n, const double *A, int lda, double *B, int ldb);
void convert_lower_to_float(hipStream_t s, int n, const double *A, int lda, float *B, int ldb);
// B = lower(diag(sc) A diag(sc) + shift I) in FP32, zeros above the diagonal
void convert_lower_scaled(hipStream_t s, int n, const double *A, int lda, const double *sc, double shift, float *B,
                          int ldb);
void add_diag(hipStream_t s, int n, double *A, int lda, const double *d, double scalar);
// out[0] = 2 * sum log A_jj
void logdet_diag(hipStream_t s, int n, const double *A, int lda, double *out);

}  // namespace hdk
