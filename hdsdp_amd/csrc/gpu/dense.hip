// dense.hip -- dense linear algebra for the dual slack S and the Schur matrix M
// on gfx950: MFMA GEMM, blocked Cholesky, recursive triangular inverse, GEMV.
//
// GEMM tiles are 64 x 64 per workgroup of four 64-wide waves; each wave owns a
// 32 x 32 quadrant as 2 x 2 MFMA 16x16 accumulators and the k-dimension is
// staged through LDS in slabs of 16.  FP64 uses v_mfma_f64_16x16x4_f64, FP32
// v_mfma_f32_16x16x4_f32; both take one A and one B element per lane,
//   A[i][k] at lane i + 16 k,  B[k][j] at lane j + 16 k,
// and return four accumulator elements per lane in different row orders
// (measured on gfx950):
//   FP32: D[i][j] at lane j + 16 (i / 4), element i % 4;
//   FP64: D[i][j] at lane j + 16 (i % 4), element i / 4.
#include "dev_common.h"

namespace hdk {

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
    typedef v4d acc_t;
    // row of accumulator element r held by lane group g = lane / 16
    __device__ static inline int row(int g, int r) { return 4 * r + g; }
    __device__ static inline acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma<float> {
    typedef v4f acc_t;
    __device__ static inline int row(int g, int r) { return 4 * g + r; }
    __device__ static inline acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <typename T, bool TA, bool TB, bool LOWER>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, T alpha, const T *A, int lda,
                                                   long sA, const T *B, int ldb, long sB, T beta,
                                                   T *C, int ldc, long sC) {
    typedef typename Mfma<T>::acc_t acc_t;
    const int i0 = blockIdx.x * TILE, j0 = blockIdx.y * TILE;
    if (LOWER && i0 + TILE - 1 < j0) return;
    A += (long) blockIdx.z * sA;
    B += (long) blockIdx.z * sB;
    C += (long) blockIdx.z * sC;
    __shared__ T As[KSTEP][TILE + 1];
    __shared__ T Bs[KSTEP][TILE + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = (acc_t){0, 0, 0, 0};

    for (int k0 = 0; k0 < K; k0 += KSTEP) {
#pragma unroll
        for (int e = tid; e < TILE * KSTEP; e += 256) {
            int i, k;
            if (!TA) { i = e & 63; k = e >> 6; } else { k = e & 15; i = e >> 4; }
            int gi = i0 + i, gk = k0 + k;
            T v = 0;
            if (gi < M && gk < K) v = TA ? A[gk + (long) gi * lda] : A[gi + (long) gk * lda];
            As[k][i] = v;
        }
#pragma unroll
        for (int e = tid; e < TILE * KSTEP; e += 256) {
            int j, k;
            if (!TB) { k = e & 15; j = e >> 4; } else { j = e & 63; k = e >> 6; }
            int gj = j0 + j, gk = k0 + k;
            T v = 0;
            if (gj < N && gk < K) v = TB ? B[gj + (long) gk * ldb] : B[gk + (long) gj * ldb];
            Bs[k][j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KSTEP; kk += 4) {
            T a0 = As[kk + lk][wi + lr], a1 = As[kk + lk][wi + 16 + lr];
            T b0 = Bs[kk + lk][wj + lr], b1 = Bs[kk + lk][wj + 16 + lr];
            acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
            acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
            acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
            acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int i = i0 + wi + a * 16 + Mfma<T>::row(lk, r);
                int j = j0 + wj + b * 16 + lr;
                if (i < M && j < N && (!LOWER || i >= j)) {
                    T v = alpha * acc[a][b][r];
                    if (beta != (T) 0) v += beta * C[i + (long) j * ldc];
                    C[i + (long) j * ldc] = v;
                }
            }
}

template <typename T>
void gemm(hipStream_t s, bool transA, bool transB, bool lower, int M, int N, int K, T alpha, const T *A,
          int lda, long strideA, const T *B, int ldb, long strideB, T beta, T *C, int ldc, long strideC,
          int batch) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    dim3 grid((M + TILE - 1) / TILE, (N + TILE - 1) / TILE, batch);
#define HDK_GEMM_CASE(ta, tb, lo)                                                                    \
    if (transA == ta && transB == tb && lower == lo) {                                               \
        gemm_kernel<T, ta, tb, lo><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, \
                                                        beta, C, ldc, strideC);                     \
        return;                                                                                      \
    }
    HDK_GEMM_CASE(false, false, false)
    HDK_GEMM_CASE(false, true, false)
    HDK_GEMM_CASE(true, false, false)
    HDK_GEMM_CASE(true, true, false)
    HDK_GEMM_CASE(false, false, true)
    HDK_GEMM_CASE(false, true, true)
    HDK_GEMM_CASE(true, false, true)
    HDK_GEMM_CASE(true, true, true)
#undef HDK_GEMM_CASE
}

template void gemm<double>(hipStream_t, bool, bool, bool, int, int, int, double, const double *, int, long,
                           const double *, int, long, double, double *, int, long, int);
template void gemm<float>(hipStream_t, bool, bool, bool, int, int, int, float, const float *, int, long,
                          const float *, int, long, float, float *, int, long, int);

// ---- Cholesky --------------------------------------------------------------

// Factor the kb x kb diagonal block at A (lower) in LDS and write the inverse of
// its factor to Dinv (NB x NB, lower, zeros above).  Sets *info = k0 + j + 1 at
// the first non-positive pivot.
template <typename T>
__global__ __launch_bounds__(256) void potf2_kernel(T *A, int lda, int k0, int kb, int *info, T *Dinv) {
    __shared__ T L[NB][NB + 1];
    __shared__ T X[NB][NB + 1];
    __shared__ int fail;
    if (*info) return;
    const int tid = threadIdx.x;
    for (int e = tid; e < kb * kb; e += 256) {
        int r = e % kb, c = e / kb;
        L[r][c] = (r >= c) ? A[r + (long) c * lda] : (T) 0;
    }
    if (tid == 0) fail = 0;
    __syncthreads();
    for (int j = 0; j < kb; ++j) {
        if (tid == 0) {
            T d = L[j][j];
            if (!(d > (T) 0)) fail = j + 1;
            else L[j][j] = sqrt(d);
        }
        __syncthreads();
        if (fail) break;
        T inv = (T) 1 / L[j][j];
        for (int r = j + 1 + tid; r < kb; r += 256) L[r][j] *= inv;
        __syncthreads();
        int rem = kb - j - 1;
        for (int e = tid; e < rem * rem; e += 256) {
            int r = j + 1 + e % rem, c = j + 1 + e / rem;
            if (r >= c) L[r][c] -= L[r][j] * L[c][j];
        }
        __syncthreads();
    }
    if (fail) {
        if (tid == 0) *info = k0 + fail;
        return;
    }
    // inverse of the block factor, one column per thread
    if (tid < kb) {
        int c = tid;
        X[c][c] = (T) 1 / L[c][c];
        for (int r = c + 1; r < kb; ++r) {
            T sum = 0;
            for (int l = c; l < r; ++l) sum += L[r][l] * X[l][c];
            X[r][c] = -sum / L[r][r];
        }
    }
    __syncthreads();
    for (int e = tid; e < kb * kb; e += 256) {
        int r = e % kb, c = e / kb;
        if (r >= c) A[r + (long) c * lda] = L[r][c];
        Dinv[r + c * NB] = (r >= c) ? X[r][c] : (T) 0;
    }
}

template <typename T>
void potrf(hipStream_t s, int n, T *A, int lda, T *work, int *info) {
    hipMemsetAsync(info, 0, sizeof(int), s);
    for (int k0 = 0; k0 < n; k0 += NB) {
        int kb = (n - k0 < NB) ? n - k0 : NB;
        potf2_kernel<T><<<1, 256, 0, s>>>(A + k0 + (long) k0 * lda, lda, k0, kb, info, work);
        int k1 = k0 + kb;
        if (k1 < n) {
            T *P = A + k1 + (long) k0 * lda;
            // panel: P = P * Dinv^T (each workgroup owns its rows, so in place is safe)
            gemm<T>(s, false, true, false, n - k1, kb, kb, (T) 1, P, lda, 0, work, NB, 0, (T) 0, P, lda, 0, 1);
            // trailing: A22 -= P P^T (lower)
            gemm<T>(s, false, true, true, n - k1, n - k1, kb, (T) -1, P, lda, 0, P, lda, 0, (T) 1,
                    A + k1 + (long) k1 * lda, lda, 0, 1);
        }
    }
}

template void potrf<double>(hipStream_t, int, double *, int, double *, int *);
template void potrf<float>(hipStream_t, int, float *, int, float *, int *);

// ---- triangular inverse ------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void trtri_diag_kernel(int n, const T *L, int ldl, T *Linv, int ldi) {
    __shared__ T Ls[NB][NB + 1];
    __shared__ T X[NB][NB + 1];
    const int o = blockIdx.x * NB;
    const int kb = (n - o < NB) ? n - o : NB;
    const int tid = threadIdx.x;
    for (int e = tid; e < kb * kb; e += 256) {
        int r = e % kb, c = e / kb;
        Ls[r][c] = (r >= c) ? L[(o + r) + (long) (o + c) * ldl] : (T) 0;
    }
    __syncthreads();
    if (tid < kb) {
        int c = tid;
        X[c][c] = (T) 1 / Ls[c][c];
        for (int r = c + 1; r < kb; ++r) {
            T sum = 0;
            for (int l = c; l < r; ++l) sum += Ls[r][l] * X[l][c];
            X[r][c] = -sum / Ls[r][r];
        }
    }
    __syncthreads();
    for (int e = tid; e < kb * kb; e += 256) {
        int r = e % kb, c = e / kb;
        Linv[(o + r) + (long) (o + c) * ldi] = (r >= c) ? X[r][c] : (T) 0;
    }
}

template <typename T>
__global__ void fill_matrix_kernel(int n, T *A, int lda, T v) {
    long total = (long) n * n;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < total; e += (long) gridDim.x * blockDim.x) {
        long i = e % n, j = e / n;
        A[i + j * lda] = v;
    }
}

// inv([[A, 0], [B, C]]) = [[A^-1, 0], [-C^-1 B A^-1, C^-1]], applied level by level
// with every pair of one level batched into two GEMMs.
template <typename T>
void trtri(hipStream_t s, int n, const T *L, int ldl, T *Linv, int ldi, T *work) {
    fill_matrix_kernel<T><<<512, 256, 0, s>>>(n, Linv, ldi, (T) 0);
    int nblk = (n + NB - 1) / NB;
    trtri_diag_kernel<T><<<nblk, 256, 0, s>>>(n, L, ldl, Linv, ldi);
    for (int sz = NB; sz < n; sz *= 2) {
        int npair = n / (2 * sz);
        long strideL = 2L * sz * (ldl + 1), strideI = 2L * sz * (ldi + 1), strideW = (long) sz * sz;
        if (npair > 0) {
            gemm<T>(s, false, false, false, sz, sz, sz, (T) 1, L + sz, ldl, strideL, Linv, ldi, strideI, (T) 0,
                    work, sz, strideW, npair);
            gemm<T>(s, false, false, false, sz, sz, sz, (T) -1, Linv + (long) sz * (ldi + 1), ldi, strideI, work,
                    sz, strideW, (T) 0, Linv + sz, ldi, strideI, npair);
        }
        int o = npair * 2 * sz, rem = n - o;
        if (rem > sz) {
            int sc = rem - sz;
            gemm<T>(s, false, false, false, sc, sz, sz, (T) 1, L + (o + sz) + (long) o * ldl, ldl, 0,
                    Linv + o + (long) o * ldi, ldi, 0, (T) 0, work, sc, 0, 1);
            gemm<T>(s, false, false, false, sc, sz, sc, (T) -1, Linv + (long) (o + sz) * (ldi + 1), ldi, 0, work,
                    sc, 0, (T) 0, Linv + (o + sz) + (long) o * ldi, ldi, 0, 1);
        }
    }
}

template void trtri<double>(hipStream_t, int, const double *, int, double *, int, double *);
template void trtri<float>(hipStream_t, int, const float *, int, float *, int, float *);

// ---- matrix-vector ------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void coldot_kernel(int m, int ncols, int tri, T alpha, const T *__restrict__ A,
                                                     int lda, const T *__restrict__ x, T beta, T *y) {
    const int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (j >= ncols) return;
    const int lo = (tri == 1) ? j : 0;
    const int hi = (tri == 2) ? j + 1 : m;
    const T *col = A + (long) j * lda;
    T sum = 0;
    for (int i = lo + lane; i < hi; i += 64) sum += col[i] * x[i];
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off);
    if (lane == 0) y[j] = alpha * sum + ((beta != (T) 0) ? beta * y[j] : (T) 0);
}

template <typename T>
void coldot(hipStream_t s, int m, int ncols, int tri, T alpha, const T *A, int lda, const T *x, T beta, T *y) {
    if (ncols <= 0) return;
    int blocks = (ncols + 3) / 4;
    coldot_kernel<T><<<blocks, 256, 0, s>>>(m, ncols, tri, alpha, A, lda, x, beta, y);
}

template void coldot<double>(hipStream_t, int, int, int, double, const double *, int, const double *, double,
                             double *);
template void coldot<float>(hipStream_t, int, int, int, float, const float *, int, const float *, float, float *);

// ---- layout helpers --------------------------------------------------------------

template <typename T>
__global__ void mirror_kernel(int n, T *A, int lda) {
    __shared__ T tile[32][33];
    const int bi = blockIdx.x, bj = blockIdx.y;
    if (bi < bj) return;
    const int i0 = bi * 32, j0 = bj * 32, tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < 32; r += 8) {
        int i = i0 + tx, j = j0 + r;
        tile[r][tx] = (i < n && j < n) ? A[i + (long) j * lda] : (T) 0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int i = i0 + r, j = j0 + tx;  // source (i, j), destination (j, i)
        if (i < n && j < n && i > j) A[j + (long) i * lda] = tile[tx][r];
    }
}

template <typename T>
void mirror_lower(hipStream_t s, int n, T *A, int lda) {
    int nb = (n + 31) / 32;
    mirror_kernel<T><<<dim3(nb, nb), dim3(32, 8), 0, s>>>(n, A, lda);
}

template void mirror_lower<double>(hipStream_t, int, double *, int);
template void mirror_lower<float>(hipStream_t, int, float *, int);

template <typename T>
__global__ void transpose_kernel(int n, const T *A, int lda, T *B, int ldb) {
    __shared__ T tile[32][33];
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32, tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < 32; r += 8) {
        int i = i0 + tx, j = j0 + r;
        tile[r][tx] = (i < n && j < n) ? A[i + (long) j * lda] : (T) 0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int i = i0 + r, j = j0 + tx;
        if (i < n && j < n) B[j + (long) i * ldb] = tile[tx][r];
    }
}

template <typename T>
void transpose(hipStream_t s, int n, const T *A, int lda, T *B, int ldb) {
    int nb = (n + 31) / 32;
    transpose_kernel<T><<<dim3(nb, nb), dim3(32, 8), 0, s>>>(n, A, lda, B, ldb);
}

template void transpose<double>(hipStream_t, int, const double *, int, double *, int);
template void transpose<float>(hipStream_t, int, const float *, int, float *, int);

template <typename T>
__global__ void zero_upper_kernel(int n, T *A, int lda) {
    long total = (long) n * n;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < total; e += (long) gridDim.x * blockDim.x) {
        long i = e % n, j = e / n;
        if (i < j) A[i + j * lda] = 0;
    }
}

template <typename T>
void zero_upper(hipStream_t s, int n, T *A, int lda) {
    zero_upper_kernel<T><<<1024, 256, 0, s>>>(n, A, lda);
}

template void zero_upper<double>(hipStream_t, int, double *, int);
template void zero_upper<float>(hipStream_t, int, float *, int);

__global__ void fill_kernel(double *x, long n, double v) {
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < n; e += (long) gridDim.x * blockDim.x) x[e] = v;
}

void fill(hipStream_t s, double *x, long n, double v) {
    if (n <= 0) return;
    long blocks = (n + 255) / 256;
    fill_kernel<<<blocks > 4096 ? 4096 : blocks, 256, 0, s>>>(x, n, v);
}

__global__ void fill_f_kernel(float *x, long n, float v) {
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < n; e += (long) gridDim.x * blockDim.x) x[e] = v;
}

void fill_f(hipStream_t s, float *x, long n, float v) {
    if (n <= 0) return;
    long blocks = (n + 255) / 256;
    fill_f_kernel<<<blocks > 4096 ? 4096 : blocks, 256, 0, s>>>(x, n, v);
}

__global__ void copy_lower_kernel(int n, const double *A, int lda, double *B, int ldb) {
    long total = (long) n * n;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < total; e += (long) gridDim.x * blockDim.x) {
        long i = e % n, j = e / n;
        if (i >= j) B[i + j * ldb] = A[i + j * lda];
    }
}

void copy_lower(hipStream_t s, int n, const double *A, int lda, double *B, int ldb) {
    copy_lower_kernel<<<2048, 256, 0, s>>>(n, A, lda, B, ldb);
}

__global__ void convert_lower_kernel(int n, const double *A, int lda, float *B, int ldb) {
    long total = (long) n * n;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < total; e += (long) gridDim.x * blockDim.x) {
        long i = e % n, j = e / n;
        B[i + j * ldb] = (i >= j) ? (float) A[i + j * lda] : 0.0f;
    }
}

void convert_lower_to_float(hipStream_t s, int n, const double *A, int lda, float *B, int ldb) {
    convert_lower_kernel<<<2048, 256, 0, s>>>(n, A, lda, B, ldb);
}

__global__ void add_diag_kernel(int n, double *A, int lda, const double *d, double scalar) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[i + (long) i * lda] += (d ? d[i] : 0.0) + scalar;
}

void add_diag(hipStream_t s, int n, double *A, int lda, const double *d, double scalar) {
    add_diag_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, A, lda, d, scalar);
}

__global__ void logdet_kernel(int n, const double *A, int lda, double *out) {
    __shared__ double red[256];
    double sum = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) sum += log(A[i + (long) i * lda]);
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int) threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = 2.0 * red[0];
}

void logdet_diag(hipStream_t s, int n, const double *A, int lda, double *out) {
    logdet_kernel<<<1, 256, 0, s>>>(n, A, lda, out);
}

}  // namespace hdk
