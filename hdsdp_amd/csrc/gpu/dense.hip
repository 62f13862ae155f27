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

// k range of one output tile when an operand is triangular: tri 1 = op(B) lower
// (B(k, j) = 0 for k < j), tri 2 = op(A) lower (A(i, k) = 0 for k > i).  Only exact
// zeros are skipped, so the result is the same as the full product.
__device__ inline void tri_krange(int tri, int i0, int j0, int tile, int K, int &kLo, int &kHi) {
    kLo = 0;
    kHi = K;
    if (tri == 1) kLo = (j0 / KSTEP) * KSTEP;
    if (tri == 2) kHi = (i0 + tile < K) ? i0 + tile : K;
}

// Output tile of this workgroup.  With lin != 0 (lower products with as many tile rows
// as tile columns) blockIdx.x runs over the lower tiles only, t = x (x + 1) / 2 + y with
// y <= x: consecutive workgroups go to consecutive XCDs, so every XCD gets the same
// share of the triangle (a 2-D grid gives tile column x to XCD x mod 8, and the
// columns near the diagonal end hold unequal work).
//
// With tri == 2 (op(A) lower: the work of a tile grows with its row) the tile row of
// a 2-D grid is rotated by the other grid coordinates: workgroup x + gridDim.x (y + ...)
// goes to XCD (x + ...) mod 8, so with 8k tile rows XCD x would receive every tile of
// row x (the heaviest rows all on one XCD); the rotation gives each XCD every row.
__device__ inline void tile_xy(int lin, int &bx, int &by, int tri = 0) {
    if (!lin) {
        bx = blockIdx.x;
        by = blockIdx.y;
        if (tri == 2) bx = (int) ((blockIdx.x + blockIdx.y + blockIdx.z) % gridDim.x);
        return;
    }
    const int t = blockIdx.x;
    int x = (int) ((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((x + 1) * (x + 2) / 2 <= t) ++x;
    while (x * (x + 1) / 2 > t) --x;
    bx = x;
    by = t - x * (x + 1) / 2;
}

// One k slab of a 64-wide tile into registers: thread t moves elements t + 256 u of
// each operand's 64 x 16 slab (zero outside the matrix).
template <typename T, bool TA, bool TB>
__device__ inline void tile_fetch(const T *A, int lda, const T *B, int ldb, int M, int N, int K, int i0, int j0,
                                  int k0, int tid, T (&ra)[4], T (&rb)[4]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = tid + 256 * u;
        int i, k;
        if (!TA) { i = e & 63; k = e >> 6; } else { k = e & 15; i = e >> 4; }
        const int gi = i0 + i, gk = k0 + k;
        T v = 0;
        if (gi < M && gk < K) v = TA ? A[gk + (long) gi * lda] : A[gi + (long) gk * lda];
        ra[u] = v;
        int j, kb;
        if (!TB) { kb = e & 15; j = e >> 4; } else { j = e & 63; kb = e >> 6; }
        const int gj = j0 + j, gkb = k0 + kb;
        T w = 0;
        if (gj < N && gkb < K) w = TB ? B[gj + (long) gkb * ldb] : B[gkb + (long) gj * ldb];
        rb[u] = w;
    }
}

template <typename T, bool TA, bool TB>
__device__ inline void tile_stash(T (*As)[TILE + 1], T (*Bs)[TILE + 1], int tid, const T (&ra)[4],
                                  const T (&rb)[4]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = tid + 256 * u;
        int i, k;
        if (!TA) { i = e & 63; k = e >> 6; } else { k = e & 15; i = e >> 4; }
        As[k][i] = ra[u];
        int j, kb;
        if (!TB) { kb = e & 15; j = e >> 4; } else { j = e & 63; kb = e >> 6; }
        Bs[kb][j] = rb[u];
    }
}

// 64 x 64 tiles: the C tile (beta != 0) is fetched before the k loop and the k slabs
// are double-buffered (slab s + 1 loads in flight while slab s is multiplied), so a
// short-k product such as a rank-64 update does not wait on memory once per slab and
// again for C.
template <typename T, bool TA, bool TB, bool LOWER>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, T alpha, const T *A, int lda,
                                                   long sA, const T *B, int ldb, long sB, T beta,
                                                   T *C, int ldc, long sC, int tri, int lin) {
    typedef typename Mfma<T>::acc_t acc_t;
    int bx, by;
    tile_xy(lin, bx, by, tri);
    const int i0 = bx * TILE, j0 = by * TILE;
    if (LOWER && i0 + TILE - 1 < j0) return;
    int kLo, kHi;
    tri_krange(tri, i0, j0, TILE, K, kLo, kHi);
    A += (long) blockIdx.z * sA;
    B += (long) blockIdx.z * sB;
    C += (long) blockIdx.z * sC;
    __shared__ T As[2][KSTEP][TILE + 1];
    __shared__ T Bs[2][KSTEP][TILE + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    T cin[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wi + a * 16 + Mfma<T>::row(lk, r);
                const int j = j0 + wj + b * 16 + lr;
                const bool ok = i < M && j < N && (!LOWER || i >= j);
                cin[a][b][r] = (ok && beta != (T) 0) ? C[i + (long) j * ldc] : (T) 0;
            }
    acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = (acc_t){0, 0, 0, 0};
    const int nslab = kHi > kLo ? (kHi - kLo + KSTEP - 1) / KSTEP : 0;
    if (nslab > 0) {
        T ra[4], rb[4];
        tile_fetch<T, TA, TB>(A, lda, B, ldb, M, N, K, i0, j0, kLo, tid, ra, rb);
        tile_stash<T, TA, TB>(As[0], Bs[0], tid, ra, rb);
        __syncthreads();
        for (int sl = 0; sl < nslab; ++sl) {
            const int cur = sl & 1;
            const bool more = sl + 1 < nslab;
            if (more) tile_fetch<T, TA, TB>(A, lda, B, ldb, M, N, K, i0, j0, kLo + (sl + 1) * KSTEP, tid, ra, rb);
#pragma unroll
            for (int kk = 0; kk < KSTEP; kk += 4) {
                T a0 = As[cur][kk + lk][wi + lr], a1 = As[cur][kk + lk][wi + 16 + lr];
                T b0 = Bs[cur][kk + lk][wj + lr], b1 = Bs[cur][kk + lk][wj + 16 + lr];
                acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
                acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
                acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
                acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
            }
            if (more) tile_stash<T, TA, TB>(As[cur ^ 1], Bs[cur ^ 1], tid, ra, rb);
            __syncthreads();
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wi + a * 16 + Mfma<T>::row(lk, r);
                const int j = j0 + wj + b * 16 + lr;
                if (i < M && j < N && (!LOWER || i >= j)) C[i + (long) j * ldc] = alpha * acc[a][b][r] + beta * cin[a][b][r];
            }
}

// Two successive symmetric rank-K updates of a 64 x 64 tile grid,
//   C = alpha A1 A1(rows j)^T + beta C,  then  C = alpha A2 A2(rows j)^T + beta C,
// with C read and written once.  Each product runs the k slabs and MFMAs of
// gemm_kernel (TA = false, TB = true, B = A) and each update rounds as gemm_kernel's
// epilogue, so the result is that of two gemm_kernel launches with half the traffic
// on C.  lin as in gemm_kernel (lower products with a square tile grid).
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void gemm_nt2_kernel(int M, int N, int K, T alpha, const T *A1, const T *A2, int lda,
                                                       T beta, T *C, int ldc, int lin) {
    typedef typename Mfma<T>::acc_t acc_t;
    int bx, by;
    tile_xy(lin, bx, by, 0);
    const int i0 = bx * TILE, j0 = by * TILE;
    if (LOWER && i0 + TILE - 1 < j0) return;
    __shared__ T As[2][KSTEP][TILE + 1];
    __shared__ T Bs[2][KSTEP][TILE + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    T cin[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wi + a * 16 + Mfma<T>::row(lk, r);
                const int j = j0 + wj + b * 16 + lr;
                const bool ok = i < M && j < N && (!LOWER || i >= j);
                cin[a][b][r] = (ok && beta != (T) 0) ? C[i + (long) j * ldc] : (T) 0;
            }
    const int nslab = K > 0 ? (K + KSTEP - 1) / KSTEP : 0;
    acc_t acc[2][2];
    auto product = [&](const T *A) {
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) acc[a][b] = (acc_t){0, 0, 0, 0};
        if (nslab == 0) return;
        T ra[4], rb[4];
        tile_fetch<T, false, true>(A, lda, A, lda, M, N, K, i0, j0, 0, tid, ra, rb);
        tile_stash<T, false, true>(As[0], Bs[0], tid, ra, rb);
        __syncthreads();
        for (int sl = 0; sl < nslab; ++sl) {
            const int cur = sl & 1;
            const bool more = sl + 1 < nslab;
            if (more) tile_fetch<T, false, true>(A, lda, A, lda, M, N, K, i0, j0, (sl + 1) * KSTEP, tid, ra, rb);
#pragma unroll
            for (int kk = 0; kk < KSTEP; kk += 4) {
                T a0 = As[cur][kk + lk][wi + lr], a1 = As[cur][kk + lk][wi + 16 + lr];
                T b0 = Bs[cur][kk + lk][wj + lr], b1 = Bs[cur][kk + lk][wj + 16 + lr];
                acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
                acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
                acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
                acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
            }
            if (more) tile_stash<T, false, true>(As[cur ^ 1], Bs[cur ^ 1], tid, ra, rb);
            __syncthreads();
        }
    };
    product(A1);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) cin[a][b][r] = alpha * acc[a][b][r] + beta * cin[a][b][r];
    product(A2);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wi + a * 16 + Mfma<T>::row(lk, r);
                const int j = j0 + wj + b * 16 + lr;
                if (i < M && j < N && (!LOWER || i >= j)) C[i + (long) j * ldc] = alpha * acc[a][b][r] + beta * cin[a][b][r];
            }
}

template <typename T>
static void gemm_nt2(hipStream_t s, int M, int N, int K, T alpha, const T *A1, const T *A2, int lda, T beta, T *C,
                     int ldc) {
    if (M <= 0 || N <= 0) return;
    const int mt = (M + TILE - 1) / TILE, nt = (N + TILE - 1) / TILE;
    const int lin = (mt == nt) ? 1 : 0;
    dim3 grid(lin ? mt * (mt + 1) / 2 : mt, lin ? 1 : nt, 1);
    gemm_nt2_kernel<T, true><<<grid, 256, 0, s>>>(M, N, K, alpha, A1, A2, lda, beta, C, ldc, lin);
}

// 128 x 128 tiles for large products: four waves, each a 64 x 64 quadrant as 4 x 4
// MFMA 16x16 accumulators; k slabs of 16 double-buffered in LDS so the global
// loads of slab s + 1 are in flight while slab s is multiplied (one barrier per slab).
constexpr int BT = 128, BPAD = 16;
static_assert(4 * 16 * 65 <= 2 * KSTEP * (BT + BPAD), "epilogue staging must fit in the k-slab buffers");

template <typename T, bool TA, bool TB, bool CHECK>
__device__ inline void big_load(const T *A, int lda, const T *B, int ldb, int M, int N, int K, int i0, int j0,
                                int k0, int tid, T (&ra)[8], T (&rb)[8]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        int i, k;
        if (!TA) { k = tid >> 4; i = (tid & 15) * 8 + u; } else { i = tid >> 1; k = (tid & 1) * 8 + u; }
        const int gi = i0 + i, gk = k0 + k;
        const T *pa = TA ? A + gk + (long) gi * lda : A + gi + (long) gk * lda;
        ra[u] = (!CHECK || (gi < M && gk < K)) ? *pa : (T) 0;
        int j, kb;
        if (TB) { kb = tid >> 4; j = (tid & 15) * 8 + u; } else { j = tid >> 1; kb = (tid & 1) * 8 + u; }
        const int gj = j0 + j, gkb = k0 + kb;
        const T *pb = TB ? B + gj + (long) gkb * ldb : B + gkb + (long) gj * ldb;
        rb[u] = (!CHECK || (gj < N && gkb < K)) ? *pb : (T) 0;
    }
}

// Interior slab with 16-byte loads (FP64; leading dimensions, strides and base
// addresses even, checked by the launcher).  Each thread moves 8 elements of each
// operand as 4 pairs:
//   k-contiguous operand (A with TA, B without TB): row tid / 2, k = 8 (tid % 2) + 2u
//   row-contiguous operand: k = tid / 16, rows 2 (tid % 16) + 32u, so that the 16
//   lanes of one k read 256 consecutive bytes per instruction.
typedef double dpair __attribute__((ext_vector_type(2)));

template <bool TA, bool TB>
__device__ inline void big_load_vec(const double *A, int lda, const double *B, int ldb, int i0, int j0, int k0,
                                    int tid, double (&ra)[8], double (&rb)[8]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        dpair va, vb;
        if (TA) {
            const int i = tid >> 1, k = (tid & 1) * 8 + 2 * u;
            va = *reinterpret_cast<const dpair *>(A + (k0 + k) + (long) (i0 + i) * lda);
        } else {
            const int k = tid >> 4, i = 2 * (tid & 15) + 32 * u;
            va = *reinterpret_cast<const dpair *>(A + (i0 + i) + (long) (k0 + k) * lda);
        }
        if (!TB) {
            const int j = tid >> 1, k = (tid & 1) * 8 + 2 * u;
            vb = *reinterpret_cast<const dpair *>(B + (k0 + k) + (long) (j0 + j) * ldb);
        } else {
            const int k = tid >> 4, j = 2 * (tid & 15) + 32 * u;
            vb = *reinterpret_cast<const dpair *>(B + (j0 + j) + (long) (k0 + k) * ldb);
        }
        ra[2 * u] = va.x;
        ra[2 * u + 1] = va.y;
        rb[2 * u] = vb.x;
        rb[2 * u + 1] = vb.y;
    }
}

template <bool TA, bool TB>
__device__ inline void big_store_vec(double (*As)[BT + BPAD], double (*Bs)[BT + BPAD], int tid, const double (&ra)[8],
                                     const double (&rb)[8]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (TA) {
            const int i = tid >> 1, k = (tid & 1) * 8 + 2 * u;
            As[k][i] = ra[2 * u];
            As[k + 1][i] = ra[2 * u + 1];
        } else {
            const int k = tid >> 4, i = 2 * (tid & 15) + 32 * u;
            *reinterpret_cast<dpair *>(&As[k][i]) = (dpair){ra[2 * u], ra[2 * u + 1]};
        }
        if (!TB) {
            const int j = tid >> 1, k = (tid & 1) * 8 + 2 * u;
            Bs[k][j] = rb[2 * u];
            Bs[k + 1][j] = rb[2 * u + 1];
        } else {
            const int k = tid >> 4, j = 2 * (tid & 15) + 32 * u;
            *reinterpret_cast<dpair *>(&Bs[k][j]) = (dpair){rb[2 * u], rb[2 * u + 1]};
        }
    }
}

// one k slab of the tile into registers (returns whether the vector layout was used);
// the bounds checks are dropped for interior slabs
template <typename T, bool TA, bool TB, bool VEC>
__device__ inline bool big_fetch(const T *A, int lda, const T *B, int ldb, int M, int N, int K, int i0, int j0,
                                 int k0, int tid, T (&ra)[8], T (&rb)[8]) {
    if (i0 + BT <= M && j0 + BT <= N && k0 + KSTEP <= K) {
        if constexpr (VEC) {
            big_load_vec<TA, TB>(A, lda, B, ldb, i0, j0, k0, tid, ra, rb);
            return true;
        }
        big_load<T, TA, TB, false>(A, lda, B, ldb, M, N, K, i0, j0, k0, tid, ra, rb);
    } else {
        big_load<T, TA, TB, true>(A, lda, B, ldb, M, N, K, i0, j0, k0, tid, ra, rb);
    }
    return false;
}

template <typename T, bool TA, bool TB>
__device__ inline void big_store(T (*As)[BT + BPAD], T (*Bs)[BT + BPAD], int tid, const T (&ra)[8],
                                 const T (&rb)[8]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        int i, k;
        if (!TA) { k = tid >> 4; i = (tid & 15) * 8 + u; } else { i = tid >> 1; k = (tid & 1) * 8 + u; }
        As[k][i] = ra[u];
        int j, kb;
        if (TB) { kb = tid >> 4; j = (tid & 15) * 8 + u; } else { j = tid >> 1; kb = (tid & 1) * 8 + u; }
        Bs[kb][j] = rb[u];
    }
}

template <typename T, bool TA, bool TB, bool LOWER, bool VEC>
__global__ __launch_bounds__(256) void gemm_big_kernel(int M, int N, int K, T alpha, const T *A, int lda, long sA,
                                                       const T *B, int ldb, long sB, T beta, T *C, int ldc,
                                                       long sC, int tri, int lin, int kLast) {
    typedef typename Mfma<T>::acc_t acc_t;
    if (kLast > 0 && blockIdx.z == gridDim.z - 1) K = kLast;  // ragged last k part of a split product
    int bx, by;
    tile_xy(lin, bx, by, tri);
    const int i0 = bx * BT, j0 = by * BT;
    if (LOWER && i0 + BT - 1 < j0) return;
    int kLo, kHi;
    tri_krange(tri, i0, j0, BT, K, kLo, kHi);
    A += (long) blockIdx.z * sA;
    B += (long) blockIdx.z * sB;
    C += (long) blockIdx.z * sC;
    __shared__ __attribute__((aligned(16))) T As[2][KSTEP][BT + BPAD];
    __shared__ __attribute__((aligned(16))) T Bs[2][KSTEP][BT + BPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 64, wj = (wave >> 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;
    acc_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (acc_t){0, 0, 0, 0};
    T ra[8], rb[8];
    auto stash = [&](int buf, bool vec) {
        if constexpr (VEC) {
            if (vec) {
                big_store_vec<TA, TB>(As[buf], Bs[buf], tid, ra, rb);
                return;
            }
        }
        big_store<T, TA, TB>(As[buf], Bs[buf], tid, ra, rb);
    };
    bool vec = big_fetch<T, TA, TB, VEC>(A, lda, B, ldb, M, N, K, i0, j0, kLo, tid, ra, rb);
    stash(0, vec);
    __syncthreads();
    const int nslab = (kHi - kLo + KSTEP - 1) / KSTEP;
    for (int sl = 0; sl < nslab; ++sl) {
        const int cur = sl & 1;
        const bool more = sl + 1 < nslab;
        if (more) vec = big_fetch<T, TA, TB, VEC>(A, lda, B, ldb, M, N, K, i0, j0, kLo + (sl + 1) * KSTEP, tid, ra, rb);
#pragma unroll
        for (int kk = 0; kk < KSTEP; kk += 4) {
            T af[4], bf[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                af[f] = As[cur][kk + lk][wi + 16 * f + lr];
                bf[f] = Bs[cur][kk + lk][wj + 16 * f + lr];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = Mfma<T>::op(af[a], bf[b], acc[a][b]);
        }
        if (more) stash(cur ^ 1, vec);
        __syncthreads();
    }
    // Epilogue: each wave stages its 64 x 64 quadrant through LDS (the k-slab buffers,
    // free after the last barrier) 16 columns at a time, so that lane q reads and
    // writes row q of C: 64 consecutive rows per column, coalesced.
    T (*stage)[65] = reinterpret_cast<T (*)[65]>(&As[0][0][0]) + wave * 16;
    const int i = i0 + wi + lane;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) stage[lr][a * 16 + Mfma<T>::row(lk, r)] = acc[a][b][r];
        // all loads of the strip are issued before any store (C may alias nothing,
        // but the compiler cannot know that)
        T cin[16];
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
            const int j = j0 + wj + b * 16 + cc;
            const bool ok = i < M && j < N && (!LOWER || i >= j);
            cin[cc] = (ok && beta != (T) 0) ? C[i + (long) j * ldc] : (T) 0;
        }
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
            const int j = j0 + wj + b * 16 + cc;
            if (i < M && j < N && (!LOWER || i >= j)) C[i + (long) j * ldc] = alpha * stage[cc][lane] + beta * cin[cc];
        }
    }
}

// 128-wide tile launch; batch z multiplies K-wide slices at strides sA, sB (the last
// one kLast wide when kLast > 0)
template <typename T>
static void gemm_big(hipStream_t s, bool transA, bool transB, bool lower, int M, int N, int K, T alpha, const T *A,
                     int lda, long strideA, const T *B, int ldb, long strideB, T beta, T *C, int ldc, long strideC,
                     int batch, int tri, int kLast) {
    const int mt = (M + BT - 1) / BT, nt = (N + BT - 1) / BT;
    const int lin = (lower && mt == nt) ? 1 : 0;
    dim3 bgrid(lin ? mt * (mt + 1) / 2 : mt, lin ? 1 : nt, batch);
    // 16-byte loads need even leading dimensions, strides and addresses
    const bool vec = sizeof(T) == 8 && lda % 2 == 0 && ldb % 2 == 0 && strideA % 2 == 0 && strideB % 2 == 0 &&
                     ((uintptr_t) A % 16) == 0 && ((uintptr_t) B % 16) == 0;
#define HDK_GEMM_BIG(ta, tb, lo)                                                                                 \
    if (transA == ta && transB == tb && lower == lo) {                                                           \
        if (vec)                                                                                                 \
            gemm_big_kernel<T, ta, tb, lo, sizeof(T) == 8><<<bgrid, 256, 0, s>>>(                                \
                M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc, strideC, tri, lin, kLast);       \
        else                                                                                                     \
            gemm_big_kernel<T, ta, tb, lo, false><<<bgrid, 256, 0, s>>>(M, N, K, alpha, A, lda, strideA, B, ldb, \
                                                                        strideB, beta, C, ldc, strideC, tri,     \
                                                                        lin, kLast);                             \
        return;                                                                                                  \
    }
    HDK_GEMM_BIG(false, false, false)
    HDK_GEMM_BIG(false, true, false)
    HDK_GEMM_BIG(true, false, false)
    HDK_GEMM_BIG(true, true, false)
    HDK_GEMM_BIG(false, false, true)
    HDK_GEMM_BIG(false, true, true)
    HDK_GEMM_BIG(true, false, true)
    HDK_GEMM_BIG(true, true, true)
#undef HDK_GEMM_BIG
}

template <typename T>
void gemm(hipStream_t s, bool transA, bool transB, bool lower, int M, int N, int K, T alpha, const T *A,
          int lda, long strideA, const T *B, int ldb, long strideB, T beta, T *C, int ldc, long strideC,
          int batch, int tri) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    const long bigTiles = (long) ((M + BT - 1) / BT) * ((N + BT - 1) / BT) * batch / (lower ? 2 : 1);
    // the 128-wide tiles pay off only when the k loop is long enough to hide their
    // prologue and epilogue (measured: rank-64 updates run faster on 64-wide tiles)
    if (bigTiles >= 192 && K >= 256) {
        gemm_big<T>(s, transA, transB, lower, M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                    strideC, batch, tri, 0);
        return;
    }
    const int mt = (M + TILE - 1) / TILE, nt = (N + TILE - 1) / TILE;
    const int lin = (lower && mt == nt) ? 1 : 0;
    dim3 grid(lin ? mt * (mt + 1) / 2 : mt, lin ? 1 : nt, batch);
#define HDK_GEMM_CASE(ta, tb, lo)                                                                    \
    if (transA == ta && transB == tb && lower == lo) {                                               \
        gemm_kernel<T, ta, tb, lo><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, \
                                                        beta, C, ldc, strideC, tri, lin);           \
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
                           const double *, int, long, double, double *, int, long, int, int);
template void gemm<float>(hipStream_t, bool, bool, bool, int, int, int, float, const float *, int, long,
                          const float *, int, long, float, float *, int, long, int, int);

// ---- split-K ---------------------------------------------------------------------

// C(i, j) = alpha sum_p P_p(i, j) + beta C(i, j) over the P partial products (ld M),
// i >= j only when lower.  Partials are summed in index order.
template <typename T>
__global__ void splitk_reduce_kernel(int M, int N, int P, T alpha, const T *W, T beta, T *C, int ldc, int lower) {
    const long total = (long) M * N, mn = total;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < total; e += (long) gridDim.x * blockDim.x) {
        const int i = (int) (e % M), j = (int) (e / M);
        if (lower && i < j) continue;
        T acc = W[e];
        for (int p = 1; p < P; ++p) acc += W[e + p * mn];
        T *c = C + i + (long) j * ldc;
        *c = (beta == (T) 0) ? alpha * acc : alpha * acc + beta * *c;
    }
}

static int cu_count() {
    static int cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cache[dev] <= 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cache[dev] = v;
    }
    return cache[dev];
}

// Parts P (k ranges of at least 1024) minimising the modelled time: the 128-wide FP64
// kernel holds two workgroups per CU, so P parts of T output tiles take
// ceil(T P / (2 CUs)) rounds of a k range K / P each, plus the partial-sum traffic
// (written by the product, read by the reduction) per part.  Ties go to fewer parts.
int splitk_parts(int M, int N, int K, bool lower) {
    const long mt = (M + BT - 1) / BT, nt = (N + BT - 1) / BT;
    const long tiles = std::max<long>(1, lower ? (mt == nt ? mt * (mt + 1) / 2 : mt * nt / 2) : mt * nt);
    const long slots = 2L * cu_count();
    const double secPerK = 2.0 * BT * BT * (double) slots / 68e12;           // one round, one k
    const double secPerPart = (double) M * N * 8.0 * (lower ? 1.0 : 2.0) / 5e12;  // partials out and in
    int best = 1;
    double tBest = 1e300;
    for (int P = 1; P <= 32 && P <= K / 1024; ++P) {
        const long kc = ((K + P - 1) / P + KSTEP - 1) / KSTEP * KSTEP;
        const int parts = (int) ((K + kc - 1) / kc);
        const long rounds = (tiles * parts + slots - 1) / slots;
        const double t = rounds * (double) kc * secPerK + (parts > 1 ? parts * secPerPart : 0.0);
        if (t < 0.97 * tBest) {
            tBest = t;
            best = parts;
        }
    }
    return best;
}

template <typename T>
void gemm_splitk(hipStream_t s, bool transA, bool transB, bool lower, int M, int N, int K, T alpha, const T *A,
                 int lda, const T *B, int ldb, T beta, T *C, int ldc, T *work, size_t workLen) {
    if (M <= 0 || N <= 0) return;
    const size_t mn = (size_t) M * N;
    long P = splitk_parts(M, N, K, lower);
    if (work) P = std::min<long>(P, (long) (workLen / mn));
    if (!work || P <= 1) {
        gemm<T>(s, transA, transB, lower, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1, 0);
        return;
    }
    // parts of kc (a multiple of the k slab), the last one K - (parts - 1) kc, in one launch
    const int kc = (int) (((K + P - 1) / P + KSTEP - 1) / KSTEP * KSTEP);
    const int parts = (K + kc - 1) / kc;
    const int kLast = K - (parts - 1) * kc;
    const long sA = transA ? kc : (long) kc * lda, sB = transB ? (long) kc * ldb : kc;
    gemm_big<T>(s, transA, transB, lower, M, N, kc, (T) 1, A, lda, sA, B, ldb, sB, (T) 0, work, M, (long) mn, parts,
                0, kLast);
    splitk_reduce_kernel<T><<<4096, 256, 0, s>>>(M, N, parts, alpha, work, beta, C, ldc, lower ? 1 : 0);
}

template void gemm_splitk<double>(hipStream_t, bool, bool, bool, int, int, int, double, const double *, int,
                                  const double *, int, double, double *, int, double *, size_t);

// ---- Cholesky ----------------------------------------------------------------

// Column r of L^-1 for the 64 x 64 lower factor held column-major in LDS,
// Ls[c][q] = L(q, c): forward substitution L x = e_r, every lane reading the same
// entry of column j (an LDS broadcast).
template <typename T>
__device__ inline void leaf_inverse(const T (&Ls)[NB][NB], int r, T (&x)[NB]) {
    // an opaque copy of r: the identity column is formed here rather than merged with
    // the callers' identical identity padding (which kept 64 values live and spilled)
    asm volatile("" : "+v"(r));
#pragma unroll
    for (int j = 0; j < NB; ++j) x[j] = (j == r) ? (T) 1 : (T) 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        x[j] = x[j] / Ls[j][j];
#pragma unroll
        for (int k = j + 1; k < NB; ++k) x[k] -= Ls[j][k] * x[j];
        // keeps the (read-only) LDS loads of later steps from being hoisted here,
        // which would hold thousands of values in registers and spill
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Eight waves factor a kb x kb diagonal block (kb <= 64) and invert its factor.
// Thread (w, r) holds row r, columns 8 w .. 8 w + 7, in registers; every column
// is published in LDS once it is final for the step that needs it:
//   step j: p = a(j, j); row r: t = a(r, j) / p; a(r, c) -= t a(c, j) for c > j;
//           column j is then scaled by 1 / sqrt(p), a(j, j) = sqrt(p).
// The wave owning column j + 1 updates that column first and publishes it before its
// other updates, and the scaling of each column waits until after the last step (no
// step reads a scaled column), so only one division, one FMA and one LDS round trip
// lie between consecutive steps.  Entries above the diagonal are scratch.
// The inverse X = L^-1 follows by forward substitution, lane r computing column r of
// X with wave w holding its rows 8 w .. 8 w + 7: after row j is published, the
// owner of row j + 1 finishes that row (x(j+1) -= L(j+1, j) x(j), then the division
// by L(j+1, j+1)) and publishes it, and every wave updates its other rows
// k > j by x(k) -= L(k, j) x(j).  Every entry sees the same operations in the same
// order as in a one-wave sweep.  Rows and columns beyond kb are the identity.  Writes
// L (lower) in place and X into Dinv (64 x 64, zeros above); sets *info = k0 + j + 1
// at the first non-positive pivot.  Eight waves of eight columns take 27 us for a
// 64 x 64 block in FP64 against 41 us for four waves of sixteen (each step between
// barriers is half as long, two waves per SIMD) and 28 us for sixteen waves of four.
constexpr int LEAF_W = 8, LEAF_CW = NB / LEAF_W, LEAF_THREADS = LEAF_W * 64;

template <typename T>
__global__ __launch_bounds__(LEAF_THREADS) void chol_leaf_kernel(T *A, int lda, int k0, int kb, int *info, T *Dinv,
                                                                 int ldd) {
    if (*info) return;
    __shared__ T Ls[NB][NB];      // Ls[c][q] = entry (q, c)
    __shared__ T Xs[NB][NB + 1];  // Xs[k][c] = X(k, c)
    const int r = threadIdx.x & 63, w = threadIdx.x >> 6, c0 = w * LEAF_CW;
    T a[LEAF_CW], sq[LEAF_CW];
#pragma unroll
    for (int i = 0; i < LEAF_CW; ++i) {
        const int c = c0 + i;
        T v = (r == c) ? (T) 1 : (T) 0;
        if (r < kb && c < kb && r >= c) v = A[r + (long) c * lda];
        a[i] = v;
        sq[i] = (T) 1;
    }
    if (w == 0) Ls[0][r] = a[0];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int wj = j / LEAF_CW, jj = j % LEAF_CW;
        const int jn = (j + 1) % LEAF_CW;  // the next column's index in its owner wave
        const bool last = j + 1 == NB;
        __syncthreads();  // column j is published
        const T p = Ls[j][j];
        if (!(p > (T) 0)) {
            if (threadIdx.x == 0) *info = k0 + j + 1;
            return;
        }
        if (w == wj) {
            const T t = a[jj] / p;
            if (jj + 1 < LEAF_CW) {  // the next column is this wave's: update and publish it first
                a[jj + 1] -= t * Ls[j][j + 1];
                Ls[j + 1][r] = a[jj + 1];
            }
#pragma unroll
            for (int i = jj + 2; i < LEAF_CW; ++i) a[i] -= t * Ls[j][c0 + i];
            sq[jj] = sqrt(p);
        } else if (w > wj) {
            const T t = Ls[j][r] / p;
            if (!last && jn == 0 && w == wj + 1) {  // this wave owns the next column
                a[0] -= t * Ls[j][c0];
                Ls[j + 1][r] = a[0];
#pragma unroll
                for (int i = 1; i < LEAF_CW; ++i) a[i] -= t * Ls[j][c0 + i];
            } else {
#pragma unroll
                for (int i = 0; i < LEAF_CW; ++i) a[i] -= t * Ls[j][c0 + i];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < LEAF_CW; ++i) a[i] = (r == c0 + i) ? sq[i] : a[i] / sq[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < LEAF_CW; ++i) {
        const int c = c0 + i;
        Ls[c][r] = a[i];
        if (r < kb && c < kb && r >= c) A[r + (long) c * lda] = a[i];
    }
    __syncthreads();
    // forward substitution: lane r builds column r of X, wave w its rows c0 .. c0 + 7
    T dg[LEAF_CW], sub[LEAF_CW];  // L(k, k) and L(k, k - 1) of this wave's rows k
#pragma unroll
    for (int i = 0; i < LEAF_CW; ++i) {
        dg[i] = Ls[c0 + i][c0 + i];
        sub[i] = (c0 + i > 0) ? Ls[c0 + i - 1][c0 + i] : (T) 0;
    }
    int rr = r;
    asm volatile("" : "+v"(rr));  // identity formed here, not merged with the padding above
    T x[LEAF_CW];
#pragma unroll
    for (int i = 0; i < LEAF_CW; ++i) x[i] = (c0 + i == rr) ? (T) 1 : (T) 0;
    if (w == 0) {
        x[0] = x[0] / dg[0];
        Xs[0][r] = x[0];
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int wj = j / LEAF_CW, jj = j % LEAF_CW;
        const int jn = (j + 1) % LEAF_CW;
        const bool last = j + 1 == NB;
        __syncthreads();  // row j of X is published
        if (w == wj) {
            const T xj = x[jj];
            if (jj + 1 < LEAF_CW) {  // finish and publish the next row first
                x[jj + 1] -= sub[jj + 1] * xj;
                x[jj + 1] = x[jj + 1] / dg[jj + 1];
                Xs[j + 1][r] = x[jj + 1];
            }
#pragma unroll
            for (int i = jj + 2; i < LEAF_CW; ++i) x[i] -= Ls[j][c0 + i] * xj;
        } else if (w > wj) {
            const T xj = Xs[j][r];
            if (!last && jn == 0 && w == wj + 1) {
                x[0] -= sub[0] * xj;
                x[0] = x[0] / dg[0];
                Xs[j + 1][r] = x[0];
#pragma unroll
                for (int i = 1; i < LEAF_CW; ++i) x[i] -= Ls[j][c0 + i] * xj;
            } else {
#pragma unroll
                for (int i = 0; i < LEAF_CW; ++i) x[i] -= Ls[j][c0 + i] * xj;
            }
        }
    }
    __syncthreads();
    // Dinv(q, c) = X(q, c) below the diagonal, stored row-coalesced: lane q writes row q
    for (int c = c0; c < c0 + LEAF_CW; ++c) Dinv[r + (long) c * ldd] = (r >= c) ? Xs[r][c] : (T) 0;
}

// Right-looking blocked Cholesky.  Per block column k (width NB):
//   leaf:  factor the diagonal block and invert its factor (work)
//   panel: P = P L11^-T as a GEMM with that inverse
//   trail: A22 -= P P^T (lower)
// With look-ahead the trailing update is split: the next block column's part runs on
// s (so that the next leaf and panel can start at once), the rest on s2; the next
// block column's own update waits for s2's previous update through an event.
template <typename T>
bool potrf(hipStream_t s, int n, T *A, int lda, T *work, size_t workLen, int *info, const LookAhead *la) {
    if (workLen < POTRF_WORK) {
        fprintf(stderr, "[hdsdp_amd] potrf: work of %zu elements, %zu needed\n", workLen, POTRF_WORK);
        return false;
    }
    hipMemsetAsync(info, 0, sizeof(int), s);
    auto at = [&](int i, int j) { return A + i + (long) j * lda; };
    auto leaf = [&](int k0) {
        const int kb = (n - k0 < NB) ? n - k0 : NB;
        chol_leaf_kernel<T><<<1, LEAF_THREADS, 0, s>>>(at(k0, k0), lda, k0, kb, info, work, NB);
        if (k0 + kb < n)
            gemm<T>(s, false, true, false, n - k0 - kb, kb, kb, (T) 1, at(k0 + kb, k0), lda, 0, work, NB, 0, (T) 0,
                    at(k0 + kb, k0), lda, 0, 1);
    };
    if (!la || n <= 4 * NB) {
        for (int k0 = 0; k0 < n; k0 += NB) {
            leaf(k0);
            const int k1 = k0 + NB;
            if (k1 < n)
                gemm<T>(s, false, true, true, n - k1, n - k1, NB, (T) -1, at(k1, k0), lda, 0, at(k1, k0), lda, 0,
                        (T) 1, at(k1, k1), lda, 0, 1);
        }
        return true;
    }
    // Block columns are taken in pairs (a, b = a + NB): the update of a's panel to the
    // blocks from a + 2 NB on waits for b's panel, and then both updates run as one pass
    // over the trailing matrix (gemm_nt2: the two updates in order, C read and written
    // once; the trailing updates are bound by that traffic).  The next block column is
    // always updated on s, ahead of its leaf; the rest on s2, overlapping the next leaves.
    int ring = 0;
    bool s2dirty = false;  // s2 holds an update that s has not yet waited for
    hipEvent_t evB = nullptr;
    int kDef = -1;  // first panel of an open pair
    leaf(0);
    for (int k0 = 0;; k0 += NB) {
        const int k1 = k0 + NB;
        if (k1 >= n) break;
        const int nb1 = (n - k1 < NB) ? n - k1 : NB, k2 = k1 + nb1;
        if (kDef < 0) {
            if (k2 < n) kDef = k0;  // this panel's update of blocks from k2 on waits for the next panel
            if (s2dirty) {
                hipStreamWaitEvent(s, evB, 0);  // the previous pair's update reached block k1
                s2dirty = false;
            }
            gemm<T>(s, false, true, true, n - k1, nb1, NB, (T) -1, at(k1, k0), lda, 0, at(k1, k0), lda, 0, (T) 1,
                    at(k1, k1), lda, 0, 1);
        } else {
            if (k2 < n) {
                hipEvent_t evP = la->ev[ring++ % LA_EVENTS];
                hipEventRecord(evP, s);
                hipStreamWaitEvent(la->s2, evP, 0);
                gemm_nt2<T>(la->s2, n - k2, n - k2, NB, (T) -1, at(k2, kDef), at(k2, k0), lda, (T) 1, at(k2, k2),
                            lda);
                evB = la->ev[ring++ % LA_EVENTS];
                hipEventRecord(evB, la->s2);
                s2dirty = true;
            }
            gemm_nt2<T>(s, n - k1, nb1, NB, (T) -1, at(k1, kDef), at(k1, k0), lda, (T) 1, at(k1, k1), lda);
            kDef = -1;
        }
        leaf(k1);
    }
    if (s2dirty) hipStreamWaitEvent(s, evB, 0);
    return true;
}

template bool potrf<double>(hipStream_t, int, double *, int, double *, size_t, int *, const LookAhead *);
template bool potrf<float>(hipStream_t, int, float *, int, float *, size_t, int *, const LookAhead *);

// ---- triangular inverse ------------------------------------------------------

// Inverse of each 64 x 64 diagonal block of the lower factor (the last one may be
// smaller, padded with the identity): one wave per block, lane q loads row q and
// stores it to LDS as column-major L for the broadcast reads of the substitution.
template <typename T>
__global__ __launch_bounds__(64) void trtri_diag_kernel(int n, const T *L, int ldl, T *Linv, int ldi) {
    const int o = blockIdx.x * NB;
    const int kb = (n - o < NB) ? n - o : NB;
    const int r = threadIdx.x;
    __shared__ T Ls[NB][NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        T v = (r == c) ? (T) 1 : (T) 0;
        if (r < kb && c < kb && r >= c) v = L[(o + r) + (long) (o + c) * ldl];
        Ls[c][r] = v;
    }
    __syncthreads();
    T x[NB];
    leaf_inverse<T>(Ls, r, x);
    __syncthreads();
    // column r of the inverse back into LDS, then stored row by row (lane q writes row q)
#pragma unroll
    for (int k = 0; k < NB; ++k) Ls[r][k] = (k >= r) ? x[k] : (T) 0;
    __syncthreads();
    for (int c = 0; c < kb; ++c)
        if (r < kb) Linv[(o + r) + (long) (o + c) * ldi] = Ls[c][r];
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
    trtri_diag_kernel<T><<<nblk, NB, 0, s>>>(n, L, ldl, Linv, ldi);
    for (int sz = NB; sz < n; sz *= 2) {
        int npair = n / (2 * sz);
        long strideL = 2L * sz * (ldl + 1), strideI = 2L * sz * (ldi + 1), strideW = (long) sz * sz;
        if (npair > 0) {
            // W = L21 Linv11 (Linv11 lower), then Linv21 = -Linv22 W (Linv22 lower)
            gemm<T>(s, false, false, false, sz, sz, sz, (T) 1, L + sz, ldl, strideL, Linv, ldi, strideI, (T) 0,
                    work, sz, strideW, npair, 1);
            gemm<T>(s, false, false, false, sz, sz, sz, (T) -1, Linv + (long) sz * (ldi + 1), ldi, strideI, work,
                    sz, strideW, (T) 0, Linv + sz, ldi, strideI, npair, 2);
        }
        int o = npair * 2 * sz, rem = n - o;
        if (rem > sz) {
            int sc = rem - sz;
            gemm<T>(s, false, false, false, sc, sz, sz, (T) 1, L + (o + sz) + (long) o * ldl, ldl, 0,
                    Linv + o + (long) o * ldi, ldi, 0, (T) 0, work, sc, 0, 1, 1);
            gemm<T>(s, false, false, false, sc, sz, sc, (T) -1, Linv + (long) (o + sz) * (ldi + 1), ldi, 0, work,
                    sc, 0, (T) 0, Linv + (o + sz) + (long) o * ldi, ldi, 0, 1, 2);
        }
    }
}

template void trtri<double>(hipStream_t, int, const double *, int, double *, int, double *);
template void trtri<float>(hipStream_t, int, const float *, int, float *, int, float *);

// ---- matrix-vector ------------------------------------------------------------

// Sums over one U-row group for NC columns: lane rows i, i + 64, ..., i + (U-1) 64,
// all 2U..(NC+1)U loads issued before the FMAs, each column's FMAs in row order.
template <typename T, int U, int NC>
__device__ inline void coldot_group(const T *const (&col)[NC], const T *__restrict__ x, int i, T (&sum)[NC]) {
    T a[NC][U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        b[u] = x[i + u * 64];
#pragma unroll
        for (int c = 0; c < NC; ++c) a[c][u] = col[c][i + u * 64];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int u = 0; u < U; ++u) sum[c] += a[c][u] * b[u];
}

// y_j = alpha <A(lo:hi, j), x(lo:hi)> + beta y_j with (lo, hi) = (0, m), (j, m) for tri 1
// and (0, j + 1) for tri 2.  One wave per NC adjacent columns (NC > 1 only for tri 0 and
// 2, where the columns share the x rows and one x load serves all of them).  Each lane
// accumulates rows lo + lane, lo + lane + 64, ... in order and the wave adds its lanes by a
// fixed shuffle tree, so the result does not depend on NC or on the unrolling.
template <typename T, int NC>
__global__ __launch_bounds__(256) void coldot_kernel(int m, int ncols, int tri, T alpha, const T *__restrict__ A,
                                                     int lda, const T *__restrict__ x, T beta, T *y) {
    const int j0 = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * NC;
    const int lane = threadIdx.x & 63;
    if (j0 >= ncols) return;
    const T *col[NC];
    int hic[NC];
    T sum[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int j = min(j0 + c, ncols - 1);  // a ragged last wave repeats the last column
        col[c] = A + (long) j * lda;
        hic[c] = (tri == 2) ? j + 1 : m;
        sum[c] = 0;
    }
    const int hmin = hic[0];  // hic is nondecreasing in c
    int i = ((NC == 1 && tri == 1) ? j0 : 0) + lane;
    for (; i + 7 * 64 < hmin; i += 8 * 64) coldot_group<T, 8, NC>(col, x, i, sum);
    if (i + 3 * 64 < hmin) {
        coldot_group<T, 4, NC>(col, x, i, sum);
        i += 4 * 64;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
        for (int k = i; k < hic[c]; k += 64) sum[c] += col[c][k] * x[k];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        for (int off = 32; off > 0; off >>= 1) sum[c] += __shfl_down(sum[c], off);
        if (lane == 0 && j0 + c < ncols) y[j0 + c] = alpha * sum[c] + ((beta != (T) 0) ? beta * y[j0 + c] : (T) 0);
    }
}

template <typename T, int NC>
static void coldot_launch(hipStream_t s, int m, int ncols, int tri, T alpha, const T *A, int lda, const T *x, T beta,
                          T *y) {
    const int waves = (ncols + NC - 1) / NC;
    coldot_kernel<T, NC><<<(waves + 3) / 4, 256, 0, s>>>(m, ncols, tri, alpha, A, lda, x, beta, y);
}

// Columns per wave: shared x loads pay once there are enough waves to fill the chip
// (measured on 20000^2 FP64: 4.4 TB/s at one column per wave, 5.8 TB/s at four).
template <typename T>
void coldot(hipStream_t s, int m, int ncols, int tri, T alpha, const T *A, int lda, const T *x, T beta, T *y) {
    if (ncols <= 0) return;
    if (tri != 1 && ncols >= 8192)
        coldot_launch<T, 4>(s, m, ncols, tri, alpha, A, lda, x, beta, y);
    else if (tri != 1 && ncols >= 1024)
        coldot_launch<T, 2>(s, m, ncols, tri, alpha, A, lda, x, beta, y);
    else
        coldot_launch<T, 1>(s, m, ncols, tri, alpha, A, lda, x, beta, y);
}

template void coldot<double>(hipStream_t, int, int, int, double, const double *, int, const double *, double,
                             double *);
template void coldot<float>(hipStream_t, int, int, int, float, const float *, int, const float *, float, float *);

// ---- symmetric product from the lower triangle -----------------------------------

// y = M x for symmetric M (column-major, ldm) reading only its lower triangle, in one
// pass.  One wave per tile of SYMV_R rows x SYMV_C columns on or below the diagonal;
// lane l owns rows r0 + l + 64 q (q < RPT = SYMV_R / 64).  For each column j of the tile the wave
// forms the column sum over its rows i >= j, sum_i M(i, j) x(i) (the (L^T x)(j)
// share, L the lower triangle), and reduces it over the lanes into P1[R][j]; each
// lane also accumulates the row sums over its columns j < i, sum_j M(i, j) x(j) (the
// strictly-lower (L x)(i) share), stored as P2[C][i].  A second kernel adds the
// partials in a fixed order, so the result is deterministic.  The triangle is read
// once: half the bytes of the column-dot product over the full matrix.  (Tile shape
// measured at m = 20000: 512 x 256 with 4 columns per pass, 0.36 ms; 512 x 128 with 2,
// 0.40 ms; wider or taller tiles leave too few waves.)
template <int RPT, int CP>
__global__ __launch_bounds__(256) void symv_lower_tile_kernel(int m, const double *__restrict__ M, long ldm,
                                                              const double *__restrict__ x, double *P1, double *P2) {
    constexpr int TR = 64 * RPT;  // rows per tile
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int R = blockIdx.x * 4 + w, C = blockIdx.y;
    const int r0 = R * TR, c0 = C * SYMV_C;
    if (r0 >= m || r0 + TR - 1 < c0) return;  // no entry of this tile on or below the diagonal
    double xi[RPT], acc2[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int i = r0 + lane + 64 * q;
        xi[q] = (i < m) ? x[i] : 0.0;
        acc2[q] = 0.0;
    }
    const int cEnd = min(min(c0 + SYMV_C, m), r0 + TR);
    // CP columns per pass, the next pass's loads issued before this pass's lane
    // reductions (software pipelining: the loop is otherwise load-latency bound)
    double v[CP][RPT];
    auto fetch = [&](int j) {
#pragma unroll
        for (int c = 0; c < CP; ++c) {
            const double *col = M + (long) (j + c) * ldm;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int i = r0 + lane + 64 * q;
                v[c][q] = (j + c < cEnd && i < m && i >= j + c) ? col[i] : 0.0;
            }
        }
    };
    fetch(c0);
    for (int j = c0; j < cEnd; j += CP) {
        double a[CP], xj[CP];
#pragma unroll
        for (int c = 0; c < CP; ++c) {
            xj[c] = (j + c < cEnd) ? x[j + c] : 0.0;
            a[c] = 0.0;
        }
#pragma unroll
        for (int c = 0; c < CP; ++c)
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int i = r0 + lane + 64 * q;
                a[c] += v[c][q] * xi[q];
                if (i > j + c) acc2[q] += v[c][q] * xj[c];
            }
        if (j + CP < cEnd) fetch(j + CP);
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int c = 0; c < CP; ++c) a[c] += __shfl_down(a[c], off);
        if (lane == 0)
#pragma unroll
            for (int c = 0; c < CP; ++c)
                if (j + c < cEnd) P1[(long) R * m + j + c] = a[c];
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int i = r0 + lane + 64 * q;
        if (i < m) P2[(long) C * m + i] = acc2[q];
    }
}

// y(i) = sum over the row chunks R >= i / SYMV_R of P1[R][i] + sum over the column
// blocks C <= (row chunk start of i + SYMV_R - 1) / SYMV_C of P2[C][i], in that order
__global__ void symv_lower_reduce_kernel(int m, int nR, int nC, const double *P1, const double *P2, double *y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double s = 0.0;
    for (int R = i / SYMV_R; R < nR; ++R) s += P1[(long) R * m + i];
    const int cLast = min(nC - 1, ((i / SYMV_R) * SYMV_R + SYMV_R - 1) / SYMV_C);
    for (int C = 0; C <= cLast; ++C) s += P2[(long) C * m + i];
    y[i] = s;
}

size_t symv_lower_work(int m) {
    const size_t nR = (m + SYMV_R - 1) / SYMV_R, nC = (m + SYMV_C - 1) / SYMV_C;
    return (nR + nC) * (size_t) m;
}

void symv_lower(hipStream_t s, int m, const double *M, long ldm, const double *x, double *y, double *work) {
    if (m <= 0) return;
    const int nR = (m + SYMV_R - 1) / SYMV_R, nC = (m + SYMV_C - 1) / SYMV_C;
    double *P1 = work, *P2 = work + (size_t) nR * m;
    symv_lower_tile_kernel<SYMV_R / 64, 4><<<dim3((nR + 3) / 4, nC), 256, 0, s>>>(m, M, ldm, x, P1, P2);
    symv_lower_reduce_kernel<<<(m + 255) / 256, 256, 0, s>>>(m, nR, nC, P1, P2, y);
}

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

__global__ void convert_lower_scaled_kernel(int n, const double *A, int lda, const double *sc, double shift, float *B,
                                            int ldb) {
    long total = (long) n * n;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < total; e += (long) gridDim.x * blockDim.x) {
        long i = e % n, j = e / n;
        float v = 0.0f;
        if (i >= j) v = (float) (A[i + j * lda] * sc[i] * sc[j] + ((i == j) ? shift : 0.0));
        B[i + j * ldb] = v;
    }
}

void convert_lower_scaled(hipStream_t s, int n, const double *A, int lda, const double *sc, double shift, float *B,
                          int ldb) {
    convert_lower_scaled_kernel<<<2048, 256, 0, s>>>(n, A, lda, sc, shift, B, ldb);
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
