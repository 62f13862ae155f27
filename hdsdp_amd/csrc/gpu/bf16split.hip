// bf16split.hip -- FP32-equivalent products on the bf16 matrix cores of gfx950.
//
// C (M x N) = alpha * X^T Y + beta * C for column-major X (K x M) and Y (K x N):
// both operands are contiguous along k, which is the natural layout of the Gram
// products of the Schur matrix (M = T^T T over vec(T_i), K = G^T G over the
// columns of G) and of the two-sided products L^-1 A L^-T written as X^T Y.
//
// Each FP32 operand x is split on its way into LDS into bf16 terms,
//   x = x0 + x1 (+ x2),  x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1),
// and the product is formed from the bf16 MFMA products whose weight is at least
// 2^-16 of the leading one:
//   NT = 2 (3 products): x0 y0 + x0 y1 + x1 y0, about 16 bits (relative ~1e-5);
//   NT = 3 (6 products): also x1 y1 + x0 y2 + x2 y0, about 24 bits (FP32-like).
// Every bf16 x bf16 product is exact in FP32.  v_mfma_f32_16x16x32_bf16 accumulates
// the leading product and the correction products in two separate FP32
// accumulators (the corrections are ~2^-8 of the leading sum and would lose their
// low bits against it), and every FLUSH k-slabs both are added into FP64
// accumulators, so long sums (K = n^2) do not build up FP32 rounding.
//
// Tiles: 128 x 128 per workgroup of four waves (2 x 2), each wave 4 x 4 MFMA
// 16 x 16 blocks; k slabs of 32 double-buffered in LDS (one plane per split term
// and operand, rows padded to 40 bf16 so a lane's 8 consecutive k are one aligned
// 16-byte read).  MFMA operand layout (measured on gfx950): A(i, k) for
// i = lane % 16 and k = 8 (lane / 16) + t in element t; B(k, j) likewise with
// j = lane % 16; D(i, j) at lane j + 16 (i / 4), element i % 4.
#include "dev_common.h"

#include <hip/hip_bf16.h>

namespace hdk {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int XT = 128;     // output tile
constexpr int XK = 32;      // k slab = one MFMA k
constexpr int XPAD = 8;     // row padding (bf16)
constexpr int FLUSH = 8;    // slabs between FP64 accumulations (k = 256)

// Load 16 consecutive k of row r (a column of X or Y) into registers.
template <bool CHECK>
__device__ inline void load16(const float *P, int ld, int nrows, int K, int row, int k0, float (&v)[16]) {
    const float *p = P + k0 + (long) row * ld;
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = (!CHECK || (row < nrows && k0 + u < K)) ? p[u] : 0.0f;
}

// Split 16 values into NT bf16 planes (plane t at pl[t] + offset).
template <int NT>
__device__ inline void store16(__bf16 *p0, __bf16 *p1, __bf16 *p2, const float (&v)[16]) {
    bf16x8 t0[2], t1[2], t2[2];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float x = v[8 * h + u];
            const __bf16 a = (__bf16) x;
            const float r1 = x - (float) a;
            const __bf16 b = (__bf16) r1;
            t0[h][u] = a;
            t1[h][u] = b;
            if (NT == 3) t2[h][u] = (__bf16) (r1 - (float) b);
        }
    *reinterpret_cast<bf16x8 *>(p0) = t0[0];
    *reinterpret_cast<bf16x8 *>(p0 + 8) = t0[1];
    *reinterpret_cast<bf16x8 *>(p1) = t1[0];
    *reinterpret_cast<bf16x8 *>(p1 + 8) = t1[1];
    if (NT == 3) {
        *reinterpret_cast<bf16x8 *>(p2) = t2[0];
        *reinterpret_cast<bf16x8 *>(p2 + 8) = t2[1];
    }
}

// tri: 1 when column j of Y is zero below row j (Y(k, j) = 0 for k > j), 2 when column i of
// X is (X(k, i) = 0 for k > i): the k range of a tile ends at its last column (row); only
// exact zeros are skipped.  With tri == 2 the tile row of a 2-D grid is rotated by the
// other grid coordinates, so that every XCD gets every row (workgroup x + gridDim.x
// (y + ...) goes to XCD (x + ...) mod 8, and the work of a tile grows with its row).
template <typename Tout, bool LOWER, int NT>
__global__ __launch_bounds__(256) void gemm_bf16s_tn_kernel(int M, int N, int K, double alpha, const float *X, int ldx,
                                                            long sX, const float *Y, int ldy, long sY, double beta,
                                                            Tout *C, int ldc, long sC, int lin, int tri) {
    // lin: blockIdx.x runs over the lower tiles only (t = x (x + 1) / 2 + y, y <= x), so
    // that consecutive workgroups, dealt to consecutive XCDs, share the triangle evenly
    int bx = blockIdx.x, by = blockIdx.y;
    if (!lin && tri == 2) bx = (int) ((blockIdx.x + blockIdx.y + blockIdx.z) % gridDim.x);
    if (lin) {
        const int t = blockIdx.x;
        int x = (int) ((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((x + 1) * (x + 2) / 2 <= t) ++x;
        while (x * (x + 1) / 2 > t) --x;
        bx = x;
        by = t - x * (x + 1) / 2;
    }
    const int i0 = bx * XT, j0 = by * XT;
    if (LOWER && i0 + XT - 1 < j0) return;
    const int Ke = (tri == 1) ? min(K, j0 + XT) : (tri == 2) ? min(K, i0 + XT) : K;
    X += (long) blockIdx.z * sX;
    Y += (long) blockIdx.z * sY;
    C += (long) blockIdx.z * sC;
    // planes [buffer][term][row][k]
    __shared__ __attribute__((aligned(16))) __bf16 As[2][NT][XT][XK + XPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][NT][XT][XK + XPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 64, wj = (wave >> 1) * 64;
    const int lr = lane & 15, lg = lane >> 4;
    // loader mapping: thread -> row tid / 2 of the tile, k half (tid & 1) * 16
    const int lrow = tid >> 1, lkh = (tid & 1) * 16;

    f32x4 acc[4][4], cor[4][4];
    double acc64[4][4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
            cor[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 4; ++r) acc64[a][b][r] = 0.0;
        }

    float vx[16], vy[16];
    auto fetch = [&](int k0) {
        const bool interior = (i0 + XT <= M) && (j0 + XT <= N) && (k0 + XK <= Ke);
        if (interior) {
            load16<false>(X, ldx, M, Ke, i0 + lrow, k0 + lkh, vx);
            load16<false>(Y, ldy, N, Ke, j0 + lrow, k0 + lkh, vy);
        } else {
            load16<true>(X, ldx, M, Ke, i0 + lrow, k0 + lkh, vx);
            load16<true>(Y, ldy, N, Ke, j0 + lrow, k0 + lkh, vy);
        }
    };
    auto stash = [&](int buf) {
        store16<NT>(&As[buf][0][lrow][lkh], &As[buf][1][lrow][lkh], &As[buf][NT - 1][lrow][lkh], vx);
        store16<NT>(&Bs[buf][0][lrow][lkh], &Bs[buf][1][lrow][lkh], &Bs[buf][NT - 1][lrow][lkh], vy);
    };

    fetch(0);
    stash(0);
    __syncthreads();
    const int nslab = (Ke + XK - 1) / XK;
    for (int sl = 0; sl < nslab; ++sl) {
        const int cur = sl & 1;
        const bool more = sl + 1 < nslab;
        if (more) fetch((sl + 1) * XK);
        bf16x8 af[NT][4], bf[NT][4];
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                af[t][f] = *reinterpret_cast<const bf16x8 *>(&As[cur][t][wi + 16 * f + lr][8 * lg]);
                bf[t][f] = *reinterpret_cast<const bf16x8 *>(&Bs[cur][t][wj + 16 * f + lr][8 * lg]);
            }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][a], bf[0][b], acc[a][b], 0, 0, 0);
                // correction terms, smallest first
                if (NT == 3) {
                    cor[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[2][a], bf[0][b], cor[a][b], 0, 0, 0);
                    cor[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][a], bf[2][b], cor[a][b], 0, 0, 0);
                    cor[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1][a], bf[1][b], cor[a][b], 0, 0, 0);
                }
                cor[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1][a], bf[0][b], cor[a][b], 0, 0, 0);
                cor[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][a], bf[1][b], cor[a][b], 0, 0, 0);
            }
        if ((sl + 1) % FLUSH == 0 || !more) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc64[a][b][r] += (double) acc[a][b][r] + (double) cor[a][b][r];
                    acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
                    cor[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
                }
        }
        if (more) stash(cur ^ 1);
        __syncthreads();
    }
    // epilogue: loads of C (beta != 0) issued before any store
    Tout cin[4][4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wi + 16 * a + 4 * lg + r, j = j0 + wj + 16 * b + lr;
                const bool ok = i < M && j < N && (!LOWER || i >= j);
                cin[a][b][r] = (ok && beta != 0.0) ? C[i + (long) j * ldc] : (Tout) 0;
            }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wi + 16 * a + 4 * lg + r, j = j0 + wj + 16 * b + lr;
                if (i < M && j < N && (!LOWER || i >= j))
                    C[i + (long) j * ldc] = (Tout) (alpha * acc64[a][b][r] + beta * (double) cin[a][b][r]);
            }
}

template <typename Tout, int NT>
void launch(hipStream_t s, bool lower, int M, int N, int K, double alpha, const float *X, int ldx, long sX,
            const float *Y, int ldy, long sY, double beta, Tout *C, int ldc, long sC, int batch, int tri) {
    const int mt = (M + XT - 1) / XT, nt = (N + XT - 1) / XT;
    const int lin = (lower && mt == nt) ? 1 : 0;
    dim3 grid(lin ? mt * (mt + 1) / 2 : mt, lin ? 1 : nt, batch);
    if (lower)
        gemm_bf16s_tn_kernel<Tout, true, NT><<<grid, 256, 0, s>>>(M, N, K, alpha, X, ldx, sX, Y, ldy, sY, beta, C, ldc,
                                                                  sC, lin, tri);
    else
        gemm_bf16s_tn_kernel<Tout, false, NT><<<grid, 256, 0, s>>>(M, N, K, alpha, X, ldx, sX, Y, ldy, sY, beta, C,
                                                                   ldc, sC, 0, tri);
}

}  // namespace

template <typename Tout>
void gemm_bf16s_tn(hipStream_t s, int terms, bool lower, int M, int N, int K, double alpha, const float *X, int ldx,
                   long sX, const float *Y, int ldy, long sY, double beta, Tout *C, int ldc, long sC, int batch,
                   int tri) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    if (terms == 3)
        launch<Tout, 3>(s, lower, M, N, K, alpha, X, ldx, sX, Y, ldy, sY, beta, C, ldc, sC, batch, tri);
    else
        launch<Tout, 2>(s, lower, M, N, K, alpha, X, ldx, sX, Y, ldy, sY, beta, C, ldc, sC, batch, tri);
}

template void gemm_bf16s_tn<float>(hipStream_t, int, bool, int, int, int, double, const float *, int, long,
                                   const float *, int, long, double, float *, int, long, int, int);
template void gemm_bf16s_tn<double>(hipStream_t, int, bool, int, int, int, double, const float *, int, long,
                                    const float *, int, long, double, double *, int, long, int, int);

}  // namespace hdk
