// bf16planes.h -- products on the bf16 matrix cores from operands stored as three bf16
// planes (x = x0 + x1 + x2, FP32-equivalent), and the kernels that write such planes.
#pragma once

#include <hip/hip_bf16.h>

#include "dev_common.h"
#include "schur.h"

namespace hdk {

constexpr int BF16P_OUT_F32 = 0, BF16P_OUT_F64 = 1, BF16P_OUT_PLANES = 2;
// padded leading dimension of a plane whose columns hold n entries (16-byte vectors)
inline long plane_ld(long n) { return (n + 7) / 8 * 8; }

// C = alpha X^T Y + beta C, X: K x M and Y: K x N column-major bf16 planes (plane t at
// X + t psX, column stride ldx, batch stride sX), six bf16 products, FP32 accumulation
// flushed into FP64 every 512 k.  K, ldx, ldy, psX, psY, sX, sY and the operand
// addresses are multiples of 8 elements; rows K_true..K of the planes are zero.
// outKind: BF16P_OUT_F32 / F64 (C of that type, ld ldc) or BF16P_OUT_PLANES (three bf16
// planes of alpha X^T Y at C + t psC; beta is ignored).  lower restricts C to i >= j;
// tri = 1 when Y(k, j) = 0 for k > j, 2 when X(k, i) = 0 for k > i (those k skipped);
// tri = 3: X holds a symmetric matrix of which only the lower triangle with the diagonal
// halved enters (X(k, i) taken as 0 for k < i, X(i, i) / 2) and Y(k, j) = 0 for k > j,
// so C is upper triangular: only tiles on or above the diagonal are written (the rest
// of C is left as it is).
void gemm_bf16p(hipStream_t s, bool lower, int M, int N, int K, double alpha, const __bf16 *X, int ldx, long psX,
                long sX, const __bf16 *Y, int ldy, long psY, long sY, double beta, void *C, int outKind, int ldc,
                long psC, long sC, int batch = 1, int tri = 0);
// Symmetric C = alpha X^T Y (n x n, only i >= j formed) written as three bf16 planes,
// packed: the strictly lower part column by column (entry (i, j), i > j, at
// j n - j (j + 1) / 2 + i - j - 1 of off; batch stride sOff, plane stride psOff) and
// the diagonal (entry i of dg; batch stride sDg, plane stride psDg).  Operands as above.
// With a second pair (X2, Y2: the same leading dimensions, their own plane and batch
// strides), C = alpha (X^T Y + X2^T Y2), the two sums run one after the other in the
// same accumulators; tri = 4 when every operand column i is zero below row i (k range
// of a tile up to its smaller index).
void gemm_bf16p_packed(hipStream_t s, int n, int K, double alpha, const __bf16 *X, int ldx, long psX, long sX,
                       const __bf16 *Y, int ldy, long psY, long sY, __bf16 *off, long sOff, long psOff, __bf16 *dg,
                       long sDg, long psDg, int batch = 1, int tri = 0, const __bf16 *X2 = nullptr, long psX2 = 0,
                       long sX2 = 0, const __bf16 *Y2 = nullptr, long psY2 = 0, long sY2 = 0);
// planes (ld, plane stride ps) of a rows x cols matrix, rows..ld-1 zero
void split_matrix(hipStream_t s, int rows, int cols, const double *A, long lda, __bf16 *P, int ld, long ps);
void split_matrix_f(hipStream_t s, int rows, int cols, const float *A, long lda, __bf16 *P, int ld, long ps);
// full n x n planes (ld ldn, plane stride ps, matrix stride stride) of dense slots
// slot0 .. slot0 + cnt - 1; the padding rows are left as they are (zero)
void unpack_dense_split(hipStream_t s, const DevCoeffs &c, int slot0, int cnt, __bf16 *out, int ldn, long ps,
                        long stride);

}  // namespace hdk
