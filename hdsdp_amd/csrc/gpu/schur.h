// schur.h -- device coefficient layout, buffer assembly and Schur kernels.
#pragma once

#include "dev_common.h"

namespace hdk {

constexpr int COEFF_ZERO = 0, COEFF_SPARSE = 1, COEFF_DENSE = 2, COEFF_SPR1 = 3, COEFF_DSR1 = 4;
constexpr int SMALL_K = 4;                                   // entries per constraint in the pair kernel
constexpr int SM_NONE = 0, SM_SP = 1, SM_R1 = 2, SM_HEAVY = 3;  // pair-kernel kinds
constexpr int PAIR1_COLS = 8, PAIR1_MAXN = 480;                 // single-entry pair kernel: columns per
                                                                 // workgroup, largest n (LDS: 16 n doubles,
                                                                 // under 64 KB with the static arrays)

// Device-resident coefficients of one SDP block (k = 0 is C, k = i + 1 is A_i).
struct DevCoeffs {
    int m, n;
    int *ctype;       // m + 1
    double *csign;    // m + 1 (rank-one signs)
    int *cBeg;        // m + 2, CSR over SPARSE / SPR1 entries
    int *cRow, *cCol;
    double *cVal;
    int nDense;       // DENSE coefficients in assembly order (C last)
    int *denseCon;    // coefficient index per dense slot
    int *dSlot;       // m + 1, slot or -1
    double *dVal;     // nDense x n(n+1)/2
    int nR1;          // DSR1 constraints (C excluded)
    int *rCon;        // coefficient index per DSR1 slot
    int *rSlot;       // m + 1, slot or -1 (C included when DSR1: slot nR1)
    double *rF;       // n x (nR1 + [C is DSR1])
    double *r1w;      // nR1 weights scratch
    int cDSR1;        // C is a dense rank-one matrix
    double cR1Sign;
    double *cR1F;     // C factor (alias into rF)
    int nPos;         // scatter map over SPARSE / SPR1 entries
    long *posIdx;
    int *posBeg;
    int *entCon;
    double *entVal;
    int *smKind, *smNnz, *smRow, *smCol;  // (m + 1) x SMALL_K
    double *smVal;
    int pairSingle;   // every small constraint is one sparse entry and n <= PAIR1_MAXN
    double *packW;    // 2 n(n+1)/2 scratch: packed buffers of the dense assembly and traces
    double *trPart;   // 2 TR_SPLIT nDense partial traces of the dense coefficients
};

constexpr int TR_SPLIT = 4;  // k-parts of the dense trace products (summed in a fixed order)

// target = sum_k coef[k] A_k (coef[0] multiplies C) + eye I on the lower triangle.
// cCoef is coef[0] on the host (needed for a dense rank-one C).  work: n * nR1.
// splitWork (splitLen elements, may be null) holds the k-split partial products of the
// rank-nR1 update F diag(w) F^T (splitk_parts(n, n, nR1, true) x n^2 used).
void assemble_buffer(hipStream_t s, const DevCoeffs &c, const double *coef, double cCoef, double *target,
                     double eye, double *work, double *splitWork = nullptr, size_t splitLen = 0);
void axpy_buffer(hipStream_t s, long nn, double step, const double *src, const double *dS, double *dst);

// out1[k - k0] = <A_k, B1>, out2[k - k0] = s2 <A_k, B2> for k in [k0, k0 + cnt); dense
// rank-one coefficients are skipped (outputs untouched) unless withR1.
void traces(hipStream_t s, const DevCoeffs &c, int k0, int cnt, const double *B1, const double *B2, double s2,
            double *out1, double *out2, bool withR1);
// The same two quantities for every dense rank-one coefficient (and C when withC) by one
// GEMM H = B F: <A_q, B> = s_q f_q^T H_q and s2 <A_q, B^2> = s2 s_q |H_q|^2.  H: n x (nR1 + 1).
void r1_quadforms(hipStream_t s, const DevCoeffs &c, const double *B, bool withC, int k0, int cnt, double s2,
                  double *out1, double *out2, double *H);
// M entries between dense rank-one constraints: K = F^T H (H = W F from r1_quadforms),
// M = s s^T o K o K.  K: nR1 x nR1 scratch (ldk).
void r1_schur(hipStream_t s, const DevCoeffs &c, const double *H, double *K, long ldk, double *M, long ldm);
// M = s s^T o K o K from a precomputed K (lower)
void r1_square(hipStream_t s, const DevCoeffs &c, const double *K, long ldk, double *M, long ldm);
// full n x n copies of dense slots slot0 .. slot0 + cnt - 1
void unpack_dense(hipStream_t s, const DevCoeffs &c, int slot0, int cnt, double *out);
void unpack_dense_f(hipStream_t s, const DevCoeffs &c, int slot0, int cnt, float *out);
// T_p (cnt matrices n x n, lower valid, stride strideIn) -> strictly lower part packed by
// columns (off + p K1, K1 = n (n - 1) / 2) and diagonal (dg + p n)
template <typename T>
void pack_lower(hipStream_t s, int n, int cnt, const T *Tin, long strideIn, T *off, T *dg);
// M(denseCon[p] - 1, denseCon[q] - 1) = sum_b P_b(p, q), p >= q
void gram_reduce(hipStream_t s, int r, int nPart, const double *P, long ldp, long strideP, const int *denseCon,
                 double *M, long ldm);
// M(i, i) += cur[i] - ref[i] for i = con[p] - 1, p < r
void add_diag_delta(hipStream_t s, int r, const int *con, const double *cur, const double *ref, double *M, long ldm);
// M_ij = <A_i, W A_j W> for pairs of small constraints (lower triangle)
void schur_pairs(hipStream_t s, const DevCoeffs &c, const double *W, double *M, long ldm);
void scatter_row(hipStream_t s, int m, int h, const double *t, double *M, long ldm);
void densify(hipStream_t s, const DevCoeffs &c, int k, double *out);
void sym_avg(hipStream_t s, int n, const double *A, double scale, double diag, double *out);
void xdots(hipStream_t s, int n, const double *X, const double *S, double *out);
void trace_diag(hipStream_t s, int n, const double *A, double *out);

}  // namespace hdk
