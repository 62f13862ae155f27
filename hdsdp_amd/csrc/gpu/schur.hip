// schur.hip -- dual-buffer assembly and Schur-complement kernels.
//
// Coefficient layout on the device (built once at load):
//   ctype[k]            kind of coefficient k (0 = C, 1..m = A_k-1)
//   cBeg/cRow/cCol/cVal CSR over coefficients for SPARSE (lower triplets) and
//                       SPR1 (cRow = factor index, cVal = factor value)
//   dSlot/dVal          DENSE packed lower columns, one n(n+1)/2 block each
//   rSlot/rF            DSR1 dense factors (n x nR1), sign in csign
//   scatter map         every SPARSE/SPR1 entry of the lower triangle grouped by
//                       position, constraints in order and C last, so buffer
//                       assembly is a deterministic gather with the reference's
//                       summation order (interface/hdsdp_conic_sdp.c:373-379).
//   small table         up to SMALL_K entries per constraint for the pair kernel.
#include "schur.h"

namespace hdk {

// ---- buffer assembly -----------------------------------------------------------

// Dense part of a buffer: y = sum_d coef[denseCon[d]] A_d over the packed lower
// entries e (one thread per entry pair e, e + 256 of a 512-entry chunk), the slots
// summed in order, then the packed sum written out as the full lower triangle (zeros
// above).  The packed columns of consecutive slots are npk apart; a wave's loads of one
// slot are contiguous, and sixteen loads per thread are issued before they are summed,
// so the kernel streams the packed coefficient matrices at HBM rate.
constexpr int ASM_CHUNK = 512;

__global__ __launch_bounds__(256) void assemble_packed_kernel(long npk, int nDense, const int *denseCon,
                                                              const double *dVal, const double *coef, double *y) {
    __shared__ double cs[ASM_CHUNK];
    const long e0 = blockIdx.x * 512L + threadIdx.x, e1 = e0 + 256;
    const bool a0 = e0 < npk, a1 = e1 < npk;
    double v0 = 0.0, v1 = 0.0;
    for (int d0 = 0; d0 < nDense; d0 += ASM_CHUNK) {
        const int cnt = min(ASM_CHUNK, nDense - d0);
        __syncthreads();
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) cs[t] = coef[denseCon[d0 + t]];
        __syncthreads();
        const double *q = dVal + (long) d0 * npk;
        int t = 0;
        for (; t + 8 <= cnt; t += 8) {
            double x0[8], x1[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                x0[u] = a0 ? q[(long) (t + u) * npk + e0] : 0.0;
                x1[u] = a1 ? q[(long) (t + u) * npk + e1] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                v0 += cs[t + u] * x0[u];
                v1 += cs[t + u] * x1[u];
            }
        }
        for (; t < cnt; ++t) {
            if (a0) v0 += cs[t] * q[(long) t * npk + e0];
            if (a1) v1 += cs[t] * q[(long) t * npk + e1];
        }
    }
    if (a0) y[e0] = v0;
    if (a1) y[e1] = v1;
}

// target(i, j) = y(packed (i, j)) for i >= j, 0 above (packed column j starts at
// j n - j (j - 1) / 2)
__global__ void unpack_packed_kernel(int n, const double *y, double *target) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    target[i + (long) j * n] = (i >= j) ? y[(long) j * n - (long) j * (j - 1) / 2 + (i - j)] : 0.0;
}

__global__ void scatter_kernel(int nPos, const long *posIdx, const int *posBeg, const int *entCon,
                               const double *entVal, const double *coef, double *target) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nPos) return;
    double v = target[posIdx[q]];
    for (int e = posBeg[q]; e < posBeg[q + 1]; ++e) {
        double c = coef[entCon[e]];
        if (c != 0.0) v += c * entVal[e];
    }
    target[posIdx[q]] = v;
}

// target(i, j) += w * f_i f_j, i >= j (one dense rank-one coefficient)
__global__ void rank1_update_kernel(int n, double w, const double *f, double *target) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || i < j) return;
    target[i + (long) j * n] += w * f[i] * f[j];
}

// Fw[:, k] = w_k F[:, k]
__global__ void scale_cols_kernel(int n, int k, const double *F, const double *w, double *Fw) {
    long total = (long) n * k;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < total; e += (long) gridDim.x * blockDim.x) {
        Fw[e] = F[e] * w[e / n];
    }
}

__global__ void dsr1_weights_kernel(int nR1, const int *rCon, const double *csign, const double *coef, double *w) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nR1) w[q] = coef[rCon[q]] * csign[rCon[q]];
}

void assemble_buffer(hipStream_t s, const DevCoeffs &c, const double *coef, double cCoef, double *target,
                     double eye, double *work, double *splitWork, size_t splitLen) {
    const int n = c.n;
    dim3 g((n + 255) / 256, n);
    const long npk = (long) n * (n + 1) / 2;
    if (c.nDense > 0) {
        assemble_packed_kernel<<<(unsigned) ((npk + 511) / 512), 256, 0, s>>>(npk, c.nDense, c.denseCon, c.dVal,
                                                                               coef, c.packW);
        unpack_packed_kernel<<<g, 256, 0, s>>>(n, c.packW, target);
    } else {
        fill(s, target, (long) n * n, 0.0);
    }
    if (c.nPos > 0) scatter_kernel<<<(c.nPos + 255) / 256, 256, 0, s>>>(c.nPos, c.posIdx, c.posBeg, c.entCon,
                                                                         c.entVal, coef, target);
    if (c.nR1 > 0) {
        // target += F diag(coef * sign) F^T (lower) as one GEMM
        dsr1_weights_kernel<<<(c.nR1 + 255) / 256, 256, 0, s>>>(c.nR1, c.rCon, c.csign, coef, c.r1w);
        scale_cols_kernel<<<4096, 256, 0, s>>>(n, c.nR1, c.rF, c.r1w, work);
        gemm_splitk<double>(s, false, true, true, n, n, c.nR1, 1.0, work, n, c.rF, n, 1.0, target, n, splitWork,
                            splitLen);
    }
    if (c.cDSR1 && cCoef != 0.0) {
        // C = sign f f^T enters last among the coefficients, as a rank-one update
        rank1_update_kernel<<<g, 256, 0, s>>>(n, cCoef * c.cR1Sign, c.cR1F, target);
    }
    if (eye != 0.0) add_diag(s, n, target, n, nullptr, eye);
}

__global__ void axpy_buffer_kernel(long nn, double step, const double *src, const double *dS, double *dst) {
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < nn; e += (long) gridDim.x * blockDim.x)
        dst[e] = src[e] + step * dS[e];
}

void axpy_buffer(hipStream_t s, long nn, double step, const double *src, const double *dS, double *dst) {
    long blocks = (nn + 255) / 256;
    axpy_buffer_kernel<<<blocks > 8192 ? 8192 : blocks, 256, 0, s>>>(nn, step, src, dS, dst);
}

// ---- traces <A_k, B> -----------------------------------------------------------------

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;
}

// Dense coefficients: <A_d, B> = sum over packed lower entries e of A_d(e) w(e) B(e),
// w = 1 on the diagonal and 2 below it.  The weighted packed B (and B2) is formed once,
// then one pass over the packed coefficients computes TR_CPB traces per workgroup over
// one of TR_SPLIT k-parts (a multi-column dot product streamed at HBM rate); the parts
// are added in a fixed order.
constexpr int TR_CPB = 4;

__global__ void pack_weighted_kernel(int n, const double *B1, const double *B2, double *w1, double *w2) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || i < j) return;
    const long e = (long) j * n - (long) j * (j - 1) / 2 + (i - j);
    const double w = (i == j) ? 1.0 : 2.0;
    w1[e] = w * B1[i + (long) j * n];
    if (B2) w2[e] = w * B2[i + (long) j * n];
}

__global__ __launch_bounds__(256) void dense_trace_kernel(long npk, int nDense, const double *dVal, const double *w1,
                                                          const double *w2, double *part1, double *part2) {
    __shared__ double red[2 * TR_CPB][4];
    const int d0 = blockIdx.x * TR_CPB, split = blockIdx.y, tid = threadIdx.x;
    const long chunk = ((npk + TR_SPLIT - 1) / TR_SPLIT + 255) / 256 * 256;
    const long ks = split * chunk, ke = min(npk, ks + chunk);
    double a1[TR_CPB], a2[TR_CPB];
    const double *A[TR_CPB];
#pragma unroll
    for (int c = 0; c < TR_CPB; ++c) {
        a1[c] = 0.0;
        a2[c] = 0.0;
        A[c] = dVal + (long) min(d0 + c, nDense - 1) * npk;
    }
    for (long e = ks + tid; e < ke; e += 256) {
        const double b1 = w1[e], b2 = w2 ? w2[e] : 0.0;
#pragma unroll
        for (int c = 0; c < TR_CPB; ++c) {
            const double av = A[c][e];
            a1[c] += av * b1;
            a2[c] += av * b2;
        }
    }
#pragma unroll
    for (int c = 0; c < TR_CPB; ++c) {
        a1[c] = wave_sum(a1[c]);
        a2[c] = wave_sum(a2[c]);
        if ((tid & 63) == 0) {
            red[c][tid >> 6] = a1[c];
            red[TR_CPB + c][tid >> 6] = a2[c];
        }
    }
    __syncthreads();
    if (tid < TR_CPB && d0 + tid < nDense) {
        const int c = tid;
        part1[(long) split * nDense + d0 + c] = red[c][0] + red[c][1] + red[c][2] + red[c][3];
        part2[(long) split * nDense + d0 + c] =
            red[TR_CPB + c][0] + red[TR_CPB + c][1] + red[TR_CPB + c][2] + red[TR_CPB + c][3];
    }
}

__global__ void dense_trace_finish_kernel(int nDense, const int *denseCon, int k0, int cnt, const double *part1,
                                          const double *part2, double s2, double *out1, double *out2) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= nDense) return;
    const int k = denseCon[d];
    if (k < k0 || k >= k0 + cnt) return;
    double t1 = 0.0, t2 = 0.0;
    for (int p = 0; p < TR_SPLIT; ++p) {
        t1 += part1[(long) p * nDense + d];
        t2 += part2[(long) p * nDense + d];
    }
    out1[k - k0] = t1;
    if (out2) out2[k - k0] = s2 * t2;
}

// out1[k] = <A_k, B1>, out2[k] = s2 <A_k, B2> (B2 may be NULL) for coefficients k0 .. k0 + cnt - 1,
// written to out[k - k0].  One workgroup of 256 threads per coefficient; dense
// coefficients are left to dense_trace_kernel.
__global__ __launch_bounds__(256) void trace_kernel(int n, DevCoeffs c, int k0, const double *B1, const double *B2,
                                                    double s2, double *out1, double *out2, int withR1) {
    __shared__ double red1[4], red2[4];
    const int k = k0 + blockIdx.x;
    const int tid = threadIdx.x;
    const int t = c.ctype[k];
    if (t == COEFF_DSR1 && !withR1) return;  // left to r1_quadforms
    if (t == COEFF_DENSE) return;            // left to dense_trace_kernel
    double a1 = 0.0, a2 = 0.0;
    if (t == COEFF_SPARSE) {
        for (int e = c.cBeg[k] + tid; e < c.cBeg[k + 1]; e += 256) {
            int r = c.cRow[e], cc = c.cCol[e];
            double w = (r == cc) ? c.cVal[e] : 2.0 * c.cVal[e];
            long idx = r + (long) cc * n;
            a1 += w * B1[idx];
            if (B2) a2 += w * B2[idx];
        }
    } else if (t == COEFF_SPR1) {
        int b = c.cBeg[k], nnz = c.cBeg[k + 1] - b;
        for (int e = tid; e < nnz * nnz; e += 256) {
            int p = e % nnz, q = e / nnz;
            double w = c.cVal[b + p] * c.cVal[b + q];
            long idx = c.cRow[b + p] + (long) c.cRow[b + q] * n;
            a1 += w * B1[idx];
            if (B2) a2 += w * B2[idx];
        }
        a1 *= c.csign[k];
        a2 *= c.csign[k];
    } else if (t == COEFF_DSR1) {
        // f^T B f with B symmetric: column i of B is read contiguously
        const double *f = c.rF + (long) c.rSlot[k] * n;
        for (int i = tid; i < n; i += 256) {
            double r1 = 0.0, r2 = 0.0;
            const double *b1 = B1 + (long) i * n;
            const double *b2 = B2 ? B2 + (long) i * n : nullptr;
            for (int j = 0; j < n; ++j) {
                r1 += b1[j] * f[j];
                if (b2) r2 += b2[j] * f[j];
            }
            a1 += f[i] * r1;
            a2 += f[i] * r2;
        }
        a1 *= c.csign[k];
        a2 *= c.csign[k];
    }
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    if ((tid & 63) == 0) {
        red1[tid >> 6] = a1;
        red2[tid >> 6] = a2;
    }
    __syncthreads();
    if (tid == 0) {
        out1[blockIdx.x] = red1[0] + red1[1] + red1[2] + red1[3];
        if (out2) out2[blockIdx.x] = s2 * (red2[0] + red2[1] + red2[2] + red2[3]);
    }
}

void traces(hipStream_t s, const DevCoeffs &c, int k0, int cnt, const double *B1, const double *B2, double s2,
            double *out1, double *out2, bool withR1) {
    if (cnt <= 0) return;
    trace_kernel<<<cnt, 256, 0, s>>>(c.n, c, k0, B1, B2, s2, out1, out2, withR1 ? 1 : 0);
    if (c.nDense > 0) {
        const int n = c.n;
        const long npk = (long) n * (n + 1) / 2;
        double *w1 = c.packW, *w2 = B2 ? c.packW + npk : nullptr;
        pack_weighted_kernel<<<dim3((n + 255) / 256, n), 256, 0, s>>>(n, B1, B2, w1, w2);
        dense_trace_kernel<<<dim3((c.nDense + TR_CPB - 1) / TR_CPB, TR_SPLIT), 256, 0, s>>>(
            npk, c.nDense, c.dVal, w1, w2, c.trPart, c.trPart + (long) TR_SPLIT * c.nDense);
        dense_trace_finish_kernel<<<(c.nDense + 255) / 256, 256, 0, s>>>(
            c.nDense, c.denseCon, k0, cnt, c.trPart, c.trPart + (long) TR_SPLIT * c.nDense, s2, out1, out2);
    }
}

// ---- dense rank-one coefficients through GEMMs ------------------------------------------

// One wave per factor column q: with h = H[:, q] = B f_q,
//   out1[con - k0] = sign f_q^T h,  out2[con - k0] = s2 sign h^T h.
// Columns 0 .. nR1 - 1 are the DSR1 constraints (coefficient rCon[q]); column nR1,
// when present, is a dense rank-one C (coefficient 0).
__global__ __launch_bounds__(256) void r1_vectors_kernel(int n, int nR1, int nCols, const int *rCon,
                                                         const double *csign, const double *F, const double *H,
                                                         int k0, int cnt, double s2, double *out1, double *out2) {
    const int q = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (q >= nCols) return;
    const int con = (q < nR1) ? rCon[q] : 0;
    const double *f = F + (long) q * n, *h = H + (long) q * n;
    double d1 = 0.0, d2 = 0.0;
    for (int i = lane; i < n; i += 64) {
        d1 += f[i] * h[i];
        d2 += h[i] * h[i];
    }
    d1 = wave_sum(d1);
    d2 = wave_sum(d2);
    if (lane == 0 && con >= k0 && con < k0 + cnt) {
        out1[con - k0] = csign[con] * d1;
        if (out2) out2[con - k0] = s2 * csign[con] * d2;
    }
}

void r1_quadforms(hipStream_t s, const DevCoeffs &c, const double *B, bool withC, int k0, int cnt, double s2,
                  double *out1, double *out2, double *H) {
    const int n = c.n;
    const int nCols = c.nR1 + ((withC && c.cDSR1) ? 1 : 0);
    if (nCols == 0) return;
    gemm<double>(s, false, false, false, n, nCols, n, 1.0, B, n, 0, c.rF, n, 0, 0.0, H, n, 0, 1);
    r1_vectors_kernel<<<(nCols + 3) / 4, 256, 0, s>>>(n, c.nR1, nCols, c.rCon, c.csign, c.rF, H, k0, cnt, s2,
                                                      out1, out2);
}

// M(rp, rq) = s_p s_q K_pq^2 for p >= q over the DSR1 slots, rp = rCon[p] - 1 (lower triangle of M).
__global__ void r1_square_kernel(int nR1, const int *rCon, const double *csign, const double *K, long ldk,
                                 double *M, long ldm) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (p >= nR1 || p < q) return;
    const int cp = rCon[p], cq = rCon[q];
    const double kv = K[p + (long) q * ldk];
    const long i = (cp > cq ? cp : cq) - 1, j = (cp > cq ? cq : cp) - 1;
    M[i + j * ldm] = csign[cp] * csign[cq] * kv * kv;
}

void r1_square(hipStream_t s, const DevCoeffs &c, const double *K, long ldk, double *M, long ldm) {
    const int r = c.nR1;
    if (r == 0) return;
    r1_square_kernel<<<dim3((r + 255) / 256, r), 256, 0, s>>>(r, c.rCon, c.csign, K, ldk, M, ldm);
}

void r1_schur(hipStream_t s, const DevCoeffs &c, const double *H, double *K, long ldk, double *M, long ldm) {
    const int n = c.n, r = c.nR1;
    if (r == 0) return;
    // K = F^T (W F), lower triangle
    gemm<double>(s, true, false, true, r, r, n, 1.0, c.rF, n, 0, H, n, 0, 0.0, K, (int) ldk, 0, 1);
    r1_square_kernel<<<dim3((r + 255) / 256, r), 256, 0, s>>>(r, c.rCon, c.csign, K, ldk, M, ldm);
}

// ---- dense coefficients through a Gram matrix -------------------------------------------

// out[b] = full n x n matrix of dense slot slot0 + b (unpacked from the lower packed column)
// Full n x n copies of the packed dense coefficients slot0 .. slot0 + cnt - 1, one
// workgroup per 64 x 64 lower tile: the tile is read from the packed columns (lane =
// row, coalesced), written to its place and, through LDS, to its mirror above the
// diagonal (lane = row of the mirror, coalesced).  Unpacking element by element read
// the upper half with a stride of a packed column per lane.
template <typename T>
__global__ __launch_bounds__(256) void unpack_dense_kernel(int n, const double *dVal, int slot0, T *out) {
    __shared__ T tile[64][65];  // tile[c][r] = A(i0 + r, j0 + c)
    const long npk = (long) n * (n + 1) / 2;
    const double *v = dVal + (long) (slot0 + blockIdx.y) * npk;
    T *o = out + (long) blockIdx.y * n * n;
    const int t = blockIdx.x;
    int bi = (int) ((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    while (bi * (bi + 1) / 2 > t) --bi;
    const int bj = t - bi * (bi + 1) / 2;
    const int i0 = bi * 64, j0 = bj * 64, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int c = ty; c < 64; c += 4) {
        const long i = i0 + tx, j = j0 + c;
        if (i < n && j < n && i >= j) {
            const T val = (T) v[j * n - j * (j + 1) / 2 + i];
            o[i + j * n] = val;
            tile[c][tx] = val;
        }
    }
    __syncthreads();
    for (int c = ty; c < 64; c += 4) {
        // mirror entry (j0 + tx, i0 + c) = A(i0 + c, j0 + tx) = tile[tx][c], above the diagonal only
        const long r = j0 + tx, q = i0 + c;
        if (r < n && q < n && q > r) o[r + q * n] = tile[tx][c];
    }
}

void unpack_dense_f(hipStream_t s, const DevCoeffs &c, int slot0, int cnt, float *out) {
    if (cnt <= 0) return;
    const int nt = (c.n + 63) / 64;
    unpack_dense_kernel<float><<<dim3(nt * (nt + 1) / 2, cnt), 256, 0, s>>>(c.n, c.dVal, slot0, out);
}

void unpack_dense(hipStream_t s, const DevCoeffs &c, int slot0, int cnt, double *out) {
    if (cnt <= 0) return;
    const int nt = (c.n + 63) / 64;
    unpack_dense_kernel<double><<<dim3(nt * (nt + 1) / 2, cnt), 256, 0, s>>>(c.n, c.dVal, slot0, out);
}

// M(i, j) (lower) = sum over the partial Gram blocks, in block order; i, j index dense slots,
// mapped to constraints through denseCon.
__global__ void gram_reduce_kernel(int r, int nPart, const double *P, long ldp, long strideP, const int *denseCon,
                                   double *M, long ldm) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (p >= r || p < q) return;
    double v = 0.0;
    for (int b = 0; b < nPart; ++b) v += P[(long) b * strideP + p + (long) q * ldp];
    const int cp = denseCon[p], cq = denseCon[q];
    const long i = (cp > cq ? cp : cq) - 1, j = (cp > cq ? cq : cp) - 1;
    M[i + j * ldm] = v;
}

// Strictly lower part of each symmetric T_p (n x n, lower triangle valid) packed column
// by column into off[p * K1 + j n - j (j + 1) / 2 + (i - j - 1)], i > j, and its diagonal
// into dg[p * n + i].
template <typename T>
__global__ void pack_lower_kernel(int n, const T *Tin, long strideIn, T *off, long K1, T *dg) {
    const int p = blockIdx.y;
    const long nn = (long) n * n;
    const T *src = Tin + p * strideIn;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < nn; e += (long) gridDim.x * blockDim.x) {
        const long i = e % n, j = e / n;
        if (i > j) off[p * K1 + j * n - j * (j + 1) / 2 + (i - j - 1)] = src[e];
        else if (i == j) dg[(long) p * n + i] = src[e];
    }
}

template <typename T>
void pack_lower(hipStream_t s, int n, int cnt, const T *Tin, long strideIn, T *off, T *dg) {
    if (cnt <= 0) return;
    const long K1 = (long) n * (n - 1) / 2;
    pack_lower_kernel<T><<<dim3(1024, cnt), 256, 0, s>>>(n, Tin, strideIn, off, K1, dg);
}

template void pack_lower<double>(hipStream_t, int, int, const double *, long, double *, double *);
template void pack_lower<float>(hipStream_t, int, int, const float *, long, float *, float *);

__global__ void add_diag_delta_kernel(int r, const int *con, const double *cur, const double *ref, double *M,
                                      long ldm) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= r) return;
    const long i = con[p] - 1;
    M[i + i * ldm] += cur[i] - ref[i];
}

void add_diag_delta(hipStream_t s, int r, const int *con, const double *cur, const double *ref, double *M, long ldm) {
    if (r <= 0) return;
    add_diag_delta_kernel<<<(r + 255) / 256, 256, 0, s>>>(r, con, cur, ref, M, ldm);
}

void gram_reduce(hipStream_t s, int r, int nPart, const double *P, long ldp, long strideP, const int *denseCon,
                 double *M, long ldm) {
    gram_reduce_kernel<<<dim3((r + 255) / 256, r), 256, 0, s>>>(r, nPart, P, ldp, strideP, denseCon, M, ldm);
}

// ---- Schur pair kernel over small sparse / rank-one constraints -----------------------

struct SmallEnt {
    int kind, nnz;
    int r[SMALL_K], c[SMALL_K];
    double v[SMALL_K];
    double sign;
};

__device__ inline double pair_value(const SmallEnt &A, const SmallEnt &B, const double *W, int n) {
#define WW(i, j) W[(i) + (long) (j) * n]
    if (A.kind == SM_R1 && B.kind == SM_R1) {
        double q = 0.0;
        for (int p = 0; p < A.nnz; ++p)
            for (int t = 0; t < B.nnz; ++t) q += A.v[p] * B.v[t] * WW(A.r[p], B.r[t]);
        return q * q * A.sign * B.sign;
    }
    const SmallEnt &S = (A.kind == SM_SP) ? A : B;   // sparse side
    const SmallEnt &O = (A.kind == SM_SP) ? B : A;   // other side
    double res = 0.0;
    for (int a = 0; a < S.nnz; ++a) {
        int ar = S.r[a], ac = S.c[a];
        double tb = 0.0;
        if (O.kind == SM_SP) {
            for (int b = 0; b < O.nnz; ++b) {
                int br = O.r[b], bc = O.c[b];
                tb += O.v[b] * WW(ar, br) * WW(ac, bc);
                if (br != bc) tb += O.v[b] * WW(ar, bc) * WW(ac, br);
            }
        } else {
            double ur = 0.0, uc = 0.0;
            for (int p = 0; p < O.nnz; ++p) {
                ur += O.v[p] * WW(ar, O.r[p]);
                uc += O.v[p] * WW(ac, O.r[p]);
            }
            tb = O.sign * ur * uc;
        }
        res += (ar == ac) ? 0.5 * S.v[a] * tb : S.v[a] * tb;
    }
    return 2.0 * res;
#undef WW
}

__device__ inline void load_small(const DevCoeffs &c, int k, SmallEnt &e) {
    e.kind = c.smKind[k];
    e.nnz = c.smNnz[k];
    e.sign = c.csign[k];
    for (int q = 0; q < SMALL_K; ++q) {
        e.r[q] = c.smRow[(long) k * SMALL_K + q];
        e.c[q] = c.smCol[(long) k * SMALL_K + q];
        e.v[q] = c.smVal[(long) k * SMALL_K + q];
    }
}

// One 64 x 64 lower tile of M per workgroup; thread (il, jg) computes rows i0 + il
// against columns j0 + jg, jg + 4, ...  Constraint k of M is coefficient k + 1.
__global__ __launch_bounds__(256) void schur_pair_kernel(int m, int n, DevCoeffs c, const double *W, double *M,
                                                         long ldm) {
    __shared__ SmallEnt Ti[64];
    __shared__ SmallEnt Tj[64];
    long t = blockIdx.x;
    long bi = (long) ((sqrt(8.0 * (double) t + 1.0) - 1.0) / 2.0);
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    while (bi * (bi + 1) / 2 > t) --bi;
    long bj = t - bi * (bi + 1) / 2;
    const int i0 = (int) bi * 64, j0 = (int) bj * 64;
    const int tid = threadIdx.x;
    if (tid < 64) {
        int i = i0 + tid;
        if (i < m) load_small(c, i + 1, Ti[tid]);
        else Ti[tid].kind = SM_NONE;
    } else if (tid < 128) {
        int j = j0 + tid - 64;
        if (j < m) load_small(c, j + 1, Tj[tid - 64]);
        else Tj[tid - 64].kind = SM_NONE;
    }
    __syncthreads();
    const int il = tid & 63, jg = tid >> 6;
    const int i = i0 + il;
    if (i >= m) return;
    const SmallEnt &A = Ti[il];
    if (A.kind != SM_SP && A.kind != SM_R1) return;
    for (int jl = jg; jl < 64; jl += 4) {
        int j = j0 + jl;
        if (j >= m || j > i) continue;
        const SmallEnt &B = Tj[jl];
        if (B.kind != SM_SP && B.kind != SM_R1) continue;
        M[i + (long) j * ldm] = pair_value(A, B, W, n);
    }
}

// The pair kernel when every small constraint is a single sparse entry (i, j) (graph
// edges, diagonal entries) and n <= PAIR1_MAXN: a workgroup takes PAIR1_COLS columns q
// of M, stages the two columns of W = S^-1 their entries (r_q, c_q) address in LDS,
// and walks the rows p >= q: M(p, q) needs W(r_p, r_q), W(c_p, c_q), W(r_p, c_q) and
// W(c_p, r_q), read from LDS instead of gathered from L2.  Each entry is formed by the
// statements of pair_value for two single-entry constraints, in the same order.
__global__ __launch_bounds__(256) void schur_pair_single_kernel(int m, int n, DevCoeffs c, const double *W, double *M,
                                                                long ldm) {
    extern __shared__ double wl[];  // [PAIR1_COLS][2][n]: W(:, r_q), W(:, c_q)
    __shared__ int qr[PAIR1_COLS], qc[PAIR1_COLS], qok[PAIR1_COLS];
    __shared__ double qv[PAIR1_COLS];
    const int j0 = blockIdx.x * PAIR1_COLS, tid = threadIdx.x;
    if (tid < PAIR1_COLS) {
        const int j = j0 + tid, k = j + 1;
        const bool ok = j < m && c.smKind[k] == SM_SP;
        qok[tid] = ok;
        qr[tid] = ok ? c.smRow[(long) k * SMALL_K] : 0;
        qc[tid] = ok ? c.smCol[(long) k * SMALL_K] : 0;
        qv[tid] = ok ? c.smVal[(long) k * SMALL_K] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < PAIR1_COLS * 2 * n; e += blockDim.x) {
        const int q = e / (2 * n), h = (e / n) & 1, row = e % n;
        const int col = h ? qc[q] : qr[q];
        wl[e] = qok[q] ? W[row + (long) col * n] : 0.0;
    }
    __syncthreads();
    for (int i = j0 + tid; i < m; i += blockDim.x) {
        const int k = i + 1;
        if (c.smKind[k] != SM_SP) continue;
        const int ar = c.smRow[(long) k * SMALL_K], ac = c.smCol[(long) k * SMALL_K];
        const double sv = c.smVal[(long) k * SMALL_K];
#pragma unroll
        for (int q = 0; q < PAIR1_COLS; ++q) {
            const int j = j0 + q;
            if (!qok[q] || j > i) continue;
            const double *wr = wl + (long) (2 * q) * n, *wc = wr + n;  // W(:, br), W(:, bc)
            const int br = qr[q], bc = qc[q];
            const double ov = qv[q];
            double res = 0.0;
            double tb = 0.0;
            tb += ov * wr[ar] * wc[ac];
            if (br != bc) tb += ov * wc[ar] * wr[ac];
            res += (ar == ac) ? 0.5 * sv * tb : sv * tb;
            M[i + (long) j * ldm] = 2.0 * res;
        }
    }
}

void schur_pairs(hipStream_t s, const DevCoeffs &c, const double *W, double *M, long ldm) {
    const int m = c.m;
    if (c.pairSingle) {
        const size_t lds = sizeof(double) * PAIR1_COLS * 2 * (size_t) c.n;
        schur_pair_single_kernel<<<(m + PAIR1_COLS - 1) / PAIR1_COLS, 256, lds, s>>>(m, c.n, c, W, M, ldm);
        return;
    }
    long nt = (m + 63) / 64;
    long tiles = nt * (nt + 1) / 2;
    schur_pair_kernel<<<tiles, 256, 0, s>>>(m, c.n, c, W, M, ldm);
}

// M(max(h, k), min(h, k)) = t[k] for all k (row/column h of M from a trace vector)
__global__ void scatter_row_kernel(int m, int h, const double *t, double *M, long ldm) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    int i = (k > h) ? k : h, j = (k > h) ? h : k;
    M[i + (long) j * ldm] = t[k];
}

void scatter_row(hipStream_t s, int m, int h, const double *t, double *M, long ldm) {
    scatter_row_kernel<<<(m + 255) / 256, 256, 0, s>>>(m, h, t, M, ldm);
}

// dense full symmetric copy of coefficient k into out (n x n)
__global__ void densify_kernel(int n, DevCoeffs c, int k, double *out) {
    int t = c.ctype[k];
    if (t == COEFF_SPARSE || t == COEFF_SPR1) {
        int b = c.cBeg[k], e1 = c.cBeg[k + 1];
        if (t == COEFF_SPARSE) {
            for (int e = b + threadIdx.x + blockIdx.x * blockDim.x; e < e1; e += blockDim.x * gridDim.x) {
                int r = c.cRow[e], cc = c.cCol[e];
                out[r + (long) cc * n] = c.cVal[e];
                out[cc + (long) r * n] = c.cVal[e];
            }
        } else {
            int nnz = e1 - b;
            for (int e = threadIdx.x + blockIdx.x * blockDim.x; e < nnz * nnz; e += blockDim.x * gridDim.x) {
                int p = e % nnz, q = e / nnz;
                out[c.cRow[b + p] + (long) c.cRow[b + q] * n] = c.csign[k] * c.cVal[b + p] * c.cVal[b + q];
            }
        }
    } else if (t == COEFF_DENSE) {
        const long npk = (long) n * (n + 1) / 2;
        const double *v = c.dVal + (long) c.dSlot[k] * npk;
        for (long e = threadIdx.x + blockIdx.x * (long) blockDim.x; e < (long) n * n; e += (long) blockDim.x * gridDim.x) {
            long i = e % n, j = e / n;
            long r = i > j ? i : j, q = i > j ? j : i;
            out[e] = v[q * n - q * (q + 1) / 2 + r];
        }
    } else if (t == COEFF_DSR1) {
        const double *f = c.rF + (long) c.rSlot[k] * n;
        const double sg = c.csign[k];
        for (long e = threadIdx.x + blockIdx.x * (long) blockDim.x; e < (long) n * n; e += (long) blockDim.x * gridDim.x) {
            long i = e % n, j = e / n;
            out[e] = sg * f[i] * f[j];
        }
    }
}

void densify(hipStream_t s, const DevCoeffs &c, int k, double *out) {
    fill(s, out, (long) c.n * c.n, 0.0);
    densify_kernel<<<64, 256, 0, s>>>(c.n, c, k, out);
}

// ---- small helpers ------------------------------------------------------------------

// out = s * (A + A^T) / 2 + d * I (A full n x n)
__global__ void sym_avg_kernel(int n, const double *A, double s, double d, double *out) {
    long total = (long) n * n;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < total; e += (long) gridDim.x * blockDim.x) {
        long i = e % n, j = e / n;
        double v = 0.5 * (A[i + j * n] + A[j + i * n]);
        if (i == j) v += d;
        out[e] = s * v;
    }
}

void sym_avg(hipStream_t s, int n, const double *A, double scale, double diag, double *out) {
    sym_avg_kernel<<<4096, 256, 0, s>>>(n, A, scale, diag, out);
}

// out[0] = sum_{i >= j} (2 - [i == j]) X_ij S_ij ; out[1] = tr(X) ; out[2] = tr(W)-like trace of A (diag)
__global__ __launch_bounds__(1024) void xdots_kernel(int n, const double *X, const double *S, double *out) {
    __shared__ double red[17];
    double a = 0.0, t = 0.0;
    long total = (long) n * n;
    for (long e = threadIdx.x; e < total; e += blockDim.x) {
        long i = e % n, j = e / n;
        if (i > j) a += 2.0 * X[e] * S[e];
        else if (i == j) {
            a += X[e] * S[e];
            t += X[e];
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        t += __shfl_down(t, off);
    }
    __shared__ double red2[17];
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = a;
        red2[threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, st = 0.0;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) {
            sa += red[w];
            st += red2[w];
        }
        out[0] = sa;
        out[1] = st;
    }
}

void xdots(hipStream_t s, int n, const double *X, const double *S, double *out) {
    xdots_kernel<<<1, 1024, 0, s>>>(n, X, S, out);
}

__global__ __launch_bounds__(1024) void trace_diag_kernel(int n, const double *A, double *out) {
    __shared__ double red[17];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a += A[i + (long) i * n];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) s += red[w];
        out[0] = s;
    }
}

void trace_diag(hipStream_t s, int n, const double *A, double *out) {
    trace_diag_kernel<<<1, 1024, 0, s>>>(n, A, out);
}

}  // namespace hdk
