// bf16planes.hip -- FP32-equivalent products on the bf16 matrix cores of gfx950 from
// operands stored split.
//
// An FP32 (or FP64, rounded to FP32 first) value x is held as three bf16 planes,
//   x = x0 + x1 + x2,  x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1),
// exact for every FP32 x in the normal range (24 significant bits in three terms of 8).
// C = alpha X^T Y + beta C is formed from the six bf16 products whose weight is at least
// 2^-16 of the leading one: x0 y0 + (x1 y0 + x0 y1) + (x1 y1 + x0 y2 + x2 y0).  Every bf16 x
// bf16 product is exact in FP32.  v_mfma_f32_16x16x32_bf16 accumulates the leading product
// and the five correction products in two FP32 accumulators (the corrections are ~2^-8 of
// the leading sum and would lose their low bits against it), and every PFLUSH k slabs both
// are added into FP64 accumulators, so long sums (K = n^2 / 2) do not build up FP32
// rounding.
//
// The split is done once per operand, by the kernel that produces it, instead of in the
// product's inner loop: the loop only moves bf16 planes global -> registers -> LDS and
// issues MFMAs.
//
// Layout: plane t of a K x M operand X at X + t * ps, column i (k-contiguous) at X + i * ld;
// ld and every column offset are multiples of 8 (16-byte vector loads), and the k rows
// from the true length up to ld are zero, so products run over whole 8-element groups.
//
// Tiles: 128 x 128 per workgroup of eight waves (4 x 2), each wave 32 x 64 (2 x 4 MFMA
// 16 x 16 blocks); k slabs of 32 double-buffered in LDS (64-byte rows, the four 16-byte
// k groups of a row swizzled by the row so that both the stores and the fragment reads
// are free of bank conflicts; a lane's 8 consecutive k are one aligned 16-byte read).
// MFMA operand layout (gfx950):
// A(i, k) for i = lane % 16 and k = 8 (lane / 16) + t in element t; B(k, j) likewise with
// j = lane % 16; D(i, j) at lane j + 16 (i / 4), element i % 4.
#include "dev_common.h"
#include "bf16planes.h"

#include <type_traits>

namespace hdk {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int PT = 128;     // output tile
constexpr int PK = 32;      // k slab = one MFMA k
constexpr int PPAD = 0;     // row padding (bf16): none, the 16-byte k groups are swizzled instead
constexpr int PFLUSH = 16;  // slabs between FP64 accumulations (k = 512)
constexpr int PTHREADS = 512;
constexpr int PSLAB = 2 * 3 * PT * (PK + PPAD);  // bf16 elements of one operand's double buffer
constexpr int CPAD = 4;                          // row padding of the staged output tile (floats)
static_assert((long) PT * (PT + CPAD) * 4 <= 2L * PSLAB * 2, "output staging must fit in the slab buffers");

__device__ inline void split3(float x, __bf16 &a, __bf16 &b, __bf16 &c) {
    a = (__bf16) x;
    const float r1 = x - (float) a;
    b = (__bf16) r1;
    c = (__bf16) (r1 - (float) b);
}

// Writes the 128 x 128 output tile staged in LDS (Ct[column][row], FP32, already scaled
// by alpha) with NT threads.  OUT 3: the packed strictly-lower part and the diagonal as
// three bf16 planes, one row per thread, so a wave's stores of one column are
// consecutive packed entries.  OUT 2 / OUT 0: three bf16 planes / FP32, each thread
// writing four consecutive rows of one column, so a wave covers whole 128-row column
// segments.  C, Dg point at this batch entry.
template <int OUT, bool LOWER, int NT>
__device__ inline void store_staged(const float (*Ct)[PT + CPAD], int tid, int i0, int j0, int M, int N, void *Cv,
                                    int ldc, long psC, __bf16 *dg, long psDg) {
    if constexpr (OUT == 3) {
        __bf16 *off = static_cast<__bf16 *>(Cv);
        const int il = tid & (PT - 1), jq = tid / PT, i = i0 + il;
        for (int jl = jq; jl < PT; jl += NT / PT) {
            const int j = j0 + jl;
            if (j >= N || i >= M || i < j) continue;
            __bf16 p0, p1, p2;
            split3(Ct[jl][il], p0, p1, p2);
            if (i > j) {
                const long e = (long) j * M - (long) j * (j + 1) / 2 + (i - j - 1);
                off[e] = p0;
                off[e + psC] = p1;
                off[e + 2 * psC] = p2;
            } else {
                dg[i] = p0;
                dg[i + psDg] = p1;
                dg[i + 2 * psDg] = p2;
            }
        }
    } else {
        const int il = (tid & 31) * 4, jt = tid >> 5;  // rows il .. il + 3, columns jt + (NT / 32) q
        const int i = i0 + il;
#pragma unroll
        for (int q = 0; q < PT / (NT / 32); ++q) {
            const int jl = jt + (NT / 32) * q, j = j0 + jl;
            if (j >= N) continue;
            const float4 v = *reinterpret_cast<const float4 *>(&Ct[jl][il]);
            const float vv[4] = {v.x, v.y, v.z, v.w};
            const bool full = i + 3 < M && (!LOWER || i >= j);
            const long e = i + (long) j * ldc;
            if constexpr (OUT == 2) {
                __bf16 *C = static_cast<__bf16 *>(Cv);
                __bf16 pl[3][4];
#pragma unroll
                for (int u = 0; u < 4; ++u) split3(vv[u], pl[0][u], pl[1][u], pl[2][u]);
                if (full && (e & 3) == 0 && ((psC & 3) == 0)) {
#pragma unroll
                    for (int t = 0; t < 3; ++t) {
                        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                        *reinterpret_cast<bf16x4 *>(C + e + t * psC) = (bf16x4){pl[t][0], pl[t][1], pl[t][2], pl[t][3]};
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (i + u < M && (!LOWER || i + u >= j)) {
                            C[e + u] = pl[0][u];
                            C[e + u + psC] = pl[1][u];
                            C[e + u + 2 * psC] = pl[2][u];
                        }
                }
            } else {
                float *C = static_cast<float *>(Cv);
                if (full && (e & 3) == 0) {
                    *reinterpret_cast<float4 *>(C + e) = v;
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (i + u < M && (!LOWER || i + u >= j)) C[e + u] = vv[u];
                }
            }
        }
    }
}

// Lower tile of an mt x mt tile grid for workgroup id (0 <= id < ntile = mt (mt + 1) / 2).
// Workgroups go to the eight XCDs in turn (XCD id % 8), and XCD c is given a run of
// consecutive tiles of the blocked order, which lists the lower tiles super-block by
// super-block (4 x 4 tiles, row-major, each row-major inside): the tiles that one XCD
// holds read few strips of the operands, so its own L2 serves their re-reads.
__device__ inline void lower_tile_xcd(int id, int ntile, int mt, int &bx, int &by) {
    const int c = id % 8, q = id / 8, base = ntile / 8, rem = ntile % 8;
    const int t = c * base + min(c, rem) + q;
    constexpr int G = 4;
    const int nb = (mt + G - 1) / G;
    int acc = 0;
    for (int I = 0; I < nb; ++I)
        for (int J = 0; J <= I; ++J) {
            const int r0 = I * G, r1 = min(mt, r0 + G), c0 = J * G, c1 = min(mt, c0 + G);
            const int w = r1 - r0, cnt = (I > J) ? w * (c1 - c0) : w * (w + 1) / 2;
            if (t < acc + cnt) {
                const int u = t - acc;
                if (I > J) {
                    bx = r0 + u / (c1 - c0);
                    by = c0 + u % (c1 - c0);
                } else {
                    int x = 0;
                    while ((x + 1) * (x + 2) / 2 <= u) ++x;
                    bx = r0 + x;
                    by = c0 + (u - x * (x + 1) / 2);
                }
                return;
            }
            acc += cnt;
        }
    bx = by = 0;  // not reached for id < ntile
}

// tri: 1 when Y(k, j) = 0 for k > j, 2 when X(k, i) = 0 for k > i (k range of a tile
// ends at its last column or row), 3 and 4 as described in bf16planes.h.  With tri 2 or 3
// on a 2-D grid the tile row is rotated by the other grid coordinates, so that every XCD
// gets every row (the work of a tile depends on its row).  lin: blockIdx.x runs over the
// lower tiles only, row by row (1) or in the XCD-grouped order of lower_tile_xcd (2); with
// 4 added, workgroups are first reordered batch-major by XCD (below).
// OUT: 0 = FP32 C, 1 = FP64 C, 2 = three bf16 planes of C (at C + t * psC, ld ldc),
// 3 = three bf16 planes of a symmetric C (M = N = n, LOWER) packed: the strictly lower
// part column by column (entry (i, j), i > j, at j n - j (j + 1) / 2 + i - j - 1 of C,
// batch stride sC, plane stride psC) and the diagonal (entry i of Dg, batch stride sDg,
// plane stride psDg).
template <int OUT, bool LOWER>
__global__ __launch_bounds__(PTHREADS) void gemm_bf16p_kernel(int M, int N, int K, double alpha,
                                                              const __bf16 *X, int ldx, long psX, long sX,
                                                              const __bf16 *Y, int ldy, long psY, long sY,
                                                              double beta, void *Cv, int ldc, long psC, long sC,
                                                              int lin, int tri, __bf16 *Dg, long sDg, long psDg,
                                                              const __bf16 *X2, long psX2, long sX2,
                                                              const __bf16 *Y2, long psY2, long sY2) {
    // lin & 4: batch-major XCD order.  Workgroups go to the eight XCDs in turn (XCD g % 8
    // for the linear workgroup index g), and XCD c is given a run of consecutive
    // (batch entry, tile) pairs in batch-major order: whole batch entries per XCD, which
    // keeps each entry's operands in one L2 and gives every XCD the same work.
    // Within a batch entry the tiles then run longest first (their k ranges differ by up
    // to 8x): tri 3 by the distance of the tile from the diagonal, j - i, descending;
    // tri 4 (lower tiles) by column, descending.  The tiles of the last entries an XCD
    // runs are then short ones, not the longest.
    int gx = blockIdx.x, gy = blockIdx.y, bz = blockIdx.z;
    const bool batchOrder = (lin & 4) != 0;
    bool mapped = false;
    int mx = 0, my = 0;
    if (batchOrder) {
        const long tiles = (long) gridDim.x * gridDim.y, total = tiles * gridDim.z;
        const long g = gx + (long) gridDim.x * (gy + (long) gridDim.y * bz);
        const long c = g % 8, q = g / 8, base = total / 8, rem = total % 8;
        const long t = c * base + (c < rem ? c : rem) + q;
        bz = (int) (t / tiles);
        gy = (int) ((t % tiles) / gridDim.x);
        gx = (int) (t % gridDim.x);
        lin &= 3;
        int tt = (int) (t % tiles);
        if (tri == 3 && lin == 0 && gridDim.x == gridDim.y) {
            const int mt = gridDim.x;
            for (int d = mt - 1; d >= 0 && !mapped; --d) {  // upper tiles, j - i = d
                if (tt < mt - d) {
                    mx = tt;
                    my = tt + d;
                    mapped = true;
                } else {
                    tt -= mt - d;
                }
            }
            for (int d = 1; d < mt && !mapped; ++d) {  // lower tiles (return at once)
                if (tt < mt - d) {
                    mx = tt + d;
                    my = tt;
                    mapped = true;
                } else {
                    tt -= mt - d;
                }
            }
        } else if (tri == 4 && lin == 1 && gridDim.y == 1) {
            const int mt = (M + PT - 1) / PT;
            for (int j = mt - 1; j >= 0 && !mapped; --j) {  // column j: rows j .. mt - 1
                if (tt < mt - j) {
                    mx = j + tt;
                    my = j;
                    mapped = true;
                } else {
                    tt -= mt - j;
                }
            }
        }
    }
    int bx = gx, by = gy;
    if (!lin && !batchOrder && (tri == 2 || tri == 3)) bx = (int) ((gx + gy + bz) % gridDim.x);
    if (mapped) {
        bx = mx;
        by = my;
    } else if (lin == 2) {
        lower_tile_xcd(gx, gridDim.x, (M + PT - 1) / PT, bx, by);
    } else if (lin) {
        const int t = gx;
        int x = (int) ((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((x + 1) * (x + 2) / 2 <= t) ++x;
        while (x * (x + 1) / 2 > t) --x;
        bx = x;
        by = t - x * (x + 1) / 2;
    }
    const int i0 = bx * PT, j0 = by * PT;
    if (LOWER && i0 + PT - 1 < j0) return;
    // k range [kb, Ke) of the tile (Ke <= K, the padded length)
    int kb = 0, Ke = (tri == 1) ? min(K, j0 + PT) : (tri == 2) ? min(K, i0 + PT) : K;
    if (tri == 3) {
        if (i0 > j0 + PT - 1) return;  // the product is upper triangular: nothing below
        kb = i0;
        Ke = min(K, j0 + PT);
    } else if (tri == 4) {
        Ke = min(K, min(i0, j0) + PT);
    }
    const int nslab1 = (Ke - kb + PK - 1) / PK;
    const int nslab = X2 ? 2 * nslab1 : nslab1;
    X += (long) bz * sX;
    Y += (long) bz * sY;
    if (X2) {
        X2 += (long) bz * sX2;
        Y2 += (long) bz * sY2;
    }
    // planes [buffer][term][row][k]; the same memory stages the output tile at the end
    __shared__ __attribute__((aligned(16))) __bf16 smem[2 * PSLAB];
    auto As = reinterpret_cast<__bf16(*)[3][PT][PK + PPAD]>(smem);
    auto Bs = reinterpret_cast<__bf16(*)[3][PT][PK + PPAD]>(smem + PSLAB);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 3) * 32, wj = (wave >> 2) * 64;
    const int lr = lane & 15, lg = lane >> 4;
    // loader: thread -> row tid / 4 of the tile, k group (tid & 3) * 8, one 16-byte
    // vector per plane and operand
    const int lrow = tid >> 2, lk = (tid & 3) * 8;
    // LDS rows are 64 bytes with the 16-byte k group g of row r stored at g ^ ((r / 4) % 4):
    // a quad of lanes stores one row and sixteen lanes four consecutive rows (64 banks),
    // and the sixteen rows of a fragment read (rows 16 q + lane % 16) meet four distinct
    // groups in each bank quarter, so neither side conflicts
    const int lks = (((tid & 3) ^ ((lrow >> 2) & 3)) * 8), lgs = ((lg ^ ((lr >> 2) & 3)) * 8);
    const bool xrow = i0 + lrow < M, yrow = j0 + lrow < N;
    const __bf16 *xp = X + (long) (i0 + lrow) * ldx + lk;
    const __bf16 *yp = Y + (long) (j0 + lrow) * ldy + lk;
    // the second pair (same layout) supplies the slabs from nslab1 on
    const __bf16 *xp2 = X2 ? X2 + (long) (i0 + lrow) * ldx + lk : xp;
    const __bf16 *yp2 = X2 ? Y2 + (long) (j0 + lrow) * ldy + lk : yp;

    f32x4 acc[2][4], cor[2][4];
    double acc64[2][4][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
            cor[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 4; ++r) acc64[a][b][r] = 0.0;
        }

    u32x4 vx[3], vy[3];
    auto fetch = [&](int sl) {
        const bool seg2 = sl >= nslab1;
        const int k0 = kb + (seg2 ? sl - nslab1 : sl) * PK;
        const __bf16 *xs = seg2 ? xp2 : xp, *ys = seg2 ? yp2 : yp;
        const long pxs = seg2 ? psX2 : psX, pys = seg2 ? psY2 : psY;
        // k0 + lk .. k0 + lk + 7 lies inside [0, ld) whenever k0 + lk < Ke (Ke <= K, the
        // padded length, a multiple of 8)
        const bool kin = k0 + lk < Ke;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            vx[t] = (xrow && kin) ? *reinterpret_cast<const u32x4 *>(xs + t * pxs + k0) : (u32x4){0u, 0u, 0u, 0u};
            vy[t] = (yrow && kin) ? *reinterpret_cast<const u32x4 *>(ys + t * pys + k0) : (u32x4){0u, 0u, 0u, 0u};
        }
        if (tri == 3 && k0 + lk <= i0 + lrow) {
            // X holds a symmetric matrix A; the product takes its lower triangle with the
            // diagonal halved, X(k, i) -> 0 for k < i and X(i, i) / 2 (exact in each plane)
            const int i = i0 + lrow;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    unsigned u = vx[t][w];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int k = k0 + lk + 2 * w + h;
                        const unsigned bits = (u >> (16 * h)) & 0xffffu;
                        unsigned nb = bits;
                        if (k < i) nb = 0u;
                        else if (k == i) nb = (unsigned) __builtin_bit_cast(unsigned short,
                                                   (__bf16) (0.5f * __builtin_bit_cast(float, bits << 16)));
                        u = (u & ~(0xffffu << (16 * h))) | (nb << (16 * h));
                    }
                    vx[t][w] = u;
                }
            }
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            *reinterpret_cast<u32x4 *>(&As[buf][t][lrow][lks]) = vx[t];
            *reinterpret_cast<u32x4 *>(&Bs[buf][t][lrow][lks]) = vy[t];
        }
    };

    fetch(0);
    stash(0);
    __syncthreads();
    for (int sl = 0; sl < nslab; ++sl) {
        const int cur = sl & 1;
        const bool more = sl + 1 < nslab;
        if (more) fetch(sl + 1);
        bf16x8 bfr[3][4];
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                bfr[t][b] = *reinterpret_cast<const bf16x8 *>(&Bs[cur][t][wj + 16 * b + lr][lgs]);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            bf16x8 afr[3];
#pragma unroll
            for (int t = 0; t < 3; ++t)
                afr[t] = *reinterpret_cast<const bf16x8 *>(&As[cur][t][wi + 16 * a + lr][lgs]);
            // term-major over the four column blocks, so consecutive MFMAs accumulate into
            // different registers; correction terms smallest first
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[0], bfr[0][b], acc[a][b], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < 4; ++b)
                cor[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[2], bfr[0][b], cor[a][b], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < 4; ++b)
                cor[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[0], bfr[2][b], cor[a][b], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < 4; ++b)
                cor[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[1], bfr[1][b], cor[a][b], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < 4; ++b)
                cor[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[1], bfr[0][b], cor[a][b], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < 4; ++b)
                cor[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[0], bfr[1][b], cor[a][b], 0, 0, 0);
        }
        if ((sl + 1) % PFLUSH == 0 || !more) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
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
    // epilogue: lane holds rows i0 + wi + 16 a + 4 lg + r of column j0 + wj + 16 b + lr
    if constexpr (OUT == 3 || OUT == 2 || OUT == 0) {
        if (OUT != 0 || beta == 0.0) {
            float(*Ct)[PT + CPAD] = reinterpret_cast<float(*)[PT + CPAD]>(smem);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int jl = wj + 16 * b + lr;
#pragma unroll
                    for (int r = 0; r < 4; ++r) Ct[jl][wi + 16 * a + 4 * lg + r] = (float) (alpha * acc64[a][b][r]);
                }
            __syncthreads();
            const size_t esz = (OUT == 0) ? sizeof(float) : sizeof(__bf16);
            store_staged<OUT, LOWER, PTHREADS>(Ct, tid, i0, j0, M, N,
                                               static_cast<char *>(Cv) + (long) bz * sC * (long) esz, ldc, psC,
                                               Dg + (long) bz * sDg, psDg);
            return;
        }
    }
    if constexpr (OUT == 0 || OUT == 1) {
        typedef typename std::conditional<OUT == 1, double, float>::type Tout;
        Tout *C = static_cast<Tout *>(Cv) + (long) bz * sC;
        Tout cin[2][4][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = i0 + wi + 16 * a + 4 * lg + r, j = j0 + wj + 16 * b + lr;
                    const bool ok = i < M && j < N && (!LOWER || i >= j);
                    cin[a][b][r] = (ok && beta != 0.0) ? C[i + (long) j * ldc] : (Tout) 0;
                }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = i0 + wi + 16 * a + 4 * lg + r, j = j0 + wj + 16 * b + lr;
                    if (i < M && j < N && (!LOWER || i >= j))
                        C[i + (long) j * ldc] = (Tout) (alpha * acc64[a][b][r] + beta * (double) cin[a][b][r]);
                }
    }
}

template <int OUT>
void launch_p(hipStream_t s, bool lower, int M, int N, int K, double alpha, const __bf16 *X, int ldx, long psX,
              long sX, const __bf16 *Y, int ldy, long psY, long sY, double beta, void *C, int ldc, long psC, long sC,
              int batch, int tri, __bf16 *Dg = nullptr, long sDg = 0, long psDg = 0, const __bf16 *X2 = nullptr,
              long psX2 = 0, long sX2 = 0, const __bf16 *Y2 = nullptr, long psY2 = 0, long sY2 = 0) {
    const int mt = (M + PT - 1) / PT, nt = (N + PT - 1) / PT;
    const int lin = (lower && mt == nt) ? 1 : 0;
    dim3 grid(lin ? mt * (mt + 1) / 2 : mt, lin ? 1 : nt, batch);
    // lower tiles of a square grid in the XCD-grouped order (lin 2); the batched products
    // of the symmetric split (tri 3, 4), whose tiles differ in work, batch-major by XCD
    const int order = (tri == 3 || tri == 4) && batch > 1 ? 4 : 0;
    if (lower)
        gemm_bf16p_kernel<OUT, true><<<grid, PTHREADS, 0, s>>>(M, N, K, alpha, X, ldx, psX, sX, Y, ldy, psY, sY, beta,
                                                               C, ldc, psC, sC, (lin ? (order ? 1 : 2) : 0) | order,
                                                               tri, Dg, sDg, psDg, X2, psX2, sX2, Y2, psY2, sY2);
    else if constexpr (OUT != 3)
        gemm_bf16p_kernel<OUT, false><<<grid, PTHREADS, 0, s>>>(M, N, K, alpha, X, ldx, psX, sX, Y, ldy, psY, sY,
                                                                beta, C, ldc, psC, sC, order, tri, Dg, sDg, psDg, X2, psX2, sX2, Y2, psY2, sY2);
}

}  // namespace

void gemm_bf16p_packed(hipStream_t s, int n, int K, double alpha, const __bf16 *X, int ldx, long psX, long sX,
                       const __bf16 *Y, int ldy, long psY, long sY, __bf16 *off, long sOff, long psOff, __bf16 *dg,
                       long sDg, long psDg, int batch, int tri, const __bf16 *X2, long psX2, long sX2,
                       const __bf16 *Y2, long psY2, long sY2) {
    if (n <= 0 || batch <= 0) return;
    const bool aligned = K % 8 == 0 && ldx % 8 == 0 && ldy % 8 == 0 && psX % 8 == 0 && psY % 8 == 0 &&
                         sX % 8 == 0 && sY % 8 == 0 && ((uintptr_t) X % 16) == 0 && ((uintptr_t) Y % 16) == 0 &&
                         K <= ldx && K <= ldy;
    if (!aligned) {
        fprintf(stderr, "[hdsdp_amd] gemm_bf16p_packed: operand layout not 8-element aligned (K %d ldx %d ldy %d); "
                        "product skipped\n", K, ldx, ldy);
        return;
    }
    if (X2 && (psX2 % 8 != 0 || psY2 % 8 != 0 || sX2 % 8 != 0 || sY2 % 8 != 0 || ((uintptr_t) X2 % 16) != 0 ||
               ((uintptr_t) Y2 % 16) != 0)) {
        fprintf(stderr, "[hdsdp_amd] gemm_bf16p_packed: second operand pair not 8-element aligned; product skipped\n");
        return;
    }
    launch_p<3>(s, true, n, n, K, alpha, X, ldx, psX, sX, Y, ldy, psY, sY, 0.0, off, 0, psOff, sOff, batch, tri, dg,
                sDg, psDg, X2, psX2, sX2, Y2, psY2, sY2);
}

namespace {

// Three bf16 planes of a rows x cols FP64 (or FP32) matrix, rows rows..ld-1 of the planes
// zero (the k padding of a product operand).
template <typename T>
__global__ void split_matrix_kernel(int rows, int cols, const T *A, long lda, __bf16 *P, int ld, long ps) {
    const long total = (long) ld * cols;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < total; e += (long) gridDim.x * blockDim.x) {
        const int i = (int) (e % ld), j = (int) (e / ld);
        __bf16 p0 = (__bf16) 0.0f, p1 = p0, p2 = p0;
        if (i < rows) split3((float) A[i + (long) j * lda], p0, p1, p2);
        P[e] = p0;
        P[e + ps] = p1;
        P[e + 2 * ps] = p2;
    }
}

// Full n x n planes (ld ldn) of the packed dense coefficients slot0 .. slot0 + cnt - 1,
// one workgroup per 64 x 64 lower tile, the mirror written through LDS (both writes
// coalesced along the rows).
__global__ __launch_bounds__(256) void unpack_dense_split_kernel(int n, const double *dVal, int slot0, __bf16 *out,
                                                                 int ldn, long ps, long stride) {
    __shared__ float tile[64][65];  // tile[c][r] = A(i0 + r, j0 + c)
    const long npk = (long) n * (n + 1) / 2;
    const double *v = dVal + (long) (slot0 + blockIdx.y) * npk;
    __bf16 *o = out + (long) blockIdx.y * stride;
    const int t = blockIdx.x;
    int bi = (int) ((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    while (bi * (bi + 1) / 2 > t) --bi;
    const int bj = t - bi * (bi + 1) / 2;
    const int i0 = bi * 64, j0 = bj * 64, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int c = ty; c < 64; c += 4) {
        const long i = i0 + tx, j = j0 + c;
        if (i < n && j < n && i >= j) {
            const float val = (float) v[j * n - j * (j + 1) / 2 + i];
            __bf16 p0, p1, p2;
            split3(val, p0, p1, p2);
            const long e = i + j * ldn;
            o[e] = p0;
            o[e + ps] = p1;
            o[e + 2 * ps] = p2;
            tile[c][tx] = val;
        }
    }
    __syncthreads();
    for (int c = ty; c < 64; c += 4) {
        // mirror entry (j0 + tx, i0 + c) = A(i0 + c, j0 + tx) = tile[tx][c], above the diagonal only
        const long r = j0 + tx, q = i0 + c;
        if (r < n && q < n && q > r) {
            __bf16 p0, p1, p2;
            split3(tile[tx][c], p0, p1, p2);
            const long e = r + q * ldn;
            o[e] = p0;
            o[e + ps] = p1;
            o[e + 2 * ps] = p2;
        }
    }
}

}  // namespace

void gemm_bf16p(hipStream_t s, bool lower, int M, int N, int K, double alpha, const __bf16 *X, int ldx, long psX,
                long sX, const __bf16 *Y, int ldy, long psY, long sY, double beta, void *C, int outKind, int ldc,
                long psC, long sC, int batch, int tri) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    // every operand load is one aligned 16-byte vector inside [0, K) of a column
    const bool aligned = K % 8 == 0 && ldx % 8 == 0 && ldy % 8 == 0 && psX % 8 == 0 && psY % 8 == 0 &&
                         sX % 8 == 0 && sY % 8 == 0 && ((uintptr_t) X % 16) == 0 && ((uintptr_t) Y % 16) == 0 &&
                         K <= ldx && K <= ldy;
    if (!aligned) {
        fprintf(stderr, "[hdsdp_amd] gemm_bf16p: operand layout not 8-element aligned (K %d ldx %d ldy %d); "
                        "product skipped\n", K, ldx, ldy);
        return;
    }
    if (outKind == BF16P_OUT_F64)
        launch_p<1>(s, lower, M, N, K, alpha, X, ldx, psX, sX, Y, ldy, psY, sY, beta, C, ldc, psC, sC, batch, tri);
    else if (outKind == BF16P_OUT_F32)
        launch_p<0>(s, lower, M, N, K, alpha, X, ldx, psX, sX, Y, ldy, psY, sY, beta, C, ldc, psC, sC, batch, tri);
    else
        launch_p<2>(s, lower, M, N, K, alpha, X, ldx, psX, sX, Y, ldy, psY, sY, beta, C, ldc, psC, sC, batch, tri);
}

void split_matrix(hipStream_t s, int rows, int cols, const double *A, long lda, __bf16 *P, int ld, long ps) {
    const long total = (long) ld * cols;
    const int blocks = (int) std::min<long>(4096, (total + 255) / 256);
    if (blocks > 0) split_matrix_kernel<double><<<blocks, 256, 0, s>>>(rows, cols, A, lda, P, ld, ps);
}

void split_matrix_f(hipStream_t s, int rows, int cols, const float *A, long lda, __bf16 *P, int ld, long ps) {
    const long total = (long) ld * cols;
    const int blocks = (int) std::min<long>(4096, (total + 255) / 256);
    if (blocks > 0) split_matrix_kernel<float><<<blocks, 256, 0, s>>>(rows, cols, A, lda, P, ld, ps);
}

void unpack_dense_split(hipStream_t s, const DevCoeffs &c, int slot0, int cnt, __bf16 *out, int ldn, long ps,
                        long stride) {
    if (cnt <= 0) return;
    const int nt = (c.n + 63) / 64;
    unpack_dense_split_kernel<<<dim3(nt * (nt + 1) / 2, cnt), 256, 0, s>>>(c.n, c.dVal, slot0, out, ldn, ps, stride);
}


}  // namespace hdk
