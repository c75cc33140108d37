// Double centering for local PCA (wot/ot/util.py:240-255 compute_pca): the
// cells of both days are stacked (x: n cells x g genes), centered per gene
// (x - mean over cells, util.py:244-246) and PCA is fitted on x^T, which
// sklearn centers once more per column of x^T, i.e. per cell over genes.
// The matrix whose singular vectors the PCA needs is therefore
//   A[c][j] = x[c][j] - mu_j - nu_c,  mu_j = mean_c x[c][j],
//   nu_c = mean_j (x[c][j] - mu_j) = mean_j x[c][j] - mean_j mu_j.
// Two streaming passes: partial column and row sums per (row block, column
// block) tile (fixed order: deterministic), then A written in one pass.  HBM
// bound: 2 reads of x + 1 write of A.  Rows 16-byte aligned take 16-byte
// loads and stores throughout.
#include "common.h"

namespace wotk {
namespace {

constexpr int kTileRows = 64;
constexpr int kTileCols = 256;

// partial sums of one 64 x 256 tile: colpart[rb][j] (over the tile's rows),
// rowpart[cb][c] (over the tile's columns)
template <typename T>
__global__ __launch_bounds__(256) void center_sums_kernel(const T *__restrict__ x, int64_t ldx, int64_t n,
                                                          int64_t g, double *__restrict__ colpart,
                                                          double *__restrict__ rowpart) {
    __shared__ double rs[kTileRows][17];
    const int rb = blockIdx.y, cb = blockIdx.x;
    const int64_t r0 = (int64_t)rb * kTileRows, c0 = (int64_t)cb * kTileCols;
    const int t = threadIdx.x;
    const int64_t j = c0 + t;
    // column sums: thread t owns column c0 + t, walks the 64 rows (coalesced per row)
    double cs = 0.0;
    for (int r = 0; r < kTileRows; ++r) {
        const int64_t row = r0 + r;
        if (row < n && j < g) cs += (double)x[row * ldx + j];
    }
    if (j < g) colpart[(int64_t)rb * g + j] = cs;
    // row sums: 16 threads per row (columns t%16, t%16 + 16, ...), 16 rows per pass
    const int lane = t & 15, rr = t >> 4;
    for (int base = 0; base < kTileRows; base += 16) {
        const int64_t row = r0 + base + rr;
        double s = 0.0;
        if (row < n)
            for (int c = lane; c < kTileCols; c += 16) {
                const int64_t col = c0 + c;
                if (col < g) s += (double)x[row * ldx + col];
            }
        rs[base + rr][lane] = s;
    }
    __syncthreads();
    if (t < kTileRows) {
        const int64_t row = r0 + t;
        double s = 0.0;
        for (int l = 0; l < 16; ++l) s += rs[t][l];
        if (row < n) rowpart[(int64_t)cb * n + row] = s;
    }
}

// Streaming form of the partial sums when rows are 16-byte aligned: each wave
// reads 32 rows of a 512-column chunk with 16-byte loads (2 KB per row for
// FP32, the wave's lanes side by side), keeps its columns' sums in registers
// and reduces each row over the wave (fixed xor tree).  One read of x;
// colpart[wrb][j] over the wave's rows, rowpart[cb][c] over the chunk's
// columns.
constexpr int kStreamRows = 32;    // rows per wave
constexpr int kStreamChunk = 512;  // columns per chunk

template <typename T>
__global__ __launch_bounds__(256) void center_stream_kernel(const T *__restrict__ x, int64_t ldx, int64_t n,
                                                            int64_t g, int n_wrb, double *__restrict__ colpart,
                                                            double *__restrict__ rowpart) {
    constexpr int VEC = 16 / (int)sizeof(T);
    constexpr int NV = kStreamChunk / (64 * VEC);
    constexpr int B = 2;  // rows per batch; the next batch's loads are in flight while one is summed
    typedef T V __attribute__((ext_vector_type(VEC)));
    const int lane = threadIdx.x & 63;
    const int wrb = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (wrb >= n_wrb) return;  // whole waves; no barriers below
    const int cb = blockIdx.x;
    const int64_t r0 = (int64_t)wrb * kStreamRows;
    const int64_t c0 = (int64_t)cb * kStreamChunk;
    double cs[NV][VEC];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int e = 0; e < VEC; ++e) cs[v][e] = 0.0;
    const int rows = (int)min<int64_t>(kStreamRows, n - r0);
    auto load_batch = [&](V (&bt)[B][NV], int rb) {
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const bool ok = rb + i < rows;
            const T *xr = x + (r0 + (ok ? rb + i : 0)) * ldx;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int64_t col = c0 + (int64_t)(v * 64 + lane) * VEC;
                if (ok && col + VEC <= g) {
                    bt[i][v] = *reinterpret_cast<const V *>(xr + col);
                } else {
#pragma unroll
                    for (int e = 0; e < VEC; ++e) bt[i][v][e] = ok && col + e < g ? xr[col + e] : T(0);
                }
            }
        }
    };
    // row sums over the wave, 8 rows at a time: each lane keeps its partial of
    // every row of the group; a butterfly halves the rows a lane holds at
    // xor 32, 16 and 8 (exchanging the other half), then sums over xor 4, 2, 1
    // -- 10 exchanges for 8 rows; row r of the group ends in lane 8 r (fixed order)
    double rp[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) rp[i] = 0.0;
    auto flush = [&](int g0) {  // rows g0 .. g0 + 7 of the wave (wave-uniform call)
        const bool h1 = lane & 32, h2 = lane & 16, h3 = lane & 8;
        double v4[4], v2[2];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double keep = h1 ? rp[4 + i] : rp[i], send = h1 ? rp[i] : rp[4 + i];
            v4[i] = keep + __shfl_xor(send, 32);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const double keep = h2 ? v4[2 + i] : v4[i], send = h2 ? v4[i] : v4[2 + i];
            v2[i] = keep + __shfl_xor(send, 16);
        }
        double x = (h3 ? v2[1] : v2[0]) + __shfl_xor(h3 ? v2[0] : v2[1], 8);
#pragma unroll
        for (int off = 4; off > 0; off >>= 1) x += __shfl_xor(x, off);
        const int r = lane >> 3;
        if ((lane & 7) == 0 && g0 + r < rows) rowpart[(int64_t)cb * n + r0 + g0 + r] = x;
#pragma unroll
        for (int i = 0; i < 8; ++i) rp[i] = 0.0;
    };
    auto sum_batch = [&](const V (&bt)[B][NV], int rb) {
#pragma unroll
        for (int i = 0; i < B; ++i) {
            if (rb + i >= rows) break;  // wave-uniform
            double rsum = 0.0;
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    cs[v][e] += (double)bt[i][v][e];
                    rsum += (double)bt[i][v][e];
                }
            rp[(rb + i) & 7] = rsum;
        }
        if (((rb + B) & 7) == 0 || rb + B >= rows) flush((rb + B - 1) & ~7);
    };
    V b0[B][NV], b1[B][NV];
    load_batch(b0, 0);
    // fully unrolled over the wave's rows, so every row's slot in rp is a constant
#pragma unroll
    for (int rb = 0; rb < kStreamRows; rb += 2 * B) {
        if (rb >= rows) break;  // wave-uniform
        if (rb + B < rows) load_batch(b1, rb + B);
        sum_batch(b0, rb);
        if (rb + B >= rows) break;
        if (rb + 2 * B < rows) load_batch(b0, rb + 2 * B);
        sum_batch(b1, rb + B);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            const int64_t col = c0 + (int64_t)(v * 64 + lane) * VEC + e;
            if (col < g) colpart[(int64_t)wrb * g + col] = cs[v][e];
        }
}

// mu_j = sum_rb colpart[rb][j] / n: 64 columns per workgroup, 4 threads per
// column each summing every 4th row block, combined in a fixed order
__global__ __launch_bounds__(256) void center_colmean_kernel(const double *__restrict__ colpart, int n_rb, int64_t n,
                                                             int64_t g, double *__restrict__ mu) {
    __shared__ double red[4][64];
    const int c = threadIdx.x & 63, part = threadIdx.x >> 6;
    const int64_t j = (int64_t)blockIdx.x * 64 + c;
    double s = 0.0;
    if (j < g)
        for (int rb = part; rb < n_rb; rb += 4) s += colpart[(int64_t)rb * g + j];
    red[part][c] = s;
    __syncthreads();
    if (part == 0 && j < g) mu[j] = (((red[0][c] + red[1][c]) + red[2][c]) + red[3][c]) / (double)n;
}

// Column means from many partial rows in two fixed-order levels:
// colseg[s][j] = sum of colpart[rb][j] over rb in segment s (64 row blocks;
// 64 columns x 4 parts per workgroup), then mu_j = sum_s colseg[s][j] / n.
constexpr int kColSeg = 64;

__global__ __launch_bounds__(256) void center_colseg_kernel(const double *__restrict__ colpart, int n_rb, int64_t g,
                                                            double *__restrict__ colseg) {
    __shared__ double red[4][64];
    const int c = threadIdx.x & 63, part = threadIdx.x >> 6;
    const int64_t j = (int64_t)blockIdx.x * 64 + c;
    const int rb0 = blockIdx.y * kColSeg, rb1 = min(n_rb, rb0 + kColSeg);
    double s = 0.0;
    if (j < g)
        for (int rb = rb0 + part; rb < rb1; rb += 4) s += colpart[(int64_t)rb * g + j];
    red[part][c] = s;
    __syncthreads();
    if (part == 0 && j < g) colseg[(int64_t)blockIdx.y * g + j] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
}

__global__ __launch_bounds__(256) void center_colfinal_kernel(const double *__restrict__ colseg, int n_seg, int64_t n,
                                                              int64_t g, double *__restrict__ mu) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= g) return;
    double s = 0.0;
    for (int k = 0; k < n_seg; ++k) s += colseg[(int64_t)k * g + j];
    mu[j] = s / (double)n;
}

// mean of mu (one workgroup, fixed order)
__global__ __launch_bounds__(256) void center_mumean_kernel(const double *__restrict__ mu, int64_t g,
                                                            double *__restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t j = threadIdx.x; j < g; j += 256) s += mu[j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / (double)g;
}

// nu_c = sum_cb rowpart[cb][c] / g - mean(mu)
__global__ __launch_bounds__(256) void center_rowmean_kernel(const double *__restrict__ rowpart, int n_cb, int64_t n,
                                                             int64_t g, const double *__restrict__ mumean,
                                                             double *__restrict__ nu) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    double s = 0.0;
    for (int cb = 0; cb < n_cb; ++cb) s += rowpart[(int64_t)cb * n + c];
    nu[c] = s / (double)g - mumean[0];
}

// A[c][j] = x[c][j] - mu_j - nu_c
template <typename T>
__global__ __launch_bounds__(256) void center_apply_kernel(const T *__restrict__ x, int64_t ldx, int64_t n, int64_t g,
                                                           const double *__restrict__ mu,
                                                           const double *__restrict__ nu, double *__restrict__ A,
                                                           int64_t lda) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * kTileRows;
    if (j >= g) return;
    const double m = mu[j];
    for (int r = 0; r < kTileRows; ++r) {
        const int64_t row = r0 + r;
        if (row >= n) break;
        A[row * lda + j] = ((double)x[row * ldx + j] - m) - nu[row];
    }
}

// the same with 2 columns per thread and 16 rows per workgroup: 8- or 16-byte
// loads of x and 16-byte stores of A, contiguous across each wave (x rows
// 8-byte aligned for FP32 / 16-byte for FP64, A rows 16-byte aligned, g even)
constexpr int kApplyRows = 16;

template <typename T>
__global__ __launch_bounds__(256) void center_apply2_kernel(const T *__restrict__ x, int64_t ldx, int64_t n,
                                                            int64_t g, const double *__restrict__ mu,
                                                            const double *__restrict__ nu, double *__restrict__ A,
                                                            int64_t lda) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    typedef T V2 __attribute__((ext_vector_type(2)));
    const int64_t j = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 2;
    if (j >= g) return;
    const d2 m = *reinterpret_cast<const d2 *>(mu + j);
    // row blocks strided over the grid's y extent (at most 65535 blocks)
    for (int64_t r0 = (int64_t)blockIdx.y * kApplyRows; r0 < n; r0 += (int64_t)gridDim.y * kApplyRows) {
        const int rows = (int)min<int64_t>(kApplyRows, n - r0);
#pragma unroll 4
        for (int r = 0; r < rows; ++r) {
            const int64_t row = r0 + r;
            const V2 v = *reinterpret_cast<const V2 *>(x + row * ldx + j);
            const double c = nu[row];
            d2 o;
            o.x = ((double)v.x - m.x) - c;
            o.y = ((double)v.y - m.y) - c;
            *reinterpret_cast<d2 *>(A + row * lda + j) = o;
        }
    }
}

struct CenterLayout {
    double *colpart, *rowpart, *mumean, *colseg;
};

CenterLayout center_carve(Arena &ar, int64_t n, int64_t g) {
    // room for the tiled pass (64 x 256 tiles) and the streaming pass (32-row
    // waves, 512-column chunks), whichever runs
    const int64_t n_rb = std::max((n + kTileRows - 1) / kTileRows, (n + kStreamRows - 1) / kStreamRows);
    const int64_t n_cb = std::max((g + kTileCols - 1) / kTileCols, (g + kStreamChunk - 1) / kStreamChunk);
    CenterLayout L;
    L.colpart = ar.take<double>((size_t)(n_rb * g));
    L.rowpart = ar.take<double>((size_t)(n_cb * n));
    L.mumean = ar.take<double>(1);
    L.colseg = ar.take<double>((size_t)(((n_rb + kColSeg - 1) / kColSeg) * g));
    return L;
}

static bool aligned16(const void *p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

template <typename T>
int center_impl(const T *x, int64_t ldx, int64_t n, int64_t g, double *A, int64_t lda, double *mu, double *nu,
                const CenterLayout &L, hipStream_t st) {
    const bool xvec = aligned16(x) && (ldx * (int64_t)sizeof(T)) % 16 == 0;
    int n_rb, n_cb;
    if (xvec) {  // one streaming read of x
        n_rb = (int)((n + kStreamRows - 1) / kStreamRows);
        n_cb = (int)((g + kStreamChunk - 1) / kStreamChunk);
        center_stream_kernel<T><<<dim3(n_cb, (n_rb + 3) / 4), 256, 0, st>>>(x, ldx, n, g, n_rb, L.colpart,
                                                                           L.rowpart);
    } else {
        n_rb = (int)((n + kTileRows - 1) / kTileRows);
        n_cb = (int)((g + kTileCols - 1) / kTileCols);
        center_sums_kernel<T><<<dim3(n_cb, n_rb), 256, 0, st>>>(x, ldx, n, g, L.colpart, L.rowpart);
    }
    WOTK_LAUNCH_CHECK();
    if (n_rb > kColSeg) {
        const int n_seg = (n_rb + kColSeg - 1) / kColSeg;
        center_colseg_kernel<<<dim3((unsigned)((g + 63) / 64), n_seg), 256, 0, st>>>(L.colpart, n_rb, g, L.colseg);
        WOTK_LAUNCH_CHECK();
        center_colfinal_kernel<<<(unsigned)((g + 255) / 256), 256, 0, st>>>(L.colseg, n_seg, n, g, mu);
    } else {
        center_colmean_kernel<<<(unsigned)((g + 63) / 64), 256, 0, st>>>(L.colpart, n_rb, n, g, mu);
    }
    WOTK_LAUNCH_CHECK();
    center_mumean_kernel<<<1, 256, 0, st>>>(mu, g, L.mumean);
    WOTK_LAUNCH_CHECK();
    center_rowmean_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(L.rowpart, n_cb, n, g, L.mumean, nu);
    WOTK_LAUNCH_CHECK();
    if (A) {
        const unsigned n_tr = (unsigned)((n + kTileRows - 1) / kTileRows);
        const bool x2 = reinterpret_cast<uintptr_t>(x) % (2 * sizeof(T)) == 0 && ldx % 2 == 0;
        if (x2 && g % 2 == 0 && aligned16(A) && lda % 2 == 0 && aligned16(mu))
            center_apply2_kernel<T><<<dim3((unsigned)((g / 2 + 255) / 256),
                                           (unsigned)std::min<int64_t>(65535, (n + kApplyRows - 1) / kApplyRows)),
                                      256, 0, st>>>(x, ldx, n, g, mu, nu, A, lda);
        else
            center_apply_kernel<T><<<dim3((unsigned)((g + 255) / 256), n_tr), 256, 0, st>>>(x, ldx, n, g, mu, nu,
                                                                                         A, lda);
        WOTK_LAUNCH_CHECK();
    }
    return WOTK_OK;
}

}  // namespace
}  // namespace wotk

using namespace wotk;

extern "C" size_t wotk_pca_center_workspace_size(int64_t n, int64_t g) {
    if (n <= 0 || g <= 0) return 0;
    Arena ar(nullptr, 0);
    center_carve(ar, n, g);
    return ar.off;
}

extern "C" int wotk_pca_center(const void *x, int32_t dtype, int64_t ldx, int64_t n, int64_t g, double *A,
                               int64_t lda, double *gene_mean, double *cell_mean, void *workspace,
                               size_t workspace_bytes, void *stream) {
    WOTK_REQUIRE(x && n > 0 && g > 0 && ldx >= g, "pca_center: bad matrix (%lld x %lld, ld %lld)", (long long)n,
                 (long long)g, (long long)ldx);
    WOTK_REQUIRE(dtype == WOTK_DTYPE_F64 || dtype == WOTK_DTYPE_F32, "pca_center: bad dtype %d", dtype);
    WOTK_REQUIRE(gene_mean && cell_mean, "pca_center: gene_mean and cell_mean are required");
    WOTK_REQUIRE(!A || lda >= g, "pca_center: lda %lld < %lld", (long long)lda, (long long)g);
    WOTK_REQUIRE((n + kTileRows - 1) / kTileRows < 65536,
                 "pca_center: too many cells (%lld)", (long long)n);
    Arena ar(workspace, workspace_bytes);
    CenterLayout L = center_carve(ar, n, g);
    WOTK_REQUIRE(ar.ok && workspace, "pca_center: workspace too small (%zu < %zu bytes)", workspace_bytes, ar.off);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == WOTK_DTYPE_F64)
        return center_impl(static_cast<const double *>(x), ldx, n, g, A, lda, gene_mean, cell_mean, L, st);
    return center_impl(static_cast<const float *>(x), ldx, n, g, A, lda, gene_mean, cell_mean, L, st);
}
