// Cost matrix for wot's OT path: C = sqeuclid(X, Y) / median(C)
// (wot/ot/ot_model.py:242-253).  FP64 direct-difference tiles, an exact
// radix-select median over the FP64 bit patterns, and a normalise+transpose
// pass that writes the FP32 pair (C, C^T) the Sinkhorn sweeps stream.
#include "common.h"

namespace wotk {

// ---------------------------------------------------------------------------
// FP64 direct difference: C_ij = sum_k (x_ik - y_jk)^2, 64x64 output tile per
// 256-thread block, 4x4 outputs per thread, coordinates staged through LDS in
// 32-wide k chunks (transposed so a thread's 4 rows are contiguous).
// ---------------------------------------------------------------------------
constexpr int kCostTile = 64;
constexpr int kCostKc = 32;

__global__ __launch_bounds__(256) void sqeuclid_f64_kernel(const double *__restrict__ X,
                                                           const double *__restrict__ Y, int n_i,
                                                           int n_j, int d, double *__restrict__ C,
                                                           int64_t ldc) {
    __shared__ double xs[kCostKc][kCostTile];
    __shared__ double ys[kCostKc][kCostTile];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int i0 = blockIdx.y * kCostTile, j0 = blockIdx.x * kCostTile;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;

    for (int k0 = 0; k0 < d; k0 += kCostKc) {
        for (int e = tid; e < kCostTile * kCostKc; e += 256) {
            const int r = e / kCostKc, kk = e % kCostKc;
            const bool kin = (k0 + kk) < d;
            xs[kk][r] = (kin && i0 + r < n_i) ? X[(int64_t)(i0 + r) * d + k0 + kk] : 0.0;
            ys[kk][r] = (kin && j0 + r < n_j) ? Y[(int64_t)(j0 + r) * d + k0 + kk] : 0.0;
        }
        __syncthreads();
        const int kmax = min(kCostKc, d - k0);
        for (int kk = 0; kk < kmax; ++kk) {
            double xv[4], yv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                xv[q] = xs[kk][ty * 4 + q];
                yv[q] = ys[kk][tx * 4 + q];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const double t = xv[a] - yv[b];
                    acc[a][b] = fma(t, t, acc[a][b]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int i = i0 + ty * 4 + a;
        if (i >= n_i) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = j0 + tx * 4 + b;
            if (j < n_j) C[(int64_t)i * ldc + j] = acc[a][b];
        }
    }
}

// ---------------------------------------------------------------------------
// Exact order statistics by MSB-first radix select on the bit patterns of
// non-negative doubles (for x >= +0 the uint64 order is the value order).
// Six digit passes (11,11,11,11,10,10 bits), each a histogram of the keys that
// match the prefix found so far; a one-block scan picks the digit.  All state
// stays on the device so the passes chain without host round trips.
// ---------------------------------------------------------------------------
struct SelectState {
    unsigned long long prefix;
    unsigned long long mask;
    long long rank;            // rank still to skip inside the current prefix
    unsigned long long eq;     // after the last pass: count of keys == prefix
    unsigned long long next;   // smallest key > prefix (for the upper median)
    double median;
    unsigned long long lo_key, hi_key;  // the two middle values (equal for odd n)
};

constexpr int kHistBins = 2048;

__global__ __launch_bounds__(256) void radix_hist_kernel(const unsigned long long *__restrict__ keys,
                                                         int64_t n, int shift, int width,
                                                         const SelectState *__restrict__ st,
                                                         unsigned long long *__restrict__ ghist) {
    __shared__ unsigned int h[kHistBins];
    for (int b = threadIdx.x; b < kHistBins; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    const unsigned long long prefix = st->prefix, mask = st->mask;
    const unsigned long long dmask = (1ull << width) - 1ull;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const unsigned long long k = keys[e];
        if ((k & mask) == prefix) atomicAdd(&h[(k >> shift) & dmask], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kHistBins; b += blockDim.x)
        if (h[b]) atomicAdd(&ghist[b], (unsigned long long)h[b]);
}

// One block: exclusive scan of the histogram, pick the digit holding `rank`,
// extend the prefix and clear the histogram for the next pass.
__global__ __launch_bounds__(256) void radix_scan_kernel(unsigned long long *__restrict__ ghist,
                                                         int shift, int width, int last,
                                                         SelectState *__restrict__ st) {
    __shared__ unsigned long long part[256];
    const int nb = 1 << width;
    const int per = (nb + 255) / 256;
    const int t = threadIdx.x;
    unsigned long long mine = 0;
    for (int b = t * per; b < min(nb, (t + 1) * per); ++b) mine += ghist[b];
    part[t] = mine;
    __syncthreads();
    if (t == 0) {
        long long rank = st->rank;
        unsigned long long cum = 0;
        int tsel = 0;
        for (; tsel < 256; ++tsel) {
            if (cum + part[tsel] > (unsigned long long)rank) break;
            cum += part[tsel];
        }
        int bsel = min(tsel, 255) * per;
        if (tsel == 256) cum -= part[255];  // rank beyond the total: stay in the last bin
        for (; bsel < nb - 1; ++bsel) {
            if (cum + ghist[bsel] > (unsigned long long)rank) break;
            cum += ghist[bsel];
        }
        st->prefix |= (unsigned long long)bsel << shift;
        st->mask |= ((1ull << width) - 1ull) << shift;
        st->rank = rank - (long long)cum;
        if (last) st->eq = ghist[bsel];
    }
    __syncthreads();
    for (int b = t; b < nb; b += blockDim.x) ghist[b] = 0ull;
}

__global__ __launch_bounds__(256) void next_above_kernel(const unsigned long long *__restrict__ keys,
                                                         int64_t n, SelectState *__restrict__ st) {
    const unsigned long long v = st->prefix;
    unsigned long long best = ~0ull;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const unsigned long long k = keys[e];
        if (k > v && k < best) best = k;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off);
        best = o < best ? o : best;
    }
    if ((threadIdx.x & 63) == 0 && best != ~0ull) atomicMin(&st->next, best);
}

__global__ void median_init_kernel(SelectState *st, long long rank, unsigned long long *hist) {
    for (int b = threadIdx.x; b < kHistBins; b += blockDim.x) hist[b] = 0ull;
    if (threadIdx.x != 0) return;
    st->prefix = 0ull;
    st->mask = 0ull;
    st->rank = rank;
    st->eq = 0ull;
    st->next = ~0ull;
    st->median = 0.0;
}

__global__ void median_finish_kernel(SelectState *st, int odd, double *out) {
    const double lo = __longlong_as_double((long long)st->prefix);
    double m = lo;
    if (!odd) {
        // the upper middle value equals the lower one if enough ties remain
        const bool tie = (unsigned long long)(st->rank + 1) < st->eq;
        const double hi = tie ? lo : __longlong_as_double((long long)st->next);
        m = (lo + hi) / 2.0;
    }
    st->median = m;
    st->lo_key = st->prefix;
    st->hi_key = odd ? st->prefix
                     : (((unsigned long long)(st->rank + 1) < st->eq) ? st->prefix : st->next);
    if (out) *out = m;
}

// Flat indices of (one of) the elements holding the lower / upper middle
// value: the smallest index with that key.  These carry the median's
// subgradient (weight 1/2 each for an even count, 1 for an odd count).
__global__ void median_index_init_kernel(unsigned long long *idx) {
    idx[0] = ~0ull;
    idx[1] = ~0ull;
}

__global__ __launch_bounds__(256) void median_index_kernel(const unsigned long long *__restrict__ keys,
                                                           int64_t n, const SelectState *__restrict__ st,
                                                           unsigned long long *__restrict__ idx) {
    const unsigned long long lo = st->lo_key, hi = st->hi_key;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const unsigned long long k = keys[e];
        if (k == lo) atomicMin(&idx[0], (unsigned long long)e);
        if (k == hi) atomicMin(&idx[1], (unsigned long long)e);
    }
}

static const int kDigitShift[6] = {53, 42, 31, 20, 10, 0};
static const int kDigitWidth[6] = {11, 11, 11, 11, 10, 10};

static size_t median_ws(Arena &ar, int64_t n, SelectState **st, unsigned long long **hist) {
    (void)n;
    *st = ar.take<SelectState>(1);
    *hist = ar.take<unsigned long long>(kHistBins);
    return ar.off;
}

static int median_run(const double *values, int64_t n, double *out, int64_t *index_out,
                      SelectState *st, unsigned long long *hist, hipStream_t s) {
    WOTK_REQUIRE(n > 0, "median of an empty array");
    const long long rank = (n % 2) ? n / 2 : n / 2 - 1;
    median_init_kernel<<<1, 256, 0, s>>>(st, rank, hist);
    WOTK_LAUNCH_CHECK();
    const auto *keys = reinterpret_cast<const unsigned long long *>(values);
    const int64_t want = (n + 255) / 256;
    const int grid = (int)std::min<int64_t>(std::max<int64_t>(want, 1), 4096);
    for (int p = 0; p < 6; ++p) {
        radix_hist_kernel<<<grid, 256, 0, s>>>(keys, n, kDigitShift[p], kDigitWidth[p], st, hist);
        WOTK_LAUNCH_CHECK();
        radix_scan_kernel<<<1, 256, 0, s>>>(hist, kDigitShift[p], kDigitWidth[p], p == 5, st);
        WOTK_LAUNCH_CHECK();
    }
    if (n % 2 == 0) {
        next_above_kernel<<<grid, 256, 0, s>>>(keys, n, st);
        WOTK_LAUNCH_CHECK();
    }
    median_finish_kernel<<<1, 1, 0, s>>>(st, (int)(n % 2), out);
    WOTK_LAUNCH_CHECK();
    if (index_out) {
        auto *idx = reinterpret_cast<unsigned long long *>(index_out);
        median_index_init_kernel<<<1, 1, 0, s>>>(idx);
        median_index_kernel<<<grid, 256, 0, s>>>(keys, n, st, idx);
        WOTK_LAUNCH_CHECK();
    }
    return WOTK_OK;
}

// ---------------------------------------------------------------------------
// FP64 (I x J) -> FP32 C (I x J) and C^T (J x I), each entry value/divisor
// (divisor read from the device, e.g. the median) or value*scale.  Padding
// columns up to the next multiple of 4 are written as 0 so the float4 sweeps
// read finite values there.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void normalise_transpose_kernel(
    const double *__restrict__ C64, int64_t ld64, int n_i, int n_j, const double *__restrict__ divisor,
    double scale, float *__restrict__ C, int64_t ldc, float *__restrict__ CT, int64_t ldct) {
    __shared__ float tile[64][65];
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    const double dv = divisor ? *divisor : 0.0;
    const int pad_j = (n_j + 3) & ~3, pad_i = (n_i + 3) & ~3;
    for (int r = ty; r < 64; r += 4) {
        const int i = i0 + r, j = j0 + tx;
        float v = 0.f;
        if (i < n_i && j < n_j) {
            const double c = C64[(int64_t)i * ld64 + j];
            v = (float)(divisor ? c / dv : c * scale);
        }
        tile[r][tx] = v;
        if (i < n_i && j < pad_j) C[(int64_t)i * ldc + j] = v;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int j = j0 + r, i = i0 + tx;
        if (j < n_j && i < pad_i) CT[(int64_t)j * ldct + i] = tile[tx][r];
    }
}

int launch_normalise_transpose(const double *C64, int64_t ld64, int64_t I, int64_t J,
                               const double *divisor, double scale, float *C, int64_t ldc,
                               float *CT, int64_t ldct, hipStream_t s) {
    dim3 grid((unsigned)((round_up(J, 4) + 63) / 64), (unsigned)((round_up(I, 4) + 63) / 64));
    normalise_transpose_kernel<<<grid, 256, 0, s>>>(C64, ld64, (int)I, (int)J, divisor, scale, C,
                                                    ldc, CT, ldct);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

int launch_sqeuclid_f64(const double *X, const double *Y, int64_t I, int64_t J, int64_t d,
                        double *C64, int64_t ld, hipStream_t s) {
    dim3 grid((unsigned)((J + kCostTile - 1) / kCostTile), (unsigned)((I + kCostTile - 1) / kCostTile));
    sqeuclid_f64_kernel<<<grid, 256, 0, s>>>(X, Y, (int)I, (int)J, (int)d, C64, ld);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

int launch_cost_bf16(const double *X, const double *Y, int64_t I, int64_t J, int64_t d, int terms,
                     double *C64, int64_t ld, void *scratch, hipStream_t s);
size_t cost_bf16_scratch(int64_t I, int64_t J, int64_t d);

static int check_cost_layout(int64_t I, int64_t J, int64_t d, const float *C, int64_t ldc,
                             const float *CT, int64_t ldct) {
    WOTK_REQUIRE(I > 0 && J > 0, "empty cost matrix (%lld x %lld)", (long long)I, (long long)J);
    WOTK_REQUIRE(I < (1ll << 31) && J < (1ll << 31), "dimension too large");
    WOTK_REQUIRE(d > 0, "coordinate dimension must be positive");
    WOTK_REQUIRE(ldc >= round_up(J, 4) && ldc % 4 == 0, "ldc must be >= J rounded up to 4 and a multiple of 4");
    WOTK_REQUIRE(ldct >= round_up(I, 4) && ldct % 4 == 0, "ldct must be >= I rounded up to 4 and a multiple of 4");
    WOTK_REQUIRE(((uintptr_t)C & 15) == 0 && ((uintptr_t)CT & 15) == 0, "C and CT must be 16-byte aligned");
    return WOTK_OK;
}

}  // namespace wotk

using namespace wotk;

extern "C" size_t wotk_median_workspace_size(int64_t n) {
    Arena ar(nullptr, 0);
    SelectState *st;
    unsigned long long *h;
    return median_ws(ar, n, &st, &h);
}

extern "C" int wotk_median_f64(const double *values, int64_t n, double *median_out,
                               int64_t *median_index_out, void *workspace, size_t workspace_bytes,
                               void *stream) {
    Arena ar(workspace, workspace_bytes);
    SelectState *st;
    unsigned long long *h;
    median_ws(ar, n, &st, &h);
    if (!ar.ok || !workspace) return fail(WOTK_EWORKSPACE, "median workspace too small");
    return median_run(values, n, median_out, median_index_out, st, h, (hipStream_t)stream);
}

extern "C" size_t wotk_cost_workspace_size(int64_t I, int64_t J, int64_t d, int32_t mode) {
    Arena ar(nullptr, 0);
    ar.take<double>((size_t)I * J);
    SelectState *st;
    unsigned long long *h;
    median_ws(ar, I * J, &st, &h);
    if (mode == WOTK_COST_BF16X3 || mode == WOTK_COST_BF16X6) ar.take<char>(cost_bf16_scratch(I, J, d));
    return ar.off;
}

extern "C" int wotk_sqeuclid_f64(const double *X, const double *Y, int64_t I, int64_t J, int64_t d,
                                 double *C64, int64_t ldc64, void *stream) {
    WOTK_REQUIRE(I > 0 && J > 0 && d > 0, "empty input");
    WOTK_REQUIRE(ldc64 >= J, "ldc64 < J");
    return launch_sqeuclid_f64(X, Y, I, J, d, C64, ldc64, (hipStream_t)stream);
}

extern "C" int wotk_cost_sqeuclid(const double *X, const double *Y, int64_t I, int64_t J, int64_t d,
                                  int32_t mode, float *C, int64_t ldc, float *CT, int64_t ldct,
                                  double *median_out, int64_t *median_index_out, void *workspace,
                                  size_t workspace_bytes, void *stream) {
    int rc = check_cost_layout(I, J, d, C, ldc, CT, ldct);
    if (rc) return rc;
    WOTK_REQUIRE(mode == WOTK_COST_F64 || mode == WOTK_COST_BF16X3 || mode == WOTK_COST_BF16X6,
                 "unknown cost mode %d", mode);
    hipStream_t s = (hipStream_t)stream;
    Arena ar(workspace, workspace_bytes);
    double *C64 = ar.take<double>((size_t)I * J);
    SelectState *st;
    unsigned long long *h;
    median_ws(ar, I * J, &st, &h);
    void *scratch = nullptr;
    if (mode != WOTK_COST_F64) scratch = ar.take<char>(cost_bf16_scratch(I, J, d));
    if (!ar.ok || !workspace) return fail(WOTK_EWORKSPACE, "cost workspace too small");
    rc = (mode == WOTK_COST_F64) ? launch_sqeuclid_f64(X, Y, I, J, d, C64, J, s)
                                 : launch_cost_bf16(X, Y, I, J, d, mode == WOTK_COST_BF16X3 ? 3 : 6, C64, J,
                                                    scratch, s);
    if (rc) return rc;
    rc = median_run(C64, I * J, nullptr, median_index_out, st, h, s);
    if (rc) return rc;
    rc = launch_normalise_transpose(C64, J, I, J, &st->median, 1.0, C, ldc, CT, ldct, s);
    if (rc) return rc;
    if (median_out) WOTK_HIP(hipMemcpyAsync(median_out, &st->median, sizeof(double), hipMemcpyDeviceToDevice, s));
    return WOTK_OK;
}

extern "C" int wotk_cost_import(const double *C64, int64_t I, int64_t J, int64_t ld64, double scale,
                                float *C, int64_t ldc, float *CT, int64_t ldct, void *stream) {
    int rc = check_cost_layout(I, J, 1, C, ldc, CT, ldct);
    if (rc) return rc;
    WOTK_REQUIRE(ld64 >= J, "ld64 < J");
    return launch_normalise_transpose(C64, ld64, I, J, nullptr, scale, C, ldc, CT, ldct,
                                      (hipStream_t)stream);
}
