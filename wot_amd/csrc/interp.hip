// Interpolated populations for OT validation (wot/ot/util.py:109-200):
// sample (i, j) cell pairs with probability proportional to a weight matrix
// and place each sample on the segment between cell i of day t0 and cell j of
// day t1: row = p0[i] (1 - f) + p1[j] f.
//   interp_weights:  W[i][j] = T[i][j] / (sum_i T[i][j])^(1 - f)   (util.py:142)
//   interp_sample:   idx[s] = first k with cdf[k] > u[s]  (numpy's
//                    cdf.searchsorted(u, side='right') in RandomState.choice)
//   interp_gather:   out[s] = p0[idx/J] (1 - f) + p1[idx%J] f
// The uniforms come from the caller (numpy's stream), so a seeded caller gets
// the reference's samples.
#include "common.h"

namespace wotk {
namespace {

constexpr int kColRows = 128;  // rows per partial column sum

// partial column sums of T (I x J) over blocks of kColRows rows (fixed order)
__global__ __launch_bounds__(256) void colsum_partial_kernel(const double *__restrict__ T, int64_t ld, int64_t I,
                                                             int64_t J, double *__restrict__ part) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * kColRows;
    if (j >= J) return;
    double s = 0.0;
    for (int r = 0; r < kColRows; ++r) {
        const int64_t i = r0 + r;
        if (i >= I) break;
        s += T[i * ld + j];
    }
    part[(int64_t)blockIdx.y * J + j] = s;
}

__global__ __launch_bounds__(256) void colsum_final_kernel(const double *__restrict__ part, int n_part, int64_t J,
                                                           double *__restrict__ colsum) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= J) return;
    double s = 0.0;
    for (int p = 0; p < n_part; ++p) s += part[(int64_t)p * J + j];
    colsum[j] = s;
}

__global__ __launch_bounds__(256) void weights_kernel(const double *__restrict__ T, int64_t ld, int64_t I, int64_t J,
                                                      const double *__restrict__ colsum, double expo,
                                                      double *__restrict__ W) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= J) return;
    W[i * J + j] = T[i * ld + j] / pow(colsum[j], expo);
}

__global__ __launch_bounds__(256) void sample_kernel(const double *__restrict__ cdf, int64_t n, const double *__restrict__ u,
                                                     int64_t size, int64_t *__restrict__ idx) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= size) return;
    const double x = u[s];
    int64_t lo = 0, hi = n;  // first k in [0, n) with cdf[k] > x, n if none
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (cdf[mid] > x) hi = mid;
        else lo = mid + 1;
    }
    idx[s] = lo;
}

__global__ __launch_bounds__(256) void gather_kernel(const int64_t *__restrict__ idx, int64_t size,
                                                     const double *__restrict__ p0, int64_t ld0, int64_t I,
                                                     const double *__restrict__ p1, int64_t ld1, int64_t J, int64_t d,
                                                     double frac, double *__restrict__ out, int64_t ldo) {
    const int64_t s = blockIdx.y;
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= size || k >= d) return;
    int64_t c = idx[s];
    c = c < 0 ? 0 : (c >= I * J ? I * J - 1 : c);
    const int64_t i = c / J, j = c % J;
    out[s * ldo + k] = p0[i * ld0 + k] * (1.0 - frac) + p1[j * ld1 + k] * frac;
}

}  // namespace
}  // namespace wotk

using namespace wotk;

extern "C" int wotk_interp_weights(const double *T, int64_t ld, int64_t I, int64_t J, double frac, double *W,
                                   double *colsum, void *stream) {
    WOTK_REQUIRE(T && W && colsum && I > 0 && J > 0 && ld >= J, "interp_weights: bad arguments");
    const int64_t n_part = (I + kColRows - 1) / kColRows;
    WOTK_REQUIRE(n_part < 65536, "interp_weights: too many rows (%lld)", (long long)I);
    hipStream_t st = (hipStream_t)stream;
    // the partial sums live in W until the weights overwrite it (W holds I x J >= n_part x J doubles)
    colsum_partial_kernel<<<dim3((unsigned)((J + 255) / 256), (unsigned)n_part), 256, 0, st>>>(T, ld, I, J, W);
    WOTK_LAUNCH_CHECK();
    colsum_final_kernel<<<(unsigned)((J + 255) / 256), 256, 0, st>>>(W, (int)n_part, J, colsum);
    WOTK_LAUNCH_CHECK();
    for (int64_t r0 = 0; r0 < I; r0 += 65535) {
        const int64_t rows = I - r0 < 65535 ? I - r0 : 65535;
        weights_kernel<<<dim3((unsigned)((J + 255) / 256), (unsigned)rows), 256, 0, st>>>(T + r0 * ld, ld, rows, J,
                                                                                         colsum, 1.0 - frac,
                                                                                         W + r0 * J);
        WOTK_LAUNCH_CHECK();
    }
    return WOTK_OK;
}

extern "C" int wotk_interp_sample(const double *cdf, int64_t n, const double *u, int64_t size, int64_t *idx,
                                  void *stream) {
    WOTK_REQUIRE(cdf && u && idx && n > 0 && size >= 0, "interp_sample: bad arguments");
    if (size == 0) return WOTK_OK;
    sample_kernel<<<(unsigned)((size + 255) / 256), 256, 0, (hipStream_t)stream>>>(cdf, n, u, size, idx);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

extern "C" int wotk_interp_gather(const int64_t *idx, int64_t size, const double *p0, int64_t ld0, int64_t I,
                                  const double *p1, int64_t ld1, int64_t J, int64_t d, double frac, double *out,
                                  int64_t ldo, void *stream) {
    WOTK_REQUIRE(idx && p0 && p1 && out && I > 0 && J > 0 && d > 0 && ld0 >= d && ld1 >= d && ldo >= d,
                 "interp_gather: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    for (int64_t s0 = 0; s0 < size; s0 += 65535) {
        const int64_t rows = size - s0 < 65535 ? size - s0 : 65535;
        gather_kernel<<<dim3((unsigned)((d + 255) / 256), (unsigned)rows), 256, 0, st>>>(
            idx + s0, rows, p0, ld0, I, p1, ld1, J, d, frac, out + s0 * ldo, ldo);
        WOTK_LAUNCH_CHECK();
    }
    return WOTK_OK;
}
