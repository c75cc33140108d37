// Squared-Euclidean cost with the cross term on the matrix cores:
//   D_ij = |x_i|^2 + |y_j|^2 - 2 x_i . y_j
// Coordinates are centred on the mean of both days (distances are
// translation invariant; centring shrinks the norms that cancel), norms are
// FP64, and the cross term is a split-BF16 product: x = x_hi + x_mid + x_lo
// with x_hi = bf16(x), x_mid = bf16(x - x_hi), x_lo = bf16(x - x_hi - x_mid),
// accumulated in FP32 as
//   3 terms: x_hi.y_hi + x_hi.y_mid + x_mid.y_hi
//   6 terms: the 3 above + x_hi.y_lo + x_lo.y_hi + x_mid.y_mid
// (every product of combined weight 2^-16 or larger; ~2^-16 and ~2^-24
// relative error of the cross term).
// v_mfma_f32_32x32x16_bf16: lane l supplies row l%32 and k = 8*(l/32)..+7 of
// A, column l%32 and the same k of B; accumulator register r holds row
// 8*(r/4) + 4*(l/32) + r%4, column l%32.  The K dimension (the PCA
// dimension, padded to a multiple of 16) is read straight from global memory.
#include "common.h"

namespace wotk {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// mean of the stacked coordinates: block k reduces column k over both days
__global__ __launch_bounds__(256) void coord_mean_kernel(const double *__restrict__ X, int n_i,
                                                         const double *__restrict__ Y, int n_j, int d,
                                                         double *__restrict__ mean) {
    __shared__ double red[4];
    const int k = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n_i; i += blockDim.x) acc += X[(int64_t)i * d + k];
    for (int j = threadIdx.x; j < n_j; j += blockDim.x) acc += Y[(int64_t)j * d + k];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) mean[k] = (red[0] + red[1] + red[2] + red[3]) / (n_i + n_j);
}

// centred FP64 norms and the bf16 hi/mid/lo split, K padded to kp with zeros
__global__ __launch_bounds__(256) void split_bf16_kernel(const double *__restrict__ X, int n, int d, int kp,
                                                         const double *__restrict__ mean, __bf16 *__restrict__ hi,
                                                         __bf16 *__restrict__ mid, __bf16 *__restrict__ lo,
                                                         double *__restrict__ norm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int k = 0; k < kp; ++k) {
        const double v = k < d ? X[(int64_t)i * d + k] - mean[k] : 0.0;
        acc = fma(v, v, acc);
        const __bf16 h = (__bf16)(float)v;
        const double r1 = v - (double)(float)h;
        const __bf16 m = (__bf16)(float)r1;
        const __bf16 l = (__bf16)(float)(r1 - (double)(float)m);
        hi[(int64_t)i * kp + k] = h;
        mid[(int64_t)i * kp + k] = m;
        lo[(int64_t)i * kp + k] = l;
    }
    norm[i] = acc;
}

struct SplitCoords {
    const __bf16 *h, *m, *l;  // hi / mid / lo parts, (n x kp)
    const double *norm;       // FP64 squared norms of the centred coordinates
};

// 128 x 128 output tile per block, 4 waves each owning 64 x 64 (2 x 2 MFMA tiles)
template <int TERMS>
__global__ __launch_bounds__(256) void cost_bf16_kernel(SplitCoords xs, SplitCoords ys, int n_i, int n_j, int kp,
                                                        double *__restrict__ D, int64_t ldd) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 128 + (wave >> 1) * 64;
    const int n0 = blockIdx.x * 128 + (wave & 1) * 64;
    const int lr = lane & 31, kg = (lane >> 5) * 8;
    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;
    for (int k0 = 0; k0 < kp; k0 += 16) {
        bf16x8 ah[2], am[2], al[2], bh[2], bm[2], bl[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int64_t o = (int64_t)min(m0 + a * 32 + lr, n_i - 1) * kp + k0 + kg;
            ah[a] = *reinterpret_cast<const bf16x8 *>(xs.h + o);
            am[a] = *reinterpret_cast<const bf16x8 *>(xs.m + o);
            if (TERMS == 6) al[a] = *reinterpret_cast<const bf16x8 *>(xs.l + o);
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int64_t o = (int64_t)min(n0 + b * 32 + lr, n_j - 1) * kp + k0 + kg;
            bh[b] = *reinterpret_cast<const bf16x8 *>(ys.h + o);
            bm[b] = *reinterpret_cast<const bf16x8 *>(ys.m + o);
            if (TERMS == 6) bl[b] = *reinterpret_cast<const bf16x8 *>(ys.l + o);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                // smallest products first: they are accumulated before the large hi.hi term
                if (TERMS == 6) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[a], bm[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                }
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[a], bh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bm[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
            }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + a * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
                const int col = n0 + b * 32 + (lane & 31);
                if (row < n_i && col < n_j)
                    D[(int64_t)row * ldd + col] = fmax(xs.norm[row] + ys.norm[col] - 2.0 * (double)acc[a][b][r], 0.0);
            }
}

static int64_t padded_k(int64_t d) { return round_up(d, 16); }

size_t cost_bf16_scratch(int64_t I, int64_t J, int64_t d) {
    Arena ar(nullptr, 0);
    const int64_t kp = padded_k(d);
    ar.take<double>(d);
    for (int p = 0; p < 3; ++p) ar.take<__bf16>((size_t)I * kp);
    for (int p = 0; p < 3; ++p) ar.take<__bf16>((size_t)J * kp);
    ar.take<double>(I);
    ar.take<double>(J);
    return ar.off + 256;
}

int launch_cost_bf16(const double *X, const double *Y, int64_t I, int64_t J, int64_t d, int terms, double *D,
                     int64_t ldd, void *scratch, hipStream_t s) {
    WOTK_REQUIRE(d <= 4096, "split-BF16 cost supports d <= 4096");
    WOTK_REQUIRE(terms == 3 || terms == 6, "split-BF16 cost has 3 or 6 terms");
    const int64_t kp = padded_k(d);
    Arena ar(scratch, cost_bf16_scratch(I, J, d));
    double *mean = ar.take<double>(d);
    __bf16 *xp[3], *yp[3];
    for (int p = 0; p < 3; ++p) xp[p] = ar.take<__bf16>((size_t)I * kp);
    for (int p = 0; p < 3; ++p) yp[p] = ar.take<__bf16>((size_t)J * kp);
    double *nx = ar.take<double>(I);
    double *ny = ar.take<double>(J);
    coord_mean_kernel<<<(unsigned)d, 256, 0, s>>>(X, (int)I, Y, (int)J, (int)d, mean);
    split_bf16_kernel<<<(unsigned)((I + 255) / 256), 256, 0, s>>>(X, (int)I, (int)d, (int)kp, mean, xp[0], xp[1],
                                                                   xp[2], nx);
    split_bf16_kernel<<<(unsigned)((J + 255) / 256), 256, 0, s>>>(Y, (int)J, (int)d, (int)kp, mean, yp[0], yp[1],
                                                                   yp[2], ny);
    const SplitCoords xs{xp[0], xp[1], xp[2], nx}, ys{yp[0], yp[1], yp[2], ny};
    dim3 grid((unsigned)((J + 127) / 128), (unsigned)((I + 127) / 128));
    if (terms == 3)
        cost_bf16_kernel<3><<<grid, 256, 0, s>>>(xs, ys, (int)I, (int)J, (int)kp, D, ldd);
    else
        cost_bf16_kernel<6><<<grid, 256, 0, s>>>(xs, ys, (int)I, (int)J, (int)kp, D, ldd);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

}  // namespace wotk
