// Squared-Euclidean cost with the cross term on the matrix cores:
//   D_ij = |x_i|^2 + |y_j|^2 - 2 x_i . y_j
// Coordinates are centred on the mean of both days (distances are
// translation invariant; centring shrinks the norms that cancel), norms are
// FP64, and the cross term is a split-BF16 product: x = x_hi + x_lo with
// x_hi = bf16(x), x_lo = bf16(x - x_hi), accumulated in FP32 as
//   x_hi.y_hi + x_hi.y_lo + x_lo.y_hi            (3 terms, x_lo.y_lo dropped).
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

// centred FP64 norms and the bf16 hi/lo split, K padded to kp with zeros
__global__ __launch_bounds__(256) void split_bf16_kernel(const double *__restrict__ X, int n, int d, int kp,
                                                         const double *__restrict__ mean, __bf16 *__restrict__ hi,
                                                         __bf16 *__restrict__ lo, double *__restrict__ norm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int k = 0; k < kp; ++k) {
        const double v = k < d ? X[(int64_t)i * d + k] - mean[k] : 0.0;
        acc = fma(v, v, acc);
        const __bf16 h = (__bf16)(float)v;
        const __bf16 l = (__bf16)(float)(v - (double)(float)h);
        hi[(int64_t)i * kp + k] = h;
        lo[(int64_t)i * kp + k] = l;
    }
    norm[i] = acc;
}

// 128 x 128 output tile per block, 4 waves each owning 64 x 64 (2 x 2 MFMA tiles)
__global__ __launch_bounds__(256) void cost_bf16x3_kernel(const __bf16 *__restrict__ Xh, const __bf16 *__restrict__ Xl,
                                                          const __bf16 *__restrict__ Yh, const __bf16 *__restrict__ Yl,
                                                          const double *__restrict__ nx, const double *__restrict__ ny,
                                                          int n_i, int n_j, int kp, double *__restrict__ D,
                                                          int64_t ldd) {
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
        bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int row = min(m0 + a * 32 + lr, n_i - 1);
            ah[a] = *reinterpret_cast<const bf16x8 *>(Xh + (int64_t)row * kp + k0 + kg);
            al[a] = *reinterpret_cast<const bf16x8 *>(Xl + (int64_t)row * kp + k0 + kg);
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int col = min(n0 + b * 32 + lr, n_j - 1);
            bh[b] = *reinterpret_cast<const bf16x8 *>(Yh + (int64_t)col * kp + k0 + kg);
            bl[b] = *reinterpret_cast<const bf16x8 *>(Yl + (int64_t)col * kp + k0 + kg);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
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
                    D[(int64_t)row * ldd + col] = fmax(nx[row] + ny[col] - 2.0 * (double)acc[a][b][r], 0.0);
            }
}

static int64_t padded_k(int64_t d) { return round_up(d, 16); }

size_t cost_bf16x3_scratch(int64_t I, int64_t J, int64_t d) {
    Arena ar(nullptr, 0);
    const int64_t kp = padded_k(d);
    ar.take<double>(d);
    ar.take<__bf16>((size_t)I * kp);
    ar.take<__bf16>((size_t)I * kp);
    ar.take<__bf16>((size_t)J * kp);
    ar.take<__bf16>((size_t)J * kp);
    ar.take<double>(I);
    ar.take<double>(J);
    return ar.off + 256;
}

int launch_cost_bf16x3(const double *X, const double *Y, int64_t I, int64_t J, int64_t d, double *D, int64_t ldd,
                       void *scratch, hipStream_t s) {
    WOTK_REQUIRE(d <= 4096, "bf16x3 cost supports d <= 4096");
    const int64_t kp = padded_k(d);
    Arena ar(scratch, cost_bf16x3_scratch(I, J, d));
    double *mean = ar.take<double>(d);
    __bf16 *xh = ar.take<__bf16>((size_t)I * kp);
    __bf16 *xl = ar.take<__bf16>((size_t)I * kp);
    __bf16 *yh = ar.take<__bf16>((size_t)J * kp);
    __bf16 *yl = ar.take<__bf16>((size_t)J * kp);
    double *nx = ar.take<double>(I);
    double *ny = ar.take<double>(J);
    coord_mean_kernel<<<(unsigned)d, 256, 0, s>>>(X, (int)I, Y, (int)J, (int)d, mean);
    split_bf16_kernel<<<(unsigned)((I + 255) / 256), 256, 0, s>>>(X, (int)I, (int)d, (int)kp, mean, xh, xl, nx);
    split_bf16_kernel<<<(unsigned)((J + 255) / 256), 256, 0, s>>>(Y, (int)J, (int)d, (int)kp, mean, yh, yl, ny);
    dim3 grid((unsigned)((J + 127) / 128), (unsigned)((I + 127) / 128));
    cost_bf16x3_kernel<<<grid, 256, 0, s>>>(xh, xl, yh, yl, nx, ny, (int)I, (int)J, (int)kp, D, ldd);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

}  // namespace wotk
