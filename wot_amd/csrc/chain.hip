// Long-range couplings (wot/tmap/util.py:74-94, wot/tmap/chaining.py:4-40):
// T_0k = T_01 T_12 ... T_(k-1)k, successive transport-map products with no
// normalisation.  FP32 GEMM on the CDNA matrix cores: v_mfma_f32_32x32x2f32,
// 128 x 128 block tile, 4 waves each owning a 64 x 64 quadrant (2 x 2 MFMA
// tiles), K staged through LDS in chunks of 16.
#include "common.h"

namespace wotk {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kGT = 128;  // block tile (rows and columns)
constexpr int kGK = 16;   // K chunk

// C (M x N, ldc) = A (M x K, lda) * B (K x N, ldb); all FP32 row-major.
// MFMA 32x32x2 operand layout (lane l of a wave64): A element (row l%32,
// k l/32), B element (k l/32, column l%32); accumulator register r holds
// row 8*(r/4) + 4*(l/32) + r%4, column l%32.
__global__ __launch_bounds__(256) void gemm_f32_mfma_kernel(const float *__restrict__ A, int64_t lda,
                                                            const float *__restrict__ B, int64_t ldb,
                                                            float *__restrict__ C, int64_t ldc, int M, int N,
                                                            int K) {
    __shared__ float As[kGK][kGT + 4];  // As[k][m]
    __shared__ float Bs[kGK][kGT + 4];  // Bs[k][n]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;

    const int lrow = lane & 31, lk = lane >> 5;
    for (int k0 = 0; k0 < K; k0 += kGK) {
        // A tile: 128 rows x 16 k; B tile: 16 k x 128 columns (each 2048 floats, 8 per thread)
        for (int e = tid; e < kGT * kGK; e += 256) {
            const int r = e / kGK, kk = e % kGK;
            const int gm = m0 + r, gk = k0 + kk;
            As[kk][r] = (gm < M && gk < K) ? A[(int64_t)gm * lda + gk] : 0.f;
            const int kb = e / kGT, c = e % kGT;
            const int gkb = k0 + kb, gn = n0 + c;
            Bs[kb][c] = (gkb < K && gn < N) ? B[(int64_t)gkb * ldb + gn] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGK; kk += 2) {
            float av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = As[kk + lk][wm + a * 32 + lrow];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = Bs[kk + lk][wn + b * 32 + lrow];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm + a * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
                const int col = n0 + wn + b * 32 + (lane & 31);
                if (row < M && col < N) C[(int64_t)row * ldc + col] = acc[a][b][r];
            }
}

}  // namespace wotk

using namespace wotk;

extern "C" int wotk_gemm_f32(const float *A, int64_t lda, const float *B, int64_t ldb, float *C, int64_t ldc,
                             int64_t M, int64_t N, int64_t K, void *stream) {
    WOTK_REQUIRE(M > 0 && N > 0 && K > 0 && M < (1ll << 31) && N < (1ll << 31) && K < (1ll << 31),
                 "bad GEMM shape %lld x %lld x %lld", (long long)M, (long long)N, (long long)K);
    WOTK_REQUIRE(lda >= K && ldb >= N && ldc >= N, "leading dimension too small");
    dim3 grid((unsigned)((N + kGT - 1) / kGT), (unsigned)((M + kGT - 1) / kGT));
    gemm_f32_mfma_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(A, lda, B, ldb, C, ldc, (int)M, (int)N, (int)K);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}
