// Instruction-rate probe for the roofline of the compute-bound sweeps.
//
// Each probe kernel runs a fixed number of independent dependency chains per
// lane over one instruction (or one instruction sequence) of the batched
// sweep's inner loop, so the measured time is issue-bound: rate = lanes x
// chains x steps / elapsed.  The chains stay in a bounded range (no inf/NaN),
// and one store per lane under a value test that never holds keeps the
// compiler from removing them.
#include <cmath>

#include "common.h"

namespace wotk {

constexpr int kProbeChains = 8;
constexpr int kProbeSteps = 2048;

enum ProbeOp {
    PROBE_EXP_F32 = 0,    // v_exp_f32
    PROBE_FMA_F64 = 1,    // v_fma_f64
    PROBE_FMA_F32 = 2,    // v_fma_f32
    PROBE_CVT_F64 = 3,    // v_add_f64 + v_cvt_f32_f64 + v_cvt_f64_f32 (a triple)
    PROBE_RNDNE_F64 = 4,  // v_rndne_f64 + v_add_f64 (a pair)
    PROBE_LDEXP_F32 = 5,  // v_ldexp_f32
    PROBE_SPLIT_EXP = 6,  // the sweep's exact split exponential of an FP64 exponent (one entry)
    PROBE_NUM = 7,
};

__device__ __forceinline__ float probe_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

template <int OP>
__global__ __launch_bounds__(256) void probe_kernel(float *sink, float seed) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    float xf[kProbeChains];
    double xd[kProbeChains];
#pragma unroll
    for (int c = 0; c < kProbeChains; ++c) {
        xf[c] = seed + 0.001f * (float)(c + (tid & 7));
        xd[c] = (double)xf[c];
    }
    for (int s = 0; s < kProbeSteps; ++s) {
#pragma unroll
        for (int c = 0; c < kProbeChains; ++c) {
            if constexpr (OP == PROBE_EXP_F32) {
                xf[c] = probe_exp2(-xf[c]);  // stays in (0.5, 1]
            } else if constexpr (OP == PROBE_FMA_F64) {
                xd[c] = fma(xd[c], 0.999, 1.0e-3);
            } else if constexpr (OP == PROBE_FMA_F32) {
                xf[c] = fmaf(xf[c], 0.999f, 1.0e-3f);
            } else if constexpr (OP == PROBE_CVT_F64) {
                xd[c] = (double)(float)(xd[c] + 1.0e-7);  // the add keeps the converts from folding
            } else if constexpr (OP == PROBE_RNDNE_F64) {
                xd[c] = rint(xd[c]) + 0.25;
            } else if constexpr (OP == PROBE_LDEXP_F32) {
                xf[c] = ldexpf(xf[c], (s & 1) ? 1 : -1);
            } else {
                // y = fma(-k, c, hk + fk) in FP64, 2^y = ldexp(exp2(y - rint y), rint y)
                const double y = fma(-3.0, xd[c], 1.5 + xd[c] * 0.25);
                const double yc = fmax(y, -16384.0);
                const double n = rint(yc);
                const float t = ldexpf(probe_exp2((float)(yc - n)), (int)n);
                xd[c] = (double)t * 0.5;
            }
        }
    }
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < kProbeChains; ++c) acc += xf[c] + (float)xd[c];
    if (acc == -1.0f) sink[tid] = acc;  // never true for these chains
}

// operations per step and chain of each probe (pairs count as two, the convert triple as three)
constexpr int kProbeOpsPerStep[PROBE_NUM] = {1, 1, 1, 3, 2, 1, 1};

template <int OP>
static int probe_time(float *sink, int grid, hipStream_t s, double *ops_per_s) {
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e == hipSuccess) {
        probe_kernel<OP><<<grid, 256, 0, s>>>(sink, 0.25f);  // warm-up (code object load)
        e = hipEventRecord(ev[0], s);
    }
    if (e == hipSuccess) {
        probe_kernel<OP><<<grid, 256, 0, s>>>(sink, 0.25f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ev[1], s);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) return fail(WOTK_EHIP, "probe %d failed: %s", OP, hipGetErrorString(e));
    const double lanes = (double)grid * 256.0;
    *ops_per_s = lanes * kProbeChains * kProbeSteps * kProbeOpsPerStep[OP] / (ms * 1e-3);
    return WOTK_OK;
}

// Issue rate of an FMA / exponential mix (29 dependent-chain FP32 FMAs and one
// v_exp_f32 per step and chain, about the batched sweep's ratio) at a given
// number of waves per SIMD: 4-wave workgroups, one per CU per wave level.
constexpr int kMixFmas = 29;

__global__ __launch_bounds__(256) void probe_mix_kernel(float *sink, float seed, int steps) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    float x[kProbeChains];
#pragma unroll
    for (int c = 0; c < kProbeChains; ++c) x[c] = seed + 0.001f * (float)(c + (tid & 7));
    for (int s = 0; s < steps; ++s) {
#pragma unroll
        for (int c = 0; c < kProbeChains; ++c) {
#pragma unroll
            for (int u = 0; u < kMixFmas; ++u) x[c] = fmaf(x[c], 0.999f, 1.0e-3f);
            x[c] = probe_exp2(-x[c]);  // stays in (0.5, 1]
        }
    }
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < kProbeChains; ++c) acc += x[c];
    if (acc == -1.0f) sink[tid] = acc;  // never true for these chains
}

constexpr int kMixWaveLevels[] = {1, 2, 4, 8};
constexpr int kMixLevels = 4;

}  // namespace wotk

using namespace wotk;

extern "C" int wotk_probe_issue_mix(double *rates_host, int32_t n_rates, void *sink, size_t sink_bytes,
                                    void *stream) {
    WOTK_REQUIRE(rates_host != nullptr && n_rates >= kMixLevels, "rates_host needs %d entries", kMixLevels);
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, ncu = 256;
    WOTK_HIP(hipGetDevice(&dev));
    WOTK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    WOTK_REQUIRE(sink != nullptr && sink_bytes >= (size_t)ncu * 8 * 256 * sizeof(float), "sink needs %zu bytes",
                 (size_t)ncu * 8 * 256 * sizeof(float));
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    for (int lv = 0; lv < kMixLevels && e == hipSuccess; ++lv) {
        const int w = kMixWaveLevels[lv], grid = ncu * w;
        const int steps = 256 * 8 / w;  // about the same duration at every level
        probe_mix_kernel<<<grid, 256, 0, s>>>((float *)sink, 0.25f, steps);  // warm-up
        e = hipEventRecord(ev[0], s);
        if (e == hipSuccess) {
            probe_mix_kernel<<<grid, 256, 0, s>>>((float *)sink, 0.25f, steps);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(ev[1], s);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (e == hipSuccess)
            rates_host[lv] = (double)grid * 256.0 * kProbeChains * steps * (kMixFmas + 1) / (ms * 1e-3);
    }
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) return fail(WOTK_EHIP, "issue-mix probe failed: %s", hipGetErrorString(e));
    return WOTK_OK;
}

extern "C" int wotk_probe_valu_rates(double *rates_host, int32_t n_rates, void *sink, size_t sink_bytes,
                                     void *stream) {
    WOTK_REQUIRE(rates_host != nullptr && n_rates >= PROBE_NUM, "rates_host needs %d entries", (int)PROBE_NUM);
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, ncu = 256;
    WOTK_HIP(hipGetDevice(&dev));
    WOTK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = ncu * 8;  // 8 blocks of 4 waves per CU: 8 waves per SIMD
    WOTK_REQUIRE(sink != nullptr && sink_bytes >= (size_t)grid * 256 * sizeof(float),
                 "sink needs %zu bytes", (size_t)grid * 256 * sizeof(float));
    float *sk = (float *)sink;
    int rc = 0;
    rc = rc ? rc : probe_time<PROBE_EXP_F32>(sk, grid, s, &rates_host[PROBE_EXP_F32]);
    rc = rc ? rc : probe_time<PROBE_FMA_F64>(sk, grid, s, &rates_host[PROBE_FMA_F64]);
    rc = rc ? rc : probe_time<PROBE_FMA_F32>(sk, grid, s, &rates_host[PROBE_FMA_F32]);
    rc = rc ? rc : probe_time<PROBE_CVT_F64>(sk, grid, s, &rates_host[PROBE_CVT_F64]);
    rc = rc ? rc : probe_time<PROBE_RNDNE_F64>(sk, grid, s, &rates_host[PROBE_RNDNE_F64]);
    rc = rc ? rc : probe_time<PROBE_LDEXP_F32>(sk, grid, s, &rates_host[PROBE_LDEXP_F32]);
    rc = rc ? rc : probe_time<PROBE_SPLIT_EXP>(sk, grid, s, &rates_host[PROBE_SPLIT_EXP]);
    return rc;
}

extern "C" size_t wotk_probe_sink_bytes(void) {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    return (size_t)ncu * 8 * 256 * sizeof(float);
}
