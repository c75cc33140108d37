// Fixed-iteration solve (transport_stablev2 semantics,
// wot/ot/optimal_transport.py:167-236) on the cost of ot_model.py:242-253
// without ever materialising the cost or the kernel: every sweep recomputes
// C_ij from the coordinates tile by tile on the matrix cores.
//
// Coordinates are centred on the mean of both days and scaled by
// 1/sqrt(median), so C = |x'|^2 + |y'|^2 - 2 x'.y' directly.  The cross term
// x'.y' is a split-BF16 MFMA product (x' = hi + mid + lo):
//   3 terms: hi.hi + hi.lo' + lo'.hi  (lo' = bf16(x' - hi); ~2^-16 relative)
//   6 terms: hi.hi + hi.mid + mid.hi + hi.lo + lo.hi + mid.mid (~2^-24 relative)
// accumulated in FP32 (v_mfma_f32_32x32x16_bf16) in two chains, hi.hi and the
// smaller products.  With 6 terms they are added in FP64 (their FP32 sum would
// round at 2^-24 of the cross term; in FP64 the marginals of a 97 x 131, d = 7
// pair at eps = 0.05 move from 1.6e-6 to 6.7e-7 of the FP64 oracle's, for 6 %
// more sweep time at 60k); with 3 terms, whose missing products already cost
// 2^-16, in FP32.  Per entry the exponent is
// formed in FP64 against the row's previous potential (a fixed shift, as in
// the batched sweep); rows whose sum under- or overflows are queued and
// recomputed by a repair kernel from FP64 direct distances with an online
// maximum.  The median of the exact FP64 distances (the reference's
// normalisation) comes from a radix select that regenerates the distances in
// every pass.
#include <type_traits>

#include "common.h"
#include "fused_problem.h"

namespace wotk {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kFThreads = 256;
constexpr int kFusedMaxSplit = 8;        // column splits per sweep, at most
constexpr int kFusedTargetBlocks = 768;  // 3 waves per SIMD x 1024 SIMDs / 4 waves per block

// ---------------------------------------------------------------------------
// preparation
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fz_mean_kernel(const double *__restrict__ X, int n_i, const double *__restrict__ Y,
                                                      int n_j, int d, double *__restrict__ mean) {
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

// x' = (x - mean) / sqrt(median): FP64 copy, FP64 norm, bf16 split into
// nterm pieces (hi, mid, lo) with K padded to kp (pitch kp)
__global__ __launch_bounds__(256) void fz_split_kernel(const double *__restrict__ X, int n, int d, int kp,
                                                       const double *__restrict__ mean,
                                                       const double *__restrict__ median, int npiece,
                                                       __bf16 *__restrict__ p0, __bf16 *__restrict__ p1,
                                                       __bf16 *__restrict__ p2, double *__restrict__ xs,
                                                       double *__restrict__ norm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sc = 1.0 / sqrt(median[0]);
    double acc = 0.0;
    for (int k = 0; k < kp; ++k) {
        const double v = k < d ? (X[(int64_t)i * d + k] - mean[k]) * sc : 0.0;
        if (k < d) xs[(int64_t)i * d + k] = v;
        acc = fma(v, v, acc);
        const __bf16 h = (__bf16)(float)v;
        const double r1 = v - (double)(float)h;
        const __bf16 m = (__bf16)(float)r1;
        const double r2 = r1 - (double)(float)m;
        p0[(int64_t)i * kp + k] = h;
        p1[(int64_t)i * kp + k] = m;
        if (npiece > 2) p2[(int64_t)i * kp + k] = (__bf16)(float)r2;
    }
    norm[i] = acc;
}

// ---------------------------------------------------------------------------
// exact median of D_ij = sum_k (x_ik - y_jk)^2 without storing D: each pass
// regenerates 64 x 64 tiles in FP64 and histograms the keys matching the
// prefix found so far (same digit schedule as the stored-matrix select)
// ---------------------------------------------------------------------------
struct FzSelect {
    unsigned long long prefix, mask;
    long long rank;
    unsigned long long eq, next;
    double median;
};

template <int MODE>  // 0: histogram digit, 1: smallest key above the prefix
__global__ __launch_bounds__(256) void fz_select_kernel(const double *__restrict__ X, const double *__restrict__ Y,
                                                        int n_i, int n_j, int d, int shift, int width,
                                                        FzSelect *__restrict__ st,
                                                        unsigned long long *__restrict__ ghist) {
    __shared__ double xs[32][64];
    __shared__ double ys[32][64];
    __shared__ unsigned int h[2048];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    if (MODE == 0)
        for (int b = tid; b < 2048; b += 256) h[b] = 0u;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < d; k0 += 32) {
        for (int e = tid; e < 64 * 32; e += 256) {
            const int r = e / 32, kk = e % 32;
            const bool kin = k0 + kk < d;
            xs[kk][r] = (kin && i0 + r < n_i) ? X[(int64_t)(i0 + r) * d + k0 + kk] : 0.0;
            ys[kk][r] = (kin && j0 + r < n_j) ? Y[(int64_t)(j0 + r) * d + k0 + kk] : 0.0;
        }
        __syncthreads();
        const int km = min(32, d - k0);
        for (int kk = 0; kk < km; ++kk)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const double t = xs[kk][ty * 4 + a] - ys[kk][tx * 4 + b];
                    acc[a][b] = fma(t, t, acc[a][b]);
                }
        __syncthreads();
    }
    const unsigned long long prefix = st->prefix, mask = st->mask;
    unsigned long long best = ~0ull;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (i0 + ty * 4 + a >= n_i || j0 + tx * 4 + b >= n_j) continue;
            const unsigned long long key = (unsigned long long)__double_as_longlong(acc[a][b]);
            if (MODE == 0) {
                if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & ((1ull << width) - 1ull)], 1u);
            } else {
                if (key > prefix && key < best) best = key;
            }
        }
    if (MODE == 0) {
        __syncthreads();
        for (int b = tid; b < 2048; b += 256)
            if (h[b]) atomicAdd(&ghist[b], (unsigned long long)h[b]);
    } else {
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(best, off);
            best = o < best ? o : best;
        }
        if ((tid & 63) == 0 && best != ~0ull) atomicMin(&st->next, best);
    }
}

__global__ void fz_select_init_kernel(FzSelect *st, long long rank, unsigned long long *hist) {
    for (int b = threadIdx.x; b < 2048; b += blockDim.x) hist[b] = 0ull;
    if (threadIdx.x) return;
    st->prefix = 0ull;
    st->mask = 0ull;
    st->rank = rank;
    st->eq = 0ull;
    st->next = ~0ull;
}

__global__ __launch_bounds__(256) void fz_select_scan_kernel(unsigned long long *__restrict__ ghist, int shift,
                                                             int width, int last, FzSelect *__restrict__ st) {
    if (threadIdx.x == 0) {
        const int nb = 1 << width;
        long long rank = st->rank;
        unsigned long long cum = 0;
        int b = 0;
        for (; b < nb - 1; ++b) {
            if (cum + ghist[b] > (unsigned long long)rank) break;
            cum += ghist[b];
        }
        st->prefix |= (unsigned long long)b << shift;
        st->mask |= ((1ull << width) - 1ull) << shift;
        st->rank = rank - (long long)cum;
        if (last) st->eq = ghist[b];
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 2048; b += blockDim.x) ghist[b] = 0ull;
}

__global__ void fz_select_finish_kernel(FzSelect *st, int odd) {
    const double lo = __longlong_as_double((long long)st->prefix);
    double m = lo;
    if (!odd) {
        const bool tie = (unsigned long long)(st->rank + 1) < st->eq;
        m = (lo + (tie ? lo : __longlong_as_double((long long)st->next))) / 2.0;
    }
    st->median = m;
}

// ---------------------------------------------------------------------------
// fused sweep: row LSE over recomputed cost tiles, potential update epilogue
// ---------------------------------------------------------------------------
struct FusedArgs {
    const __bf16 *A0, *A1, *A2;  // row points, pitch kp
    const __bf16 *B0, *B1, *B2;  // column points, pitch kp
    const double *na, *nb;       // squared norms of the scaled coordinates
    int n_rows, n_cols, kp;
    const double *h;             // column potentials
    double *pot;                 // row potentials (read as the shift, then updated)
    const double *logw;          // per-row log weight or nullptr -> *logw_dev
    const double *logw_dev;
    double log_n_other, alpha, eps, k;
    int *bad_list;               // rows whose fixed-shift sum under/overflowed
    int *n_bad;
    double *row_sums;            // row-sum mode: write sum_j R_ij / rs_div, leave pot unchanged
    double *part;                // column split (gridDim.y > 1): per-split row sums, split-major
    double rs_div;               // row-sum mode divisor (n_cols: rows of R / J; 1: of R)
};

// Row epilogue of a sweep from the row's fixed-shift sum S: new potential
// (or the plan's row sum in row-sum mode); rows whose sum under/overflowed
// are queued for the repair kernel (a NaN sum leaves the row as it is).
__device__ inline void fused_row_epilogue(const FusedArgs &fa, int row, double S) {
    const double f_old = fa.pot[row];
    if (isnan(S)) {  // NaN inputs (e.g. a NaN growth entry) propagate as in the reference's arithmetic
        if (fa.row_sums) fa.row_sums[row] = S;
        else fa.pot[row] = S;
        return;
    }
    if (!(S > 0.0) || isinf(S)) {
        fa.bad_list[atomicAdd(fa.n_bad, 1)] = row;
        return;
    }
    if (fa.row_sums) {  // the shift is the final potential: S = sum_j R_ij
        fa.row_sums[row] = S / fa.rs_div;
        return;
    }
    const double lse = log(S) - f_old / fa.eps;  // ln sum_j exp((h_j - C_ij)/eps)
    const double lw = fa.logw ? fa.logw[row] : *fa.logw_dev;
    fa.pot[row] = fa.alpha * fa.eps * (lw + fa.log_n_other - lse);
}

// EXACT: FP64 exponential (exp2 of the FP64 exponent) and FP64 lane sums, for
// the row and column sums of the duality-gap pass; otherwise v_exp_f32 of the
// exponent rounded to FP32 and FP32 lane sums (the sweeps)
template <int TERMS, int NKS, bool EXACT = false>
__global__ __launch_bounds__(kFThreads) void fused_lse_kernel(FusedArgs fa) {
    constexpr int NP = TERMS == 3 ? 2 : 3;  // pieces per point
    constexpr int KP = NKS * 16;
    constexpr int PITCH = KP + 8;            // 16-byte pad: spreads lanes over LDS banks
    constexpr int TILE = NP * 32 * PITCH;
    __shared__ __bf16 bsh[2 * TILE];         // double-buffered column tile
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 31, kg = (lane >> 5) * 8;
    const int r0 = blockIdx.x * 128 + wave * 32;
    const __bf16 *Ap[3] = {fa.A0, fa.A1, fa.A2};
    const __bf16 *Bp[3] = {fa.B0, fa.B1, fa.B2};

    // row constants for this lane's 16 accumulator rows: (f_old - |x'|^2) * k
    double wf[16];
    using acc_t = typename std::conditional<EXACT, double, float>::type;
    acc_t s[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = min(r0 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4), fa.n_rows - 1);
        wf[r] = (fa.pot[row] - fa.na[row]) * fa.k;
        s[r] = 0.f;
    }
    // this wave's 32 rows stay in registers for the whole sweep
    const int arow = min(r0 + lr, fa.n_rows - 1);
    bf16x8 afr[NKS][NP];
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
        for (int p = 0; p < NP; ++p)
            afr[ks][p] = *reinterpret_cast<const bf16x8 *>(Ap[p] + (int64_t)arow * KP + ks * 16 + kg);
    const double k2 = 2.0 * fa.k;

    // column tiles stream through two LDS buffers: the next tile's global
    // loads are issued into registers before the current tile's MFMA and
    // exponentials, and written to LDS only afterwards, so the load latency
    // hides behind the compute
    constexpr int PER_PIECE = 32 * KP / 8;                       // bf16x8 chunks per piece
    constexpr int NCHUNK = (NP * PER_PIECE + kFThreads - 1) / kFThreads;
    bf16x8 stage[NCHUNK];
    auto fetch = [&](int c0) {
#pragma unroll
        for (int u = 0; u < NCHUNK; ++u) {
            const int e = tid + u * kFThreads;
            if (e < NP * PER_PIECE) {
                const int p = e / PER_PIECE, rem = e % PER_PIECE;
                const int col = rem / (KP / 8), kk = (rem % (KP / 8)) * 8;
                const int gcol = min(c0 + col, fa.n_cols - 1);
                stage[u] = *reinterpret_cast<const bf16x8 *>(Bp[p] + (int64_t)gcol * KP + kk);
            }
        }
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int u = 0; u < NCHUNK; ++u) {
            const int e = tid + u * kFThreads;
            if (e < NP * PER_PIECE) {
                const int p = e / PER_PIECE, rem = e % PER_PIECE;
                const int col = rem / (KP / 8), kk = (rem % (KP / 8)) * 8;
                *reinterpret_cast<bf16x8 *>(bsh + buf * TILE + (p * 32 + col) * PITCH + kk) = stage[u];
            }
        }
    };

    // column split: block (x, y) takes the column tiles [t_begin, t_end) of its rows
    const int ntiles_all = (fa.n_cols + 31) / 32;
    const int per_split = (ntiles_all + (int)gridDim.y - 1) / (int)gridDim.y;
    const int t_begin = min(ntiles_all, (int)blockIdx.y * per_split);
    const int t_end = min(ntiles_all, t_begin + per_split);
    auto col_const = [&](int t, double &hv, double &nv) {  // this lane's column of tile t
        const int col = min(t * 32 + lr, fa.n_cols - 1);
        hv = fa.h[col];
        nv = fa.nb[col];
    };
    double h_cur, n_cur, h_nxt = 0.0, n_nxt = 0.0;
    col_const(t_begin, h_cur, n_cur);
    fetch(t_begin * 32);
    commit(0);
    __syncthreads();
    for (int t = t_begin; t < t_end; ++t) {
        const int buf = (t - t_begin) & 1;
        if (t + 1 < t_end) {
            fetch((t + 1) * 32);
            col_const(t + 1, h_nxt, n_nxt);
        }
        floatx16 acc, acc2;  // two independent accumulation chains
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            acc[r] = 0.f;
            acc2[r] = 0.f;
        }
        const __bf16 *bt = bsh + buf * TILE;
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
            const int ko = ks * 16 + kg;
            bf16x8 b[NP];
#pragma unroll
            for (int p = 0; p < NP; ++p) b[p] = *reinterpret_cast<const bf16x8 *>(bt + (p * 32 + lr) * PITCH + ko);
            const bf16x8 *a = afr[ks];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[0], acc, 0, 0, 0);
            acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[1], acc2, 0, 0, 0);
            acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[0], acc2, 0, 0, 0);
            if (TERMS == 6) {
                acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[NP - 1], acc2, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[NP - 1], b[0], acc2, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[1], acc2, 0, 0, 0);
            }
        }
        const int col = t * 32 + lr;
        if (col < fa.n_cols) {
            const double uc = (h_cur - n_cur) * fa.k;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                // 6 terms: the chains added in FP64; 3 terms (2^-16 of the cross term already
                // lost to the missing products): one FP32 add and one widening
                const double xy = TERMS == 6 ? (double)acc[r] + (double)acc2[r] : (double)(acc[r] + acc2[r]);
                const double y = fma(k2, xy, uc + wf[r]);
                if constexpr (EXACT)
                    s[r] += exp2(y);
                else
                    s[r] += __builtin_amdgcn_exp2f((float)y);
            }
        }
        if (t + 1 < t_end) commit(buf ^ 1);  // the other buffer was released by the previous barrier
        h_cur = h_nxt;
        n_cur = n_nxt;
        __syncthreads();
    }
    // sum over the 32 lanes that share a row set
    double sd[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        double v = (double)s[r];
#pragma unroll
        for (int off = 1; off < 32; off <<= 1) v += __shfl_xor(v, off);
        sd[r] = v;
    }
    if (lr == 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = r0 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
            if (row >= fa.n_rows) continue;
            if (gridDim.y > 1) fa.part[(int64_t)blockIdx.y * fa.n_rows + row] = sd[r];
            else fused_row_epilogue(fa, row, sd[r]);
        }
    }
}

// column-split sweeps: the splits' row sums added in a fixed order, then the
// row epilogue
__global__ __launch_bounds__(256) void fused_finalize_kernel(FusedArgs fa, int n_split) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= fa.n_rows) return;
    double S = 0.0;
    for (int q = 0; q < n_split; ++q) S += fa.part[(int64_t)q * fa.n_rows + row];
    fused_row_epilogue(fa, row, S);
}

// repair: rows whose fixed-shift sum under/overflowed, from FP64 direct
// distances of the scaled coordinates with a per-lane online maximum
__global__ __launch_bounds__(256) void fused_repair_kernel(const int *__restrict__ bad_list, const int *__restrict__ n_bad,
                                                           const double *__restrict__ xr, const double *__restrict__ xc,
                                                           int n_cols, int d, const double *__restrict__ h,
                                                           double *__restrict__ pot, const double *__restrict__ logw,
                                                           const double *__restrict__ logw_dev, double log_n_other,
                                                           double alpha, double eps, double k,
                                                           double *__restrict__ row_sums, double rs_div) {
    __shared__ double sg_sh[4], sd_sh[4];
    const int nb = *n_bad;
    for (int q = blockIdx.x; q < nb; q += gridDim.x) {
        const int row = bad_list[q];
        double sg = -1.0e6, sd = 0.0;
        for (int j = threadIdx.x; j < n_cols; j += blockDim.x) {
            double c = 0.0;
            for (int t = 0; t < d; ++t) {
                const double z = xr[(int64_t)row * d + t] - xc[(int64_t)j * d + t];
                c = fma(z, z, c);
            }
            const double y = (h[j] - c) * k;
            if (y > sg) {
                const double ns = ceil(y);
                sd = ldexp(sd, (int)(sg - ns));
                sg = ns;
            }
            sd += exp2(y - sg);
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double og = __shfl_xor(sg, off), os = __shfl_xor(sd, off);
            const double ng = fmax(sg, og);
            sd = ldexp(sd, (int)(sg - ng)) + ldexp(os, (int)(og - ng));
            sg = ng;
        }
        if ((threadIdx.x & 63) == 0) {
            sg_sh[threadIdx.x >> 6] = sg;
            sd_sh[threadIdx.x >> 6] = sd;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double g = sg_sh[0], s = sd_sh[0];
            for (int w = 1; w < 4; ++w) {
                const double ng = fmax(g, sg_sh[w]);
                s = ldexp(s, (int)(g - ng)) + ldexp(sd_sh[w], (int)(sg_sh[w] - ng));
                g = ng;
            }
            const double lse = (g + log2(s)) * kLn2;
            if (row_sums) {
                row_sums[row] = exp(pot[row] / eps + lse) / rs_div;
            } else {
                const double lw = logw ? logw[row] : *logw_dev;
                pot[row] = alpha * eps * (lw + log_n_other - lse);
            }
        }
        __syncthreads();
    }
}

__global__ void fused_clear_kernel(int *n_bad) { *n_bad = 0; }

__global__ __launch_bounds__(1024) void fz_growth_kernel(const double *__restrict__ G, int n, double *__restrict__ logp,
                                                         double *__restrict__ misc) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        logp[i] = log(G[i]);
        acc += G[i];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        misc[0] = log(t / n);
    }
}

__global__ void fz_zero_kernel(double *x, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = 0.0;
}

static const int kFzShift[6] = {53, 42, 31, 20, 10, 0};
static const int kFzWidth[6] = {11, 11, 11, 11, 10, 10};

// column splits per sweep: enough blocks for ~3 waves per SIMD (the kernel
// holds 166 VGPRs), at least 8 column tiles per split
static int fused_splits(int n_rows, int n_cols) {
    const int blocks = (n_rows + 127) / 128;
    const int ntiles = (n_cols + 31) / 32;
    const int want = (kFusedTargetBlocks + blocks - 1) / blocks;
    return std::max(1, std::min(std::min(want, kFusedMaxSplit), ntiles / 8));
}

template <int TERMS, bool EXACT>
static void launch_fused(const FusedArgs &fa, int n_split, hipStream_t s) {
    const dim3 grid((unsigned)((fa.n_rows + 127) / 128), (unsigned)n_split);
    switch (fa.kp / 16) {
        case 1: fused_lse_kernel<TERMS, 1, EXACT><<<grid, kFThreads, 0, s>>>(fa); break;
        case 2: fused_lse_kernel<TERMS, 2, EXACT><<<grid, kFThreads, 0, s>>>(fa); break;
        case 3: fused_lse_kernel<TERMS, 3, EXACT><<<grid, kFThreads, 0, s>>>(fa); break;
        case 4: fused_lse_kernel<TERMS, 4, EXACT><<<grid, kFThreads, 0, s>>>(fa); break;
        case 5: fused_lse_kernel<TERMS, 5, EXACT><<<grid, kFThreads, 0, s>>>(fa); break;
        case 6: fused_lse_kernel<TERMS, 6, EXACT><<<grid, kFThreads, 0, s>>>(fa); break;
        case 7: fused_lse_kernel<TERMS, 7, EXACT><<<grid, kFThreads, 0, s>>>(fa); break;
        default: fused_lse_kernel<TERMS, 8, EXACT><<<grid, kFThreads, 0, s>>>(fa); break;
    }
}

template <int TERMS, bool EXACT = false>
static int fused_sweep(const FusedArgs &fa, const double *xr, const double *xc, int d, hipStream_t s) {
    fused_clear_kernel<<<1, 1, 0, s>>>(fa.n_bad);
    const int n_split = fused_splits(fa.n_rows, fa.n_cols);
    launch_fused<TERMS, EXACT>(fa, n_split, s);
    if (n_split > 1) fused_finalize_kernel<<<(fa.n_rows + 255) / 256, 256, 0, s>>>(fa, n_split);
    fused_repair_kernel<<<64, 256, 0, s>>>(fa.bad_list, fa.n_bad, xr, xc, fa.n_cols, d, fa.h, fa.pot, fa.logw,
                                           fa.logw_dev, fa.log_n_other, fa.alpha, fa.eps, fa.k, fa.row_sums,
                                           fa.rs_div);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

// ---------------------------------------------------------------------------
// the problem helpers shared with the duality-gap schedule
// ---------------------------------------------------------------------------
static void problem_carve(Arena &ar, int64_t I, int64_t J, int64_t d, FusedProblem *fp) {
    const int64_t kp = round_up(d, 16);
    fp->mean = ar.take<double>(d);
    fp->xs = ar.take<double>((size_t)I * d);
    fp->ys = ar.take<double>((size_t)J * d);
    fp->nx = ar.take<double>(I);
    fp->ny = ar.take<double>(J);
    fp->zeros = ar.take<double>(std::max(I, J));
    for (int p = 0; p < 3; ++p) {
        fp->xp[p] = ar.take<__bf16>((size_t)I * kp);
        fp->yp[p] = ar.take<__bf16>((size_t)J * kp);
    }
    fp->bad = ar.take<int>(std::max(I, J));
    fp->nbad = ar.take<int>(4);
    fp->bad2 = ar.take<int>(std::max(I, J));
    fp->nbad2 = ar.take<int>(4);
    fp->part = ar.take<double>((size_t)kFusedMaxSplit * std::max(I, J));
    fp->part2 = ar.take<double>((size_t)kFusedMaxSplit * std::max(I, J));
    fp->sel = ar.take<FzSelect>(1);
    fp->hist = ar.take<unsigned long long>(2048);
}

size_t fused_problem_bytes(int64_t I, int64_t J, int64_t d) {
    Arena ar(nullptr, 0);
    FusedProblem fp;
    problem_carve(ar, I, J, d, &fp);
    return ar.off;
}

int fused_problem_setup(const double *X, const double *Y, int64_t I, int64_t J, int64_t d, int terms,
                        void *workspace, size_t workspace_bytes, FusedProblem *fp, size_t *used, hipStream_t s) {
    WOTK_REQUIRE(I > 0 && J > 0 && d > 0 && I < (1ll << 30) && J < (1ll << 30), "bad shape");
    WOTK_REQUIRE(d <= 128, "the fused solver keeps a 128-dimensional row tile in registers: d <= 128");
    WOTK_REQUIRE(terms == 3 || terms == 6, "terms must be 3 or 6");
    WOTK_REQUIRE(X && Y, "coordinates are required");
    Arena ar(workspace, workspace_bytes);
    problem_carve(ar, I, J, d, fp);
    if (!workspace || !ar.ok) return fail(WOTK_EWORKSPACE, "fused workspace too small");
    *used = ar.off;
    fp->X = X;
    fp->Y = Y;
    fp->I = (int)I;
    fp->J = (int)J;
    fp->d = (int)d;
    fp->kp = (int)round_up(d, 16);
    fp->terms = terms;

    // exact median of the FP64 distances, regenerated per pass
    const long long n = (long long)I * J;
    const long long rank = (n % 2) ? n / 2 : n / 2 - 1;
    fz_select_init_kernel<<<1, 256, 0, s>>>(fp->sel, rank, fp->hist);
    dim3 sgrid((unsigned)((J + 63) / 64), (unsigned)((I + 63) / 64));
    for (int p = 0; p < 6; ++p) {
        fz_select_kernel<0><<<sgrid, 256, 0, s>>>(X, Y, (int)I, (int)J, (int)d, kFzShift[p], kFzWidth[p], fp->sel,
                                                  fp->hist);
        fz_select_scan_kernel<<<1, 256, 0, s>>>(fp->hist, kFzShift[p], kFzWidth[p], p == 5, fp->sel);
    }
    if (n % 2 == 0) fz_select_kernel<1><<<sgrid, 256, 0, s>>>(X, Y, (int)I, (int)J, (int)d, 0, 0, fp->sel, fp->hist);
    fz_select_finish_kernel<<<1, 1, 0, s>>>(fp->sel, (int)(n % 2));
    WOTK_LAUNCH_CHECK();
    fp->median = &fp->sel->median;

    // scaled, centred coordinates and their bf16 pieces
    fz_mean_kernel<<<(unsigned)d, 256, 0, s>>>(X, (int)I, Y, (int)J, (int)d, fp->mean);
    const int np = terms == 3 ? 2 : 3;
    __bf16 *const *xp = fp->xp, *const *yp = fp->yp;
    fz_split_kernel<<<(unsigned)((I + 255) / 256), 256, 0, s>>>(X, (int)I, (int)d, fp->kp, fp->mean, fp->median, np,
                                                                xp[0], xp[1], xp[2], fp->xs, fp->nx);
    fz_split_kernel<<<(unsigned)((J + 255) / 256), 256, 0, s>>>(Y, (int)J, (int)d, fp->kp, fp->mean, fp->median, np,
                                                                yp[0], yp[1], yp[2], fp->ys, fp->ny);
    fz_zero_kernel<<<64, 256, 0, s>>>(fp->zeros, (int)std::max(I, J));
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

// FusedArgs of a sweep over source rows (over_targets = false) or target rows
static FusedArgs problem_args(const FusedProblem &fp, bool target_rows, const double *h, double *pot,
                              const double *logw, const double *logw_dev, double alpha, double eps, double *row_sums,
                              double rs_div, bool second_list) {
    __bf16 *const *A = target_rows ? fp.yp : fp.xp;
    __bf16 *const *B = target_rows ? fp.xp : fp.yp;
    const int nr = target_rows ? fp.J : fp.I, nc = target_rows ? fp.I : fp.J;
    return FusedArgs{A[0], A[1], A[2], B[0], B[1], B[2], target_rows ? fp.ny : fp.nx, target_rows ? fp.nx : fp.ny,
                     nr, nc, fp.kp, h, pot, logw, logw_dev, std::log((double)nc), alpha, eps, kLog2E / eps,
                     second_list ? fp.bad2 : fp.bad, second_list ? fp.nbad2 : fp.nbad, row_sums,
                     second_list ? fp.part2 : fp.part, rs_div};
}

template <bool EXACT>
static int problem_sweep(const FusedProblem &fp, const FusedArgs &fa, bool target_rows, hipStream_t s) {
    const double *xr = target_rows ? fp.ys : fp.xs, *xc = target_rows ? fp.xs : fp.ys;
    return fp.terms == 3 ? fused_sweep<3, EXACT>(fa, xr, xc, fp.d, s) : fused_sweep<6, EXACT>(fa, xr, xc, fp.d, s);
}

int fused_problem_iterate(const FusedProblem &fp, double *f, double *g, const double *logp,
                          const double *logq_dev, double lambda1, double lambda2, double eps, hipStream_t s) {
    const FusedArgs fr = problem_args(fp, false, g, f, logp, nullptr, lambda1 / (lambda1 + eps), eps, nullptr, 1.0,
                                      false);
    int rc = problem_sweep<false>(fp, fr, false, s);
    if (rc) return rc;
    const FusedArgs fc = problem_args(fp, true, f, g, nullptr, logq_dev, lambda2 / (lambda2 + eps), eps, nullptr, 1.0,
                                      false);
    return problem_sweep<false>(fp, fc, true, s);
}

int fused_problem_sums(const FusedProblem &fp, bool over_targets, const double *f, const double *g, double eps,
                       double div, bool exact, double *out, hipStream_t s) {
    // row-sum mode: the shift is the row's own potential, which is not updated
    const double *h = over_targets ? f : g;
    double *pot = const_cast<double *>(over_targets ? g : f);
    const FusedArgs fa = problem_args(fp, over_targets, h, pot, nullptr, nullptr, 0.0, eps, out, div, true);
    return exact ? problem_sweep<true>(fp, fa, over_targets, s) : problem_sweep<false>(fp, fa, over_targets, s);
}

}  // namespace wotk

using namespace wotk;

// solver state of the fixed-iteration schedule beside the problem buffers
struct FzState {
    double *f, *g, *logp, *misc;
};

static size_t fz_state_carve(Arena &ar, int64_t I, int64_t J, FzState *st) {
    st->f = ar.take<double>(I);
    st->g = ar.take<double>(J);
    st->logp = ar.take<double>(I);
    st->misc = ar.take<double>(8);
    return ar.off;
}

extern "C" size_t wotk_fused_workspace_size(int64_t I, int64_t J, int64_t d) {
    Arena ar(nullptr, 0);
    FzState st;
    return fused_problem_bytes(I, J, d) + 256 + fz_state_carve(ar, I, J, &st);
}

extern "C" int wotk_transport_stablev2_coords(const double *X, const double *Y, int64_t I, int64_t J, int64_t d,
                                              const double *G, double lambda1, double lambda2, double epsilon,
                                              int64_t scaling_iter, double tau, double epsilon0, int64_t extra_iter,
                                              int64_t inner_iter_max, int32_t terms, double *f_out, double *g_out,
                                              double *row_sums, double *median_out, void *workspace,
                                              size_t workspace_bytes, wotk_stats *stats, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    WOTK_REQUIRE(scaling_iter >= 0 && extra_iter >= 0 && scaling_iter + extra_iter > 0 && inner_iter_max > 0,
                 "bad iteration counts");
    WOTK_REQUIRE(G != nullptr, "G is required");
    (void)tau;
    FusedProblem fp;
    size_t used = 0;
    int rc = fused_problem_setup(X, Y, I, J, d, terms, workspace, workspace_bytes, &fp, &used, s);
    if (rc) return rc;
    used = (used + 255) & ~size_t(255);
    Arena ar(static_cast<char *>(workspace) + used, workspace_bytes > used ? workspace_bytes - used : 0);
    FzState w;
    fz_state_carve(ar, I, J, &w);
    if (!ar.ok) return fail(WOTK_EWORKSPACE, "fused workspace too small");
    if (median_out) WOTK_HIP(hipMemcpyAsync(median_out, fp.median, sizeof(double), hipMemcpyDeviceToDevice, s));
    fz_zero_kernel<<<64, 256, 0, s>>>(w.f, (int)I);
    fz_zero_kernel<<<64, 256, 0, s>>>(w.g, (int)J);
    fz_growth_kernel<<<1, 1024, 0, s>>>(G, (int)I, w.logp, w.misc);
    WOTK_LAUNCH_CHECK();

    auto reg = [&](int64_t m) { return (epsilon0 - epsilon) * std::exp(-(double)m) + epsilon; };
    hipEvent_t ev[2] = {nullptr, nullptr};
    struct Events {
        hipEvent_t *e;
        ~Events() {
            for (int i = 0; i < 2; ++i)
                if (e[i]) (void)hipEventDestroy(e[i]);
        }
    } guard{ev};
    WOTK_HIP(hipEventCreate(&ev[0]));
    WOTK_HIP(hipEventCreate(&ev[1]));
    WOTK_HIP(hipEventRecord(ev[0], s));
    const int64_t n_iter = scaling_iter + extra_iter;
    for (int64_t it = 0; it < n_iter; ++it) {
        const double eps = it < scaling_iter ? reg(it / inner_iter_max) : reg(scaling_iter / inner_iter_max);
        rc = fused_problem_iterate(fp, w.f, w.g, w.logp, w.misc, lambda1, lambda2, eps, s);
        if (rc) return rc;
    }
    WOTK_HIP(hipEventRecord(ev[1], s));
    const double e_end = reg(scaling_iter / inner_iter_max);
    if (row_sums) {
        // the returned plan is built with the eps after the last iteration
        rc = fused_problem_sums(fp, false, w.f, w.g, e_end, (double)J, false, row_sums, s);
        if (rc) return rc;
    }
    if (f_out) WOTK_HIP(hipMemcpyAsync(f_out, w.f, sizeof(double) * I, hipMemcpyDeviceToDevice, s));
    if (g_out) WOTK_HIP(hipMemcpyAsync(g_out, w.g, sizeof(double) * J, hipMemcpyDeviceToDevice, s));
    WOTK_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    WOTK_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    if (stats) {
        *stats = wotk_stats{};
        stats->iterations = n_iter;
        stats->sweeps = 2 * n_iter;
        stats->sweep_ms = ms;
        stats->epsilon_final = e_end;
    }
    return WOTK_OK;
}
