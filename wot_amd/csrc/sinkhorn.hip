// Log-domain stabilised Sinkhorn for wot's unbalanced entropic OT
// (wot/ot/optimal_transport.py:67-236) on MI355X.
//
// The reference iterates scalings a, b against a stabilised kernel
// K = exp((u - C + v)/eps) and absorbs them into the potentials u, v when they
// exceed tau or when eps changes.  In exact arithmetic every absorption is
// neutral: with f = u + eps log a and g = v + eps log b the reference's update
//   a = (p / K(b dy))^al1 exp(-u/(l1+eps)),  b = (q / K^T(a dx))^al2 exp(-v/(l2+eps))
// is exactly
//   f_i = al1 eps (log p_i + log J - LSE_j((g_j - C_ij)/eps))
//   g_j = al2 eps (log q_j + log I - LSE_i((f_i - C_ij)/eps)),
// and the plan is R = exp((f_i + g_j - C_ij)/eps).  The device keeps f, g in
// FP64 and evaluates each LSE as a streaming row reduction over FP32 cost rows
// (C for f, C^T for g): exponent arithmetic in FP64, exp2 in FP32 against a
// per-lane integer shift that tracks the running maximum, partial sums merged
// exactly (ldexp) in FP64.  Absorption events are tracked only where the
// reference's control flow observes them (the duality-gap solver's
// dual-evolution check).
#include <array>
#include <utility>

#include <mutex>

#include "common.h"
#include "fused_problem.h"

namespace wotk {

constexpr int kThreads = 256;
constexpr int kRows = 4;              // rows per block in the row kernels
constexpr double kSigInit = -1.0e6;   // "no term yet" shift (2^(y - shift) == 0)

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// Streaming read of 4 cost entries: the cost matrix is read once per sweep
// and is far larger than L2, so the loads are marked non-temporal (measured
// ~10 % faster than cached loads on a 900 MB sweep).
__device__ __forceinline__ float4 load_stream4(const float *p) {
    float4 v;
    v.x = __builtin_nontemporal_load(p);
    v.y = __builtin_nontemporal_load(p + 1);
    v.z = __builtin_nontemporal_load(p + 2);
    v.w = __builtin_nontemporal_load(p + 3);
    return v;
}

// ---------------------------------------------------------------------------
// block reductions (256 threads = 4 wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lse_merge(double &sg, double &sd, double og, double os) {
    const double ng = fmax(sg, og);
    sd = ldexp(sd, (int)(sg - ng)) + ldexp(os, (int)(og - ng));
    sg = ng;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// ---------------------------------------------------------------------------
// Row log-sum-exp sweep with the potential update fused in the epilogue.
//   lse_i = ln sum_j exp((h_j - A_ij)/eps)            (natural log)
//   pot_i = alpha * eps * (logw_i + log_n_other - lse_i)
// A: nrows x ncols FP32, lda % 4 == 0; h: FP64, padded with -inf to a
// multiple of 4.  k = log2(e)/eps.
// ---------------------------------------------------------------------------
struct RowEpilogue {
    const double *logw;      // per-row log weight, or nullptr -> *logw_dev
    const double *logw_dev;  // device scalar log weight
    double log_n_other;
    double alpha;
    double eps;
    double *pot_out;
    double *lse_out;         // optional
};

template <int ROWS, int THREADS>
__global__ __launch_bounds__(THREADS) void lse_rows_kernel(const float *__restrict__ A, int64_t lda,
                                                            int nrows, int ncols,
                                                            const double *__restrict__ h, double k,
                                                            RowEpilogue ep) {
    const int r0 = blockIdx.x * ROWS;
    const int tid = threadIdx.x;
    const float *rowp[ROWS];
    double sig[ROWS];
    float s[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        rowp[r] = A + (int64_t)min(r0 + r, nrows - 1) * lda;
        sig[r] = kSigInit;
        s[r] = 0.f;
    }
    for (int c = tid * 4; c < ncols; c += THREADS * 4) {
        const double2 h01 = *reinterpret_cast<const double2 *>(h + c);
        const double2 h23 = *reinterpret_cast<const double2 *>(h + c + 2);
        float4 a[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) a[r] = load_stream4(rowp[r] + c);
        const double hk0 = h01.x * k, hk1 = h01.y * k, hk2 = h23.x * k, hk3 = h23.y * k;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const double y0 = fma(-k, (double)a[r].x, hk0);
            const double y1 = fma(-k, (double)a[r].y, hk1);
            const double y2 = fma(-k, (double)a[r].z, hk2);
            const double y3 = fma(-k, (double)a[r].w, hk3);
            const double m = fmax(fmax(y0, y1), fmax(y2, y3));
            if (m > sig[r]) {  // new running maximum: move this lane's shift up
                const double ns = ceil(m);
                s[r] = ldexpf(s[r], (int)(sig[r] - ns));
                sig[r] = ns;
            }
            const float e0 = fast_exp2((float)(y0 - sig[r]));
            const float e1 = fast_exp2((float)(y1 - sig[r]));
            const float e2 = fast_exp2((float)(y2 - sig[r]));
            const float e3 = fast_exp2((float)(y3 - sig[r]));
            s[r] += (e0 + e1) + (e2 + e3);
        }
    }
    __shared__ double lsig[ROWS][THREADS / 64];
    __shared__ double lsum[ROWS][THREADS / 64];
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        double sg = sig[r], sd = (double)s[r];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double og = __shfl_xor(sg, off);
            const double os = __shfl_xor(sd, off);
            lse_merge(sg, sd, og, os);
        }
        if (lane == 0) {
            lsig[r][wave] = sg;
            lsum[r][wave] = sd;
        }
    }
    __syncthreads();
    if (tid < ROWS) {
        const int row = r0 + tid;
        double sg = lsig[tid][0], sd = lsum[tid][0];
#pragma unroll
        for (int w = 1; w < THREADS / 64; ++w) lse_merge(sg, sd, lsig[tid][w], lsum[tid][w]);
        if (row < nrows) {
            const double lse = (sg + log2(sd)) * kLn2;
            const double lw = ep.logw ? ep.logw[row] : *ep.logw_dev;
            ep.pot_out[row] = ep.alpha * ep.eps * (lw + ep.log_n_other - lse);
            if (ep.lse_out) ep.lse_out[row] = lse;
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 plan rows: R_ij = exp((f_i + g_j - A_ij)/eps) / div.  Optionally
// writes the plan (ldp), the row sums of R, and per-block partial sums
// [sum R*log R, sum R*A, sum exp(-A/eps)] for the duality gap.
// ---------------------------------------------------------------------------
struct PlanArgs {
    const float *A;
    int64_t lda;
    int nrows, ncols;
    const double *f;   // row potential
    const double *g;   // column potential
    double inv_eps;
    double div;        // plan divisor (J for the returned plan, 1 un-normalised)
    double *P;         // optional plan out
    int64_t ldp;
    double *row_sums;  // optional
    double *partials;  // optional: 3 doubles per block
    int want_ktilde;   // include sum exp(-A/eps) in partials[2]
};

// PLAN / PARTS / KT: compiled with or without the plan store, the gap
// partials and the K-tilde sum, so no pass carries a per-entry branch for
// another's output
template <int ROWS, bool PLAN, bool PARTS, bool KT>
__global__ __launch_bounds__(kThreads) void plan_rows_kernel(PlanArgs pa) {
    const int r0 = blockIdx.x * ROWS;
    const int tid = threadIdx.x;
    const float *rowp[ROWS];
    double fr[ROWS], acc[ROWS];
    double s_ry = 0.0, s_rc = 0.0, s_kt = 0.0;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int row = min(r0 + r, pa.nrows - 1);
        rowp[r] = pa.A + (int64_t)row * pa.lda;
        fr[r] = (r0 + r < pa.nrows) ? pa.f[row] : -INFINITY;  // duplicate tail rows contribute 0
        acc[r] = 0.0;
    }
    for (int c = tid * 4; c < pa.ncols; c += kThreads * 4) {
        double gv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) gv[q] = (c + q < pa.ncols) ? pa.g[c + q] : -INFINITY;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const float4 a4 = load_stream4(rowp[r] + c);
            const float av[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double a = (double)av[q];
                const double y = (fr[r] + gv[q] - a) * pa.inv_eps;
                const double R = exp(y);
                acc[r] += R;
                if (PLAN && r0 + r < pa.nrows && c + q < pa.ncols)
                    pa.P[(int64_t)(r0 + r) * pa.ldp + c + q] = R / pa.div;
                if (PARTS) {
                    if (R > 0.0) s_ry = fma(R, y, s_ry);
                    s_rc = fma(R, a, s_rc);
                    if (KT && r0 + r < pa.nrows && c + q < pa.ncols)
                        s_kt += exp(-a * pa.inv_eps);
                }
            }
        }
    }
    __shared__ double red[ROWS + 3][kThreads / 64];
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const double v = wave_sum(acc[r]);
        if (lane == 0) red[r][wave] = v;
    }
    if (PARTS) {
        const double a0 = wave_sum(s_ry), a1 = wave_sum(s_rc), a2 = wave_sum(s_kt);
        if (lane == 0) {
            red[ROWS][wave] = a0;
            red[ROWS + 1][wave] = a1;
            red[ROWS + 2][wave] = a2;
        }
    }
    __syncthreads();
    if (tid < ROWS + 3) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) v += red[tid][w];
        if (tid < ROWS) {
            if (pa.row_sums && r0 + tid < pa.nrows) pa.row_sums[r0 + tid] = v / pa.div;
        } else if (PARTS) {
            pa.partials[(int64_t)blockIdx.x * 3 + (tid - ROWS)] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// small vector kernels (one block unless noted)
// ---------------------------------------------------------------------------
// f = 0 on [0, n), -inf on [n, n4): potentials start at u = v = 0, a = b = 1
__global__ void fill_pot_kernel(double *x, int n, int n4, double v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += gridDim.x * blockDim.x)
        x[i] = (i < n) ? v : -INFINITY;
}

// logp_i = log G_i ; misc[0] = log(mean G) ; misc[1] = mean G
__global__ __launch_bounds__(1024) void growth_prep_kernel(const double *__restrict__ G, int n,
                                                           double *__restrict__ logp,
                                                           double *__restrict__ misc) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double gi = G[i];
        logp[i] = log(gi);
        acc += gi;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        const double mean = t / n;
        misc[0] = log(mean);
        misc[1] = mean;
    }
}

// tau-triggered absorption (optimal_transport.py:137-141): a = exp((f-u)/eps),
// b = exp((g-v)/eps); if max(a, b) > tau then u = f, v = g.  counter[0]++.
__global__ __launch_bounds__(1024) void stabilise_kernel(const double *__restrict__ f,
                                                         double *__restrict__ u, int n_i,
                                                         const double *__restrict__ g,
                                                         double *__restrict__ v, int n_j,
                                                         double inv_eps, double tau,
                                                         long long *__restrict__ counter) {
    __shared__ double red[16];
    __shared__ int fire;
    // max over the exponents, then one exp (exp is increasing)
    double m = -INFINITY;
    for (int i = threadIdx.x; i < n_i; i += blockDim.x) m = fmax(m, (f[i] - u[i]) * inv_eps);
    for (int j = threadIdx.x; j < n_j; j += blockDim.x) m = fmax(m, (g[j] - v[j]) * inv_eps);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) t = fmax(t, red[w]);
        fire = exp(t) > tau;
        if (fire) counter[0] += 1;
    }
    __syncthreads();
    if (fire) {
        for (int i = threadIdx.x; i < n_i; i += blockDim.x) u[i] = f[i];
        for (int j = threadIdx.x; j < n_j; j += blockDim.x) v[j] = g[j];
    }
}

__global__ void zero_ll_kernel(long long *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// state of the duality-gap solver's last scaling (f, g, u, v, stabilisation
// count) to its snapshot (back = 0) or from it (back = 1: rolls back a batch
// that ran ahead of its gap check)
__global__ void snapshot_kernel(double *__restrict__ f, double *__restrict__ g, double *__restrict__ u,
                                double *__restrict__ v, double *__restrict__ fs, double *__restrict__ gs,
                                double *__restrict__ us, double *__restrict__ vs, int n_i, int n_j,
                                long long *__restrict__ counter, int back) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_i; i += stride) {
        if (back) {
            f[i] = fs[i];
            u[i] = us[i];
        } else {
            fs[i] = f[i];
            us[i] = u[i];
        }
    }
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n_j; j += stride) {
        if (back) {
            g[j] = gs[j];
            v[j] = vs[j];
        } else {
            gs[j] = g[j];
            vs[j] = v[j];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (back) counter[0] = counter[1];
        else counter[1] = counter[0];
    }
}

__global__ void copy2_kernel(double *__restrict__ d0, const double *__restrict__ s0, int n0,
                             double *__restrict__ d1, const double *__restrict__ s1, int n1) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n0; i += stride) d0[i] = s0[i];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n1; i += stride) d1[i] = s1[i];
}

// dual-variable evolution check (optimal_transport.py:157-160):
//   max(||A - A_old|| / (1 + ||A||), ||B - B_old|| / (1 + ||B||))
// with A = exp(f/eps) and A_old = old_a * exp(u/eps) = exp((fp - up + u)/eps).
__global__ __launch_bounds__(1024) void evolution_kernel(const double *__restrict__ f,
                                                         const double *__restrict__ fp,
                                                         const double *__restrict__ up,
                                                         const double *__restrict__ u, int n_i,
                                                         const double *__restrict__ g,
                                                         const double *__restrict__ gp,
                                                         const double *__restrict__ vp,
                                                         const double *__restrict__ v, int n_j,
                                                         double inv_eps, double *__restrict__ out) {
    __shared__ double red[4][16];
    double d_a = 0.0, n_a = 0.0, d_b = 0.0, n_b = 0.0;
    for (int i = threadIdx.x; i < n_i; i += blockDim.x) {
        const double A = exp(f[i] * inv_eps);
        const double Ao = exp((fp[i] - up[i]) * inv_eps) * exp(u[i] * inv_eps);
        d_a += (A - Ao) * (A - Ao);
        n_a += A * A;
    }
    for (int j = threadIdx.x; j < n_j; j += blockDim.x) {
        const double B = exp(g[j] * inv_eps);
        const double Bo = exp((gp[j] - vp[j]) * inv_eps) * exp(v[j] * inv_eps);
        d_b += (B - Bo) * (B - Bo);
        n_b += B * B;
    }
    d_a = wave_sum(d_a);
    n_a = wave_sum(n_a);
    d_b = wave_sum(d_b);
    n_b = wave_sum(n_b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = d_a;
        red[1][w] = n_a;
        red[2][w] = d_b;
        red[3][w] = n_b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[4] = {0, 0, 0, 0};
        for (int q = 0; q < 4; ++q)
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t[q] += red[q][k];
        const double ga = sqrt(t[0]) / (1.0 + sqrt(t[1]));
        const double gb = sqrt(t[2]) / (1.0 + sqrt(t[3]));
        out[0] = (gb > ga) ? gb : ga;  // python max(ga, gb): first argument unless gb > ga
    }
}

// primal / dual / gap from the row sums r (I), column sums c (J) of the
// un-normalised plan and the block partials (optimal_transport.py:45-62,152-156).
struct GapArgs {
    const double *r, *c, *f, *g, *G, *misc, *partials;
    int nparts;
    int n_i, n_j;
    double lambda1, lambda2, eps;
    double sum_ktilde;
    const double *sum_ktilde_dev;  // if non-null overrides sum_ktilde
    double *out;                   // [primal, dual, gap]
};

__global__ __launch_bounds__(1024) void gap_finish_kernel(GapArgs ga) {
    __shared__ double red[7][16];
    const double q = ga.misc[1];
    const double inv_i = 1.0 / ga.n_i, inv_j = 1.0 / ga.n_j;
    double f1 = 0.0, f2 = 0.0, s1 = 0.0, s2 = 0.0, sr = 0.0, p_ry = 0.0, p_rc = 0.0;
    // block partials: thread t takes blocks t, t + blockDim, ... (fixed order: deterministic)
    for (int b = threadIdx.x; b < ga.nparts; b += blockDim.x) {
        p_ry += ga.partials[3 * b];
        p_rc += ga.partials[3 * b + 1];
    }
    for (int i = threadIdx.x; i < ga.n_i; i += blockDim.x) {
        const double p = ga.G[i];
        const double x = ga.r[i] * inv_j;  // R.dot(dy)
        f1 += inv_i * (x * log(x / p) - x + p);
        s1 += (p * inv_i) * (exp(-ga.f[i] / ga.lambda1) - 1.0);
        sr += ga.r[i];
    }
    for (int j = threadIdx.x; j < ga.n_j; j += blockDim.x) {
        const double x = ga.c[j] * inv_i;  // R.T.dot(dx)
        f2 += inv_j * (x * log(x / q) - x + q);
        s2 += (q * inv_j) * (exp(-ga.g[j] / ga.lambda2) - 1.0);
    }
    f1 = wave_sum(f1);
    f2 = wave_sum(f2);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    sr = wave_sum(sr);
    p_ry = wave_sum(p_ry);
    p_rc = wave_sum(p_rc);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = f1;
        red[1][w] = f2;
        red[2][w] = s1;
        red[3][w] = s2;
        red[4][w] = sr;
        red[5][w] = p_ry;
        red[6][w] = p_rc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 7; ++k)
            for (int x = 0; x < (int)(blockDim.x >> 6); ++x) t[k] += red[k][x];
        const double s_ry = t[5], s_rc = t[6];
        const double skt = ga.sum_ktilde_dev ? *ga.sum_ktilde_dev : ga.sum_ktilde;
        const double ij = (double)ga.n_i * (double)ga.n_j;
        const double pri = ga.lambda1 * t[0] + ga.lambda2 * t[1] +
                           (ga.eps * (s_ry - t[4] + skt) + s_rc) / ij;
        const double dua = -ga.lambda1 * t[2] - ga.lambda2 * t[3] - ga.eps * (t[4] - skt) / ij;
        ga.out[0] = pri;
        ga.out[1] = dua;
        ga.out[2] = (pri - dua) / fabs(pri);
    }
}

// sum of the K-tilde partials (slot 2 of each block) -> out
// (one block of 1024 threads; thread t takes blocks t, t + 1024, ...: fixed order)
__global__ __launch_bounds__(1024) void sum_slot_kernel(const double *__restrict__ partials, int nparts,
                                                        int slot, double *__restrict__ out) {
    __shared__ double red[16];
    double t = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) t += partials[3 * b + slot];
    t = wave_sum(t);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int x = 0; x < (int)(blockDim.x >> 6); ++x) v += red[x];
        *out = v;
    }
}

// ---------------------------------------------------------------------------
// host-side orchestration
// ---------------------------------------------------------------------------
struct Problem {
    const float *C;
    int64_t ldc;
    const float *CT;
    int64_t ldct;
    int n_i, n_j;
    const double *G;
};

struct Work {
    double *f, *g, *logp, *misc, *r, *c, *u, *v, *fp, *gp, *up, *vp, *partials, *gapout;
    double *fs, *gs, *us, *vs;  // snapshot of the last scaling's state (gap pass ahead of the next batch)
    long long *counter;
    int nparts;
};

static size_t carve(Arena &ar, int64_t I, int64_t J, Work *w) {
    const int64_t I4 = round_up(I, 4), J4 = round_up(J, 4);
    w->f = ar.take<double>(I4);
    w->g = ar.take<double>(J4);
    w->logp = ar.take<double>(I4);
    w->misc = ar.take<double>(16);
    w->r = ar.take<double>(I4);
    w->c = ar.take<double>(J4);
    w->u = ar.take<double>(I4);
    w->v = ar.take<double>(J4);
    w->fp = ar.take<double>(I4);
    w->gp = ar.take<double>(J4);
    w->up = ar.take<double>(I4);
    w->vp = ar.take<double>(J4);
    w->nparts = (int)((std::max(I, J) + kRows - 1) / kRows);
    w->partials = ar.take<double>((size_t)3 * w->nparts);
    w->gapout = ar.take<double>(8);
    w->fs = ar.take<double>(I4);
    w->gs = ar.take<double>(J4);
    w->us = ar.take<double>(I4);
    w->vs = ar.take<double>(J4);
    w->counter = ar.take<long long>(4);
    return ar.off;
}

static int check_problem(const Problem &pb) {
    WOTK_REQUIRE(pb.n_i > 0 && pb.n_j > 0, "empty cost matrix");
    WOTK_REQUIRE(pb.ldc % 4 == 0 && pb.ldc >= round_up(pb.n_j, 4), "ldc must be a multiple of 4 >= J");
    WOTK_REQUIRE(pb.ldct % 4 == 0 && pb.ldct >= round_up(pb.n_i, 4), "ldct must be a multiple of 4 >= I");
    WOTK_REQUIRE(((uintptr_t)pb.C & 15) == 0 && ((uintptr_t)pb.CT & 15) == 0,
                 "C and CT must be 16-byte aligned");
    WOTK_REQUIRE(pb.G != nullptr, "G is required");
    return WOTK_OK;
}

static int init_state(const Problem &pb, Work &w, hipStream_t s) {
    const int I4 = (int)round_up(pb.n_i, 4), J4 = (int)round_up(pb.n_j, 4);
    fill_pot_kernel<<<(I4 + 255) / 256, 256, 0, s>>>(w.f, pb.n_i, I4, 0.0);
    fill_pot_kernel<<<(J4 + 255) / 256, 256, 0, s>>>(w.g, pb.n_j, J4, 0.0);
    fill_pot_kernel<<<(I4 + 255) / 256, 256, 0, s>>>(w.u, pb.n_i, I4, 0.0);
    fill_pot_kernel<<<(J4 + 255) / 256, 256, 0, s>>>(w.v, pb.n_j, J4, 0.0);
    growth_prep_kernel<<<1, 1024, 0, s>>>(pb.G, pb.n_i, w.logp, w.misc);
    zero_ll_kernel<<<1, 64, 0, s>>>(w.counter, 4);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

// one launcher per compiled tile variant, in the order of kRowSweepTiles
typedef void (*LseLaunch)(const float *, int64_t, int, int, const double *, double, const RowEpilogue &,
                          hipStream_t);

template <int V>
static void lse_variant(const float *A, int64_t lda, int nrows, int ncols, const double *h, double k,
                        const RowEpilogue &ep, hipStream_t s) {
    constexpr int R = kRowSweepTiles[V].rows, T = kRowSweepTiles[V].threads;
    lse_rows_kernel<R, T><<<(nrows + R - 1) / R, T, 0, s>>>(A, lda, nrows, ncols, h, k, ep);
}

template <int... V>
static constexpr std::array<LseLaunch, sizeof...(V)> lse_table(std::integer_sequence<int, V...>) {
    return {lse_variant<V>...};
}

static constexpr auto kLseLaunch = lse_table(std::make_integer_sequence<int, kNumRowSweepTiles>());

static int launch_lse(const float *A, int64_t lda, int nrows, int ncols, const double *h, double eps,
                      const RowEpilogue &ep, hipStream_t s) {
    const double k = kLog2E / eps;
    kLseLaunch[tile_for(WOTK_KERNEL_ROW_SWEEP, nrows, ncols, 1)](A, lda, nrows, ncols, h, k, ep, s);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

// one (a, b) iteration at eps: f from g over C, then g from f over C^T
static int iterate(const Problem &pb, Work &w, double lambda1, double lambda2, double eps,
                   hipStream_t s) {
    RowEpilogue er{w.logp, nullptr, std::log((double)pb.n_j), lambda1 / (lambda1 + eps), eps, w.f,
                   nullptr};
    int rc = launch_lse(pb.C, pb.ldc, pb.n_i, pb.n_j, w.g, eps, er, s);
    if (rc) return rc;
    RowEpilogue ec{nullptr, w.misc, std::log((double)pb.n_i), lambda2 / (lambda2 + eps), eps, w.g,
                   nullptr};
    return launch_lse(pb.CT, pb.ldct, pb.n_j, pb.n_i, w.f, eps, ec, s);
}

static int launch_plan(const float *A, int64_t lda, int nrows, int ncols, const double *f,
                       const double *g, double eps, double div, double *P, int64_t ldp,
                       double *row_sums, double *partials, int want_kt, hipStream_t s) {
    PlanArgs pa{A, lda, nrows, ncols, f, g, 1.0 / eps, div, P, ldp, row_sums, partials, want_kt};
    const int nb = (nrows + kRows - 1) / kRows;
    if (P && partials && want_kt) plan_rows_kernel<kRows, true, true, true><<<nb, kThreads, 0, s>>>(pa);
    else if (P && partials) plan_rows_kernel<kRows, true, true, false><<<nb, kThreads, 0, s>>>(pa);
    else if (P) plan_rows_kernel<kRows, true, false, false><<<nb, kThreads, 0, s>>>(pa);
    else if (partials && want_kt) plan_rows_kernel<kRows, false, true, true><<<nb, kThreads, 0, s>>>(pa);
    else if (partials) plan_rows_kernel<kRows, false, true, false><<<nb, kThreads, 0, s>>>(pa);
    else plan_rows_kernel<kRows, false, false, false><<<nb, kThreads, 0, s>>>(pa);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

// row sums of exp((f_i + g_j - A_ij)/eps) (no plan written); shared with the adjoint
int launch_plan_sums(const float *A, int64_t lda, int nrows, int ncols, const double *f, const double *g,
                     double eps, double *row_sums, hipStream_t s) {
    return launch_plan(A, lda, nrows, ncols, f, g, eps, 1.0, nullptr, 0, row_sums, nullptr, 0, s);
}

// primal/dual/gap at potentials f, g and eps -> w.gapout[0..2] (device)
static int gap_pass(const Problem &pb, Work &w, const double *f, const double *g, double lambda1,
                    double lambda2, double eps, const double *sum_kt_dev, hipStream_t s) {
    int rc = launch_plan(pb.C, pb.ldc, pb.n_i, pb.n_j, f, g, eps, 1.0, nullptr, 0, w.r, w.partials, 0, s);
    if (rc) return rc;
    rc = launch_plan(pb.CT, pb.ldct, pb.n_j, pb.n_i, g, f, eps, 1.0, nullptr, 0, w.c, nullptr, 0, s);
    if (rc) return rc;
    GapArgs ga{w.r, w.c, f, g, pb.G, w.misc, w.partials, (pb.n_i + kRows - 1) / kRows, pb.n_i,
               pb.n_j, lambda1, lambda2, eps, 0.0, sum_kt_dev, w.gapout};
    gap_finish_kernel<<<1, 1024, 0, s>>>(ga);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

// sum_ij exp(-C_ij/eps) -> *out (device)
static int ktilde_sum(const Problem &pb, Work &w, double eps, double *out, hipStream_t s) {
    // the plan kernel accumulates sum exp(-C/eps) in partial slot 2 when asked
    int rc = launch_plan(pb.C, pb.ldc, pb.n_i, pb.n_j, w.f, w.g, eps, 1.0, nullptr, 0, nullptr,
                         w.partials, 1, s);
    if (rc) return rc;
    sum_slot_kernel<<<1, 1024, 0, s>>>(w.partials, (pb.n_i + kRows - 1) / kRows, 2, out);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

static int finish_outputs(const Problem &pb, Work &w, double eps, double div, double *plan,
                          int64_t ldp, double *row_sums, double *f_out, double *g_out,
                          hipStream_t s) {
    if (plan || row_sums) {
        int rc = launch_plan(pb.C, pb.ldc, pb.n_i, pb.n_j, w.f, w.g, eps, div, plan, ldp, row_sums,
                             nullptr, 0, s);
        if (rc) return rc;
    }
    if (f_out) WOTK_HIP(hipMemcpyAsync(f_out, w.f, sizeof(double) * pb.n_i, hipMemcpyDeviceToDevice, s));
    if (g_out) WOTK_HIP(hipMemcpyAsync(g_out, w.g, sizeof(double) * pb.n_j, hipMemcpyDeviceToDevice, s));
    return WOTK_OK;
}

}  // namespace wotk

using namespace wotk;

extern "C" size_t wotk_solver_workspace_size(int64_t I, int64_t J) {
    Arena ar(nullptr, 0);
    Work w;
    return carve(ar, I, J, &w);
}

static int setup(const float *C, int64_t ldc, const float *CT, int64_t ldct, int64_t I, int64_t J,
                 const double *G, void *workspace, size_t ws_bytes, Problem *pb, Work *w,
                 hipStream_t s) {
    WOTK_REQUIRE(I > 0 && J > 0 && I < (1ll << 31) && J < (1ll << 31), "bad shape %lld x %lld",
                 (long long)I, (long long)J);
    *pb = Problem{C, ldc, CT, ldct, (int)I, (int)J, G};
    int rc = check_problem(*pb);
    if (rc) return rc;
    Arena ar(workspace, ws_bytes);
    carve(ar, I, J, w);
    if (!workspace || !ar.ok) return fail(WOTK_EWORKSPACE, "solver workspace too small");
    return init_state(*pb, *w, s);
}

// the two events bracketing a sweep loop; released on every exit path
struct EventPair {
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~EventPair() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// the fixed-iteration schedule on a set-up problem (shared by the workspace
// and the scratch entry points); `timed` brackets the sweeps with events
static int stablev2_run(const Problem &pb, Work &w, double lambda1, double lambda2, double epsilon,
                        int64_t scaling_iter, double epsilon0, int64_t extra_iter, int64_t inner_iter_max,
                        double *plan, int64_t ldp, double *row_sums, double *f_out, double *g_out, bool timed,
                        wotk_stats *stats, hipStream_t s) {
    WOTK_REQUIRE(scaling_iter >= 0 && extra_iter >= 0 && scaling_iter + extra_iter > 0,
                 "at least one scaling iteration is required");
    WOTK_REQUIRE(inner_iter_max > 0, "inner_iter_max must be positive");
    WOTK_REQUIRE(!plan || ldp >= pb.n_j, "ldp < J");
    auto reg = [&](int64_t n) { return (epsilon0 - epsilon) * std::exp(-(double)n) + epsilon; };
    // the sweep loop is bracketed by two events on the launch stream so the
    // caller gets the in-situ device time of the 2*(iterations) LSE sweeps
    EventPair evs;
    if (timed) {
        WOTK_HIP(hipEventCreate(&evs.ev[0]));
        WOTK_HIP(hipEventCreate(&evs.ev[1]));
        WOTK_HIP(hipEventRecord(evs.ev[0], s));
    }
    int rc = 0;
    for (int64_t n = 0; n < scaling_iter; ++n) {
        rc = iterate(pb, w, lambda1, lambda2, reg(n / inner_iter_max), s);
        if (rc) return rc;
    }
    const double eps_end = reg(scaling_iter / inner_iter_max);
    for (int64_t n = 0; n < extra_iter; ++n) {
        rc = iterate(pb, w, lambda1, lambda2, eps_end, s);
        if (rc) return rc;
    }
    if (timed) WOTK_HIP(hipEventRecord(evs.ev[1], s));
    rc = finish_outputs(pb, w, eps_end, (double)pb.n_j, plan, ldp, row_sums, f_out, g_out, s);
    if (rc) return rc;
    WOTK_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    if (timed) WOTK_HIP(hipEventElapsedTime(&ms, evs.ev[0], evs.ev[1]));
    if (stats) {
        *stats = wotk_stats{};
        stats->iterations = scaling_iter + extra_iter;
        stats->epsilon_final = eps_end;
        stats->sweeps = 2 * (scaling_iter + extra_iter);
        stats->sweep_ms = timed ? ms : 0.0;
    }
    return WOTK_OK;
}

extern "C" int wotk_transport_stablev2(const float *C, int64_t ldc, const float *CT, int64_t ldct,
                                       int64_t I, int64_t J, const double *G, double lambda1,
                                       double lambda2, double epsilon, int64_t scaling_iter,
                                       double tau, double epsilon0, int64_t extra_iter,
                                       int64_t inner_iter_max, double *plan, int64_t ldp,
                                       double *row_sums, double *f_out, double *g_out,
                                       void *workspace, size_t workspace_bytes, wotk_stats *stats,
                                       void *stream) {
    hipStream_t s = (hipStream_t)stream;
    Problem pb;
    Work w;
    int rc = setup(C, ldc, CT, ldct, I, J, G, workspace, workspace_bytes, &pb, &w, s);
    if (rc) return rc;
    (void)tau;  // absorption is exact in the log domain; tau only changes the FP64 reference's rounding
    return stablev2_run(pb, w, lambda1, lambda2, epsilon, scaling_iter, epsilon0, extra_iter, inner_iter_max, plan,
                        ldp, row_sums, f_out, g_out, true, stats, s);
}

// ---------------------------------------------------------------------------
// The same solve for callers that cannot allocate: its state (potentials,
// partial sums, counters) lives in one file-scope device array per GPU
// instead of a caller workspace.  Every call on a GPU shares that GPU's array
// (calls that overlap on one GPU race); the solver initialises every part of
// the state it reads, so the previous call's contents do not matter.
constexpr size_t kScratchBytes = (size_t)16 << 20;  // 16 MiB: I, J up to ~95,000
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static void *g_scratch_ptr[kMaxDevices];  // device address of `scratch` on each GPU (host-side)
static std::mutex g_scratch_mutex;

// the device address of this GPU's `scratch`, looked up once per device with
// that device current (the array's host-side name is not a device address)
static int scratch_ptr(void **out) {
    int dev = 0;
    WOTK_HIP(hipGetDevice(&dev));
    WOTK_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        WOTK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        WOTK_REQUIRE(p != nullptr, "no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return WOTK_OK;
}

extern "C" size_t wotk_scratch_capacity(void) { return kScratchBytes; }

extern "C" int wotk_transport_stablev2_noalloc(const float *C, int64_t ldc, const float *CT, int64_t ldct,
                                               int64_t I, int64_t J, const double *G, double lambda1,
                                               double lambda2, double epsilon, int64_t scaling_iter,
                                               double tau, double epsilon0, int64_t extra_iter,
                                               int64_t inner_iter_max, double *plan, int64_t ldp,
                                               double *row_sums, double *f_out, double *g_out,
                                               wotk_stats *stats, void *stream) {
    WOTK_REQUIRE(I > 0 && J > 0 && I < (1ll << 31) && J < (1ll << 31), "bad shape %lld x %lld", (long long)I,
                 (long long)J);
    const size_t need = wotk_solver_workspace_size(I, J);
    WOTK_REQUIRE(need <= kScratchBytes,
                 "%lld x %lld needs %zu bytes of solver state; the no-allocation entry point holds %zu "
                 "(use wotk_transport_stablev2 with a workspace)",
                 (long long)I, (long long)J, need, kScratchBytes);
    hipStream_t s = (hipStream_t)stream;
    void *ws = nullptr;
    int rc = scratch_ptr(&ws);
    if (rc) return rc;
    Problem pb;
    Work w;
    rc = setup(C, ldc, CT, ldct, I, J, G, ws, kScratchBytes, &pb, &w, s);
    if (rc) return rc;
    (void)tau;
    return stablev2_run(pb, w, lambda1, lambda2, epsilon, scaling_iter, epsilon0, extra_iter, inner_iter_max, plan,
                        ldp, row_sums, f_out, g_out, false, stats, s);
}

// second stream and event of the duality-gap solver's last scaling; drained
// and released on every exit path
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t snap = nullptr;
    ~SideStream() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        if (snap) (void)hipEventDestroy(snap);
    }
};

// The duality-gap schedule (optimal_transport.py:67-164) over a backend that
// owns the cost: B::iterate (one (a, b) iteration on w.f, w.g), B::gap_pass
// (primal, dual, gap at given potentials -> w.gapout), B::ktilde_sum and
// B::finish (outputs at eps, plan divided by div).
template <class B>
static int gap_schedule(B &be, const Problem &pb, Work &w, double lambda1, double lambda2, double epsilon,
                        int64_t batch_size, double tolerance, double tau, double epsilon0, int64_t max_iter,
                        wotk_stats *stats, hipStream_t s) {
    WOTK_REQUIRE(batch_size > 0, "batch_size must be positive");
    const int n_scal = 5;
    const double factor = std::exp(-std::log(epsilon) / n_scal);
    double eps_i = epsilon0 * factor;
    int64_t it = 0, checks = 0;
    double gap = INFINITY;
    double host_out[3] = {0, 0, 0};
    double *sum_kt = w.misc + 4;
    double inv_eps = 1.0 / eps_i;
    int rc = 0;
    // one (a, b) iteration plus the tau check (optimal_transport.py:131-141)
    auto one_iteration = [&](double eps) -> int {
        int r = be.iterate(lambda1, lambda2, eps, s);
        if (r) return r;
        stabilise_kernel<<<1, 1024, 0, s>>>(w.f, w.u, pb.n_i, w.g, w.v, pb.n_j, inv_eps, tau, w.counter);
        WOTK_LAUNCH_CHECK();
        return WOTK_OK;
    };
    // early, un-normalised return (optimal_transport.py:143-145)
    auto max_iter_return = [&](double eps) -> int {
        int r = be.finish(eps, 1.0, s);
        if (r) return r;
        long long nst = 0;
        WOTK_HIP(hipMemcpyAsync(&nst, w.counter, sizeof(long long), hipMemcpyDeviceToHost, s));
        WOTK_HIP(hipStreamSynchronize(s));
        if (stats) {
            *stats = wotk_stats{};
            stats->iterations = it;
            stats->stabilisations = nst;
            stats->gap_checks = checks;
            stats->gap = gap;
            stats->epsilon_final = eps;
            stats->max_iter_reached = 1;
        }
        return WOTK_OK;
    };
    for (int e = 0; e <= n_scal; ++e) {
        gap = INFINITY;
        // absorb: with a, b folded into the potentials, u = f and v = g
        copy2_kernel<<<64, 256, 0, s>>>(w.u, w.f, pb.n_i, w.v, w.g, pb.n_j);
        WOTK_LAUNCH_CHECK();
        eps_i = eps_i / factor;
        inv_eps = 1.0 / eps_i;
        if (e == n_scal) {
            rc = be.ktilde_sum(eps_i, sum_kt, s);
            if (rc) return rc;
        }
        if (e < n_scal) {
            // dual-variable evolution checks every 5 iterations, threshold 1e-6
            while (gap > 1e-6) {
                for (int64_t t = 0; t < 5; ++t) {
                    ++it;
                    if (t == 4) {
                        copy2_kernel<<<64, 256, 0, s>>>(w.fp, w.f, pb.n_i, w.up, w.u, pb.n_i);
                        copy2_kernel<<<64, 256, 0, s>>>(w.gp, w.g, pb.n_j, w.vp, w.v, pb.n_j);
                        WOTK_LAUNCH_CHECK();
                    }
                    rc = one_iteration(eps_i);
                    if (rc) return rc;
                    if (it >= max_iter) return max_iter_return(eps_i);
                }
                ++checks;
                evolution_kernel<<<1, 1024, 0, s>>>(w.f, w.fp, w.up, w.u, pb.n_i, w.g, w.gp, w.vp, w.v,
                                                    pb.n_j, inv_eps, w.gapout);
                WOTK_LAUNCH_CHECK();
                WOTK_HIP(hipMemcpyAsync(host_out, w.gapout, sizeof(double), hipMemcpyDeviceToHost, s));
                WOTK_HIP(hipStreamSynchronize(s));
                gap = host_out[0];
            }
            continue;
        }
        // Last scaling: a primal/dual gap pass after every batch of batch_size
        // iterations.  The pass (FP64 exp per entry, VALU-bound) runs on a side
        // stream over a snapshot of the potentials while the next batch runs
        // ahead on the caller's stream; when the gap has converged, the batch
        // that ran ahead is rolled back from the snapshot, so the result is that
        // of the sequential order.  No batch runs ahead when it would reach
        // max_iter (its early return needs the state at max_iter).
        SideStream side;
        WOTK_HIP(hipStreamCreateWithFlags(&side.s, hipStreamNonBlocking));
        WOTK_HIP(hipEventCreateWithFlags(&side.snap, hipEventDisableTiming));
        const int snap_blocks = (int)std::min<int64_t>(256, (std::max(pb.n_i, pb.n_j) + 255) / 256);
        int64_t ahead = 0;  // iterations of the batch launched ahead of the last gap check
        while (gap > tolerance) {
            const int64_t nb = batch_size;
            if (ahead) {
                it += ahead;  // the batch that ran ahead is this batch
                ahead = 0;
            } else {
                for (int64_t t = 0; t < nb; ++t) {
                    ++it;
                    rc = one_iteration(eps_i);
                    if (rc) return rc;
                    if (it >= max_iter) return max_iter_return(eps_i);
                }
            }
            ++checks;
            snapshot_kernel<<<snap_blocks, 256, 0, s>>>(w.f, w.g, w.u, w.v, w.fs, w.gs, w.us, w.vs, pb.n_i,
                                                        pb.n_j, w.counter, 0);
            WOTK_LAUNCH_CHECK();
            WOTK_HIP(hipEventRecord(side.snap, s));
            WOTK_HIP(hipStreamWaitEvent(side.s, side.snap, 0));
            rc = be.gap_pass(w.fs, w.gs, lambda1, lambda2, eps_i, sum_kt, side.s);
            if (rc) return rc;
            if (it + nb < max_iter) {
                for (int64_t t = 0; t < nb; ++t) {
                    rc = one_iteration(eps_i);
                    if (rc) return rc;
                }
                ahead = nb;
            }
            WOTK_HIP(hipMemcpyAsync(host_out, w.gapout, 3 * sizeof(double), hipMemcpyDeviceToHost, side.s));
            WOTK_HIP(hipStreamSynchronize(side.s));
            gap = host_out[2];
        }
        if (ahead) {
            snapshot_kernel<<<snap_blocks, 256, 0, s>>>(w.f, w.g, w.u, w.v, w.fs, w.gs, w.us, w.vs, pb.n_i,
                                                        pb.n_j, w.counter, 1);
            WOTK_LAUNCH_CHECK();
        }
    }
    long long nst = 0;
    WOTK_HIP(hipMemcpyAsync(&nst, w.counter, sizeof(long long), hipMemcpyDeviceToHost, s));
    WOTK_HIP(hipStreamSynchronize(s));
    if (stats) {
        *stats = wotk_stats{};
        stats->iterations = it;
        stats->stabilisations = nst;
        stats->gap_checks = checks;
        stats->gap = gap;
        stats->epsilon_final = eps_i;
        stats->primal = host_out[0];
        stats->dual = host_out[1];
    }
    if (std::isnan(gap))
        return fail(WOTK_ENAN, "Overflow encountered in duality gap computation, please report this incident");
    rc = be.finish(eps_i, (double)pb.n_j, s);
    if (rc) return rc;
    WOTK_HIP(hipStreamSynchronize(s));
    return WOTK_OK;
}

// the stored FP32 cost (C, C^T) as the schedule's backend
struct StoredBackend {
    const Problem &pb;
    Work &w;
    double *plan;
    int64_t ldp;
    double *row_sums, *f_out, *g_out;
    int iterate(double l1, double l2, double eps, hipStream_t s) { return wotk::iterate(pb, w, l1, l2, eps, s); }
    int gap_pass(const double *f, const double *g, double l1, double l2, double eps, const double *skt,
                 hipStream_t s) {
        return wotk::gap_pass(pb, w, f, g, l1, l2, eps, skt, s);
    }
    int ktilde_sum(double eps, double *out, hipStream_t s) { return wotk::ktilde_sum(pb, w, eps, out, s); }
    int finish(double eps, double div, hipStream_t s) {
        return finish_outputs(pb, w, eps, div, plan, ldp, row_sums, f_out, g_out, s);
    }
};

extern "C" int wotk_transport_duality_gap(const float *C, int64_t ldc, const float *CT,
                                          int64_t ldct, int64_t I, int64_t J, const double *G,
                                          double lambda1, double lambda2, double epsilon,
                                          int64_t batch_size, double tolerance, double tau,
                                          double epsilon0, int64_t max_iter, double *plan,
                                          int64_t ldp, double *row_sums, double *f_out,
                                          double *g_out, void *workspace, size_t workspace_bytes,
                                          wotk_stats *stats, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    Problem pb;
    Work w;
    int rc = setup(C, ldc, CT, ldct, I, J, G, workspace, workspace_bytes, &pb, &w, s);
    if (rc) return rc;
    WOTK_REQUIRE(!plan || ldp >= J, "ldp < J");
    StoredBackend be{pb, w, plan, ldp, row_sums, f_out, g_out};
    return gap_schedule(be, pb, w, lambda1, lambda2, epsilon, batch_size, tolerance, tau, epsilon0, max_iter, stats,
                        s);
}

// ---------------------------------------------------------------------------
// The duality-gap schedule on the never-stored cost: every sweep regenerates
// C from the coordinates (fused.hip), the gap pass forms the row and column
// sums of R = exp((f_i + g_j - C_ij)/eps) with FP64 exponentials over the
// same regenerated C, and the primal's entropy and transport terms come from
// them: eps sum R log R + sum R C = sum_i f_i r_i + sum_j g_j c_j, since
// log R_ij = (f_i + g_j - C_ij) / eps (an entry whose exp underflows adds 0
// to both sides, as nan_to_num(log 0) * 0 does in the reference).

// partial slots the gap finish reads: [0] = sum R log R, [1] = sum R C
__global__ __launch_bounds__(1024) void gap_identity_kernel(const double *__restrict__ f, const double *__restrict__ r,
                                                            int n_i, const double *__restrict__ g,
                                                            const double *__restrict__ c, int n_j, double inv_eps,
                                                            double *__restrict__ partials) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n_i; i += blockDim.x) acc += f[i] * r[i];
    for (int j = threadIdx.x; j < n_j; j += blockDim.x) acc += g[j] * c[j];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        partials[0] = t * inv_eps;
        partials[1] = 0.0;
        partials[2] = 0.0;
    }
}

// sum of n doubles in a fixed order -> out[0]
__global__ __launch_bounds__(1024) void sum_vec_kernel(const double *__restrict__ x, int n, double *__restrict__ out) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += x[i];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        out[0] = t;
    }
}

struct FusedBackend {
    const Problem &pb;
    Work &w;
    const FusedProblem &fp;
    double *row_sums, *f_out, *g_out;
    int iterate(double l1, double l2, double eps, hipStream_t s) {
        return fused_problem_iterate(fp, w.f, w.g, w.logp, w.misc, l1, l2, eps, s);
    }
    int gap_pass(const double *f, const double *g, double l1, double l2, double eps, const double *skt,
                 hipStream_t s) {
        int rc = fused_problem_sums(fp, false, f, g, eps, 1.0, true, w.r, s);
        if (rc) return rc;
        rc = fused_problem_sums(fp, true, f, g, eps, 1.0, true, w.c, s);
        if (rc) return rc;
        gap_identity_kernel<<<1, 1024, 0, s>>>(f, w.r, pb.n_i, g, w.c, pb.n_j, 1.0 / eps, w.partials);
        GapArgs ga{w.r, w.c, f, g, pb.G, w.misc, w.partials, 1, pb.n_i, pb.n_j, l1, l2, eps, 0.0, skt, w.gapout};
        gap_finish_kernel<<<1, 1024, 0, s>>>(ga);
        WOTK_LAUNCH_CHECK();
        return WOTK_OK;
    }
    int ktilde_sum(double eps, double *out, hipStream_t s) {
        // row sums of exp(-C/eps): potentials zero, FP64 exponentials
        int rc = fused_problem_sums(fp, false, fp.zeros, fp.zeros, eps, 1.0, true, w.r, s);
        if (rc) return rc;
        sum_vec_kernel<<<1, 1024, 0, s>>>(w.r, pb.n_i, out);
        WOTK_LAUNCH_CHECK();
        return WOTK_OK;
    }
    int finish(double eps, double div, hipStream_t s) {
        if (row_sums) {
            int rc = fused_problem_sums(fp, false, w.f, w.g, eps, div, false, row_sums, s);
            if (rc) return rc;
        }
        if (f_out) WOTK_HIP(hipMemcpyAsync(f_out, w.f, sizeof(double) * pb.n_i, hipMemcpyDeviceToDevice, s));
        if (g_out) WOTK_HIP(hipMemcpyAsync(g_out, w.g, sizeof(double) * pb.n_j, hipMemcpyDeviceToDevice, s));
        return WOTK_OK;
    }
};

extern "C" size_t wotk_duality_gap_coords_workspace_size(int64_t I, int64_t J, int64_t d) {
    Arena ar(nullptr, 0);
    Work w;
    return fused_problem_bytes(I, J, d) + 256 + carve(ar, I, J, &w);
}

extern "C" int wotk_transport_duality_gap_coords(const double *X, const double *Y, int64_t I, int64_t J, int64_t d,
                                                 const double *G, double lambda1, double lambda2, double epsilon,
                                                 int64_t batch_size, double tolerance, double tau, double epsilon0,
                                                 int64_t max_iter, int32_t terms, double *f_out, double *g_out,
                                                 double *row_sums, double *median_out, void *workspace,
                                                 size_t workspace_bytes, wotk_stats *stats, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    WOTK_REQUIRE(G != nullptr, "G is required");
    FusedProblem fp;
    size_t used = 0;
    int rc = fused_problem_setup(X, Y, I, J, d, terms, workspace, workspace_bytes, &fp, &used, s);
    if (rc) return rc;
    used = (used + 255) & ~size_t(255);
    Arena ar(static_cast<char *>(workspace) + used, workspace_bytes > used ? workspace_bytes - used : 0);
    Work w;
    carve(ar, I, J, &w);
    if (!ar.ok) return fail(WOTK_EWORKSPACE, "workspace too small");
    const Problem pb{nullptr, 0, nullptr, 0, (int)I, (int)J, G};
    rc = init_state(pb, w, s);
    if (rc) return rc;
    if (median_out) WOTK_HIP(hipMemcpyAsync(median_out, fp.median, sizeof(double), hipMemcpyDeviceToDevice, s));
    FusedBackend be{pb, w, fp, row_sums, f_out, g_out};
    return gap_schedule(be, pb, w, lambda1, lambda2, epsilon, batch_size, tolerance, tau, epsilon0, max_iter, stats,
                        s);
}

extern "C" int wotk_plan(const float *C, int64_t ldc, int64_t I, int64_t J, const double *f, const double *g,
                         double epsilon, double divisor, double *plan, int64_t ldp, double *row_sums,
                         void *stream) {
    WOTK_REQUIRE(I > 0 && J > 0 && I < (1ll << 31) && J < (1ll << 31), "bad shape");
    WOTK_REQUIRE(ldc % 4 == 0 && ldc >= round_up(J, 4), "ldc must be a multiple of 4 >= J");
    WOTK_REQUIRE(!plan || ldp >= J, "ldp < J");
    return launch_plan(C, ldc, (int)I, (int)J, f, g, epsilon, divisor, plan, ldp, row_sums, nullptr, 0,
                       (hipStream_t)stream);
}

extern "C" int wotk_primal_dual(const float *C, int64_t ldc, const float *CT, int64_t ldct,
                                int64_t I, int64_t J, const double *G, const double *f,
                                const double *g, double lambda1, double lambda2, double epsilon,
                                double *out_host, void *workspace, size_t workspace_bytes,
                                void *stream) {
    hipStream_t s = (hipStream_t)stream;
    Problem pb;
    Work w;
    int rc = setup(C, ldc, CT, ldct, I, J, G, workspace, workspace_bytes, &pb, &w, s);
    if (rc) return rc;
    WOTK_REQUIRE(out_host != nullptr, "out_host is required");
    copy2_kernel<<<64, 256, 0, s>>>(w.f, f, pb.n_i, w.g, g, pb.n_j);
    WOTK_LAUNCH_CHECK();
    double *sum_kt = w.misc + 4;
    rc = ktilde_sum(pb, w, epsilon, sum_kt, s);
    if (rc) return rc;
    rc = gap_pass(pb, w, w.f, w.g, lambda1, lambda2, epsilon, sum_kt, s);
    if (rc) return rc;
    double host_out[3];
    WOTK_HIP(hipMemcpyAsync(host_out, w.gapout, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    WOTK_HIP(hipStreamSynchronize(s));
    out_host[0] = host_out[0];
    out_host[1] = host_out[1];
    return WOTK_OK;
}
