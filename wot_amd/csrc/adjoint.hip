// Gradient of the converged regularised unbalanced OT objective
// (the primal of wot/ot/optimal_transport.py:45-53 at the optimal plan) with
// respect to the source embeddings and the log growth rates.
//
// Envelope theorem: at the optimum R* of P(R; C, p, q), dOT/dtheta equals the
// partial derivative of P at fixed R*.  With K~ = exp(-C/eps):
//   dP/dC_ij = (R_ij - K~_ij) / (I J)                         =: W_ij
//   dP/dp_i  = lambda1/I (1 - x_i/p_i) + (1/I) sum_j lambda2/J (1 - y_j/q)
// where x = R 1/J, y = R^T 1/I and q = mean(p) (optimal_transport.py:108,192).
// The cost is C = D / m with D_ij = |x_i - y_j|^2 and m = median(D):
//   dOT/dx_i = (2/m) sum_j W_ij (x_i - y_j)  -  (S/m) sum_{e in mid} w_e 2 (x_i - y_{l_e}) [k_e = i]
// with S = sum W_ij C_ij and mid the middle element(s) of D (weight 1/2 each
// for an even count, 1 for an odd count: the median's subgradient).
// p = G = g^dt (ot_model.py:315), so dOT/dlog g_i = dt p_i dP/dp_i.
#include "common.h"

namespace wotk {

int launch_plan_sums(const float *A, int64_t lda, int nrows, int ncols, const double *f, const double *g,
                     double eps, double *row_sums, hipStream_t s);

constexpr int kAdjThreads = 256;
constexpr int kAdjRows = 4;

// W = (R - K~)/(IJ) written in FP64; per row: sum R, sum W; per block: sum W*C
__global__ __launch_bounds__(kAdjThreads) void adjoint_rows_kernel(
    const float *__restrict__ A, int64_t lda, int n_i, int n_j, const double *__restrict__ f,
    const double *__restrict__ g, double inv_eps, double inv_ij, double *__restrict__ W,
    double *__restrict__ r_out, double *__restrict__ sw_out, double *__restrict__ partials) {
    const int r0 = blockIdx.x * kAdjRows;
    const int tid = threadIdx.x;
    double acc_r[kAdjRows], acc_w[kAdjRows], fr[kAdjRows];
    const float *rowp[kAdjRows];
    double s_wc = 0.0;
#pragma unroll
    for (int r = 0; r < kAdjRows; ++r) {
        const int row = min(r0 + r, n_i - 1);
        rowp[r] = A + (int64_t)row * lda;
        fr[r] = f[row];
        acc_r[r] = 0.0;
        acc_w[r] = 0.0;
    }
    for (int j = tid; j < n_j; j += kAdjThreads) {
        const double gj = g[j];
#pragma unroll
        for (int r = 0; r < kAdjRows; ++r) {
            if (r0 + r >= n_i) continue;
            const double c = (double)rowp[r][j];
            const double R = exp((fr[r] + gj - c) * inv_eps);
            const double w = (R - exp(-c * inv_eps)) * inv_ij;
            W[(int64_t)(r0 + r) * n_j + j] = w;
            acc_r[r] += R;
            acc_w[r] += w;
            s_wc = fma(w, c, s_wc);
        }
    }
    __shared__ double red[2 * kAdjRows + 1][kAdjThreads / 64];
    const int lane = tid & 63, wave = tid >> 6;
    auto wsum = [](double v) {
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        return v;
    };
#pragma unroll
    for (int r = 0; r < kAdjRows; ++r) {
        const double a = wsum(acc_r[r]), b = wsum(acc_w[r]);
        if (lane == 0) {
            red[r][wave] = a;
            red[kAdjRows + r][wave] = b;
        }
    }
    const double t = wsum(s_wc);
    if (lane == 0) red[2 * kAdjRows][wave] = t;
    __syncthreads();
    if (tid < 2 * kAdjRows + 1) {
        double v = 0.0;
        for (int w = 0; w < kAdjThreads / 64; ++w) v += red[tid][w];
        if (tid < kAdjRows) {
            if (r0 + tid < n_i) r_out[r0 + tid] = v;
        } else if (tid < 2 * kAdjRows) {
            if (r0 + tid - kAdjRows < n_i) sw_out[r0 + tid - kAdjRows] = v;
        } else {
            partials[blockIdx.x] = v;
        }
    }
}

// V = W Y : (I x J) (J x d) -> (I x d), FP64, 64 x 64 output tile per block,
// J streamed through LDS in chunks of 32.
__global__ __launch_bounds__(256) void gemm_wy_kernel(const double *__restrict__ W, int n_i, int n_j,
                                                      const double *__restrict__ Y, int d,
                                                      double *__restrict__ V) {
    __shared__ double ws[32][64];
    __shared__ double ys[32][64];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int i0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    double acc[4][4] = {};
    for (int j0 = 0; j0 < n_j; j0 += 32) {
        for (int e = tid; e < 64 * 32; e += 256) {
            const int r = e / 32, jj = e % 32;  // W: row r, column j0+jj (coalesced along jj)
            ws[jj][r] = (i0 + r < n_i && j0 + jj < n_j) ? W[(int64_t)(i0 + r) * n_j + j0 + jj] : 0.0;
            const int jy = e / 64, kk = e % 64;  // Y: row j0+jy, column k0+kk
            ys[jy][kk] = (j0 + jy < n_j && k0 + kk < d) ? Y[(int64_t)(j0 + jy) * d + k0 + kk] : 0.0;
        }
        __syncthreads();
        for (int jj = 0; jj < 32; ++jj) {
            double wv[4], yv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                wv[q] = ws[jj][ty * 4 + q];
                yv[q] = ys[jj][tx * 4 + q];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fma(wv[a], yv[b], acc[a][b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int i = i0 + ty * 4 + a;
        if (i >= n_i) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int k = k0 + tx * 4 + b;
            if (k < d) V[(int64_t)i * d + k] = acc[a][b];
        }
    }
}

// scalars: [0] = S = sum W C, [1] = Q = sum_j lambda2/J (1 - y_j/q), [2] = q
__global__ __launch_bounds__(1024) void adjoint_scalars_kernel(const double *__restrict__ partials, int nparts,
                                                               const double *__restrict__ c, int n_i, int n_j,
                                                               const double *__restrict__ G, double lambda2,
                                                               double *__restrict__ out) {
    __shared__ double red[2][16];
    double sp = 0.0, sg = 0.0;
    for (int i = threadIdx.x; i < n_i; i += blockDim.x) sg += G[i];
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) sp += partials[b];
    for (int off = 32; off > 0; off >>= 1) {
        sg += __shfl_xor(sg, off);
        sp += __shfl_xor(sp, off);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = sg;
        red[1][w] = sp;
    }
    __syncthreads();
    __shared__ double q_sh;
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            a += red[0][k];
            b += red[1][k];
        }
        q_sh = a / n_i;
        out[0] = b;
        out[2] = q_sh;
    }
    __syncthreads();
    const double q = q_sh;
    double sq = 0.0;
    for (int j = threadIdx.x; j < n_j; j += blockDim.x) sq += (1.0 - (c[j] / n_i) / q);
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[0][w] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) a += red[0][k];
        out[1] = lambda2 * a / n_j;
    }
}

__global__ void adjoint_finish_kernel(const double *__restrict__ X, int n_i, int d, const double *__restrict__ V,
                                      const double *__restrict__ sw, const double *__restrict__ r,
                                      const double *__restrict__ G, double dt, double lambda1, int n_j,
                                      const double *__restrict__ median, const double *__restrict__ scal,
                                      double *__restrict__ grad_x, double *__restrict__ grad_lg) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double two_m = 2.0 / median[0];
    if (grad_x && e < (int64_t)n_i * d) {
        const int i = (int)(e / d);
        grad_x[e] = two_m * (X[e] * sw[i] - V[e]);
    }
    if (grad_lg && e < n_i) {
        const double p = G[e];
        const double x = r[e] / n_j;
        const double dp = lambda1 / n_i * (1.0 - x / p) + scal[1] / n_i;
        grad_lg[e] = dt * p * dp;
    }
}

// the median's subgradient: one thread, after adjoint_finish_kernel
__global__ void adjoint_median_kernel(const double *__restrict__ X, const double *__restrict__ Y, int n_j, int d,
                                      const double *__restrict__ median, const long long *__restrict__ idx,
                                      int odd, const double *__restrict__ scal, double *__restrict__ grad_x) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double coef = scal[0] / median[0];
    const int ne = odd ? 1 : 2;
    const double we = odd ? 1.0 : 0.5;
    for (int e = 0; e < ne; ++e) {
        const long long flat = idx[e];
        const long long k = flat / n_j, l = flat % n_j;
        for (int t = 0; t < d; ++t)
            grad_x[k * d + t] -= coef * we * 2.0 * (X[k * d + t] - Y[l * d + t]);
    }
}

struct AdjWork {
    double *W, *r, *sw, *c, *V, *partials, *scal;
    int nparts;
};

static size_t adj_carve(Arena &ar, int64_t I, int64_t J, int64_t d, AdjWork *w) {
    w->W = ar.take<double>((size_t)I * J);
    w->r = ar.take<double>(I);
    w->sw = ar.take<double>(I);
    w->c = ar.take<double>(J);
    w->V = ar.take<double>((size_t)I * d);
    w->nparts = (int)((I + kAdjRows - 1) / kAdjRows);
    w->partials = ar.take<double>(w->nparts);
    w->scal = ar.take<double>(8);
    return ar.off;
}

}  // namespace wotk

using namespace wotk;

extern "C" size_t wotk_adjoint_workspace_size(int64_t I, int64_t J, int64_t d) {
    Arena ar(nullptr, 0);
    AdjWork w;
    return adj_carve(ar, I, J, d, &w);
}

extern "C" int wotk_ot_adjoint(const float *C, int64_t ldc, const float *CT, int64_t ldct, int64_t I,
                               int64_t J, const double *X, const double *Y, int64_t d, const double *G,
                               double dt, const double *f, const double *g, double lambda1, double lambda2,
                               double epsilon, const double *median, const int64_t *median_index,
                               double *grad_x, double *grad_log_g, void *workspace, size_t workspace_bytes,
                               void *stream) {
    hipStream_t s = (hipStream_t)stream;
    WOTK_REQUIRE(I > 0 && J > 0 && d > 0 && I < (1ll << 31) && J < (1ll << 31), "bad shape");
    WOTK_REQUIRE(ldc >= J && ldct >= I, "leading dimension too small");
    WOTK_REQUIRE(X && Y && G && f && g && median, "missing input");
    Arena ar(workspace, workspace_bytes);
    AdjWork w;
    adj_carve(ar, I, J, d, &w);
    if (!workspace || !ar.ok) return fail(WOTK_EWORKSPACE, "adjoint workspace too small");
    const double inv_eps = 1.0 / epsilon;
    const double inv_ij = 1.0 / ((double)I * (double)J);
    adjoint_rows_kernel<<<w.nparts, kAdjThreads, 0, s>>>(C, ldc, (int)I, (int)J, f, g, inv_eps, inv_ij, w.W, w.r,
                                                        w.sw, w.partials);
    WOTK_LAUNCH_CHECK();
    int rc = launch_plan_sums(CT, ldct, (int)J, (int)I, g, f, epsilon, w.c, s);
    if (rc) return rc;
    adjoint_scalars_kernel<<<1, 1024, 0, s>>>(w.partials, w.nparts, w.c, (int)I, (int)J, G, lambda2, w.scal);
    WOTK_LAUNCH_CHECK();
    if (grad_x) {
        dim3 grid((unsigned)((I + 63) / 64), (unsigned)((d + 63) / 64));
        gemm_wy_kernel<<<grid, 256, 0, s>>>(w.W, (int)I, (int)J, Y, (int)d, w.V);
        WOTK_LAUNCH_CHECK();
    }
    const int64_t n = std::max<int64_t>(I * d, I);
    adjoint_finish_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(X, (int)I, (int)d, w.V, w.sw, w.r, G, dt,
                                                                      lambda1, (int)J, median, w.scal, grad_x,
                                                                      grad_log_g);
    WOTK_LAUNCH_CHECK();
    if (grad_x && median_index) {
        const int odd = (int)(((I * J) % 2) == 1);
        adjoint_median_kernel<<<1, 1, 0, s>>>(X, Y, (int)J, (int)d, median,
                                              reinterpret_cast<const long long *>(median_index), odd, w.scal,
                                              grad_x);
        WOTK_LAUNCH_CHECK();
    }
    return WOTK_OK;
}
