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
constexpr int kAdjRows = 16;  // rows per block
constexpr int kAdjCols = 32;  // columns per chunk

// One pass over the cost rows: W_ij = (R_ij - K~_ij)/(IJ) for a block of 16
// rows and a chunk of 32 columns is formed in FP64 (two exps per entry) into
// LDS next to the chunk's 32 target coordinates, and the block accumulates
// V = W Y (rows x d, FP64, registers) from LDS, so W never reaches HBM.  Per
// row it also sums R and W, and per block W * C.  Thread t forms entries
// (row t / 32 + 8 m, column t % 32) of a chunk and owns V[2 (t / 32) + a][t % 32
// + 32 q] (a < 2, q < KPT, d <= 32 KPT).
template <int KPT>
__global__ __launch_bounds__(kAdjThreads) void adjoint_rows_kernel(
    const float *__restrict__ A, int64_t lda, int n_i, int n_j, const double *__restrict__ f,
    const double *__restrict__ g, double inv_eps, double inv_ij, const double *__restrict__ Y, int d,
    double *__restrict__ V, double *__restrict__ r_out, double *__restrict__ sw_out,
    double *__restrict__ partials) {
    constexpr int DP = 32 * KPT;
    __shared__ double w_lds[kAdjRows][kAdjCols + 1];  // +1: rows of the FMA phase fall on distinct banks
    __shared__ double y_lds[kAdjCols][DP];
    __shared__ double red[2 * kAdjRows + 1][kAdjThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blockIdx.x * kAdjRows;
    const int jc = tid % kAdjCols, rq = tid / kAdjCols;  // chunk column; first row of this thread (step 8)
    constexpr int RPT = kAdjRows / (kAdjThreads / kAdjCols);  // rows formed per thread: 2
    double fr[RPT], acc_r[RPT], acc_w[RPT];
    const float *rowp[RPT];
#pragma unroll
    for (int m = 0; m < RPT; ++m) {
        const int row = min(r0 + rq + 8 * m, n_i - 1);
        rowp[m] = A + (int64_t)row * lda;
        fr[m] = f[row];
        acc_r[m] = 0.0;
        acc_w[m] = 0.0;
    }
    double s_wc = 0.0;
    double v[2][KPT];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int q = 0; q < KPT; ++q) v[a][q] = 0.0;
    const int vr = 2 * rq;  // this thread's two V rows (block-local)
    for (int j0 = 0; j0 < n_j; j0 += kAdjCols) {
        const int j = j0 + jc;
        const bool jin = j < n_j;
        const double gj = jin ? g[j] : 0.0;
#pragma unroll
        for (int m = 0; m < RPT; ++m) {
            double w = 0.0;
            if (jin && r0 + rq + 8 * m < n_i) {
                const double c = (double)rowp[m][j];
                const double R = exp((fr[m] + gj - c) * inv_eps);
                w = (R - exp(-c * inv_eps)) * inv_ij;
                acc_r[m] += R;
                acc_w[m] += w;
                s_wc = fma(w, c, s_wc);
            }
            w_lds[rq + 8 * m][jc] = w;
        }
        for (int e = tid; e < kAdjCols * DP; e += kAdjThreads) {
            const int jj = e / DP, k = e % DP;
            y_lds[jj][k] = (j0 + jj < n_j && k < d) ? Y[(int64_t)(j0 + jj) * d + k] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int jj = 0; jj < kAdjCols; ++jj) {
            const double w0 = w_lds[vr][jj], w1 = w_lds[vr + 1][jj];
#pragma unroll
            for (int q = 0; q < KPT; ++q) {
                const double y = y_lds[jj][jc + 32 * q];
                v[0][q] = fma(w0, y, v[0][q]);
                v[1][q] = fma(w1, y, v[1][q]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int i = r0 + vr + a;
        if (i >= n_i) continue;
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            const int k = jc + 32 * q;
            if (k < d) V[(int64_t)i * d + k] = v[a][q];
        }
    }
    auto wsum = [](double x) {
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        return x;
    };
    // a wave holds rows rq in {2 wave, 2 wave + 1} (+ 8 m): lanes 0-31 and 32-63
    // form the two rows' columns; reduce each half separately
    auto halfsum = [](double x) {
        for (int off = 16; off > 0; off >>= 1) x += __shfl_xor(x, off);
        return x;
    };
#pragma unroll
    for (int m = 0; m < RPT; ++m) {
        const double a = halfsum(acc_r[m]), b = halfsum(acc_w[m]);
        if ((lane & 31) == 0) {
            red[rq + 8 * m][0] = a;
            red[kAdjRows + rq + 8 * m][0] = b;
        }
    }
    const double t = wsum(s_wc);
    if (lane == 0) red[2 * kAdjRows][wave] = t;
    __syncthreads();
    if (tid < kAdjRows) {
        if (r0 + tid < n_i) r_out[r0 + tid] = red[tid][0];
    } else if (tid < 2 * kAdjRows) {
        if (r0 + tid - kAdjRows < n_i) sw_out[r0 + tid - kAdjRows] = red[tid][0];
    } else if (tid == 2 * kAdjRows) {
        double x = 0.0;
        for (int w = 0; w < kAdjThreads / 64; ++w) x += red[2 * kAdjRows][w];
        partials[blockIdx.x] = x;
    }
}

// ---------------------------------------------------------------------------
// The same rows pass on the matrix cores (d <= 64), split over columns.
//
// A wave owns RB blocks of 16 rows and the columns of one split; lane
// (m, q) = (lane % 16, lane / 16) forms W for row 16 rb + m and four columns
// chosen by q in each 16-column step: W is the A operand (16 x 4, k = q) of
// four k-steps of v_mfma_f64_16x16x4f64 and Y of the same columns (Y
// zero-padded, coordinates laid on the tiles by adj_coord) the B operand, so
// V = W Y runs on the matrix cores beside the exponentials.  Any assignment
// of columns to (step, q) works as long as W and Y follow the same one: in
// super-steps of 2 steps lane q takes the 8 contiguous columns J0 + 8 q ...
// (32 bytes of each of its rows, the wave a full line per row at once); the
// rest of a split goes in single steps (columns j + 4 q + s).  The next
// super-step's cost and the next step's column factors and Y are loaded while
// the current one computes.  No LDS and no barriers.
//
// One exponential per entry: R = exp(-C/eps) exp(f_i/eps) exp(g_j/eps), the
// row's factor formed once per lane and the column's by
// adjoint_colexp_kernel, when both factors lie in [2^-500, 2^500] and
// |C| <= 690 eps; otherwise (an underflowing K~, an overflowing factor, NaN)
// R = exp((f_i + g_j - C)/eps) directly.
//
// Outputs per split (summed over splits in a fixed order afterwards):
// V (n_i x d), the row sums of R and W, and per wave sum W C.
constexpr int kAdjMfmaE = 4;                 // columns per lane and step
constexpr int kAdjMfmaStep = 4 * kAdjMfmaE;  // columns per step
constexpr int kAdjMfmaWaves = 4;             // waves per workgroup
constexpr int kAdjMfmaTargetWaves = 4096;    // splits are chosen to reach ~4 waves per SIMD
constexpr int kAdjSuperT = 2;                // steps per super-step (each lane reads 16 T contiguous bytes of a row)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct AdjMfmaArgs {
    const float *A;
    int64_t lda;
    int n_i, n_j, d;
    const double *f, *g, *bcol;
    const double *Yp;  // Y zero-padded to 16 ND coordinates per column
    double inv_eps, inv_ij;
    float c_lim;  // |C| <= c_lim: exp(-C/eps) within [2^-1000, 2^1000] (690 eps, rounded down)
    int split_cols, n_tiles;
    bool vec;  // 16-byte cost loads (lda and the base 16-byte aligned)
    double *Vp, *rp, *swp, *wcp;  // [split][n_i][d], [split][n_i], [split][n_i], [split][n_tiles]
    double *cpart;                // [n_tiles][n_j]: column sums of R over each wave's rows
};

typedef double d2 __attribute__((ext_vector_type(2)));

// Coordinates on the MFMA tiles: tiles 2p and 2p + 1 hold coordinates 32 p +
// 2 m and 32 p + 2 m + 1 in column m (one 16-byte load per lane fetches both);
// with ND odd the last tile holds 16 (ND - 1) + m.
template <int ND>
__device__ __forceinline__ int adj_coord(int nd, int m) {
    return nd < (ND & ~1) ? 32 * (nd >> 1) + 2 * m + (nd & 1) : 16 * nd + m;
}

// column factors and Y operands of one step: columns col0 + s (s < 4) of this lane
template <int ND>
struct AdjYStep {
    double b[kAdjMfmaE];
    double y[kAdjMfmaE][ND];
};

// FULL: all four columns lie before j_end (then col0 is a multiple of 4: the
// factor loads are 16-byte aligned).  Otherwise indices are clamped into range
// and the values selected to 0, so every load is in bounds and none sits under
// a branch.
template <int ND, bool FULL>
__device__ __forceinline__ void adj_fetch_y(AdjYStep<ND> &st, const AdjMfmaArgs &a, int col0, int j_end, int m) {
    constexpr int E = kAdjMfmaE, DP = 16 * ND;
    if (FULL) {
        const d2 b01 = *reinterpret_cast<const d2 *>(a.bcol + col0);
        const d2 b23 = *reinterpret_cast<const d2 *>(a.bcol + col0 + 2);
        st.b[0] = b01.x;
        st.b[1] = b01.y;
        st.b[2] = b23.x;
        st.b[3] = b23.y;
    }
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const bool ok = FULL || col0 + s < j_end;
        const int col = FULL ? col0 + s : min(col0 + s, j_end - 1);
        if (!FULL) {
            const double b = a.bcol[col];
            st.b[s] = ok ? b : 0.0;
        }
        const double *yr = a.Yp + (int64_t)col * DP;
#pragma unroll
        for (int p = 0; p < ND / 2; ++p) {
            const d2 v = *reinterpret_cast<const d2 *>(yr + 32 * p + 2 * m);
            st.y[s][2 * p] = ok ? v.x : 0.0;
            st.y[s][2 * p + 1] = ok ? v.y : 0.0;
        }
        if (ND & 1) {
            const double v = yr[16 * (ND - 1) + m];
            st.y[s][ND - 1] = ok ? v : 0.0;
        }
    }
}

// cost entries of one step: columns col0 + s (s < 4) of this lane's rows
template <int RB, bool FULL>
__device__ __forceinline__ void adj_fetch_cost(float c[RB][kAdjMfmaE], const AdjMfmaArgs &a,
                                               const float *const rowp[RB], int col0, int j_end) {
    constexpr int E = kAdjMfmaE;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        if (FULL && a.vec) {
            const f4 v = *reinterpret_cast<const f4 *>(rowp[rb] + col0);
            c[rb][0] = v.x;
            c[rb][1] = v.y;
            c[rb][2] = v.z;
            c[rb][3] = v.w;
        } else {
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const int col = FULL ? col0 + s : min(col0 + s, j_end - 1);
                const float x = rowp[rb][col];
                c[rb][s] = FULL || col0 + s < j_end ? x : 0.f;
            }
        }
    }
}

// Entries whose row and column factors lie in [2^-500, 2^500] and whose cost
// has |C| <= 690 eps take R = exp(-C/eps) (a f) (b g); their exponential, W
// and MFMAs form one basic block per step, so the matrix-core products of one
// entry overlap the exponentials of the next.  The others (none in most
// sweeps) enter that pass with weight 0 and are added afterwards with R formed
// directly, under one branch per step.  Columns col0 + s, s < 4.
template <int RB, int ND>
__device__ __forceinline__ void adj_compute(const float cv[RB][kAdjMfmaE], const AdjYStep<ND> &st,
                                            const AdjMfmaArgs &a, const bool row_ok[RB], const bool row_in[RB],
                                            const double fr[RB], const double ar[RB], int col0, int j_end,
                                            d4 acc[RB][ND], double acc_r[RB], double acc_w[RB], double &swc,
                                            double *cprow, int m) {
    constexpr int E = kAdjMfmaE;
    const float c_lim = a.c_lim;
    bool any_slow = false;
    double csum[E];  // R summed over this lane's rows, per column
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const bool colok = col0 + s < j_end;
        const bool col_in = st.b[s] >= 0x1p-500 && st.b[s] <= 0x1p500;
        csum[s] = 0.0;
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            const float cf = cv[rb][s];
            const bool fast = row_ok[rb] && colok && row_in[rb] && col_in && fabsf(cf) <= c_lim;
            any_slow |= row_ok[rb] && colok && !fast;
            const double c = (double)cf;
            const double kt = exp(-c * a.inv_eps);
            const double R = kt * (ar[rb] * st.b[s]);
            const double w = fast ? (R - kt) * a.inv_ij : 0.0;
            acc_r[rb] += fast ? R : 0.0;
            csum[s] += fast ? R : 0.0;
            acc_w[rb] += w;
            swc = fma(w, c, swc);
#pragma unroll
            for (int nd = 0; nd < ND; ++nd)
                acc[rb][nd] = __builtin_amdgcn_mfma_f64_16x16x4f64(w, st.y[s][nd], acc[rb][nd], 0, 0, 0);
        }
    }
    if (__builtin_amdgcn_read_exec() & __ballot(any_slow)) {
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const bool colok = col0 + s < j_end;
            const bool col_in = st.b[s] >= 0x1p-500 && st.b[s] <= 0x1p500;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                const float cf = cv[rb][s];
                const bool slow = row_ok[rb] && colok && !(row_in[rb] && col_in && fabsf(cf) <= c_lim);
                double R = 0.0, w = 0.0;
                const double c = (double)cf;
                if (slow) {
                    R = exp((fr[rb] + a.g[col0 + s] - c) * a.inv_eps);
                    w = (R - exp(-c * a.inv_eps)) * a.inv_ij;
                }
                acc_r[rb] += R;
                csum[s] += R;
                acc_w[rb] += w;
                swc = fma(w, c, swc);
#pragma unroll
                for (int nd = 0; nd < ND; ++nd)
                    acc[rb][nd] = __builtin_amdgcn_mfma_f64_16x16x4f64(w, st.y[s][nd], acc[rb][nd], 0, 0, 0);
            }
        }
    }
    // the column's sum over the wave's rows: the 16 lanes of a q-group share
    // the columns (fixed xor order); lane m = 0 writes it
#pragma unroll
    for (int s = 0; s < E; ++s) {
        double v = csum[s];
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off);
        if (m == 0 && col0 + s < j_end) cprow[col0 + s] = v;
    }
}

// The wave's state over one split
template <int RB, int ND>
struct AdjWave {
    const float *rowp[RB];
    bool row_ok[RB], row_in[RB];
    double fr[RB], ar[RB], acc_r[RB], acc_w[RB];
    d4 acc[RB][ND];
    double swc;
    double *cprow;  // this row tile's column partials
};

// One super-step of 16 T columns from J0: lane q takes columns J0 + 4 T q +
// 4 t + s in step t (s < 4), so each lane's 16 T bytes of a row, and the
// wave's 64 T bytes per row, are read by one set of loads (issued a super-step
// ahead into `nc`).  Y operands go one step ahead through ya / yb (T even):
// step 0's are in ya on entry, and the next super-step's step 0 is in ya on
// exit when `more`.
template <int RB, int ND>
__device__ __forceinline__ void adj_super_step(AdjWave<RB, ND> &wv, const AdjMfmaArgs &a,
                                               const f4 cc[RB][kAdjSuperT], f4 nc[RB][kAdjSuperT],
                                               AdjYStep<ND> &ya, AdjYStep<ND> &yb, int J0, bool more, int j_end,
                                               int m, int q) {
    constexpr int T = kAdjSuperT, W = 16 * kAdjSuperT;
    static_assert(T % 2 == 0, "Y buffers alternate per step");
    if (more) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int t = 0; t < T; ++t)
                nc[rb][t] = *reinterpret_cast<const f4 *>(wv.rowp[rb] + J0 + W + 4 * T * q + 4 * t);
    }
    const int c0 = J0 + 4 * T * q;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        AdjYStep<ND> &yc = (t & 1) ? yb : ya;
        AdjYStep<ND> &yn = (t & 1) ? ya : yb;
        if (t + 1 < T)
            adj_fetch_y<ND, true>(yn, a, c0 + 4 * (t + 1), j_end, m);
        else if (more)
            adj_fetch_y<ND, true>(yn, a, J0 + W + 4 * T * q, j_end, m);
        float cv[RB][kAdjMfmaE];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            cv[rb][0] = cc[rb][t].x;
            cv[rb][1] = cc[rb][t].y;
            cv[rb][2] = cc[rb][t].z;
            cv[rb][3] = cc[rb][t].w;
        }
        adj_compute<RB, ND>(cv, yc, a, wv.row_ok, wv.row_in, wv.fr, wv.ar, c0 + 4 * t, j_end, wv.acc, wv.acc_r,
                            wv.acc_w, wv.swc, wv.cprow, m);
    }
}

template <int RB, int ND>
__global__ __launch_bounds__(kAdjMfmaWaves * 64) void adjoint_mfma_kernel(AdjMfmaArgs a) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * kAdjMfmaWaves + (threadIdx.x >> 6);
    if (tile >= a.n_tiles) return;  // whole waves only; the kernel has no barriers
    const int split = blockIdx.y;
    const int m = lane & 15, q = lane >> 4;
    const int i0 = tile * 16 * RB;
    const int j_begin = split * a.split_cols;
    const int j_end = min(a.n_j, j_begin + a.split_cols);
    AdjWave<RB, ND> wv;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int r = i0 + 16 * rb + m;
        wv.row_ok[rb] = r < a.n_i;
        wv.rowp[rb] = a.A + (int64_t)(wv.row_ok[rb] ? r : 0) * a.lda;
        wv.fr[rb] = wv.row_ok[rb] ? a.f[r] : 0.0;
        wv.ar[rb] = exp(wv.fr[rb] * a.inv_eps);
        wv.row_in[rb] = wv.ar[rb] >= 0x1p-500 && wv.ar[rb] <= 0x1p500;
        wv.acc_r[rb] = 0.0;
        wv.acc_w[rb] = 0.0;
#pragma unroll
        for (int nd = 0; nd < ND; ++nd) wv.acc[rb][nd] = d4{0.0, 0.0, 0.0, 0.0};
    }
    wv.swc = 0.0;
    wv.cprow = a.cpart + (int64_t)tile * a.n_j;
    int j = j_begin;
    // full super-steps of 16 T columns (16-byte cost loads only); two per trip
    // so the cost buffers alternate without register copies
    constexpr int SW = 16 * kAdjSuperT;
    const int n_super = a.vec ? (j_end - j_begin) / SW : 0;
    if (n_super > 0) {
        f4 ca[RB][kAdjSuperT], cb[RB][kAdjSuperT];
        AdjYStep<ND> ya, yb;
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int t = 0; t < kAdjSuperT; ++t)
                ca[rb][t] = *reinterpret_cast<const f4 *>(wv.rowp[rb] + j + 4 * kAdjSuperT * q + 4 * t);
        adj_fetch_y<ND, true>(ya, a, j + 4 * kAdjSuperT * q, j_end, m);
        int S = 0;
        for (; S + 2 <= n_super; S += 2) {
            adj_super_step<RB, ND>(wv, a, ca, cb, ya, yb, j_begin + SW * S, true, j_end, m, q);
            adj_super_step<RB, ND>(wv, a, cb, ca, ya, yb, j_begin + SW * (S + 1), S + 2 < n_super, j_end, m, q);
        }
        if (S < n_super) adj_super_step<RB, ND>(wv, a, ca, cb, ya, yb, j_begin + SW * S, false, j_end, m, q);
        j = j_begin + SW * n_super;
    }
    // the rest (and everything without 16-byte cost loads): 16-column steps,
    // lane q taking columns j + 4 q + s, the next step's loads in flight
    if (j < j_end) {
        float ca[RB][kAdjMfmaE], cb[RB][kAdjMfmaE];
        AdjYStep<ND> ya, yb;
        auto fetch = [&](float (&c)[RB][kAdjMfmaE], AdjYStep<ND> &ys, int jj) {
            if (jj + kAdjMfmaStep <= j_end) {
                adj_fetch_cost<RB, true>(c, a, wv.rowp, jj + kAdjMfmaE * q, j_end);
                adj_fetch_y<ND, true>(ys, a, jj + kAdjMfmaE * q, j_end, m);
            } else {
                adj_fetch_cost<RB, false>(c, a, wv.rowp, jj + kAdjMfmaE * q, j_end);
                adj_fetch_y<ND, false>(ys, a, jj + kAdjMfmaE * q, j_end, m);
            }
        };
        fetch(ca, ya, j);
        while (j < j_end) {
            const bool has1 = j + kAdjMfmaStep < j_end;
            if (has1) fetch(cb, yb, j + kAdjMfmaStep);
            adj_compute<RB, ND>(ca, ya, a, wv.row_ok, wv.row_in, wv.fr, wv.ar, j + kAdjMfmaE * q, j_end, wv.acc,
                                wv.acc_r, wv.acc_w, wv.swc, wv.cprow, m);
            if (!has1) break;
            const bool has2 = j + 2 * kAdjMfmaStep < j_end;
            if (has2) fetch(ca, ya, j + 2 * kAdjMfmaStep);
            adj_compute<RB, ND>(cb, yb, a, wv.row_ok, wv.row_in, wv.fr, wv.ar, j + kAdjMfmaStep + kAdjMfmaE * q,
                                j_end, wv.acc, wv.acc_r, wv.acc_w, wv.swc, wv.cprow, m);
            j += 2 * kAdjMfmaStep;
        }
    }
    // D[i][n] of a 16 x 16 tile sits in lane n + 16 (i % 4) as acc[i / 4]: row
    // q + 4 r, tile column m (coordinate adj_coord(nd, m))
    double *Vp = a.Vp + (int64_t)split * a.n_i * a.d;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int nd = 0; nd < ND; ++nd)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = i0 + 16 * rb + q + 4 * r, dd = adj_coord<ND>(nd, m);
                if (row < a.n_i && dd < a.d) Vp[(int64_t)row * a.d + dd] = wv.acc[rb][nd][r];
            }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        double x = wv.acc_r[rb], y = wv.acc_w[rb];
        x += __shfl_xor(x, 16);
        y += __shfl_xor(y, 16);
        x += __shfl_xor(x, 32);
        y += __shfl_xor(y, 32);
        const int row = i0 + 16 * rb + m;
        if (q == 0 && row < a.n_i) {
            a.rp[(int64_t)split * a.n_i + row] = x;
            a.swp[(int64_t)split * a.n_i + row] = y;
        }
    }
    double swc = wv.swc;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) swc += __shfl_xor(swc, off);
    if (lane == 0) a.wcp[(int64_t)split * a.n_tiles + tile] = swc;
}

// Y (J x d) -> J x dp, zero-padded
__global__ __launch_bounds__(256) void adjoint_ypad_kernel(const double *__restrict__ Y, int n_j, int d, int dp,
                                                           double *__restrict__ Yp) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n_j * dp) return;
    const int64_t j = e / dp;
    const int k = (int)(e % dp);
    Yp[e] = k < d ? Y[j * d + k] : 0.0;
}

// exp(g_j / eps) per column
__global__ __launch_bounds__(256) void adjoint_colexp_kernel(const double *__restrict__ g, int n_j, double inv_eps,
                                                             double *__restrict__ bcol) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n_j) bcol[j] = exp(g[j] * inv_eps);
}

// V, row sums of R and W: the splits' partials summed in split order
__global__ __launch_bounds__(256) void adjoint_split_sum_kernel(const double *__restrict__ Vp,
                                                                const double *__restrict__ rp,
                                                                const double *__restrict__ swp, int n_split,
                                                                int n_i, int d, double *__restrict__ V,
                                                                double *__restrict__ r, double *__restrict__ sw) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nv = (int64_t)n_i * d;
    if (e < nv) {
        double x = 0.0;
        for (int s = 0; s < n_split; ++s) x += Vp[s * nv + e];
        V[e] = x;
    }
    if (e < n_i) {
        double x = 0.0, y = 0.0;
        for (int s = 0; s < n_split; ++s) {
            x += rp[(int64_t)s * n_i + e];
            y += swp[(int64_t)s * n_i + e];
        }
        r[e] = x;
        sw[e] = y;
    }
}

// Column sums of R (the plan's, y_j = c_j / I) from the row tiles' partials
// in two fixed-order levels: cseg[s][j] over the 64 row tiles of segment s
// (64 columns x 4 parts per workgroup), then c_j = sum_s cseg[s][j].
constexpr int kAdjColSeg = 64;

__global__ __launch_bounds__(256) void adjoint_colseg_kernel(const double *__restrict__ cpart, int n_tiles, int n_j,
                                                             double *__restrict__ cseg) {
    __shared__ double red[4][64];
    const int c = threadIdx.x & 63, part = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + c;
    const int t0 = blockIdx.y * kAdjColSeg, t1 = min(n_tiles, t0 + kAdjColSeg);
    double x = 0.0;
    if (j < n_j)
        for (int t = t0 + part; t < t1; t += 4) x += cpart[(int64_t)t * n_j + j];
    red[part][c] = x;
    __syncthreads();
    if (part == 0 && j < n_j) cseg[(int64_t)blockIdx.y * n_j + j] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
}

__global__ __launch_bounds__(256) void adjoint_colfinal_kernel(const double *__restrict__ cseg, int n_seg, int n_j,
                                                               double *__restrict__ c) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n_j) return;
    double x = 0.0;
    for (int k = 0; k < n_seg; ++k) x += cseg[(int64_t)k * n_j + j];
    c[j] = x;
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
    double *r, *sw, *c, *V, *partials, *scal;
    int nparts;
    // matrix-core path (d <= 64): per-split partials and the column factors
    bool mfma;
    int rb, nd, n_tiles, n_split, split_cols;
    double *Vp, *rp, *swp, *bcol, *Yp, *cpart, *cseg;
};

// matrix-core path: rows per wave, splits and columns per split of an I x J x d
// adjoint (a function of the shape only, so the workspace size is too)
static void adj_mfma_shape(int64_t I, int64_t J, int64_t d, AdjWork *w) {
    w->mfma = d <= 64;
    w->nd = (int)((d + 15) / 16);
    w->rb = 2;
    w->n_tiles = (int)((I + 16 * w->rb - 1) / (16 * w->rb));
    int64_t split = (kAdjMfmaTargetWaves + w->n_tiles - 1) / w->n_tiles;
    split = std::max<int64_t>(1, std::min<int64_t>(split, (J + 255) / 256));  // >= 256 columns per split
    w->split_cols = (int)round_up((J + split - 1) / split, kAdjMfmaStep);
    w->n_split = (int)((J + w->split_cols - 1) / w->split_cols);
}

static size_t adj_carve(Arena &ar, int64_t I, int64_t J, int64_t d, AdjWork *w) {
    adj_mfma_shape(I, J, d, w);
    w->r = ar.take<double>(I);
    w->sw = ar.take<double>(I);
    w->c = ar.take<double>(J);
    w->V = ar.take<double>((size_t)I * d);
    w->nparts = w->mfma ? w->n_split * w->n_tiles : (int)((I + kAdjRows - 1) / kAdjRows);
    w->partials = ar.take<double>(w->nparts);
    w->scal = ar.take<double>(8);
    w->Vp = w->rp = w->swp = w->bcol = w->Yp = w->cpart = w->cseg = nullptr;
    if (w->mfma) {
        w->Vp = ar.take<double>((size_t)w->n_split * I * d);
        w->rp = ar.take<double>((size_t)w->n_split * I);
        w->swp = ar.take<double>((size_t)w->n_split * I);
        w->bcol = ar.take<double>(J);
        w->Yp = ar.take<double>((size_t)J * 16 * w->nd);
        w->cpart = ar.take<double>((size_t)w->n_tiles * J);
        w->cseg = ar.take<double>((size_t)((w->n_tiles + kAdjColSeg - 1) / kAdjColSeg) * J);
    }
    return ar.off;
}

template <int ND>
static void launch_adj_mfma(const AdjMfmaArgs &a, int n_split, hipStream_t s) {
    const dim3 grid((unsigned)((a.n_tiles + kAdjMfmaWaves - 1) / kAdjMfmaWaves), (unsigned)n_split);
    adjoint_mfma_kernel<2, ND><<<grid, kAdjMfmaWaves * 64, 0, s>>>(a);
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
    WOTK_REQUIRE(d <= 128, "the adjoint keeps a row's gradient in registers: d <= 128");
    Arena ar(workspace, workspace_bytes);
    AdjWork w;
    adj_carve(ar, I, J, d, &w);
    if (!workspace || !ar.ok) return fail(WOTK_EWORKSPACE, "adjoint workspace too small");
    const double inv_eps = 1.0 / epsilon;
    const double inv_ij = 1.0 / ((double)I * (double)J);
    if (w.mfma) {
        adjoint_colexp_kernel<<<(unsigned)((J + 255) / 256), 256, 0, s>>>(g, (int)J, inv_eps, w.bcol);
        WOTK_LAUNCH_CHECK();
        // |C| <= c_lim keeps C/eps within 690 < 1000 ln 2 (FP32 rounding toward zero)
        const float c_lim = std::nextafter((float)(690.0 * epsilon), 0.f);
        adjoint_ypad_kernel<<<(unsigned)((J * 16 * w.nd + 255) / 256), 256, 0, s>>>(Y, (int)J, (int)d, 16 * w.nd, w.Yp);
        WOTK_LAUNCH_CHECK();
        AdjMfmaArgs a{C, ldc, (int)I, (int)J, (int)d, f, g, w.bcol, w.Yp, inv_eps, inv_ij, c_lim, w.split_cols, w.n_tiles,
                      ldc % 4 == 0 && reinterpret_cast<uintptr_t>(C) % 16 == 0, w.Vp, w.rp, w.swp, w.partials,
                      w.cpart};
        switch (w.nd) {
            case 1: launch_adj_mfma<1>(a, w.n_split, s); break;
            case 2: launch_adj_mfma<2>(a, w.n_split, s); break;
            case 3: launch_adj_mfma<3>(a, w.n_split, s); break;
            default: launch_adj_mfma<4>(a, w.n_split, s); break;
        }
        WOTK_LAUNCH_CHECK();
        const int64_t n = std::max<int64_t>(I * d, I);
        adjoint_split_sum_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(w.Vp, w.rp, w.swp, w.n_split, (int)I,
                                                                            (int)d, w.V, w.r, w.sw);
        WOTK_LAUNCH_CHECK();
        // the plan's column sums from the same pass (C is read once)
        const int n_seg = (w.n_tiles + kAdjColSeg - 1) / kAdjColSeg;
        adjoint_colseg_kernel<<<dim3((unsigned)((J + 63) / 64), (unsigned)n_seg), 256, 0, s>>>(w.cpart, w.n_tiles,
                                                                                            (int)J, w.cseg);
        WOTK_LAUNCH_CHECK();
        adjoint_colfinal_kernel<<<(unsigned)((J + 255) / 256), 256, 0, s>>>(w.cseg, n_seg, (int)J, w.c);
        WOTK_LAUNCH_CHECK();
    } else {
        switch ((d + 31) / 32) {
            case 3: adjoint_rows_kernel<3><<<w.nparts, kAdjThreads, 0, s>>>(C, ldc, (int)I, (int)J, f, g, inv_eps,
                                                                            inv_ij, Y, (int)d, w.V, w.r, w.sw,
                                                                            w.partials); break;
            default: adjoint_rows_kernel<4><<<w.nparts, kAdjThreads, 0, s>>>(C, ldc, (int)I, (int)J, f, g, inv_eps,
                                                                             inv_ij, Y, (int)d, w.V, w.r, w.sw,
                                                                             w.partials); break;
        }
        WOTK_LAUNCH_CHECK();
        const int rc = launch_plan_sums(CT, ldct, (int)J, (int)I, g, f, epsilon, w.c, s);
        if (rc) return rc;
    }
    adjoint_scalars_kernel<<<1, 1024, 0, s>>>(w.partials, w.nparts, w.c, (int)I, (int)J, G, lambda2, w.scal);
    WOTK_LAUNCH_CHECK();
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
