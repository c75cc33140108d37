// Batched fixed-iteration solves (transport_stablev2 semantics,
// wot/ot/optimal_transport.py:167-236) for a whole time course: every day pair
// (wot/ot/ot_model.py:124-201 solves them one at a time) times a grid of
// (epsilon, lambda1) combinations (the reference's parameter sweep is one
// process per combination, workflow/optimal_transport_validation_parameter_sweep.wdl).
//
// One launch per half-iteration sweeps all pairs and all combinations: a
// persistent grid pulls (pair, row-block) tiles from a work queue ordered
// longest-row first, and every cost entry read from HBM serves all E
// combinations.  Potentials are stored combination-major ([E][n padded to 4])
// so, for each combination, a wave reads its columns' potentials as one
// contiguous, coalesced stream.  The per-iteration
// epsilon / alpha / k table lives in device memory and the kernels read the
// iteration index from a device counter, so a block of iterations can be
// captured once in a hipGraph and replayed.
//
// Exponents use the row's own previous potential as a fixed shift:
//   y_ij = (f_i_old + h_j - C_ij) * log2(e)/eps,   S_i = sum_j 2^y_ij,
//   LSE_i = ln S_i - f_i_old/eps,
// (the terms are then plan entries, bounded by the row mass), with an
// online-maximum recomputation for any row whose sum under- or overflows.
#include <algorithm>
#include <array>
#include <cmath>
#include <memory>
#include <utility>
#include <vector>

#include "common.h"

namespace wotk {

constexpr int kMaxCombos = 16;

struct PairDev {
    long long a_off;   // element offset of this pair's matrix (C or C^T) in the cost buffer
    long long lda;
    int nrows, ncols;
    long long pot_off;    // offset of this side's potentials ([E][pot_ld]) in the pot buffer
    long long other_off;  // offset of the other side's potentials ([E][other_ld])
    int pot_ld, other_ld; // row counts padded to a multiple of 4
    long long logw_off;   // offset of the pair's log weights ([E][nrows]) or -1 for its log q
    int pair;
};

struct IterParam {  // one (iteration, pair, combination)
    double k;       // log2(e) / eps
    double eps;
    double alpha1, alpha2;
    int share;      // exponentials shared across eps groups at this iteration (see derived_power)
    int pad_;
};

// Exponentials shared across eps groups.  With three eps groups whose k =
// log2(e)/eps stand in integer ratios to the last (base) group's -- once the
// schedule's eps has converged, e.g. eps {0.01, 0.05, 0.1}: k_0 = 10 k_2,
// k_1 = 2 k_2 -- 2^(k_g y) = (2^(k_2 y))^m: a derived group g takes the
// base's term to the power m instead of its own exponential, with column
// factors relative to the base's reference potentials.  For any other ratio
// r = k_g / k_base the group still reuses the base's FP64 exponent: its own is
// r times it (one FP64 product instead of the sum, FMA and clamp), with its
// own exponential.  `mode` packs the powers: group 1's in bits 0-7, group 0's
// in bits 8-15 (0: own exponent, kRealPower: r times the base's).
constexpr int kRealPower = 255;
__device__ __host__ __forceinline__ constexpr int derived_power(int mode, int g, int ng) {
    return ng != 3 ? 0 : g == 1 ? (mode & 0xff) : g == 0 ? ((mode >> 8) & 0xff) : 0;
}

// t^m for 1 <= m < 32 by binary powering (constant m: the loop folds to the
// squarings and products m needs; m = 10 is 4 FP32 products)
__device__ __forceinline__ float powi(float t, int m) {
    float r = 1.f;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        if ((m >> b) & 1) r *= t;
        if ((m >> (b + 1)) == 0) break;
        t *= t;
    }
    return r;
}

// the combination whose potential combination e's column factors are relative
// to, for a reading sweep in `mode`: its group's reference, or the base
// group's for a derived group
__device__ __forceinline__ int factor_ref(int mode, int e, int nl, int ng) {
    const int g = e / nl;
    return derived_power(mode, g, ng) ? (ng - 1) * nl + nl / 2 : g * nl + nl / 2;
}

struct Tile {
    int side_pair;  // index into the PairDev table of this side
    int row0;
};

struct BatchArgs {
    const float *cost;
    const PairDev *pairs;
    const Tile *tiles;
    int ntiles;
    const IterParam *table;    // [iteration][pair][E]
    const int *iter;           // device iteration counter
    int E;
    int n_pairs;
    int side;                  // 0: f from g over C (alpha1), 1: g from f over C^T (alpha2)
    double *pot;               // this side's potentials
    const double *other;       // the other side's potentials
    const double *logw;        // log weights of the source rows ([pair][E][rows])
    const double *logq;        // log q of the target side ([pair][E])
    const double *log_n_other; // per side-pair log(#columns)
    int *queue;                // work-queue counter
    float *fac;                // shared kernel: this side's column factors ([E][pot_ld], like pot)
    const float *fac_other;    // shared kernel: the other side's column factors
    int k_next;                // shared kernel: factors are read at iteration it + k_next
    unsigned long long *exact_rows;  // count of rows recomputed exactly
    int *bad_list;             // rows left to the repair kernel: (side_pair, row, combination mask)
    int *n_bad;                // their count (reset every iteration)
    int group;                 // shared kernels: combinations per group (factors written); 0 otherwise
    float *fmax;               // shared kernels: max(0, factor exponent) per (pair, combination) of this side
    const float *fmax_other;   // the same for the other side's factors (read)
    float *fmax_reset;         // cleared by this side's repair launch (the array this sweep consumed)
    int n_fmax;                // entries of each fmax array (pairs x combinations)
    float *psi, *msh;          // LDS kernel: this side's column shifts (fraction, integer) per group, at
                               // the group's reference slot of [E][pot_ld] (for the other side's next sweep)
    const float *psi_other, *m_other;  // the same of the other side (read)
};

// queue combinations `mask` (bit e: combination e) of row `row` of side-pair
// `sp` for the repair kernel
__device__ __forceinline__ void queue_row(const BatchArgs &ba, int sp, int row, int mask) {
    const int i = atomicAdd(ba.n_bad, 1);
    ba.bad_list[3 * i] = sp;
    ba.bad_list[3 * i + 1] = row;
    ba.bad_list[3 * i + 2] = mask;
}

// a failing reference invalidates every factor relative to it: widen the mask
// to whole groups (NL combinations per group, reference NL / 2), and with the
// base group's reference also to the groups the next sweep (in mode nmode)
// derives from the base
template <int NG, int NL>
__device__ __forceinline__ int widen_to_groups(int m, int nmode = 0) {
    int w = m;
#pragma unroll
    for (int g = 0; g < NG; ++g)
        if ((m >> (g * NL + NL / 2)) & 1) w |= ((1 << NL) - 1) << (g * NL);
    if ((m >> ((NG - 1) * NL + NL / 2)) & 1) {
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (derived_power(nmode, g, NG)) w |= ((1 << NL) - 1) << (g * NL);
    }
    return w;
}

// max of non-negative floats through their integer bit patterns
__device__ __forceinline__ void atomic_max_nonneg(float *p, float v) {
    atomicMax(reinterpret_cast<int *>(p), __float_as_int(v));
}

// factor of combination e at a row from its new potential and the reference's
// (exponent read by the other side's next sweep with k = kn), and its
// non-negative exponent for the pair's running maximum (inf when not finite)
__device__ __forceinline__ float factor_of(double kn, double pot_e, double pot_ref, float *dpos) {
    const double d = kn * (pot_e - pot_ref);
    *dpos = d <= 0.0 ? 0.f : (d < 1.0e30 ? (float)d : 1.0e30f);
    return (float)exp2(d);
}

__device__ __forceinline__ float bexp2(float x) { return __builtin_amdgcn_exp2f(x); }

// the E combinations' parameters of pair `pair` at iteration `it`
__device__ __forceinline__ const IterParam *params_at(const BatchArgs &ba, int it, int pair) {
    return ba.table + ((size_t)it * ba.n_pairs + pair) * ba.E;
}

// 2^y for an FP64 exponent y, exact up to the FP32 exp2 of a fraction: y = n + r
// with n = rint(y) and |r| <= 1/2, so the rounding of the exponent to FP32 costs
// at most 2^-25 relative whatever |y| is (rounding y itself to FP32 costs
// |y| 2^-24 ln 2).  y is clamped below at -16384 (2^y is 0 in FP32 anyway; the
// clamp keeps the -inf of padding columns from giving NaN).  Above, y stays
// below a few hundred for finite potentials; 2^y overflows to inf from 128 on,
// which the row checks catch.
__device__ __forceinline__ float bexp2_split(double y) {
    y = fmax(y, -16384.0);
    const double n = rint(y);
    return ldexpf(bexp2((float)(y - n)), (int)n);
}

// the same for an exponent already clamped below (or a product of a clamped one
// with a ratio up to 2^16: rint stays within int range)
__device__ __forceinline__ float bexp2_split_nc(double y) {
    const double n = rint(y);
    return ldexpf(bexp2((float)(y - n)), (int)n);
}

// Exponent of a term split in FP32 arithmetic only (the LDS-staged kernel):
//   y = H_j + F_i - k c,  H_j = M_j + psi_j and F_i = N_i + phi_i (integer M, N;
//   |psi|, |phi| <= 1/2), k = k1 + k2 (FP32 high and low parts).
//   P = fl(k1 c), n = rint(P): t = n - k1 c comes exactly from one FMA (|t| <=
//   1/2 + ulp(P) / 2, rounded once), then t - k2 c + psi_j, and
//   2^(y - phi_i) = ldexp(exp2(t - k2 c + psi_j), M_j + N_i - n).
// The exponent's absolute error stays below 2^-23 whatever |y| is (the FP32
// cost itself carries |k c| 2^-24); 2^phi_i is a per-row constant the
// epilogue applies to the row sum.  mn = M_j + N_i: integers below 2^23 (the
// shifts are clamped at +-2^22, where 2^y is 0 or inf in FP32 anyway).
__device__ __forceinline__ float split_exp_f32(float c, float k1, float k2, float psi, float mn) {
    const float n = rintf(k1 * c);
    const float t = fmaf(-k2, c, fmaf(-k1, c, n));
    return ldexpf(bexp2(t + psi), (int)(mn - n));
}

constexpr double kShiftClamp = 4194304.0;  // 2^22

// fraction and integer part (as FP32) of a shift H (a potential times k),
// clamped at +-2^22 (a NaN shift clamps low: its terms vanish)
__device__ __forceinline__ void split_shift(double H, float *psi, float *m) {
    H = fmin(fmax(H, -kShiftClamp), kShiftClamp);
    const double M = rint(H);
    *psi = (float)(H - M);
    *m = (float)M;
}

__device__ __forceinline__ float4 bload4(const float *p) {
    float4 v;
    v.x = __builtin_nontemporal_load(p);
    v.y = __builtin_nontemporal_load(p + 1);
    v.z = __builtin_nontemporal_load(p + 2);
    v.w = __builtin_nontemporal_load(p + 3);
    return v;
}

// Exact recomputation of every (row, combination) of a tile with a per-lane
// online maximum (integer shifts, ldexp merges in FP64); writes
// l2[r * E + e] = log2 sum_j 2^((h^e_j - C_ij) k_e).  Used when a tile's
// fixed-shift sums under- or overflow.  Every thread of the block calls it.
// One row of a tile (row < nrows), all E combinations; k of combination e is
// prm[e].k.  Every thread of the block calls it (it synchronises).
template <int E, int THREADS>
__device__ __forceinline__ void row_exact(const float *A, long long lda, int row, int ncols, const double *h,
                                          long long other_ld, const IterParam *prm, int mask, double *red_a,
                                          double *red_b, double *l2) {
    constexpr int NW = THREADS / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        const float *rowp = A + (long long)row * lda;
        double sig[E];
        float s[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            sig[e] = -1.0e6;
            s[e] = 0.f;
        }
        for (int c = tid * 4; c < ncols; c += THREADS * 4) {
            const float4 a = bload4(rowp + c);
            const double av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (!((mask >> e) & 1)) continue;  // block-uniform
                const double k = prm[e].k;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double y = fma(-k, av[q], h[(long long)e * other_ld + c + q] * k);
                    if (y > sig[e]) {
                        const double ns = ceil(y);
                        s[e] = ldexpf(s[e], (int)(sig[e] - ns));
                        sig[e] = ns;
                    }
                    s[e] += bexp2((float)(y - sig[e]));
                }
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            double sg = sig[e], sd = (double)s[e];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double og = __shfl_xor(sg, off), os = __shfl_xor(sd, off);
                const double ng = fmax(sg, og);
                sd = ldexp(sd, (int)(sg - ng)) + ldexp(os, (int)(og - ng));
                sg = ng;
            }
            if (lane == 0) {
                red_a[e * NW + wave] = sg;
                red_b[e * NW + wave] = sd;
            }
        }
        __syncthreads();
        if (tid < E) {
            double sg = red_a[tid * NW], sd = red_b[tid * NW];
#pragma unroll
            for (int w = 1; w < NW; ++w) {
                const double og = red_a[tid * NW + w], os = red_b[tid * NW + w];
                const double ng = fmax(sg, og);
                sd = ldexp(sd, (int)(sg - ng)) + ldexp(os, (int)(og - ng));
                sg = ng;
            }
            l2[tid] = sg + log2(sd);
        }
        __syncthreads();
    }
}

// sum the per-lane FP32 partials of every (row, combination) over the block:
// an FP32 shuffle tree inside each wave (positive terms: < 4e-7 relative), FP64
// across waves; result in red_a[..][0..NW) summed by the caller
template <int RE, int NW>
__device__ __forceinline__ void lane_sums(const float *s, double (*red_a)[NW]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < RE; ++i) {
        float v = s[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) red_a[i][wave] = (double)v;
    }
}

// new potential of (row, combination) from l2 = log2 sum_j 2^((h_j - C_ij) k)
__device__ __forceinline__ double new_potential(const BatchArgs &ba, const PairDev &pd, const IterParam &ip, int row,
                                                int side_pair, double l2, int e) {
    const double lse = l2 * kLn2;  // ln sum_j exp((h_j - C_ij)/eps)
    const double lw = (pd.logw_off >= 0) ? ba.logw[pd.logw_off + (long long)e * pd.nrows + row]
                                         : ba.logq[(long long)pd.pair * ba.E + e];
    const double alpha = ba.side == 0 ? ip.alpha1 : ip.alpha2;
    return alpha * ip.eps * (lw + ba.log_n_other[side_pair] - lse);
}

template <int E, int ROWS, int THREADS>
__global__ __launch_bounds__(THREADS) void batch_sweep_kernel(BatchArgs ba) {
    constexpr int NW = THREADS / 64;
    constexpr int RE = ROWS * E;
    static_assert(RE <= THREADS, "one thread per (row, combination) in the epilogue");
    __shared__ int tile_sh;
    __shared__ int bad_row[ROWS];
    __shared__ double red_a[RE][NW];
    __shared__ double l2_sh[RE];  // log2 sum_j 2^((h_j - C_ij) k) per (row, combination)
    const int tid = threadIdx.x;
    const int it = *ba.iter;

    // Every iteration ends on a barrier: lane-specific work sits right after the
    // barrier at the top, never before the back-edge (where the compiler may
    // merge it with the tile fetch into a loop nest that parks those lanes
    // outside the barriers).
    for (;;) {
        if (tid == 0) tile_sh = atomicAdd(ba.queue, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(tile_sh);
        if (tid < ROWS) bad_row[tid] = 0;  // last read before the previous iteration's final barrier
        if (t >= ba.ntiles) break;
        const Tile tl = ba.tiles[t];
        const PairDev pd = ba.pairs[tl.side_pair];
        const IterParam *prm = params_at(ba, it, pd.pair);
        double kk[E];
#pragma unroll
        for (int e = 0; e < E; ++e) kk[e] = prm[e].k;
        const float *A = ba.cost + pd.a_off;
        const double *h = ba.other + pd.other_off;
        double *pot = ba.pot + pd.pot_off;
        const float *rowp[ROWS];
        double fk[ROWS][E];
        float s[ROWS][E];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int row = min(tl.row0 + r, pd.nrows - 1);
            rowp[r] = A + (long long)row * pd.lda;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                fk[r][e] = pot[(long long)e * pd.pot_ld + row] * kk[e];
                s[r][e] = 0.f;
            }
        }
        // main pass: fixed shift f_old * k, one exp2 per (entry, combination)
        for (int c = tid * 4; c < pd.ncols; c += THREADS * 4) {
            double cv[ROWS][4];
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
                const float4 a = bload4(rowp[r] + c);
                cv[r][0] = a.x;
                cv[r][1] = a.y;
                cv[r][2] = a.z;
                cv[r][3] = a.w;
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const double *he = h + (long long)e * pd.other_ld + c;
                const double2 h01 = *reinterpret_cast<const double2 *>(he);
                const double2 h23 = *reinterpret_cast<const double2 *>(he + 2);
                const double hk[4] = {h01.x * kk[e], h01.y * kk[e], h23.x * kk[e], h23.y * kk[e]};
#pragma unroll
                for (int r = 0; r < ROWS; ++r) {
                    float acc = 0.f;
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc += bexp2((float)fma(-kk[e], cv[r][q], hk[q] + fk[r][e]));
                    s[r][e] += acc;
                }
            }
        }
        lane_sums<RE, NW>(&s[0][0], red_a);
        __syncthreads();
        if (tid < RE) {
            double S = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) S += red_a[tid][w];
            const int r = tid / E, e = tid % E;
            l2_sh[tid] = log2(S) - fk[r][e];
            if (!(S > 0.0 && S < INFINITY) && tl.row0 + r < pd.nrows) atomicOr(&bad_row[r], 1 << e);
        }
        __syncthreads();
        // rare: the fixed shift under- or overflowed; the repair kernel recomputes
        // those (row, combination) exactly.  No barrier sits under a condition.
        if (tid < RE) {
            const int r = tid / E, e = tid % E;
            const int row = tl.row0 + r;
            const int m = bad_row[r];
            if (row < pd.nrows) {
                if (!((m >> e) & 1))
                    pot[(long long)e * pd.pot_ld + row] = new_potential(ba, pd, prm[e], row, tl.side_pair, l2_sh[tid], e);
                if (m && e == 0) queue_row(ba, tl.side_pair, row, m);
            }
        }
        __syncthreads();
    }
}

// Is a factored row sum S usable?  The terms are 2^(y_ref) * B_j with 2^(y_ref)
// from bexp2_split (exact to 2^-25 for any offset between the combinations); a
// term whose 2^(y_ref) underflows FP32 (below 2^-126, flushed or denormal) is
// lost, and it was worth less than 2^-126 * B_j <= 2^(-126 + dmax) with dmax
// the largest factor exponent of the pair's columns, so the lost mass of a row
// of ncols terms is below 2^(log2 ncols + dmax - 126).  S must exceed that by
// 2^30 and be finite.
__device__ __forceinline__ bool factored_sum_ok(double S, int ncols, float dmax) {
    const double lo = exp2(log2((double)ncols) + (double)dmax - 96.0);
    return S >= lo && S < INFINITY;
}

// Shared-exponential sweep for NG groups of NL combinations of equal epsilon
// (combination e = g * NL + l).  Within a group the reference combination
// l0 = NL / 2 gives the fixed-shift exponent
//   y_ij = k (f^ref_i,old + h^ref_j - C_ij),   one exp2 per (entry, group),
// and every combination of the group sums 2^y_ij * B^e_j with the column
// factor B^e_j = 2^(k (h^e_j - h^ref_j)) (FP32, written by the other side's
// epilogue), so that
//   log2 sum_j 2^((h^e_j - C_ij) k) = log2 sum_j 2^y_ij B^e_j - k f^ref_i,old.
// The terms are combination e's previous plan entries scaled by
// 2^(-k (f^e_i - f^ref_i)); a row whose sums leave the range that
// factored_sum_ok accepts is recomputed exactly by the repair kernel.
template <int NG, int NL, int ROWS, int THREADS>
__global__ __launch_bounds__(THREADS) void batch_shared_kernel(BatchArgs ba) {
    constexpr int E = NG * NL, L0 = NL / 2, NW = THREADS / 64, RE = ROWS * E;
    static_assert(RE <= THREADS, "one thread per (row, combination) in the epilogue");
    __shared__ int tile_sh;
    __shared__ int bad_row[ROWS];
    __shared__ float fmax_sh[E];
    __shared__ double red_a[RE][NW];
    __shared__ double l2_sh[RE];
    __shared__ double pot_sh[RE];
    const int tid = threadIdx.x;
    const int it = *ba.iter;
    if (tid < E) fmax_sh[tid] = 0.f;
    int prev_pair = -1;  // pair of the previous tile, whose factor maxima are flushed at the top

    // every iteration ends on a barrier (see batch_sweep_kernel)
    for (;;) {
        if (tid == 0) tile_sh = atomicAdd(ba.queue, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(tile_sh);
        if (tid < ROWS) bad_row[tid] = 0;
        if (tid < E && prev_pair >= 0) {
            atomic_max_nonneg(&ba.fmax[prev_pair * E + tid], fmax_sh[tid]);
            fmax_sh[tid] = 0.f;
        }
        if (t >= ba.ntiles) break;
        const Tile tl = ba.tiles[t];
        const PairDev pd = ba.pairs[tl.side_pair];
        prev_pair = pd.pair;
        const IterParam *prm = params_at(ba, it, pd.pair);
        const IterParam *nxt = params_at(ba, it + ba.k_next, pd.pair);  // k the factors are read with
        double kg[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) kg[g] = prm[g * NL + L0].k;
        const float *A = ba.cost + pd.a_off;
        const double *h = ba.other + pd.other_off;
        const float *B = ba.fac_other + pd.other_off;
        double *pot = ba.pot + pd.pot_off;
        const float *rowp[ROWS];
        double fk[ROWS][NG];
        float s[ROWS][E];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int row = min(tl.row0 + r, pd.nrows - 1);
            rowp[r] = A + (long long)row * pd.lda;
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                fk[r][g] = pot[(long long)(g * NL + L0) * pd.pot_ld + row] * kg[g];
#pragma unroll
                for (int l = 0; l < NL; ++l) s[r][g * NL + l] = 0.f;
            }
        }
        for (int c = tid * 4; c < pd.ncols; c += THREADS * 4) {
            double cv[ROWS][4];
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
                const float4 a = bload4(rowp[r] + c);
                cv[r][0] = a.x;
                cv[r][1] = a.y;
                cv[r][2] = a.z;
                cv[r][3] = a.w;
            }
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const double *he = h + (long long)(g * NL + L0) * pd.other_ld + c;
                const double2 h01 = *reinterpret_cast<const double2 *>(he);
                const double2 h23 = *reinterpret_cast<const double2 *>(he + 2);
                const double hk[4] = {h01.x * kg[g], h01.y * kg[g], h23.x * kg[g], h23.y * kg[g]};
                float4 bf[NL];
#pragma unroll
                for (int l = 0; l < NL; ++l)
                    if (l != L0) bf[l] = *reinterpret_cast<const float4 *>(B + (long long)(g * NL + l) * pd.other_ld + c);
#pragma unroll
                for (int r = 0; r < ROWS; ++r) {
                    float tq[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) tq[q] = bexp2_split(fma(-kg[g], cv[r][q], hk[q] + fk[r][g]));
                    s[r][g * NL + L0] += (tq[0] + tq[1]) + (tq[2] + tq[3]);
#pragma unroll
                    for (int l = 0; l < NL; ++l)
                        if (l != L0)
                            s[r][g * NL + l] += fmaf(tq[0], bf[l].x, fmaf(tq[1], bf[l].y, fmaf(tq[2], bf[l].z, tq[3] * bf[l].w)));
                }
            }
        }
        lane_sums<RE, NW>(&s[0][0], red_a);
        __syncthreads();
        if (tid < RE) {
            double S = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) S += red_a[tid][w];
            const int r = tid / E, e = tid % E, g = e / NL;
            const int row = tl.row0 + r;
            l2_sh[tid] = log2(S) - fk[r][g];
            if (row < pd.nrows &&
                !factored_sum_ok(S, pd.ncols, ba.fmax_other[pd.pair * E + e]))
                atomicOr(&bad_row[r], 1 << e);
        }
        __syncthreads();
        // combinations out of range go to the repair kernel (whole groups when the
        // reference failed); no barrier sits under a condition
        const int nmode = nxt[0].share;  // the reading sweep's mode: which reference each factor is relative to
        if (tid < RE) {
            const int r = tid / E, e = tid % E;
            const int row = tl.row0 + r;
            const int m = widen_to_groups<NG, NL>(bad_row[r], nmode);
            const double p = new_potential(ba, pd, prm[e], min(row, pd.nrows - 1), tl.side_pair, l2_sh[tid], e);
            pot_sh[tid] = p;
            if (row < pd.nrows) {
                if (!((m >> e) & 1)) pot[(long long)e * pd.pot_ld + row] = p;
                if (m && e == 0) queue_row(ba, tl.side_pair, row, m);
            }
        }
        __syncthreads();
        if (tid < RE) {
            // this side's factors, read by the other side's next sweep with k of that sweep
            const int r = tid / E, e = tid % E, g = e / NL;
            const int row = tl.row0 + r;
            if (row < pd.nrows && !((widen_to_groups<NG, NL>(bad_row[r], nmode) >> e) & 1)) {
                float dpos;
                ba.fac[pd.pot_off + (long long)e * pd.pot_ld + row] =
                    factor_of(nxt[g * NL + L0].k, pot_sh[tid], pot_sh[r * E + factor_ref(nmode, e, NL, NG)], &dpos);
                atomic_max_nonneg(&fmax_sh[e], dpos);
            }
        }
        __syncthreads();
    }
}

// LDS-staged shared-exponential sweep: a block of NWAVES waves takes a tile of
// NWAVES * WROWS rows, each wave owning WROWS of them; the other side's
// reference potentials and column factors are staged in LDS in chunks of
// 256 columns (double-buffered, loaded once per block instead of once per
// wave), so their L2 traffic is divided by the rows per tile.  Arithmetic,
// range checks and epilogue are those of batch_shared_kernel.
// One 256-column chunk per step and three LDS buffers (one barrier per
// chunk).  Each wave loads its rows' cost entries for the next chunk into
// registers before computing the current one, so the HBM latency is hidden
// behind the chunk's ~600 VALU instructions instead of stalling every wave of
// the block after the barrier; the kernel is compiled for 2 waves per SIMD,
// which gives it 256 VGPRs and no spills.  Measured on the 38-pair, 3 x 3
// grid time course, us per sweep for 4 waves x 4 rows: no prefetch at 3 waves
// per SIMD 5850 (spills 268 bytes); no prefetch at 2 waves 6800; prefetch at
// 3 waves 6470 (spills 352 bytes); prefetch at 2 waves 4880-4940.  (Earlier,
// without prefetch: 512-column chunks at 2 waves 5.32 / 10.6 ms for 4 x 2 /
// 4 x 4 rows, 256/2 buffers at 3 waves 5.92 / 5.11 ms.)  Dropping the per-entry
// clamp of the split exponent (clamping the potentials where they are staged
// instead) measured slower again with prefetch: 5117 us.
constexpr int kLdsChunk = 256;
constexpr int kLdsBuffers = 3;

// MODE (see derived_power): a group with an integer power takes the base
// group's term to that power; a group with a real ratio r takes r times the
// base's exponent, formed in FP64 (bexp2_split) in such modes; every other
// group forms its own exponent in FP32 (split_exp_f32) from the column shifts
// the other side's epilogue wrote for it.  Groups other than those on their
// own reference (power 0) have column factors for all NL combinations,
// relative to the base's reference potentials, and their row shift is the
// base's reference potential.
template <int NG, int NL, int WROWS, int NWAVES, int CH = kLdsChunk, int NBUF = kLdsBuffers, int MODE = 0>
__global__ __launch_bounds__(NWAVES * 64) __attribute__((amdgpu_waves_per_eu(2, 8))) void batch_shared_lds_kernel(
    BatchArgs ba) {
    constexpr int E = NG * NL, L0 = NL / 2, THREADS = NWAVES * 64, RE = WROWS * E;
    constexpr int PER = (CH + THREADS - 1) / THREADS;  // staged columns per thread
    constexpr int BASE = NG - 1;
    static_assert(MODE == 0 || NG == 3, "shared exponentials across groups need three eps groups");
    static_assert(RE <= 64, "one lane per (row, combination) of a wave in the epilogue");
    static_assert(CH % 256 == 0 && (NBUF == 2 || NBUF == 3), "chunk of 64 lanes x 4 columns, 2 or 3 buffers");
    constexpr auto pw = [](int g) constexpr { return derived_power(MODE, g, NG); };
    // REAL: some group takes a real ratio of the base's FP64 exponent
    constexpr bool REAL = pw(0) == kRealPower || pw(1) == kRealPower;
    // how each group's terms are formed: FP32 split of its own exponent, the
    // base's FP64 exponent (REAL modes), r times it, or an integer power
    constexpr auto split32 = [pw](int g) constexpr { return pw(g) == 0 && !(REAL && g == NG - 1); };
    constexpr auto full64 = [](int g) constexpr { return REAL && g == NG - 1; };
    // staged slots: the column shift (fraction, integer) of every group whose
    // exponent is formed here; NL - 1 factor rows of a group on its own
    // reference, NL otherwise
    constexpr auto staged = [split32, full64](int g) constexpr { return split32(g) || full64(g); };
    constexpr auto xslot = [staged](int g) constexpr {
        int n = 0;
        for (int x = 0; x < g; ++x) n += staged(x) ? 1 : 0;
        return n;
    };
    constexpr auto bslot = [pw](int g) constexpr {
        int n = 0;
        for (int x = 0; x < g; ++x) n += pw(x) == 0 ? NL - 1 : NL;
        return n;
    };
    constexpr int NX = xslot(NG), NBS = bslot(NG);
    __shared__ float psi_lds[NBUF][NX][CH];
    __shared__ float m_lds[NBUF][NX][CH];
    __shared__ float b_lds[NBUF][NBS][CH];
    __shared__ int tile_sh;
    __shared__ int bad_row[NWAVES * WROWS];
    __shared__ float fmax_sh[E];
    __shared__ double pot_sh[NWAVES][RE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int it = *ba.iter;
    if (tid < E) fmax_sh[tid] = 0.f;
    int prev_pair = -1;  // pair of the previous tile, whose factor maxima are flushed at the top

    // every iteration ends on a barrier (see batch_sweep_kernel)
    for (;;) {
        if (tid == 0) tile_sh = atomicAdd(ba.queue, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(tile_sh);
        if (tid < NWAVES * WROWS) bad_row[tid] = 0;
        if (tid < E && prev_pair >= 0) {
            atomic_max_nonneg(&ba.fmax[prev_pair * E + tid], fmax_sh[tid]);
            fmax_sh[tid] = 0.f;
        }
        if (t >= ba.ntiles) break;
        const Tile tl = ba.tiles[t];
        const PairDev pd = ba.pairs[tl.side_pair];
        prev_pair = pd.pair;
        const IterParam *prm = params_at(ba, it, pd.pair);
        const IterParam *nxt = params_at(ba, it + ba.k_next, pd.pair);
        double kg[NG], rg[NG];
        float k1[NG], k2[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            kg[g] = prm[g * NL + L0].k;
            k1[g] = (float)kg[g];
            k2[g] = (float)(kg[g] - (double)k1[g]);
        }
#pragma unroll
        for (int g = 0; g < NG; ++g) rg[g] = pw(g) == kRealPower ? kg[g] / kg[BASE] : 1.0;
        const float *A = ba.cost + pd.a_off;
        const float *B = ba.fac_other + pd.other_off;
        const float *PS = ba.psi_other + pd.other_off;
        const float *PM = ba.m_other + pd.other_off;
        double *pot = ba.pot + pd.pot_off;
        const int ncols4 = (pd.ncols + 3) & ~3;  // potentials / factors are padded to here
        const int nchunks = (ncols4 + CH - 1) / CH;
        const float *rowp[WROWS];
        float nrow[WROWS][NG];  // integer row shifts of the FP32-split groups
        double fkb[WROWS];      // REAL modes: the base's clamped FP64 row shift
        float s[WROWS][E];
#pragma unroll
        for (int r = 0; r < WROWS; ++r) {
            const int row = min(tl.row0 + wave * WROWS + r, pd.nrows - 1);
            rowp[r] = A + (long long)row * pd.lda;
            fkb[r] = 0.0;
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                nrow[r][g] = 0.f;
                if (split32(g)) {
                    float ps_;
                    split_shift(pot[(long long)(g * NL + L0) * pd.pot_ld + row] * kg[g], &ps_, &nrow[r][g]);
                }
                if (full64(g))
                    fkb[r] = fmin(fmax(pot[(long long)(g * NL + L0) * pd.pot_ld + row] * kg[g], -kShiftClamp), kShiftClamp);
#pragma unroll
                for (int l = 0; l < NL; ++l) s[r][g * NL + l] = 0.f;
            }
        }
        // staging: thread t holds columns t + p * THREADS of a chunk (zero-mass
        // padding beyond ncols4); loaded into registers one chunk ahead
        float psv[NX][PER], pmv[NX][PER];
        float bv[NBS][PER];
#define WOTK_LDS_FETCH(k)                                                                        \
    _Pragma("unroll") for (int p = 0; p < PER; ++p) {                                           \
        /* the column is clamped into the padded row, so every load is in bounds and none */    \
        /* sits under a lane mask; staged columns at or past ncols4 are never read (a lane */   \
        /* computes its 4 columns only below ncols4), and threads past CH commit nothing */     \
        const int cc_ = min((k) * CH + tid + p * THREADS, ncols4 - 1);                          \
        _Pragma("unroll") for (int g = 0; g < NG; ++g) {                                        \
            if (staged(g)) {                                                                    \
                const long long at_ = (long long)(g * NL + L0) * pd.other_ld + cc_;             \
                psv[xslot(g)][p] = PS[at_];                                                     \
                pmv[xslot(g)][p] = PM[at_];                                                     \
            }                                                                                   \
            if (pw(g) == 0) {                                                                   \
                _Pragma("unroll") for (int i = 0; i < NL - 1; ++i) {                            \
                    const int l_ = i < L0 ? i : i + 1;                                          \
                    bv[bslot(g) + i][p] = B[(long long)(g * NL + l_) * pd.other_ld + cc_];      \
                }                                                                               \
            } else {                                                                            \
                _Pragma("unroll") for (int l = 0; l < NL; ++l)                                  \
                    bv[bslot(g) + l][p] = B[(long long)(g * NL + l) * pd.other_ld + cc_];       \
            }                                                                                   \
        }                                                                                       \
    }
#define WOTK_LDS_COMMIT(buf)                                                                     \
    _Pragma("unroll") for (int p = 0; p < PER; ++p) {                                           \
        const int j_ = tid + p * THREADS;                                                       \
        if (j_ < CH) {                                                                          \
            _Pragma("unroll") for (int x = 0; x < NX; ++x) {                                    \
                psi_lds[buf][x][j_] = psv[x][p];                                                \
                m_lds[buf][x][j_] = pmv[x][p];                                                  \
            }                                                                                   \
            _Pragma("unroll") for (int x = 0; x < NBS; ++x) b_lds[buf][x][j_] = bv[x][p];       \
        }                                                                                       \
    }
        WOTK_LDS_FETCH(0)
        WOTK_LDS_COMMIT(0)
        static_assert(CH == 256, "prefetch assumes one 256-column step per chunk");
        // lanes past the row's (padded) end re-read its last 16 bytes: every load
        // is in bounds and needs no lane mask, and those entries are never used
        float4 cnext[WROWS];
#pragma unroll
        for (int r = 0; r < WROWS; ++r) cnext[r] = bload4(rowp[r] + min(lane * 4, ncols4 - 4));
        __syncthreads();
        for (int k = 0; k < nchunks; ++k) {
            const int buf = k % NBUF;
            const int cl = lane * 4;  // column within the chunk
            const int c = k * CH + cl;
            // this chunk's cost entries leave the prefetch registers before the next
            // chunk's loads (cost and staging) are issued: nothing below waits on them
            // until the commit
            float4 ccur[WROWS];
#pragma unroll
            for (int r = 0; r < WROWS; ++r) ccur[r] = cnext[r];
            if (k + 1 < nchunks) {  // wave-uniform; in flight while this chunk computes
                const int cn = min((k + 1) * CH + cl, ncols4 - 4);
#pragma unroll
                for (int r = 0; r < WROWS; ++r) cnext[r] = bload4(rowp[r] + cn);
                WOTK_LDS_FETCH(k + 1)
            }
            if (c < ncols4) {
                // the base group first: the others take powers of its terms, or
                // multiples of its exponents
                float tb[WROWS][4];
                double yb[WROWS][REAL ? 4 : 1];
#pragma unroll
                for (int gi = 0; gi < NG; ++gi) {
                    const int g = NG - 1 - gi;
                    if (!staged(g)) {
                        float4 bf[NL];
#pragma unroll
                        for (int l = 0; l < NL; ++l) bf[l] = *reinterpret_cast<const float4 *>(&b_lds[buf][bslot(g) + l][cl]);
#pragma unroll
                        for (int r = 0; r < WROWS; ++r) {
                            float tq[4];
                            if (pw(g) == kRealPower) {
#pragma unroll
                                for (int q = 0; q < 4; ++q) tq[q] = bexp2_split_nc(yb[r][REAL ? q : 0] * rg[g]);
                            } else {
#pragma unroll
                                for (int q = 0; q < 4; ++q) tq[q] = powi(tb[r][q], pw(g));
                            }
#pragma unroll
                            for (int l = 0; l < NL; ++l)
                                s[r][g * NL + l] = fmaf(tq[3], bf[l].w, fmaf(tq[2], bf[l].z, fmaf(tq[1], bf[l].y, fmaf(tq[0], bf[l].x, s[r][g * NL + l]))));
                        }
                    } else {
                        const float4 ps4 = *reinterpret_cast<const float4 *>(&psi_lds[buf][xslot(g)][cl]);
                        const float4 pm4 = *reinterpret_cast<const float4 *>(&m_lds[buf][xslot(g)][cl]);
                        const float psq[4] = {ps4.x, ps4.y, ps4.z, ps4.w};
                        const float pmq[4] = {pm4.x, pm4.y, pm4.z, pm4.w};
                        const int nf = pw(g) == 0 ? NL - 1 : NL;  // factor rows of the group
                        float4 bf[NL];
#pragma unroll
                        for (int i = 0; i < NL; ++i)
                            if (i < nf) bf[i] = *reinterpret_cast<const float4 *>(&b_lds[buf][bslot(g) + i][cl]);
#pragma unroll
                        for (int r = 0; r < WROWS; ++r) {
                            const float cq[4] = {ccur[r].x, ccur[r].y, ccur[r].z, ccur[r].w};
                            float tq[4];
                            if (full64(g)) {
                                // y = H_j + F_i - k c in FP64 (clamped shifts; y >= -16384 as in bexp2_split)
#pragma unroll
                                for (int q = 0; q < 4; ++q) {
                                    const double y = fmax(fma(-kg[g], (double)cq[q], ((double)pmq[q] + (double)psq[q]) + fkb[r]), -16384.0);
                                    yb[r][REAL ? q : 0] = y;
                                    tq[q] = bexp2_split_nc(y);
                                }
                            } else {
#pragma unroll
                                for (int q = 0; q < 4; ++q) tq[q] = split_exp_f32(cq[q], k1[g], k2[g], psq[q], pmq[q] + nrow[r][g]);
                            }
                            if (g == BASE) {
#pragma unroll
                                for (int q = 0; q < 4; ++q) tb[r][q] = tq[q];
                            }
                            if (pw(g) == 0) {
                                s[r][g * NL + L0] += (tq[0] + tq[1]) + (tq[2] + tq[3]);
                                int i = 0;
#pragma unroll
                                for (int l = 0; l < NL; ++l)
                                    if (l != L0) {
                                        s[r][g * NL + l] = fmaf(tq[3], bf[i].w, fmaf(tq[2], bf[i].z, fmaf(tq[1], bf[i].y, fmaf(tq[0], bf[i].x, s[r][g * NL + l]))));
                                        ++i;
                                    }
                            } else {
#pragma unroll
                                for (int l = 0; l < NL; ++l)
                                    s[r][g * NL + l] = fmaf(tq[3], bf[l].w, fmaf(tq[2], bf[l].z, fmaf(tq[1], bf[l].y, fmaf(tq[0], bf[l].x, s[r][g * NL + l]))));
                            }
                        }
                    }
                }
            }
            // with two buffers, every wave must be done with chunk k - 1 (same buffer
            // as k + 1) before it is overwritten; with three, the barrier that ended
            // iteration k - 1 already ordered that
            if (NBUF == 2) __syncthreads();
            if (k + 1 < nchunks) {
                WOTK_LDS_COMMIT((k + 1) % NBUF)
            }
            __syncthreads();
        }
        // wave sums (FP32 tree over 64 lanes); lane l < RE takes (row l / E, combination l % E)
        double S = 0.0;
#pragma unroll
        for (int r = 0; r < WROWS; ++r)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                float v = s[r][e];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
                if (lane == r * E + e) S = (double)v;
            }
        const int my_r = lane / E, my_e = lane % E;
        const int my_row = tl.row0 + wave * WROWS + my_r;
        const bool mine = lane < RE && my_row < pd.nrows;
        // l2 = log2 sum_j 2^((h^e_j - C_ij) k_e) from the row's shift (its old
        // potential, read before any potential of the tile is rewritten), and
        // the sum with the row's fractional shift for the range check
        double l2 = 0.0, S_true = 0.0;
        if (lane < RE) {
            const int g = my_e / NL, prow = min(my_row, pd.nrows - 1);
            const int m = derived_power(MODE, g, NG);
            const double fb = pot[(long long)(BASE * NL + L0) * pd.pot_ld + prow];
            // mu: the part of the row's shift F left out of the terms (its fraction
            // with the FP32 split, whose terms carry rint(clamped F); the clamp's
            // excess in REAL modes, whose FP64 exponents carry the clamped F)
            double mu, fk;
            if (m == 0) {
                const double F = pot[(long long)(g * NL + L0) * pd.pot_ld + prow] * kg[g];
                const double Fc = fmin(fmax(F, -kShiftClamp), kShiftClamp);
                mu = F - (REAL && g == BASE ? Fc : rint(Fc));
                fk = F;
            } else {
                const double Fb = fb * kg[BASE];
                const double Fbc = fmin(fmax(Fb, -kShiftClamp), kShiftClamp);
                const double mub = Fb - (REAL ? Fbc : rint(Fbc));
                mu = (m == kRealPower ? rg[g] : (double)m) * mub;
                fk = fb * kg[g];
            }
            S_true = S * exp2(mu);
            l2 = log2(S) + mu - fk;
        }
        if (mine && !factored_sum_ok(S_true, pd.ncols, ba.fmax_other[pd.pair * E + my_e]))
            atomicOr(&bad_row[wave * WROWS + my_r], 1 << my_e);
        __syncthreads();
        // combinations out of range go to the repair kernel (whole groups when the
        // reference failed); no barrier sits under a condition
        const int nmode = nxt[0].share;  // the reading sweep's mode: which reference each factor is relative to
        const int m = lane < RE ? widen_to_groups<NG, NL>(bad_row[wave * WROWS + my_r], nmode) : 0;
        const bool row_ok = mine && !((m >> my_e) & 1);
        if (lane < RE) {
            const double p = new_potential(ba, pd, prm[my_e], min(my_row, pd.nrows - 1), tl.side_pair, l2, my_e);
            pot_sh[wave][lane] = p;
            if (row_ok) pot[(long long)my_e * pd.pot_ld + my_row] = p;
            if (mine && m && my_e == 0) queue_row(ba, tl.side_pair, my_row, m);
        }
        __syncthreads();
        if (row_ok) {
            const int gref = (my_e / NL) * NL + L0;  // k of the combination's own group
            float dpos;
            ba.fac[pd.pot_off + (long long)my_e * pd.pot_ld + my_row] =
                factor_of(nxt[gref].k, pot_sh[wave][lane], pot_sh[wave][my_r * E + factor_ref(nmode, my_e, NL, NG)], &dpos);
            atomic_max_nonneg(&fmax_sh[my_e], dpos);
            if (my_e == gref) {  // the group's column shift for the reading sweep
                const int sr = derived_power(nmode, my_e / NL, NG) == 0 ? my_e : BASE * NL + L0;
                split_shift(pot_sh[wave][my_r * E + sr] * nxt[gref].k, &ba.psi[pd.pot_off + (long long)my_e * pd.pot_ld + my_row],
                            &ba.msh[pd.pot_off + (long long)my_e * pd.pot_ld + my_row]);
            }
        }
        __syncthreads();
    }
}

#undef WOTK_LDS_FETCH
#undef WOTK_LDS_COMMIT

// Exact recomputation (per-lane online maximum) of the rows a sweep queued,
// with the sweep's epilogue: potentials and, for the shared-exponential
// kernels, the column factors and their maximum exponent.  Runs after every
// sweep; block 0 also clears the factor maxima the sweep just consumed (the
// other side's, rewritten by the next sweep).  With no queued row it exits.
template <int E, int THREADS>
__global__ __launch_bounds__(THREADS) void batch_repair_kernel(BatchArgs ba) {
    constexpr int NW = THREADS / 64;
    __shared__ double red_a[E * NW];
    __shared__ double red_b[E * NW];
    __shared__ double l2_sh[E];
    __shared__ double pot_sh[E];
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && ba.group)
        for (int i = tid; i < ba.n_fmax; i += THREADS) ba.fmax_reset[i] = 0.f;
    const int nb = __builtin_amdgcn_readfirstlane(*ba.n_bad);
    if ((int)blockIdx.x >= nb) return;
    const int it = *ba.iter;
    if (blockIdx.x == 0 && tid == 0) atomicAdd(ba.exact_rows, (unsigned long long)nb);
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int sp = ba.bad_list[3 * b], row = ba.bad_list[3 * b + 1];
        const int mask = __builtin_amdgcn_readfirstlane(ba.bad_list[3 * b + 2]);
        const PairDev pd = ba.pairs[sp];
        const IterParam *prm = params_at(ba, it, pd.pair);
        const IterParam *nxt = params_at(ba, it + ba.k_next, pd.pair);
        const float *A = ba.cost + pd.a_off;
        const double *h = ba.other + pd.other_off;
        double *pot = ba.pot + pd.pot_off;
        row_exact<E, THREADS>(A, pd.lda, row, pd.ncols, h, pd.other_ld, prm, mask, red_a, red_b, l2_sh);
        // recomputed combinations take their new potential; the others keep the
        // one the sweep wrote (a group's reference is recomputed with its group)
        if (tid < E) {
            const long long at = (long long)tid * pd.pot_ld + row;
            const bool redo = (mask >> tid) & 1;
            const double p = redo ? new_potential(ba, pd, prm[tid], row, sp, l2_sh[tid], tid) : pot[at];
            pot_sh[tid] = p;
            if (redo) pot[at] = p;
        }
        __syncthreads();
        if (ba.group && tid < E && ((mask >> tid) & 1)) {
            const int ng = E / ba.group;
            const int ref = factor_ref(nxt[0].share, tid, ba.group, ng);
            const int gref = (tid / ba.group) * ba.group + ba.group / 2;  // k of the combination's own group
            float dpos;
            ba.fac[pd.pot_off + (long long)tid * pd.pot_ld + row] = factor_of(nxt[gref].k, pot_sh[tid], pot_sh[ref], &dpos);
            atomic_max_nonneg(&ba.fmax[pd.pair * E + tid], dpos);
            if (tid == gref) {  // the group's column shift (the LDS-staged kernel reads it)
                const int sr = derived_power(nxt[0].share, tid / ba.group, ng) == 0 ? tid : (ng - 1) * ba.group + ba.group / 2;
                split_shift(pot_sh[sr] * nxt[gref].k, &ba.psi[pd.pot_off + (long long)tid * pd.pot_ld + row],
                            &ba.msh[pd.pot_off + (long long)tid * pd.pot_ld + row]);
            }
        }
        __syncthreads();
    }
}

// advance the iteration counter and reset the two work queues (one thread)
__global__ void batch_advance_kernel(int *iter, int *queues) {
    iter[0] += 1;
    queues[0] = 0;  // work queues of the two sides
    queues[1] = 0;
    queues[2] = 0;  // repair counts of the two sides
    queues[3] = 0;
}

__global__ void batch_reset_kernel(int *iter, int *queues, unsigned long long *exact_rows, int start) {
    iter[0] = start;
    queues[0] = 0;
    queues[1] = 0;
    queues[2] = 0;
    queues[3] = 0;
    exact_rows[0] = 0;
}

__global__ void batch_zero_f32_kernel(float *x, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = 0.f;
}

__global__ void batch_fill_kernel(double *x, long long n, double v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] = v;
}

// log weights of the source rows and log q = log(mean G) per (pair,
// combination): block (p, e); the growth of combination e starts at
// G + g_off[p] + e * n_p when every combination has its own (per_combo), else
// all share G + g_off[p]
__global__ __launch_bounds__(256) void batch_growth_kernel(const double *__restrict__ G, const long long *__restrict__ g_off,
                                                           const int *__restrict__ n_src, double *__restrict__ logw,
                                                           const long long *__restrict__ lw_off, double *__restrict__ logq,
                                                           int E, int per_combo) {
    const int p = blockIdx.x, e = blockIdx.y;
    const int n = n_src[p];
    const double *g = G + g_off[p] + (per_combo ? (long long)e * n : 0);
    double *lw = logw + lw_off[p] + (long long)e * n;
    __shared__ double red[4];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        lw[i] = log(g[i]);
        acc += g[i];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) logq[(long long)p * E + e] = log((red[0] + red[1] + red[2] + red[3]) / n);
}

// -inf on the padding rows [n, n_pad) of every (pair, combination) block:
// read as "other" potentials they make the padded columns contribute 0
__global__ void batch_pad_kernel(double *__restrict__ pot, const PairDev *__restrict__ pairs, int n_pairs, int E) {
    const int p = blockIdx.x;
    if (p >= n_pairs) return;
    const PairDev pd = pairs[p];
    for (int t = threadIdx.x; t < E * 4; t += blockDim.x) {
        const int e = t / 4, i = pd.nrows + t % 4;
        if (i < pd.pot_ld) pot[pd.pot_off + (long long)e * pd.pot_ld + i] = -INFINITY;
    }
}

// column shifts of the LDS-staged kernel: potentials start at 0 (shift 0),
// padding rows [n, n_pad) at -2^22 (their terms vanish)
__global__ void batch_shift_init_kernel(float *__restrict__ psi, float *__restrict__ msh,
                                        const PairDev *__restrict__ pairs, int n_pairs, int E) {
    const int p = blockIdx.x;
    if (p >= n_pairs) return;
    const PairDev pd = pairs[p];
    for (int t = threadIdx.x; t < E * pd.pot_ld; t += blockDim.x) {
        const int i = t % pd.pot_ld;
        psi[pd.pot_off + t] = 0.f;
        msh[pd.pot_off + t] = i < pd.nrows ? 0.f : -(float)kShiftClamp;
    }
}

// factors of the shared-exponential kernel: 1 on the rows (potentials start
// equal across combinations), 0 on the padding rows [n, n_pad)
__global__ void batch_fac_init_kernel(float *__restrict__ fac, const PairDev *__restrict__ pairs, int n_pairs, int E) {
    const int p = blockIdx.x;
    if (p >= n_pairs) return;
    const PairDev pd = pairs[p];
    for (int t = threadIdx.x; t < E * pd.pot_ld; t += blockDim.x) {
        const int i = t % pd.pot_ld;
        fac[pd.pot_off + t] = i < pd.nrows ? 1.f : 0.f;
    }
}

struct BatchWork {
    double *f, *g, *logw, *logq, *lognf, *logng;
    float *ff, *fg;  // column factors of the shared-exponential kernel
    float *psf, *msf, *psg, *msg;  // column shifts of the LDS-staged kernel
    IterParam *table;
    PairDev *pairs_f, *pairs_g;
    Tile *tiles_f, *tiles_g;
    long long *g_off, *lw_off;
    int *n_src;
    int *iter, *queues;
    unsigned long long *exact_rows;
    int *bad_f, *bad_g;      // repair lists: (side_pair, row, combination mask)
    float *fmax_f, *fmax_g;  // factor maxima per (pair, combination)
};

}  // namespace wotk

using namespace wotk;

static size_t batch_carve(Arena &ar, const wotk_batch_pair *pairs, int n_pairs, int E, int64_t n_iter,
                          int64_t *nf_out, int64_t *ng_out, BatchWork *w, int64_t *ntf, int64_t *ntg) {
    int64_t nf = 0, ng = 0, nlw = 0, tf = 0, tg = 0;
    for (int p = 0; p < n_pairs; ++p) {
        nf += round_up(pairs[p].n_i, 4);
        ng += round_up(pairs[p].n_j, 4);
        nlw += pairs[p].n_i;
        tf += pairs[p].n_i;  // tile lists sized for the smallest tile (one row)
        tg += pairs[p].n_j;
    }
    w->f = ar.take<double>((size_t)nf * E);
    w->g = ar.take<double>((size_t)ng * E);
    w->ff = ar.take<float>((size_t)nf * E);
    w->fg = ar.take<float>((size_t)ng * E);
    w->psf = ar.take<float>((size_t)nf * E);
    w->msf = ar.take<float>((size_t)nf * E);
    w->psg = ar.take<float>((size_t)ng * E);
    w->msg = ar.take<float>((size_t)ng * E);
    w->logw = ar.take<double>((size_t)nlw * E);
    w->logq = ar.take<double>((size_t)n_pairs * E);
    w->lognf = ar.take<double>(n_pairs);
    w->logng = ar.take<double>(n_pairs);
    w->table = ar.take<IterParam>((size_t)(n_iter + 1) * n_pairs * E);
    w->pairs_f = ar.take<PairDev>(n_pairs);
    w->pairs_g = ar.take<PairDev>(n_pairs);
    w->tiles_f = ar.take<Tile>(tf);
    w->tiles_g = ar.take<Tile>(tg);
    w->bad_f = ar.take<int>(3 * tf);
    w->bad_g = ar.take<int>(3 * tg);
    w->fmax_f = ar.take<float>((size_t)n_pairs * E);
    w->fmax_g = ar.take<float>((size_t)n_pairs * E);
    w->g_off = ar.take<long long>(n_pairs);
    w->lw_off = ar.take<long long>(n_pairs);
    w->n_src = ar.take<int>(n_pairs);
    w->iter = ar.take<int>(4);
    w->queues = ar.take<int>(4);
    w->exact_rows = ar.take<unsigned long long>(1);
    *nf_out = nf;
    *ng_out = ng;
    *ntf = tf;
    *ntg = tg;
    return ar.off;
}

extern "C" size_t wotk_batch_workspace_size(const wotk_batch_pair *pairs, int32_t n_pairs, int32_t n_combos,
                                            int64_t n_iterations) {
    Arena ar(nullptr, 0);
    BatchWork w;
    int64_t a, b, c, d;
    return batch_carve(ar, pairs, n_pairs, n_combos, n_iterations, &a, &b, &w, &c, &d);
}

// launchers per (combinations, tile variant); variants in the order of kBatchSweepTiles
typedef int (*SweepFn)(const BatchArgs &, const BatchArgs &, int, int, int, int *, int *, hipStream_t);
constexpr int kRepairThreads = 256;
typedef int (*OccFn)();  // resident workgroups per CU

// compiled modes (powers of groups 1 and 0, see derived_power) of the
// LDS-staged kernel for three eps groups of three combinations
constexpr int kDerivedModes[] = {2 | (10 << 8), 2 | (4 << 8), 2 | (kRealPower << 8), 1 | (1 << 8),
                                  kRealPower | (kRealPower << 8)};
constexpr int kNumDerivedModes = sizeof(kDerivedModes) / sizeof(kDerivedModes[0]);

struct BatchVariant {
    SweepFn sweep;
    OccFn occupancy;
    SweepFn derived[kNumDerivedModes];  // per kDerivedModes entry; nullptr where not compiled
};

template <int E, int V>
static int launch_sweeps(const BatchArgs &af, const BatchArgs &ag, int grid_f, int grid_g, int grid_r, int *iter,
                         int *queues, hipStream_t s) {
    constexpr int R = kBatchSweepTiles[V].rows, T = kBatchSweepTiles[V].threads;
    batch_sweep_kernel<E, R, T><<<grid_f, T, 0, s>>>(af);
    batch_repair_kernel<E, kRepairThreads><<<grid_r, kRepairThreads, 0, s>>>(af);
    batch_sweep_kernel<E, R, T><<<grid_g, T, 0, s>>>(ag);
    batch_repair_kernel<E, kRepairThreads><<<grid_r, kRepairThreads, 0, s>>>(ag);
    batch_advance_kernel<<<1, 1, 0, s>>>(iter, queues);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

template <int E, int V>
static int occupancy() {
    constexpr int R = kBatchSweepTiles[V].rows, T = kBatchSweepTiles[V].threads;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_sweep_kernel<E, R, T>, T, 0) != hipSuccess)
        return 1;
    return per < 1 ? 1 : per;
}

template <int E, int... V>
static constexpr std::array<BatchVariant, sizeof...(V)> variants_of(std::integer_sequence<int, V...>) {
    return {BatchVariant{launch_sweeps<E, V>, occupancy<E, V>, {}}...};
}

template <int E>
static const BatchVariant *variants() {
    static constexpr auto t = variants_of<E>(std::make_integer_sequence<int, kNumBatchSweepTiles>());
    return t.data();
}

static const BatchVariant *pick(int E) {
    switch (E) {
        case 1: return variants<1>();
        case 2: return variants<2>();
        case 3: return variants<3>();
        case 4: return variants<4>();
        case 6: return variants<6>();
        case 9: return variants<9>();
        case 12: return variants<12>();
        case 16: return variants<16>();
        default: return nullptr;
    }
}

// shared-exponential launchers per (groups, group size, tile variant of kBatchSharedTiles)
// the kernel of variant V: register-only (wave_rows == 0) or LDS-staged
template <int NG, int NL, int V, int MODE = 0>
static constexpr void (*shared_kernel_of())(BatchArgs) {
    constexpr TileShape ts = kBatchSharedTiles[V];
    if constexpr (ts.wave_rows == 0) {
        static_assert(MODE == 0, "derived groups need the LDS-staged kernel");
        return batch_shared_kernel<NG, NL, ts.rows, ts.threads>;
    } else {
        return batch_shared_lds_kernel<NG, NL, ts.wave_rows, ts.threads / 64, kLdsChunk, kLdsBuffers, MODE>;
    }
}

template <int NG, int NL, int V, int MODE = 0>
static int launch_shared(const BatchArgs &af, const BatchArgs &ag, int grid_f, int grid_g, int grid_r, int *iter,
                         int *queues, hipStream_t s) {
    constexpr int T = kBatchSharedTiles[V].threads;
    static_assert(kBatchSharedTiles[V].wave_rows == 0 || kBatchSharedTiles[V].rows == kBatchSharedTiles[V].wave_rows * T / 64,
                  "LDS variants: rows = waves x rows per wave");
    constexpr auto kern = shared_kernel_of<NG, NL, V, MODE>();
    kern<<<grid_f, T, 0, s>>>(af);
    batch_repair_kernel<NG * NL, kRepairThreads><<<grid_r, kRepairThreads, 0, s>>>(af);
    kern<<<grid_g, T, 0, s>>>(ag);
    batch_repair_kernel<NG * NL, kRepairThreads><<<grid_r, kRepairThreads, 0, s>>>(ag);
    batch_advance_kernel<<<1, 1, 0, s>>>(iter, queues);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

template <int NG, int NL, int V>
static int occupancy_shared() {
    constexpr int T = kBatchSharedTiles[V].threads;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, shared_kernel_of<NG, NL, V>(), T, 0) != hipSuccess)
        return 1;
    return per < 1 ? 1 : per;
}

template <int NG, int NL, int V, int M>
static constexpr SweepFn derived_sweep() {
    if constexpr (NG == 3 && NL == 3 && kBatchSharedTiles[V].wave_rows > 0)
        return launch_shared<NG, NL, V, kDerivedModes[M]>;
    else
        return nullptr;
}

template <int NG, int NL, int V, int... M>
static constexpr BatchVariant shared_variant(std::integer_sequence<int, M...>) {
    return BatchVariant{launch_shared<NG, NL, V>, occupancy_shared<NG, NL, V>, {derived_sweep<NG, NL, V, M>()...}};
}

template <int NG, int NL, int... V>
static constexpr std::array<BatchVariant, sizeof...(V)> shared_variants_of(std::integer_sequence<int, V...>) {
    return {shared_variant<NG, NL, V>(std::make_integer_sequence<int, kNumDerivedModes>())...};
}

template <int NG, int NL>
static const BatchVariant *shared_variants() {
    static constexpr auto t = shared_variants_of<NG, NL>(std::make_integer_sequence<int, kNumBatchSharedTiles>());
    return t.data();
}

// compiled (groups, group size) shapes of the shared-exponential kernel
static const BatchVariant *pick_shared(int ng, int nl) {
    switch (ng * 16 + nl) {
        case 1 * 16 + 2: return shared_variants<1, 2>();
        case 1 * 16 + 3: return shared_variants<1, 3>();
        case 2 * 16 + 2: return shared_variants<2, 2>();
        case 2 * 16 + 3: return shared_variants<2, 3>();
        case 3 * 16 + 2: return shared_variants<3, 2>();
        case 3 * 16 + 3: return shared_variants<3, 3>();
        case 4 * 16 + 4: return shared_variants<4, 4>();
        default: return nullptr;
    }
}

static int persistent_grid(const BatchVariant &v) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1024;
    return ncu * v.occupancy();
}

// Host state of one batched run between wotk_batch_begin and wotk_batch_end:
// the kernel arguments, the captured graph and the timing events.  Every HIP
// object it holds is released by its destructor, whatever path ends the run.
struct wotk_batch_run {
    BatchArgs af{}, ag{};
    SweepFn sweep[1 + kNumDerivedModes] = {};  // 0: own exponentials per group; i: kDerivedModes[i - 1]
    int grid_f = 0, grid_g = 0, grid_r = 0;
    int E = 0, n_pairs = 0, group = 0, graph_iters = 0;
    int64_t n_iter = 0, done = 0;
    BatchWork w{};
    std::vector<PairDev> pf, pg;
    std::vector<int64_t> n_i, n_j;
    std::vector<double> eps_final;  // [pair][E]
    std::vector<unsigned char> mode_ix;  // sweep index per iteration (n_iter + 1)
    hipGraph_t graph[1 + kNumDerivedModes] = {};
    hipGraphExec_t exec[1 + kNumDerivedModes] = {};
    hipStream_t cs = nullptr;  // capture stream
    bool capturing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timing;

    ~wotk_batch_run() {
        if (capturing) {
            hipGraph_t partial = nullptr;
            if (hipStreamEndCapture(cs, &partial) == hipSuccess && partial) (void)hipGraphDestroy(partial);
        }
        for (auto &ev : timing) {
            if (ev.first) (void)hipEventDestroy(ev.first);
            if (ev.second) (void)hipEventDestroy(ev.second);
        }
        for (hipGraphExec_t x : exec)
            if (x) (void)hipGraphExecDestroy(x);
        for (hipGraph_t x : graph)
            if (x) (void)hipGraphDestroy(x);
        if (cs) (void)hipStreamDestroy(cs);
    }
};

extern "C" int wotk_batch_begin(const float *cost, const wotk_batch_pair *pairs, int32_t n_pairs, const double *G,
                                int32_t growth_per_combo, int32_t n_combos, const double *epsilon,
                                const double *lambda1, const double *lambda2, int64_t scaling_iter, double epsilon0,
                                int64_t extra_iter, int64_t inner_iter_max, int32_t graph_iters, int32_t flags,
                                void *workspace, size_t workspace_bytes, wotk_batch_run **run_out, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    const int E = n_combos;
    WOTK_REQUIRE(run_out != nullptr, "run handle pointer is required");
    *run_out = nullptr;
    WOTK_REQUIRE(n_pairs > 0 && pairs && cost && G && epsilon && lambda1 && lambda2, "empty batch");
    WOTK_REQUIRE(E >= 1 && E <= kMaxCombos && pick(E), "unsupported number of combinations %d", E);
    WOTK_REQUIRE(scaling_iter >= 0 && extra_iter >= 0 && scaling_iter + extra_iter > 0 && inner_iter_max > 0,
                 "bad iteration counts");
    WOTK_REQUIRE(scaling_iter + extra_iter < (int64_t)1 << 30, "too many iterations");
    for (int p = 0; p < n_pairs; ++p) {
        const wotk_batch_pair &q = pairs[p];
        WOTK_REQUIRE(q.n_i > 0 && q.n_j > 0 && q.n_i < (1 << 30) && q.n_j < (1 << 30), "pair %d: bad shape", p);
        WOTK_REQUIRE(q.ldc % 4 == 0 && q.ldc >= round_up(q.n_j, 4) && q.ldct % 4 == 0 && q.ldct >= round_up(q.n_i, 4),
                     "pair %d: leading dimensions must be multiples of 4 covering the padded rows", p);
        WOTK_REQUIRE(q.c_offset % 4 == 0 && q.ct_offset % 4 == 0, "pair %d: offsets must be multiples of 4", p);
        for (int e = 0; e < E; ++e) {
            const double ep = epsilon[p * E + e], l1 = lambda1[p * E + e], l2 = lambda2[p * E + e];
            WOTK_REQUIRE(ep > 0.0 && l1 > 0.0 && l2 > 0.0 && std::isfinite(ep) && std::isfinite(l1) && std::isfinite(l2),
                         "pair %d, combination %d: epsilon and lambdas must be positive", p, e);
        }
    }
    const int64_t n_iter = scaling_iter + extra_iter;
    std::unique_ptr<wotk_batch_run> run(new wotk_batch_run());
    Arena ar(workspace, workspace_bytes);
    BatchWork &w = run->w;
    int64_t nf, ng, ntf, ntg;
    batch_carve(ar, pairs, n_pairs, E, n_iter, &nf, &ng, &w, &ntf, &ntg);
    if (!workspace || !ar.ok) return fail(WOTK_EWORKSPACE, "batch workspace too small");

    // tile variant from the shape bucket: combinations x mean row length over both sides
    double entries = 0.0, rows_total = 0.0;
    for (int p = 0; p < n_pairs; ++p) {
        entries += 2.0 * (double)pairs[p].n_i * (double)pairs[p].n_j;
        rows_total += (double)pairs[p].n_i + (double)pairs[p].n_j;
    }
    // shared exponentials: the largest group size of consecutive equal epsilons
    // (in every pair) with a compiled shape
    int group = 0;
    if (flags & WOTK_BATCH_SHARE_EXP)
        for (int nl = E; nl >= 2 && !group; --nl) {
            if (E % nl || !pick_shared(E / nl, nl)) continue;
            bool equal = true;
            for (int p = 0; p < n_pairs; ++p)
                for (int e = 0; e < E; ++e) equal = equal && epsilon[p * E + e] == epsilon[p * E + e - e % nl];
            if (equal) group = nl;
        }
    const int64_t mean_len = (int64_t)(entries / rows_total);
    int vi, tile_rows;
    const BatchVariant *variant;
    if (group) {
        vi = tile_for(WOTK_KERNEL_BATCH_SHARED, 0, mean_len, E);
        variant = &pick_shared(E / group, group)[vi];
        tile_rows = kBatchSharedTiles[vi].rows;
    } else {
        vi = tile_for(WOTK_KERNEL_BATCH_SWEEP, 0, mean_len, E);
        variant = &pick(E)[vi];
        tile_rows = kBatchSweepTiles[vi].rows;
    }

    // host-side tables: pairs, tiles (longest rows first), iteration parameters
    std::vector<PairDev> &pf = run->pf, &pg = run->pg;
    pf.resize(n_pairs);
    pg.resize(n_pairs);
    std::vector<long long> goff(n_pairs), lwoff(n_pairs);
    std::vector<int> nsrc(n_pairs);
    std::vector<double> lognf(n_pairs), logng(n_pairs);
    long long of = 0, og = 0, olw = 0;
    for (int p = 0; p < n_pairs; ++p) {
        const wotk_batch_pair &q = pairs[p];
        const int ldi = (int)round_up(q.n_i, 4), ldj = (int)round_up(q.n_j, 4);
        pf[p] = PairDev{q.c_offset, q.ldc, (int)q.n_i, (int)q.n_j, of * E, og * E, ldi, ldj, olw * E, p};
        pg[p] = PairDev{q.ct_offset, q.ldct, (int)q.n_j, (int)q.n_i, og * E, of * E, ldj, ldi, -1, p};
        goff[p] = q.g_offset;
        lwoff[p] = olw * E;
        nsrc[p] = (int)q.n_i;
        lognf[p] = std::log((double)q.n_j);
        logng[p] = std::log((double)q.n_i);
        run->n_i.push_back(q.n_i);
        run->n_j.push_back(q.n_j);
        of += round_up(q.n_i, 4);
        og += round_up(q.n_j, 4);
        olw += q.n_i;
    }
    auto tiles_for = [&](const std::vector<PairDev> &pd) {
        std::vector<Tile> t;
        for (int p = 0; p < n_pairs; ++p)
            for (int r = 0; r < pd[p].nrows; r += tile_rows) t.push_back(Tile{p, r});
        std::stable_sort(t.begin(), t.end(), [&](const Tile &a, const Tile &b) {
            return pd[a.side_pair].ncols > pd[b.side_pair].ncols;
        });
        return t;
    };
    std::vector<Tile> tf = tiles_for(pf), tg = tiles_for(pg);
    // eps(n) = (eps0 - eps) e^-n + eps, changed every inner_iter_max
    // iterations, then extra_iter iterations at the last value
    // (optimal_transport.py:184-185, 215-224, 231-233)
    std::vector<IterParam> table((size_t)(n_iter + 1) * n_pairs * E);
    run->eps_final.resize((size_t)n_pairs * E);
    for (int p = 0; p < n_pairs; ++p)
        for (int e = 0; e < E; ++e) {
            const double ep = epsilon[p * E + e], l1 = lambda1[p * E + e], l2 = lambda2[p * E + e];
            auto reg = [&](int64_t n) { return (epsilon0 - ep) * std::exp(-(double)n) + ep; };
            for (int64_t n = 0; n <= n_iter; ++n) {
                const double eps = n < scaling_iter ? reg(n / inner_iter_max) : reg(scaling_iter / inner_iter_max);
                table[((size_t)n * n_pairs + p) * E + e] = IterParam{kLog2E / eps, eps, l1 / (l1 + eps), l2 / (l2 + eps)};
            }
            run->eps_final[(size_t)p * E + e] = table[((size_t)n_iter * n_pairs + p) * E + e].eps;
        }
    // exponentials shared across eps groups: per iteration, the compiled mode
    // whose powers are k_g / k_base of every pair within 1e-10 relative (the
    // exponent then differs from the group's own by < 1e-10 of itself)
    run->mode_ix.assign((size_t)n_iter + 1, 0);
    bool any_derived = false;
    for (const SweepFn fn : variant->derived) any_derived = any_derived || fn;
    if ((flags & WOTK_BATCH_SHARE_EPS) && group && E / group == 3 && any_derived) {
        const int L0 = group / 2;
        for (int64_t n = 0; n <= n_iter; ++n) {
            int pw[2] = {0, 0};  // common integer power of groups 0 and 1 over every pair (0: none)
            bool in_range = true;  // real ratios below 2^16 keep r * y within int range (y >= -16384)
            for (int g = 0; g < 2; ++g) {
                int m = -1;
                for (int p = 0; p < n_pairs; ++p) {
                    const IterParam *t = &table[((size_t)n * n_pairs + p) * E];
                    const double r = t[g * group + L0].k / t[2 * group + L0].k;
                    in_range = in_range && r > 0.0 && r <= 65536.0;
                    const long mr = std::lround(r);
                    const int ok = (mr >= 1 && mr < 32 && std::fabs(r - (double)mr) <= 1e-10 * (double)mr) ? (int)mr : 0;
                    m = (m < 0 || m == ok) ? ok : 0;
                }
                pw[g] = m > 0 ? m : 0;
            }
            // the compiled mode with these powers (a group without one takes the real
            // ratio), else with group 0, then both, on the real ratio (power 1: equal
            // eps, e.g. every group at eps0 before the schedule's first step)
            const int p1 = pw[1] ? pw[1] : kRealPower, p0 = pw[0] ? pw[0] : kRealPower;
            const int wants[3] = {p1 | (p0 << 8), p1 | (kRealPower << 8), kRealPower | (kRealPower << 8)};
            int ix = 0;
            for (int a = 0; a < 3 && !ix && in_range; ++a)
                for (int i = 0; i < kNumDerivedModes; ++i)
                    if (kDerivedModes[i] == wants[a] && variant->derived[i]) {
                        ix = i + 1;
                        break;
                    }
            run->mode_ix[(size_t)n] = (unsigned char)ix;
            const int code = ix ? kDerivedModes[ix - 1] : 0;
            for (size_t q = 0; q < (size_t)n_pairs * E; ++q) table[(size_t)n * n_pairs * E + q].share = code;
        }
    }
    WOTK_HIP(hipMemcpyAsync(w.pairs_f, pf.data(), sizeof(PairDev) * n_pairs, hipMemcpyHostToDevice, s));
    WOTK_HIP(hipMemcpyAsync(w.pairs_g, pg.data(), sizeof(PairDev) * n_pairs, hipMemcpyHostToDevice, s));
    WOTK_HIP(hipMemcpyAsync(w.tiles_f, tf.data(), sizeof(Tile) * tf.size(), hipMemcpyHostToDevice, s));
    WOTK_HIP(hipMemcpyAsync(w.tiles_g, tg.data(), sizeof(Tile) * tg.size(), hipMemcpyHostToDevice, s));
    WOTK_HIP(hipMemcpyAsync(w.table, table.data(), sizeof(IterParam) * table.size(), hipMemcpyHostToDevice, s));
    WOTK_HIP(hipMemcpyAsync(w.g_off, goff.data(), sizeof(long long) * n_pairs, hipMemcpyHostToDevice, s));
    WOTK_HIP(hipMemcpyAsync(w.lw_off, lwoff.data(), sizeof(long long) * n_pairs, hipMemcpyHostToDevice, s));
    WOTK_HIP(hipMemcpyAsync(w.n_src, nsrc.data(), sizeof(int) * n_pairs, hipMemcpyHostToDevice, s));
    WOTK_HIP(hipMemcpyAsync(w.lognf, lognf.data(), sizeof(double) * n_pairs, hipMemcpyHostToDevice, s));
    WOTK_HIP(hipMemcpyAsync(w.logng, logng.data(), sizeof(double) * n_pairs, hipMemcpyHostToDevice, s));

    // potentials start at 0 (u = v = 0, a = b = 1), padding columns at -inf
    batch_fill_kernel<<<1024, 256, 0, s>>>(w.f, nf * E, 0.0);
    batch_fill_kernel<<<1024, 256, 0, s>>>(w.g, ng * E, 0.0);
    WOTK_LAUNCH_CHECK();
    batch_pad_kernel<<<n_pairs, 64, 0, s>>>(w.f, w.pairs_f, n_pairs, E);
    batch_pad_kernel<<<n_pairs, 64, 0, s>>>(w.g, w.pairs_g, n_pairs, E);
    batch_fac_init_kernel<<<n_pairs, 256, 0, s>>>(w.ff, w.pairs_f, n_pairs, E);
    batch_fac_init_kernel<<<n_pairs, 256, 0, s>>>(w.fg, w.pairs_g, n_pairs, E);
    batch_shift_init_kernel<<<n_pairs, 256, 0, s>>>(w.psf, w.msf, w.pairs_f, n_pairs, E);
    batch_shift_init_kernel<<<n_pairs, 256, 0, s>>>(w.psg, w.msg, w.pairs_g, n_pairs, E);
    WOTK_LAUNCH_CHECK();
    batch_growth_kernel<<<dim3(n_pairs, E), 256, 0, s>>>(G, w.g_off, w.n_src, w.logw, w.lw_off, w.logq, E,
                                                         growth_per_combo ? 1 : 0);
    batch_reset_kernel<<<1, 1, 0, s>>>(w.iter, w.queues, w.exact_rows, 0);
    batch_zero_f32_kernel<<<1, 256, 0, s>>>(w.fmax_f, n_pairs * E);
    batch_zero_f32_kernel<<<1, 256, 0, s>>>(w.fmax_g, n_pairs * E);
    WOTK_LAUNCH_CHECK();
    WOTK_HIP(hipStreamSynchronize(s));  // host tables above are released at return

    const int64_t n_tf = (int64_t)tf.size(), n_tg = (int64_t)tg.size();  // <= ntf, ntg (one-row tiles)
    run->af = BatchArgs{cost, w.pairs_f, w.tiles_f, (int)n_tf, w.table, w.iter, E, n_pairs, 0, w.f, w.g, w.logw,
                        w.logq, w.lognf, w.queues, w.ff, w.fg, 0, w.exact_rows, w.bad_f, w.queues + 2, group,
                        w.fmax_f, w.fmax_g, w.fmax_g, n_pairs * E, w.psf, w.msf, w.psg, w.msg};
    run->ag = BatchArgs{cost, w.pairs_g, w.tiles_g, (int)n_tg, w.table, w.iter, E, n_pairs, 1, w.g, w.f, w.logw,
                        w.logq, w.logng, w.queues + 1, w.fg, w.ff, 1, w.exact_rows, w.bad_g, w.queues + 3, group,
                        w.fmax_g, w.fmax_f, w.fmax_f, n_pairs * E, w.psg, w.msg, w.psf, w.msf};
    const int grid = persistent_grid(*variant);
    run->grid_f = (int)std::min<int64_t>(grid, n_tf);
    run->grid_g = (int)std::min<int64_t>(grid, n_tg);
    run->sweep[0] = variant->sweep;
    for (int i = 0; i < kNumDerivedModes; ++i) run->sweep[1 + i] = variant->derived[i];
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    run->grid_r = ncu;  // repair blocks: at most this many queued tiles in flight
    run->E = E;
    run->n_pairs = n_pairs;
    run->group = group;
    run->n_iter = n_iter;
    run->graph_iters = graph_iters;
    if (graph_iters > 0 && n_iter >= graph_iters) {
        // capture graph_iters iterations of each sweep the schedule uses once on
        // a private stream, replay them
        WOTK_HIP(hipStreamCreateWithFlags(&run->cs, hipStreamNonBlocking));
        for (int ix = 0; ix <= kNumDerivedModes; ++ix) {
            if (std::find(run->mode_ix.begin(), run->mode_ix.end() - 1, (unsigned char)ix) == run->mode_ix.end() - 1)
                continue;
            WOTK_HIP(hipStreamBeginCapture(run->cs, hipStreamCaptureModeThreadLocal));
            run->capturing = true;
            for (int t = 0; t < graph_iters; ++t) {
                const int rc = run->sweep[ix](run->af, run->ag, run->grid_f, run->grid_g, run->grid_r, w.iter, w.queues,
                                              run->cs);
                if (rc) return rc;  // the destructor ends the capture and drops the partial graph
            }
            run->capturing = false;
            WOTK_HIP(hipStreamEndCapture(run->cs, &run->graph[ix]));
            WOTK_HIP(hipGraphInstantiate(&run->exec[ix], run->graph[ix], nullptr, nullptr, 0));
        }
    }
    *run_out = run.release();
    return WOTK_OK;
}

extern "C" int wotk_batch_advance(wotk_batch_run *run, int64_t n_iterations, int64_t *remaining, void *stream) {
    WOTK_REQUIRE(run != nullptr, "no batch run");
    WOTK_REQUIRE(n_iterations >= 0, "negative iteration count");
    hipStream_t s = (hipStream_t)stream;
    int64_t todo = std::min<int64_t>(n_iterations, run->n_iter - run->done);
    if (todo > 0) {
        hipEvent_t a = nullptr, b = nullptr;
        WOTK_HIP(hipEventCreate(&a));
        run->timing.emplace_back(a, nullptr);
        WOTK_HIP(hipEventCreate(&b));
        run->timing.back().second = b;
        WOTK_HIP(hipEventRecord(a, s));
        while (todo > 0) {
            // a graph replay where the next graph_iters iterations share one sweep
            const int ix = run->mode_ix[(size_t)run->done];
            int64_t len = 1;
            const int64_t cap = std::min<int64_t>(todo, run->graph_iters);
            while (len < cap && run->mode_ix[(size_t)(run->done + len)] == ix) ++len;
            if (run->exec[ix] && len == run->graph_iters) {
                WOTK_HIP(hipGraphLaunch(run->exec[ix], s));
                run->done += run->graph_iters;
                todo -= run->graph_iters;
            } else {
                const int rc = run->sweep[ix](run->af, run->ag, run->grid_f, run->grid_g, run->grid_r, run->w.iter,
                                              run->w.queues, s);
                if (rc) return rc;
                ++run->done;
                --todo;
            }
        }
        WOTK_HIP(hipEventRecord(b, s));
    }
    if (remaining) *remaining = run->n_iter - run->done;
    return WOTK_OK;
}

extern "C" int wotk_batch_potentials(wotk_batch_run *run, double *f_out, double *g_out, double *eps_final,
                                     wotk_stats *stats, void *stream) {
    WOTK_REQUIRE(run != nullptr, "no batch run");
    hipStream_t s = (hipStream_t)stream;
    const int E = run->E;
    // outputs: [pair][combo][n] contiguous blocks
    long long oa = 0, ob = 0;
    for (int p = 0; p < run->n_pairs; ++p) {
        const int64_t ni = run->n_i[p], nj = run->n_j[p];
        for (int e = 0; e < E; ++e) {
            if (f_out)
                WOTK_HIP(hipMemcpyAsync(f_out + oa + (long long)e * ni,
                                        run->w.f + run->pf[p].pot_off + (long long)e * run->pf[p].pot_ld,
                                        sizeof(double) * ni, hipMemcpyDeviceToDevice, s));
            if (g_out)
                WOTK_HIP(hipMemcpyAsync(g_out + ob + (long long)e * nj,
                                        run->w.g + run->pg[p].pot_off + (long long)e * run->pg[p].pot_ld,
                                        sizeof(double) * nj, hipMemcpyDeviceToDevice, s));
        }
        oa += (long long)E * ni;
        ob += (long long)E * nj;
    }
    if (eps_final) std::copy(run->eps_final.begin(), run->eps_final.end(), eps_final);
    if (stats) {
        unsigned long long exact_rows = 0;
        WOTK_HIP(hipMemcpyAsync(&exact_rows, run->w.exact_rows, sizeof(exact_rows), hipMemcpyDeviceToHost, s));
        WOTK_HIP(hipStreamSynchronize(s));
        double ms = 0.0;
        for (auto &ev : run->timing) {
            float m = 0.f;
            WOTK_HIP(hipEventElapsedTime(&m, ev.first, ev.second));
            ms += m;
        }
        *stats = wotk_stats{};
        stats->iterations = run->done;
        stats->sweeps = 2 * run->done;
        stats->sweep_ms = ms;
        stats->epsilon_final = run->eps_final[0];
        stats->stabilisations = (int64_t)exact_rows;
        stats->gap_checks = run->group;
    }
    return WOTK_OK;
}

extern "C" int wotk_batch_modes(const wotk_batch_run *run, int32_t *modes, int64_t n) {
    WOTK_REQUIRE(run != nullptr, "no batch run");
    WOTK_REQUIRE(modes != nullptr && n >= run->n_iter, "modes needs %lld entries", (long long)run->n_iter);
    for (int64_t i = 0; i < run->n_iter; ++i) {
        const int ix = run->mode_ix[(size_t)i];
        modes[i] = ix ? kDerivedModes[ix - 1] : 0;
    }
    return WOTK_OK;
}

extern "C" void wotk_batch_end(wotk_batch_run *run) { delete run; }

extern "C" int wotk_batch_stablev2(const float *cost, const wotk_batch_pair *pairs, int32_t n_pairs, const double *G,
                                   int32_t n_combos, const double *epsilon, const double *lambda1, double lambda2,
                                   int64_t scaling_iter, double tau, double epsilon0, int64_t extra_iter,
                                   int64_t inner_iter_max, double *f_out, double *g_out, int32_t graph_iters,
                                   int32_t flags, void *workspace, size_t workspace_bytes, wotk_stats *stats,
                                   void *stream) {
    (void)tau;  // absorptions are exact in the log domain (see sinkhorn.hip)
    WOTK_REQUIRE(n_pairs > 0 && n_combos >= 1 && n_combos <= kMaxCombos && epsilon && lambda1,
                 "empty batch or unsupported number of combinations");
    // the grid is the same for every pair
    std::vector<double> ep((size_t)n_pairs * n_combos), l1(ep.size()), l2(ep.size(), lambda2);
    for (int p = 0; p < n_pairs; ++p)
        for (int e = 0; e < n_combos; ++e) {
            ep[(size_t)p * n_combos + e] = epsilon[e];
            l1[(size_t)p * n_combos + e] = lambda1[e];
        }
    wotk_batch_run *raw = nullptr;
    int rc = wotk_batch_begin(cost, pairs, n_pairs, G, 0, n_combos, ep.data(), l1.data(), l2.data(), scaling_iter,
                              epsilon0, extra_iter, inner_iter_max, graph_iters, flags, workspace, workspace_bytes,
                              &raw, stream);
    if (rc) return rc;
    std::unique_ptr<wotk_batch_run> run(raw);
    rc = wotk_batch_advance(run.get(), scaling_iter + extra_iter, nullptr, stream);
    if (rc) return rc;
    return wotk_batch_potentials(run.get(), f_out, g_out, nullptr, stats, stream);
}
