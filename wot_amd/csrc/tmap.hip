// Propagation of cell populations through stored transport maps:
//   push forward  p <- p T        (wot/tmap/transport_map_model.py:284-293)
//   pull back     p <- (T p^T)^T  (wot/tmap/transport_map_model.py:350-359)
// with the reference's optional normalisation of every population to sum 1
// after each map.  Maps stay in HBM (FP64 or FP32, row-major, any leading
// dimension); populations travel between maps in a device buffer, so a chain
// of k maps costs k streaming passes over the maps and nothing else of size.
//
// Both directions are skinny GEMMs, HBM-bound for up to ~40 populations; the
// products run on the matrix cores in FP64 (v_mfma_f64_16x16x4f64, FP32 map
// entries are widened on load), 16 populations per MFMA tile.  Per lane:
//   A (16x4): row l%16, k l/16      B (4x16): row l/16, column l%16
//   D (16x16): rows l/16 + 4r (r = 0..3), column l%16 (measured: unlike the
//   FP32 16x16x4 form, whose rows are 4*(l/16) + r)
// Push: A = a 4-row strip of T read with the columns on the lanes (coalesced
// rows), B = the 4 matching population rows.  Pull: A = 16 rows of T, every
// lane streaming its own 32 contiguous bytes of its row; the reduction index
// k is assigned to lanes so that each lane's bytes are contiguous, and B reads
// the population rows of the same columns.  Long reductions are split over
// workgroups into FP64 partials summed in a fixed order (deterministic).
#include <algorithm>
#include <vector>

#include "common.h"

namespace wotk {
namespace {

constexpr int kPopTile = 16;          // populations per MFMA tile
constexpr int kWaves = 4;             // waves per workgroup (no LDS, no barriers)
constexpr int kTargetWaves = 4096;    // ~4 waves per SIMD on 256 CUs
constexpr int kPushBlocks = 4;        // 16-column MFMA blocks per push wave (4 or 8)
constexpr int kPushCols = 16 * kPushBlocks;
constexpr int kPushSteps = 4;         // MFMA k-steps (4 rows each) per push trip
constexpr bool kPushPipelined = false; // push trips double-buffered (measured slower: fewer waves fit)
constexpr int kPullRowBlocks = 4;     // 16-row MFMA blocks per pull wave
constexpr int kPullVecs = 2;          // 16-byte loads per lane, row and pull step
constexpr bool kPullPipelined = true; // pull steps double-buffered
constexpr int kReduceCells = 256 / 16; // cells per reduce workgroup (one thread per cell and population)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// B operand scale: the reference divides every population by its sum before
// the next map (p / sum); the products use p * (1 / sum), hoisted per lane
__device__ inline double pop_scale(const double *__restrict__ scale, int group, int pop) {
    return scale ? 1.0 / scale[group * kPopTile + pop] : 1.0;
}

template <typename T> struct Vec16;  // 16 bytes of map entries
template <> struct Vec16<double> { typedef d2 type; };
template <> struct Vec16<float> { typedef f4 type; };

// Map entries are read once per walk: non-temporal loads keep them from
// evicting the population vectors (re-read by every workgroup) from L2.
template <typename V>
__device__ inline V stream_load(const V *p) {
    return __builtin_nontemporal_load(p);
}

// ---------------------------------------------------------------------------
// push: part[split][group][j][16] = sum_{i in split} T[i][j] * p[group][i][16]
// One wave owns kPushCols = 16 kPushBlocks columns of T (kPushBlocks MFMA
// blocks).  The lane reads 16-byte vectors of its row: FP64 vector v holds
// columns j0 + 32 v + 2m, +1; FP32 vector v columns j0 + 64 v + 4m .. +3.
// push_col gives the column of MFMA block b at A-row m.  4 kPushSteps rows per
// trip.  FULL: the wave's columns are inside the map and rows are 16-byte
// aligned (vector loads, no checks); otherwise element loads with column
// checks.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline int64_t push_col(int64_t j0, int b, int m) {
    if constexpr (sizeof(T) == 8) return j0 + 32 * (b >> 1) + 2 * m + (b & 1);
    else return j0 + 64 * (b >> 2) + 4 * m + (b & 3);
}

// one lane's kPushBlocks entries of one map row, as loaded
template <typename T>
struct PushRow {
    static constexpr int PER = 16 / sizeof(T);
    static constexpr int NV = kPushBlocks / PER;
    typename Vec16<T>::type v[NV];
    __device__ double get(int b) const { return (double)v[b / PER][b % PER]; }
};

template <typename T, bool FULL>
__device__ inline void push_fetch(PushRow<T> &o, const T *__restrict__ row, int64_t j0, int m, int64_t n_cols) {
    typedef typename Vec16<T>::type V;
    constexpr int PER = PushRow<T>::PER;
    constexpr int CW = 16 * PER;  // columns one vector load covers across the 16 lanes of a row
    if constexpr (FULL) {
#pragma unroll
        for (int v = 0; v < PushRow<T>::NV; ++v)
            o.v[v] = stream_load(reinterpret_cast<const V *>(row + j0 + CW * v + PER * m));
    } else {
#pragma unroll
        for (int b = 0; b < kPushBlocks; ++b) {
            const int64_t j = push_col<T>(j0, b, m);
            const T e = row[j < n_cols ? j : 0];
            o.v[b / PER][b % PER] = j < n_cols ? e : T(0);
        }
    }
}

// NGW population groups (16 populations each) per wave: one read of the map
// feeds NGW x kPushBlocks MFMA accumulators
template <typename T, int NGW>
struct PushTrip {
    PushRow<T> a[kPushSteps];
    double b[NGW][kPushSteps];
};

template <typename T, bool FULL, int NGW>
__device__ inline void push_trip_fetch(PushTrip<T, NGW> &t, const T *__restrict__ map, int64_t ld, int64_t i,
                                       const double *const P[NGW], int64_t j0, int m, int q, int64_t n_cols) {
#pragma unroll
    for (int u = 0; u < kPushSteps; ++u) {
        const int64_t r = i + 4 * u + q;
        push_fetch<T, FULL>(t.a[u], map + r * ld, j0, m, n_cols);
#pragma unroll
        for (int g = 0; g < NGW; ++g) t.b[g][u] = P[g][r * kPopTile + m];
    }
}

template <typename T, int NGW>
__device__ inline void push_trip_mfma(const PushTrip<T, NGW> &t, const double inv[NGW], d4 acc[NGW][kPushBlocks]) {
#pragma unroll
    for (int u = 0; u < kPushSteps; ++u)
#pragma unroll
        for (int g = 0; g < NGW; ++g) {
            const double bv = t.b[g][u] * inv[g];
#pragma unroll
            for (int b = 0; b < kPushBlocks; ++b) acc[g][b] = mfma_f64(t.a[u].get(b), bv, acc[g][b]);
        }
}

template <typename T, bool FULL, int NGW>
__device__ inline void push_tile(const T *__restrict__ map, int64_t ld, int64_t i_begin, int64_t i_end,
                                 int64_t n_cols, const double *const P[NGW], const double inv[NGW], int64_t j0, int m,
                                 int q, d4 acc[NGW][kPushBlocks]) {
    // full trips of 4 kPushSteps rows; pipelined: trip t+1 loads while trip t multiplies
    constexpr int TR = 4 * kPushSteps;
    const int64_t n_trips = (i_end - i_begin) / TR;
    PushTrip<T, NGW> A;
    if constexpr (kPushPipelined) {
        PushTrip<T, NGW> B;
        if (n_trips > 0) push_trip_fetch<T, FULL, NGW>(A, map, ld, i_begin, P, j0, m, q, n_cols);
        int64_t t = 0;
        for (; t + 2 <= n_trips; t += 2) {
            push_trip_fetch<T, FULL, NGW>(B, map, ld, i_begin + (t + 1) * TR, P, j0, m, q, n_cols);
            push_trip_mfma(A, inv, acc);
            if (t + 2 < n_trips)
                push_trip_fetch<T, FULL, NGW>(A, map, ld, i_begin + (t + 2) * TR, P, j0, m, q, n_cols);
            push_trip_mfma(B, inv, acc);
        }
        if (t < n_trips) push_trip_mfma(A, inv, acc);
    } else {
        for (int64_t t = 0; t < n_trips; ++t) {
            push_trip_fetch<T, FULL, NGW>(A, map, ld, i_begin + t * TR, P, j0, m, q, n_cols);
            push_trip_mfma(A, inv, acc);
        }
    }

    // remaining rows, 4 per MFMA k-step
    for (int64_t i = i_begin + n_trips * TR; i < i_end; i += 4) {
        const int64_t r = i + q;
        const bool ok = r < i_end;
        const int64_t rr = ok ? r : i_begin;
        PushRow<T> a;
        push_fetch<T, FULL>(a, map + rr * ld, j0, m, n_cols);
#pragma unroll
        for (int g = 0; g < NGW; ++g) {
            const double bv = ok ? P[g][rr * kPopTile + m] * inv[g] : 0.0;
#pragma unroll
            for (int b = 0; b < kPushBlocks; ++b) acc[g][b] = mfma_f64(ok ? a.get(b) : 0.0, bv, acc[g][b]);
        }
    }
}

template <typename T, bool VEC, int NGW>
__global__ __launch_bounds__(256) void push_partial_kernel(const T *__restrict__ map, int64_t ld, int64_t n_rows,
                                                           int64_t n_cols, const double *__restrict__ pops,
                                                           const double *__restrict__ scale,
                                                           double *__restrict__ part, int64_t split_rows,
                                                           int n_tiles, int n_groups) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (tile >= n_tiles) return;
    const int split = blockIdx.y, g0 = blockIdx.z * NGW;
    const int m = lane & 15, q = lane >> 4;
    const int64_t j0 = (int64_t)tile * kPushCols;
    const int64_t i_begin = (int64_t)split * split_rows;
    const int64_t i_end = min(n_rows, i_begin + split_rows);
    // a group past the last one multiplies group g0's values by 0 and is not written
    const double *P[NGW];
    double inv[NGW];
#pragma unroll
    for (int g = 0; g < NGW; ++g) {
        const bool valid = g0 + g < n_groups;
        P[g] = pops + (int64_t)(valid ? g0 + g : g0) * n_rows * kPopTile;
        inv[g] = valid ? pop_scale(scale, g0 + g, m) : 0.0;
    }

    d4 acc[NGW][kPushBlocks];
#pragma unroll
    for (int g = 0; g < NGW; ++g)
#pragma unroll
        for (int b = 0; b < kPushBlocks; ++b) acc[g][b] = d4{0.0, 0.0, 0.0, 0.0};
    if (VEC && j0 + kPushCols <= n_cols)
        push_tile<T, VEC, NGW>(map, ld, i_begin, i_end, n_cols, P, inv, j0, m, q, acc);
    else
        push_tile<T, false, NGW>(map, ld, i_begin, i_end, n_cols, P, inv, j0, m, q, acc);

#pragma unroll
    for (int g = 0; g < NGW; ++g) {
        if (g0 + g >= n_groups) break;
        double *out = part + ((int64_t)split * n_groups + g0 + g) * n_cols * kPopTile;
#pragma unroll
        for (int b = 0; b < kPushBlocks; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t j = push_col<T>(j0, b, q + 4 * r);
                if (j < n_cols) out[j * kPopTile + m] = acc[g][b][r];
            }
    }
}

// ---------------------------------------------------------------------------
// pull: part[split][group][i][16] = sum_{j in split} T[i][j] * p[group][j][16]
// One wave owns kPullRowBlocks x 16 rows.  Per step the lane reads E
// contiguous entries (kPullVecs x 16 bytes) of its row at columns
// j + E*q .. + E-1, so the 4 lanes of a row cover 4 E contiguous entries; MFMA s uses entry s, i.e.
// reduction index k = q <-> column j + E*q + s, and B holds population row
// j + E*q + s.  Full steps are double-buffered like the push trips; a partial
// last step (element loads with column checks) follows them.
// ---------------------------------------------------------------------------
// one lane's E = kPullVecs * 16 / sizeof(T) contiguous entries of one row
template <typename T>
struct PullRow {
    static constexpr int PER = 16 / sizeof(T);  // entries per vector
    static constexpr int E = kPullVecs * PER;
    typename Vec16<T>::type v[kPullVecs];
    __device__ double get(int s) const { return (double)v[s / PER][s % PER]; }
};

// VEC: 16-byte vector loads; CHECK: columns may reach past j_end
template <typename T, bool VEC, bool CHECK>
__device__ inline void pull_fetch(PullRow<T> &o, const T *__restrict__ row, int64_t jc, int64_t j_end) {
    typedef typename Vec16<T>::type V;
    constexpr int PER = PullRow<T>::PER;
    if constexpr (VEC && !CHECK) {
#pragma unroll
        for (int v = 0; v < kPullVecs; ++v) o.v[v] = stream_load(reinterpret_cast<const V *>(row + jc + v * PER));
    } else {
#pragma unroll
        for (int v = 0; v < kPullVecs; ++v)
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const int64_t j = jc + v * PER + e;
                if constexpr (CHECK) {
                    const T x = row[j < j_end ? j : 0];
                    o.v[v][e] = j < j_end ? x : T(0);
                } else {
                    o.v[v][e] = row[j];
                }
            }
    }
}

template <typename T>
struct PullStep {
    static constexpr int E = PullRow<T>::E;
    PullRow<T> a[kPullRowBlocks];
    double b[E];
};

template <typename T, bool VEC, bool CHECK>
__device__ inline void pull_step_fetch(PullStep<T> &st, const T *const rows[kPullRowBlocks],
                                       const bool row_ok[kPullRowBlocks], int64_t j, int64_t j_begin, int64_t j_end,
                                       const double *__restrict__ P, int m, int q) {
    constexpr int E = PullStep<T>::E;
    const int64_t jc = j + (int64_t)E * q;
#pragma unroll
    for (int rb = 0; rb < kPullRowBlocks; ++rb) {
        pull_fetch<T, VEC, CHECK>(st.a[rb], rows[rb], jc, j_end);
        if (!row_ok[rb]) st.a[rb] = PullRow<T>{};
    }
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int64_t c = jc + s;
        if constexpr (CHECK) {
            const bool ok = c < j_end;
            const double v = P[(ok ? c : j_begin) * kPopTile + m];
            st.b[s] = ok ? v : 0.0;
        } else {
            st.b[s] = P[c * kPopTile + m];
        }
    }
}

template <typename T>
__device__ inline void pull_step_mfma(const PullStep<T> &st, double inv, d4 acc[kPullRowBlocks]) {
#pragma unroll
    for (int s = 0; s < PullStep<T>::E; ++s) {
        const double bv = st.b[s] * inv;
#pragma unroll
        for (int rb = 0; rb < kPullRowBlocks; ++rb) acc[rb] = mfma_f64(st.a[rb].get(s), bv, acc[rb]);
    }
}

template <typename T, bool VEC>
__global__ __launch_bounds__(256) void pull_partial_kernel(const T *__restrict__ map, int64_t ld, int64_t n_rows,
                                                           int64_t n_cols, const double *__restrict__ pops,
                                                           const double *__restrict__ scale,
                                                           double *__restrict__ part, int64_t split_cols,
                                                           int n_tiles) {
    constexpr int KS = 4 * PullStep<T>::E;  // columns per step
    constexpr int RB = kPullRowBlocks;
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (tile >= n_tiles) return;
    const int split = blockIdx.y, group = blockIdx.z, n_groups = gridDim.z;
    const int m = lane & 15, q = lane >> 4;
    const int64_t i0 = (int64_t)tile * 16 * RB;
    const int64_t j_begin = (int64_t)split * split_cols;
    const int64_t j_end = min(n_cols, j_begin + split_cols);
    const double *P = pops + (int64_t)group * n_cols * kPopTile;
    const double inv = pop_scale(scale, group, m);

    const T *rows[RB];
    bool row_ok[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int64_t r = i0 + 16 * rb + m;
        row_ok[rb] = r < n_rows;
        rows[rb] = map + (row_ok[rb] ? r : 0) * ld;
    }
    d4 acc[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) acc[rb] = d4{0.0, 0.0, 0.0, 0.0};

    const int64_t n_full = (j_end - j_begin) / KS;
    PullStep<T> A;
    if constexpr (kPullPipelined) {
        PullStep<T> B;
        if (n_full > 0) pull_step_fetch<T, VEC, false>(A, rows, row_ok, j_begin, j_begin, j_end, P, m, q);
        int64_t t = 0;
        for (; t + 2 <= n_full; t += 2) {
            pull_step_fetch<T, VEC, false>(B, rows, row_ok, j_begin + (t + 1) * KS, j_begin, j_end, P, m, q);
            pull_step_mfma(A, inv, acc);
            if (t + 2 < n_full)
                pull_step_fetch<T, VEC, false>(A, rows, row_ok, j_begin + (t + 2) * KS, j_begin, j_end, P, m, q);
            pull_step_mfma(B, inv, acc);
        }
        if (t < n_full) pull_step_mfma(A, inv, acc);
    } else {
        for (int64_t t = 0; t < n_full; ++t) {
            pull_step_fetch<T, VEC, false>(A, rows, row_ok, j_begin + t * KS, j_begin, j_end, P, m, q);
            pull_step_mfma(A, inv, acc);
        }
    }
    if (j_begin + n_full * KS < j_end) {
        pull_step_fetch<T, VEC, true>(A, rows, row_ok, j_begin + n_full * KS, j_begin, j_end, P, m, q);
        pull_step_mfma(A, inv, acc);
    }

    double *out = part + ((int64_t)split * n_groups + group) * n_rows * kPopTile;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t i = i0 + 16 * rb + q + 4 * r;
            if (i < n_rows) out[i * kPopTile + m] = acc[rb][r];
        }
}

// ---------------------------------------------------------------------------
// glue kernels (all bounded, no inter-workgroup synchronisation)
// ---------------------------------------------------------------------------
// external cells x pops (ld) -> internal [group][cell][16], zero-padded
__global__ __launch_bounds__(256) void pops_import_kernel(const double *__restrict__ src, int64_t ld, int64_t n,
                                                          int k, double *__restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int group = blockIdx.y;
    if (t >= n * kPopTile) return;
    const int64_t cell = t / kPopTile;
    const int pop = group * kPopTile + (int)(t % kPopTile);
    dst[(int64_t)group * n * kPopTile + t] = pop < k ? src[cell * ld + pop] : 0.0;
}

// sum of the split partials in split order, one thread per (cell, population);
// per-workgroup partial sums of every population (16 cells) for the
// normalisation (fixed order: deterministic)
__global__ __launch_bounds__(256) void pops_reduce_kernel(const double *__restrict__ part, int n_split, int64_t n,
                                                          double *__restrict__ dst, double *__restrict__ block_sums) {
    __shared__ double red[256];
    const int group = blockIdx.y, n_groups = gridDim.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;  // cell * 16 + pop
    const int64_t stride = n * kPopTile;                        // one split's slice of a group
    double v = 0.0;
    if (t < stride) {
        const double *src = part + (int64_t)group * stride + t;
        const int64_t step = (int64_t)n_groups * stride;
        int s = 0;
        for (; s + 4 <= n_split; s += 4) {
            const double a = src[(int64_t)s * step], b = src[(int64_t)(s + 1) * step];
            const double c = src[(int64_t)(s + 2) * step], d = src[(int64_t)(s + 3) * step];
            v = (((v + a) + b) + c) + d;
        }
        for (; s < n_split; ++s) v += src[(int64_t)s * step];
        dst[(int64_t)group * stride + t] = v;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w >= kPopTile; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < kPopTile)
        block_sums[((int64_t)group * gridDim.x + blockIdx.x) * kPopTile + threadIdx.x] = red[threadIdx.x];
}

// totals[group*16 + pop] = sum of the block sums (padded populations: 1);
// 1024 threads, 4 independent partial sums each, combined in a fixed order
constexpr int kTotalsThreads = 1024;
__global__ __launch_bounds__(kTotalsThreads) void pops_totals_kernel(const double *__restrict__ block_sums,
                                                                     int n_blocks, int k,
                                                                     double *__restrict__ totals) {
    __shared__ double red[kTotalsThreads];
    const int group = blockIdx.x;
    const int pop = threadIdx.x % kPopTile;
    constexpr int lanes = kTotalsThreads / kPopTile;  // threads per population
    const double *src = block_sums + (int64_t)group * n_blocks * kPopTile + pop;
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    int b = threadIdx.x / kPopTile;
    for (; b + 3 * lanes < n_blocks; b += 4 * lanes)
#pragma unroll
        for (int u = 0; u < 4; ++u) t[u] += src[(int64_t)(b + u * lanes) * kPopTile];
    for (; b < n_blocks; b += lanes) t[0] += src[(int64_t)b * kPopTile];
    red[threadIdx.x] = (t[0] + t[1]) + (t[2] + t[3]);
    __syncthreads();
    for (int w = kTotalsThreads / 2; w >= kPopTile; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < kPopTile) {
        const int p = group * kPopTile + threadIdx.x;
        totals[p] = p < k ? red[threadIdx.x] : 1.0;
    }
}

// internal -> external cells x pops (ld), divided by the totals when given
__global__ __launch_bounds__(256) void pops_export_kernel(const double *__restrict__ src, int64_t n, int k,
                                                          const double *__restrict__ totals, double *__restrict__ dst,
                                                          int64_t ld) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int group = blockIdx.y;
    if (t >= n * kPopTile) return;
    const int64_t cell = t / kPopTile;
    const int pop = group * kPopTile + (int)(t % kPopTile);
    if (pop >= k) return;
    const double v = src[(int64_t)group * n * kPopTile + t];
    dst[cell * ld + pop] = totals ? v / totals[pop] : v;
}

// dst[j][i] = src[i][j] through a padded 32 x 32 LDS tile (reads and writes coalesced)
template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(const T *__restrict__ src, int64_t lds, int64_t rows,
                                                        int64_t cols, T *__restrict__ dst, int64_t ldd) {
    __shared__ T tile[32][33];
    const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int64_t r = r0 + ty + k, c = c0 + tx;
        if (r < rows && c < cols) tile[ty + k][tx] = src[r * lds + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int64_t c = c0 + ty + k, r = r0 + tx;  // dst row c, column r
        if (c < cols && r < rows) dst[c * ldd + r] = tile[tx][ty + k];
    }
}

// X[i, :k] /= sum(X[i, :k])   (fates: X /= X.sum(axis=1), transport_map_model.py:66)
__global__ __launch_bounds__(256) void rows_normalize_kernel(double *__restrict__ X, int64_t n, int k, int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *row = X + i * ld;
    double s = 0.0;
    for (int c = 0; c < k; ++c) s += row[c];
    for (int c = 0; c < k; ++c) row[c] = row[c] / s;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct StepShape {
    int64_t n_in, n_out;   // cells before / after the map
    int64_t reduce_len;    // length of the reduction (push: rows, pull: cols)
    int n_tiles, n_split;
    int64_t split_len;
};

inline int pull_step_cols(int dtype) { return 4 * kPullVecs * (dtype == WOTK_DTYPE_F64 ? 2 : 4); }

// a pull back through a map with a transposed copy is a push forward through the copy
wotk_tmap effective_map(const wotk_tmap &m, int direction, int *eff_direction) {
    if (direction == WOTK_PULL_BACK && m.transposed) {
        wotk_tmap t = m;
        t.data = m.transposed;
        t.ld = m.ld_t;
        t.rows = m.cols;
        t.cols = m.rows;
        t.transposed = nullptr;
        *eff_direction = WOTK_PUSH_FORWARD;
        return t;
    }
    *eff_direction = direction;
    return m;
}

StepShape step_shape(const wotk_tmap &m, int direction, int n_groups) {
    StepShape s;
    const bool push = direction == WOTK_PUSH_FORWARD;
    s.n_in = push ? m.rows : m.cols;
    s.n_out = push ? m.cols : m.rows;
    s.reduce_len = s.n_in;
    const int64_t tiles = push ? (m.cols + kPushCols - 1) / kPushCols
                               : (m.rows + 16 * kPullRowBlocks - 1) / (16 * kPullRowBlocks);
    s.n_tiles = (int)tiles;
    const int64_t quantum = push ? 16 : pull_step_cols(m.dtype);
    // enough splits to put ~kTargetWaves waves in flight, each reduction chunk >= 8 quanta
    int64_t want = (kTargetWaves + tiles * n_groups - 1) / (tiles * n_groups);
    int64_t most = std::max<int64_t>(1, s.reduce_len / (8 * quantum));
    int64_t splits = std::max<int64_t>(1, std::min(want, most));
    s.split_len = round_up((s.reduce_len + splits - 1) / splits, quantum);
    if (s.split_len <= 0) s.split_len = quantum;
    s.n_split = (int)std::max<int64_t>(1, (s.reduce_len + s.split_len - 1) / s.split_len);
    return s;
}

struct Layout {
    double *part = nullptr, *pop_a = nullptr, *pop_b = nullptr, *block_sums = nullptr, *totals = nullptr;
};

int check_chain(const wotk_tmap *maps, int32_t n_maps, int32_t direction, int32_t n_pops) {
    WOTK_REQUIRE(maps && n_maps > 0, "tmap: need at least one map");
    WOTK_REQUIRE(direction == WOTK_PUSH_FORWARD || direction == WOTK_PULL_BACK, "tmap: bad direction %d",
                 direction);
    WOTK_REQUIRE(n_pops > 0, "tmap: need at least one population");
    for (int s = 0; s < n_maps; ++s) {
        const wotk_tmap &m = maps[s];
        WOTK_REQUIRE(m.data && m.rows > 0 && m.cols > 0 && m.ld >= m.cols, "tmap: map %d has a bad shape", s);
        WOTK_REQUIRE(m.dtype == WOTK_DTYPE_F64 || m.dtype == WOTK_DTYPE_F32, "tmap: map %d has dtype %d", s,
                     m.dtype);
        WOTK_REQUIRE(!m.transposed || m.ld_t >= m.rows, "tmap: map %d has a transposed copy with ld %lld < %lld",
                     s, (long long)m.ld_t, (long long)m.rows);
        if (s > 0) {
            const wotk_tmap &p = maps[s - 1];
            const int64_t prev_out = direction == WOTK_PUSH_FORWARD ? p.cols : p.rows;
            const int64_t here_in = direction == WOTK_PUSH_FORWARD ? m.rows : m.cols;
            WOTK_REQUIRE(prev_out == here_in, "tmap: map %d takes %lld cells but map %d yields %lld", s,
                         (long long)here_in, s - 1, (long long)prev_out);
        }
    }
    return WOTK_OK;
}

Layout carve(Arena &ar, const wotk_tmap *maps, int32_t n_maps, int32_t direction, int32_t n_pops) {
    const int n_groups = (n_pops + kPopTile - 1) / kPopTile;
    int64_t max_cells = 0, max_part = 0, max_blocks = 0;
    for (int s = 0; s < n_maps; ++s) {
        int dir;
        const wotk_tmap m = effective_map(maps[s], direction, &dir);
        StepShape sh = step_shape(m, dir, n_groups);
        max_cells = std::max({max_cells, sh.n_in, sh.n_out});
        max_part = std::max(max_part, (int64_t)sh.n_split * sh.n_out);
        max_blocks = std::max(max_blocks, (sh.n_out + kReduceCells - 1) / kReduceCells);
    }
    Layout L;
    L.part = ar.take<double>((size_t)max_part * n_groups * kPopTile);
    L.pop_a = ar.take<double>((size_t)max_cells * n_groups * kPopTile);
    L.pop_b = ar.take<double>((size_t)max_cells * n_groups * kPopTile);
    L.block_sums = ar.take<double>((size_t)max_blocks * n_groups * kPopTile);
    L.totals = ar.take<double>((size_t)2 * n_groups * kPopTile);
    return L;
}

template <typename T>
bool vec_ok(const wotk_tmap &m) {
    return (reinterpret_cast<uintptr_t>(m.data) % 16 == 0) && ((m.ld * (int64_t)sizeof(T)) % 16 == 0);
}

template <typename T>
void launch_partial(const wotk_tmap &m, int direction, const StepShape &sh, int n_groups, const double *pops,
                    const double *scale, double *part, hipStream_t st) {
    const T *map = static_cast<const T *>(m.data);
    dim3 grid((sh.n_tiles + kWaves - 1) / kWaves, sh.n_split, n_groups);
    const bool vec = vec_ok<T>(m);
    if (direction == WOTK_PUSH_FORWARD) {
        // two population groups per wave when there are several: one read of the map for both
        if (n_groups > 1) {
            grid.z = (n_groups + 1) / 2;
            if (vec)
                push_partial_kernel<T, true, 2><<<grid, 256, 0, st>>>(map, m.ld, m.rows, m.cols, pops, scale, part,
                                                                     sh.split_len, sh.n_tiles, n_groups);
            else
                push_partial_kernel<T, false, 2><<<grid, 256, 0, st>>>(map, m.ld, m.rows, m.cols, pops, scale, part,
                                                                      sh.split_len, sh.n_tiles, n_groups);
        } else if (vec) {
            push_partial_kernel<T, true, 1><<<grid, 256, 0, st>>>(map, m.ld, m.rows, m.cols, pops, scale, part,
                                                                 sh.split_len, sh.n_tiles, n_groups);
        } else {
            push_partial_kernel<T, false, 1><<<grid, 256, 0, st>>>(map, m.ld, m.rows, m.cols, pops, scale, part,
                                                                  sh.split_len, sh.n_tiles, n_groups);
        }
    } else {
        if (vec)
            pull_partial_kernel<T, true><<<grid, 256, 0, st>>>(map, m.ld, m.rows, m.cols, pops, scale, part,
                                                              sh.split_len, sh.n_tiles);
        else
            pull_partial_kernel<T, false><<<grid, 256, 0, st>>>(map, m.ld, m.rows, m.cols, pops, scale, part,
                                                               sh.split_len, sh.n_tiles);
    }
}

}  // namespace
}  // namespace wotk

using namespace wotk;

extern "C" size_t wotk_tmap_workspace_size(const wotk_tmap *maps, int32_t n_maps, int32_t direction,
                                           int32_t n_pops) {
    if (check_chain(maps, n_maps, direction, n_pops) != WOTK_OK) return 0;
    Arena ar(nullptr, 0);
    carve(ar, maps, n_maps, direction, n_pops);
    return ar.off;
}

extern "C" int wotk_tmap_propagate(const wotk_tmap *maps, int32_t n_maps, int32_t direction, const double *pops,
                                   int64_t ldp, int32_t n_pops, int32_t normalize, double *const *outs,
                                   int64_t ldo, void *workspace, size_t workspace_bytes, wotk_stats *stats,
                                   void *stream) {
    int rc = check_chain(maps, n_maps, direction, n_pops);
    if (rc != WOTK_OK) return rc;
    WOTK_REQUIRE(pops && ldp >= n_pops, "tmap: bad population matrix (ld %lld for %d populations)",
                 (long long)ldp, n_pops);
    WOTK_REQUIRE(outs && ldo >= n_pops, "tmap: bad output (ld %lld for %d populations)", (long long)ldo, n_pops);
    WOTK_REQUIRE(outs[n_maps - 1], "tmap: the last step's output pointer is NULL");
    Arena ar(workspace, workspace_bytes);
    Layout L = carve(ar, maps, n_maps, direction, n_pops);
    WOTK_REQUIRE(ar.ok && workspace, "tmap: workspace too small (%zu < %zu bytes)", workspace_bytes, ar.off);
    hipStream_t st = (hipStream_t)stream;
    const int n_groups = (n_pops + kPopTile - 1) / kPopTile;

    const int64_t n0 = direction == WOTK_PUSH_FORWARD ? maps[0].rows : maps[0].cols;
    {
        dim3 grid((unsigned)((n0 * kPopTile + 255) / 256), n_groups);
        pops_import_kernel<<<grid, 256, 0, st>>>(pops, ldp, n0, n_pops, L.pop_a);
        WOTK_LAUNCH_CHECK();
    }
    // optional timer: one event pair around every map's product kernel
    std::vector<hipEvent_t> ev;
    if (stats) {
        ev.resize(2 * (size_t)n_maps);
        for (auto &e : ev) WOTK_HIP(hipEventCreate(&e));
    }
    double *cur = L.pop_a, *nxt = L.pop_b;
    const double *scale = nullptr;  // totals of the previous map (normalising chains)
    for (int s = 0; s < n_maps; ++s) {
        int dir;
        const wotk_tmap m = effective_map(maps[s], direction, &dir);
        StepShape sh = step_shape(m, dir, n_groups);
        if (stats) WOTK_HIP(hipEventRecord(ev[2 * s], st));
        if (m.dtype == WOTK_DTYPE_F64)
            launch_partial<double>(m, dir, sh, n_groups, cur, scale, L.part, st);
        else
            launch_partial<float>(m, dir, sh, n_groups, cur, scale, L.part, st);
        WOTK_LAUNCH_CHECK();
        if (stats) WOTK_HIP(hipEventRecord(ev[2 * s + 1], st));
        const int n_blocks = (int)((sh.n_out + kReduceCells - 1) / kReduceCells);
        pops_reduce_kernel<<<dim3(n_blocks, n_groups), 256, 0, st>>>(L.part, sh.n_split, sh.n_out, nxt,
                                                                    L.block_sums);
        WOTK_LAUNCH_CHECK();
        double *tot = nullptr;
        if (normalize) {
            tot = L.totals + (s & 1) * n_groups * kPopTile;
            pops_totals_kernel<<<n_groups, kTotalsThreads, 0, st>>>(L.block_sums, n_blocks, n_pops, tot);
            WOTK_LAUNCH_CHECK();
        }
        if (outs[s]) {
            dim3 grid((unsigned)((sh.n_out * kPopTile + 255) / 256), n_groups);
            pops_export_kernel<<<grid, 256, 0, st>>>(nxt, sh.n_out, n_pops, tot, outs[s], ldo);
            WOTK_LAUNCH_CHECK();
        }
        scale = tot;
        std::swap(cur, nxt);
    }
    if (stats) {
        *stats = wotk_stats{};
        WOTK_HIP(hipStreamSynchronize(st));
        double total = 0.0;
        for (int s = 0; s < n_maps; ++s) {
            float ms = 0.f;
            WOTK_HIP(hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]));
            total += ms;
        }
        for (auto &e : ev) WOTK_HIP(hipEventDestroy(e));
        stats->iterations = n_maps;
        stats->sweeps = n_maps;
        stats->sweep_ms = total;
    }
    return WOTK_OK;
}

extern "C" int wotk_tmap_transpose(const void *src, int32_t dtype, int64_t ld_src, int64_t rows, int64_t cols,
                                   void *dst, int64_t ld_dst, void *stream) {
    WOTK_REQUIRE(src && dst && rows > 0 && cols > 0 && ld_src >= cols && ld_dst >= rows,
                 "tmap_transpose: bad arguments");
    WOTK_REQUIRE(dtype == WOTK_DTYPE_F64 || dtype == WOTK_DTYPE_F32, "tmap_transpose: bad dtype %d", dtype);
    WOTK_REQUIRE((cols + 31) / 32 < (1ll << 31) && (rows + 31) / 32 < 65536, "tmap_transpose: map too large");
    dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
    hipStream_t st = (hipStream_t)stream;
    if (dtype == WOTK_DTYPE_F64)
        transpose_kernel<double><<<grid, 256, 0, st>>>(static_cast<const double *>(src), ld_src, rows, cols,
                                                      static_cast<double *>(dst), ld_dst);
    else
        transpose_kernel<float><<<grid, 256, 0, st>>>(static_cast<const float *>(src), ld_src, rows, cols,
                                                     static_cast<float *>(dst), ld_dst);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}

extern "C" int wotk_rows_normalize_f64(double *X, int64_t n, int32_t k, int64_t ld, void *stream) {
    WOTK_REQUIRE(X && n >= 0 && k > 0 && ld >= k, "rows_normalize: bad arguments");
    if (n == 0) return WOTK_OK;
    rows_normalize_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(X, n, k, ld);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}
