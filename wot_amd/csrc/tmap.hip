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
// lane streaming its own 64 contiguous bytes of its row; the reduction index
// k is assigned to lanes so that each lane's bytes are contiguous, and B reads
// the population rows of the same columns.  Long reductions are split over
// workgroups into FP64 partials summed in a fixed order (deterministic).
#include <algorithm>

#include "common.h"

namespace wotk {
namespace {

constexpr int kPopTile = 16;          // populations per MFMA tile
constexpr int kWaves = 4;             // waves per workgroup (no LDS, no barriers)
constexpr int kTargetWaves = 8192;    // ~8 waves per SIMD on 256 CUs
constexpr int kPullRowBlocks = 4;     // 16-row MFMA blocks per pull wave
constexpr int kReduceCells = 256;     // cells per reduce workgroup

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// B operand: population value, divided by the previous map's total when the
// chain normalises (the reference divides p by its sum before the next map)
__device__ inline double pop_value(const double *__restrict__ P, int64_t idx, const double *__restrict__ scale,
                                   int pop) {
    double v = P[idx];
    return scale ? v / scale[pop] : v;
}

// ---------------------------------------------------------------------------
// push: part[split][group][j][16] = sum_{i in split} T[i][j] * p[group][i][16]
// One wave owns 64 columns of T.  FP64: the lane reads two double2 (columns
// j0+2m.. and j0+32+2m..); FP32: one float4 (columns j0+4m..).  colmap gives
// the column of MFMA block b at A-row m.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline int64_t push_col(int64_t j0, int b, int m) {
    if constexpr (sizeof(T) == 8) return j0 + 2 * m + (b & 1) + 32 * (b >> 1);
    else return j0 + 4 * m + b;
}

template <typename T, bool VEC>
__device__ inline void push_load(const T *__restrict__ row, int64_t j0, int m, bool full, int64_t n_cols,
                                 double a[4]) {
    if (VEC && full) {
        if constexpr (sizeof(T) == 8) {
            d2 x = *reinterpret_cast<const d2 *>(row + j0 + 2 * m);
            d2 y = *reinterpret_cast<const d2 *>(row + j0 + 32 + 2 * m);
            a[0] = x[0]; a[1] = x[1]; a[2] = y[0]; a[3] = y[1];
        } else {
            f4 x = *reinterpret_cast<const f4 *>(row + j0 + 4 * m);
            a[0] = x[0]; a[1] = x[1]; a[2] = x[2]; a[3] = x[3];
        }
    } else {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            int64_t j = push_col<T>(j0, b, m);
            double v = (double)row[j < n_cols ? j : 0];
            a[b] = j < n_cols ? v : 0.0;
        }
    }
}

template <typename T, bool VEC>
__global__ __launch_bounds__(256) void push_partial_kernel(const T *__restrict__ map, int64_t ld, int64_t n_rows,
                                                           int64_t n_cols, const double *__restrict__ pops,
                                                           const double *__restrict__ scale,
                                                           double *__restrict__ part, int64_t split_rows,
                                                           int n_tiles) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (tile >= n_tiles) return;
    const int split = blockIdx.y, group = blockIdx.z, n_groups = gridDim.z;
    const int m = lane & 15, q = lane >> 4;
    const int64_t j0 = (int64_t)tile * 64;
    const int64_t i_begin = (int64_t)split * split_rows;
    const int64_t i_end = min(n_rows, i_begin + split_rows);
    const bool full = j0 + 64 <= n_cols;
    const double *P = pops + (int64_t)group * n_rows * kPopTile;
    const double *sc = scale ? scale + group * kPopTile : nullptr;

    d4 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[b] = d4{0.0, 0.0, 0.0, 0.0};

    int64_t i = i_begin;
    // main loop: 16 rows (4 MFMA k-steps) per trip, loads issued before the MFMAs
    for (; i + 16 <= i_end; i += 16) {
        double a[4][4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t r = i + 4 * u + q;
            push_load<T, VEC>(map + r * ld, j0, m, full, n_cols, a[u]);
            bv[u] = pop_value(P, r * kPopTile + m, sc, m);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[b] = mfma_f64(a[u][b], bv[u], acc[b]);
    }
    for (; i < i_end; i += 4) {
        const int64_t r = i + q;
        const bool ok = r < i_end;
        const int64_t rr = ok ? r : i_begin;
        double a[4];
        push_load<T, VEC>(map + rr * ld, j0, m, full, n_cols, a);
        const double bv = ok ? pop_value(P, rr * kPopTile + m, sc, m) : 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[b] = mfma_f64(ok ? a[b] : 0.0, bv, acc[b]);
    }

    double *out = part + ((int64_t)split * n_groups + group) * n_cols * kPopTile;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t j = push_col<T>(j0, b, q + 4 * r);
            if (j < n_cols) out[j * kPopTile + m] = acc[b][r];
        }
}

// ---------------------------------------------------------------------------
// pull: part[split][group][i][16] = sum_{j in split} T[i][j] * p[group][j][16]
// One wave owns kPullRowBlocks x 16 rows.  Per step the lane reads E
// contiguous entries (64 bytes) of its row at columns j + E*q .. + E-1; MFMA
// s uses entry s, i.e. reduction index k = q <-> column j + E*q + s, and B
// holds population row j + E*q + s.
// ---------------------------------------------------------------------------
template <typename T, bool VEC>
__device__ inline void pull_load(const T *__restrict__ row, int64_t jc, bool full, int64_t j_end,
                                 double *a /* E */) {
    constexpr int E = 64 / sizeof(T);
    if (VEC && full) {
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int v = 0; v < E / 2; ++v) {
                d2 x = *reinterpret_cast<const d2 *>(row + jc + 2 * v);
                a[2 * v] = x[0];
                a[2 * v + 1] = x[1];
            }
        } else {
#pragma unroll
            for (int v = 0; v < E / 4; ++v) {
                f4 x = *reinterpret_cast<const f4 *>(row + jc + 4 * v);
                a[4 * v] = x[0]; a[4 * v + 1] = x[1]; a[4 * v + 2] = x[2]; a[4 * v + 3] = x[3];
            }
        }
    } else {
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const int64_t j = jc + s;
            const double v = (double)row[j < j_end ? j : 0];
            a[s] = j < j_end ? v : 0.0;
        }
    }
}

template <typename T, bool VEC>
__global__ __launch_bounds__(256) void pull_partial_kernel(const T *__restrict__ map, int64_t ld, int64_t n_rows,
                                                           int64_t n_cols, const double *__restrict__ pops,
                                                           const double *__restrict__ scale,
                                                           double *__restrict__ part, int64_t split_cols,
                                                           int n_tiles) {
    constexpr int E = 64 / sizeof(T);  // entries per lane per step
    constexpr int KS = 4 * E;          // columns per step
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
    const double *sc = scale ? scale + group * kPopTile : nullptr;

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

    for (int64_t j = j_begin; j < j_end; j += KS) {
        const bool full = j + KS <= j_end;
        const int64_t jc = j + (int64_t)E * q;
        double a[RB][E], bv[E];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            pull_load<T, VEC>(rows[rb], jc, full, j_end, a[rb]);
            if (!row_ok[rb])
#pragma unroll
                for (int s = 0; s < E; ++s) a[rb][s] = 0.0;
        }
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const int64_t c = jc + s;
            const bool ok = c < j_end;
            const double v = pop_value(P, (ok ? c : j_begin) * kPopTile + m, sc, m);
            bv[s] = ok ? v : 0.0;
        }
#pragma unroll
        for (int s = 0; s < E; ++s)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) acc[rb] = mfma_f64(a[rb][s], bv[s], acc[rb]);
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

// sum of the split partials in split order; per-workgroup partial sums of
// every population for the normalisation (fixed order: deterministic)
__global__ __launch_bounds__(256) void pops_reduce_kernel(const double *__restrict__ part, int n_split, int64_t n,
                                                          double *__restrict__ dst, double *__restrict__ block_sums) {
    __shared__ double red[256];
    const int group = blockIdx.y, n_groups = gridDim.y;
    const int pop = threadIdx.x % kPopTile;
    const int64_t cell0 = (int64_t)blockIdx.x * kReduceCells;
    double tot = 0.0;
    for (int c = threadIdx.x / kPopTile; c < kReduceCells; c += 256 / kPopTile) {
        const int64_t cell = cell0 + c;
        if (cell >= n) break;
        double v = 0.0;
        for (int s = 0; s < n_split; ++s) v += part[(((int64_t)s * n_groups + group) * n + cell) * kPopTile + pop];
        dst[((int64_t)group * n + cell) * kPopTile + pop] = v;
        tot += v;
    }
    red[threadIdx.x] = tot;
    __syncthreads();
    for (int w = 128; w >= kPopTile; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < kPopTile)
        block_sums[((int64_t)group * gridDim.x + blockIdx.x) * kPopTile + threadIdx.x] = red[threadIdx.x];
}

// totals[group*16 + pop] = sum of the block sums (padded populations: 1)
__global__ __launch_bounds__(256) void pops_totals_kernel(const double *__restrict__ block_sums, int n_blocks, int k,
                                                          double *__restrict__ totals) {
    __shared__ double red[256];
    const int group = blockIdx.x;
    const int pop = threadIdx.x % kPopTile;
    double t = 0.0;
    for (int b = threadIdx.x / kPopTile; b < n_blocks; b += 256 / kPopTile)
        t += block_sums[((int64_t)group * n_blocks + b) * kPopTile + pop];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w >= kPopTile; w >>= 1) {
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

inline int pull_step_cols(int dtype) { return dtype == WOTK_DTYPE_F64 ? 32 : 64; }

StepShape step_shape(const wotk_tmap &m, int direction, int n_groups) {
    StepShape s;
    const bool push = direction == WOTK_PUSH_FORWARD;
    s.n_in = push ? m.rows : m.cols;
    s.n_out = push ? m.cols : m.rows;
    s.reduce_len = s.n_in;
    const int64_t tiles = push ? (m.cols + 63) / 64 : (m.rows + 16 * kPullRowBlocks - 1) / (16 * kPullRowBlocks);
    s.n_tiles = (int)tiles;
    const int64_t quantum = push ? 16 : pull_step_cols(m.dtype);
    // enough splits to put ~kTargetWaves waves in flight, each reduction chunk >= 4 quanta
    int64_t want = (kTargetWaves + tiles * n_groups - 1) / (tiles * n_groups);
    int64_t most = std::max<int64_t>(1, s.reduce_len / (4 * quantum));
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
        StepShape sh = step_shape(maps[s], direction, n_groups);
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
        if (vec)
            push_partial_kernel<T, true><<<grid, 256, 0, st>>>(map, m.ld, m.rows, m.cols, pops, scale, part,
                                                              sh.split_len, sh.n_tiles);
        else
            push_partial_kernel<T, false><<<grid, 256, 0, st>>>(map, m.ld, m.rows, m.cols, pops, scale, part,
                                                               sh.split_len, sh.n_tiles);
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
                                   int64_t ldo, void *workspace, size_t workspace_bytes, void *stream) {
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
    double *cur = L.pop_a, *nxt = L.pop_b;
    const double *scale = nullptr;  // totals of the previous map (normalising chains)
    for (int s = 0; s < n_maps; ++s) {
        const wotk_tmap &m = maps[s];
        StepShape sh = step_shape(m, direction, n_groups);
        if (m.dtype == WOTK_DTYPE_F64)
            launch_partial<double>(m, direction, sh, n_groups, cur, scale, L.part, st);
        else
            launch_partial<float>(m, direction, sh, n_groups, cur, scale, L.part, st);
        WOTK_LAUNCH_CHECK();
        const int n_blocks = (int)((sh.n_out + kReduceCells - 1) / kReduceCells);
        pops_reduce_kernel<<<dim3(n_blocks, n_groups), 256, 0, st>>>(L.part, sh.n_split, sh.n_out, nxt,
                                                                    L.block_sums);
        WOTK_LAUNCH_CHECK();
        double *tot = nullptr;
        if (normalize) {
            tot = L.totals + (s & 1) * n_groups * kPopTile;
            pops_totals_kernel<<<n_groups, 256, 0, st>>>(L.block_sums, n_blocks, n_pops, tot);
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
    return WOTK_OK;
}

extern "C" int wotk_rows_normalize_f64(double *X, int64_t n, int32_t k, int64_t ld, void *stream) {
    WOTK_REQUIRE(X && n >= 0 && k > 0 && ld >= k, "rows_normalize: bad arguments");
    if (n == 0) return WOTK_OK;
    rows_normalize_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(X, n, k, ld);
    WOTK_LAUNCH_CHECK();
    return WOTK_OK;
}
