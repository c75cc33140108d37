// Per-shape-bucket tile table of the sweep kernels (written by the autotuner,
// read at every launch).
#include <atomic>

#include "common.h"

namespace {

constexpr int kRowBuckets = 32 * 32;   // floor(log2 rows) x floor(log2 cols)
constexpr int kBatchBuckets = 16 * 32; // combinations 1..16 x floor(log2 mean row length)

// variant index per bucket, -1 = default (variant 0)
std::atomic<int> g_row[kRowBuckets];
std::atomic<int> g_batch[kBatchBuckets];
std::atomic<int> g_shared[kBatchBuckets];

struct TableInit {
    TableInit() {
        for (auto &v : g_row) v.store(-1);
        for (auto &v : g_batch) v.store(-1);
        for (auto &v : g_shared) v.store(-1);
    }
} g_init;

int lg(int64_t n) {
    int b = 0;
    while (n > 1 && b < 31) {
        n >>= 1;
        ++b;
    }
    return b;
}

std::atomic<int> *table(int kernel, int *n_buckets, const wotk::TileShape **tiles, int *n_tiles) {
    if (kernel == WOTK_KERNEL_ROW_SWEEP) {
        *n_buckets = kRowBuckets;
        *tiles = wotk::kRowSweepTiles;
        *n_tiles = wotk::kNumRowSweepTiles;
        return g_row;
    }
    if (kernel == WOTK_KERNEL_BATCH_SWEEP) {
        *n_buckets = kBatchBuckets;
        *tiles = wotk::kBatchSweepTiles;
        *n_tiles = wotk::kNumBatchSweepTiles;
        return g_batch;
    }
    if (kernel == WOTK_KERNEL_BATCH_SHARED) {
        *n_buckets = kBatchBuckets;
        *tiles = wotk::kBatchSharedTiles;
        *n_tiles = wotk::kNumBatchSharedTiles;
        return g_shared;
    }
    return nullptr;
}

}  // namespace

extern "C" int32_t wotk_shape_bucket(int32_t kernel, int64_t n_rows, int64_t n_cols, int32_t n_combos) {
    if (kernel == WOTK_KERNEL_ROW_SWEEP) return lg(n_rows) * 32 + lg(n_cols);
    if (kernel == WOTK_KERNEL_BATCH_SWEEP || kernel == WOTK_KERNEL_BATCH_SHARED) {
        const int e = n_combos < 1 ? 1 : (n_combos > 16 ? 16 : n_combos);
        return (e - 1) * 32 + lg(n_cols);
    }
    return -1;
}

extern "C" int32_t wotk_num_buckets(int32_t kernel) {
    int nb = 0, nt = 0;
    const wotk::TileShape *t = nullptr;
    return table(kernel, &nb, &t, &nt) ? nb : -1;
}

extern "C" int32_t wotk_tile_variants(int32_t kernel, wotk_tile *out, int32_t capacity) {
    int nb = 0, nt = 0;
    const wotk::TileShape *t = nullptr;
    if (!table(kernel, &nb, &t, &nt)) return wotk::fail(WOTK_EINVAL, "unknown kernel %d", kernel), -1;
    for (int i = 0; i < nt && out && i < capacity; ++i) out[i] = wotk_tile{t[i].rows, t[i].threads};
    return nt;
}

extern "C" int wotk_set_tile(int32_t kernel, int32_t bucket, const wotk_tile *tile) {
    int nb = 0, nt = 0;
    const wotk::TileShape *t = nullptr;
    std::atomic<int> *tab = table(kernel, &nb, &t, &nt);
    WOTK_REQUIRE(tab, "unknown kernel %d", kernel);
    WOTK_REQUIRE(bucket >= -1 && bucket < nb, "bucket %d out of range [-1, %d)", bucket, nb);
    int v = -1;
    if (tile) {
        for (int i = 0; i < nt; ++i)
            if (t[i].rows == tile->rows && t[i].threads == tile->threads) v = i;
        WOTK_REQUIRE(v >= 0, "tile (%d rows, %d threads) is not a compiled variant of kernel %d", tile->rows,
                     tile->threads, kernel);
    }
    if (bucket == -1)
        for (int b = 0; b < nb; ++b) tab[b].store(v);
    else
        tab[bucket].store(v);
    return WOTK_OK;
}

extern "C" int wotk_get_tile(int32_t kernel, int32_t bucket, wotk_tile *out) {
    int nb = 0, nt = 0;
    const wotk::TileShape *t = nullptr;
    std::atomic<int> *tab = table(kernel, &nb, &t, &nt);
    WOTK_REQUIRE(tab, "unknown kernel %d", kernel);
    WOTK_REQUIRE(bucket >= 0 && bucket < nb && out, "bucket %d out of range [0, %d)", bucket, nb);
    const int v = tab[bucket].load();
    *out = wotk_tile{t[v < 0 ? 0 : v].rows, t[v < 0 ? 0 : v].threads};
    return WOTK_OK;
}

namespace wotk {
int tile_for(int kernel, int64_t n_rows, int64_t n_cols, int n_combos) {
    int nb = 0, nt = 0;
    const TileShape *t = nullptr;
    std::atomic<int> *tab = table(kernel, &nb, &t, &nt);
    const int b = wotk_shape_bucket(kernel, n_rows, n_cols, n_combos);
    if (!tab || b < 0 || b >= nb) return 0;
    const int v = tab[b].load(std::memory_order_relaxed);
    return v < 0 ? 0 : v;
}
}  // namespace wotk
