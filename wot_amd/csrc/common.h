// Internal helpers shared by the libwotk translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>

#include "../../include/wotk.h"

namespace wotk {

int fail(int code, const char *fmt, ...);

constexpr double kLog2E = 1.4426950408889634074;  // log2(e)
constexpr double kLn2 = 0.69314718055994530942;   // ln(2)

// Bump allocator over a caller-owned workspace.  With base == nullptr it only
// counts bytes, which is how the *_workspace_size() queries are answered, so
// the size query and the real carve-up can never disagree.
struct Arena {
    char *base;
    size_t cap;
    size_t off = 0;
    bool ok = true;
    Arena(void *b, size_t c) : base(static_cast<char *>(b)), cap(c) {}
    template <typename T>
    T *take(size_t n) {
        off = (off + 255) & ~size_t(255);
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += n * sizeof(T);
        if (base && off > cap) ok = false;
        return p;
    }
};

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Compiled tile variants of the sweep kernels, (rows per workgroup, threads).
// Entry 0 of each list is the built-in default.  The kernels are instantiated
// from these lists, so a variant exists exactly when it is listed here.
struct TileShape {
    int rows, threads;
    int wave_rows = 0;  // > 0: LDS-staged variant, each wave owns wave_rows of the tile's rows
};
constexpr TileShape kRowSweepTiles[] = {{4, 256}, {1, 256}, {2, 256}, {8, 256},
                                        {2, 512}, {4, 512}, {8, 512}, {4, 128}};
constexpr int kNumRowSweepTiles = sizeof(kRowSweepTiles) / sizeof(kRowSweepTiles[0]);
constexpr TileShape kBatchSweepTiles[] = {{2, 256}, {1, 256}, {4, 256}, {1, 512}, {2, 512}};
constexpr int kNumBatchSweepTiles = sizeof(kBatchSweepTiles) / sizeof(kBatchSweepTiles[0]);
constexpr TileShape kBatchSharedTiles[] = {{8, 256, 2},  {4, 256, 1},  {16, 256, 4}, {16, 512, 2},
                                           {8, 512, 1},  {2, 256},     {1, 256},     {2, 512}};
constexpr int kNumBatchSharedTiles = sizeof(kBatchSharedTiles) / sizeof(kBatchSharedTiles[0]);

// variant index (into the lists above) a launch of this shape uses
int tile_for(int kernel, int64_t n_rows, int64_t n_cols, int n_combos);

}  // namespace wotk

#define WOTK_HIP(expr)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return wotk::fail(WOTK_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

#define WOTK_LAUNCH_CHECK()                                                                  \
    do {                                                                                     \
        hipError_t e_ = hipGetLastError();                                                   \
        if (e_ != hipSuccess)                                                                \
            return wotk::fail(WOTK_EHIP, "kernel launch failed at %s:%d: %s", __FILE__,      \
                              __LINE__, hipGetErrorString(e_));                              \
    } while (0)

#define WOTK_REQUIRE(cond, ...)                                                              \
    do {                                                                                     \
        if (!(cond)) return wotk::fail(WOTK_EINVAL, __VA_ARGS__);                            \
    } while (0)
