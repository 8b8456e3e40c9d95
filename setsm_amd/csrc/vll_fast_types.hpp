// vll_fast_types.hpp — argument blocks, launch shapes and the per-half-window
// launch entry points of the FAST kernels. The kernels themselves are in
// vll_fast_kernels.hpp, instantiated once per half-window size in
// vll_fast_h5.hip / vll_fast_h6.hip / vll_fast_h7.hip (compiled in parallel);
// the host side (vll_fast.hip) sees only these declarations.
#pragma once

#include "smx_internal.hpp"

namespace smx {

struct FastArgs {
    const int64_t *tile_run_off;   // ntiles + 1
    const int2 *tile_runs;         // (cell, first height), grouped by tile
    int64_t *tile_run_count;       // ntiles + 1
    int tiles_x, tiles_y;
    int tx, ty;                    // cells per tile
    uint32_t grid_w, grid_h;
    int lds_texels;                // float2 texels of LDS for both images' boxes
    int stage;                     // 0: read every tap from HBM
    long runs_cap;                 // capacity of tile_runs
};

// Launch modes of the FAST kernel. kSingle: the whole hypothesis in one pass
// (the level-below term of a combined iteration, if any, from HBM). A
// combined iteration in two passes over the same runs: kLevelPass computes
// the level term and keeps it per hypothesis (CellArgs::part); kBelowPass
// stages the level below in LDS, adds its term and writes the results.
constexpr int kSingle = 0, kLevelPass = 1, kBelowPass = 2;
// CellArgs::part codes besides a correlation in [-1, 1]
constexpr double kPartNoWindow = -2e300;   // a window count was zero
constexpr double kPartNoCorr = 1e300;      // windows counted, no valid correlation

// One job of a batched launch. Its VllParams is the entry of the same index
// in BatchArgs::params (a copy made for the launch).
struct BatchJob {
    CellArgs A;
    FastArgs F;
    long tile_begin;   // first tile of this job in the flattened tile space
    long ntiles;
};

struct BatchArgs {
    const BatchJob *jobs;
    const VllParams *params;    // params[j]: job j's VllParams
    int njobs;
    long total_tiles;
    unsigned long long *next;   // 8 per-XCD work counters
};

// A launch shape: cells per tile, LDS texels for both boxes, waves per SIMD
// the register budget is capped for (0: the compiler's choice), LDS staging
// on/off, whether window totals are FP64 sums of the FP32 row partials
// (wide) or FP32 sums, and the run length (heights per run: a wavefront
// evaluates 64 / run_len runs, possibly of different cells, side by side).
struct FastCfg {
    int tx, ty, lds_texels, occ, stage, wide, run_len;
};

// Launch entry points, one set per half-window size H (5, 6, 7). The
// persistent batch kernel always keeps FP64 window totals.
template <int H>
hipError_t fast_launch(int occ, bool wide, dim3 grid, size_t lds, hipStream_t s, const VllParams *dP,
                       const CellArgs &A, const FastArgs &F);
template <int H>
hipError_t fast_launch_passes(bool wide, dim3 g1, dim3 g2, size_t lds, hipStream_t s, const VllParams *dP,
                              const CellArgs &A, const FastArgs &F1, const FastArgs &F2);
template <int H>
hipError_t fast_launch_batch(int mode, bool capped, int grid, size_t lds, hipStream_t s, const BatchArgs &B);
template <int H>
int fast_batch_blocks_per_cu(int mode, bool capped, size_t lds);

#define SMX_FAST_EXTERN(H)                                                                                         \
    extern template hipError_t fast_launch<H>(int, bool, dim3, size_t, hipStream_t, const VllParams *,              \
                                              const CellArgs &, const FastArgs &);                                \
    extern template hipError_t fast_launch_passes<H>(bool, dim3, dim3, size_t, hipStream_t, const VllParams *,     \
                                                     const CellArgs &, const FastArgs &, const FastArgs &);       \
    extern template hipError_t fast_launch_batch<H>(int, bool, int, size_t, hipStream_t, const BatchArgs &);            \
    extern template int fast_batch_blocks_per_cu<H>(int, bool, size_t);
SMX_FAST_EXTERN(5)
SMX_FAST_EXTERN(6)
SMX_FAST_EXTERN(7)
#undef SMX_FAST_EXTERN

}  // namespace smx
