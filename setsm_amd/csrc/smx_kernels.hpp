// smx_kernels.hpp — argument blocks and per-hypothesis helpers shared by the
// VerticalLineLocus kernels and their host launchers.
#pragma once

#include "smx_device.hpp"
#include "smx_window.hpp"

namespace smx {

constexpr int kRun = 64;   // lanes of a wavefront: the longest run of heights one wave evaluates
// Run lengths (heights per run) a launch may use; a wavefront evaluates 64 / R
// runs side by side. The plain FP64 and RA / matching-rate passes use 64.
constexpr int kRunLens[3] = {16, 32, 64};
__host__ __device__ constexpr int run_len_slot(int R) { return R == 16 ? 0 : R == 32 ? 1 : 2; }
// RA / matching-rate passes: a NaN payload marking "no NCC at this height"
// (FindPeakNcc is not called there), distinct from an evaluated NaN.
constexpr unsigned long long kNotEvaluated = 0x7ff80000000bad01ull;

struct CellArgs {
    long ncells;
    const smx_ugrid *gridpt3;
    smx_nccresult *ncc;
    const double *grid_wgs;
    const int64_t *vox_off;        // CSR row offsets (ncells + 1)
    uint32_t *vox;                 // VOXEL words
    float *rho;                    // optional mean NCC per hypothesis
    int64_t *run_count;            // per cell (ncells + 1)
    const int64_t *run_offset;     // exclusive prefix sum of run_count
    int2 *runs;                    // (cell, first height)
    long runs_cap;                 // capacity of runs (host-known bound on the count)
    int run_len;                   // heights per run (16, 32 or 64)
    const int64_t *nruns_dev;      // the count itself, on the device (run_offset[ncells])
    unsigned long long *max_bits;  // per cell, order_key of max_WNCC
    unsigned long long *accessible;
    unsigned long long *evaluated;
    uint8_t *acc_flag;             // per cell: passed the gate (written when non-null)
    const double *gval;            // per cell: sum_GNCC_multi / count_GNCC, NaN if no GNCC term; null: none
    double ncc_alpha;              // SetNCC_alpha of the call (GNCC blend)
    double *peak_rho;              // RA / matching-rate pass: temp_rho per run lane (NaN: not evaluated);
                                   // the cube is not touched
    int2 *redo;                    // FAST: runs handed to the FP64 pass (capacity runs_cap)
    double *part;                  // FAST, combined iterations: the level term per hypothesis
    unsigned long long *redo_n;    // FAST: their count
};

// setsm_code.cpp:6481-6490: the value max_WNCC takes for one hypothesis —
// INCC blended with the cell's GNCC when that term exists (0/0 stays NaN,
// which never raises the maximum, as in the reference).
__device__ __forceinline__ double wncc_of(const CellArgs &A, long cell, double sum, double cnt, double rr)
{
#pragma clang fp contract(off)
    if (A.gval) {
        const double g = A.gval[cell];
        if (g == g) {
            const double t = sum / cnt * A.ncc_alpha + g * (1.0 - A.ncc_alpha);
            return t == t ? t : -100.0;
        }
    }
    return rr;
}

// SetOrthoImageCoord / GNCC geometry (setsm_code.cpp:6550-6670, 6240-6305).
struct OrthoArgs {
    long W, H;                     // Size_Grid2D
    long sw, sh, swn, shn;         // ortho grid at im_resolution (and at the next level's)
    double b0, b1;                 // Boundary[0], Boundary[1]
    double im_res, im_res_next, grid_res;
    int combined;                  // check_combined_WNCC
    int slot;                      // GridPT3.ortho_ncc[] index of the target image
    double ortho_th;
    const double *grid_pts;        // GridPts (x, y)
    const double *ortho_wgs;       // (lon, lat) per ortho pixel
    const smx_ugrid *gridpt3;
    double *cd[2];                 // pyramid image (x, y) per ortho pixel, image 0 / 1
    double *cdn[2];                // the same at level - 1
};

// VerticalLineLocus_blunder (setsm_code.cpp:7882-8013): the ortho grid and
// pyramid views of the blunder level, and the cell grid it scores.
struct BlunderArgs {
    long W, H;                     // Size_Grid2D
    long sw, sh;                   // ortho grid at im_res = resolution * 2^level_b
    double b0, b1, im_res, grid_res;
    int half;                      // Half_template_size
    int level_b;                   // blunder_selected_level
    int right;                     // image scored against image 0 (1; 0 when SETSM's reference is image 1)
    DevView view[2];               // the blunder level's views
    const double *grid_pts;
    const double *ortho_wgs;       // (lon, lat) per ortho pixel
    smx_ugrid *gridpt3;
    double *cd[2];                 // pyramid image (x, y) per ortho pixel, image 0 / 1
    float *out;                    // per-cell ortho NCC
};

struct InitArgs {
    long ncells;
    smx_nccresult *ncc;
    smx_ugrid *gridpt3;
    const double *grid_wgs;
    uint8_t *action;
    int64_t *new_len;              // ncells + 1
    const int64_t *old_off;
    const int64_t *new_off;
    const uint32_t *old_vox;
    uint32_t *new_vox;
    const float *old_rho;
    float *new_rho;
};

struct PeakArgs {
    long ncells;
    smx_nccresult *ncc;
    const smx_ugrid *gridpt3;
    const int64_t *vox_off;
    const uint32_t *vox;
    const float *sum;              // SGM SumCost (check_SGM_peak), or null
};

__device__ __forceinline__ unsigned long long order_key(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double order_value(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

// Everything about one (cell, height) hypothesis before the window loop.
struct Hyp {
    double rpx, rpy, tpx, tpy;     // pyramid coordinates, reference / target
    double rnx, rny, tnx, tny;     // same at level-1 (combined INCC)
    double cos0, sin0;
    float height;
    bool orient_ok;
};

// setsm_code.cpp:6329-6428 for one lane: float height, RPC pair, pyramid
// coordinates, image test, orientation difference and rotation. Returns false
// when the hypothesis is skipped without touching the cube.
__device__ __forceinline__ bool project_hypothesis(const VllParams &P, const CellArgs &A, long cell, int hidx,
                                                   const smx_nccresult &nr, int start_H, int end_H, Hyp &hy)
{
#pragma clang fp contract(off)
    if (hidx >= (int)nr.NumOfHeight) return false;
    // setsm_code.cpp:6325-6328: RA / matching-rate passes count heights from
    // the search-space bottom, the voxel passes from NCCresult.minHeight
    const int base = (P.is_ra || P.check_matching_rate) ? start_H : (int)nr.minHeight;
    const float iter_height = (float)(base + hidx * P.height_step);
    if (!(iter_height >= start_H && iter_height <= end_H)) return false;
    hy.height = iter_height;
    const double lon = A.grid_wgs[2 * cell], lat = A.grid_wgs[2 * cell + 1];
    double rix, riy, tix, tiy;
    rpc_ground_to_image(P.rpc[0], P.adjust[0], lon, lat, (double)iter_height, rix, riy);
    rpc_ground_to_image(P.rpc[1], P.adjust[1], lon, lat, (double)iter_height, tix, tiy);
    to_pyramid(rix, riy, P.view[0], P.level, hy.rpx, hy.rpy);
    to_pyramid(tix, tiy, P.view[1], P.level, hy.tpx, hy.tpy);
    if (!(py_point_inside(hy.rpx, hy.rpy, P.view[0]) && py_point_inside(hy.tpx, hy.tpy, P.view[1]))) return false;
    int diff2;
    const int od = (int)P.view[0].ori[(size_t)(int)hy.rpy * P.view[0].width + (int)hy.rpx] -
                   (int)P.view[1].ori[(size_t)(int)hy.tpy * P.view[1].width + (int)hy.tpx];
    if (P.combined_incc) {
        to_pyramid(rix, riy, P.next[0], P.level - 1, hy.rnx, hy.rny);
        to_pyramid(tix, tiy, P.next[1], P.level - 1, hy.tnx, hy.tny);
        if (!(py_point_inside(hy.rnx, hy.rny, P.next[0]) && py_point_inside(hy.tnx, hy.tny, P.next[1])))
            return false;
        const int odn = (int)P.next[0].ori[(size_t)(int)hy.rny * P.next[0].width + (int)hy.rnx] -
                        (int)P.next[1].ori[(size_t)(int)hy.tny * P.next[1].width + (int)hy.tnx];
        diff2 = od + odn;
    } else {
        hy.rnx = hy.rny = hy.tnx = hy.tny = 0.0;
        diff2 = 2 * od;
    }
    const double diff_theta = P.combined_incc ? ((double)od + (double)(diff2 - od)) / 2.0 : (double)od;
    if (P.is_ra != 1)
        hy.orient_ok = P.level <= 2 || (diff_theta * P.bin_angle < 90 && diff_theta * P.bin_angle > -90);
    else
        hy.orient_ok = true;
    hy.cos0 = P.rot_cos[diff2 + 34];
    hy.sin0 = P.rot_sin[diff2 + 34];
    return true;
}

// Max of temp_rho over each run of a wavefront (runs of R lanes) into its
// cell's max_WNCC key. Returns the wave's count of hypotheses that computed
// an NCC (the same value on every lane). Every lane of the wave must call it;
// cell_key is the key of the lane's run.
//
// The count is not added to the launch's counter here: callers sum it over
// the runs a wave evaluates and add it once (wave_count_commit). One global
// counter updated by every run would serialise tens of millions of
// device-scope atomics on one address per launch — they, not the windows,
// would set the kernel's duration.
__device__ __forceinline__ unsigned wave_max_commit(double v, unsigned evaluated, unsigned long long *cell_key,
                                                    int R = kRun)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const double o = __shfl_xor(v, m, 64);
        if (m < R) v = o > v ? o : v;
    }
    const unsigned long long any = __ballot(evaluated != 0);
    const int lane = threadIdx.x & 63;
    const unsigned long long mine = R == 64 ? any : (any >> (lane & ~(R - 1))) & ((1ull << R) - 1);
    if ((lane & (R - 1)) == 0 && mine) atomicMax(cell_key, order_key(v));
    return (unsigned)__popcll(any);
}

// Adds a wave's count (uniform over the wave) to a launch counter, once.
__device__ __forceinline__ void wave_count_commit(unsigned long long *counter, unsigned long long n)
{
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(counter, n);
}

}  // namespace smx
