// smx_internal.hpp — host-side job state and the launchers each kernel file
// exports to smx_api.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <utility>
#include <vector>

#include "smx_kernels.hpp"
#include "smx_vll.h"

namespace smx {

struct LevelView {
    uint32_t *texel = nullptr;   // intensity | magnitude << 16
    uint8_t *ori = nullptr;
    uint32_t w = 0, h = 0;
    double sx = 0, sy = 0;
};

// Scratch of one VerticalLineLocus call: the height-run list (in the FAST
// path, which reads per-tile lists instead, the queue of runs handed to the
// FP64 pass), the FAST path's per-tile run lists (two sets: a combined
// iteration's level-below pass has its own tiles), the level term of a
// combined iteration, the scan partials and the call's
// counters. A job owns one set, grown before any launch is queued; the
// no-alloc entry point carves one out of the library's static device array.
struct VllScratch {
    int2 *runs = nullptr;                                   // runs_cap
    long runs_cap = 0;
    int64_t *tile_count = nullptr, *tile_off = nullptr;     // tiles_cap + 1 each
    int2 *tile_runs = nullptr;                              // runs_cap
    long tiles_cap = 0;
    int64_t *tile2_count = nullptr, *tile2_off = nullptr;
    int2 *tile2_runs = nullptr;
    long tiles2_cap = 0;
    double *part = nullptr;                                 // part_cap
    long long part_cap = 0;
    int64_t *scan_part = nullptr;                           // scan_part_cap
    long scan_part_cap = 0;
    unsigned long long *counters = nullptr;                 // kCounters
};

// Counter slots of a call: accessible cells, evaluated hypotheses, FAST tiles
// staged in LDS, FAST waves reading taps from HBM, runs handed to the FP64
// pass, hypotheses that pass evaluated, and two scratch slots.
constexpr int kCounters = 8;

}  // namespace smx

struct smx_job {
    int device = 0;
    int num_levels = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    int64_t *h_small = nullptr;                  // pinned readback slots
    std::vector<smx::LevelView> views[2];
    long long device_bytes = 0;
    long long budget = 0;                        // 0 = unlimited
    std::vector<std::pair<void *, size_t>> sizes;

    // current level
    bool level_set = false;
    bool searchspace_set = false;
    smx_level lv{};
    smx::VllParams hp{};
    smx::VllParams *d_params = nullptr;
    long ncells = 0;
    double minmax[2] = {0, 0};
    int16_t max_ortho_ncc[2] = {0, 0};            // max GridPT3.ortho_ncc[k] over cells
    double *d_grid_wgs = nullptr;
    double *d_grid_pts = nullptr;                 // GridPts (GNCC windows)
    double *d_ortho_wgs = nullptr;                // lon/lat per ortho pixel (GNCC)
    double *d_cd = nullptr;                       // ortho image coords: 2 images x sw*sh (x, y)
    double *d_cdn = nullptr;                      // the same at level - 1
    size_t cd_bytes = 0, cdn_bytes = 0;
    double *d_gval = nullptr;                     // per-cell GNCC value (NaN: none)
    double *d_bortho_wgs = nullptr;               // VerticalLineLocus_blunder: lon/lat of its ortho grid
    int blunder_level = -1;                       // blunder_selected_level of that grid, -1 = none
    uint32_t bortho_w = 0, bortho_h = 0;
    float *d_bncc = nullptr;                      // its last output, the per-cell ortho NCC
    bool bncc_valid = false;
    double *d_peak_rho = nullptr;                 // RA / matching-rate pass: temp_rho per run lane
    long peak_cap = 0;
    smx_ugrid *d_gridpt3 = nullptr;
    smx_nccresult *d_ncc = nullptr;
    int64_t *d_vox_off = nullptr;
    uint32_t *d_vox = nullptr;
    float *d_rho = nullptr;
    long long vox_total = 0;
    float *d_sum = nullptr;                      // SGM SumCost (CSR), levels <= SGM_py
    long long sum_total = 0;
    unsigned char *d_sgm_scratch = nullptr;      // SGM path buffers when they exceed LDS
    size_t sgm_scratch_bytes = 0;

    // per-cell arrays of the launch sequences (allocated with the level)
    int64_t *d_run_count = nullptr, *d_run_off = nullptr;
    unsigned long long *d_max_bits = nullptr;
    uint8_t *d_action = nullptr;
    int64_t *d_new_len = nullptr, *d_new_off = nullptr;

    // the job's own call scratch (grown before a call queues anything;
    // js.counters is allocated with the job)
    smx::VllScratch js;
    long cap_tile_runs = 0, cap_tile2_runs = 0;   // capacities beside js.runs_cap

    int fast_cfg = -1;                           // FAST launch shape (autotuned), -1 = default
    void *d_batch = nullptr;                     // batched-launch descriptors (first job of a batch)
    size_t batch_bytes = 0;
    smx::VllParams *h_params = nullptr;          // pinned staging of hp for its upload

    // statistics of the cube layout, taken when it is (re)built: an upper bound
    // on the runs a (non-RA) VerticalLineLocus call can queue, and the longest
    // height list (SGM path buffers). Launch sequences size their scratch from
    // them instead of reading counts back mid-sequence.
    bool stats_valid = false;
    long long runs_bound[3] = {0, 0, 0};   // runs of 16, 32, 64 heights
    int max_nh = 0;
    // bumped by every call that changes what a captured launch sequence
    // refers to (level, search space, cube layout, launch shape)
    unsigned long long generation = 1;

    double last_ms = 0;
    long long last_evaluated = 0;
    long long last_tiles_lds = 0;   // FAST: tiles whose footprint was staged in LDS
    long long last_waves_hbm = 0;   // FAST: waves that read taps from HBM
    long long last_redo_runs = 0;   // FAST: runs recomputed by the FP64 pass
    long long last_redo_hyps = 0;   // FAST: hypotheses that pass evaluated
    double last_redo_ms = 0;        // FAST: its time
};

namespace smx {

int fail(int code, const char *fmt, ...);
int job_alloc(smx_job *j, void **p, size_t bytes);
void job_free(smx_job *j, void *p);
double ncc_alpha(int Pyramid_step, int iteration, int IsRA);
int vll_half_template(int level, int template_size, double resolution);

hipError_t scan_i64(const int64_t *in, int64_t *out, long n, int64_t *part, hipStream_t s);
long scan_partials(long n);
long scan_max_n();
hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, long n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, long n, hipStream_t s);
hipError_t launch_nh_stats(const smx_nccresult *ncc, long n, unsigned long long *stats, hipStream_t s);
hipError_t launch_gauss_decimate(const uint16_t *in, uint32_t w, uint32_t h, uint16_t *out, hipStream_t s);
hipError_t launch_sobel(const uint16_t *in, uint32_t w, uint32_t h, uint16_t *mag, uint32_t *ow, hipStream_t s);
hipError_t launch_orientation(const uint32_t *ow, uint32_t w, uint32_t h, uint8_t *ori, hipStream_t s);
hipError_t launch_pack(const uint16_t *img, const uint16_t *mag, uint32_t *texel, long n, hipStream_t s);
hipError_t launch_unpack(const uint32_t *texel, uint16_t *img, uint16_t *mag, long n, hipStream_t s);
hipError_t launch_dir_probe(const int *gx, const int *gy, int16_t *out, long n, hipStream_t s);
hipError_t launch_vll_gate(const VllParams *dP, const CellArgs &A, hipStream_t s);
hipError_t launch_find_peak(const VllParams *dP, const CellArgs &A, hipStream_t s);
hipError_t launch_vll_blunder(const VllParams *dP, const BlunderArgs &B, hipStream_t s);
hipError_t launch_ortho_coords(const VllParams *dP, const OrthoArgs &O, hipStream_t s);
hipError_t launch_gncc(int half, const VllParams *dP, const OrthoArgs &O, const CellArgs &A, double *gval,
                       hipStream_t s);
hipError_t launch_vll_fill_runs(const CellArgs &A, hipStream_t s);
hipError_t launch_vll_fp64(int half, const VllParams *dP, const CellArgs &A, hipStream_t s);
hipError_t launch_vll_redo(int half, const VllParams *dP, const CellArgs &A, hipStream_t s);
hipError_t launch_vll_finalize(const CellArgs &A, hipStream_t s);
hipError_t launch_init_voxel(const VllParams *dP, const InitArgs &A, hipStream_t s);
hipError_t launch_relayout(const InitArgs &A, hipStream_t s);
hipError_t launch_awncc(const VllParams *dP, const PeakArgs &A, hipStream_t s);
struct SgmLaunch {
    int W, H;
    double step;
    const smx_nccresult *ncc;
    const smx_ugrid *gridpt3;
    const int64_t *vox_off;
    const uint32_t *vox;
    float *sum;
    int cap;                      // heights per path buffer (>= max NumOfHeight, multiple of 4)
    unsigned char *gscratch;      // null: buffers in LDS (10 * cap bytes per path)
};
hipError_t launch_sgm(const SgmLaunch &L, hipStream_t s);
hipError_t launch_max_nh(const smx_nccresult *ncc, long n, unsigned long long *out, hipStream_t s);
size_t sgm_lds_limit();
// FAST launch sequences (vll_fast.hip). Scratch buffers come from S; the
// tile counts they need are fast_tiles().
int launch_vll_fast(smx_job *j, int half, const CellArgs &A, const VllScratch &S, hipStream_t s, hipEvent_t ev0,
                    hipEvent_t ev1);
// One batched launch over jobs[0..n) (same half-window size): single-pass
// jobs in one persistent launch, combined-iteration jobs in two (the level
// pass, then the level-below pass). `desc` is device memory for
// fast_batch_desc_bytes(n) bytes; the descriptors are copied there on `s`.
size_t fast_batch_desc_bytes(int n);
// Computes the descriptors of a batch and copies them to `desc` (waits on s).
int fast_batch_descriptors(smx_job *const *jobs, int n, const CellArgs *As, const VllScratch *Ss, void *desc,
                           hipStream_t s);
int launch_vll_fast_batch(smx_job *const *jobs, int n, const CellArgs *As, const VllScratch *Ss, int half,
                          void *desc, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1);
// SelectMPs and SetHeightRange (mps.hip)
constexpr int kRohBins = 100;   // SelectMPs' rohconvert
struct MpsDev {
    const smx_nccresult *ncc;
    smx_ugrid *g;
    const double *grid_pts;
    long n;
    int level, is_ra, sgm_py;
    smx_mps_args a;
    unsigned long long *hist;   // kRohBins bins, then the cell total
    double *th;                 // minimum_Th; negative: undefined for this level
    int64_t *sel;               // n + 1 selection flags
    const int64_t *off;         // their exclusive scan
    smx_point3 *pts;
};
hipError_t launch_select_mps(const MpsDev &D, const VllScratch &S, hipStream_t s);   // through the scan
hipError_t launch_write_mps(const MpsDev &D, hipStream_t s);
struct HrDev {
    smx_ugrid *g;            // in: the grid SelectMPs left; after the apply pass: claimed cells updated
    smx_ugrid *out;          // the returned grid
    uint8_t *mb;             // m_bHeight: 0 unclaimed, 1 claimed (2 is decided in the last pass)
    uint32_t *owner;         // first claiming triangle, ~0u if none
    const smx_point3 *pts;
    const smx_tri *tris;
    long npts, ntris;
    long W, H;
    double gres, b0, b1;
    double resolution;
    int level;
    double buf, th_hg, ncc_th;
    unsigned long long *zkeys;   // [0] min, [1] max of the valid triangles' heights (ordered keys)
    int *hbounds;                // [0] minH_grid, [1] maxH_grid
};
hipError_t launch_set_height_range(const HrDev &D, hipStream_t s);
hipError_t launch_set_height_range_blunder(const HrDev &D, float *heights, hipStream_t s);
constexpr int kBdMaxNodes = 30;   // triangles per point blunder_detection_TIN reads
struct BdDev {
    const smx_point3 *pts;
    long npts;
    const smx_tri *tris;
    long ntris;
    int64_t *deg;                 // per point (+1): incident triangles, then the fill cursor
    const int64_t *off;           // their exclusive scan
    uint32_t *adj;                // incident triangles grouped per point (unordered within a point)
    const float *oncc;            // per-cell ortho NCC
    smx_ugrid *g;                 // GridPT3: anchor_flag
    unsigned long long *akey;     // per cell: (highest tested point << 1) | its anchor flag
    uint32_t *state;              // a byte per point: bit 0 its own test, bit 1 a neighbour's
    int *err;                     // a point outside the grid
    long W, H;
    double gres, b0, b1;
    int level, iteration, is_ra, flag_blunder, check_match;
    double height_th, ortho_ncc_th, th_ref_ncc;
};
hipError_t launch_bd_prepare(const BdDev &D, const VllScratch &S, hipStream_t s);   // check + incidence lists
hipError_t launch_bd_points(const BdDev &D, hipStream_t s);
int fast_num_configs();
int fast_config_desc(int i, char *buf, int n);
int fast_config_index(const smx_job *j);
int fast_run_len(const smx_job *j);   // heights per run of the job's launch shape
// tiles of the job's current level for its launch shape: slot 0 the level
// pass, slot 1 the level-below pass of a combined iteration
long fast_tiles(const smx_job *j, int slot);
hipError_t job_scan(const VllScratch &S, const int64_t *in, int64_t *out, long n, hipStream_t s);

}  // namespace smx
