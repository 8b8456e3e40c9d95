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
    double *d_peak_rho = nullptr;                 // RA / matching-rate pass: temp_rho per run lane
    long peak_cap = 0;
    int2 *d_redo = nullptr;                      // FAST: runs handed to the FP64 pass
    long redo_cap = 0;
    double *d_part = nullptr;                    // FAST, combined iterations: level term per hypothesis
    long long part_cap = 0;
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

    // scratch of the launch sequences
    int64_t *d_run_count = nullptr, *d_run_off = nullptr;
    int2 *d_runs = nullptr;
    long runs_cap = 0;
    unsigned long long *d_max_bits = nullptr;
    unsigned long long *d_counters = nullptr;
    uint8_t *d_action = nullptr;
    int64_t *d_new_len = nullptr, *d_new_off = nullptr;
    void *scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;

    // fast-kernel launch shape (autotuned) and its tile run lists
    int fast_cfg = -1;
    int64_t *d_tile_count = nullptr, *d_tile_off = nullptr;
    int2 *d_tile_runs = nullptr;
    long tiles_cap = 0, tile_runs_cap = 0;
    // second tile lists: the level-below pass of a combined iteration
    int64_t *d_tile2_count = nullptr, *d_tile2_off = nullptr;
    int2 *d_tile2_runs = nullptr;
    long tiles2_cap = 0, tile2_runs_cap = 0;
    void *d_batch = nullptr;                     // batched-launch descriptors
    size_t batch_bytes = 0;

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

hipError_t scan_i64(const int64_t *in, int64_t *out, long n, void *tmp, size_t &tmp_bytes, hipStream_t s);
hipError_t launch_gauss_decimate(const uint16_t *in, uint32_t w, uint32_t h, uint16_t *out, hipStream_t s);
hipError_t launch_sobel(const uint16_t *in, uint32_t w, uint32_t h, uint16_t *mag, uint32_t *ow, hipStream_t s);
hipError_t launch_orientation(const uint32_t *ow, uint32_t w, uint32_t h, uint8_t *ori, hipStream_t s);
hipError_t launch_pack(const uint16_t *img, const uint16_t *mag, uint32_t *texel, long n, hipStream_t s);
hipError_t launch_unpack(const uint32_t *texel, uint16_t *img, uint16_t *mag, long n, hipStream_t s);
hipError_t launch_dir_probe(const int *gx, const int *gy, int16_t *out, long n, hipStream_t s);
hipError_t launch_vll_gate(const VllParams *dP, const CellArgs &A, hipStream_t s);
hipError_t launch_find_peak(const VllParams *dP, const CellArgs &A, hipStream_t s);
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
int launch_vll_fast(smx_job *j, int half, const CellArgs &A, hipEvent_t ev0, hipEvent_t ev1);
int launch_vll_fast_batch(smx_job **jobs, int n, const CellArgs *As, int half, hipEvent_t ev0, hipEvent_t ev1);
int fast_num_configs();
int fast_config_desc(int i, char *buf, int n);
int fast_scratch(smx_job *j, int slot, long ntiles, long nruns);
int job_scan(smx_job *j, const int64_t *in, int64_t *out, long n);

}  // namespace smx
