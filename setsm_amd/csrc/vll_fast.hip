// vll_fast.hip — host side of the precision form of VerticalLineLocus: the
// launch shapes the autotuner chooses from, the per-tile run lists, and the
// launch sequences (single job, combined-level passes, batched jobs). The
// kernels are in vll_fast_kernels.hpp, compiled per half-window size.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "vll_fast_types.hpp"

namespace smx {

// ------------------------------------------------------ tile run lists

__global__ void __launch_bounds__(256) k_tile_count(CellArgs A, FastArgs F)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ntiles = (long)F.tiles_x * F.tiles_y;
    if (t >= ntiles) return;
    const int tcx = (int)(t % F.tiles_x), tcy = (int)(t / F.tiles_x);
    int64_t n = 0;
    for (int y = 0; y < F.ty; y++) {
        const long row = (long)tcy * F.ty + y;
        if (row >= F.grid_h) break;
        for (int x = 0; x < F.tx; x++) {
            const long col = (long)tcx * F.tx + x;
            if (col >= F.grid_w) break;
            n += A.run_count[row * F.grid_w + col];
        }
    }
    F.tile_run_count[t] = n;
}

__global__ void __launch_bounds__(256) k_tile_fill(CellArgs A, FastArgs F, int2 *out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ntiles = (long)F.tiles_x * F.tiles_y;
    if (t >= ntiles) return;
    const int tcx = (int)(t % F.tiles_x), tcy = (int)(t / F.tiles_x);
    long k = F.tile_run_off[t];
    for (int y = 0; y < F.ty; y++) {
        const long row = (long)tcy * F.ty + y;
        if (row >= F.grid_h) break;
        for (int x = 0; x < F.tx; x++) {
            const long col = (long)tcx * F.tx + x;
            if (col >= F.grid_w) break;
            const long c = row * F.grid_w + col;
            const int n = (int)A.run_count[c];
            for (int i = 0; i < n; i++) out[k++] = make_int2((int)c, i * kRun);
        }
    }
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Launch shapes the autotuner chooses from (cells per tile, LDS rows per
// image, register cap for four waves per SIMD).
static const FastCfg kFastCfgs[] = {
    {1, 4, 3072, 0, 1, 1}, {2, 4, 3584, 0, 1, 1}, {2, 2, 3072, 0, 1, 1}, {4, 4, 4608, 0, 1, 1},
    {1, 8, 4096, 0, 1, 1}, {2, 8, 4608, 0, 1, 1}, {4, 8, 6144, 0, 1, 1}, {1, 16, 6144, 0, 1, 1},
    {1, 4, 3072, 3, 1, 1}, {2, 4, 3584, 3, 1, 1}, {4, 4, 4608, 3, 1, 1}, {1, 4, 3072, 4, 1, 1},
    {2, 4, 3584, 0, 0, 1}, {1, 4, 3072, 0, 1, 0}, {2, 4, 3584, 3, 1, 0}, {2, 8, 4608, 3, 1, 1},
    {1, 8, 4096, 3, 1, 1}, {4, 2, 3584, 3, 1, 1}, {2, 2, 3072, 3, 1, 1}, {2, 4, 3584, 4, 1, 1},
    {8, 4, 6144, 0, 1, 1}, {4, 16, 7680, 0, 1, 1}, {8, 8, 7680, 0, 1, 1}, {4, 8, 6144, 3, 1, 1},
};
constexpr int kNumFastCfgs = sizeof(kFastCfgs) / sizeof(kFastCfgs[0]);
constexpr int kDefaultFastCfg = 6;   // 4 x 8 cells, compiler's register budget, Kahan: fastest at the
                                     // 16384^2 level-0 workload when tuned

// Per-tile run lists of one job for the FAST kernel: config ci's launch
// shape, with tiles of tx x ty cells (config ci's own when 0), in tile-list
// slot `slot` (1: the level-below pass of a combined iteration).
int fast_prepare(smx_job *j, int ci, const CellArgs &A, FastArgs &F, int slot = 0, int tx = 0, int ty = 0)
{
    const FastCfg cfg = kFastCfgs[ci];
    F = FastArgs{};
    F.tx = tx > 0 ? tx : cfg.tx;
    F.ty = ty > 0 ? ty : cfg.ty;
    F.grid_w = j->lv.grid_w;
    F.grid_h = j->lv.grid_h;
    F.tiles_x = (int)((F.grid_w + F.tx - 1) / F.tx);
    F.tiles_y = (int)((F.grid_h + F.ty - 1) / F.ty);
    F.lds_texels = cfg.lds_texels;
    F.stage = cfg.stage;
    const long ntiles = (long)F.tiles_x * F.tiles_y;
    int r = fast_scratch(j, slot, ntiles, A.nruns);
    if (r) return r;
    int64_t *count = slot ? j->d_tile2_count : j->d_tile_count;
    int64_t *off = slot ? j->d_tile2_off : j->d_tile_off;
    int2 *runs = slot ? j->d_tile2_runs : j->d_tile_runs;
    F.tile_run_count = count;
    F.tile_run_off = off;
    F.tile_runs = runs;
    const unsigned nb = (unsigned)((ntiles + 255) / 256);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, j->stream, (unsigned long long *)(count + ntiles), 1);
    hipLaunchKernelGGL(k_tile_count, dim3(nb), dim3(256), 0, j->stream, A, F);
    r = job_scan(j, count, off, ntiles + 1);
    if (r) return r;
    hipLaunchKernelGGL(k_tile_fill, dim3(nb), dim3(256), 0, j->stream, A, F, runs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SMX_E_HIP, "tile lists: %s", hipGetErrorString(e));
    return 0;
}

static int cfg_index(const smx_job *j)
{
    return (j->fast_cfg >= 0 && j->fast_cfg < kNumFastCfgs) ? j->fast_cfg : kDefaultFastCfg;
}

// A combined iteration (the level-below term added): two passes over the
// same runs, the second on tiles of half the cells per side, so that the
// level below's box (twice the texels per side) fits the same LDS budget.
static int launch_vll_fast_combined(smx_job *j, int ci, int half, const CellArgs &A, hipEvent_t ev0,
                                    hipEvent_t ev1)
{
    const FastCfg cfg = kFastCfgs[ci];
    FastArgs F1, F2;
    int r = fast_prepare(j, ci, A, F1);
    if (!r) r = fast_prepare(j, ci, A, F2, 1, std::max(1, cfg.tx / 2), std::max(1, cfg.ty / 2));
    if (r) return r;
    const dim3 g1((unsigned)((long)F1.tiles_x * F1.tiles_y)), g2((unsigned)((long)F2.tiles_x * F2.tiles_y));
    const size_t lds = (size_t)cfg.lds_texels * sizeof(float2);
    hipError_t e = hipSuccess;
    if (ev0) (void)hipEventRecord(ev0, j->stream);
    switch (half) {
    case 5: e = fast_launch_passes<5>(cfg.kahan != 0, g1, g2, lds, j->stream, j->d_params, A, F1, F2); break;
    case 6: e = fast_launch_passes<6>(cfg.kahan != 0, g1, g2, lds, j->stream, j->d_params, A, F1, F2); break;
    case 7: e = fast_launch_passes<7>(cfg.kahan != 0, g1, g2, lds, j->stream, j->d_params, A, F1, F2); break;
    default: return fail(SMX_E_UNSUPPORTED, "half %d", half);
    }
    if (ev1) (void)hipEventRecord(ev1, j->stream);
    if (e != hipSuccess) return fail(SMX_E_HIP, "k_vll_fast_pass: %s", hipGetErrorString(e));
    return 0;
}

int launch_vll_fast(smx_job *j, int half, const CellArgs &A, hipEvent_t ev0, hipEvent_t ev1)
{
    const int ci = cfg_index(j);
    const FastCfg cfg = kFastCfgs[ci];
    if (j->hp.combined_incc && A.part && cfg.stage) return launch_vll_fast_combined(j, ci, half, A, ev0, ev1);
    FastArgs F;
    int r = fast_prepare(j, ci, A, F);
    if (r) return r;
    const long ntiles = (long)F.tiles_x * F.tiles_y;
    const size_t lds = cfg.stage ? (size_t)F.lds_texels * sizeof(float2) : 0;
    hipError_t e = hipSuccess;
    if (ev0) (void)hipEventRecord(ev0, j->stream);
    switch (half) {
    case 5: e = fast_launch<5>(cfg.occ, cfg.kahan != 0, dim3((unsigned)ntiles), lds, j->stream, j->d_params, A, F); break;
    case 6: e = fast_launch<6>(cfg.occ, cfg.kahan != 0, dim3((unsigned)ntiles), lds, j->stream, j->d_params, A, F); break;
    case 7: e = fast_launch<7>(cfg.occ, cfg.kahan != 0, dim3((unsigned)ntiles), lds, j->stream, j->d_params, A, F); break;
    default: return fail(SMX_E_UNSUPPORTED, "half %d", half);
    }
    if (ev1) (void)hipEventRecord(ev1, j->stream);
    if (e != hipSuccess) return fail(SMX_E_HIP, "k_vll_fast: %s", hipGetErrorString(e));
    return 0;
}

// One persistent launch over the tiles of every job (all at the same window
// half size). As[i] are the prepared cell arguments of jobs[i]; the launch
// goes on jobs[0]'s stream after every job's preparation has completed.
int launch_vll_fast_batch(smx_job **jobs, int n, const CellArgs *As, int half, hipEvent_t ev0, hipEvent_t ev1)
{
    smx_job *j0 = jobs[0];
    const int ci = cfg_index(j0);
    const FastCfg cfg = kFastCfgs[ci];
    std::vector<BatchJob> hb((size_t)n);
    long total = 0;
    for (int i = 0; i < n; i++) {
        FastArgs F;
        int r = fast_prepare(jobs[i], ci, As[i], F);
        if (r) return r;
        hb[i] = BatchJob{jobs[i]->d_params, As[i], F, total, (long)F.tiles_x * F.tiles_y};
        total += hb[i].ntiles;
    }
    for (int i = 0; i < n; i++) {
        hipError_t e = hipStreamSynchronize(jobs[i]->stream);
        if (e != hipSuccess) return fail(SMX_E_HIP, "batch prepare: %s", hipGetErrorString(e));
    }
    const size_t need = sizeof(BatchJob) * (size_t)n + 8 * sizeof(unsigned long long);
    if (need > j0->batch_bytes) {
        job_free(j0, j0->d_batch);
        j0->d_batch = nullptr;
        j0->batch_bytes = 0;
        int r = job_alloc(j0, &j0->d_batch, need);
        if (r) return r;
        j0->batch_bytes = need;
    }
    BatchJob *dj = (BatchJob *)j0->d_batch;
    unsigned long long *dnext = (unsigned long long *)(dj + n);
    hipError_t e = hipMemcpyAsync(dj, hb.data(), sizeof(BatchJob) * (size_t)n, hipMemcpyHostToDevice, j0->stream);
    if (e != hipSuccess) return fail(SMX_E_HIP, "batch upload: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, j0->stream, dnext, 8);
    BatchArgs B{dj, n, total, dnext};
    const size_t lds = (size_t)cfg.lds_texels * sizeof(float2);
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, j0->device);
    int per_cu = 0;
    switch (half) {
    case 5: per_cu = fast_batch_blocks_per_cu<5>(lds); break;
    case 6: per_cu = fast_batch_blocks_per_cu<6>(lds); break;
    case 7: per_cu = fast_batch_blocks_per_cu<7>(lds); break;
    default: return fail(SMX_E_UNSUPPORTED, "half %d", half);
    }
    int grid = std::max(1, cus) * std::max(1, per_cu);
    grid = (grid + 7) / 8 * 8;
    if (ev0) (void)hipEventRecord(ev0, j0->stream);
    switch (half) {
    case 5: e = fast_launch_batch<5>(cfg.kahan != 0, grid, lds, j0->stream, B); break;
    case 6: e = fast_launch_batch<6>(cfg.kahan != 0, grid, lds, j0->stream, B); break;
    case 7: e = fast_launch_batch<7>(cfg.kahan != 0, grid, lds, j0->stream, B); break;
    }
    if (ev1) (void)hipEventRecord(ev1, j0->stream);
    if (e != hipSuccess) return fail(SMX_E_HIP, "k_vll_fast_batch: %s", hipGetErrorString(e));
    return 0;
}

int fast_num_configs() { return kNumFastCfgs; }

int fast_config_desc(int i, char *buf, int n)
{
    if (i < 0 || i >= kNumFastCfgs) return -1;
    return snprintf(buf, (size_t)n, "%d %d %d %d %d %d", kFastCfgs[i].tx, kFastCfgs[i].ty, kFastCfgs[i].lds_texels,
                    kFastCfgs[i].occ, kFastCfgs[i].stage, kFastCfgs[i].kahan);
}

}  // namespace smx
