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
            for (int i = 0; i < n && k < F.runs_cap; i++) out[k++] = make_int2((int)c, i * A.run_len);
        }
    }
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Launch shapes the autotuner chooses from: {tx, ty, lds_texels, occ, stage,
// wide, run_len} (cells per tile, LDS for both boxes in float2 units — the
// left box takes one per texel, the right box two —, register cap for 3 / 4
// waves per SIMD or 0, LDS staging, FP64 window totals, heights per run).
// Budgets stay at or below 7680 units (60 KB, two workgroups per CU), and at
// or below 6656 (52 KB) where the register cap asks for three.
static const FastCfg kFastCfgs[] = {
    {1, 4, 4608, 0, 1, 1, 64}, {2, 4, 5120, 0, 1, 1, 64}, {2, 2, 4096, 0, 1, 1, 64}, {4, 4, 6144, 0, 1, 1, 64},
    {1, 8, 6144, 0, 1, 1, 64}, {2, 8, 7680, 0, 1, 1, 64}, {4, 8, 7680, 0, 1, 1, 64}, {1, 16, 7680, 0, 1, 1, 64},
    {1, 4, 4608, 3, 1, 1, 64}, {2, 4, 5120, 3, 1, 1, 64}, {4, 4, 6656, 3, 1, 1, 64}, {1, 4, 4608, 4, 1, 1, 64},
    {2, 4, 5120, 0, 0, 1, 64}, {1, 4, 4608, 0, 1, 0, 64}, {2, 4, 5120, 3, 1, 0, 64}, {2, 8, 6656, 3, 1, 1, 64},
    {1, 8, 6144, 3, 1, 1, 64}, {4, 2, 5120, 3, 1, 1, 64}, {2, 2, 4096, 3, 1, 1, 64}, {2, 4, 5120, 4, 1, 1, 64},
    {8, 4, 7680, 0, 1, 1, 64}, {4, 8, 6656, 3, 1, 1, 64}, {8, 2, 6656, 3, 1, 1, 64}, {4, 4, 5120, 4, 1, 1, 64},
    // runs of 32 and 16 heights
    {1, 16, 7680, 0, 1, 1, 32}, {2, 8, 7680, 0, 1, 1, 32}, {4, 4, 6656, 3, 1, 1, 32}, {2, 8, 6656, 3, 1, 1, 32},
    {1, 16, 7680, 0, 1, 1, 16}, {2, 8, 7680, 0, 1, 1, 16}, {4, 4, 6656, 3, 1, 1, 16}, {1, 8, 6144, 3, 1, 1, 16},
    {1, 8, 6144, 0, 1, 1, 32}, {2, 4, 5120, 3, 1, 1, 32}, {1, 8, 6144, 3, 1, 1, 32}, {4, 8, 7680, 0, 1, 1, 32},
};
constexpr int kNumFastCfgs = sizeof(kFastCfgs) / sizeof(kFastCfgs[0]);
constexpr int kDefaultFastCfg = 3;   // 4 x 4 cells, compiler's register budget, FP64 totals, runs of 64

int fast_config_index(const smx_job *j)
{
    return (j->fast_cfg >= 0 && j->fast_cfg < kNumFastCfgs) ? j->fast_cfg : kDefaultFastCfg;
}

// Tile shape of slot 0 (the configured one) or slot 1 (half the cells per
// side: a combined iteration's level-below pass, whose box has twice the
// texels per side, so it fits the same LDS budget).
static void tile_shape(const FastCfg &cfg, int slot, int &tx, int &ty)
{
    tx = slot ? std::max(1, cfg.tx / 2) : cfg.tx;
    ty = slot ? std::max(1, cfg.ty / 2) : cfg.ty;
}

int fast_run_len(const smx_job *j) { return kFastCfgs[fast_config_index(j)].run_len; }

long fast_tiles(const smx_job *j, int slot)
{
    int tx, ty;
    tile_shape(kFastCfgs[fast_config_index(j)], slot, tx, ty);
    return (long)((j->lv.grid_w + tx - 1) / tx) * (long)((j->lv.grid_h + ty - 1) / ty);
}

// The FastArgs of one job's level in tile slot `slot`, over S's buffers
// (host only).
static int fast_args(const smx_job *j, int slot, const VllScratch &S, FastArgs &F)
{
    const FastCfg cfg = kFastCfgs[fast_config_index(j)];
    F = FastArgs{};
    tile_shape(cfg, slot, F.tx, F.ty);
    F.grid_w = j->lv.grid_w;
    F.grid_h = j->lv.grid_h;
    F.tiles_x = (int)((F.grid_w + F.tx - 1) / F.tx);
    F.tiles_y = (int)((F.grid_h + F.ty - 1) / F.ty);
    F.lds_texels = cfg.lds_texels;
    F.stage = cfg.stage;
    F.runs_cap = S.runs_cap;
    const long ntiles = (long)F.tiles_x * F.tiles_y;
    if (ntiles > (slot ? S.tiles2_cap : S.tiles_cap))
        return fail(SMX_E_ARG, "tile lists hold %ld tiles, the launch shape needs %ld", slot ? S.tiles2_cap : S.tiles_cap,
                    ntiles);
    F.tile_run_count = slot ? S.tile2_count : S.tile_count;
    F.tile_run_off = slot ? S.tile2_off : S.tile_off;
    F.tile_runs = slot ? S.tile2_runs : S.tile_runs;
    return 0;
}

// Queue the kernels that build F's per-tile run lists (count, scan, fill).
static int fast_lists(const CellArgs &A, const FastArgs &F, const VllScratch &S, hipStream_t s)
{
    const long ntiles = (long)F.tiles_x * F.tiles_y;
    const unsigned nb = (unsigned)((ntiles + 255) / 256);
    int64_t *count = const_cast<int64_t *>(F.tile_run_count);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, (unsigned long long *)(count + ntiles), 1);
    hipLaunchKernelGGL(k_tile_count, dim3(nb), dim3(256), 0, s, A, F);
    hipError_t e = job_scan(S, count, const_cast<int64_t *>(F.tile_run_off), ntiles + 1, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_tile_fill, dim3(nb), dim3(256), 0, s, A, F, const_cast<int2 *>(F.tile_runs));
        e = hipGetLastError();
    }
    if (e != hipSuccess) return fail(SMX_E_HIP, "tile lists: %s", hipGetErrorString(e));
    return 0;
}

// A combined iteration (the level-below term added): two passes over the
// same runs, the second on tiles of half the cells per side.
static int launch_vll_fast_combined(smx_job *j, int half, const CellArgs &A, const VllScratch &S, hipStream_t s,
                                    hipEvent_t ev0, hipEvent_t ev1)
{
    const FastCfg cfg = kFastCfgs[fast_config_index(j)];
    FastArgs F1, F2;
    int r = fast_args(j, 0, S, F1);
    if (!r) r = fast_args(j, 1, S, F2);
    if (!r) r = fast_lists(A, F1, S, s);
    if (!r) r = fast_lists(A, F2, S, s);
    if (r) return r;
    const dim3 g1((unsigned)((long)F1.tiles_x * F1.tiles_y)), g2((unsigned)((long)F2.tiles_x * F2.tiles_y));
    const size_t lds = (size_t)cfg.lds_texels * sizeof(float2);
    hipError_t e = hipSuccess;
    if (ev0) (void)hipEventRecord(ev0, s);
    switch (half) {
    case 5: e = fast_launch_passes<5>(cfg.wide != 0, g1, g2, lds, s, j->d_params, A, F1, F2); break;
    case 6: e = fast_launch_passes<6>(cfg.wide != 0, g1, g2, lds, s, j->d_params, A, F1, F2); break;
    case 7: e = fast_launch_passes<7>(cfg.wide != 0, g1, g2, lds, s, j->d_params, A, F1, F2); break;
    default: return fail(SMX_E_UNSUPPORTED, "half %d", half);
    }
    if (ev1) (void)hipEventRecord(ev1, s);
    if (e != hipSuccess) return fail(SMX_E_HIP, "k_vll_fast_pass: %s", hipGetErrorString(e));
    return 0;
}

int launch_vll_fast(smx_job *j, int half, const CellArgs &A, const VllScratch &S, hipStream_t s, hipEvent_t ev0,
                    hipEvent_t ev1)
{
    const FastCfg cfg = kFastCfgs[fast_config_index(j)];
    if (j->hp.combined_incc && A.part && cfg.stage) return launch_vll_fast_combined(j, half, A, S, s, ev0, ev1);
    FastArgs F;
    int r = fast_args(j, 0, S, F);
    if (!r) r = fast_lists(A, F, S, s);
    if (r) return r;
    const long ntiles = (long)F.tiles_x * F.tiles_y;
    const size_t lds = cfg.stage ? (size_t)F.lds_texels * sizeof(float2) : 0;
    hipError_t e = hipSuccess;
    if (ev0) (void)hipEventRecord(ev0, s);
    if (ntiles > 0) {
        switch (half) {
        case 5: e = fast_launch<5>(cfg.occ, cfg.wide != 0, dim3((unsigned)ntiles), lds, s, j->d_params, A, F); break;
        case 6: e = fast_launch<6>(cfg.occ, cfg.wide != 0, dim3((unsigned)ntiles), lds, s, j->d_params, A, F); break;
        case 7: e = fast_launch<7>(cfg.occ, cfg.wide != 0, dim3((unsigned)ntiles), lds, s, j->d_params, A, F); break;
        default: return fail(SMX_E_UNSUPPORTED, "half %d", half);
        }
    }
    if (ev1) (void)hipEventRecord(ev1, s);
    if (e != hipSuccess) return fail(SMX_E_HIP, "k_vll_fast: %s", hipGetErrorString(e));
    return 0;
}

// ------------------------------------------------------------ batches
//
// Descriptor memory of a batch of n jobs: n single-pass or level-pass
// descriptors (slot-0 tiles), then n level-below-pass descriptors (slot-1
// tiles of the combined jobs), then 8 work counters per persistent launch,
// then the VllParams of each descriptor (same index).

size_t fast_batch_desc_bytes(int n)
{
    return (sizeof(BatchJob) + sizeof(VllParams)) * 2 * (size_t)n + 3 * 8 * sizeof(unsigned long long);
}

namespace {
BatchJob *desc_jobs(void *desc) { return (BatchJob *)desc; }
unsigned long long *desc_next(void *desc, int n) { return (unsigned long long *)(desc_jobs(desc) + 2 * (size_t)n); }
VllParams *desc_params(void *desc, int n) { return (VllParams *)(desc_next(desc, n) + 24); }
}  // namespace

namespace {
struct BatchLayout {
    std::vector<BatchJob> desc;   // [single..., combined slot 0...] then [combined slot 1...] at offset n
    std::vector<int> who;         // job index of each descriptor
    int nsingle = 0, ncomb = 0;
    long tiles_single = 0, tiles_level = 0, tiles_below = 0;
};

int batch_layout(smx_job *const *jobs, int n, const CellArgs *As, const VllScratch *Ss, BatchLayout &L)
{
    L.desc.assign(2 * (size_t)n, BatchJob{});
    L.who.assign(2 * (size_t)n, 0);
    std::vector<int> single, comb;
    for (int i = 0; i < n; i++) (jobs[i]->hp.combined_incc ? comb : single).push_back(i);
    L.nsingle = (int)single.size();
    L.ncomb = (int)comb.size();
    int k = 0;
    for (int i : single) {
        FastArgs F;
        int r = fast_args(jobs[i], 0, Ss[i], F);
        if (r) return r;
        const long nt = (long)F.tiles_x * F.tiles_y;
        L.who[k] = i;
        L.desc[k++] = BatchJob{As[i], F, L.tiles_single, nt};
        L.tiles_single += nt;
    }
    int k2 = n;
    for (int i : comb) {
        FastArgs F1, F2;
        int r = fast_args(jobs[i], 0, Ss[i], F1);
        if (!r) r = fast_args(jobs[i], 1, Ss[i], F2);
        if (r) return r;
        const long n1 = (long)F1.tiles_x * F1.tiles_y, n2 = (long)F2.tiles_x * F2.tiles_y;
        L.who[k] = L.who[k2] = i;
        L.desc[k++] = BatchJob{As[i], F1, L.tiles_level, n1};
        L.desc[k2++] = BatchJob{As[i], F2, L.tiles_below, n2};
        L.tiles_level += n1;
        L.tiles_below += n2;
    }
    return 0;
}
}  // namespace

int fast_batch_descriptors(smx_job *const *jobs, int n, const CellArgs *As, const VllScratch *Ss, void *desc,
                           hipStream_t s)
{
    BatchLayout L;
    int r = batch_layout(jobs, n, As, Ss, L);
    if (r) return r;
    // each descriptor's VllParams, copied from its job's current parameters
    std::vector<VllParams> params(L.desc.size());
    for (size_t k = 0; k < L.desc.size(); k++) params[k] = jobs[L.who[k]]->hp;
    hipError_t e = hipMemcpyAsync(desc_jobs(desc), L.desc.data(), sizeof(BatchJob) * L.desc.size(),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(desc_params(desc, n), params.data(), sizeof(VllParams) * params.size(),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(SMX_E_HIP, "batch descriptors: %s", hipGetErrorString(e));
    return 0;
}

// Kernels of one batched launch, on stream s, with the descriptors already
// in `desc` (fast_batch_descriptors on the same inputs).
int launch_vll_fast_batch(smx_job *const *jobs, int n, const CellArgs *As, const VllScratch *Ss, int half,
                          void *desc, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    BatchLayout L;
    int r = batch_layout(jobs, n, As, Ss, L);
    if (r) return r;
    // per-tile run lists of every job (and of the level-below tiles)
    for (int k = 0; k < 2 * n; k++) {
        if (k >= L.nsingle + L.ncomb && k < n) continue;   // unused slot-0 entries
        if (k >= n + L.ncomb) break;
        r = fast_lists(L.desc[k].A, L.desc[k].F, Ss[L.who[k]], s);
        if (r) return r;
    }
    BatchJob *dj = desc_jobs(desc);
    unsigned long long *dnext = desc_next(desc, n);
    const VllParams *dp = desc_params(desc, n);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, dnext, 24);
    // every job's box-fit test uses its own budget: allocate the largest
    int lds_texels = 0;
    for (int i = 0; i < n; i++) lds_texels = std::max(lds_texels, kFastCfgs[fast_config_index(jobs[i])].lds_texels);
    const size_t lds = (size_t)lds_texels * sizeof(float2);
    // register budget: capped when the first job's shape asks for it
    const bool capped = kFastCfgs[fast_config_index(jobs[0])].occ != 0;
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, jobs[0]->device);
    auto launch = [&](int mode, const BatchJob *jl, const VllParams *pl, int nj, long total,
                      unsigned long long *next) -> int {
        if (nj == 0 || total == 0) return 0;
        int per_cu = 0;
        switch (half) {
        case 5: per_cu = fast_batch_blocks_per_cu<5>(mode, capped, lds); break;
        case 6: per_cu = fast_batch_blocks_per_cu<6>(mode, capped, lds); break;
        case 7: per_cu = fast_batch_blocks_per_cu<7>(mode, capped, lds); break;
        default: return fail(SMX_E_UNSUPPORTED, "half %d", half);
        }
        int grid = std::max(1, cus) * std::max(1, per_cu);
        grid = (grid + 7) / 8 * 8;
        const BatchArgs B{jl, pl, nj, total, next};
        hipError_t e = hipSuccess;
        switch (half) {
        case 5: e = fast_launch_batch<5>(mode, capped, grid, lds, s, B); break;
        case 6: e = fast_launch_batch<6>(mode, capped, grid, lds, s, B); break;
        case 7: e = fast_launch_batch<7>(mode, capped, grid, lds, s, B); break;
        }
        if (e != hipSuccess) return fail(SMX_E_HIP, "k_vll_fast_batch: %s", hipGetErrorString(e));
        return 0;
    };
    if (ev0) (void)hipEventRecord(ev0, s);
    r = launch(kSingle, dj, dp, L.nsingle, L.tiles_single, dnext);
    if (!r) r = launch(kLevelPass, dj + L.nsingle, dp + L.nsingle, L.ncomb, L.tiles_level, dnext + 8);
    if (!r) r = launch(kBelowPass, dj + n, dp + n, L.ncomb, L.tiles_below, dnext + 16);
    if (ev1) (void)hipEventRecord(ev1, s);
    return r;
}

int fast_num_configs() { return kNumFastCfgs; }

int fast_config_desc(int i, char *buf, int n)
{
    if (i < 0 || i >= kNumFastCfgs) return -1;
    return snprintf(buf, (size_t)n, "%d %d %d %d %d %d %d", kFastCfgs[i].tx, kFastCfgs[i].ty, kFastCfgs[i].lds_texels,
                    kFastCfgs[i].occ, kFastCfgs[i].stage, kFastCfgs[i].wide, kFastCfgs[i].run_len);
}

}  // namespace smx
