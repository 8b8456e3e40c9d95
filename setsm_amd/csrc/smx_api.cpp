// smx_api.cpp — the C ABI of include/smx_vll.h: job state resident in HBM,
// the CSR voxel cube, and the launch sequences behind InitializeVoxel,
// VerticalLineLocus and AWNCC.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <mutex>
#include <map>
#include <thread>

#include "smx_vll.h"
#include "smx_kernels.hpp"
#include "smx_internal.hpp"

namespace smx {

static thread_local std::string g_err;
static thread_local int g_device = 0;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(SMX_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

static unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

// Device blocks are recycled across jobs and levels of one process: a job
// allocates its per-level arrays at every level (SETSM allocates them per
// level too), and hipMalloc / hipFree serialise the host threads of a
// serving loop. Blocks are kept per (device, size class); a freed block is
// cached only once its job's stream is idle, and the cache holds at most
// kCacheCap bytes, beyond which blocks go back to the runtime.
namespace {
constexpr size_t kCacheCap = 64ull << 30;
std::mutex g_cache_mu;
std::map<std::pair<int, size_t>, std::vector<void *>> g_cache;
size_t g_cached = 0;

size_t size_class(size_t b)
{
    if (b <= (1u << 20)) {
        size_t c = 256;
        while (c < b) c <<= 1;
        return c;
    }
    const size_t g = 2u << 20;   // 2 MiB granules above 1 MiB
    return (b + g - 1) / g * g;
}
}  // namespace

int job_alloc(smx_job *j, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return 0;
    const size_t cls = size_class(bytes);
    if (j->budget > 0 && j->device_bytes + (long long)cls > j->budget)
        return fail(SMX_E_NOMEM, "device budget of %lld bytes exceeded (need %zu more)", j->budget, cls);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_cache.find({j->device, cls});
        if (it != g_cache.end() && !it->second.empty()) {
            *p = it->second.back();
            it->second.pop_back();
            g_cached -= cls;
        }
    }
    if (!*p) {
        hipError_t e = hipMalloc(p, cls);
        if (e != hipSuccess) {
            *p = nullptr;
            return fail(SMX_E_NOMEM, "hipMalloc(%zu): %s", cls, hipGetErrorString(e));
        }
    }
    j->device_bytes += (long long)cls;
    j->sizes.push_back({*p, cls});
    j->generation++;   // captured launch sequences refer to the old buffers
    return 0;
}

void job_free(smx_job *j, void *p)
{
    if (!p) return;
    j->generation++;
    size_t cls = 0;
    for (size_t i = 0; i < j->sizes.size(); i++)
        if (j->sizes[i].first == p) {
            cls = j->sizes[i].second;
            j->device_bytes -= (long long)cls;
            j->sizes.erase(j->sizes.begin() + (long)i);
            break;
        }
    // kernels of this job may still read the block
    if (j->stream) (void)hipStreamSynchronize(j->stream);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        if (cls && g_cached + cls <= kCacheCap) {
            g_cache[{j->device, cls}].push_back(p);
            g_cached += cls;
            return;
        }
    }
    (void)hipFree(p);
}

template <class T>
static int jalloc(smx_job *j, T **p, size_t count)
{
    return job_alloc(j, (void **)p, count * sizeof(T));
}

template <class T>
static void jfree(smx_job *j, T *&p)
{
    job_free(j, (void *)p);
    p = nullptr;
}

hipError_t job_scan(const VllScratch &S, const int64_t *in, int64_t *out, long n, hipStream_t s)
{
    if (n > scan_max_n() || scan_partials(n) > S.scan_part_cap) return hipErrorInvalidValue;
    return scan_i64(in, out, n, S.scan_part, s);
}

// Grow a job buffer to hold `need` elements (old contents are not kept).
template <class T, class C>
static int grow(smx_job *j, T *&p, C &cap, long long need)
{
    if (p && (long long)cap >= need) return 0;
    jfree(j, p);
    cap = 0;
    int r = jalloc(j, &p, (size_t)std::max(need, 1LL));
    if (r) return r;
    cap = (C)need;
    return 0;
}

static void free_level_state(smx_job *j)
{
    jfree(j, j->d_grid_wgs);
    jfree(j, j->d_gridpt3);
    jfree(j, j->d_ncc);
    jfree(j, j->d_vox_off);
    jfree(j, j->d_vox);
    jfree(j, j->d_rho);
    jfree(j, j->d_run_count);
    jfree(j, j->d_run_off);
    jfree(j, j->d_max_bits);
    jfree(j, j->d_action);
    jfree(j, j->d_new_len);
    jfree(j, j->d_new_off);
    VllScratch &S = j->js;
    jfree(j, S.runs);
    S.runs_cap = 0;
    jfree(j, S.tile_count);
    jfree(j, S.tile_off);
    jfree(j, S.tile_runs);
    S.tiles_cap = 0;
    jfree(j, S.tile2_count);
    jfree(j, S.tile2_off);
    jfree(j, S.tile2_runs);
    S.tiles2_cap = 0;
    jfree(j, S.part);
    S.part_cap = 0;
    j->cap_tile_runs = j->cap_tile2_runs = 0;
    jfree(j, j->d_sum);
    j->sum_total = 0;
    jfree(j, j->d_grid_pts);
    jfree(j, j->d_ortho_wgs);
    jfree(j, j->d_cd);
    jfree(j, j->d_cdn);
    jfree(j, j->d_gval);
    jfree(j, j->d_bortho_wgs);
    j->blunder_level = -1;
    j->bortho_w = j->bortho_h = 0;
    jfree(j, j->d_bncc);
    j->bncc_valid = false;
    jfree(j, j->d_peak_rho);
    j->peak_cap = 0;
    j->cd_bytes = j->cdn_bytes = 0;
    j->vox_total = 0;
    j->level_set = false;
    j->searchspace_set = false;
    j->stats_valid = false;
    j->runs_bound[0] = j->runs_bound[1] = j->runs_bound[2] = 0;
    j->max_nh = 0;
}

// The per-call VllParams (host copy), for iteration `iteration`.
static void fill_params(smx_job *j, int iteration)
{
    VllParams &P = j->hp;
    const smx_level &L = j->lv;
    P.iteration = iteration;
    P.minmax[0] = j->minmax[0];
    P.minmax[1] = j->minmax[1];
    const bool combined_wncc = !L.is_ra && (L.level > L.py_combined_level);
    P.combined_incc = (combined_wncc && (L.level >= 1 && iteration % 2 == 0)) ? 1 : 0;
    P.ncc_alpha = ncc_alpha(L.level, iteration, L.is_ra);
    bool ortho = !(L.level == 4 && iteration == 1);
    if (L.pre_dem && !(L.level == 4 && iteration == 1)) ortho = true;
    P.awncc_check_ortho = ortho ? 1 : 0;
    for (int ti = 0; ti < 2; ti++) {
        const LevelView &v = j->views[ti][L.level];
        P.view[ti] = DevView{v.texel, v.ori, v.w, v.h, v.sx, v.sy};
        if (combined_wncc && L.level >= 1 && j->views[ti][L.level - 1].texel) {
            const LevelView &n = j->views[ti][L.level - 1];
            P.next[ti] = DevView{n.texel, n.ori, n.w, n.h, n.sx, n.sy};
        } else {
            P.next[ti] = DevView{nullptr, nullptr, 0, 0, 0.0, 0.0};
        }
    }
}

// Copy the host VllParams to the device through the job's pinned staging
// block (every call waits for its stream before returning, so the block is
// free again by the next call).
static int put_params(smx_job *j, hipStream_t s)
{
    *j->h_params = j->hp;
    HIPCHK(hipMemcpyAsync(j->d_params, j->h_params, sizeof(VllParams), hipMemcpyHostToDevice, s));
    return 0;
}

double ncc_alpha(int Pyramid_step, int iteration, int IsRA)
{
    // setsm_code.cpp:6014-6047 SetNCC_alpha
    double a;
    if (Pyramid_step >= 4) {
        a = 1.0 - ((4 - Pyramid_step) * 0.2 + (iteration - 1) * 0.05);
        if (a < 0.8) a = 0.8;
    } else if (Pyramid_step >= 3) {
        a = 1.0 - ((4 - Pyramid_step) * 0.2 + (iteration - 1) * 0.05);
        if (a < 0.6) a = 0.6;
    } else if (Pyramid_step == 2) {
        a = 1.0 - (0.4 + (iteration - 1) * 0.05);
        if (a < 0.4) a = 0.4;
    } else
        a = 1.0 - ((4 - Pyramid_step) * 0.2 + (iteration - 1) * 0.05);
    if (a < 0.1) a = 0.1;
    if (IsRA == 1) a = 1.0;
    return a;
}

int vll_half_template(int level, int template_size, double resolution)
{
    // setsm_code.cpp:6057-6077
    int ts = template_size;
    if (level >= 1) {
        const int t = (int)((5.0 / (resolution * (1 << level))) / 2.0) * 2 + 1;
        if (template_size < t) ts = t;
    }
    int half = ts / 2;
    if (level == 3) half -= 1;
    else if (level < 3) half -= 2;
    return half;
}

// Statistics of the current height lists (the bounds the launch sequences
// size their scratch by), recomputed after the lists changed outside
// InitializeVoxel.
static int refresh_stats(smx_job *j)
{
    unsigned long long *st = j->js.counters + 4;   // four slots between calls
    HIPCHK(launch_fill_u64(st, 0, 4, j->stream));
    HIPCHK(launch_nh_stats(j->d_ncc, j->ncells, st, j->stream));
    HIPCHK(hipMemcpyAsync(j->h_small, st, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    for (int k = 0; k < 3; k++) j->runs_bound[k] = j->h_small[k];
    j->max_nh = (int)j->h_small[3];
    j->stats_valid = true;
    return 0;
}

}  // namespace smx

// ------------------------------------------------------------ static scratch
//
// Scratch of smx_vertical_line_locus_noalloc: one array per GPU, held from
// the library's load until the process ends. Host code never passes the
// array's name to a kernel or a copy (on the host it is only a stand-in for
// the symbol); it uses the device address hipGetSymbolAddress returns, taken
// once per device at job set-up with that device current.
static constexpr size_t kStaticScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace smx {
namespace {
std::mutex g_scratch_mu;
std::map<int, unsigned char *> g_scratch_ptr;   // device -> address of `scratch` on it
std::map<int, int> g_scratch_err;               // device -> hipGetSymbolAddress error
}  // namespace

// Once per device, with that device current.
static void scratch_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(device) || g_scratch_err.count(device)) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p)
        g_scratch_ptr[device] = (unsigned char *)p;
    else
        g_scratch_err[device] = e == hipSuccess ? (int)hipErrorInvalidSymbol : (int)e;
}

static unsigned char *scratch_of(int device, int &err)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) return it->second;
    auto e = g_scratch_err.find(device);
    err = e != g_scratch_err.end() ? e->second : (int)hipErrorNotInitialized;
    return nullptr;
}

// ---------------------------------------------------- VerticalLineLocus

// Host-side decisions of one VerticalLineLocus call.
struct VllPlan {
    int precision = SMX_PREC_FP64;
    bool keep_rho = false;
    bool gncc_on = false;        // the GNCC term is computed this call
    bool combined_wncc = false;  // check_combined_WNCC (setsm_code.cpp:6081-6088)
    bool ra = false;             // RA / matching-rate pass (no cube; FindPeakNcc inline)
    bool fast = false;           // SMX_PREC_FAST kernels
    int run_len = kRun;          // heights per run of the call's run list
    int half = 5;
    int slot = 1;                // GridPT3.ortho_ncc index of the target image
    double ortho_th = 0;
};

// Validates a call and fills j->hp for it; queues nothing.
static int vll_plan(smx_job *j, int iteration, int precision, int keep_rho, VllPlan &p)
{
    if (!j || !j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    if (iteration < 1) return fail(SMX_E_ARG, "iteration starts at 1");
    if (precision != SMX_PREC_FP64 && precision != SMX_PREC_FAST) return fail(SMX_E_ARG, "unknown precision %d", precision);
    const smx_level &L = j->lv;
    p = VllPlan{};
    p.precision = precision;
    p.keep_rho = keep_rho != 0;
    // setsm_code.cpp:6100-6112 check_ortho, 6240 the per-cell GNCC gate
    bool check_ortho = !((L.level == 4 && iteration == 1) || L.is_ra);
    if (L.pre_dem && !(L.level == 4 && iteration == 1)) check_ortho = true;
    if (L.check_matchtag) check_ortho = true;
    if (L.level <= L.sgm_py) check_ortho = false;
    p.ortho_th = 0.7 - (4 - L.level) * 0.10;
    p.slot = 1 - L.reference_id;
    if (p.slot < 0 || p.slot > 1) return fail(SMX_E_ARG, "reference_id must be 0 or 1");
    p.gncc_on = check_ortho && (double)j->max_ortho_ncc[p.slot] / 1000.0 > p.ortho_th;
    p.combined_wncc = !L.is_ra && (L.level > L.py_combined_level);
    if (p.gncc_on) {
        if (!j->d_ortho_wgs || !j->d_grid_pts || !j->d_cd)
            return fail(SMX_E_ARG, "the GNCC term is active (ortho_ncc > %.2f) but the level has no ortho_wgs / "
                                   "grid_pts", p.ortho_th);
        if (p.combined_wncc) {
            for (int ti = 0; ti < 2; ti++)
                if (!j->views[ti][L.level - 1].texel)
                    return fail(SMX_E_STATE, "the combined GNCC term needs level %d", L.level - 1);
            if (!j->d_cdn) return fail(SMX_E_STATE, "no level-below ortho coordinates");
        }
    }
    p.half = j->hp.half;
    if (p.half < 5 || p.half > 7) return fail(SMX_E_UNSUPPORTED, "Half_template_size %d not built (5..7 are)", p.half);
    fill_params(j, iteration);
    if (j->hp.combined_incc)
        for (int ti = 0; ti < 2; ti++)
            if (!j->views[ti][L.level - 1].texel) return fail(SMX_E_STATE, "combined INCC needs level %d", L.level - 1);
    p.ra = L.is_ra || L.check_matching_rate;
    p.fast = precision != SMX_PREC_FP64 && !p.ra;
    p.run_len = p.fast ? fast_run_len(j) : kRun;
    return 0;
}

// Grows the job's own scratch for the call (allocations only; the one queued
// kernel is the NaN fill of a newly allocated mean-NCC array).
static int vll_reserve(smx_job *j, const VllPlan &p)
{
    int r = 0;
    if (!j->stats_valid && (r = refresh_stats(j))) return r;
    VllScratch &S = j->js;
    if ((r = grow(j, S.runs, S.runs_cap, j->runs_bound[run_len_slot(p.run_len)]))) return r;
    if (p.fast) {
        const long t1 = fast_tiles(j, 0);
        if (t1 > S.tiles_cap || !S.tile_count) {
            jfree(j, S.tile_count);
            jfree(j, S.tile_off);
            S.tiles_cap = 0;
            if ((r = jalloc(j, &S.tile_count, (size_t)t1 + 1)) || (r = jalloc(j, &S.tile_off, (size_t)t1 + 1))) return r;
            S.tiles_cap = t1;
        }
        // the hand-over queue of the FP64 pass lives in S.runs (vll_args)
        if ((r = grow(j, S.tile_runs, j->cap_tile_runs, S.runs_cap))) return r;
        if (j->hp.combined_incc) {
            const long t2 = fast_tiles(j, 1);
            if (t2 > S.tiles2_cap || !S.tile2_count) {
                jfree(j, S.tile2_count);
                jfree(j, S.tile2_off);
                S.tiles2_cap = 0;
                if ((r = jalloc(j, &S.tile2_count, (size_t)t2 + 1)) || (r = jalloc(j, &S.tile2_off, (size_t)t2 + 1)))
                    return r;
                S.tiles2_cap = t2;
            }
            if ((r = grow(j, S.tile2_runs, j->cap_tile2_runs, S.runs_cap))) return r;
            if ((r = grow(j, S.part, S.part_cap, j->vox_total))) return r;
        }
    }
    if (p.keep_rho && !j->d_rho && j->vox_total > 0) {
        if ((r = jalloc(j, &j->d_rho, (size_t)j->vox_total))) return r;
        HIPCHK(launch_fill_u32((uint32_t *)j->d_rho, 0xffffffffu, j->vox_total, j->stream));   // NaN
    }
    return 0;
}

// Carves the call's scratch out of the static array (no-alloc entry point).
static int carve_scratch(smx_job *j, const VllPlan &p, VllScratch &S)
{
    int err = 0;
    unsigned char *base = scratch_of(j->device, err);
    if (!base)
        return fail(SMX_E_HIP, "static scratch unavailable on device %d: hipGetSymbolAddress: %s", j->device,
                    hipGetErrorString((hipError_t)err));
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        const size_t at = off;
        off += std::max<size_t>(bytes, 8);
        return at;
    };
    const long runs = (long)j->runs_bound[run_len_slot(p.run_len)];
    const long nscan = scan_partials(j->ncells + 1);
    const size_t o_cnt = take(kCounters * 8), o_scan = take((size_t)nscan * 8), o_runs = take((size_t)runs * 8);
    size_t o_tc = 0, o_to = 0, o_tr = 0, o_t2c = 0, o_t2o = 0, o_t2r = 0, o_part = 0;
    long t1 = 0, t2 = 0;
    if (p.fast) {   // the FP64 hand-over queue shares o_runs (vll_args)
        t1 = fast_tiles(j, 0);
        o_tc = take(((size_t)t1 + 1) * 8);
        o_to = take(((size_t)t1 + 1) * 8);
        o_tr = take((size_t)runs * 8);
        if (j->hp.combined_incc) {
            t2 = fast_tiles(j, 1);
            o_t2c = take(((size_t)t2 + 1) * 8);
            o_t2o = take(((size_t)t2 + 1) * 8);
            o_t2r = take((size_t)runs * 8);
            o_part = take((size_t)j->vox_total * 8);
        }
    }
    if (off > kStaticScratchBytes)
        return fail(SMX_E_NOMEM, "smx_vertical_line_locus_noalloc: this level needs %zu bytes of scratch (%ld height "
                                 "runs), the static array holds %zu; use smx_vertical_line_locus",
                    off, runs, kStaticScratchBytes);
    S = VllScratch{};
    S.counters = (unsigned long long *)(base + o_cnt);
    S.scan_part = (int64_t *)(base + o_scan);
    S.scan_part_cap = nscan;
    S.runs = (int2 *)(base + o_runs);
    S.runs_cap = runs;
    if (p.fast) {
        S.tile_count = (int64_t *)(base + o_tc);
        S.tile_off = (int64_t *)(base + o_to);
        S.tile_runs = (int2 *)(base + o_tr);
        S.tiles_cap = t1;
        if (j->hp.combined_incc) {
            S.tile2_count = (int64_t *)(base + o_t2c);
            S.tile2_off = (int64_t *)(base + o_t2o);
            S.tile2_runs = (int2 *)(base + o_t2r);
            S.tiles2_cap = t2;
            S.part = (double *)(base + o_part);
            S.part_cap = j->vox_total;
        }
    }
    return 0;
}

// The kernel arguments of a call over scratch S (host only).
static void vll_args(smx_job *j, const VllPlan &p, const VllScratch &S, CellArgs &A, OrthoArgs &O)
{
    const smx_level &L = j->lv;
    A = CellArgs{};
    A.ncells = j->ncells;
    A.gridpt3 = j->d_gridpt3;
    A.ncc = j->d_ncc;
    A.grid_wgs = j->d_grid_wgs;
    A.vox_off = j->d_vox_off;
    A.vox = j->d_vox;
    A.rho = p.keep_rho ? j->d_rho : nullptr;
    A.run_count = j->d_run_count;
    A.run_offset = j->d_run_off;
    A.runs = S.runs;
    A.runs_cap = S.runs_cap;
    A.run_len = p.run_len;
    A.nruns_dev = j->d_run_off + j->ncells;
    A.max_bits = j->d_max_bits;
    A.accessible = S.counters;
    A.evaluated = S.counters + 1;
    A.ncc_alpha = j->hp.ncc_alpha;
    A.redo_n = S.counters + 4;
    if (p.fast) {
        // The FAST kernels take their runs from the per-tile lists, built
        // from the per-cell run counts: the cell-order list is not filled
        // (vll_front), and its memory (capacity = the run count) holds the
        // queue of runs handed to the FP64 pass instead.
        A.redo = S.runs;
        if (j->hp.combined_incc) A.part = S.part;
    }
    if (p.ra) A.peak_rho = j->d_peak_rho;
    O = OrthoArgs{};
    if (!p.gncc_on) return;
    // setsm_code.cpp:6121-6129: ortho image coordinates of this iteration's
    // surface (GridPT3.Height)
    const double im_res = L.resolution * (double)(1 << L.level);
    const double im_res_next = p.combined_wncc ? L.resolution * (double)(1 << (L.level - 1)) : L.resolution;
    O.W = L.grid_w;
    O.H = L.grid_h;
    O.sw = L.ortho_w;
    O.sh = L.ortho_h;
    if (p.combined_wncc) {
        O.swn = (long)((L.boundary[2] - L.boundary[0]) / im_res_next) + 1;
        O.shn = (long)((L.boundary[3] - L.boundary[1]) / im_res_next) + 1;
    }
    O.b0 = L.boundary[0];
    O.b1 = L.boundary[1];
    O.im_res = im_res;
    O.im_res_next = im_res_next;
    O.grid_res = L.grid_resolution;
    O.combined = p.combined_wncc ? 1 : 0;
    O.slot = p.slot;
    O.ortho_th = p.ortho_th;
    O.grid_pts = j->d_grid_pts;
    O.ortho_wgs = j->d_ortho_wgs;
    O.gridpt3 = j->d_gridpt3;
    O.cd[0] = j->d_cd;
    O.cd[1] = j->d_cd + (size_t)O.sw * O.sh * 2;
    O.cdn[0] = O.swn ? j->d_cdn : nullptr;
    O.cdn[1] = O.swn ? j->d_cdn + (size_t)O.swn * O.shn * 2 : nullptr;
    A.acc_flag = j->d_action;
    A.gval = j->d_gval;
}

// Queues everything up to the NCC kernel: GNCC geometry, counter clear,
// gate, GNCC term, run-list scan and fill. An RA / matching-rate pass (the
// gate recomputes NumOfHeight from the search space) reads the run count
// back and grows the job's scratch: only where may_grow (eager calls on the
// job's own scratch).
static int vll_front(smx_job *j, const VllPlan &p, VllScratch &S, bool may_grow, hipStream_t s, CellArgs &A,
                     OrthoArgs &O)
{
    if (p.gncc_on) {
        // calloc'd in the reference: pixels it skips read as (0, 0)
        HIPCHK(launch_fill_u64((unsigned long long *)j->d_cd, 0, (long)O.sw * O.sh * 4, s));
        if (O.swn) HIPCHK(launch_fill_u64((unsigned long long *)j->d_cdn, 0, (long)O.swn * O.shn * 4, s));
        HIPCHK(launch_ortho_coords(j->d_params, O, s));
    }
    HIPCHK(launch_fill_u64(S.counters, 0, kCounters, s));
    HIPCHK(launch_vll_gate(j->d_params, A, s));
    if (p.gncc_on) HIPCHK(launch_gncc(p.half, j->d_params, O, A, j->d_gval, s));
    HIPCHK(job_scan(S, j->d_run_count, j->d_run_off, j->ncells + 1, s));
    if (p.ra) {
        if (!may_grow)
            return fail(SMX_E_UNSUPPORTED, "RA and matching-rate passes size their run list from a count read back "
                                           "mid-call: use smx_vertical_line_locus");
        HIPCHK(hipMemcpyAsync(j->h_small, j->d_run_off + j->ncells, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const long nruns = (long)j->h_small[0];
        int r = grow(j, S.runs, S.runs_cap, nruns);
        if (!r && nruns > j->peak_cap) {
            jfree(j, j->d_peak_rho);
            j->peak_cap = 0;
            r = jalloc(j, &j->d_peak_rho, (size_t)std::max(nruns, 1L) * kRun);
            if (!r) j->peak_cap = nruns;
        }
        if (r) return r;
        A.runs = S.runs;
        A.runs_cap = S.runs_cap;
        A.peak_rho = j->d_peak_rho;
        j->stats_valid = false;   // NumOfHeight now follows the search space
    }
    if (!p.fast) HIPCHK(launch_vll_fill_runs(A, s));   // the FAST kernels read per-tile lists
    return 0;
}

// The NCC kernel(s) of a single-job call: FP64, or FAST plus the FP64 pass
// over the runs it hands over. With `timed`, HIP events bracket the kernel.
static int vll_kernels(smx_job *j, const VllPlan &p, const CellArgs &A, const VllScratch &S, hipStream_t s, bool timed)
{
    if (!p.fast) {
        if (timed) HIPCHK(hipEventRecord(j->ev0, s));
        HIPCHK(launch_vll_fp64(p.half, j->d_params, A, s));
        if (timed) HIPCHK(hipEventRecord(j->ev1, s));
        if (A.peak_rho && j->lv.check_matching_rate) HIPCHK(launch_find_peak(j->d_params, A, s));
        return 0;
    }
    int r = launch_vll_fast(j, p.half, A, S, s, timed ? j->ev0 : nullptr, timed ? j->ev1 : nullptr);
    if (r) return r;
    HIPCHK(launch_vll_redo(p.half, j->d_params, A, s));
    if (timed) HIPCHK(hipEventRecord(j->ev2, s));
    return 0;
}

// Waits for stream s and reads the call's counters. Returns Accessable_grid.
static long long vll_finish(smx_job *j, const unsigned long long *counters, hipStream_t s)
{
    HIPCHK(hipMemcpyAsync(j->h_small, counters, kCounters * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    j->last_evaluated = j->h_small[1] + j->h_small[5];
    j->last_redo_hyps = j->h_small[5];
    j->last_tiles_lds = j->h_small[2];
    j->last_waves_hbm = j->h_small[3];
    j->last_redo_runs = j->h_small[4];
    return (long long)j->h_small[0];
}

static int read_times(smx_job *j, bool fast)
{
    float ms = 0, rms = 0;
    HIPCHK(hipEventElapsedTime(&ms, j->ev0, j->ev1));
    if (fast) HIPCHK(hipEventElapsedTime(&rms, j->ev1, j->ev2));
    j->last_ms = ms;
    j->last_redo_ms = rms;
    return 0;
}

// One call on one job over scratch S, queued on the job's stream.
static long long vll_run(smx_job *j, const VllPlan &p, VllScratch &S, bool may_grow)
{
    int r = put_params(j, j->stream);
    if (r) return r;
    CellArgs A;
    OrthoArgs O;
    vll_args(j, p, S, A, O);
    if ((r = vll_front(j, p, S, may_grow, j->stream, A, O))) return r;
    if ((r = vll_kernels(j, p, A, S, j->stream, true))) return r;
    HIPCHK(launch_vll_finalize(A, j->stream));
    const long long acc = vll_finish(j, S.counters, j->stream);
    if (acc < 0) return acc;
    if ((r = read_times(j, p.fast))) return r;
    return acc;
}

// -------------------------------------------------------------- AWNCC

static bool awncc_sgm(const smx_job *j) { return j->lv.level <= j->lv.sgm_py; }

static int sgm_cap(const smx_job *j) { return (int)((std::max(j->max_nh, 1) + 3) & ~3); }

// Buffers of the SGM aggregation (levels <= SGM_py): SumCost and, when the
// path buffers exceed LDS, their global scratch.
static int awncc_reserve(smx_job *j)
{
    if (!awncc_sgm(j)) return 0;
    int r = 0;
    if (!j->stats_valid && (r = refresh_stats(j))) return r;
    if ((r = grow(j, j->d_sum, j->sum_total, j->vox_total))) return r;
    const int cap = sgm_cap(j);
    if ((size_t)10 * cap > sgm_lds_limit()) {
        const size_t need = (size_t)10 * cap * (size_t)std::max(j->lv.grid_w, j->lv.grid_h);
        if ((r = grow(j, j->d_sgm_scratch, j->sgm_scratch_bytes, (long long)need))) return r;
    }
    return 0;
}

// setsm_code.cpp:7017-7702: SGM (levels <= SGM_py) and the peak scan.
static int awncc_enqueue(smx_job *j, hipStream_t s)
{
    const bool sgm = awncc_sgm(j);
    if (sgm) {
        HIPCHK(launch_fill_u32((uint32_t *)j->d_sum, 0u, j->vox_total, s));
        SgmLaunch L{};
        L.W = (int)j->lv.grid_w;
        L.H = (int)j->lv.grid_h;
        L.step = j->lv.height_step;
        L.ncc = j->d_ncc;
        L.gridpt3 = j->d_gridpt3;
        L.vox_off = j->d_vox_off;
        L.vox = j->d_vox;
        L.sum = j->d_sum;
        L.cap = sgm_cap(j);
        L.gscratch = (size_t)10 * L.cap > sgm_lds_limit() ? j->d_sgm_scratch : nullptr;
        HIPCHK(launch_sgm(L, s));
    }
    PeakArgs A{j->ncells, j->d_ncc, j->d_gridpt3, j->d_vox_off, j->d_vox, sgm ? j->d_sum : nullptr};
    HIPCHK(launch_awncc(j->d_params, A, s));
    return 0;
}

}  // namespace smx

// ------------------------------------------------------------ captured graphs

struct smx_graph {
    std::vector<smx_job *> jobs;
    std::vector<unsigned long long> gens;     // each job's generation at capture
    std::vector<smx::VllParams> params;       // each job's VllParams of the captured iteration
    std::vector<int> fast;                    // whether the job's call is FAST
    hipStream_t stream = nullptr;             // the first job's stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void *desc = nullptr;                     // batched-launch descriptors owned by the graph
    double last_ms = 0;
};

using namespace smx;

extern "C" {

const char *smx_last_error(void) { return g_err.c_str(); }
int smx_version(void) { return 10100; }

int smx_set_device(int device)
{
    HIPCHK(hipSetDevice(device));
    g_device = device;
    return 0;
}

int smx_job_create(smx_job **out, int num_levels)
{
    if (!out || num_levels < 1 || num_levels > 8) return fail(SMX_E_ARG, "num_levels must be 1..8");
    smx_job *j = new smx_job();
    j->device = g_device;
    j->num_levels = num_levels;
    for (int ti = 0; ti < 2; ti++) j->views[ti].resize(num_levels);
    hipError_t e = hipSetDevice(j->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&j->ev0);
    if (e == hipSuccess) e = hipEventCreate(&j->ev1);
    if (e == hipSuccess) e = hipEventCreate(&j->ev2);
    if (e == hipSuccess) e = hipHostMalloc((void **)&j->h_small, 64 * sizeof(int64_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&j->h_params, sizeof(VllParams), hipHostMallocDefault);
    if (e != hipSuccess) {
        smx_job_destroy(j);
        return fail(SMX_E_HIP, "job setup: %s", hipGetErrorString(e));
    }
    // the no-alloc entry point's scratch address on this device
    scratch_setup(j->device);
    int r = jalloc(j, &j->d_params, 1);
    if (!r) r = jalloc(j, &j->js.counters, kCounters);
    if (r) {
        smx_job_destroy(j);
        return r;
    }
    *out = j;
    return 0;
}

int smx_job_destroy(smx_job *j)
{
    if (!j) return 0;
    (void)hipSetDevice(j->device);
    if (j->stream) (void)hipStreamSynchronize(j->stream);
    free_level_state(j);
    for (int ti = 0; ti < 2; ti++)
        for (auto &v : j->views[ti]) {
            jfree(j, v.texel);
            jfree(j, v.ori);
        }
    jfree(j, j->js.scan_part);
    j->js.scan_part_cap = 0;
    jfree(j, j->d_sgm_scratch);
    job_free(j, j->d_batch);
    j->d_batch = nullptr;
    jfree(j, j->d_params);
    jfree(j, j->js.counters);
    for (auto &s : j->sizes) (void)hipFree(s.first);
    if (j->h_small) (void)hipHostFree(j->h_small);
    if (j->h_params) (void)hipHostFree(j->h_params);
    if (j->ev0) (void)hipEventDestroy(j->ev0);
    if (j->ev1) (void)hipEventDestroy(j->ev1);
    if (j->ev2) (void)hipEventDestroy(j->ev2);
    if (j->stream) (void)hipStreamDestroy(j->stream);
    delete j;
    return 0;
}

long long smx_job_device_bytes(const smx_job *j) { return j ? j->device_bytes : 0; }

int smx_job_trim(smx_job *j)
{
    if (!j) return fail(SMX_E_ARG, "null job");
    (void)hipSetDevice(j->device);
    VllScratch &S = j->js;
    jfree(j, S.runs);
    S.runs_cap = 0;
    jfree(j, S.tile_count);
    jfree(j, S.tile_off);
    jfree(j, S.tile_runs);
    S.tiles_cap = 0;
    jfree(j, S.tile2_count);
    jfree(j, S.tile2_off);
    jfree(j, S.tile2_runs);
    S.tiles2_cap = 0;
    jfree(j, S.part);
    S.part_cap = 0;
    j->cap_tile_runs = j->cap_tile2_runs = 0;
    job_free(j, j->d_batch);
    j->d_batch = nullptr;
    j->batch_bytes = 0;
    return 0;
}

long long smx_release_cached(void)
{
    std::map<std::pair<int, size_t>, std::vector<void *>> blocks;
    long long bytes = 0;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        blocks.swap(g_cache);
        bytes = (long long)g_cached;
        g_cached = 0;
    }
    for (auto &kv : blocks) {
        (void)hipSetDevice(kv.first.first);
        for (void *p : kv.second) (void)hipFree(p);
    }
    (void)hipSetDevice(g_device);
    return bytes;
}

int smx_job_set_budget(smx_job *j, long long bytes)
{
    if (!j) return fail(SMX_E_ARG, "null job");
    j->budget = bytes;
    return 0;
}

static int put_view(smx_job *j, int img, int level, uint32_t w, uint32_t h, double sx, double sy)
{
    LevelView &v = j->views[img][level];
    jfree(j, v.texel);
    jfree(j, v.ori);
    int r = jalloc(j, &v.texel, (size_t)w * h);
    if (!r) r = jalloc(j, &v.ori, (size_t)w * h);
    if (r) return r;
    v.w = w;
    v.h = h;
    v.sx = sx;
    v.sy = sy;
    return 0;
}

int smx_job_set_view(smx_job *j, int img, int level, const uint16_t *intensity, const uint16_t *magnitude,
                     const uint8_t *orientation, uint32_t width, uint32_t height, double start_x, double start_y)
{
    if (!j || img < 0 || img > 1 || level < 0 || level >= j->num_levels || !intensity || !magnitude || !orientation ||
        width < 2 || height < 2)
        return fail(SMX_E_ARG, "smx_job_set_view: bad argument");
    HIPCHK(hipSetDevice(j->device));
    int r = put_view(j, img, level, width, height, start_x, start_y);
    if (r) return r;
    LevelView &v = j->views[img][level];
    const size_t n = (size_t)width * height;
    uint16_t *d_img = nullptr, *d_mag = nullptr;
    r = jalloc(j, &d_img, n);
    if (!r) r = jalloc(j, &d_mag, n);
    if (r) {
        jfree(j, d_img);
        return r;
    }
    HIPCHK(hipMemcpyAsync(d_img, intensity, n * 2, hipMemcpyDefault, j->stream));
    HIPCHK(hipMemcpyAsync(d_mag, magnitude, n * 2, hipMemcpyDefault, j->stream));
    HIPCHK(hipMemcpyAsync(v.ori, orientation, n, hipMemcpyDefault, j->stream));
    HIPCHK(launch_pack(d_img, d_mag, v.texel, (long)n, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    jfree(j, d_img);
    jfree(j, d_mag);
    return 0;
}

int smx_job_build_pyramid(smx_job *j, int img, const uint16_t *level0, uint32_t width, uint32_t height,
                          double start_x0, double start_y0)
{
    if (!j || img < 0 || img > 1 || !level0 || width < 2 || height < 2)
        return fail(SMX_E_ARG, "smx_job_build_pyramid: bad argument");
    HIPCHK(hipSetDevice(j->device));
    const size_t n0 = (size_t)width * height;
    uint16_t *cur = nullptr, *nxt = nullptr, *mag = nullptr;
    uint32_t *ow = nullptr;   // Sobel magnitude + direction bin per pixel
    int r = jalloc(j, &cur, n0);
    if (!r) r = jalloc(j, &nxt, n0 / 4 + 1);
    if (!r) r = jalloc(j, &mag, n0);
    if (!r) r = jalloc(j, &ow, n0);
    if (r) goto done;
    // host or device memory (hipMemcpyDefault: unified addressing)
    HIPCHK(hipMemcpyAsync(cur, level0, n0 * 2, hipMemcpyDefault, j->stream));
    {
        uint32_t w = width, h = height;
        for (int lv = 0; lv < j->num_levels; lv++) {
            if (lv > 0) {
                HIPCHK(launch_gauss_decimate(cur, w, h, nxt, j->stream));
                w /= 2;
                h /= 2;
                std::swap(cur, nxt);
            }
            // SetPyramidImages (SubFunctions.cpp:3383-3406)
            r = put_view(j, img, lv, w, h, start_x0 / (double)(1 << lv), start_y0 / (double)(1 << lv));
            if (r) goto done;
            LevelView &v = j->views[img][lv];
            HIPCHK(launch_sobel(cur, w, h, mag, ow, j->stream));
            HIPCHK(launch_orientation(ow, w, h, v.ori, j->stream));
            HIPCHK(launch_pack(cur, mag, v.texel, (long)w * h, j->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(j->stream));
done:
    jfree(j, cur);
    jfree(j, nxt);
    jfree(j, mag);
    jfree(j, ow);
    return r;
}

int smx_job_get_view(smx_job *j, int img, int level, uint16_t *intensity, uint16_t *magnitude, uint8_t *orientation)
{
    if (!j || img < 0 || img > 1 || level < 0 || level >= j->num_levels || !j->views[img][level].texel)
        return fail(SMX_E_ARG, "smx_job_get_view: no such view");
    HIPCHK(hipSetDevice(j->device));
    LevelView &v = j->views[img][level];
    const size_t n = (size_t)v.w * v.h;
    uint16_t *d_img = nullptr, *d_mag = nullptr;
    int r = jalloc(j, &d_img, n);
    if (!r) r = jalloc(j, &d_mag, n);
    if (r) {
        jfree(j, d_img);
        return r;
    }
    HIPCHK(launch_unpack(v.texel, d_img, d_mag, (long)n, j->stream));
    if (intensity) HIPCHK(hipMemcpyAsync(intensity, d_img, n * 2, hipMemcpyDeviceToHost, j->stream));
    if (magnitude) HIPCHK(hipMemcpyAsync(magnitude, d_mag, n * 2, hipMemcpyDeviceToHost, j->stream));
    if (orientation) HIPCHK(hipMemcpyAsync(orientation, v.ori, n, hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    jfree(j, d_img);
    jfree(j, d_mag);
    return 0;
}

int smx_job_set_level(smx_job *j, const smx_level *p)
{
    if (!j || !p) return fail(SMX_E_ARG, "smx_job_set_level: null argument");
    if (p->level < 0 || p->level >= j->num_levels) return fail(SMX_E_ARG, "level %d out of range", p->level);
    for (int ti = 0; ti < 2; ti++)
        if (!j->views[ti][p->level].texel) return fail(SMX_E_STATE, "image %d level %d not uploaded", ti, p->level);
    if (p->grid_w == 0 || p->grid_h == 0 || !p->grid_wgs) return fail(SMX_E_ARG, "empty grid");
    const long n = (long)p->grid_w * (long)p->grid_h;
    if (n + 1 > scan_max_n()) return fail(SMX_E_ARG, "grid of %ld cells exceeds the scan limit", n);
    HIPCHK(hipSetDevice(j->device));
    HIPCHK(hipStreamSynchronize(j->stream));
    free_level_state(j);
    j->lv = *p;
    j->ncells = n;
    int r = jalloc(j, &j->d_grid_wgs, (size_t)n * 2);
    if (!r) r = jalloc(j, &j->d_gridpt3, (size_t)n);
    if (!r) r = jalloc(j, &j->d_ncc, (size_t)n);
    if (!r) r = jalloc(j, &j->d_vox_off, (size_t)n + 1);
    if (!r) r = jalloc(j, &j->d_run_count, (size_t)n + 1);
    if (!r) r = jalloc(j, &j->d_run_off, (size_t)n + 1);
    if (!r) r = jalloc(j, &j->d_max_bits, (size_t)n);
    if (!r) r = jalloc(j, &j->d_action, (size_t)n);
    if (!r) r = jalloc(j, &j->d_new_len, (size_t)n + 1);
    if (!r) r = jalloc(j, &j->d_new_off, (size_t)n + 1);
    if (!r) r = grow(j, j->js.scan_part, j->js.scan_part_cap, scan_partials(n + 1));
    if (r) return r;
    HIPCHK(hipMemcpyAsync(j->d_grid_wgs, p->grid_wgs, (size_t)n * 16, hipMemcpyHostToDevice, j->stream));
    if (p->grid_pts) {
        r = jalloc(j, &j->d_grid_pts, (size_t)n * 2);
        if (r) return r;
        HIPCHK(hipMemcpyAsync(j->d_grid_pts, p->grid_pts, (size_t)n * 16, hipMemcpyHostToDevice, j->stream));
    }
    VllParams &P = j->hp;
    std::memset(&P, 0, sizeof P);
    P.rpc[0] = p->rpc[0];
    P.rpc[1] = p->rpc[1];
    for (int ti = 0; ti < 2; ti++)
        for (int k = 0; k < 2; k++) P.adjust[ti][k] = p->adjust[ti][k];
    P.height_step = p->height_step;
    P.bin_angle = p->bin_angle;
    P.th_height = p->dem_resolution <= 4 ? 500 : 1000;
    P.level = p->level;
    P.half = vll_half_template(p->level, p->template_size, p->resolution);
    P.is_ra = p->is_ra;
    P.check_matching_rate = p->check_matching_rate;
    P.check_matchtag = p->check_matchtag;
    P.dem_lt2 = (p->level == 0 && p->dem_resolution < 2) ? 1 : 0;
    // setsm_code.cpp:6425-6427: cos/sin(-rot_theta) for every half-step
    // orientation difference, with the host libm the CPU path uses.
    for (int k = 0; k < 69; k++) {
        const double diff_theta = (k - 34) / 2.0;
        const double rot_theta = (double)(diff_theta * p->bin_angle * 3.141592653589793 / 180.0);
        P.rot_cos[k] = cos(-rot_theta);
        P.rot_sin[k] = sin(-rot_theta);
    }
    j->lv.ortho_w = j->lv.ortho_h = 0;
    j->lv.ortho_wgs = nullptr;
    if (p->ortho_wgs) {
        j->level_set = true;
        r = smx_job_set_ortho_wgs(j, p->ortho_wgs, p->ortho_w, p->ortho_h);
        if (r) {
            j->level_set = false;
            return r;
        }
    }
    // SETSM callocs nccresult per level (setsm_code.cpp:3769): all zero.
    HIPCHK(hipMemsetAsync(j->d_ncc, 0, (size_t)n * sizeof(smx_nccresult), j->stream));
    HIPCHK(hipMemsetAsync(j->d_vox_off, 0, ((size_t)n + 1) * sizeof(int64_t), j->stream));
    HIPCHK(hipMemsetAsync(j->d_run_count, 0, ((size_t)n + 1) * sizeof(int64_t), j->stream));
    HIPCHK(hipMemsetAsync(j->d_new_len, 0, ((size_t)n + 1) * sizeof(int64_t), j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    j->stats_valid = true;   // no height lists yet
    j->runs_bound[0] = j->runs_bound[1] = j->runs_bound[2] = 0;
    j->max_nh = 0;
    j->level_set = true;
    j->generation++;
    return 0;
}

int smx_job_set_ortho_wgs(smx_job *j, const double *lonlat, uint32_t w, uint32_t h)
{
    if (!j || !lonlat) return fail(SMX_E_ARG, "smx_job_set_ortho_wgs: null argument");
    if (!j->level_set) return fail(SMX_E_STATE, "no level set");
    const smx_level &p = j->lv;
    // setsm_code.cpp:6552-6553: the ortho grid SetOrthoImageCoord walks
    const double im_res = p.resolution * (double)(1 << p.level);
    const long sw = (long)((p.boundary[2] - p.boundary[0]) / im_res) + 1;
    const long sh = (long)((p.boundary[3] - p.boundary[1]) / im_res) + 1;
    if (sw != (long)w || sh != (long)h)
        return fail(SMX_E_ARG, "ortho grid is %ldx%ld at this level and boundary, got %ux%u", sw, sh, w, h);
    HIPCHK(hipSetDevice(j->device));
    jfree(j, j->d_ortho_wgs);
    int r = jalloc(j, &j->d_ortho_wgs, (size_t)sw * sh * 2);
    if (r) return r;
    // the GNCC term's buffers: image coordinates of every ortho pixel (and of
    // the level-below grid when the level has a combined term), per-cell value
    const bool combined_wncc = !p.is_ra && (p.level > p.py_combined_level) && p.level >= 1;
    const size_t cdb = (size_t)sw * sh * 2 * 2;
    size_t cdnb = 0;
    if (combined_wncc) {
        const double im_res_next = p.resolution * (double)(1 << (p.level - 1));
        cdnb = (size_t)((long)((p.boundary[2] - p.boundary[0]) / im_res_next) + 1) *
               (size_t)((long)((p.boundary[3] - p.boundary[1]) / im_res_next) + 1) * 2 * 2;
    }
    if ((r = grow(j, j->d_cd, j->cd_bytes, (long long)cdb))) return r;
    if (cdnb && (r = grow(j, j->d_cdn, j->cdn_bytes, (long long)cdnb))) return r;
    if (!j->d_gval && (r = jalloc(j, &j->d_gval, (size_t)j->ncells))) return r;
    HIPCHK(hipMemcpyAsync(j->d_ortho_wgs, lonlat, (size_t)sw * sh * 16, hipMemcpyHostToDevice, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    j->lv.ortho_w = w;
    j->lv.ortho_h = h;
    j->generation++;
    return 0;
}

int smx_job_set_searchspace(smx_job *j, const smx_ugrid *g, const double minmax[2])
{
    if (!j || !g || !minmax) return fail(SMX_E_ARG, "smx_job_set_searchspace: null argument");
    if (!j->level_set) return fail(SMX_E_STATE, "no level set");
    HIPCHK(hipSetDevice(j->device));
    int16_t mx[2] = {-32768, -32768};
    for (long c = 0; c < j->ncells; c++)
        for (int k = 0; k < 2; k++) mx[k] = std::max(mx[k], g[c].ortho_ncc[k]);
    j->max_ortho_ncc[0] = mx[0];
    j->max_ortho_ncc[1] = mx[1];
    j->minmax[0] = minmax[0];
    j->minmax[1] = minmax[1];
    HIPCHK(hipMemcpyAsync(j->d_gridpt3, g, (size_t)j->ncells * sizeof(smx_ugrid), hipMemcpyHostToDevice, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    j->searchspace_set = true;
    j->generation++;
    return 0;
}

int smx_job_get_searchspace(smx_job *j, smx_ugrid *g)
{
    if (!j || !g || !j->searchspace_set) return fail(SMX_E_STATE, "no search space");
    HIPCHK(hipMemcpyAsync(g, j->d_gridpt3, (size_t)j->ncells * sizeof(smx_ugrid), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    return 0;
}

int smx_job_get_nccresult(smx_job *j, smx_nccresult *out)
{
    if (!j || !out || !j->level_set) return fail(SMX_E_STATE, "no level set");
    HIPCHK(hipMemcpyAsync(out, j->d_ncc, (size_t)j->ncells * sizeof(smx_nccresult), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    return 0;
}

int smx_job_set_nccresult(smx_job *j, const smx_nccresult *in)
{
    if (!j || !in || !j->level_set) return fail(SMX_E_STATE, "no level set");
    HIPCHK(hipMemcpyAsync(j->d_ncc, in, (size_t)j->ncells * sizeof(smx_nccresult), hipMemcpyHostToDevice, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    j->stats_valid = false;   // the height lists may have changed
    j->generation++;
    return 0;
}

int smx_job_half(const smx_job *j)
{
    if (!j || !j->level_set) return fail(SMX_E_STATE, "no level set");
    return j->hp.half;
}

int smx_initialize_voxel(smx_job *j, int iteration)
{
    if (!j || !j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    if (iteration < 1) return fail(SMX_E_ARG, "iteration starts at 1");
    HIPCHK(hipSetDevice(j->device));
    fill_params(j, iteration);
    int r = put_params(j, j->stream);
    if (r) return r;
    InitArgs A{};
    A.ncells = j->ncells;
    A.ncc = j->d_ncc;
    A.gridpt3 = j->d_gridpt3;
    A.grid_wgs = j->d_grid_wgs;
    A.action = j->d_action;
    A.new_len = j->d_new_len;
    A.old_off = j->d_vox_off;
    A.new_off = j->d_new_off;
    A.old_vox = j->d_vox;
    A.old_rho = j->d_rho;
    HIPCHK(launch_init_voxel(j->d_params, A, j->stream));
    HIPCHK(job_scan(j->js, j->d_new_len, j->d_new_off, j->ncells + 1, j->stream));
    // the new height lists' bounds for the launch sequences
    unsigned long long *st = j->js.counters + 4;   // four slots between calls
    HIPCHK(launch_fill_u64(st, 0, 4, j->stream));
    HIPCHK(launch_nh_stats(j->d_ncc, j->ncells, st, j->stream));
    HIPCHK(hipMemcpyAsync(j->h_small, j->d_new_off + j->ncells, sizeof(int64_t), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipMemcpyAsync(j->h_small + 1, st, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    const long long total = j->h_small[0];
    for (int k = 0; k < 3; k++) j->runs_bound[k] = j->h_small[1 + k];
    j->max_nh = (int)j->h_small[4];
    j->stats_valid = true;
    uint32_t *nv = nullptr;
    float *nr = nullptr;
    r = jalloc(j, &nv, (size_t)std::max(total, 1LL));
    if (!r && j->d_rho) r = jalloc(j, &nr, (size_t)std::max(total, 1LL));
    if (r) {
        jfree(j, nv);
        return r;
    }
    A.new_vox = nv;
    A.new_rho = nr;
    HIPCHK(launch_relayout(A, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    jfree(j, j->d_vox);
    jfree(j, j->d_rho);
    j->d_vox = nv;
    j->d_rho = nr;
    std::swap(j->d_vox_off, j->d_new_off);
    j->vox_total = total;
    j->generation++;
    return 0;
}

long long smx_vertical_line_locus(smx_job *j, int iteration, int precision, int keep_rho)
{
    VllPlan p;
    int r = vll_plan(j, iteration, precision, keep_rho, p);
    if (r) return r;
    HIPCHK(hipSetDevice(j->device));
    if ((r = vll_reserve(j, p))) return r;
    return vll_run(j, p, j->js, true);
}

long long smx_vertical_line_locus_noalloc(smx_job *j, int iteration, int precision)
{
    VllPlan p;
    int r = vll_plan(j, iteration, precision, 0, p);
    if (r) return r;
    if (p.ra)
        return fail(SMX_E_UNSUPPORTED, "smx_vertical_line_locus_noalloc: RA and matching-rate passes grow their run "
                                       "list mid-call; use smx_vertical_line_locus");
    if (!j->stats_valid)
        return fail(SMX_E_STATE, "smx_vertical_line_locus_noalloc: the height lists changed since the last "
                                 "smx_initialize_voxel / smx_job_set_voxels; run one first");
    HIPCHK(hipSetDevice(j->device));
    VllScratch S;
    if ((r = carve_scratch(j, p, S))) return r;
    return vll_run(j, p, S, false);
}

int smx_vertical_line_locus_batch(smx_job **jobs, int n, int iteration, int precision, int keep_rho,
                                  long long *accessible)
{
    if (!jobs || n < 1 || !accessible) return fail(SMX_E_ARG, "smx_vertical_line_locus_batch: bad argument");
    for (int i = 0; i < n; i++) {
        if (!jobs[i]) return fail(SMX_E_ARG, "null job %d", i);
        if (jobs[i]->device != jobs[0]->device) return fail(SMX_E_ARG, "jobs of one batch share a device");
        for (int k = 0; k < i; k++)
            if (jobs[k] == jobs[i]) return fail(SMX_E_ARG, "job %d appears twice in the batch", i);
    }
    // every check before anything is queued on any job
    std::vector<VllPlan> P((size_t)n);
    for (int i = 0; i < n; i++) {
        int r = vll_plan(jobs[i], iteration, precision, keep_rho, P[i]);
        if (r) return r;
    }
    bool ra = false;
    for (int i = 0; i < n; i++) ra = ra || P[i].ra;
    if (precision == SMX_PREC_FP64 || ra) {
        // the plain kernel: one tile per launch (RA / matching-rate passes included)
        for (int i = 0; i < n; i++) {
            const long long a = smx_vertical_line_locus(jobs[i], iteration, precision, keep_rho);
            if (a < 0) return (int)a;
            accessible[i] = a;
        }
        return 0;
    }
    for (int i = 0; i < n; i++)
        if (P[i].half != P[0].half)
            return fail(SMX_E_ARG, "a batch holds one window size (job %d has half %d, job 0 %d)", i, P[i].half,
                        P[0].half);
    smx_job *j0 = jobs[0];
    HIPCHK(hipSetDevice(j0->device));
    for (int i = 0; i < n; i++) {
        int r = vll_reserve(jobs[i], P[i]);
        if (r) return r;
    }
    const size_t need = fast_batch_desc_bytes(n);
    if (need > j0->batch_bytes) {
        job_free(j0, j0->d_batch);
        j0->d_batch = nullptr;
        j0->batch_bytes = 0;
        int r = job_alloc(j0, &j0->d_batch, need);
        if (r) return r;
        j0->batch_bytes = need;
    }
    // the jobs' own earlier work (set-up, fills) before the batch stream's
    for (int i = 0; i < n; i++) HIPCHK(hipStreamSynchronize(jobs[i]->stream));
    const hipStream_t s = j0->stream;
    std::vector<CellArgs> As((size_t)n);
    std::vector<OrthoArgs> Os((size_t)n);
    std::vector<VllScratch> Ss((size_t)n);
    for (int i = 0; i < n; i++) {
        Ss[i] = jobs[i]->js;
        vll_args(jobs[i], P[i], Ss[i], As[i], Os[i]);
        int r = put_params(jobs[i], s);
        if (!r) r = vll_front(jobs[i], P[i], Ss[i], false, s, As[i], Os[i]);
        if (r) return r;
    }
    int r = fast_batch_descriptors(jobs, n, As.data(), Ss.data(), j0->d_batch, s);
    if (!r) r = launch_vll_fast_batch(jobs, n, As.data(), Ss.data(), P[0].half, j0->d_batch, s, j0->ev0, j0->ev1);
    if (r) return r;
    for (int i = 0; i < n; i++) HIPCHK(launch_vll_redo(P[0].half, jobs[i]->d_params, As[i], s));
    HIPCHK(hipEventRecord(j0->ev2, s));
    for (int i = 0; i < n; i++) HIPCHK(launch_vll_finalize(As[i], s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0, rms = 0;
    HIPCHK(hipEventElapsedTime(&ms, j0->ev0, j0->ev1));
    HIPCHK(hipEventElapsedTime(&rms, j0->ev1, j0->ev2));
    for (int i = 0; i < n; i++) {
        const long long a = vll_finish(jobs[i], Ss[i].counters, s);
        if (a < 0) return (int)a;
        accessible[i] = a;
        jobs[i]->last_ms = ms;
        jobs[i]->last_redo_ms = rms;
    }
    return 0;
}

int smx_awncc(smx_job *j, int iteration)
{
    if (!j || !j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    if (iteration < 1) return fail(SMX_E_ARG, "iteration starts at 1");
    HIPCHK(hipSetDevice(j->device));
    fill_params(j, iteration);
    int r = awncc_reserve(j);
    if (!r) r = put_params(j, j->stream);
    if (!r) r = awncc_enqueue(j, j->stream);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(j->stream));
    return 0;
}

int smx_job_get_sumcost(smx_job *j, float *out)
{
    if (!j || !out || !j->level_set) return fail(SMX_E_STATE, "no level set");
    if (!j->d_sum || j->sum_total != j->vox_total)
        return fail(SMX_E_STATE, "no SumCost: AWNCC has not run at an SGM level on this cube");
    HIPCHK(hipSetDevice(j->device));
    if (j->sum_total)
        HIPCHK(hipMemcpyAsync(out, j->d_sum, (size_t)j->sum_total * sizeof(float), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    return 0;
}

// ------------------------------------------------------------ between iterations

}  // extern "C"

namespace {
// Temporaries of one call, returned to the job's allocator when it ends.
struct CallBuffers {
    smx_job *j;
    std::vector<void *> held;
    explicit CallBuffers(smx_job *job) : j(job) {}
    template <class T>
    int get(T **p, size_t count)
    {
        int r = jalloc(j, p, count ? count : 1);
        if (!r) held.push_back((void *)*p);
        return r;
    }
    ~CallBuffers()
    {
        for (void *p : held) job_free(j, p);
    }
};
}  // namespace

extern "C" {

int smx_select_mps(smx_job *j, const smx_mps_args *a, smx_point3 *pts, long cap, long *count, double *minimum_th)
{
    if (!j || !a || !count) return fail(SMX_E_ARG, "smx_select_mps: null argument");
    if (!j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    if (!j->d_grid_pts) return fail(SMX_E_STATE, "the level was set without grid_pts (GridPts)");
    if (cap > 0 && !pts) return fail(SMX_E_ARG, "smx_select_mps: null point buffer");
    HIPCHK(hipSetDevice(j->device));
    const long n = j->ncells;
    CallBuffers B(j);
    MpsDev D{};
    D.ncc = j->d_ncc;
    D.g = j->d_gridpt3;
    D.grid_pts = j->d_grid_pts;
    D.n = n;
    D.level = j->lv.level;
    D.is_ra = j->lv.is_ra;
    D.sgm_py = j->lv.sgm_py;
    D.a = *a;
    double *d_th = nullptr;
    int64_t *d_off = nullptr;
    int r = B.get(&D.hist, kRohBins + 1);
    if (!r) r = B.get(&d_th, 1);
    if (!r) r = B.get(&D.sel, (size_t)n + 1);
    if (!r) r = B.get(&d_off, (size_t)n + 1);
    if (r) return r;
    D.th = d_th;
    D.off = d_off;
    HIPCHK(hipEventRecord(j->ev0, j->stream));
    HIPCHK(launch_select_mps(D, j->js, j->stream));
    HIPCHK(hipEventRecord(j->ev1, j->stream));
    double th = 0;
    HIPCHK(hipMemcpyAsync(&j->h_small[0], d_off + n, sizeof(int64_t), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipMemcpyAsync(&th, d_th, sizeof(double), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    if (minimum_th) *minimum_th = th;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, j->ev0, j->ev1));
    j->last_ms = ms;   // kernel time on the job stream, host copies excluded
    if (th < 0)
        return fail(SMX_E_ARG, "SelectMPs has no peak threshold at level %d (above 4 and above SGM_py)", j->lv.level);
    j->generation++;   // GridPT3 changed
    const long m = (long)j->h_small[0];
    *count = m;
    if (m > cap) return fail(SMX_E_ARG, "%ld matched points exceed the buffer of %ld", m, cap);
    if (m > 0) {
        if ((r = B.get(&D.pts, (size_t)m))) return r;
        HIPCHK(hipEventRecord(j->ev0, j->stream));
        HIPCHK(launch_write_mps(D, j->stream));
        HIPCHK(hipEventRecord(j->ev1, j->stream));
        HIPCHK(hipMemcpyAsync(pts, D.pts, (size_t)m * sizeof(smx_point3), hipMemcpyDeviceToHost, j->stream));
        HIPCHK(hipStreamSynchronize(j->stream));
        HIPCHK(hipEventElapsedTime(&ms, j->ev0, j->ev1));
        j->last_ms += ms;
    }
    return 0;
}

int smx_set_height_range(smx_job *j, const smx_point3 *pts, long npts, const smx_tri *tris, long ntris,
                         const smx_hr_args *a, double *minH_grid, double *maxH_grid)
{
    if (!j || !a || (npts > 0 && !pts) || (ntris > 0 && !tris) || npts < 0 || ntris < 0)
        return fail(SMX_E_ARG, "smx_set_height_range: bad argument");
    if (!j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    if (npts > 0xffffffffL || ntris > (1L << 40)) return fail(SMX_E_ARG, "too many points or triangles");
    HIPCHK(hipSetDevice(j->device));
    const smx_level &L = j->lv;
    const int level = L.level, it = a->iteration;
    const long n = j->ncells;
    // setsm_code.cpp:9864-9923 and :10002-10015: the height buffer, the gap
    // limit and the ortho-NCC threshold of this level and iteration
    double buf = a->mpp * 4.0 * (1 << level);
    if (level == 1) buf = it >= 2 ? a->mpp * 2 : a->mpp * 3;
    else if (level == 0) {
        buf = it == 1 ? a->mpp * 2 : a->mpp;
        if (buf < 0.5) buf = 0.5;
    }
    if (L.pre_dem && buf > a->seed_dem_sigma) buf = a->seed_dem_sigma;
    buf = std::ceil(buf);
    if (buf > 100) buf = 100;
    double th_hg = 100000;
    if (L.dem_resolution >= 4) {
        if (level <= 1) th_hg = 1000;
    } else {
        if (!a->check_matching_rate) {
            if (level == 2 && it >= 5) th_hg = 1000;
            else if (level == 1) th_hg = 500;
            else if (level == 0) th_hg = 100;
        } else if (level <= 2) th_hg = 100;
        if (level == 0 && L.dem_resolution < 2 && it > 1) th_hg = 50;
    }
    double ncc_th;
    if (level == 4) ncc_th = 0.7;
    else if (level == 3) ncc_th = 0.6 - (it - 1) * 0.02;
    else if (level == 2) ncc_th = 0.5 - (it - 1) * 0.02;
    else if (level == 1) ncc_th = 0.4;
    else ncc_th = 0.3;
    if (ncc_th < 0.3) ncc_th = 0.3;

    CallBuffers B(j);
    HrDev D{};
    smx_point3 *d_pts = nullptr;
    smx_tri *d_tris = nullptr;
    int r = B.get(&D.out, (size_t)n);
    if (!r) r = B.get(&D.mb, (size_t)n);
    if (!r) r = B.get(&D.owner, (size_t)n);
    if (!r) r = B.get(&d_pts, (size_t)npts);
    if (!r) r = B.get(&d_tris, (size_t)ntris);
    if (!r) r = B.get(&D.zkeys, 2);
    if (!r) r = B.get(&D.hbounds, 2);
    if (r) return r;
    if (npts) HIPCHK(hipMemcpyAsync(d_pts, pts, (size_t)npts * sizeof(smx_point3), hipMemcpyHostToDevice, j->stream));
    if (ntris) HIPCHK(hipMemcpyAsync(d_tris, tris, (size_t)ntris * sizeof(smx_tri), hipMemcpyHostToDevice, j->stream));
    D.g = j->d_gridpt3;
    D.pts = d_pts;
    D.tris = d_tris;
    D.npts = npts;
    D.ntris = ntris;
    D.W = (long)L.grid_w;
    D.H = (long)L.grid_h;
    D.gres = L.grid_resolution;
    D.b0 = L.boundary[0];
    D.b1 = L.boundary[1];
    D.resolution = L.resolution;
    D.level = level;
    D.buf = buf;
    D.th_hg = th_hg;
    D.ncc_th = ncc_th;
    HIPCHK(hipEventRecord(j->ev0, j->stream));
    HIPCHK(launch_set_height_range(D, j->stream));
    HIPCHK(hipMemcpyAsync(j->d_gridpt3, D.out, (size_t)n * sizeof(smx_ugrid), hipMemcpyDeviceToDevice, j->stream));
    HIPCHK(hipEventRecord(j->ev1, j->stream));
    int hb[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&j->h_small[0], D.hbounds, 2 * sizeof(int), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, j->ev0, j->ev1));
    j->last_ms = ms;   // kernels and the grid copy on the job stream; the uploads excluded
    std::memcpy(hb, &j->h_small[0], sizeof hb);
    if (minH_grid) *minH_grid = (double)hb[0];
    if (maxH_grid) *maxH_grid = (double)hb[1];
    // the returned grid keeps ortho_ncc[1] and zeroes ortho_ncc[0]
    if (n > 0) j->max_ortho_ncc[0] = 0;
    j->generation++;
    return 0;
}

int smx_set_height_range_blunder(smx_job *j, const smx_point3 *pts, long npts, const smx_tri *tris, long ntris)
{
    if (!j || (npts > 0 && !pts) || (ntris > 0 && !tris) || npts < 0 || ntris < 0)
        return fail(SMX_E_ARG, "smx_set_height_range_blunder: bad argument");
    if (!j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    if (npts > 0xffffffffL) return fail(SMX_E_ARG, "too many points");
    HIPCHK(hipSetDevice(j->device));
    const long n = j->ncells;
    CallBuffers B(j);
    HrDev D{};
    smx_point3 *d_pts = nullptr;
    smx_tri *d_tris = nullptr;
    float *d_h = nullptr;
    int r = B.get(&d_h, (size_t)n);
    if (!r) r = B.get(&d_pts, (size_t)npts);
    if (!r) r = B.get(&d_tris, (size_t)ntris);
    if (r) return r;
    if (npts) HIPCHK(hipMemcpyAsync(d_pts, pts, (size_t)npts * sizeof(smx_point3), hipMemcpyHostToDevice, j->stream));
    if (ntris) HIPCHK(hipMemcpyAsync(d_tris, tris, (size_t)ntris * sizeof(smx_tri), hipMemcpyHostToDevice, j->stream));
    D.g = j->d_gridpt3;
    D.pts = d_pts;
    D.tris = d_tris;
    D.npts = npts;
    D.ntris = ntris;
    D.W = (long)j->lv.grid_w;
    D.H = (long)j->lv.grid_h;
    D.gres = j->lv.grid_resolution;
    D.b0 = j->lv.boundary[0];
    D.b1 = j->lv.boundary[1];
    HIPCHK(hipEventRecord(j->ev0, j->stream));
    HIPCHK(launch_set_height_range_blunder(D, d_h, j->stream));
    HIPCHK(hipEventRecord(j->ev1, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, j->ev0, j->ev1));
    j->last_ms = ms;
    j->generation++;
    return 0;
}

int smx_job_set_blunder_level(smx_job *j, int blunder_level, const double *lonlat, uint32_t w, uint32_t h)
{
    if (!j || !lonlat) return fail(SMX_E_ARG, "smx_job_set_blunder_level: null argument");
    if (!j->level_set) return fail(SMX_E_STATE, "no level set");
    if (blunder_level < 0 || blunder_level >= j->num_levels) return fail(SMX_E_ARG, "blunder level out of range");
    for (int ti = 0; ti < 2; ti++)
        if (!j->views[ti][blunder_level].texel)
            return fail(SMX_E_STATE, "image %d level %d not uploaded", ti, blunder_level);
    const smx_level &p = j->lv;
    // setsm_code.cpp:7916 and 6552-6553: the ortho grid at resolution * 2^level
    const double im_res = p.resolution * (double)(1 << blunder_level);
    const long sw = (long)((p.boundary[2] - p.boundary[0]) / im_res) + 1;
    const long sh = (long)((p.boundary[3] - p.boundary[1]) / im_res) + 1;
    if (sw != (long)w || sh != (long)h)
        return fail(SMX_E_ARG, "blunder ortho grid is %ldx%ld at level %d, got %ux%u", sw, sh, blunder_level, w, h);
    HIPCHK(hipSetDevice(j->device));
    jfree(j, j->d_bortho_wgs);
    j->blunder_level = -1;
    int r = jalloc(j, &j->d_bortho_wgs, (size_t)sw * sh * 2);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(j->d_bortho_wgs, lonlat, (size_t)sw * sh * 16, hipMemcpyHostToDevice, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    j->blunder_level = blunder_level;
    j->bortho_w = w;
    j->bortho_h = h;
    return 0;
}

int smx_vertical_line_locus_blunder(smx_job *j, int iteration, int bblunder, float *ortho_ncc)
{
    if (!j) return fail(SMX_E_ARG, "smx_vertical_line_locus_blunder: null job");
    if (!j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    if (j->blunder_level < 0) return fail(SMX_E_STATE, "no blunder level set (smx_job_set_blunder_level)");
    if (!j->d_grid_pts) return fail(SMX_E_STATE, "the level was set without grid_pts (GridPts)");
    HIPCHK(hipSetDevice(j->device));
    const smx_level &L = j->lv;
    const int bl = j->blunder_level;
    // setsm_code.cpp:7884-7911: the window half-size
    int T = (uint8_t)L.template_size;
    const int tT = (int)((5.0 / (L.resolution * (1 << bl))) / 2.0) * 2 + 1;
    if (T < tT) T = (uint8_t)tT;
    int half = (int)(T / 2.0);
    if (bblunder) {
        if (L.level == 4) half = T - iteration * 2;
        else if (L.level == 3) half = T - 2 - iteration * 2;
        if (half < (int)(T / 2.0)) half = (int)(T / 2.0);
    }
    const long n = j->ncells;
    CallBuffers CB(j);
    BlunderArgs B{};
    B.W = (long)L.grid_w;
    B.H = (long)L.grid_h;
    B.sw = (long)j->bortho_w;
    B.sh = (long)j->bortho_h;
    B.b0 = L.boundary[0];
    B.b1 = L.boundary[1];
    B.im_res = L.resolution * (double)(1 << bl);
    B.grid_res = L.grid_resolution;
    B.half = half;
    B.level_b = bl;
    B.right = L.reference_id == 1 ? 0 : 1;
    for (int ti = 0; ti < 2; ti++) {
        const LevelView &v = j->views[ti][bl];
        B.view[ti] = DevView{v.texel, v.ori, v.w, v.h, v.sx, v.sy};
    }
    B.grid_pts = j->d_grid_pts;
    B.ortho_wgs = j->d_bortho_wgs;
    B.gridpt3 = j->d_gridpt3;
    int r = CB.get(&B.cd[0], (size_t)(B.sw * B.sh) * 2);
    if (!r) r = CB.get(&B.cd[1], (size_t)(B.sw * B.sh) * 2);
    // the output stays on the device for smx_blunder_detection_tin
    if (!r && !j->d_bncc) r = jalloc(j, &j->d_bncc, (size_t)std::max(n, 1L));
    if (!r && !j->d_params) r = jalloc(j, &j->d_params, 1);
    if (r) return r;
    j->bncc_valid = false;
    B.out = j->d_bncc;
    fill_params(j, iteration);
    if ((r = put_params(j, j->stream))) return r;
    HIPCHK(hipEventRecord(j->ev0, j->stream));
    HIPCHK(launch_vll_blunder(j->d_params, B, j->stream));
    HIPCHK(hipEventRecord(j->ev1, j->stream));
    if (ortho_ncc) HIPCHK(hipMemcpyAsync(ortho_ncc, B.out, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, j->ev0, j->ev1));
    j->last_ms = ms;
    j->bncc_valid = true;
    j->generation++;   // GridPT3's ortho NCC changed
    return 0;
}

}  // extern "C"

namespace {
// setsm_code.cpp:9183-9418, the serial part of blunder_detection_TIN: the
// edge height differences of the triangles whose vertices all exist, summed
// in triangle order as the reference sums them, then the height-difference
// threshold, the level's ortho-NCC thresholds and whether the neighbour rule
// runs (:9489-9507).
struct BdThresholds {
    double height_th, ortho_ncc_th, th_ref_ncc;
    int check_match;
};

void bd_thresholds(const smx_level &L, const smx_bd_args &a, const smx_point3 *pts, long npts, const smx_tri *tris,
                   long ntris, BdThresholds &D)
{
    const int step = L.level, iteration = a.iteration;
    const uint32_t hdiffcount = (uint32_t)a.hinterval;
    std::vector<uint32_t> hdiffbin((size_t)hdiffcount + 1, 0);
    double sum_data = 0.0, sum_data2 = 0.0;
    long dh_count = 0;
    for (long t = 0; t < ntris; t++) {
        const smx_tri &tr = tris[t];
        if ((long)tr.a >= npts || (long)tr.b >= npts || (long)tr.c >= npts) continue;
        const double z0 = pts[tr.a].z, z1 = pts[tr.b].z, z2 = pts[tr.c].z;
        const double d[3] = {std::fabs(z0 - z1), std::fabs(z0 - z2), std::fabs(z1 - z2)};
        for (double v : d) {
            sum_data += v;
            sum_data2 += v * v;
            dh_count++;
        }
        for (double v : d) {
            const double b = (uint32_t)v;
            if (b < hdiffcount) hdiffbin[(uint32_t)b]++;
        }
    }
    double ortho_ncc_th;
    if (step == 4) ortho_ncc_th = 0.6 - (iteration - 1) * 0.01;
    else if (step >= 3) ortho_ncc_th = 0.5 - (iteration - 1) * 0.01;
    else if (step == 2) ortho_ncc_th = 0.4 - (iteration - 1) * 0.01;
    else if (step == 1) ortho_ncc_th = 0.3;
    else ortho_ncc_th = 0.2;
    double th_ref_ncc = 0.1 + (iteration - 1) * 0.05;
    if (th_ref_ncc > ortho_ncc_th) th_ref_ncc = ortho_ncc_th;
    bool check_dh = false;
    if (step <= 4 && step >= 3) check_dh = true;
    else if (step == 2) check_dh = iteration <= 4;
    else if (step <= 1) check_dh = iteration <= 2;
    // GSD only reaches the step >= 3 branch (the reference's pwrtwo(step - 1)
    // at step 0 is never read: the other branch sets the threshold anew)
    double GSD = L.grid_resolution;
    if (step >= 3 && L.grid_resolution < (double)(1 << (step - 1))) GSD = (double)(1 << (step - 1));
    const double height_th_1 = a.hinterval;
    double height_th_2 = GSD * 10;
    if (L.is_ra == 1) height_th_2 = GSD * 30;
    double height_th = height_th_1 < height_th_2 ? height_th_1 : height_th_2;
    if (step >= 3) {
        double up_3sigma = height_th;
        if (check_dh) {
            const double mean = sum_data / dh_count;
            const double sd = std::sqrt((sum_data2 - (sum_data) * (sum_data) / dh_count) / dh_count);
            up_3sigma = mean + 3 * sd;
        }
        if (up_3sigma < height_th) height_th = up_3sigma < height_th / 2.0 ? height_th / 2.0 : up_3sigma;
    } else {
        uint32_t total_dh = 0, th80 = 1000;
        long tcnt = 0;
        bool found = false;
        while (tcnt < (long)hdiffcount && !found) {
            total_dh += hdiffbin[tcnt];
            const double per = (double)total_dh / (double)dh_count * 100.0;
            if (per > 99) {
                th80 = (uint32_t)tcnt;
                found = true;
            }
            tcnt++;
        }
        if (!found) th80 = (uint32_t)(tcnt - 1);
        const double mean = sum_data / dh_count;
        const double sd = std::sqrt((sum_data2 - (sum_data) * (sum_data) / dh_count) / dh_count);
        const double up_3sigma = mean + 3 * sd;
        height_th = up_3sigma < th80 ? up_3sigma : (double)th80;
        if (height_th > 50) height_th = 50;
    }
    const unsigned count_bl = (uint16_t)a.count_bl;
    int check_match = 0;
    if (step > 1) check_match = 1;
    else if (step >= 0) check_match = iteration == 1 && count_bl <= 5;
    D.height_th = height_th;
    D.ortho_ncc_th = ortho_ncc_th;
    D.th_ref_ncc = th_ref_ncc;
    D.check_match = check_match;
}
}  // namespace

extern "C" {

int smx_blunder_detection_tin(smx_job *j, const smx_bd_args *a, const float *ortho_ncc, smx_point3 *pts, long npts,
                              const smx_tri *tris, long ntris, uint8_t *detected, long blunder_count[2],
                              double *minz_mp, double *maxz_mp)
{
    if (!j || !a || !blunder_count || (npts > 0 && !pts) || (ntris > 0 && !tris) || npts < 0 || ntris < 0)
        return fail(SMX_E_ARG, "smx_blunder_detection_tin: bad argument");
    if (!(a->hinterval >= 0 && a->hinterval <= 1e7)) return fail(SMX_E_ARG, "Hinterval %g out of range", a->hinterval);
    if (ntris >= 0xffffffffL) return fail(SMX_E_ARG, "too many triangles");
    if (!j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    if (!ortho_ncc && !j->bncc_valid)
        return fail(SMX_E_STATE, "no ortho NCC given and no smx_vertical_line_locus_blunder output on this level");
    if (detected && npts > 0) std::memset(detected, 0, (size_t)npts);
    const smx_level &L = j->lv;
    if (L.level == 0 && a->iteration == 3) return 0;   // setsm_code.cpp:9164
    // the serial pass over the triangles runs on its own thread while the
    // inputs go up and the incidence lists build
    BdThresholds th{};
    struct Joiner {
        std::thread t;
        ~Joiner()
        {
            if (t.joinable()) t.join();
        }
    } stats;
    try {
        stats.t = std::thread([&] { bd_thresholds(L, *a, pts, npts, tris, ntris, th); });
    } catch (...) {
        bd_thresholds(L, *a, pts, npts, tris, ntris, th);   // no thread to be had: the same pass in line
    }
    HIPCHK(hipSetDevice(j->device));
    const long n = j->ncells;
    CallBuffers B(j);
    BdDev D{};
    smx_point3 *d_pts = nullptr;
    smx_tri *d_tris = nullptr;
    int64_t *d_off = nullptr;
    float *d_on = nullptr;
    int r = B.get(&d_pts, (size_t)npts);
    if (!r) r = B.get(&d_tris, (size_t)ntris);
    if (!r) r = B.get(&D.deg, (size_t)npts + 1);
    if (!r) r = B.get(&d_off, (size_t)npts + 1);
    if (!r) r = B.get(&D.adj, (size_t)ntris * 3);
    if (!r) r = B.get(&D.akey, (size_t)n);
    if (!r) r = B.get(&D.state, (size_t)(npts + 3) / 4);
    if (!r) r = B.get(&D.err, 1);
    if (!r && ortho_ncc) r = B.get(&d_on, (size_t)n);
    if (!r) r = grow(j, j->js.scan_part, j->js.scan_part_cap, scan_partials(npts + 1));   // more points than cells
    if (r) return r;
    if (npts + 1 > scan_max_n()) return fail(SMX_E_ARG, "too many points");
    if (npts) HIPCHK(hipMemcpyAsync(d_pts, pts, (size_t)npts * sizeof(smx_point3), hipMemcpyHostToDevice, j->stream));
    if (ntris) HIPCHK(hipMemcpyAsync(d_tris, tris, (size_t)ntris * sizeof(smx_tri), hipMemcpyHostToDevice, j->stream));
    if (ortho_ncc) HIPCHK(hipMemcpyAsync(d_on, ortho_ncc, (size_t)n * sizeof(float), hipMemcpyHostToDevice, j->stream));
    D.pts = d_pts;
    D.npts = npts;
    D.tris = d_tris;
    D.ntris = ntris;
    D.off = d_off;
    D.oncc = ortho_ncc ? d_on : j->d_bncc;
    D.g = j->d_gridpt3;
    D.W = (long)L.grid_w;
    D.H = (long)L.grid_h;
    D.gres = L.grid_resolution;
    D.b0 = L.boundary[0];
    D.b1 = L.boundary[1];
    D.level = L.level;
    D.iteration = a->iteration;
    D.is_ra = L.is_ra;
    D.flag_blunder = a->flag_blunder != 0;
    HIPCHK(hipEventRecord(j->ev0, j->stream));
    HIPCHK(launch_bd_prepare(D, j->js, j->stream));
    HIPCHK(hipEventRecord(j->ev1, j->stream));
    HIPCHK(hipMemcpyAsync(&j->h_small[0], D.err, sizeof(int), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    if (stats.t.joinable()) stats.t.join();
    D.height_th = th.height_th;
    D.ortho_ncc_th = th.ortho_ncc_th;
    D.th_ref_ncc = th.th_ref_ncc;
    D.check_match = th.check_match;
    int err = 0;
    std::memcpy(&err, &j->h_small[0], sizeof err);
    float ms = 0, ms2 = 0;
    HIPCHK(hipEventElapsedTime(&ms, j->ev0, j->ev1));
    if (err) return fail(SMX_E_ARG, "a matched point lies outside the level's grid");
    std::vector<uint8_t> st((size_t)((npts + 3) / 4) * 4);
    HIPCHK(hipEventRecord(j->ev0, j->stream));
    HIPCHK(launch_bd_points(D, j->stream));
    HIPCHK(hipEventRecord(j->ev1, j->stream));
    if (npts) HIPCHK(hipMemcpyAsync(st.data(), D.state, st.size(), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    HIPCHK(hipEventElapsedTime(&ms2, j->ev0, j->ev1));
    j->last_ms = ms + ms2;   // kernels only: the host threshold pass and the copies excluded
    // setsm_code.cpp:9121-9156 and :9758-9776: a point still at flag 0 that
    // its own test or a neighbour's blames becomes a new blunder; flags 1 and
    // 3 count as blunders (3 ends as 1); the rest are kept points. In chunks
    // on host threads: each chunk's first lowest / highest kept height, then
    // the chunks in order, which is what the reference's one pass keeps.
    struct Part {
        long kept = 0, blunders = 0;
        double lo = INFINITY, hi = -INFINITY;
    };
    const long nth = std::max(1L, std::min<long>({16L, (long)std::thread::hardware_concurrency(), npts >> 16}));
    std::vector<Part> part((size_t)nth);
    auto walk = [&](long k) {
        Part &q = part[(size_t)k];
        const long i0 = npts * k / nth, i1 = npts * (k + 1) / nth;
        for (long i = i0; i < i1; i++) {
            const uint8_t f0 = pts[i].flag;
            if (f0 == 0 && st[(size_t)i]) {
                pts[i].flag = 1;
                if (detected) detected[i] = 1;
                q.blunders++;
            } else if (f0 == 1 || f0 == 3) {
                pts[i].flag = 1;
                q.blunders++;
            } else {
                q.kept++;
                if (q.lo > pts[i].z) q.lo = pts[i].z;
                if (q.hi < pts[i].z) q.hi = pts[i].z;
            }
        }
    };
    {
        std::vector<std::thread> pool;
        long k = 1;
        try {
            for (; k < nth; k++) pool.emplace_back(walk, k);
        } catch (...) {
            for (; k < nth; k++) walk(k);   // chunks no thread took, in line
        }
        walk(0);
        for (std::thread &t : pool) t.join();
    }
    for (const Part &q : part) {
        blunder_count[0] += q.kept;
        blunder_count[1] += q.blunders;
        if (minz_mp && *minz_mp > q.lo) *minz_mp = q.lo;
        if (maxz_mp && *maxz_mp < q.hi) *maxz_mp = q.hi;
    }
    j->generation++;   // anchor flags
    return 0;
}

// ------------------------------------------------------------ graphs

int smx_graph_destroy(smx_graph *g)
{
    if (!g) return 0;
    if (!g->jobs.empty()) (void)hipSetDevice(g->jobs[0]->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->desc) (void)hipFree(g->desc);
    if (g->ev0) (void)hipEventDestroy(g->ev0);
    if (g->ev1) (void)hipEventDestroy(g->ev1);
    delete g;
    return 0;
}

// Queue the captured sequence on s (called between begin and end capture).
static int graph_body(smx_graph *g, const std::vector<VllPlan> &P, std::vector<CellArgs> &As,
                      std::vector<OrthoArgs> &Os, std::vector<VllScratch> &Ss, int with_awncc)
{
    const int n = (int)g->jobs.size();
    smx_job *const *jobs = g->jobs.data();
    const hipStream_t s = g->stream;
    for (int i = 0; i < n; i++) {
        int r = vll_front(jobs[i], P[i], Ss[i], false, s, As[i], Os[i]);
        if (r) return r;
    }
    if (P[0].fast && n > 1) {
        int r = launch_vll_fast_batch(jobs, n, As.data(), Ss.data(), P[0].half, g->desc, s, nullptr, nullptr);
        if (r) return r;
        for (int i = 0; i < n; i++) HIPCHK(launch_vll_redo(P[i].half, jobs[i]->d_params, As[i], s));
    } else {
        for (int i = 0; i < n; i++) {
            int r = vll_kernels(jobs[i], P[i], As[i], Ss[i], s, false);
            if (r) return r;
        }
    }
    for (int i = 0; i < n; i++) HIPCHK(launch_vll_finalize(As[i], s));
    if (with_awncc)
        for (int i = 0; i < n; i++) {
            int r = awncc_enqueue(jobs[i], s);
            if (r) return r;
        }
    return 0;
}

int smx_graph_create(smx_graph **out, smx_job **jobs, int n, int iteration, int precision, int with_awncc)
{
    if (!out || !jobs || n < 1) return fail(SMX_E_ARG, "smx_graph_create: bad argument");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!jobs[i]) return fail(SMX_E_ARG, "null job %d", i);
        if (jobs[i]->device != jobs[0]->device) return fail(SMX_E_ARG, "jobs of one graph share a device");
        for (int k = 0; k < i; k++)
            if (jobs[k] == jobs[i]) return fail(SMX_E_ARG, "job %d appears twice", i);
    }
    std::vector<VllPlan> P((size_t)n);
    for (int i = 0; i < n; i++) {
        int r = vll_plan(jobs[i], iteration, precision, 0, P[i]);
        if (r) return r;
        if (P[i].ra)
            return fail(SMX_E_UNSUPPORTED, "RA and matching-rate passes read a count back mid-call and cannot be "
                                           "captured");
        if (P[i].fast && n > 1 && P[i].half != P[0].half)
            return fail(SMX_E_ARG, "a batched graph holds one window size (job %d has half %d, job 0 %d)", i,
                        P[i].half, P[0].half);
    }
    HIPCHK(hipSetDevice(jobs[0]->device));
    for (int i = 0; i < n; i++) {
        int r = vll_reserve(jobs[i], P[i]);
        if (!r && with_awncc) r = awncc_reserve(jobs[i]);
        if (r) return r;
    }
    smx_graph *g = new smx_graph();
    g->jobs.assign(jobs, jobs + n);
    g->stream = jobs[0]->stream;
    int r = 0;
    hipError_t e = hipEventCreate(&g->ev0);
    if (e == hipSuccess) e = hipEventCreate(&g->ev1);
    if (e == hipSuccess && P[0].fast && n > 1) e = hipMalloc(&g->desc, fast_batch_desc_bytes(n));
    for (int i = 0; i < n && e == hipSuccess; i++) e = hipStreamSynchronize(jobs[i]->stream);
    if (e != hipSuccess) {
        smx_graph_destroy(g);
        return fail(SMX_E_HIP, "smx_graph_create: %s", hipGetErrorString(e));
    }
    std::vector<CellArgs> As((size_t)n);
    std::vector<OrthoArgs> Os((size_t)n);
    std::vector<VllScratch> Ss((size_t)n);
    for (int i = 0; i < n && !r; i++) {
        Ss[i] = jobs[i]->js;
        vll_args(jobs[i], P[i], Ss[i], As[i], Os[i]);
        r = put_params(jobs[i], g->stream);
        g->params.push_back(jobs[i]->hp);
        g->fast.push_back(P[i].fast ? 1 : 0);
    }
    // descriptors of a batched launch live outside the graph (uploaded here)
    if (!r && P[0].fast && n > 1) r = fast_batch_descriptors(jobs, n, As.data(), Ss.data(), g->desc, g->stream);
    if (!r && (e = hipStreamSynchronize(g->stream)) != hipSuccess) r = fail(SMX_E_HIP, "%s", hipGetErrorString(e));
    if (r) {
        smx_graph_destroy(g);
        return r;
    }
    e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        smx_graph_destroy(g);
        return fail(SMX_E_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    r = graph_body(g, P, As, Os, Ss, with_awncc);
    const std::string body_err = r ? g_err : std::string();
    e = hipStreamEndCapture(g->stream, &g->graph);
    if (!r && e != hipSuccess) r = fail(SMX_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    if (!r && (e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess)
        r = fail(SMX_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    if (r) {
        if (!body_err.empty()) g_err = body_err;
        smx_graph_destroy(g);
        return r;
    }
    for (int i = 0; i < n; i++) g->gens.push_back(jobs[i]->generation);
    *out = g;
    return 0;
}

int smx_graph_launch(smx_graph *g, long long *accessible)
{
    if (!g || !g->exec) return fail(SMX_E_ARG, "smx_graph_launch: no graph");
    const int n = (int)g->jobs.size();
    for (int i = 0; i < n; i++)
        if (g->jobs[i]->generation != g->gens[i])
            return fail(SMX_E_STATE, "job %d changed since the capture (level, search space, cube, buffers or launch "
                                     "shape): capture again", i);
    HIPCHK(hipSetDevice(g->jobs[0]->device));
    for (int i = 0; i < n; i++) {
        g->jobs[i]->hp = g->params[i];
        int r = put_params(g->jobs[i], g->stream);
        if (r) return r;
    }
    HIPCHK(hipEventRecord(g->ev0, g->stream));
    HIPCHK(hipGraphLaunch(g->exec, g->stream));
    HIPCHK(hipEventRecord(g->ev1, g->stream));
    for (int i = 0; i < n; i++) {
        const long long a = vll_finish(g->jobs[i], g->jobs[i]->js.counters, g->stream);
        if (a < 0) return (int)a;
        if (accessible) accessible[i] = a;
    }
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    g->last_ms = ms;
    for (int i = 0; i < n; i++) {
        g->jobs[i]->last_ms = ms;
        g->jobs[i]->last_redo_ms = 0;
    }
    return 0;
}

double smx_graph_last_ms(const smx_graph *g) { return g ? g->last_ms : 0.0; }

long long smx_job_num_hypotheses(const smx_job *j) { return j ? j->vox_total : 0; }

int smx_job_get_voxels(smx_job *j, uint32_t *words, float *rho, int64_t *offsets)
{
    if (!j || !j->level_set) return fail(SMX_E_STATE, "no level set");
    HIPCHK(hipSetDevice(j->device));
    if (words && j->vox_total)
        HIPCHK(hipMemcpyAsync(words, j->d_vox, (size_t)j->vox_total * 4, hipMemcpyDeviceToHost, j->stream));
    if (rho && j->vox_total) {
        if (j->d_rho)
            HIPCHK(hipMemcpyAsync(rho, j->d_rho, (size_t)j->vox_total * 4, hipMemcpyDeviceToHost, j->stream));
        else
            for (long long k = 0; k < j->vox_total; k++) rho[k] = NAN;
    }
    if (offsets)
        HIPCHK(hipMemcpyAsync(offsets, j->d_vox_off, ((size_t)j->ncells + 1) * 8, hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    return 0;
}

int smx_job_get_voxel_rows(smx_job *j, long long cell0, long long cell1, uint32_t *words, float *rho)
{
    if (!j || !j->level_set) return fail(SMX_E_STATE, "no level set");
    if (cell0 < 0 || cell1 < cell0 || cell1 > j->ncells) return fail(SMX_E_ARG, "bad cell range");
    HIPCHK(hipSetDevice(j->device));
    HIPCHK(hipMemcpyAsync(j->h_small, j->d_vox_off + cell0, sizeof(int64_t), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipMemcpyAsync(j->h_small + 1, j->d_vox_off + cell1, sizeof(int64_t), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    const long long a = j->h_small[0], n = j->h_small[1] - j->h_small[0];
    if (n > 0 && words)
        HIPCHK(hipMemcpyAsync(words, j->d_vox + a, (size_t)n * 4, hipMemcpyDeviceToHost, j->stream));
    if (n > 0 && rho) {
        if (j->d_rho)
            HIPCHK(hipMemcpyAsync(rho, j->d_rho + a, (size_t)n * 4, hipMemcpyDeviceToHost, j->stream));
        else
            for (long long k = 0; k < n; k++) rho[k] = NAN;
    }
    HIPCHK(hipStreamSynchronize(j->stream));
    return 0;
}

int smx_job_set_voxels(smx_job *j, const uint32_t *words)
{
    if (!j || !words || !j->level_set) return fail(SMX_E_STATE, "no level set");
    HIPCHK(hipSetDevice(j->device));
    // Rebuild the CSR offsets from the NCCresult array on the device, so a
    // cube can be handed over without this job's own InitializeVoxel.
    std::vector<smx_nccresult> ncc((size_t)j->ncells);
    std::vector<int64_t> off((size_t)j->ncells + 1);
    HIPCHK(hipMemcpyAsync(ncc.data(), j->d_ncc, ncc.size() * sizeof(smx_nccresult), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    off[0] = 0;
    long long runs[3] = {0, 0, 0};
    int max_nh = 0;
    for (long c = 0; c < j->ncells; c++) {
        off[c + 1] = off[c] + ncc[c].NumOfHeight;
        for (int k = 0; k < 3; k++) runs[k] += (ncc[c].NumOfHeight + kRunLens[k] - 1) / kRunLens[k];
        max_nh = std::max(max_nh, (int)ncc[c].NumOfHeight);
    }
    const long long total = off[j->ncells];
    if (total != j->vox_total || !j->d_vox) {
        jfree(j, j->d_vox);
        int r = jalloc(j, &j->d_vox, (size_t)std::max(total, 1LL));
        if (r) return r;
        j->vox_total = total;
    }
    jfree(j, j->d_rho);   // no mean NCC for a handed-over cube
    HIPCHK(hipMemcpyAsync(j->d_vox_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice, j->stream));
    if (total)
        HIPCHK(hipMemcpyAsync(j->d_vox, words, (size_t)total * 4, hipMemcpyHostToDevice, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    for (int k = 0; k < 3; k++) j->runs_bound[k] = runs[k];
    j->max_nh = max_nh;
    j->stats_valid = true;
    j->generation++;
    return 0;
}

int smx_fast_num_configs(void) { return fast_num_configs(); }

int smx_job_set_fast_config(smx_job *j, int config)
{
    if (!j || config < -1 || config >= fast_num_configs()) return fail(SMX_E_ARG, "fast config %d out of range", config);
    if (j->fast_cfg != config) j->generation++;
    j->fast_cfg = config;
    return 0;
}

int smx_fast_config_desc(int config, char *buf, int buflen)
{
    if (!buf || fast_config_desc(config, buf, buflen) < 0) return fail(SMX_E_ARG, "fast config %d out of range", config);
    return 0;
}

double smx_job_last_kernel_ms(const smx_job *j) { return j ? j->last_ms : 0.0; }

int smx_job_last_stats(const smx_job *j, long long *out4)
{
    if (!j || !out4) return fail(SMX_E_ARG, "null argument");
    out4[0] = j->last_evaluated;
    out4[1] = j->last_tiles_lds;
    out4[2] = j->last_waves_hbm;
    out4[3] = (long long)j->fast_cfg;
    return 0;
}
int smx_job_last_redo(const smx_job *j, long long *runs, long long *hypotheses, double *ms)
{
    if (!j || !runs || !hypotheses || !ms) return fail(SMX_E_ARG, "null argument");
    *runs = j->last_redo_runs;
    *hypotheses = j->last_redo_hyps;
    *ms = j->last_redo_ms;
    return 0;
}
long long smx_job_last_evaluated(const smx_job *j) { return j ? j->last_evaluated : 0; }

int smx_job_sync(smx_job *j)
{
    if (!j) return fail(SMX_E_ARG, "null job");
    HIPCHK(hipStreamSynchronize(j->stream));
    return 0;
}

}  // extern "C"
