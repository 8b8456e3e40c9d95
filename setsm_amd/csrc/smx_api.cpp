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
    return 0;
}

void job_free(smx_job *j, void *p)
{
    if (!p) return;
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

int job_scan(smx_job *j, const int64_t *in, int64_t *out, long n)
{
    size_t need = 0;
    HIPCHK(scan_i64(in, out, n, nullptr, need, j->stream));
    if (need > j->scan_tmp_bytes) {
        jfree(j, j->scan_tmp);
        int r = jalloc(j, (char **)&j->scan_tmp, need);
        if (r) return r;
        j->scan_tmp_bytes = need;
    }
    size_t have = j->scan_tmp_bytes;
    HIPCHK(scan_i64(in, out, n, j->scan_tmp, have, j->stream));
    return 0;
}

int fast_scratch(smx_job *j, int slot, long ntiles, long nruns)
{
    int64_t *&count = slot ? j->d_tile2_count : j->d_tile_count;
    int64_t *&off = slot ? j->d_tile2_off : j->d_tile_off;
    int2 *&runs = slot ? j->d_tile2_runs : j->d_tile_runs;
    long &tcap = slot ? j->tiles2_cap : j->tiles_cap;
    long &rcap = slot ? j->tile2_runs_cap : j->tile_runs_cap;
    if (ntiles + 1 > tcap) {
        jfree(j, count);
        jfree(j, off);
        int r = jalloc(j, &count, (size_t)ntiles + 1);
        if (!r) r = jalloc(j, &off, (size_t)ntiles + 1);
        if (r) return r;
        tcap = ntiles + 1;
    }
    if (nruns > rcap) {
        jfree(j, runs);
        int r = jalloc(j, &runs, (size_t)nruns);
        if (r) return r;
        rcap = nruns;
    }
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
    jfree(j, j->d_runs);
    jfree(j, j->d_max_bits);
    jfree(j, j->d_action);
    jfree(j, j->d_new_len);
    jfree(j, j->d_new_off);
    jfree(j, j->d_tile_count);
    jfree(j, j->d_tile_off);
    jfree(j, j->d_tile_runs);
    jfree(j, j->d_tile2_count);
    jfree(j, j->d_tile2_off);
    jfree(j, j->d_tile2_runs);
    j->tiles2_cap = j->tile2_runs_cap = 0;
    jfree(j, j->d_sum);
    j->sum_total = 0;
    jfree(j, j->d_grid_pts);
    jfree(j, j->d_ortho_wgs);
    jfree(j, j->d_cd);
    jfree(j, j->d_cdn);
    jfree(j, j->d_gval);
    jfree(j, j->d_peak_rho);
    j->peak_cap = 0;
    jfree(j, j->d_redo);
    j->redo_cap = 0;
    jfree(j, j->d_part);
    j->part_cap = 0;
    j->cd_bytes = j->cdn_bytes = 0;
    j->tiles_cap = j->tile_runs_cap = 0;
    j->runs_cap = 0;
    j->vox_total = 0;
    j->level_set = false;
    j->searchspace_set = false;
}

// Fill the per-call VllParams and copy it to the device.
static int upload_params(smx_job *j, int iteration)
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
    HIPCHK(hipMemcpyAsync(j->d_params, &P, sizeof(VllParams), hipMemcpyHostToDevice, j->stream));
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

}  // namespace smx

using namespace smx;

extern "C" {

const char *smx_last_error(void) { return g_err.c_str(); }
int smx_version(void) { return 10000; }

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
    if (e != hipSuccess) {
        delete j;
        return fail(SMX_E_HIP, "job setup: %s", hipGetErrorString(e));
    }
    int r = jalloc(j, &j->d_params, 1);
    if (!r) r = jalloc(j, &j->d_counters, 8);
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
    jfree(j, j->scan_tmp);
    jfree(j, j->d_sgm_scratch);
    job_free(j, j->d_batch);
    j->d_batch = nullptr;
    jfree(j, j->d_params);
    jfree(j, j->d_counters);
    for (auto &s : j->sizes) (void)hipFree(s.first);
    if (j->h_small) (void)hipHostFree(j->h_small);
    if (j->ev0) (void)hipEventDestroy(j->ev0);
    if (j->ev1) (void)hipEventDestroy(j->ev1);
    if (j->ev2) (void)hipEventDestroy(j->ev2);
    if (j->stream) (void)hipStreamDestroy(j->stream);
    delete j;
    return 0;
}

long long smx_job_device_bytes(const smx_job *j) { return j ? j->device_bytes : 0; }

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
    HIPCHK(hipMemcpyAsync(d_img, intensity, n * 2, hipMemcpyHostToDevice, j->stream));
    HIPCHK(hipMemcpyAsync(d_mag, magnitude, n * 2, hipMemcpyHostToDevice, j->stream));
    HIPCHK(hipMemcpyAsync(v.ori, orientation, n, hipMemcpyHostToDevice, j->stream));
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
    HIPCHK(hipMemcpyAsync(cur, level0, n0 * 2, hipMemcpyHostToDevice, j->stream));
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
    HIPCHK(hipSetDevice(j->device));
    HIPCHK(hipStreamSynchronize(j->stream));
    free_level_state(j);
    j->lv = *p;
    const long n = (long)p->grid_w * (long)p->grid_h;
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
    if (r) return r;
    HIPCHK(hipMemcpyAsync(j->d_grid_wgs, p->grid_wgs, (size_t)n * 16, hipMemcpyHostToDevice, j->stream));
    if (p->grid_pts) {
        r = jalloc(j, &j->d_grid_pts, (size_t)n * 2);
        if (r) return r;
        HIPCHK(hipMemcpyAsync(j->d_grid_pts, p->grid_pts, (size_t)n * 16, hipMemcpyHostToDevice, j->stream));
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
    HIPCHK(hipStreamSynchronize(j->stream));
    j->level_set = true;
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
    HIPCHK(hipMemcpyAsync(j->d_ortho_wgs, lonlat, (size_t)sw * sh * 16, hipMemcpyHostToDevice, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    j->lv.ortho_w = w;
    j->lv.ortho_h = h;
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
    return 0;
}

int smx_initialize_voxel(smx_job *j, int iteration)
{
    if (!j || !j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    if (iteration < 1) return fail(SMX_E_ARG, "iteration starts at 1");
    HIPCHK(hipSetDevice(j->device));
    int r = upload_params(j, iteration);
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
    r = job_scan(j, j->d_new_len, j->d_new_off, j->ncells + 1);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(j->h_small, j->d_new_off + j->ncells, sizeof(int64_t), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    const long long total = j->h_small[0];
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
    return 0;
}

}  // extern "C"

namespace smx {

// VerticalLineLocus up to the NCC kernel: checks, parameters, gate, run list.
static int vll_prepare(smx_job *j, int iteration, int precision, int keep_rho, CellArgs &A)
{
    if (!j || !j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    if (precision != SMX_PREC_FP64 && precision != SMX_PREC_FAST) return fail(SMX_E_ARG, "unknown precision %d", precision);
    const smx_level &L = j->lv;
    // setsm_code.cpp:6100-6112 check_ortho, 6240 the per-cell GNCC gate
    bool check_ortho = !((L.level == 4 && iteration == 1) || L.is_ra);
    if (L.pre_dem && !(L.level == 4 && iteration == 1)) check_ortho = true;
    if (L.check_matchtag) check_ortho = true;
    if (L.level <= L.sgm_py) check_ortho = false;
    const double ortho_th = 0.7 - (4 - L.level) * 0.10;
    const int slot = 1 - L.reference_id;
    if (slot < 0 || slot > 1) return fail(SMX_E_ARG, "reference_id must be 0 or 1");
    const bool gncc_on = check_ortho && (double)j->max_ortho_ncc[slot] / 1000.0 > ortho_th;
    const bool combined_wncc = !L.is_ra && (L.level > L.py_combined_level);
    if (gncc_on) {
        if (!j->d_ortho_wgs || !j->d_grid_pts)
            return fail(SMX_E_ARG, "the GNCC term is active (ortho_ncc > %.2f) but the level has no ortho_wgs / "
                                   "grid_pts", ortho_th);
        if (combined_wncc)
            for (int ti = 0; ti < 2; ti++)
                if (!j->views[ti][L.level - 1].texel)
                    return fail(SMX_E_STATE, "the combined GNCC term needs level %d", L.level - 1);
    }
    const int half = j->hp.half;
    if (half < 5 || half > 7) return fail(SMX_E_UNSUPPORTED, "Half_template_size %d not built (5..7 are)", half);
    HIPCHK(hipSetDevice(j->device));
    int r = upload_params(j, iteration);
    if (r) return r;
    if (j->hp.combined_incc)
        for (int ti = 0; ti < 2; ti++)
            if (!j->views[ti][L.level - 1].texel) return fail(SMX_E_STATE, "combined INCC needs level %d", L.level - 1);
    if (keep_rho && !j->d_rho && j->vox_total > 0) {
        r = jalloc(j, &j->d_rho, (size_t)j->vox_total);
        if (r) return r;
        HIPCHK(hipMemsetAsync(j->d_rho, 0xff, (size_t)j->vox_total * sizeof(float), j->stream));   // NaN
    }
    A = CellArgs{};
    A.ncells = j->ncells;
    A.gridpt3 = j->d_gridpt3;
    A.ncc = j->d_ncc;
    A.grid_wgs = j->d_grid_wgs;
    A.vox_off = j->d_vox_off;
    A.vox = j->d_vox;
    A.rho = keep_rho ? j->d_rho : nullptr;
    A.run_count = j->d_run_count;
    A.run_offset = j->d_run_off;
    A.max_bits = j->d_max_bits;
    A.accessible = j->d_counters;
    A.evaluated = j->d_counters + 1;
    A.ncc_alpha = j->hp.ncc_alpha;
    OrthoArgs O{};
    if (gncc_on) {
        // setsm_code.cpp:6121-6129: ortho image coordinates of this iteration's
        // surface (GridPT3.Height)
        const double im_res = L.resolution * (double)(1 << L.level);
        const double im_res_next = combined_wncc ? L.resolution * (double)(1 << (L.level - 1)) : L.resolution;
        O.W = L.grid_w;
        O.H = L.grid_h;
        O.sw = L.ortho_w;
        O.sh = L.ortho_h;
        if (combined_wncc) {
            O.swn = (long)((L.boundary[2] - L.boundary[0]) / im_res_next) + 1;
            O.shn = (long)((L.boundary[3] - L.boundary[1]) / im_res_next) + 1;
        }
        O.b0 = L.boundary[0];
        O.b1 = L.boundary[1];
        O.im_res = im_res;
        O.im_res_next = im_res_next;
        O.grid_res = L.grid_resolution;
        O.combined = combined_wncc ? 1 : 0;
        O.slot = slot;
        O.ortho_th = ortho_th;
        O.grid_pts = j->d_grid_pts;
        O.ortho_wgs = j->d_ortho_wgs;
        O.gridpt3 = j->d_gridpt3;
        const size_t cdb = (size_t)O.sw * O.sh * 2 * 2, cdnb = (size_t)O.swn * O.shn * 2 * 2;
        if (cdb > j->cd_bytes) {
            jfree(j, j->d_cd);
            int r2 = jalloc(j, &j->d_cd, cdb);
            if (r2) return r2;
            j->cd_bytes = cdb;
        }
        if (cdnb > j->cdn_bytes) {
            jfree(j, j->d_cdn);
            int r2 = jalloc(j, &j->d_cdn, cdnb);
            if (r2) return r2;
            j->cdn_bytes = cdnb;
        }
        if (!j->d_gval) {
            int r2 = jalloc(j, &j->d_gval, (size_t)j->ncells);
            if (r2) return r2;
        }
        O.cd[0] = j->d_cd;
        O.cd[1] = j->d_cd + (size_t)O.sw * O.sh * 2;
        O.cdn[0] = cdnb ? j->d_cdn : nullptr;
        O.cdn[1] = cdnb ? j->d_cdn + (size_t)O.swn * O.shn * 2 : nullptr;
        // calloc'd in the reference: pixels it skips read as (0, 0)
        HIPCHK(hipMemsetAsync(j->d_cd, 0, cdb * sizeof(double), j->stream));
        if (cdnb) HIPCHK(hipMemsetAsync(j->d_cdn, 0, cdnb * sizeof(double), j->stream));
        HIPCHK(launch_ortho_coords(j->d_params, O, j->stream));
        A.acc_flag = j->d_action;
    }
    HIPCHK(hipMemsetAsync(j->d_counters, 0, 8 * sizeof(unsigned long long), j->stream));
    HIPCHK(launch_vll_gate(j->d_params, A, j->stream));
    if (gncc_on) {
        HIPCHK(launch_gncc(half, j->d_params, O, A, j->d_gval, j->stream));
        A.gval = j->d_gval;
    }
    r = job_scan(j, j->d_run_count, j->d_run_off, j->ncells + 1);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(j->h_small, j->d_run_off + j->ncells, sizeof(int64_t), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    const long nruns = (long)j->h_small[0];
    if (nruns > j->runs_cap) {
        jfree(j, j->d_runs);
        r = jalloc(j, &j->d_runs, (size_t)nruns);
        if (r) return r;
        j->runs_cap = nruns;
    }
    A.runs = j->d_runs;
    A.nruns = nruns;
    HIPCHK(launch_vll_fill_runs(A, j->stream));
    A.redo_n = j->d_counters + 4;
    if (precision != SMX_PREC_FP64 && !L.is_ra && !L.check_matching_rate) {
        // FAST: runs handed to the FP64 pass (each at most once)
        if (nruns > j->redo_cap) {
            jfree(j, j->d_redo);
            r = jalloc(j, &j->d_redo, (size_t)std::max(nruns, 1L));
            if (r) return r;
            j->redo_cap = nruns;
        }
        A.redo = j->d_redo;
        if (j->hp.combined_incc) {
            // the level term per hypothesis between the two passes
            if (j->vox_total > j->part_cap || !j->d_part) {
                jfree(j, j->d_part);
                r = jalloc(j, &j->d_part, (size_t)std::max(j->vox_total, 1LL));
                if (r) return r;
                j->part_cap = j->vox_total;
            }
            A.part = j->d_part;
        }
    }
    if (L.is_ra || L.check_matching_rate) {
        // RA / matching-rate pass: temp_rho per run lane for FindPeakNcc
        if (nruns > j->peak_cap) {
            jfree(j, j->d_peak_rho);
            r = jalloc(j, &j->d_peak_rho, (size_t)std::max(nruns, 1L) * kRun);
            if (r) return r;
            j->peak_cap = nruns;
        }
        A.peak_rho = j->d_peak_rho;
    }
    return 0;
}

// After the NCC kernel: max_WNCC, counters. Returns Accessable_grid.
static long long vll_finish(smx_job *j, const CellArgs &A)
{
    HIPCHK(launch_vll_finalize(A, j->stream));
    HIPCHK(hipMemcpyAsync(j->h_small, j->d_counters, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    j->last_evaluated = j->h_small[1] + j->h_small[5];
    j->last_redo_hyps = j->h_small[5];
    j->last_tiles_lds = j->h_small[2];
    j->last_waves_hbm = j->h_small[3];
    j->last_redo_runs = j->h_small[4];
    return (long long)j->h_small[0];
}

}  // namespace smx

extern "C" {

long long smx_vertical_line_locus(smx_job *j, int iteration, int precision, int keep_rho)
{
    CellArgs A;
    int r = vll_prepare(j, iteration, precision, keep_rho, A);
    if (r) return r;
    if (precision == SMX_PREC_FP64 || A.peak_rho) {
        // RA / matching-rate passes (no voxel cube) always take the FP64 kernel
        HIPCHK(hipEventRecord(j->ev0, j->stream));
        HIPCHK(launch_vll_fp64(j->hp.half, j->d_params, A, j->stream));
        HIPCHK(hipEventRecord(j->ev1, j->stream));
        if (A.peak_rho && j->lv.check_matching_rate) HIPCHK(launch_find_peak(j->d_params, A, j->stream));
    } else {
        r = launch_vll_fast(j, j->hp.half, A, j->ev0, j->ev1);
        if (r) return r;
        HIPCHK(launch_vll_redo(j->hp.half, j->d_params, A, j->stream));
        HIPCHK(hipEventRecord(j->ev2, j->stream));
    }
    const long long acc = vll_finish(j, A);
    if (acc < 0) return acc;
    float ms = 0, rms = 0;
    HIPCHK(hipEventElapsedTime(&ms, j->ev0, j->ev1));
    if (precision != SMX_PREC_FP64 && !A.peak_rho) HIPCHK(hipEventElapsedTime(&rms, j->ev1, j->ev2));
    j->last_ms = ms;
    j->last_redo_ms = rms;
    return acc;
}

int smx_vertical_line_locus_batch(smx_job **jobs, int n, int iteration, int precision, int keep_rho,
                                  long long *accessible)
{
    if (!jobs || n < 1 || !accessible) return fail(SMX_E_ARG, "smx_vertical_line_locus_batch: bad argument");
    for (int i = 0; i < n; i++) {
        if (!jobs[i]) return fail(SMX_E_ARG, "null job %d", i);
        if (jobs[i]->device != jobs[0]->device) return fail(SMX_E_ARG, "jobs of one batch share a device");
    }
    bool ra = false;
    for (int i = 0; i < n; i++) ra = ra || jobs[i]->lv.is_ra || jobs[i]->lv.check_matching_rate;
    if (precision == SMX_PREC_FP64 || ra) {
        // the plain kernel: one tile per launch (RA / matching-rate passes included)
        for (int i = 0; i < n; i++) {
            const long long a = smx_vertical_line_locus(jobs[i], iteration, precision, keep_rho);
            if (a < 0) return (int)a;
            accessible[i] = a;
        }
        return 0;
    }
    std::vector<CellArgs> As((size_t)n);
    for (int i = 0; i < n; i++) {
        int r = vll_prepare(jobs[i], iteration, precision, keep_rho, As[i]);
        if (r) return r;
        if (jobs[i]->hp.half != jobs[0]->hp.half)
            return fail(SMX_E_ARG, "a batch holds one window size (job %d has half %d, job 0 %d)", i,
                        jobs[i]->hp.half, jobs[0]->hp.half);
        if (jobs[i]->hp.combined_incc)
            return fail(SMX_E_UNSUPPORTED, "combined-level iterations are launched per job");
    }
    smx_job *j0 = jobs[0];
    int r = launch_vll_fast_batch(jobs, n, As.data(), j0->hp.half, j0->ev0, j0->ev1);
    if (r) return r;
    for (int i = 0; i < n; i++) HIPCHK(launch_vll_redo(j0->hp.half, jobs[i]->d_params, As[i], j0->stream));
    HIPCHK(hipEventRecord(j0->ev2, j0->stream));
    HIPCHK(hipStreamSynchronize(j0->stream));
    float ms = 0, rms = 0;
    HIPCHK(hipEventElapsedTime(&ms, j0->ev0, j0->ev1));
    HIPCHK(hipEventElapsedTime(&rms, j0->ev1, j0->ev2));
    for (int i = 0; i < n; i++) {
        const long long a = vll_finish(jobs[i], As[i]);
        if (a < 0) return (int)a;
        accessible[i] = a;
        jobs[i]->last_ms = ms;
        jobs[i]->last_redo_ms = rms;
    }
    return 0;
}

// setsm_code.cpp:7017-7509: SumCost of the current cube by the reference's
// 12 path loops (levels <= SGM_py).
static int sgm_aggregate(smx_job *j)
{
    const long long total = j->vox_total;
    if (total != j->sum_total || !j->d_sum) {
        jfree(j, j->d_sum);
        int r = jalloc(j, &j->d_sum, (size_t)std::max(total, 1LL));
        if (r) return r;
        j->sum_total = total;
    }
    if (total) HIPCHK(hipMemsetAsync(j->d_sum, 0, (size_t)total * sizeof(float), j->stream));
    // the longest height list sizes the per-path buffers
    HIPCHK(hipMemsetAsync(j->d_counters + 2, 0, sizeof(unsigned long long), j->stream));
    HIPCHK(launch_max_nh(j->d_ncc, j->ncells, j->d_counters + 2, j->stream));
    HIPCHK(hipMemcpyAsync(j->h_small, j->d_counters + 2, sizeof(int64_t), hipMemcpyDeviceToHost, j->stream));
    HIPCHK(hipStreamSynchronize(j->stream));
    const int cap = (int)((std::max<int64_t>(j->h_small[0], 1) + 3) & ~3LL);
    SgmLaunch L{};
    L.W = (int)j->lv.grid_w;
    L.H = (int)j->lv.grid_h;
    L.step = j->lv.height_step;
    L.ncc = j->d_ncc;
    L.gridpt3 = j->d_gridpt3;
    L.vox_off = j->d_vox_off;
    L.vox = j->d_vox;
    L.sum = j->d_sum;
    L.cap = cap;
    if ((size_t)10 * cap > sgm_lds_limit()) {
        const size_t need = (size_t)10 * cap * (size_t)std::max(L.W, L.H);
        if (need > j->sgm_scratch_bytes) {
            jfree(j, j->d_sgm_scratch);
            int r = jalloc(j, &j->d_sgm_scratch, need);
            if (r) return r;
            j->sgm_scratch_bytes = need;
        }
        L.gscratch = j->d_sgm_scratch;
    }
    HIPCHK(launch_sgm(L, j->stream));
    return 0;
}

int smx_awncc(smx_job *j, int iteration)
{
    if (!j || !j->level_set || !j->searchspace_set) return fail(SMX_E_STATE, "level and search space must be set");
    HIPCHK(hipSetDevice(j->device));
    int r = upload_params(j, iteration);
    if (r) return r;
    const bool sgm = j->lv.level <= j->lv.sgm_py;
    if (sgm) {
        r = sgm_aggregate(j);
        if (r) return r;
    }
    PeakArgs A{j->ncells, j->d_ncc, j->d_gridpt3, j->d_vox_off, j->d_vox, sgm ? j->d_sum : nullptr};
    HIPCHK(launch_awncc(j->d_params, A, j->stream));
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
    for (long c = 0; c < j->ncells; c++) off[c + 1] = off[c] + ncc[c].NumOfHeight;
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
    return 0;
}

int smx_fast_num_configs(void) { return fast_num_configs(); }

int smx_job_set_fast_config(smx_job *j, int config)
{
    if (!j || config < -1 || config >= fast_num_configs()) return fail(SMX_E_ARG, "fast config %d out of range", config);
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
