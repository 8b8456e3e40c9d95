// setsm_vll_shim.cpp — SETSM's InitializeVoxel / VerticalLineLocus / AWNCC
// (setsm_code.hpp:75, 81, 91), the SelectMPs / SetHeightRange either side of
// its TIN (setsm_code.hpp:99, 122) and DecisionMPs' VerticalLineLocus_blunder /
// SetHeightRange_blunder / blunder_detection_TIN (setsm_code.hpp:95, 120,
// 116) with their exact C++ signatures and SETSM's own types, running on the
// MI355X through libsmx (include/smx_vll.h).
//
// Drop-in use inside SETSM's build: add this file, link -lsmx, and either
//   * compile it with -DSMX_SHIM_GLOBAL_NAMES and remove the original
//     definitions from setsm_code.cpp, or
//   * keep the originals and call smx_setsm::InitializeVoxel / ... at the
//     call sites in Matching_SETSM (setsm_code.cpp:3827, 3829, 3844, 3862,
//     4192-4222) and DecisionMPs (:8961-8967, 9081-9087, 9113-9115).
//
// State: one smx job per thread mirrors SETSM's per-level state. A new job is
// made when the level's image buffers, sizes or origins change; a changed
// level record (grid, RPCs, ImageAdjust, ...) restarts the level on the
// device and re-uploads the host cube. The cube stays in HBM between the
// calls of a level, and every call hands the host arrays back exactly as the
// CPU path leaves them (grid_voxel rows re-allocated like InitializeVoxel
// does, VOXEL bytes identical to the device cube). Call
// smx_setsm::Release() when a tile is done; a later tile whose buffers land
// at the same addresses is also told apart by a sample of its pixels.
//
// Errors follow SETSM's convention (setsm_code.cpp:6571-6575): a message on
// stdout and exit(1). Inputs that need a mode libsmx lacks (frame sensors,
// more than two images, affine image adjustment) are such errors.
// The GNCC term's ortho grid is computed here with SETSM's own
// ps2wgs_single (CoordConversion.cpp), as SetOrthoImageCoord does.
//
// SMX_PRECISION=fp64 (read at every VerticalLineLocus call) selects the FP64
// kernel, bit-identical to the CPU path; the default is the FAST kernel
// (per-hypothesis mean NCC within 1e-5 of FP64).
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "CoordConversion.hpp"   // SETSM's records (Typedefine.hpp) and ps2wgs_single
#include "smx_vll.h"

static_assert(sizeof(UGRID) == sizeof(smx_ugrid), "UGRID layout");
static_assert(sizeof(NCCresult) == sizeof(smx_nccresult), "NCCresult layout");
static_assert(sizeof(VOXEL) == 4, "VOXEL layout");
static_assert(sizeof(D2DPOINT) == 2 * sizeof(double), "D2DPOINT layout");
static_assert(sizeof(D3DPOINT) == sizeof(smx_point3) && offsetof(D3DPOINT, flag) == offsetof(smx_point3, flag),
              "D3DPOINT layout");
static_assert(sizeof(UI3DPOINT) == sizeof(smx_tri), "UI3DPOINT layout");

namespace smx_setsm {

namespace {

struct State {
    smx_job *job = nullptr;
    int level = -1;
    const void *views[2][6] = {};   // the pointers and geometry the job's views came from
    double starts[2][4] = {};
    int sizes[2][4] = {};
    uint64_t pixels = 0;            // a sample of the views' pixels (see view_sample)
    smx_level lv;                   // the level record last passed to smx_job_set_level
    bool lv_set = false;
    bool ortho_set = false;         // ortho grid lon/lat uploaded for this level record
    double minmax[2] = {0, 0};      // minmaxHeight of the level's last InitializeVoxel / VerticalLineLocus
    int blunder_level = -1;         // blunder level whose views and ortho grid the job holds
    const void *bviews[2][2] = {};
    double bstarts[2][2] = {};
};

thread_local State g_state;

void die(const char *what, int rc)
{
    printf("ERROR: %s failed (%d): %s\n", what, rc, smx_last_error());
    exit(1);
}

void check(int rc, const char *what)
{
    if (rc < 0) die(what, rc);
}

void fill_rpc(smx_rpc &r, const double *const *rows)
{
    for (int k = 0; k < 5; k++) {
        r.off[k] = rows[0][k];
        r.scale[k] = rows[1][k];
    }
    for (int j = 0; j < 4; j++)
        for (int k = 0; k < 20; k++) r.coef[j][k] = rows[j + 2][k];
}

void require_pair(const ProInfo *proinfo)
{
    if (proinfo->number_of_images != 2 || !proinfo->check_selected_image[0] || !proinfo->check_selected_image[1]) {
        printf("ERROR: the MI355X VerticalLineLocus handles one stereo pair (2 selected images)\n");
        exit(1);
    }
    if (proinfo->sensor_type != SB) {
        printf("ERROR: the MI355X VerticalLineLocus handles RPC (SB) sensors only\n");
        exit(1);
    }
}

int precision_from_env()
{
    const char *pe = getenv("SMX_PRECISION");
    return (pe && !strcmp(pe, "fp64")) ? SMX_PREC_FP64 : SMX_PREC_FAST;
}

// Point the thread's job at this call's tile and level. A new tile (other
// image buffers, sizes or origins) gets a new job; a changed level record
// (level, grid, RPCs, ImageAdjust, ...) is passed to smx_job_set_level, which
// starts the level's device state afresh. Returns true when the device holds
// no cube for this level yet, so that the caller uploads the host's.
// FNV-1a over 256 evenly spaced elements of each view buffer: a later tile
// whose buffers land at the same addresses, sizes and origins (the caller
// skipped Release) still differs here, so its pixels are uploaded.
uint64_t view_sample(const void *const views[2][6], const int sizes[2][4])
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) {
        h ^= v;
        h *= 1099511628211ull;
    };
    for (int k = 0; k < 2; k++)
        for (int v = 0; v < 6; v++) {
            if (!views[k][v]) continue;
            const long n = (long)sizes[k][v < 3 ? 0 : 2] * (long)sizes[k][v < 3 ? 1 : 3];
            for (long i = 0; i < 256 && n > 0; i++) {
                const long at = i * n / 256;
                mix(v % 3 == 2 ? ((const uint8 *)views[k][v])[at] : ((const uint16 *)views[k][v])[at]);
            }
        }
    return h;
}

bool bind(const ProInfo *proinfo, LevelInfo &L)
{
    State &s = g_state;
    require_pair(proinfo);
    if (*L.NumOfIAparam != 2) {
        printf("ERROR: the MI355X VerticalLineLocus takes the 2-parameter (shift) image adjustment\n");
        exit(1);
    }
    const int level = *L.Pyramid_step;
    const bool with_next = level > *L.Py_combined_level && L.py_Images_next;
    const int ref = L.reference_id, tgt = ref == 0 ? 1 : 0;
    const int ids[2] = {ref, tgt};

    const void *views[2][6] = {};
    double starts[2][4] = {};
    int sizes[2][4] = {};
    for (int k = 0; k < 2; k++) {
        const int ti = ids[k];
        views[k][0] = L.py_Images[ti];
        views[k][1] = L.py_MagImages[ti];
        views[k][2] = L.py_OriImages[ti];
        starts[k][0] = L.py_Startpos[ti].m_X;
        starts[k][1] = L.py_Startpos[ti].m_Y;
        sizes[k][0] = L.py_Sizes[ti][level].width;
        sizes[k][1] = L.py_Sizes[ti][level].height;
        if (with_next) {
            views[k][3] = L.py_Images_next[ti];
            views[k][4] = L.py_MagImages_next[ti];
            views[k][5] = L.py_OriImages_next[ti];
            starts[k][2] = L.py_Startpos_next[ti].m_X;
            starts[k][3] = L.py_Startpos_next[ti].m_Y;
            sizes[k][2] = L.py_Sizes[ti][level - 1].width;
            sizes[k][3] = L.py_Sizes[ti][level - 1].height;
        }
    }
    const uint64_t pixels = view_sample(views, sizes);
    const bool same_views = s.job && s.level == level && !memcmp(views, s.views, sizeof views) &&
                            !memcmp(starts, s.starts, sizeof starts) && !memcmp(sizes, s.sizes, sizeof sizes) &&
                            pixels == s.pixels;
    if (!same_views) {
        if (s.job) smx_job_destroy(s.job);
        s = State{};
        // one level above for VerticalLineLocus_blunder's blunder_selected_level
        check(smx_job_create(&s.job, level + 2 > 8 ? 8 : level + 2), "smx_job_create");
        for (int k = 0; k < 2; k++) {
            check(smx_job_set_view(s.job, k, level, L.py_Images[ids[k]], L.py_MagImages[ids[k]],
                                   L.py_OriImages[ids[k]], sizes[k][0], sizes[k][1], starts[k][0], starts[k][1]),
                  "smx_job_set_view");
            if (with_next)
                check(smx_job_set_view(s.job, k, level - 1, L.py_Images_next[ids[k]], L.py_MagImages_next[ids[k]],
                                       L.py_OriImages_next[ids[k]], sizes[k][2], sizes[k][3], starts[k][2],
                                       starts[k][3]),
                      "smx_job_set_view(next)");
        }
        s.level = level;
        memcpy(s.views, views, sizeof views);
        memcpy(s.starts, starts, sizeof starts);
        memcpy(s.sizes, sizes, sizeof sizes);
        s.pixels = pixels;
    }

    smx_level lv;
    memset(&lv, 0, sizeof lv);
    for (int k = 0; k < 2; k++) {
        fill_rpc(lv.rpc[k], L.RPCs[ids[k]]);
        lv.adjust[k][0] = L.ImageAdjust[ids[k]][0];
        lv.adjust[k][1] = L.ImageAdjust[ids[k]][1];
    }
    lv.level = level;
    lv.template_size = *L.Template_size;
    lv.py_combined_level = *L.Py_combined_level;
    lv.sgm_py = proinfo->SGM_py;
    lv.is_ra = proinfo->IsRA;
    lv.check_matchtag = proinfo->check_Matchtag;
    lv.pre_dem = proinfo->pre_DEMtif;
    lv.check_matching_rate = L.check_matching_rate ? *L.check_matching_rate : 0;
    lv.resolution = proinfo->resolution;
    lv.dem_resolution = proinfo->DEM_resolution;
    lv.bin_angle = *L.bin_angle;
    lv.height_step = *L.height_step;
    lv.grid_resolution = *L.grid_resolution;
    for (int k = 0; k < 4; k++) lv.boundary[k] = L.Boundary[k];
    lv.grid_w = L.Size_Grid2D->width;
    lv.grid_h = L.Size_Grid2D->height;
    lv.grid_pts = (const double *)L.GridPts;
    lv.grid_wgs = (const double *)L.Grid_wgs;
    lv.reference_id = ref;
    if (s.lv_set && !memcmp(&lv, &s.lv, sizeof lv)) return false;
    check(smx_job_set_level(s.job, &lv), "smx_job_set_level");
    s.lv = lv;
    s.lv_set = true;
    s.ortho_set = false;
    s.blunder_level = -1;
    return true;
}

// VerticalLineLocus_blunder's inputs: the blunder level's images
// (py_BImages / py_BMagImages at py_BStartpos; the NCC reads no orientation,
// so a zero one is uploaded unless the level is the matching level itself)
// and its ortho grid's lon/lat from SETSM's ps2wgs_single.
void ensure_blunder(State &s, const ProInfo *proinfo, LevelInfo &L)
{
    const int B = *L.blunder_selected_level;
    const int ref = L.reference_id, ids[2] = {ref, ref == 0 ? 1 : 0};
    const void *bv[2][2];
    double bs[2][2];
    for (int k = 0; k < 2; k++) {
        bv[k][0] = L.py_BImages[ids[k]];
        bv[k][1] = L.py_BMagImages[ids[k]];
        bs[k][0] = L.py_BStartpos[ids[k]].m_X;
        bs[k][1] = L.py_BStartpos[ids[k]].m_Y;
    }
    if (s.blunder_level == B && !memcmp(bv, s.bviews, sizeof bv) && !memcmp(bs, s.bstarts, sizeof bs)) return;
    if (B < 0 || B >= 8) {
        printf("ERROR: blunder_selected_level %d out of range\n", B);
        exit(1);
    }
    for (int k = 0; k < 2; k++) {
        if (B == s.level && bv[k][0] == (const void *)L.py_Images[ids[k]]) continue;   // the level's own views
        const int w = L.py_Sizes[ids[k]][B].width, h = L.py_Sizes[ids[k]][B].height;
        std::vector<uint8> zero_ori((size_t)w * h, 0);
        check(smx_job_set_view(s.job, k, B, L.py_BImages[ids[k]], L.py_BMagImages[ids[k]], zero_ori.data(), w, h,
                               bs[k][0], bs[k][1]),
              "smx_job_set_view(blunder level)");
    }
    const double im_res = proinfo->resolution * (double)(1 << B);
    const long sw = (long)((L.Boundary[2] - L.Boundary[0]) / im_res) + 1;
    const long sh = (long)((L.Boundary[3] - L.Boundary[1]) / im_res) + 1;
    std::vector<double> lonlat((size_t)sw * sh * 2);
#pragma omp parallel for schedule(guided)
    for (long it = 0; it < sw * sh; it++) {
        D2DPOINT p;
        p.m_X = L.Boundary[0] + (it % sw) * im_res;
        p.m_Y = L.Boundary[1] + (it / sw) * im_res;
        const D2DPOINT w = ps2wgs_single(*L.param, p);
        lonlat[2 * it] = w.m_X;
        lonlat[2 * it + 1] = w.m_Y;
    }
    check(smx_job_set_blunder_level(s.job, B, lonlat.data(), (uint32_t)sw, (uint32_t)sh), "smx_job_set_blunder_level");
    s.blunder_level = B;
    memcpy(s.bviews, bv, sizeof bv);
    memcpy(s.bstarts, bs, sizeof bs);
}

// setsm_code.cpp:6100-6129, 6240: when some cell takes the GNCC term, hand
// the level's ortho grid to the job as SETSM's SetOrthoImageCoord sees it:
// ps2wgs_single of every ortho pixel (SB sensor).
void ensure_ortho(State &s, const ProInfo *proinfo, LevelInfo &L, const UGRID *G, int iteration)
{
    if (s.ortho_set) return;
    const int level = *L.Pyramid_step;
    bool check_ortho = !((level == 4 && iteration == 1) || proinfo->IsRA);
    if (proinfo->pre_DEMtif && !(level == 4 && iteration == 1)) check_ortho = true;
    if (proinfo->check_Matchtag) check_ortho = true;
    if (level <= proinfo->SGM_py) check_ortho = false;
    if (!check_ortho) return;
    const double ortho_th = 0.7 - (4 - level) * 0.10;
    const int ti = L.reference_id == 0 ? 1 : 0;
    const long n = *L.Grid_length;
    bool any = false;
    for (long c = 0; c < n && !any; c++) any = (double)(G[c].ortho_ncc[ti]) / 1000.0 > ortho_th;
    if (!any) return;
    const double im_res = proinfo->resolution * (double)(1 << level);
    const long sw = (long)((L.Boundary[2] - L.Boundary[0]) / im_res) + 1;
    const long sh = (long)((L.Boundary[3] - L.Boundary[1]) / im_res) + 1;
    std::vector<double> lonlat((size_t)sw * sh * 2);
#pragma omp parallel for schedule(guided)
    for (long it = 0; it < sw * sh; it++) {
        D2DPOINT p;
        p.m_X = L.Boundary[0] + (it % sw) * im_res;
        p.m_Y = L.Boundary[1] + (it / sw) * im_res;
        const D2DPOINT w = ps2wgs_single(*L.param, p);
        lonlat[2 * it] = w.m_X;
        lonlat[2 * it + 1] = w.m_Y;
    }
    check(smx_job_set_ortho_wgs(s.job, lonlat.data(), (uint32_t)sw, (uint32_t)sh), "smx_job_set_ortho_wgs");
    s.ortho_set = true;
}

// Upload the host cube (rows of NumOfHeight VOXELs) after a fresh level set.
void upload_rows(State &s, VOXEL *const *grid_voxel, const NCCresult *ncc, long n)
{
    std::vector<uint32_t> words;
    for (long c = 0; c < n; c++)
        for (int h = 0; h < ncc[c].NumOfHeight; h++) {
            uint32_t w;
            memcpy(&w, &grid_voxel[c][h], sizeof w);
            words.push_back(w);
        }
    if (!words.empty()) check(smx_job_set_voxels(s.job, words.data()), "smx_job_set_voxels");
}

// Copy the device cube into SETSM's per-cell VOXEL rows.
void download_rows(State &s, VOXEL **grid_voxel, const NCCresult *ncc, long n)
{
    const long long total = smx_job_num_hypotheses(s.job);
    std::vector<uint32_t> words((size_t)(total > 0 ? total : 1));
    std::vector<int64_t> off((size_t)n + 1);
    check(smx_job_get_voxels(s.job, words.data(), nullptr, off.data()), "smx_job_get_voxels");
    for (long c = 0; c < n; c++)
        if (ncc[c].NumOfHeight > 0 && grid_voxel[c])
            memcpy(grid_voxel[c], words.data() + off[c], sizeof(VOXEL) * ncc[c].NumOfHeight);
}

}  // namespace

void InitializeVoxel(const ProInfo *proinfo, VOXEL **grid_voxel, LevelInfo &plevelinfo, UGRID *GridPT3,
                     NCCresult *nccresult, const int iteration, const double *minmaxHeight)
{
    const bool fresh = bind(proinfo, plevelinfo);
    State &s = g_state;
    const long n = *plevelinfo.Grid_length;
    s.minmax[0] = minmaxHeight[0];
    s.minmax[1] = minmaxHeight[1];
    std::vector<unsigned short> old_n((size_t)n);
    for (long c = 0; c < n; c++) old_n[c] = nccresult[c].NumOfHeight;
    check(smx_job_set_searchspace(s.job, (const smx_ugrid *)GridPT3, minmaxHeight), "smx_job_set_searchspace");
    check(smx_job_set_nccresult(s.job, (const smx_nccresult *)nccresult), "smx_job_set_nccresult");
    if (fresh) upload_rows(s, grid_voxel, nccresult, n);
    check(smx_initialize_voxel(s.job, iteration), "smx_initialize_voxel");
    check(smx_job_get_nccresult(s.job, (smx_nccresult *)nccresult), "smx_job_get_nccresult");
    check(smx_job_get_searchspace(s.job, (smx_ugrid *)GridPT3), "smx_job_get_searchspace");
    // setsm_code.cpp:5930-5948: rows of changed cells are re-allocated with
    // flag_cal = true, INCC = -1000; rows of dropped cells are released.
    for (long c = 0; c < n; c++) {
        const NCCresult &r = nccresult[c];
        if (r.check_height_change && r.NumOfHeight > 0) {
            if (old_n[c] > 0 && grid_voxel[c]) free(grid_voxel[c]);
            grid_voxel[c] = (VOXEL *)calloc(sizeof(VOXEL), r.NumOfHeight);
            for (int h = 0; h < r.NumOfHeight; h++) {
                grid_voxel[c][h].flag_cal = true;
                grid_voxel[c][h].INCC = (short)(-1.0 * 1000.0);
            }
        } else if (r.NumOfHeight == 0 && old_n[c] > 0 && grid_voxel[c]) {
            free(grid_voxel[c]);
            grid_voxel[c] = nullptr;
        }
    }
}

int VerticalLineLocus(VOXEL **grid_voxel, const ProInfo *proinfo, NCCresult *nccresult, LevelInfo &plevelinfo,
                      const UGRID *GridPT3, const uint8 iteration, const double *minmaxHeight)
{
    const bool fresh = bind(proinfo, plevelinfo);
    State &s = g_state;
    const long n = *plevelinfo.Grid_length;
    s.minmax[0] = minmaxHeight[0];
    s.minmax[1] = minmaxHeight[1];
    check(smx_job_set_searchspace(s.job, (const smx_ugrid *)GridPT3, minmaxHeight), "smx_job_set_searchspace");
    check(smx_job_set_nccresult(s.job, (const smx_nccresult *)nccresult), "smx_job_set_nccresult");
    if (fresh) upload_rows(s, grid_voxel, nccresult, n);
    ensure_ortho(s, proinfo, plevelinfo, GridPT3, iteration);
    const long long acc = smx_vertical_line_locus(s.job, iteration, precision_from_env(), 0);
    if (acc < 0) die("smx_vertical_line_locus", (int)acc);
    check(smx_job_get_nccresult(s.job, (smx_nccresult *)nccresult), "smx_job_get_nccresult");
    // RA and matching-rate passes leave grid_voxel alone (setsm_code.cpp:6470)
    const bool cmr = plevelinfo.check_matching_rate && *plevelinfo.check_matching_rate;
    if (!proinfo->IsRA && !cmr) download_rows(s, grid_voxel, nccresult, n);
    return (int)acc;
}

void AWNCC(ProInfo *proinfo, VOXEL **grid_voxel, CSize Size_Grid2D, UGRID *GridPT3, NCCresult *nccresult,
           double step_height, uint8 Pyramid_step, uint8 iteration, int MaxNumberofHeightVoxel)
{
    (void)grid_voxel;
    (void)MaxNumberofHeightVoxel;
    State &s = g_state;
    if (!s.job || !s.lv_set || s.level != Pyramid_step || s.lv.grid_w != (unsigned)Size_Grid2D.width ||
        s.lv.grid_h != (unsigned)Size_Grid2D.height) {
        printf("ERROR: AWNCC called without the level's VerticalLineLocus on the MI355X\n");
        exit(1);
    }
    (void)step_height;
    (void)GridPT3;
    check(smx_job_set_nccresult(s.job, (const smx_nccresult *)nccresult), "smx_job_set_nccresult");
    check(smx_awncc(s.job, iteration), "smx_awncc");
    check(smx_job_get_nccresult(s.job, (smx_nccresult *)nccresult), "smx_job_get_nccresult");
    (void)proinfo;
}

long SelectMPs(const ProInfo *proinfo, LevelInfo &rlevelinfo, const NCCresult *roh_height, UGRID *GridPT3,
               const double Th_roh, const double Th_roh_min, const double Th_roh_start, const double Th_roh_next,
               const int iteration, const double MPP, const int final_level_iteration, const double MPP_stereo_angle,
               std::vector<D3DPOINT> *linkedlist)
{
    bind(proinfo, rlevelinfo);
    State &s = g_state;
    const long n = *rlevelinfo.Grid_length;
    check(smx_job_set_searchspace(s.job, (const smx_ugrid *)GridPT3, s.minmax), "smx_job_set_searchspace");
    check(smx_job_set_nccresult(s.job, (const smx_nccresult *)roh_height), "smx_job_set_nccresult");
    smx_mps_args a;
    a.th_roh = Th_roh;
    a.th_roh_min = Th_roh_min;
    a.th_roh_start = Th_roh_start;
    a.th_roh_next = Th_roh_next;
    a.mpp = MPP;
    a.mpp_stereo_angle = MPP_stereo_angle;
    a.iteration = iteration;
    a.final_level_iteration = final_level_iteration;
    std::vector<smx_point3> pts((size_t)(n > 0 ? n : 1));
    long count = 0;
    double th = 0;
    check(smx_select_mps(s.job, &a, pts.data(), (long)pts.size(), &count, &th), "smx_select_mps");
    printf("minimum TH %f\n", th);
    check(smx_job_get_searchspace(s.job, (smx_ugrid *)GridPT3), "smx_job_get_searchspace");
    for (long i = 0; i < count; i++) linkedlist->push_back(D3DPOINT(pts[i].x, pts[i].y, pts[i].z, pts[i].flag));
    return count;
}

// Like the reference, frees GridPT3 and returns the new grid (calloc'd).
UGRID *SetHeightRange(ProInfo *proinfo, LevelInfo &rlevelinfo, NCCresult *nccresult, const int numOfPts,
                      const int num_triangles, UGRID *GridPT3, const int iteration, double *minH_grid,
                      double *maxH_grid, D3DPOINT *pts, const UI3DPOINT *tris, const double MPP,
                      const bool level_check_matching_rate)
{
    (void)nccresult;
    bind(proinfo, rlevelinfo);
    State &s = g_state;
    const long n = *rlevelinfo.Grid_length;
    check(smx_job_set_searchspace(s.job, (const smx_ugrid *)GridPT3, s.minmax), "smx_job_set_searchspace");
    smx_hr_args a;
    a.mpp = MPP;
    a.seed_dem_sigma = proinfo->seedDEMsigma;
    a.iteration = iteration;
    a.check_matching_rate = level_check_matching_rate ? 1 : 0;
    check(smx_set_height_range(s.job, (const smx_point3 *)pts, numOfPts, (const smx_tri *)tris, num_triangles, &a,
                               minH_grid, maxH_grid),
          "smx_set_height_range");
    UGRID *result = (UGRID *)calloc(sizeof(UGRID), (size_t)n);
    check(smx_job_get_searchspace(s.job, (smx_ugrid *)result), "smx_job_get_searchspace");
    free(GridPT3);
    return result;
}

// SetHeightRange_blunder takes no ProInfo: it runs on the job of the level
// the thread's last call bound, which must be this LevelInfo's.
bool SetHeightRange_blunder(LevelInfo &rlevelinfo, const D3DPOINT *pts, const int numPts, UI3DPOINT *tris,
                            const long num_triangles, UGRID *GridPT3)
{
    State &s = g_state;
    if (!s.job || !s.lv_set || s.level != *rlevelinfo.Pyramid_step ||
        s.lv.grid_w != (unsigned)rlevelinfo.Size_Grid2D->width ||
        s.lv.grid_h != (unsigned)rlevelinfo.Size_Grid2D->height) {
        printf("ERROR: SetHeightRange_blunder called without the level's VerticalLineLocus on the MI355X\n");
        exit(1);
    }
    check(smx_job_set_searchspace(s.job, (const smx_ugrid *)GridPT3, s.minmax), "smx_job_set_searchspace");
    check(smx_set_height_range_blunder(s.job, (const smx_point3 *)pts, numPts, (const smx_tri *)tris, num_triangles),
          "smx_set_height_range_blunder");
    check(smx_job_get_searchspace(s.job, (smx_ugrid *)GridPT3), "smx_job_get_searchspace");
    return true;
}

bool VerticalLineLocus_blunder(const ProInfo *proinfo, LevelInfo &rlevelinfo, float *nccresult, UGRID *GridPT3,
                               uint8 iteration, bool bblunder)
{
    bind(proinfo, rlevelinfo);
    State &s = g_state;
    ensure_blunder(s, proinfo, rlevelinfo);
    check(smx_job_set_searchspace(s.job, (const smx_ugrid *)GridPT3, s.minmax), "smx_job_set_searchspace");
    check(smx_vertical_line_locus_blunder(s.job, iteration, bblunder ? 1 : 0, nccresult),
          "smx_vertical_line_locus_blunder");
    check(smx_job_get_searchspace(s.job, (smx_ugrid *)GridPT3), "smx_job_get_searchspace");
    return true;
}

// blunder_detection_TIN: on the level bound as for VerticalLineLocus_blunder,
// which DecisionMPs calls just before it (setsm_code.cpp:8962-8967).
bool blunder_detection_TIN(const ProInfo *proinfo, LevelInfo &rlevelinfo, const int iteration, float *ortho_ncc,
                           bool flag_blunder, uint16 count_bl, D3DPOINT *pts, bool *detectedBlunders,
                           long int num_points, UI3DPOINT *tris, long int num_triangles, UGRID *Gridpts,
                           long *blunder_count, double *minz_mp, double *maxz_mp)
{
    bind(proinfo, rlevelinfo);
    State &s = g_state;
    smx_bd_args a{};
    a.hinterval = *rlevelinfo.Hinterval;
    a.iteration = iteration;
    a.flag_blunder = flag_blunder ? 1 : 0;
    a.count_bl = count_bl;
    std::vector<uint8_t> det((size_t)(num_points > 0 ? num_points : 1), 0);
    check(smx_job_set_searchspace(s.job, (const smx_ugrid *)Gridpts, s.minmax), "smx_job_set_searchspace");
    check(smx_blunder_detection_tin(s.job, &a, ortho_ncc, (smx_point3 *)pts, num_points, (const smx_tri *)tris,
                                    num_triangles, det.data(), blunder_count, minz_mp, maxz_mp),
          "smx_blunder_detection_tin");
    // the reference only ever sets entries of detectedBlunders
    for (long i = 0; i < num_points; i++)
        if (det[(size_t)i]) detectedBlunders[i] = true;
    check(smx_job_get_searchspace(s.job, (smx_ugrid *)Gridpts), "smx_job_get_searchspace");
    return true;
}

// Release the thread's job (optional; it is reused across calls).
void Release()
{
    if (g_state.job) smx_job_destroy(g_state.job);
    g_state = State{};
}

}  // namespace smx_setsm

#ifdef SMX_SHIM_GLOBAL_NAMES
void InitializeVoxel(const ProInfo *proinfo, VOXEL **grid_voxel, LevelInfo &plevelinfo, UGRID *GridPT3,
                     NCCresult *nccresult, const int iteration, const double *minmaxHeight)
{
    smx_setsm::InitializeVoxel(proinfo, grid_voxel, plevelinfo, GridPT3, nccresult, iteration, minmaxHeight);
}
int VerticalLineLocus(VOXEL **grid_voxel, const ProInfo *proinfo, NCCresult *nccresult, LevelInfo &plevelinfo,
                      const UGRID *GridPT3, const uint8 iteration, const double *minmaxHeight)
{
    return smx_setsm::VerticalLineLocus(grid_voxel, proinfo, nccresult, plevelinfo, GridPT3, iteration, minmaxHeight);
}
void AWNCC(ProInfo *proinfo, VOXEL **grid_voxel, CSize Size_Grid2D, UGRID *GridPT3, NCCresult *nccresult,
           double step_height, uint8 Pyramid_step, uint8 iteration, int MaxNumberofHeightVoxel)
{
    smx_setsm::AWNCC(proinfo, grid_voxel, Size_Grid2D, GridPT3, nccresult, step_height, Pyramid_step, iteration,
                     MaxNumberofHeightVoxel);
}
long SelectMPs(const ProInfo *proinfo, LevelInfo &rlevelinfo, const NCCresult *roh_height, UGRID *GridPT3,
               const double Th_roh, const double Th_roh_min, const double Th_roh_start, const double Th_roh_next,
               const int iteration, const double MPP, const int final_level_iteration, const double MPP_stereo_angle,
               std::vector<D3DPOINT> *linkedlist)
{
    return smx_setsm::SelectMPs(proinfo, rlevelinfo, roh_height, GridPT3, Th_roh, Th_roh_min, Th_roh_start,
                                Th_roh_next, iteration, MPP, final_level_iteration, MPP_stereo_angle, linkedlist);
}
UGRID *SetHeightRange(ProInfo *proinfo, LevelInfo &rlevelinfo, NCCresult *nccresult, const int numOfPts,
                      const int num_triangles, UGRID *GridPT3, const int iteration, double *minH_grid,
                      double *maxH_grid, D3DPOINT *pts, const UI3DPOINT *tris, const double MPP,
                      const bool level_check_matching_rate)
{
    return smx_setsm::SetHeightRange(proinfo, rlevelinfo, nccresult, numOfPts, num_triangles, GridPT3, iteration,
                                     minH_grid, maxH_grid, pts, tris, MPP, level_check_matching_rate);
}
bool SetHeightRange_blunder(LevelInfo &rlevelinfo, const D3DPOINT *pts, const int numPts, UI3DPOINT *tris,
                            const long num_triangles, UGRID *GridPT3)
{
    return smx_setsm::SetHeightRange_blunder(rlevelinfo, pts, numPts, tris, num_triangles, GridPT3);
}
bool VerticalLineLocus_blunder(const ProInfo *proinfo, LevelInfo &rlevelinfo, float *nccresult, UGRID *GridPT3,
                               uint8 iteration, bool bblunder)
{
    return smx_setsm::VerticalLineLocus_blunder(proinfo, rlevelinfo, nccresult, GridPT3, iteration, bblunder);
}
bool blunder_detection_TIN(const ProInfo *proinfo, LevelInfo &rlevelinfo, const int iteration, float *ortho_ncc,
                           bool flag_blunder, uint16 count_bl, D3DPOINT *pts, bool *detectedBlunders,
                           long int num_points, UI3DPOINT *tris, long int num_triangles, UGRID *Gridpts,
                           long *blunder_count, double *minz_mp, double *maxz_mp)
{
    return smx_setsm::blunder_detection_TIN(proinfo, rlevelinfo, iteration, ortho_ncc, flag_blunder, count_bl, pts,
                                            detectedBlunders, num_points, tris, num_triangles, Gridpts,
                                            blunder_count, minz_mp, maxz_mp);
}
#endif
