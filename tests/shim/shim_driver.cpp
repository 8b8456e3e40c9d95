// Test driver for setsm_amd/shim/setsm_vll_shim.cpp: builds SETSM's own
// ProInfo and LevelInfo (Typedefine.hpp:351-419, 604-643) from an smx_level
// and drives one pyramid level the way Matching_SETSM does
// (setsm_code.cpp:3827-3844): InitializeVoxel, VerticalLineLocus, AWNCC per
// iteration, on host arrays (VOXEL **grid_voxel, NCCresult *, UGRID *), and
// SelectMPs / SetHeightRange between iterations (setsm_code.cpp:3862,
// 4192-4222) and DecisionMPs' VerticalLineLocus_blunder /
// SetHeightRange_blunder / blunder_detection_TIN. The functions are called
// through their reference signatures (setsm_code.hpp:75, 81, 91, 95, 99, 116,
// 120, 122), which libsmx_setsm.so exports.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "CoordConversion.hpp"
#include "smx_types.h"

// Test double for SETSM's ps2wgs_single (CoordConversion.cpp:396), which the
// shim calls for the GNCC ortho grid: the synthetic scenes map metres to
// degrees with a local spherical mapping, lat = lat0 + N/R*180/pi and
// lon = lon0 + E/D*180/pi (D = R cos(lat0)), evaluated in the same operation
// order as the Python scene generator.
static double g_lat0, g_lon0, g_R = 1.0, g_D = 1.0;

D2DPOINT ps2wgs_single(const TransParam _param, const D2DPOINT _ps)
{
    (void)_param;
    D2DPOINT w;
    w.m_Y = g_lat0 + _ps.m_Y / g_R * 180.0 / M_PI;
    w.m_X = g_lon0 + _ps.m_X / g_D * 180.0 / M_PI;
    return w;
}

void InitializeVoxel(const ProInfo *proinfo, VOXEL **grid_voxel, LevelInfo &plevelinfo, UGRID *GridPT3,
                     NCCresult *nccresult, const int iteration, const double *minmaxHeight);
int VerticalLineLocus(VOXEL **grid_voxel, const ProInfo *proinfo, NCCresult *nccresult, LevelInfo &plevelinfo,
                      const UGRID *GridPT3, const uint8 iteration, const double *minmaxHeight);
void AWNCC(ProInfo *proinfo, VOXEL **grid_voxel, CSize Size_Grid2D, UGRID *GridPT3, NCCresult *nccresult,
           double step_height, uint8 Pyramid_step, uint8 iteration, int MaxNumberofHeightVoxel);
long SelectMPs(const ProInfo *proinfo, LevelInfo &rlevelinfo, const NCCresult *roh_height, UGRID *GridPT3,
               const double Th_roh, const double Th_roh_min, const double Th_roh_start, const double Th_roh_next,
               const int iteration, const double MPP, const int final_level_iteration, const double MPP_stereo_angle,
               std::vector<D3DPOINT> *linkedlist);
bool VerticalLineLocus_blunder(const ProInfo *proinfo, LevelInfo &rlevelinfo, float *nccresult, UGRID *GridPT3,
                               uint8 iteration, bool bblunder);
bool SetHeightRange_blunder(LevelInfo &rlevelinfo, const D3DPOINT *pts, const int numPts, UI3DPOINT *tris,
                            const long num_triangles, UGRID *GridPT3);
bool blunder_detection_TIN(const ProInfo *proinfo, LevelInfo &rlevelinfo, const int iteration, float *ortho_ncc,
                           bool flag_blunder, uint16 count_bl, D3DPOINT *pts, bool *detectedBlunders,
                           long int num_points, UI3DPOINT *tris, long int num_triangles, UGRID *Gridpts,
                           long *blunder_count, double *minz_mp, double *maxz_mp);
UGRID *SetHeightRange(ProInfo *proinfo, LevelInfo &rlevelinfo, NCCresult *nccresult, const int numOfPts,
                      const int num_triangles, UGRID *GridPT3, const int iteration, double *minH_grid,
                      double *maxH_grid, D3DPOINT *pts, const UI3DPOINT *tris, const double MPP,
                      const bool level_check_matching_rate);

namespace {

struct Driver {
    ProInfo pro;
    LevelInfo L;
    TransParam param;
    int level, combined, reference_id;
    unsigned char template_size;
    unsigned char num_ia = 2;
    double bin_angle, height_step, grid_resolution, boundary[4];
    CSize grid_size;
    long grid_length;
    bool check_matching_rate = false;
    const uint16 *img[2], *mag[2], *img_next[2], *mag_next[2];
    const uint8 *ori[2], *ori_next[2];
    D2DPOINT start[2], start_next[2];
    std::vector<CSize> sizes[2];
    const CSize *size_rows[2];
    double rpc_rows[2][7][20];
    const double *rpc_ptr[2][7];
    const double *const *rpcs[2];
    double adjust[2][2];
    const double *adjust_ptr[2];
    std::vector<NCCresult> ncc;
    std::vector<VOXEL *> vox;
    std::vector<NCCresult> ncc_vll;
    long long acc = 0;
    int blunder_level = 0;
    const uint16 *bimg[2] = {nullptr, nullptr}, *bmag[2] = {nullptr, nullptr};
    D2DPOINT bstart[2];
    double hinterval = 0;
    bool swap = false;
};

}  // namespace

extern "C" {

void drv_set_projection(double lat0, double lon0, double R, double D)
{
    g_lat0 = lat0;
    g_lon0 = lon0;
    g_R = R;
    g_D = D;
}

// swap_ids: put the smx_level's image 0 at SETSM image index 1, so the
// shim's reference_id mapping is exercised (reference_id = 1).
void *drv_create(const smx_level *lv, int swap_ids)
{
    Driver *d = new Driver();
    memset(&d->pro, 0, sizeof d->pro);
    memset(&d->L, 0, sizeof d->L);
    d->pro.resolution = lv->resolution;
    d->pro.DEM_resolution = lv->dem_resolution;
    d->pro.number_of_images = 2;
    d->pro.sensor_type = SB;
    d->pro.IsRA = lv->is_ra;
    d->pro.pre_DEMtif = lv->pre_dem;
    d->pro.check_Matchtag = lv->check_matchtag;
    d->pro.check_selected_image[0] = d->pro.check_selected_image[1] = true;
    d->pro.SGM_py = lv->sgm_py;

    d->level = lv->level;
    d->combined = lv->py_combined_level;
    d->reference_id = swap_ids ? 1 : 0;
    d->swap = swap_ids != 0;
    d->template_size = (unsigned char)lv->template_size;
    d->bin_angle = lv->bin_angle;
    d->height_step = lv->height_step;
    d->grid_resolution = lv->grid_resolution;
    for (int k = 0; k < 4; k++) d->boundary[k] = lv->boundary[k];
    d->grid_size = CSize(lv->grid_w, lv->grid_h);
    d->grid_length = (long)lv->grid_w * lv->grid_h;
    for (int k = 0; k < 2; k++) {
        const int ti = swap_ids ? 1 - k : k;   // SETSM image index of smx image k
        d->img[ti] = (const uint16 *)lv->view[k].img;
        d->mag[ti] = (const uint16 *)lv->view[k].mag;
        d->ori[ti] = (const uint8 *)lv->view[k].ori;
        d->start[ti].m_X = lv->view[k].start_x;
        d->start[ti].m_Y = lv->view[k].start_y;
        d->img_next[ti] = (const uint16 *)lv->next[k].img;
        d->mag_next[ti] = (const uint16 *)lv->next[k].mag;
        d->ori_next[ti] = (const uint8 *)lv->next[k].ori;
        d->start_next[ti].m_X = lv->next[k].start_x;
        d->start_next[ti].m_Y = lv->next[k].start_y;
        d->sizes[ti].assign((size_t)lv->level + 1, CSize(0, 0));
        d->sizes[ti][lv->level] = CSize(lv->view[k].width, lv->view[k].height);
        if (lv->level >= 1) d->sizes[ti][lv->level - 1] = CSize(lv->next[k].width, lv->next[k].height);
        memset(d->rpc_rows[ti], 0, sizeof d->rpc_rows[ti]);
        for (int j = 0; j < 5; j++) {
            d->rpc_rows[ti][0][j] = lv->rpc[k].off[j];
            d->rpc_rows[ti][1][j] = lv->rpc[k].scale[j];
        }
        for (int p = 0; p < 4; p++)
            for (int j = 0; j < 20; j++) d->rpc_rows[ti][p + 2][j] = lv->rpc[k].coef[p][j];
        for (int r = 0; r < 7; r++) d->rpc_ptr[ti][r] = d->rpc_rows[ti][r];
        d->rpcs[ti] = d->rpc_ptr[ti];
        d->adjust[ti][0] = lv->adjust[k][0];
        d->adjust[ti][1] = lv->adjust[k][1];
        d->adjust_ptr[ti] = d->adjust[ti];
        d->size_rows[ti] = d->sizes[ti].data();
    }
    LevelInfo &L = d->L;
    L.py_Images = d->img;
    L.py_MagImages = d->mag;
    L.py_OriImages = d->ori;
    L.py_Images_next = lv->level >= 1 && lv->next[0].img ? d->img_next : nullptr;
    L.py_MagImages_next = d->mag_next;
    L.py_OriImages_next = d->ori_next;
    L.py_Startpos = d->start;
    L.py_Startpos_next = d->start_next;
    L.RPCs = d->rpcs;
    L.Boundary = d->boundary;
    L.Pyramid_step = &d->level;
    L.py_Sizes = d->size_rows;
    L.Template_size = &d->template_size;
    L.ImageAdjust = d->adjust_ptr;
    L.NumOfIAparam = &d->num_ia;
    L.bin_angle = &d->bin_angle;
    L.Size_Grid2D = &d->grid_size;
    L.Grid_length = &d->grid_length;
    L.GridPts = (const D2DPOINT *)lv->grid_pts;
    L.Grid_wgs = (const D2DPOINT *)lv->grid_wgs;
    L.reference_id = d->reference_id;
    L.height_step = &d->height_step;
    L.grid_resolution = &d->grid_resolution;
    L.Py_combined_level = &d->combined;
    L.check_matching_rate = &d->check_matching_rate;
    memset(&d->param, 0, sizeof d->param);
    L.param = &d->param;
    // setsm_code.cpp:3769-3770: calloc'd per level.
    d->ncc.assign((size_t)d->grid_length, NCCresult());
    memset(d->ncc.data(), 0, d->ncc.size() * sizeof(NCCresult));
    d->vox.assign((size_t)d->grid_length, nullptr);
    return d;
}

// One iteration: InitializeVoxel, VerticalLineLocus, AWNCC on host arrays.
// ugrid is updated in place; returns Accessable_grid.
long long drv_step(void *h, smx_ugrid *ugrid, const double *minmax, int iteration)
{
    Driver *d = (Driver *)h;
    UGRID *g = (UGRID *)ugrid;
    InitializeVoxel(&d->pro, d->vox.data(), d->L, g, d->ncc.data(), iteration, minmax);
    d->acc = VerticalLineLocus(d->vox.data(), &d->pro, d->ncc.data(), d->L, g, (uint8)iteration, minmax);
    d->ncc_vll = d->ncc;
    AWNCC(&d->pro, d->vox.data(), d->grid_size, g, d->ncc.data(), d->height_step, (uint8)d->level, (uint8)iteration,
          0);
    return d->acc;
}

long long drv_num_hyps(void *h)
{
    Driver *d = (Driver *)h;
    long long t = 0;
    for (const NCCresult &r : d->ncc) t += r.NumOfHeight;
    return t;
}

// Host grid_voxel rows in cell order, NCCresult after VerticalLineLocus and
// after AWNCC.
void drv_get(void *h, uint32_t *words, smx_nccresult *ncc_vll, smx_nccresult *ncc)
{
    Driver *d = (Driver *)h;
    long long k = 0;
    for (long c = 0; c < d->grid_length; c++)
        for (int i = 0; i < d->ncc[c].NumOfHeight; i++) memcpy(&words[k++], &d->vox[c][i], 4);
    memcpy(ncc_vll, d->ncc_vll.data(), d->ncc.size() * sizeof(NCCresult));
    memcpy(ncc, d->ncc.data(), d->ncc.size() * sizeof(NCCresult));
}

// SelectMPs after the last AWNCC: ugrid updated in place, points (in the
// order SETSM's vector receives them) into out; returns their count.
long drv_select_mps(void *h, smx_ugrid *ugrid, double th_roh, double th_roh_min, double th_roh_start,
                    double th_roh_next, int iteration, double mpp, int final_level_iteration,
                    double mpp_stereo_angle, smx_point3 *out)
{
    Driver *d = (Driver *)h;
    std::vector<D3DPOINT> pts;
    const long n = SelectMPs(&d->pro, d->L, d->ncc.data(), (UGRID *)ugrid, th_roh, th_roh_min, th_roh_start,
                             th_roh_next, iteration, mpp, final_level_iteration, mpp_stereo_angle, &pts);
    for (size_t i = 0; i < pts.size(); i++) {
        memset(&out[i], 0, sizeof out[i]);
        out[i].x = pts[i].m_X;
        out[i].y = pts[i].m_Y;
        out[i].z = pts[i].m_Z;
        out[i].flag = pts[i].flag;
    }
    return n == (long)pts.size() ? n : -1;
}

// SetHeightRange on a malloc'd copy of ugrid (the function frees its input,
// as SETSM's does); the returned grid is copied back into ugrid.
void drv_set_height_range(void *h, smx_ugrid *ugrid, smx_point3 *pts, long npts, const smx_tri *tris, long ntris,
                          int iteration, double mpp, int check_matching_rate, double *lo, double *hi)
{
    Driver *d = (Driver *)h;
    UGRID *g = (UGRID *)malloc(sizeof(UGRID) * (size_t)d->grid_length);
    memcpy(g, ugrid, sizeof(UGRID) * (size_t)d->grid_length);
    UGRID *r = SetHeightRange(&d->pro, d->L, d->ncc.data(), (int)npts, (int)ntris, g, iteration, lo, hi,
                              (D3DPOINT *)pts, (const UI3DPOINT *)tris, mpp, check_matching_rate != 0);
    memcpy(ugrid, r, sizeof(UGRID) * (size_t)d->grid_length);
    free(r);
}

// The blunder level's images (smx image order k = 0 reference, 1 target),
// as Matching_SETSM sets py_BImages / py_BMagImages / py_BStartpos
// (setsm_code.cpp:3560-3600).
void drv_set_blunder(void *h, int B, const uint16 *img0, const uint16 *mag0, const uint16 *img1, const uint16 *mag1,
                     int w0, int h0, int w1, int h1, double sx0, double sy0, double sx1, double sy1)
{
    Driver *d = (Driver *)h;
    const uint16 *im[2] = {img0, img1}, *mg[2] = {mag0, mag1};
    const int ws[2] = {w0, w1}, hs[2] = {h0, h1};
    const double sx[2] = {sx0, sx1}, sy[2] = {sy0, sy1};
    for (int k = 0; k < 2; k++) {
        const int ti = d->swap ? 1 - k : k;
        d->bimg[ti] = im[k];
        d->bmag[ti] = mg[k];
        d->bstart[ti].m_X = sx[k];
        d->bstart[ti].m_Y = sy[k];
        if ((int)d->sizes[ti].size() < B + 1) d->sizes[ti].resize((size_t)B + 1, CSize(0, 0));
        d->sizes[ti][B] = CSize(ws[k], hs[k]);
        d->size_rows[ti] = d->sizes[ti].data();
    }
    d->blunder_level = B;
    d->L.py_BImages = d->bimg;
    d->L.py_BMagImages = d->bmag;
    d->L.py_BStartpos = d->bstart;
    d->L.blunder_selected_level = &d->blunder_level;
    d->L.py_Sizes = d->size_rows;
}

void drv_set_height_range_blunder(void *h, smx_ugrid *ugrid, smx_point3 *pts, long npts, smx_tri *tris, long ntris)
{
    Driver *d = (Driver *)h;
    SetHeightRange_blunder(d->L, (const D3DPOINT *)pts, (int)npts, (UI3DPOINT *)tris, ntris, (UGRID *)ugrid);
}

void drv_vll_blunder(void *h, smx_ugrid *ugrid, int iteration, int bblunder, float *out)
{
    Driver *d = (Driver *)h;
    VerticalLineLocus_blunder(&d->pro, d->L, out, (UGRID *)ugrid, (uint8)iteration, bblunder != 0);
}

void drv_blunder_detection(void *h, smx_ugrid *ugrid, int iteration, float *ortho, int flag_blunder, int count_bl,
                           smx_point3 *pts, long npts, smx_tri *tris, long ntris, double hinterval, uint8_t *det,
                           long *blunder_count, double *minz, double *maxz)
{
    Driver *d = (Driver *)h;
    d->hinterval = hinterval;
    d->L.Hinterval = &d->hinterval;
    blunder_detection_TIN(&d->pro, d->L, iteration, ortho, flag_blunder != 0, (uint16)count_bl, (D3DPOINT *)pts,
                          (bool *)det, npts, (UI3DPOINT *)tris, ntris, (UGRID *)ugrid, blunder_count, minz, maxz);
}

void drv_destroy(void *h)
{
    Driver *d = (Driver *)h;
    for (VOXEL *v : d->vox) free(v);
    delete d;
}

}  // extern "C"
