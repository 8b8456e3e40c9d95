/*
 * smx_oracle.c — CPU FP64 restatement of SETSM 4.3.16's VerticalLineLocus
 * hot path, used as the parity checker and the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: never linked into or called by the product.
 *
 * Floating point follows the reference expression by expression (same
 * operand order, same casts, no FMA contraction: build with
 * -ffp-contract=off), so the FP64 GPU kernel can be compared bit for bit.
 *
 * Pins (checks that do not just retype these formulas) live in tests/:
 * closed-form RPCs and the real reference header code built into oracle/_ref
 * (test_oracle.py), window counts, numpy corrcoef, brute-force bilinear,
 * hand-derived tables for SetNCC_alpha, check_image_boundary, the
 * VerticalLineLocus gate, InitializeVoxel and FindPeakNcc
 * (test_oracle_pins.py), a brute-force GNCC (test_oracle_gncc.py),
 * closed-form SGM cases (test_oracle_sgm.py), and hand-derived cases of the
 * steps between iterations: SelectMPs and SetHeightRange
 * (test_oracle_mps.py), DecisionMPs' blunder helpers (test_oracle_blunder.py)
 * and blunder_detection_TIN (test_oracle_bdtin.py).
 *
 * Scope: two images (reference_id 0 or 1, ti the other) and the SB (RPC)
 * sensor. The ortho GNCC term (setsm_code.cpp:6240-6305) is computed, from
 * the ortho-grid lon/lat the caller supplies (smx_level.ortho_wgs).
 */
#include "smx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793          /* Typedefine.hpp:27 */
#define ORC_PWRTWO(x) (1 << (x))          /* Typedefine.hpp:43 */

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
int  orc_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------- voxels */

orc_voxels *orc_voxels_create(long ncells, int keep_rho)
{
    orc_voxels *v = (orc_voxels *)calloc(1, sizeof(orc_voxels));
    v->ncells = ncells;
    v->rows = (uint32_t **)calloc((size_t)ncells, sizeof(uint32_t *));
    v->len = (uint16_t *)calloc((size_t)ncells, sizeof(uint16_t));
    if (keep_rho)
        v->rho = (double **)calloc((size_t)ncells, sizeof(double *));
    return v;
}

static void vox_free_row(orc_voxels *v, long c)
{
    free(v->rows[c]);
    v->rows[c] = NULL;
    v->len[c] = 0;
    if (v->rho) {
        free(v->rho[c]);
        v->rho[c] = NULL;
    }
}

void orc_voxels_destroy(orc_voxels *v)
{
    if (!v) return;
    for (long c = 0; c < v->ncells; c++)
        vox_free_row(v, c);
    free(v->rows);
    free(v->len);
    free(v->rho);
    free(v);
}

long orc_voxels_export(const orc_voxels *v, const smx_nccresult *ncc, uint32_t *words_out, double *rho_out)
{
    long k = 0;
    for (long c = 0; c < v->ncells; c++) {
        int n = ncc[c].NumOfHeight;
        for (int h = 0; h < n; h++, k++) {
            words_out[k] = (v->rows[c] && h < v->len[c]) ? v->rows[c][h] : 0u;
            if (rho_out)
                rho_out[k] = (v->rho && v->rho[c] && h < v->len[c]) ? v->rho[c][h] : NAN;
        }
    }
    return k;
}

void orc_voxels_import(orc_voxels *v, const smx_nccresult *ncc, const uint32_t *words)
{
    long k = 0;
    for (long c = 0; c < v->ncells; c++) {
        vox_free_row(v, c);
        int n = ncc[c].NumOfHeight;
        if (n > 0) {
            v->rows[c] = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
            memcpy(v->rows[c], words + k, sizeof(uint32_t) * (size_t)n);
            v->len[c] = (uint16_t)n;
            if (v->rho) {
                v->rho[c] = (double *)malloc(sizeof(double) * (size_t)n);
                for (int h = 0; h < n; h++) v->rho[c][h] = NAN;
            }
        }
        k += n;
    }
}

/* InitializeVoxel's row allocation: flag_cal = true, INCC =
 * DoubleToSignedChar_voxel(-1) (setsm_code.cpp:5940-5948). */
static void vox_alloc_row(orc_voxels *v, long c, int n)
{
    vox_free_row(v, c);
    v->rows[c] = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    for (int h = 0; h < n; h++)
        v->rows[c][h] = SMX_VOXEL_WORD(1, -1000);
    v->len[c] = (uint16_t)n;
    if (v->rho) {
        v->rho[c] = (double *)malloc(sizeof(double) * (size_t)n);
        for (int h = 0; h < n; h++) v->rho[c][h] = NAN;
    }
}

/* ------------------------------------------------------- quantisation */

/* Template.hpp:964-967 DoubleToSignedChar_result: (short)round(v*1000).
 * The double->short conversion goes through a 32-bit integer and keeps the low
 * 16 bits, which is what GCC on x86-64 does for out-of-range values such as
 * max_WNCC's -100 sentinel (-100000 -> 31072). */
static int16_t q_result(double v) { return (int16_t)(int32_t)round(v * 1000.0); }
/* Template.hpp:984-987 DoubleToSignedChar_voxel: (short)(v*1000), truncation */
static int16_t q_voxel(double v) { return (int16_t)(int32_t)(v * 1000.0); }
/* Template.hpp:969-972 / 989-992 */
static double  dq(int16_t v) { return (double)v / 1000.0; }

/* ------------------------------------------------------------- geometry */

void orc_rpc_ground_to_image(const smx_rpc *rpc, const double adjust[2], double gx, double gy, double gz,
                             double *img_x, double *img_y)
{
    /* CoordConversion.hpp:26-87 */
    double L = (gx - rpc->off[2]) / rpc->scale[2];
    double P = (gy - rpc->off[3]) / rpc->scale[3];
    double H = (gz - rpc->off[4]) / rpc->scale[4];

    if (L < -10.0 || L > 10.0) {
        if (gx > 0) gx = gx - 360;
        else        gx = gx + 360;
        L = (gx - rpc->off[2]) / rpc->scale[2];
    }
    if (P < -10.0 || P > 10.0) {
        if (gy > 0) gy = gy - 360;
        else        gy = gy + 360;
        P = (gy - rpc->off[3]) / rpc->scale[3];
    }

    double Coeff[4];
    for (int j = 0; j < 4; j++) {
        const double *c = rpc->coef[j];
        Coeff[j] = c[0] * 1.0 + c[1] * L + c[2] * P
                 + c[3] * H + c[4] * L * P + c[5] * L * H
                 + c[6] * P * H + c[7] * L * L + c[8] * P * P
                 + c[9] * H * H + c[10] * (P * L) * H + c[11] * (L * L) * L
                 + c[12] * (L * P) * P + c[13] * (L * H) * H + c[14] * (L * L) * P
                 + c[15] * (P * P) * P + c[16] * (P * H) * H + c[17] * (L * L) * H
                 + c[18] * (P * P) * H + c[19] * (H * H) * H;
    }
    double Line = ((Coeff[0] / Coeff[1]) * rpc->scale[0] + rpc->off[0]);
    double Samp = ((Coeff[2] / Coeff[3]) * rpc->scale[1] + rpc->off[1]);

    double y = adjust[0] + Line;
    double x = adjust[1] + Samp;
    if (y < 0) y = 0;
    if (y > rpc->off[0] + rpc->scale[0] * 1.2) y = rpc->off[0] + rpc->scale[0] * 1.2;
    if (x < 0) x = 0;
    if (x > rpc->off[1] + rpc->scale[1] * 1.2) x = rpc->off[1] + rpc->scale[1] * 1.2;
    *img_x = x;
    *img_y = y;
}

/* CoordConversion.hpp:96-105 OriginalToPyramid_single */
static void to_pyramid(double x, double y, double sx, double sy, int level, double *px, double *py)
{
    *px = (x / ORC_PWRTWO(level)) - sx;
    *py = (y / ORC_PWRTWO(level)) - sy;
}

double orc_height_step(int level, double im_resolution)
{
    /* setsm_code.cpp:5741-5757 */
    const double h_divide = 2;
    im_resolution = im_resolution * ORC_PWRTWO(level);
    double HS = (double)(im_resolution / h_divide);
    double tt1 = HS * 1000.0;
    double tt2 = floor(tt1 + 0.1);
    HS = tt2 / 1000.0;
    if (HS > 3) HS = 3;
    return HS;
}

int orc_vll_half_template(int level, int template_size, double resolution)
{
    /* setsm_code.cpp:6057-6077 */
    int ts = template_size;
    if (level >= 1) {
        double template_area = 5.0;
        int t = (int)((template_area / (resolution * ORC_PWRTWO(level))) / 2.0) * 2 + 1;
        if (template_size < t) ts = t;
    }
    int half = ts / 2;
    if (level == 3) half = half - 1;
    else if (level < 3) half = half - 2;
    return half;
}

int orc_window_bits(int half, int row, int col)
{
    /* setsm_code.cpp:6434 (outer disc) and SubFunctions.cpp:3282-3312 */
    int r2 = row * row + col * col;
    if (row < -half || row > half || col < -half || col > half) return 0;
    if (r2 > (half + 1) * (half + 1)) return 0;
    int bits = 1;
    const int s1 = (int)(half / 2);
    if (r2 <= (half - s1 + 1) * (half - s1 + 1) &&
        row >= -half + s1 && row <= half - s1 && col >= -half + s1 && col <= half - s1)
        bits |= 2;
    const int s2 = s1 + (int)((s1 / 2.0) + 0.5);
    if (r2 <= (half - s2 + 1) * (half - s2 + 1) &&
        row >= -half + s2 && row <= half - s2 && col >= -half + s2 && col <= half - s2)
        bits |= 4;
    return bits;
}

int orc_check_image_boundary(const smx_level *L, long cell, double minH, double maxH, int half)
{
    /* setsm_code.cpp:11002-11083, SB sensor, images 0 and 1 both selected */
    const int buff_pixel = 1;
    int selected = 0;
    int check_reference = 1;
    const double lon = L->grid_wgs[2 * cell], lat = L->grid_wgs[2 * cell + 1];
    for (int ti = 0; ti < 2; ti++) {
        const smx_view *v = &L->view[ti];
        int ok[2];
        for (int e = 0; e < 2; e++) {
            double x, y, px, py;
            orc_rpc_ground_to_image(&L->rpc[ti], L->adjust[ti], lon, lat, e == 0 ? minH : maxH, &x, &y);
            to_pyramid(x, y, v->start_x, v->start_y, L->level, &px, &py);
            ok[e] = (px > half + buff_pixel && px < v->width - half - buff_pixel &&
                     py > half + buff_pixel && py < v->height - half - buff_pixel);
        }
        if (ok[0] && ok[1]) selected++;
        else if (ti == 0) check_reference = 0;
    }
    if (!check_reference) return 0;
    return selected >= 2;
}

double orc_ncc_alpha(int Pyramid_step, int iteration, int IsRA)
{
    /* setsm_code.cpp:6014-6047 */
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

/* --------------------------------------------------------------- NCC */

double orc_correlate(const double *L, const double *R, const int N)
{
    /* SubFunctions.cpp:974-1018, two-pass Pearson with the -99 sentinel */
    double Lmean = 0, Rmean = 0, rho;
    if (N > 0) {
        for (int i = 0; i < N; i++) {
            Lmean += L[i];
            Rmean += R[i];
        }
        Lmean = Lmean / N;
        Rmean = Rmean / N;
        double SumLR = 0, SumL2 = 0, SumR2 = 0;
        for (int i = 0; i < N; i++) {
            SumLR += (L[i] - Lmean) * (R[i] - Rmean);
            SumL2 += (L[i] - Lmean) * (L[i] - Lmean);
            SumR2 += (R[i] - Rmean) * (R[i] - Rmean);
        }
        if (SumL2 > 1e-8 && SumR2 > 1e-8)
            rho = SumLR / (sqrt(SumL2 * SumR2));
        else
            rho = (double)-99;
    } else
        rho = (double)-99;
    return rho;
}

/* Per-thread patch vectors, SetKernel (Typedefine.hpp:727-754). */
typedef struct patchvec {
    int cap;
    double *l[3], *lm[3], *r[3], *rm[3];
} patchvec;

static void pv_init(patchvec *p, int half)
{
    p->cap = (2 * half + 1) * (2 * half + 1);
    for (int k = 0; k < 3; k++) {
        p->l[k] = (double *)malloc(sizeof(double) * (size_t)p->cap);
        p->lm[k] = (double *)malloc(sizeof(double) * (size_t)p->cap);
        p->r[k] = (double *)malloc(sizeof(double) * (size_t)p->cap);
        p->rm[k] = (double *)malloc(sizeof(double) * (size_t)p->cap);
    }
}

static void pv_free(patchvec *p)
{
    for (int k = 0; k < 3; k++) {
        free(p->l[k]); free(p->lm[k]); free(p->r[k]); free(p->rm[k]);
    }
}

/* SubFunctions.cpp:3317-3326 InterpolatePatch */
static double interp(const uint16_t *I, long pos, uint32_t w, double dx, double dy)
{
    return (double)(I[pos]) * (1 - dx) * (1 - dy) +
           (double)(I[pos + 1]) * dx * (1 - dy) +
           (double)(I[pos + w]) * (1 - dx) * dy +
           (double)(I[pos + 1 + w]) * dx * dy;
}

/* SubFunctions.cpp:3249-3315 SetVecKernelValue */
static void set_vec_kernel_value(patchvec *pv, int half, const smx_view *Lv, const smx_view *Rv,
                                 int row, int col, double lx, double ly, double rx, double ry,
                                 int radius2, int *Count_N)
{
    if (ry >= 0 && ry + 1 < Rv->height && rx >= 0 && rx + 1 < Rv->width &&
        ly >= 0 && ly + 1 < Lv->height && lx >= 0 && lx + 1 < Lv->width) {
        const long position = (long)lx + (long)ly * (long)Lv->width;
        const long position_right = (long)rx + (long)ry * (long)Rv->width;
        double dx = lx - floor(lx);
        double dy = ly - floor(ly);
        double dx_r = rx - floor(rx);
        double dy_r = ry - floor(ry);
        double lp = interp(Lv->img, position, Lv->width, dx, dy);
        double lmp = interp(Lv->mag, position, Lv->width, dx, dy);
        double rp = interp(Rv->img, position_right, Rv->width, dx_r, dy_r);
        double rmp = interp(Rv->mag, position_right, Rv->width, dx_r, dy_r);
        if (lp > 0 && rp > 0) {
            pv->l[0][Count_N[0]] = lp; pv->lm[0][Count_N[0]] = lmp;
            pv->r[0][Count_N[0]] = rp; pv->rm[0][Count_N[0]] = rmp;
            Count_N[0]++;
            const int size_1 = (int)(half / 2);
            if (radius2 <= (half - size_1 + 1) * (half - size_1 + 1) &&
                row >= -half + size_1 && row <= half - size_1 &&
                col >= -half + size_1 && col <= half - size_1) {
                pv->l[1][Count_N[1]] = lp; pv->lm[1][Count_N[1]] = lmp;
                pv->r[1][Count_N[1]] = rp; pv->rm[1][Count_N[1]] = rmp;
                Count_N[1]++;
            }
            const int size_2 = size_1 + (int)((size_1 / 2.0) + 0.5);
            if (radius2 <= (half - size_2 + 1) * (half - size_2 + 1) &&
                row >= -half + size_2 && row <= half - size_2 &&
                col >= -half + size_2 && col <= half - size_2) {
                pv->l[2][Count_N[2]] = lp; pv->lm[2][Count_N[2]] = lmp;
                pv->r[2][Count_N[2]] = rp; pv->rm[2][Count_N[2]] = rmp;
                Count_N[2]++;
            }
        }
    }
}

/* SubFunctions.cpp:3328-3359 ComputeMultiNCC with Th_rho = 0 */
static void compute_multi_ncc(const patchvec *pv, const int *Count_N, double *count_NCC, double *sum_NCC_multi)
{
    if (Count_N[0] > 0 && Count_N[1] > 0 && Count_N[2] > 0) {
        double temp_roh = 0, count_roh = 0;
        for (int k = 0; k < 3; k++) {
            const double ncc = orc_correlate(pv->l[k], pv->r[k], Count_N[k]);
            if (ncc != -99) { count_roh++; temp_roh += ncc; }
            const double ncc_mag = orc_correlate(pv->lm[k], pv->rm[k], Count_N[k]);
            if (ncc_mag != -99) { count_roh++; temp_roh += ncc_mag; }
        }
        if (count_roh > 0) {
            *sum_NCC_multi += temp_roh / count_roh;
            *count_NCC += 1;
        }
    }
}

/* ------------------------------------------------------ InitializeVoxel */

void orc_initialize_voxel(const smx_level *Lv, orc_voxels *vox, smx_ugrid *GridPT3,
                          smx_nccresult *nccresult, int iteration, const double minmaxHeight[2])
{
    /* setsm_code.cpp:5759-6012 (check_image_boundary with H_template_size 7) */
    const double height_step = Lv->height_step;
    const int pyramid_step = Lv->level;
    const long n = (long)Lv->grid_w * (long)Lv->grid_h;
#pragma omp parallel for schedule(guided) num_threads(orc_get_threads())
    for (long t_i = 0; t_i < n; t_i++) {
        smx_nccresult *nr = &nccresult[t_i];
        smx_ugrid *g = &GridPT3[t_i];
        if (orc_check_image_boundary(Lv, t_i, g->minHeight, g->maxHeight, 7)) {
            int change_step_min = 0, change_step_max = 0;
            int check_blunder_cell = 1;
            double th_height = 1000;
            if (Lv->dem_resolution <= 4) th_height = 500;

            if (Lv->check_matchtag || iteration == 1) {
                nr->minHeight = g->minHeight;
                nr->maxHeight = g->maxHeight;
                nr->GNCC = q_result(-1.0);
                nr->check_height_change = 1;
                check_blunder_cell = 0;
            } else {
                if (pyramid_step >= 2)
                    check_blunder_cell = 0;
                else if (g->Matched_flag != 0) {
                    if (g->maxHeight - g->minHeight > 0) {
                        if (pyramid_step <= 1) {
                            if (iteration > 1) {
                                if (g->maxHeight <= minmaxHeight[1] && g->minHeight >= minmaxHeight[0] &&
                                    nr->maxHeight <= minmaxHeight[1] && nr->minHeight >= minmaxHeight[0])
                                    check_blunder_cell = 0;
                                else
                                    check_blunder_cell = 1;
                            } else {
                                if (g->maxHeight <= minmaxHeight[1] && g->minHeight >= minmaxHeight[0])
                                    check_blunder_cell = 0;
                                else
                                    check_blunder_cell = 1;
                            }
                        }
                        if (pyramid_step == 1) {
                            if (abs(g->maxHeight - g->minHeight) < 1000) check_blunder_cell = 0;
                            else check_blunder_cell = 1;
                        }
                        if (pyramid_step == 0) {
                            if (iteration > 1) {
                                if ((abs(nr->maxHeight - nr->minHeight) < th_height) &&
                                    (abs(g->maxHeight - g->minHeight) < th_height))
                                    check_blunder_cell = 0;
                                else
                                    check_blunder_cell = 1;
                            } else {
                                if (abs(g->maxHeight - g->minHeight) < th_height) check_blunder_cell = 0;
                                else nr->NumOfHeight = 0;
                            }
                        }
                    } else {
                        if (nr->NumOfHeight > 0) vox_free_row(vox, t_i);
                        nr->NumOfHeight = 0;
                        check_blunder_cell = 1;
                    }
                } else {
                    if (nr->NumOfHeight > 0) vox_free_row(vox, t_i);
                    nr->NumOfHeight = 0;
                    check_blunder_cell = 1;
                }
            }

            if (!check_blunder_cell) {
                if (iteration == 1 || (pyramid_step == 0 && Lv->dem_resolution < 2)) {
                    nr->minHeight = g->minHeight;
                    nr->maxHeight = g->maxHeight;
                    nr->GNCC = q_result(-1.0);
                    nr->check_height_change = 1;
                    if (pyramid_step == 1) {
                        if (abs((int)g->maxHeight - (int)g->minHeight) > 1000) {
                            g->maxHeight = -100;
                            g->minHeight = -100;
                            nr->check_height_change = 0;
                        }
                    } else if (pyramid_step == 0) {
                        if (abs((int)g->maxHeight - (int)g->minHeight) > th_height) {
                            g->maxHeight = -100;
                            g->minHeight = -100;
                            nr->check_height_change = 0;
                        }
                    }
                } else {
                    const int NumberofHeightVoxel = (int)((g->maxHeight - g->minHeight) / height_step);
                    if (nr->NumOfHeight == 0 && NumberofHeightVoxel > 0) {
                        change_step_min = 0;
                        change_step_max = 0;
                        nr->check_height_change = 1;
                    } else {
                        if (nr->minHeight > g->minHeight || nr->maxHeight < g->maxHeight) {
                            if (nr->minHeight > g->minHeight)
                                change_step_min = (int)((nr->minHeight - g->minHeight) / height_step + 0.5);
                            if (nr->maxHeight < g->maxHeight)
                                change_step_max = (int)((g->maxHeight - nr->maxHeight) / height_step + 0.5);
                            nr->minHeight = (int16_t)floor(nr->minHeight - change_step_min * height_step);
                            nr->maxHeight = (int16_t)ceil(nr->maxHeight + change_step_max * height_step);
                            nr->check_height_change = 1;
                            if (pyramid_step == 0 && abs(nr->maxHeight - nr->minHeight) > th_height)
                                nr->check_height_change = 0;
                        } else
                            nr->check_height_change = 0;
                    }
                }

                if (nr->check_height_change) {
                    nr->minHeight = (int16_t)floor(nr->minHeight - change_step_min * height_step);
                    nr->maxHeight = (int16_t)ceil(nr->maxHeight + change_step_max * height_step);
                    const int NumberofHeightVoxel = (int)((float)(nr->maxHeight - nr->minHeight) / height_step);
                    if (NumberofHeightVoxel > 0) {
                        nr->NumOfHeight = (uint16_t)NumberofHeightVoxel;
                        vox_alloc_row(vox, t_i, NumberofHeightVoxel);
                    } else {
                        if (nr->NumOfHeight > 0) vox_free_row(vox, t_i);
                        nr->NumOfHeight = 0;
                        nr->check_height_change = 0;
                    }
                }
            } else {
                if (nr->NumOfHeight > 0) vox_free_row(vox, t_i);
                nr->NumOfHeight = 0;
                nr->check_height_change = 0;
            }
        } else {
            if (nr->NumOfHeight > 0) vox_free_row(vox, t_i);
            nr->NumOfHeight = 0;
            nr->check_height_change = 0;
        }

        if (nr->minHeight == 0 || nr->maxHeight == 0) {
            if (nr->NumOfHeight > 0) vox_free_row(vox, t_i);
            nr->NumOfHeight = 0;
            nr->check_height_change = 0;
        }
        if (pyramid_step == 0 && nr->NumOfHeight > 1000) {
            if (nr->NumOfHeight > 0) vox_free_row(vox, t_i);
            nr->NumOfHeight = 0;
            nr->check_height_change = 0;
        }
    }
}

/* ---------------------------------------------------------- GNCC */

/* setsm_code.cpp:6550-6670 SetOrthoImageCoord, SB sensor, check_pyimage =
 * OR. The ortho pixels' lon/lat (ps2wgs_single of the pixel's PS position)
 * come from the caller in L->ortho_wgs. cd[k] / cdn[k] hold pyramid image
 * coordinates (x, y) of image k (0 = reference, 1 = target) at the level and
 * at level - 1; entries never written stay 0 (calloc). */
static void set_ortho_image_coord(const smx_level *Lv, const smx_ugrid *GridPT3, int combined, double im_res,
                                  double im_res_next, long sw, long sh, long swn, long shn, double *cd[2],
                                  double *cdn[2])
{
    const long total = sw * sh;
    const long W = Lv->grid_w, H = Lv->grid_h;
    const long grid_length = W * H;
#pragma omp parallel for schedule(guided) num_threads(orc_get_threads())
    for (long it = 0; it < total; it++) {
        const long pts_row = (long)floor((double)(it / sw));
        const long pts_col = it % sw;
        const double t_X = Lv->boundary[0] + pts_col * im_res;
        const double t_Y = Lv->boundary[1] + pts_row * im_res;
        const long t_col = (long)((t_X - Lv->boundary[0]) / Lv->grid_resolution);
        const long t_row = (long)((t_Y - Lv->boundary[1]) / Lv->grid_resolution);
        const long pt_index = t_row * W + t_col;
        const long pt_index_im = pts_row * sw + pts_col;
        long t_col_next = 0, t_row_next = 0, pt_index_im_next = 0;
        if (combined) {
            t_col_next = (long)((t_X - Lv->boundary[0]) / im_res_next);
            t_row_next = (long)((t_Y - Lv->boundary[1]) / im_res_next);
            pt_index_im_next = t_row_next * swn + t_col_next;
        }
        if (pt_index < grid_length && t_col < W && t_row < H) {
            if (GridPT3[pt_index].Height != -1000) {
                const double lon = Lv->ortho_wgs[2 * it], lat = Lv->ortho_wgs[2 * it + 1];
                const double Z = (double)GridPT3[pt_index].Height;
                for (int k = 0; k < 2; k++) {
                    double ix, iy, px, py;
                    orc_rpc_ground_to_image(&Lv->rpc[k], Lv->adjust[k], lon, lat, Z, &ix, &iy);
                    to_pyramid(ix, iy, Lv->view[k].start_x, Lv->view[k].start_y, Lv->level, &px, &py);
                    cd[k][2 * pt_index_im] = px;
                    cd[k][2 * pt_index_im + 1] = py;
                    if (combined) {
                        to_pyramid(ix, iy, Lv->next[k].start_x, Lv->next[k].start_y, Lv->level - 1, &px, &py);
                        if (pt_index_im_next < swn * shn && t_col_next < swn && t_row_next < shn) {
                            cdn[k][2 * pt_index_im_next] = px;
                            cdn[k][2 * pt_index_im_next + 1] = py;
                        }
                    }
                }
            }
        }
    }
}

/* setsm_code.cpp:6240-6305: the ortho multi-window NCC of one cell on the
 * fixed ortho-grid geometry. Returns count_GNCC, adds to *sum. */
static double gncc_cell(const smx_level *Lv, const smx_ugrid *GridPT3, long pt_index, int H, int combined,
                        double im_res, double im_res_next, long sw, long sh, long swn, long shn,
                        double *const cd[2], double *const cdn[2], patchvec *pv, patchvec *pvn, double *sum)
{
    const long W = Lv->grid_w, Hg = Lv->grid_h;
    const long numofpts = W * Hg;
    int Count_N_ortho[3] = {0, 0, 0}, Count_N_ortho_next[3] = {0, 0, 0};
    double count_GNCC = 0;
    for (int row = -H; row <= H; row++) {
        for (int col = -H; col <= H; col++) {
            const int radius2 = row * row + col * col;
            if (radius2 > (H + 1) * (H + 1)) continue;
            const double t_X = Lv->grid_pts[2 * pt_index] + col * im_res;
            const double t_Y = Lv->grid_pts[2 * pt_index + 1] + row * im_res;
            long t_col = (long)((t_X - Lv->boundary[0]) / im_res);
            long t_row = (long)((t_Y - Lv->boundary[1]) / im_res);
            long pt_index_temp = t_row * sw + t_col;
            const long tt_col = (long)((t_X - Lv->boundary[0]) / Lv->grid_resolution);
            const long tt_row = (long)((t_Y - Lv->boundary[1]) / Lv->grid_resolution);
            const long pt_index_dem = tt_row * W + tt_col;
            if (pt_index_temp >= 0 && pt_index_temp < sw * sh && t_col >= 0 && t_col < sw && t_row >= 0 &&
                t_row < sh && pt_index_dem >= 0 && pt_index_dem < numofpts && tt_col >= 0 && tt_col < W &&
                tt_row >= 0 && tt_row < Hg) {
                if (GridPT3[pt_index_dem].Height != -1000) {
                    set_vec_kernel_value(pv, H, &Lv->view[0], &Lv->view[1], row, col, cd[0][2 * pt_index_temp],
                                         cd[0][2 * pt_index_temp + 1], cd[1][2 * pt_index_temp],
                                         cd[1][2 * pt_index_temp + 1], radius2, Count_N_ortho);
                    if (combined) {
                        t_col = (long)((t_X - Lv->boundary[0]) / im_res_next);
                        t_row = (long)((t_Y - Lv->boundary[1]) / im_res_next);
                        pt_index_temp = t_row * swn + t_col;
                        if (pt_index_temp >= 0 && pt_index_temp < swn * shn && t_col < swn && t_row < shn)
                            set_vec_kernel_value(pvn, H, &Lv->next[0], &Lv->next[1], row, col,
                                                 cdn[0][2 * pt_index_temp], cdn[0][2 * pt_index_temp + 1],
                                                 cdn[1][2 * pt_index_temp], cdn[1][2 * pt_index_temp + 1], radius2,
                                                 Count_N_ortho_next);
                    }
                }
            }
        }
    }
    compute_multi_ncc(pv, Count_N_ortho, &count_GNCC, sum);
    if (combined) compute_multi_ncc(pvn, Count_N_ortho_next, &count_GNCC, sum);
    return count_GNCC;
}


/* setsm_code.cpp:6672-6752 FindPeakNcc: one call per evaluated height, in
 * height order, with the caller's check_rho / pre_rho / pre_height /
 * direction / max_WNCC kept in st. (The Pyramid_step / iteration branches of
 * the reference all apply the same peak test.) */
void orc_peak_state_init(orc_peak_state *st)
{
    st->check_rho = 0;
    st->pre_rho = -1.0;
    st->pre_height = 0.0f;
    st->direction = 0;
    st->max_wncc = -100.0;
}

void orc_find_peak_ncc(double temp_rho, float iter_height, orc_peak_state *st, smx_nccresult *nr)
{
    double diff_rho;
    int t_direction;
    if (temp_rho >= 0 && st->pre_rho >= 0) diff_rho = temp_rho - st->pre_rho;
    else if (temp_rho < 0 && st->pre_rho < 0) diff_rho = -(fabs(temp_rho) - fabs(st->pre_rho));
    else if (temp_rho >= 0 && st->pre_rho < 0) diff_rho = 1;
    else diff_rho = -1;
    if (diff_rho > 0) t_direction = 1;
    else if (diff_rho < 0) t_direction = -1;
    else t_direction = 0;
    if (!st->check_rho) {
        st->check_rho = 1;
        t_direction = -1;
    } else if (st->pre_rho != -1) {
        if (st->direction > 0 && t_direction < 0) {
            if (nr->result4 < 10) nr->result4 += 1;
            if (nr->result0 < q_result(st->pre_rho)) {
                double temp_1 = dq(nr->result0);
                nr->result0 = q_result(st->pre_rho);
                float temp_2 = nr->result2;
                nr->result2 = st->pre_height;
                if (nr->result1 < q_result(temp_1)) {
                    nr->result1 = q_result(temp_1);
                    nr->result3 = temp_2;
                }
            } else {
                if (nr->result1 < q_result(st->pre_rho)) {
                    nr->result1 = q_result(st->pre_rho);
                    nr->result3 = st->pre_height;
                }
            }
        }
    }
    st->pre_rho = temp_rho;
    st->pre_height = iter_height;
    st->direction = t_direction;
    if (st->max_wncc < temp_rho) st->max_wncc = temp_rho;
}

/* --------------------------------------------------- VerticalLineLocus */

long orc_vertical_line_locus(const smx_level *Lv, orc_voxels *vox, smx_nccresult *nccresult,
                             const smx_ugrid *GridPT3, int iteration, const double minmaxHeight[2])
{
    /* setsm_code.cpp:6049-6548 */
    const int check_matchtag = Lv->check_matchtag;
    const int IsRA = Lv->is_ra;
    long Accessable_grid = 0;
    const int Pyramid_step = Lv->level;
    const int Half_template_size = orc_vll_half_template(Pyramid_step, Lv->template_size, Lv->resolution);
    const long numofpts = (long)Lv->grid_w * (long)Lv->grid_h;

    int check_combined_WNCC = 0, check_combined_WNCC_INCC = 0;
    if (!IsRA && (Pyramid_step > Lv->py_combined_level)) check_combined_WNCC = 1;
    if (check_combined_WNCC && (Pyramid_step >= 1 && iteration % 2 == 0)) check_combined_WNCC_INCC = 1;

    int check_ortho;
    if ((Pyramid_step == 4 && iteration == 1) || IsRA) check_ortho = 0;
    else check_ortho = 1;
    if (Lv->pre_dem && !(Pyramid_step == 4 && iteration == 1)) check_ortho = 1;
    if (check_matchtag) check_ortho = 1;
    if (Pyramid_step <= Lv->sgm_py) check_ortho = 0;

    const double ortho_th = 0.7 - (4 - Pyramid_step) * 0.10;
    const int ti_slot = 1 - Lv->reference_id;     /* GridPT3.ortho_ncc[ti] of the target image */
    const double ncc_alpha = orc_ncc_alpha(Pyramid_step, iteration, IsRA);
    const double ncc_beta = 1.0 - ncc_alpha;
    const double im_resolution = Lv->resolution * (double)(1 << Pyramid_step);
    const double im_resolution_next = check_combined_WNCC ? Lv->resolution * (double)(1 << (Pyramid_step - 1))
                                                          : Lv->resolution;
    /* setsm_code.cpp:6121-6129 ortho image coordinates, when some cell can
     * take the GNCC term (setsm_code.cpp:6240) */
    int gncc_on = 0;
    if (check_ortho)
        for (long c = 0; c < numofpts && !gncc_on; c++)
            if (dq(GridPT3[c].ortho_ncc[ti_slot]) > ortho_th) gncc_on = 1;
    long sw = 0, sh = 0, swn = 0, shn = 0;
    double *cd[2] = {NULL, NULL}, *cdn[2] = {NULL, NULL};
    if (gncc_on) {
        if (!Lv->ortho_wgs) return -2;
        sw = (long)((Lv->boundary[2] - Lv->boundary[0]) / im_resolution) + 1;
        sh = (long)((Lv->boundary[3] - Lv->boundary[1]) / im_resolution) + 1;
        if (sw != (long)Lv->ortho_w || sh != (long)Lv->ortho_h) return -4;
        if (check_combined_WNCC) {
            if (!Lv->next[0].img || !Lv->next[1].img) return -3;
            swn = (long)((Lv->boundary[2] - Lv->boundary[0]) / im_resolution_next) + 1;
            shn = (long)((Lv->boundary[3] - Lv->boundary[1]) / im_resolution_next) + 1;
        }
        for (int k = 0; k < 2; k++) {
            cd[k] = (double *)calloc((size_t)(2 * sw * sh), sizeof(double));
            if (check_combined_WNCC) cdn[k] = (double *)calloc((size_t)(2 * swn * shn), sizeof(double));
        }
        set_ortho_image_coord(Lv, GridPT3, check_combined_WNCC, im_resolution, im_resolution_next, sw, sh, swn, shn,
                              cd, cdn);
    }
    if (check_combined_WNCC_INCC && (!Lv->next[0].img || !Lv->next[1].img)) return -3;

    const double height_step = Lv->height_step;
    const int check_matching_rate = Lv->check_matching_rate;

#pragma omp parallel num_threads(orc_get_threads())
    {
        patchvec pv, pvn;
        pv_init(&pv, Half_template_size);
        pv_init(&pvn, Half_template_size);

#pragma omp for schedule(dynamic, 1) reduction(+ : Accessable_grid)
        for (long pt_index = 0; pt_index < numofpts; pt_index++) {
            const int start_H = GridPT3[pt_index].minHeight;
            const int end_H = GridPT3[pt_index].maxHeight;
            smx_nccresult *nr = &nccresult[pt_index];

            if (!orc_check_image_boundary(Lv, pt_index, start_H, end_H, Half_template_size))
                continue;

            int check_blunder_cell = 0;
            double th_height = 1000;
            if (Lv->dem_resolution <= 4) th_height = 500;
            if (Pyramid_step >= 2)
                check_blunder_cell = 0;
            else {
                if (GridPT3[pt_index].Matched_flag == 0)
                    check_blunder_cell = 1;
                else {
                    if (abs(GridPT3[pt_index].maxHeight - GridPT3[pt_index].minHeight) < th_height)
                        check_blunder_cell = 0;
                    else
                        check_blunder_cell = 1;
                }
            }
            if (Pyramid_step <= 1) {
                if (GridPT3[pt_index].maxHeight > minmaxHeight[1] || GridPT3[pt_index].minHeight < minmaxHeight[0])
                    check_blunder_cell = 1;
            }
            if (check_matchtag && iteration <= 2) {
                check_blunder_cell = 0;
                if (GridPT3[pt_index].maxHeight > minmaxHeight[1] || GridPT3[pt_index].minHeight < minmaxHeight[0])
                    check_blunder_cell = 1;
                else if (abs(GridPT3[pt_index].maxHeight - GridPT3[pt_index].minHeight) > th_height)
                    check_blunder_cell = 1;
            }
            if (check_blunder_cell)
                continue;

            /* ti = 1 only (stereo pair) */
            Accessable_grid++;
            double sum_GNCC_multi = 0, count_GNCC = 0;
            if (check_ortho && dq(GridPT3[pt_index].ortho_ncc[ti_slot]) > ortho_th) {
                count_GNCC = gncc_cell(Lv, GridPT3, pt_index, Half_template_size, check_combined_WNCC, im_resolution,
                                       im_resolution_next, sw, sh, swn, shn, cd, cdn, &pv, &pvn, &sum_GNCC_multi);
                if (count_GNCC > 0)
                    nr->GNCC = q_result(sum_GNCC_multi / count_GNCC);
                else
                    nr->GNCC = q_result(-1.0);
            } else
                nr->GNCC = q_result(-1.0);

            const int NumOfHeights = (int)((end_H - start_H) / height_step);
            if (IsRA || check_matching_rate) {
                nr->check_height_change = 1;
                nr->NumOfHeight = (uint16_t)NumOfHeights;
            }

            if (IsRA || check_matching_rate || nr->check_height_change || (Pyramid_step == 0 && Lv->dem_resolution < 2)) {
                double max_WNCC = -100;
                nr->result0 = q_result(-1.0);
                nr->result2 = -1000;
                nr->result1 = q_result(-1.0);
                nr->result3 = -1000;
                nr->result4 = 0;

                /* FindPeakNcc state (check_matching_rate mode only) */
                orc_peak_state peak;
                orc_peak_state_init(&peak);

                for (int hidx = 0; hidx < nr->NumOfHeight; hidx++) {
                    float iter_height;
                    if (IsRA || check_matching_rate)
                        iter_height = (float)(start_H + hidx * height_step);
                    else
                        iter_height = (float)(nr->minHeight + hidx * height_step);
                    if (!(iter_height >= start_H && iter_height <= end_H))
                        continue;

                    const smx_view *Lw = &Lv->view[0], *Rw = &Lv->view[1];
                    const double lon = Lv->grid_wgs[2 * pt_index], lat = Lv->grid_wgs[2 * pt_index + 1];
                    double rix, riy, tix, tiy, rpx, rpy, tpx, tpy;
                    orc_rpc_ground_to_image(&Lv->rpc[0], Lv->adjust[0], lon, lat, (double)iter_height, &rix, &riy);
                    orc_rpc_ground_to_image(&Lv->rpc[1], Lv->adjust[1], lon, lat, (double)iter_height, &tix, &tiy);
                    to_pyramid(rix, riy, Lw->start_x, Lw->start_y, Pyramid_step, &rpx, &rpy);
                    to_pyramid(tix, tiy, Rw->start_x, Rw->start_y, Pyramid_step, &tpx, &tpy);

                    const int check_py_image_pt =
                        (int)rpy >= 0 && (int)rpy + 1 < (long)Lw->height && (int)rpx >= 0 && (int)rpx + 1 < (long)Lw->width &&
                        (int)tpy >= 0 && (int)tpy + 1 < (long)Rw->height && (int)tpx >= 0 && (int)tpx + 1 < (long)Rw->width;

                    int check_py_image_pt_next = 1;
                    double rnx = 0, rny = 0, tnx = 0, tny = 0;
                    const smx_view *Ln = &Lv->next[0], *Rn = &Lv->next[1];
                    if (check_combined_WNCC_INCC) {
                        to_pyramid(rix, riy, Ln->start_x, Ln->start_y, Pyramid_step - 1, &rnx, &rny);
                        to_pyramid(tix, tiy, Rn->start_x, Rn->start_y, Pyramid_step - 1, &tnx, &tny);
                        check_py_image_pt_next =
                            (int)rny >= 0 && (int)rny + 1 < (long)Ln->height && (int)rnx >= 0 && (int)rnx + 1 < (long)Ln->width &&
                            (int)tny >= 0 && (int)tny + 1 < (long)Rn->height && (int)tnx >= 0 && (int)tnx + 1 < (long)Rn->width;
                    }
                    if (!(check_py_image_pt && check_py_image_pt_next))
                        continue;

                    const int ori_diff = (int)Lw->ori[(int)rpy * Lw->width + (int)rpx] - (int)Rw->ori[(int)tpy * Rw->width + (int)tpx];
                    int ori_diff_next = 0;
                    if (check_combined_WNCC_INCC)
                        ori_diff_next = (int)Ln->ori[(int)rny * Ln->width + (int)rnx] - (int)Rn->ori[(int)tny * Rn->width + (int)tnx];
                    double diff_theta = (double)ori_diff;
                    if (check_combined_WNCC_INCC)
                        diff_theta = (diff_theta + (double)ori_diff_next) / 2.0;

                    int check_orientation = 0;
                    if (IsRA != 1) {
                        if (Pyramid_step >= 3 && diff_theta * Lv->bin_angle < 90 && diff_theta * Lv->bin_angle > -90)
                            check_orientation = 1;
                        else if (Pyramid_step <= 2)
                            check_orientation = 1;
                    } else
                        check_orientation = 1;

                    if (!check_orientation) {
                        if (!check_matching_rate && !IsRA)
                            vox->rows[pt_index][hidx] &= ~0xffu;   /* flag_cal = false */
                        continue;
                    }

                    double sum_INCC_multi = 0, count_INCC = 0;
                    int Count_N[3] = {0, 0, 0}, Count_N_next[3] = {0, 0, 0};
                    const double rot_theta = (double)(diff_theta * Lv->bin_angle * ORC_PI / 180.0);
                    const double cos0 = cos(-rot_theta);
                    const double sin0 = sin(-rot_theta);
                    const int H = Half_template_size;
                    for (int row = -H; row <= H; row++) {
                        for (int col = -H; col <= H; col++) {
                            int radius2 = row * row + col * col;
                            if (radius2 <= (H + 1) * (H + 1)) {
                                double lx = rpx + col, ly = rpy + row;
                                double ox = cos0 * col - sin0 * row, oy = sin0 * col + cos0 * row;
                                double rx = tpx + ox, ry = tpy + oy;
                                set_vec_kernel_value(&pv, H, Lw, Rw, row, col, lx, ly, rx, ry, radius2, Count_N);
                                if (check_combined_WNCC_INCC) {
                                    double lnx = rnx + col, lny = rny + row;
                                    double onx = cos0 * col - sin0 * row, ony = sin0 * col + cos0 * row;
                                    double rnx2 = tnx + onx, rny2 = tny + ony;
                                    set_vec_kernel_value(&pvn, H, Ln, Rn, row, col, lnx, lny, rnx2, rny2, radius2, Count_N_next);
                                }
                            }
                        }
                    }

                    compute_multi_ncc(&pv, Count_N, &count_INCC, &sum_INCC_multi);
                    if (Count_N[0] > 0 && Count_N[1] > 0 && Count_N[2] > 0) {
                        if (!check_matching_rate && !IsRA)
                            vox->rows[pt_index][hidx] |= 1u;
                    }
                    if (check_combined_WNCC_INCC)
                        compute_multi_ncc(&pvn, Count_N_next, &count_INCC, &sum_INCC_multi);

                    if (!(IsRA || check_matching_rate)) {
                        int16_t incc = (count_INCC > 0) ? q_voxel(sum_INCC_multi / count_INCC) : q_voxel(-1.0);
                        vox->rows[pt_index][hidx] = (vox->rows[pt_index][hidx] & 0xffffu) | ((uint32_t)(uint16_t)incc << 16);
                        if (vox->rho)
                            vox->rho[pt_index][hidx] = (count_INCC > 0) ? sum_INCC_multi / count_INCC : -1.0;
                    }

                    double temp_rho;
                    if (check_ortho && count_GNCC > 0)   /* setsm_code.cpp:6482-6483 */
                        temp_rho = sum_INCC_multi / count_INCC * ncc_alpha + sum_GNCC_multi / count_GNCC * ncc_beta;
                    else if (count_INCC > 0)
                        temp_rho = sum_INCC_multi / count_INCC;
                    else
                        temp_rho = -1.0;
                    if (max_WNCC < temp_rho) max_WNCC = temp_rho;

                    if (check_matching_rate) orc_find_peak_ncc(temp_rho, iter_height, &peak, nr);
                }
                if (nr->check_height_change)
                    nr->max_WNCC = q_result(max_WNCC);
            }
        }
        pv_free(&pv);
        pv_free(&pvn);
    }
    for (int k = 0; k < 2; k++) {
        free(cd[k]);
        free(cdn[k]);
    }
    return Accessable_grid;
}

/* -------------------------------------------------------- AWNCC peak */

static void awncc_impl(const smx_level *Lv, const orc_voxels *vox, const smx_ugrid *GridPT3,
                       smx_nccresult *nccresult, int iteration, float *const *SumCost)
{
    /* setsm_code.cpp:7511-7702; SumCost != NULL is check_SGM_peak (Pyramid_step <= SGM_py) */
    const int Pyramid_step = Lv->level;
    const double step_height = Lv->height_step;
    int check_ortho;
    if (Pyramid_step == 4 && iteration == 1) check_ortho = 0;
    else check_ortho = 1;
    if (Lv->pre_dem && !(Pyramid_step == 4 && iteration == 1)) check_ortho = 1;
    const double ncc_alpha = orc_ncc_alpha(Pyramid_step, iteration, Lv->is_ra);
    const double ncc_beta = 1.0 - ncc_alpha;
    const long n = (long)Lv->grid_w * (long)Lv->grid_h;

#pragma omp parallel for schedule(guided) num_threads(orc_get_threads())
    for (long pt_index = 0; pt_index < n; pt_index++) {
        smx_nccresult *nr = &nccresult[pt_index];
        double pre_rho = -1.0, pre_rho_WNCC = -1.0;
        float pre_height = 0.0f;
        int direction = 0;
        int check_rho = 0;
        double max_roh = 0;
        double temp_nccresult = -100.0, temp_nccresult_sec = -100.0;

        nr->result0 = q_result(-1.0);
        nr->result2 = -1000;
        nr->result1 = q_result(-1.0);
        nr->result3 = -1000;
        nr->result4 = 0;

        for (long hs = 0; hs < nr->NumOfHeight; hs++) {
            float iter_height = (float)(nr->minHeight + hs * step_height);
            if (!(iter_height >= GridPT3[pt_index].minHeight && iter_height <= GridPT3[pt_index].maxHeight))
                continue;
            double WNCC_sum;
            const double incc = dq(SMX_VOXEL_INCC(vox->rows[pt_index][hs]));
            if (check_ortho && dq(nr->GNCC) > -1.0)
                WNCC_sum = incc * ncc_alpha + dq(nr->GNCC) * ncc_beta;
            else
                WNCC_sum = incc;
            double WNCC_temp_rho = WNCC_sum;
            double temp_rho = SumCost ? (double)SumCost[pt_index][hs] : WNCC_sum;

            double diff_rho;
            int t_direction;
            if (temp_rho >= 0 && pre_rho >= 0) diff_rho = temp_rho - pre_rho;
            else if (temp_rho < 0 && pre_rho < 0) diff_rho = -(fabs(temp_rho) - fabs(pre_rho));
            else if (temp_rho >= 0 && pre_rho < 0) diff_rho = 1;
            else diff_rho = -1;
            if (diff_rho > 0) t_direction = 1;
            else if (diff_rho < 0) t_direction = -1;
            else t_direction = 0;

            if (!check_rho) {
                check_rho = 1;
                t_direction = -1;
            } else if (pre_rho != -1) {
                int check_peak = (direction > 0 && t_direction < 0);
                if ((iteration <= 3 && Pyramid_step == 4) || (iteration <= 2 && Pyramid_step == 3) ||
                    (iteration <= 1 && Pyramid_step == 2)) {
                    if (temp_nccresult < temp_rho) {
                        if (nr->result4 < 10) nr->result4 += 1;
                        temp_nccresult = temp_rho;
                        nr->result2 = iter_height;
                        temp_nccresult_sec = -1.0;
                        nr->result3 = -9999;
                        max_roh = WNCC_temp_rho;
                    }
                } else if (check_peak) {
                    if (nr->result4 < 10) nr->result4 += 1;
                    if (temp_nccresult < pre_rho) {
                        double temp_1 = temp_nccresult;
                        temp_nccresult = pre_rho;
                        float temp_2 = nr->result2;
                        nr->result2 = pre_height;
                        max_roh = pre_rho_WNCC;
                        if (temp_nccresult_sec < temp_1) {
                            temp_nccresult_sec = temp_1;
                            nr->result3 = temp_2;
                        }
                    } else {
                        if (temp_nccresult_sec < pre_rho) {
                            temp_nccresult_sec = pre_rho;
                            nr->result3 = pre_height;
                        }
                    }
                }
            }
            pre_rho = temp_rho;
            pre_rho_WNCC = WNCC_temp_rho;
            pre_height = iter_height;
            direction = t_direction;
        }
        if (SumCost && Pyramid_step == 0 && iteration >= 2) {
            nr->result0 = q_result(max_roh);
            nr->result1 = q_result(-1.0);
            nr->result3 = -9999;
        } else {
            nr->result0 = q_result(temp_nccresult);
            nr->result1 = q_result(temp_nccresult_sec);
        }
    }
}


void orc_awncc_peak(const smx_level *Lv, const orc_voxels *vox, const smx_ugrid *GridPT3,
                    smx_nccresult *nccresult, int iteration)
{
    awncc_impl(Lv, vox, GridPT3, nccresult, iteration, NULL);
}

/* ------------------------------------------------------------ SGM */

/* setsm_code.cpp:6755-6767 SGM_start_pos */
static void sgm_start_pos(const smx_nccresult *ncc, const orc_voxels *vox, const smx_ugrid *G, long pt,
                          float *LHcost_pre, float *const *SumCost, double step)
{
    for (int hs = 0; hs < ncc[pt].NumOfHeight; hs++) {
        float iter_height = (float)(ncc[pt].minHeight + hs * step);
        if (iter_height >= G[pt].minHeight && iter_height <= G[pt].maxHeight) {
            double WNCC_sum = dq(SMX_VOXEL_INCC(vox->rows[pt][hs]));
            LHcost_pre[hs] = (float)WNCC_sum;
            SumCost[pt][hs] += LHcost_pre[hs];
        }
    }
}

static double max2(double a, double b) { return a > b ? a : b; }

/* setsm_code.cpp:6769-6999 SGM_con_pos: the previous cell on the path is
 * (row + v_row[d], col + u_col[d]); every branch reduces to
 * max(candidates) - maxWNCC where the candidates are the previous cell's path
 * costs at the same height (V3) and one step below / above minus P1 (V1, V2),
 * each only if that voxel was computed, and maxWNCC - P2 (V4). */
static void sgm_con_pos(int col, int row, int W, int d, double step, int P_HS_step, const int *u_col,
                        const int *v_row, const smx_nccresult *ncc, const orc_voxels *vox, const smx_ugrid *G,
                        long pt, double P1, double P2, const float *LHcost_pre, float *LHcost_curr,
                        float *const *SumCost)
{
    for (int hs = 0; hs < ncc[pt].NumOfHeight; hs++) {
        const float iter_height = (float)(ncc[pt].minHeight + hs * step);
        double WNCC_sum = 0;
        if (!(iter_height >= G[pt].minHeight && iter_height <= G[pt].maxHeight)) continue;
        WNCC_sum = dq(SMX_VOXEL_INCC(vox->rows[pt][hs]));
        double t_WNCC_sum;
        const int t_col = col + u_col[d];
        const int t_row = row + v_row[d];
        const int t_index = t_row * W + t_col;
        const float maxWNCC = (float)dq(ncc[t_index].max_WNCC);
        const int nt = ncc[t_index].NumOfHeight;
        const int i0 = (int)((iter_height - (float)ncc[t_index].minHeight) / step);
        const int i1 = i0 - P_HS_step;
        const int i2 = i0 + P_HS_step;
        const uint32_t *tv = vox->rows[t_index];
        const double V4 = maxWNCC - P2;
        int use1 = 0, use2 = 0, use3 = 0, zero3 = 0;
        if (i0 == 0 && i2 >= 0 && i2 < nt) {
            use3 = SMX_VOXEL_FLAG(tv[i0]) != 0;
            use2 = SMX_VOXEL_FLAG(tv[i2]) != 0;
        } else if (i0 == nt - 1 && i1 >= 0 && i1 < nt) {
            use3 = SMX_VOXEL_FLAG(tv[i0]) != 0;
            use1 = SMX_VOXEL_FLAG(tv[i1]) != 0;
        } else if (i0 > 0 && i0 < nt - 1 && i1 >= 0 && i1 < nt && i2 >= 0 && i2 < nt) {
            use3 = SMX_VOXEL_FLAG(tv[i0]) != 0;
            use1 = SMX_VOXEL_FLAG(tv[i1]) != 0;
            use2 = SMX_VOXEL_FLAG(tv[i2]) != 0;
            /* setsm_code.cpp:6917-6929: with both neighbours' voxels computed
             * but not the centre one, the reference still compares V3, which
             * holds its initial 0. */
            zero3 = !use3 && use1 && use2;
        }
        /* The reference's branches take the maximum in the order V1, V2, V3,
         * V4 with strict '>'; max over a set is order-free for values that
         * are never NaN, so the result is the same double. */
        double m = V4;
        if (use1) m = max2(LHcost_pre[i1] - P1, m);
        if (use2) m = max2(LHcost_pre[i2] - P1, m);
        if (use3) m = max2((double)LHcost_pre[i0], m);
        if (zero3) m = max2(0.0, m);
        t_WNCC_sum = m - maxWNCC;
        LHcost_curr[hs] = (float)(WNCC_sum + t_WNCC_sum);
        SumCost[pt][hs] += LHcost_curr[hs];
    }
}

/* One path: start cell (row, col), steps (dr, dc) until it leaves the grid. */
static void sgm_path(int row, int col, int dr, int dc, int d, const smx_level *Lv, const orc_voxels *vox,
                     const smx_ugrid *G, const smx_nccresult *ncc, float *const *SumCost, float *pre, float *cur)
{
    static const int v_row[8] = {0, 0, -1, 1, -1, -1, 1, 1};
    static const int u_col[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
    const int W = (int)Lv->grid_w, H = (int)Lv->grid_h;
    const double step = Lv->height_step;
    long pt = (long)row * W + col;
    memset(pre, 0, ncc[pt].NumOfHeight * sizeof(float));
    sgm_start_pos(ncc, vox, G, pt, pre, SumCost, step);
    for (row += dr, col += dc; row >= 0 && row < H && col >= 0 && col < W; row += dr, col += dc) {
        pt = (long)row * W + col;
        memset(cur, 0, ncc[pt].NumOfHeight * sizeof(float));
        sgm_con_pos(col, row, W, d, step, 1, u_col, v_row, ncc, vox, G, pt, 0.3, 0.6, pre, cur, SumCost);
        float *t = pre;
        pre = cur;
        cur = t;
    }
}

/* setsm_code.cpp:7017-7509: the 8 path directions in the reference's order.
 * d 0..3: left->right, right->left, top->bottom, bottom->top, one path per
 * row / column. d 4..7 (diagonals): a loop over paths starting on the left or
 * right wall (one per row), then a loop over paths starting on the top or
 * bottom wall (one per column); the corner path is in both loops, so its
 * cells receive that direction twice, as in the reference. Paths of one loop
 * are disjoint, so the loop runs in parallel like the reference's omp for. */
void orc_sgm_aggregate_rows(const smx_level *Lv, const orc_voxels *vox, const smx_ugrid *G,
                            const smx_nccresult *ncc, float *const *SumCost)
{
    const int W = (int)Lv->grid_w, H = (int)Lv->grid_h;
    const long n = (long)W * H;
    int maxH = 0;
    for (long c = 0; c < n; c++)
        if (maxH < ncc[c].NumOfHeight) maxH = ncc[c].NumOfHeight;
    /* loop kinds: 0 = one path per row, 1 = one path per column */
    struct { int d, kind, dr, dc; } loops[12] = {
        {0, 0, 0, 1}, {1, 0, 0, -1}, {2, 1, 1, 0}, {3, 1, -1, 0},
        {4, 0, 1, 1}, {4, 1, 1, 1}, {5, 0, 1, -1}, {5, 1, 1, -1},
        {6, 0, -1, 1}, {6, 1, -1, 1}, {7, 0, -1, -1}, {7, 1, -1, -1},
    };
    for (int k = 0; k < 12; k++) {
        const int d = loops[k].d, dr = loops[k].dr, dc = loops[k].dc;
        const int npaths = loops[k].kind == 0 ? H : W;
#pragma omp parallel num_threads(orc_get_threads())
        {
            float *pre = (float *)calloc((size_t)maxH + 1, sizeof(float));
            float *cur = (float *)calloc((size_t)maxH + 1, sizeof(float));
#pragma omp for
            for (int p = 0; p < npaths; p++) {
                int row, col;
                if (loops[k].kind == 0) {          /* starts on the left (dc>0) or right (dc<0) wall, or a row start */
                    row = p;
                    col = dc > 0 ? 0 : (dc < 0 ? W - 1 : 0);
                } else {                           /* starts on the top (dr>0) or bottom (dr<0) wall */
                    col = p;
                    row = dr > 0 ? 0 : H - 1;
                }
                sgm_path(row, col, dr, dc, d, Lv, vox, G, ncc, SumCost, pre, cur);
            }
            free(pre);
            free(cur);
        }
    }
}

void orc_sgm_aggregate(const smx_level *Lv, const orc_voxels *vox, const smx_ugrid *G, const smx_nccresult *ncc,
                       float *sum_csr)
{
    const long n = (long)Lv->grid_w * Lv->grid_h;
    float **rows = (float **)calloc((size_t)n, sizeof(float *));
    long off = 0;
    for (long c = 0; c < n; c++) {
        rows[c] = sum_csr + off;
        for (int h = 0; h < ncc[c].NumOfHeight; h++) rows[c][h] = 0.0f;
        off += ncc[c].NumOfHeight;
    }
    orc_sgm_aggregate_rows(Lv, vox, G, ncc, rows);
    free(rows);
}

void orc_awncc(const smx_level *Lv, const orc_voxels *vox, const smx_ugrid *G, smx_nccresult *ncc, int iteration,
               float *sum_csr)
{
    /* setsm_code.cpp:7511-7533: SGM aggregation when Pyramid_step <= SGM_py */
    if (Lv->level > Lv->sgm_py) {
        awncc_impl(Lv, vox, G, ncc, iteration, NULL);
        return;
    }
    const long n = (long)Lv->grid_w * Lv->grid_h;
    long total = 0;
    for (long c = 0; c < n; c++) total += ncc[c].NumOfHeight;
    float *own = sum_csr ? NULL : (float *)calloc((size_t)total + 1, sizeof(float));
    float *buf = sum_csr ? sum_csr : own;
    orc_sgm_aggregate(Lv, vox, G, ncc, buf);
    float **rows = (float **)calloc((size_t)n, sizeof(float *));
    long off = 0;
    for (long c = 0; c < n; c++) {
        rows[c] = buf + off;
        off += ncc[c].NumOfHeight;
    }
    awncc_impl(Lv, vox, G, ncc, iteration, rows);
    free(rows);
    free(own);
}

/* ------------------------------------------------------------ pyramid */

void orc_gauss_pyramid(const uint16_t *in, uint32_t w, uint32_t h, uint16_t *out)
{
    /* Template.hpp:45-78 CreateGaussianFilter(9, 1.5) + Template.hpp:82-132 */
    const int fs = 9, half = 4;
    const double sigma = 1.5;
    double G[9][9];
    double sum = 0;
    double tmp = -1 / (2 * sigma * sigma);
    double scale = 1 / (sqrt(2 * ORC_PI) * sigma);
    for (int i = -half; i <= half; i++)
        for (int j = -half; j <= half; j++) {
            double temp = (i * i + j * j) * tmp;
            double val = exp(temp) * scale;
            G[i + half][j + half] = val;
            sum += val;
        }
    for (int i = 0; i < fs; i++)
        for (int j = 0; j < fs; j++)
            G[i][j] /= sum;

    const long ow = w / 2, oh = h / 2;
#pragma omp parallel for schedule(guided) num_threads(orc_get_threads())
    for (long r = 0; r < oh; r++) {
        for (long c = 0; c < ow; c++) {
            double temp_v = 0;
            int count = 0;
            for (int l = -half; l <= half; l++)
                for (int k = -half; k <= half; k++) {
                    if ((2 * r + l) >= 0 && (2 * c + k) >= 0 && (2 * r + l) < (long)h && (2 * c + k) < (long)w) {
                        temp_v += G[l + half][k + half] * in[(2 * r + l) * (long)w + (2 * c + k)];
                        count++;
                    }
                }
            if (count == fs * fs)
                out[r * ow + c] = (uint16_t)temp_v;
            else
                out[r * ow + c] = in[(2 * r) * (long)w + (2 * c)];
        }
    }
}

void orc_sobel(const uint16_t *src, uint32_t w, uint32_t h, uint16_t *mag, int16_t *dir)
{
    /* SubFunctions.cpp:2418-2470 */
#pragma omp parallel for num_threads(orc_get_threads())
    for (long i = 0; i < (long)h; i++) {
        for (long j = 0; j < (long)w; j++) {
            const long W = (long)w;
            if (j == 0 || j == W - 1 || i == 0 || i == (long)h - 1) {
                mag[i * W + j] = 0;
                dir[i * W + j] = 0;
            } else {
                const int a00 = src[(i - 1) * W + (j - 1)], a01 = src[(i - 1) * W + j], a02 = src[(i - 1) * W + (j + 1)];
                const int a10 = src[i * W + (j - 1)], a11 = src[i * W + j], a12 = src[i * W + (j + 1)];
                const int a20 = src[(i + 1) * W + (j - 1)], a21 = src[(i + 1) * W + j], a22 = src[(i + 1) * W + (j + 1)];
                (void)a11;
                /* integer sums of the 3x3 masks, exactly as the int products in the reference */
                double gx = (double)((-1 * a00 + 0 * a01 + 1 * a02) + (-2 * a10 + 0 * a11 + 2 * a12) + (-1 * a20 + 0 * a21 + 1 * a22));
                double gy = (double)((1 * a00 + 2 * a01 + 1 * a02) + (0 * a10 + 0 * a11 + 0 * a12) + (-1 * a20 - 2 * a21 - 1 * a22));
                double t = sqrt(gx * gx + gy * gy);
                dir[i * W + j] = (int16_t)(floor(atan2(gy, gx) * 180 / ORC_PI));
                mag[i * W + j] = (uint16_t)(t + 0.5);
            }
        }
    }
}

static const int orc_numcols[7] = {6, 5, 5, 5, 4, 3, 0};   /* SubFunctions.cpp:8 */

void orc_orientation(const uint16_t *Gmag, const int16_t *Gdir, uint32_t w, uint32_t h, int template_size, uint8_t *plhs)
{
    /* SubFunctions.cpp:2473-2571. The column range starts at
     * max(-mask_row, -dcol) exactly as written in the reference; a linear
     * index that falls before the first pixel reads as magnitude 0. */
    const int half = template_size / 2;
    const long main_col = w, main_row = h;
    double gw[13][13];
    for (int row = -half + 1; row <= half - 1; row++)
        for (int col = -half + 1; col <= half - 1; col++)
            gw[half - 1 + row][half - 1 + col] = exp(-(double)(row * row + col * col) / 4.5);
    const double SUB_RATIO = 4.6;   /* SubFunctions.hpp:33 */

#pragma omp parallel for collapse(2) num_threads(orc_get_threads())
    for (long mask_row = 0; mask_row < main_row; mask_row++) {
        for (long mask_col = 0; mask_col < main_col; mask_col++) {
            double bin[18] = {0.0};
            double sub_bin[4] = {0.0};
            long rlo = -mask_row > -half + 1 ? -mask_row : -half + 1;
            long rhi = main_row - mask_row - 1 < half - 1 ? main_row - mask_row - 1 : half - 1;
            for (long row = rlo; row <= rhi; row++) {
                int dcol = orc_numcols[labs(row)];
                long clo = -mask_row > -dcol ? -mask_row : -dcol;
                long chi = dcol < main_col - mask_col - 1 ? dcol : main_col - mask_col - 1;
                for (long col = clo; col <= chi; col++) {
                    double gu_weight = gw[half - 1 + row][half - 1 + col];
                    long idx = (mask_row + row) * main_col + (mask_col + col);
                    double mag = idx >= 0 ? (double)Gmag[idx] : 0.0;
                    double theta = idx >= 0 ? (double)Gdir[idx] : 0.0;
                    if (theta < 0) theta += 360;
                    else if (theta > 360) theta -= 360;
                    long index = (long)(theta / 20.0);
                    if (index >= 0 && index < 18) {
                        bin[index] += mag * gu_weight;
                        sub_bin[(int)(index / SUB_RATIO)] += mag * gu_weight;
                    }
                }
            }
            double max_th = -100;
            int sub_th = 0, th = 0;
            for (int i = 0; i < 4; i++)
                if (sub_bin[i] > max_th) { max_th = sub_bin[i]; sub_th = i; }
            max_th = -100;
            for (int i = (int)(sub_th * SUB_RATIO); i < (int)(sub_th * SUB_RATIO + SUB_RATIO); i++)
                if (bin[i] > max_th) { max_th = bin[i]; th = i; }
            plhs[mask_row * main_col + mask_col] = (uint8_t)th;
        }
    }
}

/* ------------------------------------------------------------------------
 * Between iterations: SelectMPs and SetHeightRange
 * ---------------------------------------------------------------------- */

/* setsm_code.cpp:8546-8614: the level's minimum peak-rho threshold from the
 * histogram of first-peak rho (cells with more than two heights). Returns a
 * negative value where the reference leaves the threshold undefined (a level
 * above 4 that is not an SGM level). */
static double mps_minimum_th(const smx_level *L, const smx_nccresult *ncc, int iteration)
{
    const int level = L->level;
    if (L->is_ra || level <= 0) return 0.2;
    long hist[100] = {0};
    long total = 0;
    const long n = (long)L->grid_w * (long)L->grid_h;
    for (long c = 0; c < n; c++) {
        if (ncc[c].NumOfHeight <= 2) continue;
        int r = (int)ceil(dq(ncc[c].result0) * 100);
        if (r > 99) r = 99;
        if (r >= 0) hist[r]++;
        total++;
    }
    if (total == 0) return 0.2;
    double frac;
    int defined = 1;
    if (level == 4) frac = 0.05 + (iteration - 1) * 0.01;
    else if (level == 3) frac = 0.15 + (iteration - 1) * 0.01;
    else if (level == 2) frac = 0.60;
    else if (level == 1) frac = 0.80;
    else { frac = 0; defined = 0; }
    if (level <= L->sgm_py) {
        frac = 0.95 - level * 0.1;
        if (frac < 0.5) frac = 0.5;
        defined = 1;
    }
    if (!defined) return -1.0;
    double th = 0.2;
    int acc = 0;
    for (int r = 99; r > 0; r--) {
        acc += (int)hist[r];
        if (acc / (double)total > frac) {
            th = (double)r / 100.0;
            break;
        }
    }
    return th > 0.95 ? 0.95 : th;
}

long orc_select_mps(const smx_level *L, const smx_nccresult *ncc, smx_ugrid *g, const smx_mps_args *a,
                    smx_point3 *pts, double *minimum_th)
{
    const int level = L->level;
    const int it = a->iteration;
    const double th = mps_minimum_th(L, ncc, it);
    if (minimum_th) *minimum_th = th;
    if (th < 0) return -1;
    /* setsm_code.cpp:8618-8643 */
    const int iter_end = (level <= 3 || L->is_ra) ? (a->th_roh - 0.10 < a->th_roh_min)
                                                   : (a->th_roh - 0.06 < a->th_roh_min);
    const double roh_th = 0.05;
    const long n = (long)L->grid_w * (long)L->grid_h;
    long count = 0;
    for (long c = 0; c < n; c++) {
        const smx_nccresult *r = &ncc[c];
        if (r->NumOfHeight <= 2) continue;
        smx_ugrid *G = &g[c];
        if (level == 4 && it == 1) G->Height = -1000.0f;
        const double r0 = dq(r->result0), r1 = dq(r->result1);
        /* setsm_code.cpp:8665-8702: ratio of the two peaks, peak tests */
        double ror;
        if ((it <= 2 && level >= 3) || (it <= 1 && level == 2)) ror = 1.0;
        else ror = fabs(r0) > 0 ? (r0 - r1) / r0 : 0;
        const double ror_min = level == 0 ? 0.1 - 0.03 * (3 - level) : 0.1;
        const int peak_ok = ror >= ror_min && r0 > th;
        const int above_prev = r0 > dq(G->roh) - roh_th;
        const int has_peak = r->result4 > 0;
        /* setsm_code.cpp:8710-8747 */
        int take;
        if (level >= 2 || (level >= 1 && !(a->mpp_stereo_angle > 5)) || (level == 0 && a->final_level_iteration < 3))
            take = has_peak && peak_ok && above_prev;
        else if (level >= 1)
            take = (above_prev && peak_ok) || (above_prev && r0 > th);
        else
            take = peak_ok && above_prev;
        /* setsm_code.cpp:8749-8781: a matched cell whose second peak is strong
         * keeps both peaks in its range */
        if (G->Matched_flag != 0) {
            const double ror_lim = level <= 0 ? 0.1 - 0.03 * (3 - level) : 0.1;
            const int both = ror < ror_lim && r1 > dq(G->roh) - roh_th;
            if (both) {
                if (r->result2 > r->result3) {
                    G->minHeight = (int16_t)floor(r->result3 - 0.5);
                    G->maxHeight = (int16_t)ceil(r->result2 + 0.5);
                } else {
                    G->minHeight = (int16_t)floor(r->result2 - 0.5);
                    G->maxHeight = (int16_t)ceil(r->result3 + 0.5);
                }
                G->Matched_flag = 4;
            }
        }
        /* setsm_code.cpp:8783-8884 */
        const double z = (double)r->result2;
        int emit = 0;
        if (take) {
            if (level < 2) {
                const double pre_h = (double)G->Height;
                const double pre_range = (double)(G->maxHeight - G->minHeight);
                const int m = G->Matched_flag;
                const double tol2 = a->mpp * 2 * (1 << level), tol3 = a->mpp * 3 * (1 << level);
                const int near_prev = fabs(pre_h - z) <= tol2;
                const int in_range = fabs(pre_h - z) <= pre_range / 2.0;
                const int near_min = fabs(z - G->minHeight) <= tol3;
                const int near_max = fabs(z - G->maxHeight) <= tol3;
                const int keep = (m == 0) || (m == 2 && near_prev) ||
                                 ((m == 1 || m == 4) && (near_prev || in_range || near_min || near_max));
                if (keep && z > -100 && z < 10000) {
                    emit = 1;
                    G->roh = r->result0;
                    if (G->roh < q_result(th)) G->roh = q_result(th);
                } else {
                    G->roh = q_result(a->th_roh);
                    if (iter_end && a->th_roh_start == dq(G->roh)) G->roh = q_result(a->th_roh_next);
                }
            } else {
                emit = z > -100 && z < 10000;
                G->roh = r->result0;
                if (G->roh < q_result(th)) G->roh = q_result(th);
            }
        } else {
            G->roh = q_result(a->th_roh);
            if (iter_end && a->th_roh_start == dq(G->roh)) G->roh = q_result(a->th_roh_next);
        }
        if (emit) {
            smx_point3 *p = &pts[count++];
            memset(p, 0, sizeof *p);
            p->x = L->grid_pts[2 * c];
            p->y = L->grid_pts[2 * c + 1];
            p->z = z;
        }
    }
    return count;
}

/* std::min / std::max as the reference calls them: the first argument on a tie */
static double smin(double a, double b) { return b < a ? b : a; }
static double smax(double a, double b) { return a < b ? b : a; }

/* SubFunctions.cpp:2946-2995 IsTinInside: inside test by the three edge
 * cross products, then the height on the triangle's plane through its unit
 * normal. */
static int tin_inside(double px, double py, const smx_point3 *p1, const smx_point3 *p2, const smx_point3 *p3,
                      float *z)
{
    int sum = 3;
    if ((p2->x - p1->x) * (py - p1->y) - (p2->y - p1->y) * (px - p1->x) <= 0) sum--;
    if ((p3->x - p2->x) * (py - p2->y) - (p3->y - p2->y) * (px - p2->x) <= 0) sum--;
    if ((p1->x - p3->x) * (py - p3->y) - (p1->y - p3->y) * (px - p3->x) <= 0) sum--;
    if (sum != 0 && sum != 3) return 0;
    const double ux = p2->x - p1->x, uy = p2->y - p1->y, uz = p2->z - p1->z;
    const double vx = p3->x - p1->x, vy = p3->y - p1->y, vz = p3->z - p1->z;
    double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    const double len = sqrt(nx * nx + ny * ny + nz * nz);
    if (!(len > 0)) return 0;
    nx /= len;
    ny /= len;
    nz /= len;
    const double d = -(nx * p1->x + ny * p1->y + nz * p1->z);
    if (nz == 0) return 0;
    *z = (float)(-1.0 * ((nx * px) + (ny * py) + d) / nz);
    return 1;
}

/* SubFunctions.cpp:2997-3024 SetNormalAngle + SetAngle: slope of the
 * triangle in degrees. */
static double tin_slope_deg(const smx_point3 *p0, const smx_point3 *p1, const smx_point3 *p2)
{
    const double ux = p1->x - p0->x, uy = p1->y - p0->y, uz = p1->z - p0->z;
    const double vx = p2->x - p0->x, vy = p2->y - p0->y, vz = p2->z - p0->z;
    const double nx = uy * vz - vy * uz, ny = -(ux * vz - vx * uz), nz = ux * vy - vx * uy;
    const double norm = sqrt(nx * nx + ny * ny + nz * nz);
    double angle = 0;
    /* RadToDeg expands to 180/PI after the product: (acos * 180) / PI */
    if (norm != 0) angle = acos(fabs(nz) / norm) * 180 / 3.141592653589793;
    if (angle <= 0 && angle >= -90) angle = fabs(angle);
    else if (angle <= -270 && angle >= -360) angle = 360 + angle;
    else if (angle >= 270 && angle <= 360) angle = 360 - angle;
    return angle;
}

int orc_set_height_range(const smx_level *L, const smx_hr_args *a, const smx_point3 *pts, long npts,
                         const smx_tri *tris, long ntris, smx_ugrid *g, double *minH_grid, double *maxH_grid)
{
    const int level = L->level, it = a->iteration;
    const long W = (long)L->grid_w, H = (long)L->grid_h, n = W * H;
    const double gres = L->grid_resolution;
    /* setsm_code.cpp:9864-9923: height buffer and gap limit */
    double buf = a->mpp * 4.0 * (1 << level);
    if (level == 1) buf = it >= 2 ? a->mpp * 2 : a->mpp * 3;
    else if (level == 0) {
        buf = it == 1 ? a->mpp * 2 : a->mpp;
        if (buf < 0.5) buf = 0.5;
    }
    if (L->pre_dem && buf > a->seed_dem_sigma) buf = a->seed_dem_sigma;
    buf = ceil(buf);
    if (buf > 100) buf = 100;
    double th_hg = 100000;
    if (L->dem_resolution >= 4) {
        if (level <= 1) th_hg = 1000;
    } else {
        if (!a->check_matching_rate) {
            if (level == 2 && it >= 5) th_hg = 1000;
            else if (level == 1) th_hg = 500;
            else if (level == 0) th_hg = 100;
        } else if (level <= 2) th_hg = 100;
        if (level == 0 && L->dem_resolution < 2 && it > 1) th_hg = 50;
    }
    double ncc_th;
    if (level == 4) ncc_th = 0.7;
    else if (level == 3) ncc_th = 0.6 - (it - 1) * 0.02;
    else if (level == 2) ncc_th = 0.5 - (it - 1) * 0.02;
    else if (level == 1) ncc_th = 0.4;
    else ncc_th = 0.3;
    if (ncc_th < 0.3) ncc_th = 0.3;

    uint8_t *mb = (uint8_t *)calloc((size_t)n, 1);
    float *newh = (float *)malloc((size_t)n * sizeof(float));
    smx_ugrid *snap = (smx_ugrid *)malloc((size_t)n * sizeof(smx_ugrid));
    if (!mb || !newh || !snap) {
        free(mb);
        free(newh);
        free(snap);
        return -1;
    }
    for (long c = 0; c < n; c++) newh[c] = -1000.0f;
    double tmin = 100000, tmax = -100000;
    /* setsm_code.cpp:9928-10091: each triangle, in list order, claims the
     * cells of its bounding box that lie inside it and are still unclaimed */
    for (long t = 0; t < ntris; t++) {
        if (tris[t].a >= (uint32_t)npts || tris[t].b >= (uint32_t)npts || tris[t].c >= (uint32_t)npts) continue;
        const smx_point3 *p1 = &pts[tris[t].a], *p2 = &pts[tris[t].b], *p3 = &pts[tris[t].c];
        /* SubFunctions.cpp:2915-2944 SetTinBoundary */
        const double zmin = smin(smin(p1->z, p2->z), p3->z), zmax = smax(smax(p1->z, p2->z), p3->z);
        if (zmin < tmin) tmin = zmin;
        if (zmax > tmax) tmax = zmax;
        const double bx0 = smin(smin(p1->x, p2->x), p3->x), by0 = smin(smin(p1->y, p2->y), p3->y);
        const double bx1 = smax(smax(p1->x, p2->x), p3->x), by1 = smax(smax(p1->y, p2->y), p3->y);
        int c0 = (int)((bx0 - L->boundary[0]) / gres + 0.5) - 1, r0 = (int)((by0 - L->boundary[1]) / gres + 0.5) - 1;
        int c1 = (int)((bx1 - L->boundary[0]) / gres + 0.5) + 1, r1 = (int)((by1 - L->boundary[1]) / gres + 0.5) + 1;
        if (c1 >= (int)W) c1 = (int)W - 1;
        if (r1 >= (int)H) r1 = (int)H - 1;
        if (c0 < 0) c0 = 0;
        if (r0 < 0) r0 = 0;
        const double slope = tin_slope_deg(p1, p2, p3);
        const double wgt = slope <= 30 ? 1.0 + 0.05 * slope / 5.0 : 1.3;
        const double bf = level >= 3 ? ceil(buf * wgt) : ceil(buf);
        for (long row = r0; row <= r1; row++)
            for (long col = c0; col <= c1; col++) {
                const long c = W * row + col;
                if (mb[c]) continue;
                const double px = col * gres + L->boundary[0], py = row * gres + L->boundary[1];
                float z = -1000.0f;
                if (!tin_inside(px, py, p1, p2, p3, &z)) continue;
                smx_ugrid *G = &g[c];
                if (level < 2) {
                    const double d1 = sqrt((px - p1->x) * (px - p1->x) + (py - p1->y) * (py - p1->y));
                    const double d2 = sqrt((px - p2->x) * (px - p2->x) + (py - p2->y) * (py - p2->y));
                    const double d3 = sqrt((px - p3->x) * (px - p3->x) + (py - p3->y) * (py - p3->y));
                    if (d1 == 0) { z = (float)p1->z; G->Matched_flag = 2; }
                    else if (d2 == 0) { z = (float)p2->z; G->Matched_flag = 2; }
                    else if (d3 == 0) { z = (float)p3->z; G->Matched_flag = 2; }
                }
                mb[c] = 1;
                G->Height = z;
                newh[c] = z;
                const double t1 = smin(zmin, (double)z), t2 = smax(zmax, (double)z);
                if (G->Matched_flag == 4) {
                    if (t1 - bf <= G->minHeight) G->minHeight = (int16_t)floor(t1 - bf);
                    if (t2 + bf >= G->maxHeight) G->maxHeight = (int16_t)ceil(t2 + bf);
                } else {
                    G->minHeight = (int16_t)floor(t1 - bf);
                    G->maxHeight = (int16_t)ceil(t2 + bf);
                }
                if (G->minHeight > G->maxHeight) {
                    const int16_t s = G->minHeight;
                    G->minHeight = G->maxHeight;
                    G->maxHeight = s;
                }
                if (G->Matched_flag == 2) {
                    G->minHeight = (int16_t)floor(z - bf);
                    G->maxHeight = (int16_t)ceil(z + bf);
                } else {
                    G->Matched_flag = 1;
                    if (G->Mean_ortho_ncc > q_result(ncc_th)) {
                        G->minHeight = (int16_t)floor(z - bf);
                        G->maxHeight = (int16_t)ceil(z + bf);
                    } else if (level >= 2) {
                        G->minHeight = (int16_t)floor(t1 - bf);
                        G->maxHeight = (int16_t)ceil(t2 + bf);
                    }
                }
                if (G->maxHeight - G->minHeight > th_hg) G->minHeight = G->maxHeight = -100;
            }
    }
    /* setsm_code.cpp:10093-10190: unclaimed cells take the range of their
     * claimed neighbours (a snapshot), or the TIN's height span at levels 0-1 */
    for (long c = 0; c < n; c++) {
        g[c].Height = newh[c];
        snap[c] = g[c];
    }
    double hmin = 100000.0, hmax = -100000.0;
    for (long row = 0; row < H; row++)
        for (long col = 0; col < W; col++) {
            const long c = W * row + col;
            smx_ugrid *G = &g[c];
            if (G->Matched_flag == 0) {
                int lo = 10000, hi = -10000, any = 0;
                for (long dr = -1; dr <= 1; dr++)
                    for (long dc = -1; dc <= 1; dc++) {
                        if (col + dc < 0 || col + dc >= W || row + dr < 0 || row + dr >= H) continue;
                        const smx_ugrid *q = &snap[(col + dc) + (row + dr) * W];
                        if (q->Matched_flag == 0) continue;
                        any = 1;
                        if (lo > q->minHeight) lo = q->minHeight;
                        if (hi < q->maxHeight) hi = q->maxHeight;
                    }
                if (any) {
                    G->minHeight = (int16_t)lo;
                    G->maxHeight = (int16_t)hi;
                    mb[c] = 2;
                } else if (level <= 1) {
                    const double span = fabs(tmax - tmin) / 3.0;
                    double bf = buf > span ? span : buf;
                    if (bf < (1 << level) * L->resolution) bf = (1 << level) * L->resolution;
                    if (G->minHeight < tmin - bf) G->minHeight = (int16_t)floor(tmin - bf);
                    if (G->maxHeight > tmax + bf) G->maxHeight = (int16_t)ceil(tmax + bf);
                }
            } else if (mb[c] == 0)
                G->Matched_flag = 0;
            if (hmin > G->minHeight) hmin = (double)G->minHeight;
            if (hmax < G->maxHeight) hmax = (double)G->maxHeight;
            if (G->maxHeight - G->minHeight > th_hg) G->minHeight = G->maxHeight = -100;
        }
    if (minH_grid) *minH_grid = hmin;
    if (maxH_grid) *maxH_grid = hmax;
    /* setsm_code.cpp:10192-10252: the returned grid */
    for (long c = 0; c < n; c++) {
        const smx_ugrid G = g[c];
        smx_ugrid *R = &g[c];
        int16_t lo = G.minHeight, hi = G.maxHeight;
        if (lo > hi) {
            const int16_t s = lo;
            lo = hi;
            hi = s;
        }
        memset(R, 0, sizeof *R);
        R->Matched_flag = (mb[c] == 2 && G.Matched_flag == 0) ? 1 : G.Matched_flag;
        R->roh = G.roh;
        R->Height = mb[c] == 0 ? -1000.0f : G.Height;
        R->anchor_flag = 0;
        R->ortho_ncc[1] = G.ortho_ncc[1];
        R->Mean_ortho_ncc = G.Mean_ortho_ncc;
        R->minHeight = lo < -100 ? -100 : lo;
        R->maxHeight = hi;
        if (R->minHeight == 0) R->minHeight = -1;
        if (R->maxHeight == 0) R->maxHeight = 1;
    }
    free(mb);
    free(newh);
    free(snap);
    return 0;
}

/* ------------------------------------------------------------------------
 * DecisionMPs' helpers: SetHeightRange_blunder, VerticalLineLocus_blunder
 * ---------------------------------------------------------------------- */

void orc_set_height_range_blunder(const smx_level *L, const smx_point3 *pts, long npts, const smx_tri *tris,
                                  long ntris, smx_ugrid *g)
{
    /* setsm_code.cpp:10393-10491: every cell takes the highest TIN height of
     * the triangles containing it (a float max; untouched cells keep theirs) */
    const long W = (long)L->grid_w, H = (long)L->grid_h, n = W * H;
    const double gres = L->grid_resolution;
    float *hts = (float *)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    if (!hts) return;
    for (long c = 0; c < n; c++) hts[c] = (float)-1000.0;
    for (long t = 0; t < ntris; t++) {
        if (tris[t].a >= (uint32_t)npts || tris[t].b >= (uint32_t)npts || tris[t].c >= (uint32_t)npts) continue;
        const smx_point3 *p1 = &pts[tris[t].a], *p2 = &pts[tris[t].b], *p3 = &pts[tris[t].c];
        const double bx0 = smin(smin(p1->x, p2->x), p3->x), by0 = smin(smin(p1->y, p2->y), p3->y);
        const double bx1 = smax(smax(p1->x, p2->x), p3->x), by1 = smax(smax(p1->y, p2->y), p3->y);
        int c0 = (int)((bx0 - L->boundary[0]) / gres + 0.5) - 1, r0 = (int)((by0 - L->boundary[1]) / gres + 0.5) - 1;
        int c1 = (int)((bx1 - L->boundary[0]) / gres + 0.5) + 1, r1 = (int)((by1 - L->boundary[1]) / gres + 0.5) + 1;
        if (c1 >= (int)W) c1 = (int)W - 1;
        if (r1 >= (int)H) r1 = (int)H - 1;
        if (c0 < 0) c0 = 0;
        if (r0 < 0) r0 = 0;
        for (long row = r0; row <= r1; row++)
            for (long col = c0; col <= c1; col++) {
                float z = -1000.0f;
                if (tin_inside(col * gres + L->boundary[0], row * gres + L->boundary[1], p1, p2, p3, &z) &&
                    z > hts[W * row + col])
                    hts[W * row + col] = z;
            }
    }
    for (long c = 0; c < n; c++)
        if (hts[c] != -1000.0) g[c].Height = hts[c];
    free(hts);
}

int orc_blunder_half_template(const smx_level *L, int blunder_level, int iteration, int bblunder)
{
    /* setsm_code.cpp:7884-7911 */
    int T = (uint8_t)L->template_size;
    const int tT = (int)((5.0 / (L->resolution * (1 << blunder_level))) / 2.0) * 2 + 1;
    if (T < tT) T = (uint8_t)tT;
    int half = (int)(T / 2.0);
    if (bblunder) {
        if (L->level == 4) half = T - iteration * 2;
        else if (L->level == 3) half = T - 2 - iteration * 2;
        if (half < (int)(T / 2.0)) half = (int)(T / 2.0);
    }
    return half;
}

int orc_vll_blunder(const smx_level *L, const smx_view bview[2], int blunder_level, const double *ortho_wgs,
                    int iteration, int bblunder, smx_ugrid *g, float *out)
{
    /* setsm_code.cpp:7882-8013 VerticalLineLocus_blunder, two images */
    const int half = orc_blunder_half_template(L, blunder_level, iteration, bblunder);
    const double im_res = L->resolution * (1 << blunder_level);
    const long sw = (long)((L->boundary[2] - L->boundary[0]) / im_res) + 1;
    const long sh = (long)((L->boundary[3] - L->boundary[1]) / im_res) + 1;
    const long W = (long)L->grid_w, Hg = (long)L->grid_h, n = W * Hg;
    double *cd[2];
    for (int k = 0; k < 2; k++) {
        cd[k] = (double *)calloc((size_t)(sw * sh) * 2, sizeof(double));
        if (!cd[k]) return -1;
    }
    /* SetOrthoImageCoord (setsm_code.cpp:6550-6670) at the blunder level,
     * check_pyimage = BD: positions in the level's pyramid from BStartpos */
    smx_level Lb = *L;
    Lb.view[0] = bview[0];
    Lb.view[1] = bview[1];
    Lb.level = blunder_level;
    Lb.ortho_wgs = ortho_wgs;
    set_ortho_image_coord(&Lb, g, 0, im_res, 0, sw, sh, 0, 0, cd, NULL);
    /* the target image of SETSM's loop (ti = 1) is the reference image when
     * reference_id is 1 */
    const int right = L->reference_id == 1 ? 0 : 1;
#pragma omp parallel num_threads(orc_get_threads())
    {
        patchvec pv;
        pv_init(&pv, half);
#pragma omp for schedule(guided)
        for (long c = 0; c < n; c++) {
            out[c] = -1.0f;
            double max_ncc = -100;
            int Count_N[3] = {0, 0, 0};
            double count_GNCC = 0, t = 0;
            for (int row = -half; row <= half; row++)
                for (int col = -half; col <= half; col++) {
                    const int radius2 = row * row + col * col;
                    if (radius2 > (half - 1) * (half - 1)) continue;
                    const double t_X = L->grid_pts[2 * c] + col * im_res;
                    const double t_Y = L->grid_pts[2 * c + 1] + row * im_res;
                    const long t_col = (long)((t_X - L->boundary[0]) / im_res);
                    const long t_row = (long)((t_Y - L->boundary[1]) / im_res);
                    const long pit = t_row * sw + t_col;
                    const long tt_col = (long)((t_X - L->boundary[0]) / L->grid_resolution);
                    const long tt_row = (long)((t_Y - L->boundary[1]) / L->grid_resolution);
                    const long pdem = tt_row * W + tt_col;
                    if (pit >= 0 && pit < sw * sh && t_col >= 0 && t_col < sw && t_row >= 0 && t_row < sh &&
                        pdem >= 0 && pdem < n && tt_col >= 0 && tt_col < W && tt_row >= 0 && tt_row < Hg &&
                        g[pdem].Height != -1000)
                        set_vec_kernel_value(&pv, half, &bview[0], &bview[right], row, col, cd[0][2 * pit],
                                             cd[0][2 * pit + 1], cd[right][2 * pit], cd[right][2 * pit + 1], radius2,
                                             Count_N);
                }
            compute_multi_ncc(&pv, Count_N, &count_GNCC, &t);
            if (max_ncc < t) max_ncc = t;
            g[c].ortho_ncc[1] = q_result(t);
            g[c].Mean_ortho_ncc = q_result(max_ncc);
            out[c] = (float)max_ncc;
        }
        pv_free(&pv);
    }
    free(cd[0]);
    free(cd[1]);
    return 0;
}

/* ------------------------------------------------------------------------
 * blunder_detection_TIN (setsm_code.cpp:9158-9782), the OpenMP loop run in
 * index order. Every schedule of that loop ends the same: a point's own test
 * only ever sets flag 1 on itself, a neighbour test only turns a flag-0 point
 * into 3, and both count as blunders at the end.
 */

/* setsm_code.cpp:9121-9156 set_blunder: only a point still at flag 0 changes,
 * and only then is it recorded as newly detected. */
static void bd_set_blunder(long index, uint8_t val, smx_point3 *pts, uint8_t *detected)
{
    if (pts[index].flag != 0) return;
    pts[index].flag = val;
    detected[index] = 1;
}

static long bd_cell(const smx_level *L, double x, double y)
{
    const long col = (long)((x - L->boundary[0]) / L->grid_resolution + 0.5);
    const long row = (long)((y - L->boundary[1]) / L->grid_resolution + 0.5);
    return (long)L->grid_w * row + col;
}

int orc_blunder_detection_tin(const smx_level *L, const smx_bd_args *a, const float *ortho_ncc, smx_point3 *pts,
                              long npts, const smx_tri *tris, long ntris, smx_ugrid *g, uint8_t *detected,
                              long blunder_count[2], double *minz, double *maxz, double *height_th_out)
{
    const int step = L->level, iteration = a->iteration, IsRA = L->is_ra;
    const int flag_blunder = a->flag_blunder != 0;
    const unsigned count_bl = (uint16_t)a->count_bl;
    const double gridspace = L->grid_resolution;
    memset(detected, 0, (size_t)(npts > 0 ? npts : 0));
    if (step == 0 && iteration == 3) return 0;
    enum { max_nodes = 30 };
    const uint32_t hdiffcount = (uint32_t)a->hinterval;
    uint32_t *savenode = (uint32_t *)malloc(sizeof(uint32_t) * max_nodes * (size_t)(npts > 0 ? npts : 1));
    uint32_t *nodecount = (uint32_t *)calloc((size_t)(npts > 0 ? npts : 1), sizeof(uint32_t));
    uint32_t *hdiffbin = (uint32_t *)calloc((size_t)hdiffcount + 1, sizeof(uint32_t));
    if (!savenode || !nodecount || !hdiffbin) {
        free(savenode);
        free(nodecount);
        free(hdiffbin);
        return -1;
    }
    /* setsm_code.cpp:9183-9287: each point's first 30 triangles, and the
     * height differences along the edges of the triangles whose vertices all
     * exist. The ortho-NCC mean and deviation gathered there feed a threshold
     * that :9304-9313 overwrites, so they are not computed. */
    double sum_data2 = 0.0, sum_data = 0.0;
    long dh_count = 0;
    for (long t = 0; t < ntris; t++) {
        const uint32_t v[3] = {tris[t].a, tris[t].b, tris[t].c};
        smx_point3 p[3] = {{0}};
        int check_pt_index = 1;
        for (int k = 0; k < 3; k++) {
            if ((long)v[k] < npts) {
                if (nodecount[v[k]] < max_nodes) {
                    savenode[(long)v[k] * max_nodes + nodecount[v[k]]] = (uint32_t)t;
                    nodecount[v[k]]++;
                }
                p[k] = pts[v[k]];
            } else
                check_pt_index = 0;
        }
        if (check_pt_index) {
            double dh1 = (double)fabs(p[0].z - p[1].z);
            sum_data += dh1;
            sum_data2 += dh1 * dh1;
            dh_count++;
            double dh2 = (double)fabs(p[0].z - p[2].z);
            sum_data += dh2;
            sum_data2 += dh2 * dh2;
            dh_count++;
            double dh3 = (double)fabs(p[1].z - p[2].z);
            sum_data += dh3;
            sum_data2 += dh3 * dh3;
            dh_count++;
            dh1 = (uint32_t)dh1;
            if (dh1 < hdiffcount) hdiffbin[(uint32_t)dh1]++;
            dh2 = (uint32_t)dh2;
            if (dh2 < hdiffcount) hdiffbin[(uint32_t)dh2]++;
            dh3 = (uint32_t)dh3;
            if (dh3 < hdiffcount) hdiffbin[(uint32_t)dh3]++;
        }
    }
    /* setsm_code.cpp:9289-9318: the ortho-NCC thresholds of the level */
    double ortho_ncc_th;
    if (step == 4) ortho_ncc_th = 0.6 - (iteration - 1) * 0.01;
    else if (step >= 3) ortho_ncc_th = 0.5 - (iteration - 1) * 0.01;
    else if (step == 2) ortho_ncc_th = 0.4 - (iteration - 1) * 0.01;
    else if (step == 1) ortho_ncc_th = 0.3;
    else ortho_ncc_th = 0.2;
    const double ortho_ancc_th = 100.;
    double th_ref_ncc = 0.1 + (iteration - 1) * 0.05;
    if (th_ref_ncc > ortho_ncc_th) th_ref_ncc = ortho_ncc_th;
    /* setsm_code.cpp:9320-9418: the height-difference threshold */
    int check_dh = 0;
    if (step <= 4 && step >= 3) check_dh = 1;
    else if (step == 2) {
        if (iteration <= 4) check_dh = 1;
    } else if (step <= 1) {
        if (iteration <= 2) check_dh = 1;
    }
    /* GSD is read only where step >= 3; at step 0 the reference's
     * pwrtwo(step - 1) is a negative shift, and the branch below for
     * step < 3 replaces the threshold it would feed. */
    double GSD = gridspace;
    if (step >= 3 && gridspace < ORC_PWRTWO(step - 1)) GSD = ORC_PWRTWO(step - 1);
    const double height_th_1 = a->hinterval;
    double height_th_2 = GSD * 10;
    if (IsRA == 1) height_th_2 = GSD * 30;
    double height_th = height_th_1 < height_th_2 ? height_th_1 : height_th_2;
    if (step >= 3) {
        double up_3sigma;
        if (check_dh) {
            const double mean = sum_data / dh_count;
            const double std = sqrt((sum_data2 - (sum_data) * (sum_data) / dh_count) / dh_count);
            up_3sigma = mean + 3 * std;
        } else
            up_3sigma = height_th;
        if (up_3sigma < height_th) {
            if (up_3sigma < height_th / 2.0) height_th = height_th / 2.0;
            else height_th = up_3sigma;
        }
    } else {
        uint32_t total_dh = 0, th80 = 1000;
        long tcnt = 0;
        int check_total_dh = 0;
        while (tcnt < (long)hdiffcount && !check_total_dh) {
            total_dh += hdiffbin[tcnt];
            const double per = (double)total_dh / (double)dh_count * 100.0;
            if (per > 99) {
                th80 = (uint32_t)tcnt;
                check_total_dh = 1;
            }
            tcnt++;
        }
        if (!check_total_dh) th80 = (uint32_t)(tcnt - 1);
        const double mean = sum_data / dh_count;
        const double std = sqrt((sum_data2 - (sum_data) * (sum_data) / dh_count) / dh_count);
        const double up_3sigma = mean + 3 * std;
        if (up_3sigma < th80) height_th = up_3sigma;
        else height_th = th80;
        if (height_th > 50) height_th = 50;
    }
    free(hdiffbin);
    if (height_th_out) *height_th_out = height_th;

    /* setsm_code.cpp:9441-9756: each point against its triangles */
    for (long index = 0; index < npts; index++) {
        if (pts[index].flag == 1) continue;
        int pt_is_blunder = 0, count_th_positive = 0, count_th_negative = 0, count = 0, check_neigh = 0;
        const int max_iter = nodecount[index] < max_nodes ? (int)nodecount[index] : max_nodes;
        const long ref_index = bd_cell(L, pts[index].x, pts[index].y);
        g[ref_index].anchor_flag = 0;
        if (!IsRA) {
            if (ortho_ncc[ref_index] < th_ref_ncc && step >= 2) pt_is_blunder = 1;
        }
        if (step >= 3 && iteration >= 4) {
            if (ortho_ncc[ref_index] >= 0.95 && flag_blunder) g[ref_index].anchor_flag = 1;
        }
        for (int it = 0; it < max_iter; it++) {
            const uint32_t tri = savenode[index * max_nodes + it];
            if (!((long)tri < ntris && tri > 0)) continue;
            long reference_index = 0, target_index_0 = 0, target_index_1 = 0;
            const uint32_t temp_tri[3] = {tris[tri].a, tris[tri].b, tris[tri].c};
            int check_index = 0;
            for (int kk = 0; kk < 3; kk++) {
                if ((long)temp_tri[kk] == index) reference_index = index;
                else if (!check_index) {
                    target_index_0 = temp_tri[kk];
                    check_index = 1;
                } else
                    target_index_1 = temp_tri[kk];
            }
            if (!(reference_index < npts && target_index_0 < npts && target_index_1 < npts)) continue;
            const smx_point3 pt0 = pts[reference_index], pt1 = pts[target_index_0], pt2 = pts[target_index_1];
            const double angle = tin_slope_deg(&pt0, &pt1, &pt2);
            const double dh1 = pt0.z - pt1.z, dh2 = pt0.z - pt2.z, dh3 = pt1.z - pt2.z;
            if (dh1 >= 0 && dh2 >= 0 && angle > 30) count_th_positive++;
            if (dh1 < 0 && dh2 < 0 && angle > 30) count_th_negative++;
            int check_match = 0;
            if (step > 1) check_match = 1;
            else if (step == 1) {
                if (iteration == 1 && count_bl <= 5) check_match = 1;
            } else if (step == 0) {
                if (iteration == 1 && count_bl <= 5) check_match = 1;
            }
            if (check_match) {
                const double ddh_1 = fabs(dh1) - fabs(dh2), ddh_2 = fabs(dh1) - fabs(dh3),
                             ddh_3 = fabs(dh2) - fabs(dh3);
                if (fabs(ddh_1) > height_th || fabs(ddh_2) > height_th || fabs(ddh_3) > height_th) {
                    const long order[3] = {reference_index, target_index_0, target_index_1};
                    const double height[3] = {pt0.z, pt1.z, pt2.z};
                    const long Index[3] = {bd_cell(L, pt0.x, pt0.y), bd_cell(L, pt1.x, pt1.y),
                                           bd_cell(L, pt2.x, pt2.y)};
                    int t_o_min, t_o_max, t_o_mid;
                    check_neigh = 1;
                    if (height[0] > height[1]) {
                        if (height[0] > height[2]) {
                            t_o_max = 0;
                            if (height[1] > height[2]) { t_o_min = 2; t_o_mid = 1; }
                            else { t_o_min = 1; t_o_mid = 2; }
                        } else { t_o_max = 2; t_o_min = 1; t_o_mid = 0; }
                    } else {
                        if (height[1] > height[2]) {
                            t_o_max = 1;
                            if (height[0] > height[2]) { t_o_min = 2; t_o_mid = 0; }
                            else { t_o_min = 0; t_o_mid = 2; }
                        } else { t_o_max = 2; t_o_mid = 1; t_o_min = 0; }
                    }
                    const double h1 = height[t_o_mid] - height[t_o_min], h2 = height[t_o_max] - height[t_o_mid];
                    const double dh = h1 - h2;
                    long nb = -1;
                    if (dh > 0 && dh > height_th) {
                        if (IsRA == 1) nb = order[t_o_min];
                        else if (flag_blunder) {
                            if (ortho_ncc[Index[t_o_min]] < ortho_ncc_th) nb = order[t_o_min];
                        } else if (ortho_ncc[Index[t_o_min]] < ortho_ancc_th)
                            nb = order[t_o_min];
                    } else if (dh < 0 && fabs(dh) > height_th) {
                        if (IsRA == 1) nb = order[t_o_max];
                        else if (flag_blunder) {
                            if (ortho_ncc[Index[t_o_max]] < ortho_ncc_th) nb = order[t_o_max];
                        } else if (ortho_ncc[Index[t_o_max]] < ortho_ancc_th)
                            nb = order[t_o_max];
                    }
                    if (nb >= 0) bd_set_blunder(nb, 3, pts, detected);
                }
            }
            count++;
        }
        if (IsRA == 1) {
            if (check_neigh && !flag_blunder) pt_is_blunder = 1;
            if ((count_th_positive >= (int)(count * 0.7 + 0.5) || count_th_negative >= (int)(count * 0.7 + 0.5)) &&
                count > 1)
                pt_is_blunder = 1;
        } else {
            if (check_neigh && !flag_blunder) {
                if (ortho_ncc[ref_index] < ortho_ancc_th) pt_is_blunder = 1;
            }
            double count_threshold = count * 0.5;
            if (step >= 2) count_threshold = count * 0.5;
            else if (step == 1) count_threshold = (int)(count * 0.6 + 0.5);
            else count_threshold = (int)(count * 0.7 + 0.5);
            if ((count_th_positive >= count_threshold || count_th_negative >= count_threshold) && count > 1) {
                if (flag_blunder) {
                    double tmp_th = 0.6 - (4 - step) * 0.1;
                    if (step == 1) tmp_th = 0.6;
                    if (step >= 1) {
                        if (ortho_ncc[ref_index] < tmp_th) pt_is_blunder = 1;
                    } else if (step == 0) {
                        if (iteration <= 1) {
                            if (ortho_ncc[ref_index] < 0.9) pt_is_blunder = 1;
                        } else if (iteration == 2)
                            pt_is_blunder = 1;
                        else if (ortho_ncc[ref_index] < ortho_ncc_th)
                            pt_is_blunder = 1;
                    }
                } else if (ortho_ncc[ref_index] < ortho_ancc_th)
                    pt_is_blunder = 1;
            }
        }
        if (pt_is_blunder) bd_set_blunder(index, 1, pts, detected);
    }
    /* setsm_code.cpp:9758-9776: counts, flag 3 -> 1, the kept points' heights */
    for (long t = 0; t < npts; t++) {
        if (pts[t].flag == 1) blunder_count[1]++;
        else if (pts[t].flag == 3) {
            blunder_count[1]++;
            pts[t].flag = 1;
        } else {
            blunder_count[0]++;
            if (*minz > pts[t].z) *minz = pts[t].z;
            if (*maxz < pts[t].z) *maxz = pts[t].z;
        }
    }
    free(savenode);
    free(nodecount);
    return 0;
}
