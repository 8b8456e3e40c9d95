/*
 * smx_oracle.h — CPU FP64 restatement of SETSM's VerticalLineLocus hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the reference the MI355X
 * kernels are compared against and the CPU baseline that bench.py times. It is
 * never linked into, loaded by or called from the product library; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every function names the reference file:line it restates.
 */
#ifndef SMX_ORACLE_H
#define SMX_ORACLE_H

#include "smx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-cell voxel rows, the oracle's model of SETSM's VOXEL **grid_voxel
 * (setsm_code.cpp:3745, rows calloc'd in InitializeVoxel). */
typedef struct orc_voxels {
    long      ncells;
    uint32_t **rows;     /* rows[c] has nccresult[c].NumOfHeight words (SMX_VOXEL_WORD) */
    uint16_t  *len;      /* allocated length of rows[c] */
    double   **rho;      /* optional per-hypothesis mean NCC before quantisation        */
} orc_voxels;

orc_voxels *orc_voxels_create(long ncells, int keep_rho);
void        orc_voxels_destroy(orc_voxels *v);
/* Pack rows into CSR order (offsets = prefix sum of NumOfHeight). Returns the
 * number of words written. rho_out may be NULL. */
long        orc_voxels_export(const orc_voxels *v, const smx_nccresult *ncc,
                              uint32_t *words_out, double *rho_out);
/* Replace all rows from CSR (used to seed a state). */
void        orc_voxels_import(orc_voxels *v, const smx_nccresult *ncc, const uint32_t *words);

/* CoordConversion.hpp:26-87 GetObjectToImageRPC_single */
void   orc_rpc_ground_to_image(const smx_rpc *rpc, const double adjust[2],
                               double lon, double lat, double h,
                               double *img_x, double *img_y);
/* setsm_code.cpp:5741-5757 GetHeightStep */
double orc_height_step(int level, double resolution);
/* setsm_code.cpp:6057-6077 — the Half_template_size VerticalLineLocus uses */
int    orc_vll_half_template(int level, int template_size, double resolution);
/* SubFunctions.cpp:3274-3312 — nested-window membership bits of an offset */
int    orc_window_bits(int half, int row, int col);
/* SubFunctions.cpp:974-1018 Correlate */
double orc_correlate(const double *L, const double *R, int N);
/* setsm_code.cpp:11002-11083 check_image_boundary (SB sensor, 2 images) */
int    orc_check_image_boundary(const smx_level *L, long cell, double minH, double maxH, int half);
/* setsm_code.cpp:6014-6047 SetNCC_alpha */
double orc_ncc_alpha(int level, int iteration, int is_ra);

/* setsm_code.cpp:6672-6752 FindPeakNcc, one evaluated height per call in
 * height order; st carries check_rho, pre_rho, pre_height, direction and
 * max_WNCC between calls (orc_peak_state_init gives the reference's initial
 * values, setsm_code.cpp:6343-6347). */
typedef struct orc_peak_state {
    int    check_rho;
    double pre_rho;
    float  pre_height;
    int    direction;
    double max_wncc;
} orc_peak_state;
void orc_peak_state_init(orc_peak_state *st);
void orc_find_peak_ncc(double temp_rho, float iter_height, orc_peak_state *st, smx_nccresult *nr);

/* setsm_code.cpp:5759-6012 InitializeVoxel */
void orc_initialize_voxel(const smx_level *L, orc_voxels *vox, smx_ugrid *gridpt3,
                          smx_nccresult *ncc, int iteration, const double minmax[2]);

/* setsm_code.cpp:6049-6548 VerticalLineLocus (INCC term). Returns
 * Accessable_grid. OpenMP dynamic,1 over cells like the reference. */
long orc_vertical_line_locus(const smx_level *L, orc_voxels *vox, smx_nccresult *ncc,
                             const smx_ugrid *gridpt3, int iteration, const double minmax[2]);

/* setsm_code.cpp:7511-7702 AWNCC peak scan without SGM aggregation
 * (check_SGM false, i.e. Pyramid_step > SGM_py). */
void orc_awncc_peak(const smx_level *L, const orc_voxels *vox, const smx_ugrid *gridpt3,
                    smx_nccresult *ncc, int iteration);
/* setsm_code.cpp:7017-7509 SGM 8-path aggregation (SGM_start_pos :6755-6767,
 * SGM_con_pos :6769-6999, P1 = 0.3, P2 = 0.6, P_HS_step = 1): SumCost per
 * hypothesis in CSR order (prefix sum of NumOfHeight), zeroed first. */
void orc_sgm_aggregate(const smx_level *L, const orc_voxels *vox, const smx_ugrid *gridpt3,
                       const smx_nccresult *ncc, float *sumcost_csr);
/* setsm_code.cpp:7511-7702 AWNCC as the reference runs it: SGM aggregation
 * and the SGM peak mode when Pyramid_step <= SGM_py, the plain scan
 * otherwise. sumcost_csr (may be NULL) receives SumCost in SGM mode. */
void orc_awncc(const smx_level *L, const orc_voxels *vox, const smx_ugrid *gridpt3,
               smx_nccresult *ncc, int iteration, float *sumcost_csr);

/* setsm_code.cpp:8535-8889 SelectMPs: updates GridPT3 (roh, ranges, flags)
 * and writes the matched points in cell order to pts (capacity: the cell
 * count). Returns their count, or -1 where the reference leaves the level's
 * threshold undefined. *minimum_th receives the histogram threshold. */
long orc_select_mps(const smx_level *L, const smx_nccresult *ncc, smx_ugrid *gridpt3, const smx_mps_args *a,
                    smx_point3 *pts, double *minimum_th);
/* setsm_code.cpp:9851-10261 SetHeightRange (SetTinBoundary, IsTinInside,
 * SetNormalAngle: SubFunctions.cpp:2915-3024): gridpt3 is replaced by the
 * returned grid. Returns 0, or -1 if host memory ran out. */
int orc_set_height_range(const smx_level *L, const smx_hr_args *a, const smx_point3 *pts, long npts,
                         const smx_tri *tris, long ntris, smx_ugrid *gridpt3, double *minH_grid,
                         double *maxH_grid);

/* setsm_code.cpp:10393-10491 SetHeightRange_blunder: each cell's Height
 * becomes the highest TIN height of the triangles containing it. */
void orc_set_height_range_blunder(const smx_level *L, const smx_point3 *pts, long npts, const smx_tri *tris,
                                  long ntris, smx_ugrid *gridpt3);
/* setsm_code.cpp:7884-7911: VerticalLineLocus_blunder's Half_template_size */
int  orc_blunder_half_template(const smx_level *L, int blunder_level, int iteration, int bblunder);
/* setsm_code.cpp:7882-8013 VerticalLineLocus_blunder: the ortho NCC of every
 * cell at its GridPT3 Height on the blunder level's images bview[0] / [1]
 * (reference / target, starts at that level), the ortho grid at
 * resolution * 2^blunder_level with its pixels' lon/lat in ortho_wgs. Writes
 * GridPT3.ortho_ncc[1], Mean_ortho_ncc and out[cell]. Returns 0 or -1. */
int  orc_vll_blunder(const smx_level *L, const smx_view bview[2], int blunder_level, const double *ortho_wgs,
                     int iteration, int bblunder, smx_ugrid *gridpt3, float *out);

/* setsm_code.cpp:9158-9782 blunder_detection_TIN: every point that is not
 * an old blunder (flag 1) is tested against its first 30 triangles; new
 * blunders end at flag 1 and are marked in detected (npts bytes, cleared
 * first), GridPT3.anchor_flag is reset or set at each tested point's cell,
 * blunder_count[0] / [1] and *minz / *maxz accumulate as in the reference.
 * *height_th (may be NULL) receives the height-difference threshold. Does
 * nothing at level 0 iteration 3. Returns 0, or -1 if host memory ran out. */
int orc_blunder_detection_tin(const smx_level *L, const smx_bd_args *a, const float *ortho_ncc, smx_point3 *pts,
                              long npts, const smx_tri *tris, long ntris, smx_ugrid *gridpt3, uint8_t *detected,
                              long blunder_count[2], double *minz, double *maxz, double *height_th);

/* Pyramid inputs. Template.hpp:45-132 (9x9, sigma 1.5 decimation),
 * SubFunctions.cpp:2418-2470 (Sobel), SubFunctions.cpp:2473-2571 (Orientation). */
void orc_gauss_pyramid(const uint16_t *in, uint32_t w, uint32_t h, uint16_t *out);
void orc_sobel(const uint16_t *in, uint32_t w, uint32_t h, uint16_t *mag, int16_t *dir);
void orc_orientation(const uint16_t *mag, const int16_t *dir, uint32_t w, uint32_t h,
                     int template_size, uint8_t *ori);

/* Threads used by the OpenMP loops (0 = OpenMP default). */
void orc_set_threads(int n);
int  orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
