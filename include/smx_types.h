/*
 * smx_types.h — plain-old-data records that cross the host <-> MI355X boundary
 * of the SETSM VerticalLineLocus engine.
 *
 * Two groups of records live here:
 *
 *  1. Byte-for-byte images of the SETSM records the matching step reads and
 *     writes, so that a caller holding SETSM arrays can hand them over without
 *     conversion (GCC x86-64 ABI; sizes/offsets are asserted below):
 *       smx_ugrid      <-> UGRID     (Typedefine.hpp:271-284, 20 bytes)
 *       smx_nccresult  <-> NCCresult (Typedefine.hpp:248-269, 24 bytes)
 *       smx_voxel      <-> VOXEL     (Typedefine.hpp:286-293,  4 bytes)
 *
 *  2. Flattened views of what SETSM passes through LevelInfo/ProInfo
 *     (Typedefine.hpp:351-419, 604-643): the per-level pyramid arrays of the
 *     stereo pair, the two RPC models and the scalar parameters that the
 *     VerticalLineLocus / InitializeVoxel / AWNCC trio consults.
 *
 * Everything is C so the records can be shared by the C++/HIP library, the C
 * test oracle and Python ctypes bindings.
 */
#ifndef SMX_TYPES_H
#define SMX_TYPES_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Rational polynomial camera model, RPC00B term order.
 * SETSM keeps it as double*[7] rows (SubFunctions.cpp:1420-1582):
 *   row 0 = offsets {line, samp, lon, lat, height}
 *   row 1 = scales  {line, samp, lon, lat, height}
 *   rows 2..5 = 20 coefficients of line-num, line-den, samp-num, samp-den.
 * Here the same numbers are stored contiguously (90 doubles). */
typedef struct smx_rpc {
    double off[5];
    double scale[5];
    double coef[4][20];
} smx_rpc;

/* One image of the stereo pair at one pyramid level (LevelInfo.py_Images /
 * py_MagImages / py_OriImages / py_Sizes / py_Startpos). Row-major, width
 * pixels per row. */
typedef struct smx_view {
    const uint16_t *img;   /* Gaussian pyramid intensity (11-bit DN, 0 = nodata) */
    const uint16_t *mag;   /* Sobel magnitude, uint16(sqrt(gx^2+gy^2)+0.5)        */
    const uint8_t  *ori;   /* dominant orientation bin 0..17                      */
    uint32_t width;
    uint32_t height;
    double   start_x;      /* pyramid-level start position of the subset          */
    double   start_y;
} smx_view;

/* Everything VerticalLineLocus reads for one (tile, level) besides the
 * per-cell search space. Image 0 is the reference image (reference_id = 0),
 * image 1 the target (ti = 1). `next` holds level-1 arrays for the combined
 * level INCC; leave next[*].img NULL when the level has no combined term. */
typedef struct smx_level {
    smx_view view[2];
    smx_view next[2];
    smx_rpc  rpc[2];
    double   adjust[2][2];        /* ImageAdjust[ti] = {deltaP (line), deltaR (samp)} */
    int32_t  level;               /* Pyramid_step                                    */
    int32_t  template_size;       /* Template_size (SETSM uses 15)                   */
    int32_t  py_combined_level;   /* Py_combined_level (SETSM uses 0)                */
    int32_t  sgm_py;              /* ProInfo.SGM_py (-1 disables SGM)                */
    int32_t  is_ra;               /* ProInfo.IsRA                                    */
    int32_t  check_matchtag;      /* ProInfo.check_Matchtag                          */
    int32_t  pre_dem;             /* ProInfo.pre_DEMtif                              */
    int32_t  check_matching_rate; /* LevelInfo.check_matching_rate                   */
    double   resolution;          /* ProInfo.resolution (image GSD, m)               */
    double   dem_resolution;      /* ProInfo.DEM_resolution                          */
    double   bin_angle;           /* 360/18                                          */
    double   height_step;         /* GetHeightStep(level, resolution)                */
    double   grid_resolution;     /* LevelInfo.grid_resolution                       */
    double   boundary[4];         /* LevelInfo.Boundary                              */
    uint32_t grid_w;              /* Size_Grid2D                                     */
    uint32_t grid_h;
    const double *grid_pts;       /* GridPts: (x, y) metres, 2 doubles per cell      */
    const double *grid_wgs;       /* Grid_wgs: (lon, lat) degrees, 2 doubles per cell*/
    int32_t  reference_id;        /* LevelInfo.reference_id: SETSM's index of image 0;
                                     the GNCC gate reads GridPT3.ortho_ncc[1 - reference_id] */
    uint32_t ortho_w;             /* ortho grid of SetOrthoImageCoord at im_resolution =
                                     resolution * 2^level: (Boundary[2]-Boundary[0])/im_res + 1 */
    uint32_t ortho_h;             /*   columns, (Boundary[3]-Boundary[1])/im_res + 1 rows   */
    const double *ortho_wgs;      /* (lon, lat) degrees of every ortho pixel (Boundary[0] +
                                     col*im_res, Boundary[1] + row*im_res), row-major, as
                                     ps2wgs_single gives them; NULL if GNCC is never active */
} smx_level;

/* UGRID (Typedefine.hpp:271-284) */
typedef struct smx_ugrid {
    float    Height;
    int16_t  minHeight;
    int16_t  maxHeight;
    int16_t  roh;
    int16_t  ortho_ncc[2];
    int16_t  Mean_ortho_ncc;
    uint8_t  Matched_flag;
    uint8_t  anchor_flag;
} smx_ugrid;

/* NCCresult (Typedefine.hpp:248-269) */
typedef struct smx_nccresult {
    float    result2;             /* first peak height  */
    float    result3;             /* second peak height */
    int16_t  minHeight;
    int16_t  maxHeight;
    int16_t  result0;             /* first peak rho x1000  */
    int16_t  result1;             /* second peak rho x1000 */
    int16_t  max_WNCC;
    int16_t  GNCC;
    uint16_t NumOfHeight;
    uint8_t  result4;             /* peak count */
    uint8_t  check_height_change; /* C++ bool   */
} smx_nccresult;

/* VOXEL (Typedefine.hpp:286-293): bool flag_cal at byte 0, short INCC at
 * byte 2. Read as a little-endian uint32: bits 0..7 flag, bits 16..31 INCC. */
typedef struct smx_voxel {
    uint8_t flag_cal;
    uint8_t pad_;
    int16_t INCC;
} smx_voxel;

#define SMX_VOXEL_WORD(flag, incc) \
    ((uint32_t)((flag) ? 1u : 0u) | ((uint32_t)(uint16_t)(int16_t)(incc) << 16))
#define SMX_VOXEL_FLAG(w) ((int)((w) & 0xffu))
#define SMX_VOXEL_INCC(w) ((int16_t)(uint16_t)((w) >> 16))

/* D3DPOINT (Typedefine.hpp:140-194): a matched point (ground x, y in metres,
 * height z) with its blunder flag; 32 bytes like the reference's struct. */
typedef struct smx_point3 {
    double  x;
    double  y;
    double  z;
    uint8_t flag;
    uint8_t pad_[7];
} smx_point3;

/* UI3DPOINT (Typedefine.hpp:73-97): a TIN triangle as three indices into
 * the matched-point list. */
typedef struct smx_tri {
    uint32_t a;
    uint32_t b;
    uint32_t c;
} smx_tri;

/* SelectMPs' per-call arguments (setsm_code.cpp:8535); the level, SGM_py and
 * IsRA come from the job's smx_level. */
typedef struct smx_mps_args {
    double  th_roh;                /* Th_roh                                   */
    double  th_roh_min;            /* Th_roh_min                               */
    double  th_roh_start;          /* Th_roh_start                             */
    double  th_roh_next;           /* Th_roh_next                              */
    double  mpp;                   /* MPP                                      */
    double  mpp_stereo_angle;      /* MPP_stereo_angle                         */
    int32_t iteration;
    int32_t final_level_iteration;
} smx_mps_args;

/* SetHeightRange's per-call arguments (setsm_code.cpp:9851); the level,
 * grid, DEM_resolution, resolution and pre_DEMtif come from smx_level. */
typedef struct smx_hr_args {
    double  mpp;                   /* MPP                                      */
    double  seed_dem_sigma;        /* ProInfo.seedDEMsigma (read if pre_dem)   */
    int32_t iteration;
    int32_t check_matching_rate;   /* level_check_matching_rate                */
} smx_hr_args;

/* blunder_detection_TIN's per-call arguments (setsm_code.cpp:9158); the
 * level, IsRA, grid and Boundary come from smx_level. */
typedef struct smx_bd_args {
    double  hinterval;             /* *LevelInfo.Hinterval                     */
    int32_t iteration;
    int32_t flag_blunder;          /* flag_blunder: 1 blunder pass, 0 anchors  */
    int32_t count_bl;              /* DecisionMPs' loop count (a uint16)       */
    int32_t pad_;
} smx_bd_args;

#if defined(__cplusplus)
static_assert(sizeof(smx_point3) == 32 && offsetof(smx_point3, flag) == 24, "D3DPOINT layout");
static_assert(sizeof(smx_tri) == 12, "UI3DPOINT layout");
static_assert(sizeof(smx_rpc) == 90 * sizeof(double), "smx_rpc layout");
static_assert(sizeof(smx_ugrid) == 20, "UGRID is 20 bytes");
static_assert(offsetof(smx_ugrid, Matched_flag) == 16, "UGRID.Matched_flag");
static_assert(sizeof(smx_nccresult) == 24, "NCCresult is 24 bytes");
static_assert(offsetof(smx_nccresult, NumOfHeight) == 20, "NCCresult.NumOfHeight");
static_assert(offsetof(smx_nccresult, check_height_change) == 23, "NCCresult.check_height_change");
static_assert(sizeof(smx_voxel) == 4 && offsetof(smx_voxel, INCC) == 2, "VOXEL layout");
#else
_Static_assert(sizeof(smx_rpc) == 90 * sizeof(double), "smx_rpc layout");
_Static_assert(sizeof(smx_ugrid) == 20, "UGRID is 20 bytes");
_Static_assert(sizeof(smx_nccresult) == 24, "NCCresult is 24 bytes");
_Static_assert(sizeof(smx_voxel) == 4, "VOXEL is 4 bytes");
#endif

#ifdef __cplusplus
}
#endif

#endif /* SMX_TYPES_H */
