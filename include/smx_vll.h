/*
 * smx_vll.h — C ABI of the MI355X VerticalLineLocus engine (libsmx.so).
 *
 * The reference has no plugin layer: its hot path is three C++ free functions
 * declared in setsm_code.hpp and called once per (tile, level, iteration) from
 * Matching_SETSM (setsm_code.cpp:3827, 3829, 3844):
 *
 *   void InitializeVoxel(const ProInfo*, VOXEL**, LevelInfo&, UGRID*, NCCresult*,
 *                        const int iteration, const double *minmaxHeight);      setsm_code.hpp:75
 *   int  VerticalLineLocus(VOXEL**, const ProInfo*, NCCresult*, LevelInfo&,
 *                          const UGRID*, const uint8 iteration,
 *                          const double *minmaxHeight);                         setsm_code.hpp:81
 *   void AWNCC(ProInfo*, VOXEL**, CSize Size_Grid2D, UGRID*, NCCresult*,
 *              double step_height, uint8 Pyramid_step, uint8 iteration,
 *              int MaxNumberofHeightVoxel);                                     setsm_code.hpp:91
 *
 * This ABI keeps their meaning but moves the state they share — the pyramid,
 * the per-cell NCCresult array and the VOXEL cube — into HBM, owned by a job
 * (one tile). The cube is stored in CSR order (cell rows back to back, offsets
 * = prefix sum of NCCresult.NumOfHeight) with one 32-bit word per hypothesis
 * laid out like VOXEL (flag_cal in bits 0..7, INCC in bits 16..31).
 *
 * All functions return 0 (or a non-negative count) on success and a negative
 * SMX_E* code on failure; smx_last_error() describes the last failure of the
 * calling thread. Calls on one job are not thread safe; different jobs may be
 * driven from different threads.
 */
#ifndef SMX_VLL_H
#define SMX_VLL_H

#include "smx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SMX_OK            0
#define SMX_E_ARG        -1   /* bad argument or unsupported configuration   */
#define SMX_E_HIP        -2   /* HIP runtime error                          */
#define SMX_E_NOMEM      -3   /* device allocation failed / budget exceeded */
#define SMX_E_UNSUPPORTED -4  /* input needs a term this library lacks       */
#define SMX_E_STATE      -5   /* call out of order (e.g. no level set)       */

/* Arithmetic of the VerticalLineLocus kernel. */
#define SMX_PREC_FP64     0   /* plain FP64, same operation order as SETSM:
                                 INCC cube bit-identical to the CPU path      */
#define SMX_PREC_FAST     1   /* 11-bit texels staged in LDS as FP32
                                 (intensity, magnitude) about a pivot (the
                                 right image with each texel's difference to
                                 its row neighbour), FP64 RPC, FP32 bilinear
                                 taps and moments, each
                                 window row's FP32 partials summed into FP64
                                 window totals, FP64 Pearson finish;
                                 |rho - rho_fp64| <= 1e-5                     */

typedef struct smx_job smx_job;
typedef struct smx_graph smx_graph;

const char *smx_last_error(void);
int  smx_version(void);
/* Select the HIP device used by jobs created afterwards on this thread. */
int  smx_set_device(int device);

/* ------------------------------------------------------------------ jobs */
/* One tile. Holds its pyramid, RPCs, per-level grid, search space,
 * NCCresult array and voxel cube in HBM until destroyed. */
int  smx_job_create(smx_job **out, int num_levels);
int  smx_job_destroy(smx_job *job);
/* Bytes of HBM the job currently holds. */
long long smx_job_device_bytes(const smx_job *job);
/* Cap the HBM the job may allocate (0 = no cap); allocations beyond it fail
 * with SMX_E_NOMEM. */
int  smx_job_set_budget(smx_job *job, long long bytes);
/* Device blocks freed by jobs are kept (up to 64 GiB per process) for later
 * jobs and levels; this returns them to the HIP runtime and reports their
 * bytes. No counterpart in the reference (library memory management). */
long long smx_release_cached(void);
/* Frees the job's per-call working lists (run and tile lists, the FP64
 * hand-over queue, level-pass partials, batch descriptors), which the next
 * VerticalLineLocus call reserves again at the size it needs: after a launch
 * shape with many small tiles or short runs, they can hold far more than the
 * job's usual shape. The cube, search space and views are kept. No
 * counterpart in the reference (library memory management). */
int  smx_job_trim(smx_job *job);

/* Upload one pyramid level of image `img` (0 = reference, 1 = target):
 * intensity, Sobel magnitude and orientation, as SetPyramidImages
 * (SubFunctions.cpp:3383-3406) leaves them. start_x/start_y are py_Startpos. */
int  smx_job_set_view(smx_job *job, int img, int level,
                      const uint16_t *intensity, const uint16_t *magnitude,
                      const uint8_t *orientation, uint32_t width, uint32_t height,
                      double start_x, double start_y);

/* Build levels 0..num_levels-1 of image `img` on the GPU from its level-0
 * intensity (9x9 sigma-1.5 Gaussian decimation, Sobel magnitude, 18-bin
 * orientation: Template.hpp:82-132, SubFunctions.cpp:2418-2571).
 * start_x0/start_y0 are the level-0 subset start (Startpos_ori). level0 may
 * be host or device memory. */
int  smx_job_build_pyramid(smx_job *job, int img, const uint16_t *level0,
                           uint32_t width, uint32_t height, double start_x0, double start_y0);

/* Download a pyramid level (any pointer may be NULL). */
int  smx_job_get_view(smx_job *job, int img, int level, uint16_t *intensity,
                      uint16_t *magnitude, uint8_t *orientation);

/* Make `params->level` current: copies the RPCs, scalars and the cell grid
 * (grid_pts / grid_wgs, 2 doubles per cell). The view pointers inside params
 * are ignored: the job's uploaded or GPU-built pyramid is used. Resets the
 * NCCresult array and voxel cube (SETSM allocates both per level,
 * setsm_code.cpp:3745, 3769). */
int  smx_job_set_level(smx_job *job, const smx_level *params);

/* Upload (or replace) the ortho grid lon/lat of the current level without
 * resetting it: w x h (lon, lat) pairs as smx_level.ortho_wgs describes
 * them. Needed before smx_vertical_line_locus whenever the GNCC term is
 * active (check_ortho and some GridPT3.ortho_ncc above ortho_th,
 * setsm_code.cpp:6100-6129, 6240). */
int  smx_job_set_ortho_wgs(smx_job *job, const double *lonlat, uint32_t w, uint32_t h);

/* Upload the search space GridPT3 for the current level (one smx_ugrid per
 * cell) and the level's minmaxHeight. */
int  smx_job_set_searchspace(smx_job *job, const smx_ugrid *gridpt3, const double minmax[2]);
/* Read back GridPT3 (InitializeVoxel may rewrite a band to -100). */
int  smx_job_get_searchspace(smx_job *job, smx_ugrid *gridpt3);
/* Read / overwrite the NCCresult array (one per cell). */
int  smx_job_get_nccresult(smx_job *job, smx_nccresult *out);
int  smx_job_set_nccresult(smx_job *job, const smx_nccresult *in);

/* ------------------------------------------------------- reference trio */
/* InitializeVoxel (setsm_code.cpp:5759-6012) on the resident state: decides
 * per cell whether its height list is (re)built, sets NumOfHeight and
 * re-lays the voxel cube (rows of unchanged cells are kept). */
int  smx_initialize_voxel(smx_job *job, int iteration);

/* VerticalLineLocus (setsm_code.cpp:6049-6548) on the resident state.
 * precision = SMX_PREC_FP64 or SMX_PREC_FAST. Returns Accessable_grid (>= 0)
 * or an error. keep_rho != 0 also stores each evaluated hypothesis' mean NCC
 * before quantisation (float) for precision studies. */
long long smx_vertical_line_locus(smx_job *job, int iteration, int precision, int keep_rho);

/* VerticalLineLocus over several jobs (tiles of different sizes, height
 * ranges, pyramids and levels) at one window size. SMX_PREC_FAST runs ONE
 * persistent launch that takes the tiles of all jobs from a work queue (jobs
 * at a combined-level iteration, setsm_code.cpp:6087-6088, in two more: the
 * level pass and the level-below pass over all such jobs); SMX_PREC_FP64
 * launches the plain kernel once per job. accessible[i] receives job i's
 * Accessable_grid. Every argument is checked before anything is queued. */
int  smx_vertical_line_locus_batch(smx_job **jobs, int n, int iteration, int precision, int keep_rho,
                                   long long *accessible);

/* smx_vertical_line_locus without any device allocation: the call's scratch
 * (height-run list, per-tile run lists, FP64 hand-over queue, level term of a
 * combined iteration, scan partials, counters) comes from the library's
 * static 64 MiB device array instead of the job. No counterpart in the
 * reference: for callers that cannot allocate (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The job must have
 * been set up by the allocating calls (set_level, set_searchspace,
 * initialize_voxel). Levels needing more scratch than the array holds (about
 * 4 M height runs in the precision form, ~260 M hypotheses at runs of 64)
 * fail with SMX_E_NOMEM; RA and
 * matching-rate passes with SMX_E_UNSUPPORTED. The mean NCC is not kept.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU race. */
long long smx_vertical_line_locus_noalloc(smx_job *job, int iteration, int precision);

/* Half_template_size of the job's current level (setsm_code.cpp:6057-6077):
 * the window size a batch groups jobs by. */
int  smx_job_half(const smx_job *job);

/* ---------------------------------------------------- captured sequences */
/* Capture one iteration's VerticalLineLocus (+ AWNCC when with_awncc) over
 * jobs[0..n) as a hipGraph: per job the GNCC geometry, gate, run-list scan and
 * fill; the NCC kernel (SMX_PREC_FAST with n > 1: the batched persistent
 * launch); the FP64 hand-over pass; max_WNCC; then SGM and the peak scan.
 * Nothing in it reads back to the host or allocates; buffers are sized from
 * the height lists InitializeVoxel left. A graph stays valid until a job's
 * level, search space, cube layout, buffers or launch shape change
 * (smx_graph_launch then fails with SMX_E_STATE). Not for RA or
 * matching-rate passes. No counterpart in the reference (its calls run
 * eagerly on the CPU). */
int  smx_graph_create(smx_graph **out, smx_job **jobs, int n, int iteration, int precision, int with_awncc);
/* Replay; waits for it and fills accessible[i] (may be NULL) and each job's
 * last_evaluated / redo statistics. */
int  smx_graph_launch(smx_graph *graph, long long *accessible);
/* Milliseconds of the last replay (HIP events around it on the stream). */
double smx_graph_last_ms(const smx_graph *graph);
int  smx_graph_destroy(smx_graph *graph);

/* AWNCC (setsm_code.cpp:7511-7702, setsm_code.hpp:91): at levels
 * <= SGM_py the SGM 8-path aggregation first (setsm_code.cpp:7017-7509,
 * P1 = 0.3, P2 = 0.6) and the peak scan over SumCost; above SGM_py the peak
 * scan over the INCC (or INCC/GNCC blend). Writes NCCresult result0..4. */
int  smx_awncc(smx_job *job, int iteration);
/* SumCost of the last SGM-level smx_awncc, float per hypothesis in CSR
 * order (SETSM keeps it in AWNCC's local float **SumCost). */
int  smx_job_get_sumcost(smx_job *job, float *sumcost);

/* ------------------------------------------------- between iterations */
/* SelectMPs (setsm_code.cpp:8535-8889, setsm_code.hpp:99): after AWNCC, the
 * level's peak threshold from the first-peak histogram, the matched points
 * and the search-space grid's update (roh, widened ranges of ambiguous
 * matched cells, Matched_flag), on the job's NCCresult and GridPT3. The
 * points go to pts in cell order, as the reference's vector receives them;
 * *count gets their number even when it exceeds cap (then SMX_E_ARG after
 * the grid update; a cap of the cell count always suffices). Needs the
 * level's grid_pts. *minimum_th (may be NULL) gets the threshold. SMX_E_ARG
 * where the reference leaves the threshold undefined (level > 4 above
 * SGM_py), with the grid unchanged. */
int  smx_select_mps(smx_job *job, const smx_mps_args *args, smx_point3 *pts, long cap, long *count,
                    double *minimum_th);
/* SetHeightRange (setsm_code.cpp:9851-10261, setsm_code.hpp:122): the
 * triangles of the TIN over the matched points (built by the caller's host
 * code, as SETSM's TINCreate_list does) give every cell its next height
 * range, height and Matched_flag; the job's GridPT3 becomes the grid the
 * reference returns. pts and tris are host arrays (a triangle indexing past
 * npts is skipped, as in the reference). minH_grid / maxH_grid (may be NULL)
 * get the reference's range bounds of the grid. */
int  smx_set_height_range(smx_job *job, const smx_point3 *pts, long npts, const smx_tri *tris, long ntris,
                          const smx_hr_args *args, double *minH_grid, double *maxH_grid);

/* DecisionMPs' blunder loop (setsm_code.cpp:8891-9093) calls these two
 * between its TIN updates; the loop and blunder_detection_TIN stay host code. */
/* SetHeightRange_blunder (setsm_code.cpp:10393-10491, setsm_code.hpp:120):
 * each cell's GridPT3.Height becomes the highest TIN height of the triangles
 * containing it; cells no triangle contains keep theirs. */
int  smx_set_height_range_blunder(smx_job *job, const smx_point3 *pts, long npts, const smx_tri *tris, long ntris);
/* The ortho grid of VerticalLineLocus_blunder: blunder_selected_level (the
 * job must hold both images' views at that level) and the (lon, lat) of
 * every ortho pixel at resolution * 2^level over the level's Boundary
 * (w = (Boundary[2]-Boundary[0])/res + 1 columns), as ps2wgs_single gives
 * them. Kept until the next smx_job_set_level. */
int  smx_job_set_blunder_level(smx_job *job, int blunder_level, const double *lonlat, uint32_t w, uint32_t h);
/* VerticalLineLocus_blunder (setsm_code.cpp:7882-8013): every cell's ortho
 * NCC at its GridPT3 Height on the blunder level's images, FP64 with the
 * reference's operation order. Writes GridPT3.ortho_ncc[1] and
 * Mean_ortho_ncc, and (if not NULL) the per-cell value to ortho_ncc. */
int  smx_vertical_line_locus_blunder(smx_job *job, int iteration, int bblunder, float *ortho_ncc);
/* blunder_detection_TIN (setsm_code.cpp:9158-9782, setsm_code.hpp:116): each
 * matched point that is not an old blunder (flag 1) is tested against its
 * first 30 triangles of the TIN (host arrays, as DecisionMPs holds them) and
 * the per-cell ortho NCC: ortho_ncc (host, one float per cell) or, if NULL,
 * the last smx_vertical_line_locus_blunder output on this level, which stays
 * on the device. New blunders end at flag 1 in pts and, if detected is not
 * NULL, at 1 in detected (npts bytes, all written); GridPT3.anchor_flag is
 * reset or set at each tested point's cell. blunder_count[0] (kept points)
 * and [1] (blunders) and *minz_mp / *maxz_mp (may be NULL: the kept points'
 * heights) continue from the values given, as DecisionMPs' running totals
 * do. A no-op at level 0 iteration 3, as in the reference. SMX_E_ARG if a
 * point lies outside the grid, before anything changes. The points are
 * tested on the GPU; the edge statistics the threshold comes from are
 * summed on the host in triangle order, as the reference sums them. */
int  smx_blunder_detection_tin(smx_job *job, const smx_bd_args *args, const float *ortho_ncc, smx_point3 *pts,
                               long npts, const smx_tri *tris, long ntris, uint8_t *detected, long blunder_count[2],
                               double *minz_mp, double *maxz_mp);
/* The five calls above leave their kernel time on the job stream (host
 * copies and host passes excluded) in smx_job_last_kernel_ms. */

/* A Delaunay TIN of matched points, host code, for callers without a TIN of
 * their own (the role of SETSM's TINCreate_list, SubFunctions.cpp:2768-2846;
 * this build's own triangulation, not the reference's). Points snap to the
 * lattice of min_max / resolution as the reference snaps them, the last
 * point of a lattice cell standing for it; triangles (indices into pts)
 * come out clockwise. Among cocircular points the split is its own.
 * Returns the triangle count (at most 2 * npts), 0 for fewer than 3 points
 * or a single line, -1 for a lattice beyond 32767 per side or a point
 * outside it, -2 if cap is too small, -3 for a bad argument, -5 if host
 * memory ran out. */
long smx_tin_create(const smx_point3 *pts, long npts, const double min_max[4], double resolution, smx_tri *out,
                    long cap);

/* Number of hypotheses in the cube (sum of NumOfHeight). */
long long smx_job_num_hypotheses(const smx_job *job);
/* Download the cube in CSR order: words[total], rho[total] (rho may be NULL,
 * NaN where not evaluated), offsets[ncells + 1] (may be NULL). */
int  smx_job_get_voxels(smx_job *job, uint32_t *words, float *rho, int64_t *offsets);
/* Download the rows of cells [cell0, cell1) only (contiguous in CSR order). */
int  smx_job_get_voxel_rows(smx_job *job, long long cell0, long long cell1, uint32_t *words, float *rho);
/* Replace the cube: words in CSR order, row lengths taken from the
 * NumOfHeight of the job's current NCCresult array (smx_job_set_nccresult),
 * offsets rebuilt from them. Hands a host cube (SETSM's grid_voxel rows,
 * setsm_code.cpp:5930-5948) to a job that did not build it. */
int  smx_job_set_voxels(smx_job *job, const uint32_t *words);

/* Launch shape of the SMX_PREC_FAST kernel (tile of cells per workgroup,
 * LDS rows per image, register budget): index in [0, smx_fast_num_configs()).
 * -1 restores the default. The autotuner in setsm_amd/autotune.py picks one. */
int  smx_fast_num_configs(void);
int  smx_job_set_fast_config(smx_job *job, int config);
/* Describe config `config` as "tx ty lds_rows o4" into buf. */
int  smx_fast_config_desc(int config, char *buf, int buflen);

/* Milliseconds the last smx_vertical_line_locus spent in its VLL kernel
 * (HIP events on the job's stream; FAST: the window kernel, without the FP64
 * pass over the runs it hands over, see smx_job_last_redo). */
double smx_job_last_kernel_ms(const smx_job *job);
/* Hypotheses the last smx_vertical_line_locus evaluated (passed every gate). */
long long smx_job_last_evaluated(const smx_job *job);
/* out4 = {evaluated hypotheses, FAST tiles staged in LDS, FAST waves that
 * read their taps from HBM, FAST launch shape index} of the last call. */
int  smx_job_last_stats(const smx_job *job, long long *out4);
/* The last SMX_PREC_FAST call's FP64 pass: runs (64 heights of one cell) the
 * FAST kernel handed over because a window's spread was not clear of FP32
 * rounding (flat patches, windows left with a few valid samples), the
 * hypotheses that pass evaluated (part of smx_job_last_evaluated) and its
 * milliseconds (not part of smx_job_last_kernel_ms). No counterpart in the
 * reference: statistics of this library's FAST path. */
int  smx_job_last_redo(const smx_job *job, long long *runs, long long *hypotheses, double *ms);

/* Wait for all work queued on the job's stream. */
int  smx_job_sync(smx_job *job);

#ifdef __cplusplus
}
#endif
#endif /* SMX_VLL_H */
