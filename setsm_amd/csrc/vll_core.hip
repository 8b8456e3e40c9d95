// vll_core.hip — per-cell gating, the plain FP64 VerticalLineLocus kernel,
// InitializeVoxel with the voxel-cube relayout, and the AWNCC peak scan.
//
// Built with -ffp-contract=off: every expression is evaluated in the operand
// order of the reference C++ so the FP64 kernel's INCC cube is bit-identical
// to SETSM's CPU path on the same inputs.
//
// Work decomposition: the cube is CSR (rows of NumOfHeight words per cell).
// The gate kernel marks cells whose height list must be evaluated; a prefix
// sum turns them into a flat list of (cell, first height) runs of 64 heights;
// one wavefront evaluates one run, one lane per height hypothesis.

#include <algorithm>

#include "smx_kernels.hpp"

namespace smx {

// ---------------------------------------------------------------- gating

__device__ __forceinline__ bool gate_cell(const VllParams *__restrict__ P, CellArgs &A, long c);

// setsm_code.cpp:6157-6204 + 6259-6279: which cells enter the INCC loop.
__global__ void __launch_bounds__(256) k_vll_gate(const VllParams *__restrict__ P, CellArgs A)
{
    // the accessible count: a ballot per wavefront, an LDS sum per block, one
    // global atomic per block
    __shared__ unsigned int s_acc;
    if (threadIdx.x == 0) s_acc = 0;
    __syncthreads();
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool accessible = false;
    if (c < A.ncells) accessible = gate_cell(P, A, c);
    const unsigned long long any = __ballot(accessible);
    if ((threadIdx.x & 63) == 0 && any) atomicAdd(&s_acc, (unsigned int)__popcll(any));
    __syncthreads();
    if (threadIdx.x == 0 && s_acc) atomicAdd(A.accessible, (unsigned long long)s_acc);
}

__device__ __forceinline__ bool gate_cell(const VllParams *__restrict__ P, CellArgs &A, long c)
{
    bool accessible = false;
    const smx_ugrid g = A.gridpt3[c];
    smx_nccresult *nr = &A.ncc[c];
    int runs = 0;
    const double lon = A.grid_wgs[2 * c], lat = A.grid_wgs[2 * c + 1];
    const int start_H = g.minHeight, end_H = g.maxHeight;
    if (check_image_boundary(P->rpc, P->adjust, P->view, P->level, lon, lat, (double)start_H, (double)end_H, P->half)) {
        bool blunder = false;
        if (P->level >= 2)
            blunder = false;
        else if (g.Matched_flag == 0)
            blunder = true;
        else
            blunder = !(abs(g.maxHeight - g.minHeight) < P->th_height);
        if (P->level <= 1 && (g.maxHeight > P->minmax[1] || g.minHeight < P->minmax[0]))
            blunder = true;
        if (P->check_matchtag && P->iteration <= 2) {
            blunder = false;
            if (g.maxHeight > P->minmax[1] || g.minHeight < P->minmax[0])
                blunder = true;
            else if (abs(g.maxHeight - g.minHeight) > P->th_height)
                blunder = true;
        }
        if (!blunder) {
            accessible = true;
            nr->GNCC = q_result(-1.0);   // k_gncc overwrites it where the ortho term applies
            const bool ra = P->is_ra || P->check_matching_rate;
            if (ra) {
                // setsm_code.cpp:6308-6313
                nr->check_height_change = 1;
                nr->NumOfHeight = (uint16_t)(int)((end_H - start_H) / P->height_step);
            }
            if (ra || nr->check_height_change || P->dem_lt2) {
                nr->result0 = q_result(-1.0);
                nr->result2 = -1000;
                nr->result1 = q_result(-1.0);
                nr->result3 = -1000;
                nr->result4 = 0;
                A.max_bits[c] = order_key(-100.0);
                runs = (nr->NumOfHeight + A.run_len - 1) / A.run_len;
            }
        }
    }
    A.run_count[c] = runs;
    if (A.acc_flag) A.acc_flag[c] = accessible ? 1 : 0;
    return accessible;
}

__global__ void __launch_bounds__(256) k_vll_fill_runs(CellArgs A)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A.ncells) return;
    const int n = A.run_count[c];
    const long base = A.run_offset[c];
    // the capacity is an upper bound on the count; never write past it
    for (int i = 0; i < n && base + i < A.runs_cap; i++)
        A.runs[base + i] = make_int2((int)c, i * A.run_len);
}

// Per-cell height lists -> the bounds the launch sequences size their
// scratch by: stats[k] += runs of kRunLens[k] heights (k = 0, 1, 2) and
// stats[3] = the longest list.
__global__ void __launch_bounds__(256) k_nh_stats(const smx_nccresult *__restrict__ ncc, long n,
                                                  unsigned long long *stats)
{
    // grid-stride sums in registers, a shuffle reduction per wavefront, LDS
    // atomics per block, then one global atomic per block and counter (the
    // four counters are shared by the whole launch)
    __shared__ unsigned long long s_runs[3];
    __shared__ unsigned int s_max;
    if (threadIdx.x == 0) {
        s_runs[0] = s_runs[1] = s_runs[2] = 0;
        s_max = 0;
    }
    unsigned long long runs[3] = {0, 0, 0};
    unsigned nh = 0;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
        const unsigned v = ncc[c].NumOfHeight;
#pragma unroll
        for (int k = 0; k < 3; k++) runs[k] += (v + kRunLens[k] - 1) / kRunLens[k];
        nh = max(nh, v);
    }
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; k++) runs[k] += (unsigned long long)__shfl_xor((long long)runs[k], o);
        nh = max(nh, (unsigned)__shfl_xor((int)nh, o));
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) atomicAdd(&s_runs[k], runs[k]);
        atomicMax(&s_max, nh);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; k++)
            if (s_runs[k]) atomicAdd(&stats[k], s_runs[k]);
        if (s_max) atomicMax(&stats[3], (unsigned long long)s_max);
    }
}

// Fills (the launch sequences clear their buffers with kernels, not runtime
// memsets, so that they can be captured in a hipGraph).
__global__ void __launch_bounds__(256) k_fill_u64(unsigned long long *p, unsigned long long v, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *p, uint32_t v, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// setsm_code.cpp:6509-6510: max_WNCC of cells whose heights were evaluated.
__global__ void __launch_bounds__(256) k_vll_finalize(CellArgs A)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A.ncells) return;
    if (A.run_count[c] > 0 && A.ncc[c].check_height_change)
        A.ncc[c].max_WNCC = q_result(order_value(A.max_bits[c]));
}

// setsm_code.cpp:6672-6752 FindPeakNcc over one cell's heights in order (the
// matching-rate pass calls it inline per evaluated hypothesis); one thread
// per cell over the temp_rho values the NCC kernel kept.
__global__ void __launch_bounds__(256) k_find_peak(const VllParams *__restrict__ P, CellArgs A)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A.ncells || A.run_count[c] == 0) return;
    smx_nccresult nr = A.ncc[c];
    const int start_H = A.gridpt3[c].minHeight;
    const double *rr = A.peak_rho + A.run_offset[c] * kRun;
    double pre_rho = -1.0;
    float pre_height = 0.0f;
    int direction = 0;
    bool check_rho = false;
    for (int h = 0; h < (int)nr.NumOfHeight; h++) {
        const double temp_rho = rr[h];
        if ((unsigned long long)__double_as_longlong(temp_rho) == kNotEvaluated) continue;
        const float iter_height = (float)(start_H + h * P->height_step);
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
            check_rho = true;
            t_direction = -1;
        } else if (pre_rho != -1) {
            if (direction > 0 && t_direction < 0) {
                if (nr.result4 < 10) nr.result4 += 1;
                if (nr.result0 < q_result(pre_rho)) {
                    const double temp_1 = dq(nr.result0);
                    nr.result0 = q_result(pre_rho);
                    const float temp_2 = nr.result2;
                    nr.result2 = pre_height;
                    if (nr.result1 < q_result(temp_1)) {
                        nr.result1 = q_result(temp_1);
                        nr.result3 = temp_2;
                    }
                } else if (nr.result1 < q_result(pre_rho)) {
                    nr.result1 = q_result(pre_rho);
                    nr.result3 = pre_height;
                }
            }
        }
        pre_rho = temp_rho;
        pre_height = iter_height;
        direction = t_direction;
    }
    A.ncc[c].result0 = nr.result0;
    A.ncc[c].result1 = nr.result1;
    A.ncc[c].result2 = nr.result2;
    A.ncc[c].result3 = nr.result3;
    A.ncc[c].result4 = nr.result4;
}

// ------------------------------------------------------ FP64 plain kernel

// SubFunctions.cpp:3317-3326 InterpolatePatch on both channels of a packed
// texel quad.
__device__ __forceinline__ void interp2(gptr_t<uint32_t> T, long pos, uint32_t w, double dx, double dy, double &v,
                                        double &m)
{
    const uint32_t a = T[pos], b = T[pos + 1], c = T[pos + w], d = T[pos + 1 + w];
    v = (double)(a & 0xffffu) * (1 - dx) * (1 - dy) + (double)(b & 0xffffu) * dx * (1 - dy) +
        (double)(c & 0xffffu) * (1 - dx) * dy + (double)(d & 0xffffu) * dx * dy;
    m = (double)(a >> 16) * (1 - dx) * (1 - dy) + (double)(b >> 16) * dx * (1 - dy) + (double)(c >> 16) * (1 - dx) * dy +
        (double)(d >> 16) * dx * dy;
}

// SubFunctions.cpp:3249-3273: bounds test, then the four bilinear samples;
// the sample is kept only if both intensities are > 0.
__device__ __forceinline__ bool sample_fp64(const DevView &Lv, const DevView &Rv, double lx, double ly, double rx,
                                            double ry, double &l, double &lm, double &r, double &rm)
{
    if (ry >= 0 && ry + 1 < Rv.height && rx >= 0 && rx + 1 < Rv.width && ly >= 0 && ly + 1 < Lv.height && lx >= 0 &&
        lx + 1 < Lv.width) {
        const long pl = (long)lx + (long)ly * (long)Lv.width;
        const long pr = (long)rx + (long)ry * (long)Rv.width;
        interp2(gptr(Lv.texel), pl, Lv.width, lx - floor(lx), ly - floor(ly), l, lm);
        interp2(gptr(Rv.texel), pr, Rv.width, rx - floor(rx), ry - floor(ry), r, rm);
        return l > 0 && r > 0;
    }
    return false;
}

// SubFunctions.cpp:974-1018 Correlate from the two passes' sums.
__device__ __forceinline__ double corr_from(double sLR, double sL2, double sR2)
{
    if (sL2 > 1e-8 && sR2 > 1e-8) return sLR / (sqrt(sL2 * sR2));
    return -99.0;
}

// One level's multi-window NCC for one hypothesis: SetVecKernelValue over the
// window (setsm_code.cpp:6429-6453) and ComputeMultiNCC (SubFunctions.cpp:
// 3328-3359), two passes like Correlate. Returns true if all 3 counts > 0.
// Sample positions of the hypothesis window: the left patch on the integer
// grid around the left centre, the right patch rotated by -theta about the
// right centre (setsm_code.cpp:6431-6436).
struct AffinePos {
    double rpx, rpy, tpx, tpy, cos0, sin0;
    __device__ __forceinline__ bool operator()(int row, int col, double &lx, double &ly, double &rx,
                                               double &ry) const
    {
        lx = rpx + col;
        ly = rpy + row;
        const double ox = cos0 * col - sin0 * row, oy = sin0 * col + cos0 * row;
        rx = tpx + ox;
        ry = tpy + oy;
        return true;
    }
};

template <int H, class Pos>
__device__ __forceinline__ bool multi_ncc_fp64_at(const DevView &Lv, const DevView &Rv, const Pos &pos,
                                                  double &sum_multi, double &count_multi);

template <int H>
__device__ __forceinline__ bool multi_ncc_fp64(const DevView &Lv, const DevView &Rv, double rpx, double rpy, double tpx,
                                               double tpy, double cos0, double sin0, double &sum_multi,
                                               double &count_multi)
{
    return multi_ncc_fp64_at<H>(Lv, Rv, AffinePos{rpx, rpy, tpx, tpy, cos0, sin0}, sum_multi, count_multi);
}

template <int H, class Pos>
__device__ __forceinline__ bool multi_ncc_fp64_at(const DevView &Lv, const DevView &Rv, const Pos &pos,
                                                  double &sum_multi, double &count_multi)
{
    double sL[3] = {0, 0, 0}, sLm[3] = {0, 0, 0}, sR[3] = {0, 0, 0}, sRm[3] = {0, 0, 0};
    int cnt[3] = {0, 0, 0};
#pragma unroll 1
    for (int row = -H; row <= H; row++) {
#pragma unroll
        for (int col = -H; col <= H; col++) {
            const int bits = window_bits_ce(H, row, col);
            double lx, ly, rx, ry;
            if (bits && pos(row, col, lx, ly, rx, ry)) {
                double l, lm, r, rm;
                if (sample_fp64(Lv, Rv, lx, ly, rx, ry, l, lm, r, rm)) {
#pragma unroll
                    for (int k = 0; k < 3; k++)
                        if (bits & (1 << k)) {
                            sL[k] += l; sLm[k] += lm; sR[k] += r; sRm[k] += rm; cnt[k]++;
                        }
                }
            }
        }
    }
    if (!(cnt[0] > 0 && cnt[1] > 0 && cnt[2] > 0)) return false;
    double mL[3], mLm[3], mR[3], mRm[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        mL[k] = sL[k] / cnt[k]; mR[k] = sR[k] / cnt[k];
        mLm[k] = sLm[k] / cnt[k]; mRm[k] = sRm[k] / cnt[k];
    }
    double LR[3] = {0, 0, 0}, L2[3] = {0, 0, 0}, R2[3] = {0, 0, 0};
    double mLR[3] = {0, 0, 0}, mL2[3] = {0, 0, 0}, mR2[3] = {0, 0, 0};
#pragma unroll 1
    for (int row = -H; row <= H; row++) {
#pragma unroll
        for (int col = -H; col <= H; col++) {
            const int bits = window_bits_ce(H, row, col);
            double lx, ly, rx, ry;
            if (bits && pos(row, col, lx, ly, rx, ry)) {
                double l, lm, r, rm;
                if (sample_fp64(Lv, Rv, lx, ly, rx, ry, l, lm, r, rm)) {
#pragma unroll
                    for (int k = 0; k < 3; k++)
                        if (bits & (1 << k)) {
                            LR[k] += (l - mL[k]) * (r - mR[k]);
                            L2[k] += (l - mL[k]) * (l - mL[k]);
                            R2[k] += (r - mR[k]) * (r - mR[k]);
                            mLR[k] += (lm - mLm[k]) * (rm - mRm[k]);
                            mL2[k] += (lm - mLm[k]) * (lm - mLm[k]);
                            mR2[k] += (rm - mRm[k]) * (rm - mRm[k]);
                        }
                }
            }
        }
    }
    double temp_roh = 0, count_roh = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double ncc = corr_from(LR[k], L2[k], R2[k]);
        if (ncc != -99) { count_roh++; temp_roh += ncc; }
        const double ncm = corr_from(mLR[k], mL2[k], mR2[k]);
        if (ncm != -99) { count_roh++; temp_roh += ncm; }
    }
    if (count_roh > 0) {
        sum_multi += temp_roh / count_roh;
        count_multi++;
    }
    return true;
}

// One run (up to 64 consecutive heights of one cell) on one wavefront, FP64.
// Returns the run's count of hypotheses that computed an NCC (uniform).
template <int H>
__device__ __forceinline__ unsigned vll_fp64_run(const VllParams *__restrict__ P, const CellArgs &A, const int2 run,
                                             const long wave, const int lane)
{
    const long cell = run.x;
    const int hidx = run.y + lane;
    const smx_nccresult nr = A.ncc[cell];
    const smx_ugrid g = A.gridpt3[cell];
    const int start_H = g.minHeight, end_H = g.maxHeight;

    // lanes past the run's length (runs of 16 or 32 handed over by the FAST
    // path) stay idle
    Hyp hy;
    bool act = lane < A.run_len && project_hypothesis(*P, A, cell, hidx, nr, start_H, end_H, hy);
    double temp_rho = -100.0;
    unsigned evaluated = 0;
    if (A.peak_rho) {
        // RA / matching-rate pass (setsm_code.cpp:6468-6496): no voxel cube;
        // temp_rho kept for the per-cell FindPeakNcc scan
        double pr = __longlong_as_double(kNotEvaluated);
        if (act && hy.orient_ok) {
            double sum_multi = 0, count_multi = 0;
            multi_ncc_fp64<H>(P->view[0], P->view[1], hy.rpx, hy.rpy, hy.tpx, hy.tpy, hy.cos0, hy.sin0, sum_multi,
                              count_multi);
            if (P->combined_incc)
                multi_ncc_fp64<H>(P->next[0], P->next[1], hy.rnx, hy.rny, hy.tnx, hy.tny, hy.cos0, hy.sin0,
                                  sum_multi, count_multi);
            const double rho = count_multi > 0 ? sum_multi / count_multi : -1.0;
            // the value FindPeakNcc receives (INCC/GNCC blend where the term
            // exists; an evaluated NaN stays NaN), and max_WNCC's
            pr = A.gval && A.gval[cell] == A.gval[cell]
                     ? sum_multi / count_multi * A.ncc_alpha + A.gval[cell] * (1.0 - A.ncc_alpha)
                     : rho;
            temp_rho = wncc_of(A, cell, sum_multi, count_multi, rho);
            evaluated = 1;
        }
        A.peak_rho[wave * kRun + lane] = pr;
        return wave_max_commit(temp_rho, evaluated, &A.max_bits[cell]);
    }
    if (act) {
        const long wi = A.vox_off[cell] + hidx;
        uint32_t word = A.vox[wi];
        if (!hy.orient_ok) {
            A.vox[wi] = word & ~0xffu;   // flag_cal = false (setsm_code.cpp:6503-6506)
        } else {
            double sum_multi = 0, count_multi = 0;
            const bool all3 = multi_ncc_fp64<H>(P->view[0], P->view[1], hy.rpx, hy.rpy, hy.tpx, hy.tpy, hy.cos0,
                                                hy.sin0, sum_multi, count_multi);
            if (P->combined_incc)
                multi_ncc_fp64<H>(P->next[0], P->next[1], hy.rnx, hy.rny, hy.tnx, hy.tny, hy.cos0, hy.sin0,
                                  sum_multi, count_multi);
            const double rho = count_multi > 0 ? sum_multi / count_multi : -1.0;
            word = (word & 0xffu) | (all3 ? 1u : 0u);
            word |= (uint32_t)(uint16_t)q_voxel(rho) << 16;
            A.vox[wi] = word;
            if (A.rho) A.rho[wi] = (float)rho;
            temp_rho = wncc_of(A, cell, sum_multi, count_multi, rho);
            evaluated = 1;
        }
    }
    return wave_max_commit(temp_rho, evaluated, &A.max_bits[cell]);
}

// One wave per run, grid-stride over the count on the device (the launch
// covers the capacity up to a cap: a grid of one wave per run would pass
// 2^32 work-items at a few 10^9 hypotheses).
template <int H>
__global__ void __launch_bounds__(256) k_vll_fp64(const VllParams *__restrict__ P, CellArgs A)
{
    const long nw = ((long)gridDim.x * blockDim.x) >> 6;
    const long n = min((long)*A.nruns_dev, A.runs_cap);
    unsigned long long cnt = 0;
    for (long q = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < n; q += nw)
        cnt += vll_fp64_run<H>(P, A, A.runs[q], q, threadIdx.x & 63);
    wave_count_commit(A.evaluated, cnt);
}

// The runs the FAST kernel handed over (a window whose spread it cannot tell
// from FP32 rounding), recomputed in full. The count is read on the device,
// so the launch needs no host round trip; an empty queue costs one load per
// wave.
template <int H>
__global__ void __launch_bounds__(256) k_vll_redo(const VllParams *__restrict__ P, CellArgs A)
{
    const long nw = ((long)gridDim.x * blockDim.x) >> 6;
    const long n = (long)*A.redo_n;
    unsigned long long cnt = 0;
    for (long q = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < n; q += nw)
        cnt += vll_fp64_run<H>(P, A, A.redo[q], q, threadIdx.x & 63);
    wave_count_commit(A.evaluated, cnt);
}

template __global__ void k_vll_fp64<5>(const VllParams *, CellArgs);
template __global__ void k_vll_fp64<6>(const VllParams *, CellArgs);
template __global__ void k_vll_fp64<7>(const VllParams *, CellArgs);

// ------------------------------------------------------------- GNCC

// setsm_code.cpp:6550-6670 SetOrthoImageCoord (SB sensor, OR pyramid): one
// thread per ortho pixel; lon/lat of the pixel come from the caller.
__global__ void __launch_bounds__(256) k_ortho_coords(const VllParams *__restrict__ P, OrthoArgs O)
{
    const long it = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= O.sw * O.sh) return;
    const long pts_row = it / O.sw;
    const long pts_col = it % O.sw;
    const double t_X = O.b0 + pts_col * O.im_res;
    const double t_Y = O.b1 + pts_row * O.im_res;
    const long t_col = (long)((t_X - O.b0) / O.grid_res);
    const long t_row = (long)((t_Y - O.b1) / O.grid_res);
    const long pt_index = t_row * O.W + t_col;
    const long pim = pts_row * O.sw + pts_col;
    long t_col_next = 0, t_row_next = 0, pim_next = 0;
    if (O.combined) {
        t_col_next = (long)((t_X - O.b0) / O.im_res_next);
        t_row_next = (long)((t_Y - O.b1) / O.im_res_next);
        pim_next = t_row_next * O.swn + t_col_next;
    }
    if (!(pt_index < O.W * O.H && t_col < O.W && t_row < O.H)) return;
    const float Hgt = O.gridpt3[pt_index].Height;
    if (Hgt == -1000) return;
    const double lon = O.ortho_wgs[2 * it], lat = O.ortho_wgs[2 * it + 1];
    const double Z = (double)Hgt;
#pragma unroll 1
    for (int k = 0; k < 2; k++) {
        double ix, iy, px, py;
        rpc_ground_to_image(P->rpc[k], P->adjust[k], lon, lat, Z, ix, iy);
        to_pyramid(ix, iy, P->view[k], P->level, px, py);
        O.cd[k][2 * pim] = px;
        O.cd[k][2 * pim + 1] = py;
        if (O.combined) {
            to_pyramid(ix, iy, P->next[k], P->level - 1, px, py);
            if (pim_next < O.swn * O.shn && t_col_next < O.swn && t_row_next < O.shn) {
                O.cdn[k][2 * pim_next] = px;
                O.cdn[k][2 * pim_next + 1] = py;
            }
        }
    }
}

// Sample positions of the GNCC window of one cell (setsm_code.cpp:
// 6250-6289): window offsets on the ortho grid, positions from the ortho
// image coordinates; absent where the reference skips SetVecKernelValue.
struct OrthoPos {
    const OrthoArgs *O;
    long cell;
    bool next;
    __device__ __forceinline__ bool operator()(int row, int col, double &lx, double &ly, double &rx,
                                               double &ry) const
    {
        const double t_X = O->grid_pts[2 * cell] + col * O->im_res;
        const double t_Y = O->grid_pts[2 * cell + 1] + row * O->im_res;
        long t_col = (long)((t_X - O->b0) / O->im_res);
        long t_row = (long)((t_Y - O->b1) / O->im_res);
        long pit = t_row * O->sw + t_col;
        const long tt_col = (long)((t_X - O->b0) / O->grid_res);
        const long tt_row = (long)((t_Y - O->b1) / O->grid_res);
        const long pdem = tt_row * O->W + tt_col;
        if (!(pit >= 0 && pit < O->sw * O->sh && t_col >= 0 && t_col < O->sw && t_row >= 0 && t_row < O->sh &&
              pdem >= 0 && pdem < O->W * O->H && tt_col >= 0 && tt_col < O->W && tt_row >= 0 && tt_row < O->H))
            return false;
        if (O->gridpt3[pdem].Height == -1000) return false;
        if (!next) {
            lx = O->cd[0][2 * pit];
            ly = O->cd[0][2 * pit + 1];
            rx = O->cd[1][2 * pit];
            ry = O->cd[1][2 * pit + 1];
            return true;
        }
        t_col = (long)((t_X - O->b0) / O->im_res_next);
        t_row = (long)((t_Y - O->b1) / O->im_res_next);
        pit = t_row * O->swn + t_col;
        if (!(pit >= 0 && pit < O->swn * O->shn && t_col < O->swn && t_row < O->shn)) return false;
        lx = O->cdn[0][2 * pit];
        ly = O->cdn[0][2 * pit + 1];
        rx = O->cdn[1][2 * pit];
        ry = O->cdn[1][2 * pit + 1];
        return true;
    }
};

// setsm_code.cpp:6240-6305: GNCC of every cell that passed the gate and whose
// previous ortho NCC clears ortho_th; one thread per cell. Writes
// NCCresult.GNCC and the unquantised value the INCC blend uses.
template <int H>
__global__ void __launch_bounds__(256) k_gncc(const VllParams *__restrict__ P, OrthoArgs O, CellArgs A, double *gval)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A.ncells) return;
    double g = __builtin_nan("");
    if (A.acc_flag[c] && dq(A.gridpt3[c].ortho_ncc[O.slot]) > O.ortho_th) {
        double sum = 0, count = 0;
        multi_ncc_fp64_at<H>(P->view[0], P->view[1], OrthoPos{&O, c, false}, sum, count);
        if (O.combined) multi_ncc_fp64_at<H>(P->next[0], P->next[1], OrthoPos{&O, c, true}, sum, count);
        if (count > 0) {
            g = sum / count;
            A.ncc[c].GNCC = q_result(g);
        } else {
            A.ncc[c].GNCC = q_result(-1.0);
        }
    }
    gval[c] = g;
}

// ------------------------------------------------------ VerticalLineLocus_blunder

// SetOrthoImageCoord (setsm_code.cpp:6550-6670) at the blunder level
// (check_pyimage = BD): the ortho pixels' positions in that level's pyramid.
__global__ void __launch_bounds__(256) k_bd_coords(const VllParams *__restrict__ P, BlunderArgs B)
{
    const long it = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= B.sw * B.sh) return;
    const long pts_row = it / B.sw, pts_col = it % B.sw;
    const double t_X = B.b0 + pts_col * B.im_res;
    const double t_Y = B.b1 + pts_row * B.im_res;
    const long t_col = (long)((t_X - B.b0) / B.grid_res);
    const long t_row = (long)((t_Y - B.b1) / B.grid_res);
    const long pt_index = t_row * B.W + t_col;
    if (!(pt_index < B.W * B.H && t_col < B.W && t_row < B.H)) return;
    const float Hgt = B.gridpt3[pt_index].Height;
    if (Hgt == -1000) return;
    const double lon = B.ortho_wgs[2 * it], lat = B.ortho_wgs[2 * it + 1];
#pragma unroll 1
    for (int k = 0; k < 2; k++) {
        double ix, iy, px, py;
        rpc_ground_to_image(P->rpc[k], P->adjust[k], lon, lat, (double)Hgt, ix, iy);
        to_pyramid(ix, iy, B.view[k], B.level_b, px, py);
        B.cd[k][2 * it] = px;
        B.cd[k][2 * it + 1] = py;
    }
}

// setsm_code.cpp:7929-8000, one thread per cell: the multi-window NCC over
// the disc of radius Half_template_size - 1 on the ortho grid, sampled on the
// blunder level's images at the cells' TIN heights, two passes like
// Correlate (the window size is a run-time value here).
__global__ void __launch_bounds__(256) k_vll_blunder(BlunderArgs B)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= B.W * B.H) return;
    const int H = B.half, r2max = (H - 1) * (H - 1);
    const double gx = B.grid_pts[2 * c], gy = B.grid_pts[2 * c + 1];
    auto pos = [&](int row, int col, double &lx, double &ly, double &rx, double &ry) -> bool {
        const double t_X = gx + col * B.im_res;
        const double t_Y = gy + row * B.im_res;
        const long t_col = (long)((t_X - B.b0) / B.im_res);
        const long t_row = (long)((t_Y - B.b1) / B.im_res);
        const long pit = t_row * B.sw + t_col;
        const long tt_col = (long)((t_X - B.b0) / B.grid_res);
        const long tt_row = (long)((t_Y - B.b1) / B.grid_res);
        const long pdem = tt_row * B.W + tt_col;
        if (!(pit >= 0 && pit < B.sw * B.sh && t_col >= 0 && t_col < B.sw && t_row >= 0 && t_row < B.sh &&
              pdem >= 0 && pdem < B.W * B.H && tt_col >= 0 && tt_col < B.W && tt_row >= 0 && tt_row < B.H))
            return false;
        if (B.gridpt3[pdem].Height == -1000) return false;
        lx = B.cd[0][2 * pit];
        ly = B.cd[0][2 * pit + 1];
        rx = B.cd[B.right][2 * pit];
        ry = B.cd[B.right][2 * pit + 1];
        return true;
    };
    const DevView &Lv = B.view[0], &Rv = B.view[B.right];
    double sL[3] = {0, 0, 0}, sLm[3] = {0, 0, 0}, sR[3] = {0, 0, 0}, sRm[3] = {0, 0, 0};
    int cnt[3] = {0, 0, 0};
    for (int row = -H; row <= H; row++)
        for (int col = -H; col <= H; col++) {
            if (row * row + col * col > r2max) continue;
            const int bits = window_bits_ce(H, row, col);
            double lx, ly, rx, ry, l, lm, r, rm;
            if (!pos(row, col, lx, ly, rx, ry) || !sample_fp64(Lv, Rv, lx, ly, rx, ry, l, lm, r, rm)) continue;
            for (int k = 0; k < 3; k++)
                if (bits & (1 << k)) {
                    sL[k] += l; sLm[k] += lm; sR[k] += r; sRm[k] += rm; cnt[k]++;
                }
        }
    double t = 0;
    if (cnt[0] > 0 && cnt[1] > 0 && cnt[2] > 0) {
        double mL[3], mLm[3], mR[3], mRm[3];
        for (int k = 0; k < 3; k++) {
            mL[k] = sL[k] / cnt[k]; mR[k] = sR[k] / cnt[k];
            mLm[k] = sLm[k] / cnt[k]; mRm[k] = sRm[k] / cnt[k];
        }
        double LR[3] = {0, 0, 0}, L2[3] = {0, 0, 0}, R2[3] = {0, 0, 0};
        double mLR[3] = {0, 0, 0}, mL2[3] = {0, 0, 0}, mR2[3] = {0, 0, 0};
        for (int row = -H; row <= H; row++)
            for (int col = -H; col <= H; col++) {
                if (row * row + col * col > r2max) continue;
                const int bits = window_bits_ce(H, row, col);
                double lx, ly, rx, ry, l, lm, r, rm;
                if (!pos(row, col, lx, ly, rx, ry) || !sample_fp64(Lv, Rv, lx, ly, rx, ry, l, lm, r, rm)) continue;
                for (int k = 0; k < 3; k++)
                    if (bits & (1 << k)) {
                        LR[k] += (l - mL[k]) * (r - mR[k]);
                        L2[k] += (l - mL[k]) * (l - mL[k]);
                        R2[k] += (r - mR[k]) * (r - mR[k]);
                        mLR[k] += (lm - mLm[k]) * (rm - mRm[k]);
                        mL2[k] += (lm - mLm[k]) * (lm - mLm[k]);
                        mR2[k] += (rm - mRm[k]) * (rm - mRm[k]);
                    }
            }
        double temp_roh = 0, count_roh = 0;
        for (int k = 0; k < 3; k++) {
            const double ncc = corr_from(LR[k], L2[k], R2[k]);
            if (ncc != -99) { count_roh++; temp_roh += ncc; }
            const double ncm = corr_from(mLR[k], mL2[k], mR2[k]);
            if (ncm != -99) { count_roh++; temp_roh += ncm; }
        }
        if (count_roh > 0) t += temp_roh / count_roh;
    }
    double max_ncc = -100;
    if (max_ncc < t) max_ncc = t;
    B.gridpt3[c].ortho_ncc[1] = q_result(t);
    B.gridpt3[c].Mean_ortho_ncc = q_result(max_ncc);
    B.out[c] = (float)max_ncc;
}

// ------------------------------------------------------ InitializeVoxel

// setsm_code.cpp:5759-6012, one thread per cell. Writes the new NCCresult
// fields, possibly rewrites GridPT3's band, and leaves an action per cell:
// 0 = no row, 1 = keep the old row, 2 = fresh row (flag 1, INCC -1000).
__global__ void __launch_bounds__(256) k_init_voxel(const VllParams *__restrict__ P, InitArgs A)
{
    const long t_i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t_i >= A.ncells) return;
    smx_nccresult nr = A.ncc[t_i];
    smx_ugrid g = A.gridpt3[t_i];
    const int iteration = P->iteration;
    const int pyramid_step = P->level;
    const double height_step = P->height_step;
    const double *minmaxHeight = P->minmax;
    int action = nr.NumOfHeight > 0 ? 1 : 0;
    const double lon = A.grid_wgs[2 * t_i], lat = A.grid_wgs[2 * t_i + 1];

    auto drop = [&]() {
        nr.NumOfHeight = 0;
        action = 0;
    };

    if (check_image_boundary(P->rpc, P->adjust, P->view, pyramid_step, lon, lat, (double)g.minHeight,
                             (double)g.maxHeight, 7)) {
        int change_step_min = 0, change_step_max = 0;
        bool check_blunder_cell = true;
        const double th_height = P->th_height;
        if (P->check_matchtag || iteration == 1) {
            nr.minHeight = g.minHeight;
            nr.maxHeight = g.maxHeight;
            nr.GNCC = q_result(-1.0);
            nr.check_height_change = 1;
            check_blunder_cell = false;
        } else {
            if (pyramid_step >= 2)
                check_blunder_cell = false;
            else if (g.Matched_flag != 0) {
                if (g.maxHeight - g.minHeight > 0) {
                    if (pyramid_step <= 1) {
                        if (iteration > 1)
                            check_blunder_cell = !(g.maxHeight <= minmaxHeight[1] && g.minHeight >= minmaxHeight[0] &&
                                                   nr.maxHeight <= minmaxHeight[1] && nr.minHeight >= minmaxHeight[0]);
                        else
                            check_blunder_cell = !(g.maxHeight <= minmaxHeight[1] && g.minHeight >= minmaxHeight[0]);
                    }
                    if (pyramid_step == 1) check_blunder_cell = !(abs(g.maxHeight - g.minHeight) < 1000);
                    if (pyramid_step == 0) {
                        if (iteration > 1)
                            check_blunder_cell = !((abs(nr.maxHeight - nr.minHeight) < th_height) &&
                                                   (abs(g.maxHeight - g.minHeight) < th_height));
                        else {
                            if (abs(g.maxHeight - g.minHeight) < th_height) check_blunder_cell = false;
                            else nr.NumOfHeight = 0;
                        }
                    }
                } else {
                    drop();
                    check_blunder_cell = true;
                }
            } else {
                drop();
                check_blunder_cell = true;
            }
        }

        if (!check_blunder_cell) {
            if (iteration == 1 || (pyramid_step == 0 && P->dem_lt2)) {
                nr.minHeight = g.minHeight;
                nr.maxHeight = g.maxHeight;
                nr.GNCC = q_result(-1.0);
                nr.check_height_change = 1;
                if (pyramid_step == 1) {
                    if (abs((int)g.maxHeight - (int)g.minHeight) > 1000) {
                        g.maxHeight = -100; g.minHeight = -100;
                        nr.check_height_change = 0;
                    }
                } else if (pyramid_step == 0) {
                    if (abs((int)g.maxHeight - (int)g.minHeight) > th_height) {
                        g.maxHeight = -100; g.minHeight = -100;
                        nr.check_height_change = 0;
                    }
                }
            } else {
                const int NumberofHeightVoxel = (int)((g.maxHeight - g.minHeight) / height_step);
                if (nr.NumOfHeight == 0 && NumberofHeightVoxel > 0) {
                    change_step_min = 0;
                    change_step_max = 0;
                    nr.check_height_change = 1;
                } else {
                    if (nr.minHeight > g.minHeight || nr.maxHeight < g.maxHeight) {
                        if (nr.minHeight > g.minHeight)
                            change_step_min = (int)((nr.minHeight - g.minHeight) / height_step + 0.5);
                        if (nr.maxHeight < g.maxHeight)
                            change_step_max = (int)((g.maxHeight - nr.maxHeight) / height_step + 0.5);
                        nr.minHeight = (int16_t)floor(nr.minHeight - change_step_min * height_step);
                        nr.maxHeight = (int16_t)ceil(nr.maxHeight + change_step_max * height_step);
                        nr.check_height_change = 1;
                        if (pyramid_step == 0 && abs(nr.maxHeight - nr.minHeight) > th_height)
                            nr.check_height_change = 0;
                    } else
                        nr.check_height_change = 0;
                }
            }
            if (nr.check_height_change) {
                nr.minHeight = (int16_t)floor(nr.minHeight - change_step_min * height_step);
                nr.maxHeight = (int16_t)ceil(nr.maxHeight + change_step_max * height_step);
                const int NumberofHeightVoxel = (int)((float)(nr.maxHeight - nr.minHeight) / height_step);
                if (NumberofHeightVoxel > 0) {
                    nr.NumOfHeight = (uint16_t)NumberofHeightVoxel;
                    action = 2;
                } else {
                    drop();
                    nr.check_height_change = 0;
                }
            }
        } else {
            drop();
            nr.check_height_change = 0;
        }
    } else {
        drop();
        nr.check_height_change = 0;
    }
    if (nr.minHeight == 0 || nr.maxHeight == 0) {
        drop();
        nr.check_height_change = 0;
    }
    if (pyramid_step == 0 && nr.NumOfHeight > 1000) {
        drop();
        nr.check_height_change = 0;
    }
    if (nr.NumOfHeight == 0) action = 0;
    A.ncc[t_i] = nr;
    A.gridpt3[t_i] = g;
    A.action[t_i] = (uint8_t)action;
    A.new_len[t_i] = nr.NumOfHeight;
}

// Re-lays the cube after InitializeVoxel: one wavefront per cell copies or
// initialises its row in 64-word strides.
__global__ void __launch_bounds__(256) k_relayout(InitArgs A)
{
    const int lane = threadIdx.x & 63;
    const long nw = ((long)gridDim.x * blockDim.x) >> 6;
    for (long c = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < A.ncells; c += nw) {
        const int act = A.action[c];
        const int n = A.new_len[c];
        if (act == 0 || n == 0) continue;
        const long dst = A.new_off[c];
        const long src = A.old_off[c];
        for (int h = lane; h < n; h += 64) {
            uint32_t w = SMX_VOXEL_WORD(1, -1000);
            float r = __builtin_nanf("");
            if (act == 1) {
                w = A.old_vox[src + h];
                if (A.old_rho) r = A.old_rho[src + h];
            }
            A.new_vox[dst + h] = w;
            if (A.new_rho) A.new_rho[dst + h] = r;
        }
    }
}

// ------------------------------------------------------------ AWNCC peak

// setsm_code.cpp:7527-7702; one thread per cell scanning its heights in
// order. With A.sum (check_SGM_peak, Pyramid_step <= SGM_py) the peaks are
// those of SumCost, and at level 0 from iteration 2 on result0 is the INCC
// (or blended) value at the chosen peak.
__global__ void __launch_bounds__(256) k_awncc(const VllParams *__restrict__ P, PeakArgs A)
{
    // The block's rows are staged through LDS in chunks of kChunk heights:
    // consecutive threads load consecutive words of one row (the rows of
    // neighbouring cells are neighbours in CSR), then each lane scans its own
    // row from LDS. Rows are padded to kChunk + 1 words so the lanes' reads
    // of one step fall in different banks.
    constexpr int kChunk = 16, kPitch = kChunk + 1;
    __shared__ uint32_t s_word[256 * kPitch];
    __shared__ float s_sum[256 * kPitch];
    __shared__ long s_base[256];
    __shared__ int s_nh[256];
    __shared__ int s_maxnh;
    const long c0 = (long)blockIdx.x * blockDim.x;
    const long c = c0 + threadIdx.x;
    const bool valid = c < A.ncells;
    smx_nccresult nr = A.ncc[valid ? c : c0];
    const smx_ugrid g = A.gridpt3[valid ? c : c0];
    const int Pyramid_step = P->level;
    const int iteration = P->iteration;
    const double ncc_alpha = P->ncc_alpha;
    const double ncc_beta = 1.0 - ncc_alpha;
    const bool check_ortho = P->awncc_check_ortho;

    double pre_rho = -1.0, pre_rho_WNCC = -1.0;
    float pre_height = 0.0f;
    int direction = 0;
    bool check_rho = false;
    double max_roh = 0;
    double temp_nccresult = -100.0, temp_nccresult_sec = -100.0;
    nr.result0 = q_result(-1.0);
    nr.result2 = -1000;
    nr.result1 = q_result(-1.0);
    nr.result3 = -1000;
    nr.result4 = 0;
    const bool max_mode = (iteration <= 3 && Pyramid_step == 4) || (iteration <= 2 && Pyramid_step == 3) ||
                          (iteration <= 1 && Pyramid_step == 2);
    const long base = valid ? A.vox_off[c] : 0;
    // an RA pass may leave NumOfHeight above the row InitializeVoxel laid
    // out; never read past the row (the reference would read past its heap
    // row there)
    const long nh = valid ? min((long)nr.NumOfHeight, (long)(A.vox_off[c + 1] - base)) : 0;
    const double gncc = dq(nr.GNCC);
    if (threadIdx.x == 0) s_maxnh = 0;
    s_base[threadIdx.x] = base;
    s_nh[threadIdx.x] = (int)nh;
    __syncthreads();
    atomicMax(&s_maxnh, (int)nh);
    __syncthreads();
    const int maxnh = s_maxnh;
    // this thread's share of a chunk: item i is word q of cell cl; the next
    // chunk's items are loaded into registers while this one is scanned
    uint32_t rw[kChunk];
    float rs[kChunk];
    auto fetch = [&](long h0) {
#pragma unroll
        for (int i = 0; i < kChunk; i++) {
            const int idx = i * 256 + (int)threadIdx.x;
            const int cl = idx / kChunk, q = idx % kChunk;
            const bool in = h0 + q < s_nh[cl];
            const long at = s_base[cl] + h0 + q;
            rw[i] = in ? A.vox[at] : 0u;
            rs[i] = (in && A.sum) ? A.sum[at] : 0.0f;
        }
    };
    if (maxnh > 0) fetch(0);
    for (long h0 = 0; h0 < maxnh; h0 += kChunk) {
        __syncthreads();   // the previous chunk has been scanned
#pragma unroll
        for (int i = 0; i < kChunk; i++) {
            const int idx = i * 256 + (int)threadIdx.x;
            const int cl = idx / kChunk, q = idx % kChunk;
            s_word[cl * kPitch + q] = rw[i];
            s_sum[cl * kPitch + q] = rs[i];
        }
        __syncthreads();
        if (h0 + kChunk < maxnh) fetch(h0 + kChunk);
        const uint32_t *wbuf = &s_word[threadIdx.x * kPitch];
        const float *sbuf = &s_sum[threadIdx.x * kPitch];
        for (int k = 0; k < kChunk; k++) {
            const long hs = h0 + k;
            if (hs >= nh) break;
            const float iter_height = (float)(nr.minHeight + hs * P->height_step);
            if (!(iter_height >= g.minHeight && iter_height <= g.maxHeight)) continue;
            const double incc = dq(SMX_VOXEL_INCC(wbuf[k]));
            double WNCC_sum;
            if (check_ortho && gncc > -1.0) WNCC_sum = incc * ncc_alpha + gncc * ncc_beta;
            else WNCC_sum = incc;
            const double WNCC_temp_rho = WNCC_sum;
            const double temp_rho = A.sum ? (double)sbuf[k] : WNCC_sum;
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
                check_rho = true;
                t_direction = -1;
            } else if (pre_rho != -1) {
                const bool check_peak = direction > 0 && t_direction < 0;
                if (max_mode) {
                    if (temp_nccresult < temp_rho) {
                        if (nr.result4 < 10) nr.result4 += 1;
                        temp_nccresult = temp_rho;
                        nr.result2 = iter_height;
                        temp_nccresult_sec = -1.0;
                        nr.result3 = -9999;
                        max_roh = WNCC_temp_rho;
                    }
                } else if (check_peak) {
                    if (nr.result4 < 10) nr.result4 += 1;
                    if (temp_nccresult < pre_rho) {
                        const double temp_1 = temp_nccresult;
                        temp_nccresult = pre_rho;
                        const float temp_2 = nr.result2;
                        nr.result2 = pre_height;
                        max_roh = pre_rho_WNCC;
                        if (temp_nccresult_sec < temp_1) {
                            temp_nccresult_sec = temp_1;
                            nr.result3 = temp_2;
                        }
                    } else if (temp_nccresult_sec < pre_rho) {
                        temp_nccresult_sec = pre_rho;
                        nr.result3 = pre_height;
                    }
                }
            }
            pre_rho = temp_rho;
            pre_rho_WNCC = WNCC_temp_rho;
            pre_height = iter_height;
            direction = t_direction;
        }
    }
    if (!valid) return;
    if (A.sum && Pyramid_step == 0 && iteration >= 2) {
        nr.result0 = q_result(max_roh);
        nr.result1 = q_result(-1.0);
        nr.result3 = -9999;
    } else {
        nr.result0 = q_result(temp_nccresult);
        nr.result1 = q_result(temp_nccresult_sec);
    }
    A.ncc[c] = nr;
}

}  // namespace smx

// Host launchers ------------------------------------------------------------
namespace smx {

static inline unsigned nblk(long n, int per = 256) { return (unsigned)((n + per - 1) / per); }

hipError_t launch_ortho_coords(const VllParams *dP, const OrthoArgs &O, hipStream_t s)
{
    hipLaunchKernelGGL(k_ortho_coords, dim3(nblk(O.sw * O.sh)), dim3(256), 0, s, dP, O);
    return hipGetLastError();
}

hipError_t launch_gncc(int half, const VllParams *dP, const OrthoArgs &O, const CellArgs &A, double *gval,
                       hipStream_t s)
{
    const dim3 g(nblk(A.ncells)), b(256);
    switch (half) {
    case 5: hipLaunchKernelGGL(k_gncc<5>, g, b, 0, s, dP, O, A, gval); break;
    case 6: hipLaunchKernelGGL(k_gncc<6>, g, b, 0, s, dP, O, A, gval); break;
    case 7: hipLaunchKernelGGL(k_gncc<7>, g, b, 0, s, dP, O, A, gval); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, long n, hipStream_t s);

hipError_t launch_vll_blunder(const VllParams *dP, const BlunderArgs &B, hipStream_t s)
{
    hipError_t e = launch_fill_u64((unsigned long long *)B.cd[0], 0, B.sw * B.sh * 2, s);
    if (e == hipSuccess) e = launch_fill_u64((unsigned long long *)B.cd[1], 0, B.sw * B.sh * 2, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bd_coords, dim3(nblk(B.sw * B.sh)), dim3(256), 0, s, dP, B);
    hipLaunchKernelGGL(k_vll_blunder, dim3(nblk(B.W * B.H)), dim3(256), 0, s, B);
    return hipGetLastError();
}

hipError_t launch_find_peak(const VllParams *dP, const CellArgs &A, hipStream_t s)
{
    hipLaunchKernelGGL(k_find_peak, dim3(nblk(A.ncells)), dim3(256), 0, s, dP, A);
    return hipGetLastError();
}

hipError_t launch_vll_gate(const VllParams *dP, const CellArgs &A, hipStream_t s)
{
    hipLaunchKernelGGL(k_vll_gate, dim3(nblk(A.ncells)), dim3(256), 0, s, dP, A);
    return hipGetLastError();
}

hipError_t launch_vll_fill_runs(const CellArgs &A, hipStream_t s)
{
    hipLaunchKernelGGL(k_vll_fill_runs, dim3(nblk(A.ncells)), dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_vll_fp64(int half, const VllParams *dP, const CellArgs &A, hipStream_t s)
{
    if (A.runs_cap == 0) return hipSuccess;
    // 4 waves of 64 lanes per block, at most 2^20 blocks
    const dim3 grid((unsigned)std::min<long>(nblk(A.runs_cap, 4), 1L << 20)), block(256);
    switch (half) {
    case 5: hipLaunchKernelGGL(k_vll_fp64<5>, grid, block, 0, s, dP, A); break;
    case 6: hipLaunchKernelGGL(k_vll_fp64<6>, grid, block, 0, s, dP, A); break;
    case 7: hipLaunchKernelGGL(k_vll_fp64<7>, grid, block, 0, s, dP, A); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_vll_redo(int half, const VllParams *dP, const CellArgs &A, hipStream_t s)
{
    if (A.runs_cap == 0 || !A.redo) return hipSuccess;
    // up to 32 waves per CU of the FP64 body over the queue; its hypotheses
    // are counted in the slot after the queue length
    CellArgs R = A;
    R.evaluated = A.redo_n + 1;
    const dim3 grid((unsigned)std::min<long>(nblk(A.runs_cap, 4), 2048)), block(256);
    switch (half) {
    case 5: hipLaunchKernelGGL(k_vll_redo<5>, grid, block, 0, s, dP, R); break;
    case 6: hipLaunchKernelGGL(k_vll_redo<6>, grid, block, 0, s, dP, R); break;
    case 7: hipLaunchKernelGGL(k_vll_redo<7>, grid, block, 0, s, dP, R); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_nh_stats(const smx_nccresult *ncc, long n, unsigned long long *stats, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_nh_stats, dim3((unsigned)std::min<long>(nblk(n), 2048)), dim3(256), 0, s, ncc, n, stats);
    return hipGetLastError();
}

static unsigned fill_blocks(long n) { return (unsigned)std::max<long>(1, std::min<long>(nblk(n), 8192)); }

hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_u64, dim3(fill_blocks(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_vll_finalize(const CellArgs &A, hipStream_t s)
{
    hipLaunchKernelGGL(k_vll_finalize, dim3(nblk(A.ncells)), dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_init_voxel(const VllParams *dP, const InitArgs &A, hipStream_t s)
{
    hipLaunchKernelGGL(k_init_voxel, dim3(nblk(A.ncells)), dim3(256), 0, s, dP, A);
    return hipGetLastError();
}

hipError_t launch_relayout(const InitArgs &A, hipStream_t s)
{
    // one wave per cell, grid-stride beyond 2^20 blocks
    hipLaunchKernelGGL(k_relayout, dim3((unsigned)std::min<long>(nblk(A.ncells, 4), 1L << 20)), dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_awncc(const VllParams *dP, const PeakArgs &A, hipStream_t s)
{
    hipLaunchKernelGGL(k_awncc, dim3(nblk(A.ncells)), dim3(256), 0, s, dP, A);
    return hipGetLastError();
}

}  // namespace smx
