// mps.hip — the data-parallel steps of SETSM's per-iteration host loop that
// sit either side of the TIN: SelectMPs (setsm_code.cpp:8535-8889), which
// turns AWNCC's peaks into matched points and updates the search-space grid,
// SetHeightRange (setsm_code.cpp:9851-10261), which rasterises the TIN's
// triangles into the next iteration's per-cell height ranges, and
// DecisionMPs' SetHeightRange_blunder (setsm_code.cpp:10393-10491) and
// blunder_detection_TIN (setsm_code.cpp:9158-9782). The TIN itself stays with
// the caller's host code.
//
// Built with -ffp-contract=off: every expression keeps the reference's
// operand order and its double rounding, so the grid, the point list and the
// height ranges are bit-identical to the CPU path.
//
// SelectMPs: a histogram of first-peak rho (LDS per block, one atomic per bin
// per block), the threshold scan on one lane, one lane per cell for the
// decision and the grid update, and an exclusive scan of the selection flags
// so the points come out in cell order, as the reference appends them.
//
// SetHeightRange: the reference walks the triangles in list order and lets
// the first triangle that contains a cell claim it. Here the claim is an
// atomicMin of the triangle index per cell (16 lanes per triangle over its
// bounding box), which picks the same triangle; a per-cell pass then applies
// the owner's update, and a last per-cell pass fills unclaimed cells from a
// snapshot of their neighbours and writes the returned grid.
//
// SetHeightRange_blunder, inside DecisionMPs' blunder loop, gives each cell
// the highest TIN height of the triangles containing it (a CAS float max);
// blunder_detection_TIN tests each point against its triangles (below).

#include <algorithm>
#include <cstring>

#include "smx_internal.hpp"

namespace smx {

// ------------------------------------------------------------ SelectMPs

__global__ void __launch_bounds__(256) k_mps_hist(MpsDev D)
{
    __shared__ unsigned int h[kRohBins + 1];
    for (int i = threadIdx.x; i <= kRohBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < D.n; c += (long)gridDim.x * blockDim.x) {
        const smx_nccresult r = D.ncc[c];
        if (r.NumOfHeight <= 2) continue;
        int b = (int)ceil(dq(r.result0) * kRohBins);
        if (b > kRohBins - 1) b = kRohBins - 1;
        if (b >= 0) atomicAdd(&h[b], 1u);
        atomicAdd(&h[kRohBins], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= kRohBins; i += blockDim.x)
        if (h[i]) atomicAdd(&D.hist[i], (unsigned long long)h[i]);
}

// setsm_code.cpp:8546-8614: the level's minimum peak rho, the highest bin at
// which the cells in it and above hold more than the level's fraction.
__global__ void k_mps_threshold(MpsDev D)
{
    if (threadIdx.x != 0) return;
    const int level = D.level;
    double th = 0.2;
    const unsigned long long total = D.hist[kRohBins];
    if (!D.is_ra && level > 0 && total > 0) {
        const int it = D.a.iteration;
        double frac = 0;
        bool defined = true;
        if (level == 4) frac = 0.05 + (it - 1) * 0.01;
        else if (level == 3) frac = 0.15 + (it - 1) * 0.01;
        else if (level == 2) frac = 0.60;
        else if (level == 1) frac = 0.80;
        else defined = false;
        if (level <= D.sgm_py) {
            frac = 0.95 - level * 0.1;
            if (frac < 0.5) frac = 0.5;
            defined = true;
        }
        if (!defined) {
            *D.th = -1.0;
            return;
        }
        int acc = 0;   // the reference sums the bins in an int
        for (int b = kRohBins - 1; b > 0; b--) {
            acc += (int)D.hist[b];
            if (acc / (double)(long)total > frac) {
                th = (double)b / (double)kRohBins;
                break;
            }
        }
        if (th > 0.95) th = 0.95;
    }
    *D.th = th;
}

// setsm_code.cpp:8646-8884, one lane per cell.
__global__ void __launch_bounds__(256) k_mps_select(MpsDev D)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.n) return;
    const double th = *D.th;
    if (th < 0) return;
    const smx_nccresult r = D.ncc[c];
    if (r.NumOfHeight <= 2) {
        D.sel[c] = 0;
        return;
    }
    const int level = D.level, it = D.a.iteration;
    smx_ugrid G = D.g[c];
    if (level == 4 && it == 1) G.Height = -1000.0f;
    const double r0 = dq(r.result0), r1 = dq(r.result1);
    double ror;
    if ((it <= 2 && level >= 3) || (it <= 1 && level == 2)) ror = 1.0;
    else ror = fabs(r0) > 0 ? (r0 - r1) / r0 : 0;
    const double roh_th = 0.05;
    const double ror_min = level == 0 ? 0.1 - 0.03 * (3 - level) : 0.1;
    const bool peak_ok = ror >= ror_min && r0 > th;
    const bool above_prev = r0 > dq(G.roh) - roh_th;
    const bool has_peak = r.result4 > 0;
    bool take;
    if (level >= 2 || (level >= 1 && !(D.a.mpp_stereo_angle > 5)) || (level == 0 && D.a.final_level_iteration < 3))
        take = has_peak && peak_ok && above_prev;
    else if (level >= 1)
        take = (above_prev && peak_ok) || (above_prev && r0 > th);
    else
        take = peak_ok && above_prev;
    // a matched cell whose second peak is strong keeps both peaks in range
    if (G.Matched_flag != 0) {
        const double ror_lim = level <= 0 ? 0.1 - 0.03 * (3 - level) : 0.1;
        if (ror < ror_lim && r1 > dq(G.roh) - roh_th) {
            const bool first_higher = r.result2 > r.result3;
            const float lo = first_higher ? r.result3 : r.result2, hi = first_higher ? r.result2 : r.result3;
            G.minHeight = (int16_t)(int)floor(lo - 0.5);
            G.maxHeight = (int16_t)(int)ceil(hi + 0.5);
            G.Matched_flag = 4;
        }
    }
    const int iter_end = (level <= 3 || D.is_ra) ? (D.a.th_roh - 0.10 < D.a.th_roh_min)
                                                   : (D.a.th_roh - 0.06 < D.a.th_roh_min);
    const double z = (double)r.result2;
    bool emit = false;
    bool keep_peak = false;
    if (take) {
        if (level < 2) {
            const double pre_h = (double)G.Height;
            const double pre_range = (double)(G.maxHeight - G.minHeight);
            const int m = G.Matched_flag;
            const double tol2 = D.a.mpp * 2 * (1 << level), tol3 = D.a.mpp * 3 * (1 << level);
            const bool near_prev = fabs(pre_h - z) <= tol2;
            const bool in_range = fabs(pre_h - z) <= pre_range / 2.0;
            const bool near_min = fabs(z - G.minHeight) <= tol3;
            const bool near_max = fabs(z - G.maxHeight) <= tol3;
            const bool keep = (m == 0) || (m == 2 && near_prev) ||
                              ((m == 1 || m == 4) && (near_prev || in_range || near_min || near_max));
            keep_peak = emit = keep && z > -100 && z < 10000;
        } else {
            emit = z > -100 && z < 10000;
            keep_peak = true;
        }
    }
    if (keep_peak) {
        G.roh = r.result0;
        if (G.roh < q_result(th)) G.roh = q_result(th);
    } else {
        G.roh = q_result(D.a.th_roh);
        if (iter_end && D.a.th_roh_start == dq(G.roh)) G.roh = q_result(D.a.th_roh_next);
    }
    D.g[c] = G;
    D.sel[c] = emit ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_mps_write(MpsDev D)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.n || !D.sel[c]) return;
    smx_point3 p;
    p.x = D.grid_pts[2 * c];
    p.y = D.grid_pts[2 * c + 1];
    p.z = (double)D.ncc[c].result2;
    p.flag = 0;
    for (int k = 0; k < 7; k++) p.pad_[k] = 0;
    D.pts[D.off[c]] = p;
}

hipError_t launch_select_mps(const MpsDev &D, const VllScratch &S, hipStream_t s)
{
    hipError_t e = launch_fill_u64(D.hist, 0, kRohBins + 1, s);
    if (e != hipSuccess) return e;
    const unsigned blocks = (unsigned)std::min<long>((D.n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_mps_hist, dim3(std::max(blocks, 1u)), dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_mps_threshold, dim3(1), dim3(64), 0, s, D);
    const unsigned cb = (unsigned)((D.n + 255) / 256);
    if ((e = launch_fill_u64((unsigned long long *)D.sel + D.n, 0, 1, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_mps_select, dim3(cb), dim3(256), 0, s, D);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return job_scan(S, D.sel, (int64_t *)D.off, D.n + 1, s);
}

hipError_t launch_write_mps(const MpsDev &D, hipStream_t s)
{
    hipLaunchKernelGGL(k_mps_write, dim3((unsigned)((D.n + 255) / 256)), dim3(256), 0, s, D);
    return hipGetLastError();
}

// ------------------------------------------------------------ SetHeightRange

// doubles as unsigned keys whose integer order is the numeric order
__device__ __forceinline__ unsigned long long okey(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double okey_value(unsigned long long k)
{
    return __longlong_as_double((long long)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
}

__device__ __forceinline__ double smin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return a < b ? b : a; }

struct TriGeom {
    const smx_point3 *p1, *p2, *p3;
    double zmin, zmax;
    int c0, r0, c1, r1;
};

// SubFunctions.cpp:2915-2944 SetTinBoundary (without the running totals)
__device__ __forceinline__ bool tri_geom(const HrDev &D, long t, TriGeom &T)
{
    const smx_tri tr = D.tris[t];
    if (tr.a >= (uint32_t)D.npts || tr.b >= (uint32_t)D.npts || tr.c >= (uint32_t)D.npts) return false;
    T.p1 = &D.pts[tr.a];
    T.p2 = &D.pts[tr.b];
    T.p3 = &D.pts[tr.c];
    const smx_point3 &a = *T.p1, &b = *T.p2, &c = *T.p3;
    T.zmin = smin(smin(a.z, b.z), c.z);
    T.zmax = smax(smax(a.z, b.z), c.z);
    const double x0 = smin(smin(a.x, b.x), c.x), y0 = smin(smin(a.y, b.y), c.y);
    const double x1 = smax(smax(a.x, b.x), c.x), y1 = smax(smax(a.y, b.y), c.y);
    T.c0 = (int)((x0 - D.b0) / D.gres + 0.5) - 1;
    T.r0 = (int)((y0 - D.b1) / D.gres + 0.5) - 1;
    T.c1 = (int)((x1 - D.b0) / D.gres + 0.5) + 1;
    T.r1 = (int)((y1 - D.b1) / D.gres + 0.5) + 1;
    if (T.c1 >= (int)D.W) T.c1 = (int)D.W - 1;
    if (T.r1 >= (int)D.H) T.r1 = (int)D.H - 1;
    if (T.c0 < 0) T.c0 = 0;
    if (T.r0 < 0) T.r0 = 0;
    return true;
}

// SubFunctions.cpp:2946-2995 IsTinInside
__device__ __forceinline__ bool tin_inside(double px, double py, const smx_point3 &p1, const smx_point3 &p2,
                                           const smx_point3 &p3, float &z)
{
    int sum = 3;
    if ((p2.x - p1.x) * (py - p1.y) - (p2.y - p1.y) * (px - p1.x) <= 0) sum--;
    if ((p3.x - p2.x) * (py - p2.y) - (p3.y - p2.y) * (px - p2.x) <= 0) sum--;
    if ((p1.x - p3.x) * (py - p3.y) - (p1.y - p3.y) * (px - p3.x) <= 0) sum--;
    if (sum != 0 && sum != 3) return false;
    const double ux = p2.x - p1.x, uy = p2.y - p1.y, uz = p2.z - p1.z;
    const double vx = p3.x - p1.x, vy = p3.y - p1.y, vz = p3.z - p1.z;
    double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    const double len = sqrt(nx * nx + ny * ny + nz * nz);
    if (!(len > 0)) return false;
    nx /= len;
    ny /= len;
    nz /= len;
    const double d = -(nx * p1.x + ny * p1.y + nz * p1.z);
    if (nz == 0) return false;
    z = (float)(-1.0 * ((nx * px) + (ny * py) + d) / nz);
    return true;
}

// SubFunctions.cpp:2997-3024 SetNormalAngle + SetAngle; RadToDeg expands to
// 180/PI after the product.
__device__ __forceinline__ double tin_slope_deg(const smx_point3 &p0, const smx_point3 &p1, const smx_point3 &p2)
{
    const double ux = p1.x - p0.x, uy = p1.y - p0.y, uz = p1.z - p0.z;
    const double vx = p2.x - p0.x, vy = p2.y - p0.y, vz = p2.z - p0.z;
    const double nx = uy * vz - vy * uz, ny = -(ux * vz - vx * uz), nz = ux * vy - vx * uy;
    const double norm = sqrt(nx * nx + ny * ny + nz * nz);
    double angle = 0;
    if (norm != 0) angle = acos(fabs(nz) / norm) * 180 / 3.141592653589793;
    if (angle <= 0 && angle >= -90) angle = fabs(angle);
    else if (angle <= -270 && angle >= -360) angle = 360 + angle;
    else if (angle >= 270 && angle <= 360) angle = 360 - angle;
    return angle;
}

// Block-wide min / max into one global atomic each: a shuffle reduction per
// wavefront, then one LDS atomic per wavefront. Every thread of the block
// must call these (neutral values for threads without one).
__device__ __forceinline__ void block_minmax_u64(unsigned long long lo, unsigned long long hi,
                                                 unsigned long long *gmin, unsigned long long *gmax)
{
    __shared__ unsigned long long s_lo, s_hi;
    if (threadIdx.x == 0) {
        s_lo = ~0ull;
        s_hi = 0ull;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&s_lo, lo);
        atomicMax(&s_hi, hi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMin(gmin, s_lo);
        atomicMax(gmax, s_hi);
    }
}

__device__ __forceinline__ void block_minmax_i32(int lo, int hi, int *gmin, int *gmax)
{
    __shared__ int s_lo, s_hi;
    if (threadIdx.x == 0) {
        s_lo = 0x7fffffff;
        s_hi = -0x7fffffff - 1;
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&s_lo, lo);
        atomicMax(&s_hi, hi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMin(gmin, s_lo);
        atomicMax(gmax, s_hi);
    }
}

// 16 lanes per triangle walk its bounding box; a cell inside the triangle
// takes the smallest claiming index. Groups stride over the triangles (the
// grid is bounded), so the valid triangles' height span reaches zkeys in one
// atomic per block.
__global__ void __launch_bounds__(256) k_hr_claim(HrDev D)
{
    const long groups = ((long)gridDim.x * blockDim.x) >> 4;
    const int sub = threadIdx.x & 15;
    unsigned long long kmin = ~0ull, kmax = 0ull;
    for (long t = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 4; t < D.ntris; t += groups) {
        TriGeom T;
        if (!tri_geom(D, t, T)) continue;
        const unsigned long long a = okey(T.zmin), b = okey(T.zmax);
        kmin = a < kmin ? a : kmin;
        kmax = b > kmax ? b : kmax;
        const long bw = (long)(T.c1 - T.c0 + 1), bh = (long)(T.r1 - T.r0 + 1);
        if (bw <= 0 || bh <= 0) continue;
        for (long k = sub; k < bw * bh; k += 16) {
            const long row = T.r0 + k / bw, col = T.c0 + k % bw;
            const double px = col * D.gres + D.b0, py = row * D.gres + D.b1;
            float z;
            if (tin_inside(px, py, *T.p1, *T.p2, *T.p3, z)) atomicMin(&D.owner[D.W * row + col], (uint32_t)t);
        }
    }
    block_minmax_u64(kmin, kmax, &D.zkeys[0], &D.zkeys[1]);
}

// setsm_code.cpp:9959-10087: the owner triangle's update of a claimed cell.
__global__ void __launch_bounds__(256) k_hr_apply(HrDev D)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.W * D.H) return;
    const uint32_t t = D.owner[c];
    smx_ugrid G = D.g[c];
    if (t == 0xffffffffu) {
        D.mb[c] = 0;
        G.Height = -1000.0f;
        D.g[c] = G;
        return;
    }
    TriGeom T;
    tri_geom(D, t, T);
    const smx_point3 &p1 = *T.p1, &p2 = *T.p2, &p3 = *T.p3;
    const double slope = tin_slope_deg(p1, p2, p3);
    const double wgt = slope <= 30 ? 1.0 + 0.05 * slope / 5.0 : 1.3;
    const double bf = D.level >= 3 ? ceil(D.buf * wgt) : ceil(D.buf);
    const long row = c / D.W, col = c % D.W;
    const double px = col * D.gres + D.b0, py = row * D.gres + D.b1;
    float z = -1000.0f;
    tin_inside(px, py, p1, p2, p3, z);
    if (D.level < 2) {
        const double d1 = sqrt((px - p1.x) * (px - p1.x) + (py - p1.y) * (py - p1.y));
        const double d2 = sqrt((px - p2.x) * (px - p2.x) + (py - p2.y) * (py - p2.y));
        const double d3 = sqrt((px - p3.x) * (px - p3.x) + (py - p3.y) * (py - p3.y));
        if (d1 == 0) { z = (float)p1.z; G.Matched_flag = 2; }
        else if (d2 == 0) { z = (float)p2.z; G.Matched_flag = 2; }
        else if (d3 == 0) { z = (float)p3.z; G.Matched_flag = 2; }
    }
    G.Height = z;
    const double t1 = smin(T.zmin, (double)z), t2 = smax(T.zmax, (double)z);
    if (G.Matched_flag == 4) {
        if (t1 - bf <= G.minHeight) G.minHeight = (int16_t)(int)floor(t1 - bf);
        if (t2 + bf >= G.maxHeight) G.maxHeight = (int16_t)(int)ceil(t2 + bf);
    } else {
        G.minHeight = (int16_t)(int)floor(t1 - bf);
        G.maxHeight = (int16_t)(int)ceil(t2 + bf);
    }
    if (G.minHeight > G.maxHeight) {
        const int16_t s = G.minHeight;
        G.minHeight = G.maxHeight;
        G.maxHeight = s;
    }
    if (G.Matched_flag == 2) {
        G.minHeight = (int16_t)(int)floor(z - bf);
        G.maxHeight = (int16_t)(int)ceil(z + bf);
    } else {
        G.Matched_flag = 1;
        if (G.Mean_ortho_ncc > q_result(D.ncc_th)) {
            G.minHeight = (int16_t)(int)floor(z - bf);
            G.maxHeight = (int16_t)(int)ceil(z + bf);
        } else if (D.level >= 2) {
            G.minHeight = (int16_t)(int)floor(t1 - bf);
            G.maxHeight = (int16_t)(int)ceil(t2 + bf);
        }
    }
    if (G.maxHeight - G.minHeight > D.th_hg) G.minHeight = G.maxHeight = -100;
    D.mb[c] = 1;
    D.g[c] = G;
}

// setsm_code.cpp:10093-10252: unclaimed cells take their claimed
// neighbours' range (D.g is the snapshot, read only here), then the
// returned grid.
__global__ void __launch_bounds__(256) k_hr_fill(HrDev D)
{
    const long c0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = c0 < D.W * D.H;
    const long c = in ? c0 : 0;
    const long row = c / D.W, col = c % D.W;
    smx_ugrid G = D.g[c];
    int mb = D.mb[c];
    if (G.Matched_flag == 0) {
        int lo = 10000, hi = -10000;
        bool any = false;
        for (long dr = -1; dr <= 1; dr++)
            for (long dc = -1; dc <= 1; dc++) {
                if (col + dc < 0 || col + dc >= D.W || row + dr < 0 || row + dr >= D.H) continue;
                const smx_ugrid q = D.g[(col + dc) + (row + dr) * D.W];
                if (q.Matched_flag == 0) continue;
                any = true;
                if (lo > q.minHeight) lo = q.minHeight;
                if (hi < q.maxHeight) hi = q.maxHeight;
            }
        if (any) {
            G.minHeight = (int16_t)lo;
            G.maxHeight = (int16_t)hi;
            mb = 2;
        } else if (D.level <= 1) {
            const double tmin = okey_value(D.zkeys[0]), tmax = okey_value(D.zkeys[1]);
            const double span = fabs(tmax - tmin) / 3.0;
            double bf = D.buf > span ? span : D.buf;
            if (bf < (1 << D.level) * D.resolution) bf = (1 << D.level) * D.resolution;
            if (G.minHeight < tmin - bf) G.minHeight = (int16_t)(int)floor(tmin - bf);
            if (G.maxHeight > tmax + bf) G.maxHeight = (int16_t)(int)ceil(tmax + bf);
        }
    } else if (mb == 0)
        G.Matched_flag = 0;
    block_minmax_i32(in ? (int)G.minHeight : 0x7fffffff, in ? (int)G.maxHeight : -0x7fffffff - 1, &D.hbounds[0],
                     &D.hbounds[1]);
    if (!in) return;
    if (G.maxHeight - G.minHeight > D.th_hg) G.minHeight = G.maxHeight = -100;
    int16_t lo = G.minHeight, hi = G.maxHeight;
    if (lo > hi) {
        const int16_t s = lo;
        lo = hi;
        hi = s;
    }
    smx_ugrid R;
    R.Height = mb == 0 ? -1000.0f : G.Height;
    R.minHeight = lo < -100 ? (int16_t)-100 : lo;
    R.maxHeight = hi;
    if (R.minHeight == 0) R.minHeight = -1;
    if (R.maxHeight == 0) R.maxHeight = 1;
    R.roh = G.roh;
    R.ortho_ncc[0] = 0;
    R.ortho_ncc[1] = G.ortho_ncc[1];
    R.Mean_ortho_ncc = G.Mean_ortho_ncc;
    R.Matched_flag = (mb == 2 && G.Matched_flag == 0) ? 1 : G.Matched_flag;
    R.anchor_flag = 0;
    D.out[c] = R;
}

__global__ void k_hr_init(HrDev D)
{
    if (threadIdx.x == 0) {
        D.zkeys[0] = okey(100000.0);
        D.zkeys[1] = okey(-100000.0);
        D.hbounds[0] = 100000;
        D.hbounds[1] = -100000;
    }
}

hipError_t launch_set_height_range(const HrDev &D, hipStream_t s)
{
    const long n = D.W * D.H;
    hipError_t e = launch_fill_u32(D.owner, 0xffffffffu, n, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_hr_init, dim3(1), dim3(64), 0, s, D);
    if (D.ntris > 0)
        hipLaunchKernelGGL(k_hr_claim, dim3((unsigned)std::min<long>((D.ntris * 16 + 255) / 256, 8192)), dim3(256),
                           0, s, D);
    const unsigned cb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_hr_apply, dim3(cb), dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_hr_fill, dim3(cb), dim3(256), 0, s, D);
    return hipGetLastError();
}

// ------------------------------------------------------------ SetHeightRange_blunder

// std::atomic<float> max as the reference's update_max does it: replace only
// while the new value is larger (a NaN never is).
__device__ __forceinline__ void float_max(float *a, float v)
{
    unsigned int *ua = (unsigned int *)a;
    unsigned int cur = *ua;
    while (v > __uint_as_float(cur)) {
        const unsigned int seen = atomicCAS(ua, cur, __float_as_uint(v));
        if (seen == cur) break;
        cur = seen;
    }
}

// setsm_code.cpp:10404-10483: 16 lanes per triangle over its bounding box,
// each inside cell keeps the highest TIN height.
__global__ void __launch_bounds__(256) k_hrb_claim(HrDev D, float *heights)
{
    const long t = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int sub = threadIdx.x & 15;
    if (t >= D.ntris) return;
    TriGeom T;
    if (!tri_geom(D, t, T)) return;
    const long bw = (long)(T.c1 - T.c0 + 1), bh = (long)(T.r1 - T.r0 + 1);
    if (bw <= 0 || bh <= 0) return;
    for (long k = sub; k < bw * bh; k += 16) {
        const long row = T.r0 + k / bw, col = T.c0 + k % bw;
        float z = -1000.0f;
        if (tin_inside(col * D.gres + D.b0, row * D.gres + D.b1, *T.p1, *T.p2, *T.p3, z))
            float_max(&heights[D.W * row + col], z);
    }
}

// setsm_code.cpp:10484-10489
__global__ void __launch_bounds__(256) k_hrb_apply(HrDev D, const float *heights)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.W * D.H) return;
    const float h = heights[c];
    if (h != -1000.0) D.g[c].Height = h;
}

hipError_t launch_set_height_range_blunder(const HrDev &D, float *heights, hipStream_t s)
{
    const long n = D.W * D.H;
    const float unset = -1000.0f;
    uint32_t bits;
    memcpy(&bits, &unset, 4);
    hipError_t e = launch_fill_u32((uint32_t *)heights, bits, n, s);
    if (e != hipSuccess) return e;
    if (D.ntris > 0)
        hipLaunchKernelGGL(k_hrb_claim, dim3((unsigned)((D.ntris * 16 + 255) / 256)), dim3(256), 0, s, D, heights);
    hipLaunchKernelGGL(k_hrb_apply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, D, (const float *)heights);
    return hipGetLastError();
}

// ------------------------------------------------------------ blunder_detection_TIN

// setsm_code.cpp:9158-9782. The reference's serial pre-pass over the triangles
// builds each point's list of its first 30 triangles and the statistics of
// the edge height differences; the statistics are summed on the host in
// triangle order (bit-exact), the lists are built here: counts, an exclusive
// scan, then the lists. A list holds a point's triangles in no set order;
// only a point with more than 30 reads them in index order, through a
// selection of its 30 smallest. The per-point pass is the reference's OpenMP
// loop with the race taken out: a point's own test and a neighbour's blame
// each set a bit, and the host turns the bits into flags.

__device__ __forceinline__ long bd_cell(const BdDev &D, double x, double y)
{
    const long col = (long)((x - D.b0) / D.gres + 0.5);
    const long row = (long)((y - D.b1) / D.gres + 0.5);
    return D.W * row + col;
}

__device__ __forceinline__ void bd_mark(const BdDev &D, long i, uint32_t bit)
{
    atomicOr(&D.state[i >> 2], bit << (8 * (i & 3)));
}

// every point's cell inside the grid (the reference indexes the ortho NCC
// with it unchecked)
__global__ void __launch_bounds__(256) k_bd_check(BdDev D)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.npts) return;
    const smx_point3 P = D.pts[i];
    const double fc = (P.x - D.b0) / D.gres + 0.5, fr = (P.y - D.b1) / D.gres + 0.5;
    if (!(fc > -1.0 && fc < (double)D.W && fr > -1.0 && fr < (double)D.H)) atomicOr(D.err, 1);
}

__global__ void __launch_bounds__(256) k_bd_degree(BdDev D)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < D.ntris; t += stride) {
        const smx_tri tr = D.tris[t];
        const uint32_t v[3] = {tr.a, tr.b, tr.c};
        for (int k = 0; k < 3; k++)
            if ((long)v[k] < D.npts) atomicAdd((unsigned long long *)&D.deg[v[k]], 1ull);
    }
}

__global__ void __launch_bounds__(256) k_bd_fill(BdDev D)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < D.ntris; t += stride) {
        const smx_tri tr = D.tris[t];
        const uint32_t v[3] = {tr.a, tr.b, tr.c};
        for (int k = 0; k < 3; k++)
            if ((long)v[k] < D.npts) {
                const unsigned long long left = atomicAdd((unsigned long long *)&D.deg[v[k]], ~0ull);
                D.adj[D.off[v[k]] + (long)left - 1] = (uint32_t)t;
            }
    }
}

// setsm_code.cpp:9441-9756, one lane per point
__global__ void __launch_bounds__(256) k_bd_points(BdDev D)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.npts) return;
    const smx_point3 P = D.pts[i];
    if (P.flag == 1) return;
    const int level = D.level;
    const long ref = bd_cell(D, P.x, P.y);
    const double on = (double)D.oncc[ref];
    bool blunder = false;
    if (!D.is_ra && on < D.th_ref_ncc && level >= 2) blunder = true;
    const unsigned anchor = (level >= 3 && D.iteration >= 4 && on >= 0.95 && D.flag_blunder) ? 1u : 0u;
    atomicMax(&D.akey[ref], ((unsigned long long)i << 1) | anchor);

    int cpos = 0, cneg = 0, count = 0;
    bool neigh = false;
    auto visit = [&](uint32_t t) {
        if (t == 0 || (long)t >= D.ntris) return;
        const smx_tri tr = D.tris[t];
        const uint32_t tt[3] = {tr.a, tr.b, tr.c};
        long refi = 0, t0 = 0, t1 = 0;
        bool ci = false;
        for (int kk = 0; kk < 3; kk++) {
            if ((long)tt[kk] == i) refi = i;
            else if (!ci) {
                t0 = tt[kk];
                ci = true;
            } else
                t1 = tt[kk];
        }
        if (!(refi < D.npts && t0 < D.npts && t1 < D.npts)) return;
        const smx_point3 p0 = D.pts[refi], p1 = D.pts[t0], p2 = D.pts[t1];
        const double angle = tin_slope_deg(p0, p1, p2);
        const double dh1 = p0.z - p1.z, dh2 = p0.z - p2.z, dh3 = p1.z - p2.z;
        if (dh1 >= 0 && dh2 >= 0 && angle > 30) cpos++;
        if (dh1 < 0 && dh2 < 0 && angle > 30) cneg++;
        if (D.check_match) {
            const double ddh_1 = fabs(dh1) - fabs(dh2), ddh_2 = fabs(dh1) - fabs(dh3), ddh_3 = fabs(dh2) - fabs(dh3);
            if (fabs(ddh_1) > D.height_th || fabs(ddh_2) > D.height_th || fabs(ddh_3) > D.height_th) {
                neigh = true;
                // setsm_code.cpp:9577-9620: the same comparisons, so ties
                // resolve to the same vertices (selects, not a local array,
                // keep the vertices in registers)
                const double h0 = p0.z, h1 = p1.z, h2 = p2.z;
                int mx, mn, md;
                if (h0 > h1) {
                    if (h0 > h2) {
                        mx = 0;
                        if (h1 > h2) { mn = 2; md = 1; }
                        else { mn = 1; md = 2; }
                    } else { mx = 2; mn = 1; md = 0; }
                } else {
                    if (h1 > h2) {
                        mx = 1;
                        if (h0 > h2) { mn = 2; md = 0; }
                        else { mn = 0; md = 2; }
                    } else { mx = 2; md = 1; mn = 0; }
                }
                auto pick = [](int k, double a, double b, double c) { return k == 0 ? a : (k == 1 ? b : c); };
                const double hmd = pick(md, h0, h1, h2);
                const double dh = (hmd - pick(mn, h0, h1, h2)) - (pick(mx, h0, h1, h2) - hmd);
                int k = -1;
                if (dh > 0 && dh > D.height_th) k = mn;
                else if (dh < 0 && fabs(dh) > D.height_th) k = mx;
                if (k >= 0) {
                    bool take = D.is_ra == 1;
                    if (!take) {
                        const double ok =
                            (double)D.oncc[bd_cell(D, pick(k, p0.x, p1.x, p2.x), pick(k, p0.y, p1.y, p2.y))];
                        take = D.flag_blunder ? ok < D.ortho_ncc_th : ok < 100.;
                    }
                    if (take) bd_mark(D, k == 0 ? refi : (k == 1 ? t0 : t1), 2u);
                }
            }
        }
        count++;
    };
    const long o = D.off[i];
    const long deg = D.off[i + 1] - o;
    if (deg <= kBdMaxNodes) {
        for (long k = 0; k < deg; k++) visit(D.adj[o + k]);
    } else {
        // the reference keeps the first 30 in triangle order: the 30
        // smallest indices, repeats included
        long last = -1;
        int taken = 0;
        while (taken < kBdMaxNodes) {
            uint32_t best = 0xffffffffu;
            int mult = 0;
            for (long k = 0; k < deg; k++) {
                const uint32_t u = D.adj[o + k];
                if ((long)u <= last) continue;
                if (u < best) {
                    best = u;
                    mult = 1;
                } else if (u == best)
                    mult++;
            }
            if (mult == 0) break;
            for (int r = 0; r < mult && taken < kBdMaxNodes; r++, taken++) visit(best);
            last = best;
        }
    }

    if (D.is_ra == 1) {
        if (neigh && !D.flag_blunder) blunder = true;
        if ((cpos >= (int)(count * 0.7 + 0.5) || cneg >= (int)(count * 0.7 + 0.5)) && count > 1) blunder = true;
    } else {
        if (neigh && !D.flag_blunder && on < 100.) blunder = true;
        double cth;
        if (level >= 2) cth = count * 0.5;
        else if (level == 1) cth = (int)(count * 0.6 + 0.5);
        else cth = (int)(count * 0.7 + 0.5);
        if ((cpos >= cth || cneg >= cth) && count > 1) {
            if (D.flag_blunder) {
                double tmp_th = 0.6 - (4 - level) * 0.1;
                if (level == 1) tmp_th = 0.6;
                if (level >= 1) {
                    if (on < tmp_th) blunder = true;
                } else if (level == 0) {
                    if (D.iteration <= 1) {
                        if (on < 0.9) blunder = true;
                    } else if (D.iteration == 2)
                        blunder = true;
                    else if (on < D.ortho_ncc_th)
                        blunder = true;
                }
            } else if (on < 100.)
                blunder = true;
        }
    }
    if (blunder) bd_mark(D, i, 1u);
}

// the anchor flag of the highest tested point of each cell, as the
// reference's last write leaves it
__global__ void __launch_bounds__(256) k_bd_anchor(BdDev D)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.npts) return;
    const smx_point3 P = D.pts[i];
    if (P.flag == 1) return;
    const long ref = bd_cell(D, P.x, P.y);
    const unsigned long long key = D.akey[ref];
    if ((long)(key >> 1) == i) D.g[ref].anchor_flag = (uint8_t)(key & 1);
}

hipError_t launch_bd_prepare(const BdDev &D, const VllScratch &S, hipStream_t s)
{
    hipError_t e = launch_fill_u64((unsigned long long *)D.deg, 0, D.npts + 1, s);
    if (e == hipSuccess) e = launch_fill_u32((uint32_t *)D.err, 0, 1, s);
    if (e == hipSuccess) e = launch_fill_u32(D.state, 0, (D.npts + 3) / 4, s);
    if (e != hipSuccess) return e;
    if (D.npts > 0) hipLaunchKernelGGL(k_bd_check, dim3((unsigned)((D.npts + 255) / 256)), dim3(256), 0, s, D);
    const unsigned tb = (unsigned)std::min<long>((D.ntris + 255) / 256, 8192);
    if (D.ntris > 0) hipLaunchKernelGGL(k_bd_degree, dim3(tb), dim3(256), 0, s, D);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = job_scan(S, D.deg, (int64_t *)D.off, D.npts + 1, s)) != hipSuccess) return e;
    if (D.ntris > 0) hipLaunchKernelGGL(k_bd_fill, dim3(tb), dim3(256), 0, s, D);
    return hipGetLastError();
}

hipError_t launch_bd_points(const BdDev &D, hipStream_t s)
{
    if (D.npts == 0) return hipSuccess;
    const long n = D.W * D.H;
    hipError_t e = launch_fill_u64(D.akey, 0, n, s);
    if (e != hipSuccess) return e;
    const unsigned pb = (unsigned)((D.npts + 255) / 256);
    hipLaunchKernelGGL(k_bd_points, dim3(pb), dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_bd_anchor, dim3(pb), dim3(256), 0, s, D);
    return hipGetLastError();
}

}  // namespace smx
