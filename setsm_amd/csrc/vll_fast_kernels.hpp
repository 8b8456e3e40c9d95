// vll_fast_kernels.hpp — the precision form of VerticalLineLocus for gfx950.
//
// Same hypotheses, gates and cube semantics as the FP64 kernel (the height,
// RPC pair, pyramid coordinates, image tests and orientation are computed in
// FP64 with identical operand order, so exactly the same (cell, height) pairs
// are evaluated). What changes is the arithmetic of the window loop:
//
//  * a workgroup owns a tile of TX x TY cells; it projects the tile's height
//    bands through both RPCs, takes the bounding box of every texel the
//    tile's windows can touch, and stages it in LDS once as (intensity -
//    pivot, magnitude - pivot): float2 for the left image, float4 with each
//    texel's difference to its right neighbour for the right image. The
//    pivot is one texel value of the box, so the staged values are small,
//    exactly representable numbers (NCC is shift invariant);
//  * one lane evaluates one height hypothesis; taps are LDS reads at
//    compile-time offsets (left window) or at rotated offsets (right window);
//  * bilinear taps and the 5 moments per channel pair are FP32, about
//    per-hypothesis pivots; each window row's partial sums (at most 2H+1
//    samples) are added into FP64 window totals (kWide) for the nested
//    outer / middle / inner windows, and the six correlations are finished
//    in FP64;
//  * a tile whose box does not fit the LDS budget reads texels straight from
//    HBM with the same arithmetic.
//
// The per-hypothesis result is the mean NCC of SETSM within 1e-5 of the FP64
// value; INCC is its (short)(rho*1000) truncation.
#pragma once

#include <type_traits>

#include "fast_window.hpp"
#include "vll_fast_types.hpp"

namespace smx {

// ------------------------------------------------------------- samplers

// 16-byte aligned: the right image's texels are read as float4
extern __shared__ __attribute__((aligned(16))) float2 smx_lds[];
// A float pair / quad in LDS addressed by its 32-bit LDS offset (address space 3).
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(3))) f32x2 lds_f2;
typedef const __attribute__((address_space(3))) f32x4 lds_f4;
__device__ __forceinline__ f32x4 lds_at4(uint32_t byte_offset) { return *(lds_f4 *)(uintptr_t)byte_offset; }

// Left image texels staged in LDS: float2 (intensity, magnitude) per texel,
// rows of the box packed back to back (pitch = box width, odd). Its window
// samples sit at integer offsets: compile-time LDS offsets from one base.
struct LdsImg {
    int base;   // float2 index of texel (0, 0) of the window origin
    int pitch;
    __device__ __forceinline__ float2 at(int row, int col) const { return smx_lds[base + row * pitch + col]; }
};

// Right image texels staged in LDS as float4 (intensity, magnitude, and
// their differences to the next texel of the row): a bilinear tap pair
// along a row is one FMA per channel. pitch and base in float4 units.
struct LdsImgD {
    int base;   // float4 index of texel (0, 0) of the window origin
    int pitch;
    __device__ __forceinline__ f32x4 texel(int row, int col) const
    {
        return *(lds_f4 *)((const float4 *)smx_lds + base + row * pitch + col);
    }
    __device__ __forceinline__ float2 at(int row, int col) const
    {
        const f32x4 v = texel(row, col);
        return make_float2(v.x, v.y);
    }
    // LDS byte address of tap (iy, ix) about the window origin (0 <= iy, ix,
    // small): one 24-bit multiply-add and the scaled base add; the next row
    // is one pitch further
    __device__ __forceinline__ uint32_t addr(int iy, int ix, int &pitch16) const
    {
        pitch16 = pitch * 16;
        // made opaque to the compiler, which would otherwise split the
        // multiply-add below into a multiply and a three-way add
        int col = (ix << 4) + (int)(uint32_t)(uintptr_t)(lds_f4 *)((const float4 *)smx_lds + base);
        asm("" : "+v"(col));
        return (uint32_t)imad24(iy, pitch16, col);
    }
    __device__ __forceinline__ void hlerp(int iy, int ix, float wx, float2 &top, float2 &bot) const
    {
        int p16;
        const uint32_t o = addr(iy, ix, p16);
        const f32x4 a = lds_at4(o), b = lds_at4(o + p16);
        top = make_float2(fmaf(wx, a.z, a.x), fmaf(wx, a.w, a.y));
        bot = make_float2(fmaf(wx, b.z, b.x), fmaf(wx, b.w, b.y));
    }
    // The 2x2 taps (iy, ix) .. (iy + 1, ix + 1): a texel plus its
    // difference is the next texel exactly (both are small integers)
    __device__ __forceinline__ void quad(int iy, int ix, float2 &t00, float2 &t01, float2 &t10, float2 &t11) const
    {
        int p16;
        const uint32_t o = addr(iy, ix, p16);
        const f32x4 a = lds_at4(o), b = lds_at4(o + p16);
        t00 = make_float2(a.x, a.y);
        t01 = make_float2(a.x + a.z, a.y + a.w);
        t10 = make_float2(b.x, b.y);
        t11 = make_float2(b.x + b.z, b.y + b.w);
    }
};

struct GlobalImg {
    gptr_t<uint32_t> texel;
    int width, height;
    int x0, y0;   // image position of at(0, 0)
    float piv_i, piv_m;
    __device__ __forceinline__ float2 at(int row, int col) const
    {
        const int x = min(max(x0 + col, 0), width - 1);
        const int y = min(max(y0 + row, 0), height - 1);
        const uint32_t t = texel[(long)y * width + x];
        return make_float2((float)(t & 0xffffu) - piv_i, (float)(t >> 16) - piv_m);
    }
    __device__ __forceinline__ void quad(int iy, int ix, float2 &t00, float2 &t01, float2 &t10, float2 &t11) const
    {
        t00 = at(iy, ix);
        t01 = at(iy, ix + 1);
        t10 = at(iy + 1, ix);
        t11 = at(iy + 1, ix + 1);
    }
    __device__ __forceinline__ void hlerp(int iy, int ix, float wx, float2 &top, float2 &bot) const
    {
        float2 t00, t01, t10, t11;
        quad(iy, ix, t00, t01, t10, t11);
        top = make_float2(fmaf(wx, t01.x - t00.x, t00.x), fmaf(wx, t01.y - t00.y, t00.y));
        bot = make_float2(fmaf(wx, t11.x - t10.x, t10.x), fmaf(wx, t11.y - t10.y, t10.y));
    }
};

struct TileBox {
    int x0[2], y0[2], w[2], h[2];
    int p[2];   // LDS row pitch: the width rounded up to odd, so that the rows of a
                // box fall on different LDS banks (float2 texels left, float4 right)
    int rbase;  // float4 index of the right box (after the left one, 16-byte aligned)
    float piv_i[2], piv_m[2];
    int fits, nodata;
};

// Window of one hypothesis with taps straight from HBM (tiles whose box
// does not fit LDS, and the level-1 term of the combined INCC). Pivots are
// the centre texels; every sample is tested like SetVecKernelValue.
template <int H, bool kWide>
__device__ __forceinline__ bool global_window(const DevView &V0, const DevView &V1, double rpx, double rpy, double tpx,
                                              double tpy, float cos0, float sin0, double &rho)
{
    constexpr int O = H + 2;
    const double flx = floor(rpx), fly = floor(rpy), ftx = floor(tpx), fty = floor(tpy);
    const int ilx = (int)flx, ily = (int)fly, itx = (int)ftx, ity = (int)fty;
    const int clx = min(max(ilx, 0), (int)V0.width - 1), cly = min(max(ily, 0), (int)V0.height - 1);
    const int ctx = min(max(itx, 0), (int)V1.width - 1), cty = min(max(ity, 0), (int)V1.height - 1);
    const uint32_t pl = V0.texel[(long)cly * V0.width + clx], pr = V1.texel[(long)cty * V1.width + ctx];
    const GlobalImg Lg{gptr(V0.texel), (int)V0.width, (int)V0.height, ilx - O, ily - O, (float)(pl & 0xffffu),
                       (float)(pl >> 16)};
    const GlobalImg Rg{gptr(V1.texel), (int)V1.width, (int)V1.height, itx - O, ity - O, (float)(pr & 0xffffu),
                       (float)(pr >> 16)};
    return fast_window<H, true, kWide>(Lg, Rg, (float)(rpx - flx), (float)(rpy - fly), (float)(tpx - ftx),
                                        (float)(tpy - fty), cos0, sin0, Lg.piv_i, Rg.piv_i, -(float)rpx,
                                        (float)((double)V0.width - 1.0 - rpx), -(float)rpy,
                                        (float)((double)V0.height - 1.0 - rpy), -(float)itx,
                                        (float)((int)V1.width - 1 - itx), -(float)ity,
                                        (float)((int)V1.height - 1 - ity), rho);
}

// Box helpers: the texels a set of hypotheses can touch in both images,
// from their band ends projected through the RPCs.
__device__ __forceinline__ void box_reset(int (&sbox)[2][4])
{
    if (threadIdx.x < 2) {
        sbox[threadIdx.x][0] = 0x7fffffff;
        sbox[threadIdx.x][1] = 0x7fffffff;
        sbox[threadIdx.x][2] = -0x7fffffff;
        sbox[threadIdx.x][3] = -0x7fffffff;
    }
}

__device__ __forceinline__ void box_add(const VllParams &P, const DevView *views, int level, int (&sbox)[2][4],
                                        double lon, double lat, double hh)
{
    for (int ti = 0; ti < 2; ti++) {
        double x, y, px, py;
        rpc_ground_to_image(P.rpc[ti], P.adjust[ti], lon, lat, hh, x, y);
        to_pyramid(x, y, views[ti], level, px, py);
        const int fx = (int)floor(px), fy = (int)floor(py);
        atomicMin(&sbox[ti][0], fx);
        atomicMin(&sbox[ti][1], fy);
        atomicMax(&sbox[ti][2], fx);
        atomicMax(&sbox[ti][3], fy);
    }
}

// Thread 0: clamp the box to the images, add the window margin, pick the
// staging pivots, decide whether it fits the LDS budget.
template <int H>
__device__ __forceinline__ void box_finalize(const DevView *views, const FastArgs &F, const int (&sbox)[2][4],
                                             TileBox &tb)
{
    constexpr int O = H + 2;
    int fits = F.stage ? 1 : 0;
    for (int ti = 0; ti < 2; ti++) {
        const DevView &v = views[ti];
        // window radius H+1, bilinear +1, 2 px for the band's curvature
        int x0 = sbox[ti][0] - O - 2, y0 = sbox[ti][1] - O - 2;
        int x1 = sbox[ti][2] + O + 3, y1 = sbox[ti][3] + O + 3;
        x0 = max(x0, 0);
        y0 = max(y0, 0);
        x1 = min(x1, (int)v.width - 1);
        y1 = min(y1, (int)v.height - 1);
        tb.x0[ti] = x0;
        tb.y0[ti] = y0;
        tb.w[ti] = max(x1 - x0 + 1, 1);
        tb.h[ti] = max(y1 - y0 + 1, 1);
        tb.p[ti] = tb.w[ti] | 1;
        const long cx = min(max((x0 + x1) / 2, 0), (int)v.width - 1);
        const long cy = min(max((y0 + y1) / 2, 0), (int)v.height - 1);
        const uint32_t t = gptr(v.texel)[cy * v.width + cx];
        tb.piv_i[ti] = (float)(t & 0xffffu);
        tb.piv_m[ti] = (float)(t >> 16);
    }
    // budget in float2 units: the right box holds two per texel
    tb.rbase = (tb.p[0] * tb.h[0] + 1) / 2;
    if (2 * tb.rbase + 2 * tb.p[1] * tb.h[1] > F.lds_texels) fits = 0;
    tb.fits = fits;
    tb.nodata = 0;
}

// All threads: copy the box into LDS as pivot-shifted texels: float2 for
// the left image, float4 with the difference to the next texel of the row
// for the right one (0 in the box's last column, which no tap pair starts
// from: the in-box test keeps ix + 1 inside).
__device__ __forceinline__ void box_stage(const DevView *views, TileBox &tb)
{
    int nod = 0;
    for (int ti = 0; ti < 2; ti++) {
        const DevView &v = views[ti];
        const int w = tb.w[ti], h = tb.h[ti], pitch = tb.p[ti];
        const float pi = tb.piv_i[ti], pm = tb.piv_m[ti];
        for (int k = threadIdx.x; k < w * h; k += blockDim.x) {
            const int rr = k / w, cc = k % w;
            const long src = (long)(tb.y0[ti] + rr) * v.width + tb.x0[ti] + cc;
            const uint32_t t = gptr(v.texel)[src];
            nod |= (t & 0xffffu) == 0;
            const float ti_ = (float)(t & 0xffffu) - pi, tm = (float)(t >> 16) - pm;
            if (ti == 0) {
                smx_lds[rr * pitch + cc] = make_float2(ti_, tm);
            } else {
                const uint32_t n = cc + 1 < w ? gptr(v.texel)[src + 1] : t;
                const float di = (float)(n & 0xffffu) - (float)(t & 0xffffu), dm = (float)(n >> 16) - (float)(t >> 16);
                ((float4 *)smx_lds)[tb.rbase + rr * pitch + cc] = make_float4(ti_, tm, di, dm);
            }
        }
    }
    if (nod) tb.nodata = 1;   // benign race: every writer stores 1
}

// log2 of the run length (16, 32 or 64 heights)
__device__ __forceinline__ int run_shift(int R) { return R == 16 ? 4 : R == 32 ? 5 : 6; }

// Whether x holds on any lane of the calling lane's run (R lanes). Every lane
// of the wave must call it.
__device__ __forceinline__ bool group_any(bool x, int R)
{
    const unsigned long long b = __ballot(x);
    if (R == 64) return b != 0;
    const int lane = threadIdx.x & 63;
    return ((b >> (lane & ~(R - 1))) & ((1ull << R) - 1)) != 0;
}

// The runs ri .. ri + 64/R - 1 of the tile's list (those below r1) on one
// wavefront, R = A.run_len lanes each. Returns the wave's count of
// hypotheses that computed an NCC here (uniform; runs handed to the FP64
// pass are counted there). kCombined: a single-pass launch may
// meet a combined-level iteration (the single-pass persistent kernel never
// does: batches launch such jobs in two passes).
template <int H, bool kWide, int kMode, bool kCombined = true>
__device__ __forceinline__ unsigned vll_fast_run(const VllParams &P, const CellArgs &A, const FastArgs &F,
                                             const TileBox &tb, const bool fits, const bool nodata, const long ri,
                                             const long r1, const int lane)
{
    constexpr int O = H + 2;
    constexpr bool kBelow = kMode == kBelowPass;   // window on the level below

        const int R = A.run_len;
        const int lr = lane & (R - 1);
        const long rl = ri + (lane >> run_shift(R));
        const bool have = rl < r1;
        const int2 run = have ? F.tile_runs[rl] : make_int2(0, 0);
        const long cell = run.x;
        const int hidx = run.y + lr;
        const smx_nccresult nr = A.ncc[cell];
        const smx_ugrid g = A.gridpt3[cell];
        Hyp hy{};
        const bool act = have && project_hypothesis(P, A, cell, hidx, nr, g.minHeight, g.maxHeight, hy);
        double temp_rho = -100.0;
        unsigned evaluated = 0;
        bool do_ncc = false;
        long wi = 0;
        uint32_t word = 0;
        if (act) {
            wi = A.vox_off[cell] + hidx;
            word = A.vox[wi];
            if (!hy.orient_ok) A.vox[wi] = word & ~0xffu;
            else do_ncc = true;
        }
        double part = 0;
        bool queued = false;   // the level pass already handed this lane's run to the FP64 pass
        if constexpr (kBelow) {
            // a run the level pass handed over (it marked every evaluated lane
            // of it) is the FP64 pass's: its lanes add nothing here. The
            // level-below tiles group runs into waves differently, so this is
            // decided per lane, not per wave.
            if (do_ncc) part = A.part[wi];
            queued = do_ncc && part == kRedoRho;
            if (queued) do_ncc = false;
        }
        const double rpx = kBelow ? hy.rnx : hy.rpx, rpy = kBelow ? hy.rny : hy.rpy;
        const double tpx = kBelow ? hy.tnx : hy.tpx, tpy = kBelow ? hy.tny : hy.tpy;
        // integer / fractional split of both centres (FP64 -> FP32 fraction)
        const double flx = floor(rpx), fly = floor(rpy), ftx = floor(tpx), fty = floor(tpy);
        const int ilx = (int)flx, ily = (int)fly, itx = (int)ftx, ity = (int)fty;
        const float cos0 = (float)hy.cos0, sin0 = (float)hy.sin0;
        const DevView &V0 = kBelow ? P.next[0] : P.view[0], &V1 = kBelow ? P.next[1] : P.view[1];
        // interior: every tap of the disc is inside both images
        const bool interior = ilx - O >= 0 && ilx + O + 1 < (int)V0.width && ily - O >= 0 &&
                              ily + O + 1 < (int)V0.height && itx - O >= 0 && itx + O + 1 < (int)V1.width &&
                              ity - O >= 0 && ity + O + 1 < (int)V1.height;
        bool in_box = fits;
        if (fits) {
            in_box = ilx - O >= tb.x0[0] && ilx + O + 1 < tb.x0[0] + tb.w[0] && ily - O >= tb.y0[0] &&
                     ily + O + 1 < tb.y0[0] + tb.h[0] && itx - O >= tb.x0[1] && itx + O + 1 < tb.x0[1] + tb.w[1] &&
                     ity - O >= tb.y0[1] && ity + O + 1 < tb.y0[1] + tb.h[1];
        }
        double rho = 0;
        bool all3 = false;
        const bool use_lds = __all(!do_ncc || in_box);
        const bool fast_ok = !nodata && __all(!do_ncc || interior);
        if (do_ncc) {
            if (use_lds) {
                const float fxl = (float)(rpx - flx), fyl = (float)(rpy - fly);
                const float fxt = (float)(tpx - ftx), fyt = (float)(tpy - fty);
                const LdsImg Lw{(ily - O - tb.y0[0]) * tb.p[0] + (ilx - O - tb.x0[0]), tb.p[0]};
                const LdsImgD Rw{tb.rbase + (ity - O - tb.y0[1]) * tb.p[1] + (itx - O - tb.x0[1]), tb.p[1]};
                if (fast_ok)
                    all3 = fast_window<H, false, kWide>(Lw, Rw, fxl, fyl, fxt, fyt, cos0, sin0, tb.piv_i[0],
                                                         tb.piv_i[1], 0, 0, 0, 0, 0, 0, 0, 0, rho);
                else
                    all3 = fast_window<H, true, kWide>(
                        Lw, Rw, fxl, fyl, fxt, fyt, cos0, sin0, tb.piv_i[0], tb.piv_i[1], -(float)rpx,
                        (float)((double)V0.width - 1.0 - rpx), -(float)rpy,
                        (float)((double)V0.height - 1.0 - rpy), -(float)itx, (float)((int)V1.width - 1 - itx),
                        -(float)ity, (float)((int)V1.height - 1 - ity), rho);
            } else {
                all3 = global_window<H, kWide>(V0, V1, rpx, rpy, tpx, tpy, cos0, sin0, rho);
            }
        }
        if (lane == 0 && !use_lds && __any(do_ncc)) atomicAdd(A.accessible + 3, 1ull);
        double rho_n = 0;
        bool all3n = false;
        if constexpr (kMode == kSingle && kCombined) {
            if (do_ncc && P.combined_incc)
                all3n = global_window<H, kWide>(P.next[0], P.next[1], hy.rnx, hy.rny, hy.tnx, hy.tny, cos0, sin0,
                                                 rho_n);
        }
        // a window this path cannot decide: the FP64 pass redoes the whole
        // run; nothing of it is written here (the cleared flags of
        // orientation-rejected lanes are rewritten identically there)
        const bool redo = (do_ncc && all3 && rho == kRedoRho) || (all3n && rho_n == kRedoRho);
        if (__any(redo)) {
            // every run of the wave with something to evaluate goes to the
            // FP64 pass, each once: the level pass marks every evaluated lane
            // of the runs it queues, and the level-below pass queues no marked
            // run (so the queue never holds more than the run count)
            const bool run_queued = group_any(queued, R), run_has = group_any(do_ncc, R);   // all lanes vote
            if (lr == 0 && have && run_has && !run_queued) A.redo[atomicAdd(A.redo_n, 1ull)] = run;
            if constexpr (kMode == kLevelPass) {
                if (do_ncc) A.part[wi] = kRedoRho;   // the level-below pass skips the run
            }
            return 0;
        }
        if constexpr (kMode == kLevelPass) {
            // the level term, kept for the level-below pass
            if (do_ncc) A.part[wi] = !all3 ? kPartNoWindow : rho == 1e300 ? kPartNoCorr : rho;
            return 0;
        }
        // setsm_code.cpp:6456-6493: average of the level term and, on even
        // iterations above level 0, the level-1 term (in that order).
        double sum = 0, cnt = 0;
        if constexpr (kBelow) {
            // the level term from the level pass, then this pass's
            if (do_ncc && part != kPartNoWindow && part != kPartNoCorr) {
                sum += part;
                cnt += 1;
            }
            rho_n = rho;
            all3n = all3;
            all3 = do_ncc && part != kPartNoWindow;
        } else if (do_ncc && all3 && rho != 1e300) {
            sum += rho;
            cnt += 1;
        }
        if (do_ncc && all3n && rho_n != 1e300) {
            sum += rho_n;
            cnt += 1;
        }
        if (do_ncc) {
            const double rr = cnt > 0 ? sum / cnt : -1.0;
            word = (word & 0xffu) | (all3 ? 1u : 0u);
            word |= (uint32_t)(uint16_t)q_voxel(rr) << 16;
            A.vox[wi] = word;
            if (A.rho) A.rho[wi] = (float)rr;
            temp_rho = wncc_of(A, cell, sum, cnt, rr);
            evaluated = 1;
        }
        return wave_max_commit(temp_rho, evaluated, &A.max_bits[cell], R);
    }

template <int H, bool kWide, int kMode = kSingle, bool kCombined = true>
__device__ __forceinline__ void vll_fast_body(const VllParams *__restrict__ Pp, const CellArgs &A, const FastArgs &F,
                                              const long tile)
{
    __shared__ int sbox[2][4];
    __shared__ TileBox tb;
    const VllParams &P = *Pp;
    // the images this launch's windows read: the level's, or the level below
    const DevView *views = kMode == kBelowPass ? P.next : P.view;
    const int level = kMode == kBelowPass ? P.level - 1 : P.level;
    const long r0 = F.tile_run_off[tile], r1 = F.tile_run_off[tile + 1];
    if (r0 == r1) return;
    const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6, lane = threadIdx.x & 63;
    const int per = kRun >> run_shift(A.run_len);   // runs per wavefront

    // 1. One box for the whole tile: project each active cell's band ends.
    box_reset(sbox);
    __syncthreads();
    const int tcx = (int)(tile % F.tiles_x), tcy = (int)(tile / F.tiles_x);
    for (int t = threadIdx.x; t < F.tx * F.ty; t += blockDim.x) {
        const long col = (long)tcx * F.tx + t % F.tx, row = (long)tcy * F.ty + t / F.tx;
        if (col >= F.grid_w || row >= F.grid_h) continue;
        const long c = row * F.grid_w + col;
        if (A.run_count[c] == 0) continue;
        const smx_ugrid g = A.gridpt3[c];
        const double lon = A.grid_wgs[2 * c], lat = A.grid_wgs[2 * c + 1];
        box_add(P, views, level, sbox, lon, lat, (double)g.minHeight);
        box_add(P, views, level, sbox, lon, lat, (double)g.maxHeight);
    }
    __syncthreads();
    if (threadIdx.x == 0) box_finalize<H>(views, F, sbox, tb);
    __syncthreads();
    if (tb.fits || !F.stage) {
        const bool fits = tb.fits != 0;
        if (threadIdx.x == 0 && fits) atomicAdd(A.accessible + 2, 1ull);
        if (fits) box_stage(views, tb);
        __syncthreads();
        const bool nodata = tb.nodata != 0;
        unsigned long long cnt = 0;
        for (long ri = r0 + (long)wave * per; ri < r1; ri += (long)nwaves * per)
            cnt += vll_fast_run<H, kWide, kMode, kCombined>(P, A, F, tb, fits, nodata, ri, r1, lane);
        wave_count_commit(A.evaluated, cnt);
        return;
    }
    // 2. Tall footprints (wide height bands, strong convergence): stage one
    // box per chunk of runs the waves evaluate together; each run spans at
    // most 64 heights, so a chunk's box stays small.
    const long cw = (long)nwaves * per;
    unsigned long long cnt = 0;
    for (long c0 = r0; c0 < r1; c0 += cw) {
        __syncthreads();
        box_reset(sbox);
        __syncthreads();
        if ((long)threadIdx.x < 2 * cw) {
            const long ri = c0 + threadIdx.x / 2;
            if (ri < r1) {
                const int2 run = F.tile_runs[ri];
                const long cell = run.x;
                const smx_nccresult nr = A.ncc[cell];
                const smx_ugrid g = A.gridpt3[cell];
                const int hl = (threadIdx.x & 1) ? min(run.y + A.run_len - 1, (int)nr.NumOfHeight - 1) : run.y;
                double hh = (double)(float)(nr.minHeight + hl * P.height_step);
                hh = fmin(fmax(hh, (double)g.minHeight), (double)g.maxHeight);
                box_add(P, views, level, sbox, A.grid_wgs[2 * cell], A.grid_wgs[2 * cell + 1], hh);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) box_finalize<H>(views, F, sbox, tb);
        __syncthreads();
        const bool fits = tb.fits != 0;
        if (threadIdx.x == 0 && fits) atomicAdd(A.accessible + 2, 1ull);
        if (fits) box_stage(views, tb);
        __syncthreads();
        const bool nodata = tb.nodata != 0;
        if (c0 + (long)wave * per < r1)
            cnt += vll_fast_run<H, kWide, kMode, kCombined>(P, A, F, tb, fits, nodata, c0 + (long)wave * per, r1,
                                                            lane);
    }
    wave_count_commit(A.evaluated, cnt);
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8), each
// with its own L2. Mapping block b to tile (b % 8) * ceil(n / 8) + b / 8 hands
// every XCD one contiguous run of spatially adjacent tiles, so the texels
// neighbouring tiles share are fetched into one L2 instead of eight.
__device__ __forceinline__ long xcd_tile(long b, long n)
{
    constexpr long kXcd = 8;
    const long per = (n + kXcd - 1) / kXcd;
    const long x = b % kXcd, i = b / kXcd;
    const long full = n - (per - 1) * kXcd;   // XCDs whose range holds `per` tiles
    const long start = x < full ? x * per : full * per + (x - full) * (per - 1);
    return start + i;
}

// The same 64-bit value on every lane, held in scalar registers: v is
// uniform over the wave, but the compiler cannot prove it for a value read
// from LDS.
__device__ __forceinline__ long uniform_long(long v)
{
    const unsigned long long u = (unsigned long long)v;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return (long)(((unsigned long long)hi << 32) | lo);
}

// Persistent form over the tiles of many jobs (tiles of different sizes,
// height ranges and pyramids): each workgroup takes the next tile of its
// XCD's contiguous range from an atomic counter and, when that range is
// drained, steals from the other XCDs' ranges. kMode: kSingle for jobs
// without a combined-level term, or one of the two passes of a combined
// iteration over every such job's tiles.
//
// jobs[j] and params[j] (job j's VllParams) are separate no-alias kernel
// arguments, indexed by a tile number made uniform: the compiler then reads
// the job's argument blocks and its RPC coefficients with scalar loads, as
// the per-tile kernels read theirs from the kernel arguments. (Reached
// through a pointer inside the batch descriptor, every such field was a
// per-lane vector load.)
template <int H, int kMode>
__device__ __forceinline__ void vll_fast_persistent(const BatchArgs &B, const BatchJob *__restrict__ jobs,
                                                    const VllParams *__restrict__ params)
{
    __shared__ long cur;
    constexpr long kXcd = 8;
    const long per = (B.total_tiles + kXcd - 1) / kXcd;
    const long home = blockIdx.x % kXcd;
    for (long q = 0; q < kXcd; q++) {
        const long x = (home + q) % kXcd;
        const long lo = x * per;
        const long hi = min(B.total_tiles, lo + per);
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) cur = lo + (long)atomicAdd(&B.next[x], 1ull);
            __syncthreads();
            const long t = uniform_long(cur);
            if (t >= hi) break;
            int j = 0;
            while (j + 1 < B.njobs && jobs[j + 1].tile_begin <= t) j++;
            vll_fast_body<H, true, kMode, false>(params + j, jobs[j].A, jobs[j].F, t - jobs[j].tile_begin);
        }
    }
}

// The persistent kernel in the same two register budgets as the per-tile
// kernels: the compiler's own choice, and one capped for three waves per
// SIMD. The batch's launch shape picks (its occupancy field).
template <int H, int kMode>
__global__ void __launch_bounds__(256) k_vll_fast_batch(BatchArgs B, const BatchJob *__restrict__ jobs,
                                                        const VllParams *__restrict__ params)
{
    vll_fast_persistent<H, kMode>(B, jobs, params);
}

template <int H, int kMode>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_vll_fast_batch_o3(BatchArgs B, const BatchJob *__restrict__ jobs, const VllParams *__restrict__ params)
{
    vll_fast_persistent<H, kMode>(B, jobs, params);
}


// Two register budgets of the same body: the compiler's own choice, and one
// capped so that four waves fit per SIMD. The autotuner picks per shape.
template <int H, bool kWide>
__global__ void __launch_bounds__(256) k_vll_fast(const VllParams *__restrict__ Pp, CellArgs A, FastArgs F)
{
    vll_fast_body<H, kWide>(Pp, A, F, xcd_tile(blockIdx.x, gridDim.x));
}

// The two passes of a combined iteration (kLevelPass, kBelowPass).
template <int H, bool kWide, int kMode>
__global__ void __launch_bounds__(256) k_vll_fast_pass(const VllParams *__restrict__ Pp, CellArgs A, FastArgs F)
{
    vll_fast_body<H, kWide, kMode>(Pp, A, F, xcd_tile(blockIdx.x, gridDim.x));
}

template <int H, bool kWide>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_vll_fast_o4(const VllParams *__restrict__ Pp, CellArgs A, FastArgs F)
{
    vll_fast_body<H, kWide>(Pp, A, F, xcd_tile(blockIdx.x, gridDim.x));
}

template <int H, bool kWide>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_vll_fast_o3(const VllParams *__restrict__ Pp, CellArgs A, FastArgs F)
{
    vll_fast_body<H, kWide>(Pp, A, F, xcd_tile(blockIdx.x, gridDim.x));
}


// Launch entry points (declared in vll_fast_types.hpp).
template <int H>
hipError_t fast_launch(int occ, bool wide, dim3 grid, size_t lds, hipStream_t s, const VllParams *dP,
                       const CellArgs &A, const FastArgs &F)
{
    if (wide) {
        if (occ == 4) hipLaunchKernelGGL((k_vll_fast_o4<H, true>), grid, dim3(256), lds, s, dP, A, F);
        else if (occ == 3) hipLaunchKernelGGL((k_vll_fast_o3<H, true>), grid, dim3(256), lds, s, dP, A, F);
        else hipLaunchKernelGGL((k_vll_fast<H, true>), grid, dim3(256), lds, s, dP, A, F);
    } else {
        if (occ == 4) hipLaunchKernelGGL((k_vll_fast_o4<H, false>), grid, dim3(256), lds, s, dP, A, F);
        else if (occ == 3) hipLaunchKernelGGL((k_vll_fast_o3<H, false>), grid, dim3(256), lds, s, dP, A, F);
        else hipLaunchKernelGGL((k_vll_fast<H, false>), grid, dim3(256), lds, s, dP, A, F);
    }
    return hipGetLastError();
}

template <int H>
hipError_t fast_launch_passes(bool wide, dim3 g1, dim3 g2, size_t lds, hipStream_t s, const VllParams *dP,
                              const CellArgs &A, const FastArgs &F1, const FastArgs &F2)
{
    if (wide) {
        hipLaunchKernelGGL((k_vll_fast_pass<H, true, kLevelPass>), g1, dim3(256), lds, s, dP, A, F1);
        hipLaunchKernelGGL((k_vll_fast_pass<H, true, kBelowPass>), g2, dim3(256), lds, s, dP, A, F2);
    } else {
        hipLaunchKernelGGL((k_vll_fast_pass<H, false, kLevelPass>), g1, dim3(256), lds, s, dP, A, F1);
        hipLaunchKernelGGL((k_vll_fast_pass<H, false, kBelowPass>), g2, dim3(256), lds, s, dP, A, F2);
    }
    return hipGetLastError();
}

template <int H, int kMode>
static auto batch_kernel(bool capped)
{
    return capped ? k_vll_fast_batch_o3<H, kMode> : k_vll_fast_batch<H, kMode>;
}

template <int H>
hipError_t fast_launch_batch(int mode, bool capped, int grid, size_t lds, hipStream_t s, const BatchArgs &B)
{
    if (mode == kLevelPass)
        hipLaunchKernelGGL((batch_kernel<H, kLevelPass>(capped)), dim3(grid), dim3(256), lds, s, B, B.jobs, B.params);
    else if (mode == kBelowPass)
        hipLaunchKernelGGL((batch_kernel<H, kBelowPass>(capped)), dim3(grid), dim3(256), lds, s, B, B.jobs, B.params);
    else
        hipLaunchKernelGGL((batch_kernel<H, kSingle>(capped)), dim3(grid), dim3(256), lds, s, B, B.jobs, B.params);
    return hipGetLastError();
}

template <int H>
int fast_batch_blocks_per_cu(int mode, bool capped, size_t lds)
{
    int per_cu = 0;
    if (mode == kLevelPass)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<H, kLevelPass>(capped), 256, lds);
    else if (mode == kBelowPass)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<H, kBelowPass>(capped), 256, lds);
    else
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<H, kSingle>(capped), 256, lds);
    return per_cu;
}

}  // namespace smx
