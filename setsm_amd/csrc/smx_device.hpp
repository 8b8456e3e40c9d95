// smx_device.hpp — device-side geometry and quantisation shared by the
// VerticalLineLocus kernels. FP64 helpers are written with contraction off so
// the image coordinates are the same bits the CPU path computes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "smx_types.h"

namespace smx {

constexpr double kPi = 3.141592653589793;   // Typedefine.hpp:27

// Pointers read out of argument structs are generic; casting them to the
// global address space lets the compiler emit global_load (32-bit offsets,
// no flat aperture check) instead of flat_load.
template <class T>
using gptr_t = const __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr_t<T> gptr(const T *p)
{
    return (gptr_t<T>)p;
}

struct DevView {
    const uint32_t *texel;   // (intensity | magnitude << 16) per pixel
    const uint8_t *ori;
    uint32_t width, height;
    double start_x, start_y;
};

// Everything a VLL launch reads besides the per-cell arrays.
struct VllParams {
    DevView view[2];
    DevView next[2];
    smx_rpc rpc[2];
    double adjust[2][2];
    double height_step;
    double bin_angle;
    double ncc_alpha;
    double th_height;
    double minmax[2];
    // cos/sin(-rot_theta) for diff_theta = (k - 34) / 2, k = 0..68 (ori_diff
    // and its combined-level average are multiples of 1/2 in [-17, 17]),
    // filled on the host with the libm the CPU path uses.
    double rot_cos[69];
    double rot_sin[69];
    int32_t level;
    int32_t half;            // Half_template_size
    int32_t iteration;
    int32_t combined_incc;   // check_combined_WNCC_INCC
    int32_t is_ra;
    int32_t check_matching_rate;
    int32_t check_matchtag;
    int32_t dem_lt2;         // Pyramid_step == 0 && DEM_resolution < 2
    int32_t awncc_check_ortho;
};

// CoordConversion.hpp:26-87 GetObjectToImageRPC_single
__device__ __forceinline__ void rpc_ground_to_image(const smx_rpc &rpc, const double adj[2], double gx, double gy,
                                                    double gz, double &ox, double &oy)
{
#pragma clang fp contract(off)
    double L = (gx - rpc.off[2]) / rpc.scale[2];
    double P = (gy - rpc.off[3]) / rpc.scale[3];
    double H = (gz - rpc.off[4]) / rpc.scale[4];
    if (L < -10.0 || L > 10.0) {
        gx = gx > 0 ? gx - 360 : gx + 360;
        L = (gx - rpc.off[2]) / rpc.scale[2];
    }
    if (P < -10.0 || P > 10.0) {
        gy = gy > 0 ? gy - 360 : gy + 360;
        P = (gy - rpc.off[3]) / rpc.scale[3];
    }
    // One polynomial at a time keeps only its 20 coefficients live.
    double Coeff[4];
#pragma unroll 1
    for (int j = 0; j < 4; j++) {
        const double *c = rpc.coef[j];
        Coeff[j] = c[0] * 1.0 + c[1] * L + c[2] * P
                 + c[3] * H + c[4] * L * P + c[5] * L * H
                 + c[6] * P * H + c[7] * L * L + c[8] * P * P
                 + c[9] * H * H + c[10] * (P * L) * H + c[11] * (L * L) * L
                 + c[12] * (L * P) * P + c[13] * (L * H) * H + c[14] * (L * L) * P
                 + c[15] * (P * P) * P + c[16] * (P * H) * H + c[17] * (L * L) * H
                 + c[18] * (P * P) * H + c[19] * (H * H) * H;
    }
    double Line = ((Coeff[0] / Coeff[1]) * rpc.scale[0] + rpc.off[0]);
    double Samp = ((Coeff[2] / Coeff[3]) * rpc.scale[1] + rpc.off[1]);
    double y = adj[0] + Line;
    double x = adj[1] + Samp;
    if (y < 0) y = 0;
    if (y > rpc.off[0] + rpc.scale[0] * 1.2) y = rpc.off[0] + rpc.scale[0] * 1.2;
    if (x < 0) x = 0;
    if (x > rpc.off[1] + rpc.scale[1] * 1.2) x = rpc.off[1] + rpc.scale[1] * 1.2;
    ox = x;
    oy = y;
}

// CoordConversion.hpp:96-105 OriginalToPyramid_single
__device__ __forceinline__ void to_pyramid(double x, double y, const DevView &v, int level, double &px, double &py)
{
#pragma clang fp contract(off)
    px = (x / (double)(1 << level)) - v.start_x;
    py = (y / (double)(1 << level)) - v.start_y;
}

// setsm_code.cpp:6373 check_py_image_pt for one image: (int) truncation, then
// the +1 neighbour must be inside.
__device__ __forceinline__ bool py_point_inside(double px, double py, const DevView &v)
{
    const int ix = (int)px, iy = (int)py;
    return iy >= 0 && (long)iy + 1 < (long)v.height && ix >= 0 && (long)ix + 1 < (long)v.width;
}

// setsm_code.cpp:11002-11083 check_image_boundary, SB sensor, both images.
__device__ __forceinline__ bool check_image_boundary(const smx_rpc *rpc, const double (*adj)[2], const DevView *views,
                                                     int level, double lon, double lat, double minH, double maxH,
                                                     int half)
{
#pragma clang fp contract(off)
    int selected = 0;
    bool check_reference = true;
#pragma unroll
    for (int ti = 0; ti < 2; ti++) {
        bool ok0, ok1;
        double x, y, px, py;
        rpc_ground_to_image(rpc[ti], adj[ti], lon, lat, minH, x, y);
        to_pyramid(x, y, views[ti], level, px, py);
        ok0 = px > half + 1 && px < views[ti].width - half - 1 && py > half + 1 && py < views[ti].height - half - 1;
        rpc_ground_to_image(rpc[ti], adj[ti], lon, lat, maxH, x, y);
        to_pyramid(x, y, views[ti], level, px, py);
        ok1 = px > half + 1 && px < views[ti].width - half - 1 && py > half + 1 && py < views[ti].height - half - 1;
        if (ok0 && ok1) selected++;
        else if (ti == 0) check_reference = false;
    }
    return check_reference && selected >= 2;
}

// Template.hpp:964-967: (short)round(v*1000), via a 32-bit integer (the
// low 16 bits are kept for out-of-range values, as GCC on x86-64 does).
__device__ __forceinline__ int16_t q_result(double v) { return (int16_t)(int32_t)round(v * 1000.0); }
// Template.hpp:984-987: (short)(v*1000), truncation
__device__ __forceinline__ int16_t q_voxel(double v) { return (int16_t)(int32_t)(v * 1000.0); }
__device__ __forceinline__ double dq(int16_t v) { return (double)v / 1000.0; }

}  // namespace smx
