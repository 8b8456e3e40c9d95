// pyramid.hip — the image pyramid VerticalLineLocus samples, built in HBM:
// 9x9 sigma-1.5 Gaussian with 2x decimation (Template.hpp:82-132), Sobel
// magnitude and direction (SubFunctions.cpp:2418-2470), 18-bin dominant
// orientation (SubFunctions.cpp:2473-2571), and the packed texel layout the
// matching kernels read (intensity | magnitude << 16, one 32-bit word per
// pixel, so one load fetches both channels of a bilinear tap).
//
// Built with -ffp-contract=off; the filter weights come from the host's libm,
// so every output is bit-identical to the CPU path. These stencils are
// HBM-bound: one thread per output pixel, row-major, coalesced.

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smx {

struct GaussW { double w[81]; };
struct OriW { double w[13 * 13]; };

__global__ void __launch_bounds__(256) k_gauss_decimate(const uint16_t *__restrict__ in, uint32_t w, uint32_t h,
                                                        uint16_t *__restrict__ out, GaussW G)
{
    const long ow = w / 2, oh = h / 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ow * oh) return;
    const long r = i / ow, c = i % ow;
    double temp_v = 0;
    int count = 0;
    for (int l = -4; l <= 4; l++) {
        for (int k = -4; k <= 4; k++) {
            if ((2 * r + l) >= 0 && (2 * c + k) >= 0 && (2 * r + l) < (long)h && (2 * c + k) < (long)w) {
                temp_v += G.w[(l + 4) * 9 + (k + 4)] * in[(2 * r + l) * (long)w + (2 * c + k)];
                count++;
            }
        }
    }
    out[i] = count == 81 ? (uint16_t)temp_v : in[(2 * r) * (long)w + (2 * c)];
}

__global__ void __launch_bounds__(256) k_sobel(const uint16_t *__restrict__ src, uint32_t w, uint32_t h,
                                               uint16_t *__restrict__ mag, int16_t *__restrict__ dir)
{
    const long W = w;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= W * (long)h) return;
    const long i = idx / W, j = idx % W;
    if (j == 0 || j == W - 1 || i == 0 || i == (long)h - 1) {
        mag[idx] = 0;
        dir[idx] = 0;
        return;
    }
    const int a00 = src[(i - 1) * W + (j - 1)], a01 = src[(i - 1) * W + j], a02 = src[(i - 1) * W + (j + 1)];
    const int a10 = src[i * W + (j - 1)], a12 = src[i * W + (j + 1)];
    const int a20 = src[(i + 1) * W + (j - 1)], a21 = src[(i + 1) * W + j], a22 = src[(i + 1) * W + (j + 1)];
    const double gx = (double)((a02 - a00) + 2 * (a12 - a10) + (a22 - a20));
    const double gy = (double)((a00 + 2 * a01 + a02) - (a20 + 2 * a21 + a22));
    const double t = sqrt(gx * gx + gy * gy);
    dir[idx] = (int16_t)(floor(atan2(gy, gx) * 180 / 3.141592653589793));
    mag[idx] = (uint16_t)(t + 0.5);
}

// One thread per pixel; the 18 + 4 bins live in LDS (a column per thread) so
// the runtime bin index costs an LDS access instead of private memory.
__global__ void __launch_bounds__(256) k_orientation(const uint16_t *__restrict__ Gmag,
                                                     const int16_t *__restrict__ Gdir, uint32_t w, uint32_t h,
                                                     uint8_t *__restrict__ ori, OriW GW)
{
    __shared__ double bins[22][256];
    const long main_col = w, main_row = h;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 22; k++) bins[k][t] = 0.0;
    if (idx >= main_col * main_row) return;
    const long mask_row = idx / main_col, mask_col = idx % main_col;
    const int half = 7;
    const int numcols[7] = {6, 5, 5, 5, 4, 3, 0};   // SubFunctions.cpp:8
    const double SUB_RATIO = 4.6;                   // SubFunctions.hpp:33
    const long rlo = -mask_row > -half + 1 ? -mask_row : -half + 1;
    const long rhi = main_row - mask_row - 1 < half - 1 ? main_row - mask_row - 1 : half - 1;
    for (long row = rlo; row <= rhi; row++) {
        const int dcol = numcols[row < 0 ? -row : row];
        const long clo = -mask_row > -dcol ? -mask_row : -dcol;   // as written in the reference
        const long chi = dcol < main_col - mask_col - 1 ? dcol : main_col - mask_col - 1;
        for (long col = clo; col <= chi; col++) {
            const double gu_weight = GW.w[(half - 1 + row) * 13 + (half - 1 + col)];
            const long p = (mask_row + row) * main_col + (mask_col + col);
            const double mag = p >= 0 ? (double)Gmag[p] : 0.0;
            double theta = p >= 0 ? (double)Gdir[p] : 0.0;
            if (theta < 0) theta += 360;
            else if (theta > 360) theta -= 360;
            const long index = (long)(theta / 20.0);
            if (index >= 0 && index < 18) {
                bins[index][t] += mag * gu_weight;
                bins[18 + (int)(index / SUB_RATIO)][t] += mag * gu_weight;
            }
        }
    }
    double max_th = -100;
    int sub_th = 0, th = 0;
    for (int i = 0; i < 4; i++)
        if (bins[18 + i][t] > max_th) { max_th = bins[18 + i][t]; sub_th = i; }
    max_th = -100;
    for (int i = (int)(sub_th * SUB_RATIO); i < (int)(sub_th * SUB_RATIO + SUB_RATIO); i++)
        if (bins[i][t] > max_th) { max_th = bins[i][t]; th = i; }
    ori[idx] = (uint8_t)th;
}

__global__ void __launch_bounds__(256) k_pack(const uint16_t *__restrict__ img, const uint16_t *__restrict__ mag,
                                              uint32_t *__restrict__ texel, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) texel[i] = (uint32_t)img[i] | ((uint32_t)mag[i] << 16);
}

__global__ void __launch_bounds__(256) k_unpack(const uint32_t *__restrict__ texel, uint16_t *__restrict__ img,
                                                uint16_t *__restrict__ mag, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t t = texel[i];
        if (img) img[i] = (uint16_t)(t & 0xffffu);
        if (mag) mag[i] = (uint16_t)(t >> 16);
    }
}

// atan2 -> floor-degree table for the direction test (integer gradients).
__global__ void __launch_bounds__(256) k_dir_probe(const int *__restrict__ gx, const int *__restrict__ gy,
                                                   int16_t *__restrict__ out, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int16_t)(floor(atan2((double)gy[i], (double)gx[i]) * 180 / 3.141592653589793));
}

}  // namespace smx

// Host launchers ------------------------------------------------------------
#include <math.h>

namespace smx {

static inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_gauss_decimate(const uint16_t *in, uint32_t w, uint32_t h, uint16_t *out, hipStream_t s)
{
    // Template.hpp:45-78 CreateGaussianFilter(9, 1.5)
    GaussW G;
    const double sigma = 1.5;
    double sum = 0;
    const double tmp = -1 / (2 * sigma * sigma);
    const double scale = 1 / (sqrt(2 * 3.141592653589793) * sigma);
    for (int i = -4; i <= 4; i++)
        for (int j = -4; j <= 4; j++) {
            const double temp = (i * i + j * j) * tmp;
            const double val = exp(temp) * scale;
            G.w[(i + 4) * 9 + (j + 4)] = val;
            sum += val;
        }
    for (int k = 0; k < 81; k++) G.w[k] /= sum;
    const long n = (long)(w / 2) * (long)(h / 2);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gauss_decimate, dim3(nblk(n)), dim3(256), 0, s, in, w, h, out, G);
    return hipGetLastError();
}

hipError_t launch_sobel(const uint16_t *in, uint32_t w, uint32_t h, uint16_t *mag, int16_t *dir, hipStream_t s)
{
    const long n = (long)w * h;
    hipLaunchKernelGGL(k_sobel, dim3(nblk(n)), dim3(256), 0, s, in, w, h, mag, dir);
    return hipGetLastError();
}

hipError_t launch_orientation(const uint16_t *mag, const int16_t *dir, uint32_t w, uint32_t h, uint8_t *ori,
                              hipStream_t s)
{
    OriW GW;
    const int half = 7;
    for (int row = -half + 1; row <= half - 1; row++)
        for (int col = -half + 1; col <= half - 1; col++)
            GW.w[(half - 1 + row) * 13 + (half - 1 + col)] = exp(-(double)(row * row + col * col) / 4.5);
    const long n = (long)w * h;
    hipLaunchKernelGGL(k_orientation, dim3(nblk(n)), dim3(256), 0, s, mag, dir, w, h, ori, GW);
    return hipGetLastError();
}

hipError_t launch_pack(const uint16_t *img, const uint16_t *mag, uint32_t *texel, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack, dim3(nblk(n)), dim3(256), 0, s, img, mag, texel, n);
    return hipGetLastError();
}

hipError_t launch_unpack(const uint32_t *texel, uint16_t *img, uint16_t *mag, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_unpack, dim3(nblk(n)), dim3(256), 0, s, texel, img, mag, n);
    return hipGetLastError();
}

hipError_t launch_dir_probe(const int *gx, const int *gy, int16_t *out, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_dir_probe, dim3(nblk(n)), dim3(256), 0, s, gx, gy, out, n);
    return hipGetLastError();
}

}  // namespace smx
