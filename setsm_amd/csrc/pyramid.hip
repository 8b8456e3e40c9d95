// pyramid.hip — the image pyramid VerticalLineLocus samples, built in HBM:
// 9x9 sigma-1.5 Gaussian with 2x decimation (Template.hpp:82-132), Sobel
// magnitude and direction (SubFunctions.cpp:2418-2470) with the direction's
// histogram bin, 18-bin dominant orientation (SubFunctions.cpp:2473-2571),
// and the packed texel layout the
// matching kernels read (intensity | magnitude << 16, one 32-bit word per
// pixel, so one load fetches both channels of a bilinear tap).
//
// Built with -ffp-contract=off; the filter weights come from the host's libm,
// so every output is bit-identical to the CPU path. One thread per output
// pixel, row-major, coalesced; the orientation histogram reads its 113 taps
// per pixel from an LDS-staged tile.

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

// Orientation word of one pixel: Sobel magnitude (bits 0-15), its 20-degree
// direction bin (bits 16-20, kNoBin when the direction falls outside the 18
// bins) and the bin's quarter (bits 21-22), computed once per pixel exactly
// as Orientation does per window tap (SubFunctions.cpp:2520-2531).
constexpr uint32_t kNoBin = 31;
constexpr double kSubRatio = 4.6;   // SubFunctions.hpp:33

__device__ __forceinline__ uint32_t orientation_word(uint16_t mag, int16_t dir)
{
    double theta = (double)dir;
    if (theta < 0) theta += 360;
    else if (theta > 360) theta -= 360;
    const long index = (long)(theta / 20.0);
    if (!(index >= 0 && index < 18)) return (uint32_t)mag | (kNoBin << 16);
    const uint32_t sub = (uint32_t)(int)((double)index / kSubRatio);
    return (uint32_t)mag | ((uint32_t)index << 16) | (sub << 21);
}

__global__ void __launch_bounds__(256) k_sobel(const uint16_t *__restrict__ src, uint32_t w, uint32_t h,
                                               uint16_t *__restrict__ mag, uint32_t *__restrict__ ow)
{
    const long W = w;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= W * (long)h) return;
    const long i = idx / W, j = idx % W;
    if (j == 0 || j == W - 1 || i == 0 || i == (long)h - 1) {
        mag[idx] = 0;
        ow[idx] = orientation_word(0, 0);
        return;
    }
    const int a00 = src[(i - 1) * W + (j - 1)], a01 = src[(i - 1) * W + j], a02 = src[(i - 1) * W + (j + 1)];
    const int a10 = src[i * W + (j - 1)], a12 = src[i * W + (j + 1)];
    const int a20 = src[(i + 1) * W + (j - 1)], a21 = src[(i + 1) * W + j], a22 = src[(i + 1) * W + (j + 1)];
    const double gx = (double)((a02 - a00) + 2 * (a12 - a10) + (a22 - a20));
    const double gy = (double)((a00 + 2 * a01 + a02) - (a20 + 2 * a21 + a22));
    const double t = sqrt(gx * gx + gy * gy);
    const int16_t dir = (int16_t)(floor(atan2(gy, gx) * 180 / 3.141592653589793));
    const uint16_t m = (uint16_t)(t + 0.5);
    mag[idx] = m;
    ow[idx] = orientation_word(m, dir);
}

// One thread per pixel of a 16x16 tile. The tile's orientation words and a
// 6-pixel halo are staged in LDS; the 18 + 4 FP64 bins live in LDS (a column
// per thread) so the runtime bin index costs an LDS access instead of private
// memory. Every bin adds mag * weight in the reference's row-major tap order,
// so the sums, and the argmax, are the CPU path's bit for bit.
constexpr int kOriTile = 16, kOriHalo = 6, kOriBox = kOriTile + 2 * kOriHalo;
constexpr int kNumCols[7] = {6, 5, 5, 5, 4, 3, 0};   // SubFunctions.cpp:8

// One tap of the histogram: bin and quarter-bin sums of thread t.
__device__ __forceinline__ void ori_tap(double (*bins)[256], int t, uint32_t word, double weight)
{
    const uint32_t index = (word >> 16) & 31u;
    if (index == kNoBin) return;
    const double v = (double)(word & 0xffffu) * weight;
    bins[index][t] += v;
    bins[18 + (word >> 21)][t] += v;
}

__global__ void __launch_bounds__(256) k_orientation(const uint32_t *__restrict__ ow, uint32_t w, uint32_t h,
                                                     uint8_t *__restrict__ ori, OriW GW)
{
    __shared__ double bins[22][256];
    __shared__ uint32_t box[kOriBox][kOriBox];
    const long main_col = w, main_row = h;
    const int t = threadIdx.x;
    const long x0 = (long)blockIdx.x * kOriTile, y0 = (long)blockIdx.y * kOriTile;
    for (int k = t; k < kOriBox * kOriBox; k += 256) {
        const long gy = y0 - kOriHalo + k / kOriBox, gx = x0 - kOriHalo + k % kOriBox;
        box[k / kOriBox][k % kOriBox] =
            (gy >= 0 && gy < main_row && gx >= 0 && gx < main_col) ? ow[gy * main_col + gx] : 0u;
    }
#pragma unroll
    for (int k = 0; k < 22; k++) bins[k][t] = 0.0;
    __syncthreads();
    const int ty = t / kOriTile, tx = t % kOriTile;
    const long mask_row = y0 + ty, mask_col = x0 + tx;
    if (mask_row >= main_row || mask_col >= main_col) return;
    const int half = 7;
    if (x0 >= kOriHalo && y0 >= kOriHalo && x0 + kOriTile + kOriHalo <= main_col &&
        y0 + kOriTile + kOriHalo <= main_row) {
        // every window of the tile is whole: the reference's clipped bounds
        // reduce to the full disc, unrolled with constant weights and offsets
        const uint32_t *b = &box[ty + kOriHalo][tx + kOriHalo];
#pragma unroll
        for (int row = -half + 1; row <= half - 1; row++) {
#pragma unroll
            for (int col = -kNumCols[row < 0 ? -row : row]; col <= kNumCols[row < 0 ? -row : row]; col++)
                ori_tap(bins, t, b[row * kOriBox + col], GW.w[(half - 1 + row) * 13 + (half - 1 + col)]);
        }
    } else {
        const long rlo = -mask_row > -half + 1 ? -mask_row : -half + 1;
        const long rhi = main_row - mask_row - 1 < half - 1 ? main_row - mask_row - 1 : half - 1;
        for (long row = rlo; row <= rhi; row++) {
            const int dcol = kNumCols[row < 0 ? -row : row];
            const long clo = -mask_row > -dcol ? -mask_row : -dcol;   // as written in the reference
            const long chi = dcol < main_col - mask_col - 1 ? dcol : main_col - mask_col - 1;
            for (long col = clo; col <= chi; col++) {
                const long pr = mask_row + row, pc = mask_col + col;
                uint32_t word;
                if (pc >= 0) {
                    word = box[pr - (y0 - kOriHalo)][pc - (x0 - kOriHalo)];
                } else {
                    // left of column 0: the row-major index wraps into the end
                    // of the previous row, as in the reference; before pixel 0
                    // the tap adds nothing (mag 0)
                    const long p = pr * main_col + pc;
                    if (p < 0) continue;
                    word = ow[p];
                }
                ori_tap(bins, t, word, GW.w[(half - 1 + row) * 13 + (half - 1 + col)]);
            }
        }
    }
    double max_th = -100;
    int sub_th = 0, th = 0;
    for (int i = 0; i < 4; i++)
        if (bins[18 + i][t] > max_th) { max_th = bins[18 + i][t]; sub_th = i; }
    max_th = -100;
    for (int i = (int)(sub_th * kSubRatio); i < (int)(sub_th * kSubRatio + kSubRatio); i++)
        if (bins[i][t] > max_th) { max_th = bins[i][t]; th = i; }
    ori[mask_row * main_col + mask_col] = (uint8_t)th;
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

hipError_t launch_sobel(const uint16_t *in, uint32_t w, uint32_t h, uint16_t *mag, uint32_t *ow, hipStream_t s)
{
    const long n = (long)w * h;
    hipLaunchKernelGGL(k_sobel, dim3(nblk(n)), dim3(256), 0, s, in, w, h, mag, ow);
    return hipGetLastError();
}

hipError_t launch_orientation(const uint32_t *ow, uint32_t w, uint32_t h, uint8_t *ori, hipStream_t s)
{
    OriW GW;
    const int half = 7;
    for (int row = -half + 1; row <= half - 1; row++)
        for (int col = -half + 1; col <= half - 1; col++)
            GW.w[(half - 1 + row) * 13 + (half - 1 + col)] = exp(-(double)(row * row + col * col) / 4.5);
    const dim3 grid((w + kOriTile - 1) / kOriTile, (h + kOriTile - 1) / kOriTile);
    hipLaunchKernelGGL(k_orientation, grid, dim3(256), 0, s, ow, w, h, ori, GW);
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
