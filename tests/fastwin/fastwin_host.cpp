// Host build of the FAST window arithmetic (setsm_amd/csrc/fast_window.hpp)
// for the CPU tests: one hypothesis' multi-window NCC from the packed texel
// images, with the HBM sampler's clamping and pivots, so the FAST numerics
// can be compared with the oracle's FP64 values without a GPU.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "fast_window.hpp"

using namespace smx;

namespace {

struct HostImg {
    const uint32_t *texel;
    int width, height;
    int x0, y0;   // image position of at(0, 0)
    float piv_i, piv_m;
    float2 at(int row, int col) const
    {
        const int x = std::min(std::max(x0 + col, 0), width - 1);
        const int y = std::min(std::max(y0 + row, 0), height - 1);
        const uint32_t t = texel[(long)y * width + x];
        return make_float2((float)(t & 0xffffu) - piv_i, (float)(t >> 16) - piv_m);
    }
    void quad(int iy, int ix, float2 &t00, float2 &t01, float2 &t10, float2 &t11) const
    {
        t00 = at(iy, ix);
        t01 = at(iy, ix + 1);
        t10 = at(iy + 1, ix);
        t11 = at(iy + 1, ix + 1);
    }
    // as the LDS sampler: a texel plus its difference to the next one
    void hlerp(int iy, int ix, float wx, float2 &top, float2 &bot) const
    {
        float2 t00, t01, t10, t11;
        quad(iy, ix, t00, t01, t10, t11);
        top = make_float2(fmaf(wx, t01.x - t00.x, t00.x), fmaf(wx, t01.y - t00.y, t00.y));
        bot = make_float2(fmaf(wx, t11.x - t10.x, t10.x), fmaf(wx, t11.y - t10.y, t10.y));
    }
};

// stage > 0: pivots from the texel `stage` pixels right of and below the
// centre, as the LDS sampler's tile pivot sits away from most of its
// hypotheses; 0: the centre texel, as the HBM sampler.
template <int H>
int window(bool checked, int stage, const uint32_t *tl, int wl, int hl, const uint32_t *tr, int wr, int hr,
           double rpx, double rpy, double tpx, double tpy, float cos0, float sin0, double *rho)
{
    constexpr int O = H + 2;
    const double flx = std::floor(rpx), fly = std::floor(rpy), ftx = std::floor(tpx), fty = std::floor(tpy);
    const int ilx = (int)flx, ily = (int)fly, itx = (int)ftx, ity = (int)fty;
    const int clx = std::min(std::max(ilx + stage, 0), wl - 1), cly = std::min(std::max(ily + stage, 0), hl - 1);
    const int ctx = std::min(std::max(itx + stage, 0), wr - 1), cty = std::min(std::max(ity + stage, 0), hr - 1);
    const uint32_t pl = tl[(long)cly * wl + clx], pr = tr[(long)cty * wr + ctx];
    const HostImg L{tl, wl, hl, ilx - O, ily - O, (float)(pl & 0xffffu), (float)(pl >> 16)};
    const HostImg R{tr, wr, hr, itx - O, ity - O, (float)(pr & 0xffffu), (float)(pr >> 16)};
    bool all3;
    if (checked)
        all3 = fast_window<H, true, true>(L, R, (float)(rpx - flx), (float)(rpy - fly), (float)(tpx - ftx),
                                          (float)(tpy - fty), cos0, sin0, L.piv_i, R.piv_i, -(float)rpx,
                                          (float)((double)wl - 1.0 - rpx), -(float)rpy,
                                          (float)((double)hl - 1.0 - rpy), -(float)itx, (float)(wr - 1 - itx),
                                          -(float)ity, (float)(hr - 1 - ity), *rho);
    else
        all3 = fast_window<H, false, true>(L, R, (float)(rpx - flx), (float)(rpy - fly), (float)(tpx - ftx),
                                           (float)(tpy - fty), cos0, sin0, L.piv_i, R.piv_i, 0, 0, 0, 0, 0, 0, 0, 0,
                                           *rho);
    return all3 ? 1 : 0;
}

}  // namespace

extern "C" int fw_window(int half, int checked, int stage, const uint32_t *tl, int wl, int hl, const uint32_t *tr,
                         int wr, int hr, double rpx, double rpy, double tpx, double tpy, float cos0, float sin0,
                         double *rho)
{
    switch (half) {
    case 5: return window<5>(checked != 0, stage, tl, wl, hl, tr, wr, hr, rpx, rpy, tpx, tpy, cos0, sin0, rho);
    case 6: return window<6>(checked != 0, stage, tl, wl, hl, tr, wr, hr, rpx, rpy, tpx, tpy, cos0, sin0, rho);
    case 7: return window<7>(checked != 0, stage, tl, wl, hl, tr, wr, hr, rpx, rpy, tpx, tpy, cos0, sin0, rho);
    default: return -1;
    }
}
