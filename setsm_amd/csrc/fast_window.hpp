// fast_window.hpp — the FAST path's window arithmetic for one hypothesis:
// bilinear taps about per-hypothesis pivots, FP32 running sums per window
// row folded into the nested inner / middle / outer window totals, and the
// FP64 Pearson finish. Written against a sampler interface (at(), quad(),
// hlerp()),
// so the same code runs in the kernels (LDS and HBM samplers) and on the
// host, where the CPU tests check it against the oracle.
#pragma once

#include <type_traits>
#include <utility>

#include "smx_window.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#ifndef SMX_FN
#define SMX_FN __device__ __forceinline__
#endif
#else
#include <math.h>
#ifndef SMX_FN
#define SMX_FN inline
#endif
struct float2 {
    float x, y;
};
inline float2 make_float2(float x, float y) { return float2{x, y}; }
#endif

namespace smx {

SMX_FN float lerp2(float a, float b, float w) { return fmaf(w, b - a, a); }

SMX_FN float2 bilerp(float2 t00, float2 t01, float2 t10, float2 t11, float wx, float wy)
{
    const float i0 = lerp2(t00.x, t01.x, wx), i1 = lerp2(t10.x, t11.x, wx);
    const float m0 = lerp2(t00.y, t01.y, wx), m1 = lerp2(t10.y, t11.y, wx);
    return make_float2(lerp2(i0, i1, wy), lerp2(m0, m1, wy));
}

// FP32 reciprocal square root (v_rsq_f32 on the GPU).
SMX_FN float rsqrt32(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

// Correlation of one window / channel from pivot-shifted moments. The
// variances and the covariance are formed in FP64 (the differences cancel);
// the ratio cv / sqrt(va vb) only needs a few units in the 7th digit and is
// taken in FP32 (|va vb| < 2^60 for 11-bit data). Returns -99 for an
// undefined correlation, like Correlate.
SMX_FN double corr_moments(double inv_n, double a, double b, double aa, double bb, double ab)
{
    const double va = aa - a * a * inv_n;
    const double vb = bb - b * b * inv_n;
    const double cv = ab - a * b * inv_n;
    if (va > 1e-8 && vb > 1e-8) return (double)((float)cv * rsqrt32((float)(va * vb)));
    return -99.0;
}

// Result of a window loop that this path cannot decide (see clear_of_noise):
// the hypothesis goes to the FP64 path.
constexpr double kRedoRho = 3e300;

// Whether one window / channel's variance (sum of squared deviations v over
// n samples, second moment ss about the pivot p) is clear of FP32 rounding.
// Each sample carries a few roundings of its taps' magnitude (about |p|), and
// the moments a few of ss; a variance not above the sum of 2^14 times that
// floor, ss / 128 and twice the 1e-8 cut of Correlate may be noise where the
// reference's two-pass FP64 Pearson (SubFunctions.cpp:974-1018) sees an
// exact zero (a flat patch, a window left with one valid sample) — or a
// spread so small that the noise would swamp the correlation. (The sum, one
// fused test per window channel, costs half the instructions of three
// separate tests.)
SMX_FN bool clear_of_noise(double n, double v, double ss, float p)
{
    const double pp = (double)p * (double)p;
    return v > fma(ss, 0x1p-7, fma(n * 0x1p-16, pp, 2e-8));
}

// Half-width of window w (0 outer disc, 1 middle, 2 inner) on row `row`:
// the largest |col| inside it, -1 if the row misses the window. Membership of
// every window is symmetric and contiguous in col on a row.
SMX_HD constexpr int half_width_ce(int H, int row, int w)
{
    int hw = -1;
    for (int c = 0; c <= H; c++)
        if (window_bits_ce(H, row, c) & (1 << w)) hw = c;
    return hw;
}

template <int H>
struct RowWidths {
    int w[2 * H + 1][3];
    constexpr RowWidths() : w()
    {
        for (int r = -H; r <= H; r++)
            for (int k = 0; k < 3; k++) w[r + H][k] = half_width_ce(H, r, k);
    }
};

// Distinct (outer, middle, inner) half-width triples over the rows of a
// window, and the triple of each row: rows sharing a triple share code.
template <int H>
struct RowSigs {
    int n;
    int w[2 * H + 1][3];
    int of_row[2 * H + 1];
    constexpr RowSigs() : n(0), w(), of_row()
    {
        for (int r = -H; r <= H; r++) {
            const int a = half_width_ce(H, r, 0), b = half_width_ce(H, r, 1), c = half_width_ce(H, r, 2);
            int k = 0;
            while (k < n && !(w[k][0] == a && w[k][1] == b && w[k][2] == c)) k++;
            if (k == n) {
                w[n][0] = a;
                w[n][1] = b;
                w[n][2] = c;
                n++;
            }
            of_row[r + H] = k;
        }
    }
};

struct Mom {
    float a, b, aa, bb, ab;   // intensity pair
    float c, d, cc, dd, cd;   // magnitude pair
};

SMX_FN void mom_add(Mom &s, float li, float ri, float lm, float rm)
{
    s.a += li;
    s.b += ri;
    s.aa = fmaf(li, li, s.aa);
    s.bb = fmaf(ri, ri, s.bb);
    s.ab = fmaf(li, ri, s.ab);
    s.c += lm;
    s.d += rm;
    s.cc = fmaf(lm, lm, s.cc);
    s.dd = fmaf(rm, rm, s.dd);
    s.cd = fmaf(lm, rm, s.cd);
}

// Window totals: sums of the row partials; the squared / product terms in
// FP64 when kWide (the row partials themselves are short FP32 sums, at most
// 2H+1 samples each), FP32 otherwise. No compensation term is carried: the
// FP64 totals make the rounding of the additions negligible.
template <bool kWide>
struct Tot {
    using Q = typename std::conditional<kWide, double, float>::type;
    float a, b, c, d, n;
    Q aa, bb, ab, cc, dd, cd;
};

template <bool kWide>
SMX_FN void tot_zero(Tot<kWide> &t)
{
    t.a = t.b = t.c = t.d = t.n = 0.f;
    t.aa = t.bb = t.ab = t.cc = t.dd = t.cd = 0;
}

template <bool kWide>
SMX_FN void mom_fold(Tot<kWide> &t, const Mom &r, float rn)
{
    using Q = typename Tot<kWide>::Q;
    t.a += r.a;
    t.b += r.b;
    t.c += r.c;
    t.d += r.d;
    t.n += rn;
    t.aa += (Q)r.aa;
    t.bb += (Q)r.bb;
    t.ab += (Q)r.ab;
    t.cc += (Q)r.cc;
    t.dd += (Q)r.dd;
    t.cd += (Q)r.cd;
}

// Fractional part of a positive coordinate (v_fract_f32 on the GPU).
SMX_FN float fract_pos(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fractf(x);
#else
    return x - floorf(x);
#endif
}

// a * b + c for |a|, |b| < 2^23 (one v_mad_i32_i24 on the GPU).
SMX_FN int imad24(int a, int b, int c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __mul24(a, b) + c;
#else
    return a * b + c;
#endif
}

// f(integral_constant<int, I>) for I = Lo .. Hi - 1, unrolled at compile time.
template <int Lo, class F, int... I>
SMX_FN void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, Lo + I>{}), ...);
}
template <int Lo, int Hi, class F>
SMX_FN void static_for(F &&f)
{
    if constexpr (Hi > Lo) static_for_impl<Lo>(f, std::make_integer_sequence<int, Hi - Lo>{});
}

// Marks the window totals as used here. The product sums are read only
// inside the Pearson finish's conditional branch, and without this the
// compiler sinks their whole accumulation chain there, holding every
// sample of the window in registers until then.
template <bool kWide>
SMX_FN void pin_totals(const Tot<kWide> &t)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" ::"v"(t.a), "v"(t.b), "v"(t.c), "v"(t.d), "v"(t.aa), "v"(t.bb), "v"(t.ab), "v"(t.cc),
                 "v"(t.dd), "v"(t.cd));
#else
    (void)t;
#endif
}

// One hypothesis' window loop. L/R are samplers positioned so that at(0,0)
// is the texel (center - (H+2)) in both axes; R.quad(iy, ix) returns the
// taps (iy, ix) .. (iy + 1, ix + 1) about that origin and R.hlerp(iy, ix,
// wx) the interpolations at wx along its two rows (the LDS sampler keeps
// each texel with its difference to the next one, so that is one FMA per
// row and channel). (fxl, fyl) is the fractional part of the left centre,
// (ftx, fty) of the right one.
// Check = per-sample image-edge and nodata tests (SubFunctions.cpp:3252-3257);
// otherwise every sample is known to be valid. Returns false if a window
// count is zero; rho = kRedoRho when a window's spread is not clear of FP32
// rounding.
//
// Each row is walked from its centre column outwards into one running sum;
// the sum is folded into the inner / middle / outer window totals when the
// walk passes that window's half-width on the row (the windows are nested
// and each is a symmetric column range on a row), so every sample updates
// one fixed set of registers. Without the per-sample test every row is
// straight-line code with its column range fixed at compile time.
//
// The right window's sample positions are taken about the window origin:
// every sample of the outer window lies within H + 1 of the centre, so its
// position is at least 1 in both axes (origin offset H + 2, centre fraction
// >= 0) — truncation is the floor and v_fract the bilinear weight.
template <int H, bool kCheck, bool kWide, class SL, class SR>
SMX_FN bool fast_window(const SL &L, const SR &R, float fxl, float fyl, float ftx, float fty,
                                         float cos0, float sin0, float piv_l, float piv_r, float lim_lx0,
                                         float lim_lx1, float lim_ly0, float lim_ly1, float lim_rx0,
                                         float lim_rx1, float lim_ry0, float lim_ry1, double &rho)
{
    constexpr int O = H + 2;   // origin offset
    constexpr RowWidths<H> kW{};
    Tot<kWide> tot[3];
#pragma unroll
    for (int k = 0; k < 3; k++) tot_zero(tot[k]);
    // Per-hypothesis pivots: the window-centre samples. Moments are taken on
    // sample - pivot, so a low-contrast patch far from the tile's staging
    // pivot does not lose its variance to cancellation.
    const float2 pvl = bilerp(L.at(O, O), L.at(O, O + 1), L.at(O + 1, O), L.at(O + 1, O + 1), fxl, fyl);
    const float2 pvr = bilerp(R.at(O, O), R.at(O, O + 1), R.at(O + 1, O), R.at(O + 1, O + 1), ftx, fty);
    // the left patch shares the centre's fractional offset: one set of weights
    const float lw11 = fxl * fyl, lw10 = fyl - lw11, lw01 = fxl - lw11, lw00 = (1.f - fxl) - lw10;
    // right centre about the window origin
    const float ftxo = ftx + (float)O, ftyo = fty + (float)O;
    Mom run = Mom{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    float rn = 0.f;
    // One sample at (row, col) of the window; xr / yr: the row's rotated
    // start about the origin. first: the row's first sample (no test).
    auto sample = [&](const int row, const int col, const float xr, const float yr, const bool first) {
        // left: integer offsets, the hypothesis' bilinear weights; the
        // pivot enters as the first term of the weighted sum
        const float2 a = L.at(O + row, O + col), b = L.at(O + row, O + col + 1);
        const float2 c = L.at(O + row + 1, O + col), d = L.at(O + row + 1, O + col + 1);
        const float li = fmaf(lw11, d.x, fmaf(lw10, c.x, fmaf(lw01, b.x, fmaf(lw00, a.x, -pvl.x))));
        const float lm = fmaf(lw11, d.y, fmaf(lw10, c.y, fmaf(lw01, b.y, fmaf(lw00, a.y, -pvl.y))));
        // right: rotated offsets (setsm_code.cpp:6437-6439),
        // x = ftx + col*cos - row*sin, y = fty + col*sin + row*cos; the
        // sampler interpolates along both tap rows at wx, then the rows are
        // interpolated at wy about the pivot
        const float x = col == 0 ? xr : fmaf((float)col, cos0, xr);
        const float y = col == 0 ? yr : fmaf((float)col, sin0, yr);
        const float wx = fract_pos(x), wy = fract_pos(y);
        float2 top, bot;
        R.hlerp((int)y, (int)x, wx, top, bot);
        const float ri = fmaf(wy, bot.x - top.x, top.x - pvr.x);
        const float rm = fmaf(wy, bot.y - top.y, top.y - pvr.y);
        bool ok = true;
        if (kCheck) {
            float2 r00, r01, r10, r11;
            R.quad((int)y, (int)x, r00, r01, r10, r11);
            // pos >= 0 && pos + 1 < size on both images, then L > 0 && R > 0
            // (SubFunctions.cpp:3252-3257). The interpolation of
            // non-negative DNs is positive exactly when a tap with a
            // non-zero weight is positive; tested on the taps, which hold
            // DN - pivot exactly, so rounding in the weighted sum cannot
            // turn a nodata sample into a valid one.
            const float frow = (float)row;
            const bool lpos = a.x > -piv_l || (fxl > 0.f && b.x > -piv_l) || (fyl > 0.f && c.x > -piv_l) ||
                              (fxl > 0.f && fyl > 0.f && d.x > -piv_l);
            const bool rpos = r00.x > -piv_r || (wx > 0.f && r01.x > -piv_r) || (wy > 0.f && r10.x > -piv_r) ||
                              (wx > 0.f && wy > 0.f && r11.x > -piv_r);
            ok = (float)col >= lim_lx0 && (float)col < lim_lx1 && frow >= lim_ly0 && frow < lim_ly1 &&
                 x >= lim_rx0 + (float)O && x < lim_rx1 + (float)O && y >= lim_ry0 + (float)O &&
                 y < lim_ry1 + (float)O && lpos && rpos;
        }
        if (!kCheck && first) {
            run = Mom{li, ri, li * li, ri * ri, li * ri, lm, rm, lm * lm, rm * rm, lm * rm};
        } else if (ok) {
            mom_add(run, li, ri, lm, rm);
            if (kCheck) rn += 1.f;
        }
    };
    if constexpr (!kCheck) {
        // every sample valid: straight-line code, rows and columns unrolled
        static_for<-H, H + 1>([&](auto RowC) {
            constexpr int row = decltype(RowC)::value;
            constexpr int W0 = kW.w[row + H][0], W1 = kW.w[row + H][1], W2 = kW.w[row + H][2];
            if constexpr (W0 >= 0) {
                const float xr = row == 0 ? ftxo : fmaf(-(float)row, sin0, ftxo);
                const float yr = row == 0 ? ftyo : fmaf((float)row, cos0, ftyo);
                static_for<0, W0 + 1>([&](auto KC) {
                    constexpr int k = decltype(KC)::value;
                    sample(row, k, xr, yr, k == 0);
                    if constexpr (k > 0) sample(row, -k, xr, yr, false);
                    if constexpr (k == W2) mom_fold(tot[2], run, rn);
                    if constexpr (k == W1) mom_fold(tot[1], run, rn);
                });
                mom_fold(tot[0], run, rn);
                pin_totals(tot[0]);
                if constexpr (W1 >= 0) pin_totals(tot[1]);
                if constexpr (W2 >= 0) pin_totals(tot[2]);
            }
        });
    } else {
#pragma unroll 1
        for (int row = -H; row <= H; row++) {
            const int w0 = kW.w[row + H][0], w1 = kW.w[row + H][1], w2 = kW.w[row + H][2];
            const float frow = (float)row;
            const float xr = fmaf(-frow, sin0, ftxo), yr = fmaf(frow, cos0, ftyo);
            run = Mom{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            rn = 0.f;
#pragma unroll
            for (int k = 0; k <= H; k++) {
                if (k <= w0) {
                    sample(row, k, xr, yr, false);
                    if (k > 0) sample(row, -k, xr, yr, false);
                }
                if (k == w2) mom_fold(tot[2], run, rn);
                if (k == w1) mom_fold(tot[1], run, rn);
            }
            if (w0 >= 0) mom_fold(tot[0], run, rn);
        }
    }
    // window totals (FP64 sums of the FP32 row partials when kWide)
    double n[3], S[3][10];
#pragma unroll
    for (int w = 0; w < 3; w++) {
        n[w] = kCheck ? (double)tot[w].n : (double)window_count_ce(H, w);
        S[w][0] = tot[w].a;
        S[w][1] = tot[w].b;
        S[w][2] = (double)tot[w].aa;
        S[w][3] = (double)tot[w].bb;
        S[w][4] = (double)tot[w].ab;
        S[w][5] = tot[w].c;
        S[w][6] = tot[w].d;
        S[w][7] = (double)tot[w].cc;
        S[w][8] = (double)tot[w].dd;
        S[w][9] = (double)tot[w].cd;
    }
    if (!(n[0] > 0 && n[1] > 0 && n[2] > 0)) return false;
    double temp_roh = 0, count_roh = 0;
    bool clear = true;
#pragma unroll
    for (int w = 0; w < 3; w++) {
        const double inv = kCheck ? 1.0 / n[w] : 1.0 / (double)window_count_ce(H, w);
        clear = clear && clear_of_noise(n[w], S[w][2] - S[w][0] * S[w][0] * inv, S[w][2], pvl.x) &&
                clear_of_noise(n[w], S[w][3] - S[w][1] * S[w][1] * inv, S[w][3], pvr.x) &&
                clear_of_noise(n[w], S[w][7] - S[w][5] * S[w][5] * inv, S[w][7], pvl.y) &&
                clear_of_noise(n[w], S[w][8] - S[w][6] * S[w][6] * inv, S[w][8], pvr.y);
        const double c1 = corr_moments(inv, S[w][0], S[w][1], S[w][2], S[w][3], S[w][4]);
        if (c1 != -99) { count_roh++; temp_roh += c1; }
        const double c2 = corr_moments(inv, S[w][5], S[w][6], S[w][7], S[w][8], S[w][9]);
        if (c2 != -99) { count_roh++; temp_roh += c2; }
    }
    rho = !clear ? kRedoRho : count_roh > 0 ? temp_roh / count_roh : 1e300;   // 1e300 = no valid correlation
    return true;
}


}  // namespace smx
