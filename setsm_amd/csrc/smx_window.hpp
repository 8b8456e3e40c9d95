// smx_window.hpp — the nested window geometry of SETSM's multi-window NCC,
// usable from device code and from plain host C++.
#pragma once

#ifndef SMX_HD
#if defined(__HIPCC__)
#define SMX_HD __host__ __device__
#else
#define SMX_HD
#endif
#endif

namespace smx {

// SubFunctions.cpp:3274-3312 plus the outer disc test of setsm_code.cpp:6434:
// bit 0 = 11x11 disc, bit 1 = middle window, bit 2 = inner window.
SMX_HD constexpr int window_bits_ce(int half, int row, int col)
{
    const int r2 = row * row + col * col;
    const int s1 = half / 2;
    const int s2 = s1 + (int)((s1 / 2.0) + 0.5);
    return (row < -half || row > half || col < -half || col > half || r2 > (half + 1) * (half + 1))
               ? 0
               : (1 | ((r2 <= (half - s1 + 1) * (half - s1 + 1) && row >= -half + s1 && row <= half - s1 &&
                        col >= -half + s1 && col <= half - s1)
                           ? 2
                           : 0) |
                  ((r2 <= (half - s2 + 1) * (half - s2 + 1) && row >= -half + s2 && row <= half - s2 &&
                    col >= -half + s2 && col <= half - s2)
                       ? 4
                       : 0));
}

// Samples of ring k (0 = outer only, 1 = middle only, 2 = inner window).
SMX_HD constexpr int ring_count_ce(int half, int ring)
{
    int n = 0;
    for (int r = -half; r <= half; r++)
        for (int c = -half; c <= half; c++) {
            const int b = window_bits_ce(half, r, c);
            const int k = (b & 4) ? 2 : ((b & 2) ? 1 : 0);
            if (b && k == ring) n++;
        }
    return n;
}

// Samples of window w (0 = outer disc, 1 = middle, 2 = inner).
SMX_HD constexpr int window_count_ce(int half, int w)
{
    int n = 0;
    for (int r = -half; r <= half; r++)
        for (int c = -half; c <= half; c++)
            if (window_bits_ce(half, r, c) & (1 << w)) n++;
    return n;
}


}  // namespace smx
