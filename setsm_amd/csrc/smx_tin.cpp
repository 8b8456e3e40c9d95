// smx_tin.cpp — a Delaunay TIN of matched points for callers without one
// (the role of SETSM's TINCreate_list, SubFunctions.cpp:2768-2846; host code,
// no GPU). The points are snapped to the level's lattice as the reference
// snaps them (0.5 + (x - min_x) / resolution, truncated), so every predicate
// is exact integer arithmetic: the orientation in 64-bit, the in-circle test
// in 128-bit.
//
// Radial sweep: the points are inserted in order of their squared distance
// from a lattice point at the middle of their bounding box. All earlier
// points lie in the closed disc through the new one, and a point on a
// circle is an extreme point of its disc, so every new point lies strictly
// outside the hull built so far: it is joined to the hull edges it sees and
// the new edges are made Delaunay by flips. An angular hash of the hull
// vertices finds a visible edge quickly. A run of collinear points at the
// start is joined to the first point off their line by a fan.
//
// The triangles come out clockwise (as the reference's TIN lists them);
// among cocircular points the split is this sweep's own, so the triangle set
// can differ from the reference's while both are Delaunay.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "smx_types.h"

namespace {

struct Pt {
    int64_t x, y;
};

inline int64_t orient(const Pt &a, const Pt &b, const Pt &c)
{
    return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
}

// > 0: d strictly inside the circle through the counter-clockwise a, b, c
inline bool in_circle(const Pt &a, const Pt &b, const Pt &c, const Pt &d)
{
    const __int128 ax = a.x - d.x, ay = a.y - d.y, bx = b.x - d.x, by = b.y - d.y, cx = c.x - d.x, cy = c.y - d.y;
    const __int128 a2 = ax * ax + ay * ay, b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
    const __int128 det = a2 * (bx * cy - cx * by) - b2 * (ax * cy - cx * ay) + c2 * (ax * by - bx * ay);
    return det > 0;
}

struct Tin {
    std::vector<Pt> p;
    std::vector<int> tv;    // 3 per triangle, counter-clockwise
    std::vector<int> he;    // twin half-edge, -1 on the hull
    std::vector<int> hnext, hprev, htri;   // hull (counter-clockwise); htri[v]: half-edge v -> hnext[v]
    std::vector<int> hash;
    Pt c{0, 0};
    int hsize = 1;
    std::vector<int> stack;

    int key(const Pt &q) const
    {
        // pseudo-angle in [0, 1) of q about the centre
        const double dx = (double)(q.x - c.x), dy = (double)(q.y - c.y);
        const double s = std::fabs(dx) + std::fabs(dy);
        const double r = s > 0 ? dx / s : 0.0;
        const double a = (dy > 0 ? 3.0 - r : 1.0 + r) / 4.0;
        int k = (int)std::floor(a * hsize);
        return k < 0 ? 0 : (k >= hsize ? hsize - 1 : k);
    }
    void link(int a, int b)
    {
        he[(size_t)a] = b;
        if (b >= 0) he[(size_t)b] = a;
    }
    int add(int i0, int i1, int i2, int a, int b, int cc)
    {
        const int t = (int)tv.size();
        tv.push_back(i0);
        tv.push_back(i1);
        tv.push_back(i2);
        he.push_back(-1);
        he.push_back(-1);
        he.push_back(-1);
        link(t, a);
        link(t + 1, b);
        link(t + 2, cc);
        return t;
    }
    // Lawson flips from half-edge a (opposite the point just inserted)
    void legalize(int a)
    {
        stack.clear();
        for (;;) {
            const int b = he[(size_t)a];
            const int a0 = a - a % 3, ar = a0 + (a + 2) % 3, al = a0 + (a + 1) % 3;
            bool flipped = false;
            if (b >= 0) {
                const int b0 = b - b % 3, bl = b0 + (b + 2) % 3, br = b0 + (b + 1) % 3;
                const int p0 = tv[(size_t)ar], pr = tv[(size_t)a], pl = tv[(size_t)al], p1 = tv[(size_t)bl];
                if (in_circle(p[(size_t)pr], p[(size_t)pl], p[(size_t)p0], p[(size_t)p1])) {
                    tv[(size_t)a] = p1;
                    tv[(size_t)b] = p0;
                    const int hbl = he[(size_t)bl], har = he[(size_t)ar];
                    if (hbl < 0) htri[(size_t)p1] = a;     // hull edge p1 -> pl now half-edge a
                    if (har < 0) htri[(size_t)p0] = b;     // hull edge p0 -> pr now half-edge b
                    link(a, hbl);
                    link(b, har);
                    link(ar, bl);
                    stack.push_back(br);
                    flipped = true;
                }
            }
            if (!flipped) {
                if (stack.empty()) return;
                a = stack.back();
                stack.pop_back();
            }
        }
    }
};

}  // namespace

namespace {
long tin_create(const smx_point3 *pts, long npts, const double min_max[4], double resolution, smx_tri *out, long cap)
{
    if (npts < 0 || (npts > 0 && (!pts || !min_max)) || !(resolution > 0)) return -3;
    if (npts <= 2) return 0;
    // SubFunctions.cpp:2780-2800: lattice positions, the last
    // point of a lattice cell standing for it
    const double w = 1 + (min_max[2] - min_max[0]) / resolution, h = 1 + (min_max[3] - min_max[1]) / resolution;
    if (!(w <= 32767 && h <= 32767)) return -1;
    std::unordered_map<int64_t, int> cell;
    cell.reserve((size_t)npts * 2);
    for (long t = 0; t < npts; t++) {
        const long col = (long)(0.5 + (pts[t].x - min_max[0]) / resolution);
        const long row = (long)(0.5 + (pts[t].y - min_max[1]) / resolution);
        if (col < 0 || row < 0 || col > 32767 || row > 32767) return -1;
        cell[row * 32768 + col] = (int)t;
    }
    Tin T;
    std::vector<int> orig;
    orig.reserve(cell.size());
    for (const auto &kv : cell) orig.push_back(kv.second);
    std::sort(orig.begin(), orig.end());
    const int n = (int)orig.size();
    T.p.resize((size_t)n);
    int64_t x0 = INT64_MAX, y0 = INT64_MAX, x1 = INT64_MIN, y1 = INT64_MIN;
    for (int i = 0; i < n; i++) {
        const smx_point3 &q = pts[orig[(size_t)i]];
        T.p[(size_t)i] = Pt{(int64_t)(0.5 + (q.x - min_max[0]) / resolution), (int64_t)(0.5 + (q.y - min_max[1]) / resolution)};
        x0 = std::min(x0, T.p[(size_t)i].x);
        x1 = std::max(x1, T.p[(size_t)i].x);
        y0 = std::min(y0, T.p[(size_t)i].y);
        y1 = std::max(y1, T.p[(size_t)i].y);
    }
    if (n < 3) return 0;
    T.c = Pt{(x0 + x1) / 2, (y0 + y1) / 2};
    std::vector<int64_t> d2((size_t)n);
    for (int i = 0; i < n; i++) {
        const int64_t dx = T.p[(size_t)i].x - T.c.x, dy = T.p[(size_t)i].y - T.c.y;
        d2[(size_t)i] = dx * dx + dy * dy;
    }
    std::vector<int> ord((size_t)n);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return d2[(size_t)a] != d2[(size_t)b] ? d2[(size_t)a] < d2[(size_t)b] : a < b; });
    // the first point off the line of the two nearest
    int k = 2;
    while (k < n && orient(T.p[(size_t)ord[0]], T.p[(size_t)ord[1]], T.p[(size_t)ord[(size_t)k]]) == 0) k++;
    if (k == n) return 0;   // all on one line
    std::vector<int> chain(ord.begin(), ord.begin() + k);
    {
        const Pt a = T.p[(size_t)chain[0]], b = T.p[(size_t)chain[1]];
        const int64_t dx = b.x - a.x, dy = b.y - a.y;
        std::sort(chain.begin(), chain.end(), [&](int u, int v) {
            return (T.p[(size_t)u].x - a.x) * dx + (T.p[(size_t)u].y - a.y) * dy <
                   (T.p[(size_t)v].x - a.x) * dx + (T.p[(size_t)v].y - a.y) * dy;
        });
    }
    const int q = ord[(size_t)k];
    if (orient(T.p[(size_t)chain[0]], T.p[(size_t)chain[1]], T.p[(size_t)q]) < 0) std::reverse(chain.begin(), chain.end());
    // now q lies left of the chain's direction: the fan (chain[j], chain[j+1], q) is counter-clockwise
    T.tv.reserve((size_t)n * 6);
    T.he.reserve((size_t)n * 6);
    T.hnext.assign((size_t)n, -1);
    T.hprev.assign((size_t)n, -1);
    T.htri.assign((size_t)n, -1);
    T.hsize = std::max(1, (int)std::ceil(std::sqrt((double)n)));
    T.hash.assign((size_t)T.hsize, -1);
    int prev_t = -1;
    const int m = (int)chain.size();
    for (int j = 0; j + 1 < m; j++) {
        const int t = T.add(chain[(size_t)j], chain[(size_t)j + 1], q, -1, -1, prev_t >= 0 ? prev_t + 1 : -1);
        // half-edges of t: t: c[j] -> c[j+1] (hull), t+1: c[j+1] -> q, t+2: q -> c[j]
        T.htri[(size_t)chain[(size_t)j]] = t;
        prev_t = t;
    }
    // hull: c0 -> c1 -> ... -> c(m-1) -> q -> c0
    for (int j = 0; j + 1 < m; j++) {
        T.hnext[(size_t)chain[(size_t)j]] = chain[(size_t)j + 1];
        T.hprev[(size_t)chain[(size_t)j + 1]] = chain[(size_t)j];
    }
    T.hnext[(size_t)chain[(size_t)m - 1]] = q;
    T.hprev[(size_t)q] = chain[(size_t)m - 1];
    T.hnext[(size_t)q] = chain[0];
    T.hprev[(size_t)chain[0]] = q;
    T.htri[(size_t)chain[(size_t)m - 1]] = prev_t + 1;   // c(m-1) -> q
    T.htri[(size_t)q] = 2;                              // q -> c0, in the first fan triangle
    // (the fan is the only triangulation of a collinear run and one point off
    // its line, so it is Delaunay as it stands)
    for (int v : chain) T.hash[(size_t)T.key(T.p[(size_t)v])] = v;
    T.hash[(size_t)T.key(T.p[(size_t)q])] = q;
    std::vector<char> on_hull((size_t)n, 0);
    for (int v : chain) on_hull[(size_t)v] = 1;
    on_hull[(size_t)q] = 1;
    // the run's inner points are hull vertices too, on straight hull edges
    for (int kk = k + 1; kk < n; kk++) {
        const int i = ord[(size_t)kk];
        const Pt &P = T.p[(size_t)i];
        int start = -1;
        const int key = T.key(P);
        for (int j = 0; j < T.hsize; j++) {
            const int s = T.hash[(size_t)((key + j) % T.hsize)];
            if (s >= 0 && on_hull[(size_t)s]) {
                start = s;
                break;
            }
        }
        if (start < 0) return -4;
        start = T.hprev[(size_t)start];
        int e = start;
        while (orient(T.p[(size_t)e], T.p[(size_t)T.hnext[(size_t)e]], P) >= 0) {
            e = T.hnext[(size_t)e];
            if (e == start) return -4;   // not outside the hull: cannot happen for distinct lattice points
        }
        // the visible chain: back from e while the previous edge is visible too
        int s0 = e;
        for (int guard = 0; orient(T.p[(size_t)T.hprev[(size_t)s0]], T.p[(size_t)s0], P) < 0; guard++) {
            if (guard > n) return -4;
            s0 = T.hprev[(size_t)s0];
        }
        // triangles over the visible edges s0 -> ... -> e1
        int v = s0, last = -1, first = -1;
        while (orient(T.p[(size_t)v], T.p[(size_t)T.hnext[(size_t)v]], P) < 0) {
            const int w = T.hnext[(size_t)v];
            // (w, v, i): half-edges t: w -> v (twin: hull edge v -> w), t+1: v -> i, t+2: i -> w
            const int t = T.add(w, v, i, T.htri[(size_t)v], last >= 0 ? last + 2 : -1, -1);
            if (first < 0) first = t;
            if (v != s0) on_hull[(size_t)v] = 0;
            last = t;
            v = w;
        }
        // hull: s0 -> i -> v
        T.hnext[(size_t)s0] = i;
        T.hprev[(size_t)i] = s0;
        T.hnext[(size_t)i] = v;
        T.hprev[(size_t)v] = i;
        T.htri[(size_t)s0] = first + 1;   // s0 -> i
        T.htri[(size_t)i] = last + 2;     // i -> v
        on_hull[(size_t)i] = 1;
        for (int t = first; t <= last; t += 3) T.legalize(t);
        T.hash[(size_t)T.key(P)] = i;
        T.hash[(size_t)T.key(T.p[(size_t)s0])] = s0;
    }
    const long nt = (long)(T.tv.size() / 3);
    if (nt > cap) return -2;
    for (long t = 0; t < nt; t++) {
        // clockwise, as the reference lists its triangles
        out[t].a = (uint32_t)orig[(size_t)T.tv[(size_t)(3 * t)]];
        out[t].b = (uint32_t)orig[(size_t)T.tv[(size_t)(3 * t + 2)]];
        out[t].c = (uint32_t)orig[(size_t)T.tv[(size_t)(3 * t + 1)]];
    }
    return nt;
}
}  // namespace

extern "C" long smx_tin_create(const smx_point3 *pts, long npts, const double min_max[4], double resolution,
                               smx_tri *out, long cap)
{
    try {
        return tin_create(pts, npts, min_max, resolution, out, cap);
    } catch (...) {   // host memory: no exception crosses the C ABI
        return -5;
    }
}
