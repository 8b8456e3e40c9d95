// ref_tin.cpp — C entry point to the reference's own TIN (grid_triangulation.cpp,
// compiled from the reference sources into oracle/_ref/libref_tin.so).
//
// TEST INFRASTRUCTURE ONLY: tests use it to produce the triangles SETSM's
// host loop would hand to SetHeightRange. It follows the steps of
// TINCreate_list (SubFunctions.cpp:2768-2846): matched points are snapped to
// the grid of the boundary at the given resolution (col, row = 0.5 + offset /
// resolution, truncated), triangulated by FullTriangulation, and the
// triangles' grid corners mapped back to point indices (the last point on a
// grid node wins, as the reference's hash map does).
#include <cstdint>
#include <cstdio>
#include <unordered_map>
#include <vector>

#include "grid_triangulation.hpp"
#include "smx_types.h"

extern "C" long ref_tin_create(const smx_point3 *pts, long npts, const double min_max[4], double resolution,
                               smx_tri *out, long cap)
{
    if (npts <= 2) return 0;
    const INDEX width = 1 + (min_max[2] - min_max[0]) / resolution;
    const INDEX height = 1 + (min_max[3] - min_max[1]) / resolution;
    if (width > 32767 || height > 32767) return -1;
    std::vector<GridPoint> gp((size_t)npts);
    std::vector<GridPoint *> ptrs((size_t)npts);
    std::unordered_map<std::size_t, std::size_t> owner;
    owner.reserve((size_t)npts);
    for (long t = 0; t < npts; t++) {
        gp[t].col = 0.5 + (pts[t].x - min_max[0]) / resolution;
        gp[t].row = 0.5 + (pts[t].y - min_max[1]) / resolution;
        owner[(std::size_t)gp[t].row * width + gp[t].col] = (std::size_t)t;
        ptrs[t] = &gp[t];
    }
    FullTriangulation tin(width, height);
    tin.Triangulate(ptrs.data(), (size_t)npts);
    std::vector<Tri> tris;
    tin.GetAllTris(&tris);
    if ((long)tris.size() > cap) return -2 - (long)tris.size();
    for (size_t t = 0; t < tris.size(); t++) {
        const GridPoint *q = tris[t].pts;
        out[t].a = (uint32_t)owner[(std::size_t)q[0].row * width + q[0].col];
        out[t].b = (uint32_t)owner[(std::size_t)q[1].row * width + q[1].col];
        out[t].c = (uint32_t)owner[(std::size_t)q[2].row * width + q[2].col];
    }
    return (long)tris.size();
}
