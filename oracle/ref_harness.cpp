/*
 * ref_harness.cpp — extern "C" entry points over the reference's own
 * header-only code, compiled straight from /root/reference by oracle/Makefile
 * into oracle/_ref/libsmx_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ to pin the C restatement in
 * smx_oracle.c against the reference sources that build here without any
 * stand-in:
 *   GetObjectToImageRPC_single  CoordConversion.hpp:26-87
 *   OriginalToPyramid_single    CoordConversion.hpp:96-105
 *   CreateImagePyramid<uint16>  Template.hpp:82-132 (9x9, sigma 1.5)
 *   DoubleToSignedChar_*        Template.hpp:964-992
 *   struct layouts              Typedefine.hpp:248-293, 351-419, 604-643
 * SubFunctions.cpp / setsm_code.cpp need libgeotiff headers and a generated
 * git_description.h that this image lacks, so they are not built.
 */
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cmath>
#include <cstring>
#include "CoordConversion.hpp"
#include "Template.hpp"

extern "C" {

void ref_rpc_ground_to_image(const double *rpc90, const double *adjust, double lon, double lat, double h,
                             double *x, double *y)
{
    double r0[5], r1[5], r2[20], r3[20], r4[20], r5[20], r6[2] = {0, 0};
    std::memcpy(r0, rpc90, sizeof r0);
    std::memcpy(r1, rpc90 + 5, sizeof r1);
    std::memcpy(r2, rpc90 + 10, sizeof r2);
    std::memcpy(r3, rpc90 + 30, sizeof r3);
    std::memcpy(r4, rpc90 + 50, sizeof r4);
    std::memcpy(r5, rpc90 + 70, sizeof r5);
    const double *rows[7] = {r0, r1, r2, r3, r4, r5, r6};
    D3DPOINT gp(lon, lat, h);
    D2DPOINT ip = GetObjectToImageRPC_single(rows, 2, adjust, gp);
    *x = ip.m_X;
    *y = ip.m_Y;
}

void ref_to_pyramid(double x, double y, double sx, double sy, int level, double *px, double *py)
{
    D2DPOINT in(x, y), st(sx, sy);
    D2DPOINT o = OriginalToPyramid_single(in, st, (uint8)level);
    *px = o.m_X;
    *py = o.m_Y;
}

void ref_gauss_pyramid(const uint16 *in, unsigned w, unsigned h, uint16 *out)
{
    CSize s(w, h);
    uint16 *r = CreateImagePyramid<uint16>(const_cast<uint16 *>(in), s, 9, 1.5);
    std::memcpy(out, r, sizeof(uint16) * (size_t)(w / 2) * (size_t)(h / 2));
    free(r);
}

short ref_q_result(double v) { return DoubleToSignedChar_result(v); }
short ref_q_voxel(double v) { return DoubleToSignedChar_voxel(v); }

/* Layout facts of the boundary records, as the reference headers define them. */
long ref_layout(int which)
{
    switch (which) {
    case 0: return (long)sizeof(UGRID);
    case 1: return (long)offsetof(UGRID, Matched_flag);
    case 2: return (long)offsetof(UGRID, ortho_ncc);
    case 3: return (long)sizeof(NCCresult);
    case 4: return (long)offsetof(NCCresult, NumOfHeight);
    case 5: return (long)offsetof(NCCresult, check_height_change);
    case 6: return (long)offsetof(NCCresult, max_WNCC);
    case 7: return (long)offsetof(NCCresult, GNCC);
    case 8: return (long)sizeof(VOXEL);
    case 9: return (long)offsetof(VOXEL, INCC);
    case 10: return (long)offsetof(NCCresult, minHeight);
    case 11: return (long)offsetof(NCCresult, result4);
    default: return -1;
    }
}

}
