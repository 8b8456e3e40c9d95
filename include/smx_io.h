/*
 * smx_io.h — C ABI of the host-side formats the matching path reads and
 * writes (libsmx_io.so): SETSM's TIFF subset reader, its GeoTIFF writer and
 * the DigitalGlobe XML RPC reader. Plain C, no GPU, no third-party library:
 * baseline and BigTIFF, uncompressed or LZW, horizontal / floating-point
 * predictors, strips or tiles.
 *
 * All functions return 0 on success and a negative SMX_IO_E* code on
 * failure; smx_io_last_error() describes the last failure of the calling
 * thread.
 */
#ifndef SMX_IO_H
#define SMX_IO_H

#include <stddef.h>
#include <stdint.h>

#include "smx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SMX_IO_OK        0
#define SMX_IO_E_OPEN   -1   /* file cannot be opened / created        */
#define SMX_IO_E_FORMAT -2   /* not a TIFF / XML this reader handles   */
#define SMX_IO_E_ARG    -3   /* bad argument                           */
#define SMX_IO_E_TYPE   -4   /* sample type differs from the request   */

/* Sample types, with the numbering of SETSM's WriteGeotiff data_type
 * argument (setsmgeo.cpp:9-11). */
#define SMX_IO_UCHAR   1
#define SMX_IO_FLOAT   4
#define SMX_IO_UINT16 12

const char *smx_io_last_error(void);

/* What a TIFF file holds (first image of the file). */
typedef struct smx_tiff_info {
    uint32_t width, height;
    uint32_t bits;            /* BitsPerSample                                   */
    uint32_t sample_format;   /* 1 uint, 2 int, 3 IEEE float                     */
    uint32_t samples;         /* SamplesPerPixel                                 */
    uint32_t compression;     /* 1 none, 5 LZW                                   */
    uint32_t predictor;       /* 1 none, 2 horizontal, 3 floating point          */
    uint32_t tiled;           /* 1: TileWidth / TileLength, 0: strips            */
    uint32_t tile_w, tile_h;  /* tile size, or width x RowsPerStrip for strips   */
    uint32_t bigtiff;
    int32_t  has_geo;         /* ModelTiepoint and ModelPixelScale present       */
    double   min_x, max_y;    /* ModelTiepoint[3], [4]                           */
    double   grid_dx, grid_dy;/* ModelPixelScale[0], [1]                         */
} smx_tiff_info;

/* ReadGeotiff_info / ReadGeotiff_info_dxy (setsmgeo.cpp:150-197) and the
 * layout Readtiff_T needs: size, sample type, tiling and geo-referencing. */
int smx_tiff_info_read(const char *path, smx_tiff_info *info);

/* Readtiff_T (Template.hpp:211-360): the pixels of columns [cols[0],
 * cols[1]) and rows [rows[0], rows[1]) of a one-sample image, row-major, into
 * out (capacity out_bytes). type is SMX_IO_UCHAR / SMX_IO_UINT16 /
 * SMX_IO_FLOAT and must match the file's samples. As in the reference, a
 * tiled file widens the window to whole tiles (clipped to the image) and
 * cols / rows are updated to the window returned; a striped file returns
 * the window asked for. *out_w / *out_h receive its size. */
int smx_tiff_read(const char *path, int type, long cols[2], long rows[2], void *out, size_t out_bytes,
                  uint32_t *out_w, uint32_t *out_h);

/* WriteGeotiff / WriteGeotiff_round (setsmgeo.cpp:22-116): a one-sample
 * BigTIFF, LZW-compressed, predictor 3 (float), 2 (uint16) or 1 (uchar),
 * rows per strip as the reference sets them, ModelTiepoint (0, 0, 0, min_x,
 * max_y, 0), ModelPixelScale (scale, scale, 0), the reference's GeoKeys
 * (projection 1: polar stereographic, north if ns_hemisphere != 0; else UTM
 * zone `zone`, north or south) and GDAL_NODATA "-9999" for float images.
 * do_round rounds float values to the nearest 1/128 as WriteGeotiff_round
 * does. */
int smx_geotiff_write(const char *path, const void *buffer, size_t width, size_t height, double scale,
                      double min_x, double max_y, int projection, int zone, int ns_hemisphere, int data_type,
                      int do_round);

/* Image metadata of a DigitalGlobe .XML besides the RPC model. */
typedef struct smx_dg_info {
    double gsd_row, gsd_col, gsd;   /* MEANCOLLECTED* or MEANPRODUCT* GSD (last one in the file) */
    double abscalfactor, effbw, tdi;
} smx_dg_info;

/* OpenXMLFile, RPC branch (SubFunctions.cpp:1302-1582): the RPC00B model of
 * a DigitalGlobe .XML, one tag per line, coefficient lists on the line after
 * their <...List> tag, with the reference's lat / long swap of the offset
 * and scale rows (smx_rpc.off / scale = {line, samp, lon, lat, height}).
 * err receives ERRBIAS, ERRRAND (may be NULL); info may be NULL. */
int smx_dg_xml_read(const char *path, smx_rpc *rpc, double err[2], smx_dg_info *info);

/* The same model written as a DigitalGlobe-style .XML that
 * smx_dg_xml_read (and the reference's reader) parses back exactly
 * (17 significant digits). No counterpart in the reference (it only reads
 * vendor files); used to store synthetic scenes. */
int smx_dg_xml_write(const char *path, const smx_rpc *rpc, const double err[2], const smx_dg_info *info);

#ifdef __cplusplus
}
#endif
#endif /* SMX_IO_H */
