/* tiffcheck.c — test helper over libtiff (the library SETSM itself reads and
 * writes TIFF with, Template.hpp:211, setsmgeo.cpp:22): writes TIFF files in
 * the layouts a reader must handle and reads files back, so the formats code
 * of libsmx_io.so is checked against an independent implementation. Test
 * infrastructure only; built where libtiff's headers are installed. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <tiffio.h>

#define GEO_PIXELSCALE 33550
#define GEO_TIEPOINTS 33922
#define GEO_KEYDIR 34735
#define GDAL_NODATA 42113

static void quiet(const char *m, const char *f, va_list ap) { (void)m; (void)f; (void)ap; }

/* Writes a one-sample image. mode: "w" classic, "w8" BigTIFF, "wb" big-endian. */
int tc_write(const char *path, const char *mode, const void *buf, uint32_t w, uint32_t h, int bits, int fmt,
             int compression, int predictor, int tiled, uint32_t tw, uint32_t th, uint32_t rps)
{
    TIFF *t = TIFFOpen(path, mode);
    if (!t) return -1;
    const int bps = bits / 8;
    TIFFSetField(t, TIFFTAG_IMAGEWIDTH, w);
    TIFFSetField(t, TIFFTAG_IMAGELENGTH, h);
    TIFFSetField(t, TIFFTAG_BITSPERSAMPLE, bits);
    TIFFSetField(t, TIFFTAG_SAMPLEFORMAT, fmt);
    TIFFSetField(t, TIFFTAG_SAMPLESPERPIXEL, 1);
    TIFFSetField(t, TIFFTAG_PHOTOMETRIC, PHOTOMETRIC_MINISBLACK);
    TIFFSetField(t, TIFFTAG_PLANARCONFIG, PLANARCONFIG_CONTIG);
    TIFFSetField(t, TIFFTAG_COMPRESSION, compression);
    if (compression != COMPRESSION_NONE) TIFFSetField(t, TIFFTAG_PREDICTOR, predictor);
    int rc = 0;
    if (tiled) {
        TIFFSetField(t, TIFFTAG_TILEWIDTH, tw);
        TIFFSetField(t, TIFFTAG_TILELENGTH, th);
        uint8_t *tile = (uint8_t *)malloc((size_t)tw * th * bps);
        for (uint32_t y = 0; y < h && !rc; y += th)
            for (uint32_t x = 0; x < w && !rc; x += tw) {
                memset(tile, 0, (size_t)tw * th * bps);
                for (uint32_t r = 0; r < th && y + r < h; r++) {
                    const uint32_t n = (x + tw <= w) ? tw : w - x;
                    memcpy(tile + (size_t)r * tw * bps, (const uint8_t *)buf + ((size_t)(y + r) * w + x) * bps,
                           (size_t)n * bps);
                }
                if (TIFFWriteTile(t, tile, x, y, 0, 0) < 0) rc = -2;
            }
        free(tile);
    } else {
        TIFFSetField(t, TIFFTAG_ROWSPERSTRIP, rps);
        uint8_t *row = (uint8_t *)malloc((size_t)w * bps);
        for (uint32_t y = 0; y < h && !rc; y++) {
            memcpy(row, (const uint8_t *)buf + (size_t)y * w * bps, (size_t)w * bps);
            if (TIFFWriteScanline(t, row, y, 0) < 0) rc = -2;
        }
        free(row);
    }
    TIFFClose(t);
    return rc;
}

/* Tags libtiff does not know are read as anonymous fields, whose value comes
 * with a count (16- or 32-bit). */
static int get_array(TIFF *t, uint32_t tag, uint32_t *count, void **data)
{
    const TIFFField *f = TIFFFindField(t, tag, TIFF_ANY);
    if (!f) return 0;
    if (!TIFFFieldPassCount(f)) {
        *count = 1;
        return TIFFGetField(t, tag, data);
    }
    if (TIFFFieldReadCount(f) == TIFF_VARIABLE2) return TIFFGetField(t, tag, count, data);
    uint16_t c16 = 0;
    const int r = TIFFGetField(t, tag, &c16, data);
    *count = c16;
    return r;
}

/* Reads a one-sample image written by anyone, and the tags a GeoTIFF of
 * SETSM carries. info[0..7] = width, height, bits, format, compression,
 * predictor, rows per strip (0 if tiled), flags (1 BigTIFF, 2 big-endian). */
int tc_read(const char *path, void *out, size_t cap, uint32_t info[8], double tie[6], double scale[3],
            char *nodata, int nodata_cap, uint16_t *keys, int keys_cap)
{
    TIFFSetWarningHandler(quiet);
    TIFF *t = TIFFOpen(path, "r");
    if (!t) return -1;
    uint32_t w = 0, h = 0, rps = 0;
    uint16_t bits = 0, fmt = 1, comp = 1, pred = 1;
    TIFFGetField(t, TIFFTAG_IMAGEWIDTH, &w);
    TIFFGetField(t, TIFFTAG_IMAGELENGTH, &h);
    TIFFGetField(t, TIFFTAG_BITSPERSAMPLE, &bits);
    TIFFGetFieldDefaulted(t, TIFFTAG_SAMPLEFORMAT, &fmt);
    TIFFGetFieldDefaulted(t, TIFFTAG_COMPRESSION, &comp);
    TIFFGetFieldDefaulted(t, TIFFTAG_PREDICTOR, &pred);
    const int tiled = TIFFIsTiled(t);
    if (!tiled) TIFFGetFieldDefaulted(t, TIFFTAG_ROWSPERSTRIP, &rps);
    info[0] = w; info[1] = h; info[2] = bits; info[3] = fmt; info[4] = comp; info[5] = pred;
    info[6] = rps; info[7] = TIFFIsBigEndian(t) ? 2 : 0;
    FILE *fh = fopen(path, "rb");
    unsigned char hb[4] = {0, 0, 0, 0};
    if (fh) {
        if (fread(hb, 1, 4, fh) != 4) hb[2] = hb[3] = 0;
        fclose(fh);
    }
    if ((hb[0] == 'I' && hb[2] == 43) || (hb[0] == 'M' && hb[3] == 43)) info[7] |= 1;
    const int bps = bits / 8;
    if ((size_t)w * h * bps > cap) { TIFFClose(t); return -3; }
    int rc = 0;
    if (tiled) {
        uint32_t tw = 0, th = 0;
        TIFFGetField(t, TIFFTAG_TILEWIDTH, &tw);
        TIFFGetField(t, TIFFTAG_TILELENGTH, &th);
        uint8_t *tile = (uint8_t *)_TIFFmalloc(TIFFTileSize(t));
        for (uint32_t y = 0; y < h && !rc; y += th)
            for (uint32_t x = 0; x < w && !rc; x += tw) {
                if (TIFFReadTile(t, tile, x, y, 0, 0) < 0) { rc = -2; break; }
                for (uint32_t r = 0; r < th && y + r < h; r++) {
                    const uint32_t n = (x + tw <= w) ? tw : w - x;
                    memcpy((uint8_t *)out + ((size_t)(y + r) * w + x) * bps, tile + (size_t)r * tw * bps,
                           (size_t)n * bps);
                }
            }
        _TIFFfree(tile);
    } else {
        for (uint32_t y = 0; y < h && !rc; y++)
            if (TIFFReadScanline(t, (uint8_t *)out + (size_t)y * w * bps, y, 0) < 0) rc = -2;
    }
    uint32_t cnt = 0;
    memset(tie, 0, 6 * sizeof(double));
    memset(scale, 0, 3 * sizeof(double));
    nodata[0] = 0;
    memset(keys, 0, (size_t)keys_cap * sizeof(uint16_t));
    double *d = NULL;
    if (get_array(t, GEO_TIEPOINTS, &cnt, (void **)&d) && d)
        for (uint32_t i = 0; i < cnt && i < 6; i++) tie[i] = d[i];
    d = NULL;
    if (get_array(t, GEO_PIXELSCALE, &cnt, (void **)&d) && d)
        for (uint32_t i = 0; i < cnt && i < 3; i++) scale[i] = d[i];
    char *s = NULL;
    if (get_array(t, GDAL_NODATA, &cnt, (void **)&s) && s) snprintf(nodata, nodata_cap, "%s", s);
    uint16_t *k = NULL;
    if (get_array(t, GEO_KEYDIR, &cnt, (void **)&k) && k)
        for (uint32_t i = 0; i < cnt && (int)i < keys_cap; i++) keys[i] = k[i];
    TIFFClose(t);
    return rc;
}
