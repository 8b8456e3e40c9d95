// smx_io.cpp — host-side formats of the matching path (libsmx_io.so):
// TIFF / BigTIFF reading (strips or tiles, uncompressed or LZW, predictors
// 1/2/3), SETSM's GeoTIFF writer, and the DigitalGlobe XML RPC reader.
//
// Written from the TIFF 6.0 specification (LZW: section 13, predictor 2:
// section 14), the BigTIFF layout, Adobe's TIFF technical note 3 (floating
// point predictor) and the GeoTIFF 1.0 key directory; the reference reaches
// the same formats through libtiff / libgeotiff (Template.hpp:211,
// setsmgeo.cpp:22-116) and its own line parser (SubFunctions.cpp:1302-1582).
#include "smx_io.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

namespace {

thread_local char g_err[512];

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

struct File {
    FILE *f = nullptr;
    explicit File(const char *path, const char *mode) : f(fopen(path, mode)) {}
    ~File()
    {
        if (f) fclose(f);
    }
    File(const File &) = delete;
    File &operator=(const File &) = delete;
};

// ------------------------------------------------------------------ TIFF

enum : uint16_t {
    T_BYTE = 1, T_ASCII = 2, T_SHORT = 3, T_LONG = 4, T_RATIONAL = 5, T_SBYTE = 6, T_UNDEF = 7, T_SSHORT = 8,
    T_SLONG = 9, T_SRATIONAL = 10, T_FLOAT = 11, T_DOUBLE = 12, T_LONG8 = 16, T_SLONG8 = 17, T_IFD8 = 18
};

int type_size(uint16_t t)
{
    switch (t) {
    case T_BYTE: case T_ASCII: case T_SBYTE: case T_UNDEF: return 1;
    case T_SHORT: case T_SSHORT: return 2;
    case T_LONG: case T_SLONG: case T_FLOAT: return 4;
    case T_RATIONAL: case T_SRATIONAL: case T_DOUBLE: case T_LONG8: case T_SLONG8: case T_IFD8: return 8;
    default: return 0;
    }
}

struct Entry {
    uint16_t type = 0;
    uint64_t count = 0;
    std::vector<uint8_t> raw;   // file byte order
};

struct Tiff {
    bool le = true, big = false;
    std::map<uint16_t, Entry> tags;

    uint64_t get(const uint8_t *p, int n) const
    {
        uint64_t v = 0;
        if (le)
            for (int i = n - 1; i >= 0; i--) v = (v << 8) | p[i];
        else
            for (int i = 0; i < n; i++) v = (v << 8) | p[i];
        return v;
    }
    bool has(uint16_t tag) const { return tags.count(tag) != 0; }
    // element k of an integer tag
    uint64_t u(uint16_t tag, uint64_t k = 0, uint64_t dflt = 0) const
    {
        auto it = tags.find(tag);
        if (it == tags.end() || k >= it->second.count) return dflt;
        const Entry &e = it->second;
        const int s = type_size(e.type);
        return get(e.raw.data() + k * s, s);
    }
    double d(uint16_t tag, uint64_t k = 0) const
    {
        auto it = tags.find(tag);
        if (it == tags.end() || k >= it->second.count) return 0.0;
        const Entry &e = it->second;
        if (e.type == T_DOUBLE) {
            const uint64_t b = get(e.raw.data() + 8 * k, 8);
            double v;
            memcpy(&v, &b, 8);
            return v;
        }
        if (e.type == T_FLOAT) {
            const uint32_t b = (uint32_t)get(e.raw.data() + 4 * k, 4);
            float v;
            memcpy(&v, &b, 4);
            return v;
        }
        return (double)u(tag, k);
    }
};

int read_at(FILE *f, uint64_t off, void *dst, size_t n)
{
    if (fseeko(f, (off_t)off, SEEK_SET) != 0) return -1;
    return fread(dst, 1, n, f) == n ? 0 : -1;
}

int parse_tiff(FILE *f, Tiff &t)
{
    uint8_t h[16];
    if (read_at(f, 0, h, 8)) return fail(SMX_IO_E_FORMAT, "short file");
    if (h[0] == 'I' && h[1] == 'I') t.le = true;
    else if (h[0] == 'M' && h[1] == 'M') t.le = false;
    else return fail(SMX_IO_E_FORMAT, "not a TIFF file");
    const uint64_t magic = t.get(h + 2, 2);
    uint64_t ifd;
    if (magic == 42) {
        ifd = t.get(h + 4, 4);
    } else if (magic == 43) {
        t.big = true;
        if (read_at(f, 0, h, 16)) return fail(SMX_IO_E_FORMAT, "short BigTIFF header");
        if (t.get(h + 4, 2) != 8) return fail(SMX_IO_E_FORMAT, "BigTIFF offset size %d", (int)t.get(h + 4, 2));
        ifd = t.get(h + 8, 8);
    } else {
        return fail(SMX_IO_E_FORMAT, "TIFF magic %d", (int)magic);
    }
    const int cn = t.big ? 8 : 2, esz = t.big ? 20 : 12, vsz = t.big ? 8 : 4;
    uint8_t nb[8];
    if (read_at(f, ifd, nb, cn)) return fail(SMX_IO_E_FORMAT, "bad IFD offset");
    const uint64_t n = t.get(nb, cn);
    if (n == 0 || n > 4096) return fail(SMX_IO_E_FORMAT, "IFD with %llu entries", (unsigned long long)n);
    std::vector<uint8_t> ents(n * esz);
    if (read_at(f, ifd + cn, ents.data(), ents.size())) return fail(SMX_IO_E_FORMAT, "truncated IFD");
    for (uint64_t i = 0; i < n; i++) {
        const uint8_t *e = ents.data() + i * esz;
        Entry en;
        const uint16_t tag = (uint16_t)t.get(e, 2);
        en.type = (uint16_t)t.get(e + 2, 2);
        en.count = t.get(e + 4, t.big ? 8 : 4);
        const int s = type_size(en.type);
        if (s == 0) continue;   // unknown type: skipped, as TIFF readers must
        const uint64_t bytes = en.count * (uint64_t)s;
        if (bytes > ((uint64_t)1 << 31)) return fail(SMX_IO_E_FORMAT, "tag %d too large", tag);
        en.raw.resize(bytes);
        const uint8_t *v = e + (t.big ? 12 : 8);
        if (bytes <= (uint64_t)vsz) {
            memcpy(en.raw.data(), v, bytes);
        } else if (read_at(f, t.get(v, vsz), en.raw.data(), bytes)) {
            return fail(SMX_IO_E_FORMAT, "tag %d values outside the file", tag);
        }
        t.tags[tag] = std::move(en);
    }
    return 0;
}

// LZW (TIFF 6.0 section 13): MSB-first codes of 9..12 bits, 256 = Clear,
// 257 = EndOfInformation, the code width grows one code early.
int lzw_decode(const uint8_t *in, size_t n, uint8_t *out, size_t cap, size_t &got)
{
    struct Ent {
        int prefix;    // previous code, -1 for a single byte
        uint8_t last;  // last byte of the string
        uint8_t first; // first byte of the string
        int len;
    };
    std::vector<Ent> tab(4096);
    for (int i = 0; i < 256; i++) tab[i] = Ent{-1, (uint8_t)i, (uint8_t)i, 1};
    int next = 258, width = 9, prev = -1;
    size_t bitpos = 0, o = 0;
    const size_t nbits = n * 8;
    auto emit = [&](int code) -> bool {
        const int len = tab[code].len;
        if (o + (size_t)len > cap) return false;
        int c = code;
        for (int k = len - 1; k >= 0; k--) {
            out[o + k] = tab[c].last;
            c = tab[c].prefix;
        }
        o += len;
        return true;
    };
    for (;;) {
        if (bitpos + width > nbits) break;   // ran out of input without EOI: accept what was decoded
        int code = 0;
        for (int k = 0; k < width; k++) {
            const size_t b = bitpos + k;
            code = (code << 1) | ((in[b >> 3] >> (7 - (b & 7))) & 1);
        }
        bitpos += width;
        if (code == 257) break;
        if (code == 256) {
            next = 258;
            width = 9;
            prev = -1;
            continue;
        }
        if (prev < 0) {
            if (code > 255) return fail(SMX_IO_E_FORMAT, "LZW: code %d after Clear", code);
            if (!emit(code)) return fail(SMX_IO_E_FORMAT, "LZW: output overflow");
            prev = code;
            continue;
        }
        if (code < next) {
            if (!emit(code)) return fail(SMX_IO_E_FORMAT, "LZW: output overflow");
            if (next < 4096) tab[next] = Ent{prev, tab[code].first, tab[prev].first, tab[prev].len + 1};
        } else if (code == next && next < 4096) {
            tab[next] = Ent{prev, tab[prev].first, tab[prev].first, tab[prev].len + 1};
            if (!emit(code)) return fail(SMX_IO_E_FORMAT, "LZW: output overflow");
        } else {
            return fail(SMX_IO_E_FORMAT, "LZW: code %d beyond the table (%d)", code, next);
        }
        if (next < 4096) next++;
        if (next + 1 >= (1 << width) && width < 12) width++;
        prev = code;
    }
    got = o;
    return 0;
}

struct BitWriter {
    std::vector<uint8_t> &out;
    uint64_t acc = 0;
    int nacc = 0;
    explicit BitWriter(std::vector<uint8_t> &o) : out(o) {}
    void put(int code, int width)
    {
        acc = (acc << width) | (uint64_t)code;
        nacc += width;
        while (nacc >= 8) {
            out.push_back((uint8_t)(acc >> (nacc - 8)));
            nacc -= 8;
        }
    }
    void flush()
    {
        if (nacc > 0) out.push_back((uint8_t)(acc << (8 - nacc)));
        nacc = 0;
        acc = 0;
    }
};

void lzw_encode(const uint8_t *in, size_t n, std::vector<uint8_t> &out)
{
    // string table as (prefix code, byte) -> code, open addressing
    const int kSlots = 1 << 14;
    std::vector<int32_t> key(kSlots), val(kSlots);
    auto reset = [&]() { std::fill(key.begin(), key.end(), -1); };
    auto slot = [&](int prefix, int byte) {
        const int32_t k = (prefix << 8) | byte;
        uint32_t h = (uint32_t)k * 2654435761u >> 18;
        while (key[h] != -1 && key[h] != k) h = (h + 1) & (kSlots - 1);
        return h;
    };
    BitWriter bw(out);
    reset();
    int width = 9, next = 258;
    bw.put(256, width);
    if (n == 0) {
        bw.put(257, width);
        bw.flush();
        return;
    }
    int w = in[0];
    for (size_t i = 1; i < n; i++) {
        const int c = in[i];
        const uint32_t h = slot(w, c);
        if (key[h] != -1) {
            w = val[h];
            continue;
        }
        bw.put(w, width);
        key[h] = (w << 8) | c;
        val[h] = next++;
        if (next > (1 << width) - 1 && width < 12) width++;
        if (next >= 4094) {   // table full: Clear at the current width, start again
            bw.put(256, width);
            reset();
            width = 9;
            next = 258;
        }
        w = c;
    }
    bw.put(w, width);
    next++;
    if (next > (1 << width) - 1 && width < 12) width++;
    bw.put(257, width);
    bw.flush();
}

int sample_bytes(int type) { return type == SMX_IO_UCHAR ? 1 : type == SMX_IO_UINT16 ? 2 : type == SMX_IO_FLOAT ? 4 : 0; }

// Undo predictor p on one decoded row of `w` samples of `bps` bytes: p == 2
// on samples already in host (little-endian) order; p == 3 on the row as
// stored (byte planes, most significant first, whatever the file's byte
// order), leaving little-endian samples.
int unpredict_row(uint8_t *row, uint32_t w, int bps, int p)
{
    if (p == 1) return 0;
    if (p == 2) {
        if (bps == 1) {
            for (uint32_t i = 1; i < w; i++) row[i] = (uint8_t)(row[i] + row[i - 1]);
        } else if (bps == 2) {
            uint16_t *r = (uint16_t *)row;
            for (uint32_t i = 1; i < w; i++) r[i] = (uint16_t)(r[i] + r[i - 1]);
        } else if (bps == 4) {
            uint32_t *r = (uint32_t *)row;
            for (uint32_t i = 1; i < w; i++) r[i] = r[i] + r[i - 1];
        } else {
            return fail(SMX_IO_E_FORMAT, "predictor 2 on %d-byte samples", bps);
        }
        return 0;
    }
    if (p == 3) {
        // bytes were differenced over the row, then stored as planes of
        // equal significance, most significant plane first
        const size_t nb = (size_t)w * bps;
        for (size_t i = 1; i < nb; i++) row[i] = (uint8_t)(row[i] + row[i - 1]);
        std::vector<uint8_t> tmp(row, row + nb);
        for (uint32_t i = 0; i < w; i++)
            for (int b = 0; b < bps; b++)   // plane b: byte b of the big-endian form
                row[(size_t)i * bps + (bps - 1 - b)] = tmp[(size_t)b * w + i];
        return 0;
    }
    return fail(SMX_IO_E_FORMAT, "predictor %d", p);
}

void swap_bytes(uint8_t *p, size_t n, int bps)
{
    if (bps <= 1) return;
    for (size_t i = 0; i < n; i += bps) std::reverse(p + i, p + i + bps);
}

int check_layout(const Tiff &t, smx_tiff_info &I)
{
    I = smx_tiff_info{};
    I.width = (uint32_t)t.u(256);
    I.height = (uint32_t)t.u(257);
    I.bits = (uint32_t)t.u(258, 0, 1);
    I.sample_format = (uint32_t)t.u(339, 0, 1);
    I.samples = (uint32_t)t.u(277, 0, 1);
    I.compression = (uint32_t)t.u(259, 0, 1);
    I.predictor = (uint32_t)t.u(317, 0, 1);
    I.tiled = t.has(322) ? 1 : 0;
    if (I.tiled) {
        I.tile_w = (uint32_t)t.u(322);
        I.tile_h = (uint32_t)t.u(323);
    } else {
        I.tile_w = I.width;
        I.tile_h = (uint32_t)std::min<uint64_t>(t.u(278, 0, 0xffffffffu), I.height);
    }
    I.bigtiff = t.big ? 1 : 0;
    I.has_geo = t.has(33922) && t.has(33550);
    if (t.has(33922)) {
        I.min_x = t.d(33922, 3);
        I.max_y = t.d(33922, 4);
    }
    if (t.has(33550)) {
        I.grid_dx = t.d(33550, 0);
        I.grid_dy = t.d(33550, 1);
    }
    if (I.width == 0 || I.height == 0) return fail(SMX_IO_E_FORMAT, "image without size");
    return 0;
}

int sample_type_of(const smx_tiff_info &I)
{
    if (I.bits == 8 && I.sample_format == 1) return SMX_IO_UCHAR;
    if (I.bits == 16 && I.sample_format == 1) return SMX_IO_UINT16;
    if (I.bits == 32 && I.sample_format == 3) return SMX_IO_FLOAT;
    return 0;
}

// Decode chunk k (strip or tile) into `dst` (rows x cols samples of bps
// bytes, host byte order).
int read_chunk(FILE *f, const Tiff &t, const smx_tiff_info &I, uint64_t k, uint32_t rows, uint32_t cols, int bps,
               std::vector<uint8_t> &dst)
{
    const uint16_t off_tag = I.tiled ? 324 : 273, cnt_tag = I.tiled ? 325 : 279;
    const uint64_t off = t.u(off_tag, k), cnt = t.u(cnt_tag, k);
    const size_t need = (size_t)rows * cols * bps;
    dst.assign(need, 0);
    if (cnt == 0) return 0;   // sparse chunk: zeros
    std::vector<uint8_t> raw(cnt);
    if (read_at(f, off, raw.data(), cnt)) return fail(SMX_IO_E_FORMAT, "chunk %llu outside the file", (unsigned long long)k);
    if (I.compression == 1) {
        memcpy(dst.data(), raw.data(), std::min<size_t>(need, raw.size()));
    } else if (I.compression == 5) {
        size_t got = 0;
        int r = lzw_decode(raw.data(), raw.size(), dst.data(), need, got);
        if (r) return r;
    } else {
        return fail(SMX_IO_E_FORMAT, "compression %u not supported", I.compression);
    }
    const bool host_le = true;
    if (I.predictor == 3) {
        for (uint32_t r = 0; r < rows; r++) {
            int e = unpredict_row(dst.data() + (size_t)r * cols * bps, cols, bps, 3);
            if (e) return e;
        }
        return 0;   // planes are most-significant first: the rows are now little-endian
    }
    if (t.le != host_le) swap_bytes(dst.data(), need, bps);
    if (I.predictor == 2)
        for (uint32_t r = 0; r < rows; r++) {
            int e = unpredict_row(dst.data() + (size_t)r * cols * bps, cols, bps, 2);
            if (e) return e;
        }
    else if (I.predictor != 1)
        return fail(SMX_IO_E_FORMAT, "predictor %u", I.predictor);
    return 0;
}

// ----------------------------------------------------------- TIFF writer

struct OutTag {
    uint16_t tag, type;
    uint64_t count;
    std::vector<uint8_t> data;   // little-endian
};

template <class T>
void put_le(std::vector<uint8_t> &v, T x, int n = sizeof(T))
{
    uint64_t u = 0;
    memcpy(&u, &x, sizeof(T));
    for (int i = 0; i < n; i++) v.push_back((uint8_t)(u >> (8 * i)));
}

OutTag tag_shorts(uint16_t tag, const std::vector<uint16_t> &vals)
{
    OutTag t{tag, T_SHORT, vals.size(), {}};
    for (uint16_t x : vals) put_le(t.data, x);
    return t;
}
OutTag tag_long(uint16_t tag, uint32_t v)
{
    OutTag t{tag, T_LONG, 1, {}};
    put_le(t.data, v);
    return t;
}
OutTag tag_long8s(uint16_t tag, const std::vector<uint64_t> &vals)
{
    OutTag t{tag, T_LONG8, vals.size(), {}};
    for (uint64_t x : vals) put_le(t.data, x);
    return t;
}
OutTag tag_doubles(uint16_t tag, const std::vector<double> &vals)
{
    OutTag t{tag, T_DOUBLE, vals.size(), {}};
    for (double x : vals) put_le(t.data, x);
    return t;
}
OutTag tag_ascii(uint16_t tag, const std::string &s)
{
    OutTag t{tag, T_ASCII, s.size() + 1, {}};
    t.data.assign(s.begin(), s.end());
    t.data.push_back(0);
    return t;
}

// GeoKeys of SetUpGeoKeys (setsmgeo.cpp:251-307), key directory order.
void geo_keys(int projection, int zone, int north, std::vector<uint16_t> &dir, std::vector<double> &dbl,
              std::string &asc)
{
    struct Key {
        uint16_t id;
        int kind;   // 0 short, 1 double, 2 ascii
        uint16_t s;
        double d;
        std::string a;
    };
    std::vector<Key> keys;
    keys.push_back({1024, 0, 1, 0, ""});          // GTModelType = ModelProjected
    keys.push_back({1025, 0, 1, 0, ""});          // GTRasterType = RasterPixelIsArea
    keys.push_back({2049, 2, 0, 0, "WGS 84"});    // GeogCitation
    keys.push_back({2054, 0, 9102, 0, ""});       // GeogAngularUnits = degree
    keys.push_back({3076, 0, 9001, 0, ""});       // ProjLinearUnits = metre
    if (projection == 1) {
        keys.push_back({1026, 2, 0, 0, "unnamed"});   // GTCitation
        keys.push_back({2048, 0, 4326, 0, ""});       // GeographicType = WGS 84
        keys.push_back({2057, 1, 0, 6378137.0, ""});  // semi-major axis
        keys.push_back({2059, 1, 0, 298.257223563, ""});
        keys.push_back({3072, 0, 32767, 0, ""});      // ProjectedCSType = user-defined
        keys.push_back({3074, 0, 32767, 0, ""});      // Projection = user-defined
        keys.push_back({3075, 0, 15, 0, ""});         // CT_PolarStereographic
        keys.push_back({3082, 1, 0, 0.0, ""});        // false easting
        keys.push_back({3083, 1, 0, 0.0, ""});        // false northing
        keys.push_back({3092, 1, 0, 1.0, ""});        // scale at natural origin
        keys.push_back({3081, 1, 0, north ? 70.0 : -71.0, ""});   // natural origin latitude
        keys.push_back({3095, 1, 0, north ? -45.0 : 0.0, ""});    // straight vertical pole longitude
    } else {
        char buf[96];
        snprintf(buf, sizeof(buf), "UTM Zone %d, %s Hemisphere", zone, north ? "Northern" : "Southern");
        keys.push_back({1026, 2, 0, 0, buf});
        keys.push_back({3072, 0, (uint16_t)((north ? 32601 : 32701) - 1 + zone), 0, ""});
    }
    std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) { return a.id < b.id; });
    dir = {1, 1, 0, (uint16_t)keys.size()};
    for (const Key &k : keys) {
        if (k.kind == 0) {
            dir.insert(dir.end(), {k.id, 0, 1, k.s});
        } else if (k.kind == 1) {
            dir.insert(dir.end(), {k.id, 34736, 1, (uint16_t)dbl.size()});
            dbl.push_back(k.d);
        } else {
            const std::string s = k.a + "|";
            dir.insert(dir.end(), {k.id, 34737, (uint16_t)s.size(), (uint16_t)asc.size()});
            asc += s;
        }
    }
}

}  // namespace

extern "C" {

const char *smx_io_last_error(void) { return g_err; }

int smx_tiff_info_read(const char *path, smx_tiff_info *info)
{
    if (!path || !info) return fail(SMX_IO_E_ARG, "smx_tiff_info_read: null argument");
    File F(path, "rb");
    if (!F.f) return fail(SMX_IO_E_OPEN, "cannot open %s", path);
    Tiff t;
    int r = parse_tiff(F.f, t);
    if (!r) r = check_layout(t, *info);
    return r;
}

int smx_tiff_read(const char *path, int type, long cols[2], long rows[2], void *out, size_t out_bytes,
                  uint32_t *out_w, uint32_t *out_h)
{
    if (!path || !cols || !rows || !out) return fail(SMX_IO_E_ARG, "smx_tiff_read: null argument");
    const int bps = sample_bytes(type);
    if (!bps) return fail(SMX_IO_E_ARG, "smx_tiff_read: type %d", type);
    File F(path, "rb");
    if (!F.f) return fail(SMX_IO_E_OPEN, "cannot open %s", path);
    Tiff t;
    smx_tiff_info I;
    int r = parse_tiff(F.f, t);
    if (!r) r = check_layout(t, I);
    if (r) return r;
    if (I.samples != 1) return fail(SMX_IO_E_FORMAT, "%u samples per pixel (one expected)", I.samples);
    if (t.u(284, 0, 1) != 1 && I.samples != 1) return fail(SMX_IO_E_FORMAT, "planar configuration");
    if (sample_type_of(I) != type)
        return fail(SMX_IO_E_TYPE, "file holds %u-bit samples of format %u, type %d asked", I.bits, I.sample_format,
                    type);
    if (cols[0] < 0 || rows[0] < 0 || cols[1] <= cols[0] || rows[1] <= rows[0] || cols[1] > (long)I.width ||
        rows[1] > (long)I.height)
        return fail(SMX_IO_E_ARG, "window [%ld,%ld) x [%ld,%ld) outside %u x %u", cols[0], cols[1], rows[0], rows[1],
                    I.width, I.height);
    long c0 = cols[0], c1 = cols[1], r0 = rows[0], r1 = rows[1];
    if (I.tiled) {
        // Template.hpp:303-322: whole tiles, the end one tile past rows[1] /
        // cols[1] as the reference computes it, clipped to the image
        c0 = (c0 / (long)I.tile_w) * I.tile_w;
        c1 = std::min<long>((c1 / (long)I.tile_w + 1) * I.tile_w, I.width);
        r0 = (r0 / (long)I.tile_h) * I.tile_h;
        r1 = std::min<long>((r1 / (long)I.tile_h + 1) * I.tile_h, I.height);
    }
    const size_t W = (size_t)(c1 - c0), Hh = (size_t)(r1 - r0);
    if (W * Hh * bps > out_bytes) return fail(SMX_IO_E_ARG, "output holds %zu bytes, window needs %zu", out_bytes,
                                              W * Hh * bps);
    uint8_t *o = (uint8_t *)out;
    std::vector<uint8_t> chunk;
    const uint32_t tw = I.tile_w, th = I.tile_h;
    const uint64_t across = I.tiled ? (I.width + tw - 1) / tw : 1;
    for (long cr = r0 / th; cr * (long)th < r1; cr++) {
        for (long cc = c0 / tw; cc * (long)tw < c1; cc++) {
            const uint64_t k = (uint64_t)cr * across + (uint64_t)cc;
            // a strip's last rows may be missing from the file
            const uint32_t crows = I.tiled ? th : (uint32_t)std::min<long>(th, (long)I.height - cr * (long)th);
            r = read_chunk(F.f, t, I, k, crows, tw, bps, chunk);
            if (r) return r;
            for (uint32_t y = 0; y < crows; y++) {
                const long gy = cr * (long)th + y;
                if (gy < r0 || gy >= r1) continue;
                const long x0 = std::max<long>(c0, cc * (long)tw), x1 = std::min<long>(c1, (cc + 1) * (long)tw);
                if (x1 <= x0) continue;
                memcpy(o + ((size_t)(gy - r0) * W + (size_t)(x0 - c0)) * bps,
                       chunk.data() + ((size_t)y * tw + (size_t)(x0 - cc * (long)tw)) * bps, (size_t)(x1 - x0) * bps);
            }
        }
    }
    cols[0] = c0;
    cols[1] = c1;
    rows[0] = r0;
    rows[1] = r1;
    if (out_w) *out_w = (uint32_t)W;
    if (out_h) *out_h = (uint32_t)Hh;
    return 0;
}

int smx_geotiff_write(const char *path, const void *buffer, size_t width, size_t height, double scale, double min_x,
                      double max_y, int projection, int zone, int ns_hemisphere, int data_type, int do_round)
{
    if (!path || !buffer || width == 0 || height == 0) return fail(SMX_IO_E_ARG, "smx_geotiff_write: bad argument");
    const int bps = sample_bytes(data_type);
    if (!bps) return fail(SMX_IO_E_ARG, "unrecognized data type: %d", data_type);
    // setsmgeo.cpp:222-246: predictor and rows per strip by type
    const size_t kStrip = 8192;
    int predictor, bits, fmt;
    size_t rps;
    if (data_type == SMX_IO_FLOAT) {
        predictor = 3, bits = 32, fmt = 3;
        rps = std::max<size_t>(1, (kStrip * 8) / (width * 32));
    } else if (data_type == SMX_IO_UCHAR) {
        predictor = 1, bits = 8, fmt = 1;
        rps = std::max<size_t>(1, (kStrip * 8) / (width * 8));
    } else {
        predictor = 2, bits = 16, fmt = 1;
        rps = std::max<size_t>(1, (kStrip * 8) / (width * 16));
    }
    rps = std::min(rps, height);
    File F(path, "wb");
    if (!F.f) return fail(SMX_IO_E_OPEN, "cannot create %s", path);
    // BigTIFF header ("w8" in XTIFFOpen), the IFD goes after the strips
    std::vector<uint8_t> hdr = {'I', 'I', 43, 0, 8, 0, 0, 0};
    put_le(hdr, (uint64_t)0);
    if (fwrite(hdr.data(), 1, hdr.size(), F.f) != hdr.size()) return fail(SMX_IO_E_OPEN, "write failed");
    uint64_t pos = hdr.size();
    std::vector<uint64_t> offs, cnts;
    std::vector<uint8_t> strip, enc;
    const uint8_t *src = (const uint8_t *)buffer;
    for (size_t r0 = 0; r0 < height; r0 += rps) {
        const size_t nr = std::min(rps, height - r0);
        strip.assign(src + r0 * width * bps, src + (r0 + nr) * width * bps);
        for (size_t y = 0; y < nr; y++) {
            uint8_t *row = strip.data() + y * width * bps;
            if (do_round && data_type == SMX_IO_FLOAT) {
                float *fr = (float *)row;
                for (size_t i = 0; i < width; i++) fr[i] = std::round(fr[i] * 128) / 128;
            }
            if (predictor == 2) {
                uint16_t *s = (uint16_t *)row;
                for (size_t i = width; i-- > 1;) s[i] = (uint16_t)(s[i] - s[i - 1]);
            } else if (predictor == 3) {
                // planes of equal significance, most significant first, then
                // byte differences along the row
                std::vector<uint8_t> tmp(row, row + width * bps);
                for (size_t i = 0; i < width; i++)
                    for (int b = 0; b < bps; b++) row[(size_t)b * width + i] = tmp[i * bps + (bps - 1 - b)];
                for (size_t i = width * bps; i-- > 1;) row[i] = (uint8_t)(row[i] - row[i - 1]);
            }
        }
        enc.clear();
        lzw_encode(strip.data(), strip.size(), enc);
        if (fwrite(enc.data(), 1, enc.size(), F.f) != enc.size()) return fail(SMX_IO_E_OPEN, "write failed");
        offs.push_back(pos);
        cnts.push_back(enc.size());
        pos += enc.size();
    }
    if (pos & 1) {   // word-align the IFD
        fputc(0, F.f);
        pos++;
    }
    std::vector<uint16_t> gdir;
    std::vector<double> gdbl;
    std::string gasc;
    geo_keys(projection, zone, ns_hemisphere, gdir, gdbl, gasc);
    std::vector<OutTag> tags;
    tags.push_back(tag_long(256, (uint32_t)width));
    tags.push_back(tag_long(257, (uint32_t)height));
    tags.push_back(tag_shorts(258, {(uint16_t)bits}));
    tags.push_back(tag_shorts(259, {5}));   // LZW
    tags.push_back(tag_shorts(262, {1}));   // min-is-black
    tags.push_back(tag_long8s(273, offs));
    tags.push_back(tag_shorts(277, {1}));
    tags.push_back(tag_long(278, (uint32_t)rps));
    tags.push_back(tag_long8s(279, cnts));
    tags.push_back(tag_shorts(284, {1}));   // contiguous
    tags.push_back(tag_shorts(317, {(uint16_t)predictor}));
    tags.push_back(tag_shorts(339, {(uint16_t)fmt}));
    tags.push_back(tag_doubles(33550, {scale, scale, 0.0}));
    tags.push_back(tag_doubles(33922, {0.0, 0.0, 0.0, min_x, max_y, 0.0}));
    tags.push_back(tag_shorts(34735, gdir));
    if (!gdbl.empty()) tags.push_back(tag_doubles(34736, gdbl));
    tags.push_back(tag_ascii(34737, gasc));
    if (data_type == SMX_IO_FLOAT) tags.push_back(tag_ascii(42113, "-9999"));
    // IFD: count, 20-byte entries, next offset; long values after it
    const uint64_t ifd = pos;
    uint64_t extra = ifd + 8 + 20 * tags.size() + 8;
    std::vector<uint8_t> dir, tail;
    put_le(dir, (uint64_t)tags.size());
    for (const OutTag &t : tags) {
        put_le(dir, t.tag);
        put_le(dir, t.type);
        put_le(dir, t.count);
        if (t.data.size() <= 8) {
            std::vector<uint8_t> v = t.data;
            v.resize(8, 0);
            dir.insert(dir.end(), v.begin(), v.end());
        } else {
            put_le(dir, (uint64_t)(extra + tail.size()));
            tail.insert(tail.end(), t.data.begin(), t.data.end());
            if (tail.size() & 1) tail.push_back(0);
        }
    }
    put_le(dir, (uint64_t)0);
    if (fwrite(dir.data(), 1, dir.size(), F.f) != dir.size() ||
        (tail.size() && fwrite(tail.data(), 1, tail.size(), F.f) != tail.size()))
        return fail(SMX_IO_E_OPEN, "write failed");
    uint8_t first[8];
    for (int i = 0; i < 8; i++) first[i] = (uint8_t)(ifd >> (8 * i));
    if (fseeko(F.f, 8, SEEK_SET) != 0 || fwrite(first, 1, 8, F.f) != 8) return fail(SMX_IO_E_OPEN, "write failed");
    if (fflush(F.f) != 0) return fail(SMX_IO_E_OPEN, "write failed");
    return 0;
}

int smx_dg_xml_read(const char *path, smx_rpc *rpc, double err[2], smx_dg_info *info)
{
    if (!path || !rpc) return fail(SMX_IO_E_ARG, "smx_dg_xml_read: null argument");
    File F(path, "r");
    if (!F.f) return fail(SMX_IO_E_OPEN, "cannot open %s", path);
    // tag -> destination, SubFunctions.cpp:1336-1488 (offsets / scales are
    // stored in file order {line, samp, lat, long, height} and swapped below)
    double off[5] = {0}, scl[5] = {0}, e2[2] = {0, 0};
    smx_dg_info in{};
    double *coef[4] = {rpc->coef[0], rpc->coef[1], rpc->coef[2], rpc->coef[3]};
    bool got[4] = {false, false, false, false};
    const struct {
        const char *tag;
        double *dst;
    } scalars[] = {{"ABSCALFACTOR", &in.abscalfactor}, {"EFFECTIVEBANDWIDTH", &in.effbw}, {"TDILEVEL", &in.tdi},
                   {"MEANCOLLECTEDROWGSD", &in.gsd_row}, {"MEANCOLLECTEDCOLGSD", &in.gsd_col},
                   {"MEANCOLLECTEDGSD", &in.gsd}, {"MEANPRODUCTROWGSD", &in.gsd_row},
                   {"MEANPRODUCTCOLGSD", &in.gsd_col}, {"MEANPRODUCTGSD", &in.gsd}, {"ERRBIAS", &e2[0]},
                   {"ERRRAND", &e2[1]}, {"LINEOFFSET", &off[0]}, {"SAMPOFFSET", &off[1]}, {"LATOFFSET", &off[2]},
                   {"LONGOFFSET", &off[3]}, {"HEIGHTOFFSET", &off[4]}, {"LINESCALE", &scl[0]},
                   {"SAMPSCALE", &scl[1]}, {"LATSCALE", &scl[2]}, {"LONGSCALE", &scl[3]}, {"HEIGHTSCALE", &scl[4]}};
    const char *lists[4] = {"LINENUMCOEFList", "LINEDENCOEFList", "SAMPNUMCOEFList", "SAMPDENCOEFList"};
    char line[4096];
    while (fgets(line, sizeof(line), F.f)) {
        // the line's first tag name: text after the first '<' up to '>'
        const char *lt = strchr(line, '<');
        if (!lt) continue;
        const char *gt = strchr(lt, '>');
        if (!gt) continue;
        const std::string name(lt + 1, gt);
        bool done = false;
        for (const auto &s : scalars)
            if (name == s.tag) {
                *s.dst = atof(gt + 1);   // the value runs to the next '<'
                done = true;
                break;
            }
        if (done) continue;
        for (int k = 0; k < 4; k++) {
            if (name != lists[k]) continue;
            // the coefficients are on the next line, between '>' and '<'
            char next[8192];
            if (!fgets(next, sizeof(next), F.f)) return fail(SMX_IO_E_FORMAT, "%s: missing coefficient line", lists[k]);
            const char *p = strchr(next, '>');
            p = p ? p + 1 : next;
            std::string body(p);
            const size_t stop = body.find('<');
            if (stop != std::string::npos) body.resize(stop);
            char *save = nullptr;
            int i = 0;
            for (char *tok = strtok_r(&body[0], " \t\r\n", &save); tok && i < 20;
                 tok = strtok_r(nullptr, " \t\r\n", &save))
                coef[k][i++] = atof(tok);
            for (; i < 20; i++) coef[k][i] = 0.0;
            got[k] = true;
        }
    }
    for (int k = 0; k < 4; k++)
        if (!got[k]) return fail(SMX_IO_E_FORMAT, "%s: no %s", path, lists[k]);
    // SubFunctions.cpp:1576-1582: lat / long swapped in both rows
    const int order[5] = {0, 1, 3, 2, 4};
    for (int i = 0; i < 5; i++) {
        rpc->off[i] = off[order[i]];
        rpc->scale[i] = scl[order[i]];
    }
    if (err) {
        err[0] = e2[0];
        err[1] = e2[1];
    }
    if (info) *info = in;
    return 0;
}

int smx_dg_xml_write(const char *path, const smx_rpc *rpc, const double err[2], const smx_dg_info *info)
{
    if (!path || !rpc) return fail(SMX_IO_E_ARG, "smx_dg_xml_write: null argument");
    File F(path, "w");
    if (!F.f) return fail(SMX_IO_E_OPEN, "cannot create %s", path);
    const smx_dg_info in = info ? *info : smx_dg_info{};
    FILE *f = F.f;
    fprintf(f, "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<isd>\n<IMD>\n");
    fprintf(f, "<BAND_P>\n<ABSCALFACTOR>%.17g</ABSCALFACTOR>\n<EFFECTIVEBANDWIDTH>%.17g</EFFECTIVEBANDWIDTH>\n"
               "<TDILEVEL>%.17g</TDILEVEL>\n</BAND_P>\n",
            in.abscalfactor, in.effbw, in.tdi);
    fprintf(f, "<IMAGE>\n<MEANPRODUCTROWGSD>%.17g</MEANPRODUCTROWGSD>\n<MEANPRODUCTCOLGSD>%.17g</MEANPRODUCTCOLGSD>\n"
               "<MEANPRODUCTGSD>%.17g</MEANPRODUCTGSD>\n</IMAGE>\n</IMD>\n",
            in.gsd_row, in.gsd_col, in.gsd);
    fprintf(f, "<RPB>\n<IMAGE>\n<ERRBIAS>%.17g</ERRBIAS>\n<ERRRAND>%.17g</ERRRAND>\n", err ? err[0] : 0.0,
            err ? err[1] : 0.0);
    const char *on[5] = {"LINEOFFSET", "SAMPOFFSET", "LONGOFFSET", "LATOFFSET", "HEIGHTOFFSET"};
    const char *sn[5] = {"LINESCALE", "SAMPSCALE", "LONGSCALE", "LATSCALE", "HEIGHTSCALE"};
    const int file_order[5] = {0, 1, 3, 2, 4};   // line, samp, lat, long, height
    for (int i : file_order) fprintf(f, "<%s>%.17g</%s>\n", on[i], rpc->off[i], on[i]);
    for (int i : file_order) fprintf(f, "<%s>%.17g</%s>\n", sn[i], rpc->scale[i], sn[i]);
    const char *ln[4] = {"LINENUMCOEF", "LINEDENCOEF", "SAMPNUMCOEF", "SAMPDENCOEF"};
    for (int k = 0; k < 4; k++) {
        fprintf(f, "<%sList>\n<%s>", ln[k], ln[k]);
        for (int i = 0; i < 20; i++) fprintf(f, "%s%.17g", i ? " " : "", rpc->coef[k][i]);
        fprintf(f, "</%s>\n</%sList>\n", ln[k], ln[k]);
    }
    fprintf(f, "</IMAGE>\n</RPB>\n</isd>\n");
    if (fflush(f) != 0) return fail(SMX_IO_E_OPEN, "write failed");
    return 0;
}

}  // extern "C"
