"""Formats of the matching path (libsmx_io.so) on the CPU.

* GeoTIFF writer (setsmgeo.cpp:22-116) checked by libtiff, the library the
  reference writes and reads TIFF with: pixels, BigTIFF, LZW, predictor,
  rows per strip, tie point, pixel scale, GDAL_NODATA and GeoKeys.
* TIFF reader (Template.hpp:211-360) on files libtiff writes in every layout
  a reader meets: classic / BigTIFF / big-endian, none / LZW, predictors
  1 / 2 / 3, strips / tiles, and the reference's widening of a window to
  whole tiles (expected windows worked out by hand from Template.hpp:303-322).
* DigitalGlobe XML reader (SubFunctions.cpp:1302-1582) on a fixture whose
  expected parse is listed by hand, including the lat / long swap.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from setsm_amd import io as sio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TC_PATH = os.path.join(ROOT, "tests", "tiffcheck", "_build", "libtiffcheck.so")


def test_io_library_exports_header():
    L = sio.lib()
    decl = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "smx_io.h")).read(), flags=re.S)
    names = set(re.findall(r"\b(smx_[a-z0-9_]+)\s*\(", decl))
    assert names == set(sio.EXPORTS)
    for n in names:
        assert hasattr(L, n), n


def _tc():
    if not os.path.exists(TC_PATH):
        pytest.skip("libtiff test helper not built (libtiff headers absent)")
    L = C.CDLL(TC_PATH)
    L.tc_write.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int,
                           C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]
    L.tc_read.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_double),
                          C.POINTER(C.c_double), C.c_char_p, C.c_int, C.POINTER(C.c_uint16), C.c_int]
    return L


def libtiff_read(path, dtype, shape_cap):
    L = _tc()
    out = np.zeros(shape_cap, dtype=dtype)
    info = (C.c_uint32 * 8)()
    tie = (C.c_double * 6)()
    scale = (C.c_double * 3)()
    nod = C.create_string_buffer(64)
    keys = (C.c_uint16 * 256)()
    r = L.tc_read(os.fsencode(path), out.ctypes.data, out.nbytes, info, tie, scale, nod, 64, keys, 256)
    assert r == 0, f"libtiff could not read {path}: {r}"
    w, h = info[0], info[1]
    return dict(pix=out.ravel()[: w * h].reshape(h, w), info=list(info), tie=list(tie), scale=list(scale),
                nodata=nod.value.decode(), keys=list(keys))


def _image(dtype, h, w, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.float32:
        a = (rng.standard_normal((h, w)) * 300 + 100).astype(np.float32)
        a.flat[:: 7] = -9999.0
        if a.size > 3:
            a.flat[1] = np.nan
            a.flat[2] = np.inf
        return a
    if dtype == np.uint16:
        # smooth ramps and noise, DN up to 2047 (11-bit imagery)
        yy, xx = np.mgrid[0:h, 0:w]
        return ((xx * 3 + yy * 5 + rng.integers(0, 40, (h, w))) % 2048).astype(np.uint16)
    return rng.integers(0, 256, (h, w)).astype(np.uint8)


@pytest.mark.parametrize("dtype", [np.float32, np.uint16, np.uint8])
@pytest.mark.parametrize("shape", [(1, 1), (5, 17), (300, 257), (64, 4000)])
def test_geotiff_writer_read_by_libtiff(tmp_path, dtype, shape):
    h, w = shape
    a = _image(dtype, h, w, seed=h * 131 + w)
    p = tmp_path / "dem.tif"
    sio.write_geotiff(p, a, scale=2.0, min_x=500010.0, max_y=3001234.0, projection=0, zone=14, north=True)
    got = libtiff_read(p, dtype, a.size)
    assert got["pix"].tobytes() == a.tobytes()   # NaN / inf bit patterns too
    bits = a.dtype.itemsize * 8
    width, height, b, fmt, comp, pred, rps, flags = got["info"]
    assert (width, height, b) == (w, h, bits)
    assert fmt == (3 if dtype == np.float32 else 1)
    assert comp == 5                                            # LZW, setsmgeo.cpp:213
    assert pred == {np.float32: 3, np.uint16: 2, np.uint8: 1}[dtype]
    assert rps == min(h, max(1, (8192 * 8) // (w * bits)))      # setsmgeo.cpp:227-245
    assert flags & 1 == 1                                       # BigTIFF ("w8")
    assert got["tie"] == [0.0, 0.0, 0.0, 500010.0, 3001234.0, 0.0]
    assert got["scale"] == [2.0, 2.0, 0.0]
    assert got["nodata"] == ("-9999" if dtype == np.float32 else "")
    # GeoKeys: version 1.1.0, then ProjectedCSType = WGS 84 / UTM zone 14N (32614)
    k = got["keys"]
    assert k[:3] == [1, 1, 0]
    entries = {k[4 + 4 * i]: k[4 + 4 * i: 8 + 4 * i] for i in range(k[3])}
    assert entries[1024][3] == 1 and entries[1025][3] == 1
    assert entries[3072][3] == 32614
    assert entries[3076][3] == 9001


def test_geotiff_polar_stereographic_keys(tmp_path):
    a = _image(np.float32, 8, 8, 3)
    for north, lat0 in ((1, 70.0), (0, -71.0)):
        p = tmp_path / f"ps{north}.tif"
        sio.write_geotiff(p, a, 8.0, -100.0, 200.0, projection=1, north=bool(north))
        got = libtiff_read(p, np.float32, a.size)
        k = got["keys"]
        entries = {k[4 + 4 * i]: k[4 + 4 * i: 8 + 4 * i] for i in range(k[3])}
        assert entries[3075][3] == 15            # CT_PolarStereographic
        assert entries[3072][3] == 32767 and entries[3074][3] == 32767
        assert entries[3081][1] == 34736         # natural origin latitude: a double parameter
        info = sio.tiff_info(p)
        assert info.has_geo and info.min_x == -100.0 and info.max_y == 200.0 and info.grid_dx == 8.0
        del lat0


def test_geotiff_round_128(tmp_path):
    a = np.array([[0.0039, 1.0 / 256, 12.3456, -3.9961]], dtype=np.float32)
    p = tmp_path / "r.tif"
    sio.write_geotiff(p, a, 1.0, 0.0, 0.0, round_128=True)
    back, _, _ = sio.read_tiff(p)
    # std::round(v * 128) / 128 in float: halves away from zero (0.5 -> 1)
    assert back.tolist() == [[0.0, 0.0078125, 12.34375, -4.0]]


def test_geotiff_roundtrip_own_reader(tmp_path):
    for dtype in (np.float32, np.uint16, np.uint8):
        a = _image(dtype, 97, 131, 9)
        p = tmp_path / f"x{a.dtype}.tif"
        sio.write_geotiff(p, a, 0.5, 1.0, 2.0)
        back, cols, rows = sio.read_tiff(p)
        assert back.tobytes() == a.tobytes() and cols == (0, 131) and rows == (0, 97)
        sub, cols, rows = sio.read_tiff(p, cols=(10, 77), rows=(3, 60))
        assert cols == (10, 77) and rows == (3, 60)
        assert sub.tobytes() == a[3:60, 10:77].tobytes()


LAYOUTS = [
    # mode, compression, predictor, tiled, tile w, tile h, rows per strip
    ("w", 1, 1, 0, 0, 0, 1), ("w", 5, 1, 0, 0, 0, 7), ("w", 5, 2, 0, 0, 0, 1000),
    ("w8", 5, 2, 0, 0, 0, 3), ("wb", 5, 2, 0, 0, 0, 5), ("wb", 1, 1, 0, 0, 0, 4),
    ("w", 1, 1, 1, 16, 16, 0), ("w", 5, 2, 1, 32, 16, 0), ("w8", 5, 2, 1, 16, 48, 0), ("wb", 5, 2, 1, 32, 32, 0),
]


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
@pytest.mark.parametrize("layout", LAYOUTS)
def test_reader_on_libtiff_files(tmp_path, dtype, layout):
    L = _tc()
    mode, comp, pred, tiled, tw, th, rps = layout
    if dtype == np.float32 and pred == 2:
        pred = 3   # floating-point predictor for float images
    h, w = 70, 100
    a = _image(dtype, h, w, seed=hash(layout) & 0xffff)
    p = tmp_path / "in.tif"
    bits = a.dtype.itemsize * 8
    fmt = 3 if dtype == np.float32 else 1
    r = L.tc_write(os.fsencode(p), mode.encode(), a.ctypes.data, w, h, bits, fmt, comp, pred, tiled, tw, th, rps)
    assert r == 0
    info = sio.tiff_info(p)
    assert (info.width, info.height, info.compression, info.tiled) == (w, h, comp, tiled)
    full, cols, rows = sio.read_tiff(p, cols=(0, w), rows=(0, h))
    # the same pixels libtiff reads from the file
    assert full.tobytes() == libtiff_read(p, dtype, a.size)["pix"].tobytes()
    # which are the pixels written, except for a big-endian file with the
    # floating-point predictor: libtiff 4.2 on a little-endian host writes
    # such a file's strips with their bytes swapped and reads it back so
    a_file = a.byteswap() if (mode == "wb" and pred == 3) else a
    if not (mode == "wb" and pred == 3):
        assert full.tobytes() == a.tobytes()
    a = a_file
    # a window: exact for strips; whole tiles for tiled files, the end
    # computed as (end / tile + 1) * tile and clipped to the image
    # (Template.hpp:303-322), worked out by hand for a 100 x 70 image
    hand = {  # (tile w, tile h): windows for cols (10, 40) rows (5, 64), and cols (40, 50) rows (20, 30)
        (16, 16): (((0, 48), (0, 70)), ((32, 64), (16, 32))),
        (32, 16): (((0, 64), (0, 70)), ((32, 64), (16, 32))),
        (16, 48): (((0, 48), (0, 70)), ((32, 64), (0, 48))),
        (32, 32): (((0, 64), (0, 70)), ((32, 64), (0, 32))),
    }
    for k, (cw, rw) in enumerate((((10, 40), (5, 64)), ((40, 50), (20, 30)))):
        sub, cols, rows = sio.read_tiff(p, cols=cw, rows=rw)
        exp_c, exp_r = hand[(tw, th)][k] if tiled else (cw, rw)
        assert cols == exp_c and rows == exp_r
        assert sub.tobytes() == a[exp_r[0]:exp_r[1], exp_c[0]:exp_c[1]].tobytes()


def test_reader_refuses_a_type_mismatch(tmp_path):
    a = _image(np.uint16, 4, 4, 1)
    p = tmp_path / "u16.tif"
    sio.write_geotiff(p, a, 1.0, 0.0, 0.0)
    with pytest.raises(IOError, match="16-bit"):
        sio.read_tiff(p, dtype_code=sio.FLOAT)
    with pytest.raises(IOError):
        sio.read_tiff(tmp_path / "missing.tif")


# hand-listed parse of tests/golden/dg_rpc_sample.xml
DG_OFF = [10240.0, 9216.0, 163.5, -71.125, 512.0]      # line, samp, lon, lat, height (swapped)
DG_SCALE = [10241.0, 9217.0, 0.1875, 0.0625, 640.0]


def test_dg_xml_fixture():
    rpc, err, info = sio.read_dg_xml(os.path.join(ROOT, "tests", "golden", "dg_rpc_sample.xml"))
    assert rpc[0:5].tolist() == DG_OFF
    assert rpc[5:10].tolist() == DG_SCALE
    c = rpc[10:].reshape(4, 20)
    assert c[0, 0] == 1.25e-3 and c[0, 1] == -2.5e-3 and c[0, 19] == -2.5e-2
    assert c[1, 0] == 1.0 and c[1, 1] == 5e-5 and c[1, 2] == -7.5e-5
    assert c[2, 0] == -1.05e-3 and c[2, 19] == -7.7e-3
    assert c[3, 0] == 1.0 and c[3, 1] == 4e-6 and c[3, 19] == 7.6e-5
    assert err.tolist() == [3.75, 0.25]
    # the product GSD lines come after the collected ones and win
    assert (info.gsd_row, info.gsd_col, info.gsd) == (0.525, 0.51, 0.5175)
    assert (info.abscalfactor, info.effbw, info.tdi) == (9.295654e-03, 2.846e-01, 16.0)


def test_dg_xml_roundtrip_preserves_projection(tmp_path):
    from oracle import oracle_py as O
    from setsm_amd.synth import SceneSpec, make_scene
    sc = make_scene(SceneSpec(image_px=256, relief=40.0, seed=5))
    for t in range(2):
        p = tmp_path / f"img{t}.xml"
        sio.write_dg_xml(p, sc.rpcs[t], err=(1.5, 0.5))
        back, err, _ = sio.read_dg_xml(p)
        assert back.tobytes() == np.asarray(sc.rpcs[t], np.float64).tobytes()
        assert err.tolist() == [1.5, 0.5]
        lon, lat = sc.to_wgs(np.array([60.0]), np.array([70.0]))
        assert O.rpc_ground_to_image(back, (0, 0), lon[0], lat[0], 210.0) == \
            O.rpc_ground_to_image(sc.rpcs[t], (0, 0), lon[0], lat[0], 210.0)


def test_dg_xml_missing_list(tmp_path):
    p = tmp_path / "bad.xml"
    p.write_text("<RPB>\n<LINEOFFSET>1</LINEOFFSET>\n</RPB>\n")
    with pytest.raises(IOError, match="LINENUMCOEFList"):
        sio.read_dg_xml(p)
