"""Raster files and the per-step statistics CSV (host code; no GPU).

Fixtures: tests/golden/rasters/ holds data files from the reference's own
test data (tests/data/tiny30_dem.tif: 30x30 float32, uncompressed,
EPSG:4326; flat_resolution_lakeland_crop_dem.tif: 128x128 float32, DEFLATE
with the floating-point predictor, UTM; sea_mask/coast_dem.tif:
uncompressed with a GDAL_NODATA tag; tiny_dem.asc and tiny_dem_nodata.asc:
ESRI ASCII grids).  GDAL is absent here, so the expected values come from
independent decoders written in this file with numpy and zlib from the
TIFF 6.0 / Technical Note 3 layouts (strip offsets, byte planes of the
floating-point predictor, LZW code stream), and from numpy.loadtxt for the
ASCII grids: a mistake in the library's tag parsing, plane order or
differencing fails one of them.
"""
import os
import struct
import zlib

import numpy as np
import pytest

from lsdtt3_amd import _native as N
from lsdtt3_amd import formats as F

GOLD = os.path.join(os.path.dirname(__file__), "golden", "rasters")


def gold(name):
    return os.path.join(GOLD, name)


# --------------------------------------------------------------------------
# independent TIFF reading (classic little-endian, strips)
# --------------------------------------------------------------------------
def tiff_tags(path):
    b = open(path, "rb").read()
    assert b[:2] == b"II"
    off = struct.unpack("<I", b[4:8])[0]
    n = struct.unpack("<H", b[off:off + 2])[0]
    sizes = {1: 1, 2: 1, 3: 2, 4: 4, 12: 8}
    fmt = {1: "B", 2: "c", 3: "H", 4: "I", 12: "d"}
    tags = {}
    for i in range(n):
        t, ty, cnt, val = struct.unpack("<HHII", b[off + 2 + 12 * i:off + 14 + 12 * i])
        nbytes = sizes[ty] * cnt
        at = off + 2 + 12 * i + 8 if nbytes <= 4 else val
        raw = b[at:at + nbytes]
        tags[t] = raw.split(b"\0")[0].decode() if ty == 2 else list(struct.unpack("<" + fmt[ty] * cnt, raw))
    return b, tags


def decode_strips(path):
    """float32 strips, compression 1 or 8, predictor 1 or 3"""
    b, t = tiff_tags(path)
    W, H = t[256][0], t[257][0]
    rps = t.get(278, [H])[0]
    comp = t.get(259, [1])[0]
    pred = t.get(317, [1])[0]
    assert t[258][0] == 32 and t[339][0] == 3
    out = np.empty((H, W), np.float32)
    for s, (o, c) in enumerate(zip(t[273], t[279])):
        data = b[o:o + c]
        if comp == 8:
            data = zlib.decompress(data)
        rows = min(rps, H - s * rps)
        if pred == 3:
            a = np.frombuffer(data, np.uint8)[:rows * W * 4].reshape(rows, W * 4).astype(np.uint32)
            a = np.cumsum(a, axis=1) & 0xFF                      # undo the byte differences
            planes = a.reshape(rows, 4, W)                          # plane 0: most significant byte
            u = (planes[:, 0] << 24) | (planes[:, 1] << 16) | (planes[:, 2] << 8) | planes[:, 3]
            vals = u.astype(np.uint32).view(np.float32)
        else:
            vals = np.frombuffer(data, "<f4")[:rows * W].reshape(rows, W)
        out[s * rps:s * rps + rows] = vals
    return out, t


@pytest.mark.parametrize("name", ["tiny30_dem.tif", "coast_dem.tif", "flat_resolution_lakeland_crop_dem.tif"])
def test_geotiff_values_match_an_independent_decoder(name):
    z, mask, meta = F.read_raster(gold(name))
    want, t = decode_strips(gold(name))
    nd = float(t[42113]) if 42113 in t else -9999.0
    bad = (want == np.float32(nd)) | ~np.isfinite(want)
    want = np.where(bad, np.float32(nd), want)
    assert z.shape == want.shape
    assert np.array_equal(z.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(mask, (~bad).astype(np.uint8))
    assert meta["no_data_value"] == nd


def test_geotiff_deflate_predictor_fixture_is_terrain():
    """the DEFLATE + floating-point predictor crop decodes to a plausible DEM
    (a wrong plane order gives noise): finite, metre-scale neighbour steps"""
    z, mask, meta = F.read_raster(gold("flat_resolution_lakeland_crop_dem.tif"))
    assert mask.all() and np.isfinite(z).all()
    assert 0.0 < float(np.median(np.abs(np.diff(z, axis=1)))) < 10.0
    assert meta["cell_size"] == pytest.approx(30.0, rel=1e-6) and meta["geographic"] == 0


@pytest.mark.parametrize("name", ["tiny30_dem.tif", "coast_dem.tif", "flat_resolution_lakeland_crop_dem.tif"])
def test_geotiff_georeference(name):
    """x_min = tiepoint X - I sx, y_min = tiepoint Y + J sy - H sy
    (lsdtt_data_loader.cpp:1782-1788), CRS as the GeoKeys' EPSG code"""
    _, t = tiff_tags(gold(name))
    sx, sy = t[33550][0], t[33550][1]
    I, J, X, Y = t[33922][0], t[33922][1], t[33922][3], t[33922][4]
    H = t[257][0]
    keys = t[34735]
    kv = {keys[4 + 4 * i]: keys[4 + 4 * i + 3] for i in range(keys[3]) if keys[4 + 4 * i + 1] == 0}
    meta = F.raster_info(gold(name))
    assert meta["n_rows"] == H and meta["n_cols"] == t[256][0]
    # PixelIsPoint (RasterTypeGeoKey 2): the tiepoint is a pixel centre, and
    # GDAL reports the corner half a pixel out
    half = 0.5 if kv.get(1025) == 2 else 0.0
    assert meta["x_min"] == pytest.approx(X - (I + half) * sx, abs=1e-9)
    assert meta["y_min"] == pytest.approx(Y + (J + half) * sy - H * sy, abs=1e-9)
    assert meta["cell_size"] == pytest.approx(sx, abs=0)
    geographic = kv.get(1024) == 2
    assert meta["geographic"] == int(geographic)
    assert meta["epsg"] == kv.get(3072, kv.get(2048, 0))


def test_geotiff_tiny30_is_epsg4326():
    meta = F.raster_info(gold("tiny30_dem.tif"))
    assert meta["epsg"] == 4326 and meta["geographic"] == 1


@pytest.mark.parametrize("compress", [0, 1])
def test_geotiff_round_trip(tmp_path, compress):
    rng = np.random.default_rng(3)
    z = (rng.normal(size=(37, 53)) * 100.0).astype(np.float32)
    z[5, 7] = np.nan
    mask = np.ones(z.shape, np.uint8)
    mask[10:12, 3:9] = 0
    meta = {"x_min": 500000.0, "y_min": 4100000.0, "cell_size": 10.0, "no_data_value": -9999.0, "epsg": 32615}
    p = tmp_path / "a.tif"
    F.write_raster(p, z, meta, mask=mask, compress=bool(compress))
    got, gmask, gmeta = F.read_raster(p)
    want_mask = mask.copy()
    want_mask[5, 7] = 0
    want = np.where(want_mask == 1, z, np.float32(-9999.0))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(gmask, want_mask)
    assert gmeta["x_min"] == 500000.0 and gmeta["y_min"] == 4100000.0 and gmeta["cell_size"] == 10.0
    assert gmeta["epsg"] == 32615 and gmeta["geographic"] == 0
    # the independent decoder reads what the library wrote: masked cells as
    # the no-data value, every other value as given (the NaN included)
    dec, _ = decode_strips(p)
    written = np.where(mask == 1, z, np.float32(-9999.0))
    assert np.array_equal(dec.view(np.uint32), written.view(np.uint32))


def _lzw_encode(data: bytes) -> bytes:
    """TIFF LZW (MSB-first codes, clear 256, end 257, early width change)"""
    table = {bytes([i]): i for i in range(256)}
    nxt, width = 258, 9
    bits, nb, out = 0, 0, bytearray()

    def emit(code):
        nonlocal bits, nb
        bits = (bits << width) | code
        nb += width
        while nb >= 8:
            out.append((bits >> (nb - 8)) & 0xFF)
            nb -= 8

    emit(256)
    w = b""
    for ch in data:
        wc = w + bytes([ch])
        if wc in table:
            w = wc
            continue
        emit(table[w])
        table[wc] = nxt
        nxt += 1
        if nxt + 1 > (1 << width) and width < 12:
            width += 1
        if nxt >= 4093:
            emit(256)
            table = {bytes([i]): i for i in range(256)}
            nxt, width = 258, 9
        w = bytes([ch])
    if w:
        emit(table[w])
    emit(257)
    if nb:
        out.append((bits << (8 - nb)) & 0xFF)
    return bytes(out)


def _tiff_uint16_lzw_pred2(a: np.ndarray) -> bytes:
    """a uint16 raster as one strip, horizontal predictor, LZW"""
    H, W = a.shape
    d = a.astype(np.int64)
    diff = np.concatenate([d[:, :1], np.diff(d, axis=1)], axis=1) & 0xFFFF
    strip = _lzw_encode(diff.astype("<u2").tobytes())
    entries = [(256, 3, 1, W), (257, 3, 1, H), (258, 3, 1, 16), (259, 3, 1, 5), (262, 3, 1, 1),
               (273, 4, 1, None), (277, 3, 1, 1), (278, 3, 1, H), (279, 4, 1, len(strip)), (317, 3, 1, 2),
               (339, 3, 1, 1)]
    ifd = 8
    data_off = ifd + 2 + 12 * len(entries) + 4
    b = bytearray(b"II" + struct.pack("<HI", 42, ifd) + struct.pack("<H", len(entries)))
    for t, ty, cnt, v in entries:
        v = data_off if v is None else v
        b += struct.pack("<HHI", t, ty, cnt) + (struct.pack("<HH", v, 0) if ty == 3 else struct.pack("<I", v))
    b += struct.pack("<I", 0) + strip
    return bytes(b)


def test_geotiff_lzw_and_horizontal_predictor(tmp_path):
    rng = np.random.default_rng(5)
    a = (np.cumsum(rng.integers(0, 40, size=(23, 61)), axis=1) % 65536).astype(np.uint16)
    a[3, :] = 7  # long runs exercise the table growth
    p = tmp_path / "lzw.tif"
    p.write_bytes(_tiff_uint16_lzw_pred2(a))
    z, mask, meta = F.read_raster(p)
    assert np.array_equal(z, a.astype(np.float32))
    assert mask.all() and meta["cell_size"] == 1.0 and meta["x_min"] == 0.0  # no georeference


def test_ascii_grid_read_and_nodata():
    z, mask, meta = F.read_raster(gold("tiny_dem.asc"))
    want = np.loadtxt(gold("tiny_dem.asc"), skiprows=6).astype(np.float32)
    assert np.array_equal(z, want) and mask.all()
    assert (meta["n_rows"], meta["n_cols"], meta["cell_size"], meta["x_min"], meta["y_min"]) == (5, 5, 1.0, 0.0, 0.0)
    z, mask, meta = F.read_raster(gold("tiny_dem_nodata.asc"))
    want = np.loadtxt(gold("tiny_dem_nodata.asc"), skiprows=6).astype(np.float32)
    assert np.array_equal(mask, (want != -9999.0).astype(np.uint8))
    assert np.array_equal(z, want) and mask.sum() == want.size - 1


def test_ascii_grid_round_trip(tmp_path):
    rng = np.random.default_rng(9)
    z = (rng.normal(size=(11, 7)) * 1e3).astype(np.float32)
    mask = (rng.random(z.shape) > 0.1).astype(np.uint8)
    p = tmp_path / "g.asc"
    F.write_raster(p, z, {"x_min": 12.5, "y_min": -3.0, "cell_size": 2.0, "no_data_value": -32768.0}, mask=mask)
    got, gmask, meta = F.read_raster(p)
    want = np.where(mask == 1, z, np.float32(-32768.0))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)) and np.array_equal(gmask, mask)
    assert (meta["x_min"], meta["y_min"], meta["cell_size"], meta["no_data_value"]) == (12.5, -3.0, 2.0, -32768.0)


def test_ascii_grid_cell_centre_header(tmp_path):
    p = tmp_path / "c.asc"
    p.write_text("ncols 2\nnrows 2\nxllcenter 10\nyllcenter 20\ncellsize 4\n1 2\n3 4\n")
    z, mask, meta = F.read_raster(p)
    assert meta["x_min"] == 8.0 and meta["y_min"] == 18.0 and meta["no_data_value"] == -9999.0
    assert np.array_equal(z, np.array([[1, 2], [3, 4]], np.float32))


def test_raster_errors(tmp_path):
    with pytest.raises(N.GpuError) as e:
        F.raster_info(tmp_path / "x.png")
    assert e.value.status == N.ERR_ARG
    with pytest.raises(N.GpuError) as e:
        F.raster_info(tmp_path / "missing.tif")
    assert e.value.status == N.ERR_RUNTIME
    bad = tmp_path / "rot.asc"
    bad.write_text("ncols 2\nnrows 2\nxllcorner 0\nyllcorner 0\ndx 1\ndy 2\n1 2\n3 4\n")
    with pytest.raises(N.GpuError) as e:
        F.read_raster(bad)
    assert e.value.status == N.ERR_ARG


# --------------------------------------------------------------------------
# the per-step statistics CSV
# --------------------------------------------------------------------------
def test_channelization_index_closed_form():
    # n = 1: U drops out; n != 1: U^(n-1) in the numerator
    assert F.channelization_index(1e-5, 0.01, 1e-3, 1000.0, 0.5, 1.0) == pytest.approx(1e-5 * 1000.0 ** 1.5 / 0.01)
    assert F.channelization_index(2e-5, 0.005, 5e-4, 3e4, 0.5, 1.5) == pytest.approx(
        2e-5 * 3e4 ** 2.0 / (0.005 ** 1.5 * 5e-4 ** -0.5))
    assert F.channelization_index(0.0, 0.01, 1e-3, 1000.0, 0.5, 1.0) == 0.0
    assert F.channelization_length(100, 200, 30.0, "onon") == 3000.0       # N-S outlets: rows
    assert F.channelization_length(100, 200, 30.0, "nono") == 6000.0       # E-W outlets: columns
    assert F.channelization_length(100, 200, 30.0, "nnon") == 6000.0       # otherwise the longer side
    assert F.channelization_length(300, 200, 30.0, "oooo") == 9000.0


def test_stats_csv_format():
    """std::to_string ("%f") of float32 values, seven significant figures for
    the parameters, time accumulated in float32 (cli/lsdtt-lem.cpp:3567-3716)"""
    t = F.ElevationStatsCSV(512, 512, 30.0, "nnon")
    st = {"delta_max_elevation": 0.123456789, "delta_mean_elevation": -2.5e-7, "max_local_elevation_change": 1.0,
          "mean_erosion_rate": 6.75e-5, "max_erosion_rate": 1e-3}
    for _ in range(3):
        t.add_row(st, dt=1000.0, uplift_rate=1e-3, baselevel_fall_rate=0.0, erodibility=6.75e-5,
                  hillslope_diffusivity=0.01, m=0.5, n=1.0)
    lines = t.text().splitlines()
    assert lines[0] == ("iteration,time_yr,delta_max_z,delta_mean_z,max_local_delta_z,mean_erosion_rate,"
                        "max_erosion_rate,uplift_rate,baselevel_fall_rate,erodibility,hillslope_diffusivity,"
                        "channelization_index")
    f = lines[3].split(",")
    # float32(6.75e-5) prints as 0.000067, the case the CLI's comment names
    # (cli/lsdtt-lem.cpp:3524-3525)
    assert f[:9] == ["3", "3000.000000", "0.123457", "-0.000000", "1.000000", "0.000067", "0.001000", "0.001000",
                     "0.000000"]
    assert f[9] == "6.75e-05" and f[10] == "0.01"
    ci = F.channelization_index(float(np.float32(6.75e-5)), float(np.float32(0.01)), float(np.float32(1e-3)),
                                15360.0, 0.5, 1.0)
    assert f[11] == "%.7g" % ci
    # float32 accumulation of the elapsed time
    t2 = F.ElevationStatsCSV(10, 10, 1.0, "oooo")
    tt = np.float32(0)
    for _ in range(1000):
        t2.add_row(st, 0.1, 0, 0, 1, 1, 0.5, 1)
        tt = np.float32(tt + np.float32(0.1))
    assert t2.rows[-1][1] == "%f" % float(tt) and t2.rows[-1][1] != "100.000000"
