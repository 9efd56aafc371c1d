// raster_io.hip -- the raster files on either side of the LEM path (host
// code only): single-band GeoTIFF and ESRI ASCII grids, read into and written
// from the row-major float32 surface + validity mask that every entry point
// of this library takes.
//
// Read semantics follow the reference's loader
// (src/core/lsdtt_data_loader.cpp):
//  * the band is converted to float32; a value equal to the no-data value
//    or not finite has mask 0 and is replaced by the no-data value
//    (:724-747);
//  * georeference: north-up only (a rotated or south-up raster is refused),
//    square pixels unless the CRS is geographic; x_min = gt[0],
//    cell_size = gt[1], y_min = gt[3] + n_rows gt[5] (:1735-1790); no
//    georeference gives x_min = y_min = 0, cell_size = 1;
//  * no-data from the band, else -9999 (:1795-1800).
// GeoTIFF: classic and BigTIFF, either byte order, strips or tiles,
// uncompressed, LZW or DEFLATE, horizontal (2) or floating-point (3)
// predictor, 8/16/32/64-bit integer or 32/64-bit float samples (the first
// sample of each pixel when there are several).  The georeference comes
// from ModelPixelScale + ModelTiepoint (PixelIsPoint shifted by half a pixel,
// as GDAL does) or a ModelTransformation without rotation; the CRS is kept
// as its EPSG code (ProjectedCSTypeGeoKey / GeographicTypeGeoKey).
// Written GeoTIFFs are classic little-endian float32 strips, uncompressed or
// DEFLATE with the floating-point predictor, with the pixel scale, tiepoint,
// GeoKeys (model type, PixelIsArea, EPSG) and GDAL_NODATA tags.
#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_util.hpp"

namespace lsdtt_gpu {
namespace {

std::string lower_ext(const std::string& path)
{
    const size_t dot = path.find_last_of('.');
    const size_t slash = path.find_last_of('/');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    std::string e = path.substr(dot + 1);
    for (auto& ch : e) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    return e;
}

std::vector<uint8_t> read_file(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    f.seekg(0, std::ios::beg);
    std::vector<uint8_t> b(static_cast<size_t>(std::max<std::streamoff>(n, 0)));
    if (n > 0 && !f.read(reinterpret_cast<char*>(b.data()), n)) throw std::runtime_error("cannot read " + path);
    return b;
}

// finishes the reference's post-read pass (lsdtt_data_loader.cpp:724-747)
void mask_nodata(float* data, uint8_t* mask, size_t n, double nodata)
{
    const float nd = static_cast<float>(nodata);
    for (size_t i = 0; i < n; ++i) {
        const float v = data[i];
        const bool ok = !(v == nd || !std::isfinite(v));
        if (mask) mask[i] = ok ? 1 : 0;
        if (!ok) data[i] = nd;
    }
}

// ---------------------------------------------------------------------------
// TIFF reading
// ---------------------------------------------------------------------------
struct Tiff {
    const std::vector<uint8_t>& b;
    bool le = true, big = false;
    explicit Tiff(const std::vector<uint8_t>& bytes) : b(bytes) {}

    void need(uint64_t off, uint64_t n) const
    {
        if (off > b.size() || n > b.size() - off) throw std::runtime_error("GeoTIFF: truncated file");
    }
    uint64_t uint(uint64_t off, int n) const
    {
        need(off, n);
        uint64_t v = 0;
        for (int i = 0; i < n; ++i) {
            const uint64_t byte = b[off + (le ? i : n - 1 - i)];
            v |= byte << (8 * i);
        }
        return v;
    }
};

struct Entry {
    uint16_t type = 0;
    uint64_t count = 0, off = 0;  // off: where the values are (inline or pointed to)
};

int type_size(uint16_t t)
{
    switch (t) {
    case 1: case 2: case 6: case 7: return 1;        // BYTE ASCII SBYTE UNDEFINED
    case 3: case 8: return 2;                        // SHORT SSHORT
    case 4: case 9: case 11: case 13: return 4;      // LONG SLONG FLOAT IFD
    case 5: case 10: case 12: case 16: case 17: case 18: return 8;  // RATIONAL SRATIONAL DOUBLE LONG8 SLONG8 IFD8
    default: return 0;
    }
}

struct TiffImage {
    Tiff t;
    std::vector<std::pair<uint16_t, Entry>> tags;
    explicit TiffImage(const std::vector<uint8_t>& bytes) : t(bytes)
    {
        if (bytes.size() < 8) throw std::runtime_error("GeoTIFF: not a TIFF file");
        if (bytes[0] == 'I' && bytes[1] == 'I') t.le = true;
        else if (bytes[0] == 'M' && bytes[1] == 'M') t.le = false;
        else throw std::runtime_error("GeoTIFF: not a TIFF file");
        const uint64_t magic = t.uint(2, 2);
        uint64_t ifd;
        if (magic == 42) {
            ifd = t.uint(4, 4);
        } else if (magic == 43) {
            t.big = true;
            if (t.uint(4, 2) != 8) throw std::runtime_error("GeoTIFF: unsupported BigTIFF offset size");
            ifd = t.uint(8, 8);
        } else {
            throw std::runtime_error("GeoTIFF: not a TIFF file");
        }
        // the first IFD is the full-resolution image (overviews follow it)
        const uint64_t n = t.uint(ifd, t.big ? 8 : 2);
        const uint64_t e0 = ifd + (t.big ? 8 : 2), esz = t.big ? 20 : 12;
        t.need(e0, n * esz);
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t e = e0 + i * esz;
            Entry en;
            const uint16_t tag = static_cast<uint16_t>(t.uint(e, 2));
            en.type = static_cast<uint16_t>(t.uint(e + 2, 2));
            en.count = t.uint(e + 4, t.big ? 8 : 4);
            const uint64_t vo = e + (t.big ? 12 : 8);
            const int ts = type_size(en.type);
            if (ts == 0) continue;
            const uint64_t bytes_ = en.count * ts;
            en.off = bytes_ <= (t.big ? 8u : 4u) ? vo : t.uint(vo, t.big ? 8 : 4);
            tags.emplace_back(tag, en);
        }
    }
    const Entry* find(uint16_t tag) const
    {
        for (const auto& kv : tags)
            if (kv.first == tag) return &kv.second;
        return nullptr;
    }
    double value(const Entry& e, uint64_t i) const
    {
        if (i >= e.count) throw std::runtime_error("GeoTIFF: tag index out of range");
        const int ts = type_size(e.type);
        const uint64_t o = e.off + i * ts;
        switch (e.type) {
        case 1: case 7: return static_cast<double>(t.uint(o, 1));
        case 6: return static_cast<double>(static_cast<int8_t>(t.uint(o, 1)));
        case 3: return static_cast<double>(t.uint(o, 2));
        case 8: return static_cast<double>(static_cast<int16_t>(t.uint(o, 2)));
        case 4: case 13: return static_cast<double>(t.uint(o, 4));
        case 9: return static_cast<double>(static_cast<int32_t>(t.uint(o, 4)));
        case 16: case 18: return static_cast<double>(t.uint(o, 8));
        case 17: return static_cast<double>(static_cast<int64_t>(t.uint(o, 8)));
        case 11: {
            const uint32_t u = static_cast<uint32_t>(t.uint(o, 4));
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        }
        case 12: {
            const uint64_t u = t.uint(o, 8);
            double d;
            std::memcpy(&d, &u, 8);
            return d;
        }
        case 5: return static_cast<double>(t.uint(o, 4)) / static_cast<double>(t.uint(o + 4, 4));
        case 10:
            return static_cast<double>(static_cast<int32_t>(t.uint(o, 4))) /
                   static_cast<double>(static_cast<int32_t>(t.uint(o + 4, 4)));
        default: throw std::runtime_error("GeoTIFF: unsupported tag type");
        }
    }
    uint64_t u(uint16_t tag, uint64_t def, bool required = false) const
    {
        const Entry* e = find(tag);
        if (!e) {
            if (required) throw std::runtime_error("GeoTIFF: missing tag " + std::to_string(tag));
            return def;
        }
        return static_cast<uint64_t>(value(*e, 0));
    }
    std::vector<uint64_t> list(uint16_t tag) const
    {
        const Entry* e = find(tag);
        if (!e) throw std::runtime_error("GeoTIFF: missing tag " + std::to_string(tag));
        std::vector<uint64_t> v(e->count);
        for (uint64_t i = 0; i < e->count; ++i) v[i] = static_cast<uint64_t>(value(*e, i));
        return v;
    }
    std::string ascii(uint16_t tag) const
    {
        const Entry* e = find(tag);
        if (!e || e->type != 2) return "";
        t.need(e->off, e->count);
        std::string s(reinterpret_cast<const char*>(t.b.data() + e->off), e->count);
        const size_t z = s.find('\0');
        return z == std::string::npos ? s : s.substr(0, z);
    }
};

// TIFF LZW (codes MSB first, 9..12 bits, 256 clear, 257 end, the code width
// grows one code early)
std::vector<uint8_t> lzw_decode(const uint8_t* in, size_t n, size_t expect)
{
    std::vector<uint8_t> out;
    out.reserve(expect);
    std::vector<std::vector<uint8_t>> dict;
    auto reset = [&] {
        dict.assign(258, {});
        for (int i = 0; i < 256; ++i) dict[i] = {static_cast<uint8_t>(i)};
    };
    reset();
    int width = 9;
    uint64_t bitpos = 0;
    const uint64_t nbits = static_cast<uint64_t>(n) * 8;
    int prev = -1;
    while (bitpos + width <= nbits) {
        int code = 0;
        for (int i = 0; i < width; ++i) {
            const uint64_t bp = bitpos + i;
            code = (code << 1) | ((in[bp >> 3] >> (7 - (bp & 7))) & 1);
        }
        bitpos += width;
        if (code == 256) {
            reset();
            width = 9;
            prev = -1;
            continue;
        }
        if (code == 257) break;
        std::vector<uint8_t> entry;
        if (code < static_cast<int>(dict.size())) {
            entry = dict[code];
            if (entry.empty()) throw std::runtime_error("GeoTIFF: corrupt LZW data");
            if (prev >= 0) {
                std::vector<uint8_t> add = dict[prev];
                add.push_back(entry[0]);
                dict.push_back(std::move(add));
            }
        } else if (code == static_cast<int>(dict.size()) && prev >= 0) {
            entry = dict[prev];
            entry.push_back(dict[prev][0]);
            dict.push_back(entry);
        } else {
            throw std::runtime_error("GeoTIFF: corrupt LZW data");
        }
        out.insert(out.end(), entry.begin(), entry.end());
        prev = code;
        const size_t next = dict.size() + 1;
        width = next >= 4096 ? 12 : (next >= 2048 ? 12 : (next >= 1024 ? 11 : (next >= 512 ? 10 : 9)));
    }
    return out;
}

std::vector<uint8_t> inflate_all(const uint8_t* in, size_t n, size_t expect)
{
    std::vector<uint8_t> out(expect);
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) throw std::runtime_error("GeoTIFF: zlib init failed");
    zs.next_in = const_cast<Bytef*>(in);
    zs.avail_in = static_cast<uInt>(n);
    zs.next_out = out.data();
    zs.avail_out = static_cast<uInt>(out.size());
    const int rc = inflate(&zs, Z_FINISH);
    const size_t got = zs.total_out;
    inflateEnd(&zs);
    if (rc != Z_STREAM_END && rc != Z_OK && rc != Z_BUF_ERROR) throw std::runtime_error("GeoTIFF: corrupt DEFLATE data");
    out.resize(got);
    return out;
}

// one decoded block (strip or tile) of bw x bh pixels, spp samples of bps
// bits, in the file's byte order -> samples as double, first sample only
void undo_predictor(std::vector<uint8_t>& d, int pred, int bw, int bh, int spp, int bps, bool le)
{
    const int bpsample = bps / 8;
    const size_t row = static_cast<size_t>(bw) * spp * bpsample;
    if (d.size() < row * bh) d.resize(row * bh, 0);
    if (pred == 2) {
        for (int y = 0; y < bh; ++y) {
            uint8_t* r = d.data() + y * row;
            for (int x = 1; x < bw; ++x)
                for (int s = 0; s < spp; ++s) {
                    uint8_t* cur = r + (static_cast<size_t>(x) * spp + s) * bpsample;
                    const uint8_t* prv = cur - static_cast<size_t>(spp) * bpsample;
                    // add as an integer of bpsample bytes in the file's order
                    uint64_t a = 0, b = 0;
                    for (int i = 0; i < bpsample; ++i) {
                        const int k = le ? i : bpsample - 1 - i;
                        a |= static_cast<uint64_t>(cur[k]) << (8 * i);
                        b |= static_cast<uint64_t>(prv[k]) << (8 * i);
                    }
                    const uint64_t v = a + b;
                    for (int i = 0; i < bpsample; ++i) cur[le ? i : bpsample - 1 - i] = static_cast<uint8_t>(v >> (8 * i));
                }
        }
    } else if (pred == 3) {
        // floating point: each row holds the samples' bytes as planes
        // (most significant first), byte-wise differenced along the row
        std::vector<uint8_t> tmp(row);
        const size_t nsamp = static_cast<size_t>(bw) * spp;
        for (int y = 0; y < bh; ++y) {
            uint8_t* r = d.data() + y * row;
            for (size_t i = static_cast<size_t>(spp); i < row; ++i) r[i] = static_cast<uint8_t>(r[i] + r[i - spp]);
            for (size_t s = 0; s < nsamp; ++s)
                for (int bidx = 0; bidx < bpsample; ++bidx) {
                    // plane bidx holds byte (bpsample - 1 - bidx) of the value, most significant first
                    const uint8_t byte = r[bidx * nsamp + s];
                    const int sig = bpsample - 1 - bidx;  // significance of this byte
                    tmp[s * bpsample + (le ? sig : bpsample - 1 - sig)] = byte;
                }
            std::memcpy(r, tmp.data(), row);
        }
    } else if (pred != 1) {
        throw std::runtime_error("GeoTIFF: unsupported predictor " + std::to_string(pred));
    }
}

struct GeoRef {
    bool has = false;
    double gt[6] = {0, 1, 0, 0, 0, -1};
    int epsg = 0;
    bool geographic = false;
    bool point = false;  // PixelIsPoint
};

GeoRef read_georef(const TiffImage& im)
{
    GeoRef g;
    if (const Entry* k = im.find(34735)) {
        if (k->count >= 4) {
            const int nkeys = static_cast<int>(im.value(*k, 3));
            for (int i = 0; i < nkeys && 4 + 4 * i + 3 < static_cast<int>(k->count); ++i) {
                const int id = static_cast<int>(im.value(*k, 4 + 4 * i));
                const int loc = static_cast<int>(im.value(*k, 4 + 4 * i + 1));
                const int val = static_cast<int>(im.value(*k, 4 + 4 * i + 3));
                if (loc != 0) continue;  // only inline SHORT values matter here
                if (id == 1024) g.geographic = val == 2;
                if (id == 1025) g.point = val == 2;
                if (id == 3072 && val > 0 && val < 32767) g.epsg = val;
                if (id == 2048 && val > 0 && val < 32767 && g.epsg == 0) g.epsg = val;
            }
        }
    }
    const Entry* sc = im.find(33550);
    const Entry* tp = im.find(33922);
    const Entry* mt = im.find(34264);
    if (sc && tp && sc->count >= 2 && tp->count >= 6) {
        const double sx = im.value(*sc, 0), sy = im.value(*sc, 1);
        const double I = im.value(*tp, 0), J = im.value(*tp, 1), X = im.value(*tp, 3), Y = im.value(*tp, 4);
        g.gt[0] = X - I * sx;
        g.gt[1] = sx;
        g.gt[2] = 0.0;
        g.gt[3] = Y + J * sy;
        g.gt[4] = 0.0;
        g.gt[5] = -sy;
        g.has = true;
    } else if (mt && mt->count >= 16) {
        g.gt[0] = im.value(*mt, 3);
        g.gt[1] = im.value(*mt, 0);
        g.gt[2] = im.value(*mt, 1);
        g.gt[3] = im.value(*mt, 7);
        g.gt[4] = im.value(*mt, 4);
        g.gt[5] = im.value(*mt, 5);
        g.has = true;
    }
    if (g.has && g.point) {
        // GDAL reports a PixelIsPoint raster's corner half a pixel out
        g.gt[0] -= 0.5 * g.gt[1] + 0.5 * g.gt[2];
        g.gt[3] -= 0.5 * g.gt[4] + 0.5 * g.gt[5];
    }
    return g;
}

// lsdtt_data_loader.cpp:1735-1790
void meta_from_georef(const GeoRef& g, const std::string& path, lsdtt_gpu_raster_meta* m)
{
    if (!g.has) {
        m->x_min = 0.0;
        m->cell_size = 1.0;
        m->y_min = 0.0;
        m->cell_size_y = 0.0;
        return;
    }
    const double* gt = g.gt;
    const double ax = std::fabs(gt[1]);
    if (std::fabs(gt[2]) > 1e-9 * ax || std::fabs(gt[4]) > 1e-9 * ax)
        throw std::invalid_argument(path + ": rotated geotransform; lsdtt3 needs north-up rasters");
    if (gt[5] >= 0.0) throw std::invalid_argument(path + ": raster is not north-up");
    if (!g.geographic && std::fabs(gt[1] + gt[5]) > 1e-6 * ax)
        throw std::invalid_argument(path + ": pixels are not square; lsdtt3 needs square pixels");
    m->x_min = gt[0];
    m->cell_size = gt[1];
    m->cell_size_y = std::fabs(gt[1] + gt[5]) > 1e-6 * ax ? -gt[5] : 0.0;
    m->y_min = gt[3] + m->n_rows * gt[5];
}

void read_geotiff(const std::string& path, float* data, uint8_t* mask, lsdtt_gpu_raster_meta* meta)
{
    const std::vector<uint8_t> bytes = read_file(path);
    const TiffImage im(bytes);
    const uint64_t W = im.u(256, 0, true), H = im.u(257, 0, true);
    if (W == 0 || H == 0 || W > (1u << 30) || H > (1u << 30) || W * H >= (1ull << 31))
        throw std::invalid_argument(path + ": unsupported raster size");
    meta->n_cols = static_cast<int>(W);
    meta->n_rows = static_cast<int>(H);
    const GeoRef g = read_georef(im);
    meta_from_georef(g, path, meta);
    meta->epsg = g.epsg;
    meta->geographic = g.geographic ? 1 : 0;
    const std::string nd = im.ascii(42113);
    meta->no_data_value = -9999.0;
    if (!nd.empty()) {
        char* end = nullptr;
        const double v = std::strtod(nd.c_str(), &end);
        if (end != nd.c_str()) meta->no_data_value = v;
    }
    if (!data) return;

    const int spp = static_cast<int>(im.u(277, 1));
    const int planar = static_cast<int>(im.u(284, 1));
    const int bps = static_cast<int>(im.u(258, 1));
    const int fmt = static_cast<int>(im.u(339, 1));
    const int comp = static_cast<int>(im.u(259, 1));
    const int pred = static_cast<int>(im.u(317, 1));
    if (!(bps == 8 || bps == 16 || bps == 32 || bps == 64))
        throw std::invalid_argument(path + ": unsupported bits per sample " + std::to_string(bps));
    if (!(fmt == 1 || fmt == 2 || fmt == 3) || (fmt == 3 && bps < 32))
        throw std::invalid_argument(path + ": unsupported sample format");
    if (!(comp == 1 || comp == 5 || comp == 8 || comp == 32946))
        throw std::invalid_argument(path + ": unsupported compression " + std::to_string(comp));
    const int spp_block = planar == 2 ? 1 : spp;  // samples interleaved in a block
    const bool tiled = im.find(322) != nullptr;
    const uint64_t bw = tiled ? im.u(322, 0, true) : W;
    const uint64_t bh = tiled ? im.u(323, 0, true) : std::min<uint64_t>(im.u(278, H), H);
    if (bw == 0 || bh == 0) throw std::invalid_argument(path + ": bad block size");
    const std::vector<uint64_t> offs = im.list(tiled ? 324 : 273);
    const std::vector<uint64_t> cnts = im.list(tiled ? 325 : 279);
    const uint64_t nbx = (W + bw - 1) / bw, nby = (H + bh - 1) / bh;
    if (offs.size() < nbx * nby || cnts.size() < nbx * nby) throw std::runtime_error(path + ": missing blocks");
    const int bpsample = bps / 8;
    const size_t block_bytes = static_cast<size_t>(bw) * bh * spp_block * bpsample;
    for (uint64_t by = 0; by < nby; ++by)
        for (uint64_t bx = 0; bx < nbx; ++bx) {
            const uint64_t k = by * nbx + bx;  // the first plane's blocks come first
            im.t.need(offs[k], cnts[k]);
            const uint8_t* src = bytes.data() + offs[k];
            std::vector<uint8_t> d;
            if (comp == 1) d.assign(src, src + std::min<uint64_t>(cnts[k], block_bytes));
            else if (comp == 5) d = lzw_decode(src, cnts[k], block_bytes);
            else d = inflate_all(src, cnts[k], block_bytes);
            // a strip at the bottom may hold fewer rows
            const int rows = static_cast<int>(std::min<uint64_t>(bh, H - by * bh));
            const int brows = tiled ? static_cast<int>(bh) : rows;
            undo_predictor(d, pred, static_cast<int>(bw), brows, spp_block, bps, im.t.le);
            for (int y = 0; y < rows; ++y) {
                const uint64_t gy = by * bh + y;
                for (uint64_t x = 0; x < bw; ++x) {
                    const uint64_t gx = bx * bw + x;
                    if (gx >= W) break;
                    const size_t o = ((static_cast<size_t>(y) * bw + x) * spp_block) * bpsample;
                    uint64_t raw = 0;
                    for (int i = 0; i < bpsample; ++i)
                        raw |= static_cast<uint64_t>(d[o + (im.t.le ? i : bpsample - 1 - i)]) << (8 * i);
                    double v;
                    if (fmt == 3) {
                        if (bps == 32) {
                            float f;
                            const uint32_t r32 = static_cast<uint32_t>(raw);
                            std::memcpy(&f, &r32, 4);
                            v = f;
                        } else {
                            std::memcpy(&v, &raw, 8);
                        }
                    } else if (fmt == 2) {
                        const int sh = 64 - bps;
                        v = static_cast<double>(static_cast<int64_t>(raw << sh) >> sh);
                    } else {
                        v = static_cast<double>(raw);
                    }
                    data[gy * W + gx] = static_cast<float>(v);
                }
            }
        }
    mask_nodata(data, mask, static_cast<size_t>(W) * H, meta->no_data_value);
}

// ---------------------------------------------------------------------------
// TIFF writing
// ---------------------------------------------------------------------------
struct TagOut {
    uint16_t tag, type;
    uint32_t count;
    std::vector<uint8_t> data;  // little-endian values
};

template <class T>
void put(std::vector<uint8_t>& v, T x)
{
    uint8_t b[sizeof(T)];
    std::memcpy(b, &x, sizeof(T));
    v.insert(v.end(), b, b + sizeof(T));
}

void write_geotiff(const std::string& path, const float* data, const uint8_t* mask, const lsdtt_gpu_raster_meta& m,
                   int compress)
{
    const int W = m.n_cols, H = m.n_rows;
    const float nd = static_cast<float>(m.no_data_value);
    const int rps = std::max(1, std::min(H, 16384 / std::max(1, W)));  // ~64 KB strips
    const int nstrips = (H + rps - 1) / rps;
    std::vector<std::vector<uint8_t>> strips(nstrips);
    for (int s = 0; s < nstrips; ++s) {
        const int r0 = s * rps, rows = std::min(rps, H - r0);
        std::vector<uint8_t> raw(static_cast<size_t>(rows) * W * 4);
        for (int y = 0; y < rows; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t i = static_cast<size_t>(r0 + y) * W + x;
                const float v = (mask && !mask[i]) ? nd : data[i];
                std::memcpy(&raw[(static_cast<size_t>(y) * W + x) * 4], &v, 4);
            }
        if (!compress) {
            strips[s] = std::move(raw);
            continue;
        }
        // floating-point predictor: per row, byte planes (most significant
        // first), then byte-wise differences along the row
        std::vector<uint8_t> pr(raw.size());
        const size_t row = static_cast<size_t>(W) * 4;
        for (int y = 0; y < rows; ++y) {
            const uint8_t* r = raw.data() + y * row;
            uint8_t* o = pr.data() + y * row;
            for (int x = 0; x < W; ++x)
                for (int b = 0; b < 4; ++b) o[static_cast<size_t>(b) * W + x] = r[static_cast<size_t>(x) * 4 + (3 - b)];
            for (size_t i = row - 1; i >= 1; --i) o[i] = static_cast<uint8_t>(o[i] - o[i - 1]);
        }
        uLongf clen = compressBound(static_cast<uLong>(pr.size()));
        std::vector<uint8_t> c(clen);
        if (compress2(c.data(), &clen, pr.data(), static_cast<uLong>(pr.size()), 6) != Z_OK)
            throw std::runtime_error("GeoTIFF: DEFLATE failed");
        c.resize(clen);
        strips[s] = std::move(c);
    }
    std::vector<TagOut> tags;
    auto tag_u16 = [&](uint16_t t, std::vector<uint16_t> v) {
        TagOut o{t, 3, static_cast<uint32_t>(v.size()), {}};
        for (auto x : v) put(o.data, x);
        tags.push_back(o);
    };
    auto tag_u32 = [&](uint16_t t, std::vector<uint32_t> v) {
        TagOut o{t, 4, static_cast<uint32_t>(v.size()), {}};
        for (auto x : v) put(o.data, x);
        tags.push_back(o);
    };
    auto tag_f64 = [&](uint16_t t, std::vector<double> v) {
        TagOut o{t, 12, static_cast<uint32_t>(v.size()), {}};
        for (auto x : v) put(o.data, x);
        tags.push_back(o);
    };
    const double cy = m.cell_size_y > 0.0 ? m.cell_size_y : m.cell_size;
    tag_u32(256, {static_cast<uint32_t>(W)});
    tag_u32(257, {static_cast<uint32_t>(H)});
    tag_u16(258, {32});
    tag_u16(259, {static_cast<uint16_t>(compress ? 8 : 1)});
    tag_u16(262, {1});
    tag_u32(273, std::vector<uint32_t>(nstrips, 0));  // offsets filled below
    tag_u16(277, {1});
    tag_u32(278, {static_cast<uint32_t>(rps)});
    std::vector<uint32_t> counts(nstrips);
    for (int s = 0; s < nstrips; ++s) counts[s] = static_cast<uint32_t>(strips[s].size());
    tag_u32(279, counts);
    tag_u16(284, {1});
    if (compress) tag_u16(317, {3});
    tag_u16(339, {3});
    tag_f64(33550, {m.cell_size, cy, 0.0});
    tag_f64(33922, {0.0, 0.0, 0.0, m.x_min, m.y_min + H * cy, 0.0});
    {
        std::vector<uint16_t> gk = {1, 1, 0, 0};
        gk.insert(gk.end(), {1024, 0, 1, static_cast<uint16_t>(m.geographic ? 2 : 1)});
        gk.insert(gk.end(), {1025, 0, 1, 1});
        if (m.epsg > 0 && m.epsg < 32767)
            gk.insert(gk.end(), {static_cast<uint16_t>(m.geographic ? 2048 : 3072), 0, 1,
                                 static_cast<uint16_t>(m.epsg)});
        gk[3] = static_cast<uint16_t>((gk.size() - 4) / 4);
        tag_u16(34735, gk);
    }
    {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%.17g", m.no_data_value);
        TagOut o{42113, 2, static_cast<uint32_t>(std::strlen(buf) + 1), {}};
        o.data.assign(buf, buf + std::strlen(buf) + 1);
        tags.push_back(o);
    }
    // layout: header, IFD, out-of-line tag values, strips
    const uint32_t ifd_off = 8;
    const uint32_t ifd_size = 2 + 12 * static_cast<uint32_t>(tags.size()) + 4;
    uint32_t cur = ifd_off + ifd_size;
    std::vector<uint32_t> value_off(tags.size(), 0);
    for (size_t i = 0; i < tags.size(); ++i)
        if (tags[i].data.size() > 4) {
            value_off[i] = cur;
            cur += static_cast<uint32_t>((tags[i].data.size() + 1) & ~size_t(1));
        }
    uint64_t total = cur;
    for (auto& st : strips) total += st.size();
    if (total >= (1ull << 32)) throw std::invalid_argument(path + ": raster too large for a classic TIFF");
    std::vector<uint32_t> soff(nstrips);
    for (int s = 0; s < nstrips; ++s) {
        soff[s] = cur;
        cur += static_cast<uint32_t>(strips[s].size());
    }
    for (auto& t : tags)
        if (t.tag == 273) {
            t.data.clear();
            for (auto x : soff) put(t.data, x);
        }
    std::vector<uint8_t> out;
    out.reserve(total);
    out.push_back('I');
    out.push_back('I');
    put<uint16_t>(out, 42);
    put<uint32_t>(out, ifd_off);
    put<uint16_t>(out, static_cast<uint16_t>(tags.size()));
    for (size_t i = 0; i < tags.size(); ++i) {
        put<uint16_t>(out, tags[i].tag);
        put<uint16_t>(out, tags[i].type);
        put<uint32_t>(out, tags[i].count);
        if (tags[i].data.size() > 4) {
            put<uint32_t>(out, value_off[i]);
        } else {
            std::vector<uint8_t> v = tags[i].data;
            v.resize(4, 0);
            out.insert(out.end(), v.begin(), v.end());
        }
    }
    put<uint32_t>(out, 0);
    for (size_t i = 0; i < tags.size(); ++i)
        if (tags[i].data.size() > 4) {
            out.insert(out.end(), tags[i].data.begin(), tags[i].data.end());
            if (tags[i].data.size() & 1) out.push_back(0);
        }
    for (auto& st : strips) out.insert(out.end(), st.begin(), st.end());
    std::ofstream f(path, std::ios::binary);
    if (!f || !f.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size())))
        throw std::runtime_error("cannot write " + path);
}

// ---------------------------------------------------------------------------
// ESRI ASCII grid
// ---------------------------------------------------------------------------
void read_asc(const std::string& path, float* data, uint8_t* mask, lsdtt_gpu_raster_meta* m)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    long long ncols = -1, nrows = -1;
    double xll = 0.0, yll = 0.0, cs = 0.0, dx = 0.0, dy = 0.0, nd = -9999.0;
    bool xc = false, yc = false;
    std::string key;
    std::streampos body = f.tellg();
    for (;;) {
        body = f.tellg();
        if (!(f >> key)) break;
        std::string k = key;
        for (auto& ch : k) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
        if (!k.empty() && (std::isdigit(static_cast<unsigned char>(k[0])) || k[0] == '-' || k[0] == '+' || k[0] == '.'))
            break;
        double v = 0.0;
        if (!(f >> v)) throw std::invalid_argument(path + ": bad header value for " + key);
        if (k == "ncols") ncols = static_cast<long long>(v);
        else if (k == "nrows") nrows = static_cast<long long>(v);
        else if (k == "xllcorner") xll = v;
        else if (k == "xllcenter") xll = v, xc = true;
        else if (k == "yllcorner") yll = v;
        else if (k == "yllcenter") yll = v, yc = true;
        else if (k == "cellsize") cs = v;
        else if (k == "dx") dx = v;
        else if (k == "dy") dy = v;
        else if (k == "nodata_value") nd = v;
        else throw std::invalid_argument(path + ": unknown header key " + key);
    }
    if (cs == 0.0) cs = dx, dy = dy == 0.0 ? dx : dy;
    if (ncols <= 0 || nrows <= 0 || ncols * nrows >= (1ll << 31) || !(cs > 0.0))
        throw std::invalid_argument(path + ": bad ESRI ASCII header");
    const double csy = dy > 0.0 ? dy : cs;
    if (std::fabs(csy - cs) > 1e-6 * cs) throw std::invalid_argument(path + ": pixels are not square");
    m->n_cols = static_cast<int>(ncols);
    m->n_rows = static_cast<int>(nrows);
    m->cell_size = cs;
    m->cell_size_y = 0.0;
    m->x_min = xc ? xll - 0.5 * cs : xll;
    m->y_min = yc ? yll - 0.5 * cs : yll;
    m->no_data_value = nd;
    m->epsg = 0;
    m->geographic = 0;
    if (!data) return;
    f.clear();
    f.seekg(body);
    const size_t n = static_cast<size_t>(ncols) * nrows;
    for (size_t i = 0; i < n; ++i) {
        std::string tok;
        if (!(f >> tok)) throw std::invalid_argument(path + ": fewer values than ncols x nrows");
        char* end = nullptr;
        errno = 0;
        const double v = std::strtod(tok.c_str(), &end);
        if (end == tok.c_str()) throw std::invalid_argument(path + ": bad value " + tok);
        data[i] = static_cast<float>(v);
    }
    mask_nodata(data, mask, n, nd);
}

void write_asc(const std::string& path, const float* data, const uint8_t* mask, const lsdtt_gpu_raster_meta& m)
{
    if (m.cell_size_y > 0.0 && std::fabs(m.cell_size_y - m.cell_size) > 1e-6 * m.cell_size)
        throw std::invalid_argument(path + ": an ESRI ASCII grid needs square pixels");
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fprintf(f, "ncols %d\nnrows %d\nxllcorner %.17g\nyllcorner %.17g\ncellsize %.17g\nNODATA_value %.17g\n",
                 m.n_cols, m.n_rows, m.x_min, m.y_min, m.cell_size, m.no_data_value);
    const float nd = static_cast<float>(m.no_data_value);
    for (int r = 0; r < m.n_rows; ++r) {
        for (int c = 0; c < m.n_cols; ++c) {
            const size_t i = static_cast<size_t>(r) * m.n_cols + c;
            const float v = (mask && !mask[i]) ? nd : data[i];
            std::fprintf(f, c ? " %.9g" : "%.9g", static_cast<double>(v));
        }
        std::fputc('\n', f);
    }
    if (std::fclose(f) != 0) throw std::runtime_error("cannot write " + path);
}

enum class Fmt { tiff, asc };

Fmt format_of(const std::string& path)
{
    const std::string e = lower_ext(path);
    if (e == "tif" || e == "tiff") return Fmt::tiff;
    if (e == "asc") return Fmt::asc;
    throw std::invalid_argument(path + ": unsupported raster format (GeoTIFF .tif/.tiff or ESRI ASCII .asc)");
}

}  // namespace
}  // namespace lsdtt_gpu

using namespace lsdtt_gpu;

extern "C" int lsdtt_gpu_raster_info(const char* path, lsdtt_gpu_raster_meta* meta)
{
    return guarded([&] {
        if (!path || !meta) return fail_arg("lsdtt_gpu_raster_info: null argument");
        const std::string p(path);
        if (format_of(p) == Fmt::tiff) read_geotiff(p, nullptr, nullptr, meta);
        else read_asc(p, nullptr, nullptr, meta);
        return LSDTT_GPU_OK;
    });
}

extern "C" int lsdtt_gpu_raster_read(const char* path, float* data, uint8_t* mask, lsdtt_gpu_raster_meta* meta)
{
    return guarded([&] {
        if (!path || !data || !meta) return fail_arg("lsdtt_gpu_raster_read: null argument");
        const std::string p(path);
        // the caller sized data from lsdtt_gpu_raster_info: read the header
        // again and refuse a file that changed shape
        lsdtt_gpu_raster_meta want = *meta;
        if (format_of(p) == Fmt::tiff) read_geotiff(p, nullptr, nullptr, meta);
        else read_asc(p, nullptr, nullptr, meta);
        if (meta->n_rows != want.n_rows || meta->n_cols != want.n_cols)
            return fail_arg("lsdtt_gpu_raster_read: meta->n_rows/n_cols must hold the shape from lsdtt_gpu_raster_info");
        if (format_of(p) == Fmt::tiff) read_geotiff(p, data, mask, meta);
        else read_asc(p, data, mask, meta);
        return LSDTT_GPU_OK;
    });
}

extern "C" int lsdtt_gpu_raster_write(const char* path, const float* data, const uint8_t* mask,
                                      const lsdtt_gpu_raster_meta* meta, int compress)
{
    return guarded([&] {
        if (!path || !data || !meta) return fail_arg("lsdtt_gpu_raster_write: null argument");
        if (meta->n_rows <= 0 || meta->n_cols <= 0 || (long long)meta->n_rows * meta->n_cols >= (1ll << 31))
            return fail_arg("lsdtt_gpu_raster_write: bad raster shape");
        if (!(meta->cell_size > 0.0)) return fail_arg("lsdtt_gpu_raster_write: cell_size must be positive");
        const std::string p(path);
        if (format_of(p) == Fmt::tiff) write_geotiff(p, data, mask, *meta, compress);
        else write_asc(p, data, mask, *meta);
        return LSDTT_GPU_OK;
    });
}
