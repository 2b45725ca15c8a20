// File formats on either side of the reconstruction: the readers that turn a
// measurement file into the buffer shrinkWrap / addTask take, and the writers
// the task queue's callbacks use (src/imresh/io/readInFuncs/readInFuncs.cpp,
// src/imresh/io/writeOutFuncs/writeOutFuncs.cpp).  Host code only.
//
//   PNG   chunk layout, CRC-32, zlib stream and the five row filters of the
//         PNG specification (ISO/IEC 15948); the compression itself is zlib's.
//         The reference goes through pngwriter, whose pixel (x, y) is 1-based
//         with y = 1 the bottom row: image row iy of a buffer is file row
//         H-1-iy, for reading and writing alike.  Reading takes the red
//         channel as a 16-bit value (8-bit samples times 256, as pngwriter's
//         read(); gray counts as red, palettes are looked up) over 65535.
//         Writing produces 16-bit RGB with R = G = B = int(v/max*65535 + 0.5f)
//         clamped to [0, 65535] (writeOutFuncs.cpp:69-114).
//   TIFF  baseline TIFF 6.0, one 32-bit IEEE float sample per pixel,
//         uncompressed strips, either byte order (the reference reads this
//         layout through libtiff, ReadTiff.cpp).
//   text  whitespace-separated floats, one image row per line (readTxt,
//         readInFuncs.cpp:56-104).
// HDF5 (readHDF5 / writeOutHDF5 through libSplash) has no counterpart: the
// HDF5 and libSplash libraries are not available to this build.
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <new>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "imresh/io/readInFuncs/readInFuncs.hpp"
#include "imresh/io/writeOutFuncs/writeOutFuncs.hpp"
#include "imresh_hip.h"

namespace {

bool read_file(const char *path, std::vector<uint8_t> &bytes)
{
    std::FILE *f = std::fopen(path, "rb");
    if (f == nullptr) return false;
    bytes.clear();
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) bytes.insert(bytes.end(), buf, buf + got);
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok;
}

uint32_t be32(const uint8_t *p) { return (uint32_t) p[0] << 24 | (uint32_t) p[1] << 16 | (uint32_t) p[2] << 8 | p[3]; }

void put_be32(std::vector<uint8_t> &v, uint32_t x)
{
    for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t) (x >> s));
}

// ------------------------------------------------------------------ PNG
const uint8_t kPngSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};

int paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// the image as floats, row iy = file row H-1-iy; false on any error
bool png_read(const char *path, std::vector<float> &out, uint32_t &W, uint32_t &H)
{
    std::vector<uint8_t> f;
    if (!read_file(path, f) || f.size() < 8 || std::memcmp(f.data(), kPngSig, 8) != 0) return false;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte;
    bool have_ihdr = false, end = false;
    for (size_t p = 8; !end;) {
        if (p + 12 > f.size()) return false;
        const uint32_t len = be32(&f[p]);
        if (len > f.size() - p - 12) return false;
        const uint8_t *type = &f[p + 4], *data = &f[p + 8];
        const uint32_t crc = be32(&f[p + 8 + len]);
        if ((uint32_t) crc32(crc32(0L, Z_NULL, 0), type, len + 4) != crc) return false;
        if (std::memcmp(type, "IHDR", 4) == 0) {
            if (len != 13) return false;
            w = be32(data);
            h = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            if (data[10] != 0 || data[11] != 0) return false;
            interlace = data[12];
            have_ihdr = true;
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            plte.assign(data, data + len);
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), data, data + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            end = true;
        } else if (!(type[0] & 0x20)) {
            return false;   // unknown critical chunk
        }
        p += 12 + (size_t) len;
    }
    // sizes come from the file: at most 2^31 pixels (the library's sizes are
    // 32-bit), so a corrupt header cannot request an absurd allocation
    if (!have_ihdr || w == 0 || h == 0 || interlace != 0 || (uint64_t) w * h > (1ull << 31)) return false;
    int channels;
    switch (ctype) {
    case 0: channels = 1; break;   // gray
    case 2: channels = 3; break;   // RGB
    case 3: channels = 1; break;   // palette
    case 4: channels = 2; break;   // gray + alpha
    case 6: channels = 4; break;   // RGBA
    default: return false;
    }
    const bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!depth_ok || (ctype == 3 && (plte.empty() || plte.size() % 3 != 0))) return false;
    const size_t bits_px = (size_t) channels * depth;
    const size_t row_bytes = ((size_t) w * bits_px + 7) / 8;
    const size_t bpp = bits_px >= 8 ? bits_px / 8 : 1;   // filter byte distance
    if ((size_t) h > std::numeric_limits<size_t>::max() / (row_bytes + 1)) return false;
    // deflate expands at most 1032:1, so a header that asks for more than
    // the compressed data can hold is corrupt: rejected before the buffer is
    // allocated (and zero-filled) rather than after inflating fails
    if ((row_bytes + 1) * h > idat.size() * 1032 + 1024) return false;
    std::vector<uint8_t> raw((row_bytes + 1) * h);
    uLongf raw_len = (uLongf) raw.size();
    if (uncompress(raw.data(), &raw_len, idat.data(), (uLong) idat.size()) != Z_OK || raw_len != raw.size())
        return false;
    std::vector<uint8_t> prev(row_bytes, 0), cur(row_bytes);
    out.assign((size_t) w * h, 0.f);
    for (uint32_t r = 0; r < h; ++r) {
        const uint8_t *src = &raw[r * (row_bytes + 1)];
        const int ft = src[0];
        for (size_t i = 0; i < row_bytes; ++i) {
            const int x = src[1 + i];
            const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v;
            switch (ft) {
            case 0: v = x; break;
            case 1: v = x + a; break;
            case 2: v = x + b; break;
            case 3: v = x + ((a + b) >> 1); break;
            case 4: v = x + paeth(a, b, c); break;
            default: return false;
            }
            cur[i] = (uint8_t) v;
        }
        float *dst = &out[(size_t) (h - 1 - r) * w];
        for (uint32_t x = 0; x < w; ++x) {
            int v16;
            if (ctype == 3 || depth < 8) {
                const size_t bit = (size_t) x * depth;
                const int s = (cur[bit / 8] >> (8 - depth - (int) (bit % 8))) & ((1 << depth) - 1);
                if (ctype == 3) {
                    if ((size_t) s * 3 + 2 >= plte.size()) return false;
                    v16 = plte[(size_t) s * 3] * 256;
                } else {
                    v16 = s * (255 / ((1 << depth) - 1)) * 256;   // gray expanded to 8 bits
                }
            } else if (depth == 8) {
                v16 = cur[(size_t) x * channels] * 256;
            } else {
                const size_t o = (size_t) x * channels * 2;
                v16 = cur[o] * 256 + cur[o + 1];
            }
            dst[x] = (float) v16 / 65535;
        }
        std::swap(prev, cur);
    }
    W = w;
    H = h;
    return true;
}

void png_chunk(std::vector<uint8_t> &out, const char *type, const uint8_t *data, size_t len)
{
    put_be32(out, (uint32_t) len);
    const size_t t0 = out.size();
    out.insert(out.end(), type, type + 4);
    if (len) out.insert(out.end(), data, data + len);
    put_be32(out, (uint32_t) crc32(crc32(0L, Z_NULL, 0), &out[t0], (uInt) (len + 4)));
}

bool png_write(const char *path, const float *mem, uint32_t W, uint32_t H)
{
    // max as algorithms::vectorMax (std::max from lowest()), 0 -> 1 (:79-83)
    float mx = std::numeric_limits<float>::lowest();
    const size_t n = (size_t) W * H;
    for (size_t i = 0; i < n; ++i) mx = (mx < mem[i]) ? mem[i] : mx;
    if (mx == 0) mx = 1;
    const size_t row_bytes = (size_t) W * 6;
    std::vector<uint8_t> raw((row_bytes + 1) * H);
    for (uint32_t r = 0; r < H; ++r) {
        uint8_t *row = &raw[r * (row_bytes + 1)];
        row[0] = 0;   // filter: none
        const float *src = mem + (size_t) (H - 1 - r) * W;
        for (uint32_t x = 0; x < W; ++x) {
            const float value = src[x] / mx;
            const float scaled = value * 65535 + 0.5f;
            int v = 0;
            if (scaled >= 65535.f) v = 65535;
            else if (scaled > 0.f) v = (int) scaled;   // NaN and negatives -> 0
            for (int c = 0; c < 3; ++c) {
                row[1 + x * 6 + c * 2] = (uint8_t) (v >> 8);
                row[1 + x * 6 + c * 2 + 1] = (uint8_t) v;
            }
        }
    }
    uLongf zlen = compressBound((uLong) raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong) raw.size(), 6) != Z_OK) return false;
    std::vector<uint8_t> file(kPngSig, kPngSig + 8);
    uint8_t ihdr[13];
    for (int i = 0; i < 4; ++i) {
        ihdr[i] = (uint8_t) (W >> (24 - 8 * i));
        ihdr[4 + i] = (uint8_t) (H >> (24 - 8 * i));
    }
    ihdr[8] = 16;   // bit depth
    ihdr[9] = 2;    // RGB
    ihdr[10] = ihdr[11] = ihdr[12] = 0;
    png_chunk(file, "IHDR", ihdr, 13);
    png_chunk(file, "IDAT", z.data(), zlen);
    png_chunk(file, "IEND", nullptr, 0);
    std::FILE *f = std::fopen(path, "wb");
    if (f == nullptr) return false;
    const bool ok = std::fwrite(file.data(), 1, file.size(), f) == file.size();
    return (std::fclose(f) == 0) && ok;
}

// ------------------------------------------------------------------ TIFF
bool tiff_read(const char *path, std::vector<float> &out, uint32_t &W, uint32_t &H)
{
    std::vector<uint8_t> f;
    if (!read_file(path, f) || f.size() < 8) return false;
    bool be;
    if (f[0] == 'I' && f[1] == 'I') be = false;
    else if (f[0] == 'M' && f[1] == 'M') be = true;
    else return false;
    auto u16 = [&](size_t o) -> uint32_t {
        return be ? (uint32_t) f[o] << 8 | f[o + 1] : (uint32_t) f[o + 1] << 8 | f[o];
    };
    auto u32 = [&](size_t o) -> uint32_t {
        return be ? be32(&f[o]) : (uint32_t) f[o + 3] << 24 | (uint32_t) f[o + 2] << 16 | (uint32_t) f[o + 1] << 8 | f[o];
    };
    if (u16(2) != 42) return false;
    const size_t ifd = u32(4);
    if (ifd + 2 > f.size()) return false;
    const uint32_t count = u16(ifd);
    if (ifd + 2 + (size_t) count * 12 > f.size()) return false;
    uint32_t w = 0, h = 0, bits = 1, comp = 1, spp = 1, planar = 1, fmt = 1, rps = 0xffffffffu;
    std::vector<uint32_t> offsets, counts;
    auto values = [&](size_t e, std::vector<uint32_t> &v) -> bool {
        const uint32_t type = u16(e + 2), n = u32(e + 4);
        const size_t size = type == 3 ? 2 : type == 4 ? 4 : 0;
        if (size == 0 || n == 0 || n > (1u << 24)) return false;
        size_t o = n * size <= 4 ? e + 8 : u32(e + 8);
        if (o + n * size > f.size()) return false;
        v.resize(n);
        for (uint32_t i = 0; i < n; ++i) v[i] = size == 2 ? u16(o + 2 * i) : u32(o + 4 * i);
        return true;
    };
    for (uint32_t i = 0; i < count; ++i) {
        const size_t e = ifd + 2 + (size_t) i * 12;
        std::vector<uint32_t> v;
        const uint32_t tag = u16(e);
        if (tag == 256 || tag == 257 || tag == 258 || tag == 259 || tag == 273 || tag == 277 || tag == 278 ||
            tag == 279 || tag == 284 || tag == 339) {
            if (!values(e, v)) return false;
        }
        switch (tag) {
        case 256: w = v[0]; break;
        case 257: h = v[0]; break;
        case 258: bits = v[0]; break;
        case 259: comp = v[0]; break;
        case 273: offsets = v; break;
        case 277: spp = v[0]; break;
        case 278: rps = v[0]; break;
        case 279: counts = v; break;
        case 284: planar = v[0]; break;
        case 339: fmt = v[0]; break;
        default: break;
        }
    }
    // the layout the reference asserts: IEEE float samples (ReadTiff.cpp),
    // one per pixel, uncompressed strips
    if (w == 0 || h == 0 || bits != 32 || fmt != 3 || spp != 1 || comp != 1 || planar != 1) return false;
    if ((uint64_t) w * h > (1ull << 31)) return false;
    if (offsets.empty() || offsets.size() != counts.size()) return false;
    const size_t need = (size_t) w * h * 4;
    size_t have = 0;   // strip bytes, all inside the file
    for (size_t s = 0; s < offsets.size(); ++s) {
        if ((size_t) offsets[s] + counts[s] > f.size()) return false;
        have += counts[s];
    }
    if (have < need) return false;
    std::vector<uint8_t> buf;
    buf.reserve(have);
    for (size_t s = 0; s < offsets.size(); ++s)
        buf.insert(buf.end(), f.begin() + offsets[s], f.begin() + offsets[s] + counts[s]);
    (void) rps;
    out.resize((size_t) w * h);
    for (size_t i = 0; i < out.size(); ++i) {
        uint8_t b[4];
        for (int k = 0; k < 4; ++k) b[k] = buf[4 * i + (be ? 3 - k : k)];
        std::memcpy(&out[i], b, 4);   // host is little-endian (x86-64)
    }
    W = w;
    H = h;
    return true;
}

// ------------------------------------------------------------------ text
bool txt_read(const char *path, std::vector<float> &out, uint32_t &W, uint32_t &H)
{
    std::ifstream file(path);
    if (!file.is_open()) return false;
    std::string line;
    if (!std::getline(file, line)) return false;
    size_t cols = 0;
    {
        std::istringstream ls(line);
        std::string tok;
        while (ls >> tok) ++cols;
    }
    if (cols == 0) return false;
    file.clear();
    file.seekg(0, std::ios::beg);
    std::vector<float> vals;
    std::string tok;
    while (file >> tok) {
        char *end = nullptr;
        const float v = std::strtof(tok.c_str(), &end);
        if (end == tok.c_str() || *end != '\0') return false;
        vals.push_back(v);
    }
    const size_t rows = vals.size() / cols;   // the reference's y = count / x
    if (rows == 0) return false;
    vals.resize(rows * cols);
    out.swap(vals);
    W = (uint32_t) cols;
    H = (uint32_t) rows;
    return true;
}

using Reader = bool (*)(const char *, std::vector<float> &, uint32_t &, uint32_t &);

int read_entry(Reader rd, const char *path, float *out, size_t capacity, uint32_t *width, uint32_t *height)
{
    if (path == nullptr || width == nullptr || height == nullptr) return IMRESH_ERR_NULL;
    std::vector<float> img;
    uint32_t w = 0, h = 0;
    try {
        if (!rd(path, img, w, h)) return IMRESH_ERR_IO;
    } catch (const std::bad_alloc &) {
        return IMRESH_ERR_OUT_OF_MEMORY;
    } catch (...) {
        return IMRESH_ERR_IO;
    }
    *width = w;
    *height = h;
    if (out == nullptr) return IMRESH_OK;   // size query
    if (capacity < img.size()) return IMRESH_ERR_INVALID;
    std::memcpy(out, img.data(), sizeof(float) * img.size());
    return IMRESH_OK;
}

imresh::io::readInFuncs::FloatImage to_image(Reader rd, const std::string &path)
{
    std::vector<float> img;
    uint32_t w = 0, h = 0;
    try {
        if (!rd(path.c_str(), img, w, h)) return {nullptr, {0, 0}};
    } catch (...) {
        return {nullptr, {0, 0}};
    }
    float *mem = new float[img.size()];
    std::memcpy(mem, img.data(), sizeof(float) * img.size());
    return {mem, {w, h}};
}

}  // namespace

extern "C" {

int imresh_read_png(const char *path, float *out, size_t capacity, uint32_t *width, uint32_t *height)
{
    return read_entry(png_read, path, out, capacity, width, height);
}

int imresh_read_tiff(const char *path, float *out, size_t capacity, uint32_t *width, uint32_t *height)
{
    return read_entry(tiff_read, path, out, capacity, width, height);
}

int imresh_read_txt(const char *path, float *out, size_t capacity, uint32_t *width, uint32_t *height)
{
    return read_entry(txt_read, path, out, capacity, width, height);
}

int imresh_write_png(const char *path, const float *data, uint32_t width, uint32_t height)
{
    if (path == nullptr || data == nullptr) return IMRESH_ERR_NULL;
    if (width == 0 || height == 0) return IMRESH_ERR_INVALID;
    try {
        return png_write(path, data, width, height) ? IMRESH_OK : IMRESH_ERR_IO;
    } catch (const std::bad_alloc &) {
        return IMRESH_ERR_OUT_OF_MEMORY;
    }
}

}  // extern "C"

namespace imresh {
namespace io {
namespace readInFuncs {

FloatImage readTxt(std::string const rFilename) { return to_image(txt_read, rFilename); }
FloatImage readPNG(std::string const rFilename) { return to_image(png_read, rFilename); }
FloatImage readTiff(std::string const rFilename) { return to_image(tiff_read, rFilename); }

}  // namespace readInFuncs

namespace writeOutFuncs {

void justFree(float *rMem, unsigned int const, unsigned int const, std::string const)
{
    delete[] rMem;
}

void writeOutPNG(float *rMem, unsigned int const rImageWidth, unsigned int const rImageHeight,
                 std::string const rFileName)
{
    if (rMem == nullptr || rImageWidth == 0 || rImageHeight == 0) return;
    try {
        png_write(rFileName.c_str(), rMem, rImageWidth, rImageHeight);
    } catch (const std::bad_alloc &) {
        // the reference's pngwriter reports nothing either; the image is not written
    }
}

void writeOutAndFreePNG(float *rMem, unsigned int const rImageWidth, unsigned int const rImageHeight,
                        std::string const rFileName)
{
    writeOutPNG(rMem, rImageWidth, rImageHeight, rFileName);
    justFree(rMem, rImageWidth, rImageHeight, rFileName);
}

}  // namespace writeOutFuncs
}  // namespace io
}  // namespace imresh
