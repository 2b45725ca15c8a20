// Structure-aware fuzzer of the PNG / TIFF / text readers (host code of the
// library, io.cpp), built with -fsanitize=address,undefined by
// tests/test_io_fuzz.py: every mutated file must be read or rejected with a
// status, and any memory error aborts the run.
//
//   io_fuzz SEED_DIR ITERATIONS RNG_SEED WORK_PREFIX
//
// PNG mutations act on the structure (IHDR fields, the inflated image data
// and its filter bytes, the palette, chunk order) with CRCs recomputed, so
// they reach the decoder instead of stopping at a checksum; every format
// also gets byte-level mutations (bit flips, overwrites, truncation, large
// 32-bit values, duplicated spans).
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <fstream>
#include <iterator>
#include <random>
#include <string>
#include <vector>

#include "imresh_hip.h"

using Bytes = std::vector<uint8_t>;
static std::mt19937_64 rng(12345);

static uint32_t be32(const uint8_t *p) { return (uint32_t) p[0] << 24 | p[1] << 16 | p[2] << 8 | p[3]; }
static void put32(uint8_t *p, uint32_t x)
{
    p[0] = x >> 24;
    p[1] = x >> 16;
    p[2] = x >> 8;
    p[3] = x;
}
static size_t rnd(size_t n) { return n ? rng() % n : 0; }

static Bytes load(const std::string &p)
{
    std::ifstream f(p, std::ios::binary);
    return Bytes((std::istreambuf_iterator<char>(f)), {});
}
static void save(const std::string &p, const Bytes &b)
{
    std::ofstream f(p, std::ios::binary);
    f.write((const char *) b.data(), (std::streamsize) b.size());
}

static void byte_mutate(Bytes &b)
{
    const int k = 1 + (int) rnd(8);
    for (int i = 0; i < k && !b.empty(); ++i) {
        switch (rnd(6)) {
        case 0: b[rnd(b.size())] ^= (uint8_t) (1u << rnd(8)); break;
        case 1: b[rnd(b.size())] = (uint8_t) rng(); break;
        case 2: b[rnd(b.size())] = rnd(2) ? 0xff : 0; break;
        case 3: b.resize(rnd(b.size() + 1)); break;
        case 4: if (b.size() >= 4) { size_t p = rnd(b.size() - 3); put32(&b[p], (uint32_t) rng() >> rnd(32)); } break;
        case 5: { size_t p = rnd(b.size()), n = rnd(64); Bytes c(b.begin() + p, b.begin() + std::min(b.size(), p + n)); b.insert(b.begin() + rnd(b.size()), c.begin(), c.end()); } break;
        }
    }
}

struct Chunk {
    std::string type;
    Bytes data;
};

static bool png_parse(const Bytes &f, std::vector<Chunk> &ch)
{
    if (f.size() < 8) return false;
    for (size_t p = 8; p + 12 <= f.size();) {
        const uint32_t len = be32(&f[p]);
        if (len > f.size() - p - 12) return false;
        ch.push_back({std::string((const char *) &f[p + 4], 4), Bytes(f.begin() + p + 8, f.begin() + p + 8 + len)});
        p += 12 + len;
    }
    return true;
}

static Bytes png_build(const std::vector<Chunk> &ch)
{
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    Bytes out(sig, sig + 8);
    for (const Chunk &c : ch) {
        uint8_t h[8];
        put32(h, (uint32_t) c.data.size());
        std::memcpy(h + 4, c.type.data(), 4);
        out.insert(out.end(), h, h + 8);
        out.insert(out.end(), c.data.begin(), c.data.end());
        uLong crc = crc32(0L, Z_NULL, 0);
        crc = crc32(crc, (const Bytef *) c.type.data(), 4);
        if (!c.data.empty()) crc = crc32(crc, c.data.data(), (uInt) c.data.size());
        uint8_t t[4];
        put32(t, (uint32_t) crc);
        out.insert(out.end(), t, t + 4);
    }
    return out;
}

// mutate IHDR fields, raw (inflated) image data, chunk order; CRCs recomputed
static Bytes png_mutate(const Bytes &seed)
{
    std::vector<Chunk> ch;
    if (!png_parse(seed, ch)) return seed;
    Bytes idat;
    for (auto &c : ch)
        if (c.type == "IDAT") idat.insert(idat.end(), c.data.begin(), c.data.end());
    uLongf rawn = 1 << 20;
    Bytes raw(rawn);
    if (uncompress(raw.data(), &rawn, idat.data(), idat.size()) != Z_OK) return seed;
    raw.resize(rawn);
    const int what = (int) rnd(5);
    if (what == 0 || what == 4) {   // IHDR fields
        for (auto &c : ch)
            if (c.type == "IHDR" && c.data.size() == 13) {
                switch (rnd(6)) {
                case 0: put32(&c.data[0], (uint32_t) rnd(40)); break;
                case 1: put32(&c.data[4], (uint32_t) rnd(40)); break;
                case 2: c.data[8] = (uint8_t) (1u << rnd(5)); break;
                case 3: c.data[9] = (uint8_t) rnd(7); break;
                case 4: put32(&c.data[rnd(2) * 4], (uint32_t) rng()); break;
                case 5: c.data[8] = (uint8_t) rng(); break;
                }
            }
    }
    if (what == 1 || what == 4) {   // raw data: filter bytes, samples, size
        byte_mutate(raw);
        if (!raw.empty() && rnd(2)) raw[rnd(raw.size())] = (uint8_t) rnd(6);
    }
    if (what == 2) {   // palette size / removal
        for (auto &c : ch)
            if (c.type == "PLTE") c.data.resize(rnd(60));
    }
    if (what == 3 && ch.size() > 2) std::swap(ch[rnd(ch.size())], ch[rnd(ch.size())]);
    uLongf cn = compressBound(raw.size());
    Bytes comp(cn);
    compress(comp.data(), &cn, raw.data(), raw.size());
    comp.resize(cn);
    std::vector<Chunk> out;
    bool put = false;
    for (auto &c : ch) {
        if (c.type == "IDAT") {
            if (!put) out.push_back({"IDAT", comp});
            put = true;
        } else {
            out.push_back(c);
        }
    }
    return png_build(out);
}

static int read_one(int kind, const char *path)
{
    uint32_t w = 0, h = 0;
    int st = kind == 0 ? imresh_read_png(path, nullptr, 0, &w, &h)
           : kind == 1 ? imresh_read_tiff(path, nullptr, 0, &w, &h)
                       : imresh_read_txt(path, nullptr, 0, &w, &h);
    if (st != IMRESH_OK) return st;
    std::vector<float> buf((size_t) w * h);
    st = kind == 0 ? imresh_read_png(path, buf.data(), buf.size(), &w, &h)
       : kind == 1 ? imresh_read_tiff(path, buf.data(), buf.size(), &w, &h)
                   : imresh_read_txt(path, buf.data(), buf.size(), &w, &h);
    return st;
}

int main(int argc, char **argv)
{
    if (argc != 5) {
        std::fprintf(stderr, "usage: io_fuzz SEED_DIR ITERATIONS RNG_SEED WORK_PREFIX\n");
        return 2;
    }
    const std::string dir = std::string(argv[1]) + "/", cur = argv[4];
    const int iters = atoi(argv[2]);
    rng.seed(strtoull(argv[3], nullptr, 10));
    const char *seeds[] = {"s0.png", "s1.png", "s2.png", "s3.png", "s4.png", "s5.png", "s6.png", "s7.png",
                           "s8.png", "s9.png", "s10.png", "s11.png", "s12.png", "s13.png", "s14.png",
                           "le.tif", "be.tif", "a.txt"};
    long ok = 0, rejected = 0;
    for (const char *s : seeds) {
        const std::string name(s);
        const std::string ext = name.substr(name.size() - 4);
        const int kind = ext == ".png" ? 0 : ext == ".tif" ? 1 : 2;
        const Bytes seed = load(dir + name);
        if (seed.empty()) {
            std::fprintf(stderr, "missing seed %s\n", name.c_str());
            return 2;
        }
        for (int i = 0; i < iters; ++i) {
            Bytes b = kind == 0 && rnd(3) ? png_mutate(seed) : seed;
            if (kind != 0 || rnd(4) == 0) byte_mutate(b);
            const std::string p = cur + (kind == 0 ? ".png" : kind == 1 ? ".tif" : ".txt");
            save(p, b);
            (read_one(kind, p.c_str()) == IMRESH_OK ? ok : rejected)++;
        }
    }
    std::printf("fuzz done: %ld read, %ld rejected\n", ok, rejected);
    return 0;
}
