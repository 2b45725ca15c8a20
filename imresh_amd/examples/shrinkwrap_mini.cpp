// The reference's minimal application flow (examples/miniExample.cpp:93-116)
// against this library's drop-in header alone: one direct call of
// imresh::algorithms::shrinkWrap on a host buffer holding the Fourier
// modulus, which it overwrites with the reconstruction.
//
//   imresh_shrinkwrap_example <modulus.f32> <width> <height> <out prefix> <cycles>
//
// The input is width*height raw float32 values.  Three calls, whose outcome
// is printed and written next to the prefix:
//   null      shrinkWrap(nullptr, W, H)                      -> prints its return (1)
//   <p>a.f32  shrinkWrap(a, W, H, cycles, 1e-6)               (miniExample's call)
//   <p>b.f32  shrinkWrap(b, W, H, cycles, 1e-6, 0, 0, 0, 0, 0, 0)
//             every parameter <= 0 / == 0 selects its default (shrinkWrap.cpp:133-140)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "imresh/algorithms/shrinkWrap.hpp"

static bool read_raw(const char *path, std::vector<float> &v)
{
    FILE *f = std::fopen(path, "rb");
    if (f == nullptr) return false;
    const size_t got = std::fread(v.data(), sizeof(float), v.size(), f);
    std::fclose(f);
    return got == v.size();
}

static bool write_raw(const std::string &path, const std::vector<float> &v)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (f == nullptr) return false;
    const size_t put = std::fwrite(v.data(), sizeof(float), v.size(), f);
    std::fclose(f);
    return put == v.size();
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s modulus.f32 width height out_prefix cycles\n", argv[0]);
        return 2;
    }
    const unsigned W = (unsigned) std::atoi(argv[2]), H = (unsigned) std::atoi(argv[3]);
    const std::string prefix = argv[4];
    const unsigned cycles = (unsigned) std::atoi(argv[5]);
    std::vector<float> a((size_t) W * H);
    if (!read_raw(argv[1], a)) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 1;
    }
    std::vector<float> b = a;

    const int rc_null = imresh::algorithms::shrinkWrap(nullptr, W, H);
    std::printf("null %d\n", rc_null);

    const int rc_a = imresh::algorithms::shrinkWrap(a.data(), W, H, cycles, 1e-6f);
    const int rc_b = imresh::algorithms::shrinkWrap(b.data(), W, H, cycles, 1e-6f, 0, 0, 0, 0, 0, 0);
    std::printf("a %d\nb %d\n", rc_a, rc_b);
    if (rc_a != 0 || rc_b != 0) return 1;
    if (!write_raw(prefix + "a.f32", a) || !write_raw(prefix + "b.f32", b)) return 1;
    std::printf("done\n");
    return 0;
}
