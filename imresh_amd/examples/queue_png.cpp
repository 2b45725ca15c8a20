// The reference's application flow (examples/threadedExample.cpp,
// examples/miniExample.cpp) written against this library's drop-in C++ API
// only: read an object image, compute its diffraction intensity, queue one
// reconstruction per copy with addTask and let the write-out callback store
// each result as a PNG and free the buffer.
//
//   imresh_queue_example <object.png> <output prefix> <tasks> [cycles]
//
// Prints "done <tasks>" when taskQueueDeinit has returned.  Built by the
// package build next to libimresh_hip.so (rpath $ORIGIN).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "imresh/io/readInFuncs/readInFuncs.hpp"
#include "imresh/io/taskQueue.hpp"
#include "imresh/io/writeOutFuncs/writeOutFuncs.hpp"
#include "imresh/libs/diffractionIntensity.hpp"

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s object.png out_prefix tasks [cycles]\n", argv[0]);
        return 2;
    }
    const std::string prefix = argv[2];
    const int tasks = std::atoi(argv[3]);
    const unsigned cycles = argc > 4 ? (unsigned) std::atoi(argv[4]) : 20u;
    auto file = imresh::io::readInFuncs::readPNG(argv[1]);
    if (file.first == nullptr) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 1;
    }
    const unsigned W = file.second.first, H = file.second.second;
    imresh::io::taskQueueInit();
    for (int i = 0; i < tasks; ++i) {
        float *io = new float[(size_t) W * H];
        std::memcpy(io, file.first, sizeof(float) * W * H);
        imresh::libs::diffractionIntensity(io, W, H);
        imresh::io::addTask(io, W, H, imresh::io::writeOutFuncs::writeOutAndFreePNG,
                            prefix + std::to_string(i) + ".png", cycles);
    }
    imresh::io::taskQueueDeinit();
    delete[] file.first;
    std::printf("done %d\n", tasks);
    return 0;
}
