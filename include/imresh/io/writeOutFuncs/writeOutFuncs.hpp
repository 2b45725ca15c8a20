// imresh::io::writeOutFuncs -- drop-in for the reference declarations
// (src/imresh/io/writeOutFuncs/writeOutFuncs.hpp): callbacks of the
// imresh::io::WriteOutCallBack shape for addTask.  justFree delete[]s the
// buffer; writeOutPNG writes a 16-bit RGB PNG with every pixel scaled by the
// image maximum (R = G = B = int(v/max*65535 + 0.5f)), image row iy as file
// row H-1-iy like pngwriter; writeOutAndFreePNG does both.  The HDF5 writers
// (libSplash) are not provided.
#pragma once

#include <string>

namespace imresh
{
namespace io
{
namespace writeOutFuncs
{

    void justFree
    (
        float * rMem,
        unsigned int const rImageWidth,
        unsigned int const rImageHeight,
        std::string const rFilename
    );

    void writeOutPNG
    (
        float * rMem,
        unsigned int const rImageWidth,
        unsigned int const rImageHeight,
        std::string const rFileName
    );

    void writeOutAndFreePNG
    (
        float * rMem,
        unsigned int const rImageWidth,
        unsigned int const rImageHeight,
        std::string const rFileName
    );

} // namespace writeOutFuncs
} // namespace io
} // namespace imresh
