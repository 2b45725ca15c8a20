// imresh::io::readInFuncs -- drop-in for the reference declarations
// (src/imresh/io/readInFuncs/readInFuncs.hpp): each reader returns a buffer
// allocated with new[] (release with delete[], e.g. by the write-out
// callbacks) and its (width, height), or {NULL, {0, 0}} when the file cannot
// be read.  readPNG: red channel / 65535, image row iy = file row H-1-iy as
// through pngwriter; readTiff: 32-bit IEEE float samples; readTxt:
// whitespace-separated floats, one row per line.  readHDF5 (libSplash) is
// not provided.
#pragma once

#include <string>
#include <utility>

namespace imresh
{
namespace io
{
namespace readInFuncs
{

    using ImageDimensions = std::pair< unsigned int, unsigned int >;
    using FloatImage = std::pair< float *, ImageDimensions >;

    FloatImage readTxt( std::string const filename );
    FloatImage readPNG( std::string const filename );
    FloatImage readTiff( std::string const filename );

} // namespace readInFuncs
} // namespace io
} // namespace imresh
