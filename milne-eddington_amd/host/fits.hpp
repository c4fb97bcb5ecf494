// fits.hpp — primary-HDU FITS images (the subset MEINV reads and writes:
// Src/IO.c rWavelength, rProfile, rProfilesll, CACHE_INIT, WRITE_RESULT),
// written without CFITSIO: 2880-byte blocks of 80-character header cards,
// big-endian data, BITPIX 8/16/32/64/-32/-64, BSCALE/BZERO on read.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mefits {

struct Image {
    std::string path;
    int bitpix = 0;
    std::vector<int64_t> naxes;  // NAXIS1 first (the fastest axis)
    double bscale = 1.0, bzero = 0.0;
    int64_t data_offset = 0;     // first data byte in the file
    int64_t npix() const;
};

// Parse the primary header.  False (with a message) if the file is not a
// readable primary image.
bool open_image(const std::string &path, Image &img, std::string &err);

// Pixels [first, first + count) in file order as doubles, BSCALE/BZERO applied.
bool read_pixels(const Image &img, int64_t first, int64_t count, double *out, std::string &err);

// Create (or replace) a BITPIX -64 image with the given axes; the data are
// zero.  On success img describes the new file.
bool create_image(const std::string &path, const std::vector<int64_t> &naxes, Image &img, std::string &err);

// Overwrite pixels [first, first + count) of a BITPIX -64 image.
bool write_pixels(const Image &img, int64_t first, int64_t count, const double *in, std::string &err);

}  // namespace mefits
