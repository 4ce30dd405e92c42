// fits.h -- a minimal FITS image writer (primary HDU, BITPIX -32) for Camera::ToFITS.
#pragma once
#include <string>
#include <utility>
#include <vector>

namespace birp_api {
struct FitsCard {
    std::string key;
    std::string value;    // already formatted (numbers as text, strings quoted by the writer)
    bool is_string = false;
    std::string comment;  // e.g. "[deg]"
};
// Writes W x H floats (row-major, row 0 first) with the given extra header cards. Returns false
// when the file cannot be written.
bool write_fits(const std::string& path, int W, int H, const float* data, const std::vector<FitsCard>& cards);
}  // namespace birp_api
