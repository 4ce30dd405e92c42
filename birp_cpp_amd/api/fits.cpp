// fits.cpp -- FITS (NASA/IAU standard 4.0) primary image: 80-character header cards in
// 2880-byte blocks, END card, then the data as big-endian IEEE floats padded to 2880 bytes.
#include "fits.h"

#include <cstdint>
#include <cstdio>
#include <cstring>

namespace birp_api {

namespace {
std::string card(const std::string& key, const std::string& value, const std::string& comment) {
    char buf[81];
    std::string k = key.substr(0, 8);
    for (auto& c : k) c = (char)std::toupper((unsigned char)c);
    if (comment.empty()) std::snprintf(buf, sizeof buf, "%-8s= %20s", k.c_str(), value.c_str());
    else std::snprintf(buf, sizeof buf, "%-8s= %20s / %s", k.c_str(), value.c_str(), comment.c_str());
    std::string s(buf);
    s.resize(80, ' ');
    return s;
}
}  // namespace

bool write_fits(const std::string& path, int W, int H, const float* data, const std::vector<FitsCard>& cards) {
    std::string hdr;
    hdr += card("SIMPLE", "T", "");
    hdr += card("BITPIX", "-32", "");
    hdr += card("NAXIS", "2", "");
    hdr += card("NAXIS1", std::to_string(W), "");
    hdr += card("NAXIS2", std::to_string(H), "");
    for (const auto& c : cards) {
        std::string v = c.value;
        if (c.is_string) {
            std::string q = "'" + v;
            while (q.size() < 9) q += ' ';
            v = q + "'";
        }
        hdr += card(c.key, v, c.comment);
    }
    std::string end = "END";
    end.resize(80, ' ');
    hdr += end;
    hdr.resize((hdr.size() + 2879) / 2880 * 2880, ' ');
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    const size_t n = (size_t)W * H;
    std::vector<unsigned char> body((n * 4 + 2879) / 2880 * 2880, 0);
    for (size_t i = 0; i < n; i++) {
        uint32_t u;
        std::memcpy(&u, &data[i], 4);
        body[4 * i] = (unsigned char)(u >> 24);
        body[4 * i + 1] = (unsigned char)(u >> 16);
        body[4 * i + 2] = (unsigned char)(u >> 8);
        body[4 * i + 3] = (unsigned char)u;
    }
    ok = ok && std::fwrite(body.data(), 1, body.size(), f) == body.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok;
}

}  // namespace birp_api
