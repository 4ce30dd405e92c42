// DataCube.cpp -- the reference's DataCube API (reference src/DataCube.cpp) over the GPU.
#include "DataCube.h"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <exception>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>

#include "birp.h"
#include "device.h"

std::vector<int> Helper_dims(const std::string& line);

DataCube::DataCube() : size{0, 0, 0}, spacing{0, 0, 0}, origin_coords{0, 0, 0} {}

DataCube::~DataCube() { DropDevice(); }

DataCube::DataCube(const DataCube& o)
    : Space(), coords_x(o.coords_x), coords_y(o.coords_y), coords_z(o.coords_z), size(o.size), slices(o.slices),
      trilinear(o.trilinear) {
    for (int k = 0; k < 3; k++) {
        spacing[k] = o.spacing[k];
        origin_coords[k] = o.origin_coords[k];
    }
}

DataCube& DataCube::operator=(const DataCube& o) {
    if (this != &o) {
        DropDevice();
        coords_x = o.coords_x;
        coords_y = o.coords_y;
        coords_z = o.coords_z;
        size = o.size;
        slices = o.slices;
        trilinear = o.trilinear;
        for (int k = 0; k < 3; k++) {
            spacing[k] = o.spacing[k];
            origin_coords[k] = o.origin_coords[k];
        }
        version_++;
    }
    return *this;
}

void DataCube::DropDevice() {
    for (auto& d : device_) {
        if (d) birp_volume_free(d);
        d = nullptr;
    }
}

void DataCube::Init() {}

void DataCube::Init(std::string& filename, bool debug) { Load(filename, debug); }

namespace {
[[noreturn]] void malformed() { throw std::invalid_argument("Datacube is invalid, inaccessible, or malformed."); }

bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

// The blank-separated numbers of text [a, b) appended to out. std::from_chars reads a decimal or
// "inf"/"nan" token correctly rounded to float, as strtof does; a token it does not take (a leading
// '+', hexadecimal, a value that under- or overflows float) goes to strtof itself. A token that
// is not a number throws.
void parse_floats(const char* a, const char* b, std::vector<float>& out) {
    const char* p = a;
    while (true) {
        while (p < b && is_blank(*p)) ++p;
        if (p >= b) return;
        float f = 0.f;
        std::from_chars_result r = std::from_chars(p, b, f);
        if (r.ec != std::errc() || (r.ptr < b && !is_blank(*r.ptr))) {
            const char* e = p;
            while (e < b && !is_blank(*e)) ++e;
            const std::string tok(p, e);
            char* end = nullptr;
            f = std::strtof(tok.c_str(), &end);
            if (end != tok.c_str() + tok.size()) malformed();
            r.ptr = e;
        }
        out.push_back(f);
        p = r.ptr;
    }
}
}  // namespace

// The .dat export (DataCube.cpp:49-125): line 1 is metadata, line 10 holds "nx ny nz"; every
// number after line 10 is, in order, the nx x-coordinates, the ny y-coordinates, the nz
// z-coordinates, then the nx*ny*nz values with x fastest, then y, then z. The reference walks the
// same numbers line by line and so also needs each axis' coordinates and each x-row of values to
// end a line; files in that layout read the same here. The file is read whole and its numbers are
// parsed on up to 16 threads, each taking a run of whole lines (a 512^3 cube is ~1.6 GB of text).
void DataCube::Load(std::string filename, bool debug) {
    std::ifstream in(filename, std::ios::binary);
    if (!in) malformed();
    in.seekg(0, std::ios::end);
    const std::streamoff len = in.tellg();
    if (len <= 0) malformed();
    std::string text((size_t)len, '\0');
    in.seekg(0);
    if (!in.read(&text[0], len)) malformed();
    const char* const beg = text.data();
    const char* const end = beg + text.size();
    // lines 1-9 are skipped; line 10 holds the dimensions
    const char* p = beg;
    for (int line = 1; line < 10; line++) {
        p = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        if (!p) malformed();
        ++p;
    }
    const char* eol = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
    if (!eol) eol = end;
    const std::vector<int> dims = Helper_dims(std::string(p, eol));
    if (dims.size() < 3 || dims[0] < 1 || dims[1] < 1 || dims[2] < 1) malformed();
    const size_t nx = dims[0], ny = dims[1], nz = dims[2];
    const size_t want = nx + ny + nz + nx * ny * nz;
    // split the rest into runs of whole lines, one per thread
    const char* body = eol < end ? eol + 1 : end;
    const size_t bytes = (size_t)(end - body);
    unsigned hw = std::thread::hardware_concurrency();
    size_t nt = std::min<size_t>(std::max(1u, std::min(hw, 16u)), std::max<size_t>(1, bytes >> 20));
    std::vector<const char*> cut(1, body);
    for (size_t t = 1; t < nt; t++) {
        const char* q = body + bytes * t / nt;
        if (q < cut.back()) q = cut.back();
        const char* nl = static_cast<const char*>(std::memchr(q, '\n', (size_t)(end - q)));
        cut.push_back(nl ? nl + 1 : end);
    }
    cut.push_back(end);
    nt = cut.size() - 1;
    std::vector<std::vector<float>> part(nt);
    std::vector<std::exception_ptr> err(nt);
    {
        std::vector<std::thread> pool;
        for (size_t t = 0; t < nt; t++)
            pool.emplace_back([&, t] {
                try {
                    part[t].reserve(want / nt + 1024);
                    parse_floats(cut[t], cut[t + 1], part[t]);
                } catch (...) {
                    err[t] = std::current_exception();
                }
            });
        for (auto& th : pool) th.join();
    }
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
    size_t got = 0;
    for (auto& v : part) got += v.size();
    if (got != want) malformed();
    // numbers in file order: the three axes, then the values
    std::vector<float> nums;
    nums.reserve(want);
    for (auto& v : part) {
        nums.insert(nums.end(), v.begin(), v.end());
        std::vector<float>().swap(v);
    }
    coords_x.assign(nums.begin(), nums.begin() + nx);
    coords_y.assign(nums.begin() + nx, nums.begin() + nx + ny);
    coords_z.assign(nums.begin() + nx + ny, nums.begin() + nx + ny + nz);
    InitSpacing();
    InitOriginCoords();
    InitSize();
    const float* v = nums.data() + nx + ny + nz;
    slices.assign(nz, std::vector<std::vector<float>>(ny));
    for (size_t z = 0; z < nz; z++)
        for (size_t y = 0; y < ny; y++) slices[z][y].assign(v + (z * ny + y) * nx, v + (z * ny + y + 1) * nx);
    if (debug) {
        // DataCube.cpp:127-133 stamps a marker line into slice 80, row 79
        if (slices.size() > 80 && slices[80].size() > 79) {
            auto& row = slices[80][79];
            for (size_t i = 0; i < 101 && i < row.size(); i++) row[i] = 1;
        }
    }
    version_++;
}

void DataCube::InitSpacing() {
    spacing[0] = GetAxisSpacing(coords_x);
    spacing[1] = GetAxisSpacing(coords_y);
    spacing[2] = GetAxisSpacing(coords_z);
}

// DataCube.cpp:150-183: the first gap |c[0] - c[1]| (the uniformity check never rejects).
float DataCube::GetAxisSpacing(const std::vector<float>& c) {
    if (c.size() < 2) throw std::invalid_argument("Datacube is invalid, inaccessible, or malformed.");
    return std::abs(c[0] - c[1]);
}

void DataCube::InitOriginCoords() {
    origin_coords[0] = coords_x[0];
    origin_coords[1] = coords_y[0];
    origin_coords[2] = coords_z[0];
}

void DataCube::InitSize() {
    size.x = (int)coords_x.size();
    size.y = (int)coords_y.size();
    size.z = (int)coords_z.size();
}

void DataCube::SetTrilinear(bool option) {
    if (option != trilinear) version_++;
    trilinear = option;
}

void DataCube::Touch() { version_++; }

void DataCube::SetValues(int nx, int ny, int nz, const float* origin, const float* spc, const float* values) {
    if (nx < 2 || ny < 2 || nz < 2) throw std::invalid_argument("DataCube::SetValues: every axis needs >= 2 points");
    coords_x.resize(nx);
    coords_y.resize(ny);
    coords_z.resize(nz);
    for (int i = 0; i < nx; i++) coords_x[i] = origin[0] + spc[0] * i;
    for (int i = 0; i < ny; i++) coords_y[i] = origin[1] + spc[1] * i;
    for (int i = 0; i < nz; i++) coords_z[i] = origin[2] + spc[2] * i;
    InitSize();
    for (int k = 0; k < 3; k++) {
        spacing[k] = spc[k];
        origin_coords[k] = origin[k];
    }
    slices.assign(nz, std::vector<std::vector<float>>(ny));
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            slices[z][y].assign(values + ((size_t)z * ny + y) * nx, values + ((size_t)z * ny + y + 1) * nx);
    version_++;
}

void DataCube::Save(const std::string& filename) const {
    FILE* f = std::fopen(filename.c_str(), "w");
    if (!f) throw std::invalid_argument("DataCube::Save: cannot write " + filename);
    std::fprintf(f, "birp DataCube export\n");
    for (int k = 0; k < 8; k++) std::fprintf(f, "#\n");
    std::fprintf(f, "%zu %zu %zu\n", coords_x.size(), coords_y.size(), coords_z.size());
    for (const auto* c : {&coords_x, &coords_y, &coords_z}) {
        for (size_t i = 0; i < c->size(); i++) std::fprintf(f, i ? " %.9g" : "%.9g", (*c)[i]);
        std::fprintf(f, "\n");
    }
    for (const auto& slice : slices) {
        for (const auto& row : slice) {
            for (size_t i = 0; i < row.size(); i++) std::fprintf(f, i ? " %.9g" : "%.9g", row[i]);
            std::fprintf(f, "\n");
        }
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

birp_volume* DataCube::DeviceVolume(int dtype) {
    if (dtype < 0 || dtype > 2) throw std::invalid_argument("DataCube::DeviceVolume: unknown dtype");
    if (device_[dtype] && device_version_[dtype] == version_) return device_[dtype];
    if (device_[dtype]) {
        birp_volume_free(device_[dtype]);
        device_[dtype] = nullptr;
    }
    const int nx = size.x, ny = size.y, nz = size.z;
    if ((int)slices.size() < nz) throw std::invalid_argument("DataCube: slices hold fewer than size.z slices");
    std::vector<float> flat((size_t)nx * ny * nz);
    for (int z = 0; z < nz; z++) {
        if ((int)slices[z].size() < ny) throw std::invalid_argument("DataCube: a slice holds fewer than size.y rows");
        for (int y = 0; y < ny; y++) {
            if ((int)slices[z][y].size() < nx) throw std::invalid_argument("DataCube: a row holds fewer than size.x values");
            std::memcpy(&flat[((size_t)z * ny + y) * nx], slices[z][y].data(), sizeof(float) * nx);
        }
    }
    birp_volume_desc d{};
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    for (int k = 0; k < 3; k++) {
        d.origin[k] = origin_coords[k];
        d.spacing[k] = spacing[k];
    }
    d.interp = trilinear ? BIRP_INTERP_TRILINEAR : BIRP_INTERP_NEAREST;
    d.dtype = dtype;
    birp_api::check(birp_volume_upload(birp_api::device(), &d, flat.data(), &device_[dtype]), "DataCube upload");
    device_version_[dtype] = version_;
    return device_[dtype];
}

float DataCube::GetSample(float x, float y, float z) {
    float p[3] = {x, y, z}, out = 0.f;
    birp_api::check(birp_volume_sample(DeviceVolume(BIRP_DTYPE_F32), p, 1, &out), "DataCube::GetSample");
    return out;
}

void DataCube::GetSamples(const float* xyz, std::size_t n, float* out) {
    birp_api::check(birp_volume_sample(DeviceVolume(BIRP_DTYPE_F32), xyz, (long long)n, out), "DataCube::GetSamples");
}

// "nx ny nz" of line 10, split on blanks (Helper::explode_int(line, ' ') in the reference).
std::vector<int> Helper_dims(const std::string& line) {
    std::vector<int> out;
    std::istringstream in(line);
    for (std::string tok; in >> tok;) out.push_back(std::stoi(tok));
    return out;
}
