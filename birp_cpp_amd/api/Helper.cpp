// Helper.cpp -- the reference's Helper routines (reference src/Helper.cpp), with the same float
// operation order: the matrix routines the render path uses (GetInverse is bit-identical), the
// string helpers the loader and the CLI use, and the pointing utilities.
#include "Helper.h"

#include <cmath>
#include <limits>
#include <sstream>
#include <stdexcept>

std::vector<float> Helper::explode_float(std::string const& s, char delim) {
    std::vector<float> out;
    std::istringstream in(s);
    for (std::string tok; std::getline(in, tok, delim);)
        if (!tok.empty()) out.push_back(std::stof(tok));
    return out;
}

std::vector<int> Helper::explode_int(std::string const& s, char delim) {
    std::vector<int> out;
    std::istringstream in(s);
    for (std::string tok; std::getline(in, tok, delim);)
        if (!tok.empty()) out.push_back(std::stoi(tok));
    return out;
}

std::vector<std::string> Helper::explode_string(std::string const& s, char delim) {
    std::vector<std::string> out;
    std::istringstream in(s);
    for (std::string tok; std::getline(in, tok, delim);)
        if (!tok.empty()) out.push_back(tok);
    return out;
}

// Helper.cpp:70-98: rows of a times columns of b, summing over a's row count, from 0.0f.
std::vector<std::vector<float>> Helper::MatrixMultiply(std::vector<std::vector<float>> a, std::vector<std::vector<float>> b) {
    const size_t n = a.size(), m = b.size();
    std::vector<std::vector<float>> r(n, std::vector<float>(m));
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < m; j++) {
            float acc = 0.0f;
            for (size_t k = 0; k < n; k++) acc = acc + (a[i][k] * b[k][j]);
            r[i][j] = acc;
        }
    return r;
}

// Helper.cpp:100-114: matrix times vector into a fixed 4-element result.
std::vector<float> Helper::MatrixMultiply(std::vector<std::vector<float>> a, std::vector<float> b) {
    float r[4] = {0, 0, 0, 0};
    for (size_t row = 0; row < a.size() && row < 4; row++)
        for (size_t col = 0; col < a[0].size(); col++) r[row] += a[row][col] * b[col];
    return std::vector<float>(r, r + 4);
}

std::vector<std::vector<float>> Helper::MatrixScalarMultiply(std::vector<std::vector<float>> m, float s) {
    for (auto& row : m)
        for (auto& x : row) x = x * s;
    return m;
}

std::vector<float> Helper::MatrixScalarMultiply(std::vector<float> m, float s) {
    for (auto& x : m) x = x * s;
    return m;
}

// Helper.cpp:136-148: each output component is row . coords, summed left to right.
std::vector<float> Helper::ApplyRotation(std::vector<std::vector<float>> rot, float* c) {
    std::vector<float> out;
    out.reserve(rot.size());
    for (const auto& row : rot) out.push_back((row[0] * c[0]) + (row[1] * c[1]) + (row[2] * c[2]));
    return out;
}

// Helper.cpp:150-191: Laplace expansion along the first row, float arithmetic.
float Helper::GetDeterminant(const std::vector<std::vector<float>> m) {
    if (m.empty()) return 1;
    if (m.size() != m[0].size()) throw std::runtime_error("Matrix is not quadratic");
    const size_t n = m.size();
    if (n == 1) return m[0][0];
    if (n == 2) return m[0][0] * m[1][1] - m[0][1] * m[1][0];
    float det = 0;
    int sign = 1;
    for (size_t col = 0; col < n; col++) {
        std::vector<std::vector<float>> minor(n - 1, std::vector<float>(n - 1));
        for (size_t r = 1; r < n; r++) {
            size_t c2 = 0;
            for (size_t c = 0; c < n; c++)
                if (c != col) minor[r - 1][c2++] = m[r][c];
        }
        det = det + sign * m[0][col] * GetDeterminant(minor);
        sign = -sign;
    }
    return det;
}

std::vector<std::vector<float>> Helper::GetTranspose(const std::vector<std::vector<float>> m) {
    std::vector<std::vector<float>> t(m[0].size(), std::vector<float>(m.size()));
    for (size_t i = 0; i < m.size(); i++)
        for (size_t j = 0; j < m[0].size(); j++) t[j][i] = m[i][j];
    return t;
}

// Helper.cpp:207-239: cofactor = (-1)^(i+j) * minor determinant (the sign is applied in double
// and rounded back, which is exact).
std::vector<std::vector<float>> Helper::GetCofactor(const std::vector<std::vector<float>> m) {
    if (m.size() != m[0].size()) throw std::runtime_error("Matrix is not quadratic");
    const size_t n = m.size();
    std::vector<std::vector<float>> out(n, std::vector<float>(n));
    std::vector<std::vector<float>> minor(n - 1, std::vector<float>(n - 1));
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < n; j++) {
            size_t p = 0;
            for (size_t x = 0; x < n; x++) {
                if (x == i) continue;
                size_t q = 0;
                for (size_t y = 0; y < n; y++) {
                    if (y == j) continue;
                    minor[p][q++] = m[x][y];
                }
                p++;
            }
            const double sgn = ((i + j) % 2) ? -1.0 : 1.0;
            out[i][j] = (float)(sgn * (double)GetDeterminant(minor));
        }
    return out;
}

// Helper.cpp:241-264: transpose(cofactor) * (1/det); throws when det == 0.
std::vector<std::vector<float>> Helper::GetInverse(const std::vector<std::vector<float>> m) {
    if (GetDeterminant(m) == 0) throw std::runtime_error("Determinant is 0");
    const float d = 1.f / GetDeterminant(m);
    std::vector<std::vector<float>> out = GetTranspose(GetCofactor(m));
    for (auto& row : out)
        for (auto& x : row) x *= d;
    return out;
}

float Helper::VectorDistance(float* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
float Helper::VectorDistance2D(float* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1]); }

// Helper.cpp:603-622 compares iterator positions (i - next(i) is always -1), so it never finds
// a difference: every list of steps is accepted, as here.
bool Helper::EqualDistance(std::vector<float> /*steps*/) { return true; }

// ------------------------------------------------------------------------------------------
// Pointing utilities (Helper.cpp:266-600): GEI -> GSE rotation for 2024-01-01, RA/Dec -> unit
// vectors, image-plane -> RA/Dec interpolation on a sky grid, 1-D interpolation and spherical
// coordinates. Same float/double types and operation order as the reference, so the results
// are bit-identical on x86-64 (checked against the reference's compiled Helper).
// ------------------------------------------------------------------------------------------
namespace {
constexpr double kPiD = 3.14159265358979323846;

// Julian date of a calendar date (the free function at Helper.cpp:266-304).
float julian_date(int year, int month, float day) {
    const bool early = month == 1 || month == 2;
    const int yp = early ? year - 1 : year;
    const int mp = early ? month + 12 : month;
    int b = 0;
    const bool julian = year < 1582 || (year == 1582 && month < 10) || (year == 1582 && month == 10 && day < 15);
    if (!julian) {
        const int a = (int)std::trunc(yp / 100.f);
        b = 2 - a + std::trunc(a / 4.f);
    }
    const int c = yp < 0 ? (int)std::trunc((356.25 * yp) - 0.75) : (int)std::trunc(365.25 * yp);
    const int d = (int)std::trunc(30.6001 * (mp + 1));
    return b + c + d + day + 1720994.5f;
}
}  // namespace

std::vector<std::vector<float>> Helper::Gei2GseTransforms() {
    const float conv = kPiD / 180;
    float date = julian_date(2024, 1, 1);
    date = date - 2400000.5f;                 // modified Julian date
    const float t0 = (date - 51544.5f) / 36525;  // Julian centuries from J2000
    const float epsilon = 23.439f - 0.013f * t0;
    const float m = 357.528 + 35999.050f * t0;
    const float lambda = 280.460 + 36000.772 * t0;
    const float lambda_sun = lambda + (1.915 - 0.0048 * t0) * std::sin(m * conv) + 0.020f * std::sin((2 * m) * conv);
    const float cl = std::cos(lambda_sun * conv), sl = std::sin(lambda_sun * conv);
    const float ce = std::cos(epsilon * conv), se = std::sin(epsilon * conv);
    std::vector<std::vector<float>> a = {{cl, sl, 0, 0}, {-sl, cl, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    std::vector<std::vector<float>> b = {{1, 0, 0, 0}, {0, ce, se, 0}, {0, -se, ce, 0}, {0, 0, 0, 1}};
    return MatrixMultiply(a, b);
}

std::vector<float> Helper::RaDecToGSI(float ra, float dec) {
    const float conv = kPiD / 180;
    const float alpha = ra * conv, delta = dec * conv;
    const float z = std::sin(delta);
    float y2 = 1 - z * z;
    if (y2 < 0) y2 = 0.f;
    const float ta = std::tan(alpha);
    float y = std::sqrt(y2) * std::tan(alpha) / std::sqrt(1 + ta * ta);
    float x2 = 1 - y * y - z * z;
    if (x2 < 0) x2 = 0.f;
    float x = std::sqrt(x2);
    if (90 < ra && ra < 180) {
        x = std::abs(x) * (-1);
        y = std::abs(y);
    } else if (180 < ra && ra < 270) {
        x = std::abs(x) * (-1);
        y = std::abs(y) * (-1);
    }
    return {x, y, z};
}

std::vector<float> Helper::RaDecGSEConversion(std::vector<float> ra_dec, std::vector<std::vector<float>> gei_to_gse) {
    std::vector<float> gei = RaDecToGSI(ra_dec[0], ra_dec[1]);
    gei.push_back(1.f);
    std::vector<float> gse = MatrixMultiply(gei_to_gse, gei);
    gse.pop_back();
    return gse;
}

// The indices (as floats) of the two smallest of arr[0], arr[1] and the masked elements among
// arr[2 .. size-2] (the last element is never considered, as in Helper.cpp:392-410).
std::vector<float> Helper::TwoMin(std::vector<float> arr, std::vector<int> mask) {
    float lo = 0, hi = 1;
    if (arr[(size_t)hi] < arr[(size_t)lo]) std::swap(lo, hi);
    for (int i = 2; i < (int)arr.size() - 1; i++) {
        if (std::find(mask.begin(), mask.end(), i) == mask.end()) continue;
        if (arr[i] < arr[(size_t)lo]) {
            hi = lo;
            lo = i;
        } else if (arr[i] < arr[(size_t)hi]) {
            hi = i;
        }
    }
    return {lo, hi};
}

std::vector<float> Helper::XYToRaDec(float ct_x, float ct_y, std::vector<float> xsky, std::vector<float> ysky,
                                     std::vector<float> sky1, std::vector<float> sky2, float skysep) {
    const float conv = 180 / kPiD;
    std::vector<float> dist;
    for (size_t i = 0; i < xsky.size(); i++)
        dist.push_back(std::sqrt(((ct_x - xsky[i]) * (ct_x - xsky[i])) + ((ct_y - ysky[i]) * (ct_y - ysky[i]))));
    const int near = (int)(std::min_element(dist.begin(), dist.end()) - dist.begin());
    // neighbours along declination: the grid points on the nearest point's RA line
    std::vector<int> sel;
    for (size_t i = 0; i < sky1.size(); i++)
        if (sky1[i] == sky1[near]) sel.push_back((int)i);
    std::vector<float> two = TwoMin(dist, sel);
    const size_t p0 = (size_t)two[0], p1 = (size_t)two[1];
    const float diff_dec = sky2[p1] - sky2[p0];
    const float diff_x = xsky[p1] - xsky[p0];
    const float diff_y = ysky[p1] - ysky[p0];
    // neighbours along right ascension: the points on the nearest point's Dec line
    sel.clear();
    for (size_t i = 0; i < sky2.size(); i++)
        if (sky2[i] == sky2[near]) sel.push_back((int)i);
    two = TwoMin(dist, sel);
    const size_t q0 = (size_t)two[0], q1 = (size_t)two[1];
    float diff_ra = sky1[q1] - sky1[q0];
    const float diff_x2 = xsky[q1] - xsky[q0];
    const float diff_y2 = ysky[q1] - ysky[q0];

    const float d_ang = conv * std::atan2(diff_y, diff_x);
    const float c_ang = conv * std::atan2((ct_y - ysky[near]), (ct_x - xsky[near]));
    float ct_dec = sky2[near] + ((diff_dec / skysep) * (dist[near] * std::cos((d_ang - c_ang) / conv)));
    if (ct_dec > 90.f) ct_dec = 90.f;
    else if (ct_dec < -90.f) ct_dec = -90.f;
    if (diff_ra == (skysep - 360.f)) diff_ra = skysep;  // across the 0/360 seam
    else if (diff_ra == (360.f - skysep)) diff_ra = -1.f * skysep;
    const float d2_ang = conv * std::atan2(diff_y2, diff_x2);
    float ct_ra;
    if (std::cos(ct_dec / conv) != 0)
        ct_ra = sky1[near] + ((diff_ra / (1.f * skysep)) * (dist[near] * std::cos((d2_ang - c_ang) / conv)) / std::cos(ct_dec / conv));
    else
        ct_ra = sky1[near];
    if (ct_ra < 0) ct_ra = ct_ra + 360.f;
    else if (ct_ra > 360) ct_ra = ct_ra - 360.f;
    return {ct_ra, ct_dec};
}

// Piecewise-linear interpolation (Helper.cpp:478-513): each new x uses the segment of the grid
// point nearest below it (strictly); the last point reuses the last segment. The reference reads
// slope[-1] when no grid point lies below a new x; this throws std::out_of_range instead.
std::vector<float> Helper::Interp1(std::vector<float>& x, std::vector<float>& y, std::vector<float> x_new) {
    const size_t n = x.size();
    std::vector<float> slope(n), intercept(n);
    for (size_t i = 0; i + 1 < n; i++) {
        const float dx = x[i + 1] - x[i], dy = y[i + 1] - y[i];
        slope[i] = dy / dx;
        intercept[i] = y[i] - x[i] * slope[i];
    }
    if (n >= 2) {
        slope[n - 1] = slope[n - 2];
        intercept[n - 1] = intercept[n - 2];
    }
    std::vector<float> out;
    out.reserve(x_new.size());
    for (float v : x_new) {
        const int k = findNearestNeighbourIndex(v, x);
        if (k < 0) throw std::out_of_range("Helper::Interp1: no grid point below x");
        out.push_back(slope[k] * v + intercept[k]);
    }
    return out;
}

// Index of the grid point nearest below value (strictly), -1 if none (Helper.cpp:515-527).
int Helper::findNearestNeighbourIndex(float value, std::vector<float>& x) {
    float best = std::numeric_limits<float>::max();
    int k = -1;
    for (int i = 0; i < (int)x.size(); ++i) {
        const float d = value - x[i];
        if (d > 0 && d < best) {
            best = d;
            k = i;
        }
    }
    return k;
}

// {phi = atan2(y, x), theta = atan2(|(x, y)|, z), r} (Helper.cpp:529-548).
std::vector<float> Helper::XYZToSpherical(float* xyz) {
    const float x = xyz[0], y = xyz[1], z = xyz[2];
    const float r = std::sqrt(x * x + y * y + z * z);
    const float theta = std::atan2(std::sqrt(x * x + y * y), z);
    const float phi = std::atan2(y, x);
    return {phi, theta, r};
}

// {phi = atan2(z, |(x, y)|), theta = atan2(y, x), r} (Helper.cpp:550-585).
std::vector<float> Helper::XYZToSphericalAlt(float* xyz) {
    const float x = xyz[0], y = xyz[1], z = xyz[2];
    const float r = std::sqrt(x * x + y * y + z * z);
    const float theta = std::atan2(y, x);
    const float phi = std::atan2(z, std::sqrt(x * x + y * y));
    return {phi, theta, r};
}
