// Helper.cpp -- the reference's Helper routines (reference src/Helper.cpp) that the render path
// and the loader use, with the same float operation order (GetInverse is bit-identical).
#include "Helper.h"

#include <cmath>
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
