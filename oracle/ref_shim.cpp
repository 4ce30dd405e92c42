// ref_shim.cpp -- a C-ABI driver over the reference's own DataCube, CMEM and Helper classes,
// compiled together with their sources (src/DataCube.cpp, src/Cmem.cpp, src/Helper.cpp) into
// oracle/_ref/libbirp_ref.so by oracle/Makefile. TEST INFRASTRUCTURE ONLY: it is used to pin
// the oracle (birp_oracle.c) against the reference, to write golden fixtures, and as the timed
// reference CPU baseline of bench.py (ref_render_rows); the product never loads it.
// src/Camera.cpp is not compiled here: it includes EleFits headers (Camera.cpp:23-27), which
// this image does not have.
#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "Cmem.h"
#include "DataCube.h"
#include "Helper.h"

extern "C" {

// DataCube::Load on a .dat file (DataCube.cpp:28-138), then SetTrilinear.
void* ref_cube_load(const char* path, int trilinear) {
    try {
        DataCube* c = new DataCube();
        std::string p(path);
        c->Init(p, false);
        c->SetTrilinear(trilinear != 0);
        return c;
    } catch (...) {
        return nullptr;
    }
}

void ref_cube_info(void* h, int* size3, float* spacing3, float* origin3) {
    DataCube* c = static_cast<DataCube*>(h);
    size3[0] = c->size.x;
    size3[1] = c->size.y;
    size3[2] = c->size.z;
    for (int i = 0; i < 3; i++) {
        spacing3[i] = c->spacing[i];
        origin3[i] = c->origin_coords[i];
    }
}

// Copies the loaded values out in [z][y][x] order (for checking the loader).
int ref_cube_values(void* h, float* out, long long n) {
    DataCube* c = static_cast<DataCube*>(h);
    long long idx = 0;
    for (int z = 0; z < c->size.z; z++)
        for (int y = 0; y < c->size.y; y++)
            for (int x = 0; x < c->size.x; x++) {
                if (idx >= n) return -1;
                out[idx++] = c->slices.at(z).at(y).at(x);
            }
    return 0;
}

// DataCube::GetSample (DataCube.cpp:203-277). *err = 1 when the reference throws.
float ref_cube_sample(void* h, float x, float y, float z, int* err) {
    *err = 0;
    try {
        return static_cast<DataCube*>(h)->GetSample(x, y, z);
    } catch (...) {
        *err = 1;
        return std::numeric_limits<float>::quiet_NaN();
    }
}

void ref_cube_free(void* h) { delete static_cast<DataCube*>(h); }

// A DataCube holding an in-memory uniform grid, set up through its public members
// (DataCube.h:28-42) to the state DataCube::Load leaves: coordinates origin + spacing * i, size,
// spacing, origin and the [z][y][x] slices.
void* ref_cube_from_grid(int nx, int ny, int nz, const float* origin3, const float* spacing3, const float* values,
                         int trilinear) {
    DataCube* c = new DataCube();
    const int n[3] = {nx, ny, nz};
    std::vector<float>* coords[3] = {&c->coords_x, &c->coords_y, &c->coords_z};
    for (int a = 0; a < 3; a++) {
        coords[a]->resize(n[a]);
        for (int i = 0; i < n[a]; i++) (*coords[a])[i] = origin3[a] + spacing3[a] * (float)i;
        c->spacing[a] = spacing3[a];
        c->origin_coords[a] = origin3[a];
    }
    c->size.x = nx;
    c->size.y = ny;
    c->size.z = nz;
    c->slices.assign(nz, std::vector<std::vector<float>>(ny, std::vector<float>(nx)));
    size_t k = 0;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) c->slices[z][y][x] = values[k++];
    c->SetTrilinear(trilinear != 0);
    return c;
}

// Image rows [row0, row1) rendered with the reference's own sampler: the march and integration
// of Camera::Render (Camera.cpp:170-229) restated around Space::GetSample, called virtually as
// Camera.cpp:191 calls it, on the reference's DataCube (its GetSample, DataCube.cpp:203-277, with
// its per-call allocations). Rays come in as unit directions (W*H*3 floats, from the oracle's
// ray generation); t_k and w follow GenerateRayDistWidth (Camera.cpp:69-81). semantics 0 stops a
// ray at its first negative scaled sample (the out-of-cube -1 included); 1 skips out-of-cube
// samples. Returns the number of samples added.
long long ref_render_rows(void* h, int W, int H, const float* dirs, const float* pos3, float L, int n, int row0,
                          int row1, int semantics, float* img) {
    Space* space = static_cast<DataCube*>(h);
    std::vector<float> t(n);
    for (int k = 0; k < n; k++) t[k] = L * ((k + 0.5f) / n);
    const float w = n > 1 ? t[1] - t[0] : 0.f;
    long long count = 0;
    for (int j = row0; j < row1 && j < H; ++j)
        for (int i = 0; i < W; ++i) {
            const float* d = dirs + ((size_t)j * W + i) * 3;
            float acc = 0.f;
            for (int k = 0; k < n; ++k) {
                float x = d[0] * t[k] + pos3[0], y = d[1] * t[k] + pos3[1], z = d[2] * t[k] + pos3[2];
                float v = space->GetSample(x, y, z);
                if (semantics == 1 && v == -1.f) continue;  // out of the cube: skipped
                float sample = v;
                sample = (float)(((double)(sample * w) * (1000 * 100 * 6378.1) / (4 / M_PI)) / 10000);
                if (sample < 0) break;
                acc = acc + sample;
                count++;
            }
            img[(size_t)j * W + i] = acc;
        }
    return count;
}

// CMEM::Init() (Cmem.cpp:18-34).
void* ref_cmem_default() {
    CMEM* m = new CMEM();
    m->Init();
    return m;
}

// CMEM::Init(hack, v, b, dipole, p0, p1, p2, p3, B, alpha, beta, bs, A1, A2, ay_bs, az_bs,
// density) (Cmem.cpp:35-115). nv/nb == 0 passes an empty vector.
void* ref_cmem_init(const int* v, int nv, const int* b, int nb, float dipole, float p0, int p1,
                    int p2, int p3, float B, float alpha, float beta, float bs, float A1, float A2,
                    float ay_bs, float az_bs, float density) {
    std::vector<int> vv(v, v + nv), bb(b, b + nb);
    CMEM* m = new CMEM();
    m->Init(true, vv, bb, dipole, p0, p1, p2, p3, B, alpha, beta, bs, A1, A2, ay_bs, az_bs, density);
    return m;
}

float ref_cmem_sample(void* h, float x, float y, float z) {
    return static_cast<CMEM*>(h)->GetSample(x, y, z);
}

void ref_cmem_free(void* h) { delete static_cast<CMEM*>(h); }

// Helper::MatrixMultiply (Helper.cpp:70-98), 3x3 row-major.
void ref_matmul3(const float* a, const float* b, float* out) {
    std::vector<std::vector<float>> A(3, std::vector<float>(3)), B(3, std::vector<float>(3));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            A[i][j] = a[i * 3 + j];
            B[i][j] = b[i * 3 + j];
        }
    auto R = Helper::MatrixMultiply(A, B);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i * 3 + j] = R[i][j];
}

// Helper::GetInverse (Helper.cpp:241-264). Returns -1 when the reference throws.
int ref_get_inverse3(const float* m, float* out) {
    std::vector<std::vector<float>> M(3, std::vector<float>(3));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[i][j] = m[i * 3 + j];
    try {
        auto R = Helper::GetInverse(M);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) out[i * 3 + j] = R[i][j];
        return 0;
    } catch (...) {
        return -1;
    }
}

// Helper::Gei2GseTransforms (Helper.cpp:306-341): the 4x4 rotation, row-major.
void ref_gei2gse(float* out16) {
    auto M = Helper::Gei2GseTransforms();
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out16[i * 4 + j] = M[i][j];
}

// Helper::RaDecToGSI (Helper.cpp:343-380).
void ref_radec_to_gsi(float ra, float dec, float* out3) {
    auto v = Helper::RaDecToGSI(ra, dec);
    for (int i = 0; i < 3; i++) out3[i] = v[i];
}

// Helper::RaDecGSEConversion (Helper.cpp:382-390) with a 4x4 row-major matrix.
void ref_radec_gse(float ra, float dec, const float* m16, float* out3) {
    std::vector<std::vector<float>> M(4, std::vector<float>(4));
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) M[i][j] = m16[i * 4 + j];
    auto v = Helper::RaDecGSEConversion({ra, dec}, M);
    for (int i = 0; i < 3; i++) out3[i] = v[i];
}

// Helper::TwoMin (Helper.cpp:392-413).
void ref_two_min(const float* arr, int n, const int* mask, int nm, float* out2) {
    auto v = Helper::TwoMin(std::vector<float>(arr, arr + n), std::vector<int>(mask, mask + nm));
    out2[0] = v[0];
    out2[1] = v[1];
}

// Helper::XYToRaDec (Helper.cpp:415-476) on a sky grid of n points.
void ref_xy_to_radec(float x, float y, const float* xsky, const float* ysky, const float* sky1, const float* sky2,
                     int n, float skysep, float* out2) {
    auto v = Helper::XYToRaDec(x, y, std::vector<float>(xsky, xsky + n), std::vector<float>(ysky, ysky + n),
                               std::vector<float>(sky1, sky1 + n), std::vector<float>(sky2, sky2 + n), skysep);
    out2[0] = v[0];
    out2[1] = v[1];
}

// Helper::Interp1 (Helper.cpp:478-513). Every x_new must lie strictly above some x (the
// reference indexes slope[-1] otherwise).
void ref_interp1(const float* x, const float* y, int n, const float* xn, int m, float* out) {
    std::vector<float> X(x, x + n), Y(y, y + n);
    auto v = Helper::Interp1(X, Y, std::vector<float>(xn, xn + m));
    for (int i = 0; i < m; i++) out[i] = v[i];
}

// Helper::findNearestNeighbourIndex (Helper.cpp:515-527).
int ref_nearest_below(float value, const float* x, int n) {
    std::vector<float> X(x, x + n);
    return Helper::findNearestNeighbourIndex(value, X);
}

// Helper::XYZToSpherical / XYZToSphericalAlt (Helper.cpp:529-585).
void ref_xyz_to_spherical(const float* xyz, int alt, float* out3) {
    float p[3] = {xyz[0], xyz[1], xyz[2]};
    auto v = alt ? Helper::XYZToSphericalAlt(p) : Helper::XYZToSpherical(p);
    for (int i = 0; i < 3; i++) out3[i] = v[i];
}

}  // extern "C"
