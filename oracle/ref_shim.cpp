// ref_shim.cpp -- a C-ABI driver over the reference's own DataCube, CMEM and Helper classes,
// compiled together with their sources (src/DataCube.cpp, src/Cmem.cpp, src/Helper.cpp) into
// oracle/_ref/libbirp_ref.so by oracle/Makefile. TEST INFRASTRUCTURE ONLY: it is used to pin
// the oracle (birp_oracle.c) against the reference and to write golden fixtures; the product
// never loads it. src/Camera.cpp is not part of this build: it includes EleFits headers
// (Camera.cpp:23-27), which this image does not have.
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

}  // extern "C"
