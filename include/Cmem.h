// Cmem.h -- the analytic CMEM emissivity model (reference include/Cmem.h:19-97).
//
// Init keeps the reference's signature and effects, quirks included: p1..p3 are ints, the NULL
// floats mean 0, p2 takes p3's value (Cmem.cpp:90) and CalcMagneticPressure overwrites B with
// |b| * 1e-9 (Cmem.cpp:310). Init runs on the host; GetSample and renders run on the GPU.
#pragma once
#include <vector>

#include "Space.h"

struct birp_volume;
struct birp_cmem_state;

class CMEM : public Space {
public:
    CMEM();
    ~CMEM() override;
    CMEM(const CMEM&) = delete;
    CMEM& operator=(const CMEM&) = delete;

    void Init(bool hack, std::vector<int> v_passed = {}, std::vector<int> b_passed = {}, float dipole_passed = 0.f,
              float p0_passed = 0.786300004f, int p1_passed = 1.f, int p2_passed = 3.f, int p3_passed = 4.f,
              float B_passed = 2.f, float alpha_passed = 2.5f, float beta_passed = -1.6f, float bs_passed = 12.6400003f,
              float A1_passed = 7.2e-06f, float A2_passed = 3.5000000000000004e-06f, float ay_bs_passed = 0.f,
              float az_bs_passed = 0.f, float density_passed = 0.f);
    void Init() override;
    float GetSample(float x, float y, float z) override;

    // ---- additions ----
    void State(birp_cmem_state* out) const;  // the post-Init members the GPU model reads
    birp_volume* DeviceVolume();             // created on first use after each Init

private:
    int temp = 200000;
    float v[3];
    float b[3];
    float pressure_dynamic;
    float pressure_magnetic;
    float dipole;
    std::vector<float> a;
    std::vector<float> beta_c;
    float charlie;
    float dn;
    float ds;
    float theta_n;
    float theta_s;
    float r0_lin;
    float p0;
    float bs;
    float A1;
    float A2;
    float B;
    float alpha;
    float beta;
    float p1;
    float p2;
    float p3;
    float ay_bs;
    float az_bs;
    float density;
    bool initialised_ = false;
    birp_volume* device_ = nullptr;

    void DefineLinearCoeffs();
    void CalcDynamicPressure();
    void CalcMagneticPressure();
    float CalcInitialAlpha();
};
