// Cmem.h -- the analytic CMEM emissivity model (reference include/Cmem.h:19-97).
//
// Init keeps the reference's signature and effects, quirks included: p1..p3 are ints, the NULL
// floats mean 0, p2 takes p3's value (Cmem.cpp:90) and CalcMagneticPressure overwrites B with
// |b| * 1e-9 (Cmem.cpp:310). Init runs on the host (birp_cmem_init); GetSample and renders run
// on the GPU.
#pragma once
#include <vector>

#include "Space.h"
#include "birp.h"

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
    // The reference's private members (Cmem.h:43-96) are replaced by the post-Init state that
    // birp_cmem_init computes (the same values, bit for bit).
    birp_cmem_state state_{};
    bool initialised_ = false;
    birp_volume* device_ = nullptr;
};
