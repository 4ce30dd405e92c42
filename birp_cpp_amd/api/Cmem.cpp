// Cmem.cpp -- CMEM over the C ABI: Init computes the post-Init state on the host with
// birp_cmem_init (reference src/Cmem.cpp:18-115, 250-325, bit-identical); the per-point model
// (Cmem.cpp:117-288) runs on the GPU.
#include "Cmem.h"

#include <stdexcept>

#include "birp.h"
#include "device.h"

CMEM::CMEM() {}

CMEM::~CMEM() {
    if (device_) birp_volume_free(device_);
}

void CMEM::Init() { Init(true); }

void CMEM::Init(bool /*hack*/, std::vector<int> v_passed, std::vector<int> b_passed, float dipole_passed, float p0_passed,
                int p1_passed, int p2_passed, int p3_passed, float B_passed, float alpha_passed, float beta_passed,
                float bs_passed, float A1_passed, float A2_passed, float ay_bs_passed, float az_bs_passed,
                float density_passed) {
    birp_cmem_params p{};
    p.has_v = !v_passed.empty();
    p.has_b = !b_passed.empty();
    for (int k = 0; k < 3; k++) {
        if (p.has_v) p.v[k] = v_passed.at(k);  // std::out_of_range for a short vector, as the reference's .at()
        if (p.has_b) p.b[k] = b_passed.at(k);
    }
    p.dipole = dipole_passed;
    p.p0 = p0_passed;
    p.p1 = p1_passed;
    p.p2 = p2_passed;
    p.p3 = p3_passed;
    p.B = B_passed;
    p.alpha = alpha_passed;
    p.beta = beta_passed;
    p.bs = bs_passed;
    p.A1 = A1_passed;
    p.A2 = A2_passed;
    p.ay_bs = ay_bs_passed;
    p.az_bs = az_bs_passed;
    p.density = density_passed;
    birp_api::check(birp_cmem_init(&p, &state_), "CMEM::Init");
    initialised_ = true;
    if (device_) {
        birp_volume_free(device_);
        device_ = nullptr;
    }
}

void CMEM::State(birp_cmem_state* s) const {
    if (!initialised_) throw std::runtime_error("CMEM: Init() has not been called");
    *s = state_;
}

birp_volume* CMEM::DeviceVolume() {
    if (!device_) {
        birp_cmem_state s;
        State(&s);
        birp_api::check(birp_cmem_create(birp_api::device(), &s, &device_), "CMEM upload");
    }
    return device_;
}

float CMEM::GetSample(float x, float y, float z) {
    float p[3] = {x, y, z}, out = 0.f;
    birp_api::check(birp_volume_sample(DeviceVolume(), p, 1, &out), "CMEM::GetSample");
    return out;
}
