// Cmem.cpp -- CMEM::Init on the host (reference src/Cmem.cpp:18-115, 250-325), with the
// reference's float/double types and order so the post-Init members are bit-identical; the
// per-point model (Cmem.cpp:117-288) runs on the GPU.
#include "Cmem.h"

#include <cmath>
#include <stdexcept>

#include "birp.h"
#include "device.h"

namespace {
constexpr double kPi = 3.14159265358979323846;
inline float norm3(const float* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
}  // namespace

CMEM::CMEM() {}

CMEM::~CMEM() {
    if (device_) birp_volume_free(device_);
}

void CMEM::Init() { Init(true); }

void CMEM::Init(bool /*hack*/, std::vector<int> v_passed, std::vector<int> b_passed, float dipole_passed, float p0_passed,
                int p1_passed, int p2_passed, int p3_passed, float B_passed, float alpha_passed, float beta_passed,
                float bs_passed, float A1_passed, float A2_passed, float ay_bs_passed, float az_bs_passed,
                float density_passed) {
    const float v_default[3] = {400, 0, 0}, b_default[3] = {0, 0, 5};
    for (int k = 0; k < 3; k++) {
        v[k] = v_passed.empty() ? v_default[k] : (float)v_passed.at(k);
        b[k] = b_passed.empty() ? b_default[k] : (float)b_passed.at(k);
    }
    dipole = dipole_passed;
    B = B_passed;
    alpha = alpha_passed;
    beta = beta_passed;
    p0 = p0_passed == 0 ? 0.786300004f : p0_passed;
    p1 = p1_passed == 0 ? 1.f : p1_passed;
    p2 = p2_passed == 0 ? 3.f : p3_passed;  // the reference reads p3 here (Cmem.cpp:90)
    p3 = p3_passed == 0 ? 4.f : p3_passed;
    bs = bs_passed == 0 ? 12.64f : bs_passed;
    A1 = A1_passed == 0 ? 7.2e-06f : A1_passed;
    A2 = A2_passed == 0 ? 3.5e-06f : A2_passed;
    density = density_passed == 0 ? 5 : density_passed;
    CalcDynamicPressure();
    CalcMagneticPressure();
    ay_bs = ay_bs_passed == 0 ? CalcInitialAlpha() : ay_bs_passed;
    az_bs = az_bs_passed == 0 ? CalcInitialAlpha() : az_bs_passed;
    DefineLinearCoeffs();
    initialised_ = true;
    if (device_) {
        birp_volume_free(device_);
        device_ = nullptr;
    }
}

// Cmem.cpp:250-276 (Lin et al. coefficients).
void CMEM::DefineLinearCoeffs() {
    a = {12.544f, -0.194f, 0.305f, 0.0573f, 2.178f, 0.0571f, -0.999f, 16.473f, 0.00152f, 0.382f, 0.0431f,
         -0.00763f, -0.210f, 0.0405f, -4.430f, -0.636f, -2.600f, 0.832f, -5.328f, 1.103f, -0.907f, 1.450f};
    beta_c = {a[6] + a[7] * ((std::exp(a[8] * b[2]) - 1) / (std::exp(a[9] * b[2]) + 1)), a[10], a[11] + a[12] * dipole, a[13]};
    charlie = a[14] * std::pow((pressure_dynamic + pressure_magnetic), a[15]);
    dn = a[16] + a[17] * dipole + a[18] * (dipole * dipole);
    ds = a[16] - a[17] * dipole + a[18] * (dipole * dipole);
    theta_n = a[19] + a[20] * dipole;
    theta_s = a[19] - a[20] * dipole;
    r0_lin = a[0] * std::pow(pressure_dynamic + pressure_magnetic, a[1]) *
             (1 + a[2] * ((std::exp(a[3] * b[2]) - 1) / (std::exp(a[4] * b[2]) + 1)));
}

// Cmem.cpp:292-301: proton mass x density x |v|^2, in nPa.
void CMEM::CalcDynamicPressure() {
    const float proton_mass = 0.00000000000000000000000000167f;
    const float vms = norm3(v) * 1000;
    const float n = density * 1000000;
    pressure_dynamic = 0.5f * proton_mass * n * (vms * vms) * 1000000000;
}

// Cmem.cpp:305-315: note that B (the model parameter) is overwritten with |b| in tesla.
void CMEM::CalcMagneticPressure() {
    B = norm3(b) * 0.000000001;
    const float mu0 = 4 * kPi * 0.0000001;
    pressure_magnetic = (B * B) / (2 * mu0) * 1000000000;
}

float CMEM::CalcInitialAlpha() { return (0.58f - 0.010 * b[2]) * (1 + 0.010f * pressure_dynamic) + 0.2f; }

void CMEM::State(birp_cmem_state* s) const {
    if (!initialised_) throw std::runtime_error("CMEM: Init() has not been called");
    s->p0 = p0;
    s->p1 = p1;
    s->p2 = p2;
    s->p3 = p3;
    s->bs = bs;
    s->A1 = A1;
    s->A2 = A2;
    s->B = B;
    s->alpha = alpha;
    s->beta = beta;
    s->ay_bs = ay_bs;
    s->az_bs = az_bs;
    s->a5 = a[5];
    s->a21 = a[21];
    for (int k = 0; k < 4; k++) s->beta_c[k] = beta_c[k];
    s->charlie = charlie;
    s->dn = dn;
    s->ds = ds;
    s->theta_n = theta_n;
    s->theta_s = theta_s;
    s->r0_lin = r0_lin;
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
