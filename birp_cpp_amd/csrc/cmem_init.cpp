// cmem_init.cpp -- CMEM::Init on the host (reference src/Cmem.cpp:18-115, 250-325): the post-Init
// model state the GPU evaluates, computed with the reference's float/double types and operation
// order (built with -ffp-contract=off) so that it is bit-identical to the reference's members.
// The quirks are kept: p1..p3 are ints, a 0 float argument means "use the default", p2 takes
// p3's value (Cmem.cpp:90), and CalcMagneticPressure overwrites B with |b| * 1e-9 (Cmem.cpp:310).
#include <cmath>

#include "../../include/birp.h"
#include "capi_internal.h"

namespace {
constexpr double kPi = 3.14159265358979323846;
inline float norm3(const float* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
}  // namespace

extern "C" void birp_cmem_default_params(birp_cmem_params* p) {
    // the default arguments of CMEM::Init (include/Cmem.h:21-39), as CMEM::Init() passes them
    *p = birp_cmem_params{};
    p->dipole = 0.f;
    p->p0 = 0.786300004f;
    p->p1 = 1;
    p->p2 = 3;
    p->p3 = 4;
    p->B = 2.f;
    p->alpha = 2.5f;
    p->beta = -1.6f;
    p->bs = 12.6400003f;
    p->A1 = 7.2e-06f;
    p->A2 = 3.5000000000000004e-06f;
}

extern "C" int birp_cmem_init(const birp_cmem_params* in, birp_cmem_state* s) {
    if (!in || !s) return birp::fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    // Cmem.cpp:35-87: solar-wind vectors, defaults for empty vectors
    const float v_default[3] = {400, 0, 0}, b_default[3] = {0, 0, 5};
    float v[3], b[3];
    for (int k = 0; k < 3; k++) {
        v[k] = in->has_v ? (float)in->v[k] : v_default[k];
        b[k] = in->has_b ? (float)in->b[k] : b_default[k];
    }
    const float dipole = in->dipole;
    float B = in->B;
    const float alpha = in->alpha, beta = in->beta;
    // Cmem.cpp:88-95
    const float p0 = in->p0 == 0 ? 0.786300004f : in->p0;
    const float p1 = in->p1 == 0 ? 1.f : (float)in->p1;
    const float p2 = in->p2 == 0 ? 3.f : (float)in->p3;  // the reference reads p3 here (Cmem.cpp:90)
    const float p3 = in->p3 == 0 ? 4.f : (float)in->p3;
    const float bs = in->bs == 0 ? 12.64f : in->bs;
    const float A1 = in->A1 == 0 ? 7.2e-06f : in->A1;
    const float A2 = in->A2 == 0 ? 3.5e-06f : in->A2;
    const float density = in->density == 0 ? 5 : in->density;
    // CalcDynamicPressure (Cmem.cpp:292-301): proton mass x density x |v|^2, in nPa
    const float proton_mass = 0.00000000000000000000000000167f;
    const float vms = norm3(v) * 1000;
    const float n = density * 1000000;
    const float pressure_dynamic = 0.5f * proton_mass * n * (vms * vms) * 1000000000;
    // CalcMagneticPressure (Cmem.cpp:305-315): B is overwritten with |b| in tesla
    B = norm3(b) * 0.000000001;
    const float mu0 = 4 * kPi * 0.0000001;
    const float pressure_magnetic = (B * B) / (2 * mu0) * 1000000000;
    // CalcInitialAlpha (Cmem.cpp:320-322) for ay_bs, az_bs passed as 0
    auto initial_alpha = [&]() -> float { return (0.58f - 0.010 * b[2]) * (1 + 0.010f * pressure_dynamic) + 0.2f; };
    const float ay_bs = in->ay_bs == 0 ? initial_alpha() : in->ay_bs;
    const float az_bs = in->az_bs == 0 ? initial_alpha() : in->az_bs;
    // DefineLinearCoeffs (Cmem.cpp:250-276): the Lin et al. coefficients
    const float a[22] = {12.544f, -0.194f, 0.305f, 0.0573f, 2.178f, 0.0571f, -0.999f, 16.473f, 0.00152f, 0.382f, 0.0431f,
                         -0.00763f, -0.210f, 0.0405f, -4.430f, -0.636f, -2.600f, 0.832f, -5.328f, 1.103f, -0.907f, 1.450f};
    const float beta_c[4] = {a[6] + a[7] * ((std::exp(a[8] * b[2]) - 1) / (std::exp(a[9] * b[2]) + 1)), a[10],
                             a[11] + a[12] * dipole, a[13]};
    s->charlie = a[14] * std::pow((pressure_dynamic + pressure_magnetic), a[15]);
    s->dn = a[16] + a[17] * dipole + a[18] * (dipole * dipole);
    s->ds = a[16] - a[17] * dipole + a[18] * (dipole * dipole);
    s->theta_n = a[19] + a[20] * dipole;
    s->theta_s = a[19] - a[20] * dipole;
    s->r0_lin = a[0] * std::pow(pressure_dynamic + pressure_magnetic, a[1]) *
                (1 + a[2] * ((std::exp(a[3] * b[2]) - 1) / (std::exp(a[4] * b[2]) + 1)));
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
    return BIRP_OK;
}
