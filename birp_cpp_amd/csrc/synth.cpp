// synth.cpp -- seeded synthetic workloads: magnetosheath-shell emissivity cubes and orbit poses.
// No reference counterpart (the reference reads MHD exports, DataCube.cpp:28-138); these stand in
// for them with the same grid conventions: coordinates are float, spacing is |c[1]-c[0]| and the
// origin is c[0], exactly what DataCube::InitSpacing/InitOriginCoords would derive.
//
// Randomness is SplitMix64 used as a counter-based stream (draw d = mix(seed + (d+1) * golden)),
// so the generator splits across threads without changing a single value: draw 0 is r0; voxel v
// (linear [z][y][x] index) takes the Box-Muller normal of pair v/2 (draws 1+2p, 2+2p), cosine
// branch for even v and sine branch for odd v.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/birp.h"

namespace {
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr double kTwoPi = 6.283185307179586476925286766559;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// Uniform double in [0, 1) from draw d of stream `seed`.
inline double uniform(uint64_t seed, uint64_t d) { return (double)(mix64(seed + (d + 1) * kGolden) >> 11) * 0x1p-53; }

inline void normal_pair(uint64_t seed, uint64_t p, double& n0, double& n1) {
    double u1 = uniform(seed, 1 + 2 * p), u2 = uniform(seed, 2 + 2 * p);
    double rad = std::sqrt(-2.0 * std::log(1.0 - u1));
    double th = kTwoPi * u2;
    n0 = rad * std::cos(th);
    n1 = rad * std::sin(th);
}

unsigned worker_count() {
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 4;
    if (n > 16) n = 16;
    return n;
}
}  // namespace

extern "C" int birp_synth_shell(const birp_shell_params* p, float* values, float* origin3, float* spacing3, float* r0_out) {
    if (!p || !values || p->n < 2 || p->n > 4096 || !(p->r0_hi >= p->r0_lo)) return BIRP_ERR_INVALID_ARGUMENT;
    const int n = p->n;
    // float coordinates, as a .dat export would list them
    std::vector<float> cx(n), cy(n);
    for (int i = 0; i < n; i++) {
        cx[i] = (float)(p->x0 + (double)(p->x1 - p->x0) * i / (n - 1));
        cy[i] = (float)(p->y0 + (double)(p->y1 - p->y0) * i / (n - 1));
    }
    if (origin3) {
        origin3[0] = cx[0];
        origin3[1] = cy[0];
        origin3[2] = cy[0];
    }
    if (spacing3) {
        spacing3[0] = std::abs(cx[0] - cx[1]);
        spacing3[1] = std::abs(cy[0] - cy[1]);
        spacing3[2] = std::abs(cy[0] - cy[1]);
    }
    const double r0 = p->r0_lo + (p->r0_hi - p->r0_lo) * uniform(p->seed, 0);
    if (r0_out) *r0_out = (float)r0;
    const uint64_t total = (uint64_t)n * n * n;
    const unsigned nt = worker_count();
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        th.emplace_back([=] {
            // pair-aligned chunk
            uint64_t pairs = (total + 1) / 2;
            uint64_t p0 = pairs * t / nt, p1 = pairs * (t + 1) / nt;
            for (uint64_t pr = p0; pr < p1; pr++) {
                double nv[2];
                normal_pair(p->seed, pr, nv[0], nv[1]);
                for (int h = 0; h < 2; h++) {
                    uint64_t v = 2 * pr + h;
                    if (v >= total) break;
                    int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((uint64_t)n * n));
                    double X = cx[x], Y = cy[y], Z = cy[z];
                    double r = std::sqrt(X * X + Y * Y + Z * Z);
                    double e = p->bg;
                    if (r > 1e-9) {
                        double ct = X / r;
                        double rmp = r0 * std::pow(2.0 / (1.0 + ct), 0.6);
                        double rbs = 1.3 * rmp;
                        if (r >= rmp && r <= rbs) {
                            double q = r / 10.0;
                            e = p->amp / (q * q * q);
                        }
                    }
                    values[v] = (float)(e * std::exp(p->sigma * nv[h]));
                }
            }
        });
    }
    for (auto& x : th) x.join();
    return BIRP_OK;
}

extern "C" int birp_synth_orbit(int count, float perigee, float apogee, float incl_deg, uint64_t seed, float* positions) {
    if (count < 0 || !(perigee > 0) || !(apogee >= perigee) || (count && !positions)) return BIRP_ERR_INVALID_ARGUMENT;
    const double a = 0.5 * ((double)perigee + apogee);
    const double e = ((double)apogee - perigee) / ((double)apogee + perigee);
    const double inc = (double)incl_deg * kTwoPi / 360.0;
    const double omega = -0.25 * kTwoPi;  // argument of perigee -90 degrees
    for (int k = 0; k < count; k++) {
        double nu = kTwoPi * uniform(seed, (uint64_t)k);
        double r = a * (1 - e * e) / (1 + e * std::cos(nu));
        double px = r * std::cos(nu + omega), py = r * std::sin(nu + omega);
        positions[3 * k + 0] = (float)px;
        positions[3 * k + 1] = (float)(py * std::cos(inc));
        positions[3 * k + 2] = (float)(py * std::sin(inc));
    }
    return BIRP_OK;
}
