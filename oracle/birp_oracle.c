/*
 * birp_oracle.c -- CPU restatement of BIRP's render path. TEST INFRASTRUCTURE ONLY: the product
 * library never links or loads this file. Build: oracle/Makefile (gcc, -ffp-contract=off).
 *
 * Every function follows the reference file:line named in its comment, keeping the
 * reference's float/double types and left-to-right operation order so that the FP32 results
 * match the reference bit for bit on x86-64 (SSE arithmetic, no FMA). The C++ overloads the
 * reference resolves to are spelled out: std::sin(float) is sinf, std::pow(float, float) is
 * powf, std::cos(double) is cos, and so on.
 */
#define _GNU_SOURCE
#include "birp_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------- std::lerp (libstdc++ 11) */
/* /usr/include/c++/11/cmath:1894-1908: sign-straddling inputs use t*b + (1-t)*a; otherwise
 * a + t*(b-a) clamped monotonically towards b. */
float or_lerpf(float a, float b, float t) {
    if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0)) return t * b + (1 - t) * a;
    if (t == 1) return b;
    const float x = a + t * (b - a);
    return ((t > 1) == (b > a)) ? (b < x ? x : b) : (b > x ? x : b);
}

double or_lerp(double a, double b, double t) {
    if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0)) return t * b + (1 - t) * a;
    if (t == 1) return b;
    const double x = a + t * (b - a);
    return ((t > 1) == (b > a)) ? (b < x ? x : b) : (b > x ? x : b);
}

/* ---------------------------------------------------------------- DataCube::GetSample */
static inline float cube_at(const or_cube* c, int z, int y, int x) {
    return c->data[((size_t)z * (size_t)c->ny + (size_t)y) * (size_t)c->nx + (size_t)x];
}

/* DataCube.cpp:214-218 (FP32). */
int or_cube_inside(const or_cube* c, float x, float y, float z) {
    float xi = x - c->origin[0];
    float xd = xi / c->spacing[0];
    float yi = y - c->origin[1];
    float yd = yi / c->spacing[1];
    float zi = z - c->origin[2];
    float zd = zi / c->spacing[2];
    if ((0 > xi || xd + 1 >= c->nx) || (0 > yi || yd + 1 >= c->ny) || (0 > zi || zd + 1 >= c->nz))
        return 0;
    /* NaN coordinates pass the reference's test (every comparison is false) and then index
     * out of range; this restatement reports them outside instead. */
    if (xd != xd || yd != yd || zd != zd) return 0;
    return 1;
}

/* DataCube.cpp:203-277. */
float or_cube_sample(const or_cube* c, float x, float y, float z) {
    float x_ingress = x - c->origin[0];
    float x_index = (x_ingress / c->spacing[0]);
    float y_ingress = y - c->origin[1];
    float y_index = (y_ingress / c->spacing[1]);
    float z_ingress = z - c->origin[2];
    float z_index = (z_ingress / c->spacing[2]);

    if ((0 > x_ingress || x_index + 1 >= c->nx) || (0 > y_ingress || y_index + 1 >= c->ny) ||
        (0 > z_ingress || z_index + 1 >= c->nz))
        return -1.f;
    if (x_index != x_index || y_index != y_index || z_index != z_index) return -1.f;

    if (c->trilinear) {
        int z_min = (int)floorf(z_index), z_max = (int)ceilf(z_index);
        int y_min = (int)floorf(y_index), y_max = (int)ceilf(y_index);
        int x_min = (int)floorf(x_index), x_max = (int)ceilf(x_index);
        float s000 = cube_at(c, z_min, y_min, x_min);
        float s001 = cube_at(c, z_min, y_min, x_max);
        float s010 = cube_at(c, z_min, y_max, x_min);
        float s011 = cube_at(c, z_min, y_max, x_max);
        float s100 = cube_at(c, z_max, y_min, x_min);
        float s101 = cube_at(c, z_max, y_min, x_max);
        float s110 = cube_at(c, z_max, y_max, x_min);
        float s111 = cube_at(c, z_max, y_max, x_max);
        float x_frac = x_index - (int)x_index;
        float y_frac = y_index - (int)y_index;
        float z_frac = z_index - (int)z_index;
        float low_x1 = or_lerpf(s000, s001, x_frac);
        float low_x2 = or_lerpf(s010, s011, x_frac);
        float high_x1 = or_lerpf(s100, s101, x_frac);
        float high_x2 = or_lerpf(s110, s111, x_frac);
        float low_y = or_lerpf(low_x1, low_x2, y_frac);
        float high_y = or_lerpf(high_x1, high_x2, y_frac);
        return or_lerpf(low_y, high_y, z_frac);
    }
    return cube_at(c, (int)z_index, (int)y_index, (int)x_index);
}

/* The same sampler with every operation in double (FP64 oracle). */
double or_cube_sample_f64(const or_cube* c, double x, double y, double z) {
    double x_ingress = x - (double)c->origin[0];
    double x_index = x_ingress / (double)c->spacing[0];
    double y_ingress = y - (double)c->origin[1];
    double y_index = y_ingress / (double)c->spacing[1];
    double z_ingress = z - (double)c->origin[2];
    double z_index = z_ingress / (double)c->spacing[2];
    if ((0 > x_ingress || x_index + 1 >= c->nx) || (0 > y_ingress || y_index + 1 >= c->ny) ||
        (0 > z_ingress || z_index + 1 >= c->nz))
        return -1.0;
    if (x_index != x_index || y_index != y_index || z_index != z_index) return -1.0;
    if (c->trilinear) {
        int z0 = (int)floor(z_index), z1 = (int)ceil(z_index);
        int y0 = (int)floor(y_index), y1 = (int)ceil(y_index);
        int x0 = (int)floor(x_index), x1 = (int)ceil(x_index);
        double fx = x_index - (int)x_index, fy = y_index - (int)y_index, fz = z_index - (int)z_index;
        double lx1 = or_lerp(cube_at(c, z0, y0, x0), cube_at(c, z0, y0, x1), fx);
        double lx2 = or_lerp(cube_at(c, z0, y1, x0), cube_at(c, z0, y1, x1), fx);
        double hx1 = or_lerp(cube_at(c, z1, y0, x0), cube_at(c, z1, y0, x1), fx);
        double hx2 = or_lerp(cube_at(c, z1, y1, x0), cube_at(c, z1, y1, x1), fx);
        return or_lerp(or_lerp(lx1, lx2, fy), or_lerp(hx1, hx2, fy), fz);
    }
    return cube_at(c, (int)z_index, (int)y_index, (int)x_index);
}

/* ---------------------------------------------------------------- CMEM */
/* Helper::VectorDistance / VectorDistance2D (Helper.cpp:591-600). */
static inline float vdist3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }
static inline float vdist2(float x, float y) { return sqrtf(x * x + y * y); }

/* CMEM::DefineLinearCoeffs (Cmem.cpp:250-276). */
static void cmem_define_linear(or_cmem* m) {
    static const float a[22] = {12.544f, -0.194f, 0.305f, 0.0573f, 2.178f,  0.0571f, -0.999f, 16.473f,
                                0.00152f, 0.382f, 0.0431f, -0.00763f, -0.210f, 0.0405f, -4.430f, -0.636f,
                                -2.600f, 0.832f, -5.328f, 1.103f,  -0.907f, 1.450f};
    memcpy(m->a, a, sizeof(a));
    const float* A = m->a;
    m->beta_c[0] = A[6] + A[7] * ((expf(A[8] * m->b[2]) - 1) / (expf(A[9] * m->b[2]) + 1));
    m->beta_c[1] = A[10];
    m->beta_c[2] = A[11] + A[12] * m->dipole;
    m->beta_c[3] = A[13];
    m->charlie = A[14] * powf((m->pressure_dynamic + m->pressure_magnetic), A[15]);
    m->dn = A[16] + A[17] * m->dipole + A[18] * (m->dipole * m->dipole);
    m->ds = A[16] - A[17] * m->dipole + A[18] * (m->dipole * m->dipole);
    m->theta_n = A[19] + A[20] * m->dipole;
    m->theta_s = A[19] - A[20] * m->dipole;
    m->r0_lin = A[0] * powf(m->pressure_dynamic + m->pressure_magnetic, A[1]) *
                (1 + A[2] * ((expf(A[3] * m->b[2]) - 1) / (expf(A[4] * m->b[2]) + 1)));
}

/* CMEM::Init (Cmem.cpp:35-115) with CalcDynamicPressure (292-301), CalcMagneticPressure
 * (305-315, which overwrites B) and CalcInitialAlpha (317-322). */
void or_cmem_init(or_cmem* m, const int* v, const int* b, float dipole, float p0, int p1, int p2,
                  int p3, float B, float alpha, float beta, float bs, float A1, float A2,
                  float ay_bs, float az_bs, float density) {
    memset(m, 0, sizeof(*m));
    if (!v) {
        m->v[0] = 400; m->v[1] = 0; m->v[2] = 0;
    } else {
        m->v[0] = (float)v[0]; m->v[1] = (float)v[1]; m->v[2] = (float)v[2];
    }
    if (!b) {
        m->b[0] = 0; m->b[1] = 0; m->b[2] = 5;
    } else {
        m->b[0] = (float)b[0]; m->b[1] = (float)b[1]; m->b[2] = (float)b[2];
    }
    m->dipole = dipole;
    m->B = B;
    m->alpha = alpha;
    m->beta = beta;
    m->p0 = p0 == 0 ? 0.786300004f : p0;
    m->p1 = p1 == 0 ? 1.f : (float)p1;
    m->p2 = p2 == 0 ? 3.f : (float)p3; /* Cmem.cpp:90 reads p3_passed */
    m->p3 = p3 == 0 ? 4.f : (float)p3;
    m->bs = bs == 0 ? 12.64f : bs;
    m->A1 = A1 == 0 ? 7.2e-06f : A1;
    m->A2 = A2 == 0 ? 3.5e-06f : A2;
    m->density = density == 0 ? 5.f : density;

    /* CalcDynamicPressure */
    {
        float proton_mass = 0.00000000000000000000000000167f;
        float vms = vdist3(m->v[0], m->v[1], m->v[2]) * 1000;
        float n = m->density * 1000000;
        m->pressure_dynamic = 0.5f * proton_mass * n * (vms * vms) * 1000000000;
    }
    /* CalcMagneticPressure */
    {
        m->B = (float)((double)vdist3(m->b[0], m->b[1], m->b[2]) * 0.000000001);
        float mu0 = (float)(4 * M_PI * 0.0000001);
        m->pressure_magnetic = (m->B * m->B) / (2 * mu0) * 1000000000;
    }
    /* CalcInitialAlpha, used when the flaring parameter is passed as 0 (NULL) */
    {
        float init_alpha = (float)((0.58f - 0.010 * m->b[2]) * (1 + 0.010f * m->pressure_dynamic) + 0.2f);
        m->ay_bs = ay_bs == 0 ? init_alpha : ay_bs;
        m->az_bs = az_bs == 0 ? init_alpha : az_bs;
    }
    cmem_define_linear(m);
}

void or_cmem_init_default(or_cmem* m) {
    /* Cmem.h:21-39 defaults; the three NULL floats are 0. */
    or_cmem_init(m, NULL, NULL, 0.f, 0.786300004f, 1, 3, 4, 2.f, 2.5f, -1.6f, 12.6400003f, 7.2e-06f,
                 3.5000000000000004e-06f, 0.f, 0.f, 0.f);
}

/* CMEM::LinScaled (Cmem.cpp:188-222). */
static float cmem_lin_scaled(const or_cmem* m, float theta, float phi) {
    float phi_n = (float)acos((double)(cosf(theta) * cosf(m->theta_n)) +
                              (double)(sinf(theta) * sinf(m->theta_n)) * cos((double)phi - M_PI / 2));
    float phi_s = (float)acos((double)(cosf(theta) * cosf(m->theta_s)) +
                              (double)(sinf(theta) * sinf(m->theta_s)) * cos((double)phi - (3 * M_PI / 2)));
    float sin_phi = sinf(phi);
    float f = powf(cosf(theta / 2) + m->a[5] * sinf(2 * theta) * (1 - expf(-theta)),
                   (m->p1 * (m->beta_c[0] + m->beta_c[1] * cosf(phi) + m->beta_c[3] * (sin_phi * sin_phi))));
    float q = m->p2 * m->charlie * expf(m->p3 * m->dn * (powf(phi_n, m->a[21]))) +
              m->p2 * m->charlie * expf(m->p3 * m->ds * (powf(phi_s, m->a[21])));
    return m->p0 * m->r0_lin * f + q;
}

/* CMEM::ShueModel (Cmem.cpp:280-288). */
static float cmem_shue_model(float theta, float phi, float r0, float ay, float az) {
    float ry = r0 * powf((2 / (1 + cosf(theta))), ay);
    float rz = r0 * powf((2 / (1 + cosf(theta))), az);
    float rz_x_cos_phi = rz * cosf(phi);
    float ry_x_sin_phi = ry * sinf(phi);
    return (ry * rz) / sqrtf((rz_x_cos_phi * rz_x_cos_phi) + (ry_x_sin_phi * ry_x_sin_phi));
}

/* CMEM::GetSample (Cmem.cpp:117-186) with ShueCoords (224-246). */
float or_cmem_sample(const or_cmem* m, float x, float y, float z) {
    float r = vdist3(x, y, z);
    float theta = (x != 0) ? acosf(x / r) : 0.f;
    float phi = (y != 0) ? acosf(y / vdist2(y, z)) : 0.f;

    float radius_mp = cmem_lin_scaled(m, theta, phi);
    float radius_bs = cmem_shue_model(theta, phi, m->bs, m->ay_bs, m->az_bs);
    float thickness = 0.05f;
    float half_thickness = thickness / 2;

    if (r < radius_mp - half_thickness) {
        return 0;
    } else if (r >= radius_mp - half_thickness && r <= radius_mp + half_thickness) {
        float shue_norm = (r - (radius_mp - half_thickness)) / (radius_mp + half_thickness);
        float b0 = 0;
        float half_shue_1 = theta / 2;
        float sin_shue_1 = sinf(theta);
        float b1 = m->A1 * (expf(-m->B * (half_shue_1 * half_shue_1 * half_shue_1 * half_shue_1))) *
                   powf(r / 10, (-m->alpha - (m->beta * (sin_shue_1 * sin_shue_1))));
        return or_lerpf(b0, b1, shue_norm);
    } else if (r > radius_mp + half_thickness && r < radius_bs - half_thickness) {
        float half_shue_1 = theta / 2;
        float sin_shue_1 = sinf(theta);
        return m->A1 * (expf(-m->B * (half_shue_1 * half_shue_1 * half_shue_1 * half_shue_1))) *
               powf(r / 10, (-m->alpha - (m->beta * (sin_shue_1 * sin_shue_1))));
    } else if (r >= radius_bs - half_thickness && r <= radius_bs + half_thickness) {
        float half_shue_1 = theta / 2;
        float sin_shue_1 = sinf(theta);
        double shue_div_10 = r / 10;
        float shue_norm = (r - (radius_bs - half_thickness)) / (radius_bs + half_thickness);
        float b1 = m->A1 * (expf(-m->B * (half_shue_1 * half_shue_1 * half_shue_1 * half_shue_1))) *
                   powf(r / 10, (-m->alpha - (m->beta * (sin_shue_1 * sin_shue_1))));
        float b2 = (float)(m->A2 * (1 / (shue_div_10 * shue_div_10 * shue_div_10)));
        return or_lerpf(b1, b2, shue_norm);
    } else {
        double shue_div_10 = r / 10;
        return (float)(m->A2 * (1 / (shue_div_10 * shue_div_10 * shue_div_10)));
    }
}

/* The same model with every operation in double (FP64 oracle for CMEM renders). */
/* The FP64 model's zone radii and inner/outer emissivities at one point (Cmem.cpp:188-288). */
typedef struct {
    double r, rmp, rbs, inner, outer;
} cmem_zones64;

static cmem_zones64 cmem_zones_f64(const or_cmem* m, double x, double y, double z) {
    cmem_zones64 o;
    double r = sqrt(x * x + y * y + z * z);
    double theta = (x != 0) ? acos(x / r) : 0.0;
    double phi = (y != 0) ? acos(y / sqrt(y * y + z * z)) : 0.0;
    const float* A = m->a;
    double phi_n = acos(cos(theta) * cos((double)m->theta_n) + sin(theta) * sin((double)m->theta_n) * cos(phi - M_PI / 2));
    double phi_s = acos(cos(theta) * cos((double)m->theta_s) + sin(theta) * sin((double)m->theta_s) * cos(phi - 3 * M_PI / 2));
    double sp = sin(phi);
    double f = pow(cos(theta / 2) + A[5] * sin(2 * theta) * (1 - exp(-theta)),
                   m->p1 * (m->beta_c[0] + m->beta_c[1] * cos(phi) + (double)m->beta_c[3] * (sp * sp)));
    double q = (double)m->p2 * m->charlie * exp((double)m->p3 * m->dn * pow(phi_n, (double)A[21])) +
               (double)m->p2 * m->charlie * exp((double)m->p3 * m->ds * pow(phi_s, (double)A[21]));
    double ry = m->bs * pow(2 / (1 + cos(theta)), (double)m->ay_bs);
    double rz = m->bs * pow(2 / (1 + cos(theta)), (double)m->az_bs);
    double hs = theta / 2, ss = sin(theta);
    o.r = r;
    o.rmp = (double)m->p0 * m->r0_lin * f + q;
    o.rbs = (ry * rz) / sqrt((rz * cos(phi)) * (rz * cos(phi)) + (ry * sin(phi)) * (ry * sin(phi)));
    o.inner = (double)m->A1 * exp(-(double)m->B * (hs * hs * hs * hs)) * pow(r / 10, -(double)m->alpha - (double)m->beta * (ss * ss));
    o.outer = (double)m->A2 / ((r / 10) * (r / 10) * (r / 10));
    return o;
}

/* The zone chain of Cmem.cpp:132-176 in FP64, with the zone chosen by r_zone and the value
 * formed at the point's own r (they differ only when asking what a neighbouring zone gives). */
static double cmem_zone_value(const cmem_zones64* o, double r_zone) {
    const double h = 0.025, r = o->r;
    if (r_zone < o->rmp - h) return 0.0;
    if (r_zone <= o->rmp + h) return or_lerp(0.0, o->inner, (r - (o->rmp - h)) / (o->rmp + h));
    if (r_zone < o->rbs - h) return o->inner;
    if (r_zone <= o->rbs + h) return or_lerp(o->inner, o->outer, (r - (o->rbs - h)) / (o->rbs + h));
    return o->outer;
}

double or_cmem_sample_f64(const or_cmem* m, double x, double y, double z) {
    cmem_zones64 o = cmem_zones_f64(m, x, y, z);
    return cmem_zone_value(&o, o.r);
}

double or_cmem_edge_jump_f64(const or_cmem* m, double x, double y, double z, double rel_band) {
    cmem_zones64 o = cmem_zones_f64(m, x, y, z);
    const double h = 0.025;
    const double e[4] = {o.rmp - h, o.rmp + h, o.rbs - h, o.rbs + h};
    double jump = 0.0;
    for (int k = 0; k < 4; k++) {
        if (!(fabs(o.r - e[k]) < rel_band * o.r)) continue;
        /* the two zones meeting at e[k]: just below it and just above it (the chain compares
         * with < and <=, so both sides are reached within a relative 1e-12) */
        const double d = fabs(e[k]) * 1e-12 + 1e-300;
        jump += fabs(cmem_zone_value(&o, e[k] - d) - cmem_zone_value(&o, e[k] + d));
    }
    return jump;
}

double or_cmem_edge_distance_f64(const or_cmem* m, double x, double y, double z) {
    cmem_zones64 o = cmem_zones_f64(m, x, y, z);
    const double h = 0.025;
    double e[4] = {o.rmp - h, o.rmp + h, o.rbs - h, o.rbs + h}, best = INFINITY;
    for (int k = 0; k < 4; k++) {
        double d = fabs(o.r - e[k]);
        if (d < best) best = d;
    }
    return o.r > 0 ? best / o.r : best;
}

/* ---------------------------------------------------------------- Camera geometry */
void or_image_dims(float pixel_deg, float fov_h, float fov_w, int* W, int* H) {
    /* Camera.cpp:38-39: std::round(float) then int. */
    *H = (int)roundf(fov_h / pixel_deg);
    *W = (int)roundf(fov_w / pixel_deg);
}

void or_ray_dist(float L, int n, float* t, float* w) {
    /* Camera.cpp:72-78: dist = max * ((i + 0.5f) / ray_samples), ray_width = dist[1] - dist[0]. */
    for (int i = 0; i < n; i++) t[i] = L * ((i + 0.5f) / n);
    if (w) *w = (n > 1) ? t[1] - t[0] : 0.f;
}

void or_ray_dist_f64(double L, int n, double* t, double* w) {
    for (int i = 0; i < n; i++) t[i] = L * ((i + 0.5) / n);
    if (w) *w = (n > 1) ? t[1] - t[0] : 0.0;
}

/* Camera::PointToPlane (Camera.cpp:419-508), FP32; returns xo, yo in out[0..1]. */
static void point_to_plane(const float obj[3], float plac, float prad, const float dv[3],
                           const float u[3], const float north[3], const float right[3],
                           const float pos[3], float out[4]) {
    float to_deg = (float)(180 / M_PI);
    float snumer = (dv[0] * obj[0]) + (dv[1] * obj[1]) + (dv[2] * obj[2]) + plac;
    float sdenom = (dv[0] * u[0]) + (dv[1] * u[1]) + (dv[2] * u[2]);
    float scalar = snumer / sdenom;
    float q[3] = {obj[0] - (scalar * u[0]), obj[1] - (scalar * u[1]), obj[2] - (scalar * u[2])};
    float ov[3] = {q[0] - pos[0], q[1] - pos[1], q[2] - pos[2]};
    float dp = sqrtf((ov[0] * ov[0]) + (ov[1] * ov[1]) + (ov[2] * ov[2]));
    float ouv[3] = {ov[0] / dp, ov[1] / dp, ov[2] / dp};
    dp = sqrtf((ouv[0] * ouv[0]) + (ouv[1] * ouv[1]) + (ouv[2] * ouv[2]));

    float t1 = (ouv[0] * north[0]) + (ouv[1] * north[1]) + (ouv[2] * north[2]);
    float t2 = dp * sqrtf((north[0] * north[0]) + (north[1] * north[1]) + (north[2] * north[2]));
    float t3 = t1 / t2;
    if (t3 > 1) t3 = 1.f;
    else if (t3 < -1) t3 = -1.f;
    float apon = to_deg * acosf(t3);

    float t4 = (ouv[0] * right[0]) + (ouv[1] * right[1]) + (ouv[2] * right[2]);
    float t5 = dp * sqrtf((right[0] * right[0]) + (right[1] * right[1]) + (right[2] * right[2]));
    float t6 = t4 / t5;
    if (t6 > 1) t6 = 1.f;
    else if (t6 < -1) t6 = -1.f;
    float apor = to_deg * acosf(t6);
    float soy = 1.f, sox = 1.f;
    if (apor > 90) sox = -1.f;

    float oov[3] = {obj[0] - pos[0], obj[1] - pos[1], obj[2] - pos[2]};
    float ood = sqrtf((oov[0] * oov[0]) + (oov[1] * oov[1]) + (oov[2] * oov[2]));
    oov[0] = oov[0] / ood;
    oov[1] = oov[1] / ood;
    oov[2] = oov[2] / ood;
    float aro = to_deg * atanf(prad / ood);
    float t7 = (oov[0] * u[0]) + (oov[1] * u[1]) + (oov[2] * u[2]);
    float t8 = sqrtf((oov[0] * oov[0]) + (oov[1] * oov[1]) + (oov[2] * oov[2]));
    float t9 = t7 / t8;
    if (t9 > 1) t9 = 1.f;
    float ooa = to_deg * acosf(t9);
    float xo = 0.f, yo = 0.f;
    if (ooa > 0.000001f) {
        xo = sox * ooa * sinf(apon / to_deg);
        yo = soy * ooa * cosf(apon / to_deg);
    }
    out[0] = xo; out[1] = yo; out[2] = aro; out[3] = ooa;
}

/* Octant sign table of Camera.cpp:368-392 applied to north[0..1]. */
#define NORTH_SIGNS(T, u, north, ABS)                                              \
    do {                                                                           \
        if (u[0] >= 0 && u[1] >= 0 && u[2] >= 0) {                                 \
            north[0] = (T)-1 * ABS(north[0]); north[1] = (T)-1 * ABS(north[1]);    \
        } else if (u[0] >= 0 && u[1] >= 0 && u[2] < 0) {                           \
            north[0] = (T)1 * ABS(north[0]); north[1] = (T)1 * ABS(north[1]);      \
        } else if (u[0] >= 0 && u[1] < 0 && u[2] >= 0) {                           \
            north[0] = (T)-1 * ABS(north[0]); north[1] = (T)1 * ABS(north[1]);     \
        } else if (u[0] >= 0 && u[1] < 0 && u[2] < 0) {                            \
            north[0] = (T)1 * ABS(north[0]); north[1] = (T)-1 * ABS(north[1]);     \
        } else if (u[0] < 0 && u[1] >= 0 && u[2] >= 0) {                           \
            north[0] = (T)1 * ABS(north[0]); north[1] = (T)-1 * ABS(north[1]);     \
        } else if (u[0] < 0 && u[1] >= 0 && u[2] < 0) {                            \
            north[0] = (T)-1 * ABS(north[0]); north[1] = (T)1 * ABS(north[1]);     \
        } else if (u[0] < 0 && u[1] < 0 && u[2] >= 0) {                            \
            north[0] = (T)1 * ABS(north[0]); north[1] = (T)1 * ABS(north[1]);      \
        } else if (u[0] < 0 && u[1] < 0 && u[2] < 0) {                             \
            north[0] = (T)-1 * ABS(north[0]); north[1] = (T)-1 * ABS(north[1]);    \
        }                                                                          \
        north[2] = ABS(north[2]);                                                  \
    } while (0)

/* Camera::Orient (Camera.cpp:309-417), FP32. */
float or_orient(const float pos[3], const float aim[3]) {
    float to_deg = (float)(180.f / M_PI);
    float dv[3] = {aim[0] - pos[0], aim[1] - pos[1], aim[2] - pos[2]};
    float distance = vdist3(dv[0], dv[1], dv[2]);
    float u[3] = {dv[0] / distance, dv[1] / distance, dv[2] / distance};
    float phi = acosf(u[2]);
    float denominator = sinf(phi);
    /* theta1 (Camera.cpp:340-349) never reaches the result. */
    float theta2;
    if (fabsf(denominator) < fabsf(u[1])) {
        if (denominator * u[1] < 0)
            theta2 = (float)(acos(-1) * to_deg); /* degrees fed to a radian cosine (Q5) */
        else
            theta2 = (float)(acos(1) * to_deg);
    } else {
        theta2 = asinf(u[1] / denominator);
    }
    float north[3] = {-1.f * cosf(theta2) * cosf(phi), -1.f * sinf(theta2) * cosf(phi), sinf(phi)};
    NORTH_SIGNS(float, u, north, fabsf);
    float right[3] = {(u[1] * north[2]) - (u[2] * north[1]), (u[2] * north[0]) - (u[0] * north[2]),
                      (u[0] * north[1]) - (u[1] * north[0])};
    float plac = (-1.f * dv[0] * pos[0]) + (-1.f * dv[1] * pos[1]) + (-1.f * dv[2] * pos[2]);
    float sun[3] = {23455.f, 0.f, 0.f};
    float p2p[4];
    point_to_plane(sun, plac, 1, dv, u, north, right, pos, p2p);
    float angle_sun = 270 - (to_deg * atan2f(0 - p2p[0], 0 - p2p[1]));
    if (angle_sun > 360) angle_sun = angle_sun - 360;
    else if (angle_sun < 0) angle_sun = angle_sun + 360;
    return angle_sun;
}

/* The same orientation with every operation in double. */
double or_orient_f64(const double pos[3], const double aim[3]) {
    double to_deg = 180.0 / M_PI;
    double dv[3] = {aim[0] - pos[0], aim[1] - pos[1], aim[2] - pos[2]};
    double distance = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    double u[3] = {dv[0] / distance, dv[1] / distance, dv[2] / distance};
    double phi = acos(u[2]);
    double den = sin(phi);
    double theta2;
    if (fabs(den) < fabs(u[1])) theta2 = (den * u[1] < 0) ? acos(-1) * to_deg : 0.0;
    else theta2 = asin(u[1] / den);
    double north[3] = {-cos(theta2) * cos(phi), -sin(theta2) * cos(phi), sin(phi)};
    NORTH_SIGNS(double, u, north, fabs);
    double right[3] = {u[1] * north[2] - u[2] * north[1], u[2] * north[0] - u[0] * north[2],
                       u[0] * north[1] - u[1] * north[0]};
    double plac = -dv[0] * pos[0] - dv[1] * pos[1] - dv[2] * pos[2];
    double obj[3] = {23455.0, 0.0, 0.0};
    double snumer = dv[0] * obj[0] + dv[1] * obj[1] + dv[2] * obj[2] + plac;
    double sdenom = dv[0] * u[0] + dv[1] * u[1] + dv[2] * u[2];
    double scalar = snumer / sdenom;
    double ov[3] = {obj[0] - scalar * u[0] - pos[0], obj[1] - scalar * u[1] - pos[1], obj[2] - scalar * u[2] - pos[2]};
    double dp = sqrt(ov[0] * ov[0] + ov[1] * ov[1] + ov[2] * ov[2]);
    double ouv[3] = {ov[0] / dp, ov[1] / dp, ov[2] / dp};
    dp = sqrt(ouv[0] * ouv[0] + ouv[1] * ouv[1] + ouv[2] * ouv[2]);
    double t3 = (ouv[0] * north[0] + ouv[1] * north[1] + ouv[2] * north[2]) /
                (dp * sqrt(north[0] * north[0] + north[1] * north[1] + north[2] * north[2]));
    t3 = t3 > 1 ? 1 : (t3 < -1 ? -1 : t3);
    double apon = to_deg * acos(t3);
    double t6 = (ouv[0] * right[0] + ouv[1] * right[1] + ouv[2] * right[2]) /
                (dp * sqrt(right[0] * right[0] + right[1] * right[1] + right[2] * right[2]));
    t6 = t6 > 1 ? 1 : (t6 < -1 ? -1 : t6);
    double apor = to_deg * acos(t6);
    double sox = apor > 90 ? -1.0 : 1.0;
    double oov[3] = {obj[0] - pos[0], obj[1] - pos[1], obj[2] - pos[2]};
    double ood = sqrt(oov[0] * oov[0] + oov[1] * oov[1] + oov[2] * oov[2]);
    for (int k = 0; k < 3; k++) oov[k] /= ood;
    double t9 = (oov[0] * u[0] + oov[1] * u[1] + oov[2] * u[2]) /
                sqrt(oov[0] * oov[0] + oov[1] * oov[1] + oov[2] * oov[2]);
    if (t9 > 1) t9 = 1;
    double ooa = to_deg * acos(t9);
    double xo = 0, yo = 0;
    if (ooa > 0.000001) {
        xo = sox * ooa * sin(apon / to_deg);
        yo = ooa * cos(apon / to_deg);
    }
    double angle_sun = 270 - to_deg * atan2(0 - xo, 0 - yo);
    if (angle_sun > 360) angle_sun -= 360;
    else if (angle_sun < 0) angle_sun += 360;
    return angle_sun;
}

/* ---------------------------------------------------------------- Helper matrix math */
void or_matmul3(const float a[9], const float b[9], float out[9]) {
    /* Helper.cpp:70-98: value = value + a[i][k]*b[k][j], starting from 0.0f. */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float value = 0.0f;
            for (int k = 0; k < 3; k++) value = value + (a[i * 3 + k] * b[k * 3 + j]);
            out[i * 3 + j] = value;
        }
}

static void matmul3_f64(const double a[9], const double b[9], double out[9]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double value = 0.0;
            for (int k = 0; k < 3; k++) value = value + a[i * 3 + k] * b[k * 3 + j];
            out[i * 3 + j] = value;
        }
}

/* Helper::GetDeterminant for 2x2 and 3x3 (Helper.cpp:150-191). The 3x3 case expands along the
 * first row: result = result + sign * v[0][i] * det2(minor), left to right. */
static float det2f(float a, float b, float c, float d) { return a * d - b * c; }
static float det3f(const float m[9]) {
    float result = 0;
    int sign = 1;
    for (int i = 0; i < 3; i++) {
        float sub[4];
        int z = 0;
        for (int n = 0; n < 3; n++)
            if (n != i) { sub[z] = m[3 + n]; sub[2 + z] = m[6 + n]; z++; }
        result = result + sign * m[i] * det2f(sub[0], sub[1], sub[2], sub[3]);
        sign = -sign;
    }
    return result;
}

/* Helper::GetInverse (Helper.cpp:241-264): transpose(cofactor) * (1/det); cofactors are
 * pow(-1, i+j) * det2(minor) (Helper.cpp:207-239), an exact sign flip. */
int or_get_inverse3(const float m[9], float out[9]) {
    if (det3f(m) == 0) return -1;
    float d = 1.f / det3f(m);
    float cof[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float sub[4];
            int p = 0;
            for (int x = 0; x < 3; x++) {
                if (x == i) continue;
                for (int y = 0; y < 3; y++) {
                    if (y == j) continue;
                    sub[p++] = m[x * 3 + y];
                }
            }
            double sgn = ((i + j) & 1) ? -1.0 : 1.0;
            cof[i * 3 + j] = (float)(sgn * (double)det2f(sub[0], sub[1], sub[2], sub[3]));
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i * 3 + j] = cof[j * 3 + i] * d;
    return 0;
}

int or_get_inverse3_f64(const double m[9], double out[9]) {
    double det = 0;
    int sign = 1;
    for (int i = 0; i < 3; i++) {
        double s[4];
        int z = 0;
        for (int n = 0; n < 3; n++)
            if (n != i) { s[z] = m[3 + n]; s[2 + z] = m[6 + n]; z++; }
        det = det + sign * m[i] * (s[0] * s[3] - s[1] * s[2]);
        sign = -sign;
    }
    if (det == 0) return -1;
    double d = 1.0 / det;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s[4];
            int p = 0;
            for (int x = 0; x < 3; x++) {
                if (x == i) continue;
                for (int y = 0; y < 3; y++) {
                    if (y == j) continue;
                    s[p++] = m[x * 3 + y];
                }
            }
            double c = (((i + j) & 1) ? -1.0 : 1.0) * (s[0] * s[3] - s[1] * s[2]);
            out[j * 3 + i] = c * d;
        }
    return 0;
}

/* ---------------------------------------------------------------- Integrate: basis */
int or_view_basis(const or_view* v, float rot[9], float rinv[9]) {
    float to_rad = (float)(M_PI / 180.f);
    float angle_sun = or_orient(v->pos, v->aim);
    float asc = 180.f - angle_sun;
    float x_diff = v->aim[0] - v->pos[0];
    float theta = (float)((2 * M_PI) - (double)atan2f(x_diff, v->pos[1]));
    float hyp = sqrtf(x_diff * x_diff + v->pos[1] * v->pos[1]);
    float phi = atan2f(hyp, v->pos[2]);
    float rz[9] = {cosf(theta), -sinf(theta), 0, sinf(theta), cosf(theta), 0, 0, 0, 1};
    float rx[9] = {1, 0, 0, 0, cosf(phi), -sinf(phi), 0, sinf(phi), cosf(phi)};
    float rz2[9] = {cosf(asc * to_rad), -sinf(asc * to_rad), 0, sinf(asc * to_rad), cosf(asc * to_rad), 0, 0, 0, 1};
    float r1[9];
    or_matmul3(rz2, rx, r1);
    or_matmul3(r1, rz, rot);
    return or_get_inverse3(rot, rinv);
}

int or_view_basis_f64(const or_view* v, double rot[9], double rinv[9]) {
    double pos[3] = {v->pos[0], v->pos[1], v->pos[2]};
    double aim[3] = {v->aim[0], v->aim[1], v->aim[2]};
    double to_rad = M_PI / 180.0;
    double asc = 180.0 - or_orient_f64(pos, aim);
    double x_diff = aim[0] - pos[0];
    double theta = 2 * M_PI - atan2(x_diff, pos[1]);
    double phi = atan2(sqrt(x_diff * x_diff + pos[1] * pos[1]), pos[2]);
    double rz[9] = {cos(theta), -sin(theta), 0, sin(theta), cos(theta), 0, 0, 0, 1};
    double rx[9] = {1, 0, 0, 0, cos(phi), -sin(phi), 0, sin(phi), cos(phi)};
    double rz2[9] = {cos(asc * to_rad), -sin(asc * to_rad), 0, sin(asc * to_rad), cos(asc * to_rad), 0, 0, 0, 1};
    double r1[9];
    matmul3_f64(rz2, rx, r1);
    matmul3_f64(r1, rz, rot);
    return or_get_inverse3_f64(rot, rinv);
}

/* ---------------------------------------------------------------- Integrate: rays */
void or_ray_dir(const or_view* v, const float rinv[9], int i, int j, float d[3]) {
    float to_rad = (float)(M_PI / 180.f);
    float aspect_ratio = (float)v->W / (float)v->H;
    float device_x = (i + 0.5f) / (v->W);
    float device_y = (j + 0.5f) / (v->H);
    float screen_x = ((2 * device_x) - 1) * tanf(to_rad * (v->fov_w / 2)) * aspect_ratio;
    float screen_y = (1 - (2 * device_y)) * tanf(to_rad * (v->fov_h / 2));
    float nd = vdist3(screen_x, screen_y, -1);
    float su[3] = {screen_x / nd, screen_y / nd, -1 / nd};
    /* Helper::ApplyRotation (Helper.cpp:136-148) */
    float wv[3];
    for (int r = 0; r < 3; r++) wv[r] = (rinv[r * 3] * su[0]) + (rinv[r * 3 + 1] * su[1]) + (rinv[r * 3 + 2] * su[2]);
    float wd = vdist3(wv[0], wv[1], wv[2]);
    d[0] = wv[0] / wd;
    d[1] = wv[1] / wd;
    d[2] = wv[2] / wd;
}

void or_ray_dir_f64(const or_view* v, const double rinv[9], int i, int j, double d[3]) {
    double to_rad = M_PI / 180.0;
    double aspect = (double)v->W / (double)v->H;
    double sx = (2 * ((i + 0.5) / v->W) - 1) * tan(to_rad * (v->fov_w / 2.0)) * aspect;
    double sy = (1 - 2 * ((j + 0.5) / v->H)) * tan(to_rad * (v->fov_h / 2.0));
    double nd = sqrt(sx * sx + sy * sy + 1);
    double su[3] = {sx / nd, sy / nd, -1 / nd};
    double wv[3];
    for (int r = 0; r < 3; r++) wv[r] = rinv[r * 3] * su[0] + rinv[r * 3 + 1] * su[1] + rinv[r * 3 + 2] * su[2];
    double wd = sqrt(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
    for (int r = 0; r < 3; r++) d[r] = wv[r] / wd;
}

float or_scale_sample(float v, float w) {
    /* Camera.cpp:201, evaluated left to right: (float)(v*w) promoted, x 637810000.0 (the product
     * 1000*100*6378.1 in double), / (4/pi), / 10000, rounded back to float. */
    float sample = v;
    sample = (float)(((double)(sample * w) * (1000 * 100 * 6378.1) / (4 / M_PI)) / 10000);
    return sample;
}

/* ---------------------------------------------------------------- render (threaded rows) */
typedef struct {
    const or_space* s;
    const or_view* v;
    int semantics, row0, row1, accumulate;
    float* img32;
    double* img64;
    const float* rinv32;
    const double* rinv64;
    const float* t32;
    const double* t64;
    float w32;
    double w64;
    long long count;
} render_job;

static void render_rows_f32(render_job* jb) {
    const or_view* v = jb->v;
    const or_space* s = jb->s;
    long long count = 0;
    for (int j = jb->row0; j < jb->row1; ++j) {
        for (int i = 0; i < v->W; ++i) {
            float d[3];
            or_ray_dir(v, jb->rinv32, i, j, d);
            float acc = jb->accumulate ? jb->img32[(size_t)j * v->W + i] : 0.f;
            for (int k = 0; k < v->n; ++k) {
                /* Helper::MatrixScalarMultiply then + position (Camera.cpp:183-189) */
                float x = d[0] * jb->t32[k] + v->pos[0];
                float y = d[1] * jb->t32[k] + v->pos[1];
                float z = d[2] * jb->t32[k] + v->pos[2];
                float val;
                if (s->kind == OR_SPACE_CUBE) {
                    if (!or_cube_inside(s->cube, x, y, z)) {
                        if (jb->semantics == OR_SEM_REFERENCE) break;
                        continue;
                    }
                    val = or_cube_sample(s->cube, x, y, z);
                } else {
                    val = or_cmem_sample(s->cmem, x, y, z);
                }
                float sample = or_scale_sample(val, jb->w32);
                if (sample < 0) break; /* Camera.cpp:207-210 */
                acc = acc + sample;
                count++;
            }
            jb->img32[(size_t)j * v->W + i] = acc;
        }
    }
    jb->count = count;
}

static void render_rows_f64(render_job* jb) {
    const or_view* v = jb->v;
    const or_space* s = jb->s;
    long long count = 0;
    double pos[3] = {v->pos[0], v->pos[1], v->pos[2]};
    for (int j = jb->row0; j < jb->row1; ++j) {
        for (int i = 0; i < v->W; ++i) {
            double d[3];
            or_ray_dir_f64(v, jb->rinv64, i, j, d);
            double acc = 0.0;
            for (int k = 0; k < v->n; ++k) {
                double x = d[0] * jb->t64[k] + pos[0];
                double y = d[1] * jb->t64[k] + pos[1];
                double z = d[2] * jb->t64[k] + pos[2];
                double val;
                if (s->kind == OR_SPACE_CUBE) {
                    val = or_cube_sample_f64(s->cube, x, y, z);
                    if (val == -1.0) { /* out of bounds (the sampler's sentinel) */
                        const or_cube* c = s->cube;
                        double xi = x - c->origin[0], yi = y - c->origin[1], zi = z - c->origin[2];
                        int oob = (0 > xi || xi / c->spacing[0] + 1 >= c->nx) || (0 > yi || yi / c->spacing[1] + 1 >= c->ny) ||
                                  (0 > zi || zi / c->spacing[2] + 1 >= c->nz);
                        if (oob) {
                            if (jb->semantics == OR_SEM_REFERENCE) break;
                            continue;
                        }
                    }
                } else {
                    val = or_cmem_sample_f64(s->cmem, x, y, z);
                }
                double sample = val * jb->w64 * (1000 * 100 * 6378.1) / (4 / M_PI) / 10000;
                if (sample < 0) break;
                acc += sample;
                count++;
            }
            jb->img64[(size_t)j * v->W + i] = acc;
        }
    }
    jb->count = count;
}

static void* job_thread32(void* p) { render_rows_f32((render_job*)p); return NULL; }
static void* job_thread64(void* p) { render_rows_f64((render_job*)p); return NULL; }

static void run_jobs(render_job* base, int nthreads, int fp64) {
    if (nthreads <= 1) {
        if (fp64) render_rows_f64(base); else render_rows_f32(base);
        return;
    }
    int rows = base->row1 - base->row0;
    if (nthreads > rows) nthreads = rows > 0 ? rows : 1;
    render_job* jobs = (render_job*)calloc((size_t)nthreads, sizeof(render_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    /* Interleave rows in chunks of 4 so that threads see similar amounts of in-cube work. */
    int r = base->row0;
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = *base;
        jobs[t].row0 = base->row0 + (int)((long long)rows * t / nthreads);
        jobs[t].row1 = base->row0 + (int)((long long)rows * (t + 1) / nthreads);
        (void)r;
        pthread_create(&th[t], NULL, fp64 ? job_thread64 : job_thread32, &jobs[t]);
    }
    long long total = 0;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        total += jobs[t].count;
    }
    base->count = total;
    free(jobs);
    free(th);
}

int or_render_f32(const or_space* s, const or_view* v, int semantics, int row0, int row1,
                  int accumulate, int nthreads, float* img, long long* n_incube) {
    float rot[9], rinv[9];
    if (or_view_basis(v, rot, rinv) != 0) return -1;
    float* t = (float*)malloc(sizeof(float) * (size_t)(v->n > 0 ? v->n : 1));
    float w = 0.f;
    or_ray_dist(v->L, v->n, t, &w);
    render_job jb;
    memset(&jb, 0, sizeof(jb));
    jb.s = s; jb.v = v; jb.semantics = semantics; jb.row0 = row0; jb.row1 = row1;
    jb.accumulate = accumulate; jb.img32 = img; jb.rinv32 = rinv; jb.t32 = t; jb.w32 = w;
    run_jobs(&jb, nthreads, 0);
    if (n_incube) *n_incube = jb.count;
    free(t);
    return 0;
}

int or_render_f64(const or_space* s, const or_view* v, int semantics, int row0, int row1,
                  int nthreads, double* img, long long* n_incube) {
    double rot[9], rinv[9];
    if (or_view_basis_f64(v, rot, rinv) != 0) return -1;
    double* t = (double*)malloc(sizeof(double) * (size_t)(v->n > 0 ? v->n : 1));
    double w = 0.0;
    or_ray_dist_f64((double)v->L, v->n, t, &w);
    render_job jb;
    memset(&jb, 0, sizeof(jb));
    jb.s = s; jb.v = v; jb.semantics = semantics; jb.row0 = row0; jb.row1 = row1;
    jb.img64 = img; jb.rinv64 = rinv; jb.t64 = t; jb.w64 = w;
    run_jobs(&jb, nthreads, 1);
    if (n_incube) *n_incube = jb.count;
    free(t);
    return 0;
}

typedef struct {
    const or_cmem* m;
    const or_view* v;
    const float* rinv;
    const float* t;
    double rel_band, wk;
    int row0, row1, count;
    unsigned char* mask;
    double* bound;
} edge_job;

static void* edge_rows(void* p) {
    edge_job* jb = (edge_job*)p;
    const or_view* v = jb->v;
    for (int j = jb->row0; j < jb->row1; ++j)
        for (int i = 0; i < v->W; ++i) {
            float d[3];
            or_ray_dir(v, jb->rinv, i, j, d);
            double b = 0.0;
            unsigned char hit = 0;
            for (int k = 0; k < v->n; ++k) {
                float x = d[0] * jb->t[k] + v->pos[0], y = d[1] * jb->t[k] + v->pos[1], z = d[2] * jb->t[k] + v->pos[2];
                if (or_cmem_edge_distance_f64(jb->m, x, y, z) < jb->rel_band) {
                    hit = 1;
                    b += or_cmem_edge_jump_f64(jb->m, x, y, z, jb->rel_band) * jb->wk;
                }
            }
            jb->mask[(size_t)(j - jb->row0) * v->W + i] = hit;
            jb->bound[(size_t)(j - jb->row0) * v->W + i] = b;
            jb->count += hit;
        }
    return NULL;
}

int or_cmem_edge_bound(const or_cmem* m, const or_view* v, double rel_band, int row0, int row1, int nthreads,
                       unsigned char* mask, double* bound) {
    float rot[9], rinv[9], w;
    if (or_view_basis(v, rot, rinv) != 0) return -1;
    float* t = (float*)malloc(sizeof(float) * (size_t)(v->n > 0 ? v->n : 1));
    or_ray_dist(v->L, v->n, t, &w);
    const int rows = row1 - row0;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > rows) nthreads = rows > 0 ? rows : 1;
    edge_job* jobs = (edge_job*)calloc((size_t)nthreads, sizeof(edge_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    /* a pixel's sample adds v * w * 637810000 / (4/pi) / 10000 to the image (Camera.cpp:201) */
    const double wk = (double)w * (1000 * 100 * 6378.1) / (4 / M_PI) / 10000;
    for (int k = 0; k < nthreads; k++) {
        edge_job* jb = &jobs[k];
        jb->m = m; jb->v = v; jb->rinv = rinv; jb->t = t; jb->rel_band = rel_band; jb->wk = wk;
        jb->row0 = row0 + (int)((long long)rows * k / nthreads);
        jb->row1 = row0 + (int)((long long)rows * (k + 1) / nthreads);
        jb->mask = mask + (size_t)(jb->row0 - row0) * v->W;
        jb->bound = bound + (size_t)(jb->row0 - row0) * v->W;
        pthread_create(&th[k], NULL, edge_rows, jb);
    }
    int count = 0;
    for (int k = 0; k < nthreads; k++) {
        pthread_join(th[k], NULL);
        count += jobs[k].count;
    }
    free(jobs);
    free(th);
    free(t);
    return count;
}

int or_cmem_edge_pixels(const or_cmem* m, const or_view* v, double rel_band, unsigned char* mask) {
    float rot[9], rinv[9], w;
    if (or_view_basis(v, rot, rinv) != 0) return -1;
    float* t = (float*)malloc(sizeof(float) * (size_t)(v->n > 0 ? v->n : 1));
    or_ray_dist(v->L, v->n, t, &w);
    int count = 0;
    for (int j = 0; j < v->H; ++j)
        for (int i = 0; i < v->W; ++i) {
            float d[3];
            or_ray_dir(v, rinv, i, j, d);
            unsigned char hit = 0;
            for (int k = 0; k < v->n && !hit; ++k) {
                float x = d[0] * t[k] + v->pos[0], y = d[1] * t[k] + v->pos[1], z = d[2] * t[k] + v->pos[2];
                if (or_cmem_edge_distance_f64(m, x, y, z) < rel_band) hit = 1;
            }
            mask[(size_t)j * v->W + i] = hit;
            count += hit;
        }
    free(t);
    return count;
}

void or_ray_dirs(const or_view* v, float* dirs) {
    float rot[9], rinv[9];
    if (or_view_basis(v, rot, rinv) != 0) return;
    for (int j = 0; j < v->H; ++j)
        for (int i = 0; i < v->W; ++i) or_ray_dir(v, rinv, i, j, dirs + ((size_t)j * v->W + i) * 3);
}
