// pose.cpp -- see pose.h. Every float expression below is written so that its evaluation order
// and types match the reference statement cited beside it (C++ overloads: std::sin(float) is the
// float function, std::acos(int) the double one, and so on). Built without FMA contraction.
#include "pose.h"

#include <cmath>

namespace birp {

namespace {
constexpr double kPi = 3.14159265358979323846;

struct V3 {
    float x, y, z;
};

// Helper::VectorDistance (Helper.cpp:591-594)
inline float norm3(float a, float b, float c) { return std::sqrt(a * a + b * b + c * c); }
inline float dot3(const V3& a, const V3& b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
inline float norm3(const V3& a) { return std::sqrt((a.x * a.x) + (a.y * a.y) + (a.z * a.z)); }

// Octant sign table of Camera.cpp:368-394 (the north vector's x and y signs, |z|).
template <class T, class V>
void north_signs(const V& u, T north[3]) {
    const T ax = std::abs(north[0]), ay = std::abs(north[1]);
    const bool px = u[0] >= 0, py = u[1] >= 0, pz = u[2] >= 0;
    // sign of north[0] is -sgn(ux)sgn(uz), of north[1] -sgn(uy)sgn(uz), with >= 0 counted positive
    const T sx = ((px == pz) ? T(-1) : T(1));
    const T sy = ((py == pz) ? T(-1) : T(1));
    north[0] = sx * ax;
    north[1] = sy * ay;
    north[2] = std::abs(north[2]);
}
}  // namespace

Mat3f mat_mul(const Mat3f& a, const Mat3f& b) {
    Mat3f r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float value = 0.0f;
            for (int k = 0; k < 3; k++) value = value + (a.m[i][k] * b.m[k][j]);
            r.m[i][j] = value;
        }
    return r;
}

static Mat3d mat_mul(const Mat3d& a, const Mat3d& b) {
    Mat3d r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double value = 0.0;
            for (int k = 0; k < 3; k++) value = value + a.m[i][k] * b.m[k][j];
            r.m[i][j] = value;
        }
    return r;
}

// 2x2 minor determinant of rows/cols other than (skip_r, skip_c): v00*v11 - v01*v10.
template <class M, class T>
static T minor2(const M& a, int skip_r, int skip_c) {
    T s[4];
    int p = 0;
    for (int x = 0; x < 3; x++) {
        if (x == skip_r) continue;
        for (int y = 0; y < 3; y++) {
            if (y == skip_c) continue;
            s[p++] = a.m[x][y];
        }
    }
    return s[0] * s[3] - s[1] * s[2];
}

float mat_det(const Mat3f& a) {
    float result = 0;
    int sign = 1;
    for (int i = 0; i < 3; i++) {
        result = result + sign * a.m[0][i] * minor2<Mat3f, float>(a, 0, i);
        sign = -sign;
    }
    return result;
}

static double mat_det(const Mat3d& a) {
    double result = 0;
    int sign = 1;
    for (int i = 0; i < 3; i++) {
        result = result + sign * a.m[0][i] * minor2<Mat3d, double>(a, 0, i);
        sign = -sign;
    }
    return result;
}

bool mat_inverse(const Mat3f& a, Mat3f& out) {
    if (mat_det(a) == 0) return false;
    float d = 1.f / mat_det(a);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            // cofactor(i,j) = pow(-1, i+j) * minor, an exact sign flip (Helper.cpp:235);
            // transposed into (j,i), then scaled by d (Helper.cpp:255-261)
            float cof = ((i + j) & 1) ? -minor2<Mat3f, float>(a, i, j) : minor2<Mat3f, float>(a, i, j);
            out.m[j][i] = cof * d;
        }
    return true;
}

bool mat_inverse(const Mat3d& a, Mat3d& out) {
    double det = mat_det(a);
    if (det == 0) return false;
    double d = 1.0 / det;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double cof = ((i + j) & 1) ? -minor2<Mat3d, double>(a, i, j) : minor2<Mat3d, double>(a, i, j);
            out.m[j][i] = cof * d;
        }
    return true;
}

// Camera::Orient (Camera.cpp:309-417) with PointToPlane (419-508) for obj = Sun, prad = 1.
float orient(const float pos[3], const float aim[3]) {
    const float to_deg = 180.f / kPi;
    const V3 P{pos[0], pos[1], pos[2]};
    const V3 dv{aim[0] - pos[0], aim[1] - pos[1], aim[2] - pos[2]};
    const float distance = norm3(dv.x, dv.y, dv.z);
    const float u[3] = {dv.x / distance, dv.y / distance, dv.z / distance};
    const V3 U{u[0], u[1], u[2]};
    const float phi = std::acos(u[2]);
    const float den = std::sin(phi);
    float theta2;
    if (std::abs(den) < std::abs(u[1]))
        theta2 = (den * u[1] < 0) ? (float)(std::acos(-1) * to_deg) : (float)(std::acos(1) * to_deg);
    else
        theta2 = std::asin(u[1] / den);
    float north[3] = {-1.f * std::cos(theta2) * std::cos(phi), -1.f * std::sin(theta2) * std::cos(phi), std::sin(phi)};
    north_signs<float>(u, north);
    const V3 N{north[0], north[1], north[2]};
    const V3 R{(u[1] * north[2]) - (u[2] * north[1]), (u[2] * north[0]) - (u[0] * north[2]),
               (u[0] * north[1]) - (u[1] * north[0])};
    const float plac = (-1.f * dv.x * P.x) + (-1.f * dv.y * P.y) + (-1.f * dv.z * P.z);

    // PointToPlane: the Sun projected onto the plane through P normal to u.
    const float ptp_to_deg = 180 / kPi;
    const V3 obj{23455.f, 0.f, 0.f};
    const float snumer = dot3(dv, obj) + plac;
    const float scalar = snumer / dot3(dv, U);
    const V3 q{obj.x - (scalar * U.x), obj.y - (scalar * U.y), obj.z - (scalar * U.z)};
    V3 ov{q.x - P.x, q.y - P.y, q.z - P.z};
    float dp = norm3(ov);
    const V3 ouv{ov.x / dp, ov.y / dp, ov.z / dp};
    dp = norm3(ouv);
    float t3 = dot3(ouv, N) / (dp * norm3(N));
    t3 = t3 > 1 ? 1.f : (t3 < -1 ? -1.f : t3);
    const float apon = ptp_to_deg * std::acos(t3);
    float t6 = dot3(ouv, R) / (dp * norm3(R));
    t6 = t6 > 1 ? 1.f : (t6 < -1 ? -1.f : t6);
    const float apor = ptp_to_deg * std::acos(t6);
    const float sox = apor > 90 ? -1.f : 1.f, soy = 1.f;
    V3 oov{obj.x - P.x, obj.y - P.y, obj.z - P.z};
    const float ood = norm3(oov);
    oov = V3{oov.x / ood, oov.y / ood, oov.z / ood};
    float t9 = dot3(oov, U) / norm3(oov);
    if (t9 > 1) t9 = 1.f;
    const float ooa = ptp_to_deg * std::acos(t9);
    float xo = 0.f, yo = 0.f;
    if (ooa > 0.000001f) {
        xo = sox * ooa * std::sin(apon / ptp_to_deg);
        yo = soy * ooa * std::cos(apon / ptp_to_deg);
    }
    float angle_sun = 270 - (to_deg * std::atan2(0 - xo, 0 - yo));
    if (angle_sun > 360) angle_sun = angle_sun - 360;
    else if (angle_sun < 0) angle_sun = angle_sun + 360;
    return angle_sun;
}

double orient64(const double pos[3], const double aim[3]) {
    const double to_deg = 180.0 / kPi;
    const double dv[3] = {aim[0] - pos[0], aim[1] - pos[1], aim[2] - pos[2]};
    const double dist = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    const double u[3] = {dv[0] / dist, dv[1] / dist, dv[2] / dist};
    const double phi = std::acos(u[2]);
    const double den = std::sin(phi);
    double theta2;
    if (std::abs(den) < std::abs(u[1])) theta2 = (den * u[1] < 0) ? std::acos(-1.0) * to_deg : 0.0;
    else theta2 = std::asin(u[1] / den);
    double north[3] = {-std::cos(theta2) * std::cos(phi), -std::sin(theta2) * std::cos(phi), std::sin(phi)};
    north_signs<double>(u, north);
    const double right[3] = {u[1] * north[2] - u[2] * north[1], u[2] * north[0] - u[0] * north[2],
                             u[0] * north[1] - u[1] * north[0]};
    const double obj[3] = {23455.0, 0.0, 0.0};
    const double plac = -dv[0] * pos[0] - dv[1] * pos[1] - dv[2] * pos[2];
    const double scalar = (dv[0] * obj[0] + dv[1] * obj[1] + dv[2] * obj[2] + plac) / (dv[0] * u[0] + dv[1] * u[1] + dv[2] * u[2]);
    double ov[3] = {obj[0] - scalar * u[0] - pos[0], obj[1] - scalar * u[1] - pos[1], obj[2] - scalar * u[2] - pos[2]};
    double dp = std::sqrt(ov[0] * ov[0] + ov[1] * ov[1] + ov[2] * ov[2]);
    for (double& c : ov) c /= dp;
    dp = std::sqrt(ov[0] * ov[0] + ov[1] * ov[1] + ov[2] * ov[2]);
    auto nrm = [](const double* a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); };
    double t3 = (ov[0] * north[0] + ov[1] * north[1] + ov[2] * north[2]) / (dp * nrm(north));
    t3 = t3 > 1 ? 1 : (t3 < -1 ? -1 : t3);
    double t6 = (ov[0] * right[0] + ov[1] * right[1] + ov[2] * right[2]) / (dp * nrm(right));
    t6 = t6 > 1 ? 1 : (t6 < -1 ? -1 : t6);
    const double apon = to_deg * std::acos(t3), apor = to_deg * std::acos(t6);
    const double sox = apor > 90 ? -1.0 : 1.0;
    double oov[3] = {obj[0] - pos[0], obj[1] - pos[1], obj[2] - pos[2]};
    const double ood = nrm(oov);
    for (double& c : oov) c /= ood;
    double t9 = (oov[0] * u[0] + oov[1] * u[1] + oov[2] * u[2]) / nrm(oov);
    if (t9 > 1) t9 = 1;
    const double ooa = to_deg * std::acos(t9);
    double xo = 0, yo = 0;
    if (ooa > 0.000001) {
        xo = sox * ooa * std::sin(apon / to_deg);
        yo = ooa * std::cos(apon / to_deg);
    }
    double angle_sun = 270 - to_deg * std::atan2(0 - xo, 0 - yo);
    if (angle_sun > 360) angle_sun -= 360;
    else if (angle_sun < 0) angle_sun += 360;
    return angle_sun;
}

bool view_basis(const float pos[3], const float aim[3], float& angle_sun, Mat3f& rot, Mat3f& rinv) {
    const float to_rad = kPi / 180.f;
    angle_sun = orient(pos, aim);
    const float asc = 180.f - angle_sun;                         // Camera.cpp:89
    const float x_diff = aim[0] - pos[0];                         // :92
    const float theta = (2 * kPi) - std::atan2(x_diff, pos[1]);  // :93
    const float hyp = std::sqrt(x_diff * x_diff + pos[1] * pos[1]);
    const float phi = std::atan2(hyp, pos[2]);                    // :98
    const Mat3f rz{{{std::cos(theta), -std::sin(theta), 0}, {std::sin(theta), std::cos(theta), 0}, {0, 0, 1}}};
    const Mat3f rx{{{1, 0, 0}, {0, std::cos(phi), -std::sin(phi)}, {0, std::sin(phi), std::cos(phi)}}};
    const Mat3f rz2{{{std::cos(asc * to_rad), -std::sin(asc * to_rad), 0},
                     {std::sin(asc * to_rad), std::cos(asc * to_rad), 0},
                     {0, 0, 1}}};
    rot = mat_mul(mat_mul(rz2, rx), rz);  // :133-134
    return mat_inverse(rot, rinv);        // :135
}

bool view_basis64(const float posf[3], const float aimf[3], Mat3d& rot, Mat3d& rinv) {
    const double pos[3] = {posf[0], posf[1], posf[2]}, aim[3] = {aimf[0], aimf[1], aimf[2]};
    const double to_rad = kPi / 180.0;
    const double asc = 180.0 - orient64(pos, aim);
    const double x_diff = aim[0] - pos[0];
    const double theta = 2 * kPi - std::atan2(x_diff, pos[1]);
    const double phi = std::atan2(std::sqrt(x_diff * x_diff + pos[1] * pos[1]), pos[2]);
    const Mat3d rz{{{std::cos(theta), -std::sin(theta), 0}, {std::sin(theta), std::cos(theta), 0}, {0, 0, 1}}};
    const Mat3d rx{{{1, 0, 0}, {0, std::cos(phi), -std::sin(phi)}, {0, std::sin(phi), std::cos(phi)}}};
    const Mat3d rz2{{{std::cos(asc * to_rad), -std::sin(asc * to_rad), 0},
                     {std::sin(asc * to_rad), std::cos(asc * to_rad), 0},
                     {0, 0, 1}}};
    rot = mat_mul(mat_mul(rz2, rx), rz);
    return mat_inverse(rot, rinv);
}

}  // namespace birp
