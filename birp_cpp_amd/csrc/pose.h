// pose.h -- host-side per-pose math of Camera::Integrate (reference Camera.cpp:83-137, 309-508)
// and the Helper matrix routines it calls (Helper.cpp:70-264). FP32 versions keep the
// reference's operation order so that the GPU renders see bit-identical camera constants; FP64
// versions carry the same steps in double for the fast path and the FP64 render.
#pragma once

namespace birp {

struct Mat3f {
    float m[3][3];
};
struct Mat3d {
    double m[3][3];
};

// Helper::MatrixMultiply (Helper.cpp:70-98).
Mat3f mat_mul(const Mat3f& a, const Mat3f& b);
// Helper::GetDeterminant (Helper.cpp:150-191), 3x3 by first-row cofactor expansion.
float mat_det(const Mat3f& a);
// Helper::GetInverse (Helper.cpp:241-264); returns false when the determinant is 0.
bool mat_inverse(const Mat3f& a, Mat3f& out);
bool mat_inverse(const Mat3d& a, Mat3d& out);

// Camera::Orient + PointToPlane: the image roll angle_sun, degrees.
float orient(const float pos[3], const float aim[3]);
double orient64(const double pos[3], const double aim[3]);

// Integrate's basis: R = (Rz2 Rx) Rz and R^-1 (Camera.cpp:85-135). false on det == 0.
bool view_basis(const float pos[3], const float aim[3], float& angle_sun, Mat3f& rot, Mat3f& rinv);
bool view_basis64(const float pos[3], const float aim[3], Mat3d& rot, Mat3d& rinv);

}  // namespace birp
