/*
 * birp_oracle.h -- CPU restatement of BIRP's render path (TEST INFRASTRUCTURE ONLY).
 *
 * This is the oracle the HIP renderer is checked against. It is never linked into, loaded
 * by, or called from the product library; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, and only to check results or as the timed CPU baseline.
 *
 * Two renderers live here:
 *   or_render_f32  -- follows the reference's operation order and float/double types
 *                     (src/Camera.cpp:83-233, src/DataCube.cpp:203-277, src/Cmem.cpp:214-422),
 *                     so that on x86-64 (no FMA, -ffp-contract=off) it reproduces the
 *                     reference image bit for bit;
 *   or_render_f64  -- the same sample grid with every operation in double (the FP64 oracle).
 *
 * Semantics flag (see OR_SEM_*): the reference stops a ray at its first out-of-bounds sample
 * (Camera.cpp:207-210 with DataCube.cpp:214-218). The extended semantics lets out-of-bounds
 * samples contribute nothing and keeps marching; for a camera whose first sample is inside
 * the cube the two agree sample for sample.
 *
 * Pinning (what checks each function, beyond this file's own logic):
 *   or_cube_sample        -- bit-exact against the reference's own DataCube::GetSample compiled
 *                            from its sources (oracle/_ref) on random and edge points, committed
 *                            as tests/golden/datacube_samples.json; linear-field identity.
 *   or_cmem_init/sample   -- the 12 known answers of tests/CMEM_test.cpp:14-90, plus the
 *                            compiled reference CMEM on a random point set (golden fixture).
 *   or_get_inverse3       -- bit-exact against the compiled reference Helper::GetInverse.
 *   or_orient             -- closed form: angle_sun == wrap(90 - ang) with ang the signed angle
 *                            of the projected Sun from the projected +z axis (tests).
 *   or_view_basis         -- invariants: R^-1 * (0,0,-1) == (A-P)/|A-P| for aim (a,0,0);
 *                            det == 1; README.md:75-146 worked angles.
 *   or_render_f32         -- constant field (CubeFake, tests/Camera_test.cpp:25-29) closed form;
 *                            linear field line integral; camera-inside == clipped extended.
 *   or_render_f64         -- agrees with or_render_f32 to FP32 rounding; linear-field integral.
 */
#ifndef BIRP_ORACLE_H
#define BIRP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_SEM_REFERENCE 0 /* stop the ray at the first out-of-bounds or negative sample */
#define OR_SEM_EXTENDED 1  /* out-of-bounds samples add 0 and the ray continues */

#define OR_SPACE_CUBE 0
#define OR_SPACE_CMEM 1

/* DataCube grid metadata and storage (include/DataCube.h:22-50): values [z][y][x], x fastest. */
typedef struct {
    int nx, ny, nz;
    float origin[3];  /* coords_{x,y,z}[0]  (DataCube.cpp:185-191) */
    float spacing[3]; /* |c[0]-c[1]| per axis (DataCube.cpp:150-183) */
    const float* data;
    int trilinear; /* 1 = trilinear, 0 = nearest (truncation) */
} or_cube;

/* CMEM post-Init state (include/Cmem.h:43-96). */
typedef struct {
    float v[3], b[3];
    float pressure_dynamic, pressure_magnetic, dipole;
    float a[22];
    float beta_c[4];
    float charlie, dn, ds, theta_n, theta_s, r0_lin;
    float p0, bs, A1, A2, B, alpha, beta, p1, p2, p3, ay_bs, az_bs, density;
} or_cmem;

typedef struct {
    int kind; /* OR_SPACE_* */
    const or_cube* cube;
    const or_cmem* cmem;
} or_space;

/* One camera pose and ray grid. */
typedef struct {
    int W, H;            /* image columns, rows (row 0 = top, Camera.cpp:151) */
    float fov_w, fov_h;  /* degrees */
    float pos[3];        /* camera position (RE) */
    float aim[3];        /* aim point; only aim[0] enters the rotation (Camera.cpp:92-98) */
    float L;             /* ray length: Camera::Render passes 44 (Camera.cpp:65) */
    int n;               /* samples per ray: 200 in the reference (Camera.cpp:31) */
} or_view;

/* libstdc++ std::lerp (float and double), the form DataCube.cpp:260-270 and Cmem.cpp:246,267 call. */
float or_lerpf(float a, float b, float t);
double or_lerp(double a, double b, double t);

float or_cube_sample(const or_cube* c, float x, float y, float z);
double or_cube_sample_f64(const or_cube* c, double x, double y, double z);
/* 1 if (x,y,z) passes DataCube.cpp:214-218's bounds test in FP32. */
int or_cube_inside(const or_cube* c, float x, float y, float z);

/* CMEM::Init with the reference's positional parameters (Cmem.cpp:35-115); pass v/b NULL for
 * "empty vector". or_cmem_init_default is CMEM::Init() (Cmem.cpp:18-34). */
void or_cmem_init(or_cmem* m, const int* v, const int* b, float dipole, float p0, int p1, int p2,
                  int p3, float B, float alpha, float beta, float bs, float A1, float A2,
                  float ay_bs, float az_bs, float density);
void or_cmem_init_default(or_cmem* m);
float or_cmem_sample(const or_cmem* m, float x, float y, float z);
double or_cmem_sample_f64(const or_cmem* m, double x, double y, double z);
/* |r - e| / r for the nearest zone edge e of r = rmp +- h, r = rbs +- h (Cmem.cpp:214-288), in
 * FP64. The model jumps at these edges (the ramps are normalised by rmp + h, not 2h), so two
 * FP32 evaluations whose libm results differ by an ulp can fall on opposite sides of one. */
double or_cmem_edge_distance_f64(const or_cmem* m, double x, double y, double z);
/* mask[j*W+i] = 1 when pixel (i,j)'s FP32 ray (traced as or_render_f32 traces it) has a sample
 * within rel_band of a zone edge; returns the number of such pixels, -1 for a singular basis. */
int or_cmem_edge_pixels(const or_cmem* m, const or_view* v, double rel_band, unsigned char* mask);
/* The change in the model's value at (x,y,z) when its r lands on the other side of each zone edge
 * e within rel_band (|r - e| < rel_band * r): the sum over those edges of |zone below e - zone above
 * e|, both formed at the point's own r, in FP64 (0 when no edge is that close). A sample whose
 * FP32 evaluations (under two libms) fall on opposite sides of an edge differ by about this. */
double or_cmem_edge_jump_f64(const or_cmem* m, double x, double y, double z, double rel_band);
/* Rows [row0, row1) of a view: mask as or_cmem_edge_pixels, and bound = the sum over the pixel's
 * near-edge samples of or_cmem_edge_jump_f64 * w * 637810000 / (4/pi) / 10000 (the largest image
 * change the zone flips of those samples can make, Camera.cpp:201). Both arrays hold
 * (row1 - row0) * W entries; threaded over rows. Returns the count of masked pixels. */
int or_cmem_edge_bound(const or_cmem* m, const or_view* v, double rel_band, int row0, int row1, int nthreads,
                       unsigned char* mask, double* bound);

/* Camera::Camera image size (Camera.cpp:38-39). */
void or_image_dims(float pixel_deg, float fov_h, float fov_w, int* W, int* H);
/* GenerateRayDistWidth (Camera.cpp:69-81): t[k], k<n, and the step weight w = t[1]-t[0]. */
void or_ray_dist(float L, int n, float* t, float* w);
void or_ray_dist_f64(double L, int n, double* t, double* w);

/* Camera::Orient + PointToPlane (Camera.cpp:309-508): the roll angle angle_sun in degrees. */
float or_orient(const float pos[3], const float aim[3]);
double or_orient_f64(const double pos[3], const double aim[3]);

/* Helper::MatrixMultiply (Helper.cpp:70-98) and Helper::GetInverse (Helper.cpp:150-264), 3x3
 * row-major. or_get_inverse3 returns -1 when the determinant is 0 (Helper.cpp:242-244). */
void or_matmul3(const float a[9], const float b[9], float out[9]);
int or_get_inverse3(const float m[9], float out[9]);
int or_get_inverse3_f64(const double m[9], double out[9]);

/* Integrate's basis (Camera.cpp:85-137): R = (Rz2 Rx) Rz and R^-1; returns -1 on det == 0. */
int or_view_basis(const or_view* v, float rot[9], float rinv[9]);
int or_view_basis_f64(const or_view* v, double rot[9], double rinv[9]);

/* Per-pixel world unit ray (Camera.cpp:148-168). */
void or_ray_dir(const or_view* v, const float rinv[9], int i, int j, float d[3]);
void or_ray_dir_f64(const or_view* v, const double rinv[9], int i, int j, double d[3]);

/* All W*H unit rays of a view (x, y, z per pixel, row-major), as or_render_f32 traces them. */
void or_ray_dirs(const or_view* v, float* dirs);

/* The reference's per-sample scale (Camera.cpp:201): f32(f64(v*w) * 637810000 / (4/pi) / 1e4). */
float or_scale_sample(float v, float w);

/* Render rows [row0, row1) of the image into img (H*W floats, row-major, row 0 = top; rows
 * outside the range are not touched). When accumulate != 0 the sums start from img's current
 * values, as Camera::Render does (Camera.cpp:223, the image is never cleared). n_incube (may be
 * NULL) receives the number of samples that were added. nthreads <= 1 runs serially.
 * Returns 0, or -1 when the basis is singular. */
int or_render_f32(const or_space* s, const or_view* v, int semantics, int row0, int row1,
                  int accumulate, int nthreads, float* img, long long* n_incube);
int or_render_f64(const or_space* s, const or_view* v, int semantics, int row0, int row1,
                  int nthreads, double* img, long long* n_incube);

#ifdef __cplusplus
}
#endif
#endif
