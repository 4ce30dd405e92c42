// birp_internal.h -- device-side descriptors shared by the HIP kernels (render.hip) and the host
// library (capi.cpp, server.cpp). Plain structs only: passed by value as kernel arguments or
// stored in device arrays.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace birp {

// Storage layouts kept on the device for a cube.
//  linear: the reference's [z][y][x] FP32 values (DataCube::slices flattened).
//  quad:   FP32 "quad" records, one per cell (x,y,z) with x<nx-1, y<ny-1, z<nz:
//          {v(x,y,z), v(x+1,y,z), v(x,y+1,z), v(x+1,y+1,z)} (16 B). Two records (z and z+1)
//          hold a trilinear sample's eight corners. Records are grouped in bricks of 8x8 cells
//          in x,y and 9 layers in z (the 9th is the next brick's first layer; 9216 B per brick),
//          records inside a brick [lz][ly][lx].
//  oct:    16-bit "octet" records, one per cell x<nx-1, y<ny-1, z<nz-1 holding all eight corners
//          (16 B: the four z-layer corners in quad order, then the four at z+1), bricks of 8^3
//          cells (8192 B); F16 values are divided by a power-of-two per-brick scale.
//  Bricks are placed by a code that is a sum of one term per axis: the Morton code of
//  (bx,by,bz) when the Morton index space wastes at most 2x, else bz*nby*nbx + by*nbx + bx.
//  Since Morton codes of different axes occupy disjoint bits, both are sums, so a record's
//  byte offset is tab_x[x] + tab_y[y] + tab_z[z] with one small table per axis:
//    tab_a[i] = code_a(i >> 3) * brick_bytes + (i & 7) * {16, 128, 1024}[a].
//  The kernels keep these tables in LDS: one add-chain instead of bit interleaving per sample.
enum VolumeClass : int {
    VC_POSITIVE = 0,  // every value finite and > 0: the lerps never straddle zero
    VC_FINITE = 1,    // finite, any sign
    VC_NONFINITE = 2  // some inf/nan: only the linear layout with ceil-corner semantics is exact
};

enum SpaceKind : int { SK_CUBE = 0, SK_CMEM = 1 };

struct CmemParams {
    float p0, p1, p2, p3;
    float bs, A1, A2, B, alpha, beta, ay_bs, az_bs;
    float a5, a21;
    float beta_c[4];
    float charlie, dn, ds, theta_n, theta_s, r0_lin;
    // cosf/sinf of theta_n and theta_s, constant per model: evaluated once on the host with the
    // same libm as the reference (Cmem.cpp:205, 209)
    float cos_theta_n, sin_theta_n, cos_theta_s, sin_theta_s;
};

struct DevVolume {
    int kind;    // SpaceKind
    int nx, ny, nz;
    float ox, oy, oz;     // origin (coords[0])
    float sx, sy, sz;     // spacing
    float rsx, rsy, rsz;  // RN(1/spacing), host-computed
    double rdx, rdy, rdz; // RN64(1/spacing), host-computed (exact-path divisions)
    float fnx, fny, fnz;  // (float) n per axis, as the reference's bound test compares
    int interp;           // 0 nearest, 1 trilinear
    int dtype;            // BIRP_DTYPE_*
    int vclass;           // VolumeClass
    int nbx, nby, nbz;    // bricks per axis
    const float* linear;  // may be null for 16-bit volumes
    const void* quad;     // float4 records
    const void* oct;      // uint4 records (bf16 or f16)
    const float* bscale;  // per-brick scale (F16), indexed by byte offset >> 13
    const unsigned* qtab; // quad offset tables [nx | ny | nz] (bytes)
    const unsigned* otab; // oct offset tables [nx | ny | nz] (bytes)
    unsigned quad_bytes;  // buffer range of the quad records (hardware bounds check)
    unsigned oct_bytes;
    int origin_far;       // every |origin| >= 2^-30: a nonzero ingress is never below 2^-56
    float vmin, vmax;     // smallest and largest |v| over the nonzero finite values
    int f16_uniform;      // F16 records all share one power-of-two scale (f16_scale)
    float f16_scale;      // that scale: records hold v / f16_scale
    CmemParams cmem;
};

// Per-render constants of the exact path, computed on the host in the reference's float order.
struct ExactView {
    float rinv[9];
    float px, py, pz;
    float tanx, tany, aspect;
    float L;
    float w;  // ray_width = t[1] - t[0]
    int W, H, n;
    int semantics;    // BIRP_SEM_*
    int vol;          // multi-view launches: index of the view's volume in the volume table
    int accumulate;   // start from out's values
    float margin;     // FP64 slab expansion (RE) that covers FP32 position error
    int fast_guards;  // positive cube whose scaled samples stay normal floats: only the
                      // scale's midpoint guard is needed
    float* out;       // device image H*W
    unsigned long long* counter;  // optional: samples added (one atomic per wave)
    unsigned* wave_counts;  // optional, instead of counter: each wave stores its samples added at
                            // [block * 4 + wave] (exact_wave_slots(W, H, lanes) slots)
};

// Per-pose constants of the fast path: FP64 basis and origin, FP32 march.
struct FastPose {
    double rinv[9];
    double px, py, pz;
    double tanx_aspect, tany;  // tan(fov_w/2)*aspect, tan(fov_h/2), in double
    double L;
    double inv_w, inv_h;  // 1/W, 1/H
    double sc, dt;        // n/L (samples per unit length), L/n (the step)
    int W, H, n;
    int semantics;
    int vol;          // index into the volume table
    int tiles_x;      // wave tiles across
    float w;          // step weight (t1 - t0 in double, rounded)
    float pad;
    float* out;       // device image
};

// FP64 render constants (every operation in double).
struct F64View {
    double rinv[9];
    double px, py, pz, tanx, tany, aspect, L, w;
    int W, H, n, semantics;
    double* out;
};

}  // namespace birp

// Launch wrappers implemented in render.hip (host side of the HIP translation unit).
namespace birp {
// Tile order of a single exact render (see k_exact_plan): `order` is the device's plan buffer of
// BIRP_PLAN_INTS ints (the order first, then the plan kernel's tile estimates and its ticket,
// which starts at zero); `replan` launches the plan kernel into it first, otherwise it holds the
// plan of the same view geometry, cube box and tile shape. ncu: the device's CU count.
#define BIRP_PLAN_MAX_BLOCKS 16384
#define BIRP_PLAN_BUCKETS 1024
#define BIRP_PLAN_INTS (2 * BIRP_PLAN_MAX_BLOCKS + 64)
struct ExactOrder {
    int* order;
    int ncu;
    bool replan;
    int slot;  // the host's plan slot (bookkeeping only)
};
int launch_exact(const DevVolume& v, const ExactView& ev, int lanes_per_ray, int pipe, void* stream,
                 const ExactOrder* ord = nullptr);
// the plan alone, into order (test hook)
int launch_exact_plan(const DevVolume& v, const ExactView& ev, int lanes_per_ray, int* order, void* stream);
// Slots of ExactView::wave_counts for a W x H image at `lanes_per_ray` (4 waves per block).
int exact_wave_slots(int W, int H, int lanes_per_ray);
// Sample modes of the exact kernels. SM_QUAD_POS: positive finite cube, |origin| >= 2^-30 on every
// axis and scaled samples that stay normal floats, so the only guard left is the scale's midpoint
// test; SM_QUAD_GEN: any finite cube (std::lerp's branches); SM_LIN_TRI / SM_LIN_NEAREST: the
// [z][y][x] copy (any data, nearest sampling); SM_CMEM: the analytic model.
enum SampleMode { SM_QUAD_POS = 0, SM_QUAD_GEN = 1, SM_LIN_TRI = 2, SM_LIN_NEAREST = 3, SM_CMEM = 4 };
int exact_mode_of(const DevVolume& v, const ExactView& ev);
// Lanes per ray of multi-view exact launches of many tiles: 4 for CMEM, 2 for cubes (the measured
// best of each). Exact batches of few small images take 4 for cubes too (exact_batch_setup).
inline int exact_multi_lanes(int mode) { return mode == SM_CMEM ? 4 : 2; }
// Dynamic LDS bytes of an exact block: the t table, plus the brick offset tables in quad modes.
size_t exact_lds_bytes(const DevVolume& v, int n, int mode);
// 2x2-wave tiles of a W x H image at `lanes_per_ray`.
int exact_tiles(int W, int H, int lanes_per_ray);
// Views [0, nview) of one sample mode, tiles view-major from prefix (nview + 1 entries, device),
// at `lanes` per ray (2 or 4; SM_CMEM always takes 4) with `pipe` samples in flight per lane in the
// positive-cube mode (2 or 4); block b renders item order_dev[b] when order_dev is given.
int launch_exact_multi(const DevVolume* vols_dev, const ExactView* views_dev, const int* prefix_dev, int nview, int total,
                       int mode, size_t lds, void* stream, const int* order_dev = nullptr, int lanes = 2, int pipe = 2);
// Per-item work estimates of such a list (not for SM_CMEM) at `lanes` per ray: cost_dev[g] for g < total.
int launch_plan_multi(const DevVolume* vols_dev, const ExactView* views_dev, const int* prefix_dev, int nview, int total,
                      int mode, int* cost_dev, void* stream, int lanes = 2);
int launch_f64(const DevVolume& v, const F64View& fv, void* stream);
int launch_sample_points(const DevVolume& v, const float* xyz, long long n, float* out, void* stream);
// Any other Space: the sample points of rows [row0, row0 + rows) of a view (3 floats per (pixel, k),
// k fastest), and the integration of the caller's values (one float per (pixel, k)) into the rows.
int launch_space_points(const ExactView& ev, int row0, int rows, float* xyz, void* stream);
int launch_space_integrate(const ExactView& ev, int row0, int rows, const float* vals, float* img, void* stream);
// tab_entries: the largest nx + ny + nz of the volumes the launch reads (<= 2048; any value for CMEM)
int launch_fast_batch(const DevVolume* vols_dev, int dtype, const FastPose* poses_dev, const int* tile_prefix_dev,
                      int npose, int total_tiles, unsigned* work_counter, int tile_w, int tile_h,
                      int lanes_per_ray, int unroll, int grid, void* stream, int tab_entries);
int launch_reset_counter(unsigned* counter, void* stream);
int launch_zero_f32(float* p, size_t n, void* stream);
int launch_scan_class(const float* values, size_t n, unsigned* flags_dev, void* stream);
int launch_build_quad(const DevVolume& v, void* quad, void* stream);
int launch_build_oct(const DevVolume& v, void* oct, float* bscale, void* stream);

// Fast-path kernel variants: the storage dtype (BIRP_DTYPE_*), or F16 whose records share one
// scale for the whole volume, which the kernel then applies once per ray instead of per sample;
// plus BIRP_KDT_SIGNED for volumes with values <= 0, whose kernels stop each ray at its first
// negative sample (Camera.cpp:207-210) and take one sample per lane per trip. Positive volumes
// use the kernels without the stop test.
// BIRP_KDT_CMEM is the analytic model in the fast form (FP32 with the hardware exp2/log2 and
// polynomial arc cosines; it also stops rays at negative samples).
#define BIRP_KDT_F16_UNIFORM 3
#define BIRP_KDT_SIGNED 4
#define BIRP_KDT_CMEM 8
#define BIRP_KDT_COUNT 9
inline int fast_kernel_dtype(const DevVolume& d) {
    if (d.kind == SK_CMEM) return BIRP_KDT_CMEM;
    const int base = (d.dtype == BIRP_DTYPE_F16 && d.f16_uniform) ? BIRP_KDT_F16_UNIFORM : d.dtype;
    return base | (d.vclass != VC_POSITIVE ? BIRP_KDT_SIGNED : 0);
}
int fast_grid_size(int dtype, int lanes_per_ray, int unroll, int tab_entries);
// per-wave LDS table words of a fast launch whose largest table has tab_entries entries
int fast_tab_stride(int dtype, int tab_entries);
inline int volume_tab_entries(const DevVolume& d) { return d.kind == SK_CMEM ? 0 : d.nx + d.ny + d.nz; }

// birp_render_views_noalloc: its fast launches read their descriptors from a fixed device array
// (render.hip) laid out as [DevVolume | FastPose x BIRP_NOALLOC_MAX_VIEWS | tile prefix sums of the
// three lanes-per-ray groups | work-queue counters of the three groups, 256 each].
#define BIRP_NOALLOC_MAX_VIEWS 64
#define BIRP_NOALLOC_POSE_OFF 1024
#define BIRP_NOALLOC_PREFIX_OFF (BIRP_NOALLOC_POSE_OFF + 256 * BIRP_NOALLOC_MAX_VIEWS)
#define BIRP_NOALLOC_CTR_OFF (BIRP_NOALLOC_PREFIX_OFF + 1024)
#define BIRP_NOALLOC_SCRATCH_BYTES (BIRP_NOALLOC_CTR_OFF + 3 * 256 * 4)
// The device address of this GPU's array (the current device's), from hipGetSymbolAddress.
int noalloc_scratch_address(void** out);
// Stage the volume, the prefix sums (3 groups x (BIRP_NOALLOC_MAX_VIEWS + 1) ints) and cleared
// counters, then poses [first, first + count) (count <= 4 per call), into the array at `scratch`
// by kernels on `stream` (their arguments carry the data: no copy from host memory).
int noalloc_stage_head(const DevVolume& v, const int* prefix, void* scratch, void* stream);
int noalloc_stage_poses(const FastPose* poses, int first, int count, void* scratch, void* stream);
// The grid of one persistent launch over `tiles` tiles: the occupancy-limited grid, or about one
// wave per tile when that is smaller (idle waves cost serialized work-queue claims on exit).
// One pose over one volume, both by value, as a plain grid of its tiles (k_render_fast_one): no
// descriptors on the device and no work counters; the same per-tile arithmetic as launch_fast_batch.
int launch_fast_one(const DevVolume& v, int dtype, const FastPose& pose, int total_tiles, int tile_w, int tile_h, int lanes_per_ray,
                    int unroll, void* stream, int tab_entries);
int fast_grid_for(int dtype, int lanes_per_ray, int unroll, int tiles, int tab_entries);
}  // namespace birp
