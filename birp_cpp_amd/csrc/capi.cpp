// capi.cpp -- the extern "C" boundary (include/birp.h): devices, volumes, single renders and
// batched renders. Host code only; kernels are reached through the launch wrappers of render.hip.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/birp.h"
#include "birp_internal.h"
#include "capi_internal.h"
#include "pose.h"

using namespace birp;

#define BIRP_SCALE_K_HOST (1000 * 100 * 6378.1)
#define BIRP_SCALE_C_HOST (4 / 3.14159265358979323846)

namespace birp {

static thread_local std::string g_last_error;

int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

int hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? BIRP_ERR_OUT_OF_MEMORY : BIRP_ERR_DEVICE;
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

int validate_view(const birp_view* v, bool exact) {
    if (!v) return fail(BIRP_ERR_INVALID_ARGUMENT, "view is null");
    if (v->width <= 0 || v->height <= 0 || (long long)v->width * v->height > (1ll << 28))
        return fail(BIRP_ERR_INVALID_ARGUMENT, "image size out of range");
    if (v->samples < 2) return fail(BIRP_ERR_INVALID_ARGUMENT, "samples must be >= 2 (ray_width reads t[1])");
    if (exact && v->samples > 8192) return fail(BIRP_ERR_UNSUPPORTED, "exact renders take at most 8192 samples per ray");
    if (v->samples > (1 << 20)) return fail(BIRP_ERR_INVALID_ARGUMENT, "samples out of range");
    if (!(v->ray_length > 0) || !std::isfinite(v->ray_length) || v->ray_length > 1e6f)
        return fail(BIRP_ERR_INVALID_ARGUMENT, "ray_length must be finite and > 0");
    if (!finite3(v->position) || !finite3(v->aim)) return fail(BIRP_ERR_INVALID_ARGUMENT, "non-finite pose");
    for (int k = 0; k < 3; k++)
        if (std::abs(v->position[k]) > 1e6f || std::abs(v->aim[k]) > 1e6f)
            return fail(BIRP_ERR_INVALID_ARGUMENT, "pose coordinate out of range");
    if (!std::isfinite(v->fov_w) || !std::isfinite(v->fov_h)) return fail(BIRP_ERR_INVALID_ARGUMENT, "non-finite fov");
    if (v->semantics != BIRP_SEM_REFERENCE && v->semantics != BIRP_SEM_EXTENDED)
        return fail(BIRP_ERR_INVALID_ARGUMENT, "unknown semantics");
    if (v->math != BIRP_MATH_EXACT && v->math != BIRP_MATH_FAST) return fail(BIRP_ERR_INVALID_ARGUMENT, "unknown math");
    return BIRP_OK;
}

// t_k and the step weight exactly as GenerateRayDistWidth (Camera.cpp:69-81).
float ray_width(float L, int n) {
    volatile float t0 = L * ((0 + 0.5f) / n);
    volatile float t1 = L * ((1 + 0.5f) / n);
    return t1 - t0;
}

int exact_view(const birp_volume* vol, const birp_view* v, float* out_dev, ExactView& ev) {
    float angle_sun;
    Mat3f rot, rinv;
    if (!view_basis(v->position, v->aim, angle_sun, rot, rinv)) return fail(BIRP_ERR_SINGULAR, "Determinant is 0");
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            if (!std::isfinite(rinv.m[r][c])) return fail(BIRP_ERR_INVALID_ARGUMENT, "degenerate pose (camera at the aim point?)");
            ev.rinv[r * 3 + c] = rinv.m[r][c];
        }
    const float to_rad = 3.14159265358979323846 / 180.f;  // Camera.cpp:85
    ev.px = v->position[0];
    ev.py = v->position[1];
    ev.pz = v->position[2];
    ev.tanx = std::tan(to_rad * (v->fov_w / 2));
    ev.tany = std::tan(to_rad * (v->fov_h / 2));
    ev.aspect = (float)v->width / (float)v->height;  // Camera.cpp:137
    ev.L = v->ray_length;
    ev.w = ray_width(v->ray_length, v->samples);
    ev.W = v->width;
    ev.H = v->height;
    ev.n = v->samples;
    ev.semantics = v->semantics;
    ev.accumulate = (v->flags & BIRP_RENDER_ACCUMULATE) ? 1 : 0;
    const DevVolume& d = vol->dv;
    double scale = std::abs((double)v->position[0]) + std::abs((double)v->position[1]) + std::abs((double)v->position[2]) +
                   v->ray_length + std::abs((double)d.ox) + std::abs((double)d.oy) + std::abs((double)d.oz) +
                   (double)d.nx * d.sx + (double)d.ny * d.sy + (double)d.nz * d.sz + 1.0;
    ev.margin = (float)(scale * 0x1p-16);
    // positive cube whose scaled samples v*w*K stay normal floats: the scale needs only its
    // midpoint guard (interpolated values stay within [vmin, vmax])
    const double kf = (1.0 * BIRP_SCALE_K_HOST / BIRP_SCALE_C_HOST) / 10000;
    ev.fast_guards = (d.vclass == VC_POSITIVE && (double)d.vmin * ev.w * kf >= 0x1p-124 && (double)d.vmax * ev.w * kf <= 0x1p126) ? 1 : 0;
    ev.out = out_dev;
    ev.vol = 0;
    ev.counter = nullptr;
    ev.wave_counts = nullptr;
    return BIRP_OK;
}

int fast_pose(const birp_volume* vol, const birp_view* v, int vol_index, float* out_dev, int tile_w, int tile_h, FastPose& fp) {
    if (vol->dv.kind == SK_CUBE && ((vol->desc.dtype == BIRP_DTYPE_F32 ? !vol->dv.quad : !vol->dv.oct) ||
                                    vol->dv.nx + vol->dv.ny + vol->dv.nz > 2048))
        return fail(BIRP_ERR_UNSUPPORTED, "fast renders need a bricked volume with nx+ny+nz <= 2048");
    Mat3d rot, rinv;
    if (!view_basis64(v->position, v->aim, rot, rinv)) return fail(BIRP_ERR_SINGULAR, "Determinant is 0");
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            if (!std::isfinite(rinv.m[r][c])) return fail(BIRP_ERR_INVALID_ARGUMENT, "degenerate pose (camera at the aim point?)");
            fp.rinv[r * 3 + c] = rinv.m[r][c];
        }
    const double to_rad = 3.14159265358979323846 / 180.0;
    fp.px = v->position[0];
    fp.py = v->position[1];
    fp.pz = v->position[2];
    fp.tanx_aspect = std::tan(to_rad * (v->fov_w / 2.0)) * ((double)v->width / v->height);
    fp.tany = std::tan(to_rad * (v->fov_h / 2.0));
    fp.L = v->ray_length;
    fp.inv_w = 1.0 / v->width;
    fp.inv_h = 1.0 / v->height;
    fp.sc = (double)v->samples / v->ray_length;
    fp.dt = (double)v->ray_length / v->samples;
    fp.W = v->width;
    fp.H = v->height;
    fp.n = v->samples;
    fp.semantics = v->semantics;
    fp.vol = vol_index;
    fp.tiles_x = (v->width + tile_w - 1) / tile_w;
    fp.w = (float)((double)v->ray_length / v->samples);
    if ((fast_kernel_dtype(vol->dv) & ~BIRP_KDT_SIGNED) == BIRP_KDT_F16_UNIFORM) fp.w *= vol->dv.f16_scale;  // power of two: exact
    fp.pad = 0.f;
    fp.out = out_dev;
    (void)tile_h;
    return BIRP_OK;
}

int tiles_of(const birp_view& v, int tile_w, int tile_h) {
    return ((v.width + tile_w - 1) / tile_w) * ((v.height + tile_h - 1) / tile_h);
}

int device_scratch(birp_device* dev, size_t bytes, void** out) {
    if (dev->scratch_bytes < bytes) {
        if (dev->scratch) (void)hipFree(dev->scratch);
        dev->scratch = nullptr;
        dev->scratch_bytes = 0;
        hipError_t e = hipMalloc(&dev->scratch, bytes);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(scratch)");
        dev->scratch_bytes = bytes;
    }
    *out = dev->scratch;
    return BIRP_OK;
}

int pinned_staging(birp_device* dev, size_t bytes) {
    if (dev->pinned_bytes < bytes) {
        if (dev->pinned) (void)hipHostFree(dev->pinned);
        dev->pinned = nullptr;
        dev->pinned_dev = nullptr;
        dev->pinned_bytes = 0;
        hipError_t e = hipHostMalloc(&dev->pinned, bytes, hipHostMallocDefault);
        if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(staging)");
        memset(dev->pinned, 0, bytes);  // fault the pages in now, not on the first render's copy
        dev->pinned_bytes = bytes;
        if (hipHostGetDevicePointer(&dev->pinned_dev, dev->pinned, 0) != hipSuccess) {
            (void)hipGetLastError();
            dev->pinned_dev = nullptr;  // not mapped: images come back by DMA
        }
    }
    return BIRP_OK;
}

bool take_injected_failure(birp_device* dev) {
    int c = dev->inject_failures.load();
    while (c > 0) {
        if (dev->inject_failures.compare_exchange_weak(c, c - 1)) return true;
    }
    return false;
}

// Tuning knobs of the exact path, read once: BIRP_EXACT_LANES (lanes per ray: 1, 2, 4, 8, 16),
// BIRP_EXACT_PIPE (samples in flight per lane in the positive-cube loop: 1..4) and
// BIRP_EXACT_LAYOUT ("linear" reads the [z][y][x] copy instead of the quad bricks).
// By default a single exact render takes 16 lanes per ray for images of at most 8,000 rays, 8 for
// at most 20,000, 4 for CMEM and for images of at most 60,000 rays, and 2 above: a small image's
// few waves are bound by the chain of each ray's samples, which more lanes cut (ms per render over
// 6 orbit poses at n = 800 with 4 samples in flight, 2 / 4 / 8 lanes: 144x84 0.032 / 0.020 / 0.016,
// 288x168 0.038 / 0.027 / 0.030, 576x336 0.071 / 0.075 / 0.089; 8 / 16 lanes at n = 800: 72x42
// 0.0133 / 0.0110, 100x60 0.0138 / 0.0122, 144x84 0.0152 / 0.0143, 200x120 0.0182 / 0.0202, and
// at n = 320 144x84 0.0104 / 0.0106).
int default_exact_lanes(bool cmem, int W, int H) {
    static int lanes = [] {
        const char* e = getenv("BIRP_EXACT_LANES");
        int v = e ? atoi(e) : 0;
        return (v == 1 || v == 2 || v == 4 || v == 8 || v == 16) ? v : 0;
    }();
    if (lanes) return lanes;
    if (cmem) return 4;
    const long long rays = (long long)W * H;
    return rays <= 8000 ? 16 : (rays <= 20000 ? 8 : (rays <= 60000 ? 4 : 2));
}
// Single positive-cube renders keep 4 samples in flight per lane: their waves are few and wait
// on each sample's loads (the 256 orbit views rendered one by one: 8.04 ms against 8.85 with 2;
// the 576x336 x 800 view 0.0900 against 0.0916 ms). Exact batches of more blocks than the GPU
// holds at once keep 2 (exact_batch_setup): their many waves hide the loads already, and 4 made
// the orbit batch 1-2% slower.
int default_exact_pipe() {
    static int pipe = [] {
        const char* e = getenv("BIRP_EXACT_PIPE");
        int v = e ? atoi(e) : 4;
        return (v >= 1 && v <= 4) ? v : 4;
    }();
    return pipe;
}
// BIRP_EXACT_ORDER: "natural" renders exact views in their natural tile order (block b renders
// tile b), "replan" plans the tile order on every single render, and by default a plan is reused
// while the view's geometry, the cube's box and the tile shape stay the same. 0 natural,
// 1 cached, 2 replan.
// birp_debug_exact_order overrides it (-1: no override).
static std::atomic<int> g_exact_order_override{-1};
int exact_order_policy() {
    static int policy = [] {
        const char* e = getenv("BIRP_EXACT_ORDER");
        if (e && strcmp(e, "natural") == 0) return 0;
        if (e && strcmp(e, "replan") == 0) return 2;
        return 1;
    }();
    const int o = g_exact_order_override.load();
    return o >= 0 ? o : policy;
}
bool exact_force_linear() {
    static bool lin = [] {
        const char* e = getenv("BIRP_EXACT_LAYOUT");
        return e && strcmp(e, "linear") == 0;
    }();
    return lin;
}

}  // namespace birp

// ============================================================================ library / device
extern "C" const char* birp_strerror(int status) {
    switch (status) {
        case BIRP_OK: return "ok";
        case BIRP_ERR_INVALID_ARGUMENT: return "invalid argument";
        case BIRP_ERR_SINGULAR: return "Determinant is 0";
        case BIRP_ERR_OUT_OF_MEMORY: return "out of device memory";
        case BIRP_ERR_DEVICE: return "HIP runtime error";
        case BIRP_ERR_NO_DEVICE: return "no GPU device";
        case BIRP_ERR_ADMISSION: return "request exceeds the serving memory budget";
        case BIRP_ERR_NOT_FOUND: return "not found";
        case BIRP_ERR_TIMEOUT: return "timed out";
        case BIRP_ERR_INJECTED: return "injected failure";
        case BIRP_ERR_UNSUPPORTED: return "unsupported";
        case BIRP_ERR_CANCELLED: return "cancelled";
        default: return "unknown status";
    }
}

extern "C" const char* birp_last_error(void) { return g_last_error.c_str(); }
extern "C" int birp_version(void) { return 10000; }

extern "C" int birp_device_open(int ordinal, birp_device** out) {
    if (!out) return fail(BIRP_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(BIRP_ERR_NO_DEVICE, "no HIP device visible");
    }
    if (ordinal < 0 || ordinal >= count) return fail(BIRP_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    if ((e = hipSetDevice(ordinal)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    birp_device* d = new birp_device();
    d->ordinal = ordinal;
    const char* what = nullptr;
    if ((e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking)) != hipSuccess) what = "hipStreamCreate";
    else if ((e = hipEventCreate(&d->ev0)) != hipSuccess || (e = hipEventCreate(&d->ev1)) != hipSuccess ||
             (e = hipEventCreate(&d->ev_a)) != hipSuccess || (e = hipEventCreate(&d->ev_b)) != hipSuccess)
        what = "hipEventCreate";
    else if ((e = hipMalloc(&d->d_counter, 64)) != hipSuccess) what = "hipMalloc(counter)";
    else if ((e = hipMalloc(&d->plan_order, birp_device::kPlanSlots * BIRP_PLAN_INTS * sizeof(int))) != hipSuccess)
        what = "hipMalloc(tile plans)";
    else if ((e = hipMemsetAsync(d->plan_order, 0, birp_device::kPlanSlots * BIRP_PLAN_INTS * sizeof(int), d->stream)) != hipSuccess)
        what = "hipMemsetAsync(tile plans)";  // the plan kernel's tickets start at zero
    else if ((e = hipDeviceGetAttribute(&d->ncu, hipDeviceAttributeMultiprocessorCount, ordinal)) != hipSuccess)
        what = "hipDeviceGetAttribute(CUs)";
    else if (noalloc_scratch_address(&d->noalloc_scratch) != BIRP_OK) {
        e = hipErrorInvalidSymbol;
        what = "hipGetSymbolAddress(scratch)";
    }
    if (what) {
        int rc = hip_fail(e, what);
        birp_device_close(d);
        return rc;
    }
    *out = d;
    return BIRP_OK;
}

extern "C" int birp_device_close(birp_device* dev) {
    if (!dev) return BIRP_OK;
    (void)hipSetDevice(dev->ordinal);
    if (dev->stream) (void)hipStreamSynchronize(dev->stream);
    if (dev->scratch) (void)hipFree(dev->scratch);
    if (dev->pinned) (void)hipHostFree(dev->pinned);
    if (dev->rbuf) (void)hipFree(dev->rbuf);
    if (dev->d_counter) (void)hipFree(dev->d_counter);
    if (dev->plan_order) (void)hipFree(dev->plan_order);
    for (hipEvent_t ev : {dev->ev0, dev->ev1, dev->ev_a, dev->ev_b})
        if (ev) (void)hipEventDestroy(ev);
    if (dev->stream) (void)hipStreamDestroy(dev->stream);
    delete dev;
    return BIRP_OK;
}

extern "C" int birp_device_synchronize(birp_device* dev) {
    if (!dev) return fail(BIRP_ERR_INVALID_ARGUMENT, "device is null");
    hipError_t e = hipStreamSynchronize(dev->stream);
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "hipStreamSynchronize");
}

extern "C" int birp_timer_begin(birp_device* dev) {
    if (!dev) return fail(BIRP_ERR_INVALID_ARGUMENT, "device is null");
    hipError_t e = hipEventRecord(dev->ev_a, dev->stream);
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "hipEventRecord");
}

extern "C" int birp_timer_end(birp_device* dev, double* ms) {
    if (!dev || !ms) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    hipError_t e = hipEventRecord(dev->ev_b, dev->stream);
    if (e == hipSuccess) e = hipEventSynchronize(dev->ev_b);
    float f = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, dev->ev_a, dev->ev_b);
    if (e != hipSuccess) return hip_fail(e, "timer");
    *ms = f;
    return BIRP_OK;
}

extern "C" int birp_debug_inject_failure(birp_device* dev, int count) {
    if (!dev || count < 0) return fail(BIRP_ERR_INVALID_ARGUMENT, "bad argument");
    dev->inject_failures.store(count);
    return BIRP_OK;
}

extern "C" int birp_debug_exact_order(int policy, int* previous) {
    if (policy < -1 || policy > 2) return fail(BIRP_ERR_INVALID_ARGUMENT, "policy must be -1, 0, 1 or 2");
    const int was = exact_order_policy();
    g_exact_order_override.store(policy);
    if (previous) *previous = was;
    return BIRP_OK;
}

extern "C" int birp_debug_exact_plan(birp_volume* vol, const birp_view* view, int* order, int capacity, int* nblocks) {
    if (!vol || !view || !order || !nblocks) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    int rc = validate_view(view, true);
    if (rc) return rc;
    if (vol->dv.kind != SK_CUBE) return fail(BIRP_ERR_UNSUPPORTED, "tile plans are made for cubes");
    birp_device* dev = vol->dev;
    (void)hipSetDevice(dev->ordinal);
    ExactView ev;
    if ((rc = exact_view(vol, view, nullptr, ev))) return rc;
    const int lanes = default_exact_lanes(false, ev.W, ev.H);
    const int nb = exact_tiles(ev.W, ev.H, lanes);
    *nblocks = nb;
    if (nb > BIRP_PLAN_MAX_BLOCKS) return fail(BIRP_ERR_UNSUPPORTED, "view has more tiles than a plan holds");
    if (capacity < nb) return fail(BIRP_ERR_INVALID_ARGUMENT, "order capacity below the view's tile count");
    dev->plan_valid[0] = false;  // plan slot 0 is reused here
    rc = launch_exact_plan(vol->dv, ev, lanes, dev->plan_order, dev->stream);
    if (rc) return fail(rc, "plan launch failed");
    hipError_t e = hipMemcpyAsync(order, dev->plan_order, (size_t)nb * sizeof(int), hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "plan copy");
}

// ============================================================================ volumes
static void free_volume_buffers(birp_volume* v) {
    if (v->linear) (void)hipFree(v->linear);
    if (v->quad) (void)hipFree(v->quad);
    if (v->oct) (void)hipFree(v->oct);
    if (v->bscale) (void)hipFree(v->bscale);
    if (v->qtab) (void)hipFree(v->qtab);
    if (v->otab) (void)hipFree(v->otab);
    v->linear = nullptr;
    v->quad = nullptr;
    v->oct = nullptr;
    v->bscale = nullptr;
    v->qtab = nullptr;
    v->otab = nullptr;
}

namespace birp {
// Spread the low 10 bits of b to every third bit (Morton interleave of one axis).
static uint32_t spread3(uint32_t b) {
    uint32_t x = b & 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Per-axis byte-offset tables of a bricked record layout (see birp_internal.h): returns the
// number of brick slots the layout spans, or 0 when the byte offsets would not fit 32 bits.
static size_t brick_tables(int nx, int ny, int nz, unsigned brick_bytes, std::vector<unsigned>& tab, bool* morton) {
    const int nbx = (nx - 1 + 7) / 8, nby = (ny - 1 + 7) / 8, nbz = (nz - 1 + 7) / 8;
    int bits = 0;
    while ((1 << bits) < std::max(nbx, std::max(nby, nbz))) bits++;
    const double dense = (double)nbx * nby * nbz;
    const double mspace = std::ldexp(1.0, 3 * bits);
    *morton = bits <= 10 && mspace <= 2.0 * dense;
    const double slots = *morton ? mspace : dense;
    if (slots * brick_bytes >= 4294967296.0) return 0;
    tab.assign((size_t)nx + ny + nz, 0u);
    const unsigned sub[3] = {16u, 128u, 1024u};
    const int dims[3] = {nx, ny, nz};
    size_t base = 0;
    for (int a = 0; a < 3; a++) {
        for (int i = 0; i < dims[a]; i++) {
            uint32_t b = (uint32_t)(i >> 3), code;
            if (*morton) code = spread3(b) << a;
            else code = a == 0 ? b : (a == 1 ? b * (uint32_t)nbx : b * (uint32_t)nbx * (uint32_t)nby);
            tab[base + i] = code * brick_bytes + (uint32_t)(i & 7) * sub[a];
        }
        base += dims[a];
    }
    return (size_t)slots;
}
}  // namespace birp

extern "C" int birp_volume_upload(birp_device* dev, const birp_volume_desc* desc, const float* values, birp_volume** out) {
    if (!dev || !desc || !values || !out) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const int nx = desc->nx, ny = desc->ny, nz = desc->nz;
    if (nx < 2 || ny < 2 || nz < 2 || nx > 16384 || ny > 16384 || nz > 16384)
        return fail(BIRP_ERR_INVALID_ARGUMENT, "cube dimensions must be in [2, 16384]");
    const size_t nvox = (size_t)nx * ny * nz;
    if (nvox > (size_t)1 << 31) return fail(BIRP_ERR_INVALID_ARGUMENT, "cube larger than 2^31 voxels");
    for (int k = 0; k < 3; k++) {
        float s = desc->spacing[k], o = desc->origin[k];
        if (!std::isfinite(s) || !(s >= 0x1p-30f) || !(s <= 0x1p30f))
            return fail(BIRP_ERR_INVALID_ARGUMENT, "spacing must be finite and in [2^-30, 2^30]");
        if (!std::isfinite(o) || std::abs(o) > 1e6f) return fail(BIRP_ERR_INVALID_ARGUMENT, "origin out of range");
    }
    if (desc->dtype != BIRP_DTYPE_F32 && desc->dtype != BIRP_DTYPE_BF16 && desc->dtype != BIRP_DTYPE_F16)
        return fail(BIRP_ERR_INVALID_ARGUMENT, "unknown dtype");
    if (desc->interp != BIRP_INTERP_NEAREST && desc->interp != BIRP_INTERP_TRILINEAR)
        return fail(BIRP_ERR_INVALID_ARGUMENT, "unknown interpolation");
    if (desc->dtype != BIRP_DTYPE_F32 && desc->interp != BIRP_INTERP_TRILINEAR)
        return fail(BIRP_ERR_UNSUPPORTED, "16-bit storage supports trilinear only");
    (void)hipSetDevice(dev->ordinal);

    birp_volume* v = new birp_volume();
    v->dev = dev;
    v->desc = *desc;
    DevVolume& d = v->dv;
    memset(&d, 0, sizeof(d));
    d.kind = SK_CUBE;
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.ox = desc->origin[0];
    d.oy = desc->origin[1];
    d.oz = desc->origin[2];
    d.sx = desc->spacing[0];
    d.sy = desc->spacing[1];
    d.sz = desc->spacing[2];
    {
        volatile float one = 1.0f;
        d.rsx = one / d.sx;
        d.rsy = one / d.sy;
        d.rsz = one / d.sz;
        volatile double done = 1.0;
        d.rdx = done / (double)d.sx;
        d.rdy = done / (double)d.sy;
        d.rdz = done / (double)d.sz;
    }
    d.fnx = (float)nx;
    d.fny = (float)ny;
    d.fnz = (float)nz;
    d.interp = desc->interp;
    d.dtype = desc->dtype;
    d.nbx = (nx - 1 + 7) / 8;
    d.nby = (ny - 1 + 7) / 8;
    d.nbz = (nz - 1 + 7) / 8;

    d.origin_far = (std::abs(d.ox) >= 0x1p-30f && std::abs(d.oy) >= 0x1p-30f && std::abs(d.oz) >= 0x1p-30f) ? 1 : 0;

    hipError_t e;
    int rc = BIRP_OK;
    unsigned flags[3] = {0u, 0x7F7FFFFFu, 0u};
    unsigned* dflags = (unsigned*)dev->d_counter;
    size_t quad_bytes = 0, oct_bytes = 0, slots = 0;
    bool morton = false;
    std::vector<unsigned> tab;
    if ((e = hipMalloc(&v->linear, nvox * sizeof(float))) != hipSuccess) { rc = hip_fail(e, "hipMalloc(cube)"); goto fail_out; }
    if ((e = hipMemcpyAsync(v->linear, values, nvox * sizeof(float), hipMemcpyHostToDevice, dev->stream)) != hipSuccess) {
        rc = hip_fail(e, "hipMemcpy(cube)");
        goto fail_out;
    }
    d.linear = v->linear;
    if ((e = hipMemcpyAsync(dflags, flags, sizeof(flags), hipMemcpyHostToDevice, dev->stream)) != hipSuccess) {
        rc = hip_fail(e, "class scan init");
        goto fail_out;
    }
    if ((rc = launch_scan_class(v->linear, nvox, dflags, dev->stream)) != BIRP_OK) goto fail_out;
    if ((e = hipMemcpyAsync(flags, dflags, sizeof(flags), hipMemcpyDeviceToHost, dev->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(dev->stream)) != hipSuccess) {
        rc = hip_fail(e, "class scan");
        goto fail_out;
    }
    d.vclass = (flags[0] & 2u) ? VC_NONFINITE : (flags[0] & 1u) ? VC_FINITE : VC_POSITIVE;
    memcpy(&d.vmin, &flags[1], 4);
    memcpy(&d.vmax, &flags[2], 4);
    if (desc->dtype == BIRP_DTYPE_F32) {
        if (desc->interp == BIRP_INTERP_TRILINEAR && d.vclass != VC_NONFINITE &&
            (slots = brick_tables(nx, ny, nz, 576u * 16u, tab, &morton)) != 0 && (size_t)nx + ny + nz <= 8192) {
            quad_bytes = slots * 576u * 16u;
            if ((e = hipMalloc(&v->qtab, tab.size() * 4)) != hipSuccess) { rc = hip_fail(e, "hipMalloc(tables)"); goto fail_out; }
            if ((e = hipMemcpy(v->qtab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) { rc = hip_fail(e, "tables"); goto fail_out; }
            if ((e = hipMalloc(&v->quad, quad_bytes)) != hipSuccess) { rc = hip_fail(e, "hipMalloc(quad)"); goto fail_out; }
            if ((e = hipMemsetAsync(v->quad, 0, quad_bytes, dev->stream)) != hipSuccess) { rc = hip_fail(e, "memset(quad)"); goto fail_out; }
            d.quad = v->quad;
            d.qtab = v->qtab;
            d.quad_bytes = (unsigned)quad_bytes;
            if ((rc = launch_build_quad(d, v->quad, dev->stream)) != BIRP_OK) goto fail_out;
        }
    } else {
        if (d.vclass == VC_NONFINITE) { rc = fail(BIRP_ERR_UNSUPPORTED, "16-bit storage needs finite values"); goto fail_out; }
        if ((slots = brick_tables(nx, ny, nz, 512u * 16u, tab, &morton)) == 0 || (size_t)nx + ny + nz > 2048) {
            rc = fail(BIRP_ERR_UNSUPPORTED, "16-bit storage supports cubes with nx+ny+nz <= 2048");
            goto fail_out;
        }
        oct_bytes = slots * 512u * 16u;
        if ((e = hipMalloc(&v->otab, tab.size() * 4)) != hipSuccess) { rc = hip_fail(e, "hipMalloc(tables)"); goto fail_out; }
        if ((e = hipMemcpy(v->otab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) { rc = hip_fail(e, "tables"); goto fail_out; }
        if ((e = hipMalloc(&v->oct, oct_bytes)) != hipSuccess) { rc = hip_fail(e, "hipMalloc(oct)"); goto fail_out; }
        if ((e = hipMalloc(&v->bscale, slots * sizeof(float))) != hipSuccess) { rc = hip_fail(e, "hipMalloc(scale)"); goto fail_out; }
        if ((e = hipMemsetAsync(v->oct, 0, oct_bytes, dev->stream)) != hipSuccess) { rc = hip_fail(e, "memset(oct)"); goto fail_out; }
        d.oct = v->oct;
        d.otab = v->otab;
        d.oct_bytes = (unsigned)oct_bytes;
        d.bscale = v->bscale;
        if (desc->dtype == BIRP_DTYPE_F16) {
            // One power-of-two scale for the whole volume when every nonzero |v| then lands in
            // F16's normal range [2^-14, 2^15): no per-sample scale lookup in the kernel.
            double s = 1.0;
            if (d.vmax > 0.f) {
                int ex = 0;
                std::frexp((double)d.vmax, &ex);  // vmax = f * 2^ex, f in [0.5, 1)
                s = std::ldexp(1.0, ex - 15);     // vmax / s < 2^15
            }
            d.f16_uniform = (d.vmax == 0.f || (double)d.vmin / s >= 0x1p-14) && s >= 0x1p-126 && s <= 0x1p120;
            d.f16_scale = d.f16_uniform ? (float)s : 1.f;
        }
        if ((rc = launch_build_oct(d, v->oct, v->bscale, dev->stream)) != BIRP_OK) goto fail_out;
    }
    if ((e = hipStreamSynchronize(dev->stream)) != hipSuccess) { rc = hip_fail(e, "layout build"); goto fail_out; }
    if (desc->dtype != BIRP_DTYPE_F32) {
        // 16-bit volumes keep only their 16-bit records on the device
        (void)hipFree(v->linear);
        v->linear = nullptr;
        d.linear = nullptr;
    }
    v->morton = morton;
    v->bytes = (v->linear ? nvox * 4 : 0) + quad_bytes + oct_bytes + (v->bscale ? slots * 4 : 0) + tab.size() * 4;
    *out = v;
    return BIRP_OK;
fail_out:
    (void)hipStreamSynchronize(dev->stream);
    free_volume_buffers(v);
    delete v;
    return rc;
}

extern "C" int birp_cmem_create(birp_device* dev, const birp_cmem_state* s, birp_volume** out) {
    if (!dev || !s || !out) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    birp_volume* v = new birp_volume();
    v->dev = dev;
    memset(&v->desc, 0, sizeof(v->desc));
    v->desc.dtype = BIRP_DTYPE_F32;
    memset(&v->dv, 0, sizeof(v->dv));
    v->dv.kind = SK_CMEM;
    CmemParams& c = v->dv.cmem;
    c.p0 = s->p0; c.p1 = s->p1; c.p2 = s->p2; c.p3 = s->p3;
    c.bs = s->bs; c.A1 = s->A1; c.A2 = s->A2; c.B = s->B; c.alpha = s->alpha; c.beta = s->beta;
    c.ay_bs = s->ay_bs; c.az_bs = s->az_bs; c.a5 = s->a5; c.a21 = s->a21;
    for (int k = 0; k < 4; k++) c.beta_c[k] = s->beta_c[k];
    c.charlie = s->charlie; c.dn = s->dn; c.ds = s->ds; c.theta_n = s->theta_n; c.theta_s = s->theta_s; c.r0_lin = s->r0_lin;
    c.cos_theta_n = std::cos(c.theta_n);  // float overloads: cosf/sinf, as the reference calls them
    c.sin_theta_n = std::sin(c.theta_n);
    c.cos_theta_s = std::cos(c.theta_s);
    c.sin_theta_s = std::sin(c.theta_s);
    v->dv.dtype = BIRP_DTYPE_F32;
    v->dv.interp = BIRP_INTERP_TRILINEAR;
    v->bytes = 0;
    *out = v;
    return BIRP_OK;
}

extern "C" int birp_volume_free(birp_volume* v) {
    if (!v) return BIRP_OK;
    (void)hipSetDevice(v->dev->ordinal);
    (void)hipStreamSynchronize(v->dev->stream);
    free_volume_buffers(v);
    delete v;
    return BIRP_OK;
}

extern "C" int birp_volume_bytes(const birp_volume* v, size_t* bytes) {
    if (!v || !bytes) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    *bytes = v->bytes;
    return BIRP_OK;
}

extern "C" int birp_volume_sample(birp_volume* v, const float* xyz, long long n, float* out) {
    if (!v || (n > 0 && (!xyz || !out)) || n < 0) return fail(BIRP_ERR_INVALID_ARGUMENT, "bad argument");
    if (n > (1ll << 32)) return fail(BIRP_ERR_INVALID_ARGUMENT, "at most 2^32 points per call");
    if (n == 0) return BIRP_OK;
    if (v->dv.kind == SK_CUBE && !v->dv.linear) return fail(BIRP_ERR_UNSUPPORTED, "point sampling needs an FP32 volume");
    birp_device* dev = v->dev;
    (void)hipSetDevice(dev->ordinal);
    void* buf = nullptr;
    size_t bytes = (size_t)n * 16;
    int rc = device_scratch(dev, bytes, &buf);
    if (rc) return rc;
    float* dxyz = (float*)buf;
    float* dout = dxyz + 3 * n;
    hipError_t e = hipMemcpyAsync(dxyz, xyz, (size_t)n * 12, hipMemcpyHostToDevice, dev->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(points)");
    if ((rc = launch_sample_points(v->dv, dxyz, n, dout, dev->stream)) != BIRP_OK) return fail(rc, "sample launch failed");
    e = hipMemcpyAsync(out, dout, (size_t)n * 4, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "sample");
}

// ============================================================================ any other Space
namespace birp {
// The exact-path view constants of a view without a volume (the geometry does not depend on it).
static int space_view(const birp_view* view, int row0, int row1, ExactView& ev) {
    birp_view v = *view;
    v.math = BIRP_MATH_EXACT;
    int rc = validate_view(&v, false);
    if (rc) return rc;
    if (row0 < 0 || row1 > view->height || row0 >= row1) return fail(BIRP_ERR_INVALID_ARGUMENT, "rows out of range");
    if ((long long)(row1 - row0) * view->width * view->samples > (1ll << 28))
        return fail(BIRP_ERR_INVALID_ARGUMENT, "at most 2^28 samples per call: pass fewer rows");
    birp_volume none;
    none.dv.kind = SK_CMEM;  // no cube: margin and guards are unused here
    none.dv.vclass = VC_FINITE;
    return exact_view(&none, &v, nullptr, ev);
}
}  // namespace birp

extern "C" int birp_space_points(birp_device* dev, const birp_view* view, int row0, int row1, float* xyz) {
    if (!dev || !view || !xyz) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    ExactView ev;
    int rc = space_view(view, row0, row1, ev);
    if (rc) return rc;
    (void)hipSetDevice(dev->ordinal);
    const size_t bytes = (size_t)(row1 - row0) * view->width * view->samples * 12;
    void* buf = nullptr;
    if ((rc = device_scratch(dev, bytes, &buf))) return rc;
    if ((rc = launch_space_points(ev, row0, row1 - row0, (float*)buf, dev->stream))) return fail(rc, "points launch failed");
    hipError_t e = hipMemcpyAsync(xyz, buf, bytes, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "space points");
}

extern "C" int birp_space_integrate(birp_device* dev, const birp_view* view, int row0, int row1, const float* values, float* image) {
    if (!dev || !view || !values || !image) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    ExactView ev;
    int rc = space_view(view, row0, row1, ev);
    if (rc) return rc;
    (void)hipSetDevice(dev->ordinal);
    const size_t nval = (size_t)(row1 - row0) * view->width * view->samples, npix = (size_t)(row1 - row0) * view->width;
    void* buf = nullptr;
    if ((rc = device_scratch(dev, (nval + npix) * 4, &buf))) return rc;
    float* dvals = (float*)buf;
    float* dimg = dvals + nval;
    hipError_t e = hipMemcpyAsync(dvals, values, nval * 4, hipMemcpyHostToDevice, dev->stream);
    if (e == hipSuccess && ev.accumulate) e = hipMemcpyAsync(dimg, image, npix * 4, hipMemcpyHostToDevice, dev->stream);
    if (e != hipSuccess) return hip_fail(e, "space values upload");
    if ((rc = launch_space_integrate(ev, row0, row1 - row0, dvals, dimg, dev->stream))) return fail(rc, "integrate launch failed");
    e = hipMemcpyAsync(image, dimg, npix * 4, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "space integrate");
}

// ============================================================================ single renders
extern "C" int birp_view_basis(const birp_view* view, float* angle_sun, float rot[9], float rinv[9]) {
    if (!view || !finite3(view->position) || !finite3(view->aim)) return fail(BIRP_ERR_INVALID_ARGUMENT, "bad view");
    float a;
    Mat3f r, ri;
    bool ok = view_basis(view->position, view->aim, a, r, ri);
    if (angle_sun) *angle_sun = a;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            if (rot) rot[i * 3 + j] = r.m[i][j];
            if (rinv && ok) rinv[i * 3 + j] = ri.m[i][j];
        }
    return ok ? BIRP_OK : fail(BIRP_ERR_SINGULAR, "Determinant is 0");
}

// The buffer of birp_render and birp_count_samples: a u64 in-cube sample counter at offset 0 and
// the image at offset 256, so that one copy brings both back. The counter is not cleared per
// render: each render reads how far it advanced. It is cleared when the buffer is (re)allocated
// and after a render that failed on the device side.
static int render_buffer(birp_device* dev, size_t npix, float** img, unsigned long long** cnt) {
    const size_t need = 256 + npix * 4;
    if (dev->rbuf_bytes < need) {
        if (dev->rbuf) (void)hipFree(dev->rbuf);
        dev->rbuf = nullptr;
        dev->rbuf_bytes = 0;
        dev->count_known = false;
        hipError_t e = hipMalloc(&dev->rbuf, need);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(render buffer)");
        dev->rbuf_bytes = need;
    }
    if (!dev->count_known) {
        int rc = launch_zero_f32((float*)dev->rbuf, 2, dev->stream);
        if (rc) return fail(rc, "counter reset");
        dev->count_base = 0;
        dev->count_known = true;
    }
    *cnt = (unsigned long long*)dev->rbuf;
    *img = (float*)((char*)dev->rbuf + 256);
    return BIRP_OK;
}

// The tile order of a single exact render (k_exact_plan). A view (its geometry, the cube's box and
// the tile shape) that one of the device's plan slots was made for reuses that plan. A view seen
// before among the last few rendered without a plan gets one, in the least recently used slot.
// A view seen for the first time renders in the natural order (null): a plan costs about 10 us,
// which the render of a new view often does not repay (a stream of 256 distinct orbit views took
// 12.25 ms planning each, 10.98 ms in the natural order), while a view that comes back repays it
// on every later render (576x336 at n = 800: 0.121 ms unplanned, 0.105 ms planned). Views of
// fewer tiles than CUs or more than BIRP_PLAN_MAX_BLOCKS, and CMEM views (every ray takes all n
// samples), keep the natural order. BIRP_EXACT_ORDER=replan plans every render in slot 0.
static const ExactOrder* exact_order(birp_device* dev, const DevVolume& dv, const ExactView& ev, int lanes, ExactOrder& ord) {
    const int policy = exact_order_policy();
    if (policy == 0 || !dev->plan_order || dev->ncu <= 0 || dv.kind == SK_CMEM) return nullptr;
    const int nblocks = exact_tiles(ev.W, ev.H, lanes);
    if (nblocks < dev->ncu || nblocks > BIRP_PLAN_MAX_BLOCKS) return nullptr;
    ExactPlanKey k;
    memset(&k, 0, sizeof k);
    for (int q = 0; q < 9; ++q) k.rinv[q] = ev.rinv[q];
    k.px = ev.px, k.py = ev.py, k.pz = ev.pz;
    k.tanx = ev.tanx, k.tany = ev.tany, k.aspect = ev.aspect, k.L = ev.L;
    k.ox = dv.ox, k.oy = dv.oy, k.oz = dv.oz, k.sx = dv.sx, k.sy = dv.sy, k.sz = dv.sz;
    k.nx = dv.nx, k.ny = dv.ny, k.nz = dv.nz;
    k.W = ev.W, k.H = ev.H, k.n = ev.n, k.semantics = ev.semantics, k.lanes = lanes;
    ord.ncu = dev->ncu;
    dev->plan_clock++;
    int slot = -1;
    if (policy == 2) {
        slot = 0;
        ord.replan = true;
    } else {
        for (int q = 0; q < birp_device::kPlanSlots; q++)
            if (dev->plan_valid[q] && memcmp(&k, &dev->plan_key[q], sizeof k) == 0) slot = q;
        ord.replan = slot < 0;
        if (slot < 0) {
            bool seen = false;
            for (int q = 0; q < birp_device::kSeen; q++)
                if (dev->seen_valid[q] && memcmp(&k, &dev->seen_key[q], sizeof k) == 0) seen = true;
            if (!seen) {  // first sight: natural order, remember the view
                dev->seen_key[dev->seen_next] = k;
                dev->seen_valid[dev->seen_next] = true;
                dev->seen_next = (dev->seen_next + 1) % birp_device::kSeen;
                return nullptr;
            }
            slot = 0;  // the least recently used slot (an empty one first)
            for (int q = 1; q < birp_device::kPlanSlots; q++)
                if (dev->plan_used[q] < dev->plan_used[slot]) slot = q;
        }
    }
    ord.order = dev->plan_order + (size_t)slot * BIRP_PLAN_INTS;
    dev->plan_used[slot] = dev->plan_clock;
    if (ord.replan) {
        dev->plan_key[slot] = k;
        dev->plan_valid[slot] = false;  // until the plan's launch is enqueued
    }
    ord.slot = slot;
    return &ord;
}

// `counter` (a device u64 that each wave adds its in-cube samples to) or `wave_counts` (one slot
// per wave, exact_wave_slots of the image, each stored by its wave), or neither.
static int render_exact_into(birp_volume* vol, const birp_view* view, float* dimg, unsigned long long* counter,
                             unsigned* wave_counts = nullptr) {
    birp_device* dev = vol->dev;
    if (vol->dv.kind == SK_CUBE && !vol->dv.linear) return fail(BIRP_ERR_UNSUPPORTED, "exact renders need an FP32 volume");
    ExactView ev;
    int rc = exact_view(vol, view, dimg, ev);
    if (rc) return rc;
    ev.counter = wave_counts ? nullptr : counter;
    ev.wave_counts = wave_counts;
    if (take_injected_failure(dev)) return fail(BIRP_ERR_INJECTED, "injected failure");
    DevVolume dv = vol->dv;
    if (exact_force_linear() && dv.vclass != VC_NONFINITE && dv.kind == SK_CUBE) dv.vclass = VC_NONFINITE;
    const int lanes = default_exact_lanes(dv.kind == SK_CMEM, ev.W, ev.H);
    ExactOrder ord{};
    const ExactOrder* use = exact_order(dev, dv, ev, lanes, ord);
    rc = launch_exact(dv, ev, lanes, default_exact_pipe(), dev->stream, use);
    if (use && use->replan) dev->plan_valid[use->slot] = rc == BIRP_OK;
    return rc ? fail(rc, "exact render launch failed") : BIRP_OK;
}

static int render_fast_into(birp_volume* vol, const birp_view* view, float* dimg);

extern "C" int birp_render(birp_volume* vol, const birp_view* view, float* image, birp_render_stats* stats) {
    if (!vol || !image) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    int rc = validate_view(view, view && view->math == BIRP_MATH_EXACT);
    if (rc) return rc;
    birp_device* dev = vol->dev;
    (void)hipSetDevice(dev->ordinal);
    const size_t npix = (size_t)view->width * view->height;
    float* dimg = nullptr;
    unsigned long long* cnt = nullptr;
    if ((rc = render_buffer(dev, npix, &dimg, &cnt))) return rc;
    // a birp_host_alloc image is read and written by the kernel in place, through its device
    // address; any other image goes through the device's pinned staging, which the kernel also
    // writes in place when the device maps it. With a mapped staging an exact render's waves also
    // store their in-cube sample counts there (after the image's place), and nothing comes back by
    // DMA: the host sums the counts after the stream synchronises.
    const bool exact = view->math == BIRP_MATH_EXACT;
    float* mapped = pinned_device_ptr(image, npix * 4);
    const size_t counts_at = mapped ? 256 : 256 + ((npix * 4 + 255) & ~(size_t)255);
    const int nslots = exact ? exact_wave_slots(view->width, view->height,
                                                default_exact_lanes(vol->dv.kind == SK_CMEM, view->width, view->height)) : 0;
    if ((rc = pinned_staging(dev, counts_at + (size_t)nslots * 4))) return rc;
    float* staged = nullptr;  // the image's place in the pinned staging, when the kernel writes it
    unsigned* counts_dev = nullptr;
    const unsigned* counts_host = nullptr;
    if (mapped) {
        dimg = mapped;
    } else if (dev->pinned_dev) {
        dimg = (float*)((char*)dev->pinned_dev + 256);
        staged = (float*)((char*)dev->pinned + 256);
    }
    if (exact && dev->pinned_dev) {
        counts_dev = (unsigned*)((char*)dev->pinned_dev + counts_at);
        counts_host = (const unsigned*)((const char*)dev->pinned + counts_at);
    }
    hipError_t e;
    birp_view vv = *view;
    if (vv.flags & BIRP_RENDER_ACCUMULATE) {
        // an image of +0.0 bits adds nothing: render as a fresh image instead of uploading it
        const uint32_t* bits = reinterpret_cast<const uint32_t*>(image);
        size_t k = 0;
        while (k < npix && bits[k] == 0u) k++;
        if (k == npix) vv.flags &= ~BIRP_RENDER_ACCUMULATE;
        else if (staged) memcpy(staged, image, npix * 4);
        else if (!mapped && (e = hipMemcpyAsync(dimg, image, npix * 4, hipMemcpyHostToDevice, dev->stream)) != hipSuccess)
            return hip_fail(e, "hipMemcpy(image)");
    }
    view = &vv;
    (void)hipEventRecord(dev->ev0, dev->stream);
    if (exact) rc = render_exact_into(vol, view, dimg, cnt, counts_dev);
    else rc = render_fast_into(vol, view, dimg);
    if (rc) {
        (void)hipStreamSynchronize(dev->stream);
        dev->count_known = false;
        return rc;
    }
    (void)hipEventRecord(dev->ev1, dev->stream);
    bool counter_back = true;  // the device counter's value is at the staging's offset 0
    if (mapped || staged || pinned_range(image, npix * 4)) {
        // the image is in place (in the caller's pinned buffer or in the staging), or goes by one DMA
        // straight into a pinned caller buffer; an exact render's counter, unless its waves stored
        // their counts in the staging, comes back by an 8-byte DMA
        counter_back = exact && !counts_host;
        if ((counter_back && (e = hipMemcpyAsync(dev->pinned, dev->rbuf, 8, hipMemcpyDeviceToHost, dev->stream)) != hipSuccess) ||
            (!mapped && !staged && (e = hipMemcpyAsync(image, dimg, npix * 4, hipMemcpyDeviceToHost, dev->stream)) != hipSuccess) ||
            (e = hipStreamSynchronize(dev->stream)) != hipSuccess) {
            dev->count_known = false;
            return hip_fail(e, "render");
        }
        if (staged) memcpy(image, staged, npix * 4);
    } else {
        // device -> pinned staging -> caller: one DMA of the counter and the image, one host copy
        if ((rc = pinned_staging(dev, 256 + npix * 4))) return rc;
        if ((e = hipMemcpyAsync(dev->pinned, dev->rbuf, 256 + npix * 4, hipMemcpyDeviceToHost, dev->stream)) != hipSuccess ||
            (e = hipStreamSynchronize(dev->stream)) != hipSuccess) {
            dev->count_known = false;
            return hip_fail(e, "render");
        }
        memcpy(image, (char*)dev->pinned + 256, npix * 4);
    }
    unsigned long long added = 0;
    if (counts_host) {
        for (int k = 0; k < nslots; k++) added += counts_host[k];
    } else if (counter_back) {
        const unsigned long long count = *(const unsigned long long*)dev->pinned;
        added = count - dev->count_base;
        dev->count_base = count;
    }
    if (stats) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, dev->ev0, dev->ev1);
        stats->kernel_ms = ms;
        stats->samples_nominal = (long long)npix * view->samples;
        stats->samples_in_cube = (view->math == BIRP_MATH_EXACT) ? (long long)added : -1;
    }
    return BIRP_OK;
}

extern "C" int birp_image_create(birp_device* dev, int width, int height, birp_image** out) {
    if (!dev || !out || width <= 0 || height <= 0) return fail(BIRP_ERR_INVALID_ARGUMENT, "bad argument");
    (void)hipSetDevice(dev->ordinal);
    birp_image* img = new birp_image();
    img->dev = dev;
    img->W = width;
    img->H = height;
    hipError_t e = hipMalloc(&img->d, (size_t)width * height * sizeof(float));
    if (e != hipSuccess) {
        delete img;
        return hip_fail(e, "hipMalloc(image)");
    }
    int rc = launch_zero_f32(img->d, (size_t)width * height, dev->stream);
    if (rc) {
        (void)hipFree(img->d);
        delete img;
        return rc;
    }
    *out = img;
    return BIRP_OK;
}

extern "C" int birp_image_free(birp_image* img) {
    if (!img) return BIRP_OK;
    (void)hipStreamSynchronize(img->dev->stream);
    (void)hipFree(img->d);
    delete img;
    return BIRP_OK;
}

extern "C" int birp_image_download(birp_image* img, float* host) {
    if (!img || !host) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    hipError_t e = hipMemcpyAsync(host, img->d, (size_t)img->W * img->H * 4, hipMemcpyDeviceToHost, img->dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(img->dev->stream);
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "image download");
}

extern "C" int birp_image_upload(birp_image* img, const float* host) {
    if (!img || !host) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    hipError_t e = hipMemcpyAsync(img->d, host, (size_t)img->W * img->H * 4, hipMemcpyHostToDevice, img->dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(img->dev->stream);
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "image upload");
}

extern "C" int birp_render_async(birp_volume* vol, const birp_view* view, birp_image* img) {
    if (!vol || !img) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    int rc = validate_view(view, view && view->math == BIRP_MATH_EXACT);
    if (rc) return rc;
    if (img->W != view->width || img->H != view->height) return fail(BIRP_ERR_INVALID_ARGUMENT, "image size mismatch");
    if (img->dev != vol->dev) return fail(BIRP_ERR_INVALID_ARGUMENT, "image and volume on different devices");
    if (view->math == BIRP_MATH_EXACT) return render_exact_into(vol, view, img->d, nullptr);
    return render_fast_into(vol, view, img->d);
}

extern "C" int birp_count_samples(birp_volume* vol, const birp_view* view, long long* in_cube) {
    if (!vol || !view || !in_cube) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    birp_view v = *view;
    v.math = BIRP_MATH_EXACT;
    v.flags = 0;
    int rc = validate_view(&v, true);
    if (rc) return rc;
    birp_device* dev = vol->dev;
    (void)hipSetDevice(dev->ordinal);
    float* dimg = nullptr;
    unsigned long long* cnt = nullptr;
    if ((rc = render_buffer(dev, (size_t)v.width * v.height, &dimg, &cnt))) return rc;
    if ((rc = render_exact_into(vol, &v, dimg, cnt))) {
        (void)hipStreamSynchronize(dev->stream);
        dev->count_known = false;
        return rc;
    }
    if ((rc = pinned_staging(dev, 256))) return rc;
    hipError_t e = hipMemcpyAsync(dev->pinned, cnt, 8, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess) {
        dev->count_known = false;
        return hip_fail(e, "count");
    }
    const unsigned long long count = *(const unsigned long long*)dev->pinned;
    *in_cube = (long long)(count - dev->count_base);
    dev->count_base = count;
    return BIRP_OK;
}

extern "C" int birp_render_f64(birp_volume* vol, const birp_view* view, double* image) {
    if (!vol || !image) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    int rc = validate_view(view, false);
    if (rc) return rc;
    if (vol->dv.kind == SK_CUBE && !vol->dv.linear) return fail(BIRP_ERR_UNSUPPORTED, "FP64 renders need an FP32 volume");
    birp_device* dev = vol->dev;
    (void)hipSetDevice(dev->ordinal);
    Mat3d rot, rinv;
    if (!view_basis64(view->position, view->aim, rot, rinv)) return fail(BIRP_ERR_SINGULAR, "Determinant is 0");
    F64View fv;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) fv.rinv[r * 3 + c] = rinv.m[r][c];
    const double to_rad = 3.14159265358979323846 / 180.0;
    fv.px = view->position[0];
    fv.py = view->position[1];
    fv.pz = view->position[2];
    fv.tanx = std::tan(to_rad * (view->fov_w / 2.0));
    fv.tany = std::tan(to_rad * (view->fov_h / 2.0));
    fv.aspect = (double)view->width / view->height;
    fv.L = view->ray_length;
    {
        double t0 = fv.L * (0.5 / view->samples), t1 = fv.L * (1.5 / view->samples);
        fv.w = t1 - t0;
    }
    fv.W = view->width;
    fv.H = view->height;
    fv.n = view->samples;
    fv.semantics = view->semantics;
    const size_t npix = (size_t)view->width * view->height;
    void* buf = nullptr;
    if ((rc = device_scratch(dev, npix * 8, &buf))) return rc;
    fv.out = (double*)buf;
    if ((rc = launch_f64(vol->dv, fv, dev->stream))) return fail(rc, "f64 launch failed");
    hipError_t e = hipMemcpyAsync(image, buf, npix * 8, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "f64 render");
}

// ============================================================================ batches
namespace birp {

int batch_setup(birp_batch* b);
int exact_view(const birp_volume* vol, const birp_view* v, float* out_dev, ExactView& ev);

int choose_tuning(const birp_volume* vol, const std::vector<birp_view>& views, birp_tuning& t) {
    // the tuning of the batch's dominant (largest) image class
    size_t best = 0;
    int bi = 0;
    for (size_t k = 0; k < views.size(); k++) {
        size_t w = (size_t)views[k].width * views[k].height * views[k].samples;
        if (w > best) { best = w; bi = (int)k; }
    }
    if (!views.empty() && birp_autotune_lookup(vol->desc.dtype, views[bi].width, views[bi].height, views[bi].samples, &t) == BIRP_OK)
        return BIRP_OK;
    // untuned default: the tuner's most frequent winner on the standard workloads
    t.tile_w = 8;
    t.tile_h = 8;
    t.lanes_per_ray = 1;
    t.unroll = 4;
    return BIRP_OK;
}

// The tuning of a fast render launched alone or in a server batch, chosen for latency from the
// view and the volume's storage alone (so a request's result never depends on what it is batched
// with): an image with few rays splits each ray over 2 or 4 lanes, shortening every lane's chain
// of dependent sample gathers. Measured on back-to-back single launches of 6 orbit poses (256^3
// cube), ms per render for 1 / 2 / 4 lanes per ray:
//   144x84,  n = 160: F32 0.028 / 0.030 / 0.030, F16 0.027 / 0.024 / 0.027
//   144x84,  n = 320: F32 0.036 / 0.039 / 0.033, F16 0.034 / 0.027 / 0.028
//   144x84,  n = 800: F32 0.060 / 0.063 / 0.044, F16 0.054 / 0.037 / 0.032
//   288x168, n = 800: F32 0.059 / 0.067 / 0.073, F16 0.054 / 0.049 / 0.063
//   576x336 and 288x168 at n <= 320: 1 lane, by 20-70%
// (BF16 as F32). A CMEM model costs about 25 transcendentals per sample, whose chains want more
// waves per SIMD than a cube's gathers: its views take 4 lanes per ray over 4x4 rays whatever
// their size (576x336 at n = 800: 0.724 ms, against 0.817 ms with 1 lane and 0.733 ms with 2).
birp_tuning latency_tuning(const birp_view& v, const DevVolume& vol) {
    const bool cmem = vol.kind == SK_CMEM;
    const bool f16 = vol.dtype == BIRP_DTYPE_F16;
    birp_tuning t;
    if (cmem) {
        t.lanes_per_ray = 4;
        t.unroll = 1;
        t.tile_w = 4;
        t.tile_h = 4;
        return t;
    }
    const long long rays = (long long)v.width * v.height;
    int lanes = 1;
    if (rays <= 20000) lanes = v.samples >= 600 ? 4 : (f16 ? 2 : (v.samples >= 240 ? 4 : 1));
    else if (rays <= 60000 && v.samples >= 600 && f16) lanes = 2;
    t.lanes_per_ray = lanes;
    t.unroll = 4;
    t.tile_w = 8;
    t.tile_h = 64 / (8 * lanes);
    return t;
}

void batch_release(birp_batch* b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    if (b->d_poses) (void)hipFree(b->d_poses);
    if (b->d_prefix) (void)hipFree(b->d_prefix);
    if (b->d_vols) (void)hipFree(b->d_vols);
    if (b->d_evs) (void)hipFree(b->d_evs);
    if (b->d_plan) (void)hipFree(b->d_plan);
    b->d_poses = nullptr;
    b->d_prefix = nullptr;
    b->d_vols = nullptr;
    b->d_evs = nullptr;
    b->d_plan = nullptr;
}

// The kernels of an exact batch on stream st: one multi-view launch per sample-mode group.
int exact_batch_launch(birp_batch* b, hipStream_t st) {
    int base = 0;
    for (size_t g = 0; g < b->groups.size(); g++) {
        const birp_batch::ExactGroup& G = b->groups[g];
        const int* plan = G.plan_at >= 0 ? b->d_plan + G.plan_at : nullptr;
        int rc = launch_exact_multi(b->d_vols, b->d_evs + G.first, b->d_prefix + base, G.count, G.total, G.mode, G.lds, st, plan,
                                    G.lanes, G.pipe);
        if (rc) return rc;
        base += G.count + 1;
    }
    return BIRP_OK;
}

// The tile plans of an exact batch, made once when the batch is set up: each cube group's tiles
// are estimated on the device (k_plan_multi) and counting-sorted on the host, heaviest first.
// A group of more tiles than the GPU holds at once sorts each view's tiles in place of their
// natural order and keeps the views' order, so the launch still walks the batch view by view: the
// multi-view launch hands blocks to CUs as earlier ones finish, and a view's light tiles fill the
// gaps its heavy tiles leave. A group whose blocks are all resident from the start (`resident`)
// gets no such dynamic balance, so its tiles are sorted together and dealt in a snake over the
// CUs, as a single render's plan is (k_render_exact): slot b = r ncu + c takes sorted tile
// r ncu + c in even rounds r and r ncu + nr - 1 - c in odd ones (nr slots in the round). CMEM groups (every ray
// takes all n samples), groups of fewer tiles than CUs, and BIRP_EXACT_ORDER=natural keep the
// natural order.
static int exact_batch_plan(birp_batch* b, const std::vector<int>& prefix) {
    birp_device* dev = b->vol->dev;
    if (exact_order_policy() == 0 || dev->ncu <= 0) return BIRP_OK;
    long long need = 0;
    for (const birp_batch::ExactGroup& G : b->groups)
        if (G.mode != SM_CMEM && G.total >= dev->ncu) need += G.total;
    if (!need) return BIRP_OK;
    hipError_t e;
    int* d_cost = nullptr;
    if ((e = hipMalloc(&b->d_plan, (size_t)need * sizeof(int))) != hipSuccess) return hip_fail(e, "hipMalloc(batch plan)");
    if ((e = hipMalloc(&d_cost, (size_t)need * sizeof(int))) != hipSuccess) return hip_fail(e, "hipMalloc(plan estimates)");
    std::vector<int> cost((size_t)need), plan((size_t)need), bucket;
    int rc = BIRP_OK;
    long long at = 0;
    int base = 0;  // the group's first entry of prefix (count + 1 entries per group)
    for (birp_batch::ExactGroup& G : b->groups) {
        const int pbase = base;
        base += G.count + 1;
        if (G.mode == SM_CMEM || G.total < dev->ncu) continue;
        rc = launch_plan_multi(b->d_vols, b->d_evs + G.first, b->d_prefix + pbase, G.count, G.total, G.mode, d_cost + at, dev->stream,
                               G.lanes);
        if (rc) break;
        e = hipMemcpyAsync(cost.data() + at, d_cost + at, (size_t)G.total * sizeof(int), hipMemcpyDeviceToHost, dev->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
        if (e != hipSuccess) { rc = hip_fail(e, "plan estimates"); break; }
        int* c = cost.data() + at;
        for (int g = 0; g < G.total; g++) c[g] = std::min(std::max(c[g], 0), 65535);  // estimates are at most 4n
        if (G.resident) {
            std::vector<int> sorted((size_t)G.total);
            int top = 0;
            for (int g = 0; g < G.total; g++) top = std::max(top, c[g]);
            bucket.assign((size_t)top + 2, 0);
            for (int g = 0; g < G.total; g++) bucket[top - c[g] + 1]++;
            for (size_t k = 1; k < bucket.size(); k++) bucket[k] += bucket[k - 1];
            for (int g = 0; g < G.total; g++) sorted[bucket[top - c[g]]++] = g;
            const int ncu = dev->ncu;
            for (int slot = 0; slot < G.total; slot++) {
                const int r = slot / ncu, r0 = r * ncu, col = slot - r0, nr = std::min(ncu, G.total - r0);
                plan[at + slot] = sorted[r0 + ((r & 1) ? nr - 1 - col : col)];
            }
            G.plan_at = at;
            at += G.total;
            continue;
        }
        for (int q = 0; q < G.count; q++) {  // view q of the group: items prefix[pbase + q] .. prefix[pbase + q + 1]
            const int i0 = prefix[pbase + q], i1 = prefix[pbase + q + 1];
            int top = 0;
            for (int g = i0; g < i1; g++) top = std::max(top, c[g]);
            bucket.assign((size_t)top + 2, 0);
            for (int g = i0; g < i1; g++) bucket[top - c[g] + 1]++;  // heaviest first: key top - c
            for (size_t k = 1; k < bucket.size(); k++) bucket[k] += bucket[k - 1];
            for (int g = i0; g < i1; g++) plan[at + i0 + bucket[top - c[g]]++] = g;
        }
        G.plan_at = at;
        at += G.total;
    }
    if (rc == BIRP_OK) {
        e = hipMemcpyAsync(b->d_plan, plan.data(), (size_t)at * sizeof(int), hipMemcpyHostToDevice, dev->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
        if (e != hipSuccess) rc = hip_fail(e, "batch plan upload");
    }
    (void)hipFree(d_cost);
    if (rc != BIRP_OK)
        for (birp_batch::ExactGroup& G : b->groups) G.plan_at = -1;
    return rc;
}

// Exact batches: per-view ExactView descriptors (the arithmetic of birp_render's exact path), grouped
// by sample mode, the tile prefix sums of each group, the volume table, the tile plans; then the graph.
int exact_batch_setup(birp_batch* b) {
    birp_device* dev = b->vol->dev;
    int rc;
    batch_release(b);
    const int n = (int)b->views.size();
    std::vector<ExactView> evs(n);
    std::vector<int> mode(n);
    for (int k = 0; k < n; k++) {
        const birp_volume* vol = b->evols[b->vol_of[k]];
        float* out = (b->host_out_dev ? b->host_out_dev : b->d_images) + b->offsets[k];
        if ((rc = exact_view(vol, &b->views[k], out, evs[k]))) return rc;
        evs[k].vol = b->vol_of[k];
        mode[k] = exact_mode_of(vol->dv, evs[k]);
    }
    // group by mode; each view's sample-count slots (4 per tile at the group's lanes per ray)
    b->groups.clear();
    b->order.clear();
    b->count_off.assign(n, 0);
    b->count_slots = 0;
    std::vector<int> prefix;
    std::vector<ExactView> packed;
    for (int m = SM_QUAD_POS; m <= SM_CMEM; m++) {
        birp_batch::ExactGroup G;
        G.mode = m;
        G.first = (int)b->order.size();
        G.lanes = exact_multi_lanes(m);
        // A cube group small enough that the GPU holds all its blocks at once (7 blocks of 4 waves
        // per CU, the occupancy of the 4-samples-in-flight kernel) renders like single views:
        // 8 lanes per ray when every image has at most 20,000 rays and the 8-lane tiles fit, else 4
        // when every image has at most 60,000 rays, 4 samples in flight, and one snake deal over
        // the CUs (exact_batch_plan). Larger groups keep 2 and 2 (the orbit batch measured 1-2%
        // slower with 4 in flight and 10% slower with 4 lanes).
        if (m != SM_CMEM && dev->ncu > 0) {
            const long long cap = 7LL * dev->ncu;
            long long t2 = 0, t4 = 0, t8 = 0;
            bool small = true, tiny = true;
            for (int k = 0; k < n; k++) {
                if (mode[k] != m) continue;
                t2 += exact_tiles(b->views[k].width, b->views[k].height, 2);
                t4 += exact_tiles(b->views[k].width, b->views[k].height, 4);
                t8 += exact_tiles(b->views[k].width, b->views[k].height, 8);
                const long long rays = (long long)b->views[k].width * b->views[k].height;
                small = small && rays <= 60000;
                tiny = tiny && rays <= 20000;
            }
            if (tiny && t8 <= cap) G.lanes = 8;
            else if (small && t4 <= cap) G.lanes = 4;
            G.resident = (G.lanes == 8 ? t8 : G.lanes == 4 ? t4 : t2) <= cap;
            if (G.resident && m == SM_QUAD_POS) G.pipe = 4;
        }
        prefix.push_back(0);
        for (int k = 0; k < n; k++) {
            if (mode[k] != m) continue;
            const int lanes = G.lanes;
            const int tiles = exact_tiles(b->views[k].width, b->views[k].height, lanes);
            const long long next = (long long)prefix.back() + tiles;
            if (next > INT32_MAX) return fail(BIRP_ERR_UNSUPPORTED, "exact batch of more than 2^31 - 1 tiles: split it");
            prefix.push_back((int)next);
            b->count_off[k] = b->count_slots;
            b->count_slots += 4 * (size_t)tiles;
            G.lds = std::max(G.lds, exact_lds_bytes(b->evols[b->vol_of[k]]->dv, b->views[k].samples, m));
            b->order.push_back(k);
            packed.push_back(evs[k]);
            G.count++;
        }
        if (!G.count) {
            prefix.pop_back();
            continue;
        }
        G.total = prefix.back();
        if (G.lds > 64 * 1024) return fail(BIRP_ERR_UNSUPPORTED, "exact batch view needs more than 64 KiB of LDS");
        b->groups.push_back(G);
    }
    hipError_t e;
    if (b->d_counts) (void)hipFree(b->d_counts);
    b->d_counts = nullptr;
    if ((e = hipMalloc(&b->d_counts, std::max<size_t>(b->count_slots, 1) * 4)) != hipSuccess) return hip_fail(e, "hipMalloc(counts)");
    for (size_t q = 0; q < packed.size(); q++) packed[q].wave_counts = b->d_counts + b->count_off[b->order[q]];
    std::vector<DevVolume> vt(b->evols.size());
    for (size_t k = 0; k < vt.size(); k++) vt[k] = b->evols[k]->dv;
    if ((e = hipMalloc(&b->d_evs, sizeof(ExactView) * n)) != hipSuccess) return hip_fail(e, "hipMalloc(views)");
    if ((e = hipMalloc(&b->d_prefix, sizeof(int) * prefix.size())) != hipSuccess) return hip_fail(e, "hipMalloc(prefix)");
    if ((e = hipMalloc(&b->d_vols, sizeof(DevVolume) * vt.size())) != hipSuccess) return hip_fail(e, "hipMalloc(vols)");
    e = hipMemcpyAsync(b->d_evs, packed.data(), sizeof(ExactView) * n, hipMemcpyHostToDevice, dev->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->d_prefix, prefix.data(), sizeof(int) * prefix.size(), hipMemcpyHostToDevice, dev->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->d_vols, vt.data(), sizeof(DevVolume) * vt.size(), hipMemcpyHostToDevice, dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess) return hip_fail(e, "exact batch upload");
    if ((rc = exact_batch_plan(b, prefix)) != BIRP_OK) return rc;
    hipStream_t cs = b->capture_stream;
    if ((e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) != hipSuccess) return hip_fail(e, "begin capture");
    rc = exact_batch_launch(b, cs);
    e = hipStreamEndCapture(cs, &b->graph);
    if (rc != BIRP_OK) return fail(rc, "exact batch capture launch failed");
    if (e != hipSuccess) return hip_fail(e, "end capture");
    if ((e = hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0)) != hipSuccess) return hip_fail(e, "graph instantiate");
    return BIRP_OK;
}

int batch_setup(birp_batch* b) {
    if (b->exact) return exact_batch_setup(b);
    birp_device* dev = b->vol->dev;
    batch_release(b);
    const birp_tuning& t = b->tuning;
    const int n = (int)b->views.size();
    std::vector<FastPose> poses(n);
    std::vector<int> prefix(n + 1, 0);
    for (int k = 0; k < n; k++) {
        float* out = (b->host_out_dev ? b->host_out_dev : b->d_images) + b->offsets[k];
        int rc = fast_pose(b->vol, &b->views[k], 0, out, t.tile_w, t.tile_h, poses[k]);
        if (rc) return rc;
        const long long next = (long long)prefix[k] + tiles_of(b->views[k], t.tile_w, t.tile_h);
        if (next > INT32_MAX) return fail(BIRP_ERR_UNSUPPORTED, "batch of more than 2^31 - 1 tiles: split it");
        prefix[k + 1] = (int)next;
    }
    b->total_tiles = prefix[n];
    hipError_t e;
    if ((e = hipMalloc(&b->d_poses, sizeof(FastPose) * n)) != hipSuccess) return hip_fail(e, "hipMalloc(poses)");
    if ((e = hipMalloc(&b->d_prefix, sizeof(int) * (n + 1))) != hipSuccess) return hip_fail(e, "hipMalloc(prefix)");
    if ((e = hipMalloc(&b->d_vols, sizeof(DevVolume))) != hipSuccess) return hip_fail(e, "hipMalloc(vols)");
    e = hipMemcpyAsync(b->d_poses, poses.data(), sizeof(FastPose) * n, hipMemcpyHostToDevice, dev->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->d_prefix, prefix.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, dev->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->d_vols, &b->vol->dv, sizeof(DevVolume), hipMemcpyHostToDevice, dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess) return hip_fail(e, "batch upload");
    b->grid = fast_grid_for(fast_kernel_dtype(b->vol->dv), t.lanes_per_ray, t.unroll, b->total_tiles, volume_tab_entries(b->vol->dv));
    // capture: reset the work counter (a kernel, not a memset node), then the persistent kernel
    hipStream_t cs = b->capture_stream;
    if ((e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) != hipSuccess) return hip_fail(e, "begin capture");
    int rc = launch_reset_counter(b->d_work, cs);
    if (rc == BIRP_OK)
        rc = launch_fast_batch(b->d_vols, fast_kernel_dtype(b->vol->dv), b->d_poses, b->d_prefix, n, b->total_tiles, b->d_work, t.tile_w,
                               t.tile_h, t.lanes_per_ray, t.unroll, b->grid, cs, volume_tab_entries(b->vol->dv));
    e = hipStreamEndCapture(cs, &b->graph);
    if (rc != BIRP_OK) return fail(rc, "batch capture launch failed");
    if (e != hipSuccess) return hip_fail(e, "end capture");
    if ((e = hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0)) != hipSuccess) return hip_fail(e, "graph instantiate");
    return BIRP_OK;
}

}  // namespace birp

extern "C" int birp_batch_create(birp_volume* vol, const birp_view* views, int count, birp_batch** out) {
    if (!vol || !views || count <= 0 || !out) return fail(BIRP_ERR_INVALID_ARGUMENT, "bad argument");
    if (count > (1 << 20)) return fail(BIRP_ERR_INVALID_ARGUMENT, "too many views");
    *out = nullptr;
    for (int k = 0; k < count; k++) {
        int rc = validate_view(&views[k], false);
        if (rc) return rc;
    }
    birp_device* dev = vol->dev;
    (void)hipSetDevice(dev->ordinal);
    birp_batch* b = new birp_batch();
    b->vol = vol;
    b->views.assign(views, views + count);
    size_t off = 0;
    for (int k = 0; k < count; k++) {
        b->views[k].math = BIRP_MATH_FAST;
        b->offsets.push_back(off);
        off += (size_t)views[k].width * views[k].height;
        b->nominal += (long long)views[k].width * views[k].height * views[k].samples;
    }
    b->total_pixels = off;
    hipError_t e;
    int rc;
    if ((e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking)) != hipSuccess) { rc = hip_fail(e, "stream"); goto bad; }
    if ((e = hipMalloc(&b->d_images, off * sizeof(float))) != hipSuccess) { rc = hip_fail(e, "hipMalloc(images)"); goto bad; }
    if ((e = hipMalloc(&b->d_work, 1024)) != hipSuccess) { rc = hip_fail(e, "hipMalloc(counter)"); goto bad; }
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess) { rc = hip_fail(e, "event"); goto bad; }
    choose_tuning(vol, b->views, b->tuning);
    if ((rc = batch_setup(b)) != BIRP_OK) goto bad;
    *out = b;
    return BIRP_OK;
bad:
    birp_batch_free(b);
    return rc;
}

extern "C" int birp_batch_create_exact(birp_volume* const* volumes, int nvolumes, const birp_view* views, const int* volume_of,
                                       int count, birp_batch** out) {
    if (!volumes || nvolumes <= 0 || !views || count <= 0 || !out || (nvolumes > 1 && !volume_of))
        return fail(BIRP_ERR_INVALID_ARGUMENT, "bad argument");
    if (count > (1 << 20) || nvolumes > (1 << 16)) return fail(BIRP_ERR_INVALID_ARGUMENT, "too many views or volumes");
    *out = nullptr;
    for (int k = 0; k < nvolumes; k++) {
        if (!volumes[k]) return fail(BIRP_ERR_INVALID_ARGUMENT, "null volume");
        if (volumes[k]->dev != volumes[0]->dev) return fail(BIRP_ERR_INVALID_ARGUMENT, "volumes on different devices");
        if (volumes[k]->dv.kind == SK_CUBE && !volumes[k]->dv.linear) return fail(BIRP_ERR_UNSUPPORTED, "exact renders need an FP32 volume");
    }
    for (int k = 0; k < count; k++) {
        birp_view v = views[k];
        v.math = BIRP_MATH_EXACT;
        int rc = validate_view(&v, true);
        if (rc) return rc;
        const int vi = volume_of ? volume_of[k] : 0;
        if (vi < 0 || vi >= nvolumes) return fail(BIRP_ERR_INVALID_ARGUMENT, "volume index out of range");
    }
    birp_device* dev = volumes[0]->dev;
    (void)hipSetDevice(dev->ordinal);
    birp_batch* b = new birp_batch();
    b->exact = true;
    b->vol = volumes[0];
    b->evols.assign(volumes, volumes + nvolumes);
    b->views.assign(views, views + count);
    size_t off = 0;
    for (int k = 0; k < count; k++) {
        b->views[k].math = BIRP_MATH_EXACT;
        b->vol_of.push_back(volume_of ? volume_of[k] : 0);
        b->offsets.push_back(off);
        off += (size_t)views[k].width * views[k].height;
        b->nominal += (long long)views[k].width * views[k].height * views[k].samples;
    }
    b->total_pixels = off;
    hipError_t e;
    int rc;
    if ((e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking)) != hipSuccess) { rc = hip_fail(e, "stream"); goto bad; }
    if ((e = hipMalloc(&b->d_images, off * sizeof(float))) != hipSuccess) { rc = hip_fail(e, "hipMalloc(images)"); goto bad; }
    if ((rc = launch_zero_f32(b->d_images, off, dev->stream)) != BIRP_OK) goto bad;  // accumulating views start from 0
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess) { rc = hip_fail(e, "event"); goto bad; }
    if ((rc = batch_setup(b)) != BIRP_OK) goto bad;
    *out = b;
    return BIRP_OK;
bad:
    birp_batch_free(b);
    return rc;
}

extern "C" int birp_batch_samples(birp_batch* b, long long* in_cube) {
    if (!b || !in_cube) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    if (!b->exact) return fail(BIRP_ERR_UNSUPPORTED, "sample counts are kept by exact batches");
    std::vector<unsigned> c(b->count_slots);
    hipError_t e = hipStreamSynchronize(b->vol->dev->stream);
    if (e == hipSuccess && !c.empty()) e = hipMemcpy(c.data(), b->d_counts, c.size() * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "batch sample counts");
    for (size_t q = 0; q < b->order.size(); q++) {
        const int k = b->order[q];
        const size_t first = b->count_off[k];
        const size_t last = (q + 1 < b->order.size()) ? b->count_off[b->order[q + 1]] : b->count_slots;
        long long s = 0;
        for (size_t i = first; i < last; i++) s += c[i];
        in_cube[k] = s;
    }
    return BIRP_OK;
}

extern "C" int birp_batch_free(birp_batch* b) {
    if (!b) return BIRP_OK;
    if (b->capture_stream) (void)hipStreamSynchronize(b->capture_stream);
    (void)hipStreamSynchronize(b->vol->dev->stream);
    batch_release(b);
    if (b->d_counts) (void)hipFree(b->d_counts);
    if (b->d_images) (void)hipFree(b->d_images);
    if (b->d_work) (void)hipFree(b->d_work);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    delete b;
    return BIRP_OK;
}

extern "C" int birp_batch_set_tuning(birp_batch* b, const birp_tuning* t) {
    if (!b) return fail(BIRP_ERR_INVALID_ARGUMENT, "null batch");
    if (b->exact) return fail(BIRP_ERR_UNSUPPORTED, "exact batches have no tuning");
    birp_tuning nt;
    if (t) {
        nt = *t;
        if (!(nt.lanes_per_ray == 1 || nt.lanes_per_ray == 2 || nt.lanes_per_ray == 4) ||
            !(nt.unroll == 1 || nt.unroll == 2 || nt.unroll == 4) || nt.tile_w <= 0 || nt.tile_h <= 0 ||
            nt.tile_w * nt.tile_h * nt.lanes_per_ray != 64 || (nt.tile_w & (nt.tile_w - 1)))
            return fail(BIRP_ERR_INVALID_ARGUMENT, "tuning: tile_w*tile_h*lanes_per_ray must be 64, tile_w a power of two");
    } else {
        choose_tuning(b->vol, b->views, nt);
    }
    (void)hipStreamSynchronize(b->vol->dev->stream);
    b->tuning = nt;
    return batch_setup(b);
}

// The batch's kernels go straight onto the device stream: on ROCm 7 a replay of the captured graph
// took longer per launch (an exact batch of one 144x84 view 0.0231 against 0.0158 ms on the
// device, 64 fast orbit views 0.961 against 0.953 ms), so the graph is kept for callers that ask
// for it (birp_batch_launch_graph).
extern "C" int birp_batch_launch(birp_batch* b) { return birp_batch_launch_eager(b); }

extern "C" int birp_batch_launch_graph(birp_batch* b) {
    if (!b || !b->graph_exec) return fail(BIRP_ERR_INVALID_ARGUMENT, "batch not ready");
    birp_device* dev = b->vol->dev;
    if (take_injected_failure(dev)) return fail(BIRP_ERR_INJECTED, "injected failure");
    (void)hipEventRecord(b->ev0, dev->stream);
    hipError_t e = hipGraphLaunch(b->graph_exec, dev->stream);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    (void)hipEventRecord(b->ev1, dev->stream);
    b->launched = true;
    return BIRP_OK;
}

extern "C" int birp_batch_launch_eager(birp_batch* b) {
    if (!b || (!b->d_poses && !b->d_evs)) return fail(BIRP_ERR_INVALID_ARGUMENT, "batch not ready");
    birp_device* dev = b->vol->dev;
    if (take_injected_failure(dev)) return fail(BIRP_ERR_INJECTED, "injected failure");
    const birp_tuning& t = b->tuning;
    (void)hipEventRecord(b->ev0, dev->stream);
    if (b->exact) {
        int rc = exact_batch_launch(b, dev->stream);
        if (rc) return fail(rc, "exact batch launch failed");
        (void)hipEventRecord(b->ev1, dev->stream);
        b->launched = true;
        return BIRP_OK;
    }
    int rc = launch_reset_counter(b->d_work, dev->stream);
    if (rc == BIRP_OK)
        rc = launch_fast_batch(b->d_vols, fast_kernel_dtype(b->vol->dv), b->d_poses, b->d_prefix, (int)b->views.size(),
                               b->total_tiles, b->d_work, t.tile_w, t.tile_h, t.lanes_per_ray, t.unroll, b->grid, dev->stream,
                               volume_tab_entries(b->vol->dv));
    if (rc) return fail(rc, "batch launch failed");
    (void)hipEventRecord(b->ev1, dev->stream);
    b->launched = true;
    return BIRP_OK;
}

extern "C" int birp_batch_wait(birp_batch* b) {
    if (!b) return fail(BIRP_ERR_INVALID_ARGUMENT, "null batch");
    hipError_t e = hipStreamSynchronize(b->vol->dev->stream);
    if (e != hipSuccess) return hip_fail(e, "batch wait");
    if (b->launched) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) b->last_ms = ms;
        b->launched = false;
    }
    return BIRP_OK;
}

extern "C" int birp_batch_download(birp_batch* b, int index, float* host) {
    if (!b || !host || index < 0 || index >= (int)b->views.size()) return fail(BIRP_ERR_INVALID_ARGUMENT, "bad argument");
    const birp_view& v = b->views[index];
    const size_t bytes = (size_t)v.width * v.height * 4;
    hipError_t e = hipSuccess;
    if (b->host_out) {  // already in host memory once the launch is done
        e = hipStreamSynchronize(b->vol->dev->stream);
        if (e == hipSuccess && host != b->host_out + b->offsets[index]) memcpy(host, b->host_out + b->offsets[index], bytes);
    } else {
        e = hipMemcpyAsync(host, b->d_images + b->offsets[index], bytes, hipMemcpyDeviceToHost, b->vol->dev->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->vol->dev->stream);
    }
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "batch download");
}

extern "C" int birp_batch_set_output(birp_batch* b, float* host) {
    if (!b) return fail(BIRP_ERR_INVALID_ARGUMENT, "null batch");
    float* dev_addr = nullptr;
    if (host) {
        dev_addr = pinned_device_ptr(host, b->total_pixels * 4);
        if (!dev_addr) return fail(BIRP_ERR_INVALID_ARGUMENT, "output must be a device-mapped birp_host_alloc buffer of the batch's pixels");
    }
    (void)hipStreamSynchronize(b->vol->dev->stream);
    b->host_out = host;
    b->host_out_dev = dev_addr;
    return batch_setup(b);
}

extern "C" int birp_batch_stats(birp_batch* b, long long* nominal, double* last_ms) {
    if (!b) return fail(BIRP_ERR_INVALID_ARGUMENT, "null batch");
    if (nominal) *nominal = b->nominal;
    if (last_ms) *last_ms = b->last_ms;
    return BIRP_OK;
}

// birp_render_views_noalloc: birp_render_async of each view without allocating. Exact views launch
// the exact kernel (it needs no scratch); fast views are grouped by the lanes per ray of their
// latency tuning (the tuning birp_render_async uses, so the images are the same bit for bit) and
// each group is one persistent launch whose volume, poses, prefix sums and work counters are
// staged by kernels into the device's fixed scratch array.
extern "C" int birp_render_views_noalloc(birp_volume* vol, const birp_view* views, int count, float* const* images_dev) {
    if (!vol || !views || !images_dev || count <= 0) return fail(BIRP_ERR_INVALID_ARGUMENT, "bad argument");
    if (count > BIRP_NOALLOC_MAX_VIEWS)
        return fail(BIRP_ERR_UNSUPPORTED, "birp_render_views_noalloc takes at most 64 views per call (its fixed scratch holds 64 poses)");
    birp_device* dev = vol->dev;
    if (!dev->noalloc_scratch) return fail(BIRP_ERR_DEVICE, "the device's scratch array was not set up");
    for (int k = 0; k < count; k++) {
        int rc = validate_view(&views[k], views[k].math == BIRP_MATH_EXACT);
        if (rc) return rc;
        if (!images_dev[k]) return fail(BIRP_ERR_INVALID_ARGUMENT, "null image");
        if (views[k].math == BIRP_MATH_FAST && (views[k].flags & BIRP_RENDER_ACCUMULATE))
            return fail(BIRP_ERR_UNSUPPORTED, "fast renders do not accumulate");
    }
    (void)hipSetDevice(dev->ordinal);
    FastPose poses[BIRP_NOALLOC_MAX_VIEWS];
    int prefix[3 * (BIRP_NOALLOC_MAX_VIEWS + 1)] = {};
    int nfast[3] = {0, 0, 0};
    birp_tuning gt[3] = {};
    int nf = 0;
    for (int g = 0; g < 3; g++) {  // lanes per ray 1, 2, 4: poses of group g are contiguous
        for (int k = 0; k < count; k++) {
            if (views[k].math != BIRP_MATH_FAST) continue;
            const birp_tuning t = latency_tuning(views[k], vol->dv);
            if (t.lanes_per_ray != (1 << g)) continue;
            int rc = fast_pose(vol, &views[k], 0, images_dev[k], t.tile_w, t.tile_h, poses[nf]);
            if (rc) return rc;
            int* pre = prefix + g * (BIRP_NOALLOC_MAX_VIEWS + 1);
            const long long next = (long long)pre[nfast[g]] + tiles_of(views[k], t.tile_w, t.tile_h);
            if (next > INT32_MAX) return fail(BIRP_ERR_UNSUPPORTED, "more than 2^31 - 1 tiles");
            pre[nfast[g] + 1] = (int)next;
            nfast[g]++;
            gt[g] = t;
            nf++;
        }
    }
    if (take_injected_failure(dev)) return fail(BIRP_ERR_INJECTED, "injected failure");
    int rc = BIRP_OK;
    for (int k = 0; k < count && rc == BIRP_OK; k++)
        if (views[k].math == BIRP_MATH_EXACT) rc = render_exact_into(vol, &views[k], images_dev[k], nullptr);
    if (rc || nf == 0) return rc;
    char* s = static_cast<char*>(dev->noalloc_scratch);
    if ((rc = noalloc_stage_head(vol->dv, prefix, s, dev->stream))) return fail(rc, "scratch staging failed");
    for (int q = 0; q < nf; q += 4)
        if ((rc = noalloc_stage_poses(poses + q, q, std::min(4, nf - q), s, dev->stream))) return fail(rc, "scratch staging failed");
    const int kdt = fast_kernel_dtype(vol->dv);
    int first = 0;
    for (int g = 0; g < 3; g++) {
        if (!nfast[g]) continue;
        const int* pre_dev = reinterpret_cast<const int*>(s + BIRP_NOALLOC_PREFIX_OFF) + g * (BIRP_NOALLOC_MAX_VIEWS + 1);
        const int total = prefix[g * (BIRP_NOALLOC_MAX_VIEWS + 1) + nfast[g]];
        const birp_tuning& t = gt[g];
        const int grid = fast_grid_for(kdt, t.lanes_per_ray, t.unroll, total, volume_tab_entries(vol->dv));
        rc = launch_fast_batch(reinterpret_cast<const DevVolume*>(s), kdt, reinterpret_cast<const FastPose*>(s + BIRP_NOALLOC_POSE_OFF) + first,
                               pre_dev, nfast[g], total, reinterpret_cast<unsigned*>(s + BIRP_NOALLOC_CTR_OFF) + g * 256, t.tile_w, t.tile_h,
                               t.lanes_per_ray, t.unroll, grid, dev->stream, volume_tab_entries(vol->dv));
        if (rc) return fail(rc, "fast launch failed");
        first += nfast[g];
    }
    return BIRP_OK;
}

// A single fast render: a one-pose launch of the persistent kernel (no graph). Its descriptors go
// through the device's ring of pinned slots, so the host waits only for the upload issued
// kFastRing renders earlier, not for the work queued before this render.
// A fast single render: its pose and volume go to the one-shot kernel by value, so the launch
// needs nothing uploaded and no counters reset.
static int render_fast_into(birp_volume* vol, const birp_view* view, float* dimg) {
    birp_device* dev = vol->dev;
    if (view->flags & BIRP_RENDER_ACCUMULATE) return fail(BIRP_ERR_UNSUPPORTED, "fast renders do not accumulate");
    const birp_tuning t = latency_tuning(*view, vol->dv);
    FastPose fp;
    int rc = fast_pose(vol, view, 0, dimg, t.tile_w, t.tile_h, fp);
    if (rc) return rc;
    if (take_injected_failure(dev)) return fail(BIRP_ERR_INJECTED, "injected failure");
    rc = launch_fast_one(vol->dv, fast_kernel_dtype(vol->dv), fp, tiles_of(*view, t.tile_w, t.tile_h), t.tile_w, t.tile_h,
                         t.lanes_per_ray, t.unroll, dev->stream, volume_tab_entries(vol->dv));
    return rc ? fail(rc, "fast render launch failed") : BIRP_OK;
}
