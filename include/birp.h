/*
 * birp.h -- C ABI of the MI355X BIRP renderer (libbirp_hip.so).
 *
 * The reference exposes no C ABI: its boundary is the C++ class API (include/Space.h,
 * include/DataCube.h, include/Camera.h, include/Cmem.h in the reference). The C++ API in this
 * repository (include/Space.h ... include/Camera.h, built into libbirp_api.so) keeps those
 * signatures and reaches the GPU only through the functions below. Each entry point names the
 * reference interface it replaces.
 *
 * Conventions: every function returns an int status (BIRP_OK == 0), never throws, and borrows
 * host pointers for the duration of the call. Images are H x W float32, row-major, row 0 = the
 * top of the image (reference Camera.cpp:151). Cube values are [z][y][x] with x fastest
 * (reference DataCube.cpp:99-122). One birp_device is used by one host thread at a time.
 */
#ifndef BIRP_H
#define BIRP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status codes */
#define BIRP_OK 0
#define BIRP_ERR_INVALID_ARGUMENT 1 /* reference: std::invalid_argument (DataCube.cpp:136,179) */
#define BIRP_ERR_SINGULAR 2         /* reference: std::runtime_error("Determinant is 0") (Helper.cpp:243) */
#define BIRP_ERR_OUT_OF_MEMORY 3
#define BIRP_ERR_DEVICE 4           /* a HIP runtime call failed */
#define BIRP_ERR_NO_DEVICE 5        /* no GPU, or the library was built without one visible */
#define BIRP_ERR_ADMISSION 6        /* request larger than the serving memory budget */
#define BIRP_ERR_NOT_FOUND 7
#define BIRP_ERR_TIMEOUT 8
#define BIRP_ERR_INJECTED 9         /* failure forced by birp_debug_inject_failure */
#define BIRP_ERR_UNSUPPORTED 10
#define BIRP_ERR_CANCELLED 11

/* ------------------------------------------------------------------ enums */
#define BIRP_DTYPE_F32 0
#define BIRP_DTYPE_BF16 1
#define BIRP_DTYPE_F16 2 /* stored with a power-of-two scale per 8^3 brick */

#define BIRP_INTERP_NEAREST 0   /* reference DataCube.cpp:273 (truncation) */
#define BIRP_INTERP_TRILINEAR 1 /* reference DataCube.cpp:221-271 */

/* Ray semantics. REFERENCE stops a ray at its first out-of-bounds sample, exactly as
 * Camera.cpp:207-210 does with DataCube's -1 sentinel (so a camera outside the cube renders
 * zeros). EXTENDED lets out-of-bounds samples add nothing and keeps marching. Both stop a ray at
 * a negative scaled sample. For a camera whose first sample is inside the cube they agree. */
#define BIRP_SEM_REFERENCE 0
#define BIRP_SEM_EXTENDED 1

/* Arithmetic of a render.
 * EXACT: FP32 with the reference's operation order (no FMA contraction, correctly rounded
 *        division and square root, std::lerp's branches, the double scale chain); bit-identical
 *        to the reference CPU render for finite cubes.
 * FAST:  FP32 with FMA, FP64 camera basis / ray origin / cube entry and FP32 Kahan (compensated)
 *        accumulation; the only form for BF16/F16 storage. For CMEM models: the model in FP32
 *        with the hardware exp2/log2 and polynomial arc cosines (within the reference FP32
 *        model's own error of the FP64 model). */
#define BIRP_MATH_EXACT 0
#define BIRP_MATH_FAST 1

/* birp_view.flags */
#define BIRP_RENDER_ACCUMULATE 1 /* add into the output's current values, as repeated
                                    Camera::Render calls do (Camera.cpp:223 never clears) */

/* ------------------------------------------------------------------ handles */
typedef struct birp_device birp_device;
typedef struct birp_volume birp_volume;
typedef struct birp_image birp_image;
typedef struct birp_batch birp_batch;
typedef struct birp_server birp_server;

/* Grid description (reference DataCube public members, DataCube.h:30-41). spacing is the
 * per-axis |coords[0]-coords[1]| (DataCube.cpp:150-183); origin is coords[0] (:185-191). */
typedef struct birp_volume_desc {
    int nx, ny, nz;
    float origin[3];
    float spacing[3];
    int interp; /* BIRP_INTERP_* (reference DataCube::SetTrilinear, DataCube.cpp:199-201) */
    int dtype;  /* BIRP_DTYPE_*: storage precision on the device */
} birp_volume_desc;

/* CMEM post-Init state (reference include/Cmem.h:43-76): the values CMEM::Init leaves in the
 * object, including B as overwritten by CalcMagneticPressure (Cmem.cpp:310). */
typedef struct birp_cmem_state {
    float p0, p1, p2, p3;
    float bs, A1, A2, B, alpha, beta, ay_bs, az_bs;
    float a5, a21;
    float beta_c[4];
    float charlie, dn, ds, theta_n, theta_s, r0_lin;
} birp_cmem_state;

/* CMEM::Init's positional arguments (reference include/Cmem.h:21-39). has_v / has_b = 0 stands for
 * the reference's empty vector (defaults (400,0,0) and (0,0,5)); a 0 float means "default". */
typedef struct birp_cmem_params {
    int has_v, v[3];
    int has_b, b[3];
    float dipole, p0;
    int p1, p2, p3;
    float B, alpha, beta, bs, A1, A2, ay_bs, az_bs, density;
} birp_cmem_params;

/* One camera pose and ray grid (reference Camera members, Camera.h:23-53). */
typedef struct birp_view {
    int width, height;   /* W x H pixels; the reference derives them from the FOV (Camera.cpp:38-39) */
    float fov_w, fov_h;  /* degrees */
    float position[3];   /* camera position, RE (Camera::SetPosition) */
    float aim[3];        /* aim point (Camera::SetAim); only aim[0] enters the rotation (Camera.cpp:92-98) */
    float ray_length;    /* L: Camera::Render passes 44 (Camera.cpp:65) */
    int samples;         /* n >= 2: the reference uses 200 (Camera.cpp:31) */
    int semantics;       /* BIRP_SEM_* */
    int math;            /* BIRP_MATH_* */
    int flags;           /* BIRP_RENDER_* */
} birp_view;

typedef struct birp_render_stats {
    long long samples_in_cube; /* exact renders: samples added into the image (cube), or n per ray (CMEM);
                                  fast renders: -1 (not counted) */
    long long samples_nominal; /* W * H * n */
    double kernel_ms;          /* device time of the render kernel(s), from HIP events */
} birp_render_stats;

/* ------------------------------------------------------------------ library / device */
const char* birp_strerror(int status);
/* Most recent error detail for the calling thread (a HIP message, the failing check). */
const char* birp_last_error(void);
int birp_version(void);

int birp_device_open(int ordinal, birp_device** out);
int birp_device_close(birp_device* dev);
int birp_device_synchronize(birp_device* dev);
/* Device-time bracket on the device's render stream: begin records an event, end records a
 * second one, waits for it and returns the elapsed milliseconds between them. */
int birp_timer_begin(birp_device* dev);
int birp_timer_end(birp_device* dev, double* elapsed_ms);

/* ------------------------------------------------------------------ volumes (Space) */
/* Replaces DataCube::Load + the DataCube object as a Space (DataCube.h:22-50): copies the
 * values to the GPU and lays them out for the render kernels. values holds nx*ny*nz floats
 * [z][y][x]. The values are checked once for finiteness and sign to pick the kernel variant. */
int birp_volume_upload(birp_device* dev, const birp_volume_desc* desc, const float* values,
                       birp_volume** out);
/* Replaces CMEM::Init (Cmem.cpp:35-115, 250-325): the post-Init model state, bit-identical to the
 * reference's members. Host-only. birp_cmem_default_params fills the arguments CMEM::Init()
 * passes (Cmem.cpp:18-34, the defaults of Cmem.h:21-39). */
void birp_cmem_default_params(birp_cmem_params* params);
int birp_cmem_init(const birp_cmem_params* params, birp_cmem_state* state);
/* Replaces a CMEM object as a Space (Cmem.h:19-97). */
int birp_cmem_create(birp_device* dev, const birp_cmem_state* state, birp_volume** out);
int birp_volume_free(birp_volume* vol);
/* Device bytes held by the volume (all layouts it keeps). */
int birp_volume_bytes(const birp_volume* vol, size_t* bytes);
/* Space::GetSample (Space.h:22; DataCube.cpp:203-277 / Cmem.cpp:117-186) at n points xyz
 * (host, 3n floats), evaluated on the GPU with the exact arithmetic; out-of-bounds cube points
 * give -1 as in the reference. */
int birp_volume_sample(birp_volume* vol, const float* xyz, long long n, float* out);

/* Any other Space (a subclass whose virtual GetSample is the caller's host code, such as the
 * reference tests' CubeFake, tests/Camera_test.cpp:25-29): replaces the per-pixel loop of
 * Camera::Render around the virtual call (Camera.cpp:140-229, the call at :191) by two device
 * steps around the caller's own sampling. birp_space_points writes the FP32 sample points
 * X = d t_k + P of image rows [row0, row1), exactly as the exact renders form them, into xyz:
 * 3 floats per (pixel, k), k fastest, pixels row-major from (0, row0). The caller evaluates its
 * GetSample there (the reference stops calling for a ray once a value v has v * w < 0, with w the
 * view's ray width) and birp_space_integrate adds the values (one float per (pixel, k), same order;
 * values after a ray's stop are not read) into image rows [row0, row1) of `image` ((row1 - row0) *
 * W floats; read first when the view has BIRP_RENDER_ACCUMULATE): the scale chain of
 * Camera.cpp:201, the stop at the first negative sample (:207-210) and the FP32 sum in k order
 * (:223), bit-identical to the reference. The view's semantics has no effect (a generic Space has
 * no bounds); at most 2^28 samples per call. Synchronous. */
int birp_space_points(birp_device* dev, const birp_view* view, int row0, int row1, float* xyz);
int birp_space_integrate(birp_device* dev, const birp_view* view, int row0, int row1, const float* values, float* image);

/* ------------------------------------------------------------------ render (Camera) */
/* Replaces Camera::Render (Camera.cpp:63-233): ray generation, march, integration. Host image
 * (H*W floats): read first when BIRP_RENDER_ACCUMULATE is set, always written. Synchronous. An
 * image inside a birp_host_alloc allocation is read and written by the kernel in place. */
int birp_render(birp_volume* vol, const birp_view* view, float* image, birp_render_stats* stats);

/* Device-resident images for pipelines that keep results on the GPU. */
int birp_image_create(birp_device* dev, int width, int height, birp_image** out);
int birp_image_free(birp_image* img);
int birp_image_download(birp_image* img, float* host); /* waits for the device's work */
int birp_image_upload(birp_image* img, const float* host);
/* Enqueue one render into img on the device's stream; no host synchronisation. */
int birp_render_async(birp_volume* vol, const birp_view* view, birp_image* img);
/* No reference counterpart. For callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): birp_render_async of each of `count` views of
 * one volume into images_dev[k] (device memory the caller owns, views[k].height * width floats),
 * the same images bit for bit, with no allocation of any kind. Exact views run the exact kernel;
 * fast views are grouped by lanes per ray and each group is one persistent launch whose volume,
 * poses, tile prefix sums and work-queue counters live in a fixed device array of the library,
 * BIRP_NOALLOC_SCRATCH_KIB per GPU, set up by birp_device_open and held until the process ends. At
 * most BIRP_NOALLOC_MAX_VIEWS views per call (BIRP_ERR_UNSUPPORTED beyond: the array holds that
 * many poses). Asynchronous on
 * the device's stream, like birp_render_async. Every call on a GPU shares that GPU's array, which
 * keeps the last call's descriptors (the counters are cleared by each call): calls on one device
 * from different threads or streams that overlap race on it. */
#define BIRP_NOALLOC_SCRATCH_KIB 21
#define BIRP_NOALLOC_MAX_VIEWS 64
int birp_render_views_noalloc(birp_volume* vol, const birp_view* views, int count, float* const* images_dev);
/* The exact in-cube sample count of a view (for throughput accounting; runs a counting kernel). */
int birp_count_samples(birp_volume* vol, const birp_view* view, long long* in_cube);

/* The FP64 render of the same sample grid (every operation in double; the FP64 oracle's
 * arithmetic evaluated on the GPU) into a host double image. */
int birp_render_f64(birp_volume* vol, const birp_view* view, double* image);

/* Host-side per-pose math the renders use (Camera.cpp:85-137, 309-508), exported for tests:
 * angle_sun (degrees), R (rotation) and R^-1, all FP32 in the reference's operation order. */
int birp_view_basis(const birp_view* view, float* angle_sun, float rot[9], float rinv[9]);

/* ------------------------------------------------------------------ batched renders */
/* Many poses (and image sizes, sample counts) over one volume in one persistent-kernel launch,
 * also captured in a HIP graph that can be replayed. Images are written into one device buffer; image i
 * starts at offset[i] floats. No reference counterpart: the reference renders one pose per
 * Camera (Camera.cpp:63) and batches cubes serially (main.cpp:242-264). */
typedef struct birp_tuning {
    int tile_w, tile_h; /* rays per wave: tile_w * tile_h == 64 */
    int lanes_per_ray;  /* 1, 2 or 4 lanes share one ray's samples */
    int unroll;         /* samples per lane per loop trip: 1, 2 or 4 */
} birp_tuning;

int birp_batch_create(birp_volume* vol, const birp_view* views, int count, birp_batch** out);
int birp_batch_free(birp_batch* b);
/* Selects kernel tuning for the batch: an explicit choice, or NULL for the autotuner's. */
int birp_batch_set_tuning(birp_batch* b, const birp_tuning* tuning);
/* Launch without waiting (birp_batch_wait waits for completion): the batch's kernels go straight
 * onto the device stream, which measured faster per launch than replaying the captured graph. */
int birp_batch_launch(birp_batch* b);
/* The same launch; kept as the name of the direct launch beside the graph replay below. */
int birp_batch_launch_eager(birp_batch* b);
/* The same kernels as one replay of the HIP graph captured when the batch was made (identical
 * images; for callers that want a graph launch). */
int birp_batch_launch_graph(birp_batch* b);
int birp_batch_wait(birp_batch* b);
/* Copy image i (views[i].height * views[i].width floats) to host memory. */
int birp_batch_download(birp_batch* b, int index, float* host);
/* Directs the batch's images into `host`, a birp_host_alloc buffer of at least the batch's total
 * pixels (image i at the sum of the earlier images' W*H floats): the kernel writes them there
 * itself, so a finished launch has every image in host memory. NULL returns the images to the
 * device buffer. Re-captures the launch. */
int birp_batch_set_output(birp_batch* b, float* host);
int birp_batch_stats(birp_batch* b, long long* samples_nominal, double* last_ms);

/* Exact batches: replaces the reference's serial loop of Camera::Render over the cubes of a batch
 * directory (main.cpp:242-264; Camera.cpp:63-233 per cube) with one launch per sample mode. View
 * k renders volume volumes[volume_of[k]] (volume_of may be NULL when nvolumes == 1) with the
 * BIRP_MATH_EXACT arithmetic of birp_render, so every image is bit-identical to birp_render of
 * that view and volume; volumes may be FP32 cubes (any interpolation and sign) and CMEM models, on
 * one device, and views may differ in size, pose, ray and semantics. A view with
 * BIRP_RENDER_ACCUMULATE adds into its image's current contents (zero at creation; the previous
 * launch's result after that). The tiles of all views of one sample mode share one launch, so SIMDs
 * stay fed while a view's longest rays finish. Launch, wait, download, set_output, stats and free
 * work as for fast batches; birp_batch_set_tuning returns BIRP_ERR_UNSUPPORTED. */
int birp_batch_create_exact(birp_volume* const* volumes, int nvolumes, const birp_view* views, const int* volume_of,
                            int count, birp_batch** out);
/* Per-view in-cube sample counts of an exact batch's last finished launch (count entries, as
 * birp_render_stats.samples_in_cube); BIRP_ERR_UNSUPPORTED for fast batches. */
int birp_batch_samples(birp_batch* b, long long* in_cube);

/* Autotuner: times every candidate birp_tuning on the given views (few launches each), keeps
 * the fastest per (dtype, image-size class, samples class) in an in-process table and, when
 * path != NULL, writes the table as JSON there. birp_autotune_load reads such a file. */
int birp_autotune(birp_volume* vol, const birp_view* views, int count, const char* path);
int birp_autotune_load(const char* path);
int birp_autotune_lookup(int dtype, int width, int height, int samples, birp_tuning* out);

/* ------------------------------------------------------------------ serving */
typedef struct birp_server_config {
    size_t memory_budget;  /* bytes for the resident volumes (device memory) plus the image
                              buffers the slots hold (4 bytes per pixel, kept between batches:
                              pinned host staging that the kernels write through the device
                              mapping, plus device images where the device does not map it; a
                              slot grows its buffer within the budget, releasing idle slots'
                              buffers when a request needs the room). */
    int max_batch;         /* most requests one launch may carry */
    int max_inflight;      /* launches in flight at once (each on its own stream) */
    double batch_window_ms;/* wait this long for more arrivals before launching a non-full batch */
    size_t reserve_pixels; /* image pixels each slot allocates at creation (device + pinned
                              staging), so that batches up to that size allocate nothing while
                              serving; max_inflight * reserve_pixels * 4 must fit the budget
                              (else BIRP_ERR_ADMISSION); 0: slots grow on demand */
} birp_server_config;

typedef struct birp_request {
    int volume;     /* index returned by birp_server_add_volume */
    birp_view view; /* rendered in the view's math: BIRP_MATH_FAST, or else BIRP_MATH_EXACT
                       (FP32 cubes and CMEM models; BIRP_ERR_UNSUPPORTED for 16-bit cubes),
                       bit-identical to birp_render of that view in that math; flags are ignored */
} birp_request;

typedef struct birp_result_info {
    int status;          /* BIRP_OK or the error of this request */
    double submit_s;     /* host clock at submission (seconds, steady clock) */
    double start_s;      /* launch time of the batch that carried it */
    double done_s;       /* completion time observed by the server */
    int batch_size;      /* requests in that launch */
} birp_result_info;

int birp_server_create(birp_device* dev, const birp_server_config* cfg, birp_server** out);
int birp_server_destroy(birp_server* s);
/* Registers a resident volume, a cube or a CMEM model (owned by the caller; must outlive the
 * server). Its device bytes
 * count against the budget; fails with BIRP_ERR_ADMISSION when it does not fit beside the
 * volumes already registered and the image buffers the slots hold. A request whose image alone
 * exceeds the budget left after the volumes completes with BIRP_ERR_ADMISSION. */
int birp_server_add_volume(birp_server* s, birp_volume* vol, int* index);
/* Queues a request; the result buffer (H*W floats, host) must stay valid until it completes. */
int birp_server_submit(birp_server* s, const birp_request* req, float* result, uint64_t* id);
/* Waits up to timeout_ms (< 0: forever) for request id; fills info and forgets the request.
 * One thread at a time may wait for a given id (a second concurrent waiter gets
 * BIRP_ERR_INVALID_ARGUMENT); after BIRP_ERR_TIMEOUT the id can be waited for again. */
int birp_server_wait(birp_server* s, uint64_t id, double timeout_ms, birp_result_info* info);
/* Steady-clock seconds as used in birp_result_info. */
double birp_clock_s(void);
/* Pinned (page-locked) host memory for response buffers, mapped into the device's address space
 * (no reference counterpart: the reference keeps its image in a std::vector, Camera.h:53).
 * A served request whose result buffer lies inside one of these allocations has its image
 * written there by the render kernel itself (no device image, DMA or host copy), and
 * birp_render reads and writes such an image in place the same way. Free only after every
 * request writing into it has been waited for. */
int birp_host_alloc(size_t bytes, void** out);
int birp_host_free(void* p);

/* ------------------------------------------------------------------ synthetic workloads */
/* Seeded magnetosheath-shell emissivity cube on the uniform grid x in [x0,x1], y,z in [y0,y1]:
 * r_mp = r0 (2/(1+cos th))^0.6 with r0 ~ U(r0_lo, r0_hi), r_bs = 1.3 r_mp, shell value
 * amp (r/10)^-3, background bg elsewhere, times exp(sigma N(0,1)) per voxel (SplitMix64 +
 * Box-Muller, one normal per voxel in [z][y][x] order after the r0 draw). Host, threaded. */
typedef struct birp_shell_params {
    int n;            /* voxels per axis */
    float x0, x1, y0, y1;
    float r0_lo, r0_hi;
    float amp, bg, sigma;
    uint64_t seed;
} birp_shell_params;
int birp_synth_shell(const birp_shell_params* p, float* values, float* origin3, float* spacing3,
                     float* r0_out);
/* Poses on a Kepler ellipse (perigee, apogee in RE; inclination deg; argument of perigee -90 deg),
 * true anomaly ~ U(0, 2 pi) from seed: 3 floats per pose. */
int birp_synth_orbit(int count, float perigee, float apogee, float incl_deg, uint64_t seed,
                     float* positions);

/* ------------------------------------------------------------------ test hooks */
/* Makes the next `count` render launches of this device fail with BIRP_ERR_INJECTED before
 * they reach the GPU (host-side failure injection for the serving tests). */
int birp_debug_inject_failure(birp_device* dev, int count);
/* Tile order of single exact renders (birp_render, birp_render_async, birp_count_samples):
 * 0 renders in the natural order (block b renders tile b; exact batches are then made without
 * plans); 1, the default, renders a view's first time in the natural order and, when the view
 * comes back, deals the tiles heaviest first over the CUs from a plan that is kept (four per
 * device) and reused while the view's geometry, the cube's box and the tile shape stay the same;
 * 2 plans on every render. -1 returns to the default or to what BIRP_EXACT_ORDER ("natural",
 * "replan") asks. The order never changes an image. `previous` (may be null) receives the policy
 * in force before the call. Process-wide. */
int birp_debug_exact_order(int policy, int* previous);
/* Plans the tile order of `view` over a cube on the device and copies it into order[0..n),
 * n = *nblocks (the view's tile count, also set when the call fails for capacity < n). */
int birp_debug_exact_plan(birp_volume* vol, const birp_view* view, int* order, int capacity, int* nblocks);

#ifdef __cplusplus
}
#endif
#endif /* BIRP_H */
