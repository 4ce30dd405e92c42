// server.cpp -- continuous-batching render server (include/birp.h "serving").
//
// One dispatcher thread owns `max_inflight` launch slots, each with its own HIP stream (default
// priority), descriptor buffers and an image buffer: pinned staging that the kernels write through
// its device address (plus device images where the device does not map it). The memory budget
// bounds the resident volumes plus the image buffers every slot holds (they are kept between
// batches). Requests queue on submission; whenever a slot is free the
// dispatcher admits queued requests in arrival order while the slot's image buffer can hold them,
// or can grow to hold them within the budget (releasing idle slots' buffers when that makes the
// room), and launches them as one persistent-kernel batch (one launch per storage type and lanes
// per ray). A request that arrives while other batches are on the GPU is launched in the next
// free slot, so its work overlaps theirs on the device. A result buffer from birp_host_alloc is
// written by the kernel itself through its device address, the other images into the slot's
// pinned staging the same way, and when a slot's device work is done a completion thread copies
// those into the callers' buffers (split over a pool of helper threads for large batches) and
// releases the slot, so the dispatcher keeps launching meanwhile.
// Every request in BIRP_MATH_FAST is rendered with the tuning its own view selects (the one fast
// single renders use) and the same per-tile arithmetic whatever it is batched with, so results do
// not depend on the mix and equal birp_render's. Requests in BIRP_MATH_EXACT (FP32 cubes of any
// interpolation and sign, CMEM models) render with the exact arithmetic through the multi-view
// exact kernel, one launch per sample mode in the same slot, and equal birp_render's exact images
// bit for bit.
// No reference counterpart: the reference renders one pose per process run (main.cpp:266-316).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/birp.h"
#include "capi_internal.h"

using namespace birp;

extern "C" double birp_clock_s(void) {
    using namespace std::chrono;
    return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

namespace {
// Pinned response buffers handed out by birp_host_alloc: base address -> size.
std::mutex g_pinned_mu;
std::map<uintptr_t, std::pair<size_t, char*>> g_pinned;  // base -> (size, device address of base)

}  // namespace

// The device address of [p, p + bytes) when it lies inside one birp_host_alloc allocation that
// the device maps, else null.
float* birp::pinned_device_ptr(const void* p, size_t bytes) {
    const uintptr_t a = (uintptr_t)p;
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    auto it = g_pinned.upper_bound(a);
    if (it == g_pinned.begin()) return nullptr;
    --it;
    const size_t size = it->second.first;
    if (!it->second.second || a - it->first > size || bytes > size - (a - it->first)) return nullptr;
    return (float*)(it->second.second + (a - it->first));
}

bool birp::pinned_range(const void* p, size_t bytes) {
    const uintptr_t a = (uintptr_t)p;
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    auto it = g_pinned.upper_bound(a);
    if (it == g_pinned.begin()) return false;
    --it;
    return a >= it->first && a - it->first <= it->second.first && bytes <= it->second.first - (a - it->first);
}

extern "C" int birp_host_alloc(size_t bytes, void** out) {
    if (!out || bytes == 0 || bytes > ((size_t)1 << 40)) return fail(BIRP_ERR_INVALID_ARGUMENT, "bytes in [1, 2^40] and out required");
    *out = nullptr;
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(response buffer)");
    // the device-side address of the allocation, for kernels that write results into it
    void* d = nullptr;
    if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
        (void)hipGetLastError();
        d = nullptr;  // not mapped: requests into it take the staging path
    }
    {
        std::lock_guard<std::mutex> lk(g_pinned_mu);
        g_pinned[(uintptr_t)p] = {bytes, (char*)d};
    }
    *out = p;
    return BIRP_OK;
}

extern "C" int birp_host_free(void* p) {
    if (!p) return BIRP_OK;
    {
        std::lock_guard<std::mutex> lk(g_pinned_mu);
        auto it = g_pinned.find((uintptr_t)p);
        if (it == g_pinned.end()) return fail(BIRP_ERR_NOT_FOUND, "not a birp_host_alloc pointer");
        g_pinned.erase(it);
    }
    hipError_t e = hipHostFree(p);
    return e == hipSuccess ? BIRP_OK : hip_fail(e, "hipHostFree(response buffer)");
}

namespace {

struct Request {
    uint64_t id = 0;
    birp_request req{};
    birp_volume* vol = nullptr;  // captured at submission
    float* result = nullptr;
    size_t pixels = 0;
    float* dev_out = nullptr;  // result lies in a birp_host_alloc buffer: its device address
    double submit_s = 0, start_s = 0, done_s = 0;
    int batch_size = 0;
    int status = BIRP_OK;
    bool finished = false;
    bool waited = false;  // a thread is in birp_server_wait for it
};

// Fork-join pool for the completion thread's copy-out: helper threads wait for a job, copy one
// part each, and the caller copies part 0 and waits for the rest (no thread start per batch).
class CopyPool {
public:
    void start(int helpers) {
        for (int k = 0; k < helpers; k++) th_.emplace_back([this] { loop(); });
    }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        cv_job_.notify_all();
        for (auto& t : th_) t.join();
        th_.clear();
    }
    int helpers() const { return (int)th_.size(); }
    // fn(part) for every part in [0, parts); part 0 runs on the calling thread
    void run(int parts, const std::function<void(int)>& fn) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &fn;
            parts_ = parts;
            next_ = 1;
            pending_ = parts - 1;
        }
        cv_job_.notify_all();
        fn(0);
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
    }

private:
    void loop() {
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            cv_job_.wait(lk, [&] { return quit_ || (job_ && next_ < parts_); });
            if (quit_) return;
            const int part = next_++;
            const std::function<void(int)>* fn = job_;
            lk.unlock();
            (*fn)(part);
            lk.lock();
            if (--pending_ == 0) cv_done_.notify_one();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_job_, cv_done_;
    std::vector<std::thread> th_;
    const std::function<void(int)>* job_ = nullptr;
    int parts_ = 0, next_ = 0, pending_ = 0;
    bool quit_ = false;
};

// A launch's descriptors live in one block, pinned on the host and mirrored on the device, and go
// to the device in one copy (desc_layout): the work-queue counters first, kept zero on the host so
// that the copy resets them, then the tile prefix sums, the fast poses and the exact views, each
// region sized for the launch's requests.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    unsigned char* h_desc = nullptr;  // pinned descriptor block, sized for max_batch requests
    unsigned char* d_desc = nullptr;  // its device copy
    unsigned* d_work = nullptr;       // work-queue counters, 256 per kernel group, at d_desc
    // this launch's regions in h_desc / d_desc
    FastPose* h_poses = nullptr;  // poses packed by kernel group
    FastPose* d_poses = nullptr;
    int* h_prefix = nullptr;      // each kernel group's tile prefix sums
    int* d_prefix = nullptr;
    ExactView* h_evs = nullptr;   // the exact requests' view constants, grouped by sample mode
    ExactView* d_evs = nullptr;
    int* h_eprefix = nullptr;     // each mode group's tile prefix sums
    int* d_eprefix = nullptr;
    float* d_images = nullptr;
    float* h_images = nullptr;    // pinned
    float* h_images_dev = nullptr;  // its device address: kernels write the staged images there
    size_t image_cap = 0;         // pixels allocated (dispatcher thread)
    size_t held = 0;              // device bytes charged to the budget for the images (under mu)
    std::vector<Request*> reqs;
    std::vector<size_t> offsets;
    std::atomic<bool> busy{false};     // admitted or launched, and not yet released
};

}  // namespace

struct birp_server {
    birp_device* dev = nullptr;
    birp_server_config cfg{};
    std::vector<birp_volume*> vols;
    DevVolume* d_vols = nullptr;
    int vol_cap = 64;
    size_t resident = 0;  // device bytes of the registered volumes (under mu)
    size_t held = 0;      // device bytes of the slots' image buffers, allocated or promised (under mu)
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    std::deque<Slot> slots;  // deque: slots hold atomics and never move
    std::thread worker;
    std::thread completer;
    CopyPool copy_pool;  // helpers of the completion thread
    std::deque<Slot*> launched;  // guarded by mu: launched slots not yet taken by the completion thread
    std::condition_variable cv_launched;
    uint64_t wake_gen = 0;  // bumped (under mu) by submissions, released slots and shutdown
    bool stop = false;
    bool stop_copy = false;
};

namespace {

void slot_release_images(Slot& s) {
    if (s.d_images) (void)hipFree(s.d_images);
    if (s.h_images) (void)hipHostFree(s.h_images);
    s.d_images = nullptr;
    s.h_images = nullptr;
    s.h_images_dev = nullptr;
    s.image_cap = 0;
}

// Reallocates slot s's image buffers to exactly `pixels` when they hold fewer (hipFree
// synchronises the device, so growth is made rare by the headroom chosen at admission): the
// pinned staging, written by the kernels through its device address, plus device images only
// when the device does not map the staging.
int slot_alloc_images(Slot& s, size_t pixels) {
    if (s.image_cap >= pixels) return BIRP_OK;
    slot_release_images(s);
    hipError_t e = hipHostMalloc(&s.h_images, pixels * 4, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(server images)");
    // touch the staging now: the first host access of each page of a fresh pinned allocation
    // faults it into the process, which would otherwise land on the first batches' copy-out
    memset(s.h_images, 0, pixels * 4);
    void* d = nullptr;
    if (hipHostGetDevicePointer(&d, s.h_images, 0) != hipSuccess) {
        (void)hipGetLastError();
        d = nullptr;
    }
    s.h_images_dev = (float*)d;
    if (!s.h_images_dev) {  // not mapped: the kernels write device images, which come back by DMA
        e = hipMalloc(&s.d_images, pixels * 4);
        if (e != hipSuccess) {
            slot_release_images(s);
            return hip_fail(e, "hipMalloc(server images)");
        }
    }
    s.image_cap = pixels;
    return BIRP_OK;
}

// Kernel groups of a batch: storage kernel x lanes per ray (1, 2, 4). Every request's tuning comes
// from its own view (latency_tuning), so a group has one tuning and a request renders the same
// whatever shares its launch.
constexpr int kGroups = BIRP_KDT_COUNT * 3;
int lane_index(int lanes) { return lanes == 4 ? 2 : (lanes == 2 ? 1 : 0); }
int request_group(const Request* r, birp_tuning* t) {
    *t = latency_tuning(r->req.view, r->vol->dv);
    return fast_kernel_dtype(r->vol->dv) * 3 + lane_index(t->lanes_per_ray);
}
// Requests in BIRP_MATH_EXACT render with the exact arithmetic of birp_render through the
// multi-view exact kernel, one launch per sample mode present; the others take the fast kernels.
bool exact_request(const Request* r) { return r->req.view.math == BIRP_MATH_EXACT; }
constexpr int kModes = SM_CMEM + 1;
constexpr size_t kCounterBytes = sizeof(unsigned) * kGroups * 256;

// Offsets of a launch's descriptor regions for n fast and nexact exact requests, 256-byte aligned,
// and the span that one copy uploads: the counters only when fast requests use them.
struct DescLayout {
    size_t prefix, eprefix, poses, evs, lo, hi;
};
DescLayout desc_layout(int n, int nexact) {
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    DescLayout L;
    L.prefix = kCounterBytes;
    L.eprefix = L.prefix + up(sizeof(int) * (n + kGroups));
    L.poses = L.eprefix + up(sizeof(int) * (nexact + kModes));
    L.evs = L.poses + up(sizeof(FastPose) * n);
    L.hi = L.evs + up(sizeof(ExactView) * nexact);
    L.lo = n ? 0 : L.prefix;
    return L;
}

// Launch the admitted requests of slot s (dispatcher thread, lock not held). Group g's poses are
// h_poses[base[g], base[g+1]) and its tile prefix sums h_prefix[base[g] + g, ... + count_g].
int launch_slot(birp_server* S, Slot& s) {
    size_t total = 0;
    s.offsets.clear();
    for (Request* r : s.reqs) {
        s.offsets.push_back(total);
        total += r->pixels;
    }
    int rc = slot_alloc_images(s, s.held / 4);  // admission charged held >= 4 * total
    if (rc) return rc;
    int base[kGroups + 1] = {}, fill[kGroups];
    birp_tuning gt[kGroups];
    for (Request* r : s.reqs) {
        if (exact_request(r)) continue;
        birp_tuning t;
        const int g = request_group(r, &t);
        base[g + 1]++;
        gt[g] = t;
    }
    for (int g = 0; g < kGroups; g++) {
        base[g + 1] += base[g];
        fill[g] = base[g];
    }
    // exact requests: group m (sample mode) holds views h_evs[estart[m], + ecount[m]) and prefix
    // sums h_eprefix[estart[m] + m, + ecount[m] + 1)
    int ecount[kModes] = {}, estart[kModes] = {}, efill[kModes] = {};
    size_t elds[kModes] = {};
    long long et2[kModes] = {}, et4[kModes] = {}, et8[kModes] = {};
    bool esmall[kModes], etiny[kModes];
    for (int m = 0; m < kModes; m++) esmall[m] = etiny[m] = true;
    for (Request* r : s.reqs) {
        if (!exact_request(r)) continue;
        ExactView ev;
        if ((rc = exact_view(r->vol, &r->req.view, nullptr, ev))) return rc;
        const int m = exact_mode_of(r->vol->dv, ev);
        ecount[m]++;
        et2[m] += exact_tiles(r->req.view.width, r->req.view.height, 2);
        et4[m] += exact_tiles(r->req.view.width, r->req.view.height, 4);
        et8[m] += exact_tiles(r->req.view.width, r->req.view.height, 8);
        const long long rays = (long long)r->req.view.width * r->req.view.height;
        esmall[m] = esmall[m] && rays <= 60000;
        etiny[m] = etiny[m] && rays <= 20000;
    }
    // each mode's launch shape, as exact batches choose it: a cube launch whose blocks the GPU holds
    // all at once takes 4 samples in flight, and 8 lanes per ray when every image has at most
    // 20,000 rays and those tiles fit, else 4 when every image has at most 60,000 rays
    int elanes[kModes], epipe[kModes];
    for (int m = 0; m < kModes; m++) {
        elanes[m] = exact_multi_lanes(m);
        epipe[m] = 2;
        if (m == SM_CMEM || S->dev->ncu <= 0) continue;
        const long long cap = 7LL * S->dev->ncu;
        if (etiny[m] && et8[m] <= cap) elanes[m] = 8;
        else if (esmall[m] && et4[m] <= cap) elanes[m] = 4;
        if (m == SM_QUAD_POS && (elanes[m] == 8 ? et8[m] : elanes[m] == 4 ? et4[m] : et2[m]) <= cap) epipe[m] = 4;
    }
    int nexact = 0;
    for (int m = 0; m < kModes; m++) nexact += ecount[m];
    const int n = base[kGroups];
    const DescLayout L = desc_layout(n, nexact);
    s.h_prefix = (int*)(s.h_desc + L.prefix);
    s.d_prefix = (int*)(s.d_desc + L.prefix);
    s.h_eprefix = (int*)(s.h_desc + L.eprefix);
    s.d_eprefix = (int*)(s.d_desc + L.eprefix);
    s.h_poses = (FastPose*)(s.h_desc + L.poses);
    s.d_poses = (FastPose*)(s.d_desc + L.poses);
    s.h_evs = (ExactView*)(s.h_desc + L.evs);
    s.d_evs = (ExactView*)(s.d_desc + L.evs);
    for (int m = 0, at = 0; m < kModes; m++) {
        estart[m] = efill[m] = at;
        at += ecount[m];
        s.h_eprefix[estart[m] + m] = 0;
    }
    int tabs[kGroups] = {};  // each group's largest volume table (nx + ny + nz)
    for (size_t k = 0; k < s.reqs.size(); k++) {
        Request* r = s.reqs[k];
        // a result in a birp_host_alloc buffer is written there by the kernel (zero-copy), the others
        // into the slot's pinned staging (or its device images when the staging is not mapped)
        float* out = r->dev_out ? r->dev_out : (s.h_images_dev ? s.h_images_dev : s.d_images) + s.offsets[k];
        if (exact_request(r)) {
            ExactView ev;
            if ((rc = exact_view(r->vol, &r->req.view, out, ev))) return rc;
            ev.vol = r->req.volume;
            const int m = exact_mode_of(r->vol->dv, ev);
            const int q = efill[m]++;
            s.h_evs[q] = ev;
            elds[m] = std::max(elds[m], exact_lds_bytes(r->vol->dv, r->req.view.samples, m));
            int* pre = s.h_eprefix + estart[m] + m;
            const long long next = (long long)pre[q - estart[m]] + exact_tiles(r->req.view.width, r->req.view.height, elanes[m]);
            if (next > INT32_MAX) return fail(BIRP_ERR_UNSUPPORTED, "launch of more than 2^31 - 1 tiles");
            pre[q - estart[m] + 1] = (int)next;
            continue;
        }
        birp_tuning t;
        const int g = request_group(r, &t);
        tabs[g] = std::max(tabs[g], volume_tab_entries(r->vol->dv));
        rc = fast_pose(r->vol, &r->req.view, r->req.volume, out, t.tile_w, t.tile_h, s.h_poses[fill[g]++]);
        if (rc) return rc;
    }
    for (int g = 0; g < kGroups; g++) {
        int* pre = s.h_prefix + base[g] + g;
        pre[0] = 0;
        for (int k = base[g]; k < base[g + 1]; k++) {
            birp_view tmp{};
            tmp.width = s.h_poses[k].W;
            tmp.height = s.h_poses[k].H;
            const long long next = (long long)pre[k - base[g]] + tiles_of(tmp, gt[g].tile_w, gt[g].tile_h);
            if (next > INT32_MAX) return fail(BIRP_ERR_UNSUPPORTED, "launch of more than 2^31 - 1 tiles");
            pre[k - base[g] + 1] = (int)next;
        }
    }
    // A lone exact request renders as its single exact render does: its view goes to the kernel by
    // value, with the single renders' lanes per ray and samples in flight. A lone fast request's
    // pose and volume go by value to the one-shot fast kernel, as a single fast render's do. Either
    // way the request gives the same bits as in a multi-view launch, and a launch of lone requests
    // uploads nothing.
    const Request* lone = nullptr;
    const Request* lone_fast = nullptr;
    for (size_t k = 0; k < s.reqs.size(); k++) {
        if (exact_request(s.reqs[k])) lone = nexact == 1 ? s.reqs[k] : nullptr;
        else lone_fast = n == 1 ? s.reqs[k] : nullptr;
    }
    hipError_t e = hipSuccess;
    if ((n && !lone_fast) || (nexact && !lone))
        e = hipMemcpyAsync(s.d_desc + L.lo, s.h_desc + L.lo, L.hi - L.lo, hipMemcpyHostToDevice, s.stream);
    if (e != hipSuccess) return hip_fail(e, "server descriptor upload");
    if (lone) {
        const DevVolume& dv = lone->vol->dv;
        rc = launch_exact(dv, s.h_evs[0], default_exact_lanes(dv.kind == SK_CMEM, lone->req.view.width, lone->req.view.height),
                          default_exact_pipe(), s.stream, nullptr);
        if (rc) return fail(rc, "exact launch failed");
    }
    for (int m = 0; m < kModes && !lone; m++) {
        if (!ecount[m]) continue;
        const int first = estart[m] + m;
        rc = launch_exact_multi(S->d_vols, s.d_evs + estart[m], s.d_eprefix + first, ecount[m], s.h_eprefix[first + ecount[m]], m,
                                elds[m], s.stream, nullptr, elanes[m], epipe[m]);
        if (rc) return fail(rc, "exact launch failed");
    }
    for (int g = 0; g < kGroups; g++) {
        const int count = base[g + 1] - base[g];
        if (!count) continue;
        const birp_tuning& t = gt[g];
        const int kdt = g / 3;
        const int total_tiles = s.h_prefix[base[g] + g + count];
        if (lone_fast) {
            rc = launch_fast_one(lone_fast->vol->dv, kdt, s.h_poses[base[g]], total_tiles, t.tile_w, t.tile_h, t.lanes_per_ray,
                                 t.unroll, s.stream, tabs[g]);
            if (rc) return fail(rc, "fast launch failed");
            continue;
        }
        const int grid = fast_grid_for(kdt, t.lanes_per_ray, t.unroll, total_tiles, tabs[g]);
        rc = launch_fast_batch(S->d_vols, kdt, s.d_poses + base[g], s.d_prefix + base[g] + g, count, total_tiles, s.d_work + g * 256,
                               t.tile_w, t.tile_h, t.lanes_per_ray, t.unroll, grid, s.stream, tabs[g]);
        if (rc) return rc;
    }
    // without a mapped staging, the other results come back to it in one DMA; the completion thread
    // copies them out
    size_t lo = total, hi = 0;
    for (size_t k = 0; k < s.reqs.size() && !s.h_images_dev; k++) {
        if (s.reqs[k]->dev_out) continue;
        lo = std::min(lo, s.offsets[k]);
        hi = std::max(hi, s.offsets[k] + s.reqs[k]->pixels);
    }
    if (e == hipSuccess && hi > lo) e = hipMemcpyAsync(s.h_images + lo, s.d_images + lo, (hi - lo) * 4, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
    if (e != hipSuccess) return hip_fail(e, "server launch");
    return BIRP_OK;
}

// Copy the slot's images out of pinned staging, split by bytes into parts of at least 2 MB over
// the copy pool's threads.
void copy_out(CopyPool& pool, Slot& s) {
    size_t total = 0;
    auto staged = [&s](size_t k) { return !s.reqs[k]->dev_out; };
    const size_t n = s.reqs.size();
    for (size_t k = 0; k < n; k++) total += staged(k) ? s.reqs[k]->pixels * 4 : 0;
    if (!total) return;
    auto run = [&s, &staged](size_t a, size_t b) {
        for (size_t k = a; k < b; k++)
            if (staged(k)) memcpy(s.reqs[k]->result, s.h_images + s.offsets[k], s.reqs[k]->pixels * 4);
    };
    const size_t parts = std::min<size_t>({n, (size_t)pool.helpers() + 1, std::max<size_t>(1, total >> 21)});
    if (parts <= 1) {
        run(0, n);
        return;
    }
    std::vector<size_t> cut(1, 0);
    size_t acc = 0, next = total / parts;
    for (size_t k = 0; k < n && cut.size() < parts; k++) {
        acc += staged(k) ? s.reqs[k]->pixels * 4 : 0;
        if (acc >= next) {
            cut.push_back(k + 1);
            next += total / parts;
        }
    }
    cut.push_back(n);
    const std::function<void(int)> part = [&](int p) {
        if ((size_t)p + 1 < cut.size()) run(cut[p], cut[p + 1]);
    };
    pool.run((int)cut.size() - 1, part);
}

void finish_slot(birp_server* S, Slot& s, int status) {
    double now = birp_clock_s();
    std::lock_guard<std::mutex> lk(S->mu);
    for (Request* r : s.reqs) {
        r->status = status;
        r->done_s = now;
        r->finished = true;
    }
    s.reqs.clear();
    s.busy = false;
    S->cv_done.notify_all();
    S->wake_gen++;  // a slot is free
    S->cv_work.notify_one();
}

// The completion thread polls the events of the launched slots while any is on the GPU (it
// spins, yielding, only then), copies each finished slot out and releases it.
void completer(birp_server* S) {
    (void)hipSetDevice(S->dev->ordinal);
    std::vector<Slot*> mine;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (mine.empty()) S->cv_launched.wait(lk, [&] { return S->stop_copy || !S->launched.empty(); });
            while (!S->launched.empty()) {
                mine.push_back(S->launched.front());
                S->launched.pop_front();
            }
            if (mine.empty()) return;  // stop_copy with nothing in flight
        }
        bool progressed = false;
        for (size_t k = 0; k < mine.size();) {
            Slot* s = mine[k];
            const hipError_t q = hipEventQuery(s->done);
            if (q == hipErrorNotReady) {
                k++;
                continue;
            }
            mine.erase(mine.begin() + (long)k);
            if (q == hipSuccess) {
                copy_out(S->copy_pool, *s);
                finish_slot(S, *s, BIRP_OK);
            } else {
                const int rc = hip_fail(q, "server batch");
                (void)hipStreamSynchronize(s->stream);
                finish_slot(S, *s, rc);
            }
            progressed = true;
        }
        if (!progressed) std::this_thread::yield();
    }
}

// Device bytes to charge slot s for a batch of `pixels` image pixels: its current charge when that
// holds them, else the batch plus headroom (half again, at least 4 Mi pixels) as far as `room`
// (the budget left for slot s) allows, or 0 when even the batch does not fit.
size_t slot_charge_for(const Slot& s, size_t pixels, size_t room) {
    const size_t need = pixels * 4;
    if (need <= s.held) return s.held;
    if (need > room) return 0;
    return std::min(std::max({need + need / 2, (size_t)16 << 20}), room);
}

// The dispatcher sleeps on cv_work until a submission, a released slot or shutdown bumps
// wake_gen, admits queued requests into the free slot with the largest image buffer and launches
// them; launched slots go to the completion thread.
void dispatcher(birp_server* S) {
    (void)hipSetDevice(S->dev->ordinal);
    const size_t budget = S->cfg.memory_budget;
    std::unique_lock<std::mutex> lk(S->mu);
    while (true) {
        Slot* free_slot = nullptr;
        bool any_busy = false;
        for (Slot& s : S->slots) {
            if (s.busy) any_busy = true;
            else if (!free_slot || s.held > free_slot->held) free_slot = &s;
        }
        if (S->stop && S->pending.empty() && !any_busy) break;
        const uint64_t seen = S->wake_gen;
        if (S->pending.empty() || !free_slot) {
            S->cv_work.wait(lk, [&] { return S->wake_gen != seen; });
            continue;
        }
        // a short window lets a trickle of arrivals share a launch while the GPU is busy
        if (any_busy && (int)S->pending.size() < S->cfg.max_batch && S->cfg.batch_window_ms > 0 && !S->stop) {
            const double left = S->pending.front()->submit_s + S->cfg.batch_window_ms * 1e-3 - birp_clock_s();
            if (left > 0) {
                S->cv_work.wait_for(lk, std::chrono::duration<double>(left), [&] { return S->wake_gen != seen; });
                continue;
            }
        }
        const double now = birp_clock_s();
        // device bytes for image buffers: all of it for this slot once the others are released
        const size_t images = budget > S->resident ? budget - S->resident : 0;
        const size_t room = images - std::min(images, S->held - free_slot->held);
        size_t pixels = 0, charge = free_slot->held;
        bool release_idle = false, failed = false;
        while (!S->pending.empty() && (int)free_slot->reqs.size() < S->cfg.max_batch) {
            Request* r = S->pending.front();
            if (r->pixels * 4 > images) {  // can never fit
                S->pending.pop_front();
                r->status = BIRP_ERR_ADMISSION;
                r->done_s = now;
                r->finished = failed = true;
                continue;
            }
            const size_t c = slot_charge_for(*free_slot, pixels + r->pixels, room);
            if (!c) {
                // an empty batch waits for the other slots' buffers: released below when idle,
                // else when their launches finish
                release_idle = free_slot->reqs.empty();
                break;
            }
            S->pending.pop_front();
            if (take_injected_failure(S->dev)) {
                r->status = BIRP_ERR_INJECTED;
                r->done_s = now;
                r->finished = failed = true;
                continue;
            }
            r->start_s = now;
            free_slot->reqs.push_back(r);
            pixels += r->pixels;
            charge = c;
        }
        for (Request* r : free_slot->reqs) r->batch_size = (int)free_slot->reqs.size();
        S->held += charge - free_slot->held;
        free_slot->held = charge;
        if (failed) S->cv_done.notify_all();
        if (!free_slot->reqs.empty()) {
            free_slot->busy = true;
            lk.unlock();
            const int rc = launch_slot(S, *free_slot);
            if (rc) {
                // work queued before the failure may still write the callers' pinned buffers:
                // let it drain before the requests are reported failed
                (void)hipStreamSynchronize(free_slot->stream);
                finish_slot(S, *free_slot, rc);
                lk.lock();
            } else {
                lk.lock();
                S->launched.push_back(free_slot);
                S->cv_launched.notify_one();
            }
            continue;
        }
        if (release_idle) {
            // give the budget of the idle slots' buffers back (no work of theirs is in flight)
            std::vector<Slot*> idle;
            for (Slot& s : S->slots)
                if (&s != free_slot && !s.busy && s.held) idle.push_back(&s);
            if (!idle.empty()) {
                lk.unlock();
                for (Slot* s : idle) slot_release_images(*s);
                lk.lock();
                for (Slot* s : idle) {
                    S->held -= s->held;
                    s->held = 0;
                }
                continue;
            }
            // the room is held by running launches: wait for one to finish
            S->cv_work.wait(lk, [&] { return S->wake_gen != seen; });
        }
    }
}

}  // namespace

extern "C" int birp_server_create(birp_device* dev, const birp_server_config* cfg, birp_server** out) {
    if (!dev || !cfg || !out) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    if (cfg->max_batch < 1 || cfg->max_batch > 4096 || cfg->max_inflight < 1 || cfg->max_inflight > 8)
        return fail(BIRP_ERR_INVALID_ARGUMENT, "max_batch in [1,4096], max_inflight in [1,8]");
    if (cfg->reserve_pixels > ((size_t)1 << 32)) return fail(BIRP_ERR_INVALID_ARGUMENT, "reserve_pixels above 2^32");
    if (cfg->reserve_pixels * 4 * cfg->max_inflight > cfg->memory_budget)
        return fail(BIRP_ERR_ADMISSION, "the slots' reserved images exceed the memory budget");
    (void)hipSetDevice(dev->ordinal);
    birp_server* S = new birp_server();
    S->dev = dev;
    S->cfg = *cfg;
    hipError_t e = hipMalloc(&S->d_vols, sizeof(DevVolume) * S->vol_cap);
    int rc = BIRP_OK;
    if (e != hipSuccess) rc = hip_fail(e, "hipMalloc(server volumes)");
    for (int k = 0; k < cfg->max_inflight; k++) S->slots.emplace_back();
    const int nb = cfg->max_batch;
    for (Slot& s : S->slots) {
        if (rc) break;
        if ((e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking)) != hipSuccess) { rc = hip_fail(e, "stream"); break; }
        if ((e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming)) != hipSuccess) { rc = hip_fail(e, "event"); break; }
        // every region of a launch fits: each is largest when all max_batch requests fill it
        const size_t bytes = desc_layout(nb, nb).hi;
        if ((e = hipHostMalloc(&s.h_desc, bytes, hipHostMallocDefault)) != hipSuccess) { rc = hip_fail(e, "pinned"); break; }
        if ((e = hipMalloc(&s.d_desc, bytes)) != hipSuccess) { rc = hip_fail(e, "descriptors"); break; }
        memset(s.h_desc, 0, kCounterBytes);  // the counters stay zero on the host
        s.d_work = (unsigned*)s.d_desc;
        if (cfg->reserve_pixels) {
            if ((rc = slot_alloc_images(s, cfg->reserve_pixels))) break;
            s.held = cfg->reserve_pixels * 4;
            S->held += s.held;
        }
    }
    if (rc) {
        birp_server_destroy(S);
        return rc;
    }
    S->copy_pool.start(7);
    S->completer = std::thread(completer, S);
    S->worker = std::thread(dispatcher, S);
    *out = S;
    return BIRP_OK;
}

extern "C" int birp_server_destroy(birp_server* S) {
    if (!S) return BIRP_OK;
    if (S->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = true;
            S->wake_gen++;
        }
        S->cv_work.notify_all();
        S->worker.join();
    }
    if (S->completer.joinable()) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop_copy = true;
        }
        S->cv_launched.notify_all();
        S->completer.join();
    }
    S->copy_pool.stop();
    for (Slot& s : S->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.d_images) (void)hipFree(s.d_images);
        if (s.h_images) (void)hipHostFree(s.h_images);
        if (s.h_desc) (void)hipHostFree(s.h_desc);
        if (s.d_desc) (void)hipFree(s.d_desc);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    if (S->d_vols) (void)hipFree(S->d_vols);
    for (auto& kv : S->reqs) delete kv.second;
    delete S;
    return BIRP_OK;
}

extern "C" int birp_server_add_volume(birp_server* S, birp_volume* vol, int* index) {
    if (!S || !vol || !index) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    if (vol->dev != S->dev) return fail(BIRP_ERR_INVALID_ARGUMENT, "volume on another device");
    std::lock_guard<std::mutex> lk(S->mu);
    if ((int)S->vols.size() >= S->vol_cap) return fail(BIRP_ERR_OUT_OF_MEMORY, "volume table full");
    if (S->resident + S->held + vol->bytes > S->cfg.memory_budget)
        return fail(BIRP_ERR_ADMISSION, "volume does not fit the memory budget");
    hipError_t e = hipMemcpy(S->d_vols + S->vols.size(), &vol->dv, sizeof(DevVolume), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "volume table upload");
    S->resident += vol->bytes;
    *index = (int)S->vols.size();
    S->vols.push_back(vol);
    return BIRP_OK;
}

extern "C" int birp_server_submit(birp_server* S, const birp_request* req, float* result, uint64_t* id) {
    if (!S || !req || !result || !id) return fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    bool exact_ok = true;  // the exact arithmetic reads an FP32 cube (or a CMEM model)
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (req->volume >= 0 && req->volume < (int)S->vols.size()) {
            const DevVolume& d = S->vols[req->volume]->dv;
            exact_ok = d.kind == SK_CMEM || d.linear != nullptr;
        }
    }
    birp_view v = req->view;
    // a request renders in the math its view names: BIRP_MATH_FAST through the fast kernels, any
    // other value as BIRP_MATH_EXACT through the multi-view exact kernel
    v.math = req->view.math == BIRP_MATH_FAST ? BIRP_MATH_FAST : BIRP_MATH_EXACT;
    v.flags = 0;
    if (v.math == BIRP_MATH_EXACT && !exact_ok) return fail(BIRP_ERR_UNSUPPORTED, "exact requests need an FP32 cube");
    int rc = validate_view(&v, v.math == BIRP_MATH_EXACT);
    if (rc) return rc;
    Request* r = new Request();
    r->req = *req;
    r->req.view = v;
    r->result = result;
    r->pixels = (size_t)v.width * v.height;
    r->dev_out = pinned_device_ptr(result, r->pixels * 4);
    r->submit_s = birp_clock_s();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {  // destroy has begun: the dispatcher may already have exited
            delete r;
            return fail(BIRP_ERR_INVALID_ARGUMENT, "server is being destroyed");
        }
        if (req->volume < 0 || req->volume >= (int)S->vols.size()) {
            delete r;
            return fail(BIRP_ERR_NOT_FOUND, "unknown volume index");
        }
        r->vol = S->vols[req->volume];
        r->id = S->next_id++;
        S->reqs[r->id] = r;
        S->pending.push_back(r);
        S->wake_gen++;
    }
    S->cv_work.notify_one();
    *id = r->id;
    return BIRP_OK;
}

extern "C" int birp_server_wait(birp_server* S, uint64_t id, double timeout_ms, birp_result_info* info) {
    if (!S) return fail(BIRP_ERR_INVALID_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(BIRP_ERR_NOT_FOUND, "unknown request id");
    Request* r = it->second;
    // one waiter per request: the waiter that sees it finish frees it
    if (r->waited) return fail(BIRP_ERR_INVALID_ARGUMENT, "another thread is waiting for this request");
    r->waited = true;
    auto ready = [&] { return r->finished; };
    if (timeout_ms < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
        r->waited = false;
        return fail(BIRP_ERR_TIMEOUT, "request not finished");
    }
    if (info) {
        info->status = r->status;
        info->submit_s = r->submit_s;
        info->start_s = r->start_s;
        info->done_s = r->done_s;
        info->batch_size = r->batch_size;
    }
    int status = r->status;
    S->reqs.erase(it);
    delete r;
    return status == BIRP_OK ? BIRP_OK : fail(status, birp_strerror(status));
}
