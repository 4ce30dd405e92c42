// capi_internal.h -- definitions of the opaque C-ABI handles (include/birp.h) and helpers shared
// by capi.cpp, server.cpp and autotune.cpp.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <string>
#include <vector>

#include "../../include/birp.h"
#include "birp_internal.h"

// What a tile plan depends on: the rays' geometry, the cube's box and the tile shape. Filled
// after a memset, so that two keys compare with memcmp.
struct ExactPlanKey {
    float rinv[9];
    float px, py, pz, tanx, tany, aspect, L;
    float ox, oy, oz, sx, sy, sz;
    int nx, ny, nz;
    int W, H, n, semantics, lanes;
};

struct birp_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;      // render stream (default priority)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // per-render timing
    hipEvent_t ev_a = nullptr, ev_b = nullptr; // birp_timer_begin/end
    void* d_counter = nullptr;         // 64 B: the upload's class-scan flags
    void* scratch = nullptr;           // image / point staging, grown on demand
    size_t scratch_bytes = 0;
    void* pinned = nullptr;            // pinned host staging for birp_render's image copy
    void* pinned_dev = nullptr;        // its device address (null when the device does not map it)
    size_t pinned_bytes = 0;
    void* rbuf = nullptr;              // birp_render: [u64 sample counter | pad to 256 B | image]
    size_t rbuf_bytes = 0;
    unsigned long long count_base = 0; // the counter's value at its last read (it only grows)
    bool count_known = false;          // false: clear the counter before its next use
    std::atomic<int> inject_failures{0};
    void* noalloc_scratch = nullptr;   // this GPU's fixed scratch array (birp_render_views_noalloc)
    // single exact renders: kPlanSlots tile plans (BIRP_PLAN_INTS ints each, the order first) and
    // what each was planned for, and the last kSeen views rendered without a plan
    static constexpr int kPlanSlots = 4, kSeen = 4;
    int* plan_order = nullptr;
    int ncu = 0;
    ExactPlanKey plan_key[kPlanSlots] = {};
    bool plan_valid[kPlanSlots] = {};
    unsigned long long plan_used[kPlanSlots] = {};  // last use (plan_clock), for LRU replacement
    unsigned long long plan_clock = 0;
    ExactPlanKey seen_key[kSeen] = {};
    bool seen_valid[kSeen] = {};
    int seen_next = 0;
};

struct birp_volume {
    birp_device* dev = nullptr;
    birp_volume_desc desc{};
    birp::DevVolume dv{};
    float* linear = nullptr;
    void* quad = nullptr;
    void* oct = nullptr;
    float* bscale = nullptr;
    unsigned* qtab = nullptr;
    unsigned* otab = nullptr;
    bool morton = false;
    size_t bytes = 0;
};

struct birp_image {
    birp_device* dev = nullptr;
    int W = 0, H = 0;
    float* d = nullptr;
};

struct birp_batch {
    birp_volume* vol = nullptr;
    std::vector<birp_view> views;
    std::vector<size_t> offsets;
    size_t total_pixels = 0;
    long long nominal = 0;
    float* d_images = nullptr;
    float* host_out = nullptr;      // birp_batch_set_output: images go to this pinned host arena
    float* host_out_dev = nullptr;  // its device address
    birp::FastPose* d_poses = nullptr;
    int* d_prefix = nullptr;
    birp::DevVolume* d_vols = nullptr;
    unsigned* d_work = nullptr;
    int total_tiles = 0;
    int grid = 0;
    birp_tuning tuning{4, 4, 4, 2};
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool launched = false;
    double last_ms = 0.0;
    // exact batches (birp_batch_create_exact): views over several volumes, grouped by sample mode,
    // one multi-view launch per group
    bool exact = false;
    std::vector<birp_volume*> evols;  // the batch's volumes (evols[vol_of[k]] renders view k)
    std::vector<int> vol_of;
    struct ExactGroup {
        int mode = 0, first = 0, count = 0, total = 0;  // views order[first, first + count)
        int lanes = 2, pipe = 2;  // lanes per ray and samples in flight per lane of its launch
        bool resident = false;    // every block resident at once: one snake deal over the CUs
        size_t lds = 0;
        long long plan_at = -1;  // the group's tile plan in d_plan (total entries), or -1: none
    };
    std::vector<ExactGroup> groups;
    std::vector<int> order;           // view indices, group by group
    std::vector<size_t> count_off;    // each view's first sample-count slot
    size_t count_slots = 0;
    birp::ExactView* d_evs = nullptr;
    int* d_plan = nullptr;            // tile plans of the groups (each view's tiles, heaviest first)
    unsigned* d_counts = nullptr;     // one slot per wave: its in-cube samples of the last launch
};

namespace birp {
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int validate_view(const birp_view* v, bool exact);
int fast_pose(const birp_volume* vol, const birp_view* v, int vol_index, float* out_dev, int tile_w, int tile_h, FastPose& fp);
int exact_view(const birp_volume* vol, const birp_view* v, float* out_dev, ExactView& ev);
int default_exact_lanes(bool cmem, int W, int H);
int default_exact_pipe();
int tiles_of(const birp_view& v, int tile_w, int tile_h);
bool take_injected_failure(birp_device* dev);
int choose_tuning(const birp_volume* vol, const std::vector<birp_view>& views, birp_tuning& t);
birp_tuning latency_tuning(const birp_view& v, const birp::DevVolume& vol);
// true when [p, p + bytes) lies inside one birp_host_alloc allocation (pinned host memory)
bool pinned_range(const void* p, size_t bytes);
// its device address when the allocation is mapped into the device's address space, else null
float* pinned_device_ptr(const void* p, size_t bytes);
}  // namespace birp
