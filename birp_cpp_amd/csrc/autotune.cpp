// autotune.cpp -- kernel-tuning search for the batched (fast) renders: every candidate
// (tile shape, lanes per ray, unroll) is timed on the caller's views with HIP events and the
// fastest is kept per (dtype, image-size class, samples class). The table persists as JSON lines.
// No reference counterpart: the reference has one serial loop (Camera.cpp:139-232).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/birp.h"
#include "capi_internal.h"

namespace {
std::mutex g_mu;
std::map<std::tuple<int, int, int>, std::pair<birp_tuning, double>> g_table;

int size_class(int w, int h) {
    long long p = (long long)w * h;
    return p <= 20000 ? 0 : (p <= 80000 ? 1 : 2);
}
int samples_class(int n) { return n <= 400 ? 0 : (n <= 1000 ? 1 : 2); }

std::vector<birp_tuning> candidates() {
    std::vector<birp_tuning> c;
    const int shapes[][3] = {{8, 8, 1}, {16, 4, 1}, {8, 4, 2}, {4, 8, 2}, {4, 4, 4}, {8, 2, 4}};
    for (auto& s : shapes)
        for (int u : {1, 2, 4}) c.push_back(birp_tuning{s[0], s[1], s[2], u});
    return c;
}
}  // namespace

extern "C" int birp_autotune_lookup(int dtype, int width, int height, int samples, birp_tuning* out) {
    if (!out) return birp::fail(BIRP_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find({dtype, size_class(width, height), samples_class(samples)});
    if (it == g_table.end()) return BIRP_ERR_NOT_FOUND;
    *out = it->second.first;
    return BIRP_OK;
}

static int write_table(const char* path) {
    FILE* f = fopen(path, "w");
    if (!f) return birp::fail(BIRP_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& kv : g_table) {
        const birp_tuning& t = kv.second.first;
        fprintf(f,
                "{\"dtype\": %d, \"size_class\": %d, \"samples_class\": %d, \"tile_w\": %d, \"tile_h\": %d, "
                "\"lanes_per_ray\": %d, \"unroll\": %d, \"ms\": %.6f}\n",
                std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), t.tile_w, t.tile_h, t.lanes_per_ray, t.unroll,
                kv.second.second);
    }
    fclose(f);
    return BIRP_OK;
}

extern "C" int birp_autotune_load(const char* path) {
    if (!path) return birp::fail(BIRP_ERR_INVALID_ARGUMENT, "null path");
    FILE* f = fopen(path, "r");
    if (!f) return birp::fail(BIRP_ERR_NOT_FOUND, std::string("cannot read ") + path);
    char line[512];
    int loaded = 0;
    while (fgets(line, sizeof line, f)) {
        int dt, sc, nc;
        birp_tuning t;
        double ms;
        if (sscanf(line,
                   "{\"dtype\": %d, \"size_class\": %d, \"samples_class\": %d, \"tile_w\": %d, \"tile_h\": %d, "
                   "\"lanes_per_ray\": %d, \"unroll\": %d, \"ms\": %lf}",
                   &dt, &sc, &nc, &t.tile_w, &t.tile_h, &t.lanes_per_ray, &t.unroll, &ms) == 8) {
            const bool ok = (t.lanes_per_ray == 1 || t.lanes_per_ray == 2 || t.lanes_per_ray == 4) &&
                            (t.unroll == 1 || t.unroll == 2 || t.unroll == 4) && t.tile_w > 0 && t.tile_h > 0 &&
                            (t.tile_w & (t.tile_w - 1)) == 0 && t.tile_w * t.tile_h * t.lanes_per_ray == 64 &&
                            (dt == BIRP_DTYPE_F32 || dt == BIRP_DTYPE_BF16 || dt == BIRP_DTYPE_F16);
            if (!ok) continue;  // the same rule as birp_batch_set_tuning
            std::lock_guard<std::mutex> lk(g_mu);
            g_table[{dt, sc, nc}] = {t, ms};
            loaded++;
        }
    }
    fclose(f);
    return loaded > 0 ? BIRP_OK : birp::fail(BIRP_ERR_NOT_FOUND, "no tuning entries in file");
}

extern "C" int birp_autotune(birp_volume* vol, const birp_view* views, int count, const char* path) {
    if (!vol || !views || count <= 0) return birp::fail(BIRP_ERR_INVALID_ARGUMENT, "bad argument");
    // group the views by class and tune each group separately
    std::map<std::tuple<int, int, int>, std::vector<birp_view>> groups;
    for (int k = 0; k < count; k++)
        groups[{vol->desc.dtype, size_class(views[k].width, views[k].height), samples_class(views[k].samples)}].push_back(views[k]);
    for (auto& g : groups) {
        birp_batch* b = nullptr;
        int rc = birp_batch_create(vol, g.second.data(), (int)g.second.size(), &b);
        if (rc) return rc;
        birp_tuning best{};
        double best_ms = 1e30;
        for (const birp_tuning& t : candidates()) {
            if ((rc = birp_batch_set_tuning(b, &t))) break;
            double ms = 1e30;
            for (int rep = 0; rep < 4 && rc == BIRP_OK; rep++) {
                if ((rc = birp_batch_launch(b))) break;
                if ((rc = birp_batch_wait(b))) break;
                double m = 0;
                birp_batch_stats(b, nullptr, &m);
                if (rep > 0) ms = std::min(ms, m);  // first launch warms caches
            }
            if (rc) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
        birp_batch_free(b);
        if (rc) return rc;
        std::lock_guard<std::mutex> lk(g_mu);
        g_table[g.first] = {best, best_ms};
    }
    return path ? write_table(path) : BIRP_OK;
}
