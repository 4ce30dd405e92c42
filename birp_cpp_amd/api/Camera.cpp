// Camera.cpp -- the reference's Camera API (reference src/Camera.cpp) with Render on the GPU.
#include "Camera.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>

#include "Cmem.h"
#include "birp.h"
#include "device.h"
#include "fits.h"

Camera::Camera(Space& cube, float pixel_size_degrees, float plot_fov_h, float plot_fov_w) {
    fov_x = plot_fov_w;
    fov_y = plot_fov_h;
    pixel_size_deg = pixel_size_degrees;
    dataCube = &cube;
    image_dimension_y = (int)std::round(plot_fov_h / pixel_size_deg);  // Camera.cpp:38-39
    image_dimension_x = (int)std::round(plot_fov_w / pixel_size_deg);
    image.assign(image_dimension_y, std::vector<float>(image_dimension_x, 0.f));
    position = {0, 0, 0};
    aim = {0, 0, 0};
}

Camera::Camera(Space& cube, int width, int height, float plot_fov_h, float plot_fov_w) {
    if (width <= 0 || height <= 0) throw std::invalid_argument("Camera: image size must be positive");
    fov_x = plot_fov_w;
    fov_y = plot_fov_h;
    pixel_size_deg = plot_fov_w / width;
    dataCube = &cube;
    image_dimension_x = width;
    image_dimension_y = height;
    image.assign(height, std::vector<float>(width, 0.f));
    position = {0, 0, 0};
    aim = {0, 0, 0};
}

Camera::~Camera() {}

Camera::PinnedImage::~PinnedImage() {
    if (p_) birp_host_free(p_);
}

float* Camera::PinnedImage::get(std::size_t n) {
    if (n_ < n) {
        if (p_) birp_host_free(p_);
        p_ = nullptr;
        n_ = 0;
        void* q = nullptr;
        birp_api::check(birp_host_alloc(n * sizeof(float), &q), "Camera::Render staging");
        p_ = static_cast<float*>(q);
        n_ = n;
    }
    return p_;
}

void Camera::SetPosition(float& x, float& y, float& z) { position = {x, y, z}; }
void Camera::SetAim(float x, float y, float z) { aim = {x, y, z}; }

void Camera::SetRay(float length, int samples) {
    if (!(length > 0) || samples < 2) throw std::invalid_argument("Camera::SetRay: length > 0 and samples >= 2");
    ray_length_ = length;
    ray_samples = samples;
}
void Camera::SetSemantics(int s) { semantics_ = s; }
void Camera::SetMath(int m) { math_ = m; }

namespace {
// The step weight w = t_1 - t_0 in FP32 (GenerateRayDistWidth, Camera.cpp:69-81).
float ray_width_of(float L, int n) {
    volatile float t0 = L * ((0 + 0.5f) / n);
    volatile float t1 = L * ((1 + 0.5f) / n);
    return t1 - t0;
}
}  // namespace

birp_view Camera::View() const {
    birp_view v{};
    v.width = image_dimension_x;
    v.height = image_dimension_y;
    v.fov_w = fov_x;
    v.fov_h = fov_y;
    v.position[0] = position.x;
    v.position[1] = position.y;
    v.position[2] = position.z;
    v.aim[0] = aim.x;
    v.aim[1] = aim.y;
    v.aim[2] = aim.z;
    v.ray_length = ray_length_;
    v.samples = ray_samples;
    v.semantics = semantics_;
    v.math = math_;
    return v;
}

void Camera::CheckImage() const {
    const int W = image_dimension_x, H = image_dimension_y;
    if (W <= 0 || H <= 0) throw std::invalid_argument("Camera::Render: empty image");
    if ((int)image.size() != H) throw std::invalid_argument("Camera::Render: image has the wrong number of rows");
    for (const auto& row : image)
        if ((int)row.size() != W) throw std::invalid_argument("Camera::Render: image row of the wrong length");
}

birp_volume* Camera::DeviceSpace() const {
    if (auto* cube = dynamic_cast<DataCube*>(dataCube)) return cube->DeviceVolume(BIRP_DTYPE_F32);
    if (auto* cmem = dynamic_cast<CMEM*>(dataCube)) return cmem->DeviceVolume();
    return nullptr;
}

// Any other Space: its GetSample is host code (a virtual call, Camera.cpp:191). The GPU forms the
// rays' sample points and integrates; the host only samples, a block of image rows at a time,
// and stops calling for a ray at its first negative scaled sample, as the reference's loop does
// (Camera.cpp:207-210: s < 0 exactly when the FP32 product v * w is negative).
void Camera::RenderOtherSpace() {
    const int W = image_dimension_x, H = image_dimension_y, n = ray_samples;
    birp_view v = View();
    v.math = BIRP_MATH_EXACT;
    v.flags = BIRP_RENDER_ACCUMULATE;  // Render() adds into the image (Camera.cpp:223)
    const float w = ray_width_of(ray_length_, n);
    const size_t per_row = (size_t)W * n;
    const int rows = (int)std::max<size_t>(1, std::min<size_t>((size_t)H, ((size_t)1 << 22) / per_row));
    std::vector<float> xyz((size_t)rows * per_row * 3), vals((size_t)rows * per_row), img((size_t)rows * W);
    birp_device* dev = birp_api::device();
    for (int r0 = 0; r0 < H; r0 += rows) {
        const int r1 = std::min(H, r0 + rows);
        birp_api::check(birp_space_points(dev, &v, r0, r1, xyz.data()), "Camera::Render");
        const size_t npix = (size_t)(r1 - r0) * W;
        for (size_t p = 0; p < npix; p++) {
            const float* X = xyz.data() + p * n * 3;
            float* V = vals.data() + p * n;
            for (int k = 0; k < n; k++) {
                const float s = dataCube->GetSample(X[3 * k], X[3 * k + 1], X[3 * k + 2]);
                V[k] = s;
                if (s * w < 0.f) break;
            }
        }
        for (int j = r0; j < r1; j++) std::memcpy(img.data() + (size_t)(j - r0) * W, image[j].data(), (size_t)W * 4);
        birp_api::check(birp_space_integrate(dev, &v, r0, r1, vals.data(), img.data()), "Camera::Render");
        for (int j = r0; j < r1; j++) std::memcpy(image[j].data(), img.data() + (size_t)(j - r0) * W, (size_t)W * 4);
    }
    last_in_cube_ = -1;
    last_ms_ = 0.0;
}

void Camera::Render() {
    CheckImage();
    birp_volume* vol = DeviceSpace();
    if (!vol) {
        RenderOtherSpace();
        return;
    }
    const int W = image_dimension_x, H = image_dimension_y;
    birp_view v = View();
    float* buf = staging_.get((size_t)W * H);
    v.flags = 0;
    if (math_ == BIRP_MATH_EXACT) {
        // the exact render adds to the image on the device (Camera.cpp:223); an image of +0.0
        // bits (a fresh camera) is the same as starting from zero, so it is not uploaded
        uint32_t any = 0;
        for (int j = 0; j < H; j++) {
            std::memcpy(buf + (size_t)j * W, image[j].data(), (size_t)W * 4);
            for (int i = 0; i < W; i++) {
                uint32_t u;
                std::memcpy(&u, &buf[(size_t)j * W + i], 4);
                any |= u;
            }
        }
        if (any) v.flags = BIRP_RENDER_ACCUMULATE;
    }
    birp_render_stats st{};
    birp_api::check(birp_render(vol, &v, buf, &st), "Camera::Render");
    for (int j = 0; j < H; j++) {
        if (math_ == BIRP_MATH_EXACT) {
            std::memcpy(image[j].data(), buf + (size_t)j * W, (size_t)W * 4);
        } else {
            // the fast form renders a fresh image; keep the accumulate contract on the host
            for (int i = 0; i < W; i++) image[j][i] += buf[(size_t)j * W + i];
        }
    }
    last_in_cube_ = st.samples_in_cube;
    last_ms_ = st.kernel_ms;
}

// Exact renders of several cameras (each with its own DataCube or CMEM) in one exact batch: the
// images are added into each camera's image in place (pinned host memory the kernels write), so
// the result equals calling Render() on each camera in turn. Cameras whose Space is another type,
// or whose math is FAST, are rendered by their own Render().
void Camera::RenderAll(const std::vector<Camera*>& cameras) {
    std::vector<Camera*> batch;
    std::vector<birp_volume*> vols;
    std::vector<int> vol_of;
    std::vector<birp_view> views;
    for (Camera* c : cameras) {
        if (!c) throw std::invalid_argument("Camera::RenderAll: null camera");
        c->CheckImage();
        birp_volume* vol = c->math_ == BIRP_MATH_EXACT ? c->DeviceSpace() : nullptr;
        if (!vol) {
            c->Render();
            continue;
        }
        size_t k = 0;
        while (k < vols.size() && vols[k] != vol) k++;
        if (k == vols.size()) vols.push_back(vol);
        vol_of.push_back((int)k);
        birp_view v = c->View();
        v.flags = BIRP_RENDER_ACCUMULATE;  // the kernels start from the staged images
        views.push_back(v);
        batch.push_back(c);
    }
    if (batch.empty()) return;
    size_t total = 0;
    for (const birp_view& v : views) total += (size_t)v.width * v.height;
    void* arena = nullptr;
    birp_api::check(birp_host_alloc(total * sizeof(float), &arena), "Camera::RenderAll");
    birp_batch* b = nullptr;
    try {
        float* out = static_cast<float*>(arena);
        size_t off = 0;
        for (Camera* c : batch) {
            for (int j = 0; j < c->image_dimension_y; j++)
                std::memcpy(out + off + (size_t)j * c->image_dimension_x, c->image[j].data(), (size_t)c->image_dimension_x * 4);
            off += (size_t)c->image_dimension_x * c->image_dimension_y;
        }
        birp_api::check(birp_batch_create_exact(vols.data(), (int)vols.size(), views.data(), vol_of.data(), (int)views.size(), &b),
                        "Camera::RenderAll");
        birp_api::check(birp_batch_set_output(b, out), "Camera::RenderAll");
        birp_api::check(birp_batch_launch(b), "Camera::RenderAll");
        birp_api::check(birp_batch_wait(b), "Camera::RenderAll");
        std::vector<long long> counts(batch.size());
        birp_api::check(birp_batch_samples(b, counts.data()), "Camera::RenderAll");
        double ms = 0.0;
        birp_batch_stats(b, nullptr, &ms);
        off = 0;
        for (size_t q = 0; q < batch.size(); q++) {
            Camera* c = batch[q];
            for (int j = 0; j < c->image_dimension_y; j++)
                std::memcpy(c->image[j].data(), out + off + (size_t)j * c->image_dimension_x, (size_t)c->image_dimension_x * 4);
            off += (size_t)c->image_dimension_x * c->image_dimension_y;
            c->last_in_cube_ = counts[q];
            c->last_ms_ = ms;
        }
    } catch (...) {
        birp_batch_free(b);
        birp_host_free(arena);
        throw;
    }
    birp_batch_free(b);
    birp_host_free(arena);
}

// Camera.cpp:235-250: std::to_string of every value followed by a comma, one image row per line.
int Camera::ToDat(std::string filename) {
    std::ofstream out(filename + ".dat");
    for (int j = 0; j < image_dimension_y; ++j) {
        for (int i = 0; i < image_dimension_x; ++i) {
            out << std::to_string(image[j][i]);
            out << ",";
        }
        out << "\n";
    }
    return 1;
}

// Camera.cpp:252-307: the raster W x H (row 0 first) with the header cards CRVAL1/CDELT1,
// CRVAL2/CDELT2, POS_*/AIM_* and the string record RHO; an existing file is overwritten.
int Camera::ToFITS(std::string filename) {
    std::vector<float> out((size_t)image_dimension_x * image_dimension_y);
    for (int j = 0; j < image_dimension_y; j++)
        for (int i = 0; i < image_dimension_x; i++) out[(size_t)j * image_dimension_x + i] = image[j][i];
    auto num = [](float x) {
        char b[32];
        std::snprintf(b, sizeof b, "%.9G", x);
        return std::string(b);
    };
    std::vector<birp_api::FitsCard> cards = {
        {"CRVAL1", num(-(fov_x / 2)), false, "[deg]"},
        {"CDELT1", num(pixel_size_deg), false, "[deg]"},
        {"CRVAL2", num(-(fov_y / 2)), false, "[deg]"},
        {"CDELT2", num(pixel_size_deg), false, "[deg]"},
        {"POS_X", num(position.x), false, "[float]"},
        {"POS_Y", num(position.y), false, "[float]"},
        {"POS_Z", num(position.z), false, "[float]"},
        {"AIM_X", num(aim.x), false, "[float]"},
        {"AIM_Y", num(aim.y), false, "[float]"},
        {"AIM_Z", num(aim.z), false, "[float]"},
        {"rho", "0", true, "[t] comment"},
    };
    if (!birp_api::write_fits(filename + ".fits", image_dimension_x, image_dimension_y, out.data(), cards))
        throw std::runtime_error("Camera::ToFITS: cannot write " + filename + ".fits");
    return 1;
}
