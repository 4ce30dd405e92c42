// Camera.cpp -- the reference's Camera API (reference src/Camera.cpp) with Render on the GPU.
#include "Camera.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
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

void Camera::Render() {
    birp_volume* vol = nullptr;
    if (auto* cube = dynamic_cast<DataCube*>(dataCube)) vol = cube->DeviceVolume(BIRP_DTYPE_F32);
    else if (auto* cmem = dynamic_cast<CMEM*>(dataCube)) vol = cmem->DeviceVolume();
    else throw std::invalid_argument("Camera::Render: the GPU renders DataCube and CMEM spaces only");
    const int W = image_dimension_x, H = image_dimension_y;
    if (W <= 0 || H <= 0) throw std::invalid_argument("Camera::Render: empty image");
    birp_view v{};
    v.width = W;
    v.height = H;
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
    if ((int)image.size() != H) throw std::invalid_argument("Camera::Render: image has the wrong number of rows");
    for (const auto& row : image)
        if ((int)row.size() != W) throw std::invalid_argument("Camera::Render: image row of the wrong length");
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
