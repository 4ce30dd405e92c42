// Camera.h -- the SXI camera: ray generation, march and integration into an image
// (reference include/Camera.h:20-63). Render runs on the GPU for DataCube and CMEM spaces.
#pragma once
#include <cstddef>
#include <string>
#include <vector>

#include "DataCube.h"
#include "birp.h"

class Camera {
public:
    // W = round(fov_w / pixel), H = round(fov_h / pixel) (Camera.cpp:38-39); 200 samples per ray
    // over 44 RE (Camera.cpp:31, 65).
    Camera(Space& cube, float pixel_size_deg, float plot_fov_h, float plot_fov_w);
    // Addition: an explicit W x H (576x336 has no pixel size under the rule above).
    Camera(Space& cube, int width, int height, float plot_fov_h, float plot_fov_w);
    ~Camera();

    void SetPosition(float&, float&, float&);
    void SetAim(float, float, float);
    // Adds one render into `image` (repeated calls accumulate, as Camera.cpp:223 does). DataCube and
    // CMEM spaces render on the GPU; any other Space subclass is sampled through its own virtual
    // GetSample on the host at the sample points the GPU forms, and the GPU integrates the values.
    // Throws std::invalid_argument for a malformed image and std::runtime_error for a device failure.
    void Render();
    int ToDat(std::string filename);   // CSV, one row per line, trailing commas (Camera.cpp:235-250)
    int ToFITS(std::string filename);  // FITS, BITPIX -32, row 0 first (Camera.cpp:252-307)

    Space* dataCube;
    std::vector<std::vector<float>> image;

    // ---- additions ----
    void SetRay(float length, int samples);  // default 44 RE, 200 samples
    void SetSemantics(int semantics);        // birp.h BIRP_SEM_*, default BIRP_SEM_REFERENCE
    void SetMath(int math);                  // birp.h BIRP_MATH_*, default BIRP_MATH_EXACT
    int Width() const { return image_dimension_x; }
    int Height() const { return image_dimension_y; }
    long long LastSamplesInCube() const { return last_in_cube_; }  // -1 for other Space types
    double LastKernelMs() const { return last_ms_; }
    // Render() of every camera, the exact DataCube / CMEM ones together in one batched launch
    // (the same images as rendering them one after another).
    static void RenderAll(const std::vector<Camera*>& cameras);

private:
    birp_view View() const;
    void CheckImage() const;
    birp_volume* DeviceSpace() const;
    void RenderOtherSpace();
    int image_dimension_x;
    int image_dimension_y;
    float fov_x;
    float fov_y;
    struct {
        float x;
        float y;
        float z;
    } position;
    struct {
        float x;
        float y;
        float z;
    } aim;
    float pixel_size_deg;
    float ray_length_ = 44.f;
    int ray_samples = 200;
    int semantics_ = 0;
    int math_ = 0;
    long long last_in_cube_ = 0;
    double last_ms_ = 0.0;
    // The flat H x W image handed to the C ABI, kept between renders, in pinned host memory
    // (birp_host_alloc) so that a render arrives in it by DMA. A copied Camera starts without one.
    class PinnedImage {
    public:
        PinnedImage() = default;
        PinnedImage(const PinnedImage&) {}
        PinnedImage& operator=(const PinnedImage&) { return *this; }
        ~PinnedImage();
        float* get(std::size_t n);  // at least n floats
    private:
        float* p_ = nullptr;
        std::size_t n_ = 0;
    };
    PinnedImage staging_;
};
