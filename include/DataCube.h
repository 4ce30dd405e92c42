// DataCube.h -- an emissivity cube on a uniform grid (reference include/DataCube.h:22-50).
//
// Same public members and methods as the reference. The values live on the host in `slices`
// ([z][y][x]) exactly as there; the first render or GetSample uploads them to the GPU and keeps
// the device copy. Additions (marked below) cover what the GPU copy needs: Touch() after
// changing slices/coords by hand, batched sampling, a .dat writer and array setters.
#pragma once
#include <cstddef>
#include <string>
#include <vector>

#include "Space.h"

struct birp_volume;

class DataCube : public Space {
public:
    DataCube();
    ~DataCube() override;
    DataCube(const DataCube& other);             // copies the host data, not the device copy
    DataCube& operator=(const DataCube& other);

    void Init() override;
    void Init(std::string& filename, bool debug);
    // Parse an exported MHD .dat file (reference DataCube.cpp:28-138): line 1 metadata, line 10
    // "nx ny nz", then the x, y and z coordinates, then the values x fastest, then y, then z.
    // Throws std::invalid_argument for a file that cannot be read or does not hold nx*ny*nz values.
    void Load(std::string filename, bool debug);
    // Trilinear or nearest sample at (x,y,z), -1 outside the cube (DataCube.cpp:203-277),
    // evaluated on the GPU with the reference's FP32 arithmetic.
    float GetSample(float x, float y, float z) override;
    void SetTrilinear(bool);

    std::vector<float> coords_x;
    std::vector<float> coords_y;
    std::vector<float> coords_z;
    struct {
        int x;
        int y;
        int z;
    } size;
    float spacing[3];
    float origin_coords[3];
    std::vector<std::vector<std::vector<float>>> slices;

    // ---- additions ----
    // Mark the host data as changed (after writing to slices/coords directly): the next render
    // re-uploads the cube.
    void Touch();
    bool IsTrilinear() const { return trilinear; }
    // GetSample at n points (xyz: 3n floats) in one GPU call.
    void GetSamples(const float* xyz, std::size_t n, float* out);
    // Fill from a [z][y][x] array on a uniform grid starting at origin with the given spacing.
    void SetValues(int nx, int ny, int nz, const float* origin, const float* spacing_in, const float* values_zyx);
    // Write the cube in the .dat layout Load reads.
    void Save(const std::string& filename) const;
    // The device copy at the given storage precision (birp.h BIRP_DTYPE_*), uploaded if stale.
    birp_volume* DeviceVolume(int dtype = 0);

private:
    void InitSpacing();
    void InitOriginCoords();
    void InitSize();
    float GetAxisSpacing(const std::vector<float>& coords);
    void DropDevice();
    // The reference leaves this uninitialised until SetTrilinear (DataCube.h:49); its CLI
    // default is nearest (main.cpp:56), which is used here.
    bool trilinear = false;
    birp_volume* device_[3] = {nullptr, nullptr, nullptr};
    unsigned long long version_ = 1, device_version_[3] = {0, 0, 0};
};
