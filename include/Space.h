// Space.h -- the sampler interface the Camera renders (reference include/Space.h:18-23).
// DataCube and CMEM implement it; Camera::Render sends those two to the GPU through the C ABI
// of birp.h. The virtual destructor is an addition (owners may delete through Space*).
#pragma once

class Space {
public:
    virtual ~Space() = default;
    virtual void Init() = 0;
    virtual float GetSample(float x, float y, float z) = 0;
};
