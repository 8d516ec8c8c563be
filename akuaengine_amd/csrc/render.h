// render.h -- headless point-sprite rendering (render.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace akua {

struct RenderParams {
    float vp[16];        // row-major view-projection (glm::perspective * glm::lookAt)
    float eye[3], front[3], right[3], up[3];
    float tanHalf, aspect;
    int width, height;
    float pointSize;     // > 0: every sprite's size; otherwise each particle's own
    int drawFloor;
};

// The camera of the reference (Camera.cpp:76-87, glm::lookAt, glm::perspective) for an image of
// width x height; fills everything but pointSize and drawFloor.
void camera_setup(const float position[3], float yawDeg, float pitchDeg, float fovYDeg, float nearZ,
                  float farZ, int width, int height, RenderParams& R);

// Slots 0..n-1 of one scene: X (physical order), colour and size by particle id, the logical
// (draw) index of each slot and the slot of each logical index.  zbuf: width * height keys;
// rgb: width * height * 3 bytes (device memory).
void launch_render(hipStream_t st, const float4* X, const float4* color, const float* size,
                   const uint32_t* id, const uint32_t* lid, const uint32_t* phys, uint32_t n,
                   const RenderParams& R, unsigned long long* zbuf, unsigned char* rgb);

}  // namespace akua
