// render.hip -- headless point-sprite rendering of the particles (akua_render, include/akua_io.h).
//
// The reference draws its particle VBO with GL point sprites (Renderer.cpp:28-56;
// fluidParticle.vert: gl_PointSize = size / |world - camera|; fluidParticle.frag: discard where
// |gl_PointCoord - 0.5| > 0.5, colour * exp(-3 d^2)) over a floor mesh (floor.frag: checker of
// floor(x) + floor(z)), depth test GL_LESS, particles in buffer order.  Here:
//   k_render_clear   every pixel key to the maximum (no sprite);
//   k_render_splat   one thread per particle: the clip-volume test of a GL point, the sprite
//                    square, and for each fragment centre inside the disc an atomic minimum of
//                    (window depth bits << 32 | draw index) -- the smallest depth wins and, at
//                    equal depth, the earlier particle in buffer order, which is what GL_LESS
//                    keeps when the sprites are drawn in that order;
//   k_render_resolve one thread per pixel: the winning sprite's colour at this fragment, or the
//                    floor where a ray through the pixel centre meets it nearer, or the clear
//                    colour; written as RGB8 rows from the top.
#include "render.h"

#include <cmath>

namespace akua {

namespace {

constexpr float kMaxPointSize = 256.0f;  // GL clamps sprites to an implementation range
constexpr float kFloorY = -0.02f;        // FLOOR_LEVEL (Application.cpp:16)
constexpr float kFloorExtent = 100.0f;   // floor quad x, z in [0, 100] (Application.cpp:130-135)

__global__ void k_render_clear(unsigned long long* __restrict__ z, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = ~0ull;
}

struct Sprite {
    float xw, yw, zw, s;  // window centre (pixels, origin bottom-left), depth in [0, 1], size
};

// The sprite of the particle at world position p, or false when GL would clip the point.
__device__ __forceinline__ bool sprite_of(const RenderParams& R, float px, float py, float pz,
                                          float size, Sprite& out) {
    const float* m = R.vp;
    const float cx = ((m[0] * px + m[1] * py) + m[2] * pz) + m[3];
    const float cy = ((m[4] * px + m[5] * py) + m[6] * pz) + m[7];
    const float cz = ((m[8] * px + m[9] * py) + m[10] * pz) + m[11];
    const float cw = ((m[12] * px + m[13] * py) + m[14] * pz) + m[15];
    if (!(cw > 0.0f) || !(fabsf(cx) <= cw) || !(fabsf(cy) <= cw) || !(fabsf(cz) <= cw)) return false;
    const float nx = cx / cw, ny = cy / cw, nz = cz / cw;
    out.xw = (nx * 0.5f + 0.5f) * (float)R.width;
    out.yw = (ny * 0.5f + 0.5f) * (float)R.height;
    out.zw = nz * 0.5f + 0.5f;
    const float dx = px - R.eye[0], dy = py - R.eye[1], dz = pz - R.eye[2];
    const float dist = sqrtf((dx * dx + dy * dy) + dz * dz);
    float s = size / dist;
    s = fminf(fmaxf(s, 1.0f), kMaxPointSize);
    if (!(s == s)) return false;
    out.s = s;
    return true;
}

// Squared distance of fragment (fx, fy) from the sprite centre in sprite coordinates (0.25 is the
// disc's edge).
__device__ __forceinline__ float sprite_d2(const Sprite& sp, int fx, int fy) {
    const float x0 = sp.xw - 0.5f * sp.s, y0 = sp.yw - 0.5f * sp.s;
    const float ox = ((float)fx + 0.5f - x0) / sp.s - 0.5f;
    const float oy = ((float)fy + 0.5f - y0) / sp.s - 0.5f;
    return ox * ox + oy * oy;
}

__global__ void k_render_splat(const float4* __restrict__ X, const float4* __restrict__ color,
                               const float* __restrict__ size, const uint32_t* __restrict__ id,
                               const uint32_t* __restrict__ lid, uint32_t n, RenderParams R,
                               unsigned long long* __restrict__ zbuf) {
    (void)color;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = X[i];
    const float sz = R.pointSize > 0.0f ? R.pointSize : size[id[i]];
    Sprite sp;
    if (!sprite_of(R, x.x, x.y, x.z, sz, sp)) return;
    const float x0 = sp.xw - 0.5f * sp.s, y0 = sp.yw - 0.5f * sp.s;
    // fragments whose centres f + 0.5 lie in [x0, x0 + s)
    int fx0 = (int)ceilf(x0 - 0.5f), fx1 = (int)ceilf(x0 + sp.s - 0.5f) - 1;
    int fy0 = (int)ceilf(y0 - 0.5f), fy1 = (int)ceilf(y0 + sp.s - 0.5f) - 1;
    fx0 = max(fx0, 0);
    fy0 = max(fy0, 0);
    fx1 = min(fx1, R.width - 1);
    fy1 = min(fy1, R.height - 1);
    const unsigned long long key =
        ((unsigned long long)__float_as_uint(sp.zw) << 32) | (unsigned long long)lid[i];
    for (int fy = fy0; fy <= fy1; ++fy)
        for (int fx = fx0; fx <= fx1; ++fx) {
            if (sprite_d2(sp, fx, fy) > 0.25f) continue;
            atomicMin(&zbuf[(size_t)(R.height - 1 - fy) * (size_t)R.width + (size_t)fx], key);
        }
}

__device__ __forceinline__ unsigned char to_u8(float c) {
    c = fminf(fmaxf(c, 0.0f), 1.0f);
    return (unsigned char)__float2int_rn(c * 255.0f);
}

__global__ void k_render_resolve(const float4* __restrict__ X, const float4* __restrict__ color,
                                 const float* __restrict__ size, const uint32_t* __restrict__ id,
                                 const uint32_t* __restrict__ phys, RenderParams R,
                                 const unsigned long long* __restrict__ zbuf,
                                 unsigned char* __restrict__ rgb) {
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= (uint32_t)(R.width * R.height)) return;
    const int row = (int)(pix / (uint32_t)R.width), fx = (int)(pix % (uint32_t)R.width);
    const int fy = R.height - 1 - row;
    float r = 0.15f, g = 0.15f, b = 0.18f;  // glClearColor (Application.cpp:72)
    bool floorHit = false;
    float zf = 1.0f;
    if (R.drawFloor) {
        // the ray through the pixel centre against the floor plane
        const float nx = ((float)fx + 0.5f) / (float)R.width * 2.0f - 1.0f;
        const float ny = ((float)fy + 0.5f) / (float)R.height * 2.0f - 1.0f;
        const float a = nx * R.tanHalf * R.aspect, c = ny * R.tanHalf;
        const float dx = (R.front[0] + a * R.right[0]) + c * R.up[0];
        const float dy = (R.front[1] + a * R.right[1]) + c * R.up[1];
        const float dz = (R.front[2] + a * R.right[2]) + c * R.up[2];
        if (dy != 0.0f) {
            const float t = (kFloorY - R.eye[1]) / dy;
            const float hx = R.eye[0] + t * dx, hz = R.eye[2] + t * dz;
            if (t > 0.0f && hx >= 0.0f && hx <= kFloorExtent && hz >= 0.0f && hz <= kFloorExtent) {
                const float* m = R.vp;
                const float cz = ((m[8] * hx + m[9] * kFloorY) + m[10] * hz) + m[11];
                const float cw = ((m[12] * hx + m[13] * kFloorY) + m[14] * hz) + m[15];
                if (cw > 0.0f && fabsf(cz) <= cw) {
                    floorHit = true;
                    zf = (cz / cw) * 0.5f + 0.5f;
                    const int check = ((int)floorf(hx) + (int)floorf(hz)) & 1;
                    r = g = b = check ? 0.5f : 0.8f;  // mix(color1, color2, check)
                }
            }
        }
    }
    const unsigned long long key = zbuf[pix];
    if (key != ~0ull) {
        const float zp = __uint_as_float((uint32_t)(key >> 32));
        if (!floorHit || zp < zf) {
            const uint32_t slot = phys[(uint32_t)(key & 0xFFFFFFFFull)];  // draw index -> slot
            const float4 x = X[slot];
            const float sz = R.pointSize > 0.0f ? R.pointSize : size[id[slot]];
            Sprite sp;
            if (sprite_of(R, x.x, x.y, x.z, sz, sp)) {
                const float d2 = sprite_d2(sp, fx, fy);
                const float alpha = expf(-3.0f * d2);
                const float4 col = color[id[slot]];
                r = col.x * alpha;
                g = col.y * alpha;
                b = col.z * alpha;
            }
        }
    }
    rgb[3 * (size_t)pix + 0] = to_u8(r);
    rgb[3 * (size_t)pix + 1] = to_u8(g);
    rgb[3 * (size_t)pix + 2] = to_u8(b);
}

}  // namespace

void camera_setup(const float position[3], float yawDeg, float pitchDeg, float fovYDeg, float nearZ,
                  float farZ, int width, int height, RenderParams& R) {
    const double d2r = 3.14159265358979323846 / 180.0;
    const double yaw = yawDeg * d2r, pitch = pitchDeg * d2r;
    // Camera::updateVectors (Camera.cpp:76-87), then glm::lookAt's basis
    double f[3] = {std::cos(yaw) * std::cos(pitch), std::sin(pitch), std::sin(yaw) * std::cos(pitch)};
    auto norm = [](double* v) {
        const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; ++k) v[k] /= l;
    };
    auto cross = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    norm(f);
    const double worldUp[3] = {0.0, 1.0, 0.0};
    double rt[3], up[3];
    cross(f, worldUp, rt);
    norm(rt);
    cross(rt, f, up);
    norm(up);
    const double eye[3] = {position[0], position[1], position[2]};
    // glm::lookAt(eye, eye + f, up): rows s, u, -f with translations
    double s[3], u[3];
    cross(f, up, s);
    norm(s);
    cross(s, f, u);
    double V[16] = {s[0], s[1], s[2], -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]),
                    u[0], u[1], u[2], -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]),
                    -f[0], -f[1], -f[2], (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]),
                    0.0, 0.0, 0.0, 1.0};
    // glm::perspective (right-handed, depth in [-1, 1])
    const double aspect = (double)width / (double)height;
    const double th = std::tan(fovYDeg * d2r / 2.0);
    const double n = nearZ, fz = farZ;
    double P[16] = {1.0 / (aspect * th), 0, 0, 0,
                    0, 1.0 / th, 0, 0,
                    0, 0, -(fz + n) / (fz - n), -(2.0 * fz * n) / (fz - n),
                    0, 0, -1.0, 0};
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += P[4 * r + k] * V[4 * k + c];
            R.vp[4 * r + c] = (float)acc;
        }
    for (int k = 0; k < 3; ++k) {
        R.eye[k] = (float)eye[k];
        R.front[k] = (float)f[k];
        R.right[k] = (float)s[k];
        R.up[k] = (float)u[k];
    }
    R.tanHalf = (float)th;
    R.aspect = (float)aspect;
    R.width = width;
    R.height = height;
}

void launch_render(hipStream_t st, const float4* X, const float4* color, const float* size,
                   const uint32_t* id, const uint32_t* lid, const uint32_t* phys, uint32_t n,
                   const RenderParams& R, unsigned long long* zbuf, unsigned char* rgb) {
    const uint32_t pixels = (uint32_t)(R.width * R.height);
    k_render_clear<<<(pixels + 255) / 256, 256, 0, st>>>(zbuf, pixels);
    if (n) k_render_splat<<<(n + 255) / 256, 256, 0, st>>>(X, color, size, id, lid, n, R, zbuf);
    k_render_resolve<<<(pixels + 255) / 256, 256, 0, st>>>(X, color, size, id, phys, R, zbuf, rgb);
}

}  // namespace akua
