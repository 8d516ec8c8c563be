// scene_io.cpp -- deterministic scenes and trajectory files (include/akua_io.h).  Host code only:
// nothing here launches a kernel, except that akua_traj_append_solver reads a solver's state
// through akua_solver_get_state.
//
// The scenes replace Application::prepareDamBreak (Application.cpp:126-204) for a headless run.
// Every position is formed with the same FP32 operations as the Python generators in
// akuaengine_amd/scenes.py (origin + index * spacing, then + (2u - 1) * jitter), so the two
// produce identical bits; the library is compiled with -ffp-contract=off, so no product is fused
// into the following sum.
#include <sys/types.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/akua_io.h"
#include "last_error.h"

namespace {

struct IoError : std::runtime_error {
    akua_status code;
    IoError(akua_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void io_fail(akua_status c, const std::string& m) { throw IoError(c, m); }

template <class F>
akua_status io_guarded(F&& f) {
    try {
        f();
        akua::set_last_error("");
        return AKUA_OK;
    } catch (const IoError& e) {
        akua::set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        akua::set_last_error("host allocation failed");
        return AKUA_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        akua::set_last_error(e.what());
        return AKUA_ERR_INVALID_ARG;
    }
}

// ---- scenes ---------------------------------------------------------------------------------

// SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number generators", 2014):
// output k (k = 1, 2, ...) of the stream seeded with `seed`.
uint64_t splitmix64(uint64_t seed, uint64_t k) {
    uint64_t z = seed + k * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Lattice {
    int nx, ny, nz;
    float spacing;
    float origin[3];
    float jitter;    // 0: none
    uint64_t seed;
    int64_t count() const { return (int64_t)nx * ny * nz; }
};

// Positions of a lattice, x-major then y then z (Application.cpp:170-176), written at pos.
void fill_lattice(const Lattice& L, float* pos) {
    int64_t p = 0;
    const float two = 2.0f, one = 1.0f;
    for (int ix = 0; ix < L.nx; ++ix)
        for (int iy = 0; iy < L.ny; ++iy)
            for (int iz = 0; iz < L.nz; ++iz, ++p) {
                const float idx[3] = {(float)ix, (float)iy, (float)iz};
                for (int a = 0; a < 3; ++a) {
                    float x = L.origin[a] + idx[a] * L.spacing;
                    if (L.jitter > 0.0f) {
                        // u: the top 24 bits of output 3p + a + 1, times 2^-24 (exact in FP32)
                        const uint64_t z = splitmix64(L.seed, (uint64_t)(3 * p + a + 1));
                        const float u = (float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
                        x = x + (u * two - one) * L.jitter;
                    }
                    pos[3 * p + a] = x;
                }
            }
}

struct SceneSpec {
    std::vector<Lattice> parts;
    akua_scene_info info;
};

constexpr float kDamBreakDt = 0.0083f;
constexpr float kAppDt = 0.016f;      // Application::_deltaTime (Application.h:32)
constexpr float kJitter = 0.005f;

SceneSpec make_spec(const char* nameC, float spacing, int64_t seed) {
    if (!nameC) io_fail(AKUA_ERR_INVALID_ARG, "NULL scene name");
    const std::string name(nameC);
    SceneSpec S{};
    akua_scene_info& I = S.info;
    I.deltaTime = kDamBreakDt;
    I.referenceParams = 0;
    auto box = [&](float x0, float y0, float z0, float x1, float y1, float z1) {
        I.boxMin[0] = x0; I.boxMin[1] = y0; I.boxMin[2] = z0;
        I.boxMax[0] = x1; I.boxMax[1] = y1; I.boxMax[2] = z1;
    };
    if (name == "app_dam_break") {
        // Application.cpp:159-176: numX = int((maxPos - minPos).x / particle_spacing), FP32
        const float s = spacing > 0.0f ? spacing : 0.05f;
        const float minPos[3] = {2.0f, 1.0f, 2.0f}, maxPos[3] = {3.5f, 2.5f, 3.5f};
        int n[3];
        for (int a = 0; a < 3; ++a) {
            const float q = (maxPos[a] - minPos[a]) / s;
            if (!(q >= 1.0f && q < 1024.0f)) io_fail(AKUA_ERR_INVALID_ARG, "app_dam_break: spacing gives no lattice of 1..1023 per axis");
            n[a] = (int)q;
        }
        S.parts.push_back(Lattice{n[0], n[1], n[2], s, {minPos[0], minPos[1], minPos[2]}, 0.0f, 0});
        box(1.5f, 0.0f, 1.5f, 4.5f, 4.0f, 4.5f);  // BOX_MIN, BOX_MAX (Application.cpp:14-15)
        I.deltaTime = kAppDt;
        I.referenceParams = 1;
    } else if (name == "dam_break_27k") {
        S.parts.push_back(Lattice{30, 30, 30, 0.05f, {0.025f, 0.025f, 0.025f}, 0.0f, 0});
        box(0.0f, 0.0f, 0.0f, 2.0f, 1.5f, 1.0f);
    } else if (name == "dam_break_1m") {
        const uint64_t sd = seed < 0 ? 1u : (uint64_t)seed;
        S.parts.push_back(Lattice{100, 100, 100, 0.05f, {0.025f, 0.025f, 0.025f}, kJitter, sd});
        box(0.0f, 0.0f, 0.0f, 10.0f, 6.0f, 5.0f);
    } else if (name == "dam_break_8m") {
        const uint64_t sd = seed < 0 ? 2u : (uint64_t)seed;
        S.parts.push_back(Lattice{200, 200, 100, 0.05f, {0.025f, 0.025f, 0.025f}, kJitter, sd});
        S.parts.push_back(Lattice{200, 200, 100, 0.05f, {20.025f, 0.025f, 0.025f}, kJitter, sd + 1000u});
        box(0.0f, 0.0f, 0.0f, 30.0f, 12.0f, 5.0f);
    } else if (name.rfind("cube:", 0) == 0) {
        char* end = nullptr;
        const long k = std::strtol(name.c_str() + 5, &end, 10);
        if (end == name.c_str() + 5 || *end != '\0' || k < 1 || k > 1000)
            io_fail(AKUA_ERR_INVALID_ARG, "cube:<k> needs an integer k in 1..1000");
        const uint64_t sd = seed < 0 ? 1u : (uint64_t)seed;
        S.parts.push_back(Lattice{(int)k, (int)k, (int)k, 0.05f, {0.025f, 0.025f, 0.025f}, kJitter, sd});
        const double e = (double)k * 0.05;
        box(0.0f, 0.0f, 0.0f, (float)(2.0 * e), (float)(1.2 * e), (float)e);
    } else {
        io_fail(AKUA_ERR_INVALID_ARG, "unknown scene '" + name +
                                          "' (app_dam_break, dam_break_27k, dam_break_1m, dam_break_8m, cube:<k>)");
    }
    int64_t n = 0;
    for (const Lattice& L : S.parts) n += L.count();
    if (n > 0x7FFFFFFF) io_fail(AKUA_ERR_INVALID_ARG, "scene too large");
    I.numParticles = (int32_t)n;
    return S;
}

// ---- trajectory files -----------------------------------------------------------------------

struct TrajHeader {
    char magic[8];
    uint32_t version;
    uint32_t headerBytes;
    int32_t n;
    uint32_t fields;
    float dt;
    float boxMin[3];
    float boxMax[3];
    float gravity[3];
    uint64_t frameBytes;
    uint8_t zero[AKUA_TRAJ_HEADER_BYTES - 72];
};
static_assert(sizeof(TrajHeader) == AKUA_TRAJ_HEADER_BYTES, "header is 128 bytes");
static_assert(offsetof(TrajHeader, fields) == 20 && offsetof(TrajHeader, dt) == 24 &&
                  offsetof(TrajHeader, boxMin) == 28 && offsetof(TrajHeader, boxMax) == 40 &&
                  offsetof(TrajHeader, gravity) == 52 && offsetof(TrajHeader, frameBytes) == 64,
              "header layout");

constexpr char kMagic[8] = {'A', 'K', 'U', 'A', 'T', 'R', 'J', '\0'};
constexpr int kFrameHead = 16;  // int64 step, double time

// Bytes per particle of each field, in bit order.
constexpr int kFieldBytes[6] = {12, 12, 4, 4, 4, 4};

uint64_t frame_bytes(int32_t n, uint32_t fields) {
    uint64_t b = kFrameHead;
    for (int f = 0; f < 6; ++f)
        if (fields & (1u << f)) b += (uint64_t)kFieldBytes[f] * (uint64_t)n;
    return b;
}

}  // namespace

struct akua_traj {
    FILE* f = nullptr;
    bool writing = false;
    TrajHeader h{};
    int64_t frames = 0;
    int32_t partial = 0;
    std::vector<float> pos, vel, mass, dens, lam;  // staging of append_solver
    std::vector<uint32_t> ids;
};

namespace {

void write_all(akua_traj* t, const void* p, size_t bytes) {
    if (bytes && std::fwrite(p, 1, bytes, t->f) != bytes)
        io_fail(AKUA_ERR_INVALID_ARG, "trajectory write failed (disk full or file closed?)");
}

void read_all(akua_traj* t, void* p, size_t bytes) {
    if (bytes && std::fread(p, 1, bytes, t->f) != bytes) io_fail(AKUA_ERR_INVALID_ARG, "trajectory read failed");
}

}  // namespace

extern "C" {

akua_status akua_scene_describe(const char* name, float spacing, int64_t seed, akua_scene_info* info) {
    return io_guarded([&] {
        if (!info) io_fail(AKUA_ERR_INVALID_ARG, "NULL info");
        *info = make_spec(name, spacing, seed).info;
    });
}

akua_status akua_scene_generate(const char* name, float spacing, int64_t seed, float* pos, int32_t n) {
    return io_guarded([&] {
        const SceneSpec S = make_spec(name, spacing, seed);
        if (!pos) io_fail(AKUA_ERR_INVALID_ARG, "NULL pos");
        if (n != S.info.numParticles)
            io_fail(AKUA_ERR_INVALID_ARG, "n = " + std::to_string(n) + " but the scene has " +
                                              std::to_string(S.info.numParticles) + " particles");
        int64_t off = 0;
        for (const Lattice& L : S.parts) {
            fill_lattice(L, pos + 3 * off);
            off += L.count();
        }
    });
}

akua_status akua_traj_create(const char* path, int32_t n, uint32_t fields, float deltaTime,
                             const float boxMin[3], const float boxMax[3], const float gravity[3],
                             akua_traj** out) {
    return io_guarded([&] {
        if (!path || !out) io_fail(AKUA_ERR_INVALID_ARG, "NULL path or out");
        *out = nullptr;
        if (n < 1) io_fail(AKUA_ERR_INVALID_ARG, "a trajectory needs at least one particle");
        if (fields == 0 || (fields & ~(uint32_t)AKUA_FIELD_ALL))
            io_fail(AKUA_ERR_INVALID_ARG, "fields must be a nonzero combination of AKUA_FIELD_* bits");
        auto t = new akua_traj();
        t->writing = true;
        std::memcpy(t->h.magic, kMagic, 8);
        t->h.version = AKUA_TRAJ_VERSION;
        t->h.headerBytes = AKUA_TRAJ_HEADER_BYTES;
        t->h.n = n;
        t->h.fields = fields;
        t->h.dt = deltaTime;
        for (int a = 0; a < 3; ++a) {
            t->h.boxMin[a] = boxMin ? boxMin[a] : 0.0f;
            t->h.boxMax[a] = boxMax ? boxMax[a] : 0.0f;
            t->h.gravity[a] = gravity ? gravity[a] : 0.0f;
        }
        t->h.frameBytes = frame_bytes(n, fields);
        t->f = std::fopen(path, "wb");
        if (!t->f) {
            delete t;
            io_fail(AKUA_ERR_INVALID_ARG, std::string("cannot create ") + path);
        }
        try {
            write_all(t, &t->h, sizeof(TrajHeader));
        } catch (...) {
            std::fclose(t->f);
            delete t;
            throw;
        }
        *out = t;
    });
}

akua_status akua_traj_append(akua_traj* t, int64_t step, double time, const float* pos,
                             const float* vel, const float* mass, const float* density,
                             const float* lambda, const uint32_t* ids) {
    return io_guarded([&] {
        if (!t || !t->f || !t->writing) io_fail(AKUA_ERR_STATE, "not a trajectory open for writing");
        const void* arr[6] = {pos, vel, mass, density, lambda, ids};
        static const char* names[6] = {"position", "velocity", "mass", "density", "lambda", "id"};
        for (int f = 0; f < 6; ++f)
            if ((t->h.fields & (1u << f)) && !arr[f])
                io_fail(AKUA_ERR_INVALID_ARG, std::string("the file holds ") + names[f] + " but it is NULL");
        write_all(t, &step, 8);
        write_all(t, &time, 8);
        for (int f = 0; f < 6; ++f)
            if (t->h.fields & (1u << f)) write_all(t, arr[f], (size_t)kFieldBytes[f] * (size_t)t->h.n);
        t->frames += 1;
    });
}

akua_status akua_traj_append_solver(akua_traj* t, akua_solver* s, int64_t step, double time,
                                    const uint32_t* idMap) {
    if (!t || !t->f || !t->writing) {
        akua::set_last_error("not a trajectory open for writing");
        return AKUA_ERR_STATE;
    }
    const int32_t n = t->h.n;
    const uint32_t F = t->h.fields;
    try {
        if (F & AKUA_FIELD_POSITION) t->pos.resize(3 * (size_t)n);
        if (F & AKUA_FIELD_VELOCITY) t->vel.resize(3 * (size_t)n);
        if (F & AKUA_FIELD_MASS) t->mass.resize(n);
        if (F & AKUA_FIELD_DENSITY) t->dens.resize(n);
        if (F & AKUA_FIELD_LAMBDA) t->lam.resize(n);
        if (F & AKUA_FIELD_ID) t->ids.resize(n);
    } catch (const std::bad_alloc&) {
        akua::set_last_error("host allocation failed");
        return AKUA_ERR_OUT_OF_MEMORY;
    }
    auto ptr = [&](std::vector<float>& v, uint32_t bit) { return (F & bit) ? v.data() : nullptr; };
    akua_status st = akua_solver_get_state(s, ptr(t->pos, AKUA_FIELD_POSITION), ptr(t->vel, AKUA_FIELD_VELOCITY),
                                           ptr(t->mass, AKUA_FIELD_MASS), ptr(t->dens, AKUA_FIELD_DENSITY),
                                           ptr(t->lam, AKUA_FIELD_LAMBDA),
                                           (F & AKUA_FIELD_ID) ? t->ids.data() : nullptr, n);
    if (st != AKUA_OK) return st;
    if (idMap && (F & AKUA_FIELD_ID)) {
        for (int32_t k = 0; k < n; ++k) {
            if (t->ids[k] >= (uint32_t)n) {
                akua::set_last_error("solver returned an id outside 0..n-1");
                return AKUA_ERR_STATE;
            }
            t->ids[k] = idMap[t->ids[k]];
        }
    }
    return akua_traj_append(t, step, time, t->pos.data(), t->vel.data(), t->mass.data(), t->dens.data(),
                            t->lam.data(), t->ids.data());
}

akua_status akua_traj_open(const char* path, akua_traj_info* info, akua_traj** out) {
    return io_guarded([&] {
        if (!path || !out) io_fail(AKUA_ERR_INVALID_ARG, "NULL path or out");
        *out = nullptr;
        FILE* f = std::fopen(path, "rb");
        if (!f) io_fail(AKUA_ERR_INVALID_ARG, std::string("cannot open ") + path);
        auto t = new akua_traj();
        t->f = f;
        try {
            if (std::fread(&t->h, 1, sizeof(TrajHeader), f) != sizeof(TrajHeader))
                io_fail(AKUA_ERR_INVALID_ARG, "not a trajectory file: shorter than its header");
            if (std::memcmp(t->h.magic, kMagic, 8) != 0) io_fail(AKUA_ERR_INVALID_ARG, "not a trajectory file: bad magic");
            if (t->h.version != AKUA_TRAJ_VERSION)
                io_fail(AKUA_ERR_UNSUPPORTED, "trajectory version " + std::to_string(t->h.version) +
                                                  " (this library reads version " +
                                                  std::to_string(AKUA_TRAJ_VERSION) + ")");
            if (t->h.headerBytes != AKUA_TRAJ_HEADER_BYTES || t->h.n < 1 || t->h.fields == 0 ||
                (t->h.fields & ~(uint32_t)AKUA_FIELD_ALL) || t->h.frameBytes != frame_bytes(t->h.n, t->h.fields))
                io_fail(AKUA_ERR_INVALID_ARG, "inconsistent trajectory header");
            if (fseeko(f, 0, SEEK_END) != 0) io_fail(AKUA_ERR_INVALID_ARG, "cannot seek");
            const off_t size = ftello(f);
            const uint64_t body = (uint64_t)size - AKUA_TRAJ_HEADER_BYTES;
            t->frames = (int64_t)(body / t->h.frameBytes);
            t->partial = (body % t->h.frameBytes) != 0 ? 1 : 0;
        } catch (...) {
            std::fclose(f);
            delete t;
            throw;
        }
        if (info) {
            info->version = t->h.version;
            info->numParticles = t->h.n;
            info->fields = t->h.fields;
            info->deltaTime = t->h.dt;
            for (int a = 0; a < 3; ++a) {
                info->boxMin[a] = t->h.boxMin[a];
                info->boxMax[a] = t->h.boxMax[a];
                info->gravity[a] = t->h.gravity[a];
            }
            info->frameBytes = t->h.frameBytes;
            info->frames = t->frames;
            info->partialFrame = t->partial;
        }
        *out = t;
    });
}

akua_status akua_traj_read(akua_traj* t, int64_t frame, int64_t* step, double* time, float* pos,
                           float* vel, float* mass, float* density, float* lambda, uint32_t* ids) {
    return io_guarded([&] {
        if (!t || !t->f || t->writing) io_fail(AKUA_ERR_STATE, "not a trajectory open for reading");
        const int64_t k = frame < 0 ? t->frames + frame : frame;
        if (k < 0 || k >= t->frames)
            io_fail(AKUA_ERR_INVALID_ARG, "frame " + std::to_string(frame) + " of " + std::to_string(t->frames));
        void* arr[6] = {pos, vel, mass, density, lambda, ids};
        static const char* names[6] = {"position", "velocity", "mass", "density", "lambda", "id"};
        for (int f = 0; f < 6; ++f)
            if (arr[f] && !(t->h.fields & (1u << f)))
                io_fail(AKUA_ERR_INVALID_ARG, std::string("the file holds no ") + names[f]);
        const off_t at = (off_t)AKUA_TRAJ_HEADER_BYTES + (off_t)k * (off_t)t->h.frameBytes;
        if (fseeko(t->f, at, SEEK_SET) != 0) io_fail(AKUA_ERR_INVALID_ARG, "cannot seek");
        int64_t st = 0;
        double tm = 0.0;
        read_all(t, &st, 8);
        read_all(t, &tm, 8);
        if (step) *step = st;
        if (time) *time = tm;
        for (int f = 0; f < 6; ++f) {
            if (!(t->h.fields & (1u << f))) continue;
            const size_t bytes = (size_t)kFieldBytes[f] * (size_t)t->h.n;
            if (arr[f]) {
                read_all(t, arr[f], bytes);
            } else if (fseeko(t->f, (off_t)bytes, SEEK_CUR) != 0) {
                io_fail(AKUA_ERR_INVALID_ARG, "cannot seek");
            }
        }
    });
}

akua_status akua_traj_close(akua_traj* t) {
    return io_guarded([&] {
        if (!t) io_fail(AKUA_ERR_INVALID_ARG, "NULL trajectory");
        const bool ok = !t->f || std::fclose(t->f) == 0;
        const bool wrote = t->writing;
        delete t;
        if (!ok && wrote) io_fail(AKUA_ERR_INVALID_ARG, "trajectory close failed (data may be lost)");
    });
}

}  // extern "C"
