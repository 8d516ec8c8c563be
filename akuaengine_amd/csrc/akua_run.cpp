// akua_run -- headless driver of libakua.so: the counterpart of the reference application's
// scene set-up and fixed-step loop (Application::prepareDamBreak, Application.cpp:126-204;
// Application::run, Application.cpp:37-70) without a window.  It generates a scene (or resumes
// from a trajectory frame), steps it with the Application's accumulator rule, writes frames to
// a trajectory file and prints one JSON line with the throughput and the final density error.
//
// The loop: every frame adds the frame time to an FP32 accumulator, then runs steps while the
// accumulator holds at least one time step and fewer than --max-steps-per-frame steps ran in
// the frame (MAX_STEPS_PER_FRAME = 3 in the reference, Application.cpp:20, :63-70).  A headless
// frame lasts exactly --frame-time seconds, so the schedule is deterministic.  --steps N is N
// frames of one time step each: one step per frame.
//
// Built by akuaengine_amd/build.py with g++ against include/akua.h and include/akua_io.h only.
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "akua.h"
#include "akua_io.h"

namespace {

const char* kUsage =
    "usage: akua_run [options]\n"
    "  --scene NAME            app_dam_break (default), dam_break_27k, dam_break_1m, dam_break_8m,\n"
    "                          cube:<k>\n"
    "  --seed S                jitter seed (default: the scene's own)\n"
    "  --init FILE[@FRAME]     start from a trajectory frame (default the last) instead of a scene\n"
    "  --params reference|dam_break\n"
    "                          parameter set (default: reference for app_dam_break, else dam_break)\n"
    "  --spacing S             particle_spacing (also the app_dam_break lattice spacing)\n"
    "  --iterations I          solver iterations (default 4)\n"
    "  --dt DT                 time step (default: the scene's, or the trajectory's)\n"
    "  --steps N               N steps, one per frame\n"
    "  --frames F              F frames of the Application loop ...\n"
    "  --frame-time T          ... each T seconds long (default: dt)\n"
    "  --max-steps-per-frame M at most M steps per frame (default 3; 1 with --steps)\n"
    "  --out FILE              write a trajectory: the initial state, then every --every steps\n"
    "  --every K               (default 1)\n"
    "  --fields LIST           pos,vel,mass,density,lambda,id or all (default pos,vel,density,id)\n"
    "  --render-every K        draw the particles every K steps (and the initial state) as the\n"
    "                          reference's window would show them, as PPM images ...\n"
    "  --render-dir DIR        ... named DIR/frame_<step>.ppm (default: .)\n"
    "  --width W --height H    image size (default 1280 x 720, the reference's window)\n"
    "  --point-size S          sprite size (default 50, the application's particles)\n"
    "  --boundary L            L layers of wall samples (0 = off, the reference's walls; 1..4)\n"
    "  --boundary-spacing S    their spacing (default: particle_spacing)\n"
    "  --autotune              pick launch and layout settings before the run (results unchanged)\n"
    "  --tuning LIST           set them instead: block,level,hashBits,layoutBits,interval,\n"
    "                          balanceTile,slabBatch,fuseSweep,laneGroup,workQueue (akua_tuning's\n"
    "                          fields in order; missing trailing fields are 0)\n"
    "  --network FILE.npy      learned warm start with this parameter blob (float32 .npy)\n"
    "  --learned-iterations J  (default 1)\n"
    "  --gate-ratio R          fallback when r(x*_w) > R r(x*_pred) + tolerance (default 1)\n"
    "  --gate-tolerance E      (default 0.01)\n"
    "  --no-graph              launch kernels one by one instead of one graph per step\n"
    "  --dry-run               generate the scene, write the initial frame, print the step plan,\n"
    "                          and stop before touching a GPU\n";

struct Options {
    std::string scene = "app_dam_break";
    int64_t seed = -1;
    std::string init;
    std::string params;
    float spacing = 0.0f;
    int iterations = 4;
    float dt = 0.0f;
    int64_t steps = -1;
    int64_t frames = -1;
    float frameTime = 0.0f;
    int maxPerFrame = -1;
    std::string out;
    int64_t every = 1;
    uint32_t fields = AKUA_FIELD_POSITION | AKUA_FIELD_VELOCITY | AKUA_FIELD_DENSITY | AKUA_FIELD_ID;
    bool autotune = false;
    std::vector<int32_t> tuning;  // --tuning fields, in akua_tuning's order
    int64_t renderEvery = 0;
    std::string renderDir = ".";
    int width = 1280, height = 720;
    float pointSize = 50.0f;
    int boundary = 0;
    float boundarySpacing = 0.0f;
    std::string network;
    int learnedIters = 1;
    float gateRatio = 1.0f;
    float gateTol = 0.01f;
    bool graph = true;
    bool dryRun = false;
};

[[noreturn]] void usage_error(const std::string& m) {
    std::fprintf(stderr, "akua_run: %s\n%s", m.c_str(), kUsage);
    std::exit(2);
}

[[noreturn]] void fatal(const std::string& what, akua_status st) {
    std::fprintf(stderr, "akua_run: %s failed (status %d): %s\n", what.c_str(), (int)st, akua_last_error());
    std::exit(1);
}

void check(akua_status st, const char* what) {
    if (st != AKUA_OK) fatal(what, st);
}

double parse_num(const std::string& flag, const char* v) {
    char* end = nullptr;
    const double x = std::strtod(v, &end);
    if (end == v || *end != '\0' || !std::isfinite(x)) usage_error(flag + " needs a number, got '" + v + "'");
    return x;
}

int64_t parse_int(const std::string& flag, const char* v, int64_t lo, int64_t hi) {
    char* end = nullptr;
    const long long x = std::strtoll(v, &end, 10);
    if (end == v || *end != '\0' || x < lo || x > hi)
        usage_error(flag + " needs an integer in " + std::to_string(lo) + ".." + std::to_string(hi) + ", got '" + v + "'");
    return x;
}

uint32_t parse_fields(const std::string& v) {
    if (v == "all") return AKUA_FIELD_ALL;
    uint32_t f = 0;
    std::stringstream ss(v);
    std::string item;
    while (std::getline(ss, item, ',')) {
        if (item == "pos") f |= AKUA_FIELD_POSITION;
        else if (item == "vel") f |= AKUA_FIELD_VELOCITY;
        else if (item == "mass") f |= AKUA_FIELD_MASS;
        else if (item == "density") f |= AKUA_FIELD_DENSITY;
        else if (item == "lambda") f |= AKUA_FIELD_LAMBDA;
        else if (item == "id") f |= AKUA_FIELD_ID;
        else usage_error("unknown field '" + item + "'");
    }
    if (!f) usage_error("--fields is empty");
    return f;
}

Options parse(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) usage_error(a + " needs a value");
            return argv[++i];
        };
        if (a == "-h" || a == "--help") {
            std::fputs(kUsage, stdout);
            std::exit(0);
        } else if (a == "--scene") o.scene = val();
        else if (a == "--seed") o.seed = parse_int(a, val(), 0, INT64_MAX);
        else if (a == "--init") o.init = val();
        else if (a == "--params") {
            o.params = val();
            if (o.params != "reference" && o.params != "dam_break") usage_error("--params is reference or dam_break");
        } else if (a == "--spacing") o.spacing = (float)parse_num(a, val());
        else if (a == "--iterations") o.iterations = (int)parse_int(a, val(), 0, 64);
        else if (a == "--dt") o.dt = (float)parse_num(a, val());
        else if (a == "--steps") o.steps = parse_int(a, val(), 0, INT32_MAX);
        else if (a == "--frames") o.frames = parse_int(a, val(), 0, INT32_MAX);
        else if (a == "--frame-time") o.frameTime = (float)parse_num(a, val());
        else if (a == "--max-steps-per-frame") o.maxPerFrame = (int)parse_int(a, val(), 1, 1000000);
        else if (a == "--out") o.out = val();
        else if (a == "--every") o.every = parse_int(a, val(), 1, INT32_MAX);
        else if (a == "--fields") o.fields = parse_fields(val());
        else if (a == "--autotune") o.autotune = true;
        else if (a == "--tuning") {
            std::string list = val();
            size_t pos = 0;
            while (pos <= list.size()) {
                const size_t comma = list.find(',', pos);
                const std::string item = list.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
                o.tuning.push_back((int32_t)parse_int("--tuning", item.c_str(), 0, 1 << 20));
                if (comma == std::string::npos) break;
                pos = comma + 1;
            }
            if (o.tuning.size() < 5 || o.tuning.size() > 10) usage_error("--tuning takes 5 to 10 values");
        }
        else if (a == "--render-every") o.renderEvery = parse_int(a, val(), 1, INT32_MAX);
        else if (a == "--render-dir") o.renderDir = val();
        else if (a == "--width") o.width = (int)parse_int(a, val(), 1, 16384);
        else if (a == "--height") o.height = (int)parse_int(a, val(), 1, 16384);
        else if (a == "--point-size") o.pointSize = (float)parse_num(a, val());
        else if (a == "--boundary") o.boundary = (int)parse_int(a, val(), 0, 4);
        else if (a == "--boundary-spacing") o.boundarySpacing = (float)parse_num(a, val());
        else if (a == "--network") o.network = val();
        else if (a == "--learned-iterations") o.learnedIters = (int)parse_int(a, val(), 1, 64);
        else if (a == "--gate-ratio") o.gateRatio = (float)parse_num(a, val());
        else if (a == "--gate-tolerance") o.gateTol = (float)parse_num(a, val());
        else if (a == "--no-graph") o.graph = false;
        else if (a == "--dry-run") o.dryRun = true;
        else usage_error("unknown option '" + a + "'");
    }
    if (o.steps >= 0 && o.frames >= 0) usage_error("give --steps or --frames, not both");
    if (o.steps >= 0 && (o.frameTime != 0.0f || o.maxPerFrame > 0))
        usage_error("--frame-time and --max-steps-per-frame go with --frames");
    if (o.steps < 0 && o.frames < 0) o.steps = 0;
    if (o.spacing < 0.0f) usage_error("--spacing must be positive");
    if (o.autotune && !o.tuning.empty()) usage_error("give --autotune or --tuning, not both");
    return o;
}

// A float32 little-endian one-dimensional .npy array (NumPy format 1.0-3.0).
std::vector<float> load_npy_f32(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) usage_error("cannot open " + path);
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) usage_error(path + " is not a .npy file");
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        uint16_t h16 = 0;
        f.read(reinterpret_cast<char*>(&h16), 2);
        hlen = h16;
    } else {
        f.read(reinterpret_cast<char*>(&hlen), 4);
    }
    std::string header(hlen, '\0');
    f.read(&header[0], hlen);
    if (!f) usage_error(path + ": truncated header");
    if (header.find("'descr': '<f4'") == std::string::npos || header.find("'fortran_order': False") == std::string::npos)
        usage_error(path + ": expected a little-endian float32 C-order array");
    const size_t sp = header.find("'shape': (");
    if (sp == std::string::npos) usage_error(path + ": no shape");
    const long long count = std::strtoll(header.c_str() + sp + 10, nullptr, 10);
    const size_t close = header.find(')', sp);
    const std::string shape = header.substr(sp + 10, close - (sp + 10));
    if (count <= 0 || shape.find(',') != shape.rfind(',') || (shape.find(',') != std::string::npos &&
                                                               shape.find_first_not_of(" ", shape.find(',') + 1) != std::string::npos))
        usage_error(path + ": expected a one-dimensional array");
    std::vector<float> v((size_t)count);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(count * 4));
    if (!f) usage_error(path + ": truncated data");
    return v;
}

// Renders the solver's particles with the reference's camera and writes a binary PPM.
double render_frame(akua_solver* s, const Options& o, int64_t step, std::vector<uint8_t>& img) {
    const auto t0 = std::chrono::steady_clock::now();
    akua_camera cam;
    akua_default_camera(&cam);
    img.resize((size_t)o.width * (size_t)o.height * 3);
    const akua_status st = akua_render(s, &cam, o.width, o.height, o.pointSize, 1, img.data());
    if (st != AKUA_OK) fatal("rendering", st);
    char name[64];
    std::snprintf(name, sizeof(name), "/frame_%08lld.ppm", (long long)step);
    const std::string path = o.renderDir + name;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "akua_run: cannot write %s\n", path.c_str());
        std::exit(1);
    }
    std::fprintf(f, "P6\n%d %d\n255\n", o.width, o.height);
    const bool ok = std::fwrite(img.data(), 1, img.size(), f) == img.size();
    if (std::fclose(f) != 0 || !ok) {
        std::fprintf(stderr, "akua_run: writing %s failed\n", path.c_str());
        std::exit(1);
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

std::string json_str(const std::string& s) {
    std::string r = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') r += '\\';
        if ((unsigned char)c < 0x20) continue;
        r += c;
    }
    return r + "\"";
}

}  // namespace

int main(int argc, char** argv) {
    Options o = parse(argc, argv);

    // ---- the initial state (host only) ---------------------------------------------------
    akua_pbf_config cfg;
    akua_lambda_corr corr;
    akua_default_config(&cfg, &corr);
    int32_t n = 0;
    float boxMin[3], boxMax[3], dt = 0.0f;
    bool referenceParams = false;
    std::vector<float> pos, vel, mass, dens, lam;
    std::vector<uint32_t> fileIds;
    int64_t step0 = 0;
    double time0 = 0.0;
    std::string source;
    if (!o.init.empty()) {
        std::string path = o.init;
        int64_t frame = -1;
        const size_t at = path.rfind('@');
        if (at != std::string::npos) {
            frame = parse_int("--init", path.c_str() + at + 1, INT64_MIN / 2, INT64_MAX / 2);
            path = path.substr(0, at);
        }
        akua_traj* t = nullptr;
        akua_traj_info info;
        check(akua_traj_open(path.c_str(), &info, &t), "opening the --init trajectory");
        if (!(info.fields & AKUA_FIELD_POSITION)) usage_error(path + " holds no positions");
        if (info.frames < 1) usage_error(path + " holds no whole frame");
        n = info.numParticles;
        pos.resize(3 * (size_t)n);
        vel.assign(3 * (size_t)n, 0.0f);
        mass.assign(n, 1.0f);
        dens.assign(n, 0.0f);
        lam.assign(n, 0.0f);
        if (info.fields & AKUA_FIELD_ID) fileIds.resize(n);
        check(akua_traj_read(t, frame, &step0, &time0, pos.data(),
                             (info.fields & AKUA_FIELD_VELOCITY) ? vel.data() : nullptr,
                             (info.fields & AKUA_FIELD_MASS) ? mass.data() : nullptr,
                             (info.fields & AKUA_FIELD_DENSITY) ? dens.data() : nullptr,
                             (info.fields & AKUA_FIELD_LAMBDA) ? lam.data() : nullptr,
                             (info.fields & AKUA_FIELD_ID) ? fileIds.data() : nullptr),
              "reading the --init frame");
        check(akua_traj_close(t), "closing the --init trajectory");
        for (int a = 0; a < 3; ++a) {
            boxMin[a] = info.boxMin[a];
            boxMax[a] = info.boxMax[a];
            cfg.gravity[a] = info.gravity[a];
        }
        dt = info.deltaTime;
        source = path + "@" + std::to_string(step0);
        if (!fileIds.empty()) {
            std::vector<char> seen(n, 0);
            for (uint32_t id : fileIds) {
                if (id >= (uint32_t)n || seen[id]) usage_error(path + ": ids are not a permutation of 0..n-1");
                seen[id] = 1;
            }
        }
    } else {
        akua_scene_info info;
        check(akua_scene_describe(o.scene.c_str(), o.spacing, o.seed, &info), "describing the scene");
        n = info.numParticles;
        pos.resize(3 * (size_t)n);
        check(akua_scene_generate(o.scene.c_str(), o.spacing, o.seed, pos.data(), n), "generating the scene");
        vel.assign(3 * (size_t)n, 0.0f);
        mass.assign(n, 1.0f);
        dens.assign(n, 0.0f);
        lam.assign(n, 0.0f);
        for (int a = 0; a < 3; ++a) {
            boxMin[a] = info.boxMin[a];
            boxMax[a] = info.boxMax[a];
        }
        dt = info.deltaTime;
        referenceParams = info.referenceParams != 0;
        source = o.scene;
    }
    if (!o.params.empty()) referenceParams = o.params == "reference";
    if (!referenceParams) {
        // the dam-break parameter set (rest density 6378, s_corr k 0.1 and delta_q 0.2 h)
        cfg.restDensity = 6378.0f;
        corr.k = 0.1f;
        corr.n = 4.0f;
        corr.delta_q = (float)(0.2 * 0.1);
    }
    if (o.spacing > 0.0f) cfg.particle_spacing = o.spacing;
    cfg.solverIterations = o.iterations;
    if (o.dt != 0.0f) dt = o.dt;
    if (!(dt > 0.0f)) usage_error("the time step must be positive (--dt)");

    // ---- the step plan: the Application's accumulator rule ---------------------------------
    int64_t frames = o.frames;
    float frameTime = o.frameTime;
    int maxPerFrame = o.maxPerFrame;
    if (o.steps >= 0) {
        frames = o.steps;
        frameTime = dt;
        maxPerFrame = 1;
    } else {
        if (frameTime == 0.0f) frameTime = dt;
        if (maxPerFrame < 0) maxPerFrame = 3;
        if (!(frameTime > 0.0f)) usage_error("--frame-time must be positive");
    }
    std::vector<int> plan((size_t)frames);
    int64_t totalSteps = 0;
    {
        float accumulator = 0.0f;
        for (int64_t f = 0; f < frames; ++f) {
            accumulator += frameTime;
            int k = 0;
            while (accumulator >= dt && k < maxPerFrame) {
                accumulator -= dt;
                ++k;
            }
            plan[(size_t)f] = k;
            totalSteps += k;
        }
    }

    // ---- the initial frame ------------------------------------------------------------------
    akua_traj* out = nullptr;
    int64_t written = 0;
    std::vector<uint32_t> ids0(n);
    for (int32_t k = 0; k < n; ++k) ids0[k] = fileIds.empty() ? (uint32_t)k : fileIds[k];
    if (!o.out.empty()) {
        check(akua_traj_create(o.out.c_str(), n, o.fields, dt, boxMin, boxMax, cfg.gravity, &out),
              "creating the trajectory");
        check(akua_traj_append(out, step0, time0, pos.data(), vel.data(), mass.data(), dens.data(), lam.data(),
                               ids0.data()),
              "writing the initial frame");
        ++written;
    }
    if (o.dryRun) {
        if (out) check(akua_traj_close(out), "closing the trajectory");
        std::printf("{\"dry_run\": true, \"source\": %s, \"particles\": %d, \"dt\": %.9g, \"frames\": %lld, "
                    "\"steps\": %lld, \"steps_per_frame\": [",
                    json_str(source).c_str(), n, dt, (long long)frames, (long long)totalSteps);
        for (size_t f = 0; f < plan.size(); ++f) std::printf("%s%d", f ? ", " : "", plan[f]);
        std::printf("], \"frames_written\": %lld}\n", (long long)written);
        return 0;
    }

    // ---- the run ----------------------------------------------------------------------------
    akua_solver* s = nullptr;
    check(akua_solver_create(n, &cfg, &corr, &s), "creating the solver");
    check(akua_solver_set_graph(s, o.graph ? 1 : 0), "set_graph");
    if (o.boundary > 0) check(akua_solver_set_boundary(s, o.boundary, o.boundarySpacing), "set_boundary");
    check(akua_solver_set_state(s, pos.data(), vel.data(), mass.data(), dens.data(), lam.data(), n),
          "uploading the particles");
    if (!o.network.empty()) {
        const std::vector<float> net = load_npy_f32(o.network);
        check(akua_solver_set_network(s, net.data(), (int32_t)net.size()), "set_network");
        check(akua_solver_set_learned(s, 1, o.learnedIters, 0, o.gateRatio, o.gateTol), "set_learned");
    }
    akua_tuning tuning{};
    if (!o.tuning.empty()) {
        int32_t* f[] = {&tuning.block, &tuning.layoutLevel, &tuning.hashDigitBits, &tuning.layoutDigitBits,
                        &tuning.layoutInterval, &tuning.balanceTile, &tuning.slabBatch, &tuning.fuseSweep,
                        &tuning.laneGroup, &tuning.workQueue};
        for (size_t k = 0; k < o.tuning.size(); ++k) *f[k] = o.tuning[k];
        check(akua_solver_set_tuning(s, &tuning), "set_tuning");
    }
    if (o.autotune) check(akua_solver_autotune(s, dt, boxMin, boxMax, 3, nullptr, &tuning), "autotune");
    check(akua_solver_get_tuning(s, &tuning), "get_tuning");

    using clock = std::chrono::steady_clock;
    double ioSeconds = 0.0, renderSeconds = 0.0;
    int64_t images = 0;
    std::vector<uint8_t> img;
    if (o.renderEvery > 0) {
        renderSeconds += render_frame(s, o, step0, img);
        ++images;
    }
    const auto t0 = clock::now();
    int64_t step = step0;
    double time = time0;
    int64_t done = 0;
    const uint32_t* idMap = fileIds.empty() ? nullptr : fileIds.data();
    for (size_t f = 0; f < plan.size(); ++f) {
        for (int k = 0; k < plan[f]; ++k) {
            check(akua_solver_step(s, dt, boxMin, boxMax), "step");
            ++step;
            ++done;
            time += (double)dt;
            if (out && done % o.every == 0) {
                check(akua_solver_synchronize(s), "synchronize");
                const auto a = clock::now();
                check(akua_traj_append_solver(out, s, step, time, idMap), "writing a frame");
                ioSeconds += std::chrono::duration<double>(clock::now() - a).count();
                ++written;
            }
            if (o.renderEvery > 0 && done % o.renderEvery == 0) {
                check(akua_solver_synchronize(s), "synchronize");
                const double r = render_frame(s, o, step, img);
                ioSeconds += r;
                renderSeconds += r;
                ++images;
            }
        }
    }
    check(akua_solver_synchronize(s), "synchronize");
    const double total = std::chrono::duration<double>(clock::now() - t0).count();
    float maxErr = 0.0f;
    double meanErr = 0.0;
    check(akua_solver_density_error(s, &maxErr, &meanErr), "density error");
    int64_t learnedSteps = 0, fallbacks = 0;
    float resid[6] = {0, 0, 0, 0, 0, 0};
    if (!o.network.empty()) check(akua_solver_learned_stats(s, &learnedSteps, &fallbacks, resid), "learned stats");
    if (out) check(akua_traj_close(out), "closing the trajectory");
    check(akua_solver_destroy(s), "destroying the solver");

    const double stepSeconds = total - ioSeconds;
    std::printf("{\"source\": %s, \"particles\": %d, \"params\": \"%s\", \"iterations\": %d, \"dt\": %.9g, "
                "\"frames\": %lld, \"steps\": %lld, \"seconds\": %.6f, \"io_seconds\": %.6f, "
                "\"particle_substeps_per_s\": %.6g, \"max_density_error\": %.6g, \"mean_density_error\": %.6g, "
                "\"tuning\": [%d, %d, %d, %d, %d, %d, %d, %d, %d, %d], \"graph\": %s, \"boundary_layers\": %d",
                json_str(source).c_str(), n, referenceParams ? "reference" : "dam_break", o.iterations, dt,
                (long long)frames, (long long)done, total, ioSeconds,
                stepSeconds > 0.0 ? (double)n * (double)done / stepSeconds : 0.0, maxErr, meanErr, tuning.block,
                tuning.layoutLevel, tuning.hashDigitBits, tuning.layoutDigitBits, tuning.layoutInterval,
                tuning.balanceTile, tuning.slabBatch, tuning.fuseSweep, tuning.laneGroup, tuning.workQueue,
                o.graph ? "true" : "false", o.boundary);
    if (!o.network.empty())
        std::printf(", \"learned\": {\"iterations\": %d, \"learned_steps\": %lld, \"fallbacks\": %lld}", o.learnedIters,
                    (long long)learnedSteps, (long long)fallbacks);
    if (out) std::printf(", \"out\": %s, \"frames_written\": %lld", json_str(o.out).c_str(), (long long)written);
    if (o.renderEvery > 0)
        std::printf(", \"images\": %lld, \"render_seconds\": %.6f", (long long)images, renderSeconds);
    std::printf("}\n");
    return 0;
}
