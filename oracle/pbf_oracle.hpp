// pbf_oracle.hpp -- CPU restatement of the AkuaEngine Position-Based Fluids step.
//
// TEST INFRASTRUCTURE ONLY.  This header is the correctness oracle: it is compiled into
// oracle/liboracle.so and used by tests/, by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg, always as the checker or the timed CPU baseline, never as a product path.
//
// It restates, per function, the reference's CUDA kernels (file:line citations below point into
// the Kevin22888/AkuaEngine sources) as a plain serial/OpenMP C++ template on T = float
// (parity) or T = double (error reference).  Semantics fixed here, where the reference leaves
// them to the hardware or to a library:
//   * stable sort by hash over the current order (thrust::sort's tie order is unspecified);
//   * the neighbour walk follows the reference's stencil order (dx, dy, dz nested -1..1),
//     walks a bucket once per stencil cell that hashes to it (duplicates kept), caps at
//     maxNeighbours;
//   * XSPH is Jacobi: it reads the velocities from before the pass (the reference races);
//   * no FMA contraction anywhere (build with -ffp-contract=off);
//   * pi is 3.14f, as in SmoothingKernelsCUDA.h;
//   * powf(x, k) with a small integer exponent k inside the per-pair loops (the cubes and
//     squares of the SPH kernels, and s_corr's exponent n when it is an integer) is the k-fold
//     product x*x*...*x evaluated left to right; a non-integer n calls the libm powf.  The
//     per-call constants powf(h, 9) and powf(h, 6) call the libm powf as written.
//   * boundary samples (opt-in, no reference counterpart; the reference plans them as "virtual
//     particles", ConstraintSolverCUDA.cu:132-135): the definition above make_boundary().
// Parity status of each function is listed in the comment above it ("pinned by ...").
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace pbf_oracle {

// Planted mistakes, for the tests that show each pin fails when the function it pins is wrong.
// Only the build with ORACLE_PLANTS (oracle/liboracle_plant.so) carries the run-time switch; the
// regular build compiles every PLANT(k) to false.  Mistake k, where it is planted:
//   1 vorticities: v_i - v_j in place of v_j - v_i
//   2 vorticities: cross(grad W, v_ij) in place of cross(v_ij, grad W)
//   3 vorticities: m_i in place of m_j
//   4 vorticity_confinement: |w_j| - |w_i| in place of |w_i| - |w_j|
//   5 vorticity_confinement: cross(w_i, N) in place of cross(N, w_i)
//   6 vorticity_confinement: the factor 1 / rho_i dropped
//   7 vorticity_confinement: m_i in place of m_j
//   8 xsph: m_i / rho_i in place of m_j / rho_j
//   9 xsph: v_i - v_j in place of v_j - v_i
//  10 xsph: velocities updated in place (Gauss-Seidel) instead of Jacobi
//  11 boundary_damping: the normal of the y-max wall is +y
//  12 boundary_damping: every wall acts on the velocity from before the damping
//  13 boundary_damping: the x-max wall placed at boxMin.x
//  14 boundary_damping: the friction factor applied to the normal part, not the tangential one
#ifdef ORACLE_PLANTS
inline int& planted_mistake() {
    static int k = 0;
    return k;
}
#define PLANT(k) (::pbf_oracle::planted_mistake() == (k))
#else
#define PLANT(k) false
#endif

// Mirrors AkuaEngine::PBFConfig (PBFConfig.h:18-29).  Plain C layout, no glm.
struct Config {
    float restDensity;
    float particle_spacing;
    float smoothRadius;
    float spatialHashCellSize;
    float relaxation;
    float vorticityEpsilon;
    float viscosity;
    int32_t maxNeighbours;
    int32_t solverIterations;
    float gravity[3];
};

// Mirrors AkuaEngine::LambdaCorrParams (PBFConfig.h:10-15).  `enabled` == 0 drops the s_corr term
// from the position delta (and from the learned features).  The reference never reads the flag
// (ConstraintSolverCUDA.cu:99-130 always applies the term), so the two agree whenever it is set.
struct LambdaCorr {
    int32_t enabled;
    float k;
    float n;
    float delta_q;
};

template <class T>
struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// scalar * vec and vec * scalar, each component rounded once (glm operator*, type_vec3.inl)
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
// glm dot for vec3: tmp = a*b; tmp.x + tmp.y + tmp.z (glm/detail/func_geometric.inl:48-54)
template <class T> inline T dot(V3<T> a, V3<T> b) {
    T tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
    return (tx + ty) + tz;
}
// glm length = sqrt(dot(v, v)) (glm/detail/func_geometric.inl:8-13)
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// MathUtilsCUDA::cross (MathUtilsCUDA.h:12-18)
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <class T> inline T powT(T x, T e) { return std::pow(x, e); }
template <> inline float powT<float>(float x, float e) { return powf(x, e); }

// powf with the exponent given as a float: a small integer exponent k (0..16) is the k-fold
// product evaluated left to right, otherwise libm pow.
template <class T>
inline T ipow(T x, T e) {
    if (e >= T(0) && e <= T(16) && e == std::floor(e)) {
        int k = (int)e;
        if (k == 0) return T(1);
        T r = x;
        for (int i = 1; i < k; ++i) r = r * x;
        return r;
    }
    return powT<T>(x, e);
}

// The constant factors of the two SPH kernels, as written in SmoothingKernelsCUDA.h:20 and :27.
template <class T>
inline T poly6_coef(T h) { return T(315.0f) / (T(64.0f) * T(3.14f) * powT<T>(h, T(9.0f))); }
template <class T>
inline T spiky_coef(T h) { return -T(45.0f) / (T(3.14f) * powT<T>(h, T(6.0f))); }

// SmoothingKernels::device_poly6 (SmoothingKernelsCUDA.h:16-21).  Pinned by the closed form
// (known-answer tests in tests/test_oracle.py) and by the interior-lattice density.
// Kern caches h, h*h and the two constant factors; the value of each is the one the reference
// computes inline on every call.
template <class T>
struct Kern {
    T h, h2, cp, cs;
    explicit Kern(T hh) : h(hh), h2(hh * hh), cp(poly6_coef<T>(hh)), cs(spiky_coef<T>(hh)) {}
};

template <class T>
inline T poly6(T d2, const Kern<T>& k) {
    if (d2 > k.h2) return T(0);
    return k.cp * ipow<T>(k.h2 - d2, T(3.0f));
}
template <class T>
inline T poly6(T d2, T h) { return poly6<T>(d2, Kern<T>(h)); }

// SmoothingKernels::device_gradient_spiky (SmoothingKernelsCUDA.h:23-28).  Pinned by a central
// finite difference of the spiky kernel and by its closed form along an axis.
template <class T>
inline V3<T> grad_spiky(V3<T> r, const Kern<T>& k) {
    T rm = length(r);
    if (rm > k.h || rm < T(1e-5f)) return {T(0), T(0), T(0)};
    T s = k.cs * ipow<T>(k.h - rm, T(2.0f));
    return (s * r) / rm;
}
template <class T>
inline V3<T> grad_spiky(V3<T> r, T h) { return grad_spiky<T>(r, Kern<T>(h)); }

// static_cast<int>(float) as the GPU executes it: NaN gives 0 and out-of-range values saturate
// (C++ leaves both undefined; x86 would give INT_MIN).
template <class T>
inline int32_t to_int_sat(T x) {
    if (!(x == x)) return 0;
    if (x >= T(2147483648.0)) return INT32_MAX;
    if (x < T(-2147483648.0)) return INT32_MIN;
    return static_cast<int32_t>(x);
}

// device_discretize_position (NeighbourSearchCUDA.cu:15-21): floor(x / cellSize) per axis.
template <class T>
inline void discretize(V3<T> p, T cs, int32_t c[3]) {
    c[0] = to_int_sat<T>(std::floor(p.x / cs));
    c[1] = to_int_sat<T>(std::floor(p.y / cs));
    c[2] = to_int_sat<T>(std::floor(p.z / cs));
}

// device_get_hash (NeighbourSearchCUDA.cu:23-27).  The reference multiplies in int and casts to
// uint32; the wrap is restated as a uint32 multiply (no signed-overflow UB).  `% tableSize`
// converts the int table size to unsigned.  Pinned by hand-computed known answers.
inline uint32_t cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t tableSize) {
    uint32_t a = static_cast<uint32_t>(cx) * 73856093u;
    uint32_t b = static_cast<uint32_t>(cy) * 19349663u;
    uint32_t c = static_cast<uint32_t>(cz) * 83492791u;
    return (a ^ b ^ c) % tableSize;
}

// Per-particle state, in the reference's physical (sorted) order.  The reference keeps it in the
// 108-byte Particle record (Particle.h:8-31); only the fields the step reads or writes are here.
template <class T>
struct State {
    std::vector<V3<T>> x;     // position
    std::vector<V3<T>> v;     // velocity
    std::vector<V3<T>> xs;    // new_position (x*)
    std::vector<V3<T>> vs;    // new_velocity
    std::vector<V3<T>> dp;    // position_delta
    std::vector<V3<T>> w;     // vorticity
    std::vector<T> mass, density, lambda;
    std::vector<uint32_t> hash, id;
    std::vector<T> rhoB;      // boundary samples' density contribution at x* (boundary on)
    std::vector<V3<T>> gB;    // sum_b psi_b grad W(x*_i - x_b) (boundary on)
    size_t size() const { return x.size(); }
    void resize(size_t n) {
        x.resize(n); v.resize(n); xs.resize(n); vs.resize(n); dp.resize(n); w.resize(n);
        mass.resize(n); density.resize(n); lambda.resize(n); hash.resize(n); id.resize(n);
    }
};

// Everything one step leaves behind besides the state: the parity taps.
struct Taps {
    std::vector<uint32_t> perm;          // sorted slot k came from pre-sort slot perm[k]
    std::vector<uint32_t> nbr;           // row-major [i * maxNeighbours + k] as the reference
    std::vector<uint32_t> cnt;
    std::vector<uint32_t> cell_hash;     // occupied buckets in sorted order
    std::vector<uint32_t> cell_start;    // first sorted index of each bucket
    std::vector<uint32_t> cell_end;      // one past the last sorted index
};

// ---- boundary samples (opt-in; no reference counterpart) --------------------------------------
// The reference keeps particles in the box with a soft clamp and plans "virtual particles" in its
// place (ConstraintSolverCUDA.cu:132-135).  This is that extension, after Akinci et al. 2012
// ("Versatile rigid-fluid coupling for incompressible SPH") in the PBF form of Bender's
// PositionBasedDynamics library: static samples b on the six walls, each with the volume weight
// psi_b = rho0 / sum_k W(x_b - x_k) over the samples k, add
//   psi_b W(x*_i - x_b) to rho_i,  psi_b grad W(x*_i - x_b) to the sum behind grad_i C_i,
//   and lambda_i psi_b grad W(x*_i - x_b) to the sum behind dp_i (both sums times 1/rho0).
// The samples do not move and have no lambda of their own; the clamp and damping stay.
// Definition (the library implements the same):
//   per axis a: n_a = max(2, int(ceil((hi_a - lo_a) / s)) + 1), d_a = (hi_a - lo_a) / (n_a - 1);
//   wall w = 0..5: normal axis a = w / 2, side w % 2 (lo, hi), tangential axes b < c;
//   layer l = 0..L-1 at lo_a - l s (side 0) or hi_a + l s (side 1);
//   tangential index t along axis e: t in [-(L-1), n_e - 1 + (L-1)] when e > a (the wall owns
//   that edge), t in [1, n_e - 2] when e < a; coordinate lo_e + t d_e;
//   sample index base_w + (l N_b + (t_b - t0_b)) N_c + (t_c - t0_c), base_w the samples of the
//   walls before w.
//   For a point x, the samples that count are those of the index window below with
//   |x - x_b|^2 < h^2 (dot in the usual order); sums run in ascending sample index.  The window
//   per wall and layer: layers with (x_a - plane)^2 < h^2, then t_e from
//   floor((x_e - lo_e - h) / d_e) - 1 to floor((x_e - lo_e + h) / d_e) + 1 (clamped to the wall),
//   rows and columns with (x_e - coord)^2 >= h^2 skipped.  Pinned by brute force over all
//   samples, by psi's invariant sum_k psi_k W_bk = rho0 on a flat wall, and by closed forms.
template <class T>
struct Boundary {
    int L = 0;
    T s{}, lo[3]{}, hi[3]{}, d[3]{};
    int n[3]{};
    int t0[6][2]{}, tn[6][2]{};
    size_t base[7]{};
    std::vector<T> psi;
    size_t size() const { return base[6]; }
};

inline void wall_axes(int w, int& a, int& b, int& c) {
    a = w / 2;
    b = a == 0 ? 1 : 0;
    c = a == 2 ? 1 : 2;
}

template <class T>
inline long long floor_ll(T v) {
    if (!(v == v)) return 0;
    if (v >= T(1e15)) return (long long)1e15;
    if (v <= T(-1e15)) return -(long long)1e15;
    return (long long)std::floor(v);
}

// Returns false when the box or the spacing admits no boundary (a non-positive extent, or more
// than 2^26 samples).
template <class T>
bool make_boundary(Boundary<T>& B, const T lo[3], const T hi[3], T s, int L) {
    B = Boundary<T>{};
    if (!(s > T(0)) || L < 1 || L > 4) return false;
    B.L = L;
    B.s = s;
    for (int a = 0; a < 3; ++a) {
        B.lo[a] = lo[a];
        B.hi[a] = hi[a];
        const T E = hi[a] - lo[a];
        if (!(E > T(0))) return false;
        const T q = std::ceil(E / s);
        if (!(q < T(1 << 20))) return false;
        B.n[a] = std::max(2, (int)q + 1);
        B.d[a] = E / T(B.n[a] - 1);
    }
    B.base[0] = 0;
    for (int w = 0; w < 6; ++w) {
        int a, b, c;
        wall_axes(w, a, b, c);
        const int e2[2] = {b, c};
        for (int k = 0; k < 2; ++k) {
            const int e = e2[k];
            if (e > a) {
                B.t0[w][k] = -(L - 1);
                B.tn[w][k] = B.n[e] + 2 * (L - 1);
            } else {
                B.t0[w][k] = 1;
                B.tn[w][k] = std::max(0, B.n[e] - 2);
            }
        }
        B.base[w + 1] = B.base[w] + (size_t)L * (size_t)B.tn[w][0] * (size_t)B.tn[w][1];
    }
    return B.size() <= (size_t(1) << 26);
}

template <class T>
V3<T> boundary_sample(const Boundary<T>& B, size_t idx) {
    int w = 0;
    while (w < 5 && idx >= B.base[w + 1]) ++w;
    int a, b, c;
    wall_axes(w, a, b, c);
    size_t r = idx - B.base[w];
    const size_t nc = (size_t)B.tn[w][1], nb = (size_t)B.tn[w][0];
    const int tc = (int)(r % nc) + B.t0[w][1];
    r /= nc;
    const int tb = (int)(r % nb) + B.t0[w][0];
    const int l = (int)(r / nb);
    T sp[3];
    sp[a] = (w % 2) ? B.hi[a] + T(l) * B.s : B.lo[a] - T(l) * B.s;
    sp[b] = B.lo[b] + T(tb) * B.d[b];
    sp[c] = B.lo[c] + T(tc) * B.d[c];
    return {sp[0], sp[1], sp[2]};
}

// Calls f(index, x - x_b, |x - x_b|^2) for every sample that counts for x, in ascending index.
template <class T, class F>
void for_boundary_near(const Boundary<T>& B, V3<T> x, const Kern<T>& K, F&& f) {
    const T xa[3] = {x.x, x.y, x.z};
    for (int w = 0; w < 6; ++w) {
        int a, b, c;
        wall_axes(w, a, b, c);
        if (B.tn[w][0] == 0 || B.tn[w][1] == 0) continue;
        auto window = [&](int e, int k, long long& lo, long long& hi) {
            lo = std::max<long long>(B.t0[w][k], floor_ll<T>((xa[e] - B.lo[e] - K.h) / B.d[e]) - 1);
            hi = std::min<long long>(B.t0[w][k] + B.tn[w][k] - 1,
                                     floor_ll<T>((xa[e] - B.lo[e] + K.h) / B.d[e]) + 1);
        };
        long long b0, b1, c0, c1;
        window(b, 0, b0, b1);
        window(c, 1, c0, c1);
        for (int l = 0; l < B.L; ++l) {
            const T na = (w % 2) ? B.hi[a] + T(l) * B.s : B.lo[a] - T(l) * B.s;
            const T ra = xa[a] - na;
            if (ra * ra >= K.h2) continue;
            for (long long tb = b0; tb <= b1; ++tb) {
                const T cb = B.lo[b] + T(tb) * B.d[b];
                const T rb = xa[b] - cb;
                if (rb * rb >= K.h2) continue;
                for (long long tc = c0; tc <= c1; ++tc) {
                    const T cc = B.lo[c] + T(tc) * B.d[c];
                    const T rc = xa[c] - cc;
                    if (rc * rc >= K.h2) continue;
                    T sp[3];
                    sp[a] = na;
                    sp[b] = cb;
                    sp[c] = cc;
                    const V3<T> sep = x - V3<T>{sp[0], sp[1], sp[2]};
                    const T d2 = dot(sep, sep);
                    if (!(d2 < K.h2)) continue;
                    const size_t idx = B.base[w] +
                                       ((size_t)l * (size_t)B.tn[w][0] + (size_t)(tb - B.t0[w][0])) *
                                           (size_t)B.tn[w][1] +
                                       (size_t)(tc - B.t0[w][1]);
                    f(idx, sep, d2);
                }
            }
        }
    }
}

// psi_b = rho0 / sum_k W(x_b - x_k), the sum over the samples that count for x_b (b included).
template <class T>
void boundary_psi(Boundary<T>& B, const Kern<T>& K, T rest) {
    const long m = (long)B.size();
    B.psi.assign((size_t)m, T(0));
#pragma omp parallel for schedule(static)
    for (long k = 0; k < m; ++k) {
        T delta = T(0);
        for_boundary_near(B, boundary_sample(B, (size_t)k), K,
                          [&](size_t, V3<T>, T d2) { delta = delta + poly6<T>(d2, K); });
        B.psi[(size_t)k] = rest / delta;
    }
}

template <class T>
struct Params {
    const Boundary<T>* bnd = nullptr;  // boundary samples (null: the reference's walls only)
    Kern<T> K{T(0.1f)};
    T h, h2, cellHash, cellSearch, invRest, relax, vortEps, visc, dt;
    T g[3], bmin[3], bmax[3];
    T corrK, corrN, corrDq;
    bool corrOn;
    uint32_t tableSize;
    int32_t maxN, iters;
};

template <class T>
Params<T> make_params(const Config& c, const LambdaCorr& lc, float dt, const float bmin[3],
                      const float bmax[3], uint32_t tableSize) {
    Params<T> p;
    p.K = Kern<T>(T(c.smoothRadius));
    p.h = T(c.smoothRadius);
    p.h2 = p.h * p.h;
    // NeighbourSearchCUDA.cu:163 passes smoothRadius as the hashing cell size; the search itself
    // (NeighbourSearchCUDA.cu:177-178) uses spatialHashCellSize.
    p.cellHash = T(c.smoothRadius);
    p.cellSearch = T(c.spatialHashCellSize);
    // ConstraintSolverCUDA.cu:201: inverseRestDensity = 1.0f / restDensity (in float).
    p.invRest = T(1.0f / c.restDensity);
    if (sizeof(T) == 8) p.invRest = T(1.0) / T(c.restDensity);
    p.relax = T(c.relaxation);
    p.vortEps = T(c.vorticityEpsilon);
    p.visc = T(c.viscosity);
    p.dt = T(dt);
    for (int k = 0; k < 3; ++k) {
        p.g[k] = T(c.gravity[k]);
        p.bmin[k] = T(bmin[k]);
        p.bmax[k] = T(bmax[k]);
    }
    p.corrK = T(lc.k);
    p.corrN = T(lc.n);
    p.corrDq = T(lc.delta_q);
    p.corrOn = lc.enabled != 0;
    p.tableSize = tableSize;
    p.maxN = c.maxNeighbours;
    p.iters = c.solverIterations;
    return p;
}

// kernel_predict_position (IntegrationCUDA.cu:27-36): v* = v + g*dt; x* = x + v*·dt.
template <class T>
void predict(State<T>& s, const Params<T>& p) {
    const V3<T> g{p.g[0], p.g[1], p.g[2]};
    const long n = (long)s.size();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        s.vs[i] = s.v[i] + g * p.dt;
        s.xs[i] = s.x[i] + s.vs[i] * p.dt;
    }
}

// kernel_compute_hashes (NeighbourSearchCUDA.cu:36-44), launched with smoothRadius as cellSize
// (NeighbourSearchCUDA.cu:163).
template <class T>
void compute_hashes(State<T>& s, const Params<T>& p) {
    const long n = (long)s.size();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        int32_t c[3];
        discretize(s.xs[i], p.cellHash, c);
        s.hash[i] = cell_hash(c[0], c[1], c[2], p.tableSize);
    }
}

template <class A>
static void permute(std::vector<A>& a, const std::vector<uint32_t>& perm, std::vector<A>& tmp) {
    tmp.resize(a.size());
    for (size_t k = 0; k < a.size(); ++k) tmp[k] = a[perm[k]];
    a.swap(tmp);
}

// thrust::sort by hash (NeighbourSearchCUDA.cu:166-170), restated as a STABLE sort over the
// current order; the whole record moves, so the permutation accumulates across steps.
// Pinned by std::stable_sort itself and by the sortedness/stability invariants in the tests.
template <class T>
void sort_by_hash(State<T>& s, std::vector<uint32_t>& perm) {
    const size_t n = s.size();
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return s.hash[a] < s.hash[b]; });
    std::vector<V3<T>> t3;
    std::vector<T> t1;
    std::vector<uint32_t> tu;
    permute(s.x, perm, t3); permute(s.v, perm, t3); permute(s.xs, perm, t3);
    permute(s.vs, perm, t3); permute(s.dp, perm, t3); permute(s.w, perm, t3);
    permute(s.mass, perm, t1); permute(s.density, perm, t1); permute(s.lambda, perm, t1);
    permute(s.hash, perm, tu); permute(s.id, perm, tu);
}

// kernel_build_hash_table (NeighbourSearchCUDA.cu:52-65) stores the first sorted index of each
// hash; a bucket ends where the hash changes (the walk in NeighbourSearchCUDA.cu:103-124).  The
// oracle keeps the occupied buckets as sorted (hash, start, end) triples and finds a bucket by
// binary search, which yields the same first index the reference's table holds.
inline void cell_ranges(const std::vector<uint32_t>& sortedHash, std::vector<uint32_t>& ch,
                        std::vector<uint32_t>& cs, std::vector<uint32_t>& ce) {
    ch.clear(); cs.clear(); ce.clear();
    const size_t n = sortedHash.size();
    for (size_t i = 0; i < n; ++i) {
        if (i == 0 || sortedHash[i] != sortedHash[i - 1]) {
            ch.push_back(sortedHash[i]);
            cs.push_back((uint32_t)i);
            if (i > 0) ce.push_back((uint32_t)i);
        }
    }
    if (n > 0) ce.push_back((uint32_t)n);
}

// kernel_find_neighbours (NeighbourSearchCUDA.cu:72-130).  Pinned by O(N^2) brute force as
// multisets (duplicates from stencil hash collisions computed from the hash) in the tests.
template <class T>
void find_neighbours(const State<T>& s, const Params<T>& p, const std::vector<uint32_t>& ch,
                     const std::vector<uint32_t>& cs, const std::vector<uint32_t>& ce,
                     std::vector<uint32_t>& nbr, std::vector<uint32_t>& cnt) {
    const long n = (long)s.size();
    const int maxN = p.maxN;
    nbr.assign((size_t)n * (size_t)maxN, 0u);
    cnt.assign((size_t)n, 0u);
    // NeighbourSearchCUDA.cu:90: float smoothRadiusSquared = smoothRadius * smoothRadius
    const T h2 = p.h * p.h;
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < n; ++i) {
        int32_t c[3];
        discretize(s.xs[i], p.cellSearch, c);
        int count = 0;
        const size_t start = (size_t)i * maxN;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    // glm ivec3 addition wraps like int; restate in uint32 then back.
                    int32_t qx = (int32_t)((uint32_t)c[0] + (uint32_t)dx);
                    int32_t qy = (int32_t)((uint32_t)c[1] + (uint32_t)dy);
                    int32_t qz = (int32_t)((uint32_t)c[2] + (uint32_t)dz);
                    uint32_t hh = cell_hash(qx, qy, qz, p.tableSize);
                    auto it = std::lower_bound(ch.begin(), ch.end(), hh);
                    if (it == ch.end() || *it != hh) continue;  // table entry UINT32_MAX
                    size_t b = (size_t)(it - ch.begin());
                    long k = (long)cs[b];
                    while (k < n && count < maxN) {
                        if (k == i) { ++k; continue; }
                        if (s.hash[k] != hh) break;
                        V3<T> sep = s.xs[i] - s.xs[k];
                        T d2 = dot(sep, sep);
                        if (d2 < h2) nbr[start + count++] = (uint32_t)k;
                        ++k;
                    }
                }
        cnt[i] = (uint32_t)count;
    }
}

// kernel_calculate_densities (ConstraintSolverCUDA.cu:16-42).  Pinned by the interior-lattice
// known answer and by brute-force sums over all particles within h.
template <class T>
void densities(State<T>& s, const Params<T>& p, const std::vector<uint32_t>& nbr,
               const std::vector<uint32_t>& cnt) {
    const long n = (long)s.size();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        T rho = s.mass[i] * poly6<T>(T(0), p.K);
        const size_t st = (size_t)i * p.maxN;
        for (uint32_t o = 0; o < cnt[i]; ++o) {
            uint32_t j = nbr[st + o];
            V3<T> sep = s.xs[i] - s.xs[j];
            T d2 = dot(sep, sep);
            rho = rho + s.mass[j] * poly6<T>(d2, p.K);
        }
        if (p.bnd) rho = rho + s.rhoB[i];
        s.density[i] = rho;
    }
}

// kernel_calculate_lambdas (ConstraintSolverCUDA.cu:51-97).  Pinned by the interior-lattice
// known answer and the two-particle closed form.
template <class T>
void lambdas(State<T>& s, const Params<T>& p, const std::vector<uint32_t>& nbr,
             const std::vector<uint32_t>& cnt) {
    const long n = (long)s.size();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const size_t st = (size_t)i * p.maxN;
        V3<T> gi{T(0), T(0), T(0)};
        for (uint32_t o = 0; o < cnt[i]; ++o) {
            uint32_t j = nbr[st + o];
            V3<T> sep = s.xs[i] - s.xs[j];
            gi = gi + s.mass[j] * grad_spiky<T>(sep, p.K);
        }
        if (p.bnd) gi = gi + s.gB[i];
        gi = gi * p.invRest;  // grad_pi_Ci *= inverseRestDensity
        T sum = T(0);
        for (uint32_t o = 0; o < cnt[i]; ++o) {
            uint32_t j = nbr[st + o];
            V3<T> sep = s.xs[i] - s.xs[j];
            V3<T> gj = (-p.invRest * s.mass[j]) * grad_spiky<T>(sep, p.K);
            sum = sum + dot(gj, gj);
        }
        T C = s.density[i] * p.invRest - T(1);
        s.lambda[i] = -C / (sum + dot(gi, gi) + p.relax);
    }
}

// kernel_calculate_position_delta (ConstraintSolverCUDA.cu:99-130).  Pinned by
// sum_i m_i * dp_i = 0 for symmetric neighbour sets and by the two-particle closed form.
template <class T>
void position_deltas(State<T>& s, const Params<T>& p, const std::vector<uint32_t>& nbr,
                     const std::vector<uint32_t>& cnt) {
    const long n = (long)s.size();
    const T dq2 = p.corrDq * p.corrDq;
    const T wdq = poly6<T>(dq2, p.K);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const size_t st = (size_t)i * p.maxN;
        V3<T> d{T(0), T(0), T(0)};
        for (uint32_t o = 0; o < cnt[i]; ++o) {
            uint32_t j = nbr[st + o];
            V3<T> sep = s.xs[i] - s.xs[j];
            T d2 = dot(sep, sep);
            T coef = s.lambda[i] + s.lambda[j];
            if (p.corrOn) coef = coef + -p.corrK * ipow<T>(poly6<T>(d2, p.K) / wdq, p.corrN);
            d = d + (coef * s.mass[j]) * grad_spiky<T>(sep, p.K);
        }
        if (p.bnd) d = d + s.lambda[i] * s.gB[i];
        s.dp[i] = d * p.invRest;
    }
}

// handle_particle_collision + kernel_correct_position (ConstraintSolverCUDA.cu:136-169).
// Pinned by per-wall special cases.
template <class T>
void correct_positions(State<T>& s, const Params<T>& p) {
    const long n = (long)s.size();
    const T minDist = T(0.025f), stiff = T(0.5f);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        V3<T> q = s.xs[i] + s.dp[i];
        T c[3] = {T(0), T(0), T(0)};
        T qa[3] = {q.x, q.y, q.z};
        for (int a = 0; a < 3; ++a) {
            if (qa[a] < p.bmin[a] + minDist) c[a] = c[a] + stiff * (p.bmin[a] + minDist - qa[a]);
            if (qa[a] > p.bmax[a] - minDist) c[a] = c[a] + stiff * (p.bmax[a] - minDist - qa[a]);
        }
        s.xs[i] = q + V3<T>{c[0], c[1], c[2]};
    }
}

// kernel_update_position_and_velocity (IntegrationCUDA.cu:38-49).  Pinned by the lone-particle
// ballistic recursion.
template <class T>
void update_position_velocity(State<T>& s, const Params<T>& p) {
    const long n = (long)s.size();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        s.vs[i] = (s.xs[i] - s.x[i]) / p.dt;
        s.x[i] = s.xs[i];
        s.v[i] = s.vs[i];
    }
}

// resolve_collision + kernel_apply_boundary_velocity_damping (IntegrationCUDA.cu:51-102), called
// with restitution 0 and friction 0.95 (PBFSolver.cpp:64).  Planes are applied in order, each
// seeing the velocity the previous one left.  Pinned by known answers on each of the six walls
// of an asymmetric box (approaching, resting, leaving), an interior particle left untouched and a
// corner that needs the sequential order (tests/test_oracle_postpass.py).
template <class T>
void boundary_damping(State<T>& s, const Params<T>& p, T restitution, T friction) {
    const long n = (long)s.size();
    const T minDist = T(0.025f);
    const T xMax = PLANT(13) ? p.bmin[0] : p.bmax[0];
    const T nyMax = PLANT(11) ? T(1) : T(-1);
    const V3<T> pt[6] = {{p.bmin[0], 0, 0}, {xMax, 0, 0},        {0, p.bmin[1], 0},
                         {0, p.bmax[1], 0}, {0, 0, p.bmin[2]}, {0, 0, p.bmax[2]}};
    const V3<T> nm[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, nyMax, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const V3<T> x = s.x[i], v0 = s.v[i];
        V3<T> v = v0;
        for (int k = 0; k < 6; ++k) {
            const V3<T> vin = PLANT(12) ? v0 : v;
            T distance = dot(x - pt[k], nm[k]);
            T approaching = dot(vin, nm[k]);
            if (distance < minDist) {
                V3<T> vn = approaching * nm[k];
                V3<T> vt = vin - vn;
                if (PLANT(14)) std::swap(vn, vt);
                if (approaching < T(0)) {
                    v = (-restitution) * vn + (T(1) - friction) * vt;
                } else if (std::fabs(approaching) < T(1e-5f)) {
                    v = (T(1) - friction) * vt;
                }
            }
        }
        s.v[i] = v;
    }
}

// kernel_compute_vorticities (IntegrationCUDA.cu:104-128).  Pinned by zero vorticity under a
// uniform velocity field, and by rigid rotation v = Omega x r on an interior lattice with
// checkerboard masses: w_i = (2/3) Omega sum_j m_j |grad W(r_j)| |r_j|, summed over the shells
// of the 26 neighbours (tests/test_oracle_postpass.py).
template <class T>
void vorticities(State<T>& s, const Params<T>& p, const std::vector<uint32_t>& nbr,
                 const std::vector<uint32_t>& cnt) {
    const long n = (long)s.size();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const size_t st = (size_t)i * p.maxN;
        V3<T> w{T(0), T(0), T(0)};
        for (uint32_t o = 0; o < cnt[i]; ++o) {
            uint32_t j = nbr[st + o];
            V3<T> vij = PLANT(1) ? s.v[i] - s.v[j] : s.v[j] - s.v[i];
            V3<T> sep = s.xs[i] - s.xs[j];
            const V3<T> g = grad_spiky<T>(sep, p.K);
            const T mj = PLANT(3) ? s.mass[i] : s.mass[j];
            w = w + (-mj) * (PLANT(2) ? cross(g, vij) : cross(vij, g));
        }
        s.w[i] = w;
    }
}

// kernel_apply_vorticity_confinement (IntegrationCUDA.cu:130-165).  Pinned by closed forms
// (tests/test_oracle_postpass.py): on an interior lattice whose |w| grows along x with w along z,
// eta points along -x and the velocity gains dt * eps * |w_i| along +y; the 1e-5 threshold on
// |eta| decides at the value the 1 / rho_i factor gives; three particles with unequal masses fix
// the direction of eta.
template <class T>
void vorticity_confinement(State<T>& s, const Params<T>& p, const std::vector<uint32_t>& nbr,
                           const std::vector<uint32_t>& cnt) {
    const long n = (long)s.size();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const size_t st = (size_t)i * p.maxN;
        T invD = PLANT(6) ? T(1) : T(1) / s.density[i];
        V3<T> eta{T(0), T(0), T(0)};
        const T li = length(s.w[i]);
        for (uint32_t o = 0; o < cnt[i]; ++o) {
            uint32_t j = nbr[st + o];
            T diff = PLANT(4) ? length(s.w[j]) - li : li - length(s.w[j]);
            V3<T> sep = s.xs[i] - s.xs[j];
            const T mj = PLANT(7) ? s.mass[i] : s.mass[j];
            eta = eta + (mj * diff) * grad_spiky<T>(sep, p.K);
        }
        eta = eta * invD;
        T el = length(eta);
        if (el < T(1e-5f)) continue;
        V3<T> N = eta / el;
        V3<T> f = p.vortEps * (PLANT(5) ? cross(s.w[i], N) : cross(N, s.w[i]));
        s.v[i] = s.v[i] + p.dt * f;
    }
}

// kernel_apply_xsph_viscosity (IntegrationCUDA.cu:167-195), Jacobi: every particle reads the
// velocities from before the pass.  Pinned by "no change under a uniform velocity field", by the
// two-particle closed form with unequal masses and densities, and by a three-particle chain whose
// middle particle must see its neighbours' velocities from before the pass
// (tests/test_oracle_postpass.py).
template <class T>
void xsph(State<T>& s, const Params<T>& p, const std::vector<uint32_t>& nbr,
          const std::vector<uint32_t>& cnt, std::vector<V3<T>>& vnew) {
    const long n = (long)s.size();
    vnew.resize(n);
    auto dv_of = [&](long i, const std::vector<V3<T>>& v) {
        const size_t st = (size_t)i * p.maxN;
        V3<T> dv{T(0), T(0), T(0)};
        for (uint32_t o = 0; o < cnt[i]; ++o) {
            uint32_t j = nbr[st + o];
            V3<T> vij = PLANT(9) ? v[i] - v[j] : v[j] - v[i];
            V3<T> sep = s.xs[i] - s.xs[j];
            T w = poly6<T>(dot(sep, sep), p.K);
            const T mr = PLANT(8) ? s.mass[i] / s.density[i] : s.mass[j] / s.density[j];
            dv = dv + (mr * vij) * w;
        }
        return dv;
    };
    if (PLANT(10)) {  // in place, in slot order
        for (long i = 0; i < n; ++i) s.v[i] = s.v[i] + p.visc * dv_of(i, s.v);
        return;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) vnew[i] = s.v[i] + p.visc * dv_of(i, s.v);
    s.v.swap(vnew);
}

// Features of the learned warm start (no reference counterpart), per particle at x* = x*_pred
// right after the neighbour search, in list order:
//   0 C = rho / rho0 - 1, 1 lambda (both as densities() + lambdas() compute them), 2 neighbour
//   count, 3 the lambda carried from the previous step, 4-6 (x* - x) / h,
//   7-9 2 lambda (1/rho0) sum_j m_j grad W_ij / h, 10-12 (1/rho0) sum_j m_j s_corr,ij grad W_ij / h,
//   13-15 per axis clamp((x* - lo) / h, 0, 2) - clamp((hi - x*) / h, 0, 2).
// Pinned by its first two entries equalling densities()/lambdas(), by the closed forms of the
// others on hand-placed particles, and by the device kernel's independent implementation.
template <class T>
void features(const State<T>& s, const Params<T>& p, const std::vector<uint32_t>& nbr,
              const std::vector<uint32_t>& cnt, std::vector<T>& F) {
    const long n = (long)s.size();
    F.assign((size_t)n * 16, T(0));
    const T dq2 = p.corrDq * p.corrDq;
    const T wdq = poly6<T>(dq2, p.K);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const size_t st = (size_t)i * p.maxN;
        T rho = s.mass[i] * poly6<T>(T(0), p.K);
        V3<T> gi{T(0), T(0), T(0)}, ac{T(0), T(0), T(0)};
        T sum = T(0);
        for (uint32_t o = 0; o < cnt[i]; ++o) {
            const uint32_t j = nbr[st + o];
            const V3<T> sep = s.xs[i] - s.xs[j];
            const T d2 = dot(sep, sep);
            const T w = poly6<T>(d2, p.K);
            rho = rho + s.mass[j] * w;
            const V3<T> g = grad_spiky<T>(sep, p.K);
            gi = gi + s.mass[j] * g;
            const V3<T> gj = (-p.invRest * s.mass[j]) * g;
            sum = sum + dot(gj, gj);
            if (p.corrOn) {
                const T lc = -p.corrK * ipow<T>(w / wdq, p.corrN);
                ac = ac + (lc * s.mass[j]) * g;
            }
        }
        gi = gi * p.invRest;
        const T C = rho * p.invRest - T(1);
        const T lam1 = -C / (sum + dot(gi, gi) + p.relax);
        const T invH = T(1) / p.h;
        const V3<T> sg = gi * (T(2) * lam1 * invH);
        const V3<T> sa = ac * (p.invRest * invH);
        const T xa[3] = {s.xs[i].x, s.xs[i].y, s.xs[i].z};
        const T x0[3] = {s.x[i].x, s.x[i].y, s.x[i].z};
        T* f = &F[(size_t)i * 16];
        f[0] = C;
        f[1] = lam1;
        f[2] = T(cnt[i]);
        f[3] = s.lambda[i];
        for (int a = 0; a < 3; ++a) f[4 + a] = (xa[a] - x0[a]) * invH;
        f[7] = sg.x; f[8] = sg.y; f[9] = sg.z;
        f[10] = sa.x; f[11] = sa.y; f[12] = sa.z;
        for (int a = 0; a < 3; ++a) {
            const T lo = std::min(std::max((xa[a] - p.bmin[a]) * invH, T(0)), T(2));
            const T hi = std::min(std::max((p.bmax[a] - xa[a]) * invH, T(0)), T(2));
            f[13 + a] = lo - hi;
        }
    }
}

// Boundary sums at x* (boundary on): rhoB_i = sum_b psi_b W(x*_i - x_b) and
// gB_i = sum_b psi_b grad W(x*_i - x_b), in ascending sample index.
template <class T>
void boundary_sums(State<T>& s, const Params<T>& p) {
    const long n = (long)s.size();
    s.rhoB.assign((size_t)n, T(0));
    s.gB.assign((size_t)n, V3<T>{T(0), T(0), T(0)});
    const Boundary<T>& B = *p.bnd;
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < n; ++i) {
        T rb = T(0);
        V3<T> gb{T(0), T(0), T(0)};
        for_boundary_near(B, s.xs[i], p.K, [&](size_t idx, V3<T> sep, T d2) {
            const T psi = B.psi[idx];
            rb = rb + psi * poly6<T>(d2, p.K);
            gb = gb + psi * grad_spiky<T>(sep, p.K);
        });
        s.rhoB[i] = rb;
        s.gB[i] = gb;
    }
}

// One iteration of runConstraintSolverCUDA's loop (ConstraintSolverCUDA.cu:206-218).
template <class T>
void constraint_iteration(State<T>& s, const Params<T>& p, const std::vector<uint32_t>& nbr,
                          const std::vector<uint32_t>& cnt) {
    if (p.bnd) boundary_sums(s, p);
    densities(s, p, nbr, cnt);
    lambdas(s, p, nbr, cnt);
    position_deltas(s, p, nbr, cnt);
    correct_positions(s, p);
}

// PBFSolver::step (PBFSolver.cpp:22-78): predict, hash, sort, neighbours, solverIterations x
// constraint iteration, update, damping(0, 0.95), vorticity, confinement, XSPH.
template <class T>
void step(State<T>& s, const Params<T>& p, Taps* taps) {
    Taps local;
    Taps& t = taps ? *taps : local;
    predict(s, p);
    compute_hashes(s, p);
    sort_by_hash(s, t.perm);
    cell_ranges(s.hash, t.cell_hash, t.cell_start, t.cell_end);
    find_neighbours(s, p, t.cell_hash, t.cell_start, t.cell_end, t.nbr, t.cnt);
    for (int it = 0; it < p.iters; ++it) constraint_iteration(s, p, t.nbr, t.cnt);
    update_position_velocity(s, p);
    boundary_damping(s, p, T(0.0f), T(0.95f));
    vorticities(s, p, t.nbr, t.cnt);
    vorticity_confinement(s, p, t.nbr, t.cnt);
    std::vector<V3<T>> vnew;
    xsph(s, p, t.nbr, t.cnt, vnew);
}

// Density of every particle at the current positions (x) over all particles within h, each
// counted once: particles are hashed with the hashing cell size and stably sorted as a step would
// sort them; each distinct stencil bucket is walked once (no duplicates, no cap).  rhoSorted[k] is
// the density of the particle in sorted slot k; rhoById[id] the same value by particle id.  This
// is the quantity the max-density-error metric reads.
template <class T>
void true_density(const State<T>& s, const Params<T>& p, std::vector<T>& rhoSorted,
                  std::vector<T>& rhoById) {
    State<T> t = s;
    t.xs = t.x;
    compute_hashes(t, p);
    std::vector<uint32_t> perm;
    sort_by_hash(t, perm);
    std::vector<uint32_t> ch, cs, ce;
    cell_ranges(t.hash, ch, cs, ce);
    const long n = (long)t.size();
    rhoSorted.assign(n, T(0));
    rhoById.assign(n, T(0));
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < n; ++i) {
        int32_t c[3];
        discretize(t.xs[i], p.cellHash, c);
        T rho = t.mass[i] * poly6<T>(T(0), p.K);
        uint32_t seen[27];
        int nseen = 0;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    uint32_t hh = cell_hash((int32_t)((uint32_t)c[0] + (uint32_t)dx),
                                            (int32_t)((uint32_t)c[1] + (uint32_t)dy),
                                            (int32_t)((uint32_t)c[2] + (uint32_t)dz), p.tableSize);
                    bool dup = false;
                    for (int q = 0; q < nseen; ++q) dup = dup || seen[q] == hh;
                    if (dup) continue;
                    seen[nseen++] = hh;
                    auto it = std::lower_bound(ch.begin(), ch.end(), hh);
                    if (it == ch.end() || *it != hh) continue;
                    size_t b = (size_t)(it - ch.begin());
                    for (long k = (long)cs[b]; k < (long)ce[b]; ++k) {
                        if (k == i) continue;
                        V3<T> sep = t.xs[i] - t.xs[k];
                        T d2 = dot(sep, sep);
                        if (d2 < p.h2) rho = rho + t.mass[k] * poly6<T>(d2, p.K);
                    }
                }
        rhoSorted[i] = rho;
        rhoById[t.id[i]] = rho;
    }
}

}  // namespace pbf_oracle
