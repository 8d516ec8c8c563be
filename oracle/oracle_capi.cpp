// oracle_capi.cpp -- C ABI over pbf_oracle.hpp for the tests and the CPU baseline (ctypes).
//
// TEST INFRASTRUCTURE ONLY: the oracle is the checker.  Nothing in the product library links or
// loads this file.  One handle holds one scene's state in the reference's physical (sorted) order
// in FP32 (prec = 32, the parity oracle) or FP64 (prec = 64, the error reference).
#include "pbf_oracle.hpp"

#include <memory>

using namespace pbf_oracle;

namespace {

struct Handle {
    int prec;
    Config cfg;
    LambdaCorr corr;
    uint32_t tableSize;
    State<float> sf;
    State<double> sd;
    Taps taps;
    std::vector<uint32_t> preSortHash;
    std::vector<float> featF;
    std::vector<double> featD;
    // boundary samples (opt-in): settings, and the samples built for the last box seen
    int bndLayers = 0;          // 0: off
    float bndSpacing = 0.0f;
    bool bndBuilt = false;
    float bndBox[6] = {0, 0, 0, 0, 0, 0};
    Boundary<float> bf;
    Boundary<double> bd;
};

template <class T>
Boundary<T>& bnd(Handle* h);
template <>
Boundary<float>& bnd<float>(Handle* h) { return h->bf; }
template <>
Boundary<double>& bnd<double>(Handle* h) { return h->bd; }

// The boundary for this box (rebuilt when the box changes); null when off or not admissible.
template <class T>
const Boundary<T>* boundary_for(Handle* h, const float* bmin, const float* bmax) {
    if (h->bndLayers <= 0) return nullptr;
    const float box[6] = {bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2]};
    Boundary<T>& B = bnd<T>(h);
    if (!h->bndBuilt || std::memcmp(box, h->bndBox, sizeof(box)) != 0) {
        const T lo[3] = {T(bmin[0]), T(bmin[1]), T(bmin[2])};
        const T hi[3] = {T(bmax[0]), T(bmax[1]), T(bmax[2])};
        const T sp = T(h->bndSpacing > 0.0f ? h->bndSpacing : h->cfg.particle_spacing);
        if (!make_boundary<T>(B, lo, hi, sp, h->bndLayers)) return nullptr;
        boundary_psi<T>(B, Kern<T>(T(h->cfg.smoothRadius)), T(h->cfg.restDensity));
        std::memcpy(h->bndBox, box, sizeof(box));
        h->bndBuilt = true;
    }
    return &B;
}

template <class T>
State<T>& st(Handle* h);
template <>
State<float>& st<float>(Handle* h) { return h->sf; }
template <>
State<double>& st<double>(Handle* h) { return h->sd; }

template <class T>
std::vector<T>& feat(Handle* h);
template <>
std::vector<float>& feat<float>(Handle* h) { return h->featF; }
template <>
std::vector<double>& feat<double>(Handle* h) { return h->featD; }

template <class T>
void set_state(Handle* h, long n, const float* x, const float* v, const float* m, const float* d,
               const float* l, const uint32_t* id) {
    State<T>& s = st<T>(h);
    s.resize(n);
    for (long i = 0; i < n; ++i) {
        s.x[i] = {T(x[3 * i]), T(x[3 * i + 1]), T(x[3 * i + 2])};
        s.v[i] = {T(v[3 * i]), T(v[3 * i + 1]), T(v[3 * i + 2])};
        s.xs[i] = s.x[i];
        s.vs[i] = s.v[i];
        s.dp[i] = {T(0), T(0), T(0)};
        s.w[i] = {T(0), T(0), T(0)};
        s.mass[i] = m ? T(m[i]) : T(1);
        s.density[i] = d ? T(d[i]) : T(0);
        s.lambda[i] = l ? T(l[i]) : T(0);
        s.id[i] = id ? id[i] : (uint32_t)i;
        s.hash[i] = 0;
    }
}

template <class T>
void get_state(Handle* h, double* x, double* v, double* d, double* l, uint32_t* id, double* m) {
    State<T>& s = st<T>(h);
    const long n = (long)s.size();
    for (long i = 0; i < n; ++i) {
        if (x) { x[3 * i] = s.x[i].x; x[3 * i + 1] = s.x[i].y; x[3 * i + 2] = s.x[i].z; }
        if (v) { v[3 * i] = s.v[i].x; v[3 * i + 1] = s.v[i].y; v[3 * i + 2] = s.v[i].z; }
        if (d) d[i] = s.density[i];
        if (l) l[i] = s.lambda[i];
        if (id) id[i] = s.id[i];
        if (m) m[i] = s.mass[i];
    }
}

template <class T>
void get_stage(Handle* h, double* xs, double* vs, double* dp, double* w) {
    State<T>& s = st<T>(h);
    const long n = (long)s.size();
    for (long i = 0; i < n; ++i) {
        if (xs) { xs[3 * i] = s.xs[i].x; xs[3 * i + 1] = s.xs[i].y; xs[3 * i + 2] = s.xs[i].z; }
        if (vs) { vs[3 * i] = s.vs[i].x; vs[3 * i + 1] = s.vs[i].y; vs[3 * i + 2] = s.vs[i].z; }
        if (dp) { dp[3 * i] = s.dp[i].x; dp[3 * i + 1] = s.dp[i].y; dp[3 * i + 2] = s.dp[i].z; }
        if (w) { w[3 * i] = s.w[i].x; w[3 * i + 1] = s.w[i].y; w[3 * i + 2] = s.w[i].z; }
    }
}

template <class T>
void run_steps(Handle* h, float dt, const float* bmin, const float* bmax, int nsteps) {
    Params<T> p = make_params<T>(h->cfg, h->corr, dt, bmin, bmax, h->tableSize);
    p.bnd = boundary_for<T>(h, bmin, bmax);
    for (int k = 0; k < nsteps; ++k) step(st<T>(h), p, &h->taps);
}

// Stage ids follow PBFSolver::step's order (PBFSolver.cpp:29-77).
template <class T>
void run_stage(Handle* h, int stage, float dt, const float* bmin, const float* bmax) {
    Params<T> p = make_params<T>(h->cfg, h->corr, dt, bmin, bmax, h->tableSize);
    p.bnd = boundary_for<T>(h, bmin, bmax);
    State<T>& s = st<T>(h);
    Taps& t = h->taps;
    if (p.bnd && stage == 4) boundary_sums(s, p);
    switch (stage) {
        case 0: predict(s, p); break;
        case 1: compute_hashes(s, p); h->preSortHash = s.hash; break;
        case 2:
            sort_by_hash(s, t.perm);
            cell_ranges(s.hash, t.cell_hash, t.cell_start, t.cell_end);
            break;
        case 3: find_neighbours(s, p, t.cell_hash, t.cell_start, t.cell_end, t.nbr, t.cnt); break;
        case 4: densities(s, p, t.nbr, t.cnt); break;
        case 5: lambdas(s, p, t.nbr, t.cnt); break;
        case 6: position_deltas(s, p, t.nbr, t.cnt); break;
        case 7: correct_positions(s, p); break;
        case 8: update_position_velocity(s, p); break;
        case 9: boundary_damping(s, p, T(0.0f), T(0.95f)); break;
        case 10: vorticities(s, p, t.nbr, t.cnt); break;
        case 11: vorticity_confinement(s, p, t.nbr, t.cnt); break;
        case 12: {
            std::vector<V3<T>> vn;
            xsph(s, p, t.nbr, t.cnt, vn);
            break;
        }
        case 13: features(s, p, t.nbr, t.cnt, feat<T>(h)); break;
        default: break;
    }
}

}  // namespace

extern "C" {

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void* orc_create(int prec, long n, const Config* cfg, const LambdaCorr* corr, uint32_t tableSize) {
    auto* h = new Handle();
    h->prec = prec;
    h->cfg = *cfg;
    h->corr = *corr;
    h->tableSize = tableSize ? tableSize : (uint32_t)((uint64_t)cfg->maxNeighbours * (uint64_t)n);
    if (prec == 64) h->sd.resize(n); else h->sf.resize(n);
    return h;
}

void orc_destroy(void* hp) { delete static_cast<Handle*>(hp); }

void orc_set_gravity(void* hp, const float g[3]) {
    auto* h = static_cast<Handle*>(hp);
    for (int k = 0; k < 3; ++k) h->cfg.gravity[k] = g[k];
}

void orc_set_state(void* hp, long n, const float* x, const float* v, const float* m, const float* d,
                   const float* l, const uint32_t* id) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) set_state<double>(h, n, x, v, m, d, l, id);
    else set_state<float>(h, n, x, v, m, d, l, id);
}

void orc_get_state(void* hp, double* x, double* v, double* d, double* l, uint32_t* id, double* m) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) get_state<double>(h, x, v, d, l, id, m);
    else get_state<float>(h, x, v, d, l, id, m);
}

void orc_get_stage(void* hp, double* xs, double* vs, double* dp, double* w) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) get_stage<double>(h, xs, vs, dp, w);
    else get_stage<float>(h, xs, vs, dp, w);
}

void orc_step(void* hp, float dt, const float bmin[3], const float bmax[3], int nsteps) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) run_steps<double>(h, dt, bmin, bmax, nsteps);
    else run_steps<float>(h, dt, bmin, bmax, nsteps);
}

void orc_stage(void* hp, int stage, float dt, const float bmin[3], const float bmax[3]) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) run_stage<double>(h, stage, dt, bmin, bmax);
    else run_stage<float>(h, stage, dt, bmin, bmax);
}

// Features of the last "features" stage (16 per particle, sorted order).
void orc_get_features(void* hp, double* out) {
    auto* h = static_cast<Handle*>(hp);
    if (h->prec == 64) std::copy(h->featD.begin(), h->featD.end(), out);
    else for (size_t i = 0; i < h->featF.size(); ++i) out[i] = h->featF[i];
}

// Replaces x* (sorted order, 3 floats per particle): the learned step's warm start.
void orc_set_xs(void* hp, const float* xs) {
    auto* h = static_cast<Handle*>(hp);
    const long n = (long)(h->prec == 64 ? h->sd.size() : h->sf.size());
    for (long i = 0; i < n; ++i) {
        if (h->prec == 64) h->sd.xs[i] = {xs[3 * i], xs[3 * i + 1], xs[3 * i + 2]};
        else h->sf.xs[i] = {xs[3 * i], xs[3 * i + 1], xs[3 * i + 2]};
    }
}

// Replaces the vorticity (sorted order, 3 floats per particle): a confinement pass from a given
// vorticity field.
void orc_set_vort(void* hp, const float* w) {
    auto* h = static_cast<Handle*>(hp);
    const long n = (long)(h->prec == 64 ? h->sd.size() : h->sf.size());
    for (long i = 0; i < n; ++i) {
        if (h->prec == 64) h->sd.w[i] = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
        else h->sf.w[i] = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
    }
}

#ifdef ORACLE_PLANTS
// Turns planted mistake k on (0: none); see the list in pbf_oracle.hpp.
void orc_set_plant(int k) { planted_mistake() = k; }
#endif

// Boundary samples (layers 0: off; spacing <= 0: the configuration's particle spacing).
void orc_set_boundary(void* hp, int layers, float spacing) {
    auto* h = static_cast<Handle*>(hp);
    h->bndLayers = layers;
    h->bndSpacing = spacing;
    h->bndBuilt = false;
}

// The samples for this box: returns their count (-1: not admissible); pos (3 per sample) and psi
// receive up to `cap` samples in index order.  Either may be NULL.
long orc_boundary_samples(void* hp, const float bmin[3], const float bmax[3], double* pos, double* psi,
                          long cap) {
    auto* h = static_cast<Handle*>(hp);
    auto out = [&](auto* B) -> long {
        if (!B) return -1;
        const long m = (long)B->size();
        for (long k = 0; k < std::min(m, cap); ++k) {
            const auto x = boundary_sample(*B, (size_t)k);
            if (pos) { pos[3 * k] = x.x; pos[3 * k + 1] = x.y; pos[3 * k + 2] = x.z; }
            if (psi) psi[k] = B->psi[(size_t)k];
        }
        return m;
    };
    if (h->prec == 64) return out(boundary_for<double>(h, bmin, bmax));
    return out(boundary_for<float>(h, bmin, bmax));
}

// Boundary sums at x* of the current state (rhoB per particle, gB 3 per particle, sorted order).
void orc_boundary_sums(void* hp, const float bmin[3], const float bmax[3], double* rhoB, double* gB) {
    auto* h = static_cast<Handle*>(hp);
    auto run = [&](auto& s, const auto* B, auto tag) {
        using T = decltype(tag);
        if (!B) return;
        Params<T> p = make_params<T>(h->cfg, h->corr, 0.0f, bmin, bmax, h->tableSize);
        p.bnd = B;
        boundary_sums(s, p);
        for (size_t i = 0; i < s.size(); ++i) {
            if (rhoB) rhoB[i] = s.rhoB[i];
            if (gB) { gB[3 * i] = s.gB[i].x; gB[3 * i + 1] = s.gB[i].y; gB[3 * i + 2] = s.gB[i].z; }
        }
    };
    if (h->prec == 64) run(h->sd, boundary_for<double>(h, bmin, bmax), 0.0);
    else run(h->sf, boundary_for<float>(h, bmin, bmax), 0.0f);
}

long orc_ncells(void* hp) { return (long)static_cast<Handle*>(hp)->taps.cell_hash.size(); }

// Taps of the last step (or of the last stages run).  Any pointer may be NULL.
void orc_get_taps(void* hp, uint32_t* sorted_hash, uint32_t* perm, uint32_t* nbr, uint32_t* cnt,
                  uint32_t* cell_hash, uint32_t* cell_start, uint32_t* cell_end) {
    auto* h = static_cast<Handle*>(hp);
    const Taps& t = h->taps;
    const std::vector<uint32_t>& hs = (h->prec == 64) ? h->sd.hash : h->sf.hash;
    if (sorted_hash) std::copy(hs.begin(), hs.end(), sorted_hash);
    if (perm) std::copy(t.perm.begin(), t.perm.end(), perm);
    if (nbr) std::copy(t.nbr.begin(), t.nbr.end(), nbr);
    if (cnt) std::copy(t.cnt.begin(), t.cnt.end(), cnt);
    if (cell_hash) std::copy(t.cell_hash.begin(), t.cell_hash.end(), cell_hash);
    if (cell_start) std::copy(t.cell_start.begin(), t.cell_start.end(), cell_start);
    if (cell_end) std::copy(t.cell_end.begin(), t.cell_end.end(), cell_end);
}

// Exact density at the current positions (see pbf_oracle::true_density).  rho_sorted[k] is in
// the order a step would sort the particles; rho_by_id[id] by particle id.  Either may be NULL.
void orc_true_density(void* hp, double* rho_sorted, double* rho_by_id) {
    auto* h = static_cast<Handle*>(hp);
    const float z[3] = {0.0f, 0.0f, 0.0f};
    if (h->prec == 64) {
        Params<double> p = make_params<double>(h->cfg, h->corr, 0.0f, z, z, h->tableSize);
        std::vector<double> a, b;
        true_density(h->sd, p, a, b);
        if (rho_sorted) std::copy(a.begin(), a.end(), rho_sorted);
        if (rho_by_id) std::copy(b.begin(), b.end(), rho_by_id);
    } else {
        Params<float> p = make_params<float>(h->cfg, h->corr, 0.0f, z, z, h->tableSize);
        std::vector<float> a, b;
        true_density(h->sf, p, a, b);
        if (rho_sorted) for (size_t i = 0; i < a.size(); ++i) rho_sorted[i] = a[i];
        if (rho_by_id) for (size_t i = 0; i < b.size(); ++i) rho_by_id[i] = b[i];
    }
}

double orc_poly6(int prec, double d2, double h) {
    if (prec == 64) return poly6<double>(d2, h);
    return (double)poly6<float>((float)d2, (float)h);
}

void orc_grad_spiky(int prec, const double r[3], double h, double out[3]) {
    if (prec == 64) {
        V3<double> g = grad_spiky<double>({r[0], r[1], r[2]}, h);
        out[0] = g.x; out[1] = g.y; out[2] = g.z;
    } else {
        V3<float> g = grad_spiky<float>({(float)r[0], (float)r[1], (float)r[2]}, (float)h);
        out[0] = g.x; out[1] = g.y; out[2] = g.z;
    }
}

uint32_t orc_cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t tableSize) {
    return cell_hash(cx, cy, cz, tableSize);
}

}  // extern "C"
