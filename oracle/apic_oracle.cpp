// apic_oracle.cpp -- CPU restatement of the reference APIC step (src/apic.rs).
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the GPU library is
// checked against, and the CPU baseline timed beside it.  Only tests/, the
// smoke check and the benchmark's cpu_baseline leg load it; the product
// (libapic_hip) never links or calls it.
//
// One template restates every stage of Apic::update for R = float (the
// reference's own precision; "reference mode") and R = double (the FP64
// oracle).  Order of operations follows the reference line by line: it is
// compiled with -ffp-contract=off and uses std::fma exactly where the
// reference calls f32::mul_add (src/apic.rs:493,496,498,501,648,651,654,657,
// 1836,1839,1843,1846,1916).  wide's fast_max/fast_min are the x86
// maxps/minps selects `a > b ? a : b` / `a < b ? a : b`.
//
// Where the reference is nondeterministic (parallel CAS scatters, rayon
// reduction trees, rdst's unstable sort) this restatement fixes one valid
// order: chunks in ascending order for the P2G scatter, left-to-right sums of
// 8192-cell chunk partials for s.As, std::stable_sort for the Morton sort.
//
// Parity pins for each function are listed in tests/test_oracle_*.py; the
// functions without a closed-form or invariant check are named in the header
// of tests/test_oracle_invariants.py as "parity unpinned".
#include "../include/apic.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include <parallel/algorithm>

namespace {

constexpr float kSafety = 1e-5f;  // SAFTY, src/apic.rs:8

// Threaded mode (the CPU baseline): the reference's rayon loops as OpenMP
// loops over the same units (particle chunks, fluid-list slices, grid rows),
// parallel only above the reference's with_min_len thresholds, P2G scattering
// through CAS adds as the reference's atomic_add_f32/add_to_mac do.  With one
// thread (the default, the checker) every loop runs serially in the order the
// parity tests pin; with several the sums of the scatter and of the relative
// mode's reductions run in a thread-dependent order, as the reference's do.
#define ORC_PRAGMA(...) _Pragma(#__VA_ARGS__)
#define ORC_PAR(cond) ORC_PRAGMA(omp parallel for num_threads(nthreads) if(cond) schedule(static))
#define ORC_PAR_RED(cond, op, ...) \
    ORC_PRAGMA(omp parallel for num_threads(nthreads) if(cond) schedule(static) reduction(op : __VA_ARGS__))

// dst += v by compare-and-swap on the value's bits (atomic_add_f32, src/apic.rs:271-292)
template <class R> inline void cas_add(R& dst, R v) {
    using U = typename std::conditional<sizeof(R) == 4, uint32_t, uint64_t>::type;
    U* p = reinterpret_cast<U*>(&dst);
    U old = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (true) {
        R cur;
        std::memcpy(&cur, &old, sizeof(R));
        const R nv = cur + v;
        U nb;
        std::memcpy(&nb, &nv, sizeof(R));
        if (__atomic_compare_exchange_n(p, &old, nb, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return;
    }
}

template <class R> inline R fast_max(R a, R b) { return a > b ? a : b; }
template <class R> inline R fast_min(R a, R b) { return a < b ? a : b; }
// Rust f32::max/min (NaN-ignoring; operands here are never NaN).
template <class R> inline R rmax(R a, R b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a > b ? a : b)); }
template <class R> inline R rmin(R a, R b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a < b ? a : b)); }
template <class R> inline R rclamp(R x, R lo, R hi) { if (x < lo) x = lo; if (x > hi) x = hi; return x; }
inline int32_t trunc_i(double x) { return (int32_t)x; }  // cvttps2dq for in-range values

// expand_bits_simd / morton_code_simd, src/apic.rs:175-198
inline uint32_t expand_bits(uint32_t v) {
    v &= 0x0000FFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
inline uint32_t morton(uint32_t x, uint32_t y) { return expand_bits(x) | (expand_bits(y) << 1); }

// random_simd lane value, src/apic.rs:946-975 (lowbias32 hash of index ^ seed)
inline float random_lane(uint32_t index, uint32_t seed) {
    uint32_t m = index ^ seed;
    m ^= m >> 16;
    m *= 0x7feb352du;
    m ^= m >> 15;
    m *= 0x846ca68bu;
    m ^= m >> 16;
    uint32_t bits = (m & 0x007FFFFFu) | 0x3F800000u;
    float f;
    std::memcpy(&f, &bits, 4);
    return f - 1.5f;
}

// temporal_kernel_w_t, src/apic.rs:31-46
template <class R> inline R temporal_w(R tau) {
    R ts = tau - R(0.5);
    R tsq = ts * ts;
    R inner = fast_max(R(1) - tsq, R(0));
    R w = inner * inner * inner;
    w = w * R(2.1875);
    return fast_max(w, R(1e-5f));
}

// calculate_theta, src/apic.rs:1628-1638
template <class R> inline R theta(R rho, R rho_center, R target) {
    R threshold = target * R(0.5);
    R diff = rho_center - rho;
    if (std::fabs(diff) <= R(1e-5f)) return R(1);
    return rclamp((rho_center - threshold) / diff, R(0.1f), R(1.0f));
}

// dot_product, src/apic.rs:11-29: eight lane-strided partial sums, wide's
// reduce_add ((l0+l4)+(l2+l6))+((l1+l5)+(l3+l7)), then the scalar tail.
template <class R> R dot_product(const R* a, const R* b, size_t n) {
    R lane[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    size_t full = n / 8;
    for (size_t c = 0; c < full; ++c)
        for (int l = 0; l < 8; ++l) lane[l] = lane[l] + a[c * 8 + l] * b[c * 8 + l];
    R s04 = lane[0] + lane[4], s15 = lane[1] + lane[5], s26 = lane[2] + lane[6], s37 = lane[3] + lane[7];
    R result = (s04 + s26) + (s15 + s37);
    for (size_t i = full * 8; i < n; ++i) result += a[i] * b[i];
    return result;
}

template <class R> struct Level {
    uint32_t w = 0, h = 0, cells = 0;
    size_t alloc = 0;
    std::vector<uint8_t> fluid;  // type_fluid bit per raw cell
    std::vector<uint32_t> fluid_idx, red, black;  // padded indices
    std::vector<R> diag, inv_diag, px, py, r, e, cg_s, cg_ap, cg_r;
    int nthreads = 1;

    void init(uint32_t W, uint32_t H) {  // MultiGridLevel::new, src/apic.rs:428-453
        w = W; h = H; cells = W * H;
        size_t padded = (size_t)cells + 2 * (size_t)W;
        alloc = (padded + 7) / 8 * 8;
        fluid.assign(cells, 0);
        diag.assign(alloc, 0); inv_diag.assign(alloc, 0); px.assign(alloc, 0); py.assign(alloc, 0);
        r.assign(alloc, 0); e.assign(alloc, 0);
    }
    void allocate_cg() { cg_s.assign(alloc, 0); cg_ap.assign(alloc, 0); cg_r.assign(alloc, 0); }  // :455-461
    bool is_fluid(size_t raw) const { return fluid[raw] != 0; }

    // smooth_level, src/apic.rs:471-535 (one iteration, omega = 1.2)
    void smooth_list(const std::vector<uint32_t>& list, R omega) {
        const long n = (long)list.size();
        ORC_PAR(nthreads > 1 && n > 4096)
        for (long li = 0; li < n; ++li) {
            size_t c = list[(size_t)li];
            R id = inv_diag[c];
            if (id != R(0)) {
                R sum = r[c];
                sum = std::fma(-px[c], e[c + 1], sum);
                sum = std::fma(-py[c], e[c + w], sum);
                sum = std::fma(-px[c - 1], e[c - 1], sum);
                sum = std::fma(-py[c - w], e[c - w], sum);
                R gs = sum * id;
                R old = e[c];
                e[c] = old + (omega * (gs - old));
            }
        }
    }
    void smooth(bool reverse, R omega) {
        if (reverse) { smooth_list(black, omega); smooth_list(red, omega); }
        else { smooth_list(red, omega); smooth_list(black, omega); }
    }

    // restrict, src/apic.rs:537-594.  A child past the edge of an odd-sided
    // level (the relative mode's levels below the reference's last, whose
    // sides are the finer ones' halves rounded up) is skipped, as
    // update_multilevel_grid skips it (:1566); the reference's levels have
    // even sides and never meet one.
    void restrict_to(Level& coarse) const {
        std::fill(coarse.e.begin() + coarse.w, coarse.e.begin() + coarse.w + coarse.cells, R(0));
        const long n = (long)coarse.fluid_idx.size();
        ORC_PAR(nthreads > 1 && n > 4096)
        for (long li = 0; li < n; ++li) {
            const uint32_t ci = coarse.fluid_idx[(size_t)li];
            size_t raw_c = ci - coarse.w;
            size_t cx = raw_c % coarse.w, cy = raw_c / coarse.w;
            R sum = 0;
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    size_t fx = cx * 2 + dx, fy = cy * 2 + dy;
                    if (fx >= w || fy >= h) continue;
                    size_t raw_f = fy * w + fx, pf = raw_f + w;
                    if (is_fluid(raw_f)) {
                        R b = r[pf];
                        R ax = diag[pf] * e[pf];
                        ax += px[pf] * e[pf + 1];
                        ax += py[pf] * e[pf + w];
                        ax += px[pf - 1] * e[pf - 1];
                        ax += py[pf - w] * e[pf - w];
                        sum += b - ax;
                    }
                }
            coarse.r[ci] = sum;
        }
    }

    // prolongate, src/apic.rs:596-616 (iterates the fluid list including the
    // base level's padding duplicates)
    void prolongate_from(const Level& coarse) {
        const long n = (long)fluid_idx.size();
        // (the padded list's duplicates of the last cell are adjacent: one thread takes them)
        ORC_PAR(nthreads > 1 && n > 4096)
        for (long li = 0; li < n; ++li) {
            const uint32_t fi = fluid_idx[(size_t)li];
            size_t raw_f = fi - w;
            size_t cx = (raw_f % w) / 2, cy = (raw_f / w) / 2;
            size_t rc = cy * coarse.w + cx + coarse.w;
            e[fi] += coarse.e[rc];
        }
    }

    // solve_coarse_exact, src/apic.rs:618-704
    void solve_coarse_exact() {
        size_t start = w, end = start + cells;
        std::fill(e.begin() + start, e.begin() + end, R(0));
        std::copy(r.begin() + start, r.begin() + end, cg_r.begin() + start);
        std::copy(r.begin() + start, r.begin() + end, cg_s.begin() + start);
        R gamma = 0;
        for (uint32_t i : fluid_idx) gamma += cg_r[i] * cg_r[i];
        if (gamma < R(1e-12f)) return;
        size_t maxit = std::min<size_t>(fluid_idx.size(), 100);
        for (size_t it = 0; it < maxit; ++it) {
            R delta = 0;
            for (uint32_t i : fluid_idx) {
                R v = diag[i] * cg_s[i];
                v = std::fma(px[i], cg_s[i + 1], v);
                v = std::fma(px[i - 1], cg_s[i - 1], v);
                v = std::fma(py[i], cg_s[i + w], v);
                v = std::fma(py[i - w], cg_s[i - w], v);
                cg_ap[i] = v;
                delta += cg_s[i] * v;
            }
            if (std::fabs(delta) < R(kSafety)) break;
            R alpha = gamma / delta;
            R maxe = 0;
            for (uint32_t i : fluid_idx) {
                e[i] += alpha * cg_s[i];
                cg_r[i] -= alpha * cg_ap[i];
                maxe = rmax(maxe, std::fabs(cg_r[i]));
            }
            if (maxe < R(1e-6f)) break;
            R ng = 0;
            for (uint32_t i : fluid_idx) ng += cg_r[i] * cg_r[i];
            R beta = ng / gamma;
            gamma = ng;
            for (uint32_t i : fluid_idx) cg_s[i] = cg_r[i] + (beta * cg_s[i]);
        }
    }

    // The relative mode's coarse CG: the iteration of solve_coarse_exact
    // above (FP32 vectors, the same SpMV chain and cap) with FP64 dot
    // products, a search vector that is zero on non-fluid cells (their
    // couplings are zero), and exits relative to the level's right-hand side
    // instead of absolute ones: r.r = 0 at the start, s.As <= 0, and
    // ||r|| <= 1e-2 ||r0||.  With absolute exits the coarse correction
    // vanishes once the residual is small (|s.As| < 1e-5 at the first step),
    // and the outer CG slows down, or stalls on a large coarsest level (the
    // reference's coarsest on a grid side of 2 x odd).  The GPU sums in
    // another order (a wave's or a workgroup's tree), which changes the FP32
    // step lengths only where an FP64 rounding tips them.
    void solve_coarse_relative() {
        const size_t start = w, end = start + cells;
        std::fill(e.begin() + start, e.begin() + end, R(0));
        std::copy(r.begin() + start, r.begin() + end, cg_r.begin() + start);
        std::fill(cg_s.begin(), cg_s.end(), R(0));
        double gamma = 0;
        for (uint32_t i : fluid_idx) gamma += double(cg_r[i]) * double(cg_r[i]);
        if (!(gamma > 0)) return;
        const double stop = 1e-4 * gamma;
        const size_t maxit = std::min<size_t>(fluid_idx.size(), 100);
        R beta = 0;
        for (size_t it = 0; it < maxit; ++it) {
            for (uint32_t i : fluid_idx) cg_s[i] = it == 0 ? cg_r[i] : cg_r[i] + (beta * cg_s[i]);
            double delta = 0;
            for (uint32_t i : fluid_idx) {
                R v = diag[i] * cg_s[i];
                v = std::fma(px[i], cg_s[i + 1], v);
                v = std::fma(px[i - 1], cg_s[i - 1], v);
                v = std::fma(py[i], cg_s[i + w], v);
                v = std::fma(py[i - w], cg_s[i - w], v);
                cg_ap[i] = v;
                delta += double(cg_s[i]) * double(v);
            }
            if (!(delta > 0)) break;
            const R alpha = R(gamma / delta);
            double ng = 0;
            for (uint32_t i : fluid_idx) {
                e[i] += alpha * cg_s[i];
                cg_r[i] = cg_r[i] - alpha * cg_ap[i];
                ng += double(cg_r[i]) * double(cg_r[i]);
            }
            if (ng <= stop) break;
            beta = R(ng / gamma);
            gamma = ng;
        }
    }
};

enum Pf { PX_ = 0, PY_, VX_, VY_, CUX_, CUY_, CVX_, CVY_, TR_, NPF_ };

template <class R> struct Sim {
    R gravity, border_damping, cfl, solve_error;
    R gravity_x = R(0);  // horizontal gravity: the library's extension (zero in the reference)
    uint32_t W = 0, H = 0, cells = 0;
    R cell = 1;
    uint32_t num_chunks = 0, num_particles = 0;
    size_t nalloc = 0;  // num_chunks * 8

    std::vector<R> ext_u, ext_v, post_u, post_v;
    std::vector<R> smoothed, old_density, pressure, density;
    std::vector<uint8_t> obstacle;
    std::vector<uint8_t> valid_u, valid_v, next_valid_u, next_valid_v;
    std::vector<R> mac_u, mac_v, old_u, old_v;
    std::vector<R> u_mom, u_w, v_mom, v_w;  // the two halves of mac_{u,v}_and_weight
    std::vector<uint8_t> open_u, open_v;     // mac_{u,v}_blocked bits (1 = open)
    std::vector<R> search, As;
    std::vector<R> p[NPF_], ps[NPF_];
    std::vector<uint32_t> skey, sval;
    uint32_t fluid_cells = 0;
    Level<R> base;
    std::vector<Level<R>> levels;  // the reference's levels, then the relative mode's deeper ones
    size_t ref_levels = 0;         // how many of them the reference's halving makes
    // the levels the current solver mode builds and cycles through
    size_t active_levels() const { return solver_mode == 1 ? levels.size() : ref_levels; }
    uint32_t timestamp = 0;
    R prev_dt = R(1.0f / 60.0f);

    // statistics
    apic_step_stats stats{};
    uint32_t p2g_shuffle_seed = 0;  // nonzero: scatter chunks in a seeded random order
    int nthreads = 1;               // threaded mode (the CPU baseline) when > 1
    void set_threads(int n) {
        nthreads = n < 1 ? 1 : n;
        base.nthreads = nthreads;
        for (auto& L : levels) L.nthreads = nthreads;
    }

    size_t nu() const { return (size_t)cells + H; }
    size_t nv() const { return (size_t)cells + W; }

    uint32_t u_index(int32_t i, int32_t j) const {  // :977-982
        uint32_t ci = (uint32_t)std::max(std::min(i, (int32_t)W), 0);
        uint32_t cj = (uint32_t)std::max(std::min(j, (int32_t)H - 1), 0);
        return cj * (W + 1) + ci;
    }
    uint32_t v_index(int32_t i, int32_t j) const {  // :985-991
        uint32_t ci = (uint32_t)std::max(std::min(i, (int32_t)W - 1), 0);
        uint32_t cj = (uint32_t)std::max(std::min(j, (int32_t)H), 0);
        return cj * W + ci;
    }
    int is_obstacle(int32_t i, int32_t j) const {  // :993-1003
        if (i < 0 || i >= (int32_t)W || j < 0 || j >= (int32_t)H) return 1;
        return obstacle[(size_t)j * W + i];
    }
    int is_fluid(int32_t i, int32_t j) const {  // :1014-1022
        if (i < 0 || i >= (int32_t)W || j < 0 || j >= (int32_t)H) return 0;
        return base.fluid[(size_t)j * W + i];
    }

    // Apic::new, src/apic.rs:779-916 (+ explicit particle replacement)
    int create(uint32_t w, uint32_t h, float cell_size, const apic_world_properties& wp,
               const float* xy, uint32_t n) {
        gravity = wp.gravity; border_damping = wp.border_damping; cfl = wp.cfl; solve_error = wp.solve_error;
        W = w; H = h; cells = w * h; cell = cell_size;
        base.init(W, H);
        size_t c8 = ((size_t)cells + 7) / 8 * 8;
        pressure.assign(c8, 0); density.assign(c8, 0); smoothed.assign(c8, 0); old_density.assign(c8, 0);
        obstacle.assign(cells, 0);
        valid_u.assign(nu(), 0); valid_v.assign(nv(), 0); next_valid_u.assign(nu(), 0); next_valid_v.assign(nv(), 0);
        mac_u.assign(nu(), 0); mac_v.assign(nv(), 0); old_u.assign(nu(), 0); old_v.assign(nv(), 0);
        u_mom.assign(nu(), 0); u_w.assign(nu(), 0); v_mom.assign(nv(), 0); v_w.assign(nv(), 0);
        open_u.assign(nu(), 0); open_v.assign(nv(), 0);
        search.assign(base.alloc, 0); As.assign(base.alloc, 0);
        ext_u.assign(nu(), 0); ext_v.assign(nv(), 0); post_u.assign(nu(), 0); post_v.assign(nv(), 0);
        prev_dt = R(1.0f / 60.0f);
        timestamp = 0;

        std::vector<float> pos;  // initial positions are computed in f32 for both precisions
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                if (x == 0 || x == W - 1 || y == 0 || y == H - 1) {
                    obstacle[(size_t)y * W + x] = 1;
                } else if (!xy) {
                    if (x > 4 && x < W - 4 && y > 4 && y < H / 2) {
                        for (int py = 0; py < 3; ++py)
                            for (int px = 0; px < 3; ++px) {
                                float fx = ((float)x + ((float)px + 0.5f) / 3.0f) * cell_size;
                                float fy = ((float)y + ((float)py + 0.5f) / 3.0f) * cell_size;
                                pos.push_back(fx);
                                pos.push_back(fy);
                            }
                    }
                }
            }
        if (xy) pos.assign(xy, xy + 2 * (size_t)n);
        num_particles = (uint32_t)(pos.size() / 2);
        if (num_particles == 0) {  // :851-855 returns before levels and update_world
            num_chunks = 0; nalloc = 0;
            for (int f = 0; f < NPF_; ++f) { p[f].clear(); ps[f].clear(); }
            return 0;
        }
        num_chunks = (num_particles + 7) / 8;
        nalloc = (size_t)num_chunks * 8;
        for (int f = 0; f < NPF_; ++f) { p[f].assign(nalloc, 0); ps[f].assign(nalloc, 0); }
        skey.assign(num_particles, 0); sval.assign(num_particles, 0);
        for (size_t i = 0; i < nalloc; ++i) {
            size_t s = std::min<size_t>(i, num_particles - 1);  // tail lanes repeat the last particle (:887-898)
            p[PX_][i] = (R)pos[2 * s];
            p[PY_][i] = (R)pos[2 * s + 1];
        }
        uint32_t cw = W, ch = H;
        levels.clear();
        while (cw % 2 == 0 && ch % 2 == 0 && cw > 4 && ch > 4) {  // :900-908
            cw /= 2; ch /= 2;
            levels.emplace_back();
            levels.back().init(cw, ch);
        }
        if (levels.empty()) return APIC_ERR_UNSUPPORTED_SIZE;  // :910 panics
        levels.back().allocate_cg();
        // The relative mode's hierarchy goes on below the reference's last
        // level while it has more than 64 cells and both sides exceed 2, each
        // side halved and rounded up (a grid side of 2 x odd stops the
        // reference's halving at up to 511^2 cells)
        ref_levels = levels.size();
        while ((size_t)cw * ch > 64 && cw > 2 && ch > 2) {
            cw = (cw + 1) / 2; ch = (ch + 1) / 2;
            levels.emplace_back();
            levels.back().init(cw, ch);
        }
        if (levels.size() > ref_levels) levels.back().allocate_cg();
        update_world();
        return 0;
    }

    void update_world() {  // :918-944
        for (size_t idx = 0; idx < nu(); ++idx) {
            int32_t i = (int32_t)(idx % (W + 1)), j = (int32_t)(idx / (W + 1));
            open_u[idx] = (is_obstacle(i - 1, j) | is_obstacle(i, j)) == 0;
        }
        for (size_t idx = 0; idx < nv(); ++idx) {
            int32_t i = (int32_t)(idx % W), j = (int32_t)(idx / W);
            open_v[idx] = (is_obstacle(i, j - 1) | is_obstacle(i, j)) == 0;
        }
    }

    // update_spatial_lookup, src/apic.rs:1043-1111
    void compute_keys() {
        R max_x = (R)(float)(W > 0 ? W - 1 : 0), max_y = (R)(float)(H > 0 ? H - 1 : 0);
        ORC_PAR(nthreads > 1 && num_particles > 4096)
        for (uint32_t i = 0; i < num_particles; ++i) {
            R gx = p[PX_][i] / cell, gy = p[PY_][i] / cell;
            uint32_t ix = (uint32_t)trunc_i(fast_min(fast_max(gx, R(0)), max_x));
            uint32_t iy = (uint32_t)trunc_i(fast_min(fast_max(gy, R(0)), max_y));
            skey[i] = morton(ix, iy);
            sval[i] = i;
        }
    }
    void sort_records() {
        std::vector<uint32_t> order(num_particles);
        std::iota(order.begin(), order.end(), 0u);
        auto less = [&](uint32_t a, uint32_t b) { return skey[a] < skey[b]; };
        if (nthreads > 1)  // a parallel multiway merge sort: the same stable order
            __gnu_parallel::stable_sort(order.begin(), order.end(), less,
                                        __gnu_parallel::multiway_mergesort_tag(nthreads));
        else
            std::stable_sort(order.begin(), order.end(), less);
        std::vector<uint32_t> k(num_particles), v(num_particles);
        ORC_PAR(nthreads > 1 && num_particles > 4096)
        for (uint32_t i = 0; i < num_particles; ++i) { k[i] = skey[order[i]]; v[i] = sval[order[i]]; }
        skey.swap(k); sval.swap(v);
    }
    void permute() {
        for (int f = 0; f < NPF_; ++f) {
            ORC_PAR(nthreads > 1 && nalloc > 8192)
            for (size_t i = 0; i < nalloc; ++i) ps[f][i] = i < num_particles ? p[f][sval[i]] : R(0);
        }
        for (int f = 0; f < NPF_; ++f) p[f].swap(ps[f]);
    }
    void update_spatial_lookup() { compute_keys(); sort_records(); permute(); }

    // transfer_particles_to_grid, src/apic.rs:1218-1497
    void transfer_particles_to_grid() {
        std::fill(base.fluid.begin(), base.fluid.end(), 0);
        base.red.clear(); base.black.clear(); base.fluid_idx.clear();
        std::fill(mac_u.begin(), mac_u.end(), R(0)); std::fill(mac_v.begin(), mac_v.end(), R(0));
        std::fill(u_mom.begin(), u_mom.end(), R(0)); std::fill(u_w.begin(), u_w.end(), R(0));
        std::fill(v_mom.begin(), v_mom.end(), R(0)); std::fill(v_w.begin(), v_w.end(), R(0));
        std::fill(density.begin(), density.end(), R(0));
        std::fill(valid_u.begin(), valid_u.end(), 0); std::fill(valid_v.begin(), valid_v.end(), 0);

        std::vector<uint32_t> chunk_order(num_chunks);
        std::iota(chunk_order.begin(), chunk_order.end(), 0u);
        if (p2g_shuffle_seed) {
            std::mt19937 rng(p2g_shuffle_seed);
            std::shuffle(chunk_order.begin(), chunk_order.end(), rng);
        }
        const R half = R(0.5), one = R(1);
        const bool par = nthreads > 1 && num_chunks > 256;  // par_chunks, with_min_len(256) (:1237-1240)
        ORC_PAR(par)
        for (long ci = 0; ci < (long)num_chunks; ++ci) {
            const uint32_t chunk = chunk_order[(size_t)ci];
            uint32_t idx_u[4][8], idx_v[4][8], idx_c[4][8], idx_f[8];
            R add_u[4][8], wt_u[4][8], add_v[4][8], wt_v[4][8], wt_c[4][8];
            for (int lane = 0; lane < 8; ++lane) {
                size_t i = (size_t)chunk * 8 + lane;
                R x = p[PX_][i], y = p[PY_][i];
                R vx = p[VX_][i], vy = p[VY_][i];
                R tau = (-p[TR_][i]) / prev_dt;
                R wt = temporal_w(tau);
                R cux = p[CUX_][i], cuy = p[CUY_][i], cvx = p[CVX_][i], cvy = p[CVY_][i];
                R gx = x / cell, gy = y / cell;
                // density stencil (:1260-1284)
                R cgx = gx - half, cgy = gy - half;
                R cbx = std::floor(cgx), cby = std::floor(cgy);
                R ctx = cgx - cbx, cty = cgy - cby;
                wt_c[0][lane] = ((one - ctx) * (one - cty)) * wt;
                wt_c[1][lane] = (ctx * (one - cty)) * wt;
                wt_c[2][lane] = ((one - ctx) * cty) * wt;
                wt_c[3][lane] = (ctx * cty) * wt;
                int32_t mx = (int32_t)W - 1, my = (int32_t)H - 1;
                int32_t cx0 = std::max(std::min(trunc_i(cbx), mx), 0);
                int32_t cx1 = std::max(std::min(trunc_i(cbx) + 1, mx), 0);
                int32_t cy0 = std::max(std::min(trunc_i(cby), my), 0);
                int32_t cy1 = std::max(std::min(trunc_i(cby) + 1, my), 0);
                idx_c[0][lane] = (uint32_t)(cy0 * (int32_t)W + cx0);
                idx_c[1][lane] = (uint32_t)(cy0 * (int32_t)W + cx1);
                idx_c[2][lane] = (uint32_t)(cy1 * (int32_t)W + cx0);
                idx_c[3][lane] = (uint32_t)(cy1 * (int32_t)W + cx1);
                // u grid (get_u_grid :1113-1154, get_grid_distances :1199-1216)
                {
                    R ugx = gx, ugy = gy - half;
                    R bx = std::floor(ugx), by = std::floor(ugy);
                    int32_t ix = trunc_i(bx), iy = trunc_i(by);
                    idx_u[0][lane] = u_index(ix, iy); idx_u[1][lane] = u_index(ix + 1, iy);
                    idx_u[2][lane] = u_index(ix, iy + 1); idx_u[3][lane] = u_index(ix + 1, iy + 1);
                    R tx = ugx - bx, ty = ugy - by;
                    R itx = one - tx, ity = one - ty;
                    R w4[4] = {itx * ity, tx * ity, itx * ty, tx * ty};
                    R dxs[4] = {(-tx) * cell, itx * cell, (-tx) * cell, itx * cell};
                    R dys[4] = {(-ty) * cell, (-ty) * cell, ity * cell, ity * cell};
                    for (int k = 0; k < 4; ++k) {
                        R ww = w4[k] * wt;
                        wt_u[k][lane] = ww;
                        add_u[k][lane] = ((vx + cux * dxs[k]) + cuy * dys[k]) * ww;
                    }
                }
                // v grid (get_v_grid :1156-1197)
                {
                    R vgx = gx - half, vgy = gy;
                    R bx = std::floor(vgx), by = std::floor(vgy);
                    R tx = vgx - bx, ty = vgy - by;
                    R itx = one - tx, ity = one - ty;
                    R w4[4] = {itx * ity, tx * ity, itx * ty, tx * ty};
                    int32_t ix = trunc_i(bx), iy = trunc_i(by);
                    idx_v[0][lane] = v_index(ix, iy); idx_v[1][lane] = v_index(ix + 1, iy);
                    idx_v[2][lane] = v_index(ix, iy + 1); idx_v[3][lane] = v_index(ix + 1, iy + 1);
                    R dxs[4] = {(-tx) * cell, itx * cell, (-tx) * cell, itx * cell};
                    R dys[4] = {(-ty) * cell, (-ty) * cell, ity * cell, ity * cell};
                    for (int k = 0; k < 4; ++k) {
                        R ww = w4[k] * wt;
                        wt_v[k][lane] = ww;
                        add_v[k][lane] = ((vy + cvx * dxs[k]) + cvy * dys[k]) * ww;
                    }
                }
                // fluid cell (:1358-1360)
                R clx = fast_max(fast_min(gx, (R)((float)W - 1.0f)), R(0));
                R cly = fast_max(fast_min(gy, (R)((float)H - 1.0f)), R(0));
                idx_f[lane] = (uint32_t)(trunc_i(cly) * (int32_t)W + trunc_i(clx));
            }
            int active = (chunk == num_chunks - 1 && num_particles % 8 != 0) ? (int)(num_particles % 8) : 8;
            for (int l = 0; l < active; ++l)  // add_to_fluid :371-395 (fetch_or)
                if (par) __atomic_store_n(&base.fluid[idx_f[l]], (uint8_t)1, __ATOMIC_RELAXED);
                else base.fluid[idx_f[l]] = 1;
            for (int k = 0; k < 4; ++k) add_to_mac(u_mom, u_w, idx_u[k], add_u[k], wt_u[k], active, par);
            for (int k = 0; k < 4; ++k) add_to_mac(v_mom, v_w, idx_v[k], add_v[k], wt_v[k], active, par);
            for (int k = 0; k < 4; ++k) add_to_vec(density, idx_c[k], wt_c[k], active, par);
        }
        // normalisation (:1428-1459; rayon::join of u and v, with_min_len(4096))
        {
            const size_t nu_ = nu(), nv_ = nv();
            ORC_PAR(nthreads > 1 && nu_ > 4096)
            for (size_t i = 0; i < nu_; ++i)
                if (u_w[i] > R(0)) { valid_u[i] = 1; mac_u[i] = u_mom[i] / u_w[i]; }
            ORC_PAR(nthreads > 1 && nv_ > 4096)
            for (size_t i = 0; i < nv_; ++i)
                if (v_w[i] > R(0)) { valid_v[i] = 1; mac_v[i] = v_mom[i] / v_w[i]; }
        }
        // fluid lists (:1461-1494)
        fluid_cells = 0;
        for (size_t raw = 0; raw < cells; ++raw) {
            if (obstacle[raw]) base.fluid[raw] = 0;
            if (!base.fluid[raw]) continue;
            uint32_t pidx = (uint32_t)(raw + W);
            base.fluid_idx.push_back(pidx);
            ++fluid_cells;
            size_t x = raw % W, y = raw / W;
            ((x + y) % 2 == 0 ? base.red : base.black).push_back(pidx);
        }
        uint32_t rem = fluid_cells % 8;
        if (rem) {
            uint32_t last = base.fluid_idx[fluid_cells - 1];
            for (uint32_t k = 0; k < 8 - rem; ++k) base.fluid_idx.push_back(last);
        }
    }
    // add_to_mac, src/apic.rs:295-344: lanes hitting the same face are summed
    // in lane order first, then added to the face.
    static void add_to_mac(std::vector<R>& mom, std::vector<R>& wsum, const uint32_t* idx, const R* val,
                           const R* wt, int active, bool atomic) {
        unsigned processed = 0;
        for (int l = 0; l < active; ++l) {
            if (processed & (1u << l)) continue;
            uint32_t t = idx[l];
            R value = val[l], weight = wt[l];
            for (int o = l + 1; o < active; ++o)
                if (idx[o] == t) { value += val[o]; weight += wt[o]; processed |= 1u << o; }
            if (value == R(0) && weight == R(0)) continue;
            if (atomic) { cas_add(mom[t], value); cas_add(wsum[t], weight); }
            else { mom[t] = mom[t] + value; wsum[t] = wsum[t] + weight; }
        }
    }
    // add_to_vec_f32 + atomic_add_f32, src/apic.rs:271-292, 346-369
    static void add_to_vec(std::vector<R>& dst, const uint32_t* idx, const R* wt, int active, bool atomic) {
        unsigned processed = 0;
        for (int l = 0; l < active; ++l) {
            if (processed & (1u << l)) continue;
            uint32_t t = idx[l];
            R weight = wt[l];
            for (int o = l + 1; o < active; ++o)
                if (idx[o] == t) { weight += wt[o]; processed |= 1u << o; }
            if (weight == R(0)) continue;
            if (atomic) cas_add(dst[t], weight);
            else dst[t] = dst[t] + weight;
        }
    }

    // update_multilevel_grid, src/apic.rs:1499-1602
    void update_multilevel_grid() {
        uint32_t width = W;
        for (size_t gi = 0; gi < active_levels(); ++gi) {
            const Level<R>& fine = gi == 0 ? base : levels[gi - 1];
            Level<R>& coarse = levels[gi];
            coarse.fluid_idx.clear(); coarse.red.clear(); coarse.black.clear();
            std::fill(coarse.fluid.begin(), coarse.fluid.end(), 0);
            size_t cs = coarse.w, ce = cs + coarse.cells;
            std::fill(coarse.diag.begin() + cs, coarse.diag.begin() + ce, R(0));
            std::fill(coarse.inv_diag.begin() + cs, coarse.inv_diag.begin() + ce, R(0));
            std::fill(coarse.px.begin() + cs, coarse.px.begin() + ce, R(0));
            std::fill(coarse.py.begin() + cs, coarse.py.begin() + ce, R(0));
            uint32_t nw = coarse.w;
            for (uint32_t ci : fine.fluid_idx) {
                uint32_t raw = ci - width;
                uint32_t nx = (raw % width) / 2, ny = (raw / width) / 2;
                coarse.fluid[(size_t)ny * nw + nx] = 1;
            }
            for (uint32_t raw = 0; raw < coarse.cells; ++raw) {
                if (!coarse.fluid[raw]) continue;
                uint32_t pidx = raw + nw;
                coarse.fluid_idx.push_back(pidx);
                uint32_t x = raw % nw, y = raw / nw;
                ((x + y) % 2 == 0 ? coarse.red : coarse.black).push_back(pidx);
            }
            for (uint32_t ci : coarse.fluid_idx) {
                size_t cx = (ci - coarse.w) % coarse.w, cy = (ci - coarse.w) / coarse.w;
                R diag = 0, plus_x = 0, plus_y = 0;
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        size_t fx = cx * 2 + dx, fy = cy * 2 + dy;
                        if (fx >= fine.w || fy >= fine.h) continue;
                        size_t fr = fy * fine.w + fx, fp = fr + fine.w;
                        if (fine.is_fluid(fr)) {
                            diag += fine.diag[fp];
                            if (dx == 0) diag += R(2) * fine.px[fp];
                            if (dy == 0) diag += R(2) * fine.py[fp];
                            if (dx == 1) plus_x += fine.px[fp];
                            if (dy == 1) plus_y += fine.py[fp];
                        }
                    }
                R fd = diag * R(0.5);
                coarse.diag[ci] = fd;
                coarse.inv_diag[ci] = fd > R(1e-6f) ? R(1) / fd : R(0);
                coarse.px[ci] = plus_x * R(0.5);
                coarse.py[ci] = plus_y * R(0.5);
            }
            width = nw;
        }
    }

    // apply_external_forces, src/apic.rs:1604-1626
    void apply_external_forces(R dt) {
        R g = gravity * dt;
        for (size_t i = 0; i < nu(); ++i) mac_u[i] += ext_u[i] * dt;
        for (size_t i = 0; i < nv(); ++i) mac_v[i] += ext_v[i] * dt;
        for (int32_t y = 0; y <= (int32_t)H; ++y)
            for (int32_t x = 0; x < (int32_t)W; ++x) {
                uint32_t f = v_index(x, y);
                if (is_fluid(x, y - 1) == 1 || is_fluid(x, y) == 1) mac_v[f] += g;
            }
        // the library's horizontal gravity (an extension): the same rule on u faces
        if (gravity_x != R(0)) {
            R gx = gravity_x * dt;
            for (int32_t y = 0; y < (int32_t)H; ++y)
                for (int32_t x = 0; x <= (int32_t)W; ++x) {
                    uint32_t f = u_index(x, y);
                    if (is_fluid(x - 1, y) == 1 || is_fluid(x, y) == 1) mac_u[f] += gx;
                }
        }
    }

    // build_pressure_system, src/apic.rs:1640-1785
    void build_pressure_system() {
        size_t start = W, end = start + cells;
        std::fill(base.diag.begin() + start, base.diag.begin() + end, R(0));
        std::fill(base.inv_diag.begin() + start, base.inv_diag.begin() + end, R(0));
        std::fill(base.px.begin() + start, base.px.begin() + end, R(0));
        std::fill(base.py.begin() + start, base.py.begin() + end, R(0));
        smoothed = density;
        old_density = density;
        const size_t w = W;
        const bool par = nthreads > 1 && base.fluid_idx.size() > 4096;  // (:1683-1684, 1779-1781)
        for (int pass = 0; pass < 3; ++pass) {
            smoothed.swap(old_density);
            ORC_PAR(par)
            for (uint32_t k = 0; k < fluid_cells; ++k) {
                size_t raw = base.fluid_idx[k] - w;
                R sum = old_density[raw] * R(4);
                R weight = 4;
                R fl = (R)base.fluid[raw - 1], fr = (R)base.fluid[raw + 1];
                R fb = (R)base.fluid[raw - w], ft = (R)base.fluid[raw + w];
                sum += old_density[raw - 1] * fl;
                sum += old_density[raw + 1] * fr;
                sum += old_density[raw - w] * fb;
                sum += old_density[raw + w] * ft;
                weight += ((fl + fr) + fb) + ft;
                smoothed[raw] = sum / weight;
            }
        }
        const R target = 9, noise = target * R(1.15f), rate = R(0.1f);
        ORC_PAR(par)
        for (uint32_t k = 0; k < fluid_cells; ++k) {
            size_t fi = base.fluid_idx[k], raw = fi - w;
            size_t gx = raw % w, gy = raw / w, ui = gy * (w + 1) + gx;
            R ul = mac_u[ui], ur = mac_u[ui + 1], vb = mac_v[raw], vt = mac_v[raw + w];
            int nol = 1 - obstacle[raw - 1], nor = 1 - obstacle[raw + 1];
            int nob = 1 - obstacle[raw - w], not_ = 1 - obstacle[raw + w];
            ul *= (R)nol; ur *= (R)nor; vb *= (R)nob; vt *= (R)not_;
            R div = (ur - ul) + (vt - vb);
            R rc = smoothed[raw];
            R excess = rmin(rmax(rc - noise, R(0)), target);
            int fl = base.fluid[raw - 1], fr = base.fluid[raw + 1], fb = base.fluid[raw - w], ft = base.fluid[raw + w];
            R diag = 0, px = 0, py = 0;
            if (nol == 1) diag += fl == 1 ? R(1) : R(1) / theta(smoothed[raw - 1], rc, target);
            if (nor == 1) {
                if (fr == 1) { diag += R(1); px = R(-1); }
                else diag += R(1) / theta(smoothed[raw + 1], rc, target);
            }
            if (nob == 1) diag += fb == 1 ? R(1) : R(1) / theta(smoothed[raw - w], rc, target);
            if (not_ == 1) {
                if (ft == 1) { diag += R(1); py = R(-1); }
                else diag += R(1) / theta(smoothed[raw + w], rc, target);
            }
            base.r[fi] = (-div) * cell + ((excess * rate) * cell) * cell;
            base.px[fi] = px;
            base.py[fi] = py;
            base.diag[fi] = diag;
            base.inv_diag[fi] = diag > R(1e-6f) ? R(1) / diag : R(0);
        }
    }

    // apply_preconditioner, src/apic.rs:1788-1821
    void apply_preconditioner() {
        std::fill(base.e.begin() + W, base.e.begin() + W + cells, R(0));
        const R omega = R(1.2f);
        base.smooth(false, omega);
        if (!levels.empty()) {
            base.restrict_to(levels[0]);
            // the reference's levels, or the relative mode's deeper hierarchy
            size_t last = active_levels() - 1;
            for (size_t l = 0; l < last; ++l) {
                levels[l].smooth(false, omega);
                levels[l].restrict_to(levels[l + 1]);
            }
            if (solver_mode == 1) levels[last].solve_coarse_relative();
            else levels[last].solve_coarse_exact();
            for (size_t l = last; l-- > 0;) {
                levels[l].prolongate_from(levels[l + 1]);
                levels[l].smooth(true, omega);
            }
            base.prolongate_from(levels[0]);
        }
        base.smooth(true, omega);
    }

    // apply_matrix_and_dot, src/apic.rs:1823-1863
    R apply_matrix_and_dot(const std::vector<R>& in, std::vector<R>& out) {
        const size_t w = W;
        auto chunk = [&](size_t b, size_t e) {
            R sum = 0;
            for (size_t k = b; k < e; ++k) {
                size_t i = base.fluid_idx[k];
                R v = base.diag[i] * in[i];
                v = std::fma(base.px[i], in[i + 1], v);
                v = std::fma(base.px[i - 1], in[i - 1], v);
                v = std::fma(base.py[i], in[i + w], v);
                v = std::fma(base.py[i - w], in[i - w], v);
                out[i] = v;
                sum += v * in[i];
            }
            return sum;
        };
        size_t n = fluid_cells;
        if (n >= 8192) {
            const long nch = (long)((n + 8191) / 8192);
            std::vector<R> part((size_t)nch);
            ORC_PAR(nthreads > 1)
            for (long c = 0; c < nch; ++c) part[(size_t)c] = chunk((size_t)c * 8192, std::min(n, (size_t)c * 8192 + 8192));
            R total = 0;
            for (long c = 0; c < nch; ++c) total += part[(size_t)c];
            return total;
        }
        return chunk(0, n);
    }

    // solve_pcg, src/apic.rs:1865-1956
    void solve_pcg() {
        size_t start = W, end = start + cells;
        stats.pcg_iterations = 0;
        stats.pcg_exit_reason = 0;
        stats.pcg_max_residual = 0;
        for (uint32_t k = 0; k < fluid_cells; ++k) pressure[base.fluid_idx[k] - W] = 0;
        apply_preconditioner();
        std::copy(base.e.begin() + start, base.e.begin() + end, search.begin() + start);
        R gamma = dot_product(base.r.data() + start, base.e.data() + start, cells);
        if (gamma < R(kSafety)) { stats.pcg_exit_reason = 1; return; }
        for (int it = 0; it < 100; ++it) {
            stats.pcg_iterations = it + 1;
            R delta = apply_matrix_and_dot(search, As);
            if (std::fabs(delta) < R(kSafety)) { stats.pcg_exit_reason = 2; return; }
            R alpha = gamma / delta;
            R maxe = 0;
            if (nthreads > 1) {  // par_iter().map().reduce(max) (:1910-1928): max is order-free
                ORC_PAR_RED(true, max, maxe)
                for (uint32_t k = 0; k < fluid_cells; ++k) {
                    size_t pi = base.fluid_idx[k], raw = pi - W;
                    pressure[raw] += alpha * search[pi];
                    R nr = std::fma(-alpha, As[pi], base.r[pi]);
                    base.r[pi] = nr;
                    maxe = std::max(maxe, std::fabs(nr));
                }
            } else {
                for (uint32_t k = 0; k < fluid_cells; ++k) {
                    size_t pi = base.fluid_idx[k], raw = pi - W;
                    pressure[raw] += alpha * search[pi];
                    R nr = std::fma(-alpha, As[pi], base.r[pi]);
                    base.r[pi] = nr;
                    maxe = rmax(maxe, std::fabs(nr));
                }
            }
            stats.pcg_max_residual = (float)maxe;
            if (maxe < solve_error) { stats.pcg_exit_reason = 3; return; }
            apply_preconditioner();
            R ng = dot_product(base.r.data() + start, base.e.data() + start, cells);
            if (ng < R(kSafety)) { stats.pcg_exit_reason = 4; return; }
            R beta = ng / gamma;
            gamma = ng;
            ORC_PAR(nthreads > 1 && fluid_cells > 4096)
            for (uint32_t k = 0; k < fluid_cells; ++k) {
                size_t i = base.fluid_idx[k];
                search[i] = base.e[i] + (beta * search[i]);
            }
        }
        stats.pcg_exit_reason = 5;
    }

    // Solver rules of the "relative" mode, an extension the reference cannot
    // express (its solve_pcg, src/apic.rs:1865-1956, hardcodes absolute exits
    // and 100 iterations): preconditioned CG with the same V-cycle and FP32
    // vectors s, r, z; A s summed per row and kept in FP64; the
    // pressure is accumulated in FP64 (written to the FP32 pressure as it
    // goes); dot products, norms and step scalars in A (FP64 when fp64); stop
    // when the residual meets ||r||_2 <= rel_tol * ||b||_2, or after
    // max_iterations; every recompute_every iterations r is replaced by the
    // true residual b - A p evaluated in FP64; the residual is the pair
    // (r, r_lo) of the update below, whose sum follows b - A p to FP64
    // rounding; breakdown when s.As <= 0 or r.z <= 0.  Sums run in ascending
    // cell order.  The FP32 pressure itself cannot have a true residual much
    // below eps32 || |A| |p| || / ||b||.
    std::vector<double> pressure_d;  // the relative mode's FP64 pressure of the last solve
    std::vector<double> as_d;        // the relative mode's A s (FP64; As holds its rounding to R)
    std::vector<R> as_lo;            // and as_d - As in R
    std::vector<R> r_lo;             // the relative mode's residual low part (padded like r)
    int solver_mode = 0;
    double solver_rel_tol = 1e-6;
    int solver_max_it = 500, solver_fp64 = 1, solver_recompute = 25;
    template <class A>
    void solve_pcg_relative_acc() {
        const size_t w = W, start = W, end = start + cells;
        stats.pcg_iterations = 0;
        stats.pcg_exit_reason = 0;
        stats.pcg_max_residual = 0;
        stats.pcg_rel_residual = 0;
        for (uint32_t k = 0; k < fluid_cells; ++k) pressure[base.fluid_idx[k] - W] = 0;
        std::vector<double>& pd = pressure_d;
        pd.assign(pressure.begin(), pressure.end());
        std::vector<R> b(base.r);
        const bool par = nthreads > 1;
        const long ncells = (long)cells;
        auto fluid_sum = [&](auto term) {
            A acc = 0;
            if (par) {
                ORC_PAR_RED(true, +, acc)
                for (long raw = 0; raw < ncells; ++raw)
                    if (base.fluid[(size_t)raw]) acc += term((size_t)raw + w);
            } else {
                for (size_t raw = 0; raw < cells; ++raw)
                    if (base.fluid[raw]) acc += term(raw + w);
            }
            return acc;
        };
        // r.z over the dense padded interior, and (A s).z beside it
        auto dense_dots = [&](A& ng_out, A* asz) {
            A ng = 0, az = 0;
            if (par) {
                ORC_PAR_RED(true, +, ng, az)
                for (long i = (long)start; i < (long)end; ++i) {
                    ng += A(base.r[(size_t)i]) * A(base.e[(size_t)i]);
                    if (asz) az += A(As[(size_t)i]) * A(base.e[(size_t)i]);
                }
            } else {
                for (size_t i = start; i < end; ++i) {
                    ng += A(base.r[i]) * A(base.e[i]);
                    if (asz) az += A(As[i]) * A(base.e[i]);
                }
            }
            ng_out = ng;
            if (asz) *asz = az;
        };
        const A bnorm2 = fluid_sum([&](size_t i) { return A(base.r[i]) * A(base.r[i]); });
        if (!(bnorm2 > 0)) { stats.pcg_exit_reason = 1; return; }
        apply_preconditioner();
        std::copy(base.e.begin() + start, base.e.begin() + end, search.begin() + start);
        A gamma = 0;
        dense_dots(gamma, nullptr);
        if (!(gamma > 0)) { stats.pcg_exit_reason = 1; return; }
        const A tol2 = A(solver_rel_tol) * A(solver_rel_tol);
        auto pd_at = [&](int64_t raw) { return raw >= 0 && raw < (int64_t)cells ? pd[(size_t)raw] : 0.0; };
        // r = b - A p of the FP64 pressure (FP64, the fused chain of
        // apply_matrix_and_dot); returns ||r||^2
        auto true_residual = [&]() {
            double acc = 0;
            ORC_PAR_RED(par, +, acc)
            for (size_t raw = 0; raw < cells; ++raw) {
                if (!base.fluid[raw]) continue;
                size_t pi = raw + w;
                const int64_t c = (int64_t)raw, ww = (int64_t)w;
                double v = double(base.diag[pi]) * pd[raw];
                v = std::fma(double(base.px[pi]), pd_at(c + 1), v);
                v = std::fma(double(base.px[pi - 1]), pd_at(c - 1), v);
                v = std::fma(double(base.py[pi]), pd_at(c + ww), v);
                v = std::fma(double(base.py[pi - w]), pd_at(c - ww), v);
                const double t = double(b[pi]) - v;
                base.r[pi] = R(t);
                r_lo[pi] = R(t - double(base.r[pi]));
                acc += t * t;
            }
            return acc;
        };
        std::vector<double>& As_d = as_d;
        As_d.assign(As.size(), 0.0);
        as_lo.assign(As.size(), R(0));
        std::fill(As.begin(), As.end(), R(0));
        r_lo.assign(base.r.size(), R(0));
        // the FP64 residual r + r_lo of the last update, per padded cell
        std::vector<double> rv(base.r.size(), 0.0);
        for (int it = 0; it < solver_max_it; ++it) {
            stats.pcg_iterations = it + 1;
            // A s with FP64 row sums (for R = float: the products of two
            // floats are exact in double), kept as its rounding As and the
            // rest as_lo
            ORC_PAR(par)
            for (uint32_t k = 0; k < fluid_cells; ++k) {
                const size_t i = base.fluid_idx[k];
                double v = double(base.diag[i]) * double(search[i]);
                v = std::fma(double(base.px[i]), double(search[i + 1]), v);
                v = std::fma(double(base.px[i - 1]), double(search[i - 1]), v);
                v = std::fma(double(base.py[i]), double(search[i + w]), v);
                v = std::fma(double(base.py[i - w]), double(search[i - w]), v);
                As_d[i] = v;
                As[i] = R(v);
                as_lo[i] = R(v - double(As[i]));
            }
            const A delta = fluid_sum([&](size_t i) { return A(search[i]) * A(As_d[i]); });
            if (!(delta > 0)) { stats.pcg_exit_reason = 2; return; }
            const A alpha_a = gamma / delta;
            ORC_PAR(par)
            for (size_t raw = 0; raw < cells; ++raw) {
                if (!base.fluid[raw]) continue;
                size_t pi = raw + w;
                pd[raw] += double(alpha_a) * double(search[pi]);
                pressure[raw] = R(pd[raw]);
                // r keeps the rounding of r - alpha As (the preconditioner's
                // input); r_lo gathers what that drops, alpha as_lo included
                const double t = std::fma(-double(alpha_a), double(As[pi]), double(base.r[pi]));
                const double v = std::fma(-double(alpha_a), double(as_lo[pi]), t + double(r_lo[pi]));
                base.r[pi] = R(t);
                r_lo[pi] = R(v - double(base.r[pi]));
                rv[pi] = v;
            }
            A rr = fluid_sum([&](size_t i) { return A(rv[i]) * A(rv[i]); });
            const bool due = solver_recompute > 0 && (it + 1) % solver_recompute == 0;
            if (due) rr = A(true_residual());  // true residual, FP64
            stats.pcg_rel_residual = (double)std::sqrt((double)rr / (double)bnorm2);
            if (rr <= tol2 * bnorm2) { stats.pcg_exit_reason = 3; return; }
            apply_preconditioner();
            // flexible (Polak-Ribiere) beta = r_{k+1}.(z_{k+1} - z_k) / r_k.z_k, with
            // r_k.z_{k+1} = r_{k+1}.z_{k+1} + alpha (A s).z_{k+1}: robust when the
            // V-cycle (a coarse CG with absolute exits) is not a fixed linear operator
            A ng = 0, asz = 0;
            dense_dots(ng, &asz);  // (A s).z with As
            if (!(ng > 0)) { stats.pcg_exit_reason = 4; return; }
            const R beta = R(-alpha_a * asz / gamma);
            gamma = ng;
            ORC_PAR(par)
            for (size_t raw = 0; raw < cells; ++raw) {
                if (!base.fluid[raw]) continue;
                size_t i = raw + w;
                search[i] = base.e[i] + (beta * search[i]);
            }
        }
        stats.pcg_exit_reason = 5;
    }
    void solve_pressure() {
        if (solver_mode == 0) solve_pcg();
        else if (solver_fp64) solve_pcg_relative_acc<double>();
        else solve_pcg_relative_acc<R>();
    }

    // apply_pressure_gradient, src/apic.rs:1958-2055
    void apply_pressure_gradient() {
        const R target = 9, inv_cell = R(1) / cell;
        ORC_PAR(nthreads > 1 && H > 32)
        for (int32_t y = 0; y < (int32_t)H; ++y)
            for (int32_t x = 0; x <= (int32_t)W; ++x) {
                size_t f = u_index(x, y);
                int32_t lx = x - 1;
                bool lo = is_obstacle(lx, y) == 1, ro = is_obstacle(x, y) == 1;
                if (lo != ro) { mac_u[f] = 0; valid_u[f] = 1; continue; }
                bool lf = is_fluid(lx, y) == 1, rf = is_fluid(x, y) == 1;
                R pd = 0;
                size_t li = (size_t)(y * (int32_t)W + lx), ri = (size_t)(y * (int32_t)W + x);
                if (lf && rf) pd = pressure[ri] - pressure[li];
                else if (lf) pd = -pressure[li] / theta(smoothed[ri], smoothed[li], target);
                else if (rf) pd = pressure[ri] / theta(smoothed[li], smoothed[ri], target);
                mac_u[f] -= pd * inv_cell;
            }
        ORC_PAR(nthreads > 1 && H > 32)
        for (int32_t y = 0; y < (int32_t)H + 1; ++y)
            for (int32_t x = 0; x < (int32_t)W; ++x) {
                size_t f = v_index(x, y);
                int32_t by = y - 1;
                bool bo = is_obstacle(x, by) == 1, to = is_obstacle(x, y) == 1;
                if (bo != to) { mac_v[f] = 0; valid_v[f] = 1; continue; }
                bool bf = is_fluid(x, by) == 1, tf = is_fluid(x, y) == 1;
                R pd = 0;
                size_t bi = (size_t)(by * (int32_t)W + x), ti = (size_t)(y * (int32_t)W + x);
                if (bf && tf) pd = pressure[ti] - pressure[bi];
                else if (bf) pd = -pressure[bi] / theta(smoothed[ti], smoothed[bi], target);
                else if (tf) pd = pressure[ti] / theta(smoothed[bi], smoothed[ti], target);
                mac_v[f] -= pd * inv_cell;
            }
    }

    void apply_post_forces(R dt) {  // src/apic.rs:2453-2458
        for (size_t i = 0; i < nu(); ++i) mac_u[i] += post_u[i] * dt;
        for (size_t i = 0; i < nv(); ++i) mac_v[i] += post_v[i] * dt;
    }

    // extrapolate_velocity, src/apic.rs:2266-2350
    void extrapolate_velocity() {
        const size_t w = W, h = H;
        for (int pass = 0; pass < 4; ++pass) {
            mac_u.swap(old_u);
            valid_u.swap(next_valid_u);
            ORC_PAR(nthreads > 1 && h > 32)
            for (size_t y = 0; y < h; ++y)
                for (size_t x = 0; x <= w; ++x) {
                    size_t i = y * (w + 1) + x;
                    if (next_valid_u[i]) { mac_u[i] = old_u[i]; valid_u[i] = 1; continue; }
                    R sum = 0; uint32_t cnt = 0;
                    if (x > 0 && next_valid_u[i - 1]) { sum += old_u[i - 1]; ++cnt; }
                    if (x < w && next_valid_u[i + 1]) { sum += old_u[i + 1]; ++cnt; }
                    if (y > 0 && next_valid_u[i - (w + 1)]) { sum += old_u[i - (w + 1)]; ++cnt; }
                    if (y < h - 1 && next_valid_u[i + w + 1]) { sum += old_u[i + w + 1]; ++cnt; }
                    if (cnt > 0) { mac_u[i] = sum / (R)cnt; valid_u[i] = 1; }
                    else { mac_u[i] = old_u[i]; valid_u[i] = 0; }
                }
        }
        for (int pass = 0; pass < 4; ++pass) {
            mac_v.swap(old_v);
            valid_v.swap(next_valid_v);
            ORC_PAR(nthreads > 1 && h > 32)
            for (size_t y = 0; y <= h; ++y)
                for (size_t x = 0; x < w; ++x) {
                    size_t i = y * w + x;
                    if (next_valid_v[i]) { mac_v[i] = old_v[i]; valid_v[i] = 1; continue; }
                    R sum = 0; uint32_t cnt = 0;
                    if (x > 0 && next_valid_v[i - 1]) { sum += old_v[i - 1]; ++cnt; }
                    if (x < w - 1 && next_valid_v[i + 1]) { sum += old_v[i + 1]; ++cnt; }
                    if (y > 0 && next_valid_v[i - w]) { sum += old_v[i - w]; ++cnt; }
                    if (y < h && next_valid_v[i + w]) { sum += old_v[i + w]; ++cnt; }
                    if (cnt > 0) { mac_v[i] = sum / (R)cnt; valid_v[i] = 1; }
                    else { mac_v[i] = old_v[i]; valid_v[i] = 0; }
                }
        }
    }

    // transfer_grid_to_particles_and_advect, src/apic.rs:2057-2264
    void transfer_grid_to_particles_and_advect(R dt) {
        const R min_x = R(1.001f) * cell, max_x = ((R)(float)W - R(1.001f)) * cell;
        const R min_y = R(1.001f) * cell, max_y = ((R)(float)H - R(1.001f)) * cell;
        const R h2 = cell * cell, one = R(1), eps = R(1e-6f), half = R(0.5), zero = R(0);
        const R bd = -border_damping;
        ORC_PAR(nthreads > 1 && nalloc > 8 * 256)  // par_iter over chunks (:2079-2085)
        for (size_t i = 0; i < nalloc; ++i) {
            R x = p[PX_][i], y = p[PY_][i];
            R gx = x / cell, gy = y / cell;
            // u grid
            R ubx = std::floor(gx), uby = std::floor(gy - half);
            R utx = gx - ubx, uty = (gy - half) - uby;
            int32_t uix = trunc_i(ubx), uiy = trunc_i(uby);
            uint32_t ui[4] = {u_index(uix, uiy), u_index(uix + 1, uiy), u_index(uix, uiy + 1), u_index(uix + 1, uiy + 1)};
            R uitx = one - utx, uity = one - uty;
            R uw[4] = {uitx * uity, utx * uity, uitx * uty, utx * uty};
            R udx[4] = {(-utx) * cell, uitx * cell, (-utx) * cell, uitx * cell};
            R udy[4] = {(-uty) * cell, (-uty) * cell, uity * cell, uity * cell};
            // v grid
            R vgx = gx - half;
            R vbx = std::floor(vgx), vby = std::floor(gy);
            R vtx = vgx - vbx, vty = gy - vby;
            R vitx = one - vtx, vity = one - vty;
            R vw[4] = {vitx * vity, vtx * vity, vitx * vty, vtx * vty};
            int32_t vix = trunc_i(vbx), viy = trunc_i(vby);
            uint32_t vi[4] = {v_index(vix, viy), v_index(vix + 1, viy), v_index(vix, viy + 1), v_index(vix + 1, viy + 1)};
            R vdx[4] = {(-vtx) * cell, vitx * cell, (-vtx) * cell, vitx * cell};
            R vdy[4] = {(-vty) * cell, (-vty) * cell, vity * cell, vity * cell};
            R un[4], vn[4];
            for (int k = 0; k < 4; ++k) { un[k] = mac_u[ui[k]]; vn[k] = mac_v[vi[k]]; }
            auto interp = [&](const R* wts, const uint8_t* open, const uint32_t* ids, const R* nodes) {
                R mw[4];
                for (int k = 0; k < 4; ++k) mw[k] = open[ids[k]] ? wts[k] : R(0);
                R ws = ((mw[0] + mw[1]) + mw[2]) + mw[3];
                R inv = one / fast_max(ws, R(1e-6f));
                R vel = 0;
                for (int k = 0; k < 4; ++k) vel += (nodes[k] * mw[k]) * inv;
                return vel;
            };
            R nvx = interp(uw, open_u.data(), ui, un);
            R nvy = interp(vw, open_v.data(), vi, vn);
            R res = p[TR_][i];
            R speed = std::sqrt(nvx * nvx + nvy * nvy);
            R cfl_local = (speed * dt) / cell;
            R scale = fast_min(fast_max(cfl_local, zero), R(1));
            R gam = (scale * scale) * (R(3) - (R(2) * scale));
            R jitter = (gam * (R)random_lane((uint32_t)i, timestamp)) * dt;
            R dt_raw = (dt + res) + jitter;
            R dt_act = fast_min(fast_max(dt_raw, zero), dt * R(2));
            p[TR_][i] = (dt + res) - dt_act;
            R nx = x + dt_act * nvx, ny = y + dt_act * nvy;
            bool oxmin = nx < min_x, oxmax = nx > max_x, oymin = ny < min_y, oymax = ny > max_y;
            bool ovxmin = oxmin && nvx < zero, ovxmax = oxmax && nvx > zero;
            bool ovymin = oymin && nvy < zero, ovymax = oymax && nvy > zero;
            if (oxmin) nx = min_x;
            if (oxmax) nx = max_x;
            if (ovxmin) nvx = nvx * bd;
            if (ovxmax) nvx = nvx * bd;
            if (oymin) ny = min_y;
            if (oymax) ny = max_y;
            if (ovymin) nvy = nvy * bd;
            if (ovymax) nvy = nvy * bd;
            R udx_ = (h2 * utx) * (one - utx), udy_ = (h2 * uty) * (one - uty);
            R vdx_ = (h2 * vtx) * (one - vtx), vdy_ = (h2 * vty) * (one - vty);
            R uix_ = udx_ > eps ? one / udx_ : R(0), uiy_ = udy_ > eps ? one / udy_ : R(0);
            R vix_ = vdx_ > eps ? one / vdx_ : R(0), viy_ = vdy_ > eps ? one / vdy_ : R(0);
            R cux = 0, cuy = 0, cvx = 0, cvy = 0;
            for (int k = 0; k < 4; ++k) {
                cux += (uw[k] * un[k]) * udx[k];
                cuy += (uw[k] * un[k]) * udy[k];
            }
            for (int k = 0; k < 4; ++k) {
                cvx += (vw[k] * vn[k]) * vdx[k];
                cvy += (vw[k] * vn[k]) * vdy[k];
            }
            R c_ux = cux * uix_, c_uy = cuy * uiy_, c_vx = cvx * vix_, c_vy = cvy * viy_;
            bool rx = ovxmin || ovxmax, ry = ovymin || ovymax;
            if (rx) { c_uy = c_uy * R(-1); c_vx = c_vx * R(-1); }
            if (ry) { c_uy = c_uy * R(-1); c_vx = c_vx * R(-1); }
            p[PX_][i] = nx; p[PY_][i] = ny; p[VX_][i] = nvx; p[VY_][i] = nvy;
            p[CUX_][i] = c_ux; p[CUY_][i] = c_uy; p[CVX_][i] = c_vx; p[CVY_][i] = c_vy;
        }
    }

    // get_max_particle_velocity, src/apic.rs:2412-2427 (all lanes, including tail lanes)
    R max_particle_velocity() const {
        R m = 0;
        for (size_t i = 0; i < nalloc; ++i) {
            R vx = p[VX_][i], vy = p[VY_][i];
            m = fast_max(m, std::sqrt(vx * vx + vy * vy));
        }
        return m;
    }

    // add_radial_force, src/apic.rs:2352-2410
    void add_radial_force(R wx, R wy, R radius, R strength) {
        int32_t mnx = std::max((int32_t)std::floor((wx - radius) / cell), 0);
        int32_t mxx = std::min((int32_t)std::ceil((wx + radius) / cell), (int32_t)W);
        int32_t mny = std::max((int32_t)std::floor((wy - radius) / cell), 0);
        int32_t mxy = std::min((int32_t)std::ceil((wy + radius) / cell), (int32_t)H);
        R r2 = radius * radius;
        bool repel = strength > R(0);
        // The reference casts the bounds to u32; a negative upper bound wraps
        // to a huge value, and its inner range checks then stop at W/H, so
        // the upper bounds are capped at W and H here.
        int64_t xe = std::min<int64_t>((uint32_t)mxx, W), ye = std::min<int64_t>((uint32_t)mxy, H);
        for (int64_t y = (uint32_t)mny; y <= ye; ++y)
            for (int64_t x = (uint32_t)mnx; x <= xe; ++x) {
                if (x <= (int64_t)W && y < (int64_t)H) {
                    R dx = (R)(float)x * cell - wx, dy = ((R)(float)y + R(0.5)) * cell - wy;
                    R d2 = dx * dx + dy * dy;
                    if (d2 < r2 && d2 > R(1e-5f)) {
                        R d = std::sqrt(d2);
                        R fall = R(1) - (d / radius);
                        R f = ((dx / d) * strength) * fall;
                        size_t id = u_index((int32_t)x, (int32_t)y);
                        (repel ? post_u : ext_u)[id] += f;
                    }
                }
                if (x < (int64_t)W && y <= (int64_t)H) {
                    R dx = ((R)(float)x + R(0.5)) * cell - wx, dy = (R)(float)y * cell - wy;
                    R d2 = dx * dx + dy * dy;
                    if (d2 < r2 && d2 > R(1e-5f)) {
                        R d = std::sqrt(d2);
                        R fall = R(1) - (d / radius);
                        R f = ((dy / d) * strength) * fall;
                        size_t id = v_index((int32_t)x, (int32_t)y);
                        (repel ? post_v : ext_v)[id] += f;
                    }
                }
            }
    }

    // Apic::update, src/apic.rs:2429-2474
    void update(bool disable_sort, float frame_dt_f) {
        R frame_dt = frame_dt_f;
        R t = 0;
        uint32_t substeps = 0, total_it = 0;
        R last_maxv = 0;
        while (t < frame_dt) {
            R maxv = max_particle_velocity();
            last_maxv = maxv;
            R safe = maxv > R(1e-5f) ? (cfl * cell) / maxv : frame_dt;
            R dt = rmin(safe, frame_dt - t);
            if (!disable_sort && t == R(0)) update_spatial_lookup();
            transfer_particles_to_grid();
            apply_external_forces(dt);
            build_pressure_system();
            update_multilevel_grid();
            solve_pressure();
            apply_pressure_gradient();
            apply_post_forces(dt);
            extrapolate_velocity();
            transfer_grid_to_particles_and_advect(dt);
            t += dt;
            prev_dt = dt;
            ++timestamp;
            ++substeps;
            total_it += stats.pcg_iterations;
        }
        std::fill(ext_u.begin(), ext_u.end(), R(0)); std::fill(ext_v.begin(), ext_v.end(), R(0));
        std::fill(post_u.begin(), post_u.end(), R(0)); std::fill(post_v.begin(), post_v.end(), R(0));
        stats.substeps = substeps;
        stats.pcg_iterations_total = total_it;
        stats.fluid_cells = fluid_cells;
        stats.timestamp = timestamp;
        stats.prev_deltatime = (float)prev_dt;
        stats.max_particle_velocity = (float)last_maxv;
        if (solver_mode == 0) stats.pcg_rel_residual = -1;
    }

    // ---- field access ----------------------------------------------------
    Level<R>* level_at(uint32_t l) { return l == 0 ? &base : (l <= levels.size() ? &levels[l - 1] : nullptr); }
    std::vector<R>* real_field(int f, uint32_t l) {
        if (f >= APIC_POS_X && f <= APIC_TIME_RESIDUAL) return &p[f - APIC_POS_X];
        Level<R>* L = level_at(l);
        switch (f) {
            case APIC_MAC_U: return &mac_u;
            case APIC_MAC_V: return &mac_v;
            case APIC_DENSITY: return &density;
            case APIC_SMOOTHED_DENSITY: return &smoothed;
            case APIC_PRESSURE: return &pressure;
            case APIC_DIAG: return L ? &L->diag : nullptr;
            case APIC_INV_DIAG: return L ? &L->inv_diag : nullptr;
            case APIC_PLUS_X: return L ? &L->px : nullptr;
            case APIC_PLUS_Y: return L ? &L->py : nullptr;
            case APIC_RESIDUAL: return L ? &L->r : nullptr;
            case APIC_ERROR: return L ? &L->e : nullptr;
            case APIC_SEARCH: return &search;
            case APIC_MATRIX_TIMES_SEARCH: return &As;
            case APIC_EXT_FORCE_U: return &ext_u;
            case APIC_EXT_FORCE_V: return &ext_v;
            case APIC_POST_FORCE_U: return &post_u;
            case APIC_POST_FORCE_V: return &post_v;
            default: return nullptr;
        }
    }
    std::vector<uint8_t>* flag_field(int f, uint32_t l) {
        Level<R>* L = level_at(l);
        switch (f) {
            case APIC_VALID_U: return &valid_u;
            case APIC_VALID_V: return &valid_v;
            case APIC_FLUID: return L ? &L->fluid : nullptr;
            case APIC_OBSTACLE: return &obstacle;
            case APIC_OPEN_U: return &open_u;
            case APIC_OPEN_V: return &open_v;
            default: return nullptr;
        }
    }
    size_t field_length(int f, uint32_t l) {
        if (f == APIC_SORT_KEY || f == APIC_SORT_VAL) return num_particles;
        if (f == APIC_DENSITY || f == APIC_SMOOTHED_DENSITY || f == APIC_PRESSURE) return cells;
        if (f == APIC_PRESSURE_F64) return pressure_d.size();
        if (auto* v = real_field(f, l)) return v->size();
        if (auto* v = flag_field(f, l)) return v->size();
        return (size_t)-1;
    }
    // Fields whose lists depend on flags: rebuild the fluid/red/black lists
    // from the flags after a write so that stage entry points see them.
    void rebuild_lists_from_flags() {
        base.fluid_idx.clear(); base.red.clear(); base.black.clear();
        fluid_cells = 0;
        for (size_t raw = 0; raw < cells; ++raw) {
            if (!base.fluid[raw]) continue;
            uint32_t pidx = (uint32_t)(raw + W);
            base.fluid_idx.push_back(pidx);
            ++fluid_cells;
            size_t x = raw % W, y = raw / W;
            ((x + y) % 2 == 0 ? base.red : base.black).push_back(pidx);
        }
        if (fluid_cells % 8) {
            uint32_t last = base.fluid_idx[fluid_cells - 1];
            for (uint32_t k = 0; k < 8 - fluid_cells % 8; ++k) base.fluid_idx.push_back(last);
        }
        for (auto& L : levels) {
            L.fluid_idx.clear(); L.red.clear(); L.black.clear();
            for (uint32_t raw = 0; raw < L.cells; ++raw) {
                if (!L.fluid[raw]) continue;
                uint32_t pidx = raw + L.w;
                L.fluid_idx.push_back(pidx);
                ((raw % L.w + raw / L.w) % 2 == 0 ? L.red : L.black).push_back(pidx);
            }
        }
    }
};

struct Handle {
    int precision;
    void* sim;
};

template <class R> Sim<R>* S(Handle* h) { return static_cast<Sim<R>*>(h->sim); }

template <class R> int read_field_t(Sim<R>* s, int f, uint32_t l, void* dst) {
    if (f == APIC_SORT_KEY || f == APIC_SORT_VAL) {
        auto& v = f == APIC_SORT_KEY ? s->skey : s->sval;
        std::memcpy(dst, v.data(), v.size() * 4);
        return 0;
    }
    if (f == APIC_PRESSURE_F64) {
        std::memcpy(dst, s->pressure_d.data(), s->pressure_d.size() * sizeof(double));
        return 0;
    }
    if (auto* v = s->real_field(f, l)) {
        size_t n = s->field_length(f, l);
        double* d = static_cast<double*>(dst);
        for (size_t i = 0; i < n; ++i) d[i] = (double)(*v)[i];
        return 0;
    }
    if (auto* v = s->flag_field(f, l)) {
        std::memcpy(dst, v->data(), v->size());
        return 0;
    }
    return APIC_ERR_INVALID_ARGUMENT;
}

template <class R> int write_field_t(Sim<R>* s, int f, uint32_t l, const void* src) {
    if (f == APIC_SORT_KEY || f == APIC_SORT_VAL) {
        auto& v = f == APIC_SORT_KEY ? s->skey : s->sval;
        std::memcpy(v.data(), src, v.size() * 4);
        return 0;
    }
    if (auto* v = s->real_field(f, l)) {
        size_t n = s->field_length(f, l);
        const double* d = static_cast<const double*>(src);
        for (size_t i = 0; i < n; ++i) (*v)[i] = (R)d[i];
        return 0;
    }
    if (auto* v = s->flag_field(f, l)) {
        std::memcpy(v->data(), src, v->size());
        if (f == APIC_FLUID) s->rebuild_lists_from_flags();
        return 0;
    }
    return APIC_ERR_INVALID_ARGUMENT;
}

template <class R> int stage_t(Sim<R>* s, int stage, float dt, uint32_t ts) {
    switch (stage) {
        case 0: s->update_spatial_lookup(); break;
        case 1: s->transfer_particles_to_grid(); (void)ts; break;
        case 2: s->apply_external_forces((R)dt); break;
        case 3: s->build_pressure_system(); break;
        case 4: s->update_multilevel_grid(); break;
        case 5: s->apply_preconditioner(); break;
        case 6: s->solve_pressure(); break;
        case 7: s->apply_pressure_gradient(); break;
        case 8: s->apply_post_forces((R)dt); break;
        case 9: s->extrapolate_velocity(); break;
        case 10: s->transfer_grid_to_particles_and_advect((R)dt); break;
        case 11: s->compute_keys(); break;
        // single multigrid operations between the base grid and level 1, for known-answer tests
        case 12: if (s->levels.empty()) return APIC_ERR_INVALID_ARGUMENT; s->base.prolongate_from(s->levels[0]); break;
        case 13: if (s->levels.empty()) return APIC_ERR_INVALID_ARGUMENT; s->base.restrict_to(s->levels[0]); break;
        case 14: s->base.smooth(false, 1.2f); break;
        case 15: s->base.smooth(true, 1.2f); break;
        // the coarsest level's CG alone (solve_coarse_exact), for the dense-solve pins
        case 16: if (s->levels.empty()) return APIC_ERR_INVALID_ARGUMENT; s->levels[s->ref_levels - 1].solve_coarse_exact(); break;
        // the relative mode's coarse CG
        case 17: if (s->levels.empty()) return APIC_ERR_INVALID_ARGUMENT; s->levels[s->ref_levels - 1].solve_coarse_relative(); break;
        // restrict / prolongate between level `timestamp` and the next, for the
        // pins of the odd-sided levels
        case 18: case 19: {
            auto* f = s->level_at(ts);
            auto* c = s->level_at(ts + 1);
            if (!f || !c) return APIC_ERR_INVALID_ARGUMENT;
            if (stage == 18) f->restrict_to(*c); else f->prolongate_from(*c);
            break;
        }
        default: return APIC_ERR_INVALID_ARGUMENT;
    }
    return 0;
}

}  // namespace

#define DISPATCH(h, expr_f, expr_d) ((h)->precision == 64 ? (expr_d) : (expr_f))

extern "C" {

void* orc_create(int precision, uint32_t w, uint32_t hgt, float cell, const apic_world_properties* wp,
                 const float* xy, uint32_t n, int* status) {
    Handle* h = new Handle{precision == 64 ? 64 : 32, nullptr};
    int st;
    if (h->precision == 64) { auto* s = new Sim<double>(); st = s->create(w, hgt, cell, *wp, xy, n); h->sim = s; }
    else { auto* s = new Sim<float>(); st = s->create(w, hgt, cell, *wp, xy, n); h->sim = s; }
    if (status) *status = st;
    return h;
}

void orc_destroy(void* hv) {
    Handle* h = static_cast<Handle*>(hv);
    if (!h) return;
    if (h->precision == 64) delete S<double>(h); else delete S<float>(h);
    delete h;
}

int orc_update(void* hv, int disable_sort, float frame_dt) {
    Handle* h = static_cast<Handle*>(hv);
    DISPATCH(h, S<float>(h)->update(disable_sort != 0, frame_dt), S<double>(h)->update(disable_sort != 0, frame_dt));
    return 0;
}

int orc_stage(void* hv, int stage, float dt, uint32_t ts) {
    Handle* h = static_cast<Handle*>(hv);
    return DISPATCH(h, stage_t(S<float>(h), stage, dt, ts), stage_t(S<double>(h), stage, dt, ts));
}

long long orc_field_length(void* hv, int f, uint32_t l) {
    Handle* h = static_cast<Handle*>(hv);
    size_t n = DISPATCH(h, S<float>(h)->field_length(f, l), S<double>(h)->field_length(f, l));
    return n == (size_t)-1 ? -1 : (long long)n;
}

int orc_read_field(void* hv, int f, uint32_t l, void* dst) {
    Handle* h = static_cast<Handle*>(hv);
    return DISPATCH(h, read_field_t(S<float>(h), f, l, dst), read_field_t(S<double>(h), f, l, dst));
}

int orc_write_field(void* hv, int f, uint32_t l, const void* src) {
    Handle* h = static_cast<Handle*>(hv);
    return DISPATCH(h, write_field_t(S<float>(h), f, l, src), write_field_t(S<double>(h), f, l, src));
}

void orc_get_dims(void* hv, uint32_t* w, uint32_t* hh, uint32_t* np, uint32_t* nc, uint32_t* nl, uint32_t* fc) {
    Handle* h = static_cast<Handle*>(hv);
    if (h->precision == 64) { auto* s = S<double>(h); *w = s->W; *hh = s->H; *np = s->num_particles; *nc = s->num_chunks; *nl = (uint32_t)s->ref_levels; *fc = s->fluid_cells; }
    else { auto* s = S<float>(h); *w = s->W; *hh = s->H; *np = s->num_particles; *nc = s->num_chunks; *nl = (uint32_t)s->ref_levels; *fc = s->fluid_cells; }
}

// levels of the reference's halving and with the relative mode's deeper ones
void orc_get_level_counts(void* hv, uint32_t* ref, uint32_t* all) {
    Handle* h = static_cast<Handle*>(hv);
    if (h->precision == 64) { *ref = (uint32_t)S<double>(h)->ref_levels; *all = (uint32_t)S<double>(h)->levels.size(); }
    else { *ref = (uint32_t)S<float>(h)->ref_levels; *all = (uint32_t)S<float>(h)->levels.size(); }
}

void orc_get_timestamp(void* hv, uint32_t* ts, double* prev_dt) {
    Handle* h = static_cast<Handle*>(hv);
    if (h->precision == 64) { *ts = S<double>(h)->timestamp; *prev_dt = S<double>(h)->prev_dt; }
    else { *ts = S<float>(h)->timestamp; *prev_dt = S<float>(h)->prev_dt; }
}

void orc_set_timestamp(void* hv, uint32_t ts, double prev_dt) {
    Handle* h = static_cast<Handle*>(hv);
    if (h->precision == 64) { S<double>(h)->timestamp = ts; S<double>(h)->prev_dt = prev_dt; }
    else { S<float>(h)->timestamp = ts; S<float>(h)->prev_dt = (float)prev_dt; }
}

void orc_set_solver(void* hv, int mode, double rel_tol, int max_it, int fp64, int recompute) {
    auto* h = static_cast<Handle*>(hv);
    auto set = [&](auto* sim) {
        sim->solver_mode = mode; sim->solver_rel_tol = rel_tol; sim->solver_max_it = max_it;
        sim->solver_fp64 = fp64; sim->solver_recompute = recompute;
    };
    if (h->precision == 64) set(S<double>(h)); else set(S<float>(h));
}

void orc_set_gravity_x(void* hv, double gx) {
    auto* h = static_cast<Handle*>(hv);
    if (h->precision == 64) S<double>(h)->gravity_x = gx; else S<float>(h)->gravity_x = (float)gx;
}

void orc_get_stats(void* hv, apic_step_stats* out) {
    Handle* h = static_cast<Handle*>(hv);
    *out = DISPATCH(h, S<float>(h)->stats, S<double>(h)->stats);
}

// Threads of the threaded mode (the CPU baseline); 1, the default, is the checker.
void orc_set_threads(void* hv, int threads) {
    Handle* h = static_cast<Handle*>(hv);
    if (h->precision == 64) S<double>(h)->set_threads(threads); else S<float>(h)->set_threads(threads);
}

void orc_set_p2g_shuffle(void* hv, uint32_t seed) {
    Handle* h = static_cast<Handle*>(hv);
    if (h->precision == 64) S<double>(h)->p2g_shuffle_seed = seed; else S<float>(h)->p2g_shuffle_seed = seed;
}

double orc_max_particle_velocity(void* hv) {
    Handle* h = static_cast<Handle*>(hv);
    return DISPATCH(h, (double)S<float>(h)->max_particle_velocity(), S<double>(h)->max_particle_velocity());
}

void orc_add_radial_force(void* hv, float x, float y, float r, float s) {
    Handle* h = static_cast<Handle*>(hv);
    if (h->precision == 64) S<double>(h)->add_radial_force(x, y, r, s); else S<float>(h)->add_radial_force(x, y, r, s);
}

// Scalar helpers for known-answer tests.
uint32_t orc_morton(uint32_t x, uint32_t y) { return morton(x, y); }
void orc_random_chunk(uint32_t chunk, uint32_t seed, float* out8) {
    for (uint32_t l = 0; l < 8; ++l) out8[l] = random_lane(chunk * 8 + l, seed);
}
float orc_temporal_w(float tau) { return temporal_w<float>(tau); }
float orc_theta(float rho, float rc, float target) { return theta<float>(rho, rc, target); }
float orc_dot(const float* a, const float* b, size_t n) { return dot_product<float>(a, b, n); }

}  // extern "C"
