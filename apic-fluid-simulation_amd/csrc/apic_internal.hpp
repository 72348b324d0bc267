// apic_internal.hpp -- device state of one simulation and the launchers of
// every gfx950 kernel.  Host code only (plus plain-old-data structs shared
// with the kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/apic.h"
#include "launch.hpp"

namespace apic {

// Particle fields in one SoA block: field f of particle i is at f*nalloc + i.
enum ParticleField { P_X = 0, P_Y, V_X, V_Y, CU_X, CU_Y, CV_X, CV_Y, T_RES, NUM_PFIELDS };

// One level of the multigrid hierarchy.  Level 0 is the base pressure grid.
// Solver arrays are padded by one row on each side (index = raw + w), as in
// MultiGridLevel (src/apic.rs:402-453), so that the 5-point stencil of any
// interior row reads zeros past the grid instead of branching.
struct LevelView {
    uint32_t w, h, cells;
    float* diag;
    float* inv_diag;
    float* px;
    float* py;
    float* r;
    float* e;
    uint8_t* fluid;  // raw index
};

// The one-workgroup bottom of the V-cycle keeps its levels in LDS when they
// fit: six float arrays over the padded range per level and the flag bytes,
// plus the coarsest level's three CG vectors and its serial-sum products.
// 144 KB of gfx950's 160 KB per workgroup: room for the levels from 64 x 64
// cells down (the bottom starts at kBottomFirstCells cells or fewer).
constexpr size_t kBottomLdsFloats = 34816;
constexpr size_t kBottomLdsFlags = 8192;
// A level's padded array length (ceil8(cells + 2w), that of every level
// array) and its flag bytes rounded to a multiple of 4.  The LDS image of the
// bottom levels and their arrays in global memory share this layout.
__host__ __device__ inline size_t bottom_padded(const LevelView& L) {
    return ((size_t)L.cells + 2 * (size_t)L.w + 7) / 8 * 8;
}
__host__ __device__ inline size_t bottom_flags(const LevelView& L) { return ((size_t)L.cells + 3) / 4 * 4; }
constexpr uint32_t kBottomFirstCells = 1024;  // cells of the one-workgroup bottom's finest level

// Whether levels [first, count) fit the one-workgroup bottom's LDS.
inline bool bottom_fits_lds(const std::vector<LevelView>& lv, size_t first, size_t count) {
    size_t fl = 0, fb = 0;
    for (size_t l = first; l < count; ++l) {
        fl += 6 * bottom_padded(lv[l]);
        fb += bottom_flags(lv[l]);
    }
    fl += 3 * bottom_padded(lv[count - 1]) + lv[count - 1].cells;
    return fl <= kBottomLdsFloats && fb <= kBottomLdsFlags;
}

// P2G works on Morton-aligned blocks of 16 x 16 cells (transfer.hip); each
// block keeps a halo record of kP2gHaloFloats partial sums for the targets
// around its edge.
constexpr uint32_t kP2gBlockSide = 16;
constexpr size_t kP2gHaloFloats = 512;
inline uint32_t p2g_blocks(uint32_t side) { return (side + kP2gBlockSide - 1) / kP2gBlockSide; }

// Scalars of one pressure solve.  The reductions write per-workgroup
// partials; the kernel that needs a reduced scalar sums the partials itself,
// in the same fixed order in every workgroup, and applies the reference's exit
// test.  Workgroup 0 records the decision here for later kernels and the host.
constexpr int kRedBlocks = 512;          // most workgroups (and partials) of a solver reduction
constexpr int kRedCells = 2048;          // cells per reduction workgroup below that
inline uint32_t reduction_workgroups(uint32_t cells) {
    return std::max<uint32_t>(1, std::min<uint32_t>(kRedBlocks, (cells + kRedCells - 1) / kRedCells));
}
constexpr int kMaxPcgIterations = 100;   // solve_pcg's cap (src/apic.rs:1894)
struct PcgState {
    int32_t done;         // nonzero: every kernel of the PCG loop returns at once
    int32_t iters;        // iterations entered
    int32_t exit_reason;  // as apic_step_stats.pcg_exit_reason
    int32_t r_in_alt;     // the solve ended with the base residual in Sim::r_alt (copied home by the up-sweep)
    float maxerr, delta, alpha, pad1;
    float gamma[kMaxPcgIterations + 4];  // gamma[k]: the r.z used by iteration k
    // relative mode (apic_solver_config.mode 1): FP64 scalars
    double bnorm2, rnorm2;   // ||b||^2 and the last ||r||^2
    double tol2;             // rel_tol^2
    double gamma_d[2];       // r.z of iteration k in gamma_d[k & 1]
    double alpha_d;          // the last step length in FP64 (the flexible beta uses it)
    int32_t max_iters;
};

// Base-level fluid statistics needed by the padded-list quirk of prolongate
// (src/apic.rs:1487-1494, 599-615): count and largest raw index.
struct FluidStats {
    uint32_t count;
    uint32_t last_raw;  // largest fluid raw index + 1 (0 when there is none)
};

// Time-step scalars of the current substep, kept on the device so that the
// host can queue a substep before the previous one has finished.
struct StepState {
    float dt;        // this substep's dt: min(cfl * dx / max|v|, time left in the frame)
    float prev_dt;   // the previous substep's dt (the temporal kernel's tau uses it)
    float elapsed;   // time advanced in the current frame
    float maxv;      // max |v| the dt was computed from
};

// The end of a substep (src/apic.rs:2444-2461): prev_dt = dt, elapsed += dt.
__device__ inline void substep_end(StepState* ss) {
    StepState s = *ss;
    s.prev_dt = s.dt;
    s.elapsed = s.elapsed + s.dt;
    *ss = s;
}

struct Sim;

// ---- launchers (kernels live in sort.hip, transfer.hip, pressure.hip, grid.hip)
void launch_fill_u32(Sim& sim, uint32_t* p, uint32_t v, size_t n);
void launch_fill_f32(Sim& sim, float* p, float v, size_t n);

// sort.hip
void launch_keygen(Sim& sim);
void launch_radix_sort(Sim& sim);  // keys/vals[0] -> sorted in keys/vals[0]
void launch_permute(Sim& sim);
size_t radix_scratch_words(size_t n);

// transfer.hip
size_t cell_key_count(const Sim& sim);            // Morton keys of the grid: morton(W-1, H-1) + 1
void launch_cell_ranges(Sim& sim);                // per-cell [begin, end) of keys[0]'s order
void launch_p2g(Sim& sim, bool forces, float dt);  // P2G + normalisation (+ external forces and gravity)
void launch_forces(Sim& sim, float dt);            // apply_external_forces alone
void launch_g2p(Sim& sim, float dt, bool write_keys, bool end_substep = false);  // then launch_substep_end's work
void launch_max_velocity(Sim& sim);

// grid.hip
void launch_gradient(Sim& sim, bool post_force, float dt);
void launch_post_force(Sim& sim, float dt);
void launch_extrapolate(Sim& sim);
void launch_radial_force(Sim& sim, float wx, float wy, float radius, float strength);
void launch_open_faces(Sim& sim);

// pressure.hip
// defer_fstats: the fluid statistics are finished by the launch_mg_build
// that follows at once (its bottom build), not by a launch of their own
void launch_build_pressure(Sim& sim, bool zero_pressure, bool defer_fstats = false);
void launch_mg_build(Sim& sim);
// mode 0: plain, 1: PCG start (gamma), 2: after PCG iteration k (residual test, gamma')
void launch_vcycle(Sim& sim, int mode, int k = 0);
void launch_pcg_reset(Sim& sim);                 // also the relative mode's ||b|| prologue
int pcg_cap(const Sim& sim);                      // iteration cap of the current solver rules
size_t gamma_partials(const Sim& sim);            // partials of r.z the V-cycle writes

// mirror.hip
size_t aosoa_length(const Sim& sim, int group);   // floats of a group in the reference's chunk layout
void launch_pack_aosoa(Sim& sim, int group, float* staging);
void launch_unpack_aosoa(Sim& sim, int group, const float* staging);
void launch_pcg_iteration_head(Sim& sim, int k);  // s update + A s + delta, then p/r update + max|r|
void launch_pcg_decide(Sim& sim, int k);          // gamma test after k iterations; mirrors the state to the host
void launch_substep_begin(Sim& sim, bool first, float frame_dt);  // dt of the substep, on the device
void launch_substep_end(Sim& sim);                                // prev_dt = dt, elapsed += dt
void launch_zero_pressure(Sim& sim);
void launch_fluid_stats(Sim& sim);

struct Sim {
    // dimensions and parameters
    uint32_t W = 0, H = 0, cells = 0;
    float cell = 1.0f;
    apic_world_properties props{};
    float gravity_x = 0.0f;  // horizontal gravity (an extension; the reference has none)
    apic_solver_config solver{};
    uint32_t num_particles = 0, num_chunks = 0;
    size_t nalloc = 0;  // num_chunks * 8
    uint32_t timestamp = 0;
    float prev_dt = 1.0f / 60.0f;

    hipStream_t stream = nullptr;
    bool own_stream = true;   // false: the stream belongs to a batch this simulation is part of
    int device = 0;
    Recorder* rec = nullptr;  // set while a batch records this simulation's launches
    uint32_t nred = 1;        // workgroups (and partials) of the base-level solver reductions

    // particles: two SoA blocks (sort double buffer)
    float* pbuf[2] = {nullptr, nullptr};
    int cur = 0;
    // sort records and radix scratch
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* radix_scratch = nullptr;
    bool keys_valid = false;  // keys[0]/vals[0] hold the current positions' keys
    uint2* cell_range = nullptr;  // [begin, end) of every cell's particles in cell order, by Morton key
    float* p2g_halo = nullptr;  // P2G: per 16 x 16 block, the partial sums of the targets around its edge
    bool cells_valid = false;     // the particles are in cell order and cell_range describes it
    // The sort's gather is deferred: the particles lie in pbuf[cur] in their
    // order before the sort, vals[0] maps a sorted position to its record,
    // P2G reads through it and G2P writes the sorted order into the other
    // buffer (within one update: its first substep's G2P ends it).
    bool perm_pending = false;

    // MAC grid
    float* mac_u = nullptr;   float* mac_v = nullptr;
    float* old_u = nullptr;   float* old_v = nullptr;
    uint8_t* valid_u = nullptr; uint8_t* valid_v = nullptr;
    uint8_t* nvalid_u = nullptr; uint8_t* nvalid_v = nullptr;
    float* density = nullptr;
    float* smooth_a = nullptr;  // smoothed density
    float* pressure = nullptr;
    uint8_t* obstacle = nullptr;
    uint8_t* open_u = nullptr; uint8_t* open_v = nullptr;
    float* ext_u = nullptr; float* ext_v = nullptr;
    float* post_u = nullptr; float* post_v = nullptr;
    bool forces_active = false;

    // solver
    // [0] = base grid, [k] = multigrid_levels[k-1]; after the reference's
    // levels, the relative mode's deeper ones (odd sides halved, rounded up)
    std::vector<LevelView> levels;
    size_t ref_levels = 0;          // levels of the reference's halving, the base included
    // Each solver mode's V-cycle: the levels it runs through (the reference's
    // rules: the reference's levels; the relative mode: all) and where its
    // bottom solver starts.  select_hierarchy() copies the current mode's
    // into level_count, bottom_level and bottom_in_lds.
    struct Hierarchy {
        size_t count = 0, bottom = 0;
        bool in_lds = false;
    } hier[2];
    size_t level_count = 0;         // levels the current mode's V-cycle runs through
    size_t bottom_level = 0;        // first level solved by the one-workgroup kernel
    // the V-cycle's down-sweep errors of the levels above the bottom (the
    // up-sweep reads them and writes the level's e, so that no launch reads
    // a tile ring that another workgroup of the same launch writes)
    std::vector<float*> down_e;
    bool bottom_in_lds = false;     // that kernel holds its levels in LDS
    // The base tiles' launch order (TileOrder, pressure.hip): per base tile a
    // mark that it holds fluid (set by MgBuild), and the order itself, valid
    // from the hierarchy's build until a field is written
    uint8_t* tile_flag = nullptr;
    uint32_t* tile_order = nullptr;
    bool tile_order_valid = false;
    bool fstats_pending = false;    // launch_build_pressure left the fluid statistics to launch_mg_build
    float* search[2] = {nullptr, nullptr};
    float* As = nullptr;
    float* As_lo = nullptr;          // relative mode: A s (FP64) - As
    float* r_lo = nullptr;           // relative mode: the base residual's low part (r + r_lo in FP64)
    float* cg_s = nullptr; float* cg_ap = nullptr; float* cg_r = nullptr;  // coarsest level
    float* rhs = nullptr;            // copy of b for the relative mode's residual recompute
    float* r_alt = nullptr;          // the base residual's second buffer (the fused update writes it)
    PcgState* pcg = nullptr;
    float* part_gamma = nullptr;     // kRedBlocks partials of r.z
    float* part_delta = nullptr;     // kRedBlocks partials of s.As
    float* part_max = nullptr;       // kRedBlocks partials of max |r|
    double* part_gamma_d = nullptr;  // relative mode: FP64 partials of r.z, s.As and |r|^2
    double* part_delta_d = nullptr;
    double* part_rr_d = nullptr;     // partials of ||b||^2 (the relative mode's prologue)
    double* part_rr_t = nullptr;     // ||r||^2 partials of the recurrence residual, one per base tile
    double* part_rr2_d = nullptr;    // ... after the true-residual check
    double* part_asz_d = nullptr;    // relative mode: partials of (A s).z for the flexible beta
    double* pressure_d = nullptr;    // relative mode: the pressure accumulated in FP64
    float* aosoa_staging = nullptr;  // chunk-layout mirrors (2 * nalloc floats, allocated on first use)
    uint32_t* part_fluid = nullptr;  // kRedBlocks (count, last+1) pairs of the base fluid cells
    float* part_maxv = nullptr;      // per-workgroup max |v| of the G2P kernel
    size_t part_maxv_cap = 0;
    FluidStats* fstats = nullptr;
    uint32_t* maxv_bits = nullptr;   // max |v| over all lanes (float bits)
    bool maxv_valid = false;

    // Host-pinned mirrors, written by the kernels themselves (the *_d
    // pointers are their device addresses), read by the host after an event.
    PcgState* h_pcg = nullptr;
    uint32_t* h_maxv = nullptr;
    FluidStats* h_fstats = nullptr;
    StepState* h_step = nullptr;
    PcgState* h_pcg_d = nullptr;
    uint32_t* h_maxv_d = nullptr;
    FluidStats* h_fstats_d = nullptr;
    StepState* h_step_d = nullptr;
    StepState* step = nullptr;          // device StepState
    hipEvent_t e_sync = nullptr;        // host waits (timing disabled)
    // Set while apic_update queues a substep: kernels take dt and prev_dt
    // from `step_src`, and the kernels after the pressure solve return at
    // once unless *tail_guard (the solve's done flag) is set.
    const StepState* step_src = nullptr;
    const int32_t* tail_guard = nullptr;

    apic_step_stats stats{};
    int pcg_batch = 4;

    // stage timers: pairs of HIP events on `stream`, collected once complete
    struct TimerMark {
        int t;
        hipEvent_t a, b;
    };
    int profiling = 0;  // 0 off, 1 the transfer kernels only, 2 every timer
    apic_tuning tuning{11u, 1024u, 0.0f, 0.0f};  // launch shapes (apic_set_tuning, apic_autotune)
    std::vector<hipEvent_t> ev_pool;
    std::vector<TimerMark> ev_pending;
    size_t ev_first = 0;  // ev_pending[ev_first..] are not collected yet
    hipEvent_t ev_open[APIC_TIMER_COUNT] = {};
    double timer_ms[APIC_TIMER_COUNT] = {};
    uint32_t timer_count[APIC_TIMER_COUNT] = {};
    hipEvent_t take_event() {
        if (!ev_pool.empty()) { hipEvent_t e = ev_pool.back(); ev_pool.pop_back(); return e; }
        // timing only: no system-scope fence (no cache writeback around the kernels timed)
        hipEvent_t e = nullptr;
        (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
        return e;
    }
    void mark(int t, int which) {
        if (rec || profiling == 0 || (profiling == 1 && t < APIC_T_K_P2G)) return;
        hipEvent_t e = take_event();
        (void)hipEventRecord(e, stream);
        if (which == 0) {
            ev_open[t] = e;
        } else if (ev_open[t]) {
            ev_pending.push_back(TimerMark{t, ev_open[t], e});
            ev_open[t] = nullptr;
        } else {
            ev_pool.push_back(e);
        }
    }
    // Marks from ev_pending[from..] are dropped (their kernels were queued
    // but will not run).
    void drop_marks(size_t from) {
        for (size_t i = std::max(from, ev_first); i < ev_pending.size(); ++i) {
            ev_pool.push_back(ev_pending[i].a);
            ev_pool.push_back(ev_pending[i].b);
        }
        if (ev_pending.size() > from) ev_pending.resize(std::max(from, ev_first));
    }
    // Adds the completed marks to the timers; with wait, waits for all.
    void collect_timers(bool wait) {
        while (ev_first < ev_pending.size()) {
            TimerMark& m = ev_pending[ev_first];
            if (wait) (void)hipEventSynchronize(m.b);
            else if (hipEventQuery(m.b) != hipSuccess) break;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
                timer_ms[m.t] += ms;
                timer_count[m.t] += 1;
            }
            ev_pool.push_back(m.a);
            ev_pool.push_back(m.b);
            ++ev_first;
        }
        if (ev_first == ev_pending.size()) { ev_pending.clear(); ev_first = 0; }
    }
    void free_events() {
        for (auto& m : ev_pending) { (void)hipEventDestroy(m.a); (void)hipEventDestroy(m.b); }
        for (size_t i = 0; i < ev_first && i < ev_pending.size(); ++i) {}
        ev_pending.clear();
        ev_first = 0;
        for (int t = 0; t < APIC_TIMER_COUNT; ++t)
            if (ev_open[t]) { (void)hipEventDestroy(ev_open[t]); ev_open[t] = nullptr; }
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        ev_pool.clear();
    }

    // device allocations to free
    std::vector<void*> allocations;
    size_t device_bytes = 0;
    uint8_t* slab = nullptr;  // the allocation the construction-time arrays are carved from
    size_t slab_size = 0, slab_used = 0;
    int slab_pass = 0;        // 1: sizing, 2: assigning, 0: separate allocations
    uint8_t* h_block = nullptr;  // the pinned allocation holding the host mirrors

    size_t nu() const { return (size_t)cells + H; }
    size_t nv() const { return (size_t)cells + W; }
    void select_hierarchy() {
        const Hierarchy& hz = hier[solver.mode == 1 ? 1 : 0];
        level_count = hz.count;
        bottom_level = hz.bottom;
        bottom_in_lds = hz.in_lds;
    }
    size_t padded(uint32_t w, uint32_t h) const { return (((size_t)w * h + 2 * (size_t)w) + 7) / 8 * 8; }
    float* pfield(int f, int which = -1) const {
        return pbuf[which < 0 ? cur : which] + (size_t)f * nalloc;
    }
};

// Launches kernel K for `s`: at once on its stream, or into the recorder of
// the batch it is being stepped in.
template <class K>
inline void emit(Sim& s, const typename K::Args& a, uint32_t nwg, int k = 0) {
    if (nwg == 0) return;
    if (s.rec) {
        s.rec->add(reinterpret_cast<const void*>(&batch_kernel<K>), sizeof(a), &a, K::kThreads, nwg, k,
                   !round_robin<K>::value, deal_chunk<K>(a));
        return;
    }
    hipLaunchKernelGGL(one_kernel<K>, dim3(nwg), dim3(K::kThreads), 0, s.stream, a, k);
}

}  // namespace apic
