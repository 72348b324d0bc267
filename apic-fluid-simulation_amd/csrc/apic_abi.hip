// apic_abi.hip -- the extern "C" boundary (include/apic.h) and the substep
// driver of Apic::update (src/apic.rs:2429-2474) on one MI355X.
//
// The host side sequences kernels on one HIP stream and keeps every array of
// the simulation resident in HBM between calls.  It synchronises with the
// device only where the reference's control flow needs a reduced scalar:
// max |v| for the substep dt (once per substep), and the PCG exit flag,
// read every few iterations while the device-side state machine predicates
// the queued iterations.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "apic_internal.hpp"
#include "device_math.hpp"

using namespace apic;

struct apic_handle {
    Sim sim;
    apic_batch* batch = nullptr;  // the batch whose stream the simulation shares
};

// Several simulations on one stream, stepped together (launch.hpp).  The
// launch tables of a substep go to the device in one copy before its first
// kernel; the pinned staging buffer is rewritten only after that copy.
struct apic_batch {
    hipStream_t stream = nullptr;
    uint32_t members = 0;
    hipEvent_t e_sync = nullptr;
    hipEvent_t e_copied = nullptr;
    bool copy_pending = false;
    apic::Recorder rec;
    std::vector<uint8_t> host;
    uint8_t* pinned = nullptr;
    size_t pinned_cap = 0;
    uint8_t* dev = nullptr;
    size_t dev_cap = 0;
    int device = 0;
    hipEvent_t e_arena = nullptr;  // end of this batch's last no-allocation update
    // Set while apic_batch_update_noalloc runs: the launch tables go to this
    // fixed device arena (the library's static scratch) instead of `dev`.
    uint8_t* fixed = nullptr;
    size_t fixed_cap = 0;
    uint8_t* arena() const { return fixed ? fixed : dev; }
};

namespace {

thread_local std::string g_last_error;

apic_status fail(apic_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? APIC_ERR_OUT_OF_MEMORY : APIC_ERR_HIP,    \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

#define CHECK_LAUNCH()                                                                       \
    do {                                                                                     \
        hipError_t e_ = hipGetLastError();                                                   \
        if (e_ != hipSuccess) return fail(APIC_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

// Device arrays of a simulation are carved out of one allocation (a sizing
// pass, then an assigning pass over the same sequence); arrays allocated
// later get their own.
template <class T>
apic_status dalloc(Sim& s, T*& p, size_t count) {
    size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    bytes = (bytes + 255) / 256 * 256;
    if (s.slab_pass == 1) {
        s.slab_size += bytes;
        p = nullptr;
        return APIC_OK;
    }
    if (s.slab_pass == 2) {
        p = reinterpret_cast<T*>(s.slab + s.slab_used);
        s.slab_used += bytes;
        return APIC_OK;
    }
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? APIC_ERR_OUT_OF_MEMORY : APIC_ERR_HIP,
                    std::string("hipMalloc: ") + hipGetErrorString(e));
    s.allocations.push_back(q);
    s.device_bytes += bytes;
    e = hipMemset(q, 0, bytes);
    if (e != hipSuccess) return fail(APIC_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    p = static_cast<T*>(q);
    return APIC_OK;
}

#define ALLOC(ptr, n)                                     \
    do {                                                  \
        apic_status st_ = dalloc(s, ptr, (size_t)(n));    \
        if (st_ != APIC_OK) return st_;                   \
    } while (0)

void release(Sim& s) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    s.free_events();
    if (s.e_sync) { (void)hipEventDestroy(s.e_sync); s.e_sync = nullptr; }
    for (void* p : s.allocations) (void)hipFree(p);
    s.allocations.clear();
    if (s.h_block) (void)hipHostFree(s.h_block);
    s.h_block = nullptr;
    s.h_pcg = nullptr; s.h_maxv = nullptr; s.h_fstats = nullptr; s.h_step = nullptr;
    if (s.stream && s.own_stream) (void)hipStreamDestroy(s.stream);
    s.stream = nullptr;
}

// Apic::new (src/apic.rs:779-916) with optional explicit particles.
apic_status build(Sim& s, uint32_t W, uint32_t H, float cell, const apic_world_properties& wp, bool explicit_set,
                  const float* xy, uint32_t n, hipStream_t shared = nullptr) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(APIC_ERR_NO_DEVICE, "no HIP device found");
    if (W < 3 || H < 3) return fail(APIC_ERR_INVALID_ARGUMENT, "grid sides must be at least 3");
    if (!(cell > 0.0f)) return fail(APIC_ERR_INVALID_ARGUMENT, "cell_size must be positive");
    if ((uint64_t)W * H > (1ull << 31)) return fail(APIC_ERR_INVALID_ARGUMENT, "grid too large");
    HIP_TRY(hipGetDevice(&s.device));
    s.W = W; s.H = H; s.cells = W * H; s.cell = cell; s.props = wp;
    s.nred = reduction_workgroups(s.cells);
    s.solver = apic_solver_config{0, 1e-6f, 500, 1, 25};
    s.timestamp = 0;
    s.prev_dt = 1.0f / 60.0f;

    // scene: walls one cell thick and the 3x3 stratified block (:831-849)
    std::vector<uint8_t> obstacle(s.cells, 0);
    std::vector<float> pos;
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            if (x == 0 || x == W - 1 || y == 0 || y == H - 1) {
                obstacle[(size_t)y * W + x] = 1;
            } else if (!explicit_set && x > 4 && x < W - 4 && y > 4 && y < H / 2) {
                for (int py = 0; py < 3; ++py)
                    for (int px = 0; px < 3; ++px) {
                        pos.push_back(((float)x + ((float)px + 0.5f) / 3.0f) * cell);
                        pos.push_back(((float)y + ((float)py + 0.5f) / 3.0f) * cell);
                    }
            }
        }
    if (explicit_set && n) pos.assign(xy, xy + 2 * (size_t)n);
    s.num_particles = (uint32_t)(pos.size() / 2);
    s.num_chunks = (s.num_particles + 7) / 8;
    s.nalloc = (size_t)s.num_chunks * 8;

    // multigrid levels while both sides are even and > 4 (:900-911); the
    // reference builds none when the scene has no particle (:852-855)
    std::vector<std::pair<uint32_t, uint32_t>> dims{{W, H}};
    if (s.num_particles > 0) {
        uint32_t cw = W, ch = H;
        while (cw % 2 == 0 && ch % 2 == 0 && cw > 4 && ch > 4) {
            cw /= 2; ch /= 2;
            dims.push_back({cw, ch});
        }
        if (dims.size() == 1)
            return fail(APIC_ERR_UNSUPPORTED_SIZE,
                        "no multigrid level: both grid sides must be even and > 4 (the reference panics here)");
    }
    // the relative mode's levels below: while the coarsest has more than 64
    // cells and both sides exceed 2, sides halved and rounded up (a side of
    // 2 x odd stops the reference's halving at up to 511^2 cells)
    const size_t ref_n = dims.size();
    if (ref_n > 1)
        while ((size_t)dims.back().first * dims.back().second > 64 && dims.back().first > 2 &&
               dims.back().second > 2)
            dims.push_back({(dims.back().first + 1) / 2, (dims.back().second + 1) / 2});

    if (shared) {
        s.stream = shared;
        s.own_stream = false;
    } else {
        HIP_TRY(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    }
    const size_t np = std::max<uint32_t>(s.num_particles, 1);
    const size_t na = std::max<size_t>(s.nalloc, 8);
    auto carve = [&]() -> apic_status {
    ALLOC(s.pbuf[0], NUM_PFIELDS * na);
    ALLOC(s.pbuf[1], NUM_PFIELDS * na);
    ALLOC(s.keys[0], np); ALLOC(s.keys[1], np);
    ALLOC(s.vals[0], np); ALLOC(s.vals[1], np);
    ALLOC(s.radix_scratch, radix_scratch_words(np));
    ALLOC(s.cell_range, cell_key_count(s));
    {
        const size_t nblk = (size_t)p2g_blocks(s.W) * p2g_blocks(s.H);
        ALLOC(s.p2g_halo, kP2gHaloFloats * nblk);
    }
    const size_t nu = s.nu(), nv = s.nv(), nc = s.cells;
    ALLOC(s.mac_u, nu); ALLOC(s.mac_v, nv); ALLOC(s.old_u, nu); ALLOC(s.old_v, nv);
    ALLOC(s.valid_u, nu); ALLOC(s.valid_v, nv); ALLOC(s.nvalid_u, nu); ALLOC(s.nvalid_v, nv);
    ALLOC(s.density, nc); ALLOC(s.smooth_a, nc); ALLOC(s.pressure, nc);
    ALLOC(s.obstacle, nc); ALLOC(s.open_u, nu); ALLOC(s.open_v, nv);
    ALLOC(s.ext_u, nu); ALLOC(s.ext_v, nv); ALLOC(s.post_u, nu); ALLOC(s.post_v, nv);
    s.levels.clear();
    for (auto& d : dims) {
        LevelView L{};
        L.w = d.first; L.h = d.second; L.cells = d.first * d.second;
        s.levels.push_back(L);
    }
    if (s.levels.size() > 16) return fail(APIC_ERR_UNSUPPORTED_SIZE, "more than 15 multigrid levels");
    // Each mode's bottom: the one-workgroup bottom solver takes the coarse
    // levels of at most kBottomFirstCells cells that fit in LDS (a finer level
    // runs faster as tiles over several CUs); when even the coarsest does not
    // fit, it works in global memory on every level with <= 4096 cells.
    auto plan = [&](size_t count) {
        Sim::Hierarchy hz;
        hz.count = count;
        hz.bottom = count > 1 ? count - 1 : 0;
        for (size_t l = 1; l < count; ++l)
            if (s.levels[l].cells <= kBottomFirstCells && bottom_fits_lds(s.levels, l, count)) {
                hz.bottom = l;
                hz.in_lds = true;
                break;
            }
        if (!hz.in_lds)
            for (size_t l = 1; l < count; ++l)
                if (s.levels[l].cells <= 4096) { hz.bottom = l; break; }
        return hz;
    };
    s.ref_levels = ref_n;
    s.hier[0] = plan(ref_n);
    s.hier[1] = plan(s.levels.size());
    s.select_hierarchy();
    // The arrays of the LDS bottoms' levels lie in two blocks laid out as the
    // bottom's LDS image (per level: diag, inv_diag, plus_x, plus_y, r, e over
    // the padded range; the flag bytes apart), so that it copies them in and
    // out in one stride; every other level has its own arrays.  One image
    // from the shallower of the two modes' LDS bottoms to the last level
    // holds both (a mode's bottom levels are a contiguous run of it).
    size_t lds_first = s.levels.size();
    for (const auto& hz : s.hier)
        if (hz.in_lds) lds_first = std::min(lds_first, hz.bottom);
    for (size_t l = 0; l < lds_first; ++l) {
        LevelView& L = s.levels[l];
        const size_t pa = s.padded(L.w, L.h);
        ALLOC(L.diag, pa); ALLOC(L.inv_diag, pa); ALLOC(L.px, pa); ALLOC(L.py, pa);
        ALLOC(L.r, pa); ALLOC(L.e, pa);
        ALLOC(L.fluid, bottom_flags(L));
    }
    if (lds_first < s.levels.size()) {
        size_t nf = 0, nb = 0;
        for (size_t l = lds_first; l < s.levels.size(); ++l) {
            nf += 6 * bottom_padded(s.levels[l]);
            nb += bottom_flags(s.levels[l]);
        }
        float* img = nullptr;
        uint8_t* flags = nullptr;
        ALLOC(img, nf);
        ALLOC(flags, nb);
        for (size_t l = lds_first; l < s.levels.size(); ++l) {
            LevelView& L = s.levels[l];
            const size_t pa = bottom_padded(L);
            if (img) {
                L.diag = img; L.inv_diag = img + pa; L.px = img + 2 * pa; L.py = img + 3 * pa;
                L.r = img + 4 * pa; L.e = img + 5 * pa;
                L.fluid = flags;
                img += 6 * pa;
                flags += bottom_flags(L);
            }
        }
    }
    const size_t ndown = std::max(s.hier[0].bottom, s.hier[1].bottom);
    s.down_e.assign(ndown, nullptr);
    for (size_t l = 0; l < ndown; ++l) ALLOC(s.down_e[l], s.padded(s.levels[l].w, s.levels[l].h));
    const LevelView& B = s.levels[0];
    size_t pb = s.padded(B.w, B.h);
    ALLOC(s.search[0], pb); ALLOC(s.search[1], pb); ALLOC(s.As, pb); ALLOC(s.rhs, pb); ALLOC(s.r_alt, pb);
    ALLOC(s.As_lo, pb); ALLOC(s.r_lo, pb);
    // the coarse CG vectors of either mode's coarsest level
    size_t pcs = 8;
    for (const auto& hz : s.hier)
        if (hz.count > 0) pcs = std::max(pcs, s.padded(s.levels[hz.count - 1].w, s.levels[hz.count - 1].h));
    ALLOC(s.cg_s, pcs); ALLOC(s.cg_ap, pcs); ALLOC(s.cg_r, pcs);
    ALLOC(s.pcg, 1);
    const size_t ngam = std::max<size_t>(kRedBlocks, gamma_partials(s));
    ALLOC(s.part_gamma, ngam); ALLOC(s.part_delta, kRedBlocks); ALLOC(s.part_max, kRedBlocks);
    ALLOC(s.part_fluid, 2 * kRedBlocks);
    ALLOC(s.part_gamma_d, ngam); ALLOC(s.part_asz_d, ngam); ALLOC(s.part_rr_t, ngam);
    ALLOC(s.tile_flag, gamma_partials(s)); ALLOC(s.tile_order, gamma_partials(s));
    ALLOC(s.part_delta_d, kRedBlocks); ALLOC(s.part_rr_d, kRedBlocks);
    ALLOC(s.part_rr2_d, kRedBlocks); ALLOC(s.pressure_d, nc);
    s.part_maxv_cap = 4096;  // G2P's workgroups: 2 particles per thread and step up to 2M particles
    ALLOC(s.part_maxv, s.part_maxv_cap);
    ALLOC(s.fstats, 1);
    ALLOC(s.maxv_bits, 1);
    ALLOC(s.step, 1);
    return APIC_OK;
    };
    apic_status cst;
    s.slab_pass = 1;
    s.slab_size = 0;
    if ((cst = carve()) != APIC_OK) return cst;
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, s.slab_size);
        if (e != hipSuccess)
            return fail(e == hipErrorOutOfMemory ? APIC_ERR_OUT_OF_MEMORY : APIC_ERR_HIP,
                        std::string("hipMalloc: ") + hipGetErrorString(e));
        s.allocations.push_back(q);
        s.device_bytes += s.slab_size;
        s.slab = static_cast<uint8_t*>(q);
        HIP_TRY(hipMemset(q, 0, s.slab_size));
    }
    s.slab_pass = 2;
    s.slab_used = 0;
    cst = carve();
    s.slab_pass = 0;
    if (cst != APIC_OK) return cst;
    // coherent, mapped host memory (one block): kernels write these mirrors directly
    {
        const unsigned hflags = hipHostMallocMapped | hipHostMallocCoherent;
        auto al = [](size_t v) { return (v + 63) / 64 * 64; };
        const size_t o_maxv = al(sizeof(PcgState)), o_fs = o_maxv + al(sizeof(uint32_t));
        const size_t o_step = o_fs + al(sizeof(FluidStats)), total = o_step + al(sizeof(StepState));
        HIP_TRY(hipHostMalloc((void**)&s.h_block, total, hflags));
        std::memset(s.h_block, 0, total);
        uint8_t* hd = nullptr;
        HIP_TRY(hipHostGetDevicePointer((void**)&hd, s.h_block, 0));
        s.h_pcg = reinterpret_cast<PcgState*>(s.h_block);
        s.h_maxv = reinterpret_cast<uint32_t*>(s.h_block + o_maxv);
        s.h_fstats = reinterpret_cast<FluidStats*>(s.h_block + o_fs);
        s.h_step = reinterpret_cast<StepState*>(s.h_block + o_step);
        s.h_pcg_d = reinterpret_cast<PcgState*>(hd);
        s.h_maxv_d = reinterpret_cast<uint32_t*>(hd + o_maxv);
        s.h_fstats_d = reinterpret_cast<FluidStats*>(hd + o_fs);
        s.h_step_d = reinterpret_cast<StepState*>(hd + o_step);
    }
    HIP_TRY(hipEventCreateWithFlags(&s.e_sync, hipEventDisableTiming));
    {
        StepState st0{};
        st0.prev_dt = s.prev_dt;  // 1/60 before the first update (:824)
        HIP_TRY(hipMemcpy(s.step, &st0, sizeof(StepState), hipMemcpyHostToDevice));
    }

    // upload particles (tail lanes of the last chunk repeat the last particle, :887-898)
    if (s.nalloc) {
        std::vector<float> px(s.nalloc), py(s.nalloc);
        for (size_t i = 0; i < s.nalloc; ++i) {
            size_t k = std::min<size_t>(i, s.num_particles - 1);
            px[i] = pos[2 * k];
            py[i] = pos[2 * k + 1];
        }
        HIP_TRY(hipMemcpy(s.pfield(P_X, 0), px.data(), s.nalloc * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s.pfield(P_Y, 0), py.data(), s.nalloc * 4, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(s.obstacle, obstacle.data(), s.cells, hipMemcpyHostToDevice));
    // update_world (:913) runs only when the scene has particles
    if (s.num_particles > 0) launch_open_faces(s);
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(s.stream));
    s.cur = 0;
    s.keys_valid = false;
    s.maxv_valid = true;  // velocities start at zero
    *s.h_maxv = 0;
    return APIC_OK;
}

float host_maxv(Sim& s) {
    float f;
    std::memcpy(&f, s.h_maxv, 4);
    return f;
}

apic_status refresh_maxv(Sim& s) {
    if (!s.maxv_valid) {
        launch_max_velocity(s);  // the finalize kernel also writes the host mirror
        CHECK_LAUNCH();
        s.maxv_valid = true;
    }
    HIP_TRY(hipStreamSynchronize(s.stream));
    return APIC_OK;
}

// ||r||_2 / ||b||_2 at the end of a relative-mode solve, else -1.
double rel_residual(const Sim& s) {
    if (s.solver.mode != 1) return -1.0;
    const PcgState& p = *s.h_pcg;
    return p.bnorm2 > 0.0 ? std::sqrt(p.rnorm2 / p.bnorm2) : 0.0;
}

// After a failed update: a deferred gather is done, so that the records are
// in sorted order as the reference leaves them (best effort: the stream may
// hold the error).
void settle_particles(Sim& s) {
    if (!s.perm_pending) return;
    launch_permute(s);
    s.perm_pending = false;
    (void)hipGetLastError();
}

// The sort of update_spatial_lookup (src/apic.rs:1049-1103).  `defer`: the
// gather of the particle records into sorted order is left to the substep's
// P2G (which reads through the permutation) and G2P (which writes the sorted
// order), saving a pass over every record.
apic_status do_sort(Sim& s, bool defer) {
    if (s.num_particles == 0) return APIC_OK;
    if (!s.keys_valid) launch_keygen(s);
    launch_radix_sort(s);
    if (defer) s.perm_pending = true;
    else launch_permute(s);
    launch_cell_ranges(s);
    s.cells_valid = true;
    s.keys_valid = false;
    CHECK_LAUNCH();
    return APIC_OK;
}

// Iterations to queue before the first check of the next solve: as many as
// this solve took.
int next_batch(const Sim& s) { return s.h_pcg->iters; }

// The solve's counters, from the host mirror.
apic_status pcg_outcome(Sim& s) {
    s.stats.pcg_iterations = (uint32_t)s.h_pcg->iters;
    s.stats.pcg_exit_reason = (uint32_t)s.h_pcg->exit_reason;
    s.stats.pcg_max_residual = s.h_pcg->maxerr;
    s.stats.pcg_rel_residual = rel_residual(s);
    s.pcg_batch = next_batch(s);
    return APIC_OK;
}

// One PCG iteration (src/apic.rs:1899-1955): the SpMV and dot, the update
// and the V-cycle with the next r.z.  The kernel timer brackets iteration 1
// of each solve (the first with every operand).
void pcg_iteration(Sim& s, int k) {
    const bool timed = k == 1;
    if (timed) s.mark(APIC_T_K_PCG_ITER, 0);
    launch_pcg_iteration_head(s, k);
    launch_vcycle(s, 2, k);
    if (timed) s.mark(APIC_T_K_PCG_ITER, 1);
}

// solve_pcg (src/apic.rs:1865-1956) with the reference's rules.  Iterations
// are queued in batches; the device-side state machine turns the kernels of
// a finished solve into no-ops, and the host reads the state only at the end
// of a batch.  The first batch is as long as the previous solve.
apic_status do_pcg(Sim& s, bool zero_pressure) {
    if (zero_pressure) launch_zero_pressure(s);
    launch_pcg_reset(s);
    launch_vcycle(s, 1);
    CHECK_LAUNCH();
    int k = 0;
    int check_at = std::max(0, s.pcg_batch);
    while (true) {
        while (k < check_at && k < pcg_cap(s)) pcg_iteration(s, k++);
        launch_pcg_decide(s, k);  // also writes the host mirror
        CHECK_LAUNCH();
        HIP_TRY(hipStreamSynchronize(s.stream));
        if (s.h_pcg->done || k >= pcg_cap(s)) break;
        check_at = k + 2;
    }
    if (apic_status st_ = pcg_outcome(s); st_ != APIC_OK) return st_;
    return APIC_OK;
}

// One substep of Apic::update (src/apic.rs:2444-2461), queued without
// waiting for the device except once, for the pressure solve's state: the
// substep's dt is computed on the device, and the stages after the solve are
// queued right behind the first batch of PCG iterations, guarded by the
// solve's done flag.  The host waits on the batch while the device goes on;
// if the solve needs more iterations, the guarded stages have turned into
// no-ops, the host-side bookkeeping they did is undone, and they are queued
// again after the remaining iterations.
apic_status substep(Sim& s, bool sort, bool first, float frame_dt) {
    apic_status st;
    s.step_src = s.step;
    launch_substep_begin(s, first, frame_dt);
    if (sort) {
        s.mark(APIC_T_SORT, 0);
        if ((st = do_sort(s, true)) != APIC_OK) return st;
        s.mark(APIC_T_SORT, 1);
    }
    s.mark(APIC_T_P2G, 0);
    launch_p2g(s, true, 0.0f);
    s.mark(APIC_T_P2G, 1);
    s.mark(APIC_T_PRESSURE_BUILD, 0);
    launch_build_pressure(s, true, true);
    s.mark(APIC_T_PRESSURE_BUILD, 1);
    s.mark(APIC_T_MG_BUILD, 0);
    launch_mg_build(s);
    s.mark(APIC_T_MG_BUILD, 1);
    CHECK_LAUNCH();
    // pressure solve: first batch as long as the previous solve
    s.mark(APIC_T_PCG, 0);
    launch_pcg_reset(s);
    launch_vcycle(s, 1);
    int k = 0;
    for (; k < std::max(0, s.pcg_batch) && k < pcg_cap(s); ++k) pcg_iteration(s, k);
    launch_pcg_decide(s, k);
    s.mark(APIC_T_PCG, 1);
    HIP_TRY(hipEventRecord(s.e_sync, s.stream));
    CHECK_LAUNCH();
    // the rest of the substep, speculatively
    struct Saved {
        float *mac_u, *mac_v, *old_u, *old_v;
        uint8_t *valid_u, *valid_v, *nvalid_u, *nvalid_v;
        bool cells_valid, keys_valid, maxv_valid, perm_pending;
        int cur;
        size_t marks;
    } saved{s.mac_u, s.mac_v, s.old_u, s.old_v, s.valid_u, s.valid_v, s.nvalid_u, s.nvalid_v,
            s.cells_valid, s.keys_valid, s.maxv_valid, s.perm_pending, s.cur, s.ev_pending.size()};
    auto tail = [&]() -> apic_status {
        s.mark(APIC_T_GRADIENT, 0);
        launch_gradient(s, true, 0.0f);
        s.mark(APIC_T_GRADIENT, 1);
        s.mark(APIC_T_EXTRAPOLATE, 0);
        launch_extrapolate(s);
        s.mark(APIC_T_EXTRAPOLATE, 1);
        s.mark(APIC_T_G2P, 0);
        launch_g2p(s, 0.0f, true, true);  // and the substep's end
        s.mark(APIC_T_G2P, 1);
        s.cells_valid = false;
        s.keys_valid = true;
        s.maxv_valid = true;
        CHECK_LAUNCH();
        return APIC_OK;
    };
    s.tail_guard = &s.pcg->done;
    st = tail();
    s.tail_guard = nullptr;
    if (st != APIC_OK) { s.step_src = nullptr; return st; }
    HIP_TRY(hipEventSynchronize(s.e_sync));
    if (!s.h_pcg->done && k < pcg_cap(s)) {
        s.mac_u = saved.mac_u; s.mac_v = saved.mac_v; s.old_u = saved.old_u; s.old_v = saved.old_v;
        s.valid_u = saved.valid_u; s.valid_v = saved.valid_v; s.nvalid_u = saved.nvalid_u; s.nvalid_v = saved.nvalid_v;
        s.cells_valid = saved.cells_valid; s.keys_valid = saved.keys_valid; s.maxv_valid = saved.maxv_valid;
        s.perm_pending = saved.perm_pending; s.cur = saved.cur;
        s.drop_marks(saved.marks);
        int check_at = k + 2;
        while (true) {
            while (k < check_at && k < pcg_cap(s)) pcg_iteration(s, k++);
            launch_pcg_decide(s, k);
            HIP_TRY(hipEventRecord(s.e_sync, s.stream));
            CHECK_LAUNCH();
            HIP_TRY(hipEventSynchronize(s.e_sync));
            if (s.h_pcg->done || k >= pcg_cap(s)) break;
            check_at = k + 2;
        }
        if ((st = tail()) != APIC_OK) { s.step_src = nullptr; return st; }
    }
    s.step_src = nullptr;
    if (apic_status st_ = pcg_outcome(s); st_ != APIC_OK) return st_;
    s.collect_timers(false);
    return APIC_OK;
}

struct FieldDesc {
    void* ptr;
    size_t count;
    size_t elem;
};

bool field_desc(Sim& s, int f, uint32_t level, FieldDesc& d) {
    d = FieldDesc{nullptr, 0, 4};
    if (f >= APIC_POS_X && f <= APIC_TIME_RESIDUAL) {
        d.ptr = s.pfield(f - APIC_POS_X);
        d.count = s.nalloc;
        return true;
    }
    const LevelView* L = level < s.levels.size() ? &s.levels[level] : nullptr;
    switch (f) {
        case APIC_SORT_KEY: d.ptr = s.keys[0]; d.count = s.num_particles; return true;
        case APIC_SORT_VAL: d.ptr = s.vals[0]; d.count = s.num_particles; return true;
        case APIC_MAC_U: d.ptr = s.mac_u; d.count = s.nu(); return true;
        case APIC_MAC_V: d.ptr = s.mac_v; d.count = s.nv(); return true;
        case APIC_VALID_U: d.ptr = s.valid_u; d.count = s.nu(); d.elem = 1; return true;
        case APIC_VALID_V: d.ptr = s.valid_v; d.count = s.nv(); d.elem = 1; return true;
        case APIC_DENSITY: d.ptr = s.density; d.count = s.cells; return true;
        case APIC_SMOOTHED_DENSITY: d.ptr = s.smooth_a; d.count = s.cells; return true;
        case APIC_PRESSURE: d.ptr = s.pressure; d.count = s.cells; return true;
        case APIC_OBSTACLE: d.ptr = s.obstacle; d.count = s.cells; d.elem = 1; return true;
        case APIC_OPEN_U: d.ptr = s.open_u; d.count = s.nu(); d.elem = 1; return true;
        case APIC_OPEN_V: d.ptr = s.open_v; d.count = s.nv(); d.elem = 1; return true;
        case APIC_EXT_FORCE_U: d.ptr = s.ext_u; d.count = s.nu(); return true;
        case APIC_EXT_FORCE_V: d.ptr = s.ext_v; d.count = s.nv(); return true;
        case APIC_POST_FORCE_U: d.ptr = s.post_u; d.count = s.nu(); return true;
        case APIC_POST_FORCE_V: d.ptr = s.post_v; d.count = s.nv(); return true;
        case APIC_SEARCH: d.ptr = s.search[s.h_pcg ? (s.h_pcg->iters & 1) : 0]; d.count = s.padded(s.W, s.H); return true;
        case APIC_MATRIX_TIMES_SEARCH: d.ptr = s.As; d.count = s.padded(s.W, s.H); return true;
        case APIC_PRESSURE_F64: d.ptr = s.pressure_d; d.count = s.cells; d.elem = 8; return true;
        case APIC_RHS: d.ptr = s.rhs; d.count = s.padded(s.W, s.H); return true;
        default: break;
    }
    if (!L) return false;
    const size_t pa = s.padded(L->w, L->h);
    switch (f) {
        case APIC_FLUID: d.ptr = L->fluid; d.count = L->cells; d.elem = 1; return true;
        case APIC_DIAG: d.ptr = L->diag; d.count = pa; return true;
        case APIC_INV_DIAG: d.ptr = L->inv_diag; d.count = pa; return true;
        case APIC_PLUS_X: d.ptr = L->px; d.count = pa; return true;
        case APIC_PLUS_Y: d.ptr = L->py; d.count = pa; return true;
        case APIC_RESIDUAL: d.ptr = L->r; d.count = pa; return true;
        case APIC_ERROR: d.ptr = L->e; d.count = pa; return true;
        default: return false;
    }
}

apic_status create(uint32_t width, uint32_t height, float cell_size, const apic_world_properties* world_properties,
                   bool explicit_set, const float* xy, uint32_t n, apic_handle** out, apic_batch* batch = nullptr) {
    if (!out || !world_properties) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (xy == nullptr && n != 0) return fail(APIC_ERR_INVALID_ARGUMENT, "xy is null but n > 0");
    apic_handle* h = new (std::nothrow) apic_handle();
    if (!h) return fail(APIC_ERR_OUT_OF_MEMORY, "host allocation failed");
    h->batch = batch;
    apic_status st = build(h->sim, width, height, cell_size, *world_properties, explicit_set, xy, n,
                           batch ? batch->stream : nullptr);
    if (st != APIC_OK) {
        release(h->sim);
        delete h;
        return st;
    }
    *out = h;
    if (batch) batch->members += 1;
    return APIC_OK;
}

// ---- batched stepping

// The arena of apic_batch_update_noalloc: the launch tables and workgroup
// maps of one substep of a batched update, in a static device array instead
// of memory the batch allocates.  Every GPU the library runs on holds its own
// copy until the process ends.  Every no-allocation call on that GPU shares
// it: the calls are serialised on the host (g_noalloc_call), and a call's
// stream waits for the event that ended the previous call (another batch's
// stream, perhaps) before its first copy into the arena.  Nothing in it has
// to start at zero: a substep writes its tables before its kernels read them.
constexpr size_t kNoAllocArenaBytes = (size_t)APIC_NOALLOC_ARENA_BYTES;
static __device__ __align__(256) uint8_t scratch[kNoAllocArenaBytes];

// Device address of `scratch` on each device, looked up once per device
// with that device current (host code never uses the symbol itself as a
// device pointer), and the per-device order of the calls.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;                 // guards the two arrays below
uint8_t* g_scratch_ptr[kMaxDevices] = {};
apic_batch* g_arena_last[kMaxDevices] = {};  // batch whose e_arena ends the last call
std::mutex g_noalloc_call[kMaxDevices];

apic_status noalloc_arena(int device, uint8_t** out) {
    if (device < 0 || device >= kMaxDevices) return fail(APIC_ERR_INVALID_ARGUMENT, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int prev = 0;
        HIP_TRY(hipGetDevice(&prev));
        if (prev != device) HIP_TRY(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) HIP_TRY(hipSetDevice(prev));
        if (e != hipSuccess || !p)
            return fail(APIC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<uint8_t*>(p);
    }
    *out = g_scratch_ptr[device];
    return APIC_OK;
}

// Uploads host[0, used) to the batch's device arena (in stream order, before
// the kernels that read it).
// Uploads the launch tables host[from, used) to the same offsets of the
// device arena, in stream order: the kernels already issued read the tables
// before `from`, which stay as they are (a new arena receives all of them).
apic_status batch_upload(apic_batch& B, size_t from, size_t used) {
    if (from >= used) return APIC_OK;
    if (B.fixed) {
        // the no-allocation path: a copy from pageable host memory (the
        // source can be reused when the call returns) into the fixed arena,
        // in stream order after every kernel that still reads the arena
        if (used > B.fixed_cap)
            return fail(APIC_ERR_INVALID_ARGUMENT,
                        "batched launch tables need " + std::to_string(used) + " bytes, more than the " +
                            std::to_string(B.fixed_cap) + " bytes of the no-allocation arena");
        HIP_TRY(hipMemcpyAsync(B.fixed + from, B.host.data() + from, used - from, hipMemcpyHostToDevice, B.stream));
        return APIC_OK;
    }
    if (B.copy_pending) {
        HIP_TRY(hipEventSynchronize(B.e_copied));
        B.copy_pending = false;
    }
    if (used > B.pinned_cap) {
        if (B.pinned) HIP_TRY(hipHostFree(B.pinned));
        B.pinned = nullptr;
        B.pinned_cap = 0;
        const size_t cap = std::max<size_t>(used + used / 2, 1 << 20);
        HIP_TRY(hipHostMalloc((void**)&B.pinned, cap, hipHostMallocDefault));
        B.pinned_cap = cap;
    }
    if (used > B.dev_cap) {
        HIP_TRY(hipStreamSynchronize(B.stream));  // kernels may still read the old arena
        if (B.dev) HIP_TRY(hipFree(B.dev));
        B.dev = nullptr;
        B.dev_cap = 0;
        const size_t cap = std::max<size_t>(used + used / 2, 1 << 20);
        HIP_TRY(hipMalloc((void**)&B.dev, cap));
        B.dev_cap = cap;
        from = 0;
    }
    std::memcpy(B.pinned + from, B.host.data() + from, used - from);
    HIP_TRY(hipMemcpyAsync(B.dev + from, B.pinned + from, used - from, hipMemcpyHostToDevice, B.stream));
    HIP_TRY(hipEventRecord(B.e_copied, B.stream));
    B.copy_pending = true;
    return APIC_OK;
}

// Compiles what the recorder holds into B.host at `used`.
apic_status batch_compile(apic_batch& B, size_t& used, std::vector<Recorder::Launch>& out) {
    if (B.rec.overflow()) {
        B.rec.clear();
        return fail(APIC_ERR_UNSUPPORTED_SIZE, "batched launches: a simulation needs more than 2^20 - 1 workgroups");
    }
    out = B.rec.compile(B.host, used);
    if (out.empty() && !B.rec.empty())
        return fail(APIC_ERR_HIP, "batched launches: the simulations' launch orders conflict");
    B.rec.clear();
    return APIC_OK;
}

#define HIP_LAUNCH_TRY(expr)                                                                   \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(APIC_ERR_HIP, std::string("batched launch: ") + hipGetErrorString(e_)); \
    } while (0)

// Apic::update for several simulations at once: the substeps of all of them
// that still have frame time left are stepped together.  Per substep:
// the head (dt, sort, P2G, pressure system, hierarchy, PCG start), the PCG
// iterations in lockstep until every solve has ended (each simulation's
// kernels return at once after its own exit), then the tail (gradient,
// extrapolation, G2P).  The launches of one simulation are those of substep()
// in the same order, so it computes the same bits as when stepped alone.
apic_status batch_update(apic_batch& B, const std::vector<Sim*>& sims, bool disable_sort,
                         const std::vector<float>& frame_dt) {
    const size_t n = sims.size();
    for (Sim* s : sims)
        if (!s->maxv_valid) {
            launch_max_velocity(*s);
            s->maxv_valid = true;
        }
    CHECK_LAUNCH();
    std::vector<float> t(n, 0.0f), last_maxv(n, 0.0f);
    std::vector<uint32_t> substeps(n, 0), total_it(n, 0);
    for (uint32_t round = 0;; ++round) {
        std::vector<size_t> act;
        for (size_t i = 0; i < n; ++i)
            if (t[i] < frame_dt[i]) act.push_back(i);
        if (act.empty()) break;
        const bool first = round == 0;
        const bool sort = !disable_sort && first;
        std::vector<Recorder::Launch> head, iter, decide, tail;
        size_t used = 0;
        apic_status st;
        // head
        for (size_t r = 0; r < act.size(); ++r) {
            Sim& s = *sims[act[r]];
            B.rec.begin((uint32_t)r);
            s.rec = &B.rec;
            s.step_src = s.step;
            launch_substep_begin(s, first, frame_dt[act[r]]);
            if (sort) do_sort(s, true);
            launch_p2g(s, true, 0.0f);
            launch_build_pressure(s, true, true);
            launch_mg_build(s);
            launch_pcg_reset(s);
            launch_vcycle(s, 1);
            s.rec = nullptr;
        }
        if ((st = batch_compile(B, used, head)) != APIC_OK) return st;
        // the head runs while the rest is recorded
        if ((st = batch_upload(B, 0, used)) != APIC_OK) return st;
        HIP_LAUNCH_TRY(issue(head, B.arena(), B.stream));
        const size_t used_head = used;
        // one PCG iteration (issued with k = 0, 1, ...) and the check
        for (size_t r = 0; r < act.size(); ++r) {
            Sim& s = *sims[act[r]];
            B.rec.begin((uint32_t)r);
            s.rec = &B.rec;
            launch_pcg_iteration_head(s, 0);
            launch_vcycle(s, 2);
            s.rec = nullptr;
        }
        if ((st = batch_compile(B, used, iter)) != APIC_OK) return st;
        for (size_t r = 0; r < act.size(); ++r) {
            Sim& s = *sims[act[r]];
            B.rec.begin((uint32_t)r);
            s.rec = &B.rec;
            launch_pcg_decide(s, 0);
            s.rec = nullptr;
        }
        if ((st = batch_compile(B, used, decide)) != APIC_OK) return st;
        // tail (issued once every solve has ended)
        for (size_t r = 0; r < act.size(); ++r) {
            Sim& s = *sims[act[r]];
            B.rec.begin((uint32_t)r);
            s.rec = &B.rec;
            s.tail_guard = nullptr;
            launch_gradient(s, true, 0.0f);
            launch_extrapolate(s);
            launch_g2p(s, 0.0f, true, true);  // and the substep's end
            s.cells_valid = false;
            s.keys_valid = true;
            s.maxv_valid = true;
            s.rec = nullptr;
        }
        if ((st = batch_compile(B, used, tail)) != APIC_OK) return st;
        if ((st = batch_upload(B, used_head, used)) != APIC_OK) return st;
        int batch_len = 0, cap = 0;
        std::vector<int> prev;
        for (size_t i : act) {
            batch_len = std::max(batch_len, sims[i]->pcg_batch);
            cap = std::max(cap, pcg_cap(*sims[i]));
            prev.push_back(sims[i]->pcg_batch);
        }
        // Checks: where half of the solves ended in the last substep (so that
        // the ended ones can leave the program early), where the last one did,
        // then every two iterations.
        std::nth_element(prev.begin(), prev.begin() + prev.size() / 2, prev.end());
        const int median = prev[prev.size() / 2];
        int k = 0, check_at = median > 0 && median < batch_len ? median : batch_len;
        size_t in_program = act.size();
        while (true) {
            for (; k < check_at && k < cap; ++k) HIP_LAUNCH_TRY(issue(iter, B.arena(), B.stream, k));
            HIP_LAUNCH_TRY(issue(decide, B.arena(), B.stream, k));
            HIP_TRY(hipEventRecord(B.e_sync, B.stream));
            HIP_TRY(hipEventSynchronize(B.e_sync));
            std::vector<size_t> live;
            for (size_t i : act)
                if (!sims[i]->h_pcg->done && k < pcg_cap(*sims[i])) live.push_back(i);
            if (live.empty() || k >= cap) break;
            check_at = k < batch_len ? batch_len : k + 2;
            // Once a quarter of the program's solves have ended, the iteration
            // and check programs are recorded again for the others alone: an
            // ended solve's workgroups would still be dispatched, each to
            // return after one load.  The next iteration of the program in
            // hand keeps the GPU busy meanwhile, and a simulation's launches
            // are the same in either program, so its bits are too.
            if (live.size() * 4 <= in_program * 3) {
                if (k < check_at && k < cap) HIP_LAUNCH_TRY(issue(iter, B.arena(), B.stream, k++));
                const size_t used0 = used;
                std::vector<Recorder::Launch> iter0 = iter, decide0 = decide;
                for (size_t r = 0; r < live.size(); ++r) {
                    Sim& s = *sims[live[r]];
                    B.rec.begin((uint32_t)r);
                    s.rec = &B.rec;
                    launch_pcg_iteration_head(s, 0);
                    launch_vcycle(s, 2);
                    s.rec = nullptr;
                }
                if ((st = batch_compile(B, used, iter)) != APIC_OK) return st;
                for (size_t r = 0; r < live.size(); ++r) {
                    Sim& s = *sims[live[r]];
                    B.rec.begin((uint32_t)r);
                    s.rec = &B.rec;
                    launch_pcg_decide(s, 0);
                    s.rec = nullptr;
                }
                if ((st = batch_compile(B, used, decide)) != APIC_OK) return st;
                if (B.fixed && used > B.fixed_cap) {
                    // the fixed arena has no room for the shorter program:
                    // keep the full one (its ended solves' kernels return at once)
                    used = used0;
                    iter.swap(iter0);
                    decide.swap(decide0);
                } else if ((st = batch_upload(B, used0, used)) != APIC_OK) {
                    return st;
                }
                in_program = live.size();
            }
        }
        HIP_LAUNCH_TRY(issue(tail, B.arena(), B.stream));
        for (size_t i : act) {
            Sim& s = *sims[i];
            s.step_src = nullptr;
            if ((st = pcg_outcome(s)) != APIC_OK) return st;
            const StepState hs = *s.h_step;  // written by the head, complete after the wait
            last_maxv[i] = hs.maxv;
            t[i] += hs.dt;
            s.prev_dt = hs.dt;
            s.timestamp += 1;
            substeps[i] += 1;
            total_it[i] += s.stats.pcg_iterations;
        }
    }
    for (size_t i = 0; i < n; ++i) {
        Sim& s = *sims[i];
        if (s.forces_active) {
            launch_fill_f32(s, s.ext_u, 0.0f, s.nu());
            launch_fill_f32(s, s.ext_v, 0.0f, s.nv());
            launch_fill_f32(s, s.post_u, 0.0f, s.nu());
            launch_fill_f32(s, s.post_v, 0.0f, s.nv());
            s.forces_active = false;
        }
        s.stats.substeps = substeps[i];
        s.stats.pcg_iterations_total = total_it[i];
        s.stats.fluid_cells = s.h_fstats->count;
        s.stats.timestamp = s.timestamp;
        s.stats.prev_deltatime = s.prev_dt;
        s.stats.max_particle_velocity = last_maxv[i];
    }
    CHECK_LAUNCH();
    return APIC_OK;
}

}  // namespace

extern "C" {

const char* apic_last_error(void) { return g_last_error.c_str(); }

apic_status apic_create_with_particles(uint32_t width, uint32_t height, float cell_size,
                                       const apic_world_properties* world_properties, const float* xy, uint32_t n,
                                       apic_handle** out) {
    return create(width, height, cell_size, world_properties, true, xy, n, out);
}

apic_status apic_create(uint32_t width, uint32_t height, float cell_size,
                        const apic_world_properties* world_properties, apic_handle** out) {
    return create(width, height, cell_size, world_properties, false, nullptr, 0, out);
}

void apic_destroy(apic_handle* h) {
    if (!h) return;
    release(h->sim);
    if (h->batch) h->batch->members -= 1;
    delete h;
}

apic_status apic_set_gravity_x(apic_handle* h, float gx) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    if (!std::isfinite(gx)) return fail(APIC_ERR_INVALID_ARGUMENT, "gx must be finite");
    h->sim.gravity_x = gx;
    return APIC_OK;
}

apic_status apic_device_bytes(apic_handle* h, size_t* bytes) {
    if (!h || !bytes) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    *bytes = h->sim.device_bytes;
    return APIC_OK;
}

apic_status apic_device_memory(size_t* free_bytes, size_t* total_bytes) {
    size_t f = 0, t = 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(APIC_ERR_NO_DEVICE, "no HIP device found");
    HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return APIC_OK;
}

apic_status apic_batch_create(apic_batch** out) {
    if (!out) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(APIC_ERR_NO_DEVICE, "no HIP device found");
    apic_batch* b = new (std::nothrow) apic_batch();
    if (!b) return fail(APIC_ERR_OUT_OF_MEMORY, "host allocation failed");
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->e_sync, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->e_copied, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->e_arena, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (b->e_copied) (void)hipEventDestroy(b->e_copied);
        if (b->e_sync) (void)hipEventDestroy(b->e_sync);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return fail(APIC_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
    }
    *out = b;
    return APIC_OK;
}

apic_status apic_batch_destroy(apic_batch* b) {
    if (!b) return APIC_OK;
    if (b->members) return fail(APIC_ERR_INVALID_ARGUMENT, "the batch still has simulations");
    (void)hipStreamSynchronize(b->stream);
    if (b->device >= 0 && b->device < kMaxDevices) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (g_arena_last[b->device] == b) g_arena_last[b->device] = nullptr;
    }
    if (b->pinned) (void)hipHostFree(b->pinned);
    if (b->dev) (void)hipFree(b->dev);
    (void)hipEventDestroy(b->e_sync);
    (void)hipEventDestroy(b->e_copied);
    (void)hipEventDestroy(b->e_arena);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return APIC_OK;
}

apic_status apic_batch_create_sim(apic_batch* b, uint32_t width, uint32_t height, float cell_size,
                                  const apic_world_properties* world_properties, const float* xy, uint32_t n,
                                  apic_handle** out) {
    if (!b) return fail(APIC_ERR_INVALID_ARGUMENT, "null batch");
    return create(width, height, cell_size, world_properties, xy != nullptr, xy, xy ? n : 0, out, b);
}

// The simulations of a batch update, checked.
apic_status batch_members(apic_batch* b, apic_handle* const* sims, uint32_t count, const float* frame_dt,
                          std::vector<Sim*>& v) {
    if (!sims || !frame_dt) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    if (count > kMaxBatchRows) return fail(APIC_ERR_INVALID_ARGUMENT, "more than 4096 simulations in one update");
    v.assign(count, nullptr);
    for (uint32_t i = 0; i < count; ++i) {
        if (!sims[i] || sims[i]->batch != b)
            return fail(APIC_ERR_INVALID_ARGUMENT, "a simulation is not part of this batch");
        v[i] = &sims[i]->sim;
        for (uint32_t j = 0; j < i; ++j)
            if (v[j] == v[i]) return fail(APIC_ERR_INVALID_ARGUMENT, "a simulation is listed twice");
    }
    return APIC_OK;
}

apic_status apic_batch_update(apic_batch* b, apic_handle* const* sims, uint32_t count,
                              const apic_debug_options* opts, const float* frame_dt) {
    if (!b) return fail(APIC_ERR_INVALID_ARGUMENT, "null batch");
    if (count == 0) return APIC_OK;
    std::vector<Sim*> v;
    apic_status st = batch_members(b, sims, count, frame_dt, v);
    if (st != APIC_OK) return st;
    std::vector<float> fdt(frame_dt, frame_dt + count);
    return batch_update(*b, v, opts && opts->disable_particle_sort, fdt);
}

apic_status apic_batch_update_noalloc(apic_batch* b, apic_handle* const* sims, uint32_t count,
                                      const apic_debug_options* opts, const float* frame_dt) {
    if (!b) return fail(APIC_ERR_INVALID_ARGUMENT, "null batch");
    if (count == 0) return APIC_OK;
    std::vector<Sim*> v;
    apic_status st = batch_members(b, sims, count, frame_dt, v);
    if (st != APIC_OK) return st;
    uint8_t* arena = nullptr;
    if ((st = noalloc_arena(b->device, &arena)) != APIC_OK) return st;
    std::vector<float> fdt(frame_dt, frame_dt + count);
    std::lock_guard<std::mutex> call(g_noalloc_call[b->device]);
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        apic_batch* last = g_arena_last[b->device];
        if (last && last != b) HIP_TRY(hipStreamWaitEvent(b->stream, last->e_arena, 0));
    }
    b->fixed = arena;
    b->fixed_cap = kNoAllocArenaBytes;
    st = batch_update(*b, v, opts && opts->disable_particle_sort, fdt);
    b->fixed = nullptr;
    b->fixed_cap = 0;
    HIP_TRY(hipEventRecord(b->e_arena, b->stream));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    g_arena_last[b->device] = b;
    return st;
}

apic_status apic_update(apic_handle* h, const apic_debug_options* opts, float frame_dt) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    Sim& s = h->sim;
    const bool disable_sort = opts && opts->disable_particle_sort;
    // The substep loop of Apic::update (src/apic.rs:2431-2467).  The host keeps
    // the same f32 sum of the dts the device computes, to know when the frame
    // is done; it returns while the last substep may still run on the device.
    float t = 0.0f;
    uint32_t substeps = 0, total_it = 0;
    float last_maxv = 0.0f;
    apic_status st;
    if (!s.maxv_valid) {
        launch_max_velocity(s);
        CHECK_LAUNCH();
        s.maxv_valid = true;
    }
    while (t < frame_dt) {
        if ((st = substep(s, !disable_sort && substeps == 0, substeps == 0, frame_dt)) != APIC_OK) {
            settle_particles(s);
            return st;
        }
        const StepState hs = *s.h_step;  // complete: it was written before the solve the host waited for
        last_maxv = hs.maxv;
        t += hs.dt;
        s.prev_dt = hs.dt;
        s.timestamp += 1;
        substeps += 1;
        total_it += s.stats.pcg_iterations;
    }
    if (s.forces_active) {
        launch_fill_f32(s, s.ext_u, 0.0f, s.nu());
        launch_fill_f32(s, s.ext_v, 0.0f, s.nv());
        launch_fill_f32(s, s.post_u, 0.0f, s.nu());
        launch_fill_f32(s, s.post_v, 0.0f, s.nv());
        CHECK_LAUNCH();
        s.forces_active = false;
    }
    s.stats.substeps = substeps;
    s.stats.pcg_iterations_total = total_it;
    s.stats.fluid_cells = s.h_fstats->count;
    s.stats.timestamp = s.timestamp;
    s.stats.prev_deltatime = s.prev_dt;
    s.stats.max_particle_velocity = last_maxv;
    return APIC_OK;
}

apic_status apic_add_radial_force(apic_handle* h, float wx, float wy, float radius, float strength) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    launch_radial_force(h->sim, wx, wy, radius, strength);
    CHECK_LAUNCH();
    return APIC_OK;
}

apic_status apic_get_max_particle_velocity(apic_handle* h, float* out) {
    if (!h || !out) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    apic_status st = refresh_maxv(h->sim);
    if (st != APIC_OK) return st;
    *out = host_maxv(h->sim);
    return APIC_OK;
}

apic_status apic_is_fluid(apic_handle* h, int32_t i, int32_t j, uint64_t* out) {
    if (!h || !out) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    Sim& s = h->sim;
    if (i < 0 || i >= (int32_t)s.W || j < 0 || j >= (int32_t)s.H) { *out = 0; return APIC_OK; }
    uint8_t v = 0;
    HIP_TRY(hipStreamSynchronize(s.stream));
    HIP_TRY(hipMemcpy(&v, s.levels[0].fluid + (size_t)j * s.W + i, 1, hipMemcpyDeviceToHost));
    *out = v ? 1 : 0;
    return APIC_OK;
}

apic_status apic_set_solver_config(apic_handle* h, const apic_solver_config* cfg) {
    if (!h || !cfg) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    if (cfg->mode != 0 && cfg->mode != 1) return fail(APIC_ERR_INVALID_ARGUMENT, "solver mode must be 0 or 1");
    if (cfg->mode == 1) {
        if (!(cfg->rel_tol > 0.0f)) return fail(APIC_ERR_INVALID_ARGUMENT, "rel_tol must be positive");
        if (cfg->max_iterations < 1 || cfg->max_iterations > 100000)
            return fail(APIC_ERR_INVALID_ARGUMENT, "max_iterations must be in [1, 100000]");
        if (cfg->recompute_every < 0) return fail(APIC_ERR_INVALID_ARGUMENT, "recompute_every must be >= 0");
        if (cfg->fp64_reductions == 0)
            return fail(APIC_ERR_INVALID_ARGUMENT, "the relative mode runs its reductions in FP64 (fp64_reductions = 1)");
    }
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    h->sim.solver = *cfg;
    h->sim.select_hierarchy();
    return APIC_OK;
}

apic_status apic_set_profiling(apic_handle* h, int enable) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    Sim& s = h->sim;
    HIP_TRY(hipStreamSynchronize(s.stream));
    s.collect_timers(true);
    s.profiling = enable < 0 ? 0 : (enable > 2 ? 2 : enable);
    for (int t = 0; t < APIC_TIMER_COUNT; ++t) { s.timer_ms[t] = 0.0; s.timer_count[t] = 0; }
    return APIC_OK;
}

apic_status apic_get_timers(apic_handle* h, double* ms, uint32_t* counts, uint32_t n) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    if (n > APIC_TIMER_COUNT) return fail(APIC_ERR_INVALID_ARGUMENT, "n exceeds APIC_TIMER_COUNT");
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    h->sim.collect_timers(true);
    for (uint32_t t = 0; t < n; ++t) {
        if (ms) ms[t] = h->sim.timer_ms[t];
        if (counts) counts[t] = h->sim.timer_count[t];
    }
    return APIC_OK;
}

apic_status apic_get_stats(apic_handle* h, apic_step_stats* out) {
    if (!h || !out) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    *out = h->sim.stats;
    return APIC_OK;
}

apic_status apic_get_dims(apic_handle* h, uint32_t* width, uint32_t* height, uint32_t* num_particles,
                          uint32_t* num_chunks, uint32_t* num_levels) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    Sim& s = h->sim;
    if (width) *width = s.W;
    if (height) *height = s.H;
    if (num_particles) *num_particles = s.num_particles;
    if (num_chunks) *num_chunks = s.num_chunks;
    if (num_levels) *num_levels = (uint32_t)(s.ref_levels > 0 ? s.ref_levels - 1 : 0);
    return APIC_OK;
}

apic_status apic_get_level_counts(apic_handle* h, uint32_t* reference, uint32_t* relative) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    const Sim& s = h->sim;
    if (reference) *reference = (uint32_t)(s.ref_levels > 0 ? s.ref_levels - 1 : 0);
    if (relative) *relative = (uint32_t)(s.levels.size() - 1);
    return APIC_OK;
}

apic_status apic_get_timestamp(apic_handle* h, uint32_t* timestamp, float* prev_deltatime) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    if (timestamp) *timestamp = h->sim.timestamp;
    if (prev_deltatime) *prev_deltatime = h->sim.prev_dt;
    return APIC_OK;
}

apic_status apic_set_timestamp(apic_handle* h, uint32_t timestamp, float prev_deltatime) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    Sim& s = h->sim;
    HIP_TRY(hipStreamSynchronize(s.stream));
    s.timestamp = timestamp;
    s.prev_dt = prev_deltatime;
    StepState st{};
    HIP_TRY(hipMemcpy(&st, s.step, sizeof(StepState), hipMemcpyDeviceToHost));
    st.prev_dt = prev_deltatime;
    HIP_TRY(hipMemcpy(s.step, &st, sizeof(StepState), hipMemcpyHostToDevice));
    return APIC_OK;
}

apic_status apic_field_length(apic_handle* h, apic_field field, uint32_t level, size_t* count) {
    if (!h || !count) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    FieldDesc d;
    if (!field_desc(h->sim, field, level, d)) return fail(APIC_ERR_INVALID_ARGUMENT, "unknown field or level");
    *count = d.count;
    return APIC_OK;
}

apic_status apic_read_field(apic_handle* h, apic_field field, uint32_t level, void* dst, size_t count) {
    if (!h || (!dst && count)) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    Sim& s = h->sim;
    FieldDesc d;
    if (!field_desc(s, field, level, d)) return fail(APIC_ERR_INVALID_ARGUMENT, "unknown field or level");
    if (count != d.count) return fail(APIC_ERR_INVALID_ARGUMENT, "count does not match the field length");
    HIP_TRY(hipStreamSynchronize(s.stream));
    if (count) HIP_TRY(hipMemcpy(dst, d.ptr, count * d.elem, hipMemcpyDeviceToHost));
    return APIC_OK;
}

apic_status apic_write_field(apic_handle* h, apic_field field, uint32_t level, const void* src, size_t count) {
    if (!h || (!src && count)) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    Sim& s = h->sim;
    FieldDesc d;
    if (!field_desc(s, field, level, d)) return fail(APIC_ERR_INVALID_ARGUMENT, "unknown field or level");
    if (count != d.count) return fail(APIC_ERR_INVALID_ARGUMENT, "count does not match the field length");
    HIP_TRY(hipStreamSynchronize(s.stream));
    if (count) HIP_TRY(hipMemcpy(d.ptr, src, count * d.elem, hipMemcpyHostToDevice));
    s.tile_order_valid = false;  // the solver's flags may have changed: base tiles in their own order
    if (field >= APIC_POS_X && field <= APIC_TIME_RESIDUAL) s.cells_valid = false;
    if (field == APIC_POS_X || field == APIC_POS_Y) s.keys_valid = false;
    if (field == APIC_VEL_X || field == APIC_VEL_Y) s.maxv_valid = false;
    // the reference recomputes the keys from the positions at every sort (:1050-1063)
    if (field == APIC_SORT_KEY || field == APIC_SORT_VAL) s.keys_valid = false;
    if (field >= APIC_EXT_FORCE_U && field <= APIC_POST_FORCE_V) s.forces_active = true;
    if (field == APIC_FLUID && level == 0) {
        launch_fluid_stats(s);
        CHECK_LAUNCH();
        HIP_TRY(hipStreamSynchronize(s.stream));
    }
    return APIC_OK;
}

static apic_status aosoa_prepare(Sim& s, int group, size_t count) {
    if (group < APIC_PART_POSITIONS || group > APIC_PART_TIME_RESIDUAL)
        return fail(APIC_ERR_INVALID_ARGUMENT, "unknown particle group");
    if (count != aosoa_length(s, group)) return fail(APIC_ERR_INVALID_ARGUMENT, "count does not match the group length");
    if (!s.aosoa_staging && s.nalloc) {
        apic_status st = dalloc(s, s.aosoa_staging, 2 * s.nalloc);
        if (st != APIC_OK) return st;
    }
    return APIC_OK;
}

apic_status apic_read_particles_aosoa(apic_handle* h, apic_particle_group group, float* dst, size_t count) {
    if (!h || (!dst && count)) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    Sim& s = h->sim;
    apic_status st = aosoa_prepare(s, (int)group, count);
    if (st != APIC_OK) return st;
    if (!count) return APIC_OK;
    launch_pack_aosoa(s, (int)group, s.aosoa_staging);
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(s.stream));
    HIP_TRY(hipMemcpy(dst, s.aosoa_staging, count * 4, hipMemcpyDeviceToHost));
    return APIC_OK;
}

apic_status apic_write_particles_aosoa(apic_handle* h, apic_particle_group group, const float* src, size_t count) {
    if (!h || (!src && count)) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    Sim& s = h->sim;
    apic_status st = aosoa_prepare(s, (int)group, count);
    if (st != APIC_OK) return st;
    if (!count) return APIC_OK;
    HIP_TRY(hipStreamSynchronize(s.stream));
    HIP_TRY(hipMemcpy(s.aosoa_staging, src, count * 4, hipMemcpyHostToDevice));
    launch_unpack_aosoa(s, (int)group, s.aosoa_staging);
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(s.stream));
    // the same bookkeeping as a field write (apic_write_field)
    s.cells_valid = false;
    if (group == APIC_PART_POSITIONS) s.keys_valid = false;
    if (group == APIC_PART_VELOCITIES) s.maxv_valid = false;
    return APIC_OK;
}

apic_status apic_update_spatial_lookup(apic_handle* h) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    apic_status st = do_sort(h->sim, false);
    if (st != APIC_OK) return st;
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    return APIC_OK;
}

apic_status apic_transfer_particles_to_grid(apic_handle* h, float deltatime, uint32_t timestamp) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    (void)deltatime; (void)timestamp;  // unused by the reference as well (:1218)
    launch_p2g(h->sim, false, 0.0f);
    launch_fluid_stats(h->sim);
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    return APIC_OK;
}

apic_status apic_apply_external_forces(apic_handle* h, float deltatime) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    launch_forces(h->sim, deltatime);
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    return APIC_OK;
}

apic_status apic_build_pressure_system(apic_handle* h) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    launch_build_pressure(h->sim, false);
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    return APIC_OK;
}

apic_status apic_update_multilevel_grid(apic_handle* h) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    launch_mg_build(h->sim);
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    return APIC_OK;
}

apic_status apic_get_tuning(apic_handle* h, apic_tuning* out) {
    if (!h || !out) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    *out = h->sim.tuning;
    return APIC_OK;
}

apic_status apic_set_tuning(apic_handle* h, const apic_tuning* t) {
    if (!h || !t) return fail(APIC_ERR_INVALID_ARGUMENT, "null argument");
    if (t->radix_bits < 8 || t->radix_bits > 11) return fail(APIC_ERR_INVALID_ARGUMENT, "radix_bits must be in [8, 11]");
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    h->sim.tuning = apic_tuning{t->radix_bits, t->few_tiles, 0.0f, 0.0f};
    return APIC_OK;
}

// Each candidate runs once untimed, then reps times between two events on
// the simulation's stream; the kernel timers are off meanwhile.
apic_status apic_autotune(apic_handle* h, uint32_t reps, apic_tuning* out) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    if (reps == 0 || reps > 1000) return fail(APIC_ERR_INVALID_ARGUMENT, "reps must be in [1, 1000]");
    Sim& s = h->sim;
    HIP_TRY(hipStreamSynchronize(s.stream));
    const int prof = s.profiling;
    s.profiling = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    apic_tuning best = s.tuning;
    apic_status st = APIC_OK;
    auto time_ms = [&](auto&& work, float& ms) -> apic_status {
        work();
        HIP_TRY(hipEventRecord(e0, s.stream));
        for (uint32_t i = 0; i < reps; ++i) work();
        HIP_TRY(hipEventRecord(e1, s.stream));
        CHECK_LAUNCH();
        HIP_TRY(hipEventSynchronize(e1));
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        ms /= (float)reps;
        return APIC_OK;
    };
    // the sort of the keys of the current positions; the particles are not
    // permuted, and the next sort makes its keys again
    if (s.num_particles > 1) {
        float best_ms = 0.0f;
        for (uint32_t b = 8; b <= 11 && st == APIC_OK; ++b) {
            s.tuning.radix_bits = b;
            float ms = 0.0f;
            st = time_ms([&]() { launch_keygen(s); launch_radix_sort(s); }, ms);
            if (st == APIC_OK && (b == 8 || ms < best_ms)) { best_ms = ms; best.radix_bits = b; }
        }
        s.keys_valid = false;
        best.sort_ms = best_ms;
    }
    // apply_preconditioner on the current system
    if (st == APIC_OK && s.levels.size() > 1) {
        launch_pcg_reset(s);
        float best_ms = 0.0f;
        const uint32_t cand[] = {0u, 64u, 256u, 1024u, 4096u, 16384u};
        for (size_t i = 0; i < sizeof(cand) / sizeof(cand[0]) && st == APIC_OK; ++i) {
            s.tuning.few_tiles = cand[i];
            float ms = 0.0f;
            st = time_ms([&]() { launch_vcycle(s, 0); }, ms);
            if (st == APIC_OK && (i == 0 || ms < best_ms)) { best_ms = ms; best.few_tiles = cand[i]; }
        }
        best.vcycle_ms = best_ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s.profiling = prof;
    s.tuning = best;
    if (st != APIC_OK) return st;
    if (out) *out = best;
    return APIC_OK;
}

apic_status apic_apply_preconditioner(apic_handle* h) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    launch_pcg_reset(h->sim);
    launch_vcycle(h->sim, 0);
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    return APIC_OK;
}

apic_status apic_solve_pcg(apic_handle* h) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    apic_status st = do_pcg(h->sim, true);
    if (st != APIC_OK) return st;
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    return APIC_OK;
}

apic_status apic_apply_pressure_gradient(apic_handle* h) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    launch_gradient(h->sim, false, 0.0f);
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    return APIC_OK;
}

apic_status apic_extrapolate_velocity(apic_handle* h) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    launch_extrapolate(h->sim);
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(h->sim.stream));
    return APIC_OK;
}

apic_status apic_transfer_grid_to_particles_and_advect(apic_handle* h, float deltatime) {
    if (!h) return fail(APIC_ERR_INVALID_ARGUMENT, "null handle");
    Sim& s = h->sim;
    launch_g2p(s, deltatime, true);
    s.cells_valid = false;
    CHECK_LAUNCH();
    HIP_TRY(hipStreamSynchronize(s.stream));
    s.keys_valid = true;
    s.maxv_valid = true;
    return APIC_OK;
}

}  // extern "C"
