// pressure.hip -- pressure system, multigrid hierarchy and MGPCG solve.
//
//  smooth_density_kernel   the three density-smoothing passes of
//                          build_pressure_system (src/apic.rs:1649-1685)
//  assemble_kernel         ghost-fluid Laplacian + RHS (:1698-1784)
//  mg_build_kernel         one coarse level of update_multilevel_grid
//                          (:1499-1602), dense over coarse cells
//  V-cycle (apply_preconditioner, :1788-1821) with red-black SOR, omega=1.2,
//  one launch per level each way, on LDS tiles with a two-cell ring:
//    vdown_kernel          (base: e = 0 and the first red sweep) the black
//                          sweep of smooth_level (:471-535), then restrict
//                          (:537-594) into the next level fused with its
//                          first red sweep; on the base level in a PCG
//                          iteration it first applies the residual test
//    bottom_lds_kernel     one workgroup runs the coarse levels that fit in
//    (bottom_kernel)       LDS (else every level with <= 4096 cells, in global
//                          memory) and the coarsest-level CG (:618-704)
//    vup_kernel            prolongate (:596-616) fused with both colours of the
//                          reverse sweep; on the base level it also computes
//                          r.z (gamma), one partial per tile
//  PCG (solve_pcg, :1865-1956), reference rules:
//    spmv_dot_kernel       s = z + beta*s fused with A s and s.As
//    axpy_max_kernel       p += alpha s, r -= alpha As, max |r|
//  and the relative mode (FP64 scalars and sums, FP64 pressure accumulation,
//  periodic true residual: pcg_bnorm_kernel, true_residual_kernel).
//  Scalars live in a PcgState in device memory.  A reduction kernel writes
//  one partial per workgroup (plain stores, no device-wide fence or atomic);
//  the next kernel of the chain sums those partials itself, in the same fixed
//  order in every workgroup, applies the reference's exit test and skips its
//  work when the test ends the solve.  Every kernel of the loop returns at once
//  when `done` is set, so the host queues several iterations between reads.
//
// Every stencil keeps the reference's operation order and its fused
// multiply-adds (mul_add -> fmaf); the library is compiled with
// -ffp-contract=off.  Colour sweeps, restriction and prolongation are
// per-cell deterministic, so the V-cycle equals the CPU restatement bit for
// bit; only the PCG dot products are summed in a different order.
#include "apic_internal.hpp"
#include "device_math.hpp"

namespace apic {

namespace {

constexpr float kSafety = 1e-5f;  // SAFTY, src/apic.rs:8
constexpr float kOmega = 1.2f;
constexpr int kThreads = 256;
constexpr int kSolverThreads = kThreads;
constexpr int kBottomThreads = 1024;
constexpr int kSerialCap = 4096;
// The relative mode's coarse CG ends when r.r <= kCoarseRelStop2 r0.r0 (the
// level's residual reduced a hundredfold).
constexpr double kCoarseRelStop2 = 1e-4;

__device__ inline float atomic_load_f(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- assembly

// Per-workgroup (count, last raw index + 1) of the base fluid cells, reduced
// by fstats_finalize_kernel: the padded-list quirk of prolongate needs both.
__device__ inline void fluid_partial(bool fl, uint32_t c, uint32_t& cnt, uint32_t& last) {
    if (fl) { cnt += 1; last = max(last, c + 1); }
}

__device__ inline void write_fluid_partial(uint32_t cnt, uint32_t last, uint32_t* part, uint32_t wg) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o, 64);
        last = max(last, (uint32_t)__shfl_xor(last, o, 64));
    }
    __shared__ uint32_t s_cnt[1024 / 64], s_last[1024 / 64];  // (up to 1024 threads)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { s_cnt[wave] = cnt; s_last[wave] = last; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0, l = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { c += s_cnt[w]; l = max(l, s_last[w]); }
        part[2 * wg] = c;
        part[2 * wg + 1] = l;
    }
}

__device__ __forceinline__ void assemble(
    const float* __restrict__ mac_u, const float* __restrict__ mac_v, const float* __restrict__ smoothed,
    const uint8_t* __restrict__ fluid, const uint8_t* __restrict__ obstacle, uint32_t W, uint32_t cells, float cell,
    const LevelView& L, float* __restrict__ pressure, int zero_pressure, uint32_t* __restrict__ part_fluid, uint32_t wg,
    uint32_t nwg) {
    uint32_t cnt = 0, last = 0;
    for (uint32_t c = wg * blockDim.x + threadIdx.x; c < cells; c += nwg * blockDim.x) {
        size_t pc = (size_t)c + W;
        bool fl = fluid[c];
        fluid_partial(fl, c, cnt, last);
        if (fl) {
            uint32_t gx = c % W, gy = c / W;
            size_t ui = (size_t)gy * (W + 1) + gx;
            float ul = mac_u[ui], ur = mac_u[ui + 1], vb = mac_v[c], vt = mac_v[c + W];
            int nol = 1 - obstacle[c - 1], nor = 1 - obstacle[c + 1], nob = 1 - obstacle[c - W], not_ = 1 - obstacle[c + W];
            ul *= (float)nol; ur *= (float)nor; vb *= (float)nob; vt *= (float)not_;
            float div = (ur - ul) + (vt - vb);
            float rc = smoothed[c];
            const float target = 9.0f, noise = target * 1.15f, rate = 0.1f;
            float excess = fminf(fmaxf(rc - noise, 0.0f), target);
            int fl_ = fluid[c - 1], fr = fluid[c + 1], fb = fluid[c - W], ft = fluid[c + W];
            float diag = 0.0f, px = 0.0f, py = 0.0f;
            if (nol == 1) diag += fl_ ? 1.0f : 1.0f / theta(smoothed[c - 1], rc);
            if (nor == 1) {
                if (fr) { diag += 1.0f; px = -1.0f; }
                else diag += 1.0f / theta(smoothed[c + 1], rc);
            }
            if (nob == 1) diag += fb ? 1.0f : 1.0f / theta(smoothed[c - W], rc);
            if (not_ == 1) {
                if (ft) { diag += 1.0f; py = -1.0f; }
                else diag += 1.0f / theta(smoothed[c + W], rc);
            }
            L.r[pc] = (-div) * cell + ((excess * rate) * cell) * cell;
            L.px[pc] = px;
            L.py[pc] = py;
            L.diag[pc] = diag;
            L.inv_diag[pc] = diag > 1e-6f ? 1.0f / diag : 0.0f;
            if (zero_pressure) pressure[c] = 0.0f;
        } else {
            L.r[pc] = 0.0f;
            L.px[pc] = 0.0f;
            L.py[pc] = 0.0f;
            L.diag[pc] = 0.0f;
            L.inv_diag[pc] = 0.0f;
        }
    }
    write_fluid_partial(cnt, last, part_fluid, wg);
}

// 1024 threads a workgroup: a cell's loads depend on its fluid flag, so the
// loop needs many waves in flight (the partials of the fluid statistics are
// sums and maxima, the same for any partition of the cells)
struct Assemble {
    static constexpr int kThreads = 1024;
    struct Args {
        const float* mac_u;
        const float* mac_v;
        const float* smoothed;
        const uint8_t* fluid;
        const uint8_t* obstacle;
        uint32_t W, cells;
        float cell;
        LevelView L;
        float* pressure;
        int zero_pressure;
        uint32_t* part_fluid;
        uint32_t nwg;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        assemble(a.mac_u, a.mac_v, a.smoothed, a.fluid, a.obstacle, a.W, a.cells, a.cell, a.L, a.pressure,
                 a.zero_pressure, a.part_fluid, wg, a.nwg);
    }
};

__global__ __launch_bounds__(kThreads) void fluid_stats_kernel(const uint8_t* __restrict__ fluid, uint32_t cells,
                                                                uint32_t* __restrict__ part_fluid) {
    uint32_t cnt = 0, last = 0;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x)
        fluid_partial(fluid[c] != 0, c, cnt, last);
    write_fluid_partial(cnt, last, part_fluid, blockIdx.x);
}

struct FstatsArgs {
    const uint32_t* part_fluid;
    uint32_t n;
    FluidStats* fstats;
    FluidStats* host;
};
// The fluid statistics from the assembly's per-workgroup (count, last + 1):
// every thread of the workgroup takes part (integer sums and maxima, so the
// order does not matter).
__device__ __forceinline__ void fstats_finalize(const FstatsArgs& a) {
    __shared__ uint32_t s_cnt[1024 / 64], s_last[1024 / 64];
    uint32_t cnt = 0, last = 0;
    for (uint32_t b = threadIdx.x; b < a.n; b += blockDim.x) {
        cnt += a.part_fluid[2 * b];
        last = max(last, a.part_fluid[2 * b + 1]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o, 64);
        last = max(last, (uint32_t)__shfl_xor(last, o, 64));
    }
    if ((threadIdx.x & 63) == 0) { s_cnt[threadIdx.x >> 6] = cnt; s_last[threadIdx.x >> 6] = last; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0, l = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { c += s_cnt[w]; l = max(l, s_last[w]); }
        a.fstats->count = c;
        a.fstats->last_raw = l;
        if (a.host) { a.host->count = c; a.host->last_raw = l; }
    }
}
struct FstatsFinalize {
    static constexpr int kThreads = kSolverThreads;
    using Args = FstatsArgs;
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) { fstats_finalize(a); }
};

__device__ inline bool mg_build_cell(const LevelView& F, const LevelView& C, uint32_t c) {
    uint32_t cx = c % C.w, cy = c / C.w;
    size_t pcc = (size_t)c + C.w;
    bool any = false;
    float diag = 0.0f, plus_x = 0.0f, plus_y = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            uint32_t fx = cx * 2 + dx, fy = cy * 2 + dy;
            if (fx >= F.w || fy >= F.h) continue;
            uint32_t fr = fy * F.w + fx;
            size_t fp = (size_t)fr + F.w;
            if (F.fluid[fr]) {
                any = true;
                diag += F.diag[fp];
                if (dx == 0) diag += 2.0f * F.px[fp];
                if (dy == 0) diag += 2.0f * F.py[fp];
                if (dx == 1) plus_x += F.px[fp];
                if (dy == 1) plus_y += F.py[fp];
            }
        }
    C.fluid[c] = any;
    if (any) {
        float fd = diag * 0.5f;
        C.diag[pcc] = fd;
        C.inv_diag[pcc] = fd > 1e-6f ? 1.0f / fd : 0.0f;
        C.px[pcc] = plus_x * 0.5f;
        C.py[pcc] = plus_y * 0.5f;
    } else {
        C.diag[pcc] = 0.0f; C.inv_diag[pcc] = 0.0f; C.px[pcc] = 0.0f; C.py[pcc] = 0.0f;
    }
    return any;
}

// tile_flag (level 1 only, else null): a fluid coarse cell marks the base
// tile above it (a kBaseTile x kBaseTile tile holds kBaseTile / 2 coarse
// cells a side); TileOrder reads the marks and clears them.
struct MgBuild {
    static constexpr int kThreads = kSolverThreads;
    struct Args {
        LevelView F, C;
        uint8_t* tile_flag;
        uint32_t tiles_x;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        const uint32_t c = wg * kThreads + threadIdx.x;
        if (c >= a.C.cells) return;
        const bool fl = mg_build_cell(a.F, a.C, c);
        if (fl && a.tile_flag) a.tile_flag[(c / a.C.w) / 16 * a.tiles_x + (c % a.C.w) / 16] = 1;
    }
};

struct LevelSet {
    LevelView lv[16];
    int count;  // number of levels in lv (lv[0] = base)
};

// The base tiles' launch order for a simulation stepped alone.  A base-level
// launch has a workgroup per tile, and most tiles of a scene hold no fluid:
// such a tile only takes part in the solve's decisions and writes zero
// partials.  In launch order those tiles would hold their CU slots while
// fluid tiles behind them wait.  TileOrder keeps each tile on the workgroups
// of its XCD (launch index i runs on XCD (i + c) mod 8, and one_kernel deals
// tile rows to the XCDs) but gives the fluid tiles of each XCD its first
// workgroups, in their order; it zeroes the partials of the tiles without
// fluid (they stay zero for the whole solve), and order[l] = the tile that
// launch-logical workgroup l runs, bit 31 set for a tile without fluid.
// One wave per XCD residue class, up to kOrderMaxPer tiles a class.  The
// marks come from MgBuild (level 1) and are cleared here.
// Threads 0 .. 64 kXcds - 1 of a workgroup run it: MgBuildBottom's after
// the bottom levels, or TileOrder's own launch.
constexpr int kOrderMaxPer = 32;  // tiles per lane and class
constexpr uint32_t kOrderMaxTiles = kXcds * 64 * kOrderMaxPer;
constexpr int kOrderThreads = 64 * kXcds;
struct TileOrderArgs {
    uint8_t* flag;
    uint32_t* order;
    uint32_t n, tiles_x;
    uint32_t cmul;  // ceil(2^32 / tiles_x) for tiles_x > 1 (a quotient by tiles_x as a multiply), else 0
    double* part_rr_t;
    double* part_gamma_d;
    double* part_asz_d;
    float* part_gamma;
};
__device__ __forceinline__ void tile_order(const TileOrderArgs& a) {
    if (threadIdx.x >= (unsigned)kOrderThreads) return;
    const uint32_t x = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t n = a.n, c = a.tiles_x;
    const uint32_t per = x < n ? (n - x + kXcds - 1) / kXcds : 0u;  // launch indices x, x + 8, ... below n
    const uint32_t full = n / (kXcds * c) * (kXcds * c);
    // xcd_deal(hw, n, c), its quotient by c from the multiplier (exact:
    // the slots and c are below 2^16)
    auto deal = [&](uint32_t hw) -> uint32_t {
        if (hw < full) {
            const uint32_t xcd = hw % kXcds, slot = hw / kXcds;
            const uint32_t q = a.cmul ? __umulhi(slot, a.cmul) : slot;
            return (q * kXcds + xcd) * c + (slot - q * c);
        }
        return full + xcd_band(hw - full, n - full);
    };
    uint32_t tl[kOrderMaxPer];
    bool fl[kOrderMaxPer];
#pragma unroll
    for (int m = 0; m < kOrderMaxPer; ++m) {
        tl[m] = 0u;
        fl[m] = false;
        const uint32_t s = lane + 64u * m;
        if (64u * m < per && s < per) {
            tl[m] = deal(x + kXcds * s);
            fl[m] = a.flag[tl[m]] != 0;
        }
    }
    uint32_t nw = 0;
#pragma unroll
    for (int m = 0; m < kOrderMaxPer; ++m)
        if (64u * m < per) nw += (uint32_t)__popcll(__ballot(fl[m]));
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t rw = 0, rn = 0;
#pragma unroll
    for (int m = 0; m < kOrderMaxPer; ++m) {
        if (64u * m >= per) break;  // (uniform in the wave)
        const uint32_t s = lane + 64u * m;
        const uint64_t bw = __ballot(fl[m]), bv = __ballot(s < per);
        const uint64_t bn = bv & ~bw;
        if (s < per) {
            const uint32_t pos = fl[m] ? rw + (uint32_t)__popcll(bw & below)
                                       : nw + rn + (uint32_t)__popcll(bn & below);
            a.order[deal(x + kXcds * pos)] = tl[m] | (fl[m] ? 0u : 0x80000000u);
            a.flag[tl[m]] = 0;
            if (!fl[m]) {
                a.part_rr_t[tl[m]] = 0.0;
                a.part_gamma_d[tl[m]] = 0.0;
                a.part_asz_d[tl[m]] = 0.0;
                a.part_gamma[tl[m]] = 0.0f;
            }
        }
        rw += (uint32_t)__popcll(bw);
        rn += (uint32_t)__popcll(bn);
    }
}

struct TileOrder {
    static constexpr int kThreads = kOrderThreads;
    using Args = TileOrderArgs;
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) { tile_order(a); }
};

// with_fstats: FstatsFinalize's work (the assembly's partials are complete);
// with_order: TileOrder's; both after the levels, launches fewer
struct MgBuildBottom {
    static constexpr int kThreads = kBottomThreads;
    static_assert(kBottomThreads >= kOrderThreads && kBottomThreads <= 1024, "the tail's waves");
    struct Args {
        LevelSet ls;
        int first;
        int with_fstats;
        FstatsArgs fstats;
        int with_order;
        TileOrderArgs order;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) {
        for (int l = a.first; l < a.ls.count; ++l) {
            const LevelView F = a.ls.lv[l - 1];
            const LevelView C = a.ls.lv[l];
            for (uint32_t c = threadIdx.x; c < C.cells; c += blockDim.x) mg_build_cell(F, C, c);
            __syncthreads();
        }
        if (a.with_fstats) fstats_finalize(a.fstats);  // (every thread reaches its barrier)
        if (a.with_order) tile_order(a.order);
    }
};

// ------------------------------------------------------------------ V-cycle

struct Prolong {
    LevelView C;            // coarse level whose error is injected
    int has;                // 0: no prolongation (no coarser level)
    const FluidStats* fs;   // base level only: count and last index of its fluid list
};

// e + e_coarse(parent) for a fluid cell.  The base fluid list is padded to a
// multiple of 8 with copies of its last index (src/apic.rs:1487-1494) and
// prolongate walks the padded list (:599-615), so the last fluid cell of the
// base level receives the correction 1 + pad times, added one at a time.
__device__ inline float prolonged(const LevelView& L, const Prolong& P, uint32_t raw) {
    float v = L.e[(size_t)raw + L.w];
    if (!P.has || !L.fluid[raw]) return v;
    uint32_t x = raw % L.w, y = raw / L.w;
    float pc = P.C.e[(size_t)(y / 2) * P.C.w + x / 2 + P.C.w];
    int reps = 1;
    if (P.fs) {
        uint32_t last1 = P.fs->last_raw, cnt = P.fs->count;
        if (last1 != 0 && raw == last1 - 1 && (cnt & 7u) != 0) reps = 1 + (int)(8u - (cnt & 7u));
    }
    for (int k = 0; k < reps; ++k) v = v + pc;
    return v;
}

// neighbour value seen by an up-sweep of the first colour: a neighbour of the
// other colour has not been prolongated in memory yet
__device__ inline float nb_prolonged(const LevelView& L, const Prolong& P, int64_t raw) {
    if (raw < 0 || raw >= (int64_t)L.cells) return L.e[raw + L.w];
    return prolonged(L, P, (uint32_t)raw);
}

// mode 0: down sweep (plain), 1: up sweep, first colour (prolong own and
// neighbours), 2: up sweep, second colour (prolong own; neighbours final)
__device__ inline void sor_cell(const LevelView& L, const Prolong& P, uint32_t c, int mode) {
    if (!L.fluid[c]) return;
    const size_t w = L.w, pc = (size_t)c + w;
    float own = mode == 0 ? L.e[pc] : prolonged(L, P, c);
    float id = L.inv_diag[pc];
    if (id != 0.0f) {
        float er, et, el, eb;
        if (mode == 1) {
            er = nb_prolonged(L, P, (int64_t)c + 1);
            et = nb_prolonged(L, P, (int64_t)c + (int64_t)w);
            el = nb_prolonged(L, P, (int64_t)c - 1);
            eb = nb_prolonged(L, P, (int64_t)c - (int64_t)w);
        } else {
            er = L.e[pc + 1]; et = L.e[pc + w]; el = L.e[pc - 1]; eb = L.e[pc - w];
        }
        float sum = L.r[pc];
        sum = fmaf(-L.px[pc], er, sum);
        sum = fmaf(-L.py[pc], et, sum);
        sum = fmaf(-L.px[pc - 1], el, sum);
        sum = fmaf(-L.py[pc - w], eb, sum);
        float gs = sum * id;
        own = own + (kOmega * (gs - own));
    }
    L.e[pc] = own;
}

__device__ inline void red_zero_cell(const LevelView& L, uint32_t c) {
    size_t pc = (size_t)c + L.w;
    uint32_t x = c % L.w, y = c / L.w;
    float id = L.inv_diag[pc];
    bool red = ((x + y) & 1u) == 0;
    L.e[pc] = (L.fluid[c] && red && id != 0.0f) ? kOmega * (L.r[pc] * id) : 0.0f;
}

__device__ inline void restrict_cell(const LevelView& F, const LevelView& C, uint32_t c, bool fuse_red) {
    size_t pcc = (size_t)c + C.w;
    if (!C.fluid[c]) { C.r[pcc] = 0.0f; C.e[pcc] = 0.0f; return; }
    uint32_t cx = c % C.w, cy = c / C.w;
    const size_t fw = F.w;
    float sum = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            uint32_t fx = cx * 2 + dx, fy = cy * 2 + dy;
            if (fx >= F.w || fy >= F.h) continue;  // past the edge of an odd-sided level
            uint32_t rf = fy * F.w + fx;
            size_t pf = (size_t)rf + fw;
            if (F.fluid[rf]) {
                float b = F.r[pf];
                float ax = F.diag[pf] * F.e[pf];
                ax += F.px[pf] * F.e[pf + 1];
                ax += F.py[pf] * F.e[pf + fw];
                ax += F.px[pf - 1] * F.e[pf - 1];
                ax += F.py[pf - fw] * F.e[pf - fw];
                sum += b - ax;
            }
        }
    C.r[pcc] = sum;
    float id = C.inv_diag[pcc];
    bool red = ((cx + cy) & 1u) == 0;
    C.e[pcc] = (fuse_red && red && id != 0.0f) ? kOmega * (sum * id) : 0.0f;
}

__device__ inline float block_max(float v, float* scratch);
template <typename T>
__device__ inline T block_sum(T v, T* scratch);

// ---- block-level reductions (fixed order, deterministic)

// Whether pred holds on any thread of the workgroup: a ballot per wave, one
// LDS slot per wave and one barrier (which also orders the LDS stores made
// before it, as __syncthreads does).  The caller owns the slots (NT / 64
// words of LDS): a wave that has returned may write them again only after a
// barrier that every wave reaches after its read, so each call site declares
// its own, and a call site reached more than once needs such a barrier
// between the reads and the next call.
template <int NT>
__device__ __forceinline__ bool block_any(bool pred, uint32_t* slots) {
    const bool wave_any = __ballot(pred) != 0ull;
    if ((threadIdx.x & 63) == 0) slots[threadIdx.x >> 6] = wave_any;
    __syncthreads();
    uint32_t any = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) any |= slots[w];
    return any != 0;
}

// The three smoothing passes of build_pressure_system (src/apic.rs:1653-1685)
// in one launch.  A workgroup loads a kSmT x kSmT tile of the density and the
// fluid flags with a kSmR-cell ring into LDS, runs the first pass over the
// tile and a two-cell ring, the second over the tile and a one-cell ring, the
// third over the tile, and writes the tile: each pass reads only values the
// pass before computed inside the workgroup.  Per cell it is the reference's
// arithmetic: a fluid cell becomes (4 rho + the fluid neighbours' rho) / (4 +
// their count), every other cell keeps its value.  Cells past the grid enter
// as non-fluid zeros (no fluid cell borders the grid's edge: the edge cells
// are walls).
constexpr int kSmT = 32;
constexpr int kSmR = 3;
constexpr int kSmN = kSmT + 2 * kSmR;
struct Smooth3 {
    static constexpr int kThreads = 256;
    static constexpr bool kDealRows = true;  // its tile rows are dealt to the XCDs in turn
    struct Args {
        const float* in;   // density
        float* out;        // smoothed density
        const uint8_t* fluid;
        uint32_t W, H, tiles_x;
    };
    static __device__ __forceinline__ float cell(const float* src, const uint8_t* f, int i) {
        const float own = src[i];
        if (!f[i]) return own;
        float sum = own * 4.0f;
        float weight = 4.0f;
        const float fl = (float)f[i - 1], fr = (float)f[i + 1], fb = (float)f[i - kSmN], ft = (float)f[i + kSmN];
        sum += src[i - 1] * fl;
        sum += src[i + 1] * fr;
        sum += src[i - kSmN] * fb;
        sum += src[i + kSmN] * ft;
        weight += ((fl + fr) + fb) + ft;
        return sum / weight;
    }
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        __shared__ float s_a[kSmN * kSmN], s_b[kSmN * kSmN];
        __shared__ uint8_t s_f[kSmN * kSmN];
        __shared__ uint32_t any_slots[kThreads / 64];
        const int32_t W = (int32_t)a.W, H = (int32_t)a.H;
        const int32_t X0 = (int32_t)(wg % a.tiles_x) * kSmT - kSmR, Y0 = (int32_t)(wg / a.tiles_x) * kSmT - kSmR;
        constexpr int kPer = (kSmN * kSmN + kThreads - 1) / kThreads;
        float v[kPer];
        uint8_t fv[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = (int)threadIdx.x + kThreads * q;
            const int32_t x = X0 + i % kSmN, y = Y0 + i / kSmN;
            const bool in = i < kSmN * kSmN && x >= 0 && x < W && y >= 0 && y < H;
            const size_t c = in ? (size_t)y * W + x : 0;
            v[q] = in ? a.in[c] : 0.0f;
            fv[q] = in ? a.fluid[c] : 0;
        }
        bool any = false;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = (int)threadIdx.x + kThreads * q;
            if (i < kSmN * kSmN) {
                s_a[i] = v[q];
                s_f[i] = fv[q];
                any = any || fv[q] != 0;
            }
        }
        const bool work = block_any<kThreads>(any, any_slots);  // (its barrier orders the stores)
        if (work) {
            for (int i = threadIdx.x; i < kSmN * kSmN; i += kThreads) {
                const int lx = i % kSmN, ly = i / kSmN;
                if (lx >= 1 && lx < kSmN - 1 && ly >= 1 && ly < kSmN - 1) s_b[i] = cell(s_a, s_f, i);
            }
            __syncthreads();
            for (int i = threadIdx.x; i < kSmN * kSmN; i += kThreads) {
                const int lx = i % kSmN, ly = i / kSmN;
                if (lx >= 2 && lx < kSmN - 2 && ly >= 2 && ly < kSmN - 2) s_a[i] = cell(s_b, s_f, i);
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < kSmT * kSmT; i += kThreads) {
            const int tx = i % kSmT, ty = i / kSmT;
            const int32_t x = X0 + kSmR + tx, y = Y0 + kSmR + ty;
            if (x >= W || y >= H) continue;
            const int li = (ty + kSmR) * kSmN + tx + kSmR;
            a.out[(size_t)y * W + x] = work ? cell(s_a, s_f, li) : s_a[li];
        }
    }
};

// The xor butterfly over a wave, 32, 16, 8, 4, 2, 1 in that order.  The last
// four steps stay inside a 16-lane row and take their partner from a DPP row
// rotation: rotating by 8 is xor 8; after the steps before it, lanes whose
// indices differ by 8 (then 4, then 2) hold equal values, so rotating by 4,
// 2 and 1 delivers exactly the value xor 4, 2 and 1 would.  Every lane adds
// the same operands in the same order as with shuffles alone.
template <int N>
__device__ __forceinline__ int row_ror(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x120 + N, 0xF, 0xF, false);
}
template <int N>
__device__ __forceinline__ float row_ror(float v) {
    return __int_as_float(row_ror<N>(__float_as_int(v)));
}
template <int N>
__device__ __forceinline__ double row_ror(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = row_ror<N>((int)(b & 0xffffffffll)), hi = row_ror<N>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
// The partners of the first two steps (lane xor 32, lane xor 16) come from
// gfx950's permlane swaps, VALU moves instead of two LDS permutes.
// permlane32_swap(x, x) returns {x[l] below lane 32 and x[l - 32] above,
// x[l + 32] below and x[l] above}; permlane16_swap(x, x) the same for the
// rows pairs (0, 1) and (2, 3).  A lane takes the half that is not its own.
template <int X>
__device__ __forceinline__ uint32_t lane_xor_word(uint32_t x) {
    static_assert(X == 32 || X == 16, "permlane swaps exchange halves or row pairs");
    const bool upper = (__lane_id() & X) != 0;
    if constexpr (X == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        return upper ? r[0] : r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
        return upper ? r[0] : r[1];
    }
}
template <int X>
__device__ __forceinline__ float lane_xor(float v) {
    return __uint_as_float(lane_xor_word<X>(__float_as_uint(v)));
}
template <int X>
__device__ __forceinline__ int lane_xor(int v) {
    return (int)lane_xor_word<X>((uint32_t)v);
}
template <int X>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
    return lane_xor_word<X>(v);
}
template <int X>
__device__ __forceinline__ double lane_xor(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = lane_xor_word<X>((uint32_t)b), hi = lane_xor_word<X>((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T, class Op>
__device__ __forceinline__ T wave_butterfly(T v, Op op) {
    v = op(v, lane_xor<32>(v));
    v = op(v, lane_xor<16>(v));
    v = op(v, row_ror<8>(v));
    v = op(v, row_ror<4>(v));
    v = op(v, row_ror<2>(v));
    v = op(v, row_ror<1>(v));
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    return wave_butterfly(v, [](T a, T b) { return a + b; });
}

template <typename T>
__device__ inline T block_sum(T v, T* scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    T s = scratch[0];
    for (int w = 1; w < nw; ++w) s += scratch[w];
    return s;
}

// block_sum of two values at once: each gets exactly block_sum's operations
// (the same shuffle tree and wave order), the two chains overlapping.
// scratch holds 2 * (blockDim.x / 64) values.
template <typename T>
__device__ inline void block_sum2(T& a, T& b, T* scratch) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) { scratch[wave] = a; scratch[nw + wave] = b; }
    __syncthreads();
    T sa = scratch[0], sb = scratch[nw];
    for (int w = 1; w < nw; ++w) { sa += scratch[w]; sb += scratch[nw + w]; }
    a = sa;
    b = sb;
}

__device__ inline float block_max(float v, float* scratch) {
    v = wave_butterfly(v, [](float a, float b) { return fast_max(a, b); });
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    float s = scratch[0];
    for (int w = 1; w < nw; ++w) s = fast_max(s, scratch[w]);
    return s;
}


// ---- the one-workgroup bottom of the V-cycle

__device__ void coarse_cg_block(const LevelView& L, float* cg_s, float* cg_ap, float* cg_r, float* sprod,
                                float* scratch) {
    __shared__ float s_bcast[2];
    __shared__ uint32_t s_nf;
    const uint32_t n = L.cells;
    const size_t w = L.w;
    for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
        size_t pc = c + w;
        L.e[pc] = 0.0f;
        cg_r[pc] = L.r[pc];
        cg_s[pc] = L.r[pc];
    }
    if (threadIdx.x == 0) s_nf = 0;
    __syncthreads();
    // products go to LDS when they fit, else to cg_ap's padding-free twin
    float* prod = n <= (uint32_t)kSerialCap ? sprod : cg_ap;  // cg_ap is rewritten below when used
    const bool prod_in_lds = n <= (uint32_t)kSerialCap;
    uint32_t nf_local = 0;
    for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
        size_t pc = c + w;
        float rr = L.fluid[c] ? cg_r[pc] * cg_r[pc] : 0.0f;
        nf_local += L.fluid[c] ? 1u : 0u;
        if (prod_in_lds) prod[c] = rr;
    }
    atomicAdd(&s_nf, nf_local);
    __syncthreads();
    // Sequential sums in ascending cell order, as the reference's serial
    // loops.  Terms of non-fluid cells are replaced by +0, which leaves the
    // running sum unchanged (it starts at +0 and never becomes -0), so eight
    // terms at a time are loaded before they are added in order.
    int bc = 0;
    auto serial_sum = [&](auto term) -> float {
        if (threadIdx.x == 0) {
            float s = 0.0f;
            uint32_t c = 0;
            for (; c + 8 <= n; c += 8) {
                float t[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) t[k] = L.fluid[c + k] ? term(c + k) : 0.0f;
#pragma unroll
                for (int k = 0; k < 8; ++k) s += t[k];
            }
            for (; c < n; ++c) s += L.fluid[c] ? term(c) : 0.0f;
            s_bcast[bc] = s;
        }
        __syncthreads();
        float r = s_bcast[bc];
        bc ^= 1;  // the next sum writes the other slot, so no second barrier is needed
        return r;
    };
    float gamma;
    if (prod_in_lds) gamma = serial_sum([&](uint32_t c) { return prod[c]; });
    else gamma = serial_sum([&](uint32_t c) { size_t pc = c + w; return cg_r[pc] * cg_r[pc]; });
    if (gamma < 1e-12f) return;
    const uint32_t maxit = s_nf < 100u ? s_nf : 100u;
    for (uint32_t it = 0; it < maxit; ++it) {
        for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
            if (!L.fluid[c]) continue;
            size_t pc = c + w;
            float v = L.diag[pc] * cg_s[pc];
            v = fmaf(L.px[pc], cg_s[pc + 1], v);
            v = fmaf(L.px[pc - 1], cg_s[pc - 1], v);
            v = fmaf(L.py[pc], cg_s[pc + w], v);
            v = fmaf(L.py[pc - w], cg_s[pc - w], v);
            cg_ap[pc] = v;
            if (prod_in_lds) sprod[c] = cg_s[pc] * v;
        }
        __syncthreads();
        float delta;
        if (prod_in_lds) delta = serial_sum([&](uint32_t c) { return sprod[c]; });
        else delta = serial_sum([&](uint32_t c) { size_t pc = c + w; return cg_s[pc] * cg_ap[pc]; });
        if (fabsf(delta) < kSafety) break;
        float alpha = gamma / delta;
        float m = 0.0f;
        for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
            if (!L.fluid[c]) continue;
            size_t pc = c + w;
            L.e[pc] += alpha * cg_s[pc];
            float nr = cg_r[pc] - alpha * cg_ap[pc];
            cg_r[pc] = nr;
            m = fast_max(m, fabsf(nr));
            if (prod_in_lds) sprod[c] = nr * nr;
        }
        m = block_max(m, scratch);
        __syncthreads();
        if (m < 1e-6f) break;
        float ng;
        if (prod_in_lds) ng = serial_sum([&](uint32_t c) { return sprod[c]; });
        else ng = serial_sum([&](uint32_t c) { size_t pc = c + w; return cg_r[pc] * cg_r[pc]; });
        float beta = ng / gamma;
        gamma = ng;
        for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
            if (!L.fluid[c]) continue;
            size_t pc = c + w;
            cg_s[pc] = cg_r[pc] + (beta * cg_s[pc]);
        }
        __syncthreads();
    }
}

// Level views of the one-workgroup bottom in global memory (the LDS bottom
// copies levels [first, count) in, laid out level after level as diag,
// inv_diag, px, py, r, e over the padded range [0, cells + 2w), fluid flags
// after them, then the coarse CG vectors).
struct GlobalViews {
    const LevelSet* ls;
    __device__ LevelView view(int l) const { return ls->lv[l]; }
};

// solve_coarse_exact (src/apic.rs:618-704) for a level of at most 64 cells,
// run by one wave with one lane per cell: the vectors stay in registers,
// neighbours come from other lanes, and the serial sums read the lanes in
// ascending cell order, so the arithmetic is that of coarse_cg_block without a
// workgroup barrier per phase.  Only e of the level is written.
__device__ void coarse_cg_wave(const LevelView& L) {
    const uint32_t n = L.cells;
    const int w = (int)L.w;
    const int c = (int)(threadIdx.x & 63);
    const bool in = (uint32_t)c < n;
    const size_t pc = (size_t)c + (size_t)w;
    const bool fl = in && L.fluid[c];
    const float diag = in ? L.diag[pc] : 0.0f, px = in ? L.px[pc] : 0.0f, py = in ? L.py[pc] : 0.0f;
    const float pxm = in ? L.px[pc - 1] : 0.0f, pym = in ? L.py[pc - w] : 0.0f;
    float r = in ? L.r[pc] : 0.0f;
    float sv = r, e = 0.0f, ap = 0.0f;
    // the padded vectors are zero outside [0, n)
    auto nb = [&](int d) {
        const int j = c + d;
        const float v = __shfl(sv, j & 63, 64);
        return (j >= 0 && (uint32_t)j < n) ? v : 0.0f;
    };
    auto serial_sum = [&](float term) {
        const float t = fl ? term : 0.0f;  // +0 leaves the running sum unchanged
        float acc = 0.0f;
        for (uint32_t k = 0; k < n; ++k) acc += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t), (int)k));
        return acc;
    };
    uint32_t nf = 0;
    for (uint32_t k = 0; k < n; ++k) nf += (uint32_t)__builtin_amdgcn_readlane((int)fl, (int)k);
    float gamma = serial_sum(r * r);
    if (!(gamma < 1e-12f)) {
        const uint32_t maxit = nf < 100u ? nf : 100u;
        for (uint32_t it = 0; it < maxit; ++it) {
            const float sr = nb(1), sl = nb(-1), st = nb(w), sb = nb(-w);
            if (fl) {
                float v = diag * sv;
                v = fmaf(px, sr, v);
                v = fmaf(pxm, sl, v);
                v = fmaf(py, st, v);
                v = fmaf(pym, sb, v);
                ap = v;
            }
            const float delta = serial_sum(sv * ap);
            if (fabsf(delta) < kSafety) break;
            const float alpha = gamma / delta;
            float m = 0.0f;
            if (fl) {
                e += alpha * sv;
                r = r - alpha * ap;
                m = fabsf(r);
            }
            m = wave_butterfly(m, [](float a, float b) { return fast_max(a, b); });
            if (m < 1e-6f) break;
            const float ng = serial_sum(r * r);
            const float beta = ng / gamma;
            gamma = ng;
            if (fl) sv = r + (beta * sv);
        }
    }
    if (in) L.e[pc] = e;
}

// The relative mode's coarse CG on the one-workgroup bottom:
// solve_coarse_exact's iteration (src/apic.rs:618-704: the same FP32 vector
// arithmetic and fused SpMV chain, at most min(fluid cells, 100) iterations)
// with FP64 dot products, a search vector that is zero off the fluid cells
// (its matrix couplings are zero: the reference's copies of r there are only
// ever multiplied by them), and exits relative to the level's right-hand
// side: r.r = 0 at the start, s.As <= 0, ||r|| <= 1e-2 ||r0||.  (With the
// reference's absolute exits the coarse correction vanishes once the
// residual is small, |s.As| < 1e-5 at the first step, and the outer CG slows
// down.)  The FP64 sums are those of one wave (<= 64 cells, one lane per
// cell) or of the workgroup.
__device__ void coarse_cg_rel_wave(const LevelView& L) {
    const uint32_t n = L.cells;
    const int w = (int)L.w;
    const int c = (int)(threadIdx.x & 63);
    const bool in = (uint32_t)c < n;
    const size_t pc = (size_t)c + (size_t)w;
    const bool fl = in && L.fluid[c];
    const float diag = fl ? L.diag[pc] : 0.0f, px = fl ? L.px[pc] : 0.0f, py = fl ? L.py[pc] : 0.0f;
    const float pxm = fl ? L.px[pc - 1] : 0.0f, pym = fl ? L.py[pc - w] : 0.0f;
    float r = fl ? L.r[pc] : 0.0f;
    float sv = 0.0f, sn = 0.0f, e = 0.0f;
    auto nb = [&](int d) {
        const int j = c + d;
        const float v = __shfl(sn, j & 63, 64);
        return (j >= 0 && (uint32_t)j < n) ? v : 0.0f;
    };
    double gamma = wave_sum(fl ? (double)r * (double)r : 0.0);
    const double nf = wave_sum(fl ? 1.0 : 0.0);
    if (gamma > 0.0) {
        const double stop = kCoarseRelStop2 * gamma;
        const int maxit = (int)fmin(nf, 100.0);
        float beta = 0.0f;
        for (int it = 0; it < maxit; ++it) {
            sn = fl ? (it == 0 ? r : r + (beta * sv)) : 0.0f;
            const float sr = nb(1), sl = nb(-1), st = nb(w), sb = nb(-w);
            float ap = 0.0f;
            if (fl) {
                ap = diag * sn;
                ap = fmaf(px, sr, ap);
                ap = fmaf(pxm, sl, ap);
                ap = fmaf(py, st, ap);
                ap = fmaf(pym, sb, ap);
            }
            const double delta = wave_sum(fl ? (double)sn * (double)ap : 0.0);
            if (!(delta > 0.0)) break;
            const float alpha = (float)(gamma / delta);
            if (fl) {
                e += alpha * sn;
                r = r - alpha * ap;
            }
            sv = sn;
            const double ng = wave_sum(fl ? (double)r * (double)r : 0.0);
            if (ng <= stop) break;
            beta = (float)(ng / gamma);
            gamma = ng;
        }
    }
    if (in) L.e[pc] = e;
}

__device__ void coarse_cg_rel_block(const LevelView& L, float* cg_s, float* cg_ap, float* cg_r) {
    __shared__ double dscratch[2 * (kBottomThreads / 64)];
    const uint32_t n = L.cells;
    const size_t w = L.w;
    // the search vector: zero on the padded range, then on the fluid cells only
    for (size_t i = threadIdx.x; i < (size_t)n + 2 * w; i += blockDim.x) cg_s[i] = 0.0f;
    double g = 0.0, nf = 0.0;
    for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
        const size_t pc = c + w;
        const bool f = L.fluid[c];
        const float rv = f ? L.r[pc] : 0.0f;
        L.e[pc] = 0.0f;
        cg_r[pc] = rv;
        if (f) {
            g += (double)rv * (double)rv;
            nf += 1.0;
        }
    }
    block_sum2(g, nf, dscratch);  // (its barriers also order the stores above)
    if (!(g > 0.0)) return;
    double gamma = g;
    const double stop = kCoarseRelStop2 * gamma;
    const int maxit = (int)fmin(nf, 100.0);
    float beta = 0.0f;
    for (int it = 0; it < maxit; ++it) {
        for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
            if (!L.fluid[c]) continue;
            const size_t pc = c + w;
            cg_s[pc] = it == 0 ? cg_r[pc] : cg_r[pc] + (beta * cg_s[pc]);
        }
        __syncthreads();
        double dp = 0.0;
        for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
            if (!L.fluid[c]) continue;
            const size_t pc = c + w;
            float v = L.diag[pc] * cg_s[pc];
            v = fmaf(L.px[pc], cg_s[pc + 1], v);
            v = fmaf(L.px[pc - 1], cg_s[pc - 1], v);
            v = fmaf(L.py[pc], cg_s[pc + w], v);
            v = fmaf(L.py[pc - w], cg_s[pc - w], v);
            cg_ap[pc] = v;
            dp += (double)cg_s[pc] * (double)v;
        }
        const double delta = block_sum(dp, dscratch);
        if (!(delta > 0.0)) break;
        const float alpha = (float)(gamma / delta);
        double ng = 0.0;
        for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
            if (!L.fluid[c]) continue;
            const size_t pc = c + w;
            L.e[pc] += alpha * cg_s[pc];
            const float nr = cg_r[pc] - alpha * cg_ap[pc];
            cg_r[pc] = nr;
            ng += (double)nr * (double)nr;
        }
        ng = block_sum(ng, dscratch);
        if (ng <= stop) break;
        beta = (float)(ng / gamma);
        gamma = ng;
    }
}

// Levels [first, count) of the V-cycle inside one workgroup.  On entry
// level `first` holds r (restricted) and e after its fused first red sweep
// (or e = 0 when it is the coarsest level).  On exit level `first` holds its
// final post-smoothed error.
template <class Views>
__device__ void bottom_body(const Views& V, int first, int last, float* cg_s, float* cg_ap, float* cg_r,
                            float* sprod, float* scratch, bool rel) {
    Prolong none{};
    // down
    for (int l = first; l < last; ++l) {
        const LevelView L = V.view(l);
        if (l > first) {
            const LevelView F = V.view(l - 1);
            for (uint32_t c = threadIdx.x; c < L.cells; c += blockDim.x) restrict_cell(F, L, c, true);
            __syncthreads();
        }
        for (uint32_t c = threadIdx.x; c < L.cells; c += blockDim.x) {
            uint32_t x = c % L.w, y = c / L.w;
            if (((x + y) & 1u) == 1) sor_cell(L, none, c, 0);
        }
        __syncthreads();
    }
    const LevelView C = V.view(last);
    if (last > first) {
        const LevelView F = V.view(last - 1);
        for (uint32_t c = threadIdx.x; c < C.cells; c += blockDim.x) restrict_cell(F, C, c, false);
        __syncthreads();
    }
    if (C.cells <= 64) {
        if (threadIdx.x < 64) {
            if (rel) coarse_cg_rel_wave(C);
            else coarse_cg_wave(C);
        }
    } else if (rel) {
        coarse_cg_rel_block(C, cg_s, cg_ap, cg_r);
    } else {
        coarse_cg_block(C, cg_s, cg_ap, cg_r, sprod, scratch);
    }
    __syncthreads();
    // up
    for (int l = last - 1; l >= first; --l) {
        const LevelView L = V.view(l);
        Prolong P{};
        P.C = V.view(l + 1);
        P.has = 1;
        for (uint32_t c = threadIdx.x; c < L.cells; c += blockDim.x) {
            uint32_t x = c % L.w, y = c / L.w;
            if (((x + y) & 1u) == 1) sor_cell(L, P, c, 1);
        }
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < L.cells; c += blockDim.x) {
            uint32_t x = c % L.w, y = c / L.w;
            if (((x + y) & 1u) == 0) sor_cell(L, P, c, 2);
        }
        __syncthreads();
    }
}

// Bottom levels in global memory (used when they do not fit in LDS).
struct BottomGlobal {
    static constexpr int kThreads = kBottomThreads;
    struct Args {
        LevelSet ls;
        int first;
        float* cg_s;
        float* cg_ap;
        float* cg_r;
        const int32_t* done;
        int rel;  // the relative mode's coarse exits
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) {
        __shared__ float sprod[kSerialCap];
        __shared__ float scratch[kBottomThreads / 64];
        if (*a.done) return;
        GlobalViews V{&a.ls};
        bottom_body(V, a.first, a.ls.count - 1, a.cg_s, a.cg_ap, a.cg_r, sprod, scratch, a.rel != 0);
    }
};

// Bottom levels copied to LDS: the same arithmetic without a global-memory
// round trip between the dependent phases.  Every level of an LDS bottom has
// at most kBottomFirstCells <= kBottomThreads cells, so thread t owns cell t
// of each level: the sweeps and restrictions of sor_cell, restrict_cell and
// prolonged run with the cell's coordinates computed once per level and
// direct LDS indexing.  The errors and residuals of every bottom level are
// copied back (the finer level's prolongation reads them).
static_assert(kBottomFirstCells <= (uint32_t)kBottomThreads, "one cell per thread on an LDS bottom");

// Offsets of a bottom level's arrays in the LDS image (the host computes them
// from the level sizes; the bottom's levels lie in global memory in the same
// layout, from the first bottom level's diag and flags).
struct BLev {
    int32_t w, h, n;                // sides and cells
    uint32_t dg, id, px, py, r, e;  // float offsets of the padded arrays
    uint32_t fl;                    // offset of the fluid flags
    uint32_t wmul;                  // ceil(2^32 / w): y = umulhi(c, wmul) for w > 1
};

// The row of cell c of a bottom level: c / w by a multiply-high, exact for
// c * w < 2^32 (a bottom level has at most kBottomFirstCells cells).
__device__ __forceinline__ int bottom_row(const BLev& L, int c) {
    return L.w == 1 ? c : (int)__umulhi((uint32_t)c, L.wmul);
}

// Makes the compiler produce the given loaded values at this point: their
// loads are issued before the first wait and cannot be moved into the
// branches that use them (one LDS round trip for a phase instead of one per
// branch).
#define APIC_TIE4(a, b, c, d) asm volatile("" : "+v"(a), "+v"(b), "+v"(c), "+v"(d))
#define APIC_TIE2(a, b) asm volatile("" : "+v"(a), "+v"(b))
#define APIC_TIE1(a) asm volatile("" : "+v"(a))

// sor_cell's update from the cell's coefficients, own value and neighbour
// values: own when the cell has no inverse diagonal (the same value sor_cell
// leaves), computed without a branch.
__device__ __forceinline__ float bottom_sor(float own, float id, float r, float px, float py, float pxm, float pym,
                                            float er, float et, float el, float eb) {
    float sum = r;
    sum = fmaf(-px, er, sum);
    sum = fmaf(-py, et, sum);
    sum = fmaf(-pxm, el, sum);
    sum = fmaf(-pym, eb, sum);
    const float gs = sum * id;
    const float nv = own + (kOmega * (gs - own));
    return id != 0.0f ? nv : own;
}

struct BottomLdsArgs {
    BLev tab[16];          // by level index
    float* img;            // the levels' arrays in global memory (the LDS image's layout)
    const uint32_t* fimg;  // their flag bytes, as words
    uint32_t total;        // floats of the image
    uint32_t fwords;       // flag words of the image
    uint32_t cg;           // float offset of the coarse CG vectors (after the image)
    int first, last;       // the bottom's levels
    const int32_t* done;
    int rel;               // the relative mode's coarse exits
};

__device__ __forceinline__ void bottom_lds(const BottomLdsArgs& a) {
    __shared__ float f[kBottomLdsFloats];
    __shared__ __attribute__((aligned(16))) uint32_t fw[kBottomLdsFlags / 4];  // copied in as words
    __shared__ float scratch[kBottomThreads / 64];
    const uint8_t* fl = reinterpret_cast<const uint8_t*>(fw);
    const int32_t dn = *a.done;
    const int first = a.first, last = a.last;
    const int c = (int)threadIdx.x;
    // Copy-in: one stride of floats and one of flag words; a thread's loads
    // all go out before its stores.
    {
        const uint32_t total = a.total, fwords = a.fwords;
        const float* __restrict__ img = a.img;
        constexpr int kPer = 12;
        for (uint32_t i0 = threadIdx.x; i0 < total; i0 += kPer * kBottomThreads) {
            float v[kPer];
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const uint32_t i = i0 + j * kBottomThreads;
                v[j] = img[i < total ? i : 0];
            }
            uint32_t fv = 0;
            if (i0 == threadIdx.x && threadIdx.x < fwords) fv = a.fimg[threadIdx.x];
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const uint32_t i = i0 + j * kBottomThreads;
                if (i < total) f[i] = v[j];
            }
            if (i0 == threadIdx.x && threadIdx.x < fwords) fw[threadIdx.x] = fv;
        }
        for (uint32_t i = threadIdx.x + kBottomThreads; i < fwords; i += kBottomThreads) fw[i] = a.fimg[i];
    }
    if (dn) return;
    const BLev& CL = a.tab[last];
    const uint32_t pl = (uint32_t)(((uint32_t)CL.n + 2u * (uint32_t)CL.w + 7u) / 8u * 8u);
    float* cg_s = f + a.cg;  // the coarse CG vectors
    float* cg_ap = cg_s + pl;
    float* cg_r = cg_ap + pl;
    float* sprod = cg_r + pl;
    // the CG vectors' padding rows are read as neighbours: zero, as in global memory
    for (uint32_t i = threadIdx.x; i < 3 * pl + (uint32_t)CL.n; i += blockDim.x) cg_s[i] = 0.0f;
    __syncthreads();

    // restrict_cell: coarse cell c of C from level F, with C's red sweep
    // fused when asked.  Every child's operands load unconditionally (a
    // child past the edge of an odd-sided level loads the first child's and
    // adds nothing; a non-fluid child adds +0, which leaves the sum as
    // skipping it would: the sum starts at +0 and never becomes -0).
    auto restrict_to = [&](const BLev& F, const BLev& C, bool fuse_red) {
        if ((c & ~63) >= C.n) return;  // a wave without cells of C
        const bool mine = c < C.n;
        const int cc = mine ? c : 0;
        const int cy = bottom_row(C, cc), cx = cc - cy * C.w;
        const int pcc = cc + C.w;
        float b[4], dg[4], e0[4], pxa[4], e1[4], pya[4], ew[4], pxm[4], em[4], pym[4], es[4];
        uint32_t cfl[4];
        bool in[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int dx = k & 1, dy = k >> 1;
            in[k] = cx * 2 + dx < F.w && cy * 2 + dy < F.h;
            const int rf = in[k] ? (cy * 2 + dy) * F.w + cx * 2 + dx : cy * 2 * F.w + cx * 2, pf = rf + F.w;
            b[k] = f[F.r + pf];
            dg[k] = f[F.dg + pf];
            e0[k] = f[F.e + pf];
            pxa[k] = f[F.px + pf];
            e1[k] = f[F.e + pf + 1];
            pya[k] = f[F.py + pf];
            ew[k] = f[F.e + pf + F.w];
            pxm[k] = f[F.px + pf - 1];
            em[k] = f[F.e + pf - 1];
            pym[k] = f[F.py + pf - F.w];
            es[k] = f[F.e + pf - F.w];
            cfl[k] = fl[F.fl + rf];
        }
        float id = f[C.id + pcc];
        uint32_t own_fl = fl[C.fl + cc];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            APIC_TIE4(b[k], dg[k], e0[k], pxa[k]);
            APIC_TIE4(e1[k], pya[k], ew[k], pxm[k]);
            APIC_TIE4(em[k], pym[k], es[k], cfl[k]);
        }
        APIC_TIE2(id, own_fl);
        float sum = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float ax = dg[k] * e0[k];
            ax += pxa[k] * e1[k];
            ax += pya[k] * ew[k];
            ax += pxm[k] * em[k];
            ax += pym[k] * es[k];
            sum += (in[k] && cfl[k]) ? b[k] - ax : 0.0f;
        }
        if (mine) {
            const bool red = ((cx + cy) & 1) == 0;
            f[C.r + pcc] = own_fl ? sum : 0.0f;
            f[C.e + pcc] = (own_fl && fuse_red && red && id != 0.0f) ? kOmega * (sum * id) : 0.0f;
        }
    };

    // down: restriction into level l, then its black sweep (the red one was
    // fused into the restriction, or into the kernel above for the first)
    for (int l = first; l < last; ++l) {
        const BLev L = a.tab[l];
        if (l > first) {
            restrict_to(a.tab[l - 1], L, true);
            __syncthreads();
        }
        if ((c & ~63) < L.n) {  // (a wave without cells of the level skips the sweep)
            const bool mine = c < L.n;
            const int cc = mine ? c : 0;
            const int y = bottom_row(L, cc), x = cc - y * L.w;
            const int pc = cc + L.w;
            float own = f[L.e + pc], id = f[L.id + pc], r = f[L.r + pc], px = f[L.px + pc], py = f[L.py + pc];
            float pxm = f[L.px + pc - 1], pym = f[L.py + pc - L.w];
            float er = f[L.e + pc + 1], et = f[L.e + pc + L.w], el = f[L.e + pc - 1], eb = f[L.e + pc - L.w];
            uint32_t flg = fl[L.fl + cc];
            APIC_TIE4(own, id, r, px);
            APIC_TIE4(py, pxm, pym, er);
            APIC_TIE4(et, el, eb, flg);
            const float v = bottom_sor(own, id, r, px, py, pxm, pym, er, et, el, eb);
            if (mine && ((x + y) & 1) == 1 && flg) f[L.e + pc] = v;
        }
        __syncthreads();
    }
    if (last > first) {
        restrict_to(a.tab[last - 1], CL, false);
        __syncthreads();
    }
    LevelView C{};
    C.w = (uint32_t)CL.w; C.h = (uint32_t)CL.h; C.cells = (uint32_t)CL.n;
    C.diag = f + CL.dg; C.inv_diag = f + CL.id; C.px = f + CL.px; C.py = f + CL.py;
    C.r = f + CL.r; C.e = f + CL.e;
    C.fluid = const_cast<uint8_t*>(fl) + CL.fl;
    if (C.cells <= 64) {
        if (threadIdx.x < 64) {
            if (a.rel) coarse_cg_rel_wave(C);
            else coarse_cg_wave(C);
        }
    } else if (a.rel) {
        coarse_cg_rel_block(C, cg_s, cg_ap, cg_r);
    } else {
        coarse_cg_block(C, cg_s, cg_ap, cg_r, sprod, scratch);
    }
    __syncthreads();
    // up: the two colours of each level with the parent's error prolongated
    for (int l = last - 1; l >= first; --l) {
        const BLev L = a.tab[l], P = a.tab[l + 1];
        const bool act = (c & ~63) < L.n;  // a wave without cells of the level skips both colours
        const bool mine = c < L.n;
        const int cc = mine ? c : 0;
        const int y = bottom_row(L, cc), x = cc - y * L.w;
        const int pc = cc + L.w;
        // neighbour raw indices and coordinates (rows wrap as in the flat arrays)
        const int rr = cc + 1, rt = cc + L.w, rl = cc - 1, rb = cc - L.w;
        const int xr = x + 1 < L.w ? x + 1 : 0, yr = x + 1 < L.w ? y : y + 1;
        const int xl = x > 0 ? x - 1 : L.w - 1, yl = x > 0 ? y : y - 1;
        // e + the parent's error at raw index raw = (nx, ny) when it is a
        // fluid cell of the level; e alone outside [0, n) or for other cells.
        // The loads are unconditional (clamped), so that they go out together.
        const bool inr = rr < L.n, int_ = rt < L.n, inl = rl >= 0, inb = rb >= 0;
        float own = 0.0f, id = 0.0f, r = 0.0f, px = 0.0f, py = 0.0f, pxm = 0.0f, pym = 0.0f;
        bool work = false, black = false;
        if (act) {
            float vr = f[L.e + rr + L.w], vt = f[L.e + rt + L.w], vl = f[L.e + rl + L.w], vb = f[L.e + rb + L.w];
            uint32_t fr = fl[L.fl + (inr ? rr : cc)], ft = fl[L.fl + (int_ ? rt : cc)];
            uint32_t fl_l = fl[L.fl + (inl ? rl : cc)], fb = fl[L.fl + (inb ? rb : cc)];
            float per = f[P.e + ((inr ? yr : y) / 2) * P.w + (inr ? xr : x) / 2 + P.w];
            float pet = f[P.e + ((int_ ? y + 1 : y) / 2) * P.w + x / 2 + P.w];
            float pel = f[P.e + ((inl ? yl : y) / 2) * P.w + (inl ? xl : x) / 2 + P.w];
            float peb = f[P.e + ((inb ? y - 1 : y) / 2) * P.w + x / 2 + P.w];
            // the cell's own prolongated value and coefficients, before the first
            // colour (it leaves the second colour's cells as they are)
            float e_own = f[L.e + pc], pe_own = f[P.e + (y / 2) * P.w + x / 2 + P.w];
            id = f[L.id + pc]; r = f[L.r + pc]; px = f[L.px + pc]; py = f[L.py + pc];
            pxm = f[L.px + pc - 1]; pym = f[L.py + pc - L.w];
            uint32_t wfl = fl[L.fl + cc];
            APIC_TIE4(vr, vt, vl, vb);
            APIC_TIE4(fr, ft, fl_l, fb);
            APIC_TIE4(per, pet, pel, peb);
            APIC_TIE4(e_own, pe_own, id, r);
            APIC_TIE4(px, py, pxm, pym);
            APIC_TIE1(wfl);
            work = mine && wfl;
            own = e_own + pe_own;
            black = ((x + y) & 1) == 1;
            {
                const float er = (inr && fr) ? vr + per : vr;
                const float et = (int_ && ft) ? vt + pet : vt;
                const float el = (inl && fl_l) ? vl + pel : vl;
                const float eb = (inb && fb) ? vb + peb : vb;
                const float v = bottom_sor(own, id, r, px, py, pxm, pym, er, et, el, eb);
                if (work && black) f[L.e + pc] = v;
            }
        }
        __syncthreads();
        if (act) {
            float er = f[L.e + pc + 1], et = f[L.e + pc + L.w], el = f[L.e + pc - 1], eb = f[L.e + pc - L.w];
            APIC_TIE4(er, et, el, eb);
            const float v = bottom_sor(own, id, r, px, py, pxm, pym, er, et, el, eb);
            if (work && !black) f[L.e + pc] = v;
        }
        __syncthreads();
    }
    // copy-back: each level's r and e are one stride of the image
    for (int l = first; l <= last; ++l) {
        const BLev& L = a.tab[l];
        float* __restrict__ dst = a.img + L.r;  // e follows r
        const uint32_t n2 = 2 * (L.e - L.r);
        for (uint32_t i = threadIdx.x; i < n2; i += kBottomThreads) dst[i] = f[L.r + i];
    }
}
#undef APIC_TIE4
#undef APIC_TIE2
#undef APIC_TIE1

struct BottomLds {
    static constexpr int kThreads = kBottomThreads;
    using Args = BottomLdsArgs;
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) { bottom_lds(a); }
};

// ---------------------------------------------------------------------- PCG

struct PcgReset {
    static constexpr int kThreads = 64;
    struct Args {
        PcgState* st;
        double tol2;
        int max_iters;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) {
        if (threadIdx.x != 0) return;
        PcgState* st = a.st;
        st->done = 0;
        st->iters = 0;
        st->exit_reason = 0;
        st->r_in_alt = 0;
        st->maxerr = st->delta = st->alpha = 0.0f;
        st->bnorm2 = st->rnorm2 = 0.0;
        st->tol2 = a.tol2;
        st->gamma_d[0] = st->gamma_d[1] = 0.0;
        st->alpha_d = 0.0;
        st->max_iters = a.max_iters;
    }
};

// Relative mode, before the first V-cycle: b = r is kept for the residual
// recomputation, the residual's low part starts at zero, the FP64 pressure
// starts from the FP32 one, and the partials of ||b||^2 over the fluid cells.
__device__ __forceinline__ void pcg_bnorm(const LevelView& L, float* __restrict__ rhs, float* __restrict__ r_lo,
                                 double* __restrict__ part, const float* __restrict__ pressure,
                                 double* __restrict__ pd, uint32_t wg, uint32_t nwg) {
    __shared__ double dscratch[kThreads / 64];
    double acc = 0.0;
    for (uint32_t c = wg * blockDim.x + threadIdx.x; c < L.cells; c += nwg * blockDim.x) {
        pd[c] = (double)pressure[c];
        if (!L.fluid[c]) continue;
        const size_t pc = (size_t)c + L.w;
        const float b = L.r[pc];
        rhs[pc] = b;
        r_lo[pc] = 0.0f;
        acc += (double)b * (double)b;
    }
    acc = block_sum(acc, dscratch);
    if (threadIdx.x == 0) part[wg] = acc;
}

struct PcgBnorm {
    static constexpr int kThreads = kSolverThreads;
    struct Args {
        LevelView L;
        float* rhs;
        float* r_lo;
        double* part;
        const float* pressure;
        double* pd;
        uint32_t nwg;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        pcg_bnorm(a.L, a.rhs, a.r_lo, a.part, a.pressure, a.pd, wg, a.nwg);
    }
};

// ||b||^2 from PcgBnorm's partials; it also resets the state (PcgReset's
// work: no kernel before it reads the state), one launch instead of two.
struct PcgBnormFinalize {
    static constexpr int kThreads = kSolverThreads;
    struct Args {
        PcgReset::Args reset;
        const double* part;
        uint32_t n;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) {
        __shared__ double dscratch[kThreads / 64];
        double b = 0.0;
        for (uint32_t i = threadIdx.x; i < a.n; i += blockDim.x) b += a.part[i];
        b = block_sum(b, dscratch);
        PcgReset::run(a.reset, 0, 0);  // (thread 0 alone)
        if (threadIdx.x == 0) {
            PcgState* st = a.reset.st;
            st->bnorm2 = b;
            if (!(b > 0.0)) { st->done = 1; st->exit_reason = 1; }  // b = 0: p = 0 is the solution
        }
    }
};

// Relative mode, after iteration k's update, every recompute_every
// iterations (due): r is replaced by the true residual b - A p evaluated in
// FP64 against the FP64 pressure (the product with the same fused operations
// as apply_matrix_and_dot, zero outside the grid), split into its FP32
// rounding r and the low part r_lo.  part_out receives the partials of
// ||r||^2 that the residual test reads on those iterations (on the others it
// reads the update's per-tile partials).
__device__ __forceinline__ void true_residual(const LevelView& L, const float* __restrict__ rhs,
                                     float* __restrict__ r_lo, const double* __restrict__ pd, const PcgState* st,
                                     double* __restrict__ part_out, int due, uint32_t wg, uint32_t nwg) {
    __shared__ double dscratch[kThreads / 64];
    if (st->done || !due) return;
    const int64_t w = L.w, n = L.cells;
    auto pa = [&](int64_t raw) { return raw >= 0 && raw < n ? pd[raw] : 0.0; };
    double acc = 0.0;
    for (uint32_t c = wg * blockDim.x + threadIdx.x; c < L.cells; c += nwg * blockDim.x) {
        if (!L.fluid[c]) continue;
        const size_t pc = (size_t)c + L.w;
        const int64_t ci = c;
        double v = (double)L.diag[pc] * pd[c];
        v = fma((double)L.px[pc], pa(ci + 1), v);
        v = fma((double)L.px[pc - 1], pa(ci - 1), v);
        v = fma((double)L.py[pc], pa(ci + w), v);
        v = fma((double)L.py[pc - L.w], pa(ci - w), v);
        const double t = (double)rhs[pc] - v;
        const float hi = (float)t;
        L.r[pc] = hi;
        r_lo[pc] = (float)(t - (double)hi);
        acc += t * t;
    }
    acc = block_sum(acc, dscratch);
    if (threadIdx.x == 0) part_out[wg] = acc;
}

// Relative mode, after the update of iteration k (the uniform argument).
struct TrueResidual {
    static constexpr int kThreads = kSolverThreads;
    struct Args {
        LevelView L;
        const float* rhs;
        float* r_lo;
        const double* pd;
        const PcgState* st;
        double* part_out;
        int every;  // recompute every `every` iterations (0: never)
        uint32_t nwg;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int k) {
        const int due = a.every > 0 && (k + 1) % a.every == 0;
        true_residual(a.L, a.rhs, a.r_lo, a.pd, a.st, a.part_out, due, wg, a.nwg);
    }
};

__global__ void zero_pressure_kernel(const uint8_t* __restrict__ fluid, float* __restrict__ pressure,
                                     uint32_t cells) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cells && fluid[c]) pressure[c] = 0.0f;
}

// Relative mode: breakdown when r.z <= 0 (or not finite), the iteration cap.
__device__ inline bool gamma_exit_rel(PcgState* st, double g, int k, bool lead) {
    if (!(g > 0.0)) {
        if (lead) { st->done = 1; st->exit_reason = k == 0 ? 1 : 4; st->iters = k; }
        return true;
    }
    if (k >= st->max_iters) {
        if (lead) { st->done = 1; st->exit_reason = 5; st->iters = k; }
        return true;
    }
    return false;
}

// Exit tests on gamma = r.z before iteration k (src/apic.rs:1890, 1938) and
// the iteration cap.  Returns true when the solve ends here.
__device__ inline bool gamma_exit(PcgState* st, float g, int k, bool lead) {
    if (g < kSafety) {
        if (lead) { st->done = 1; st->exit_reason = k == 0 ? 1 : 4; st->iters = k; }
        return true;
    }
    if (k >= kMaxPcgIterations) {
        if (lead) { st->done = 1; st->exit_reason = 5; st->iters = k; }
        return true;
    }
    return false;
}

// SpmvDot and AxpyMax take their cells in batches: a thread first loads the
// fluid flags of a batch of its cells (with the reduction partials it needs),
// then every operand of the batch's fluid cells at once (a non-fluid cell
// reads the row's first cell, a line every lane of the wave shares), so that
// a workgroup waits twice per batch instead of twice per cell.  With FP64
// reductions on a grid of at least kWideMinRed reduction workgroups SpmvDot
// runs kWideThreads threads per workgroup (about two cells per thread): one
// such grid's launch is bound by latency, which four times the waves hide.
// Smaller grids, and batches of them (whose launches fill the GPU with the
// workgroups of all their simulations), keep 256 threads and more cells per
// wave, and so do the FP32 reductions of the reference's rules (AxpyMax,
// which keeps each partial's cells and summation order).  The choice depends
// on the simulation alone, so a batched simulation computes what it computes
// alone.  (The relative mode's update is axpy_tile, over the base tiles.)
constexpr int kWideThreads = 1024;
constexpr int kWideBatch = 2;
constexpr uint32_t kWideMinRed = 256;

// The n partials of one reduction summed in the order of a 256-thread
// workgroup (thread t < 256 adds part[t], part[t + 256], ... in turn; the
// four wave sums are added in wave order), whatever the size of the calling
// workgroup: every kernel that reduces the same partials obtains the same
// value.  Four loads per thread are in flight at a time.
template <typename T>
__device__ inline T sum_partials256(const T* __restrict__ part, int n, T* scratch) {
    T v = 0;
    if (threadIdx.x < 256) {
        for (int i0 = (int)threadIdx.x; i0 < n; i0 += 1024) {
            T x[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int i = i0 + 256 * j;
                x[j] = part[i < n ? i : 0];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (i0 + 256 * j < n) v += x[j];
        }
    }
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && threadIdx.x < 256) scratch[threadIdx.x >> 6] = v;
    __syncthreads();
    T s = scratch[0];
    for (int w = 1; w < 4; ++w) s += scratch[w];
    return s;
}

// sum_partials256 of two partial arrays of n values at once (each value is
// exactly sum_partials256's); scratch holds 8 values.
template <typename T>
__device__ inline void sum_partials256x2(const T* __restrict__ pa, const T* __restrict__ pb, int n, T* scratch,
                                         T& ra, T& rb) {
    T va = 0, vb = 0;
    if (threadIdx.x < 256) {
        for (int i0 = (int)threadIdx.x; i0 < n; i0 += 1024) {
            T xa[4], xb[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int i = i0 + 256 * j;
                xa[j] = pa[i < n ? i : 0];
                xb[j] = pb[i < n ? i : 0];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (i0 + 256 * j < n) { va += xa[j]; vb += xb[j]; }
        }
    }
    va = wave_sum(va);
    vb = wave_sum(vb);
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && threadIdx.x < 256) {
        scratch[threadIdx.x >> 6] = va;
        scratch[4 + (threadIdx.x >> 6)] = vb;
    }
    __syncthreads();
    T sa = scratch[0], sb = scratch[4];
    for (int w = 1; w < 4; ++w) { sa += scratch[w]; sb += scratch[4 + w]; }
    ra = sa;
    rb = sb;
}

// Iteration k: gamma from the V-cycle's partials, s = z + beta*s (s = z when
// k = 0), As = A s and the partials of delta = s.As (apply_matrix_and_dot,
// src/apic.rs:1823-1863; search update :1945-1954).
template <int NT>
__device__ __forceinline__ void spmv_dot(const LevelView& L, const float* __restrict__ s_old, float* __restrict__ s_new,
                                float* __restrict__ As, float* __restrict__ As_lo, PcgState* st,
                                const float* __restrict__ part_gamma,
                                float* __restrict__ part_delta, int k, const double* __restrict__ part_gamma_d,
                                const double* __restrict__ part_asz_d, double* __restrict__ part_delta_d, int ngam,
                                uint32_t wg, uint32_t nwg) {
    const bool lead = wg == 0 && threadIdx.x == 0;
    __shared__ float scratch[NT / 64];
    __shared__ double dscratch[NT / 64];
    if (st->done) return;
    const bool first = k == 0;
    const size_t w = L.w;
    const uint32_t cells = L.cells, stride = nwg * NT;
    uint32_t c0 = wg * NT + threadIdx.x;
    // the first batch's flags go out with the partials
    uint8_t fl[kWideBatch];
    auto fetch_flags = [&]() {
#pragma unroll
        for (int j = 0; j < kWideBatch; ++j) {
            const uint32_t c = c0 + j * stride;
            fl[j] = c < cells ? L.fluid[c < cells ? c : 0] : 0;
        }
    };
    fetch_flags();
    float beta;
    if (part_gamma_d) {  // relative mode: FP64 scalars, flexible beta
        // gamma and (A s_{k-1}).z_k in one pass (the second is unused when k = 0)
        __shared__ double dscratch2[8];
        double g, asz;
        sum_partials256x2(part_gamma_d, part_asz_d, ngam, dscratch2, g, asz);
        if (gamma_exit_rel(st, g, k, lead)) return;
        // Polak-Ribiere: r_{k}.(z_{k} - z_{k-1}) / r_{k-1}.z_{k-1}, where
        // r_{k-1}.z_k = r_k.z_k + alpha_{k-1} (A s_{k-1}).z_k
        if (!first) {
            beta = (float)(-st->alpha_d * asz / st->gamma_d[(k - 1) & 1]);
        } else {
            beta = 0.0f;
        }
        if (lead) { st->gamma_d[k & 1] = g; st->iters = k + 1; }
    } else {
        const float g = sum_partials256(part_gamma, ngam, scratch);
        if (gamma_exit(st, g, k, lead)) return;
        beta = first ? 0.0f : g / st->gamma[k - 1];
        if (lead) { st->gamma[k] = g; st->iters = k + 1; }
    }
    float prod = 0.0f;
    double dprod = 0.0;
    for (; c0 < cells; c0 += kWideBatch * stride) {
        // operands of the batch's cells: own, right, left, top, bottom
        float e[kWideBatch][5], so[kWideBatch][5], dg[kWideBatch], px[kWideBatch][2], py[kWideBatch][2];
#pragma unroll
        for (int j = 0; j < kWideBatch; ++j) {
            const size_t pc = fl[j] ? (size_t)(c0 + j * stride) + w : w;
            const size_t nb[5] = {pc, pc + 1, pc - 1, pc + w, pc - w};
#pragma unroll
            for (int d = 0; d < 5; ++d) {
                e[j][d] = L.e[nb[d]];
                so[j][d] = first ? 0.0f : s_old[nb[d]];
            }
            dg[j] = L.diag[pc];
            px[j][0] = L.px[pc]; px[j][1] = L.px[pc - 1];
            py[j][0] = L.py[pc]; py[j][1] = L.py[pc - w];
        }
        uint8_t fl_now[kWideBatch];
#pragma unroll
        for (int j = 0; j < kWideBatch; ++j) fl_now[j] = fl[j];
        // the next batch's flags, in flight during this batch's arithmetic
        c0 += kWideBatch * stride;
        fetch_flags();
        c0 -= kWideBatch * stride;
#pragma unroll
        for (int j = 0; j < kWideBatch; ++j) {
            if (!fl_now[j]) continue;
            const size_t pc = (size_t)(c0 + j * stride) + w;
            float sv[5];
#pragma unroll
            for (int d = 0; d < 5; ++d) sv[d] = first ? e[j][d] : e[j][d] + (beta * so[j][d]);
            s_new[pc] = sv[0];
            if (part_delta_d) {
                // relative mode: the products of FP32 operands are exact in
                // FP64; the row is summed in FP64 (the same fused chain) and
                // kept as an FP32 pair, its rounding and the rest (48 bits),
                // so that the update of r loses nothing but its own rounding
                // to FP32, which r_lo keeps
                double vd = (double)dg[j] * (double)sv[0];
                vd = fma((double)px[j][0], (double)sv[1], vd);
                vd = fma((double)px[j][1], (double)sv[2], vd);
                vd = fma((double)py[j][0], (double)sv[3], vd);
                vd = fma((double)py[j][1], (double)sv[4], vd);
                const float hi = (float)vd;
                As[pc] = hi;
                As_lo[pc] = (float)(vd - (double)hi);
                dprod += (double)sv[0] * vd;
            } else {
                float v = dg[j] * sv[0];
                v = fmaf(px[j][0], sv[1], v);
                v = fmaf(px[j][1], sv[2], v);
                v = fmaf(py[j][0], sv[3], v);
                v = fmaf(py[j][1], sv[4], v);
                As[pc] = v;
                prod += v * sv[0];
            }
        }
    }
    if (part_delta_d) {
        const double bs = block_sum(dprod, dscratch);
        if (threadIdx.x == 0) part_delta_d[wg] = bs;
        return;
    }
    const float bs = block_sum(prod, scratch);
    if (threadIdx.x == 0) part_delta[wg] = bs;
}

// Iteration k (the uniform argument): the search vectors alternate between
// search[0] and search[1].
struct SpmvDotArgs {
    LevelView L;
    float* search[2];
    float* As;
    float* As_lo;  // relative mode: A s - As (A s in FP64 is As + As_lo)
    PcgState* st;
    const float* part_gamma;
    float* part_delta;
    const double* part_gamma_d;  // relative mode (else null)
    const double* part_asz_d;
    double* part_delta_d;
    int ngam;
    uint32_t nwg;
};

template <int NT>
struct SpmvDot {
    static constexpr int kThreads = NT;
    using Args = SpmvDotArgs;
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int k) {
        spmv_dot<NT>(a.L, a.search[k & 1], a.search[(k + 1) & 1], a.As, a.As_lo, a.st, a.part_gamma, a.part_delta, k,
                     a.part_gamma_d, a.part_asz_d, a.part_delta_d, a.ngam, wg, a.nwg);
    }
};

// The reference's rules (FP32): delta from the partials (exit when |delta| <
// 1e-5, src/apic.rs:1903), then p += alpha*s; r = fma(-alpha, As, r);
// partials of max |r| (:1907-1928).  The relative mode's update is
// axpy_tile.
template <int NT>
__device__ __forceinline__ void axpy_max(const LevelView& L, const float* __restrict__ s, const float* __restrict__ As,
                                float* __restrict__ pressure, PcgState* st, const float* __restrict__ part_delta,
                                float* __restrict__ part_max, int k, uint32_t wg, uint32_t nwg) {
    const bool lead = wg == 0 && threadIdx.x == 0;
    __shared__ float scratch[NT / 64];
    if (st->done) return;
    const uint32_t cells = L.cells, stride = nwg * NT;
    const size_t w = L.w;
    uint32_t c0 = wg * NT + threadIdx.x;
    uint8_t fl[kWideBatch];
    auto fetch_flags = [&]() {
#pragma unroll
        for (int j = 0; j < kWideBatch; ++j) {
            const uint32_t c = c0 + j * stride;
            fl[j] = c < cells ? L.fluid[c < cells ? c : 0] : 0;
        }
    };
    fetch_flags();  // with the partials
    const float d = sum_partials256(part_delta, (int)nwg, scratch);
    if (fabsf(d) < kSafety) {
        if (lead) { st->done = 1; st->exit_reason = 2; st->delta = d; }
        return;
    }
    const float alpha = st->gamma[k] / d;
    if (lead) { st->delta = d; st->alpha = alpha; }
    float m = 0.0f;
    for (; c0 < cells; c0 += kWideBatch * stride) {
        float p[kWideBatch], sv[kWideBatch], av[kWideBatch], rv[kWideBatch];
        uint8_t fl_now[kWideBatch];
#pragma unroll
        for (int j = 0; j < kWideBatch; ++j) {
            const uint32_t c = fl[j] ? c0 + j * stride : 0;
            p[j] = pressure[c];
            sv[j] = s[c + w];
            av[j] = As[c + w];
            rv[j] = L.r[c + w];
            fl_now[j] = fl[j];
        }
        c0 += kWideBatch * stride;
        fetch_flags();  // the next batch's, in flight during this one
        c0 -= kWideBatch * stride;
#pragma unroll
        for (int j = 0; j < kWideBatch; ++j) {
            if (!fl_now[j]) continue;
            const uint32_t c = c0 + j * stride;
            const size_t pc = (size_t)c + w;
            pressure[c] = p[j] + alpha * sv[j];
            const float nr = fmaf(-alpha, av[j], rv[j]);
            L.r[pc] = nr;
            m = fast_max(m, fabsf(nr));
        }
    }
    const float bm = block_max(m, scratch);
    if (threadIdx.x == 0) part_max[wg] = bm;
}

struct AxpyMaxArgs {
    LevelView L;
    float* search[2];
    const float* As;
    float* pressure;
    PcgState* st;
    const float* part_delta;
    float* part_max;
    uint32_t nwg;
};

template <int NT>
struct AxpyMax {
    static constexpr int kThreads = NT;
    using Args = AxpyMaxArgs;
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int k) {
        axpy_max<NT>(a.L, a.search[(k + 1) & 1], a.As, a.pressure, a.st, a.part_delta, a.part_max, k, wg, a.nwg);
    }
};

// The gamma test after k iterations, run at the host's check points so that
// the recorded state is final when the host reads it.
struct Decide {
    static constexpr int kThreads = kSolverThreads;
    struct Args {
        PcgState* st;
        const float* part_gamma;
        PcgState* host;
        const double* part_gamma_d;  // relative mode (else null)
        int ngam;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int k) {
        __shared__ float scratch[kThreads / 64];
        __shared__ double dscratch[kThreads / 64];
        PcgState* st = a.st;
        const bool lead = threadIdx.x == 0;
        if (!st->done) {
            if (a.part_gamma_d) {
                const double g = sum_partials256(a.part_gamma_d, a.ngam, dscratch);
                (void)gamma_exit_rel(st, g, k, lead);
            } else {
                const float g = sum_partials256(a.part_gamma, a.ngam, scratch);
                (void)gamma_exit(st, g, k, lead);
            }
        }
        if (a.host && threadIdx.x == 0) {  // the scalars the host reads (not the gamma history)
            PcgState* host = a.host;
            host->done = st->done;
            host->iters = st->iters;
            host->exit_reason = st->exit_reason;
            host->maxerr = st->maxerr;
            host->delta = st->delta;
            host->alpha = st->alpha;
            host->bnorm2 = st->bnorm2;
            host->rnorm2 = st->rnorm2;
        }
    }
};

// ---- fused V-cycle kernels: one launch per level on the way down and one on
// the way up.  A workgroup takes a T x T tile of the level and loads it with a
// kVH-cell ring into LDS, then recomputes what the ring needs, so that the
// per-cell arithmetic (operation order, fused multiply-adds) is exactly that
// of red_zero_cell, sor_cell and restrict_cell above, whatever the tile side.
// Ring positions outside the grid are loaded by their padded flat index, the
// address the global kernels read for a row's first and last cell; their
// coefficients there are zero.
//
// Tiles are 32 x 32 on a level with many of them and 16 x 16 on a level with
// few: there a launch fills a fraction of the CUs with one workgroup each, its
// time is the latency of the workgroup's phases, and a smaller tile spreads
// the level over four times as many CUs.

constexpr int kVH = 2;         // ring: the first colour of a sweep needs two
constexpr int kBaseTile = 32;  // tile side on the base level (one partial of r.z per tile)
static_assert(kBaseTile / 2 == 16, "MgBuild marks base tiles of 16 x 16 coarse cells");
constexpr int kWideTileThreads = 512;

template <int T, int NT = 256>
struct VGeom {
    static constexpr int kT = T;                        // tile side (cells of the level)
    static constexpr int kR = T + 2 * kVH;              // region side
    static constexpr int kN = kR * kR;                  // region cells
    static constexpr int kC = T / 2 + 2;                // parent cells across the region
    static constexpr int kPer = (kN + NT - 1) / NT;     // region cells per thread
    static constexpr int kCN = (T / 2) * (T / 2);       // the tile's coarse cells
    static constexpr int kCPer = (kCN + NT - 1) / NT;
    static constexpr int kTPer = (T * T + NT - 1) / NT;  // tile cells per thread
    static constexpr int kEPer = (kC * kC + NT - 1) / NT;  // parent errors per thread
    static_assert(kPer <= 32, "region slots are masks of 32 bits");
};

template <int T>
struct VTile {
    using G = VGeom<T>;
    float e[G::kN], r[G::kN], id[G::kN], px[G::kN], py[G::kN], dg[G::kN];
    uint8_t fl[G::kN];  // fluid flag; 0 outside the grid
};

// A thread's share of a region, held in registers between the loads and the
// LDS stores, so that all of the loads go out before any of them is waited
// for (compile-time stride: the workgroup has NT threads).
template <int T, int NT>
struct VRegs {
    static constexpr int kPer = VGeom<T, NT>::kPer;
    float r[kPer], id[kPer], px[kPer], py[kPer], dg[kPer], e[kPer];
    uint8_t fl[kPer];
    uint32_t inb, real;  // bit k: slot k is in the padded range / a cell of the grid
};

// Loads at clamped addresses (no branch per load, so that they all issue back
// to back); the values outside the range are replaced by zeros in vtile_put,
// after the loads the caller issues next.
template <int T, int NT>
__device__ __forceinline__ void vtile_fetch(const LevelView& L, int32_t X0, int32_t Y0, VRegs<T, NT>& g, bool load_e,
                                            bool load_dg) {
    using G = VGeom<T, NT>;
    const int64_t w = L.w, h = L.h, np = (int64_t)L.cells + 2 * w;
    g.inb = g.real = 0u;
#pragma unroll
    for (int k = 0; k < G::kPer; ++k) {
        const int i = (int)threadIdx.x + NT * k;
        const int32_t x = X0 - kVH + i % G::kR, y = Y0 - kVH + i / G::kR;
        const int64_t p = (int64_t)(y + 1) * w + x;
        const bool inb = i < G::kN && p >= 0 && p < np;
        const bool real = i < G::kN && x >= 0 && x < w && y >= 0 && y < h;
        g.inb |= (uint32_t)inb << k;
        g.real |= (uint32_t)real << k;
        const int64_t q = inb ? p : 0;
        g.r[k] = L.r[q];
        g.id[k] = L.inv_diag[q];
        g.px[k] = L.px[q];
        g.py[k] = L.py[q];
        g.dg[k] = load_dg ? L.diag[q] : 0.0f;
        g.e[k] = load_e ? L.e[q] : 0.0f;
        g.fl[k] = L.fluid[real ? (size_t)y * w + x : 0];
    }
}

template <int T, int NT>
__device__ __forceinline__ void vtile_put(const VRegs<T, NT>& g, VTile<T>& t) {
    using G = VGeom<T, NT>;
#pragma unroll
    for (int k = 0; k < G::kPer; ++k) {
        const int i = (int)threadIdx.x + NT * k;
        if (i >= G::kN) continue;
        const bool inb = (g.inb >> k) & 1u, real = (g.real >> k) & 1u;
        t.r[i] = inb ? g.r[k] : 0.0f;
        t.id[i] = inb ? g.id[k] : 0.0f;
        t.px[i] = inb ? g.px[k] : 0.0f;
        t.py[i] = inb ? g.py[k] : 0.0f;
        t.dg[i] = inb ? g.dg[k] : 0.0f;
        t.e[i] = inb ? g.e[k] : 0.0f;
        t.fl[i] = real ? g.fl[k] : 0;
    }
}

// SOR update of cell i from its own value and four neighbour values (sor_cell).
template <int T>
__device__ inline float vsor(const VTile<T>& t, int i, float own, float er, float et, float el, float eb) {
    // every operand loads at once; a cell with id = 0 keeps its value (the
    // select discards what was computed from its operands)
    const float id = t.id[i];
    float sum = t.r[i];
    sum = fmaf(-t.px[i], er, sum);
    sum = fmaf(-t.py[i], et, sum);
    sum = fmaf(-t.px[i - 1], el, sum);
    sum = fmaf(-t.py[i - VGeom<T>::kR], eb, sum);
    const float gs = sum * id;
    return id != 0.0f ? own + (kOmega * (gs - own)) : own;
}

// alpha = gamma / delta of iteration k in FP64 (src/apic.rs:1903-1907; the
// relative mode ends the solve when delta <= 0); false when the solve ends.
__device__ __forceinline__ bool pcg_alpha(PcgState* st, double d, double gam, bool lead, double& alpha_d) {
    if (!(d > 0.0)) {
        if (lead) { st->done = 1; st->exit_reason = 2; st->delta = (float)d; }
        return false;
    }
    alpha_d = gam / d;
    if (lead) { st->delta = (float)d; st->alpha = (float)alpha_d; st->alpha_d = alpha_d; }
    return true;
}

// Iteration k's update in the relative mode (src/apic.rs:1907-1928 with FP64
// scalars and pressure): p += alpha s, r = fma(-alpha, As, r) on the fluid
// cells, and the partials of ||r||^2, one per 32 x 32 base tile: a workgroup
// of NT threads takes one tile, thread t its cells t, t + NT, ... in turn,
// and the workgroup adds its threads' sums with block_sum.  The base
// down-sweep that fuses the update (VDownAx, the same NT) forms the same
// partials in the same order, so both paths give the same residual norm.
struct AxIn {
    float* search[2];
    const float* As;            // A s: its FP32 rounding,
    const float* As_lo;         // and the rest
    float* pressure;
    double* pd;                 // the FP64 pressure
    const double* part_delta_d; // SpmvDot's partials of delta
    uint32_t ndelta;
    float* r_out;               // the updated residual (the old one is read from the level's r)
    float* r_lo;                // the residual's low part, read and written in place
    double* part_rr_out;        // one partial of ||r + r_lo||^2 per tile
};

// The relative mode's residual is the pair (r, r_lo): r evolves as the FP32
// rounding of r - alpha As in FP64 (As the FP32 rounding of A s: the
// V-cycle's input, the same value wherever a tile's halo recomputes a
// neighbour's cell), and r_lo, one word per cell kept by the cell's own
// tile, takes what that drops, alpha As_lo included, so that r + r_lo follows
// b - A p to FP64 rounding (A s is summed in FP64 and p accumulates in FP64)
// and the residual test measures the true residual without recomputing it.
__device__ __forceinline__ float update_r(float r, double alpha_d, float as) {
    return (float)fma(-alpha_d, (double)as, (double)r);
}
// A tile's own cell: the new (r, r_lo) and the FP64 residual r + r_lo (returned).
__device__ __forceinline__ double update_r_pair(float r, float lo, double alpha_d, float as, float as_lo,
                                                float& r_new, float& lo_new) {
    const double t = fma(-alpha_d, (double)as, (double)r);
    const double v = fma(-alpha_d, (double)as_lo, t + (double)lo);
    r_new = (float)t;
    lo_new = (float)(v - (double)r_new);
    return v;
}

template <int NT>
__device__ __forceinline__ void axpy_tile(const LevelView& L, const AxIn& ax, PcgState* st, uint32_t tiles_x, int k,
                                          uint32_t wg) {
    constexpr int T = kBaseTile, kPer = (T * T + NT - 1) / NT;
    __shared__ double dscratch[NT / 64];
    const bool lead = wg == 0 && threadIdx.x == 0;
    const int32_t w = (int32_t)L.w, h = (int32_t)L.h;
    const int32_t X0 = (int32_t)(wg % tiles_x) * T, Y0 = (int32_t)(wg / tiles_x) * T;
    const float* __restrict__ s = ax.search[(k + 1) & 1];
    const int32_t dn = st->done;
    const double gam = st->gamma_d[k & 1];
    uint8_t fl[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int i = (int)threadIdx.x + NT * q;
        const int32_t x = X0 + i % T, y = Y0 + i / T;
        const bool in = i < T * T && x < w && y < h;
        const uint8_t f = L.fluid[in ? (size_t)y * w + x : 0];
        fl[q] = in ? f : 0;
    }
    const double d = sum_partials256(ax.part_delta_d, (int)ax.ndelta, dscratch);
    if (dn) return;
    double alpha_d;
    if (!pcg_alpha(st, d, gam, lead, alpha_d)) return;
    double pv[kPer];
    float sv[kPer], av[kPer], al[kPer], rv[kPer], lv[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int i = (int)threadIdx.x + NT * q;
        const size_t c = fl[q] ? (size_t)(Y0 + i / T) * w + X0 + i % T : 0;
        pv[q] = ax.pd[c];
        sv[q] = s[c + w];
        av[q] = ax.As[c + w];
        al[q] = ax.As_lo[c + w];
        rv[q] = L.r[c + w];
        lv[q] = ax.r_lo[c + w];
    }
    double rr = 0.0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        if (!fl[q]) continue;
        const int i = (int)threadIdx.x + NT * q;
        const size_t c = (size_t)(Y0 + i / T) * w + X0 + i % T;
        const double np = pv[q] + alpha_d * (double)sv[q];
        ax.pd[c] = np;
        ax.pressure[c] = (float)np;
        float nr, nl;
        const double v = update_r_pair(rv[q], lv[q], alpha_d, av[q], al[q], nr, nl);
        ax.r_out[c + w] = nr;
        ax.r_lo[c + w] = nl;
        rr += v * v;
    }
    rr = block_sum(rr, dscratch);
    if (threadIdx.x == 0) ax.part_rr_out[wg] = rr;
}

struct AxpyTileArgs {
    LevelView L;  // the base level (r read and written in place)
    AxIn ax;
    PcgState* st;
    uint32_t tiles_x;
};

template <int NT>
struct AxpyTile {
    static constexpr int kThreads = NT;
    using Args = AxpyTileArgs;
    static constexpr bool kDealRows = true;  // its tile rows are dealt to the XCDs in turn
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int k) {
        axpy_tile<NT>(a.L, a.ax, a.st, a.tiles_x, k, wg);
    }
};

// The relative mode's residual test after iteration k (the kernel that
// applies it: every workgroup computes the same decision from the same
// partials).  rr = ||r + r_lo||^2: the recurrence's, or, with true_r, the
// true residual b - A p that replaced it this iteration.  The solve ends
// when ||r + r_lo|| <= rel_tol ||b||.  true: the solve ends.  The scalars it
// reads, loaded by the caller with its other early loads.
struct RelTestIn {
    double tol2, bn2;
};
__device__ __forceinline__ RelTestIn rel_test_in(const PcgState* st) { return RelTestIn{st->tol2, st->bnorm2}; }
__device__ __forceinline__ bool rel_residual_test(PcgState* st, const RelTestIn& in, double rr, bool lead) {
    if (lead) st->rnorm2 = rr;
    if (rr <= in.tol2 * in.bn2) {
        if (lead) { st->done = 1; st->exit_reason = 3; }
        return true;
    }
    return false;
}

// Down: (base: e = 0 and the first red sweep) the black sweep of level F,
// then the restriction of b - A e into C fused with C's first red sweep.  On
// the base level, in a PCG iteration, every workgroup first applies the
// residual test of the previous update (max |r| < solve_error,
// src/apic.rs:1930, or the relative test).
template <int T, int NT, bool AX = false>
__device__ __forceinline__ void vdown(const LevelView& F, const LevelView& C, int base, int fuse_red, int spec,
                             const int32_t* __restrict__ done, uint32_t tiles_x, PcgState* st, int check,
                             float solve_error, const float* __restrict__ part_max, const double* __restrict__ part_rr,
                             const double* __restrict__ part_rr_t, uint32_t ntiles, uint32_t nred, const AxIn& ax,
                             int k, uint32_t wg, uint32_t lwg) {
    // wg: the tile; lwg: the workgroup's index in the launch (workgroup 0
    // records the solve's decisions)
    using G = VGeom<T, NT>;
    constexpr int kR = G::kR;
    __shared__ VTile<T> t;
    __shared__ float scratch[NT / 64];
    __shared__ double dscratch[NT / 64];
    const bool lead = lwg == 0 && threadIdx.x == 0;
    const int32_t X0 = (int32_t)(wg % tiles_x) * T, Y0 = (int32_t)(wg / tiles_x) * T;
    const int32_t w = (int32_t)F.w, h = (int32_t)F.h;
    const int32_t cw = (int32_t)C.w, ch = (int32_t)C.h;
    // Loads that do not depend on each other go out together, so that the
    // workgroup waits for them once: the done flag, the coarse cells' flags
    // and inverse diagonals, the residual test's partials and scalars, with
    // spec the region itself before it is known whether the tile has work (a
    // coarse level has few tiles: its time is the latency of these loads, not
    // their bytes), and with AX the flags of the tile and its region and the
    // partials of delta.
    const int32_t dn = *done;
    VRegs<T, NT> g;
    if (spec) vtile_fetch<T, NT>(F, X0, Y0, g, !base, true);
    uint8_t cfl[G::kCPer];
    float cid[G::kCPer];
    bool cin[G::kCPer];
#pragma unroll
    for (int q = 0; q < G::kCPer; ++q) {
        const int k = (int)threadIdx.x + NT * q;
        const int32_t cx = X0 / 2 + k % (T / 2), cy = Y0 / 2 + k / (T / 2);
        const bool in = k < G::kCN && cx < cw && cy < ch;  // (the restriction below skips the others)
        const size_t c = in ? (size_t)cy * cw + cx : 0;
        cfl[q] = C.fluid[c];
        cid[q] = C.inv_diag[c + cw];
        cin[q] = in;
    }
    constexpr int kRPer = (kRedBlocks + NT - 1) / NT;
    double prr[kRPer];
    float pmx[kRPer];
#pragma unroll
    for (int q = 0; q < kRPer; ++q) {
        const uint32_t i = threadIdx.x + (uint32_t)NT * q, ic = i < nred ? i : 0u;
        prr[q] = 0.0;
        pmx[q] = 0.0f;
        if (check && part_rr) prr[q] = part_rr[ic];
        else if (check && !part_rr_t) pmx[q] = part_max[ic];
    }
    const bool rel_test = check && (part_rr || part_rr_t);
    const RelTestIn rti = rel_test ? rel_test_in(st) : RelTestIn{0.0, 0.0};
    double gam = 0.0;
    uint8_t ofl[G::kTPer], rfl[G::kPer];
    if constexpr (AX) {
        gam = st->gamma_d[k & 1];
#pragma unroll
        for (int q = 0; q < G::kTPer; ++q) {
            const int i = (int)threadIdx.x + NT * q;
            const int32_t x = X0 + i % T, y = Y0 + i / T;
            const bool in = i < T * T && x < w && y < h;
            const uint8_t f = F.fluid[in ? (size_t)y * w + x : 0];
            ofl[q] = in ? f : 0;
        }
#pragma unroll
        for (int q = 0; q < G::kPer; ++q) {
            const int i = (int)threadIdx.x + NT * q;
            const int32_t x = X0 - kVH + i % kR, y = Y0 - kVH + i / kR;
            const bool real = i < G::kN && x >= 0 && x < w && y >= 0 && y < h;
            const uint8_t f = F.fluid[real ? (size_t)y * w + x : 0];
            rfl[q] = real ? f : 0;
        }
    }
    const double rr_t = (check && part_rr_t) ? sum_partials256(part_rr_t, (int)ntiles, dscratch) : 0.0;
    if (spec) vtile_put<T, NT>(g, t);  // read after the barrier of the tile test below
    if (dn) return;
    if (check && part_rr_t) {  // the update's partials, one per base tile
        if (rel_residual_test(st, rti, rr_t, lead)) return;
    } else if (check && part_rr) {  // the true residual's
        double rr = 0.0;
#pragma unroll
        for (int q = 0; q < kRPer; ++q)
            if (threadIdx.x + (uint32_t)NT * q < nred) rr += prr[q];
        rr = block_sum(rr, dscratch);
        if (rel_residual_test(st, rti, rr, lead)) return;
    } else if (check) {
        float m = 0.0f;
#pragma unroll
        for (int q = 0; q < kRPer; ++q)
            if (threadIdx.x + (uint32_t)NT * q < nred) m = fast_max(m, pmx[q]);
        m = block_max(m, scratch);
        if (lead) st->maxerr = m;
        if (m < solve_error) {
            if (lead) { st->done = 1; st->exit_reason = 3; }
            return;
        }
    }
    // A tile without a fluid cell (none of its coarse cells is fluid) has
    // nothing to compute: the values of non-fluid cells are only ever read
    // times zero coefficients, so it leaves them as they are.  It leaves
    // before it loads anything more, so that its slot on the CU frees early.
    double alpha_d = 0.0;
    bool any = false;
#pragma unroll
    for (int q = 0; q < G::kCPer; ++q) any = any || (cin[q] && cfl[q] != 0);
    __shared__ uint32_t any_slots[NT / 64];  // (one call per launch)
    const bool has_fluid = block_any<NT>(any, any_slots);
    if constexpr (AX) {
        // no fluid cell in the tile: a zero partial, as axpy_tile's (workgroup
        // 0 goes on to record alpha)
        if (!has_fluid && lwg != 0) {
            if (threadIdx.x == 0) ax.part_rr_out[wg] = 0.0;
            return;
        }
    } else {
        if (!has_fluid) return;
    }
    if (!spec) {
        if (has_fluid) vtile_fetch<T, NT>(F, X0, Y0, g, !base, true);
        if constexpr (AX) {
            // the update's operands: A s over the region's fluid cells, and
            // the tile's fluid cells' p, s, r and A s (non-fluid slots read
            // cell 0)
            const float* __restrict__ sv_in = ax.search[(k + 1) & 1];
            float asr[G::kPer];
#pragma unroll
            for (int q = 0; q < G::kPer; ++q) {
                const int i = (int)threadIdx.x + NT * q;
                const int32_t x = X0 - kVH + i % kR, y = Y0 - kVH + i / kR;
                asr[q] = ax.As[rfl[q] ? (size_t)(y + 1) * w + x : 0];
            }
            double pv[G::kTPer];
            float sv[G::kTPer], av[G::kTPer], al[G::kTPer], rv[G::kTPer], lv[G::kTPer];
#pragma unroll
            for (int q = 0; q < G::kTPer; ++q) {
                const int i = (int)threadIdx.x + NT * q;
                const size_t c = ofl[q] ? (size_t)(Y0 + i / T) * w + X0 + i % T : 0;
                pv[q] = ax.pd[c];
                sv[q] = sv_in[c + w];
                av[q] = ax.As[c + w];
                al[q] = ax.As_lo[c + w];
                rv[q] = F.r[c + w];
                lv[q] = ax.r_lo[c + w];
            }
            // delta's partials, their loads in flight with the tile's, then
            // alpha (every workgroup forms the same sum)
            const double dsum = sum_partials256(ax.part_delta_d, (int)ax.ndelta, dscratch);
            if (!pcg_alpha(st, dsum, gam, lead, alpha_d)) return;
            if (!has_fluid) {  // workgroup 0 of a tile without a fluid cell
                if (threadIdx.x == 0) ax.part_rr_out[wg] = 0.0;
                return;
            }
#pragma unroll
            for (int q = 0; q < G::kPer; ++q)
                if (rfl[q]) g.r[q] = update_r(g.r[q], alpha_d, asr[q]);
            vtile_put<T, NT>(g, t);
            // the tile's own cells, as axpy_tile updates them
            double rr = 0.0;
#pragma unroll
            for (int q = 0; q < G::kTPer; ++q) {
                if (!ofl[q]) continue;
                const int i = (int)threadIdx.x + NT * q;
                const size_t c = (size_t)(Y0 + i / T) * w + X0 + i % T;
                const double np = pv[q] + alpha_d * (double)sv[q];
                ax.pd[c] = np;
                ax.pressure[c] = (float)np;
                float nr, nl;
                const double v = update_r_pair(rv[q], lv[q], alpha_d, av[q], al[q], nr, nl);
                ax.r_out[c + w] = nr;
                ax.r_lo[c + w] = nl;
                rr += v * v;
            }
            rr = block_sum(rr, dscratch);  // (its barriers also order the region's LDS stores)
            if (threadIdx.x == 0) ax.part_rr_out[wg] = rr;
        } else {
            vtile_put<T, NT>(g, t);
            __syncthreads();
        }
    }
    if (base) {  // red_zero_cell over the region
        for (int i = threadIdx.x; i < G::kN; i += blockDim.x) {
            const int32_t x = X0 - kVH + i % kR, y = Y0 - kVH + i / kR;
            const bool red = ((x + y) & 1) == 0;
            const uint8_t fl = t.fl[i];  // (the three loads go out together)
            const float id = t.id[i], r = t.r[i];
            t.e[i] = (fl && red && id != 0.0f) ? kOmega * (r * id) : 0.0f;
        }
        __syncthreads();
    }
    // black cells of the tile and its one-cell ring (they read red cells
    // only).  A black cell's own value before its sweep is zero (the red
    // sweep before it leaves it as the restriction set it); it is not taken
    // from memory, where a neighbouring tile may already have written its
    // result for the ring's cells.
    for (int i = threadIdx.x; i < G::kN; i += blockDim.x) {
        const int lx = i % kR, ly = i / kR;
        if (lx < 1 || lx >= kR - 1 || ly < 1 || ly >= kR - 1) continue;
        const int32_t x = X0 - kVH + lx, y = Y0 - kVH + ly;
        if (((x + y) & 1) == 0) continue;  // (the colour alone gates the loads: they go out together)
        const float v = vsor<T>(t, i, 0.0f, t.e[i + 1], t.e[i + kR], t.e[i - 1], t.e[i - kR]);
        if (t.fl[i]) t.e[i] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < G::kTPer; ++q) {
        const int i = (int)threadIdx.x + NT * q;
        const int32_t x = X0 + i % T, y = Y0 + i / T;
        if (i >= T * T || x >= w || y >= h) continue;
        if (!base && ((x + y) & 1) == 0) continue;  // red values are already in memory
        const int li = (y - Y0 + kVH) * kR + (x - X0 + kVH);
        F.e[(size_t)(y + 1) * w + x] = t.e[li];
    }
    // restriction into the tile's coarse cells (restrict_cell)
#pragma unroll
    for (int q = 0; q < G::kCPer; ++q) {
        const int k = (int)threadIdx.x + NT * q;
        const int32_t cx = X0 / 2 + k % (T / 2), cy = Y0 / 2 + k / (T / 2);
        if (!cin[q]) continue;
        const size_t c = (size_t)cy * cw + cx, pcc = c + cw;
        if (!cfl[q]) { C.r[pcc] = 0.0f; C.e[pcc] = 0.0f; continue; }
        float sum = 0.0f;
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int li = (2 * cy + dy - Y0 + kVH) * kR + (2 * cx + dx - X0 + kVH);
            // every child's operands load unconditionally (they go out
            // together); a non-fluid child adds +0, which leaves the sum as
            // skipping it would: the sum starts at +0 and never becomes -0
                const float b = t.r[li];
                float ax = t.dg[li] * t.e[li];
                ax += t.px[li] * t.e[li + 1];
                ax += t.py[li] * t.e[li + kR];
                ax += t.px[li - 1] * t.e[li - 1];
                ax += t.py[li - kR] * t.e[li - kR];
                sum += t.fl[li] ? b - ax : 0.0f;
            }
        C.r[pcc] = sum;
        const float id = cid[q];
        const bool red = ((cx + cy) & 1) == 0;
        C.e[pcc] = (fuse_red && red && id != 0.0f) ? kOmega * (sum * id) : 0.0f;
    }
}

// One level on the way down.  On the base level of a PCG iteration (check)
// the residual test reads the nred partials of the previous update.
struct VDownArgs {
    LevelView F, C;
    int base, fuse_red;
    int spec;                // load the region with the flags (levels of few tiles)
    const int32_t* done;
    uint32_t tiles_x;
    PcgState* st;
    int check;
    float solve_error;
    const float* part_max;
    const double* part_rr_t;  // relative mode (else null): ||r||^2 partials of the update, one per base tile
    uint32_t ntiles;
    const double* part_rr2;   // ... and of the true residual (nred), on the iterations that recompute it
    int every;                // the true residual is recomputed after iteration k when (k + 1) % every == 0
    uint32_t nred;
    AxIn ax;                  // VDownAx: the iteration's update
    const uint32_t* order;    // base level of a simulation stepped alone (else null): TileOrder's tile order
};

// The tile a base-level workgroup runs: its own, or with TileOrder's order
// the one that order gives it, and whether it may leave at once (a tile
// without fluid whose partials TileOrder zeroed; workgroup 0 always runs).
__device__ __forceinline__ bool ordered_tile(const uint32_t* order, uint32_t wg, uint32_t& tile) {
    tile = wg;
    if (!order) return true;
    const uint32_t o = order[wg];
    tile = o & 0x7FFFFFFFu;
    return (o >> 31) == 0u || wg == 0;
}

template <int T, int NT = 256, bool AX = false>
struct VDown {
    static constexpr int kThreads = NT;
    using Args = VDownArgs;
    static constexpr bool kDealRows = true;  // its tile rows are dealt to the XCDs in turn
    // k: the PCG iteration whose update the residual test checks
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int k) {
        uint32_t tile;
        if (!ordered_tile(a.order, wg, tile)) return;
        const bool due = a.every > 0 && (k + 1) % a.every == 0;
        vdown<T, NT, AX>(a.F, a.C, a.base, a.fuse_red, a.spec, a.done, a.tiles_x, a.st, a.check, a.solve_error,
                         a.part_max, due ? a.part_rr2 : nullptr, due ? nullptr : a.part_rr_t, a.ntiles, a.nred, a.ax,
                         k, tile, wg);
    }
};
// The base level's down-sweep with iteration k's update (axpy_tile) applied
// on the way in: the relative mode's iterations that do not recompute the
// true residual, for a simulation stepped alone.  The residual test of the
// update moves to the next launch (VDown2's check).
template <int NT>
using VDownAx = VDown<kBaseTile, NT, true>;

// ---- two levels of the way down in one launch: levels l and l + 1, both
// above the bottom and neither the base.  A workgroup takes a kP2T x kP2T
// tile of level l + 1 and the 2 kP2T x 2 kP2T tile of level l under it.  It
// loads level l with a kP2FR-cell ring, so that everything the coarse tile's
// black sweep and restriction read is computed inside the workgroup: level
// l's black sweep out to ring kP2FR - 1, the restriction into level l + 1's
// cells out to ring kP2CR (with their fused red sweep), level l + 1's black
// sweep out to ring 1, and the restriction into level l + 2 under the coarse
// tile.  Each value is computed with exactly the arithmetic of vdown, so the
// result is that of the two launches.  Cells of level l + 1 outside its grid
// enter only times zero coefficients (the grid's boundary couplings are zero:
// its outermost cells are walls on the base level), so they are taken as
// zero.  The tile is skipped as a whole when none of its level l + 1 cells
// is fluid (both launches skip it then).
constexpr int kP2T = 16;                    // tile side on level l + 1
constexpr int kP2FR = 6;                    // ring on level l
constexpr int kP2F = 2 * kP2T + 2 * kP2FR;  // region side on level l
constexpr int kP2CR = 2;                    // ring on level l + 1
constexpr int kP2C = kP2T + 2 * kP2CR;      // region side on level l + 1
// A level pair's tiles are few and most are fluid-free: the active ones run
// alone on their CUs, and 16 waves hide more of their latency than 8.
constexpr int kP2Threads = 1024;

template <int S>
struct Region {
    static constexpr int kN = S * S;
    float e[kN], r[kN], id[kN], px[kN], py[kN], dg[kN];
    uint8_t fl[kN];  // fluid flag; 0 outside the grid
};

// A thread's share of a region of side S (stride NT) between its loads and
// its LDS stores, as VRegs for the tiles.
template <int S, int NT>
struct RegionRegs {
    static constexpr int kPer = (S * S + NT - 1) / NT;
    float r[kPer], id[kPer], px[kPer], py[kPer], dg[kPer], e[kPer];
    uint8_t fl[kPer];
    uint32_t inb, real;
};

// Clamped loads of the region at (X0, Y0) by padded flat index (as
// vtile_fetch); r and e only when asked.
template <int S, int NT>
__device__ __forceinline__ void region_fetch(const LevelView& L, int32_t X0, int32_t Y0, RegionRegs<S, NT>& g,
                                             bool load_r, bool load_e) {
    const int64_t w = L.w, h = L.h, np = (int64_t)L.cells + 2 * w;
    g.inb = g.real = 0u;
#pragma unroll
    for (int k = 0; k < RegionRegs<S, NT>::kPer; ++k) {
        const int i = (int)threadIdx.x + NT * k;
        const int32_t x = X0 + i % S, y = Y0 + i / S;
        const int64_t p = (int64_t)(y + 1) * w + x;
        const bool inb = i < S * S && p >= 0 && p < np;
        const bool real = i < S * S && x >= 0 && x < w && y >= 0 && y < h;
        g.inb |= (uint32_t)inb << k;
        g.real |= (uint32_t)real << k;
        const int64_t q = inb ? p : 0;
        g.r[k] = load_r ? L.r[q] : 0.0f;
        g.id[k] = L.inv_diag[q];
        g.px[k] = L.px[q];
        g.py[k] = L.py[q];
        g.dg[k] = L.diag[q];
        g.e[k] = load_e ? L.e[q] : 0.0f;
        g.fl[k] = L.fluid[real ? (size_t)y * w + x : 0];
    }
}

template <int S, int NT>
__device__ __forceinline__ void region_put(const RegionRegs<S, NT>& g, Region<S>& t) {
#pragma unroll
    for (int k = 0; k < RegionRegs<S, NT>::kPer; ++k) {
        const int i = (int)threadIdx.x + NT * k;
        if (i >= S * S) continue;
        const bool inb = (g.inb >> k) & 1u, real = (g.real >> k) & 1u;
        t.r[i] = inb ? g.r[k] : 0.0f;
        t.id[i] = inb ? g.id[k] : 0.0f;
        t.px[i] = inb ? g.px[k] : 0.0f;
        t.py[i] = inb ? g.py[k] : 0.0f;
        t.dg[i] = inb ? g.dg[k] : 0.0f;
        t.e[i] = inb ? g.e[k] : 0.0f;
        t.fl[i] = real ? g.fl[k] : 0;
    }
}

// Whether one of the thread's fetched cells of a region of side S is a fluid
// cell of the square [lo, lo + T)^2 of region coordinates (from its own
// registers: the region's LDS copy is complete only after a barrier).
template <int S, int NT, int T>
__device__ __forceinline__ bool region_any_fluid(const RegionRegs<S, NT>& g, int lo) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < RegionRegs<S, NT>::kPer; ++k) {
        const int i = (int)threadIdx.x + NT * k;
        const int lx = i % S, ly = i / S;
        const bool tile = i < S * S && lx >= lo && lx < lo + T && ly >= lo && ly < lo + T;
        any = any || (tile && ((g.real >> k) & 1u) && g.fl[k] != 0);
    }
    return any;
}

// sor_cell's update of region cell i of side S
template <int S>
__device__ __forceinline__ float rsor(const Region<S>& t, int i, float own, float er, float et, float el, float eb) {
    const float id = t.id[i];  // (as vsor: operands at once, a select at the end)
    float sum = t.r[i];
    sum = fmaf(-t.px[i], er, sum);
    sum = fmaf(-t.py[i], et, sum);
    sum = fmaf(-t.px[i - 1], el, sum);
    sum = fmaf(-t.py[i - S], eb, sum);
    const float gs = sum * id;
    return id != 0.0f ? own + (kOmega * (gs - own)) : own;
}

// restrict_cell's sum b - A e over the four children of a coarse cell whose
// lower-left child is region cell i0 of side S
template <int S>
__device__ __forceinline__ float rrestrict(const Region<S>& t, int i0) {
    float sum = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const int li = i0 + dy * S + dx;
            // (as vdown's restriction: unconditional loads, +0 for a
            // non-fluid child)
            const float b = t.r[li];
            float ax = t.dg[li] * t.e[li];
            ax += t.px[li] * t.e[li + 1];
            ax += t.py[li] * t.e[li + S];
            ax += t.px[li - 1] * t.e[li - 1];
            ax += t.py[li - S] * t.e[li - S];
            sum += t.fl[li] ? b - ax : 0.0f;
        }
    return sum;
}

__device__ __forceinline__ void vdown2(const LevelView& F, const LevelView& C, const LevelView& D, int fuse_red_d,
                                       const int32_t* __restrict__ done, uint32_t tiles_x, PcgState* st,
                                       const double* __restrict__ part_rr_t, uint32_t ntiles, uint32_t wg) {
    constexpr int NT = kP2Threads;
    __shared__ Region<kP2F> tf;
    __shared__ Region<kP2C> tc;
    __shared__ double dscratch[NT / 64];
    const int32_t w = (int32_t)F.w, h = (int32_t)F.h, cw = (int32_t)C.w, ch = (int32_t)C.h;
    const int32_t dw = (int32_t)D.w, dh = (int32_t)D.h;
    const int32_t cx0 = (int32_t)(wg % tiles_x) * kP2T, cy0 = (int32_t)(wg / tiles_x) * kP2T;
    const int32_t FX0 = 2 * cx0 - kP2FR, FY0 = 2 * cy0 - kP2FR;  // region origins
    const int32_t CX0 = cx0 - kP2CR, CY0 = cy0 - kP2CR;
    // one batch of loads: the done flag, both regions, the coarse tile's
    // children's flags and inverse diagonals on level l + 2
    const int32_t dn = *done;
    RegionRegs<kP2F, NT> gf;
    RegionRegs<kP2C, NT> gc;
    region_fetch<kP2F, NT>(F, FX0, FY0, gf, true, true);
    region_fetch<kP2C, NT>(C, CX0, CY0, gc, false, false);
    constexpr int kDT = kP2T / 2;
    const int32_t ddx = (int32_t)threadIdx.x % kDT, ddy = (int32_t)threadIdx.x / kDT;
    const int32_t dcx = cx0 / 2 + ddx, dcy = cy0 / 2 + ddy;
    const bool din = (int)threadIdx.x < kDT * kDT && dcx < dw && dcy < dh;
    const size_t dc = din ? (size_t)dcy * dw + dcx : 0;
    const uint8_t dfl = D.fluid[dc];
    const float did = D.inv_diag[dc + dw];
    // with part_rr_t, the residual test of the update that the base level's
    // launch before this one fused (as vdown's)
    const RelTestIn rti = part_rr_t ? rel_test_in(st) : RelTestIn{0.0, 0.0};
    const double rr = part_rr_t ? sum_partials256(part_rr_t, (int)ntiles, dscratch) : 0.0;
    region_put<kP2F, NT>(gf, tf);
    region_put<kP2C, NT>(gc, tc);
    if (dn) return;
    if (part_rr_t) {
        const bool lead = wg == 0 && threadIdx.x == 0;
        if (rel_residual_test(st, rti, rr, lead)) {
            // the new residual is in the base level's second buffer: the
            // base up-sweep copies it home
            if (lead) st->r_in_alt = 1;
            return;
        }
    }
    // skip: no fluid cell of level l + 1 in the tile
    __shared__ uint32_t any_slots[NT / 64];  // (one call per launch)
    if (!block_any<NT>(region_any_fluid<kP2C, NT, kP2T>(gc, kP2CR), any_slots)) return;
    // level l: black cells out to ring kP2FR - 1 (own value zero before the sweep)
    for (int i = threadIdx.x; i < kP2F * kP2F; i += NT) {
        const int lx = i % kP2F, ly = i / kP2F;
        if (lx < 1 || lx >= kP2F - 1 || ly < 1 || ly >= kP2F - 1) continue;
        const int32_t x = FX0 + lx, y = FY0 + ly;
        if (((x + y) & 1) == 0) continue;
        const float v = rsor<kP2F>(tf, i, 0.0f, tf.e[i + 1], tf.e[i + kP2F], tf.e[i - 1], tf.e[i - kP2F]);
        if (tf.fl[i]) tf.e[i] = v;
    }
    __syncthreads();
    // level l + 1 out to ring kP2CR: restriction with the fused red sweep
    for (int i = threadIdx.x; i < kP2C * kP2C; i += NT) {
        const int lx = i % kP2C, ly = i / kP2C;
        const int32_t cx = CX0 + lx, cy = CY0 + ly;
        // (every region cell's children lie in tf: the sum needs no flag
        // first; a cell outside the grid or not fluid takes zeros)
        const float rs = rrestrict<kP2F>(tf, (2 * ly + kP2FR - 2 * kP2CR) * kP2F + 2 * lx + kP2FR - 2 * kP2CR);
        const float id = tc.id[i];
        const bool ok = cx >= 0 && cx < cw && cy >= 0 && cy < ch && tc.fl[i];
        const bool red = ((cx + cy) & 1) == 0;
        tc.r[i] = ok ? rs : 0.0f;
        tc.e[i] = (ok && red && id != 0.0f) ? kOmega * (rs * id) : 0.0f;
    }
    __syncthreads();
    // level l + 1: black cells of the tile and its one-cell ring
    for (int i = threadIdx.x; i < kP2C * kP2C; i += NT) {
        const int lx = i % kP2C, ly = i / kP2C;
        if (lx < 1 || lx >= kP2C - 1 || ly < 1 || ly >= kP2C - 1) continue;
        const int32_t cx = CX0 + lx, cy = CY0 + ly;
        if (((cx + cy) & 1) == 0) continue;
        const float v = rsor<kP2C>(tc, i, 0.0f, tc.e[i + 1], tc.e[i + kP2C], tc.e[i - 1], tc.e[i - kP2C]);
        if (tc.fl[i]) tc.e[i] = v;
    }
    // level l: the tile's black cells (its red cells are already in memory)
    for (int i = threadIdx.x; i < 4 * kP2T * kP2T; i += NT) {
        const int32_t x = 2 * cx0 + i % (2 * kP2T), y = 2 * cy0 + i / (2 * kP2T);
        if (x >= w || y >= h || ((x + y) & 1) == 0) continue;
        const int li = (y - FY0) * kP2F + (x - FX0);
        F.e[(size_t)(y + 1) * w + x] = tf.e[li];
    }
    __syncthreads();
    // level l + 1: the tile's residual and error
    for (int i = threadIdx.x; i < kP2T * kP2T; i += NT) {
        const int32_t cx = cx0 + i % kP2T, cy = cy0 + i / kP2T;
        if (cx >= cw || cy >= ch) continue;
        const int li = (cy - CY0) * kP2C + (cx - CX0);
        const size_t pc = (size_t)(cy + 1) * cw + cx;
        C.r[pc] = tc.r[li];
        C.e[pc] = tc.e[li];
    }
    // level l + 2 under the tile (restrict_cell with its fused red sweep)
    if (din) {
        const size_t pdc = dc + dw;
        if (!dfl) {
            D.r[pdc] = 0.0f;
            D.e[pdc] = 0.0f;
        } else {
            const float sum = rrestrict<kP2C>(tc, (2 * dcy - CY0) * kP2C + 2 * dcx - CX0);
            D.r[pdc] = sum;
            const bool red = ((dcx + dcy) & 1) == 0;
            D.e[pdc] = (fuse_red_d && red && did != 0.0f) ? kOmega * (sum * did) : 0.0f;
        }
    }
}

struct VDown2Args {
    LevelView F, C, D;  // levels l, l + 1, l + 2, each with its down-sweep error
    int fuse_red_d;
    const int32_t* done;
    uint32_t tiles_x;  // tiles of level l + 1 in a row
    PcgState* st;
    const double* part_rr_t;  // after VDownAx: the update's ||r||^2 partials (else null)
    uint32_t ntiles;
};

struct VDown2 {
    static constexpr int kThreads = kP2Threads;
    using Args = VDown2Args;
    static constexpr bool kDealRows = true;  // its tile rows are dealt to the XCDs in turn
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        vdown2(a.F, a.C, a.D, a.fuse_red_d, a.done, a.tiles_x, a.st, a.part_rr_t, a.ntiles, wg);
    }
};

// ---- two levels of the way up in one launch: level l + 1 with level l + 2's
// final error prolongated, then level l with level l + 1's.  A workgroup
// takes a kP2T x kP2T tile of level l + 1 and the tile of level l under it.
// Level l + 1 is loaded with a three-cell ring: its black sweep out to ring 2
// and its red sweep out to ring 1 give the final error of every cell whose
// value level l's region prolongates; level l then runs as in vup.  Every
// value is computed with the arithmetic of vup, so the result is that of the
// two launches; the tile is skipped as a whole when none of its level l + 1
// cells is fluid (both launches skip it then).
constexpr int kU2CR = 3;                    // ring on level l + 1
constexpr int kU2C = kP2T + 2 * kU2CR;      // its region side
constexpr int kU2D = kP2T / 2 + 4;          // level l + 2 cells across that region
constexpr int kU2F = 2 * kP2T + 2 * kVH;    // level l's region side (ring kVH)

__device__ __forceinline__ void vup2(const LevelView& F, const LevelView& C, const LevelView& D,
                                     float* __restrict__ f_out, float* __restrict__ c_out,
                                     const int32_t* __restrict__ done, uint32_t tiles_x, uint32_t wg) {
    constexpr int NT = kP2Threads;
    __shared__ Region<kU2F> tf;
    __shared__ Region<kU2C> tc;
    __shared__ float de[kU2D * kU2D];
    const int32_t w = (int32_t)F.w, h = (int32_t)F.h, cw = (int32_t)C.w, ch = (int32_t)C.h;
    const int32_t dw = (int32_t)D.w, dh = (int32_t)D.h;
    const int32_t cx0 = (int32_t)(wg % tiles_x) * kP2T, cy0 = (int32_t)(wg / tiles_x) * kP2T;
    const int32_t CX0 = cx0 - kU2CR, CY0 = cy0 - kU2CR;        // level l + 1 region
    const int32_t DX0 = CX0 >> 1, DY0 = CY0 >> 1;                // level l + 2 parents
    const int32_t FX0 = 2 * cx0 - kVH, FY0 = 2 * cy0 - kVH;    // level l region
    // one batch of loads: the done flag, both regions (down-sweep errors) and
    // level l + 2's final error under level l + 1's region
    const int32_t dn = *done;
    RegionRegs<kU2F, NT> gf;
    RegionRegs<kU2C, NT> gc;
    region_fetch<kU2F, NT>(F, FX0, FY0, gf, true, true);
    region_fetch<kU2C, NT>(C, CX0, CY0, gc, true, true);
    constexpr int kDPer = (kU2D * kU2D + NT - 1) / NT;
    float dv[kDPer];
    uint32_t dmask = 0u;
#pragma unroll
    for (int q = 0; q < kDPer; ++q) {
        const int j = (int)threadIdx.x + NT * q;
        const int32_t dx = DX0 + j % kU2D, dy = DY0 + j / kU2D;
        const bool in = j < kU2D * kU2D && dx >= 0 && dx < dw && dy >= 0 && dy < dh;
        dmask |= (uint32_t)in << q;
        dv[q] = D.e[in ? (size_t)(dy + 1) * dw + dx : 0];
    }
    region_put<kU2F, NT>(gf, tf);
    region_put<kU2C, NT>(gc, tc);
#pragma unroll
    for (int q = 0; q < kDPer; ++q) {
        const int j = (int)threadIdx.x + NT * q;
        if (j < kU2D * kU2D) de[j] = ((dmask >> q) & 1u) ? dv[q] : 0.0f;
    }
    if (dn) return;
    // skip: no fluid cell of level l + 1 in the tile
    __shared__ uint32_t any_slots[NT / 64];  // (one call per launch)
    if (!block_any<NT>(region_any_fluid<kU2C, NT, kP2T>(gc, kU2CR), any_slots)) return;
    // level l + 1: e plus the parent's error over the region (the diagonal's slot)
    for (int i = threadIdx.x; i < kU2C * kU2C; i += NT) {
        // (every region cell's parent lies in de: the load needs no flag)
        const int32_t x = CX0 + i % kU2C, y = CY0 + i / kU2C;
        const float pe = de[((y >> 1) - DY0) * kU2D + ((x >> 1) - DX0)];
        const float v = tc.e[i];
        tc.dg[i] = tc.fl[i] ? v + pe : v;
    }
    __syncthreads();
    // level l + 1: black cells out to ring 2
    for (int i = threadIdx.x; i < kU2C * kU2C; i += NT) {
        const int lx = i % kU2C, ly = i / kU2C;
        if (lx < 1 || lx >= kU2C - 1 || ly < 1 || ly >= kU2C - 1) continue;
        const int32_t x = CX0 + lx, y = CY0 + ly;
        if (((x + y) & 1) == 0) continue;
        const float v = rsor<kU2C>(tc, i, tc.dg[i], tc.dg[i + 1], tc.dg[i + kU2C], tc.dg[i - 1], tc.dg[i - kU2C]);
        if (tc.fl[i]) tc.e[i] = v;
    }
    __syncthreads();
    // level l + 1: red cells out to ring 1, written in place (red cells read
    // black neighbours only); non-fluid cells keep e
    for (int i = threadIdx.x; i < kU2C * kU2C; i += NT) {
        const int lx = i % kU2C, ly = i / kU2C;
        if (lx < 2 || lx >= kU2C - 2 || ly < 2 || ly >= kU2C - 2) continue;
        const int32_t x = CX0 + lx, y = CY0 + ly;
        if (((x + y) & 1) != 0) continue;
        const float v = rsor<kU2C>(tc, i, tc.dg[i], tc.e[i + 1], tc.e[i + kU2C], tc.e[i - 1], tc.e[i - kU2C]);
        if (tc.fl[i]) tc.e[i] = v;
    }
    __syncthreads();
    // level l + 1: the tile's final error
    for (int i = threadIdx.x; i < kP2T * kP2T; i += NT) {
        const int32_t x = cx0 + i % kP2T, y = cy0 + i / kP2T;
        if (x >= cw || y >= ch) continue;
        c_out[(size_t)(y + 1) * cw + x] = tc.e[(y - CY0) * kU2C + (x - CX0)];
    }
    // level l: e plus the parent's final error (the parent of a fluid cell is
    // a cell of the grid, within ring 1 of the level l + 1 tile)
    for (int i = threadIdx.x; i < kU2F * kU2F; i += NT) {
        // (every region cell's parent lies in the level l + 1 region)
        const int32_t x = FX0 + i % kU2F, y = FY0 + i / kU2F;
        const float pe = tc.e[((y >> 1) - CY0) * kU2C + ((x >> 1) - CX0)];
        const float v = tf.e[i];
        tf.dg[i] = tf.fl[i] ? v + pe : v;
    }
    __syncthreads();
    // level l: black cells of the tile and its one-cell ring
    for (int i = threadIdx.x; i < kU2F * kU2F; i += NT) {
        const int lx = i % kU2F, ly = i / kU2F;
        if (lx < 1 || lx >= kU2F - 1 || ly < 1 || ly >= kU2F - 1) continue;
        const int32_t x = FX0 + lx, y = FY0 + ly;
        if (((x + y) & 1) == 0) continue;
        const float v = rsor<kU2F>(tf, i, tf.dg[i], tf.dg[i + 1], tf.dg[i + kU2F], tf.dg[i - 1], tf.dg[i - kU2F]);
        if (tf.fl[i]) tf.e[i] = v;
    }
    __syncthreads();
    // level l: red cells of the tile, then the tile's final error
    for (int i = threadIdx.x; i < 4 * kP2T * kP2T; i += NT) {
        const int32_t x = 2 * cx0 + i % (2 * kP2T), y = 2 * cy0 + i / (2 * kP2T);
        if (x >= w || y >= h) continue;
        const int li = (y - FY0) * kU2F + (x - FX0);
        float e = tf.e[li];
        if (((x + y) & 1) == 0) {
            const float v = rsor<kU2F>(tf, li, tf.dg[li], tf.e[li + 1], tf.e[li + kU2F], tf.e[li - 1], tf.e[li - kU2F]);
            if (tf.fl[li]) e = v;
        }
        f_out[(size_t)(y + 1) * w + x] = e;
    }
}

struct VUp2Args {
    LevelView F, C, D;  // levels l, l + 1 (their down-sweep errors) and l + 2 (its final error)
    float* f_out;       // the final errors of levels l and l + 1
    float* c_out;
    const int32_t* done;
    uint32_t tiles_x;  // tiles of level l + 1 in a row
};

struct VUp2 {
    static constexpr int kThreads = kP2Threads;
    using Args = VUp2Args;
    static constexpr bool kDealRows = true;  // its tile rows are dealt to the XCDs in turn
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        vup2(a.F, a.C, a.D, a.f_out, a.c_out, a.done, a.tiles_x, wg);
    }
};

// Up: the two colour sweeps of smooth_level's reverse pass on level L with
// C's error prolongated on the fly (prolongate + sor_cell modes 1 and 2); on
// the base level, with the padded-list repeats of prolongate and the partials
// of r.z over the tiles a workgroup takes (dot: 1 FP32, 2 FP64).  L.e is the
// level's down-sweep buffer and the result goes to e_out: a workgroup's ring
// must not see a neighbouring tile's result.
template <int T, int NT>
__device__ __forceinline__ void vup(const LevelView& L, const LevelView& C, const FluidStats* __restrict__ fs,
                           const int32_t* __restrict__ done, uint32_t tiles_x, int dot, float* __restrict__ part_gamma,
                           double* __restrict__ part_gamma_d, const float* __restrict__ As,
                           double* __restrict__ part_asz_d, int spec, float* __restrict__ e_out,
                           float* __restrict__ r_home, const int32_t* __restrict__ r_in_alt, uint32_t wg) {
    using G = VGeom<T, NT>;
    constexpr int kR = G::kR, kC = G::kC;
    __shared__ VTile<T> t;
    __shared__ float ce[kC * kC];
    __shared__ float scratch[NT / 64];
    __shared__ double dscratch[2 * (NT / 64)];  // (block_sum2)
    const int32_t w = (int32_t)L.w, h = (int32_t)L.h, cw = (int32_t)C.w, chh = (int32_t)C.h;
    const int32_t X0 = (int32_t)(wg % tiles_x) * T, Y0 = (int32_t)(wg / tiles_x) * T;
    const int32_t CX0 = X0 / 2 - 1, CY0 = Y0 / 2 - 1;
    // Loads that do not depend on each other go out together (as in vdown):
    // the done flag, the region (with spec) and A s, the parent's error over
    // the region, the tile's coarse flags and the fluid statistics.
    const int32_t dn = *done;
    // A s of the thread's tile cells (the partials of (A s).z)
    float asv[G::kTPer];
    auto fetch_as = [&]() {
#pragma unroll
        for (int q = 0; q < G::kTPer; ++q) {
            const int i = (int)threadIdx.x + NT * q;
            const int32_t x = X0 + i % T, y = Y0 + i / T;
            const bool in = i < T * T && x < w && y < h;  // (the sweep below skips the others)
            asv[q] = dot == 2 ? As[in ? (size_t)(y + 1) * w + x : 0] : 0.0f;
        }
    };
    VRegs<T, NT> g;
    // the parent's error over the region (read after a barrier)
    float cev[G::kEPer];
    uint32_t cin_mask = 0u;
    auto fetch_ce = [&]() {
#pragma unroll
        for (int q = 0; q < G::kEPer; ++q) {
            const int j = (int)threadIdx.x + NT * q;
            const int32_t cx = CX0 + j % kC, cy = CY0 + j / kC;
            const bool in = j < kC * kC && cx >= 0 && cx < cw && cy >= 0 && cy < chh;
            cin_mask |= (uint32_t)in << q;
            cev[q] = C.e[in ? (size_t)(cy + 1) * cw + cx : 0];
        }
    };
    auto put_ce = [&]() {
#pragma unroll
        for (int q = 0; q < G::kEPer; ++q) {
            const int j = (int)threadIdx.x + NT * q;
            if (j < kC * kC) ce[j] = ((cin_mask >> q) & 1u) ? cev[q] : 0.0f;
        }
    };
    if (spec) {
        vtile_fetch<T, NT>(L, X0, Y0, g, true, false);
        fetch_as();
        fetch_ce();
    }
    const int32_t ccx = X0 / 2 + (int32_t)(threadIdx.x % (T / 2)), ccy = Y0 / 2 + (int32_t)(threadIdx.x / (T / 2));
    const bool cin = threadIdx.x < G::kCN && ccx < cw && ccy < chh;
    const uint8_t cfv = C.fluid[cin ? (size_t)ccy * cw + ccx : 0];
    const uint32_t last1 = fs ? fs->last_raw : 0u, cnt = fs ? fs->count : 0u;
    const int32_t ria = r_home ? *r_in_alt : 0;
    if (spec) {  // read after the barrier of the tile test below
        vtile_put<T, NT>(g, t);
        put_ce();
    }
    if (dn) {
        // the solve ended at the residual test of a fused update: its
        // residual (L.r, the second buffer) goes home over the fluid cells
        if (ria) {
#pragma unroll
            for (int q = 0; q < G::kTPer; ++q) {
                const int i = (int)threadIdx.x + NT * q;
                const int32_t x = X0 + i % T, y = Y0 + i / T;
                if (i >= T * T || x >= w || y >= h || !L.fluid[(size_t)y * w + x]) continue;
                const size_t pc = (size_t)(y + 1) * w + x;
                r_home[pc] = L.r[pc];
            }
        }
        return;
    }
    // a tile without a fluid cell (see vdown) only reports a zero partial,
    // before it loads anything more
    __shared__ uint32_t any_slots[NT / 64];  // (one call per launch)
    if (!block_any<NT>(cin && cfv != 0, any_slots)) {
        if (threadIdx.x == 0) {
            if (dot == 2) { part_gamma_d[wg] = 0.0; part_asz_d[wg] = 0.0; }
            else if (dot == 1) part_gamma[wg] = 0.0f;
        }
        return;
    }
    if (!spec) {
        vtile_fetch<T, NT>(L, X0, Y0, g, true, false);
        fetch_as();
        fetch_ce();
        put_ce();
        if (r_home) {
            // r from the second buffer, whose non-fluid cells are never
            // written: zero there (they enter the r.z partials times e = 0)
#pragma unroll
            for (int q = 0; q < G::kPer; ++q)
                if (!(((g.real >> q) & 1u) && g.fl[q])) g.r[q] = 0.0f;
        }
        vtile_put<T, NT>(g, t);
        __syncthreads();
    }
    // e plus the parent's error (prolonged) of every region cell, once, into
    // the diagonal's slot (not read on the way up); non-fluid cells keep e
    for (int i = threadIdx.x; i < G::kN; i += blockDim.x) {
        // (every region cell's parent lies in ce: the load needs no flag)
        const int32_t x = X0 - kVH + i % kR, y = Y0 - kVH + i / kR;
        const float pc = ce[((y >> 1) - CY0) * kC + ((x >> 1) - CX0)];
        float v = t.e[i];
        int reps = t.fl[i] ? 1 : 0;
        if (fs && reps) {
            const uint32_t raw = (uint32_t)(y * w + x);
            if (last1 != 0 && raw == last1 - 1 && (cnt & 7u) != 0) reps = 1 + (int)(8u - (cnt & 7u));
        }
        for (int k = 0; k < reps; ++k) v = v + pc;
        t.dg[i] = v;
    }
    __syncthreads();
    // first colour (black): own and neighbours prolongated
    for (int i = threadIdx.x; i < G::kN; i += blockDim.x) {
        const int lx = i % kR, ly = i / kR;
        if (lx < 1 || lx >= kR - 1 || ly < 1 || ly >= kR - 1) continue;
        const int32_t x = X0 - kVH + lx, y = Y0 - kVH + ly;
        if (((x + y) & 1) == 0) continue;
        const float v = vsor<T>(t, i, t.dg[i], t.dg[i + 1], t.dg[i + kR], t.dg[i - 1], t.dg[i - kR]);
        if (t.fl[i]) t.e[i] = v;
    }
    __syncthreads();
    // second colour (red) of the tile: own prolongated (the first colour left
    // red cells as they were), neighbours final
    float prod = 0.0f;
    double dprod = 0.0, dasz = 0.0;
#pragma unroll
    for (int q = 0; q < G::kTPer; ++q) {
        const int i = (int)threadIdx.x + NT * q;
        const int32_t x = X0 + i % T, y = Y0 + i / T;
        if (i >= T * T || x >= w || y >= h) continue;
        const int li = (y - Y0 + kVH) * kR + (x - X0 + kVH);
        float e = t.e[li];
        if (((x + y) & 1) == 0) {
            const float v = vsor<T>(t, li, t.dg[li], t.e[li + 1], t.e[li + kR], t.e[li - 1], t.e[li - kR]);
            if (t.fl[li]) e = v;
        }
        const size_t pc = (size_t)(y + 1) * w + x;
        e_out[pc] = e;
        if (r_home && t.fl[li]) r_home[pc] = t.r[li];
        if (dot == 2) {
            dprod += (double)t.r[li] * (double)e;
            if (t.fl[li]) dasz += (double)asv[q] * (double)e;
        } else if (dot == 1) {
            prod += t.r[li] * e;
        }
    }
    if (dot == 2) {
        block_sum2(dprod, dasz, dscratch);  // (each as block_sum alone)
        if (threadIdx.x == 0) {
            part_gamma_d[wg] = dprod;
            part_asz_d[wg] = dasz;
        }
    } else if (dot == 1) {
        const float bs = block_sum(prod, scratch);
        if (threadIdx.x == 0) part_gamma[wg] = bs;
    }
}

struct VUpArgs {
    LevelView L, C;
    const FluidStats* fs;  // base level only
    const int32_t* done;
    uint32_t tiles_x;
    int dot;
    float* part_gamma;
    double* part_gamma_d;
    const float* As;      // relative mode: A s of the last iteration (its FP32 rounding) and
    double* part_asz_d;   // the partials of (A s).z
    int spec;             // load the region with the flags (levels of few tiles)
    float* e_out;         // the level's e; L.e is its down-sweep buffer
    float* r_home;        // base level after VDownAx (L.r its second buffer): the residual's own buffer
    const int32_t* r_in_alt;
    const uint32_t* order;  // base level of a simulation stepped alone (else null): TileOrder's tile order
};

template <int T, int NT = 256>
struct VUp {
    static constexpr int kThreads = NT;
    using Args = VUpArgs;
    static constexpr bool kDealRows = true;  // its tile rows are dealt to the XCDs in turn
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        uint32_t tile;
        if (!ordered_tile(a.order, wg, tile)) return;
        vup<T, NT>(a.L, a.C, a.fs, a.done, a.tiles_x, a.dot, a.part_gamma, a.part_gamma_d, a.As, a.part_asz_d, a.spec,
               a.e_out, a.r_home, a.r_in_alt, tile);
    }
};



unsigned blocks_for(size_t n, unsigned threads = kThreads) { return (unsigned)((n + threads - 1) / threads); }

LevelSet level_set(const Sim& sim) {
    LevelSet ls{};
    ls.count = (int)sim.level_count;
    for (int l = 0; l < ls.count && l < 16; ++l) ls.lv[l] = sim.levels[l];
    return ls;
}

// The LDS bottom's level table: each level's array offsets in the image
// (per level diag, inv_diag, plus_x, plus_y, r, e over the padded range; the
// flag bytes apart), which is also the layout of the levels' arrays in
// global memory from the first bottom level on.
BottomLdsArgs bottom_lds_args(const Sim& sim, int first, const int32_t* done) {
    BottomLdsArgs a{};
    const int last = (int)sim.level_count - 1;
    uint32_t total = 0, ftotal = 0;
    for (int l = first; l <= last; ++l) {
        const LevelView& G = sim.levels[l];
        const uint32_t p = (uint32_t)bottom_padded(G);
        const uint32_t wmul = G.w > 1 ? (uint32_t)(((1ull << 32) + G.w - 1) / G.w) : 0u;
        a.tab[l] = {(int32_t)G.w, (int32_t)G.h, (int32_t)G.cells, total, total + p, total + 2 * p, total + 3 * p,
                    total + 4 * p, total + 5 * p, ftotal, wmul};
        total += 6 * p;
        ftotal += (uint32_t)bottom_flags(G);
    }
    a.img = sim.levels[first].diag;
    a.fimg = reinterpret_cast<const uint32_t*>(sim.levels[first].fluid);
    a.total = total;
    a.fwords = ftotal / 4;
    a.cg = a.tab[last].e + (uint32_t)bottom_padded(sim.levels[last]);
    a.first = first;
    a.last = last;
    a.done = done;
    a.rel = (int)(sim.solver.mode == 1);
    return a;
}

Prolong base_prolong(const Sim& sim) {
    Prolong P{};
    P.has = sim.level_count > 1;
    if (P.has) P.C = sim.levels[1];
    P.fs = sim.fstats;
    return P;
}

}  // namespace

void launch_build_pressure(Sim& sim, bool zero_pressure, bool defer_fstats) {
    const uint32_t cells = sim.cells;
    const uint8_t* fluid = sim.levels[0].fluid;
    // the three smoothing passes, density -> smooth_a, in one launch
    const uint32_t stx = (sim.W + kSmT - 1) / kSmT, sty = (sim.H + kSmT - 1) / kSmT;
    emit<Smooth3>(sim, {sim.density, sim.smooth_a, fluid, sim.W, sim.H, stx}, stx * sty);
    Assemble::Args a{sim.mac_u, sim.mac_v, sim.smooth_a, fluid, sim.obstacle, sim.W, cells, sim.cell, sim.levels[0],
                     sim.pressure, (int)zero_pressure, sim.part_fluid, sim.nred};
    emit<Assemble>(sim, a, sim.nred);
    // (the hierarchy's bottom build finishes the statistics when it follows)
    if (defer_fstats) sim.fstats_pending = true;
    else emit<FstatsFinalize>(sim, {sim.part_fluid, sim.nred, sim.fstats, sim.h_fstats_d}, 1);
}

void launch_fluid_stats(Sim& sim) {
    hipLaunchKernelGGL(fluid_stats_kernel, dim3(sim.nred), dim3(kThreads), 0, sim.stream, sim.levels[0].fluid,
                       sim.cells, sim.part_fluid);
    emit<FstatsFinalize>(sim, {sim.part_fluid, sim.nred, sim.fstats, sim.h_fstats_d}, 1);
}



// The base level's up-sweep writes one partial of r.z per tile.
size_t gamma_partials(const Sim& sim) {
    const LevelView& B = sim.levels[0];
    return (size_t)((B.w + kBaseTile - 1) / kBaseTile) * ((B.h + kBaseTile - 1) / kBaseTile);
}

int pcg_cap(const Sim& sim) { return sim.solver.mode == 1 ? sim.solver.max_iterations : kMaxPcgIterations; }

namespace {

// 16 x 16 tiles on a level above the base with at most tuning.few_tiles
// 32 x 32 ones (the base keeps one partial of r.z per 32 x 32 tile)
uint32_t tiles_of(const LevelView& L, uint32_t side, uint32_t& tx) {
    tx = (L.w + side - 1) / side;
    return tx * ((L.h + side - 1) / side);
}

bool few_tiles(const Sim& sim, int l) {
    uint32_t tx;
    return l > 0 && tiles_of(sim.levels[l], kBaseTile, tx) <= sim.tuning.few_tiles;
}

// Two levels of few tiles above the bottom (not the base), l and l + 1, in
// one launch each way when the simulation steps alone, where launches are
// latency-bound.  A batch's launches are bound by throughput, which the
// pair's recomputed ring costs more than the launch saves; the pair computes
// the bits of the two launches, so either path gives the same result.
bool level_pair(const Sim& sim, int l) {
    return !sim.rec && l >= 1 && l + 1 < (int)sim.bottom_level && few_tiles(sim, l);
}

// The relative mode's update of iteration k fused into the base level's
// down-sweep (VDownAx, one launch fewer): a simulation stepped alone whose
// level pair (1, 2) follows the base (it takes over the residual test), on
// the iterations that do not recompute the true residual.
bool fused_update(const Sim& sim, int k) {
    const int every = sim.solver.recompute_every;
    const bool due = every > 0 && (k + 1) % every == 0;
    return sim.solver.mode == 1 && !due && level_pair(sim, 1);
}

// The base tiles' thread count (see kWideMinRed): four tiles a CU are then 32
// waves.  AxpyTile takes the same count, so that its partials are VDownAx's.
bool wide_base(const Sim& sim) { return sim.nred >= kWideMinRed; }

// Whether the base-level V-cycle kernels of this simulation run their tiles
// in TileOrder's order: a simulation stepped alone whose base level takes the
// wide kernels and whose level 1 MgBuild builds.
bool ordered_base(const Sim& sim) {
    return !sim.rec && wide_base(sim) && 1 < (int)sim.bottom_level && (uint32_t)gamma_partials(sim) <= kOrderMaxTiles;
}

// The update's operands (AxpyTile, VDownAx), r_out receiving the residual.
AxIn update_operands(const Sim& sim, float* r_out) {
    return AxIn{{sim.search[0], sim.search[1]}, sim.As, sim.As_lo, sim.pressure, sim.pressure_d, sim.part_delta_d, sim.nred,
                r_out, sim.r_lo, sim.part_rr_t};
}

}  // namespace

// Before the solve: the state, and in the relative mode b = r (kept for the
// residual recomputation) and ||b||^2.
void launch_mg_build(Sim& sim) {
    const int n = (int)sim.level_count;
    const bool ordered = ordered_base(sim);
    uint32_t btx = 0;
    const uint32_t nt = tiles_of(sim.levels[0], kBaseTile, btx);
    int l = 1;
    for (; l < n && l < (int)sim.bottom_level; ++l)
        emit<MgBuild>(sim, {sim.levels[l - 1], sim.levels[l], (ordered && l == 1) ? sim.tile_flag : nullptr, btx},
                      blocks_for(sim.levels[l].cells));
    const uint32_t cmul = btx > 1 ? (uint32_t)(((1ull << 32) + btx - 1) / btx) : 0u;
    const TileOrderArgs order{sim.tile_flag, sim.tile_order, nt, btx, cmul, sim.part_rr_t, sim.part_gamma_d,
                              sim.part_asz_d, sim.part_gamma};
    const FstatsArgs fst{sim.part_fluid, sim.nred, sim.fstats, sim.h_fstats_d};
    const bool fstats = sim.fstats_pending;
    sim.fstats_pending = false;
    if (l < n) {
        emit<MgBuildBottom>(sim, {level_set(sim), l, (int)fstats, fst, (int)ordered, order}, 1);
    } else {
        if (fstats) emit<FstatsFinalize>(sim, fst, 1);
        if (ordered) emit<TileOrder>(sim, order, 1);
    }
    sim.tile_order_valid = ordered;
}

void launch_pcg_reset(Sim& sim) {
    const bool rel = sim.solver.mode == 1;
    const double tol = (double)sim.solver.rel_tol;
    const PcgReset::Args reset{sim.pcg, tol * tol, pcg_cap(sim)};
    if (rel) {
        emit<PcgBnorm>(sim, {sim.levels[0], sim.rhs, sim.r_lo, sim.part_rr_d, sim.pressure, sim.pressure_d, sim.nred}, sim.nred);
        emit<PcgBnormFinalize>(sim, {reset, sim.part_rr_d, sim.nred}, 1);
    } else {
        emit<PcgReset>(sim, reset, 1);
    }
}

void launch_zero_pressure(Sim& sim) {
    hipLaunchKernelGGL(zero_pressure_kernel, dim3(blocks_for(sim.cells)), dim3(kThreads), 0, sim.stream,
                       sim.levels[0].fluid, sim.pressure, sim.cells);
}

// mode 0: apply_preconditioner alone; 1: PCG start (gamma); 2: PCG iteration
// (residual test in the first kernel, gamma' partials in the last).
void launch_vcycle(Sim& sim, int mode, int k) {
    const int32_t* done = &sim.pcg->done;
    const int last = (int)sim.level_count - 1;
    const int bottom = (int)sim.bottom_level;  // 1 <= bottom <= last
    const bool rel = sim.solver.mode == 1 && mode != 0;
    auto tiles = tiles_of;
    auto small = [&](int l) { return few_tiles(sim, l); };
    // The base level of a large grid runs kWideTileThreads threads per tile.
    // The FP32 partials of r.z under the reference's rules (dot 1) keep 256
    // threads and their summation order.
    const bool wide = wide_base(sim);
    const bool fused = mode == 2 && fused_update(sim, k);
    const uint32_t ntiles = (uint32_t)gamma_partials(sim);
    const AxIn ax = update_operands(sim, fused ? sim.r_alt : sim.levels[0].r);
    // TileOrder's order for the wide base kernels (built with the hierarchy
    // of this substep; not after fields were written since)
    const uint32_t* base_order = (sim.tile_order_valid && ordered_base(sim)) ? sim.tile_order : nullptr;
    // A level above the bottom with its down-sweep buffer as e: the way down
    // reads and writes it, the way up reads it and writes the level's e
    // (the bottom's levels keep theirs: one workgroup runs them)
    auto down = [&](int l) {
        LevelView v = sim.levels[l];
        if (l < bottom) v.e = sim.down_e[l];
        return v;
    };
    for (int l = 0; l < bottom; ++l) {
        if (level_pair(sim, l)) {
            uint32_t tx;
            const uint32_t nt = tiles(sim.levels[l + 1], kP2T, tx);
            const int fuse_red_d = l + 2 < bottom ? 1 : (int)(bottom < last);
            const bool test = fused && l == 1;  // the fused update's residual test
            emit<VDown2>(sim, VDown2Args{down(l), down(l + 1), down(l + 2), fuse_red_d, done, tx, sim.pcg,
                                         test ? sim.part_rr_t : nullptr, ntiles},
                         nt, k);
            ++l;
            continue;
        }
        const LevelView F = down(l);
        const int fuse_red = l + 1 < bottom ? 1 : (int)(bottom < last);
        const bool sm = small(l);
        uint32_t tx;
        const uint32_t nt = tiles(F, sm ? 16 : kBaseTile, tx);
        const bool rel_check = l == 0 && rel;
        const bool ax_here = l == 0 && fused;
        VDownArgs a{F, down(l + 1), (int)(l == 0), fuse_red, (int)(l > 0), done, tx, sim.pcg,
                          (int)(l == 0 && mode == 2 && !fused), sim.props.solve_error, sim.part_max,
                          rel_check ? sim.part_rr_t : nullptr, ntiles, rel_check ? sim.part_rr2_d : nullptr,
                          rel_check ? sim.solver.recompute_every : 0, sim.nred, ax};
        if (l == 0 && wide) a.order = base_order;
        if (ax_here && wide) emit<VDownAx<kWideTileThreads>>(sim, a, nt, k);
        else if (ax_here) emit<VDownAx<kSolverThreads>>(sim, a, nt, k);
        else if (sm) emit<VDown<16>>(sim, a, nt, k);
        else if (l == 0 && wide) emit<VDown<kBaseTile, kWideTileThreads>>(sim, a, nt, k);
        else emit<VDown<kBaseTile>>(sim, a, nt, k);
    }
    if (sim.bottom_in_lds) {
        emit<BottomLds>(sim, bottom_lds_args(sim, bottom, done), 1);
    } else {
        emit<BottomGlobal>(sim, {level_set(sim), bottom, sim.cg_s, sim.cg_ap, sim.cg_r, done,
                                 (int)(sim.solver.mode == 1)},
                           1);
    }
    for (int l = bottom - 1; l >= 0; --l) {
        // levels l - 1 and l in one launch, as on the way down
        if (level_pair(sim, l - 1)) {
            uint32_t tx;
            const uint32_t nt = tiles(sim.levels[l], kP2T, tx);
            emit<VUp2>(sim, VUp2Args{down(l - 1), down(l), sim.levels[l + 1], sim.levels[l - 1].e, sim.levels[l].e,
                                     done, tx},
                       nt);
            --l;
            continue;
        }
        LevelView L = down(l);
        const bool home = l == 0 && fused;  // the residual is in the second buffer
        if (home) L.r = sim.r_alt;
        const bool sm = small(l);
        uint32_t tx;
        const uint32_t nt = tiles(L, sm ? 16 : kBaseTile, tx);
        const int dot = (l == 0 && mode != 0) ? (rel ? 2 : 1) : 0;  // one partial of r.z per tile
        VUpArgs a{L, sim.levels[l + 1], l == 0 ? (const FluidStats*)sim.fstats : nullptr, done, tx, dot,
                        sim.part_gamma, sim.part_gamma_d, sim.As, sim.part_asz_d, (int)(l > 0), sim.levels[l].e,
                        home ? sim.levels[0].r : nullptr, &sim.pcg->r_in_alt};
        if (sm) emit<VUp<16>>(sim, a, nt);
        else if (l == 0 && wide && dot != 1) {
            a.order = base_order;
            emit<VUp<kBaseTile, kWideTileThreads>>(sim, a, nt);
        }
        else emit<VUp<kBaseTile>>(sim, a, nt);
    }
}

void launch_pcg_iteration_head(Sim& sim, int k) {
    const LevelView& B = sim.levels[0];
    const bool rel = sim.solver.mode == 1;
    // the kernel timer samples iteration 1 of each solve (the first with
    // every operand; events around every launch would add their own cost
    // to the step they time)
    const bool timed = k == 1;
    if (timed) sim.mark(APIC_T_K_SPMV, 0);
    SpmvDotArgs a{B, {sim.search[0], sim.search[1]}, sim.As, rel ? sim.As_lo : nullptr, sim.pcg, sim.part_gamma, sim.part_delta,
                  rel ? sim.part_gamma_d : nullptr, sim.part_asz_d, rel ? sim.part_delta_d : nullptr,
                  (int)gamma_partials(sim), sim.nred};
    const bool wide = rel && sim.nred >= kWideMinRed;
    if (wide) emit<SpmvDot<kWideThreads>>(sim, a, sim.nred, k);
    else emit<SpmvDot<kSolverThreads>>(sim, a, sim.nred, k);
    if (timed) sim.mark(APIC_T_K_SPMV, 1);
    if (rel) {
        // the update, unless the base down-sweep fuses it
        if (!fused_update(sim, k)) {
            uint32_t tx;
            const uint32_t nt = tiles_of(B, kBaseTile, tx);
            AxpyTileArgs b{B, update_operands(sim, B.r), sim.pcg, tx};
            if (wide_base(sim)) emit<AxpyTile<kWideTileThreads>>(sim, b, nt, k);
            else emit<AxpyTile<kSolverThreads>>(sim, b, nt, k);
        }
    } else {
        AxpyMaxArgs b{B, {sim.search[0], sim.search[1]}, sim.As, sim.pressure, sim.pcg, sim.part_delta, sim.part_max,
                      sim.nred};
        emit<AxpyMax<kSolverThreads>>(sim, b, sim.nred, k);
    }
    // the true residual, on the iterations that recompute it (a batch records
    // one iteration for all k: there the kernel itself skips the others)
    const int every = sim.solver.recompute_every;
    if (rel && (sim.rec || (every > 0 && (k + 1) % every == 0))) {
        TrueResidual::Args c{B, sim.rhs, sim.r_lo, sim.pressure_d, sim.pcg, sim.part_rr2_d, every, sim.nred};
        emit<TrueResidual>(sim, c, sim.nred, k);
    }
}

void launch_pcg_decide(Sim& sim, int k) {
    emit<Decide>(sim, {sim.pcg, sim.part_gamma, sim.h_pcg_d, sim.solver.mode == 1 ? sim.part_gamma_d : nullptr,
                       (int)gamma_partials(sim)},
                 1, k);
}

}  // namespace apic
