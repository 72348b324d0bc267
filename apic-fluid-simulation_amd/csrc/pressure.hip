// pressure.hip -- pressure system, multigrid hierarchy and MGPCG solve.
//
//  smooth_density_kernel   the three density-smoothing passes of
//                          build_pressure_system (src/apic.rs:1649-1685)
//  assemble_kernel         ghost-fluid Laplacian + RHS (:1698-1784)
//  mg_build_kernel         one coarse level of update_multilevel_grid
//                          (:1499-1602), dense over coarse cells
//  V-cycle (apply_preconditioner, :1788-1821) with red-black SOR, omega=1.2:
//    red_zero_kernel       zero e + first red sweep (e starts at 0, so the
//                          sweep reduces to e = omega*(r*inv_diag))
//    sweep_kernel          one colour of smooth_level (:471-535), optionally
//                          fused with prolongate (:596-616) on the way up
//    restrict_kernel       restrict (:537-594), fused with the next level's
//                          first red sweep
//    bottom_kernel         one workgroup runs every level with <= 4096 cells
//                          and the coarsest-level CG (:618-704)
//  PCG (solve_pcg, :1865-1956), reference rules:
//    spmv_dot_kernel       s = z + beta*s fused with A s and s.As
//    axpy_max_kernel       p += alpha s, r -= alpha As, max |r|
//    the base level's final red sweep also computes r.z (gamma)
//  Scalars live in a PcgState in device memory.  Each reduction finishes in
//  the last workgroup to arrive, which sums the per-workgroup partials in a
//  fixed order, decides the reference's exit tests and sets `done`; every
//  kernel of the loop returns at once when `done` is set, so the host can
//  queue several iterations between synchronisations.
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
constexpr int kBottomThreads = 1024;
constexpr int kSerialCap = 4096;
constexpr int kMaxIterations = 100;

__device__ inline float atomic_load_f(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- assembly

__global__ void smooth_density_kernel(const float* __restrict__ in, float* __restrict__ out,
                                      const uint8_t* __restrict__ fluid, uint32_t W, uint32_t cells) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    if (!fluid[c]) { out[c] = in[c]; return; }
    float sum = in[c] * 4.0f;
    float weight = 4.0f;
    float fl = (float)fluid[c - 1], fr = (float)fluid[c + 1], fb = (float)fluid[c - W], ft = (float)fluid[c + W];
    sum += in[c - 1] * fl;
    sum += in[c + 1] * fr;
    sum += in[c - W] * fb;
    sum += in[c + W] * ft;
    weight += ((fl + fr) + fb) + ft;
    out[c] = sum / weight;
}

__global__ void assemble_kernel(const float* __restrict__ mac_u, const float* __restrict__ mac_v,
                                const float* __restrict__ smoothed, const uint8_t* __restrict__ fluid,
                                const uint8_t* __restrict__ obstacle, uint32_t W, uint32_t cells, float cell,
                                LevelView L, float* __restrict__ pressure, int zero_pressure, FluidStats* fstats) {
    __shared__ uint32_t s_cnt, s_last;
    if (threadIdx.x == 0) { s_cnt = 0; s_last = 0; }
    __syncthreads();
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cells) {
        size_t pc = (size_t)c + W;
        if (fluid[c]) {
            atomicAdd(&s_cnt, 1u);
            atomicMax(&s_last, c + 1);
            uint32_t gx = c % W, gy = c / W;
            size_t ui = (size_t)gy * (W + 1) + gx;
            float ul = mac_u[ui], ur = mac_u[ui + 1], vb = mac_v[c], vt = mac_v[c + W];
            int nol = 1 - obstacle[c - 1], nor = 1 - obstacle[c + 1], nob = 1 - obstacle[c - W], not_ = 1 - obstacle[c + W];
            ul *= (float)nol; ur *= (float)nor; vb *= (float)nob; vt *= (float)not_;
            float div = (ur - ul) + (vt - vb);
            float rc = smoothed[c];
            const float target = 9.0f, noise = target * 1.15f, rate = 0.1f;
            float excess = fminf(fmaxf(rc - noise, 0.0f), target);
            int fl = fluid[c - 1], fr = fluid[c + 1], fb = fluid[c - W], ft = fluid[c + W];
            float diag = 0.0f, px = 0.0f, py = 0.0f;
            if (nol == 1) diag += fl ? 1.0f : 1.0f / theta(smoothed[c - 1], rc);
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
    __syncthreads();
    if (threadIdx.x == 0 && s_cnt) {
        atomicAdd(&fstats->count, s_cnt);
        // last_raw holds raw+1 while accumulating (0 = none); finalized on read
        atomicMax(&fstats->last_raw, s_last);
    }
}

__global__ void fluid_stats_kernel(const uint8_t* __restrict__ fluid, uint32_t cells, FluidStats* fstats) {
    __shared__ uint32_t s_cnt, s_last;
    if (threadIdx.x == 0) { s_cnt = 0; s_last = 0; }
    __syncthreads();
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cells && fluid[c]) { atomicAdd(&s_cnt, 1u); atomicMax(&s_last, c + 1); }
    __syncthreads();
    if (threadIdx.x == 0 && s_cnt) { atomicAdd(&fstats->count, s_cnt); atomicMax(&fstats->last_raw, s_last); }
}

__global__ void reset_fstats_kernel(FluidStats* fstats) { fstats->count = 0; fstats->last_raw = 0; }

__device__ inline void mg_build_cell(const LevelView& F, const LevelView& C, uint32_t c) {
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
}

__global__ void mg_build_kernel(LevelView F, LevelView C) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < C.cells) mg_build_cell(F, C, c);
}

struct LevelSet {
    LevelView lv[16];
    int count;  // number of levels in lv (lv[0] = base)
};

__global__ __launch_bounds__(kBottomThreads) void mg_build_bottom_kernel(LevelSet ls, int first) {
    for (int l = first; l < ls.count; ++l) {
        const LevelView F = ls.lv[l - 1];
        const LevelView C = ls.lv[l];
        for (uint32_t c = threadIdx.x; c < C.cells; c += blockDim.x) mg_build_cell(F, C, c);
        __syncthreads();
    }
}

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

// cell of colour `color` addressed by thread t (half the cells per colour)
__device__ inline bool colour_cell(const LevelView& L, uint32_t t, int color, uint32_t& c) {
    uint32_t hw = (L.w + 1) / 2;
    uint32_t y = t / hw, k = t % hw;
    if (y >= L.h) return false;
    uint32_t x = 2 * k + ((y + (uint32_t)color) & 1u);
    if (x >= L.w) return false;
    c = y * L.w + x;
    return true;
}

__global__ void red_zero_kernel(LevelView L, const int32_t* __restrict__ done) {
    if (*done) return;
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < L.cells) red_zero_cell(L, c);
}

__global__ void sweep_kernel(LevelView L, Prolong P, int color, int mode, const int32_t* __restrict__ done) {
    if (*done) return;
    uint32_t c;
    if (colour_cell(L, blockIdx.x * blockDim.x + threadIdx.x, color, c)) sor_cell(L, P, c, mode);
}

__global__ void restrict_kernel(LevelView F, LevelView C, int fuse_red, const int32_t* __restrict__ done) {
    if (*done) return;
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < C.cells) restrict_cell(F, C, c, fuse_red != 0);
}

// ---- block-level reductions (fixed order, deterministic)

template <typename T>
__device__ inline T block_sum(T v, T* scratch) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    T s = scratch[0];
    for (int w = 1; w < nw; ++w) s += scratch[w];
    return s;
}

__device__ inline float block_max(float v, float* scratch) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fast_max(v, __shfl_xor(v, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    float s = scratch[0];
    for (int w = 1; w < nw; ++w) s = fast_max(s, scratch[w]);
    return s;
}

// Last-workgroup finish: returns true in the workgroup that arrives last,
// after which the partials of every workgroup are visible.
__device__ inline bool arrive_last(uint32_t* ticket) {
    __shared__ bool s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        uint32_t t = atomicAdd(ticket, 1u);
        s_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (s_last) __threadfence();
    return s_last;
}

// Sum of partials[0..n) in a fixed order by one workgroup.
__device__ inline float sum_partials(const float* partials, uint32_t n, float* scratch) {
    float v = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) v += atomic_load_f(&partials[i]);
    return block_sum(v, scratch);
}
__device__ inline float max_partials(const float* partials, uint32_t n, float* scratch) {
    float v = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) v = fast_max(v, atomic_load_f(&partials[i]));
    return block_max(v, scratch);
}

// Base level, up sweep of the second colour (red), fused with r.z over the
// whole padded interior (dot_product over [W, W+W*H), src/apic.rs:1888,1936).
// mode 0: V-cycle only; 1: PCG start; 2: PCG iteration.
__global__ __launch_bounds__(kThreads) void base_red_up_dot_kernel(LevelView L, Prolong P, int mode,
                                                                    PcgState* st, float* __restrict__ partials) {
    __shared__ float scratch[kThreads / 64];
    if (st->done) return;
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    float prod = 0.0f;
    if (c < L.cells) {
        uint32_t x = c % L.w, y = c / L.w;
        if (((x + y) & 1u) == 0) sor_cell(L, P, c, 2);
        size_t pc = (size_t)c + L.w;
        prod = L.r[pc] * L.e[pc];
    }
    if (mode == 0) return;
    float bs = block_sum(prod, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
    if (!arrive_last(&st->ticket[2])) return;
    float g = sum_partials(partials, gridDim.x, scratch);
    if (threadIdx.x == 0) {
        st->ticket[2] = 0;
        if (mode == 1) {
            if (g < kSafety) { st->done = 1; st->exit_reason = 1; }
            else { st->gamma = g; st->first = 1; }
        } else {
            if (g < kSafety) { st->done = 1; st->exit_reason = 4; }
            else {
                st->beta = g / st->gamma;
                st->gamma = g;
                st->first = 0;
                if (st->iters >= kMaxIterations) { st->done = 1; st->exit_reason = 5; }
            }
        }
    }
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
    // sequential sums in ascending cell order, as the reference's serial loops
    auto serial_sum = [&](auto term) -> float {
        if (threadIdx.x == 0) {
            float s = 0.0f;
            for (uint32_t c = 0; c < n; ++c)
                if (L.fluid[c]) s += term(c);
            s_bcast[0] = s;
        }
        __syncthreads();
        float r = s_bcast[0];
        __syncthreads();
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

// Levels [first, count) of the V-cycle inside one workgroup.  On entry
// level `first` holds r (restricted) and e after its fused first red sweep
// (or e = 0 when it is the coarsest level).  On exit level `first` holds its
// final post-smoothed error.
__global__ __launch_bounds__(kBottomThreads) void bottom_kernel(LevelSet ls, int first, float* cg_s, float* cg_ap,
                                                               float* cg_r, const int32_t* __restrict__ done) {
    __shared__ float sprod[kSerialCap];
    __shared__ float scratch[kBottomThreads / 64];
    if (*done) return;
    const int last = ls.count - 1;
    Prolong none{};
    // down
    for (int l = first; l < last; ++l) {
        const LevelView& L = ls.lv[l];
        if (l > first) {
            for (uint32_t c = threadIdx.x; c < L.cells; c += blockDim.x) restrict_cell(ls.lv[l - 1], L, c, true);
            __syncthreads();
        }
        for (uint32_t c = threadIdx.x; c < L.cells; c += blockDim.x) {
            uint32_t x = c % L.w, y = c / L.w;
            if (((x + y) & 1u) == 1) sor_cell(L, none, c, 0);
        }
        __syncthreads();
    }
    if (last > first) {
        for (uint32_t c = threadIdx.x; c < ls.lv[last].cells; c += blockDim.x)
            restrict_cell(ls.lv[last - 1], ls.lv[last], c, false);
        __syncthreads();
    }
    coarse_cg_block(ls.lv[last], cg_s, cg_ap, cg_r, sprod, scratch);
    __syncthreads();
    // up
    for (int l = last - 1; l >= first; --l) {
        const LevelView& L = ls.lv[l];
        Prolong P{};
        P.C = ls.lv[l + 1];
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

// ---------------------------------------------------------------------- PCG

__global__ void pcg_reset_kernel(PcgState* st) {
    st->gamma = st->delta = st->alpha = st->beta = st->maxerr = 0.0f;
    st->done = 0;
    st->iters = 0;
    st->exit_reason = 0;
    st->first = 1;
    for (int k = 0; k < 4; ++k) st->ticket[k] = 0;
    st->dgamma = st->ddelta = st->bnorm2 = st->rnorm2 = 0.0;
}

__global__ void zero_pressure_kernel(const uint8_t* __restrict__ fluid, float* __restrict__ pressure,
                                     uint32_t cells) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cells && fluid[c]) pressure[c] = 0.0f;
}

// s = z + beta*s (s = z on the first iteration), As = A s, delta = s.As
// (apply_matrix_and_dot, src/apic.rs:1823-1863; search update :1945-1954)
__global__ __launch_bounds__(kThreads) void spmv_dot_kernel(LevelView L, const float* __restrict__ s_old,
                                                             float* __restrict__ s_new, float* __restrict__ As,
                                                             PcgState* st, float* __restrict__ partials) {
    __shared__ float scratch[kThreads / 64];
    if (st->done) return;
    const int first = st->first;
    const float beta = st->beta;
    const size_t w = L.w;
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    float prod = 0.0f;
    if (c < L.cells && L.fluid[c]) {
        size_t pc = (size_t)c + w;
        auto sv = [&](size_t i) { return first ? L.e[i] : L.e[i] + (beta * s_old[i]); };
        float s0 = sv(pc);
        float v = L.diag[pc] * s0;
        v = fmaf(L.px[pc], sv(pc + 1), v);
        v = fmaf(L.px[pc - 1], sv(pc - 1), v);
        v = fmaf(L.py[pc], sv(pc + w), v);
        v = fmaf(L.py[pc - w], sv(pc - w), v);
        s_new[pc] = s0;
        As[pc] = v;
        prod = v * s0;
    }
    float bs = block_sum(prod, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
    if (!arrive_last(&st->ticket[0])) return;
    float d = sum_partials(partials, gridDim.x, scratch);
    if (threadIdx.x == 0) {
        st->ticket[0] = 0;
        st->iters += 1;
        st->delta = d;
        if (fabsf(d) < kSafety) { st->done = 1; st->exit_reason = 2; }
        else st->alpha = st->gamma / d;
    }
}

// p += alpha*s; r = fma(-alpha, As, r); max |r| (src/apic.rs:1907-1932)
__global__ __launch_bounds__(kThreads) void axpy_max_kernel(LevelView L, const float* __restrict__ s,
                                                             const float* __restrict__ As,
                                                             float* __restrict__ pressure, float solve_error,
                                                             PcgState* st, float* __restrict__ partials) {
    __shared__ float scratch[kThreads / 64];
    if (st->done) return;
    const float alpha = st->alpha;
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    if (c < L.cells && L.fluid[c]) {
        size_t pc = (size_t)c + L.w;
        pressure[c] += alpha * s[pc];
        float nr = fmaf(-alpha, As[pc], L.r[pc]);
        L.r[pc] = nr;
        m = fabsf(nr);
    }
    float bm = block_max(m, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = bm;
    if (!arrive_last(&st->ticket[1])) return;
    float mx = max_partials(partials, gridDim.x, scratch);
    if (threadIdx.x == 0) {
        st->ticket[1] = 0;
        st->maxerr = mx;
        if (mx < solve_error) { st->done = 1; st->exit_reason = 3; }
    }
}

unsigned blocks_for(size_t n, unsigned threads = kThreads) { return (unsigned)((n + threads - 1) / threads); }

LevelSet level_set(const Sim& sim) {
    LevelSet ls{};
    ls.count = (int)sim.levels.size();
    for (int l = 0; l < ls.count && l < 16; ++l) ls.lv[l] = sim.levels[l];
    return ls;
}

Prolong base_prolong(const Sim& sim) {
    Prolong P{};
    P.has = sim.levels.size() > 1;
    if (P.has) P.C = sim.levels[1];
    P.fs = sim.fstats;
    return P;
}

}  // namespace

void launch_build_pressure(Sim& sim, bool zero_pressure) {
    const uint32_t cells = sim.cells;
    const uint8_t* fluid = sim.levels[0].fluid;
    hipLaunchKernelGGL(reset_fstats_kernel, dim3(1), dim3(1), 0, sim.stream, sim.fstats);
    // three smoothing passes: density -> b -> a -> b ... ending in smooth_a
    hipLaunchKernelGGL(smooth_density_kernel, dim3(blocks_for(cells)), dim3(kThreads), 0, sim.stream, sim.density,
                       sim.smooth_a, fluid, sim.W, cells);
    hipLaunchKernelGGL(smooth_density_kernel, dim3(blocks_for(cells)), dim3(kThreads), 0, sim.stream, sim.smooth_a,
                       sim.smooth_b, fluid, sim.W, cells);
    hipLaunchKernelGGL(smooth_density_kernel, dim3(blocks_for(cells)), dim3(kThreads), 0, sim.stream, sim.smooth_b,
                       sim.smooth_a, fluid, sim.W, cells);
    hipLaunchKernelGGL(assemble_kernel, dim3(blocks_for(cells)), dim3(kThreads), 0, sim.stream, sim.mac_u, sim.mac_v,
                       sim.smooth_a, fluid, sim.obstacle, sim.W, cells, sim.cell, sim.levels[0], sim.pressure,
                       (int)zero_pressure, sim.fstats);
    (void)hipMemcpyAsync(sim.h_fstats, sim.fstats, sizeof(FluidStats), hipMemcpyDeviceToHost, sim.stream);
}

void launch_fluid_stats(Sim& sim) {
    hipLaunchKernelGGL(reset_fstats_kernel, dim3(1), dim3(1), 0, sim.stream, sim.fstats);
    hipLaunchKernelGGL(fluid_stats_kernel, dim3(blocks_for(sim.cells)), dim3(kThreads), 0, sim.stream,
                       sim.levels[0].fluid, sim.cells, sim.fstats);
    (void)hipMemcpyAsync(sim.h_fstats, sim.fstats, sizeof(FluidStats), hipMemcpyDeviceToHost, sim.stream);
}

void launch_mg_build(Sim& sim) {
    const int n = (int)sim.levels.size();
    int l = 1;
    for (; l < n && l < (int)sim.bottom_level; ++l)
        hipLaunchKernelGGL(mg_build_kernel, dim3(blocks_for(sim.levels[l].cells)), dim3(kThreads), 0, sim.stream,
                           sim.levels[l - 1], sim.levels[l]);
    if (l < n)
        hipLaunchKernelGGL(mg_build_bottom_kernel, dim3(1), dim3(kBottomThreads), 0, sim.stream, level_set(sim), l);
}

void launch_pcg_reset(Sim& sim) { hipLaunchKernelGGL(pcg_reset_kernel, dim3(1), dim3(1), 0, sim.stream, sim.pcg); }

void launch_zero_pressure(Sim& sim) {
    hipLaunchKernelGGL(zero_pressure_kernel, dim3(blocks_for(sim.cells)), dim3(kThreads), 0, sim.stream,
                       sim.levels[0].fluid, sim.pressure, sim.cells);
}

// mode 0: apply_preconditioner alone; 1: PCG start (gamma); 2: PCG iteration (gamma')
void launch_vcycle(Sim& sim, int mode) {
    const int32_t* done = &sim.pcg->done;
    const int n = (int)sim.levels.size();
    const LevelView& B = sim.levels[0];
    hipStream_t s = sim.stream;
    // down: base
    hipLaunchKernelGGL(red_zero_kernel, dim3(blocks_for(B.cells)), dim3(kThreads), 0, s, B, done);
    Prolong none{};
    auto half_blocks = [](const LevelView& L) { return blocks_for((size_t)((L.w + 1) / 2) * L.h); };
    hipLaunchKernelGGL(sweep_kernel, dim3(half_blocks(B)), dim3(kThreads), 0, s, B, none, 1, 0, done);
    if (n > 1) {
        const int last = n - 1;
        const int bottom = (int)sim.bottom_level;  // 1 <= bottom <= last
        for (int l = 1; l < bottom; ++l) {
            hipLaunchKernelGGL(restrict_kernel, dim3(blocks_for(sim.levels[l].cells)), dim3(kThreads), 0, s,
                               sim.levels[l - 1], sim.levels[l], 1, done);
            hipLaunchKernelGGL(sweep_kernel, dim3(half_blocks(sim.levels[l])), dim3(kThreads), 0, s, sim.levels[l],
                               none, 1, 0, done);
        }
        hipLaunchKernelGGL(restrict_kernel, dim3(blocks_for(sim.levels[bottom].cells)), dim3(kThreads), 0, s,
                           sim.levels[bottom - 1], sim.levels[bottom], (int)(bottom < last), done);
        hipLaunchKernelGGL(bottom_kernel, dim3(1), dim3(kBottomThreads), 0, s, level_set(sim), bottom, sim.cg_s,
                           sim.cg_ap, sim.cg_r, done);
        for (int l = bottom - 1; l >= 1; --l) {
            Prolong P{};
            P.C = sim.levels[l + 1]; P.has = 1;
            hipLaunchKernelGGL(sweep_kernel, dim3(half_blocks(sim.levels[l])), dim3(kThreads), 0, s, sim.levels[l], P,
                               1, 1, done);
            hipLaunchKernelGGL(sweep_kernel, dim3(half_blocks(sim.levels[l])), dim3(kThreads), 0, s, sim.levels[l], P,
                               0, 2, done);
        }
    }
    Prolong P = base_prolong(sim);
    hipLaunchKernelGGL(sweep_kernel, dim3(half_blocks(B)), dim3(kThreads), 0, s, B, P, 1, 1, done);
    hipLaunchKernelGGL(base_red_up_dot_kernel, dim3(blocks_for(B.cells)), dim3(kThreads), 0, s, B, P, mode, sim.pcg,
                       static_cast<float*>(sim.partials));
}

void launch_pcg_iteration_head(Sim& sim, int k) {
    const LevelView& B = sim.levels[0];
    float* s_old = sim.search[k & 1];
    float* s_new = sim.search[(k + 1) & 1];
    float* partials = static_cast<float*>(sim.partials);
    hipLaunchKernelGGL(spmv_dot_kernel, dim3(blocks_for(B.cells)), dim3(kThreads), 0, sim.stream, B, s_old, s_new,
                       sim.As, sim.pcg, partials);
    hipLaunchKernelGGL(axpy_max_kernel, dim3(blocks_for(B.cells)), dim3(kThreads), 0, sim.stream, B, s_new, sim.As,
                       sim.pressure, sim.props.solve_error, sim.pcg, partials);
}

}  // namespace apic
