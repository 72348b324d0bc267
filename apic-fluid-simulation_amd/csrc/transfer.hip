// transfer.hip -- particle<->grid transfers of the APIC step.
//
//  p2g_owner_kernel
//                  transfer_particles_to_grid (src/apic.rs:1218-1497) over
//                  particles grouped by cell: one thread per cell sums its own
//                  particles into register partials, each face and cell adds
//                  the partials of its neighbour cells, and the same kernel
//                  normalises the faces and applies the external forces
//                  (apply_external_forces, :1604-1626).  Deterministic, no
//                  atomics.  The cell grouping comes from the Morton sort
//                  (the particles themselves are sorted) or, when the sort is
//                  disabled, from a sorted index.
//  forces_kernel   apply_external_forces on its own (the stage entry point).
//  g2p_kernel      transfer_grid_to_particles_and_advect (:2057-2264): masked
//                  bilinear gather, jittered time step, wall clamp/reflect,
//                  affine C rebuild; fused with the next sort's keygen and the
//                  max |v| reduction that sets the next substep's dt
//                  (get_max_particle_velocity, :2412-2427).
#include "apic_internal.hpp"
#include "device_math.hpp"

namespace apic {

namespace {

// ---- cell-owner P2G -----------------------------------------------------------
//
// In cell order (the Morton-sorted particles, or the particles through a
// sorted index `order`), the particles of one cell are the contiguous range
// cell_range[key] = [begin, end).  A workgroup
// takes a kOwnWin x kOwnWin window of cells: a kOwnTile x kOwnTile tile and
// its one-cell ring.  Particles of the window are staged in LDS (coalesced
// loads), then one thread per window cell sums the contributions of its own
// particles into 33 register partials: the 2x3 u faces, 3x2 v faces and 3x3
// cells that a particle of that cell can reach (src/apic.rs:1237-1426, with
// u_index / v_index clamping, :1029-1041).  Every face and cell of the tile
// then adds the partials of its 6 (faces) or 9 (cells) source cells in a
// fixed order, so the result is deterministic and needs no atomics, and
// finishes the target in the same kernel: u = momentum / weight and the valid
// flag (:1428-1459), the fluid flag (:1461-1464) and, when asked,
// apply_external_forces (:1604-1626).  The faces x = W and y = H belong to the
// cells x = W-1 and y = H-1.

constexpr int kOwnWin = 16;
constexpr int kOwnTile = kOwnWin - 2;
constexpr int kOwnCells = kOwnWin * kOwnWin;  // one thread per window cell
constexpr int kOwnCap = 1024;                 // particles staged per pass
constexpr int kOwnFields = 9;
constexpr int kOwnSlots = 33;                 // 6 u (mom, w), 6 v (mom, w), 9 density
static_assert(kOwnFields * kOwnCap >= kOwnSlots * kOwnCells, "partials alias the staging buffer");

struct OwnerAcc {
    float um[6], uw[6];  // u slot kx * 3 + sy: face (cx + kx, cy - 1 + sy)
    float vm[6], vw[6];  // v slot sx * 2 + ky: face (cx - 1 + sx, cy + ky)
    float d[9];          // cell slot sx * 3 + sy: cell (cx - 1 + sx, cy - 1 + sy)
};

template <int N>
__device__ inline void slot_add(float (&acc)[N], int s, float v) {
#pragma unroll
    for (int t = 0; t < N; ++t)
        if (s == t) acc[t] += v;
}

// Two stencil corners c0 (offset 0) and c1 (offset 1) added to three slots
// whose first one is the stencil's start when h is false, one before it when
// h is true.
__device__ inline void shift3(bool h, float c0, float c1, float& s0, float& s1, float& s2) {
    const float z = 0.0f;
    s0 += h ? z : c0;
    s1 += h ? c0 : c1;
    s2 += h ? c1 : z;
}

// One particle of cell (cx, cy): the same arithmetic as p2g_contrib.
__device__ inline void owner_accumulate(const float* __restrict__ sp, int k, float prev_dt, float cell, int32_t Wi,
                                        int32_t Hi, int32_t cx, int32_t cy, OwnerAcc& a) {
    const float x = sp[0 * kOwnCap + k], y = sp[1 * kOwnCap + k];
    const float vx = sp[2 * kOwnCap + k], vy = sp[3 * kOwnCap + k];
    const float cux = sp[4 * kOwnCap + k], cuy = sp[5 * kOwnCap + k];
    const float cvx = sp[6 * kOwnCap + k], cvy = sp[7 * kOwnCap + k];
    const float tres = sp[8 * kOwnCap + k];
    const float tau = (-tres) / prev_dt;
    const float wt = temporal_w(tau);
    const float gx = x / cell, gy = y / cell;
    const float cgx = gx - 0.5f, cgy = gy - 0.5f;
    const float cbx = floorf(cgx), cby = floorf(cgy);
    const float ctx = cgx - cbx, cty = cgy - cby;
    float cw[4];
    cw[0] = ((1.0f - ctx) * (1.0f - cty)) * wt;
    cw[1] = (ctx * (1.0f - cty)) * wt;
    cw[2] = ((1.0f - ctx) * cty) * wt;
    cw[3] = (ctx * cty) * wt;
    float uadd[4], uwt[4], vadd[4], vwt[4];
    Stencil su, sv;
    make_stencil(gx, gy - 0.5f, cell, su);
    make_stencil(gx - 0.5f, gy, cell, sv);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float ww = su.w[c] * wt;
        uwt[c] = ww;
        uadd[c] = ((vx + cux * su.dx[c]) + cuy * su.dy[c]) * ww;
        ww = sv.w[c] * wt;
        vwt[c] = ww;
        vadd[c] = ((vy + cvx * sv.dx[c]) + cvy * sv.dy[c]) * ww;
    }
    const bool regular = gx >= 1.0f && gx < (float)(Wi - 1) && gy >= 1.0f && gy < (float)(Hi - 1);
    if (regular) {
        // No index is clamped and the particle lies in (cx, cy): the stencils
        // start at column cx - 1 + hx and row cy - 1 + hy.
        const bool hx = (int32_t)cbx - (cx - 1) != 0;
        const bool hy = (int32_t)cby - (cy - 1) != 0;
        shift3(hy, uadd[0], uadd[2], a.um[0], a.um[1], a.um[2]);
        shift3(hy, uwt[0], uwt[2], a.uw[0], a.uw[1], a.uw[2]);
        shift3(hy, uadd[1], uadd[3], a.um[3], a.um[4], a.um[5]);
        shift3(hy, uwt[1], uwt[3], a.uw[3], a.uw[4], a.uw[5]);
        shift3(hx, vadd[0], vadd[1], a.vm[0], a.vm[2], a.vm[4]);
        shift3(hx, vwt[0], vwt[1], a.vw[0], a.vw[2], a.vw[4]);
        shift3(hx, vadd[2], vadd[3], a.vm[1], a.vm[3], a.vm[5]);
        shift3(hx, vwt[2], vwt[3], a.vw[1], a.vw[3], a.vw[5]);
        const float r00 = hx ? 0.0f : cw[0], r01 = hx ? cw[0] : cw[1], r02 = hx ? cw[1] : 0.0f;
        const float r10 = hx ? 0.0f : cw[2], r11 = hx ? cw[2] : cw[3], r12 = hx ? cw[3] : 0.0f;
        shift3(hy, r00, r10, a.d[0], a.d[1], a.d[2]);
        shift3(hy, r01, r11, a.d[3], a.d[4], a.d[5]);
        shift3(hy, r02, r12, a.d[6], a.d[7], a.d[8]);
    } else {
        // Clamped corners (a particle near or beyond the walls): slots from
        // the clamped indices, which always stay within the cell's reach.
        const int32_t ux0 = clampi(su.ix, 0, Wi) - cx, ux1 = clampi(su.ix + 1, 0, Wi) - cx;
        const int32_t uy0 = clampi(su.iy, 0, Hi - 1) - (cy - 1), uy1 = clampi(su.iy + 1, 0, Hi - 1) - (cy - 1);
        const int32_t vx0 = clampi(sv.ix, 0, Wi - 1) - (cx - 1), vx1 = clampi(sv.ix + 1, 0, Wi - 1) - (cx - 1);
        const int32_t vy0 = clampi(sv.iy, 0, Hi) - cy, vy1 = clampi(sv.iy + 1, 0, Hi) - cy;
        const int32_t bx = (int32_t)cbx, by = (int32_t)cby;
        const int32_t dx0 = clampi(bx, 0, Wi - 1) - (cx - 1), dx1 = clampi(bx + 1, 0, Wi - 1) - (cx - 1);
        const int32_t dy0 = clampi(by, 0, Hi - 1) - (cy - 1), dy1 = clampi(by + 1, 0, Hi - 1) - (cy - 1);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int32_t us = ((c & 1) ? ux1 : ux0) * 3 + ((c >> 1) ? uy1 : uy0);
            const int32_t vs = ((c & 1) ? vx1 : vx0) * 2 + ((c >> 1) ? vy1 : vy0);
            const int32_t ds = ((c & 1) ? dx1 : dx0) * 3 + ((c >> 1) ? dy1 : dy0);
            slot_add(a.um, us, uadd[c]);
            slot_add(a.uw, us, uwt[c]);
            slot_add(a.vm, vs, vadd[c]);
            slot_add(a.vw, vs, vwt[c]);
            slot_add(a.d, ds, cw[c]);
        }
    }
}

__device__ __forceinline__ void p2g_owner(
    const float* __restrict__ pb, size_t nalloc, uint32_t np, const uint32_t* __restrict__ order, float prev_dt_arg,
    float cell, uint32_t W, uint32_t H, const uint2* __restrict__ cell_range, uint32_t tiles_x,
    float* __restrict__ mac_u, float* __restrict__ mac_v, uint8_t* __restrict__ valid_u,
    uint8_t* __restrict__ valid_v, float* __restrict__ density, uint8_t* __restrict__ fluid,
    const uint8_t* __restrict__ obstacle, const float* __restrict__ ext_u, const float* __restrict__ ext_v,
    int forces, int ext, float dt_arg, float gravity, float gravity_x, const StepState* __restrict__ ss,
    uint32_t wg) {
    __shared__ float s_buf[kOwnFields * kOwnCap];
    const float prev_dt = ss ? ss->prev_dt : prev_dt_arg;
    const float dt = ss ? ss->dt : dt_arg;
    const float g = gravity * dt;
    const float gx = gravity_x * dt;  // horizontal gravity (an extension; zero in the reference)
    __shared__ uint32_t s_off[kOwnCells + 1];      // window-order prefix of the cell counts
    __shared__ uint32_t s_src[kOwnCells];          // sorted-order begin of each window cell
    __shared__ uint32_t s_wave[kOwnCells / 64];
    __shared__ uint8_t s_fl[kOwnCells];
    __shared__ uint8_t s_cell[kOwnCap];            // the window cell of each staged particle
    static_assert(kOwnCells <= 256, "window cells are bytes");
    const int32_t Wi = (int32_t)W, Hi = (int32_t)H;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wx = tid % kOwnWin, wy = tid / kOwnWin;
    const int32_t cx = (int32_t)(wg % tiles_x) * kOwnTile - 1 + wx;
    const int32_t cy = (int32_t)(wg / tiles_x) * kOwnTile - 1 + wy;
    const bool inside = cx >= 0 && cx < Wi && cy >= 0 && cy < Hi;
    uint32_t b = 0, cnt = 0;
    bool fl = false;
    if (inside) {
        const uint2 r = cell_range[morton((uint32_t)cx, (uint32_t)cy)];
        b = r.x;
        cnt = (r.y > r.x && r.y <= np) ? r.y - r.x : 0u;
        fl = cnt > 0 && !obstacle[(size_t)cy * W + cx];
    }
    // exclusive prefix of the counts in window order
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    s_src[tid] = b;
    s_fl[tid] = fl;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kOwnCells / 64; ++w) {
        if (w < wave) wbase += s_wave[w];
        total += s_wave[w];
    }
    const uint32_t my0 = wbase + incl - cnt, my1 = wbase + incl;
    s_off[tid] = my0;
    if (tid == 0) s_off[kOwnCells] = total;
    __syncthreads();

    OwnerAcc a;
#pragma unroll
    for (int s = 0; s < 6; ++s) { a.um[s] = 0.0f; a.uw[s] = 0.0f; a.vm[s] = 0.0f; a.vw[s] = 0.0f; }
#pragma unroll
    for (int s = 0; s < 9; ++s) a.d[s] = 0.0f;
    for (uint32_t base = 0; base < total; base += kOwnCap) {
        const uint32_t n = min((uint32_t)kOwnCap, total - base);
        {  // each cell marks its particles' slots of this pass
            const uint32_t lo = max(my0, base), hi = min(my1, base + n);
            for (uint32_t idx = lo; idx < hi; ++idx) s_cell[idx - base] = (uint8_t)tid;
        }
        __syncthreads();
        // a thread's kOwnCap / kOwnCells slots: their positions in cell order,
        // then the permutation entries, then the fields, each stage's loads
        // issued together
        constexpr int kPer = kOwnCap / kOwnCells;
        static_assert(kPer * kOwnCells == kOwnCap, "whole slots per thread");
        uint32_t src[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t k = (uint32_t)tid + (uint32_t)(j * kOwnCells);
            const uint32_t kk = k < n ? k : 0u;
            const int lo = s_cell[kk];  // the window cell holding base + kk
            src[j] = s_src[lo] + (base + kk - s_off[lo]);  // position in cell order
        }
        if (order) {
            // (a permutation entry past the records, written through the
            // field API, reads the last record, as the sort's gather does)
#pragma unroll
            for (int j = 0; j < kPer; ++j) src[j] = min(order[src[j]], np - 1);
        }
        float v[kPer][kOwnFields];
#pragma unroll
        for (int j = 0; j < kPer; ++j)
#pragma unroll
            for (int f = 0; f < kOwnFields; ++f) v[j][f] = pb[(size_t)f * nalloc + src[j]];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t k = (uint32_t)tid + (uint32_t)(j * kOwnCells);
            if (k < n)
#pragma unroll
                for (int f = 0; f < kOwnFields; ++f) s_buf[f * kOwnCap + k] = v[j][f];
        }
        __syncthreads();
        const uint32_t lo = max(my0, base), hi = min(my1, base + n);
        for (uint32_t idx = lo; idx < hi; ++idx)
            owner_accumulate(s_buf, (int)(idx - base), prev_dt, cell, Wi, Hi, cx, cy, a);
        __syncthreads();
    }
    // partials of every window cell, slot-major
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        s_buf[(0 + s) * kOwnCells + tid] = a.um[s];
        s_buf[(6 + s) * kOwnCells + tid] = a.uw[s];
        s_buf[(12 + s) * kOwnCells + tid] = a.vm[s];
        s_buf[(18 + s) * kOwnCells + tid] = a.vw[s];
    }
#pragma unroll
    for (int s = 0; s < 9; ++s) s_buf[(24 + s) * kOwnCells + tid] = a.d[s];
    __syncthreads();
    if (!inside || wx < 1 || wx > kOwnTile || wy < 1 || wy > kOwnTile) return;
#define PART(slot, dwx, dwy) s_buf[(slot) * kOwnCells + tid + (dwy) * kOwnWin + (dwx)]
    // u face (cx, cy): cells (cx - 1 + dx, cy - 1 + dy), slot (1 - dx) * 3 + (2 - dy)
    {
        float m = 0.0f, w = 0.0f;
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int s = (1 - dx) * 3 + (2 - dy);
                m += PART(0 + s, dx - 1, dy - 1);
                w += PART(6 + s, dx - 1, dy - 1);
            }
        const size_t f = (size_t)cy * (W + 1) + cx;
        const bool ok = w > 0.0f;
        float u = ok ? m / w : 0.0f;
        if (ext) u += ext_u[f] * dt;
        if (forces && gravity_x != 0.0f && (s_fl[tid - 1] || s_fl[tid])) u += gx;
        mac_u[f] = u;
        valid_u[f] = ok;
        if (cx == Wi - 1) {  // face x = W: only the cells x = W - 1 reach it
            m = 0.0f; w = 0.0f;
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                m += PART(0 + 3 + (2 - dy), 0, dy - 1);
                w += PART(6 + 3 + (2 - dy), 0, dy - 1);
            }
            const bool ok2 = w > 0.0f;
            u = ok2 ? m / w : 0.0f;
            if (ext) u += ext_u[f + 1] * dt;
            if (forces && gravity_x != 0.0f && s_fl[tid]) u += gx;
            mac_u[f + 1] = u;
            valid_u[f + 1] = ok2;
        }
    }
    // v face (cx, cy): cells (cx - 1 + dx, cy - 1 + dy), slot (2 - dx) * 2 + (1 - dy)
    {
        float m = 0.0f, w = 0.0f;
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int s = (2 - dx) * 2 + (1 - dy);
                m += PART(12 + s, dx - 1, dy - 1);
                w += PART(18 + s, dx - 1, dy - 1);
            }
        const size_t f = (size_t)cy * W + cx;
        const bool ok = w > 0.0f;
        float v = ok ? m / w : 0.0f;
        if (ext) v += ext_v[f] * dt;
        if (forces && ((cy >= 1 && s_fl[tid - kOwnWin]) || s_fl[tid])) v += g;
        mac_v[f] = v;
        valid_v[f] = ok;
        if (cy == Hi - 1) {  // face y = H: only the cells y = H - 1 reach it
            m = 0.0f; w = 0.0f;
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                m += PART(12 + (2 - dx) * 2 + 1, dx - 1, 0);
                w += PART(18 + (2 - dx) * 2 + 1, dx - 1, 0);
            }
            const bool ok2 = w > 0.0f;
            v = ok2 ? m / w : 0.0f;
            if (ext) v += ext_v[f + W] * dt;
            if (forces && s_fl[tid]) v += g;
            mac_v[f + W] = v;
            valid_v[f + W] = ok2;
        }
    }
    // cell (cx, cy): cells (cx - 1 + dx, cy - 1 + dy), slot (2 - dx) * 3 + (2 - dy)
    {
        float dsum = 0.0f;
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) dsum += PART(24 + (2 - dx) * 3 + (2 - dy), dx - 1, dy - 1);
        const size_t c = (size_t)cy * W + cx;
        density[c] = dsum;
        fluid[c] = s_fl[tid];
    }
#undef PART
}

struct P2gOwner {
    static constexpr int kThreads = kOwnCells;
    static constexpr bool kRoundRobin = true;  // work follows the particles
    struct Args {
        const float* pb;
        size_t nalloc;
        uint32_t np;
        const uint32_t* order;
        float prev_dt;
        float cell;
        uint32_t W, H;
        const uint2* cell_range;
        uint32_t tiles_x;
        float* mac_u;
        float* mac_v;
        uint8_t* valid_u;
        uint8_t* valid_v;
        float* density;
        uint8_t* fluid;
        const uint8_t* obstacle;
        const float* ext_u;
        const float* ext_v;
        int forces, ext;
        float dt, gravity, gravity_x;
        const StepState* ss;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        p2g_owner(a.pb, a.nalloc, a.np, a.order, a.prev_dt, a.cell, a.W, a.H, a.cell_range, a.tiles_x, a.mac_u,
                  a.mac_v, a.valid_u, a.valid_v, a.density, a.fluid, a.obstacle, a.ext_u, a.ext_v, a.forces, a.ext,
                  a.dt, a.gravity, a.gravity_x, a.ss, wg);
    }
};

// [begin, end) of every cell's particles in cell order, indexed by key
struct CellRanges {
    static constexpr int kThreads = 256;
    struct Args {
        const uint32_t* keys;
        uint32_t np;
        uint2* range;
        uint32_t nkeys;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        const uint32_t i = wg * kThreads + threadIdx.x;
        const uint32_t np = a.np;
        if (i >= np) return;
        const uint32_t* __restrict__ keys = a.keys;
        const uint32_t k = keys[i];
        if (k >= a.nkeys) return;  // keys come from clamped cell coordinates; kept as a guard
        if (i == 0 || keys[i - 1] != k) a.range[k].x = i;
        if (i + 1 == np || keys[i + 1] != k) a.range[k].y = i + 1;
    }
};

// apply_external_forces (src/apic.rs:1604-1626): u, v += force * dt on every
// face, v += g * dt on faces with a fluid cell below or above.
__global__ void forces_kernel(float* __restrict__ mac_u, float* __restrict__ mac_v, const float* __restrict__ ext_u,
                              const float* __restrict__ ext_v, const uint8_t* __restrict__ fluid, uint32_t W,
                              uint32_t H, int ext_forces, float dt, float g, float gx) {
    const size_t nu = (size_t)W * H + H, nv = (size_t)W * H + W;
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nu) {
        float u = mac_u[f];
        if (ext_forces) u = u + ext_u[f] * dt;
        if (gx != 0.0f) {  // horizontal gravity (an extension): faces with a fluid cell left or right
            uint32_t x = (uint32_t)(f % (W + 1)), y = (uint32_t)(f / (W + 1));
            bool left = x >= 1 && fluid[(size_t)y * W + x - 1];
            bool right = x < W && fluid[(size_t)y * W + x];
            if (left || right) u += gx;
        }
        mac_u[f] = u;
    }
    if (f < nv) {
        float v = mac_v[f];
        if (ext_forces) v += ext_v[f] * dt;
        uint32_t x = (uint32_t)(f % W), y = (uint32_t)(f / W);
        bool below = y >= 1 && fluid[(size_t)(y - 1) * W + x];
        bool above = y < H && fluid[(size_t)y * W + x];
        if (below || above) v += g;
        mac_v[f] = v;
    }
}

// One particle of G2P between its loads: the record's position and time
// residual, its two stencils and the grid values and open flags of their
// corners.
struct G2pLane {
    float x, y, res;
    Stencil su, sv;
    uint32_t ui[4], vi[4];
    float un[4], vn[4];
    uint8_t ou[4], ov[4];
};

__device__ __forceinline__ float masked_interp_loaded(const float* w, const uint8_t* open, const float* nodes) {
    float mw[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) mw[k] = open[k] ? w[k] : 0.0f;
    // bilinear_interpolate (src/apic.rs:138-172)
    float ws = ((mw[0] + mw[1]) + mw[2]) + mw[3];
    float inv = 1.0f / fast_max(ws, 1e-6f);
    float vel = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) vel += (nodes[k] * mw[k]) * inv;
    return vel;
}

// pin / pout: the records before and after (the same buffer, or with a
// deferred sort gather the other one, pin read through perm: record perm[i]
// becomes record i, and lanes past the records read zeros, as the gather
// leaves them).  Each thread takes U particles per step, nwg * blockDim
// apart (each wave's lanes stay consecutive for every one of them), and
// issues their loads stage by stage: the permutation entries, the records,
// then the grid corners of all U, so that a step waits three times for U
// particles instead of once per load chain of each.
template <int U>
__device__ __forceinline__ void g2p(
    const float* pin, float* pout, const uint32_t* perm, size_t nalloc, uint32_t np, float dt_arg,
    float cell, uint32_t W, uint32_t H,
    uint32_t timestamp, float border_damping, const float* __restrict__ mac_u, const float* __restrict__ mac_v,
    const uint8_t* __restrict__ open_u, const uint8_t* __restrict__ open_v,
    int write_keys, uint32_t* __restrict__ keys, uint32_t* vals, float* __restrict__ part_maxv,
    const StepState* __restrict__ ss, const int32_t* __restrict__ guard, uint32_t wg, uint32_t nwg) {
    __shared__ float s_max[256 / 64];
    if (guard && *guard == 0) return;
    const float dt = ss ? ss->dt : dt_arg;
    float speed_new = 0.0f;
    const size_t stride = (size_t)nwg * blockDim.x;
    for (size_t i0 = (size_t)wg * blockDim.x + threadIdx.x; i0 < nalloc; i0 += (size_t)U * stride) {
        G2pLane p[U];
        size_t src[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = i0 + (size_t)u * stride;
            src[u] = i < nalloc ? i : 0;
            if (perm) {
                const uint32_t j = perm[i < np ? i : 0];
                src[u] = j >= np ? np - 1 : j;  // (as the sort's gather)
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = i0 + (size_t)u * stride;
            const bool rec = i < nalloc && (!perm || i < np);
            const float x = pin[(size_t)P_X * nalloc + src[u]];
            const float y = pin[(size_t)P_Y * nalloc + src[u]];
            const float res = pin[(size_t)T_RES * nalloc + src[u]];
            p[u].x = rec ? x : 0.0f;
            p[u].y = rec ? y : 0.0f;
            p[u].res = rec ? res : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float gx = p[u].x / cell, gy = p[u].y / cell;
            make_stencil(gx, gy - 0.5f, cell, p[u].su);
            make_stencil(gx - 0.5f, gy, cell, p[u].sv);
            const Stencil& su = p[u].su;
            const Stencil& sv = p[u].sv;
            p[u].ui[0] = u_index(su.ix, su.iy, W, H); p[u].ui[1] = u_index(su.ix + 1, su.iy, W, H);
            p[u].ui[2] = u_index(su.ix, su.iy + 1, W, H); p[u].ui[3] = u_index(su.ix + 1, su.iy + 1, W, H);
            p[u].vi[0] = v_index(sv.ix, sv.iy, W, H); p[u].vi[1] = v_index(sv.ix + 1, sv.iy, W, H);
            p[u].vi[2] = v_index(sv.ix, sv.iy + 1, W, H); p[u].vi[3] = v_index(sv.ix + 1, sv.iy + 1, W, H);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                p[u].un[k] = mac_u[p[u].ui[k]];
                p[u].vn[k] = mac_v[p[u].vi[k]];
                p[u].ou[k] = open_u[p[u].ui[k]];
                p[u].ov[k] = open_v[p[u].vi[k]];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = i0 + (size_t)u * stride;
            if (i >= nalloc) continue;
            const Stencil& su = p[u].su;
            const Stencil& sv = p[u].sv;
            const float* un = p[u].un;
            const float* vn = p[u].vn;
            const float x = p[u].x, y = p[u].y, res = p[u].res;
            float nvx = masked_interp_loaded(su.w, p[u].ou, un);
            float nvy = masked_interp_loaded(sv.w, p[u].ov, vn);
            // jittered time step (:2173-2186)
            float speed = sqrtf(nvx * nvx + nvy * nvy);
            float cfl_local = (speed * dt) / cell;
            float scale = fast_min(fast_max(cfl_local, 0.0f), 1.0f);
            float gam = (scale * scale) * (3.0f - (2.0f * scale));
            float jitter = (gam * random_lane((uint32_t)i, timestamp)) * dt;
            float dt_raw = (dt + res) + jitter;
            float dt_act = fast_min(fast_max(dt_raw, 0.0f), dt * 2.0f);
            float new_res = (dt + res) - dt_act;
            float nx = x + dt_act * nvx, ny = y + dt_act * nvy;
            // wall clamp and reflection (:2188-2209)
            const float min_x = 1.001f * cell, max_x = ((float)W - 1.001f) * cell;
            const float min_y = 1.001f * cell, max_y = ((float)H - 1.001f) * cell;
            const float bd = -border_damping;
            bool oxmin = nx < min_x, oxmax = nx > max_x, oymin = ny < min_y, oymax = ny > max_y;
            bool ovxmin = oxmin && nvx < 0.0f, ovxmax = oxmax && nvx > 0.0f;
            bool ovymin = oymin && nvy < 0.0f, ovymax = oymax && nvy > 0.0f;
            if (oxmin) nx = min_x;
            if (oxmax) nx = max_x;
            if (ovxmin) nvx = nvx * bd;
            if (ovxmax) nvx = nvx * bd;
            if (oymin) ny = min_y;
            if (oymax) ny = max_y;
            if (ovymin) nvy = nvy * bd;
            if (ovymax) nvy = nvy * bd;
            // affine matrix C with unmasked weights (:2211-2247)
            const float h2 = cell * cell;
            float udx = (h2 * su.tx) * (1.0f - su.tx), udy = (h2 * su.ty) * (1.0f - su.ty);
            float vdx = (h2 * sv.tx) * (1.0f - sv.tx), vdy = (h2 * sv.ty) * (1.0f - sv.ty);
            float uix = udx > 1e-6f ? 1.0f / udx : 0.0f, uiy = udy > 1e-6f ? 1.0f / udy : 0.0f;
            float vix = vdx > 1e-6f ? 1.0f / vdx : 0.0f, viy = vdy > 1e-6f ? 1.0f / vdy : 0.0f;
            float cux = 0.0f, cuy = 0.0f, cvx = 0.0f, cvy = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                cux += (su.w[k] * un[k]) * su.dx[k];
                cuy += (su.w[k] * un[k]) * su.dy[k];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                cvx += (sv.w[k] * vn[k]) * sv.dx[k];
                cvy += (sv.w[k] * vn[k]) * sv.dy[k];
            }
            float c_ux = cux * uix, c_uy = cuy * uiy, c_vx = cvx * vix, c_vy = cvy * viy;
            bool rx = ovxmin || ovxmax, ry = ovymin || ovymax;
            if (rx) { c_uy = c_uy * -1.0f; c_vx = c_vx * -1.0f; }
            if (ry) { c_uy = c_uy * -1.0f; c_vx = c_vx * -1.0f; }
            pout[(size_t)P_X * nalloc + i] = nx;
            pout[(size_t)P_Y * nalloc + i] = ny;
            pout[(size_t)V_X * nalloc + i] = nvx;
            pout[(size_t)V_Y * nalloc + i] = nvy;
            pout[(size_t)CU_X * nalloc + i] = c_ux;
            pout[(size_t)CU_Y * nalloc + i] = c_uy;
            pout[(size_t)CV_X * nalloc + i] = c_vx;
            pout[(size_t)CV_Y * nalloc + i] = c_vy;
            pout[(size_t)T_RES * nalloc + i] = new_res;
            speed_new = fast_max(speed_new, sqrtf(nvx * nvx + nvy * nvy));
            if (write_keys && i < np) {
                // keygen of the next update_spatial_lookup (:1049-1064)
                float kx = nx / cell, ky = ny / cell;
                uint32_t ix = (uint32_t)(int32_t)fast_min(fast_max(kx, 0.0f), (float)(W - 1));
                uint32_t iy = (uint32_t)(int32_t)fast_min(fast_max(ky, 0.0f), (float)(H - 1));
                keys[i] = morton(ix, iy);
                vals[i] = (uint32_t)i;  // (perm, when given, is this array: every lane read its entries above)
            }
        }
    }
    // max |v| over all lanes including the tail lanes of the last chunk
    float m = speed_new;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fast_max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = s_max[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = fast_max(b, s_max[w]);
        part_maxv[wg] = b;
    }
}

template <int U>
struct G2p {
    static constexpr int kThreads = 256;
    struct Args {
        const float* pin;
        float* pout;
        const uint32_t* perm;
        size_t nalloc;
        uint32_t np;
        float dt;
        float cell;
        uint32_t W, H, timestamp;
        float border_damping;
        const float* mac_u;
        const float* mac_v;
        const uint8_t* open_u;
        const uint8_t* open_v;
        int write_keys;
        uint32_t* keys;
        uint32_t* vals;
        float* part_maxv;
        const StepState* ss;
        const int32_t* guard;
        uint32_t nwg;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        g2p<U>(a.pin, a.pout, a.perm, a.nalloc, a.np, a.dt, a.cell, a.W, a.H, a.timestamp, a.border_damping,
               a.mac_u, a.mac_v, a.open_u, a.open_v, a.write_keys, a.keys, a.vals, a.part_maxv, a.ss, a.guard, wg,
               a.nwg);
    }
};

__global__ void max_velocity_kernel(const float* __restrict__ vx, const float* __restrict__ vy, size_t n,
                                    float* __restrict__ part_maxv) {
    __shared__ float s_max[256 / 64];
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fast_max(m, sqrtf(vx[i] * vx[i] + vy[i] * vy[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fast_max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = s_max[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = fast_max(b, s_max[w]);
        part_maxv[blockIdx.x] = b;
    }
}

// max over the per-workgroup maxima into the float bits read by the host;
// with ss, also the end of the substep (SubstepEnd's bookkeeping, under the
// same guard), one launch instead of two
struct MaxvFinalize {
    static constexpr int kThreads = 1024;
    struct Args {
        const float* part_maxv;
        uint32_t n;
        uint32_t* maxv_bits;
        uint32_t* host_bits;
        const int32_t* guard;
        StepState* ss;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) {
        __shared__ float s_max[1024 / 64];
        if (a.guard && *a.guard == 0) return;
        float m = 0.0f;
        for (uint32_t i = threadIdx.x; i < a.n; i += blockDim.x) m = fast_max(m, a.part_maxv[i]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fast_max(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            float b = 0.0f;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) b = fast_max(b, s_max[w]);
            *a.maxv_bits = __float_as_uint(b);
            if (a.host_bits) *a.host_bits = __float_as_uint(b);
            if (a.ss) substep_end(a.ss);
        }
    }
};

constexpr int kG2pPer = 2;  // particles per thread and loop step of G2P

unsigned particle_blocks(const Sim& sim) {
    size_t b = (sim.nalloc + 255) / 256;
    return (unsigned)std::min<size_t>(std::max<size_t>(b, 1), sim.part_maxv_cap);
}

}  // namespace

size_t cell_key_count(const Sim& sim) {
    return (size_t)morton(sim.W > 0 ? sim.W - 1 : 0, sim.H > 0 ? sim.H - 1 : 0) + 1;
}

void launch_cell_ranges(Sim& sim) {
    const size_t nk = cell_key_count(sim);
    launch_fill_u32(sim, reinterpret_cast<uint32_t*>(sim.cell_range), 0u, 2 * nk);
    if (sim.num_particles == 0) return;
    emit<CellRanges>(sim, {sim.keys[0], sim.num_particles, sim.cell_range, (uint32_t)nk},
                     (sim.num_particles + 255) / 256);
}

void launch_p2g(Sim& sim, bool forces, float dt) {
    const uint32_t* order = nullptr;
    if (!sim.cells_valid) {
        // The particles are not in cell order (sort disabled, or state written
        // through the field API): sort an index by cell instead of the particles.
        if (sim.num_particles > 0) {
            if (!sim.keys_valid) launch_keygen(sim);
            launch_radix_sort(sim);
            sim.keys_valid = false;  // keys[0] now holds the sorted keys
            order = sim.vals[0];
        }
        launch_cell_ranges(sim);
    } else if (sim.perm_pending) {
        order = sim.vals[0];  // the sort's gather is deferred: read through its permutation
    }
    const uint32_t tx = (sim.W + kOwnTile - 1) / kOwnTile, ty = (sim.H + kOwnTile - 1) / kOwnTile;
    const int ext = forces && sim.forces_active;
    sim.mark(APIC_T_K_P2G, 0);
    P2gOwner::Args a{sim.pbuf[sim.cur], sim.nalloc, sim.num_particles, order, sim.prev_dt, sim.cell, sim.W, sim.H,
                     sim.cell_range, tx, sim.mac_u, sim.mac_v, sim.valid_u, sim.valid_v, sim.density,
                     sim.levels[0].fluid, sim.obstacle, sim.ext_u, sim.ext_v, (int)forces, ext, dt,
                     sim.props.gravity, sim.gravity_x, sim.step_src};
    emit<P2gOwner>(sim, a, tx * ty);
    sim.mark(APIC_T_K_P2G, 1);
}

void launch_forces(Sim& sim, float dt) {
    size_t n = std::max(sim.nu(), sim.nv());
    float g = sim.props.gravity * dt;
    int ext = sim.forces_active;
    hipLaunchKernelGGL(forces_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, sim.stream, sim.mac_u,
                       sim.mac_v, sim.ext_u, sim.ext_v, sim.levels[0].fluid, sim.W, sim.H, ext, dt, g,
                       sim.gravity_x * dt);
}

void launch_g2p(Sim& sim, float dt, bool write_keys, bool end_substep) {
    if (sim.nalloc == 0) {
        launch_fill_u32(sim, sim.maxv_bits, 0u, 1);
        if (sim.h_maxv) *sim.h_maxv = 0u;
        if (end_substep) launch_substep_end(sim);
        return;
    }
    const unsigned blocks = particle_blocks(sim);
    sim.mark(APIC_T_K_G2P, 0);
    // a deferred sort gather: read the records through the permutation and
    // write the sorted order into the other buffer
    const bool gather = sim.perm_pending;
    float* pout = sim.pbuf[gather ? 1 - sim.cur : sim.cur];
    G2p<kG2pPer>::Args a{sim.pbuf[sim.cur], pout, gather ? sim.vals[0] : nullptr, sim.nalloc, sim.num_particles, dt,
                sim.cell, sim.W, sim.H, sim.timestamp, sim.props.border_damping, sim.mac_u, sim.mac_v, sim.open_u,
                sim.open_v, (int)write_keys, sim.keys[0], sim.vals[0], sim.part_maxv, sim.step_src, sim.tail_guard,
                blocks};
    emit<G2p<kG2pPer>>(sim, a, blocks);
    if (gather) {
        sim.cur = 1 - sim.cur;
        sim.perm_pending = false;
    }
    sim.mark(APIC_T_K_G2P, 1);
    emit<MaxvFinalize>(sim, {sim.part_maxv, blocks, sim.maxv_bits, sim.h_maxv_d, sim.tail_guard,
                             end_substep ? sim.step : nullptr},
                       1);
}

void launch_max_velocity(Sim& sim) {
    if (sim.nalloc == 0) {
        launch_fill_u32(sim, sim.maxv_bits, 0u, 1);
        if (sim.h_maxv) *sim.h_maxv = 0u;
        return;
    }
    const unsigned blocks = particle_blocks(sim);
    hipLaunchKernelGGL(max_velocity_kernel, dim3(blocks), dim3(256), 0, sim.stream, sim.pfield(V_X),
                       sim.pfield(V_Y), sim.nalloc, sim.part_maxv);
    emit<MaxvFinalize>(sim, {sim.part_maxv, blocks, sim.maxv_bits, sim.h_maxv_d, nullptr}, 1);
}

}  // namespace apic
