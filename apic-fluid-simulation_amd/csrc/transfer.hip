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

// ---- block P2G ----------------------------------------------------------------
//
// In cell order (the Morton-sorted particles, or the particles through a
// sorted index `order`), the particles of one cell are the contiguous range
// cell_range[key] = [begin, end), and the 256 cells of a Morton-aligned
// 16 x 16 block (keys (block << 8) + 0..255) hold one contiguous run.  A
// workgroup takes one block: it stages the run in LDS with coalesced loads,
// then one thread per cell sums the contributions of its own particles into
// 33 register partials: the 2x3 u faces, 3x2 v faces and 3x3 cells that a
// particle of that cell can reach (src/apic.rs:1237-1426, with u_index /
// v_index clamping, :1029-1041).  A face or cell whose source cells (the
// existing cells that can reach it) all lie in the block is finished by that
// workgroup: it adds the partials of its sources in a fixed order, then
// u = momentum / weight and the valid flag (:1428-1459), the fluid flag
// (:1461-1464) and, when asked, apply_external_forces (:1604-1626).  For every
// other target it reaches (a band around its edge) the block writes its
// partial sum into its halo record, and P2gSeam adds the records of the blocks
// that reach such a target, in a fixed block order, and finishes it.  Each
// particle is read once, nothing needs atomics and the result is
// deterministic.  The faces x = W and y = H belong to the cells x = W-1 and
// y = H-1.

constexpr int kBlk = (int)kP2gBlockSide;        // block side in cells
constexpr int kBlkCells = kBlk * kBlk;          // one thread per block cell
constexpr int kOwnCap = 1024;                   // particles staged per pass
constexpr int kOwnFields = 9;
constexpr int kOwnSlots = 33;                   // 6 u (mom, w), 6 v (mom, w), 9 density
static_assert(kOwnFields * kOwnCap >= kOwnSlots * kBlkCells, "partials alias the staging buffer");
static_assert(kBlkCells == 256, "Morton offsets of a block are one byte");

// A block's halo record: the partial sums of the targets it reaches outside
// its interior, as five arrays (u momentum, u weight, v momentum, v weight,
// density).  Positions are relative to the block's origin (x0, y0):
//   u faces (x0 + i, y0 - 1 + j), i in [0,16], j in [0,17]: rows 0-1, rows
//     16-17, then columns 0 and 16 of rows 2-15 (96);
//   v faces (x0 - 1 + i, y0 + j), i in [0,17], j in [0,16]: the transpose
//     of the u layout (96);
//   cells (x0 - 1 + i, y0 - 1 + j), i, j in [0,17]: rows 0-1, rows 16-17,
//     then columns 0, 1, 16 and 17 of rows 2-15 (128).
constexpr int kHaloUm = 0, kHaloUw = 96, kHaloVm = 192, kHaloVw = 288, kHaloDen = 384;
static_assert(kHaloDen + 128 == (int)kP2gHaloFloats, "halo record size");

__device__ inline int halo_u_index(int i, int j) {
    if (j < 2) return j * 17 + i;
    if (j >= 16) return 34 + (j - 16) * 17 + i;
    return 68 + (j - 2) * 2 + (i == 16 ? 1 : 0);
}
__device__ inline void halo_u_pos(int e, int& i, int& j) {
    if (e < 34) { j = e / 17; i = e % 17; }
    else if (e < 68) { j = 16 + (e - 34) / 17; i = (e - 34) % 17; }
    else { j = 2 + (e - 68) / 2; i = ((e - 68) & 1) ? 16 : 0; }
}
__device__ inline int halo_v_index(int i, int j) { return halo_u_index(j, i); }
__device__ inline void halo_v_pos(int e, int& i, int& j) { halo_u_pos(e, j, i); }
__device__ inline int halo_d_index(int i, int j) {
    if (j < 2) return j * 18 + i;
    if (j >= 16) return 36 + (j - 16) * 18 + i;
    return 72 + (j - 2) * 4 + (i < 2 ? i : i - 14);
}
__device__ inline void halo_d_pos(int e, int& i, int& j) {
    if (e < 36) { j = e / 18; i = e % 18; }
    else if (e < 72) { j = 16 + (e - 36) / 18; i = (e - 36) % 18; }
    else { j = 2 + (e - 72) / 4; const int c = (e - 72) & 3; i = c < 2 ? c : c + 14; }
}

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

// The staged form of a particle: its record with the position in cells and
// the temporal weight in place of the position and the time residual (the
// divisions are done once per particle while staging, not in the per-cell
// loop, whose lanes wait for the cell with the most particles).
__device__ inline void owner_stage(float (&v)[kOwnFields], float prev_dt, float cell) {
    const float tau = (-v[8]) / prev_dt;
    v[8] = temporal_w(tau);
    v[0] = v[0] / cell;
    v[1] = v[1] / cell;
}

// One particle of cell (cx, cy): the same arithmetic as p2g_contrib.
__device__ inline void owner_accumulate(const float* __restrict__ sp, int k, float cell, int32_t Wi,
                                        int32_t Hi, int32_t cx, int32_t cy, OwnerAcc& a) {
    const float gx = sp[0 * kOwnCap + k], gy = sp[1 * kOwnCap + k];
    const float vx = sp[2 * kOwnCap + k], vy = sp[3 * kOwnCap + k];
    const float cux = sp[4 * kOwnCap + k], cuy = sp[5 * kOwnCap + k];
    const float cvx = sp[6 * kOwnCap + k], cvy = sp[7 * kOwnCap + k];
    const float wt = sp[8 * kOwnCap + k];
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

struct P2gArgs {
    const float* pb;
    size_t nalloc;
    uint32_t np;
    const uint32_t* order;
    float prev_dt;
    float cell;
    uint32_t W, H;
    const uint2* cell_range;
    uint32_t tiles_x, tiles_y;  // blocks per row and column
    float* mac_u;
    float* mac_v;
    uint8_t* valid_u;
    uint8_t* valid_v;
    float* density;
    uint8_t* fluid;
    const uint8_t* obstacle;
    const float* ext_u;
    const float* ext_v;
    float* halo;  // kP2gHaloFloats per block
    int forces, ext;
    float dt, gravity, gravity_x;
    const StepState* ss;
};

// Whether every existing source cell of a target, columns [sx0, sx1] and rows
// [sy0, sy1] clipped to the grid, lies in the block at (x0, y0).
__device__ inline bool sources_in_block(int32_t sx0, int32_t sx1, int32_t sy0, int32_t sy1, int32_t Wi,
                                        int32_t Hi, int32_t x0, int32_t y0) {
    sx0 = max(sx0, 0); sx1 = min(sx1, Wi - 1);
    sy0 = max(sy0, 0); sy1 = min(sy1, Hi - 1);
    return sx0 >= x0 && sx1 < x0 + kBlk && sy0 >= y0 && sy1 < y0 + kBlk;
}

// Partial sums of one target over its source cells that lie in the block at
// (x0, y0) and in the grid, from the block's cell partials in LDS (slot-major,
// cells row-major), in the fixed order of the source offsets.
__device__ inline bool block_cell(int32_t sx, int32_t sy, int32_t Wi, int32_t Hi, int32_t x0, int32_t y0, int& at) {
    if (sx < 0 || sx >= Wi || sy < 0 || sy >= Hi) return false;
    const int32_t lx = sx - x0, ly = sy - y0;
    if (lx < 0 || lx >= kBlk || ly < 0 || ly >= kBlk) return false;
    at = ly * kBlk + lx;
    return true;
}
// u face (fx, fy): cells (fx - 1 + dx, fy - 1 + dy), slot (1 - dx) * 3 + (2 - dy)
__device__ inline void part_u(const float* sb, int32_t fx, int32_t fy, int32_t Wi, int32_t Hi, int32_t x0,
                              int32_t y0, float& m, float& w) {
    m = 0.0f; w = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            int at;
            if (!block_cell(fx - 1 + dx, fy - 1 + dy, Wi, Hi, x0, y0, at)) continue;
            const int s = (1 - dx) * 3 + (2 - dy);
            m += sb[(0 + s) * kBlkCells + at];
            w += sb[(6 + s) * kBlkCells + at];
        }
}
// v face (fx, fy): cells (fx - 1 + dx, fy - 1 + dy), slot (2 - dx) * 2 + (1 - dy)
__device__ inline void part_v(const float* sb, int32_t fx, int32_t fy, int32_t Wi, int32_t Hi, int32_t x0,
                              int32_t y0, float& m, float& w) {
    m = 0.0f; w = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            int at;
            if (!block_cell(fx - 1 + dx, fy - 1 + dy, Wi, Hi, x0, y0, at)) continue;
            const int s = (2 - dx) * 2 + (1 - dy);
            m += sb[(12 + s) * kBlkCells + at];
            w += sb[(18 + s) * kBlkCells + at];
        }
}
// cell (fx, fy): cells (fx - 1 + dx, fy - 1 + dy), slot (2 - dx) * 3 + (2 - dy)
__device__ inline float part_d(const float* sb, int32_t fx, int32_t fy, int32_t Wi, int32_t Hi, int32_t x0,
                               int32_t y0) {
    float d = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            int at;
            if (!block_cell(fx - 1 + dx, fy - 1 + dy, Wi, Hi, x0, y0, at)) continue;
            d += sb[(24 + (2 - dx) * 3 + (2 - dy)) * kBlkCells + at];
        }
    return d;
}

// The end of a face: the normalisation and valid flag (src/apic.rs:1428-1459)
// and the external forces (:1604-1626; horizontal gravity is an extension,
// zero in the reference).
__device__ inline void finish_u(const P2gArgs& a, int32_t fx, int32_t fy, float m, float w, float dt, bool left,
                                bool right) {
    const size_t f = (size_t)fy * (a.W + 1) + fx;
    const bool ok = w > 0.0f;
    float u = ok ? m / w : 0.0f;
    if (a.ext) u += a.ext_u[f] * dt;
    if (a.forces && a.gravity_x != 0.0f && (left || right)) u += a.gravity_x * dt;
    a.mac_u[f] = u;
    a.valid_u[f] = ok;
}
__device__ inline void finish_v(const P2gArgs& a, int32_t fx, int32_t fy, float m, float w, float dt, bool below,
                                bool above) {
    const size_t f = (size_t)fy * a.W + fx;
    const bool ok = w > 0.0f;
    float v = ok ? m / w : 0.0f;
    if (a.ext) v += a.ext_v[f] * dt;
    if (a.forces && (below || above)) v += a.gravity * dt;
    a.mac_v[f] = v;
    a.valid_v[f] = ok;
}

// the x (even) bits of a Morton offset within a block
__device__ inline int morton_even4(int t) { return (t & 1) | ((t >> 1) & 2) | ((t >> 2) & 4) | ((t >> 3) & 8); }

__device__ __forceinline__ void p2g_block(const P2gArgs& a, uint32_t wg) {
    __shared__ float s_buf[kOwnFields * kOwnCap];
    __shared__ uint32_t s_red[2 * (kBlkCells / 64)];
    __shared__ uint8_t s_fl[kBlkCells];  // fluid flags of the block's cells, row-major
    const float prev_dt = a.ss ? a.ss->prev_dt : a.prev_dt;
    const float dt = a.ss ? a.ss->dt : a.dt;
    const int32_t Wi = (int32_t)a.W, Hi = (int32_t)a.H;
    const uint32_t np = a.np;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t bx = wg % a.tiles_x, by = wg / a.tiles_x;
    const int32_t x0 = (int32_t)bx * kBlk, y0 = (int32_t)by * kBlk;
    // Accumulation: thread tid takes the cell of Morton offset tid (its cell
    // range is the tid-th of the block's 256 consecutive keys).
    const int lx = morton_even4(tid), ly = morton_even4(tid >> 1);
    const int32_t cx = x0 + lx, cy = y0 + ly;
    uint32_t b = 0, e = 0;
    bool fl = false;
    if (cx < Wi && cy < Hi) {
        const uint2 r = a.cell_range[(morton(bx, by) << 8) | (uint32_t)tid];
        if (r.y > r.x && r.y <= np) { b = r.x; e = r.y; }
        fl = e > b && !a.obstacle[(size_t)cy * a.W + cx];
    }
    // the block's run [run0, run1) in cell order
    uint32_t mn = e > b ? b : 0xFFFFFFFFu, mx = e > b ? e : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor((int)mn, o, 64));
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, 64));
    }
    if (lane == 0) { s_red[wave] = mn; s_red[kBlkCells / 64 + wave] = mx; }
    s_fl[ly * kBlk + lx] = fl;
    __syncthreads();
    uint32_t run0 = s_red[0], run1 = s_red[kBlkCells / 64];
#pragma unroll
    for (int w = 1; w < kBlkCells / 64; ++w) {
        run0 = min(run0, s_red[w]);
        run1 = max(run1, s_red[kBlkCells / 64 + w]);
    }
    const uint32_t total = run1 > run0 ? run1 - run0 : 0u;
    const uint32_t my0 = e > b ? b - run0 : 0u, my1 = e > b ? e - run0 : 0u;

    OwnerAcc acc;
#pragma unroll
    for (int s = 0; s < 6; ++s) { acc.um[s] = 0.0f; acc.uw[s] = 0.0f; acc.vm[s] = 0.0f; acc.vw[s] = 0.0f; }
#pragma unroll
    for (int s = 0; s < 9; ++s) acc.d[s] = 0.0f;
    for (uint32_t base = 0; base < total; base += kOwnCap) {
        const uint32_t n = min((uint32_t)kOwnCap, total - base);
        // a thread's kOwnCap / kBlkCells slots: their positions in cell order,
        // then the permutation entries, then the fields, each stage's loads
        // issued together
        constexpr int kPer = kOwnCap / kBlkCells;
        static_assert(kPer * kBlkCells == kOwnCap, "whole slots per thread");
        uint32_t src[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t k = (uint32_t)tid + (uint32_t)(j * kBlkCells);
            src[j] = run0 + base + (k < n ? k : 0u);
        }
        if (a.order) {
            // (a permutation entry past the records, written through the
            // field API, reads the last record, as the sort's gather does)
#pragma unroll
            for (int j = 0; j < kPer; ++j) src[j] = min(a.order[src[j]], np - 1);
        }
        float v[kPer][kOwnFields];
#pragma unroll
        for (int j = 0; j < kPer; ++j)
#pragma unroll
            for (int f = 0; f < kOwnFields; ++f) v[j][f] = a.pb[(size_t)f * a.nalloc + src[j]];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t k = (uint32_t)tid + (uint32_t)(j * kBlkCells);
            if (k < n) {
                owner_stage(v[j], prev_dt, a.cell);
#pragma unroll
                for (int f = 0; f < kOwnFields; ++f) s_buf[f * kOwnCap + k] = v[j][f];
            }
        }
        __syncthreads();
        const uint32_t lo = max(my0, base), hi = min(my1, base + n);
        for (uint32_t idx = lo; idx < hi; ++idx)
            owner_accumulate(s_buf, (int)(idx - base), a.cell, Wi, Hi, cx, cy, acc);
        __syncthreads();
    }
    // partials of every block cell, slot-major, cells row-major
    const int at = ly * kBlk + lx;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        s_buf[(0 + s) * kBlkCells + at] = acc.um[s];
        s_buf[(6 + s) * kBlkCells + at] = acc.uw[s];
        s_buf[(12 + s) * kBlkCells + at] = acc.vm[s];
        s_buf[(18 + s) * kBlkCells + at] = acc.vw[s];
    }
#pragma unroll
    for (int s = 0; s < 9; ++s) s_buf[(24 + s) * kBlkCells + at] = acc.d[s];
    __syncthreads();

    // Targets: thread tid takes the cell (x0 + tid % 16, y0 + tid / 16), its u
    // and v faces (and the faces x = W, y = H of the last column and row).
    const int tx = tid % kBlk, ty = tid / kBlk;
    const int32_t gx = x0 + tx, gy = y0 + ty;
    auto fl_at = [&](int32_t x, int32_t y) -> bool {  // (x, y) in the block or outside the grid
        return x >= 0 && x < Wi && y >= 0 && y < Hi && s_fl[(y - y0) * kBlk + (x - x0)];
    };
    if (gx < Wi && gy < Hi) {
        a.fluid[(size_t)gy * a.W + gx] = s_fl[tid];
        float m, w;
        for (int32_t fx = gx; fx <= (gx == Wi - 1 ? Wi : gx); ++fx)
            if (sources_in_block(fx - 1, fx, gy - 1, gy + 1, Wi, Hi, x0, y0)) {
                part_u(s_buf, fx, gy, Wi, Hi, x0, y0, m, w);
                finish_u(a, fx, gy, m, w, dt, fl_at(fx - 1, gy), fl_at(fx, gy));
            }
        for (int32_t fy = gy; fy <= (gy == Hi - 1 ? Hi : gy); ++fy)
            if (sources_in_block(gx - 1, gx + 1, fy - 1, fy, Wi, Hi, x0, y0)) {
                part_v(s_buf, gx, fy, Wi, Hi, x0, y0, m, w);
                finish_v(a, gx, fy, m, w, dt, fl_at(gx, fy - 1), fl_at(gx, fy));
            }
        if (sources_in_block(gx - 1, gx + 1, gy - 1, gy + 1, Wi, Hi, x0, y0))
            a.density[(size_t)gy * a.W + gx] = part_d(s_buf, gx, gy, Wi, Hi, x0, y0);
    }
    // the halo record (zeros for a block without particles)
    float* rec = a.halo + (size_t)wg * kP2gHaloFloats;
    if (total == 0) {
        for (int q = tid; q < (int)kP2gHaloFloats; q += kBlkCells) rec[q] = 0.0f;
        return;
    }
    for (int q = tid; q < 320; q += kBlkCells) {
        int i, j;
        if (q < 96) {
            halo_u_pos(q, i, j);
            float m, w;
            part_u(s_buf, x0 + i, y0 - 1 + j, Wi, Hi, x0, y0, m, w);
            rec[kHaloUm + q] = m;
            rec[kHaloUw + q] = w;
        } else if (q < 192) {
            halo_v_pos(q - 96, i, j);
            float m, w;
            part_v(s_buf, x0 - 1 + i, y0 + j, Wi, Hi, x0, y0, m, w);
            rec[kHaloVm + q - 96] = m;
            rec[kHaloVw + q - 96] = w;
        } else {
            halo_d_pos(q - 192, i, j);
            rec[kHaloDen + q - 192] = part_d(s_buf, x0 - 1 + i, y0 - 1 + j, Wi, Hi, x0, y0);
        }
    }
}

// The targets of a block that P2gBlock left open: the band around its edge
// whose sources reach into other blocks.  Each wave takes one kind of target:
//   wave 0, lane q < 62: u faces: column 0, column 16 (the face x = W), rows
//                        0 and 15 of columns 1-15;
//   wave 1, lane q < 62: v faces: columns 0 and 15, row 0 of columns 1-14,
//                        row 16 (the face y = H);
//   wave 2, lane q < 60: cells: columns 0 and 15, rows 0 and 15 of columns
//                        1-14;
// and sums the halo records of the blocks that reach the target, in the order
// of their block rows, then columns.
constexpr int kSeamThreads = 192;
template <int Kind>
__device__ __forceinline__ void seam_target(const P2gArgs& a, uint32_t wg, int q) {
    const int32_t Wi = (int32_t)a.W, Hi = (int32_t)a.H;
    const int32_t bx = (int32_t)(wg % a.tiles_x), by = (int32_t)(wg / a.tiles_x);
    const int32_t x0 = bx * kBlk, y0 = by * kBlk;
    int32_t fx, fy, sx0, sx1, sy0, sy1;
    bool valid;
    if (Kind == 0) {
        int i, t;
        if (q < 16) { i = 0; t = q; }
        else if (q < 32) { i = 16; t = q - 16; }
        else if (q < 47) { i = q - 31; t = 0; }
        else { i = q - 46; t = 15; }
        fx = x0 + i; fy = y0 + t;
        sx0 = fx - 1; sx1 = fx; sy0 = fy - 1; sy1 = fy + 1;
        valid = q < 62 && fy < Hi && fx <= Wi && (i < 16 || fx == Wi);
    } else if (Kind == 1) {
        int t, j;
        if (q < 16) { t = 0; j = q; }
        else if (q < 32) { t = 15; j = q - 16; }
        else if (q < 46) { t = q - 31; j = 0; }
        else { t = q - 46; j = 16; }
        fx = x0 + t; fy = y0 + j;
        sx0 = fx - 1; sx1 = fx + 1; sy0 = fy - 1; sy1 = fy;
        valid = q < 62 && fx < Wi && fy <= Hi && (j < 16 || fy == Hi);
    } else {
        int t, u;
        if (q < 16) { t = 0; u = q; }
        else if (q < 32) { t = 15; u = q - 16; }
        else if (q < 46) { t = q - 31; u = 0; }
        else { t = q - 45; u = 15; }
        fx = x0 + t; fy = y0 + u;
        sx0 = fx - 1; sx1 = fx + 1; sy0 = fy - 1; sy1 = fy + 1;
        valid = q < 60 && fx < Wi && fy < Hi;
    }
    if (!valid || sources_in_block(sx0, sx1, sy0, sy1, Wi, Hi, x0, y0)) return;  // P2gBlock finished it
    const float dt = a.ss ? a.ss->dt : a.dt;
    // The blocks that reach the target hold one of its source cells (the
    // grid's or one past it): block rows r0..r1 and columns c0..c1, at most
    // two each.  Their records and the fluid flags beside a face are loaded
    // at once.
    const int32_t c0 = sx0 >> 4, c1 = sx1 >> 4, r0 = sy0 >> 4, r1 = sy1 >> 4;  // (arithmetic shifts: -1 -> -1)
    static_assert(kBlk == 16, "block index shifts");
    constexpr int hm = Kind == 0 ? kHaloUm : Kind == 1 ? kHaloVm : kHaloDen;
    constexpr int hw = Kind == 0 ? kHaloUw : kHaloVw;
    bool ok[4];
    float vm[4], vw[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t ry = (k >> 1) ? r1 : r0, rx = (k & 1) ? c1 : c0;
        ok[k] = ((k & 1) == 0 || c1 != c0) && ((k >> 1) == 0 || r1 != r0) && rx >= 0 &&
                rx < (int32_t)a.tiles_x && ry >= 0 && ry < (int32_t)a.tiles_y;
        const uint32_t nb = ok[k] ? (uint32_t)ry * a.tiles_x + (uint32_t)rx : wg;
        const int32_t nx0 = rx * kBlk, ny0 = ry * kBlk;
        int h;
        if (Kind == 0) h = halo_u_index(fx - nx0, fy - (ny0 - 1));
        else if (Kind == 1) h = halo_v_index(fx - (nx0 - 1), fy - ny0);
        else h = halo_d_index(fx - (nx0 - 1), fy - (ny0 - 1));
        const float* rec = a.halo + (size_t)nb * kP2gHaloFloats + (ok[k] ? h : 0);
        vm[k] = rec[hm];
        if (Kind != 2) vw[k] = rec[hw];
    }
    if (Kind == 2) {
        float d = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (ok[k]) d += vm[k];
        a.density[(size_t)fy * a.W + fx] = d;
        return;
    }
    // fluid flags beside a face: (fx - 1, fy) and (fx, fy) for u, (fx, fy - 1)
    // and (fx, fy) for v (cells outside the grid are not fluid)
    const int32_t ax = Kind == 0 ? fx - 1 : fx, ay = Kind == 1 ? fy - 1 : fy;
    const bool a_in = ax >= 0 && ax < Wi && ay >= 0 && ay < Hi, b_in = fx < Wi && fy < Hi;
    const uint8_t fa = a.fluid[a_in ? (size_t)ay * a.W + ax : 0];
    const uint8_t fb = a.fluid[b_in ? (size_t)fy * a.W + fx : 0];
    float m = 0.0f, w = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (ok[k]) {
            m += vm[k];
            w += vw[k];
        }
    }
    if (Kind == 0) finish_u(a, fx, fy, m, w, dt, a_in && fa, b_in && fb);
    else finish_v(a, fx, fy, m, w, dt, a_in && fa, b_in && fb);
}

__device__ __forceinline__ void p2g_seam(const P2gArgs& a, uint32_t wg) {
    const int wave = threadIdx.x >> 6, q = threadIdx.x & 63;
    if (wave == 0) seam_target<0>(a, wg, q);
    else if (wave == 1) seam_target<1>(a, wg, q);
    else seam_target<2>(a, wg, q);
}

struct P2gBlock {
    static constexpr int kThreads = kBlkCells;
    static constexpr bool kRoundRobin = true;  // work follows the particles
    using Args = P2gArgs;
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) { p2g_block(a, wg); }
};

struct P2gSeam {
    static constexpr int kThreads = kSeamThreads;
    static constexpr bool kDealRows = true;
    using Args = P2gArgs;
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) { p2g_seam(a, wg); }
};

// [begin, end) of every cell's particles in cell order, indexed by key
struct CellRanges {
    static constexpr int kThreads = 256;
    struct Args {
        const uint32_t* keys;
        uint32_t np;
        uint2* range;
        uint32_t nkeys;
        uint32_t nwg;
    };
    // (a grid-stride loop over at most kMaxWorkgroups workgroups: one
    // workgroup per 256 keys would be dispatch-bound at tens of millions)
    static constexpr uint32_t kMaxWorkgroups = 8192;
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        const uint32_t np = a.np;
        const uint32_t* __restrict__ keys = a.keys;
        const uint32_t stride = a.nwg * kThreads;
        for (uint32_t i = wg * kThreads + threadIdx.x; i < np; i += stride) {
            const uint32_t k = keys[i];
            const uint32_t kp = i > 0 ? keys[i - 1] : ~0u, kn = i + 1 < np ? keys[i + 1] : ~0u;
            if (k >= a.nkeys) continue;  // keys come from clamped cell coordinates; kept as a guard
            if (i == 0 || kp != k) a.range[k].x = i;
            if (i + 1 == np || kn != k) a.range[k].y = i + 1;
        }
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
    const uint32_t nwg = std::min<uint32_t>((sim.num_particles + 255) / 256, CellRanges::kMaxWorkgroups);
    emit<CellRanges>(sim, {sim.keys[0], sim.num_particles, sim.cell_range, (uint32_t)nk, nwg}, nwg);
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
    const uint32_t tx = p2g_blocks(sim.W), ty = p2g_blocks(sim.H);
    const int ext = forces && sim.forces_active;
    sim.mark(APIC_T_K_P2G, 0);
    P2gArgs a{sim.pbuf[sim.cur], sim.nalloc, sim.num_particles, order, sim.prev_dt, sim.cell, sim.W, sim.H,
              sim.cell_range, tx, ty, sim.mac_u, sim.mac_v, sim.valid_u, sim.valid_v, sim.density,
              sim.levels[0].fluid, sim.obstacle, sim.ext_u, sim.ext_v, sim.p2g_halo, (int)forces, ext,
              dt, sim.props.gravity, sim.gravity_x, sim.step_src};
    emit<P2gBlock>(sim, a, tx * ty);
    emit<P2gSeam>(sim, a, tx * ty);
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
