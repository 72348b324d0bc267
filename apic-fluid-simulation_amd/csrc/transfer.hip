// transfer.hip -- particle<->grid transfers of the APIC step.
//
//  p2g_kernel      transfer_particles_to_grid's scatter (src/apic.rs:1237-1426).
//                  A workgroup takes 256 consecutive particles.  After the
//                  Morton sort, particles of one cell are consecutive, so each
//                  lane writes its contributions into a fixed window of slots
//                  around its cell (2x3 u faces, 3x2 v faces, 3x3 density
//                  cells) and a 64-lane segmented scan sums the window over
//                  each run of same-cell lanes; only the last lane of a run
//                  adds the run's sums into LDS tiles covering the
//                  workgroup's bounding box.  The tiles are flushed with one
//                  global float atomic per nonzero entry.  Particles next to
//                  the domain edge (where the reference clamps stencil
//                  indices) add their own contributions; when the box does
//                  not fit the tiles (unsorted particles, sort disabled) the
//                  workgroup scatters with global atomics.
//  normalize_forces_kernel
//                  u = momentum / weight and the valid flags (:1428-1459),
//                  fused with apply_external_forces (:1604-1626).
//  g2p_kernel      transfer_grid_to_particles_and_advect (:2057-2264): masked
//                  bilinear gather, jittered time step, wall clamp/reflect,
//                  affine C rebuild; fused with the next sort's keygen and the
//                  max |v| reduction that sets the next substep's dt
//                  (get_max_particle_velocity, :2412-2427).
#include "apic_internal.hpp"
#include "device_math.hpp"

namespace apic {

namespace {

constexpr int kP2GThreads = 256;
constexpr int kTileCap = 1024;

__device__ inline int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ void clear_grid_kernel(float2* __restrict__ u_acc, size_t nu, float2* __restrict__ v_acc, size_t nv,
                                  float* __restrict__ density, uint32_t* __restrict__ fluid_words, size_t cells,
                                  size_t fluid_words_n, FluidStats* fstats) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = t; i < nu; i += stride) u_acc[i] = make_float2(0.0f, 0.0f);
    for (size_t i = t; i < nv; i += stride) v_acc[i] = make_float2(0.0f, 0.0f);
    for (size_t i = t; i < cells; i += stride) density[i] = 0.0f;
    for (size_t i = t; i < fluid_words_n; i += stride) fluid_words[i] = 0u;
    if (t == 0) { fstats->count = 0; fstats->last_raw = 0; }
}

// Value of lane - n within the same 16-lane row (DPP row_shr:n, no LDS
// traffic); lanes whose source would cross the row start get 0.
template <int N>
__device__ inline float row_shr(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 | N, 0xF, 0xF, true));
}
template <int N>
__device__ inline int row_shr_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x110 | N, 0xF, 0xF, true);
}

// Segmented inclusive sum inside each 16-lane row: after the call, v holds
// the sum of its segment's values from the segment head up to this lane.
// Segments never cross a row boundary (the caller starts one at every row).
template <int N>
__device__ inline void row_segmented_scan(float (&v)[N], bool head, int lane) {
    int f = head ? 1 : 0;
    const int r = lane & 15;
#define APIC_SEG_STEP(O)                                              \
    {                                                                 \
        int tf = row_shr_i<O>(f);                                     \
        _Pragma("unroll") for (int k = 0; k < N; ++k) {               \
            float t = row_shr<O>(v[k]);                               \
            if (r >= O && !f) v[k] += t;                              \
        }                                                             \
        if (r >= O) f |= tf;                                          \
    }
    APIC_SEG_STEP(1)
    APIC_SEG_STEP(2)
    APIC_SEG_STEP(4)
    APIC_SEG_STEP(8)
#undef APIC_SEG_STEP
}

__device__ inline float pick2(int s, int base, float a, float b) { return s == base ? a : (s == base + 1 ? b : 0.0f); }

__global__ __launch_bounds__(kP2GThreads) void p2g_kernel(
    const float* __restrict__ pb, size_t nalloc, uint32_t np, float prev_dt, float cell, uint32_t W, uint32_t H,
    float2* __restrict__ u_acc, float2* __restrict__ v_acc, float* __restrict__ density,
    uint8_t* __restrict__ fluid, const uint8_t* __restrict__ obstacle) {
    __shared__ float s_umom[kTileCap], s_uw[kTileCap], s_vmom[kTileCap], s_vw[kTileCap], s_den[kTileCap];
    __shared__ int s_box[4][kP2GThreads / 64];

    const uint32_t i = blockIdx.x * kP2GThreads + threadIdx.x;
    const bool active = i < np;
    const int32_t Wi = (int32_t)W, Hi = (int32_t)H;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    float uadd[4], uwt[4], vadd[4], vwt[4], cw[4];
    int32_t ux[2], uy[2], vx_[2], vy_[2], cx[2], cy[2];
    int32_t kx = 0, ky = 0;
    bool regular = false;
    int lo_x = INT32_MAX, hi_x = INT32_MIN, lo_y = INT32_MAX, hi_y = INT32_MIN;
    if (active) {
        const float x = pb[(size_t)P_X * nalloc + i], y = pb[(size_t)P_Y * nalloc + i];
        const float vx = pb[(size_t)V_X * nalloc + i], vy = pb[(size_t)V_Y * nalloc + i];
        const float cux = pb[(size_t)CU_X * nalloc + i], cuy = pb[(size_t)CU_Y * nalloc + i];
        const float cvx = pb[(size_t)CV_X * nalloc + i], cvy = pb[(size_t)CV_Y * nalloc + i];
        const float tres = pb[(size_t)T_RES * nalloc + i];
        const float tau = (-tres) / prev_dt;
        const float wt = temporal_w(tau);
        const float gx = x / cell, gy = y / cell;
        // density stencil (:1260-1284)
        {
            float cgx = gx - 0.5f, cgy = gy - 0.5f;
            float cbx = floorf(cgx), cby = floorf(cgy);
            float ctx = cgx - cbx, cty = cgy - cby;
            cw[0] = ((1.0f - ctx) * (1.0f - cty)) * wt;
            cw[1] = (ctx * (1.0f - cty)) * wt;
            cw[2] = ((1.0f - ctx) * cty) * wt;
            cw[3] = (ctx * cty) * wt;
            int32_t bx = (int32_t)cbx, by = (int32_t)cby;
            cx[0] = clampi(bx, 0, Wi - 1); cx[1] = clampi(bx + 1, 0, Wi - 1);
            cy[0] = clampi(by, 0, Hi - 1); cy[1] = clampi(by + 1, 0, Hi - 1);
        }
        // u faces
        {
            Stencil s;
            make_stencil(gx, gy - 0.5f, cell, s);
            ux[0] = clampi(s.ix, 0, Wi); ux[1] = clampi(s.ix + 1, 0, Wi);
            uy[0] = clampi(s.iy, 0, Hi - 1); uy[1] = clampi(s.iy + 1, 0, Hi - 1);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float ww = s.w[k] * wt;
                uwt[k] = ww;
                uadd[k] = ((vx + cux * s.dx[k]) + cuy * s.dy[k]) * ww;
            }
        }
        // v faces
        {
            Stencil s;
            make_stencil(gx - 0.5f, gy, cell, s);
            vx_[0] = clampi(s.ix, 0, Wi - 1); vx_[1] = clampi(s.ix + 1, 0, Wi - 1);
            vy_[0] = clampi(s.iy, 0, Hi); vy_[1] = clampi(s.iy + 1, 0, Hi);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float ww = s.w[k] * wt;
                vwt[k] = ww;
                vadd[k] = ((vy + cvx * s.dx[k]) + cvy * s.dy[k]) * ww;
            }
        }
        // fluid bit of the particle's cell (:1358-1360, 1409; obstacle cells are
        // cleared from the fluid mask afterwards, :1464)
        {
            float clx = fast_max(fast_min(gx, (float)W - 1.0f), 0.0f);
            float cly = fast_max(fast_min(gy, (float)H - 1.0f), 0.0f);
            uint32_t fi = (uint32_t)((int32_t)cly * Wi + (int32_t)clx);
            if (!obstacle[fi]) fluid[fi] = 1;
        }
        // inside [1, W-1) x [1, H-1) no stencil index is clamped, and every
        // target lies in the 3x3 window around the particle's cell
        regular = gx >= 1.0f && gx < (float)(Wi - 1) && gy >= 1.0f && gy < (float)(Hi - 1);
        kx = (int32_t)floorf(gx);
        ky = (int32_t)floorf(gy);
        lo_x = min(min(ux[0], vx_[0]), cx[0]); hi_x = max(max(ux[1], vx_[1]), cx[1]);
        lo_y = min(min(uy[0], vy_[0]), cy[0]); hi_y = max(max(uy[1], vy_[1]), cy[1]);
    }
    // workgroup bounding box of all targets
    lo_x = wave_min(lo_x); hi_x = wave_max(hi_x); lo_y = wave_min(lo_y); hi_y = wave_max(hi_y);
    if (lane == 0) { s_box[0][wave] = lo_x; s_box[1][wave] = hi_x; s_box[2][wave] = lo_y; s_box[3][wave] = hi_y; }
    __syncthreads();
    lo_x = s_box[0][0]; hi_x = s_box[1][0]; lo_y = s_box[2][0]; hi_y = s_box[3][0];
#pragma unroll
    for (int w = 1; w < kP2GThreads / 64; ++w) {
        lo_x = min(lo_x, s_box[0][w]); hi_x = max(hi_x, s_box[1][w]);
        lo_y = min(lo_y, s_box[2][w]); hi_y = max(hi_y, s_box[3][w]);
    }
    if (hi_x < lo_x) return;  // no active particle in this workgroup (uniform)
    const int tw = hi_x - lo_x + 1, th = hi_y - lo_y + 1;
    const uint32_t W1 = W + 1;
    if ((long long)tw * th <= kTileCap) {
        const int n = tw * th;
        for (int e = threadIdx.x; e < n; e += kP2GThreads) {
            s_umom[e] = 0.0f; s_uw[e] = 0.0f; s_vmom[e] = 0.0f; s_vw[e] = 0.0f; s_den[e] = 0.0f;
        }
        __syncthreads();
        // window slots around cell (kx, ky):
        //   u faces (kx + sx, ky - 1 + sy), sx in 0..1, sy in 0..2   -> slot sx*3 + sy
        //   v faces (kx - 1 + sx, ky + sy), sx in 0..2, sy in 0..1   -> slot sx*2 + sy
        //   cells   (kx - 1 + sx, ky - 1 + sy), sx, sy in 0..2        -> slot sx*3 + sy
        float v[33];
        {
            const int du = uy[0] - (ky - 1);   // 0 or 1
            const int dv = vx_[0] - (kx - 1);  // 0 or 1
            const int dcx = cx[0] - (kx - 1), dcy = cy[0] - (ky - 1);
#pragma unroll
            for (int sx = 0; sx < 2; ++sx)
#pragma unroll
                for (int sy = 0; sy < 3; ++sy) {
                    v[sx * 3 + sy] = regular ? pick2(sy, du, uadd[sx], uadd[2 + sx]) : 0.0f;
                    v[6 + sx * 3 + sy] = regular ? pick2(sy, du, uwt[sx], uwt[2 + sx]) : 0.0f;
                }
#pragma unroll
            for (int sx = 0; sx < 3; ++sx)
#pragma unroll
                for (int sy = 0; sy < 2; ++sy) {
                    v[12 + sx * 2 + sy] = regular ? pick2(sx, dv, vadd[2 * sy], vadd[2 * sy + 1]) : 0.0f;
                    v[18 + sx * 2 + sy] = regular ? pick2(sx, dv, vwt[2 * sy], vwt[2 * sy + 1]) : 0.0f;
                }
#pragma unroll
            for (int sx = 0; sx < 3; ++sx)
#pragma unroll
                for (int sy = 0; sy < 3; ++sy) {
                    float a = pick2(sy, dcy, cw[0], cw[2]);
                    float b = pick2(sy, dcy, cw[1], cw[3]);
                    v[24 + sx * 3 + sy] = regular ? pick2(sx, dcx, a, b) : 0.0f;
                }
        }
        // runs of same-cell regular lanes, cut at 16-lane row boundaries, form the segments
        const uint32_t key = regular ? (uint32_t)(ky * Wi + kx) : (0x80000000u | (uint32_t)threadIdx.x);
        const uint32_t prev_key = (uint32_t)row_shr_i<1>((int)key);
        const uint32_t next_key = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)key, 0x100 | 1, 0xF, 0xF, true);
        const bool head = (lane & 15) == 0 || prev_key != key;
        const bool tail = (lane & 15) == 15 || next_key != key;
        row_segmented_scan(v, head, lane);
        if (regular && tail) {
#pragma unroll
            for (int sx = 0; sx < 2; ++sx)
#pragma unroll
                for (int sy = 0; sy < 3; ++sy) {
                    int e = (ky - 1 + sy - lo_y) * tw + (kx + sx - lo_x);
                    float m = v[sx * 3 + sy], w = v[6 + sx * 3 + sy];
                    if (m != 0.0f) atomicAdd(&s_umom[e], m);
                    if (w != 0.0f) atomicAdd(&s_uw[e], w);
                }
#pragma unroll
            for (int sx = 0; sx < 3; ++sx)
#pragma unroll
                for (int sy = 0; sy < 2; ++sy) {
                    int e = (ky + sy - lo_y) * tw + (kx - 1 + sx - lo_x);
                    float m = v[12 + sx * 2 + sy], w = v[18 + sx * 2 + sy];
                    if (m != 0.0f) atomicAdd(&s_vmom[e], m);
                    if (w != 0.0f) atomicAdd(&s_vw[e], w);
                }
#pragma unroll
            for (int sx = 0; sx < 3; ++sx)
#pragma unroll
                for (int sy = 0; sy < 3; ++sy) {
                    int e = (ky - 1 + sy - lo_y) * tw + (kx - 1 + sx - lo_x);
                    float d = v[24 + sx * 3 + sy];
                    if (d != 0.0f) atomicAdd(&s_den[e], d);
                }
        }
        if (active && !regular) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int e = (uy[k >> 1] - lo_y) * tw + (ux[k & 1] - lo_x);
                atomicAdd(&s_umom[e], uadd[k]);
                atomicAdd(&s_uw[e], uwt[k]);
                e = (vy_[k >> 1] - lo_y) * tw + (vx_[k & 1] - lo_x);
                atomicAdd(&s_vmom[e], vadd[k]);
                atomicAdd(&s_vw[e], vwt[k]);
                e = (cy[k >> 1] - lo_y) * tw + (cx[k & 1] - lo_x);
                atomicAdd(&s_den[e], cw[k]);
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < n; e += kP2GThreads) {
            const int ex = lo_x + e % tw, ey = lo_y + e / tw;
            float m = s_umom[e], w = s_uw[e];
            if ((m != 0.0f || w != 0.0f) && ex <= Wi && ey < Hi) {
                float2* a = &u_acc[(size_t)ey * W1 + ex];
                unsafeAtomicAdd(&a->x, m);
                unsafeAtomicAdd(&a->y, w);
            }
            m = s_vmom[e]; w = s_vw[e];
            if ((m != 0.0f || w != 0.0f) && ex < Wi && ey <= Hi) {
                float2* a = &v_acc[(size_t)ey * W + ex];
                unsafeAtomicAdd(&a->x, m);
                unsafeAtomicAdd(&a->y, w);
            }
            float d = s_den[e];
            if (d != 0.0f && ex < Wi && ey < Hi) unsafeAtomicAdd(&density[(size_t)ey * W + ex], d);
        }
    } else if (active) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (uadd[k] != 0.0f || uwt[k] != 0.0f) {
                float2* a = &u_acc[(size_t)uy[k >> 1] * W1 + ux[k & 1]];
                unsafeAtomicAdd(&a->x, uadd[k]);
                unsafeAtomicAdd(&a->y, uwt[k]);
            }
            if (vadd[k] != 0.0f || vwt[k] != 0.0f) {
                float2* a = &v_acc[(size_t)vy_[k >> 1] * W + vx_[k & 1]];
                unsafeAtomicAdd(&a->x, vadd[k]);
                unsafeAtomicAdd(&a->y, vwt[k]);
            }
            if (cw[k] != 0.0f) unsafeAtomicAdd(&density[(size_t)cy[k >> 1] * W + cx[k & 1]], cw[k]);
        }
    }
}

__global__ void normalize_forces_kernel(const float2* __restrict__ u_acc, const float2* __restrict__ v_acc,
                                        float* __restrict__ mac_u, float* __restrict__ mac_v,
                                        uint8_t* __restrict__ valid_u, uint8_t* __restrict__ valid_v,
                                        const float* __restrict__ ext_u, const float* __restrict__ ext_v,
                                        const uint8_t* __restrict__ fluid, uint32_t W, uint32_t H, int normalize,
                                        int ext_forces, int gravity_on, float dt, float g) {
    const size_t nu = (size_t)W * H + H, nv = (size_t)W * H + W;
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nu) {
        float u;
        if (normalize) {
            float2 a = u_acc[f];
            bool ok = a.y > 0.0f;
            u = ok ? a.x / a.y : 0.0f;
            valid_u[f] = ok;
        } else {
            u = mac_u[f];
        }
        if (ext_forces) u += ext_u[f] * dt;
        mac_u[f] = u;
    }
    if (f < nv) {
        float v;
        if (normalize) {
            float2 a = v_acc[f];
            bool ok = a.y > 0.0f;
            v = ok ? a.x / a.y : 0.0f;
            valid_v[f] = ok;
        } else {
            v = mac_v[f];
        }
        if (ext_forces) v += ext_v[f] * dt;
        if (gravity_on) {
            uint32_t x = (uint32_t)(f % W), y = (uint32_t)(f / W);
            bool below = y >= 1 && fluid[(size_t)(y - 1) * W + x];
            bool above = y < H && fluid[(size_t)y * W + x];
            if (below || above) v += g;
        }
        mac_v[f] = v;
    }
}

__device__ inline float masked_interp(const float* w, const uint8_t* open, const uint32_t* ids, const float* nodes) {
    float mw[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) mw[k] = open[ids[k]] ? w[k] : 0.0f;
    // bilinear_interpolate (src/apic.rs:138-172)
    float ws = ((mw[0] + mw[1]) + mw[2]) + mw[3];
    float inv = 1.0f / fast_max(ws, 1e-6f);
    float vel = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) vel += (nodes[k] * mw[k]) * inv;
    return vel;
}

__global__ __launch_bounds__(256) void g2p_kernel(
    float* __restrict__ pb, size_t nalloc, uint32_t np, float dt, float cell, uint32_t W, uint32_t H,
    uint32_t timestamp, float border_damping, const float* __restrict__ mac_u, const float* __restrict__ mac_v,
    const uint8_t* __restrict__ open_u, const uint8_t* __restrict__ open_v,
    int write_keys, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, float* __restrict__ part_maxv) {
    __shared__ float s_max[256 / 64];
    float speed_new = 0.0f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nalloc; i += (size_t)gridDim.x * blockDim.x) {
        const float x = pb[(size_t)P_X * nalloc + i], y = pb[(size_t)P_Y * nalloc + i];
        const float res = pb[(size_t)T_RES * nalloc + i];
        const float gx = x / cell, gy = y / cell;
        Stencil su, sv;
        make_stencil(gx, gy - 0.5f, cell, su);
        make_stencil(gx - 0.5f, gy, cell, sv);
        uint32_t ui[4] = {u_index(su.ix, su.iy, W, H), u_index(su.ix + 1, su.iy, W, H),
                          u_index(su.ix, su.iy + 1, W, H), u_index(su.ix + 1, su.iy + 1, W, H)};
        uint32_t vi[4] = {v_index(sv.ix, sv.iy, W, H), v_index(sv.ix + 1, sv.iy, W, H),
                          v_index(sv.ix, sv.iy + 1, W, H), v_index(sv.ix + 1, sv.iy + 1, W, H)};
        float un[4], vn[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) { un[k] = mac_u[ui[k]]; vn[k] = mac_v[vi[k]]; }
        float nvx = masked_interp(su.w, open_u, ui, un);
        float nvy = masked_interp(sv.w, open_v, vi, vn);
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
        pb[(size_t)P_X * nalloc + i] = nx;
        pb[(size_t)P_Y * nalloc + i] = ny;
        pb[(size_t)V_X * nalloc + i] = nvx;
        pb[(size_t)V_Y * nalloc + i] = nvy;
        pb[(size_t)CU_X * nalloc + i] = c_ux;
        pb[(size_t)CU_Y * nalloc + i] = c_uy;
        pb[(size_t)CV_X * nalloc + i] = c_vx;
        pb[(size_t)CV_Y * nalloc + i] = c_vy;
        pb[(size_t)T_RES * nalloc + i] = new_res;
        speed_new = fast_max(speed_new, sqrtf(nvx * nvx + nvy * nvy));
        if (write_keys && i < np) {
            // keygen of the next update_spatial_lookup (:1049-1064)
            float kx = nx / cell, ky = ny / cell;
            uint32_t ix = (uint32_t)(int32_t)fast_min(fast_max(kx, 0.0f), (float)(W - 1));
            uint32_t iy = (uint32_t)(int32_t)fast_min(fast_max(ky, 0.0f), (float)(H - 1));
            keys[i] = morton(ix, iy);
            vals[i] = (uint32_t)i;
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
        part_maxv[blockIdx.x] = b;
    }
}

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

// max over the per-workgroup maxima into the float bits read by the host
__global__ __launch_bounds__(1024) void maxv_finalize_kernel(const float* __restrict__ part_maxv, uint32_t n,
                                                             uint32_t* __restrict__ maxv_bits) {
    __shared__ float s_max[1024 / 64];
    float m = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) m = fast_max(m, part_maxv[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fast_max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) b = fast_max(b, s_max[w]);
        *maxv_bits = __float_as_uint(b);
    }
}

unsigned particle_blocks(const Sim& sim) {
    size_t b = (sim.nalloc + 255) / 256;
    return (unsigned)std::min<size_t>(std::max<size_t>(b, 1), sim.part_maxv_cap);
}

}  // namespace

void launch_p2g(Sim& sim) {
    size_t fw = (sim.cells + 3) / 4;
    hipLaunchKernelGGL(clear_grid_kernel, dim3(1024), dim3(256), 0, sim.stream, sim.u_acc, sim.nu(), sim.v_acc,
                       sim.nv(), sim.density, reinterpret_cast<uint32_t*>(sim.levels[0].fluid), (size_t)sim.cells, fw,
                       sim.fstats);
    if (sim.num_particles == 0) return;
    uint32_t blocks = (sim.num_particles + kP2GThreads - 1) / kP2GThreads;
    sim.mark(APIC_T_K_P2G, 0);
    hipLaunchKernelGGL(p2g_kernel, dim3(blocks), dim3(kP2GThreads), 0, sim.stream, sim.pbuf[sim.cur], sim.nalloc,
                       sim.num_particles, sim.prev_dt, sim.cell, sim.W, sim.H, sim.u_acc, sim.v_acc, sim.density,
                       sim.levels[0].fluid, sim.obstacle);
    sim.mark(APIC_T_K_P2G, 1);
}

void launch_normalize_forces(Sim& sim, bool normalize, bool forces, float dt) {
    size_t n = std::max(sim.nu(), sim.nv());
    float g = sim.props.gravity * dt;
    int ext = forces && sim.forces_active;
    hipLaunchKernelGGL(normalize_forces_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, sim.stream,
                       sim.u_acc, sim.v_acc, sim.mac_u, sim.mac_v, sim.valid_u, sim.valid_v, sim.ext_u, sim.ext_v,
                       sim.levels[0].fluid, sim.W, sim.H, (int)normalize, ext, (int)forces, dt, g);
}

void launch_g2p(Sim& sim, float dt, bool write_keys) {
    if (sim.nalloc == 0) {
        launch_fill_u32(sim.stream, sim.maxv_bits, 0u, 1);
        return;
    }
    const unsigned blocks = particle_blocks(sim);
    sim.mark(APIC_T_K_G2P, 0);
    hipLaunchKernelGGL(g2p_kernel, dim3(blocks), dim3(256), 0, sim.stream, sim.pbuf[sim.cur], sim.nalloc,
                       sim.num_particles, dt, sim.cell, sim.W, sim.H, sim.timestamp, sim.props.border_damping,
                       sim.mac_u, sim.mac_v, sim.open_u, sim.open_v, (int)write_keys, sim.keys[0], sim.vals[0],
                       sim.part_maxv);
    sim.mark(APIC_T_K_G2P, 1);
    hipLaunchKernelGGL(maxv_finalize_kernel, dim3(1), dim3(1024), 0, sim.stream, sim.part_maxv, blocks, sim.maxv_bits);
}

void launch_max_velocity(Sim& sim) {
    if (sim.nalloc == 0) {
        launch_fill_u32(sim.stream, sim.maxv_bits, 0u, 1);
        return;
    }
    const unsigned blocks = particle_blocks(sim);
    hipLaunchKernelGGL(max_velocity_kernel, dim3(blocks), dim3(256), 0, sim.stream, sim.pfield(V_X),
                       sim.pfield(V_Y), sim.nalloc, sim.part_maxv);
    hipLaunchKernelGGL(maxv_finalize_kernel, dim3(1), dim3(1024), 0, sim.stream, sim.part_maxv, blocks, sim.maxv_bits);
}

}  // namespace apic
