// grid.hip -- face-centred stencils of the APIC step.
//
//  gradient_kernel     apply_pressure_gradient (src/apic.rs:1958-2055), fused
//                      with the post-force add of Apic::update (:2453-2458)
//  extrapolate_kernel  extrapolate_velocity (:2266-2350): the four Jacobi
//                      passes of both components in one launch, on LDS tiles
//                      with a four-face halo (each pass needs one more ring)
//  radial_force_kernel add_radial_force (:2352-2410)
//  open_faces_kernel   update_world's open-face masks (:918-944)
#include "apic_internal.hpp"
#include "device_math.hpp"

namespace apic {

namespace {

template <typename T>
struct Fill {
    static constexpr int kThreads = 256;
    struct Args {
        T* p;
        T v;
        size_t n;
        uint32_t nwg;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        const size_t stride = (size_t)a.nwg * kThreads;
        for (size_t i = (size_t)wg * kThreads + threadIdx.x; i < a.n; i += stride) a.p[i] = a.v;
    }
};

__device__ inline bool obst(const uint8_t* o, int32_t i, int32_t j, uint32_t W, uint32_t H) {
    if (i < 0 || i >= (int32_t)W || j < 0 || j >= (int32_t)H) return true;
    return o[(size_t)j * W + i] != 0;
}
__device__ inline bool fluid_at(const uint8_t* f, int32_t i, int32_t j, uint32_t W, uint32_t H) {
    if (i < 0 || i >= (int32_t)W || j < 0 || j >= (int32_t)H) return false;
    return f[(size_t)j * W + i] != 0;
}

// The substep's dt (the time-step rule of Apic::update, src/apic.rs:2433-2441):
// min(cfl * dx / max|v|, frame time left), with cfl * dx / max|v| replaced by
// the frame dt when max|v| <= 1e-5.  The host keeps the same sum of dts in f32
// to know whether another substep follows.
struct SubstepBegin {
    static constexpr int kThreads = 64;
    struct Args {
        StepState* ss;
        StepState* host;
        const uint32_t* maxv_bits;
        float cfl_cell, frame_dt;
        int first;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) {
        if (threadIdx.x != 0) return;
        StepState s = *a.ss;
        if (a.first) s.elapsed = 0.0f;
        const float maxv = __uint_as_float(*a.maxv_bits);
        const float safe = maxv > 1e-5f ? a.cfl_cell / maxv : a.frame_dt;
        const float rem = a.frame_dt - s.elapsed;
        s.dt = safe < rem ? safe : rem;
        s.maxv = maxv;
        *a.ss = s;
        if (a.host) *a.host = s;
    }
};

struct SubstepEnd {
    static constexpr int kThreads = 64;
    struct Args {
        StepState* ss;
        const int32_t* guard;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) {
        if (threadIdx.x != 0) return;
        if (a.guard && *a.guard == 0) return;
        substep_end(a.ss);
    }
};

struct Gradient {
    static constexpr int kThreads = 256;
    struct Args {
        float* mac_u;
        float* mac_v;
        uint8_t* valid_u;
        uint8_t* valid_v;
        const float* pressure;
        const float* smoothed;
        const uint8_t* obstacle;
        const uint8_t* fluid;
        uint32_t W, H;
        float inv_cell;
        int post;
        const float* post_u;
        const float* post_v;
        float dt_arg;
        const StepState* ss;
        const int32_t* guard;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        if (a.guard && *a.guard == 0) return;
        const uint32_t W = a.W, H = a.H;
        const uint8_t* __restrict__ obstacle = a.obstacle;
        const uint8_t* __restrict__ fluid = a.fluid;
        const float* __restrict__ pressure = a.pressure;
        const float* __restrict__ smoothed = a.smoothed;
        const float dt = a.ss ? a.ss->dt : a.dt_arg;
        // (face indices fit 32 bits: 32-bit quotients, far cheaper than 64-bit ones)
        const uint32_t nu = W * H + H, nv = W * H + W;
        const uint32_t f = wg * kThreads + threadIdx.x;
        // Every operand of a face is loaded before the first test (cells past
        // the grid read cell 0 and are not used): one wait instead of a chain
        // of obstacle, fluid and pressure loads.
        if (f < nu) {
            const int32_t y = (int32_t)(f / (W + 1)), x = (int32_t)(f % (W + 1));
            const bool in_l = x >= 1, in_r = x < (int32_t)W;
            const size_t li = in_l ? (size_t)y * W + (x - 1) : 0, ri = in_r ? (size_t)y * W + x : 0;
            const uint8_t ol = obstacle[li], orr = obstacle[ri], fl = fluid[li], fr = fluid[ri];
            const float pl = pressure[li], pr = pressure[ri], sl = smoothed[li], sr = smoothed[ri];
            const float mu = a.mac_u[f];
            const float pu = a.post ? a.post_u[f] : 0.0f;
            float u;
            const bool lo = !in_l || ol != 0, ro = !in_r || orr != 0;
            if (lo != ro) {
                u = 0.0f;
                a.valid_u[f] = 1;
            } else {
                const bool lf = in_l && fl != 0, rf = in_r && fr != 0;
                float pd = 0.0f;
                if (lf && rf) pd = pr - pl;
                else if (lf) pd = (-pl) / theta(sr, sl);
                else if (rf) pd = pr / theta(sl, sr);
                u = mu - pd * a.inv_cell;
            }
            if (a.post) u += pu * dt;
            a.mac_u[f] = u;
        }
        if (f < nv) {
            const int32_t y = (int32_t)(f / W), x = (int32_t)(f % W);
            const bool in_b = y >= 1, in_t = y < (int32_t)H;
            const size_t bi = in_b ? (size_t)(y - 1) * W + x : 0, ti = in_t ? (size_t)y * W + x : 0;
            const uint8_t ob = obstacle[bi], ot = obstacle[ti], fb = fluid[bi], ft = fluid[ti];
            const float pb = pressure[bi], pt = pressure[ti], sb = smoothed[bi], st = smoothed[ti];
            const float mv = a.mac_v[f];
            const float pv = a.post ? a.post_v[f] : 0.0f;
            float v;
            const bool bo = !in_b || ob != 0, to = !in_t || ot != 0;
            if (bo != to) {
                v = 0.0f;
                a.valid_v[f] = 1;
            } else {
                const bool bf = in_b && fb != 0, tf = in_t && ft != 0;
                float pd = 0.0f;
                if (bf && tf) pd = pt - pb;
                else if (bf) pd = (-pb) / theta(st, sb);
                else if (tf) pd = pt / theta(sb, st);
                v = mv - pd * a.inv_cell;
            }
            if (a.post) v += pv * dt;
            a.mac_v[f] = v;
        }
    }
};

__global__ void post_force_kernel(float* __restrict__ mac_u, float* __restrict__ mac_v,
                                  const float* __restrict__ post_u, const float* __restrict__ post_v, size_t nu,
                                  size_t nv, float dt) {
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nu) mac_u[f] += post_u[f] * dt;
    if (f < nv) mac_v[f] += post_v[f] * dt;
}

constexpr int kExTX = 64, kExTY = 16, kExHalo = 4;  // four passes need a four-face halo
constexpr int kExLW = kExTX + 2 * kExHalo, kExLH = kExTY + 2 * kExHalo;

// One tile of one component (nx x ny faces: u is (W+1) x H, v is W x (H+1)).
// Pass p recomputes the faces at least p + 1 faces inside the loaded region,
// so after four passes the tile itself holds exactly what four whole-grid
// passes give.  Per face: a valid face keeps its value; an invalid one takes
// the mean of its valid neighbours, summed left, right, below, above
// (src/apic.rs:2276-2310), and becomes valid.
struct Extrapolate {
    static constexpr int kThreads = 256;
    struct Args {
        const float* in_u;
        const uint8_t* in_vu;
        float* out_u;
        uint8_t* out_vu;
        const float* in_v;
        const uint8_t* in_vv;
        float* out_v;
        uint8_t* out_vv;
        uint32_t W, H, tiles_u_x, tiles_u, tiles_v_x;
        const int32_t* guard;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        __shared__ float val[2][kExLH * kExLW];
        __shared__ uint8_t ok[2][kExLH * kExLW];
        if (a.guard && *a.guard == 0) return;
        const uint32_t W = a.W, H = a.H;
        const bool is_u = wg < a.tiles_u;
        const uint32_t b = is_u ? wg : wg - a.tiles_u;
        const uint32_t tx = is_u ? a.tiles_u_x : a.tiles_v_x;
        const int32_t nx = (int32_t)(is_u ? W + 1 : W), ny = (int32_t)(is_u ? H : H + 1);
        const float* __restrict__ in = is_u ? a.in_u : a.in_v;
        const uint8_t* __restrict__ in_ok = is_u ? a.in_vu : a.in_vv;
        float* __restrict__ out = is_u ? a.out_u : a.out_v;
        uint8_t* __restrict__ out_ok = is_u ? a.out_vu : a.out_vv;
        const int32_t X0 = (int32_t)(b % tx) * kExTX - kExHalo, Y0 = (int32_t)(b / tx) * kExTY - kExHalo;
        for (int i = threadIdx.x; i < kExLH * kExLW; i += blockDim.x) {
            const int32_t gx = X0 + i % kExLW, gy = Y0 + i / kExLW;
            const bool inside = gx >= 0 && gx < nx && gy >= 0 && gy < ny;
            const size_t g = (size_t)gy * nx + gx;
            val[0][i] = inside ? in[g] : 0.0f;
            ok[0][i] = inside ? in_ok[g] : 0;
        }
        __syncthreads();
#pragma unroll 1
        for (int p = 0; p < 4; ++p) {
            const int src = p & 1, dst = src ^ 1;
            for (int i = threadIdx.x; i < kExLH * kExLW; i += blockDim.x) {
                const int lx = i % kExLW, ly = i / kExLW;
                if (lx <= p || lx >= kExLW - 1 - p || ly <= p || ly >= kExLH - 1 - p) continue;
                const int32_t gx = X0 + lx, gy = Y0 + ly;
                if (gx < 0 || gx >= nx || gy < 0 || gy >= ny) continue;
                // the face's and its neighbours' flags and values load at
                // once; an invalid neighbour adds +0 (the sum starts at +0
                // and never becomes -0, so it is the sum of the valid ones)
                const uint8_t o = ok[src][i];
                const uint8_t ol = ok[src][i - 1], orr = ok[src][i + 1];
                const uint8_t ob = ok[src][i - kExLW], ot = ok[src][i + kExLW];
                const float v = val[src][i];
                const float vl = val[src][i - 1], vr = val[src][i + 1];
                const float vb = val[src][i - kExLW], vt = val[src][i + kExLW];
                const bool hl = gx > 0 && ol, hr = gx < nx - 1 && orr;
                const bool hb = gy > 0 && ob, ht = gy < ny - 1 && ot;
                float sum = 0.0f;
                sum += hl ? vl : 0.0f;
                sum += hr ? vr : 0.0f;
                sum += hb ? vb : 0.0f;
                sum += ht ? vt : 0.0f;
                const uint32_t cnt = (uint32_t)hl + (uint32_t)hr + (uint32_t)hb + (uint32_t)ht;
                if (o) {
                    val[dst][i] = v;
                    ok[dst][i] = 1;
                } else if (cnt > 0) {
                    val[dst][i] = sum / (float)cnt;
                    ok[dst][i] = 1;
                } else {
                    val[dst][i] = v;
                    ok[dst][i] = 0;
                }
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < kExTY * kExTX; i += blockDim.x) {
            const int lx = kExHalo + i % kExTX, ly = kExHalo + i / kExTX;
            const int32_t gx = X0 + lx, gy = Y0 + ly;
            if (gx >= nx || gy >= ny) continue;
            const size_t g = (size_t)gy * nx + gx;
            out[g] = val[0][ly * kExLW + lx];
            out_ok[g] = ok[0][ly * kExLW + lx];
        }
    }
};

__global__ void radial_force_kernel(float* __restrict__ ext_u, float* __restrict__ ext_v,
                                    float* __restrict__ post_u, float* __restrict__ post_v, uint32_t W,
                                    uint32_t H, float cell, uint32_t x0, uint32_t y0, uint32_t nx, uint32_t ny,
                                    float wx, float wy, float radius, float strength) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    uint32_t x = x0 + t % nx, y = y0 + t / nx;
    const float r2 = radius * radius;
    const bool repel = strength > 0.0f;
    if (x <= W && y < H) {
        float dx = (float)x * cell - wx, dy = ((float)y + 0.5f) * cell - wy;
        float d2 = dx * dx + dy * dy;
        if (d2 < r2 && d2 > 1e-5f) {
            float d = sqrtf(d2);
            float fall = 1.0f - (d / radius);
            float f = ((dx / d) * strength) * fall;
            size_t id = u_index((int32_t)x, (int32_t)y, W, H);
            if (repel) post_u[id] += f; else ext_u[id] += f;
        }
    }
    if (x < W && y <= H) {
        float dx = ((float)x + 0.5f) * cell - wx, dy = (float)y * cell - wy;
        float d2 = dx * dx + dy * dy;
        if (d2 < r2 && d2 > 1e-5f) {
            float d = sqrtf(d2);
            float fall = 1.0f - (d / radius);
            float f = ((dy / d) * strength) * fall;
            size_t id = v_index((int32_t)x, (int32_t)y, W, H);
            if (repel) post_v[id] += f; else ext_v[id] += f;
        }
    }
}

__global__ void open_faces_kernel(const uint8_t* __restrict__ obstacle, uint8_t* __restrict__ open_u,
                                  uint8_t* __restrict__ open_v, uint32_t W, uint32_t H) {
    const size_t nu = (size_t)W * H + H, nv = (size_t)W * H + W;
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nu) {
        int32_t i = (int32_t)(f % (W + 1)), j = (int32_t)(f / (W + 1));
        open_u[f] = !(obst(obstacle, i - 1, j, W, H) || obst(obstacle, i, j, W, H));
    }
    if (f < nv) {
        int32_t i = (int32_t)(f % W), j = (int32_t)(f / W);
        open_v[f] = !(obst(obstacle, i, j - 1, W, H) || obst(obstacle, i, j, W, H));
    }
}

unsigned blocks_for(size_t n, unsigned threads = 256) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace

void launch_substep_begin(Sim& sim, bool first, float frame_dt) {
    emit<SubstepBegin>(sim, {sim.step, sim.h_step_d, sim.maxv_bits, sim.props.cfl * sim.cell, frame_dt, (int)first}, 1);
}

void launch_substep_end(Sim& sim) { emit<SubstepEnd>(sim, {sim.step, sim.tail_guard}, 1); }

void launch_fill_u32(Sim& sim, uint32_t* p, uint32_t v, size_t n) {
    if (n == 0) return;
    const uint32_t b = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
    emit<Fill<uint32_t>>(sim, {p, v, n, b}, b);
}

void launch_fill_f32(Sim& sim, float* p, float v, size_t n) {
    if (n == 0) return;
    const uint32_t b = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
    emit<Fill<float>>(sim, {p, v, n, b}, b);
}

void launch_gradient(Sim& sim, bool post_force, float dt) {
    size_t n = std::max(sim.nu(), sim.nv());
    Gradient::Args a{sim.mac_u, sim.mac_v, sim.valid_u, sim.valid_v, sim.pressure, sim.smooth_a, sim.obstacle,
                     sim.levels[0].fluid, sim.W, sim.H, 1.0f / sim.cell, (int)(post_force && sim.forces_active),
                     sim.post_u, sim.post_v, dt, sim.step_src, sim.tail_guard};
    emit<Gradient>(sim, a, blocks_for(n));
}

void launch_post_force(Sim& sim, float dt) {
    size_t n = std::max(sim.nu(), sim.nv());
    hipLaunchKernelGGL(post_force_kernel, dim3(blocks_for(n)), dim3(256), 0, sim.stream, sim.mac_u, sim.mac_v,
                       sim.post_u, sim.post_v, sim.nu(), sim.nv(), dt);
}

void launch_extrapolate(Sim& sim) {
    // into the spare buffers, which then become the current ones
    const uint32_t tux = (sim.W + 1 + kExTX - 1) / kExTX, tuy = (sim.H + kExTY - 1) / kExTY;
    const uint32_t tvx = (sim.W + kExTX - 1) / kExTX, tvy = (sim.H + 1 + kExTY - 1) / kExTY;
    Extrapolate::Args a{sim.mac_u, sim.valid_u, sim.old_u, sim.nvalid_u, sim.mac_v, sim.valid_v, sim.old_v,
                        sim.nvalid_v, sim.W, sim.H, tux, tux * tuy, tvx, sim.tail_guard};
    emit<Extrapolate>(sim, a, tux * tuy + tvx * tvy);
    std::swap(sim.mac_u, sim.old_u);
    std::swap(sim.mac_v, sim.old_v);
    std::swap(sim.valid_u, sim.nvalid_u);
    std::swap(sim.valid_v, sim.nvalid_v);
}

void launch_radial_force(Sim& sim, float wx, float wy, float radius, float strength) {
    // bounds as the reference computes them (casts to i32 then u32)
    auto to_i32 = [](float v) -> int32_t {
        if (std::isnan(v)) return 0;
        if (v >= 2147483647.0f) return INT32_MAX;
        if (v <= -2147483648.0f) return INT32_MIN;
        return (int32_t)v;
    };
    int32_t mnx = std::max(to_i32(std::floor((wx - radius) / sim.cell)), 0);
    int32_t mxx = std::min(to_i32(std::ceil((wx + radius) / sim.cell)), (int32_t)sim.W);
    int32_t mny = std::max(to_i32(std::floor((wy - radius) / sim.cell)), 0);
    int32_t mxy = std::min(to_i32(std::ceil((wy + radius) / sim.cell)), (int32_t)sim.H);
    // a negative upper bound wraps to a huge u32 in the reference; its inner
    // range checks then stop at W/H, so capping at W/H is equivalent
    uint32_t xe = std::min<uint32_t>((uint32_t)mxx, sim.W), ye = std::min<uint32_t>((uint32_t)mxy, sim.H);
    uint32_t xs = (uint32_t)mnx, ys = (uint32_t)mny;
    if (xs > xe || ys > ye) return;
    uint32_t nx = xe - xs + 1, ny = ye - ys + 1;
    hipLaunchKernelGGL(radial_force_kernel, dim3(blocks_for((size_t)nx * ny)), dim3(256), 0, sim.stream, sim.ext_u,
                       sim.ext_v, sim.post_u, sim.post_v, sim.W, sim.H, sim.cell, xs, ys, nx, ny, wx, wy, radius,
                       strength);
    sim.forces_active = true;
}

void launch_open_faces(Sim& sim) {
    size_t n = std::max(sim.nu(), sim.nv());
    hipLaunchKernelGGL(open_faces_kernel, dim3(blocks_for(n)), dim3(256), 0, sim.stream, sim.obstacle, sim.open_u,
                       sim.open_v, sim.W, sim.H);
}

}  // namespace apic
