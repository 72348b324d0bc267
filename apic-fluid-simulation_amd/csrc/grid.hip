// grid.hip -- face-centred stencils of the APIC step.
//
//  gradient_kernel     apply_pressure_gradient (src/apic.rs:1958-2055), fused
//                      with the post-force add of Apic::update (:2453-2458)
//  extrapolate_kernel  one Jacobi pass of extrapolate_velocity (:2266-2350)
//                      over u and v faces at once; four ping-pong launches
//  radial_force_kernel add_radial_force (:2352-2410)
//  open_faces_kernel   update_world's open-face masks (:918-944)
#include "apic_internal.hpp"
#include "device_math.hpp"

namespace apic {

namespace {

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void fill_f32_kernel(float* __restrict__ p, float v, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__device__ inline bool obst(const uint8_t* o, int32_t i, int32_t j, uint32_t W, uint32_t H) {
    if (i < 0 || i >= (int32_t)W || j < 0 || j >= (int32_t)H) return true;
    return o[(size_t)j * W + i] != 0;
}
__device__ inline bool fluid_at(const uint8_t* f, int32_t i, int32_t j, uint32_t W, uint32_t H) {
    if (i < 0 || i >= (int32_t)W || j < 0 || j >= (int32_t)H) return false;
    return f[(size_t)j * W + i] != 0;
}

__global__ void gradient_kernel(float* __restrict__ mac_u, float* __restrict__ mac_v, uint8_t* __restrict__ valid_u,
                                uint8_t* __restrict__ valid_v, const float* __restrict__ pressure,
                                const float* __restrict__ smoothed, const uint8_t* __restrict__ obstacle,
                                const uint8_t* __restrict__ fluid, uint32_t W, uint32_t H, float inv_cell,
                                int post, const float* __restrict__ post_u, const float* __restrict__ post_v,
                                float dt) {
    const size_t nu = (size_t)W * H + H, nv = (size_t)W * H + W;
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nu) {
        int32_t y = (int32_t)(f / (W + 1)), x = (int32_t)(f % (W + 1));
        float u;
        bool lo = obst(obstacle, x - 1, y, W, H), ro = obst(obstacle, x, y, W, H);
        if (lo != ro) {
            u = 0.0f;
            valid_u[f] = 1;
        } else {
            bool lf = fluid_at(fluid, x - 1, y, W, H), rf = fluid_at(fluid, x, y, W, H);
            float pd = 0.0f;
            size_t li = (size_t)y * W + (x - 1), ri = (size_t)y * W + x;
            if (lf && rf) pd = pressure[ri] - pressure[li];
            else if (lf) pd = (-pressure[li]) / theta(smoothed[ri], smoothed[li]);
            else if (rf) pd = pressure[ri] / theta(smoothed[li], smoothed[ri]);
            u = mac_u[f] - pd * inv_cell;
        }
        if (post) u += post_u[f] * dt;
        mac_u[f] = u;
    }
    if (f < nv) {
        int32_t y = (int32_t)(f / W), x = (int32_t)(f % W);
        float v;
        bool bo = obst(obstacle, x, y - 1, W, H), to = obst(obstacle, x, y, W, H);
        if (bo != to) {
            v = 0.0f;
            valid_v[f] = 1;
        } else {
            bool bf = fluid_at(fluid, x, y - 1, W, H), tf = fluid_at(fluid, x, y, W, H);
            float pd = 0.0f;
            size_t bi = (size_t)(y - 1) * W + x, ti = (size_t)y * W + x;
            if (bf && tf) pd = pressure[ti] - pressure[bi];
            else if (bf) pd = (-pressure[bi]) / theta(smoothed[ti], smoothed[bi]);
            else if (tf) pd = pressure[ti] / theta(smoothed[bi], smoothed[ti]);
            v = mac_v[f] - pd * inv_cell;
        }
        if (post) v += post_v[f] * dt;
        mac_v[f] = v;
    }
}

__global__ void post_force_kernel(float* __restrict__ mac_u, float* __restrict__ mac_v,
                                  const float* __restrict__ post_u, const float* __restrict__ post_v, size_t nu,
                                  size_t nv, float dt) {
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nu) mac_u[f] += post_u[f] * dt;
    if (f < nv) mac_v[f] += post_v[f] * dt;
}

__global__ void extrapolate_kernel(const float* __restrict__ in_u, const uint8_t* __restrict__ in_vu,
                                   float* __restrict__ out_u, uint8_t* __restrict__ out_vu,
                                   const float* __restrict__ in_v, const uint8_t* __restrict__ in_vv,
                                   float* __restrict__ out_v, uint8_t* __restrict__ out_vv, uint32_t W, uint32_t H) {
    const size_t nu = (size_t)W * H + H, nv = (size_t)W * H + W;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nu) {
        const size_t W1 = W + 1;
        size_t x = i % W1, y = i / W1;
        if (in_vu[i]) {
            out_u[i] = in_u[i];
            out_vu[i] = 1;
        } else {
            float sum = 0.0f;
            uint32_t cnt = 0;
            if (x > 0 && in_vu[i - 1]) { sum += in_u[i - 1]; ++cnt; }
            if (x < W && in_vu[i + 1]) { sum += in_u[i + 1]; ++cnt; }
            if (y > 0 && in_vu[i - W1]) { sum += in_u[i - W1]; ++cnt; }
            if (y < H - 1 && in_vu[i + W1]) { sum += in_u[i + W1]; ++cnt; }
            if (cnt > 0) { out_u[i] = sum / (float)cnt; out_vu[i] = 1; }
            else { out_u[i] = in_u[i]; out_vu[i] = 0; }
        }
    }
    if (i < nv) {
        size_t x = i % W, y = i / W;
        if (in_vv[i]) {
            out_v[i] = in_v[i];
            out_vv[i] = 1;
        } else {
            float sum = 0.0f;
            uint32_t cnt = 0;
            if (x > 0 && in_vv[i - 1]) { sum += in_v[i - 1]; ++cnt; }
            if (x < W - 1 && in_vv[i + 1]) { sum += in_v[i + 1]; ++cnt; }
            if (y > 0 && in_vv[i - W]) { sum += in_v[i - W]; ++cnt; }
            if (y < H && in_vv[i + W]) { sum += in_v[i + W]; ++cnt; }
            if (cnt > 0) { out_v[i] = sum / (float)cnt; out_vv[i] = 1; }
            else { out_v[i] = in_v[i]; out_vv[i] = 0; }
        }
    }
}

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

void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n) {
    if (n == 0) return;
    unsigned b = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(b), dim3(256), 0, s, p, v, n);
}

void launch_fill_f32(hipStream_t s, float* p, float v, size_t n) {
    if (n == 0) return;
    unsigned b = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_f32_kernel, dim3(b), dim3(256), 0, s, p, v, n);
}

void launch_gradient(Sim& sim, bool post_force, float dt) {
    size_t n = std::max(sim.nu(), sim.nv());
    float inv_cell = 1.0f / sim.cell;
    int post = post_force && sim.forces_active;
    hipLaunchKernelGGL(gradient_kernel, dim3(blocks_for(n)), dim3(256), 0, sim.stream, sim.mac_u, sim.mac_v,
                       sim.valid_u, sim.valid_v, sim.pressure, sim.smooth_a, sim.obstacle, sim.levels[0].fluid,
                       sim.W, sim.H, inv_cell, post, sim.post_u, sim.post_v, dt);
}

void launch_post_force(Sim& sim, float dt) {
    size_t n = std::max(sim.nu(), sim.nv());
    hipLaunchKernelGGL(post_force_kernel, dim3(blocks_for(n)), dim3(256), 0, sim.stream, sim.mac_u, sim.mac_v,
                       sim.post_u, sim.post_v, sim.nu(), sim.nv(), dt);
}

void launch_extrapolate(Sim& sim) {
    size_t n = std::max(sim.nu(), sim.nv());
    for (int pass = 0; pass < 4; ++pass) {
        bool fwd = (pass % 2) == 0;
        float* iu = fwd ? sim.mac_u : sim.old_u;
        float* ou = fwd ? sim.old_u : sim.mac_u;
        uint8_t* ivu = fwd ? sim.valid_u : sim.nvalid_u;
        uint8_t* ovu = fwd ? sim.nvalid_u : sim.valid_u;
        float* iv = fwd ? sim.mac_v : sim.old_v;
        float* ov = fwd ? sim.old_v : sim.mac_v;
        uint8_t* ivv = fwd ? sim.valid_v : sim.nvalid_v;
        uint8_t* ovv = fwd ? sim.nvalid_v : sim.valid_v;
        hipLaunchKernelGGL(extrapolate_kernel, dim3(blocks_for(n)), dim3(256), 0, sim.stream, iu, ivu, ou, ovu, iv,
                           ivv, ov, ovv, sim.W, sim.H);
    }
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
