// device_math.hpp -- per-element arithmetic of the APIC step, shared by the
// gfx950 kernels.  Every helper keeps the reference's order of operations
// (src/apic.rs); the library is compiled with -ffp-contract=off so that a*b+c
// is never fused except where the reference calls mul_add (written fmaf here).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace apic {

// wide's fast_max / fast_min are x86 maxps / minps: `a > b ? a : b`.
__host__ __device__ inline float fast_max(float a, float b) { return a > b ? a : b; }
__host__ __device__ inline float fast_min(float a, float b) { return a < b ? a : b; }

// expand_bits_simd / morton_code_simd, src/apic.rs:175-198
__host__ __device__ inline uint32_t expand_bits(uint32_t v) {
    v &= 0x0000FFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
__host__ __device__ inline uint32_t morton(uint32_t x, uint32_t y) {
    return expand_bits(x) | (expand_bits(y) << 1);
}

// random_simd lane value (lowbias32 of index ^ seed), src/apic.rs:946-975
__device__ inline float random_lane(uint32_t index, uint32_t seed) {
    uint32_t m = index ^ seed;
    m ^= m >> 16;
    m *= 0x7feb352du;
    m ^= m >> 15;
    m *= 0x846ca68bu;
    m ^= m >> 16;
    return __uint_as_float((m & 0x007FFFFFu) | 0x3F800000u) - 1.5f;
}

// temporal_kernel_w_t, src/apic.rs:31-46
__device__ inline float temporal_w(float tau) {
    float ts = tau - 0.5f;
    float tsq = ts * ts;
    float inner = fast_max(1.0f - tsq, 0.0f);
    float w = inner * inner * inner;
    w = w * 2.1875f;
    return fast_max(w, 1e-5f);
}

// calculate_theta, src/apic.rs:1628-1638 (target density 9)
__device__ inline float theta(float rho, float rho_center) {
    const float threshold = 9.0f * 0.5f;
    float diff = rho_center - rho;
    if (fabsf(diff) <= 1e-5f) return 1.0f;
    float t = (rho_center - threshold) / diff;
    if (t < 0.1f) t = 0.1f;
    if (t > 1.0f) t = 1.0f;
    return t;
}

__device__ inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v > hi ? hi : (v < lo ? lo : v); }

// u_index / v_index (src/apic.rs:977-991, 1029-1041)
__device__ inline uint32_t u_index(int32_t i, int32_t j, uint32_t W, uint32_t H) {
    uint32_t ci = (uint32_t)clampi(i, 0, (int32_t)W);
    uint32_t cj = (uint32_t)clampi(j, 0, (int32_t)H - 1);
    return cj * (W + 1) + ci;
}
__device__ inline uint32_t v_index(int32_t i, int32_t j, uint32_t W, uint32_t H) {
    uint32_t ci = (uint32_t)clampi(i, 0, (int32_t)W - 1);
    uint32_t cj = (uint32_t)clampi(j, 0, (int32_t)H);
    return cj * W + ci;
}

// Bilinear stencil of one particle on one staggered grid: get_u_grid /
// get_v_grid (src/apic.rs:1113-1197) and get_grid_distances (:1199-1216).
struct Stencil {
    int32_t ix, iy;   // base node (before clamping)
    float tx, ty;     // fractional offsets
    float w[4];       // bl, br, tl, tr
    float dx[4], dy[4];
};

__device__ inline void make_stencil(float gx, float gy, float cell, Stencil& s) {
    float bx = floorf(gx), by = floorf(gy);
    s.ix = (int32_t)bx;
    s.iy = (int32_t)by;
    s.tx = gx - bx;
    s.ty = gy - by;
    float itx = 1.0f - s.tx, ity = 1.0f - s.ty;
    s.w[0] = itx * ity; s.w[1] = s.tx * ity; s.w[2] = itx * s.ty; s.w[3] = s.tx * s.ty;
    float ntx = (-s.tx) * cell, nty = (-s.ty) * cell, ptx = itx * cell, pty = ity * cell;
    s.dx[0] = ntx; s.dy[0] = nty;
    s.dx[1] = ptx; s.dy[1] = nty;
    s.dx[2] = ntx; s.dy[2] = pty;
    s.dx[3] = ptx; s.dy[3] = pty;
}

}  // namespace apic
