// mirror.hip -- the reference's pub particle fields in its own layout.
//
// Apic keeps pos, vel, c_u and c_v as Vec<Vec2x8> and time_residual as
// Vec<f32x8> (src/apic.rs:757-767): chunk k holds particles 8k..8k+7, a
// Vec2x8 as eight x followed by eight y (64 bytes), an f32x8 as eight
// floats.  The library keeps structure-of-arrays fields; these kernels pack
// a group into that chunk layout (or unpack it) through a staging buffer,
// so that a host mirror is one copy.
#include "apic_internal.hpp"

namespace apic {

namespace {

// Vec2 groups: staging[16k + l] = x[8k + l], staging[16k + 8 + l] = y[8k + l].
__global__ void pack_vec2x8_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                   float* __restrict__ out, size_t nalloc) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nalloc) return;
    const size_t k = i >> 3, l = i & 7;
    out[16 * k + l] = x[i];
    out[16 * k + 8 + l] = y[i];
}

__global__ void unpack_vec2x8_kernel(const float* __restrict__ in, float* __restrict__ x, float* __restrict__ y,
                                     size_t nalloc) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nalloc) return;
    const size_t k = i >> 3, l = i & 7;
    x[i] = in[16 * k + l];
    y[i] = in[16 * k + 8 + l];
}

}  // namespace

size_t aosoa_length(const Sim& sim, int group) {
    return group == APIC_PART_TIME_RESIDUAL ? sim.nalloc : 2 * sim.nalloc;
}

// The two SoA fields of a Vec2 group (group 0..3), or the time residual.
static void group_fields(int group, int& fx, int& fy) {
    static const int first[4] = {P_X, V_X, CU_X, CV_X};
    if (group == APIC_PART_TIME_RESIDUAL) { fx = T_RES; fy = -1; return; }
    fx = first[group];
    fy = fx + 1;
}

void launch_pack_aosoa(Sim& sim, int group, float* staging) {
    if (sim.nalloc == 0) return;
    int fx, fy;
    group_fields(group, fx, fy);
    if (fy < 0) {  // an f32x8 vector is the SoA field itself
        (void)hipMemcpyAsync(staging, sim.pfield(fx), sim.nalloc * 4, hipMemcpyDeviceToDevice, sim.stream);
        return;
    }
    hipLaunchKernelGGL(pack_vec2x8_kernel, dim3((unsigned)((sim.nalloc + 255) / 256)), dim3(256), 0, sim.stream,
                       sim.pfield(fx), sim.pfield(fy), staging, sim.nalloc);
}

void launch_unpack_aosoa(Sim& sim, int group, const float* staging) {
    if (sim.nalloc == 0) return;
    int fx, fy;
    group_fields(group, fx, fy);
    if (fy < 0) {
        (void)hipMemcpyAsync(sim.pfield(fx), staging, sim.nalloc * 4, hipMemcpyDeviceToDevice, sim.stream);
        return;
    }
    hipLaunchKernelGGL(unpack_vec2x8_kernel, dim3((unsigned)((sim.nalloc + 255) / 256)), dim3(256), 0, sim.stream,
                       staging, sim.pfield(fx), sim.pfield(fy), sim.nalloc);
}

}  // namespace apic
