// sort.hip -- Morton-key particle sort (update_spatial_lookup,
// src/apic.rs:1043-1111) as keygen + stable LSD radix sort + SoA gather.
//
// The reference sorts (key, index) records with rdst's unstable radix sort;
// this one is stable, so equal keys keep ascending particle index and the
// permutation equals std::stable_sort's.  Each pass ranks one 8-bit digit:
//   upsweep   per-tile digit histogram in LDS  -> counts[digit][tile]
//   scan      exclusive prefix over the digit-major counts (3 kernels)
//   downsweep per-tile stable ranks: 64-lane ballots give each lane the mask
//             of lower lanes holding the same digit, LDS carries per-wave and
//             running per-digit counts, and the record is scattered to
//             offset[digit][tile] + rank.
// Only the digits that can be nonzero are sorted: keys are Morton codes of
// cells clamped into the grid, so key <= morton(W-1, H-1).
#include "apic_internal.hpp"
#include "device_math.hpp"

namespace apic {

namespace {

constexpr int kRadixThreads = 256;   // 4 waves of 64
constexpr int kRadixItems = 16;      // keys per thread per tile
constexpr int kRadixTile = kRadixThreads * kRadixItems;
constexpr int kRadixBits = 8;
constexpr int kRadixBuckets = 1 << kRadixBits;
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void keygen_kernel(const float* __restrict__ px, const float* __restrict__ py, uint32_t n,
                              float cell, float max_x, float max_y, uint32_t* __restrict__ keys,
                              uint32_t* __restrict__ vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gx = px[i] / cell, gy = py[i] / cell;
    uint32_t ix = (uint32_t)(int32_t)fast_min(fast_max(gx, 0.0f), max_x);
    uint32_t iy = (uint32_t)(int32_t)fast_min(fast_max(gy, 0.0f), max_y);
    keys[i] = morton(ix, iy);
    vals[i] = i;
}

__global__ __launch_bounds__(kRadixThreads) void radix_upsweep(const uint32_t* __restrict__ keys, uint32_t n,
                                                               int shift, uint32_t* __restrict__ counts,
                                                               uint32_t ntiles) {
    __shared__ uint32_t hist[kRadixBuckets];
    const int tid = threadIdx.x;
    hist[tid] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * kRadixTile;
#pragma unroll
    for (int j = 0; j < kRadixItems; ++j) {
        size_t idx = base + (size_t)j * kRadixThreads + tid;
        if (idx < n) atomicAdd(&hist[(keys[idx] >> shift) & (kRadixBuckets - 1)], 1u);
    }
    __syncthreads();
    counts[(size_t)tid * ntiles + blockIdx.x] = hist[tid];
}

// Block-wide exclusive scan of one value per thread (1024 threads = 16 waves).
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* wave_sums, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wave_sums[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        const int nw = blockDim.x >> 6;
        for (int w = 0; w < nw; ++w) { uint32_t t = wave_sums[w]; wave_sums[w] = run; run += t; }
        wave_sums[nw] = run;
    }
    __syncthreads();
    total = wave_sums[blockDim.x >> 6];
    uint32_t r = wave_sums[wave] + incl - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanThreads) void scan_tiles(uint32_t* __restrict__ data, size_t n,
                                                           uint32_t* __restrict__ tile_sums) {
    __shared__ uint32_t wave_sums[kScanThreads / 64 + 1];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? data[base + k] : 0u;
        local += v[k];
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan(local, wave_sums, total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) data[base + k] = off;
        off += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Exclusive scan of the tile sums by one block (looping over 1024-wide chunks).
__global__ __launch_bounds__(kScanThreads) void scan_tile_sums(uint32_t* __restrict__ tile_sums, uint32_t n) {
    __shared__ uint32_t wave_sums[kScanThreads / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < n; b += kScanThreads) {
        uint32_t i = b + threadIdx.x;
        uint32_t v = i < n ? tile_sums[i] : 0u;
        uint32_t total;
        uint32_t off = block_exclusive_scan(v, wave_sums, total);
        if (i < n) tile_sums[i] = carry + off;
        carry += total;
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_add(uint32_t* __restrict__ data, size_t n,
                                                         const uint32_t* __restrict__ tile_sums) {
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t add = tile_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) data[base + k] += add;
}

__global__ __launch_bounds__(kRadixThreads) void radix_downsweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n, int shift, const uint32_t* __restrict__ offsets, uint32_t ntiles) {
    constexpr int kWaves = kRadixThreads / 64;
    __shared__ uint32_t goff[kRadixBuckets];
    __shared__ uint32_t run[kRadixBuckets];
    __shared__ uint32_t whist[kWaves][kRadixBuckets];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    goff[tid] = offsets[(size_t)tid * ntiles + blockIdx.x];
    run[tid] = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) whist[w][tid] = 0;
    __syncthreads();
    const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    size_t base = (size_t)blockIdx.x * kRadixTile;
    for (int j = 0; j < kRadixItems; ++j) {
        size_t idx = base + (size_t)j * kRadixThreads + tid;
        bool valid = idx < n;
        uint32_t k = valid ? keys_in[idx] : 0u;
        uint32_t v = valid ? vals_in[idx] : 0u;
        uint32_t d = (k >> shift) & (kRadixBuckets - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lanes_below);
        if (valid && rank == 0) whist[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int w = 0; w < wave; ++w) pos += whist[w][d];
            uint32_t dst = goff[d] + pos;
            keys_out[dst] = k;
            vals_out[dst] = v;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) { add += whist[w][tid]; whist[w][tid] = 0; }
        run[tid] += add;
        __syncthreads();
    }
}

__global__ void permute_kernel(const float* __restrict__ in, float* __restrict__ out,
                               const uint32_t* __restrict__ vals, uint32_t np, size_t nalloc) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nalloc) return;
    if (i < np) {
        size_t j = vals[i];
        if (j >= np) j = np - 1;  // records written through the field API are not trusted
#pragma unroll
        for (int f = 0; f < NUM_PFIELDS; ++f) out[(size_t)f * nalloc + i] = in[(size_t)f * nalloc + j];
    } else {
        // the reference's gather leaves tail lanes of the last chunk zero (src/apic.rs:1077-1103)
#pragma unroll
        for (int f = 0; f < NUM_PFIELDS; ++f) out[(size_t)f * nalloc + i] = 0.0f;
    }
}

uint32_t key_bits(const Sim& sim) {
    uint32_t mk = morton(sim.W > 0 ? sim.W - 1 : 0, sim.H > 0 ? sim.H - 1 : 0);
    uint32_t bits = 0;
    while (bits < 32 && (mk >> bits) != 0) ++bits;
    return bits == 0 ? 1 : bits;
}

}  // namespace

size_t radix_scratch_words(size_t n) {
    size_t ntiles = (n + kRadixTile - 1) / kRadixTile;
    size_t counts = (size_t)kRadixBuckets * ntiles;
    size_t scan_tiles_n = (counts + kScanTile - 1) / kScanTile;
    return counts + scan_tiles_n + 64;
}

void launch_keygen(Sim& sim) {
    if (sim.num_particles == 0) return;
    uint32_t n = sim.num_particles;
    float max_x = (float)(sim.W > 0 ? sim.W - 1 : 0), max_y = (float)(sim.H > 0 ? sim.H - 1 : 0);
    hipLaunchKernelGGL(keygen_kernel, dim3((n + 255) / 256), dim3(256), 0, sim.stream, sim.pfield(P_X),
                       sim.pfield(P_Y), n, sim.cell, max_x, max_y, sim.keys[0], sim.vals[0]);
}

void launch_radix_sort(Sim& sim) {
    uint32_t n = sim.num_particles;
    if (n <= 1) return;
    uint32_t ntiles = (n + kRadixTile - 1) / kRadixTile;
    size_t ncounts = (size_t)kRadixBuckets * ntiles;
    uint32_t nscan = (uint32_t)((ncounts + kScanTile - 1) / kScanTile);
    uint32_t* counts = sim.radix_scratch;
    uint32_t* tile_sums = counts + ncounts;
    int passes = (int)((key_bits(sim) + kRadixBits - 1) / kRadixBits);
    int src = 0;
    for (int p = 0; p < passes; ++p) {
        int shift = p * kRadixBits;
        hipLaunchKernelGGL(radix_upsweep, dim3(ntiles), dim3(kRadixThreads), 0, sim.stream, sim.keys[src], n, shift,
                           counts, ntiles);
        hipLaunchKernelGGL(scan_tiles, dim3(nscan), dim3(kScanThreads), 0, sim.stream, counts, ncounts, tile_sums);
        hipLaunchKernelGGL(scan_tile_sums, dim3(1), dim3(kScanThreads), 0, sim.stream, tile_sums, nscan);
        hipLaunchKernelGGL(scan_add, dim3(nscan), dim3(kScanThreads), 0, sim.stream, counts, ncounts, tile_sums);
        hipLaunchKernelGGL(radix_downsweep, dim3(ntiles), dim3(kRadixThreads), 0, sim.stream, sim.keys[src],
                           sim.vals[src], sim.keys[1 - src], sim.vals[1 - src], n, shift, counts, ntiles);
        src = 1 - src;
    }
    if (src != 0) {
        std::swap(sim.keys[0], sim.keys[1]);
        std::swap(sim.vals[0], sim.vals[1]);
    }
}

void launch_permute(Sim& sim) {
    if (sim.nalloc == 0) return;
    size_t n = sim.nalloc;
    hipLaunchKernelGGL(permute_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, sim.stream, sim.pbuf[sim.cur],
                       sim.pbuf[1 - sim.cur], sim.vals[0], sim.num_particles, n);
    sim.cur = 1 - sim.cur;
}

}  // namespace apic
