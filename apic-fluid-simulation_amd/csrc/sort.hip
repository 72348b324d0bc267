// sort.hip -- Morton-key particle sort (update_spatial_lookup,
// src/apic.rs:1043-1111) as keygen + stable LSD radix sort + SoA gather.
//
// The reference sorts (key, index) records with rdst's unstable radix sort;
// this one is stable, so equal keys keep ascending particle index and the
// permutation equals std::stable_sort's.  Keys are Morton codes of cells
// clamped into the grid (key <= morton(W-1, H-1)), so only that many bits are
// sorted, in as few passes of at most 11-bit digits as cover them (9 + 11
// bits for a 1024^2 grid).  Each pass:
//   upsweep   per-tile digit histogram in LDS, written as one contiguous row
//             counts[tile][digit]
//   scan      per digit, the exclusive prefix over tiles in chunks of 32
//             tiles (coalesced along the rows), then one workgroup scans the
//             chunk totals and the digit totals
//   downsweep each wave ranks its own 1024 consecutive records (ballot peer
//             groups and a wave-private LDS count per digit, no workgroup
//             barrier per record), the workgroup adds the counts of lower
//             waves, stages the tile in LDS in digit order and writes it out
//             as runs: record -> start(digit, tile) + rank.
#include "apic_internal.hpp"
#include "device_math.hpp"

namespace apic {

namespace {

// A tile of 4096 keys per workgroup of 8 waves (a 2M-particle sort is about
// 500 tiles, two a CU: more waves per tile hide more of the scatter's latency
// than more, smaller tiles, whose count and scan arrays grow)
constexpr int kRadixThreads = 512;   // 8 waves of 64
constexpr int kRadixWaves = kRadixThreads / 64;
constexpr int kRadixItems = 8;       // keys per thread per tile
constexpr int kRadixTile = kRadixThreads * kRadixItems;
constexpr int kRadixWaveItems = kRadixTile / kRadixWaves;
constexpr int kRadixMaxBits = 11;
constexpr int kRadixMaxBuckets = 1 << kRadixMaxBits;
constexpr int kScanThreads = 1024;
constexpr int kScanChunkTiles = 32;  // tiles per chunk of the two-level column scan

struct Keygen {
    static constexpr int kThreads = 256;
    struct Args {
        const float* px;
        const float* py;
        uint32_t n;
        float cell, max_x, max_y;
        uint32_t* keys;
        uint32_t* vals;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        uint32_t i = wg * kThreads + threadIdx.x;
        if (i >= a.n) return;
        float gx = a.px[i] / a.cell, gy = a.py[i] / a.cell;
        uint32_t ix = (uint32_t)(int32_t)fast_min(fast_max(gx, 0.0f), a.max_x);
        uint32_t iy = (uint32_t)(int32_t)fast_min(fast_max(gy, 0.0f), a.max_y);
        a.keys[i] = morton(ix, iy);
        a.vals[i] = i;
    }
};

// Lanes of the wave whose digit equals this lane's (valid lanes only).  When
// the wave's digits form few runs (the usual case for Morton-sorted particles
// that moved little since the last sort), one ballot per distinct digit;
// otherwise one ballot per digit bit.
__device__ inline uint64_t digit_peers(uint32_t d, bool valid, int bits) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t prev = (uint32_t)__shfl((int)d, (lane + 63) & 63, 64);
    const uint64_t act = __ballot(valid);
    const int runs = __popcll(__ballot(valid && (lane == 0 || prev != d)));
    if (runs <= 12) {
        uint64_t remaining = act, peers = 0;
        while (remaining) {
            const int leader = (int)__builtin_ctzll(remaining);
            const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)d, leader);
            const uint64_t m = __ballot(d == v) & remaining;
            if (valid && d == v) peers = m;
            remaining &= ~m;
        }
        return peers;
    }
    uint64_t peers = act;
    for (int b = 0; b < bits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return valid ? peers : 0ull;
}

// A wave whose digits are all equal adds its count once; otherwise each lane
// adds itself (LDS atomics).
struct RadixUpsweep {
    static constexpr int kThreads = kRadixThreads;
    struct Args {
        const uint32_t* keys;
        uint32_t n;
        int shift, bits;
        uint32_t* counts;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        __shared__ uint32_t hist[kRadixMaxBuckets];
        const int tid = threadIdx.x, lane = tid & 63;
        const uint32_t n = a.n;
        const int shift = a.shift;
        const uint32_t nb = 1u << a.bits, mask = nb - 1u;
        for (uint32_t d = tid; d < nb; d += kRadixThreads) hist[d] = 0;
        const size_t base = (size_t)wg * kRadixTile;
        uint32_t k[kRadixItems];
#pragma unroll
        for (int j = 0; j < kRadixItems; ++j) {
            const size_t idx = base + (size_t)j * kRadixThreads + tid;
            k[j] = idx < n ? a.keys[idx] : 0u;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kRadixItems; ++j) {
            const size_t idx = base + (size_t)j * kRadixThreads + tid;
            const bool valid = idx < n;
            const uint32_t d = (k[j] >> shift) & mask;
            const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
            const uint64_t act = __ballot(valid);
            const uint64_t same = __ballot(valid && d == d0);
            if (same == act) {
                if (act != 0 && lane == (int)__builtin_ctzll(act)) atomicAdd(&hist[d0], (uint32_t)__popcll(act));
            } else if (valid) {
                atomicAdd(&hist[d], 1u);
            }
        }
        __syncthreads();
        for (uint32_t d = tid; d < nb; d += kRadixThreads) a.counts[(size_t)wg * nb + d] = hist[d];
    }
};

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

// Column scan, level 1: for each chunk of kScanChunkTiles tiles and each
// digit, the exclusive prefix of the tile counts within the chunk (in place)
// and the chunk's total.  Threads of a block take consecutive digits, so every
// access is coalesced along the tile-major rows.
struct ScanColumns {
    static constexpr int kThreads = 256;
    struct Args {
        uint32_t* counts;
        uint32_t ntiles, nb, nchunks;
        uint32_t* chunk_sum;
    };
    // workgroup wg: chunk wg % nchunks, digits from 256 * (wg / nchunks)
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        const uint32_t nb = a.nb, ntiles = a.ntiles;
        uint32_t* __restrict__ counts = a.counts;
        const uint32_t chunk = wg % a.nchunks;
        const uint32_t t0 = chunk * kScanChunkTiles, t1 = min(t0 + (uint32_t)kScanChunkTiles, ntiles);
        const uint32_t d = (wg / a.nchunks) * kThreads + threadIdx.x;
        if (d >= nb) return;
        uint32_t c[kScanChunkTiles];
#pragma unroll
        for (int q = 0; q < kScanChunkTiles; ++q) c[q] = t0 + q < t1 ? counts[(size_t)(t0 + q) * nb + d] : 0u;
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < kScanChunkTiles; ++q) {
            if (t0 + q < t1) counts[(size_t)(t0 + q) * nb + d] = run;
            run += c[q];
        }
        a.chunk_sum[(size_t)chunk * nb + d] = run;
    }
};

// Column scan, level 2 (one workgroup): per digit the exclusive prefix over
// chunks (in place), and each digit's start (exclusive scan over digits of
// the digit totals).  A thread takes up to two consecutive digits.
struct ScanChunks {
    static constexpr int kThreads = kScanThreads;
    struct Args {
        uint32_t* chunk_sum;
        uint32_t nchunks, nb;
        uint32_t* digit_start;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t, int) {
        __shared__ uint32_t wave_sums[kScanThreads / 64 + 1];
        const uint32_t nb = a.nb, nchunks = a.nchunks;
        uint32_t* __restrict__ chunk_sum = a.chunk_sum;
        const uint32_t per = (nb + kScanThreads - 1) / kScanThreads;  // 1 or 2
        const uint32_t d0 = threadIdx.x * per;
        uint32_t run[2] = {0u, 0u};
        // Each step loads a batch of chunk totals, then rewrites them: one
        // wait per batch (32 chunks with one digit a thread when there are
        // more than 16, else 16; larger batches spill)
        if (per == 1 && nchunks > 16) {
            for (uint32_t c0 = 0; c0 < nchunks; c0 += 32) {
                uint32_t v[32];
#pragma unroll
                for (int c = 0; c < 32; ++c) {
                    const bool ok = d0 < nb && c0 + c < nchunks;
                    v[c] = ok ? chunk_sum[(size_t)(c0 + c) * nb + d0] : 0u;
                }
#pragma unroll
                for (int c = 0; c < 32; ++c) {
                    if (d0 < nb && c0 + c < nchunks) chunk_sum[(size_t)(c0 + c) * nb + d0] = run[0];
                    run[0] += v[c];
                }
            }
        } else {
            for (uint32_t c0 = 0; c0 < nchunks; c0 += 16) {
                uint32_t v[2][16];
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int c = 0; c < 16; ++c) {
                        const bool ok = (uint32_t)q < per && d0 + q < nb && c0 + c < nchunks;
                        v[q][c] = ok ? chunk_sum[(size_t)(c0 + c) * nb + d0 + q] : 0u;
                    }
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int c = 0; c < 16; ++c) {
                        if ((uint32_t)q < per && d0 + q < nb && c0 + c < nchunks)
                            chunk_sum[(size_t)(c0 + c) * nb + d0 + q] = run[q];
                        run[q] += v[q][c];
                    }
            }
        }
        uint32_t total;
        uint32_t start = block_exclusive_scan(run[0] + run[1], wave_sums, total);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if ((uint32_t)q < per && d0 + q < nb) a.digit_start[d0 + q] = start;
            start += run[q];
        }
    }
};

// prefix[tile][digit] + chunk_base[chunk of tile][digit] + digit_start[digit]
// is the tile's first destination for the digit.
// After ranking, the tile is reordered by digit in LDS, so that consecutive
// threads write consecutive destinations of one digit.
__device__ __forceinline__ void radix_downsweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n, int shift, int bits, const uint32_t* __restrict__ prefix,
    const uint32_t* __restrict__ chunk_base, const uint32_t* __restrict__ digit_start, uint32_t wg) {
    constexpr int kPerThread = kRadixMaxBuckets / kRadixThreads;  // digits per thread in the tile scan
    __shared__ uint32_t whist[kRadixWaves][kRadixMaxBuckets];      // per-wave counts, then the staged tile
    __shared__ uint32_t gdelta[kRadixMaxBuckets];                  // global start - tile start, per digit
    __shared__ uint32_t wsum[kRadixWaves + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nb = 1u << bits, mask = nb - 1u;
    for (uint32_t d = tid; d < nb; d += kRadixThreads) {
#pragma unroll
        for (int w = 0; w < kRadixWaves; ++w) whist[w][d] = 0;
    }
    const size_t tile0 = (size_t)wg * kRadixTile;
    const size_t base = tile0 + (size_t)wave * kRadixWaveItems;
    uint32_t k[kRadixItems], v[kRadixItems], r[kRadixItems];
#pragma unroll
    for (int j = 0; j < kRadixItems; ++j) {
        const size_t idx = base + (size_t)j * 64 + lane;
        const bool valid = idx < n;
        k[j] = valid ? keys_in[idx] : 0u;
        v[j] = valid ? vals_in[idx] : 0u;
    }
    __syncthreads();
    // ranks within the wave, in record order
    const uint64_t lanes_below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t* wh = whist[wave];
#pragma unroll
    for (int j = 0; j < kRadixItems; ++j) {
        const size_t idx = base + (size_t)j * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = (k[j] >> shift) & mask;
        const uint64_t peers = digit_peers(d, valid, bits);
        const uint32_t below = (uint32_t)__popcll(peers & lanes_below);
        const uint32_t before = valid ? wh[d] : 0u;
        r[j] = before + below;
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) wh[d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // tile-local start of every digit (exclusive scan over digits of the
    // tile's totals); whist[w][d] becomes the tile-local start of wave w's
    // records of digit d
    uint32_t tot[kPerThread];
    uint32_t local = 0;
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t d = (uint32_t)tid * kPerThread + q;
        uint32_t t = 0;
        if (d < nb) {
#pragma unroll
            for (int w = 0; w < kRadixWaves; ++w) t += whist[w][d];
        }
        tot[q] = t;
        local += t;
    }
    uint32_t incl = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t start = incl - local;
    for (int w = 0; w < wave; ++w) start += wsum[w];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t d = (uint32_t)tid * kPerThread + q;
        if (d < nb) {
            gdelta[d] = prefix[(size_t)wg * nb + d] +
                        chunk_base[(size_t)(wg / kScanChunkTiles) * nb + d] + digit_start[d] - start;
            uint32_t run = start;
#pragma unroll
            for (int w = 0; w < kRadixWaves; ++w) {
                const uint32_t c = whist[w][d];
                whist[w][d] = run;
                run += c;
            }
        }
        start += tot[q];
    }
    __syncthreads();
    uint32_t pos[kRadixItems];
#pragma unroll
    for (int j = 0; j < kRadixItems; ++j) pos[j] = wh[(k[j] >> shift) & mask] + r[j];
    __syncthreads();
    // stage the tile in digit order (the count arrays are free now)
    uint32_t* sk = &whist[0][0];
    uint32_t* sv = &whist[kRadixWaves / 2][0];
#pragma unroll
    for (int j = 0; j < kRadixItems; ++j) {
        const size_t idx = base + (size_t)j * 64 + lane;
        if (idx < n) { sk[pos[j]] = k[j]; sv[pos[j]] = v[j]; }
    }
    __syncthreads();
    const uint32_t count = (uint32_t)min((size_t)kRadixTile, (size_t)n - tile0);
    for (uint32_t t = tid; t < count; t += kRadixThreads) {
        const uint32_t key = sk[t];
        const uint32_t dst = gdelta[(key >> shift) & mask] + t;
        keys_out[dst] = key;
        vals_out[dst] = sv[t];
    }
}

struct RadixDownsweep {
    static constexpr int kThreads = kRadixThreads;
    struct Args {
        const uint32_t* keys_in;
        const uint32_t* vals_in;
        uint32_t* keys_out;
        uint32_t* vals_out;
        uint32_t n;
        int shift, bits;
        const uint32_t* prefix;
        const uint32_t* chunk_base;
        const uint32_t* digit_start;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        radix_downsweep(a.keys_in, a.vals_in, a.keys_out, a.vals_out, a.n, a.shift, a.bits, a.prefix, a.chunk_base,
                        a.digit_start, wg);
    }
};

struct Permute {
    static constexpr int kThreads = 256;
    struct Args {
        const float* in;
        float* out;
        const uint32_t* vals;
        uint32_t np;
        size_t nalloc;
    };
    static __device__ __forceinline__ void run(const Args& a, uint32_t wg, int) {
        const size_t nalloc = a.nalloc;
        const uint32_t np = a.np;
        size_t i = (size_t)wg * kThreads + threadIdx.x;
        if (i >= nalloc) return;
        if (i < np) {
            size_t j = a.vals[i];
            if (j >= np) j = np - 1;  // records written through the field API are not trusted
#pragma unroll
            for (int f = 0; f < NUM_PFIELDS; ++f) a.out[(size_t)f * nalloc + i] = a.in[(size_t)f * nalloc + j];
        } else {
            // the reference's gather leaves tail lanes of the last chunk zero (src/apic.rs:1077-1103)
#pragma unroll
            for (int f = 0; f < NUM_PFIELDS; ++f) a.out[(size_t)f * nalloc + i] = 0.0f;
        }
    }
};

uint32_t key_bits(const Sim& sim) {
    uint32_t mk = morton(sim.W > 0 ? sim.W - 1 : 0, sim.H > 0 ? sim.H - 1 : 0);
    uint32_t bits = 0;
    while (bits < 32 && (mk >> bits) != 0) ++bits;
    return bits == 0 ? 1 : bits;
}

}  // namespace

size_t radix_scratch_words(size_t n) {
    size_t ntiles = (n + kRadixTile - 1) / kRadixTile;
    size_t nchunks = (ntiles + kScanChunkTiles - 1) / kScanChunkTiles;
    return (size_t)kRadixMaxBuckets * (ntiles + nchunks + 1) + 64;
}

void launch_keygen(Sim& sim) {
    if (sim.num_particles == 0) return;
    uint32_t n = sim.num_particles;
    float max_x = (float)(sim.W > 0 ? sim.W - 1 : 0), max_y = (float)(sim.H > 0 ? sim.H - 1 : 0);
    emit<Keygen>(sim, {sim.pfield(P_X), sim.pfield(P_Y), n, sim.cell, max_x, max_y, sim.keys[0], sim.vals[0]},
                 (n + 255) / 256);
}

void launch_radix_sort(Sim& sim) {
    uint32_t n = sim.num_particles;
    if (n <= 1) return;
    // Passes of at most dmax <= kRadixMaxBits bits (the tuning's); the first
    // (least significant) pass takes the remainder, so the later passes get
    // the widest digits.
    const int dmax = (int)std::min<uint32_t>(std::max<uint32_t>(sim.tuning.radix_bits, 1u), (uint32_t)kRadixMaxBits);
    const int total_bits = (int)key_bits(sim);
    const int passes = (total_bits + dmax - 1) / dmax;
    const uint32_t ntiles = (n + kRadixTile - 1) / kRadixTile;
    const uint32_t nchunks = (ntiles + kScanChunkTiles - 1) / kScanChunkTiles;
    uint32_t* counts = sim.radix_scratch;                               // [tile][digit]
    uint32_t* chunk_sum = counts + (size_t)kRadixMaxBuckets * ntiles;   // [chunk][digit]
    uint32_t* digit_start = chunk_sum + (size_t)kRadixMaxBuckets * nchunks;
    int src = 0, shift = 0;
    for (int p = 0; p < passes; ++p) {
        const int bits = p == 0 ? total_bits - (passes - 1) * dmax : dmax;
        const uint32_t nb = 1u << bits;
        emit<RadixUpsweep>(sim, {sim.keys[src], n, shift, bits, counts}, ntiles);
        emit<ScanColumns>(sim, {counts, ntiles, nb, nchunks, chunk_sum}, nchunks * ((nb + 255) / 256));
        emit<ScanChunks>(sim, {chunk_sum, nchunks, nb, digit_start}, 1);
        emit<RadixDownsweep>(sim, {sim.keys[src], sim.vals[src], sim.keys[1 - src], sim.vals[1 - src], n, shift, bits,
                                   counts, chunk_sum, digit_start},
                             ntiles);
        src = 1 - src;
        shift += bits;
    }
    if (src != 0) {
        std::swap(sim.keys[0], sim.keys[1]);
        std::swap(sim.vals[0], sim.vals[1]);
    }
}

void launch_permute(Sim& sim) {
    if (sim.nalloc == 0) return;
    size_t n = sim.nalloc;
    emit<Permute>(sim, {sim.pbuf[sim.cur], sim.pbuf[1 - sim.cur], sim.vals[0], sim.num_particles, n},
                  (uint32_t)((n + 255) / 256));
    sim.cur = 1 - sim.cur;
}

}  // namespace apic
