// launch.hpp -- one kernel body for one simulation or for many at once.
//
// Every kernel of the library is a struct K with
//   static constexpr int kThreads;     workgroup size
//   struct Args;                       all that one simulation's launch reads
//   static __device__ void run(const Args& a, uint32_t wg, int k);
//                                      the body for workgroup wg of that launch
//                                      (k: the PCG iteration, else unused)
// one_kernel<K> runs one simulation's launch with the arguments by value.
// batch_kernel<K> runs the launches of many simulations as one grid: an
// argument table holds one Args per simulation and a map gives, per
// workgroup of the grid, (table row << kMapWgBits) | workgroup of that
// simulation's launch.  The body cannot tell the two apart, so a simulation
// computes the same bits alone and in a batch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <type_traits>
#include <vector>

namespace apic {

constexpr int kMapWgBits = 20;                           // workgroups per simulation launch: < 2^20
constexpr uint32_t kMapWgMask = (1u << kMapWgBits) - 1;
constexpr uint32_t kMaxBatchRows = 1u << (32 - kMapWgBits);  // simulations per batched launch

// Workgroup i of a grid runs on XCD i % kXcds (round-robin dispatch), and
// every XCD has its own L2.  The logical workgroup index gives each XCD a
// contiguous band of the grid's logical workgroups, so that the neighbouring
// rows a stencil reads mostly come from the same L2.  A bijection of
// [0, n): every logical workgroup still runs exactly once.
constexpr uint32_t kXcds = 8;
__host__ __device__ __forceinline__ uint32_t xcd_band(uint32_t hw, uint32_t n) {
    const uint32_t xcd = hw % kXcds, slot = hw / kXcds;
    const uint32_t q = n / kXcds, r = n % kXcds;
    return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// A kernel over a row-major grid of tiles declares kDealRows = true (and has
// Args::tiles_x): its tile rows are dealt to the XCDs in turn (row i to XCD
// i % kXcds), so that a fluid region covering only some rows of the grid (a
// dam break's lower part) still spreads over every XCD, while a tile's left
// and right neighbours stay in its L2.  Workgroups past the last whole round
// of kXcds rows are laid out in bands.  A bijection of [0, n).
__host__ __device__ __forceinline__ uint32_t xcd_deal(uint32_t hw, uint32_t n, uint32_t c) {
    const uint32_t full = c ? n / (kXcds * c) * (kXcds * c) : 0;
    if (hw < full) {
        const uint32_t xcd = hw % kXcds, slot = hw / kXcds;
        return ((slot / c) * kXcds + xcd) * c + slot % c;
    }
    return full + xcd_band(hw - full, n - full);
}

// A kernel whose work per workgroup follows the particles (uneven over the
// grid) declares kRoundRobin = true and keeps the dispatch order: bands would
// leave the XCDs that hold empty regions idle.
template <class K, class = void>
struct round_robin : std::false_type {};
template <class K>
struct round_robin<K, std::void_t<decltype(K::kRoundRobin)>> : std::bool_constant<K::kRoundRobin> {};

template <class K, class = void>
struct deal_rows : std::false_type {};
template <class K>
struct deal_rows<K, std::void_t<decltype(K::kDealRows)>> : std::bool_constant<K::kDealRows> {};

// The workgroups of one tile row for a kDealRows kernel, else 0 (bands).
template <class K>
__host__ __device__ __forceinline__ uint32_t deal_chunk(const typename K::Args& a) {
    if constexpr (deal_rows<K>::value) return a.tiles_x;
    else return 0;
}

template <class K>
__global__ __launch_bounds__(K::kThreads) void one_kernel(typename K::Args a, int k) {
    uint32_t wg;
    if constexpr (round_robin<K>::value) wg = blockIdx.x;
    else if constexpr (deal_rows<K>::value) wg = xcd_deal(blockIdx.x, gridDim.x, a.tiles_x);
    else wg = xcd_band(blockIdx.x, gridDim.x);
    K::run(a, wg, k);
}

// The argument table is read through the constant address space, as kernel
// arguments are: the rows are never written while a launch runs, and the
// loads of a workgroup's (uniform) row then go through the scalar unit.
constexpr uint32_t kMapSkip = 0xFFFFFFFFu;  // a map entry that runs nothing (XCD queue padding)

template <class K>
__global__ __launch_bounds__(K::kThreads) void batch_kernel(const typename K::Args* __restrict__ tab,
                                                            const uint32_t* __restrict__ map, int k) {
    using ConstArgs = const __attribute__((address_space(4))) typename K::Args;
    const uint32_t m = map[blockIdx.x];
    if (m == kMapSkip) return;
    const ConstArgs* row = (const ConstArgs*)tab + (m >> kMapWgBits);
    K::run(*(const typename K::Args*)row, m & kMapWgMask, k);
}

// Collects the launches that the host-side launchers make for several
// simulations, then issues each kind of launch once for all of them.
//
// Launches are grouped by (kernel, how many launches of that kernel the
// simulation made before): the n-th smoothing pass of every simulation is one
// group, the n-th V-cycle level on the way down another.  Groups are issued
// in an order that keeps every simulation's own launch order (a topological
// order of the groups, ties broken by first appearance), so a group depends
// only on groups issued before it.
class Recorder {
public:
    struct Launch {
        const void* fn;      // batch_kernel<K>
        uint32_t threads;
        uint32_t grid;       // workgroups over all simulations of the group
        size_t tab_off;      // argument table: offset in the uploaded arena
        size_t map_off;      // workgroup map: offset in the uploaded arena
        int k;               // uniform argument recorded with the group
    };

    void begin(uint32_t sim);  // the following launches belong to simulation `sim`
    // bands: lay each simulation's workgroups out on the XCDs as one_kernel
    // does (in bands, or dealt in rows of `chunk` workgroups when chunk > 0),
    // else in dispatch order
    void add(const void* fn, size_t arg_size, const void* args, uint32_t threads, uint32_t nwg, int k, bool bands,
             uint32_t chunk = 0);
    bool empty() const { return groups_.empty(); }
    // a simulation's launch had more workgroups than a map entry can name
    bool overflow() const { return overflow_; }
    // Appends the tables and maps to host[used, ...) (growing it as needed,
    // advancing `used`) and returns the launches in issue order, with offsets
    // from host[0].  An empty result with launches recorded means that the
    // simulations' launch orders conflict.  The caller uploads host[0, used)
    // to a device arena and issues the launches with `issue`.
    std::vector<Launch> compile(std::vector<uint8_t>& host, size_t& used);
    void clear();

private:
    struct Group {
        const void* fn;
        uint32_t occurrence;
        size_t arg_size;
        uint32_t threads;
        int k;
        bool bands;
        std::vector<uint8_t> args;
        std::vector<uint32_t> nwg;  // per table row
        std::vector<uint32_t> chunk;  // per table row: workgroups of a dealt row (0: bands)
    };
    std::vector<Group> groups_;
    std::vector<std::vector<int>> chains_;  // per simulation: its groups in launch order
    std::vector<std::vector<const void*>> chain_fns_;  // and their kernels (the occurrence count scans them)
    std::vector<int> hint_;  // per chain position: the group the last simulation's launch there joined
    uint32_t cur_ = 0;
    bool overflow_ = false;
    int find_or_add(const void* fn, uint32_t occurrence, size_t arg_size, uint32_t threads, int k, bool bands);
};

// Issues compiled launches whose tables were uploaded to `dev`.  k >= 0
// replaces the recorded uniform argument (the PCG iteration).
hipError_t issue(const std::vector<Recorder::Launch>& launches, const uint8_t* dev, hipStream_t stream, int k = -1);

}  // namespace apic
