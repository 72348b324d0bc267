// batch.hip -- the recorder that turns the launches of several simulations
// into one launch per kind (launch.hpp), and the issue of compiled launches.
#include <algorithm>
#include <queue>

#include "launch.hpp"

namespace apic {

void Recorder::clear() {
    groups_.clear();
    chains_.clear();
    chain_fns_.clear();
    hint_.clear();
    cur_ = 0;
    overflow_ = false;
}

void Recorder::begin(uint32_t sim) {
    cur_ = sim;
    if (chains_.size() <= sim) {
        chains_.resize(sim + 1);
        chain_fns_.resize(sim + 1);
    }
}

int Recorder::find_or_add(const void* fn, uint32_t occurrence, size_t arg_size, uint32_t threads, int k,
                          bool bands) {
    // simulations mostly record the same launches in the same order: try the
    // group the previous one's launch at this position joined first
    const size_t pos = chains_[cur_].size();
    if (pos < hint_.size()) {
        const int h = hint_[pos];
        const Group& G = groups_[h];
        if (G.fn == fn && G.occurrence == occurrence && G.k == k) return h;
    }
    for (size_t g = 0; g < groups_.size(); ++g) {
        const Group& G = groups_[g];
        if (G.fn == fn && G.occurrence == occurrence && G.k == k) return (int)g;
    }
    Group G;
    G.fn = fn;
    G.occurrence = occurrence;
    G.arg_size = arg_size;
    G.threads = threads;
    G.k = k;
    G.bands = bands;
    groups_.push_back(std::move(G));
    return (int)groups_.size() - 1;
}

void Recorder::add(const void* fn, size_t arg_size, const void* args, uint32_t threads, uint32_t nwg, int k,
                   bool bands, uint32_t chunk) {
    if (nwg > kMapWgMask || cur_ >= kMaxBatchRows) overflow_ = true;  // kMapWgMask itself stays unused (kMapSkip)
    std::vector<int>& chain = chains_[cur_];
    std::vector<const void*>& fns = chain_fns_[cur_];
    uint32_t occurrence = 0;
    for (const void* f : fns) occurrence += f == fn;
    const int g = find_or_add(fn, occurrence, arg_size, threads, k, bands);
    const size_t pos = chain.size();
    if (hint_.size() <= pos) hint_.resize(pos + 1);
    hint_[pos] = g;
    fns.push_back(fn);
    Group& G = groups_[g];
    const uint8_t* p = static_cast<const uint8_t*>(args);
    G.args.insert(G.args.end(), p, p + arg_size);
    G.nwg.push_back(nwg);
    G.chunk.push_back(chunk);
    chain.push_back(g);
}

std::vector<Recorder::Launch> Recorder::compile(std::vector<uint8_t>& host, size_t& used) {
    std::vector<Launch> out;
    const size_t ng = groups_.size();
    if (ng == 0 || overflow_) return out;
    // topological order of the groups over every simulation's chain, ties by
    // first appearance (lowest group index)
    std::vector<std::vector<int>> succ(ng);
    std::vector<int> indeg(ng, 0);
    for (const auto& chain : chains_)
        for (size_t i = 1; i < chain.size(); ++i) {
            succ[chain[i - 1]].push_back(chain[i]);
            ++indeg[chain[i]];
        }
    std::priority_queue<int, std::vector<int>, std::greater<int>> ready;
    for (size_t g = 0; g < ng; ++g)
        if (indeg[g] == 0) ready.push((int)g);
    std::vector<int> order;
    order.reserve(ng);
    while (!ready.empty()) {
        const int g = ready.top();
        ready.pop();
        order.push_back(g);
        for (int s : succ[g])
            if (--indeg[s] == 0) ready.push(s);
    }
    if (order.size() != ng) return out;
    auto align = [](size_t v) { return (v + 255) / 256 * 256; };
    // XCD bands: workgroup w of a simulation with n workgroups goes to the
    // XCD whose band holds w (as xcd_band maps them for one simulation),
    // rotated by the workgroups laid out before it so that the queues stay
    // even; a simulation whose launch deals rows (chunk > 0) sends its row i
    // to XCD (i + rotation) % kXcds instead, rotated by its rows.  The grid
    // interleaves the eight queues (position 8 i + x runs the i-th entry of
    // XCD x), padded with entries that run nothing.
    std::vector<std::vector<uint32_t>> maps(ng);
    for (int g : order) {
        const Group& G = groups_[g];
        std::vector<uint32_t>& m = maps[g];
        size_t total_wg = 0;
        for (uint32_t n : G.nwg) total_wg += n;
        if (!G.bands) {
            m.reserve(total_wg);
            for (size_t row = 0; row < G.nwg.size(); ++row)
                for (uint32_t wg = 0; wg < G.nwg[row]; ++wg) m.push_back(((uint32_t)row << kMapWgBits) | wg);
            continue;
        }
        std::vector<uint32_t> q[kXcds];
        for (uint32_t x = 0; x < kXcds; ++x) q[x].reserve(total_wg / kXcds + G.nwg.size());
        uint32_t rot = 0;
        for (size_t row = 0; row < G.nwg.size(); ++row) {
            const uint32_t n = G.nwg[row], qn = n / kXcds, rn = n % kXcds, c = G.chunk[row];
            if (c > 0) {
                const uint32_t rows = (n + c - 1) / c;
                for (uint32_t i = 0; i < rows; ++i) {
                    std::vector<uint32_t>& dst = q[(i + rot) % kXcds];
                    for (uint32_t j = i * c; j < n && j < (i + 1) * c; ++j)
                        dst.push_back(((uint32_t)row << kMapWgBits) | j);
                }
                rot = (rot + rows) % kXcds;
                continue;
            }
            for (uint32_t x = 0; x < kXcds; ++x) {
                const uint32_t len = qn + (x < rn ? 1u : 0u), first = x < rn ? x * (qn + 1) : rn * (qn + 1) + (x - rn) * qn;
                std::vector<uint32_t>& dst = q[(x + rot) % kXcds];
                for (uint32_t j = 0; j < len; ++j) dst.push_back(((uint32_t)row << kMapWgBits) | (first + j));
            }
            rot = (rot + rn) % kXcds;
        }
        size_t longest = 0;
        for (uint32_t x = 0; x < kXcds; ++x) longest = std::max(longest, q[x].size());
        m.assign(longest * kXcds, kMapSkip);
        for (uint32_t x = 0; x < kXcds; ++x)
            for (size_t i = 0; i < q[x].size(); ++i) m[i * kXcds + x] = q[x][i];
    }
    size_t total = align(used);
    for (int g : order) {
        total = align(total) + groups_[g].args.size();
        total = align(total) + maps[g].size() * sizeof(uint32_t);
    }
    total = align(total);
    if (host.size() < total) host.resize(total);
    size_t off = used;
    for (int g : order) {
        const Group& G = groups_[g];
        Launch L{};
        L.fn = G.fn;
        L.threads = G.threads;
        L.k = G.k;
        off = align(off);
        L.tab_off = off;
        std::memcpy(host.data() + off, G.args.data(), G.args.size());
        off += G.args.size();
        off = align(off);
        L.map_off = off;
        std::memcpy(host.data() + off, maps[g].data(), maps[g].size() * sizeof(uint32_t));
        off += maps[g].size() * sizeof(uint32_t);
        L.grid = (uint32_t)maps[g].size();
        out.push_back(L);
    }
    used = align(off);
    return out;
}

hipError_t issue(const std::vector<Recorder::Launch>& launches, const uint8_t* dev, hipStream_t stream, int k) {
    for (const Recorder::Launch& L : launches) {
        const void* tab = dev + L.tab_off;
        const void* map = dev + L.map_off;
        int kk = k >= 0 ? k : L.k;
        void* params[3] = {&tab, &map, &kk};
        hipError_t e = hipLaunchKernel(L.fn, dim3(L.grid), dim3(L.threads), params, 0, stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace apic
