// State storage, device-wide scans, relayout and reductions.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "blockscan.hpp"
#include "engine.hpp"
#include "keyops.hpp"

namespace sparq {

// ------------------------------------------------------------------ memory
// Caching device allocator.  Requests are rounded up to a size class (sixteen steps
// per power of two, so at most 12.5 % slack) and a released block goes onto a
// free list instead of back to the runtime.  Every block remembers the stream that
// released it: that stream may reuse it at once (its queue orders the old and new
// uses); any other stream first waits on an event recorded on the releasing stream.
// A stream that is being destroyed is drained and its blocks are untagged.  The
// runtime is called only on a miss (hipMalloc) and when a miss runs out of memory
// (the cache is then synchronised, returned and the miss retried).
namespace {
struct CachedBlock {
    void* p;
    hipStream_t stream;   // last releasing stream; nullptr = no pending use
};
std::mutex g_mem_mu;
std::unordered_map<void*, size_t> g_live;              // pointer -> class bytes
std::multimap<size_t, CachedBlock> g_free_blocks;      // class bytes -> free blocks
size_t g_cached_bytes = 0, g_live_bytes = 0, g_peak_bytes = 0;
uint64_t g_misses = 0, g_hits = 0;

size_t size_class(size_t b) {
    if (b <= 4096) return 4096;
    int e = 63 - __builtin_clzll((unsigned long long)(b - 1));   // 2^e < b <= 2^(e+1)
    size_t step = size_t(1) << (e - 3 > 12 ? e - 3 : 12);        // eighth of 2^e
    return (b + step - 1) / step * step;
}

void release_cache_locked() {
    (void)hipDeviceSynchronize();
    for (auto& kv : g_free_blocks) (void)hipFree(kv.second.p);
    g_free_blocks.clear();
    g_cached_bytes = 0;
}

// Live + cached bytes above which a miss first returns cached blocks (largest
// first) to the runtime: 85 % of the device memory.
size_t cache_budget() {
    static size_t budget = [] {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || total_b == 0) return ~size_t(0);
        return total_b / 20 * 17;
    }();
    return budget;
}

void trim_cache_locked(size_t need) {
    const size_t budget = cache_budget();
    if (g_live_bytes + g_cached_bytes + need <= budget || g_free_blocks.empty()) return;
    (void)hipDeviceSynchronize();
    while (!g_free_blocks.empty() && g_live_bytes + g_cached_bytes + need > budget) {
        auto it = std::prev(g_free_blocks.end());
        (void)hipFree(it->second.p);
        g_cached_bytes -= it->first;
        g_free_blocks.erase(it);
    }
}
}  // namespace

void* dmalloc(size_t bytes, hipStream_t s) {
    const size_t cls0 = size_class(bytes == 0 ? 1 : bytes);
    std::lock_guard<std::mutex> lock(g_mem_mu);
    // A cached block of the request's class, else the smallest one of up to twice
    // its size (sizes drift from operator to operator); a block last used on the
    // same stream (or on none) is preferred.
    auto pick = g_free_blocks.end();
    for (auto it = g_free_blocks.lower_bound(cls0); it != g_free_blocks.end() && it->first <= 2 * cls0; ++it) {
        if (it->second.stream == s || it->second.stream == nullptr) {
            pick = it;
            break;
        }
        if (pick == g_free_blocks.end()) pick = it;
        if (it->first != pick->first) break;
    }
    void* p = nullptr;
    size_t cls = cls0;
    if (pick != g_free_blocks.end()) {
        cls = pick->first;
        CachedBlock blk = pick->second;
        g_free_blocks.erase(pick);
        g_cached_bytes -= cls;
        if (blk.stream != s && blk.stream != nullptr) {
            hipEvent_t ev;
            SPQ_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            SPQ_HIP(hipEventRecord(ev, blk.stream));
            SPQ_HIP(hipStreamWaitEvent(s, ev, 0));
            (void)hipEventDestroy(ev);
        }
        p = blk.p;
        ++g_hits;
    } else {
        trim_cache_locked(cls);
        hipError_t e = hipMalloc(&p, cls);
        if (e == hipErrorOutOfMemory && !g_free_blocks.empty()) {
            (void)hipGetLastError();
            release_cache_locked();
            e = hipMalloc(&p, cls);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(e == hipErrorOutOfMemory ? SPARQ_OOM : SPARQ_RUNTIME,
                        std::string("device allocation of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
        }
        ++g_misses;
    }
    g_live[p] = cls;
    g_live_bytes += cls;
    g_peak_bytes = std::max(g_peak_bytes, g_live_bytes);
    return p;
}

void dfree(void* p, hipStream_t s) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(g_mem_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) return;   // not ours: never happens for engine buffers
    const size_t cls = it->second;
    g_live.erase(it);
    g_live_bytes -= cls;
    g_free_blocks.emplace(cls, CachedBlock{p, s});
    g_cached_bytes += cls;
}

void forget_stream(hipStream_t s) {
    std::lock_guard<std::mutex> lock(g_mem_mu);
    for (auto& kv : g_free_blocks)
        if (kv.second.stream == s) kv.second.stream = nullptr;
}

void memory_stats(uint64_t out[6]) {
    std::lock_guard<std::mutex> lock(g_mem_mu);
    out[0] = g_live_bytes;
    out[1] = g_cached_bytes;
    out[2] = g_peak_bytes;
    out[3] = g_hits;
    out[4] = g_misses;
    out[5] = g_free_blocks.size();
}

void release_cached_memory() {
    std::lock_guard<std::mutex> lock(g_mem_mu);
    release_cache_locked();
}

void alloc_arrays(State& st, uint32_t W, uint64_t cap, uint64_t** keys, double2** amps) {
    if (cap == 0) cap = 1;
    uint64_t* k = dalloc<uint64_t>((size_t)W * cap, st.stream);
    double2* a = nullptr;
    try {
        a = dalloc<double2>(cap, st.stream);
    } catch (...) {
        dfree(k, st.stream);
        throw;
    }
    *keys = k;
    *amps = a;
}

void install_arrays(State& st, uint32_t W, uint64_t* keys, double2* amps, uint64_t n, uint64_t cap) {
    if (st.keys != keys) dfree(st.keys, st.stream);
    if (st.amps != amps) dfree(st.amps, st.stream);
    st.keys = keys;
    st.amps = amps;
    st.W = W;
    st.Wcap = 0;
    st.n = n;
    st.cap = cap;
}

void resize_words(State& st, uint32_t new_words) {
    SPQ_REQUIRE(new_words >= 1 && new_words <= SPARQ_MAX_WORDS, "key width out of range");
    const uint32_t have = std::max(st.W, st.Wcap);
    if (new_words <= have) {   // the allocation already holds the words: no data motion
        if (new_words > st.W)
            SPQ_HIP(hipMemsetAsync(st.keys + (uint64_t)st.W * st.cap, 0, (size_t)(new_words - st.W) * st.cap * sizeof(uint64_t), st.stream));
        st.Wcap = have;
        st.W = new_words;
        return;
    }
    uint64_t* k = dalloc<uint64_t>((size_t)new_words * st.cap, st.stream);
    SPQ_HIP(hipMemcpyAsync(k, st.keys, (size_t)st.W * st.cap * sizeof(uint64_t), hipMemcpyDeviceToDevice, st.stream));
    SPQ_HIP(hipMemsetAsync(k + (uint64_t)st.W * st.cap, 0, (size_t)(new_words - st.W) * st.cap * sizeof(uint64_t), st.stream));
    dfree(st.keys, st.stream);
    st.keys = k;
    st.W = new_words;
    st.Wcap = 0;
}

// Readbacks of up to kPinnedBytes go through the state's pinned words (a copy
// into pageable memory is staged by the runtime and costs more per call).
constexpr size_t kPinnedBytes = 64 * sizeof(uint64_t);

void read_back(State& st, const void* dptr, size_t bytes, void* host) {
    const bool on = prof_enabled();
    const auto t0 = std::chrono::steady_clock::now();
    if (bytes <= kPinnedBytes) {
        SPQ_HIP(hipMemcpyAsync(st.pinned, dptr, bytes, hipMemcpyDeviceToHost, st.stream));
        SPQ_HIP(hipStreamSynchronize(st.stream));
        std::memcpy(host, st.pinned, bytes);
    } else {
        SPQ_HIP(hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, st.stream));
        SPQ_HIP(hipStreamSynchronize(st.stream));
    }
    if (on) prof_host_wait("host:readback", std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
}

uint64_t read_u64(State& st, const uint64_t* dptr) {
    uint64_t v = 0;
    read_back(st, dptr, sizeof(v), &v);
    return v;
}

// ------------------------------------------------------------------ scans
// 2048 elements per workgroup (256 threads x 8 items).
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

// The kScanItems consecutive inputs of a lane.  VEC (input 16-byte aligned):
// a lane whose items all lie below n reads them with 16-byte loads (its items
// start at a multiple of 32 bytes); the others read item by item.
template <class TIn, bool VEC>
__device__ __forceinline__ void scan_load(const TIn* __restrict__ in, uint64_t n, uint64_t base, uint64_t (&v)[kScanItems]) {
    static_assert(kScanItems * sizeof(TIn) % 16 == 0, "items must fill whole 16-byte words");
    if (VEC && base + kScanItems <= n) {
        constexpr int kWords = kScanItems * (int)sizeof(TIn) / 16;
        constexpr int kPerWord = 16 / (int)sizeof(TIn);
        const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int q = 0; q < kWords; ++q) {
            const uint4 w = p[q];
            const uint32_t u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int j = 0; j < kPerWord; ++j) {
                if constexpr (sizeof(TIn) == 4) v[q * kPerWord + j] = u[j];
                else v[q * kPerWord + j] = (uint64_t)u[2 * j] | ((uint64_t)u[2 * j + 1] << 32);   // little-endian
            }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) v[j] = (base + j < n) ? (uint64_t)in[base + j] : 0;
}

template <class TIn, bool VEC>
__global__ void __launch_bounds__(kBlock) scan_reduce_kernel(const TIn* __restrict__ in, uint64_t n, uint64_t* __restrict__ sums) {
    __shared__ uint64_t lds[kBlock / 64];
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems];
    scan_load<TIn, VEC>(in, n, base, v);
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) s += v[j];
    uint64_t tot;
    (void)block_excl_scan<uint64_t>(s, lds, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// VEC: input and output 16-byte aligned; a lane whose items all lie below n
// writes its 8 prefix values with 16-byte stores.
template <class TIn, bool VEC>
__global__ void __launch_bounds__(kBlock) scan_apply_kernel(const TIn* __restrict__ in, uint64_t n, const uint64_t* __restrict__ offs,
                                                       uint64_t* __restrict__ out, uint64_t* __restrict__ total) {
    __shared__ uint64_t lds[kBlock / 64];
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems];
    scan_load<TIn, VEC>(in, n, base, v);
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) s += v[j];
    uint64_t tot;
    uint64_t pre = block_excl_scan<uint64_t>(s, lds, tot) + (offs ? offs[blockIdx.x] : 0);
    if (VEC && base + kScanItems <= n) {
        uint64_t o[kScanItems];
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            o[j] = pre;
            pre += v[j];
        }
        ulonglong2* q = reinterpret_cast<ulonglong2*>(out + base);
#pragma unroll
        for (int j = 0; j < kScanItems / 2; ++j) q[j] = make_ulonglong2(o[2 * j], o[2 * j + 1]);
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            if (base + j < n) out[base + j] = pre;
            pre += v[j];
        }
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) total[0] = pre;
}

template <class TIn>
static void excl_scan_impl(State& st, const TIn* in, uint64_t* out, uint64_t n, uint64_t* total_dev) {
    if (n == 0) {
        if (total_dev) SPQ_HIP(hipMemsetAsync(total_dev, 0, sizeof(uint64_t), st.stream));
        return;
    }
    unsigned nb = grid_for(n, kScanTile);
    const bool vec = (((uintptr_t)in | (uintptr_t)out) & 15) == 0;
    if (nb == 1) {
        {
            SPQ_PROF(st.stream, "scan_apply_kernel", n * ((double)sizeof(TIn) + 8));
            if (vec) scan_apply_kernel<TIn, true><<<1, kBlock, 0, st.stream>>>(in, n, nullptr, out, total_dev);
            else scan_apply_kernel<TIn, false><<<1, kBlock, 0, st.stream>>>(in, n, nullptr, out, total_dev);
        }
        SPQ_HIP(hipGetLastError());
        return;
    }
    uint64_t* sums = dalloc<uint64_t>(nb, st.stream);
    uint64_t* offs = dalloc<uint64_t>(nb, st.stream);
    {
        SPQ_PROF(st.stream, "scan_reduce_kernel", n * (double)sizeof(TIn));
        if (vec) scan_reduce_kernel<TIn, true><<<nb, kBlock, 0, st.stream>>>(in, n, sums);
        else scan_reduce_kernel<TIn, false><<<nb, kBlock, 0, st.stream>>>(in, n, sums);
    }
    SPQ_HIP(hipGetLastError());
    excl_scan_impl<uint64_t>(st, sums, offs, nb, nullptr);
    {
        SPQ_PROF(st.stream, "scan_apply_kernel", n * ((double)sizeof(TIn) + 8));
        if (vec) scan_apply_kernel<TIn, true><<<nb, kBlock, 0, st.stream>>>(in, n, offs, out, total_dev);
        else scan_apply_kernel<TIn, false><<<nb, kBlock, 0, st.stream>>>(in, n, offs, out, total_dev);
    }
    SPQ_HIP(hipGetLastError());
    dfree(sums, st.stream);
    dfree(offs, st.stream);
}

// The same scan with its per-level partial sums in caller-provided device
// memory (2 words per kScanTile inputs at every level above the first) and no
// profiling events: nothing is allocated.  n >= 1.
template <class TIn>
static void excl_scan_scratch_impl(hipStream_t s, const TIn* in, uint64_t* out, uint64_t n, uint64_t* total_dev, uint64_t* scratch,
                                   uint64_t scratch_words) {
    const unsigned nb = grid_for(n, kScanTile);
    const bool vec = (((uintptr_t)in | (uintptr_t)out) & 15) == 0;
    if (nb == 1) {
        if (vec) scan_apply_kernel<TIn, true><<<1, kBlock, 0, s>>>(in, n, nullptr, out, total_dev);
        else scan_apply_kernel<TIn, false><<<1, kBlock, 0, s>>>(in, n, nullptr, out, total_dev);
        SPQ_HIP(hipGetLastError());
        return;
    }
    SPQ_REQUIRE(2 * (uint64_t)nb <= scratch_words, "scan scratch too small");
    uint64_t* sums = scratch;
    uint64_t* offs = scratch + nb;
    if (vec) scan_reduce_kernel<TIn, true><<<nb, kBlock, 0, s>>>(in, n, sums);
    else scan_reduce_kernel<TIn, false><<<nb, kBlock, 0, s>>>(in, n, sums);
    SPQ_HIP(hipGetLastError());
    excl_scan_scratch_impl<uint64_t>(s, sums, offs, nb, nullptr, scratch + 2 * (uint64_t)nb, scratch_words - 2 * (uint64_t)nb);
    if (vec) scan_apply_kernel<TIn, true><<<nb, kBlock, 0, s>>>(in, n, offs, out, total_dev);
    else scan_apply_kernel<TIn, false><<<nb, kBlock, 0, s>>>(in, n, offs, out, total_dev);
    SPQ_HIP(hipGetLastError());
}

void exclusive_scan_u32_scratch(hipStream_t s, const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* total_dev, uint64_t* scratch,
                                uint64_t scratch_words) {
    SPQ_REQUIRE(n >= 1, "empty scan");
    excl_scan_scratch_impl<uint32_t>(s, in, out, n, total_dev, scratch, scratch_words);
}

void exclusive_scan_u32(State& st, const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* total_dev) {
    excl_scan_impl<uint32_t>(st, in, out, n, total_dev);
}
void exclusive_scan_u64(State& st, const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* total_dev) {
    excl_scan_impl<uint64_t>(st, in, out, n, total_dev);
}

// ------------------------------------------------------------------ relayout
struct FieldMoves {
    uint32_t nf;
    sparq_field_t from[64];
    sparq_field_t to[64];
};

// kRelayoutItems entries per lane (striped by the block size), all loads
// issued before the first store: a key word is 8 bytes, so one entry per lane
// keeps too few bytes in flight to cover the memory latency.
constexpr int kRelayoutItems = 4;
template <int WI, int WO>
__global__ void __launch_bounds__(kBlock) relayout_kernel(const uint64_t* __restrict__ kin, uint64_t cap_in, uint64_t* __restrict__ kout,
                                                     uint64_t cap_out, uint64_t n, FieldMoves mv) {
    const uint64_t base = (uint64_t)blockIdx.x * (kBlock * kRelayoutItems) + threadIdx.x;
    uint64_t a[kRelayoutItems][WI];
#pragma unroll
    for (int it = 0; it < kRelayoutItems; ++it) {
        const uint64_t i = base + (uint64_t)it * kBlock;
        load_key<WI>(kin, cap_in, i < n ? i : n - 1, a[it]);   // clamped: in bounds, used only when i < n
    }
#pragma unroll
    for (int it = 0; it < kRelayoutItems; ++it) {
        const uint64_t i = base + (uint64_t)it * kBlock;
        if (i >= n) break;
        uint64_t b[WO];
#pragma unroll
        for (int w = 0; w < WO; ++w) b[w] = 0;
        for (uint32_t f = 0; f < mv.nf; ++f) fput<WO>(b, mv.to[f], fget<WI>(a[it], mv.from[f]));
        store_key<WO>(kout, cap_out, i, b);
    }
}

template <int WI>
static void launch_relayout_in(State& st, uint32_t WO, uint64_t* kout, uint64_t cap_out, const FieldMoves& mv) {
    unsigned g = grid_for(st.n, kBlock * kRelayoutItems);
#define SPQ_RL(WOV)                                                                                          \
    case WOV: {                                                                                              \
        SPQ_PROF(st.stream, "relayout_kernel", st.n * 8.0 * (WI + WOV));                                     \
        relayout_kernel<WI, WOV><<<g, kBlock, 0, st.stream>>>(st.keys, st.cap, kout, cap_out, st.n, mv);     \
    } break;
    switch (WO) {
        SPQ_RL(1) SPQ_RL(2) SPQ_RL(3) SPQ_RL(4) SPQ_RL(5) SPQ_RL(6) SPQ_RL(7) SPQ_RL(8)
        default: throw Error(SPARQ_INVALID, "key width out of range");
    }
#undef SPQ_RL
    SPQ_HIP(hipGetLastError());
}

void relayout(State& st, uint32_t new_words, const sparq_field_t* from, const sparq_field_t* to, uint32_t nf) {
    SPQ_REQUIRE(new_words >= 1 && new_words <= SPARQ_MAX_WORDS, "key width out of range");
    SPQ_REQUIRE(nf <= 64, "too many fields");
    FieldMoves mv{};
    mv.nf = nf;
    for (uint32_t f = 0; f < nf; ++f) {
        SPQ_REQUIRE(from[f].word < st.W && to[f].word < new_words, "field word out of range");
        SPQ_REQUIRE(from[f].width == to[f].width && from[f].width >= 1 && from[f].width <= 64, "field width mismatch");
        SPQ_REQUIRE((uint32_t)from[f].shift + from[f].width <= 64 && (uint32_t)to[f].shift + to[f].width <= 64, "field crosses a word");
        mv.from[f] = from[f];
        mv.to[f] = to[f];
    }
    uint64_t* kout = dalloc<uint64_t>((size_t)new_words * std::max<uint64_t>(st.cap, 1), st.stream);
    if (st.n) SPQ_DISPATCH_W(st.W, launch_relayout_in, st, new_words, kout, st.cap, mv);
    dfree(st.keys, st.stream);
    st.keys = kout;
    st.W = new_words;
    st.Wcap = 0;
}

// ------------------------------------------------------------------ basis state
__global__ void basis_kernel(uint64_t* __restrict__ keys, double2* __restrict__ amps, uint64_t cap, uint32_t W, BasisKey k) {
    if (threadIdx.x < W) keys[(uint64_t)threadIdx.x * cap] = k.w[threadIdx.x];
    if (threadIdx.x == 0) amps[0] = make_double2(1.0, 0.0);
}

void write_basis(State& st, const BasisKey& k) {
    {
        SPQ_PROF(st.stream, "basis_kernel", 8.0 * st.W + 16);
        basis_kernel<<<1, 64, 0, st.stream>>>(st.keys, st.amps, st.cap, st.W, k);
    }
    SPQ_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ reductions
__global__ void __launch_bounds__(kBlock) sum_abs2_kernel(const double2* __restrict__ amps, uint64_t n, double* __restrict__ partial) {
    __shared__ double lds[kBlock / 64];
    double s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) s += cabs2(amps[i]);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
        partial[blockIdx.x] = t;
    }
}

double sum_abs2(State& st) {
    if (st.n == 0) return 0.0;
    unsigned g = std::min<unsigned>(grid_for(st.n, kBlock), 1024);
    double* partial = dalloc<double>(g, st.stream);
    {
        SPQ_PROF(st.stream, "sum_abs2_kernel", st.n * 16.0);
        sum_abs2_kernel<<<g, kBlock, 0, st.stream>>>(st.amps, st.n, partial);
    }
    SPQ_HIP(hipGetLastError());
    std::vector<double> h(g);
    read_back(st, partial, g * sizeof(double), h.data());
    dfree(partial, st.stream);
    double s = 0;
    for (double v : h) s += v;
    return s;
}

// Sum of |a|^2 over each tile of `tile` consecutive entries (in entry order):
// the coarse level of sampling an entry with probability |a|^2 (MeasureZ).
__global__ void __launch_bounds__(kBlock) tile_abs2_kernel(const double2* __restrict__ amps, uint64_t n, uint64_t tile,
                                                     double* __restrict__ out) {
    __shared__ double lds[kBlock / 64];
    const uint64_t b = (uint64_t)blockIdx.x * tile, e = min(n, b + tile);
    double s = 0;
    for (uint64_t i = b + threadIdx.x; i < e; i += kBlock) s += cabs2(amps[i]);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
        out[blockIdx.x] = t;
    }
}

void tile_abs2(State& st, uint64_t tile, double* host_out, uint64_t ntiles) {
    SPQ_REQUIRE(tile >= 1 && ntiles == (st.n + tile - 1) / tile, "tile count must be ceil(n / tile)");
    if (ntiles == 0) return;
    SPQ_REQUIRE(ntiles <= (1u << 31), "too many tiles");
    double* out = dalloc<double>(ntiles, st.stream);
    {
        SPQ_PROF(st.stream, "tile_abs2_kernel", st.n * 16.0);
        tile_abs2_kernel<<<(unsigned)ntiles, kBlock, 0, st.stream>>>(st.amps, st.n, tile, out);
    }
    SPQ_HIP(hipGetLastError());
    read_back(st, out, ntiles * sizeof(double), host_out);
    dfree(out, st.stream);
}

// Sum of |a|^2 over the entries satisfying every condition (Probability).
template <int W>
__global__ void __launch_bounds__(kBlock) cond_abs2_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps, uint64_t cap,
                                                      uint64_t n, const CondList cl, double* __restrict__ partial) {
    __shared__ double lds[kBlock / 64];
    double s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t k[W];
        load_key<W>(keys, cap, i, k);
        bool ok = true;
        for (uint32_t c = 0; c < cl.n; ++c) ok = ok && cond_ok<W>(k, cl.c[c]);
        if (ok) s += cabs2(amps[i]);
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
        partial[blockIdx.x] = t;
    }
}

template <int W>
static void launch_cond_abs2(State& st, const CondList& cl, double* partial, unsigned g) {
    {
        SPQ_PROF(st.stream, "cond_abs2_kernel", st.n * (8.0 * W + 16));
        cond_abs2_kernel<W><<<g, kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, st.n, cl, partial);
    }
    SPQ_HIP(hipGetLastError());
}

double cond_abs2(State& st, const sparq_cond_t* conds, uint32_t ncond) {
    SPQ_REQUIRE(ncond <= SPARQ_MAX_PROGRAM_CONDS, "too many conditions");
    CondList cl{};
    cl.n = ncond;
    for (uint32_t c = 0; c < ncond; ++c) {
        SPQ_REQUIRE(conds[c].field.word < st.W, "condition field out of range");
        cl.c[c] = conds[c];
    }
    if (st.n == 0) return 0.0;
    const unsigned g = std::min<unsigned>(grid_for(st.n, kBlock), 1024);
    double* partial = dalloc<double>(g, st.stream);
    SPQ_DISPATCH_W(st.W, launch_cond_abs2, st, cl, partial, g);
    std::vector<double> h(g);
    read_back(st, partial, g * sizeof(double), h.data());
    dfree(partial, st.stream);
    double s = 0;
    for (double v : h) s += v;
    return s;
}

// ------------------------------------------------------------------ dense post-selection
// Entries whose `zero` fields are all 0 and whose index (the concatenation of
// the `index` fields, field 0 least significant) is below len are written to
// out[index]; per-block partial sums of |a|^2 of the selected entries.
struct FieldList {
    uint32_t n;
    sparq_field_t f[8];
};

template <int W>
__global__ void __launch_bounds__(kBlock) post_select_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps, uint64_t cap,
                                                        uint64_t n, const FieldList zero, const FieldList index, double2* __restrict__ out,
                                                        uint64_t len, double* __restrict__ partial) {
    __shared__ double lds[kBlock / 64];
    double s = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t k[W];
        load_key<W>(keys, cap, i, k);
        bool sel = true;
        for (uint32_t z = 0; z < zero.n; ++z) sel = sel && fget<W>(k, zero.f[z]) == 0;
        if (!sel) continue;
        uint64_t idx = 0;
        uint32_t at = 0;
        for (uint32_t f = 0; f < index.n; ++f) {
            idx |= fget<W>(k, index.f[f]) << at;
            at += index.f[f].width;
        }
        if (idx >= len) continue;   // outside the requested range: not selected
        const double2 a = amps[i];
        out[idx] = a;
        s += cabs2(a);
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
        partial[blockIdx.x] = t;
    }
}

template <int W>
static void launch_post_select(State& st, const FieldList& zero, const FieldList& index, double2* out, uint64_t len, double* partial,
                               unsigned g) {
    {
        SPQ_PROF(st.stream, "post_select_kernel", st.n * (8.0 * W + 16));
        post_select_kernel<W><<<g, kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, st.n, zero, index, out, len, partial);
    }
    SPQ_HIP(hipGetLastError());
}

double post_select_dense(State& st, const sparq_field_t* zero, uint32_t nzero, const sparq_field_t* index, uint32_t nindex,
                         double* out_host, uint64_t len) {
    SPQ_REQUIRE(nzero <= 8 && nindex >= 1 && nindex <= 8, "at most 8 zero fields and 1..8 index fields");
    FieldList zl{}, il{};
    uint32_t bits = 0;
    zl.n = nzero;
    il.n = nindex;
    for (uint32_t i = 0; i < nzero; ++i) {
        SPQ_REQUIRE(zero[i].word < st.W, "field out of range");
        zl.f[i] = zero[i];
    }
    for (uint32_t i = 0; i < nindex; ++i) {
        SPQ_REQUIRE(index[i].word < st.W, "field out of range");
        il.f[i] = index[i];
        bits += index[i].width;
    }
    SPQ_REQUIRE(bits <= 64 && len >= 1 && len <= (uint64_t(1) << 40) && (bits == 64 || len <= (uint64_t(1) << bits)),
                "output length must be in [1, min(2^(total index bits), 2^40)]");
    double2* out = dalloc<double2>(len, st.stream);
    SPQ_HIP(hipMemsetAsync(out, 0, len * sizeof(double2), st.stream));
    const unsigned g = st.n ? std::min<unsigned>(grid_for(st.n, kBlock), 1024) : 1;
    double* partial = dalloc<double>(g, st.stream);
    SPQ_HIP(hipMemsetAsync(partial, 0, g * sizeof(double), st.stream));
    if (st.n) SPQ_DISPATCH_W(st.W, launch_post_select, st, zl, il, out, len, partial, g);
    std::vector<double> h(g);
    SPQ_HIP(hipMemcpyAsync(out_host, out, len * sizeof(double2), hipMemcpyDeviceToHost, st.stream));
    read_back(st, partial, g * sizeof(double), h.data());
    dfree(out, st.stream);
    dfree(partial, st.stream);
    double s = 0;
    for (double v : h) s += v;
    return s;
}

__global__ void __launch_bounds__(kBlock) scale_kernel(double2* __restrict__ amps, uint64_t n, double s) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) amps[i] = cscale(amps[i], s);
}

void scale_amps(State& st, double s) {
    if (!st.n) return;
    {
        SPQ_PROF(st.stream, "scale_kernel", st.n * 32.0);
        scale_kernel<<<grid_for(st.n, kBlock), kBlock, 0, st.stream>>>(st.amps, st.n, s);
    }
    SPQ_HIP(hipGetLastError());
}

// Min and max of one field over all entries: reads only the field's key word
// (8 bytes per entry), reduces in registers and across the waves of a block,
// then one atomic pair per block.
__global__ void __launch_bounds__(kBlock) field_minmax_kernel(const uint64_t* __restrict__ word, uint64_t n, sparq_field_t f,
                                                         unsigned long long* mm) {
    __shared__ unsigned long long lmin[kBlock / 64], lmax[kBlock / 64];
    const unsigned long long mask = fmask(f.width);
    unsigned long long vmin = ~0ull, vmax = 0ull;
    // kMinmaxItems striped loads per lane per tile, issued together; indices
    // past n read entry n-1 again, which changes neither minimum nor maximum.
    constexpr int kMinmaxItems = 8;
    const uint64_t tile = (uint64_t)kBlock * kMinmaxItems;
    for (uint64_t t = (uint64_t)blockIdx.x * tile; t < n; t += (uint64_t)gridDim.x * tile) {
        unsigned long long v[kMinmaxItems];
#pragma unroll
        for (int j = 0; j < kMinmaxItems; ++j) {
            const uint64_t i = t + (uint64_t)j * kBlock + threadIdx.x;
            v[j] = word[i < n ? i : n - 1];
        }
#pragma unroll
        for (int j = 0; j < kMinmaxItems; ++j) {
            const unsigned long long x = (v[j] >> f.shift) & mask;
            vmin = x < vmin ? x : vmin;
            vmax = x > vmax ? x : vmax;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long a = __shfl_xor(vmin, off, 64), b = __shfl_xor(vmax, off, 64);
        vmin = a < vmin ? a : vmin;
        vmax = b > vmax ? b : vmax;
    }
    if ((threadIdx.x & 63) == 0) {
        lmin[threadIdx.x >> 6] = vmin;
        lmax[threadIdx.x >> 6] = vmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) {
            vmin = lmin[w] < vmin ? lmin[w] : vmin;
            vmax = lmax[w] > vmax ? lmax[w] : vmax;
        }
        atomicMin(&mm[0], vmin);
        atomicMax(&mm[1], vmax);
    }
}

bool field_uniform(State& st, sparq_field_t f, uint64_t* value) {
    if (st.n == 0) {
        *value = 0;
        return true;
    }
    unsigned long long* mm = dalloc<unsigned long long>(2, st.stream);
    unsigned long long init[2] = {~0ull, 0ull};
    SPQ_HIP(hipMemcpyAsync(mm, init, sizeof(init), hipMemcpyHostToDevice, st.stream));
    {
        SPQ_PROF(st.stream, "field_minmax_kernel", st.n * 8.0);
        const unsigned g = std::min<unsigned>(grid_for(st.n, kBlock * 8), 2048);
        field_minmax_kernel<<<g, kBlock, 0, st.stream>>>(st.keys + (uint64_t)f.word * st.cap, st.n, f, mm);
    }
    SPQ_HIP(hipGetLastError());
    unsigned long long h[2];
    read_back(st, mm, sizeof(h), h);
    dfree(mm, st.stream);
    *value = h[0];
    return h[0] == h[1];
}

}  // namespace sparq
