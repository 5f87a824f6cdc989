// Stream compaction: ClearZero, condition partition, and the duplicate merge
// of a sorted state.  Each is count -> scan of tile counts -> ordered scatter.
// ClearZero bytes per entry: (8W+16) read + (8W+16) per kept entry written
// (+16 for the amplitude re-read of the counting pass).
#include <algorithm>
#include <mutex>

#include "blockscan.hpp"
#include "engine.hpp"
#include "keyops.hpp"

namespace sparq {

constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;   // 2048 entries per workgroup, blocked per thread

enum Mode { kClearZero = 0, kPartition = 1, kMerge = 2, kSelect = 3 };

// Predicate of entry i for the given mode; for kMerge also returns the run sum.
template <int W, int MODE>
__device__ __forceinline__ bool keep_entry(const uint64_t* __restrict__ keys, const double2* __restrict__ amps, uint64_t cap,
                                           uint64_t n, uint64_t i, double eps, const CondList& cl, double2& sum) {
    if (MODE == kClearZero) {
        sum = amps[i];
        return !(cabs2(sum) < eps);
    } else if (MODE == kPartition || MODE == kSelect) {
        uint64_t k[W];
        load_key<W>(keys, cap, i, k);
        bool ok = true;
        for (uint32_t c = 0; c < cl.n; ++c) ok = ok && cond_ok<W>(k, cl.c[c]);
        if (MODE == kSelect && ok) sum = amps[i];
        return ok;
    } else {
        uint64_t k[W];
        load_key<W>(keys, cap, i, k);
        if (i > 0) {
            bool same = true;
#pragma unroll
            for (int w = 0; w < W; ++w) same = same && (keys[(uint64_t)w * cap + i - 1] == k[w]);
            if (same) return false;   // not the head of its run
        }
        double2 s = amps[i];
        for (uint64_t j = i + 1; j < n; ++j) {
            bool same = true;
#pragma unroll
            for (int w = 0; w < W; ++w) same = same && (keys[(uint64_t)w * cap + j] == k[w]);
            if (!same) break;
            s = cadd(s, amps[j]);
        }
        sum = s;
        return !(cabs2(s) < eps);
    }
}

// Tiles are striped: item j of thread t is entry tile + j*kBlock + t, so every
// load and store instruction of a wave touches 64 consecutive entries, and the
// kept entries of one row land in one contiguous output run.
template <int W, int MODE>
__global__ void __launch_bounds__(kBlock) compact_count_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps, uint64_t cap,
                                                          uint64_t n, const uint64_t* __restrict__ n_dev, double eps, const CondList cl,
                                                          uint32_t* __restrict__ counts) {
    __shared__ uint32_t lds[kBlock / 64];
    if (n_dev) n = min(n, n_dev[0]);   // entry count known only on the device
    const uint64_t tile = (uint64_t)blockIdx.x * kTile;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        uint64_t i = tile + (uint64_t)j * kBlock + threadIdx.x;
        double2 s;
        if (i < n && keep_entry<W, MODE>(keys, amps, cap, n, i, eps, cl, s)) ++c;
    }
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < kBlock / 64; ++w) tot += lds[w];
        counts[blockIdx.x] = tot;
    }
}

// Kept entries go to kout/aout at offs[tile] + rank in entry order; for
// kPartition the others go after all kept ones (at n_kept + rejects before them).
template <int W, int MODE>
__global__ void __launch_bounds__(kBlock) compact_scatter_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps,
                                                            uint64_t cap, uint64_t n, const uint64_t* __restrict__ n_dev, double eps,
                                                            const CondList cl,
                                                            const uint64_t* __restrict__ offs, const uint64_t* __restrict__ total,
                                                            uint64_t* __restrict__ kout, double2* __restrict__ aout, uint64_t cap_out) {
    constexpr int kWaves = kBlock / 64;
    __shared__ uint32_t row_wave[kItems * kWaves];   // kept count per (row, wave), then its exclusive prefix
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t tile = (uint64_t)blockIdx.x * kTile;
    if (n_dev) n = min(n, n_dev[0]);
    uint32_t flags = 0;
    uint32_t lane_pre[kItems];
    double2 sums[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const uint64_t i = tile + (uint64_t)j * kBlock + threadIdx.x;
        sums[j] = make_double2(0.0, 0.0);
        const bool k = i < n && keep_entry<W, MODE>(keys, amps, cap, n, i, eps, cl, sums[j]);
        const uint64_t m = __ballot(k);
        flags |= uint32_t(k) << j;
        lane_pre[j] = (uint32_t)__popcll(m & lanemask_lt());
        if (lane == 0) row_wave[j * kWaves + wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < 64) {   // exclusive scan of the 32 (row, wave) counts in entry order
        uint32_t v = threadIdx.x < kItems * kWaves ? row_wave[threadIdx.x] : 0;
        uint32_t inc = wave_incl_scan(v);
        if (threadIdx.x < kItems * kWaves) row_wave[threadIdx.x] = inc - v;
    }
    __syncthreads();
    const uint64_t kept_before_tile = offs[blockIdx.x];
    const uint64_t rej_base = (MODE == kPartition) ? total[0] + (tile - kept_before_tile) : 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const uint64_t i = tile + (uint64_t)j * kBlock + threadIdx.x;
        if (i >= n) break;
        const uint32_t kept_rank = row_wave[j * kWaves + wave] + lane_pre[j];   // kept entries before i in the tile
        uint64_t dst;
        if ((flags >> j) & 1) dst = kept_before_tile + kept_rank;
        else if (MODE == kPartition) dst = rej_base + ((uint64_t)j * kBlock + threadIdx.x - kept_rank);
        else continue;
#pragma unroll
        for (int w = 0; w < W; ++w) kout[(uint64_t)w * cap_out + dst] = keys[(uint64_t)w * cap + i];
        aout[dst] = (MODE == kPartition) ? amps[i] : sums[j];
    }
}

template <int W, int MODE>
static uint64_t run_compact(State& st, double eps, const CondList& cl) {
    const uint64_t n = st.n;
    const unsigned nb = grid_for(n, kTile);
    uint32_t* counts = dalloc<uint32_t>(nb, st.stream);
    uint64_t* offs = dalloc<uint64_t>(nb + 1, st.stream);
    {
        SPQ_PROF(st.stream, "compact_count_kernel", n * (MODE == kClearZero ? 16.0 : MODE == kMerge ? 8.0 * W + 16 : 8.0 * W));
        compact_count_kernel<W, MODE><<<nb, kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, n, nullptr, eps, cl, counts);
    }
    SPQ_HIP(hipGetLastError());
    exclusive_scan_u32(st, counts, offs, nb, offs + nb);
    uint64_t kept = read_u64(st, offs + nb);
    uint64_t out_n = (MODE == kPartition) ? n : kept;
    if (MODE != kPartition && kept == n) {   // nothing removed
        dfree(counts, st.stream);
        dfree(offs, st.stream);
        return kept;
    }
    uint64_t cap_out = std::max<uint64_t>(out_n, 1);
    uint64_t* kout;
    double2* aout;
    alloc_arrays(st, st.W, cap_out, &kout, &aout);
    {
        SPQ_PROF(st.stream, "compact_scatter_kernel", n * (MODE == kClearZero ? 16.0 : MODE == kMerge ? 8.0 * W + 16 : 8.0 * W) + (double)out_n * (8.0 * W + 16));
        compact_scatter_kernel<W, MODE><<<nb, kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, n, nullptr, eps, cl, offs, offs + nb, kout,
                                                                   aout, cap_out);
    }
    SPQ_HIP(hipGetLastError());
    install_arrays(st, st.W, kout, aout, out_n, cap_out);
    dfree(counts, st.stream);
    dfree(offs, st.stream);
    return kept;
}

template <int W>
static void clear_zero_w(State& st, double eps, uint64_t* kept) {
    *kept = run_compact<W, kClearZero>(st, eps, CondList{});
}
template <int W>
static void partition_w(State& st, const CondList& cl, uint64_t* kept) {
    *kept = run_compact<W, kPartition>(st, 0.0, cl);
}
template <int W>
static void select_w(State& st, const CondList& cl, uint64_t* kept) {
    *kept = run_compact<W, kSelect>(st, 0.0, cl);
}
template <int W>
static void merge_w(State& st, double eps, uint64_t* kept) {
    *kept = run_compact<W, kMerge>(st, eps, CondList{});
}

void clear_zero(State& st, double eps) {
    if (st.n == 0) return;
    uint64_t kept;
    SPQ_DISPATCH_W(st.W, clear_zero_w, st, eps, &kept);
}

// ClearZero of entries [0, counts[2]) of (keys, amps, cap), with the entry count
// known only on the device.  counts[3] receives the kept count and counts[0..4)
// is read back into info in the same transfer.  The result is (keys, amps)
// itself when nothing is pruned, otherwise freshly allocated arrays.
template <int W>
static void clear_zero_sized_w(State& st, uint64_t* keys, double2* amps, uint64_t cap, uint64_t* counts, double eps, uint64_t* info,
                               uint64_t** kout, double2** aout) {
    const unsigned nb = grid_for(cap, kTile);
    uint32_t* tile_counts = dalloc<uint32_t>(nb, st.stream);
    uint64_t* offs = dalloc<uint64_t>(nb + 1, st.stream);
    {
        SPQ_PROF(st.stream, "compact_count_kernel", cap * 16.0);
        compact_count_kernel<W, kClearZero><<<nb, kBlock, 0, st.stream>>>(keys, amps, cap, cap, counts + 2, eps, CondList{}, tile_counts);
    }
    SPQ_HIP(hipGetLastError());
    exclusive_scan_u32(st, tile_counts, offs, nb, offs + nb);
    SPQ_HIP(hipMemcpyAsync(counts + 3, offs + nb, sizeof(uint64_t), hipMemcpyDeviceToDevice, st.stream));
    read_back(st, counts, 4 * sizeof(uint64_t), info);
    const uint64_t valid = info[2], kept = info[3];
    *kout = keys;
    *aout = amps;
    if (kept != valid) {
        uint64_t* kk;
        double2* aa;
        alloc_arrays(st, st.W, std::max<uint64_t>(kept, 1), &kk, &aa);
        {
            SPQ_PROF(st.stream, "compact_scatter_kernel", valid * 16.0 + (double)kept * (8.0 * W + 16));
            compact_scatter_kernel<W, kClearZero><<<nb, kBlock, 0, st.stream>>>(keys, amps, cap, cap, counts + 2, eps, CondList{}, offs,
                                                                              offs + nb, kk, aa, std::max<uint64_t>(kept, 1));
        }
        SPQ_HIP(hipGetLastError());
        *kout = kk;
        *aout = aa;
    }
    dfree(tile_counts, st.stream);
    dfree(offs, st.stream);
}

void clear_zero_sized(State& st, uint64_t* keys, double2* amps, uint64_t cap, uint64_t* counts, double eps, uint64_t info[4],
                      uint64_t** kout, double2** aout) {
    SPQ_DISPATCH_W(st.W, clear_zero_sized_w, st, keys, amps, cap, counts, eps, info, kout, aout);
}

void merge_sorted_duplicates(State& st, double eps) {
    if (st.n == 0) return;
    uint64_t kept;
    SPQ_DISPATCH_W(st.W, merge_w, st, eps, &kept);
}

uint64_t partition_by_conditions(State& st, const sparq_cond_t* conds, uint32_t ncond) {
    if (ncond == 0) return st.n;
    SPQ_REQUIRE(ncond <= SPARQ_MAX_PROGRAM_CONDS, "too many conditions");
    CondList cl{};
    cl.n = ncond;
    for (uint32_t c = 0; c < ncond; ++c) {
        SPQ_REQUIRE(conds[c].field.word < st.W, "condition field out of range");
        cl.c[c] = conds[c];
    }
    if (st.n == 0) return 0;
    uint64_t kept;
    SPQ_DISPATCH_W(st.W, partition_w, st, cl, &kept);
    return kept;
}

uint64_t select_by_conditions(State& st, const sparq_cond_t* conds, uint32_t ncond) {
    SPQ_REQUIRE(ncond <= SPARQ_MAX_PROGRAM_CONDS, "too many conditions");
    CondList cl{};
    cl.n = ncond;
    for (uint32_t c = 0; c < ncond; ++c) {
        SPQ_REQUIRE(conds[c].field.word < st.W, "condition field out of range");
        cl.c[c] = conds[c];
    }
    if (st.n == 0 || ncond == 0) return st.n;
    uint64_t kept;
    SPQ_DISPATCH_W(st.W, select_w, st, cl, &kept);
    return kept;
}

// ---- duplicate detection on a sorted state
template <int W>
__global__ void __launch_bounds__(kBlock) dup_kernel(const uint64_t* __restrict__ keys, uint64_t cap, uint64_t n, unsigned int* flag) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x + 1;
    if (i >= n) return;
    bool same = true;
#pragma unroll
    for (int w = 0; w < W; ++w) same = same && (keys[(uint64_t)w * cap + i] == keys[(uint64_t)w * cap + i - 1]);
    if (same) atomicOr(flag, 1u);
}

template <int W>
static void launch_dup(State& st, unsigned int* flag) {
    {
        SPQ_PROF(st.stream, "dup_kernel", st.n * 8.0 * W);
        dup_kernel<W><<<grid_for(st.n, kBlock), kBlock, 0, st.stream>>>(st.keys, st.cap, st.n, flag);
    }
    SPQ_HIP(hipGetLastError());
}

bool has_sorted_duplicates(State& st) {
    if (st.n <= 1) return false;
    unsigned int* flag = dalloc<unsigned int>(1, st.stream);
    SPQ_HIP(hipMemsetAsync(flag, 0, sizeof(unsigned int), st.stream));
    SPQ_DISPATCH_W(st.W, launch_dup, st, flag);
    unsigned int h = 0;
    read_back(st, flag, sizeof(h), &h);
    dfree(flag, st.stream);
    return h != 0;
}

}  // namespace sparq

// ---------------------------------------------------------------- ClearZero without allocation
// sparq_clear_zero_noalloc: the ClearZero compaction of sparq_clear_zero over
// caller-owned arrays, for callers that cannot allocate.  Its per-tile counts,
// their offsets and the scan's partial sums live in the file-scope array
// `scratch` below (one copy per GPU, held until the process ends); host code
// reaches it only through the address hipGetSymbolAddress returns on each
// device.  Nothing here allocates, synchronizes or reads back.
namespace sparq {
namespace {
constexpr uint64_t kNoAllocMaxEntries = uint64_t(1) << 28;
constexpr uint64_t kNoAllocTiles = kNoAllocMaxEntries / kTile;                      // 131072
constexpr uint64_t kNoAllocCountBytes = kNoAllocTiles * sizeof(uint32_t);           // 512 KiB
constexpr uint64_t kNoAllocOffsWords = kNoAllocTiles + 1;
constexpr uint64_t kNoAllocScanWords = 2 * ((kNoAllocTiles + 2047) / 2048) + 64;   // partial sums above level 0
constexpr uint64_t kNoAllocScratchBytes = kNoAllocCountBytes + 8 * (kNoAllocOffsWords + kNoAllocScanWords);
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};
}  // namespace
}  // namespace sparq

static __device__ __align__(256) unsigned char scratch[sparq::kNoAllocScratchBytes];

namespace sparq {
namespace {
// Device address of `scratch` on the current device, resolved once per device.
uint8_t* scratch_on_current_device() {
    int dev = 0;
    SPQ_HIP(hipGetDevice(&dev));
    SPQ_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SPQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SPQ_REQUIRE(p != nullptr, "no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<uint8_t*>(p);
    }
    return g_scratch_ptr[dev];
}

__global__ void store_zero_kernel(uint64_t* __restrict__ out) { out[0] = 0; }

template <int W>
void clear_zero_noalloc_w(hipStream_t s, const uint64_t* keys, const double2* amps, uint64_t n, uint64_t cap_in, double eps,
                          uint64_t* kout, double2* aout, uint64_t cap_out, uint64_t* kept_dev, uint8_t* base) {
    const unsigned nb = grid_for(n, kTile);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base);
    uint64_t* offs = reinterpret_cast<uint64_t*>(base + kNoAllocCountBytes);
    uint64_t* scan_scratch = offs + kNoAllocOffsWords;
    compact_count_kernel<W, kClearZero><<<nb, kBlock, 0, s>>>(keys, amps, cap_in, n, nullptr, eps, CondList{}, counts);
    SPQ_HIP(hipGetLastError());
    // the scan writes every offset and the kept total: nothing in scratch needs clearing
    exclusive_scan_u32_scratch(s, counts, offs, nb, kept_dev, scan_scratch, kNoAllocScanWords);
    compact_scatter_kernel<W, kClearZero><<<nb, kBlock, 0, s>>>(keys, amps, cap_in, n, nullptr, eps, CondList{}, offs, kept_dev, kout,
                                                             aout, cap_out);
    SPQ_HIP(hipGetLastError());
}
}  // namespace
}  // namespace sparq

extern "C" int sparq_clear_zero_noalloc(const uint64_t* keys_in, const double* amps_in, uint32_t words, uint64_t n, uint64_t cap_in,
                                        double eps, uint64_t* keys_out, double* amps_out, uint64_t cap_out, uint64_t* kept_dev,
                                        void* stream) {
    using namespace sparq;
    return guarded_call([&] {
        SPQ_REQUIRE(words >= 1 && words <= SPARQ_MAX_WORDS, "key width out of range");
        SPQ_REQUIRE(n <= kNoAllocMaxEntries, "sparq_clear_zero_noalloc: more than 2^28 entries (the scratch limit)");
        SPQ_REQUIRE(kept_dev != nullptr, "null kept-count output");
        SPQ_REQUIRE(n == 0 || (keys_in && amps_in && keys_out && amps_out), "null array");
        SPQ_REQUIRE(cap_in >= n && cap_out >= n, "array capacities must hold n entries");
        SPQ_REQUIRE((reinterpret_cast<uintptr_t>(amps_in) | reinterpret_cast<uintptr_t>(amps_out)) % 16 == 0,
                    "amplitude arrays must be 16-byte aligned");
        ensure_device();
        uint8_t* base = scratch_on_current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (n == 0) {
            store_zero_kernel<<<1, 1, 0, s>>>(kept_dev);
            SPQ_HIP(hipGetLastError());
            return;
        }
        const double2* a = reinterpret_cast<const double2*>(amps_in);
        double2* ao = reinterpret_cast<double2*>(amps_out);
        SPQ_DISPATCH_W(words, clear_zero_noalloc_w, s, keys_in, a, n, cap_in, eps, keys_out, ao, cap_out, kept_dev, base);
    });
}
