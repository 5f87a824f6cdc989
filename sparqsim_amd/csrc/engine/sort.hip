// LSD radix sort of the state by packed key (32-bit index payload), with
// per-wave ballot ranking and an LDS-staged, digit-grouped write-out.
//
// Per digit pass (N entries): histogram reads the 8-byte key word; the scatter
// reads key word + index and writes both (2*(8+4) bytes); the final gather
// moves W key words and the 16-byte amplitude.  Only bits that vary across
// the state are sorted (an OR/AND reduction finds them first).
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <unordered_map>

#include "blockscan.hpp"
#include "engine.hpp"
#include "keyops.hpp"

namespace sparq {

constexpr int kSortTile = 4096;   // entries per workgroup of a digit pass
// Digit-pass workgroup sizes: 8 waves of 8 entries per lane for 8-bit digits;
// 4 waves of 16 for 11-bit digits (whose per-wave digit counters fill the LDS).
constexpr int kSortBlock8 = 512;
constexpr int kSortBlock11 = 256;

// Lanes of the wave holding the same value as the calling lane, for values of
// `nbits` digit bits plus the invalid marker at bit RB.
template <int RB>
__device__ __forceinline__ uint64_t match_any(uint32_t d, int nbits) {
    uint64_t m = ~0ull;
    for (int b = 0; b < nbits; ++b) {
        uint64_t bal = __ballot((d >> b) & 1);
        m &= ((d >> b) & 1) ? bal : ~bal;
    }
    uint64_t bal = __ballot((d >> RB) & 1);
    m &= ((d >> RB) & 1) ? bal : ~bal;
    return m;
}

// ---------------------------------------------------------------- key statistics
// One read pass over the keys: bitwise OR and AND per word (bits that vary
// across the state), whether the state is already ordered by
// (all bits except `last`, then `last`), and for every adjacent pair out of
// that order the most significant bit where the pair differs (one mask per
// key word): the bits above the most significant such bit are a prefix the
// state is already sorted by.  Per-block partial records are combined by a
// one-block second kernel.
struct StatsOrder {
    uint64_t others[SPARQ_MAX_WORDS];   // per-word mask of the primary comparison bits
    sparq_field_t last;                 // width 0: none
};

template <int W>
__device__ __forceinline__ bool out_of_order(const uint64_t (&p)[W], const uint64_t (&c)[W], const StatsOrder& so) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
        uint64_t a = p[w] & so.others[w], b = c[w] & so.others[w];
        if (a != b) return a > b;
    }
    if (so.last.width) return fget<W>(p, so.last) > fget<W>(c, so.last);
    return false;
}

// The first bit (in the order of StatsOrder) where two keys differ, as a
// one-bit mask in its word's slot of m: the most significant differing bit of
// the first word that differs outside `last`; nothing when they differ only in
// `last`.
template <int W>
__device__ __forceinline__ void first_difference(const uint64_t (&p)[W], const uint64_t (&c)[W], const StatsOrder& so, uint64_t (&m)[W]) {
    bool found = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint64_t x = (p[w] ^ c[w]) & so.others[w];
        if (!found && x) m[w] |= 1ull << (63 - __clzll(x));
        found = found || x != 0;
    }
}

// Record layout: stats_words(W) (engine.hpp); the histogram of `hist` is zero
// without one.
template <int W>
__device__ __forceinline__ uint64_t stats_combine(int j, uint64_t x, uint64_t y) {
    if (j < W) return x | y;
    if (j < 2 * W) return x & y;
    if (j < 3 * W + 2) return x | y;
    return x + y;
}

template <int W>
__global__ void __launch_bounds__(kBlock) key_stats_kernel(const uint64_t* __restrict__ keys, uint64_t cap, uint64_t n, StatsOrder so,
                                                      sparq_field_t hist, uint64_t* __restrict__ partial) {
    constexpr int R = stats_words(W);
    constexpr int kI = W <= 2 ? 8 : (W <= 4 ? 4 : 2);   // entries per thread per tile, loads issued together
    __shared__ uint64_t lds[kBlock / 64][R];
    uint64_t r[R];
#pragma unroll
    for (int j = 0; j < R; ++j) r[j] = (j >= W && j < 2 * W) ? ~0ull : 0;
    const uint64_t ntiles = (n + (uint64_t)kBlock * kI - 1) / ((uint64_t)kBlock * kI);
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint64_t k[kI][W], p[kI][W];
#pragma unroll
        for (int j = 0; j < kI; ++j) {
            const uint64_t i = t * kBlock * kI + (uint64_t)j * kBlock + threadIdx.x;
            const uint64_t ic = i < n ? i : n - 1;            // clamped: duplicates do not change OR/AND/order
            const uint64_t ip = ic > 0 ? ic - 1 : 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                k[j][w] = keys[(uint64_t)w * cap + ic];
                p[j][w] = keys[(uint64_t)w * cap + ip];
            }
        }
#pragma unroll
        for (int j = 0; j < kI; ++j) {
            const uint64_t i = t * kBlock * kI + (uint64_t)j * kBlock + threadIdx.x;
#pragma unroll
            for (int w = 0; w < W; ++w) { r[w] |= k[j][w]; r[W + w] &= k[j][w]; }
            const bool bad = out_of_order<W>(p[j], k[j], so);
            r[2 * W] |= bad ? 1 : 0;
            if (so.last.width) r[2 * W + 1] |= (fget<W>(p[j], so.last) > fget<W>(k[j], so.last)) ? 1 : 0;
            if (bad) {
                uint64_t m[W];
#pragma unroll
                for (int w = 0; w < W; ++w) m[w] = 0;
                first_difference<W>(p[j], k[j], so, m);
#pragma unroll
                for (int w = 0; w < W; ++w) r[2 * W + 2 + w] |= m[w];
            }
            if (hist.width && i < n) {
                const int b = min(__popcll(fget<W>(k[j], hist)), kStatsHistBins - 1);
#pragma unroll
                for (int h = 0; h < kStatsHistBins; ++h) r[3 * W + 2 + h] += (h == b);
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int j = 0; j < R; ++j) r[j] = stats_combine<W>(j, r[j], __shfl_xor(r[j], off, 64));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int j = 0; j < R; ++j) lds[wave][j] = r[j];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < kBlock / 64; ++v)
            for (int j = 0; j < R; ++j) r[j] = stats_combine<W>(j, r[j], lds[v][j]);
        for (int j = 0; j < R; ++j) partial[(uint64_t)blockIdx.x * R + j] = r[j];
    }
}

// Combines the per-block records (the kernel boundary makes them visible; no
// device-scope fences, which cost an L2 write-back per block on this part).
template <int W>
__global__ void __launch_bounds__(kBlock) key_stats_final_kernel(const uint64_t* __restrict__ partial, uint32_t blocks,
                                                            uint64_t* __restrict__ out) {
    constexpr int R = stats_words(W);
    __shared__ uint64_t lds[kBlock / 64][R];
    uint64_t r[R];
#pragma unroll
    for (int j = 0; j < R; ++j) r[j] = (j >= W && j < 2 * W) ? ~0ull : 0;
    for (uint32_t b = threadIdx.x; b < blocks; b += kBlock)
#pragma unroll
        for (int j = 0; j < R; ++j) r[j] = stats_combine<W>(j, r[j], partial[(uint64_t)b * R + j]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int j = 0; j < R; ++j) r[j] = stats_combine<W>(j, r[j], __shfl_xor(r[j], off, 64));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int j = 0; j < R; ++j) lds[wave][j] = r[j];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < kBlock / 64; ++v)
            for (int j = 0; j < R; ++j) r[j] = stats_combine<W>(j, r[j], lds[v][j]);
        for (int j = 0; j < R; ++j) out[j] = r[j];
    }
}

constexpr unsigned kStatsMaxGrid = 1024;

template <int W>
static void launch_key_stats(State& st, const StatsOrder& so, sparq_field_t hist, uint64_t* partial, uint64_t* out) {
    constexpr int kI = W <= 2 ? 8 : (W <= 4 ? 4 : 2);
    unsigned g = std::min<unsigned>(grid_for(st.n, kBlock * kI), kStatsMaxGrid);
    {
        SPQ_PROF(st.stream, "key_stats_kernel", st.n * 8.0 * W);
        key_stats_kernel<W><<<g, kBlock, 0, st.stream>>>(st.keys, st.cap, st.n, so, hist, partial);
    }
    SPQ_HIP(hipGetLastError());
    {
        SPQ_PROF(st.stream, "key_stats_final_kernel", (double)g * stats_words(W) * 8.0);
        key_stats_final_kernel<W><<<1, kBlock, 0, st.stream>>>(partial, g, out);
    }
    SPQ_HIP(hipGetLastError());
}

KeyStats key_stats(State& st, const sparq_field_t* last, const SizerQuery* q) {
    KeyStats ks;
    ks.vary.assign(st.W, 0);
    ks.sorted = true;
    if (st.n == 0) return ks;
    StatsOrder so{};
    for (uint32_t w = 0; w < SPARQ_MAX_WORDS; ++w) so.others[w] = ~0ull;
    if (last) {
        so.last = *last;
        so.others[last->word] &= ~(fmask_host(last->width) << last->shift);
    }
    bool predict = false;
    const bool sized = q && last && sizer_wanted(predict);
    sparq_field_t hist{0, 0, 0, 0};
    if (sized) hist = *last;
    const size_t R = stats_words((int)st.W);
    const size_t fwords = (SPARQ_SIZER_FEATURES + 2) / 2;   // features + prediction, as floats
    uint64_t* partial = dalloc<uint64_t>((size_t)kStatsMaxGrid * R, st.stream);
    uint64_t* out = dalloc<uint64_t>(R + fwords, st.stream);
    SPQ_DISPATCH_W(st.W, launch_key_stats, st, so, hist, partial, out);
    if (sized) launch_sizer_features(st, out, *q, predict, reinterpret_cast<float*>(out + R));
    std::vector<uint64_t> h(R + (sized ? fwords : 0));
    read_back(st, out, h.size() * 8, h.data());
    dfree(partial, st.stream);
    dfree(out, st.stream);
    for (uint32_t w = 0; w < st.W; ++w) ks.vary[w] = h[w] ^ h[st.W + w];
    ks.ones.assign(h.begin(), h.begin() + st.W);
    ks.sorted = h[2 * st.W] == 0;
    ks.last_sorted = last != nullptr && h[2 * st.W + 1] == 0;
    ks.disorder.assign(h.begin() + 2 * st.W + 2, h.begin() + 3 * st.W + 2);
    if (sized) {
        const float* f = reinterpret_cast<const float*>(h.data() + R);
        std::copy(f, f + SPARQ_SIZER_FEATURES, ks.features);
        ks.have_features = true;
        ks.prediction = f[SPARQ_SIZER_FEATURES];
        ks.have_prediction = predict;
    }
    return ks;
}

// Consecutive positions of one word share a digit of up to digit_bits bits
// made of at most three contiguous ranges.  `bits` is in LSD order.
static std::vector<Digit> digits_of(const std::vector<std::pair<int, int>>& bits, int digit_bits) {
    std::vector<Digit> plan;
    for (size_t i = 0; i < bits.size();) {
        Digit d{};
        d.word = (uint8_t)bits[i].first;
        while (i < bits.size() && bits[i].first == d.word && d.bits < digit_bits) {
            const int p = bits[i].second;
            const bool extends = d.nseg > 0 && p == d.shift[d.nseg - 1] + d.width[d.nseg - 1];
            if (!extends) {
                if (d.nseg == 3) break;
                d.shift[d.nseg] = (uint8_t)p;
                d.width[d.nseg] = 0;
                ++d.nseg;
            }
            ++d.width[d.nseg - 1];
            ++d.bits;
            ++i;
        }
        plan.push_back(d);
    }
    return plan;
}

// Varying bit positions, in LSD order, of the ordering
// (major fields in the given order, the remaining key bits in canonical order,
// minor fields in the given order): the minor fields' bits first (the last
// minor field least significant), each field from its lowest bit up, then the
// unlisted bits from the least significant word's lowest bit upward, then the
// major fields.  `skip_minor` leaves the minor fields out (already ordered).
static std::vector<std::pair<int, int>> order_bits(const std::vector<uint64_t>& vary, uint32_t W, const std::vector<sparq_field_t>& major,
                                                    const std::vector<sparq_field_t>& minor, bool skip_minor) {
    std::vector<uint64_t> listed(W, 0);
    for (const auto* fs : {&major, &minor})
        for (const sparq_field_t& f : *fs) listed[f.word] |= fmask_host(f.width) << f.shift;
    std::vector<std::pair<int, int>> bits;
    auto push_field = [&](const sparq_field_t& f) {
        for (int p = f.shift; p < f.shift + f.width; ++p)
            if ((vary[f.word] >> p) & 1) bits.emplace_back(f.word, p);
    };
    if (!skip_minor)
        for (size_t i = minor.size(); i-- > 0;) push_field(minor[i]);
    for (int w = (int)W - 1; w >= 0; --w)
        for (int p = 0; p < 64; ++p)
            if (((vary[w] >> p) & 1) && !((listed[w] >> p) & 1)) bits.emplace_back(w, p);
    for (size_t i = major.size(); i-- > 0;) push_field(major[i]);
    return bits;
}

std::vector<Digit> make_plan(const KeyStats& ks, uint32_t W, const sparq_field_t* last, int digit_bits) {
    // the digits of `last` are left out when that field is already non-decreasing
    std::vector<sparq_field_t> minor;
    if (last) minor.push_back(*last);
    return digits_of(order_bits(ks.vary, W, {}, minor, last && ks.last_sorted), digit_bits);
}

// ---------------------------------------------------------------- digit-width autotuner
// The plan is built for 8- and 11-bit digits; when the wider digits save
// passes, the first sorts of each size class (log2 N) try each width twice --
// an unmeasured warm-up (first allocations of the class, cold caches) and then
// a run whose digit passes alone are timed with HIP events, normalised per
// entry and pass (the two timed sorts have different sizes and pass counts).
// Later sorts of that class take 11-bit digits only when passes x measured
// cost per entry-pass is at least 10 % lower.
namespace {
std::mutex g_tune_mu;
float g_ms_per_pass[2][64];        // [8-bit, 11-bit][log2 N]: ms per pass per 2^20 entries; < 0: not measured
int g_tune_phase[64];              // per class: 0, 1 (8-bit warm-up, measured), 2, 3 (11-bit), 4 done
bool g_tune_init = false;
int g_forced_bits = 0;
}  // namespace

void set_radix_bits(int bits) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_forced_bits = bits;
}

std::string radix_tuning_report() {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    std::string out = "[";
    for (int lg = 0; lg < 64 && g_tune_init; ++lg) {
        if (g_tune_phase[lg] == 0) continue;
        const float c8 = g_ms_per_pass[0][lg], c11 = g_ms_per_pass[1][lg];
        char buf[160];
        snprintf(buf, sizeof(buf), "%s{\"log2_n\":%d,\"ms_per_pass_per_2p20\":[%.5f,%.5f],\"done\":%d}", out.size() > 1 ? "," : "",
                 lg, c8, c11, g_tune_phase[lg] >= 4 ? 1 : 0);
        out += buf;
    }
    return out + "]";
}

// Sort by one of two plans of the same ordering (8- and 11-bit digits), the
// digit width picked by the autotuner.
static void sort_with_plans(State& st, const std::vector<Digit>& p8, const std::vector<Digit>& p11) {
    if (p8.empty()) return;
    int lg = 0;
    while ((uint64_t(1) << (lg + 1)) <= st.n && lg < 63) ++lg;
    int use = 0, measure = -1;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        if (!g_tune_init) {
            for (auto& row : g_ms_per_pass)
                for (float& v : row) v = -1.f;
            for (int& ph : g_tune_phase) ph = 0;
            g_tune_init = true;
        }
        if (g_forced_bits == 8 || g_forced_bits == 11) {
            use = g_forced_bits == 11 ? 1 : 0;
        } else if (p11.size() < p8.size() && st.n >= (1u << 16)) {
            const int ph = g_tune_phase[lg];
            if (ph < 4) {
                use = ph >= 2 ? 1 : 0;
                if (ph & 1) measure = use;
                g_tune_phase[lg] = ph + 1;
            } else {
                // a measurement still running on another thread reads as unmeasured
                const float c8 = g_ms_per_pass[0][lg], c11 = g_ms_per_pass[1][lg];
                use = (c8 > 0.f && c11 > 0.f && p11.size() * c11 < 0.9f * p8.size() * c8) ? 1 : 0;
            }
        }
    }
    const std::vector<Digit>& plan = use ? p11 : p8;
    if (measure < 0) {
        sort_by_plan(st, plan, use ? 11 : 8);
        return;
    }
    // events of this call only: sorts of other states may be timed concurrently
    hipEvent_t ev[2];
    SPQ_HIP(hipEventCreate(&ev[0]));
    SPQ_HIP(hipEventCreate(&ev[1]));
    const uint64_t n = st.n;
    sort_by_plan(st, plan, use ? 11 : 8, ev[0], ev[1]);
    SPQ_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    SPQ_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    SPQ_HIP(hipEventDestroy(ev[0]));
    SPQ_HIP(hipEventDestroy(ev[1]));
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_ms_per_pass[measure][lg] = std::max(ms, 1e-6f) / (float)plan.size() / (float)((double)n / 1048576.0);
}

// Sort by an ordering given as its varying bit positions in LSD order: the
// prefix + local path when it applies (sort_local), else the LSD radix passes
// with the autotuned digit width.
static bool sort_local(State& st, const std::vector<std::pair<int, int>>& bits, uint32_t presorted);
static int g_sort_mode = 0;   // 0: prefix + local sort where it applies, 1: LSD passes only
// sorts by path and the entries each path sorted (sort_path_stats)
enum { kStatLocal, kStatLocalEntries, kStatOverflow, kStatLsd, kStatLsdEntries, kStatPresorted, kStatCount };
static std::atomic<uint64_t> g_sort_stats[kStatCount];

void set_sort_mode(int mode) { g_sort_mode = mode; }

// `presorted`: the top that many bits of the ordering are a prefix the state is
// already sorted by (0: unknown).
static void sort_ordering(State& st, const std::vector<std::pair<int, int>>& bits, uint32_t presorted = 0) {
    if (bits.empty() || st.n <= 1) return;
    if (g_sort_mode == 0 && sort_local(st, bits, presorted)) return;
    g_sort_stats[kStatLsd] += 1;
    g_sort_stats[kStatLsdEntries] += st.n;
    sort_with_plans(st, digits_of(bits, 8), digits_of(bits, 11));
}

// How many of the top ordering bits (LSD list `bits` of the ordering
// (others, last)) lie above the most significant first-difference of the
// pairs out of order: the state is already sorted by them.
static uint32_t presorted_top(const std::vector<std::pair<int, int>>& bits, const KeyStats& ks, const sparq_field_t* last) {
    int wv = -1, pv = -1;
    for (size_t w = 0; w < ks.disorder.size(); ++w)
        if (ks.disorder[w]) {
            wv = (int)w;
            pv = 63 - __builtin_clzll(ks.disorder[w]);
            break;
        }
    uint32_t count = 0;
    for (size_t i = bits.size(); i-- > 0;) {
        const int w = bits[i].first, p = bits[i].second;
        if (last && w == (int)last->word && p >= (int)last->shift && p < (int)(last->shift + last->width)) break;
        if (wv >= 0 && !(w < wv || (w == wv && p > pv))) break;
        ++count;
    }
    return count;
}

void sort_with_stats(State& st, const KeyStats& ks, const sparq_field_t* last) {
    std::vector<sparq_field_t> minor;
    if (last) minor.push_back(*last);
    const std::vector<std::pair<int, int>> bits = order_bits(ks.vary, st.W, {}, minor, last && ks.last_sorted);
    sort_ordering(st, bits, presorted_top(bits, ks, last));
}

void sort_state_by(State& st, const std::vector<sparq_field_t>& major, const std::vector<sparq_field_t>& minor) {
    for (const auto* fs : {&major, &minor})
        for (const sparq_field_t& f : *fs)
            SPQ_REQUIRE(f.word < st.W && f.width >= 1 && (uint32_t)f.shift + f.width <= 64, "sort field out of range");
    std::vector<uint64_t> listed(st.W, 0);
    for (const auto* fs : {&major, &minor})
        for (const sparq_field_t& f : *fs) {
            const uint64_t m = fmask_host(f.width) << f.shift;
            SPQ_REQUIRE((listed[f.word] & m) == 0, "sort fields overlap");
            listed[f.word] |= m;
        }
    if (major.empty() && minor.size() <= 1) {
        sort_state(st, minor.empty() ? nullptr : &minor[0]);
        return;
    }
    if (st.n <= 1) return;
    const KeyStats ks = key_stats(st, nullptr);
    sort_ordering(st, order_bits(ks.vary, st.W, major, minor, false));
}

// Sort key of SortByAmplitude: |a|^2 descending.  A non-negative double's bit
// pattern orders like its value, so the complement orders descending.
__global__ void __launch_bounds__(kBlock) amp_order_key_kernel(const double2* __restrict__ amps, uint64_t n, uint64_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double2 a = amps[i];
    const double p = a.x * a.x + a.y * a.y;
    out[i] = ~(uint64_t)__double_as_longlong(p);
}

void sort_by_amplitude(State& st) {
    if (st.n <= 1) return;
    uint64_t* aux = dalloc<uint64_t>(st.n, st.stream);
    {
        SPQ_PROF(st.stream, "amp_order_key_kernel", st.n * 24.0);
        amp_order_key_kernel<<<grid_for(st.n, kBlock), kBlock, 0, st.stream>>>(st.amps, st.n, aux);
    }
    SPQ_HIP(hipGetLastError());
    State view = st;   // one-word key view of the order keys, for their statistics
    view.W = 1;
    view.keys = aux;
    view.cap = st.n;
    const KeyStats ks = key_stats(view, nullptr);
    if (!ks.sorted) {
        std::vector<std::pair<int, int>> bits;
        for (int p = 0; p < 64; ++p)
            if ((ks.vary[0] >> p) & 1) bits.emplace_back(0, p);
        sort_by_plan(st, digits_of(bits, 8), 8, nullptr, nullptr, aux, st.n);
    }
    dfree(aux, st.stream);
}

bool sort_state(State& st, const sparq_field_t* last) {
    if (st.n <= 1) return false;
    KeyStats ks = key_stats(st, last);
    if (ks.sorted) return false;
    sort_with_stats(st, ks, last);
    return true;
}

// ---------------------------------------------------------------- passes
__device__ __forceinline__ uint32_t digit_of(uint64_t k, const Digit& dg) {
    uint32_t v = 0, at = 0;
    for (int s = 0; s < dg.nseg; ++s) {
        v |= (uint32_t)((k >> dg.shift[s]) & ((1ull << dg.width[s]) - 1)) << at;
        at += dg.width[s];
    }
    return v;
}

// Block histogram, block-major: block_hist[b * 2^RB + d] for digits of up to RB
// bits (one contiguous row per block; a digit-major layout would spread each
// block's counts over 2^RB lines shared with other blocks on other XCDs).
template <int RB, int BS>
__global__ void __launch_bounds__(BS) radix_hist_kernel(const uint64_t* __restrict__ kw, uint64_t n, Digit dg,
                                                   uint32_t* __restrict__ block_hist, uint32_t nb) {
    constexpr int kR = 1 << RB;
    constexpr uint32_t kInvalid = 1u << RB;
    constexpr int NW = BS / 64, NI = kSortTile / BS;
    __shared__ uint32_t hist[kR];
    for (int d = threadIdx.x; d < kR; d += BS) hist[d] = 0;
    __syncthreads();
    const uint64_t tile = (uint64_t)blockIdx.x * kSortTile;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // all loads of the tile first (clamped in-bounds indices: no per-item wait)
    uint64_t key[NI];
#pragma unroll
    for (int it = 0; it < NI; ++it) {
        const uint64_t e = tile + (uint64_t)wave * (kSortTile / NW) + it * 64 + lane;
        key[it] = kw[e < n ? e : n - 1];
    }
#pragma unroll
    for (int it = 0; it < NI; ++it) {
        const uint64_t e = tile + (uint64_t)wave * (kSortTile / NW) + it * 64 + lane;
        const uint32_t d = e < n ? digit_of(key[it], dg) : kInvalid;
        const uint64_t peers = match_any<RB>(d, dg.bits);
        const int leader = __ffsll((unsigned long long)peers) - 1;
        if (lane == leader && d != kInvalid) atomicAdd(&hist[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kR; d += BS) block_hist[(uint64_t)blockIdx.x * kR + d] = hist[d];
}

// out[c * rows + r] = in[r * cols + c] through 32 x 32 LDS tiles (both sides
// coalesced); turns the block-major histograms into the digit-major order of
// the offset scan.
__global__ void __launch_bounds__(kBlock) transpose_u32_kernel(const uint32_t* __restrict__ in, uint32_t rows, uint32_t cols,
                                                          uint32_t* __restrict__ out) {
    __shared__ uint32_t t[32][33];
    const uint32_t c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
#pragma unroll
    for (uint32_t k = 0; k < 32; k += 8) {
        const uint32_t r = r0 + ty + k, c = c0 + tx;
        if (r < rows && c < cols) t[ty + k][tx] = in[(uint64_t)r * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < 32; k += 8) {
        const uint32_t c = c0 + ty + k, r = r0 + tx;
        if (r < rows && c < cols) out[(uint64_t)c * rows + r] = t[tx][ty + k];
    }
}

// Stable scatter of one digit pass.  idx_in == nullptr means identity order.
// BS threads (BS / 64 waves) rank a tile of kSortTile entries; the digit scan
// gives each of the first kR threads (or each thread kR / BS digits) its digits.
template <int RB, int BS>
__global__ void __launch_bounds__(BS) radix_scatter_kernel(const uint64_t* __restrict__ kw_in, const uint32_t* __restrict__ idx_in,
                                                      uint64_t n, Digit dg, const uint64_t* __restrict__ offs, uint32_t nb,
                                                      uint64_t* __restrict__ kw_out, uint32_t* __restrict__ idx_out,
                                                      int write_keys) {
    constexpr int kR = 1 << RB;
    constexpr uint32_t kInvalid = 1u << RB;
    constexpr int NW = BS / 64, NI = kSortTile / BS;
    constexpr int kPerThread = kR >= BS ? kR / BS : 1;
    static_assert(kR % BS == 0 || BS % kR == 0, "radix and block size must divide one another");
    __shared__ uint32_t wcnt[NW][kR];          // per-wave running digit counts
    __shared__ uint32_t dstart[kR];            // first tile position of each digit
    __shared__ uint64_t goff[kR];              // global position of tile position 0 of each digit
    __shared__ uint64_t skey[kSortTile];
    __shared__ uint32_t sidx[kSortTile];
    __shared__ uint32_t scan_lds[NW];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int d = threadIdx.x; d < kR; d += BS) {
#pragma unroll
        for (int w = 0; w < NW; ++w) wcnt[w][d] = 0;
    }
    __syncthreads();

    // Each digit's output region holds the runs of tile 0, 1, 2, ... in order;
    // neighbouring tiles' runs share lines, so they are written from one XCD.
    const uint32_t tile_id = xcd_tile(blockIdx.x, nb);
    const uint64_t tile = (uint64_t)tile_id * kSortTile;
    uint64_t key[NI];
    uint32_t idx[NI];
    uint32_t rank[NI];
    uint32_t dig[NI];
    // all loads of the tile first (clamped in-bounds indices: no per-item wait)
#pragma unroll
    for (int it = 0; it < NI; ++it) {
        const uint64_t e = tile + (uint64_t)wave * (kSortTile / NW) + it * 64 + lane;
        const uint64_t ec = e < n ? e : n - 1;
        key[it] = kw_in[ec];
        idx[it] = idx_in ? idx_in[ec] : (uint32_t)ec;
    }
    // Stable per-wave ranking: items in (it, lane) order, which is entry order.
    // The leader of each group of equal digits adds the group's size to the
    // wave's digit counter with one LDS atomic that returns the old count; a
    // wave's LDS operations complete in issue order, so item it+1 sees item
    // it's additions and the 16 atomics need no wait between them.
    uint32_t before[NI], got[NI];
    int leader[NI];
#pragma unroll
    for (int it = 0; it < NI; ++it) {
        const uint64_t e = tile + (uint64_t)wave * (kSortTile / NW) + it * 64 + lane;
        const bool valid = e < n;
        const uint32_t d = valid ? digit_of(key[it], dg) : kInvalid;
        dig[it] = d;
        const uint64_t peers = match_any<RB>(d, dg.bits);
        before[it] = (uint32_t)__popcll(peers & lanemask_lt());
        leader[it] = __ffsll((unsigned long long)peers) - 1;
        got[it] = (valid && before[it] == 0) ? atomicAdd(&wcnt[wave][d & (kR - 1)], (uint32_t)__popcll(peers)) : 0u;
    }
#pragma unroll
    for (int it = 0; it < NI; ++it) rank[it] = (uint32_t)__shfl((int)got[it], leader[it], 64) + before[it];
    __syncthreads();
    // Digit totals and their exclusive scan over digits; per-wave offsets.
    const bool owner = kR >= BS || (int)threadIdx.x < kR;
    uint32_t c[kPerThread][NW];
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const int d = owner ? threadIdx.x * kPerThread + q : 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            c[q][w] = owner ? wcnt[w][d] : 0;
            sum += c[q][w];
        }
    }
    uint32_t tot;
    uint32_t start = block_excl_scan<uint32_t>(sum, scan_lds, tot);
    if (owner) {
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const int d = threadIdx.x * kPerThread + q;
            dstart[d] = start;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                wcnt[w][d] = start;
                start += c[q][w];
            }
            goff[d] = offs[(uint64_t)d * nb + tile_id];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NI; ++it) {
        if (dig[it] != kInvalid) {
            uint32_t pos = wcnt[wave][dig[it]] + rank[it];
            skey[pos] = key[it];
            sidx[pos] = idx[it];
        }
    }
    __syncthreads();
    const uint32_t count = (uint32_t)min<uint64_t>((uint64_t)kSortTile, n - tile);
    for (uint32_t p = threadIdx.x; p < count; p += BS) {
        uint64_t k = skey[p];
        uint32_t dd = digit_of(k, dg);
        uint64_t dst = goff[dd] + (p - dstart[dd]);
        if (write_keys) kw_out[dst] = k;
        idx_out[dst] = sidx[p];
    }
}

constexpr int kGatherItems = 4;

// dst[i] = src[idx[i]], kGatherItems entries per lane (striped by the block
// size): the index loads, then the word loads, are in flight together.
__global__ void __launch_bounds__(kBlock) gather_word_kernel(const uint64_t* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                                                        uint64_t* __restrict__ dst) {
    const uint64_t base = (uint64_t)blockIdx.x * kBlock * kGatherItems + threadIdx.x;
    uint32_t s[kGatherItems];
#pragma unroll
    for (int it = 0; it < kGatherItems; ++it) {
        const uint64_t i = base + (uint64_t)it * kBlock;
        s[it] = idx[i < n ? i : n - 1];
    }
    uint64_t v[kGatherItems];
#pragma unroll
    for (int it = 0; it < kGatherItems; ++it) v[it] = src[s[it]];
#pragma unroll
    for (int it = 0; it < kGatherItems; ++it) {
        const uint64_t i = base + (uint64_t)it * kBlock;
        if (i < n) dst[i] = v[it];
    }
}

template <int W>
__global__ void __launch_bounds__(kBlock) gather_state_kernel(const uint64_t* __restrict__ kin, const double2* __restrict__ ain, uint64_t cap_in,
                                                         const uint32_t* __restrict__ idx, uint64_t n, uint64_t* __restrict__ kout,
                                                         double2* __restrict__ aout, uint64_t cap_out) {
    // kGatherItems entries per thread (striped by the block size): their index
    // loads, then their key and amplitude loads, are in flight together.
    const uint64_t base = (uint64_t)blockIdx.x * kBlock * kGatherItems + threadIdx.x;
    uint32_t s[kGatherItems];
#pragma unroll
    for (int it = 0; it < kGatherItems; ++it) {
        const uint64_t i = base + (uint64_t)it * kBlock;
        s[it] = idx[i < n ? i : n - 1];
    }
    uint64_t k[kGatherItems][W];
    double2 a[kGatherItems];
#pragma unroll
    for (int it = 0; it < kGatherItems; ++it) {
#pragma unroll
        for (int w = 0; w < W; ++w) k[it][w] = kin[(uint64_t)w * cap_in + s[it]];
        a[it] = ain[s[it]];
    }
#pragma unroll
    for (int it = 0; it < kGatherItems; ++it) {
        const uint64_t i = base + (uint64_t)it * kBlock;
        if (i >= n) break;
#pragma unroll
        for (int w = 0; w < W; ++w) kout[(uint64_t)w * cap_out + i] = k[it][w];
        aout[i] = a[it];
    }
}

template <int W>
static void launch_gather_state(State& st, const uint32_t* idx, uint64_t* kout, double2* aout, uint64_t cap_out) {
    {
        SPQ_PROF(st.stream, "gather_state_kernel", st.n * (4.0 + 2 * (8.0 * W + 16)));
        gather_state_kernel<W><<<grid_for(st.n, kBlock * kGatherItems), kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, idx, st.n, kout,
                                                                                            aout, cap_out);
    }
    SPQ_HIP(hipGetLastError());
}

void sort_by_plan(State& st, const std::vector<Digit>& plan, int radix_bits, hipEvent_t pass_begin, hipEvent_t pass_end,
                  const uint64_t* word_src, uint64_t word_cap) {
    const uint64_t n = st.n;
    // the digits read key words from the state, or from word_src when given
    const uint64_t* wbase = word_src ? word_src : st.keys;
    const uint64_t wcap = word_src ? word_cap : st.cap;
    if (plan.empty() || n <= 1) return;
    SPQ_REQUIRE(radix_bits == 8 || radix_bits == 11, "digit width must be 8 or 11");
    static const bool trace = std::getenv("SPARQ_TRACE_SORT") != nullptr;   // diagnostic: one line per sort
    if (trace) {
        std::fprintf(stderr, "[sort] n=%llu radix=%d passes=%zu:", (unsigned long long)n, radix_bits, plan.size());
        for (const Digit& d : plan) {
            std::fprintf(stderr, " w%u[", d.word);
            for (int s = 0; s < d.nseg; ++s) std::fprintf(stderr, "%s%u+%u", s ? "," : "", d.shift[s], d.width[s]);
            std::fprintf(stderr, "]");
        }
        std::fprintf(stderr, "\n");
    }
    for (const Digit& d : plan) SPQ_REQUIRE(d.bits >= 1 && d.bits <= radix_bits && d.nseg >= 1 && d.nseg <= 3, "bad digit");
    SPQ_REQUIRE(n < (1ull << 32), "sort supports fewer than 2^32 entries");
    const uint32_t nb = grid_for(n, kSortTile);
    const uint64_t radix = uint64_t(1) << radix_bits;
    uint64_t* kwA = dalloc<uint64_t>(n, st.stream);
    uint64_t* kwB = dalloc<uint64_t>(n, st.stream);
    uint32_t* idxA = dalloc<uint32_t>(n, st.stream);
    uint32_t* idxB = dalloc<uint32_t>(n, st.stream);
    uint32_t* bh = dalloc<uint32_t>((size_t)radix * nb, st.stream);      // digit-major
    uint32_t* bh_rows = dalloc<uint32_t>((size_t)radix * nb, st.stream); // block-major
    uint64_t* offs = dalloc<uint64_t>((size_t)radix * nb, st.stream);

    bool identity = true;
    int cur_word = -1;
    const uint64_t* src = nullptr;
    if (pass_begin) SPQ_HIP(hipEventRecord(pass_begin, st.stream));   // the digit passes only
    for (size_t p = 0; p < plan.size(); ++p) {
        const Digit dg = plan[p];
        if (dg.word != cur_word) {
            if (identity) {
                src = wbase + (uint64_t)dg.word * wcap;
            } else {
                {
                    SPQ_PROF(st.stream, "gather_word_kernel", n * 20.0);
                    gather_word_kernel<<<grid_for(n, kBlock * kGatherItems), kBlock, 0, st.stream>>>(wbase + (uint64_t)dg.word * wcap,
                                                                                                   idxA, n, kwA);
                }
                SPQ_HIP(hipGetLastError());
                src = kwA;
            }
            cur_word = dg.word;
        }
        {
            SPQ_PROF(st.stream, "radix_hist_kernel", n * 8.0);
            if (radix_bits == 11) radix_hist_kernel<11, kSortBlock11><<<nb, kSortBlock11, 0, st.stream>>>(src, n, dg, bh_rows, nb);
            else radix_hist_kernel<8, kSortBlock8><<<nb, kSortBlock8, 0, st.stream>>>(src, n, dg, bh_rows, nb);
        }
        SPQ_HIP(hipGetLastError());
        {
            SPQ_PROF(st.stream, "transpose_u32_kernel", (double)radix * nb * 8.0);
            const dim3 tg((unsigned)((radix + 31) / 32), (nb + 31) / 32);
            transpose_u32_kernel<<<tg, kBlock, 0, st.stream>>>(bh_rows, nb, (uint32_t)radix, bh);
        }
        SPQ_HIP(hipGetLastError());
        exclusive_scan_u32(st, bh, offs, radix * nb, nullptr);
        // The key word of the last pass of a word is not needed afterwards.
        bool need_keys = (p + 1 < plan.size()) && plan[p + 1].word == dg.word;
        {
            SPQ_PROF(st.stream, "radix_scatter_kernel", n * (8.0 + (identity ? 0.0 : 4.0) + 4.0 + (need_keys ? 8.0 : 0.0)));
            if (radix_bits == 11)
                radix_scatter_kernel<11, kSortBlock11><<<nb, kSortBlock11, 0, st.stream>>>(src, identity ? nullptr : idxA, n, dg, offs, nb,
                                                                                        kwB, idxB, need_keys ? 1 : 0);
            else
                radix_scatter_kernel<8, kSortBlock8><<<nb, kSortBlock8, 0, st.stream>>>(src, identity ? nullptr : idxA, n, dg, offs, nb,
                                                                                      kwB, idxB, need_keys ? 1 : 0);
        }
        SPQ_HIP(hipGetLastError());
        std::swap(kwA, kwB);
        std::swap(idxA, idxB);
        identity = false;
        src = kwA;
    }
    if (pass_end) SPQ_HIP(hipEventRecord(pass_end, st.stream));
    uint64_t* kout;
    double2* aout;
    alloc_arrays(st, st.W, st.cap, &kout, &aout);
    SPQ_DISPATCH_W(st.W, launch_gather_state, st, idxA, kout, aout, st.cap);
    install_arrays(st, st.W, kout, aout, n, st.cap);
    dfree(kwA, st.stream);
    dfree(kwB, st.stream);
    dfree(idxA, st.stream);
    dfree(idxB, st.stream);
    dfree(bh, st.stream);
    dfree(bh_rows, st.stream);
    dfree(offs, st.stream);
}


// ================================================================ prefix + local sort
// For states that fit (key words W <= kLocalMaxW, at most 128 ordering bits
// below the prefix): a few LSD passes over the top B ordering bits move whole
// entries (key words and amplitude, 8-bit digits, tile-staged writes), which
// leaves buckets of equal prefix contiguous; then one workgroup per chunk of
// consecutive buckets (<= kLocalCap entries) sorts the chunk in LDS by the
// remaining bits (compacted into <= 2 words) and its bucket ordinal, and
// writes the entries in order.  No index permutation and no random gather over
// the whole state: the chunk's own entries are re-read from L2.  A chunk
// larger than kLocalCap (a skewed key distribution) sets a device flag; the
// host then discards the output and runs the LSD path on the untouched input.
constexpr int kLocalCap = 2048;       // entries per chunk
constexpr int kLocalBlock = 256;      // 4 waves of 8 items (three workgroups per CU)
constexpr int kLocalMaxW = 5;
constexpr int kMaxSegs = 40;
constexpr int kPayloadTile = 4096;    // entries per workgroup of a payload pass
constexpr int kPayloadBlock = 512;

// MSB-first bit ranges of the key whose concatenation is a compact key.
struct BitSegs {
    uint32_t n;
    uint8_t word[kMaxSegs], shift[kMaxSegs], width[kMaxSegs];
};
struct LocalArgs {
    BitSegs pre;        // the top B ordering bits (bucket prefix)
    BitSegs rem;        // the remaining R ordering bits
    uint32_t rem_bits;  // R
    uint32_t h;         // chunk spacing: chunk j starts at the first bucket start >= j*h
    uint64_t n;
};

template <int W>
__device__ __forceinline__ uint64_t prefix_of(const uint64_t (&k)[W], const BitSegs& sg) {
    uint64_t v = 0;
    for (uint32_t i = 0; i < sg.n; ++i) {
        const uint32_t w = sg.width[i];
        const uint64_t bits = (kword<W>(k, sg.word[i]) >> sg.shift[i]) & fmask(w);
        v = (w >= 64 ? 0 : (v << w)) | bits;
    }
    return v;
}

// Compact key of the remaining bits: Q words, word 0 least significant.
template <int W, int Q>
__device__ __forceinline__ void compact_of(const uint64_t (&k)[W], const BitSegs& sg, uint64_t (&c)[Q]) {
#pragma unroll
    for (int q = 0; q < Q; ++q) c[q] = 0;
    for (uint32_t i = 0; i < sg.n; ++i) {
        const uint32_t w = sg.width[i];
        const uint64_t bits = (kword<W>(k, sg.word[i]) >> sg.shift[i]) & fmask(w);
        if (Q == 2) c[Q - 1] = w >= 64 ? c[0] : ((c[Q - 1] << w) | (w ? (c[0] >> (64 - w)) : 0));
        c[0] = (w >= 64 ? 0 : (c[0] << w)) | bits;
    }
}

// First bucket start at or after q0 (bucket = run of equal prefix), or n.
// Called by the whole workgroup; the result is uniform.  The scan stops after
// kLocalCap + 1 positions: finding no head there means entries q0 - 1 ..
// q0 + kLocalCap share one prefix, a bucket longer than a chunk, and the
// result is kNoHead (the caller reports an overflow).
constexpr uint64_t kNoHead = ~0ull;
template <int W>
__device__ uint64_t chunk_start(uint64_t q0, const uint64_t* __restrict__ keys, uint64_t cap, const LocalArgs& la, uint32_t* lds_min) {
    if (q0 == 0) return 0;
    const uint64_t limit = min(la.n, q0 + (uint64_t)kLocalCap + 1);
    for (uint64_t base = q0; base < limit; base += kLocalBlock) {
        const uint64_t q = base + threadIdx.x;
        bool head = false;
        if (q < limit) {   // q >= 1: q0 >= 1
            uint64_t a[W], b[W];
            load_key<W>(keys, cap, q, a);
            load_key<W>(keys, cap, q - 1, b);
            head = prefix_of<W>(a, la.pre) != prefix_of<W>(b, la.pre);
        }
        if (threadIdx.x == 0) *lds_min = ~0u;
        __syncthreads();
        if (head) atomicMin(lds_min, (uint32_t)threadIdx.x);
        __syncthreads();
        const uint32_t m = *lds_min;
        __syncthreads();
        if (m != ~0u) return base + m;
    }
    return limit == la.n ? la.n : kNoHead;
}

template <int W, int Q>
__global__ void __launch_bounds__(kLocalBlock) local_sort_kernel(const uint64_t* __restrict__ kin, const double2* __restrict__ ain,
                                                            uint64_t cap_in, LocalArgs la, uint64_t* __restrict__ kout,
                                                            double2* __restrict__ aout, uint64_t cap_out,
                                                            unsigned int* __restrict__ overflow) {
    constexpr int NW = kLocalBlock / 64, IT = kLocalCap / kLocalBlock, R = 256;
    constexpr uint32_t kInvalid = 1u << 8;
    __shared__ uint64_t sck[Q][kLocalCap];
    __shared__ uint32_t soi[kLocalCap];       // bucket ordinal << 16 | chunk position of the entry
    __shared__ uint32_t wcnt[NW][R];
    uint64_t* spre = &sck[0][0];              // prefixes, before the passes use sck
    __shared__ uint32_t rowpre[NW * IT];
    __shared__ uint32_t scan_lds[NW];
    __shared__ uint64_t red[2 * Q][NW];
    __shared__ uint32_t misc[2];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t s = la.h ? chunk_start<W>((uint64_t)blockIdx.x * la.h, kin, cap_in, la, &misc[0]) : 0;
    const uint64_t e = la.h && s != kNoHead ? chunk_start<W>((uint64_t)(blockIdx.x + 1) * la.h, kin, cap_in, la, &misc[0]) : la.n;
    if (s == kNoHead || e == kNoHead || e - s > (uint64_t)kLocalCap) {
        if (threadIdx.x == 0) atomicOr(overflow, 1u);
        return;
    }
    const uint64_t len = e - s;
    if (len == 0) return;

    // ---- load: item it of (wave, lane) is chunk position p = wave*(cap/NW) + it*64 + lane
    uint64_t ck[IT][Q];
    uint32_t oi[IT];   // bucket ordinal << 16 | chunk position (both < kLocalCap)
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const uint32_t p = wave * (kLocalCap / NW) + it * 64 + lane;
        const uint64_t src = s + (p < len ? p : len - 1);
        uint64_t a[W];
        load_key<W>(kin, cap_in, src, a);
        compact_of<W, Q>(a, la.rem, ck[it]);
        if (p < len) spre[p] = prefix_of<W>(a, la.pre);
        oi[it] = p;
    }
    __syncthreads();
    // ---- bucket ordinal: heads where the prefix changes (position 0 starts the chunk's first bucket)
    uint32_t mine[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const uint32_t p = wave * (kLocalCap / NW) + it * 64 + lane;
        const bool head = p < len && (p == 0 || spre[p] != spre[p - 1]);
        const uint64_t m = __ballot(head);
        mine[it] = (uint32_t)__popcll(m & (lanemask_lt() | (1ull << lane)));   // heads at or before p in this row
        if (lane == 0) rowpre[wave * IT + it] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < 64) {   // exclusive scan of the NW*IT row counts, rows in position order
        const uint32_t v = threadIdx.x < NW * IT ? rowpre[threadIdx.x] : 0;
        const uint32_t inc = wave_incl_scan(v);
        if (threadIdx.x < NW * IT) rowpre[threadIdx.x] = inc - v;
        if (threadIdx.x == 63) misc[1] = inc;     // bucket count
    }
    __syncthreads();
    const uint32_t nbuckets = misc[1];
#pragma unroll
    for (int it = 0; it < IT; ++it) oi[it] |= (rowpre[wave * IT + it] + mine[it] - 1) << 16;
    // ---- bits of the compact key that vary inside the chunk
    uint64_t vo[Q], va[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        vo[q] = 0;
        va[q] = ~0ull;
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const uint32_t p = wave * (kLocalCap / NW) + it * 64 + lane;
        if (p < len) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                vo[q] |= ck[it][q];
                va[q] &= ck[it][q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            vo[q] |= __shfl_xor(vo[q], off, 64);
            va[q] &= __shfl_xor(va[q], off, 64);
        }
        if (lane == 0) {
            red[2 * q][wave] = vo[q];
            red[2 * q + 1][wave] = va[q];
        }
    }
    __syncthreads();
    uint64_t vary[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        uint64_t o = 0, an = ~0ull;
        for (int w = 0; w < NW; ++w) {
            o |= red[2 * q][w];
            an &= red[2 * q + 1][w];
        }
        vary[q] = o ^ an;
    }
    const uint32_t ord_bits = nbuckets <= 1 ? 0 : 32 - __clz(nbuckets - 1);

    // ---- LSD passes over 8-bit windows: compact key windows that vary, then the ordinal
    const uint32_t key_windows = (la.rem_bits + 7) / 8;
    const uint32_t ord_windows = (ord_bits + 7) / 8;
    for (uint32_t win = 0; win < key_windows + ord_windows; ++win) {
        const bool of_ord = win >= key_windows;
        const uint32_t wq = of_ord ? 0 : (win * 8) / 64, wsh = of_ord ? (win - key_windows) * 8 : (win * 8) % 64;
        if (!of_ord) {
            uint64_t vw = 0;
#pragma unroll
            for (int q = 0; q < Q; ++q) vw |= (q == (int)wq) ? vary[q] : 0;
            if (((vw >> wsh) & 0xff) == 0) continue;   // uniform: the window does not vary
        }
        for (int d = threadIdx.x; d < R; d += kLocalBlock)
#pragma unroll
            for (int w = 0; w < NW; ++w) wcnt[w][d] = 0;
        __syncthreads();
        uint32_t dig[IT], before[IT], got[IT];
        int leader[IT];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const uint32_t p = wave * (kLocalCap / NW) + it * 64 + lane;
            uint64_t src = 0;
#pragma unroll
            for (int q = 0; q < Q; ++q) src |= (q == (int)wq) ? ck[it][q] : 0;
            const uint32_t d = p < len ? (uint32_t)(((of_ord ? (uint64_t)(oi[it] >> 16) : src) >> wsh) & 0xff) : kInvalid;
            dig[it] = d;
            const uint64_t peers = match_any<8>(d, 8);
            before[it] = (uint32_t)__popcll(peers & lanemask_lt());
            leader[it] = __ffsll((unsigned long long)peers) - 1;
            got[it] = (p < len && before[it] == 0) ? atomicAdd(&wcnt[wave][d & (R - 1)], (uint32_t)__popcll(peers)) : 0u;
        }
        uint32_t rank[IT];
#pragma unroll
        for (int it = 0; it < IT; ++it) rank[it] = (uint32_t)__shfl((int)got[it], leader[it], 64) + before[it];
        __syncthreads();
        // digit totals -> exclusive scan over digits -> per-wave starts
        uint32_t c[NW], sum = 0;
        const bool owner = (int)threadIdx.x < R;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            c[w] = owner ? wcnt[w][threadIdx.x] : 0;
            sum += c[w];
        }
        uint32_t tot;
        uint32_t start = block_excl_scan<uint32_t>(sum, scan_lds, tot);
        if (owner) {
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                wcnt[w][threadIdx.x] = start;
                start += c[w];
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            if (dig[it] != kInvalid) {
                const uint32_t pos = wcnt[wave][dig[it]] + rank[it];
#pragma unroll
                for (int q = 0; q < Q; ++q) sck[q][pos] = ck[it][q];
                soi[pos] = oi[it];
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const uint32_t p = wave * (kLocalCap / NW) + it * 64 + lane;
            if (p < len) {
#pragma unroll
                for (int q = 0; q < Q; ++q) ck[it][q] = sck[q][p];
                oi[it] = soi[p];
            }
        }
        __syncthreads();
    }
    // ---- write out: position p takes the entry idx of the chunk (re-read from L2)
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const uint32_t p = wave * (kLocalCap / NW) + it * 64 + lane;
        if (p < len) {
            const uint64_t src = s + (oi[it] & 0xffffu);
            uint64_t a[W];
            load_key<W>(kin, cap_in, src, a);
            const double2 v = ain[src];
            store_key<W>(kout, cap_out, s + p, a);
            aout[s + p] = v;
        }
    }
}

// One stable LSD pass over the whole entries (W key words and the amplitude)
// by an 8-bit digit: per-wave ballot ranking, then every plane (each key word,
// then the two amplitude halves) is staged through LDS in digit order and
// written out in runs.
template <int W>
__global__ void __launch_bounds__(kPayloadBlock) payload_scatter_kernel(const uint64_t* __restrict__ kin, const double2* __restrict__ ain,
                                                                   uint64_t cap_in, uint64_t n, Digit dg, const uint64_t* __restrict__ offs,
                                                                   uint32_t nb, uint64_t* __restrict__ kout, double2* __restrict__ aout,
                                                                   uint64_t cap_out) {
    constexpr int R = 256, NW = kPayloadBlock / 64, NI = kPayloadTile / kPayloadBlock;
    constexpr uint32_t kInvalid = 1u << 8;
    __shared__ uint32_t wcnt[NW][R];
    __shared__ uint32_t dstart[R];
    __shared__ uint64_t goff[R];
    __shared__ uint64_t stage[kPayloadTile];
    __shared__ uint16_t sdig[kPayloadTile];
    __shared__ uint32_t scan_lds[NW];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int d = threadIdx.x; d < R; d += kPayloadBlock)
#pragma unroll
        for (int w = 0; w < NW; ++w) wcnt[w][d] = 0;
    __syncthreads();
    const uint32_t tile_id = xcd_tile(blockIdx.x, nb);
    const uint64_t tile = (uint64_t)tile_id * kPayloadTile;
    uint32_t dig[NI], rank[NI];
    {
        uint32_t before[NI], got[NI];
        int leader[NI];
#pragma unroll
        for (int it = 0; it < NI; ++it) {
            const uint64_t e = tile + (uint64_t)wave * (kPayloadTile / NW) + it * 64 + lane;
            const uint64_t word = kin[(uint64_t)dg.word * cap_in + (e < n ? e : n - 1)];
            const uint32_t d = e < n ? digit_of(word, dg) : kInvalid;
            dig[it] = d;
            const uint64_t peers = match_any<8>(d, dg.bits);
            before[it] = (uint32_t)__popcll(peers & lanemask_lt());
            leader[it] = __ffsll((unsigned long long)peers) - 1;
            got[it] = (e < n && before[it] == 0) ? atomicAdd(&wcnt[wave][d & (R - 1)], (uint32_t)__popcll(peers)) : 0u;
        }
#pragma unroll
        for (int it = 0; it < NI; ++it) rank[it] = (uint32_t)__shfl((int)got[it], leader[it], 64) + before[it];
    }
    __syncthreads();
    {
        uint32_t c[NW], sum = 0;
        const bool owner = (int)threadIdx.x < R;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            c[w] = owner ? wcnt[w][threadIdx.x] : 0;
            sum += c[w];
        }
        uint32_t tot;
        uint32_t start = block_excl_scan<uint32_t>(sum, scan_lds, tot);
        if (owner) {
            dstart[threadIdx.x] = start;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                wcnt[w][threadIdx.x] = start;
                start += c[w];
            }
            goff[threadIdx.x] = offs[(uint64_t)threadIdx.x * nb + tile_id];
        }
    }
    __syncthreads();
    uint32_t pos[NI];
#pragma unroll
    for (int it = 0; it < NI; ++it) {
        pos[it] = dig[it] != kInvalid ? wcnt[wave][dig[it]] + rank[it] : 0u;
        if (dig[it] != kInvalid) sdig[pos[it]] = (uint16_t)dig[it];
    }
    const uint32_t count = (uint32_t)min<uint64_t>((uint64_t)kPayloadTile, n - tile);
    // planes: key words 0..W-1, then the amplitude's real and imaginary halves
    double im[NI];
    for (int plane = 0; plane < W + 2; ++plane) {
#pragma unroll
        for (int it = 0; it < NI; ++it) {
            const uint64_t e = tile + (uint64_t)wave * (kPayloadTile / NW) + it * 64 + lane;
            if (dig[it] != kInvalid) {
                uint64_t v;
                if (plane < W) {
                    v = kin[(uint64_t)plane * cap_in + e];
                } else if (plane == W) {
                    const double2 a = ain[e];   // one load; the imaginary half waits in a register
                    im[it] = a.y;
                    v = (uint64_t)__double_as_longlong(a.x);
                } else {
                    v = (uint64_t)__double_as_longlong(im[it]);
                }
                stage[pos[it]] = v;
            }
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < count; p += kPayloadBlock) {
            const uint32_t d = sdig[p];
            const uint64_t dst = goff[d] + (p - dstart[d]);
            const uint64_t v = stage[p];
            if (plane < W) kout[(uint64_t)plane * cap_out + dst] = v;
            else reinterpret_cast<double*>(aout + dst)[plane - W] = __longlong_as_double((long long)v);
        }
        __syncthreads();
    }
}

template <int W>
static void launch_payload_pass(State& st, const uint64_t* kin, const double2* ain, uint64_t cap_in, const Digit& dg, const uint64_t* offs,
                                uint32_t nb, uint64_t* kout, double2* aout, uint64_t cap_out) {
    {
        SPQ_PROF(st.stream, "payload_scatter_kernel", st.n * 2.0 * (8.0 * W + 16));
        payload_scatter_kernel<W><<<nb, kPayloadBlock, 0, st.stream>>>(kin, ain, cap_in, st.n, dg, offs, nb, kout, aout, cap_out);
    }
    SPQ_HIP(hipGetLastError());
}

template <int W>
static void launch_local(State& st, const uint64_t* kin, const double2* ain, uint64_t cap_in, const LocalArgs& la, unsigned grid, int Q,
                         uint64_t* kout, double2* aout, uint64_t cap_out, unsigned int* overflow) {
    SPQ_PROF(st.stream, "local_sort_kernel", st.n * (8.0 * W + 2.0 * (8.0 * W + 16)));
    if (Q == 1)
        local_sort_kernel<W, 1><<<grid, kLocalBlock, 0, st.stream>>>(kin, ain, cap_in, la, kout, aout, cap_out, overflow);
    else
        local_sort_kernel<W, 2><<<grid, kLocalBlock, 0, st.stream>>>(kin, ain, cap_in, la, kout, aout, cap_out, overflow);
    SPQ_HIP(hipGetLastError());
}

// MSB-first segments of a run of bit positions given in LSD order.
static bool segments_of(const std::vector<std::pair<int, int>>& lsd, BitSegs& sg) {
    sg = BitSegs{};
    for (size_t i = lsd.size(); i-- > 0;) {
        const int w = lsd[i].first, p = lsd[i].second;
        if (sg.n > 0 && sg.word[sg.n - 1] == w && sg.shift[sg.n - 1] == p + 1) {
            --sg.shift[sg.n - 1];
            ++sg.width[sg.n - 1];
            continue;
        }
        if (sg.n == (uint32_t)kMaxSegs) return false;
        sg.word[sg.n] = (uint8_t)w;
        sg.shift[sg.n] = (uint8_t)p;
        sg.width[sg.n] = 1;
        ++sg.n;
    }
    return true;
}

// Retry level per ordering and size class: a walk repeats the same sorts
// every step, so an ordering whose buckets overflowed a chunk starts the next
// sort at the level that worked (kGiveUp: straight to the LSD passes).
static std::mutex g_local_mu;
static std::unordered_map<uint64_t, int> g_local_level;
constexpr int kGiveUp = -1;

void sort_path_stats(uint64_t* out, int reset) {
    for (int i = 0; i < kStatCount; ++i) out[i] = reset ? g_sort_stats[i].exchange(0) : g_sort_stats[i].load();
}

static uint64_t ordering_signature(const std::vector<std::pair<int, int>>& bits, uint32_t W, uint64_t n) {
    uint64_t h = 0x9E3779B97F4A7C15ull ^ W;
    int lg = 0;
    while ((uint64_t(1) << (lg + 1)) <= n && lg < 63) ++lg;
    auto mix = [&h](uint64_t v) { h = (h ^ v) * 0x100000001B3ull; h ^= h >> 29; };
    mix((uint64_t)lg);
    for (const auto& b : bits) mix((uint64_t)b.first << 8 | (uint64_t)b.second);
    return h;
}

// One attempt with a B-bit prefix: returns true when the sorted entries are
// installed, false when a bucket overflowed a chunk (the state is untouched).
// B equal to the number of ordering bits: the prefix passes sort the whole
// ordering and no chunk sort follows.  presorted: the state is already sorted
// by the B-bit prefix, so no prefix pass runs.
static bool local_attempt(State& st, const std::vector<std::pair<int, int>>& bits, uint32_t B, uint32_t h, bool presorted = false) {
    const uint64_t n = st.n;
    const uint32_t R = (uint32_t)bits.size() - B;
    const int Q = R > 64 ? 2 : 1;
    const std::vector<std::pair<int, int>> low(bits.begin(), bits.begin() + R), top(bits.begin() + R, bits.end());
    LocalArgs la{};
    if (!segments_of(low, la.rem) || !segments_of(top, la.pre)) return false;
    la.rem_bits = R;
    la.n = n;
    la.h = B ? h : 0;   // chunk = h entries + the bucket straddling its end
    const unsigned grid = B ? grid_for(n, la.h) : 1;

    // prefix passes: whole entries, ping-pong between two scratch sets
    const std::vector<Digit> plan = presorted ? std::vector<Digit>{} : digits_of(top, 8);
    const uint64_t* src_k = st.keys;
    const double2* src_a = st.amps;
    uint64_t src_cap = st.cap;
    uint64_t* setk[2] = {nullptr, nullptr};
    double2* seta[2] = {nullptr, nullptr};
    int cur = -1;   // set holding the current data (-1: the state's arrays)
    const uint32_t nb = grid_for(n, kPayloadTile);
    uint32_t* bh = nullptr;
    uint32_t* bh_rows = nullptr;
    uint64_t* offs = nullptr;
    if (!plan.empty()) {
        bh = dalloc<uint32_t>((size_t)256 * nb, st.stream);
        bh_rows = dalloc<uint32_t>((size_t)256 * nb, st.stream);
        offs = dalloc<uint64_t>((size_t)256 * nb, st.stream);
    }
    for (const Digit& dg : plan) {
        const int dst = cur == 0 ? 1 : 0;
        if (!setk[dst]) alloc_arrays(st, st.W, n, &setk[dst], &seta[dst]);
        {
            SPQ_PROF(st.stream, "radix_hist_kernel", n * 8.0);
            radix_hist_kernel<8, kPayloadBlock><<<nb, kPayloadBlock, 0, st.stream>>>(src_k + (uint64_t)dg.word * src_cap, n, dg, bh_rows, nb);
        }
        SPQ_HIP(hipGetLastError());
        {
            SPQ_PROF(st.stream, "transpose_u32_kernel", 256.0 * nb * 8.0);
            const dim3 tg(256 / 32, (nb + 31) / 32);
            transpose_u32_kernel<<<tg, kBlock, 0, st.stream>>>(bh_rows, nb, 256, bh);
        }
        SPQ_HIP(hipGetLastError());
        exclusive_scan_u32(st, bh, offs, (uint64_t)256 * nb, nullptr);
        SPQ_DISPATCH_W(st.W, launch_payload_pass, st, src_k, src_a, src_cap, dg, offs, nb, setk[dst], seta[dst], n);
        cur = dst;
        src_k = setk[dst];
        src_a = seta[dst];
        src_cap = n;
    }
    // local sort into the other set (or a fresh one)
    const bool passes_only = R == 0 && cur >= 0;
    const int out = passes_only ? cur : (cur == 0 ? 1 : 0);
    if (!setk[out]) alloc_arrays(st, st.W, n, &setk[out], &seta[out]);
    unsigned int* flag = B ? st.counter + 8 : nullptr;   // zero between kernels; cleared again below
    if (!passes_only)
        SPQ_DISPATCH_W(st.W, launch_local, st, src_k, src_a, src_cap, la, grid, Q, setk[out], seta[out], n, flag);
    bool ok = true;
    if (B && !passes_only) {
        unsigned int f = 0;
        read_back(st, flag, sizeof(f), &f);
        if (f) {
            SPQ_HIP(hipMemsetAsync(flag, 0, sizeof(unsigned int), st.stream));
            ok = false;
        }
    }
    if (bh) {
        dfree(bh, st.stream);
        dfree(bh_rows, st.stream);
        dfree(offs, st.stream);
    }
    if (ok) {
        install_arrays(st, st.W, setk[out], seta[out], n, n);
        setk[out] = nullptr;
        seta[out] = nullptr;
    }
    for (int i = 0; i < 2; ++i)
        if (setk[i]) {
            dfree(setk[i], st.stream);
            dfree(seta[i], st.stream);
        }
    return ok;
}

// The prefix + local path.  Level 0: B = the smallest multiple of 8 ordering
// bits that leaves about n / 2^B <= kLocalCap / 8 entries per bucket, chunk
// spacing h = 3/4 of a chunk (a chunk holds h entries plus the bucket that
// straddles its end, so buckets of up to kLocalCap / 4 always fit).  Each level
// after an overflow adds 8 prefix bits, up to 32 (the prefix is compared as one
// 32-bit word) or all ordering bits; past that it lowers the spacing to 1/2,
// then 1/4 of a chunk (buckets up to 1/2, then 3/4 of a chunk always fit).
// An ordering of at most 32 bits that the prefix covers, or leaves at most 8
// bits of, is sorted by the prefix passes alone: at most four whole-entry
// passes cost less than index passes plus the random gather, and one pass
// less than a chunk sort.
static bool level_shape(uint32_t base, uint32_t max_b, int level, uint32_t& B, uint32_t& h) {
    B = base;
    h = kLocalCap * 3 / 4;
    for (int l = 0; l < level; ++l) {
        if (B < max_b)
            B = std::min<uint32_t>(B + 8, max_b);
        else if (h > (uint32_t)kLocalCap / 4)
            h = h == (uint32_t)kLocalCap * 3 / 4 ? kLocalCap / 2 : kLocalCap / 4;
        else
            return false;
    }
    return true;
}

static bool sort_local(State& st, const std::vector<std::pair<int, int>>& bits, uint32_t presorted) {
    const uint64_t n = st.n;
    const uint32_t T = (uint32_t)bits.size();
    if (st.W > (uint32_t)kLocalMaxW || n >= (1ull << 32)) return false;
    if (n > (uint64_t)kLocalCap && presorted > 0) {
        // Already sorted by a prefix at least as long as the passes would
        // make (up to 64 bits of it serve as the bucket prefix): the chunk
        // sort alone.  Buckets too large for it fall through to the passes.
        uint32_t need = 0;
        while ((n >> need) > (uint64_t)kLocalCap / 8) ++need;
        const uint32_t B = std::min<uint32_t>(presorted, 64);
        if (B >= need && B < T && T - B <= 128) {
            if (local_attempt(st, bits, B, kLocalCap * 3 / 4, true)) {
                g_sort_stats[kStatLocal] += 1;
                g_sort_stats[kStatLocalEntries] += n;
                g_sort_stats[kStatPresorted] += 1;
                return true;
            }
            g_sort_stats[kStatOverflow] += 1;
        }
    }
    const uint64_t sig = ordering_signature(bits, st.W, n);
    int level = 0;
    {
        std::lock_guard<std::mutex> lock(g_local_mu);
        auto it = g_local_level.find(sig);
        if (it != g_local_level.end()) level = it->second;
    }
    if (level == kGiveUp) return false;
    uint32_t base = 0;
    if (n > (uint64_t)kLocalCap) {
        uint32_t need = 0;
        while ((n >> need) > (uint64_t)kLocalCap / 8) ++need;
        base = (need + 7) / 8 * 8;
    }
    const uint32_t max_b = std::min<uint32_t>(T, 32);
    base = std::min(base, max_b);
    if (base > 0 && T <= 32 && T - base <= 8) base = T;   // one more whole-entry pass costs less than a chunk sort
    for (;;) {
        uint32_t B = 0, h = 0;
        bool ok = level_shape(base, max_b, level, B, h) && T - B <= 128 && (B > 0) == (n > (uint64_t)kLocalCap);
        if (ok) {
            if (local_attempt(st, bits, B, h)) {
                g_sort_stats[kStatLocal] += 1;
                g_sort_stats[kStatLocalEntries] += n;
                return true;
            }
            g_sort_stats[kStatOverflow] += 1;
        }
        // B == 0: the whole state is one chunk and never overflows; nothing to refine
        level = ok && B > 0 ? level + 1 : kGiveUp;
        {
            std::lock_guard<std::mutex> lock(g_local_mu);
            g_local_level[sig] = level;
        }
        if (level == kGiveUp) return false;
    }
}

}  // namespace sparq
