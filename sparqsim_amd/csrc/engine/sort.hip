// LSD radix sort of the state by packed key (32-bit index payload), with
// per-wave ballot ranking and an LDS-staged, digit-grouped write-out.
//
// Per digit pass (N entries): histogram reads the 8-byte key word; the scatter
// reads key word + index and writes both (2*(8+4) bytes); the final gather
// moves W key words and the 16-byte amplitude.  Only bits that vary across
// the state are sorted (an OR/AND reduction finds them first).
#include <algorithm>

#include "blockscan.hpp"
#include "engine.hpp"
#include "keyops.hpp"

namespace sparq {

constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;   // 4096 entries per workgroup
constexpr int kRadix = 256;                      // digits of up to 8 bits
constexpr uint32_t kInvalidDigit = 0x100;

// Lanes of the wave holding the same 9-bit value as the calling lane.
__device__ __forceinline__ uint64_t match_any(uint32_t d, int nbits) {
    uint64_t m = ~0ull;
    for (int b = 0; b < nbits; ++b) {
        uint64_t bal = __ballot((d >> b) & 1);
        m &= ((d >> b) & 1) ? bal : ~bal;
    }
    uint64_t bal = __ballot((d >> 8) & 1);
    m &= ((d >> 8) & 1) ? bal : ~bal;
    return m;
}

// ---------------------------------------------------------------- OR / AND
template <int W>
__global__ void __launch_bounds__(kBlock) or_and_kernel(const uint64_t* __restrict__ keys, uint64_t cap, uint64_t n,
                                                   unsigned long long* __restrict__ out /* [W] or, [W] and */) {
    uint64_t o[W], a[W];
#pragma unroll
    for (int w = 0; w < W; ++w) { o[w] = 0; a[w] = ~0ull; }
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            uint64_t v = keys[(uint64_t)w * cap + i];
            o[w] |= v;
            a[w] &= v;
        }
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            o[w] |= __shfl_xor(o[w], off, 64);
            a[w] &= __shfl_xor(a[w], off, 64);
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            atomicOr(&out[w], (unsigned long long)o[w]);
            atomicAnd(&out[W + w], (unsigned long long)a[w]);
        }
    }
}

template <int W>
static void launch_or_and(State& st, unsigned long long* out) {
    unsigned g = std::min<unsigned>(grid_for(st.n, kBlock * 4), 2048);
    {
        SPQ_PROF(st.stream, "or_and_kernel", st.n * 8.0 * W);
        or_and_kernel<W><<<g, kBlock, 0, st.stream>>>(st.keys, st.cap, st.n, out);
    }
    SPQ_HIP(hipGetLastError());
}

void key_or_and(State& st, std::vector<uint64_t>& or_mask, std::vector<uint64_t>& and_mask) {
    or_mask.assign(st.W, 0);
    and_mask.assign(st.W, ~0ull);
    if (st.n == 0) return;
    unsigned long long* d = dalloc<unsigned long long>(2 * st.W, st.stream);
    std::vector<unsigned long long> init(2 * st.W);
    for (uint32_t w = 0; w < st.W; ++w) { init[w] = 0; init[st.W + w] = ~0ull; }
    SPQ_HIP(hipMemcpyAsync(d, init.data(), init.size() * 8, hipMemcpyHostToDevice, st.stream));
    SPQ_DISPATCH_W(st.W, launch_or_and, st, d);
    std::vector<unsigned long long> h(2 * st.W);
    read_back(st, d, h.size() * 8, h.data());
    dfree(d, st.stream);
    for (uint32_t w = 0; w < st.W; ++w) { or_mask[w] = h[w]; and_mask[w] = h[st.W + w]; }
}

std::vector<Digit> make_plan(State& st, const sparq_field_t* last, int digit_bits) {
    std::vector<uint64_t> om, am;
    key_or_and(st, om, am);
    std::vector<uint64_t> vary(st.W);
    for (uint32_t w = 0; w < st.W; ++w) vary[w] = om[w] ^ am[w];
    // Segments of eligible bits in LSD order: (word, lo, hi).
    struct Seg { int word, lo, hi; };
    std::vector<Seg> segs;
    if (last) segs.push_back({last->word, last->shift, last->shift + last->width});
    for (int w = (int)st.W - 1; w >= 0; --w) {
        if (last && last->word == w) {
            if (last->shift > 0) segs.push_back({w, 0, last->shift});
            if (last->shift + last->width < 64) segs.push_back({w, last->shift + last->width, 64});
        } else {
            segs.push_back({w, 0, 64});
        }
    }
    std::vector<Digit> plan;
    for (const Seg& s : segs) {
        int p = s.lo;
        while (p < s.hi) {
            if (!((vary[s.word] >> p) & 1)) { ++p; continue; }
            int b = std::min(digit_bits, s.hi - p);
            plan.push_back(Digit{(uint8_t)s.word, (uint8_t)p, (uint8_t)b, 0});
            p += b;
        }
    }
    return plan;
}

// ---------------------------------------------------------------- passes
__device__ __forceinline__ uint32_t digit_of(uint64_t k, Digit dg) {
    return (uint32_t)((k >> dg.shift) & ((1ull << dg.bits) - 1));
}

// Block histogram: block_hist[d * nb + b].
__global__ void __launch_bounds__(kBlock) radix_hist_kernel(const uint64_t* __restrict__ kw, uint64_t n, Digit dg,
                                                       uint32_t* __restrict__ block_hist, uint32_t nb) {
    __shared__ uint32_t hist[kRadix];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t tile = (uint64_t)blockIdx.x * kSortTile;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int it = 0; it < kSortItems; ++it) {
        uint64_t e = tile + (uint64_t)wave * (kSortTile / 4) + it * 64 + lane;
        uint32_t d = e < n ? digit_of(kw[e], dg) : kInvalidDigit;
        uint64_t peers = match_any(d, dg.bits);
        int leader = __ffsll((unsigned long long)peers) - 1;
        if (lane == leader && d != kInvalidDigit) atomicAdd(&hist[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    block_hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = hist[threadIdx.x];
}

// Stable scatter of one digit pass.  idx_in == nullptr means identity order.
__global__ void __launch_bounds__(kBlock) radix_scatter_kernel(const uint64_t* __restrict__ kw_in, const uint32_t* __restrict__ idx_in,
                                                          uint64_t n, Digit dg, const uint64_t* __restrict__ offs, uint32_t nb,
                                                          uint64_t* __restrict__ kw_out, uint32_t* __restrict__ idx_out,
                                                          int write_keys) {
    __shared__ uint32_t wcnt[4][kRadix];       // per-wave running digit counts
    __shared__ uint32_t dstart[kRadix];        // first tile position of each digit
    __shared__ uint64_t goff[kRadix];          // global position of tile position 0 of each digit
    __shared__ uint64_t skey[kSortTile];
    __shared__ uint32_t sidx[kSortTile];
    __shared__ uint32_t scan_lds[kBlock / 64];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int w = 0; w < 4; ++w) wcnt[w][threadIdx.x] = 0;
    __syncthreads();

    const uint64_t tile = (uint64_t)blockIdx.x * kSortTile;
    uint64_t key[kSortItems];
    uint32_t idx[kSortItems];
    uint32_t rank[kSortItems];
    uint32_t dig[kSortItems];
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        uint64_t e = tile + (uint64_t)wave * (kSortTile / 4) + it * 64 + lane;
        bool valid = e < n;
        key[it] = valid ? kw_in[e] : 0;
        idx[it] = valid ? (idx_in ? idx_in[e] : (uint32_t)e) : 0;
        uint32_t d = valid ? digit_of(key[it], dg) : kInvalidDigit;
        dig[it] = d;
        uint64_t peers = match_any(d, dg.bits);
        uint32_t before = (uint32_t)__popcll(peers & lanemask_lt());
        uint32_t base = valid ? wcnt[wave][d & (kRadix - 1)] : 0;
        rank[it] = base + before;
        if (valid && before == 0) wcnt[wave][d & (kRadix - 1)] = base + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // Digit totals and their exclusive scan over digits; per-wave offsets.
    uint32_t d = threadIdx.x;
    uint32_t c0 = wcnt[0][d], c1 = wcnt[1][d], c2 = wcnt[2][d], c3 = wcnt[3][d];
    uint32_t tot;
    uint32_t start = block_excl_scan<uint32_t>(c0 + c1 + c2 + c3, scan_lds, tot);
    dstart[d] = start;
    wcnt[0][d] = start;
    wcnt[1][d] = start + c0;
    wcnt[2][d] = start + c0 + c1;
    wcnt[3][d] = start + c0 + c1 + c2;
    goff[d] = offs[(uint64_t)d * nb + blockIdx.x];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        if (dig[it] != kInvalidDigit) {
            uint32_t pos = wcnt[wave][dig[it]] + rank[it];
            skey[pos] = key[it];
            sidx[pos] = idx[it];
        }
    }
    __syncthreads();
    const uint32_t count = (uint32_t)min<uint64_t>((uint64_t)kSortTile, n - tile);
    for (uint32_t p = threadIdx.x; p < count; p += kBlock) {
        uint64_t k = skey[p];
        uint32_t dd = digit_of(k, dg);
        uint64_t dst = goff[dd] + (p - dstart[dd]);
        if (write_keys) kw_out[dst] = k;
        idx_out[dst] = sidx[p];
    }
}

__global__ void __launch_bounds__(kBlock) gather_word_kernel(const uint64_t* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                                                        uint64_t* __restrict__ dst) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

template <int W>
__global__ void __launch_bounds__(kBlock) gather_state_kernel(const uint64_t* __restrict__ kin, const double2* __restrict__ ain, uint64_t cap_in,
                                                         const uint32_t* __restrict__ idx, uint64_t n, uint64_t* __restrict__ kout,
                                                         double2* __restrict__ aout, uint64_t cap_out) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t s = idx[i];
#pragma unroll
    for (int w = 0; w < W; ++w) kout[(uint64_t)w * cap_out + i] = kin[(uint64_t)w * cap_in + s];
    aout[i] = ain[s];
}

template <int W>
static void launch_gather_state(State& st, const uint32_t* idx, uint64_t* kout, double2* aout, uint64_t cap_out) {
    {
        SPQ_PROF(st.stream, "gather_state_kernel", st.n * (4.0 + 2 * (8.0 * W + 16)));
        gather_state_kernel<W><<<grid_for(st.n, kBlock), kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, idx, st.n, kout, aout, cap_out);
    }
    SPQ_HIP(hipGetLastError());
}

void sort_by_plan(State& st, const std::vector<Digit>& plan) {
    const uint64_t n = st.n;
    if (plan.empty() || n <= 1) return;
    SPQ_REQUIRE(n < (1ull << 32), "sort supports fewer than 2^32 entries");
    const uint32_t nb = grid_for(n, kSortTile);
    uint64_t* kwA = dalloc<uint64_t>(n, st.stream);
    uint64_t* kwB = dalloc<uint64_t>(n, st.stream);
    uint32_t* idxA = dalloc<uint32_t>(n, st.stream);
    uint32_t* idxB = dalloc<uint32_t>(n, st.stream);
    uint32_t* bh = dalloc<uint32_t>((size_t)kRadix * nb, st.stream);
    uint64_t* offs = dalloc<uint64_t>((size_t)kRadix * nb, st.stream);

    bool identity = true;
    int cur_word = -1;
    const uint64_t* src = nullptr;
    for (size_t p = 0; p < plan.size(); ++p) {
        const Digit dg = plan[p];
        if (dg.word != cur_word) {
            if (identity) {
                src = st.keys + (uint64_t)dg.word * st.cap;
            } else {
                {
                    SPQ_PROF(st.stream, "gather_word_kernel", n * 20.0);
                    gather_word_kernel<<<grid_for(n, kBlock), kBlock, 0, st.stream>>>(st.keys + (uint64_t)dg.word * st.cap, idxA, n, kwA);
                }
                SPQ_HIP(hipGetLastError());
                src = kwA;
            }
            cur_word = dg.word;
        }
        {
            SPQ_PROF(st.stream, "radix_hist_kernel", n * 8.0);
            radix_hist_kernel<<<nb, kBlock, 0, st.stream>>>(src, n, dg, bh, nb);
        }
        SPQ_HIP(hipGetLastError());
        exclusive_scan_u32(st, bh, offs, (uint64_t)kRadix * nb, nullptr);
        // The key word of the last pass of a word is not needed afterwards.
        bool need_keys = (p + 1 < plan.size()) && plan[p + 1].word == dg.word;
        {
            SPQ_PROF(st.stream, "radix_scatter_kernel", n * (8.0 + (identity ? 0.0 : 4.0) + 4.0 + (need_keys ? 8.0 : 0.0)));
            radix_scatter_kernel<<<nb, kBlock, 0, st.stream>>>(src, identity ? nullptr : idxA, n, dg, offs, nb, kwB, idxB,
                                                            need_keys ? 1 : 0);
        }
        SPQ_HIP(hipGetLastError());
        std::swap(kwA, kwB);
        std::swap(idxA, idxB);
        identity = false;
        src = kwA;
    }
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
    dfree(offs, st.stream);
}

}  // namespace sparq
