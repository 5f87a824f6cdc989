// Branching operators: Hadamard_Int / Hadamard_Int_Full (fan-out 2^n per
// group) and the two-branch pair operators (Hadamard_Bool, CondRot_*,
// Rot_Bool general).  Groups come from one LSD radix sort that places the
// target bits least significant; every group then owns a contiguous output
// block, so the expansion is a pure streaming write and the Walsh-Hadamard
// butterflies run tile-by-tile in LDS.
//
// Bytes per output entry of the expansion: 8W+16 written (+16 read/written per
// WHT pass); pair ops: (8W+16) read + (8W+16) written per output entry.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "blockscan.hpp"
#include "engine.hpp"
#include "keyops.hpp"

namespace sparq {

// kPass: a group of a conditioned group unitary whose entries fail the
// conditions (all members agree: the conditions read no target bit); its
// members pass through unchanged.
enum GroupClass : uint8_t { kDirect = 0, kSparseSum = 1, kWht = 2, kDense = 3, kPass = 4 };

static bool g_dense_transform = true;   // the dense-domain Hadamard path (below) where it applies
static bool g_canonical = true;         // transform outputs in canonical (target-major) order where that applies
void set_dense_transform(bool on) { g_dense_transform = on; }
void set_canonical_output(bool on) { g_canonical = on; }

// ---------------------------------------------------------------- groups
template <int W>
__global__ void __launch_bounds__(kBlock) head_flag_kernel(const uint64_t* __restrict__ keys, uint64_t cap, uint64_t n, sparq_field_t low,
                                                      uint32_t* __restrict__ head) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (i == 0) { head[0] = 1; return; }
    const uint64_t lm = fmask(low.width) << low.shift;
    bool diff = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        uint64_t m = (w == low.word) ? ~lm : ~0ull;
        diff = diff || ((keys[(uint64_t)w * cap + i] & m) != (keys[(uint64_t)w * cap + i - 1] & m));
    }
    head[i] = diff ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) group_start_kernel(const uint32_t* __restrict__ head, const uint64_t* __restrict__ pos, uint64_t n,
                                                        uint64_t* __restrict__ start, const uint64_t* __restrict__ G) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n && head[i]) start[pos[i]] = i;
    if (i == 0) start[G[0]] = n;
}

// Class of each group (grid covers an upper bound of the group count; the
// actual count is read from device memory).  counts[1] receives the number of
// WHT groups so the butterfly passes can be skipped when there are none.
// Grid-stride over the groups; the WHT-group count is reduced per block and
// added with one atomic per block.
__global__ void __launch_bounds__(kBlock) classify_kernel(const uint64_t* __restrict__ start, const uint64_t* __restrict__ G_dev, uint64_t F,
                                                     uint32_t nd, int full, uint8_t* __restrict__ cls,
                                                     unsigned long long* __restrict__ n_wht) {
    __shared__ uint32_t lds[kBlock / 64];
    const uint64_t G = G_dev[0];
    uint32_t wht = 0;
    for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < G; g += (uint64_t)gridDim.x * kBlock) {
        const uint64_t cnt = start[g + 1] - start[g];
        uint8_t c;
        if (full == 2) c = kDense;                                         // rot.cpp:395-455
        else if (!full) c = (cnt == 1) ? kDirect : kWht;                   // hadamard.cpp:26-131
        else c = (cnt == F) ? kWht : (cnt + 1 <= nd ? kSparseSum : kWht);  // hadamard.cpp:220-256 (few_threshold = n-1)
        cls[g] = c;
        wht += c == kWht;
    }
    wht = wave_sum(wht);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = wht;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
        if (t) atomicAdd(n_wht, (unsigned long long)t);
    }
}

// Groups whose head fails the conditions pass through (kPass).
template <int W>
__global__ void __launch_bounds__(kBlock) pass_classify_kernel(const uint64_t* __restrict__ keys, uint64_t cap,
                                                          const uint64_t* __restrict__ start, const uint64_t* __restrict__ G_dev,
                                                          const CondList cl, uint8_t* __restrict__ cls) {
    const uint64_t G = G_dev[0];
    for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < G; g += (uint64_t)gridDim.x * kBlock) {
        uint64_t k[W];
        load_key<W>(keys, cap, start[g], k);
        bool ok = true;
        for (uint32_t c = 0; c < cl.n; ++c) ok = ok && cond_ok<W>(k, cl.c[c]);
        if (!ok) cls[g] = kPass;
    }
}

// Outputs of each group: its members when it passes through, else 2^nd.
__global__ void __launch_bounds__(kBlock) group_out_size_kernel(const uint64_t* __restrict__ start, const uint8_t* __restrict__ cls,
                                                           uint64_t G, uint32_t nd, uint32_t* __restrict__ size) {
    const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= G) return;
    size[g] = cls[g] == kPass ? (uint32_t)(start[g + 1] - start[g]) : (1u << nd);
}

constexpr int kExpandItems = 4;   // outputs per thread (strided by the block size)

// Largest output (bytes of keys + amplitudes) a learned capacity may allocate.
static double learned_budget_bytes() {
    static double budget = [] {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || total_b == 0) return 4.0e9;
        return (double)total_b / 8.0;
    }();
    return budget;
}

// Output entry o = g*F + j: key of the group head with the low target bits = j.
// DIRECT and SPARSE_SUM amplitudes are produced here; WHT blocks start at 0.
// The group count G is read from the device; the output holds G_cap groups
// and nothing is written when G > G_cap (a too small predicted capacity).
// With group offsets `goff` (a conditioned group unitary) group g's outputs
// start at goff[g] instead: 2^nd of them, or its members unchanged (kPass).
// With tG = G (no WHT groups, no offsets) output o holds j = o / G of group
// g = o mod G: the transposed order, canonical when every others bit lies
// below the target bits.
template <int W>
__global__ void __launch_bounds__(kBlock) expand_init_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps, uint64_t cap,
                                                        const uint64_t* __restrict__ start, const uint8_t* __restrict__ cls, uint32_t nd,
                                                        sparq_field_t low, double extra, const uint64_t* __restrict__ G_dev, uint64_t G_cap,
                                                        uint64_t* __restrict__ kout, double2* __restrict__ aout, uint64_t cap_out,
                                                        const double2* __restrict__ dense, const uint64_t* __restrict__ goff,
                                                        uint64_t tG) {
    const uint64_t G = G_dev[0];
    if (G > G_cap) return;
    const uint64_t M = G << nd;
#pragma unroll
    for (int it = 0; it < kExpandItems; ++it) {
    const uint64_t o = ((uint64_t)blockIdx.x * kExpandItems + it) * kBlock + threadIdx.x;
    if (o >= M) return;
    uint64_t g, j;
    if (tG) {
        if (M <= 0xffffffffull) {
            g = (uint32_t)o % (uint32_t)tG;
            j = (uint32_t)o / (uint32_t)tG;
        } else {
            g = o % tG;
            j = o / tG;
        }
    } else {
        g = o >> nd;
        j = o & ((uint64_t(1) << nd) - 1);
    }
    const uint64_t h = start[g];
    const uint8_t c = cls[g];
    if (c == kPass) {   // members j, j + 2^nd, ... to their place in the group's output
        const uint64_t e = start[g + 1];
        for (uint64_t m = h + j; m < e; m += uint64_t(1) << nd) {
            uint64_t km[W];
            load_key<W>(keys, cap, m, km);
            store_key<W>(kout, cap_out, goff[g] + (m - h), km);
            aout[goff[g] + (m - h)] = amps[m];
        }
        continue;
    }
    const uint64_t out = goff ? goff[g] + j : o;
    uint64_t k[W];
    load_key<W>(keys, cap, h, k);
    const uint64_t vlow = fget<W>(k, low);
    fput<W>(k, low, j);
    store_key<W>(kout, cap_out, out, k);
    double2 r = make_double2(0.0, 0.0);
    if (c == kDirect) {
        // amp * 2^{-n/2}, then * -1 on odd parity (hadamard.cpp:30-39)
        double2 t = cscale(amps[h], extra);
        r = (__popcll(j & vlow) & 1) ? make_double2(t.x * -1.0, t.y * -1.0) : t;
    } else if (c == kSparseSum) {
        // sum over members in ascending value of (amp * (+-1)) * 2^{-n/2} (hadamard.cpp:269-304)
        const uint64_t e = start[g + 1];
        for (uint64_t m = h; m < e; ++m) {
            uint64_t km[W];
            load_key<W>(keys, cap, m, km);
            uint64_t v = fget<W>(km, low);
            double sg = (__popcll(j & v) & 1) ? -1. : 1.;
            r = cadd(r, cscale(cscale(amps[m], sg), extra));
        }
    } else if (c == kDense) {
        // row j of the dense matrix against the members, ascending column (rot.cpp:351-358)
        const uint64_t e = start[g + 1];
        const uint64_t F = uint64_t(1) << nd;
        for (uint64_t m = h; m < e; ++m) {
            uint64_t km[W];
            load_key<W>(keys, cap, m, km);
            r = cadd(r, cmul(dense[j * F + fget<W>(km, low)], amps[m]));
        }
    }
    aout[out] = r;
    }
}

// The DIRECT case of a state with one entry (H on a basis state): pure
// streaming writes of F = 2^nd outputs, 8 per thread, no per-output loads.
template <int W>
__global__ void __launch_bounds__(kBlock) expand_single_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps,
                                                          uint64_t cap, sparq_field_t low, uint32_t nd, double extra,
                                                          uint64_t* __restrict__ kout, double2* __restrict__ aout, uint64_t cap_out,
                                                          double eps, uint64_t* __restrict__ pruned) {
    uint64_t k[W];
    load_key<W>(keys, cap, 0, k);
    const uint64_t vlow = fget<W>(k, low);
    const double2 t = cscale(amps[0], extra);                   // hadamard.cpp:30-39
    const double2 tn = make_double2(t.x * -1.0, t.y * -1.0);
    const uint64_t F = uint64_t(1) << nd;
    // every output is t or -t: ClearZero keeps all of them or none
    if (blockIdx.x == 0 && threadIdx.x == 0) pruned[0] = (cabs2(t) < eps) ? 1 : 0;
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const uint64_t o = ((uint64_t)blockIdx.x * 8 + it) * kBlock + threadIdx.x;
        if (o >= F) return;
        uint64_t ko[W];
#pragma unroll
        for (int w = 0; w < W; ++w) ko[w] = k[w];
        fput<W>(ko, low, o);
        store_key<W>(kout, cap_out, o, ko);
        const bool neg = __popcll(o & vlow) & 1;
        aout[o] = make_double2(neg ? tn.x : t.x, neg ? tn.y : t.y);   // per component: no indexed pair
    }
}

// Place the members of WHT groups at their slot of the zeroed block.
template <int W>
__global__ void __launch_bounds__(kBlock) expand_scatter_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps, uint64_t cap,
                                                           uint64_t n, const uint32_t* __restrict__ head, const uint64_t* __restrict__ pos,
                                                           const uint8_t* __restrict__ cls, uint32_t nd, sparq_field_t low,
                                                           const uint64_t* __restrict__ G_dev, uint64_t G_cap, double2* __restrict__ aout) {
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // the loads that do not depend on the group go out together (all in bounds
    // for i < n); only the group's class waits for the group id
    const uint64_t G = G_dev[0];
    const uint64_t p = head ? pos[i] : 0;
    const uint32_t h = head ? head[i] : 1;
    uint64_t k[W];
    load_key<W>(keys, cap, i, k);
    const double2 a = amps[i];
    // The grouping sort is stable, so equal keys (a state uploaded with
    // duplicates) are neighbours: only the last of them is placed, so the
    // slot's value does not depend on the order of the writes.
    const uint64_t inext = i + 1 < n ? i + 1 : i;
    const uint64_t wnext = keys[(uint64_t)low.word * cap + inext];
    const uint32_t hnext = head ? head[inext] : 0u;
    if (G > G_cap) return;
    const uint64_t vlow = fget<W>(k, low);
    if (inext != i && hnext == 0 && ((wnext >> low.shift) & fmask(low.width)) == vlow) return;
    // group id of entry i = inclusive count of heads up to i, minus one
    const uint64_t g = p + h - 1;
    if (cls[g] != kWht) return;
    aout[(g << nd) + vlow] = a;
}

// Walsh-Hadamard butterflies.  Stage q pairs indices i, i + 2^q of a group's
// block and maps (a0, a1) -> ((a0+a1)*s, (a0-a1)*s), stages in ascending order
// (hadamard.cpp:44-60, 118-130); every element sees the same operation
// sequence as in the reference, so the results are bit-identical.
//
// A tile of up to 2^11 complex doubles (32 KiB) sits in LDS; each thread
// applies up to 3 consecutive stages to 8 elements in registers per round.
constexpr int kWhtTileLog = 11;

// LDS index of tile element e: one pad slot per 16 elements, so that the
// 8-element register groups of neighbouring lanes (stride 8 or 64 elements)
// fall in different banks.
__host__ __device__ constexpr uint32_t lpad(uint32_t e) { return e + (e >> 4); }

// Stages [s0, s0+cnt) of tile bits, applied in rounds of <= 3 stages.  Groups
// of 2^k elements differing only in the round's stage bits are handled by one
// thread; `base_index` is the tile's first index in the output array (for the
// class lookup when a tile holds several groups).
__device__ __forceinline__ void lds_wht_rounds(double2* buf, uint32_t T, uint32_t s0, uint32_t cnt, bool check_class,
                                               uint64_t base_index, uint32_t nd, const uint8_t* __restrict__ cls) {
    for (uint32_t done = 0; done < cnt;) {
        const uint32_t k = min(3u, cnt - done);
        const uint32_t sb = s0 + done;
        const uint32_t ngroups = T >> k;
        for (uint32_t q = threadIdx.x; q < ngroups; q += kBlock) {
            const uint32_t lowq = q & ((1u << sb) - 1);
            const uint32_t first = ((q >> sb) << (sb + k)) | lowq;
            if (check_class && cls[(base_index + first) >> nd] != 2) continue;
            double2 v[8];
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                if (j < (1u << k)) v[j] = buf[lpad(first | (j << sb))];
#pragma unroll
            for (uint32_t r = 0; r < 3; ++r) {
                if (r >= k) break;
#pragma unroll
                for (uint32_t j = 0; j < 8; ++j) {
                    if (j >= (1u << k) || (j >> r) & 1) continue;
                    const uint32_t j1 = j | (1u << r);
                    double2 a0 = v[j], a1 = v[j1];
                    v[j] = cscale(cadd(a0, a1), kSqrt2Inv);
                    v[j1] = cscale(csub(a0, a1), kSqrt2Inv);
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                if (j < (1u << k)) buf[lpad(first | (j << sb))] = v[j];
        }
        __syncthreads();
        done += k;
    }
}

// First pass: stages [0, min(nd, 11)) on contiguous tiles of 2^11 elements of the
// output array (a tile may hold many small groups; non-WHT groups are skipped).
// G_dev = {group count, WHT group count} on the device (tiles past G << nd do
// nothing), or null when all G_cap groups are WHT groups.  The tile's loads are
// issued before the device-side counts are checked: every tile lies inside the
// G_cap << nd allocated elements, and a tile that fails the check discards them.
__global__ void __launch_bounds__(kBlock) wht_first_pass_kernel(double2* __restrict__ a, const uint64_t* __restrict__ G_dev, uint64_t G_cap,
                                                           uint32_t nd, uint32_t cnt, const uint8_t* __restrict__ cls) {
    __shared__ double2 buf[lpad(1 << kWhtTileLog)];
    const uint64_t base = (uint64_t)blockIdx.x << kWhtTileLog;
    const uint64_t M_cap = G_cap << nd;
    const uint32_t T_cap = (uint32_t)min<uint64_t>(uint64_t(1) << kWhtTileLog, M_cap - base);
    // all loads of the tile are issued before the first LDS store; a partial
    // tile loads its element 0 in place of the missing ones (unconditional
    // loads keep v in registers)
    constexpr int kPer = (1 << kWhtTileLog) / kBlock;
    double2 v[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t e = r * kBlock + threadIdx.x;
        v[r] = a[base + (e < T_cap ? e : 0)];
    }
    uint32_t T = T_cap;
    if (G_dev) {
        const uint64_t G = G_dev[0];
        if (G > G_cap || G_dev[1] == 0) return;
        const uint64_t M = G << nd;
        if (base >= M) return;
        if (nd >= (uint32_t)kWhtTileLog && cls[base >> nd] != kWht) return;   // the tile lies in one non-WHT group
        T = (uint32_t)min<uint64_t>(uint64_t(1) << kWhtTileLog, M - base);
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t e = r * kBlock + threadIdx.x;
        if (e < T) buf[lpad(e)] = v[r];
    }
    __syncthreads();
    // a tile inside one group was checked above; smaller groups are checked per round
    lds_wht_rounds(buf, T, 0, cnt, G_dev != nullptr && nd < (uint32_t)kWhtTileLog, base, nd, cls);
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t e = r * kBlock + threadIdx.x;
        if (e < T) a[base + e] = buf[lpad(e)];
    }
}

// Later passes: stages [lo, lo+cnt) of one group, on tiles of 2^cnt x 2^c2
// elements (c2 contiguous low index bits, c2 <= lo) of up to 2^kWhtPassTileLog
// elements (64 KiB of LDS).
constexpr int kWhtPassTileLog = 11;
__global__ void __launch_bounds__(kBlock) wht_pass_kernel(double2* __restrict__ a, const uint64_t* __restrict__ G_dev, uint64_t G_cap,
                                                     uint32_t nd, uint32_t lo, uint32_t cnt, uint32_t c2, const uint8_t* __restrict__ cls) {
    __shared__ double2 buf[lpad(1 << kWhtPassTileLog)];
    const uint32_t T = 1u << (cnt + c2);
    // Workgroups are dispatched round-robin over the 8 XCDs, each with its own
    // L2.  With rows of 2^c2 < 8 elements (shorter than a 128-byte line)
    // neighbouring tiles share the lines of their rows, so consecutive tiles go
    // to the workgroups of one XCD (blockIdx.x, blockIdx.x + 8, ...): the 2^20
    // pass then reads its 16 MiB once instead of twice.  Longer rows keep the
    // plain order, which spreads the XCDs' accesses over the array.
    const uint64_t tile = c2 < 3 ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint64_t per_group = (uint64_t(1) << nd) >> (cnt + c2);
    const uint64_t g = tile / per_group;
    const uint64_t r = tile % per_group;
    // r enumerates (high bits above lo+cnt) x (low chunk index below lo, in units of 2^c2)
    const uint64_t nlowchunks = (uint64_t(1) << lo) >> c2;
    const uint64_t lowchunk = r % nlowchunks, high = r / nlowchunks;
    const uint64_t base = (g << nd) + (high << (lo + cnt)) + (lowchunk << c2);
    // element e of the tile: sub = e >> c2 (stage index bits), l = e & (2^c2-1);
    // the whole tile is loaded before its LDS stores, and before the group is
    // checked (g < G_cap: the tile lies inside the allocated elements)
    constexpr int kPer = (1 << kWhtPassTileLog) / kBlock;   // T <= kPer * kBlock
    {
        double2 v[kPer];
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint32_t e0 = r * kBlock + threadIdx.x;
            const uint32_t e = e0 < T ? e0 : 0;   // unconditional loads keep v in registers
            v[r] = a[base + ((uint64_t)(e >> c2) << lo) + (e & ((1u << c2) - 1))];
        }
        if (G_dev) {
            const uint64_t G = G_dev[0];
            if (G > G_cap || g >= G || cls[g] != kWht) return;
        }
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint32_t e = r * kBlock + threadIdx.x;
            if (e < T) buf[lpad(e)] = v[r];
        }
    }
    __syncthreads();
    lds_wht_rounds(buf, T, c2, cnt, false, 0, nd, cls);
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t e = r * kBlock + threadIdx.x;
        if (e < T) a[base + ((uint64_t)(e >> c2) << lo) + (e & ((1u << c2) - 1))] = buf[lpad(e)];
    }
}

__global__ void single_group_kernel(uint64_t* __restrict__ start, uint8_t* __restrict__ cls, uint64_t n, uint8_t c,
                                    uint64_t* __restrict__ counts) {
    start[0] = 0;
    start[1] = n;
    cls[0] = c;
    counts[0] = 1;
    counts[1] = (c == kWht) ? 1 : 0;
}

// counts = {G, WHT groups, valid entries, -}: valid entries = G << nd, or 0
// when G exceeds the capacity the output was sized for.
__global__ void sized_total_kernel(uint64_t* __restrict__ counts, uint64_t G_cap, uint32_t nd) {
    const uint64_t G = counts[0];
    counts[2] = (G > G_cap) ? 0 : (G << nd);
}

// Butterfly passes over G_cap blocks of 2^nd amplitudes (the WHT groups among
// the first counts[0] blocks; counts null: all G_cap blocks): the first pass covers min(nd, 11) stages on
// contiguous tiles; the rest are split evenly into passes of <= 9 stages on
// column tiles of one group (2^cnt rows of 2^c2 contiguous elements, up to
// 2^12 elements).
static void launch_wht(State& st, double2* a, const uint64_t* counts, uint64_t G_cap, uint32_t nd, const uint8_t* cls, double bytes) {
    const uint64_t M_cap = G_cap << nd;
    uint32_t lo = std::min<uint32_t>(nd, kWhtTileLog);
    {
        SPQ_PROF(st.stream, "wht_first_pass_kernel", bytes);
        wht_first_pass_kernel<<<grid_for(M_cap, 1u << kWhtTileLog), kBlock, 0, st.stream>>>(a, counts, G_cap, nd, lo, cls);
    }
    SPQ_HIP(hipGetLastError());
    while (lo < nd) {
        const uint32_t rem = nd - lo;
        const uint32_t passes = (rem + 8) / 9;
        const uint32_t cnt = (rem + passes - 1) / passes;
        const uint32_t c2 = std::min<uint32_t>(lo, kWhtPassTileLog - cnt);
        const uint64_t tiles = M_cap >> (cnt + c2);
        {
            SPQ_PROF(st.stream, "wht_pass_kernel", bytes);
            wht_pass_kernel<<<(unsigned)tiles, kBlock, 0, st.stream>>>(a, counts, G_cap, nd, lo, cnt, c2, cls);
        }
        SPQ_HIP(hipGetLastError());
        lo += cnt;
    }
}

// ---------------------------------------------------------------- dense-domain transform
// When the low target bits and the key bits varying outside them span a small
// domain of 2^T slots that the state fills to at least an eighth, the
// transform runs on a dense amplitude array instead of on groups found by a
// sort: slot = (others index << nd) | low target value, the others index
// packing the varying bits outside the low target bits in canonical order.
// The array is group-major like the grouped path's output, so its butterfly
// passes run unchanged, on the groups that hold an entry (absent groups stay
// zero and are skipped).  The kept entries are written in slot order, or in
// the transposed (canonical) order below when that applies and the output is
// not concentrated, keys rebuilt from the slot and the bits every key
// shares.  The scatter marks
// each slot it writes; a state with duplicate keys (fewer marks than
// entries) keeps the grouped path.
constexpr uint32_t kDenseMaxLog = 28;   // 2^28 slots: 4 GiB of amplitudes
constexpr uint32_t kTransposeMaxLog = 8;   // widest others index emitted in transposed (canonical) order
constexpr int kDenseTile = 4096;        // slots per workgroup of the count / emit kernels
constexpr int kDenseItems = kDenseTile / kBlock;

struct DenseMap {
    uint32_t nseg;                      // runs of the others bits, least significant first
    uint8_t word[kDenseMaxLog], shift[kDenseMaxLog], width[kDenseMaxLog];
    uint32_t nd;
    sparq_field_t low;
    uint64_t base[SPARQ_MAX_WORDS];     // the bits every key shares (others and low target bits cleared)
};

template <int W>
__device__ __forceinline__ uint64_t dense_others(const uint64_t (&k)[W], const DenseMap& dm) {
    uint64_t v = 0;
    uint32_t at = 0;
    for (uint32_t s = 0; s < dm.nseg; ++s) {
        v |= ((kword<W>(k, dm.word[s]) >> dm.shift[s]) & fmask(dm.width[s])) << at;
        at += dm.width[s];
    }
    return v;
}

// Entry i to its slot; its slot's mark and its group's class (plain byte
// stores: every writer of a word stores the same value).  Each XCD takes a
// contiguous range of entries (xcd_tile): in a state ordered by the target
// value its writes then stay in 1/8 of every group's row, a smaller set of
// pages and lines per XCD than with round-robin blocks.
template <int W>
__global__ void __launch_bounds__(kBlock) dense_scatter_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps,
                                                          uint64_t cap, uint64_t n, DenseMap dm, double2* __restrict__ A,
                                                          uint8_t* __restrict__ mark, uint8_t* __restrict__ cls) {
    const uint64_t i = (uint64_t)xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    uint64_t k[W];
    load_key<W>(keys, cap, i, k);
    const double2 a = amps[i];
    const uint64_t g = dense_others<W>(k, dm);
    const uint64_t slot = (g << dm.nd) | fget<W>(k, dm.low);
    A[slot] = a;
    mark[slot] = 1;
    cls[g] = kWht;
}

// Per tile of slots: kept slots (|a|^2 not below eps, as ClearZero), and the
// marks added to marks_total (one atomic per workgroup); slots of absent
// groups are skipped.
__global__ void __launch_bounds__(kBlock) dense_count_kernel(const double2* __restrict__ A, const uint8_t* __restrict__ mark,
                                                        const uint8_t* __restrict__ cls, uint32_t nd, double eps,
                                                        uint32_t* __restrict__ kept_out, unsigned long long* __restrict__ marks_total) {
    __shared__ uint32_t lds[2][kBlock / 64];
    const uint64_t tile = (uint64_t)blockIdx.x * kDenseTile;
    uint32_t c = 0, m = 0;
#pragma unroll
    for (int j = 0; j < kDenseItems; ++j) {
        const uint64_t s = tile + (uint64_t)j * kBlock + threadIdx.x;
        if (cls[s >> nd] != kWht) continue;
        c += !(cabs2(A[s]) < eps);
        m += mark[s];
    }
    c = wave_sum(c);
    m = wave_sum(m);
    if ((threadIdx.x & 63) == 0) {
        lds[0][threadIdx.x >> 6] = c;
        lds[1][threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tc = 0, tm = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            tc += lds[0][w];
            tm += lds[1][w];
        }
        kept_out[blockIdx.x] = tc;
        if (tm) atomicAdd(marks_total, (unsigned long long)tm);
    }
}

// Kept slots in slot order: key rebuilt from the slot, amplitude copied.
template <int W>
__global__ void __launch_bounds__(kBlock) dense_emit_kernel(const double2* __restrict__ A, const uint8_t* __restrict__ cls, DenseMap dm,
                                                       double eps, const uint64_t* __restrict__ offs, uint64_t* __restrict__ kout,
                                                       double2* __restrict__ aout, uint64_t cap_out) {
    constexpr int kWaves = kBlock / 64;
    __shared__ uint32_t row_wave[kDenseItems * kWaves];   // kept count per (item row, wave), then its exclusive prefix
    const uint64_t tile = (uint64_t)blockIdx.x * kDenseTile;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double2 v[kDenseItems];
    uint32_t flags = 0, lane_pre[kDenseItems];
#pragma unroll
    for (int j = 0; j < kDenseItems; ++j) {
        const uint64_t s = tile + (uint64_t)j * kBlock + threadIdx.x;
        const bool present = cls[s >> dm.nd] == kWht;
        v[j] = present ? A[s] : make_double2(0.0, 0.0);
        const bool keep = present && !(cabs2(v[j]) < eps);
        const uint64_t m = __ballot(keep);
        flags |= uint32_t(keep) << j;
        lane_pre[j] = (uint32_t)__popcll(m & lanemask_lt());
        if (lane == 0) row_wave[j * kWaves + wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < 64) {   // exclusive scan of the (row, wave) counts in slot order
        const uint32_t x = threadIdx.x < kDenseItems * kWaves ? row_wave[threadIdx.x] : 0;
        const uint32_t inc = wave_incl_scan(x);
        if (threadIdx.x < kDenseItems * kWaves) row_wave[threadIdx.x] = inc - x;
    }
    __syncthreads();
    const uint64_t before = offs[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kDenseItems; ++j) {
        if (!((flags >> j) & 1)) continue;
        const uint64_t s = tile + (uint64_t)j * kBlock + threadIdx.x;
        const uint64_t others = s >> dm.nd;
        uint64_t k[W];
#pragma unroll
        for (int w = 0; w < W; ++w) k[w] = dm.base[w];
        uint32_t at = 0;
        for (uint32_t q = 0; q < dm.nseg; ++q) {
            const uint64_t bits = (others >> at) & fmask(dm.width[q]);
            kset_word<W>(k, dm.word[q], kword<W>(k, dm.word[q]) | (bits << dm.shift[q]));
            at += dm.width[q];
        }
        fput<W>(k, dm.low, s & fmask(dm.nd));
        const uint64_t dst = before + row_wave[j * kWaves + wave] + lane_pre[j];
        store_key<W>(kout, cap_out, dst, k);
        aout[dst] = v[j];
    }
}

// ---- transposed (canonical) emit order
// A group-major array of Gp = 2^tg groups of 2^nd slots (group g at
// [g << nd, (g + 1) << nd)) read in the transposed order: position
// p = (j << tg) | g holds slot (g << nd) | j.  When every varying key bit
// outside the low target bits lies below them, this order is canonical.
// A tile of kTxTile positions covers every group at C = kTxTile >> tg
// consecutive j: its amplitudes come in as C-long row segments (coalesced)
// through LDS and leave in position order.  Groups g >= G, or whose class is
// not kWht when `cls` is given, are empty.
constexpr int kTxTile = 2048;
constexpr int kTxItems = kTxTile / kBlock;

// Kept slots counted in slot order (streaming), per tile of kDenseTile slots
// (`tile_kept`, optional) and per chunk of C = 2^lc slots of one group
// (`chunk_cnt`, [Gp][2^nd / C], optional): tile t of the transposed order
// holds chunk t of every group, so its count is a column sum of chunk_cnt.
// Totals of kept slots and of marks (optional) go to totals[0], totals[1].
// Groups g >= G, or whose class is not kWht when `cls` is given, are empty.
// Needs nd >= 12: a tile of slots lies in one group.
__global__ void __launch_bounds__(kBlock) chunk_count_kernel(const double2* __restrict__ A, const uint8_t* __restrict__ cls,
                                                        const uint8_t* __restrict__ mark, uint64_t G, uint32_t nd, uint32_t lc,
                                                        double eps, uint32_t* __restrict__ tile_kept,
                                                        uint32_t* __restrict__ chunk_cnt, unsigned long long* __restrict__ totals) {
    __shared__ uint32_t cc[kDenseTile / 8];
    __shared__ uint32_t red[2][kBlock / 64];
    const uint64_t s0 = (uint64_t)blockIdx.x * kDenseTile;
    const uint64_t g = s0 >> nd;
    const uint32_t nch = chunk_cnt ? (uint32_t)kDenseTile >> lc : 0;   // <= kDenseTile / 8: lc >= 3
    for (uint32_t x = threadIdx.x; x < nch; x += kBlock) cc[x] = 0;
    __syncthreads();
    const bool present = g < G && (!cls || cls[g] == kWht);
    uint32_t c = 0, m = 0;
    if (present) {
#pragma unroll
        for (int it = 0; it < kDenseItems; ++it) {
            const uint32_t off = (uint32_t)it * kBlock + threadIdx.x;
            const bool keep = !(cabs2(A[s0 + off]) < eps);
            if (chunk_cnt && keep) atomicAdd(&cc[off >> lc], 1u);
            c += keep;
            if (mark) m += mark[s0 + off];
        }
    }
    c = wave_sum(c);
    m = wave_sum(m);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = c;
        red[1][threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tc = 0, tm = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            tc += red[0][w];
            tm += red[1][w];
        }
        if (tile_kept) tile_kept[blockIdx.x] = tc;
        if (tc) atomicAdd(&totals[0], (unsigned long long)tc);
        if (mark && tm) atomicAdd(&totals[1], (unsigned long long)tm);
    }
    if (chunk_cnt) {
        const uint64_t T = (uint64_t(1) << nd) >> lc;
        const uint64_t t0 = (s0 & ((uint64_t(1) << nd) - 1)) >> lc;
        for (uint32_t x = threadIdx.x; x < nch; x += kBlock) chunk_cnt[g * T + t0 + x] = cc[x];
    }
}

// Tile t of the transposed order: the sum over the Gp groups of chunk t.
__global__ void __launch_bounds__(kBlock) chunk_column_kernel(const uint32_t* __restrict__ chunk_cnt, uint64_t Gp, uint64_t T,
                                                         uint32_t* __restrict__ tile_cnt) {
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= T) return;
    uint32_t sum = 0;
    for (uint64_t g = 0; g < Gp; ++g) sum += chunk_cnt[g * T + t];
    tile_cnt[t] = sum;
}

// Kept positions in position order.  The tile's amplitudes (and, with KEYS,
// the keys from K, group-major like A) come in as row segments through LDS;
// without KEYS the keys are rebuilt from the slot by the dense map.
constexpr uint32_t kTxMaxKeyWords = 2;   // LDS: 36 KiB of amplitudes + 18 KiB per key word
template <int W, bool KEYS>
__global__ void __launch_bounds__(kBlock) tx_emit_kernel(const double2* __restrict__ A, const uint8_t* __restrict__ cls,
                                                    const uint64_t* __restrict__ K, uint64_t cap_in, DenseMap dm, uint64_t G,
                                                    uint32_t tg, uint32_t nd, double eps, const uint64_t* __restrict__ offs,
                                                    uint64_t* __restrict__ kout, double2* __restrict__ aout, uint64_t cap_out) {
    constexpr int kWaves = kBlock / 64;
    constexpr int kPadded = kTxTile + 256;   // row g at g * (C + 1): C + 1 odd, so a column read spreads over the banks
    __shared__ double2 la[kPadded];
    __shared__ uint64_t lk[KEYS ? W : 1][KEYS ? kPadded : 1];
    __shared__ uint32_t row_wave[kTxItems * kWaves];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t lc = (uint32_t)__builtin_ctz(kTxTile) - tg;
    const uint32_t C = 1u << lc;
    // each XCD reads a contiguous range of every row (xcd_tile)
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint64_t j0 = (uint64_t)tile << lc;
    const uint64_t F = uint64_t(1) << nd;
#pragma unroll
    for (int r = 0; r < kTxItems; ++r) {   // row segments in
        const uint32_t q = (uint32_t)r * kBlock + threadIdx.x;
        const uint32_t g = q >> lc, jj = q & (C - 1);
        const bool ok = g < G && j0 + jj < F && (!cls || cls[g] == kWht);
        const uint64_t slot = ((uint64_t)g << nd) | (j0 + jj);
        la[g * (C + 1) + jj] = ok ? A[slot] : make_double2(0.0, 0.0);
        if (KEYS && ok)
#pragma unroll
            for (int w = 0; w < W; ++w) lk[w][g * (C + 1) + jj] = K[(uint64_t)w * cap_in + slot];
    }
    __syncthreads();
    double2 v[kTxItems];
    uint32_t flags = 0, lane_pre[kTxItems];
#pragma unroll
    for (int r = 0; r < kTxItems; ++r) {   // position order out
        const uint32_t x = (uint32_t)r * kBlock + threadIdx.x;
        const uint32_t g = x & ((1u << tg) - 1), jj = x >> tg;
        const bool present = g < G && j0 + jj < F && (!cls || cls[g] == kWht);
        v[r] = la[g * (C + 1) + jj];
        const bool keep = present && !(cabs2(v[r]) < eps);
        const uint64_t m = __ballot(keep);
        flags |= uint32_t(keep) << r;
        lane_pre[r] = (uint32_t)__popcll(m & lanemask_lt());
        if (lane == 0) row_wave[r * kWaves + wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < 64) {   // exclusive scan of the (row, wave) counts in position order
        const uint32_t x = threadIdx.x < kTxItems * kWaves ? row_wave[threadIdx.x] : 0;
        const uint32_t inc = wave_incl_scan(x);
        if (threadIdx.x < kTxItems * kWaves) row_wave[threadIdx.x] = inc - x;
    }
    __syncthreads();
    const uint64_t before = offs[tile];
#pragma unroll
    for (int r = 0; r < kTxItems; ++r) {
        if (!((flags >> r) & 1)) continue;
        const uint32_t x = (uint32_t)r * kBlock + threadIdx.x;
        const uint32_t g = x & ((1u << tg) - 1), jj = x >> tg;
        const uint64_t dst = before + row_wave[r * kWaves + wave] + lane_pre[r];
        uint64_t k[W];
        if (KEYS) {
#pragma unroll
            for (int w = 0; w < W; ++w) k[w] = lk[KEYS ? w : 0][KEYS ? g * (C + 1) + jj : 0];
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) k[w] = dm.base[w];
            uint32_t at = 0;
            for (uint32_t q = 0; q < dm.nseg; ++q) {
                const uint64_t bits = ((uint64_t)g >> at) & fmask(dm.width[q]);
                kset_word<W>(k, dm.word[q], kword<W>(k, dm.word[q]) | (bits << dm.shift[q]));
                at += dm.width[q];
            }
            fput<W>(k, dm.low, j0 + jj);
        }
        store_key<W>(kout, cap_out, dst, k);
        aout[dst] = v[r];
    }
}

// Tiles of the transposed order over Gp = 2^tg groups of 2^nd slots.
static uint64_t tx_tiles(uint32_t tg, uint32_t nd) {
    const uint64_t P = uint64_t(1) << (tg + nd);
    return (P + kTxTile - 1) / kTxTile;
}

// The state holds a group-major expansion of G groups of 2^nd entries:
// ClearZero it into the transposed order.
template <int W>
static void transposed_clear_zero(State& st, uint64_t G, uint32_t nd, double eps) {
    if constexpr (W > (int)kTxMaxKeyWords) {
        clear_zero(st, eps);
        return;
    }
    uint32_t tg = 0;
    while ((uint64_t(1) << tg) < G) ++tg;
    const uint32_t lc = (uint32_t)__builtin_ctz(kTxTile) - tg;
    const uint64_t tiles = tx_tiles(tg, nd);                 // = 2^nd / C
    const uint64_t slot_tiles = (G << nd) / kDenseTile;      // nd >= 12
    uint32_t* chunk = dalloc<uint32_t>(G * tiles, st.stream);
    uint32_t* kept = dalloc<uint32_t>(tiles, st.stream);
    uint64_t* offs = dalloc<uint64_t>(tiles + 3, st.stream);   // [tiles] total, [tiles + 1, + 3) atomic totals
    SPQ_HIP(hipMemsetAsync(offs + tiles + 1, 0, 2 * sizeof(uint64_t), st.stream));
    {
        SPQ_PROF(st.stream, "chunk_count_kernel", (double)(G << nd) * 16.0);
        chunk_count_kernel<<<(unsigned)slot_tiles, kBlock, 0, st.stream>>>(st.amps, nullptr, nullptr, G, nd, lc, eps, nullptr, chunk,
                                                                           (unsigned long long*)(offs + tiles + 1));
    }
    SPQ_HIP(hipGetLastError());
    chunk_column_kernel<<<grid_for(tiles, kBlock), kBlock, 0, st.stream>>>(chunk, G, tiles, kept);
    SPQ_HIP(hipGetLastError());
    exclusive_scan_u32(st, kept, offs, tiles, offs + tiles);
    const uint64_t total = read_u64(st, offs + tiles);
    dfree(chunk, st.stream);
    const uint64_t cap_out = std::max<uint64_t>(total, 1);
    uint64_t* kout;
    double2* aout;
    alloc_arrays(st, st.W, cap_out, &kout, &aout);
    {
        SPQ_PROF(st.stream, "tx_emit_kernel", (double)(G << nd) * 16.0 + (double)total * (8.0 * W + 16));
        tx_emit_kernel<W, (W <= (int)kTxMaxKeyWords)><<<(unsigned)tiles, kBlock, 0, st.stream>>>(st.amps, nullptr, st.keys, st.cap, DenseMap{}, G, tg, nd, eps,
                                                                           offs, kout, aout, cap_out);
    }
    SPQ_HIP(hipGetLastError());
    install_arrays(st, st.W, kout, aout, total, cap_out);
    dfree(kept, st.stream);
    dfree(offs, st.stream);
}

// {kept total = 0, marks = 0, groups, WHT groups = 1}.
__global__ void dense_init_kernel(uint64_t* __restrict__ tail, uint64_t groups) {
    tail[0] = 0;
    tail[1] = 0;
    tail[2] = groups;
    tail[3] = 1;
}

// The dense-domain transform, when it applies: returns false (state untouched)
// when the domain is too large or too sparsely filled, or the state holds
// duplicate keys.
template <int W>
static bool hadamard_dense(State& st, const KeyStats& ks, sparq_field_t low, uint32_t nd, double eps) {
    const uint64_t n = st.n;
    if (eps <= 0.0 || n < (uint64_t(1) << 14) || ks.vary.size() != st.W || ks.ones.size() != st.W) return false;
    const uint64_t lowmask = fmask_host(nd) << low.shift;
    uint32_t t_o = 0;
    for (uint32_t w = 0; w < st.W; ++w) t_o += (uint32_t)__builtin_popcountll(ks.vary[w] & (w == low.word ? ~lowmask : ~0ull));
    const uint32_t T = t_o + nd;
    if (t_o == 0 || T > kDenseMaxLog || T < 12 || n * 8 < (uint64_t(1) << T)) return false;
    DenseMap dm{};
    dm.nd = nd;
    dm.low = low;
    for (int w = (int)st.W - 1; w >= 0; --w) {   // least significant word first, bit 0 up
        const uint64_t o = ks.vary[w] & ((uint32_t)w == low.word ? ~lowmask : ~0ull);
        for (int p = 0; p < 64; ++p) {
            if (!((o >> p) & 1)) continue;
            if (dm.nseg > 0 && dm.word[dm.nseg - 1] == w && dm.shift[dm.nseg - 1] + dm.width[dm.nseg - 1] == p) {
                ++dm.width[dm.nseg - 1];
                continue;
            }
            dm.word[dm.nseg] = (uint8_t)w;
            dm.shift[dm.nseg] = (uint8_t)p;
            dm.width[dm.nseg] = 1;
            ++dm.nseg;
        }
    }
    for (uint32_t w = 0; w < st.W; ++w)
        dm.base[w] = ks.ones[w] & ~ks.vary[w] & (w == low.word ? ~lowmask : ~0ull);
    // The output may go in canonical order (low target value major): when
    // every others bit lies below the low target bits and the others index is
    // narrow enough for a tile of the transposed order to cover every group.
    // It does when the output is not much smaller than the input (a spread
    // state, which the next sort would otherwise reorder in full).
    bool tx_ok = t_o <= kTransposeMaxLog && nd >= 12 && g_canonical;
    for (uint32_t q = 0; q < dm.nseg; ++q)
        tx_ok = tx_ok && (dm.word[q] > low.word || (dm.word[q] == low.word && dm.shift[q] + dm.width[q] <= low.shift));
    const uint64_t slots = uint64_t(1) << T, groups = uint64_t(1) << t_o;
    const uint32_t tiles = (uint32_t)(slots / kDenseTile);
    const uint32_t lc = tx_ok ? (uint32_t)__builtin_ctz(kTxTile) - t_o : 0;
    const uint64_t txt = tx_ok ? tx_tiles(t_o, nd) : 0;   // tiles of the transposed order
    double2* A = dalloc<double2>(slots, st.stream);
    uint8_t* mark = dalloc<uint8_t>(slots, st.stream);
    uint8_t* cls = dalloc<uint8_t>(groups, st.stream);
    uint32_t* kept = dalloc<uint32_t>(std::max<uint64_t>(tiles, txt), st.stream);
    uint32_t* chunk = tx_ok ? dalloc<uint32_t>(groups * txt, st.stream) : nullptr;
    uint64_t* offs = dalloc<uint64_t>(std::max<uint64_t>(tiles, txt) + 1, st.stream);
    uint64_t* tail = dalloc<uint64_t>(4, st.stream);   // {kept total, marks, groups, WHT groups}
    uint64_t* gcount = tail + 2;
    SPQ_HIP(hipMemsetAsync(A, 0, slots * sizeof(double2), st.stream));
    SPQ_HIP(hipMemsetAsync(mark, 0, slots, st.stream));
    SPQ_HIP(hipMemsetAsync(cls, kDirect, groups, st.stream));
    dense_init_kernel<<<1, 1, 0, st.stream>>>(tail, groups);
    SPQ_HIP(hipGetLastError());
    {
        SPQ_PROF(st.stream, "dense_scatter_kernel", n * (8.0 * W + 16 + 16 + 2));
        dense_scatter_kernel<W><<<grid_for(n, kBlock), kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, n, dm, A, mark, cls);
    }
    SPQ_HIP(hipGetLastError());
    // profiled bytes: 32 per entry, a lower bound (the slots of occupied groups
    // move; which groups are occupied is known only on the device)
    launch_wht(st, A, gcount, groups, nd, cls, (double)n * 32.0);
    // counts per tile of slots (and per chunk for the transposed order); a
    // tile lies in one group from nd = 12 on, else the per-slot class lookup
    const bool chunked = nd >= 12;
    if (chunked) {
        SPQ_PROF(st.stream, "chunk_count_kernel", (double)n * 17.0);   // lower bound, as the butterflies
        chunk_count_kernel<<<tiles, kBlock, 0, st.stream>>>(A, cls, mark, groups, nd, lc, eps, kept, chunk, (unsigned long long*)tail);
    } else {
        {
            SPQ_PROF(st.stream, "dense_count_kernel", (double)n * 17.0);
            dense_count_kernel<<<tiles, kBlock, 0, st.stream>>>(A, mark, cls, nd, eps, kept, (unsigned long long*)(tail + 1));
        }
        SPQ_HIP(hipGetLastError());
        exclusive_scan_u32(st, kept, offs, tiles, tail);
    }
    SPQ_HIP(hipGetLastError());
    uint64_t info[2];   // kept total, marks
    read_back(st, tail, sizeof(info), info);
    const bool ok = info[1] == n;
    if (ok) {
        const uint64_t total = info[0];
        const bool tx = tx_ok && total * 4 >= n;
        if (tx) {
            chunk_column_kernel<<<grid_for(txt, kBlock), kBlock, 0, st.stream>>>(chunk, groups, txt, kept);
            SPQ_HIP(hipGetLastError());
        }
        const uint64_t et = tx ? txt : tiles;
        if (chunked) exclusive_scan_u32(st, kept, offs, et, offs + et);
        const uint64_t cap_out = std::max<uint64_t>(total, 1);
        uint64_t* kout;
        double2* aout;
        alloc_arrays(st, st.W, cap_out, &kout, &aout);
        {
            SPQ_PROF(st.stream, tx ? "tx_emit_kernel" : "dense_emit_kernel", (double)n * 16.0 + (double)total * (8.0 * W + 16));   // lower bound
            if (tx)
                tx_emit_kernel<W, false><<<(unsigned)txt, kBlock, 0, st.stream>>>(A, cls, nullptr, 0, dm, groups, t_o, nd, eps, offs, kout,
                                                                                  aout, cap_out);
            else
                dense_emit_kernel<W><<<tiles, kBlock, 0, st.stream>>>(A, cls, dm, eps, offs, kout, aout, cap_out);
        }
        SPQ_HIP(hipGetLastError());
        install_arrays(st, st.W, kout, aout, total, cap_out);
    }
    if (chunk) dfree(chunk, st.stream);
    dfree(tail, st.stream);
    dfree(A, st.stream);
    dfree(mark, st.stream);
    dfree(cls, st.stream);
    dfree(kept, st.stream);
    dfree(offs, st.stream);
    return ok;
}

template <int W>
static void hadamard_impl(State& st, sparq_field_t target, uint32_t nd, bool full, double eps, const double2* dense,
                          bool allow_learned = true, bool unit_basis = false, bool check_dups = false,
                          const CondList* pass = nullptr) {
    const uint64_t n = st.n;
    const uint64_t F = uint64_t(1) << nd;
    sparq_field_t low{target.word, target.shift, (uint8_t)nd, 0};
    // Sort by (everything else, low target bits) unless already in that order;
    // when no bit outside the low target bits varies there is a single group.
    // Conditions (group unitaries only) take the general path, which finds the
    // groups that pass through.
    if (pass && pass->n == 0) pass = nullptr;
    bool single = (n == 1) && !pass;
    bool others_below = true;   // every varying bit outside the low target bits lies below them
    KeyStats ks;
    if (n > 1) {
        const SizerQuery q{n, st.W, nd, dense ? 2u : (full ? 1u : 0u), low.word, low.shift};
        ks = key_stats(st, &low, allow_learned ? &q : nullptr);
        // a completed dense-domain transform saw every entry on its own slot:
        // the keys were distinct
        if (!dense && g_dense_transform && hadamard_dense<W>(st, ks, low, nd, eps)) return;
        if (!ks.sorted) sort_with_stats(st, ks, &low);
        if (check_dups && has_sorted_duplicates(st)) throw Error(SPARQ_RUNTIME, "Hadamard_Int_Full: duplicate key in state");
        single = !pass;
        for (uint32_t w = 0; w < st.W; ++w) {
            uint64_t v = ks.vary[w];
            if (w == low.word) v &= ~(fmask_host(nd) << low.shift);
            if (v) single = false;
            // an others bit above the low target bits: group-major order is
            // not canonical transposed
            if (v && (w < low.word || (w == low.word && (v >> low.shift) != 0))) others_below = false;
        }
    }
    uint32_t* head = nullptr;
    uint64_t* pos = nullptr;
    uint64_t* start = nullptr;
    uint8_t* cls = nullptr;
    // counts = {G, WHT groups, valid output entries, kept entries} on the device
    uint64_t* counts = dalloc<uint64_t>(4, st.stream);
    uint64_t G = 0, n_wht = 0, G_cap = 0;
    bool learned = false;
    if (single) {
        uint8_t c;
        if (dense) c = kDense;
        else if (!full) c = (n == 1) ? kDirect : kWht;
        else c = (n == F) ? kWht : (n + 1 <= nd ? kSparseSum : kWht);
        if (c == kDirect) {
            uint64_t* kout;
            double2* aout;
            alloc_arrays(st, st.W, F, &kout, &aout);
            {
                SPQ_PROF(st.stream, "expand_single_kernel", F * (8.0 * W + 16));
                expand_single_kernel<W><<<grid_for(F, kBlock * 8), kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, low, nd,
                                                                                         1.0 / std::sqrt((double)F), kout, aout, F,
                                                                                         eps, counts);
            }
            SPQ_HIP(hipGetLastError());
            // The kernel's comparison cabs2(a * extra) < eps, done here when
            // a = (1, 0) is known: no readback.
            const double extra = 1.0 / std::sqrt((double)F);
            const uint64_t pruned = unit_basis ? ((extra * extra + 0.0 * 0.0 < eps) ? 1 : 0) : read_u64(st, counts);
            dfree(counts, st.stream);
            install_arrays(st, st.W, kout, aout, pruned ? 0 : F, F);
            return;
        }
        if (c == kWht && n == F) {
            // One full group in target order: the output keys are the input keys,
            // so the butterflies run in place on the amplitudes (one WHT group,
            // known on the host).
            dfree(counts, st.stream);
            launch_wht(st, st.amps, nullptr, 1, nd, nullptr, (double)F * 32.0);
            clear_zero(st, eps);
            return;
        }
        G = G_cap = 1;
        n_wht = (c == kWht) ? 1 : 0;
        start = dalloc<uint64_t>(2, st.stream);
        cls = dalloc<uint8_t>(1, st.stream);
        single_group_kernel<<<1, 1, 0, st.stream>>>(start, cls, n, c, counts);
        SPQ_HIP(hipGetLastError());
    } else {
        head = dalloc<uint32_t>(n, st.stream);
        pos = dalloc<uint64_t>(n + 2, st.stream);   // [n] = group count, [n+1] = WHT group count
        {
            SPQ_PROF(st.stream, "head_flag_kernel", n * (8.0 * W + 4));
            head_flag_kernel<W><<<grid_for(n, kBlock), kBlock, 0, st.stream>>>(st.keys, st.cap, n, low, head);
        }
        SPQ_HIP(hipGetLastError());
        exclusive_scan_u32(st, head, pos, n, pos + n);
        start = dalloc<uint64_t>(n + 1, st.stream);
        {
            SPQ_PROF(st.stream, "group_start_kernel", n * 12.0);
            group_start_kernel<<<grid_for(n, kBlock), kBlock, 0, st.stream>>>(head, pos, n, start, pos + n);
        }
        SPQ_HIP(hipGetLastError());
        cls = dalloc<uint8_t>(n, st.stream);
        SPQ_HIP(hipMemsetAsync(pos + n + 1, 0, sizeof(uint64_t), st.stream));
        {
            SPQ_PROF(st.stream, "classify_kernel", n * 17.0);
            classify_kernel<<<std::min<unsigned>(grid_for(n, kBlock * 4), 2048), kBlock, 0, st.stream>>>(start, pos + n, F, nd,
                                                                                                     dense ? 2 : (full ? 1 : 0), cls,
                                                                          (unsigned long long*)(pos + n + 1));
        }
        SPQ_HIP(hipGetLastError());
        SPQ_HIP(hipMemcpyAsync(counts, pos + n, 2 * sizeof(uint64_t), hipMemcpyDeviceToDevice, st.stream));
        if (pass) {
            pass_classify_kernel<W><<<std::min<unsigned>(grid_for(n, kBlock), 2048), kBlock, 0, st.stream>>>(st.keys, st.cap, start,
                                                                                                       pos + n, *pass, cls);
            SPQ_HIP(hipGetLastError());
        }
        // With the sizer active the output is sized by the predicted group count and
        // the exact count stays on the device until the final readback.
        if (ks.have_prediction && !pass) {
            // The prediction refines the exact bound G <= min(n, 2^(bits varying
            // outside the low target bits)) and never exceeds it.
            int vo = 0;
            for (uint32_t w = 0; w < st.W; ++w) {
                uint64_t v = ks.vary[w];
                if (w == low.word) v &= ~(fmask_host(nd) << low.shift);
                vo += __builtin_popcountll(v);
            }
            const uint64_t bound = vo >= 63 ? n : std::min<uint64_t>(n, uint64_t(1) << vo);
            const double lg = std::min(40.0, (double)ks.prediction * 32.0 + sizer_margin());
            const double g = std::ceil(std::exp2(std::max(0.0, lg)));
            G_cap = (uint64_t)std::min<double>((double)bound, g);
            // A predicted output beyond the learned-allocation budget (1/8 of the
            // device memory) is not trusted: a misprediction there would fail the
            // allocation instead of costing a rerun, so count exactly.
            learned = G_cap >= 1 && G_cap <= (~0ull >> nd) && (G_cap << nd) <= (uint64_t(1) << 33) &&
                      (double)(G_cap << nd) * (8.0 * st.W + 16.0) <= learned_budget_bytes();
        }
        if (!learned) {
            uint64_t h2[2];
            read_back(st, counts, sizeof(h2), h2);
            G = G_cap = h2[0];
            n_wht = h2[1];
            if (ks.have_features) sizer_record(ks.features, (float)(std::log2((double)G) / 32.0));
            sizer_count(0);
        }
    }
    SPQ_REQUIRE(G_cap <= (~0ull >> nd), "Hadamard output size overflows");
    // no butterflies: the outputs can be written straight in the transposed
    // (canonical) order
    const uint64_t tG = (g_canonical && !pass && !learned && n_wht == 0 && others_below && G > 1) ? G : 0;
    // conditioned: groups that pass through keep their members; the output
    // stays in group order
    uint64_t* goff = nullptr;
    uint64_t M_out = G_cap << nd;
    if (pass) {
        uint32_t* gsize = dalloc<uint32_t>(G, st.stream);
        goff = dalloc<uint64_t>(G + 1, st.stream);
        group_out_size_kernel<<<grid_for(G, kBlock), kBlock, 0, st.stream>>>(start, cls, G, nd, gsize);
        SPQ_HIP(hipGetLastError());
        exclusive_scan_u32(st, gsize, goff, G, goff + G);
        M_out = read_u64(st, goff + G);
        dfree(gsize, st.stream);
    }
    const uint64_t M_cap = G_cap << nd;
    uint64_t* kout;
    double2* aout;
    alloc_arrays(st, st.W, std::max<uint64_t>(M_out, 1), &kout, &aout);
    const double extra = 1.0 / std::sqrt((double)F);
    {
        SPQ_PROF(st.stream, "expand_init_kernel", M_out * (8.0 * W + 16));
        expand_init_kernel<W><<<grid_for(M_cap, kBlock * kExpandItems), kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, start, cls, nd, low, extra, counts,
                                                                             G_cap, kout, aout, std::max<uint64_t>(M_out, 1), dense, goff,
                                                                             tG);
    }
    SPQ_HIP(hipGetLastError());
    if (goff) dfree(goff, st.stream);
    if (n_wht || learned) {
        {
            SPQ_PROF(st.stream, "expand_scatter_kernel", n * (8.0 * W + 16));
            expand_scatter_kernel<W><<<grid_for(n, kBlock), kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, n, head, pos, cls, nd, low,
                                                                                     counts, G_cap, aout);
        }
        SPQ_HIP(hipGetLastError());
        launch_wht(st, aout, counts, G_cap, nd, cls, (double)(learned ? G_cap : n_wht) * F * 32.0);
    }
    dfree(head, st.stream);
    dfree(pos, st.stream);
    dfree(start, st.stream);
    dfree(cls, st.stream);
    if (!learned) {
        dfree(counts, st.stream);
        install_arrays(st, st.W, kout, aout, pass ? M_out : G << nd, std::max<uint64_t>(M_out, 1));
        // output in canonical order when the groups' others bits lie below
        // the target bits (the next sort then finds it sorted)
        if (g_canonical && !tG && !pass && others_below && G > 1 && G <= (uint64_t(1) << kTransposeMaxLog) && st.W <= kTxMaxKeyWords &&
            nd >= 12)
            transposed_clear_zero<W>(st, G, nd, eps);
        else
            clear_zero(st, eps);
        return;
    }
    // One readback for the exact group count, the overflow check and ClearZero.
    sized_total_kernel<<<1, 1, 0, st.stream>>>(counts, G_cap, nd);
    SPQ_HIP(hipGetLastError());
    uint64_t info[4];
    uint64_t* kz = nullptr;
    double2* az = nullptr;
    clear_zero_sized(st, kout, aout, M_cap, counts, eps, info, &kz, &az);
    G = info[0];
    if (ks.have_features) sizer_record(ks.features, (float)(std::log2((double)G) / 32.0));
    if (G > G_cap) {   // under-predicted: nothing was written; rerun with the exact count
        sizer_count(2);
        dfree(kout, st.stream);
        dfree(aout, st.stream);
        dfree(counts, st.stream);
        hadamard_impl<W>(st, target, nd, full, eps, dense, false);
        return;
    }
    sizer_count(1);
    dfree(counts, st.stream);
    if (kz == kout) {
        install_arrays(st, st.W, kout, aout, info[3], M_cap);
    } else {
        dfree(kout, st.stream);
        dfree(aout, st.stream);
        install_arrays(st, st.W, kz, az, info[3], std::max<uint64_t>(info[3], 1));
    }
}

void hadamard_int(State& st, sparq_field_t target, uint32_t nd, bool full, double eps, bool unit_basis, bool check_dups) {
    SPQ_REQUIRE(target.word < st.W && target.width >= 1 && (uint32_t)target.shift + target.width <= 64, "bad target field");
    SPQ_REQUIRE(nd >= 1 && nd <= target.width, "n_digits must be in [1, width]");
    SPQ_REQUIRE(nd <= 40, "n_digits too large for a sparse expansion");
    if (st.n == 0) return;
    SPQ_DISPATCH_W(st.W, hadamard_impl, st, target, nd, full, eps, (const double2*)nullptr, true, unit_basis && st.n == 1,
                   check_dups);
}

void group_unitary(State& st, sparq_field_t target, const double* mat, double eps, const CondList* cl) {
    SPQ_REQUIRE(target.word < st.W && target.width >= 1 && target.width <= 6 && (uint32_t)target.shift + target.width <= 64,
                "group unitary needs a register of 1..6 bits");
    if (st.n == 0) return;
    const uint64_t F = uint64_t(1) << target.width;
    double2* dense = dalloc<double2>(F * F, st.stream);
    SPQ_HIP(hipMemcpyAsync(dense, mat, F * F * sizeof(double2), hipMemcpyHostToDevice, st.stream));
    SPQ_HIP(hipStreamSynchronize(st.stream));   // the host matrix may be released after return
    SPQ_DISPATCH_W(st.W, hadamard_impl, st, target, (uint32_t)target.width, false, eps, (const double2*)dense, true, false, false, cl);
    dfree(dense, st.stream);
}

// ---------------------------------------------------------------- replicate
template <int W>
__global__ void __launch_bounds__(kBlock) replicate_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps, uint64_t cap,
                                                      uint32_t nd, sparq_field_t f, uint64_t M, uint64_t* __restrict__ kout,
                                                      double2* __restrict__ aout) {
    uint64_t o = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o >= M) return;
    const uint64_t g = o >> nd, j = o & ((uint64_t(1) << nd) - 1);
    uint64_t k[W];
    load_key<W>(keys, cap, g, k);
    fput<W>(k, f, j);
    store_key<W>(kout, M, o, k);
    aout[o] = amps[g];
}

template <int W>
static void replicate_impl(State& st, sparq_field_t f) {
    const uint32_t nd = f.width;
    SPQ_REQUIRE(st.n <= (~0ull >> nd), "replicated state too large");
    const uint64_t M = st.n << nd;
    uint64_t* kout;
    double2* aout;
    alloc_arrays(st, st.W, M, &kout, &aout);
    {
        SPQ_PROF(st.stream, "replicate_kernel", M * (8.0 * W + 16) + st.n * (8.0 * W + 16));
        replicate_kernel<W><<<grid_for(M, kBlock), kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, nd, f, M, kout, aout);
    }
    SPQ_HIP(hipGetLastError());
    install_arrays(st, st.W, kout, aout, M, M);
}

void replicate(State& st, sparq_field_t f) {
    SPQ_REQUIRE(f.word < st.W && f.width >= 1 && f.width <= 40, "bad replicate field");
    if (st.n == 0) return;
    SPQ_DISPATCH_W(st.W, replicate_impl, st, f);
}

// ---------------------------------------------------------------- pair ops
// Entries sorted with the target bit least significant (or known to hold their
// partners adjacent, bit clear first): a pair (i, i + 1) differs only in that
// bit.  Both kernels evaluate the same device functions, so the count and the
// write agree bit for bit; ClearZero (|a|^2 not below eps kept, as
// compact.hip) is applied on the way: only kept outputs are counted and
// written, the bit-clear output of a pair before its bit-set one.

// Whether entry i (key k) has its partner adjacent; `one`: i's bit is set.
template <int W>
__device__ __forceinline__ bool pair_partner(const uint64_t* __restrict__ keys, uint64_t cap, uint64_t n, uint64_t i,
                                             const uint64_t (&k)[W], uint32_t word, uint64_t bm, bool& one) {
    one = kword<W>(k, word) & bm;
    const uint64_t j = one ? i - 1 : i + 1;
    if (!((one && i > 0) || (!one && i + 1 < n))) return false;
    bool paired = true;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint64_t expect = (w == (int)word) ? (k[w] ^ bm) : k[w];
        paired = paired && (keys[(uint64_t)w * cap + j] == expect);
    }
    return paired;
}

// The two outputs (|bit clear>, |bit set>) of entry i, the bit-clear partner
// of a pair or an unpaired entry.
template <int W>
__device__ __forceinline__ void pair_values(const double2* __restrict__ amps, uint64_t i, bool paired, bool one,
                                            const uint64_t (&k)[W], const sparq_pair_params_t& p, double2& z2, double2& o2) {
    double2 z, on;                            // amplitudes of the |0> and |1> branches
    const double2 zero = make_double2(0.0, 0.0);
    if (paired) { z = amps[i]; on = amps[i + 1]; }
    else if (one) { z = zero; on = amps[i]; }
    else { z = amps[i]; on = zero; }
    switch (p.kind) {
        case SPARQ_PAIR_HADAMARD_BOOL:
            if (paired) {
                z2 = cadd(cscale(z, kSqrt2Inv), cscale(on, kSqrt2Inv));
                o2 = csub(cscale(z, kSqrt2Inv), cscale(on, kSqrt2Inv));
            } else if (!one) {
                z2 = cscale(z, kSqrt2Inv);
                o2 = cscale(z, kSqrt2Inv);
            } else {
                z2 = cscale(on, kSqrt2Inv);
                o2 = cscale(on, -kSqrt2Inv);
            }
            break;
        case SPARQ_PAIR_CONDROT_RATIONAL: {
            uint64_t v = fget<W>(k, p.in);
            double frac = (p.in.width == 64) ? v * 1.0 / 2 / 9223372036854775808.0 : 1.0 * v / (double)(uint64_t(1) << p.in.width);
            double rot = frac * (p.mat[0] * 2 * 3.141592653589793238462643383279502884);
            double c = cos(rot), s = sin(rot);
            z2 = csub(cscale(z, c), cscale(on, s));
            o2 = cadd(cscale(z, s), cscale(on, c));
            break;
        }
        default: {
            double2 m0, m1, m2, m3;
            if (p.kind == SPARQ_PAIR_CONDROT_TABLE) {
                const double* t = p.table + 8 * fget<W>(k, p.in);
                m0 = make_double2(t[0], t[1]); m1 = make_double2(t[2], t[3]);
                m2 = make_double2(t[4], t[5]); m3 = make_double2(t[6], t[7]);
            } else {
                m0 = make_double2(p.mat[0], p.mat[1]); m1 = make_double2(p.mat[2], p.mat[3]);
                m2 = make_double2(p.mat[4], p.mat[5]); m3 = make_double2(p.mat[6], p.mat[7]);
            }
            if (paired) {
                z2 = cadd(cmul(z, m0), cmul(on, m1));
                o2 = cadd(cmul(z, m2), cmul(on, m3));
            } else if (!one) {
                z2 = cmul(z, m0);
                o2 = cmul(z, m2);
            } else {
                z2 = cmul(on, m1);
                o2 = cmul(on, m3);
            }
            break;
        }
    }
}

// Kept outputs of entry i: (kept bit-clear output, kept bit-set output), or
// for an entry failing the conditions whether it stays; the bit-set partner
// of a pair contributes nothing (its pair's bit-clear partner writes both).
template <int W>
__device__ __forceinline__ void pair_entry(const uint64_t* __restrict__ keys, const double2* __restrict__ amps, uint64_t cap, uint64_t n,
                                           uint64_t i, uint32_t word, uint64_t bm, const CondList& cl, const sparq_pair_params_t& p,
                                           double eps, uint64_t (&k)[W], bool& pass, bool& keep_z, bool& keep_o, double2& z2,
                                           double2& o2) {
    load_key<W>(keys, cap, i, k);
    bool ok = true;
    for (uint32_t c = 0; c < cl.n; ++c) ok = ok && cond_ok<W>(k, cl.c[c]);
    pass = !ok;
    if (pass) {                                   // passes through unchanged, ClearZero'd
        z2 = amps[i];
        keep_z = !(cabs2(z2) < eps);
        keep_o = false;
        return;
    }
    bool one;
    const bool paired = pair_partner<W>(keys, cap, n, i, k, word, bm, one);
    if (paired && one) {
        keep_z = keep_o = false;
        return;
    }
    pair_values<W>(amps, i, paired, one, k, p, z2, o2);
    keep_z = !(cabs2(z2) < eps);
    keep_o = !(cabs2(o2) < eps);
}

template <int W>
__global__ void __launch_bounds__(kBlock) pair_count_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps, uint64_t cap,
                                                       uint64_t n, uint32_t word, uint64_t bm, const CondList cl,
                                                       const sparq_pair_params_t p, double eps, uint32_t* __restrict__ cnt) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint64_t k[W];
    bool pass, kz, ko;
    double2 z2, o2;
    pair_entry<W>(keys, amps, cap, n, i, word, bm, cl, p, eps, k, pass, kz, ko, z2, o2);
    cnt[i] = (uint32_t)kz + (uint32_t)ko;
}

template <int W>
__global__ void __launch_bounds__(kBlock) pair_apply_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps, uint64_t cap,
                                                       uint64_t n, uint32_t word, uint64_t bm, const CondList cl,
                                                       const uint64_t* __restrict__ off, const sparq_pair_params_t p, double eps,
                                                       uint64_t* __restrict__ kout, double2* __restrict__ aout, uint64_t cap_out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint64_t k[W];
    bool pass, kz, ko;
    double2 z2, o2;
    pair_entry<W>(keys, amps, cap, n, i, word, bm, cl, p, eps, k, pass, kz, ko, z2, o2);
    uint64_t o = off[i];
    if (pass) {
        if (kz) {
            store_key<W>(kout, cap_out, o, k);
            aout[o] = z2;
        }
        return;
    }
    if (kz) {
        uint64_t kk[W];
#pragma unroll
        for (int w = 0; w < W; ++w) kk[w] = (w == (int)word) ? (k[w] & ~bm) : k[w];
        store_key<W>(kout, cap_out, o, kk);
        aout[o] = z2;
        ++o;
    }
    if (ko) {
        uint64_t kk[W];
#pragma unroll
        for (int w = 0; w < W; ++w) kk[w] = (w == (int)word) ? (k[w] | bm) : k[w];
        store_key<W>(kout, cap_out, o, kk);
        aout[o] = o2;
    }
}

template <int W>
static void pair_impl(State& st, sparq_field_t target, uint32_t bit, const sparq_pair_params_t& p, const CondList& cl, double eps) {
    const uint64_t n = st.n;
    sparq_field_t b{target.word, (uint8_t)(target.shift + bit), 1, 0};
    // partners adjacent, bit-clear first: sorted with the bit least significant,
    // or already grouped so across it
    const int32_t pair_bit = (int32_t)(b.word * 64 + b.shift);
    if (st.pair_bit != pair_bit) sort_state(st, &b);
    st.pair_bit = -1;
    const uint64_t bm = uint64_t(1) << b.shift;
    uint32_t* cnt = dalloc<uint32_t>(n, st.stream);
    uint64_t* off = dalloc<uint64_t>(n + 1, st.stream);
    {
        SPQ_PROF(st.stream, "pair_count_kernel", n * (8.0 * W + 16 + 4));
        pair_count_kernel<W><<<grid_for(n, kBlock), kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, n, b.word, bm, cl, p, eps, cnt);
    }
    SPQ_HIP(hipGetLastError());
    exclusive_scan_u32(st, cnt, off, n, off + n);
    const uint64_t M = read_u64(st, off + n);
    uint64_t* kout;
    double2* aout;
    alloc_arrays(st, st.W, std::max<uint64_t>(M, 1), &kout, &aout);
    {
        SPQ_PROF(st.stream, "pair_apply_kernel", n * (8.0 * W + 16 + 8) + M * (8.0 * W + 16));
        pair_apply_kernel<W><<<grid_for(n, kBlock), kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, n, b.word, bm, cl, off, p, eps, kout,
                                                                         aout, std::max<uint64_t>(M, 1));
    }
    SPQ_HIP(hipGetLastError());
    install_arrays(st, st.W, kout, aout, M, std::max<uint64_t>(M, 1));
    dfree(cnt, st.stream);
    dfree(off, st.stream);
    // every kept output pair is written as (bit clear, bit set) at adjacent
    // slots; entries failing the conditions pass through beside their
    // partners; ClearZero only drops entries
    st.pair_bit = pair_bit;
}

void pair_op(State& st, sparq_field_t target, uint32_t bit, const sparq_pair_params_t& p, const CondList& cl, double eps) {
    SPQ_REQUIRE(target.word < st.W && bit < target.width, "bad target bit");
    SPQ_REQUIRE(p.kind <= SPARQ_PAIR_ROT_GENERAL, "unknown pair kind");
    if (p.kind == SPARQ_PAIR_CONDROT_RATIONAL || p.kind == SPARQ_PAIR_CONDROT_TABLE)
        SPQ_REQUIRE(p.in.word < st.W && p.in.width >= 1, "bad input register");
    if (p.kind == SPARQ_PAIR_CONDROT_TABLE) SPQ_REQUIRE(p.table != nullptr && p.in.width <= 24, "table-driven rotation needs a table");
    if (st.n == 0) return;
    SPQ_DISPATCH_W(st.W, pair_impl, st, target, bit, p, cl, eps);
}

// ---------------------------------------------------------------- captured transform
// Hadamard_Int inside a captured sequence: the dense-domain transform with a
// plan fixed when the sequence was planned.  The satisfying entries' others
// value (the key bits outside the low target bits that vary across the planned
// state, packed in canonical order) maps through a lookup table to a segment
// number: one segment per value present, numbered in ascending order, segment
// g owning slots [g << nd, (g + 1) << nd) -- for a batch of trajectories, one
// segment per trajectory.  Every kernel reads the entry count from the device
// (grid-stride, persistent grids); nothing is read back or allocated.  The
// output goes back into the state's own arrays (the scatter has consumed the
// input): the transformed entries in slot order (canonical order), then the
// entries that fail the conditions, in their order.  What the plan assumed
// (constant bits, segments present, no duplicate keys, the capacity) is
// checked on the device into the capture's flags word.
struct CapHadamard {
    sparq_field_t target{}, low{};
    uint32_t nd = 0;
    double eps = 0.0;
    uint32_t ncond = 0;
    sparq_cond_t conds[SPARQ_MAX_PROGRAM_CONDS] = {};
    DenseMap dm{};                          // others bits (runs); base = the constant bits' values
    uint64_t cmask[SPARQ_MAX_WORDS] = {};   // bits outside the others and the low target bits
    uint32_t t_o = 0;
    uint64_t G = 0, slots = 0, slots_pad = 0, rest = 0;
    std::vector<uint32_t> lut;              // 2^t_o: segment number, or ~0u
    std::vector<uint64_t> gkey;             // [G][W]: the segment's key, low target bits zero
    std::vector<uint8_t> cls;               // slots_pad >> nd: kWht for segments, kDirect for padding
    uint32_t* d_lut = nullptr;
    uint64_t* d_gkey = nullptr;
    uint8_t* d_cls = nullptr;
};

struct KeyMask {
    uint64_t m[SPARQ_MAX_WORDS];
};

constexpr int kCapTile = 4096;              // entries per tile of the scatter and rest kernels
constexpr int kCapItems = kCapTile / kBlock;
constexpr uint32_t kCapMaxOthers = 24;      // lookup table of at most 2^24 segment numbers

static bool host_cond_ok(const uint64_t* k, const sparq_cond_t& c) {
    const uint64_t v = (k[c.field.word] >> c.field.shift) & fmask_host(c.field.width);
    switch (c.kind) {
        case SPARQ_COND_NONZERO: return v != 0;
        case SPARQ_COND_ALL_ONES: return v == fmask_host(c.field.width);
        case SPARQ_COND_BIT: return (v >> c.value) & 1;
        case SPARQ_COND_BIT_CLEAR: return !((v >> c.value) & 1);
        case SPARQ_COND_GE: return v >= c.value;
        case SPARQ_COND_LE: return v <= c.value;
        default: return v == c.value;
    }
}

// Zero the dense domain and the op's counters.
__global__ void __launch_bounds__(kBlock) cap_fill_kernel(double2* __restrict__ A, uint8_t* __restrict__ mark, uint64_t slots,
                                                     uint64_t* __restrict__ words) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < slots; i += (uint64_t)gridDim.x * kBlock) {
        A[i] = make_double2(0.0, 0.0);
        mark[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        words[kCapSat] = 0;
        words[kCapMarks] = 0;
        words[kCapRest] = 0;
    }
}

// Entries failing the conditions, per tile of kCapTile entries (tiles past the
// entry count give 0).
template <int W>
__global__ void __launch_bounds__(kBlock) cap_rest_count_kernel(const uint64_t* __restrict__ keys, uint64_t cap,
                                                           const uint64_t* __restrict__ words, CondList cl,
                                                           uint32_t* __restrict__ rest_cnt, uint64_t ntiles) {
    __shared__ uint32_t lds[kBlock / 64];
    const uint64_t n = words[kCapN];
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint32_t c = 0;
#pragma unroll 4
        for (int j = 0; j < kCapItems; ++j) {
            const uint64_t i = tile * kCapTile + (uint64_t)j * kBlock + threadIdx.x;
            if (i >= n) break;
            uint64_t k[W];
            load_key<W>(keys, cap, i, k);
            bool ok = true;
            for (uint32_t q = 0; q < cl.n; ++q) ok = ok && cond_ok<W>(k, cl.c[q]);
            c += ok ? 0 : 1;
        }
        c = wave_sum(c);
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
            rest_cnt[tile] = t;
        }
        __syncthreads();
    }
}

// Satisfying entries to their slot (and mark); the others, in order, to the
// rest arrays at rest_off[tile] (when the op has conditions).
template <int W>
__global__ void __launch_bounds__(kBlock) cap_scatter_kernel(const uint64_t* __restrict__ keys, const double2* __restrict__ amps,
                                                        uint64_t cap, uint64_t* __restrict__ words, DenseMap dm,
                                                        KeyMask cmask_in, const uint32_t* __restrict__ lut,
                                                        CondList cl, double2* __restrict__ A, uint8_t* __restrict__ mark,
                                                        uint64_t* __restrict__ rest_keys, double2* __restrict__ rest_amps,
                                                        uint64_t rest_cap, const uint64_t* __restrict__ rest_off, uint64_t ntiles) {
    constexpr int kWaves = kBlock / 64;
    __shared__ uint32_t row_wave[kCapItems * kWaves];
    __shared__ uint32_t sat_lds[kWaves];
    const uint64_t n = words[kCapN];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint64_t cmask[W];
#pragma unroll
    for (int w = 0; w < W; ++w) cmask[w] = cmask_in.m[w];
    uint32_t sat_count = 0, bad = 0;
    for (uint64_t tile = blockIdx.x; tile * kCapTile < n && tile < ntiles; tile += gridDim.x) {
        uint32_t rest_flags = 0, lane_pre[kCapItems];
#pragma unroll
        for (int j = 0; j < kCapItems; ++j) {
            const uint64_t i = tile * kCapTile + (uint64_t)j * kBlock + threadIdx.x;
            bool rest = false;
            if (i < n) {
                uint64_t k[W];
                load_key<W>(keys, cap, i, k);
                bool ok = true;
                for (uint32_t q = 0; q < cl.n; ++q) ok = ok && cond_ok<W>(k, cl.c[q]);
                if (ok) {
                    bool in_plan = true;
#pragma unroll
                    for (int w = 0; w < W; ++w) in_plan = in_plan && ((k[w] & cmask[w]) == dm.base[w]);
                    const uint32_t g = in_plan ? lut[dense_others<W>(k, dm)] : ~0u;
                    if (g != ~0u) {
                        const uint64_t slot = ((uint64_t)g << dm.nd) | fget<W>(k, dm.low);
                        A[slot] = amps[i];
                        mark[slot] = 1;
                        ++sat_count;
                    } else {
                        bad = 1;
                    }
                } else {
                    rest = true;
                }
            }
            const uint64_t m = __ballot(rest);
            rest_flags |= uint32_t(rest) << j;
            lane_pre[j] = (uint32_t)__popcll(m & lanemask_lt());
            if (lane == 0) row_wave[j * kWaves + wave] = (uint32_t)__popcll(m);
        }
        if (rest_keys) {
            __syncthreads();
            if (threadIdx.x < 64) {   // exclusive scan of the (row, wave) counts in entry order
                const uint32_t x = row_wave[threadIdx.x];
                const uint32_t inc = wave_incl_scan(x);
                row_wave[threadIdx.x] = inc - x;
            }
            __syncthreads();
            const uint64_t before = rest_off[tile];
#pragma unroll
            for (int j = 0; j < kCapItems; ++j) {
                if (!((rest_flags >> j) & 1)) continue;
                const uint64_t i = tile * kCapTile + (uint64_t)j * kBlock + threadIdx.x;
                const uint64_t dst = before + row_wave[j * kWaves + wave] + lane_pre[j];
                if (dst >= rest_cap) continue;
#pragma unroll
                for (int w = 0; w < W; ++w) rest_keys[(uint64_t)w * rest_cap + dst] = keys[(uint64_t)w * cap + i];
                rest_amps[dst] = amps[i];
            }
            __syncthreads();
        }
    }
    sat_count = wave_sum(sat_count);
    if (lane == 0) sat_lds[wave] = sat_count;
    if (__any(bad) && lane == 0) atomicOr((unsigned long long*)(words + kCapFlags), (unsigned long long)kCapOutsidePlan);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kWaves; ++w) t += sat_lds[w];
        if (t) atomicAdd((unsigned long long*)(words + kCapSat), (unsigned long long)t);
    }
}

// The op's output size: kept slots + rest entries, bounded by the capacity;
// duplicate keys (fewer marks than satisfying entries) and overflow flagged.
__global__ void cap_finish_kernel(uint64_t* __restrict__ words, const uint64_t* __restrict__ kept_total, uint64_t capacity) {
    const uint64_t kept = kept_total ? kept_total[0] : 0;
    if (words[kCapMarks] != words[kCapSat]) words[kCapFlags] |= kCapDuplicate;
    uint64_t total = kept + words[kCapRest];
    if (total > capacity) {
        words[kCapFlags] |= kCapOverflow;
        total = capacity;
    }
    words[kCapN] = total;
}

// Kept slots in slot order: key = segment key with the low target bits set.
template <int W>
__global__ void __launch_bounds__(kBlock) cap_emit_kernel(const double2* __restrict__ A, const uint8_t* __restrict__ cls, uint32_t nd,
                                                     sparq_field_t low, double eps, const uint64_t* __restrict__ offs,
                                                     const uint64_t* __restrict__ gkey, uint64_t* __restrict__ kout,
                                                     double2* __restrict__ aout, uint64_t cap_out, uint64_t capacity) {
    constexpr int kWaves = kBlock / 64;
    __shared__ uint32_t row_wave[kDenseItems * kWaves];
    const uint64_t tile = (uint64_t)blockIdx.x * kDenseTile;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double2 v[kDenseItems];
    uint32_t flags = 0, lane_pre[kDenseItems];
#pragma unroll
    for (int j = 0; j < kDenseItems; ++j) {
        const uint64_t s = tile + (uint64_t)j * kBlock + threadIdx.x;
        const bool present = cls[s >> nd] == kWht;
        v[j] = present ? A[s] : make_double2(0.0, 0.0);
        const bool keep = present && !(cabs2(v[j]) < eps);
        const uint64_t m = __ballot(keep);
        flags |= uint32_t(keep) << j;
        lane_pre[j] = (uint32_t)__popcll(m & lanemask_lt());
        if (lane == 0) row_wave[j * kWaves + wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t x = threadIdx.x < kDenseItems * kWaves ? row_wave[threadIdx.x] : 0;
        const uint32_t inc = wave_incl_scan(x);
        if (threadIdx.x < kDenseItems * kWaves) row_wave[threadIdx.x] = inc - x;
    }
    __syncthreads();
    const uint64_t before = offs[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kDenseItems; ++j) {
        if (!((flags >> j) & 1)) continue;
        const uint64_t s = tile + (uint64_t)j * kBlock + threadIdx.x;
        const uint64_t dst = before + row_wave[j * kWaves + wave] + lane_pre[j];
        if (dst >= capacity) continue;
        const uint64_t g = s >> nd;
        uint64_t k[W];
#pragma unroll
        for (int w = 0; w < W; ++w) k[w] = gkey[g * W + w];
        fput<W>(k, low, s & fmask(nd));
        store_key<W>(kout, cap_out, dst, k);
        aout[dst] = v[j];
    }
}

// The rest entries after the kept ones.
template <int W>
__global__ void __launch_bounds__(kBlock) cap_rest_copy_kernel(const uint64_t* __restrict__ rest_keys, const double2* __restrict__ rest_amps,
                                                          uint64_t rest_cap, const uint64_t* __restrict__ words,
                                                          const uint64_t* __restrict__ kept_total, uint64_t* __restrict__ kout,
                                                          double2* __restrict__ aout, uint64_t cap_out, uint64_t capacity) {
    const uint64_t r = words[kCapRest], kept = kept_total ? kept_total[0] : 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < r; i += (uint64_t)gridDim.x * kBlock) {
        const uint64_t dst = kept + i;
        if (dst >= capacity || i >= rest_cap) break;
#pragma unroll
        for (int w = 0; w < W; ++w) kout[(uint64_t)w * cap_out + dst] = rest_keys[(uint64_t)w * rest_cap + i];
        aout[dst] = rest_amps[i];
    }
}

// Phase 1: the plan of one op from the snapshot.
static CapHadamard* plan_hadamard(const Capture& c, sparq_field_t target, uint32_t nd, double eps, const sparq_cond_t* conds,
                                  uint32_t ncond) {
    const uint32_t W = c.W;
    auto* p = new CapHadamard();
    std::unique_ptr<CapHadamard> guard(p);
    p->target = target;
    p->low = sparq_field_t{target.word, target.shift, (uint8_t)nd, 0};
    p->nd = nd;
    p->eps = eps;
    p->ncond = ncond;
    for (uint32_t q = 0; q < ncond; ++q) p->conds[q] = conds[q];
    const uint64_t lowmask = fmask_host(nd) << target.shift;
    // the key bits that vary among the entries the op transforms, and their OR
    std::vector<uint64_t> k(W), any(W, 0), all(W, ~0ull);
    uint64_t sat = 0;
    auto satisfies = [&](uint64_t i) {
        for (uint32_t w = 0; w < W; ++w) k[w] = c.snap[(uint64_t)w * c.n0 + i];
        bool ok = true;
        for (uint32_t q = 0; q < ncond && ok; ++q) ok = host_cond_ok(k.data(), conds[q]);
        return ok;
    };
    for (uint64_t i = 0; i < c.n0; ++i) {
        if (!satisfies(i)) continue;
        ++sat;
        for (uint32_t w = 0; w < W; ++w) {
            any[w] |= k[w];
            all[w] &= k[w];
        }
    }
    std::vector<uint64_t> vary(W, 0);
    for (uint32_t w = 0; w < W && sat; ++w) vary[w] = any[w] ^ all[w];
    DenseMap& dm = p->dm;
    dm.nd = nd;
    dm.low = p->low;
    for (int w = (int)W - 1; w >= 0; --w) {   // least significant word first, bit 0 up
        const uint64_t o = vary[w] & ((uint32_t)w == target.word ? ~lowmask : ~0ull);
        for (int b = 0; b < 64; ++b) {
            if (!((o >> b) & 1)) continue;
            ++p->t_o;
            SPQ_REQUIRE(p->t_o <= kCapMaxOthers, "captured Hadamard_Int: more than 24 key bits vary outside its target");
            if (dm.nseg > 0 && dm.word[dm.nseg - 1] == w && dm.shift[dm.nseg - 1] + dm.width[dm.nseg - 1] == b) {
                ++dm.width[dm.nseg - 1];
                continue;
            }
            dm.word[dm.nseg] = (uint8_t)w;
            dm.shift[dm.nseg] = (uint8_t)b;
            dm.width[dm.nseg] = 1;
            ++dm.nseg;
        }
    }
    for (uint32_t w = 0; w < W; ++w) {
        p->cmask[w] = ~vary[w] & (w == target.word ? ~lowmask : ~0ull);
        dm.base[w] = any[w] & p->cmask[w];
    }
    auto others = [&](const uint64_t* k) {
        uint64_t v = 0;
        uint32_t at = 0;
        for (uint32_t s = 0; s < dm.nseg; ++s) {
            v |= ((k[dm.word[s]] >> dm.shift[s]) & fmask_host(dm.width[s])) << at;
            at += dm.width[s];
        }
        return v;
    };
    p->lut.assign(size_t(1) << p->t_o, 0u);
    for (uint64_t i = 0; i < c.n0; ++i)
        if (satisfies(i)) p->lut[others(k.data())] = 1;
    for (uint64_t o = 0; o < p->lut.size(); ++o) {
        if (!p->lut[o]) {
            p->lut[o] = ~0u;
            continue;
        }
        p->lut[o] = (uint32_t)p->G++;
        for (uint32_t w = 0; w < W; ++w) p->gkey.push_back(dm.base[w]);
        uint64_t* gk = p->gkey.data() + (p->G - 1) * W;
        uint32_t at = 0;
        for (uint32_t s = 0; s < dm.nseg; ++s) {
            gk[dm.word[s]] |= ((o >> at) & fmask_host(dm.width[s])) << dm.shift[s];
            at += dm.width[s];
        }
    }
    SPQ_REQUIRE(p->G <= (uint64_t(1) << (32 - std::min<uint32_t>(nd, 32))) && nd <= 30,
                "captured Hadamard_Int: domain beyond 2^32 slots");
    p->slots = p->G << nd;
    p->slots_pad = (p->slots + kDenseTile - 1) / kDenseTile * kDenseTile;
    SPQ_REQUIRE(p->slots_pad <= (uint64_t(1) << 32), "captured Hadamard_Int: domain beyond 2^32 slots");
    p->cls.assign(p->slots_pad >> nd, (uint8_t)kDirect);
    for (uint64_t g = 0; g < p->G; ++g) p->cls[g] = kWht;
    p->rest = c.n0 - sat;
    return guard.release();
}

uint64_t capture_plan_bytes(const Capture& c) {
    uint64_t b = 0;
    for (const CapHadamard* p : c.hs) {
        b += (p->lut.size() * 4 + 255) / 256 * 256;
        b += (p->gkey.size() * 8 + 255) / 256 * 256;
        b += (p->cls.size() + 255) / 256 * 256;
    }
    return b;
}

void capture_plan_upload(Capture& c, char*& at) {
    auto put = [&](const void* src, size_t bytes) {
        void* dst = at;
        if (bytes) SPQ_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        at += (bytes + 255) / 256 * 256;
        return dst;
    };
    for (CapHadamard* p : c.hs) {
        p->d_lut = (uint32_t*)put(p->lut.data(), p->lut.size() * 4);
        p->d_gkey = (uint64_t*)put(p->gkey.data(), p->gkey.size() * 8);
        p->d_cls = (uint8_t*)put(p->cls.data(), p->cls.size());
    }
}

uint64_t capture_plan_slots(const CapHadamard& p) { return p.slots; }
uint64_t capture_plan_slots_padded(const CapHadamard& p) { return p.slots_pad; }
uint64_t capture_plan_rest(const CapHadamard& p) { return p.rest; }
bool capture_plan_has_conditions(const CapHadamard& p) { return p.ncond > 0; }

void capture_plan_free(Capture& c) {
    for (CapHadamard* p : c.hs) delete p;
    c.hs.clear();
}

// Phase 2: the op's kernels on the capturing stream.
template <int W>
static void record_hadamard(State& st, const CapHadamard& p) {
    Capture& c = *st.capture;
    hipStream_t s = st.stream;
    uint64_t* words = c.words;
    CondList cl{};
    cl.n = p.ncond;
    for (uint32_t q = 0; q < p.ncond; ++q) cl.c[q] = p.conds[q];
    const bool rest = p.ncond > 0;
    const uint64_t ntiles = (c.capacity + kCapTile - 1) / kCapTile;
    cap_fill_kernel<<<persistent_grid(std::max<uint64_t>(p.slots_pad, 1), kBlock), kBlock, 0, s>>>(c.A, c.mark, p.slots_pad, words);
    SPQ_HIP(hipGetLastError());
    if (rest) {
        cap_rest_count_kernel<W><<<persistent_grid(c.capacity, kCapTile), kBlock, 0, s>>>(st.keys, st.cap, words, cl, c.rest_cnt, ntiles);
        SPQ_HIP(hipGetLastError());
        exclusive_scan_u32_scratch(s, c.rest_cnt, c.rest_off, ntiles, words + kCapRest, c.scan_scratch, c.scan_words);
    }
    KeyMask cm{};
    for (uint32_t w = 0; w < st.W; ++w) cm.m[w] = p.cmask[w];
    cap_scatter_kernel<W><<<persistent_grid(c.capacity, kCapTile), kBlock, 0, s>>>(
        st.keys, st.amps, st.cap, words, p.dm, cm, p.d_lut, cl, c.A, c.mark, rest ? c.rest_keys : nullptr,
        c.rest_amps, c.capacity, c.rest_off, ntiles);
    SPQ_HIP(hipGetLastError());
    const uint64_t stiles = p.slots_pad / kDenseTile;
    if (p.G > 0) {
        launch_wht(st, c.A, nullptr, p.G, p.nd, nullptr, (double)p.slots * 32.0);
        dense_count_kernel<<<(unsigned)stiles, kBlock, 0, s>>>(c.A, c.mark, p.d_cls, p.nd, p.eps, c.kept,
                                                               (unsigned long long*)(words + kCapMarks));
        SPQ_HIP(hipGetLastError());
        exclusive_scan_u32_scratch(s, c.kept, c.offs, stiles, c.offs + stiles, c.scan_scratch, c.scan_words);
    }
    const uint64_t* kept_total = p.G > 0 ? c.offs + stiles : nullptr;
    cap_finish_kernel<<<1, 1, 0, s>>>(words, kept_total, c.capacity);
    SPQ_HIP(hipGetLastError());
    if (p.G > 0) {
        cap_emit_kernel<W><<<(unsigned)stiles, kBlock, 0, s>>>(c.A, p.d_cls, p.nd, p.low, p.eps, c.offs, p.d_gkey, st.keys, st.amps, st.cap,
                                                             c.capacity);
        SPQ_HIP(hipGetLastError());
    }
    if (rest) {
        cap_rest_copy_kernel<W><<<persistent_grid(c.capacity, kBlock), kBlock, 0, s>>>(c.rest_keys, c.rest_amps, c.capacity, words,
                                                                                     kept_total, st.keys, st.amps, st.cap, c.capacity);
        SPQ_HIP(hipGetLastError());
    }
}

void hadamard_captured(State& st, sparq_field_t target, uint32_t nd, double eps, const sparq_cond_t* conds, uint32_t ncond) {
    Capture& c = *st.capture;
    SPQ_REQUIRE(target.word < st.W && target.width >= 1 && (uint32_t)target.shift + target.width <= 64, "bad target field");
    SPQ_REQUIRE(nd >= 1 && nd <= target.width, "n_digits must be in [1, width]");
    SPQ_REQUIRE(eps > 0.0, "a captured Hadamard_Int needs a positive ClearZero threshold");
    SPQ_REQUIRE(ncond <= SPARQ_MAX_PROGRAM_CONDS, "too many conditions");
    if (c.phase == 1) {
        c.hs.push_back(plan_hadamard(c, target, nd, eps, conds, ncond));
        return;
    }
    SPQ_REQUIRE(c.next_h < c.hs.size(), "the captured sequence runs more branching operators than its plan");
    const CapHadamard& p = *c.hs[c.next_h++];
    bool same = std::memcmp(&p.target, &target, sizeof(target)) == 0 && p.nd == nd && p.eps == eps && p.ncond == ncond;
    for (uint32_t q = 0; same && q < ncond; ++q)
        same = p.conds[q].kind == conds[q].kind && std::memcmp(&p.conds[q].field, &conds[q].field, sizeof(sparq_field_t)) == 0 &&
               p.conds[q].value == conds[q].value;
    SPQ_REQUIRE(same, "the captured sequence differs from its plan");
    SPQ_DISPATCH_W(st.W, record_hadamard, st, p);
}

}  // namespace sparq
