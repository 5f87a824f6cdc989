// Device helpers: register-resident packed keys, field access, conditions and
// FP64 complex arithmetic written out explicitly (the library is compiled with
// -ffp-contract=off so every product and sum rounds exactly like the host
// reference's std::complex<double> arithmetic).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/sparq_hip.h"

namespace sparq {

__device__ __forceinline__ uint64_t fmask(uint32_t width) {
    return width >= 64 ? ~uint64_t(0) : ((uint64_t(1) << width) - 1);
}

// Select word w of a register-resident key without dynamic register indexing.
// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx.x % 8), each
// with its own L2.  xcd_tile maps block b to a tile index so that each XCD gets
// a contiguous range of tiles (XCD x: tiles [x*q + min(x, r), ...) with
// q = nb / 8, r = nb % 8; a bijection on [0, nb)).  Tiles whose outputs or
// inputs share cache lines with their neighbours then meet in one L2.
constexpr uint32_t kXcds = 8;
__host__ __device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
    const uint32_t x = b % kXcds, j = b / kXcds, q = nb / kXcds, r = nb % kXcds;
    return x * q + min(x, r) + j;
}

// Word w of a register-resident key.  Masks rather than selects: a select chain
// over the words can be folded back into an indexed array, which lives in
// scratch memory.
__device__ __forceinline__ uint64_t kword_mask(int w, int i) { return uint64_t(0) - (uint64_t)(w == i); }
template <int W>
__device__ __forceinline__ uint64_t kword(const uint64_t (&k)[W], int w) {
    if (W == 1) return k[0];   // a field of a one-word key is in word 0
    uint64_t r = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) r |= k[i] & kword_mask(w, i);
    return r;
}
template <int W>
__device__ __forceinline__ void kset_word(uint64_t (&k)[W], int w, uint64_t v) {
    if (W == 1) {
        k[0] = v;
        return;
    }
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const uint64_t m = kword_mask(w, i);
        k[i] = (k[i] & ~m) | (v & m);
    }
}
template <int W>
__device__ __forceinline__ uint64_t fget(const uint64_t (&k)[W], sparq_field_t f) {
    return (kword<W>(k, f.word) >> f.shift) & fmask(f.width);
}
template <int W>
__device__ __forceinline__ void fput(uint64_t (&k)[W], sparq_field_t f, uint64_t v) {
    uint64_t m = fmask(f.width) << f.shift;
    uint64_t w = kword<W>(k, f.word);
    w = (w & ~m) | ((v << f.shift) & m);
    kset_word<W>(k, f.word, w);
}

template <int W>
__device__ __forceinline__ void load_key(const uint64_t* __restrict__ keys, uint64_t cap, uint64_t i, uint64_t (&k)[W]) {
#pragma unroll
    for (int w = 0; w < W; ++w) k[w] = keys[(uint64_t)w * cap + i];
}
template <int W>
__device__ __forceinline__ void store_key(uint64_t* __restrict__ keys, uint64_t cap, uint64_t i, const uint64_t (&k)[W]) {
#pragma unroll
    for (int w = 0; w < W; ++w) keys[(uint64_t)w * cap + i] = k[w];
}

__device__ __forceinline__ int64_t sext(uint64_t v, uint32_t w) {
    if (w == 0) return 0;
    if (w >= 64) return (int64_t)v;
    return (int64_t)(v << (64 - w)) >> (64 - w);
}

template <int W>
__device__ __forceinline__ bool cond_ok(const uint64_t (&k)[W], const sparq_cond_t& c) {
    uint64_t v = fget<W>(k, c.field);
    switch (c.kind) {
        case SPARQ_COND_NONZERO: return v != 0;
        case SPARQ_COND_ALL_ONES: return v == fmask(c.field.width);
        case SPARQ_COND_BIT: return (v >> c.value) & 1;
        case SPARQ_COND_BIT_CLEAR: return !((v >> c.value) & 1);
        case SPARQ_COND_GE: return v >= c.value;
        case SPARQ_COND_LE: return v <= c.value;
        default: return v == c.value;
    }
}

// ---- complex FP64 (no contraction: see file header)
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

// 1/sqrt(2), correctly rounded.
constexpr double kSqrt2Inv = 0.7071067811865475244008443621048490;

}  // namespace sparq
