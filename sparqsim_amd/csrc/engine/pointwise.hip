// Fused per-entry program: key rewrites (arithmetic, QRAM XOR-load, bit ops)
// and diagonal amplitude phases applied to every entry in one streaming pass.
// HBM bytes per entry: 2*8*W (+32 when any op touches the amplitude)
// (+ the table word for each QRAM load, served mostly from L2/MALL).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "engine.hpp"
#include "keyops.hpp"

namespace sparq {

// A QRAM table word read before the program's ops run (PwProgram::pre).
struct PwPrefetch {
    const void* table;
    uint64_t mask;
    sparq_field_t f;
    uint32_t bytes;
};
constexpr int kPwPrefetch = 2;

struct PwProgram {
    uint32_t nops, nconds, npre, pad;
    sparq_pw_op_t ops[SPARQ_MAX_PROGRAM_OPS];
    sparq_cond_t conds[SPARQ_MAX_PROGRAM_CONDS];
    PwPrefetch pre[kPwPrefetch];
};

// Fixed-point encoding of x in [0,1) into w bits by truncation; identical to the
// host layer's rational encoding.
__device__ __forceinline__ uint64_t rational_bits(double x, uint32_t w) {
    double scaled = (w >= 64) ? x * 2.0 * 9223372036854775808.0 : x * (double)(uint64_t(1) << w);
    uint64_t v = scaled <= 0 ? 0 : (uint64_t)scaled;
    return v & fmask(w);
}

__device__ __forceinline__ uint64_t isqrt_u64(uint64_t v) {
    uint64_t r = (uint64_t)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

__device__ __forceinline__ uint64_t table_word(const void* t, uint32_t bytes, uint64_t i) {
    switch (bytes) {
        case 1: return ((const uint8_t*)t)[i];
        case 2: return ((const uint16_t*)t)[i];
        case 4: return ((const uint32_t*)t)[i];
        default: return ((const uint64_t*)t)[i];
    }
}

constexpr double kPi = 3.141592653589793238462643383279502884;

// Binary search of `target` in table[l, l+len) with `steps` halvings; the hit
// address, or 0 when no probed word equals the target
// (QuantumBinarySearch_Fast::binary_search, SparQ_Algorithm/src/hamiltonian_simulation.cpp:182-199).
__device__ __forceinline__ uint64_t table_search(const void* t, uint32_t bytes, uint64_t mask, uint64_t l, uint64_t len,
                                                 uint64_t steps, uint64_t target) {
    uint64_t r = l + len;
    for (uint64_t s = 0; s < steps; ++s) {
        const uint64_t mid = (l + r) / 2;
        const uint64_t v = table_word(t, bytes, mid & mask);
        if (v == target) return mid;
        if (v < target) l = mid;
        else r = mid;
    }
    return 0;
}

// The tridiagonal matrix of the direct walk oracles: four slots per row over
// the column window [base(j), base(j)+4), and the combined memory
// [values of all rows | columns of all rows] of 8n words
// (TridiagonalSpec, SparQ_Algorithm/include/hamiltonian_simulation_tridiagonal.h:104-190).
struct Tridiag {
    uint64_t n, alpha, beta, kmask;
    __device__ __forceinline__ explicit Tridiag(const sparq_pw_op_t& op)
        : n(op.c[0] & ((uint64_t(1) << 48) - 1)), alpha(op.c[1]), beta(op.c[2]), kmask(fmask((uint32_t)(op.c[0] >> 48))) {}
    __device__ __forceinline__ uint64_t base(uint64_t j) const {
        const uint64_t b = j > 0 ? j - 1 : 0;
        return b < n - 4 ? b : n - 4;
    }
    __device__ __forceinline__ uint64_t value(uint64_t j, uint64_t s) const {
        const uint64_t c = base(j) + s;
        if (c == j) return alpha;
        if (c + 1 == j || c == j + 1) return beta;
        return 0;
    }
    __device__ __forceinline__ uint64_t mem(uint64_t i) const {
        i &= 8 * n - 1;
        if (i < 4 * n) return value(i >> 2, i & 3);
        const uint64_t t = i - 4 * n;
        return base(t >> 2) + (t & 3);
    }
};

// One multiplication by the phase per set bit of `bits`, lowest first, as one
// Phase_Bool per digit does (basic_gates.cpp:11-31).  For a phase (-1, s), Z's
// (cos pi, sin pi), a*(-1, s) = (-a.x - a.y*s, a.x*s - a.y) is the same IEEE
// result as the general complex product (x*(-1) is exact and x + (-y) ==
// x - y) with two fewer multiplications.
__device__ __forceinline__ void phase_per_bit(double2& amp, uint64_t bits, const double2 ph) {
    if (ph.x == -1.0) {
        const double sn = ph.y;
        while (bits) {
            amp = make_double2(-amp.x - amp.y * sn, amp.x * sn - amp.y);
            bits &= bits - 1;
        }
    } else {
        while (bits) {
            amp = cmul(amp, ph);
            bits &= bits - 1;
        }
    }
}

// LIGHT: the heavy opcodes (is_heavy_opcode) compile to nothing, so a program
// of light opcodes runs with far fewer registers per lane.
// pre0 / pre1: the program's prefetched table words (an op with pad = s + 1
// reads word s instead of loading it).
template <int W, bool LIGHT>
__device__ __forceinline__ void run_op(const sparq_pw_op_t& op, uint64_t (&k)[W], double2& amp, uint64_t pre0, uint64_t pre1) {
    const sparq_field_t f0 = op.f[0], f1 = op.f[1], f2 = op.f[2], f3 = op.f[3];
    switch (op.opcode) {
        case SPARQ_PW_FLIP_BIT: fput<W>(k, f0, fget<W>(k, f0) ^ (uint64_t(1) << op.c[0])); break;
        case SPARQ_PW_FLIP_ALL: fput<W>(k, f0, ~fget<W>(k, f0)); break;
        case SPARQ_PW_SWAP_BITS: {
            uint64_t a = fget<W>(k, f0), b = fget<W>(k, f1);
            uint64_t ba = (a >> op.c[0]) & 1, bb = (b >> op.c[1]) & 1;
            a = (a & ~(uint64_t(1) << op.c[0])) | (bb << op.c[0]);
            b = (b & ~(uint64_t(1) << op.c[1])) | (ba << op.c[1]);
            fput<W>(k, f0, a);
            fput<W>(k, f1, b);
            break;
        }
        case SPARQ_PW_SWAP_FIELDS: {
            uint64_t a = fget<W>(k, f0), b = fget<W>(k, f1);
            fput<W>(k, f0, b);
            fput<W>(k, f1, a);
            break;
        }
        case SPARQ_PW_ROTL: {
            uint32_t size = f0.width, d = (uint32_t)op.c[0];
            uint64_t v = fget<W>(k, f0);
            uint64_t high = (size - d) >= 64 ? 0 : v >> (size - d);
            uint64_t low = v - ((size - d) >= 64 ? 0 : (high << (size - d)));
            fput<W>(k, f0, (d >= 64 ? 0 : (low << d)) + high);
            break;
        }
        case SPARQ_PW_ROTR: {
            uint32_t size = f0.width, d = (uint32_t)op.c[0];
            uint64_t v = fget<W>(k, f0);
            uint64_t high = d >= 64 ? 0 : v >> d;
            uint64_t low = v - (d >= 64 ? 0 : (high << d));
            fput<W>(k, f0, ((size - d) >= 64 ? 0 : (low << (size - d))) + high);
            break;
        }
        case SPARQ_PW_XOR_MUL_CONST: fput<W>(k, f1, fget<W>(k, f1) ^ (fget<W>(k, f0) * op.c[0])); break;
        case SPARQ_PW_ADD_MUL_CONST: fput<W>(k, f1, fget<W>(k, f1) + op.c[0] * fget<W>(k, f0)); break;
        case SPARQ_PW_MODMUL_CONST:
            if constexpr (!LIGHT) {
                fput<W>(k, f0, (fget<W>(k, f0) * op.c[0]) % op.c[1]);
            }
            break;
        case SPARQ_PW_XOR_ADD: fput<W>(k, f2, fget<W>(k, f2) ^ (fget<W>(k, f0) + fget<W>(k, f1))); break;
        case SPARQ_PW_ADD_FIELD: fput<W>(k, f1, fget<W>(k, f1) + fget<W>(k, f0)); break;
        case SPARQ_PW_SUB_FIELD: fput<W>(k, f1, fget<W>(k, f1) - fget<W>(k, f0)); break;
        case SPARQ_PW_XOR_ADD_CONST: fput<W>(k, f1, fget<W>(k, f1) ^ (fget<W>(k, f0) + op.c[0])); break;
        case SPARQ_PW_ADD_CONST: fput<W>(k, f0, fget<W>(k, f0) + op.c[0]); break;
        case SPARQ_PW_XOR_SUB: fput<W>(k, f2, fget<W>(k, f2) ^ (fget<W>(k, f0) - fget<W>(k, f1))); break;
        case SPARQ_PW_XOR_NEG: fput<W>(k, f1, fget<W>(k, f1) ^ (uint64_t(0) - fget<W>(k, f0))); break;
        case SPARQ_PW_XOR_ABS: {
            int64_t v = sext(fget<W>(k, f0), f0.width);
            uint64_t mag = v < 0 ? (uint64_t)(-v) : (uint64_t)v;
            fput<W>(k, f1, fget<W>(k, f1) ^ mag);
            break;
        }
        case SPARQ_PW_XOR_MUL: fput<W>(k, f2, fget<W>(k, f2) ^ (fget<W>(k, f0) * fget<W>(k, f1))); break;
        case SPARQ_PW_XOR_DIV:
            if constexpr (!LIGHT) {
                uint64_t a = fget<W>(k, f0), b = fget<W>(k, f1);
                fput<W>(k, f2, fget<W>(k, f2) ^ (b == 0 ? 0 : a / b));
            }
            break;
        case SPARQ_PW_XOR_ISQRT:
            if constexpr (!LIGHT) {
                fput<W>(k, f1, fget<W>(k, f1) ^ isqrt_u64(fget<W>(k, f0)));
            }
            break;
        case SPARQ_PW_XOR_SELECT: {
            uint64_t sel = (fget<W>(k, f0) & 1) ? fget<W>(k, f1) : fget<W>(k, f2);
            fput<W>(k, f3, fget<W>(k, f3) ^ sel);
            break;
        }
        case SPARQ_PW_XOR_AND: fput<W>(k, f2, fget<W>(k, f2) ^ (fget<W>(k, f0) & fget<W>(k, f1))); break;
        case SPARQ_PW_XOR_OR: fput<W>(k, f2, fget<W>(k, f2) ^ (fget<W>(k, f0) | fget<W>(k, f1))); break;
        case SPARQ_PW_XOR_XOR: fput<W>(k, f2, fget<W>(k, f2) ^ (fget<W>(k, f0) ^ fget<W>(k, f1))); break;
        case SPARQ_PW_FLAG_LESS_S: {
            bool p = sext(fget<W>(k, f0), f0.width) < sext(fget<W>(k, f1), f1.width);
            fput<W>(k, f2, fget<W>(k, f2) ^ (p ? 1 : 0));
            break;
        }
        case SPARQ_PW_FLAG_CARRY: {
            uint64_t a = fget<W>(k, f0), b = fget<W>(k, f1);
            uint32_t w = f2.width;
            bool p = w >= 64 ? (a + b < a) : (a >= (uint64_t(1) << w)) || (b >= (uint64_t(1) << w) - a);
            fput<W>(k, f3, fget<W>(k, f3) ^ (p ? 1 : 0));
            break;
        }
        case SPARQ_PW_FLAG_OVERFLOW_S: {
            uint32_t w = f2.width;
            uint64_t m = fmask(w);
            uint64_t A = (uint64_t)sext(fget<W>(k, f0), f0.width) & m;
            uint64_t B = (uint64_t)sext(fget<W>(k, f1), f1.width) & m;
            uint64_t S = (A + B) & m;
            uint64_t sa = (A >> (w - 1)) & 1, sb = (B >> (w - 1)) & 1, ss = (S >> (w - 1)) & 1;
            fput<W>(k, f3, fget<W>(k, f3) ^ (((sa == sb) && (ss != sa)) ? 1 : 0));
            break;
        }
        case SPARQ_PW_FLAG_MULOVERFLOW:
            if constexpr (!LIGHT) {
                uint64_t a = fget<W>(k, f0), b = fget<W>(k, f1);
                uint64_t lo = a * b, hi = __umul64hi(a, b);
                uint32_t w = f2.width;
                bool p = hi != 0 || (w < 64 && lo >= (uint64_t(1) << w));
                fput<W>(k, f3, fget<W>(k, f3) ^ (p ? 1 : 0));
            }
            break;
        case SPARQ_PW_FLAG_ISZERO: fput<W>(k, f1, fget<W>(k, f1) ^ (fget<W>(k, f0) == 0 ? 1 : 0)); break;
        case SPARQ_PW_FLAG_NEGATIVE: fput<W>(k, f1, fget<W>(k, f1) ^ (sext(fget<W>(k, f0), f0.width) < 0 ? 1 : 0)); break;
        case SPARQ_PW_ADD_ANY:
        case SPARQ_PW_SUB_ANY: {
            uint64_t r = fget<W>(k, f1);
            uint64_t e = (f1.flags & 1) ? (uint64_t)sext(r, f1.width) : r;
            uint64_t v = fget<W>(k, f0);
            fput<W>(k, f0, op.opcode == SPARQ_PW_ADD_ANY ? v + e : v - e);
            break;
        }
        case SPARQ_PW_XOR_ASSIGN: fput<W>(k, f1, fget<W>(k, f1) ^ fget<W>(k, f0)); break;
        case SPARQ_PW_COMPARE: {
            uint64_t a = fget<W>(k, f0), b = fget<W>(k, f1);
            if (a == b) fput<W>(k, f3, fget<W>(k, f3) ^ 1);
            else fput<W>(k, f2, fget<W>(k, f2) ^ (a < b ? 1 : 0));
            break;
        }
        case SPARQ_PW_FLAG_LESS_U: fput<W>(k, f2, fget<W>(k, f2) ^ (fget<W>(k, f0) < fget<W>(k, f1) ? 1 : 0)); break;
        case SPARQ_PW_XOR_MID: fput<W>(k, f2, fget<W>(k, f2) ^ ((fget<W>(k, f0) + fget<W>(k, f1)) / 2)); break;
        case SPARQ_PW_XOR_DIVSQRT_ACOS:
            if constexpr (!LIGHT) {
                double o = acos(sqrt(1.0 * fget<W>(k, f0) / fget<W>(k, f1))) / kPi / 2;
                fput<W>(k, f2, fget<W>(k, f2) ^ rational_bits(o, f2.width));
            }
            break;
        case SPARQ_PW_XOR_SQRTDIV_ACOS:
            if constexpr (!LIGHT) {
                double o = acos(sext(fget<W>(k, f0), f0.width) / sqrt((double)fget<W>(k, f1))) / kPi / 2;
                fput<W>(k, f2, fget<W>(k, f2) ^ rational_bits(o, f2.width));
            }
            break;
        case SPARQ_PW_XOR_ROTATE_ANGLE:
            if constexpr (!LIGHT) {
                int64_t l = sext(fget<W>(k, f0), f0.width), r = sext(fget<W>(k, f1), f1.width);
                double o;
                if (l == 0 && r >= 0) o = 0.25;
                else if (l == 0 && r < 0) o = 0.75;
                else {
                    o = atan2((double)r, (double)l) / kPi / 2;
                    if (o < 0) o += 1;
                }
                fput<W>(k, f2, fget<W>(k, f2) ^ rational_bits(o, f2.width));
            }
            break;
        case SPARQ_PW_QRAM_LOAD: {
            uint64_t w;
            if (op.pad == 1) w = pre0;
            else if (op.pad == 2) w = pre1;
            else w = table_word(op.table, op.table_word_bytes, fget<W>(k, f0) & fmask((uint32_t)op.c[0]));
            fput<W>(k, f1, fget<W>(k, f1) ^ w);
            break;
        }
        case SPARQ_PW_XOR_AFFINE: {
            uint64_t a = fget<W>(k, f0) + op.c[0] * fget<W>(k, f1);
            if (f2.width) a += fget<W>(k, f2);
            fput<W>(k, f3, fget<W>(k, f3) ^ a);
            break;
        }
        case SPARQ_PW_XOR_BSEARCH:
            if constexpr (!LIGHT) {
                const uint64_t hit = table_search(op.table, op.table_word_bytes, op.c[2], fget<W>(k, f0), op.c[0], op.c[1], fget<W>(k, f1));
                fput<W>(k, f2, fget<W>(k, f2) ^ hit);
            }
            break;
        case SPARQ_PW_XOR_QW_ANGLE:
            if constexpr (!LIGHT) {
                const uint32_t ds = (uint32_t)op.c[0];
                uint64_t v = fget<W>(k, f0);
                double ratio;
                if (op.c[1]) {
                    const uint64_t amax = fmask(ds);
                    ratio = (double)(v & amax) * 1.0 / (double)amax;
                } else {
                    const int64_t vr = sext(v & fmask(ds), ds);
                    ratio = (double)(vr < 0 ? -vr : vr) * 1.0 / (double)fmask(ds - 1);
                }
                ratio = fmax(0.0, fmin(1.0, ratio));
                const double o = acos(sqrt(ratio)) / kPi / 2;
                fput<W>(k, f1, fget<W>(k, f1) ^ rational_bits(o, f1.width));
            }
            break;
        case SPARQ_PW_TRIDIAG_LOAD:
            if constexpr (!LIGHT) {
                const Tridiag m(op);
                fput<W>(k, f3, fget<W>(k, f3) ^ m.mem(fget<W>(k, f0) + 4 * fget<W>(k, f1) + fget<W>(k, f2)));
            }
            break;
        case SPARQ_PW_TRIDIAG_COL2SLOT:
            if constexpr (!LIGHT) {
                const Tridiag m(op);
                const uint64_t j = fget<W>(k, f1), col = fget<W>(k, f2), sr = fget<W>(k, f3);
                const uint64_t base = m.base(j), row_addr = fget<W>(k, f0) + 4 * j;
                const bool hit = col >= base && col < base + 4;
                const uint64_t sr1 = sr ^ (hit ? row_addr + (col - base) : 0);
                const uint64_t k1 = col ^ m.mem(sr1);
                fput<W>(k, f2, (sr1 - row_addr) & m.kmask);
                fput<W>(k, f3, k1 & m.kmask);
            }
            break;
        case SPARQ_PW_TRIDIAG_SLOT2COL:
            if constexpr (!LIGHT) {
                const Tridiag m(op);
                const uint64_t j = fget<W>(k, f1), slot = fget<W>(k, f2), sr = fget<W>(k, f3);
                const uint64_t base = m.base(j), row_addr = fget<W>(k, f0) + 4 * j;
                const uint64_t k1 = (slot + row_addr) & m.kmask;
                const uint64_t k2 = sr ^ m.mem(k1);
                const bool hit = k2 >= base && k2 < base + 4;
                fput<W>(k, f2, k2);
                fput<W>(k, f3, k1 ^ (hit ? row_addr + (k2 - base) : 0));
            }
            break;
        case SPARQ_PW_PHASE_BITS: phase_per_bit(amp, fget<W>(k, f0) & op.c[0], make_double2(op.d[0], op.d[1])); break;
        case SPARQ_PW_ASSERT_ZERO:
            if (fget<W>(k, f0) != 0) atomicOr((unsigned long long*)op.table, (unsigned long long)kCapNonzeroRemoved);
            fput<W>(k, f0, 0);
            break;
        case SPARQ_PW_PAULI_TABLE: {
            const uint64_t* t = (const uint64_t*)op.table;
            const uint64_t it = t[0], s = fget<W>(k, f0);
            if (it < op.c[1] && s < op.c[0]) {
                const uint64_t row = 1 + 2 * (it * op.c[0] + s);
                const uint64_t v = fget<W>(k, f1) ^ t[row];
                fput<W>(k, f1, v);
                phase_per_bit(amp, v & t[row + 1], make_double2(op.d[0], op.d[1]));
            }
            break;
        }
        case SPARQ_PW_RZ: {
            bool b = (fget<W>(k, f0) >> op.c[0]) & 1;
            amp = cmul(amp, b ? make_double2(op.d[0], op.d[1]) : make_double2(op.d[2], op.d[3]));
            break;
        }
        case SPARQ_PW_Y: {
            uint64_t v = fget<W>(k, f0) ^ (uint64_t(1) << op.c[0]);
            fput<W>(k, f0, v);
            amp = cmul(amp, ((v >> op.c[0]) & 1) ? make_double2(0.0, 1.0) : make_double2(0.0, -1.0));
            break;
        }
        case SPARQ_PW_ROT_DIAG: {
            bool b = (fget<W>(k, f0) >> op.c[0]) & 1;
            amp = cmul(amp, b ? make_double2(op.d[2], op.d[3]) : make_double2(op.d[0], op.d[1]));
            break;
        }
        case SPARQ_PW_ROT_OFFDIAG: {
            uint64_t v = fget<W>(k, f0);
            bool b = (v >> op.c[0]) & 1;
            amp = cmul(amp, b ? make_double2(op.d[2], op.d[3]) : make_double2(op.d[0], op.d[1]));
            fput<W>(k, f0, v ^ (uint64_t(1) << op.c[0]));
            break;
        }
        case SPARQ_PW_NEG_AMP: amp = make_double2(amp.x * -1.0, amp.y * -1.0); break;
        case SPARQ_PW_MUL_AMP: amp = cmul(amp, make_double2(op.d[0], op.d[1])); break;
        case SPARQ_PW_SET_FIELD: fput<W>(k, f0, op.c[0]); break;
        case SPARQ_PW_INC_WRAP: {
            uint64_t v = fget<W>(k, f0);
            if (v == fmask(f0.width)) {
                fput<W>(k, f1, (fget<W>(k, f1) + 1) % 2);
                fput<W>(k, f0, 0);
            } else {
                fput<W>(k, f0, v + 1);
            }
            break;
        }
        case SPARQ_PW_DEC_WRAP: {
            uint64_t v = fget<W>(k, f0);
            if (v == 0) {
                fput<W>(k, f1, (fget<W>(k, f1) + 1) % 2);
                fput<W>(k, f0, fmask(f0.width));
            } else {
                fput<W>(k, f0, v - 1);
            }
            break;
        }
        default: break;
    }
}

// Entries per thread when the launch-shape autotuner has no measurement (and
// for captured sequences): one for one- and two-word keys, where the
// interpreter's registers for several entries cut the waves per SIMD more than
// the extra loads in flight gain (23 vs 27 us for the 2^20 QRAM-load +
// 32-phase program with the full interpreter), two for three and four words.
template <int W>
constexpr int kPwItems() { return (W == 3 || W == 4) ? 2 : 1; }
// Entries per thread the autotuner may try.
constexpr bool pw_items_allowed(int W, int items) { return items == 1 || items == 2 || (items == 4 && W <= 2); }

// DEV: the entry count is read from n_dev and the grid strides over it (a
// persistent grid; captured sequences); otherwise one round over n.
template <int W, bool AMP, bool LIGHT, bool DEV, int ITEMS>
__global__ void __launch_bounds__(kBlock) pointwise_kernel(uint64_t* __restrict__ keys, double2* __restrict__ amps, uint64_t cap,
                                                      uint64_t n_arg, const uint64_t* __restrict__ n_dev, const PwProgram prog) {
    // ITEMS entries per thread, striped by the block size.
    constexpr int kI = ITEMS;
    const uint64_t n = DEV ? n_dev[0] : n_arg;
    for (uint64_t blk = blockIdx.x; blk * kBlock * kI < n; blk += gridDim.x) {
    const uint64_t base = blk * kBlock * kI + threadIdx.x;
    uint64_t k[kI][W];
    double2 a[kI];
#pragma unroll
    for (int it = 0; it < kI; ++it) {
        const uint64_t i = base + (uint64_t)it * kBlock;
        a[it] = make_double2(0.0, 0.0);
        if (i < n) {
            load_key<W>(keys, cap, i, k[it]);
            if (AMP) a[it] = amps[i];
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) k[it][w] = 0;
        }
    }
    // prefetched table words: every load of the entry in flight together
    uint64_t p0[kI], p1[kI];
#pragma unroll
    for (int it = 0; it < kI; ++it) {
        p0[it] = prog.npre > 0 ? table_word(prog.pre[0].table, prog.pre[0].bytes, fget<W>(k[it], prog.pre[0].f) & prog.pre[0].mask) : 0;
        p1[it] = prog.npre > 1 ? table_word(prog.pre[1].table, prog.pre[1].bytes, fget<W>(k[it], prog.pre[1].f) & prog.pre[1].mask) : 0;
    }
    for (uint32_t o = 0; o < prog.nops; ++o) {
        const sparq_pw_op_t& op = prog.ops[o];
#pragma unroll
        for (int it = 0; it < kI; ++it) {
            bool ok = true;
            for (uint32_t c = 0; c < op.ncond; ++c) ok = ok && cond_ok<W>(k[it], prog.conds[op.cond_first + c]);
            if (ok) run_op<W, LIGHT>(op, k[it], a[it], p0[it], p1[it]);
        }
    }
#pragma unroll
    for (int it = 0; it < kI; ++it) {
        const uint64_t i = base + (uint64_t)it * kBlock;
        if (i >= n) break;
        store_key<W>(keys, cap, i, k[it]);
        if (AMP) amps[i] = a[it];
    }
    if (!DEV) break;
    }
}

// Opcodes with 64-bit division, square roots, arc cosines or table searches:
// their registers set the occupancy of every program that may contain them.
static bool is_heavy_opcode(uint32_t opcode) {
    switch (opcode) {
        case SPARQ_PW_MODMUL_CONST: case SPARQ_PW_XOR_DIV: case SPARQ_PW_XOR_ISQRT: case SPARQ_PW_FLAG_MULOVERFLOW:
        case SPARQ_PW_XOR_DIVSQRT_ACOS: case SPARQ_PW_XOR_SQRTDIV_ACOS: case SPARQ_PW_XOR_ROTATE_ANGLE:
        case SPARQ_PW_XOR_BSEARCH: case SPARQ_PW_XOR_QW_ANGLE: case SPARQ_PW_TRIDIAG_LOAD: case SPARQ_PW_TRIDIAG_COL2SLOT:
        case SPARQ_PW_TRIDIAG_SLOT2COL:
            return true;
        default:
            return false;
    }
}

// Whether an op may write its operand field f (conservatively: every operand
// field of an op not known to write less).
static bool may_write_field(const sparq_pw_op_t& op, int f) {
    switch (op.opcode) {
        case SPARQ_PW_PHASE_BITS: case SPARQ_PW_RZ: case SPARQ_PW_ROT_DIAG: case SPARQ_PW_NEG_AMP: case SPARQ_PW_MUL_AMP:
            return false;
        case SPARQ_PW_QRAM_LOAD:
            return f == 1;
        default:
            return op.f[f].width != 0;
    }
}
static bool fields_overlap(sparq_field_t a, sparq_field_t b) {
    return a.word == b.word && a.shift < (uint32_t)b.shift + b.width && b.shift < (uint32_t)a.shift + a.width;
}

// QRAM loads whose address bits no earlier op of the program may write read
// their table word before the first op runs (up to kPwPrefetch distinct
// words; loads of the same word share one): every gather of an entry is then
// in flight together instead of one after another.
static void plan_prefetch(PwProgram& prog) {
    prog.npre = 0;
    for (uint32_t o = 0; o < prog.nops; ++o) {
        sparq_pw_op_t& op = prog.ops[o];
        op.pad = 0;
        if (op.opcode != SPARQ_PW_QRAM_LOAD) continue;
        bool stable = true;
        for (uint32_t e = 0; e < o && stable; ++e)
            for (int f = 0; f < 4 && stable; ++f)
                if (may_write_field(prog.ops[e], f) && fields_overlap(prog.ops[e].f[f], op.f[0])) stable = false;
        if (!stable) continue;
        const uint64_t mask = fmask_host((uint32_t)op.c[0]);
        uint32_t slot = prog.npre;
        for (uint32_t s = 0; s < prog.npre; ++s)
            if (prog.pre[s].table == op.table && prog.pre[s].bytes == op.table_word_bytes && prog.pre[s].mask == mask &&
                std::memcmp(&prog.pre[s].f, &op.f[0], sizeof(sparq_field_t)) == 0)
                slot = s;
        if (slot == (uint32_t)kPwPrefetch) continue;
        if (slot == prog.npre) {
            prog.pre[slot] = PwPrefetch{op.table, mask, op.f[0], op.table_word_bytes};
            ++prog.npre;
        }
        op.pad = slot + 1;
    }
}

// Whether a program keeps the entries' grouping in pairs across key bit
// `bit` (word * 64 + shift): every op that may change keys reads and writes
// only fields, and evaluates only conditions, that leave the bit out.  Such
// an op maps a key and its partner to a key and its partner, in place and in
// the same order; ops that change amplitudes alone keep it whatever they read.
static bool keeps_pairs(const PwProgram& prog, int32_t bit) {
    const sparq_field_t b{(uint8_t)(bit / 64), (uint8_t)(bit % 64), 1, 0};
    for (uint32_t o = 0; o < prog.nops; ++o) {
        const sparq_pw_op_t& op = prog.ops[o];
        bool keys_change = false;
        for (int f = 0; f < 4; ++f) keys_change = keys_change || may_write_field(op, f);
        if (!keys_change) continue;
        for (int f = 0; f < 4; ++f)
            if (op.f[f].width && fields_overlap(op.f[f], b)) return false;
        for (uint32_t c = 0; c < op.ncond; ++c)
            if (fields_overlap(prog.conds[op.cond_first + c].field, b)) return false;
    }
    return true;
}

static bool touches_amp(uint32_t opcode) {
    switch (opcode) {
        case SPARQ_PW_PHASE_BITS: case SPARQ_PW_RZ: case SPARQ_PW_Y: case SPARQ_PW_ROT_DIAG:
        case SPARQ_PW_ROT_OFFDIAG: case SPARQ_PW_NEG_AMP: case SPARQ_PW_MUL_AMP: case SPARQ_PW_PAULI_TABLE:
            return true;
        default:
            return false;
    }
}

template <int W, bool DEV, int ITEMS>
static void launch_pointwise_items(State& st, const PwProgram& prog, bool amp, bool light, unsigned g, const uint64_t* n_dev) {
    if (light) {
        if (amp) pointwise_kernel<W, true, true, DEV, ITEMS><<<g, kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, st.n, n_dev, prog);
        else pointwise_kernel<W, false, true, DEV, ITEMS><<<g, kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, st.n, n_dev, prog);
    } else {
        if (amp) pointwise_kernel<W, true, false, DEV, ITEMS><<<g, kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, st.n, n_dev, prog);
        else pointwise_kernel<W, false, false, DEV, ITEMS><<<g, kBlock, 0, st.stream>>>(st.keys, st.amps, st.cap, st.n, n_dev, prog);
    }
    SPQ_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- launch-shape autotuner
// Entries per thread (1, 2 or 4) are chosen per class (key words, amplitude
// access, light / full interpreter, log2 N) from N >= 2^16 on: the first two
// calls of a class run the default shape unmeasured (first allocations, cold
// caches), the next ones cycle twice through the candidates with the kernel
// alone timed by HIP events and normalised per entry, keeping each
// candidate's faster sample; later calls take the fastest, the default unless
// another is at least 5 % faster.  Every call does its real work, whichever
// shape runs.
namespace {
constexpr int kPwCandidates[3] = {1, 2, 4};
constexpr int kPwWarmCalls = 2, kPwRounds = 2;
struct PwTune {
    int calls = 0;
    float ms[3] = {-1.f, -1.f, -1.f};   // ms per 2^20 entries, the faster sample
    int best = -1;
};
std::mutex g_pw_mu;
std::map<uint32_t, PwTune> g_pw_tune;
int g_pw_forced = 0;
}  // namespace

void set_pointwise_items(int items) {
    std::lock_guard<std::mutex> lock(g_pw_mu);
    g_pw_forced = items;
}

std::string pointwise_tuning_report() {
    std::lock_guard<std::mutex> lock(g_pw_mu);
    std::string out = "[";
    for (auto& kv : g_pw_tune) {
        const uint32_t k = kv.first;
        char buf[256];
        snprintf(buf, sizeof(buf), "%s{\"words\":%u,\"amp\":%u,\"light\":%u,\"log2_n\":%u,\"items\":%d,\"ms_per_2p20\":[%.5f,%.5f,%.5f]}",
                 out.size() > 1 ? "," : "", k & 15, (k >> 4) & 1, (k >> 5) & 1, k >> 8,
                 kv.second.best >= 0 ? kPwCandidates[kv.second.best] : 0, kv.second.ms[0], kv.second.ms[1], kv.second.ms[2]);
        out += buf;
    }
    return out + "]";
}

template <int W, bool DEV>
static void launch_pointwise_impl(State& st, const PwProgram& prog, bool amp, bool light, int items, const uint64_t* n_dev,
                                  uint64_t span) {
    const unsigned g = DEV ? persistent_grid(span, kBlock * items) : grid_for(span, kBlock * items);
    if (items == 4) {
        if constexpr (pw_items_allowed(W, 4)) {
            launch_pointwise_items<W, DEV, 4>(st, prog, amp, light, g, n_dev);
            return;
        }
    }
    if (items == 2) launch_pointwise_items<W, DEV, 2>(st, prog, amp, light, g, n_dev);
    else launch_pointwise_items<W, DEV, 1>(st, prog, amp, light, g, n_dev);
}

template <int W>
static void launch_pointwise(State& st, const PwProgram& prog, bool amp, bool light) {
    if (st.capture) {   // recording a captured sequence: the count is on the device
        const uint64_t* n_dev = capture_words(st) + kCapN;
        launch_pointwise_impl<W, true>(st, prog, amp, light, kPwItems<W>(), n_dev, st.capture->capacity);
        return;
    }
    int items = kPwItems<W>(), measure = -1;
    {
        std::lock_guard<std::mutex> lock(g_pw_mu);
        if (g_pw_forced) {
            items = pw_items_allowed(W, g_pw_forced) ? g_pw_forced : kPwItems<W>();
        } else if (st.n >= (uint64_t(1) << 16)) {
            int lg = 0;
            while ((uint64_t(1) << (lg + 1)) <= st.n && lg < 62) ++lg;
            const uint32_t key = (uint32_t)W | (amp ? 16u : 0u) | (light ? 32u : 0u) | ((uint32_t)lg << 8);
            PwTune& t = g_pw_tune[key];
            const int ncand = pw_items_allowed(W, 4) ? 3 : 2;
            if (t.calls < kPwWarmCalls) {
                ++t.calls;
            } else if (t.calls < kPwWarmCalls + kPwRounds * ncand) {
                measure = (t.calls - kPwWarmCalls) % ncand;
                items = kPwCandidates[measure];
                ++t.calls;
            } else if (t.best >= 0) {
                items = kPwCandidates[t.best];
            }
        }
    }
    SPQ_PROF(st.stream, "pointwise_kernel", st.n * (16.0 * W + (amp ? 32.0 : 0.0)));
    if (measure < 0) {
        launch_pointwise_impl<W, false>(st, prog, amp, light, items, nullptr, st.n);
        return;
    }
    // events of this call only: other states may be timed concurrently
    hipEvent_t ev[2];
    SPQ_HIP(hipEventCreate(&ev[0]));
    SPQ_HIP(hipEventCreate(&ev[1]));
    SPQ_HIP(hipEventRecord(ev[0], st.stream));
    launch_pointwise_impl<W, false>(st, prog, amp, light, items, nullptr, st.n);
    SPQ_HIP(hipEventRecord(ev[1], st.stream));
    SPQ_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    SPQ_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    SPQ_HIP(hipEventDestroy(ev[0]));
    SPQ_HIP(hipEventDestroy(ev[1]));
    int lg = 0;
    while ((uint64_t(1) << (lg + 1)) <= st.n && lg < 62) ++lg;
    const uint32_t key = (uint32_t)W | (amp ? 16u : 0u) | (light ? 32u : 0u) | ((uint32_t)lg << 8);
    std::lock_guard<std::mutex> lock(g_pw_mu);
    PwTune& t = g_pw_tune[key];
    const float per = std::max(ms, 1e-6f) / (float)((double)st.n / 1048576.0);
    t.ms[measure] = t.ms[measure] > 0.f ? std::min(t.ms[measure], per) : per;
    const int ncand = pw_items_allowed(W, 4) ? 3 : 2;
    bool all = t.calls >= kPwWarmCalls + kPwRounds * ncand;
    for (int c = 0; c < ncand; ++c) all = all && t.ms[c] > 0.f;
    if (all) {   // a call of another thread still measuring may refine a sample afterwards
        int def = 0;
        while (kPwCandidates[def] != kPwItems<W>()) ++def;
        int best = def;
        for (int c = 0; c < ncand; ++c)
            if (c != def && t.ms[c] < 0.95f * t.ms[def] && t.ms[c] < t.ms[best]) best = c;
        t.best = best;
    }
}

static void check_field(const State& st, sparq_field_t f) {
    SPQ_REQUIRE(f.word < st.W, "field word out of range");
    SPQ_REQUIRE(f.width >= 1 && f.width <= 64 && (uint32_t)f.shift + f.width <= 64, "bad field");
}

void pointwise(State& st, const sparq_pw_op_t* ops, uint32_t nops, const sparq_cond_t* conds, uint32_t nconds) {
    SPQ_REQUIRE(nops <= SPARQ_MAX_PROGRAM_OPS, "too many ops in one pointwise program");
    SPQ_REQUIRE(nconds <= SPARQ_MAX_PROGRAM_CONDS, "too many conditions in one pointwise program");
    PwProgram prog{};
    prog.nops = nops;
    prog.nconds = nconds;
    bool amp = false, light = true;
    for (uint32_t o = 0; o < nops; ++o) {
        const sparq_pw_op_t& op = ops[o];
        SPQ_REQUIRE(op.opcode >= SPARQ_PW_FLIP_BIT && op.opcode < SPARQ_PW_OPCODE_END, "unknown pointwise opcode");
        SPQ_REQUIRE((uint32_t)op.cond_first + op.ncond <= nconds, "condition range out of bounds");
        for (int f = 0; f < 4; ++f)
            if (op.f[f].width) check_field(st, op.f[f]);
        if (op.opcode == SPARQ_PW_TRIDIAG_LOAD || op.opcode == SPARQ_PW_TRIDIAG_COL2SLOT || op.opcode == SPARQ_PW_TRIDIAG_SLOT2COL) {
            const uint64_t n = op.c[0] & ((uint64_t(1) << 48) - 1), kw = op.c[0] >> 48;
            SPQ_REQUIRE(n >= 4 && (n & (n - 1)) == 0, "tridiagonal oracle needs a power-of-two dimension >= 4");
            SPQ_REQUIRE(kw >= 1 && kw <= 64, "tridiagonal oracle: bad column width");
        }
        if (op.opcode == SPARQ_PW_XOR_QW_ANGLE) SPQ_REQUIRE(op.c[0] >= 2 && op.c[0] <= 63, "walk angle: data size out of range");
        if (op.opcode == SPARQ_PW_ASSERT_ZERO) SPQ_REQUIRE(st.capture, "the zero assertion runs only inside a captured sequence");
        if (op.opcode == SPARQ_PW_PAULI_TABLE) {
            SPQ_REQUIRE(op.table != nullptr, "Pauli noise without table");
            SPQ_REQUIRE(op.c[0] >= 1 && op.c[1] >= 1, "Pauli noise table needs rows and iterations");
        }
        if (op.opcode == SPARQ_PW_QRAM_LOAD || op.opcode == SPARQ_PW_XOR_BSEARCH) {
            SPQ_REQUIRE(op.table != nullptr, "QRAM load without table");
            SPQ_REQUIRE(op.table_word_bytes == 1 || op.table_word_bytes == 2 || op.table_word_bytes == 4 ||
                            op.table_word_bytes == 8, "bad table word size");
        }
        amp = amp || touches_amp(op.opcode);
        light = light && !is_heavy_opcode(op.opcode);
        prog.ops[o] = op;
        if (op.opcode == SPARQ_PW_ASSERT_ZERO && capture_words(st)) prog.ops[o].table = capture_words(st) + kCapFlags;
    }
    for (uint32_t c = 0; c < nconds; ++c) {
        check_field(st, conds[c].field);
        prog.conds[c] = conds[c];
    }
    plan_prefetch(prog);
    if (st.pair_bit >= 0 && (st.capture || !keeps_pairs(prog, st.pair_bit))) st.pair_bit = -1;
    if (st.capture && st.capture->phase == 1) return;   // planning: nothing runs
    if ((st.n == 0 && !st.capture) || nops == 0) return;
    SPQ_DISPATCH_W(st.W, launch_pointwise, st, prog, amp, light);
}

}  // namespace sparq
