// CPU model of one LSD digit pass of the engine's radix sort: the block
// histogram (radix_hist_kernel), the digit-major exclusive scan of the block
// counts, and the stable scatter (radix_scatter_kernel), all in
// sparqsim_amd/csrc/engine/sort.hip.  The model restates the kernels' indexing
// statement for statement on 64-lane waves executed in lockstep, with
//   - every LDS and global access bounds-checked,
//   - a race detector on the LDS arrays: between two barriers, a word that one
//     wave writes (plain store) may not be read or written by another wave, and
//     LDS atomics may not meet plain accesses of another wave,
//   - the waves of a block run phase by phase (barrier to barrier) in several
//     orders (forward, reverse, shuffled); every order must give the same
//     output,
//   - every output position written exactly once,
// and it checks the result against std::stable_sort by the digit (compiled
// with PLANT_MISSING_BARRIER or PLANT_RANK_OFF_BY_ONE it must fail).  Two
// re-orderings of the scatter kernel are modelled too: "early_goff" (the
// tile's global digit offsets loaded at kernel start instead of after the
// ranking) and "reads_first" (for full tiles, every LDS read of the write-out
// before the first store).  Built with -fsanitize=address,undefined by
// tests/test_cpu_models.py.
#include <algorithm>
#include <cinttypes>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <string>
#include <vector>

namespace {

constexpr int kTile = 4096;   // kSortTile
constexpr int kWave = 64;

struct Digit {
    uint8_t word, nseg, bits, pad;
    uint8_t shift[3], width[3];
};

uint32_t digit_of(uint64_t k, const Digit& dg) {
    uint32_t v = 0, at = 0;
    for (int s = 0; s < dg.nseg; ++s) {
        v |= (uint32_t)((k >> dg.shift[s]) & ((1ull << dg.width[s]) - 1)) << at;
        at += dg.width[s];
    }
    return v;
}

uint32_t xcd_tile(uint32_t b, uint32_t nb) {
    const uint32_t x = b % 8, j = b / 8, q = nb / 8, r = nb % 8;
    return x * q + std::min(x, r) + j;
}

[[noreturn]] void fail(const std::string& what) {
    std::fprintf(stderr, "MODEL FAILURE: %s\n", what.c_str());
    std::exit(1);
}

// An LDS array with bounds checks and the per-epoch race detector.
template <class T>
struct Lds {
    std::string name;
    std::vector<T> v;
    std::vector<int> writer;            // wave that plain-stored the word this epoch, -1: none
    std::vector<uint64_t> readers;      // waves that read it this epoch (bit mask)
    std::vector<uint64_t> atomics;      // waves that used an atomic on it this epoch
    Lds(const char* n, size_t size) : name(n), v(size), writer(size, -1), readers(size, 0), atomics(size, 0) {}
    void check(size_t i) const {
        if (i >= v.size()) fail(name + ": LDS index " + std::to_string(i) + " out of bounds " + std::to_string(v.size()));
    }
    void barrier() {
        std::fill(writer.begin(), writer.end(), -1);
        std::fill(readers.begin(), readers.end(), 0);
        std::fill(atomics.begin(), atomics.end(), 0);
    }
    T load(size_t i, int wave) {
        check(i);
        if (writer[i] >= 0 && writer[i] != wave) fail(name + ": read of a word another wave stored since the last barrier");
        if (atomics[i] & ~(1ull << wave)) fail(name + ": read of a word another wave updated atomically since the last barrier");
        readers[i] |= 1ull << wave;
        return v[i];
    }
    void store(size_t i, T x, int wave) {
        check(i);
        if (writer[i] >= 0 && writer[i] != wave) fail(name + ": two waves store one word between barriers");
        if (readers[i] & ~(1ull << wave)) fail(name + ": store to a word another wave read since the last barrier");
        if (atomics[i] & ~(1ull << wave)) fail(name + ": store to a word another wave updated atomically");
        writer[i] = wave;
        v[i] = x;
    }
    T atomic_add(size_t i, T x, int wave) {
        check(i);
        if ((writer[i] >= 0 && writer[i] != wave) || (readers[i] & ~(1ull << wave)))
            fail(name + ": atomic on a word another wave accessed plainly since the last barrier");
        atomics[i] |= 1ull << wave;
        const T old = v[i];
        v[i] = old + x;
        return old;
    }
};

template <class T>
struct Global {
    std::string name;
    std::vector<T> v;
    std::vector<uint8_t> written;
    Global(const char* n, size_t size) : name(n), v(size), written(size, 0) {}
    const T& load(uint64_t i) const {
        if (i >= v.size()) fail(name + ": global read out of bounds");
        return v[i];
    }
    void store(uint64_t i, T x) {
        if (i >= v.size()) fail(name + ": global write out of bounds at " + std::to_string(i));
        if (written[i]) fail(name + ": output position written twice");
        written[i] = 1;
        v[i] = x;
    }
};

// One wave's 64 lanes: match_any / ballot / shfl over lane values.
uint64_t match_any(const uint32_t (&d)[kWave], int lane) {
    uint64_t m = 0;
    for (int l = 0; l < kWave; ++l)
        if (d[l] == d[lane]) m |= 1ull << l;
    return m;
}
uint64_t lanemask_lt(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

enum Variant { kCommitted, kEarlyGoff, kReadsFirst };

struct PassArgs {
    int RB, BS;
    Digit dg;
    std::vector<int> wave_order;   // execution order of the waves in every phase
};

// radix_hist_kernel: block-major counts hist[b * kR + d].
std::vector<uint32_t> model_hist(const std::vector<uint64_t>& keys, const PassArgs& a, uint32_t nb) {
    const int kR = 1 << a.RB, NW = a.BS / kWave, NI = kTile / a.BS;
    const uint32_t kInvalid = 1u << a.RB;
    const uint64_t n = keys.size();
    std::vector<uint32_t> out((size_t)nb * kR);
    for (uint32_t b = 0; b < nb; ++b) {
        Lds<uint32_t> hist("hist", kR);
        for (int wave : a.wave_order)
            for (int lane = 0; lane < kWave; ++lane)
                for (int d = wave * kWave + lane; d < kR; d += a.BS) hist.store(d, 0, wave);
        hist.barrier();
        const uint64_t tile = (uint64_t)b * kTile;
        for (int wave : a.wave_order)
            for (int it = 0; it < NI; ++it) {
                uint32_t d[kWave];
                for (int lane = 0; lane < kWave; ++lane) {
                    const uint64_t e = tile + (uint64_t)wave * (kTile / NW) + it * kWave + lane;
                    d[lane] = e < n ? digit_of(keys.at(e < n ? e : n - 1), a.dg) : kInvalid;
                }
                for (int lane = 0; lane < kWave; ++lane) {
                    const uint64_t peers = match_any(d, lane);
                    const int leader = __builtin_ffsll((long long)peers) - 1;
                    if (lane == leader && d[lane] != kInvalid) hist.atomic_add(d[lane], (uint32_t)__builtin_popcountll(peers), wave);
                }
            }
        hist.barrier();
        for (int wave : a.wave_order)
            for (int lane = 0; lane < kWave; ++lane)
                for (int d = wave * kWave + lane; d < kR; d += a.BS) out[(size_t)b * kR + d] = hist.load(d, wave);
    }
    return out;
}

// Transpose to digit-major and exclusive scan: offs[d * nb + b].
std::vector<uint64_t> model_offsets(const std::vector<uint32_t>& block_hist, int kR, uint32_t nb) {
    std::vector<uint64_t> offs((size_t)kR * nb);
    uint64_t run = 0;
    for (int d = 0; d < kR; ++d)
        for (uint32_t b = 0; b < nb; ++b) {
            offs[(size_t)d * nb + b] = run;
            run += block_hist[(size_t)b * kR + d];
        }
    return offs;
}

// radix_scatter_kernel for one block, phase by phase.
void model_scatter_block(uint32_t block, uint32_t nb, const Global<uint64_t>& kin, const Global<uint32_t>* idx_in,
                         const Global<uint64_t>& offs, Global<uint64_t>& kout, Global<uint32_t>& iout, const PassArgs& a,
                         Variant var) {
    const int RB = a.RB, BS = a.BS, kR = 1 << RB, NW = BS / kWave, NI = kTile / BS;
    const int kPerThread = kR >= BS ? kR / BS : 1;
    const uint32_t kInvalid = 1u << RB;
    const uint64_t n = kin.v.size();
    Lds<uint32_t> wcnt("wcnt", (size_t)NW * kR);
    Lds<uint32_t> dstart("dstart", kR);
    Lds<uint64_t> goff("goff", kR);
    Lds<uint64_t> skey("skey", kTile);
    Lds<uint32_t> sidx("sidx", kTile);
    Lds<uint32_t> scan_lds("scan_lds", NW);
    auto barrier = [&] {
        wcnt.barrier(); dstart.barrier(); goff.barrier(); skey.barrier(); sidx.barrier(); scan_lds.barrier();
    };
    const uint32_t tile_id = xcd_tile(block, nb);
    if (tile_id >= nb) fail("xcd_tile out of range");
    const uint64_t tile = (uint64_t)tile_id * kTile;
    auto owner_digit = [&](int tid, int q) { return tid * kPerThread + q; };
    const bool all_owners = kR >= BS;

    // ---- phase 1: clear the wave counters (early_goff: also load the tile's digit offsets)
    for (int wave : a.wave_order)
        for (int lane = 0; lane < kWave; ++lane) {
            const int tid = wave * kWave + lane;
            for (int d = tid; d < kR; d += BS)
                for (int w = 0; w < NW; ++w) wcnt.store((size_t)w * kR + d, 0, wave);
            if (var == kEarlyGoff && (all_owners || tid < kR))
                for (int q = 0; q < kPerThread; ++q) {
                    const int d = owner_digit(tid, q);
                    goff.store(d, offs.load((uint64_t)d * nb + tile_id), wave);
                }
        }
    barrier();

    // ---- phase 2: loads and the stable per-wave ranking
    std::vector<std::vector<uint64_t>> key(BS, std::vector<uint64_t>(NI));
    std::vector<std::vector<uint32_t>> idx(BS, std::vector<uint32_t>(NI)), rank(BS, std::vector<uint32_t>(NI)),
        dig(BS, std::vector<uint32_t>(NI));
    for (int wave : a.wave_order) {
        for (int lane = 0; lane < kWave; ++lane) {
            const int tid = wave * kWave + lane;
            for (int it = 0; it < NI; ++it) {
                const uint64_t e = tile + (uint64_t)wave * (kTile / NW) + it * kWave + lane;
                const uint64_t ec = e < n ? e : n - 1;
                key[tid][it] = kin.load(ec);
                idx[tid][it] = idx_in ? idx_in->load(ec) : (uint32_t)ec;
            }
        }
        std::vector<std::vector<uint32_t>> got(kWave, std::vector<uint32_t>(NI));
        std::vector<std::vector<uint32_t>> before(kWave, std::vector<uint32_t>(NI));
        std::vector<std::vector<int>> leader(kWave, std::vector<int>(NI));
        for (int it = 0; it < NI; ++it) {
            uint32_t d[kWave];
            bool valid[kWave];
            for (int lane = 0; lane < kWave; ++lane) {
                const int tid = wave * kWave + lane;
                const uint64_t e = tile + (uint64_t)wave * (kTile / NW) + it * kWave + lane;
                valid[lane] = e < n;
                d[lane] = valid[lane] ? digit_of(key[tid][it], a.dg) : kInvalid;
                dig[tid][it] = d[lane];
            }
            // one LDS atomic instruction: the leader lanes, in lane order
            for (int lane = 0; lane < kWave; ++lane) {
                const uint64_t peers = match_any(d, lane);
                before[lane][it] = (uint32_t)__builtin_popcountll(peers & lanemask_lt(lane));
                leader[lane][it] = __builtin_ffsll((long long)peers) - 1;
                got[lane][it] = (valid[lane] && before[lane][it] == 0)
                                    ? wcnt.atomic_add((size_t)wave * kR + (d[lane] & (kR - 1)), (uint32_t)__builtin_popcountll(peers), wave)
                                    : 0u;
            }
        }
        for (int it = 0; it < NI; ++it)
            for (int lane = 0; lane < kWave; ++lane)
                rank[wave * kWave + lane][it] = got[leader[lane][it]][it] + before[lane][it];
    }
    barrier();

    // ---- phase 3: digit totals, block exclusive scan (part 1: wave scans, lane 63 publishes)
    std::vector<std::vector<uint32_t>> c(BS, std::vector<uint32_t>((size_t)kPerThread * NW));
    std::vector<uint32_t> sum(BS, 0), inc(BS, 0);
    for (int wave : a.wave_order) {
        for (int lane = 0; lane < kWave; ++lane) {
            const int tid = wave * kWave + lane;
            const bool owner = all_owners || tid < kR;
            for (int q = 0; q < kPerThread; ++q) {
                const int d = owner ? owner_digit(tid, q) : 0;
                for (int w = 0; w < NW; ++w) {
                    c[tid][(size_t)q * NW + w] = owner ? wcnt.load((size_t)w * kR + d, wave) : 0;
                    sum[tid] += c[tid][(size_t)q * NW + w];
                }
            }
        }
        uint32_t run = 0;
        for (int lane = 0; lane < kWave; ++lane) {
            run += sum[wave * kWave + lane];
            inc[wave * kWave + lane] = run;
        }
        scan_lds.store(wave, run, wave);
    }
    barrier();
    // ---- phase 4: every thread reads the wave totals
    std::vector<uint32_t> start(BS);
    for (int wave : a.wave_order)
        for (int lane = 0; lane < kWave; ++lane) {
            const int tid = wave * kWave + lane;
            uint32_t bef = 0;
            for (int w = 0; w < NW; ++w) {
                const uint32_t s = scan_lds.load(w, wave);
                bef += (w < wave) ? s : 0;
            }
            start[tid] = bef + inc[tid] - sum[tid];
        }
    barrier();
    // ---- phase 5: digit starts, per-wave starts, global digit offsets
    for (int wave : a.wave_order)
        for (int lane = 0; lane < kWave; ++lane) {
            const int tid = wave * kWave + lane;
            if (!(all_owners || tid < kR)) continue;
            uint32_t st = start[tid];
            for (int q = 0; q < kPerThread; ++q) {
                const int d = owner_digit(tid, q);
                dstart.store(d, st, wave);
                for (int w = 0; w < NW; ++w) {
                    wcnt.store((size_t)w * kR + d, st, wave);
                    st += c[tid][(size_t)q * NW + w];
                }
                if (var != kEarlyGoff) goff.store(d, offs.load((uint64_t)d * nb + tile_id), wave);
            }
        }
#ifndef PLANT_MISSING_BARRIER
    barrier();
#endif
    // ---- phase 6: items to their tile positions
    for (int wave : a.wave_order)
        for (int lane = 0; lane < kWave; ++lane) {
            const int tid = wave * kWave + lane;
            for (int it = 0; it < NI; ++it) {
                if (dig[tid][it] == kInvalid) continue;
#ifdef PLANT_RANK_OFF_BY_ONE
                const uint32_t pos = wcnt.load((size_t)wave * kR + dig[tid][it], wave) + rank[tid][it] + (it == NI - 1 ? 1u : 0u);
#else
                const uint32_t pos = wcnt.load((size_t)wave * kR + dig[tid][it], wave) + rank[tid][it];
#endif
                skey.store(pos, key[tid][it], wave);
                sidx.store(pos, idx[tid][it], wave);
            }
        }
    barrier();
    // ---- phase 7: write-out
    const uint32_t count = (uint32_t)std::min<uint64_t>((uint64_t)kTile, n - tile);
    for (int wave : a.wave_order)
        for (int lane = 0; lane < kWave; ++lane) {
            const int tid = wave * kWave + lane;
            if (var == kReadsFirst && count == (uint32_t)kTile) {
                std::vector<uint64_t> k(NI);
                std::vector<uint32_t> ix(NI);
                for (int j = 0; j < NI; ++j) {
                    k[j] = skey.load((size_t)tid + (size_t)j * BS, wave);
                    ix[j] = sidx.load((size_t)tid + (size_t)j * BS, wave);
                }
                for (int j = 0; j < NI; ++j) {
                    const uint32_t p = tid + j * BS;
                    const uint32_t dd = digit_of(k[j], a.dg);
                    const uint64_t dst = goff.load(dd, wave) + (p - dstart.load(dd, wave));
                    kout.store(dst, k[j]);
                    iout.store(dst, ix[j]);
                }
                continue;
            }
            for (uint32_t p = tid; p < count; p += BS) {
                const uint64_t k = skey.load(p, wave);
                const uint32_t dd = digit_of(k, a.dg);
                const uint64_t dst = goff.load(dd, wave) + (p - dstart.load(dd, wave));
                kout.store(dst, k);
                iout.store(dst, sidx.load(p, wave));
            }
        }
}

struct Result {
    std::vector<uint64_t> keys;
    std::vector<uint32_t> idx;
};

Result model_pass(const std::vector<uint64_t>& keys, const std::vector<uint32_t>* idx_in, const PassArgs& a, Variant var) {
    const uint64_t n = keys.size();
    const uint32_t nb = (uint32_t)((n + kTile - 1) / kTile);
    const std::vector<uint32_t> bh = model_hist(keys, a, nb);
    Global<uint64_t> offs("offs", 0);
    offs.v = model_offsets(bh, 1 << a.RB, nb);
    Global<uint64_t> kin("keys_in", 0), kout("keys_out", n);
    kin.v = keys;
    Global<uint32_t> iin("idx_in", 0), iout("idx_out", n);
    if (idx_in) iin.v = *idx_in;
    for (uint32_t b = 0; b < nb; ++b) model_scatter_block(b, nb, kin, idx_in ? &iin : nullptr, offs, kout, iout, a, var);
    for (uint64_t i = 0; i < n; ++i)
        if (!kout.written[i]) fail("output position " + std::to_string(i) + " never written");
    return {kout.v, iout.v};
}

Digit make_digit(std::initializer_list<std::pair<int, int>> segs) {
    Digit d{};
    for (auto [shift, width] : segs) {
        d.shift[d.nseg] = (uint8_t)shift;
        d.width[d.nseg] = (uint8_t)width;
        d.bits = (uint8_t)(d.bits + width);
        ++d.nseg;
    }
    return d;
}

}  // namespace

int main() {
    std::mt19937_64 rng(0x5EED2026);
    const uint64_t sizes[] = {1, 2, 63, 64, 65, 1000, 3000, 4095, 4096, 4097, 8192 + 5, 20011};
    struct Config {
        int RB, BS;
        Digit dg;
    };
    const Config configs[] = {
        {8, 512, make_digit({{0, 8}})},
        {8, 512, make_digit({{3, 2}, {17, 4}, {40, 2}})},   // three ranges of one word
        {11, 256, make_digit({{5, 11}})},
        {11, 256, make_digit({{0, 3}, {30, 8}})},
    };
    const char* dists[] = {"uniform", "few distinct", "walk-like"};
    uint64_t cases = 0;
    for (const Config& cf : configs)
        for (uint64_t n : sizes)
            for (int dist = 0; dist < 3; ++dist) {
                std::vector<uint64_t> keys(n);
                for (auto& k : keys) {
                    const uint64_t r = rng();
                    if (dist == 0) k = r;
                    else if (dist == 1) k = (r % 5) * 0x0101010101010101ull;
                    else k = (r & 0x0000000F0000C00Full) | 0x8000000000000000ull;   // a handful of varying bits
                }
                std::vector<uint32_t> perm(n);
                std::iota(perm.begin(), perm.end(), 0u);
                std::shuffle(perm.begin(), perm.end(), rng);
                const int NW = cf.BS / kWave;
                std::vector<std::vector<int>> orders;
                std::vector<int> fwd(NW);
                std::iota(fwd.begin(), fwd.end(), 0);
                orders.push_back(fwd);
                orders.push_back(std::vector<int>(fwd.rbegin(), fwd.rend()));
                std::vector<int> sh = fwd;
                std::shuffle(sh.begin(), sh.end(), rng);
                orders.push_back(sh);
                // expected: stable sort of (key, idx) by digit
                std::vector<uint32_t> order(n);
                std::iota(order.begin(), order.end(), 0u);
                std::stable_sort(order.begin(), order.end(),
                                 [&](uint32_t x, uint32_t y) { return digit_of(keys[x], cf.dg) < digit_of(keys[y], cf.dg); });
                for (int with_idx = 0; with_idx < 2; ++with_idx)
                    for (const auto& ord : orders)
                        for (Variant var : {kCommitted, kEarlyGoff, kReadsFirst}) {
                            PassArgs a{cf.RB, cf.BS, cf.dg, ord};
                            const Result r = model_pass(keys, with_idx ? &perm : nullptr, a, var);
                            for (uint64_t i = 0; i < n; ++i) {
                                const uint32_t src = order[i];
                                if (r.keys[i] != keys[src] || r.idx[i] != (with_idx ? perm[src] : src))
                                    fail(std::string("wrong order: ") + dists[dist] + " n=" + std::to_string(n) +
                                         " RB=" + std::to_string(cf.RB) + " variant=" + std::to_string((int)var) +
                                         " at " + std::to_string(i));
                            }
                            ++cases;
                        }
            }
    std::printf("radix pass model: %" PRIu64 " cases, every access in bounds, no LDS race, stable order in every wave order and variant\n", cases);
    return 0;
}
