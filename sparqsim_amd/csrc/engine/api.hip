// extern "C" entry points of the engine (declared in include/sparq_hip.h).
// Every call converts exceptions into status codes and records a message.
#include <algorithm>
#include <cmath>
#include <limits>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

using namespace sparq;

struct sparq_state_s {
    State st;
};
struct sparq_table_s {
    void* dev = nullptr;
    uint64_t nwords = 0;
    uint32_t word_bytes = 0;
};

// A captured sequence: the graph, its capture (device memory, plans) and the
// state's arrays it was recorded on.
struct sparq_graph_s {
    sparq_state_t state = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Capture* cap = nullptr;
    uint64_t* keys = nullptr;
    double2* amps = nullptr;
    uint64_t arr_cap = 0;
    uint32_t W = 0;
    size_t nodes = 0;
};

static thread_local std::string g_last_error;
static int g_device = -1;

void sparq::set_last_error(const std::string& msg) { g_last_error = msg; }

template <class F>
static int guarded(F&& f) {
    return guarded_call(std::forward<F>(f));
}

void sparq::ensure_device() {
    if (g_device >= 0) return;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        throw Error(SPARQ_NO_DEVICE, "no HIP device available");
    }
    SPQ_HIP(hipSetDevice(0));
    // Every branching operator waits for a few words of its result; spinning
    // wakes the host sooner than a blocking wait.  A process that created the
    // device's context already keeps its own setting.
    if (hipSetDeviceFlags(hipDeviceScheduleSpin) != hipSuccess) (void)hipGetLastError();
    g_device = 0;
}

// Per-state resources (stream, pinned readback words, timing events, zeroed device
// counter) are recycled through a free list: creating and destroying streams and
// pinned memory costs far more than a small state's operators.  A recycled stream is
// drained first; its counter is zero, as every kernel that uses it leaves it.
struct StateResources {
    hipStream_t stream;
    uint64_t* pinned;
    hipEvent_t ev[8];
    unsigned int* counter;
};
static std::mutex g_res_mu;
static std::vector<StateResources> g_res_pool;

static void acquire_resources(State& st) {
    {
        std::lock_guard<std::mutex> lock(g_res_mu);
        if (!g_res_pool.empty()) {
            StateResources r = g_res_pool.back();
            g_res_pool.pop_back();
            st.stream = r.stream;
            st.pinned = r.pinned;
            for (int i = 0; i < 8; ++i) st.ev[i] = r.ev[i];
            st.counter = r.counter;
            return;
        }
    }
    SPQ_HIP(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
    SPQ_HIP(hipHostMalloc((void**)&st.pinned, 64 * sizeof(uint64_t), hipHostMallocDefault));
    for (auto& e : st.ev) SPQ_HIP(hipEventCreate(&e));
    SPQ_HIP(hipMalloc(&st.counter, 256));
    SPQ_HIP(hipMemset(st.counter, 0, 256));
}

static void recycle_resources(State& st) {
    if (!st.stream || !st.pinned || !st.counter) {   // partially built: return what exists
        if (st.stream) (void)hipStreamSynchronize(st.stream);
        for (auto& e : st.ev)
            if (e) (void)hipEventDestroy(e);
        if (st.pinned) (void)hipHostFree(st.pinned);
        if (st.counter) (void)hipFree(st.counter);
        if (st.stream) {
            forget_stream(st.stream);
            (void)hipStreamDestroy(st.stream);
        }
        return;
    }
    (void)hipStreamSynchronize(st.stream);
    StateResources r{st.stream, st.pinned, {}, st.counter};
    for (int i = 0; i < 8; ++i) r.ev[i] = st.ev[i];
    std::lock_guard<std::mutex> lock(g_res_mu);
    g_res_pool.push_back(r);
}

// Replays of a captured sequence leave the entry count on the device: the
// first access after them reads it with the violation flags.  A violated
// plan assumption leaves the content undefined (the count stays within the
// arrays) and is reported once.
static void settle(State& st) {
    if (!st.pending_words) return;
    uint64_t w[2];
    read_back(st, st.pending_words, sizeof(w), w);
    uint64_t* words = const_cast<uint64_t*>(st.pending_words);
    st.pending_words = nullptr;
    st.n = w[kCapN];
    st.pair_bit = -1;
    if (w[kCapFlags]) {
        SPQ_HIP(hipMemsetAsync(words + kCapFlags, 0, sizeof(uint64_t), st.stream));
        std::string why;
        if (w[kCapFlags] & kCapOutsidePlan) why += " an entry outside the planned segments or constant bits;";
        if (w[kCapFlags] & kCapDuplicate) why += " duplicate keys;";
        if (w[kCapFlags] & kCapOverflow) why += " more entries than the captured capacity;";
        if (w[kCapFlags] & kCapNonzeroRemoved) why += " a removed register was not zero;";
        throw Error(SPARQ_RUNTIME, "a replay of the captured sequence broke its plan (state content undefined):" + why);
    }
}

// Every access through S may change the state: it clears basis_amp_one and
// pair_bit.  No access but the capture-aware ones (kAllowCapture) is allowed
// while a sequence is being captured on the state.  kKeepPairs is for the
// accesses that keep the pair grouping or decide about it themselves (size,
// pointwise programs, ClearZero, Normalize, the pair operators).
enum : unsigned { kAllowCapture = 1, kKeepBasis = 2, kKeepPairs = 4 };
static State& access(sparq_state_t h, unsigned how) {
    if (!h) throw Error(SPARQ_INVALID, "null state handle");
    if (h->st.capture && !(how & kAllowCapture))
        throw Error(SPARQ_INVALID, "not available while a sequence is being captured on this state");
    settle(h->st);
    if (!(how & kKeepBasis)) h->st.basis_amp_one = false;
    if (!(how & kKeepPairs) || h->st.capture || h->st.arrays_escaped) h->st.pair_bit = -1;
    return h->st;
}
static State& S(sparq_state_t h) { return access(h, 0); }
static State& S_keep_basis(sparq_state_t h) { return access(h, kKeepBasis); }
static State& S_cap(sparq_state_t h) { return access(h, kAllowCapture); }
static State& S_pairs(sparq_state_t h) { return access(h, kKeepPairs); }

// One lane per entry: all W key words and the amplitude of entry b+i of src
// go to entry at+i of dst.  One launch instead of W+1 copies.
template <int W>
__global__ void __launch_bounds__(kBlock) copy_entries_kernel(uint64_t* __restrict__ dk, double2* __restrict__ da, uint64_t dcap,
                                                              const uint64_t* __restrict__ sk, const double2* __restrict__ sa,
                                                              uint64_t scap, uint64_t cnt) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < cnt; i += stride) {
#pragma unroll
        for (int w = 0; w < W; ++w) dk[(uint64_t)w * dcap + i] = sk[(uint64_t)w * scap + i];
        da[i] = sa[i];
    }
}

template <int W>
static void launch_copy_entries(State& dst, uint64_t at, const State& src, uint64_t b, uint64_t cnt) {
    SPQ_PROF(dst.stream, "copy_entries", (double)cnt * 2.0 * (8.0 * W + 16.0));
    copy_entries_kernel<W><<<std::min<unsigned>(grid_for(cnt, kBlock), 8192), kBlock, 0, dst.stream>>>(
        dst.keys + at, dst.amps + at, dst.cap, src.keys + b, src.amps + b, src.cap, cnt);
    SPQ_HIP(hipGetLastError());
}

// Copy entries [b, b+cnt) of src into dst at position at (same W).
static void copy_entries(State& dst, uint64_t at, const State& src, uint64_t b, uint64_t cnt) {
    if (!cnt) return;
    if (src.W != dst.W) throw Error(SPARQ_RUNTIME, "copy_entries between states of different key widths");
    if (at + cnt > dst.cap || b + cnt > src.cap) throw Error(SPARQ_RUNTIME, "copy_entries out of range");
    SPQ_DISPATCH_W(src.W, launch_copy_entries, dst, at, src, b, cnt);
}

static State sub_state(State& parent, uint64_t b, uint64_t cnt) {
    State s;
    s.W = parent.W;
    s.stream = parent.stream;
    s.pinned = parent.pinned;
    s.counter = parent.counter;
    s.device = parent.device;
    alloc_arrays(s, s.W, std::max<uint64_t>(cnt, 1), &s.keys, &s.amps);
    s.cap = std::max<uint64_t>(cnt, 1);
    s.n = cnt;
    copy_entries(s, 0, parent, b, cnt);
    return s;
}

static void release(State& s) {
    dfree(s.keys, s.stream);
    dfree(s.amps, s.stream);
    s.keys = nullptr;
    s.amps = nullptr;
}

// Append the entries of `tail` behind those of st: in place when st's arrays
// have room, else into fresh arrays.
static void append_entries(State& st, const State& tail) {
    if (!tail.n) return;
    const uint64_t total = st.n + tail.n;
    if (total <= st.cap) {
        copy_entries(st, st.n, tail, 0, tail.n);
        st.n = total;
        return;
    }
    uint64_t* k;
    double2* a;
    alloc_arrays(st, st.W, total, &k, &a);
    State out = st;
    out.keys = k;
    out.amps = a;
    out.cap = total;
    copy_entries(out, 0, st, 0, st.n);
    copy_entries(out, st.n, tail, 0, tail.n);
    install_arrays(st, st.W, k, a, total, total);
}

// Run `op` on the entries satisfying `conds`; the rest are kept (and ClearZero'd,
// as the reference's ClearZero covers the whole state).  After the partition
// the rest moves to arrays of its own, the op runs on the state itself holding
// only the satisfying entries, and the rest is appended behind its result.
template <class F>
static void with_conditions(State& st, const sparq_cond_t* conds, uint32_t ncond, double eps, F op) {
    if (ncond == 0) {
        op(st);
        return;
    }
    const uint64_t S_ = partition_by_conditions(st, conds, ncond);
    if (S_ == st.n) {
        op(st);
        return;
    }
    State rest = sub_state(st, S_, st.n - S_);
    st.n = S_;
    bool appended = false;
    try {
        if (S_) op(st);
        clear_zero(rest, eps);
        append_entries(st, rest);
        appended = true;
    } catch (...) {
        // keep the rest's entries behind whatever the failed op left
        if (!appended) {
            try {
                append_entries(st, rest);
            } catch (...) {
            }
        }
        release(rest);
        throw;
    }
    release(rest);
}

extern "C" {

const char* sparq_last_error(void) { return g_last_error.c_str(); }

int sparq_device_init(int device) {
    return guarded([&] {
        SPQ_REQUIRE(device == 0, "only device 0 is supported");
        ensure_device();
    });
}

int sparq_device_name(char* buf, size_t len) {
    return guarded([&] {
        ensure_device();
        hipDeviceProp_t p;
        SPQ_HIP(hipGetDeviceProperties(&p, 0));
        std::snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    });
}

int sparq_state_create(uint32_t words, uint64_t capacity, sparq_state_t* out) {
    return guarded([&] {
        SPQ_REQUIRE(out != nullptr, "null output");
        SPQ_REQUIRE(words >= 1 && words <= SPARQ_MAX_WORDS, "key width out of range");
        ensure_device();
        auto* h = new sparq_state_s();
        State& st = h->st;
        st.W = words;
        try {
            acquire_resources(st);
            st.cap = std::max<uint64_t>(capacity, 1);
            alloc_arrays(st, words, st.cap, &st.keys, &st.amps);
            st.n = 0;
        } catch (...) {
            release(st);
            recycle_resources(st);
            delete h;
            throw;
        }
        *out = h;
    });
}

int sparq_state_destroy(sparq_state_t h) {
    return guarded([&] {
        if (!h) return;
        State& st = h->st;
        capture_abort(st);
        release(st);
        recycle_resources(st);
        delete h;
    });
}

int sparq_state_clone(sparq_state_t src, sparq_state_t* out) {
    return guarded([&] {
        State& s = S(src);
        sparq_state_t h = nullptr;
        int rc = sparq_state_create(s.W, std::max<uint64_t>(s.n, 1), &h);
        if (rc != SPARQ_OK) throw Error(rc, g_last_error);
        State& d = h->st;
        // order the copy after pending work on the source stream
        SPQ_HIP(hipEventRecord(s.ev[7], s.stream));
        SPQ_HIP(hipStreamWaitEvent(d.stream, s.ev[7], 0));
        copy_entries(d, 0, s, 0, s.n);
        d.n = s.n;
        *out = h;
    });
}

int sparq_state_words(sparq_state_t h, uint32_t* words) {
    return guarded([&] { *words = access(h, kAllowCapture | kKeepPairs).W; });
}

int sparq_state_size(sparq_state_t h, uint64_t* n) {
    return guarded([&] {
        State& st = access(h, kAllowCapture | kKeepPairs);
        // while a sequence is captured the count of the plan's starting state
        if (!st.capture) SPQ_HIP(hipStreamSynchronize(st.stream));
        *n = st.n;
    });
}

int sparq_state_set_basis(sparq_state_t h, const uint64_t* key) {
    return guarded([&] {
        State& st = S(h);
        BasisKey k{};
        if (key) std::memcpy(k.w, key, st.W * 8);
        if (st.cap < 1 || !st.keys) {
            uint64_t* kd;
            double2* ad;
            alloc_arrays(st, st.W, 1, &kd, &ad);
            install_arrays(st, st.W, kd, ad, 0, 1);
        }
        write_basis(st, k);   // stream-ordered; no host synchronization
        st.n = 1;
        st.basis_amp_one = !st.arrays_escaped;
    });
}

int sparq_state_upload(sparq_state_t h, const uint64_t* keys, const double* amps, uint64_t n) {
    return guarded([&] {
        State& st = S(h);
        SPQ_REQUIRE(n == 0 || (keys && amps), "null host buffer");
        uint64_t cap = std::max<uint64_t>(n, 1);
        std::vector<uint64_t> planes((size_t)st.W * n);
        for (uint64_t i = 0; i < n; ++i)
            for (uint32_t w = 0; w < st.W; ++w) planes[(size_t)w * n + i] = keys[i * st.W + w];
        uint64_t* kd;
        double2* ad;
        alloc_arrays(st, st.W, cap, &kd, &ad);
        for (uint32_t w = 0; w < st.W && n; ++w)
            SPQ_HIP(hipMemcpyAsync(kd + (uint64_t)w * cap, planes.data() + (size_t)w * n, n * 8, hipMemcpyHostToDevice, st.stream));
        if (n) SPQ_HIP(hipMemcpyAsync(ad, amps, n * 16, hipMemcpyHostToDevice, st.stream));
        SPQ_HIP(hipStreamSynchronize(st.stream));
        install_arrays(st, st.W, kd, ad, n, cap);
    });
}

int sparq_state_download(sparq_state_t h, uint64_t* keys, double* amps, uint64_t capacity, uint64_t* n) {
    return guarded([&] {
        State& st = S(h);
        SPQ_HIP(hipStreamSynchronize(st.stream));
        if (n) *n = st.n;
        if (!keys && !amps) return;
        SPQ_REQUIRE(capacity >= st.n, "download buffer too small");
        if (st.n == 0) return;
        std::vector<uint64_t> planes((size_t)st.W * st.n);
        for (uint32_t w = 0; w < st.W; ++w)
            SPQ_HIP(hipMemcpyAsync(planes.data() + (size_t)w * st.n, st.keys + (uint64_t)w * st.cap, st.n * 8, hipMemcpyDeviceToHost,
                                   st.stream));
        if (amps) SPQ_HIP(hipMemcpyAsync(amps, st.amps, st.n * 16, hipMemcpyDeviceToHost, st.stream));
        SPQ_HIP(hipStreamSynchronize(st.stream));
        if (keys)
            for (uint64_t i = 0; i < st.n; ++i)
                for (uint32_t w = 0; w < st.W; ++w) keys[i * st.W + w] = planes[(size_t)w * st.n + i];
    });
}

int sparq_state_relayout(sparq_state_t h, uint32_t new_words, const sparq_field_t* from, const sparq_field_t* to, uint32_t nfields) {
    return guarded([&] {
        SPQ_REQUIRE(nfields == 0 || (from && to), "null field list");
        relayout(S(h), new_words, from, to, nfields);
    });
}

int sparq_state_resize_words(sparq_state_t h, uint32_t new_words) {
    return guarded([&] { resize_words(S(h), new_words); });
}

int sparq_state_sync(sparq_state_t h) {
    return guarded([&] { SPQ_HIP(hipStreamSynchronize(S(h).stream)); });
}

int sparq_memory_stats(uint64_t out[6]) {
    return guarded([&] {
        SPQ_REQUIRE(out != nullptr, "null output");
        memory_stats(out);
    });
}

int sparq_post_select_dense(sparq_state_t h, const sparq_field_t* zero, uint32_t nzero, const sparq_field_t* index, uint32_t nindex,
                            double* out, uint64_t len, double* weight) {
    return guarded([&] {
        SPQ_REQUIRE(out != nullptr && (zero != nullptr || nzero == 0) && index != nullptr, "null argument");
        const double w = post_select_dense(S(h), zero, nzero, index, nindex, out, len);
        if (weight) *weight = w;
    });
}

int sparq_select(sparq_state_t h, const sparq_cond_t* conds, uint32_t nconds, double* factor) {
    return guarded([&] {
        SPQ_REQUIRE(conds != nullptr || nconds == 0, "null argument");
        State& st = S(h);
        double f = 0.0;
        if (st.n != 0) {
            select_by_conditions(st, conds, nconds);
            const double s = sum_abs2(st);
            if (s != 0) {
                f = 1.0 / std::sqrt(s);
                scale_amps(st, f);
            } else {
                f = std::numeric_limits<double>::infinity();
            }
        }
        if (factor) *factor = f;
    });
}

int sparq_set_radix_bits(int bits) {
    return guarded([&] {
        SPQ_REQUIRE(bits == 0 || bits == 8 || bits == 11, "digit width must be 0 (auto), 8 or 11");
        set_radix_bits(bits);
    });
}

int sparq_set_sort_mode(int mode) {
    return guarded([&] {
        SPQ_REQUIRE(mode == 0 || mode == 1, "sort mode must be 0 (prefix + local) or 1 (LSD passes only)");
        set_sort_mode(mode);
    });
}

int sparq_set_pointwise_items(int items) {
    return guarded([&] {
        SPQ_REQUIRE(items == 0 || items == 1 || items == 2 || items == 4, "entries per thread must be 0 (autotuned), 1, 2 or 4");
        set_pointwise_items(items);
    });
}

int sparq_autotune_report(char* buf, size_t len, size_t* needed) {
    return guarded([&] {
        const std::string r = "{\"pointwise\":" + pointwise_tuning_report() + ",\"radix\":" + radix_tuning_report() + "}";
        if (needed) *needed = r.size() + 1;
        if (buf && len) {
            const size_t k = std::min(len - 1, r.size());
            std::memcpy(buf, r.data(), k);
            buf[k] = 0;
        }
    });
}

int sparq_set_dense_transform(int on) {
    return guarded([&] { set_dense_transform(on != 0); });
}

int sparq_set_canonical_output(int on) {
    return guarded([&] { set_canonical_output(on != 0); });
}

static bool g_group_pass = false;
int sparq_set_group_pass_through(int on) {
    return guarded([&] { g_group_pass = on != 0; });
}

int sparq_sort_stats(uint64_t* out, uint32_t count, int reset) {
    return guarded([&] {
        SPQ_REQUIRE(out && count >= 6, "sparq_sort_stats needs room for 6 counters");
        sort_path_stats(out, reset);
    });
}

int sparq_release_cached_memory(void) {
    return guarded([&] {
        if (g_device < 0) return;
        release_cached_memory();
    });
}

int sparq_state_device_arrays(sparq_state_t h, uint64_t** keys, double** amps, uint64_t* capacity) {
    return guarded([&] {
        State& st = S(h);
        st.arrays_escaped = true;
        if (keys) *keys = st.keys;
        if (amps) *amps = reinterpret_cast<double*>(st.amps);
        if (capacity) *capacity = st.cap;
    });
}

int sparq_state_stream(sparq_state_t h, void** stream) {
    return guarded([&] { *stream = (void*)S(h).stream; });
}

int sparq_table_create(const void* host_words, uint64_t nwords, uint32_t word_bytes, sparq_table_t* out) {
    return guarded([&] {
        SPQ_REQUIRE(out && (host_words || nwords == 0), "null argument");
        SPQ_REQUIRE(word_bytes == 1 || word_bytes == 2 || word_bytes == 4 || word_bytes == 8, "word_bytes must be 1, 2, 4 or 8");
        ensure_device();
        auto* t = new sparq_table_s();
        t->nwords = nwords;
        t->word_bytes = word_bytes;
        hipError_t e = hipMalloc(&t->dev, std::max<uint64_t>(nwords * word_bytes, 8));
        if (e != hipSuccess) {
            delete t;
            throw Error(SPARQ_OOM, "table allocation failed");
        }
        if (nwords) {
            e = hipMemcpy(t->dev, host_words, nwords * word_bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                (void)hipFree(t->dev);
                delete t;
                throw Error(SPARQ_RUNTIME, "table upload failed");
            }
        }
        *out = t;
    });
}

int sparq_table_destroy(sparq_table_t t) {
    return guarded([&] {
        if (!t) return;
        (void)hipDeviceSynchronize();
        (void)hipFree(t->dev);
        delete t;
    });
}

int sparq_table_device_ptr(sparq_table_t t, const void** ptr) {
    return guarded([&] {
        SPQ_REQUIRE(t && ptr, "null argument");
        *ptr = t->dev;
    });
}

int sparq_pointwise(sparq_state_t h, const sparq_pw_op_t* ops, uint32_t nops, const sparq_cond_t* conds, uint32_t nconds) {
    return guarded([&] {
        SPQ_REQUIRE(nops == 0 || ops, "null op list");
        SPQ_REQUIRE(nconds == 0 || conds, "null condition list");
        pointwise(access(h, kAllowCapture | kKeepPairs), ops, nops, conds, nconds);
    });
}

int sparq_clear_zero(sparq_state_t h, double eps) {
    return guarded([&] { clear_zero(S_pairs(h), eps); });
}

int sparq_sort(sparq_state_t h, const sparq_field_t* last) {
    return guarded([&] {
        State& st = S(h);
        if (last) SPQ_REQUIRE(last->word < st.W && last->width >= 1 && (uint32_t)last->shift + last->width <= 64, "bad field");
        sort_state(st, last);
    });
}

int sparq_sort_by(sparq_state_t h, const sparq_field_t* major, uint32_t nmajor, const sparq_field_t* minor, uint32_t nminor) {
    return guarded([&] {
        SPQ_REQUIRE((major != nullptr || nmajor == 0) && (minor != nullptr || nminor == 0), "null field list");
        State& st = S(h);
        sort_state_by(st, std::vector<sparq_field_t>(major, major + nmajor), std::vector<sparq_field_t>(minor, minor + nminor));
    });
}

int sparq_sort_by_amplitude(sparq_state_t h) {
    return guarded([&] { sort_by_amplitude(S(h)); });
}

// Make `dst`'s stream wait for the work queued so far on `src`'s stream.
static void stream_after(State& dst, State& src) {
    if (dst.stream == src.stream) return;
    SPQ_HIP(hipEventRecord(src.ev[7], src.stream));
    SPQ_HIP(hipStreamWaitEvent(dst.stream, src.ev[7], 0));
}

int sparq_partition(sparq_state_t h, const sparq_cond_t* conds, uint32_t ncond, sparq_state_t hrest) {
    return guarded([&] {
        SPQ_REQUIRE(conds != nullptr || ncond == 0, "null conditions");
        SPQ_REQUIRE(h != hrest, "the remainder must be another state");
        State& st = S(h);
        State& rest = S(hrest);
        SPQ_REQUIRE(rest.W == st.W, "states have different key widths");
        for (uint32_t c = 0; c < ncond; ++c) SPQ_REQUIRE(conds[c].field.word < st.W, "condition field out of range");
        const uint64_t kept = ncond ? partition_by_conditions(st, conds, ncond) : st.n;
        const uint64_t moved = st.n - kept;
        // the remainder's new arrays are written on st's stream after rest's pending work
        stream_after(st, rest);
        uint64_t* k;
        double2* a;
        alloc_arrays(rest, rest.W, std::max<uint64_t>(moved, 1), &k, &a);
        State out = rest;
        out.keys = k;
        out.amps = a;
        out.cap = std::max<uint64_t>(moved, 1);
        out.stream = st.stream;
        copy_entries(out, 0, st, kept, moved);
        stream_after(rest, st);
        install_arrays(rest, rest.W, k, a, moved, out.cap);
        st.n = kept;
    });
}

int sparq_append(sparq_state_t hto, sparq_state_t hfrom) {
    return guarded([&] {
        State& to = S(hto);
        State& from = S(hfrom);
        SPQ_REQUIRE(to.W == from.W, "states have different key widths");
        if (from.n == 0) return;
        stream_after(to, from);
        const uint64_t orig = to.n, add_n = from.n, total = orig + add_n;
        uint64_t* k;
        double2* a;
        alloc_arrays(to, to.W, total, &k, &a);
        State out = to;
        out.keys = k;
        out.amps = a;
        out.cap = total;
        copy_entries(out, 0, to, 0, orig);
        copy_entries(out, orig, from, 0, add_n);
        install_arrays(to, to.W, k, a, total, total);
        stream_after(from, to);   // `from` stays unchanged until the copy has run
    });
}

int sparq_merge_duplicates(sparq_state_t h, double eps) {
    return guarded([&] {
        State& st = S(h);
        sort_state(st, nullptr);
        merge_sorted_duplicates(st, eps);
    });
}

int sparq_add_systems(sparq_state_t hc, sparq_state_t ha, double coef, double eps) {
    return guarded([&] {
        State& cur = S(hc);
        State& add = S(ha);
        SPQ_REQUIRE(cur.W == add.W, "states have different key widths");
        if (add.n == 0) return;
        SPQ_HIP(hipEventRecord(add.ev[7], add.stream));
        SPQ_HIP(hipStreamWaitEvent(cur.stream, add.ev[7], 0));
        // `add` may be `cur` itself: its size is read before the arrays are replaced
        const uint64_t orig = cur.n, add_n = add.n, total = orig + add_n;
        uint64_t* k;
        double2* a;
        alloc_arrays(cur, cur.W, total, &k, &a);
        State out = cur;
        out.keys = k;
        out.amps = a;
        out.cap = total;
        copy_entries(out, 0, cur, 0, orig);
        copy_entries(out, orig, add, 0, add_n);
        install_arrays(cur, cur.W, k, a, total, total);
        if (std::abs(coef - 1.0) > 1e-14) {
            State tail = cur;   // view of the appended entries
            tail.amps = cur.amps + orig;
            tail.n = add_n;
            scale_amps(tail, coef);
        }
        if (orig != 0) {
            sort_state(cur, nullptr);
            merge_sorted_duplicates(cur, eps);
        }
        // keep `add` alive until the copy has run
        SPQ_HIP(hipEventRecord(cur.ev[6], cur.stream));
        SPQ_HIP(hipStreamWaitEvent(add.stream, cur.ev[6], 0));
    });
}

int sparq_check_duplicates(sparq_state_t h, int* has_dup) {
    return guarded([&] {
        State& st = S(h);
        sort_state(st, nullptr);
        *has_dup = has_sorted_duplicates(st) ? 1 : 0;
    });
}

// The group transforms (Hadamard_Int, Hadamard_Int_Full, Rot_GeneralUnitary)
// evaluate their conditions per entry; the reference evaluates them on a group's
// first entry (hadamard.cpp:28,43,80).  The two agree unless a condition reads
// the bits the transform changes, which is rejected.
static void require_conditions_off_target(sparq_field_t target, uint32_t n_digits, const sparq_cond_t* conds, uint32_t ncond) {
    SPQ_REQUIRE(conds != nullptr || ncond == 0, "null conditions");
    const uint32_t lo = target.shift, hi = (uint32_t)target.shift + n_digits;
    for (uint32_t c = 0; c < ncond; ++c) {
        const sparq_field_t f = conds[c].field;
        // a bit condition reads one bit of its register, the others the whole field
        const bool one_bit = conds[c].kind == SPARQ_COND_BIT || conds[c].kind == SPARQ_COND_BIT_CLEAR;
        const uint32_t flo = one_bit ? f.shift + (uint32_t)conds[c].value : f.shift;
        const uint32_t fhi = one_bit ? flo + 1 : (uint32_t)f.shift + f.width;
        const bool overlap = f.word == target.word && flo < hi && fhi > lo;
        SPQ_REQUIRE(!overlap, "a condition of a group transform reads the transformed bits of its target register");
    }
}

int sparq_hadamard_int(sparq_state_t h, sparq_field_t target, uint32_t n_digits, const sparq_cond_t* conds, uint32_t ncond,
                       double eps) {
    return guarded([&] {
        require_conditions_off_target(target, n_digits, conds, ncond);
        if (h && h->st.capture) {
            hadamard_captured(S_cap(h), target, n_digits, eps, conds, ncond);
            return;
        }
        State& st = S_keep_basis(h);
        const bool unit = st.basis_amp_one && ncond == 0;
        st.basis_amp_one = false;
        with_conditions(st, conds, ncond, eps, [&](State& s) { hadamard_int(s, target, n_digits, false, eps, unit); });
    });
}

int sparq_hadamard_int_full(sparq_state_t h, sparq_field_t target, const sparq_cond_t* conds, uint32_t ncond, double eps) {
    return guarded([&] {
        require_conditions_off_target(target, target.width, conds, ncond);
        if (h && h->st.capture) {   // duplicate keys are flagged by the captured transform itself
            hadamard_captured(S_cap(h), target, target.width, eps, conds, ncond);
            return;
        }
        State& st = S(h);
        // CheckDuplicateKey precedes the transform (hadamard.cpp:183).  Without
        // conditions the transform checks on its way (before it changes
        // anything); with them the whole state is checked first.
        if (ncond == 0) {
            hadamard_int(st, target, target.width, true, eps, false, true);
            return;
        }
        if (st.n > 1) {
            // any total order of the keys makes duplicates adjacent; sorting with the
            // target least significant also prepares the grouping sort below
            sort_state(st, &target);
            if (has_sorted_duplicates(st)) throw Error(SPARQ_RUNTIME, "Hadamard_Int_Full: duplicate key in state");
        }
        with_conditions(st, conds, ncond, eps, [&](State& s) { hadamard_int(s, target, target.width, true, eps); });
    });
}

int sparq_group_unitary(sparq_state_t h, sparq_field_t target, const double* mat, const sparq_cond_t* conds, uint32_t ncond,
                        double eps) {
    return guarded([&] {
        SPQ_REQUIRE(mat != nullptr, "null matrix");
        require_conditions_off_target(target, target.width, conds, ncond);
        SPQ_REQUIRE(ncond <= SPARQ_MAX_PROGRAM_CONDS, "too many conditions");
        State& st = S(h);
        // Partition first (the default): only the entries that satisfy the
        // conditions are grouped and transformed.
        if (ncond && !g_group_pass) {
            with_conditions(st, conds, ncond, eps, [&](State& s) { group_unitary(s, target, mat, eps); });
            return;
        }
        // The conditions read no target bit, so a group's members agree on
        // them: groups that fail pass through inside the transform.
        CondList cl{};
        cl.n = ncond;
        for (uint32_t c = 0; c < ncond; ++c) {
            SPQ_REQUIRE(conds[c].field.word < st.W, "condition field out of range");
            cl.c[c] = conds[c];
        }
        group_unitary(st, target, mat, eps, &cl);
    });
}

int sparq_pair_op(sparq_state_t h, sparq_field_t target, uint32_t bit, const sparq_pair_params_t* params, const sparq_cond_t* conds,
                  uint32_t ncond, double eps) {
    return guarded([&] {
        SPQ_REQUIRE(params != nullptr, "null params");
        SPQ_REQUIRE(ncond <= SPARQ_MAX_PROGRAM_CONDS, "too many conditions");
        State& st = S_pairs(h);
        // Conditions that do not read the target bit are evaluated inside the
        // pair kernels (partners agree on them; failing entries pass through);
        // otherwise the state is partitioned first.
        bool inline_ok = true;
        CondList cl{};
        cl.n = ncond;
        for (uint32_t c = 0; c < ncond; ++c) {
            SPQ_REQUIRE(conds[c].field.word < st.W, "condition field out of range");
            cl.c[c] = conds[c];
            const sparq_field_t f = conds[c].field;
            const uint32_t tb = (uint32_t)target.shift + bit;
            if (f.word == target.word && tb >= f.shift && tb < (uint32_t)f.shift + f.width) inline_ok = false;
        }
        if (inline_ok) {
            pair_op(st, target, bit, *params, cl, eps);
            return;
        }
        st.pair_bit = -1;
        with_conditions(st, conds, ncond, eps, [&](State& s) { pair_op(s, target, bit, *params, CondList{}, eps); });
        st.pair_bit = -1;   // the rest is appended behind the pairs
    });
}

int sparq_replicate(sparq_state_t h, sparq_field_t target) {
    return guarded([&] { replicate(S(h), target); });
}

int sparq_normalize(sparq_state_t h, double* norm2) {
    return guarded([&] {
        State& st = S_pairs(h);
        double s = sum_abs2(st);
        if (norm2) *norm2 = s;
        double f = 1.0 / std::sqrt(s);
        if (std::isnan(f)) throw Error(SPARQ_RUNTIME, "Normalize: bad result (NaN norm)");
        scale_amps(st, f);
    });
}

int sparq_tile_norm2(sparq_state_t h, uint64_t tile, double* out, uint64_t ntiles) {
    return guarded([&] {
        SPQ_REQUIRE(out != nullptr || ntiles == 0, "null argument");
        tile_abs2(S(h), tile, out, ntiles);
    });
}

int sparq_cond_norm2(sparq_state_t h, const sparq_cond_t* conds, uint32_t nconds, double* out) {
    return guarded([&] {
        SPQ_REQUIRE(out != nullptr && (conds != nullptr || nconds == 0), "null argument");
        *out = cond_abs2(S(h), conds, nconds);
    });
}

int sparq_state_read_range(sparq_state_t h, uint64_t first, uint64_t count, uint64_t* keys, double* amps) {
    return guarded([&] {
        State& st = S(h);
        SPQ_REQUIRE(first <= st.n && count <= st.n - first, "range outside the state");
        if (count == 0) return;
        if (keys) {
            std::vector<uint64_t> word(count);
            for (uint32_t w = 0; w < st.W; ++w) {
                read_back(st, st.keys + (uint64_t)w * st.cap + first, count * sizeof(uint64_t), word.data());
                for (uint64_t i = 0; i < count; ++i) keys[i * st.W + w] = word[i];
            }
        }
        if (amps) read_back(st, st.amps + first, count * sizeof(double2), amps);
    });
}

int sparq_norm2(sparq_state_t h, double* norm2) {
    return guarded([&] {
        SPQ_REQUIRE(norm2 != nullptr, "null argument");
        *norm2 = sum_abs2(S_pairs(h));
    });
}

int sparq_field_uniform(sparq_state_t h, sparq_field_t f, int* uniform, uint64_t* value) {
    return guarded([&] {
        State& st = S_pairs(h);
        SPQ_REQUIRE(f.word < st.W, "bad field");
        uint64_t v = 0;
        bool u = field_uniform(st, f, &v);
        if (uniform) *uniform = u ? 1 : 0;
        if (value) *value = v;
    });
}

int sparq_capture_plan(sparq_state_t h) {
    return guarded([&] { capture_plan_begin(S(h)); });
}

int sparq_capture_record(sparq_state_t h) {
    return guarded([&] {
        State& st = S_cap(h);
        try {
            capture_record_begin(st);
        } catch (...) {
            capture_abort(st);
            throw;
        }
    });
}

int sparq_capture_end(sparq_state_t h, sparq_graph_t* out) {
    return guarded([&] {
        SPQ_REQUIRE(out != nullptr, "null output");
        State& st = S_cap(h);
        hipGraph_t graph = nullptr;
        std::unique_ptr<Capture> cap(capture_end(st, &graph));
        auto g = std::make_unique<sparq_graph_s>();
        g->state = h;
        g->graph = graph;
        g->cap = cap.get();
        const hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            throw Error(SPARQ_RUNTIME, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(e));
        }
        (void)hipGraphGetNodes(graph, nullptr, &g->nodes);
        g->keys = st.keys;
        g->amps = st.amps;
        g->arr_cap = st.cap;
        g->W = st.W;
        cap.release();
        *out = g.release();
    });
}

int sparq_capture_abort(sparq_state_t h) {
    return guarded([&] {
        SPQ_REQUIRE(h != nullptr, "null state handle");
        capture_abort(h->st);
    });
}

int sparq_capture_phase(sparq_state_t h, int* phase) {
    return guarded([&] {
        SPQ_REQUIRE(h && phase, "null argument");
        *phase = h->st.capture ? h->st.capture->phase : 0;
    });
}

int sparq_graph_launch(sparq_graph_t g, uint32_t times) {
    return guarded([&] {
        SPQ_REQUIRE(g && g->state, "null graph");
        State& st = g->state->st;
        SPQ_REQUIRE(!st.capture, "a sequence is being captured on the graph's state");
        SPQ_REQUIRE(st.keys == g->keys && st.amps == g->amps && st.cap == g->arr_cap && st.W == g->W,
                    "the state's arrays changed since the sequence was captured");
        if (st.pending_words != g->cap->words) {
            settle(st);
            SPQ_REQUIRE(st.n <= g->cap->capacity, "the state holds more entries than the captured capacity");
            capture_set_count(st, g->cap->words);
        }
        st.basis_amp_one = false;
        st.pair_bit = -1;
        for (uint32_t t = 0; t < times; ++t) SPQ_HIP(hipGraphLaunch(g->exec, st.stream));
        st.pending_words = g->cap->words;
    });
}

int sparq_graph_info(sparq_graph_t g, uint64_t* capacity, uint64_t* nodes) {
    return guarded([&] {
        SPQ_REQUIRE(g != nullptr, "null graph");
        if (capacity) *capacity = g->cap->capacity;
        if (nodes) *nodes = g->nodes;
    });
}

int sparq_graph_destroy(sparq_graph_t g) {
    return guarded([&] {
        if (!g) return;
        // replays may still run on the state's stream: the device memory they use goes last
        (void)hipDeviceSynchronize();
        if (g->state && g->state->st.pending_words == g->cap->words) {
            State& st = g->state->st;
            uint64_t w[2] = {0, 0};
            if (hipMemcpy(w, st.pending_words, sizeof(w), hipMemcpyDeviceToHost) == hipSuccess) st.n = std::min(w[kCapN], st.cap);
            st.pending_words = nullptr;
        }
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g->cap;
        delete g;
    });
}

int sparq_table_advance(sparq_state_t h, sparq_table_t t) {
    return guarded([&] {
        SPQ_REQUIRE(t && t->word_bytes == 8 && t->nwords >= 1, "the table needs uint64 words");
        table_advance(S_cap(h), (uint64_t*)t->dev);
    });
}

int sparq_table_write(sparq_table_t t, uint64_t first, const void* host_words, uint64_t count) {
    return guarded([&] {
        SPQ_REQUIRE(t && (host_words || count == 0), "null argument");
        SPQ_REQUIRE(first <= t->nwords && count <= t->nwords - first, "table write out of range");
        if (count) {
            SPQ_HIP(hipDeviceSynchronize());   // after queued work that reads the table
            SPQ_HIP(hipMemcpy((char*)t->dev + first * t->word_bytes, host_words, count * t->word_bytes, hipMemcpyHostToDevice));
        }
    });
}

int sparq_timer_record(sparq_state_t h, int slot) {
    return guarded([&] {
        SPQ_REQUIRE(slot >= 0 && slot < 6, "timer slot out of range");
        State& st = S(h);
        SPQ_HIP(hipEventRecord(st.ev[slot], st.stream));
    });
}

int sparq_timer_elapsed(sparq_state_t h, int b, int e, float* ms) {
    return guarded([&] {
        SPQ_REQUIRE(b >= 0 && b < 6 && e >= 0 && e < 6 && ms, "timer slot out of range");
        State& st = S(h);
        SPQ_HIP(hipEventSynchronize(st.ev[e]));
        SPQ_HIP(hipEventElapsedTime(ms, st.ev[b], st.ev[e]));
    });
}

}  // extern "C"
