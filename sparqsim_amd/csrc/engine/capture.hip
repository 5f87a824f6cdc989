// Recording a fixed operator sequence on one state as a HIP graph.
//
// Phase 1 (plan) snapshots the state's keys to the host and runs the
// sequence through the API with nothing launched: every branching operator
// plans its transform from the snapshot (branch.hip: segments, lookup table,
// constant bits).  Phase 2 (record) sizes the device memory the plans need
// (one allocation: the device words, the dense domain, scan partials, the
// rest arrays, the plans' tables), grows the state's arrays to the capacity
// the plans imply, and starts capturing the state's stream; the same sequence
// then records its kernels.  Every recorded kernel reads the entry count from
// the capture's device words and strides over it, so replays need no host
// synchronisation; assumptions of the plan are checked into a flags word.
#include <algorithm>
#include <atomic>
#include <memory>
#include <vector>

#include "engine.hpp"

namespace sparq {

static std::atomic<int> g_active_captures{0};

bool capture_active() { return g_active_captures.load(std::memory_order_relaxed) > 0; }

// Persistent grids: at most 8 workgroups per CU of the 256-CU part, fewer when
// the capacity needs fewer; the kernels stride over the device entry count.
unsigned persistent_grid(uint64_t cap, unsigned per_block) {
    const uint64_t need = (cap + per_block - 1) / per_block;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(need, 2048));
}

Capture::~Capture() {
    capture_plan_free(*this);
    if (arena) (void)hipFree(arena);
}

// Words of partial sums exclusive_scan_u32_scratch needs for n inputs.
static uint64_t scan_scratch_words(uint64_t n) {
    constexpr uint64_t kTile = 2048;   // inputs per workgroup of the scan
    uint64_t w = 0;
    while (n > kTile) {
        n = (n + kTile - 1) / kTile;
        w += 2 * n;
    }
    return w + 8;
}

void capture_plan_begin(State& st) {
    SPQ_REQUIRE(!st.capture, "a sequence is already being captured on this state");
    auto c = std::make_unique<Capture>();
    c->W = st.W;
    c->n0 = st.n;
    c->snap.resize((size_t)st.W * st.n);
    for (uint32_t w = 0; w < st.W && st.n; ++w)
        SPQ_HIP(hipMemcpyAsync(c->snap.data() + (size_t)w * st.n, st.keys + (uint64_t)w * st.cap, st.n * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, st.stream));
    SPQ_HIP(hipStreamSynchronize(st.stream));
    st.capture = c.release();
    g_active_captures.fetch_add(1);
}

void capture_record_begin(State& st) {
    SPQ_REQUIRE(st.capture && st.capture->phase == 1, "no sequence is being planned on this state");
    Capture& c = *st.capture;
    SPQ_REQUIRE(st.W == c.W && st.n == c.n0, "the state changed while the sequence was planned");
    c.snap.clear();
    c.snap.shrink_to_fit();
    c.capacity = std::max<uint64_t>(c.n0, 1);
    c.max_slots = 0;
    c.any_conditions = false;
    for (size_t i = 0; i < c.hs.size(); ++i) {
        c.capacity = std::max(c.capacity, capture_plan_slots(*c.hs[i]) + capture_plan_rest(*c.hs[i]));
        c.max_slots = std::max(c.max_slots, capture_plan_slots_padded(*c.hs[i]));
        c.any_conditions = c.any_conditions || capture_plan_has_conditions(*c.hs[i]);
    }
    // the state's arrays hold the capacity (the recorded kernels write in place)
    if (st.cap < c.capacity) {
        uint64_t* k;
        double2* a;
        alloc_arrays(st, st.W, c.capacity, &k, &a);
        for (uint32_t w = 0; w < st.W && st.n; ++w)
            SPQ_HIP(hipMemcpyAsync(k + (uint64_t)w * c.capacity, st.keys + (uint64_t)w * st.cap, st.n * sizeof(uint64_t),
                                   hipMemcpyDeviceToDevice, st.stream));
        if (st.n) SPQ_HIP(hipMemcpyAsync(a, st.amps, st.n * sizeof(double2), hipMemcpyDeviceToDevice, st.stream));
        install_arrays(st, st.W, k, a, st.n, c.capacity);
    }
    const uint64_t stiles = c.max_slots / 4096, ntiles = (c.capacity + 4095) / 4096;
    c.scan_words = scan_scratch_words(std::max<uint64_t>(std::max(stiles, ntiles), 1));
    auto round = [](uint64_t b) { return (b + 255) / 256 * 256; };
    uint64_t bytes = round(kCapWords * 8) + round(c.max_slots * 16) + round(c.max_slots) + round(stiles * 4) + round((stiles + 1) * 8) +
                     round(ntiles * 4) + round((ntiles + 1) * 8) + round(c.scan_words * 8) + capture_plan_bytes(c);
    if (c.any_conditions) bytes += round((uint64_t)st.W * c.capacity * 8) + round(c.capacity * 16);
    SPQ_HIP(hipStreamSynchronize(st.stream));
    SPQ_HIP(hipMalloc((void**)&c.arena, bytes));
    char* at = c.arena;
    auto take = [&](uint64_t b) {
        char* p = at;
        at += round(b);
        return p;
    };
    c.words = (uint64_t*)take(kCapWords * 8);
    c.A = (double2*)take(c.max_slots * 16);
    c.mark = (uint8_t*)take(c.max_slots);
    c.kept = (uint32_t*)take(stiles * 4);
    c.offs = (uint64_t*)take((stiles + 1) * 8);
    c.rest_cnt = (uint32_t*)take(ntiles * 4);
    c.rest_off = (uint64_t*)take((ntiles + 1) * 8);
    c.scan_scratch = (uint64_t*)take(c.scan_words * 8);
    if (c.any_conditions) {
        c.rest_keys = (uint64_t*)take((uint64_t)st.W * c.capacity * 8);
        c.rest_amps = (double2*)take(c.capacity * 16);
    }
    capture_plan_upload(c, at);
    uint64_t init[kCapWords] = {};
    init[kCapN] = st.n;
    SPQ_HIP(hipMemcpy(c.words, init, sizeof(init), hipMemcpyHostToDevice));
    c.phase = 2;
    c.next_h = 0;
    SPQ_HIP(hipStreamBeginCapture(st.stream, hipStreamCaptureModeRelaxed));
}

Capture* capture_end(State& st, hipGraph_t* graph) {
    SPQ_REQUIRE(st.capture && st.capture->phase == 2, "no sequence is being recorded on this state");
    std::unique_ptr<Capture> c(st.capture);
    st.capture = nullptr;
    g_active_captures.fetch_sub(1);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st.stream, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        throw Error(SPARQ_RUNTIME, std::string("hipStreamEndCapture failed: ") + hipGetErrorString(e));
    }
    if (c->next_h != c->hs.size()) {
        (void)hipGraphDestroy(g);
        throw Error(SPARQ_INVALID, "the captured sequence ran fewer branching operators than its plan");
    }
    *graph = g;
    return c.release();
}

void capture_abort(State& st) {
    if (!st.capture) return;
    std::unique_ptr<Capture> c(st.capture);
    st.capture = nullptr;
    g_active_captures.fetch_sub(1);
    if (c->phase == 2) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
    }
}

__global__ void set_word_kernel(uint64_t* __restrict__ w, uint64_t v) { w[0] = v; }
__global__ void add_word_kernel(uint64_t* __restrict__ w, uint64_t v) { w[0] += v; }

void capture_set_count(State& st, uint64_t* words) {
    set_word_kernel<<<1, 1, 0, st.stream>>>(words + kCapN, st.n);
    SPQ_HIP(hipGetLastError());
}

void table_advance(State& st, uint64_t* word) {
    if (st.capture && st.capture->phase == 1) return;   // planning: nothing runs
    add_word_kernel<<<1, 1, 0, st.stream>>>(word, 1);
    SPQ_HIP(hipGetLastError());
}

}  // namespace sparq
