// Internal declarations of the MI355X sparse-state engine (gfx950, wave64).
//
// Data layout in HBM (one state):
//   keys : uint64 [W][cap]  word-major, word 0 most significant; entry i's
//          word w at keys[w*cap + i], so every word plane is read coalesced.
//   amps : double2 [cap]    FP64 complex amplitude (re, im).
// Registers are bit fields of the key (sparq_field_t) assigned by the host
// layer in register-id order, most significant first, so unsigned
// lexicographic order of the words is the reference canonical order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/sparq_hip.h"

namespace sparq {

struct Error : std::exception {
    int status;
    std::string msg;
    Error(int s, std::string m) : status(s), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

#define SPQ_HIP(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::sparq::Error(e_ == hipErrorOutOfMemory ? SPARQ_OOM : SPARQ_RUNTIME,       \
                                 std::string(#expr " failed: ") + hipGetErrorString(e_));     \
    } while (0)

#define SPQ_REQUIRE(cond, msg)                                  \
    do {                                                        \
        if (!(cond)) throw ::sparq::Error(SPARQ_INVALID, msg);  \
    } while (0)

// C-ABI wrapper: runs f, maps exceptions to status codes and records the
// message for sparq_last_error (api.hip).
void set_last_error(const std::string& msg);
template <class F>
int guarded_call(F&& f) {
    try {
        f();
        return SPARQ_OK;
    } catch (const Error& e) {
        set_last_error(e.msg);
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return SPARQ_OOM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return SPARQ_RUNTIME;
    } catch (...) {
        set_last_error("unknown error");
        return SPARQ_RUNTIME;
    }
}
// Selects device 0 on first use; throws SPARQ_NO_DEVICE without a GPU.
void ensure_device();

constexpr int kBlock = 256;           // threads per workgroup (4 waves of 64)
constexpr int kWave = 64;

struct Capture;

struct State {
    uint32_t W = 1;
    uint32_t Wcap = 0;          // key words the allocation holds when more than W (0: exactly W)
    uint64_t n = 0, cap = 0;
    uint64_t* keys = nullptr;   // [W][cap]
    double2* amps = nullptr;    // [cap]
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    uint64_t* pinned = nullptr; // 64 words of pinned host memory for small readbacks
    unsigned int* counter = nullptr;   // device word, zero between kernels (last-block reductions)
    int device = 0;
    // One entry of amplitude exactly (1, 0), written by set_basis and untouched
    // since (every API access but Hadamard_Int's clears it).  Never set once the
    // device arrays have been handed to a caller, who may write them directly.
    bool basis_amp_one = false;
    bool arrays_escaped = false;
    // Entries are grouped in pairs across key bit pair_bit (word * 64 + shift;
    // -1: not known): wherever a key and the key with that bit flipped are both
    // present they are adjacent, the one with the bit clear first.  Set by the
    // pair operators (whose output is laid out so); kept by accesses that only
    // drop entries in order, change amplitudes, or change keys without reading
    // or writing that bit; cleared by every other access (api.hip's accessors),
    // and never kept once the arrays have been handed to a caller.
    int32_t pair_bit = -1;
    // Graph capture in progress on this state (capture.hip), or null.
    Capture* capture = nullptr;
    // Replays of a captured sequence are queued and n is on the device until
    // the next access settles it: {entry count, violation flags} (capture.hip).
    const uint64_t* pending_words = nullptr;
};

// ---- memory (caching device allocator)
void* dmalloc(size_t bytes, hipStream_t s);
void dfree(void* p, hipStream_t s);
void forget_stream(hipStream_t s);          // untag cached blocks of a drained stream
void memory_stats(uint64_t out[6]);         // live, cached, peak bytes; hits, misses, cached blocks
void release_cached_memory();               // synchronise the device and return cached blocks
template <class T> T* dalloc(size_t count, hipStream_t s) { return static_cast<T*>(dmalloc(count * sizeof(T), s)); }

// Allocate fresh key/amp arrays of capacity `cap` for `st` (content undefined).
void alloc_arrays(State& st, uint32_t W, uint64_t cap, uint64_t** keys, double2** amps);
// Install new arrays as the state's content, freeing the old ones.
void install_arrays(State& st, uint32_t W, uint64_t* keys, double2* amps, uint64_t n, uint64_t cap);
// Change the key width in place: words below min(old, new) keep their content,
// added words read as zero, removed words are dropped.
void resize_words(State& st, uint32_t new_words);
// Read `count` uint64 values from device memory (synchronizes the stream).
void read_back(State& st, const void* dptr, size_t bytes, void* host);
uint64_t read_u64(State& st, const uint64_t* dptr);

inline uint64_t fmask_host(uint32_t width) { return width >= 64 ? ~uint64_t(0) : ((uint64_t(1) << width) - 1); }

inline unsigned grid_for(uint64_t n, unsigned per_block) {
    uint64_t g = (n + per_block - 1) / per_block;
    if (g == 0) g = 1;
    return (unsigned)g;
}

// ---- device-wide primitives (scan.hip)
// Exclusive scan of `n` uint32 values into uint64 offsets; returns the total (syncs).
void exclusive_scan_u32(State& st, const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* total_dev);
void exclusive_scan_u64(State& st, const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* total_dev);
// The u32 scan on stream `s` with its partial sums in `scratch` (2 words per
// 2048 inputs per level above the first); allocates nothing.  n >= 1.
void exclusive_scan_u32_scratch(hipStream_t s, const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* total_dev, uint64_t* scratch,
                                uint64_t scratch_words);

// ---- sort (sort.hip)
// A digit of the LSD plan: up to three bit ranges of key word `word`,
// concatenated with range 0 least significant; `bits` is their total width.
struct Digit {
    uint8_t word, nseg, bits, pad;
    uint8_t shift[3], width[3];
};
// Stable LSD radix sort of the state by the given digit plan (least significant
// digit first, every digit at most radix_bits wide).  Reorders keys and amps.
// pass_begin / pass_end (optional) are recorded around the digit passes alone.
// word_src (optional, [words][word_cap]) replaces the state's keys as the source
// of the digits; the state's keys and amplitudes are permuted either way.
void sort_by_plan(State& st, const std::vector<Digit>& plan, int radix_bits = 8, hipEvent_t pass_begin = nullptr,
                  hipEvent_t pass_end = nullptr, const uint64_t* word_src = nullptr, uint64_t word_cap = 0);
// Digit width forced for every sort (8 or 11), or 0 for the autotuner.
void set_radix_bits(int bits);
// 0: prefix passes + LDS-local chunk sort where it applies, 1: LSD passes only.
void set_sort_mode(int mode);
// Hadamard_Int on a small, well filled key domain through a dense amplitude
// array (on by default); off: always the grouped path.
void set_dense_transform(bool on);
void set_canonical_output(bool on);
// Entries per thread of the pointwise kernel: 1, 2 or 4 forced, 0 autotuned.
void set_pointwise_items(int items);
// The autotuners' measurements and choices as JSON arrays.
std::string pointwise_tuning_report();
std::string radix_tuning_report();
// out[6]: sorts by the prefix + local path, their entries, local attempts that
// overflowed a chunk, sorts by the LSD passes, their entries, and the local
// sorts that ran no prefix pass (the state was already sorted by the prefix);
// reset clears them.
void sort_path_stats(uint64_t* out, int reset);
// One pass over the keys: bits that vary across the state, and whether the
// state is already ordered with the optional field `last` least significant.
// With a SizerQuery the same pass also histograms popcount(value of `last`)
// into kStatsHistBins bins and the sizer's feature kernel runs on the result
// (and its network when active) before the single readback.
constexpr int kStatsHistBins = 6;
// Key-statistics record: [0, W) OR of the keys, [W, 2W) AND, [2W] order
// violations of (others, last), [2W+1] violations of `last` alone being
// non-decreasing, [2W+2, 3W+2) per word the first differing bits of the pairs
// out of order, [3W+2, +kStatsHistBins) popcount histogram of `last`.
constexpr int stats_words(int W) { return 3 * W + 2 + kStatsHistBins; }
struct SizerQuery {
    uint64_t n;
    uint32_t words, nd, kind;          // kind: 0 Hadamard_Int, 1 Hadamard_Int_Full, 2 dense group unitary
    uint32_t target_word, target_shift;
};
struct KeyStats {
    std::vector<uint64_t> vary;
    std::vector<uint64_t> ones;        // OR of the keys (the constant bits' values where vary is 0)
    bool sorted = true;
    bool last_sorted = true;           // the `last` field alone is non-decreasing
    // per word, the first differing bits of adjacent pairs out of (others, last)
    // order: the state is sorted by the bits above the most significant one
    std::vector<uint64_t> disorder;
    bool have_features = false, have_prediction = false;
    float features[SPARQ_SIZER_FEATURES] = {};
    float prediction = 0.f;            // network output: log2(G) / 32
};
KeyStats key_stats(State& st, const sparq_field_t* last, const SizerQuery* q = nullptr);

// ---- learned sizer (sizer.hip)
bool sizer_wanted(bool& predict);     // logging or prediction requested; predict = network active
void launch_sizer_features(State& st, const uint64_t* stats, const SizerQuery& q, bool predict, float* out);
float sizer_margin();
void sizer_record(const float* features, float target);
void sizer_count(int what);           // 0 exact path, 1 predicted, 2 overflow fallback
// LSD plan over the varying bits with `last` moved to the least significant
// place; the digits of `last` are left out when it is already non-decreasing.
std::vector<Digit> make_plan(const KeyStats& ks, uint32_t W, const sparq_field_t* last, int digit_bits);
// Sort by (others, last) given the statistics (digit width by the autotuner).
void sort_with_stats(State& st, const KeyStats& ks, const sparq_field_t* last);
// Sort unless already ordered; returns whether a sort ran.
bool sort_state(State& st, const sparq_field_t* last);
// Stable sort by (major fields in order, the other key bits canonically, minor
// fields in order); the fields must not overlap.
void sort_state_by(State& st, const std::vector<sparq_field_t>& major, const std::vector<sparq_field_t>& minor);
// Stable sort by |amplitude|^2, largest first.
void sort_by_amplitude(State& st);

// ---- compaction (compact.hip)
void clear_zero(State& st, double eps);
// Dense post-selection into out_host[2*len] (state.hip); returns the selected weight.
double post_select_dense(State& st, const sparq_field_t* zero, uint32_t nzero, const sparq_field_t* index, uint32_t nindex,
                         double* out_host, uint64_t len);
// ClearZero of (keys, amps, cap) holding counts[2] entries (a device word); reads
// counts[0..4) back into info with the kept count in info[3].
void clear_zero_sized(State& st, uint64_t* keys, double2* amps, uint64_t cap, uint64_t* counts, double eps, uint64_t info[4],
                      uint64_t** kout, double2** aout);
void merge_sorted_duplicates(State& st, double eps);
bool has_sorted_duplicates(State& st);

// ---- condition split (compact.hip): stable partition so that entries that
// satisfy `conds` come first; returns their count.
uint64_t partition_by_conditions(State& st, const sparq_cond_t* conds, uint32_t ncond);
// Keep only the entries satisfying every condition (in order); returns the kept count.
uint64_t select_by_conditions(State& st, const sparq_cond_t* conds, uint32_t ncond);

// ---- branching (branch.hip)
// unit_basis: st is one entry of amplitude exactly (1, 0) (lets the one-entry
// expansion decide its ClearZero on the host).  check_dups: throw on duplicate
// keys before changing anything (the dense-domain path finds them by its slot
// marks, the grouping path after its sort).
void hadamard_int(State& st, sparq_field_t target, uint32_t n_digits, bool full_variant, double eps,
                  bool unit_basis = false, bool check_dups = false);
void replicate(State& st, sparq_field_t f);
// With conditions (which read no target bit), groups whose entries fail them
// pass through unchanged, in place in the group-ordered output.
struct CondList;
void group_unitary(State& st, sparq_field_t target, const double* mat, double eps, const CondList* cl = nullptr);
// Conditions as a kernel argument (AND of all).
struct CondList {
    uint32_t n;
    sparq_cond_t c[SPARQ_MAX_PROGRAM_CONDS];
};
// With conditions, entries that fail them pass through unchanged; the
// conditions must not read the target bit (the caller checks).  The result is
// ClearZero'd with eps (every entry, as the reference's ClearZero after the
// operator).
void pair_op(State& st, sparq_field_t target, uint32_t bit, const sparq_pair_params_t& p, const CondList& cl, double eps);

// ---- pointwise (pointwise.hip)
void pointwise(State& st, const sparq_pw_op_t* ops, uint32_t nops, const sparq_cond_t* conds, uint32_t nconds);
void relayout(State& st, uint32_t new_words, const sparq_field_t* from, const sparq_field_t* to, uint32_t nf);
struct BasisKey { uint64_t w[SPARQ_MAX_WORDS]; };
// Entry 0 := |key>, amplitude 1 (the state's arrays must hold at least one entry).
void write_basis(State& st, const BasisKey& k);
double sum_abs2(State& st);
// |a|^2 summed per tile of `tile` entries, read back into host_out[ntiles].
void tile_abs2(State& st, uint64_t tile, double* host_out, uint64_t ntiles);
// |a|^2 summed over the entries satisfying every condition.
double cond_abs2(State& st, const sparq_cond_t* conds, uint32_t ncond);
void scale_amps(State& st, double s);
bool field_uniform(State& st, sparq_field_t f, uint64_t* value);

// ---- graph capture (capture.hip)
// A fixed operator sequence is recorded once on a state and replayed as a HIP
// graph.  Phase 1 (plan) runs the sequence without launching anything: each
// branching operator plans its transform from a host snapshot of the state
// (the key bits that vary, the segments its conditions select).  Phase 2
// (capture) runs it again while the state's stream is capturing: every kernel
// reads the entry count from device memory and grid-strides over it, so no
// size crosses to the host and nothing is allocated.  What the plan assumed is
// checked on the device into a flags word, read when the replays settle.
enum CaptureFlag : uint64_t {
    kCapOutsidePlan = 1,     // an entry of a branching op fell outside its planned segments or constant bits
    kCapDuplicate = 2,       // duplicate keys in a branching op's domain
    kCapOverflow = 4,        // more entries than the captured capacity
    kCapNonzeroRemoved = 8,  // a removed register was not zero
};
// Device words of a capture: entry count, flags, then per-branching-op counters.
enum CaptureWord { kCapN = 0, kCapFlags = 1, kCapSat = 2, kCapMarks = 3, kCapRest = 4, kCapWords = 8 };
struct CapHadamard;   // one planned Hadamard_Int (branch.hip)
struct Capture {
    int phase = 1;                        // 1 plan, 2 capture
    uint32_t W = 1;
    uint64_t n0 = 0;                      // entries when the plan began
    std::vector<uint64_t> snap;           // phase 1: the state's keys, word-major [W][n0]
    std::vector<CapHadamard*> hs;         // planned branching ops, in sequence order
    size_t next_h = 0;                    // phase 2: the next planned op to record
    uint64_t capacity = 0;                // entries the sequence may hold (the state's arrays hold as many)
    // phase 2 device memory: one allocation, owned by the graph once captured
    char* arena = nullptr;
    uint64_t* words = nullptr;            // kCapWords words (CaptureWord)
    double2* A = nullptr;                 // dense amplitudes of the largest planned domain
    uint8_t* mark = nullptr;
    uint32_t* kept = nullptr;             // kept slots per slot tile
    uint64_t* offs = nullptr;             // their exclusive scan (+ total)
    uint32_t* rest_cnt = nullptr;         // entries failing a branching op's conditions, per entry tile
    uint64_t* rest_off = nullptr;
    uint64_t* rest_keys = nullptr;        // [W][capacity]
    double2* rest_amps = nullptr;
    uint64_t* scan_scratch = nullptr;
    uint64_t scan_words = 0;
    uint64_t max_slots = 0;               // largest planned domain (slots, padded to whole slot tiles)
    bool any_conditions = false;
    ~Capture();
};
inline uint64_t* capture_words(State& st) { return st.capture ? st.capture->words : nullptr; }
// Branching ops in a capture (branch.hip): plan (phase 1) or record (phase 2).
void hadamard_captured(State& st, sparq_field_t target, uint32_t nd, double eps, const sparq_cond_t* conds, uint32_t ncond);
// Device memory and tables of the planned ops (phase 1 -> 2), sizes in bytes.
uint64_t capture_plan_bytes(const Capture& c);
void capture_plan_upload(Capture& c, char*& at);
void capture_plan_free(Capture& c);
uint64_t capture_plan_slots(const CapHadamard& p);          // planned domain (segments << digits)
uint64_t capture_plan_slots_padded(const CapHadamard& p);   // ... in whole slot tiles
uint64_t capture_plan_rest(const CapHadamard& p);           // planned entries failing its conditions
bool capture_plan_has_conditions(const CapHadamard& p);
// Lifecycle (capture.hip): plan, record (starts stream capture), end (returns
// the capture, which the graph keeps), abort.
void capture_plan_begin(State& st);
void capture_record_begin(State& st);
Capture* capture_end(State& st, hipGraph_t* graph);
void capture_abort(State& st);
// Queue "entry count := st.n" on the state's stream (before replays).
void capture_set_count(State& st, uint64_t* words);
// word += 1 on the state's stream (capturable; nothing while planning).
void table_advance(State& st, uint64_t* word);
// Grid of a persistent grid-stride kernel over up to `cap` items of `per_block`.
unsigned persistent_grid(uint64_t cap, unsigned per_block);
// Graph capture is active somewhere in the process (profiling events stay out of it).
bool capture_active();

// ---- profiling (prof.hip)
bool prof_enabled();
void prof_begin(hipStream_t s, const char* name, double bytes);
void prof_end(hipStream_t s);
// A host-side wait (a readback's copy + stream synchronisation), recorded under
// a name starting with "host:" with its count and blocked milliseconds.
void prof_host_wait(const char* name, double ms);
struct ProfScope {
    bool on;
    hipStream_t s;
    ProfScope(hipStream_t s_, const char* name, double bytes) : on(prof_enabled()), s(s_) {
        if (on) prof_begin(s, name, bytes);
    }
    ~ProfScope() {
        if (on) prof_end(s);
    }
};

}  // namespace sparq

#define SPQ_CAT2(a, b) a##b
#define SPQ_CAT(a, b) SPQ_CAT2(a, b)
// Bracket the next kernel launch with profiling events (no-op unless enabled).
#define SPQ_PROF(stream, name, bytes) ::sparq::ProfScope SPQ_CAT(spq_prof_, __LINE__)((stream), (name), (double)(bytes))

// Dispatch a templated launcher on the runtime key width W (1..8).
#define SPQ_DISPATCH_W(Wv, FN, ...)                                         \
    switch (Wv) {                                                           \
        case 1: FN<1>(__VA_ARGS__); break;                                  \
        case 2: FN<2>(__VA_ARGS__); break;                                  \
        case 3: FN<3>(__VA_ARGS__); break;                                  \
        case 4: FN<4>(__VA_ARGS__); break;                                  \
        case 5: FN<5>(__VA_ARGS__); break;                                  \
        case 6: FN<6>(__VA_ARGS__); break;                                  \
        case 7: FN<7>(__VA_ARGS__); break;                                  \
        case 8: FN<8>(__VA_ARGS__); break;                                  \
        default: throw ::sparq::Error(SPARQ_INVALID, "key width out of range"); \
    }
