// Per-kernel event profiling: events recorded on the launch stream around
// each kernel, credited with the kernel's algorithmic HBM bytes.
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

namespace sparq {
namespace {
struct Pending {
    const char* name;
    double bytes;
    hipEvent_t ev0, ev1;
};
struct Acc {
    uint64_t launches = 0;
    double ms = 0, bytes = 0;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_pool;
std::map<std::string, Acc> g_acc;

hipEvent_t take() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    SPQ_HIP(hipEventCreate(&e));
    return e;
}

void drain() {
    for (auto& p : g_pending) {
        SPQ_HIP(hipEventSynchronize(p.ev1));
        float ms = 0;
        SPQ_HIP(hipEventElapsedTime(&ms, p.ev0, p.ev1));
        Acc& a = g_acc[p.name];
        a.launches += 1;
        a.ms += ms;
        a.bytes += p.bytes;
        g_pool.push_back(p.ev0);
        g_pool.push_back(p.ev1);
    }
    g_pending.clear();
}
}  // namespace

bool prof_enabled() { return g_on && !capture_active(); }

void prof_begin(hipStream_t s, const char* name, double bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_pending.size() > 4096) drain();
    Pending p{name, bytes, take(), take()};
    SPQ_HIP(hipEventRecord(p.ev0, s));
    g_pending.push_back(p);
}

void prof_host_wait(const char* name, double ms) {
    std::lock_guard<std::mutex> lk(g_mu);
    Acc& a = g_acc[name];
    a.launches += 1;
    a.ms += ms;
}

void prof_end(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_pending.empty()) SPQ_HIP(hipEventRecord(g_pending.back().ev1, s));
}

}  // namespace sparq

using namespace sparq;

extern "C" {
int sparq_prof_enable(int on) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = on != 0;
    return SPARQ_OK;
}
int sparq_prof_reset(void) {
    try {
        std::lock_guard<std::mutex> lk(g_mu);
        drain();
        g_acc.clear();
        return SPARQ_OK;
    } catch (const Error& e) {
        return e.status;
    }
}
int sparq_prof_report(sparq_prof_record_t* out, uint32_t capacity, uint32_t* count) {
    try {
        std::lock_guard<std::mutex> lk(g_mu);
        drain();
        uint32_t i = 0;
        for (auto& kv : g_acc) {
            if (out && i < capacity) {
                std::memset(&out[i], 0, sizeof(out[i]));
                std::strncpy(out[i].name, kv.first.c_str(), sizeof(out[i].name) - 1);
                out[i].launches = kv.second.launches;
                out[i].total_ms = kv.second.ms;
                out[i].total_bytes = kv.second.bytes;
            }
            ++i;
        }
        if (count) *count = i;
        return SPARQ_OK;
    } catch (const Error& e) {
        return e.status;
    }
}
}
