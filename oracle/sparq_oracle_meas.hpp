// MeasureZ / Reset / Probability of the CPU oracle, restated from
// SparQ/src/measurement.cpp, with the seedable engine of this build
// (splitmix64; the reference's random_engine lives in the absent submodule).
// The walk over entries runs in canonical order (SortUnconditional first), the
// order the engine samples in.  TEST INFRASTRUCTURE ONLY -- see sparq_oracle.hpp.
#pragma once

#include <map>

#include "sparq_oracle_cks.hpp"

namespace oracle {

struct RandomEngine {
    uint64_t seed = 0x5EED2026ull, state = 0x5EED2026ull;
    static RandomEngine& get() { static RandomEngine e; return e; }
    void set_seed(uint64_t s) { seed = state = s; }
    double uniform01() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

// measurement.cpp:127-182
inline std::pair<std::vector<uint64_t>, double> MeasureZ(State& st, const std::vector<size_t>& regs) {
    if (st.empty()) throw_invalid_input("MeasureZ: cannot measure an empty state.");
    SortUnconditional(st);
    double total = 0;
    for (auto& s : st) total += abs_sqr(s.amplitude);
    if (!std::isfinite(total) || total <= 0.0 || !(std::fabs(total - 1.0) < 1e-5))
        throw_runtime("MeasureZ: input state is not normalized");
    if (regs.empty()) return {{}, 1.0};
    double r = RandomEngine::get().uniform01();
    std::vector<uint64_t> outcome(regs.size(), 0);
    bool selected = false;
    for (auto& s : st) {
        double p = abs_sqr(s.amplitude);
        if (r < p) {
            for (size_t i = 0; i < regs.size(); ++i) outcome[i] = s.get(regs[i]);
            selected = true;
            break;
        }
        r -= p;
    }
    if (!selected)
        for (size_t i = 0; i < regs.size(); ++i) outcome[i] = st.back().get(regs[i]);
    double sum = 0;
    auto drop = [&](const System& s) {
        for (size_t i = 0; i < regs.size(); ++i)
            if (s.get(regs[i]) != outcome[i]) return true;
        sum += abs_sqr(s.amplitude);
        return false;
    };
    st.erase(std::remove_if(st.begin(), st.end(), drop), st.end());
    if (sum <= 0 || st.empty()) throw_runtime("MeasureZ: sampled outcome has zero support");
    double inv = 1.0 / std::sqrt(sum);
    for (auto& s : st) s.amplitude *= inv;
    return {outcome, sum};
}
// measurement.cpp:213-228
inline std::vector<uint64_t> Reset(State& st, const std::vector<size_t>& regs, const std::vector<uint64_t>& targets) {
    if (regs.empty()) return {};
    auto res = MeasureZ(st, regs);
    for (auto& s : st)
        for (size_t i = 0; i < regs.size(); ++i) s.get(regs[i]) = targets[i];
    return res.first;
}
// measurement.cpp:281-318
inline double Probability(const State& st, const std::vector<size_t>& regs, const std::vector<uint64_t>& values) {
    if (regs.empty()) return 1.0;
    double sum = 0;
    for (auto& s : st) {
        bool m = true;
        for (size_t i = 0; i < regs.size() && m; ++i) m = s.get(regs[i]) == values[i];
        if (m) sum += abs_sqr(s.amplitude);
    }
    return sum;
}
inline std::map<uint64_t, double> Distribution(const State& st, size_t reg) {
    std::map<uint64_t, double> d;
    for (auto& s : st) d[s.get(reg)] += abs_sqr(s.amplitude);
    return d;
}

}  // namespace oracle
