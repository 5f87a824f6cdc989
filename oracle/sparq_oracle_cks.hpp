// CKS quantum-walk primitives of the CPU oracle, restated from
// SparQ_Algorithm/src/hamiltonian_simulation.cpp, SparQ_Algorithm/src/hamiltonian_tridiagonal.cpp,
// SparQ_Algorithm/include/hamiltonian_simulation_tridiagonal.h, SparQ/src/partial_trace.cpp and
// SparQ/src/debugger.cpp.  TEST INFRASTRUCTURE ONLY -- see sparq_oracle.hpp.
#pragma once

#include <limits>

#include "sparq_oracle_ops.hpp"

namespace oracle {

// ---- sparse matrix in compact QRAM storage.  The memory is [elements | sparsity],
// zero-padded to a power of two; the tridiagonal case reproduces
// TridiagonalSpec::combined_mem (hamiltonian_simulation_tridiagonal.h:169-178).
// The float quantisation lives in the absent submodule; the reference's Python
// front end fixes its scale and range by clipping every value to
// (2^ds-1)/2^ds (positive only) or +-(2^(ds-1)-1)/2^(ds-1)
// (PySparQ/pysparq/algorithms/cks_solver.py:325-342).  Here: floor(x*2^ds) for
// positive-only matrices, the two's complement of trunc(x*2^(ds-1)) otherwise;
// both within one code step and exact on dyadics (tests/test_oracle.py).
struct SparseMatrix {
    std::vector<uint64_t> elements;
    std::vector<size_t> sparsity;
    size_t data_size = 0, nnz_col = 0, n_row = 0;
    bool positive_only = true;
    std::vector<uint64_t> get_data() const {
        size_t total = 1;
        while (total < elements.size() + sparsity.size()) total <<= 1;
        std::vector<uint64_t> mem(total, 0);
        for (size_t i = 0; i < elements.size(); ++i) mem[i] = elements[i];
        for (size_t i = 0; i < sparsity.size(); ++i) mem[elements.size() + i] = sparsity[i];
        return mem;
    }
    size_t get_sparsity_offset() const { return elements.size(); }
};
inline uint64_t quantise_element(double x, size_t ds, bool positive_only) {
    if (positive_only) {
        double v = std::floor(x * std::ldexp(1.0, (int)ds));
        v = std::min(std::max(v, 0.0), double(width_mask(ds)));
        return uint64_t(v);
    }
    const double top = double(width_mask(ds - 1));
    double v = std::trunc(x * std::ldexp(1.0, (int)ds - 1));
    v = std::min(std::max(v, -top), top);
    return uint64_t(int64_t(v)) & width_mask(ds);
}

// ---- GetRowAddr / GetDataAddr (hamiltonian_simulation.cpp:222-255): XOR of an affine address
inline OpPtr GetRowAddr(size_t offset, size_t row, size_t row_sz, size_t out) {
    return pw("GetRowAddr", [=](System& s) { s.get(out) ^= s.as_uint(offset) + row_sz * s.as_uint(row); });
}
inline OpPtr GetDataAddr(size_t offset, size_t row, size_t col, size_t row_sz, size_t out) {
    return pw("GetDataAddr", [=](System& s) { s.get(out) ^= s.as_uint(offset) + row_sz * s.as_uint(row) + s.as_uint(col); });
}

// ---- QuantumBinarySearch_Fast (hamiltonian_simulation.cpp:159-220)
inline size_t binary_search_memory(const std::vector<uint64_t>& mem, size_t offset, size_t total_length, size_t max_step,
                                   uint64_t target) {
    size_t l = offset, r = l + total_length;
    for (size_t i = 0; i < max_step; ++i) {
        size_t mid = (l + r) / 2;
        if (mid >= mem.size()) throw_runtime("QuantumBinarySearch_Fast: probe outside the memory");
        if (mem[mid] == target) return mid;
        if (mem[mid] < target) l = mid;
        else r = mid;
    }
    return 0;
}
inline OpPtr QuantumBinarySearch_Fast(std::shared_ptr<QRAMCircuit> q, size_t offset, size_t total_length, size_t target,
                                      size_t result) {
    if (!q || total_length == 0) throw_invalid_input();
    const size_t max_step = size_t(std::log2(double(total_length))) + 1;
    return pw("QuantumBinarySearch_Fast", [=](System& s) {
        s.get(result) ^= binary_search_memory(q->memory, s.as_uint(offset), total_length, max_step, s.as_uint(target));
    });
}

// ---- GetQWRotateAngle_Int_Int_Int (hamiltonian_simulation.cpp:257-290)
inline OpPtr GetQWRotateAngle_Int_Int_Int(size_t data, size_t row, size_t col, size_t out, const SparseMatrix& mat) {
    (void)row;
    (void)col;
    const size_t ds = mat.data_size;
    const bool pos = mat.positive_only;
    return pw("GetQWRotateAngle_Int_Int_Int", [=](System& s) {
        uint64_t v = s.as_uint(data);
        double ratio;
        if (pos) {
            uint64_t amax = pow2(ds) - 1;
            v = v % (amax + 1);
            ratio = v * 1.0 / amax;
        } else {
            uint64_t amax = pow2(ds - 1) - 1;
            v = v % pow2(ds);
            int64_t vr = get_complement(v, ds);
            ratio = std::abs(vr) * 1.0 / amax;
        }
        ratio = std::max(0.0, std::min(1.0, ratio));
        double o = std::acos(std::sqrt(ratio)) / pi / 2;
        s.get(out) ^= get_rational(o, Registry::size_of(out));
    });
}

// ---- the tridiagonal walk's direct oracles (hamiltonian_simulation_tridiagonal.h:104-396,
//      hamiltonian_tridiagonal.cpp:16-105)
struct TridiagonalSpec {
    size_t n_row = 0, data_size = 0;
    uint64_t alpha_enc = 0, beta_enc = 0;
    size_t sparse_offset = 0;
    uint64_t addr_mask = 0;
    size_t k_width = 0;
    static constexpr size_t slot_count = 4;
    size_t window_base(size_t j) const {
        size_t b = (j > 0) ? (j - 1) : 0;
        size_t bmax = n_row - slot_count;
        return b < bmax ? b : bmax;
    }
    size_t col_of(size_t j, size_t s) const { return window_base(j) + s; }
    uint64_t value_of(size_t j, size_t s) const {
        size_t c = col_of(j, s);
        if (c == j) return alpha_enc;
        if (c + 1 == j || c == j + 1) return beta_enc;
        return 0;
    }
    uint64_t combined_mem(uint64_t idx) const {
        idx &= addr_mask;
        uint64_t half = uint64_t(n_row) * slot_count;
        if (idx < half) return value_of(size_t(idx >> 2), size_t(idx & 3));
        uint64_t t = idx - half;
        return uint64_t(col_of(size_t(t >> 2), size_t(t & 3)));
    }
};
inline std::pair<SparseMatrix, TridiagonalSpec> make_tridiagonal_matrix_spec(double alpha, double beta, size_t n, size_t ds) {
    if (n < 4 || (n & (n - 1))) throw_invalid_input("tridiagonal walk needs a power-of-two dimension >= 4");
    double lam_max = std::fabs(alpha) + 2.0 * std::fabs(beta);
    double scale = 1.0 / lam_max;
    TridiagonalSpec spec;
    spec.n_row = n;
    spec.data_size = ds;
    spec.alpha_enc = uint64_t(std::floor(std::fabs(alpha) * scale * double(pow2(ds))));
    spec.beta_enc = uint64_t(std::floor(std::fabs(beta) * scale * double(pow2(ds))));
    spec.sparse_offset = n * TridiagonalSpec::slot_count;
    spec.addr_mask = 2 * spec.sparse_offset - 1;
    spec.k_width = std::max(size_t(std::log2(double(n))) + 1, ds);
    SparseMatrix mat;
    mat.data_size = ds;
    mat.nnz_col = TridiagonalSpec::slot_count;
    mat.n_row = n;
    mat.positive_only = true;
    for (size_t i = 0; i < n; ++i)
        for (size_t s = 0; s < TridiagonalSpec::slot_count; ++s) {
            mat.elements.push_back(spec.value_of(i, s));
            mat.sparsity.push_back(spec.col_of(i, s));
        }
    return {mat, spec};
}
inline OpPtr TridiagOracle1(size_t offset, size_t row, size_t slot, size_t out, const TridiagonalSpec& spec) {
    return pw("TridiagOracle1", [=](System& s) {
        s.get(out) ^= spec.combined_mem(s.as_uint(offset) + TridiagonalSpec::slot_count * s.as_uint(row) + s.as_uint(slot));
    });
}
inline OpPtr TridiagOracle2(size_t offset, size_t row, size_t k_id, size_t sr_id, const TridiagonalSpec& spec) {
    const uint64_t kmask = pow2(spec.k_width) - 1;
    auto fwd = [=](System& s) {   // column -> slot
        uint64_t j = s.as_uint(row), off = s.as_uint(offset), k = s.as_uint(k_id), sr = s.as_uint(sr_id);
        uint64_t base = spec.window_base(size_t(j));
        uint64_t row_addr = off + TridiagonalSpec::slot_count * j;
        bool hit = k >= base && k < base + TridiagonalSpec::slot_count;
        uint64_t sr1 = sr ^ (hit ? row_addr + (k - base) : 0);
        uint64_t k1 = k ^ spec.combined_mem(sr1);
        s.get(k_id) = (sr1 - row_addr) & kmask;
        s.get(sr_id) = k1 & kmask;
    };
    auto bwd = [=](System& s) {   // slot -> column
        uint64_t j = s.as_uint(row), off = s.as_uint(offset), k = s.as_uint(k_id), sr = s.as_uint(sr_id);
        uint64_t base = spec.window_base(size_t(j));
        uint64_t row_addr = off + TridiagonalSpec::slot_count * j;
        uint64_t k1 = (k + row_addr) & kmask;
        uint64_t k2 = sr ^ spec.combined_mem(k1);
        bool hit = k2 >= base && k2 < base + TridiagonalSpec::slot_count;
        s.get(k_id) = k2;
        s.get(sr_id) = k1 ^ (hit ? row_addr + (k2 - base) : 0);
    };
    return pw("TridiagOracle2", fwd, bwd);
}
inline OpPtr TridiagParityPhase(size_t reg) {
    return pw("TridiagParityPhase", [=](System& s) {
        if (s.as_uint(reg) & 1) s.amplitude = -s.amplitude;
    });
}

// ---- PartialTraceSelect / PartialTraceSelectRange (partial_trace.cpp:148-225)
inline double PartialTraceSelect(State& st, const std::vector<size_t>& regs, const std::vector<uint64_t>& values) {
    if (regs.empty() || st.empty()) return 0;
    double sum = 0;
    auto pred = [&](const System& s) {
        for (size_t i = 0; i < regs.size(); ++i)
            if (values[i] != s.get(regs[i])) return true;
        sum += abs_sqr(s.amplitude);
        return false;
    };
    st.erase(std::remove_if(st.begin(), st.end(), pred), st.end());
    if (sum == 0) return std::numeric_limits<double>::infinity();
    sum = 1.0 / std::sqrt(sum);
    for (auto& s : st) s.amplitude *= sum;
    return sum;
}
inline double PartialTraceSelectRange(State& st, size_t reg, size_t lo, size_t hi) {
    if (st.empty()) return 0;
    double sum = 0;
    auto pred = [&](const System& s) {
        uint64_t v = s.as_uint(reg);
        if (v >= lo && v <= hi) {
            sum += abs_sqr(s.amplitude);
            return false;
        }
        return true;
    };
    st.erase(std::remove_if(st.begin(), st.end(), pred), st.end());
    if (sum == 0) return std::numeric_limits<double>::infinity();
    sum = 1.0 / std::sqrt(sum);
    for (auto& s : st) s.amplitude *= sum;
    return sum;
}

// ---- CheckNormalization / CheckNan (debugger.cpp:13-70), always active in the oracle
inline void CheckNormalization(const State& st, double threshold = 1e-5) {
    double f = 0;
    for (auto& s : st) f += abs_sqr(s.amplitude);
    if (!(std::fabs(f - 1.0) < threshold) || std::isnan(f)) throw_runtime("State is not normalized.");
}
inline void CheckNan(const State& st) {
    for (auto& s : st)
        if (std::isnan(s.amplitude.real()) || std::isnan(s.amplitude.imag())) throw_runtime("bad result");
}

// ---- ChebyshevPolynomialCoefficient (hamiltonian_simulation.cpp:44-99)
struct ChebyshevPolynomialCoefficient {
    size_t b;
    explicit ChebyshevPolynomialCoefficient(size_t b_) : b(b_) {}
    double C(size_t Big, size_t Small) const {
        double ret = 1;
        bool flag = true;
        double pow2_b = std::pow(2, b);
        for (size_t i = 0; i < Small; ++i) {
            ret /= (Small - i);
            ret *= (Big - i);
            if (flag && ret > pow2_b) {
                flag = false;
                ret /= pow2_b;
            }
        }
        if (flag) ret /= pow2_b;
        ret /= pow2_b;
        return ret;
    }
    double coef(size_t j) const {
        if (b > 100) return std::erfc((j + 0.5) / std::sqrt(b)) * 2;
        double ret = 0;
        for (size_t i = j + 1; i <= b; ++i) ret += C(2 * b, b + i);
        return ret * 4;
    }
    bool sign(size_t j) const { return j & 1; }
    size_t step(size_t j) const { return 2 * j + 1; }
};

}  // namespace oracle
