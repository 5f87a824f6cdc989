// Per-entry operators of the CPU oracle (key rewrites and diagonal phases),
// restated from the reference, plus a uniform operator object used by the
// Python binding.  TEST INFRASTRUCTURE ONLY -- see sparq_oracle.hpp.
#pragma once

#include <memory>

#include "sparq_oracle.hpp"

namespace oracle {

// Uniform operator object: forward / dagger over a State plus conditions.
struct Op {
    Conditions cond;
    std::string name;
    virtual ~Op() = default;
    virtual void apply(State& st) const = 0;
    virtual void apply_dag(State& st) const { apply(st); }
};
using OpPtr = std::shared_ptr<Op>;

// Per-entry op: f runs on condition-satisfied entries (the shared loop shape of
// every operator in quantum_arithmetic.cpp / qram.cpp:15-35 / basic_gates.cpp).
struct PointwiseOp : Op {
    std::function<void(System&)> fwd, bwd;
    void apply(State& st) const override { for_each_satisfied(st, cond, fwd); }
    void apply_dag(State& st) const override { for_each_satisfied(st, cond, bwd ? bwd : fwd); }
};
inline OpPtr pw(const char* name, std::function<void(System&)> f, std::function<void(System&)> b = nullptr) {
    auto p = std::make_shared<PointwiseOp>();
    p->name = name; p->fwd = std::move(f); p->bwd = std::move(b);
    return p;
}
inline uint64_t wm(size_t id) { return width_mask(Registry::size_of(id)); }

// ---- bit-level (quantum_arithmetic.cpp:11-61, basic_gates.cpp:383-400)
inline OpPtr X_Bool(size_t id, size_t digit) {
    if (digit >= Registry::size_of(id)) throw_invalid_input();
    return pw("X_Bool", [=](System& s) { s.get(id) = flip_digit(s.get(id), digit); });
}
inline OpPtr FlipBools(size_t id) {
    return pw("FlipBools", [=](System& s) { s.get(id) = (~s.get(id)) & wm(id); });
}
inline OpPtr Swap_Bool_Bool(size_t lhs, size_t d1, size_t rhs, size_t d2) {
    return pw("Swap_Bool_Bool", [=](System& s) {
        uint64_t& r1 = s.get(lhs); uint64_t& r2 = s.get(rhs);
        bool v1 = get_digit(r1, d1), v2 = get_digit(r2, d2);
        if (v1 && !v2) { r1 -= pow2(d1); r2 += pow2(d2); }
        if (v2 && !v1) { r1 += pow2(d1); r2 -= pow2(d2); }
    });
}
inline OpPtr Swap_General_General(size_t a, size_t b) {          // :982-999
    return pw("Swap_General_General", [=](System& s) { std::swap(s.get(a), s.get(b)); });
}
// ---- shifts (:63-135)
inline void shl(System& s, size_t r, size_t d) {
    size_t size = Registry::size_of(r);
    uint64_t v = s.as_uint(r);
    uint64_t high = v >> (size - d);
    uint64_t low = v - (high << (size - d));
    s.get(r) = (low << d) + high;
}
inline void shr(System& s, size_t r, size_t d) {
    size_t size = Registry::size_of(r);
    uint64_t v = s.as_uint(r);
    uint64_t high = v >> d;
    uint64_t low = v - (high << d);
    s.get(r) = (low << (size - d)) + high;
}
inline OpPtr ShiftLeft_InPlace(size_t r, size_t d) {
    if (d > Registry::size_of(r)) throw_invalid_input();
    return pw("ShiftLeft_InPlace", [=](System& s) { shl(s, r, d); }, [=](System& s) { shr(s, r, d); });
}
inline OpPtr ShiftRight_InPlace(size_t r, size_t d) {
    if (d > Registry::size_of(r)) throw_invalid_input();
    return pw("ShiftRight_InPlace", [=](System& s) { shr(s, r, d); }, [=](System& s) { shl(s, r, d); });
}
// ---- multiplication by constants (:137-285)
inline OpPtr Mult_UInt_ConstUInt(size_t lhs, uint64_t m, size_t res) {
    return pw("Mult_UInt_ConstUInt", [=](System& s) { s.get(res) = (s.get(res) ^ (s.as_uint(lhs) * m)) & wm(res); });
}
inline OpPtr Add_Mult_UInt_ConstUInt_InPlace(size_t lhs, uint64_t m, size_t res) {
    return pw("Add_Mult_UInt_ConstUInt_InPlace",
        [=](System& s) { s.get(res) = (s.get(res) + m * s.as_uint(lhs)) & wm(res); },
        [=](System& s) { s.get(res) = (s.get(res) - s.as_uint(lhs) * m) & wm(res); });
}
inline OpPtr Mod_Mult_UInt_ConstUInt_InPlace(size_t reg, uint64_t a, uint64_t x, uint64_t N) {
    if (std::gcd(a, N) > 1) throw_invalid_input("a and N must be coprime");
    uint64_t opnum = a % N;
    for (size_t i = 0; i < x; ++i) opnum = (opnum * opnum) % N;
    int64_t t = 0, nt = 1, r = (int64_t)N, nr = (int64_t)opnum;
    while (nr != 0) { int64_t q = r / nr, tt = t - q * nt, tr = r - q * nr; t = nt; r = nr; nt = tt; nr = tr; }
    uint64_t inv = 1;
    if (r == 1) { int64_t tn = t % (int64_t)N; if (tn < 0) tn += (int64_t)N; inv = (uint64_t)tn; }
    return pw("Mod_Mult_UInt_ConstUInt_InPlace",
        [=](System& s) { s.get(reg) = (s.as_uint(reg) * opnum) % N; },
        [=](System& s) { s.get(reg) = (s.as_uint(reg) * inv) % N; });
}
// ---- additions (:287-411)
inline OpPtr Add_UInt_UInt(size_t lhs, size_t rhs, size_t res) {
    return pw("Add_UInt_UInt", [=](System& s) { s.get(res) = (s.get(res) ^ (s.as_uint(lhs) + s.as_uint(rhs))) & wm(res); });
}
inline OpPtr Add_UInt_UInt_InPlace(size_t lhs, size_t rhs) {
    return pw("Add_UInt_UInt_InPlace",
        [=](System& s) { s.get(rhs) = (s.get(rhs) + s.as_uint(lhs)) & wm(rhs); },
        [=](System& s) { s.get(rhs) = (s.get(rhs) - s.as_uint(lhs)) & wm(rhs); });
}
inline OpPtr Add_UInt_ConstUInt(size_t lhs, uint64_t add, size_t res) {
    return pw("Add_UInt_ConstUInt", [=](System& s) { s.get(res) = (s.get(res) ^ (s.as_uint(lhs) + add)) & wm(res); });
}
inline OpPtr Add_ConstUInt_InPlace(size_t reg, uint64_t add) {
    return pw("Add_ConstUInt_InPlace",
        [=](System& s) { s.get(reg) = (s.get(reg) + add) & wm(reg); },
        [=](System& s) { s.get(reg) = (s.get(reg) - add) & wm(reg); });
}
// ---- rational outputs (:413-488)
inline OpPtr Div_Sqrt_Arccos_UInt_UInt(size_t l, size_t r, size_t out) {
    return pw("Div_Sqrt_Arccos_UInt_UInt", [=](System& s) {
        double o = std::acos(std::sqrt(1.0 * s.as_uint(l) / s.as_uint(r))) / pi / 2;
        s.get(out) ^= get_rational(o, Registry::size_of(out));
    });
}
inline OpPtr Sqrt_Div_Arccos_Int_UInt(size_t l, size_t r, size_t out) {
    return pw("Sqrt_Div_Arccos_Int_UInt", [=](System& s) {
        double o = std::acos(s.as_int(l) / std::sqrt((double)s.as_uint(r))) / pi / 2;
        s.get(out) ^= get_rational(o, Registry::size_of(out));
    });
}
inline OpPtr GetRotateAngle_Int_Int(size_t l, size_t r, size_t out) {
    return pw("GetRotateAngle_Int_Int", [=](System& s) {
        int64_t lv = get_complement(s.as_uint(l), Registry::size_of(l));
        int64_t rv = get_complement(s.as_uint(r), Registry::size_of(r));
        double o;
        if (lv == 0 && rv >= 0) o = 0.25;
        else if (lv == 0 && rv < 0) o = 0.75;
        else { o = std::atan2((double)rv, (double)lv) / pi / 2; if (o < 0) o += 1; }
        s.get(out) ^= get_rational(o, Registry::size_of(out));
    });
}
// ---- XOR-into binary ops (:490-707)
inline OpPtr Sub_UInt_UInt(size_t l, size_t r, size_t res) {
    return pw("Sub_UInt_UInt", [=](System& s) { s.get(res) = (s.get(res) ^ (s.as_uint(l) - s.as_uint(r))) & wm(res); });
}
inline OpPtr Neg_UInt(size_t reg, size_t res) {
    return pw("Neg_UInt", [=](System& s) { s.get(res) = (s.get(res) ^ (uint64_t(0) - s.as_uint(reg))) & wm(res); });
}
inline OpPtr Abs_SInt(size_t reg, size_t res) {
    return pw("Abs_SInt", [=](System& s) {
        int64_t v = get_complement(s.as_uint(reg), Registry::size_of(reg));
        uint64_t mag = v < 0 ? (uint64_t)(-v) : (uint64_t)v;
        s.get(res) = (s.get(res) ^ mag) & wm(res);
    });
}
inline OpPtr Mul_UInt_UInt(size_t l, size_t r, size_t res) {
    return pw("Mul_UInt_UInt", [=](System& s) { s.get(res) = (s.get(res) ^ (s.as_uint(l) * s.as_uint(r))) & wm(res); });
}
inline OpPtr Div_UInt_UInt(size_t l, size_t r, size_t res) {
    return pw("Div_UInt_UInt", [=](System& s) {
        uint64_t a = s.as_uint(l), b = s.as_uint(r);
        s.get(res) = (s.get(res) ^ (b == 0 ? 0 : a / b)) & wm(res);
    });
}
inline OpPtr Sqrt_UInt(size_t reg, size_t res) {
    return pw("Sqrt_UInt", [=](System& s) { s.get(res) = (s.get(res) ^ isqrt_u64(s.as_uint(reg))) & wm(res); });
}
inline OpPtr Select_Bool_UInt_UInt(size_t c, size_t l, size_t r, size_t res) {
    return pw("Select_Bool_UInt_UInt", [=](System& s) {
        uint64_t sel = (s.as_uint(c) & 1) ? s.as_uint(l) : s.as_uint(r);
        s.get(res) = (s.get(res) ^ sel) & wm(res);
    });
}
inline OpPtr And_UInt_UInt(size_t l, size_t r, size_t res) {
    return pw("And_UInt_UInt", [=](System& s) { s.get(res) = (s.get(res) ^ (s.as_uint(l) & s.as_uint(r))) & wm(res); });
}
inline OpPtr Or_UInt_UInt(size_t l, size_t r, size_t res) {
    return pw("Or_UInt_UInt", [=](System& s) { s.get(res) = (s.get(res) ^ (s.as_uint(l) | s.as_uint(r))) & wm(res); });
}
inline OpPtr Xor_UInt_UInt(size_t l, size_t r, size_t res) {
    return pw("Xor_UInt_UInt", [=](System& s) { s.get(res) = (s.get(res) ^ (s.as_uint(l) ^ s.as_uint(r))) & wm(res); });
}
// ---- predicate flags (:709-867)
inline OpPtr Less_SInt_SInt(size_t l, size_t r, size_t flag) {
    return pw("Less_SInt_SInt", [=](System& s) {
        bool p = get_complement(s.as_uint(l), Registry::size_of(l)) < get_complement(s.as_uint(r), Registry::size_of(r));
        s.get(flag) ^= (p ? 1 : 0);
    });
}
inline OpPtr Carry_UInt_UInt(size_t l, size_t r, size_t res, size_t flag) {
    return pw("Carry_UInt_UInt", [=](System& s) {
        uint64_t a = s.as_uint(l), b = s.as_uint(r);
        size_t w = Registry::size_of(res);
        bool p = w >= 64 ? (a + b < a) : (a >= pow2(w)) || (b >= pow2(w) - a);
        s.get(flag) ^= (p ? 1 : 0);
    });
}
inline OpPtr Overflow_SInt_SInt(size_t l, size_t r, size_t res, size_t flag) {
    return pw("Overflow_SInt_SInt", [=](System& s) {
        size_t w = Registry::size_of(res);
        uint64_t m = width_mask(w);
        uint64_t A = (uint64_t)get_complement(s.as_uint(l), Registry::size_of(l)) & m;
        uint64_t B = (uint64_t)get_complement(s.as_uint(r), Registry::size_of(r)) & m;
        uint64_t S = (A + B) & m;
        uint64_t sa = (A >> (w - 1)) & 1, sb = (B >> (w - 1)) & 1, ss = (S >> (w - 1)) & 1;
        s.get(flag) ^= ((sa == sb) && (ss != sa)) ? 1 : 0;
    });
}
inline OpPtr MulOverflow_UInt_UInt(size_t l, size_t r, size_t res, size_t flag) {
    return pw("MulOverflow_UInt_UInt", [=](System& s) {
        uint64_t a = s.as_uint(l), b = s.as_uint(r);
        unsigned __int128 p = (unsigned __int128)a * b;
        uint64_t lo = (uint64_t)p, hi = (uint64_t)(p >> 64);
        size_t w = Registry::size_of(res);
        s.get(flag) ^= (hi != 0 || (w < 64 && lo >= pow2(w))) ? 1 : 0;
    });
}
inline OpPtr IsZero_UInt(size_t reg, size_t flag) {
    return pw("IsZero_UInt", [=](System& s) { s.get(flag) ^= (s.as_uint(reg) == 0) ? 1 : 0; });
}
inline OpPtr Negative_SInt(size_t reg, size_t flag) {
    return pw("Negative_SInt", [=](System& s) {
        s.get(flag) ^= (get_complement(s.as_uint(reg), Registry::size_of(reg)) < 0) ? 1 : 0;
    });
}
inline uint64_t extended_rhs(const System& s, size_t id) {     // :858-867
    size_t w = Registry::size_of(id);
    uint64_t raw = s.as_uint(id);
    if (Registry::type_of(id) == SignedInteger) return (uint64_t)get_complement(raw, w);
    return raw;
}
inline OpPtr Add_AnyInt_AnyInt_InPlace(size_t lhs, size_t rhs) {
    return pw("Add_AnyInt_AnyInt_InPlace",
        [=](System& s) { s.get(lhs) = (s.get(lhs) + extended_rhs(s, rhs)) & wm(lhs); },
        [=](System& s) { s.get(lhs) = (s.get(lhs) - extended_rhs(s, rhs)) & wm(lhs); });
}
inline OpPtr Assign(size_t r1, size_t r2) {                       // :909-930
    if (r1 == r2) throw_invalid_input();
    return pw("Assign", [=](System& s) { s.get(r2) = (s.get(r2) ^ s.get(r1)) & wm(r2); });
}
inline OpPtr Compare_UInt_UInt(size_t l, size_t r, size_t lt, size_t eq) {   // :932-959
    if (lt == eq || lt == l || lt == r || eq == l || eq == r) throw_invalid_input();
    return pw("Compare_UInt_UInt", [=](System& s) {
        uint64_t a = s.as_uint(l), b = s.as_uint(r);
        if (a == b) s.get(eq) ^= 1; else s.get(lt) ^= (a < b) ? 1 : 0;
    });
}
inline OpPtr Less_UInt_UInt(size_t l, size_t r, size_t lt) {       // :961-980
    return pw("Less_UInt_UInt", [=](System& s) { s.get(lt) ^= (s.as_uint(l) < s.as_uint(r)) ? 1 : 0; });
}
inline OpPtr GetMid_UInt_UInt(size_t l, size_t r, size_t mid) {     // :1001-1020
    return pw("GetMid_UInt_UInt", [=](System& s) { s.get(mid) ^= (s.as_uint(l) + s.as_uint(r)) / 2; });
}

// ---- QRAM (qram.cpp:15-35); the memory table of a noise-free QRAM circuit.
struct QRAMCircuit {
    size_t address_size, data_size;
    std::vector<uint64_t> memory;
};
inline OpPtr QRAMLoad(std::shared_ptr<QRAMCircuit> q, size_t addr, size_t data) {
    if (!q || addr == data) throw_invalid_input();
    return pw("QRAMLoad", [=](System& s) {
        uint64_t a = s.as_uint(addr) % pow2(q->address_size);
        s.get(data) ^= q->memory[a];
    });
}

// ---- diagonal phases (basic_gates.cpp:11-55, 402-420, 486-520;
//      parallel_phase_operations.cpp:11-108)
inline OpPtr Phase_Bool(size_t id, size_t digit, double lambda) {
    if (digit >= Registry::size_of(id)) throw_invalid_input();
    complex_t ph(std::cos(lambda), std::sin(lambda)), phd(std::cos(lambda), -std::sin(lambda));
    return pw("Phase_Bool",
        [=](System& s) { if (digit1(s.get(id), digit)) s.amplitude *= ph; },
        [=](System& s) { if (digit1(s.get(id), digit)) s.amplitude *= phd; });
}
inline OpPtr Y_Bool(size_t id, size_t digit) {
    if (digit >= Registry::size_of(id)) throw_invalid_input();
    return pw("Y_Bool", [=](System& s) {
        s.get(id) = flip_digit(s.get(id), digit);
        if (digit1(s.get(id), digit)) s.amplitude *= complex_t(0, 1); else s.amplitude *= complex_t(0, -1);
    });
}
inline OpPtr RZ_Bool(size_t id, size_t digit, double angle) {
    if (digit >= Registry::size_of(id)) throw_invalid_input();
    auto mk = [=](double a) {
        complex_t pp(std::cos(a / 2), std::sin(a / 2)), pn(std::cos(a / 2), -std::sin(a / 2));
        return std::function<void(System&)>([=](System& s) {
            if (digit1(s.get(id), digit)) s.amplitude *= pp; else s.amplitude *= pn;
        });
    };
    return pw("RZ_Bool", mk(angle), mk(-angle));
}
inline OpPtr ZeroConditionalPhaseFlip(std::vector<size_t> ids) {
    return pw("ZeroConditionalPhaseFlip", [=](System& s) {
        for (size_t id : ids) if (s.get(id) != 0) return;
        s.amplitude *= -1;
    });
}
inline OpPtr Reflection_Bool(std::vector<size_t> regs, bool inverse) {
    return pw("Reflection_Bool", [=](System& s) {
        bool z = true;
        for (size_t id : regs) if (s.as_uint(id) != 0) { z = false; break; }
        if (!(inverse ^ z)) s.amplitude *= -1.0;
    });
}
inline OpPtr GlobalPhase(complex_t c) {
    return pw("GlobalPhase", [=](System& s) { s.amplitude *= c; }, [=](System& s) { s.amplitude *= std::conj(c); });
}

// ---- PlusOneAndOverflow (SparQ_Algorithm/src/BlockEncoding/block_encoding_tridiagonal.cpp:20-84)
inline OpPtr PlusOneAndOverflow(size_t main, size_t overflow) {
    return pw("PlusOneAndOverflow",
        [=](System& s) {
            uint64_t& v = s.get(main);
            uint64_t& o = s.get(overflow);
            if (v == pow2(Registry::size_of(main)) - 1) { o = (o + 1) % 2; v = 0; }
            else v += 1;
        },
        [=](System& s) {
            uint64_t& v = s.get(main);
            uint64_t& o = s.get(overflow);
            if (v == 0) { o = (o + 1) % 2; v = pow2(Registry::size_of(main)) - 1; }
            else v -= 1;
        });
}

// ---- Rot_GeneralUnitary (SparQ/src/rot.cpp:340-467): every bucket (entries equal on
// all registers but `id`, conditions per entry) is completed with zero-amplitude
// clones and multiplied by the dense 2^n x 2^n matrix (row i: sum over j ascending).
struct Rot_GeneralUnitary {
    size_t id, full_size;
    std::vector<complex_t> mat;   // row-major full_size x full_size
    Conditions cond;
    Rot_GeneralUnitary(size_t id_, std::vector<complex_t> m) : id(id_), mat(std::move(m)) {
        full_size = pow2(Registry::size_of(id));
        if (mat.size() != full_size * full_size) throw_invalid_input("Matrix size does not match the register's size.");
    }
    void bucket(const std::vector<size_t>& pos, State& st, bool dagger) const {
        std::vector<complex_t> amps(pos.size()), out(pos.size(), 0);
        for (size_t i = 0; i < pos.size(); ++i) amps[i] = st[pos[i]].amplitude;
        for (size_t i = 0; i < full_size; ++i)
            for (size_t j = 0; j < full_size; ++j)
                out[i] += (dagger ? std::conj(mat[j * full_size + i]) : mat[i * full_size + j]) * amps[j];
        for (size_t i = 0; i < pos.size(); ++i) st[pos[i]].amplitude = out[i];
    }
    void apply(State& st, bool dagger) const {
        if (st.empty()) return;
        size_t cur = st.size();
        std::map<System, std::pair<size_t, std::vector<size_t>>, StateLessExceptKey> buckets(StateLessExceptKey{id});
        for (size_t i = 0; i < cur; ++i) {
            if (!cond.satisfied(st[i])) continue;
            const System& s = st[i];
            auto res = buckets.insert({s, {0, std::vector<size_t>(full_size, size_t(-1))}});
            auto it = res.first;
            it->second.second[s.as_uint(id)] = i;
            it->second.first += 1;
            if (it->second.first == full_size) { bucket(it->second.second, st, dagger); buckets.erase(it); }
        }
        for (auto& kv : buckets) {
            auto& pos = kv.second.second;
            for (size_t i = 0; i < pos.size(); ++i)
                if (pos[i] == size_t(-1)) {
                    st.push_back(kv.first);
                    st.back().get(id) = i;
                    st.back().amplitude = 0;
                    pos[i] = st.size() - 1;
                }
            bucket(pos, st, dagger);
        }
        ClearZero(st);
        Registry::update_max_size(st.size());
    }
    void operator()(State& st) const { apply(st, false); }
    void dag(State& st) const { apply(st, true); }
};

// stateprep_unitary_build_schmidt (SparQ/src/rot.cpp:469-503): Householder
// reflection taking |0> to the normalised target, with the first column rephased.
inline std::vector<complex_t> stateprep_unitary(const std::vector<complex_t>& target) {
    const size_t N = target.size();
    if (N == 0) throw_invalid_input("Cannot prepare an empty vector.");
    double scale = 0;
    for (auto& v : target) {
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag())) throw_invalid_input("State amplitudes must be finite.");
        scale = std::max(scale, std::max(std::abs(v.real()), std::abs(v.imag())));
    }
    if (scale == 0) throw_invalid_input("Cannot prepare the zero vector.");
    std::vector<complex_t> v(N);
    double n2 = 0;
    for (size_t i = 0; i < N; ++i) { v[i] = target[i] / scale; n2 += abs_sqr(v[i]); }
    const double norm = std::sqrt(n2);
    for (auto& x : v) x /= norm;
    const double a0 = std::abs(v[0]);
    const complex_t phase = a0 == 0 ? complex_t(1) : v[0] / a0;
    v[0] += phase;
    double vn2 = 0;
    for (auto& x : v) vn2 += abs_sqr(x);
    std::vector<complex_t> U(N * N);
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) {
            U[i * N + j] = complex_t(i == j ? 1 : 0) - (2.0 / vn2) * v[i] * std::conj(v[j]);
            if (j == 0) U[i * N + j] *= -phase;
        }
    return U;
}

// ---- wrappers making the structural/branching operators uniform
template <class T, class = void> struct has_cond : std::false_type {};
template <class T> struct has_cond<T, std::void_t<decltype(std::declval<T&>().cond)>> : std::true_type {};
template <class T, class = void> struct has_dag : std::false_type {};
template <class T> struct has_dag<T, std::void_t<decltype(std::declval<const T&>().dag(std::declval<State&>()))>> : std::true_type {};

template <class T> struct Wrapped : Op {
    T op;
    Wrapped(T o, const char* n) : op(std::move(o)) { name = n; }
    T configured() const {
        T c = op;
        if constexpr (has_cond<T>::value) c.cond = cond;
        return c;
    }
    void apply(State& st) const override { configured()(st); }
    void apply_dag(State& st) const override {
        if constexpr (has_dag<T>::value) configured().dag(st);
        else throw std::runtime_error("Dagger is not implemented.");   // basic_components.h:756-759
    }
};
template <class T> inline OpPtr wrap(T o, const char* n) { return std::make_shared<Wrapped<T>>(std::move(o), n); }

}  // namespace oracle
