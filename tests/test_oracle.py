"""CPU checks that pin the oracle (oracle/sparq_oracle*.hpp) before it is trusted.

Each test names what fixes the expected value: a known answer from the
reference's own tests, a closed form, an invariant, or brute force with a dense
numpy model on small registers.  None of them re-types the oracle's formula.
"""
import itertools

import numpy as np
import pytest

from parity import check_condition_truth_tables


def reg_state(o, spec):
    st = o.SparseState()
    ids = [o.AddRegister(n, getattr(o, t), w)(st) for n, t, w in spec]
    return st, ids


def amp_map(st):
    ids, vals, amps = st.export()
    return {tuple(int(x) for x in v): a for v, a in zip(vals, amps)}


# ---- known answers of SparQ/test/hadamard.cpp:14-111
def test_hadamard_known_answers(oracle):
    st, _ = reg_state(oracle, [("q", "Boolean", 1)])
    oracle.Hadamard_Int_Full("q")(st)
    m = amp_map(st)
    assert abs(m[(0,)] - 2 ** -0.5) < 1e-9 and abs(m[(1,)] - 2 ** -0.5) < 1e-9
    oracle.System.clear()
    st, _ = reg_state(oracle, [("q", "Boolean", 1)])
    oracle.Init_Unsafe("q", 1)(st)
    oracle.Hadamard_Int_Full("q")(st)
    m = amp_map(st)
    assert abs(m[(0,)] - 2 ** -0.5) < 1e-9 and abs(m[(1,)] + 2 ** -0.5) < 1e-9
    oracle.Hadamard_Int_Full("q")(st)
    m = amp_map(st)
    assert list(m) == [(1,)] and abs(m[(1,)] - 1) < 1e-9     # H^2 = I
    oracle.System.clear()
    st, _ = reg_state(oracle, [("q", "UnsignedInteger", 2)])
    oracle.Hadamard_Int_Full("q")(st)
    assert st.size() == 4 and all(abs(abs(a) - 0.5) < 1e-9 for a in amp_map(st).values())


def dense_wht(n):
    h = np.array([[1.0]])
    for _ in range(n):
        h = np.kron(h, np.array([[1, 1], [1, -1]]) / np.sqrt(2))
    return h


@pytest.mark.parametrize("variant", ["Int", "Full"])
@pytest.mark.parametrize("k", [1, 2, 5, 16])
def test_hadamard_closed_form(oracle, variant, k):
    """H^{(x)n} on a random k-sparse vector equals the dense Walsh-Hadamard matrix product
    (H|v> = 2^{-n/2} sum_j (-1)^{popcount(j&v)} |j>, hadamard.cpp:11-17)."""
    n = 4
    rng = np.random.default_rng(k)
    st, _ = reg_state(oracle, [("x", "UnsignedInteger", n), ("y", "UnsignedInteger", 2)])
    xs = rng.choice(1 << n, size=k, replace=False)
    vec = np.zeros(1 << n, complex)
    vec[xs] = rng.standard_normal(k) + 1j * rng.standard_normal(k)
    st.load(np.stack([xs, np.full(k, 2)], 1).astype(np.uint64), vec[xs])
    (oracle.Hadamard_Int("x", n) if variant == "Int" else oracle.Hadamard_Int_Full("x"))(st)
    out = dense_wht(n) @ vec
    m = amp_map(st)
    for j in range(1 << n):
        got = m.get((j, 2), 0)
        assert abs(got - out[j]) < 1e-12
    assert all(key[1] == 2 for key in m)


def test_hadamard_partial_digits(oracle):
    """n_digits < width acts on the low bits only; upper bits keep their value."""
    st, _ = reg_state(oracle, [("x", "UnsignedInteger", 6)])
    oracle.Init_Unsafe("x", 0b101100)(st)
    oracle.Hadamard_Int("x", 2)(st)
    m = amp_map(st)
    assert sorted(m) == [(0b101100 + j,) for j in range(4)]
    assert all(abs(a - 0.5) < 1e-15 for a in m.values())


# ---- reference arithmetic expectations (SparQ/test/quantum_arithmetic.cpp:155-1181)
def run_classical(o, spec, init, op_factory, dagger=False):
    o.System.clear()
    st, ids = reg_state(o, spec)
    for name, v in init.items():
        o.Init_Unsafe(name, v)(st)
    op = op_factory()
    (op.dag if dagger else op)(st)
    ids_, vals, amps = st.export()
    assert len(amps) == 1 and amps[0] == 1
    return dict(zip([n for n, _, _ in spec], (int(x) for x in vals[0])))


def test_add_inplace_truth_table(oracle):
    # (lhs + rhs) & 0xf, quantum_arithmetic.cpp test :175-191
    for a, b in itertools.product(range(16), repeat=2):
        r = run_classical(oracle, [("l", "UnsignedInteger", 4), ("r", "UnsignedInteger", 4)], {"l": a, "r": b},
                          lambda: oracle.Add_UInt_UInt_InPlace("l", "r"))
        assert r["r"] == (a + b) & 0xF and r["l"] == a
        r = run_classical(oracle, [("l", "UnsignedInteger", 4), ("r", "UnsignedInteger", 4)], {"l": a, "r": b},
                          lambda: oracle.Add_UInt_UInt_InPlace("l", "r"), dagger=True)
        assert r["r"] == (b - a) & 0xF


def test_xor_into_truth_tables(oracle):
    spec = [("a", "UnsignedInteger", 3), ("b", "UnsignedInteger", 3), ("c", "UnsignedInteger", 4)]
    cases = {
        "Add_UInt_UInt": lambda a, b: (a + b) & 0xF, "Sub_UInt_UInt": lambda a, b: (a - b) & 0xF,
        "Mul_UInt_UInt": lambda a, b: (a * b) & 0xF, "Div_UInt_UInt": lambda a, b: 0 if b == 0 else a // b,
        "And_UInt_UInt": lambda a, b: a & b, "Or_UInt_UInt": lambda a, b: a | b, "Xor_UInt_UInt": lambda a, b: a ^ b,
        "GetMid_UInt_UInt": lambda a, b: (a + b) // 2,
    }
    for name, f in cases.items():
        for a, b in itertools.product(range(8), repeat=2):
            for c0 in (0, 9):
                r = run_classical(oracle, spec, {"a": a, "b": b, "c": c0}, lambda: getattr(oracle, name)("a", "b", "c"))
                assert r["c"] == c0 ^ f(a, b), (name, a, b, c0)


def test_compare_and_flags(oracle):
    spec = [("a", "UnsignedInteger", 3), ("b", "UnsignedInteger", 3), ("lt", "Boolean", 1), ("eq", "Boolean", 1)]
    for a, b in itertools.product(range(8), repeat=2):
        r = run_classical(oracle, spec, {"a": a, "b": b}, lambda: oracle.Compare_UInt_UInt("a", "b", "lt", "eq"))
        assert (r["lt"], r["eq"]) == (int(a < b), int(a == b))
    spec = [("s", "SignedInteger", 4), ("t", "SignedInteger", 4), ("f", "Boolean", 1)]
    sx = lambda v: v - 16 if v >= 8 else v  # noqa: E731
    for a, b in itertools.product(range(16), repeat=2):
        r = run_classical(oracle, spec, {"s": a, "t": b}, lambda: oracle.Less_SInt_SInt("s", "t", "f"))
        assert r["f"] == int(sx(a) < sx(b))


def test_shift_rotations_roundtrip(oracle):
    for v in range(32):
        r = run_classical(oracle, [("x", "UnsignedInteger", 5)], {"x": v}, lambda: oracle.ShiftLeft_InPlace("x", 2))
        assert r["x"] == ((v << 2) | (v >> 3)) & 31          # rotate left by 2 in 5 bits
        r = run_classical(oracle, [("x", "UnsignedInteger", 5)], {"x": v}, lambda: oracle.ShiftRight_InPlace("x", 2))
        assert r["x"] == ((v >> 2) | (v << 3)) & 31


def test_modmul_inverse(oracle):
    for v in range(21):
        r = run_classical(oracle, [("x", "UnsignedInteger", 5)], {"x": v}, lambda: oracle.Mod_Mult_UInt_ConstUInt_InPlace("x", 2, 1, 21))
        assert r["x"] == (v * 4) % 21      # a^(2^x) with a=2, x=1
        r2 = run_classical(oracle, [("x", "UnsignedInteger", 5)], {"x": r["x"]}, lambda: oracle.Mod_Mult_UInt_ConstUInt_InPlace("x", 2, 1, 21), dagger=True)
        assert r2["x"] == v


def test_qram_xor_model(oracle):
    """data ^= memory[address] (docs/verification/v1b-runtime.md:14)."""
    mem = [(7 * i + 3) % 64 for i in range(16)]
    q = oracle.QRAMCircuit_qutrit(4, 6, mem)
    for a in range(16):
        for d0 in (0, 21):
            r = run_classical(oracle, [("a", "UnsignedInteger", 4), ("d", "UnsignedInteger", 6)], {"a": a, "d": d0},
                              lambda: oracle.QRAMLoad(q, "a", "d"))
            assert r["d"] == d0 ^ mem[a]


# ---- rotations: CudaBenchmark analytic answer (test/GPUTest/Benchmark/CudaBenchmark.cu:13-18,188-217)
def test_cudabenchmark_analytic(oracle):
    n = 6
    st, _ = reg_state(oracle, [("addr", "UnsignedInteger", n), ("sum", "UnsignedInteger", n + 1),
                               ("sq", "UnsignedInteger", n // 2 + 1), ("b", "UnsignedInteger", 1)])
    th = np.pi / 3
    rot = [np.cos(th), -np.sin(th), np.sin(th), np.cos(th)]
    phase = complex(np.cos(np.pi / 5), np.sin(np.pi / 5))
    oracle.Hadamard_Int("addr", n)(st)
    oracle.Add_UInt_UInt("addr", "addr", "sum")(st)
    oracle.Sqrt_UInt("addr", "sq")(st)
    oracle.X_Bool("b", 0)(st)
    oracle.Rot_Bool("b", 0, rot)(st)
    oracle.GlobalPhase(phase)(st)
    oracle.Normalize()(st)
    inv = 1 / np.sqrt(2 ** n)
    m = amp_map(st)
    assert len(m) == 2 * 2 ** n
    for k in range(2 ** n):
        r = int(np.floor(np.sqrt(k)))
        assert abs(m[(k, 2 * k, r, 1)] - 0.5 * phase * inv) < 1e-12
        assert abs(m[(k, 2 * k, r, 0)] + np.sqrt(3) / 2 * phase * inv) < 1e-12


def test_condrot_closed_form(oracle):
    """CondRot_Rational_Bool: |v>|0> -> cos(2 pi v/2^w)|v>|0> + sin(2 pi v/2^w)|v>|1> (condrot.cpp:27-33)."""
    w = 5
    st, _ = reg_state(oracle, [("r", "Rational", w), ("q", "Boolean", 1)])
    oracle.Hadamard_Int("r", w)(st)
    oracle.CondRot_Rational_Bool("r", "q")(st)
    m = amp_map(st)
    for v in range(2 ** w):
        th = 2 * np.pi * v / 2 ** w
        a = 2 ** (-w / 2)
        assert abs(m.get((v, 0), 0) - a * np.cos(th)) < 1e-12
        assert abs(m.get((v, 1), 0) - a * np.sin(th)) < 1e-12
    oracle.CondRot_Rational_Bool("r", "q").dag(st)
    m = amp_map(st)
    assert all(k[1] == 0 for k in m)            # dag undoes the rotation (|1> branch pruned)


def test_hadamard_bool_pairs(oracle):
    """Dense check of H on a 1-bit register over a random superposition."""
    st, _ = reg_state(oracle, [("x", "UnsignedInteger", 3), ("q", "Boolean", 1)])
    rng = np.random.default_rng(3)
    keys = [(x, q) for x in range(8) for q in range(2)]
    pick = rng.choice(16, size=9, replace=False)
    vals = np.array([keys[i] for i in pick], np.uint64)
    amps = rng.standard_normal(9) + 1j * rng.standard_normal(9)
    st.load(vals, amps)
    dense = np.zeros(16, complex)
    for (x, q), a in zip(vals, amps):
        dense[int(x) * 2 + int(q)] = a
    h = np.kron(np.eye(8), dense_wht(1))
    out = h @ dense
    m = amp_map(st)
    oracle.Hadamard_Bool("q")(st)
    m = amp_map(st)
    for i in range(16):
        assert abs(m.get((i // 2, i % 2), 0) - out[i]) < 1e-12


def test_sort_merge_invariants(oracle):
    """add_systems: result sorted, unique keys, duplicates summed, |a|^2 < 1e-14 dropped."""
    spec = [("a", "UnsignedInteger", 6)]
    s1, _ = reg_state(oracle, spec)
    s1.load(np.array([[5], [1], [9], [3]], np.uint64), np.array([1, 2, 3, 1e-8]))
    s2 = oracle.SparseState()
    s2.load(np.array([[9], [4], [3]], np.uint64), np.array([-3, 5, -1e-8]))
    oracle.add_systems(s1, s2, 1.0)
    ids, vals, amps = s1.export()
    assert [int(v) for v in vals[:, 0]] == [1, 4, 5]
    assert list(amps) == [2, 5, 1]


def test_reflection_bool_closed_form(oracle):
    """Reflection_Bool(regs, inverse) negates when !(inverse XOR all-zero)
    (parallel_phase_operations.cpp:46-70): inverse=False flips the non-zero
    entries, inverse=True the all-zero ones; conditions gate the flip."""
    st, _ = reg_state(oracle, [("x", "UnsignedInteger", 2), ("c", "Boolean", 1)])
    vals = np.array([[x, c] for x in range(4) for c in range(2)], np.uint64)
    st.load(vals, np.ones(8))
    oracle.Reflection_Bool(["x"], False)(st)
    m = amp_map(st)
    assert all(m[(x, c)] == (1 if x == 0 else -1) for x in range(4) for c in range(2))
    oracle.Reflection_Bool(["x"], True).conditioned_by_all_ones("c")(st)
    m = amp_map(st)
    assert all(m[(x, c)] == (1 if x == 0 else -1) * (-1 if (x == 0 and c == 1) else 1) for x in range(4) for c in range(2))


def test_clear_zero_is_strict(oracle):
    st, _ = reg_state(oracle, [("a", "UnsignedInteger", 2)])
    st.load(np.array([[0], [1], [2]], np.uint64), np.array([0.5, 0.49999, 0.5j]))
    oracle.ClearZero(0.25)(st)                  # |0.5|^2 = 0.25 is not < 0.25: kept (strict <)
    assert sorted(int(v[0]) for v in st.export()[1]) == [0, 2]


def test_register_reuse_lifo(oracle):
    """Freed ids are reused last-in-first-out (basic_components.cpp:344-356)."""
    st = oracle.SparseState()
    ids = [oracle.AddRegister(f"r{i}", oracle.UnsignedInteger, 3)(st) for i in range(4)]
    oracle.RemoveRegister("r1")(st)
    oracle.RemoveRegister("r3")(st)
    assert oracle.AddRegister("x", oracle.UnsignedInteger, 2)(st) == ids[3]
    assert oracle.AddRegister("y", oracle.UnsignedInteger, 2)(st) == ids[1]


def test_split_combine_roundtrip(oracle):
    for v in (0, 1, 0b1011011, 127):
        oracle.System.clear()
        st, _ = reg_state(oracle, [("a", "UnsignedInteger", 7)])
        oracle.Init_Unsafe("a", v)(st)
        oracle.SplitRegister("a", "lo", 3)(st)
        m = amp_map(st)
        assert list(m) == [(v >> 3, v & 7)]
        oracle.CombineRegister("a", "lo")(st)
        assert list(amp_map(st)) == [(v,)]


# ---- flags, select, abs: exhaustive truth tables against Python integer arithmetic
# (definitions of SparQ/src/quantum_arithmetic.cpp:532-554, 623-644, 731-816)
def _sx(v, w):
    return v - (1 << w) if v >= 1 << (w - 1) else v


def test_flag_truth_tables(oracle):
    for w in (2, 3, 4):
        spec = [("a", "UnsignedInteger", 4), ("b", "UnsignedInteger", 4), ("r", "UnsignedInteger", w), ("f", "Boolean", 1)]
        for a, b in itertools.product(range(16), repeat=2):
            for f0 in (0, 1):
                got = run_classical(oracle, spec, {"a": a, "b": b, "f": f0}, lambda: oracle.Carry_UInt_UInt("a", "b", "r", "f"))
                assert got["f"] == f0 ^ int(a + b >= 1 << w), ("carry", w, a, b)   # carry out of a w-bit sum
                got = run_classical(oracle, spec, {"a": a, "b": b, "f": f0},
                                    lambda: oracle.MulOverflow_UInt_UInt("a", "b", "r", "f"))
                assert got["f"] == f0 ^ int(a * b >= 1 << w), ("muloverflow", w, a, b)
        spec = [("a", "SignedInteger", 4), ("b", "SignedInteger", 4), ("r", "UnsignedInteger", w), ("f", "Boolean", 1)]
        for a, b in itertools.product(range(16), repeat=2):
            # both operands reduced to w-bit two's complement; overflow = the true sum leaves the w-bit range
            sa, sb = _sx(_sx(a, 4) % (1 << w), w), _sx(_sx(b, 4) % (1 << w), w)
            expect = int(not (-(1 << (w - 1)) <= sa + sb < 1 << (w - 1)))
            got = run_classical(oracle, spec, {"a": a, "b": b}, lambda: oracle.Overflow_SInt_SInt("a", "b", "r", "f"))
            assert got["f"] == expect, ("overflow", w, a, b)


def test_select_and_abs_truth_tables(oracle):
    spec = [("c", "UnsignedInteger", 2), ("x", "UnsignedInteger", 3), ("y", "UnsignedInteger", 3), ("r", "UnsignedInteger", 2)]
    for c, x, y, r0 in itertools.product(range(4), range(8), range(8), (0, 2)):
        got = run_classical(oracle, spec, {"c": c, "x": x, "y": y, "r": r0},
                            lambda: oracle.Select_Bool_UInt_UInt("c", "x", "y", "r"))
        assert got["r"] == (r0 ^ (x if c & 1 else y)) & 3, (c, x, y, r0)   # bit 0 selects, result truncated
    spec = [("v", "SignedInteger", 4), ("r", "UnsignedInteger", 3)]
    for v, r0 in itertools.product(range(16), (0, 5)):
        got = run_classical(oracle, spec, {"v": v, "r": r0}, lambda: oracle.Abs_SInt("v", "r"))
        assert got["r"] == (r0 ^ abs(_sx(v, 4))) & 7, (v, r0)


# ---- rational-output ops: the angle formula to the resolution of the output register
# (quantum_arithmetic.cpp:413-456, 458-490; get_rational is pinned only to 2^-w)
def test_rational_angle_ops(oracle):
    import math

    w = 12

    def decode(bits):
        return bits / float(1 << w)

    def close(d, x):
        e = abs(d - x)
        return min(e, 1 - e) <= 2.0 ** -w

    spec = [("a", "UnsignedInteger", 5), ("b", "UnsignedInteger", 5), ("o", "Rational", w)]
    for b in range(1, 32):
        for a in range(0, b + 1):
            got = run_classical(oracle, spec, {"a": a, "b": b}, lambda: oracle.Div_Sqrt_Arccos_UInt_UInt("a", "b", "o"))
            assert close(decode(got["o"]), math.acos(math.sqrt(a / b)) / (2 * math.pi)), (a, b)
    spec = [("l", "SignedInteger", 5), ("r", "UnsignedInteger", 6), ("o", "Rational", w)]
    for r in range(1, 64):
        for l in range(0, int(math.isqrt(r)) + 1):
            got = run_classical(oracle, spec, {"l": l, "r": r}, lambda: oracle.Sqrt_Div_Arccos_Int_UInt("l", "r", "o"))
            assert close(decode(got["o"]), math.acos(l / math.sqrt(r)) / (2 * math.pi)), (l, r)
    spec = [("l", "SignedInteger", 4), ("r", "SignedInteger", 4), ("o", "Rational", w)]
    for l, r in itertools.product(range(16), repeat=2):
        sl, sr = _sx(l, 4), _sx(r, 4)
        x = 0.25 if sl == 0 and sr == 0 else (math.atan2(sr, sl) / (2 * math.pi)) % 1.0
        got = run_classical(oracle, spec, {"l": l, "r": r}, lambda: oracle.GetRotateAngle_Int_Int("l", "r", "o"))
        assert close(decode(got["o"]), x), (sl, sr)


# ---- one-bit gates against a dense numpy model on random superpositions
# (basic_gates.cpp:11-55 Phase, 418-441 Y, 623-653 RZ; parallel_phase_operations.cpp:11-31)
def _dense_one_bit(amps, a_vals, digit, U):
    out = {}
    for (a, b), amp in zip(a_vals, amps):
        bit = (a >> digit) & 1
        for nb in (0, 1):
            na = (a & ~(1 << digit)) | (nb << digit)
            if U[nb][bit] != 0:
                out[(na, b)] = out.get((na, b), 0) + U[nb][bit] * amp
    return out


def test_one_bit_gates_dense(oracle):
    rng = np.random.default_rng(11)
    lam, theta = 0.7, 1.9
    gates = {
        "Phase_Bool": (lambda: oracle.Phase_Bool("a", 1, lam), [[1, 0], [0, np.exp(1j * lam)]]),
        "RZ_Bool": (lambda: oracle.RZ_Bool("a", 1, theta), [[np.exp(-0.5j * theta), 0], [0, np.exp(0.5j * theta)]]),
        "Y_Bool": (lambda: oracle.Y_Bool("a", 1), [[0, -1j], [1j, 0]]),
    }
    keys = [(a, b) for a in range(8) for b in range(4) if rng.random() < 0.6]
    amps = rng.standard_normal(len(keys)) + 1j * rng.standard_normal(len(keys))
    for name, (make, U) in gates.items():
        oracle.System.clear()
        st, _ = reg_state(oracle, [("a", "UnsignedInteger", 3), ("b", "UnsignedInteger", 2)])
        st.load(np.array(keys, dtype=np.uint64), amps)
        make()(st)
        got = amp_map(st)
        want = _dense_one_bit(amps, keys, 1, U)
        assert set(got) == {k for k, v in want.items() if v != 0}, name
        for k, v in want.items():
            assert abs(got[k] - v) <= 1e-14 * abs(v), (name, k)
    oracle.System.clear()
    st, _ = reg_state(oracle, [("a", "UnsignedInteger", 3), ("b", "UnsignedInteger", 2)])
    st.load(np.array(keys, dtype=np.uint64), amps)
    oracle.ZeroConditionalPhaseFlip(["a", "b"])(st)
    got = amp_map(st)
    for (a, b), amp in zip(keys, amps):
        assert got[(a, b)] == (-amp if a == 0 and b == 0 else amp)


# ---- one-line key rewrites (quantum_arithmetic.cpp:511-530, 819-857, 859-888, 909-930)
def test_simple_key_rewrite_truth_tables(oracle):
    spec = [("a", "UnsignedInteger", 4), ("r", "UnsignedInteger", 3), ("f", "Boolean", 1)]
    for a, r0, f0 in itertools.product(range(16), (0, 6), (0, 1)):
        got = run_classical(oracle, spec, {"a": a, "r": r0}, lambda: oracle.Neg_UInt("a", "r"))
        assert got["r"] == (r0 ^ (-a & 7)), ("neg", a, r0)                  # two's-complement negation, truncated
        got = run_classical(oracle, spec, {"a": a, "r": r0}, lambda: oracle.Assign("a", "r"))
        assert got["r"] == (r0 ^ a) & 7 and got["a"] == a, ("assign", a, r0)  # XOR-copy, truncated
        got = run_classical(oracle, spec, {"a": a, "f": f0}, lambda: oracle.IsZero_UInt("a", "f"))
        assert got["f"] == f0 ^ int(a == 0), ("iszero", a)
    spec = [("s", "SignedInteger", 4), ("f", "Boolean", 1)]
    for s, f0 in itertools.product(range(16), (0, 1)):
        got = run_classical(oracle, spec, {"s": s, "f": f0}, lambda: oracle.Negative_SInt("s", "f"))
        assert got["f"] == f0 ^ int(_sx(s, 4) < 0), ("negative", s)
    for rtype in ("UnsignedInteger", "SignedInteger"):
        spec = [("l", "UnsignedInteger", 4), ("x", rtype, 3)]
        for l, x in itertools.product(range(16), range(8)):
            ext = _sx(x, 3) if rtype == "SignedInteger" else x                 # rhs read by its declared type
            got = run_classical(oracle, spec, {"l": l, "x": x}, lambda: oracle.Add_AnyInt_AnyInt_InPlace("l", "x"))
            assert got["l"] == (l + ext) % 16 and got["x"] == x, ("add_any", rtype, l, x)
            got = run_classical(oracle, spec, {"l": l, "x": x}, lambda: oracle.Add_AnyInt_AnyInt_InPlace("l", "x"), dagger=True)
            assert got["l"] == (l - ext) % 16, ("add_any dag", rtype, l, x)



# ---- condition evaluation (global_macros.h:13-17, 82-86): brute force over every
# value of three small registers against Python predicates (tests/parity.py).
def test_condition_truth_tables(oracle):
    check_condition_truth_tables(oracle)


# ---- fixed-point encodings: the reference's own bounds
# get_rational (test/CPUTest/CommonTest/CorrectnessTest_Common.inl:171-195):
# |decode(encode(x)) - x| <= 2^-w for w = 3..32 at x = 0.1231231.  Checked here
# at every width 3..64 on that value, 400 seeded uniform values and the exact
# dyadics k / 2^w (which must encode to k), with exact rational arithmetic.  The
# rule is truncation, floor(x 2^w): within the reference's bound, the rounding
# mode itself is this build's (stated in the oracle's header).
def test_get_rational_reference_bound_every_width(oracle):
    from fractions import Fraction

    rng = np.random.default_rng(0x5EED2026)
    xs = [0.1231231, 0.0, 0.5, 0.999999999] + [float(x) for x in rng.uniform(0.0, 1.0, 400)]
    for w in range(3, 65):
        one = 1 << w
        for x in xs:
            v = oracle.get_rational(x, w)
            assert 0 <= v < one
            err = abs(Fraction(v, one) - Fraction(x))
            assert err <= Fraction(1, one), (w, x, v)
            assert v == (Fraction(x) * one).__floor__(), (w, x, v)
        # exact dyadics: every k/2^w that a double holds exactly
        ks = [0, 1, 2, 3, one // 2, one - 1] + [int(k) for k in rng.integers(0, min(one, 2 ** 53), 50)]
        for k in ks:
            x = k / one
            if Fraction(x) != Fraction(k, one):
                continue
            assert oracle.get_rational(x, w) == k, (w, k)


# SparseMatrix quantisation (the reference's Python front end clips every value
# to [0, (2^ds - 1)/2^ds] (positive_only) or +-(2^(ds-1) - 1)/2^(ds-1) and hands
# it to the C++ SparseMatrix, PySparQ/pysparq/algorithms/cks_solver.py:325-342;
# the encoder itself lives in the absent submodule).  What those bounds fix:
# the clip ends encode to the largest code (2^ds - 1, and 2^(ds-1) - 1 / its
# two's complement negative), exact dyadics encode exactly, every value
# decodes to within one code step of its input with its sign kept, and the
# oracle and the engine's host layer encode identically.
@pytest.mark.parametrize("positive_only", [True, False])
def test_sparse_matrix_quantisation_against_float_input(oracle, positive_only):
    from fractions import Fraction

    import sparqsim_amd as ps

    rng = np.random.default_rng(11)
    for ds in (4, 8, 16, 31, 32, 40):
        step = Fraction(1, 1 << ds) if positive_only else Fraction(1, 1 << (ds - 1))
        top = 1 - step                                   # max_encoded of the reference front end
        lo = Fraction(0) if positive_only else -top
        xs = [float(top), float(lo), 0.0, 0.5, float(step), float(lo + step)]
        xs += [float(x) for x in rng.uniform(float(lo), float(top), 200)]
        dy = [int(k) for k in rng.integers(0, (1 << ds) if positive_only else (1 << (ds - 1)), 40)]
        xs += [float(k * step) for k in dy] + ([] if positive_only else [float(-k * step) for k in dy])
        xs = [x for x in xs if lo <= Fraction(x) <= top]
        n = len(xs)
        enc_o = list(oracle.SparseMatrix(xs, list(range(n)), ds, n, 1, positive_only).elements)
        enc_e = list(ps.SparseMatrix(xs, list(range(n)), ds, n, 1, positive_only).elements)
        assert enc_o == enc_e, ds
        mask = (1 << ds) - 1
        for x, v in zip(xs, enc_o):
            assert 0 <= v <= mask
            s = v if positive_only or v < (1 << (ds - 1)) else v - (1 << ds)
            dec = s * step
            assert abs(dec - Fraction(x)) < step, (ds, x, v)
            if Fraction(x) % step == 0:
                assert dec == Fraction(x), (ds, x, v)           # exact dyadics
            assert (s < 0) <= (x < 0) and (s > 0) <= (x > 0)    # sign kept (or zero)
        assert enc_o[0] == (mask if positive_only else (1 << (ds - 1)) - 1)
        if not positive_only:
            assert enc_o[1] == (1 << (ds - 1)) + 1             # -(2^(ds-1) - 1) in two's complement
