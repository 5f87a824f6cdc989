"""GPU parity: the HIP engine against the CPU oracle on the same seeded inputs.

Keys are compared bit-exact after canonical sort; amplitudes within
parity.RTOL (1e-12 relative to max(|a|, max|a| of the state)); per-entry ops
that only move keys or multiply by host-computed constants are checked
bit-identical.
"""
import numpy as np
import pytest

from parity import assert_same_state, check_condition_truth_tables, random_state, run_both

pytestmark = pytest.mark.gpu

U = None  # set per module (UnsignedInteger etc. are identical names in both modules)


def regs(mod, spec):
    st = mod.SparseState()
    for name, typ, w in spec:
        mod.AddRegister(name, getattr(mod, typ), w)(st)
    return st


def loaded(mod, spec, n, seed):
    st = regs(mod, spec)
    vals, amps = random_state(mod, n, seed)
    st.load(vals, amps)
    return st


# ---------------------------------------------------------------- configs (reduced sizes)
@pytest.mark.parametrize("w", [4, 8])
def test_config1_plain_circuit(ps, oracle, w):
    """H(a); b += a; d ^= M[a]; H(a); ClearZero -- the plain CPU-runnable circuit at reduced width."""
    rng = np.random.default_rng(0x5EED2026 + 0)
    mem = [int(x) for x in rng.integers(0, 1 << w, size=1 << w)]

    def build(mod):
        st = regs(mod, [("a", "UnsignedInteger", w), ("b", "UnsignedInteger", w), ("d", "UnsignedInteger", w)])
        q = mod.QRAMCircuit_qutrit(w, w, mem)
        mod.Hadamard_Int("a", w)(st)
        mod.Add_UInt_UInt_InPlace("a", "b")(st)
        mod.QRAMLoad(q, "a", "d")(st)
        mod.Hadamard_Int("a", w)(st)
        mod.ClearZero(1e-14)(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert eg[1].shape[0] == (1 << (2 * w))
    assert_same_state(eg, eo, exact_amps=True, what="config1")


def test_config1_16bit_first_steps(ps, oracle):
    """Configuration 1 at its own 16-bit width, steps 1-3 (H(a); b += a;
    d ^= M[a]: 65 536 entries) against the oracle, amplitudes bit-exact.  The
    literal final H(a) makes 2^32 entries; it is checked by its closed form at
    12 bits below."""
    w = 16
    rng = np.random.default_rng(0x5EED2026 + 0)
    mem = [int(x) for x in rng.integers(0, 1 << w, size=1 << w)]

    def build(mod):
        st = regs(mod, [("a", "UnsignedInteger", w), ("b", "UnsignedInteger", w), ("d", "UnsignedInteger", w)])
        q = mod.QRAMCircuit_qutrit(w, w, mem)
        mod.Hadamard_Int("a", w)(st)
        mod.Add_UInt_UInt_InPlace("a", "b")(st)
        mod.QRAMLoad(q, "a", "d")(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert eg[1].shape[0] == 1 << w
    assert_same_state(eg, eo, exact_amps=True, what="config1 steps 1-3 at 16 bits")


def test_config1_12bit_full_circuit_closed_form(ps):
    """The whole configuration-1 circuit at 12 bits (2^24 entries after the final
    H(a)) against its closed form: every (a', b) pair once, d = M[b], amplitude
    2^-12 (-1)^popcount(a' & b) (b holds the first superposition's a)."""
    w = 12
    rng = np.random.default_rng(0x5EED2026 + 0)
    mem = rng.integers(0, 1 << w, size=1 << w).astype(np.uint64)
    ps.System.clear()
    st = regs(ps, [("a", "UnsignedInteger", w), ("b", "UnsignedInteger", w), ("d", "UnsignedInteger", w)])
    q = ps.QRAMCircuit_qutrit(w, w, [int(x) for x in mem])
    ps.Hadamard_Int("a", w)(st)
    ps.Add_UInt_UInt_InPlace("a", "b")(st)
    ps.QRAMLoad(q, "a", "d")(st)
    ps.Hadamard_Int("a", w)(st)
    ps.ClearZero(1e-14)(st)
    _, vals, amps = st.export()
    assert vals.shape[0] == 1 << (2 * w)
    a, b, d = vals[:, 0], vals[:, 1], vals[:, 2]
    pair = np.sort(b * (1 << w) + a)
    assert np.array_equal(pair, np.arange(1 << (2 * w), dtype=np.uint64))
    assert np.array_equal(d, mem[b])
    sign = 1.0 - 2.0 * (np.bitwise_count(a & b) & 1)
    assert np.max(np.abs(amps - sign * 2.0 ** -w)) <= 1e-12 * 2.0 ** -w


@pytest.mark.parametrize("A", [6, 12, 16])
def test_config2_qram_interference(ps, oracle, A):
    """H(addr); data ^= M[addr]; 32 Z on data bits; uncompute; H(addr)."""
    D = 32
    rng = np.random.default_rng(0x5EED2026 + 1)
    mem = [int(x) for x in rng.integers(0, 1 << D, size=1 << A, dtype=np.uint64)]

    def build(mod):
        st = regs(mod, [("addr", "UnsignedInteger", A), ("data", "UnsignedInteger", D)])
        q = mod.QRAMCircuit_qutrit(A, D, mem)
        mod.Hadamard_Int("addr", A)(st)
        mod.QRAMLoad(q, "addr", "data")(st)
        for i in range(D):
            mod.Phase_Bool("data", i, np.pi)(st)
        mod.QRAMLoad(q, "addr", "data")(st)
        mod.Hadamard_Int("addr", A)(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what="config2")


def _fwht(v):
    """Unnormalised Walsh-Hadamard transform, in place over the last axis (numpy)."""
    v = v.copy()
    h = 1
    n = v.shape[-1]
    while h < n:
        v = v.reshape(-1, 2, h)
        a, b = v[:, 0, :].copy(), v[:, 1, :].copy()
        v[:, 0, :], v[:, 1, :] = a + b, a - b
        v = v.reshape(n)
        h *= 2
    return v


def test_config2_full_size_closed_form(ps):
    """The benchmarked QRAM-interference pass at 2^20 entries (20-bit addr, 32-bit
    data, the benchmark's seeded memory): H U H with U = (-1)^parity(M[addr]) on
    |0, 0> leaves amplitude WHT(s)_y / 2^20 on |y, 0>, s_x = (-1)^popcount(M[x]).
    Checked against a numpy transform: the kept entries (those clear of the
    ClearZero threshold by a factor 2 either way) and every amplitude to 1e-12
    of the largest (Phase_Bool(pi) multiplies by (-1, 1.2e-16) per set bit)."""
    from sparqsim_amd.workloads import qram_memory

    A, D = 20, 32
    mem = qram_memory(A, D, seed_offset=1)
    st = regs(ps, [("addr", "UnsignedInteger", A), ("data", "UnsignedInteger", D)])
    q = ps.QRAMCircuit_qutrit(A, D, mem)
    ps.Hadamard_Int("addr", A)(st)
    ps.QRAMLoad(q, "addr", "data")(st)
    for i in range(D):
        ps.Phase_Bool("data", i, np.pi)(st)
    ps.QRAMLoad(q, "addr", "data")(st)
    ps.Hadamard_Int("addr", A)(st)
    ids, vals, amps = st.export()
    assert np.all(vals[:, 1] == 0), "data register not uncomputed"
    words = np.array(mem, dtype=np.uint64)
    parity = np.zeros(words.shape, dtype=np.uint64)
    for b in range(D):
        parity ^= (words >> np.uint64(b)) & np.uint64(1)
    expect = _fwht(1.0 - 2.0 * parity.astype(np.float64)) / float(1 << A)
    eps = ps.get_clearzero_default_eps()
    p2 = expect ** 2
    got = dict(zip((int(y) for y in vals[:, 0]), amps))
    must = np.nonzero(p2 >= 2 * eps)[0]
    never = np.nonzero(p2 < eps / 2)[0]
    assert all(int(y) in got for y in must), "entries above the ClearZero threshold missing"
    assert not any(int(y) in got for y in never), "entries below the ClearZero threshold kept"
    scale = np.max(np.abs(expect))
    err = max(abs(got[y] - expect[y]) for y in got)
    assert err <= 1e-12 * scale, err
    assert abs(float(np.sum(np.abs(amps) ** 2)) - float(np.sum(p2[list(got)]))) <= 1e-12


# ---------------------------------------------------------------- per-entry ops
SPEC3 = [("a", "UnsignedInteger", 7), ("b", "UnsignedInteger", 9), ("c", "UnsignedInteger", 6),
         ("s", "SignedInteger", 8), ("f", "Boolean", 1), ("g", "Boolean", 1)]

ARITH = [
    ("Add_UInt_UInt", ("a", "b", "c")), ("Sub_UInt_UInt", ("a", "b", "c")), ("Mul_UInt_UInt", ("a", "b", "c")),
    ("Div_UInt_UInt", ("b", "a", "c")), ("And_UInt_UInt", ("a", "b", "c")), ("Or_UInt_UInt", ("a", "b", "c")),
    ("Xor_UInt_UInt", ("a", "b", "c")), ("Less_SInt_SInt", ("s", "a", "f")), ("Less_UInt_UInt", ("a", "b", "f")),
    ("GetMid_UInt_UInt", ("a", "c", "b")), ("Add_UInt_UInt_InPlace", ("a", "b")), ("Neg_UInt", ("a", "c")),
    ("Abs_SInt", ("s", "c")), ("Sqrt_UInt", ("b", "c")), ("IsZero_UInt", ("c", "f")), ("Negative_SInt", ("s", "g")),
    ("Add_AnyInt_AnyInt_InPlace", ("b", "s")), ("Assign", ("c", "a")),
    ("Select_Bool_UInt_UInt", ("f", "a", "c", "b")), ("Carry_UInt_UInt", ("a", "c", "a", "f")),
    ("Overflow_SInt_SInt", ("s", "a", "s", "g")), ("MulOverflow_UInt_UInt", ("a", "c", "a", "f")),
    ("Compare_UInt_UInt", ("a", "c", "f", "g")), ("Mult_UInt_ConstUInt", ("a", 5, "b")),
    ("Add_Mult_UInt_ConstUInt_InPlace", ("a", 3, "b")), ("Add_UInt_ConstUInt", ("a", 77, "b")),
    ("Add_ConstUInt_InPlace", ("b", 300)), ("ShiftLeft_InPlace", ("b", 3)), ("ShiftRight_InPlace", ("a", 2)),
    ("X_Bool", ("b", 4)), ("FlipBools", ("c",)), ("Mod_Mult_UInt_ConstUInt_InPlace", ("b", 7, 2, 509)),
    ("Swap_Bool_Bool", ("a", 1, "b", 7)), ("Init_Unsafe", ("c", 13)),
]


@pytest.mark.parametrize("name,args", ARITH, ids=[a[0] for a in ARITH])
def test_key_rewrite_ops(ps, oracle, name, args):
    def build(mod):
        st = loaded(mod, SPEC3, 3000, seed=1)
        op = getattr(mod, name)(*args)
        op(st)
        op(st)              # twice: exercises XOR-into involution / in-place composition
        if name not in ("Init_Unsafe", "Mod_Mult_UInt_ConstUInt_InPlace"):
            try:
                op.dag(st)
            except RuntimeError:
                pass
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, exact_amps=True, what=name)


def test_conditions_on_pointwise(ps, oracle):
    def build(mod):
        st = loaded(mod, SPEC3, 4000, seed=2)
        mod.Add_UInt_UInt("a", "b", "c").conditioned_by_nonzeros("f")(st)
        mod.X_Bool("a", 3).conditioned_by_bit("b", 2)(st)
        mod.Add_ConstUInt_InPlace("b", 11).conditioned_by_value("c", 5)(st)
        mod.FlipBools("c").conditioned_by_all_ones("g")(st)
        mod.Phase_Bool("a", 1, 0.37).conditioned_by_bit([("b", 0), ("c", 1)])(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, exact_amps=True, what="conditions")


DIAGONAL = [
    ("Phase_Bool", lambda m: m.Phase_Bool("a", 2, 0.913), False),
    ("Phase_Bool.dag", lambda m: m.Phase_Bool("b", 5, 1.7), True),
    ("Z_Bool", lambda m: m.Z_Bool("c", 1), False),
    ("S_Bool", lambda m: m.S_Bool("a", 0), False),
    ("T_Bool", lambda m: m.T_Bool("b", 3), False),
    ("RZ_Bool", lambda m: m.RZ_Bool("b", 1, 0.77), False),
    ("RZ_Bool.dag", lambda m: m.RZ_Bool("b", 1, 0.77), True),
    ("Y_Bool", lambda m: m.Y_Bool("c", 2), False),
    ("GlobalPhase", lambda m: m.GlobalPhase(complex(0.6, 0.8)), False),
    ("GlobalPhase.dag", lambda m: m.GlobalPhase(complex(0.6, 0.8)), True),
    ("ZeroConditionalPhaseFlip", lambda m: m.ZeroConditionalPhaseFlip(["f", "g"]), False),
    ("Reflection_Bool", lambda m: m.Reflection_Bool(["f"], False), False),
    ("Reflection_Bool.inverse", lambda m: m.Reflection_Bool(["g", "f"], True), False),
    ("Reflection_Bool.cond", lambda m: m.Reflection_Bool("c", True).conditioned_by_all_ones(["f", "g"]), False),
    ("Reflection_Bool.cond2", lambda m: m.Reflection_Bool("c", False).conditioned_by_all_ones("f"), False),
]


@pytest.mark.parametrize("name,make,dagger", DIAGONAL, ids=[d[0] for d in DIAGONAL])
def test_diagonal_ops(ps, oracle, name, make, dagger):
    """Each diagonal operator on its own (compensating errors cannot cancel)."""
    def build(mod):
        st = loaded(mod, SPEC3, 4000, seed=3)
        op = make(mod)
        (op.dag if dagger else op)(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, exact_amps=True, what=name)


def test_qram_load_conditioned(ps, oracle):
    rng = np.random.default_rng(5)
    mem = [int(x) for x in rng.integers(0, 1 << 9, size=1 << 7)]

    def build(mod):
        st = loaded(mod, SPEC3, 3000, seed=4)
        q = mod.QRAMCircuit_qutrit(7, 9, mem)
        mod.QRAMLoad(q, "a", "b").conditioned_by_nonzeros("f")(st)
        mod.QRAMLoad(q, "c", "b")(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, exact_amps=True, what="qram")


@pytest.mark.parametrize("case", ["address_rewritten", "three_tables", "shared_word", "phases_between"])
def test_qram_load_fused_programs(ps, oracle, case):
    """QRAM loads fused into one per-entry program: a load whose address an
    earlier op rewrites, three tables at once (two table words are read ahead
    of the ops, the third in place), two loads of one word, and phases between
    a load and its uncompute (the QRAM-interference pattern)."""
    rng = np.random.default_rng(9)
    m1 = [int(x) for x in rng.integers(0, 1 << 9, size=1 << 7)]
    m2 = [int(x) for x in rng.integers(0, 1 << 6, size=1 << 6)]
    m3 = [int(x) for x in rng.integers(0, 1 << 7, size=1 << 6)]

    def build(mod):
        st = loaded(mod, SPEC3, 3000, seed=6)
        q1 = mod.QRAMCircuit_qutrit(7, 9, m1)
        q2 = mod.QRAMCircuit_qutrit(6, 6, m2)
        q3 = mod.QRAMCircuit_qutrit(6, 7, m3)
        if case == "address_rewritten":
            ops = [mod.QRAMLoad(q1, "a", "b"), mod.Add_ConstUInt_InPlace("a", 5), mod.QRAMLoad(q1, "a", "b"),
                   mod.X_Bool("a", 2), mod.QRAMLoad(q1, "a", "b").conditioned_by_nonzeros("f")]
        elif case == "three_tables":
            ops = [mod.QRAMLoad(q1, "a", "b"), mod.QRAMLoad(q2, "c", "s"), mod.QRAMLoad(q3, "c", "a"),
                   mod.QRAMLoad(q1, "a", "b")]
        elif case == "shared_word":
            ops = [mod.QRAMLoad(q1, "a", "b"), mod.Z_Bool("b", 3), mod.QRAMLoad(q1, "a", "b"), mod.QRAMLoad(q1, "a", "b")]
        else:
            ops = [mod.QRAMLoad(q1, "a", "b")] + [mod.Z_Bool("b", i) for i in range(9)] + [mod.QRAMLoad(q1, "a", "b")]
        for o in ops:
            o(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, exact_amps=True, what=case)


# ---------------------------------------------------------------- branching
@pytest.mark.parametrize("nd", [1, 3, 6])
def test_hadamard_int_mixed_groups(ps, oracle, nd):
    """Random state: groups of size 1, partial and full occur together; n_digits <= width."""
    spec = [("x", "UnsignedInteger", 6), ("y", "UnsignedInteger", 5)]

    def build(mod):
        st = loaded(mod, spec, 700, seed=10 + nd)
        mod.Hadamard_Int("x", nd)(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what=f"Hadamard_Int nd={nd}")


def test_hadamard_int_conditioned(ps, oracle):
    spec = [("x", "UnsignedInteger", 5), ("y", "UnsignedInteger", 4), ("c", "Boolean", 1)]

    def build(mod):
        st = loaded(mod, spec, 300, seed=21)
        mod.Hadamard_Int("x", 5).conditioned_by_nonzeros("c")(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what="Hadamard_Int conditioned")


def test_group_transform_condition_on_target_rejected(ps):
    """A condition that reads the transformed bits of a group transform's own
    target is refused (the reference decides such a group by its first entry,
    hadamard.cpp:28,43,80, the engine per entry); a condition on the target's
    untouched upper bits is accepted.  The state is left unchanged."""
    st = loaded(ps, [("x", "UnsignedInteger", 5), ("y", "UnsignedInteger", 4)], 200, seed=22)
    before = st.export()
    with pytest.raises(ValueError):
        ps.Hadamard_Int("x", 5).conditioned_by_bit("x", 2)(st)
    with pytest.raises(ValueError):
        ps.Hadamard_Int_Full("x").conditioned_by_nonzeros("x")(st)
    assert_same_state(st.export(), before, exact_amps=True, what="state after a refused operator")


def test_hadamard_int_conditioned_on_upper_target_bits(ps, oracle):
    """Hadamard_Int on the low 3 bits of x, conditioned on bit 4 of x: every
    member of a group shares the upper bits, so per-entry and per-group
    evaluation agree and the engine matches the oracle."""
    spec = [("x", "UnsignedInteger", 5), ("y", "UnsignedInteger", 4)]

    def build(mod):
        st = loaded(mod, spec, 300, seed=23)
        mod.Hadamard_Int("x", 3).conditioned_by_bit("x", 4)(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what="Hadamard_Int conditioned on its upper bits")


@pytest.mark.parametrize("n", [1, 3, 40, 256])
def test_hadamard_int_full(ps, oracle, n):
    spec = [("x", "UnsignedInteger", 8), ("y", "UnsignedInteger", 3)]

    def build(mod):
        st = regs(mod, spec)
        rng = np.random.default_rng(n)
        xs = rng.choice(256, size=n, replace=False)
        vals = np.stack([xs, np.full(n, 5)], axis=1).astype(np.uint64)
        amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        st.load(vals, amps)
        mod.Hadamard_Int_Full("x")(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what=f"Hadamard_Int_Full n={n}")


def test_hadamard_bool_and_rotations(ps, oracle):
    spec = [("a", "UnsignedInteger", 5), ("r", "Rational", 6), ("q", "Boolean", 1), ("m", "UnsignedInteger", 3)]
    mat = [complex(0.6, 0.1), complex(-0.3, 0.2), complex(0.5, -0.4), complex(0.2, 0.7)]

    def build(mod):
        st = loaded(mod, spec, 1500, seed=30)
        mod.Hadamard_Bool("q")(st)
        mod.CondRot_Rational_Bool("r", "q")(st)
        mod.CondRot_Rational_Bool("r", "q").dag(st)
        mod.Rot_Bool("m", 1, mat)(st)
        mod.Rot_Bool("m", 2, [1, 0, 0, complex(0, 1)])(st)            # diagonal, in place
        mod.Rot_Bool("m", 0, [0, complex(0, -1), complex(0, 1), 0])(st)  # off-diagonal
        mod.Hadamard_Bool("q").conditioned_by_nonzeros("a")(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what="pair ops")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pair_ops_between_pointwise_programs(ps, oracle, seed):
    """A pair operator leaves partners across its bit adjacent; the next pair
    operator on that bit skips its grouping sort when only ops that neither read
    nor write the bit ran in between (flips and adds of other registers, phases
    and reflections reading it, ClearZero, Normalize).  Random sequences mixing
    those with ops that do read or write the bit (a flip of it, a flip
    conditioned on it, an add into a register holding it, a Hadamard on another
    bit), compared with the oracle after every operator."""
    spec = [("a", "Boolean", 1), ("b", "UnsignedInteger", 3), ("c", "Boolean", 1), ("d", "UnsignedInteger", 4)]
    mat = [complex(0.6, 0.1), complex(-0.3, 0.2), complex(0.5, -0.4), complex(0.2, 0.7)]
    ops = [
        lambda m: m.Hadamard_Bool("c"),
        lambda m: m.Hadamard_Bool("c").conditioned_by_all_ones("a"),
        lambda m: m.Rot_Bool("c", 0, mat).conditioned_by_all_ones("a"),
        lambda m: m.Rot_Bool("c", 0, mat),
        lambda m: m.X_Bool("a", 0),                                     # keeps the grouping on c
        lambda m: m.Add_UInt_UInt_InPlace("b", "d"),                    # keeps
        lambda m: m.Reflection_Bool("c", True),                          # amplitude only: keeps
        lambda m: m.GlobalPhase(complex(0.0, 1.0)),
        lambda m: m.X_Bool("a", 0).conditioned_by_all_ones("c"),        # reads c
        lambda m: m.X_Bool("c", 0),                                     # writes c
        lambda m: m.Hadamard_Bool("a"),                                 # pairs across another bit
        lambda m: m.ClearZero(1e-14),
        lambda m: m.Normalize(),
    ]
    rng = np.random.default_rng(0xA11 + seed)
    seq = [int(x) for x in rng.integers(0, len(ops), size=24)]
    sts = {}
    for mod in (ps, oracle):
        mod.System.clear()
        sts[mod.__name__] = loaded(mod, spec, 90, seed=50 + seed)
    for step, k in enumerate(seq):
        for mod in (ps, oracle):
            ops[k](mod)(sts[mod.__name__])
        assert_same_state(sts["sparqsim_amd"].export(), sts["_oracle"].export(), what=f"seed {seed} step {step} op {k}")


def test_condrot_general(ps, oracle):
    spec = [("r", "UnsignedInteger", 5), ("q", "Boolean", 1), ("z", "UnsignedInteger", 4)]

    def build(mod):
        st = loaded(mod, spec, 400, seed=31)
        mod.CondRot_General_Bool("r", "q", lambda v: mod.make_func(v, 5))(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what="CondRot_General_Bool")


# ---------------------------------------------------------------- structural
def test_register_structure(ps, oracle):
    def build(mod):
        st = loaded(mod, [("a", "UnsignedInteger", 10), ("b", "UnsignedInteger", 12)], 2000, seed=40)
        mod.AddRegister("t", mod.UnsignedInteger, 9)(st)
        mod.Add_UInt_UInt("a", "b", "t")(st)
        mod.SplitRegister("b", "b_lo", 5)(st)
        mod.Add_UInt_UInt_InPlace("b_lo", "a")(st)
        mod.CombineRegister("b", "b_lo")(st)
        mod.Add_UInt_UInt("a", "b", "t")(st)    # uncompute t ...
        mod.Add_UInt_UInt("a", "b", "t")(st)
        mod.Add_UInt_UInt("a", "b", "t")(st)    # ... t = a+b again
        mod.Init_Unsafe("t", 0)(st)
        mod.RemoveRegister("t")(st)
        mod.AddRegister("u", mod.UnsignedInteger, 3)(st)      # reuses the freed id
        mod.X_Bool("u", 1)(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, exact_amps=True, what="register structure")


def test_remove_entangled_register_raises(ps):
    st = ps.SparseState()
    ps.AddRegister("a", ps.UnsignedInteger, 4)(st)
    ps.Hadamard_Int("a", 4)(st)
    with pytest.raises(RuntimeError):
        ps.RemoveRegister("a")(st)


def test_add_systems_and_normalize(ps, oracle):
    spec = [("a", "UnsignedInteger", 9), ("b", "UnsignedInteger", 4)]

    def build(mod):
        s1 = loaded(mod, spec, 900, seed=50)
        s2 = mod.SparseState()
        vals, amps = random_state(mod, 900, seed=51)
        s2.load(vals, amps)
        mod.add_systems(s1, s2, 0.5)
        mod.Normalize()(s1)
        return s1

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what="add_systems")


def test_hadamard_int_duplicate_key_keeps_last(ps):
    """A state uploaded with a duplicated key (the reference's general-group
    Hadamard writes every member into its slot, hadamard.cpp:109-116, so one
    copy overwrites the other): the engine's grouping sort is stable and it
    places the later copy, every run the same.  Group b=1 holds a=3 (twice,
    amplitudes x then z) and a=5 (y); group b=2 holds only duplicates."""
    st = regs(ps, [("a", "UnsignedInteger", 4), ("b", "UnsignedInteger", 3)])
    x, y, z, u, v = 0.3 + 0.1j, -0.2 + 0.4j, 0.5 - 0.25j, 0.125j, -0.75
    vals = np.array([[3, 1], [5, 1], [3, 1], [6, 2], [6, 2]], dtype=np.uint64)
    st.load(vals, np.array([x, y, z, u, v]))
    ps.Hadamard_Int("a", 4)(st)
    ids, got_v, got_a = st.export()
    got = {(int(a), int(b)): amp for (a, b), amp in zip(got_v, got_a)}
    sign = lambda j, m: -1.0 if bin(j & m).count("1") & 1 else 1.0  # noqa: E731
    want = {}
    for j in range(16):
        want[(j, 1)] = (z * sign(j, 3) + y * sign(j, 5)) / 4
        want[(j, 2)] = v * sign(j, 6) / 4
    want = {k: a for k, a in want.items() if abs(a) ** 2 >= 1e-14}
    assert got.keys() == want.keys()
    assert max(abs(got[k] - want[k]) for k in want) <= 1e-15


def test_add_systems_to_itself(ps):
    """add_systems(s, s, c) appends s scaled by c and merges every key with its
    copy: each amplitude becomes a*(1+c) (one sum of two terms, so bit-exact
    against numpy); nothing falls under the merge threshold."""
    spec = [("a", "UnsignedInteger", 9), ("b", "UnsignedInteger", 4)]
    st = loaded(ps, spec, 900, seed=52)
    ids, vals, amps = st.export()
    ps.add_systems(st, st, 0.5)
    ps.add_systems(st, st, 1.0)
    want = (amps + amps * 0.5) * 2.0
    assert_same_state(st.export(), (ids, vals, want), exact_amps=True, what="add_systems(s, s)")


def test_sort_unconditional_order(ps, oracle):
    spec = [("a", "UnsignedInteger", 20), ("b", "UnsignedInteger", 30), ("c", "UnsignedInteger", 30)]

    def build(mod):
        st = loaded(mod, spec, 50000, seed=60)
        mod.SortUnconditional()(st)
        return st

    ps.System.clear()
    oracle.System.clear()
    eg = build(ps).export()
    eo = build(oracle).export()
    # both are in canonical order already: compare row by row without re-sorting
    assert np.array_equal(eg[1], eo[1])
    assert np.array_equal(eg[2], eo[2])


def test_clear_zero_threshold(ps, oracle):
    def build(mod):
        st = regs(mod, [("a", "UnsignedInteger", 10)])
        n = 1024
        vals = np.arange(n, dtype=np.uint64).reshape(n, 1)
        amps = np.logspace(-9, 0, n) * np.exp(1j * np.arange(n))
        st.load(vals, amps)
        mod.ClearZero(1e-8)(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, exact_amps=True, what="ClearZero")


@pytest.mark.parametrize("touched", [False, True], ids=["fresh_basis", "after_phase"])
def test_single_entry_expansion_clear_zero(ps, oracle, touched):
    """H(a, 20) on one entry: every output is +-a/2^10, so ClearZero keeps all of
    them or none.  A fresh basis state (amplitude exactly 1) is decided on the
    host; after another operator the decision is read from the device.  Both
    sides of |a|^2 / 2^20 = 2^-20 against the oracle."""
    old = ps.get_clearzero_default_eps()
    try:
        for eps, expect in ((2.0 ** -20 * 1.0001, 0), (2.0 ** -20 * 0.9999, 1 << 20)):
            sizes = []
            for mod in (ps, oracle):
                mod.set_clearzero_default_eps(eps)
                mod.System.clear()
                st = mod.SparseState()
                mod.AddRegister("a", mod.UnsignedInteger, 20)(st)
                mod.reset_systems(st)
                if touched:
                    mod.GlobalPhase(complex(1.0, 0.0))(st)
                mod.Hadamard_Int("a", 20)(st)
                sizes.append(st.size())
                mod.set_clearzero_default_eps(old)
            assert sizes == [expect, expect], (eps, sizes)
    finally:
        ps.set_clearzero_default_eps(old)
        oracle.set_clearzero_default_eps(old)


@pytest.mark.parametrize("op", ["Div_Sqrt_Arccos_UInt_UInt", "Sqrt_Div_Arccos_Int_UInt", "GetRotateAngle_Int_Int"])
def test_rational_angle_ops(ps, oracle, op):
    """Arc-cosine / arc-tangent ops into a 12-bit Rational register over every
    input pair of small registers (acos arguments kept in [0, 1]), as one
    superposition: keys bit-exact against the oracle."""
    if op == "Div_Sqrt_Arccos_UInt_UInt":
        spec = [("x", "UnsignedInteger", 5), ("y", "UnsignedInteger", 5)]
        pairs = [(a, b) for b in range(1, 32) for a in range(b + 1)]
    elif op == "Sqrt_Div_Arccos_Int_UInt":
        spec = [("x", "SignedInteger", 5), ("y", "UnsignedInteger", 6)]
        pairs = [(l, r) for r in range(1, 64) for l in range(int(np.sqrt(r)) + 1)]
    else:
        spec = [("x", "SignedInteger", 4), ("y", "SignedInteger", 4)]
        pairs = [(l, r) for l in range(16) for r in range(16)]

    def build(mod):
        st = regs(mod, spec + [("o", "Rational", 12)])
        vals = np.array([(a, b, 0) for a, b in pairs], dtype=np.uint64)
        st.load(vals, np.full(len(pairs), 1.0 / np.sqrt(len(pairs)), dtype=np.complex128))
        getattr(mod, op)("x", "y", "o")(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, exact_amps=True, what=op)


@pytest.mark.parametrize("w", [3, 7, 31, 32, 33, 48, 52, 53, 63, 64])
def test_rational_encoding_every_width(ps, oracle, w):
    """The device's fraction -> w-bit Rational encoding equals the oracle's
    (pinned against the reference's 2^-w bound and exact dyadics in
    test_oracle.py): arc-tangent and arc-cosine outputs of every input pair,
    bit-exact up to 48 bits.  Wider outputs resolve the double result of atan2 /
    acos itself, where the device math library and the host's may differ by an
    ulp (neither is correctly rounded): there a difference of up to
    2 ulp -> 2 * 2^(w-53) + 1 codes is allowed (ulp of a value in [0, 1) <= 2^-53)."""
    spec = [("x", "SignedInteger", 4), ("y", "SignedInteger", 4)]
    pairs = [(l, r) for l in range(16) for r in range(16)]

    for op in ("GetRotateAngle_Int_Int", "Div_Sqrt_Arccos_UInt_UInt"):
        def build(mod):
            sp = spec if op == "GetRotateAngle_Int_Int" else [("x", "UnsignedInteger", 4), ("y", "UnsignedInteger", 4)]
            st = regs(mod, sp + [("o", "Rational", w)])
            use = pairs if op == "GetRotateAngle_Int_Int" else [(a, b) for b in range(1, 16) for a in range(b + 1)]
            vals = np.array([(a, b, 0) for a, b in use], dtype=np.uint64)
            st.load(vals, np.full(len(use), 1.0 / np.sqrt(len(use)), dtype=np.complex128))
            getattr(mod, op)("x", "y", "o")(st)
            return st

        eg, eo = run_both(ps, oracle, build)
        if w <= 48:
            assert_same_state(eg, eo, exact_amps=True, what=f"{op} into {w} bits")
            continue
        from parity import canonical

        (_, vg, ag), (_, vr, ar) = canonical(eg), canonical(eo)   # inputs sort first, so rows pair up
        assert np.array_equal(vg[:, :2], vr[:, :2]) and np.array_equal(ag, ar)
        tol = 2 * 2 ** (w - 53) + 1
        diff = np.abs(vg[:, 2].astype(object) - vr[:, 2].astype(object))
        assert max(diff) <= tol, (op, w, max(diff))


def test_operator_sequence(ps, oracle):
    """OperatorSequence applies its operators in order (one native call) and its
    dag applies their adjoints in reverse: the same state as the operators one
    by one, and the identity after op; op.dag."""
    spec = [("a", "UnsignedInteger", 6), ("b", "UnsignedInteger", 6), ("f", "Boolean", 1)]

    def ops(mod):
        return [mod.Hadamard_Int("a", 6), mod.Add_UInt_UInt_InPlace("a", "b"), mod.Phase_Bool("b", 2, 0.3),
                mod.RZ_Bool("f", 0, 1.1), mod.Hadamard_Bool("f"), mod.X_Bool("b", 5)]

    def build(mod):
        st = regs(mod, spec)
        for op in ops(mod):
            op(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    ps.System.clear()
    st = regs(ps, spec)
    seq = ps.OperatorSequence(ops(ps))
    assert len(seq) == 6
    seq(st)
    assert_same_state(st.export(), eo, what="sequence vs oracle")
    seq.dag(st)
    ids, vals, amps = st.export()
    assert vals.shape[0] == 1 and np.all(vals == 0) and abs(amps[0] - 1) < 1e-12


def test_empty_state_ops(ps):
    st = ps.SparseState()
    ps.AddRegister("a", ps.UnsignedInteger, 4)(st)
    ps.X_Bool("a", 0)(st)
    ps.ClearZero(2.0)(st)      # removes the only entry
    assert st.size() == 0
    ps.Hadamard_Int("a", 4)(st)
    ps.X_Bool("a", 1)(st)
    ps.Hadamard_Bool  # noqa: B018
    assert st.size() == 0


def test_condition_truth_tables(ps):
    """Condition evaluation on the engine against the predicates of
    global_macros.h:13-17, 82-86, brute force over three small registers (the
    same table pins the oracle); moved keys keep their amplitudes bit for bit."""
    check_condition_truth_tables(ps)
