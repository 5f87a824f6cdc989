"""The rest of the pysparq._core surface on the operator-apply path
(PySparQ/core.cpp): SparseState.basis_states, StateStorage, the grouping
sorts (sort_state.cpp:10-171), Hadamard_Partial (hadamard.cpp:470-712),
MoveBackRegister / Push / Pop (system_operations.cpp:223-369),
split_systems / combine_systems (:53-106) and merge_system / remove_system
(basic_components.cpp:517-526).

Orders are checked against the reference's comparison predicates restated in
Python (an order with ties is checked by "no neighbour pair out of order" plus
"same entries"); transforms against dense numpy models; structure operations
by round trips."""
import itertools

import numpy as np
import pytest

from parity import assert_same_state, random_state

SPEC = [("a", "UnsignedInteger", 5), ("b", "UnsignedInteger", 4), ("c", "UnsignedInteger", 3)]


def regs(mod, spec=SPEC):
    st = mod.SparseState()
    for name, typ, w in spec:
        mod.AddRegister(name, getattr(mod, typ), w)(st)
    return st


def loaded(mod, n, seed, spec=SPEC):
    st = regs(mod, spec)
    vals, amps = random_state(mod, n, seed)
    st.load(vals, amps)
    return st


def rows(st):
    ids, vals, amps = st.export()
    return list(ids), [tuple(int(x) for x in v) for v in vals], list(amps)


def same_entries(before, after):
    """Same (key, amplitude) multiset."""
    a = sorted(zip(before[1], [(z.real, z.imag) for z in before[2]]))
    b = sorted(zip(after[1], [(z.real, z.imag) for z in after[2]]))
    return a == b


def check_order(r, less):
    """No neighbour pair (x, y) with y < x under the reference predicate."""
    keys = r[1]
    amps = r[2]
    for i in range(len(keys) - 1):
        assert not less((keys[i + 1], amps[i + 1]), (keys[i], amps[i])), (i, keys[i], keys[i + 1])


# ---------------------------------------------------------------- CPU
def test_merge_system_and_remove_system():
    """merge_system adds b's amplitude into a and zeroes b; remove_system is
    |a|^2 < 1e-14 (basic_components.cpp:517-526)."""
    import sparqsim_amd as ps

    ps.System.clear()
    ps.System.add_register("r", ps.UnsignedInteger, 3)
    a, b = ps.System(), ps.System()
    a.amplitude, b.amplitude = 0.25 + 0.5j, -0.125 + 0.0j
    ps.merge_system(a, b)
    assert a.amplitude == 0.125 + 0.5j and b.amplitude == 0
    assert ps.remove_system(b)
    b.amplitude = 0.9e-7                 # |a|^2 = 8.1e-15 < 1e-14
    assert ps.remove_system(b)
    b.amplitude = 2e-7
    assert not ps.remove_system(b)
    assert isinstance(ps.StateStorage().value, int) and ps.StateStorage().value == 0
    ps.System.clear()


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_basis_states(ps):
    st = loaded(ps, 60, seed=1)
    ids, vals, amps = st.export()
    got = st.basis_states
    assert len(got) == len(amps)
    for s, v, a in zip(got, vals, amps):
        assert s.amplitude == a
        assert [s.get(i).value for i in ids] == [int(x) for x in v]
        assert [s.registers[i].value for i in ids] == [int(x) for x in v]
        assert s.get("b").value == int(v[1])
    # canonical order of the System objects (System::operator<)
    srt = sorted(got, key=lambda s: tuple(s.get(i).value for i in ids))
    assert all(srt[k] < srt[k + 1] for k in range(len(srt) - 1))
    assert srt[0] == srt[0] and not (srt[0] != srt[0])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["SortExceptKey", "SortByKey", "SortExceptBit", "SortExceptKeyHadamard", "SortByKey2",
                                  "SortByAmplitude"])
def test_grouping_sorts_follow_reference_predicates(ps, case):
    st = loaded(ps, 500, seed=2)
    before = rows(st)
    ia, ib, ic = (ps.System.get_id(n) for n in ("a", "b", "c"))
    pos = {i: k for k, i in enumerate(before[0])}
    A, B, C = pos[ia], pos[ib], pos[ic]

    def others_except(x, skip):
        return tuple(v for k, v in enumerate(x) if k not in skip)

    if case == "SortExceptKey":          # sort_state.cpp:10-36
        op, less = ps.SortExceptKey("b"), lambda x, y: (others_except(x[0], {B}), x[0][B]) < (others_except(y[0], {B}), y[0][B])
    elif case == "SortByKey":            # :46-53
        op, less = ps.SortByKey("c"), lambda x, y: x[0][C] < y[0][C]
    elif case == "SortExceptBit":        # :55-82, digit 2 of a masked
        m = ~(1 << 2)
        op, less = ps.SortExceptBit("a", 2), lambda x, y: (others_except(x[0], {A}), x[0][A] & m) < (others_except(y[0], {A}), y[0][A] & m)
    elif case == "SortExceptKeyHadamard":   # :138-171, qubits {0, 3} of a
        m = ~((1 << 0) | (1 << 3))

        def key(x):
            return tuple((v & m) if k == A else v for k, v in enumerate(x[0])) + (x[0][A],)
        op, less = ps.SortExceptKeyHadamard("a", {0, 3}), lambda x, y: key(x) < key(y)
    elif case == "SortByKey2":           # :101-131
        op, less = ps.SortByKey2("c", "a"), lambda x, y: (x[0][C], x[0][A], others_except(x[0], {C, A})) < \
            (y[0][C], y[0][A], others_except(y[0], {C, A}))
    else:                                # :90-99
        op, less = ps.SortByAmplitude(), lambda x, y: abs(x[1]) ** 2 > abs(y[1]) ** 2
    op(st)
    after = rows(st)
    assert same_entries(before, after)
    check_order(after, less)
    op.dag(st)                           # self-adjoint: sorting again changes nothing
    assert rows(st) == after


@pytest.mark.gpu
def test_grouping_sort_is_stable(ps):
    """Entries equal under the ordering keep their order (the engine's sorts
    are stable; the reference leaves it open)."""
    st = loaded(ps, 400, seed=3)
    ps.SortUnconditional()(st)
    canon = rows(st)
    ps.SortByKey("c")(st)
    got = rows(st)
    C = canon[0].index(ps.System.get_id("c"))
    want = sorted(zip(canon[1], canon[2]), key=lambda e: e[0][C])   # Python's sort is stable
    assert got[1] == [k for k, _ in want]


def _dense_partial_h(keys, amps, col, positions, width):
    out = {}
    k = len(positions)
    for key, amp in zip(keys, amps):
        v = key[col]
        sub = sum(((v >> q) & 1) << j for j, q in enumerate(positions))
        base = v & ~sum(1 << q for q in positions)
        for t in range(1 << k):
            nv = base | sum(((t >> j) & 1) << q for j, q in enumerate(positions))
            sign = -1.0 if bin(t & sub).count("1") & 1 else 1.0
            nk = key[:col] + (nv,) + key[col + 1:]
            out[nk] = out.get(nk, 0) + sign * amp / np.sqrt(2.0 ** k)
    return {kk: a for kk, a in out.items() if abs(a) ** 2 >= 1e-14}


@pytest.mark.gpu
@pytest.mark.parametrize("positions", [{2}, {0, 3}, {1, 2, 4}])
def test_hadamard_partial_against_dense_model(ps, positions):
    st = loaded(ps, 120, seed=4)
    ids, keys, amps = rows(st)
    col = ids.index(ps.System.get_id("a"))
    want = _dense_partial_h(keys, amps, col, sorted(positions), 5)
    ps.Hadamard_Partial("a", positions)(st)
    _, gk, ga = rows(st)
    got = dict(zip(gk, ga))
    assert got.keys() == want.keys()
    scale = max(abs(a) for a in want.values())
    assert max(abs(got[k] - want[k]) for k in want) <= 1e-12 * scale
    ps.Hadamard_Partial("a", positions).dag(st)      # H is its own inverse
    _, bk, ba = rows(st)
    back = dict(zip(bk, ba))
    orig = dict(zip(keys, amps))
    assert back.keys() == orig.keys()
    assert max(abs(back[k] - orig[k]) for k in orig) <= 1e-12 * max(abs(a) for a in orig.values())


@pytest.mark.gpu
def test_hadamard_partial_conditioned(ps):
    """Entries with c == 5 get H on bit 1 of a; the others are unchanged."""
    st = loaded(ps, 200, seed=5)
    ids, keys, amps = rows(st)
    col, cc = ids.index(ps.System.get_id("a")), ids.index(ps.System.get_id("c"))
    sel = [(k, a) for k, a in zip(keys, amps) if k[cc] == 5]
    rest = {k: a for k, a in zip(keys, amps) if k[cc] != 5}
    want = _dense_partial_h([k for k, _ in sel], [a for _, a in sel], col, [1], 5)
    want.update(rest)
    ps.Hadamard_Partial("a", [1]).conditioned_by_value("c", 5)(st)
    _, gk, ga = rows(st)
    got = dict(zip(gk, ga))
    assert got.keys() == want.keys()
    assert max(abs(got[k] - want[k]) for k in want) <= 1e-12 * max(abs(a) for a in want.values())


@pytest.mark.gpu
def test_move_back_register(ps):
    """MoveBackRegister(a): a takes the last id, the register there takes a's
    id, the values travel with their registers and the canonical order
    follows the new ids."""
    st = loaded(ps, 150, seed=6)
    ids, keys, amps = rows(st)
    by_name = {ps.System.name_of(i): [k[j] for k in keys] for j, i in enumerate(ids)}
    ia, ic = ps.System.get_id("a"), ps.System.get_id("c")
    ps.MoveBackRegister("a")(st)
    assert ps.System.get_id("a") == ic and ps.System.get_id("c") == ia
    ids2, keys2, amps2 = rows(st)
    assert ids2 == ids
    got = {ps.System.name_of(i): [k[j] for k in keys2] for j, i in enumerate(ids2)}
    perm = sorted(range(len(amps)), key=lambda e: tuple(by_name[ps.System.name_of(i)][e] for i in ids2))
    ps.SortUnconditional()(st)
    ids3, keys3, amps3 = rows(st)
    for name in ("a", "b", "c"):
        assert [k[ids3.index(ps.System.get_id(name))] for k in keys3] == [by_name[name][e] for e in perm], name
    assert amps3 == [amps[e] for e in perm]
    assert sorted(got["a"]) == sorted(by_name["a"])
    ps.MoveBackRegister("c")(st)                 # c is at a's old id; moving it back restores the table
    assert ps.System.get_id("a") == ia and ps.System.get_id("c") == ic


@pytest.mark.gpu
def test_push_pop_round_trip(ps):
    st = loaded(ps, 100, seed=7)
    before = rows(st)
    b = before[0].index(ps.System.get_id("b"))
    ps.Push("b", "b_saved")(st)
    mid = rows(st)
    g = mid[0].index(ps.System.get_id("b_saved"))
    bb = mid[0].index(ps.System.get_id("b"))
    assert all(k[bb] == 0 for k in mid[1])
    assert sorted(k[g] for k in mid[1]) == sorted(k[b] for k in before[1])
    ps.X_Bool("b", 1)(st)                        # work on the fresh register, then undo it
    ps.X_Bool("b", 1)(st)
    ps.Pop("b")(st)
    after = rows(st)
    assert after[0] == before[0] and same_entries(before, after)
    assert "b_saved" not in [ps.System.name_of(i) for i in after[0]]     # the stacked register is gone
    push = ps.Push("b", "tmp")
    push(st)
    push.dag(st)                                 # Push.dag is Pop
    assert same_entries(before, rows(st))
    with pytest.raises(RuntimeError):
        ps.Pop("b")(st)                          # nothing stacked


@pytest.mark.gpu
def test_split_and_combine_systems(ps):
    st = loaded(ps, 300, seed=8)
    before = rows(st)
    B, C = (before[0].index(ps.System.get_id(n)) for n in ("b", "c"))
    rest = ps.split_systems(st, ["b"], [], [("c", 0)], [])
    kept, moved = rows(st), rows(rest)
    sat = [(k, a) for k, a in zip(before[1], before[2]) if k[B] != 0 and k[C] & 1]
    uns = [(k, a) for k, a in zip(before[1], before[2]) if not (k[B] != 0 and k[C] & 1)]
    assert list(zip(kept[1], kept[2])) == sat                  # in order (stable)
    assert list(zip(moved[1], moved[2])) == uns
    ps.combine_systems(st, rest)
    assert list(zip(*rows(st)[1:])) == sat + uns
    assert rest.size() == len(uns)                             # from_ is unchanged
