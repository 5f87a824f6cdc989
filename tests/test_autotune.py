"""Launch-shape autotuner of the fused pointwise kernel (entries per thread 1, 2
or 4): every shape gives the same state bit for bit, and the tuner settles on a
shape per class after its timed first calls.  Key widths 1, 2, 3 and 5 words,
light and full interpreter programs, with and without amplitude access."""
import numpy as np
import pytest

from parity import assert_same_state, random_state


def test_report_without_calls_is_empty_json():
    import sparqsim_amd as ps

    r = ps.autotune_report()
    assert set(r) == {"pointwise", "radix"}
    assert isinstance(r["pointwise"], list) and isinstance(r["radix"], list)


def _program(ps, heavy):
    """Key rewrites over every word, a QRAM load and phases; `heavy` adds a
    square root and a division (the full interpreter)."""
    ops = [ps.Add_UInt_UInt("a", "b", "c"), ps.X_Bool("d", 3), ps.Xor_UInt_UInt("d", "a", "c"),
           ps.Z_Bool("c", 1), ps.Phase_Bool("a", 2, 0.37)]
    if heavy:
        ops += [ps.Sqrt_UInt("a", "e"), ps.Div_UInt_UInt("d", "b", "e")]
    return ops


def _run(ps, widths, n, items, heavy, phases):
    ps.System.clear()
    ps.set_pointwise_items(items)
    st = ps.SparseState()
    for name, w in widths:
        ps.AddRegister(name, ps.UnsignedInteger, w)(st)
    vals, amps = random_state(ps, n, seed=11)
    st.load(vals, amps)
    ops = _program(ps, heavy)
    if not phases:
        ops = [o for o in ops if not isinstance(o, (ps.Z_Bool, ps.Phase_Bool))]
    for _ in range(3):
        for o in ops:
            o(st)
        st.size()                  # one pointwise launch per round
    out = st.export()
    ps.set_pointwise_items(0)
    return out


# 1, 2, 3 and 5 key words (fields are packed in register order, none
# straddles a word)
LAYOUTS = {
    1: [("a", 12), ("b", 12), ("c", 13), ("d", 12), ("e", 12)],
    2: [("a", 30), ("b", 30), ("c", 31), ("d", 30), ("e", 3)],
    3: [("a", 40), ("b", 40), ("c", 41), ("d", 20), ("e", 3)],
    5: [("a", 60), ("b", 60), ("c", 61), ("d", 60), ("e", 60)],
}


def test_layouts_have_their_word_counts():
    for words, regs in LAYOUTS.items():
        used, room = 1, 64
        for _, w in regs:
            if w > room:
                used, room = used + 1, 64
            room -= w
        assert used == words, (words, used)


@pytest.mark.gpu
@pytest.mark.parametrize("words", sorted(LAYOUTS))
@pytest.mark.parametrize("heavy", [False, True], ids=["light", "full"])
def test_every_shape_gives_the_same_state(ps, words, heavy):
    n = 70000
    ref = _run(ps, LAYOUTS[words], n, 1, heavy, True)
    for items in (2, 4):
        got = _run(ps, LAYOUTS[words], n, items, heavy, True)
        assert_same_state(got, ref, exact_amps=True, what=f"{words} words, {items} entries per thread")
    keys_only = _run(ps, LAYOUTS[words], n, 2, heavy, False)
    assert len(keys_only[2]) == n


@pytest.mark.gpu
def test_tuner_settles_and_keeps_results(ps):
    """Enough calls of one class finish its measurements; the chosen shape is
    one of the candidates and the state equals a forced-shape run's."""
    widths = LAYOUTS[1]
    ref = _run(ps, widths, 1 << 17, 1, False, True)
    ps.System.clear()
    st = ps.SparseState()
    for name, w in widths:
        ps.AddRegister(name, ps.UnsignedInteger, w)(st)
    vals, amps = random_state(ps, 1 << 17, seed=11)
    st.load(vals, amps)
    ops = _program(ps, False)
    for _ in range(3):
        for o in ops:
            o(st)
        st.size()
    assert_same_state(st.export(), ref, exact_amps=True, what="autotuned")
    # 3 + 6 launches of the class: two warm-up calls, then two per candidate
    for _ in range(6):
        for o in ops:
            o(st)
        st.size()
    rows = [r for r in ps.autotune_report()["pointwise"] if r["words"] == 1 and r["log2_n"] == 17 and r["amp"] == 1 and r["light"] == 1]
    assert rows and rows[0]["items"] in (1, 2, 4), rows
    assert all(v > 0 for v in rows[0]["ms_per_2p20"]), rows
