"""Radix-sort variants against the oracle: 8- and 11-bit digits (forced), digits
made of several bit ranges, the skipped passes of an already non-decreasing
target field, and conditioned pair operators evaluated inside the pair kernels.
Keys bit-exact, amplitudes to 1e-12 relative (parity.py)."""
import numpy as np
import pytest

from parity import assert_same_state, random_state, run_both

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[8, 11, 0], ids=["radix8", "radix11", "auto"])
def radix(request, ps):
    ps.set_radix_bits(request.param)
    yield request.param
    ps.set_radix_bits(0)


def _regs(mod, widths):
    st = mod.SparseState()
    for name, w in widths:
        mod.AddRegister(name, mod.UnsignedInteger, w)(st)
    return st


@pytest.mark.parametrize("n", [300, 9000, 70000])
def test_sort_unconditional(ps, oracle, radix, n):
    widths = [("a", 13), ("b", 9), ("c", 1), ("d", 17), ("e", 5)]

    def build(mod):
        st = _regs(mod, widths)
        vals, amps = random_state(mod, n, seed=n)
        st.load(vals, amps)
        mod.SortUnconditional()(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    # after SortUnconditional both are in canonical order: compare without re-sorting
    assert np.array_equal(eg[1], eo[1])
    assert_same_state(eg, eo, exact_amps=True, what="SortUnconditional")


@pytest.mark.parametrize("target", ["a", "c", "d"])
def test_hadamard_full_on_sorted_and_unsorted(ps, oracle, radix, target):
    """`a` is the most significant register: on a canonically sorted state its
    values are non-decreasing and its digit passes are skipped."""
    widths = [("a", 6), ("b", 10), ("c", 5), ("d", 4)]

    def build(mod):
        st = _regs(mod, widths)
        vals, amps = random_state(mod, 6000, seed=7)
        st.load(vals, amps)
        mod.SortUnconditional()(st)
        mod.Hadamard_Int_Full(target)(st)
        mod.X_Bool("c", 1)(st)
        mod.Hadamard_Int_Full(target)(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what=f"Hadamard_Int_Full({target})")


def test_two_word_keys(ps, oracle, radix):
    widths = [("a", 40), ("b", 30), ("c", 12), ("d", 1)]

    def build(mod):
        st = _regs(mod, widths)
        vals, amps = random_state(mod, 12000, seed=3)
        st.load(vals, amps)
        mod.Hadamard_Int("c", 5)(st)
        mod.Hadamard_Bool("d")(st)
        mod.SortUnconditional()(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what="two-word keys")


@pytest.mark.parametrize("kind", ["hadamard_bool", "rot_bool"])
def test_conditioned_pair_ops_inline(ps, oracle, radix, kind):
    widths = [("m", 9), ("f", 2), ("t", 3), ("g", 1), ("h", 1)]

    def build(mod):
        st = _regs(mod, widths)
        vals, amps = random_state(mod, 1500, seed=11)
        st.load(vals, amps)
        if kind == "hadamard_bool":
            op = mod.Hadamard_Bool("h")
        else:
            op = mod.Rot_Bool("t", 1, [0.6, 0.8j, 0.8j, 0.6])
        op.conditioned_by_all_ones("f")(st)
        mod.CondRot_Rational_Bool("m", "h")(st)           # not controllable in the reference
        mod.Hadamard_Bool("g").conditioned_by_value("t", 5)(st)
        return st

    eg, eo = run_both(ps, oracle, build)
    assert_same_state(eg, eo, what=f"conditioned {kind}")
