"""The dense-domain Hadamard path (Hadamard_Int / Hadamard_Int_Full through a
dense amplitude array when the target bits and the other varying key bits
span a small, well filled domain) against the grouped path and the oracle:
keys bit-exact, amplitudes to 1e-12 relative.
Cases: target above, below and between the other varying bits, n_digits
below the register width, the full variant, a sparsely filled domain (grouped
path), and duplicate keys (detected, grouped path)."""
import numpy as np
import pytest

from parity import assert_same_state, run_both

pytestmark = pytest.mark.gpu


def _state(mod, spec, vals, amps):
    st = mod.SparseState()
    for name, w in spec:
        mod.AddRegister(name, mod.UnsignedInteger, w)(st)
    st.load(vals, amps)
    return st


def _random_rows(rng, spec, n, fill=0.9):
    """n distinct rows over the registers' domain (drawn without replacement
    from a product domain filled to about `fill`)."""
    sizes = [1 << w for _, w in spec]
    total = int(np.prod(sizes, dtype=np.float64))
    idx = rng.choice(total, size=min(n, int(total * fill)), replace=False)
    cols, rest = [], idx.astype(np.uint64)
    for s in reversed(sizes):
        cols.append(rest % np.uint64(s))
        rest //= np.uint64(s)
    vals = np.stack(list(reversed(cols)), axis=1).astype(np.uint64)
    amps = rng.standard_normal(len(vals)) + 1j * rng.standard_normal(len(vals))
    return vals, amps / np.linalg.norm(amps)


def _kernels(ps, fn):
    from sparqsim_amd.profiling import KernelProfile

    with KernelProfile(ps) as prof:
        out = fn()
    return out, {r["name"] for r in prof.records}


def _both_paths(ps, spec, vals, amps, op):
    out = []
    for on in (True, False):
        ps.set_dense_transform(on)
        try:
            ps.System.clear()
            st = _state(ps, spec, vals, amps)
            (e, names) = _kernels(ps, lambda: (op(st), st.export())[1])
            out.append((e, names))
        finally:
            ps.set_dense_transform(True)
    return out


CASES = [
    # (registers, target, n_digits, full)
    ([("main", 14), ("anc", 3), ("q", 1)], "main", 14, False),        # target most significant (QDA's main register)
    ([("hi", 3), ("t", 12), ("lo", 2)], "t", 12, False),               # target between the other varying bits
    ([("a", 5), ("t", 13)], "t", 13, True),                            # target least significant, the full variant
    ([("t", 16), ("b", 2)], "t", 11, False),                           # n_digits below the width: upper target bits group
]


@pytest.mark.parametrize("spec,target,nd,full", CASES)
def test_dense_path_matches_grouped_path(ps, spec, target, nd, full):
    rng = np.random.default_rng(nd * 7 + len(spec))
    total_bits = sum(w for _, w in spec)
    vals, amps = _random_rows(rng, spec, 1 << (total_bits - 1), fill=0.6)
    op = (lambda st: ps.Hadamard_Int_Full(target)(st)) if full else (lambda st: ps.Hadamard_Int(target, nd)(st))
    (dense, dn), (grouped, gn) = _both_paths(ps, spec, vals, amps, op)
    assert "dense_scatter_kernel" in dn and "dense_scatter_kernel" not in gn
    assert_same_state(dense, grouped, what="dense vs grouped")


def test_dense_path_matches_oracle(ps, oracle):
    spec = [("main", 13), ("anc", 2), ("q", 1)]
    rng = np.random.default_rng(5)
    vals, amps = _random_rows(rng, spec, 1 << 15, fill=0.75)

    def build(mod):
        st = _state(mod, spec, vals, amps)
        mod.Hadamard_Int("main", 13)(st)
        mod.Hadamard_Bool("q")(st)
        mod.Hadamard_Int_Full("main")(st)
        return st

    ps.sort_stats(reset=True)
    (eg, eo), names = _kernels(ps, lambda: run_both(ps, oracle, build))
    assert "dense_scatter_kernel" in names
    assert_same_state(eg, eo, what="dense path vs oracle")


def test_sparse_domain_keeps_grouped_path(ps):
    """A domain filled below an eighth takes the grouped path."""
    spec = [("main", 12), ("anc", 7)]
    rng = np.random.default_rng(6)
    vals, amps = _random_rows(rng, spec, 1 << 15, fill=0.0625)
    (dense, dn), (grouped, _) = _both_paths(ps, spec, vals, amps, lambda st: ps.Hadamard_Int("main", 12)(st))
    assert "dense_scatter_kernel" not in dn
    assert_same_state(dense, grouped, what="sparse domain")


def test_duplicate_keys_fall_back(ps):
    """An uploaded state with a repeated key: the dense scatter sees the
    collision and the grouped path (which keeps the last duplicate) runs."""
    spec = [("main", 13), ("anc", 2)]
    rng = np.random.default_rng(7)
    vals, amps = _random_rows(rng, spec, 1 << 14, fill=0.9)
    vals = np.concatenate([vals, vals[10:11]])
    amps = np.concatenate([amps, [0.25 + 0.5j]])
    (dense, dn), (grouped, _) = _both_paths(ps, spec, vals, amps, lambda st: ps.Hadamard_Int("main", 13)(st))
    assert "dense_scatter_kernel" in dn and "group_start_kernel" in dn
    assert_same_state(dense, grouped, exact_amps=True, what="duplicate keys")


@pytest.mark.parametrize("fill,cond", [(0.9, False), (0.05, False), (0.9, True)])
def test_full_variant_rejects_duplicate_keys(ps, fill, cond):
    """Hadamard_Int_Full checks for duplicate keys before it changes anything
    (CheckDuplicateKey, hadamard.cpp:183): through the dense-domain path (its
    slot marks), the grouped path (after its sort) and, with a condition, on
    the whole state first.  The state keeps its entries."""
    spec = [("main", 11), ("anc", 3)]
    rng = np.random.default_rng(8)
    vals, amps = _random_rows(rng, spec, 1 << 14, fill=fill)
    vals = np.concatenate([vals, vals[3:4]])
    amps = np.concatenate([amps, [0.5 - 0.25j]])
    ps.System.clear()
    st = _state(ps, spec, vals, amps)
    op = ps.Hadamard_Int_Full("main")
    if cond:
        op = op.conditioned_by_nonzeros("anc")
    with pytest.raises(RuntimeError, match="duplicate"):
        op(st)
    assert st.size() == len(vals)
    # without the repeated row the same transform goes through
    ps.System.clear()
    st = _state(ps, spec, vals[:-1], amps[:-1])
    ps.Hadamard_Int_Full("main")(st)
    assert st.size() > 0


@pytest.mark.parametrize("fill,kernel", [(0.9, "tx_emit_kernel"), (0.05, "tx_emit_kernel")])
def test_output_in_canonical_order(ps, oracle, fill, kernel):
    """With every other varying bit below the target, Hadamard_Int_Full emits
    in canonical order (target value major): through the dense-domain path
    (fill 0.9) and through the grouped expansion (fill 0.05, a few groups).
    The rows come back sorted, and equal the oracle's."""
    spec = [("main", 12), ("anc", 3)]
    rng = np.random.default_rng(9)
    vals, amps = _random_rows(rng, spec, 1 << 14, fill=fill)
    ps.System.clear()
    st = _state(ps, spec, vals, amps)
    eg, names = _kernels(ps, lambda: (ps.Hadamard_Int_Full("main")(st), st.export())[1])
    assert kernel in names and (("dense_scatter_kernel" in names) == (fill > 0.5))
    oracle.System.clear()
    so = _state(oracle, spec, vals, amps)
    oracle.Hadamard_Int_Full("main")(so)
    assert_same_state(eg, so.export(), what=f"fill {fill}")
    v = np.asarray(eg[1], dtype=np.uint64)
    order = np.lexsort(tuple(v[:, k] for k in reversed(range(v.shape[1]))))
    assert np.array_equal(order, np.arange(len(order))), "output not in canonical order"


def test_concentrating_transform_keeps_slot_order(ps, oracle):
    """A transform whose output is far smaller than its input (uniform rows of
    main per anc value go to main = 0) takes the dense path with the slot-order
    emit; the result equals the oracle's."""
    spec = [("main", 12), ("anc", 3)]
    rows = np.array([[m, a] for a in (1, 4, 6, 7) for m in range(1 << 12)], dtype=np.uint64)
    amps = np.repeat(np.array([0.5, -0.25j, 0.125 + 0.5j, 0.3]), 1 << 12) / np.sqrt(1 << 12)
    amps = amps / np.linalg.norm(amps)
    ps.System.clear()
    st = _state(ps, spec, rows, amps)
    eg, names = _kernels(ps, lambda: (ps.Hadamard_Int_Full("main")(st), st.export())[1])
    assert "dense_scatter_kernel" in names and "dense_emit_kernel" in names
    assert len(eg[2]) == 4
    oracle.System.clear()
    so = _state(oracle, spec, rows, amps)
    oracle.Hadamard_Int_Full("main")(so)
    assert_same_state(eg, so.export(), what="concentrating transform")


@pytest.mark.parametrize("full", [True, False])
def test_expansion_without_butterflies_in_canonical_order(ps, oracle, full):
    """Groups of one entry (Hadamard_Int: direct expansion) or of a few
    (Hadamard_Int_Full: sparse sums) need no butterflies: their outputs are
    written straight in canonical order.  Rows come back sorted and equal the
    oracle's."""
    spec = [("main", 12), ("anc", 3)]
    rng = np.random.default_rng(10)
    rows = []
    for a in range(8):
        for m in rng.choice(1 << 12, size=1 if not full else 1 + a % 3, replace=False):
            rows.append([int(m), a])
    rows = np.array(rows, dtype=np.uint64)
    amps = rng.standard_normal(len(rows)) + 1j * rng.standard_normal(len(rows))
    amps /= np.linalg.norm(amps)

    def op(mod, st):
        if full:
            mod.Hadamard_Int_Full("main")(st)
        else:
            mod.Hadamard_Int("main", 12)(st)

    ps.System.clear()
    st = _state(ps, spec, rows, amps)
    eg, names = _kernels(ps, lambda: (op(ps, st), st.export())[1])
    assert "expand_init_kernel" in names and "tx_emit_kernel" not in names and "wht_first_pass_kernel" not in names
    oracle.System.clear()
    so = _state(oracle, spec, rows, amps)
    op(oracle, so)
    assert_same_state(eg, so.export(), what=f"full={full}")
    v = np.asarray(eg[1], dtype=np.uint64)
    order = np.lexsort(tuple(v[:, k] for k in reversed(range(v.shape[1]))))
    assert np.array_equal(order, np.arange(len(order))), "output not in canonical order"


@pytest.mark.parametrize("fill", [0.9, 0.05])
def test_canonical_output_switch_keeps_the_state(ps, fill):
    """set_canonical_output(False) keeps the group-major order; the state is the
    same entry for entry (amplitudes bit-identical)."""
    spec = [("main", 12), ("anc", 3)]
    rng = np.random.default_rng(12)
    vals, amps = _random_rows(rng, spec, 1 << 14, fill=fill)
    out = []
    for on in (True, False):
        ps.set_canonical_output(on)
        try:
            ps.System.clear()
            st = _state(ps, spec, vals, amps)
            ps.Hadamard_Int_Full("main")(st)
            ps.Hadamard_Int("anc", 3)(st)
            out.append(st.export())
        finally:
            ps.set_canonical_output(True)
    assert_same_state(out[0], out[1], exact_amps=True, what=f"canonical on/off, fill {fill}")
