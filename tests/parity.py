"""Parity helpers: run one circuit on the HIP engine and on the CPU oracle and
compare the states after canonical sort (the reference's CudaBenchmark pattern,
test/GPUTest/Benchmark/CudaBenchmark.cu:40-74): register values bit-exact,
amplitudes within a stated tolerance."""
from __future__ import annotations

import itertools

import numpy as np

# Amplitude tolerance of the parity tests: |a_gpu - a_ref| <= RTOL * max(|a_ref|, scale)
# where scale = max |a_ref| over the state (1e-12 relative, FP64).
RTOL = 1e-12


def canonical(export):
    ids, vals, amps = export
    vals = np.asarray(vals, dtype=np.uint64)
    amps = np.asarray(amps, dtype=np.complex128)
    if vals.shape[0] == 0:
        return list(ids), vals, amps
    order = np.lexsort(tuple(vals[:, k] for k in reversed(range(vals.shape[1])))) if vals.shape[1] else np.arange(len(amps))
    return list(ids), vals[order], amps[order]


def assert_same_state(exp_gpu, exp_ref, rtol=RTOL, exact_amps=False, what=""):
    ids_g, v_g, a_g = canonical(exp_gpu)
    ids_r, v_r, a_r = canonical(exp_ref)
    assert ids_g == ids_r, f"{what}: register ids differ {ids_g} vs {ids_r}"
    assert v_g.shape == v_r.shape, f"{what}: entry count {v_g.shape[0]} vs {v_r.shape[0]}"
    if v_g.size:
        bad = np.nonzero(np.any(v_g != v_r, axis=1))[0]
        assert bad.size == 0, f"{what}: keys differ at canonical row {bad[0]}: {v_g[bad[0]]} vs {v_r[bad[0]]}"
    if a_r.size == 0:
        return
    if exact_amps:
        same = (a_g.real == a_r.real) & (a_g.imag == a_r.imag)
        assert same.all(), f"{what}: amplitudes not bit-identical, first {np.argmin(same)}: {a_g[~same][0]} vs {a_r[~same][0]}"
        return
    scale = np.max(np.abs(a_r))
    err = np.abs(a_g - a_r)
    tol = rtol * np.maximum(np.abs(a_r), scale)
    worst = int(np.argmax(err - tol))
    assert np.all(err <= tol), f"{what}: amplitude error {err[worst]:.3e} > {tol[worst]:.3e} at row {worst}"


def random_state(mod, n, seed, amp_scale=1.0):
    """Random superposition over the active registers: n distinct keys, N(0,1)+iN(0,1) amplitudes."""
    rng = np.random.default_rng(seed)
    ids = [i for i in range(mod.System.register_count()) if mod.System.status_of(i)]
    widths = [mod.System.size_of(i) for i in ids]
    total = sum(widths)
    n = min(n, 1 << min(total, 62))
    seen = set()
    rows = []
    while len(rows) < n:
        row = tuple(int(rng.integers(0, 1 << w, dtype=np.uint64)) if w < 64 else int(rng.integers(0, 2**63, dtype=np.uint64)) for w in widths)
        if row not in seen:
            seen.add(row)
            rows.append(row)
    vals = np.array(rows, dtype=np.uint64).reshape(n, len(ids))
    amps = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * amp_scale
    return vals, amps


def run_both(ps, oracle, build, seed=0):
    """build(mod) -> state; returns (gpu_export, oracle_export)."""
    ps.System.clear()
    oracle.System.clear()
    sg = build(ps)
    eg = sg.export()
    so = build(oracle)
    eo = so.export()
    return eg, eo


# ---- condition evaluation (global_macros.h:13-17, 82-86): an operator carrying
# conditions acts only on entries where every listed predicate holds -- register
# nonzero, all bits set, a given bit set, equal to a given value -- and leaves the
# others untouched.  Brute force over every value of three small registers, with
# a flip of a fourth as the conditioned operator.
def check_condition_truth_tables(mod):
    spec = [("c", "UnsignedInteger", 3), ("d", "UnsignedInteger", 2), ("s", "SignedInteger", 2), ("t", "Boolean", 1)]
    combos = list(itertools.product(range(8), range(4), range(4)))
    cases = [
        (lambda op: op.conditioned_by_value("c", 5), lambda c, d, s: c == 5),
        (lambda op: op.conditioned_by_value([("c", 3), ("d", 2)]), lambda c, d, s: c == 3 and d == 2),
        (lambda op: op.conditioned_by_nonzeros("d"), lambda c, d, s: d != 0),
        (lambda op: op.conditioned_by_nonzeros(["s", "c"]), lambda c, d, s: s != 0 and c != 0),
        (lambda op: op.conditioned_by_all_ones("c"), lambda c, d, s: c == 7),
        (lambda op: op.conditioned_by_all_ones(["d", "s"]), lambda c, d, s: d == 3 and s == 3),
        (lambda op: op.conditioned_by_bit("c", 1), lambda c, d, s: (c >> 1) & 1 == 1),
        (lambda op: op.conditioned_by_bit([("c", 0), ("d", 1)]), lambda c, d, s: c & 1 and d & 2),
        (lambda op: op.conditioned_by_bit("c", 2).conditioned_by_nonzeros("s"), lambda c, d, s: c >= 4 and s != 0),
        (lambda op: op.conditioned_by_value("d", 1).conditioned_by_all_ones("s").conditioned_by_bit("c", 0),
         lambda c, d, s: d == 1 and s == 3 and c & 1),
        (lambda op: op.conditioned_by_value("c", 6).conditioned_by_value("c", 2), lambda c, d, s: c == 2),   # replaces
    ]
    for k, (cond, pred) in enumerate(cases):
        mod.System.clear()
        st, _ = _regs(mod, spec)
        vals = np.array([[c, d, s, 0] for c, d, s in combos], dtype=np.uint64)
        amps = np.arange(1, len(combos) + 1, dtype=complex) * (1 + 0.5j)
        st.load(vals, amps)
        cond(mod.X_Bool("t", 0))(st)
        got = _amp_map(st)
        want = {(c, d, s, int(bool(pred(c, d, s)))): a for (c, d, s), a in zip(combos, amps)}
        assert got.keys() == want.keys(), ("condition case", k)
        assert all(got[key] == want[key] for key in want), ("condition case", k)


def _regs(mod, spec):
    st = mod.SparseState()
    ids = [mod.AddRegister(n, getattr(mod, t), w)(st) for n, t, w in spec]
    return st, ids


def _amp_map(st):
    ids, vals, amps = st.export()
    return {tuple(int(x) for x in v): a for v, a in zip(vals, amps)}
