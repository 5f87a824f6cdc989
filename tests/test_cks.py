"""CKS Chebyshev quantum-walk linear solver (SparQ_Algorithm/include/hamiltonian_simulation.h,
hamiltonian_simulation_tridiagonal.h): the oracle against the reference's published
anchors, the engine against the oracle."""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from parity import assert_same_state, random_state
from sparqsim_amd import cks

ALPHA, BETA = 2.2666666666564934, -1.0

# Published anchors, docs/verification/v1e-cks-direct-oracle-linsolver.md:
#   :63  dim16 --path both --coef stock: res_dir=3.962820e-01 res_opt=1.522807e-01
#        fid=0.95515840 p_succ=2.288431e-01 (direct == QRAM path)
#   :122 dim16 --coef true --J 640: succP=1.004941e-04, res_dir ~ 0, fid=1.0, max_system_size=312
#   :123 dim64 --coef true --J 1024: succP=5.101613e-02, res_dir <= 1e-7, fid=1.0, max_system_size=1272
#   :138 2^20: walk state after FirstStep = 20,971,512 systems
STOCK16 = {"res_dir": 3.962820e-01, "res_opt": 1.522807e-01, "fidelity": 0.95515840, "p_succ": 2.288431e-01}


def _check_stock16(r):
    assert r["j0"] == 137 and r["terms"] == 138
    for key in ("res_dir", "res_opt", "p_succ"):
        assert abs(r[key] - STOCK16[key]) <= 5e-7 * STOCK16[key], (key, r[key])
    assert abs(r["fidelity"] - STOCK16["fidelity"]) <= 5e-9
    assert r["max_system_size"] == 312


# ---------------------------------------------------------------- oracle vs published anchors
@pytest.mark.parametrize("path", ["direct", "qram"])
def test_oracle_reproduces_published_stock_anchors(oracle, path):
    _check_stock16(cks.run_probe(oracle, nqubit=4, path=path, coef="stock"))


def test_oracle_fitted_coefficient_anchors(oracle):
    # The fitted 1/x coefficients are a rank-truncated least-squares solution; the
    # truncation depends on the SVD routine (Eigen BDCSVD in the reference, LAPACK
    # here), so p_succ = P/(sum|d|)^2 agrees to 1e-3 relative, the solution itself exactly.
    r = cks.run_probe(oracle, nqubit=4, path="direct", coef="true", J=640)
    assert r["res_dir"] <= 1e-7 and r["fidelity"] >= 1 - 1e-8
    assert abs(r["p_succ"] - 1.004941e-04) <= 1e-3 * 1.004941e-04
    assert r["max_system_size"] == 312
    r = cks.run_probe(oracle, nqubit=6, path="direct", coef="true", J=1024)
    assert r["res_dir"] <= 1e-7 and r["fidelity"] >= 1 - 1e-8
    assert abs(r["p_succ"] - 5.101613e-02) <= 1e-3 * 5.101613e-02
    assert r["max_system_size"] == 1272


def _walk_states(mod, path, nqubit=3, steps=4):
    mod.System.clear()
    mat, spec = mod.make_tridiagonal_matrix_spec(ALPHA, BETA, 1 << nqubit, 32)
    w = cks.QuantumWalkNSteps(mod, mat, spec if path == "direct" else None)
    w.InitEnvironment()
    st = w.CreateSys()
    mod.Hadamard_Int(w.j, w.get_init_size())(st)
    mod.ClearZero()(st)
    out = []
    for step in range(steps):
        if step == 0:
            w.FirstStep(st)
        else:
            w.StepImplOneStep(st)
        out.append(st.export())
    return out


def test_direct_oracles_mirror_qram_walk_and_keep_norm(oracle):
    # v1e:37-59: the direct oracles reproduce the QRAM composite register maps bit
    # for bit (same entries after every step) and the walk keeps the norm at 1.
    direct = _walk_states(oracle, "direct")
    qram = _walk_states(oracle, "qram")
    for k, (d, q) in enumerate(zip(direct, qram)):
        assert_same_state(d, q, exact_amps=True, what=f"step {k}")
        assert abs(float(np.sum(np.abs(d[2]) ** 2)) - 1.0) < 1e-12


def test_chebyshev_coefficients_closed_form(oracle):
    # b <= 100: coef(j) = 4 sum_{i=j+1}^{b} binom(2b, b+i) / 4^b (hamiltonian_simulation.cpp:44-88)
    for b in (1, 10, 37, 100):
        cheb = cks.ChebyshevPolynomialCoefficient(b)
        ref = oracle.ChebyshevPolynomialCoefficient(b)
        for j in range(0, b + 1, max(1, b // 7)):
            exact = Fraction(4 * sum(math.comb(2 * b, b + i) for i in range(j + 1, b + 1)), 4 ** b)
            assert abs(cheb.coef(j) - float(exact)) <= 1e-12 * max(float(exact), 1e-300)
            assert cheb.coef(j) == ref.coef(j)
    # b > 100: 2 erfc((j + 1/2)/sqrt(b))
    cheb = cks.ChebyshevPolynomialCoefficient(1436)
    for j in (0, 5, 137):
        assert cheb.coef(j) == math.erfc((j + 0.5) / math.sqrt(1436)) * 2
        assert cheb.coef(j) == oracle.ChebyshevPolynomialCoefficient(1436).coef(j)
    # PySparQ/test/algorithms/test_cks_solver.py:29-66: b kept, alternating signs, odd steps, decreasing
    cheb = cks.ChebyshevPolynomialCoefficient(10)
    assert cheb.b == 10
    assert [cheb.sign(j) for j in range(4)] == [False, True, False, True]
    assert [cheb.step(j) for j in range(4)] == [1, 3, 5, 7]
    assert all(cheb.coef(j + 1) <= cheb.coef(j) + 1e-10 for j in range(10))
    assert cks.kappa_schedule(14.074907851087548, 0.01) == (1436, 137)


def test_fitted_coefficients_approximate_inverse():
    # sum_j d_j T_(2j+1)(x) ~ 1/x on [m, M] (hamiltonian_simulation_tridiagonal.h:54-100),
    # evaluated independently by numpy's Clenshaw recurrence
    m, M = 1.0 / 64, 0.25
    d, a = cks.chebyshev_inverse_coefficients(m, M, 640)
    x = np.linspace(m, M, 997)
    c = np.zeros(2 * len(d))
    c[1::2] = d
    assert np.max(np.abs(np.polynomial.chebyshev.chebval(x, c) * x - 1.0)) < 1e-7
    assert a == pytest.approx(float(np.sum(np.abs(d))))


def test_tridiagonal_memory_layout(ps, oracle):
    # The combined memory [elements | sparsity] of make_tridiagonal_matrix_spec
    # (hamiltonian_simulation_tridiagonal.h:169-225) from both implementations.
    for n, ds in ((4, 8), (16, 32), (64, 20), (1 << 10, 32)):
        mat, spec = ps.make_tridiagonal_matrix_spec(ALPHA, BETA, n, ds)
        omat, ospec = oracle.make_tridiagonal_matrix_spec(ALPHA, BETA, n, ds)
        data = np.array(mat.get_data(), dtype=np.uint64)
        assert len(data) == 8 * n and mat.get_sparsity_offset() == 4 * n == spec.sparse_offset
        assert np.array_equal(data, np.array(omat.get_data(), dtype=np.uint64))
        assert all(int(data[i]) == spec.combined_mem(i) == ospec.combined_mem(i) for i in range(0, 8 * n, max(1, n // 64)))
        # closed forms: alpha/(|alpha|+2|beta|) just below 17/32, 1/(...) just above 15/64
        assert spec.alpha_enc == (17 * 2 ** ds) // 32 - 1 and spec.beta_enc == (15 * 2 ** ds) // 64
        assert spec.k_width == max(int(math.log2(n)) + 1, ds) == ospec.k_width
        for j in (0, 1, n // 2, n - 2, n - 1):   # columns ascend inside the 4-wide window
            cols = [spec.col_of(j, s) for s in range(4)]
            assert cols == sorted(cols) and cols[0] == min(max(j - 1, 0), n - 4)


def test_binary_search_and_angle_closed_forms(oracle):
    # QuantumBinarySearch_Fast (hamiltonian_simulation.cpp:182-220) on ascending rows,
    # GetQWRotateAngle_Int_Int_Int (:257-290) against acos(sqrt(v/Amax))/2pi.
    rng = np.random.default_rng(5)
    rows, L = 8, 8
    mem = np.zeros(2 * rows * L, dtype=np.uint64)
    for r in range(rows):
        mem[64 + r * L: 64 + (r + 1) * L] = np.sort(rng.choice(200, L, replace=False))
    oracle.System.clear()
    for name, w in (("off", 8), ("tgt", 8), ("res", 8)):
        oracle.System.add_register(name, oracle.UnsignedInteger, w)
    q = oracle.QRAMCircuit_qutrit(7, 8, [int(v) for v in mem])
    st = oracle.SparseState()
    vals = []
    for r in range(rows):
        for t in list(mem[64 + r * L: 64 + (r + 1) * L][::3]) + [201, 0]:
            vals.append((64 + r * L, int(t), 0))
    st.load(np.array(vals, dtype=np.uint64), np.ones(len(vals), dtype=np.complex128))
    oracle.QuantumBinarySearch_Fast(q, "off", L, "tgt", "res")(st)
    _, v, _ = st.export()
    for off, t, res in v:   # distinct ascending row: every present target is found, a miss gives 0
        hit = np.nonzero(mem[off: off + L] == t)[0]
        assert res == (off + hit[0] if hit.size else 0)
    oracle.System.clear()
    oracle.System.add_register("d", oracle.UnsignedInteger, 12)
    oracle.System.add_register("o", oracle.Rational, 32)
    for pos in (True, False):
        mat = oracle.SparseMatrix.from_encoded([0] * 4, [0, 1, 2, 3], 12, 4, 1, pos)
        st = oracle.SparseState()
        vs = [0, 1, 100, 2047, 2048, 4095]
        st.load(np.array([(v, 0) for v in vs], dtype=np.uint64), np.ones(len(vs), dtype=np.complex128))
        oracle.GetQWRotateAngle_Int_Int_Int("d", "d", "d", "o", mat)(st)
        _, v, _ = st.export()
        for d, o in v.astype(object):
            if pos:
                ratio = d / 4095
            else:
                sv = d - 4096 if d >= 2048 else d
                ratio = min(1.0, abs(sv) / 2047)
            expect = math.acos(math.sqrt(ratio)) / math.pi / 2
            assert abs(int(o) / 2 ** 32 - expect) <= 2 ** -32


def test_batched_rhs_equals_single_rhs(oracle):
    rng = np.random.default_rng(3)
    rhs = np.stack([np.ones(16), rng.standard_normal(16), rng.standard_normal(16)])
    kappa = 14.074907851087548
    batch = cks.BatchedPoiseuilleCKS(oracle, 4, rhs, kappa=kappa, coef="stock")
    batch.prepare_input()
    assert batch.lcu.step_state.size() == 3 * 312
    while batch.Step():
        pass
    blocks, res = batch.extract()
    for key in ("res_dir", "res_opt", "p_succ"):
        assert abs(res[0][key] - STOCK16[key]) <= 5e-7 * STOCK16[key]
    for r in (1, 2):
        one = cks.BatchedPoiseuilleCKS(oracle, 4, rhs[r:r + 1], kappa=kappa, coef="stock")
        one.prepare_input()
        while one.Step():
            pass
        b1, r1 = one.extract()
        x = blocks[r] / np.linalg.norm(blocks[r])
        y = b1[0] / np.linalg.norm(b1[0])
        assert np.max(np.abs(x - y)) < 1e-12
        assert abs(res[r]["p_succ"] - r1[0]["p_succ"]) <= 1e-12 * r1[0]["p_succ"]


def test_partial_trace_select_semantics(oracle):
    # partial_trace.cpp:148-225: keep matching entries in order, renormalise, return 1/sqrt(P)
    oracle.System.clear()
    oracle.System.add_register("a", oracle.UnsignedInteger, 3)
    oracle.System.add_register("b", oracle.UnsignedInteger, 3)
    st = oracle.SparseState()
    vals = np.array([(a, b) for a in range(8) for b in range(8)], dtype=np.uint64)
    amps = np.arange(1, 65, dtype=np.float64).astype(np.complex128)
    st.load(vals, amps)
    f = oracle.PartialTraceSelect({"a": 2})(st)
    p = float(np.sum(np.abs(amps[16:24]) ** 2))
    assert f == pytest.approx(1 / math.sqrt(p), rel=1e-15)
    _, v, a = st.export()
    assert np.array_equal(v[:, 1], np.arange(8)) and np.allclose(a, amps[16:24] / math.sqrt(p), rtol=1e-15)
    f2 = oracle.PartialTraceSelectRange("b", (3, 5))(st)
    assert st.size() == 3 and f2 == pytest.approx(1 / math.sqrt(float(np.sum(np.abs(a[3:6]) ** 2))), rel=1e-15)
    assert oracle.PartialTraceSelect({"a": 7})(st) == math.inf and st.size() == 0
    assert oracle.PartialTraceSelectRange("b", (0, 7))(st) == 0


# ---------------------------------------------------------------- engine vs oracle
def _setup_regs(mod, regs):
    mod.System.clear()
    for name, t, w in regs:
        mod.System.add_register(name, getattr(mod, t), w)


def _run_pair(ps, oracle, regs, vals, amps, apply):
    outs = []
    for mod in (ps, oracle):
        _setup_regs(mod, regs)
        st = mod.SparseState()
        st.load(vals, amps)
        apply(mod, st)
        outs.append(st.export())
    return outs


@pytest.mark.gpu
def test_walk_primitives_match_oracle(ps, oracle):
    rng = np.random.default_rng(11)
    n = 64
    # "addr" is wide enough for every address the XOR-affine ops form (values are
    # stored masked to the register width; the reference keeps over-wide values)
    regs = [("off", "UnsignedInteger", 12), ("row", "UnsignedInteger", 7), ("k", "UnsignedInteger", 12),
            ("sr", "UnsignedInteger", 12), ("d", "UnsignedInteger", 32), ("rot", "Rational", 32),
            ("addr", "UnsignedInteger", 14)]
    N = 3000
    vals = np.zeros((N, 7), dtype=np.uint64)
    vals[:, 0] = rng.choice([0, 4 * n, 17], N)
    vals[:, 1] = rng.integers(0, n, N)
    vals[:, 2] = rng.integers(0, 1 << 12, N)
    vals[:, 2][: N // 2] = rng.integers(0, n, N // 2)          # half hit the window
    vals[:, 3] = rng.integers(0, 1 << 12, N) * (rng.random(N) < 0.3)
    vals[:, 4] = rng.integers(0, 1 << 32, N, dtype=np.uint64)
    vals[:, 5] = rng.integers(0, 1 << 32, N, dtype=np.uint64)
    vals[:, 6] = rng.integers(0, 1 << 14, N)
    vals = np.unique(vals, axis=0)
    amps = rng.standard_normal(len(vals)) + 1j * rng.standard_normal(len(vals))
    cases = {
        "GetRowAddr": lambda m, s: m.GetRowAddr("off", "row", 4, "addr")(s),
        "GetDataAddr": lambda m, s: m.GetDataAddr("off", "row", "k", 4, "addr")(s),
        "TridiagOracle1": lambda m, s: m.TridiagOracle1("off", "row", "k", "d", m.make_tridiagonal_matrix_spec(ALPHA, BETA, n, 12)[1])(s),
        "TridiagOracle2": lambda m, s: m.TridiagOracle2("off", "row", "k", "sr", m.make_tridiagonal_matrix_spec(ALPHA, BETA, n, 12)[1])(s),
        "TridiagOracle2.dag": lambda m, s: m.TridiagOracle2("off", "row", "k", "sr", m.make_tridiagonal_matrix_spec(ALPHA, BETA, n, 12)[1]).dag(s),
        "TridiagParityPhase": lambda m, s: m.TridiagParityPhase("k")(s),
        "GetQWRotateAngle(+)": lambda m, s: m.GetQWRotateAngle_Int_Int_Int("d", "row", "k", "rot", m.SparseMatrix.from_encoded([0] * 4, [0, 1, 2, 3], 32, 4, 1, True))(s),
        "GetQWRotateAngle(+-)": lambda m, s: m.GetQWRotateAngle_Int_Int_Int("d", "row", "k", "rot", m.SparseMatrix.from_encoded([0] * 4, [0, 1, 2, 3], 32, 4, 1, False))(s),
    }
    for what, fn in cases.items():
        g, o = _run_pair(ps, oracle, regs, vals, amps, fn)
        assert_same_state(g, o, exact_amps=True, what=what)


@pytest.mark.gpu
def test_binary_search_matches_oracle(ps, oracle):
    rng = np.random.default_rng(12)
    rows, L = 32, 4
    mem = np.zeros(2 * rows * L, dtype=np.uint64)
    mem[: rows * L] = rng.integers(0, 1 << 20, rows * L)
    for r in range(rows):
        mem[rows * L + r * L: rows * L + (r + 1) * L] = np.sort(rng.choice(rows, L, replace=False))
    regs = [("off", "UnsignedInteger", 9), ("tgt", "UnsignedInteger", 9), ("res", "UnsignedInteger", 9)]
    vals = np.array([(rows * L + r * L, t, z) for r in range(rows) for t in range(rows + 3) for z in (0, 5)], dtype=np.uint64)
    amps = np.ones(len(vals), dtype=np.complex128)
    fn = lambda m, s: m.QuantumBinarySearch_Fast(m.QRAMCircuit_qutrit(8, 20, [int(v) for v in mem]), "off", L, "tgt", "res")(s)
    g, o = _run_pair(ps, oracle, regs, vals, amps, fn)
    assert_same_state(g, o, exact_amps=True, what="QuantumBinarySearch_Fast")


@pytest.mark.gpu
def test_partial_trace_select_matches_oracle(ps, oracle):
    regs = [("a", "UnsignedInteger", 3), ("b", "UnsignedInteger", 5), ("c", "Boolean", 1)]
    vals, amps = random_state(_setup(ps, regs), 200, seed=4)
    for sel in ({"a": 2}, {"a": 2, "c": 1}, {"a": 9}):
        outs, facs = [], []
        for mod in (ps, oracle):
            _setup_regs(mod, regs)
            st = mod.SparseState()
            st.load(vals, amps)
            facs.append(mod.PartialTraceSelect(dict(sel))(st))
            facs.append(mod.PartialTraceSelectRange("b", (3, 20))(st))
            outs.append(st.export())
        assert_same_state(outs[0], outs[1], what=str(sel))
        for fg, fo in ((facs[0], facs[2]), (facs[1], facs[3])):
            assert (fg == fo) or abs(fg - fo) <= 1e-12 * abs(fo)


def _setup(mod, regs):
    _setup_regs(mod, regs)
    return mod


@pytest.mark.gpu
def test_debug_checks(ps):
    ps.set_debug_checks(True)
    try:
        ps.System.add_register("q", ps.UnsignedInteger, 4)
        st = ps.SparseState()
        ps.Hadamard_Int("q", 4)(st)
        ps.CheckNormalization()(st)
        ps.CheckNan()(st)
        ps.GlobalPhase(2.0)(st)
        with pytest.raises(RuntimeError):
            ps.CheckNormalization()(st)
    finally:
        ps.set_debug_checks(False)


@pytest.mark.gpu
@pytest.mark.parametrize("path", ["direct", "qram"])
def test_engine_reproduces_published_stock_anchors(ps, oracle, path):
    ps.set_debug_checks(True)
    try:
        r = cks.run_probe(ps, nqubit=4, path=path, coef="stock")
    finally:
        ps.set_debug_checks(False)
    _check_stock16(r)
    ref = cks.run_probe(oracle, nqubit=4, path=path, coef="stock")
    for key in ("res_dir", "res_opt", "fidelity", "p_succ"):
        assert abs(r[key] - ref[key]) <= 1e-12 * max(abs(ref[key]), 1e-3), key


@pytest.mark.gpu
def test_engine_walk_states_match_oracle(ps, oracle):
    for path in ("direct", "qram"):
        g = _walk_states(ps, path, nqubit=4, steps=5)
        o = _walk_states(oracle, path, nqubit=4, steps=5)
        for k, (eg, eo) in enumerate(zip(g, o)):
            assert_same_state(eg, eo, what=f"{path} step {k}")


@pytest.mark.gpu
def test_engine_fitted_coefficients_dim64(ps, oracle):
    r = cks.run_probe(ps, nqubit=6, path="direct", coef="true", J=1024)
    ref = cks.run_probe(oracle, nqubit=6, path="direct", coef="true", J=1024)
    assert r["max_system_size"] == 1272 and r["res_dir"] <= 1e-7 and r["fidelity"] >= 1 - 1e-8
    assert abs(r["p_succ"] - ref["p_succ"]) <= 1e-10 * ref["p_succ"]


@pytest.mark.gpu
def test_engine_batched_rhs_matches_oracle(ps, oracle):
    rng = np.random.default_rng(21)
    rhs = np.stack([np.ones(32)] + [rng.standard_normal(32) for _ in range(3)])
    res = []
    for mod in (ps, oracle):
        b = cks.BatchedPoiseuilleCKS(mod, 5, rhs, kappa=16.0, coef="stock")
        b.prepare_input()
        for _ in range(12):
            b.Step()
        res.append(b.extract())
    for r in range(4):
        assert np.max(np.abs(res[0][0][r] - res[1][0][r])) <= 1e-12 * np.max(np.abs(res[1][0][r]))
        assert abs(res[0][1][r]["p_succ"] - res[1][1][r]["p_succ"]) <= 1e-12 * res[1][1][r]["p_succ"]


@pytest.mark.gpu
def test_engine_2p20_state_after_first_step(ps):
    # v1e:138: the 2^20 walk holds 20,971,512 systems after FirstStep.
    ps.System.clear()
    mat, spec = ps.make_tridiagonal_matrix_spec(ALPHA, BETA, 1 << 20, 32)
    w = cks.QuantumWalkNSteps(ps, mat, spec)
    w.InitEnvironment()
    st = w.CreateSys()
    ps.Hadamard_Int(w.j, 20)(st)
    ps.TridiagParityPhase(w.j)(st)
    ps.ClearZero()(st)
    w.FirstStep(st)
    assert st.size() == 20971512


# ---------------------------------------------------------------- band-plus-random matrix, QRAM-backed walk
def test_band_random_matrix_and_walk_spectrum():
    v, s, A = cks.band_random_matrix(64, 50.0)
    D = A.toarray()
    ev = np.linalg.eigvalsh(D)
    assert np.allclose(D, D.T) and D.min() >= 0 and D.max() == pytest.approx(1.0)
    assert ev[-1] / ev[0] == pytest.approx(50.0, rel=1e-6)
    cols = np.array(s).reshape(64, 4)
    assert all(len(set(r)) == 4 and list(r) == sorted(r) for r in cols)          # ascending, distinct slots
    assert all(np.count_nonzero(D[i]) == (3 if i in (0, 63) else 4) for i in range(64))
    # the walk block v_jk / (4 (2^32 - 1)) of the quantised matrix lies inside the fitted interval
    q = np.minimum(np.floor(np.array(v) * 2.0 ** 32), 2.0 ** 32 - 1).reshape(64, 4)
    W = np.zeros((64, 64))
    for i in range(64):
        W[i, cols[i]] = q[i] / (4 * (2.0 ** 32 - 1))
    wev = np.linalg.eigvalsh(W)
    lo, hi = cks.walk_spectrum(A, 32)
    assert lo <= wev[0] and wev[-1] <= hi and hi / lo < 50.01


def test_walk_fit_interval_bounds_coefficients():
    """The 1/x fit for the sparse walk: the interval [lo, 1] encloses the walk
    block's exact spectrum (dense eigenvalues at dim 64); on it the fit matches
    1/x to 1e-7 at every eigenvalue with sum|d| < 1e3, while the fit on the
    estimated spectrum [m, M] (M about 0.49) needs sum|d| > 1e4 for the same
    accuracy -- the factor by which every per-term error of the walk state is
    amplified in the LCU output."""
    v, s, A = cks.band_random_matrix(64, 50.0)
    cols = np.array(s).reshape(64, 4)
    q = np.minimum(np.floor(np.array(v) * 2.0 ** 32), 2.0 ** 32 - 1).reshape(64, 4)
    W = np.zeros((64, 64))
    for i in range(64):
        W[i, cols[i]] = q[i] / (4 * (2.0 ** 32 - 1))
    wev = np.linalg.eigvalsh(W)
    lo, hi = cks.walk_fit_interval(A, 32)
    assert lo <= wev[0] and hi == 1.0
    m, M = cks.walk_fit_interval(A, 32, top="spectrum")
    assert m <= wev[0] and wev[-1] <= M < 0.6

    def p(d, x):   # sum_j d_j T_(2j+1)(x), T_k(x) = cos(k arccos x)
        return sum(dj * np.cos((2 * j + 1) * np.arccos(x)) for j, dj in enumerate(d))

    d_unit, s_unit = cks.chebyshev_inverse_coefficients(lo, hi, 1024)
    d_spec, s_spec = cks.chebyshev_inverse_coefficients(m, M, 1024)
    assert np.max(np.abs(wev * p(d_unit, wev) - 1)) < 1e-7
    assert np.max(np.abs(wev * p(d_spec, wev) - 1)) < 1e-7
    assert s_unit < 1e3 < 1e4 < s_spec


def test_oracle_band_random_pruning_amplification(oracle):
    """The cause of the 2^20 band-plus-random non-convergence, reproduced at dim
    32: ClearZero / merge thresholds scaled to the mean |a|^2 of the 2^20 x 16
    right-hand-side state (1e-14 x 1.3e5) leave res_dir ~1e-2 with coefficients
    fitted on [m, M] (sum|d| ~ 3e4) and ~1e-4 with the fit on [lo, 1]."""
    rng = np.random.default_rng(7)
    rhs = np.stack([np.ones(32), rng.standard_normal(32)])
    out = {}
    try:
        for top in ("spectrum", "unit"):
            oracle.set_clearzero_default_eps(1.3e-9)
            b = cks.BatchedSparseCKS(oracle, 5, rhs, kappa=50.0, J=1024, fit_top=top)
            b.lcu.merge_eps = 1.3e-9
            b.prepare_input()
            while b.Step():
                pass
            out[top] = max(r["res_dir"] for r in b.extract()[1])
    finally:
        oracle.set_clearzero_default_eps(1e-14)
    assert out["spectrum"] > 3e-3 and out["unit"] < 1e-3


def test_oracle_band_random_solve_converges(oracle):
    rng = np.random.default_rng(8)
    b = cks.BatchedSparseCKS(oracle, 5, np.stack([np.ones(32), rng.standard_normal(32)]), kappa=50.0, J=1024)
    b.prepare_input()
    while b.Step():
        pass
    _, res = b.extract()
    for r in res:
        assert r["fidelity"] >= 1 - 1e-9 and r["res_dir"] <= 1e-4


@pytest.mark.gpu
def test_engine_band_random_walk_matches_oracle(ps, oracle):
    rng = np.random.default_rng(9)
    rhs = np.stack([np.ones(32)] + [rng.standard_normal(32) for _ in range(2)])
    out = []
    for mod in (ps, oracle):
        b = cks.BatchedSparseCKS(mod, 5, rhs, kappa=50.0, J=1024)
        b.prepare_input()
        first = b.lcu.step_state.export()
        for _ in range(20):
            b.Step()
        out.append((first, b.extract(), b.lcu.current_state.export()))
    assert_same_state(out[0][0], out[1][0], what="after FirstStep")
    assert_same_state(out[0][2], out[1][2], what="accumulated state after 20 terms")
    for r in range(3):
        assert np.max(np.abs(out[0][1][0][r] - out[1][1][0][r])) <= 1e-12 * np.max(np.abs(out[1][1][0][r]))


@pytest.mark.gpu
def test_engine_band_random_solve_converges(ps):
    rng = np.random.default_rng(10)
    b = cks.BatchedSparseCKS(ps, 6, np.stack([np.ones(64), rng.standard_normal(64)]), kappa=50.0, J=1024)
    b.prepare_input()
    while b.Step():
        pass
    _, res = b.extract()
    for r in res:
        assert r["fidelity"] >= 1 - 1e-9 and r["res_dir"] <= 1e-4
