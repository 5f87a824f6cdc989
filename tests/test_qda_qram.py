"""QDA over the QRAM block encoding (Block_Encoding_via_QRAM + State_Prep_via_QRAM)
against the reference's golden fidelity lists Neg_QRAM / Pos_QRAM
(test/CPUTest/CommonTest/CorrectnessTest_QDA_CompareList.inl:3-4, driver
:5-100, :182, :198; 371 fidelities each at dim 8, tolerance 1e-5): on the
oracle (CPU) and on the engine (GPU), plus engine-oracle parity entry for
entry after the first steps and the tree and quantisation helpers against
their definitions."""
import json
import math
import os

import numpy as np
import pytest

from parity import assert_same_state
from sparqsim_amd import qda
from sparqsim_amd import qda_qram as Q

GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "qda_fidelity_lists.json")))
TOL = GOLDEN["meta"]["tolerance"]


def run_qram_golden(mod, name, beta, max_points=None):
    ref = GOLDEN["lists"][name]
    st, qA, qb, a, b, kappa = Q.setup_qram_poiseuille(mod, 3, 2.0, beta)
    _, _, _, steps = qda.qda_schedule(3, 0.01, 2.0, beta)
    assert steps == 2 * len(ref)
    worst, k = 0.0, 0
    for n in range(steps):
        w = Q.WalkSViaQRAM(mod, qA, qb, "main_reg", "anc_UA", "anc_1", "anc_2", "anc_3", "anc_4", n / steps, kappa, 1.3,
                           Q.DATA_SIZE, Q.RATIONAL_SIZE)
        w(st)
        mod.ClearZero()(st)
        if (n + 1) % 2 == 0:
            vec, _ = qda.get_output(st, mod)
            f = qda.fidelity(qda.mid_eigenstate(a, b, w.fs), vec)
            worst = max(worst, abs(f - ref[k]))
            k += 1
            if max_points and k >= max_points:
                break
    return worst, k


def test_quantisation_and_tree():
    """Rounding halves away from zero into two's complement; the tree holds the
    squared-norm sums level by level, root first, padded with one 0 word."""
    assert Q.scale_and_convert([0.5 / 2**3, -0.5 / 2**3, 1.25, -1.25, 0.0], 3, 8) == [1, 255, 10, 246, 0]
    leaves = Q.scale_and_convert([0.5, -0.25, 0.125, 0.0], 4, 10)
    v = [Q.signed(x, 10) for x in leaves]
    assert v == [8, -4, 2, 0]
    tree = Q.make_vector_tree(leaves, 10)
    assert tree == [8 * 8 + 4 * 4 + 2 * 2, 8 * 8 + 4 * 4, 2 * 2] + leaves + [0]
    assert len(Q.make_vector_tree(Q.scale_and_convert(np.ones(64) / 8, 20, 50), 50)) == 128
    m = np.arange(16.0).reshape(4, 4)
    assert np.array_equal(Q.column_flatten(m), m.T.reshape(-1))


@pytest.mark.parametrize("name,beta", [("Neg_QRAM", -1.0), ("Pos_QRAM", 1.0)])
def test_oracle_reproduces_qram_golden_fidelities(oracle, name, beta):
    worst, k = run_qram_golden(oracle, name, beta)
    assert k == 371
    assert worst < TOL, worst


def test_oracle_qram_encoding_is_A_over_frobenius_norm(oracle):
    """<i|_col <0|_row U_A |j>_col |0>_row = A_ij / |A|_F to the quantisation's
    resolution: the encoding applied to every basis column |j>."""
    m = oracle
    n, nq = 8, 3
    a = qda.poiseuille_matrix(n, 2.0, -1.0)
    a = a / np.linalg.norm(a, "fro")
    tree = Q.make_vector_tree(Q.scale_and_convert(Q.column_flatten(a), Q.EXPONENT, Q.DATA_SIZE), Q.DATA_SIZE)
    got = np.zeros((n, n), dtype=complex)
    for j in range(n):
        m.System.clear()
        q = m.QRAMCircuit_qutrit(2 * nq + 1, Q.DATA_SIZE, tree)
        st = m.SparseState()
        m.AddRegister("col", m.UnsignedInteger, nq)(st)
        m.AddRegister("row", m.UnsignedInteger, nq)(st)
        m.Init_Unsafe("col", j)(st)
        Q.BlockEncodingViaQRAM(m, q, "col", "row", Q.DATA_SIZE, Q.RATIONAL_SIZE).apply(st)
        ids, vals, amps = st.export()
        for v, amp in zip(vals, amps):
            if v[1] == 0:
                got[v[0], j] += amp
    assert np.max(np.abs(got - a)) < 1e-5, np.max(np.abs(got - a))


@pytest.mark.gpu
def test_engine_qram_walk_matches_oracle(ps, oracle):
    """Every entry of the state after each of the first 6 walk steps."""
    states = {}
    for mod in (ps, oracle):
        st, qA, qb, a, b, kappa = Q.setup_qram_poiseuille(mod, 3, 2.0, -1.0)
        _, _, _, steps = qda.qda_schedule(3, 0.01, 2.0, -1.0)
        out = []
        for n in range(6):
            Q.WalkSViaQRAM(mod, qA, qb, "main_reg", "anc_UA", "anc_1", "anc_2", "anc_3", "anc_4", n / steps, kappa, 1.3,
                           Q.DATA_SIZE, Q.RATIONAL_SIZE)(st)
            mod.ClearZero()(st)
            out.append(st.export())
        states[mod.__name__] = out
    for n, (eg, eo) in enumerate(zip(states["sparqsim_amd"], states["_oracle"])):
        assert_same_state(eg, eo, what=f"QRAM walk step {n}")


@pytest.mark.gpu
@pytest.mark.parametrize("name,beta", [("Neg_QRAM", -1.0), ("Pos_QRAM", 1.0)])
def test_engine_reproduces_qram_golden_fidelities(ps, name, beta):
    worst, k = run_qram_golden(ps, name, beta)
    assert k == 371
    assert worst < TOL, worst
