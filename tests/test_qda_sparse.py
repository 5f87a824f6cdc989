"""QDA over the sparse block encoding composed from the CKS oracles: the block
encoding against the matrix it encodes, the engine against the oracle, and the
solve against conjugate gradients."""
from __future__ import annotations

import numpy as np
import pytest

from parity import assert_same_state
from sparqsim_amd import cks, qda_sparse


def _encoded_column(mod, nq, i):
    st, walk, enc, A, kappa = qda_sparse.setup(mod, nq, 30.0, 8)
    st = walk.CreateSys()
    mod.Init_Unsafe(walk.j, i)(st)
    enc.apply(st)
    n = 1 << nq
    zero = [walk.b1, walk.k, walk.b2, walk.j_comp, walk.k_comp, walk.data_offset]
    out, _ = cks.select_dense(mod, st, zero, [walk.j], n)
    q = np.array(enc.matrix.elements, dtype=float).reshape(n, 8)
    cols = np.array(enc.matrix.sparsity).reshape(n, 8)
    ref = np.zeros(n)
    for r in range(n):
        for s, c in enumerate(cols[r]):
            if c == i:
                ref[r] += q[r, s] / (8 * (2.0 ** 32 - 1))
    return out, ref


def test_sparse_block_encoding_extracts_the_matrix(oracle):
    # <r, 0| T^dag S T |i, 0> = A_ri / (nnz (2^ds - 1)), up to the 32-bit rotation angle
    # (cos(2 pi theta_q) differs from sqrt(v/Amax) by at most 2 pi 2^-32 relative)
    for i in (0, 7, 31):
        out, ref = _encoded_column(oracle, 5, i)
        assert np.max(np.abs(out - ref)) < 1e-9
        assert np.max(np.abs(out.imag)) < 1e-9


def test_sparse_qda_short_schedule_converges(oracle):
    met, *_ = qda_sparse.solve(oracle, 5, 200)
    assert met["fidelity"] > 0.99 and met["kappa_encoded"] == pytest.approx(100.6, rel=1e-2)


@pytest.mark.gpu
def test_engine_sparse_block_encoding_matches(ps, oracle):
    for i in (0, 7, 31):
        g, ref = _encoded_column(ps, 5, i)
        o, _ = _encoded_column(oracle, 5, i)
        assert np.max(np.abs(g - o)) <= 1e-12 * np.max(np.abs(o))
        assert np.max(np.abs(g - ref)) < 1e-9


@pytest.mark.gpu
def test_engine_sparse_qda_steps_match_oracle(ps, oracle):
    outs = []
    for mod in (ps, oracle):
        met, st, walk, blk = qda_sparse.solve(mod, 5, 2000, stop_after=6)
        outs.append(st.export())
    assert_same_state(outs[0], outs[1], what="6 QDA steps")


@pytest.mark.gpu
def test_engine_sparse_qda_solve_converges(ps):
    met, *_ = qda_sparse.solve(ps, 5, 2000)
    assert met["fidelity"] > 0.9999 and met["res_dir"] < 5e-4
