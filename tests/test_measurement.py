"""MeasureZ / Reset / Probability (SparQ/include/measurement.h, SparQ/src/measurement.cpp):
the oracle against the Born rule and the definitions, the engine against the oracle
for the same seed."""
from __future__ import annotations

import math

import numpy as np
import pytest

from parity import assert_same_state, random_state


def _state(mod, seed=3, n=40):
    mod.System.clear()
    mod.System.add_register("a", mod.UnsignedInteger, 3)
    mod.System.add_register("b", mod.UnsignedInteger, 4)
    mod.System.add_register("f", mod.Boolean, 1)
    vals, amps = random_state(mod, n, seed)
    amps = amps / np.linalg.norm(amps)
    st = mod.SparseState()
    st.load(vals, amps)
    return st, vals, amps


def test_oracle_born_rule_frequencies(oracle):
    st0, vals, amps = _state(oracle)
    p_a = np.zeros(8)
    for v, a in zip(vals[:, 0], amps):
        p_a[int(v)] += abs(a) ** 2
    counts = np.zeros(8)
    trials = 4000
    for t in range(trials):
        oracle.set_seed(1000 + t)
        st = st0.clone()
        (out,), prob = oracle.MeasureZ("a")(st)
        counts[out] += 1
        assert prob == pytest.approx(p_a[out], rel=1e-12)
    sigma = np.sqrt(trials * p_a * (1 - p_a))
    assert np.all(np.abs(counts - trials * p_a) <= 5 * sigma + 1)


def test_oracle_collapse_reset_probability(oracle):
    st, vals, amps = _state(oracle)
    oracle.set_seed(7)
    (oa, ob), prob = oracle.MeasureZ(["a", "b"])(st)
    sel = (vals[:, 0] == oa) & (vals[:, 1] == ob)
    assert prob == pytest.approx(float(np.sum(np.abs(amps[sel]) ** 2)), rel=1e-12)
    _, v, a = st.export()
    assert np.all(v[:, 0] == oa) and np.all(v[:, 1] == ob)
    assert float(np.sum(np.abs(a) ** 2)) == pytest.approx(1.0, abs=1e-12)
    st, vals, amps = _state(oracle)
    p = oracle.Probability({"a": 2, "f": 1})(st)
    assert p == pytest.approx(float(np.sum(np.abs(amps[(vals[:, 0] == 2) & (vals[:, 2] == 1)]) ** 2)), rel=1e-12)
    dist = oracle.Probability.distribution(st, "b")
    assert sum(dist.values()) == pytest.approx(1.0, abs=1e-12)
    for k, pk in dist.items():
        assert pk == pytest.approx(float(np.sum(np.abs(amps[vals[:, 1] == k]) ** 2)), rel=1e-12)
    oracle.set_seed(9)
    measured = oracle.Reset(["b"], [5])(st)
    _, v, a = st.export()
    assert np.all(v[:, 1] == 5) and len(measured) == 1
    assert float(np.sum(np.abs(a) ** 2)) == pytest.approx(1.0, abs=1e-12)


@pytest.mark.gpu
def test_engine_measurement_matches_oracle(ps, oracle):
    for seed in (1, 2, 3, 4, 5):
        outs = []
        for mod in (ps, oracle):
            st, _, _ = _state(mod, seed=seed, n=300)
            mod.set_seed(100 + seed)
            res = mod.MeasureZ(["a", "f"])(st)
            outs.append((list(res[0]), res[1], st.export()))
        assert outs[0][0] == outs[1][0]
        assert outs[0][1] == pytest.approx(outs[1][1], rel=1e-12)
        assert_same_state(outs[0][2], outs[1][2], what=f"after MeasureZ seed {seed}")
    outs = []
    for mod in (ps, oracle):
        st, _, _ = _state(mod, seed=11, n=300)
        mod.set_seed(42)
        m = mod.Reset(["b"], [3])(st)
        outs.append((list(m), st.export(), mod.Probability({"a": 1})(st), mod.Probability.distribution(st, "a")))
    assert outs[0][0] == outs[1][0]
    assert_same_state(outs[0][1], outs[1][1], what="after Reset")
    assert outs[0][2] == pytest.approx(outs[1][2], rel=1e-12)
    assert set(outs[0][3]) == set(outs[1][3])
    for k in outs[1][3]:
        assert outs[0][3][k] == pytest.approx(outs[1][3][k], rel=1e-12)


@pytest.mark.gpu
def test_engine_measurement_errors(ps):
    st, _, _ = _state(ps)
    with pytest.raises(ValueError):
        ps.Probability("a", 9)          # 9 does not fit 3 bits
    with pytest.raises(ValueError):
        ps.MeasureZ(["a", "a"])
    ps.GlobalPhase(2.0)(st)
    with pytest.raises(RuntimeError):
        ps.MeasureZ("a")(st)             # not normalised
    empty = ps.SparseState()
    ps.PartialTraceSelect({"a": 7, "b": 15, "f": 1})(empty)
    if empty.size() == 0:
        with pytest.raises(ValueError):
            ps.MeasureZ("a")(empty)


# ---------------------------------------------------------------- text state format
def test_state_file_round_trip_and_format(oracle, tmp_path):
    from sparqsim_amd import state_io

    st, vals, amps = _state(oracle, seed=5, n=30)
    lines = state_io.state_to_lines(oracle, st, 16)
    # System::to_string(16) (basic_components.cpp:497-515): "{: 1.16f} {:+1.16f}i" then |name:value>
    ids, v, a = st.export()
    first = f"{a[0].real: 1.16f} {a[0].imag:+1.16f}i" + "".join(
        f"|{oracle.System.name_of(i)}:{int(x)}>" for i, x in zip(ids, v[0]))
    assert lines[0] == first
    path = str(tmp_path / "state.txt")
    state_io.print_state_to_file(oracle, st, path, 16)
    assert open(path).readline() == "Print State To File:\n"
    back = state_io.load_state_file(oracle, path, check_norm=1e-12)
    assert_same_state(back.export(), st.export(), rtol=1e-15, what="text round trip")


def test_state_load_qda_registers(oracle, tmp_path):
    from sparqsim_amd import qda, state_io

    st = qda.setup_poiseuille(oracle, 3)
    qda.WalkSTridiagonal(oracle, "main_reg", "anc_UA", "anc_1", "anc_2", "anc_3", "anc_4", 0.3, 10.0, 1.3,
                         2.2666666666564934, -1.0)(st)
    path = str(tmp_path / "qda_state.txt")
    state_io.print_state_to_file(oracle, st, path, 17)
    back = state_io.StateLoad(oracle)(path)
    assert_same_state(back.export(), st.export(), rtol=1e-15, what="StateLoad")
    with pytest.raises(FileNotFoundError):
        state_io.StateLoad(oracle)(str(tmp_path / "missing.txt"))
    (tmp_path / "empty.txt").write_text("no entries here\n")
    with pytest.raises(ValueError):
        state_io.StateLoad(oracle)(str(tmp_path / "empty.txt"))


def test_artifact_files(tmp_path):
    from sparqsim_amd import state_io

    state_io.write_residual_curve(str(tmp_path / "residual_curve.csv"), [(10, 0.5, 0.25, 0.125, 0.9), (20, 0.1, 0.05, 0.2, 0.99)])
    text = (tmp_path / "residual_curve.csv").read_text().splitlines()
    assert text[0] == "step,res_dir,res_opt,success_prob,fidelity" and text[1] == "10,0.5,0.25,0.125,0.90000000000000002"
    state_io.write_summary_json(str(tmp_path / "qda_summary_cpp.json"), {"final_res_dir": float("nan"), "n": 3})
    import json

    assert json.load(open(tmp_path / "qda_summary_cpp.json")) == {"final_res_dir": None, "n": 3}


@pytest.mark.gpu
def test_engine_state_file_round_trip(ps, oracle, tmp_path):
    from sparqsim_amd import state_io

    st, vals, amps = _state(ps, seed=6, n=200)
    path = str(tmp_path / "engine_state.txt")
    state_io.print_state_to_file(ps, st, path, 17)
    back = state_io.load_state_file(ps, path)
    assert_same_state(back.export(), st.export(), rtol=1e-15, what="engine text round trip")
    so, _, _ = _state(oracle, seed=6, n=200)
    assert sorted(state_io.state_to_lines(ps, st)) == sorted(state_io.state_to_lines(oracle, so))
