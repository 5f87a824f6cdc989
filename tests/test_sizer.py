"""The learned output sizer.

CPU: the numpy statement of the network (oracle/sizer_mlp.py) is pinned by
finite differences; the packed-parameter count agrees with the C header and
the committed weights.  GPU: the HIP forward and backward kernels against the
numpy statement; GPU training lowers the loss; the learned path leaves every
state bit-identical to the exact count-then-allocate path, including when the
network under-predicts (forced) and the exact path reruns; and an ablation on
fresh logs: what the trained network saves over sizing without it.
"""
import math
import os
import re

import numpy as np
import pytest

import sizer_mlp as ref
from parity import assert_same_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- CPU
def test_numpy_network_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.standard_normal(ref.P) * 0.2
    x = rng.uniform(0, 1, size=(12, ref.F))
    y = rng.uniform(-0.5, 0.5, size=12)
    loss, g = ref.grad(p, x, y, alpha=4.0)
    assert abs(loss - ref.loss(p, x, y, 4.0)) < 1e-12
    idx = rng.choice(ref.P, size=60, replace=False)
    idx = np.concatenate([idx, [ref.P - 1, ref.P - 2]])   # b4 and the last W4 weight
    h = 1e-6
    for i in idx:
        e = np.zeros(ref.P)
        e[i] = h
        fd = (ref.loss(p + e, x, y, 4.0) - ref.loss(p - e, x, y, 4.0)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_asymmetric_weight_plant():
    """A plant: with alpha = 1 the loss is plain MSE; alpha scales only under-predictions."""
    rng = np.random.default_rng(4)
    p = rng.standard_normal(ref.P) * 0.2
    x = rng.uniform(0, 1, size=(20, ref.F))
    yh, _ = ref.forward(p, x)
    y_over = yh - 0.1      # every prediction is above its target
    y_under = yh + 0.1
    assert abs(ref.loss(p, x, y_over, 7.0) - 0.01) < 1e-12
    assert abs(ref.loss(p, x, y_under, 7.0) - 0.07) < 1e-12


def test_parameter_count_matches_header_and_package():
    import sparqsim_amd.sizer as sz

    text = open(os.path.join(ROOT, "include", "sparq_hip.h")).read()
    assert int(re.search(r"#define SPARQ_SIZER_PARAMS (\d+)", text).group(1)) == ref.P == sz.PARAMS
    assert int(re.search(r"#define SPARQ_SIZER_FEATURES (\d+)", text).group(1)) == ref.F == sz.FEATURES


def test_committed_weights_load():
    import sparqsim_amd.sizer as sz

    if not os.path.exists(sz.DEFAULT_WEIGHTS):
        pytest.skip("no committed sizer weights")
    p, m = sz.load_file()
    assert p.shape == (ref.P,) and p.dtype == np.float32 and np.all(np.isfinite(p))
    assert 0.0 <= m < 20.0


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_hip_forward_matches_numpy():
    import sparqsim_amd.sizer as sz

    p = sz.init_params(5)
    x = np.random.default_rng(5).uniform(0, 1, size=(100, ref.F)).astype(np.float32)
    got = sz.forward(p, x)
    want, _ = ref.forward(p.astype(np.float64), x.astype(np.float64))
    assert np.allclose(got, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 37, 300])
def test_hip_gradient_matches_numpy(n):
    import sparqsim_amd.sizer as sz

    rng = np.random.default_rng(n)
    p = sz.init_params(n)
    x = rng.uniform(0, 1, size=(n, ref.F)).astype(np.float32)
    y = rng.uniform(-0.3, 0.6, size=n).astype(np.float32)
    loss, g = sz.grad(p, x, y, alpha=4.0)
    lr, gr = ref.grad(p.astype(np.float64), x.astype(np.float64), y.astype(np.float64), 4.0)
    assert abs(loss - lr) <= 1e-5 * max(1.0, lr)
    scale = np.abs(gr).max()
    assert np.max(np.abs(g - gr)) <= 2e-5 * scale + 1e-7, np.max(np.abs(g - gr)) / scale


@pytest.mark.gpu
def test_hip_training_lowers_loss():
    import sparqsim_amd.sizer as sz

    rng = np.random.default_rng(11)
    x = rng.uniform(0, 1, size=(2000, ref.F)).astype(np.float32)
    y = (0.4 * x[:, 0] - 0.2 * x[:, 3] + 0.3 * np.maximum(x[:, 5] - 0.5, 0)).astype(np.float32)
    p0 = sz.init_params(2)
    l0, _ = sz.grad(p0, x, y)
    p, hist = sz.train(p0, x, y, steps=800, batch=128, lr=1e-3, seed=3)
    l1, _ = sz.grad(p, x, y)
    assert np.all(np.isfinite(hist))
    assert l1 < 0.05 * l0, (l0, l1)
    # deterministic: the same seed gives the same parameters
    p2, _ = sz.train(p0, x, y, steps=800, batch=128, lr=1e-3, seed=3)
    assert np.array_equal(p, p2)


def _sized_workload(ps, seed):
    """A few states hitting Hadamard_Int, Hadamard_Int_Full and a dense group unitary
    with many groups; returns their exports."""
    from sparqsim_amd.grover import NoisyGrover

    out = []
    rng = np.random.default_rng(seed)
    for t in range(6):
        ps.System.clear()
        st = ps.SparseState()
        ps.AddRegister("a", ps.UnsignedInteger, 6)(st)
        ps.AddRegister("b", ps.UnsignedInteger, 7)(st)
        rows = np.unique(np.stack([rng.integers(0, 8, 300), rng.integers(0, 90, 300)], axis=1).astype(np.uint64), axis=0)
        amps = rng.standard_normal(len(rows)) + 1j * rng.standard_normal(len(rows))
        st.load(rows, amps / np.linalg.norm(amps))
        if t % 3 == 0:
            ps.Hadamard_Int("a", 4)(st)
        elif t % 3 == 1:
            ps.Hadamard_Int_Full("a")(st)
        else:
            ps.Hadamard_Int("a", 3)(st)
            z = rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64))
            q, _ = np.linalg.qr(z)
            ps.Rot_GeneralUnitary("a", q.tolist())(st)
        out.append(st.export())
    out.append(NoisyGrover(ps, 8, 7, 7, 2, 0.05).run_batched().export())
    return out


def _trained(ps):
    import sparqsim_amd.sizer as sz
    from sparqsim_amd.train_sizer import collect

    if os.path.exists(sz.DEFAULT_WEIGHTS):
        return sz.load_file()
    x, y = collect(ps, seed=1, trials=150)
    p, _ = sz.train(sz.init_params(0), x, y, steps=1500, batch=128, seed=1)
    return p, sz.fit_margin(p, x, y)


@pytest.mark.gpu
def test_learned_path_is_bit_identical(ps):
    import sparqsim_amd.sizer as sz

    p, margin = _trained(ps)
    sz.enable(False)
    exact = _sized_workload(ps, 21)
    before = sz.stats()
    sz.activate(p, margin)
    try:
        learned = _sized_workload(ps, 21)
    finally:
        sz.enable(False)
    after = sz.stats()
    assert after["predicted"] + after["reruns"] > before["predicted"] + before["reruns"]
    for i, (a, b) in enumerate(zip(learned, exact)):
        assert_same_state(a, b, exact_amps=True, what=f"state {i}")


@pytest.mark.gpu
def test_underprediction_reruns_the_exact_path(ps):
    import sparqsim_amd.sizer as sz

    sz.enable(False)
    exact = _sized_workload(ps, 22)
    p = np.zeros(ref.P, dtype=np.float32)
    p[-1] = -1.0          # predicts log2(G) = -32: capacity of one group
    before = sz.stats()
    sz.activate(p, 0.0)
    try:
        learned = _sized_workload(ps, 22)
    finally:
        sz.enable(False)
    after = sz.stats()
    assert after["reruns"] > before["reruns"]
    for i, (a, b) in enumerate(zip(learned, exact)):
        assert_same_state(a, b, exact_amps=True, what=f"state {i}")


@pytest.mark.gpu
def test_ablation_trained_network_vs_no_network(ps):
    """On fresh logs (another seed): sizing without the count readback by the
    bound the features give (G <= min(N, 2^(bits varying outside the target)))
    never reruns but over-allocates; the trained network, capped by that bound,
    rarely reruns and allocates less in total and per operator on average."""
    from sparqsim_amd.train_sizer import collect, evaluate, evaluate_bound

    p, margin = _trained(ps)
    x, y = collect(ps, seed=99, trials=120)
    assert len(y) > 100
    net = evaluate(p, margin, x, y)
    bound = evaluate_bound(x, y)
    print("network", net, "bound", bound)
    assert bound["rerun_rate"] <= 0.01          # float32 rounding of log2(N) only
    assert net["rerun_rate"] < 0.1
    assert net["allocated_over_needed_total"] <= bound["allocated_over_needed_total"]
    assert net["capacity_over_need_mean"] < bound["capacity_over_need_mean"]
