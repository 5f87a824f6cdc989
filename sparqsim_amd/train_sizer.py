"""Train the learned output sizer on logs the simulator produces itself.

    python -m sparqsim_amd.train_sizer [--out PATH] [--steps N]

Runs a mix of workloads with sizer logging on -- QDA walk steps on the
Poiseuille problem (Hadamard_Int_Full on the main register, dense state
preparation on the ancillas), batched noisy Grover trajectories
(Hadamard_Int_Full per trajectory segment), the QDA over the sparse block
encoding and the CKS walks, and seeded random states under
Hadamard_Int / Hadamard_Int_Full / Rot_GeneralUnitary -- then fits the network
with the engine's GPU kernels, picks the safety margin on the training split,
reports the held-out split, and saves parameters + margin (npz).  Needs a GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

from . import sizer
from .workloads import BASE_SEED


def random_structured_states(ps, rng, trials: int):
    """States whose group structure varies: K distinct values of the other
    registers, target values drawn from a sub-range, then one sized operator."""
    for t in range(trials):
        ps.System.clear()
        wa = int(rng.integers(2, 13))
        wb = int(rng.integers(1, 16))
        wc = int(rng.integers(1, 12))
        st = ps.SparseState()
        ps.AddRegister("a", ps.UnsignedInteger, wa)(st)
        ps.AddRegister("b", ps.UnsignedInteger, wb)(st)
        ps.AddRegister("c", ps.UnsignedInteger, wc)(st)
        n = int(min(2 ** int(rng.integers(1, 17)), 2 ** (wa + wb + wc)))
        k = max(1, int(n / 2 ** rng.uniform(0, min(wa, 12))))
        others = np.unique(rng.integers(0, 2 ** (wb + wc), size=k, dtype=np.uint64))
        tbits = int(rng.integers(0, wa + 1))
        pick = others[rng.integers(0, len(others), size=n)]
        tval = rng.integers(0, 2 ** tbits, size=n, dtype=np.uint64) if tbits else np.zeros(n, dtype=np.uint64)
        rows = np.unique(np.stack([tval, pick >> np.uint64(wc), pick & np.uint64(2 ** wc - 1)], axis=1), axis=0)
        amps = rng.standard_normal(len(rows)) + 1j * rng.standard_normal(len(rows))
        amps /= np.linalg.norm(amps)
        st.load(rows.astype(np.uint64), amps)
        kind = int(rng.integers(0, 3))
        if kind == 0:
            ps.Hadamard_Int("a", int(rng.integers(1, wa + 1)))(st)
        elif kind == 1:
            ps.Hadamard_Int_Full("a")(st)
        else:
            w = min(wa, 5)
            if w != wa:
                ps.Hadamard_Int("a", int(rng.integers(1, wa + 1)))(st)
            else:
                z = rng.standard_normal((2 ** w, 2 ** w)) + 1j * rng.standard_normal((2 ** w, 2 ** w))
                q, _ = np.linalg.qr(z)
                ps.Rot_GeneralUnitary("a", q.tolist())(st)
        st.size()


def qda_steps(ps, sizes, steps):
    from . import qda

    old = ps.get_clearzero_default_eps()
    ps.set_clearzero_default_eps(1e-16)
    try:
        for nq in sizes:
            st = qda.setup_poiseuille(ps, nq)
            for s in range(steps):
                qda.WalkSTridiagonal(ps, "main_reg", "anc_UA", "anc_1", "anc_2", "anc_3", "anc_4", s / 2000, 16.0, 1.3,
                                     2.2666666666564934, -1.0)(st)
                ps.ClearZero()(st)
            st.size()
    finally:
        ps.set_clearzero_default_eps(old)


def sparse_qda_steps(ps, sizes, steps, starts=(0.0, 0.5, 0.9)):
    """Steps of the QDA over the sparse block encoding (qda_sparse) from fresh
    states at several points of the schedule."""
    from . import qda_sparse

    for nq in sizes:
        for s0 in starts:
            st, walk, enc, A, kappa = qda_sparse.setup(ps, nq)
            for i in range(steps):
                qda_sparse.WalkSSparse(ps, enc, "anc_1", "anc_2", "anc_3", "anc_4", s0 + i / 2000, kappa, 1.3)(st)
                ps.ClearZero()(st)
                ps.Normalize()(st)
            st.size()


def cks_steps(ps, sizes, terms):
    """LCU terms of the CKS walk: the tridiagonal walk and the QRAM-backed walk."""
    from . import cks_solve

    for nq in sizes:
        cks_solve.solve(ps, nqubit=nq, R=2, terms=terms, log_every=10 ** 9, log=lambda s: None)
        cks_solve.solve(ps, nqubit=max(nq, 5), R=2, kappa=50.0, J=64, terms=terms, log_every=10 ** 9,
                        log=lambda s: None, matrix="band-random")


def grover_runs(ps, configs):
    from .grover import NoisyGrover

    for traj, bits in configs:
        NoisyGrover(ps, traj, bits, bits, 2, 0.05).run_batched().size()


def collect(ps, seed: int = 0, trials: int = 600, qda_sizes=(4, 6, 8, 10, 12, 14), qda_step_count: int = 6,
            sparse_qda_sizes=(), cks_sizes=()):
    rng = np.random.default_rng(BASE_SEED + seed)
    with sizer.logging():
        qda_steps(ps, qda_sizes, qda_step_count)
        grover_runs(ps, ((1, 6), (4, 8), (16, 8), (8, 10), (32, 6), (64, 9)))
        random_structured_states(ps, rng, trials)
        if sparse_qda_sizes:
            sparse_qda_steps(ps, sparse_qda_sizes, 4)
        if cks_sizes:
            cks_steps(ps, cks_sizes, 3)
    return sizer.fetch_logs()


def capacity_report(cap, need):
    """Reruns (capacity below need) and how much the others over-allocate."""
    rerun = cap < need - 0.5
    ok = ~rerun
    ratio = cap[ok] / need[ok]
    return {"samples": int(len(need)), "rerun_rate": float(rerun.mean()) if len(need) else 0.0,
            "capacity_over_need_mean": float(ratio.mean()) if ok.any() else None,
            "capacity_over_need_p99": float(np.quantile(ratio, 0.99)) if ok.any() else None,
            "capacity_over_need_max": float(ratio.max()) if ok.any() else None,
            "allocated_over_needed_total": float(cap[ok].sum() / need[ok].sum()) if ok.any() else None}


def _bound(x):
    """G <= min(N, 2^(bits varying outside the low target bits)), from the features."""
    return np.ceil(np.minimum(np.exp2(x[:, 0].astype(np.float64) * 32.0),
                              np.exp2(np.minimum(x[:, 3].astype(np.float64) * 32.0, 40.0))))


def evaluate(params, margin, x, y):
    """The engine's capacities (network + margin, capped by the bound) against the
    exact group counts."""
    pred = sizer.forward(params, x).astype(np.float64) * 32.0 + margin
    need = np.exp2(y.astype(np.float64) * 32.0)
    cap = np.minimum(np.ceil(np.exp2(np.maximum(pred, 0.0))), _bound(x))
    return capacity_report(cap, need)


def evaluate_bound(x, y):
    """Sizing without the network: the bound alone."""
    return capacity_report(_bound(x), np.exp2(y.astype(np.float64) * 32.0))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=sizer.DEFAULT_WEIGHTS)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--trials", type=int, default=3000)
    args = ap.parse_args()
    import sparqsim_amd as ps

    t0 = time.time()
    x, y = collect(ps, seed=0, trials=args.trials, qda_sizes=(4, 6, 8, 10, 12, 14, 16, 18, 20), qda_step_count=8,
                   sparse_qda_sizes=(5, 6, 8, 10, 12), cks_sizes=(4, 6, 8, 10, 12))
    t_collect = time.time() - t0
    # train / validation (picks the margin) / test (reported) = 70 / 15 / 15
    order = np.random.default_rng(7).permutation(len(y))
    a, b = int(0.7 * len(y)), int(0.85 * len(y))
    tr, va, te = order[:a], order[a:b], order[b:]
    p0 = sizer.init_params(0)
    t0 = time.time()
    params, hist = sizer.train(p0, x[tr], y[tr], steps=args.steps, batch=256, lr=1e-3, alpha=4.0, seed=1)
    t_train = time.time() - t0
    margin = sizer.fit_margin(params, x[va], y[va], quantile=0.99, slack=0.1)
    report = {"samples": int(len(y)), "collect_s": t_collect, "train_s": t_train, "steps": args.steps,
              "loss_first": float(hist[:50].mean()), "loss_last": float(hist[-50:].mean()), "margin_log2": margin,
              "train": evaluate(params, margin, x[tr], y[tr]), "test": evaluate(params, margin, x[te], y[te]),
              "test_bound_without_network": evaluate_bound(x[te], y[te])}
    sizer.save(args.out, params, margin, train_samples=len(tr), validation_samples=len(va), test_samples=len(te),
               loss_first=report["loss_first"], loss_last=report["loss_last"])
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
