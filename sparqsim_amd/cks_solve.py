"""Poiseuille CKS linear solve on the engine -- the counterpart of the reference's
production benchmark (Experiments/CKS/CKSPoiseuilleBench.cu), extended to many
right-hand sides in one state.

    python -m sparqsim_amd.cks_solve --nqubit 20 --rhs 16 --terms 8

Solves tridiag(beta, alpha, beta) x_r = b_r for R right-hand sides (r = 0 is
ones/sqrt(n), the others seeded N(0,1) vectors) with the Chebyshev quantum walk
and direct tridiagonal oracles (--matrix band-random: a seeded symmetric
positive matrix with the diagonal, the +-1 band and one random partner per row,
condition number --kappa, through the QRAM-backed sparse-access oracles): input D|b_r> (D = diag((-1)^j)), FirstStep, then
one LCU term per Step (two walk steps each); at every log point the dense
post-selection gives, per right-hand side, res_dir, res_opt, fidelity against the
classical tridiagonal solve and p_succ.  Coefficients: the fitted 1/x family on
the quantised walk spectrum (--coef true, J terms) or the binomial-tail family
(--coef stock).  Prints one line per log point and a JSON summary.
"""
from __future__ import annotations

import argparse
import contextlib
import json
import time

import numpy as np

from . import cks


def make_rhs(n: int, R: int, seed: int = 0x5EED2026) -> np.ndarray:
    rng = np.random.default_rng(seed)
    rows = [np.ones(n)] + [rng.standard_normal(n) for _ in range(R - 1)]
    return np.stack(rows)


def solve(ps, nqubit=20, R=1, kappa=16.0, eps=0.01, ds=32, coef="true", J=640, terms=None, log_every=64, tol=1e-6,
          alpha=2.2666666666564934, beta=-1.0, log=print, matrix="poiseuille", fit_top=None, clearzero_eps=None,
          merge_eps=None):
    """clearzero_eps: the ClearZero threshold of every walk operator (the process default is restored
    afterwards); merge_eps: the accumulator's add_systems threshold.  None keeps 1e-14 for both."""
    n = 1 << nqubit
    saved_eps = ps.get_clearzero_default_eps()
    if clearzero_eps is not None:
        ps.set_clearzero_default_eps(clearzero_eps)
    try:
        return _solve(ps, nqubit, R, kappa, eps, ds, coef, J, terms, log_every, tol, alpha, beta, log, matrix, fit_top,
                      merge_eps)
    finally:
        ps.set_clearzero_default_eps(saved_eps)


def _solve(ps, nqubit, R, kappa, eps, ds, coef, J, terms, log_every, tol, alpha, beta, log, matrix, fit_top, merge_eps):
    n = 1 << nqubit
    t_setup = time.perf_counter()
    if matrix == "poiseuille":
        batch = cks.BatchedPoiseuilleCKS(ps, nqubit, make_rhs(n, R), kappa=kappa, eps=eps, ds=ds, alpha=alpha, beta=beta,
                                         coef=coef, J=J, fit_top=fit_top or "spectrum")
    elif matrix == "band-random":
        batch = cks.BatchedSparseCKS(ps, nqubit, make_rhs(n, R), kappa=kappa, ds=ds, J=J, eps=eps,
                                     fit_top=fit_top or "unit")
    else:
        raise ValueError("matrix must be 'poiseuille' or 'band-random'")
    batch.lcu.merge_eps = merge_eps
    batch.prepare_input()
    after_first = batch.lcu.step_state.size()
    t_setup = time.perf_counter() - t_setup
    log(f"[setup] n={n} R={R} j0={batch.lcu.j0} entries after FirstStep={after_first} ({t_setup:.1f}s)")
    budget = batch.lcu.j0 + 1 if terms is None else min(terms, batch.lcu.j0 + 1)
    curve, termination = [], "full_lcu_expansion" if terms is None else "term_budget"
    t0 = time.perf_counter()
    t_extract = 0.0
    j = 0
    while j < budget and batch.Step():
        j += 1
        if j % log_every == 0 or j == budget:
            te = time.perf_counter()
            _, res = batch.extract()
            t_extract += time.perf_counter() - te
            worst = max(r["res_dir"] for r in res)
            curve.append({"term": j, "walk_steps": 2 * j + 1, "entries": batch.lcu.step_state.size(),
                          "accumulated": batch.lcu.current_state.size(), "wall_s": time.perf_counter() - t0,
                          "per_rhs": res})
            log(f"term {j:5d}/{batch.lcu.j0 + 1} worst res_dir={worst:.3e} rhs0 res_dir={res[0]['res_dir']:.6e} "
                f"p_succ={res[0]['p_succ']:.6e} fid={res[0]['fidelity']:.8f} entries={curve[-1]['entries']} "
                f"wall={curve[-1]['wall_s']:.1f}s")
            if worst < tol:
                termination = f"residual_below_tol({tol:g})"
                break
    wall = time.perf_counter() - t0
    last = curve[-1]
    return {"n": n, "nqubit": nqubit, "rhs": R, "matrix": matrix, "kappa": kappa, "eps": eps, "ds": ds, "coef": coef,
            "fit_interval": getattr(batch, "fit_interval", None), "clearzero_eps": ps.get_clearzero_default_eps(),
            "merge_eps": merge_eps if merge_eps is not None else 1e-14,
            "sum_abs_coef": float(np.sum(np.abs(batch.lcu.inv_coef))) if batch.lcu.inv_coef is not None else None,
            "J": J if coef == "true" else None, "lcu_terms_budget": batch.lcu.j0 + 1, "terms_run": j,
            "entries_after_first_step": after_first, "termination": termination, "wall_s": wall, "walk_s": wall - t_extract,
            "s_per_term": (wall - t_extract) / max(j, 1), "setup_s": t_setup,
            "entries_per_rhs": after_first / R, "final": last["per_rhs"], "max_system_size": ps.System.max_system_size(),
            "curve": curve}


def write_artifacts(outdir, res, args):
    """residual_curve.csv (walk step = 2 term + 1) + qda_summary_cpp.json with the keys of
    CKSPoiseuilleBench.cu:305-420, for right-hand side 0."""
    import os

    from .state_io import write_residual_curve, write_summary_json

    os.makedirs(outdir, exist_ok=True)
    rows = [(2 * c["term"] + 1, c["per_rhs"][0]["res_dir"], c["per_rhs"][0]["res_opt"], c["per_rhs"][0]["p_succ"],
             c["per_rhs"][0]["fidelity"]) for c in res["curve"]]
    write_residual_curve(os.path.join(outdir, "residual_curve.csv"), rows)
    last = res["final"][0]
    write_summary_json(os.path.join(outdir, "qda_summary_cpp.json"), {
        "engine": "cks (sparqsim_amd Chebyshev quantum walk + LCU)", "backend": "HIP (MI355X engine)",
        "dimension": res["n"], "nqubit": res["nqubit"], "kappa": res["kappa"], "eps": res["eps"], "ds": res["ds"],
        "coef_mode": res["coef"], "J": res["J"], "matrix": res["matrix"], "rhs": res["rhs"],
        "lcu_terms": res["lcu_terms_budget"], "walk_steps": 2 * res["lcu_terms_budget"],
        "termination": res["termination"], "tol": args.tol, "final_res_dir": last["res_dir"],
        "final_res_opt": last["res_opt"], "final_success_prob": last["p_succ"],
        "final_fidelity_vs_classical": last["fidelity"], "wall_time_s": res["wall_s"],
        "time_per_term_s": res["s_per_term"], "extract_time_s": res["wall_s"] - res["walk_s"],
        "max_system_size": res["max_system_size"], "note": "sparqsim_amd engine run"})


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--nqubit", type=int, default=20)
    ap.add_argument("--rhs", type=int, default=1)
    ap.add_argument("--kappa", type=float, default=16.0)
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--ds", type=int, default=32)
    ap.add_argument("--coef", choices=["true", "stock"], default="true")
    ap.add_argument("--J", type=int, default=640)
    ap.add_argument("--terms", type=int, default=None, help="stop after this many LCU terms")
    ap.add_argument("--log-every", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--json", default=None, help="write the summary (with the curve) to this file")
    ap.add_argument("--kernel-profile", default=None,
                    help="write the engine's per-kernel profile (launches, time, algorithmic bytes) and the sort-path "
                         "counters of the whole run to this JSON file")
    ap.add_argument("--outdir", default=None,
                    help="write residual_curve.csv and qda_summary_cpp.json here (the reference bench's artifacts, "
                         "right-hand side 0)")
    ap.add_argument("--fit-top", choices=["unit", "spectrum"], default=None,
                    help="upper end of the 1/x fit: 1 (default for band-random) or the largest walk eigenvalue "
                         "(default for the Poiseuille case, as the reference fits)")
    ap.add_argument("--clearzero-eps", type=float, default=None,
                    help="ClearZero threshold of the walk operators on |a|^2 (default 1e-14)")
    ap.add_argument("--merge-eps", type=float, default=None,
                    help="threshold of the accumulator's add_systems merge on |a|^2 (default 1e-14)")
    ap.add_argument("--matrix", choices=["poiseuille", "band-random"], default="poiseuille",
                    help="tridiagonal Poiseuille case (direct oracles) or a seeded band-plus-random matrix "
                         "(QRAM-backed sparse-access oracles, fitted coefficients; --kappa sets its condition number)")
    args = ap.parse_args()
    import sparqsim_amd as ps

    from .profiling import KernelProfile

    with KernelProfile(ps) if args.kernel_profile else contextlib.nullcontext() as prof:
        res = solve(ps, args.nqubit, args.rhs, args.kappa, args.eps, args.ds, args.coef, args.J, args.terms, args.log_every,
                    args.tol, log=lambda s: print(s, flush=True), matrix=args.matrix, fit_top=args.fit_top,
                    clearzero_eps=args.clearzero_eps, merge_eps=args.merge_eps)
    if args.kernel_profile:
        prof.write(args.kernel_profile, terms=res["curve"][-1]["term"] if res["curve"] else None,
                   entries_after_first_step=res.get("entries_after_first_step"))
    summary = {k: v for k, v in res.items() if k not in ("curve", "final")}
    summary["final_worst_res_dir"] = max(r["res_dir"] for r in res["final"])
    summary["final_rhs0"] = res["final"][0]
    print(json.dumps(summary), flush=True)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f)
    if args.outdir:
        write_artifacts(args.outdir, res, args)


if __name__ == "__main__":
    main()
