"""Poiseuille QDA linear solve on the engine -- the counterpart of the reference's
production benchmark (Experiments/QDA/QDA_Poiseuille/PoiseuilleQDABench.cu).

    python -m sparqsim_amd.qda_solve --nqubit 20 --max-steps 40 --cz-eps 1e-16

Solves tridiag(beta, alpha, beta) x = ones/sqrt(n) with the discrete-adiabatic
walk: steps = ceil(steprate * 2305 * kappa) (kappa analytic unless given), capped
by --max-steps and made even (--stop-after N runs only its first N steps); each
step is W(s = i/steps), ClearZero(cz_eps),
Normalize; at every log point the post-selected output gives res_dir, res_opt,
success probability and fidelity against the Thomas solution, exactly as the
reference bench computes them.  Prints one line per log point and a JSON
summary; --learned enables the learned output sizer.
"""
from __future__ import annotations

import argparse
import contextlib
import json
import math
import time

from . import qda


def solve(ps, nqubit=20, steprate=1.1, p=1.3, alpha=2.2666666666564934, beta=-1.0, kappa=0.0, log_every=250,
          tol=1e-6, max_steps=10 ** 9, cz_eps=1e-14, ancilla_order=("anc_4", "anc_3", "anc_2", "anc_1"), log=print,
          stop_after=None, step_hook=None, time_limit=None):
    n = 1 << nqubit
    kappa_analytic = qda.analytic_kappa(alpha, beta, n)
    kappa = kappa if kappa > 0 else kappa_analytic
    steps = min(int(math.ceil(steprate * qda.QDA_STEP_CONSTANT * kappa)), max_steps)
    if steps % 2:
        steps -= 1
    steps = max(steps, 2)
    if log_every % 2:
        log_every += 1
    xref = qda.thomas_solution(alpha, beta, n)
    old_eps = ps.get_clearzero_default_eps()
    ps.set_clearzero_default_eps(cz_eps)       # walk-internal ClearZero() calls follow the threshold
    curve, peak, termination = [], 0, "fixed_step_budget"
    t0 = time.perf_counter()
    t_extract = 0.0
    last_walk = (0.0, 0)
    try:
        st = qda.setup_poiseuille(ps, nqubit, ancilla_order)
        clear = ps.ClearZero(cz_eps)
        norm = ps.Normalize()
        for i in range(steps if stop_after is None else min(steps, stop_after)):
            if step_hook is not None:
                step_hook(i)
            qda.WalkSTridiagonal(ps, "main_reg", "anc_UA", "anc_1", "anc_2", "anc_3", "anc_4", i / steps, kappa, p,
                                 alpha, beta)(st)
            clear(st)
            norm(st)
            out_of_time = time_limit is not None and time.perf_counter() - t0 - t_extract > time_limit
            if (i + 1) % log_every == 0 or i == steps - 1 or (stop_after is not None and i + 1 == stop_after) or out_of_time:
                size = st.size()
                peak = max(peak, size)
                te = time.perf_counter()
                walk_now = te - t0 - t_extract
                interval_ms = (walk_now - last_walk[0]) / (i + 1 - last_walk[1]) * 1e3
                last_walk = (walk_now, i + 1)
                out, _ = qda.get_output(st, ps)
                m = qda.solution_metrics(out, n, alpha, beta, xref)
                t_extract += time.perf_counter() - te
                m.update(step=i + 1, entries=size, wall_s=time.perf_counter() - t0, interval_ms_per_step=interval_ms)
                curve.append(m)
                log(f"step {i + 1:7d}/{steps} res_dir={m['res_dir']:.6e} res_opt={m['res_opt']:.6e} "
                    f"succP={m['success_prob']:.4f} fid={m['fidelity']:.6f} entries={size} wall={m['wall_s']:.1f}s "
                    f"ms/step={interval_ms:.1f}")
                if not math.isnan(m["res_dir"]) and m["res_dir"] < tol:
                    termination = f"residual_below_tol({tol:g})"
                    break
                if out_of_time:
                    termination = f"time_limit({time_limit:g}s)"
                    break
    finally:
        ps.set_clearzero_default_eps(old_eps)
    wall = time.perf_counter() - t0
    last = curve[-1]
    return {"nqubit": nqubit, "n": n, "kappa": kappa, "kappa_analytic": kappa_analytic, "steps_budget": steps,
            "steps_run": last["step"], "cz_eps": cz_eps, "p": p, "alpha": alpha, "beta": beta,
            "final_res_dir": last["res_dir"], "final_res_opt": last["res_opt"], "final_success_prob": last["success_prob"],
            "final_fidelity": last["fidelity"], "wall_s": wall, "walk_s": wall - t_extract,
            "ms_per_step": (wall - t_extract) / last["step"] * 1e3, "peak_logged_entries": peak,
            "termination": termination, "curve": curve}


def write_artifacts(outdir, res, args, ps):
    """residual_curve.csv + qda_summary_cpp.json with the keys of PoiseuilleQDABench.cu:170-293."""
    import os

    from .state_io import write_residual_curve, write_summary_json

    os.makedirs(outdir, exist_ok=True)
    write_residual_curve(os.path.join(outdir, "residual_curve.csv"),
                         [(m["step"], m["res_dir"], m["res_opt"], m["success_prob"], m["fidelity"]) for m in res["curve"]])
    write_summary_json(os.path.join(outdir, "qda_summary_cpp.json"), {
        "engine": "qda (sparqsim_amd WalkSTridiagonal + device GetOutput post-selection)",
        "backend": "HIP (MI355X engine)", "dimension": res["n"], "nqubit": res["nqubit"], "kappa": res["kappa"],
        "kappa_analytic": res["kappa_analytic"], "eps": 0.01, "alpha": res["alpha"], "beta": res["beta"], "p": res["p"],
        "clearzero_eps": res["cz_eps"], "problem": "1D Poisson tridiagonal, b=ones/sqrt(n)",
        "step_budget": res["steps_budget"], "steps_executed": res["steps_run"], "step_rate": args.steprate,
        "step_constant": qda.QDA_STEP_CONSTANT, "termination": res["termination"], "tol": args.tol,
        "final_res_dir": res["final_res_dir"], "final_res_opt": res["final_res_opt"],
        "final_success_prob": res["final_success_prob"], "final_fidelity_vs_classical": res["final_fidelity"],
        "wall_time_s": res["wall_s"], "time_per_step_s": res["wall_s"] / max(res["steps_run"], 1),
        "extract_time_s": res["wall_s"] - res["walk_s"], "max_system_size": ps.System.max_system_size(),
        "note": "sparqsim_amd engine run"})


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--nqubit", type=int, default=20)
    ap.add_argument("--steprate", type=float, default=1.1)
    ap.add_argument("--p", type=float, default=1.3)
    ap.add_argument("--alpha", type=float, default=2.2666666666564934)
    ap.add_argument("--beta", type=float, default=-1.0)
    ap.add_argument("--kappa", type=float, default=0.0)
    ap.add_argument("--log-every", type=int, default=250)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-steps", type=int, default=10 ** 9)
    ap.add_argument("--cz-eps", type=float, default=1e-14)
    ap.add_argument("--stop-after", type=int, default=None, help="run only the first N steps of the schedule")
    ap.add_argument("--time-limit", type=float, default=None,
                    help="stop after the first step that ends past this many seconds of walk time (extraction excluded)")
    ap.add_argument("--learned", action="store_true", help="size branching operators with the learned sizer")
    ap.add_argument("--json", default=None, help="write the summary (with the curve) to this file")
    ap.add_argument("--kernel-profile", default=None,
                    help="write the engine's per-kernel profile (launches, time, algorithmic bytes) and the sort-path "
                         "counters of the run (from --profile-from on) to this JSON file")
    ap.add_argument("--profile-from", type=int, default=0, help="start the kernel profile at this step")
    ap.add_argument("--outdir", default=None,
                    help="write residual_curve.csv and qda_summary_cpp.json here (the reference bench's artifacts)")
    args = ap.parse_args()
    import sparqsim_amd as ps
    from . import sizer

    if args.learned and not sizer.activate_default():
        raise SystemExit("no sizer weights found")
    from .profiling import KernelProfile

    prof = KernelProfile(ps, start_now=args.profile_from == 0) if args.kernel_profile else None
    hook = (lambda i: prof.start() if i == args.profile_from else None) if prof and args.profile_from else None
    with prof or contextlib.nullcontext():
        res = solve(ps, args.nqubit, args.steprate, args.p, args.alpha, args.beta, args.kappa, args.log_every, args.tol,
                    args.max_steps, args.cz_eps, log=lambda s: print(s, flush=True), stop_after=args.stop_after,
                    step_hook=hook, time_limit=args.time_limit)
    if prof:
        prof.write(args.kernel_profile, steps=res["steps_run"] - args.profile_from, first_step=args.profile_from,
                   peak_logged_entries=res["peak_logged_entries"])
    if args.learned:
        res["sizer"] = sizer.stats()
    res["max_system_size"] = ps.System.max_system_size()
    summary = {k: v for k, v in res.items() if k != "curve"}
    print(json.dumps(summary), flush=True)
    if args.outdir:
        write_artifacts(args.outdir, res, args, ps)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f)


if __name__ == "__main__":
    main()
