"""Benchmark: single-state QRAM interference at 2^20 entries on one MI355X.

One step = one pass of the operator-apply path over the workload: H(addr,20) on
|0> (1 -> 2^20 entries), QRAMLoad data ^= M[addr] (M = 2^20 uint32 words),
32 Z phases on the data bits, the load's uncompute, and the inverse H(addr,20)
(radix sort + 2^20-point Walsh-Hadamard + ClearZero).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sparse entries/s + HBM GB/s; 2^20 QDA solve time (batched+learned vs plain)"
HBM_PEAK_GBS = 8000.0

# Measured HBM bytes per launch at 2^20 entries / 32-bit data (rocprofv3 --pmc FETCH_SIZE x2.0 +
# WRITE_SIZE x1.0, factors calibrated on a 512 MiB device copy); profiles/qram_interference_2p20_traffic.md.
PMC_TRAFFIC_MIB_2P20 = {
    "pointwise_kernel": 28.1 + 24.0, "compact_scatter_kernel": 24.1 + 24.0, "compact_count_kernel": 16.0,
    "key_stats_kernel": 8.2, "expand_single_kernel": 0.0 + 24.0, "wht_first_pass_kernel": 16.1 + 16.0,
    "wht_pass_kernel": 16.1 + 16.0,
}


def kernel_profile(lib):
    """Per-kernel records of the engine's profiler (launches, ms, algorithmic bytes)."""
    from sparqsim_amd import profiling

    return profiling.report()


def cpu_baseline(address_bits, data_bits, mem_words, budget_s=10.0, module="_oracle"):
    """The oracle (restatement of the reference CPU path; `_oracle_omp` has the
    reference's own OpenMP loops enabled) on a bounded sample: the same circuit,
    repeated until `budget_s` seconds have elapsed."""
    import importlib

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    o = importlib.import_module(module)
    from sparqsim_amd.workloads import QRAMInterference

    o.System.clear()
    st = o.SparseState()
    o.AddRegister("addr", o.UnsignedInteger, address_bits)(st)
    o.AddRegister("data", o.UnsignedInteger, data_bits)(st)
    q = o.QRAMCircuit_qutrit(address_bits, data_bits, mem_words)
    circ = QRAMInterference(o, q, address_bits, data_bits)
    reps, total = 0, 0.0
    while total < budget_s and reps < 50:
        o.reset_systems(st)
        t0 = time.perf_counter()
        circ(st)
        total += time.perf_counter() - t0
        reps += 1
    o.System.clear()
    return circ.entry_ops * reps / total, total, reps


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_process(module, threads, address_bits, seconds):
    """cpu_baseline in a child process (the two oracle builds cannot share one
    interpreter); OMP_NUM_THREADS = threads.  Returns (entries/s, seconds, passes)."""
    import subprocess

    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", module,
                          "--cpu-address-bits", str(address_bits), "--cpu-seconds", str(seconds)],
                         env=env, capture_output=True, text=True, timeout=600, check=True)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return r["entries_per_s"], r["seconds"], r["passes"]


def qda_walk_run(ps, nqubit, nsteps, total_steps):
    """Steps 0..nsteps-1 of the 2^nqubit Poiseuille walk; returns (sizes, seconds per step)."""
    from sparqsim_amd import qda

    alpha, beta, kappa, p = 2.2666666666564934, -1.0, 16.0, 1.3
    old_eps = ps.get_clearzero_default_eps()
    ps.set_clearzero_default_eps(1e-16)
    st = qda.setup_poiseuille(ps, nqubit)
    st.size()
    sizes, times = [], []
    try:
        for n in range(nsteps):
            t0 = time.perf_counter()
            qda.WalkSTridiagonal(ps, "main_reg", "anc_UA", "anc_1", "anc_2", "anc_3", "anc_4",
                                 n / total_steps, kappa, p, alpha, beta)(st)
            ps.ClearZero()(st)
            sizes.append(st.size())
            times.append(time.perf_counter() - t0)
    finally:
        ps.set_clearzero_default_eps(old_eps)
    return sizes, times


def qda_walk_leg(ps, nqubit, nsteps, total_steps):
    """The 2^nqubit Poiseuille QDA solve (A = tridiag(-1, 2.2666666666564934, -1),
    b = ones/sqrt(n), kappa 16, p 1.3, ClearZero eps 1e-16) on the engine, timed over
    its first `nsteps` walk steps of a `total_steps` schedule (one step = W(s) +
    ClearZero), once on the plain path and once with the learned sizer."""
    from sparqsim_amd import sizer

    sizer.enable(False)
    sizes, times = qda_walk_run(ps, nqubit, nsteps, total_steps)
    learned = None
    if sizer.activate_default():
        s0 = sizer.stats()
        lsizes, ltimes = qda_walk_run(ps, nqubit, nsteps, total_steps)
        s1 = sizer.stats()
        sizer.enable(False)
        assert lsizes == sizes, "the learned sizer changed the walk"
        learned = {"wall_s": sum(ltimes), "ms_per_step": sum(ltimes) / nsteps * 1e3,
                   "sized_ops_predicted": s1["predicted"] - s0["predicted"], "reruns": s1["reruns"] - s0["reruns"],
                   "speedup_vs_plain": sum(times) / sum(ltimes)}
    wall = sum(times)
    return {
        "workload": f"2^{nqubit} Poiseuille QDA walk (tridiag(-1, 2.2666666666564934, -1), kappa 16, p 1.3, "
                    f"czEps 1e-16), steps 0..{nsteps - 1} of a {total_steps}-step schedule",
        "mode": "plain single-state",
        "steps": nsteps, "wall_s": wall, "ms_per_step": wall / nsteps * 1e3,
        "entries_first": sizes[0], "entries_last": sizes[-1], "entries_peak": max(sizes),
        "entry_steps_per_s": sum(sizes) / wall,
        "last_step_ms": times[-1] * 1e3,
        "learned_sizer": learned,
        "reference_published": "1.33 s/step at ~33.5M entries, 1x A800 (docs/verification/v1d-qda-linsolver-a800.md:52)",
    }


def grover_leg(ps, n_traj, address_bits, iterations, p, plain_traj):
    """Batched noisy Grover: n_traj trajectories in one state keyed by `traj` vs the same
    trajectories run one state at a time (timed on the first `plain_traj` of them and
    scaled to n_traj; the trajectories are independent and of equal size)."""
    from sparqsim_amd.grover import NoisyGrover

    g = NoisyGrover(ps, n_traj, address_bits, address_bits, iterations, p)
    n_events = sum(len(e) for e in g.events)
    g.run_batched().size()                      # warm-up (allocator pool, code objects)
    t0 = time.perf_counter()
    st = g.run_batched()
    n_final = st.size()
    t_batched = time.perf_counter() - t0
    del st
    g.new_plain_registers()
    g.run_plain(0).size()
    t0 = time.perf_counter()
    for t in range(plain_traj):
        g.run_plain(t).size()
    t_plain = (time.perf_counter() - t0) * n_traj / plain_traj
    t_learned = None
    from sparqsim_amd import sizer

    if sizer.activate_default():
        t0 = time.perf_counter()
        g.run_batched().size()
        t_learned = time.perf_counter() - t0
        sizer.enable(False)
    captured = grover_captured_timing(g, iterations)
    # each iteration applies 12 reference operators to every entry of every trajectory
    entry_ops = 12 * iterations * n_final
    return {
        "workload": f"{n_traj} Grover trajectories, {address_bits}-bit addr/data/search, {iterations} iterations, "
                    f"Pauli X/Z insertions p={p} per addr qubit per iteration ({n_events} events)",
        "entries_final": n_final, "batched_s": t_batched, "batched_ms_per_iteration": t_batched / iterations * 1e3,
        "batched_entries_per_s": entry_ops / t_batched,
        "batched_learned_s": t_learned,
        "plain_s": t_plain, "plain_sample_trajectories": plain_traj,
        "batched_speedup_vs_plain": t_plain / t_batched,
        "batched_learned_speedup_vs_plain": (t_plain / t_learned) if t_learned else None,
        "captured": captured,
        "captured_speedup_vs_batched": t_batched / captured["replays_s"],
        "uneven": grover_uneven_leg(ps, n_traj, address_bits, iterations, p),
    }


def grover_captured_timing(g, iterations):
    """The batch with every iteration replayed from one captured HIP graph
    (oracle, diffusion, this iteration's noise from a device table): the
    replays are queued back to back with no host synchronisation; setup is the
    one-off planning and recording of the sequence."""
    st, seq = g.capture_batched()               # warm-up
    seq.launch(iterations)
    seq.sync()
    del seq, st
    t0 = time.perf_counter()
    st, seq = g.capture_batched()
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    seq.launch(iterations)
    seq.sync()
    t_replays = time.perf_counter() - t0
    out = {"setup_s": t_setup, "replays_s": t_replays, "ms_per_iteration": t_replays / iterations * 1e3,
           "graph_nodes_per_iteration": seq.nodes, "capacity": seq.capacity, "entries_final": st.size()}
    del seq, st
    return out


def grover_uneven_leg(ps, n_traj, address_bits, iterations, p):
    """The same batch with uneven nnz: trajectory t searches the first
    2^(t mod (address_bits + 1)) addresses (1 to 2^address_bits entries), one
    conditioned Hadamard per width in each diffusion; against every trajectory
    run one state at a time (timed on one trajectory of each width and weighted
    by the number of trajectories of that width)."""
    from sparqsim_amd.grover import NoisyGrover, uneven_widths

    widths = uneven_widths(n_traj, address_bits)
    g = NoisyGrover(ps, n_traj, address_bits, address_bits, iterations, p, widths=widths)
    g.run_batched().size()                      # warm-up
    t0 = time.perf_counter()
    st = g.run_batched()
    n_final = st.size()
    t_batched = time.perf_counter() - t0
    counts = np.bincount(st.export()[1][:, 0].astype(np.int64), minlength=n_traj)
    del st
    g.new_plain_registers()
    g.run_plain(0).size()
    t_plain = 0.0
    for c in sorted(set(widths)):
        t = widths.index(c)
        t0 = time.perf_counter()
        g.run_plain(t).size()
        t_plain += (time.perf_counter() - t0) * widths.count(c)
    captured = grover_captured_timing(g, iterations)
    return {"workload": f"{n_traj} trajectories searching 2^0..2^{address_bits} addresses (t mod {address_bits + 1} bits)",
            "entries_final": n_final, "entries_per_trajectory_min": int(counts.min()),
            "entries_per_trajectory_max": int(counts.max()), "batched_s": t_batched,
            "batched_entries_per_s": 12 * iterations * n_final / t_batched, "plain_s": t_plain,
            "batched_speedup_vs_plain": t_plain / t_batched,
            "captured": captured, "captured_speedup_vs_batched": t_batched / captured["replays_s"],
            "captured_speedup_vs_plain": t_plain / captured["replays_s"]}


def qda_sparse_leg(ps, nqubit, steps):
    """QDA over the sparse block encoding of the configuration's complex
    Hermitian matrix (8 N(0,1) + i N(0,1) entries per row at random columns,
    phase oracle, encoded kappa 100): the first `steps` steps of the 2000-step
    schedule, plain and with the learned sizer."""
    from sparqsim_amd import qda_sparse, sizer

    out = {"workload": f"QDA over a sparse block encoding (T^dag S P T of the CKS walk with the phase oracle) of a "
                       f"seeded complex Hermitian matrix with 8 random-column slots per row at 2^{nqubit}, encoded "
                       f"kappa 100, steps 0..{steps - 1} of a 2000-step schedule"}
    sizer.enable(False)
    for mode in ("plain", "learned"):
        if mode == "learned" and not sizer.activate_default():
            break
        s0 = sizer.stats()
        met, st, *_ = qda_sparse.solve(ps, nqubit, 2000, stop_after=steps, hermitian=True, kappa_encoded=100.0)
        out[mode] = {"ms_per_step": met["ms_per_step"], "entries": met["entries"], "fidelity": met["fidelity"]}
        if mode == "learned":
            s1 = sizer.stats()
            out[mode]["sized_ops_predicted"] = s1["predicted"] - s0["predicted"]
            out[mode]["reruns"] = s1["reruns"] - s0["reruns"]
            sizer.enable(False)
        del st
        ps.System.clear()
        ps.release_cached_memory()
    if "learned" in out:
        out["learned_speedup_vs_plain"] = out["plain"]["ms_per_step"] / out["learned"]["ms_per_step"]
    return out


SORT_KERNELS = ("radix_hist_kernel", "radix_scatter_kernel", "transpose_u32_kernel", "scan_reduce_kernel",
                "scan_apply_kernel", "gather_state_kernel", "gather_word_kernel", "key_stats_kernel",
                "key_stats_final_kernel", "payload_scatter_kernel", "local_sort_kernel")


def sort_leg(ps, lib, address_bits, reps, mode=0):
    """SortUnconditional of 2^address_bits entries whose 48-bit keys (b: 24 bits,
    most significant, then a: address_bits) are scrambled before each sort by a
    QRAM load b ^= M[a] with M uniform from splitmix64.  mode 0: whole entries
    through prefix passes, then the LDS chunk sort; mode 1: LSD passes over an
    index payload and the permutation gather.  Reports each sort kernel's
    algorithmic bytes / time (HIP events on the launch stream) against the HBM peak."""
    ps.set_sort_mode(mode)
    from sparqsim_amd.workloads import qram_memory

    ps.System.clear()
    st = ps.SparseState()
    ps.AddRegister("b", ps.UnsignedInteger, 24)(st)
    ps.AddRegister("a", ps.UnsignedInteger, address_bits)(st)
    q = ps.QRAMCircuit_qutrit(address_bits, 24, qram_memory(address_bits, 24, seed_offset=7))
    ps.Hadamard_Int("a", address_bits)(st)
    load, sort = ps.QRAMLoad(q, "a", "b"), ps.SortUnconditional()

    def once():
        load(st)
        sort(st)
        st.size()

    for _ in range(3):
        once()
    lib.sparq_prof_reset()
    ps.sort_stats(reset=True)
    lib.sparq_prof_enable(1)
    for _ in range(reps):
        once()
    lib.sparq_prof_enable(0)
    prof = kernel_profile(lib)
    paths = ps.sort_stats(reset=True)
    ps.set_sort_mode(0)
    rows = [r for r in prof if r["name"] in SORT_KERNELS and r["ms"] > 0]
    ms = sum(r["ms"] for r in rows)
    byt = sum(r["bytes"] for r in rows)
    n = st.size()
    ps.System.clear()
    return {
        "workload": f"SortUnconditional of 2^{address_bits} entries, 48-bit keys scrambled by a QRAM load before each "
                    f"sort, {reps} sorts, " + ("prefix passes + LDS chunk sort" if mode == 0 else "LSD passes + gather"),
        "entries": n, "sort_paths": paths, "ms_per_sort": ms / reps, "entries_per_s": n * reps / (ms * 1e-3),
        "GBs": byt / (ms * 1e-3) / 1e9, "frac_of_peak": byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "kernels": [{"name": r["name"], "launches_per_sort": r["launches"] / reps, "us_per_launch": r["ms"] / r["launches"] * 1e3,
                     "GBs": r["bytes"] / (r["ms"] * 1e-3) / 1e9,
                     "frac_of_peak": r["bytes"] / (r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS} for r in rows],
    }


def cks_leg(ps, nqubit, terms, R, band_terms=0):
    """The 2^nqubit Poiseuille CKS walk (direct tridiagonal oracles, fitted
    coefficients J = 640): LCU terms timed for one right-hand side and for R in
    one state (entries from the walk itself; inputs built on the device)."""
    from sparqsim_amd import cks_solve

    out = {"workload": f"CKS Chebyshev walk, tridiag(-1, 2.2666666666564934, -1) at 2^{nqubit}, kappa 16, "
                       f"ds 32, J 640, {terms} LCU terms (2 walk steps each) after FirstStep",
           "reference_published": "2^20 on A800: ~4.2 s per LCU term, 20,971,512 entries, 2681.9 s to converge "
                                  "(docs/verification/v1e-cks-direct-oracle-linsolver.md:136-145)"}
    for r in (1, R):
        res = cks_solve.solve(ps, nqubit=nqubit, R=r, terms=terms, log_every=10 ** 9, log=lambda s: None)
        entries = res["entries_after_first_step"]
        out[f"rhs{r}"] = {"s_per_term": res["s_per_term"], "s_per_rhs_term": res["s_per_term"] / r,
                          "entries": entries, "walk_entry_steps_per_s": entries * 2 / res["s_per_term"]}
        ps.System.clear()
        ps.release_cached_memory()
    out["batched_speedup_per_rhs"] = out["rhs1"]["s_per_term"] / out[f"rhs{R}"]["s_per_rhs_term"]
    if band_terms > 0:
        t0 = time.perf_counter()
        res = cks_solve.solve(ps, nqubit=nqubit, R=R, kappa=50.0, J=1024, terms=band_terms, log_every=10 ** 9,
                              log=lambda s: None, matrix="band-random")
        entries = res["entries_after_first_step"]
        out[f"band_random_rhs{R}"] = {
            "workload": f"seeded band-plus-random symmetric positive matrix at 2^{nqubit} (diagonal, +-1 band, one "
                        f"random partner per row; kappa 50), QRAM-backed sparse-access oracles (binary search), "
                        f"{R} right-hand sides in one state, {band_terms} LCU terms",
            "s_per_term": res["s_per_term"], "s_per_rhs_term": res["s_per_term"] / R, "entries": entries,
            "walk_entry_steps_per_s": entries * 2 / res["s_per_term"], "setup_s": res["setup_s"],
            "wall_s": time.perf_counter() - t0}
        ps.System.clear()
        ps.release_cached_memory()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--address-bits", type=int, default=20)
    ap.add_argument("--data-bits", type=int, default=32)
    ap.add_argument("--cpu-address-bits", type=int, default=20, help="size of the CPU-baseline sample")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="time budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--autotune-passes", type=int, default=10,
                    help="untimed passes before the warm-up, for the launch-shape autotuner's first calls")
    ap.add_argument("--qda-qubits", type=int, default=20)
    ap.add_argument("--qda-steps", type=int, default=40, help="walk steps timed in the QDA leg (0 skips it)")
    ap.add_argument("--qda-total-steps", type=int, default=40568, help="length of the schedule s = n/total")
    ap.add_argument("--qda-solve-steps", type=int, default=40, help="complete solve on a schedule of this length (0 skips)")
    ap.add_argument("--grover-traj", type=int, default=1024, help="trajectories of the batched Grover leg (0 skips it)")
    ap.add_argument("--grover-bits", type=int, default=14)
    ap.add_argument("--grover-iterations", type=int, default=40)
    ap.add_argument("--grover-plain-traj", type=int, default=16, help="trajectories timed one state at a time")
    ap.add_argument("--qda-sparse-qubits", type=int, default=16)
    ap.add_argument("--qda-sparse-steps", type=int, default=40, help="steps timed on the sparse-encoding QDA (0 skips)")
    ap.add_argument("--cks-qubits", type=int, default=20)
    ap.add_argument("--cks-terms", type=int, default=6, help="LCU terms timed in the CKS leg (0 skips it)")
    ap.add_argument("--cks-rhs", type=int, default=16, help="right-hand sides of the batched CKS run")
    ap.add_argument("--cks-band-terms", type=int, default=3,
                    help="LCU terms timed on the band-plus-random matrix with the batched right-hand sides (0 skips)")
    ap.add_argument("--sort-bits", type=int, default=24, help="entries (log2) of the sort leg (0 skips it)")
    ap.add_argument("--sort-reps", type=int, default=6)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 512 MiB device copies (the reference kernel for scaling FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.cpu_baseline_child:   # child of cpu_baseline_process: CPU only
        from sparqsim_amd.workloads import qram_memory

        eps, secs, reps = cpu_baseline(args.cpu_address_bits, args.data_bits,
                                       qram_memory(args.cpu_address_bits, args.data_bits, seed_offset=1),
                                       args.cpu_seconds, args.cpu_baseline_child)
        print(json.dumps({"entries_per_s": eps, "seconds": secs, "passes": reps}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this engine has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch

        a = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": a.numel() * 8, "copy_bytes_written_per_launch": a.numel() * 8}))
        return

    import sparqsim_amd as ps
    from sparqsim_amd.workloads import QRAMInterference, qram_memory

    A, D = args.address_bits, args.data_bits
    lib = ctypes.CDLL(ps.ENGINE_LIBRARY)
    mem = qram_memory(A, D, seed_offset=1)
    ps.System.clear()
    st = ps.SparseState()
    ps.AddRegister("addr", ps.UnsignedInteger, A)(st)
    ps.AddRegister("data", ps.UnsignedInteger, D)(st)
    q = ps.QRAMCircuit_qutrit(A, D, mem)
    circ = QRAMInterference(ps, q, A, D)
    ops = circ.entry_ops
    print(f"# device: {ps.device_name()}", file=sys.stderr, flush=True)

    def step():
        ps.reset_systems(st)          # |0,0>, amplitude 1
        circ(st)
        return st.size()              # synchronizes the state's stream

    # the launch-shape autotuner times each candidate twice on the first calls
    # of a size class: those calls run before the warm-up
    for _ in range(args.autotune_passes):
        step()
    for _ in range(args.warmup):
        step()
    lib.sparq_prof_reset()
    lib.sparq_prof_enable(1)
    for _ in range(args.steps):
        step()
    lib.sparq_prof_enable(0)
    prof = kernel_profile(lib)
    # timed region (no event bracketing)
    t2 = time.perf_counter()
    for _ in range(args.steps):
        n_out = step()
    t3 = time.perf_counter()

    ms_per_step = (t3 - t2) / args.steps * 1e3
    value = ops / (ms_per_step / 1e3)
    dom = prof[0]
    avg_ms = dom["ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    kernels_ms = sum(r["ms"] for r in prof) / args.steps
    hbm_total = sum(r["bytes"] for r in prof) / args.steps
    result = {
        "metric": METRIC, "value": value, "unit": "entry-ops/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (splitmix64 seed 0x5EED2026+1)",
        "config": {"workload": f"single-state QRAM interference, 2^{A} entries ({A}-bit addr, {D}-bit data, "
                               f"2^{A} uint32 memory words), H->QRAMLoad->{D}xZ->QRAMLoad->H",
                   "entries": 1 << A, "entry_ops_per_step": ops, "final_entries": n_out, "parallelism": "single GPU",
                   "value_counts": "entry-operations: the entries each of the 36 reference operators touches, per second",
                   "entries_per_s_per_pass": (1 << A) / (ms_per_step / 1e3),
                   "autotune_passes": args.autotune_passes},
        "roofline": {"bound": "hbm", "kernel": dom["name"], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (PMC_TRAFFIC_MIB_2P20[dom["name"]] * 2**20
                                 if (A, D) == (20, 32) and dom["name"] in PMC_TRAFFIC_MIB_2P20 else None),
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms},
        "step_hbm_GBs": hbm_total / (ms_per_step * 1e-3) / 1e9,
        "kernel_ms_per_step": kernels_ms,
        "kernels": [{"name": r["name"], "launches_per_step": r["launches"] / args.steps,
                     "ms_per_step": r["ms"] / args.steps,
                     "GBs": (r["bytes"] / (r["ms"] * 1e-3) / 1e9) if r["ms"] > 0 else None} for r in prof[:12]],
    }
    if args.qda_steps > 0:
        result["qda_solve"] = qda_walk_leg(ps, args.qda_qubits, args.qda_steps, args.qda_total_steps)
    if args.qda_solve_steps > 0:
        from sparqsim_amd import qda_solve

        t0 = time.perf_counter()
        r = qda_solve.solve(ps, nqubit=args.qda_qubits, max_steps=args.qda_solve_steps, cz_eps=1e-16, log=lambda s: None)
        result["qda_compressed_solve"] = {
            "workload": f"complete 2^{args.qda_qubits} Poiseuille QDA solve on a {r['steps_run']}-step schedule "
                        "(walk + ClearZero(1e-16) + Normalize per step, as PoiseuilleQDABench)",
            "wall_s": time.perf_counter() - t0, "ms_per_step": r["ms_per_step"], "final_res_dir": r["final_res_dir"],
            "final_fidelity": r["final_fidelity"], "final_entries": r["peak_logged_entries"],
            "reference_published": "2^20, 40-step, czEps 1e-16: res_dir 5.178906e-3, fid 0.999999 "
                                   "(docs/verification/v1d-qda-linsolver-a800.md:123)",
        }
    if args.qda_sparse_steps > 0:
        result["qda_sparse"] = qda_sparse_leg(ps, args.qda_sparse_qubits, args.qda_sparse_steps)
    if args.cks_terms > 0:
        result["cks"] = cks_leg(ps, args.cks_qubits, args.cks_terms, args.cks_rhs, args.cks_band_terms)
    if args.sort_bits > 0:
        result["sort"] = sort_leg(ps, lib, args.sort_bits, args.sort_reps, mode=0)
        result["sort_lsd"] = sort_leg(ps, lib, args.sort_bits, args.sort_reps, mode=1)
    if args.grover_traj > 0:
        result["noisy_grover"] = grover_leg(ps, args.grover_traj, args.grover_bits, args.grover_iterations, 1e-3,
                                            min(args.grover_plain_traj, args.grover_traj))
    if not args.no_cpu_baseline:
        # BASELINE mode (b): the reference's OpenMP loops on the host's cores (sorts stay
        # sequential, as in the reference); mode (a), one thread as shipped, beside it
        ca = args.cpu_address_bits
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        half = args.cpu_seconds / 2
        eps_omp, secs_omp, reps_omp = cpu_baseline_process("_oracle_omp", threads, ca, half)
        eps_1, secs_1, reps_1 = cpu_baseline_process("_oracle", 1, ca, half)
        model = cpu_model()
        result["cpu_baseline"] = {"value": eps_omp, "unit": "entry-ops/s", "cores": threads, "kind": "port",
                                  "sample": f"same circuit at 2^{ca} entries, {reps_omp} passes in {secs_omp:.1f} s, "
                                            f"oracle (restated reference CPU path) with the reference's OpenMP loops, "
                                            f"{threads} threads on {model}"}
        result["cpu_baseline_single_thread"] = {
            "value": eps_1, "unit": "entry-ops/s", "cores": 1, "kind": "port",
            "sample": f"same circuit at 2^{ca} entries, {reps_1} passes in {secs_1:.1f} s, one thread "
                      f"(the reference as shipped, SINGLE_THREAD) on {model}"}
    result["autotune"] = ps.autotune_report()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
