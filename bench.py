#!/usr/bin/env python
"""Benchmark of the MI355X MPS backend on the 40-qubit SPSA workload.

Workload (one step): analytic energies of 256 parameter vectors (128 SPSA
+/- pairs) of a 40-qubit, 4-layer RY ansatz with the reversed CX chain
(CX(39,38) ... CX(1,0)), chi <= 16, zero SVD cutoffs, on a 4000-term Z-string
Hamiltonian (weights 1-6), all 256 circuits in one batched launch sequence.
The circuit is real, so the state is exactly real and the forests run in real
FP64 arithmetic; the complex128 kernels (the path of complex circuits) are
timed on the same workload beside it.  Inputs are resident in HBM before
timing starts.

Printed: one JSON line (rank 0, N=1) with the batched rate as `value`, the
plain one-circuit-per-launch rate beside it, the FP64 roofline of the
dominant kernel (forest chains) from HIP-event timing on the launch stream,
and the CPU oracle timed on a bounded sample on this host.  The `adjoint`
object reports the second half of the metric: value + dE/dx of the same 256
vectors by one batched reverse pass per step, the plain one-vector-per-call
rate, phase times and the torch-CPU reverse oracle on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'graph-vqe_amd'))
sys.path.insert(0, ROOT)

METRIC = "energy evals/s + adjoint grads/s at 40q chi=16, batched vs plain, 1 MI355X"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def _oracle_energy_worker(args):
    """Energy evaluations of the 40-qubit workload by the numpy oracle, one
    BLAS thread, for `budget` seconds (the batch's vectors round robin)."""
    worker, nworkers, budget = args
    from threadpoolctl import threadpool_limits
    from gvmps import workloads
    from oracle import mps_ref
    w = workloads.config(1)
    done, t0 = 0, time.perf_counter()
    with threadpool_limits(limits=1):
        while True:
            x = w['x'][(worker + done * nworkers) % len(w['x'])]
            tape = [(g, q, (x[p] if p is not None else None)) for g, q, p in w['ops']]
            sim = mps_ref.RefSimulator(w['n'], max_bond=16, abs_cut=0.0, rel_cut=0.0).run(tape)
            sim.observe(w['words'], w['coeffs'])
            done += 1
            if time.perf_counter() - t0 > budget:
                break
    return done, time.perf_counter() - t0


def _oracle_grad_worker(args):
    """Value + gradient evaluations by the torch-CPU reverse oracle, one
    thread, for `budget` seconds."""
    worker, nworkers, budget = args
    import torch
    from gvmps import workloads
    from oracle import reverse_ref
    torch.set_num_threads(1)
    w = workloads.config(1)
    done, t0 = 0, time.perf_counter()
    while True:
        x = w['x'][(worker + done * nworkers) % len(w['x'])]
        reverse_ref.value_gradient(w['n'], w['ops'], w['words'], w['coeffs'], x)
        done += 1
        if time.perf_counter() - t0 > budget:
            break
    return done, time.perf_counter() - t0


def cpu_cores():
    """Worker processes of the all-cores baseline: the affinity set, capped by
    the host's thread budget for one GPU (OMP_NUM_THREADS where it is set),
    less this process, which stays alive beside the workers (a process that
    imports torch counts as a user of the GPU, and one GPU admits 16)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get('OMP_NUM_THREADS')
    if cap and cap.isdigit():
        n = min(n, int(cap))
    return max(1, min(n, 16) - 1) if n > 1 else 1


def _timed_pool(fn, nworkers, budget_s):
    """Throughput of `nworkers` single-threaded worker processes (started
    fresh, before this process touches the GPU) over about budget_s."""
    import multiprocessing as mp
    if nworkers == 1:
        done, dt = fn((0, 1, budget_s))
        return done, dt
    ctx = mp.get_context('spawn')
    t0 = time.perf_counter()
    with ctx.Pool(nworkers) as pool:
        res = pool.map(fn, [(i, nworkers, budget_s) for i in range(nworkers)])
    wall = time.perf_counter() - t0
    # each worker's own clock excludes its interpreter start-up
    rate = sum(d / t for d, t in res)
    return sum(d for d, _ in res), sum(d for d, _ in res) / rate if rate > 0 else wall


def cpu_baselines(budget_s, with_grad):
    """The CPU path timed on this host: the numpy oracle (explicit-SVD MPS +
    chunk-128 prefix DAG, the reference's algorithm) and the torch-CPU
    reverse oracle, on one core and on every core of the host (one
    single-threaded process per core, independent evaluations of the batch)."""
    out = {}
    ncores = cpu_cores()
    for key, fn, unit, what in (
            ('energy', _oracle_energy_worker, 'energy evals/s',
             'energy evaluation(s) of the 40 q, 4000-term workload by the numpy oracle (explicit-SVD MPS + '
             'chunk-128 prefix DAG)'),
            ('grad', _oracle_grad_worker, 'adjoint grads/s',
             'value+gradient evaluation(s) (40 q, 160 params, 4000 terms) by the torch-CPU restatement of the '
             'reverse engine (bond doubling, chunk 128)')):
        if key == 'grad' and not with_grad:
            continue
        d1, t1 = _timed_pool(fn, 1, budget_s)
        dn, tn = _timed_pool(fn, ncores, budget_s)
        out[key] = dict(value=dn / tn, unit=unit, cores=ncores, kind="port",
                        sample=f"{dn} {what}, {ncores} single-threaded processes, {budget_s:.0f} s each",
                        single_core=dict(value=d1 / t1, cores=1, sample=f"{d1} {what}, 1 thread, {t1:.1f} s"))
    return out


def timed_calls(fn, steps, batch):
    """`steps` calls of fn with the batch's HIP-event instrumentation off (the
    wall time: throughput), then `steps` more with it on (per-phase and
    per-launch device times: the breakdown and the roofline).  The events
    recorded around every launch cost GPU time of their own, so the two are
    measured apart.  Returns (last result, seconds per call, phase_ms, calls)."""
    import torch
    batch.set_timing(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = fn()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    batch.set_timing(True)
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    ph, calls = batch.timing()
    batch.set_timing(False)
    return r, dt, ph, calls


def measured_traffic(key='hbm_bytes_per_launch:'):
    """HBM bytes per launch of the chain kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(ROOT, 'profiles', 'rchains_energy_spsa40q_pmc.txt')
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split(':')[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--plain-sample', type=int, default=32, help='circuits timed one per launch')
    ap.add_argument('--cpu-budget', type=float, default=15.0, help='seconds of CPU oracle work')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-grad', action='store_true')
    ap.add_argument('--no-uneven', action='store_true')
    ap.add_argument('--uneven-reps', type=int, default=20)
    ap.add_argument('--grad-steps', type=int, default=10)
    ap.add_argument('--grad-checkpoint', type=int, default=0,
                    help='40-qubit adjoint: keep every k-th forward step environment (the rest are recomputed); '
                         '0 (default): the gradient autotuner picks it with the workgroup size')
    ap.add_argument('--plain-grad-sample', type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps(dict(error="this build measures one GPU; --gpus must be 1",
                              requested_gpus=args.gpus)))
        sys.exit(2)

    # the CPU path first, in fresh single-threaded processes, before this
    # process touches the GPU
    cpu = None if args.no_cpu else cpu_baselines(args.cpu_budget, not args.no_grad)

    import numpy as np
    import torch
    from gvmps import Batch, Problem, workcount, workloads

    torch.cuda.set_device(0)
    w = workloads.config(1)
    B = len(w['x'])
    prob = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0.0, rel_cut=0.0)
    info = prob.info()
    batch = Batch([prob], [B])
    x = torch.tensor(w['x'], dtype=torch.float64, device='cuda')
    for _ in range(args.warmup):
        e = batch.energies(x)
    e, dt_call, phase_ms, calls = timed_calls(lambda: batch.energies(x), args.steps, batch)
    dt = dt_call * args.steps
    batched_rate = B / dt_call

    # plain comparator: the same kernels, one circuit per launch sequence
    plain_rate, same = plain_rate_of(lambda b, v: b.energies(v), Batch([prob], [1]), x[: args.plain_sample],
                                     lambda i, r: torch.equal(r, e[i:i + 1]))

    # dominant kernel: the forest-chain kernel (one launch per generation,
    # 8 per 40-qubit step), timed per launch by HIP events around each launch on the launch
    # stream (slots 6, 7); the chains phase adds the site packing and the
    # leaf term-end launch
    chains_ms = phase_ms[2] / max(calls, 1)
    kern_ms, kern_launches = phase_ms[6], int(phase_ms[7])
    kern_ms_step = kern_ms / max(calls, 1)
    # executed MFMA flops of the chain kernel, counted by the library per step
    # of the last call: the products each step issues (one real product per
    # complex product on the real path; fewer where the sites' zero structure
    # lets a step skip all-zero row groups or merge its two products)
    real = bool(info['real_arith'])
    exec_per_step = batch.chain_flops()
    exec_per_eval = exec_per_step / B
    full_per_eval = info['flops_energy_executed'] * (0.25 if real else 0.75)
    achieved = exec_per_step / (kern_ms_step * 1e-3) / 1e12
    # the algorithm's own work: left/right environments with shared prefixes,
    # counted from the workload (gvmps.workcount); the kernels do less
    # arithmetic than that (real state, prefix and suffix forests with
    # sibling sharing): reported as a reduction, not as a roofline fraction
    wc = workcount.energy_work(w['n'], w['ops'], w['words'])
    algorithmic = {
        "flops_per_eval": wc['flops_per_eval'], "node_site_transfers": wc['node_site_transfers'],
        "reference_chunk128_dag_flops_per_eval": wc['reference_dag_flops_per_eval'],
        "executed_flops_per_eval": exec_per_eval,
        "executed_flops_per_eval_without_zero_skips": full_per_eval,
        "reduction_vs_algorithmic": round(wc['flops_per_eval'] / exec_per_eval, 2),
        "basis": "16 (Dl^2 Dr + Dl Dr^2) flops per environment transfer, N_s distinct prefixes per site over "
                 "the words' spans, plus the left and right identity environments",
    }
    traffic = measured_traffic()

    # the same workload on the complex kernels (the path of complex circuits)
    complex_leg = None
    if real:
        batch.set_arith('complex')
        batch.energies(x)
        e_cx, cx_dt, cx_ms, cx_calls = timed_calls(lambda: batch.energies(x), args.steps, batch)
        cx_rate = B / cx_dt
        cx_chains = cx_ms[2] / max(cx_calls, 1)
        cx_kern = cx_ms[6] / max(cx_calls, 1)
        cx_flops = batch.chain_flops() / (cx_kern * 1e-3) / 1e12
        complex_leg = {"evals_per_s": round(cx_rate, 2), "forest_chains_ms": round(cx_chains, 3),
                       "chain_kernel_ms": round(cx_kern, 3),
                       "achieved_tflops": round(cx_flops, 3), "frac": round(cx_flops / PEAK_FP64_TFLOPS, 4),
                       "flops_basis": "executed MFMA flops (3 real products per complex product)",
                       "max_rel_diff_vs_real": float(torch.max(torch.abs(e_cx - e) / torch.clamp(torch.abs(e), min=1.0))),
                       "kernel": ("k_chains_energy_iz" if info['iz_forest'] else "k_chains_energy")
                                 + " + k_leaf_energy (complex128, Gauss-form MFMA products)"}
        batch.set_arith('auto')

    # the same energies with the state built by exact CNOT bond doubling (no
    # SVD; the reverse engine's representation) instead of gates + SVDs
    batch.set_prep('doubling')
    e_dbl = batch.energies(x)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        e_dbl = batch.energies(x)
    torch.cuda.synchronize()
    dbl_rate = B * args.steps / (time.perf_counter() - t2)
    dbl_dev = float(torch.max(torch.abs(e_dbl - e) / torch.clamp(torch.abs(e), min=1.0)))
    # the same energies with the pivoted QR / SVD run at every gate, although
    # no split of this workload cuts rank (the default takes the exact split)
    batch.set_prep('factorize')
    batch.energies(x)
    e_fac, fac_dt, fac_ms, fac_calls = timed_calls(lambda: batch.energies(x), args.steps, batch)
    fac_rate = B / fac_dt
    fac_dev = float(torch.max(torch.abs(e_fac - e) / torch.clamp(torch.abs(e), min=1.0)))
    batch.set_prep('svd')

    grad = None if args.no_grad else bench_grads(args, prob, info, x, B)
    out = {
        "metric": METRIC,
        "value": round(batched_rate, 2),
        "unit": "energy evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if real else "complex128",
        "data": "synthetic",
        "config": {"workload": "spsa-40q: 40 qubits, 4x(RY layer + CX 39->38->...->0), chi<=16, zero cutoffs, "
                               "4000 Z-strings (weight 1-6) + constant, 256 SPSA vectors per step",
                   "batch": B, "qubits": w['n'], "terms": len(w['words']), "params": w['nparams']},
        "plain_evals_per_s": plain_rate and round(plain_rate, 2),
        "batched_vs_plain": plain_rate and round(batched_rate / plain_rate, 2),
        "plain_matches_batched_bitwise": same,
        "doubling_prep": {"evals_per_s": round(dbl_rate, 2), "max_rel_diff_vs_svd": dbl_dev,
                          "note": "same energies, state built by exact CNOT bond doubling instead of SVDs"},
        "factorized_prep": {"evals_per_s": round(fac_rate, 2), "max_rel_diff": fac_dev,
                            "state_prep_ms": round(fac_ms[0] / max(fac_calls, 1), 3),
                            "note": "same energies, pivoted QR (SVD gauge) at every gate instead of the exact split"},
        "exact_split": bool(info['exact_split']),
        "timing": "value and ms_per_step: uninstrumented calls; phase_ms_per_step and roofline: as many calls "
                  "again with HIP events around every launch on the launch stream",
        "phase_ms_per_step": {"state_prep": round(phase_ms[0] / max(calls, 1), 3),
                              "forest_chains": round(chains_ms, 3),
                              "chain_kernel": round(kern_ms_step, 3),
                              "reduce": round(phase_ms[3] / max(calls, 1), 3)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                     "traffic": traffic,
                     "traffic_basis": "HBM bytes per launch of the chain kernel (mean of its generation launches), "
                                      "rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated",
                     "kernel": ("k_rchains_energy_izm" if info['iz_forest'] else "k_rchains_energy") if real
                               else ("k_chains_energy_iz" if info['iz_forest'] else "k_chains_energy"),
                     "launches_per_step": kern_launches // max(calls, 1),
                     "avg_launch_us": round(kern_ms / max(kern_launches, 1) * 1e3, 1),
                     "flops_basis": "executed MFMA flops of the kernel per step, counted by the library "
                                    "(gvmps_batch_chain_flops): " + (
                         "one real 16x16x16 product per complex product (real state), fewer where a step "
                         "merges its two products or skips all-zero row groups of a site (merged site "
                         "fragments)" if real else
                         "three real products per complex product (Gauss form)"),
                     "executed_flops_per_step": exec_per_step},
        "algorithmic": algorithmic,
        "complex_path": complex_leg,
    }
    if grad is not None:
        out["adjoint"] = grad
        out["adjoint_60q"] = bench_adjoint60(args)
    if not args.no_uneven:
        out["uneven_batch"] = bench_uneven(args)
        out["sampling"] = bench_sampling(args)
    if cpu is not None:
        out["cpu_baseline"] = cpu['energy']
        if grad is not None and 'grad' in cpu:
            out["adjoint"]["cpu_baseline"] = cpu['grad']
    print(json.dumps(out), flush=True)


def bench_adjoint60(args):
    """60 qubits, 4 layers of RY+RZ with linear CX chains (480 parameters),
    8000 Z-strings, 32 vectors per reverse pass, step environments
    checkpointed every other step (the rest recomputed)."""
    import torch
    from gvmps import Batch, Problem, workloads
    w = workloads.config(2)
    prob = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0.0, rel_cut=0.0)
    info = prob.info()
    x = torch.tensor(w['x'], dtype=torch.float64, device='cuda')
    B = x.shape[0]
    batch = Batch([prob], [B])
    batch.set_checkpoint(2)
    batch.value_grad(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.grad_steps):
        batch.value_grad(x)
    torch.cuda.synchronize()
    rate = B * args.grad_steps / (time.perf_counter() - t0)
    del batch
    torch.cuda.empty_cache()
    single = Batch([prob], [1])
    single.set_checkpoint(2)
    plain, _ = plain_rate_of(lambda b, v: b.value_grad(v), single, x[:4], lambda i, r: True)
    env_bytes = info['n_steps'] * 4096
    return {"workload": "adjoint-60q: 60 qubits, 4x(RY,RZ + linear CX), 480 params, 8000 Z-strings, batch 32",
            "grads_per_s": round(rate, 2), "plain_grads_per_s": plain and round(plain, 2),
            "batched_vs_plain": plain and round(rate / plain, 2), "checkpoint_stride": 2,
            "step_env_bytes_per_instance_without_checkpoints": env_bytes}


def bench_uneven(args):
    """64 partitions (8-32 qubits, 200-1500 terms): energies + 1000 shots each
    from one captured HIP graph (autotuned launch shape) vs sequential
    per-partition energy + sample calls."""
    import numpy as np
    import torch
    from gvmps import Batch, Problem, autotune, seeds, workloads
    parts = workloads.config(3)['parts']
    probs = [Problem(q['n'], q['ops'], q['nparams'], q['words'], q['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
             for q in parts]
    x = np.concatenate([q['x'][0] for q in parts])
    sd = [seeds.shot_shard_seed(seeds.INITIAL_SAMPLE_SEED + j, 0) for j in range(len(parts))]
    shots = 1000
    batch = Batch(probs, [1] * len(probs))
    best, table = autotune.tune(batch, x, shots, sd, reps=2, use_cache=False)
    graph = batch.capture(x, shots, sd)
    graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.uneven_reps):
        graph.replay()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) / args.uneven_reps * 1e3
    singles = [Batch([p], [1]) for p in probs]
    for b, q, s in zip(singles, parts, sd):  # warm
        b.energies(q['x'])
        b.sample(q['x'], shots, [s])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    seq_e = []
    for b, q, s in zip(singles, parts, sd):
        seq_e.append(b.energies(q['x']))
        b.sample(q['x'], shots, [s])
    torch.cuda.synchronize()
    seq_ms = (time.perf_counter() - t1) * 1e3
    same = bool(torch.equal(torch.cat(seq_e), graph.energies))
    ns = [int(q['n']) for q in parts]
    return {"partitions": len(parts), "qubits": f"{min(ns)}-{max(ns)}",
            "shots_per_partition": shots, "graph_ms": round(graph_ms, 3), "sequential_ms": round(seq_ms, 3),
            "speedup": round(seq_ms / graph_ms, 2), "energies_match_sequential_bitwise": same,
            "autotuned_launch": {"waves_per_group": best[0], "order": ["instance-major", "longest-first-per-xcd"][best[1]]},
            "tuning_ms": {f"{k[0]}w/{'len' if k[1] else 'inst'}": round(v, 3) for k, v in table.items()},
            "shots_per_s": round(len(parts) * shots / (graph_ms * 1e-3), 1)}


def bench_sampling(args):
    """Shots/s of perfect sampling: the 12-qubit configuration (one circuit,
    1000 shots per call, preparation included) and the 64 uneven partitions
    (1000 shots each, one batched call), with the FP64 roofline of the
    sampling kernels from the algorithmic count (per site: shots * 8 (2 Dl Dr
    + 2 Dr^2) for the conditional probabilities plus 8 (2 Dl Dr^2 + 2 Dl^2 Dr)
    for the right environment; a quarter of that in real arithmetic) and the
    numpy oracle's sampler on one core beside it."""
    import numpy as np
    import torch
    from gvmps import Batch, Problem, seeds, workcount, workloads
    out = {}
    reps = 20
    for name, parts in (('plain-12q', [workloads.config(0)]), ('uneven-64', workloads.config(3)['parts'])):
        probs = [Problem(q['n'], q['ops'], q['nparams'], q['words'], q['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
                 for q in parts]
        x = torch.tensor(np.concatenate([q['x'][0] for q in parts]), dtype=torch.float64, device='cuda')
        sd = [seeds.shot_shard_seed(seeds.INITIAL_SAMPLE_SEED + j, 0) for j in range(len(parts))]
        b = Batch(probs, [1] * len(probs))
        real = all(p.info()['real_arith'] for p in probs)
        shots = 1000
        b.sample(x, shots, sd)
        _, dt, ph, calls = timed_calls(lambda: b.sample(x, shots, sd), reps, b)
        c = max(calls, 1)
        envs_ms, sample_ms = ph[1] / c, ph[2] / c
        flops = sum(workcount.sampling_flops(workcount.allocated_bonds(q['n'], q['ops'])[0], shots) for q in parts)
        flops *= 0.25 if real else 1.0
        achieved = flops / ((envs_ms + sample_ms) * 1e-3) / 1e12
        out[name] = {"instances": len(parts), "shots_per_instance": shots, "ms_per_call": round(dt * 1e3, 3),
                     "shots_per_s": round(len(parts) * shots / dt, 1),
                     "phase_ms": {"state_prep": round(ph[0] / c, 3), "pack_and_right_envs": round(envs_ms, 3),
                                  "sample_kernel": round(sample_ms, 3)},
                     "arithmetic": "real" if real else "complex",
                     "roofline": {"bound": "latency (one wavefront per 16 shots, n dependent site steps)",
                                  "achieved_tflops": round(achieved, 4),
                                  "frac_of_fp64_peak": round(achieved / PEAK_FP64_TFLOPS, 5),
                                  "algorithmic_flops_per_call": flops}}
    if not args.no_cpu:
        from oracle import mps_ref  # the CPU baseline only
        w = workloads.config(0)
        tape = [(g, q, (w['x'][0][p] if p is not None else None)) for g, q, p in w['ops']]
        sim = mps_ref.RefSimulator(w['n'], max_bond=16, abs_cut=0.0, rel_cut=0.0).run(tape)
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            done = 0
            while time.perf_counter() - t0 < 3.0:
                sim.mps.sample_bits(1000, 12345 + done)
                done += 1
            dt = time.perf_counter() - t0
        out["plain-12q"]["cpu_baseline"] = {"value": round(done * 1000 / dt, 1), "unit": "shots/s", "cores": 1,
                                            "kind": "port", "sample": f"{done} x 1000 shots of the prepared 12-qubit "
                                                                      f"state by the numpy oracle, {dt:.1f} s"}
    return out


def plain_rate_of(call, single, xs, matches):
    """Rate of `call` on a one-instance batch over the vectors xs, one call
    each, and whether every result equals the batched one; (None, None) for an
    empty sample."""
    import torch
    if len(xs) == 0:
        return None, None
    call(single, xs[0])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    res = [call(single, xs[i]) for i in range(len(xs))]
    torch.cuda.synchronize()
    rate = len(xs) / (time.perf_counter() - t1)
    return rate, all(bool(matches(i, r)) for i, r in enumerate(res))


def bench_grads(args, prob, info, x, B):
    """Adjoint gradients (value + dE/dx) of the same 256 vectors: one batched
    reverse pass per step vs one vector per call."""
    import torch
    from gvmps import Batch, autotune
    batch = Batch([prob], [B])
    # every k-th forward step keeps its input environment; the backward pass
    # recomputes the rest (bitwise the same gradients at every stride); by
    # default the autotuner times strides 1, 2 x 4, 2 waves per workgroup
    tuned = None
    if args.grad_checkpoint > 0:
        batch.set_checkpoint(args.grad_checkpoint)
        stride, wpb = args.grad_checkpoint, 4
    else:
        (wpb, stride), table = autotune.tune_grad(batch, x, reps=2, use_cache=False)
        tuned = {"choice": {"waves_per_group": wpb, "checkpoint_stride": stride},
                 "ms": {f"{k[0]}w/stride{k[1]}": round(v, 3) for k, v in table.items()}}
    for _ in range(args.warmup):
        batch.value_grad(x)
    (e, g), dt_call, phase_ms, calls = timed_calls(lambda: batch.value_grad(x), args.grad_steps, batch)
    dt = dt_call * args.grad_steps
    rate = B / dt_call
    # the products the last value_grad call's chain launches executed (the
    # forward's zero-structure skips counted on the real path)
    flops = batch.grad_flops()
    del batch
    torch.cuda.empty_cache()
    single = Batch([prob], [1])
    single.set_checkpoint(stride)
    plain_rate, same = plain_rate_of(
        lambda b, v: b.value_grad(v), single, x[: args.plain_grad_sample],
        lambda i, r: torch.equal(r[0], e[i:i + 1]) and torch.equal(r[1], g.view(B, -1)[i]))
    c = max(calls, 1)
    names = ["bond_doubling_build", "forward_chains", "energy_reduce", "backward_chains", "site_reduce",
             "local_backward"]
    fwd_ms, bwd_ms = phase_ms[1] / c, phase_ms[3] / c
    real = bool(info['real_arith'])
    achieved = flops / ((fwd_ms + bwd_ms) * 1e-3) / 1e12
    return {
        "value": round(rate, 2), "unit": "adjoint grads/s", "steps": args.grad_steps,
        "ms_per_step": round(dt / args.grad_steps * 1e3, 3),
        "plain_grads_per_s": plain_rate and round(plain_rate, 2),
        "batched_vs_plain": plain_rate and round(rate / plain_rate, 2),
        "plain_matches_batched_bitwise": same,
        "phase_ms_per_step": {n: round(phase_ms[i] / c, 3) for i, n in enumerate(names)},
        "arithmetic": "real" if real else "complex",
        "checkpoint_stride": stride,
        "waves_per_group": wpb,
        "autotuned": tuned,
        "chains_tflops": round(achieved, 3),
        "chains_frac_of_fp64_peak": round(achieved / PEAK_FP64_TFLOPS, 4),
        "backward_kernel_ms_per_step": round(phase_ms[6] / c, 3),
    }


if __name__ == '__main__':
    main()
