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


def cpu_baseline(w, budget_s):
    """Oracle (numpy restatement of the reference algorithm) on this host."""
    from threadpoolctl import threadpool_limits
    from oracle import mps_ref
    tape = [(g, q, (w['x'][0][p] if p is not None else None)) for g, q, p in w['ops']]
    done, t0 = 0, time.perf_counter()
    with threadpool_limits(limits=1):
        while True:
            sim = mps_ref.RefSimulator(w['n'], max_bond=16, abs_cut=0.0, rel_cut=0.0).run(tape)
            sim.observe(w['words'], w['coeffs'])
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
    dt = time.perf_counter() - t0
    return dict(value=done / dt, unit="energy evals/s", cores=1, kind="port",
                sample=f"{done} full energy evaluation(s) of instance 0 (40 q, 4000 terms) by the numpy "
                       f"oracle (explicit-SVD MPS + chunk-128 prefix DAG), BLAS limited to 1 thread, {dt:.1f} s")


def measured_traffic():
    """HBM bytes per step of the chain kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(ROOT, 'profiles', 'rchains_energy_spsa40q_pmc.txt')
    try:
        with open(path) as f:
            for line in f:
                if line.startswith('hbm_bytes_per_step:'):
                    return float(line.split(':')[1])
    except OSError:
        pass
    return None


def cpu_grad_baseline(w, budget_s):
    """Torch-CPU restatement of the reference's reverse engine on this host."""
    import torch
    from oracle import reverse_ref
    torch.set_num_threads(1)
    done, t0 = 0, time.perf_counter()
    while True:
        reverse_ref.value_gradient(w['n'], w['ops'], w['words'], w['coeffs'], w['x'][done % len(w['x'])])
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return dict(value=done / dt, unit="adjoint grads/s", cores=1, kind="port",
                sample=f"{done} value+gradient evaluation(s) (40 q, 160 params, 4000 terms) by the torch-CPU "
                       f"restatement of the reverse engine (bond doubling, chunk 128), 1 thread, {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--plain-sample', type=int, default=32, help='circuits timed one per launch')
    ap.add_argument('--cpu-budget', type=float, default=15.0, help='seconds of CPU oracle work')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-grad', action='store_true')
    ap.add_argument('--no-uneven', action='store_true')
    ap.add_argument('--uneven-reps', type=int, default=5)
    ap.add_argument('--grad-steps', type=int, default=3)
    ap.add_argument('--grad-checkpoint', type=int, default=1,
                    help='40-qubit adjoint: keep every k-th forward step environment (the rest are recomputed)')
    ap.add_argument('--plain-grad-sample', type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps(dict(error="this build measures one GPU; --gpus must be 1",
                              requested_gpus=args.gpus)))
        sys.exit(2)

    import numpy as np
    import torch
    from gvmps import Batch, Problem, workloads

    torch.cuda.set_device(0)
    w = workloads.config(1)
    B = len(w['x'])
    prob = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0.0, rel_cut=0.0)
    info = prob.info()
    batch = Batch([prob], [B])
    x = torch.tensor(w['x'], dtype=torch.float64, device='cuda')
    for _ in range(args.warmup):
        e = batch.energies(x)
    torch.cuda.synchronize()
    batch.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e = batch.energies(x)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    phase_ms, calls = batch.timing()
    batched_rate = B * args.steps / dt

    # plain comparator: the same kernels, one circuit per launch sequence
    plain_rate, same = plain_rate_of(lambda b, v: b.energies(v), Batch([prob], [1]), x[: args.plain_sample],
                                     lambda i, r: torch.equal(r, e[i:i + 1]))

    # dominant kernel: forest chains (all generations of one step)
    chains_ms = phase_ms[2] / max(calls, 1)
    # algorithmic flops: 16 (Dl^2 Dr + Dl Dr^2) per forest node in complex
    # arithmetic (prefix and suffix forests, the planner's node count); the
    # circuit is real (RY + CX, Z-strings), so the forests run in real
    # arithmetic and the count is 4 (Dl^2 Dr + Dl Dr^2): a quarter
    real = bool(info['real_arith'])
    arith_div = 4.0 if real else 1.0
    node_flops = info['flops_energy_nodes'] / arith_div
    achieved = (node_flops * B) / (chains_ms * 1e-3) / 1e12
    executed = (info['flops_energy_executed'] / arith_div * B) / (chains_ms * 1e-3) / 1e12

    # the same workload on the complex kernels (the path of complex circuits)
    complex_leg = None
    if real:
        batch.set_arith('complex')
        e_cx = batch.energies(x)
        torch.cuda.synchronize()
        batch.set_timing(False)
        batch.set_timing(True)
        t1 = time.perf_counter()
        for _ in range(args.steps):
            e_cx = batch.energies(x)
        torch.cuda.synchronize()
        cx_rate = B * args.steps / (time.perf_counter() - t1)
        cx_ms, cx_calls = batch.timing()
        cx_chains = cx_ms[2] / max(cx_calls, 1)
        cx_flops = info['flops_energy_nodes'] * B / (cx_chains * 1e-3) / 1e12
        complex_leg = {"evals_per_s": round(cx_rate, 2), "forest_chains_ms": round(cx_chains, 3),
                       "achieved_tflops": round(cx_flops, 3), "frac": round(cx_flops / PEAK_FP64_TFLOPS, 4),
                       "max_rel_diff_vs_real": float(torch.max(torch.abs(e_cx - e) / torch.clamp(torch.abs(e), min=1.0))),
                       "kernel": ("k_chains_energy_iz" if info['iz_forest'] else "k_chains_energy")
                                 + " + k_leaf_energy (complex128, Gauss-form MFMA products)"}
        batch.set_arith('auto')
        batch.set_timing(False)

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
        "phase_ms_per_step": {"state_prep": round(phase_ms[0] / max(calls, 1), 3),
                              "forest_chains": round(chains_ms, 3),
                              "reduce": round(phase_ms[3] / max(calls, 1), 3)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                     "traffic": measured_traffic(),
                     "traffic_basis": "HBM bytes per step (all generations), rocprofv3 PMC, copy-calibrated",
                     "kernel": (("k_rchains_energy_iz (real arithmetic, I/Z-only forest)" if info['iz_forest']
                                 else "k_rchains_energy (real arithmetic)") if real
                                else "k_chains_energy_iz" if info['iz_forest'] else "k_chains_energy")
                               + " (all generations of one step) + leaf term-end kernel",
                     "algorithmic_flops_per_step": node_flops * B,
                     "executed_tflops": round(executed, 3)},
        "complex_path": complex_leg,
    }
    if grad is not None:
        out["adjoint"] = grad
        out["adjoint_60q"] = bench_adjoint60(args)
    if not args.no_uneven:
        out["uneven_batch"] = bench_uneven(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
        if grad is not None:
            out["adjoint"]["cpu_baseline"] = cpu_grad_baseline(w, args.cpu_budget)
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
    from gvmps import Batch
    batch = Batch([prob], [B])
    # every k-th forward step keeps its input environment; the backward pass
    # recomputes the rest (bitwise the same gradients at every stride)
    batch.set_checkpoint(args.grad_checkpoint)
    for _ in range(args.warmup):
        batch.value_grad(x)
    torch.cuda.synchronize()
    batch.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.grad_steps):
        e, g = batch.value_grad(x)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    phase_ms, calls = batch.timing()
    rate = B * args.grad_steps / dt
    del batch
    torch.cuda.empty_cache()
    single = Batch([prob], [1])
    single.set_checkpoint(args.grad_checkpoint)
    plain_rate, same = plain_rate_of(
        lambda b, v: b.value_grad(v), single, x[: args.plain_grad_sample],
        lambda i, r: torch.equal(r[0], e[i:i + 1]) and torch.equal(r[1], g.view(B, -1)[i]))
    c = max(calls, 1)
    names = ["bond_doubling_build", "forward_chains", "energy_reduce", "backward_chains", "site_reduce",
             "local_backward"]
    fwd_ms, bwd_ms = phase_ms[1] / c, phase_ms[3] / c
    # complex count: 4 real products per complex product; the real path
    # (real tape, no Y words) executes one
    real = bool(info['real_arith'])
    flops = info['flops_grad_executed'] * B / (4.0 if real else 1.0)
    achieved = flops / ((fwd_ms + bwd_ms) * 1e-3) / 1e12
    return {
        "value": round(rate, 2), "unit": "adjoint grads/s", "steps": args.grad_steps,
        "ms_per_step": round(dt / args.grad_steps * 1e3, 3),
        "plain_grads_per_s": plain_rate and round(plain_rate, 2),
        "batched_vs_plain": plain_rate and round(rate / plain_rate, 2),
        "plain_matches_batched_bitwise": same,
        "phase_ms_per_step": {n: round(phase_ms[i] / c, 3) for i, n in enumerate(names)},
        "arithmetic": "real" if real else "complex",
        "checkpoint_stride": args.grad_checkpoint,
        "chains_tflops": round(achieved, 3),
        "chains_frac_of_fp64_peak": round(achieved / PEAK_FP64_TFLOPS, 4),
    }


if __name__ == '__main__':
    main()
