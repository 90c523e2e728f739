"""Launch-shape autotuner for a batch.

The forest-chain kernels run one wavefront per (instance, chain).  Two knobs
change only how that work is packed onto the chip, never the results: the
wavefronts per workgroup (1, 2 or 4) and the work-item order (instance-major,
or the longest chains first within each XCD's share of the instance-major
items, which shortens a generation's tail while every XCD keeps the same
instances in its L2).  The
tuner times every combination on the batch's own workload with HIP events on
the launch stream and keeps the fastest.  Choices are cached per batch
signature, in memory and optionally in a JSON file (GVMPS_TUNE_CACHE).
"""
import json
import os

CONFIGS = [(wpb, order) for order in (0, 1) for wpb in (4, 2, 1)]
_memory = {}


def signature(batch):
    parts = []
    for p, c in zip(batch.problems, batch.counts):
        info = p.info()
        parts.append(f"{info['n_qubits']}x{info['n_steps']}x{info['n_chains']}x{c}")
    return '|'.join(parts)


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(batch, params, shots=0, seeds=None, reps=3, configs=CONFIGS, use_cache=True):
    """Pick (waves_per_group, order) for `batch`; returns (choice, {config: ms})."""
    import torch
    key = signature(batch) + f'|shots={shots}'
    path = os.environ.get('GVMPS_TUNE_CACHE')
    if use_cache:
        hit = _memory.get(key)
        if hit is None and path:
            hit = _load(path).get(key)
        if hit is not None:
            batch.set_launch(*hit)
            return tuple(hit), {}
    times = {}
    stream = torch.cuda.current_stream()
    for cfg in configs:
        batch.set_launch(*cfg)
        batch.evaluate(params, shots, seeds)  # warm
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(reps):
            batch.evaluate(params, shots, seeds)
        end.record(stream)
        end.synchronize()
        times[cfg] = start.elapsed_time(end) / reps
    best = min(times, key=times.get)
    batch.set_launch(*best)
    _memory[key] = list(best)
    if path:
        table = _load(path)
        table[key] = list(best)
        tmp = path + '.tmp'
        with open(tmp, 'w') as f:
            json.dump(table, f)
        os.replace(tmp, path)
    return best, times
