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

For reverse gradients (tune_grad) the knobs are the wavefronts per workgroup
of the forward and backward chain launches and the checkpoint stride (every
stride-th forward step keeps its input environment, the backward pass
recomputes the rest): memory and recomputation traded against each other,
with gradients bitwise the same for every choice.  The launch shape is shared
by a batch's energy and gradient calls, so tune the call type it serves.
"""
import json
import os

CONFIGS = [(wpb, order) for order in (0, 1) for wpb in (4, 2, 1)]
GRAD_CONFIGS = [(wpb, stride) for stride in (1, 2) for wpb in (4, 2)]
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


def _cached(key):
    path = os.environ.get('GVMPS_TUNE_CACHE')
    hit = _memory.get(key)
    if hit is None and path:
        hit = _load(path).get(key)
    return hit


def _remember(key, best):
    _memory[key] = list(best)
    path = os.environ.get('GVMPS_TUNE_CACHE')
    if path:
        table = _load(path)
        table[key] = list(best)
        tmp = path + '.tmp'
        with open(tmp, 'w') as f:
            json.dump(table, f)
        os.replace(tmp, path)


def _time(call, reps):
    """Milliseconds per call, HIP events on the current stream after a warm call."""
    import torch
    call()
    stream = torch.cuda.current_stream()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        call()
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / reps


def tune(batch, params, shots=0, seeds=None, reps=3, configs=CONFIGS, use_cache=True):
    """Pick (waves_per_group, order) for `batch`; returns (choice, {config: ms})."""
    key = signature(batch) + f'|shots={shots}'
    if use_cache:
        hit = _cached(key)
        if hit is not None:
            batch.set_launch(*hit)
            return tuple(hit), {}
    times = {}
    for cfg in configs:
        batch.set_launch(*cfg)
        times[cfg] = _time(lambda: batch.evaluate(params, shots, seeds), reps)
    best = min(times, key=times.get)
    batch.set_launch(*best)
    _remember(key, best)
    return best, times


def tune_grad(batch, params, reps=2, configs=GRAD_CONFIGS, use_cache=True):
    """Pick (waves_per_group, checkpoint stride) for `batch`'s value_grad
    calls (longest-first item order); returns (choice, {config: ms})."""
    key = signature(batch) + '|grad'
    if use_cache:
        hit = _cached(key)
        if hit is not None:
            batch.set_launch(hit[0], 1)
            batch.set_checkpoint(hit[1])
            return tuple(hit), {}
    times = {}
    for wpb, stride in configs:
        batch.set_launch(wpb, 1)
        batch.set_checkpoint(stride)
        times[(wpb, stride)] = _time(lambda: batch.value_grad(params), reps)
    best = min(times, key=times.get)
    batch.set_launch(best[0], 1)
    batch.set_checkpoint(best[1])
    _remember(key, best)
    return best, times
