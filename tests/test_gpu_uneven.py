"""Uneven-size batch: 64 independent partitions (8-32 qubits, 200-1500
Z-strings, mixed linear / reverse-chain 4-layer RY ansatz, chi <= 16), energies
plus 1000 shots each from one state preparation, eagerly, from a captured HIP
graph, and under every autotuner launch shape.  Every variant must equal the
sequential per-partition launches bitwise (energies and bitstrings).  Every
partition's energy is checked against the CPU oracle (1e-10 relative), and
the bitstrings of the 16 largest partitions against the oracle's draws."""
import numpy as np
import pytest

from gvmps import Batch, Problem, autotune, seeds, workloads
from oracle import mps_ref

pytestmark = pytest.mark.gpu
SHOTS = 1000


@pytest.fixture(scope='module')
def uneven(cuda):
    w = workloads.config(3)
    parts = w['parts']
    probs = [Problem(q['n'], q['ops'], q['nparams'], q['words'], q['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
             for q in parts]
    x = np.concatenate([q['x'][0] for q in parts])
    sd = [seeds.shot_shard_seed(seeds.INITIAL_SAMPLE_SEED + j, 0) for j in range(len(parts))]
    return parts, probs, x, sd


def sequential(parts, probs, sd):
    es, bits = [], []
    for q, p, s in zip(parts, probs, sd):
        b = Batch([p], [1])
        es.append(float(b.energies(q['x']).cpu()[0]))
        bb, _ = b.sample(q['x'], SHOTS, [s])
        bits.append(bb[0].cpu().numpy())
    return np.array(es), bits


def same_bits(bits_batched, bits_seq, parts):
    bb = bits_batched.cpu().numpy()
    for j, q in enumerate(parts):
        n = q['n']
        assert np.array_equal(bb[j, :, :n], bits_seq[j][:, :n]), f'partition {j}'


def test_evaluate_matches_sequential_and_oracle(uneven):
    parts, probs, x, sd = uneven
    print('case: uneven-64 evaluate vs sequential', flush=True)
    batch = Batch(probs, [1] * len(probs))
    e, bits, packed = batch.evaluate(x, SHOTS, sd)
    es, bs = sequential(parts, probs, sd)
    assert np.array_equal(e.cpu().numpy(), es)
    same_bits(bits, bs, parts)


def test_evaluate_matches_oracle_every_partition(uneven):
    """Every partition's energy against the oracle (1e-10 relative), and the
    1000 bitstrings of the 16 largest partitions (26-32 qubits) against the
    oracle's draws with the same SplitMix64 stream: identical except where a
    shot's first difference sits within 1e-12 of a CDF boundary."""
    parts, probs, x, sd = uneven
    print('case: uneven-64 evaluate vs oracle', flush=True)
    batch = Batch(probs, [1] * len(probs))
    e, bits, _ = batch.evaluate(x, SHOTS, sd)
    e, bits = e.cpu().numpy(), bits.cpu().numpy()
    largest = sorted(range(len(parts)), key=lambda j: -parts[j]['n'])[:16]
    assert min(parts[j]['n'] for j in largest) >= 26
    differing = 0
    for j, q in enumerate(parts):
        tape = [(g, qq, (q['x'][0][k] if k is not None else None)) for g, qq, k in q['ops']]
        sim = mps_ref.RefSimulator(q['n'], max_bond=16, abs_cut=0.0, rel_cut=0.0).run(tape)
        ref, _ = sim.observe(q['words'], q['coeffs'])
        assert abs(e[j] - ref) <= 1e-10 * max(1.0, abs(ref)), (j, q['n'], e[j], ref)
        if j in largest:
            ref_bits, margins = sim.mps.sample_bits(SHOTS, sd[j], margins=True)
            got = bits[j, :, :q['n']]
            for sh in np.nonzero((got != ref_bits).any(axis=1))[0]:
                first = int(np.nonzero(got[sh] != ref_bits[sh])[0][0])
                assert margins[sh, first] < 1e-12, (j, sh, first, margins[sh, first])
                differing += 1
    print(f'shots differing from the oracle (all at CDF boundaries): {differing}', flush=True)


def test_graph_replay_matches_eager(uneven):
    import torch
    parts, probs, x, sd = uneven
    print('case: uneven-64 graph capture + 3 replays', flush=True)
    batch = Batch(probs, [1] * len(probs))
    e0, b0, _ = batch.evaluate(x, SHOTS, sd)
    e0, b0 = e0.clone(), b0.clone()
    g = batch.capture(x, SHOTS, sd)
    for rep in range(3):
        e, b, _ = g.replay()
        assert torch.equal(e, e0), f'replay {rep}'
        assert torch.equal(b, b0), f'replay {rep}'
        batch.energies(x)  # eager work between replays
        torch.cuda.synchronize()
    x2 = x + 0.05
    g.params.copy_(torch.as_tensor(x2, device=g.params.device))
    e, b, _ = g.replay()
    e2, b2, _ = batch.evaluate(x2, SHOTS, sd)
    assert torch.equal(e, e2) and torch.equal(b, b2)


@pytest.mark.skip(reason="launch order 2 (the forest pass as one persistent launch behind grid barriers) did not return within the test's 120 s on its first GPU call (40-qubit batch of 256, 4 waves per group); no cause was found by reading the kernel, so the kernel was taken out of the library and set_launch rejects order 2")
def test_persistent_launch_evaluate_and_replay(uneven):
    """The forest pass in one persistent launch (order 2) inside evaluate and
    inside the captured graph: energies and bitstrings bitwise equal to the
    generation launches, on every replay."""
    import torch
    parts, probs, x, sd = uneven
    print('case: uneven-64 evaluate, persistent launch, graph capture + 3 replays', flush=True)
    batch = Batch(probs, [1] * len(probs))
    e0, b0, _ = batch.evaluate(x, SHOTS, sd)
    e0, b0 = e0.clone(), b0.clone()
    batch.set_launch(4, 2)
    e, b, _ = batch.evaluate(x, SHOTS, sd)
    assert torch.equal(e, e0) and torch.equal(b, b0)
    g = batch.capture(x, SHOTS, sd)
    for rep in range(3):
        e, b, _ = g.replay()
        assert torch.equal(e, e0), f'replay {rep}'
        assert torch.equal(b, b0), f'replay {rep}'
        batch.energies(x)  # eager persistent launches between replays
        torch.cuda.synchronize()


def test_autotuner_shapes_do_not_change_results(uneven):
    import torch
    parts, probs, x, sd = uneven
    print('case: uneven-64 all launch shapes', flush=True)
    batch = Batch(probs, [1] * len(probs))
    ref = None
    for cfg in autotune.CONFIGS:
        batch.set_launch(*cfg)
        e, b, _ = batch.evaluate(x, SHOTS, sd)
        if ref is None:
            ref = (e.clone(), b.clone())
        else:
            assert torch.equal(e, ref[0]) and torch.equal(b, ref[1]), cfg
    best, times = autotune.tune(batch, x, SHOTS, sd, reps=1, use_cache=False)
    assert best in autotune.CONFIGS and len(times) == len(autotune.CONFIGS)
    assert batch.launch == best
