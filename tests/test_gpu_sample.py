"""GPU parity of perfect sampling and the finite-shot estimator.

Bar: bitstrings identical to the oracle's for the same uniform stream, except
for draws within 1e-12 of a CDF boundary (where the first difference of a
shot must sit); shot energies equal to the oracle's estimator on the same bits.
"""
import numpy as np
import pytest

from gvmps import Batch, Problem, workloads
from oracle import dense, mps_ref, rng

pytestmark = pytest.mark.gpu


def tapev(ops, x):
    return [(g, q, (x[p] if isinstance(p, (int, np.integer)) else p)) for g, q, p in ops]


def assert_same_bits(gpu_bits, ref_bits, margins, boundary=1e-12):
    diff = np.nonzero((gpu_bits != ref_bits).any(axis=1))[0]
    for sh in diff:
        first = int(np.nonzero(gpu_bits[sh] != ref_bits[sh])[0][0])
        assert margins[sh, first] < boundary, (sh, first, margins[sh, first])
    return len(diff)


def test_12q_sampling_bit_identical(cuda):
    w = workloads.config(0)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    seed = rng.shot_shard_seed(rng.INITIAL_SAMPLE_SEED, 0)
    print('case: 12q 1000 shots', flush=True)
    bits, packed = Batch([p], [1]).sample(w['x'], 1000, [seed])
    bits = bits.cpu().numpy()[0]
    sim = mps_ref.RefSimulator(w['n'], 16, 0.0, 0.0).run(tapev(w['ops'], w['x'][0]))
    ref, margins = sim.mps.sample_bits(1000, seed, margins=True)
    assert assert_same_bits(bits, ref, margins) == 0
    # packed words carry the same bits (bit s of word 0 = site s)
    pk = packed.cpu().numpy()[0].view(np.uint64)
    for sh in range(0, 1000, 97):
        assert all(((int(pk[sh, 0]) >> s) & 1) == bits[sh, s] for s in range(w['n']))


@pytest.mark.parametrize('seed', range(3))
def test_complex_circuits_vs_dense_sequential_cdf(cuda, seed):
    r = np.random.default_rng(300 + seed)
    n = 7
    ops, P = [], 0
    for _ in range(50):
        kind = r.integers(0, 4)
        q = int(r.integers(0, n))
        if kind == 0:
            ops.append((str(r.choice(['rx', 'ry', 'rz'])), (q,), P))
            P += 1
        elif kind == 1:
            ops.append((str(r.choice(['h', 's', 'sdg'])), (q,), None))
        else:
            t = int(r.integers(0, n - 1))
            ops.append(('cx', (q, t + 1 if t >= q else t), None))
    x = r.uniform(-np.pi, np.pi, P)
    s = int(r.integers(0, 2 ** 32))
    print(f'case: sampling random circuit seed={seed}', flush=True)
    bits, _ = Batch([Problem(n, ops, P, ['Z' * n], [1.0], max_bond=16, abs_cut=0, rel_cut=0)], [1]).sample(
        x, 400, [s])
    d = dense.Dense(n).run(tapev(ops, x))
    ref = d.sample_sites(400, s)
    sim = mps_ref.RefSimulator(n, 16, 0.0, 0.0).run(tapev(ops, x))
    _, margins = sim.mps.sample_bits(400, s, margins=True)
    assert assert_same_bits(bits.cpu().numpy()[0], ref, margins) == 0


def test_shot_energy_matches_oracle_estimator(cuda):
    w = workloads.config(0)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    seed = rng.shot_shard_seed(rng.INITIAL_SAMPLE_SEED, 0)
    print('case: 12q shot energy', flush=True)
    e, tv = Batch([p], [1]).shot_energies(w['x'], 1000, [seed], terms=True)
    sim = mps_ref.RefSimulator(w['n'], 16, 0.0, 0.0).run(tapev(w['ops'], w['x'][0]))
    eo, to = sim.observe(w['words'], w['coeffs'], shots=1000)   # sampleSeed 1234567 -> same seed
    np.testing.assert_allclose(tv.cpu().numpy().real, to, atol=0)
    assert abs(e.cpu().numpy()[0] - eo) <= 1e-12 * max(1.0, abs(eo))


def test_uneven_batch_sampling(cuda):
    w = workloads.config(3)
    parts = w['parts'][:5]
    probs = [Problem(q['n'], q['ops'], q['nparams'], q['words'], q['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
             for q in parts]
    seeds = [rng.shot_shard_seed(1000 + j, 0) for j in range(len(parts))]
    x = np.concatenate([q['x'][0] for q in parts])
    print('case: uneven sampling', flush=True)
    bits, _ = Batch(probs, [1] * len(parts)).sample(x, 300, seeds)
    bits = bits.cpu().numpy()
    for j, q in enumerate(parts):
        sim = mps_ref.RefSimulator(q['n'], 16, 0.0, 0.0).run(tapev(q['ops'], q['x'][0]))
        ref, margins = sim.mps.sample_bits(300, seeds[j], margins=True)
        assert assert_same_bits(bits[j, :, :q['n']], ref, margins) == 0


def test_six_patterns_shot_error_bound(cuda):
    # validate_mps.py:33-56: 8192-shot energies within 0.08 of the exact value (Z words only here)
    n = 6
    words = ['ZIIIII', 'ZZIIII', 'IIIZII', 'IIIIII']
    coeffs = [0.3, -0.7, 0.2, 0.4]
    r = np.random.default_rng(17)
    for pattern in ('linear', 'reverse_linear', 'circular', 'full', 'pairwise', 'sca'):
        ops, P = workloads.su2_ansatz(n, 2, ('ry', 'rz'), pattern)
        x = r.uniform(-1, 1, P)
        b = Batch([Problem(n, ops, P, words, coeffs, max_bond=8, abs_cut=0, rel_cut=0)], [1])
        exact = b.energies(x).cpu().numpy()[0]
        print(f'case: shots pattern {pattern}', flush=True)
        shot = b.shot_energies(x, 8192, [rng.shot_shard_seed(rng.INITIAL_SAMPLE_SEED, 0)]).cpu().numpy()[0]
        assert abs(shot - exact) < 0.08
        assert abs(exact - dense.Dense(n).run(tapev(ops, x)).energy(words, coeffs)) < 1e-10


def test_real_sampling_kernels_match_complex_and_oracle(cuda):
    """Real circuits sample on the real kernels (real right environments and
    shot vectors on the packed real arena).  Their bits equal the complex
    kernels' bit for bit (the complex kernels only add exact zeros to the
    same accumulations) and the oracle's draws, on the 12-qubit
    configuration and on the 16 largest uneven partitions (26-32 qubits),
    through sample, evaluate and the shot-energy path."""
    w = workloads.config(0)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    assert p.info()['real_arith'] == 1
    seed = rng.shot_shard_seed(rng.INITIAL_SAMPLE_SEED + 3, 0)
    print('case: 12q real vs complex sampling kernels', flush=True)
    b = Batch([p], [1])
    bits_r, packed_r = b.sample(w['x'], 1000, [seed])
    e_r = b.shot_energies(w['x'], 1000, [seed]).cpu().numpy()
    b.set_arith('complex')
    bits_c, packed_c = b.sample(w['x'], 1000, [seed])
    e_c = b.shot_energies(w['x'], 1000, [seed]).cpu().numpy()
    assert np.array_equal(bits_r.cpu().numpy(), bits_c.cpu().numpy())
    assert np.array_equal(packed_r.cpu().numpy(), packed_c.cpu().numpy()) and np.array_equal(e_r, e_c)
    sim = mps_ref.RefSimulator(w['n'], 16, 0.0, 0.0).run(tapev(w['ops'], w['x'][0]))
    ref, margins = sim.mps.sample_bits(1000, seed, margins=True)
    assert assert_same_bits(bits_r.cpu().numpy()[0], ref, margins) == 0
    parts = sorted(workloads.config(3)['parts'], key=lambda q: -q['n'])[:16]
    probs = [Problem(q['n'], q['ops'], q['nparams'], q['words'], q['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
             for q in parts]
    x = np.concatenate([q['x'][0] for q in parts])
    sd = [rng.shot_shard_seed(2000 + j, 0) for j in range(len(parts))]
    print('case: 16 largest partitions, evaluate on real vs complex kernels', flush=True)
    bb = Batch(probs, [1] * len(parts))
    e1, b1, _ = bb.evaluate(x, 1000, sd)
    bb.set_arith('complex')
    e2, b2, _ = bb.evaluate(x, 1000, sd)
    assert np.array_equal(b1.cpu().numpy(), b2.cpu().numpy())
    assert np.all(np.abs(e1.cpu().numpy() - e2.cpu().numpy()) <= 1e-12 * np.abs(e2.cpu().numpy()))
    b1 = b1.cpu().numpy()
    for j, q in enumerate(parts[:4]):
        sim = mps_ref.RefSimulator(q['n'], 16, 0.0, 0.0).run(tapev(q['ops'], q['x'][0]))
        ref, margins = sim.mps.sample_bits(1000, sd[j], margins=True)
        assert assert_same_bits(b1[j, :, :q['n']], ref, margins) == 0
