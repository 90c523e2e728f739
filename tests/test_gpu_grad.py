"""GPU parity of the reverse (adjoint) gradient, gvmps_batch_value_grad.

Checks, each against something the kernels were not derived from:
  * the reference's own reverse-engine cases (test/tests/test_reverse_mps.py:
    10-78, golden/reference_cases.json) against oracle/reverse_ref and the
    dense statevector: energy within 1e-12, gradient within 1e-8 absolute;
  * random tapes with X/Y/Z words against reverse_ref (1e-8 absolute);
  * the 60-qubit adjoint workload (8000 Z-strings, batch 32, checkpoint
    stride 2) against the exact parameter-shift rule evaluated with the
    independent SVD energy path (E(x + pi/2 e_p) - E(x - pi/2 e_p)) / 2 on
    32 random parameters, against central differences (h = 1e-6) with a
    tolerance set by the difference's rounding error, and against the
    reverse oracle;
  * the value equals the SVD path's energy (1e-10 relative);
  * batched and one-at-a-time gradients are bitwise equal.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from gvmps import Batch, Problem, workloads
from oracle import dense, reverse_ref

pytestmark = pytest.mark.gpu

SVD = dict(max_bond=16, abs_cut=0.0, rel_cut=0.0)
GOLD = json.loads((Path(__file__).parent / 'golden' / 'reference_cases.json').read_text())


def _case_ops(case, pattern):
    edges = [tuple(e) for e in case['edges_linear']]
    if pattern == 'reverse_linear':
        edges = edges[::-1]
    ops, p = [], 0
    for layer in range(case['reps'] + 1):
        for q in range(case['n']):
            for g in case['gates']:
                ops.append((g, (q,), p))
                p += 1
        if layer < case['reps']:
            ops += [('cx', e, None) for e in edges]
    return ops, p


def _dense_energy(n, ops, x, words, coeffs):
    d = dense.Dense(n)
    d.run([(g, q, (x[p] if isinstance(p, (int, np.integer)) else None)) for g, q, p in ops])
    return d.energy(words, coeffs)


@pytest.mark.parametrize('pattern', ['linear', 'reverse_linear'])
def test_reference_reverse_case(cuda, pattern):
    case = GOLD['reverse_mps_dense']
    ops, P = _case_ops(case, pattern)
    x = np.random.default_rng(52).normal(size=P)
    print(f'case: reverse_mps_dense {pattern}', flush=True)
    e, g = Batch([Problem(case['n'], ops, P, case['words'], case['coeffs'], **SVD)], [1]).value_grad(x)
    e, g = float(e.cpu()[0]), g.cpu().numpy()
    eo, go = reverse_ref.value_gradient(case['n'], ops, case['words'], case['coeffs'], x)
    assert abs(e - _dense_energy(case['n'], ops, x, case['words'], case['coeffs'])) < 0.5e-12
    assert abs(e - eo) < 1e-12
    np.testing.assert_allclose(g, go, atol=1e-8, rtol=0)


def test_reference_complex_case(cuda):
    """Initial |+++> (reverse_mps.py initial='+') is an H layer; Y/X words."""
    case = GOLD['reverse_mps_complex']
    n = case['n']
    ops = [(g, (q, t) if g == 'cx' else (q,), None if g == 'cx' else t) for g, q, t in case['ops']]
    x = np.array(case['x'])
    print('case: reverse_mps_complex', flush=True)
    tape = [('h', (q,), None) for q in range(n)] + ops
    e, g = Batch([Problem(n, tape, len(x), case['words'], case['coeffs'], **SVD)], [1]).value_grad(x)
    eo, go = reverse_ref.value_gradient(n, ops, case['words'], case['coeffs'], x, initial=case['initial'])
    assert abs(float(e.cpu()[0]) - eo) < 0.5 * 10 ** -case['energy_places']
    np.testing.assert_allclose(g.cpu().numpy(), go, atol=0.5 * 10 ** -case['gradient_places'], rtol=0)


@pytest.mark.parametrize('seed', range(4))
def test_random_tapes_xyz_words(cuda, seed):
    r = np.random.default_rng(300 + seed)
    n = int(r.integers(3, 9))
    ops, P, bond = [], 0, [1] * (n + 1)
    for _ in range(50):
        kind = r.integers(0, 4)
        q = int(r.integers(0, n))
        if kind == 0:
            ops.append((str(r.choice(['rx', 'ry', 'rz'])), (q,), P))
            P += 1
        elif kind == 1:
            ops.append((str(r.choice(['h', 's', 'sdg', 'x', 'y', 'z'])), (q,), None))
        elif q + 1 < n and bond[q + 1] < 16:
            bond[q + 1] *= 2
            ops.append(('cx', (q, q + 1) if r.integers(0, 2) else (q + 1, q), None))
    words = [''.join(r.choice(list('IXYZ'), n)) for _ in range(40)] + ['I' * n]
    coeffs = r.normal(size=len(words))
    X = r.uniform(-np.pi, np.pi, (3, P))
    print(f'case: random tape seed={seed} n={n} P={P}', flush=True)
    e, g = Batch([Problem(n, ops, P, words, coeffs, **SVD)], [3]).value_grad(X)
    e, g = e.cpu().numpy(), g.cpu().numpy().reshape(3, P)
    for i in range(3):
        eo, go = reverse_ref.value_gradient(n, ops, words, coeffs, X[i])
        assert abs(e[i] - eo) < 1e-11 * max(1.0, abs(eo))
        np.testing.assert_allclose(g[i], go, atol=1e-8, rtol=0)


@pytest.mark.parametrize('seed', range(3))
def test_real_tapes_real_path(cuda, seed):
    """Real tapes (RY, H, X, Z, nearest-neighbour CX) with I/X/Z words take the
    real-arithmetic forward/backward kernels: against the reverse oracle and
    against the complex kernels of the same batch."""
    r = np.random.default_rng(700 + seed)
    n = int(r.integers(3, 10))
    ops, P, bond = [], 0, [1] * (n + 1)
    for _ in range(60):
        kind = r.integers(0, 4)
        q = int(r.integers(0, n))
        if kind == 0:
            ops.append(('ry', (q,), P))
            P += 1
        elif kind == 1:
            ops.append((str(r.choice(['h', 'x', 'z'])), (q,), None))
        elif q + 1 < n and bond[q + 1] < 16:
            bond[q + 1] *= 2
            ops.append(('cx', (q, q + 1) if r.integers(0, 2) else (q + 1, q), None))
    words = [''.join(r.choice(list('IXZ'), n)) for _ in range(40)] + ['I' * n]
    coeffs = r.normal(size=len(words))
    X = r.uniform(-np.pi, np.pi, (3, P))
    print(f'case: real tape seed={seed} n={n} P={P}', flush=True)
    p = Problem(n, ops, P, words, coeffs, **SVD)
    assert p.info()['real_arith'] == 1
    b = Batch([p], [3])
    e, g = b.value_grad(X)
    e, g = e.cpu().numpy(), g.cpu().numpy().reshape(3, P)
    b.set_arith('complex')
    ec, gc = b.value_grad(X)
    ec, gc = ec.cpu().numpy(), gc.cpu().numpy().reshape(3, P)
    np.testing.assert_allclose(e, ec, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g, gc, atol=1e-11, rtol=0)
    for i in range(3):
        eo, go = reverse_ref.value_gradient(n, ops, words, coeffs, X[i])
        assert abs(e[i] - eo) < 1e-11 * max(1.0, abs(eo))
        np.testing.assert_allclose(g[i], go, atol=1e-8, rtol=0)


def test_real_path_40q_matches_complex(cuda):
    """The 40-qubit workload's gradients, real kernels against complex kernels
    (checkpoint stride 1 and 2)."""
    w = workloads.config(1)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], **SVD)
    X = w['x'][:2]
    print('case: spsa-40q value_grad real vs complex kernels', flush=True)
    b = Batch([p], [2])
    e, g = b.value_grad(X)
    b.set_checkpoint(2)
    e2, g2 = b.value_grad(X)
    assert np.array_equal(e.cpu().numpy(), e2.cpu().numpy()) and np.array_equal(g.cpu().numpy(), g2.cpu().numpy())
    b.set_arith('complex')
    ec, gc = b.value_grad(X)
    np.testing.assert_allclose(e.cpu().numpy(), ec.cpu().numpy(), rtol=1e-12, atol=0)
    np.testing.assert_allclose(g.cpu().numpy(), gc.cpu().numpy(), atol=1e-9, rtol=0)


def test_adjoint_60q_config(cuda):
    """The 60-qubit adjoint configuration as stated: 4 layers of RY+RZ with
    linear CX chains (480 parameters), 8000 Z-strings, a batch of 32 vectors
    with step environments checkpointed every other step (recomputed in the
    backward pass), checked against
      * the batch at checkpoint stride 1, bitwise;
      * exact parameter shift on 32 random parameters of two vectors (1e-8);
      * central differences, h = 1e-6, on the same 32 parameters;
      * the reverse oracle (reverse_mps.py:84-99 restated) on one vector at
        1e-8 absolute, its energy at 1e-10 relative."""
    w = workloads.config(2)
    n, P = w['n'], w['nparams']
    p = Problem(n, w['ops'], P, w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    assert p.info()['grad_ok'] == 1 and P == 480 and len(w['x']) == 32
    X = w['x']
    print('case: adjoint-60q value_grad batch 32, checkpoint stride 2 and 1', flush=True)
    b = Batch([p], [32])
    b.set_checkpoint(2)
    e, g = b.value_grad(X)
    e, g = e.cpu().numpy(), g.cpu().numpy().reshape(32, P)
    b.set_checkpoint(1)
    e1, g1 = b.value_grad(X)
    assert np.array_equal(e1.cpu().numpy(), e) and np.array_equal(g1.cpu().numpy().reshape(32, P), g)
    del b
    idx = np.random.default_rng(7).choice(P, 32, replace=False)
    h = 1e-6
    pts = []
    for row in (0, 1):
        for sgn in (+1, -1):
            for k in idx:
                xs = X[row].copy()
                xs[k] += sgn * np.pi / 2
                pts.append(xs)
    for sgn in (+1, -1):
        for k in idx:
            xs = X[0].copy()
            xs[k] += sgn * h
            pts.append(xs)
    print(f'case: adjoint-60q SVD-path energies for parameter shift + FD ({len(pts) + 2} vectors)', flush=True)
    es = Batch([p], [len(pts) + 2]).energies(np.vstack(pts + [X[0], X[1]])).cpu().numpy()
    # value: the SVD state and the bond-doubling state are the same state
    assert np.all(np.abs(e[:2] - es[-2:]) <= 1e-10 * np.abs(es[-2:]))
    for row in (0, 1):
        base = 64 * row
        ps = (es[base:base + 32] - es[base + 32:base + 64]) / 2
        # parameter shift is exact for single-use rotations (rounding of E only)
        np.testing.assert_allclose(g[row, idx], ps, atol=1e-8, rtol=0)
    fd = (es[128:160] - es[160:192]) / (2 * h)
    # central difference: the truncation term h^2 E'''/6 is negligible; the
    # rounding of the two energies is bounded by ~10 eps sum|c_k|, over 2h
    tol = 10 * np.finfo(float).eps * np.abs(w['coeffs']).sum() / h
    np.testing.assert_allclose(g[0, idx], fd, atol=tol, rtol=0)
    print('case: adjoint-60q reverse oracle on vector 0 (CPU)', flush=True)
    eo, go = reverse_ref.value_gradient(n, w['ops'], w['words'], np.real(w['coeffs']), X[0])
    assert abs(e[0] - eo) <= 1e-10 * abs(eo)
    np.testing.assert_allclose(g[0], go, atol=1e-8, rtol=0)


def test_batched_equals_plain_bitwise(cuda):
    w = workloads.config(1)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    X = w['x'][:4]
    print('case: spsa-40q value_grad batch 4 vs plain', flush=True)
    eb, gb = Batch([p], [4]).value_grad(X)
    eb, gb = eb.cpu().numpy(), gb.cpu().numpy().reshape(4, -1)
    single = Batch([p], [1])
    for i in range(4):
        e1, g1 = single.value_grad(X[i])
        assert e1.cpu().numpy()[0] == eb[i]
        assert np.array_equal(g1.cpu().numpy(), gb[i])


def test_mixed_problem_batch(cuda):
    """Instances of different problems in one batch match their own batches.
    A batch runs real arithmetic only when all its problems qualify, so the
    real 12-qubit problem runs complex here: it matches its own batch run
    complex bitwise, and its real-arithmetic run to rounding."""
    a = workloads.config(0)
    r = np.random.default_rng(5)
    n = 9
    ops, P = workloads.su2_ansatz(n, 3, ('rx', 'rz'), 'reverse_linear')
    words = [''.join(r.choice(list('IXYZ'), n)) for _ in range(60)]
    coeffs = r.normal(size=60)
    pa = Problem(a['n'], a['ops'], a['nparams'], a['words'], a['coeffs'], **SVD)
    pb = Problem(n, ops, P, words, coeffs, **SVD)
    xa, xb = a['x'][0], r.uniform(-np.pi, np.pi, P)
    print('case: mixed 12q/9q value_grad batch', flush=True)
    e, g = Batch([pa, pb], [1, 1]).value_grad(np.concatenate([xa, xb]))
    e, g = e.cpu().numpy(), g.cpu().numpy()
    ba = Batch([pa], [1])
    ea_real, ga_real = ba.value_grad(xa)
    ba.set_arith('complex')
    ea, ga = ba.value_grad(xa)
    eb, gb = Batch([pb], [1]).value_grad(xb)
    assert e[0] == ea.cpu().numpy()[0] and e[1] == eb.cpu().numpy()[0]
    assert abs(e[0] - float(ea_real.cpu()[0])) < 1e-12 * abs(e[0])
    np.testing.assert_allclose(g[:a['nparams']], ga_real.cpu().numpy(), atol=1e-11, rtol=0)
    assert np.array_equal(g[:a['nparams']], ga.cpu().numpy())
    assert np.array_equal(g[a['nparams']:], gb.cpu().numpy())
    eo, go = reverse_ref.value_gradient(n, ops, words, coeffs, xb)
    np.testing.assert_allclose(g[a['nparams']:], go, atol=1e-8, rtol=0)


def test_unsupported_tapes_raise(cuda):
    n = 4
    ops = [('ry', (0,), 0), ('swap', (0, 1), None)]
    p = Problem(n, ops, 1, ['ZIII'], [1.0])
    assert p.info()['grad_ok'] == 0
    with pytest.raises(RuntimeError, match='SWAP'):
        Batch([p], [1]).value_grad([0.3])
    p2 = Problem(n, [('ry', (0,), 0), ('cx', (0, 2), None)], 1, ['ZIII'], [1.0])
    with pytest.raises(RuntimeError, match='neighbouring'):
        Batch([p2], [1]).value_grad([0.3])


def test_checkpoint_strides_are_bitwise_identical(cuda):
    """Keeping every 2nd / 3rd step environment and recomputing the others in
    the backward pass gives bitwise the same energies and gradients."""
    w = workloads.config(1)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], **SVD)
    X = w['x'][:3]
    for arith in ('auto', 'complex'):  # real-arithmetic and complex kernels
        b = Batch([p], [3])
        b.set_arith(arith)
        print(f'case: spsa-40q value_grad checkpoint strides 1/2/3 ({arith})', flush=True)
        e1, g1 = b.value_grad(X)
        e1, g1 = e1.cpu().numpy(), g1.cpu().numpy()
        for stride in (2, 3):
            b.set_checkpoint(stride)
            e, g = b.value_grad(X)
            assert np.array_equal(e.cpu().numpy(), e1), (arith, stride)
            assert np.array_equal(g.cpu().numpy(), g1), (arith, stride)


def test_grad_flops_count_the_forward_skips(cuda):
    """The executed-flop count of value_grad: with the zero structure the
    40-qubit forward issues half its products on the bulk steps (2 of the
    10 real products per step), so the count lies between 70% and 90% of
    the full-product count; with skipping off it is the full count, and
    the gradients agree with the skipping run to rounding."""
    w = workloads.config(1)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], **SVD)
    X = w['x'][:4]
    b = Batch([p], [4])
    print('case: spsa-40q x 4, gradient flop count with and without skips', flush=True)
    e, g = b.value_grad(X)
    full = p.info()['flops_grad_executed'] * 4 / 4.0  # one real product per complex one
    assert 0.7 * full < b.grad_flops() < 0.9 * full
    b.set_sparsity(False)
    e2, g2 = b.value_grad(X)
    assert b.grad_flops() == pytest.approx(full, rel=1e-12)
    np.testing.assert_allclose(e2.cpu().numpy(), e.cpu().numpy(), rtol=1e-12, atol=0)
    np.testing.assert_allclose(g2.cpu().numpy(), g.cpu().numpy(), rtol=0, atol=1e-9)


def test_gradient_autotuner_configs_are_bitwise_identical(cuda):
    """Every launch shape and checkpoint stride the gradient autotuner tries
    gives bitwise the same energies and gradients, and the tuner leaves the
    batch on its choice."""
    from gvmps import autotune
    w = workloads.config(1)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], **SVD)
    X = w['x'][:8]
    b = Batch([p], [8])
    print('case: spsa-40q x 8, gradient autotuner configurations', flush=True)
    ref = None
    for wpb, stride in autotune.GRAD_CONFIGS:
        b.set_launch(wpb, 1)
        b.set_checkpoint(stride)
        e, g = b.value_grad(X)
        e, g = e.cpu().numpy(), g.cpu().numpy()
        if ref is None:
            ref = (e, g)
        else:
            assert np.array_equal(e, ref[0]) and np.array_equal(g, ref[1]), (wpb, stride)
    best, times = autotune.tune_grad(b, X, reps=1, use_cache=False)
    assert best in autotune.GRAD_CONFIGS and len(times) == len(autotune.GRAD_CONFIGS)
    assert b.launch == (best[0], 1)
    e, g = b.value_grad(X)
    assert np.array_equal(g.cpu().numpy(), ref[1])
