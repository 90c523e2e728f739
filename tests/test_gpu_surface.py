"""GPU checks of the CUDA-Q-compatible surface (gvmps.cudaq / gvmps.optimize)
against the oracle's simulator semantics (oracle/mps_ref.RefSimulator runs the
same call sequence with its own sampleSeed counter) and the reverse oracle.

Bars: energies 1e-10 relative, gradients 1e-8 absolute, shot results and
counts identical (same uniforms; a draw within 1e-12 of a CDF boundary would
be the only allowed difference and does not occur for these seeds)."""
import numpy as np
import pytest

from gvmps import cudaq, optimize, workloads
from oracle import mps_ref, reverse_ref, rng

pytestmark = pytest.mark.gpu


def build(n, reps, gates, pattern, spread=False):
    count = n * (reps + 1) * len(gates)
    if spread:
        kernel, *params = cudaq.make_kernel(*([float] * count))
        par = lambda i: params[i]
    else:
        kernel, thetas = cudaq.make_kernel(list)
        par = lambda i: thetas[i]
    q = kernel.qalloc(n)
    ops, idx = [], 0
    for layer in range(reps + 1):
        for j in range(n):
            for g in gates:
                getattr(kernel, g)(par(idx), q[j])
                ops.append((g, (j,), idx))
                idx += 1
        if layer < reps:
            for a, b in workloads.entangling_pairs(n, layer, pattern):
                kernel.cx(q[a], q[b])
                ops.append(('cx', (a, b), None))
    return kernel, ops, count


def hamiltonian(n, rng_, T, xy=True):
    s = cudaq.spin
    letters = 'IXYZ' if xy else 'IZ'
    H = 0.7 * s.i(0)
    words, coeffs = [], []
    for _ in range(T):
        w = ''.join(rng_.choice(list(letters), n))
        c = float(rng_.normal())
        term = c * s.i(0)
        for qq, ch in enumerate(w):
            if ch != 'I':
                term = term * getattr(s, ch.lower())(qq)
        H = H + term
    words, coeffs = H.words(n)
    return H, words, coeffs


def tape_values(ops, x):
    return [(g, q, (x[p] if p is not None else None)) for g, q, p in ops]


@pytest.fixture(autouse=True)
def fresh_engine(monkeypatch):
    for k in ('CUDAQ_MPS_MAX_BOND', 'CUDAQ_MPS_ABS_CUTOFF', 'CUDAQ_MPS_RELATIVE_CUTOFF'):
        monkeypatch.setenv(k, '1')  # restored after the test
    cudaq.configure_full_bond(8, [dict(entanglement='linear', reps=3)])
    cudaq._engine.seeds.reset()
    yield


def test_observe_analytic_and_terms(cuda):
    n = 7
    r = np.random.default_rng(11)
    kernel, ops, P = build(n, 2, ('ry', 'rz'), 'linear', spread=True)
    H, words, coeffs = hamiltonian(n, r, 40)
    x = r.uniform(-np.pi, np.pi, P)
    print('case: observe analytic 7q', flush=True)
    res = cudaq.observe(kernel, H, *x)
    res_async = cudaq.observe_async(kernel, H, *x, qpu_id=0).get()
    sim = mps_ref.RefSimulator(n, max_bond=16, abs_cut=0.0, rel_cut=0.0).run(tape_values(ops, x))
    e_ref, tv_ref = sim.observe(words, coeffs)
    assert abs(res.expectation() - e_ref) <= 1e-10 * max(1.0, abs(e_ref))
    assert res_async.expectation() == res.expectation()
    np.testing.assert_allclose(res.term_values, tv_ref, atol=1e-11)
    w0 = next(w for w in words if w != 'I' * n)
    assert res.expectation(w0) == res.get_term_values()[w0]


def test_shot_observe_and_sample_follow_seed_counter(cuda):
    n = 8
    r = np.random.default_rng(12)
    kernel, ops, P = build(n, 3, ('ry',), 'reverse_linear')
    H, words, coeffs = hamiltonian(n, r, 60, xy=False)
    x = r.uniform(-np.pi, np.pi, P)
    sim = mps_ref.RefSimulator(n, max_bond=8, abs_cut=0.0, rel_cut=0.0).run(tape_values(ops, x))
    print('case: shot observe x2 + sample x2, 8q', flush=True)
    for _ in range(2):
        res = cudaq.observe(kernel, H, x, shots_count=500)
        e_ref, tv_ref = sim.observe(words, coeffs, shots=500)
        np.testing.assert_array_equal(res.term_values, np.asarray(tv_ref))
        assert abs(res.expectation() - e_ref) <= 1e-12 * max(1.0, abs(e_ref))
    for shots in (300, 0):
        counts = cudaq.sample(kernel, x, shots_count=shots)
        assert dict(counts) == sim.sample(shots)
        assert sum(counts.values()) == (shots if shots > 0 else 1000)


def test_mixed_basis_shot_observe(cuda):
    n = 6
    r = np.random.default_rng(13)
    kernel, ops, P = build(n, 2, ('ry', 'rz'), 'linear')
    H, words, coeffs = hamiltonian(n, r, 25)
    x = r.uniform(-np.pi, np.pi, P)
    print('case: mixed-basis shot observe 6q', flush=True)
    res = cudaq.observe(kernel, H, x, shots_count=400)
    sim = mps_ref.RefSimulator(n, max_bond=8, abs_cut=0.0, rel_cut=0.0).run(tape_values(ops, x))
    exact = sim.observe_terms(words)
    for k, w in enumerate(words):
        sw = w[::-1]
        if set(w) == {'I'}:
            assert res.term_values[k] == 1.0
        elif set(w) <= {'I', 'Z'}:
            want = sim.mps.sample_z_expectations([sw], 400, rng.word_seed(rng.INITIAL_SAMPLE_SEED, sw))[0]
            assert res.term_values[k] == want
        else:
            assert abs(res.term_values[k] - exact[k].real) < 1e-11


def test_reverse_problem_and_autograd(cuda):
    import torch
    n = 6
    r = np.random.default_rng(14)
    kernel, ops, P = build(n, 2, ('ry', 'rz'), 'reverse_linear')
    H, words, coeffs = hamiltonian(n, r, 30)
    x = r.uniform(-np.pi, np.pi, P)
    print('case: reverse problem 6q', flush=True)
    prob = cudaq.make_problem(kernel, H)
    assert prob.engine == 'gvmps_reverse'
    e, g = prob.value_gradient(x)
    eo, go = reverse_ref.value_gradient(n, ops, words, np.real(coeffs), x)
    assert abs(e - eo) < 1e-11
    np.testing.assert_allclose(g, go, atol=1e-8, rtol=0)
    np.testing.assert_allclose(prob.values(x), reverse_ref.values(n, ops, words, x), atol=1e-11)
    X = np.stack([x, x + 0.1, x - 0.2])
    xt = torch.tensor(X, dtype=torch.float64, requires_grad=True)
    E = prob.energy(xt)
    (grad,) = torch.autograd.grad((E * torch.tensor([1.0, 2.0, -1.0], dtype=torch.float64)).sum(), xt)
    eb, gb = prob.value_gradient_batch(X)
    np.testing.assert_array_equal(E.detach().numpy(), eb)
    np.testing.assert_allclose(grad.numpy(), gb * np.array([1.0, 2.0, -1.0])[:, None], rtol=0, atol=0)
    fun = lambda v: 0.0
    fun.reverse_problem = (kernel, H)
    assert isinstance(cudaq.for_objective(fun, 'auto'), cudaq.ReverseProblem)


def test_batched_adapters_equal_sequential_calls(cuda):
    n = 6
    r = np.random.default_rng(15)
    kernel, ops, P = build(n, 2, ('ry',), 'linear')
    H, words, coeffs = hamiltonian(n, r, 20, xy=False)
    x = r.uniform(-np.pi, np.pi, P)
    obj = optimize.BatchedObjective(kernel, H)
    print('case: batched adapters 6q', flush=True)
    seq = lambda v: cudaq.observe(kernel, H, v).expectation()
    g_ps = optimize.parameter_shift_gradient(obj, x)
    want = np.array([(seq(x + np.pi / 2 * e) - seq(x - np.pi / 2 * e)) / 2 for e in np.eye(P)])
    np.testing.assert_array_equal(g_ps, want)
    _, go = reverse_ref.value_gradient(n, ops, words, np.real(coeffs), x)
    np.testing.assert_allclose(g_ps, go, atol=1e-9)
    g_fd = optimize.central_gradient(obj, x, 1e-6)
    np.testing.assert_allclose(g_fd, go, atol=1e-6)
    D = r.choice([-1.0, 1.0], size=(4, P))
    est = optimize.perturbation_gradients(obj, x, 0.1, D)
    for d, row in zip(D, est):
        np.testing.assert_array_equal(row, (seq(x + 0.1 * d) - seq(x - 0.1 * d)) / 0.2 * d)


def test_lbfgs_with_adjoint_descends(cuda):
    n = 6
    r = np.random.default_rng(16)
    kernel, ops, P = build(n, 2, ('ry',), 'linear')
    H, words, coeffs = hamiltonian(n, r, 20, xy=False)
    x0 = r.uniform(-np.pi, np.pi, P)
    print('case: L-BFGS-B with adjoint 6q', flush=True)
    res = optimize.lbfgs_minimize(kernel, H, x0, maxiter=30, reverse_mode='required')
    e0 = cudaq.observe(kernel, H, x0).expectation()
    assert res.fun < e0 - 1e-6
    assert res.njev_reverse > 0
    res_c = optimize.cobyla_minimize(optimize.BatchedObjective(kernel, H), x0, rhobeg=0.5, maxiter=40)
    assert res_c.fun < e0


def test_backend_metrics_wrapper_composes_with_observe(tmp_path):
    """The evaluation recorder wraps this module's observe/observe_async the
    way the campaign wraps CUDA-Q's (kernel, Hamiltonian, then parameters):
    same energies as the unwrapped calls, one row per evaluation."""
    import json
    import types
    from gvmps import readout
    n = 6
    kernel, ops, count = build(n, 2, ('ry',), 'linear', spread=True)
    H, words, coeffs = hamiltonian(n, np.random.default_rng(5), 20)
    x = np.random.default_rng(6).uniform(-1, 1, count).tolist()
    ns = types.SimpleNamespace(observe=cudaq.observe, observe_async=cudaq.observe_async,
                               sample=cudaq.sample, sample_async=cudaq.sample_async)
    log = tmp_path / 'rows.jsonl'
    readout.install_backend_metrics(ns, path=str(log))
    e0 = cudaq.observe(kernel, H, *x).expectation()
    e1 = ns.observe(kernel, H, *x).expectation()
    e2 = ns.observe_async(kernel, H, *x, qpu_id=0).get().expectation()
    assert e1 == e0 and e2 == e0
    rows = [json.loads(l) for l in log.read_text().splitlines()]
    assert [r['operation'] for r in rows] == ['observe', 'observe_async_until_get']
    assert all(r['parameters'] == x and r['parameter_count'] == count for r in rows)


def test_sampled_readout_and_rows_against_oracle(tmp_path):
    """A sampled campaign run on this backend (sampled_execution.py:10-41 over
    this module): sample and sample_async through the recorder give the
    oracle's counts for the same sampleSeed sequence, the protein decoder's
    probabilities are the reversed keys over 1000 shots (readout.py:4-9),
    every call leaves one row with its parameters, and state extraction is
    refused."""
    import json
    import types
    from gvmps import readout
    n = 8
    r = np.random.default_rng(21)
    kernel, ops, count = build(n, 3, ('ry',), 'reverse_linear', spread=True)
    xs = [r.uniform(-np.pi, np.pi, count).tolist() for _ in range(3)]
    ns = types.SimpleNamespace(observe=cudaq.observe, observe_async=cudaq.observe_async,
                               sample=cudaq.sample, sample_async=cudaq.sample_async, get_state=cudaq.get_state)
    log = tmp_path / 'sampled.jsonl'
    readout.install_backend_metrics(ns, path=str(log), shots=1000)
    print('case: sampled readout 8q, 3 sample calls', flush=True)
    got = [ns.sample(kernel, *xs[0]), ns.sample_async(kernel, *xs[1]).get(), ns.sample(kernel, *xs[2])]
    with pytest.raises(RuntimeError, match='Dense state extraction is disabled'):
        ns.get_state(kernel, *xs[0])
    sim_seed = rng.INITIAL_SAMPLE_SEED
    for x, counts in zip(xs, got):
        sim = mps_ref.RefSimulator(n, max_bond=8, abs_cut=0.0, rel_cut=0.0).run(tape_values(ops, x))
        sim.sample_seed = sim_seed
        want = sim.sample(1000)
        sim_seed += 1
        assert dict(counts) == want
        probs = readout.protein_probabilities(counts)
        assert probs == {k[::-1]: v / 1000 for k, v in want.items()}
        assert abs(sum(probs.values()) - 1.0) < 1e-12
    rows = [json.loads(l) for l in log.read_text().splitlines()]
    assert [r_['operation'] for r_ in rows] == ['sample', 'sample_async', 'sample']
    assert [r_['parameters'] for r_ in rows] == xs
    assert all(r_['shots'] == 1000 and r_['parameter_count'] == count for r_ in rows)


def test_spsa_batched_equals_sequential_observe_loop(cuda):
    """SPSA (spsa_optimizer.py:66-184): the batched loop, one launch per
    update, against the sequential restatement making one scalar observe call
    per probe: identical iterates and result; the best energy equals the
    oracle's energy of the best point (1e-10 relative)."""
    from oracle import spsa_ref
    n = 8
    r = np.random.default_rng(31)
    kernel, ops, P = build(n, 3, ('ry',), 'linear')
    H, words, coeffs = hamiltonian(n, r, 40, xy=False)
    x0 = r.uniform(-np.pi, np.pi, P)
    c = optimize.spsa_settings({})
    c.update(energy_tol=1e-6)
    print('case: SPSA 8q batched vs sequential observe', flush=True)
    seq_it, bat_it = [], []
    want = spsa_ref.minimize(lambda v: cudaq.observe(kernel, H, v).expectation(), x0, c, maxiter=25,
                             callback=lambda v: seq_it.append(v.copy()))
    obj = optimize.BatchedObjective(kernel, H)
    got = optimize.spsa_minimize(obj, x0, maxiter=25, settings=c, callback=lambda v: bat_it.append(v.copy()))
    assert want['nit'] == 25 and len(seq_it) == len(bat_it) == want['nit']
    assert all(np.array_equal(a, b) for a, b in zip(seq_it, bat_it))
    assert np.array_equal(got.x, want['x']) and got.fun == want['fun']
    assert (got.nit, got.nfev, got.njev, got.message) == (want['nit'], want['nfev'], want['njev'], want['message'])
    assert obj.launches == got.nit + 1
    sim = mps_ref.RefSimulator(n, max_bond=8, abs_cut=0.0, rel_cut=0.0).run(tape_values(ops, got.x))
    e_ref, _ = sim.observe(words, coeffs)
    assert abs(got.fun - e_ref) <= 1e-10 * max(1.0, abs(e_ref))


def test_spsa_many_runs_in_one_launch_per_round(cuda):
    """Several independent SPSA runs (partitions of different sizes) share
    each round's launch; every run equals its own batched loop."""
    from gvmps import Batch, Problem

    class ProblemObjective:
        def __init__(self, p):
            self.p = p

        def prefetch(self, pts):
            pts = np.atleast_2d(pts)
            return Batch([self.p], [len(pts)]).energies(pts).cpu().numpy()

    r = np.random.default_rng(32)
    probs, x0s = [], []
    for n in (6, 9, 12):
        ops, P = workloads.layered_ansatz(n, 3, ('ry',), 'linear')
        words, coeffs = workloads.z_hamiltonian(n, 60, 3, r)
        probs.append(Problem(n, ops, P, words, coeffs, max_bond=16, abs_cut=0, rel_cut=0))
        x0s.append(r.uniform(-np.pi, np.pi, P))
    c = optimize.spsa_settings({})
    c.update(energy_tol=1e-3)
    print('case: spsa_many over 6/9/12-qubit problems', flush=True)
    res, rounds = optimize.spsa_many(probs, x0s, maxiter=40, settings=c)
    assert rounds == max(rr.nit for rr in res) + 1
    for p, x0, rr in zip(probs, x0s, res):
        one = optimize.spsa_minimize(ProblemObjective(p), x0, maxiter=40, settings=c)
        assert np.array_equal(one.x, rr.x) and one.fun == rr.fun and one.nit == rr.nit and one.nfev == rr.nfev
