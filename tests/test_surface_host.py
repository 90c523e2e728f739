"""Host-side behaviour of the CUDA-Q-compatible surface (no GPU): kernel
tapes, spin-operator algebra, seed mixers, bond configuration."""
import json
import os
from pathlib import Path

import numpy as np
import pytest

from gvmps import cudaq, seeds, workloads
from oracle import rng

GOLD = json.loads((Path(__file__).parent / 'golden' / 'reference_cases.json').read_text())


def test_kernel_tape_and_affine_parameters():
    k, a, b = cudaq.make_kernel(float, float)
    q = k.qalloc(3)
    k.ry(a, q[0])
    k.rz(2 * b - 0.5, q[1])
    k.rx(0.25, q[2])
    k.cx(q[0], q[1])
    k.h(q[2])
    assert k.width == 3 and k.argument_count == 2 and k.n_params() == 2
    assert k.tape() == [('ry', (0,), (0, 1.0, 0.0)), ('rz', (1,), (1, 2.0, -0.5)), ('rx', (2,), 0.25),
                        ('cx', (0, 1), None), ('h', (2,), None)]
    with pytest.raises(TypeError):
        k.ry(a * b, q[0])
    with pytest.raises(TypeError):
        k.ry(a + b, q[0])
    with pytest.raises(ValueError):
        k.qalloc(2)
    with pytest.raises(RuntimeError):
        k.cx(q[1], q[1])


def test_list_kernel_and_argument_parsing():
    k, thetas = cudaq.make_kernel(list)
    q = k.qalloc(2)
    k.ry(thetas[0], q[0])
    k.ry(thetas[3], q[1])
    assert k.n_params() == 4
    assert cudaq._params(k, ([1, 2, 3, 4],)).tolist() == [1, 2, 3, 4]
    with pytest.raises(RuntimeError):
        cudaq._params(k, ([1, 2],))
    f, x, y = cudaq.make_kernel(float, float)
    assert cudaq._params(f, (0.1, 0.2)).tolist() == [0.1, 0.2]
    assert cudaq._params(f, (np.array([0.1, 0.2]),)).tolist() == [0.1, 0.2]
    with pytest.raises(RuntimeError):
        cudaq._params(f, (0.1,))


def test_spin_algebra_matches_pauli_matrices():
    s = cudaq.spin
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0 + 0j, -1.0])
    I2 = np.eye(2)
    mats = {'I': I2, 'X': X, 'Y': Y, 'Z': Z}

    def dense(op, n):
        words, coeffs = op.words(n)
        out = np.zeros((2 ** n, 2 ** n), dtype=complex)
        for w, c in zip(words, coeffs):
            m = np.eye(1)
            for ch in reversed(w):  # qubit 0 is the least significant factor
                m = np.kron(m, mats[ch])
            out += c * m
        return out

    def kron_q(ops):  # ops[q] acting on qubit q
        m = np.eye(1)
        for o in reversed(ops):
            m = np.kron(m, o)
        return m

    H = 0.5 * s.x(0) * s.y(1) + s.z(0) * s.z(0) - 2.0 * s.y(0) * s.x(0) + 3 + s.z(1) * 0.25
    ref = 0.5 * kron_q([X, Y]) + np.eye(4) - 2.0 * kron_q([Y @ X, I2]) + 3 * np.eye(4) + 0.25 * kron_q([I2, Z])
    np.testing.assert_allclose(dense(H, 2), ref, atol=1e-14)
    words, coeffs = (s.x(0) * s.z(2)).words(3)
    assert words == ['XIZ'] and coeffs.tolist() == [1.0]
    with pytest.raises(RuntimeError):
        s.z(3).words(3)
    terms = []
    H.for_each_term(lambda t: terms.append((t.to_string(False), t.get_coefficient())))
    assert len(terms) == H.get_term_count()


def test_seed_mixers_match_golden_and_oracle():
    for text, h in GOLD['fnv1a32_kat']['cases']:
        assert seeds.word_seed(0, text) == h
    r = np.random.default_rng(3)
    for v in r.integers(0, 2 ** 32, 100).tolist() + [seeds.INITIAL_SAMPLE_SEED]:
        assert seeds.shot_shard_seed(v, 0) == rng.shot_shard_seed(v, 0)
        assert seeds.word_seed(v, 'IZZI') == rng.word_seed(v, 'IZZI')
    c = seeds.SeedCounter()
    assert [c.next() for _ in range(3)] == [1234567, 1234568, 1234569]


def test_configure_full_bond_cases(monkeypatch):
    for n, pattern, reps, expected in GOLD['full_bond_caps']['cases']:
        monkeypatch.setenv('CUDAQ_MPS_MAX_BOND', '1')
        stages = [dict(entanglement=pattern, reps=reps)] if pattern else None
        assert cudaq.configure_full_bond(n, stages) == expected
        assert os.environ['CUDAQ_MPS_MAX_BOND'] == str(expected)
        assert cudaq.engine_config() == (expected, 0.0, 0.0)


def test_engine_config_defaults(monkeypatch):
    for k in ('CUDAQ_MPS_MAX_BOND', 'CUDAQ_MPS_ABS_CUTOFF', 'CUDAQ_MPS_RELATIVE_CUTOFF'):
        monkeypatch.delenv(k, raising=False)
    assert cudaq.engine_config() == (64, 1e-5, 1e-5)
    monkeypatch.setenv('CUDAQ_MPS_MAX_BOND', '0')
    assert cudaq.engine_config()[0] == 1


def test_reverse_factory_lookup():
    k, a = cudaq.make_kernel(float)
    q = k.qalloc(2)
    k.ry(a, q[0])
    assert cudaq.for_objective(lambda x: 0.0, 'off') is None
    with pytest.raises(ValueError):
        cudaq.for_objective(lambda x: 0.0, 'bogus')

    def fun(x):
        return 0.0
    fun.reverse_factory = lambda: 'made'
    assert cudaq.for_objective(fun, 'auto') == 'made'
    assert cudaq.make_problem(object(), cudaq.spin.z(0)) is None


def test_entangling_patterns_match_reference_definitions():
    n = 5
    assert workloads.entangling_pairs(n, 0, 'linear') == [(i, i + 1) for i in range(n - 1)]
    assert workloads.entangling_pairs(n, 0, 'reverse_linear') == [(i, i + 1) for i in range(n - 2, -1, -1)]
    assert workloads.entangling_pairs(n, 1, 'pairwise') == [(1, 2), (3, 4)]
    assert workloads.entangling_pairs(n, 0, 'circular') == [(i, i + 1) for i in range(n - 1)] + [(n - 1, 0)]


class _FakeSparsePauliOp:
    """The two members of qiskit's SparsePauliOp the converter reads."""

    def __init__(self, pairs, n):
        self._pairs, self.num_qubits = pairs, n

    def to_list(self):
        return list(self._pairs)


def test_sparsepauliop_conversion_order_and_coefficients():
    """Qiskit_VQE/cudaq.py:25-47: Qiskit labels are big-endian (character idx
    acts on qubit n-1-idx), zero coefficients are dropped, only real parts
    are kept, repeated labels add up and an all-I label is a constant."""
    pairs = [('ZIXY', 0.5 + 0.2j), ('IIII', 1.25), ('IZII', 0.0), ('XIIZ', -0.75), ('XIIZ', 0.25)]
    H = cudaq.qiskit_sparsepauliop_to_cudaq_spin(_FakeSparsePauliOp(pairs, 4))
    words, coeffs = H.words(4)
    got = {w: c for w, c in zip(words, coeffs)}
    assert got == {'YXIZ': 0.5, 'IIII': 1.25, 'ZIIX': -0.5}
    # plain (label, coeff) lists, and the unreversed order
    H2 = cudaq.qiskit_sparsepauliop_to_cudaq_spin(pairs, reverse_qubit_order=False)
    words2, coeffs2 = H2.words(4)
    assert {w: c for w, c in zip(words2, coeffs2)} == {'ZIXY': 0.5, 'IIII': 1.25, 'XIIZ': -0.5}
    with pytest.raises(ValueError):
        cudaq.qiskit_sparsepauliop_to_cudaq_spin([('ZZ', 1.0)], num_qubits=3)
    with pytest.raises(ValueError):
        cudaq.qiskit_sparsepauliop_to_cudaq_spin([('ZA', 1.0)])


def test_sparsepauliop_conversion_energy_closed_form():
    """After RY(t_q) on every qubit of |0...0>, <Z_q> = cos t_q, <X_q> = sin t_q
    and <Y_q> = 0, so a converted operator's energy on the dense state must
    equal the product formula read off the Qiskit labels (qubit n-1-idx for
    character idx).  A reversed qubit order fails this."""
    from oracle import dense
    n = 5
    t = np.random.default_rng(3).uniform(-np.pi, np.pi, n)
    labels = ['ZIIIZ', 'IXIZI', 'YIIII', 'IIXXI', 'ZZZZZ', 'IIIII']
    coeffs = [0.7, -1.3, 2.0, 0.4, -0.9, 0.35]
    H = cudaq.qiskit_sparsepauliop_to_cudaq_spin(_FakeSparsePauliOp(list(zip(labels, coeffs)), n))
    words, cw = H.words(n)
    state = dense.Dense(n).run([('ry', (q,), t[q]) for q in range(n)])
    single = {'I': lambda a: 1.0, 'Z': np.cos, 'X': np.sin, 'Y': lambda a: 0.0}
    want = sum(c * np.prod([single[ch](t[n - 1 - idx]) for idx, ch in enumerate(lab)])
               for lab, c in zip(labels, coeffs))
    assert abs(state.energy(words, cw) - want) < 1e-12


def test_protein_probabilities_reverses_and_normalises():
    """test/tools/readout.py:4-9: q0-left counts to q0-right probabilities."""
    from gvmps import readout
    assert readout.protein_probabilities({'0011': 600, '1000': 400}) == {'1100': 0.6, '0001': 0.4}
    with pytest.raises(ValueError):
        readout.protein_probabilities({'01': 0})


def test_backend_metrics_rows(tmp_path, monkeypatch):
    """Rows of the campaign's analytic and sampled recorders: one per call,
    async calls logged once at their first get(), shots forced, parameters
    after the kernel (sample) or the kernel and Hamiltonian (observe)."""
    import types
    from gvmps import readout
    calls = []

    class Fut:
        def get(self):
            return 'async'

    def fake(name):
        def f(*args, **kw):
            calls.append((name, kw.get('shots_count')))
            return Fut() if name.endswith('_async') else name
        return f
    mod = types.SimpleNamespace(**{n: fake(n) for n in ('observe', 'observe_async', 'sample', 'sample_async')})
    monkeypatch.setenv('CUDAQ_MPS_MAX_BOND', '16')
    log = tmp_path / 'an.jsonl'
    assert readout.install_backend_metrics(mod, path=str(log)) == ['observe', 'observe_async']
    assert mod.observe('k', 'H', 0.5, 1.5) == 'observe'
    fut = mod.observe_async('k', 'H', [0.25, 0.75], qpu_id=0)
    assert fut.get() == 'async' and fut.get() == 'async'
    rows = [json.loads(l) for l in log.read_text().splitlines()]
    assert [r['operation'] for r in rows] == ['observe', 'observe_async_until_get']
    assert calls == [('observe', -1), ('observe_async', -1)]
    assert rows[0]['parameters'] == [0.5, 1.5] and rows[1]['parameters'] == [0.25, 0.75]
    assert all(r['shots'] == -1 and r['physical_shots'] == 0 and r['mps_max_bond'] == '16'
               and r['target'] == 'tensornet-mps-mqpu-shots' for r in rows)
    mod2 = types.SimpleNamespace(**{n: fake(n) for n in ('observe', 'observe_async', 'sample', 'sample_async')})
    log2 = tmp_path / 'sh.jsonl'
    readout.install_backend_metrics(mod2, path=str(log2), shots=1000)
    mod2.sample('k', 0.1, 0.2)
    mod2.sample_async('k', 0.3).get()
    mod2.observe('k', 'H', 0.4)
    rows = [json.loads(l) for l in log2.read_text().splitlines()]
    assert [r['operation'] for r in rows] == ['sample', 'sample_async', 'observe']
    assert [r['parameters'] for r in rows] == [[0.1, 0.2], [0.3], [0.4]]
    assert all(r['shots'] == 1000 and r['physical_shots_per_circuit_or_term'] == 1000
               and r['added_noise_model'] is False for r in rows)


def test_backend_metrics_guards(tmp_path):
    """The recorders' guards (analytic_execution.py:63-66, 77-79;
    sampled_execution.py:13-14, 40-41): an analytic run rejects finite shots
    and noise models and forbids sampling and state extraction; a sampled run
    needs exactly 1000 shots and forbids state extraction."""
    import types
    from gvmps import readout

    def fake(name):
        return lambda *a, **kw: name
    names = ('observe', 'observe_async', 'sample', 'sample_async', 'get_state')
    mod = types.SimpleNamespace(**{n: fake(n) for n in names})
    readout.install_backend_metrics(mod, path=str(tmp_path / 'a.jsonl'))
    with pytest.raises(RuntimeError, match='Finite-shot observe forbidden in analytic campaign'):
        mod.observe('k', 'H', 0.5, shots_count=1000)
    with pytest.raises(RuntimeError, match='Noise model forbidden in analytic campaign'):
        mod.observe_async('k', 'H', 0.5, noise_model=object())
    for name in ('sample', 'sample_async', 'get_state'):
        with pytest.raises(RuntimeError, match='Sampling/statevector readout forbidden'):
            getattr(mod, name)('k', 0.5)
    assert mod.observe('k', 'H', 0.5, shots_count=-1) == 'observe'
    assert not (tmp_path / 'b.jsonl').exists()
    for bad in (999, 1, 4096):
        with pytest.raises(ValueError, match='Matched sampled suite requires 1000 shots'):
            readout.install_backend_metrics(types.SimpleNamespace(**{n: fake(n) for n in names}),
                                            path=str(tmp_path / 'b.jsonl'), shots=bad)
    mod2 = types.SimpleNamespace(**{n: fake(n) for n in names})
    readout.install_backend_metrics(mod2, path=str(tmp_path / 's.jsonl'), shots=1000)
    with pytest.raises(RuntimeError, match='Dense state extraction is disabled; use sampled readout'):
        mod2.get_state('k', 0.5)
    assert mod2.sample('k', 0.5) == 'sample'
