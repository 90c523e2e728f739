"""Seeded synthetic workloads of the benchmark configurations.

Hamiltonians are Z-only Pauli strings (qubit-order words, word[q] acts on
qubit q): T distinct strings, weight w ~ U{1..wmax} on distinct qubits drawn
uniformly, 20% of the coefficients ~ U[50, 500] (penalties), 80% ~ N(-2, 1.5^2)
(contacts), plus a constant identity term c0 ~ U[0, 100].

Ansatz tapes are lists of (name, qubits, param_index | None), parameter order
layer / qubit / gate as in the reference's global ansatz
(Qiskit_VQE/independent_vqe.py:104-110, 160-174).
"""
import numpy as np

SEED_H = 20260


def z_hamiltonian(n, T, wmax, rng):
    """T distinct Z-strings (capped at the number of distinct supports of
    weight 1..wmax on n qubits, e.g. 246 for n = 8, wmax = 6)."""
    from math import comb
    T = min(T, sum(comb(n, w) for w in range(1, min(wmax, n) + 1)))
    words, coeffs, seen = [], [], set()
    while len(words) < T:
        w = int(rng.integers(1, wmax + 1))
        support = tuple(sorted(rng.choice(n, size=w, replace=False).tolist()))
        if support in seen:
            continue
        seen.add(support)
        word = ['I'] * n
        for q in support:
            word[q] = 'Z'
        words.append(''.join(word))
        coeffs.append(float(rng.uniform(50, 500)) if rng.random() < 0.2 else float(rng.normal(-2.0, 1.5)))
    words.append('I' * n)
    coeffs.append(float(rng.uniform(0, 100)))
    return words, coeffs


def chain_pairs(n, pattern):
    """CX pairs of one entangling layer.

    'linear'          CX(0,1), CX(1,2), ..., CX(n-2,n-1)  (efficientsu2.py:37-38)
    'reverse_linear'  CX(n-2,n-1), ..., CX(0,1)           (efficientsu2.py:41-43)
    'reverse_chain'   CX(n-1,n-2), ..., CX(1,0): the "39->38->...->0" chain of
                      the 40-qubit batched configuration read literally
    """
    if pattern == 'linear':
        return [(i, i + 1) for i in range(n - 1)]
    if pattern == 'reverse_linear':
        return [(i, i + 1) for i in range(n - 2, -1, -1)]
    if pattern == 'reverse_chain':
        return [(i, i - 1) for i in range(n - 1, 0, -1)]
    raise ValueError(pattern)


def entangling_pairs(n, layer, pattern):
    """CX (control, target) pairs of layer `layer` for the SU(2) ansatz patterns
    of Qiskit_VQE/efficientsu2.py:23-75 (linear, reverse_linear, circular,
    full, pairwise, sca)."""
    if pattern == 'full':
        return [(i, j) for i in range(n) for j in range(i + 1, n)]
    if pattern in ('linear', 'reverse_linear'):
        return chain_pairs(n, pattern)
    if pattern == 'circular':
        return chain_pairs(n, 'linear') + [(n - 1, 0)]
    if pattern == 'pairwise':
        return [(i, i + 1) for i in range(layer % 2, n - 1, 2)]
    if pattern == 'sca':
        shift = layer % n
        if shift % 2 == 0:
            out = [(i, i + 1) for i in range(shift - 1, -1, -1)]
            out.append((n - 1, 0))
            out += [(i - 1, i) for i in range(n - 1, shift, -1)]
        else:
            out = [(i + 1, i) for i in range(n - shift - 1, n - 1)]
            out.append((0, n - 1))
            out += [(i + 1, i) for i in range(n - shift - 1)]
        return out
    raise ValueError(pattern)


def su2_ansatz(n, reps, gates=('ry', 'rz'), pattern='linear'):
    """reps+1 rotation layers with entangling layers between them, parameters
    in layer / qubit / gate order (Qiskit_VQE/independent_vqe.py:160-174)."""
    ops, p = [], 0
    for layer in range(reps + 1):
        for q in range(n):
            for g in gates:
                ops.append((g, (q,), p))
                p += 1
        if layer < reps:
            for c, t in entangling_pairs(n, layer, pattern):
                ops.append(('cx', (c, t), None))
    return ops, p


def layered_ansatz(n, layers, gates=('ry',), pattern='linear', final_rotations=False):
    """`layers` x (rotations on every qubit -> CX layer) [+ final rotation layer]."""
    ops, p = [], 0
    reps = layers + (1 if final_rotations else 0)
    for layer in range(reps):
        for q in range(n):
            for g in gates:
                ops.append((g, (q,), p))
                p += 1
        if layer < layers:
            for c, t in chain_pairs(n, pattern):
                ops.append(('cx', (c, t), None))
    return ops, p


def config(index, seed_offset=0):
    """Workload of the benchmark configuration `index` (0-based)."""
    rng = np.random.default_rng(SEED_H + index + seed_offset)
    if index == 0:
        n = 12
        words, coeffs = z_hamiltonian(n, 300, 4, rng)
        ops, P = layered_ansatz(n, 4, ('ry',), 'linear')
        x = rng.uniform(-np.pi, np.pi, P)
        return dict(name='plain-12q', n=n, ops=ops, nparams=P, words=words, coeffs=coeffs, x=x[None, :],
                    max_bond=16, shots=1000)
    if index == 1:
        n = 40
        words, coeffs = z_hamiltonian(n, 4000, 6, rng)
        ops, P = layered_ansatz(n, 4, ('ry',), 'reverse_chain')
        x0 = rng.uniform(-np.pi, np.pi, P)
        delta = rng.choice([-1.0, 1.0], size=(128, P))
        x = np.empty((256, P))
        x[0::2] = x0 + 0.1 * delta
        x[1::2] = x0 - 0.1 * delta
        return dict(name='spsa-40q', n=n, ops=ops, nparams=P, words=words, coeffs=coeffs, x=x, max_bond=16)
    if index == 2:
        n = 60
        words, coeffs = z_hamiltonian(n, 8000, 6, rng)
        ops, P = layered_ansatz(n, 4, ('ry', 'rz'), 'linear')
        x = rng.uniform(-np.pi, np.pi, (32, P))
        return dict(name='adjoint-60q', n=n, ops=ops, nparams=P, words=words, coeffs=coeffs, x=x, max_bond=16)
    if index == 3:
        parts = []
        for j in range(64):
            n = int(rng.integers(8, 33))
            T = int(rng.integers(200, 1501))
            words, coeffs = z_hamiltonian(n, T, 6, rng)
            pattern = 'linear' if j % 2 == 0 else 'reverse_chain'
            ops, P = layered_ansatz(n, 4, ('ry',), pattern)
            x = rng.uniform(-np.pi, np.pi, P)
            parts.append(dict(n=n, ops=ops, nparams=P, words=words, coeffs=coeffs, x=x[None, :], max_bond=16))
        return dict(name='uneven-64', parts=parts, shots=1000)
    raise ValueError(index)
