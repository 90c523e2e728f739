"""Host checks of the partition warm start (gvmps.warmstart) against the
semantics of the reference's IndependentSubspaceVQE
(Qiskit_VQE/independent_vqe.py):

  * interaction graph: every pair of qubits of a term's support gains |c|
    (:57-64);
  * blocks: Louvain communities per connected component, Kernighan-Lin
    bisection only of blocks above max_block_qubits; blocks cover every
    qubit once (:65-68, 140-158);
  * local Hamiltonians: a term goes to the block owning its whole support,
    the constant to the first block, crossing terms stay global, local width
    max(2, |block|) (:70-85);
  * the local -> global parameter map (layer, qubit, gate) (:86-100, 276-277).

The operator identity sum_b embed(H_b) + H_cross = H is checked on dense
matrices, so a wrong character order in the local words fails it."""
import numpy as np
import pytest

from gvmps import warmstart, workloads
from oracle.dense import PAULI


def random_words(n, T, rng, letters='IXYZ'):
    words, coeffs = [], []
    seen = set()
    while len(words) < T:
        w = ''.join(rng.choice(list(letters), n))
        if w in seen or set(w) == {'I'}:
            continue
        seen.add(w)
        words.append(w)
        coeffs.append(float(rng.normal()))
    words.append('I' * n)
    coeffs.append(0.75)
    return words, coeffs


def dense_operator(words, coeffs, n):
    """sum_k c_k P_k with word[q] on qubit q (qubit 0 = most significant axis)."""
    H = np.zeros((2 ** n, 2 ** n), dtype=complex)
    for w, c in zip(words, coeffs):
        M = np.ones((1, 1), dtype=complex)
        for q in range(n):
            M = np.kron(M, PAULI[w[q]])
        H += c * M
    return H


def embed(word, block, n):
    """Local word (local qubit j = block[j], padding beyond |block|) on n qubits."""
    out = ['I'] * n
    for j, q in enumerate(block):
        out[q] = word[j]
    assert set(word[len(block):]) <= {'I'}, 'padding qubits carry no Pauli'
    return ''.join(out)


def test_interaction_graph_weights():
    words = ['ZZII', 'ZIZI', 'IZZZ', 'IIIZ', 'ZZII']
    coeffs = [2.0, -3.0, 0.5, 7.0, -1.0]
    g = warmstart.interaction_graph(4, words, coeffs)
    assert sorted(g.nodes) == [0, 1, 2, 3]
    w = {tuple(sorted(e)): d['weight'] for *e, d in g.edges(data=True)}
    assert w == {(0, 1): 3.0, (0, 2): 3.0, (1, 2): 0.5, (1, 3): 0.5, (2, 3): 0.5}
    # zero coefficients add no edge (independent_vqe.py:62-63)
    g0 = warmstart.interaction_graph(3, ['ZZI'], [0.0])
    assert g0.number_of_edges() == 0 and g0.number_of_nodes() == 3


@pytest.mark.parametrize('seed', range(3))
def test_blocks_cover_and_respect_limit(seed):
    r = np.random.default_rng(seed)
    n = 30
    words, coeffs = workloads.z_hamiltonian(n, 400, 4, r)
    g = warmstart.interaction_graph(n, words, coeffs)
    for limit in (12, 5, 3):
        blocks = warmstart.partition(g, limit, seed)
        flat = sorted(q for b in blocks for q in b)
        assert flat == list(range(n))
        assert all(1 <= len(b) <= limit for b in blocks)
        assert all(b == sorted(b) for b in blocks)
        assert [b[0] for b in blocks] == sorted(b[0] for b in blocks)
    # no limit: plain Louvain communities of each component
    unbounded = warmstart.partition(g, None, seed)
    assert sorted(q for b in unbounded for q in b) == list(range(n))


def test_kernighan_lin_only_above_limit():
    """Two 8-cliques joined by one weak edge: Louvain gives the two cliques;
    a limit of 8 keeps them, a limit of 4 bisects each clique into halves of
    4 (KL), and isolated qubits stay singleton blocks."""
    n = 18
    words, coeffs = [], []
    for base in (0, 8):
        for a in range(base, base + 8):
            for b in range(a + 1, base + 8):
                w = ['I'] * n
                w[a] = w[b] = 'Z'
                words.append(''.join(w))
                coeffs.append(5.0)
    w = ['I'] * n
    w[7] = w[8] = 'Z'
    words.append(''.join(w))
    coeffs.append(0.01)
    g = warmstart.interaction_graph(n, words, coeffs)
    b8 = warmstart.partition(g, 8, 0)
    assert b8 == [list(range(8)), list(range(8, 16)), [16], [17]]
    b4 = warmstart.partition(g, 4, 0)
    assert sorted(len(b) for b in b4) == [1, 1, 4, 4, 4, 4]
    for half in b4:
        if len(half) == 4:
            assert set(half) <= set(range(8)) or set(half) <= set(range(8, 16))


@pytest.mark.parametrize('seed', range(3))
def test_local_models_decompose_hamiltonian(seed):
    r = np.random.default_rng(10 + seed)
    n = 7
    words, coeffs = random_words(n, 40, r)
    # sparse supports so that some terms stay inside blocks
    words = [''.join(c if r.random() < 0.35 else 'I' for c in w) if set(w) != {'I'} else w for w in words]
    blocks = [[0, 3], [1, 2, 5], [4], [6]]
    models, cross = warmstart.local_models(n, words, coeffs, blocks, 2, ('ry', 'rz'), 'linear')
    assert len(models) == len(blocks)
    parts_w, parts_c = [], []
    for i, (m, b) in enumerate(zip(models, blocks)):
        assert m['qubits'] == b and m['width'] == max(2, len(b))
        assert all(len(w) == m['width'] for w in m['words'])
        for w, c in zip(m['words'], m['coeffs']):
            parts_w.append(embed(w, b, n))
            parts_c.append(c)
        const = [c for w, c in zip(m['words'], m['coeffs']) if set(w) == {'I'}]
        # the constant term goes to the first block only (independent_vqe.py:79)
        assert (0.75 in const) == (i == 0)
    for w, c in cross:
        owners = {next(j for j, b in enumerate(blocks) if q in b) for q, p in enumerate(w) if p != 'I'}
        assert len(owners) > 1
        parts_w.append(w)
        parts_c.append(c)
    nz = [(w, c) for w, c in zip(words, coeffs) if c != 0]
    assert len(parts_w) == len(nz) + sum(1 for m in models if m['words'] == ['I' * m['width']]
                                        and m['coeffs'] == [0.0])
    np.testing.assert_allclose(dense_operator(parts_w, parts_c, n), dense_operator(words, coeffs, n),
                               atol=1e-13)


def test_parameter_map_follows_layer_qubit_gate_order():
    """Local rotation j of block b is the global rotation of the same gate on
    qubit b[q_local] in the same layer: index (layer*n + q)*G + g."""
    n, reps, gates = 9, 3, ('ry', 'rz')
    ops_g, P = workloads.su2_ansatz(n, reps, gates, 'linear')
    assert P == (reps + 1) * n * len(gates)
    where = {}
    layer = 0
    for name, qs, p in ops_g:
        if name == 'cx':
            continue
        where[p] = (p // (n * len(gates)), qs[0], name)
    blocks = [[0, 1, 2, 5], [3, 4], [6, 7, 8]]
    words = ['Z' * n]
    models, _ = warmstart.local_models(n, words, [1.0], blocks, reps, gates, 'linear')
    seen = []
    for m, b in zip(models, blocks):
        assert m['nparams'] == (reps + 1) * len(b) * len(gates)
        assert len(m['parameter_indices']) == m['nparams']
        for name, qs, p in m['ops']:
            if name == 'cx':
                assert all(q < len(b) for q in qs)
                continue
            layer = p // (len(b) * len(gates))
            gp = m['parameter_indices'][p]
            assert where[gp] == (layer, b[qs[0]], name), (b, p)
            seen.append(gp)
    assert sorted(seen) == list(range(P))


def test_single_qubit_block_keeps_a_spectator():
    models, _ = warmstart.local_models(3, ['ZII', 'IZZ'], [1.0, 2.0], [[0], [1, 2]], 1, ('ry',), 'linear')
    m = models[0]
    assert m['width'] == 2 and m['words'] == ['ZI'] and m['nparams'] == 2
    assert all(qs == (0,) for _, qs, _ in m['ops'])
    assert models[1]['words'] == ['ZZ'] and models[1]['coeffs'] == [2.0]
