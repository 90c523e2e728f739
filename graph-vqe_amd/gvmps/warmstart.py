"""Partition warm start: interaction graph, communities, independent local
VQEs, assembled global parameters (Qiskit_VQE/independent_vqe.py).

  interaction_graph   qubits as nodes; every pair of qubits in a term's
                      support gains |c| of that term (:57-64)
  partition           Louvain communities per connected component (seeded),
                      then Kernighan-Lin bisection until every block has at
                      most max_block_qubits qubits (:140-158)
  local_models        each block's terms (support inside the block; the
                      constant goes to the first block), on max(2, |block|)
                      qubits, and the global indices of the block's
                      rotations in layer / qubit / gate order (:65-100)
  partition_warm_start  local L-BFGS-B solves on the GPU (adjoint gradient),
                      global parameters assembled per qubit (:257-305)

The global parameter order is (layer, qubit, gate), as in su2_ansatz.
"""
from itertools import combinations

import numpy as np

from . import workloads


def interaction_graph(n, words, coeffs):
    import networkx as nx
    g = nx.Graph()
    g.add_nodes_from(range(n))
    for w, c in zip(words, coeffs):
        c = complex(c)
        if not abs(c):
            continue
        support = [q for q, p in enumerate(w) if p != 'I']
        for a, b in combinations(support, 2):
            weight = g.get_edge_data(a, b, {}).get('weight', 0.0)
            g.add_edge(a, b, weight=weight + float(abs(c)))
    return g


def partition(graph, max_block_qubits=12, seed=0):
    import networkx as nx
    communities = []
    for component in nx.connected_components(graph):
        sub = graph.subgraph(component)
        if sub.number_of_edges():
            communities.extend(nx.community.louvain_communities(sub, weight='weight', seed=seed))
        else:
            communities.extend({q} for q in component)
    pending, result = list(communities), []
    while pending:
        block = pending.pop()
        if max_block_qubits is None or len(block) <= max_block_qubits:
            result.append(sorted(block))
        else:
            left, right = nx.community.kernighan_lin_bisection(graph.subgraph(sorted(block)), weight='weight',
                                                               seed=seed)
            pending.extend((left, right))
    return sorted(result, key=lambda b: b[0])


def local_models(n, words, coeffs, blocks, reps, gates, pattern):
    owner = {q: i for i, b in enumerate(blocks) for q in b}
    local = [[] for _ in blocks]
    cross = []
    for w, c in zip(words, coeffs):
        c = complex(c)
        if c == 0:
            continue
        support = [q for q, p in enumerate(w) if p != 'I']
        owners = {owner[q] for q in support}
        if len(owners) > 1:
            cross.append((w, c.real))
            continue
        i = next(iter(owners)) if owners else 0
        b = blocks[i]
        width = max(2, len(b))
        local[i].append((''.join(w[q] for q in b) + 'I' * (width - len(b)), c.real))
    models = []
    for i, b in enumerate(blocks):
        width = max(2, len(b))
        terms = local[i] or [('I' * width, 0.0)]
        ops, P = workloads.su2_ansatz(len(b), reps, gates, pattern) if len(b) > 1 else \
            _single_qubit_ansatz(reps, gates)
        models.append(dict(
            qubits=list(b), width=width, ops=ops, nparams=P,
            words=[t[0] for t in terms], coeffs=[t[1] for t in terms],
            constant=all(set(t[0]) <= {'I'} for t in terms),
            parameter_indices=[(layer * n + q) * len(gates) + g
                               for layer in range(reps + 1) for q in b for g in range(len(gates))]))
    return models, cross


def _single_qubit_ansatz(reps, gates):
    """A one-qubit block keeps an idle spectator (the backend needs 2 qubits)."""
    ops, p = [], 0
    for _ in range(reps + 1):
        for g in gates:
            ops.append((g, (0,), p))
            p += 1
    return ops, p


def partition_warm_start(n, words, coeffs, reps=3, gates=('ry',), pattern='linear', theta0=None, seed=0,
                         max_block_qubits=12, maxiter=100):
    """Warm-start parameters: local L-BFGS-B solves (adjoint gradients on the
    GPU) from theta0's block entries, assembled into the global vector."""
    from scipy.optimize import minimize
    from .engine import Batch, Problem
    P = (reps + 1) * n * len(gates)
    rng = np.random.default_rng(seed)
    theta = rng.uniform(-np.pi, np.pi, P) if theta0 is None else np.asarray(theta0, dtype=float).copy()
    blocks = partition(interaction_graph(n, words, coeffs), max_block_qubits, seed)
    models, cross = local_models(n, words, coeffs, blocks, reps, gates, pattern)
    local = []
    for m in models:
        x0 = theta[m['parameter_indices']]
        if m['constant'] or not maxiter:
            local.append(dict(params=x0, nit=0))
            continue
        prob = Problem(m['width'], m['ops'], m['nparams'], m['words'], m['coeffs'], max_bond=16, abs_cut=0,
                       rel_cut=0)
        batch = Batch([prob], [1])

        def fg(x, batch=batch):
            e, g = batch.value_grad(x)
            return float(e.cpu()[0]), g.cpu().numpy()
        res = minimize(fg, x0, jac=True, method='L-BFGS-B',
                       options=dict(maxiter=maxiter, ftol=1e-12, gtol=1e-8))
        local.append(dict(params=res.x, nit=int(res.nit), energy=float(res.fun)))
    for m, r in zip(models, local):
        theta[m['parameter_indices']] = r['params']
    return theta, dict(blocks=blocks, crossing_terms=len(cross), local=local)
