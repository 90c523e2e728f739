"""Algorithmic work of an energy evaluation, a sampling pass and a state
preparation, counted from the workload alone (no kernel is run).

Bonds are the zero-cutoff allocation bound: every two-site gate on sites
(s, s+1) sets bond s+1 to min(2 D_s, 2 D_{s+2}, chi), SWAP-routed pairs
included (test/tools/graph_resources.py:12-32, restated in site order).

Energy, the left/right-environment algorithm with shared prefixes:
  f_s = 16 (Dl^2 Dr + Dl Dr^2) real flops per environment transfer at site s
        (two physical indices, Dl^2 Dr + Dl Dr^2 complex multiply-adds each,
        8 flops per complex multiply-add);
  F_E = sum_s N_s f_s + 2 sum_s f_s,
  where N_s counts the distinct prefixes word[0..s] at site s over the
  distinct words whose span [first non-I site, last non-I site] contains s,
  and the last term is the left and right identity environments.
The reference's own chunked prefix DAG (sharded_mps.cuh:614-687: sorted
words, chunks of at most 128, every site transferred for every distinct
prefix of the chunk) is counted beside it.

Sampling, per site: shots * 8 (2 Dl Dr + 2 Dr^2) for the conditional
probabilities plus 8 (2 Dl Dr^2 + 2 Dl^2 Dr) for the right environment.
State preparation: 4 (4 m^2 n + 8 m n^2 + 9 n^3) per two-site gate with
m >= n the sides of Theta (the Golub-Reinsch nominal count).
"""
import numpy as np


def allocated_bonds(n, ops, max_bond=16):
    """Site-order bond dimensions (n + 1 values) after the tape's two-qubit
    gates under the zero-cutoff allocation rule; also the (m, n) shapes of
    every Theta block, in application order."""
    bonds = [1] * (n + 1)
    shapes = []
    for name, qubits, _ in ops:
        if name not in ('cx', 'swap'):
            continue
        a, b = sorted((n - 1 - qubits[0], n - 1 - qubits[1]))
        for cut in [*range(b - 1, a, -1), a, *range(a + 1, b)]:
            m, k = 2 * bonds[cut], 2 * bonds[cut + 2]
            shapes.append((m, k))
            bonds[cut + 1] = min(m, k, max_bond)
    return bonds, shapes


def _site_words(words):
    return sorted({w[::-1] for w in words if set(w) != {'I'}})


def transfer_flops(bonds):
    return np.array([16.0 * (bonds[s] ** 2 * bonds[s + 1] + bonds[s] * bonds[s + 1] ** 2)
                     for s in range(len(bonds) - 1)])


def energy_work(n, ops, words, max_bond=16, chunk=128):
    """Algorithmic flops of one energy evaluation (qubit-order words)."""
    bonds, _ = allocated_bonds(n, ops, max_bond)
    f = transfer_flops(bonds)
    sw = _site_words(words)
    nodes = np.zeros(n, dtype=np.int64)
    for s in range(n):
        seen = set()
        for w in sw:
            core = [i for i, c in enumerate(w) if c != 'I']
            if core[0] <= s <= core[-1]:
                seen.add(w[:s + 1])
        nodes[s] = len(seen)
    f_e = float(nodes @ f + 2.0 * f.sum())
    dag = 0
    for a in range(0, len(sw), chunk):
        part = sw[a:a + chunk]
        dag_nodes = [len({w[:s + 1] for w in part}) for s in range(n)]
        dag += float(np.asarray(dag_nodes) @ f)
    return dict(bonds=bonds, node_site_transfers=int(nodes.sum()), flops_per_eval=f_e,
                reference_dag_flops_per_eval=dag)


def sampling_flops(bonds, shots):
    out = 0.0
    for s in range(len(bonds) - 1):
        dl, dr = bonds[s], bonds[s + 1]
        out += shots * 8.0 * (2 * dl * dr + 2 * dr * dr) + 8.0 * (2 * dl * dr * dr + 2 * dl * dl * dr)
    return out


def svd_flops(shapes):
    out = 0.0
    for m, k in shapes:
        M, N = max(m, k), min(m, k)
        out += 4.0 * (4.0 * M * M * N + 8.0 * M * N * N + 9.0 * N ** 3)
    return out
