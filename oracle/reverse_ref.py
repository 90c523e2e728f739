"""Reverse-mode gradient restatement of the reference's ReverseMPS (test infrastructure).

Follows test/tools/reverse_mps.py:28-99 (and the static-gate/ordered-CX tape
variant :141-184): the circuit is re-simulated on an MPS whose tensors are
indexed by QUBIT (tensor q = qubit q, no reversal), a CNOT on neighbouring
qubits doubles the bond exactly (control tensor A -> A (x) |s>, target tensor
B -> stack(B, flip(B)) -- no SVD), terms are contracted in chunks with a
prefix DAG, and torch autograd returns dE/dtheta.  torch on the CPU, float64 /
complex128.

A tape op is (name, qubits, angle_ref) where angle_ref is an int parameter
index, a float constant, or None for fixed gates.
"""
import numpy as np
import torch

_FIXED = {
    'h': [[2 ** -0.5, 2 ** -0.5], [2 ** -0.5, -2 ** -0.5]], 'x': [[0, 1], [1, 0]],
    'y': [[0, -1j], [1j, 0]], 'z': [[1, 0], [0, -1]], 's': [[1, 0], [0, 1j]], 'sdg': [[1, 0], [0, -1j]],
}
DT = torch.complex128


def _rot(name, angle):
    c, s = torch.cos(angle / 2), torch.sin(angle / 2)
    if name == 'ry':
        return torch.stack((c, -s, s, c)).reshape(2, 2).to(DT)
    if name == 'rx':
        return torch.stack((c + 0j, -1j * s, -1j * s, c + 0j)).reshape(2, 2)
    return torch.diag(torch.stack((torch.exp(-0.5j * angle), torch.exp(0.5j * angle))))


def build_sites(n, ops, x, initial=None):
    """Bond-doubling site tensors.  reverse_mps.py:39-57, 152-184."""
    vec = {None: [1, 0], '0': [1, 0], '1': [0, 1], '+': [2 ** -.5, 2 ** -.5], '-': [2 ** -.5, -2 ** -.5]}[initial]
    sites = [torch.tensor(vec, dtype=DT).reshape(1, 2, 1) for _ in range(n)]
    eye = torch.eye(2, dtype=DT)
    for name, qubits, ref in ops:
        if name == 'cx':
            c, t = qubits
            lo, hi = sorted(qubits)
            if hi != lo + 1:
                raise ValueError("reverse restatement handles neighbouring CX only")
            a, b = sites[lo], sites[hi]
            if c == lo:     # control on the left tensor
                aa = torch.einsum('aib,is->aibs', a, eye)
                bb = torch.stack((b, b.flip(1)), dim=1)
            else:           # control on the right tensor
                aa = torch.stack((a, a.flip(1)), dim=-1)
                bb = torch.einsum('aib,is->asib', b, eye)
            sites[lo] = aa.reshape(a.shape[0], 2, a.shape[2] * 2)
            sites[hi] = bb.reshape(b.shape[0] * 2, 2, b.shape[2])
        else:
            q = qubits[0]
            if ref is None:
                m = torch.tensor(_FIXED[name], dtype=DT)
            else:
                angle = x[ref] if isinstance(ref, int) else torch.tensor(float(ref), dtype=torch.float64)
                m = _rot(name, angle)
            sites[q] = torch.einsum('ij,ajb->aib', m, sites[q])
    return sites


def energy_chunk(sites, words, coeffs):
    """Prefix-shared contraction of one chunk.  reverse_mps.py:59-82."""
    env = torch.ones((1, 1, 1), dtype=DT)
    previous = [0] * len(words)
    for q, a in enumerate(sites):
        nodes, parents, paulis, current = {}, [], [], []
        for prev, word in zip(previous, words):
            key = (prev, word[q])
            if key not in nodes:
                nodes[key] = len(nodes)
                parents.append(prev)
                paulis.append(word[q])
            current.append(nodes[key])
        inc = env[torch.tensor(parents)]
        tmp = torch.einsum('tab,bic->taic', inc, a)                      # E[bra,ket] A[ket,i,r]
        flip = torch.tensor([[1, 0] if p in 'XY' else [0, 1] for p in paulis])
        tmp = tmp.gather(2, flip[:, None, :, None].expand(-1, tmp.shape[1], -1, tmp.shape[3]))
        ph = torch.tensor([[1, -1] if p == 'Z' else [-1j, 1j] if p == 'Y' else [1, 1] for p in paulis], dtype=DT)
        tmp = tmp * ph[:, None, :, None]
        env = torch.einsum('aib,taic->tbc', a.conj(), tmp)
        previous = current
    vals = env[torch.tensor(previous), 0, 0].real
    return vals if coeffs is None else (vals * torch.as_tensor(coeffs, dtype=torch.float64)).sum()


def value_gradient(n, ops, words, coeffs, x, chunk=128, initial=None):
    """(E, dE/dx).  reverse_mps.py:84-99: chunked autograd to the site leaves,
    then one backward through the site construction."""
    xt = torch.tensor(np.asarray(x, dtype=float), dtype=torch.float64, requires_grad=True)
    actual = build_sites(n, ops, xt, initial)
    leaves = [a.detach().requires_grad_(True) for a in actual]
    grads = [torch.zeros_like(a) for a in leaves]
    value = 0.0
    coeffs = np.asarray(coeffs, dtype=float)
    for start in range(0, len(words), chunk):
        e = energy_chunk(leaves, words[start:start + chunk], coeffs[start:start + chunk])
        g = torch.autograd.grad(e, leaves)
        value += float(e.detach())
        for tot, gi in zip(grads, g):
            tot.add_(gi)
    active = [(a, g) for a, g in zip(actual, grads) if a.requires_grad]
    if active:
        grad = torch.autograd.grad([a for a, _ in active], xt, [g for _, g in active])[0]
    else:
        grad = torch.zeros_like(xt)
    return value, grad.detach().numpy()


def values(n, ops, words, x, chunk=128, initial=None):
    """Unweighted <P_k> (ReverseTapeMPS.values, reverse_mps.py:186-194)."""
    with torch.no_grad():
        sites = build_sites(n, ops, torch.as_tensor(np.asarray(x, dtype=float)), initial)
        out = [energy_chunk(sites, words[i:i + chunk], None).numpy() for i in range(0, len(words), chunk)]
    return np.concatenate(out) if out else np.zeros(0)
