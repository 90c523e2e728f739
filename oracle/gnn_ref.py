"""CPU restatement of the warm-start graph network (test infrastructure).

The network has no reference counterpart (the reference's warm start is the
partition heuristic, Qiskit_VQE/independent_vqe.py:140-305); this module is the
checker for the library's FP32 HIP kernels (csrc/kernels_gnn.hip): numpy
float64 forward, hand-written backward, same flat parameter layout
[W_in, b_in, (Ws, Wn, b) per layer, W_out, b_out].  It is pinned against
torch autograd and central differences in tests/test_oracle.py.
"""
import numpy as np


def layout(F, H, L, O):
    off, parts = 0, {}

    def take(name, shape):
        nonlocal off
        size = int(np.prod(shape))
        parts[name] = (off, shape)
        off += size
    take('w_in', (H, F))
    take('b_in', (H,))
    for l in range(L):
        take(f'ws{l}', (H, H))
        take(f'wn{l}', (H, H))
        take(f'b{l}', (H,))
    take('w_out', (O, H))
    take('b_out', (O,))
    return parts, off


def unpack(theta, F, H, L, O):
    parts, total = layout(F, H, L, O)
    assert theta.size == total
    return {k: theta[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in parts.items()}


def aggregate(rows, cols, w, Hm):
    """A[v] = sum_{e in row v} w[e] H[cols[e]]."""
    A = np.zeros_like(Hm)
    for v in range(len(rows) - 1):
        for e in range(rows[v], rows[v + 1]):
            A[v] += w[e] * Hm[cols[e]]
    return A


def forward(theta, X, graph, F, H, L, O):
    p = unpack(np.asarray(theta, dtype=np.float64), F, H, L, O)
    rows, cols, w = graph['rows'], graph['cols'], graph['weights']
    cache = {}
    Z = X @ p['w_in'].T + p['b_in']
    Hm = np.maximum(Z, 0)
    cache['Z0'], cache['H0'] = Z, Hm
    for l in range(L):
        A = aggregate(rows, cols, w, Hm)
        Z = Hm @ p[f'ws{l}'].T + A @ p[f'wn{l}'].T + p[f'b{l}']
        cache[f'A{l + 1}'] = A
        Hm = np.maximum(Z, 0)
        cache[f'Z{l + 1}'], cache[f'H{l + 1}'] = Z, Hm
    Y = Hm @ p['w_out'].T + p['b_out']
    return Y, cache


def loss_and_grad(theta, X, graph, targets, F, H, L, O):
    """(mean(1 - cos(Y - T)), d loss / d theta)."""
    theta = np.asarray(theta, dtype=np.float64)
    p = unpack(theta, F, H, L, O)
    Y, c = forward(theta, X, graph, F, H, L, O)
    D = Y - targets
    loss = float(np.mean(1 - np.cos(D)))
    dY = np.sin(D) / D.size
    g = {}
    g['w_out'] = dY.T @ c[f'H{L}']
    g['b_out'] = dY.sum(0)
    dH = dY @ p['w_out']
    rows_t, cols_t, w_t = graph['rows_t'], graph['cols_t'], graph['weights_t']
    for l in range(L - 1, -1, -1):
        dZ = dH * (c[f'Z{l + 1}'] > 0)
        g[f'ws{l}'] = dZ.T @ c[f'H{l}']
        g[f'wn{l}'] = dZ.T @ c[f'A{l + 1}']
        g[f'b{l}'] = dZ.sum(0)
        dA = dZ @ p[f'wn{l}']
        dH = dZ @ p[f'ws{l}'] + aggregate(rows_t, cols_t, w_t, dA)
    dZ = dH * (c['Z0'] > 0)
    g['w_in'] = dZ.T @ X
    g['b_in'] = dZ.sum(0)
    parts, total = layout(F, H, L, O)
    flat = np.zeros(total)
    for k, (o, s) in parts.items():
        flat[o:o + int(np.prod(s))] = g[k].reshape(-1)
    return loss, flat
