"""Learned warm start: node features, interaction-graph batches and the
graph network (HIP kernels behind gvmps_gnn_*).

Nodes are qubits; an edge (u, v) carries the summed |c| of the Z-strings that
contain both qubits (the partition warm start's interaction graph,
Qiskit_VQE/independent_vqe.py:57-64), row-normalised per destination.  Node v
gets 8 features of its terms (scaled by the Hamiltonian's RMS coefficient):
signed and absolute coefficient sums, its single-Z field, its share of terms,
even- and odd-weight signed sums, its chain position and a constant.  The
network predicts the `outputs` rotation angles of qubit v, one per ansatz
layer (global parameter index layer * n + v).
"""
import ctypes

import numpy as np

from ._lib import Graph, check, lib

N_FEATURES = 8


def node_features(n, words, coeffs):
    c = np.real(np.asarray(coeffs, dtype=complex))
    support = [[q for q, p in enumerate(w) if p != 'I'] for w in words]
    active = [k for k, s in enumerate(support) if s]
    scale = float(np.sqrt(np.mean(c[active] ** 2))) if active else 1.0
    scale = scale if scale > 0 else 1.0
    F = np.zeros((n, N_FEATURES))
    cnt = np.zeros(n)
    for k in active:
        s = support[k]
        for q in s:
            cnt[q] += 1
            F[q, 0] += c[k]
            F[q, 1] += abs(c[k])
            F[q, 4 if len(s) % 2 == 0 else 5] += c[k]
        if len(s) == 1:
            F[s[0], 2] += c[k]
    norm = np.sqrt(cnt + 1.0)
    F[:, 0] /= scale * norm
    F[:, 1] /= scale * (cnt + 1.0)
    F[:, 2] /= scale
    F[:, 3] = cnt / max(1, len(active))
    F[:, 4] /= scale * norm
    F[:, 5] /= scale * norm
    F[:, 6] = np.arange(n) / max(1, n - 1)
    F[:, 7] = 1.0
    return F.astype(np.float32)


def edges(n, words, coeffs):
    """Coupling weights W[u, v] = sum |c| over terms containing u and v."""
    W = np.zeros((n, n))
    for w, cf in zip(words, coeffs):
        s = [q for q, p in enumerate(w) if p != 'I']
        if len(s) > 1:
            a = abs(complex(cf))
            idx = np.array(s)
            W[np.ix_(idx, idx)] += a
    np.fill_diagonal(W, 0.0)
    return W


class GraphBatch:
    """Concatenated graphs on the device, CSR by destination and by source."""

    def __init__(self, graphs, device=None):
        import torch
        self.torch = torch
        dev = f'cuda:{torch.cuda.current_device() if device is None else device}'
        feats, rows, cols, ws = [], [0], [], []
        rows_t, cols_t, ws_t = [0], [], []
        self.offsets = [0]
        base = 0
        for F, W in graphs:
            n = F.shape[0]
            feats.append(F)
            deg = W.sum(axis=0)  # incoming weight per destination
            Wn = np.divide(W, deg[None, :], out=np.zeros_like(W), where=deg[None, :] > 0)  # Wn[u, v]
            for v in range(n):
                src = np.nonzero(Wn[:, v])[0]
                cols.extend((src + base).tolist())
                ws.extend(Wn[src, v].tolist())
                rows.append(len(cols))
            for u in range(n):
                dst = np.nonzero(Wn[u, :])[0]
                cols_t.extend((dst + base).tolist())
                ws_t.extend(Wn[u, dst].tolist())
                rows_t.append(len(cols_t))
            base += n
            self.offsets.append(base)
        self.n_nodes = base
        T = lambda a, dt: torch.as_tensor(np.asarray(a, dtype=dt), device=dev).contiguous()
        self.features = T(np.concatenate(feats) if feats else np.zeros((0, N_FEATURES)), np.float32)
        self.rows, self.cols, self.w = T(rows, np.int32), T(cols or [0], np.int32), T(ws or [0.0], np.float32)
        self.rows_t, self.cols_t = T(rows_t, np.int32), T(cols_t or [0], np.int32)
        self.w_t = T(ws_t or [0.0], np.float32)
        self.struct = Graph(self.n_nodes, len(cols), self.features.data_ptr(), self.rows.data_ptr(),
                            self.cols.data_ptr(), self.w.data_ptr(), self.rows_t.data_ptr(), self.cols_t.data_ptr(),
                            self.w_t.data_ptr())

    def host_graph(self):
        """numpy copies (for the CPU oracle)."""
        g = lambda t: t.cpu().numpy()
        return dict(rows=g(self.rows), cols=g(self.cols), weights=g(self.w).astype(np.float64),
                    rows_t=g(self.rows_t), cols_t=g(self.cols_t), weights_t=g(self.w_t).astype(np.float64))


class WarmStartGNN:
    def __init__(self, hidden=64, layers=3, outputs=4, seed=0, device=None):
        import torch
        self.torch = torch
        self.device = torch.cuda.current_device() if device is None else int(device)
        self.shape = (N_FEATURES, hidden, layers, outputs)
        h = ctypes.c_void_p()
        check(lib().gvmps_gnn_create(ctypes.byref(h), N_FEATURES, hidden, layers, outputs, int(seed), self.device))
        self._h = h
        self.P = lib().gvmps_gnn_param_count(h)

    def params(self):
        out = np.zeros(self.P, dtype=np.float32)
        check(lib().gvmps_gnn_params(self._h, out.ctypes.data_as(ctypes.c_void_p), None))
        return out

    def set_params(self, theta):
        v = np.ascontiguousarray(np.asarray(theta, dtype=np.float32))
        if v.size != self.P:
            raise ValueError(f'expected {self.P} parameters')
        check(lib().gvmps_gnn_params(self._h, None, v.ctypes.data_as(ctypes.c_void_p)))

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def forward(self, gb):
        t = self.torch
        out = t.empty((gb.n_nodes, self.shape[3]), dtype=t.float32, device=gb.features.device)
        check(lib().gvmps_gnn_forward(self._h, ctypes.byref(gb.struct), ctypes.c_void_p(out.data_ptr()),
                                      self._stream()))
        return out

    def loss_grad(self, gb, targets, want_grad=False):
        t = self.torch
        tt = t.as_tensor(targets, dtype=t.float32, device=gb.features.device).contiguous()
        loss = np.zeros(1, dtype=np.float32)
        grad = np.zeros(self.P, dtype=np.float32) if want_grad else None
        check(lib().gvmps_gnn_loss_grad(self._h, ctypes.byref(gb.struct), ctypes.c_void_p(tt.data_ptr()),
                                        loss.ctypes.data_as(ctypes.c_void_p),
                                        grad.ctypes.data_as(ctypes.c_void_p) if want_grad else None,
                                        self._stream()))
        return (float(loss[0]), grad) if want_grad else float(loss[0])

    def backward(self, gb, dy, want_grad=False):
        """Gradient of the parameters from dL/dY (n_nodes x outputs)."""
        t = self.torch
        d = t.as_tensor(dy, dtype=t.float32, device=gb.features.device).contiguous()
        grad = np.zeros(self.P, dtype=np.float32) if want_grad else None
        check(lib().gvmps_gnn_backward(self._h, ctypes.byref(gb.struct), ctypes.c_void_p(d.data_ptr()),
                                       grad.ctypes.data_as(ctypes.c_void_p) if want_grad else None,
                                       self._stream()))
        return grad

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib().gvmps_gnn_adam(self._h, float(lr), float(beta1), float(beta2), float(eps), self._stream()))

    def predict(self, n, words, coeffs):
        """Initial angles of an n-qubit Hamiltonian, global order layer * n + v."""
        gb = GraphBatch([(node_features(n, words, coeffs), edges(n, words, coeffs))], self.device)
        y = self.forward(gb).cpu().numpy().astype(np.float64)  # (n, layers)
        return y.T.reshape(-1)

    def save(self, path):
        np.savez(path, theta=self.params(), shape=np.array(self.shape))

    @classmethod
    def pretrained(cls, device=None):
        """The network trained by `python -m gvmps.learned_warmstart` (2000
        Hamiltonians of 20-40 qubits)."""
        import os
        net = cls(device=device)
        net.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', 'warmstart_gnn.npz'))
        return net

    def load(self, path):
        d = np.load(path, allow_pickle=False)
        if tuple(d['shape']) != self.shape:
            raise ValueError(f'network shape {tuple(d["shape"])} != {self.shape}')
        self.set_params(d['theta'])

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            from . import _lib
            if _lib._lib is not None:
                _lib._lib.gvmps_gnn_destroy(h)
            self._h = None
