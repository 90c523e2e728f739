"""Dense FP64 statevector reference, <= 24 qubits (test infrastructure).

Written from the textbook definitions, not from the reference: the state is
an n-axis complex128 tensor psi[b_0, ..., b_{n-1}] indexed by qubit.  It pins
the MPS restatement (mps_ref) and the reverse restatement (reverse_ref) the
same way the reference's own tests pin theirs against a Qiskit statevector
(test/tests/test_reverse_mps.py:28-36, 54-71).
"""
import numpy as np

from .mps_ref import FIXED, rotation
from . import rng

PAULI = {'I': np.eye(2, dtype=complex), 'X': np.array([[0, 1], [1, 0]], dtype=complex),
         'Y': np.array([[0, -1j], [1j, 0]], dtype=complex), 'Z': np.diag([1.0, -1.0]).astype(complex)}


class Dense:
    def __init__(self, n, initial=None):
        if n > 24:
            raise ValueError("dense reference limited to 24 qubits")
        self.n = n
        self.psi = np.zeros((2,) * n, dtype=complex)
        self.psi[(0,) * n] = 1.0
        if initial is not None:
            for q in range(n):
                self.apply1(np.asarray(initial, dtype=complex).reshape(2, 1) @ np.array([[1, 0]]), q)

    def apply1(self, U, q):
        self.psi = np.moveaxis(np.tensordot(U, self.psi, axes=([1], [q])), 0, q)

    def apply2(self, G4x4, a, b):
        """G indexed [(Ia*2+Ib), (ia*2+ib)] acting on qubits (a, b)."""
        G = np.asarray(G4x4, dtype=complex).reshape(2, 2, 2, 2)
        out = np.tensordot(G, self.psi, axes=([2, 3], [a, b]))
        self.psi = np.moveaxis(out, [0, 1], [a, b])

    def apply(self, name, qubits, angle=None):
        if name in ('rx', 'ry', 'rz'):
            self.apply1(rotation(name, angle), qubits[0])
        elif name in FIXED:
            self.apply1(FIXED[name], qubits[0])
        elif name == 'cx':
            CX = np.eye(4, dtype=complex)[[0, 1, 3, 2]]
            self.apply2(CX, qubits[0], qubits[1])
        elif name == 'swap':
            SW = np.eye(4, dtype=complex)[[0, 2, 1, 3]]
            self.apply2(SW, qubits[0], qubits[1])
        else:
            raise ValueError(name)

    def run(self, ops):
        for name, qubits, angle in ops:
            self.apply(name, qubits, angle)
        return self

    def expect(self, word):
        """<psi|P|psi>; word[q] is the Pauli on qubit q."""
        phi = self.psi
        for q, p in enumerate(word):
            if p != 'I':
                phi = np.moveaxis(np.tensordot(PAULI[p], phi, axes=([1], [q])), 0, q)
        return complex(np.vdot(self.psi, phi))

    def energy(self, words, coeffs):
        return sum(complex(c) * self.expect(w) for w, c in zip(words, coeffs)).real

    def z_diagonal_energy(self, words, coeffs):
        """Fast path for Z/I words: sum_x |psi_x|^2 sum_k c_k (-1)^{x.mask_k}."""
        prob = np.abs(self.psi.reshape(-1)) ** 2
        idx = np.arange(prob.size, dtype=np.int64)
        total = 0.0
        for w, c in zip(words, coeffs):
            mask = 0
            for q, p in enumerate(w):
                if p == 'Z':
                    mask |= 1 << (self.n - 1 - q)   # axis q is bit n-1-q of the flat index
                elif p != 'I':
                    raise ValueError("z_diagonal_energy needs Z/I words")
            par = np.zeros(prob.size, dtype=np.int64)
            m = idx & mask
            while m.any():
                par ^= m & 1
                m >>= 1
            total += complex(c).real * float(np.sum(prob * (1 - 2 * par)))
        return total

    def sample_sites(self, shots, seed=None, uniforms=None):
        """Sequential inverse-CDF sampling in MPS site order (qubit n-1 first).

        Same decision rule as the reference sampler: bit 0 iff u < p0/(p0+p1)
        (test/backend/source/sharded_mps.cuh:164-176), with u either the
        SplitMix64 draw for (seed, site, shot) or the given uniforms[shot, site].
        Returns bits[shot, site].
        """
        n = self.n
        prob = np.abs(self.psi) ** 2            # axes = qubits
        # reorder axes into site order: site s = qubit n-1-s
        prob = np.transpose(prob, list(range(n - 1, -1, -1)))
        bits = np.zeros((shots, n), dtype=np.int8)
        for sh in range(shots):
            cur = prob
            for k in range(n):
                marg = cur.reshape(2, -1).sum(axis=1)
                p0, p1 = marg
                total = p0 + p1
                if total <= 0:
                    p0, p1, total = 0.5, 0.5, 1.0
                u = rng.draw_uniform(seed, k, sh) if uniforms is None else uniforms[sh, k]
                b = 0 if u < p0 / total else 1
                bits[sh, k] = b
                cur = cur[b]
        return bits
