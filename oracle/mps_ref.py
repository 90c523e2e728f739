"""Explicit-SVD MPS restatement of the reference backend (test infrastructure).

Follows test/backend/source/sharded_mps.cuh, dist_svd.cuh, analytic_batch.cuh
and the single-GPU branches of simulator_mps_dist.cu.  complex128 throughout.

Conventions reproduced (each cited where it is used):
  * site s = qubit n-1-s; Pauli words are reversed into site order;
  * site tensors are (Dl, 2, Dr) arrays, |0> = n sites of shape (1,2,1) = [1,0];
  * a two-qubit gate is a row-major 4x4 [(I*2+J)*4+(i*2+j)], first operand =
    left site; controlled-U has the control first; non-adjacent pairs are
    SWAP-routed;
  * each two-site contraction is split by an SVD with the reference rank rule
    and zero-cutoff guard: left = U_k, right = diag(s_k) Vh_k.

The SVD is LAPACK's (numpy.linalg.svd, gesdd); the reference calls cuSOLVER's
Zgesvd (dist_svd.cuh:255-259).  Both return the same singular values up to
rounding; the singular-vector gauge differs and no gauge-invariant output
(energies, probabilities, sampled bits away from CDF boundaries) depends on it.
"""
import numpy as np

from . import rng

EPS = 2.2204460492503131e-16


# ---------------------------------------------------------------- gate matrices
def rotation(name, theta):
    """RX/RY/RZ matrices, c = cos(theta/2), s = sin(theta/2).

    test/tools/reverse_mps.py:52-55 (the same CUDA-Q conventions the backend
    receives as row-major 2x2 task matrices).
    """
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    if name == 'rx':
        return np.array([[c, -1j * s], [-1j * s, c]], dtype=complex)
    if name == 'ry':
        return np.array([[c, -s], [s, c]], dtype=complex)
    if name == 'rz':
        return np.array([[np.exp(-0.5j * theta), 0], [0, np.exp(0.5j * theta)]], dtype=complex)
    raise ValueError(name)


FIXED = {
    'h': np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2),
    'x': np.array([[0, 1], [1, 0]], dtype=complex),
    'y': np.array([[0, -1j], [1j, 0]], dtype=complex),
    'z': np.array([[1, 0], [0, -1]], dtype=complex),
    's': np.array([[1, 0], [0, 1j]], dtype=complex),
    'sdg': np.array([[1, 0], [0, -1j]], dtype=complex),
}  # test/tools/reverse_mps.py:155-157


def controlled_gate(U):
    """controlled-U, control first.  simulator_mps_dist.cu:374-384."""
    G = np.zeros(16, dtype=complex)
    for C in range(2):
        for T in range(2):
            for c in range(2):
                for t in range(2):
                    if C != c:
                        continue
                    G[(C * 2 + T) * 4 + (c * 2 + t)] = (1.0 if T == t else 0.0) if c == 0 else U[T, t]
    return G


def swap_mat():
    """sharded_mps.cuh:551-556."""
    S = np.zeros(16, dtype=complex)
    for I in range(2):
        for J in range(2):
            for i in range(2):
                for j in range(2):
                    S[(I * 2 + J) * 4 + (i * 2 + j)] = 1.0 if (I == j and J == i) else 0.0
    return S


def swap_operands(G):
    """sharded_mps.cuh:557-562."""
    H = np.zeros(16, dtype=complex)
    for I in range(2):
        for J in range(2):
            for i in range(2):
                for j in range(2):
                    H[(I * 2 + J) * 4 + (i * 2 + j)] = G[(J * 2 + I) * 4 + (j * 2 + i)]
    return H


# ------------------------------------------------------------------------ SVD
class BondGuardError(RuntimeError):
    pass


def select_rank(singular, max_bond, abs_cut, rel_cut):
    """selectDirectSvdRank.  dist_svd.cuh:194-212.  Returns k, -1 or -2."""
    count = len(singular)
    threshold = max(abs_cut, rel_cut * singular[0])
    selected = 0
    for i in range(count):
        if not np.isfinite(singular[i]) or singular[i] < 0.0:
            return -1
        if (abs_cut == 0.0 and rel_cut == 0.0) or singular[i] > threshold:
            selected += 1
    if (abs_cut == 0.0 and rel_cut == 0.0 and max_bond < count
            and singular[max_bond] > 64.0 * EPS * count * singular[0]):
        return -2
    return max(1, min(max_bond, selected))


def svd_factors(M, max_bond, abs_cut, rel_cut):
    """singleGpuSvdFactors + packDirectSvdFactors.  dist_svd.cuh:214-283.

    M is (m, n).  Returns (left (m,k) = U_k, right (k,n) = diag(s_k) Vh_k, s).
    """
    U, s, Vh = np.linalg.svd(M, full_matrices=False)
    k = select_rank(s, max_bond, abs_cut, rel_cut)
    if k == -2:
        raise BondGuardError(
            f"CUDAQ_MPS_MAX_BOND={max_bond} is insufficient for zero-cutoff simulation: "
            "resolved singular values would be discarded")
    if k < 1:
        raise RuntimeError("Direct MPS SVD failed or produced nonfinite singular values; devInfo=0")
    return U[:, :k], s[:k, None] * Vh[:k, :], s


# ------------------------------------------------------------------------ MPS
class RefMPS:
    """ShardedMPS restricted to one device.  sharded_mps.cuh:309-1043."""

    def __init__(self, n, max_bond=64, abs_cut=1e-5, rel_cut=1e-5):
        # ctor / resetZero: sharded_mps.cuh:340-353, 440-447
        self.n = n
        self.max_bond, self.abs_cut, self.rel_cut = max_bond, abs_cut, rel_cut
        self.sites = [np.array([1.0, 0.0], dtype=complex).reshape(1, 2, 1) for _ in range(n)]
        self.singular_log = []

    def bonds(self):
        return [1] + [a.shape[2] for a in self.sites]

    def apply1q(self, U, s):
        """k_apply1qInPlace: A[l,:,r] <- U A[l,:,r].  sharded_mps.cuh:16-26, 465-474."""
        self.sites[s] = np.einsum('ij,ajb->aib', U, self.sites[s])

    def build_theta(self, G4, s):
        """k_buildM: Theta[(l,I),(J,r)].  sharded_mps.cuh:30-45 (row-major here)."""
        A, B = self.sites[s], self.sites[s + 1]
        Dl, Dr = A.shape[0], B.shape[2]
        G = np.asarray(G4, dtype=complex).reshape(2, 2, 2, 2)   # [I,J,i,j]
        AB = np.einsum('aic,cjb->aijb', A, B)
        T = np.einsum('IJij,aijb->aIJb', G, AB)
        return T.reshape(2 * Dl, 2 * Dr)

    def apply2q(self, G4, s):
        """apply2q on adjacent sites (s, s+1).  sharded_mps.cuh:477-516."""
        Dl, Dr = self.sites[s].shape[0], self.sites[s + 1].shape[2]
        M = self.build_theta(G4, s)
        left, right, sv = svd_factors(M, self.max_bond, self.abs_cut, self.rel_cut)
        self.singular_log.append(sv)
        k = left.shape[1]
        self.sites[s] = left.reshape(Dl, 2, k)
        self.sites[s + 1] = right.reshape(k, 2, Dr)

    def apply_two_qubit(self, G4, a, b):
        """SWAP-routed arbitrary pair.  sharded_mps.cuh:565-572."""
        if a == b:
            raise RuntimeError("2q gate on same qubit")
        if a > b:
            return self.apply_two_qubit(swap_operands(G4), b, a)
        SW = swap_mat()
        for x in range(b, a + 1, -1):
            self.apply2q(SW, x - 1)
        self.apply2q(G4, a)
        for x in range(a + 1, b):
            self.apply2q(SW, x)

    # ------------------------------------------------------------- expectations
    @staticmethod
    def pauli2(p):
        """P[i][ip] (bra i, ket ip).  sharded_mps.cuh:255-264, 767-772."""
        m = {'I': [[1, 0], [0, 1]], 'X': [[0, 1], [1, 0]], 'Y': [[0, -1j], [1j, 0]],
             'Z': [[1, 0], [0, -1]]}
        return np.array(m.get(p, m['I']), dtype=complex)

    def expect_pauli_host(self, word):
        """expectPauliHost: left-to-right transfer.  sharded_mps.cuh:765-791."""
        E = np.ones((1, 1), dtype=complex)
        for s, A in enumerate(self.sites):
            P = self.pauli2(word[s] if word else 'I')
            # newE[b,bp] = sum conj(A[a,i,b]) P[i,ip] A[ap,ip,bp] E[a,ap]
            E = np.einsum('aib,ij,cjd,ac->bd', A.conj(), P, A, E)
        return E[0, 0]

    def expect_pauli_shared(self, words, chunk=None):
        """expectPauliShared: sorted words, bounded chunks, prefix DAG per chunk.

        sharded_mps.cuh:614-687 with the node kernels analyticLeft/Right/Gather
        (analytic_batch.cuh:4-36): X/Y flip the ket index, Z gives -1 on i=1,
        Y gives -i on i=0 and +i on i=1.
        """
        n = self.n
        if not words:
            return []
        bond = max(max(a.shape[0], a.shape[2]) for a in self.sites)
        stride = bond * bond
        if chunk is None:
            budget = 64 * 1024 * 1024
            chunk = max(1, min(128, budget // (4 * 16 * stride)))
            chunk = min(chunk, len(words))
        order = sorted(range(len(words)), key=lambda k: words[k])
        result = [0j] * len(words)
        for start in range(0, len(order), chunk):
            sel = order[start:start + chunk]
            envs = np.ones((1, 1, 1), dtype=complex)
            previous = [0] * len(sel)
            for q in range(n):
                nodes, parents, paulis, current = {}, [], [], []
                for j, k in enumerate(sel):
                    word = words[k]
                    if len(word) != n:
                        raise RuntimeError("Invalid analytic Pauli width")
                    p = word[q]
                    if p not in 'IXYZ':
                        raise RuntimeError("Invalid analytic Pauli")
                    key = (previous[j], p)
                    if key not in nodes:
                        nodes[key] = len(nodes)
                        parents.append(previous[j])
                        paulis.append(p)
                    current.append(nodes[key])
                A = self.sites[q]
                new = []
                for parent, p in zip(parents, paulis):
                    new.append(self.transfer(envs[parent], A, p))
                envs = np.stack(new)
                previous = current
            for j, k in enumerate(sel):
                result[k] = envs[previous[j]][0, 0]
        return result

    @staticmethod
    def transfer(E, A, p):
        """E' = sum_i phase(p,i) A_i^H E A_{flip(p,i)}.  analytic_batch.cuh:10-30."""
        out = np.zeros((A.shape[2], A.shape[2]), dtype=complex)
        for i in range(2):
            ip = 1 - i if p in 'XY' else i
            phase = 1.0
            if p == 'Z' and i:
                phase = -1.0
            if p == 'Y':
                phase = 1j if i else -1j
            S = phase * (E @ A[:, ip, :])
            out += A[:, i, :].conj().T @ S
        return out

    # ------------------------------------------------------------------ sampling
    def right_envs(self):
        """prepareSamplingRightEnvironments: R[k][a,a'].  sharded_mps.cuh:61-90, 814-836."""
        R = [None] * (self.n + 1)
        R[self.n] = np.ones((1, 1), dtype=complex)
        for k in range(self.n - 1, -1, -1):
            A = self.sites[k]
            T = np.einsum('aib,bc->aic', A, R[k + 1])           # k_rightStage
            R[k] = np.einsum('aic,dic->ad', T, A.conj())         # k_rightFinish
        return R

    def sample_bits(self, shots, seed, uniforms=None, margins=False):
        """Perfect sampling with the GPU uniform stream.

        sampleBitsFromPreparedRight / k_sampleW / k_sampleProbabilities /
        k_chooseSample (sharded_mps.cuh:103-188, 838-877).  Returns an int8
        array bits[shot, site] (site order).  `uniforms`, if given, is a
        (shots, n) array used instead of the SplitMix64 stream.
        """
        n = self.n
        R = self.right_envs()
        bits = np.zeros((shots, n), dtype=np.int8)
        margin = np.zeros((shots, n))
        v = np.ones((shots, 1), dtype=complex)
        for k in range(n):
            A = self.sites[k]
            w = np.einsum('sa,aib->sib', v, A)                     # k_sampleW
            Rn = R[k + 1]
            # p[shot,i] = max(0, Re sum_{b,bp} w_b R[b,bp] conj(w_bp))
            p = np.einsum('sib,bc,sic->si', w, Rn, w.conj()).real
            p = np.maximum(p, 0.0)
            nxt = np.empty((shots, A.shape[2]), dtype=complex)
            for sh in range(shots):
                p0, p1 = p[sh, 0], p[sh, 1]
                total = p0 + p1
                if total <= 0.0:
                    p0 = p1 = 0.5
                    total = 1.0
                u = rng.draw_uniform(seed, k, sh) if uniforms is None else uniforms[sh, k]
                sel = 0 if u < p0 / total else 1
                margin[sh, k] = abs(u - p0 / total)
                ps = p1 if sel else p0
                inv = 1.0 / np.sqrt(ps) if ps > 0.0 else 1.0
                bits[sh, k] = sel
                nxt[sh] = w[sh, sel] * inv
            v = nxt
        return (bits, margin) if margins else bits

    def sample_strings(self, shots, seed):
        """sampleBitsGpu: qubit-order strings (q0 leftmost).  sharded_mps.cuh:887-902."""
        bits = self.sample_bits(shots, seed)
        return [''.join('1' if bits[sh, self.n - 1 - q] else '0' for q in range(self.n))
                for sh in range(shots)]

    def sample_z_expectations(self, words, shots, seed):
        """sampleZExpectationsGpu + k_zSampleSums.  sharded_mps.cuh:193-214, 906-935."""
        bits = self.sample_bits(shots, seed)
        out = []
        for w in words:
            mask = np.array([c == 'Z' for c in w], dtype=bool)
            parity = (bits[:, mask].sum(axis=1) & 1)
            out.append(float(np.sum(1 - 2 * parity)) / shots)
        return out


# ----------------------------------------------------------- simulator surface
class RefSimulator:
    """Single-GPU semantics of SimulatorMPSDist (simulator_mps_dist.cu).

    Gates arrive in qubit indices and are routed to site n-1-q (:386-416);
    observe reverses each word into site order (:571-583), returns 1 for
    identity words (:1024-1037), the analytic path uses the shared prefix DAG
    (:641-660), the finite-shot diagonal path draws one shared sample set with
    seed hbatchSeedForShotShard(sampleSeed++, 0) (:1157-1171, 610-619), and
    sample() draws with the same seed rule (:1587-1612).
    """

    def __init__(self, n, max_bond=64, abs_cut=1e-5, rel_cut=1e-5):
        if n < 2:
            raise RuntimeError("tensornet-mps needs >= 2 qubits")  # :235-237
        self.n = n
        self.mps = RefMPS(n, max_bond, abs_cut, rel_cut)
        self.sample_seed = rng.INITIAL_SAMPLE_SEED

    def site(self, q):
        return self.n - 1 - q

    def apply(self, name, qubits, angle=None):
        """applyGateToState.  simulator_mps_dist.cu:392-416."""
        if name in ('rx', 'ry', 'rz'):
            self.mps.apply1q(rotation(name, angle), self.site(qubits[0]))
        elif name in FIXED:
            self.mps.apply1q(FIXED[name], self.site(qubits[0]))
        elif name == 'cx':
            c, t = qubits
            self.mps.apply_two_qubit(controlled_gate(FIXED['x']), self.site(c), self.site(t))
        elif name == 'swap':
            a, b = qubits
            self.mps.apply_two_qubit(swap_mat(), self.site(a), self.site(b))
        else:
            raise RuntimeError(f"unsupported gate {name}")

    def run(self, ops):
        for name, qubits, angle in ops:
            self.apply(name, qubits, angle)
        return self

    def site_words(self, words):
        return [w[::-1] for w in words]

    def observe_terms(self, words, shots=-1):
        """Per-term <P_k> (qubit-order words in, as CUDA-Q get_pauli_word)."""
        sw = self.site_words(words)
        values = [1.0 + 0j] * len(sw)
        active = [k for k, w in enumerate(sw) if any(c != 'I' for c in w)]
        if shots is None or shots < 0:
            if active:
                got = self.mps.expect_pauli_shared([sw[k] for k in active])
                for k, v in zip(active, got):
                    values[k] = v
            return values
        evaluation_seed = self.sample_seed
        self.sample_seed += 1
        if not all(set(w) <= {'I', 'Z'} for w in sw):
            raise NotImplementedError("mixed-basis finite-shot words")
        seed = rng.shot_shard_seed(evaluation_seed, 0)
        got = self.mps.sample_z_expectations(sw, shots, seed)
        return [complex(g) for g in got]

    def observe(self, words, coeffs, shots=-1):
        """Energy = Re sum_k c_k <P_k>.  simulator_mps_dist.cu:1467-1481."""
        ev = self.observe_terms(words, shots)
        total = sum(complex(c) * v for c, v in zip(coeffs, ev))
        return total.real, [v.real for v in ev]

    def sample(self, shots=1000, qubits=None):
        """counts keyed by the requested qubits.  simulator_mps_dist.cu:1587-1612."""
        shots = shots if shots > 0 else 1000
        evaluation_seed = self.sample_seed
        self.sample_seed += 1
        strings = self.mps.sample_strings(shots, rng.shot_shard_seed(evaluation_seed, 0))
        qubits = list(range(self.n)) if qubits is None else qubits
        counts = {}
        for full in strings:
            key = ''.join(full[q] for q in qubits)
            counts[key] = counts.get(key, 0) + 1
        return counts
