"""Protein-folding Hamiltonians of the campaign's jobs without qiskit.

The campaign builds each job's Hamiltonian with the Protein_Folding package
(test/tools/run_case.py:83-86): a peptide of the job's residue sequence with
no side chains, Miyazawa-Jernigan pair energies and penalty parameters
(10, 10, 10), compressed to the qubits it uses.  Every operator of that
construction is diagonal, a sum of Z-strings, so it is restated here on an
ordered list of (Z-string, coefficient) terms; a Z-string is an integer bit
mask, qubit q = bit q (Qiskit's order).  The steps follow the package the
campaign imports (test/sources/independent-vqe/, paths below relative to it)

  Protein_Folding/peptide/chains/base_chain.py:58-75    turn qubits (I + Z_k) / 2
  Protein_Folding/peptide/beads/main_bead.py:59-74      turn indicator functions
  Protein_Folding/bead_contacts/contact_map_builder.py  contact qubits (I - Z_c) / 2
  Protein_Folding/bead_distances/distance_map_builder.py:79-131  squared distances
  Protein_Folding/bead_distances/distance_map.py:48-133 first / second neighbour terms
  Protein_Folding/qubit_op_builder.py:53-170, 290-382   H_back + H_bbbb, tensor layout
  Protein_Folding/qubit_utils/qubit_fixing.py           _fix_qubits
  Protein_Folding/qubit_utils/qubit_number_reducer.py   unused qubits removed
  Protein_Folding/interactions/miyazawa_jernigan_interaction.py  pair energies

Two properties of that code shape the Hamiltonian, and with it the
campaign's recorded energies, so they are kept:
  * _fix_qubits writes False into rows 0, 1, 2, 3 and 5 of the operator's
    Z table: it clears the whole Z-string of the terms at those list
    positions (not those qubits' Z), and then rebuilds every term from a
    label written qubit 0 first, which Pauli() reads with qubit 0 last, so
    the operator's qubit order is reversed at every call;
  * which terms those are depends on the term order of every intermediate
    operator, so the algebra keeps SparsePauliOp's orders: sums concatenate,
    products and tensor products enumerate pairs with the left factor
    outermost, and simplify() merges equal strings in the order of their
    first appearance (adding in list order) and drops |c| <= 1e-8, leaving
    one zero identity term when nothing is left.
Side chains, chirality and the side-chain terms are absent without side
chains and are not built; a sequence with side chains is rejected.
"""
import os

import numpy as np

MJ_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', 'mj_matrix.txt')
ATOL = 1e-8            # SparsePauliOp.simplify: coefficients within 1e-8 of 0 are dropped
FIXED_ROWS = (0, 1, 2, 3, 5)   # _preset_binary_vals without a second side-chain bead


class Op:
    """Ordered sum of diagonal Pauli terms on n qubits."""
    __slots__ = ('n', 'm', 'c')

    def __init__(self, n, masks, coeffs):
        self.n, self.m, self.c = n, list(masks), [float(x) for x in coeffs]

    @staticmethod
    def identity(n, coeff=1.0):
        return Op(n, [0], [coeff])

    def _same(self, other):
        if self.n != other.n:
            raise ValueError(f'qubit counts differ: {self.n} and {other.n}')

    def __add__(self, other):
        if isinstance(other, (int, float)):
            if other == 0:
                return self
            raise TypeError('only 0 adds to an operator')
        self._same(other)
        return Op(self.n, self.m + other.m, self.c + other.c)

    __radd__ = __add__

    def __neg__(self):
        return Op(self.n, self.m, [-x for x in self.c])

    def __sub__(self, other):
        return self + (-other)

    def __rsub__(self, other):
        return (-self) + other

    def __mul__(self, s):
        s = float(s)
        return Op(self.n, self.m, [s * x for x in self.c])

    __rmul__ = __mul__

    def __truediv__(self, s):
        return Op(self.n, self.m, [x / float(s) for x in self.c])

    def __matmul__(self, other):
        """Product, every pair (left term outer): Z^a Z^b = Z^(a xor b)."""
        self._same(other)
        return Op(self.n, [a ^ b for a in self.m for b in other.m], [x * y for x in self.c for y in other.c])

    def __xor__(self, other):
        """Tensor product self (x) other, `other` on the low qubits, left term outer."""
        sh = other.n
        return Op(self.n + other.n, [(a << sh) | b for a in self.m for b in other.m],
                  [x * y for x in self.c for y in other.c])

    def simplify(self):
        acc = {}
        for m, c in zip(self.m, self.c):
            acc[m] = acc.get(m, 0.0) + c
        keep = [(m, c) for m, c in acc.items() if abs(c) > ATOL]
        if not keep:
            return Op(self.n, [0], [0.0])
        return Op(self.n, [m for m, _ in keep], [c for _, c in keep])

    def reversed_qubits(self):
        n = self.n
        out = []
        for m in self.m:
            r = 0
            for q in range(n):
                if (m >> q) & 1:
                    r |= 1 << (n - 1 - q)
            out.append(r)
        return Op(n, out, self.c)


def fix_qubits(op):
    """qubit_fixing.py:13-55 as it executes on a SparsePauliOp: the Z-strings
    of the terms at list positions 0, 1, 2, 3 and 5 are cleared, the qubit
    order is reversed, and the result is simplified.  Integers pass through."""
    if isinstance(op, (int, float)):
        return op
    m = list(op.m)
    for row in FIXED_ROWS:
        if row < len(m):
            m[row] = 0
    return Op(op.n, m, op.c).reversed_qubits().simplify()


def mj_pair_energies(sequence):
    """pair_energies[i, 0, j, 0] (1 <= i < j <= N) from the Miyazawa-Jernigan
    table (Miyazawa & Jernigan, J. Mol. Biol. 256, 623-644 (1996), Table 3):
    the upper triangle of the 20 x 20 residue table, symbols in its header
    (energy_matrix_loader.py:10-44, miyazawa_jernigan_interaction.py:21-47)."""
    with open(MJ_PATH) as f:
        rows = [line.split() for line in f if line.strip()]
    symbols = rows[0]
    table = np.zeros((len(symbols), len(symbols)))
    for r, row in enumerate(rows[1:]):
        for c in range(r, len(symbols)):
            table[r, c] = float(row[c])
    for ch in sequence:
        if ch not in symbols:
            raise ValueError(f'unknown residue {ch!r}')
    n = len(sequence)
    pe = np.zeros((n + 1, 2, n + 1, 2))
    for i in range(1, n + 1):
        for j in range(i + 1, n + 1):
            a, b = symbols.index(sequence[i - 1]), symbols.index(sequence[j - 1])
            pe[i, 0, j, 0] = table[min(a, b), max(a, b)]
    return pe


class ProteinHamiltonian:
    """The campaign's protein-folding Hamiltonian of a residue sequence (main
    chain only; penalty_back and penalty_1 as PenaltyParameters)."""

    def __init__(self, sequence, side_chains=None, penalty_back=10.0, penalty_1=10.0):
        if side_chains is not None and any(side_chains):
            raise ValueError('side chains are not supported (the campaign builds main chains only)')
        if len(sequence) < 3:
            raise ValueError('a main chain of at least 3 residues is required')
        self.seq = sequence
        self.N = N = len(sequence)
        self.back, self.l1 = float(penalty_back), float(penalty_1)
        self.nturn = 2 * (N - 1)          # turn qubits of the main chain
        self.nconf = 2 * self.nturn       # indicator width: turn register + an identity register
        self.ncontact = (N - 1) ** 2      # one block of contact qubits
        self.pe = mj_pair_energies(sequence)
        self.ind = [self._indicators(b) for b in range(N - 1)]
        self.dist = {}

    def _indicators(self, bead):
        """Turn indicators of main bead `bead` (main_bead.py:59-74): the
        identity composed with q or (I - q) for the turn vectors 00, 01, 10,
        11, q = (I + Z_k) / 2 (base_chain.py:58-75), tensored below an
        identity register and simplified."""
        n = self.nturn
        one = Op.identity(n)
        qs = [Op(n, [0, 1 << (2 * bead)], [0.5, 0.5]), Op(n, [0, 1 << (2 * bead + 1)], [0.5, 0.5])]
        out = []
        for vec in ((0, 0), (0, 1), (1, 0), (1, 1)):
            op = one
            for q, v in zip(qs, vec):
                op = op @ (q if v else one - q)
            out.append((Op.identity(n) ^ op).simplify())
        return out

    def turn_operator(self, lower, upper):
        """qubit_op_builder.py:117-150."""
        s = 0
        for a in range(4):
            s = s + (self.ind[lower][a] @ self.ind[upper][a])
        return fix_qubits(s)

    def distance(self, lower, upper):
        """Squared distance of main beads lower < upper (1-based): per axis
        d_a = fix(sum_{k=lower}^{upper-1} (-1)^k indicator_a(bead k)), then
        fix(sum_a d_a @ d_a) (distance_map_builder.py:79-131, 105-131)."""
        key = (lower, upper)
        if key not in self.dist:
            total = 0
            for a in range(4):
                d = 0
                for k in range(lower, upper):
                    d = d + ((-1) ** k) * self.ind[k - 1][a]
                d = fix_qubits(d)
                total = total + (d @ d)
            self.dist[key] = fix_qubits(total)
        return self.dist[key]

    def first_neighbor(self, i, j):
        """distance_map.py:48-92: fix(lambda_0 (x - I) + 0.1 e_ij I),
        lambda_0 = 7 (j - i + 1) lambda_1."""
        x = self.distance(i, j)
        lam0 = 7 * (j - i + 1) * self.l1
        e = self.pe[i][0][j][0]
        one = Op.identity(x.n)
        return fix_qubits(lam0 * (x - one) + (0.1 * e) * one)

    def second_neighbor(self, i, j):
        """distance_map.py:94-133: fix(lambda_1 (2 I - x) + 0.1 e_ij I); None
        where the reference's lookup raises and the term is skipped
        (qubit_op_builder.py:312-359): i = 0 wraps to the last bead, which
        has no distances, and j = N + 1 is past the pair-energy table."""
        if i < 1 or j > self.N:
            return None
        x = self.distance(i, j)
        e = self.pe[i][0][j][0]
        one = Op.identity(x.n)
        return fix_qubits(self.l1 * ((2 * one) - x) + (0.1 * e) * one)

    def contact(self, i, j):
        """Main-main contact of beads (i, j): (I - Z_c) / 2 with Z_c in the
        highest of four (N-1)^2-qubit blocks (contact_map_builder.py:170-200,
        235-240)."""
        nb = self.ncontact
        idx = (i - 1) * (self.N - 1) + (j - 1)
        return Op(4 * nb, [0, 1 << (3 * nb + idx)], [0.5, -0.5])

    def h_back(self):
        """qubit_op_builder.py:152-170: turns along one axis in a row."""
        h = 0
        for b in range(self.N - 2):
            h = h + self.back * self.turn_operator(b, b + 1)
        return fix_qubits(h)

    def h_bbbb(self):
        """qubit_op_builder.py:290-382: first and second neighbour terms of
        every main-main contact (j - i odd, j - i >= 5), each tensored below
        its contact qubit; _fix_qubits after every contact."""
        h = 0
        N = self.N
        for i in range(1, N - 3):
            for j in range(i + 5, N + 1):
                if (j - i) % 2 == 0:
                    continue
                c = self.contact(i, j)
                h = h + (c ^ self.first_neighbor(i, j))
                for lo, up in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)):
                    s = self.second_neighbor(lo, up)
                    if s is not None:
                        h = h + (c ^ s)
                h = fix_qubits(h)
        return h

    def full(self):
        """H_back (below the contact register) + H_bbbb, simplified
        (qubit_op_builder.py:53-115 without side chains)."""
        h = Op.identity(4 * self.ncontact) ^ self.h_back()
        return (h + self.h_bbbb()).simplify()

    def qubit_op(self):
        """(labels, coefficients, unused qubits): the Hamiltonian on the qubits
        its terms use (qubit_number_reducer.py:12-110), Qiskit labels
        (leftmost character = highest qubit), in term order."""
        h = self.full()
        used = 0
        for m in h.m:
            used |= m
        keep = [q for q in range(h.n) if (used >> q) & 1]
        unused = [q for q in range(h.n) if not (used >> q) & 1]
        masks = []
        for m in h.m:
            nm = 0
            for k, q in enumerate(keep):
                if (m >> q) & 1:
                    nm |= 1 << k
            masks.append(nm)
        red = Op(len(keep), masks, h.c).simplify()
        n = red.n
        labels = [''.join('Z' if (m >> (n - 1 - k)) & 1 else 'I' for k in range(n)) for m in red.m]
        return labels, np.array(red.c), unused


def campaign_hamiltonian(sequence):
    """(n_qubits, qubit-order words, coefficients) of a campaign job's
    Hamiltonian: word[q] acts on qubit q (Qiskit qubit q, which
    qiskit_sparsepauliop_to_cudaq_spin maps to CUDA-Q qubit q)."""
    labels, coeffs, _ = ProteinHamiltonian(sequence).qubit_op()
    n = len(labels[0])
    return n, [lab[::-1] for lab in labels], coeffs


def job_hamiltonian(job):
    """Hamiltonian of a campaign job record (test/jobs/**/job.json: sequence,
    simulation_qubits, ...): (n, words, coefficients), checked against the
    job's qubit count as run_case.py:83-85 does."""
    n, words, coeffs = campaign_hamiltonian(job['sequence'])
    if 'simulation_qubits' in job and n != int(job['simulation_qubits']):
        raise ValueError('Protein qubit count differs from job')
    return n, words, coeffs
