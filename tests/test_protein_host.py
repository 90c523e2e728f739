"""Host checks of the protein-folding Hamiltonian builder (gvmps.protein).

Pinned by the reference's own records, which the builder was not fitted to:
  * the compressed qubit count of every residue length in the campaign
    catalog (test/protein_catalog.json, 5-14 residues, 7-97 qubits);
  * the 97-qubit campaign Hamiltonian (3d7z) through the oracle MPS: the
    analytic energies of test/VALIDATION.json global_97 for the linear and
    reverse-linear 4-layer RY+RZ ansatz at x ~ default_rng(534) U[-0.4, 0.4]
    (1e-10 relative), and its 1000-shot energies with the sampleSeed values
    1234567 and 1234569, exactly;
and by hand cases of _fix_qubits' behaviour (terms 0, 1, 2, 3, 5 cleared,
qubit order reversed) and of SparsePauliOp's term order."""
import json
from pathlib import Path

import numpy as np
import pytest

from gvmps import protein, workloads
from oracle import mps_ref

GOLD = Path(__file__).parent / 'golden'


def test_fix_qubits_clears_rows_and_reverses():
    op = protein.Op(4, [0b0001, 0b0010, 0b0100, 0b1000, 0b0011, 0b0110, 0b1100], [1, 2, 3, 4, 5, 6, 7])
    out = protein.fix_qubits(op)
    # rows 0,1,2,3,5 become identity (merged, first position), row 4 (0011)
    # and row 6 (1100) keep their strings, reversed over 4 qubits
    assert out.m == [0, 0b1100, 0b0011] and out.c == [1 + 2 + 3 + 4 + 6, 5, 7]
    assert protein.fix_qubits(0) == 0


def test_simplify_order_and_tolerance():
    op = protein.Op(3, [4, 1, 4, 2, 0], [1.0, 5e-9, 2.0, -1e-3, 0.5])
    s = op.simplify()
    assert s.m == [4, 2, 0] and s.c == [3.0, -1e-3, 0.5]
    z = protein.Op(2, [1, 1], [1.0, -1.0]).simplify()
    assert z.m == [0] and z.c == [0.0]
    # products and tensor products enumerate pairs left term outermost
    a, b = protein.Op(2, [0, 1], [1, 2]), protein.Op(2, [0, 2], [3, 4])
    assert (a @ b).m == [0, 2, 1, 3] and (a @ b).c == [3, 4, 6, 8]
    assert (a ^ b).m == [0, 2, 4, 6] and (a ^ b).n == 4


def test_catalog_qubit_counts():
    cat = json.loads((GOLD / 'protein_catalog.json').read_text())['proteins']
    by_len = {}
    for j in cat:
        by_len.setdefault(j['residues'], j)
    assert sorted(by_len) == list(range(5, 15))
    for length, j in sorted(by_len.items()):
        labels, coeffs, _ = protein.ProteinHamiltonian(j['sequence']).qubit_op()
        assert len(labels[0]) == j['simulation_qubits'], (j['protein_id'], len(labels[0]))
        assert np.all(np.isfinite(coeffs)) and set(''.join(labels)) <= {'I', 'Z'}
    with pytest.raises(ValueError, match='Protein qubit count differs from job'):
        protein.job_hamiltonian(dict(sequence=cat[0]['sequence'], simulation_qubits=cat[0]['simulation_qubits'] + 1))


@pytest.fixture(scope='module')
def h97():
    g = json.loads((GOLD / 'validation_97.json').read_text())
    n, words, coeffs = protein.campaign_hamiltonian(g['sequence'])
    assert n == g['qubits'] == 97
    return g, words, coeffs


@pytest.mark.parametrize('row', [0, 1], ids=['linear', 'reverse_linear'])
def test_97q_golden_energies_through_the_oracle(h97, row):
    g, words, coeffs = h97
    r = g['rows'][row]
    ops, P = workloads.su2_ansatz(97, g['reps'], tuple(g['gates']), r['entanglement'])
    assert P == g['params']
    x = np.random.default_rng(534).uniform(-0.4, 0.4, P)
    tape = [(gt, q, (x[p] if p is not None else None)) for gt, q, p in ops]
    sim = mps_ref.RefSimulator(97, max_bond=16, abs_cut=0.0, rel_cut=0.0).run(tape)
    e, _ = sim.observe(words, coeffs)
    assert abs(e - r['energy']) <= 1e-10 * abs(r['energy']), (e, r['energy'])
    sim.sample_seed = g['sample_seeds'][r['entanglement']]
    es, _ = sim.observe(words, coeffs, shots=r['readout_shots'])
    assert es == r['sampled_energy']
