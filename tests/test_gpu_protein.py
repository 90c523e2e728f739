"""The campaign's largest recorded validation on the GPU: the 97-qubit 3d7z
protein Hamiltonian (rebuilt by gvmps.protein, 3016 Z-strings), the 4-layer
RY+RZ ansatz (970 parameters), linear and reverse-linear, at
x ~ default_rng(534) U[-0.4, 0.4] (test/tools/check_largest_graph.py:24-68).

Bars, against test/VALIDATION.json global_97: analytic energies within 1e-10
relative; 1000-shot energies from the sampleSeed values 1234567 / 1234569
within 1e-9 relative (one changed bit would move them by far more); the
adjoint gradient against exact parameter shift at the three indices the
reference checks (1e-8 absolute), its value against the analytic energy."""
import json
from pathlib import Path

import numpy as np
import pytest

from gvmps import Batch, Problem, protein, seeds, workloads

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / 'golden' / 'validation_97.json').read_text())


@pytest.fixture(scope='module')
def h97(cuda):
    n, words, coeffs = protein.campaign_hamiltonian(GOLD['sequence'])
    assert n == 97 and len(words) == 3016
    return n, words, coeffs


@pytest.mark.parametrize('row', [0, 1], ids=['linear', 'reverse_linear'])
def test_97q_protein_golden_energies_samples_and_gradient(h97, row):
    n, words, coeffs = h97
    r = GOLD['rows'][row]
    ops, P = workloads.su2_ansatz(n, GOLD['reps'], tuple(GOLD['gates']), r['entanglement'])
    x = np.random.default_rng(534).uniform(-0.4, 0.4, P)
    p = Problem(n, ops, P, words, coeffs, max_bond=16, abs_cut=0.0, rel_cut=0.0)
    info = p.info()
    assert info['max_alloc_bond'] == 16 and info['grad_ok'] == 1
    print(f'case: 3d7z 97q {r["entanglement"]} energy + 1000 shots + adjoint', flush=True)
    b = Batch([p], [1])
    e = float(b.energies(x).cpu()[0])
    assert abs(e - r['energy']) <= 1e-10 * abs(r['energy']), (e, r['energy'])
    seed = seeds.shot_shard_seed(GOLD['sample_seeds'][r['entanglement']], 0)
    es = float(b.shot_energies(x, r['readout_shots'], [seed]).cpu()[0])
    assert abs(es - r['sampled_energy']) <= 1e-9 * abs(r['sampled_energy']), (es, r['sampled_energy'])
    ev, g = b.value_grad(x)
    assert abs(float(ev.cpu()[0]) - e) <= 1e-10 * abs(e)
    g = g.cpu().numpy()
    idx = [0, P // 2, P - 1]
    pts = []
    for i in idx:
        for sgn in (+1, -1):
            xs = x.copy()
            xs[i] += sgn * np.pi / 2
            pts.append(xs)
    es6 = Batch([p], [6]).energies(np.vstack(pts)).cpu().numpy()
    ps = (es6[0::2] - es6[1::2]) / 2
    np.testing.assert_allclose(g[idx], ps, atol=1e-8, rtol=0)
