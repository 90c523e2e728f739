"""Host checks of the algorithmic work counts (gvmps.workcount).

The bond profile is pinned against the oracle MPS with zero cutoffs (which
keeps min(chi, m, n) columns, the allocation rule), the prefix-node counts
against a hand count on a four-qubit example, and the 40-qubit workload
against its recorded count (40,111 node-site transfers, 4.854 GFLOP)."""
import numpy as np
import pytest

from gvmps import workcount, workloads
from oracle import mps_ref


@pytest.mark.parametrize('seed', range(4))
def test_allocated_bonds_match_oracle_mps(seed):
    r = np.random.default_rng(seed)
    n = int(r.integers(4, 9))
    # rotations with a bond cap the circuit never reaches, or a product state
    # (every discarded singular value is zero, so the guard stays quiet)
    # with a cap that clamps
    rotations = seed % 2 == 0
    chi = 64 if rotations else int(r.choice([2, 4]))
    ops = []
    for _ in range(30):
        if rotations and r.random() < 0.4:
            ops.append(('ry', (int(r.integers(0, n)),), float(r.uniform(-3, 3))))
        else:
            a, b = r.choice(n, 2, replace=False)
            ops.append((str(r.choice(['cx', 'swap'])), (int(a), int(b)), None))
    bonds, shapes = workcount.allocated_bonds(n, ops, chi)
    sim = mps_ref.RefSimulator(n, max_bond=chi, abs_cut=0.0, rel_cut=0.0)
    sim.run(ops)
    assert bonds == sim.mps.bonds()
    assert len(shapes) == len(sim.mps.singular_log)
    assert [min(m, k) for m, k in shapes] == [len(s) for s in sim.mps.singular_log]


def test_transfer_flops_closed_form():
    f = workcount.transfer_flops([1, 2, 4, 2, 1])
    np.testing.assert_array_equal(f, [16 * (2 + 4), 16 * (16 + 32), 16 * (32 + 16), 16 * (4 + 2)])


def test_prefix_nodes_hand_count():
    """Site-order words IIIZ, IIZZ, IZZI, ZIII, IZIZ (qubit order reversed):
    spans [3,3], [2,3], [1,2], [0,0], [1,3]; distinct prefixes inside the
    spans: 1, 1, 3, 3 at sites 0..3.  The chunked DAG (chunks of 2 of the
    sorted words) transfers every site: 6 + 6 + 4."""
    words = ['ZIII', 'ZZII', 'IZZI', 'IIIZ', 'ZIZI', 'IIII']
    ops = [('cx', (0, 1), None), ('cx', (1, 2), None), ('cx', (2, 3), None)]
    bonds, _ = workcount.allocated_bonds(4, ops, 16)
    f = workcount.transfer_flops(bonds)
    r = workcount.energy_work(4, ops, words, chunk=2)
    assert r['node_site_transfers'] == 8
    assert r['flops_per_eval'] == pytest.approx(float(np.array([1, 1, 3, 3]) @ f + 2 * f.sum()))
    assert r['reference_dag_flops_per_eval'] == pytest.approx(float(np.array([3, 3, 5, 5]) @ f))


def test_forty_qubit_workload_count():
    w = workloads.config(1)
    r = workcount.energy_work(w['n'], w['ops'], w['words'])
    assert r['node_site_transfers'] == 40111
    assert r['flops_per_eval'] == pytest.approx(4.854111776e9, rel=1e-9)
    assert r['bonds'] == [min(2 ** s, 2 ** (40 - s), 16) for s in range(41)]
    _, shapes = workcount.allocated_bonds(w['n'], w['ops'])
    assert len(shapes) == 4 * 39
    assert workcount.svd_flops(shapes) == pytest.approx(25.58e6, rel=1e-3)


def test_sampling_and_svd_closed_forms():
    assert workcount.sampling_flops([1, 2, 1], 10) == pytest.approx(
        10 * 8 * (2 * 2 + 2 * 4) + 8 * (2 * 4 + 2 * 2) + 10 * 8 * (2 * 2 + 2 * 1) + 8 * (2 * 2 + 2 * 4))
    assert workcount.svd_flops([(4, 2)]) == 4 * (4 * 16 * 2 + 8 * 4 * 4 + 9 * 8)
