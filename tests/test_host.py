"""Host-side checks that need no GPU: the C-ABI library loads and exports
every function include/gvmps.h declares, and the host planner (problem
compilation, forests, reverse-gradient programs) behaves as specified."""
import re
from pathlib import Path

import numpy as np
import pytest

from gvmps import Problem, workloads
from gvmps._lib import declared_symbols, lib

HEADER = Path(__file__).resolve().parents[1] / 'include' / 'gvmps.h'


def header_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r'^\s*(?:const\s+)?\w[\w\s\*]*?\b(gvmps_\w+)\s*\(', text, re.M)))


def test_library_exports_every_declared_symbol():
    L = lib()
    names = header_functions()
    assert 'gvmps_batch_value_grad' in names and 'gvmps_problem_create' in names
    for name in names:
        assert hasattr(L, name), name
    assert sorted(declared_symbols()) == names


def test_problem_info_of_workloads():
    for idx, (n, steps_min) in enumerate([(12, 1), (40, 1), (60, 1)]):
        w = workloads.config(idx)
        p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
        info = p.info()
        assert info['n_qubits'] == n
        assert info['n_params'] == w['nparams']
        assert info['max_alloc_bond'] == 16
        assert info['n_steps'] >= steps_min
        assert info['grad_ok'] == 1
        assert info['n_rotations'] == w['nparams']
        assert info['flops_grad_executed'] > info['flops_energy_executed'] > 0
        gens = p.generations()
        assert sum(c for c, _ in gens) == info['n_chains']
        assert sum(s for _, s in gens) == info['n_steps']


def test_reverse_gradient_support_rules():
    words, coeffs = ['ZIII'], [1.0]
    ok = Problem(4, [('ry', (0,), 0), ('cx', (1, 0), None), ('rz', (1,), 0)], 1, words, coeffs, max_bond=16)
    assert ok.info()['grad_ok'] == 1 and ok.info()['n_rotations'] == 2
    assert Problem(4, [('swap', (0, 1), None)], 0, words, coeffs).info()['grad_ok'] == 0
    assert Problem(4, [('cx', (0, 3), None)], 0, words, coeffs).info()['grad_ok'] == 0
    # five CNOTs on one bond would double it to 32 > 16
    many = [('cx', (1, 2), None)] * 5
    assert Problem(4, many, 0, words, coeffs, max_bond=16).info()['grad_ok'] == 0
    assert Problem(4, many[:4], 0, words, coeffs, max_bond=16).info()['grad_ok'] == 1


def test_real_arithmetic_rules():
    """real_arith: every gate matrix real (RY, H, X, Z, CX, SWAP; RX/RZ only as
    constant identities) and no Y in any word."""
    words, coeffs = ['ZXIZ', 'IIII'], [1.0, 0.5]
    real_ops = [('ry', (0,), 0), ('h', (1,), None), ('cx', (1, 0), None), ('swap', (2, 3), None),
                ('x', (2,), None), ('z', (3,), None), ('rz', (1,), 0.0)]
    assert Problem(4, real_ops, 1, words, coeffs).info()['real_arith'] == 1
    for extra in [('rx', (0,), 0), ('rz', (2,), 0), ('s', (1,), None), ('y', (0,), None), ('rx', (3,), 0.3)]:
        assert Problem(4, real_ops + [extra], 1, words, coeffs).info()['real_arith'] == 0, extra
    assert Problem(4, real_ops, 1, ['ZYII'], [1.0]).info()['real_arith'] == 0
    # iz_forest: no X or Y letter in any word (the I/Z chain kernels)
    assert Problem(4, real_ops, 1, words, coeffs).info()['iz_forest'] == 0
    assert Problem(4, real_ops, 1, ['ZIIZ', 'IZZI', 'IIII'], [1.0, 2.0, 0.5]).info()['iz_forest'] == 1
    assert Problem(4, real_ops, 1, ['ZIIZ', 'IIXI'], [1.0, 2.0]).info()['iz_forest'] == 0


def test_planner_rejects_bad_input():
    with pytest.raises(RuntimeError, match='>= 2 qubits'):
        Problem(1, [], 0, ['Z'], [1.0])
    with pytest.raises(RuntimeError, match='Invalid analytic Pauli'):
        Problem(2, [], 0, ['ZQ'], [1.0])
    with pytest.raises(RuntimeError, match='out of range'):
        Problem(2, [('ry', (0,), 3)], 1, ['ZZ'], [1.0])


def test_forest_dedup_counts():
    # words sharing prefixes share nodes: 3 words over 4 sites
    n = 4
    words = ['ZZII', 'ZZZI', 'IIIZ']
    p = Problem(n, [('ry', (q,), q) for q in range(n)], n, words, np.ones(3), max_bond=16)
    info = p.info()
    assert info['n_active_terms'] == 3
    assert info['n_steps'] <= 3 * n


def test_exact_split_rule():
    """exact_split: zero cutoffs and no two-site gate cut below
    min(2 Dl, 2 Dr) by max_bond (then every split keeps all of Theta)."""
    from gvmps import Problem, workloads
    for i in (0, 1, 2):
        w = workloads.config(i)
        p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
        assert p.info()['exact_split'] == 1
    ops, P = workloads.su2_ansatz(10, 3, ('ry', 'rz'), 'linear')
    words = ['Z' * 10]
    assert Problem(10, ops, P, words, [1.0], max_bond=64, abs_cut=0, rel_cut=0).info()['exact_split'] == 1
    # cutoffs compare singular values with a threshold at every gate
    assert Problem(10, ops, P, words, [1.0], max_bond=64, abs_cut=1e-5, rel_cut=0).info()['exact_split'] == 0
    assert Problem(10, ops, P, words, [1.0], max_bond=64, abs_cut=0, rel_cut=1e-5).info()['exact_split'] == 0
    # a bond cap below the circuit's allocation bond cuts rank (and arms the guard)
    assert Problem(10, ops, P, words, [1.0], max_bond=4, abs_cut=0, rel_cut=0).info()['exact_split'] == 0
    assert Problem(10, ops, P, words, [1.0], max_bond=8, abs_cut=0, rel_cut=0).info()['exact_split'] == 1
