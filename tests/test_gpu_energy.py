"""GPU parity of the analytic energy path (state preparation + prefix chains)
against the CPU oracle and the dense statevector.  Tolerance: energies within
1e-10 relative (absolute 1e-10 for |E| < 1), per-term <P_k> within 1e-11."""
import numpy as np
import pytest

from gvmps import Batch, Problem, workloads
from oracle import dense, mps_ref

pytestmark = pytest.mark.gpu


def rel_close(a, b, rtol=1e-10):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return np.all(np.abs(a - b) <= rtol * np.maximum(1.0, np.abs(b)))


def tape_with_values(ops, x):
    return [(g, q, (x[p] if isinstance(p, (int, np.integer)) else p)) for g, q, p in ops]


def oracle_terms(n, ops, x, words, coeffs, max_bond=16, cut=0.0):
    sim = mps_ref.RefSimulator(n, max_bond=max_bond, abs_cut=cut, rel_cut=cut).run(tape_with_values(ops, x))
    e, _ = sim.observe(words, coeffs)
    return e, np.array(sim.observe_terms(words))


def test_plain_12q_config_energy_and_terms(cuda):
    w = workloads.config(0)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    b = Batch([p], [1])
    print('case: 12q plain config', flush=True)
    e, tv = b.energies(w['x'], terms=True)
    e = e.cpu().numpy()
    tv = tv.cpu().numpy()
    eo, to = oracle_terms(w['n'], w['ops'], w['x'][0], w['words'], w['coeffs'])
    d = dense.Dense(w['n']).run(tape_with_values(w['ops'], w['x'][0]))
    assert rel_close(e[0], eo)
    assert rel_close(e[0], d.z_diagonal_energy(w['words'], w['coeffs']))
    np.testing.assert_allclose(tv, to, atol=1e-11)


@pytest.mark.parametrize('seed', range(4))
def test_random_circuits_xyz_words_vs_dense(cuda, seed):
    r = np.random.default_rng(100 + seed)
    n = int(r.integers(3, 9))
    ops = []
    P = 0
    for _ in range(60):
        kind = r.integers(0, 4)
        q = int(r.integers(0, n))
        if kind == 0:
            ops.append((str(r.choice(['rx', 'ry', 'rz'])), (q,), P))
            P += 1
        elif kind == 1:
            ops.append((str(r.choice(['h', 's', 'sdg', 'x', 'y', 'z'])), (q,), None))
        else:
            t = int(r.integers(0, n - 1))
            ops.append(('cx', (q, t + 1 if t >= q else t), None))
    words = [''.join(r.choice(list('IXYZ'), n)) for _ in range(40)] + ['I' * n]
    coeffs = r.normal(size=len(words))
    X = r.uniform(-np.pi, np.pi, (3, P))
    print(f'case: random circuit seed={seed} n={n} P={P}', flush=True)
    p = Problem(n, ops, P, words, coeffs, max_bond=16, abs_cut=0, rel_cut=0)
    e, tv = Batch([p], [3]).energies(X, terms=True)
    e, tv = e.cpu().numpy(), tv.cpu().numpy().reshape(3, -1)
    for i in range(3):
        d = dense.Dense(n).run(tape_with_values(ops, X[i]))
        exact = np.array([d.expect(wd) for wd in words])
        np.testing.assert_allclose(tv[i], exact, atol=1e-11)
        assert rel_close(e[i], d.energy(words, coeffs))


@pytest.mark.parametrize('alphabet', ['IXZ', 'IZ'])
@pytest.mark.parametrize('seed', range(3))
def test_real_circuits_real_path_vs_complex_and_dense(cuda, seed, alphabet):
    """Real gates on |0..0> and I/X/Z words run the real-arithmetic forests
    (I/Z-only words: the I/Z chain kernels); they must agree with the complex
    kernels and the dense statevector."""
    r = np.random.default_rng(300 + seed)
    n = int(r.integers(4, 10))
    ops, P = [], 0
    for _ in range(70):
        kind = r.integers(0, 4)
        q = int(r.integers(0, n))
        if kind == 0:
            ops.append(('ry', (q,), P))
            P += 1
        elif kind == 1:
            ops.append((str(r.choice(['h', 'x', 'z'])), (q,), None))
        else:
            t = int(r.integers(0, n - 1))
            ops.append((str(r.choice(['cx', 'cx', 'swap'])), (q, t + 1 if t >= q else t), None))
    words = [''.join(r.choice(list(alphabet), n)) for _ in range(50)] + ['I' * n]
    coeffs = r.normal(size=len(words))
    X = r.uniform(-np.pi, np.pi, (4, P))
    print(f'case: real circuit seed={seed} n={n} P={P} words {alphabet}', flush=True)
    p = Problem(n, ops, P, words, coeffs, max_bond=16, abs_cut=0, rel_cut=0)
    assert p.info()['real_arith'] == 1
    b = Batch([p], [4])
    e_re, t_re = b.energies(X, terms=True)
    e_plain = b.energies(X).cpu().numpy()
    b.set_arith('complex')
    e_cx, t_cx = b.energies(X, terms=True)
    e_re, e_cx = e_re.cpu().numpy(), e_cx.cpu().numpy()
    t_re, t_cx = t_re.cpu().numpy().reshape(4, -1), t_cx.cpu().numpy().reshape(4, -1)
    assert np.all(t_re.imag == 0.0)
    np.testing.assert_allclose(t_re.real, t_cx.real, atol=1e-12)
    assert rel_close(e_re, e_cx, 1e-12) and rel_close(e_plain, e_cx, 1e-12)
    for i in range(4):
        d = dense.Dense(n).run(tape_with_values(ops, X[i]))
        exact = np.array([d.expect(wd) for wd in words])
        np.testing.assert_allclose(t_re[i].real, exact.real, atol=1e-11)
        assert rel_close(e_re[i], d.energy(words, coeffs))
        assert rel_close(e_plain[i], d.energy(words, coeffs))


def test_six_entanglement_patterns_vs_dense(cuda):
    n = 6
    labels = [("IIIIIZ", 0.3), ("IIIIZZ", -0.7), ("IIZIII", 0.2), ("IIIIIX", 0.15), ("IIIYII", -0.1),
              ("IIIIII", 0.4)]
    words = [lab[::-1] for lab, _ in labels]
    coeffs = [c for _, c in labels]
    r = np.random.default_rng(17)
    for pattern in ('linear', 'reverse_linear', 'circular', 'full', 'pairwise', 'sca'):
        ops, P = workloads.su2_ansatz(n, 2, ('ry', 'rz'), pattern)
        x = r.uniform(-1, 1, P)
        print(f'case: pattern {pattern}', flush=True)
        e = Batch([Problem(n, ops, P, words, coeffs, max_bond=8, abs_cut=0, rel_cut=0)], [1]).energies(x)
        ref = dense.Dense(n).run(tape_with_values(ops, x)).energy(words, coeffs)
        assert rel_close(e.cpu().numpy()[0], ref), pattern


def test_uneven_multi_problem_batch(cuda):
    w = workloads.config(3)
    parts = w['parts'][:6]
    probs = [Problem(q['n'], q['ops'], q['nparams'], q['words'], q['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
             for q in parts]
    x = np.concatenate([q['x'][0] for q in parts])
    print('case: uneven batch of 6 partitions', flush=True)
    e = Batch(probs, [1] * len(probs)).energies(x).cpu().numpy()
    for q, ei in zip(parts, e):
        eo, _ = oracle_terms(q['n'], q['ops'], q['x'][0], q['words'], q['coeffs'])
        assert rel_close(ei, eo)


def test_nonzero_cutoffs_match_oracle(cuda):
    # reference defaults: max bond 64, cutoffs 1e-5 (simulator_mps_dist.cu:363-366)
    n = 10
    ops, P = workloads.su2_ansatz(n, 3, ('ry', 'rz'), 'linear')
    r = np.random.default_rng(4)
    x = r.uniform(-np.pi, np.pi, P)
    words, coeffs = workloads.z_hamiltonian(n, 80, 4, r)
    print('case: nonzero cutoffs', flush=True)
    e = Batch([Problem(n, ops, P, words, coeffs, max_bond=64, abs_cut=1e-5, rel_cut=1e-5)], [1]).energies(x)
    eo, _ = oracle_terms(n, ops, x, words, coeffs, max_bond=64, cut=1e-5)
    assert rel_close(e.cpu().numpy()[0], eo, 1e-9)


def truncation_margins(sim, max_bond, cut):
    """(truncating gates, smallest relative gap between the last kept and the
    first dropped singular value, smallest relative distance of any singular
    value from the cutoff threshold) over the oracle's SVDs."""
    trunc, gap, thr_margin = 0, np.inf, np.inf
    for s in sim.mps.singular_log:
        k = mps_ref.select_rank(s, max_bond, cut, cut)
        thr = max(cut, cut * s[0])
        if k < len(s):
            trunc += 1
            if s[k] > 1e-13 * s[0]:
                gap = min(gap, (s[k - 1] - s[k]) / s[0])
        if cut > 0:
            thr_margin = min(thr_margin, np.min(np.abs(s - thr)) / thr)
    return trunc, gap, thr_margin


# (qubits, reps, rotations, pattern, max_bond, cutoff, seed, truncates): the
# surface's default engine (max bond 64, cutoffs 1e-5) on a real RY + CX
# campaign circuit (every singular value above the threshold); a 1e-3 cutoff
# that drops singular values at 5-6 gates; 22-qubit circuits whose 32 x 32
# Theta blocks are cut to chi = 16 at 26 gates, one real and one complex
TRUNCATING = [(14, 4, ('ry',), 'linear', 64, 1e-5, 1, False),
              (12, 4, ('ry',), 'linear', 64, 1e-3, 2, True),
              (22, 6, ('ry',), 'linear', 16, 1e-9, 3, True),
              (22, 6, ('ry', 'rz'), 'reverse_linear', 16, 1e-9, 4, True)]


@pytest.mark.parametrize('case', TRUNCATING, ids=['14q-default-engine', '12q-cutoff-1e-3', '22q-real-chi16',
                                                  '22q-complex-chi16'])
def test_truncating_jacobi_svd_matches_oracle(cuda, case):
    """Nonzero cutoffs send every two-site gate through the one-sided Jacobi
    SVD and the cutoff rank rule (dist_svd.cuh:194-231); with chi = 16 the
    32 x 32 Theta blocks are truncated.  The oracle (LAPACK SVD) must show
    the truncations well separated (singular values away from the threshold,
    kept and dropped values apart) so that both sides cut the same subspace;
    energies then agree to 1e-9 relative on the real and the complex
    kernels."""
    n, reps, gates, pattern, max_bond, cut, seed, truncates = case
    r = np.random.default_rng(seed)
    ops, P = workloads.su2_ansatz(n, reps, gates, pattern)
    words, coeffs = workloads.z_hamiltonian(n, 200, 4, r)
    X = np.stack([r.uniform(-np.pi, np.pi, P) for _ in range(2)])
    print(f'case: truncating SVD n={n} reps={reps} gates={gates} max_bond={max_bond} cut={cut}', flush=True)
    p = Problem(n, ops, P, words, coeffs, max_bond=max_bond, abs_cut=cut, rel_cut=cut)
    b = Batch([p], [2])
    e_auto = b.energies(X).cpu().numpy()
    sweeps = b.prep_stats()[:, 0]
    assert np.all(sweeps > 0), 'the Jacobi SVD ran'
    b.set_arith('complex')
    e_cx = b.energies(X).cpu().numpy()
    assert np.all(b.prep_stats()[:, 0] > 0)
    for i in range(2):
        sim = mps_ref.RefSimulator(n, max_bond=max_bond, abs_cut=cut, rel_cut=cut).run(tape_with_values(ops, X[i]))
        trunc, gap, thr_margin = truncation_margins(sim, max_bond, cut)
        assert (trunc > 0) == truncates and gap > 1e-6 and thr_margin > 1e-2, (trunc, gap, thr_margin)
        eo, _ = sim.observe(words, coeffs)
        assert rel_close(e_auto[i], eo, 1e-9), (i, e_auto[i], eo)
        assert rel_close(e_cx[i], eo, 1e-9), (i, e_cx[i], eo)
    if set(gates) == {'ry'}:
        assert p.info()['real_arith'] == 1


def test_bond_guard_error(cuda):
    n = 6
    ops, P = workloads.su2_ansatz(n, 3, ('ry',), 'linear')
    x = np.random.default_rng(1).uniform(-np.pi, np.pi, P)
    b = Batch([Problem(n, ops, P, ['Z' * n], [1.0], max_bond=2, abs_cut=0, rel_cut=0)], [1])
    print('case: bond guard', flush=True)
    with pytest.raises(RuntimeError, match='CUDAQ_MPS_MAX_BOND=2 is insufficient for zero-cutoff simulation'):
        b.energies(x)


def test_spsa_40q_config_parity_and_batch_invariance(cuda):
    w = workloads.config(1)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    print('case: 40q SPSA batch of 256', flush=True)
    e = Batch([p], [256]).energies(w['x']).cpu().numpy()
    assert np.all(np.isfinite(e))
    # batched and one-instance launches run the same per-instance code path
    single = Batch([p], [1])
    for i in (0, 1, 77, 255):
        assert single.energies(w['x'][i]).cpu().numpy()[0] == e[i]
    # oracle parity on every 16th instance (both members of the first pair)
    for i in [0, 1] + list(range(16, 256, 16)):
        eo, _ = oracle_terms(w['n'], w['ops'], w['x'][i], w['words'], w['coeffs'])
        assert rel_close(e[i], eo), (i, e[i], eo)
    # SPSA pairs differ (the +/- perturbations are not degenerate)
    assert np.all(e[0::2] != e[1::2])
    # energy-only chains (lane-partial sums) agree with the per-term path
    small = Batch([p], [4])
    e_only = small.energies(w['x'][:4]).cpu().numpy()
    e_terms, tv = small.energies(w['x'][:4], terms=True)
    e_terms = e_terms.cpu().numpy()
    assert np.all(np.abs(e_only - e_terms) <= 1e-12 * np.maximum(1.0, np.abs(e_terms)))
    c = np.asarray(w['coeffs'])
    tv = tv.cpu().numpy().reshape(4, -1)
    assert np.all(np.abs((tv * c).sum(axis=1).real - e_terms) <= 1e-10 * np.abs(e_terms))


@pytest.mark.skip(reason="launch order 2 (the forest pass as one persistent launch behind grid barriers) did not return within the test's 120 s on its first GPU call (40-qubit batch of 256, 4 waves per group); no cause was found by reading the kernel, so the kernel was taken out of the library and set_launch rejects order 2")
def test_persistent_forest_launch_bitwise(cuda):
    """Launch order 2 (the forest pass of an energy call in one persistent
    launch: every generation's chains, the deferred leaves and the energy
    reduction as phases behind grid barriers) gives bitwise the energies of
    the generation launches, on the 40-qubit workload (I/Z forest) for every
    workgroup size and on a 14-qubit problem with X words (general forest);
    repeated calls reuse the self-resetting barrier words, and per-term calls
    fall back to the generation launches."""
    w = workloads.config(1)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    assert p.info()['iz_forest'] == 1
    b = Batch([p], [256])
    ref = b.energies(w['x']).cpu().numpy()
    for wpb in (4, 2, 1):
        print(f'case: 40q SPSA batch of 256, persistent launch, {wpb} waves per group', flush=True)
        b.set_launch(wpb, 2)
        for rep in range(2):
            assert np.array_equal(b.energies(w['x']).cpu().numpy(), ref), (wpb, rep)
    r = np.random.default_rng(11)
    n = 14
    ops, P = workloads.su2_ansatz(n, 3, ('ry',), 'linear')
    words = [''.join(r.choice(list('IIXZ'), n)) for _ in range(300)] + ['I' * n]
    coeffs = r.normal(size=len(words))
    q = Problem(n, ops, P, words, coeffs, max_bond=16, abs_cut=0, rel_cut=0)
    assert q.info()['iz_forest'] == 0 and q.info()['real_arith'] == 1
    X = r.uniform(-np.pi, np.pi, (64, P))
    print(f'case: {n}q, 300 I/X/Z words, batch 64, orders 1/2', flush=True)
    bq = Batch([q], [64])
    e1 = bq.energies(X).cpu().numpy()
    bq.set_launch(4, 2)
    assert np.array_equal(bq.energies(X).cpu().numpy(), e1)
    e_t, _ = bq.energies(X, terms=True)
    assert np.all(np.abs(e_t.cpu().numpy() - e1) <= 1e-12 * np.maximum(1.0, np.abs(e1)))
    d = dense.Dense(n).run(tape_with_values(ops, X[0]))
    assert rel_close(e1[0], d.energy(words, coeffs))


def test_zero_structure_skipping_is_exact(cuda):
    """The chain steps that skip all-zero row groups and merge the two
    products where the sites' components have disjoint columns give the
    energies and per-term values of the full products, bitwise: on the
    40-qubit workload (I/Z forest; its bulk sites are column-disjoint in the
    transposed copy, row-split in the direct one), on real circuits with X
    words (general forest: the merged product's off-diagonal blocks feed the
    X children) for linear, reverse-linear and random CNOT patterns, and on
    an uneven multi-problem batch."""
    w = workloads.config(1)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    b = Batch([p], [64])
    print('case: 40q x 64, zero-structure skipping on/off', flush=True)
    on = b.energies(w['x'][:64]).cpu().numpy()
    fl_on = b.chain_flops()
    b.set_sparsity(False)
    off = b.energies(w['x'][:64]).cpu().numpy()
    assert np.array_equal(on, off)
    # executed-flop count: the planner's full count without skips; with them,
    # every bulk step issues half its products (only sites 0-4 keep full ones)
    full = p.info()['flops_energy_executed'] * 0.25 * 64
    assert b.chain_flops() == pytest.approx(full, rel=1e-12)
    assert 0.5 * full < fl_on < 0.6 * full
    r = np.random.default_rng(21)
    probs, X, tapes = [], [], []
    for n, pattern in ((12, 'linear'), (15, 'reverse_linear'), (10, 'linear')):
        ops, P = workloads.su2_ansatz(n, 4, ('ry',), pattern)
        if n == 10:  # random neighbouring CNOT directions between the RY layers
            ops = [o for o in ops if o[0] != 'cx']
            for _ in range(30):
                a = int(r.integers(0, n - 1))
                ops.append(('cx', (a, a + 1) if r.random() < 0.5 else (a + 1, a), None))
        words = [''.join(r.choice(list('IIXZ'), n)) for _ in range(200)] + ['I' * n]
        q = Problem(n, ops, P, words, r.normal(size=len(words)), max_bond=16, abs_cut=0, rel_cut=0)
        assert q.info()['real_arith'] == 1 and q.info()['iz_forest'] == 0
        probs.append(q)
        tapes.append(ops)
        X.append(r.uniform(-np.pi, np.pi, (4, P)))
    x = np.concatenate([xx.reshape(-1) for xx in X])
    print('case: 3 real problems with X words, batch of 4 each, skipping on/off', flush=True)
    bu = Batch(probs, [4, 4, 4])
    e_on, t_on = bu.energies(x, terms=True)
    e_on, t_on = e_on.cpu().numpy(), t_on.cpu().numpy()
    bu.set_sparsity(False)
    e_off, t_off = bu.energies(x, terms=True)
    assert np.array_equal(e_on, e_off.cpu().numpy())
    assert np.array_equal(t_on, t_off.cpu().numpy())
    for j, (q, ops, xx) in enumerate(zip(probs, tapes, X)):
        for k in range(4):
            d = dense.Dense(q.n).run(tape_with_values(ops, xx[k]))
            assert rel_close(e_on[4 * j + k], d.energy(q.words, q.coeffs))


def test_bond_doubling_preparation_matches_svd(cuda):
    """Preparation mode 'doubling' (exact CNOT bond doubling, no SVD) gives the
    same energies (1e-10 relative) and the same samples as the SVD path."""
    w = workloads.config(1)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    X = w['x'][:4]
    print('case: spsa-40q doubling vs svd preparation', flush=True)
    b = Batch([p], [4])
    e_svd = b.energies(X).cpu().numpy()
    b.set_prep('doubling')
    e_dbl = b.energies(X).cpu().numpy()
    assert rel_close(e_dbl, e_svd)
    w0 = workloads.config(0)
    p0 = Problem(w0['n'], w0['ops'], w0['nparams'], w0['words'], w0['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    b0 = Batch([p0], [1])
    bits_svd, _ = b0.sample(w0['x'], 1000, [12345])
    b0.set_prep('doubling')
    bits_dbl, _ = b0.sample(w0['x'], 1000, [12345])
    assert np.array_equal(bits_svd.cpu().numpy(), bits_dbl.cpu().numpy())


def test_exact_split_matches_factorisation(cuda):
    """Problems whose splits all keep the whole of Theta (zero cutoffs, no
    rank reduction) skip the decomposition (Theta = I Theta or Theta I).
    The state is the same as with the pivoted QR / SVD at every gate
    (preparation 'factorize'): energies agree to rounding, on the real and
    the complex kernels, with the oracle too, and the sampled bitstrings are
    the same."""
    w = workloads.config(1)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    assert p.info()['exact_split'] == 1
    X = w['x'][:8]
    print('case: spsa-40q exact split vs factorisation, real and complex kernels', flush=True)
    b = Batch([p], [8])
    for arith in ('auto', 'complex'):
        b.set_arith(arith)
        b.set_prep('svd')
        e_exact = b.energies(X).cpu().numpy()
        b.set_prep('factorize')
        e_fact = b.energies(X).cpu().numpy()
        assert rel_close(e_exact, e_fact, 1e-12), (arith, e_exact - e_fact)
    for i in (0, 5):
        eo, _ = oracle_terms(w['n'], w['ops'], X[i], w['words'], w['coeffs'])
        assert rel_close(e_exact[i], eo)
    r = np.random.default_rng(41)
    n = 10
    ops, P = workloads.su2_ansatz(n, 3, ('ry', 'rz'), 'reverse_linear')
    words = [''.join(r.choice(list('IXYZ'), n)) for _ in range(60)] + ['I' * n]
    coeffs = r.normal(size=len(words))
    Xc = r.uniform(-np.pi, np.pi, (3, P))
    q = Problem(n, ops, P, words, coeffs, max_bond=16, abs_cut=0, rel_cut=0)
    assert q.info()['exact_split'] == 1 and q.info()['real_arith'] == 0
    print('case: 10q complex circuit, X/Y/Z words, exact split vs factorisation vs dense', flush=True)
    bq = Batch([q], [3])
    e1 = bq.energies(Xc).cpu().numpy()
    bq.set_prep('factorize')
    e2 = bq.energies(Xc).cpu().numpy()
    for i in range(3):
        d = dense.Dense(n).run(tape_with_values(ops, Xc[i])).energy(words, coeffs)
        assert rel_close(e1[i], d) and rel_close(e2[i], d)
    w0 = workloads.config(0)
    p0 = Problem(w0['n'], w0['ops'], w0['nparams'], w0['words'], w0['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    print('case: 12q samples, exact split vs factorisation', flush=True)
    b0 = Batch([p0], [1])
    bits_exact, _ = b0.sample(w0['x'], 1000, [9876])
    b0.set_prep('factorize')
    bits_fact, _ = b0.sample(w0['x'], 1000, [9876])
    assert np.array_equal(bits_exact.cpu().numpy(), bits_fact.cpu().numpy())
