"""Host checks of SPSA.

* oracle/spsa_ref (the sequential restatement of test/tools/spsa_optimizer.py)
  is pinned by closed forms: on a quadratic the central difference is exact,
  so every update is x - a_k (grad f . delta_k) delta_k with the calibrated
  gain a = target / mean|g| and the draws of default_rng(seed) in the
  reference's order (calibration pairs, then one per update).
* gvmps.optimize.spsa_minimize (the batched loop: one objective launch per
  update) reproduces the sequential loop exactly on CPU objectives: same
  iterates, result, counts and stop reason.
"""
import numpy as np
import pytest

from gvmps import optimize
from oracle import spsa_ref


def cfg(**kw):
    c = optimize.spsa_settings({})
    c.update(kw)
    return c


class CountingObjective:
    """A CPU objective with the BatchedObjective interface."""

    def __init__(self, f):
        self.f, self.launches, self.points = f, 0, 0

    def prefetch(self, pts):
        self.launches += 1
        self.points += len(pts)
        return np.array([self.f(p) for p in np.atleast_2d(pts)])


def test_defaults_match_the_reference_settings():
    c = optimize.spsa_settings({})
    assert c['perturbation'] == 0.2 and c['learning_rate'] is None and c['alpha'] == 0.602
    assert c['gamma'] == 0.101 and c['stability'] == 0.0 and c['calibration_pairs'] == 5
    assert c['target_update'] == 2 * np.pi / 10 and c['seed'] == 0 and c['energy_tol'] == 1.0
    env = {'ANALYTIC_MPS_SPSA_LEARNING_RATE': '0.05', 'ANALYTIC_MPS_SPSA_CALIBRATION_PAIRS': '3',
           'ANALYTIC_MPS_IMMEDIATE_ENERGY_TOL': '0.5'}
    c2 = optimize.spsa_settings(env)
    assert c2['learning_rate'] == 0.05 and c2['calibration_pairs'] == 3 and c2['energy_tol'] == 0.5
    for bad in ({'ANALYTIC_MPS_SPSA_PERTURBATION': '0'}, {'ANALYTIC_MPS_SPSA_STABILITY': '-1'},
                {'ANALYTIC_MPS_SPSA_CALIBRATION_PAIRS': '0'}, {'ANALYTIC_MPS_SPSA_LEARNING_RATE': 'nan'}):
        with pytest.raises(ValueError):
            optimize.spsa_settings(bad)


def test_oracle_one_dimensional_closed_form():
    """f = x^2 from x0 = 1: every calibration estimate is 2 (the central
    difference of a quadratic is exact), a = (2 pi / 10) / 2, the first
    update moves to 1 - 2 a = 1 - pi/5, and |f_1 - f_0| < 1 stops at once."""
    r = spsa_ref.minimize(lambda x: float(x[0] ** 2), [1.0], cfg())
    assert r['learning_rate'] == pytest.approx(np.pi / 10, rel=1e-15)
    assert r['x'][0] == pytest.approx(1 - np.pi / 5, rel=1e-14)
    assert r['fun'] == pytest.approx((1 - np.pi / 5) ** 2, rel=1e-14)
    assert (r['nit'], r['nfev'], r['njev'], r['success']) == (1, 14, 1, True)
    assert r['message'] == 'Immediate energy stopping: |delta E| < 1.0; no patience/window'


def test_oracle_quadratic_updates_closed_form():
    w = np.array([1.0, 3.0, 0.5, 2.0])
    f = lambda x: float(np.sum(w * x ** 2))
    x0 = np.array([1.0, -0.5, 2.0, 0.25])
    c = cfg(energy_tol=1e-300, seed=11, calibration_pairs=3, stability=2.0)
    cap = 6
    r = spsa_ref.minimize(f, x0, c, maxiter=cap)
    rng = np.random.default_rng(11)
    grad = lambda x: 2 * w * x
    mags = []
    for _ in range(3):
        d = rng.choice(np.array([-1.0, 1.0]), size=4)
        mags.append(np.mean(np.abs((grad(x0) @ d) * d)))
    a = c['target_update'] * 3.0 ** c['alpha'] / np.mean(mags)
    assert r['learning_rate'] == pytest.approx(a, rel=1e-12)
    x, best = x0.copy(), (f(x0), x0.copy())
    for k in range(1, cap + 1):
        d = rng.choice(np.array([-1.0, 1.0]), size=4)
        x = x - a / (k + 2.0) ** c['alpha'] * (grad(x) @ d) * d
        if f(x) < best[0]:
            best = (f(x), x.copy())
    np.testing.assert_allclose(r['x'], best[1], rtol=1e-11)
    assert r['nit'] == cap and r['message'] == 'SPSA iteration cap reached' and not r['success']
    assert r['nfev'] == 1 + 6 + 2 * cap + cap and r['njev'] == cap


def rugged(x):
    return float(40 * np.sum(np.cos(x)) + 7 * np.sum(np.sin(2 * x[:-1] + x[1:])) + 3 * np.sum(x ** 2))


@pytest.mark.parametrize('case', ['auto', 'fixed-rate', 'bounds', 'cap', 'tight-tol'])
def test_batched_loop_equals_sequential_loop(case):
    rng = np.random.default_rng(3)
    x0 = rng.uniform(-np.pi, np.pi, 9)
    kw, bounds, maxiter = {}, None, 1000000
    if case == 'fixed-rate':
        kw = dict(learning_rate=0.03)
    if case == 'bounds':
        bounds = [(-1.0, 1.0)] * 9
        x0[0] = 3.0   # clipped into the box first
    if case == 'cap':
        maxiter, kw = 4, dict(energy_tol=1e-12)
    if case == 'tight-tol':
        kw = dict(energy_tol=1e-3, seed=5)
    c = cfg(**kw)
    seq_iterates, bat_iterates = [], []
    want = spsa_ref.minimize(rugged, x0, c, maxiter=maxiter, bounds=bounds,
                             callback=lambda x: seq_iterates.append(x.copy()))
    obj = CountingObjective(rugged)
    got = optimize.spsa_minimize(obj, x0, maxiter=maxiter, bounds=bounds, settings=c,
                                 callback=lambda x: bat_iterates.append(x.copy()))
    assert len(seq_iterates) == len(bat_iterates) == want['nit']
    for a, b in zip(seq_iterates, bat_iterates):
        assert np.array_equal(a, b)
    assert np.array_equal(got.x, want['x']) and got.fun == want['fun']
    for key in ('nit', 'nfev', 'njev', 'success', 'message', 'learning_rate'):
        assert got[key] == want[key], key
    seq_energies = [v for kind, _, v in want['trace'] if kind == 'objective']
    assert got.energies == seq_energies
    # one launch for f(x0) + calibration + the first pair, one per update
    assert obj.launches == want['nit'] + 1
