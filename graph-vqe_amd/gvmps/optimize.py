"""Batched forms of the gradient / objective consumers that drive the backend.

The reference's optimizer loops ask the backend for many independent energies
one call at a time: SPSA probes x +/- c*Delta (test/tools/spsa_optimizer.py:
59-63, 129-178), central differences (test/tools/lbfgs_optimizer.py:124-135),
parameter-shift gradients (Qiskit_VQE/independent_vqe.py:203-211), COBYLA's
initial simplex, and L-BFGS-B with the reverse gradient cross-checked against
the objective (lbfgs_optimizer.py:137-169, rtol 1e-9, atol 1e-7).  Here each of
these is one batched launch over all its points; results equal the
sequential calls' (same kernels, deterministic reductions).
"""
import numpy as np

from . import cudaq as _cq


class BatchedObjective:
    """Energy of (kernel, H) with a point cache; ``prefetch`` evaluates many
    points in one launch so that later scalar calls are cache hits."""

    def __init__(self, kernel, hamiltonian, shots_count=-1):
        self.kernel, self.hamiltonian, self.shots = kernel, hamiltonian, shots_count
        self.cache = {}
        self.nfev = 0
        self.launches = 0

    def prefetch(self, points):
        X = np.atleast_2d(np.asarray(points, dtype=float))
        todo = [x for x in X if x.tobytes() not in self.cache]
        if todo:
            e = _cq.observe_batch(self.kernel, self.hamiltonian, np.array(todo), shots_count=self.shots)
            self.launches += 1
            for x, v in zip(todo, e):
                self.cache[x.tobytes()] = float(v)
        return np.array([self.cache[x.tobytes()] for x in X])

    def __call__(self, x):
        x = np.asarray(x, dtype=float)
        self.nfev += 1
        v = self.cache.get(x.tobytes())
        if v is None:
            v = float(self.prefetch(x[None, :])[0])
        return v


def perturbation_gradients(objective, x, c, deltas):
    """SPSA estimates ((f(x+cD) - f(x-cD)) / 2c) D for every row D of `deltas`
    (spsa_optimizer.perturbation_gradient), all 2K probes in one launch."""
    x = np.asarray(x, dtype=float)
    D = np.atleast_2d(np.asarray(deltas, dtype=float))
    pts = np.concatenate([x + c * D, x - c * D])
    e = objective.prefetch(pts)
    k = len(D)
    return ((e[:k] - e[k:]) / (2 * c))[:, None] * D


def central_gradient(objective, x, step, bounds=None):
    """lbfgs_optimizer.central_gradient (one-sided at explicit bounds) with
    every probe in one launch."""
    x = np.asarray(x, dtype=float)
    a = np.repeat(x[None, :], x.size, axis=0)
    b = a.copy()
    for i in range(x.size):
        lo, hi = (None, None) if bounds is None else bounds[i]
        a[i, i] = max(x[i] - step, -np.inf if lo is None else lo)
        b[i, i] = min(x[i] + step, np.inf if hi is None else hi)
    e = objective.prefetch(np.concatenate([b, a]))
    den = np.diag(b) - np.diag(a)
    out = np.zeros_like(x)
    nz = den != 0
    out[nz] = (e[:x.size][nz] - e[x.size:][nz]) / den[nz]
    return out


def parameter_shift_gradient(objective, x, shift=np.pi / 2):
    """independent_vqe gradient: (f(x + s e_i) - f(x - s e_i)) / 2 for every i."""
    x = np.asarray(x, dtype=float)
    P = np.repeat(x[None, :], x.size, axis=0) + shift * np.eye(x.size)
    M = np.repeat(x[None, :], x.size, axis=0) - shift * np.eye(x.size)
    e = objective.prefetch(np.concatenate([P, M]))
    return (e[:x.size] - e[x.size:]) / 2


def cobyla_minimize(objective, x0, rhobeg=1.0, **options):
    """scipy COBYLA with its initial simplex (x0 and x0 + rhobeg e_i)
    evaluated in one launch beforehand."""
    from scipy.optimize import minimize
    x0 = np.asarray(x0, dtype=float)
    objective.prefetch(np.concatenate([x0[None, :], x0 + rhobeg * np.eye(x0.size)]))
    opts = dict(options)
    opts['rhobeg'] = rhobeg
    return minimize(objective, x0, method='COBYLA', options=opts)


class _RoundRobin:
    """Gathers one value+gradient request from every running optimizer and
    serves the round with a single batched reverse pass.  A finished
    optimizer's instance keeps its last point (its result is discarded)."""

    def __init__(self, batch, x0s):
        import threading
        self.batch = batch
        self.cond = threading.Condition()
        self.sizes = [len(x) for x in x0s]
        self.last = [np.asarray(x, dtype=float).copy() for x in x0s]
        self.pending, self.results = {}, {}
        self.alive = len(x0s)
        self.rounds = 0
        self.error = None

    def _serve(self):
        try:
            for i, x in self.pending.items():
                self.last[i] = x
            e, g = self.batch.value_grad(np.concatenate(self.last))
            e, g = e.cpu().numpy(), g.cpu().numpy()
            off = np.cumsum([0] + self.sizes)
            for i in self.pending:
                self.results[i] = (float(e[i]), g[off[i]:off[i + 1]].copy())
        except Exception as ex:  # wake every waiter with the error
            self.error = ex
        self.pending = {}
        self.rounds += 1
        self.cond.notify_all()

    def request(self, i, x):
        with self.cond:
            if self.error:
                raise self.error
            self.pending[i] = np.asarray(x, dtype=float).copy()
            if len(self.pending) == self.alive:
                self._serve()
            while i not in self.results and self.error is None:
                self.cond.wait()
            if self.error:
                raise self.error
            return self.results.pop(i)

    def finish(self):
        with self.cond:
            self.alive -= 1
            if self.pending and len(self.pending) == self.alive:
                self._serve()


def lbfgs_many(problems, x0s, maxiter=200, ftol=1e-9, gtol=1e-3, maxls=40, maxcor=10):
    """scipy L-BFGS-B on every (problem, x0) at once, with exact adjoint
    gradients: each optimizer runs in its own thread, and all their requests
    of one round are evaluated by one batched reverse pass.  Every iterate
    equals that of a sequential scipy run (the batched values are bitwise the
    single-instance values).  Returns (results, rounds)."""
    import threading
    from scipy.optimize import minimize
    from .engine import Batch
    batch = Batch(list(problems), [1] * len(problems))
    rr = _RoundRobin(batch, x0s)
    results = [None] * len(problems)
    errors = []

    def run(i):
        try:
            results[i] = minimize(lambda x: rr.request(i, x), np.asarray(x0s[i], dtype=float), jac=True,
                                  method='L-BFGS-B',
                                  options=dict(maxiter=maxiter, ftol=ftol, gtol=gtol, maxls=maxls, maxcor=maxcor))
        except Exception as ex:
            errors.append(ex)
        finally:
            rr.finish()

    threads = [threading.Thread(target=run, args=(i,)) for i in range(len(problems))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return results, rr.rounds


def lbfgs_minimize(kernel, hamiltonian, x0, bounds=None, maxiter=100, ftol=1e-9, gtol=1e-3,
                   reverse_mode='auto', step=1e-6, callback=None):
    """L-BFGS-B with the adjoint gradient.  The first reverse energy is checked
    against the objective (rtol 1e-9, atol 1e-7); on a mismatch 'required'
    raises and 'auto' falls back to batched central differences."""
    from scipy.optimize import minimize
    objective = BatchedObjective(kernel, hamiltonian)
    reverse = _cq.make_problem(kernel, hamiltonian) if reverse_mode != 'off' else None
    state = dict(checked=False, reverse=reverse, njev=0)

    def jac(x):
        state['njev'] += 1
        rev = state['reverse']
        if rev is None:
            return central_gradient(objective, x, step, bounds)
        e, g = rev.value_gradient(x)
        if not state['checked']:
            observed = objective(x)
            if not np.isclose(e, observed, rtol=1e-9, atol=1e-7):
                if reverse_mode == 'required':
                    raise RuntimeError(f'Reverse energy {e} differs from the objective {observed}')
                state['reverse'] = None
                return central_gradient(objective, x, step, bounds)
            state['checked'] = True
        objective.cache.setdefault(np.asarray(x, dtype=float).tobytes(), float(e))
        return g

    res = minimize(objective, np.asarray(x0, dtype=float), method='L-BFGS-B', jac=jac, bounds=bounds,
                   callback=callback, options=dict(maxiter=maxiter, ftol=ftol, gtol=gtol, maxls=40, maxcor=10))
    res.njev_reverse = state['njev']
    return res
