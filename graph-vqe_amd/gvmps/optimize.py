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


class _ImmediateStop(Exception):
    pass


def lbfgs_many(problems, x0s, maxiter=200, ftol=1e-9, gtol=1e-3, maxls=40, maxcor=10, immediate_tol=None):
    """scipy L-BFGS-B on every (problem, x0) at once, with exact adjoint
    gradients: each optimizer runs in its own thread, and all their requests
    of one round are evaluated by one batched reverse pass.  Every iterate
    equals that of a sequential scipy run (the batched values are bitwise the
    single-instance values).

    immediate_tol: the campaign adapter's immediate energy stop
    (test/tools/lbfgs_optimizer.py:64-67, 98-117, 172-186): a run ends as
    soon as two consecutive objective evaluations differ by less than the
    tolerance, and returns the best point evaluated.  Every result carries
    `energies` (every objective evaluation) and `iterate_energies` (the start,
    then each accepted iterate).  Returns (results, rounds)."""
    import threading
    from scipy.optimize import OptimizeResult, minimize
    from .engine import Batch
    batch = Batch(list(problems), [1] * len(problems))
    rr = _RoundRobin(batch, x0s)
    results = [None] * len(problems)
    errors = []

    def run(i):
        x0 = np.asarray(x0s[i], dtype=float)
        st = dict(prev=None, best=(np.inf, x0.copy()), evals=[], iters=[], nit=0)

        def fg(x):
            e, g = rr.request(i, x)
            st['evals'].append(e)
            if e < st['best'][0]:
                st['best'] = (e, np.asarray(x, dtype=float).copy())
            if not st['iters']:
                st['iters'].append(e)
            delta = None if st['prev'] is None else abs(e - st['prev'])
            st['prev'] = e
            if immediate_tol is not None and delta is not None and delta < immediate_tol:
                raise _ImmediateStop()
            return e, g

        def accepted(xk):
            st['nit'] += 1
            st['iters'].append(st['evals'][-1])

        try:
            try:
                res = minimize(fg, x0, jac=True, method='L-BFGS-B', callback=accepted,
                               options=dict(maxiter=maxiter, ftol=ftol, gtol=gtol, maxls=maxls, maxcor=maxcor))
                res.energies, res.iterate_energies = st['evals'], st['iters']
            except _ImmediateStop:
                res = OptimizeResult(x=st['best'][1], fun=st['best'][0], nit=st['nit'], nfev=len(st['evals']),
                                     njev=len(st['evals']), success=True,
                                     message=f'Immediate energy stopping: consecutive |delta E| < {immediate_tol:g} '
                                             '(no patience/window)',
                                     energies=st['evals'], iterate_energies=st['iters'])
            results[i] = res
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


# ------------------------------------------------------------------- SPSA
def spsa_settings(environ=None):
    """First-order SPSA hyperparameters from the campaign's environment
    (test/tools/spsa_optimizer.py:41-56): perturbation c (0.2), learning
    rate a ('auto': calibrated), alpha (0.602), gamma (0.101), stability A
    (0), calibration pairs (5), target update (2 pi / 10), seed (0), plus the
    immediate-energy stop tolerance (1, spsa_optimizer.py:172) and the
    iteration cap."""
    import os
    env = os.environ if environ is None else environ
    lr = env.get('ANALYTIC_MPS_SPSA_LEARNING_RATE', 'auto')
    values = dict(perturbation=float(env.get('ANALYTIC_MPS_SPSA_PERTURBATION', '0.2')),
                  learning_rate=None if lr == 'auto' else float(lr),
                  alpha=float(env.get('ANALYTIC_MPS_SPSA_ALPHA', '0.602')),
                  gamma=float(env.get('ANALYTIC_MPS_SPSA_GAMMA', '0.101')),
                  stability=float(env.get('ANALYTIC_MPS_SPSA_STABILITY', '0')),
                  calibration_pairs=int(env.get('ANALYTIC_MPS_SPSA_CALIBRATION_PAIRS', '5')),
                  target_update=float(env.get('ANALYTIC_MPS_SPSA_TARGET_UPDATE', str(2 * np.pi / 10))),
                  seed=int(env.get('ANALYTIC_MPS_SPSA_SEED', '0')),
                  energy_tol=float(env.get('ANALYTIC_MPS_IMMEDIATE_ENERGY_TOL', '1')),
                  max_iter=int(env.get('ANALYTIC_MPS_SPSA_MAX_ITER', '1000000')))
    for key in ('perturbation', 'alpha', 'gamma', 'target_update'):
        if not np.isfinite(values[key]) or values[key] <= 0:
            raise ValueError(f'Invalid SPSA {key}')
    if not np.isfinite(values['stability']) or values['stability'] < 0:
        raise ValueError('Invalid SPSA stability')
    if values['learning_rate'] is not None and (not np.isfinite(values['learning_rate'])
                                                or values['learning_rate'] <= 0):
        raise ValueError('Invalid SPSA learning rate')
    if values['calibration_pairs'] < 1:
        raise ValueError('SPSA calibration_pairs must be positive')
    return values


class _SpsaRun:
    """One SPSA minimisation (spsa_optimizer.py:66-184) as a state machine
    that names the points it needs next instead of calling the objective.

    Round 0 asks for f(x0), the calibration pairs and the probe pair of
    update 1; round k >= 1 asks for f(candidate_k) and, speculatively, the
    probe pair of update k+1 around candidate_k (a candidate is always
    accepted, so x_k = candidate_k is known before its energy is).  The
    Rademacher draws are made in the reference's order (calibration pairs,
    then one per update); a stop after f(candidate_k) discards the
    speculative pair, which the evaluation counts leave out.  The iterates
    are those of the sequential loop, with one objective launch per update."""

    def __init__(self, x0, config, maxiter, bounds):
        x = np.asarray(x0, dtype=float).copy()
        if x.ndim != 1 or not np.isfinite(x).all():
            raise ValueError('Expected finite one-dimensional initial parameters')
        cap = min(int(config['max_iter']), int(maxiter))
        if cap < 1:
            raise ValueError('Positive SPSA iteration cap required')
        if bounds is None:
            lower, upper = np.full(x.size, -np.inf), np.full(x.size, np.inf)
        else:
            lower = np.array([-np.inf if b[0] is None else b[0] for b in bounds], dtype=float)
            upper = np.array([np.inf if b[1] is None else b[1] for b in bounds], dtype=float)
        if lower.shape != x.shape or upper.shape != x.shape or np.any(lower > upper):
            raise ValueError('Invalid bounds')
        self.cfg, self.cap, self.lower, self.upper = config, cap, lower, upper
        self.x = np.clip(x, lower, upper)
        # a local generator, as the reference's (independent of global state)
        self.rng = np.random.default_rng(config['seed'])
        self.counts = dict(objective=0, gradient_probe=0, calibration_probe=0, gradient=0, accepted_iteration=0)
        self.k, self.nit, self.done, self.result = 0, 0, False, None
        self.history = []
        self.a = config['learning_rate']
        self.cal_deltas = []
        if self.x.size and self.a is None:
            self.cal_deltas = [self._draw() for _ in range(config['calibration_pairs'])]
        self.delta = self._draw() if self.x.size else None   # the delta of update k+1

    def _draw(self):
        return self.rng.choice(np.array([-1.0, 1.0]), size=self.x.size)

    def _c(self, k):
        return self.cfg['perturbation'] / k ** self.cfg['gamma']

    def points(self):
        """Points of the next round."""
        center = self.x if self.k == 0 else self.candidate
        pts = [center]
        if self.k == 0:
            c = self.cfg['perturbation']
            for d in self.cal_deltas:
                pts += [center + c * d, center - c * d]
        if self.x.size:
            c = self._c(self.k + 1)
            pts += [center + c * self.delta, center - c * self.delta]
        return pts

    def _advance(self, plus, minus):
        """Update k+1 from its probe pair: gradient estimate, candidate, and
        the delta of the update after it."""
        k = self.k + 1
        g = ((plus - minus) / (2 * self._c(k))) * self.delta
        self.counts['gradient_probe'] += 2
        self.counts['gradient'] += 1
        ak = self.a / (k + self.cfg['stability']) ** self.cfg['alpha']
        self.candidate = np.clip(self.x - ak * g, self.lower, self.upper)
        self.k = k
        self.delta = self._draw()

    def _finish(self, status, success):
        from scipy.optimize import OptimizeResult
        self.done = True
        nfev = sum(self.counts[v] for v in ('objective', 'gradient_probe', 'calibration_probe'))
        self.result = OptimizeResult(x=self.best_x, fun=self.best_f, nit=self.nit, nfev=nfev,
                                     njev=self.counts['gradient'], success=success, message=status,
                                     learning_rate=self.a, energies=list(self.history))

    def feed(self, values, callback=None):
        v = [float(e) for e in values]
        if not all(np.isfinite(v)):
            raise ValueError(f'Nonfinite SPSA energy: {v}')
        if self.k == 0:
            self.fx = v[0]
            self.counts['objective'] += 1
            self.best_x, self.best_f = self.x.copy(), self.fx
            self.history.append(self.fx)
            if not self.x.size:
                self._finish('No variational parameters', True)
                return
            pos = 1
            if self.a is None:
                # target initial component update over the mean magnitude of
                # the calibration pairs' directional differences (:114-123)
                c = self.cfg['perturbation']
                mags = []
                for d in self.cal_deltas:
                    g = ((v[pos] - v[pos + 1]) / (2 * c)) * d
                    mags.append(float(np.mean(np.abs(g))))
                    pos += 2
                self.counts['calibration_probe'] += 2 * len(self.cal_deltas)
                magnitude = float(np.mean(mags))
                self.a = (self.cfg['target_update'] * (self.cfg['stability'] + 1) ** self.cfg['alpha'] / magnitude
                          if magnitude > 1e-12 else 1.0)
            self._advance(v[pos], v[pos + 1])
            return
        previous = self.fx
        self.x, self.fx = self.candidate, v[0]
        self.counts['objective'] += 1
        self.counts['accepted_iteration'] += 1
        self.nit = self.k
        self.history.append(self.fx)
        if self.fx < self.best_f:
            self.best_x, self.best_f = self.x.copy(), self.fx
        if callback is not None:
            callback(self.x)
        tol = self.cfg['energy_tol']
        if abs(self.fx - previous) < tol:
            self._finish(f'Immediate energy stopping: |delta E| < {tol}; no patience/window', True)
        elif self.k >= self.cap:
            self._finish('SPSA iteration cap reached', False)
        else:
            self._advance(v[1], v[2])


def spsa_minimize(objective, x0, maxiter=1000000, bounds=None, callback=None, settings=None):
    """First-order SPSA (test/tools/spsa_optimizer.py:66-184) over a
    batched objective: the initial energy, every calibration pair and the
    first probe pair go in one launch, then each update needs one launch
    (f(candidate) with the next update's probe pair).  Returns the
    reference's OptimizeResult fields: best x and energy, nit, nfev, njev,
    success (immediate-energy stop), message.  `objective` needs a
    ``prefetch(points) -> energies`` method (BatchedObjective)."""
    run = _SpsaRun(x0, settings or spsa_settings(), maxiter, bounds)
    while not run.done:
        run.feed(objective.prefetch(np.asarray(run.points())), callback)
    return run.result


def spsa_many(problems, x0s, maxiter=1000000, bounds=None, settings=None):
    """spsa_minimize on every (problem, x0) at once: each round's points of all
    runs are evaluated by one batched energy launch (finished runs repeat
    their last points, whose energies are ignored).  Every run's iterates
    equal those of its own sequential SPSA loop.  Returns (results, rounds)."""
    from .engine import Batch
    cfg = settings or spsa_settings()
    runs = [_SpsaRun(x, cfg, maxiter, bounds) for x in x0s]
    first = Batch(list(problems), [len(r.points()) for r in runs])
    steady = Batch(list(problems), [3] * len(runs))
    rounds = 0
    last = [None] * len(runs)
    while not all(r.done for r in runs):
        pts = []
        for i, r in enumerate(runs):
            if not r.done:
                last[i] = r.points()
            pts.append(last[i] if len(last[i]) == 3 or rounds == 0 else [last[i][0]] * 3)
        batch = first if rounds == 0 else steady
        e = batch.energies(np.concatenate([np.concatenate(p) for p in pts])).cpu().numpy()
        off = 0
        for r, p in zip(runs, pts):
            if not r.done:
                r.feed(e[off:off + len(p)])
            off += len(p)
        rounds += 1
    return [r.result for r in runs], rounds
