"""Sequential first-order SPSA, restated from the reference (test
infrastructure: the checker of the batched SPSA in gvmps.optimize).

Follows test/tools/spsa_optimizer.py:59-63 (perturbation_gradient) and
:66-184 (minimize) without the journal, work-trace and file logging: a local
default_rng(seed); f(x0); with learning_rate 'auto', `calibration_pairs`
perturbation gradients at c and a = target * (A + 1)^alpha / mean_i
mean|g_i| (1 when that mean is below 1e-12); then for k = 1..cap:
a_k = a / (k + A)^alpha, c_k = c / k^gamma, a Rademacher delta, the probes
f(x + c_k delta) and f(x - c_k delta) in that order, the candidate
clip(x - a_k g) always accepted, f(candidate), best tracking, callback,
and the immediate-energy stop |f_k - f_{k-1}| < tol.  Every objective call
is a separate scalar call, as in the reference.
"""
import numpy as np


def perturbation_gradient(fun, x, c, rng):
    """spsa_optimizer.py:59-63."""
    delta = rng.choice(np.array([-1., 1.]), size=x.size)
    plus = float(fun(x + c * delta))
    minus = float(fun(x - c * delta))
    return ((plus - minus) / (2 * c)) * delta


def minimize(fun, x0, config, maxiter=1000000, bounds=None, callback=None):
    """spsa_optimizer.py:66-184; returns a dict with the result fields."""
    cap = min(int(config['max_iter']), int(maxiter))
    x = np.asarray(x0, dtype=float).copy()
    if bounds is None:
        lower, upper = np.full(x.size, -np.inf), np.full(x.size, np.inf)
    else:
        lower = np.array([-np.inf if b[0] is None else b[0] for b in bounds])
        upper = np.array([np.inf if b[1] is None else b[1] for b in bounds])
    x = np.clip(x, lower, upper)
    rng = np.random.default_rng(config['seed'])
    counts = dict(objective=0, gradient_probe=0, calibration_probe=0, gradient=0)
    trace = []

    def evaluate(point, kind):
        value = float(fun(point))
        counts[kind] += 1
        trace.append((kind, point.copy(), value))
        return value

    fx = evaluate(x, 'objective')
    best_x, best_f = x.copy(), fx
    nit = 0
    status = 'SPSA iteration cap reached'
    if not x.size:
        return dict(x=x, fun=fx, nit=0, nfev=1, njev=0, success=True, message='No variational parameters',
                    learning_rate=None, trace=trace)
    c, a = config['perturbation'], config['learning_rate']
    if a is None:
        mags = []
        for _ in range(config['calibration_pairs']):
            g = perturbation_gradient(lambda p: evaluate(p, 'calibration_probe'), x, c, rng)
            mags.append(float(np.mean(np.abs(g))))
        magnitude = float(np.mean(mags))
        a = config['target_update'] * (config['stability'] + 1) ** config['alpha'] / magnitude \
            if magnitude > 1e-12 else 1.0
    for k in range(1, 1 + cap):
        previous = fx
        ak = a / (k + config['stability']) ** config['alpha']
        ck = c / k ** config['gamma']
        g = perturbation_gradient(lambda p: evaluate(p, 'gradient_probe'), x, ck, rng)
        counts['gradient'] += 1
        candidate = np.clip(x - ak * g, lower, upper)
        fx = evaluate(candidate, 'objective')
        x = candidate
        nit = k
        if fx < best_f:
            best_x, best_f = x.copy(), fx
        if callback is not None:
            callback(x)
        if abs(fx - previous) < config['energy_tol']:
            status = f"Immediate energy stopping: |delta E| < {config['energy_tol']}; no patience/window"
            break
    nfev = counts['objective'] + counts['gradient_probe'] + counts['calibration_probe']
    return dict(x=best_x, fun=best_f, nit=nit, nfev=nfev, njev=counts['gradient'],
                success=status.startswith('Immediate energy'), message=status, learning_rate=a, trace=trace)
