"""Learned warm start for the global L-BFGS-B solve.

    python -m gvmps.learned_warmstart --train 2000 --eval 100 --out ws.json
    python -m gvmps.learned_warmstart --evaluate-pretrained --eval 100 --seed 4049 --out eval.json

Pipeline:
  1. Training data are (Hamiltonian, converged parameters) pairs. The plain
     GPU VQE produces them: a partition warm start (local L-BFGS-B on every
     block, gvmps.warmstart), then a global L-BFGS-B solve with adjoint
     gradients. The Hamiltonians are synthetic Z-string Hamiltonians of
     20-40 qubits with 1000-4000 terms. The ansatz has 4 RY layers with
     linear CX chains between them, which gives 4n angles.
  2. The graph network (gvmps.gnn, HIP kernels) is trained with Adam on the
     periodic angle loss. It can then be fine-tuned on the MPS energy of its
     own predictions: dE/dangle comes from the adjoint kernel and is
     back-propagated through the network.
  3. Evaluation uses held-out 40-qubit Hamiltonians. For each start, the
     global L-BFGS-B iterations and the final energy are recorded, from the
     network's angles and from the partition warm start. The MPS energy
     judges both.  compare_starts counts, per Hamiltonian, the learned
     route's iterations until it reaches the partition route's final
     energy (charging a run that never does its own iterations plus the
     partition route's), under the campaign's L-BFGS-B rule.
All optimizer runs of one stage are batched by optimize.lbfgs_many.
"""
import argparse
import json
import time

import numpy as np

from . import gnn as G
from . import optimize, warmstart, workloads
from .engine import Problem

REPS, GATES, PATTERN = 3, ('ry',), 'linear'
LBFGS = dict(maxiter=300, ftol=1e-9, gtol=1e-3)
# The campaign's L-BFGS-B adapter (test/tools/lbfgs_optimizer.py:58-67,
# 172-186 with run_case.py's energy_tolerance 1): ftol 1e-9, gtol 1e-3, no
# iteration cap, and the immediate energy stop as soon as two consecutive
# objective evaluations differ by less than 1; local and global solves alike.
CAMPAIGN_LBFGS = dict(maxiter=100000, ftol=1e-9, gtol=1e-3, immediate_tol=1.0)


def log(msg):
    print(f'[{time.strftime("%H:%M:%S")}] {msg}', flush=True)


def hamiltonians(count, nmin, nmax, tmin, tmax, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        T = int(rng.integers(tmin, tmax + 1))
        words, coeffs = workloads.z_hamiltonian(n, T, 6, rng)
        out.append((n, words, np.real(np.asarray(coeffs, dtype=complex))))
    return out


def ansatz(n):
    return workloads.su2_ansatz(n, REPS, GATES, PATTERN)


def problem(n, words, coeffs):
    ops, P = ansatz(n)
    return Problem(n, ops, P, words, coeffs, max_bond=16, abs_cut=0.0, rel_cut=0.0)


def partition_starts(hams, seed, chunk=256, local_maxiter=100, theta0s=None, rule=None):
    """Partition warm starts of every Hamiltonian; local solves batched.  The
    local solves start from theta0s[h] (default: the reference's seeded
    uniform draw, independent_vqe.py:261-263); `rule` replaces their
    L-BFGS-B options (default ftol 1e-12, gtol 1e-8, local_maxiter, as in
    independent_vqe.py:233-241)."""
    opts = dict(maxiter=local_maxiter, ftol=1e-12, gtol=1e-8) if rule is None else dict(rule)
    starts, local_problems, local_x0, owners = [], [], [], []
    for h, (n, words, coeffs) in enumerate(hams):
        P = (REPS + 1) * n * len(GATES)
        theta = np.random.default_rng(seed + h).uniform(-np.pi, np.pi, P) if theta0s is None else \
            np.asarray(theta0s[h], dtype=float).copy()
        blocks = warmstart.partition(warmstart.interaction_graph(n, words, coeffs), 12, seed)
        models, _ = warmstart.local_models(n, words, coeffs, blocks, REPS, GATES, PATTERN)
        for m in models:
            if m['constant']:
                continue
            local_problems.append(Problem(m['width'], m['ops'], m['nparams'], m['words'], m['coeffs'],
                                          max_bond=16, abs_cut=0.0, rel_cut=0.0))
            local_x0.append(theta[m['parameter_indices']])
            owners.append((h, m['parameter_indices']))
        starts.append(theta)
    local_nit = np.zeros(len(hams), dtype=int)
    for a in range(0, len(local_problems), chunk):
        res, _ = optimize.lbfgs_many(local_problems[a:a + chunk], local_x0[a:a + chunk], **opts)
        for r, (h, idx) in zip(res, owners[a:a + chunk]):
            starts[h][idx] = r.x
            local_nit[h] += r.nit
    return starts, local_nit


def solve(hams, starts, chunk=128, label=None, rule=None, full=False):
    """Global L-BFGS-B from `starts`: [(x, energy, nit, nfev)] (full: the
    scipy results, with their per-iterate energies)."""
    out = []
    for a in range(0, len(hams), chunk):
        probs = [problem(*h) for h in hams[a:a + chunk]]
        res, rounds = optimize.lbfgs_many(probs, starts[a:a + chunk], **(LBFGS if rule is None else rule))
        out.extend(res if full else [(r.x, float(r.fun), int(r.nit), int(r.nfev)) for r in res])
        if label:
            log(f'{label}: {len(out)}/{len(hams)} solved ({rounds} batched rounds)')
    return out


def iterations_to_reach(iterate_energies, target, tol):
    """First accepted iteration whose energy is within tol of target or below
    (0 = the start), or None."""
    for k, e in enumerate(iterate_energies):
        if e <= target + tol:
            return k
    return None


def compare_starts(net, hams, seed, rule=CAMPAIGN_LBFGS, reach_rtol=1e-6):
    """Learned start against the partition warm start under one L-BFGS-B
    rule, at equal final energy.

    For every Hamiltonian both routes run the global L-BFGS-B (the partition
    route after its local solves).  E_p is the partition route's final
    energy and N_p its global iterations.  The learned route's cost is the
    number of its iterations until an iterate is within reach_rtol * sum|c|
    of E_p or below; when it never gets there, it is charged its own
    iterations plus N_p (it ends in a worse basin and the partition route
    has to be run as well).  cut = 1 - mean(charged) / mean(N_p)."""
    part_starts, local_nit = partition_starts(hams, seed, rule=rule)
    part = solve(hams, part_starts, rule=rule, full=True)
    learned = solve(hams, predict_all(net, hams), rule=rule, full=True)
    rows = []
    for h, (p, l) in enumerate(zip(part, learned)):
        scale = float(np.sum(np.abs(hams[h][2])))
        reach = iterations_to_reach(l.iterate_energies, float(p.fun), reach_rtol * scale)
        rows.append(dict(n=hams[h][0], terms=len(hams[h][1]), e_partition=float(p.fun), nit_partition=int(p.nit),
                         local_nit_partition=int(local_nit[h]), e_learned=float(l.fun), nit_learned=int(l.nit),
                         learned_reach=reach, charged_learned=reach if reach is not None else int(l.nit) + int(p.nit),
                         e_learned_start=float(l.iterate_energies[0]), scale=scale))
    nit_p = np.array([r['nit_partition'] for r in rows], dtype=float)
    charged = np.array([r['charged_learned'] for r in rows], dtype=float)
    reached = [r for r in rows if r['learned_reach'] is not None]
    e_p = np.array([r['e_partition'] for r in rows])
    e_l = np.array([r['e_learned'] for r in rows])
    scale = np.array([r['scale'] for r in rows])
    summary = dict(
        hamiltonians=len(rows), rule={k: v for k, v in rule.items()},
        mean_nit_partition=float(nit_p.mean()), mean_local_nit_partition=float(np.mean(local_nit)),
        mean_charged_learned=float(charged.mean()), iteration_cut_at_equal_energy=float(1 - charged.mean() / nit_p.mean()),
        reached_fraction=len(reached) / len(rows),
        mean_reach_when_reached=float(np.mean([r['learned_reach'] for r in reached])) if reached else None,
        mean_nit_learned=float(np.mean([r['nit_learned'] for r in rows])),
        mean_e_partition=float(e_p.mean()), mean_e_learned=float(e_l.mean()),
        mean_rel_energy_diff=float(np.mean((e_l - e_p) / scale)),
        learned_not_worse_fraction=float(np.mean(e_l <= e_p + reach_rtol * scale)))
    return summary, rows


def magnetization_targets(hams, thetas):
    """Per-node targets from converged parameters.  Converged angles are not
    unique (many angle settings give the same state), so the label is the
    state's magnetisation m_v = <Z_v>, a property of the state.  The product
    state with the same magnetisations is reached by the last RY layer alone:
    targets (0, ..., 0, arccos m_v)."""
    from .engine import Batch
    probs = []
    for n, _, _ in hams:
        words = ['I' * v + 'Z' + 'I' * (n - v - 1) for v in range(n)]
        ops, P = ansatz(n)
        probs.append(Problem(n, ops, P, words, np.ones(n), max_bond=16, abs_cut=0.0, rel_cut=0.0))
    out = []
    for a in range(0, len(hams), 256):
        b = Batch(probs[a:a + 256], [1] * len(probs[a:a + 256]))
        _, tv = b.energies(np.concatenate(thetas[a:a + 256]), terms=True)
        m, off = tv.real.cpu().numpy(), 0
        for n, _, _ in hams[a:a + 256]:
            t = np.zeros((n, REPS + 1))
            t[:, REPS] = np.arccos(np.clip(m[off:off + n], -1.0, 1.0))
            out.append(t)
            off += n
    return out


def graph_batch(hams):
    return G.GraphBatch([(G.node_features(n, w, c), G.edges(n, w, c)) for n, w, c in hams])


def train(net, hams, targets, epochs=200, graphs_per_step=32, lr=2e-3, seed=0, log_every=20):
    """Adam on the periodic angle loss; targets[i] is (n_i, layers)."""
    rng = np.random.default_rng(seed)
    order = np.arange(len(hams))
    batches = []
    for a in range(0, len(order), graphs_per_step):
        idx = order[a:a + graphs_per_step]
        batches.append((graph_batch([hams[i] for i in idx]),
                        np.concatenate([targets[i] for i in idx]).astype(np.float32)))
    history = []
    for ep in range(epochs):
        tot = 0.0
        for b in rng.permutation(len(batches)):
            gb, tg = batches[b]
            tot += net.loss_grad(gb, tg)
            net.adam(lr)
        history.append(tot / len(batches))
        if log_every and (ep % log_every == 0 or ep == epochs - 1):
            log(f'epoch {ep}: angle loss {history[-1]:.5f}')
    return history


def energy_finetune(net, hams, steps=50, graphs_per_step=32, lr=5e-4, seed=0, log_every=10):
    """Minimise the mean scaled MPS energy of the network's own predictions:
    dE/dangle from the adjoint kernel, back-propagated through the network."""
    from .engine import Batch
    rng = np.random.default_rng(seed)
    scales = [float(np.sum(np.abs(c))) for _, _, c in hams]
    hist = []
    for st in range(steps):
        idx = rng.choice(len(hams), size=min(graphs_per_step, len(hams)), replace=False)
        sub = [hams[i] for i in idx]
        gb = graph_batch(sub)
        y = net.forward(gb).cpu().numpy().astype(np.float64)
        X, off = [], 0
        for n, _, _ in sub:
            X.append(y[off:off + n].T.reshape(-1))
            off += n
        b = Batch([problem(*h) for h in sub], [1] * len(sub))
        e, g = b.value_grad(np.concatenate(X))
        e, g = e.cpu().numpy(), g.cpu().numpy()
        dy, goff, noff = np.zeros_like(y), 0, 0
        for k, (n, _, _) in enumerate(sub):
            P = (REPS + 1) * n
            dy[noff:noff + n] = g[goff:goff + P].reshape(REPS + 1, n).T / (scales[idx[k]] * len(sub))
            goff += P
            noff += n
        net.backward(gb, dy.astype(np.float32))
        net.adam(lr)
        hist.append(float(np.mean(e / np.array([scales[i] for i in idx]))))
        if log_every and (st % log_every == 0 or st == steps - 1):
            log(f'energy fine-tune step {st}: mean E/sum|c| {hist[-1]:.5f}')
    return hist


def predict_all(net, hams):
    gb = graph_batch(hams)
    y = net.forward(gb).cpu().numpy().astype(np.float64)
    out, off = [], 0
    for n, _, _ in hams:
        out.append(y[off:off + n].T.reshape(-1))
        off += n
    return out


def evaluate(net, hams, seed):
    part_starts, local_nit = partition_starts(hams, seed)
    part = solve(hams, part_starts, label='held-out, partition start')
    learned_starts = predict_all(net, hams)
    learned = solve(hams, learned_starts, label='held-out, learned start')
    rows = []
    for h, (p, l) in enumerate(zip(part, learned)):
        rows.append(dict(n=hams[h][0], terms=len(hams[h][1]), nit_partition=p[2], e_partition=p[1],
                         nit_learned=l[2], e_learned=l[1], local_nit_partition=int(local_nit[h])))
    nit_p = np.array([r['nit_partition'] for r in rows])
    nit_l = np.array([r['nit_learned'] for r in rows])
    e_p = np.array([r['e_partition'] for r in rows])
    e_l = np.array([r['e_learned'] for r in rows])
    scale = np.array([np.sum(np.abs(c)) for _, _, c in hams])
    summary = dict(
        hamiltonians=len(rows), mean_nit_partition=float(nit_p.mean()), mean_nit_learned=float(nit_l.mean()),
        iteration_cut=float(1 - nit_l.mean() / nit_p.mean()) if nit_p.mean() > 0 else 0.0,
        mean_local_nit_partition=float(np.mean([r['local_nit_partition'] for r in rows])),
        mean_e_partition=float(e_p.mean()), mean_e_learned=float(e_l.mean()),
        mean_rel_energy_diff=float(np.mean((e_l - e_p) / scale)),
        learned_not_worse_fraction=float(np.mean(e_l <= e_p + 1e-6 * scale)))
    return summary, rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--train', type=int, default=2000)
    ap.add_argument('--eval', type=int, default=100)
    ap.add_argument('--nmin', type=int, default=20)
    ap.add_argument('--nmax', type=int, default=40)
    ap.add_argument('--tmin', type=int, default=1000)
    ap.add_argument('--tmax', type=int, default=4000)
    ap.add_argument('--eval-n', type=int, default=40)
    ap.add_argument('--epochs', type=int, default=300)
    ap.add_argument('--finetune', type=int, default=0, help='energy fine-tuning steps after pretraining')
    ap.add_argument('--seed', type=int, default=2026)
    ap.add_argument('--out', default='learned_warmstart.json')
    ap.add_argument('--weights', default='warmstart_gnn.npz')
    ap.add_argument('--evaluate-pretrained', action='store_true',
                    help='only compare the shipped network with the partition start on --eval held-out '
                         'Hamiltonians of --eval-n qubits (seed --seed), under the campaign rule and without '
                         'the immediate energy stop')
    args = ap.parse_args(argv)
    t0 = time.time()
    if args.evaluate_pretrained:
        eval_h = hamiltonians(args.eval, args.eval_n, args.eval_n, args.tmin, args.tmax, args.seed)
        net = G.WarmStartGNN.pretrained()
        report = dict(args=vars(args))
        for name, rule in (('campaign_rule', CAMPAIGN_LBFGS), ('no_immediate_stop', LBFGS)):
            summary, rows = compare_starts(net, eval_h, args.seed + 7, rule=rule)
            log(f'{name}: {json.dumps(summary)}')
            report[name] = dict(summary=summary, rows=rows)
        report['seconds'] = time.time() - t0
        with open(args.out, 'w') as f:
            json.dump(report, f)
        return report
    train_h = hamiltonians(args.train, args.nmin, args.nmax, args.tmin, args.tmax, args.seed)
    eval_h = hamiltonians(args.eval, args.eval_n, args.eval_n, args.tmin, args.tmax, args.seed + 1)
    log(f'{len(train_h)} training / {len(eval_h)} held-out Hamiltonians')
    starts = []
    for a in range(0, len(train_h), 256):
        starts += partition_starts(train_h[a:a + 256], args.seed + a)[0]
        log(f'partition warm starts: {len(starts)}/{len(train_h)}')
    solved = solve(train_h, starts, label='plain VQE (training data)')
    log(f'plain VQE converged: mean nit {np.mean([s[2] for s in solved]):.1f}')
    net = G.WarmStartGNN(seed=args.seed)
    hist = train(net, train_h, magnetization_targets(train_h, [s[0] for s in solved]), epochs=args.epochs)
    ft = energy_finetune(net, train_h, steps=args.finetune) if args.finetune else []
    summary, rows = evaluate(net, eval_h, args.seed + 7)
    summary.update(train_loss_first=hist[0], train_loss_last=hist[-1], finetune_last=ft[-1] if ft else None,
                   seconds=time.time() - t0, train=args.train, eval=args.eval)
    log(json.dumps(summary))
    with open(args.out, 'w') as f:
        json.dump(dict(summary=summary, rows=rows, args=vars(args)), f)
    net.save(args.weights)
    return summary


if __name__ == '__main__':
    main()
