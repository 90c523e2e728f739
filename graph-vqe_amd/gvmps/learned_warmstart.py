"""Learned warm start for the global L-BFGS-B solve.

    python -m gvmps.learned_warmstart --train 2000 --eval 100 --out ws.json

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
     (complex128) judges both.
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


def partition_starts(hams, seed, chunk=256, local_maxiter=100, theta0s=None):
    """Partition warm starts of every Hamiltonian; local solves batched.  The
    local solves start from theta0s[h] (default: the reference's seeded
    uniform draw, independent_vqe.py:261-263)."""
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
        res, _ = optimize.lbfgs_many(local_problems[a:a + chunk], local_x0[a:a + chunk], maxiter=local_maxiter,
                                     ftol=1e-12, gtol=1e-8)
        for r, (h, idx) in zip(res, owners[a:a + chunk]):
            starts[h][idx] = r.x
            local_nit[h] += r.nit
    return starts, local_nit


def solve(hams, starts, chunk=128, label=None):
    """Global L-BFGS-B from `starts`: [(x, energy, nit, nfev)]."""
    out = []
    for a in range(0, len(hams), chunk):
        probs = [problem(*h) for h in hams[a:a + chunk]]
        res, rounds = optimize.lbfgs_many(probs, starts[a:a + chunk], **LBFGS)
        out.extend((r.x, float(r.fun), int(r.nit), int(r.nfev)) for r in res)
        if label:
            log(f'{label}: {len(out)}/{len(hams)} solved ({rounds} batched rounds)')
    return out


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
    args = ap.parse_args(argv)
    t0 = time.time()
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
