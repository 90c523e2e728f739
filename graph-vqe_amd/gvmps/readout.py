"""Readout decoding and per-evaluation backend records.

* ``protein_probabilities`` turns ``sample`` counts (q0 leftmost, the CUDA-Q
  key order) into the Qiskit-ordered probabilities the protein decoder reads
  (test/tools/readout.py:4-9).
* ``install_backend_metrics`` wraps ``observe``/``observe_async`` (and, for a
  sampled run, ``sample``/``sample_async``) of a CUDA-Q-like module so that
  each evaluation appends one JSON row to the backend metrics file, in the
  rows the campaign's recorders write (test/tools/analytic_execution.py:53-62
  for analytic runs, test/tools/sampled_execution.py:20-27 for 1000-shot
  runs) and its record validator audits.  The wrapped functions keep the
  call shape: kernel first, then the Hamiltonian (observe only), then the
  parameters.
"""
import json
import os
import threading
import time
from pathlib import Path

import numpy as np

TARGET = 'tensornet-mps-mqpu-shots'
_lock = threading.Lock()


def protein_probabilities(counts):
    """{bitstring reversed: count / total}; raises on an empty histogram."""
    total = sum(counts.values())
    if total <= 0:
        raise ValueError("Cannot decode an empty measurement histogram")
    return {bits[::-1]: count / total for bits, count in counts.items()}


def _append(path, row):
    with _lock, Path(path).open('a') as f:
        f.write(json.dumps(row) + '\n')


def install_backend_metrics(module, path=None, shots=None):
    """Wrap ``module``'s evaluation calls to log one row per evaluation.

    ``shots=None`` is an analytic run: observe calls run with
    ``shots_count=-1`` and rows carry ``shots=-1, physical_shots=0`` with the
    engine's bond and cutoff settings.  An integer ``shots`` is a sampled run:
    observe and sample calls are forced to that shot count and rows carry
    ``shots`` and ``physical_shots_per_circuit_or_term``.  The path defaults
    to ``$ANALYTIC_MPS_BACKEND_METRICS`` (else backend_evaluations.jsonl).
    Async calls log when ``get()`` first returns.  Returns the wrapped names.

    The campaign recorders' guards come with them.  An analytic run raises on
    a finite ``shots_count`` or a noise model and replaces ``sample``,
    ``sample_async`` and ``get_state`` by functions that raise
    (analytic_execution.py:63-66, 77-79); a sampled run accepts exactly 1000
    shots and replaces ``get_state`` (sampled_execution.py:13-14, 40-41).
    """
    if shots is not None and int(shots) != 1000:
        raise ValueError('Matched sampled suite requires 1000 shots')
    log_path = path or os.environ.get('ANALYTIC_MPS_BACKEND_METRICS', 'backend_evaluations.jsonl')

    def row(operation, start, args):
        if shots is None:
            data = dict(operation=operation, elapsed_seconds=time.perf_counter() - start, shots=-1,
                        physical_shots=0, target=TARGET,
                        mps_max_bond=os.environ.get('CUDAQ_MPS_MAX_BOND'),
                        abs_cutoff=os.environ.get('CUDAQ_MPS_ABS_CUTOFF'),
                        relative_cutoff=os.environ.get('CUDAQ_MPS_RELATIVE_CUTOFF'))
        else:
            data = dict(operation=operation, elapsed_seconds=time.perf_counter() - start, shots=int(shots),
                        physical_shots_per_circuit_or_term=int(shots), target=TARGET,
                        added_noise_model=False)
        values = args[1:] if operation.startswith('sample') else args[2:]
        parameters = np.asarray(values, dtype=float).reshape(-1).tolist()
        data.update(parameters=parameters, parameter_count=len(parameters))
        _append(log_path, data)

    class _Logged:
        def __init__(self, result, start, name, args):
            self.result, self.start, self.name, self.args, self.logged = result, start, name, args, False

        def get(self):
            out = self.result.get()
            if not self.logged:
                row(self.name, self.start, self.args)
                self.logged = True
            return out

    def wrap(name, logged_name):
        original = getattr(module, name)

        def call(*args, **kwargs):
            if shots is None:
                if kwargs.get('shots_count', -1) != -1:
                    raise RuntimeError('Finite-shot observe forbidden in analytic campaign')
                if kwargs.get('noise_model') is not None:
                    raise RuntimeError('Noise model forbidden in analytic campaign')
            kwargs['shots_count'] = -1 if shots is None else int(shots)
            start = time.perf_counter()
            result = original(*args, **kwargs)
            if name.endswith('_async'):
                return _Logged(result, start, logged_name, args)
            row(logged_name, start, args)
            return result

        setattr(module, name, call)

    wrapped = [('observe', 'observe'), ('observe_async', 'observe_async_until_get')]
    if shots is not None:
        wrapped = [('observe', 'observe'), ('observe_async', 'observe_async'),
                   ('sample', 'sample'), ('sample_async', 'sample_async')]
    for name, logged in wrapped:
        wrap(name, logged)

    def forbidden(message):
        def call(*args, **kwargs):
            raise RuntimeError(message)
        return call

    if shots is None:
        message = 'Sampling/statevector readout forbidden: this campaign records analytic optimization only'
        for name in ('sample', 'sample_async', 'get_state'):
            setattr(module, name, forbidden(message))
    else:
        setattr(module, 'get_state', forbidden('Dense state extraction is disabled; use sampled readout'))
    return [name for name, _ in wrapped]
