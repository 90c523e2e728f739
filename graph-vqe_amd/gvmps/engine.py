"""Host mirror of the batched MPS engine over the C ABI.

`Problem` = one circuit tape + one Hamiltonian compiled by the native planner;
`Batch` = instances of one or more problems evaluated by single launches.
Device buffers are torch tensors on the current HIP device; the stream is
torch's current stream.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

# reference defaults: CUDAQ_MPS_MAX_BOND = 64, cutoffs 1e-5
# (test/backend/source/simulator_mps_dist.cu:363-366, class defaults)
DEFAULT_MAX_BOND = 64
DEFAULT_CUTOFF = 1e-5


def encode_ops(ops):
    """Tape -> gvmps_op array.  An op is (name, qubits, ref) where ref is None
    (fixed gate), an int parameter index, a float constant angle, or a tuple
    (param_index, scale, offset) for affine parameter expressions."""
    arr = (_lib.Op * max(1, len(ops)))()
    for k, (name, qubits, ref) in enumerate(ops):
        name = name.lower()
        if name not in _lib.GATE:
            raise RuntimeError(f"tensornet-mps: unsupported gate '{name}'")
        o = arr[k]
        o.kind = _lib.GATE[name]
        o.q0 = int(qubits[0])
        o.q1 = int(qubits[1]) if len(qubits) > 1 else -1
        o.param, o.scale, o.offset = -1, 0.0, 0.0
        if name in ('rx', 'ry', 'rz'):
            if isinstance(ref, (int, np.integer)):
                o.param, o.scale = int(ref), 1.0
            elif isinstance(ref, tuple):
                o.param, o.scale, o.offset = int(ref[0]), float(ref[1]), float(ref[2])
            else:
                o.offset = float(ref)
    return arr


class Problem:
    def __init__(self, n_qubits, ops, n_params, words, coeffs, max_bond=DEFAULT_MAX_BOND,
                 abs_cut=DEFAULT_CUTOFF, rel_cut=DEFAULT_CUTOFF):
        self.n = int(n_qubits)
        self.n_params = int(n_params)
        self.words = list(words)
        c = np.asarray(coeffs, dtype=complex).reshape(-1)
        if len(c) != len(self.words):
            raise ValueError("words/coeffs length mismatch")
        for w in self.words:
            if len(w) != self.n:
                raise ValueError("Invalid analytic Pauli width")
        self.coeffs = c
        cbuf = np.ascontiguousarray(np.stack([c.real, c.imag], axis=1).reshape(-1), dtype=np.float64)
        wbuf = ''.join(self.words).encode()
        opa = encode_ops(ops)
        h = ctypes.c_void_p()
        check(lib().gvmps_problem_create(ctypes.byref(h), self.n, opa, len(ops), self.n_params, int(max_bond),
                                         float(abs_cut), float(rel_cut), wbuf,
                                         cbuf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(self.words)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def info(self):
        inf = _lib.ProblemInfo()
        check(lib().gvmps_problem_get_info(self._h, ctypes.byref(inf)))
        return {name: getattr(inf, name) for name, _ in inf._fields_}

    def generations(self, detail=False):
        """[(chains, expansion steps)] per chain generation; with detail,
        (chains, steps, term ends, environment stores)."""
        g = self.info()['n_generations']
        out = (ctypes.c_int * (4 * max(g, 1)))()
        check(lib().gvmps_problem_get_generations(self._h, out, g))
        k = 4 if detail else 2
        return [tuple(out[4 * i + j] for j in range(k)) for i in range(g)]

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.gvmps_problem_destroy(h)
            self._h = None


class CapturedEvaluation:
    """A captured evaluate() sequence and the device buffers it is bound to."""

    def __init__(self, batch, params, seeds, energies, bits, packed):
        self.batch, self.params, self.seeds = batch, params, seeds
        self.energies, self.bits, self.packed = energies, bits, packed

    def replay(self):
        check(lib().gvmps_batch_replay(self.batch._h, _stream_ptr(self.batch.torch)))
        return self.energies, self.bits, self.packed


def _stream_ptr(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Batch:
    """counts[j] instances of problems[j]; parameters are concatenated in order."""

    def __init__(self, problems, counts, device=None):
        import torch
        self.torch = torch
        self.problems = list(problems)
        self.counts = [int(c) for c in counts]
        self.device = torch.cuda.current_device() if device is None else int(device)
        harr = (ctypes.c_void_p * len(self.problems))(*[p.handle.value for p in self.problems])
        carr = (ctypes.c_int * len(self.counts))(*self.counts)
        h = ctypes.c_void_p()
        check(lib().gvmps_batch_create(ctypes.byref(h), harr, carr, len(self.problems), self.device))
        self._h = h
        self.size = sum(self.counts)
        self.n_params_total = sum(p.n_params * c for p, c in zip(self.problems, self.counts))
        self.n_terms_total = sum(len(p.words) * c for p, c in zip(self.problems, self.counts))

    def _params(self, params):
        t = self.torch
        x = t.as_tensor(params, dtype=t.float64, device=f'cuda:{self.device}').reshape(-1).contiguous()
        if x.numel() != self.n_params_total:
            raise ValueError(f"expected {self.n_params_total} parameters, got {x.numel()}")
        return x

    def energies(self, params, terms=False):
        """Analytic energies (and optionally per-term <P_k> as complex)."""
        t = self.torch
        x = self._params(params)
        out = t.empty(self.size, dtype=t.float64, device=x.device)
        tv = t.empty(2 * self.n_terms_total, dtype=t.float64, device=x.device) if terms else None
        check(lib().gvmps_batch_energy(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(tv.data_ptr()) if terms else None, _stream_ptr(t)))
        if terms:
            return out, t.view_as_complex(tv.view(-1, 2))
        return out

    def value_grad(self, params):
        """(energies, dE/dx) of every instance by one reverse pass over the
        bond-doubling state (reverse_mps.py:84-99).  Gradients are returned
        flat, n_params per instance in batch order."""
        t = self.torch
        x = self._params(params)
        out = t.empty(self.size, dtype=t.float64, device=x.device)
        g = t.empty(max(self.n_params_total, 1), dtype=t.float64, device=x.device)
        check(lib().gvmps_batch_value_grad(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                           ctypes.c_void_p(g.data_ptr()), _stream_ptr(t)))
        return out, g[:self.n_params_total]

    def set_launch(self, waves_per_group=4, order=1):
        """Chain-kernel launch shape (results do not depend on it): waves per
        workgroup, and item order 0 (instance-major), 1 (longest chains
        first within each XCD's share of the instance-major items).  The
        library has no persistent forest launch; other orders are
        rejected."""
        check(lib().gvmps_batch_set_launch(self._h, int(waves_per_group), int(order)))
        self.launch = (int(waves_per_group), int(order))

    def set_sparsity(self, on=True):
        """Real energy forests skip exact-zero products of the sites' zero
        structure (default on); energies are the same either way.  The real
        gradient's forward chains use it too (suffix chains then contract
        their rows in even-then-odd order), so gradients with it off agree
        to rounding, not bit for bit."""
        check(lib().gvmps_batch_set_sparsity(self._h, int(bool(on))))

    def chain_flops(self):
        """FP64 flops the forest-chain launches of the last energy call executed
        (the products each step issued, zero-structure skips counted)."""
        out = ctypes.c_double()
        check(lib().gvmps_batch_chain_flops(self._h, ctypes.byref(out)))
        return out.value

    def grad_flops(self):
        """FP64 flops the chain launches of the last value_grad call executed
        (forward and backward; the forward's zero-structure skips counted)."""
        out = ctypes.c_double()
        check(lib().gvmps_batch_grad_flops(self._h, ctypes.byref(out)))
        return out.value

    def set_checkpoint(self, stride):
        """value_grad keeps every stride-th step environment of a chain and
        recomputes the rest in the backward pass (same results)."""
        check(lib().gvmps_batch_set_checkpoint(self._h, int(stride)))

    def set_prep(self, mode):
        """'svd' (gates + the plugin's exact-bond split: a decomposition
        wherever the rank rule can cut, the exact split Theta = I Theta or
        Theta I where no gate of the problem is cut), 'factorize' (a
        decomposition at every gate) or 'doubling' (exact CNOT bond doubling,
        no SVD; needs a doubling-compatible tape)."""
        m = {'svd': 0, 'doubling': 1, 'factorize': 2}[mode]
        check(lib().gvmps_batch_set_prep(self._h, m))
        self.prep = mode

    def set_arith(self, mode):
        """'auto' (real-arithmetic forests when every problem has real gates
        and no Y words) or 'complex' (complex kernels always)."""
        m = {'auto': 0, 'complex': 1}[mode]
        check(lib().gvmps_batch_set_arith(self._h, m))
        self.arith = mode

    def _eval_buffers(self, shots):
        t = self.torch
        dev = f'cuda:{self.device}'
        nmax = max(p.n for p in self.problems)
        e = t.empty(self.size, dtype=t.float64, device=dev)
        if shots > 0:
            bits = t.empty((self.size, int(shots), nmax), dtype=t.uint8, device=dev)
            packed = t.empty((self.size, int(shots), 2), dtype=t.int64, device=dev)
        else:
            bits = packed = None
        return e, bits, packed

    def evaluate(self, params, shots=0, seeds=None):
        """Energies and (shots > 0) samples from one state preparation:
        (energies, bits[site order], packed)."""
        t = self.torch
        x = self._params(params)
        e, bits, packed = self._eval_buffers(shots)
        sd = None
        if shots > 0:
            sd = t.as_tensor(self._seeds(seeds).view(np.int32), device=x.device)  # same 32 bits
        ptr = lambda v: ctypes.c_void_p(v.data_ptr()) if v is not None else None
        check(lib().gvmps_batch_evaluate(self._h, ptr(x), int(shots), ptr(sd), ptr(e), ptr(bits), ptr(packed),
                                         _stream_ptr(t)))
        return e, bits, packed

    def capture(self, params, shots=0, seeds=None):
        """Record evaluate() into a HIP graph bound to device buffers; returns
        a CapturedEvaluation whose replay() re-runs it (update .params /
        .seeds in place between replays)."""
        t = self.torch
        x = self._params(params).clone()
        e, bits, packed = self._eval_buffers(shots)
        sd = None
        if shots > 0:
            sd = t.as_tensor(self._seeds(seeds).view(np.int32), device=x.device)  # same 32 bits
        t.cuda.synchronize()
        ptr = lambda v: ctypes.c_void_p(v.data_ptr()) if v is not None else None
        check(lib().gvmps_batch_capture(self._h, ptr(x), int(shots), ptr(sd), ptr(e), ptr(bits), ptr(packed)))
        return CapturedEvaluation(self, x, sd, e, bits, packed)

    def _seeds(self, seeds):
        s = np.asarray(seeds, dtype=np.uint32).reshape(-1)
        if s.size == 1 and self.size > 1:
            s = np.full(self.size, s[0], dtype=np.uint32)
        if s.size != self.size:
            raise ValueError("one seed per instance is required")
        return np.ascontiguousarray(s)

    def sample(self, params, shots, seeds):
        """Perfect sampling.  Returns (bits, packed): bits[inst, shot, site] in
        {0,1} (site order, row stride = largest n of the batch) and packed
        uint64 words [inst, shot, 2] (bit s = site s), both on the device."""
        t = self.torch
        x = self._params(params)
        s = self._seeds(seeds)
        nmax = max(p.n for p in self.problems)
        bits = t.empty((self.size, int(shots), nmax), dtype=t.uint8, device=x.device)
        packed = t.empty((self.size, int(shots), 2), dtype=t.int64, device=x.device)
        check(lib().gvmps_batch_sample(self._h, ctypes.c_void_p(x.data_ptr()), int(shots),
                                       s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                       ctypes.c_void_p(bits.data_ptr()), ctypes.c_void_p(packed.data_ptr()),
                                       _stream_ptr(t)))
        return bits, packed

    def shot_energies(self, params, shots, seeds, terms=False):
        """Finite-shot energies from one shared sample set per instance."""
        t = self.torch
        x = self._params(params)
        s = self._seeds(seeds)
        out = t.empty(self.size, dtype=t.float64, device=x.device)
        tv = t.empty(2 * self.n_terms_total, dtype=t.float64, device=x.device) if terms else None
        check(lib().gvmps_batch_shot_energy(self._h, ctypes.c_void_p(x.data_ptr()), int(shots),
                                            s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                            ctypes.c_void_p(out.data_ptr()),
                                            ctypes.c_void_p(tv.data_ptr()) if terms else None, _stream_ptr(t)))
        if terms:
            return out, t.view_as_complex(tv.view(-1, 2))
        return out

    def prepare(self, params):
        t = self.torch
        x = self._params(params)
        check(lib().gvmps_batch_prepare(self._h, ctypes.c_void_p(x.data_ptr()), _stream_ptr(t)))

    def set_timing(self, enable=True):
        check(lib().gvmps_batch_set_timing(self._h, 1 if enable else 0))

    def timing(self):
        """Accumulated per-phase device milliseconds and the number of calls."""
        ms = (ctypes.c_double * 8)()
        calls = ctypes.c_longlong()
        check(lib().gvmps_batch_get_timing(self._h, ms, ctypes.byref(calls)))
        return list(ms), calls.value

    def prep_stats(self):
        """Per instance: [sweeps, build, qr, jacobi, tail clocks/16, ...] of the last preparation."""
        out = (ctypes.c_int * (8 * self.size))()
        check(lib().gvmps_batch_prep_stats(self._h, out))
        return np.array(out, dtype=np.int64).reshape(self.size, 8)

    def read_sites(self, inst):
        """Prepared site tensors of instance `inst` as numpy (n, 16, 2, 16) + bonds."""
        j, k = 0, inst
        while k >= self.counts[j]:
            k -= self.counts[j]
            j += 1
        n = self.problems[j].n
        buf = np.zeros(n * 16 * 2 * 16 * 2, dtype=np.float64)
        bonds = np.zeros(n + 1, dtype=np.int32)
        check(lib().gvmps_batch_read_sites(self._h, int(inst), buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           bonds.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        z = buf.view(np.complex128).reshape(n, 16, 2, 16)
        return z, bonds

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.gvmps_batch_destroy(h)
            self._h = None
