"""CUDA-Q-compatible Python call surface of the MPS backend.

The reference drives its custom NVQIR plugin (target
``tensornet-mps-mqpu-shots``, simulator_mps_dist.cu:1664-1667) through a small
part of CUDA-Q: ``make_kernel`` builders with ``qalloc`` and rx/ry/rz/h/x/y/z/
s/sdg/cx gates (Qiskit_VQE/independent_vqe.py:160-174), ``spin`` operators,
``observe(kernel, H, *params, shots_count=...)``, ``observe_async(...).get()``
(Qiskit_VQE/subspace_vqe.py:580-585) and ``sample`` (test/tools/
run_case.py:107-108).  This module offers the same calls over the gvmps HIP
library, plus the reverse-gradient factory of test/tools/reverse_mps.py
(``make_problem`` / ``for_objective``), a ``torch.autograd.Function`` over the
adjoint kernel, and batched entry points (``observe_batch``,
``value_grad_batch``) that evaluate many parameter vectors in one launch.

Semantics kept from the plugin:
  * qubit q is MPS site n-1-q (:386-387); words and bitstrings use qubit order;
  * shots_count <= 0 (or the SIZE_MAX sentinel) is analytic (:1438-1444);
  * finite-shot observe of an all-I/Z Hamiltonian shares one sample set for
    every word (runSharedDiagonalShotBatch, :1157-1171, 1502-1514); with X/Y
    words present the Z words are sampled independently per word and the
    others are exact (evaluateHbatchTermBatch, :641-697);
  * sample() with shots <= 0 draws 1000 shots (:1592); counts are keyed by
    the register's qubits in order, q0 leftmost (:1603-1611);
  * every finite-shot observe or sample consumes one value of the per-thread
    ``sampleSeed`` counter (1234567, 1235568, ...; :214, 1495, 1601);
  * engine configuration from CUDAQ_MPS_MAX_BOND (default 64),
    CUDAQ_MPS_ABS_CUTOFF and CUDAQ_MPS_RELATIVE_CUTOFF (default 1e-5)
    (:363-366); CUDAQ_MPS_ANALYTIC_MODE in {shared, batch, reference}
    (:648-657) is accepted -- every mode evaluates the same <P_k> here;
  * fewer than 2 qubits, or >2-qubit gates, raise as the plugin does.
Noise models and multi-GPU execution are outside this backend.
"""
import os
import threading

import numpy as np

from .engine import DEFAULT_CUTOFF, DEFAULT_MAX_BOND, Batch, Problem
from .seeds import SeedCounter, shot_shard_seed, word_seed

TARGET = 'tensornet-mps-mqpu-shots'
_SIZE_MAX = (1 << 64) - 1


# ----------------------------------------------------------------- kernels
class Parameter:
    """Kernel argument expression, affine in one argument: scale * x[index] + offset."""
    __slots__ = ('index', 'scale', 'offset')

    def __init__(self, index, scale=1.0, offset=0.0):
        self.index, self.scale, self.offset = index, float(scale), float(offset)

    def _combine(self, other, sign):
        if isinstance(other, Parameter):
            if other.index != self.index and other.scale != 0 and self.scale != 0:
                raise TypeError('tensornet-mps: rotation angles must be affine in one kernel argument')
            idx = self.index if self.scale != 0 else other.index
            return Parameter(idx, self.scale + sign * other.scale, self.offset + sign * other.offset)
        return Parameter(self.index, self.scale, self.offset + sign * float(other))

    def __add__(self, other): return self._combine(other, 1.0)
    def __radd__(self, other): return self._combine(other, 1.0)
    def __sub__(self, other): return self._combine(other, -1.0)
    def __rsub__(self, other): return (-self)._combine(other, 1.0)
    def __neg__(self): return Parameter(self.index, -self.scale, -self.offset)

    def __mul__(self, other):
        if isinstance(other, Parameter):
            raise TypeError('tensornet-mps: rotation angles must be affine in one kernel argument')
        return Parameter(self.index, self.scale * float(other), self.offset * float(other))

    __rmul__ = __mul__

    def __truediv__(self, other): return self * (1.0 / float(other))

    def ref(self):
        return (self.index, self.scale, self.offset)


class ParameterList:
    """A ``list[float]`` kernel argument: element i is parameter i."""

    def __init__(self):
        self.size = 0

    def __getitem__(self, i):
        i = int(i)
        if i < 0:
            raise IndexError('negative parameter index')
        self.size = max(self.size, i + 1)
        return Parameter(i)


class Qubit:
    __slots__ = ('index',)

    def __init__(self, index):
        self.index = index


class QuakeRegister(list):
    pass


class Kernel:
    """Circuit builder recording a tape of (name, qubits, angle ref)."""

    def __init__(self, arg_types):
        self.arg_types = list(arg_types)
        self.arguments = []
        for i, t in enumerate(self.arg_types):
            self.arguments.append(ParameterList() if t is list else Parameter(i))
        self.width = 0
        self.operations = []
        self.version = 0

    @property
    def argument_count(self):
        return len(self.arg_types)

    @property
    def name(self):
        return f'gvmps_kernel_{id(self):x}'

    def qalloc(self, count=1):
        if self.width:
            raise ValueError('tensornet-mps kernels allocate one register')
        self.width = int(count)
        return QuakeRegister(Qubit(i) for i in range(self.width))

    def _qubit(self, q):
        idx = q.index if isinstance(q, Qubit) else int(q)
        if idx < 0 or idx >= self.width:
            raise RuntimeError('qubit index out of range')
        return idx

    def _emit(self, name, qubits, ref):
        self.operations.append((name, tuple(qubits), ref))
        self.version += 1

    def _rotation(self, name, angle, qubit):
        if isinstance(angle, Parameter):
            ref = angle.ref()
        else:
            ref = float(angle)
        self._emit(name, (self._qubit(qubit),), ref)

    def rx(self, angle, qubit): self._rotation('rx', angle, qubit)
    def ry(self, angle, qubit): self._rotation('ry', angle, qubit)
    def rz(self, angle, qubit): self._rotation('rz', angle, qubit)
    def h(self, qubit): self._emit('h', (self._qubit(qubit),), None)
    def x(self, qubit): self._emit('x', (self._qubit(qubit),), None)
    def y(self, qubit): self._emit('y', (self._qubit(qubit),), None)
    def z(self, qubit): self._emit('z', (self._qubit(qubit),), None)
    def s(self, qubit): self._emit('s', (self._qubit(qubit),), None)
    def sdg(self, qubit): self._emit('sdg', (self._qubit(qubit),), None)

    def cx(self, control, target):
        c, t = self._qubit(control), self._qubit(target)
        if c == t:
            raise RuntimeError('2q gate on same qubit')
        self._emit('cx', (c, t), None)

    def swap(self, a, b):
        self._emit('swap', (self._qubit(a), self._qubit(b)), None)

    def mz(self, *args, **kwargs):
        """Measurement marker: sample() reads every qubit of the register."""

    def compile(self):
        """Nothing to JIT: the tape is compiled per (kernel, Hamiltonian)."""

    def n_params(self):
        n = 0
        for a in self.arguments:
            n = max(n, a.size if isinstance(a, ParameterList) else 0)
        scalars = sum(1 for a in self.arguments if isinstance(a, Parameter))
        return scalars if scalars else n

    def tape(self):
        return list(self.operations)


def make_kernel(*types):
    """``kernel, *params = make_kernel(float, float, ...)`` or ``make_kernel(list)``."""
    for t in types:
        if t is not float and t is not list:
            raise TypeError('tensornet-mps kernels take float or list[float] arguments')
    if types.count(list) and len(types) > 1:
        raise TypeError('a list[float] kernel takes exactly one argument')
    k = Kernel(types)
    return (k, *k.arguments) if types else k


# ----------------------------------------------------------- spin operators
_PAULI_MUL = {  # (a, b) -> (phase, product)
    ('X', 'X'): (1, 'I'), ('Y', 'Y'): (1, 'I'), ('Z', 'Z'): (1, 'I'),
    ('X', 'Y'): (1j, 'Z'), ('Y', 'X'): (-1j, 'Z'), ('Y', 'Z'): (1j, 'X'),
    ('Z', 'Y'): (-1j, 'X'), ('Z', 'X'): (1j, 'Y'), ('X', 'Z'): (-1j, 'Y'),
}


def _mul_strings(a, b):
    ops = dict(a)
    phase = 1 + 0j
    for q, p in b:
        if q in ops:
            ph, r = _PAULI_MUL[(ops[q], p)]
            phase *= ph
            if r == 'I':
                del ops[q]
            else:
                ops[q] = r
        else:
            ops[q] = p
    return phase, tuple(sorted(ops.items()))


class SpinTerm:
    def __init__(self, paulis, coeff, width):
        self.paulis, self.coeff, self.width = paulis, complex(coeff), width

    def get_coefficient(self):
        return self.coeff

    def to_string(self, with_coefficient=True):
        w = ['I'] * self.width
        for q, p in self.paulis:
            w[q] = p
        word = ''.join(w)
        return f'{self.coeff}{word}' if with_coefficient else word

    def is_identity(self):
        return not self.paulis


class SpinOperator:
    """Sum of Pauli strings with complex coefficients (qubit-indexed)."""

    def __init__(self, terms=None):
        self.terms = dict(terms or {})

    @staticmethod
    def _lift(v):
        if isinstance(v, SpinOperator):
            return v
        return SpinOperator({(): complex(v)})

    def __add__(self, other):
        o = self._lift(other)
        t = dict(self.terms)
        for k, c in o.terms.items():
            t[k] = t.get(k, 0) + c
        return SpinOperator(t)

    __radd__ = __add__

    def __neg__(self):
        return SpinOperator({k: -c for k, c in self.terms.items()})

    def __sub__(self, other): return self + (-self._lift(other))
    def __rsub__(self, other): return self._lift(other) + (-self)

    def __mul__(self, other):
        if not isinstance(other, SpinOperator):
            return SpinOperator({k: c * complex(other) for k, c in self.terms.items()})
        t = {}
        for ka, ca in self.terms.items():
            for kb, cb in other.terms.items():
                ph, k = _mul_strings(ka, kb)
                t[k] = t.get(k, 0) + ph * ca * cb
        return SpinOperator(t)

    def __rmul__(self, other):
        return self * other

    def get_qubit_count(self):
        return 1 + max((q for k in self.terms for q, _ in k), default=-1)

    def get_term_count(self):
        return len(self.terms)

    @property
    def term_count(self):
        return len(self.terms)

    def for_each_term(self, fn):
        w = self.get_qubit_count()
        for k, c in self.terms.items():
            fn(SpinTerm(k, c, w))

    def words(self, width):
        """(qubit-order words of length `width`, complex coefficients)."""
        words, coeffs = [], []
        for k, c in self.terms.items():
            w = ['I'] * width
            for q, p in k:
                if q >= width:
                    raise RuntimeError(f'spin operator acts on qubit {q} outside the {width}-qubit register')
                w[q] = p
            words.append(''.join(w))
            coeffs.append(c)
        return words, np.array(coeffs, dtype=complex)

    def __str__(self):
        return ' + '.join(f'{c}*{SpinTerm(k, c, self.get_qubit_count()).to_string(False)}'
                          for k, c in self.terms.items())


class _Spin:
    @staticmethod
    def i(q): return SpinOperator({(): 1.0})

    @staticmethod
    def x(q): return SpinOperator({((int(q), 'X'),): 1.0})

    @staticmethod
    def y(q): return SpinOperator({((int(q), 'Y'),): 1.0})

    @staticmethod
    def z(q): return SpinOperator({((int(q), 'Z'),): 1.0})


spin = _Spin()


def qiskit_sparsepauliop_to_cudaq_spin(op, reverse_qubit_order=True, num_qubits=None):
    """SpinOperator of a Qiskit ``SparsePauliOp`` (Qiskit_VQE/cudaq.py:25-47).

    ``op`` is anything with ``to_list()`` (and ``num_qubits``), such as a
    SparsePauliOp, or a plain list of ``(label, coeff)`` pairs, which is what
    ``SparsePauliOp.to_list()`` returns; qiskit itself is not needed.  As in
    the reference:
      * labels are Qiskit strings, leftmost character = highest qubit, so with
        ``reverse_qubit_order`` character idx acts on qubit n-1-idx (n = the
        operator's qubit count, else the label length); without it, on idx;
      * zero coefficients are skipped and only the real part of the others is
        kept;
      * an all-I label adds a constant; repeated labels add up.
    Labels must hold only I, X, Y, Z.
    """
    pairs = list(op.to_list()) if hasattr(op, 'to_list') else list(op)
    if num_qubits is None:
        num_qubits = getattr(op, 'num_qubits', None)
    H = SpinOperator()
    for label, coeff in pairs:
        label = str(label)
        n = len(label) if num_qubits is None else int(num_qubits)
        if len(label) != n:
            raise ValueError(f'Pauli label {label!r} has {len(label)} characters for {n} qubits')
        if coeff == 0:
            continue
        term = SpinOperator({(): 1.0})
        for idx, ch in enumerate(label):
            q = (n - 1 - idx) if reverse_qubit_order else idx
            if ch == 'I':
                continue
            if ch not in 'XYZ':
                raise ValueError(f'Pauli label {label!r} has {ch!r}; expected I, X, Y or Z')
            term = term * getattr(spin, ch.lower())(q)
        H = H + float(complex(coeff).real) * term
    return H


class _Parallel:
    thread = 'thread'
    mpi = 'mpi'


parallel = _Parallel()


# ------------------------------------------------------------------ targets
class Target:
    def __init__(self, name, options=None):
        self.name, self.options = name, dict(options or {})

    def num_qpus(self):
        return 1


_target = Target(TARGET)


def set_target(name, **options):
    global _target
    _target = Target(str(name), options)


def get_target(name=None):
    return _target if name is None else Target(name)


def num_available_gpus():
    import torch
    return torch.cuda.device_count()


def set_random_seed(seed):
    """CUDA-Q's simulator seed.  The MPS plugin draws its shots from its own
    ``sampleSeed`` counter and ignores this seed (simulator_mps_dist.cu:214)."""


def configure_full_bond(num_qubits, stages=None):
    """Exact bond bound of the campaign ansatz with zero cutoffs
    (Qiskit_VQE/mps_config.py:8-43): each CNOT crossing a cut at most doubles
    its rank, so linear / reverse_linear chains give 2^min(n//2, reps),
    pairwise 2^min(n//2, ceil(reps/2)), anything else 2^(n//2)."""
    if num_qubits < 2:
        raise ValueError(f'{TARGET} requires at least two qubits')
    exponents = []
    for stage in stages or [{'entanglement': None}]:
        exponent = num_qubits // 2
        pattern, reps = stage.get('entanglement'), stage.get('reps')
        if isinstance(pattern, str) and reps is not None:
            reps = int(reps)
            if reps < 0:
                raise ValueError('reps must be nonnegative')
            if pattern.lower() in ('linear', 'reverse_linear'):
                exponent = min(exponent, reps)
            elif pattern.lower() == 'pairwise':
                exponent = min(exponent, (reps + 1) // 2)
        exponents.append(exponent)
    exponent = max(exponents)
    bond = (1 << exponent) if exponent < 31 else 2 ** 31 - 1
    os.environ['CUDAQ_MPS_MAX_BOND'] = str(bond)
    os.environ['CUDAQ_MPS_ABS_CUTOFF'] = '0'
    os.environ['CUDAQ_MPS_RELATIVE_CUTOFF'] = '0'
    return bond


# ------------------------------------------------------------------ engine
def engine_config():
    max_bond = max(1, int(os.environ.get('CUDAQ_MPS_MAX_BOND', DEFAULT_MAX_BOND)))
    abs_cut = float(os.environ.get('CUDAQ_MPS_ABS_CUTOFF', DEFAULT_CUTOFF))
    rel_cut = float(os.environ.get('CUDAQ_MPS_RELATIVE_CUTOFF', DEFAULT_CUTOFF))
    mode = os.environ.get('CUDAQ_MPS_ANALYTIC_MODE', 'shared')
    if mode not in ('shared', 'batch', 'reference'):
        mode = 'shared'  # unknown values fall through to the shared path (:648-657)
    return max_bond, abs_cut, rel_cut


class _Engine:
    """Compiled problems and device batches, cached per thread (the plugin is
    one thread_local simulator per calling thread)."""

    def __init__(self):
        self.tls = threading.local()
        self.seeds = SeedCounter()

    def _caches(self):
        if not hasattr(self.tls, 'problems'):
            self.tls.problems, self.tls.batches = {}, {}
        return self.tls.problems, self.tls.batches

    def problem(self, kernel, words, coeffs, config=None, n_params=None):
        problems, _ = self._caches()
        config = config or engine_config()
        n = kernel.width
        if n < 2:
            raise RuntimeError('tensornet-mps needs >= 2 qubits')
        P = kernel.n_params() if n_params is None else int(n_params)
        key = (id(kernel), kernel.version, P, tuple(words), np.asarray(coeffs).tobytes(), config)
        hit = problems.get(key)
        if hit is not None and hit[0] is kernel:
            return hit[1]
        p = Problem(n, kernel.tape(), P, words, coeffs, max_bond=config[0], abs_cut=config[1], rel_cut=config[2])
        if len(problems) > 64:
            problems.clear()
        problems[key] = (kernel, p)
        return p

    def batch(self, problem, count):
        _, batches = self._caches()
        key = (id(problem), int(count))
        b = batches.get(key)
        if b is None or b.problems[0] is not problem:
            if len(batches) > 16:
                batches.clear()
            b = Batch([problem], [int(count)])
            batches[key] = b
        return b


_engine = _Engine()


def _params(kernel, args):
    """Kernel arguments -> flat float64 vector.  A list[float] kernel takes one
    sequence; a float kernel takes its floats spread or as one sequence."""
    if len(args) == 1 and np.ndim(args[0]) > 0:
        x = np.asarray(args[0], dtype=float).reshape(-1)
    else:
        x = np.asarray([float(a) for a in args], dtype=float)
    if kernel.arg_types == [list]:
        if len(x) < kernel.n_params():
            raise RuntimeError(f'kernel uses {kernel.n_params()} parameters, got {len(x)}')
    elif len(x) != kernel.argument_count:
        raise RuntimeError(f'kernel expects {kernel.argument_count} arguments, got {len(x)}')
    return x


def _is_finite_shots(shots_count):
    return shots_count is not None and 0 < int(shots_count) != _SIZE_MAX


class ObserveResult:
    """``expectation()`` = Re sum_k c_k <P_k>; per-term raw data are <P_k>
    (simulator_mps_dist.cu:1467-1481)."""

    def __init__(self, energy, words, coeffs, term_values, hamiltonian, shots=-1):
        self._energy = float(energy)
        self.words, self.coeffs = words, coeffs
        self._terms = term_values  # array, or a callable evaluated on first use
        self.hamiltonian = hamiltonian
        self.shots = shots

    @property
    def term_values(self):
        if callable(self._terms):
            self._terms = self._terms()
        return np.asarray(self._terms, dtype=float)

    def expectation(self, term=None):
        if term is None:
            return self._energy
        if isinstance(term, SpinOperator):
            if term.get_term_count() != 1:
                raise ValueError('expectation(term) takes a single Pauli term')
            w = term.words(len(self.words[0]) if self.words else 0)[0][0]
        else:
            w = str(term)
        for word, v in zip(self.words, self.term_values):
            if word == w:
                return float(v)
        raise KeyError(f'term {w} not in the observed Hamiltonian')

    def get_term_values(self):
        return dict(zip(self.words, self.term_values.tolist()))

    def dump(self):
        print(f'<H> = {self._energy}')


class AsyncResult:
    def __init__(self, value):
        self._value = value

    def get(self):
        return self._value


def observe(kernel, hamiltonian, *args, shots_count=-1, execution=None, qpu_id=None, noise_model=None):
    if noise_model is not None:
        raise RuntimeError('noise models are not supported by this backend')
    x = _params(kernel, args)
    H = hamiltonian if isinstance(hamiltonian, SpinOperator) else SpinOperator._lift(hamiltonian)
    words, coeffs = H.words(kernel.width)
    problem = _engine.problem(kernel, words, coeffs, n_params=len(x))
    batch = _engine.batch(problem, 1)
    if not _is_finite_shots(shots_count):
        # the energy comes from the energy-only reduction, as in observe_batch
        # (so batched and sequential calls agree bitwise); per-term <P_k> are
        # evaluated when first asked for
        e = float(batch.energies(x).cpu()[0])
        terms = lambda: batch.energies(x, terms=True)[1].real.cpu().numpy()
        return ObserveResult(e, words, coeffs, terms, H)
    shots = int(shots_count)
    evaluation_seed = _engine.seeds.next()
    diagonal = all(set(w) <= {'I', 'Z'} for w in words)
    if diagonal:
        e, tv = batch.shot_energies(x, shots, [shot_shard_seed(evaluation_seed, 0)], terms=True)
        return ObserveResult(float(e.cpu()[0]), words, coeffs, tv.real.cpu().numpy(), H, shots)
    values = _mixed_shot_values(kernel, problem, x, words, shots, evaluation_seed)
    energy = float(np.real(np.sum(coeffs * values)))
    return ObserveResult(energy, words, coeffs, values, H, shots)


def _mixed_shot_values(kernel, problem, x, words, shots, evaluation_seed):
    """Non-diagonal finite-shot observe: X/Y words exact, every I/Z word
    sampled with its own seed word_seed(evaluation_seed, site word)."""
    import torch
    _, tv = _engine.batch(problem, 1).energies(x, terms=True)
    values = tv.real.cpu().numpy().copy()
    n = kernel.width
    zidx = [k for k, w in enumerate(words) if set(w) <= {'I', 'Z'} and set(w) != {'I'}]
    if zidx:
        site_words = [words[k][::-1] for k in zidx]
        seeds = [word_seed(evaluation_seed, sw) for sw in site_words]
        b = _engine.batch(problem, len(zidx))
        bits, _ = b.sample(np.tile(x, len(zidx)), shots, seeds)
        masks = torch.tensor([[c == 'Z' for c in sw] for sw in site_words], dtype=torch.int32,
                             device=bits.device)
        par = (bits[:, :, :n].to(torch.int32) * masks[:, None, :]).sum(dim=2) & 1
        vals = (1.0 - 2.0 * par.to(torch.float64)).mean(dim=1).cpu().numpy()
        values[zidx] = vals
    return values


def observe_async(kernel, hamiltonian, *args, qpu_id=0, shots_count=-1, noise_model=None):
    return AsyncResult(observe(kernel, hamiltonian, *args, shots_count=shots_count, noise_model=noise_model))


class SampleResult(dict):
    """Counts keyed by bitstrings over the register's qubits, q0 leftmost."""

    def most_probable(self):
        return max(self.items(), key=lambda kv: kv[1])[0] if self else None

    def count(self, key):
        return self.get(key, 0)

    def probability(self, key):
        total = sum(self.values())
        return self.get(key, 0) / total if total else 0.0

    def dump(self):
        print('{ ' + ' '.join(f'{k}:{v}' for k, v in sorted(self.items())) + ' }')


def sample(kernel, *args, shots_count=1000, noise_model=None):
    if noise_model is not None:
        raise RuntimeError('noise models are not supported by this backend')
    shots = int(shots_count) if shots_count and int(shots_count) > 0 else 1000
    x = _params(kernel, args)
    n = kernel.width
    problem = _engine.problem(kernel, ['Z' * n], [0.0], n_params=len(x))
    seed = shot_shard_seed(_engine.seeds.next(), 0)
    bits, _ = _engine.batch(problem, 1).sample(x, shots, [seed])
    b = bits[0, :, :n].flip(1).cpu().numpy()  # site order -> qubit order
    keys, counts = np.unique(b, axis=0, return_counts=True)
    return SampleResult({''.join('1' if v else '0' for v in k): int(c) for k, c in zip(keys, counts)})


def sample_async(kernel, *args, shots_count=1000, qpu_id=0):
    return AsyncResult(sample(kernel, *args, shots_count=shots_count))


def get_state(*args, **kwargs):
    raise NotImplementedError('tensornet-mps: state-vector retrieval is not provided by this backend')


# --------------------------------------------------------- batched calls
def observe_batch(kernel, hamiltonian, params, shots_count=-1, seeds=None):
    """Energies of every row of `params` in one batched launch (numpy array).
    Finite shots need an all-I/Z Hamiltonian; seeds default to consecutive
    ``sampleSeed`` values, one per row, as sequential observe calls would use."""
    X = np.atleast_2d(np.asarray(params, dtype=float))
    H = hamiltonian if isinstance(hamiltonian, SpinOperator) else SpinOperator._lift(hamiltonian)
    words, coeffs = H.words(kernel.width)
    problem = _engine.problem(kernel, words, coeffs, n_params=X.shape[1])
    batch = _engine.batch(problem, X.shape[0])
    if not _is_finite_shots(shots_count):
        return batch.energies(X).cpu().numpy()
    if seeds is None:
        seeds = [shot_shard_seed(_engine.seeds.next(), 0) for _ in range(X.shape[0])]
    return batch.shot_energies(X, int(shots_count), seeds).cpu().numpy()


def value_grad_batch(kernel, hamiltonian, params):
    """(energies, gradients) of every row by one batched reverse pass."""
    X = np.atleast_2d(np.asarray(params, dtype=float))
    H = hamiltonian if isinstance(hamiltonian, SpinOperator) else SpinOperator._lift(hamiltonian)
    words, coeffs = H.words(kernel.width)
    problem = _engine.problem(kernel, words, np.real(coeffs), config=(16, 0.0, 0.0), n_params=X.shape[1])
    e, g = _engine.batch(problem, X.shape[0]).value_grad(X)
    return e.cpu().numpy(), g.cpu().numpy().reshape(X.shape[0], X.shape[1])


# --------------------------------------------------------- reverse gradient
class ReverseProblem:
    """Exact value and gradient of sum_k c_k <P_k> (reverse_mps.py:84-121):
    the adjoint kernel over the bond-doubling state."""
    engine = 'gvmps_reverse'

    def __init__(self, kernel, words, coeffs):
        cf = np.asarray(coeffs, dtype=complex)
        if np.any(np.abs(cf.imag) > 1e-12):
            raise ValueError('Reverse mode requires Hermitian real Pauli coefficients')
        self.kernel, self.words, self.coeffs = kernel, list(words), cf.real.copy()
        self.n_params = kernel.n_params()
        p = self._problem(self.coeffs)
        if not p.info()['grad_ok']:
            raise ValueError('circuit has no bond-doubling reverse gradient within 16-wide bonds')

    def _problem(self, coeffs):
        return _engine.problem(self.kernel, self.words, coeffs, config=(16, 0.0, 0.0), n_params=self.n_params)

    def value_gradient(self, parameters, coefficients=None):
        c = self.coeffs if coefficients is None else np.asarray(coefficients, dtype=float)
        e, g = _engine.batch(self._problem(c), 1).value_grad(np.asarray(parameters, dtype=float))
        return float(e.cpu()[0]), g.cpu().numpy()

    def value_gradient_batch(self, params):
        X = np.atleast_2d(np.asarray(params, dtype=float))
        e, g = _engine.batch(self._problem(self.coeffs), X.shape[0]).value_grad(X)
        return e.cpu().numpy(), g.cpu().numpy().reshape(X.shape[0], -1)

    def values(self, parameters):
        """Unweighted <P_k> in word order (ReverseTapeMPS.values, :186-194)."""
        _, tv = _engine.batch(self._problem(self.coeffs), 1).energies(np.asarray(parameters, dtype=float),
                                                                      terms=True)
        return tv.real.cpu().numpy()

    def energy(self, x):
        """Differentiable energies of x (P,) or (B, P) as a torch tensor."""
        return ReverseEnergy.apply(x, self)


def make_problem(kernel, hamiltonian):
    """reverse_mps.make_problem: None unless `kernel` is a gvmps kernel."""
    if not isinstance(kernel, Kernel):
        return None
    H = hamiltonian if isinstance(hamiltonian, SpinOperator) else SpinOperator._lift(hamiltonian)
    words, coeffs = H.words(kernel.width)
    return ReverseProblem(kernel, words, coeffs)


def for_objective(fun, mode):
    """reverse_mps.for_objective: reverse_factory, then reverse_problem=(kernel, H)."""
    if mode == 'off':
        return None
    if mode not in ('auto', 'hybrid', 'required'):
        raise ValueError('Reverse mode must be off, auto, hybrid or required')
    factory = getattr(fun, 'reverse_factory', None)
    if factory is not None:
        return factory()
    pair = getattr(fun, 'reverse_problem', None)
    owner = getattr(fun, '__self__', None)
    if pair is None and owner is not None and getattr(fun, '__name__', '') == 'cost_func' and \
            hasattr(owner, 'kernel') and hasattr(owner, 'hamiltonian'):
        pair = (owner.kernel, owner.hamiltonian)
    return make_problem(*pair) if pair is not None else None


def _autograd_function():
    import torch

    class _ReverseEnergy(torch.autograd.Function):
        """E(x) with dE/dx from the adjoint kernel; x is (P,) or (B, P) float64."""

        @staticmethod
        def forward(ctx, x, problem):
            X = x.detach().to(torch.float64)
            single = X.dim() == 1
            X2 = X.reshape(1, -1) if single else X
            b = _engine.batch(problem._problem(problem.coeffs), X2.shape[0])
            e, g = b.value_grad(X2.to(f'cuda:{b.device}').contiguous())
            g = g.reshape(X2.shape).to(x.device)
            ctx.save_for_backward(g)
            ctx.single = single
            e = e.to(x.device)
            return e[0] if single else e

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            if ctx.single:
                return grad_out * g[0], None
            return grad_out[:, None] * g, None

    return _ReverseEnergy


class _LazyFunction:
    _fn = None

    def apply(self, *args):
        if _LazyFunction._fn is None:
            _LazyFunction._fn = _autograd_function()
        return _LazyFunction._fn.apply(*args)


ReverseEnergy = _LazyFunction()
