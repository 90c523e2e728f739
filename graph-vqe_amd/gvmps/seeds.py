"""Per-call sampling seeds of the MPS backend.

The plugin keeps one counter per simulator (``sampleSeed``, starting at
1234567, simulator_mps_dist.cu:214) and consumes one value per finite-shot
observe or sample call (:1495, :1601).  The sampling kernel receives
``shot_shard_seed(counter, 0)`` (hbatchSeedForShotShard, :596-608); the
independent per-word estimator of non-diagonal finite-shot observes seeds each
Z word with ``word_seed(counter, word)`` (hbatchSeedForWord, :586-594, word in
site order).  Both are 32-bit integer mixers; the formulas are the
reference's, as they must be for identical bitstrings.
"""
import threading

INITIAL_SAMPLE_SEED = 1234567
_M32 = 0xFFFFFFFF


def shot_shard_seed(evaluation_seed, worker_index=0):
    v = (evaluation_seed + ((0x9E3779B9 * (worker_index + 1)) & _M32)) & _M32
    v ^= v >> 16
    v = (v * 0x7FEB352D) & _M32
    v ^= v >> 15
    v = (v * 0x846CA68B) & _M32
    v ^= v >> 16
    return v


def word_seed(evaluation_seed, site_word):
    h = (2166136261 ^ evaluation_seed) & _M32
    for ch in site_word.encode():
        h ^= ch
        h = (h * 16777619) & _M32
    return h


class SeedCounter:
    """One ``sampleSeed`` per thread: the plugin is a thread_local singleton
    (simulator_mps_dist.cu:1656-1659), so every calling thread has its own."""

    def __init__(self):
        self._tls = threading.local()

    def next(self):
        v = getattr(self._tls, 'value', INITIAL_SAMPLE_SEED)
        self._tls.value = (v + 1) & _M32
        return v

    def reset(self, value=INITIAL_SAMPLE_SEED):
        self._tls.value = value & _M32
