"""Uniform stream and seed mixers of the reference sampler (test infrastructure).

All arithmetic is done on Python ints masked to the stated width, so the
results are bit-exact restatements of the reference's unsigned C arithmetic.
"""
M64 = (1 << 64) - 1
M32 = (1 << 32) - 1


def splitmix64(x):
    """SplitMix64 finalizer.  test/backend/source/sharded_mps.cuh:147-152."""
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def draw_uniform(seed, site, shot):
    """Per-draw uniform in [0,1).  test/backend/source/sharded_mps.cuh:171-174.

    raw = SplitMix64((seed << 32) ^ (site << 20) ^ shot); u = (raw >> 11) * 2^-53.
    """
    raw = splitmix64((((seed & M32) << 32) ^ (site << 20) ^ shot) & M64)
    return (raw >> 11) * (2.0 ** -53)


def shot_shard_seed(evaluation_seed, worker_index=0):
    """hbatchSeedForShotShard.  test/backend/source/simulator_mps_dist.cu:596-608."""
    v = (evaluation_seed + ((0x9E3779B9 * (worker_index + 1)) & M32)) & M32
    v ^= v >> 16
    v = (v * 0x7FEB352D) & M32
    v ^= v >> 15
    v = (v * 0x846CA68B) & M32
    v ^= v >> 16
    return v


def word_seed(evaluation_seed, word):
    """hbatchSeedForWord (FNV-1a over the site-order word).

    test/backend/source/simulator_mps_dist.cu:586-594 (same mixing inline at
    :1525-1529 for the non-MQPU fallback).
    """
    h = (2166136261 ^ evaluation_seed) & M32
    for ch in word.encode():
        h ^= ch
        h = (h * 16777619) & M32
    return h


INITIAL_SAMPLE_SEED = 1234567  # simulator_mps_dist.cu:214
