"""gvmps_batch_shot_energy_noalloc against gvmps_batch_shot_energy, through
the C ABI: the same energies and per-term values, bitwise, on two calls in a
row (the second with fewer instances and fewer shots), and the size limit of
its scratch rejected with an error before anything runs."""
import ctypes

import numpy as np
import pytest

from gvmps import Batch, Problem, seeds, workloads
from gvmps._lib import lib

pytestmark = pytest.mark.gpu


def shot_call(name, batch, x, shots, sd):
    import torch
    L = lib()
    xt = torch.as_tensor(np.asarray(x, dtype=np.float64).reshape(-1), device='cuda')
    e = torch.full((batch.size,), float('nan'), dtype=torch.float64, device='cuda')
    tv = torch.full((2 * batch.n_terms_total,), float('nan'), dtype=torch.float64, device='cuda')
    s = np.ascontiguousarray(np.asarray(sd, dtype=np.uint32))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    status = getattr(L, name)(batch._h, ctypes.c_void_p(xt.data_ptr()), int(shots),
                              s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.c_void_p(e.data_ptr()),
                              ctypes.c_void_p(tv.data_ptr()), stream)
    if status != 0:
        raise RuntimeError(L.gvmps_last_error().decode())
    torch.cuda.synchronize()
    return e.cpu().numpy(), tv.cpu().numpy()


def test_noalloc_shot_energy_matches_allocating_entry_point(cuda):
    w = workloads.config(0)
    p = Problem(w['n'], w['ops'], w['nparams'], w['words'], w['coeffs'], max_bond=16, abs_cut=0, rel_cut=0)
    r = np.random.default_rng(5)
    X = r.uniform(-np.pi, np.pi, (8, w['nparams']))
    sd = [seeds.shot_shard_seed(seeds.INITIAL_SAMPLE_SEED + j, 0) for j in range(8)]
    print('case: no-alloc shot energies, 12q, 8 instances x 1000 shots', flush=True)
    big = Batch([p], [8])
    e_ref, t_ref = shot_call('gvmps_batch_shot_energy', big, X, 1000, sd)
    e_na, t_na = shot_call('gvmps_batch_shot_energy_noalloc', big, X, 1000, sd)
    assert np.all(np.isfinite(e_ref)) and np.array_equal(e_na, e_ref) and np.array_equal(t_na, t_ref)
    print('case: no-alloc shot energies, second call: 2 instances x 300 shots', flush=True)
    small = Batch([p], [2])
    e_ref2, t_ref2 = shot_call('gvmps_batch_shot_energy', small, X[5:7], 300, sd[5:7])
    e_na2, t_na2 = shot_call('gvmps_batch_shot_energy_noalloc', small, X[5:7], 300, sd[5:7])
    assert np.array_equal(e_na2, e_ref2) and np.array_equal(t_na2, t_ref2)
    # per-term values are (sum of +-1) / shots and the energy their weighted sum
    assert np.all(np.abs(t_na2.reshape(2, -1, 2)[:, :, 1]) == 0.0)
    print('case: no-alloc scratch limit', flush=True)
    with pytest.raises(RuntimeError, match='exceeds the scratch maximum of 2097152 shot words'):
        shot_call('gvmps_batch_shot_energy_noalloc', big, X, 262145, sd)
    # the failed call left the batch usable
    e_na3, _ = shot_call('gvmps_batch_shot_energy_noalloc', small, X[5:7], 300, sd[5:7])
    assert np.array_equal(e_na3, e_ref2)
