"""ctypes binding of libgvmps.so (include/gvmps.h).  Loads the in-tree build or
fails loudly: there is no CPU fallback for the product path."""
import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / 'libgvmps.so'

GATE = dict(rx=0, ry=1, rz=2, h=3, x=4, y=5, z=6, s=7, sdg=8, cx=9, swap=10)


class Op(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('q0', ctypes.c_int32), ('q1', ctypes.c_int32),
                ('param', ctypes.c_int32), ('scale', ctypes.c_double), ('offset', ctypes.c_double)]


class Graph(ctypes.Structure):
    _fields_ = [('n_nodes', ctypes.c_int32), ('n_edges', ctypes.c_int32), ('features', ctypes.c_void_p),
                ('rows', ctypes.c_void_p), ('cols', ctypes.c_void_p), ('weights', ctypes.c_void_p),
                ('rows_t', ctypes.c_void_p), ('cols_t', ctypes.c_void_p), ('weights_t', ctypes.c_void_p)]


class ProblemInfo(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int32) for name in (
        'n_qubits', 'n_params', 'n_terms', 'n_active_terms', 'n_prog_ops', 'n_two_site_gates',
        'max_alloc_bond', 'n_chains', 'n_steps', 'n_spill', 'n_generations', 'n_child_nodes',
        'n_sched_steps', 'n_right_nodes')] + [
        ('flops_energy_executed', ctypes.c_double), ('flops_energy_nodes', ctypes.c_double),
        ('flops_svd_nominal', ctypes.c_double), ('flops_grad_executed', ctypes.c_double),
        ('n_rotations', ctypes.c_int32), ('grad_ok', ctypes.c_int32),
        ('real_arith', ctypes.c_int32),
        ('iz_forest', ctypes.c_int32),
        ('exact_split', ctypes.c_int32)]


# (name, argtypes); every function returns int status except where noted
_PROTOS = {
    'gvmps_problem_create': [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(Op), ctypes.c_int,
                             ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_char_p,
                             ctypes.POINTER(ctypes.c_double), ctypes.c_int],
    'gvmps_problem_destroy': [ctypes.c_void_p],
    'gvmps_problem_get_info': [ctypes.c_void_p, ctypes.POINTER(ProblemInfo)],
    'gvmps_problem_get_generations': [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int],
    'gvmps_batch_create': [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                           ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int],
    'gvmps_batch_destroy': [ctypes.c_void_p],
    'gvmps_batch_size': [ctypes.c_void_p],
    'gvmps_batch_energy': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'gvmps_batch_prepare': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'gvmps_batch_sample': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'gvmps_batch_shot_energy': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'gvmps_batch_shot_energy_noalloc': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_uint32), ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p],
    'gvmps_batch_set_timing': [ctypes.c_void_p, ctypes.c_int],
    'gvmps_batch_get_timing': [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)],
    'gvmps_batch_prep_stats': [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)],
    'gvmps_batch_value_grad': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p],
    'gvmps_batch_set_launch': [ctypes.c_void_p, ctypes.c_int, ctypes.c_int],
    'gvmps_batch_set_prep': [ctypes.c_void_p, ctypes.c_int],
    'gvmps_batch_set_arith': [ctypes.c_void_p, ctypes.c_int],
    'gvmps_batch_set_checkpoint': [ctypes.c_void_p, ctypes.c_int],
    'gvmps_batch_set_sparsity': [ctypes.c_void_p, ctypes.c_int],
    'gvmps_batch_chain_flops': [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)],
    'gvmps_batch_grad_flops': [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)],
    'gvmps_batch_evaluate': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'gvmps_batch_capture': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                            ctypes.c_void_p, ctypes.c_void_p],
    'gvmps_batch_replay': [ctypes.c_void_p, ctypes.c_void_p],
    'gvmps_gnn_create': [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                         ctypes.c_uint64, ctypes.c_int],
    'gvmps_gnn_destroy': [ctypes.c_void_p],
    'gvmps_gnn_param_count': [ctypes.c_void_p],
    'gvmps_gnn_params': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'gvmps_gnn_forward': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'gvmps_gnn_loss_grad': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                            ctypes.c_void_p],
    'gvmps_gnn_adam': [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                       ctypes.c_void_p],
    'gvmps_gnn_backward': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'gvmps_batch_read_sites': [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_int)],
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"gvmps native library missing: {LIB_PATH} (build it with "
                               "`python -m gvmps.build` or __graft_entry__.build())")
        L = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)
        for name, args in _PROTOS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.gvmps_last_error.argtypes = []
        L.gvmps_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(status):
    if status != 0:
        raise RuntimeError(lib().gvmps_last_error().decode())
    return status


def declared_symbols():
    """Names declared in include/gvmps.h (for the export test)."""
    return sorted(list(_PROTOS) + ['gvmps_last_error'])
