"""ctypes binding of libflowket_hip.so (the C ABI declared in
include/flowket_hip.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises.  There is no CPU fallback.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libflowket_hip.so')

FK_OK = 0
FK_CONN_SWAP = 0
FK_CONN_SWAP_ZERO_PAD = 1
FK_CONN_FLIP = 2
FK_DIAG_SUM_F64 = 0
FK_DIAG_ISING_F32 = 1
FK_ACT_RELU = 0
FK_ACT_ELU = 1
FK_PREC_FP32 = 0
FK_PREC_BF16 = 1
FK_PREC_FP16 = 2

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64

# name -> (restype, argtypes)
SIGNATURES = {
    'fk_version': (ctypes.c_int, []),
    'fk_last_error': (ctypes.c_char_p, []),
    'fk_find_conn': (ctypes.c_int, [_P, _I64, _I32, _I32, _P, _P, _P, _I32, _P, _P, _P, _I32, _I32, _P, _P, _P, _P,
                                    _P]),
    'fk_compact_scratch_bytes': (_I64, [_I64]),
    'fk_compact_connections': (ctypes.c_int, [_P, _I64, _I32, _P, _P, _P, _P]),
    'fk_local_energy': (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _P, _P]),
    'fk_local_energy_dense': (ctypes.c_int, [_P, _P, _P, _I32, _I64, _P, _P]),
    'fk_energy_moments': (ctypes.c_int, [_P, _I64, _P, _P]),
    'fk_sigma_z': (ctypes.c_int, [_P, _I64, _I32, _I32, _P, _P]),
    'fk_convar_packed_sizes': (ctypes.c_int, [_I32, _I32, _I32, _P, _P]),
    'fk_convar_pack_maps': (ctypes.c_int, [_I32, _I32, _I32, _P, _P]),
    'fk_pack_gather': (ctypes.c_int, [_P, _P, _I64, _I32, _P, _P]),
    'fk_convar_forward': (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _I32,
                                         _P, _P, _P, _P, _P]),
    'fk_convar_forward_ws': (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _I32,
                                            _P, _P, _P, _P, _P, _I64, _P]),
    'fk_bucket_scratch_bytes': (_I64, [_I64, _I32]),
    'fk_bucket_connections': (ctypes.c_int, [_P, _I64, _I32, _P, _I32, _P, _P, _P, _P, _P, _P]),
    'fk_convar_rcache_bytes': (_I64, [_I32, _I32, _I32, _I32, _I32]),
    'fk_convar_forward_rows': (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _I32,
                                              _P, _P, _I32, _P, _I32, _P, _P, _P, _P, _I64, _P]),
    'fk_convar_forward_noalloc': (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _I64,
                                                 _I32, _P, _P, _P, _P, _P]),
    'fk_convar_forward_noalloc_scratch_bytes': (_I64, []),
    'fk_convar_forward_workspace_bytes': (_I64, [_I32, _I32, _I32, _I32, _I32, _I32]),
    'fk_convar_num_schedules': (_I32, [_I32, _I32]),
    'fk_convar_schedule_info': (ctypes.c_int, [_I32, _I32, _I32, _P, _P]),
    'fk_convar_set_schedule': (ctypes.c_int, [_I32, _I32, _I32, _I32, _I32, _I32, _I32]),
    'fk_convar_get_schedule': (_I32, [_I32, _I32, _I32, _I32, _I32, _I32]),
    'fk_ar_draw_site': (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _P, _P]),
    'fk_fast_sampler_workspace_bytes': (_I64, [_I32, _I32, _I32, _I32, _I32, _I64]),
    'fk_fast_sample': (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I64, _P, _P, _P, _P, _P]),
    'fk_convar_grad_workspace_bytes': (_I64, [_I32, _I32, _I32, _I32, _I64]),
    'fk_convar_grad': (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _P, _I64, _P, _P, _P, _I64, _P, _P]),
    'fk_cg_state_doubles': (_I64, []),
    'fk_cg_init': (ctypes.c_int, [_P, _P, _P, _P, _I64, ctypes.c_double, _I64, _P, _P, _P]),
    'fk_cg_step': (ctypes.c_int, [_P, _P, _P, _P, _I64, _P, _P, _P]),
    'fk_center_pairs': (ctypes.c_int, [_P, _I64, _P, _P]),
    'fk_adam_step': (ctypes.c_int, [_P, _P, _P, _P, _I64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, _I64, _P]),
    'fk_convar_vjp_workspace_bytes': (_I64, [_I32, _I32, _I32, _I32, _I64]),
    'fk_convar_vjp': (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _P, _I64, _P, ctypes.c_double, _P, _P, _P,
                                     _P]),
    'fk_convar_jvp_workspace_bytes': (_I64, [_I32, _I32, _I32, _I32, _I64]),
    'fk_convar_jvp': (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _P, _I64, _P, _P, _P, _P]),
    'fk_wn_grad_rows': (ctypes.c_int, [_P, _I64, _I64, _P, _P, _P, _I32, _I32, _I32, _I64, _P]),
    'fk_wn_tangent': (ctypes.c_int, [_P, _P, _P, _I32, _I32, _I32, _P, _P, _I64, _P]),
    'fk_o_mean': (ctypes.c_int, [_P, _P, _I64, _I64, _P, _P, _P]),
    'fk_o_times_v': (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _I64, _P, _P, _P]),
    'fk_oh_times_u': (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _I64, ctypes.c_double, ctypes.c_double, _P, _P,
                                     _P, _P]),
    'fk_sr_scratch_doubles': (_I64, []),
    'fk_sr_matvec_workspace_bytes': (_I64, [_I64, _I64]),
    'fk_sr_matvec': (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _I64, ctypes.c_double, ctypes.c_double, _P, _P, _I32,
                                    _P, _I64, _P, _P]),
    'fk_dot_f64': (ctypes.c_int, [_P, _P, _I64, _P, _P, _P]),
    'fk_axpy': (ctypes.c_int, [ctypes.c_double, _P, _P, _I64, _P]),
    'fk_xpby': (ctypes.c_int, [_P, ctypes.c_double, _P, _I64, _P]),
    'fk_rbm_log_psi': (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I32, _I32, _P, _P, _P, _I64, _I32, _P, _P, _P, _P]),
    'fk_mh_rbm': (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _I32, _I64, _I64, _I64, _I64, _P,
                                 _P, _P, _P, _P, _P, _P, ctypes.c_uint64, _P, _P]),
    'fk_rbm_grad': (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I32, _I32, _P, _I64, _P, _P, _I64, _P]),
    'fk_mh_propose': (ctypes.c_int, [_P, _I64, _I32, _I32, _I32, _P, _P, _P, _P]),
    'fk_mh_accept': (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _I32, _P, _P, _P]),
    'fk_mh_accept_independent': (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _I64, _I32, _P, _P, _P]),
    'fk_made_packed_bytes': (_I64, [_I32, _I32]),
    'fk_made_pack': (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I32, _I32, _P, _P]),
    'fk_made_log_prob': (ctypes.c_int, [_P, _I32, _I32, _P, _I64, _P, _P]),
    'fk_made_sample': (ctypes.c_int, [_P, _I32, _I32, _P, _I64, _P, _P, _P]),
    'fk_made_fit_workspace_floats': (_I64, [_I32, _I32, _I64]),
    'fk_made_fit_step': (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _P, _I64, _P,
                                        ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _I64, _P]),
    'fk_exact_states': (ctypes.c_int, [_I32, _I64, _I64, _P, _P]),
    'fk_exact_workspace_doubles': (_I64, [_I64]),
    'fk_exact_normalize': (ctypes.c_int, [_P, _I64, _P, _P, _P, _P]),
    'fk_exact_local_energy': (ctypes.c_int, [_P, _I64, _I32, _I32, _P, _P, _P, _P, _I32, _P, _P]),
    'fk_exact_energy': (ctypes.c_int, [_P, _P, _I64, _P, _P, _P, _P]),
    'fk_exact_sample': (ctypes.c_int, [_P, _I64, _I32, _P, _I64, _P, _P, _P]),
    'fk_transform_spins': (ctypes.c_int, [_P, _I64, _I32, _P, _I32, _P, _P]),
    'fk_ensemble_combine': (ctypes.c_int, [_P, _I32, _I64, _I64, _P, _P, _P]),
}

_lib = None


class FlowketHipError(RuntimeError):
    pass


def load():
    """Load the library (once).  Raises FlowketHipError when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FlowketHipError(
            f'{LIB_PATH} not found: build it with `python -m flowket_amd.build` '
            '(flowket_amd has no CPU fallback)')
    # torch is imported first so that its HIP runtime (same soname) is the one
    # the library binds to.
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(t):
    """Device (or host) pointer of a tensor, None for None."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def call(name, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != FK_OK:
        msg = lib.fk_last_error().decode(errors='replace')
        raise FlowketHipError(f'{name} failed (code {rc}): {msg}')
    return rc


def require_gpu():
    if not torch.cuda.is_available():
        raise FlowketHipError('flowket_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
    load()
