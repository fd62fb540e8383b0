"""ctypes binding of libmi3hip.so (C ABI declared in include/mi3hip.h).

The library is the only compute path of this package: importing the binding
fails loudly if it has not been built, and there is no CPU fallback.
"""
import ctypes
from pathlib import Path

import numpy as np

LIBPATH = Path(__file__).resolve().parent / 'libmi3hip.so'

MI3_OK = 0
MI3_EINVAL = -1
MI3_ENOBUF = -2
MI3_ESTATE = -3
MI3_EHIP = -4
MI3_ENOMEM = -5
MI3_EUNSUP = -6

# every symbol the header declares, with (restype, argtypes)
_c = ctypes
_P = _c.c_void_p
_S = _c.c_char_p
_u32, _u64, _i, _f, _sz = _c.c_uint32, _c.c_uint64, _c.c_int, _c.c_float, _c.c_size_t
SIGNATURES = {
    'mi3_version': (_S, []),
    'mi3_device_count': (_i, [_c.POINTER(_i)]),
    'mi3_device_info': (_i, [_i, _c.c_char_p, _sz]),
    'mi3_create': (_i, [_i, _u32, _u32, _u32, _u64, _i, _c.POINTER(_P)]),
    'mi3_destroy': (_i, [_P]),
    'mi3_last_error': (_S, [_P]),
    'mi3_buf_alloc': (_i, [_P, _S, _sz]),
    'mi3_buf_exists': (_i, [_P, _S]),
    'mi3_buf_size': (_i, [_P, _S, _c.POINTER(_sz)]),
    'mi3_buf_write': (_i, [_P, _S, _P, _sz, _sz]),
    'mi3_buf_copy': (_i, [_P, _S, _S, _sz]),
    'mi3_stream_copy': (_i, [_P, _S, _S, _sz]),
    'mi3_stream_read': (_i, [_P, _S, _sz, _S]),
    'mi3_buf_read_async': (_i, [_P, _S, _sz, _sz, _c.POINTER(_i)]),
    'mi3_ticket_wait': (_i, [_P, _i, _P, _sz]),
    'mi3_buf_fill_u32': (_i, [_P, _S, _u32]),
    'mi3_synchronize': (_i, [_P]),
    'mi3_fill_seqs': (_i, [_P, _S, _P]),
    'mi3_set_wgsize': (_i, [_P, _u32]),
    'mi3_init_mcmc': (_i, [_P, _u32, _u64, _u64]),
    'mi3_run_mcmc': (_i, [_P, _f]),
    'mi3_stage_probe': (_i, [_P, _i]),
    'mi3_wm_atomic_probe': (_i, [_P, _u32, _u32]),
    'mi3_mcmc_step_count': (_i, [_P, _c.POINTER(_u64)]),
    'mi3_gen_indep': (_i, [_P, _S]),
    'mi3_calc_energies': (_i, [_P, _S, _S, _S, _u32]),
    'mi3_calc_bicounts': (_i, [_P, _S, _u32]),
    'mi3_bicount_split': (_i, [_P, _u32, _c.POINTER(_u32)]),
    'mi3_bicounts_to_bimarg': (_i, [_P, _u32]),
    'mi3_weighted_marg': (_i, [_P, _S, _S, _u32]),
    'mi3_weighted_marg_noalloc': (_i, [_P, _P, _u32, _P, _u32, _P]),
    'mi3_renormalize_bimarg': (_i, [_P]),
    'mi3_min_buf': (_i, [_P, _S, _u32]),
    'mi3_weight_stats': (_i, [_P, _S, _u32]),
    'mi3_dE_to_weights': (_i, [_P, _S, _S, _u32, _f]),
    'mi3_fixed_beta_weights': (_i, [_P, _S, _S, _u32, _f, _f]),
    'mi3_add_float_bufs': (_i, [_P, _S, _S, _u32]),
    'mi3_update_J': (_i, [_P, _f, _f, _S]),
    'mi3_reg_l1z': (_i, [_P, _f, _f, _f]),
    'mi3_reg_l2z': (_i, [_P, _f, _f, _f]),
    'mi3_update_J_reg': (_i, [_P, _f, _f, _i, _f]),
    'mi3_copy_seqs': (_i, [_P, _S, _u32, _S, _u32, _u32]),
    'mi3_buf_device_ptr': (_i, [_P, _S, _c.POINTER(_P)]),
    'mi3_buf_mark_written': (_i, [_P, _S]),
    'mi3_stream': (_i, [_P, _c.POINTER(_P)]),
    'mi3_made_last_error': (_S, []),
    'mi3_gemm_bf16': (_i, [_P, _i, _i, _i, _i, _i, _f, _P, _i, _P, _i, _f, _P, _i, _P, _i, _P,
                           _i]),
    'mi3_made_embed': (_i, [_P, _P, _u32, _u32, _u32, _u32, _u32, _P, _P, _P, _P]),
    'mi3_made_embed_grad': (_i, [_P, _P, _u32, _u32, _u32, _u32, _u32, _P, _P]),
    'mi3_made_xent': (_i, [_P, _P, _P, _u32, _u32, _u32, _u32, _f, _P, _P, _P]),
    'mi3_relu_grad': (_i, [_P, _u64, _P, _P, _P, _P]),
    'mi3_made_fused': (_i, [_P, _u32, _u32, _u32, _u32, _P, _P, _P, _P, _P, _P, _u64, _u32, _u32,
                            _P, _u32, _P, _i]),
    'mi3_adam_step': (_i, [_P, _u64, _P, _P, _P, _P, _P, _f, _f, _f, _f, _u32, _P]),
    'mi3_colsum': (_i, [_P, _u32, _u32, _P, _P]),
    'mi3_to_bf16': (_i, [_P, _u64, _P, _P]),
    'mi3_mh_accept': (_i, [_P, _u32, _u32, _P, _u32, _P, _u32, _P, _P, _P, _P, _u64, _u64, _u32,
                           _P]),
    'mi3_batch_create': (_i, [_i, _u32, _c.POINTER(_u32), _c.POINTER(_u32), _c.POINTER(_u32),
                              _u64, _c.POINTER(_P)]),
    'mi3_batch_destroy': (_i, [_P]),
    'mi3_batch_last_error': (_S, [_P]),
    'mi3_batch_set_couplings': (_i, [_P, _u32, _P]),
    'mi3_batch_set_seqs': (_i, [_P, _u32, _P]),
    'mi3_batch_get_seqs': (_i, [_P, _u32, _P]),
    'mi3_batch_run': (_i, [_P, _u32]),
    'mi3_batch_synchronize': (_i, [_P]),
    'mi3_batch_tune': (_i, [_P, _u32]),
    'mi3_batch_tile': (_i, [_P, _u32, _c.POINTER(_u32)]),
    'mi3_batch_set_tile': (_i, [_P, _u32, _u32]),
    'mi3_batch_step_count': (_i, [_P, _u32, _c.POINTER(_u64)]),
    'mi3_copy_subseq': (_i, [_P, _S, _u32, _S, _S]),
    'mi3_reg_SCADJ': (_i, [_P, _f, _f, _f, _f]),
    'mi3_reg_X': (_i, [_P, _f, _f, _f]),
    'mi3_reg_Xij': (_i, [_P, _f, _f]),
    'mi3_reg_SCADX': (_i, [_P, _f, _f, _f, _f, _f]),
    'mi3_reg_expX': (_i, [_P, _f, _f, _f]),
    'mi3_reg_ddE': (_i, [_P, _f, _f, _f]),
    'mi3_reg_SCADddE': (_i, [_P, _f, _f, _f, _f]),
    'mi3_profile_flush': (_i, [_P, _S]),
    'mi3_profile_kernel_ms': (_i, [_P, _S, _c.POINTER(_c.c_double), _c.POINTER(_i)]),
    'mi3_profile_reset': (_i, [_P]),
}

_lib = None


class Mi3Error(RuntimeError):
    """A failure reported by libmi3hip (status code + message)."""

    def __init__(self, code, msg):
        super().__init__(f"{msg} (libmi3hip status {code})")
        self.code = code


def lib():
    """Load libmi3hip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise ImportError(f"{LIBPATH} is missing: build it with "
                              "`python mi3-gpu_amd/build.py` (this package has no CPU path)")
        L = ctypes.CDLL(str(LIBPATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def enc(name):
    return name.encode() if isinstance(name, str) else name


def check(ctx, rc):
    """Map a libmi3hip status to the exception types of the reference wrapper
    (ValueError for bad dtype/shape/size, Exception for state errors)."""
    if rc >= 0:
        return rc
    msg = lib().mi3_last_error(ctx)
    msg = msg.decode(errors='replace') if msg else ''
    if rc in (MI3_EINVAL, MI3_EUNSUP):
        raise ValueError(msg)
    raise Mi3Error(rc, msg)


def device_count():
    n = ctypes.c_int(0)
    rc = lib().mi3_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


def device_info(dev):
    buf = ctypes.create_string_buffer(1024)
    check(None, lib().mi3_device_info(dev, buf, len(buf)))
    return buf.value.decode()


def as_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


__all__ = ['lib', 'check', 'enc', 'Mi3Error', 'device_count', 'device_info', 'as_ptr', 'np']
