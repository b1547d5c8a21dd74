"""ctypes binding of libmbpt_hip.so, the package's gfx950 kernels.

The library is loaded after ``import torch`` so that it binds to the HIP
runtime torch already loaded (both carry the soname libamdhip64.so.7): device
pointers and streams then come from one runtime.  There is no fallback: if
the library is missing, ``lib()`` raises.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        'libmbpt_hip.so')

FP64, FP32, BF16X3 = 0, 1, 2
PRECISIONS = {'fp64': FP64, 'fp32': FP32, 'bf16x3': BF16X3}
OK, ERR_VALUE, ERR_HIP, ERR_UNSUPPORTED = 0, 1, 2, 3
CHI0_MAX_TAU = 64
GW_MAX_STATES = 4096

#: Every symbol include/mbpt_hip.h declares.
EXPORTS = ('mbpt_last_error', 'mbpt_abi_version', 'mbpt_sigma_create',
           'mbpt_sigma_apply', 'mbpt_sigma_apply_ab', 'mbpt_sigma_screened_diag',
           'mbpt_sigma_held_bytes', 'mbpt_sigma_workspace_bytes', 'mbpt_sigma_profile',
           'mbpt_sigma_profile_read', 'mbpt_sigma_destroy', 'mbpt_chi0_create',
           'mbpt_chi0_build', 'mbpt_chi0_terms', 'mbpt_chi0_transform',
           'mbpt_chi0_transform_profile_read', 'mbpt_chi0_profile', 'mbpt_chi0_profile_read',
           'mbpt_chi0_sigma', 'mbpt_chi0_sigma_profile_read', 'mbpt_chi0_destroy', 'mbpt_gemm_tn_f64_work', 'mbpt_gemm_tn_f64',
           'mbpt_gemm_tn_f64_noalloc_max', 'mbpt_gemm_tn_f64_noalloc',
           'mbpt_gemm_tn_seg_f64_work', 'mbpt_gemm_tn_seg_f64', 'mbpt_gemm_nn_seg_f64',
           'mbpt_ritz_residual_f64', 'mbpt_block_scale_f64', 'mbpt_jd_correction_f64_work',
           'mbpt_jd_correction_f64', 'mbpt_ritz_max_m', 'mbpt_ritz_eigh_f64_work',
           'mbpt_ritz_eigh_f64',
           'mbpt_mlp_nparams', 'mbpt_mlp_create', 'mbpt_mlp_get_params',
           'mbpt_mlp_forward', 'mbpt_mlp_gradient', 'mbpt_mlp_train_step',
           'mbpt_mlp_correction', 'mbpt_mlp_correction_batched', 'mbpt_mlp_destroy',
           'mbpt_syev_batched_f64', 'mbpt_bnn_f64', 'mbpt_btn_f64_work', 'mbpt_btn_f64',
           'mbpt_batch_create', 'mbpt_batch_plan', 'mbpt_batch_screened_diag',
           'mbpt_batch_npair_max',
           'mbpt_batch_held_bytes', 'mbpt_batch_apply', 'mbpt_batch_destroy')

_lib = None

_vp, _i, _d, _dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_SIGNATURES = {
    'mbpt_last_error': (ctypes.c_char_p, []),
    'mbpt_abi_version': (_i, []),
    'mbpt_sigma_create': (_i, [ctypes.POINTER(_vp), _i, _i, _i, _i, _d, _vp, _vp,
                               _vp, _dp, _vp]),
    'mbpt_sigma_apply': (_i, [_vp, _i, _vp, _vp, _vp]),
    'mbpt_sigma_apply_ab': (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    'mbpt_sigma_screened_diag': (_i, [_vp, _vp, _vp]),
    'mbpt_sigma_held_bytes': (ctypes.c_size_t, [_vp]),
    'mbpt_sigma_workspace_bytes': (ctypes.c_size_t, [_vp, _i, _i]),
    'mbpt_sigma_profile': (_i, [_vp, _i]),
    'mbpt_sigma_profile_read': (_i, [_vp, _dp, ctypes.POINTER(ctypes.c_long), _dp]),
    'mbpt_sigma_destroy': (_i, [_vp]),
    'mbpt_chi0_create': (_i, [ctypes.POINTER(_vp), _i, _i, _i, _i, _vp, _dp, _d,
                              _vp]),
    'mbpt_chi0_build': (_i, [_vp, _i, _dp, _dp, _vp, _i, _vp]),
    'mbpt_chi0_terms': (_i, [_vp, _i, _dp, _vp]),
    'mbpt_chi0_transform': (_i, [_vp, _i, _dp, _vp, _i, _vp]),
    'mbpt_chi0_transform_profile_read': (_i, [_vp, _dp, ctypes.POINTER(ctypes.c_long)]),
    'mbpt_chi0_profile': (_i, [_vp, _i]),
    'mbpt_chi0_profile_read': (_i, [_vp, _dp, ctypes.POINTER(ctypes.c_long), _dp]),
    'mbpt_chi0_sigma': (_i, [_vp, _i, _dp, _vp, _i, _ip, _vp, _vp]),
    'mbpt_chi0_sigma_profile_read': (_i, [_vp, _dp, ctypes.POINTER(ctypes.c_long), _dp]),
    'mbpt_chi0_destroy': (_i, [_vp]),
    'mbpt_gemm_tn_f64_work': (ctypes.c_long, [_i, _i, _i]),
    'mbpt_gemm_tn_f64': (_i, [_i, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, ctypes.c_long,
                              _vp]),
    'mbpt_gemm_tn_f64_noalloc_max': (ctypes.c_long, []),
    'mbpt_gemm_tn_f64_noalloc': (_i, [_i, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp]),
    'mbpt_gemm_tn_seg_f64_work': (ctypes.c_long, [_i, _i, _i]),
    'mbpt_gemm_tn_seg_f64': (_i, [_i, _i, _i, _vp, _i, _i, _i, _i, _vp, _i, _i, _i, _i, _vp, _i,
                                  _vp, ctypes.c_long, _vp]),
    'mbpt_gemm_nn_seg_f64': (_i, [_i, _i, _i, _vp, _i, _i, _i, _i, _vp, _i, _vp, _i, _i, _d, _d,
                                  _vp]),
    'mbpt_ritz_residual_f64': (_i, [_i, _i, _i, _vp, _i, _i, _vp, _i, _i, _vp, _i, _vp, _vp, _i,
                                    _vp, _i, _vp]),
    'mbpt_block_scale_f64': (_i, [_i, _i, _i, _d, _vp, _vp, _i, _vp]),
    'mbpt_ritz_max_m': (_i, []),
    'mbpt_ritz_eigh_f64_work': (ctypes.c_long, [_i, _i]),
    'mbpt_ritz_eigh_f64': (_i, [_i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _vp, ctypes.c_long, _vp]),
    'mbpt_jd_correction_f64_work': (ctypes.c_long, [_i, _i]),
    'mbpt_jd_correction_f64': (_i, [_i, _i, _vp, _i, _vp, _i, _vp, _vp, _d, _vp, _i, _vp,
                                    ctypes.c_long, _vp]),
    'mbpt_mlp_nparams': (_i, []),
    'mbpt_mlp_create': (_i, [ctypes.POINTER(_vp), _fp, _vp]),
    'mbpt_mlp_get_params': (_i, [_vp, _fp, _vp]),
    'mbpt_mlp_forward': (_i, [_vp, ctypes.c_long, _vp, _vp, _vp]),
    'mbpt_mlp_gradient': (_i, [_vp, ctypes.c_long, _vp, _vp, _vp, _fp, _vp]),
    'mbpt_mlp_train_step': (_i, [_vp, ctypes.c_long, _vp, _vp, ctypes.c_float, _fp, _vp]),
    'mbpt_mlp_correction': (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _vp, _i, _vp]),
    'mbpt_mlp_correction_batched': (_i, [_vp, _i, _i, _i, _vp, ctypes.c_long, _vp, _i, _vp,
                                         _vp, _vp, _i, _vp]),
    'mbpt_mlp_destroy': (_i, [_vp]),
    'mbpt_syev_batched_f64': (_i, [_i, _i, _vp, _i, ctypes.c_long, _vp, ctypes.c_long,
                                   _vp, _i, ctypes.c_long, _i, _d, _vp]),
    'mbpt_bnn_f64': (_i, [_i, _i, _i, _i, _vp, _vp, ctypes.c_long, _i, _vp, ctypes.c_long, _i,
                          _vp, ctypes.c_long, _i, _i, _vp]),
    'mbpt_btn_f64_work': (ctypes.c_long, [_i, _i, _i, _i]),
    'mbpt_btn_f64': (_i, [_i, _i, _i, _i, _vp, _vp, ctypes.c_long, _i, _vp, ctypes.c_long, _i,
                          _vp, ctypes.c_long, _i, _vp, ctypes.c_long, _vp]),
    'mbpt_batch_create': (_i, [ctypes.POINTER(_vp), _i, _ip, _ip, _ip, _i, _d,
                               ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                               ctypes.POINTER(_vp), _dp, _i, _vp]),
    'mbpt_batch_plan': (_i, [_vp, _i, _i, _i, _i, _i, _vp]),
    'mbpt_batch_screened_diag': (_i, [_vp, _vp, _vp]),
    'mbpt_batch_npair_max': (_i, [_vp]),
    'mbpt_batch_held_bytes': (ctypes.c_size_t, [_vp]),
    'mbpt_batch_apply': (_i, [_vp, _vp, _vp, _vp, _vp]),
    'mbpt_batch_destroy': (_i, [_vp]),
}


class KernelError(RuntimeError):
    """A HIP runtime failure inside libmbpt_hip."""


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f'{LIB_PATH} is missing: build it with '
                '`python -c "import __graft_entry__ as g; g.build()"` '
                '(hipcc --offload-arch=gfx950). There is no CPU fallback.')
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc, what):
    """Map a status code to the reference's exception classes."""
    if rc == OK:
        return
    msg = f'{what}: {lib().mbpt_last_error().decode(errors="replace")}'
    if rc == ERR_VALUE:
        raise ValueError(msg)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise KernelError(msg)


def precision_code(precision):
    if isinstance(precision, int):
        return precision
    try:
        return PRECISIONS[precision.lower()]
    except KeyError:
        raise ValueError(f'precision {precision!r}: one of {sorted(PRECISIONS)}') from None


def stream_ptr(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def dptr(t):
    """Device pointer of a contiguous CUDA tensor."""
    if not t.is_cuda:
        raise ValueError('expected a device tensor')
    if not t.is_contiguous():
        raise ValueError('expected a contiguous tensor')
    return ctypes.c_void_p(t.data_ptr())


def host_doubles(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)
