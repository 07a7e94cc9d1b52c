"""ctypes binding of the oopao_amd C-ABI (include/oopao_amd.h).

The shared library is built in-tree by ``python -m oopao_amd.build``. There is no
CPU fallback: importing a GPU entry point without the library raises.
"""
import ctypes
import os

from .tools import OopaoError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get('OOPAO_AMD_LIB', os.path.join(_HERE, 'liboopao_amd.so'))

_lib = None

c_int = ctypes.c_int
c_float = ctypes.c_float
c_size_t = ctypes.c_size_t
c_void_p = ctypes.c_void_p
c_double = ctypes.c_double

# name -> (restype, argtypes); the single source of the binding's signatures
SIGNATURES = {
    'oopao_last_error': (ctypes.c_char_p, []),
    'oopao_prepare_tables': (c_int, [c_int]),
    'oopao_pwfs_supported': (c_int, [c_int, c_int, c_int]),
    'oopao_pwfs_plans': (c_int, [c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    'oopao_pwfs_workspace_bytes': (c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_pwfs_frames': (c_int, [c_void_p, c_float, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                  c_void_p, c_void_p, c_size_t, c_void_p, c_void_p]),
    'oopao_pwfs_slopes_workspace_bytes': (c_int, [c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_pwfs_slopes': (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_float,
                                  c_void_p, c_int, c_int, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p]),
    'oopao_pwfs_slopes_static': (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_float,
                                         c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    'oopao_pwfs_intermediate_bytes': (c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_size_t),
                                              ctypes.POINTER(c_size_t)]),
    'oopao_pwfs_rows': (c_int, [c_void_p, c_float, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int,
                                c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    'oopao_pwfs_cols': (c_int, [c_void_p, c_size_t, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_void_p,
                                c_void_p, c_size_t, c_void_p, c_void_p]),
    'oopao_pwfs_raw_bytes': (c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_pwfs_raw': (c_int, [c_void_p, c_float, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_int, c_int, c_int, c_int, c_int, c_int, c_void_p, c_size_t, c_void_p, c_void_p]),
    'oopao_bin_area': (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    'oopao_pwfs_focal_bytes': (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_pwfs_focal': (c_int, [c_void_p, c_float, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int,
                                 c_int, c_int, c_int, c_int, c_int, c_void_p, c_size_t, c_void_p, c_void_p]),
    'oopao_pwfs_adjoint_bytes': (c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_pwfs_adjoint': (c_int, [c_void_p, c_float, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_void_p,
                                   c_void_p, c_size_t, c_void_p, c_void_p]),
    'oopao_pwfs_generic_supported': (c_int, [c_int, c_int, c_int]),
    'oopao_pwfs_generic_bytes': (c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_pwfs_generic_frames': (c_int, [c_void_p, c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                          c_int, c_int, c_int, c_int, c_void_p, c_size_t, c_void_p, c_void_p,
                                          c_void_p]),
    'oopao_pwfs_generic_focal_bytes': (c_int, [c_int, c_int, c_int, c_int, c_int, c_int,
                                               ctypes.POINTER(c_size_t)]),
    'oopao_pwfs_generic_focal': (c_int, [c_void_p, c_float, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                         c_int, c_int, c_int, c_void_p, c_size_t, c_void_p, c_void_p]),
    'oopao_pwfs_generic_adjoint_bytes': (c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_pwfs_generic_adjoint': (c_int, [c_void_p, c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_int, c_int, c_int, c_int, c_int, c_int, c_void_p, c_size_t, c_void_p,
                                           c_void_p]),
    'oopao_loop_residual_workspace_bytes': (c_int, [ctypes.POINTER(c_size_t)]),
    'oopao_loop_residual': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_double, c_void_p, c_void_p,
                                    c_void_p, c_void_p]),
    'oopao_atm_add_row':(c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'oopao_atm_minmax': (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    'oopao_atm_boil_bytes': (c_int, [c_int, ctypes.POINTER(c_size_t)]),
    'oopao_atm_boil_subharmonics': (c_int, [c_void_p, ctypes.c_longlong, c_void_p, c_int, c_void_p, c_void_p]),
    'oopao_normal_philox': (c_int, [c_void_p, ctypes.c_longlong, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                    ctypes.c_ulonglong, c_void_p]),
    'oopao_atm_boil': (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_double,
                               c_double, c_void_p, c_size_t, c_void_p]),
    'oopao_atm_layer_opd': (c_int, [c_void_p, c_int, c_void_p, c_double, c_double, c_int, c_int, c_double, c_void_p,
                                    c_int, c_int, c_void_p]),
    'oopao_bilinear_affine': (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_double, c_double, c_double,
                                      c_double, c_int, c_void_p]),
    'oopao_field_phase': (c_int, [c_void_p, c_float, c_void_p, c_int, c_int, c_void_p]),
    'oopao_asm_bytes': (c_int, [c_int, ctypes.POINTER(c_size_t)]),
    'oopao_asm_propagate': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_size_t, c_void_p]),
    'oopao_field_readout': (c_int, [c_void_p, c_void_p, c_float, c_void_p, c_void_p, c_int, c_void_p]),
    'oopao_dm_gaussian_modes': (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_double, c_double, c_double,
                                        c_float, c_void_p]),
    'oopao_gemv': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_float, c_float, c_void_p]),
    'oopao_bf16_packed_bytes': (c_int, [c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_bf16_pack': (c_int, [c_void_p, ctypes.c_longlong, ctypes.c_longlong, c_int, c_int, c_void_p, c_void_p,
                                c_void_p]),
    'oopao_gemm_bf16': (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                ctypes.c_longlong, c_float, c_int, c_void_p]),
    'oopao_gemv_bf16_workspace_bytes': (c_int, [c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_gemv_bf16': (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_float, c_float, c_void_p,
                                c_size_t, c_void_p]),
    'oopao_fft2_supported': (c_int, [c_int]),
    'oopao_fft2_bytes': (c_int, [c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_fft2': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_float, c_void_p, c_size_t, c_void_p]),
    'oopao_gsc_workspace_bytes': (c_int, [c_int, c_int, c_int, ctypes.POINTER(c_size_t)]),
    'oopao_gsc_impulse_response': (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_size_t,
                                           c_void_p]),
    'oopao_gsc_basis_product': (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_size_t,
                                        c_void_p]),
    'oopao_gsc_sensitivity': (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_size_t,
                                      c_void_p]),
    'oopao_dm_separable_opd': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                       c_int, c_void_p]),
}


def lib():
    """The loaded library (raises OopaoError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OopaoError('oopao_amd HIP library not found at %s; build it with '
                             '`python -m oopao_amd.build` (there is no CPU fallback)' % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(handle, name)
            f.restype = res
            f.argtypes = args
        _lib = handle
    return _lib


def check(rc):
    """Raise OopaoError with the library's message on a non-zero return code."""
    if rc != 0:
        msg = lib().oopao_last_error().decode(errors='replace')
        raise OopaoError('oopao_amd error %d: %s' % (rc, msg))


def call(name, *args):
    check(getattr(lib(), name)(*args))


def prepare_tables(N, device):
    """Build the twiddle table of length N on ``device`` now (not inside a later, maybe captured, call)."""
    import torch
    device = torch.device(device)
    if device.type != 'cuda':
        return
    with torch.cuda.device(device):
        call('oopao_prepare_tables', int(N))
