"""ctypes binding of libpgj.so (the C ABI declared in include/pgj.h).

The library is loaded from this package directory. There is no fallback: if
the library is missing or fails to load, every entry point raises.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libpgj.so')

c_void_p, c_int, c_double, c_size_t = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
c_int64 = ctypes.c_int64
c_dp = ctypes.POINTER(ctypes.c_double)


class Grid(ctypes.Structure):
    """pgj_grid: periodic grid (nz, ny, nx), voxel (hx, hy, hz), medium (lam, mu)."""
    _fields_ = [('nz', ctypes.c_int32), ('ny', ctypes.c_int32), ('nx', ctypes.c_int32),
                ('reserved', ctypes.c_int32), ('hx', c_double), ('hy', c_double), ('hz', c_double),
                ('lam', c_double), ('mu', c_double)]


GP = ctypes.POINTER(Grid)

# name -> (restype, argtypes)
SIGNATURES = {
    'pgj_version': (c_int, []),
    'pgj_last_error': (ctypes.c_char_p, []),
    'pgj_hex8_apply': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_hex8_apply32': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_hex8_apply32_mfma': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    'pgj_hex8_apply_reference': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_hex8_rhs': (c_int, [GP, c_void_p, c_dp, c_void_p, c_void_p]),
    'pgj_hex8_energy': (c_int, [GP, c_int, c_void_p, c_dp, c_void_p, c_void_p]),
    'pgj_effective_tensor_ws_bytes': (c_size_t, [GP]),
    'pgj_effective_tensor': (c_int, [GP, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_effective_tensor_noalloc_max_elements': (c_int64, []),
    'pgj_effective_tensor_noalloc': (c_int, [GP, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_avg_stress_ws_bytes': (c_size_t, [GP]),
    'pgj_hex8_avg_stress': (c_int, [GP, c_void_p, c_void_p, c_dp, c_void_p, c_void_p, c_void_p]),
    'pgj_scale_ws_bytes': (c_size_t, [GP]),
    'pgj_diagonal_scale': (c_int, [GP, c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_hex8_set_shape': (c_int, [GP, c_int, c_int, c_int, c_int, c_int, c_int]),
    'pgj_hex8_get_shape': (c_int, [GP, c_int, c_int, ctypes.POINTER(c_int)]),
    'pgj_fft_set_plan': (c_int, [GP, c_int, c_int, c_int]),
    'pgj_zpass_wide': (c_int, [GP, c_int, c_int]),
    'pgj_fft_get_plan': (c_int, [GP, c_int, ctypes.POINTER(c_int)]),
    'pgj_aos_to_soa': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p]),
    'pgj_soa_to_aos': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p]),
    'pgj_resident_scale': (c_int, [GP, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_project_ws_bytes': (c_size_t, [GP, c_int]),
    'pgj_project': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p]),
    'pgj_plan_bytes': (c_size_t, [GP]),
    'pgj_plan_init': (c_int, [GP, c_void_p, c_void_p]),
    'pgj_plan_symbol': (c_int, [GP, c_void_p, c_void_p, c_void_p]),
    'pgj_fi_coeff': (c_int, [GP, c_void_p, c_void_p, c_void_p]),
    'pgj_precond_ws_bytes': (c_size_t, [GP, c_int]),
    'pgj_scale_spectrum': (c_int, [GP, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_precond_apply_mixed': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    'pgj_precond_apply': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p]),
    'pgj_resident_state_bytes': (c_size_t, [c_int]),
    'pgj_resident_state_offsets': (c_int, [c_int, ctypes.POINTER(c_size_t)]),
    'pgj_resident_ws_bytes': (c_size_t, [GP, c_int]),
    'pgj_resident_setup': (c_int, [GP, c_int, c_int, c_int, c_int, c_void_p, c_dp, c_double, c_int64, c_void_p,
                                   c_void_p, c_void_p, c_void_p]),
    'pgj_resident_certify': (c_int, [GP, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p]),
    'pgj_resident_init': (c_int, [GP, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_resident_iterate': (c_int, [GP, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_int, c_void_p]),
    'pgj_resident_stage': (c_int, [GP, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_int, c_void_p]),
    'pgj_resident_scale_dev': (c_int, [GP, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_alpha_ws_bytes': (c_size_t, [GP]),
    'pgj_alpha_select': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_hybrid_apply': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_hybrid_apply_gather': (c_int, [GP, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_hybrid_rhs': (c_int, [GP, c_void_p, c_void_p, c_dp, c_void_p, c_void_p]),
    'pgj_hybrid_stress_ws_bytes': (c_size_t, [GP]),
    'pgj_hybrid_avg_stress': (c_int, [GP, c_void_p, c_void_p, c_void_p, c_dp, c_void_p, c_void_p, c_void_p]),
    'pgj_hybrid_resident_setup': (c_int, [GP, c_int, c_int, c_int, c_void_p, c_void_p, c_dp, c_double, c_int64,
                                          c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_hybrid_resident_certify': (c_int, [GP, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_hybrid_resident_stage': (c_int, [GP, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_int, c_void_p]),
    'pgj_hybrid_resident_iterate': (c_int, [GP, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_void_p, c_int, c_void_p]),
    'pgj_oc_ws_bytes': (c_size_t, [c_int]),
    'pgj_cone_filter': (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p]),
    'pgj_cone_project': (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                 c_double, c_double, c_int, c_void_p, c_void_p]),
    'pgj_cone_sensitivity': (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p,
                                     c_void_p]),
    'pgj_cone_ratio': (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_cone_candidate': (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    'pgj_cone_bisect': (c_int, [c_int, c_void_p, c_void_p, c_double, c_void_p]),
    'pgj_absmax_diff': (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_gauss_filter': (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_gauss_project': (c_int, [c_int, c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                                  c_void_p]),
    'pgj_gauss_sensitivity': (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_double, c_double, c_void_p,
                                      c_void_p]),
    'pgj_gauss_ratio': (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_gauss_candidate': (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    'pgj_gauss_bisect': (c_int, [c_int, c_void_p, c_void_p, c_double, c_void_p]),
    'pgj_plan32_bytes': (c_size_t, [GP]),
    'pgj_plan32_init': (c_int, [GP, c_void_p, c_void_p, c_void_p]),
    'pgj_to_f32': (c_int, [c_int64, c_void_p, c_void_p, c_void_p]),
    'pgj_inner_ws_bytes': (c_size_t, [GP, c_int]),
    'pgj_inner_begin': (c_int, [GP, c_int, c_int, c_int, c_void_p, c_void_p, c_double, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p]),
    'pgj_inner_init': (c_int, [GP, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p]),
    'pgj_inner_stage': (c_int, [GP, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                c_void_p]),
    'pgj_inner_iterate': (c_int, [GP, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                  c_void_p]),
    'pgj_inner_end': (c_int, [GP, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_pcg_ws_bytes': (c_size_t, [c_int, c_int64]),
    'pgj_pcg_start': (c_int, [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_pcg_init': (c_int, [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_pcg_pre_step': (c_int, [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_int64,
                                 c_int, c_void_p, c_void_p]),
    'pgj_pcg_post_step': (c_int, [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                  c_void_p, c_void_p]),
    'pgj_pcg32_start': (c_int, [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_pcg32_init': (c_int, [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'pgj_pcg32_pre_step': (c_int, [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_int64,
                                   c_int, c_void_p, c_void_p]),
    'pgj_pcg32_post_step': (c_int, [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                    c_void_p, c_void_p]),
}

_lib = None


def lib():
    """The loaded library; raises RuntimeError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f'libpgj.so not built at {LIB_PATH}; run __graft_entry__.build()')
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(code):
    if code != 0:
        msg = lib().pgj_last_error().decode(errors='replace')
        if code == 1:
            raise ValueError(msg)
        raise RuntimeError(msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))


def ptr(t):
    """Device pointer of a tensor (None passes NULL); requires a contiguous CUDA tensor."""
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError('contiguous device tensor required')
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def host_doubles(values):
    values = [float(v) for v in values]
    return (ctypes.c_double * len(values))(*values)


def make_grid(shape, voxel, lam, mu):
    nz, ny, nx = map(int, shape)
    hx, hy, hz = map(float, voxel)
    return Grid(nz, ny, nx, 0, hx, hy, hz, float(lam), float(mu))
