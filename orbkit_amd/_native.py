'''ctypes binding to liborbkit_amd.so (the C ABI of include/orbkit_amd.h).

The library is loaded from this package directory only; there is no other
implementation of the compute path, so every entry point raises if the
library or a HIP device is missing.
'''
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liborbkit_amd.so')

OKMI_F64, OKMI_F32 = 0, 1
PREC = {'f64': 0, 'f32': 1, 'bf16x3': 2, 'bf16x6': 3}

c_int, c_i64, c_dbl, c_vp, c_sz = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
P_int = ctypes.POINTER(ctypes.c_int)
P_dbl = ctypes.POINTER(ctypes.c_double)


class OkmiGrid(ctypes.Structure):
  _fields_ = [('d_x', c_vp), ('d_y', c_vp), ('d_z', c_vp),
              ('nx', c_i64), ('ny', c_i64), ('nz', c_i64),
              ('regular', c_int), ('pad', c_int)]


# name -> (restype, argtypes)
SIGNATURES = {
    'okmi_last_error': (ctypes.c_char_p, []),
    'okmi_basis_create': (c_int, [c_int, P_int, P_int, P_int, P_dbl, c_int, P_dbl, P_int, c_int,
                                  P_dbl, P_int, P_dbl, ctypes.POINTER(c_vp)]),
    'okmi_basis_destroy': (c_int, [c_vp]),
    'okmi_basis_nao': (c_int, [c_vp]),
    'okmi_ao_eval': (c_int, [c_vp, ctypes.POINTER(OkmiGrid), c_i64, c_i64, P_int, c_int, c_int,
                             c_vp, c_i64, c_vp]),
    'okmi_mocreator': (c_int, [c_vp, c_int, c_i64, c_i64, c_vp, c_int, c_vp, c_i64, c_vp]),
    'okmi_workspace_bytes': (c_sz, [c_vp, c_int, c_i64]),
    'okmi_mo_eval': (c_int, [c_vp, ctypes.POINTER(OkmiGrid), c_i64, c_i64, P_int, c_int, c_vp, c_int,
                             c_vp, c_i64, c_vp, c_sz, c_vp]),
    'okmi_rho': (c_int, [c_vp, ctypes.POINTER(OkmiGrid), c_i64, c_i64, c_vp, c_vp, c_int, P_int, c_int,
                         c_vp, c_vp, c_i64, c_vp, c_vp, c_sz, c_vp]),
    'okmi_rho_staged': (c_int, [c_vp, ctypes.POINTER(OkmiGrid), c_i64, c_i64, c_vp, c_vp, c_int, P_int, c_int,
                                c_vp, c_vp, c_i64, c_vp, c_vp, c_sz, c_vp]),
    'okmi_rho_noalloc_scratch_bytes': (c_sz, []),
    'okmi_rho_noalloc': (c_int, [c_vp, ctypes.POINTER(OkmiGrid), c_i64, c_i64, c_vp, c_vp, c_int, P_int, c_int,
                                 c_vp, c_vp, c_i64, c_vp, c_vp]),
    'okmi_profile_enable': (c_int, [c_int]),
    'okmi_fused_live_rows': (c_int, [c_vp, ctypes.POINTER(c_dbl), c_i64, ctypes.POINTER(c_dbl), c_i64,
                                     ctypes.POINTER(c_dbl), c_i64, c_int, ctypes.POINTER(c_dbl)]),
    'okmi_profile_query': (c_int, [ctypes.c_char_p, ctypes.POINTER(c_int), ctypes.POINTER(c_dbl)]),
    'okmi_rho_backward_workspace_bytes': (c_sz, [c_vp, c_int, c_i64]),
    'okmi_rho_backward': (c_int, [c_vp, ctypes.POINTER(OkmiGrid), c_i64, c_i64, c_vp, c_vp, c_int, c_vp, c_vp, c_vp,
                                  c_vp, c_vp, c_sz, c_vp]),
    'okmi_rho_backward_f32_workspace_bytes': (c_sz, [c_vp, c_int, c_i64]),
    'okmi_rho_backward_f32': (c_int, [c_vp, ctypes.POINTER(OkmiGrid), c_i64, c_i64, c_vp, c_vp, c_int, c_vp, c_vp,
                                      c_vp, c_vp, c_vp, c_sz, c_vp]),
    'okmi_electron_count_workspace_bytes': (c_sz, []),
    'okmi_electron_count': (c_int, [c_vp, c_i64, c_dbl, c_vp, c_sz, c_vp, c_vp]),
    'okmi_gross_atomic_workspace_bytes': (c_sz, [c_vp, c_i64, c_int, c_int]),
    'okmi_gross_atomic_density': (c_int, [c_vp, ctypes.POINTER(OkmiGrid), c_i64, c_i64, c_int, c_vp, c_vp, c_int,
                                          P_int, P_int, c_int, c_vp, c_i64, c_vp, c_sz, c_vp]),
    'okmi_masked_mo_product': (c_int, [c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_i64, c_i64, c_vp, c_int, c_vp, c_vp]),
    'okmi_mo_pair_products': (c_int, [c_vp, c_int, c_vp, c_int, c_int, c_i64, c_i64, c_vp, c_i64, c_vp]),
    'okmi_transition_flux': (c_int, [c_vp, c_int, c_int, c_i64, c_i64, c_vp, c_int, c_vp, c_i64, c_vp]),
    'okmi_cube_format': (c_sz, [ctypes.POINTER(c_dbl), c_int, c_i64, c_i64, c_i64, ctypes.c_char_p, c_sz]),
    'okmi_ci_workspace_bytes': (c_sz, [c_int]),
    'okmi_ci_reduce': (c_int, [c_vp, c_vp, c_int, c_int, c_i64, c_i64, ctypes.POINTER(c_dbl), P_int, P_int, c_int,
                               c_int, c_vp, c_i64, c_vp, c_sz, c_vp]),
    'okmi_batch_create': (c_int, [c_int, ctypes.POINTER(c_vp), ctypes.POINTER(OkmiGrid), ctypes.POINTER(c_vp),
                                  ctypes.POINTER(c_vp), P_int, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), c_int, c_int,
                                  ctypes.POINTER(c_vp)]),
    'okmi_batch_tiles': (c_i64, [c_vp]),
    'okmi_batch_set_workgroups': (c_int, [c_vp, c_int]),
    'okmi_batch_set_tile_shape': (c_int, [c_vp, c_int, c_int, c_int]),
    'okmi_batch_launch': (c_int, [c_vp, c_int, c_vp]),
    'okmi_batch_destroy': (c_int, [c_vp]),
    'okmi_rho_fused_workspace_bytes': (c_sz, [c_vp, c_int, c_int, c_i64]),
    'okmi_rho_fused': (c_int, [c_vp, ctypes.POINTER(OkmiGrid), c_i64, c_i64, c_vp, c_vp, c_int, P_int, c_int,
                               c_int, c_vp, c_vp, c_i64, c_vp, c_sz, c_vp]),
}

_lib = None


def lib():
  '''Loads the library once; raises RuntimeError if it has not been built.'''
  global _lib
  if _lib is None:
    if not os.path.exists(LIB_PATH):
      raise RuntimeError('liborbkit_amd.so is not built (run `python -m orbkit_amd.build`); '
                         'orbkit_amd has no other compute path')
    l = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
      f = getattr(l, name)
      f.restype = res
      f.argtypes = args
    _lib = l
  return _lib


def check(rc, what=''):
  if rc != 0:
    msg = lib().okmi_last_error().decode()
    raise RuntimeError('%s failed (status %d): %s' % (what or 'orbkit_amd', rc, msg))


def int_array(a):
  import numpy
  a = numpy.ascontiguousarray(a, dtype=numpy.intc)
  return a, a.ctypes.data_as(P_int)


def dbl_array(a):
  import numpy
  a = numpy.ascontiguousarray(a, dtype=numpy.float64)
  return a, a.ctypes.data_as(P_dbl)
