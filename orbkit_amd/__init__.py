'''MI355X-native implementation of orbkit's grid-density path.

``orbkit_amd.core`` exposes ``ao_creator``, ``mo_creator``,
``cartesian2spherical``, ``slice_rho``, ``rho_compute`` and
``rho_compute_no_slice`` with orbkit's signatures; the work runs in HIP
kernels for gfx950 reached through the C ABI in ``liborbkit_amd.so``.
'''
from . import grid, tools, units
from .qcinfo import QCinfo
from .orbitals import AOClass, MOClass
from . import read

__all__ = ['grid', 'tools', 'units', 'QCinfo', 'AOClass', 'MOClass', 'read', 'core']


def __getattr__(name):
  if name == 'core':
    import importlib
    return importlib.import_module(__name__ + '.core')
  raise AttributeError(name)
