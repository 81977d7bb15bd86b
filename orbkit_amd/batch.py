'''Density (and gradient) of many molecules in one launch (okmi_batch_*).

orbkit computes one molecule per ``rho_compute`` call (orbkit/core.py:314);
``BatchPlan`` binds many (qc, grid) pairs to one native batch whose launch
runs the fused reduced-precision kernel over the tiles of every molecule
(one persistent kernel, optionally replayed from a HIP graph).

    plan = BatchPlan([(qc0, axes0), (qc1, axes1), ...], precision='bf16x6')
    plan.launch()                       # or plan.launch(use_graph=True)
    rho, grad = plan.result(0)          # molecule 0: N and (3,) + N
'''
import ctypes

import numpy

from . import _native
from .basis import DeviceBasis, flatten
from .core import DeviceGrid


class BatchPlan:
  '''Device-resident inputs and outputs of a batch of molecules.

  ``molecules``: sequence of ``(qc, (x, y, z))`` (regular grid axes) or
  ``(qc, (x, y, z), is_vector)``.'''

  def __init__(self, molecules, precision='bf16x6', grad=True, order_by_cost=True):
    import torch
    if not torch.cuda.is_available():
      raise RuntimeError('orbkit_amd needs a HIP device (torch.cuda.is_available() is False); '
                         'it has no CPU compute path')
    if precision not in ('f32', 'bf16x3', 'bf16x6'):
      raise ValueError('precision must be f32, bf16x3 or bf16x6, not %r' % (precision,))
    self.L = _native.lib()
    self.precision = precision
    self.grad = bool(grad)
    dev = torch.device('cuda', torch.cuda.current_device())
    self.bases, self.grids, self.C, self.occ, self.rho, self.dgrad, self.shapes = [], [], [], [], [], [], []
    for item in molecules:
      qc, axes = item[0], item[1]
      is_vector = bool(item[2]) if len(item) > 2 else False
      b = DeviceBasis(flatten(qc.geo_spec, qc.ao_spec))
      g = DeviceGrid(*axes, regular=not is_vector)
      C = numpy.ascontiguousarray(qc.mo_spec.get_coeffs(), dtype=numpy.float64)
      if C.shape[1] != b.nao:
        raise ValueError('MO coefficients have shape %s but the basis has %d AOs' % (C.shape, b.nao))
      self.bases.append(b)
      self.grids.append(g)
      self.C.append(torch.as_tensor(C, device=dev))
      self.occ.append(torch.as_tensor(numpy.ascontiguousarray(qc.mo_spec.get_occ(), dtype=numpy.float64),
                                      device=dev))
      self.rho.append(torch.empty(g.npts, dtype=torch.float64, device=dev))
      self.dgrad.append(torch.empty((3, g.npts), dtype=torch.float64, device=dev) if self.grad else None)
      self.shapes.append(g.shape)
    n = len(self.bases)
    self.nmol = n
    self.npts = sum(g.npts for g in self.grids)
    P = lambda xs: (ctypes.c_void_p * n)(*xs)
    self._keep = dict(
        bases=P([b.handle.value for b in self.bases]),
        grids=(_native.OkmiGrid * n)(*[g.desc for g in self.grids]),
        C=P([t.data_ptr() for t in self.C]), occ=P([t.data_ptr() for t in self.occ]),
        nmo=(ctypes.c_int * n)(*[t.shape[0] for t in self.C]),
        rho=P([t.data_ptr() for t in self.rho]),
        grad=P([t.data_ptr() for t in self.dgrad]) if self.grad else None)
    k = self._keep
    h = ctypes.c_void_p()
    _native.check(self.L.okmi_batch_create(n, k['bases'], k['grids'], k['C'], k['occ'], k['nmo'], k['rho'],
                                           k['grad'], _native.PREC[precision], int(bool(order_by_cost)),
                                           ctypes.byref(h)), 'okmi_batch_create')
    self.handle = h
    self.ntile = int(self.L.okmi_batch_tiles(h))
    self.workgroups = 0
    self.tile_shape = (2, 2, 2)

  def __del__(self):
    h = getattr(self, 'handle', None)
    if h is not None and h.value:
      try:
        self.L.okmi_batch_destroy(h)
      except Exception:
        pass
      self.handle = None

  def set_workgroups(self, n):
    '''Persistent workgroup count (0: one workgroup per 64-point tile).'''
    _native.check(self.L.okmi_batch_set_workgroups(self.handle, int(n)), 'okmi_batch_set_workgroups')
    self.workgroups = int(n)

  def set_tile_shape(self, lx, ly, lz):
    '''Regular-grid tiles of 2^lx x 2^ly x 2^lz points (lx + ly + lz = 6).'''
    _native.check(self.L.okmi_batch_set_tile_shape(self.handle, int(lx), int(ly), int(lz)),
                  'okmi_batch_set_tile_shape')
    self.tile_shape = (int(lx), int(ly), int(lz))
    self.ntile = int(self.L.okmi_batch_tiles(self.handle))

  def launch(self, use_graph=False, stream=None):
    import torch
    st = torch.cuda.current_stream().cuda_stream if stream is None else stream
    if use_graph and not st:
      raise ValueError('graph launches need a non-default stream (e.g. torch.cuda.Stream())')
    _native.check(self.L.okmi_batch_launch(self.handle, int(bool(use_graph)), ctypes.c_void_p(st)),
                  'okmi_batch_launch')

  def result(self, m):
    '''(rho N, grad (3,)+N) of molecule m (grad None without gradients).'''
    rho = self.rho[m].cpu().numpy().reshape(self.shapes[m])
    if not self.grad:
      return rho, None
    return rho, self.dgrad[m].cpu().numpy().reshape((3,) + self.shapes[m])
