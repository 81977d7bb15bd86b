'''Fused density path (okmi_rho_fused): AO evaluation, MO contraction and
density assembly in one kernel, the AO values never stored in HBM.

``precision``: 'f64' (FP64 AO values and FP64 MFMA, FP64 accumulation; up to
64 MOs), 'f32' (FP32 MFMA), 'bf16x3' (two-part bf16 split, 3 MFMA products)
or 'bf16x6' (three-part split, 6 products); the reduced precisions
accumulate in FP32, every precision assembles the density in FP64.  See
include/orbkit_amd.h.
'''
import ctypes

import numpy

from . import _native

#: device bytes for the fused path's per-point term buffers
WORKSPACE_BYTES = 256 << 20
#: points per pass of the many-MO path (more than 64 MOs): AO sets of a pass
#: are staged in HBM (FP32), e.g. 2,950 AOs x 7 sets x 131,072 points = 10.8 GB
MANY_MO_CHUNK = 131072


class FusedPlan:
  '''Device-resident inputs and outputs for repeated fused evaluations
  (MO coefficients, occupations, output and workspace buffers).'''

  def __init__(self, basis, dgrid, coeffs, occ, codes, precision='bf16x6', p0=0, np_=None):
    import torch
    if precision not in ('f64', 'f32', 'bf16x3', 'bf16x6'):
      raise ValueError('precision must be f64, f32, bf16x3 or bf16x6, not %r' % (precision,))
    self.L = _native.lib()
    self.basis, self.dgrid = basis, dgrid
    self.prec = _native.PREC[precision]
    self.p0 = p0
    self.np = dgrid.npts - p0 if np_ is None else np_
    dev = dgrid.x.device
    C = numpy.ascontiguousarray(coeffs, dtype=numpy.float64)
    if C.ndim != 2 or C.shape[1] != basis.nao:
      raise ValueError('MO coefficients have shape %s but the basis has %d AOs' % (C.shape, basis.nao))
    self.nmo = C.shape[0]
    self.C = torch.as_tensor(C, device=dev)
    self.occ = torch.as_tensor(numpy.ascontiguousarray(occ, dtype=numpy.float64), device=dev)
    self.codes = numpy.ascontiguousarray(codes, dtype=numpy.intc)
    self.nd = len(self.codes)
    n = max(self.np, 1)
    self.rho = torch.empty(n, dtype=torch.float64, device=dev)
    self.delta = torch.empty((max(self.nd, 1), n), dtype=torch.float64, device=dev)
    chunk = min(n, MANY_MO_CHUNK if self.nmo > 64 else max(64, WORKSPACE_BYTES // 160))
    nbytes = self.L.okmi_rho_fused_workspace_bytes(basis.handle, self.nmo, self.prec, chunk)
    self.ws = torch.empty(nbytes // 8 + 1, dtype=torch.float64, device=dev)

  def launch(self, stream=None):
    import torch
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream if stream is None else stream)
    _native.check(self.L.okmi_rho_fused(
        self.basis.handle, ctypes.byref(self.dgrid.desc), self.p0, self.np,
        ctypes.c_void_p(self.C.data_ptr()), ctypes.c_void_p(self.occ.data_ptr()), self.nmo,
        self.codes.ctypes.data_as(_native.P_int), self.nd, self.prec,
        ctypes.c_void_p(self.rho.data_ptr()), ctypes.c_void_p(self.delta.data_ptr()),
        self.delta.shape[1], ctypes.c_void_p(self.ws.data_ptr()), self.ws.numel() * 8, st),
        'okmi_rho_fused')

  def result(self):
    return self.rho[:self.np].cpu().numpy(), self.delta[:self.nd, :self.np].cpu().numpy()


def rho_fused(basis, dgrid, coeffs, occ, codes, precision='bf16x6', device=False):
  '''Density (np,) and derivative rows (len(codes), np) on the whole grid
  (device tensors with ``device``, else host arrays).'''
  plan = FusedPlan(basis, dgrid, coeffs, occ, codes, precision)
  plan.launch()
  if device:
    return plan.rho[:plan.np], plan.delta[:plan.nd, :plan.np]
  return plan.result()
