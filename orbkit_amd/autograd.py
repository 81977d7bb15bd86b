'''Differentiable electron density on the device.

``DensityModel`` evaluates rho(r) = sum_i occ_i (sum_mu C_i,mu chi_mu(r))^2 on
a fixed grid (the forward of orbkit/core.py:rho_compute,
orbkit/core.py:529-607) and differentiates it with respect to the MO
coefficients C, the primitive exponents alpha and the nuclear positions R.
The backward pass is native: AO values and their gradients are recomputed
per chunk of points, dL/dC is an MFMA contraction over points and the
exponent / position gradients are reduced per shell in a fixed order.

``precision``:
  'f64'  FP64 throughout: forward okmi_rho, backward okmi_rho_backward;
         parameters and gradients are float64 tensors.
  'f32'  FP32: forward okmi_rho_fused with the FP32 MFMA contraction (FP64
         exp arguments, FP64 density assembly), backward
         okmi_rho_backward_f32 (FP32 values and MFMA products; FP64 grid
         geometry, exp arguments and reductions); parameters, rho and
         gradients are float32 tensors.

    model = DensityModel(qc, axes)                  # regular grid axes (or x, y, z with is_vector)
    C, alpha, geo = model.parameters()              # leaf tensors on the device
    rho = model(C, alpha, geo)                      # (npts,) in the model's precision
    loss = ((rho - target) ** 2 * w).sum()
    loss.backward()                                 # C.grad, alpha.grad, geo.grad

orbkit has no derivative code; the FP64 forward values equal okmi_rho's (and
so orbkit's to 1e-12).  Row factors of the basis (orbkit's 'N' row
renormalisation, the Cartesian-to-spherical map, contraction coefficients)
are constants of the model.
'''
import ctypes

import numpy

from . import _native
from .basis import DeviceBasis, flatten
from .core import DeviceGrid, _stream

#: device workspace of the backward pass (per chunk of points: AO values, MO
#: values, dL/dA, about 7 KB per point at 400 AOs and 40 MOs in FP64): 4 GiB
#: holds a 512,000-point grid in one chunk, enough waves to fill the GPU
WORKSPACE_BYTES = 4 << 30
PRECISIONS = ('f64', 'f32')


def _torch():
  import torch
  if not torch.cuda.is_available():
    raise RuntimeError('orbkit_amd needs a HIP device (torch.cuda.is_available() is False); '
                       'it has no CPU compute path')
  return torch


class DensityModel:
  '''rho on a fixed grid as a differentiable function of (C, alpha, geo).'''

  def __init__(self, qc, axes=None, is_vector=False, occ=None, workspace_bytes=WORKSPACE_BYTES, precision='f64'):
    torch = _torch()
    if precision not in PRECISIONS:
      raise ValueError('precision must be one of %s, not %r' % (PRECISIONS, precision))
    self.precision = precision
    self.dtype = torch.float64 if precision == 'f64' else torch.float32
    self.flat = flatten(qc.geo_spec, qc.ao_spec)
    if axes is None:
      from . import grid
      axes = (grid.x, grid.y, grid.z)
      is_vector = grid.is_vector
    self.grid = DeviceGrid(*axes, regular=not is_vector)
    self.npts = self.grid.npts
    dev = self.grid.x.device
    self.device = dev
    occ = qc.mo_spec.get_occ() if occ is None else occ
    self.occ = torch.as_tensor(numpy.ascontiguousarray(occ, dtype=numpy.float64), device=dev)
    self.occ_r = self.occ.to(self.dtype)
    self.C0 = numpy.ascontiguousarray(qc.mo_spec.get_coeffs(), dtype=numpy.float64)
    self.nprim = len(self.flat['prims'])
    self.natoms = len(self.flat['geo'])
    self.ws_bytes = int(workspace_bytes)
    self._ws = None
    self._basis_key = None
    self._basis = None

  def parameters(self, requires_grad=True):
    '''Leaf tensors (C, alpha, geo) holding the model's initial values.'''
    torch = _torch()
    t = lambda a: torch.tensor(numpy.array(a, dtype=numpy.float64), device=self.device, dtype=self.dtype,
                               requires_grad=requires_grad)
    return t(self.C0), t(self.flat['prims'][:, 0]), t(self.flat['geo'])

  def basis(self, alpha, geo):
    '''DeviceBasis with these exponents and centres (rebuilt when they change).'''
    a = numpy.ascontiguousarray(alpha.detach().cpu().numpy() if hasattr(alpha, 'detach') else alpha,
                                dtype=numpy.float64).reshape(-1)
    g = numpy.ascontiguousarray(geo.detach().cpu().numpy() if hasattr(geo, 'detach') else geo,
                                dtype=numpy.float64).reshape(-1, 3)
    if a.shape != (self.nprim,) or g.shape != (self.natoms, 3):
      raise ValueError('alpha must have %d entries and geo shape (%d, 3)' % (self.nprim, self.natoms))
    if numpy.any(a <= 0):
      raise ValueError('primitive exponents must be positive')
    key = (a.tobytes(), g.tobytes())
    if key != self._basis_key:
      flat = dict(self.flat)
      prims = numpy.array(self.flat['prims'], dtype=numpy.float64, copy=True)
      prims[:, 0] = a
      flat['prims'] = prims
      flat['geo'] = g
      self._basis = DeviceBasis(flat)
      self._basis_key = key
    return self._basis

  def workspace(self):
    torch = _torch()
    if self._ws is None:
      self._ws = torch.empty(self.ws_bytes // 8, dtype=torch.float64, device=self.device)
    return self._ws

  def rho(self, C, alpha, geo):
    '''Forward only (no graph): density (npts,) in the model's precision.'''
    torch = _torch()
    L = _native.lib()
    b = self.basis(alpha, geo)
    # C enters the native call in FP64 (the fused FP32 path rounds it to FP32
    # itself, so an FP32 C passes through exactly)
    C = C.detach().to(device=self.device, dtype=torch.float64).contiguous()
    nmo = C.shape[0]
    if C.shape != (nmo, b.nao) or nmo != self.occ.numel():
      raise ValueError('C has shape %s; expected (%d, %d)' % (tuple(C.shape), self.occ.numel(), b.nao))
    rho = torch.empty(self.npts, dtype=torch.float64, device=self.device)
    ws = self.workspace()
    args = (b.handle, ctypes.byref(self.grid.desc), 0, self.npts, ctypes.c_void_p(C.data_ptr()),
            ctypes.c_void_p(self.occ.data_ptr()), nmo)
    if self.precision == 'f64':
      _native.check(L.okmi_rho(*args, None, 0, ctypes.c_void_p(rho.data_ptr()), None, self.npts, None,
                               ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream()), 'okmi_rho')
      return rho
    _native.check(L.okmi_rho_fused(*args, None, 0, _native.PREC['f32'], ctypes.c_void_p(rho.data_ptr()), None,
                                   self.npts, ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream()),
                  'okmi_rho_fused')
    return rho.to(torch.float32)

  def vjp(self, C, alpha, geo, grad_rho):
    '''(dL/dC, dL/dalpha, dL/dgeo) for dL/drho = grad_rho.'''
    torch = _torch()
    L = _native.lib()
    b = self.basis(alpha, geo)
    C = C.detach().to(device=self.device, dtype=self.dtype).contiguous()
    g = grad_rho.detach().to(device=self.device, dtype=self.dtype).contiguous().reshape(-1)
    if g.numel() != self.npts:
      raise ValueError('grad_rho has %d entries for %d grid points' % (g.numel(), self.npts))
    nmo = C.shape[0]
    if C.shape != (nmo, b.nao) or nmo != self.occ.numel():
      raise ValueError('C has shape %s; expected (%d, %d)' % (tuple(C.shape), self.occ.numel(), b.nao))
    dC = torch.zeros_like(C)
    da = torch.zeros(self.nprim, dtype=self.dtype, device=self.device)
    dg = torch.zeros((self.natoms, 3), dtype=self.dtype, device=self.device)
    ws = self.workspace()
    fn, name = ((L.okmi_rho_backward, 'okmi_rho_backward') if self.precision == 'f64'
                else (L.okmi_rho_backward_f32, 'okmi_rho_backward_f32'))
    _native.check(fn(b.handle, ctypes.byref(self.grid.desc), 0, self.npts,
                     ctypes.c_void_p(C.data_ptr()), ctypes.c_void_p(self.occ_r.data_ptr()), nmo,
                     ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(dC.data_ptr()),
                     ctypes.c_void_p(da.data_ptr()), ctypes.c_void_p(dg.data_ptr()),
                     ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream()), name)
    return dC, da, dg

  def __call__(self, C, alpha, geo):
    return _DensityFunction.apply(C, alpha, geo, self)


def _function_class():
  torch = _torch()

  class DensityFunction(torch.autograd.Function):
    '''rho(C, alpha, geo) with the native backward.'''

    @staticmethod
    def forward(ctx, C, alpha, geo, model):
      ctx.model = model
      ctx.save_for_backward(C, alpha, geo)
      return model.rho(C, alpha, geo)

    @staticmethod
    def backward(ctx, grad_rho):
      C, alpha, geo = ctx.saved_tensors
      dC, da, dg = ctx.model.vjp(C, alpha, geo, grad_rho)
      return (dC.to(C.device, C.dtype) if ctx.needs_input_grad[0] else None,
              da.to(alpha.device, alpha.dtype) if ctx.needs_input_grad[1] else None,
              dg.to(geo.device, geo.dtype).reshape(geo.shape) if ctx.needs_input_grad[2] else None,
              None)

  return DensityFunction


class _Lazy:
  '''Defers building the autograd.Function subclass until first use.'''
  _cls = None

  def apply(self, *args):
    if _Lazy._cls is None:
      _Lazy._cls = _function_class()
    return _Lazy._cls.apply(*args)


_DensityFunction = _Lazy()
