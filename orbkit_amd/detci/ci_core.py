'''CI (transition) densities and fluxes on the grid (orbkit/detci/ci_core.py).

``rho``, ``jab`` and ``a_nabla_b`` take orbkit's arguments: ``zero`` (two
lists: prefactors and occupied-orbital indices of identical determinants),
``sing`` (two lists: coefficient products and orbital pairs of effective
single excitations) and MO values on a grid (``molist`` (NMO,) + N,
``molistdrv`` (3, NMO) + N).  The reduction runs on the device
(okmi_ci_reduce); ``slice_length`` and ``numproc`` are accepted for
signature compatibility (every point is independent on the device).

``get_rho_full``, ``get_j_full`` and ``get_jab_full`` mirror
orbkit/detci/cy_ci.pyx:96-193: the time-dependent combinations are FP64
GEMMs (okmi_mocreator) and the basis-level flux a quadratic form.
'''
import ctypes

import numpy

from .. import _native


def _torch():
  import torch
  if not torch.cuda.is_available():
    raise RuntimeError('orbkit_amd needs a HIP device (torch.cuda.is_available() is False); '
                       'it has no CPU compute path')
  return torch


def _stream():
  return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _terms(zero, sing):
  '''(coef, a, b) arrays: zero terms (a = b) first, then single excitations.'''
  c, a, b = [], [], []
  if zero is not None and len(zero) == 2:
    for pref, occ in zip(zero[0], zero[1]):
      for cz, iz in zip(pref, occ):
        c.append(float(cz)); a.append(int(iz)); b.append(int(iz))
  sc, sa, sb = [], [], []
  if sing is not None and len(sing) == 2:
    for cs, (i, j) in zip(sing[0], sing[1]):
      sc.append(float(cs)); sa.append(int(i)); sb.append(int(j))
  return c, a, b, sc, sa, sb


def _reduce(molist, molistdrv, c, a, b, mode, nd=3):
  torch = _torch()
  L = _native.lib()
  mo = numpy.ascontiguousarray(molist, dtype=numpy.float64)
  shape = mo.shape
  mo = mo.reshape(shape[0], -1)
  npts = mo.shape[1]
  dev = torch.device('cuda', torch.cuda.current_device())
  dmo = torch.as_tensor(mo, device=dev)
  ddrv = None
  if molistdrv is not None:
    drv = numpy.ascontiguousarray(molistdrv, dtype=numpy.float64).reshape(nd, shape[0], -1)
    ddrv = torch.as_tensor(drv, device=dev)
  nout = 1 if mode < 0 else nd
  out = torch.empty((nout, max(npts, 1)), dtype=torch.float64, device=dev)
  coef = numpy.ascontiguousarray(c, dtype=numpy.float64)
  ia = numpy.ascontiguousarray(a, dtype=numpy.intc)
  ib = numpy.ascontiguousarray(b, dtype=numpy.intc)
  ws = torch.empty(L.okmi_ci_workspace_bytes(len(coef)) // 8 + 1, dtype=torch.float64, device=dev)
  _native.check(L.okmi_ci_reduce(ctypes.c_void_p(dmo.data_ptr()),
                                 ctypes.c_void_p(ddrv.data_ptr()) if ddrv is not None else None,
                                 shape[0], nd, npts, npts,
                                 coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ia.ctypes.data_as(_native.P_int),
                                 ib.ctypes.data_as(_native.P_int), len(coef), mode,
                                 ctypes.c_void_p(out.data_ptr()), out.shape[1],
                                 ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream()), 'okmi_ci_reduce')
  r = out[:, :npts].cpu().numpy()
  return r[0].reshape(shape[1:]) if mode < 0 else r.reshape((nd,) + shape[1:])


def rho(zero, sing, molist, slice_length=1e4, numproc=1):
  '''Electron (transition) density sum_zero c phi_a^2 + sum_sing c phi_a phi_b.'''
  c, a, b, sc, sa, sb = _terms(zero, sing)
  return _reduce(molist, None, c + sc, a + sa, b + sb, -1)


def jab(zero, sing, molist, molistdrv, slice_length=1e4, numproc=1):
  '''Imaginary part of the electronic (transition) flux density,
  -1/2 sum_sing c (phi_a grad phi_b - phi_b grad phi_a), shape (3,) + N.'''
  _, _, _, sc, sa, sb = _terms(zero, sing)
  return _reduce(molist, molistdrv, sc, sa, sb, 0)


def a_nabla_b(zero, sing, molist, molistdrv, slice_length=1e4, numproc=1):
  '''sum_sing c phi_a grad phi_b, shape (3,) + N (orbkit uses the single
  excitations only, orbkit/detci/cy_ci.pyx:222-227).'''
  _, _, _, sc, sa, sb = _terms(zero, sing)
  return _reduce(molist, molistdrv, sc, sa, sb, 1)


def get_rho_full(ReS, rho):
  '''tdrho[t] = sum_{n <= m} (ReS[t,m,m] if m == n else 2 ReS[t,m,n]) rho[count]
  (orbkit/detci/cy_ci.pyx:96-117), count running over (n, m >= n).'''
  from ..core import gemm
  ReS = numpy.asarray(ReS, dtype=numpy.float64)
  nt, ns = ReS.shape[0], ReS.shape[1]
  W = numpy.zeros((nt, ns * (ns + 1) // 2))
  count = 0
  for n in range(ns):
    for m in range(n, ns):
      W[:, count] = ReS[:, m, m] if m == n else 2.0 * ReS[:, m, n]
      count += 1
  return gemm(W, numpy.asarray(rho, dtype=numpy.float64))


def get_j_full(ImS, j):
  '''tdj[t, d] = -sum_{n < m} 2 ImS[t,n,m] j[count, d] (orbkit/detci/cy_ci.pyx:119-141).'''
  from ..core import gemm
  ImS = numpy.asarray(ImS, dtype=numpy.float64)
  j = numpy.asarray(j, dtype=numpy.float64)
  nt, ns = ImS.shape[0], ImS.shape[1]
  W = numpy.zeros((nt, ns * (ns + 1) // 2))
  count = 0
  for n in range(ns):
    for m in range(n, ns):
      if m != n:
        W[:, count] = -2.0 * ImS[:, n, m]
      count += 1
  return numpy.stack([gemm(W, numpy.ascontiguousarray(j[:, d])) for d in range(3)], axis=1)


def get_jab_full(ImS, chi_n, nabla_chi_n, mu):
  '''j[c] = sum_{n > m} ImS[n,m] / mu (chi_n d_c chi_m - chi_m d_c chi_n)
  (orbkit/detci/cy_ci.pyx:168-193).'''
  ImS = numpy.asarray(ImS, dtype=numpy.float64)
  nabla = numpy.asarray(nabla_chi_n, dtype=numpy.float64)
  pairs = [(n, m) for n in range(ImS.shape[0]) for m in range(n)]
  c = [-2.0 * ImS[n, m] / mu for n, m in pairs]
  return _reduce(chi_n, nabla, c, [n for n, _ in pairs], [m for _, m in pairs], 0, nd=nabla.shape[0])
