'''Drop-in for orbkit's grid path (orbkit/core.py:38-839) running on MI355X.

Same names, signatures, defaults, return shapes and exceptions as orbkit's
``ao_creator``, ``mo_creator``, ``cartesian2spherical``, ``slice_rho``,
``initializer``, ``rho_compute`` and ``rho_compute_no_slice``.  The numbers
come from HIP kernels reached through liborbkit_amd.so; inputs are host
numpy arrays and so are the results (float64, C order, owned by the caller).

Differences from orbkit, all on the execution side:
  * ``numproc`` is accepted and ignored (one GPU does the slices); a value
    <= 0 still selects ``rho_compute_no_slice`` as in orbkit.
  * ``slice_length`` keeps its meaning for the printed slice plan; the device
    processes points in chunks sized to its memory, which gives the same
    values because every point is independent.
  * integer codes in ``drv`` are accepted in the density branch too (orbkit
    raises TypeError there because it calls ``len()`` on them).
  * ``drv`` lists of any length are taken, as in orbkit; more than 10 codes
    (AO / MO values) or 16 density outputs run as several native calls.
  * extra keyword ``precision`` of ``rho_compute``: 'f64' (default) or one of
    the reduced-precision fused paths ('f32', 'bf16x3', 'bf16x6').
'''
import ctypes

import numpy

from . import grid, _native
from .basis import device_basis, spherical_matrix
from .display import display
from .qcinfo import QCinfo
from .tools import validate_drv, require, convert

#: device bytes for one staged chunk of AO values (FP64 path)
WORKSPACE_BYTES = 1 << 30
#: device bytes for one chunk of returned values
OUTPUT_CHUNK_BYTES = 1 << 30
#: derivative codes per native AO / MO call and density outputs per native
#: density call; longer ``drv`` lists (orbkit takes any length) are split
MAX_CODES_PER_CALL = 10
MAX_OUTPUTS_PER_CALL = 16
#: points per block when a result already on the host is written to a file
OUTPUT_SLICE = 1 << 16


# --------------------------------------------------------------------------
# device plumbing
def _torch():
  import torch
  if not torch.cuda.is_available():
    raise RuntimeError('orbkit_amd needs a HIP device (torch.cuda.is_available() is False); '
                       'it has no CPU compute path')
  _native.lib()
  return torch


def _stream():
  torch = _torch()
  return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class DeviceGrid:
  '''Points on the device: a regular grid (axes only) or a vector grid.'''

  def __init__(self, x, y, z, regular):
    torch = _torch()
    dev = torch.device('cuda', torch.cuda.current_device())
    t = lambda a: torch.as_tensor(numpy.ascontiguousarray(a, dtype=numpy.float64).reshape(-1), device=dev)
    self.x, self.y, self.z = t(x), t(y), t(z)
    self.regular = bool(regular)
    if self.regular:
      self.shape = (len(self.x), len(self.y), len(self.z))
      self.npts = self.shape[0] * self.shape[1] * self.shape[2]
    else:
      if not (len(self.x) == len(self.y) == len(self.z)):
        raise ValueError('Dimensions of x-, y-, and z- coordinate differ!')
      self.shape = (len(self.x),)
      self.npts = len(self.x)
    g = _native.OkmiGrid()
    g.d_x, g.d_y, g.d_z = self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr()
    if self.regular:
      g.nx, g.ny, g.nz, g.regular = self.shape[0], self.shape[1], self.shape[2], 1
    else:
      g.nx, g.ny, g.nz, g.regular = self.npts, 0, 0, 0
    self.desc = g


def _codes(drv_list):
  return numpy.array([validate_drv(d) for d in drv_list], dtype=numpy.intc)


def _fill(out, blocks):
  """Copies each (first point, host block) into out[..., range] and passes
  it on (for a file sink that writes the same blocks)."""
  for c0, b in blocks:
    out[..., c0:c0 + b.shape[-1]] = b
    yield c0, b


def _ao_blocks(basis, dgrid, codes, p0, np_):
  """AO values (ncodes, nao, n) of consecutive point chunks, each copied to
  the host once: (first point relative to p0, block)."""
  torch = _torch()
  L = _native.lib()
  nd = len(codes)
  chunk = max(64, min(np_, OUTPUT_CHUNK_BYTES // max(1, nd * basis.nao * 8)))
  buf = torch.empty((nd, basis.nao, min(chunk, max(np_, 1))), dtype=torch.float64,
                    device=dgrid.x.device)
  for c0 in range(0, np_, chunk):
    n = min(chunk, np_ - c0)
    _native.check(L.okmi_ao_eval(basis.handle, ctypes.byref(dgrid.desc), p0 + c0, n,
                                 codes.ctypes.data_as(_native.P_int), nd, _native.OKMI_F64,
                                 ctypes.c_void_p(buf.data_ptr()), buf.shape[2], _stream()),
                  'okmi_ao_eval')
    yield c0, buf[:, :, :n].cpu().numpy()


def _mo_blocks(basis, dgrid, codes, C, p0, np_):
  """MO values (ncodes, nmo, n) of consecutive point chunks, as _ao_blocks."""
  torch = _torch()
  L = _native.lib()
  nd, nmo = len(codes), C.shape[0]
  dev = dgrid.x.device
  dC = torch.as_tensor(C, device=dev)
  chunk = max(64, min(np_, OUTPUT_CHUNK_BYTES // max(1, nd * nmo * 8)))
  buf = torch.empty((nd, nmo, min(chunk, max(np_, 1))), dtype=torch.float64, device=dev)
  ws_chunk = max(64, WORKSPACE_BYTES // max(1, nd * basis.nao * 8))
  ws = torch.empty(L.okmi_workspace_bytes(basis.handle, nd, min(ws_chunk, max(chunk, 64))) // 8 + 1,
                   dtype=torch.float64, device=dev)
  for c0 in range(0, np_, chunk):
    n = min(chunk, np_ - c0)
    _native.check(L.okmi_mo_eval(basis.handle, ctypes.byref(dgrid.desc), p0 + c0, n,
                                 codes.ctypes.data_as(_native.P_int), nd, ctypes.c_void_p(dC.data_ptr()),
                                 nmo, ctypes.c_void_p(buf.data_ptr()), buf.shape[2],
                                 ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream()),
                  'okmi_mo_eval')
    yield c0, buf[:, :, :n].cpu().numpy()


def _consume(out, blocks, sink):
  """Fills out from blocks; sink(shape, blocks) writes them to a file too."""
  blocks = _fill(out, blocks)
  if sink is None:
    for _ in blocks:
      pass
  else:
    sink(out.shape, blocks)
  return out


def eval_aos(basis, dgrid, codes, p0=0, np_=None, out=None, sink=None):
  '''AO values for derivative codes on points [p0, p0+np): host array
  (ncodes, nao, np).  sink(shape, blocks): optional writer of the same
  blocks (e.g. a result file), called with a generator of (first point,
  host block) in point order.'''
  np_ = dgrid.npts - p0 if np_ is None else np_
  codes = numpy.ascontiguousarray(codes, dtype=numpy.intc)
  nd = len(codes)
  if out is None:
    out = numpy.empty((nd, basis.nao, np_))
  if nd > MAX_CODES_PER_CALL:
    for g in range(0, nd, MAX_CODES_PER_CALL):
      eval_aos(basis, dgrid, codes[g:g + MAX_CODES_PER_CALL], p0, np_, out=out[g:g + MAX_CODES_PER_CALL])
    if sink is not None:   # all codes of a point range together: from the host array
      sink(out.shape, ((c0, out[..., c0:c0 + OUTPUT_SLICE]) for c0 in range(0, np_, OUTPUT_SLICE)))
    return out
  return _consume(out, _ao_blocks(basis, dgrid, codes, p0, np_), sink)


def eval_mos(basis, dgrid, codes, coeffs, p0=0, np_=None, sink=None):
  '''MO values (ncodes, nmo, np) for derivative codes (sink as eval_aos).'''
  np_ = dgrid.npts - p0 if np_ is None else np_
  codes = numpy.ascontiguousarray(codes, dtype=numpy.intc)
  nd = len(codes)
  C = numpy.ascontiguousarray(coeffs, dtype=numpy.float64)
  nmo = C.shape[0]
  if C.ndim != 2 or C.shape[1] != basis.nao:
    raise ValueError('MO coefficients have shape %s but the basis has %d AOs' % (C.shape, basis.nao))
  if nd > MAX_CODES_PER_CALL:
    out = numpy.concatenate([eval_mos(basis, dgrid, codes[g:g + MAX_CODES_PER_CALL], C, p0, np_)
                             for g in range(0, nd, MAX_CODES_PER_CALL)])
    if sink is not None:
      sink(out.shape, ((c0, out[..., c0:c0 + OUTPUT_SLICE]) for c0 in range(0, np_, OUTPUT_SLICE)))
    return out
  return _consume(numpy.empty((nd, nmo, np_)), _mo_blocks(basis, dgrid, codes, C, p0, np_), sink)


def eval_rho(basis, dgrid, coeffs, occ, codes, p0=0, np_=None, mo_norm=False, staged=False, device=False):
  '''Density (np,), derivative outputs (ncodes, np) and optionally the MO
  norms sum_p phi^2 (nmo,), all FP64: okmi_rho (the fused FP64 kernel up to
  64 MOs), or with ``staged`` okmi_rho_staged (AO sets through device memory,
  then the FP64 GEMM).  ``device``: density and derivative rows as device
  tensors instead of host arrays.'''
  torch = _torch()
  L = _native.lib()
  np_ = dgrid.npts - p0 if np_ is None else np_
  codes = numpy.ascontiguousarray(codes, dtype=numpy.intc)
  nd = len(codes)
  C = numpy.ascontiguousarray(coeffs, dtype=numpy.float64)
  nmo = C.shape[0]
  if C.ndim != 2 or C.shape[1] != basis.nao:
    raise ValueError('MO coefficients have shape %s but the basis has %d AOs' % (C.shape, basis.nao))
  if nd > MAX_OUTPUTS_PER_CALL:
    k = MAX_OUTPUTS_PER_CALL
    r, d, nrm = eval_rho(basis, dgrid, C, occ, codes[:k], p0, np_, mo_norm, staged, device)
    rest = [eval_rho(basis, dgrid, C, occ, codes[g:g + k], p0, np_, False, staged, device)[1]
            for g in range(k, nd, k)]
    return r, (torch.cat([d] + rest) if device else numpy.concatenate([d] + rest)), nrm
  dev = dgrid.x.device
  dC = torch.as_tensor(C, device=dev)
  docc = torch.as_tensor(numpy.ascontiguousarray(occ, dtype=numpy.float64), device=dev)
  rho = torch.empty(max(np_, 1), dtype=torch.float64, device=dev)
  delta = torch.empty((max(nd, 1), max(np_, 1)), dtype=torch.float64, device=dev)
  nrm = torch.zeros(nmo, dtype=torch.float64, device=dev) if mo_norm else None
  nsets = 1 + 3 + nd
  ws_chunk = max(64, min(np_, WORKSPACE_BYTES // max(1, nsets * basis.nao * 8)))
  ws = torch.empty(L.okmi_workspace_bytes(basis.handle, nsets, ws_chunk) // 8 + 1,
                   dtype=torch.float64, device=dev)
  fn = L.okmi_rho_staged if staged else L.okmi_rho
  _native.check(fn(basis.handle, ctypes.byref(dgrid.desc), p0, np_, ctypes.c_void_p(dC.data_ptr()),
                   ctypes.c_void_p(docc.data_ptr()), nmo, codes.ctypes.data_as(_native.P_int), nd,
                   ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(delta.data_ptr()),
                   delta.shape[1], ctypes.c_void_p(nrm.data_ptr()) if mo_norm else None,
                   ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream()), 'okmi_rho')
  nrm = nrm.cpu().numpy() if mo_norm else None
  if device:
    return rho[:np_], delta[:nd, :np_], nrm
  return rho[:np_].cpu().numpy(), delta[:nd, :np_].cpu().numpy(), nrm


def electron_count(rho, d3r):
  '''d3r * sum(rho) in FP64 with compensated summation on the device
  (okmi_electron_count); ``rho`` is a numpy array or a device tensor.'''
  torch = _torch()
  L = _native.lib()
  if not isinstance(rho, torch.Tensor):
    rho = torch.as_tensor(numpy.ascontiguousarray(rho, dtype=numpy.float64).reshape(-1),
                          device=torch.device('cuda', torch.cuda.current_device()))
  rho = rho.reshape(-1).contiguous()
  ws = torch.empty(L.okmi_electron_count_workspace_bytes() // 8 + 1, dtype=torch.float64, device=rho.device)
  out = torch.empty(1, dtype=torch.float64, device=rho.device)
  _native.check(L.okmi_electron_count(ctypes.c_void_p(rho.data_ptr()), rho.numel(), float(d3r),
                                      ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8,
                                      ctypes.c_void_p(out.data_ptr()), _stream()), 'okmi_electron_count')
  return float(out.item())


def gemm(A, B):
  '''A (m x k) @ B (k x n) in FP64 on the device (the MO kernel).'''
  torch = _torch()
  L = _native.lib()
  A = numpy.ascontiguousarray(A, dtype=numpy.float64)
  B = numpy.ascontiguousarray(B, dtype=numpy.float64)
  m, k = A.shape
  if B.shape[0] != k:
    raise ValueError('shape mismatch: (%d, %d) @ %s' % (m, k, B.shape))
  n = B.shape[1]
  out = numpy.empty((m, n))
  if n == 0 or m == 0:
    return out
  dev = torch.device('cuda', torch.cuda.current_device())
  dA = torch.as_tensor(A, device=dev)
  chunk = max(64, min(n, OUTPUT_CHUNK_BYTES // (8 * max(m, k))))
  dB = torch.empty((k, chunk), dtype=torch.float64, device=dev)
  dO = torch.empty((m, chunk), dtype=torch.float64, device=dev)
  for c0 in range(0, n, chunk):
    w = min(chunk, n - c0)
    dB[:, :w].copy_(torch.from_numpy(B[:, c0:c0 + w]))
    _native.check(L.okmi_mocreator(ctypes.c_void_p(dB.data_ptr()), k, w, chunk,
                                   ctypes.c_void_p(dA.data_ptr()), m, ctypes.c_void_p(dO.data_ptr()),
                                   chunk, _stream()), 'okmi_mocreator')
    out[:, c0:c0 + w] = dO[:, :w].cpu().numpy()
  return out


def _specs(qc):
  if isinstance(qc, dict):
    return qc['geo_spec'], qc['ao_spec'], qc['mo_spec']
  return qc.geo_spec, qc.ao_spec, qc.mo_spec


# --------------------------------------------------------------------------
# orbkit's API
def ao_creator(geo_spec, ao_spec, drv=None, x=None, y=None, z=None, is_vector=None):
  '''All contracted AOs (or a derivative) on the grid: shape (NAO,)+N
  (orbkit/core.py:38-105).'''
  if all(v is None for v in [x, y, z, is_vector]) and not grid.is_initialized:
    display('\nSetting up the grid...')
    grid.grid_init(is_vector=True)
    display(grid.get_grid())
  if x is None: x = grid.x
  if y is None: y = grid.y
  if z is None: z = grid.z
  if is_vector is None: is_vector = grid.is_vector
  was_vector = is_vector
  if not is_vector:
    N = (len(x), len(y), len(z))
  else:
    if len(x) != len(y) or len(x) != len(z):
      raise ValueError('Dimensions of x-, y-, and z- coordinate differ!')
    N = (len(x),)
  code = validate_drv(drv)
  geo_spec = require(geo_spec, dtype='f')
  basis = device_basis(geo_spec, ao_spec)
  dg = DeviceGrid(x, y, z, regular=not was_vector)
  ao_list = eval_aos(basis, dg, [code])[0]
  if not was_vector:
    return ao_list.reshape((len(ao_list),) + N, order='C')
  return ao_list


def mo_creator(ao_list, mo_spec):
  '''MOs from AOs: shape (NMO,)+N (orbkit/core.py:107-132).'''
  ao_list = require(ao_list, dtype='f')
  shape = ao_list.shape
  mo_coeffs = mo_spec.get_coeffs()
  mo = gemm(mo_coeffs, ao_list.reshape((shape[0], -1)))
  return mo.reshape(((len(mo_coeffs),) + shape[1:]), order='C')


def cartesian2spherical(ao_list, ao_spec):
  '''Cartesian AO rows to real spherical rows (orbkit/core.py:135-176).'''
  ao_list = require(ao_list, dtype='f')
  shape = ao_list.shape
  T = spherical_matrix(ao_spec)
  out = gemm(T, ao_list.reshape((shape[0], -1)))
  return out.reshape((T.shape[0],) + shape[1:])


Spec = None


def initializer(global_args):
  global Spec
  Spec = global_args


def slice_rho(xx):
  '''One slice [xx[0], xx[1]) of the vector grid in ``grid`` (orbkit/core.py:179-308).'''
  try:
    geo_spec, ao_spec, mo_spec = _specs(Spec['qc'])
    drv = Spec['derivative']
    calc_mo = Spec['calc_mo']
    x = grid.x[xx[0]:xx[1]]
    y = grid.y[xx[0]:xx[1]]
    z = grid.z[xx[0]:xx[1]]
    basis = device_basis(require(geo_spec, dtype='f'), ao_spec)
    dg = DeviceGrid(x, y, z, regular=False)
    codes = [0] if drv is None else _codes(drv)
    if calc_mo:
      if Spec['calc_ao']:
        vals = eval_aos(basis, dg, codes)
      else:
        vals = eval_mos(basis, dg, codes, mo_spec.get_coeffs())
      return vals if drv is not None else vals[0]
    occ = mo_spec.get_occ()
    rho, delta, mo_norm = eval_rho(basis, dg, mo_spec.get_coeffs(), occ,
                                   [] if drv is None else codes, mo_norm=True)
    if drv is None:
      return rho, mo_norm
    return rho, mo_norm, delta
  except KeyboardInterrupt:
    return 0


def _drv_list(drv, laplacian):
  if laplacian:
    if not (drv is None or drv == ['xx', 'yy', 'zz'] or drv == ['x2', 'y2', 'z2']):
      display('Note: You have set the option `laplacian` and specified values\n' +
              'for `drv`. Both options are not compatible.\n' +
              'The option `drv` has been changed to `drv=["xx","yy","zz"]`.')
    drv = ['xx', 'yy', 'zz']
  if drv is not None:
    try:
      drv = list(drv)
    except TypeError:
      drv = [drv]
  return drv


def rho_compute(qc, calc_ao=False, calc_mo=False, drv=None, laplacian=False,
                numproc=1, slice_length=1e4, vector=None, save_hdf5=False, **kwargs):
  '''Density, MOs or AOs (and derivatives) on the grid module's grid
  (orbkit/core.py:314-605).  Returns, as orbkit does:
    calc_mo/calc_ao: mo_list (NMO,)+N, or (NDRV, NMO)+N with drv
    density:         rho N; (rho, delta_rho) with drv;
                     (rho, delta_rho, laplacian_rho) with laplacian
  '''
  precision = kwargs.pop('precision', 'f64')
  if calc_ao and calc_mo:
    raise ValueError('Choose either calc_ao=True or calc_mo=True')
  elif calc_ao:
    calc_mo = True
  slice_length = slice_length if not vector else vector
  if numproc <= 0:
    return rho_compute_no_slice(qc, calc_ao=calc_ao, calc_mo=calc_mo, drv=drv,
                                laplacian=laplacian, **kwargs)
  drv = _drv_list(drv, laplacian)
  is_drv = drv is not None
  if isinstance(qc, dict):
    qc = QCinfo(qc)
  geo_spec, ao_spec, mo_spec = qc.geo_spec, qc.ao_spec, qc.mo_spec
  codes = None if drv is None else _codes(drv)

  if not grid.is_initialized:
    display('\nSetting up the grid...')
    grid.grid_init()
    display(grid.get_grid())
  was_vector = grid.is_vector
  N = (len(grid.x),) if was_vector else (len(grid.x), len(grid.y), len(grid.z))
  npts = int(numpy.prod(N))
  if slice_length <= 0:
    slice_length = numpy.ceil(npts / float(max(numproc, 1))) + 1
  sNum = int(numpy.floor(npts / slice_length) + 1)
  if slice_length >= npts:
    slice_length, sNum = npts, 1
  display('\nStarting the calculation of the %s...' % ('molecular orbitals' if calc_mo else 'density'))
  display('The grid has been separated into %d slices each having %.2e grid points.' % (sNum, slice_length))

  # result file: written slice by slice as orbkit's HDF5 sink does (a name
  # ending in .npz: a NumPy archive, no h5py needed)
  out_file = _open_result_file(save_hdf5, slice_length, was_vector) if save_hdf5 else None

  basis = device_basis(require(geo_spec, dtype='f'), ao_spec)
  dg = DeviceGrid(grid.x, grid.y, grid.z, regular=not was_vector)

  if calc_mo:
    c = [0] if codes is None else codes
    name = 'ao_list' if calc_ao else 'mo_list'
    sink = None
    if out_file is not None:
      # the MO axis alone without drv: (nmo, npts), else (ndrv, nmo, npts)
      sink = lambda shape, blocks: _write_dataset(out_file, name, shape if is_drv else shape[1:],
                                                  blocks if is_drv else ((c0, b[0]) for c0, b in blocks),
                                                  ((len(drv),) if is_drv else ()) + (shape[1],) + N)
    try:
      vals = eval_aos(basis, dg, c, sink=sink) if calc_ao else eval_mos(basis, dg, c, mo_spec.get_coeffs(),
                                                                          sink=sink)
    finally:
      if out_file is not None:
        out_file.close()
    mo_list = vals if is_drv else vals[0]
    if not was_vector and drv is None:
      display('\nNorm of the MOs:')
      for i in range(mo_list.shape[0]):
        display('\t%.6f\t%s %d' % (numpy.sum(numpy.square(mo_list[i])) * grid.d3r,
                                   'AO' if calc_ao else 'MO', i))
    shape = (mo_list.shape[0],) if drv is None else (len(drv), mo_list.shape[1])
    return mo_list.reshape(shape + N)

  torch = _torch()
  if precision != 'f64':
    from .fused import rho_fused
    cl = [] if codes is None else list(codes)
    k = MAX_OUTPUTS_PER_CALL
    parts = [rho_fused(basis, dg, mo_spec.get_coeffs(), mo_spec.get_occ(), cl[g:g + k], precision=precision,
                       device=True) for g in range(0, max(len(cl), 1), k)]
    rho_d, delta_d = parts[0][0], torch.cat([p[1] for p in parts])
    mo_norm = None
  else:
    rho_d, delta_d, mo_norm = eval_rho(basis, dg, mo_spec.get_coeffs(), mo_spec.get_occ(),
                                       [] if codes is None else codes, mo_norm=not display_quiet(), device=True)
  # to the host one slice at a time; with a result file each slice is also
  # written there (rho, then delta_rho)
  try:
    rho = _drain(rho_d, npts, slice_length, out_file, 'rho', N)
    delta = _drain(delta_d, npts, slice_length, out_file, 'delta_rho', (len(drv),) + N) if is_drv else None
  finally:
    if out_file is not None:
      out_file.close()
  if not was_vector and drv is None and mo_norm is not None:
    display('\nNorm of the MOs:')
    for i, v in enumerate(mo_norm):
      display('\t%.6f\tMO %d' % (v * grid.d3r, i))
  if not was_vector:
    display('We have ' + str(electron_count(rho_d, grid.d3r)) + ' electrons.')
  rho = rho.reshape(N)
  if not is_drv:
    return rho
  delta_rho = delta.reshape((len(drv),) + N)
  if laplacian:
    return rho, delta_rho, delta_rho.sum(axis=0)
  return rho, delta_rho


def display_quiet():
  from . import display as _d
  return _d.quiet


def _open_result_file(target, slice_length, was_vector):
  '''Result file of rho_compute with its grid group as orbkit writes it
  (orbkit/core.py:478-485): the coordinates of the vector form of the grid,
  is_vector false and is_regular set to whether the input grid was a vector
  grid (orbkit's values, kept as they are).'''
  from .output.sink import open_sink, write_grid
  f = open_sink(target, slice_length)
  if was_vector:
    x, y, z = grid.x, grid.y, grid.z
  else:
    X, Y, Z = numpy.meshgrid(grid.x, grid.y, grid.z, indexing='ij')
    x, y, z = X.ravel(), Y.ravel(), Z.ravel()
  write_grid(f, x, y, z, False, was_vector)
  return f


def _write_dataset(f, name, shape, blocks, grid_shape):
  f.stream(name, shape, blocks)
  f.set_attr(name, 'shape', numpy.asarray(grid_shape))


def _drain(t, npts, slice_length, f, name, grid_shape):
  '''Host copy of a device result (..., npts) made one slice of points at a
  time; with a result file f, each slice also goes to dataset `name`.'''
  out = numpy.empty(tuple(t.shape[:-1]) + (npts,))
  step = max(1, int(slice_length))
  blocks = ((p0, t[..., p0:p0 + step].cpu().numpy()) for p0 in range(0, npts, step))
  if f is None:
    _consume(out, blocks, None)
  else:
    _consume(out, blocks, lambda shape, b: _write_dataset(f, name, shape, b, grid_shape))
  return out


def rho_compute_no_slice(qc, calc_ao=False, calc_mo=False, drv=None, laplacian=False,
                         return_components=False, x=None, y=None, z=None, is_vector=None, **kwargs):
  '''Same quantities on the whole grid at once, optionally with the AO and MO
  arrays (orbkit/core.py:607-839).'''
  if calc_ao and calc_mo:
    raise ValueError('calc_ao and calc_mo are mutually exclusive arguments.' +
                     'Use calc_mo and return_components instead!')
  if all(v is None for v in [x, y, z, is_vector]) and not grid.is_initialized:
    display('\nSetting up the grid...')
    grid.grid_init()
    display(grid.get_grid())
  if x is None: x = grid.x
  if y is None: y = grid.y
  if z is None: z = grid.z
  if is_vector is None: is_vector = grid.is_vector
  was_vector = is_vector
  if not is_vector:
    N = (len(x), len(y), len(z))
    d3r = 1.0
    for a in (x, y, z):
      if len(a) > 1:
        d3r *= (a[1] - a[0])
  else:
    if len(x) != len(y) or len(x) != len(z):
      raise ValueError('Dimensions of x-, y-, and z- coordinate differ!')
    N = (len(x),)
  if laplacian:
    drv = _drv_list(None, True)
  geo_spec, ao_spec, mo_spec = _specs(qc)
  basis = device_basis(require(geo_spec, dtype='f'), ao_spec)
  dg = DeviceGrid(x, y, z, regular=not was_vector)
  C = mo_spec.get_coeffs()

  if drv is not None:
    drv = _drv_list(drv, False)
    codes = _codes(drv)
    delta_ao = eval_aos(basis, dg, codes)
    delta_ao_list = convert(delta_ao, was_vector, N)
    if calc_ao:
      return delta_ao_list
    delta_mo = gemm_sets(C, delta_ao)
    delta_mo_list = convert(delta_mo, was_vector, N)
    if calc_mo:
      return (delta_ao_list, delta_mo_list) if return_components else delta_mo_list

  ao = eval_aos(basis, dg, [0])[0]
  ao_list = convert(ao, was_vector, N)
  if calc_ao:
    return ao_list
  mo = gemm(C, ao)
  mo_list = convert(mo, was_vector, N)
  if calc_mo:
    return (ao_list, mo_list) if return_components else mo_list
  occ = mo_spec.get_occ()
  codes = [] if drv is None else _codes(drv)
  rho, delta, _ = eval_rho(basis, dg, C, occ, codes)
  rho = rho.reshape(N)
  if not was_vector:
    display('We have ' + str(electron_count(rho, d3r)) + ' electrons.')
  if drv is None:
    return (ao_list, mo_list, rho) if return_components else rho
  delta_rho = delta.reshape((len(drv),) + N)
  delta = (delta_rho, delta_rho.sum(axis=0)) if laplacian else (delta_rho,)
  return ((ao_list, mo_list, rho, delta_ao_list, delta_mo_list) + delta
          if return_components else (rho,) + delta)


def gemm_sets(C, sets):
  '''C @ sets[j] for each derivative set j.'''
  return numpy.array([gemm(C, s) for s in sets])

def gross_atomic(basis, dgrid, coeffs, occ, code, rows_per_sel, p0=0, np_=None):
  '''Gross atomic densities (len(rows_per_sel), np) for AO derivative code
  `code`; rows_per_sel[j] are the AO rows of selection j (okmi_gross_atomic_density).'''
  torch = _torch()
  L = _native.lib()
  np_ = dgrid.npts - p0 if np_ is None else np_
  C = numpy.ascontiguousarray(coeffs, dtype=numpy.float64)
  dev = dgrid.x.device
  dC = torch.as_tensor(C, device=dev)
  docc = torch.as_tensor(numpy.ascontiguousarray(occ, dtype=numpy.float64), device=dev)
  start = numpy.zeros(len(rows_per_sel) + 1, dtype=numpy.intc)
  start[1:] = numpy.cumsum([len(r) for r in rows_per_sel])
  flat = numpy.ascontiguousarray(numpy.concatenate([numpy.asarray(r, dtype=numpy.intc) for r in rows_per_sel])
                                 if len(rows_per_sel) else numpy.zeros(0, numpy.intc), dtype=numpy.intc)
  nsel = len(rows_per_sel)
  out = torch.empty((max(nsel, 1), max(np_, 1)), dtype=torch.float64, device=dev)
  chunk = max(64, min(np_, WORKSPACE_BYTES // max(1, 16 * basis.nao)))
  ws = torch.empty(L.okmi_gross_atomic_workspace_bytes(basis.handle, chunk, int(start[-1]), nsel) // 8 + 1,
                   dtype=torch.float64, device=dev)
  _native.check(L.okmi_gross_atomic_density(
      basis.handle, ctypes.byref(dgrid.desc), p0, np_, int(code), ctypes.c_void_p(dC.data_ptr()),
      ctypes.c_void_p(docc.data_ptr()), C.shape[0], start.ctypes.data_as(_native.P_int),
      flat.ctypes.data_as(_native.P_int) if len(flat) else None, nsel, ctypes.c_void_p(out.data_ptr()),
      out.shape[1], ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream()), 'okmi_gross_atomic_density')
  return out[:nsel, :np_].cpu().numpy()


def masked_mo_product(coeffs, occ, ao, mo, rows):
  '''sum_i occ_i mo_i sum_{r in rows} C[i][r] ao_r for host AO (nao, n) and
  MO (nmo, n) arrays (okmi_masked_mo_product).'''
  torch = _torch()
  L = _native.lib()
  dev = torch.device('cuda', torch.cuda.current_device())
  t = lambda a: torch.as_tensor(numpy.ascontiguousarray(a, dtype=numpy.float64), device=dev)
  rows = numpy.ascontiguousarray(rows, dtype=numpy.intc).reshape(-1)
  nmo, nao = numpy.shape(coeffs)
  if numpy.shape(ao)[0] != nao or numpy.shape(mo)[0] != nmo or numpy.shape(ao)[1] != numpy.shape(mo)[1]:
    raise ValueError('masked_mo_product: C %s, AO %s and MO %s do not match'
                     % (numpy.shape(coeffs), numpy.shape(ao), numpy.shape(mo)))
  if rows.size and (rows.min() < 0 or rows.max() >= nao):
    raise ValueError('masked_mo_product: AO rows must lie in [0, %d)' % nao)
  C, docc, A, M = t(coeffs), t(occ), t(ao), t(mo)
  n = A.shape[1]
  r = torch.as_tensor(rows, device=dev)
  out = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
  _native.check(L.okmi_masked_mo_product(ctypes.c_void_p(C.data_ptr()), ctypes.c_void_p(docc.data_ptr()),
                                         C.shape[0], C.shape[1], ctypes.c_void_p(A.data_ptr()),
                                         ctypes.c_void_p(M.data_ptr()), n, n,
                                         ctypes.c_void_p(r.data_ptr()) if r.numel() else None, r.numel(),
                                         ctypes.c_void_p(out.data_ptr()), _stream()), 'okmi_masked_mo_product')
  return out[:n].cpu().numpy()


def _mo_values(qc, codes):
  '''Device MO values (len(codes), nmo, npts) on the grid module's grid.'''
  torch = _torch()
  L = _native.lib()
  if not grid.is_initialized:
    grid.grid_init()
  basis = device_basis(require(qc.geo_spec, dtype='f'), qc.ao_spec)
  dg = DeviceGrid(grid.x, grid.y, grid.z, regular=not grid.is_vector)
  codes = numpy.ascontiguousarray(codes, dtype=numpy.intc)
  C = numpy.ascontiguousarray(qc.mo_spec.get_coeffs(), dtype=numpy.float64)
  dC = torch.as_tensor(C, device=dg.x.device)
  mo = torch.empty((len(codes), C.shape[0], max(dg.npts, 1)), dtype=torch.float64, device=dg.x.device)
  chunk = max(64, min(dg.npts, WORKSPACE_BYTES // max(1, len(codes) * basis.nao * 8)))
  ws = torch.empty(L.okmi_workspace_bytes(basis.handle, len(codes), chunk) // 8 + 1, dtype=torch.float64,
                   device=dg.x.device)
  _native.check(L.okmi_mo_eval(basis.handle, ctypes.byref(dg.desc), 0, dg.npts, codes.ctypes.data_as(_native.P_int),
                               len(codes), ctypes.c_void_p(dC.data_ptr()), C.shape[0], ctypes.c_void_p(mo.data_ptr()),
                               mo.shape[2], ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, _stream()), 'okmi_mo_eval')
  N = (len(grid.x),) if grid.is_vector else (len(grid.x), len(grid.y), len(grid.z))
  return mo, dg.npts, N


def calc_mo_matrix(qc_a, qc_b=None, drv=None, numproc=1, slice_length=1e4, save_hdf5=False, **kwargs):
  '''Products of the bra MOs of qc_a with the ket MOs (or their derivatives)
  of qc_b (default qc_a) on the grid (orbkit/core.py:841-941):
  mo_matrix[d, n, m] = mo_a[n] * d_d mo_b[m], shape (NDRV, NMO_a, NMO_b) + N
  (NDRV = 1 without drv).'''
  torch = _torch()
  L = _native.lib()
  if isinstance(qc_a, dict):
    qc_a = QCinfo(qc_a)
  if drv is None:
    kets = [None]
  elif not isinstance(drv, list):
    kets = [drv]
  else:
    kets = list(drv)
  kcodes = [0 if d is None else _codes([d])[0] for d in kets]
  if qc_b is None or qc_b is qc_a:
    mo, npts, N = _mo_values(qc_a, [0] + kcodes)
    bra, ket = mo[0], mo[1:]
  else:
    bra, npts, N = _mo_values(qc_a, [0])
    bra = bra[0]
    ket, _, _ = _mo_values(qc_b if not isinstance(qc_b, dict) else QCinfo(qc_b), kcodes)
  na, nb, nd = bra.shape[0], ket.shape[1], ket.shape[0]
  out = torch.empty((nd, na, nb, max(npts, 1)), dtype=torch.float64, device=bra.device)
  _native.check(L.okmi_mo_pair_products(ctypes.c_void_p(bra.data_ptr()), na, ctypes.c_void_p(ket.data_ptr()), nb, nd,
                                        npts, bra.shape[-1], ctypes.c_void_p(out.data_ptr()), out.shape[-1],
                                        _stream()), 'okmi_mo_pair_products')
  if not save_hdf5:
    return out[..., :npts].cpu().numpy().reshape((nd, na, nb) + N)
  # result files as orbkit writes them (orbkit/core.py:874-924): the bra
  # (and, with qc_b, the ket) MO values in their own files, then the grid
  # and mo_matrix, slice by slice; mo_matrix is stored flat (NDRV, NMO_a,
  # NMO_b, npts) with its grid shape in the attribute 'shape', as
  # rho_compute's datasets are (orbkit stores this one in the grid shape)
  from .output.sink import open_sink, write_grid
  target = str(save_hdf5)
  side = (lambda k: target[:-4] + '.%s.npz' % k) if target.endswith('.npz') else (lambda k: target + '.' + k)
  if qc_b is None or qc_b is qc_a:   # one file with the values and every ket derivative
    mo_files = [('bra', bra[None] if drv is None else torch.cat([bra[None], ket]))]
  else:
    mo_files = [('bra', bra[None]), ('ket', ket)]
  for kind, mo in mo_files:
    f = _open_result_file(side(kind), slice_length, grid.is_vector)
    try:
      _drain(mo[..., :npts], npts, slice_length, f, 'mo_list', (mo.shape[0], mo.shape[1]) + N)
    finally:
      f.close()
  f = open_sink(target, slice_length)
  try:
    write_grid(f, grid.x, grid.y, grid.z, grid.is_vector, None)
    res = _drain(out[..., :npts], npts, slice_length, f, 'mo_matrix', (nd, na, nb) + N)
  finally:
    f.close()
  return res.reshape((nd, na, nb) + N)


def transition_flux(qc, indices, drv):
  '''-1/2 (phi_i d phi_j - phi_j d phi_i) for MO index pairs (NDRV, NPAIR) + N.'''
  torch = _torch()
  L = _native.lib()
  drv = [drv] if not isinstance(drv, (list, tuple)) else list(drv)
  codes = [0] + [_codes([d])[0] for d in drv]
  mo, npts, N = _mo_values(qc, codes)
  idx = numpy.ascontiguousarray(indices, dtype=numpy.intc)
  if idx.size % 2 or (idx.ndim > 1 and idx.shape[-1] != 2):
    raise ValueError('transition_flux: indices must be pairs (i, j), shape (npair, 2)')
  if idx.size and (idx.min() < 0 or idx.max() >= mo.shape[1]):
    raise ValueError('transition_flux: MO indices must lie in [0, %d)' % mo.shape[1])
  pairs = torch.as_tensor(idx.reshape(-1), device=mo.device)
  npair = pairs.numel() // 2
  out = torch.empty((len(drv), npair, max(npts, 1)), dtype=torch.float64, device=mo.device)
  _native.check(L.okmi_transition_flux(ctypes.c_void_p(mo.data_ptr()), mo.shape[1], len(drv), npts, mo.shape[-1],
                                       ctypes.c_void_p(pairs.data_ptr()) if npair else None, npair,
                                       ctypes.c_void_p(out.data_ptr()), out.shape[-1], _stream()),
                'okmi_transition_flux')
  return out[..., :npts].cpu().numpy().reshape((len(drv), npair) + N)
