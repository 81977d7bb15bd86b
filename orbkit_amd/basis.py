'''Device basis tables built from orbkit-style ``geo_spec`` / ``ao_spec``.

``flatten`` gathers the arrays orbkit hands its AO kernel
(orbkit/core.py:89-97: lxlylz, components per contraction, primitive
(exponent, coefficient) pairs, primitives per contraction, contraction atoms,
the pre-normalised flag), plus the two post-processing steps of ao_creator
that this implementation folds into the kernel: the per-row factor
``ao_spec[0]['N']`` (orbkit/core.py:98-100) and the Cartesian-to-spherical map
(orbkit/core.py:135-176).
'''
import ctypes
import hashlib

import numpy

from . import _native
from .tools import get_cart2sph


def spherical_blocks(ao_spec):
  '''Per-contraction dense maps (nout x ncart) from Cartesian rows to the
  spherical rows of ``ao_spec.get_old_ao_spherical()``.

  For spherical AO (j0, (l, m)) each term (exponents, coefficient) of
  ``get_cart2sph(l, m)`` adds coefficient*global to the column of the row of
  contraction j0 with those exponents (the last such row, as the loop in
  orbkit/core.py:166-174 leaves it).'''
  lxlylz = ao_spec.get_lxlylz()
  assign = ao_spec.get_assign_lxlylz_to_cont()
  ncont = len(ao_spec)
  rows = [[] for _ in range(ncont)]
  for i, j in enumerate(assign):
    rows[j].append(i)
  blocks = [[] for _ in range(ncont)]
  for j0, lm in ao_spec.get_old_ao_spherical():
    l, m = int(lm[0]), int(lm[1])
    exps, coefs, glob = get_cart2sph(l, m)
    r = numpy.zeros(len(rows[j0]))
    for e, c in zip(exps, coefs):
      hit = [k for k, i in enumerate(rows[j0]) if tuple(lxlylz[i]) == tuple(e)]
      if not hit:
        raise ValueError('contraction %d has no Cartesian component %s needed for (l,m)=(%d,%d)'
                         % (j0, e, l, m))
      r[hit[-1]] += c * glob
    blocks[j0].append(r)
  return [numpy.array(b).reshape(len(b), len(rows[j])) for j, b in enumerate(blocks)]


def spherical_matrix(ao_spec):
  '''Dense (NAO_sph x NAO_cart) Cartesian-to-spherical matrix.'''
  blocks = spherical_blocks(ao_spec)
  ncart = ao_spec.get_nlxlylz_per_cont()
  T = numpy.zeros((sum(len(b) for b in blocks), int(ncart.sum())))
  r = c = 0
  for b, n in zip(blocks, ncart):
    T[r:r + len(b), c:c + n] = b
    r += len(b)
    c += n
  return T


def flatten(geo_spec, ao_spec):
  geo = numpy.ascontiguousarray(geo_spec, dtype=numpy.float64).reshape(-1, 3)
  f = dict(lxlylz=ao_spec.get_lxlylz(), ncart=ao_spec.get_nlxlylz_per_cont(),
           prims=ao_spec.get_prim_coeffs(), nprim=ao_spec.get_nprim_per_cont(),
           atoms=ao_spec.get_assign_cont_to_atoms(), normalized=int(ao_spec.get_normalized()),
           geo=geo, scale=None, nout=None, transform=None)
  if 'N' in ao_spec[0]:
    f['scale'] = numpy.ascontiguousarray(ao_spec[0]['N'], dtype=numpy.float64).reshape(-1)
  if ao_spec.spherical:
    blocks = spherical_blocks(ao_spec)
    f['nout'] = numpy.array([len(b) for b in blocks], dtype=numpy.intc)
    f['transform'] = numpy.concatenate([b.reshape(-1) for b in blocks]) if blocks else numpy.zeros(0)
  return f


class DeviceBasis:
  '''Owns an okmi_basis handle (device tables on the current HIP device).'''

  def __init__(self, flat):
    L = _native.lib()
    self.flat = flat
    keep = []

    def I(a):
      arr, p = _native.int_array(a)
      keep.append(arr)
      return p

    def D(a):
      if a is None:
        return None
      arr, p = _native.dbl_array(a)
      keep.append(arr)
      return p
    h = ctypes.c_void_p()
    rc = L.okmi_basis_create(len(flat['ncart']), I(flat['ncart']), I(flat['lxlylz']), I(flat['nprim']),
                             D(flat['prims']), len(flat['geo']), D(flat['geo']), I(flat['atoms']),
                             flat['normalized'], D(flat['scale']),
                             I(flat['nout']) if flat['nout'] is not None else None,
                             D(flat['transform']), ctypes.byref(h))
    _native.check(rc, 'okmi_basis_create')
    self.handle = h
    self.nao = L.okmi_basis_nao(h)

  def __del__(self):
    h = getattr(self, 'handle', None)
    if h is not None and h.value:
      try:
        _native.lib().okmi_basis_destroy(h)
      except Exception:
        pass
      self.handle = None


_cache = {}


def _key(flat):
  h = hashlib.sha1()
  for k in ('lxlylz', 'ncart', 'prims', 'nprim', 'atoms', 'geo', 'scale', 'nout', 'transform'):
    v = flat[k]
    h.update(k.encode())
    if v is not None:
      h.update(numpy.ascontiguousarray(v).tobytes())
  h.update(bytes([flat['normalized']]))
  return h.hexdigest()


def device_basis(geo_spec, ao_spec):
  '''Cached DeviceBasis for this geometry and basis (keyed by content).'''
  import torch
  if not torch.cuda.is_available():
    raise RuntimeError('orbkit_amd needs a HIP device (torch.cuda.is_available() is False); '
                       'it has no CPU compute path')
  flat = flatten(geo_spec, ao_spec)
  key = (_key(flat), torch.cuda.current_device())
  b = _cache.get(key)
  if b is None:
    if len(_cache) > 16:
      _cache.clear()
    b = _cache[key] = DeviceBasis(flat)
  return b
