'''numpy restatement of orbkit's grid orchestration.  TEST INFRASTRUCTURE ONLY.

Functions and the reference lines they follow:

  ao_creator            orbkit/core.py:38-105 (vector grid): C AO kernel,
                        ORCA row factor 'N', then cartesian2spherical
  cartesian2spherical   orbkit/core.py:135-176 (row loop with numpy axpys)
  mo_creator            orbkit/core.py:107-132 -> orc_mocreator (C triple loop)
  slice_rho             orbkit/core.py:179-308 (including the separate AO
                        evaluations for second-derivative product terms)
  rho_compute           orbkit/core.py:314-605 (argument rules, slice plan,
                        assembly and return shapes) on explicit x, y, z

Pinning (tests/test_oracle.py):
  AO values / derivatives   bit-level vs orbkit's own c_lcreator (oracle/_ref);
                            central finite differences; closed-form norm
  cart2sph table            unit norm and orthogonality of d/f rows; fit of
                            real solid harmonics (scipy) up to row sign;
                            Gaussian cubegen cubes (d shells)
  whole chain               refdata_rho_compute.npz (Cartesian H2O, rho,
                            laplacian, MOs with 10 derivative sets) and the
                            Gaussian cubes (spherical H2O: rho, grad, laplacian,
                            valence rho, HOMO)
  mocreator                 numpy FP64 matmul
'''
import ctypes
import os
import subprocess

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, 'liborbkit_oracle.so')
TUNED_LIB = os.path.join(HERE, 'libcpu_tuned.so')
REF_LIB = os.path.join(HERE, '_ref', 'libref_orbkit.so')

_P_i = ctypes.POINTER(ctypes.c_int)
_P_d = ctypes.POINTER(ctypes.c_double)
_lib = None
_ref = None
_tuned = None


def build():
  '''Compiles the oracle (and oracle/_ref where the reference sources exist).'''
  subprocess.run(['make', '-s', '-C', HERE], check=True)


def clib():
  global _lib
  if _lib is None:
    if not os.path.exists(ORACLE_LIB):
      build()
    l = ctypes.CDLL(ORACLE_LIB)
    l.orc_aocreator.argtypes = [_P_d, _P_i, _P_i, _P_d, _P_i, _P_d, _P_i, ctypes.c_int,
                                _P_d, _P_d, _P_d, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
    l.orc_mocreator.argtypes = [_P_d, _P_d, _P_d, ctypes.c_int, ctypes.c_int64, ctypes.c_int]
    l.orc_ao_xyz.restype = ctypes.c_double
    l.orc_ao_xyz.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int]
    l.orc_ao_norm.restype = ctypes.c_double
    l.orc_ao_norm.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int]
    l.orc_lcreator.argtypes = [_P_d, ctypes.c_int64, _P_i, _P_d, _P_d, _P_d, _P_d, _P_d,
                               ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    l.orc_rho_slices.restype = ctypes.c_int
    l.orc_rho_slices.argtypes = [_P_d, _P_d, _P_d, _P_i, ctypes.c_int, _P_i, _P_i, _P_d, _P_i, _P_d, _P_i,
                                 ctypes.c_int, ctypes.c_int, _P_d, _P_d, ctypes.c_int, _P_d, _P_d,
                                 ctypes.c_int, _P_d, _P_d, _P_d, ctypes.c_int64, ctypes.c_int64,
                                 ctypes.c_int, ctypes.c_int]
    l.orc_max_threads.restype = ctypes.c_int
    _lib = l
  return _lib


def ref_lib():
  '''orbkit's own c_lcreator (None when oracle/_ref was not built).'''
  global _ref
  if _ref is None and os.path.exists(REF_LIB):
    r = ctypes.CDLL(REF_LIB)
    r.c_lcreator.argtypes = [_P_d, _P_i, _P_d, _P_d, _P_d, _P_d, _P_d,
                             ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    r.get_ao_xyz.restype = ctypes.c_double
    r.get_ao_xyz.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int]
    r.ao_norm.restype = ctypes.c_double
    r.ao_norm.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int]
    _ref = r
  return _ref


def _i(a):
  a = numpy.ascontiguousarray(a, dtype=numpy.intc)
  return a, a.ctypes.data_as(_P_i)


def _d(a):
  a = numpy.ascontiguousarray(a, dtype=numpy.float64)
  return a, a.ctypes.data_as(_P_d)


def flat_basis(geo_spec, ao_spec):
  return dict(lxlylz=ao_spec.get_lxlylz(), ncart=ao_spec.get_nlxlylz_per_cont(),
              prims=ao_spec.get_prim_coeffs(), nprim=ao_spec.get_nprim_per_cont(),
              atoms=ao_spec.get_assign_cont_to_atoms(), normalized=int(ao_spec.get_normalized()),
              geo=numpy.ascontiguousarray(geo_spec, dtype=numpy.float64).reshape(-1, 3))


def aocreator(f, x, y, z, drv):
  '''Cartesian AO rows (NAO_cart, npts): orbkit/cy_core.pyx:49-78.'''
  L = clib()
  x, px = _d(x); y, py = _d(y); z, pz = _d(z)
  n = len(x)
  out = numpy.zeros((int(numpy.sum(f['ncart'])), n))
  keep = [_i(f['lxlylz']), _i(f['ncart']), _d(f['prims']), _i(f['nprim']), _d(f['geo']), _i(f['atoms'])]
  L.orc_aocreator(out.ctypes.data_as(_P_d), keep[0][1], keep[1][1], keep[2][1], keep[3][1], keep[4][1],
                  keep[5][1], len(f['ncart']), px, py, pz, n, int(drv), f['normalized'])
  return out


def ref_aocreator(f, x, y, z, drv):
  '''Same rows from orbkit's own c_lcreator, looped as cy_core.aocreator does.'''
  R = ref_lib()
  x, px = _d(x); y, py = _d(y); z, pz = _d(z)
  n = len(x)
  out = numpy.zeros((int(numpy.sum(f['ncart'])), n))
  lx = numpy.ascontiguousarray(f['lxlylz'], dtype=numpy.intc)
  pr = numpy.ascontiguousarray(f['prims'], dtype=numpy.float64)
  geo = numpy.ascontiguousarray(f['geo'], dtype=numpy.float64)
  row = prim = 0
  for c, (nc, npr) in enumerate(zip(f['ncart'], f['nprim'])):
    at = numpy.ascontiguousarray(geo[f['atoms'][c]])
    sub = numpy.zeros((nc, n))
    R.c_lcreator(sub.ctypes.data_as(_P_d), _i(lx[row:row + nc])[1], _d(pr[prim:prim + npr])[1],
                 at.ctypes.data_as(_P_d), px, py, pz, n, int(nc), int(npr), int(drv), f['normalized'])
    out[row:row + nc] = sub
    row += nc
    prim += npr
  return out


def mocreator(ao, C):
  L = clib()
  ao, pa = _d(ao)
  C, pc = _d(C)
  nao, n = ao.shape
  nmo = C.shape[0]
  out = numpy.zeros((nmo, n))
  L.orc_mocreator(out.ctypes.data_as(_P_d), pa, pc, nao, n, nmo)
  return out


def cartesian2spherical(ao_list, ao_spec, get_cart2sph):
  '''orbkit/core.py:135-176: for each spherical AO, sum coefficient*global*
  Cartesian row over the table terms, the row found by a linear search over
  the contraction's rows (last match).'''
  lxlylz = ao_spec.get_lxlylz()
  assign = ao_spec.get_assign_lxlylz_to_cont()
  rows = [[] for _ in range(len(ao_spec))]
  for i, j in enumerate(assign):
    rows[j].append(i)
  sph = ao_spec.get_old_ao_spherical()
  out = numpy.zeros((len(sph),) + ao_list.shape[1:])
  for i0, (j0, k0) in enumerate(sph):
    exps, coefs, glob = get_cart2sph(*k0)
    for c0 in range(len(exps)):
      index0 = None
      for i, j in enumerate(rows[j0]):
        if tuple(lxlylz[j]) == tuple(exps[c0]):
          index0 = i + rows[j0][0]
      out[i0] += coefs[c0] * glob * ao_list[index0]
  return out


def ao_creator(geo_spec, ao_spec, drv, x, y, z, get_cart2sph):
  from orbkit_amd.tools import validate_drv  # label -> code table (orbkit/tools.py:225-239)
  f = flat_basis(geo_spec, ao_spec)
  ao = aocreator(f, x, y, z, validate_drv(drv))
  if 'N' in ao_spec[0]:
    ao = ao * ao_spec[0]['N']
  if ao_spec.spherical:
    ao = cartesian2spherical(ao, ao_spec, get_cart2sph)
  return ao


def slice_rho(geo_spec, ao_spec, mo_spec, x, y, z, drv, calc_mo, calc_ao, get_cart2sph):
  '''Outputs of orbkit's slice_rho for one slice of a vector grid.'''
  aoc = lambda d: ao_creator(geo_spec, ao_spec, d, x, y, z, get_cart2sph)
  C = mo_spec.get_coeffs()
  moc = (lambda a: a) if calc_ao else (lambda a: mocreator(a, C))
  if drv is not None and calc_mo:
    return numpy.array([moc(aoc(d)) for d in drv])
  mo_list = moc(aoc(None))
  if calc_mo:
    return numpy.array(mo_list)
  occ = mo_spec.get_occ()
  rho = numpy.zeros(len(x))
  mo_norm = numpy.zeros(len(mo_list))
  for i in range(len(mo_list)):
    mo_norm[i] = numpy.sum(numpy.square(mo_list[i]))
    rho += occ[i] * numpy.square(numpy.abs(mo_list[i]))
  if drv is None:
    return rho, mo_norm
  delta = numpy.zeros((len(drv), len(x)))
  for i, d in enumerate(drv):
    dmo = mocreator(aoc(d), C)
    two = isinstance(d, str) and len(d) == 2
    if two:
      m0 = mocreator(aoc(d[0]), C)
      if '2' in d or d[0] == d[1]:
        d2 = m0 ** 2
      else:
        d2 = m0 * mocreator(aoc(d[1]), C)
    for m in range(len(mo_list)):
      delta[i] += occ[m] * 2 * dmo[m] * mo_list[m]
      if two:
        delta[i] += occ[m] * 2 * d2[m]
  return rho, mo_norm, delta


def expand(x, y, z):
  '''Regular axes -> vector grid, x slowest (orbkit/cy_grid.pyx:12-30), loop form.'''
  out = numpy.zeros((3, len(x) * len(y) * len(z)))
  c = 0
  for i in range(len(x)):
    for j in range(len(y)):
      for k in range(len(z)):
        out[:, c] = (x[i], y[j], z[k])
        c += 1
  return out


def rho_compute(qc, x, y, z, is_vector, calc_ao=False, calc_mo=False, drv=None, laplacian=False,
                slice_length=1e4, get_cart2sph=None):
  '''orbkit/core.py:314-605 on an explicit grid (no grid module state).'''
  if get_cart2sph is None:
    from orbkit_amd.tools import get_cart2sph
  if calc_ao and calc_mo:
    raise ValueError('Choose either calc_ao=True or calc_mo=True')
  elif calc_ao:
    calc_mo = True
  if laplacian:
    drv = ['xx', 'yy', 'zz']
  if drv is not None:
    try:
      drv = list(drv)
    except TypeError:
      drv = [drv]
  N = (len(x),) if is_vector else (len(x), len(y), len(z))
  if not is_vector:
    x, y, z = expand(x, y, z)
  npts = len(x)
  geo, ao_spec, mo_spec = qc.geo_spec, qc.ao_spec, qc.mo_spec
  if slice_length <= 0:
    slice_length = numpy.ceil(npts / 1.0) + 1
  sNum = int(numpy.floor(npts / slice_length) + 1)
  if slice_length >= npts:
    slice_length, sNum = npts, 1
  xx, i = [], 0
  for s in range(sNum):
    if i == npts:
      break
    xx.append((int(i), int(min(i + slice_length, npts))))
    i += slice_length
  nmo = ao_spec.get_ao_num() if calc_ao else len(mo_spec)
  if calc_mo:
    shape = (nmo, npts) if drv is None else (len(drv), nmo, npts)
    out = numpy.zeros(shape)
  else:
    rho = numpy.zeros(npts)
    delta = numpy.zeros((len(drv), npts)) if drv is not None else None
  for a, b in xx:
    r = slice_rho(geo, ao_spec, mo_spec, x[a:b], y[a:b], z[a:b], drv, calc_mo, calc_ao, get_cart2sph)
    if calc_mo:
      out[..., a:b] = r
    else:
      rho[a:b] = r[0]
      if drv is not None:
        delta[:, a:b] = r[2]
  if calc_mo:
    return out.reshape(shape[:-1] + N)
  rho = rho.reshape(N)
  if drv is None:
    return rho
  delta = delta.reshape((len(drv),) + N)
  if laplacian:
    return rho, delta, delta.sum(axis=0)
  return rho, delta


def rho_slices_c(qc, x, y, z, drv_codes=(), slice_length=10000, nthreads=1):
  """Density (and first / diagonal second derivatives, codes 1-6) on a
  vector grid with the C oracle's sliced, OpenMP-parallel driver.  Returns
  (rho, delta (len(codes), n), mo_norm)."""
  from orbkit_amd.tools import get_cart2sph
  L = clib()
  geo, ao_spec, mo_spec = qc.geo_spec, qc.ao_spec, qc.mo_spec
  f = flat_basis(geo, ao_spec)
  nao_cart = int(numpy.sum(f['ncart']))
  scale = numpy.ascontiguousarray(ao_spec[0]['N'], dtype=float).reshape(-1) if 'N' in ao_spec[0] else None
  if ao_spec.spherical:
    eye = numpy.eye(nao_cart)
    T = cartesian2spherical(eye, ao_spec, get_cart2sph)   # (nao_sph, nao_cart)
  else:
    T = None
  nao_out = T.shape[0] if T is not None else nao_cart
  C = numpy.ascontiguousarray(mo_spec.get_coeffs())
  occ = numpy.ascontiguousarray(mo_spec.get_occ())
  x, px = _d(x); y, py = _d(y); z, pz = _d(z)
  n = len(x)
  codes = numpy.ascontiguousarray(drv_codes, dtype=numpy.intc)
  rho = numpy.zeros(n)
  delta = numpy.zeros((max(len(codes), 1), n))
  mo_norm = numpy.zeros(len(C))
  keep = [_i(f['lxlylz']), _i(f['ncart']), _d(f['prims']), _i(f['nprim']), _d(f['geo']), _i(f['atoms'])]
  sc = _d(scale) if scale is not None else (None, None)
  tt = _d(T) if T is not None else (None, None)
  rc = L.orc_rho_slices(rho.ctypes.data_as(_P_d), mo_norm.ctypes.data_as(_P_d), delta.ctypes.data_as(_P_d),
                        codes.ctypes.data_as(_P_i), len(codes), keep[0][1], keep[1][1], keep[2][1], keep[3][1],
                        keep[4][1], keep[5][1], len(f['ncart']), nao_cart, sc[1], tt[1], nao_out,
                        _d(C)[1], _d(occ)[1], len(C), px, py, pz, n, int(slice_length), f['normalized'],
                        int(nthreads))
  if rc:
    raise RuntimeError('orc_rho_slices failed (%d)' % rc)
  return rho, delta[:len(codes)], mo_norm


def tuned_lib():
  global _tuned
  if _tuned is None:
    if not os.path.exists(TUNED_LIB):
      build()
    t = ctypes.CDLL(TUNED_LIB)
    t.orc_rho_slices_tuned.restype = ctypes.c_int
    t.orc_rho_slices_tuned.argtypes = [_P_d, _P_d, _P_i, ctypes.c_int, _P_i, _P_i, _P_d, _P_i, _P_d, _P_i,
                                       ctypes.c_int, ctypes.c_int, _P_d, _P_d, ctypes.c_int, _P_d, _P_d, _P_d,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
    _tuned = t
  return _tuned


def rho_slices_tuned(qc, x, y, z, drv_codes=(), slice_length=10000, nthreads=1):
  """rho_slices_c's results from the tuned CPU form (oracle/cpu_tuned.c):
  exponentials shared across components and AO sets, the row map folded into
  C, a blocked MO contraction.  Returns (rho, delta (len(codes), n))."""
  from orbkit_amd.tools import get_cart2sph
  L = tuned_lib()
  geo, ao_spec, mo_spec = qc.geo_spec, qc.ao_spec, qc.mo_spec
  f = flat_basis(geo, ao_spec)
  nao_cart = int(numpy.sum(f['ncart']))
  N = numpy.asarray(ao_spec[0]['N'], dtype=float).reshape(-1) if 'N' in ao_spec[0] else numpy.ones(nao_cart)
  T = cartesian2spherical(numpy.diag(N), ao_spec, get_cart2sph) if ao_spec.spherical else numpy.diag(N)
  TC = numpy.ascontiguousarray(mo_spec.get_coeffs() @ T)
  occ = numpy.ascontiguousarray(mo_spec.get_occ())
  x, px = _d(x); y, py = _d(y); z, pz = _d(z)
  n = len(x)
  codes = numpy.ascontiguousarray(drv_codes, dtype=numpy.intc)
  rho = numpy.zeros(n)
  delta = numpy.zeros((max(len(codes), 1), n))
  keep = [_i(f['lxlylz']), _i(f['ncart']), _d(f['prims']), _i(f['nprim']), _d(f['geo']), _i(f['atoms'])]
  rc = L.orc_rho_slices_tuned(rho.ctypes.data_as(_P_d), delta.ctypes.data_as(_P_d), codes.ctypes.data_as(_P_i),
                              len(codes), keep[0][1], keep[1][1], keep[2][1], keep[3][1], keep[4][1], keep[5][1],
                              len(f['ncart']), nao_cart, _d(TC)[1], _d(occ)[1], len(TC), px, py, pz, n,
                              int(slice_length), f['normalized'], int(nthreads))
  if rc:
    raise RuntimeError('orc_rho_slices_tuned failed (%d)' % rc)
  return rho, delta[:len(codes)]
