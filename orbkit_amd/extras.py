'''Grid-based analysis helpers of orbkit/extras.py on the device path.

Same names, arguments and return values as orbkit:

  calc_mo, calc_ao, mo_set          thin wrappers over rho_compute
                                    (orbkit/extras.py:40-260)
  atom2index                        orbkit/extras.py:262-304
  gross_atomic_density              orbkit/extras.py:306-385 (native: density
                                    matrix, FP64 MFMA GEMM, per-atom reduction)
  numerical_mulliken_charges        orbkit/extras.py:386-439
  calc_jmo                          orbkit/extras.py:441-492 (native pair kernel)

Output files (``otype``): 'cube' / 'cb' are written by orbkit_amd.output.cube;
other orbkit output types are not part of this package.
'''
import ctypes

import numpy

from . import _native, core, grid
from .basis import device_basis
from .display import display
from .qcinfo import QCinfo
from .tools import require


def _qc(qc):
  return QCinfo(qc) if isinstance(qc, dict) else qc


def _copy_with_mos(qc, mo_spec):
  q = QCinfo()
  q.geo_spec, q.geo_info, q.ao_spec = qc.geo_spec, getattr(qc, 'geo_info', None), qc.ao_spec
  q.mo_spec = mo_spec
  return q


def _write(data, qc, ofid, otype, labels=None):
  if otype is None:
    return
  types = [otype] if isinstance(otype, str) else list(otype)
  from .output.cube import cube_creator
  for t in types:
    if t not in ('cube', 'cb'):
      raise NotImplementedError('output type %r (orbkit_amd writes cube files only)' % t)
    if grid.is_vector:
      raise ValueError('cube files need a regular grid')
    cube_creator(data, ofid, qc.geo_info, qc.geo_spec, comments='orbkit_amd', labels=None)


def calc_mo(qc, fid_mo_list, drv=None, otype=None, ofid=None, numproc=None, slice_length=None):
  '''Selected MOs (or derivatives) on the grid: (NMO,) + N, or (NDRV, NMO) + N.'''
  qc = _qc(qc)
  qc_select = _copy_with_mos(qc, qc.mo_spec[fid_mo_list])
  mo_list = core.rho_compute(qc_select, calc_mo=True, drv=drv,
                             slice_length=1e4 if slice_length is None else slice_length,
                             numproc=1 if numproc is None else numproc)
  if otype is not None:
    _write(mo_list, qc_select, ofid or 'orbkit_MO', otype)
  return mo_list


def calc_ao(qc, drv=None, otype=None, ofid=None, numproc=None, slice_length=None):
  '''All AOs (or derivatives) on the grid: (NAO,) + N.'''
  qc = _qc(qc)
  ao_list = core.rho_compute(qc, calc_ao=True, drv=drv)
  if otype is not None:
    _write(ao_list, qc, ofid or 'orbkit_AO', otype)
  return ao_list


def mo_set(qc, fid_mo_list, drv=None, laplacian=None, otype=None, ofid=None, return_all=True,
           numproc=None, slice_length=None):
  '''Density (and derivatives) of each MO set: datasets (NSET,) + N and, with
  derivatives, delta_datasets (NSET*NDRV (+1 with laplacian),) + N.'''
  qc = _qc(qc)
  sets = fid_mo_list if isinstance(fid_mo_list, (list, tuple)) and fid_mo_list and \
      not isinstance(fid_mo_list[0], (int, numpy.integer)) else [fid_mo_list]
  datasets, delta_datasets = [], []
  for sel in sets:
    qc_select = _copy_with_mos(qc, qc.mo_spec[sel])
    data = core.rho_compute(qc_select, drv=drv, laplacian=bool(laplacian),
                            slice_length=1e4 if slice_length is None else slice_length,
                            numproc=1 if numproc is None else numproc)
    if drv is None and not laplacian:
      rho = data
    elif laplacian:
      rho, delta_rho, laplacian_rho = data
      delta_datasets.extend(delta_rho)
      delta_datasets.append(laplacian_rho)
    else:
      rho, delta_rho = data
      delta_datasets.extend(delta_rho)
    datasets.append(rho)
  datasets = numpy.array(datasets)
  if otype is not None:
    _write(datasets, qc, ofid or 'orbkit_mo_set', otype)
  if not return_all:
    return None
  if drv is None and not laplacian:
    return datasets
  return datasets, numpy.array(delta_datasets)


def atom2index(atom, geo_info=None):
  '''Atom numbers (counting from one) -> indices into geo_info (from zero).'''
  if not isinstance(atom, (list, numpy.ndarray, range)):
    atom = [atom]
  atom = list(atom)
  if geo_info is not None:
    labels = numpy.array(geo_info)[:, 1]
    index = []
    for a in atom:
      i = numpy.argwhere(labels.astype(int) == a)
      if len(i) != 1:
        raise ValueError('No or multiple occurence of the atom number %d in geo_info!' % a)
      index.append(int(i[0, 0]))
    return atom, numpy.array(index, dtype=int)
  try:
    return atom, numpy.array(atom, dtype=int)
  except ValueError:
    raise ValueError('Cannot convert atom to integer array!')


def _row_atoms(ao_spec):
  '''Atom (0-based) of every output AO row.'''
  cont_atom = ao_spec.get_assign_cont_to_atoms()
  if ao_spec.spherical:
    return numpy.array([cont_atom[j0] for j0, _ in ao_spec.get_old_ao_spherical()], dtype=int)
  return numpy.asarray(cont_atom)[ao_spec.get_assign_lxlylz_to_cont()]


def gross_atomic_density(atom, qc, bReturnmo=False, ao_list=None, mo_list=None, drv=None):
  '''rho^a = sum_i occ_i phi_i phi_i^a for the selected atoms (counting from
  one; 'all' or -1 for every atom), phi_i^a the MO contracted over atom a's
  AO rows only.  Returns a list of arrays N (and, with bReturnmo, the
  per-atom MO lists (NMO,) + N).'''
  import torch
  qc = _qc(qc)
  if atom == 'all' or atom == -1:
    atom = range(1, len(qc.geo_info) + 1)
  atom, index = atom2index(atom, geo_info=qc.geo_info)
  display('Computing the gross atomic density with respect to the atom(s) (internal numbering)')
  ra = _row_atoms(qc.ao_spec)
  C = numpy.ascontiguousarray(qc.mo_spec.get_coeffs(), dtype=numpy.float64)
  occ = numpy.ascontiguousarray(qc.mo_spec.get_occ(), dtype=numpy.float64)
  dev = torch.device('cuda', torch.cuda.current_device())
  rows = [numpy.nonzero(ra == a)[0] for a in index]
  if ao_list is None and mo_list is None:
    basis = device_basis(require(qc.geo_spec, dtype='f'), qc.ao_spec)
    if not grid.is_initialized:
      grid.grid_init()
    dg = core.DeviceGrid(grid.x, grid.y, grid.z, regular=not grid.is_vector)
    N = (len(grid.x),) if grid.is_vector else (len(grid.x), len(grid.y), len(grid.z))
    code = core._codes([drv])[0] if drv is not None else 0
    out = core.gross_atomic(basis, dg, C, occ, code, rows)
    rho_atom = [out[j].reshape(N) for j in range(len(rows))]
  else:
    if ao_list is None:
      ao_list = core.ao_creator(qc.geo_spec, qc.ao_spec, drv=drv)
    ao = numpy.asarray(ao_list, dtype=numpy.float64)
    N = ao.shape[1:]
    A = ao.reshape(ao.shape[0], -1)
    M = core.gemm(C, A) if mo_list is None else numpy.asarray(mo_list, dtype=numpy.float64).reshape(len(C), -1)
    rho_atom = [core.masked_mo_product(C, occ, A, M, r).reshape(N) for r in rows]
  if bReturnmo:
    mo_atom = []
    ao = core.ao_creator(qc.geo_spec, qc.ao_spec, drv=drv) if ao_list is None else numpy.asarray(ao_list)
    A = ao.reshape(ao.shape[0], -1)
    for r in rows:
      Cm = numpy.zeros_like(C)
      Cm[:, r] = C[:, r]
      mo_atom.append(core.gemm(Cm, A).reshape((len(C),) + ao.shape[1:]))
    return rho_atom, mo_atom
  return rho_atom


def numerical_mulliken_charges(atom, qc, ao_list=None, mo_list=None, rho_atom=None):
  '''Gross atomic populations P^a = integral of rho^a (compensated FP64 sum
  times the volume element) and charges Z_a - P^a.  orbkit calls a
  `core.integration` that its core module does not define
  (orbkit/extras.py:429); the integral here is the grid sum times d3r.'''
  qc = _qc(qc)
  if atom == 'all' or atom == -1 or atom == [-1]:
    atom = range(1, len(qc.geo_info) + 1)
  atom, index = atom2index(atom, geo_info=qc.geo_info)
  if rho_atom is None:
    rho_atom = gross_atomic_density(atom, qc, ao_list=ao_list, mo_list=mo_list)
  if grid.is_vector:
    display('Warning: You have applied a vector grid.\nThe integration is not implemented for such a grid!\n'
            '\nReturning only the gross atomic densities...')
    return rho_atom
  GP_A = numpy.array([core.electron_count(r, grid.d3r) for r in rho_atom])
  mulliken_num = {'population': GP_A,
                  'charge': numpy.array(numpy.array(qc.geo_info)[index, 2], dtype=float) - GP_A}
  display('\nNumerical Mulliken Charges Q and Gross Atomic Populations GP:')
  for i, a in enumerate(index):
    display('\tAtom %s (%s):\tQ = %+0.4f ( GP = %0.4f )' % (qc.geo_info[a][1], qc.geo_info[a][0],
                                                            mulliken_num['charge'][i],
                                                            mulliken_num['population'][i]))
  return rho_atom, mulliken_num


def calc_jmo(qc, ij, drv=['x', 'y', 'z'], numproc=1, otype=None, ofid='', **kwargs):
  '''Transition electronic flux density components between MO pairs ij:
  -1/2 (phi_i d phi_j - phi_j d phi_i), shape (NDRV, NPAIR) + N.'''
  qc = _qc(qc)
  ij = numpy.asarray(ij)
  if ij.ndim == 1 and len(ij) == 2:
    ij = ij.reshape(1, 2)
  assert ij.ndim == 2 and ij.shape[1] == 2
  u, indices = numpy.unique(ij, return_inverse=True)
  indices = indices.reshape(-1, 2)
  qc_select = _copy_with_mos(qc, qc.mo_spec[u])
  jmo = core.transition_flux(qc_select, indices, drv)
  if otype is not None:
    _write(jmo, qc_select, ofid or 'orbkit_jmo', otype)
  return jmo
