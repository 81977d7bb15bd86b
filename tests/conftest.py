import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
  config.addinivalue_line('markers', 'gpu: needs a HIP device and liborbkit_amd.so')


@pytest.fixture(scope='session')
def golden():
  return GOLDEN


def rel_err(a, b):
  '''max |a-b| / max |b| (norm-wise relative error).'''
  a = numpy.asarray(a, dtype=float)
  b = numpy.asarray(b, dtype=float)
  den = numpy.max(numpy.abs(b)) if b.size else 1.0
  return float(numpy.max(numpy.abs(a - b)) / den) if a.size else 0.0


def gpu_available():
  try:
    import torch
    return torch.cuda.is_available()
  except Exception:
    return False


def long_list_case(nmo=70):
  '''300 atoms in a 15 A box, 6,000 spherical AOs (more than 128 chunks of 32
  rows), nmo MOs with random coefficients, and the axes of a 6^3 regular grid
  whose first brick reaches about 170 chunks: live-chunk lists past 128
  entries.  Returns (qc, axes).'''
  from orbkit_amd import workloads as W
  from orbkit_amd.orbitals import MOClass
  from orbkit_amd.qcinfo import QCinfo
  rng = numpy.random.default_rng(17)
  geo = W._atoms(rng, 300, 15.0)
  ao = W._basis(rng, geo, [['s', 's', 'p', 'p', 'd', 'f']] * 300, 0.05, 1e5, True)
  assert ao.get_ao_num() == 6000
  C = 0.05 * rng.standard_normal((nmo, 6000))
  qc = QCinfo()
  qc.geo_spec = geo
  qc.geo_info = numpy.array([['X', str(i + 1), '0'] for i in range(len(geo))])
  qc.ao_spec = ao
  qc.mo_spec = MOClass([{'coeffs': C[i], 'occ_num': 2.0, 'energy': -1.0, 'sym': '%d.1' % (i + 1)} for i in range(nmo)])
  return qc, W._grid(geo, 6)
