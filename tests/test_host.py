'''Host-side logic, argument rules and the C ABI library (no GPU needed).'''
import ctypes
import os
import re
import subprocess
import sys

import numpy
import pytest

from conftest import ROOT, rel_err
from orbkit_amd import grid, read, tools, workloads
from orbkit_amd.basis import spherical_matrix, flatten
from oracle import core_oracle as O


def test_validate_drv_table():
  '''orbkit/tools.py:225-239.'''
  want = {None: 0, 'None': 0, '': 0, 'x': 1, 'y': 2, 'z': 3, 'xx': 4, 'x2': 4, 'yy': 5, 'y2': 5,
          'zz': 6, 'z2': 6, 'xy': 7, 'yx': 7, 'xz': 8, 'zx': 8, 'yz': 9, 'zy': 9, 0: 0, 9: 9}
  for k, v in want.items():
    assert tools.validate_drv(k) == v
  for bad in ('q', 'xyz', 10, -1, 'x3'):
    with pytest.raises(ValueError):
      tools.validate_drv(bad)


def test_cart2sph_matches_orbkit_table(golden):
  """Every (l, m) row of the cart2sph table, l <= 4, equals orbkit's own
  table (tests/golden/cart2sph_table.json, read from orbkit/tools.py by
  make_cart2sph_fixture.py): the same Cartesian components in the same
  order, the same coefficients with their signs and the same global factor.
  Orthonormality checks alone would pass a row with its sign flipped."""
  import json
  rows = json.load(open(os.path.join(golden, 'cart2sph_table.json')))['rows']
  assert len(rows) == sum(2 * l + 1 for l in range(5))
  for r in rows:
    exps, coeffs, glob = tools.get_cart2sph(r['l'], r['m'])
    assert [tuple(e) for e in exps] == [tuple(e) for e in r['exps']], (r['l'], r['m'])
    assert numpy.allclose(coeffs, r['coeffs'], rtol=1e-15, atol=0), (r['l'], r['m'])
    assert numpy.sign(coeffs).tolist() == numpy.sign(r['coeffs']).tolist()
    assert abs(glob - r['global']) <= 1e-15 * abs(r['global']), (r['l'], r['m'])


def test_molden_single_label_renumbered(golden, tmp_path):
  """Files that give every MO the same label (as ORCA writes them) get the
  labels '1.a', '2.a', ... in file order (orbkit/read/molden.py:364-368);
  distinct labels are kept."""
  import re as _re
  text = open(os.path.join(golden, 'h2o_rhf_sph.molden')).read()
  same = _re.sub(r'(?im)^(\s*sym\s*=\s*)\S+', r'\g<1>a', text)
  p = tmp_path / 'same_label.molden'
  p.write_text(same)
  qc = read.main_read(str(p), all_mo=True)
  assert [mo['sym'] for mo in qc.mo_spec] == ['%d.a' % (i + 1) for i in range(len(qc.mo_spec))]
  orig = read.main_read(os.path.join(golden, 'h2o_rhf_sph.molden'), all_mo=True)
  assert len({mo['sym'] for mo in orig.mo_spec}) == len(orig.mo_spec)


def test_readers(golden):
  qc = read.main_read(os.path.join(golden, 'h2o_rhf_sph.molden'), all_mo=True)
  assert not qc.ao_spec.spherical and qc.ao_spec.get_ao_num() == 25 and len(qc.mo_spec) == 24
  assert numpy.allclose(qc.mo_spec.get_occ()[:5], 2.0) and not qc.mo_spec.get_occ()[5:].any()
  occ_only = read.main_read(os.path.join(golden, 'h2o_rhf_sph.molden'))
  assert len(occ_only.mo_spec) == 5
  # [Atoms] Angs converted to bohr
  assert abs(qc.geo_spec[0, 1] + 0.8901392951 * 1.8897261) < 1e-6
  f = read.main_read(os.path.join(golden, 'h2o_rhf_sph.fchk'))
  assert f.ao_spec.spherical and f.ao_spec.get_ao_num() == 24
  assert [c['type'] for c in f.ao_spec][:4] == ['s', 's', 'p', 's']
  assert f.ao_spec.get_old_ao_spherical()[2][1] == (1, 1)


def test_grid_module():
  grid.reset_grid()
  grid.min_ = [-1.0, 0.0, 2.0]
  grid.max_ = [1.0, 0.0, 3.0]
  grid.N_ = [5, 1, 3]
  grid.delta_ = [0, 0, 0]
  grid.grid_init()
  assert len(grid.x) == 5 and len(grid.y) == 1 and len(grid.z) == 3 and not grid.is_vector
  X, Y, Z = grid.x.copy(), grid.y.copy(), grid.z.copy()
  grid.grid2vector()
  assert grid.is_vector and len(grid.x) == 15
  # x slowest, z fastest (orbkit/cy_grid.pyx:12-30)
  v = O.expand(X, Y, Z)
  assert numpy.array_equal(numpy.stack([grid.x, grid.y, grid.z]), v)
  grid.vector2grid(5, 1, 3)
  assert numpy.array_equal(grid.x, X) and numpy.array_equal(grid.z, Z)
  m = grid.mv2g(a=numpy.arange(2 * 15).reshape(2, 15))
  assert m.shape == (2, 5, 1, 3)
  with pytest.raises(ValueError):
    grid.vector2grid(2, 2, 2)
  qc = read.main_read(os.path.join(ROOT, 'tests', 'golden', 'h2o_rhf_sph.molden'))
  grid.adjust_to_geo(qc, extend=2.0, step=1)
  grid.grid_init(force=True)
  assert (len(grid.x), len(grid.y), len(grid.z)) == (5, 9, 7)


def test_spherical_blocks_match_reference_loop():
  '''The transform folded into the kernel equals orbkit's row loop
  (orbkit/core.py:135-176) applied to the identity.'''
  qc, axes, _ = workloads.make('medium', npts_side=4)
  T = spherical_matrix(qc.ao_spec)
  ncart = int(qc.ao_spec.get_nlxlylz_per_cont().sum())
  R = O.cartesian2spherical(numpy.eye(ncart), qc.ao_spec, tools.get_cart2sph)
  assert T.shape == (400, 480) and numpy.array_equal(T, R)
  f = flatten(qc.geo_spec, qc.ao_spec)
  assert f['nout'].sum() == 400 and len(f['transform']) == sum(n * c for n, c in zip(f['nout'], f['ncart']))


def test_workload_overlap_matches_quadrature():
  '''Analytic multi-centre overlap used for the synthetic MOs vs quadrature.'''
  from orbkit_amd.orbitals import AOClass
  ao = AOClass()
  ao.append({'atom': 0, 'type': 'p', 'pnum': 2, 'coeffs': numpy.array([[1.2, 0.6], [0.4, 0.5]])})
  ao.append({'atom': 1, 'type': 'd', 'pnum': 1, 'coeffs': numpy.array([[0.9, 1.0]])})
  ao.update()
  geo = numpy.array([[0.0, 0.0, 0.0], [0.7, -0.4, 0.5]])
  S = workloads.cartesian_overlap(geo, ao)
  ax = numpy.linspace(-9, 9, 121)
  v = O.expand(ax, ax, ax)
  A = O.ao_creator(geo, ao, None, v[0], v[1], v[2], tools.get_cart2sph)
  Sq = A @ A.T * (ax[1] - ax[0]) ** 3
  assert numpy.allclose(S, Sq, atol=1e-8)


def test_argument_errors_before_device():
  from orbkit_amd import core
  qc, axes, _ = workloads.make('small', npts_side=4)
  with pytest.raises(ValueError):
    core.rho_compute(qc, calc_ao=True, calc_mo=True)
  with pytest.raises(ValueError):
    core.rho_compute_no_slice(qc, calc_ao=True, calc_mo=True)
  with pytest.raises(ValueError):
    core.ao_creator(qc.geo_spec, qc.ao_spec, x=numpy.zeros(3), y=numpy.zeros(2), z=numpy.zeros(3),
                    is_vector=True)
  grid.set_grid(*axes, is_vector=False)
  with pytest.raises(ValueError):
    core.rho_compute(qc, drv=['q'])


def _gpu():
  try:
    import torch
    return torch.cuda.is_available()
  except Exception:
    return False


@pytest.mark.skipif(_gpu(), reason='checks the no-GPU behaviour')
def test_no_cpu_fallback():
  '''Without a HIP device the product raises instead of computing.'''
  from orbkit_amd import core
  qc, axes, _ = workloads.make('small', npts_side=4)
  with pytest.raises(RuntimeError, match='HIP device'):
    core.ao_creator(qc.geo_spec, qc.ao_spec, x=axes[0], y=axes[1], z=axes[2], is_vector=False)


def _header_functions():
  txt = open(os.path.join(ROOT, 'include', 'orbkit_amd.h')).read()
  txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
  return re.findall(r'^\s*(?:const\s+)?[a-z_0-9]+\*?\s+\*?(okmi_[a-z_0-9]+)\s*\(', txt, flags=re.M)


def test_library_exports_header_symbols():
  from orbkit_amd import _native
  if not os.path.exists(_native.LIB_PATH):
    pytest.skip('liborbkit_amd.so not built')
  names = _header_functions()
  assert len(names) >= 12
  lib = ctypes.CDLL(_native.LIB_PATH)
  for n in names:
    assert hasattr(lib, n), n
  assert set(names) <= set(_native.SIGNATURES), set(names) - set(_native.SIGNATURES)
  assert lib.okmi_basis_nao(None) == -1


def test_bench_refuses_multi_gpu():
  r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'], capture_output=True,
                     text=True, timeout=120)
  assert r.returncode == 2 and 'one GPU' in r.stderr
