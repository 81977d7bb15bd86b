'''Parity of the HIP path with the CPU oracle and orbkit's golden data (FP64).

Bar: norm-wise relative error <= 1e-12 (max|gpu - oracle| / max|oracle|)
for every AO / MO / density array, and the reference test's own comparator
(numpy.allclose rtol=1e-3, atol=1e-5) against the golden files.
'''
import os

import numpy
import pytest

from conftest import long_list_case, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope='module')
def ok():
  import torch
  if not torch.cuda.is_available():
    pytest.skip('no HIP device')
  from orbkit_amd import core, grid, read, tools, workloads
  from oracle import core_oracle
  return dict(core=core, grid=grid, read=read, tools=tools, W=workloads, O=core_oracle)


def _pts(axes, n, seed=0):
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  P = numpy.stack([X.ravel(), Y.ravel(), Z.ravel()])
  idx = numpy.random.default_rng(seed).choice(P.shape[1], size=min(n, P.shape[1]), replace=False)
  return P[0, idx].copy(), P[1, idx].copy(), P[2, idx].copy()


@pytest.mark.parametrize('name', ['small', 'medium'])
def test_ao_all_drv(ok, name):
  qc, axes, _ = ok['W'].make(name, npts_side=30)
  x, y, z = _pts(axes, 3001)
  print('case', name, 'npts', len(x), flush=True)
  for drv in [None, 'x', 'y', 'z', 'xx', 'yy', 'zz', 'xy', 'xz', 'yz']:
    got = ok['core'].ao_creator(qc.geo_spec, qc.ao_spec, drv=drv, x=x, y=y, z=z, is_vector=True)
    ref = ok['O'].ao_creator(qc.geo_spec, qc.ao_spec, drv, x, y, z, ok['tools'].get_cart2sph)
    assert got.shape == ref.shape
    e = rel_err(got, ref)
    print(name, drv, e)
    assert e <= TOL, (name, drv, e)


def test_ao_regular_grid_shape(ok):
  qc, axes, _ = ok['W'].make('small', npts_side=11)
  ax = [a[:7] for a in axes]
  got = ok['core'].ao_creator(qc.geo_spec, qc.ao_spec, x=ax[0], y=ax[1], z=ax[2][:5], is_vector=False)
  assert got.shape == (25, 7, 7, 5)
  v = ok['O'].expand(ax[0], ax[1], ax[2][:5])
  ref = ok['O'].ao_creator(qc.geo_spec, qc.ao_spec, None, v[0], v[1], v[2], ok['tools'].get_cart2sph)
  assert rel_err(got.reshape(25, -1), ref) <= TOL


def test_golden_rho_compute(ok, golden):
  '''orbkit/test/grid_based/rho_compute.py on the GPU path.'''
  core, grid = ok['core'], ok['grid']
  qc = ok['read'].main_read(os.path.join(golden, 'h2o_rhf_sph.molden'), all_mo=True)
  grid.adjust_to_geo(qc, extend=2.0, step=1)
  grid.grid_init(is_vector=False, force=True)
  drv = [None, 'x', 'y', 'z', 'xx', 'xy', 'xz', 'yy', 'yz', 'zz']
  data = []
  for i in range(2):
    if i:
      grid.grid2vector()
    data.append([core.rho_compute(qc, slice_length=0),
                 core.rho_compute(qc, numproc=2),
                 core.rho_compute(qc, laplacian=True, slice_length=0)[-1],
                 core.rho_compute(qc, laplacian=True, numproc=2)[-1],
                 core.rho_compute(qc, calc_mo=True, drv=drv, slice_length=0),
                 core.rho_compute(qc, calc_mo=True, drv=drv, numproc=2)])
  data[1] = [grid.mv2g(d=i) for i in data[1]]
  for a, b in zip(data[0], data[1]):
    assert numpy.allclose(a, b, rtol=1e-3, atol=1e-5)
  ref = numpy.load(os.path.join(golden, 'refdata_rho_compute.npz'))
  for i, k in enumerate(['zero', 'one', 'two', 'three', 'four']):
    assert data[0][i].shape == ref[k].shape
    assert numpy.allclose(data[0][i], ref[k], rtol=1e-3, atol=1e-5)
    e = rel_err(data[0][i], ref[k])
    print('golden', k, e)
    assert e <= 1e-12, (k, e)


def test_cube_kat(ok, golden):
  '''orbkit/test/grid_based/cube_files.py (Gaussian cubegen, spherical basis).'''
  core, grid = ok['core'], ok['grid']
  dx = 4.970736
  grid.x = numpy.arange(3) * dx - 4.970736
  grid.y = numpy.arange(3) * dx - 4.970736
  grid.z = numpy.arange(3) * dx - 4.732975
  grid.is_initialized = True
  grid.is_regular = True
  grid.is_vector = False
  qc = ok['read'].main_read(os.path.join(golden, 'h2o_rhf_sph.fchk'), all_mo=False)
  rho, drho = core.rho_compute(qc, drv='xyz')
  rho2, drho2, lap = core.rho_compute(qc, laplacian=True)
  assert numpy.allclose(rho, rho2, rtol=1e-3, atol=1e-5)
  g = numpy.genfromtxt(os.path.join(golden, 'h2o_rhf_sph_grad.cube'), skip_header=9).reshape(-1)
  refrho = g[0::4].reshape(3, 3, 3)
  refd = numpy.stack([g[1::4], g[2::4], g[3::4]]).reshape(3, 3, 3, 3)
  reflap = numpy.genfromtxt(os.path.join(golden, 'h2o_rhf_sph_laplace.cube'), skip_header=9).reshape(3, 3, 3)
  assert numpy.allclose(rho, refrho, rtol=1e-3, atol=1e-5)
  assert numpy.allclose(drho, refd, rtol=1e-1, atol=1e-3)
  assert numpy.allclose(lap, reflap, rtol=1e-3, atol=1e-5)
  qc.mo_spec = qc.mo_spec[1:]
  val = core.rho_compute(qc)
  refv = numpy.genfromtxt(os.path.join(golden, 'h2o_rhf_sph.cube'), skip_header=9).reshape(val.shape)
  assert numpy.allclose(val, refv, rtol=1e-3, atol=1e-5)
  qc.mo_spec = qc.mo_spec['homo']
  homo = core.rho_compute(qc, calc_mo=True)
  refh = numpy.genfromtxt(os.path.join(golden, 'h2o_rhf_sph_mo.cube'), skip_header=10).reshape(homo.shape)
  assert numpy.allclose(homo, refh, rtol=1e-3, atol=1e-5)


@pytest.mark.parametrize('name', ['small', 'medium'])
def test_rho_grad_lap_vs_oracle(ok, name):
  core, grid, O = ok['core'], ok['grid'], ok['O']
  qc, axes, _ = ok['W'].make(name, npts_side=40)
  x, y, z = _pts(axes, 4099, seed=1)
  print('case', name, flush=True)
  grid.set_grid(x, y, z, is_vector=True)
  rho, d1 = core.rho_compute(qc, drv=['x', 'y', 'z'], slice_length=1000)
  _, d2, lap = core.rho_compute(qc, laplacian=True)
  r_rho, r_d1 = O.rho_compute(qc, x, y, z, True, drv=['x', 'y', 'z'], slice_length=1000)
  _, r_d2, r_lap = O.rho_compute(qc, x, y, z, True, laplacian=True)
  for a, b, what in [(rho, r_rho, 'rho'), (d1, r_d1, 'grad'), (d2, r_d2, 'dd'), (lap, r_lap, 'lap')]:
    e = rel_err(a, b)
    print(name, what, e)
    assert e <= TOL, (what, e)


def test_mixed_and_mo_derivatives(ok):
  core, grid, O = ok['core'], ok['grid'], ok['O']
  qc, axes, _ = ok['W'].make('medium', npts_side=30)
  x, y, z = _pts(axes, 1500, seed=2)
  grid.set_grid(x, y, z, is_vector=True)
  drv = [None, 'x', 'xy', 'yz', 'zz', 'xz']
  got = core.rho_compute(qc, calc_mo=True, drv=drv)
  ref = O.rho_compute(qc, x, y, z, True, calc_mo=True, drv=drv)
  assert got.shape == ref.shape == (6, 40, len(x))
  for j in range(len(drv)):
    assert rel_err(got[j], ref[j]) <= TOL
  rho, dmix = core.rho_compute(qc, drv=['xy', 'xz', 'yz', 'x2'])
  _, rmix = O.rho_compute(qc, x, y, z, True, drv=['xy', 'xz', 'yz', 'x2'])
  assert rel_err(dmix, rmix) <= TOL
  ao = core.rho_compute(qc, calc_ao=True, drv=['x', 'yy'])
  rao = O.rho_compute(qc, x, y, z, True, calc_ao=True, drv=['x', 'yy'])
  assert ao.shape == rao.shape and rel_err(ao, rao) <= TOL


def test_repeated_derivative_codes(ok):
  '''A derivative listed twice (drv='xx' iterates to ['x', 'x'],
  orbkit/core.py:401-406; or an explicit list) yields equal sets: the AO
  kernel writes each code once and the repeats are device copies.'''
  core, grid, O = ok['core'], ok['grid'], ok['O']
  qc, axes, _ = ok['W'].make('medium', npts_side=30)
  x, y, z = _pts(axes, 1200, seed=5)
  grid.set_grid(x, y, z, is_vector=True)
  print('case repeated drv', flush=True)
  for drv in ('xx', ['y', None, 'y', 'xz', 'xz']):
    ao = core.rho_compute(qc, calc_ao=True, drv=drv)
    rao = O.rho_compute(qc, x, y, z, True, calc_ao=True, drv=drv)
    assert ao.shape == rao.shape and rel_err(ao, rao) <= TOL, drv
    assert numpy.array_equal(ao[0], ao[2 if len(ao) > 2 else 1])
    rho, d = core.rho_compute(qc, drv=drv)
    _, rd = O.rho_compute(qc, x, y, z, True, drv=drv)
    assert rel_err(d, rd) <= TOL, drv


def test_long_derivative_lists(ok):
  """drv lists longer than one native call takes (10 AO / MO codes, 16
  density outputs) are split into several calls, as orbkit takes any length:
  24 labels for the AOs, MOs and the density (FP64 and bf16x6)."""
  core, grid, O = ok['core'], ok['grid'], ok['O']
  qc, axes, _ = ok['W'].make('small', npts_side=20)
  x, y, z = _pts(axes, 700, seed=9)
  grid.set_grid(x, y, z, is_vector=True)
  drv = ['x', 'y', 'z', 'xx', 'yy', 'zz', 'xy', 'xz', 'yz', None, 'z', 'y'] * 2
  print('case long drv', len(drv), flush=True)
  ao = core.rho_compute(qc, calc_ao=True, drv=drv)
  assert ao.shape[0] == len(drv) and rel_err(ao, O.rho_compute(qc, x, y, z, True, calc_ao=True, drv=drv)) <= TOL
  mo = core.rho_compute(qc, calc_mo=True, drv=drv)
  assert mo.shape[0] == len(drv) and rel_err(mo, O.rho_compute(qc, x, y, z, True, calc_mo=True, drv=drv)) <= TOL
  rho, d = core.rho_compute(qc, drv=drv)
  r_rho, rd = O.rho_compute(qc, x, y, z, True, drv=drv)
  assert d.shape[0] == len(drv) and rel_err(rho, r_rho) <= TOL and rel_err(d, rd) <= TOL
  frho, fd = core.rho_compute(qc, drv=drv, precision='bf16x6')
  assert fd.shape == d.shape and numpy.linalg.norm(fd - d) <= 1e-6 * numpy.linalg.norm(d)


def test_mo_creator_and_cart2sph(ok):
  core, O = ok['core'], ok['O']
  qc, axes, _ = ok['W'].make('medium', npts_side=20)
  x, y, z = _pts(axes, 777, seed=3)
  ao = core.ao_creator(qc.geo_spec, qc.ao_spec, x=x, y=y, z=z, is_vector=True)
  mo = core.mo_creator(ao, qc.mo_spec)
  assert rel_err(mo, O.mocreator(ao, qc.mo_spec.get_coeffs())) <= TOL
  assert rel_err(mo, qc.mo_spec.get_coeffs() @ ao) <= TOL
  qc.ao_spec.spherical = False
  qc.ao_spec.update()
  aoc = core.ao_creator(qc.geo_spec, qc.ao_spec, x=x, y=y, z=z, is_vector=True)
  assert aoc.shape[0] == 480
  qc.ao_spec.spherical = True
  qc.ao_spec.update()
  sph = core.cartesian2spherical(aoc, qc.ao_spec)
  assert rel_err(sph, O.cartesian2spherical(aoc, qc.ao_spec, ok['tools'].get_cart2sph)) <= TOL
  assert rel_err(sph, ao) <= TOL


def test_renorm_prenormalised_and_custom_order(ok):
  '''ORCA-style row factor N, pre-normalised primitives (pnum < 0) and a
  permuted Cartesian component order.'''
  core, O, tools = ok['core'], ok['O'], ok['tools']
  qc, axes, _ = ok['W'].make('small', npts_side=20)
  x, y, z = _pts(axes, 500, seed=4)
  rng = numpy.random.default_rng(7)
  qc.ao_spec[0]['N'] = rng.uniform(0.5, 1.5, (25, 1))
  for c in qc.ao_spec:
    c['pnum'] = -abs(c['pnum'])
    if c['type'] == 'd':
      c['lxlylz'] = [(0, 1, 1), (2, 0, 0), (1, 0, 1), (0, 2, 0), (1, 1, 0), (0, 0, 2)]
  qc.ao_spec.update()
  for drv in [None, 'y', 'zz', 'xy']:
    got = core.ao_creator(qc.geo_spec, qc.ao_spec, drv=drv, x=x, y=y, z=z, is_vector=True)
    ref = O.ao_creator(qc.geo_spec, qc.ao_spec, drv, x, y, z, tools.get_cart2sph)
    assert rel_err(got, ref) <= TOL, drv


def test_no_slice_components(ok):
  core, grid, O = ok['core'], ok['grid'], ok['O']
  qc, axes, _ = ok['W'].make('small', npts_side=9)
  out = core.rho_compute_no_slice(qc, drv=['x', 'zz'], return_components=True,
                                  x=axes[0], y=axes[1], z=axes[2], is_vector=False)
  ao, mo, rho, dao, dmo, drho = out
  assert ao.shape == (25, 9, 9, 9) and mo.shape == (5, 9, 9, 9) and rho.shape == (9, 9, 9)
  assert dao.shape == (2, 25, 9, 9, 9) and dmo.shape == (2, 5, 9, 9, 9) and drho.shape == (2, 9, 9, 9)
  r_rho, r_d = O.rho_compute(qc, axes[0], axes[1], axes[2], False, drv=['x', 'zz'])
  assert rel_err(rho, r_rho) <= TOL and rel_err(drho, r_d) <= TOL
  grid.set_grid(*axes, is_vector=False)
  lap = core.rho_compute(qc, laplacian=True, numproc=0)
  r_lap = O.rho_compute(qc, axes[0], axes[1], axes[2], False, laplacian=True)
  for a, b in zip(lap, r_lap):
    assert rel_err(a, b) <= TOL


@pytest.mark.parametrize('npts', [1, 63, 64, 65, 1000])
def test_ragged_sizes(ok, npts):
  core, O, tools = ok['core'], ok['O'], ok['tools']
  qc, axes, _ = ok['W'].make('small', npts_side=20)
  x, y, z = _pts(axes, npts, seed=npts)
  got = core.ao_creator(qc.geo_spec, qc.ao_spec, drv='x', x=x, y=y, z=z, is_vector=True)
  ref = O.ao_creator(qc.geo_spec, qc.ao_spec, 'x', x, y, z, tools.get_cart2sph)
  assert rel_err(got, ref) <= TOL
  ok['grid'].set_grid(x, y, z, is_vector=True)
  r = core.rho_compute(qc)
  assert rel_err(r, O.rho_compute(qc, x, y, z, True)) <= TOL


def test_empty_grid(ok):
  core = ok['core']
  qc, axes, _ = ok['W'].make('small', npts_side=5)
  e = numpy.zeros(0)
  ao = core.ao_creator(qc.geo_spec, qc.ao_spec, x=e, y=e, z=e, is_vector=True)
  assert ao.shape == (25, 0)


@pytest.mark.parametrize('drv', [['x', 'yy'], ['x', 'y', 'z'], ['x', 'y', 'z', 'xx', 'yy', 'zz']])
def test_slice_rho_direct(ok, drv):
  '''slice_rho's three outputs, MO norms included, over an 800-point slice
  (the last 64-point tile is partial); ['x', 'y', 'z'] and the six-label
  list use every code 0..3 / 0..6, the FP64 kernel without set branches.'''
  core, grid, O = ok['core'], ok['grid'], ok['O']
  qc, axes, _ = ok['W'].make('small', npts_side=12)
  grid.set_grid(*axes, is_vector=False)
  grid.grid2vector()
  print('case slice_rho', drv, flush=True)
  core.initializer({'qc': qc, 'calc_ao': False, 'calc_mo': False, 'derivative': drv})
  rho, mo_norm, delta = core.slice_rho([100, 900])
  r = O.slice_rho(qc.geo_spec, qc.ao_spec, qc.mo_spec, grid.x[100:900], grid.y[100:900], grid.z[100:900],
                  drv, False, False, ok['tools'].get_cart2sph)
  assert rel_err(rho, r[0]) <= TOL and rel_err(mo_norm, r[1]) <= TOL and rel_err(delta, r[2]) <= TOL


@pytest.mark.parametrize('regular,ws_points', [(True, None), (False, None), (False, 256), (True, 320)])
def test_okmi_rho_fused_and_staged(ok, regular, ws_points):
  """okmi_rho (fused FP64 kernel up to 64 MOs) against okmi_rho_staged and
  the oracle: rho, first and second derivatives and the MO norms, on a
  regular grid (bricks) and a vector grid, with a workspace that holds the
  whole grid or only a few hundred points (several passes, the MO norms
  accumulated pass by pass; a regular grid then runs as vector passes)."""
  import ctypes
  import torch
  from orbkit_amd import _native
  from orbkit_amd.basis import device_basis
  from orbkit_amd.core import DeviceGrid
  O = ok['O']
  qc, axes, _ = ok['W'].make('medium', npts_side=13, nmo=20)
  basis = device_basis(qc.geo_spec, qc.ao_spec)
  C, occ = qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ()
  if regular:
    dg = DeviceGrid(*axes, regular=True)
    x, y, z = O.expand(*axes)
  else:
    x, y, z = _pts(axes, 1777, seed=7)
    dg = DeviceGrid(x, y, z, regular=False)
  codes = numpy.array([1, 2, 3, 4, 6, 8], dtype=numpy.intc)
  L, dev = _native.lib(), dg.x.device
  dC, docc = torch.as_tensor(C, device=dev), torch.as_tensor(occ, device=dev)
  st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
  out = {}
  for name in ('okmi_rho', 'okmi_rho_staged'):
    print('case', name, 'regular' if regular else 'vector', ws_points, flush=True)
    if name == 'okmi_rho' and ws_points is not None:
      nb = L.okmi_rho_fused_workspace_bytes(basis.handle, C.shape[0], 0, ws_points)
    else:
      nb = L.okmi_workspace_bytes(basis.handle, 7, dg.npts if ws_points is None else ws_points)
    ws = torch.empty(nb // 8 + 1, dtype=torch.float64, device=dev)
    rho = torch.empty(dg.npts, dtype=torch.float64, device=dev)
    delta = torch.empty((len(codes), dg.npts), dtype=torch.float64, device=dev)
    nrm = torch.full((C.shape[0],), 0.5, dtype=torch.float64, device=dev)   # accumulated onto
    _native.check(getattr(L, name)(basis.handle, ctypes.byref(dg.desc), 0, dg.npts, ctypes.c_void_p(dC.data_ptr()),
                                   ctypes.c_void_p(docc.data_ptr()), C.shape[0], codes.ctypes.data_as(_native.P_int),
                                   len(codes), ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(delta.data_ptr()),
                                   dg.npts, ctypes.c_void_p(nrm.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                                   nb, st), name)
    out[name] = (rho.cpu().numpy(), delta.cpu().numpy(), nrm.cpu().numpy() - 0.5)
  r = O.slice_rho(qc.geo_spec, qc.ao_spec, qc.mo_spec, x, y, z, ['x', 'y', 'z', 'xx', 'zz', 'xz'], False, False,
                  ok['tools'].get_cart2sph)
  for name, (rho, d, nrm) in out.items():
    e = (rel_err(rho, r[0]), rel_err(d, r[2]), rel_err(nrm, r[1]))
    print(name, e)
    assert max(e) <= TOL, (name, e)
  f, s_ = out['okmi_rho'], out['okmi_rho_staged']
  assert rel_err(f[0], s_[0]) <= TOL and rel_err(f[1], s_[1]) <= TOL and rel_err(f[2], s_[2]) <= TOL


@pytest.mark.parametrize('name,nmo,side', [('small', None, 21), ('medium', 5, 17), ('medium', 12, 14),
                                           ('medium', 20, 15), ('medium', 40, 19), ('medium', 64, 13)])
def test_fused_f64_vs_oracle(ok, name, nmo, side):
  """FP64 fused kernel (AO values in FP64 in LDS, FP64 MFMA, FP64
  accumulation) against the oracle at the FP64 bar: rho, the three first
  derivatives and every second-derivative code (all four passes), on a
  regular grid tiled in bricks (side not a multiple of 4) and on a vector
  grid of ragged length; MOs in 16-MO tiles (16x16x4 MFMA) and 4-MO groups
  (4x4x4 MFMA): 5 -> 2 groups, 12 -> 3 groups, 20 -> 1 tile + 1 group,
  40 -> 2 tiles + 2 groups, 64 -> 4 tiles."""
  from orbkit_amd.basis import device_basis
  from orbkit_amd.core import DeviceGrid
  from orbkit_amd.fused import rho_fused
  O = ok['O']
  kw = {} if nmo is None else {'nmo': nmo}
  qc, axes, _ = ok['W'].make(name, npts_side=side, **kw)
  basis = device_basis(qc.geo_spec, qc.ao_spec)
  C, occ = qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ()
  codes = [1, 2, 3, 4, 5, 6, 7, 8, 9]
  drv = ['x', 'y', 'z', 'xx', 'yy', 'zz', 'xy', 'xz', 'yz']
  x, y, z = _pts(axes, 2113, seed=5)
  for regular in (True, False):
    print('case fused f64', name, nmo, 'regular' if regular else 'vector', flush=True)
    if regular:
      dg = DeviceGrid(*axes, regular=True)
      rx, ry, rz = O.expand(*axes)
    else:
      dg = DeviceGrid(x, y, z, regular=False)
      rx, ry, rz = x, y, z
    rho, d = rho_fused(basis, dg, C, occ, codes, precision='f64')
    r_rho, r_d = O.rho_compute(qc, rx, ry, rz, True, drv=drv)
    e_rho, e_d = rel_err(rho, r_rho), max(rel_err(d[i], r_d[i]) for i in range(len(drv)))
    print(name, nmo, regular, e_rho, e_d)
    assert e_rho <= TOL and e_d <= TOL, (regular, e_rho, e_d)
    rho1, _ = rho_fused(basis, dg, C, occ, [], precision='f64')
    assert rel_err(rho1, r_rho) <= TOL


# The precision target: <= 1e-7 relative density error against FP64 (norm-wise)
# for the FP32-operand paths at the medium workload's full size.  Measured:
# bf16x6 rho 3.2e-8, grad 3.6e-8, Laplacian 2.2e-8; f32 rho 4.0e-8, grad
# 3.0e-8, Laplacian 3.6e-8; the largest pointwise rho error (rho > 1e-10 max
# rho) 4.6e-7 / 5.2e-7, which FP32 operands cannot bring below 1e-7
# (test_fp32_operand_floor in test_precision_floor.py).  The errors depend on
# the order of the FP32 accumulation over the AO rows (with the rows ordered
# by screening radius alone bf16x6 measured rho 7.5e-8, grad 5.9e-8).
# bf16x3 (two bf16 parts, 16 significant bits): rho 1.9e-6, pointwise 2.0e-5.
@pytest.mark.parametrize('prec,l2_bar,max_bar', [('f32', 1e-7, 1e-6), ('bf16x6', 1e-7, 1e-6), ('bf16x3', 5e-6, 5e-5)])
def test_fused_precision(ok, prec, l2_bar, max_bar):
  """Reduced-precision fused path against the FP64 path (which equals the
  oracle to 1e-12 in the tests above) on the medium workload's full 100^3
  grid: rho, grad rho and the Laplacian, norm-wise, and rho pointwise."""
  core, grid = ok['core'], ok['grid']
  qc, axes, _ = ok['W'].make('medium')
  grid.set_grid(*axes, is_vector=False)
  print('case fused precision', prec, flush=True)
  rho, g = core.rho_compute(qc, drv=['x', 'y', 'z'])
  frho, fg = core.rho_compute(qc, drv=['x', 'y', 'z'], precision=prec)
  _, _, lap = core.rho_compute(qc, laplacian=True)
  _, _, flap = core.rho_compute(qc, laplacian=True, precision=prec)
  mask = rho > 1e-10 * rho.max()
  errs = {'rho': numpy.linalg.norm(frho - rho) / numpy.linalg.norm(rho),
          'grad': numpy.linalg.norm(fg - g) / numpy.linalg.norm(g),
          'lap': numpy.linalg.norm(flap - lap) / numpy.linalg.norm(lap),
          'rho_max': float(numpy.max(numpy.abs(frho - rho)[mask] / rho[mask]))}
  print(prec, errs)
  for k in ('rho', 'grad', 'lap'):
    assert errs[k] <= l2_bar, (k, errs[k])
  assert errs['rho_max'] <= max_bar, errs['rho_max']
  frho2, fd2 = core.rho_compute(qc, drv=['x', 'y', 'z', 'xy'], precision=prec)   # mixed-derivative pass
  rho2, d2 = core.rho_compute(qc, drv=['x', 'y', 'z', 'xy'])
  assert numpy.linalg.norm(fd2 - d2) / numpy.linalg.norm(d2) <= l2_bar


# bars just above the measured errors (bf16x6 rho / grad: 5 MOs 1.0e-7 /
# 8.8e-8, 20 MOs 2.9e-8 / 4.7e-8, 40 MOs 2.7e-8 / 4.4e-8; bf16x3 up to
# 4.4e-6 / 1.1e-5): fewer MOs average the rounding over fewer terms
@pytest.mark.parametrize('nmo', [5, 20, 40])
@pytest.mark.parametrize('prec', ['bf16x6', 'bf16x3'])
def test_fused_mo_tiles(ok, prec, nmo):
  '''rho and grad rho (the schedule that shares the C operand through LDS)
  with one, two and three 16-MO tiles and two or three bf16 planes, against
  the FP64 path.'''
  core, grid = ok['core'], ok['grid']
  qc, axes, _ = ok['W'].make('medium', npts_side=17, nmo=nmo)
  grid.set_grid(*axes, is_vector=False)
  print('case fused tiles', prec, nmo, flush=True)
  rho, g = core.rho_compute(qc, drv=['x', 'y', 'z'])
  frho, fg = core.rho_compute(qc, drv=['x', 'y', 'z'], precision=prec)
  e_rho = numpy.linalg.norm(frho - rho) / numpy.linalg.norm(rho)
  e_g = numpy.linalg.norm(fg - g) / numpy.linalg.norm(g)
  print(prec, nmo, e_rho, e_g)
  tol = {'bf16x6': {5: 2e-7, 20: 1.5e-7, 40: 1e-7}[nmo], 'bf16x3': 2e-5}[prec]
  assert e_rho <= tol and e_g <= tol, (e_rho, e_g)


# measured (largest of the seven errors): bf16x6 6.5e-8 (the Laplacian), f32
# 1.4e-7 (the vector grid's xx row), bf16x3 5.0e-6
@pytest.mark.parametrize('prec,tol', [('f32', 2.5e-7), ('bf16x6', 2.5e-7), ('bf16x3', 2e-5)])
def test_many_mo_precision(ok, prec, tol):
  '''More MOs than the fused kernel holds (130: two 128-MO tiles, the second
  ragged): FP32 AO chunks + MFMA GEMM path vs the FP64 path, all four AO-set
  widths (rho; rho+grad; +Laplacian; +mixed second derivatives), regular grid
  and a vector grid whose length is no multiple of any tile.'''
  core, grid = ok['core'], ok['grid']
  qc, axes, _ = ok['W'].make('medium', npts_side=19, nmo=130)   # 19^3: bricks padded at the faces
  print('case many-MO', prec, flush=True)
  grid.set_grid(*axes, is_vector=False)
  errs = {}
  r64 = core.rho_compute(qc)
  rlo = core.rho_compute(qc, precision=prec)
  errs['rho'] = numpy.linalg.norm(rlo - r64) / numpy.linalg.norm(r64)
  rho, d = core.rho_compute(qc, drv=['x', 'y', 'z'])
  frho, fd = core.rho_compute(qc, drv=['x', 'y', 'z'], precision=prec)
  errs['rho4'] = numpy.linalg.norm(frho - rho) / numpy.linalg.norm(rho)
  errs['grad'] = numpy.linalg.norm(fd - d) / numpy.linalg.norm(d)
  rho, d, lap = core.rho_compute(qc, laplacian=True)
  frho, fd, flap = core.rho_compute(qc, laplacian=True, precision=prec)
  errs['lap'] = numpy.linalg.norm(flap - lap) / numpy.linalg.norm(lap)
  rho, d = core.rho_compute(qc, drv=['xy', 'z', None])
  frho, fd = core.rho_compute(qc, drv=['xy', 'z', None], precision=prec)
  errs['mixed'] = numpy.linalg.norm(fd - d) / numpy.linalg.norm(d)
  x, y, z = _pts(axes, 3001, seed=3)
  grid.set_grid(x, y, z, is_vector=True)
  rho, d = core.rho_compute(qc, drv=['x', 'zz'])
  frho, fd = core.rho_compute(qc, drv=['x', 'zz'], precision=prec)
  errs['vec_rho'] = numpy.linalg.norm(frho - rho) / numpy.linalg.norm(rho)
  errs['vec_d'] = numpy.linalg.norm(fd - d) / numpy.linalg.norm(d)
  print(prec, errs)
  for k, v in errs.items():
    assert v <= tol, (k, v)


def test_large_basis_slices(ok):
  '''2,950 spherical AOs (s, p, d, f on 200 atoms), 400 MOs: the many-MO
  bf16x6 path and the FP64 path on a 20,000-point vector sample of the
  128^3 grid (FP64 vs oracle on 2,000 of them), and all 400 MOs on an 8^3
  subgrid vs the oracle.'''
  core, grid, O = ok['core'], ok['grid'], ok['O']
  qc, axes, info = ok['W'].make('large')
  assert info['nao'] == 2950 and info['nmo'] == 400
  print('case large', flush=True)
  x, y, z = _pts(axes, 20000, seed=11)
  grid.set_grid(x, y, z, is_vector=True)
  rho, d = core.rho_compute(qc, drv=['x', 'y', 'z'])
  frho, fd = core.rho_compute(qc, drv=['x', 'y', 'z'], precision='bf16x6')
  e_rho = numpy.linalg.norm(frho - rho) / numpy.linalg.norm(rho)
  e_d = numpy.linalg.norm(fd - d) / numpy.linalg.norm(d)
  print('large bf16x6 vs f64', e_rho, e_d)
  assert e_rho <= 1e-7 and e_d <= 1e-7
  r_o = O.rho_compute(qc, x[:2000], y[:2000], z[:2000], True, drv=['x', 'y', 'z'])
  assert rel_err(rho[:2000], r_o[0]) <= TOL and rel_err(d[:, :2000], r_o[1]) <= TOL
  sub = [a[60:68] for a in axes]
  grid.set_grid(*sub, is_vector=False)
  mo = core.rho_compute(qc, calc_mo=True)
  X, Y, Z = numpy.meshgrid(*sub, indexing='ij')
  mo_o = O.rho_compute(qc, X.ravel(), Y.ravel(), Z.ravel(), True, calc_mo=True)
  assert mo.shape == (400, 8, 8, 8)
  assert rel_err(mo.reshape(400, -1), mo_o) <= TOL


@pytest.mark.parametrize('name,nmo', [('medium', 70), ('medium', 260), ('large', 520)])
def test_staged_mo_shares(ok, name, nmo):
  '''The FP64 path beyond 64 MOs splits the 16-MO tiles between the waves of
  a workgroup in contiguous shares, each run in passes of at most four tiles
  with the exact count of the last one: 5 tiles over 2 waves (passes of 2
  and 3), 17 over 8 waves (2 and 3), 33 over 4 waves (4 + 4 and 4 + 4 + 1).
  rho and rho + grad rho on a screened regular grid and on a vector grid,
  against the oracle.'''
  core, grid, O = ok['core'], ok['grid'], ok['O']
  qc, axes, _ = ok['W'].make(name, npts_side=9, nmo=nmo)
  print('case staged MO shares', name, nmo, flush=True)
  grid.set_grid(*axes, is_vector=False)
  rho = core.rho_compute(qc)
  rho4, d = core.rho_compute(qc, drv=['x', 'y', 'z'])
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  r_o = O.rho_compute(qc, X.ravel(), Y.ravel(), Z.ravel(), True, drv=['x', 'y', 'z'])
  assert rel_err(rho.ravel(), r_o[0]) <= TOL and rel_err(rho4.ravel(), r_o[0]) <= TOL
  assert rel_err(d.reshape(3, -1), r_o[1]) <= TOL
  x, y, z = _pts(axes, 777, seed=5)
  grid.set_grid(x, y, z, is_vector=True)
  rho, d = core.rho_compute(qc, drv=['x', 'y', 'z'])
  r_o = O.rho_compute(qc, x, y, z, True, drv=['x', 'y', 'z'])
  assert rel_err(rho, r_o[0]) <= TOL and rel_err(d, r_o[1]) <= TOL


def test_long_live_lists(ok):
  '''A basis of more than 128 chunks of 32 rows (300 atoms in a 15 A box,
  6,000 spherical AOs) on a coarse grid whose first brick holds the whole
  molecule and reaches nearly every chunk: the live-chunk
  lists run past 128 entries, so the FP64 staged contraction and the many-MO
  bf16x6 contraction take their long-list instances (entries past 128 read
  from memory).  70 MOs (random coefficients: the density needs no
  orthonormality); FP64 against the oracle, bf16x6 against FP64.'''
  core, grid, O = ok['core'], ok['grid'], ok['O']
  qc, axes = long_list_case()
  print('case long live lists', flush=True)
  # the first brick's list runs past 128 chunks, with the FP64 radii and with
  # the smaller ones of the FP32-ending paths (okmi_fused_live_rows counts
  # live rows x points)
  import ctypes
  from orbkit_amd import _native
  from orbkit_amd.basis import device_basis
  b = device_basis(qc.geo_spec, qc.ao_spec)
  L = _native.lib()
  ax = [numpy.ascontiguousarray(a[:4], dtype=numpy.float64) for a in axes]
  P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
  for prec in ('f64', 'bf16x6'):
    out = ctypes.c_double()
    _native.check(L.okmi_fused_live_rows(b.handle, P(ax[0]), 4, P(ax[1]), 4, P(ax[2]), 4, _native.PREC[prec],
                                         ctypes.byref(out)), 'okmi_fused_live_rows')
    print(prec, 'live chunks of the first brick', out.value / 64 / 32)
    assert out.value > 128 * 32 * 64, (prec, out.value)
  grid.set_grid(*axes, is_vector=False)
  rho, d = core.rho_compute(qc, drv=['x', 'y', 'z'])
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  r_o = O.rho_compute(qc, X.ravel(), Y.ravel(), Z.ravel(), True, drv=['x', 'y', 'z'])
  assert rel_err(rho.ravel(), r_o[0]) <= TOL and rel_err(d.reshape(3, -1), r_o[1]) <= TOL
  frho, fd = core.rho_compute(qc, drv=['x', 'y', 'z'], precision='bf16x6')
  e = [numpy.linalg.norm(frho - rho) / numpy.linalg.norm(rho), numpy.linalg.norm(fd - d) / numpy.linalg.norm(d)]
  print('long lists bf16x6 vs f64', e)
  assert max(e) <= 1e-6, e


def test_large_full_grid_laplacian(ok):
  """The large workload at its full size with the Laplacian: 2,950 AOs, 400
  MOs, the whole 128^3 grid (2.1e6 points), rho + grad rho + the three
  diagonal second derivatives (seven AO sets): the many-MO bf16x6 path
  against the FP64 path (which equals the oracle on samples above), and the
  electron count.  Measured: rho 1.5e-8, grad 1.8e-8, Laplacian 3.9e-8."""
  core, grid = ok['core'], ok['grid']
  qc, axes, info = ok['W'].make('large')
  grid.set_grid(*axes, is_vector=False)
  print('case large full grid laplacian', flush=True)
  rho, d, lap = core.rho_compute(qc, laplacian=True)
  frho, fd, flap = core.rho_compute(qc, laplacian=True, precision='bf16x6')
  assert rho.shape == (128, 128, 128) and d.shape == (3, 128, 128, 128)
  errs = [numpy.linalg.norm(a - b) / numpy.linalg.norm(b) for a, b in ((frho, rho), (fd, d), (flap, lap))]
  print('large full grid bf16x6 vs f64 (rho, xx/yy/zz, laplacian)', errs)
  assert max(errs) <= 1e-7, errs
  n64 = core.electron_count(rho, grid.d3r)
  assert abs(core.electron_count(frho, grid.d3r) - n64) <= 1e-6 * n64


def test_electron_count(ok):
  '''Compensated FP64 device sum vs the exactly rounded math.fsum, on data
  where plain summation loses digits (a few huge values among many small).'''
  import math
  rng = numpy.random.default_rng(5)
  v = rng.uniform(0.0, 1.0, (1 << 22) + 17) * 10.0 ** rng.integers(-12, 1, (1 << 22) + 17)
  v[rng.choice(v.size, 5, replace=False)] = 1e15
  d3r = 0.037
  got = ok['core'].electron_count(v, d3r)
  exact = math.fsum(v) * d3r
  print('electron count', got, exact, abs(got - exact) / exact, flush=True)
  assert abs(got - exact) <= 4e-16 * exact
  assert ok['core'].electron_count(numpy.zeros(0), 1.0) == 0.0


def test_fused_live_rows(ok):
  '''okmi_fused_live_rows (the screened-work measure of bench.py): within
  [0, padded rows x points], zero on a grid far from every atom, additive
  over bricks (a grid split at a brick boundary), on the medium grid between
  a third and all of the dense count for the FP64 radii, and no more with
  the smaller radii of the paths whose AO values end in FP32.'''
  import ctypes
  from orbkit_amd import _native
  from orbkit_amd.basis import device_basis
  qc, axes, info = ok['W'].make('medium', npts_side=24)
  b = device_basis(qc.geo_spec, qc.ao_spec)
  L = _native.lib()

  def rows(x, y, z, prec='f64'):
    ax = [numpy.ascontiguousarray(a, dtype=numpy.float64) for a in (x, y, z)]
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    out = ctypes.c_double()
    _native.check(L.okmi_fused_live_rows(b.handle, P(ax[0]), len(ax[0]), P(ax[1]), len(ax[1]), P(ax[2]), len(ax[2]),
                                         _native.PREC[prec], ctypes.byref(out)), 'okmi_fused_live_rows')
    return out.value
  x, y, z = axes
  npts = len(x) * len(y) * len(z)
  full = rows(x, y, z)
  assert 0.33 * b.nao * npts < full <= 416 * npts
  assert rows(x + 1e4, y, z) == 0.0
  assert rows(x[:12], y, z) + rows(x[12:], y, z) == full
  split = rows(x, y, z, 'bf16x6')
  assert 0 < split <= full and split == rows(x, y, z, 'f32')
  print('live rows x points: FP64 radii %.4g, FP32 radii %.4g' % (full, split))


def test_okmi_rho_staged_subrange(ok):
  '''okmi_rho_staged on points [p0, p0 + np) of a regular grid (the screened
  path's point-order tiles, not bricks), 130 MOs (beyond the fused kernel),
  against the oracle on the same points.'''
  import ctypes
  import torch
  from orbkit_amd import _native
  from orbkit_amd.basis import device_basis
  from orbkit_amd.core import DeviceGrid
  O = ok['O']
  qc, axes, _ = ok['W'].make('medium', npts_side=11, nmo=130)
  basis = device_basis(qc.geo_spec, qc.ao_spec)
  C, occ = qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ()
  dg = DeviceGrid(*axes, regular=True)
  x, y, z = O.expand(*axes)
  p0, np_ = 100, 1000
  codes = numpy.array([1, 2, 3], dtype=numpy.intc)
  L, dev = _native.lib(), dg.x.device
  dC, docc = torch.as_tensor(C, device=dev), torch.as_tensor(occ, device=dev)
  nb = L.okmi_workspace_bytes(basis.handle, 4, 256) + C.shape[0] * 512 * 8
  ws = torch.empty(nb // 8 + 1, dtype=torch.float64, device=dev)
  rho = torch.empty(np_, dtype=torch.float64, device=dev)
  delta = torch.empty((3, np_), dtype=torch.float64, device=dev)
  print('case staged subrange', p0, np_, flush=True)
  _native.check(L.okmi_rho_staged(basis.handle, ctypes.byref(dg.desc), p0, np_, ctypes.c_void_p(dC.data_ptr()),
                                  ctypes.c_void_p(docc.data_ptr()), C.shape[0], codes.ctypes.data_as(_native.P_int),
                                  3, ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(delta.data_ptr()), np_, None,
                                  ctypes.c_void_p(ws.data_ptr()), nb, ctypes.c_void_p(torch.cuda.current_stream()
                                                                                      .cuda_stream)),
                'okmi_rho_staged')
  s = slice(p0, p0 + np_)
  r = O.slice_rho(qc.geo_spec, qc.ao_spec, qc.mo_spec, x[s], y[s], z[s], ['x', 'y', 'z'], False, False,
                  ok['tools'].get_cart2sph)
  assert rel_err(rho.cpu().numpy(), r[0]) <= TOL
  assert rel_err(delta.cpu().numpy(), r[2]) <= TOL
