'''Result files written slice by slice (orbkit_amd.output.sink): the NumPy
archive format on the CPU, and rho_compute / calc_mo_matrix with
save_hdf5 on the GPU against their returned arrays (orbkit's layout:
orbkit/core.py:478-501, :874-924).'''
import os

import numpy
import pytest

from orbkit_amd.output import sink as S


def test_npz_stream_roundtrip(tmp_path):
  rng = numpy.random.default_rng(1)
  a1 = rng.normal(size=1003)
  a3 = rng.normal(size=(2, 3, 1003))
  path = str(tmp_path / 'r.npz')
  f = S.open_sink(path, 100)
  S.write_grid(f, numpy.arange(3.0), numpy.arange(4.0), numpy.arange(5.0), False, True)
  for name, a in (('rho', a1), ('delta_rho', a3)):
    f.stream(name, a.shape, ((p0, a[..., p0:p0 + 100]) for p0 in range(0, a.shape[-1], 100)))
    f.set_attr(name, 'shape', numpy.array([17, 59]))
  f.close()
  got = S.read_npz(path)
  assert numpy.array_equal(got['rho'], a1) and numpy.array_equal(got['delta_rho'], a3)
  assert got['delta_rho'].shape == a3.shape
  assert got['rho_attrs/shape'].tolist() == [17, 59]
  assert got['grid/is_vector'] == False and got['grid/is_regular'] == True   # noqa: E712
  assert numpy.array_equal(got['grid/z'], numpy.arange(5.0))


def test_npz_stream_rejects_gaps(tmp_path):
  f = S.open_sink(str(tmp_path / 'g.npz'), 10)
  a = numpy.zeros((2, 30))
  with pytest.raises(ValueError):
    f.stream('x', a.shape, [(0, a[:, :10]), (20, a[:, 20:])])
  f.close()


def test_hdf5_needs_h5py(tmp_path):
  try:
    import h5py  # noqa: F401
  except ImportError:
    with pytest.raises(ImportError):
      S.open_sink(str(tmp_path / 'r.h5'), 10)
  else:
    f = S.open_sink(str(tmp_path / 'r.h5'), 10)
    f.stream('rho', (25,), [(0, numpy.ones(10)), (10, numpy.ones(15))])
    f.close()


@pytest.mark.gpu
def test_rho_compute_result_files(tmp_path):
  import torch
  if not torch.cuda.is_available():
    pytest.skip('no HIP device')
  from orbkit_amd import core, grid, workloads
  qc, axes, _ = workloads.make('small', npts_side=13)
  grid.set_grid(*axes, is_vector=False)
  N = tuple(len(a) for a in axes)
  path = str(tmp_path / 'rho.npz')
  print('case result file rho', flush=True)
  rho, d = core.rho_compute(qc, drv=['x', 'zz'], slice_length=500, save_hdf5=path)
  rho0, d0 = core.rho_compute(qc, drv=['x', 'zz'])
  assert numpy.array_equal(rho, rho0) and numpy.array_equal(d, d0)
  f = S.read_npz(path)
  assert numpy.array_equal(f['rho'], rho.ravel()) and numpy.array_equal(f['delta_rho'], d.reshape(2, -1))
  assert tuple(f['rho_attrs/shape']) == N and tuple(f['delta_rho_attrs/shape']) == (2,) + N
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  assert numpy.array_equal(f['grid/x'], X.ravel()) and numpy.array_equal(f['grid/z'], Z.ravel())
  assert f['grid/is_vector'] == False and f['grid/is_regular'] == False   # noqa: E712  (orbkit's values)
  print('case result file mo', flush=True)
  path = str(tmp_path / 'mo.npz')
  mo = core.rho_compute(qc, calc_mo=True, slice_length=333, save_hdf5=path)
  f = S.read_npz(path)
  assert numpy.array_equal(f['mo_list'], mo.reshape(mo.shape[0], -1))
  assert tuple(f['mo_list_attrs/shape']) == mo.shape
  ao = core.rho_compute(qc, calc_ao=True, drv=['y'], save_hdf5=str(tmp_path / 'ao.npz'))
  assert numpy.array_equal(S.read_npz(str(tmp_path / 'ao.npz'))['ao_list'], ao.reshape(1, ao.shape[1], -1))
  print('case result file mo_matrix', flush=True)
  path = str(tmp_path / 'mm.npz')
  mm = core.calc_mo_matrix(qc, drv=['x'], slice_length=400, save_hdf5=path)
  assert numpy.array_equal(mm, core.calc_mo_matrix(qc, drv=['x']))
  f = S.read_npz(path)
  assert numpy.array_equal(f['mo_matrix'], mm.reshape(mm.shape[:3] + (-1,)))
  bra = S.read_npz(str(tmp_path / 'mm.bra.npz'))['mo_list']
  assert bra.shape == (2, mo.shape[0], mo[0].size)
  assert numpy.array_equal(bra[0], mo.reshape(mo.shape[0], -1))
  try:
    import h5py  # noqa: F401
  except ImportError:
    with pytest.raises(ImportError):
      core.rho_compute(qc, save_hdf5=str(tmp_path / 'rho.h5'))
