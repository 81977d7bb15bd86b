'''okmi_rho_noalloc (workspace in a per-GPU device array, no allocation)
against okmi_rho with a caller workspace, at the FP64 bar (max |a - b| /
max |b| <= 1e-12), on two calls in a row, the second with a smaller input;
both paths (fused up to 64 MOs, staged beyond), a grid that needs several
passes of the scratch, and a request that exceeds the scratch.'''
import ctypes

import numpy
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope='module')
def env():
  import torch
  if not torch.cuda.is_available():
    pytest.skip('no HIP device')
  from orbkit_amd import _native, workloads
  from orbkit_amd.basis import device_basis
  from orbkit_amd.core import DeviceGrid
  return dict(torch=torch, N=_native, W=workloads, device_basis=device_basis, DeviceGrid=DeviceGrid)


def _call(env, name, basis, dg, C, occ, codes, ws_bytes=None):
  '''rho, derivative rows and MO norms (accumulated onto 0.25) from one entry point.'''
  torch, N = env['torch'], env['N']
  L, dev = N.lib(), dg.x.device
  codes = numpy.ascontiguousarray(codes, dtype=numpy.intc)
  dC, docc = torch.as_tensor(C, device=dev), torch.as_tensor(occ, device=dev)
  rho = torch.empty(dg.npts, dtype=torch.float64, device=dev)
  delta = torch.empty((max(len(codes), 1), dg.npts), dtype=torch.float64, device=dev)
  nrm = torch.full((C.shape[0],), 0.25, dtype=torch.float64, device=dev)
  st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
  args = [basis.handle, ctypes.byref(dg.desc), 0, dg.npts, ctypes.c_void_p(dC.data_ptr()),
          ctypes.c_void_p(docc.data_ptr()), C.shape[0], codes.ctypes.data_as(N.P_int), len(codes),
          ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(delta.data_ptr()), dg.npts, ctypes.c_void_p(nrm.data_ptr())]
  if name == 'okmi_rho':
    ws = torch.empty(ws_bytes // 8 + 1, dtype=torch.float64, device=dev)
    N.check(L.okmi_rho(*args, ctypes.c_void_p(ws.data_ptr()), ws_bytes, st), name)
  else:
    N.check(L.okmi_rho_noalloc(*args, st), name)
  torch.cuda.synchronize()
  return rho.cpu().numpy(), delta[:len(codes)].cpu().numpy(), nrm.cpu().numpy() - 0.25


def _compare(env, basis, dg, C, occ, codes, what):
  L = env['N'].lib()
  ws_bytes = max(L.okmi_rho_fused_workspace_bytes(basis.handle, C.shape[0], 0, dg.npts),
                 L.okmi_workspace_bytes(basis.handle, 10, 4096))
  print('case', what, 'npts', dg.npts, 'nmo', C.shape[0], flush=True)
  ref = _call(env, 'okmi_rho', basis, dg, C, occ, codes, ws_bytes)
  got = _call(env, 'okmi_rho_noalloc', basis, dg, C, occ, codes)
  errs = [rel_err(g, r) for g, r in zip(got, ref)]
  print(what, errs)
  assert max(errs) <= TOL, (what, errs)
  assert numpy.all(numpy.isfinite(got[0])) and got[0].max() > 0


@pytest.mark.parametrize('nmo', [40, 80])
def test_noalloc_matches_okmi_rho_twice(env, nmo):
  '''Call 1: a 80^3 regular grid (512,000 points: the fused path needs two
  passes of the 64 MiB scratch) with rho, gradient and Laplacian terms;
  call 2: a smaller vector grid with other derivative codes, so that what
  call 1 left in the scratch cannot pass for call 2's results.  80 MOs run
  the staged path (AO sets of 2,900-point passes in the scratch).'''
  W, DeviceGrid = env['W'], env['DeviceGrid']
  qc, axes, _ = W.make('medium', npts_side=80 if nmo <= 64 else 40, nmo=nmo)
  basis = env['device_basis'](qc.geo_spec, qc.ao_spec)
  C, occ = qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ()
  _compare(env, basis, DeviceGrid(*axes, regular=True), C, occ, [1, 2, 3, 4, 5, 6], 'regular')
  rng = numpy.random.default_rng(12)
  pts = [rng.uniform(a[0], a[-1], 1234) for a in axes]
  _compare(env, basis, DeviceGrid(*pts, regular=False), C, occ, [9, 1], 'vector')


def test_noalloc_rejects_beyond_scratch(env):
  '''A request whose fixed buffers plus one 64-point pass exceed the scratch
  is refused with an error naming the limit, before any kernel runs.'''
  torch, N = env['torch'], env['N']
  from orbkit_amd.orbitals import AOClass
  rng = numpy.random.default_rng(3)
  geo = rng.uniform(-3, 3, (4, 3))
  ao = AOClass()
  for i in range(20000):   # 20,000 s shells: 7 staged AO sets x 64 points = 72 MB
    ao.append({'atom': i % 4, 'type': 's', 'pnum': 1, 'coeffs': numpy.array([[1.0 + (i % 7), 1.0]])})
  ao.update()
  basis = env['device_basis'](geo, ao)
  nmo = 65
  C = rng.normal(size=(nmo, basis.nao))
  occ = numpy.full(nmo, 2.0)
  dg = env['DeviceGrid'](*[numpy.linspace(-1, 1, 4)] * 3, regular=True)
  L = N.lib()
  assert L.okmi_rho_noalloc_scratch_bytes() == 64 << 20
  with pytest.raises(RuntimeError, match='scratch'):
    _call(env, 'okmi_rho_noalloc', basis, dg, C, occ, [1, 2, 3, 4, 5, 6])
