'''Gradients of a density loss L = sum_r w_r (rho_r - t_r)^2 with respect to
MO coefficients, primitive exponents and nuclear positions.

CPU tests pin the FP64 numpy gradient oracle (oracle/grad_oracle.py) against
central finite differences of the oracle's forward and against two exact
identities.  GPU tests compare the native backward (okmi_rho_backward via
orbkit_amd.autograd) with that oracle (relative 1e-10), with central finite
differences of the GPU forward (relative 1e-6) and with the identities on a
full-size grid.
'''
import numpy
import pytest

from conftest import rel_err


def _points_near_atoms(geo, n, seed, spread=1.5):
  rng = numpy.random.default_rng(seed)
  a = rng.integers(0, len(geo), n)
  P = geo[a] + rng.normal(0.0, spread, (n, 3))
  return P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy()


def _loss_parts(rho0, seed):
  rng = numpy.random.default_rng(seed)
  w = rng.uniform(0.5, 1.5, rho0.shape)
  t = rho0 * rng.uniform(0.8, 1.0, rho0.shape)
  return w, t


def _case(name, prenormalized=False, npts=600, seed=0):
  from orbkit_amd import workloads
  qc, _, _ = workloads.make(name)
  if prenormalized:
    for c in qc.ao_spec.data:
      c['pnum'] = -abs(c['pnum'])
    qc.ao_spec.update()
  x, y, z = _points_near_atoms(numpy.asarray(qc.geo_spec), npts, seed)
  return qc, x, y, z


def _fd(fun, theta, idx, h):
  out = []
  for i in idx:
    tp = theta.copy()
    tm = theta.copy()
    tp.flat[i] += h(theta.flat[i])
    tm.flat[i] -= h(theta.flat[i])
    out.append((fun(tp) - fun(tm)) / (tp.flat[i] - tm.flat[i]))
  return numpy.array(out)


CASES = [('small', False), ('small', True), ('medium', False)]


@pytest.mark.parametrize('name,prenorm', CASES)
def test_grad_oracle_fd(name, prenorm):
  from oracle import grad_oracle as go
  qc, x, y, z = _case(name, prenorm, npts=400)
  C, occ = qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ()
  alpha = qc.ao_spec.get_prim_coeffs()[:, 0].copy()
  geo = numpy.array(qc.geo_spec, dtype=float)
  rho0 = go.density(qc, x, y, z)
  w, t = _loss_parts(rho0, 1)
  loss = lambda r: float(numpy.sum(w * (r - t) ** 2))
  g = 2.0 * w * (rho0 - t)
  rho, dC, da, dg = go.vjp(qc, x, y, z, g)
  assert rel_err(rho, rho0) < 1e-14
  rng = numpy.random.default_rng(3)
  ia = rng.choice(len(alpha), size=min(24, len(alpha)), replace=False)
  fa = _fd(lambda a: loss(go.density(qc, x, y, z, alpha=a)), alpha, ia, lambda v: 1e-5 * v)
  assert numpy.max(numpy.abs(fa - da[ia])) <= 1e-6 * numpy.max(numpy.abs(da)), (fa, da[ia])
  ig = rng.choice(geo.size, size=min(12, geo.size), replace=False)
  fg = _fd(lambda R: loss(go.density(qc, x, y, z, geo=R)), geo, ig, lambda v: 1e-5)
  assert numpy.max(numpy.abs(fg - dg.flat[ig])) <= 1e-6 * numpy.max(numpy.abs(dg)), (fg, dg.flat[ig])
  ic = rng.choice(C.size, size=12, replace=False)
  fc = _fd(lambda Cm: loss(go.density(qc, x, y, z, C=Cm)), C, ic, lambda v: 1e-6)
  assert numpy.max(numpy.abs(fc - dC.flat[ic])) <= 1e-6 * numpy.max(numpy.abs(dC)), (fc, dC.flat[ic])


@pytest.mark.parametrize('name,prenorm', CASES)
def test_grad_oracle_identities(name, prenorm):
  '''rho is homogeneous of degree 2 in C: sum C dL/dC = 2 sum g rho.
  Moving every nucleus by d moves rho by d: sum_a dL/dR_a = -sum g grad rho.'''
  from oracle import core_oracle, grad_oracle as go
  qc, x, y, z = _case(name, prenorm, npts=500, seed=5)
  C, occ = qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ()
  g = numpy.random.default_rng(2).normal(size=len(x))
  rho, dC, da, dg = go.vjp(qc, x, y, z, g)
  assert abs(numpy.sum(C * dC) - 2 * numpy.sum(g * rho)) <= 1e-12 * numpy.sum(numpy.abs(g * rho))
  _, dgrad = core_oracle.rho_compute(qc, x, y, z, True, drv=['x', 'y', 'z'])
  lhs = dg.sum(axis=0)
  rhs = -(dgrad * g[None, :]).sum(axis=1)
  assert numpy.max(numpy.abs(lhs - rhs)) <= 1e-12 * numpy.sum(numpy.abs(dgrad * g[None, :]))


# ---------------------------------------------------------------- GPU tests

def _gpu():
  import torch
  if not torch.cuda.is_available():
    pytest.skip('no HIP device')
  return torch


@pytest.mark.gpu
@pytest.mark.parametrize('name,prenorm', CASES)
def test_backward_vs_oracle(name, prenorm):
  torch = _gpu()
  from orbkit_amd.autograd import DensityModel
  from oracle import grad_oracle as go
  qc, x, y, z = _case(name, prenorm, npts=3001, seed=7)
  print('case', name, 'prenormalized', prenorm, 'npts', len(x), flush=True)
  model = DensityModel(qc, (x, y, z), is_vector=True)
  C, alpha, geo = model.parameters()
  rho = model(C, alpha, geo)
  rho0 = rho.detach().cpu().numpy()
  w, t = _loss_parts(rho0, 1)
  wt, tt = torch.as_tensor(w, device=rho.device), torch.as_tensor(t, device=rho.device)
  loss = (wt * (rho - tt) ** 2).sum()
  loss.backward()
  g = 2.0 * w * (rho0 - t)
  r_o, dC_o, da_o, dg_o = go.vjp(qc, x, y, z, g)
  assert rel_err(rho0, r_o) < 1e-12
  assert rel_err(C.grad.cpu().numpy(), dC_o) < 1e-10
  assert rel_err(alpha.grad.cpu().numpy(), da_o) < 1e-10
  assert rel_err(geo.grad.cpu().numpy(), dg_o) < 1e-10


@pytest.mark.gpu
def test_backward_fd_regular_grid():
  '''Spherical basis with f shells on a regular grid; every exponent and
  position component against central differences of the GPU forward.'''
  torch = _gpu()
  from orbkit_amd import workloads
  from orbkit_amd.autograd import DensityModel
  qc, axes, _ = workloads.make('gradient', npts_side=24)
  print('case gradient 24^3', flush=True)
  model = DensityModel(qc, axes)
  C, alpha, geo = model.parameters()
  rho = model(C, alpha, geo)
  rho0 = rho.detach()
  w = torch.rand(rho0.shape, dtype=torch.float64, device=rho0.device, generator=torch.Generator(
      device=rho0.device).manual_seed(0)) + 0.5
  t = 0.9 * rho0
  loss_of = lambda r: float(((w * (r - t)) * (r - t)).sum())
  loss = (w * (rho - t) ** 2).sum()
  loss.backward()
  ga, gg, gc = alpha.grad.cpu().numpy(), geo.grad.cpu().numpy(), C.grad.cpu().numpy()
  a0 = alpha.detach().cpu().numpy()
  R0 = geo.detach().cpu().numpy()
  C0 = C.detach()
  f = lambda a=a0, R=R0, Cm=C0: loss_of(model.rho(Cm, a, R))
  fa = _fd(lambda a: f(a=a), a0, range(len(a0)), lambda v: 1e-5 * v)
  assert numpy.max(numpy.abs(fa - ga)) <= 1e-6 * numpy.max(numpy.abs(ga))
  fg = _fd(lambda R: f(R=R), R0, range(R0.size), lambda v: 1e-5)
  assert numpy.max(numpy.abs(fg - gg.reshape(-1))) <= 1e-6 * numpy.max(numpy.abs(gg))
  rng = numpy.random.default_rng(0)
  ic = rng.choice(gc.size, size=40, replace=False)
  Cn = C0.cpu().numpy()
  fc = _fd(lambda Cm: f(Cm=torch.as_tensor(Cm, device=C0.device)), Cn, ic, lambda v: 1e-6)
  assert numpy.max(numpy.abs(fc - gc.reshape(-1)[ic])) <= 1e-6 * numpy.max(numpy.abs(gc))


@pytest.mark.gpu
def test_backward_identities_and_chunking():
  '''Full 80^3 grid of the gradient workload: homogeneity and translation
  identities against the forward's own grad rho, and the same gradients
  from a workspace small enough to force many chunks.'''
  torch = _gpu()
  from orbkit_amd import workloads
  from orbkit_amd.autograd import DensityModel
  from orbkit_amd.basis import device_basis
  from orbkit_amd.core import DeviceGrid, eval_rho
  qc, axes, _ = workloads.make('gradient')
  print('case gradient 80^3', flush=True)
  model = DensityModel(qc, axes)
  C, alpha, geo = model.parameters()
  rho = model(C, alpha, geo)
  g = torch.randn(rho.shape, dtype=torch.float64, device=rho.device,
                  generator=torch.Generator(device=rho.device).manual_seed(1))
  (rho * g).sum().backward()
  gn, rn = g.cpu().numpy(), rho.detach().cpu().numpy()
  Cn = C.detach().cpu().numpy()
  assert abs(numpy.sum(Cn * C.grad.cpu().numpy()) - 2 * numpy.sum(gn * rn)) <= 1e-11 * numpy.sum(numpy.abs(gn * rn))
  b = device_basis(qc.geo_spec, qc.ao_spec)
  _, d, _ = eval_rho(b, DeviceGrid(*axes, regular=True), Cn, qc.mo_spec.get_occ(), [1, 2, 3])
  lhs = geo.grad.cpu().numpy().sum(axis=0)
  rhs = -(d * gn[None, :]).sum(axis=1)
  assert numpy.max(numpy.abs(lhs - rhs)) <= 1e-11 * numpy.sum(numpy.abs(d * gn[None, :]))
  small = DensityModel(qc, axes, workspace_bytes=24 << 20)
  dC2, da2, dg2 = small.vjp(C, alpha, geo, g)
  assert rel_err(dC2.cpu().numpy(), C.grad.cpu().numpy()) < 1e-12
  assert rel_err(da2.cpu().numpy(), alpha.grad.cpu().numpy()) < 1e-12
  assert rel_err(dg2.cpu().numpy(), geo.grad.cpu().numpy()) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize('precision,tol', [('f64', 1e-12), ('f32', 2e-6)])
def test_backward_screened_equals_dense(precision, tol):
  """A whole regular grid takes the screened backward (points in 4x4x4
  bricks, AO rows in 32-row chunks, only the chunks in reach of a brick
  evaluated and contracted); the same points given as a vector grid take
  the dense one.  The gradients agree to rounding (the dropped terms are
  below 1e-24).  Side 30: bricks straddling the grid's end."""
  torch = _gpu()
  from orbkit_amd import workloads
  from orbkit_amd.autograd import DensityModel
  qc, axes, _ = workloads.make('gradient', npts_side=30)
  print('case screened vs dense backward', precision, flush=True)
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  vec = tuple(numpy.ascontiguousarray(v.reshape(-1)) for v in (X, Y, Z))
  reg = DensityModel(qc, axes, precision=precision)
  den = DensityModel(qc, vec, is_vector=True, precision=precision)
  C, alpha, geo = reg.parameters(requires_grad=False)
  g = torch.randn(reg.npts, dtype=torch.float64, device=C.device,
                  generator=torch.Generator(device=C.device).manual_seed(3)).to(C.dtype)
  out_r = reg.vjp(C, alpha, geo, g)
  out_d = den.vjp(C, alpha, geo, g)
  for a, b in zip(out_r, out_d):
    a, b = a.double().cpu().numpy(), b.double().cpu().numpy()
    assert numpy.max(numpy.abs(a - b)) <= tol * numpy.max(numpy.abs(b))


@pytest.mark.gpu
@pytest.mark.parametrize('precision,tol,nmo', [('f64', 1e-12, 40), ('f64', 1e-12, 70), ('f32', 2e-6, 70)])
def test_backward_long_live_lists(precision, tol, nmo):
  """The screened backward with live-chunk lists past 128 entries (a
  6,000-AO basis whose first brick reaches about 170 chunks of 32 rows: the
  long-list instances of the C products; 40 MOs take the short-K product
  for G, 70 the streaming one) equals the dense backward on the same points
  given as a vector grid."""
  torch = _gpu()
  from conftest import long_list_case
  from orbkit_amd.autograd import DensityModel
  qc, axes = long_list_case(nmo)
  print('case long-list backward', precision, nmo, flush=True)
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  vec = tuple(numpy.ascontiguousarray(v.reshape(-1)) for v in (X, Y, Z))
  reg = DensityModel(qc, axes, precision=precision)
  den = DensityModel(qc, vec, is_vector=True, precision=precision)
  C, alpha, geo = reg.parameters(requires_grad=False)
  g = torch.randn(reg.npts, dtype=torch.float64, device=C.device,
                  generator=torch.Generator(device=C.device).manual_seed(5)).to(C.dtype)
  out_r = reg.vjp(C, alpha, geo, g)
  out_d = den.vjp(C, alpha, geo, g)
  for a, b in zip(out_r, out_d):
    a, b = a.double().cpu().numpy(), b.double().cpu().numpy()
    assert numpy.max(numpy.abs(a - b)) <= tol * numpy.max(numpy.abs(b))


# ------------------------------------------------------------ FP32 backward
#
# The FP32 model rounds its parameters (C, exponents, positions) to FP32; the
# FP64 checks below evaluate the oracle at those same rounded values, so the
# differences measure the FP32 arithmetic alone.  Errors are norm-wise
# (max |a - b| / max |b|).

def _f32_params(torch, model):
  C, alpha, geo = model.parameters()
  assert C.dtype == alpha.dtype == geo.dtype == torch.float32
  return C, alpha, geo, tuple(t.detach().double().cpu().numpy() for t in (C, alpha, geo))


@pytest.mark.gpu
@pytest.mark.parametrize('name,prenorm', CASES)
def test_backward_f32_vs_oracle(name, prenorm):
  '''FP32 forward and FP32 backward (okmi_rho_backward_f32) against the FP64
  oracle at the same FP32-rounded parameters and the same upstream gradient.'''
  torch = _gpu()
  from orbkit_amd.autograd import DensityModel
  from oracle import grad_oracle as go
  qc, x, y, z = _case(name, prenorm, npts=3001, seed=7)
  print('case', name, 'prenormalized', prenorm, 'npts', len(x), 'f32', flush=True)
  model = DensityModel(qc, (x, y, z), is_vector=True, precision='f32')
  C, alpha, geo, (Cd, ad, gd) = _f32_params(torch, model)
  rho = model(C, alpha, geo)
  assert rho.dtype == torch.float32
  r_o = go.density(qc, x, y, z, C=Cd, alpha=ad, geo=gd)
  w, t = _loss_parts(r_o, 1)
  g = 2.0 * w * (r_o - t)
  _, dC_o, da_o, dg_o = go.vjp(qc, x, y, z, g, C=Cd, alpha=ad, geo=gd)
  dC, da, dg = model.vjp(C, alpha, geo, torch.as_tensor(g, device=C.device))
  assert dC.dtype == da.dtype == dg.dtype == torch.float32
  errs = [rel_err(rho.detach().cpu().numpy(), r_o), rel_err(dC.cpu().numpy(), dC_o), rel_err(da.cpu().numpy(), da_o),
          rel_err(dg.cpu().numpy(), dg_o)]
  print('f32 errors rho %.2e dC %.2e dalpha %.2e dgeo %.2e' % tuple(errs), flush=True)
  # measured: rho <= 2.4e-7, gradients <= 4.6e-7
  assert errs[0] < 1e-6
  assert max(errs[1:]) < 2e-6
  # through torch.autograd: the FP32 loss of the FP32 forward
  wt = torch.as_tensor(w, device=rho.device, dtype=torch.float32)
  tt = torch.as_tensor(t, device=rho.device, dtype=torch.float32)
  (wt * (rho - tt) ** 2).sum().backward()
  assert C.grad.dtype == alpha.grad.dtype == geo.grad.dtype == torch.float32
  assert rel_err(C.grad.cpu().numpy(), dC_o) < 1e-4
  assert rel_err(alpha.grad.cpu().numpy(), da_o) < 1e-4
  assert rel_err(geo.grad.cpu().numpy(), dg_o) < 1e-4


@pytest.mark.gpu
def test_backward_f32_full_grid():
  '''Full 80^3 grid of the gradient workload: FP32 gradients against the FP64
  backward for the same upstream gradient and FP32-rounded parameters, the
  homogeneity identity in FP32, and the same gradients from a workspace
  small enough to force many chunks.'''
  torch = _gpu()
  from orbkit_amd import workloads
  from orbkit_amd.autograd import DensityModel
  qc, axes, _ = workloads.make('gradient')
  print('case gradient 80^3 f32', flush=True)
  m32 = DensityModel(qc, axes, precision='f32')
  m64 = DensityModel(qc, axes)
  C, alpha, geo, _ = _f32_params(torch, m32)
  C64, a64, g64 = (t.detach().double() for t in (C, alpha, geo))
  rho64 = m64.rho(C64, a64, g64)
  g = torch.randn(rho64.shape, dtype=torch.float64, device=rho64.device,
                  generator=torch.Generator(device=rho64.device).manual_seed(1))
  ref = m64.vjp(C64, a64, g64, g)
  got = m32.vjp(C, alpha, geo, g.float())
  errs = [rel_err(a.cpu().numpy(), b.cpu().numpy()) for a, b in zip(got, ref)]
  print('f32 vs f64 full grid: dC %.2e dalpha %.2e dgeo %.2e' % tuple(errs), flush=True)
  assert max(errs) < 2e-6   # measured 4.0e-7
  rho32 = m32.rho(C, alpha, geo)
  assert rel_err(rho32.cpu().numpy(), rho64.cpu().numpy()) < 1e-6
  lhs = float((C64 * got[0].double()).sum())
  rhs = 2.0 * float((g * rho64).sum())
  assert abs(lhs - rhs) <= 1e-5 * float((g * rho64).abs().sum())
  small = DensityModel(qc, axes, precision='f32', workspace_bytes=24 << 20)
  again = small.vjp(C, alpha, geo, g.float())
  for a, b in zip(again, got):
    assert rel_err(a.cpu().numpy(), b.cpu().numpy()) < 1e-6
