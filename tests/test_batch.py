'''Many molecules in one launch (okmi_batch_*) and the launch-shape tuner.

The batch runs the single-molecule fused kernel's tile body over the tiles of
every molecule, so each molecule's rho and gradient must equal the
single-molecule fused path bit for bit (same arithmetic, same order), and
the FP64 path within the precision's bar.  Graph replays are checked against
the eager launch, with eager work and synchronisation between replays, and
after the MO coefficients change in place.
'''
import numpy
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mols():
  import torch
  if not torch.cuda.is_available():
    pytest.skip('no HIP device')
  from orbkit_amd import workloads
  ms = workloads.make_batch(n_mol=10, seed=7, side_range=(12, 24))
  # one molecule on a vector grid whose length is no multiple of 64
  qc, axes = ms[3]
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  idx = numpy.random.default_rng(0).choice(X.size, size=1001, replace=False)
  ms[3] = (qc, (X.ravel()[idx].copy(), Y.ravel()[idx].copy(), Z.ravel()[idx].copy()), True)
  return ms


def _single(item, precision):
  from orbkit_amd.basis import device_basis
  from orbkit_amd.core import DeviceGrid
  from orbkit_amd.fused import FusedPlan
  qc, axes = item[0], item[1]
  vec = len(item) > 2 and item[2]
  b = device_basis(qc.geo_spec, qc.ao_spec)
  g = DeviceGrid(*axes, regular=not vec)
  p = FusedPlan(b, g, qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ(), [1, 2, 3], precision=precision)
  p.launch()
  r, d = p.result()
  return r, d, g.shape


@pytest.mark.parametrize('precision', ['bf16x6', 'f32'])
def test_batch_matches_single(mols, precision):
  from orbkit_amd.batch import BatchPlan
  print('case batch', precision, len(mols), flush=True)
  plan = BatchPlan(mols, precision=precision)
  plan.launch()
  for m, item in enumerate(mols):
    r, d, shape = _single(item, precision)
    rb, db = plan.result(m)
    assert numpy.array_equal(rb.reshape(-1), r), m
    assert numpy.array_equal(db.reshape(3, -1), d), m


def test_batch_vs_f64_and_order(mols):
  from orbkit_amd.basis import device_basis
  from orbkit_amd.batch import BatchPlan
  from orbkit_amd.core import DeviceGrid, eval_rho
  print('case batch f64', flush=True)
  plan = BatchPlan(mols, precision='bf16x6', order_by_cost=False)
  plan.launch()
  for m, item in enumerate(mols):
    qc, axes = item[0], item[1]
    vec = len(item) > 2 and item[2]
    r, d, _ = eval_rho(device_basis(qc.geo_spec, qc.ao_spec), DeviceGrid(*axes, regular=not vec),
                       qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ(), [1, 2, 3])
    rb, db = plan.result(m)
    assert numpy.linalg.norm(rb.reshape(-1) - r) <= 1e-6 * numpy.linalg.norm(r), m
    assert numpy.linalg.norm(db.reshape(3, -1) - d) <= 1e-6 * numpy.linalg.norm(d), m


def test_batch_density_only_and_workgroups(mols):
  from orbkit_amd.batch import BatchPlan
  print('case batch rho-only, persistent grids', flush=True)
  ref = BatchPlan(mols, precision='bf16x6')
  ref.launch()
  plan = BatchPlan(mols, precision='bf16x6', grad=False)
  for w in (0, 7, 256, 1024):
    plan.set_workgroups(w)
    plan.launch()
    for m in range(len(mols)):
      assert numpy.array_equal(plan.result(m)[0], ref.result(m)[0]), (w, m)


def test_batch_graph_replays(mols):
  import torch
  from orbkit_amd.batch import BatchPlan
  print('case batch graph', flush=True)
  plan = BatchPlan(mols, precision='bf16x6')
  plan.set_workgroups(512)
  plan.launch()
  torch.cuda.synchronize()
  eager = [plan.result(m) for m in range(len(mols))]
  s = torch.cuda.Stream()
  for rep in range(3):
    for t in plan.rho:
      t.fill_(-1.0)
    for t in plan.dgrad:
      t.fill_(-1.0)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
      plan.launch(use_graph=True, stream=s.cuda_stream)
    s.synchronize()
    junk = torch.randn(1 << 20, device='cuda').sum().item()   # eager work between replays
    assert junk == junk
    for m in range(len(mols)):
      r, d = plan.result(m)
      assert numpy.array_equal(r, eager[m][0]), (rep, m)
      assert numpy.array_equal(d, eager[m][1]), (rep, m)
  # C and occupations are re-read on every replay: scaling C by 1.5 scales rho
  # by 2.25 (up to the bf16 split of the scaled values)
  for c in plan.C:
    c.mul_(1.5)
  with torch.cuda.stream(s):
    plan.launch(use_graph=True, stream=s.cuda_stream)
  s.synchronize()
  for m in range(len(mols)):
    r, _ = plan.result(m)
    assert numpy.linalg.norm(r - 2.25 * eager[m][0]) <= 1e-6 * numpy.linalg.norm(eager[m][0]), m


def test_tuner(mols, tmp_path, monkeypatch):
  from orbkit_amd import autotune
  from orbkit_amd.batch import BatchPlan
  print('case tuner', flush=True)
  monkeypatch.setenv('ORBKIT_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
  plan = BatchPlan(mols, precision='bf16x6')
  r1 = autotune.tune_batch(plan, reps=2)
  assert not r1['cached'] and tuple(r1['shape']) in autotune.SHAPES
  assert plan.workgroups == r1['workgroups'] and plan.tile_shape == tuple(r1['shape'])
  plan2 = BatchPlan(mols, precision='bf16x6')
  r2 = autotune.tune_batch(plan2, reps=2)
  assert r2['cached'] and r2['workgroups'] == r1['workgroups'] and plan2.tile_shape == plan.tile_shape
  # results do not depend on the launch shape beyond screened (< 1e-40) terms
  ref = BatchPlan(mols, precision='bf16x6')
  ref.launch()
  plan2.launch()
  for m in range(len(mols)):
    _close(ref.result(m), plan2.result(m), m)


def _close(a, b, what):
  for x, y in zip(a, b):
    assert numpy.linalg.norm(x - y) <= 1e-12 * numpy.linalg.norm(x), what


@pytest.mark.parametrize('shape', [(1, 2, 3), (0, 3, 3), (0, 0, 6), (3, 3, 0)])
def test_batch_tile_shapes(mols, shape):
  '''Other bricks screen shells per tile differently; only terms below the
  screening bound (1e-40) may change.'''
  from orbkit_amd.batch import BatchPlan
  print('case batch tile shape', shape, flush=True)
  ref = BatchPlan(mols, precision='bf16x6')
  ref.launch()
  plan = BatchPlan(mols, precision='bf16x6')
  plan.set_tile_shape(*shape)
  plan.launch()
  for m in range(len(mols)):
    _close(ref.result(m), plan.result(m), (shape, m))


def _g_molecule(nmo=20, side=14, seed=11):
  '''Four atoms with Cartesian [s, p, d, f, g] shells (l = 4: the fused
  kernels' general instances) and nmo orthonormal MOs on a side^3 grid.'''
  from orbkit_amd import workloads as W
  from orbkit_amd.qcinfo import QCinfo
  rng = numpy.random.default_rng(seed)
  geo = W._atoms(rng, 4, 5.0)
  ao = W._basis(rng, geo, [['s', 'p', 'd', 'f', 'g']] * 4, 0.1, 1e3, False)
  qc = QCinfo()
  qc.geo_spec = geo
  qc.geo_info = numpy.array([['X', str(i + 1), '0'] for i in range(len(geo))])
  qc.ao_spec = ao
  qc.mo_spec = W._mos(rng, W.ao_overlap(geo, ao), nmo)
  return qc, W._grid(geo, side)


def test_g_shells_and_many_tiles(mols):
  '''Fused instances other than the one-buffer l <= 3 ones: a basis with g
  shells (fused path and a batch holding it next to l <= 3 molecules) and
  64 MOs (4 MO tiles) on the medium basis; each against the FP64 path.'''
  import torch
  from orbkit_amd import core, grid, workloads
  from orbkit_amd.batch import BatchPlan
  qc, axes = _g_molecule()
  print('case g shells', flush=True)
  grid.set_grid(*axes, is_vector=False)
  r64, d64 = core.rho_compute(qc, drv=['x', 'y', 'z'])
  for prec in ('bf16x6', 'f32'):
    r, d = core.rho_compute(qc, drv=['x', 'y', 'z'], precision=prec)
    assert numpy.linalg.norm(r - r64) / numpy.linalg.norm(r64) <= 1e-5, prec
    assert numpy.linalg.norm(d - d64) / numpy.linalg.norm(d64) <= 1e-5, prec
  items = [mols[0], (qc, axes), mols[1]]
  plan = BatchPlan(items, precision='bf16x6')
  plan.launch()
  torch.cuda.synchronize()
  for m, item in enumerate(items):
    rs, ds, _ = _single(item, 'bf16x6')
    rb, db = plan.result(m)
    assert numpy.array_equal(rb.reshape(-1), rs), m
    assert numpy.array_equal(db.reshape(3, -1), ds), m
  print('case 64 MOs', flush=True)
  qc, axes, _ = workloads.make('medium', npts_side=16, nmo=64)
  grid.set_grid(*axes, is_vector=False)
  r64, d64 = core.rho_compute(qc, drv=['x', 'y', 'z'])
  r, d = core.rho_compute(qc, drv=['x', 'y', 'z'], precision='bf16x6')
  assert numpy.linalg.norm(r - r64) / numpy.linalg.norm(r64) <= 1e-5
  assert numpy.linalg.norm(d - d64) / numpy.linalg.norm(d64) <= 1e-5


# --------------------------------------------------------------------------
# The batched workload at its real size: 256 molecules of 5-60 atoms
# ([s,s,p,d] spherical, 8-30 MOs), regular grids with sides 32-80 (5.5e7
# points), one persistent launch replayed from a HIP graph.  Each molecule is
# checked against the oracle (orbkit's algorithm, rho_compute once per
# molecule: orbkit/core.py:314-605) on 1,000 points of its own grid.

C4_SAMPLE = 1000


@pytest.fixture(scope='module')
def c4():
  import torch
  if not torch.cuda.is_available():
    pytest.skip('no HIP device')
  from orbkit_amd import workloads
  from oracle import core_oracle as O
  mols = workloads.make_batch(n_mol=256)
  samples = []
  for m, (qc, axes) in enumerate(mols):
    n = len(axes[0]) * len(axes[1]) * len(axes[2])
    idx = numpy.sort(numpy.random.default_rng(100 + m).choice(n, size=C4_SAMPLE, replace=False))
    ix, rest = numpy.divmod(idx, len(axes[1]) * len(axes[2]))   # x slowest, z fastest
    iy, iz = numpy.divmod(rest, len(axes[2]))
    x, y, z = axes[0][ix], axes[1][iy], axes[2][iz]
    r, d, _ = O.rho_slices_c(qc, x, y, z, [1, 2, 3], slice_length=64, nthreads=16)
    samples.append((idx, (x, y, z), r, d))
  return mols, samples


def _errs(rho, grad, r, d):
  mask = r > 1e-10 * r.max()
  return (float(numpy.linalg.norm(rho - r) / numpy.linalg.norm(r)),
          float(numpy.linalg.norm(grad - d) / numpy.linalg.norm(d)),
          float(numpy.max(numpy.abs(rho - r)[mask] / r[mask])))


# bars just above the measured worst molecule (bf16x6: rho 1.9e-7, grad
# 3.0e-7, pointwise 6.5e-7; f32: 2.4e-7, 3.3e-7, 9.3e-7): FP32 rounding of
# the AO values and coefficients and FP32 accumulation over 50-600 AOs, with
# fewer MOs (8-30) to average over than the 40 of the medium workload
@pytest.mark.parametrize('precision,bar,bar_max', [('bf16x6', 4e-7, 1.5e-6), ('f32', 5e-7, 1.5e-6)])
def test_c4_batch_vs_oracle(c4, precision, bar, bar_max):
  """Every molecule's rho and grad rho from the batched launch (persistent
  workgroups, 4x4x4 bricks, HIP graph) against the oracle on its sample:
  norm-wise relative error <= bar for rho and for grad rho, and the
  pointwise relative rho error (rho > 1e-10 max rho) <= bar_max."""
  import torch
  from orbkit_amd.batch import BatchPlan
  mols, samples = c4
  print('case C4 batch', precision, len(mols), flush=True)
  plan = BatchPlan(mols, precision=precision)
  plan.set_workgroups(2048)
  s = torch.cuda.Stream()
  plan.launch(use_graph=True, stream=s.cuda_stream)
  plan.launch(use_graph=True, stream=s.cuda_stream)
  s.synchronize()
  worst = [0.0, 0.0, 0.0]
  for m in range(len(mols)):
    idx, _, r, d = samples[m]
    rho, grad = plan.result(m)
    e = _errs(rho.reshape(-1)[idx], grad.reshape(3, -1)[:, idx], r, d)
    worst = [max(a, b) for a, b in zip(worst, e)]
    assert e[0] <= bar and e[1] <= bar and e[2] <= bar_max, (m, e)
  print(precision, 'worst rho L2 %.2e, grad L2 %.2e, pointwise rho %.2e' % tuple(worst))
  del plan
  torch.cuda.empty_cache()


def test_c4_f64_subset_vs_oracle(c4):
  """The FP64 path (okmi_rho) for every 16th molecule on its sample points
  against the oracle at the FP64 bar (max |a - b| / max |b| <= 1e-12)."""
  from conftest import rel_err
  from orbkit_amd.basis import device_basis
  from orbkit_amd.core import DeviceGrid, eval_rho
  mols, samples = c4
  for m in range(0, len(mols), 16):
    qc = mols[m][0]
    idx, (x, y, z), r, d = samples[m]
    print('case C4 f64 molecule', m, flush=True)
    rho, grad, _ = eval_rho(device_basis(qc.geo_spec, qc.ao_spec), DeviceGrid(x, y, z, regular=False),
                            qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ(), [1, 2, 3])
    assert rel_err(rho, r) <= 1e-12 and rel_err(grad, d) <= 1e-12, (m, rel_err(rho, r), rel_err(grad, d))
