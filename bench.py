'''Throughput of the grid-density hot path on one MI355X.

Default (the headline line): step = one evaluation of rho and grad(rho) over
the whole 100^3 regular grid of the medium workload (20 atoms, spherical
[s,s,p,p,d,f] shells, 400 AOs, 3 primitives per shell, exponents in
[0.05, 1e5], 40 doubly occupied MOs), inputs resident in HBM, outputs written
to HBM, all arithmetic in FP64 (the fused FP64 kernel).  Beside it, in the
same JSON line under "legs": the staged FP64 path (AO sets through HBM, then
the FP64 GEMM) and the bf16x6 fused path, each with its time, roofline
fraction, error against the FP64 result and speed-up over the headline.
Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision f64|bf16x6|bf16x3|f32]
                    [--workload medium|small|large|gradient|batch] [--laplacian] [--no-legs]

Other workloads (parity cases of the same path, measured the same way):
  large     200 atoms, 2,950 spherical AOs, 400 MOs, 128^3 grid (many-MO path)
  gradient  forward + backward of L = sum d3r (rho - rho_t)^2 on the medium
            basis, 80^3 grid (FP64; dL/dC, dL/dalpha, dL/dR)
  batch     256 molecules of 5-60 atoms, uneven grids 32^3..80^3, one tuned
            persistent launch replayed from a HIP graph
'''
import argparse
import ctypes
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published dense peaks
PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {'bf16': 2500.0, 'f32': 157.3, 'f64': 78.6}
PRODUCTS = {'bf16x6': 6, 'bf16x3': 3, 'f32': 1}
METRIC = 'grid points/s for rho_compute (+grad rho)'


def parse():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=20)
  ap.add_argument('--warmup', type=int, default=3)
  ap.add_argument('--precision', default='f64', choices=['bf16x6', 'bf16x3', 'f32', 'f64'])
  ap.add_argument('--workload', default='medium', choices=['medium', 'small', 'large', 'gradient', 'batch'])
  ap.add_argument('--side', type=int, default=None, help='grid points per axis (default: workload)')
  ap.add_argument('--laplacian', action='store_true', help='also the Laplacian (7 AO sets)')
  ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
  ap.add_argument('--no-legs', action='store_true', help='time only the headline path')
  ap.add_argument('--cpu-threads', type=int, default=None)
  ap.add_argument('--errors', action='store_true', help='also report FP64-relative errors')
  ap.add_argument('--calibrate-copy', type=int, default=0, metavar='MIB',
                  help='before the timed steps, copy MIB MiB device-to-device once (a kernel of known '
                       'traffic for scaling FETCH_SIZE / WRITE_SIZE counters)')
  return ap.parse_args()


def measured_traffic(kernel, workload, precision):
  '''HBM bytes per launch from the committed counter summary (FETCH_SIZE +
  WRITE_SIZE, scaled by a plain copy of known size), or None.'''
  path = os.path.join(ROOT, 'profiles', 'traffic.json')
  if not os.path.exists(path):
    return None
  with open(path) as f:
    t = json.load(f)
  v = t.get('%s/%s/%s' % (kernel, workload, precision))
  return None if v is None else v['bytes_per_launch']


def cpu_info():
  """CPU model, logical CPUs of the machine, CPUs this process may run on,
  and the thread budget of the job (OMP_NUM_THREADS where the environment
  sets one: the CPU share that comes with one GPU)."""
  model = 'unknown'
  try:
    with open('/proc/cpuinfo') as f:
      for line in f:
        if line.startswith('model name'):
          model = line.split(':', 1)[1].strip()
          break
  except OSError:
    pass
  nproc = os.cpu_count() or 1
  avail = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else nproc
  env = os.environ.get('OMP_NUM_THREADS')
  share = min(avail, int(env)) if env and env.isdigit() and int(env) > 0 else avail
  return {'model': model, 'nproc': nproc, 'affinity': avail, 'threads': share}


def cpu_baseline(qc, axes, threads, slice_length, codes, slices_per_thread=2, reps=5):
  """orbkit's algorithm (C restatement in oracle/: orbkit's AO kernel, the
  MO triple loop and slice_rho's assembly, OpenMP over slices standing in
  for orbkit's Pool over slices) on a bounded sample of the grid: a
  contiguous run from the middle of the flattened grid, slices_per_thread
  slices of slice_length points per thread.  One warm-up, then the median of
  `reps` timed runs; the same on one thread (one warm-up, median of 3) on
  a sample of one slice; and the tuned CPU form (oracle/cpu_tuned.c) on the
  first sample with the same threads."""
  from oracle import core_oracle
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  P = numpy.stack([X.ravel(), Y.ravel(), Z.ravel()])

  def sample(n):
    n = int(min(P.shape[1], n))
    a = max(0, P.shape[1] // 2 - n // 2)
    return n, P[0, a:a + n].copy(), P[1, a:a + n].copy(), P[2, a:a + n].copy()

  def run(nthreads, n, reps, fn=core_oracle.rho_slices_c):
    n, x, y, z = sample(n)
    fn(qc, x[:slice_length // 4], y[:slice_length // 4], z[:slice_length // 4], codes,
       slice_length=slice_length, nthreads=nthreads)
    ts = []
    for _ in range(reps):
      t = time.perf_counter()
      fn(qc, x, y, z, codes, slice_length=slice_length, nthreads=nthreads)
      ts.append(time.perf_counter() - t)
    return n, float(numpy.median(ts)), ts

  info = cpu_info()
  n, dt, ts = run(threads, threads * slice_length * slices_per_thread, reps)
  n1, dt1, _ = run(1, slice_length, 3)
  nt, dtt, tst = run(threads, threads * slice_length * slices_per_thread, reps, core_oracle.rho_slices_tuned)
  return {'value': n / dt, 'unit': 'grid points/s', 'cores': threads, 'kind': 'port',
          'sample': '%d of the %d grid points (%d slices of %d, rho + derivative codes %s, FP64); median of %d '
                    'runs after one warm-up: %.2f s (runs %s)'
                    % (n, P.shape[1], -(-n // slice_length), slice_length, [int(c) for c in codes], reps, dt,
                       ', '.join('%.2f' % t for t in ts)),
          'one_core': {'value': n1 / dt1, 'sample': '%d points, median of 3 after one warm-up: %.2f s' % (n1, dt1)},
          'tuned': {'value': nt / dtt, 'cores': threads,
                    'sample': 'the same %d points and threads; orbkit\'s redundant work removed (one exponential '
                              'per primitive and point shared by all components and AO sets, row map folded into '
                              'C, blocked MO contraction: oracle/cpu_tuned.c); median of %d runs after one '
                              'warm-up: %.2f s (runs %s)' % (nt, reps, dtt, ', '.join('%.2f' % t for t in tst))},
          'cpu_model': info['model'], 'nproc': info['nproc'], 'affinity_cpus': info['affinity'],
          'threads_note': 'threads = the CPUs this job may use (affinity, capped by OMP_NUM_THREADS where set)'}


def cpu_baseline_grad(qc, axes, n=3000):
  """FP64 numpy/C gradient oracle (forward + dL/dC, dL/dalpha, dL/dR) on a
  bounded sample of n grid points from the middle of the grid, one thread."""
  from threadpoolctl import threadpool_limits
  from oracle import grad_oracle
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  a = X.size // 2 - n // 2
  x, y, z = (V.ravel()[a:a + n].copy() for V in (X, Y, Z))
  g = numpy.ones(n)
  with threadpool_limits(1):
    t = time.perf_counter()
    grad_oracle.vjp(qc, x, y, z, g)
    dt = time.perf_counter() - t
  return {'value': n / dt, 'unit': 'grid points/s', 'cores': 1, 'kind': 'port',
          'sample': '%d of the %d grid points, forward + backward, FP64, in %.1f s' % (n, X.size, dt)}


def prof_times(L, names):
  from orbkit_amd import _native
  out = {}
  for k in names:
    cnt, ms = ctypes.c_int(), ctypes.c_double()
    _native.check(L.okmi_profile_query(k.encode(), ctypes.byref(cnt), ctypes.byref(ms)), 'profile_query')
    out[k] = (cnt.value, ms.value)
  return out


def timed(step, args, L, names, sync):
  for _ in range(args.warmup):
    step()
  sync()
  L.okmi_profile_enable(1)
  sync()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    step()
  sync()
  dt = time.perf_counter() - t0
  times = prof_times(L, names)
  L.okmi_profile_enable(0)
  return dt, times


# ---------------------------------------------------------------- workloads

def screened_rows(L, basis, axes, precision='f64'):
  """Sum over the grid's 4x4x4 bricks of (AO rows of the chunks the kernels
  contract there) x (points in the brick): okmi_fused_live_rows, with the
  screening radii of the precision ('f64', or one whose AO values end in
  FP32)."""
  from orbkit_amd import _native
  ax = [numpy.ascontiguousarray(a, dtype=numpy.float64) for a in axes]
  P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
  out = ctypes.c_double()
  _native.check(L.okmi_fused_live_rows(basis.handle, P(ax[0]), len(ax[0]), P(ax[1]), len(ax[1]), P(ax[2]), len(ax[2]),
                                       _native.PREC[precision], ctypes.byref(out)), 'okmi_fused_live_rows')
  return out.value


def density_leg(args, torch, info, basis, dg, C, occ, codes, path, axes=None):
  """One way of computing rho (+ grad, + Laplacian) over the whole grid,
  timed: path 'f64' (fused FP64), 'f64_staged' (AO sets through HBM, then
  the FP64 GEMM with density epilogue), or a reduced precision of the fused
  kernel ('bf16x6', 'bf16x3', 'f32'; more than 64 MOs: the many-MO path).
  Returns (measurements, result getter)."""
  from orbkit_amd import _native
  from orbkit_amd.fused import FusedPlan
  L = _native.lib()
  npts, nao, nmo = dg.npts, info['nao'], info['nmo']
  nsets = 7 if args.laplacian else 4
  if path == 'f64_staged' or (path == 'f64' and nmo > 64):
    dev = dg.x.device
    dC, docc = torch.as_tensor(C, device=dev), torch.as_tensor(occ, device=dev)
    rho = torch.empty(npts, dtype=torch.float64, device=dev)
    delta = torch.empty((len(codes), npts), dtype=torch.float64, device=dev)
    ws = torch.empty(L.okmi_workspace_bytes(basis.handle, nsets, 1 << 17) // 8 + 1, dtype=torch.float64, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def step():
      _native.check(L.okmi_rho_staged(basis.handle, ctypes.byref(dg.desc), 0, npts, ctypes.c_void_p(dC.data_ptr()),
                                      ctypes.c_void_p(docc.data_ptr()), nmo, codes.ctypes.data_as(_native.P_int),
                                      len(codes), ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(delta.data_ptr()),
                                      npts, None, ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, st),
                    'okmi_rho_staged')
    names, result = ['ao_eval', 'rho_f64'], (lambda: (rho.cpu().numpy(), delta.cpu().numpy()))
    dom = ('rho_f64', 'k_rho_f64', 'f64', 1)
  else:
    plan = FusedPlan(basis, dg, C, occ, codes, precision=path)
    step, result = plan.launch, plan.result
    if path == 'f64':
      names = ['rho_fused_f64']
      dom = ('rho_fused_f64', 'k_rho_fused', 'f64', 1)
    elif nmo > 64:
      names = ['ao_eval', 'rho_big']
      dom = ('rho_big', 'k_rho_big', 'f32' if path == 'f32' else 'bf16', PRODUCTS[path])
    else:
      names = ['rho_fused']
      dom = ('rho_fused', 'k_rho_fused', 'f32' if path == 'f32' else 'bf16', PRODUCTS[path])
  dt, times = timed(step, args, L, names, torch.cuda.synchronize)
  flops = 2.0 * nmo * nao * nsets * npts   # algorithmic, one pass over the grid
  cnt, ms = times[dom[0]]
  t_launch = ms / 1e3 / args.steps
  achieved = flops / t_launch / 1e12
  peak = PEAK_TFLOPS[dom[2]] / dom[3]
  roof = {'bound': 'mfma', 'achieved': achieved, 'peak': peak, 'unit': 'TFLOP/s', 'frac': achieved / peak,
          'traffic': measured_traffic(dom[1], args.workload, 'f64' if path == 'f64_staged' else path),
          'kernel': dom[1] + ('<f64>' if path == 'f64' else ''),
          'kernel_ms_per_step': t_launch * 1e3, 'launches_per_step': cnt / args.steps,
          'algorithmic_flop_per_step': flops}
  if dom[3] > 1:
    roof.update(executed_tflops=achieved * dom[3], mfma_dtype=dom[2], products=dom[3])
  if axes is not None:
    # the screened contraction (every path: fused, many-MO and staged): per
    # brick only the AO chunks with a shell in reach (the work the kernel
    # does besides MO-tile padding)
    sflops = 2.0 * nmo * nsets * screened_rows(L, basis, axes, 'f64' if path == 'f64_staged' else path)
    roof['screened'] = {'flop_per_step': sflops, 'achieved': sflops / t_launch / 1e12,
                        'frac': sflops / t_launch / 1e12 / peak, 'share_of_algorithmic': sflops / flops,
                        'note': 'algorithmic flops count every AO at every point; the kernel skips the AO chunks '
                                'whose shells do not reach a brick (values below 1e-24), so frac (algorithmic) can '
                                'exceed 1 and screened.frac is the hardware utilisation'}
  roof['other_kernels_ms_per_step'] = {k: v[1] / args.steps for k, v in times.items() if k != dom[0]}
  return {'value': npts * args.steps / dt, 'ms_per_step': dt / args.steps * 1e3, 'roofline': roof}, result


def error_vs(r, d, rr, dd, axes, laplacian):
  """Errors of (r, d) against the FP64 result (rr, dd): rho pointwise max
  over rho > 1e-10 max rho and L2, gradient L2 and max, electron counts."""
  mask = rr > 1e-10 * rr.max()
  d3r = numpy.prod([a[1] - a[0] for a in axes])
  err = {'rho_max_rel': float(numpy.max(numpy.abs(r - rr)[mask] / rr[mask])),
         'rho_l2_rel': float(numpy.linalg.norm(r - rr) / numpy.linalg.norm(rr)),
         'grad_l2_rel': float(numpy.linalg.norm(d[:3] - dd[:3]) / numpy.linalg.norm(dd[:3])),
         'grad_max_rel': float(numpy.max(numpy.abs(d[:3] - dd[:3])) / numpy.max(numpy.abs(dd[:3]))),
         'electrons': float(numpy.sum(r) * d3r), 'electrons_f64': float(numpy.sum(rr) * d3r)}
  if laplacian:
    lap, lap0 = d[3:].sum(axis=0), dd[3:].sum(axis=0)
    err['lap_l2_rel'] = float(numpy.linalg.norm(lap - lap0) / numpy.linalg.norm(lap0))
    err['lap_max_rel'] = float(numpy.max(numpy.abs(lap - lap0)) / numpy.max(numpy.abs(lap0)))
  return err


def density_step(args, torch, info, qc, axes):
  """rho (+ grad, + Laplacian) over one regular grid.  The headline is the
  precision asked for (default FP64, fused); beside it, unless --no-legs,
  the staged FP64 path and the bf16x6 fused path are timed the same way,
  with their errors against the FP64 result and their speed-up over it."""
  from orbkit_amd.basis import device_basis
  from orbkit_amd.core import DeviceGrid
  basis = device_basis(qc.geo_spec, qc.ao_spec)
  dg = DeviceGrid(*axes, regular=True)
  C, occ = qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ()
  codes = numpy.array([1, 2, 3] + ([4, 5, 6] if args.laplacian else []), dtype=numpy.intc)
  nsets = 7 if args.laplacian else 4
  nao, nmo = info['nao'], info['nmo']
  head, head_result = density_leg(args, torch, info, basis, dg, C, occ, codes, args.precision, axes)
  out = {'value': head['value'], 'ms_per_step': head['ms_per_step'],
         'dtype': {'bf16x6': 'bf16x6 (fp32 acc, fp64 epilogue)', 'bf16x3': 'bf16x3 (fp32 acc, fp64 epilogue)',
                   'f32': 'f32 (fp64 epilogue)', 'f64': 'f64'}[args.precision],
         'config': {'workload': '%s: %d atoms, %d spherical AOs (l<=3, 3 prims/shell), %d MOs, %d^3 regular grid, '
                                'rho + grad rho%s' % (args.workload, info['natoms'], nao, nmo, info['side'],
                                                      ' + laplacian' if args.laplacian else ''),
                    'npts': dg.npts, 'nao': nao, 'nmo': nmo, 'ao_sets': nsets,
                    'path': 'fused' if args.precision != 'f64' or nmo <= 64 else 'staged'},
         'roofline': head['roofline']}
  legs = [] if args.no_legs else [p for p in ('f64_staged', 'bf16x6') if p != args.precision]
  ref = None
  if args.precision != 'f64' or legs or args.errors:
    from orbkit_amd.core import eval_rho
    rr, dd, _ = eval_rho(basis, dg, C, occ, codes, staged=True)   # the staged FP64 path
    ref = (rr, dd)
  if ref is not None:
    r, d = head_result()
    out['error_vs_f64'] = error_vs(r, d, rr, dd, axes, args.laplacian)
  out['legs'] = {}
  for p in legs:
    leg, res = density_leg(args, torch, info, basis, dg, C, occ, codes, p, axes)
    r, d = res()
    leg['error_vs_f64'] = error_vs(r, d, rr, dd, axes, args.laplacian)
    leg['speedup_vs_headline'] = head['ms_per_step'] / leg['ms_per_step']
    out['legs'][p] = leg
  return out, codes


def gradient_step(args, torch, info, qc, axes):
  '''Forward rho and backward of L = sum d3r (rho - rho_t)^2, in FP64
  (default) or, with --precision f32, in FP32 (FP32 forward contraction and
  FP32 backward; their gradients are then compared with FP64 ones).'''
  from orbkit_amd import _native
  from orbkit_amd.autograd import DensityModel
  L = _native.lib()
  prec = 'f32' if args.precision == 'f32' else 'f64'
  model = DensityModel(qc, axes, precision=prec)
  C, alpha, geo = model.parameters()
  # target from perturbed coefficients C + 0.01 N(0,1) (seed 33)
  Ct = C.detach() + 0.01 * torch.as_tensor(numpy.random.default_rng(33).normal(size=tuple(C.shape)),
                                           device=C.device, dtype=C.dtype)
  rho_t = model.rho(Ct, alpha, geo)
  d3r = float(numpy.prod([a[1] - a[0] for a in axes]))

  def step():
    for p in (C, alpha, geo):
      p.grad = None
    rho = model(C, alpha, geo)
    loss = (d3r * (rho - rho_t) ** 2).sum()
    loss.backward()
  sfx = '' if prec == 'f64' else '_f32'
  fwd = 'rho_fused_f64' if prec == 'f64' else 'rho_fused'
  names = [fwd, 'ao_eval', 'gemm_nn' + sfx, 'gemm_tn' + sfx, 'shell_grad' + sfx, 'rho_backward' + sfx]
  dt, times = timed(step, args, L, names, torch.cuda.synchronize)
  npts, nao, nmo = model.npts, info['nao'], info['nmo']
  es = 8 if prec == 'f64' else 4
  # the MFMA contractions of the backward: per point, 2 nmo nao flop each
  # for Phi = C A, dL/dC += B A^T and G = C^T B; algorithmic bytes: A and B
  # read by the dL/dC contraction; A read and Phi written, B read and G
  # written by the two C products
  cost = {'gemm_tn' + sfx: ('k_gemm_tn', 2.0 * nmo * nao * npts, es * (nao + nmo) * npts),
          'gemm_nn' + sfx: ('k_gemm_nn', 4.0 * nmo * nao * npts, 2 * es * (nao + nmo) * npts)}
  roofs = {}
  for k, (kern, flops, nbytes) in cost.items():
    t = times[k][1] / 1e3 / args.steps
    t_mfma = flops / (PEAK_TFLOPS[prec] * 1e12)
    t_hbm = nbytes / (PEAK_HBM_GBS * 1e9)
    if t_hbm >= t_mfma:
      r = {'bound': 'hbm', 'achieved': nbytes / t / 1e9, 'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
           'frac': nbytes / t / 1e9 / PEAK_HBM_GBS}
    else:
      r = {'bound': 'mfma', 'achieved': flops / t / 1e12, 'peak': PEAK_TFLOPS[prec], 'unit': 'TFLOP/s',
           'frac': flops / t / 1e12 / PEAK_TFLOPS[prec]}
    r.update({'traffic': measured_traffic(kern, 'gradient', prec), 'kernel': kern + '<%s>' % prec,
              'kernel_ms_per_step': t * 1e3, 'algorithmic_flop_per_step': flops,
              'algorithmic_bytes_per_step': nbytes, 'launches_per_step': times[k][0] / args.steps})
    roofs[k] = r
  dom = max(roofs, key=lambda k: roofs[k]['kernel_ms_per_step'])
  roof = roofs.pop(dom)
  roof['other_kernels_ms_per_step'] = {k: v[1] / args.steps for k, v in times.items() if k != dom}
  roof['other_rooflines'] = roofs
  out = {'value': npts * args.steps / dt, 'ms_per_step': dt / args.steps * 1e3, 'dtype': prec,
         'metric_note': 'forward rho + backward (dL/dC, dL/dalpha, dL/dR) per grid point',
         'config': {'workload': 'gradient: %d atoms, %d spherical AOs, %d MOs, %d^3 regular grid, '
                                'L = sum d3r (rho - rho_t)^2' % (info['natoms'], nao, nmo, info['side']),
                    'npts': npts, 'nao': nao, 'nmo': nmo, 'nprim': info['nprim']},
         'roofline': roof}
  if prec == 'f32':
    # the same loss and parameter values in FP64: norm-wise relative errors
    ref = DensityModel(qc, axes, precision='f64')
    C64, a64, g64 = (t.detach().double().requires_grad_(True) for t in (C, alpha, geo))
    rt64 = ref.rho(Ct.double(), a64, g64)
    (d3r * (ref(C64, a64, g64) - rt64) ** 2).sum().backward()
    err = lambda a, b: float(torch.linalg.norm(a.double() - b) / torch.linalg.norm(b))
    out['error_vs_f64'] = {'dC_l2_rel': err(C.grad, C64.grad), 'dalpha_l2_rel': err(alpha.grad, a64.grad),
                           'dgeo_l2_rel': err(geo.grad, g64.grad)}
  return out, numpy.array([], dtype=numpy.intc)


def batch_step(args, torch):
  from orbkit_amd import _native, autotune, workloads
  from orbkit_amd.batch import BatchPlan
  L = _native.lib()
  prec = 'bf16x6' if args.precision == 'f64' else args.precision
  mols = workloads.make_batch(n_mol=256)
  plan = BatchPlan(mols, precision=prec)
  tune = autotune.tune_batch(plan, reps=2, use_cache=False)
  stream = torch.cuda.Stream()

  def step():
    plan.launch(use_graph=True, stream=stream.cuda_stream)
  dt, times = timed(step, args, L, ['rho_batch'], stream.synchronize)
  flops = sum(2.0 * c.shape[0] * b.nao * 4 * g.npts for c, b, g in zip(plan.C, plan.bases, plan.grids))
  t = times['rho_batch'][1] / 1e3 / args.steps
  peak = PEAK_TFLOPS['f32' if prec == 'f32' else 'bf16'] / PRODUCTS[prec]
  roof = {'bound': 'mfma', 'achieved': flops / t / 1e12, 'peak': peak, 'unit': 'TFLOP/s',
          'frac': flops / t / 1e12 / peak, 'traffic': measured_traffic('k_rho_batch', 'batch', prec),
          'kernel': 'k_rho_batch', 'kernel_ms_per_step': t * 1e3, 'algorithmic_flop_per_step': flops,
          'tuned': {'tile_shape': tune['shape'], 'workgroups': tune['workgroups']}}
  # the screened contraction (AO chunks in reach of each brick), per molecule
  sflops = sum(2.0 * c.shape[0] * 4 * screened_rows(L, b, axes, args.precision)
               for c, b, (_, axes) in zip(plan.C, plan.bases, mols))
  roof['screened'] = {'flop_per_step': sflops, 'achieved': sflops / t / 1e12, 'frac': sflops / t / 1e12 / peak,
                      'share_of_algorithmic': sflops / flops}
  nao = [b.nao for b in plan.bases]
  return {'value': plan.npts * args.steps / dt, 'ms_per_step': dt / args.steps * 1e3,
          'dtype': '%s (fp32 acc, fp64 epilogue)' % prec,
          'config': {'workload': 'batch: 256 molecules, 5-60 atoms x [s,s,p,d] spherical (NAO %d-%d), 8-30 MOs, '
                                 'regular grids 32^3..80^3, rho + grad rho, one persistent launch in a HIP graph'
                                 % (min(nao), max(nao)),
                     'npts': plan.npts, 'nmol': plan.nmol, 'tiles': plan.ntile, 'precision': prec},
          'roofline': roof}, None


def main():
  args = parse()
  if args.gpus != 1:
    print('bench.py runs on one GPU: this build has no multi-GPU path (--gpus %d requested)' % args.gpus,
          file=sys.stderr)
    sys.exit(2)
  import torch
  from orbkit_amd import workloads
  torch.cuda.set_device(0)
  if args.calibrate_copy:
    src = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
  qc = axes = info = None
  if args.workload == 'batch':
    res, codes = batch_step(args, torch)
  else:
    qc, axes, info = workloads.make(args.workload, npts_side=args.side)
    if args.workload == 'gradient':
      res, codes = gradient_step(args, torch, info, qc, axes)
    else:
      res, codes = density_step(args, torch, info, qc, axes)
  out = {'metric': METRIC, 'value': res.pop('value'), 'unit': 'grid points/s', 'n_gpus': 1, 'steps': args.steps,
         'warmup': args.warmup, 'ms_per_step': res.pop('ms_per_step'), 'higher_is_better': True, 'scaling': 'weak',
         'vs_baseline': None, 'dtype': res.pop('dtype'),
         'data': 'synthetic (seeded random molecules, bases and orthonormal MOs)'}
  out.update(res)
  if not args.no_cpu and args.workload == 'gradient':
    out['cpu_baseline'] = cpu_baseline_grad(qc, axes)
  elif not args.no_cpu and qc is not None:
    threads = args.cpu_threads or cpu_info()['threads']
    sl = int(info['slice_length'])
    per = 2 if info['nao'] <= 1000 else 1
    if args.workload == 'large':
      sl = 512
    out['cpu_baseline'] = cpu_baseline(qc, axes, threads, sl, list(codes) if len(codes) else [1, 2, 3],
                                       slices_per_thread=per, reps=5 if info['nao'] <= 1000 else 3)
  print(json.dumps(out))


if __name__ == '__main__':
  main()
