'''Throughput of the grid-density hot path on one MI355X.

Step = one evaluation of rho and grad(rho) over the whole 100^3 regular grid
of the medium workload (20 atoms, spherical [s,s,p,p,d,f] shells, 400 AOs,
3 primitives per shell, exponents in [0.05, 1e5], 40 doubly occupied MOs),
inputs resident in HBM, outputs written to HBM (FP64).  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision bf16x6|bf16x3|f32|f64]
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


def parse():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=20)
  ap.add_argument('--warmup', type=int, default=3)
  ap.add_argument('--precision', default='bf16x6', choices=['bf16x6', 'bf16x3', 'f32', 'f64'])
  ap.add_argument('--workload', default='medium')
  ap.add_argument('--side', type=int, default=None, help='grid points per axis (default: workload)')
  ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
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


def cpu_baseline(qc, axes, threads, slice_length, slices_per_thread=4):
  '''orbkit's algorithm (C restatement in oracle/, OpenMP over slices) on a
  bounded sample: slices_per_thread slices of slice_length points per
  thread, a contiguous run of the flattened grid from its middle.'''
  from oracle import core_oracle
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  P = numpy.stack([X.ravel(), Y.ravel(), Z.ravel()])
  n = int(threads * slice_length * slices_per_thread)
  a = max(0, P.shape[1] // 2 - n // 2)
  x, y, z = P[0, a:a + n].copy(), P[1, a:a + n].copy(), P[2, a:a + n].copy()
  core_oracle.rho_slices_c(qc, x[:64], y[:64], z[:64], [1, 2, 3], slice_length=64, nthreads=1)
  t = time.perf_counter()
  core_oracle.rho_slices_c(qc, x, y, z, [1, 2, 3], slice_length=slice_length, nthreads=threads)
  dt = time.perf_counter() - t
  return {'value': n / dt, 'unit': 'grid points/s', 'cores': threads, 'kind': 'port',
          'sample': '%d of the %d grid points (%d slices of %d, rho + grad rho, FP64) in %.1f s'
                    % (n, P.shape[1], n // slice_length, slice_length, dt)}


def main():
  args = parse()
  if args.gpus != 1:
    print('bench.py runs on one GPU: this build has no multi-GPU path (--gpus %d requested)' % args.gpus,
          file=sys.stderr)
    sys.exit(2)
  import torch
  from orbkit_amd import _native, workloads
  from orbkit_amd.basis import device_basis
  from orbkit_amd.core import DeviceGrid
  from orbkit_amd.fused import FusedPlan

  torch.cuda.set_device(0)
  qc, axes, info = workloads.make(args.workload, npts_side=args.side)
  basis = device_basis(qc.geo_spec, qc.ao_spec)
  dg = DeviceGrid(*axes, regular=True)
  C, occ = qc.mo_spec.get_coeffs(), qc.mo_spec.get_occ()
  codes = numpy.array([1, 2, 3], dtype=numpy.intc)
  L = _native.lib()
  npts = dg.npts
  nao, nmo, nsets = info['nao'], info['nmo'], 4

  if args.precision == 'f64':
    dev = dg.x.device
    dC = torch.as_tensor(C, device=dev)
    docc = torch.as_tensor(occ, device=dev)
    rho = torch.empty(npts, dtype=torch.float64, device=dev)
    delta = torch.empty((3, npts), dtype=torch.float64, device=dev)
    ws = torch.empty(L.okmi_workspace_bytes(basis.handle, nsets, 1 << 17) // 8 + 1, dtype=torch.float64,
                     device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def step():
      _native.check(L.okmi_rho(basis.handle, ctypes.byref(dg.desc), 0, npts, ctypes.c_void_p(dC.data_ptr()),
                               ctypes.c_void_p(docc.data_ptr()), nmo, codes.ctypes.data_as(_native.P_int), 3,
                               ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(delta.data_ptr()), npts, None,
                               ctypes.c_void_p(ws.data_ptr()), ws.numel() * 8, st), 'okmi_rho')
    kernels = ['ao_eval', 'rho_f64']
    result = lambda: (rho.cpu().numpy(), delta.cpu().numpy())
  else:
    plan = FusedPlan(basis, dg, C, occ, codes, precision=args.precision)
    step = plan.launch
    kernels = ['rho_fused']
    result = plan.result

  if args.calibrate_copy:
    src = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device=dg.x.device)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
  for _ in range(args.warmup):
    step()
  torch.cuda.synchronize()
  L.okmi_profile_enable(1)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    step()
  torch.cuda.synchronize()
  dt = time.perf_counter() - t0
  times = {}
  for k in kernels:
    cnt, ms = ctypes.c_int(), ctypes.c_double()
    _native.check(L.okmi_profile_query(k.encode(), ctypes.byref(cnt), ctypes.byref(ms)), 'profile_query')
    times[k] = (cnt.value, ms.value)
  L.okmi_profile_enable(0)

  value = npts * args.steps / dt
  gemm_flops = 2.0 * nmo * nao * nsets * npts   # algorithmic, per launch (one pass over the grid)
  if args.precision == 'f64':
    # dominant kernel of the FP64 path: the MFMA contraction with density epilogue
    cnt, ms = times['rho_f64']
    t_launch = ms / 1e3 / args.steps
    achieved = gemm_flops / t_launch / 1e12
    roof = {'bound': 'mfma', 'achieved': achieved, 'peak': PEAK_TFLOPS['f64'], 'unit': 'TFLOP/s',
            'frac': achieved / PEAK_TFLOPS['f64'],
            'traffic': measured_traffic('k_rho_f64', args.workload, args.precision), 'kernel': 'k_rho_f64',
            'launch_ms': t_launch * 1e3, 'other_kernels_ms': {'k_ao_eval': times['ao_eval'][1] / args.steps}}
    dtype = 'f64'
  else:
    cnt, ms = times['rho_fused']
    t_launch = ms / 1e3 / args.steps
    achieved = gemm_flops / t_launch / 1e12
    nprod = {'bf16x6': 6, 'bf16x3': 3, 'f32': 1}[args.precision]
    unit_dtype = 'f32' if args.precision == 'f32' else 'bf16'
    peak = PEAK_TFLOPS[unit_dtype] / nprod
    roof = {'bound': 'mfma', 'achieved': achieved, 'peak': peak, 'unit': 'TFLOP/s', 'frac': achieved / peak,
            'traffic': measured_traffic('k_rho_fused', args.workload, args.precision), 'kernel': 'k_rho_fused',
            'launch_ms': t_launch * 1e3,
            'executed_tflops': achieved * nprod, 'mfma_dtype': unit_dtype, 'products': nprod}
    dtype = {'bf16x6': 'bf16x6 (fp32 acc, fp64 epilogue)', 'bf16x3': 'bf16x3 (fp32 acc, fp64 epilogue)',
             'f32': 'f32'}[args.precision]

  out = {'metric': 'grid points/s for rho_compute (+grad rho)', 'value': value, 'unit': 'grid points/s',
         'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': dt / args.steps * 1e3,
         'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': dtype,
         'data': 'synthetic (seeded random molecule, basis and orthonormal MOs)',
         'config': {'workload': '%s: %d atoms, %d spherical AOs (l<=3, 3 prims/shell), %d MOs, %d^3 regular '
                                'grid, rho + grad rho' % (args.workload, info['natoms'], nao, nmo, info['side']),
                    'npts': npts, 'nao': nao, 'nmo': nmo, 'ao_sets': nsets, 'precision': args.precision},
         'roofline': roof}
  if args.errors or args.precision != 'f64':
    r, d = result()
    from orbkit_amd.core import eval_rho
    rr, dd, _ = eval_rho(basis, dg, C, occ, codes)
    mask = rr > 1e-10 * rr.max()
    out['error_vs_f64'] = {
        'rho_max_rel': float(numpy.max(numpy.abs(r - rr)[mask] / rr[mask])),
        'rho_l2_rel': float(numpy.linalg.norm(r - rr) / numpy.linalg.norm(rr)),
        'grad_l2_rel': float(numpy.linalg.norm(d - dd) / numpy.linalg.norm(dd)),
        'grad_max_rel': float(numpy.max(numpy.abs(d - dd)) / numpy.max(numpy.abs(dd))),
        'electrons': float(numpy.sum(r) * numpy.prod([a[1] - a[0] for a in axes])),
        'electrons_f64': float(numpy.sum(rr) * numpy.prod([a[1] - a[0] for a in axes]))}
  if not args.no_cpu:
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    out['cpu_baseline'] = cpu_baseline(qc, axes, threads, int(info['slice_length']))
  print(json.dumps(out))


if __name__ == '__main__':
  main()
