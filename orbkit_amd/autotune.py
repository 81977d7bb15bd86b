'''Launch-shape tuning by measurement.

``tune_batch(plan)`` times a ``BatchPlan`` over candidate tile shapes
(bricks of 4x4x4, 2x4x8, 1x8x8 and 1x1x64 regular-grid points) and
persistent workgroup counts (0 = one workgroup per 64-point tile, and
multiples of the device's resident capacity: compute units x 2 workgroups)
with HIP events on the launch stream, keeps the fastest, and caches the
choice in a JSON file keyed by device, precision, derivative pass, MO tiles
and the point count (rounded to a power of two), so later plans of the same
shape skip the search.  The cache file is ``$ORBKIT_AMD_TUNE_CACHE`` or
``~/.cache/orbkit_amd/tune.json``.
'''
import json
import os


def _cache_path():
  return os.environ.get('ORBKIT_AMD_TUNE_CACHE',
                        os.path.join(os.path.expanduser('~'), '.cache', 'orbkit_amd', 'tune.json'))


def _load():
  try:
    with open(_cache_path()) as f:
      return json.load(f)
  except (OSError, ValueError):
    return {}


def _store(d):
  p = _cache_path()
  try:
    os.makedirs(os.path.dirname(p), exist_ok=True)
    tmp = p + '.tmp%d' % os.getpid()
    with open(tmp, 'w') as f:
      json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, p)
  except OSError:
    pass


def batch_key(plan):
  import torch
  name = torch.cuda.get_device_properties(torch.cuda.current_device()).name
  mt = max((c.shape[0] + 15) // 16 for c in plan.C)
  bucket = 1 << max(0, int(plan.npts - 1).bit_length())
  return 'batch|%s|%s|grad=%d|mt=%d|points<=%d' % (name, plan.precision, int(plan.grad), mt, bucket)


SHAPES = [(2, 2, 2), (1, 2, 3), (0, 3, 3), (0, 0, 6)]


def candidates(plan):
  import torch
  cu = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
  out = [0]
  for k in (1, 2, 4, 8, 16):
    w = k * 2 * cu
    if w < plan.ntile:
      out.append(w)
  return out


def time_plan(plan, reps=3, use_graph=False, stream=None):
  '''Median milliseconds of `reps` launches (after one warm-up).'''
  import torch
  s = stream or torch.cuda.current_stream()
  with torch.cuda.stream(s):
    plan.launch(use_graph=use_graph, stream=s.cuda_stream)
    ts = []
    for _ in range(reps):
      a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
      a.record(s)
      plan.launch(use_graph=use_graph, stream=s.cuda_stream)
      b.record(s)
      b.synchronize()
      ts.append(a.elapsed_time(b))
  ts.sort()
  return ts[len(ts) // 2]


def tune_batch(plan, reps=3, use_cache=True, stream=None):
  '''Sets the fastest tile shape and workgroup count on `plan`; returns
  {'shape': [lx, ly, lz], 'workgroups': w, 'ms': t, 'timings': {...}, 'cached': bool}.'''
  key = batch_key(plan)
  cache = _load() if use_cache else {}
  if key in cache:
    rec = cache[key]
    plan.set_tile_shape(*rec['shape'])
    plan.set_workgroups(int(rec['workgroups']))
    return dict(rec, cached=True)
  timings = {}
  for shape in SHAPES:
    plan.set_tile_shape(*shape)
    for w in candidates(plan):
      plan.set_workgroups(w)
      timings[(shape, w)] = time_plan(plan, reps=reps, stream=stream)
  (shape, w) = min(timings, key=timings.get)
  plan.set_tile_shape(*shape)
  plan.set_workgroups(w)
  rec = {'shape': list(shape), 'workgroups': w, 'ms': timings[(shape, w)],
         'timings': {'%d%d%d/%d' % (k[0] + (k[1],)): v for k, v in timings.items()}}
  if use_cache:
    cache = _load()
    cache[key] = rec
    _store(cache)
  return dict(rec, cached=False)
