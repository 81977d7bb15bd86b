'''Gaussian cube files of grid data, in the layout orbkit writes
(orbkit/output/cube.py:5-101), so that files from either program compare
byte for byte.

The format, as written here:

  line 1   title ('orbkit calculation')
  line 2   one space and the caller's comment
  line 3   number of atoms (%5d; negative when a line of data-set labels
           follows the atoms), grid origin x y z (%12.6f each) and, for more
           than one data set, the number of sets (%12d)
  4 .. 6   per axis: number of points (%5d) and the step vector (%12.6f x 3,
           the step on the axis's own component, zeros elsewhere)
  atoms    per atom: the rounded third geo_info column (%5d), the second
           geo_info column as a float (%12.6f), the position (%12.6f x 3)
  labels   only with labels: the number of sets and one integer label per
           set, %5d each, nine fields to a line (the count opens the first),
           then a line break (also after a full line)
  values   x slowest, then y, then z, all sets of a point together, each
           value %13.5E, a line break after every sixth value of an (x, y)
           run and at the end of the run

The value text comes from the library's native formatter
(okmi_cube_format); names ending in '.gz' are gzip-compressed.
'''
import ctypes
import gzip

import numpy

from .. import _native, grid

_EXTENSIONS = ('cube', 'cb', 'cube.gz', 'cb.gz')


def _label_block(labels):
  fields = ['%5d' % len(labels)] + ['%5d' % int(v) for v in labels]
  return ''.join(f + ('\n' if i % 9 == 8 else '') for i, f in enumerate(fields)) + '\n'


def header(ndata, geo_info, geo_spec, comments='', labels=None):
  '''Every line before the values (see the module docstring).'''
  natoms = len(geo_info) if labels is None else -len(geo_info)
  first = '%5d' % natoms + ''.join('%12.6f' % grid.min_[a] for a in range(3))
  if ndata > 1:
    first += '%12d' % ndata
  lines = ['orbkit calculation', ' %s' % comments, first]
  for a in range(3):
    step = [grid.delta_[a] if b == a else 0.0 for b in range(3)]
    lines.append('%5d' % grid.N_[a] + ''.join('%12.6f' % v for v in step))
  for info, pos in zip(geo_info, geo_spec):
    lines.append('%5d%12.6f' % (round(float(info[2])), float(info[1])) + ''.join('%12.6f' % v for v in pos[:3]))
  text = '\n'.join(lines) + '\n'
  return text + (_label_block(labels) if labels is not None else '')


def body(data):
  '''Value text of data (NDATA, Nx, Ny, Nz), x slowest.'''
  L = _native.lib()
  d = numpy.ascontiguousarray(data, dtype=numpy.float64)
  nd, nx, ny, nz = d.shape
  ptr = d.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
  cap = L.okmi_cube_format(ptr, nd, nx, ny, nz, None, 0)
  buf = ctypes.create_string_buffer(cap)
  n = L.okmi_cube_format(ptr, nd, nx, ny, nz, buf, cap)
  if n == 0 and d.size:
    raise RuntimeError('okmi_cube_format: buffer too small')
  return buf.raw[:n].decode('ascii')


def _checked(data, labels):
  '''(data as (NDATA,) + N, integer labels or None); AssertionError on bad
  input, as orbkit's writer raises.'''
  data = numpy.array(data)
  if data.ndim not in (3, 4):
    raise AssertionError('cube data must be one grid (3 axes) or a stack of grids (4 axes), not %d axes'
                         % data.ndim)
  if data.ndim == 3:
    data = data[numpy.newaxis]
  if data.shape[1:] != tuple(grid.N_):
    raise AssertionError('data grid %s differs from the grid module\'s %s' % (data.shape[1:], tuple(grid.N_)))
  if labels is None:
    return data, None
  if labels is True or labels == 'auto':
    labels = range(len(data))
  labels = list(labels)
  if len(labels) != len(data):
    raise AssertionError('%d labels for %d data sets' % (len(labels), len(data)))
  try:
    return data, [int(v) for v in labels]
  except (TypeError, ValueError):
    raise AssertionError('cube labels must be integers (or convertible to them)')


def cube_creator(data, filename, geo_info, geo_spec, comments='', labels=None, **kwargs):
  '''Writes grid data (N or (NDATA,) + N on the grid module's regular grid)
  as a Gaussian cube file and returns the file name ('.cube' appended
  unless the name already ends in a cube extension).'''
  data, labels = _checked(data, labels)
  if not filename.endswith(_EXTENSIONS):
    filename += '.cube'
  text = header(len(data), geo_info, geo_spec, comments, labels) + body(data)
  if filename.endswith('gz'):
    with gzip.open(filename, 'wb') as f:
      f.write(text.encode('ascii'))
  else:
    with open(filename, 'w') as f:
      f.write(text)
  return filename
