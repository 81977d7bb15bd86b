'''Gaussian cube files as orbkit writes them (orbkit/output/cube.py:5-101).

``cube_creator(data, filename, geo_info, geo_spec, comments='', labels=None)``
takes data of shape N or (NDATA,) + N on the grid module's regular grid and
writes the same text as orbkit: the header (title line, comment line, atom
count, grid origin, axis vectors, atoms), optional data-set labels and the
values, x slowest, '%13.5E', six per line plus a break after each z run.
The value text is produced by the library's native formatter
(okmi_cube_format); '.gz' file names are gzip-compressed.
'''
import ctypes
import gzip

import numpy

from .. import _native, grid


def header(ndata, geo_info, geo_spec, comments='', labels=None):
  '''Header lines (orbkit/output/cube.py:45-82).'''
  string = 'orbkit calculation\n'
  string += ' %(f)s\n' % {'f': comments}
  string += ('%(at)d' % {'at': (-1) ** (labels is not None) * len(geo_info)}).rjust(5)
  for ii in range(3):
    string += ('%(min)0.6f' % {'min': grid.min_[ii]}).rjust(12)
  if ndata > 1:
    string += ('%d' % ndata).rjust(12)
  for ii in range(3):
    string += '\n'
    string += ('%(N)d' % {'N': grid.N_[ii]}).rjust(5)
    for jj in range(3):
      if jj == ii:
        string += ('%(dr)0.6f' % {'dr': grid.delta_[ii]}).rjust(12)
      else:
        string += ('%(dr)0.6f' % {'dr': 0}).rjust(12)
  string += '\n'
  for ii in range(len(geo_info)):
    string += ('%(N)d' % {'N': round(float(geo_info[ii][2]))}).rjust(5)
    string += ('%(ch)0.6f' % {'ch': float(geo_info[ii][1])}).rjust(12)
    for jj in range(3):
      string += ('%(r)0.6f' % {'r': geo_spec[ii][jj]}).rjust(12)
    string += '\n'
  if labels is not None:
    string += ('%(N)d' % {'N': ndata}).rjust(5)
    c = 0
    for i, j in enumerate(labels):
      c += 1
      string += str(j).rjust(5)
      if c % 9 == 8:
        string += '\n'
    string += '\n'
  return string


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


def cube_creator(data, filename, geo_info, geo_spec, comments='', labels=None, **kwargs):
  '''Writes a plain text Gaussian cube file (see module docstring).'''
  data = numpy.array(data)
  dims = 3
  if data.ndim < dims:
    raise AssertionError('data.ndim < ndim of grid')
  elif data.ndim == dims:
    data = data[numpy.newaxis]
  elif data.ndim > dims + 1:
    raise AssertionError('data.ndim > (ndim of grid) +2')
  if labels is not None:
    if labels is True or labels == 'auto':
      labels = list(range(len(data)))
    assert len(labels) == len(data)
    try:
      labels = [int(v) for v in labels]
    except ValueError:
      raise AssertionError('labels has to be list of integers.')
  assert data.shape[1:] == tuple(grid.N_), 'The grid does not fit the data.'
  if not any(filename.endswith(ext) for ext in ['cube', 'cb', 'cube.gz', 'cb.gz']):
    filename += '.cube'
  string = header(len(data), geo_info, geo_spec, comments, labels) + body(data)
  if filename.endswith('gz'):
    with gzip.open(filename, 'wb') as f:
      f.write(string.encode('utf-8'))
  else:
    with open(filename, 'w') as f:
      f.write(string)
  return filename
