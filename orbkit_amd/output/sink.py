'''Result files written slice by slice from device memory: the ``save_hdf5``
argument of ``rho_compute`` and ``calc_mo_matrix`` (orbkit/core.py:478-501,
:874-924, orbkit/tools.py:303-314).

orbkit opens an HDF5 file (h5py), stores the grid (``grid/x``, ``grid/y``,
``grid/z``, ``grid/is_vector``, ``grid/is_regular``), creates one flat
dataset per result (``rho``, ``delta_rho``, ``mo_list``, ``ao_list``,
``mo_matrix``; last axis = grid points, chunked by ``slice_length``) that each
slice is written into as it is computed, and tags each with its grid shape
(attribute ``shape``).  The same layout is written here from the device
buffers, one slice at a time, so that no second host copy of a result is
made for the file:

  * ``save_hdf5='name.npz'``: a NumPy zip archive (no h5py needed).  Each
    dataset is a member ``<name>.npy`` streamed slice by slice; arrays with
    more than one axis are stored in Fortran order (a slice of grid points
    is then one contiguous run of the file) and load as the same array with
    ``numpy.load``.  Attributes go to ``<name>_attrs/<key>.npy``, as orbkit's
    npz writer names them (orbkit/output/hdf5.py:56-103).
  * any other name: an HDF5 file through h5py, as orbkit writes it; h5py
    missing raises ImportError, as in orbkit.
'''
import os
import zipfile

import numpy


class NpzSink:
  '''NumPy zip archive; datasets are streamed one after another (a zip
  archive takes one member at a time).'''

  def __init__(self, path):
    self.path = path
    self.zf = zipfile.ZipFile(path, 'w', compression=zipfile.ZIP_STORED, allowZip64=True)

  def put(self, name, value):
    with self.zf.open(name + '.npy', 'w', force_zip64=True) as f:
      numpy.lib.format.write_array(f, numpy.asanyarray(value), allow_pickle=False)

  def stream(self, name, shape, blocks):
    '''Dataset `name` of shape (..., npts) from `blocks`: an iterable of
    (first point, array of shape[:-1] + (n,)) in point order.'''
    shape = tuple(int(v) for v in shape)
    header = {'descr': numpy.lib.format.dtype_to_descr(numpy.dtype(numpy.float64)),
              'fortran_order': len(shape) > 1, 'shape': shape}
    done = 0
    with self.zf.open(name + '.npy', 'w', force_zip64=True) as f:
      numpy.lib.format.write_array_header_2_0(f, header)
      for p0, block in blocks:
        if p0 != done or block.shape[:-1] != shape[:-1]:
          raise ValueError('%s: block at point %d (expected %d) of shape %s' % (name, p0, done, block.shape))
        # Fortran order: for each point, the leading axes first-fastest
        f.write(numpy.ascontiguousarray(numpy.asarray(block, dtype=numpy.float64).T).tobytes())
        done += block.shape[-1]
    if done != shape[-1]:
      raise ValueError('%s: %d of %d points written' % (name, done, shape[-1]))

  def set_attr(self, name, key, value):
    self.put('%s_attrs/%s' % (name, key), value)

  def close(self):
    self.zf.close()


class Hdf5Sink:
  '''HDF5 file through h5py, datasets chunked by slice.'''

  def __init__(self, path, slice_length):
    import h5py   # as orbkit: the HDF5 sink needs h5py
    self.f = h5py.File(str(path), 'w')
    self.slice_length = int(slice_length)

  def put(self, name, value):
    self.f[name] = value

  def stream(self, name, shape, blocks):
    shape = tuple(int(v) for v in shape)
    chunks = shape[:-1] + (max(1, min(self.slice_length, shape[-1])),)
    d = self.f.create_dataset(name, shape, dtype=numpy.float64, chunks=chunks if shape[-1] else None)
    for p0, block in blocks:
      d[..., p0:p0 + block.shape[-1]] = block

  def set_attr(self, name, key, value):
    self.f[name].attrs[key] = value

  def close(self):
    self.f.close()


def open_sink(target, slice_length):
  '''Sink for a ``save_hdf5`` argument (see the module docstring).'''
  target = str(target)
  if target.endswith('.npz'):
    return NpzSink(target)
  return Hdf5Sink(target, slice_length)


def device_blocks(t, npts, slice_length):
  '''(first point, host block) pairs of a device tensor (..., >= npts), one
  slice of grid points at a time.'''
  step = max(1, int(slice_length))
  for p0 in range(0, npts, step):
    n = min(step, npts - p0)
    yield p0, t[..., p0:p0 + n].cpu().numpy()


def write_grid(sink, x, y, z, is_vector, is_regular):
  '''The grid group as orbkit writes it.'''
  sink.put('grid/x', numpy.asarray(x))
  sink.put('grid/y', numpy.asarray(y))
  sink.put('grid/z', numpy.asarray(z))
  sink.put('grid/is_vector', numpy.asarray(bool(is_vector)))
  if is_regular is not None:
    sink.put('grid/is_regular', numpy.asarray(bool(is_regular)))


def read_npz(path):
  '''The datasets and attributes of an archive written by NpzSink, as a dict
  of name -> array (``numpy.load`` without pickles).'''
  with numpy.load(path, allow_pickle=False) as z:
    return {os.path.splitext(k)[0] if k.endswith('.npy') else k: z[k] for k in z.files}
