'''Module-global grid state with orbkit's semantics (orbkit/grid.py:31-291, :553-677).

A regular grid is three 1-D axes ``x, y, z``; ``grid2vector`` expands it to
``Nx*Ny*Nz`` points with x slowest and z fastest (orbkit/cy_grid.pyx:12-30)
and ``vector2grid`` takes the axes back (orbkit/cy_grid.pyx:32-55).
'''
import sys

import numpy

_mod = sys.modules[__name__]


def _defaults():
  _mod.min_ = [-8.0, -8.0, -8.0]
  _mod.max_ = [8.0, 8.0, 8.0]
  _mod.N_ = [101, 101, 101]
  _mod.delta_ = numpy.zeros((3, 1))
  _mod.is_initialized = False
  _mod.is_vector = True
  _mod.is_regular = False


x = numpy.array([0.0])
y = numpy.array([0.0])
z = numpy.array([0.0])
d3r = 0.0
_defaults()


def grid_init(is_vector=False, force=False):
  '''Builds the regular axes from ``min_``, ``max_`` and ``delta_`` (arange)
  or ``N_`` (linspace) as orbkit/grid.py:31-84 does.'''
  global x, y, z, d3r, is_initialized, is_regular
  if is_initialized and not force:
    return 0
  axes = [None, None, None]
  step = lambda i: float(numpy.asarray(delta_[i]).reshape(-1)[0])
  for i in range(3):
    if max_[i] == min_[i]:
      axes[i] = numpy.array([min_[i]], dtype=numpy.float64)
      delta_[i] = 1
      N_[i] = 1
    elif step(i):
      axes[i] = numpy.arange(min_[i], max_[i] + step(i), step(i), dtype=numpy.float64)
      N_[i] = len(axes[i])
    else:
      axes[i] = numpy.array(numpy.linspace(min_[i], max_[i], N_[i]), dtype=numpy.float64)
      delta_[i] = axes[i][1] - axes[i][0]
  x, y, z = axes
  d3r = float(numpy.prod([step(i) for i in range(3)]))
  is_initialized = True
  is_regular = True
  if is_vector:
    grid2vector()
  else:
    _mod.is_vector = False


init = grid_init
init_grid = grid_init


def expand(xa, ya, za):
  '''Regular axes -> vector grid, x slowest, z fastest.'''
  xa, ya, za = (numpy.asarray(v, dtype=numpy.float64).reshape(-1) for v in (xa, ya, za))
  X, Y, Z = numpy.meshgrid(xa, ya, za, indexing='ij')
  return X.reshape(-1).copy(), Y.reshape(-1).copy(), Z.reshape(-1).copy()


def grid2vector():
  global x, y, z, is_regular
  if not is_initialized:
    raise ValueError('You have to initialize a grid before calling '
                     ' `grid.grid2vector`, i.e., `grid.is_initialized=True`.')
  x, y, z = expand(x, y, z)
  _mod.is_vector = True
  is_regular = True


def vector2grid(Nx, Ny, Nz):
  global x, y, z
  if not is_initialized:
    raise ValueError('You have to initialize a grid before calling `grid.vector2grid`. '
                     '(`grid.is_initialized == True`)')
  if not is_regular:
    raise ValueError('The grid has to regular. (`grid.is_regular == True`)')
  if not (len(x) == len(y) == len(z)):
    raise ValueError('Not a valid vector grid, i.e., dimensions of '
                     'x-, y-, and z- coordinate differ.')
  if Nx * Ny * Nz != len(x):
    raise ValueError('It has to hold that `len(x) = (N_x * N_y * N_z)`')
  x = numpy.array(x[::Ny * Nz][:Nx])
  y = numpy.array(y[:Ny * Nz:Nz][:Ny])
  z = numpy.array(z[:Nz])
  _mod.is_vector = False


def set_grid(xnew, ynew, znew, is_vector=None):
  '''Sets the axes (regular) or point lists (vector) directly.'''
  global x, y, z, is_initialized, is_regular
  reset_grid()
  vals = []
  for c in (xnew, ynew, znew):
    if isinstance(c, (int, float)):
      c = [c]
    vals.append(numpy.asarray(c, dtype=numpy.float64).reshape(-1))
  x, y, z = vals
  is_initialized = True
  if isinstance(is_vector, bool):
    _mod.is_vector = is_vector
  is_regular = is_vector is False
  set_boundaries(is_regular)


def set_boundaries(is_regular, Nx=None, Ny=None, Nz=None):
  _mod.min_ = [x.min(), y.min(), z.min()]
  _mod.max_ = [x.max(), y.max(), z.max()]
  _mod.N_ = [len(x), len(y), len(z)]
  if is_regular:
    f = lambda a: 1.0 if len(a) <= 1 else a[1] - a[0]
    _mod.delta_ = [f(x), f(y), f(z)]


def adjust_to_geo(qc, extend=5.0, step=0.1):
  '''Box around the nuclei extended by ``extend`` with spacing ``step``
  (orbkit/grid.py:553-581).'''
  geo = numpy.asarray(qc.geo_spec, dtype=float)
  for i in range(3):
    min_[i] = min(geo[:, i]) - extend
    max_[i] = max(geo[:, i]) + extend
    dist = max_[i] - min_[i]
    N_[i] = int(numpy.ceil(dist / step)) + 1
    rest = (N_[i] - 1) * step - dist
    min_[i] -= rest / 2.
    max_[i] += rest / 2.
    delta_[i] = step
  _mod.is_vector = False
  _mod.is_initialized = False


def get_shape():
  if not is_initialized:
    raise ValueError('`grid.get_shape` requires the grid to be initialized.')
  return (len(x),) if is_vector else tuple(N_)


def get_grid(start='\t'):
  out = ''
  for c, a, i in zip('xyz', (x, y, z), range(3)):
    out += '%s%s[0] = %.2f %s[-1] = %.2f N%s = %d ' % (start, c, a[0], c, a[-1], c, len(a))
    if max_[i] != min_[i] and numpy.all(delta_[i] != 0.):
      out += 'd%s = %.3f' % (c, float(numpy.asarray(delta_[i]).reshape(-1)[0]))
    out += '\n'
  return out


def tolist():
  return [numpy.copy(x), numpy.copy(y), numpy.copy(z)]


def todict():
  return {'x': x, 'y': y, 'z': z}


def mv2g(**kwargs):
  '''Reshapes arrays with one axis of length ``prod(N_)`` to (Nx, Ny, Nz)
  in place of that axis (orbkit/grid.py:268-291).'''
  out = {}
  n = int(numpy.prod(N_))
  for k, v in kwargs.items():
    v = numpy.asarray(v, dtype=float)
    where = list(v.shape).index(n)
    out[k] = v.reshape(v.shape[:where] + tuple(N_) + v.shape[where + 1:])
  return list(out.values())[0] if len(out) == 1 else out


def reset_grid():
  _defaults()
