'''Tables and small helpers used by the grid path.

Mirrors the parts of ``orbkit/tools.py`` that the density path reads:

* ``lquant`` / ``l_deg``               orbkit/tools.py:88-115
* ``exp`` (Cartesian component order)  orbkit/tools.py:117-135 (molden order)
* ``cart2sph`` / ``get_cart2sph``      orbkit/tools.py:148-220
* ``validate_drv``                     orbkit/tools.py:225-239
* ``require`` / ``convert``            orbkit/tools.py:290-302 (array helpers)

The numeric content of ``exp`` and ``cart2sph`` has to equal the reference's
(they define which AO row is which and the real-solid-harmonic combinations),
including the two g-shell rows whose norm is not one in the reference; the
representation here is a per-(l, m) list of (lx, ly, lz, coefficient) terms.
'''
from math import sqrt

import numpy

#: shell letter of each angular momentum l = 0, 1, 2, ...: s, p, d, then the
#: alphabet from f on without the letters already taken (s, p)
orbit = 'spdfghijklmnoqrtuvwxyz'
lquant = {c: i for i, c in enumerate(orbit)}


def l_deg(l=0, ao=None, cartesian_basis=True):
  '''Number of components of a shell of angular momentum ``l``.'''
  if ao is not None:
    l = 0 if ao == 's' else len(ao)
  elif not isinstance(l, int):
    l = lquant[l]
  return (l + 1) * (l + 2) // 2 if cartesian_basis else 2 * l + 1


def _cart_order():
  s = [(0, 0, 0)]
  p = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
  d = [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
  f = [(3, 0, 0), (0, 3, 0), (0, 0, 3), (1, 2, 0), (2, 1, 0), (2, 0, 1),
       (1, 0, 2), (0, 1, 2), (0, 2, 1), (1, 1, 1)]
  g = [(4, 0, 0), (0, 4, 0), (0, 0, 4), (3, 1, 0), (3, 0, 1), (1, 3, 0),
       (0, 3, 1), (1, 0, 3), (0, 1, 3), (2, 2, 0), (2, 0, 2), (0, 2, 2),
       (2, 1, 1), (1, 2, 1), (1, 1, 2)]
  return [s, p, d, f, g]


#: Cartesian component order per l (molden convention), l = 0..4.
exp = _cart_order()


def _sph_table():
  '''Real solid harmonics as Cartesian combinations (Schlegel & Frisch,
  IJQC 54, 83 (1995)), keyed by (l, m): (terms, global_factor) with terms a
  list of ((lx, ly, lz), coefficient).'''
  r3, r5 = sqrt(3.), sqrt(5.)
  t = {}
  t[0, 0] = ([((0, 0, 0), 1.)], 1.)
  t[1, -1] = ([((0, 1, 0), 1.)], 1.)
  t[1, 0] = ([((0, 0, 1), 1.)], 1.)
  t[1, 1] = ([((1, 0, 0), 1.)], 1.)
  t[2, -2] = ([((1, 1, 0), 1.)], 1.)
  t[2, -1] = ([((0, 1, 1), 1.)], 1.)
  t[2, 0] = ([((0, 0, 2), 1.), ((2, 0, 0), -.5), ((0, 2, 0), -.5)], 1.)
  t[2, 1] = ([((1, 0, 1), 1.)], 1.)
  t[2, 2] = ([((2, 0, 0), 1.), ((0, 2, 0), -1.)], r3 / 2.)
  t[3, -3] = ([((0, 3, 0), -r5), ((2, 1, 0), 3.)], 1 / (2. * sqrt(2)))
  t[3, -2] = ([((1, 1, 1), 1.)], 1.)
  t[3, -1] = ([((0, 1, 2), sqrt(3 / 5.)), ((0, 3, 0), -r3 / 4.),
               ((2, 1, 0), -r3 / (4. * r5))], sqrt(2))
  t[3, 0] = ([((0, 0, 3), 1.), ((2, 0, 1), -3 / (2 * r5)),
              ((0, 2, 1), -3 / (2 * r5))], 1.)
  t[3, 1] = ([((1, 0, 2), sqrt(3 / 5.)), ((3, 0, 0), -r3 / 4.),
              ((1, 2, 0), -r3 / (4. * r5))], sqrt(2))
  t[3, 2] = ([((2, 0, 1), 1.), ((0, 2, 1), -1.)], r3 / 2.)
  t[3, 3] = ([((3, 0, 0), r5), ((1, 2, 0), -3.)], 1 / (2. * sqrt(2)))
  s7, s14 = sqrt(7.), sqrt(14.)
  t[4, -4] = ([((3, 1, 0), 1.), ((1, 3, 0), -1.)], sqrt(2) * sqrt(5 / 8.))
  t[4, -3] = ([((0, 3, 1), -r5 / 4.), ((2, 1, 1), 3 / 4.)], sqrt(2))
  t[4, -2] = ([((1, 1, 2), 3 / s14), ((3, 1, 0), -r5 / (2 * s14)),
               ((1, 3, 0), -r5 / (2 * s14))], sqrt(2))
  # the reference lists (0,3,1) twice here and has no (0,1,3) term
  t[4, -1] = ([((0, 3, 1), sqrt(5 / 7.)), ((0, 3, 1), -3 * r5 / (4. * s7)),
               ((2, 1, 1), -3 / (4. * s7))], sqrt(2))
  # the reference's global factor here is sqrt(2)
  t[4, 0] = ([((0, 0, 4), 1.), ((4, 0, 0), 3 / 8.), ((0, 4, 0), 3 / 8.),
              ((2, 0, 2), -3 * r3 / sqrt(35)), ((0, 2, 2), -3 * r3 / sqrt(35)),
              ((2, 2, 0), -1 / 4.)], sqrt(2))
  t[4, 1] = ([((1, 0, 3), sqrt(5 / 7.)), ((3, 0, 1), -3 * r5 / (4. * s7)),
              ((1, 2, 1), -3 / (4. * s7))], sqrt(2))
  t[4, 2] = ([((2, 0, 2), 3 * r3 / (2. * s14)), ((0, 2, 2), -3 * r3 / (2. * s14)),
              ((4, 0, 0), -r5 / (4. * sqrt(2))), ((0, 4, 0), r5 / (4. * sqrt(2)))], sqrt(2))
  t[4, 3] = ([((3, 0, 1), r5 / 4.), ((1, 2, 1), -3 / 4.)], sqrt(2))
  t[4, 4] = ([((4, 0, 0), sqrt(35) / (8. * sqrt(2))), ((0, 4, 0), sqrt(35) / (8. * sqrt(2))),
              ((2, 2, 0), -3 * r3 / (4. * sqrt(2)))], sqrt(2))
  return t


_SPH = _sph_table()


def get_cart2sph(l, m):
  '''Returns ``[exponents, coefficients, global_factor]`` for the real solid
  harmonic (l, m), the same triple as orbkit/tools.py:193-220.'''
  terms, glob = _SPH[int(l), int(m)]
  return [[c for c, _ in terms], [v for _, v in terms], glob]


def validate_drv(drv):
  '''Maps a derivative label to the integer code 0-9 (orbkit/tools.py:225-239).'''
  if drv is None or drv == 'None' or drv == '':
    return 0
  codes = {'x': 1, 'y': 2, 'z': 3, 'xx': 4, 'x2': 4, 'yy': 5, 'y2': 5,
           'zz': 6, 'z2': 6, 'xy': 7, 'yx': 7, 'xz': 8, 'zx': 8, 'yz': 9, 'zy': 9}
  if isinstance(drv, str) and drv in codes:
    return codes[drv]
  if isinstance(drv, (int, numpy.integer)) and 0 <= drv <= 9:
    return int(drv)
  raise ValueError('The selection `drv=%s` is not valid!' % drv)


_DTYPES = {'f': numpy.float64, 'i': numpy.intc}


def require(data, dtype='f', requirements='CA'):
  """C-contiguous, aligned array of float64 ('f'), C int ('i') or any
  numpy dtype, copied only when ``data`` is not one already."""
  return numpy.require(data, dtype=_DTYPES.get(dtype, dtype), requirements=requirements)


def convert(data, was_vector, N):
  """A fresh C-ordered array; for a regular grid (``was_vector`` false) the
  last axis is unfolded into the grid shape ``N``."""
  out = numpy.array(data, order='C')
  return out if was_vector else out.reshape(out.shape[:-1] + tuple(N))
