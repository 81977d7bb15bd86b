'''Readers that build the qc object consumed by ``core``.'''
from .molden import read_molden
from .gaussian_fchk import read_gaussian_fchk


def main_read(fname, itype='auto', all_mo=False, spin=None, **kwargs):
  '''Reads a molden or Gaussian fchk file (orbkit/read/high_level.py:33-79,
  restricted to these two formats).'''
  if itype == 'auto':
    low = str(fname).lower()
    if low.endswith('.fchk') or low.endswith('.fch'):
      itype = 'gaussian.fchk'
    else:
      itype = 'molden'
  if itype == 'molden':
    return read_molden(fname, all_mo=all_mo, spin=spin, **kwargs)
  if itype in ('gaussian.fchk', 'fchk'):
    return read_gaussian_fchk(fname, all_mo=all_mo, spin=spin, **kwargs)
  raise NotImplementedError('reader %r is not provided' % itype)
