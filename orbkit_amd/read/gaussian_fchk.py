'''Gaussian formatted-checkpoint reader (orbkit/read/gaussian_fchk.py:11-324).

Reads geometry, shells (types, primitives per shell, shell-to-atom map,
exponents, contraction and SP coefficients) and alpha/beta MO coefficients
and energies.  Shell type < 0 means a pure (spherical) shell; the (l, m) order
per shell is (l,0),(l,1),(l,-1),(l,2),(l,-2)... and for p (1,1),(1,-1),(1,0).
Occupations follow from the electron counts (2 per orbital when restricted).
'''
from copy import deepcopy

import numpy

from ..orbitals import AOClass, MOClass
from ..qcinfo import QCinfo
from ..tools import orbit, lquant


def _sections(lines):
  '''Yields (label, kind, count, values) for every header of the file.'''
  i, n = 0, len(lines)
  while i < n:
    line = lines[i]
    i += 1
    if len(line) < 43 or line[43:45].strip() not in ('I', 'R', 'C', 'L', 'H'):
      continue
    label = line[:43].strip()
    kind = line[43:45].strip()
    rest = line[45:].split()
    if rest and rest[0] == 'N=':
      count = int(rest[1])
      vals = []
      while len(vals) < count and i < n:
        vals.extend(lines[i].split())
        i += 1
      if kind == 'I':
        vals = [int(v) for v in vals]
      elif kind == 'R':
        vals = [float(v) for v in vals]
      yield label, kind, count, vals
    else:
      v = rest[0] if rest else ''
      yield label, kind, None, (int(v) if kind == 'I' else float(v) if kind == 'R' else v)


def read_gaussian_fchk(fname, all_mo=False, spin=None, **kwargs):
  lines = open(fname, encoding='iso-8859-1').read().splitlines()
  sec = {label: vals for label, _, _, vals in _sections(lines)}
  is6d = sec.get('Pure/Cartesian d shells', 0) == 1
  is10f = sec.get('Pure/Cartesian f shells', 0) == 1
  cartesian = is6d and is10f
  if is6d != is10f:
    raise IOError('Please apply a Spherical Harmonics (5D, 7F) or a Cartesian Gaussian Basis Set (6D, 10F)!')
  has_beta = 'Beta MO coefficients' in sec
  if spin is not None and not has_beta:
    raise IOError('The keyword `spin` is only supported for unrestricted calculations.')
  restricted = not has_beta
  nel = [sec['Number of alpha electrons'], sec['Number of beta electrons']]
  nbas = sec['Number of basis functions']

  qc = QCinfo()
  qc.ao_spec.spherical = not cartesian
  types, npr, amap = sec['Shell types'], sec['Number of primitives per shell'], sec['Shell to atom map']
  ex, co = sec['Primitive exponents'], sec['Contraction coefficients']
  sp = sec.get('P(S=P) Contraction coefficients')
  p = 0
  shells = []
  for t, n, a in zip(types, npr, amap):
    l = abs(t)
    coeffs = numpy.array([ex[p:p + n], co[p:p + n]]).T.copy()
    sh = {'atom': a - 1, 'type': orbit[l], 'pnum': n, 'coeffs': coeffs}
    if not cartesian:
      sh['lm'] = []
      for m in (range(0, l + 1) if l != 1 else [1, 0]):
        sh['lm'].append((l, m))
        if m != 0:
          sh['lm'].append((l, -m))
    if t == -1 and sp is not None:
      # SP shell: an s shell and a p shell sharing exponents
      s_sh = deepcopy(sh)
      s_sh['type'] = 's'
      if not cartesian:
        s_sh['lm'] = [(0, 0)]
        sh['lm'] = [(1, 1), (1, -1), (1, 0)]
      sh['type'] = 'p'
      sh['coeffs'][:, 1] = sp[p:p + n]
      shells.extend([s_sh, sh])
    else:
      shells.append(sh)
    p += n
  for s in shells:
    qc.ao_spec.append(s)

  mos = []
  for what, ne in (('Alpha', nel[0]), ('Beta', nel[1])):
    key = '%s MO coefficients' % what
    if key not in sec:
      continue
    c = numpy.array(sec[key]).reshape(-1, nbas)
    e = numpy.array(sec['%s Orbital Energies' % what])
    if restricted:
      occ, nocc = (2.0, nel[0]) if nel[0] == nel[1] else (1.0, ne)
    else:
      occ, nocc = 1.0, ne
    for i in range(len(c)):
      mos.append({'coeffs': c[i].copy(), 'energy': float(e[i]),
                  'occ_num': float(occ if i < nocc else 0), 'sym': '%i.1' % (i + 1),
                  'spin': what.lower()})
  if not all_mo:
    mos = [m for m in mos if m['occ_num'] >= 0.0000001]
  if spin is not None:
    mos = [m for m in mos if m['spin'] == spin]
  for m in mos:
    if restricted:
      del m['spin']
    else:
      m['sym'] += '_%s' % m['spin'][0]
  qc.mo_spec = MOClass(mos)
  qc.geo_spec = numpy.array(sec['Current cartesian coordinates']).reshape(-1, 3)
  qc.geo_info = numpy.array([[str(z), str(i + 1), str(float(q))] for i, (z, q) in
                             enumerate(zip(sec['Atomic numbers'], sec['Nuclear charges']))])
  qc.format_geo(is_angstrom=False)
  qc.ao_spec.update()
  return qc
