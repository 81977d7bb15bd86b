'''Molden reader producing a ``QCinfo`` (orbkit/read/molden.py:46-402).

Reads [Atoms], [GTO], the [5D]/[7F]/[9G]/[6D]/[10F]/[15G] flags and [MO].
After reading, each contraction's self-overlap under orbkit's primitive
normalisation is computed in closed form; if any differs from one by more than
1e-5 the contraction coefficients are rescaled (ORCA files instead get the
per-row factor ``ao_spec[0]['N']``), as orbkit/read/molden.py:377-398 does.
'''
import re

import numpy

from ..orbitals import AOClass, MOClass
from ..qcinfo import QCinfo
from ..tools import l_deg, lquant, orbit
from ..display import display
from .overlap import self_overlap_rows, omitted_cca_norm

_F = r'([-+]?\d+\.?\d*[eEdD]?[+-]?\d+|NaN)'
_re_molden = re.compile(r'\[\s*molden\s+format\s*\]', re.I)
_re_atoms = re.compile(r'\[atoms\]\s*\(?(angs|au)\)?', re.I)
_re_atom = re.compile(r'\s*([a-z]+)\s+(\d+)\s+(\d+)\s+' + r'\s+'.join((_F,) * 3), re.I)
_re_center = re.compile(r'\s*(\d+)\s+(\d+)$')
_re_shell = re.compile(r'\s*([a-z]+)\s+(\d+)\s+(\d+(\.\d+)?)\s*$', re.I)
_re_prim = re.compile(r'\s*' + r'\s+'.join((_F,) * 2), re.I)
_SPH_FLAGS = ('5d', '7f', '9g')
_CART_FLAGS = ('6d', '10f', '15g')
_re_flagline = re.compile(r'\[((' + '|'.join(_SPH_FLAGS + _CART_FLAGS) + r')+)\]', re.I)
_re_flag = re.compile(r'(\d+[dfg])', re.I)
_re_sym = re.compile(r'\s*sym\s*=\s*(\S+)', re.I)
_re_ene = re.compile(r'\s*ene(?:rgy)?\s*=\s*' + _F, re.I)
_re_spin = re.compile(r'\s*spin\s*=\s*(alpha|beta)', re.I)
_re_occ = re.compile(r'\s*occup\s*=\s*' + _F, re.I)
_re_coeff = re.compile(r'\s*(\d+)\s+' + _F, re.I)


def _f(s):
  return float(s.lower().replace('d', 'e'))


def read_molden(fname, all_mo=False, spin=None, i_md=-1, **kwargs):
  text = open(fname).read() if isinstance(fname, str) else fname.read()
  if isinstance(text, bytes):
    text = text.decode()
  starts = [m.start() for m in _re_molden.finditer(text)]
  if not starts:
    raise IOError('The input file %s is no valid molden file!' % fname)
  k = list(range(len(starts)))[i_md]
  text = text[starts[k]:(starts + [None])[k + 1]]
  lines = text.splitlines()

  qc = QCinfo()
  geo_info, geo_spec = [], []
  sph_flags, cart_flags, used_l = [], [], []
  by_orca = False
  angstrom = False
  at = 0
  shell_types, nprim_left = '', 0
  iline = 0
  for iline, line in enumerate(lines):
    low = line.lower()
    if 'orca' in low:
      by_orca = True
      continue
    m = _re_atoms.match(line)
    if m:
      angstrom = m.group(1).lower() == 'angs'
      continue
    m = _re_atom.match(line)
    if m:
      geo_info.append(list(m.groups()[:3]))
      geo_spec.append([_f(v) for v in m.groups()[3:]])
      continue
    if '[sto]' in low:
      raise IOError('orbkit does not work for STOs!')
    m = _re_center.match(line)
    if m:
      at = int(m.group(1)) - 1
      continue
    m = _re_flagline.match(low)
    if m:
      for fl in _re_flag.findall(m.group(1)):
        (sph_flags if fl in _SPH_FLAGS else cart_flags).append(fl)
    m = _re_shell.match(line)
    if m:
      shell_types = m.group(1).lower()
      pnum = int(m.group(2))
      nprim_left = 0
      for t in shell_types:
        qc.ao_spec.append({'atom': at, 'type': t, 'pnum': -pnum if by_orca else pnum,
                           'coeffs': numpy.zeros((pnum, 2))})
        if t not in used_l:
          used_l.append(t)
      continue
    m = _re_prim.match(line)
    if m:
      vals = [_f(v) for v in line.split()]
      for j in range(len(shell_types)):
        qc.ao_spec[-len(shell_types) + j]['coeffs'][nprim_left, :] = [vals[0], vals[1 + j]]
      nprim_left += 1
      continue
    if '[mo]' in low:
      break

  max_l = max(lquant[t] for t in used_l)
  cartesian = True
  if max_l >= 2:
    ls = orbit[2:max_l + 1]
    sph = [f for f in sph_flags if f[-1] in ls]
    cart = [f for f in cart_flags if f[-1] in ls]
    if sph and cart:
      raise IOError('mixed spherical and Cartesian functions are not supported')
    cartesian = not sph
  nbasis = sum(l_deg(lquant[c['type']], cartesian_basis=cartesian) for c in qc.ao_spec)

  mos = []
  cur = {}
  new = False
  for line in lines[iline:]:
    m = _re_coeff.match(line)
    if m:
      if new:
        mos.append({'coeffs': numpy.zeros(nbasis),
                    'sym': cur.get('sym') or '%d.1' % (len(mos) + 1),
                    'energy': cur.get('energy'), 'occ_num': cur.get('occ'),
                    'spin': cur.get('spin') or 'alpha'})
        cur, new = {}, False
      c = _f(m.group(2))
      if not numpy.isnan(c):
        mos[-1]['coeffs'][int(m.group(1)) - 1] = c
      continue
    new = True
    for key, rx, conv in (('sym', _re_sym, str), ('energy', _re_ene, _f),
                          ('spin', _re_spin, str.lower), ('occ', _re_occ, _f)):
      m = rx.match(line)
      if m:
        cur[key] = conv(m.group(1))
        break
  restricted = any(mo['occ_num'] is not None and mo['occ_num'] > 1.0001 for mo in mos)

  if sph_flags:
    qc.ao_spec.set_lm_dict(p=[1, 0])
  if not all_mo:
    mos = [mo for mo in mos if mo['occ_num'] >= 0.0000001]
  if spin is not None:
    mos = [mo for mo in mos if mo['spin'] == spin]
  for mo in mos:
    if restricted:
      del mo['spin']
    else:
      mo['sym'] += '_%s' % mo['spin'][0]
  # ORCA writes one label for every MO: number them in file order, keeping
  # the part after the last '.' (orbkit/read/molden.py:364-368)
  if mos and len({mo['sym'] for mo in mos}) == 1:
    suffix = mos[0]['sym'].split('.')[-1]
    for i, mo in enumerate(mos):
      mo['sym'] = '%d.%s' % (i + 1, suffix)
  qc.mo_spec = MOClass(mos)
  qc.geo_info, qc.geo_spec = geo_info, geo_spec
  qc.format_geo(is_angstrom=angstrom)

  norm = self_overlap_rows(qc.ao_spec)
  if numpy.max(numpy.abs(norm - 1.)) > 1e-5:
    display('The atomic orbitals are not normalized correctly, renormalizing...\n')
    if not by_orca:
      j = 0
      for c in qc.ao_spec:
        c['coeffs'][:, 1] /= numpy.sqrt(norm[j])
        j += l_deg(lquant[c['type']])
    else:
      qc.ao_spec[0]['N'] = 1 / numpy.sqrt(norm[:, numpy.newaxis])
    if cart_flags:
      cca = omitted_cca_norm(qc.ao_spec.get_lxlylz())
      for mo in qc.mo_spec:
        mo['coeffs'] *= cca
  qc.ao_spec.update()
  qc.mo_spec.update()
  return qc
