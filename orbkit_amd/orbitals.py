'''Basis-set and MO containers with the interface the grid path reads.

``AOClass`` and ``MOClass`` mirror the getters of orbkit's classes
(orbkit/orbitals.py:25-409 and :411-832) that ``core`` uses: the flattened
basis arrays handed to the AO kernel (``get_lxlylz``, ``get_nlxlylz_per_cont``,
``get_prim_coeffs``, ``get_nprim_per_cont``, ``get_assign_cont_to_atoms``,
``get_normalized``), the spherical (l, m) list (``set_lm_dict``,
``get_old_ao_spherical``) and the MO coefficient/occupation arrays.  Any object
with these methods (orbkit's own classes included) is accepted by ``core``.
'''
from collections import UserList
from copy import copy
import warnings

import numpy

from .tools import exp, lquant, require


class AOClass(UserList):
  '''List of contractions; each entry is a dict with keys ``atom``, ``type``,
  ``pnum``, ``coeffs`` (pnum x 2: exponent, coefficient) and optionally
  ``lxlylz`` (custom Cartesian component order) and ``lm`` (spherical).'''

  def __init__(self, data=None, seq=(), restart=None):
    if isinstance(data, list):
      seq = data
    elif isinstance(data, dict):
      restart = data
    UserList.__init__(self, list(seq))
    self._up_to_date = False
    self.normalized = False
    self.spherical = False
    if restart is not None:
      self._from_restart(restart)

  # -- flattening (orbkit/orbitals.py:226-316) ------------------------------
  def _from_restart(self, r):
    self.normalized = bool(r['normalized'])
    self.spherical = bool(r['spherical'])
    atoms = numpy.asarray(r['_assign_cont_to_atoms'])
    types = list(r['_cont_types'])
    pc = numpy.asarray(r['_prim_coeffs'], dtype=float)
    p2c = numpy.asarray(r['_assign_prim_to_cont'])
    lxlylz = numpy.asarray(r['_lxlylz'])
    l2c = numpy.asarray(r['_assign_lxlylz_to_cont'])
    lm = r.get('_lm', None)
    lm2c = r.get('_assign_lm_to_cont', None)
    data = []
    for ic in range(len(atoms)):
      e = {'atom': int(atoms[ic]), 'type': types[ic],
           'pnum': int(r['_nprim_per_cont'][ic]), 'coeffs': pc[p2c == ic],
           'lxlylz': [tuple(int(v) for v in t) for t in lxlylz[l2c == ic]]}
      if self.spherical:
        e['lm'] = [tuple(int(v) for v in t) for t in numpy.asarray(lm)[numpy.asarray(lm2c) == ic]]
      data.append(e)
    self.data = data
    self.update()

  def todict(self):
    self.update()
    return {'normalized': self.normalized, 'spherical': self.spherical,
            '_assign_cont_to_atoms': self._assign_cont_to_atoms,
            '_cont_types': self._cont_types, '_nprim_per_cont': self._nprim_per_cont,
            '_prim_coeffs': self._prim_coeffs, '_assign_prim_to_cont': self._assign_prim_to_cont,
            '_lxlylz': self._lxlylz, '_assign_lxlylz_to_cont': self._assign_lxlylz_to_cont,
            '_nlxlylz_per_cont': self._nlxlylz_per_cont, '_lm': self._lm,
            '_assign_lm_to_cont': self._assign_lm_to_cont}

  def __getitem__(self, item):
    if isinstance(item, (int, numpy.integer)):
      return self.data[item]
    out = AOClass(seq=self.data[item])
    out.spherical = self.spherical
    out.update()
    return out

  def __setitem__(self, i, item):
    self.data[i] = item
    self._up_to_date = False

  def __delitem__(self, i):
    del self.data[i]
    self._up_to_date = False

  def append(self, item):
    UserList.append(self, item)
    self._up_to_date = False

  def extend(self, item):
    UserList.extend(self, item)
    self._up_to_date = False

  def update(self):
    self._up_to_date = False
    if not self.data:
      raise ValueError('ao_spec not initialized')
    for i, c in enumerate(self.data):
      if not isinstance(c, dict):
        raise ValueError('ao_spec[{0}] has to be a dictionary'.format(i))
      missing = [k for k in ('atom', 'type', 'pnum', 'coeffs') if k not in c]
      if self.spherical and 'lm' not in c:
        missing.append('lm')
      if missing:
        raise ValueError('ao_spec[{0}] misses {1}'.format(i, str(missing)))
    atoms, types, npc, pcs, p2c = [], [], [], [], []
    lxl, l2c = [], []
    for i, c in enumerate(self.data):
      atoms.append(c['atom'])
      types.append(c['type'])
      co = numpy.asarray(c['coeffs'], dtype=float).reshape(-1, 2)
      npc.append(len(co))
      pcs.append(co)
      p2c.extend([i] * len(co))
      comps = c['lxlylz'] if 'lxlylz' in c else exp[lquant[c['type']]]
      lxl.extend([tuple(int(v) for v in t) for t in comps])
      l2c.extend([i] * len(comps))
    self._assign_cont_to_atoms = require(atoms, dtype='i')
    self._cont_types = types
    self._nprim_per_cont = require(npc, dtype='i')
    self._prim_coeffs = require(numpy.concatenate(pcs, axis=0) if pcs else numpy.zeros((0, 2)), dtype='f')
    self._assign_prim_to_cont = require(p2c, dtype='i')
    self._lxlylz = require(numpy.array(lxl, dtype=numpy.intc).reshape(-1, 3), dtype='i')
    self._assign_lxlylz_to_cont = require(l2c, dtype='i')
    self._nlxlylz_per_cont = require(numpy.bincount(self._assign_lxlylz_to_cont,
                                                    minlength=len(self.data)), dtype='i')
    self._lm, self._assign_lm_to_cont = None, None
    if self.spherical:
      lm, lm2c = [], []
      for i, c in enumerate(self.data):
        for t in c['lm']:
          lm.append(tuple(int(v) for v in t))
          lm2c.append(i)
      self._lm = numpy.array(lm, dtype=numpy.intc).reshape(-1, 2)
      self._assign_lm_to_cont = require(lm2c, dtype='i')
    flags = [c['pnum'] < 0 for c in self.data]
    if all(flags) != any(flags):
      raise ValueError('Either all or none of the atomic orbitals have to be normalized!')
    self.normalized = all(flags)
    self._up_to_date = True

  def set_lm_dict(self, p=(1, 0)):
    '''Spherical (l, m) list per contraction in orbkit's order
    (l,0),(l,1),(l,-1),(l,2),(l,-2)...; p shells use ``p`` for m.'''
    for c in self.data:
      l = lquant[c['type']]
      c['lm'] = []
      for m in (range(0, l + 1) if l != 1 else p):
        c['lm'].append((l, m))
        if m != 0:
          c['lm'].append((l, -m))
    self.spherical = True
    self._up_to_date = False

  def _get(self, name):
    if not self._up_to_date:
      self.update()
    return copy(getattr(self, name))

  def get_assign_cont_to_atoms(self): return self._get('_assign_cont_to_atoms')
  def get_cont_types(self): return self._get('_cont_types')
  def get_nprim_per_cont(self): return self._get('_nprim_per_cont')
  def get_prim_coeffs(self): return self._get('_prim_coeffs')
  def get_assign_prim_to_cont(self): return self._get('_assign_prim_to_cont')
  def get_lxlylz(self): return self._get('_lxlylz')
  def get_assign_lxlylz_to_cont(self): return self._get('_assign_lxlylz_to_cont')
  def get_nlxlylz_per_cont(self): return self._get('_nlxlylz_per_cont')
  def get_lm(self): return self._get('_lm')
  def get_assign_lm_to_cont(self): return self._get('_assign_lm_to_cont')

  def get_normalized(self):
    if not self._up_to_date:
      self.update()
    return int(self.normalized)

  def get_old_ao_spherical(self):
    if not self._up_to_date:
      self.update()
    if not self.spherical:
      return []
    return list(zip(self.get_assign_lm_to_cont(), [tuple(t) for t in self.get_lm()]))

  def get_labels(self):
    if not self._up_to_date:
      self.update()
    atoms = self._assign_cont_to_atoms
    if self.spherical:
      return ['l,m=%s,atom=%d' % (tuple(self._lm[i]), atoms[j])
              for i, j in enumerate(self._assign_lm_to_cont)]
    return ['lxlylz=%s,atom=%d' % (self._lxlylz[i], atoms[j])
            for i, j in enumerate(self._assign_lxlylz_to_cont)]

  def get_ao_num(self):
    if not self._up_to_date:
      self.update()
    return len(self._lm) if self.spherical else len(self._lxlylz)


class MOClass(UserList):
  '''List of MOs; each entry is a dict with keys ``coeffs`` (NAO),
  ``occ_num``, ``energy``, ``sym`` and optionally ``spin``.'''

  def __init__(self, seq=(), restart=None, data=None):
    if isinstance(data, list):
      seq = data
    elif isinstance(data, dict):
      restart = data
    UserList.__init__(self, list(seq))
    self._up_to_date = False
    self.selected_mo = None
    if restart is not None:
      coeffs = numpy.asarray(restart['coeffs'], dtype=float)
      occ = numpy.asarray(restart['occ'], dtype=float)
      eig = numpy.asarray(restart['eig'], dtype=float)
      sym = list(restart['sym'])
      self.data = [{'coeffs': coeffs[i], 'occ_num': occ[i], 'energy': eig[i], 'sym': sym[i]}
                   for i in range(len(occ))]

  def todict(self):
    return {'coeffs': self.get_coeffs(), 'occ': self.get_occ(), 'eig': self.get_eig(),
            'sym': self.get_sym(), 'spinpolarized': False, 'alpha_index': None,
            'beta_index': None}

  def __getitem__(self, item):
    if isinstance(item, (int, numpy.integer)):
      return self.data[item]
    if isinstance(item, slice):
      return MOClass(self.data[item])
    if isinstance(item, str):
      return self.select(item)
    idx = numpy.asarray(item)
    if idx.dtype == bool:
      idx = numpy.nonzero(idx)[0]
    out = MOClass([self.data[int(i)] for i in idx])
    out.selected_mo = idx
    return out

  def __setitem__(self, i, item):
    self.data[i] = item
    self._up_to_date = False

  def __delitem__(self, i):
    del self.data[i]
    self._up_to_date = False

  def update(self):
    self._up_to_date = True

  def select(self, key):
    '''Minimal string selection: 'homo', 'lumo', 'all_mo' or an MO label.
    (The full selection grammar of orbkit/orbitals.py:896-1210 is outside the
    grid path.)'''
    k = key.strip().lower()
    occ, eig = self.get_occ(), self.get_eig()
    order = numpy.argsort(eig, kind='stable')
    if k == 'all_mo':
      return MOClass(list(self.data))
    if k in ('homo', 'lumo'):
      sel = [i for i in order if (occ[i] > 1e-5) == (k == 'homo')]
      if not sel:
        raise ValueError('no %s found' % k.upper())
      return self[[sel[-1] if k == 'homo' else sel[0]]]
    labels = [m['sym'] for m in self.data]
    if key in labels:
      return self[[labels.index(key)]]
    raise ValueError('MO selection %r is not supported' % key)

  def get_coeffs(self):
    return numpy.array([numpy.asarray(m['coeffs'], dtype=float) for m in self.data], dtype=float)

  def get_occ(self):
    return numpy.array([m['occ_num'] for m in self.data], dtype=float)

  def get_eig(self):
    return numpy.array([0.0 if m.get('energy') is None else float(m['energy']) for m in self.data])

  def get_sym(self):
    return numpy.array([str(m.get('sym', '%d.1' % (i + 1))) for i, m in enumerate(self.data)])

  def get_labels(self, format='default'):
    def fmt(m, i):
      d = {'sym': m.get('sym', '%d.1' % (i + 1)), 'occ_num': float(m['occ_num']),
           'energy': 0.0 if m.get('energy') is None else float(m['energy'])}
      if format == 'short':
        return '%(sym)s' % d
      if format == 'print':
        return '%(sym)s (Occ = %(occ_num).2f, E = %(energy)+.4f E_h)' % d
      if format in ('cb', 'cube', 'vmd'):
        return '%(sym)s,Occ=%(occ_num).1f,E=%(energy)+.2f' % d
      return '%(sym)s, Occ=%(occ_num).2f, E=%(energy)+.4f E_h' % d
    return [fmt(m, i) for i, m in enumerate(self.data)]

  def get_homo(self, tol=1e-5):
    occ = self.get_occ()
    idx = numpy.nonzero(occ > tol)[0]
    return None if not len(idx) else idx[-1]

  def sort_by_energy(self):
    eig = self.get_eig()
    if not all(numpy.diff(eig) >= 0):
      warnings.warn('MOs are not sorted by energy. Sorting them...', UserWarning)
      self.data = [self.data[i] for i in numpy.argsort(eig)]
