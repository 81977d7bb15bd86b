'''Container for one quantum-chemistry result (orbkit/qcinfo.py:35-110).'''
import numpy

from .orbitals import AOClass, MOClass
from .units import aa_to_a0


class QCinfo:
  '''Holds ``geo_spec`` (NATOMS x 3, Bohr), ``geo_info``, ``ao_spec`` and
  ``mo_spec``.  Built empty, or from a dict with those keys, where the two
  specs may be container objects, lists of dicts or restart dicts.'''

  def __init__(self, data=None):
    self.geo_info = []
    self.geo_spec = []
    self.etot = None
    self.ao_spec = AOClass()
    self.mo_spec = MOClass()
    if data:
      self.geo_spec = numpy.array(data['geo_spec'], dtype=float)
      self.geo_info = numpy.array(data.get('geo_info', []))
      ao, mo = data['ao_spec'], data['mo_spec']
      if isinstance(ao, numpy.ndarray):
        ao = ao[numpy.newaxis][0]
      if isinstance(mo, numpy.ndarray):
        mo = mo[numpy.newaxis][0]
      self.ao_spec = ao if hasattr(ao, 'get_lxlylz') else AOClass(ao)
      self.mo_spec = mo if hasattr(mo, 'get_coeffs') else MOClass(mo) if isinstance(mo, list) \
          else MOClass(restart=mo)

  def update(self):
    self.ao_spec.update()
    self.mo_spec.update()

  def format_geo(self, is_angstrom=False):
    self.geo_info = numpy.array(self.geo_info)
    self.geo_spec = numpy.array(self.geo_spec, dtype=float)
    if is_angstrom:
      self.geo_spec *= aa_to_a0

  def copy(self):
    from copy import deepcopy
    out = deepcopy(self)
    out.update()
    return out
