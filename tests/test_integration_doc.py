'''INTEGRATION.md's stand-ins for orbkit's Cython calls, run as written.

The first python block of the document's section 2 defines `aocreator` and
`mocreator` with the argument lists of cy_core.aocreator and
cy_core.mocreator (orbkit/cy_core.pyx:49-101).  The block is executed from
the document itself, then called the way orbkit's core.ao_creator and
mo_creator call the Cython functions (orbkit/core.py:87-97, :125-131),
against the oracle at the FP64 bar.
'''
import os
import re

import numpy
import pytest

from conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu


def _binding_source():
  text = open(os.path.join(ROOT, 'INTEGRATION.md')).read()
  sec = text[text.index('## 2.'):]
  return re.search(r'```python\n(.*?)```', sec, re.S).group(1)


@pytest.fixture(scope='module')
def binding():
  import torch
  if not torch.cuda.is_available():
    pytest.skip('no HIP device')
  ns = {}
  exec(compile(_binding_source(), 'INTEGRATION.md section 2', 'exec'), ns)
  return ns


@pytest.mark.parametrize('name', ['small', 'medium'])
def test_documented_aocreator_and_mocreator(binding, name):
  from orbkit_amd import workloads
  from oracle import core_oracle as O
  qc, axes, _ = workloads.make(name, npts_side=16)
  ao = qc.ao_spec
  X, Y, Z = numpy.meshgrid(*axes, indexing='ij')
  idx = numpy.random.default_rng(4).choice(X.size, size=777, replace=False)
  x, y, z = X.ravel()[idx].copy(), Y.ravel()[idx].copy(), Z.ravel()[idx].copy()
  f = O.flat_basis(qc.geo_spec, ao)
  for drv in range(10):
    print('case documented aocreator', name, drv, flush=True)
    got = binding['aocreator'](ao.get_lxlylz(), ao.get_nlxlylz_per_cont(), ao.get_prim_coeffs(),
                               ao.get_nprim_per_cont(), qc.geo_spec, ao.get_assign_cont_to_atoms(),
                               x, y, z, drv, ao.get_normalized())
    ref = O.aocreator(f, x, y, z, drv)
    assert got.shape == ref.shape
    assert rel_err(got, ref) <= 1e-12, (drv, rel_err(got, ref))
  cart = O.aocreator(f, x, y, z, 0)
  C = numpy.random.default_rng(5).normal(size=(7, cart.shape[0]))
  mo = binding['mocreator'](cart, C)
  assert rel_err(mo, C @ cart) <= 1e-12
