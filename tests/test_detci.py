'''detCI grid reductions (orbkit/detci/ci_core.py, orbkit/detci/cy_ci.pyx).

CPU: the numpy restatement (oracle/ci_oracle.py) against invariants tied to
the pinned density oracle.  GPU: the device reductions (okmi_ci_reduce and
the FP64 GEMM for the time-dependent combinations) against the oracle on
random CI term lists (relative 1e-12).
'''
import numpy
import pytest

from conftest import rel_err


def _mos(n=700, seed=0):
  from oracle import core_oracle
  from orbkit_amd import workloads
  from orbkit_amd.tools import get_cart2sph
  qc, _, _ = workloads.make('small')
  geo = numpy.asarray(qc.geo_spec)
  rng = numpy.random.default_rng(seed)
  P = geo[rng.integers(0, len(geo), n)] + rng.normal(0.0, 1.5, (n, 3))
  x, y, z = (P[:, i].copy() for i in range(3))
  C = qc.mo_spec.get_coeffs()
  mo = C @ core_oracle.ao_creator(qc.geo_spec, qc.ao_spec, None, x, y, z, get_cart2sph)
  dmo = numpy.array([C @ core_oracle.ao_creator(qc.geo_spec, qc.ao_spec, d, x, y, z, get_cart2sph)
                     for d in ('x', 'y', 'z')])
  return qc, (x, y, z), mo, dmo


def _terms(nmo, seed=1):
  rng = numpy.random.default_rng(seed)
  zero = [[list(rng.uniform(0.1, 2.0, 3)) for _ in range(4)],
          [list(rng.choice(nmo, 3, replace=False)) for _ in range(4)]]
  sing = [list(rng.normal(size=9)), [list(rng.choice(nmo, 2, replace=False)) for _ in range(9)]]
  return zero, sing


def test_ci_oracle_invariants():
  from oracle import ci_oracle as co, core_oracle
  qc, (x, y, z), mo, dmo = _mos()
  occ = qc.mo_spec.get_occ()
  zero = [[list(occ)], [list(range(len(occ)))]]
  assert rel_err(co.rho(zero, [[], []], mo), core_oracle.rho_compute(qc, x, y, z, True)) < 1e-13
  zero, sing = _terms(len(occ))
  swapped = [sing[0], [[b, a] for a, b in sing[1]]]
  j = co.jab(zero, sing, mo, dmo)
  assert rel_err(co.jab(zero, swapped, mo, dmo), -j) < 1e-14
  anb = co.a_nabla_b(zero, sing, mo, dmo)
  bna = co.a_nabla_b(zero, swapped, mo, dmo)
  assert rel_err(-0.5 * (anb - bna), j) < 1e-13
  # time-dependent combinations: one unit coefficient selects one pair
  ns, nt = 3, 2
  npair = ns * (ns + 1) // 2
  rho = numpy.random.default_rng(2).normal(size=(npair, 50))
  ReS = numpy.zeros((nt, ns, ns))
  ReS[1, 2, 0] = 1.0   # pair (n=0, m=2) is count 2
  out = co.rho_full(ReS, rho)
  assert numpy.array_equal(out[1], 2.0 * rho[2]) and not numpy.any(out[0])
  jj = numpy.random.default_rng(3).normal(size=(npair, 3, 50))
  ImS = numpy.zeros((nt, ns, ns))
  ImS[0, 1, 2] = 0.5   # pair (n=1, m=2) is count 4
  out = co.j_full(ImS, jj)
  assert numpy.array_equal(out[0], -jj[4]) and not numpy.any(out[1])
  # jab_full with a single basis pair is jab with coefficient -2 ImS / mu
  Ib = numpy.zeros((len(occ), len(occ)))
  Ib[3, 1] = 0.7
  jf = co.jab_full(Ib, mo, dmo, 2.0)
  assert rel_err(jf, co.jab(None, [[-2.0 * 0.7 / 2.0], [[3, 1]]], mo, dmo)) < 1e-14


@pytest.mark.gpu
def test_ci_device_vs_oracle():
  import torch
  if not torch.cuda.is_available():
    pytest.skip('no HIP device')
  from oracle import ci_oracle as co
  from orbkit_amd import detci
  print('case detci', flush=True)
  qc, _, mo, dmo = _mos(n=1001)
  zero, sing = _terms(mo.shape[0])
  assert rel_err(detci.rho(zero, sing, mo), co.rho(zero, sing, mo)) < 1e-12
  assert rel_err(detci.jab(zero, sing, mo, dmo), co.jab(zero, sing, mo, dmo)) < 1e-12
  assert rel_err(detci.a_nabla_b(zero, sing, mo, dmo), co.a_nabla_b(zero, sing, mo, dmo)) < 1e-12
  # grid-shaped input keeps its shape
  g = mo[:, :1000].reshape(mo.shape[0], 10, 10, 10)
  assert detci.rho(zero, sing, g).shape == (10, 10, 10)
  rng = numpy.random.default_rng(4)
  ns, nt = 4, 3
  npair = ns * (ns + 1) // 2
  rho = rng.normal(size=(npair, 1001))
  jj = rng.normal(size=(npair, 3, 1001))
  ReS, ImS = rng.normal(size=(nt, ns, ns)), rng.normal(size=(nt, ns, ns))
  assert rel_err(detci.get_rho_full(ReS, rho), co.rho_full(ReS, rho)) < 1e-12
  assert rel_err(detci.get_j_full(ImS, jj), co.j_full(ImS, jj)) < 1e-12
  Ib = rng.normal(size=(mo.shape[0], mo.shape[0]))
  assert rel_err(detci.get_jab_full(Ib, mo, dmo, 1.7), co.jab_full(Ib, mo, dmo, 1.7)) < 1e-12
  with pytest.raises(RuntimeError):
    detci.rho([[[1.0]], [[mo.shape[0]]]], [[], []], mo)   # orbital index out of range
