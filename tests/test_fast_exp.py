'''The reduced-precision exponential of the FP32-output AO paths
(gauss_f64_fast in orbkit_amd/csrc/ao_device.h) against exp itself: its
constants are read from the header and the function is replayed on the CPU
(numpy float64, the same operation order), so a wrong or mistyped
coefficient or reduction constant fails here.'''
import math
import os
import re

import numpy
import pytest

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'orbkit_amd', 'csrc', 'ao_device.h')


def _constants():
  src = open(HDR).read()
  body = src[src.index('double gauss_f64_fast(double t)'):]
  body = body[:body.index('\n}\n')]
  num = r'([0-9.]+(?:e[-+]?[0-9]+)?)'
  red = [float(x) for x in re.findall(r'fma\(-k, ' + num, body)]
  inv_ln2 = float(re.search(r'rint\(t \* ' + num, body).group(1))
  first = float(re.search(r'double p = ' + num, body).group(1))
  rest = [float(x) for x in re.findall(r'p = fma\(p, s, ' + num, body)]
  return inv_ln2, red, [first] + rest


def _fast_exp(t):
  inv_ln2, (ln2_hi, ln2_lo), coef = _constants()
  t = numpy.minimum(t, 1.0e4)
  k = numpy.rint(t * inv_ln2)
  r = (t - k * ln2_hi) - k * ln2_lo
  s = -r
  p = numpy.full_like(s, coef[0])
  for c in coef[1:]:
    p = p * s + c
  return numpy.ldexp(p, -k.astype(numpy.int64))


def test_constants_shape():
  inv_ln2, red, coef = _constants()
  assert inv_ln2 == pytest.approx(1.0 / math.log(2.0), rel=1e-15)
  assert len(red) == 2 and red[0] + red[1] == pytest.approx(math.log(2.0), rel=1e-16)
  assert len(coef) == 7   # degree 6


def test_relative_error_below_fp32_half_ulp():
  # FP32 half-ulp is 2^-24 = 6.0e-8; the polynomial keeps below 3e-9
  t = numpy.concatenate([numpy.linspace(0.0, 2.0, 200001), numpy.linspace(2.0, 700.0, 200001),
                         numpy.array([0.5 * math.log(2.0) * j for j in range(1, 2000)])])
  ref = numpy.exp(-t)
  got = _fast_exp(t)
  rel = numpy.abs(got - ref) / ref
  assert rel.max() < 3e-9, rel.max()


def test_underflow_is_zero():
  got = _fast_exp(numpy.array([800.0, 1e4, 1e6, numpy.inf]))
  assert (got == 0.0).all()
