import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
  config.addinivalue_line('markers', 'gpu: needs a HIP device and liborbkit_amd.so')


@pytest.fixture(scope='session')
def golden():
  return GOLDEN


def rel_err(a, b):
  '''max |a-b| / max |b| (norm-wise relative error).'''
  a = numpy.asarray(a, dtype=float)
  b = numpy.asarray(b, dtype=float)
  den = numpy.max(numpy.abs(b)) if b.size else 1.0
  return float(numpy.max(numpy.abs(a - b)) / den) if a.size else 0.0


def gpu_available():
  try:
    import torch
    return torch.cuda.is_available()
  except Exception:
    return False
